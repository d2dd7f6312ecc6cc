// peps_oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's explicit-C PEPS row-boundary contraction
// (crate `nqueens-peps-naive`, src/lib.rs and src/dfs_bitmask.rs).  It is the
// checker for the HIP library and the timed CPU baseline; nothing in the
// product library links it.  Every function cites the reference lines it
// follows.  Arithmetic is exact: checked unsigned __int128 like the
// reference's checked u128.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/peps_mi355x.h"

namespace oracle {

using u128 = unsigned __int128;

struct OracleError : std::runtime_error {
  explicit OracleError(const std::string& m) : std::runtime_error(m) {}
};

// VirtualLegs, src/lib.rs:19-29.
struct Legs {
  uint8_t column_in, column_out, row_in, row_out, diag_dr_in, diag_dr_out, diag_dl_in, diag_dl_out;
  bool operator==(const Legs& o) const;
};

struct BEntry {  // src/lib.rs:31-36
  uint8_t alpha;
  Legs legs;
  u128 value;
};

struct CEntry {  // src/lib.rs:89-93
  Legs legs;
  u128 value;
};

// SiteTensorB::sec_vi, src/lib.rs:43-83.
std::vector<BEntry> site_tensor_b();

// SiteTensorC::from_b / sec_vi / matching_entries, src/lib.rs:100-157.
struct SiteTensorC {
  std::vector<CEntry> entries;                    // sorted by legs_key
  std::array<std::vector<CEntry>, 16> by_input;  // input_key buckets
  static SiteTensorC from_b(const std::vector<BEntry>& b);
  static SiteTensorC sec_vi();
  // A caller-supplied C (mutation tests): entries kept in the given order and
  // bucketed by input_key like from_b.
  static SiteTensorC from_entries(const std::vector<CEntry>& entries);
  const std::vector<CEntry>& matching(uint8_t col, uint8_t row, uint8_t dr, uint8_t dl) const;
};

unsigned input_key(uint8_t col, uint8_t row, uint8_t dr, uint8_t dl);  // src/lib.rs:159-164
unsigned legs_key(const Legs& l);                                       // src/lib.rs:166-175

struct Boundary {  // BoundaryState, src/lib.rs:170-178
  uint64_t columns, diag_dr, diag_dl;
  bool operator==(const Boundary& o) const {
    return columns == o.columns && diag_dr == o.diag_dr && diag_dl == o.diag_dl;
  }
};

// PackedBoundary::pack / unpack, src/lib.rs:348-377.
u128 pack(const Boundary& s, unsigned n);
Boundary unpack(u128 key, unsigned n);

// CompiledRowOperator::compile, src/lib.rs:425-485 (fail-closed).
struct CompiledRowOperator {
  CEntry occupied;
  static CompiledRowOperator compile(const SiteTensorC& c);
};

struct RowCounters {  // src/lib.rs:417-423
  u128 tensor_examined = 0, tensor_matched = 0, operator_candidates = 0, operator_matched = 0;
};

using Term = std::pair<Boundary, u128>;

// contract_one_row_sitewise, src/lib.rs:502-592.  `linear_scan_counters`
// reproduces the counters of the pre-index revision (17 entries examined per
// partial row per site), which the published naive CSVs were measured with.
std::vector<Term> contract_one_row_sitewise(unsigned n, const SiteTensorC& c, Boundary parent,
                                            u128 weight, RowCounters& k,
                                            bool linear_scan_counters = false);
// contract_one_row_compiled, src/lib.rs:594-638.
std::vector<Term> contract_one_row_compiled(unsigned n, const CompiledRowOperator& op,
                                            Boundary parent, u128 weight, RowCounters& k);
// contract_one_row_compiled_sparse, src/lib.rs:640-689.
std::vector<Term> contract_one_row_compiled_sparse(unsigned n, const CompiledRowOperator& op,
                                                   Boundary parent, u128 weight, RowCounters& k);

// top_row_vertical_orbit_weight, src/lib.rs:721-732; 0 means "not kept".
u128 top_row_vertical_orbit_weight(unsigned n, const Boundary& successor);

enum class Symmetry { None, TopRowOrbits };
enum class Position { Dense, Sparse };
enum class Backend { SortReduce, Hash, Sitewise, ParallelSortReduce };

struct Options {
  Symmetry symmetry = Symmetry::TopRowOrbits;
  Position position = Position::Dense;
  Backend backend = Backend::SortReduce;
  unsigned threads = 1;
  bool linear_scan_counters = false;  // sitewise only
  // Optional start boundary (k placed rows); used by prefix counts.  When
  // set, the first contracted row is `start_row` and no orbit slicing applies.
  bool has_start = false;
  Boundary start{0, 0, 0};
  unsigned start_row = 0;
  u128 start_weight = 1;
  // Sitewise backend only: contract this C instead of Sec. VI's (no compile).
  const SiteTensorC* custom_tensor = nullptr;
  // Called after every row with the merged layer (packed key, weight).
  std::function<void(unsigned row, const std::vector<std::pair<u128, u128>>& layer)> on_layer;
};

// contract_rows_sort_reduce_with_modes src/lib.rs:5287-5415,
// contract_rows_parallel_sort_reduce_with_modes src/lib.rs:5469-5626,
// contract_rows_with_backend src/lib.rs:5633-5750.
void contract_rows(unsigned n, const Options& opt, peps_contraction_result* out);

// known_count, src/lib.rs:5752-5784.
bool known_count(unsigned n, u128* out);

// CRT: wide_crt_plan src/lib.rs:3859-3886, reconstruct_crt src/lib.rs:4134-4163.
struct CrtPlan {
  unsigned n;
  u128 factorial_bound;
  std::vector<uint64_t> primes;
  u128 modulus_product;
};
CrtPlan wide_crt_plan(unsigned n);
u128 reconstruct_crt(const std::vector<uint64_t>& residues, const std::vector<uint64_t>& primes);
uint64_t mod_pow(uint64_t base, uint64_t exp, uint64_t m);  // src/lib.rs:4122-4132
bool is_prime_u32_range(uint64_t v);                        // src/lib.rs:3834-3849

// build_wide_crt_tasks, src/lib.rs:3888-3956, with a fixed split depth.
struct PrefixTask {
  Boundary state;
  uint64_t orbit_weight;
};
std::vector<PrefixTask> prefix_tasks(unsigned n, unsigned depth);

// DFS comparator, src/dfs_bitmask.rs:67-352.
struct DfsResult {
  u128 count;
  unsigned split_depth;
  size_t tasks;
  u128 recursive_nodes, candidate_placements;
};
DfsResult count_dfs(unsigned n, unsigned threads, bool track_metrics);
// Completions of the prefix cols[0..k) (queen of row i in column cols[i]).
u128 count_prefix_dfs(unsigned n, const uint32_t* cols, unsigned k);
// Boundary reached after placing the prefix (0 weight if it attacks).
bool prefix_boundary(unsigned n, const uint32_t* cols, unsigned k, Boundary* out);

uint64_t peak_rss_bytes();

std::string u128_to_string(u128 v);

// layer_census.cpp: per-depth orbit-weighted prefix counts (index r = r rows
// placed, r = 0..n) and the unweighted per-row state counts of the default
// row loop, by a DFS and by a breadth-first pass grouped by column set.
std::vector<u128> depth_counts(unsigned n, bool orbits, unsigned threads);
u128 count_prefix_parallel(unsigned n, const uint32_t* cols, unsigned k, unsigned threads);
struct CensusRow {
  unsigned row;
  uint64_t input_states, matched, completed, output_states;
};
std::vector<CensusRow> frontier_census(unsigned n, bool orbits, unsigned threads);

}  // namespace oracle
