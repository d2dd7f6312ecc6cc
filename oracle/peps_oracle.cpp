// peps_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
//
// Restates the reference's explicit-C PEPS row contraction on the CPU with
// checked unsigned __int128 arithmetic.  Used by tests/ as the checker and by
// bench.py's cpu_baseline leg as the timed CPU path; the product library
// libpeps_mi355x.so never links or loads it.
#include "peps_oracle.hpp"

#include <sys/resource.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
#include <parallel/algorithm>
#include <thread>
#include <unordered_map>

namespace oracle {

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

u128 checked_add(u128 a, u128 b, const std::string& what) {
  u128 r;
  if (__builtin_add_overflow(a, b, &r)) throw OracleError(what);
  return r;
}

u128 checked_mul(u128 a, u128 b, const std::string& what) {
  u128 r;
  if (__builtin_mul_overflow(a, b, &r)) throw OracleError(what);
  return r;
}

peps_u128 to_abi(u128 v) { return peps_u128{(uint64_t)v, (uint64_t)(v >> 64)}; }

uint8_t bit(uint64_t mask, unsigned index) { return (uint8_t)((mask >> index) & 1u); }

void set_bit(uint64_t* mask, unsigned index, uint8_t value) {
  if (value == 1) *mask |= (uint64_t)1 << index;
}

uint64_t replace_bit(uint64_t mask, unsigned index, uint8_t value) {
  uint64_t selected = (uint64_t)1 << index;
  return (mask & ~selected) | ((uint64_t)value << index);
}

struct U128Hash {
  size_t operator()(u128 v) const {
    uint64_t x = (uint64_t)v ^ ((uint64_t)(v >> 64) * 0x9e3779b97f4a7c15ull);
    x ^= x >> 31;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 29;
    return (size_t)x;
  }
};

}  // namespace

bool Legs::operator==(const Legs& o) const {
  return column_in == o.column_in && column_out == o.column_out && row_in == o.row_in &&
         row_out == o.row_out && diag_dr_in == o.diag_dr_in && diag_dr_out == o.diag_dr_out &&
         diag_dl_in == o.diag_dl_in && diag_dl_out == o.diag_dl_out;
}

// src/lib.rs:43-83: 16 alpha=0 identity pass-throughs (signals = 4 bits:
// column, row, diag_dr, diag_dl) plus the alpha=1 all-channel 0->1 emission.
std::vector<BEntry> site_tensor_b() {
  std::vector<BEntry> e;
  e.reserve(17);
  for (uint8_t s = 0; s < 16; ++s) {
    uint8_t c = s & 1, r = (s >> 1) & 1, dr = (s >> 2) & 1, dl = (s >> 3) & 1;
    e.push_back(BEntry{0, Legs{c, c, r, r, dr, dr, dl, dl}, 1});
  }
  e.push_back(BEntry{1, Legs{0, 1, 0, 1, 0, 1, 0, 1}, 1});
  return e;
}

unsigned input_key(uint8_t col, uint8_t row, uint8_t dr, uint8_t dl) {
  return (unsigned)col | ((unsigned)row << 1) | ((unsigned)dr << 2) | ((unsigned)dl << 3);
}

unsigned legs_key(const Legs& l) {
  return (unsigned)l.column_in | ((unsigned)l.column_out << 1) | ((unsigned)l.row_in << 2) |
         ((unsigned)l.row_out << 3) | ((unsigned)l.diag_dr_in << 4) |
         ((unsigned)l.diag_dr_out << 5) | ((unsigned)l.diag_dl_in << 6) |
         ((unsigned)l.diag_dl_out << 7);
}

// SiteTensorC::from_b, src/lib.rs:102-131: sum B over alpha (checked), sort
// by legs key, bucket by the four incoming signals.
SiteTensorC SiteTensorC::from_b(const std::vector<BEntry>& b) {
  std::map<unsigned, CEntry> acc;  // keyed by legs_key => already sorted
  for (const BEntry& be : b) {
    unsigned k = legs_key(be.legs);
    auto it = acc.find(k);
    if (it == acc.end()) {
      acc.emplace(k, CEntry{be.legs, be.value});
    } else {
      it->second.value = checked_add(it->second.value, be.value, "local tensor coefficient overflow");
    }
  }
  SiteTensorC c;
  for (auto& kv : acc) c.entries.push_back(kv.second);
  for (const CEntry& ce : c.entries) {
    c.by_input[input_key(ce.legs.column_in, ce.legs.row_in, ce.legs.diag_dr_in, ce.legs.diag_dl_in)]
        .push_back(ce);
  }
  return c;
}

SiteTensorC SiteTensorC::sec_vi() { return from_b(site_tensor_b()); }

SiteTensorC SiteTensorC::from_entries(const std::vector<CEntry>& entries) {
  SiteTensorC c;
  c.entries = entries;
  for (const CEntry& ce : c.entries) {
    for (uint8_t v : {ce.legs.column_in, ce.legs.column_out, ce.legs.row_in, ce.legs.row_out,
                      ce.legs.diag_dr_in, ce.legs.diag_dr_out, ce.legs.diag_dl_in, ce.legs.diag_dl_out})
      if (v > 1) throw OracleError("C entry with a non-binary virtual leg");
    c.by_input[input_key(ce.legs.column_in, ce.legs.row_in, ce.legs.diag_dr_in, ce.legs.diag_dl_in)]
        .push_back(ce);
  }
  return c;
}

const std::vector<CEntry>& SiteTensorC::matching(uint8_t col, uint8_t row, uint8_t dr,
                                                 uint8_t dl) const {
  return by_input[input_key(col, row, dr, dl)];
}

// PackedBoundary::pack/unpack, src/lib.rs:357-371: bits [0,n) columns,
// [n,2n) diag_dr, [2n,3n) diag_dl.
u128 pack(const Boundary& s, unsigned n) {
  return (u128)s.columns | ((u128)s.diag_dr << n) | ((u128)s.diag_dl << (2 * n));
}

Boundary unpack(u128 key, unsigned n) {
  u128 mask = ((u128)1 << n) - 1;
  return Boundary{(uint64_t)(key & mask), (uint64_t)((key >> n) & mask),
                  (uint64_t)((key >> (2 * n)) & mask)};
}

// CompiledRowOperator::compile, src/lib.rs:437-485.
CompiledRowOperator CompiledRowOperator::compile(const SiteTensorC& c) {
  bool passthrough[16] = {false};
  bool have_occupied = false;
  CEntry occupied{};
  for (const CEntry& e : c.entries) {
    const Legs& l = e.legs;
    bool is_pass = l.column_in == l.column_out && l.row_in == l.row_out &&
                   l.diag_dr_in == l.diag_dr_out && l.diag_dl_in == l.diag_dl_out;
    bool is_occ = l.column_in == 0 && l.column_out == 1 && l.row_in == 0 && l.row_out == 1 &&
                  l.diag_dr_in == 0 && l.diag_dr_out == 1 && l.diag_dl_in == 0 &&
                  l.diag_dl_out == 1;
    if (e.value != 1) throw OracleError("compiled row operator requires unit Sec. VI coefficients");
    if (is_pass) {
      unsigned key = input_key(l.column_in, l.row_in, l.diag_dr_in, l.diag_dl_in);
      if (passthrough[key]) throw OracleError("duplicate C pass-through signature");
      passthrough[key] = true;
    } else if (is_occ) {
      if (have_occupied) throw OracleError("multiple occupied C entries");
      have_occupied = true;
      occupied = e;
    } else {
      throw OracleError("C contains an entry outside the Sec. VI row automaton");
    }
  }
  for (bool p : passthrough)
    if (!p) throw OracleError("C is missing a pass-through signature");
  if (!have_occupied) throw OracleError("C is missing the occupied entry");
  return CompiledRowOperator{occupied};
}

// contract_one_row_sitewise, src/lib.rs:513-592.  Left row boundary v0
// (incoming row signal 0); right boundary v1 keeps row_signal == 1; diagonal
// signals leaving the board are contracted with v2 and dropped.
std::vector<Term> contract_one_row_sitewise(unsigned n, const SiteTensorC& c, Boundary parent,
                                            u128 weight, RowCounters& k,
                                            bool linear_scan_counters) {
  struct Partial {
    uint64_t cols, dr, dl;
    uint8_t row_signal;
    u128 weight;
  };
  std::vector<Partial> partials{Partial{0, 0, 0, 0, weight}}, next;
  for (unsigned col = 0; col < n; ++col) {
    next.clear();
    uint8_t cin = bit(parent.columns, col), drin = bit(parent.diag_dr, col),
            dlin = bit(parent.diag_dl, col);
    for (const Partial& p : partials) {
      const auto& m = c.matching(cin, p.row_signal, drin, dlin);
      k.tensor_examined += linear_scan_counters ? (u128)c.entries.size() : (u128)m.size();
      for (const CEntry& e : m) {
        k.tensor_matched += 1;
        Partial s{p.cols, p.dr, p.dl, e.legs.row_out,
                  checked_mul(p.weight, e.value, "coefficient overflow in local C contraction")};
        set_bit(&s.cols, col, e.legs.column_out);
        if (col + 1 < n) set_bit(&s.dr, col + 1, e.legs.diag_dr_out);
        if (col > 0) set_bit(&s.dl, col - 1, e.legs.diag_dl_out);
        next.push_back(s);
      }
    }
    std::swap(partials, next);
  }
  std::vector<Term> out;
  for (const Partial& p : partials)
    if (p.row_signal == 1) out.push_back(Term{Boundary{p.cols, p.dr, p.dl}, p.weight});
  return out;
}

// contract_one_row_compiled, src/lib.rs:600-638.
std::vector<Term> contract_one_row_compiled(unsigned n, const CompiledRowOperator& op,
                                            Boundary parent, u128 weight, RowCounters& k) {
  const Legs& l = op.occupied.legs;
  uint64_t board_mask = (n == 64) ? ~0ull : (((uint64_t)1 << n) - 1);
  std::vector<Term> out;
  for (unsigned col = 0; col < n; ++col) {
    k.operator_candidates += 1;
    if (bit(parent.columns, col) != l.column_in || bit(parent.diag_dr, col) != l.diag_dr_in ||
        bit(parent.diag_dl, col) != l.diag_dl_in)
      continue;
    k.operator_matched += 1;
    uint64_t cols = replace_bit(parent.columns, col, l.column_out);
    uint64_t dr = replace_bit(parent.diag_dr, col, l.diag_dr_out);
    uint64_t dl = replace_bit(parent.diag_dl, col, l.diag_dl_out);
    u128 w = checked_mul(weight, op.occupied.value, "coefficient overflow in compiled row operator");
    out.push_back(Term{Boundary{cols, (dr << 1) & board_mask, dl >> 1}, w});
  }
  return out;
}

// contract_one_row_compiled_sparse, src/lib.rs:640-689.
std::vector<Term> contract_one_row_compiled_sparse(unsigned n, const CompiledRowOperator& op,
                                                   Boundary parent, u128 weight, RowCounters& k) {
  const Legs& l = op.occupied.legs;
  if (l.row_in != 0)
    throw OracleError("sparse row iterator requires occupied row_in to match the left v0 boundary");
  uint64_t board_mask = ((uint64_t)1 << n) - 1;
  auto matching_bits = [&](uint64_t mask, uint8_t required) -> uint64_t {
    if (required == 0) return (~mask) & board_mask;
    if (required == 1) return mask & board_mask;
    throw OracleError("compiled C entry contains a non-binary incoming signal");
  };
  uint64_t positions = matching_bits(parent.columns, l.column_in) &
                       matching_bits(parent.diag_dr, l.diag_dr_in) &
                       matching_bits(parent.diag_dl, l.diag_dl_in);
  u128 w = checked_mul(weight, op.occupied.value,
                       "coefficient overflow in sparse compiled row operator");
  std::vector<Term> out;
  while (positions) {
    unsigned col = (unsigned)__builtin_ctzll(positions);
    positions &= positions - 1;
    k.operator_candidates += 1;
    k.operator_matched += 1;
    uint64_t cols = replace_bit(parent.columns, col, l.column_out);
    uint64_t dr = replace_bit(parent.diag_dr, col, l.diag_dr_out);
    uint64_t dl = replace_bit(parent.diag_dl, col, l.diag_dl_out);
    out.push_back(Term{Boundary{cols, (dr << 1) & board_mask, dl >> 1}, w});
  }
  return out;
}

// top_row_vertical_orbit_weight, src/lib.rs:721-732.
u128 top_row_vertical_orbit_weight(unsigned n, const Boundary& s) {
  unsigned column = (unsigned)__builtin_ctzll(s.columns);
  unsigned mirror = n - 1 - column;
  if (column > mirror) return 0;
  if (column == mirror) return 1;
  return 2;
}

namespace {

// apply_top_row_symmetry, src/lib.rs:734-749.
std::vector<Term> apply_top_row_symmetry(unsigned n, std::vector<Term> terms) {
  std::vector<Term> out;
  for (const Term& t : terms) {
    u128 m = top_row_vertical_orbit_weight(n, t.first);
    if (m == 0) continue;
    out.push_back(Term{t.first, checked_mul(t.second, m, "coefficient overflow in D4 orbit weighting")});
  }
  return out;
}

std::vector<Term> row_terms(unsigned n, const Options& opt, const SiteTensorC& c,
                            const CompiledRowOperator& op, Boundary parent, u128 w,
                            RowCounters& k) {
  if (opt.backend == Backend::Sitewise)
    return contract_one_row_sitewise(n, c, parent, w, k, opt.linear_scan_counters);
  if (opt.position == Position::Sparse) return contract_one_row_compiled_sparse(n, op, parent, w, k);
  return contract_one_row_compiled(n, op, parent, w, k);
}

void fill_layer(peps_layer_metric* L, unsigned row, uint64_t input_states, const RowCounters& k,
                u128 completed, uint64_t output_states, u128 output_weight, double elapsed) {
  std::memset(L, 0, sizeof(*L));
  L->row = row;
  L->input_states = input_states;
  L->tensor_entries_examined = to_abi(k.tensor_examined);
  L->tensor_entries_matched = to_abi(k.tensor_matched);
  L->row_operator_candidates = to_abi(k.operator_candidates);
  L->row_operator_matched = to_abi(k.operator_matched);
  L->completed_row_terms = to_abi(completed);
  L->output_states = output_states;
  L->output_weight = to_abi(output_weight);
  L->elapsed_s = elapsed;
  L->peak_rss_bytes = peak_rss_bytes();
}

using Record = std::pair<u128, u128>;

// Adjacent checked reduce of sorted candidates, src/lib.rs:5354-5367.
void reduce_sorted(std::vector<Record>& cand, unsigned row) {
  size_t write = 0;
  for (size_t read = 0; read < cand.size(); ++read) {
    if (write > 0 && cand[write - 1].first == cand[read].first) {
      cand[write - 1].second = checked_add(cand[write - 1].second, cand[read].second,
                                           "coefficient overflow after row " + std::to_string(row + 1));
    } else {
      cand[write++] = cand[read];
    }
  }
  cand.resize(write);
}

}  // namespace

// contract_rows_sort_reduce_with_modes (src/lib.rs:5287-5415), the parallel
// form (src/lib.rs:5469-5626) and the HashMap form (src/lib.rs:5633-5750).
void contract_rows(unsigned n, const Options& opt, peps_contraction_result* out) {
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  if (opt.backend == Backend::ParallelSortReduce && opt.threads == 0)
    throw OracleError("parallel sort-reduce requires at least one thread");
  if (n > 42) throw OracleError("the packed u128 virtual-boundary backend supports N <= 42");
  if (n == 0) {
    out->count = to_abi(1);
    out->peak_states = 1;
    out->peak_rss_bytes = peak_rss_bytes();
    return;
  }
  const bool hash = opt.backend == Backend::Hash || opt.backend == Backend::Sitewise;
  if (opt.custom_tensor && opt.backend != Backend::Sitewise)
    throw OracleError("a custom C tensor needs the sitewise backend");
  SiteTensorC tensor = opt.custom_tensor ? *opt.custom_tensor : SiteTensorC::sec_vi();
  CompiledRowOperator op = CompiledRowOperator::compile(SiteTensorC::sec_vi());
  const unsigned first_row = opt.has_start ? opt.start_row : 0;
  const bool orbits = opt.symmetry == Symmetry::TopRowOrbits && !opt.has_start;

  std::vector<Record> boundary{Record{pack(opt.has_start ? opt.start : Boundary{0, 0, 0}, n),
                                      opt.has_start ? opt.start_weight : (u128)1}};
  uint64_t peak_states = 1;
  u128 total_cand = 0, total_matched = 0, total_examined = 0, total_tmatched = 0;
  if (hash && opt.backend == Backend::Hash) total_examined = total_tmatched = 17;
  auto t0 = Clock::now();

  for (unsigned row = first_row; row < n; ++row) {
    auto lt0 = Clock::now();
    uint64_t input_states = boundary.size();
    RowCounters k;
    u128 completed = 0;
    std::vector<Record> next;

    if (hash) {
      // HashMap accumulate, src/lib.rs:5680-5699.
      std::unordered_map<u128, u128, U128Hash> acc;
      for (const Record& p : boundary) {
        auto terms = row_terms(n, opt, tensor, op, unpack(p.first, n), p.second, k);
        if (row == 0 && orbits) terms = apply_top_row_symmetry(n, std::move(terms));
        for (const Term& t : terms) {
          completed += 1;
          u128& v = acc[pack(t.first, n)];
          v = checked_add(v, t.second, "coefficient overflow after row " + std::to_string(row + 1));
        }
      }
      next.reserve(acc.size());
      for (auto& kv : acc) next.push_back(Record{kv.first, kv.second});
    } else if (opt.backend == Backend::SortReduce) {
      for (const Record& p : boundary) {
        auto terms = row_terms(n, opt, tensor, op, unpack(p.first, n), p.second, k);
        if (row == 0 && orbits) terms = apply_top_row_symmetry(n, std::move(terms));
        completed += terms.size();
        for (const Term& t : terms) next.push_back(Record{pack(t.first, n), t.second});
      }
      boundary.clear();
      boundary.shrink_to_fit();
      std::sort(next.begin(), next.end(),
                [](const Record& a, const Record& b) { return a.first < b.first; });
      reduce_sorted(next, row);
    } else {
      // Parallel: chunk = ceil(S / (4T)), src/lib.rs:5519-5545.
      const unsigned T = opt.threads;
      size_t target = (size_t)T * 4;
      size_t chunk = std::max<size_t>(1, (input_states + target - 1) / target);
      size_t nchunks = (input_states + chunk - 1) / chunk;
      std::vector<std::vector<Record>> locals(nchunks);
      std::vector<RowCounters> lk(nchunks);
      std::atomic<size_t> next_chunk{0};
      std::vector<std::string> errors(T);
      auto worker = [&](unsigned tid) {
        try {
          for (;;) {
            size_t ci = next_chunk.fetch_add(1);
            if (ci >= nchunks) break;
            size_t lo = ci * chunk, hi = std::min(input_states, lo + chunk);
            for (size_t i = lo; i < hi; ++i) {
              auto terms = row_terms(n, opt, tensor, op, unpack(boundary[i].first, n),
                                     boundary[i].second, lk[ci]);
              if (row == 0 && orbits) terms = apply_top_row_symmetry(n, std::move(terms));
              for (const Term& t : terms) locals[ci].push_back(Record{pack(t.first, n), t.second});
            }
          }
        } catch (const std::exception& e) {
          errors[tid] = e.what();
        }
      };
      std::vector<std::thread> pool;
      for (unsigned t = 1; t < T; ++t) pool.emplace_back(worker, t);
      worker(0);
      for (auto& th : pool) th.join();
      for (auto& e : errors)
        if (!e.empty()) throw OracleError(e);
      size_t total = 0;
      for (auto& l : locals) total += l.size();
      next.reserve(total);
      for (size_t ci = 0; ci < nchunks; ++ci) {
        next.insert(next.end(), locals[ci].begin(), locals[ci].end());
        std::vector<Record>().swap(locals[ci]);
        k.tensor_examined += lk[ci].tensor_examined;
        k.tensor_matched += lk[ci].tensor_matched;
        k.operator_candidates += lk[ci].operator_candidates;
        k.operator_matched += lk[ci].operator_matched;
      }
      completed = total;
      boundary.clear();
      boundary.shrink_to_fit();
      __gnu_parallel::sort(next.begin(), next.end(),
                           [](const Record& a, const Record& b) { return a.first < b.first; },
                           __gnu_parallel::multiway_mergesort_tag((int)T));
      reduce_sorted(next, row);
    }
    boundary.swap(next);
    if (opt.on_layer) opt.on_layer(row, boundary);

    u128 output_weight = 0;
    for (const Record& r : boundary)
      output_weight = checked_add(output_weight, r.second,
                                  "coefficient sum overflow after row " + std::to_string(row + 1));
    peak_states = std::max<uint64_t>(peak_states, boundary.size());
    total_cand += k.operator_candidates;
    total_matched += k.operator_matched;
    total_examined += k.tensor_examined;
    total_tmatched += k.tensor_matched;
    if (out->layer_count < PEPS_MAX_N) {
      fill_layer(&out->layers[out->layer_count], row, input_states, k, completed, boundary.size(),
                 output_weight, seconds_since(lt0));
      out->layer_count++;
    }
  }

  // Final v1/v2 contraction, src/lib.rs:5393-5401.
  uint64_t board_mask = ((uint64_t)1 << n) - 1;
  u128 count = 0;
  for (const Record& r : boundary)
    if (unpack(r.first, n).columns == board_mask)
      count = checked_add(count, r.second, "final coefficient sum overflow");

  out->count = to_abi(count);
  out->elapsed_s = seconds_since(t0);
  out->peak_states = peak_states;
  if (hash) {
    out->tensor_entries_examined = to_abi(total_examined);
    out->tensor_entries_matched = to_abi(total_tmatched);
  } else {
    out->tensor_entries_examined = to_abi(17);
    out->tensor_entries_matched = to_abi(17);
  }
  out->row_operator_candidates = to_abi(total_cand);
  out->row_operator_matched = to_abi(total_matched);
  out->peak_rss_bytes = peak_rss_bytes();
}

// known_count, src/lib.rs:5752-5784 (OEIS A000170).
bool known_count(unsigned n, u128* out) {
  static const char* kCounts[28] = {"1",
                                    "1",
                                    "0",
                                    "0",
                                    "2",
                                    "10",
                                    "4",
                                    "40",
                                    "92",
                                    "352",
                                    "724",
                                    "2680",
                                    "14200",
                                    "73712",
                                    "365596",
                                    "2279184",
                                    "14772512",
                                    "95815104",
                                    "666090624",
                                    "4968057848",
                                    "39029188884",
                                    "314666222712",
                                    "2691008701644",
                                    "24233937684440",
                                    "227514171973736",
                                    "2207893435808352",
                                    "22317699616364044",
                                    "234907967154122528"};
  if (n >= 28) return false;
  u128 v = 0;
  for (const char* p = kCounts[n]; *p; ++p) v = v * 10 + (u128)(*p - '0');
  *out = v;
  return true;
}

// is_prime_u32_range, src/lib.rs:3834-3849 (trial division).
bool is_prime_u32_range(uint64_t v) {
  if (v < 2) return false;
  if (v % 2 == 0) return v == 2;
  for (uint64_t d = 3; d * d <= v; d += 2)
    if (v % d == 0) return false;
  return true;
}

static const uint64_t kPrimes[4] = {4294967291ull, 4294967279ull, 4294967231ull, 4294967197ull};

// wide_crt_plan, src/lib.rs:3859-3886.
CrtPlan wide_crt_plan(unsigned n) {
  if (n > 34) throw OracleError("wide CRT backend supports N <= 34 because 35! exceeds u128");
  u128 fact = 1;
  for (unsigned f = 2; f <= n; ++f) fact = checked_mul(fact, f, "N! exactness bound does not fit u128");
  CrtPlan plan{n, fact, {}, 1};
  for (uint64_t p : kPrimes) {
    if (!is_prime_u32_range(p)) throw OracleError("CRT modulus failed deterministic primality check");
    plan.primes.push_back(p);
    plan.modulus_product = checked_mul(plan.modulus_product, p, "CRT modulus product overflow");
    if (plan.modulus_product > fact) return plan;
  }
  throw OracleError("available certified CRT product does not exceed N! bound");
}

// mod_pow, src/lib.rs:4122-4132.
uint64_t mod_pow(uint64_t base, uint64_t exp, uint64_t m) {
  uint64_t result = 1;
  while (exp) {
    if (exp & 1) result = (uint64_t)(((u128)result * base) % m);
    base = (uint64_t)(((u128)base * base) % m);
    exp >>= 1;
  }
  return result;
}

// reconstruct_crt, src/lib.rs:4134-4163 (Garner step with Fermat inverse).
u128 reconstruct_crt(const std::vector<uint64_t>& residues, const std::vector<uint64_t>& primes) {
  if (residues.size() != primes.size() || residues.empty())
    throw OracleError("CRT residue/modulus length mismatch");
  u128 value = 0, prod = 1;
  for (size_t i = 0; i < residues.size(); ++i) {
    uint64_t p = primes[i], r = residues[i];
    uint64_t vm = (uint64_t)(value % p);
    uint64_t delta = r >= vm ? r - vm : r + p - vm;
    uint64_t pm = (uint64_t)(prod % p);
    uint64_t inv = mod_pow(pm, p - 2, p);
    uint64_t step = (uint64_t)(((u128)delta * inv) % p);
    value = checked_add(value, checked_mul(prod, step, "CRT reconstruction product overflow"),
                        "CRT reconstruction sum overflow");
    prod = checked_mul(prod, p, "CRT reconstruction modulus overflow");
  }
  return value;
}

// build_wide_crt_tasks, src/lib.rs:3888-3956, expanding exactly `depth` rows.
std::vector<PrefixTask> prefix_tasks(unsigned n, unsigned depth) {
  SiteTensorC tensor = SiteTensorC::sec_vi();
  CompiledRowOperator op = CompiledRowOperator::compile(tensor);
  const Legs& l = op.occupied.legs;
  uint64_t board_mask = n ? (((uint64_t)1 << n) - 1) : 0;
  std::vector<PrefixTask> tasks{PrefixTask{Boundary{0, 0, 0}, 1}};
  for (unsigned d = 0; d < depth && d < n; ++d) {
    std::vector<PrefixTask> next;
    for (const PrefixTask& t : tasks) {
      uint64_t positions = ((l.column_in ? t.state.columns : ~t.state.columns) & board_mask) &
                           ((l.diag_dr_in ? t.state.diag_dr : ~t.state.diag_dr) & board_mask) &
                           ((l.diag_dl_in ? t.state.diag_dl : ~t.state.diag_dl) & board_mask);
      while (positions) {
        unsigned col = (unsigned)__builtin_ctzll(positions);
        positions &= positions - 1;
        Boundary s{replace_bit(t.state.columns, col, l.column_out),
                   (replace_bit(t.state.diag_dr, col, l.diag_dr_out) << 1) & board_mask,
                   replace_bit(t.state.diag_dl, col, l.diag_dl_out) >> 1};
        uint64_t w = 1;
        if (d == 0) {
          u128 ow = top_row_vertical_orbit_weight(n, s);
          if (ow == 0) continue;
          w = (uint64_t)ow;
        }
        next.push_back(PrefixTask{s, t.orbit_weight * w * (uint64_t)op.occupied.value});
      }
    }
    tasks.swap(next);
  }
  return tasks;
}

namespace {

// search, src/dfs_bitmask.rs:67-109.
template <bool TRACK>
uint64_t dfs_search(uint64_t board_mask, uint64_t columns, uint64_t dl, uint64_t dr,
                    unsigned rows_left, u128* nodes, u128* cands) {
  if (TRACK) *nodes += 1;
  uint64_t available = board_mask & ~(columns | dl | dr);
  if (rows_left == 1) {
    if (TRACK) *cands += available != 0;
    return available != 0;
  }
  if (TRACK) *cands += (unsigned)__builtin_popcountll(available);
  uint64_t count = 0;
  while (available) {
    uint64_t b = available & (~available + 1);
    available ^= b;
    count += dfs_search<TRACK>(board_mask, columns | b, ((dl | b) << 1) & board_mask, (dr | b) >> 1,
                               rows_left - 1, nodes, cands);
  }
  return count;
}

struct DfsTask {
  uint64_t columns, dl, dr;
  uint8_t weight, rows_placed;
};

}  // namespace

// count_dfs_bitmask_impl, src/dfs_bitmask.rs:229-338 (seed_tasks :111-138,
// split_once :140-157, make_tasks :159-198).
DfsResult count_dfs(unsigned n, unsigned threads, bool track) {
  if (n > 27) throw OracleError("DFS bitmask baseline supports N <= 27");
  if (threads == 0) throw OracleError("thread count must be positive");
  DfsResult res{0, 0, 1, 0, 0};
  if (n == 0) {
    res.count = 1;
    res.recursive_nodes = 1;
    return res;
  }
  uint64_t board_mask = ((uint64_t)1 << n) - 1;
  std::vector<DfsTask> tasks;
  for (unsigned c = 0; c < n / 2; ++c) {
    uint64_t b = (uint64_t)1 << c;
    tasks.push_back(DfsTask{b, b << 1, b >> 1, 2, 1});
  }
  if (n % 2 == 1) {
    uint64_t b = (uint64_t)1 << (n / 2);
    tasks.push_back(DfsTask{b, b << 1, b >> 1, 1, 1});
  }
  u128 nodes = 1, cands = tasks.size();
  size_t target = threads == 1 ? 1 : (size_t)threads * 64;
  unsigned split_depth = tasks.empty() ? 0 : tasks[0].rows_placed;
  while (tasks.size() < target && split_depth + 1 < n) {
    nodes += tasks.size();
    std::vector<DfsTask> children;
    for (const DfsTask& t : tasks) {
      uint64_t avail = board_mask & ~(t.columns | t.dl | t.dr);
      while (avail) {
        uint64_t b = avail & (~avail + 1);
        avail ^= b;
        children.push_back(DfsTask{t.columns | b, ((t.dl | b) << 1) & board_mask, (t.dr | b) >> 1,
                                   t.weight, (uint8_t)(t.rows_placed + 1)});
      }
    }
    tasks.swap(children);
    cands += tasks.size();
    split_depth++;
    if (tasks.empty()) break;
  }
  std::stable_sort(tasks.begin(), tasks.end(), [&](const DfsTask& a, const DfsTask& b) {
    return __builtin_popcountll(board_mask & ~(a.columns | a.dl | a.dr)) >
           __builtin_popcountll(board_mask & ~(b.columns | b.dl | b.dr));
  });
  res.split_depth = split_depth;
  res.tasks = tasks.size();
  unsigned workers = std::max<unsigned>(1, std::min<size_t>(threads, tasks.size()));
  std::atomic<size_t> next{0};
  std::vector<u128> sub(workers, 0), wn(workers, 0), wc(workers, 0);
  auto run = [&](unsigned w) {
    for (;;) {
      size_t i = next.fetch_add(1);
      if (i >= tasks.size()) break;
      const DfsTask& t = tasks[i];
      unsigned rows_left = n - t.rows_placed;
      uint64_t raw;
      if (rows_left == 0)
        raw = 1;
      else if (track)
        raw = dfs_search<true>(board_mask, t.columns, t.dl, t.dr, rows_left, &wn[w], &wc[w]);
      else
        raw = dfs_search<false>(board_mask, t.columns, t.dl, t.dr, rows_left, &wn[w], &wc[w]);
      sub[w] += (u128)raw * t.weight;
    }
  };
  std::vector<std::thread> pool;
  for (unsigned w = 1; w < workers; ++w) pool.emplace_back(run, w);
  run(0);
  for (auto& th : pool) th.join();
  for (unsigned w = 0; w < workers; ++w) {
    res.count += sub[w];
    nodes += wn[w];
    cands += wc[w];
  }
  if (track) {
    res.recursive_nodes = nodes;
    res.candidate_placements = cands;
  }
  return res;
}

// The boundary after the prefix: per placed queen in column c of the current
// row, the compiled successor (src/lib.rs:622-635).  False if it attacks.
bool prefix_boundary(unsigned n, const uint32_t* cols, unsigned k, Boundary* out) {
  uint64_t board_mask = ((uint64_t)1 << n) - 1;
  Boundary s{0, 0, 0};
  for (unsigned i = 0; i < k; ++i) {
    if (cols[i] >= n) throw OracleError("prefix column outside the board");
    uint64_t b = (uint64_t)1 << cols[i];
    if ((s.columns | s.diag_dr | s.diag_dl) & b) return false;
    s = Boundary{s.columns | b, ((s.diag_dr | b) << 1) & board_mask, (s.diag_dl | b) >> 1};
  }
  *out = s;
  return true;
}

// DFS completion count of a prefix (src/dfs_bitmask.rs:67-109 from the
// prefix's attack masks).
u128 count_prefix_dfs(unsigned n, const uint32_t* cols, unsigned k) {
  if (k > n) throw OracleError("prefix longer than the board");
  if (n == 0) return 1;
  Boundary s;
  if (!prefix_boundary(n, cols, k, &s)) return 0;
  unsigned rows_left = n - k;
  if (rows_left == 0) return 1;
  uint64_t board_mask = ((uint64_t)1 << n) - 1;
  u128 a = 0, b = 0;
  // diag_dr of the boundary shifts left each row like the DFS's diag_left.
  return dfs_search<false>(board_mask, s.columns, s.diag_dr, s.diag_dl, rows_left, &a, &b);
}

uint64_t peak_rss_bytes() {
  struct rusage ru;
  if (getrusage(RUSAGE_SELF, &ru) != 0) return 0;
  return (uint64_t)ru.ru_maxrss * 1024ull;
}

std::string u128_to_string(u128 v) {
  if (v == 0) return "0";
  std::string s;
  while (v) {
    s.push_back((char)('0' + (int)(v % 10)));
    v /= 10;
  }
  std::reverse(s.begin(), s.end());
  return s;
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// extern "C" surface for the Python tests (ctypes) and bench.py's CPU leg.
// ---------------------------------------------------------------------------
namespace {
int fail(const std::exception& e, char* err, size_t cap) {
  if (err && cap) {
    std::strncpy(err, e.what(), cap - 1);
    err[cap - 1] = 0;
  }
  return 1;
}
}  // namespace

extern "C" {

// backend: 0 sort-reduce, 1 hash (compiled), 2 sitewise (hash), 3 parallel sort-reduce.
int oracle_contract_rows(uint32_t n, uint32_t symmetry, uint32_t sparse, uint32_t backend,
                         uint32_t threads, uint32_t linear_scan_counters,
                         peps_contraction_result* out, char* err, size_t cap) {
  try {
    oracle::Options o;
    o.symmetry = symmetry ? oracle::Symmetry::TopRowOrbits : oracle::Symmetry::None;
    o.position = sparse ? oracle::Position::Sparse : oracle::Position::Dense;
    o.backend = (oracle::Backend)backend;
    o.threads = threads;
    o.linear_scan_counters = linear_scan_counters != 0;
    oracle::contract_rows(n, o, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, cap);
  }
}

// Row contraction from a k-row prefix (no orbit slicing), sort-reduce.
int oracle_contract_prefix(uint32_t n, const uint32_t* cols, uint32_t k, uint32_t threads,
                           peps_contraction_result* out, char* err, size_t cap) {
  try {
    oracle::Options o;
    o.symmetry = oracle::Symmetry::None;
    o.backend = threads > 1 ? oracle::Backend::ParallelSortReduce : oracle::Backend::SortReduce;
    o.threads = threads ? threads : 1;
    oracle::Boundary s;
    if (k > n) throw oracle::OracleError("prefix longer than the board");
    if (!oracle::prefix_boundary(n, cols, k, &s)) {
      std::memset(out, 0, sizeof(*out));
      out->n = n;
      return 0;
    }
    if (k == n || n == 0) {
      std::memset(out, 0, sizeof(*out));
      out->n = n;
      out->count = peps_u128{1, 0};
      return 0;
    }
    o.has_start = true;
    o.start = s;
    o.start_row = k;
    oracle::contract_rows(n, o, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, cap);
  }
}

int oracle_known_count(uint32_t n, peps_u128* out) {
  oracle::u128 v;
  if (!oracle::known_count(n, &v)) return 5;
  *out = peps_u128{(uint64_t)v, (uint64_t)(v >> 64)};
  return 0;
}

int oracle_crt_plan(uint32_t n, uint64_t* primes, uint32_t* count, peps_u128* product,
                    peps_u128* factorial, char* err, size_t cap) {
  try {
    oracle::CrtPlan p = oracle::wide_crt_plan(n);
    for (size_t i = 0; i < p.primes.size(); ++i) primes[i] = p.primes[i];
    *count = (uint32_t)p.primes.size();
    *product = peps_u128{(uint64_t)p.modulus_product, (uint64_t)(p.modulus_product >> 64)};
    *factorial = peps_u128{(uint64_t)p.factorial_bound, (uint64_t)(p.factorial_bound >> 64)};
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, cap);
  }
}

int oracle_crt_reconstruct(const uint64_t* residues, const uint64_t* primes, uint32_t count,
                           peps_u128* out, char* err, size_t cap) {
  try {
    std::vector<uint64_t> r(residues, residues + count), p(primes, primes + count);
    oracle::u128 v = oracle::reconstruct_crt(r, p);
    *out = peps_u128{(uint64_t)v, (uint64_t)(v >> 64)};
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, cap);
  }
}

int oracle_is_prime(uint64_t v) { return oracle::is_prime_u32_range(v) ? 1 : 0; }

int oracle_prefix_tasks(uint32_t n, uint32_t depth, uint64_t* tasks, uint64_t cap_tasks,
                        uint64_t* count) {
  auto t = oracle::prefix_tasks(n, depth);
  *count = t.size();
  for (uint64_t i = 0; i < t.size() && i < cap_tasks; ++i) {
    tasks[4 * i + 0] = t[i].state.columns;
    tasks[4 * i + 1] = t[i].state.diag_dr;
    tasks[4 * i + 2] = t[i].state.diag_dl;
    tasks[4 * i + 3] = t[i].orbit_weight;
  }
  return 0;
}

int oracle_dfs(uint32_t n, uint32_t threads, uint32_t track, peps_u128* count, peps_u128* nodes,
               peps_u128* cands, uint32_t* split_depth, uint64_t* ntasks, char* err, size_t cap) {
  try {
    oracle::DfsResult r = oracle::count_dfs(n, threads, track != 0);
    *count = peps_u128{(uint64_t)r.count, (uint64_t)(r.count >> 64)};
    *nodes = peps_u128{(uint64_t)r.recursive_nodes, (uint64_t)(r.recursive_nodes >> 64)};
    *cands = peps_u128{(uint64_t)r.candidate_placements, (uint64_t)(r.candidate_placements >> 64)};
    *split_depth = r.split_depth;
    *ntasks = r.tasks;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, cap);
  }
}

int oracle_count_prefix_dfs(uint32_t n, const uint32_t* cols, uint32_t k, peps_u128* out,
                            char* err, size_t cap) {
  try {
    oracle::u128 v = oracle::count_prefix_dfs(n, cols, k);
    *out = peps_u128{(uint64_t)v, (uint64_t)(v >> 64)};
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, cap);
  }
}

// Tensor tables for the tensor-level tests: B as 17 x (alpha, 8 legs, value),
// C as 17 x (8 legs, value), and the 16 input buckets' sizes.
int oracle_site_tensors(uint8_t* b_rows, uint8_t* c_rows, uint32_t* bucket_sizes) {
  auto b = oracle::site_tensor_b();
  for (size_t i = 0; i < b.size(); ++i) {
    const auto& l = b[i].legs;
    uint8_t row[10] = {b[i].alpha, l.column_in, l.column_out, l.row_in, l.row_out,
                       l.diag_dr_in, l.diag_dr_out, l.diag_dl_in, l.diag_dl_out,
                       (uint8_t)b[i].value};
    std::memcpy(b_rows + 10 * i, row, 10);
  }
  auto c = oracle::SiteTensorC::sec_vi();
  for (size_t i = 0; i < c.entries.size(); ++i) {
    const auto& l = c.entries[i].legs;
    uint8_t row[9] = {l.column_in, l.column_out, l.row_in, l.row_out, l.diag_dr_in,
                      l.diag_dr_out, l.diag_dl_in, l.diag_dl_out, (uint8_t)c.entries[i].value};
    std::memcpy(c_rows + 9 * i, row, 9);
  }
  for (unsigned k = 0; k < 16; ++k) bucket_sizes[k] = (uint32_t)c.by_input[k].size();
  return (int)c.entries.size();
}

// One-row transfer of a single parent through the sitewise and compiled
// operators (used by the operator-equivalence test): writes up to `cap`
// successors as (cols, dr, dl, weight_lo) for each operator.
int oracle_row_transfer(uint32_t n, uint64_t cols, uint64_t dr, uint64_t dl, uint32_t sitewise,
                        uint64_t* outs, uint32_t cap, uint32_t* count) {
  try {
    auto c = oracle::SiteTensorC::sec_vi();
    auto op = oracle::CompiledRowOperator::compile(c);
    oracle::RowCounters k;
    auto terms = sitewise ? oracle::contract_one_row_sitewise(n, c, {cols, dr, dl}, 1, k)
                          : oracle::contract_one_row_compiled(n, op, {cols, dr, dl}, 1, k);
    *count = (uint32_t)terms.size();
    for (uint32_t i = 0; i < terms.size() && i < cap; ++i) {
      outs[4 * i + 0] = terms[i].first.columns;
      outs[4 * i + 1] = terms[i].first.diag_dr;
      outs[4 * i + 2] = terms[i].first.diag_dl;
      outs[4 * i + 3] = (uint64_t)terms[i].second;
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

uint64_t oracle_peak_rss_bytes(void) { return oracle::peak_rss_bytes(); }

// Sitewise contraction of a caller-supplied C (peps_tensor_entry layout);
// layer(user, row, count, states) receives every merged layer as count
// quintuples (columns, diag_dr, diag_dl, weight_lo, weight_hi) in no order.
// c == NULL: Sec. VI's C.  backend: 0 sort-reduce (compiled), 2 sitewise.
typedef void (*oracle_layer_fn)(void* user, uint32_t row, uint64_t count, const uint64_t* states);

int oracle_contract_rows_tensor(uint32_t n, uint32_t symmetry, uint32_t backend, const peps_tensor_entry* c,
                                uint32_t count, uint32_t linear_scan_counters, oracle_layer_fn layer,
                                void* user, peps_contraction_result* out, char* err, size_t cap) {
  try {
    oracle::Options o;
    o.symmetry = symmetry ? oracle::Symmetry::TopRowOrbits : oracle::Symmetry::None;
    o.backend = backend == 2 ? oracle::Backend::Sitewise : oracle::Backend::SortReduce;
    o.linear_scan_counters = linear_scan_counters != 0;
    oracle::SiteTensorC custom;
    if (c) {
      std::vector<oracle::CEntry> entries;
      for (uint32_t i = 0; i < count; ++i)
        entries.push_back(oracle::CEntry{
            oracle::Legs{c[i].column_in, c[i].column_out, c[i].row_in, c[i].row_out, c[i].diag_dr_in,
                         c[i].diag_dr_out, c[i].diag_dl_in, c[i].diag_dl_out},
            (oracle::u128)c[i].value});
      custom = oracle::SiteTensorC::from_entries(entries);
      o.custom_tensor = &custom;
    }
    if (layer) {
      o.on_layer = [&](unsigned row, const std::vector<std::pair<oracle::u128, oracle::u128>>& recs) {
        std::vector<uint64_t> flat;
        flat.reserve(recs.size() * 5);
        for (const auto& r : recs) {
          oracle::Boundary b = oracle::unpack(r.first, n);
          flat.push_back(b.columns);
          flat.push_back(b.diag_dr);
          flat.push_back(b.diag_dl);
          flat.push_back((uint64_t)r.second);
          flat.push_back((uint64_t)(r.second >> 64));
        }
        layer(user, row, recs.size(), flat.data());
      };
    }
    oracle::contract_rows(n, o, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, cap);
  }
}

}  // extern "C"
