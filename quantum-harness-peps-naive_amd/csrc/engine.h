// engine.h -- device-resident row contraction driver (one per device).
//
// The engine owns a HIP stream and an HBM arena of two candidate buffers
// whose roles alternate every row.  The seed frontier (SoA: u64 keys, then
// the exact u64 weights for packed records or one u32 residue array per CRT
// prime for lane records) is expanded into the first row's hash-radix
// buckets; from then on each row's buckets are merged and expanded straight
// into the next row's buckets (merge_expand), so the frontier is never
// materialised in HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "host_math.h"
#include "kernels.h"

namespace peps {

struct Seed {
  uint64_t key;  // packed boundary, slice id above bit 3n
  u128 weight;
};

// Candidate record format of the fused pipeline.
enum RecordFormat : int {
  kRecordsAuto = 0,    // packed 12-byte records; lane records if a weight outgrows them
  kRecordsLanes = 1,   // key + one u32 residue per CRT prime (16 or 24 bytes)
  kRecordsPacked = 2,  // packed only: a weight overflow fails the run
};

// Fallback paths a run can be made to take (the results never change):
// exact histogram partition levels, SCATTER runs sized by the upper-bound
// count, fixed-capacity buckets too small (every such level overflows and is
// repeated exactly), caps too small (capped runs overflow and the SCATTER is
// repeated with upper-bound offsets).
enum ForcePath : uint32_t {
  kForceExactLevels = 1,
  kForceUpperRuns = 2,
  kForceLevelOverflow = 4,
  kForceCapOverflow = 8,
};

struct RunSpec {
  unsigned n = 0;
  unsigned first_row = 0;          // contract rows [first_row, n)
  bool orbit_first_row = false;    // row-0 vertical-reflection orbit weights
  std::vector<Seed> seeds;         // frontier entering first_row (distinct keys)
  unsigned slice_bits = 0;         // slice ids occupy bits [3n, 3n + slice_bits)
  std::vector<uint64_t> primes;    // CRT lanes (1..4)
  uint32_t tile_states = 2048;     // target records per LDS merge bucket
  uint32_t radix_bits = 8;         // max hash-radix digit width per partition level
  uint32_t expand_bits = 10;       // max digit width written by the merge-expand scatter
  uint32_t merge_passes = 1;       // target bucket = merge_passes x one LDS table load
  uint32_t merge_waves = 1;        // merge grid = CUs x resident workgroups x merge_waves
  uint64_t budget_bytes = 0;       // arena cap (0: 85% of free HBM at first use)
  bool per_slice = false;          // report a count per slice id
  bool use_graph = false;          // record the size plan; replay repeats as one HIP graph
  // Materialised mode: every row's merged frontier is written to HBM (keys and
  // residues) and expanded from there, through the compiled relation or, when
  // `site` is set, the literal sitewise C tensor.
  bool materialize = false;
  const SiteTable* site = nullptr;
  bool bound_check = true;         // per-row sums and the count must not exceed N!
  int records = kRecordsAuto;      // fused pipeline record format (materialised: lanes)
  uint32_t packed_weight_bits = kPackedMaxWeightBits;  // narrower packed weights (tests)
  uint32_t force = 0;              // kForce* bits: take the fallback paths (tests)
  // Materialised mode: called after every row with the merged layer
  // (keys[S], residues[l * S + i]); a nonzero return aborts the run.
  std::function<int(unsigned row, const std::vector<uint64_t>& keys, const std::vector<uint32_t>& res,
                    std::string* err)>
      on_layer;
};

struct LayerOut {
  unsigned row;
  uint64_t input_states;
  u128 candidates;  // dense row operator candidates, n per parent
  u128 matched;     // legal (parent, column) pairs before the orbit filter
  u128 completed;   // successor terms after the orbit filter
  uint64_t output_states;
  u128 output_weight;
  double elapsed_s;
  uint64_t device_bytes;
  uint32_t levels;
  uint32_t buckets;
  u128 site_visits = 0;   // literal sitewise rows: partial-row visits
  u128 site_matched = 0;  // literal sitewise rows: C entries applied
};

struct RunOut {
  u128 count = 0;
  std::vector<uint64_t> residues;
  std::vector<u128> slice_counts;  // per slice id when slice_bits > 0
  std::vector<LayerOut> layers;
  uint64_t peak_states = 1;
  u128 total_candidates = 0, total_matched = 0;
  double elapsed_s = 0, kernel_s = 0;
  uint64_t device_peak_bytes = 0;
  uint64_t peak_write_bound = 0;   // most records one SCATTER output needed (>= the exact count)
  bool replayed = false;           // the run was a HIP-graph replay of a recorded plan
  uint32_t record_bytes = 0;       // candidate record size the run used
  double alloc_s = 0;              // host time spent growing device buffers
};

// Data-dependent sizes of one run, recorded by an eager run (which reads them
// back from the device row by row).  The contraction is deterministic, so a
// repeat of the same spec has the same sizes and its whole row pipeline can be
// captured once into a HIP graph and replayed without host synchronisation.
struct RowPlan {
  // How the SCATTER finds its write positions: exact (COUNT merge + scan),
  // capped runs from the fanout sums, or runs sized by the upper-bound
  // count of every record's successors.  Capped and upper-bound runs leave
  // gaps, which the next row's first partition level reads in run mode.
  enum Mode : int { kExact = 0, kCapped = 1, kUpper = 2 };
  int mode = kExact;
  uint64_t C_next = 0;             // candidates of the next row (0: none)
  uint64_t C_count = 0;            // records the SCATTER's layout spans (C_next, or more with gaps)
  uint32_t W = 0;                  // merge-expand workgroups for this row
  uint32_t d1 = 0;                 // next row's level-1 digit width
  std::vector<unsigned> pdigits;   // next row's partition digit widths
  std::vector<uint64_t> pcap;      // their bucket capacities (0: exact histogram level)
};

struct Plan {
  uint64_t C_first = 0;
  u128 matched_first = 0;
  uint32_t d1_first = 0;
  std::vector<unsigned> pdigits_first;
  std::vector<uint64_t> pcap_first;
  uint64_t tiles = 1, per_tile = 1;
  std::vector<RowPlan> rows;       // rows [first_row, n)
  size_t region_need = 0, hist_need = 0, offsets_need = 0, partial_need = 0;
  size_t segs_need[2] = {0, 0};
  size_t run_start_need = 0, run_len_need = 0, fan_need = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t generation = 0;         // buffer generation the graph was captured against
};

struct KernelStat {
  uint64_t launches = 0;
  double seconds = 0;
  double bytes = 0;
};

// Device bytes of the no-allocation engine's scratch array (one per device).
constexpr size_t kScratchBytes = (size_t)256 << 20;

class Engine {
 public:
  static Engine* get(int device, std::string* err);  // one engine per device
  // The no-allocation engine of `device`: every device buffer is a fixed
  // slice of the library's static scratch array, kernels run on the device's
  // null stream, and nothing is allocated on the device or as pinned host
  // memory, at set-up or later.  A row that does not fit its slice fails
  // with PEPS_E_MEMORY.  Fused pipeline only (no graphs, no slices, no
  // materialised layers, no kernel timing).
  static Engine* get_scratch(int device, std::string* err);
  // A run in kRecordsAuto format whose weights outgrow the packed records is
  // repeated with lane records (same result, larger records).
  int run(const RunSpec& spec, RunOut* out, std::string* err);
  // Runs the fused pipelines of `specs` one after another.  The first call
  // with a given `key` runs them eagerly and records every size plan; later
  // calls with the same key (and the same specs) replay all of them as one
  // HIP graph, each group with its own seeds, counters and slice sums.
  int run_many(const std::string& key, const std::vector<RunSpec>& specs, std::vector<RunOut>* outs,
               std::string* err);
  std::mutex& mutex() { return mu_; }
  int device() const { return device_; }
  // Records one candidate buffer can hold (0 before the first run sets the
  // budget), in the record format `records` selects.
  uint64_t candidate_capacity(int lanes, int records) const;
  // Kernel timing with HIP events on the engine stream (off by default).
  void set_profiling(bool on);
  std::vector<std::pair<std::string, KernelStat>> kernel_stats() const;

 private:
  explicit Engine(int device) : device_(device) {}
  int init(std::string* err);
  int init_scratch(std::string* err);
  bool fixed_ = false;                  // buffers are fixed scratch slices
  uint64_t h_local_[kNumCounters] = {}; // host counter mirror of the scratch engine
  int ensure_region(int idx, size_t bytes, std::string* err);
  template <typename T>
  int ensure_buffer(T** ptr, size_t* cap, size_t count, std::string* err);
  uint64_t arena_bytes() const;
  uint64_t region_cap() const;     // bytes of a candidate region allocated at its final size

  int device_;
  std::mutex mu_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_start_ = nullptr, ev_stop_ = nullptr;
  int prepare(const RunSpec& spec, std::string* err);
  int run_once(const RunSpec& spec, RunOut* out, std::string* err);
  int run_many_once(const std::string& key, const std::vector<RunSpec>& specs, std::vector<RunOut>* outs,
                    std::string* err);
  int pipeline(const RunSpec& spec, Plan* plan, bool eager, RunOut* out, std::string* err);
  struct GroupBufs {
    uint64_t* seed = nullptr;
    uint64_t* counters = nullptr;
    uint64_t* slices = nullptr;
    size_t seed_words = 0, counter_words = 0, slice_words = 0;
  };
  struct BatchGraph {
    std::vector<Plan> plans;
    std::vector<GroupBufs> bufs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t generation = 0;
  };
  std::map<std::string, BatchGraph> batches_;
  int materialized(const RunSpec& spec, RunOut* out, std::string* err);
  int finish(const RunSpec& spec, const Plan& plan, RunOut* out, std::string* err);
  std::map<std::string, Plan> plans_;
  uint64_t generation_ = 1;        // bumped whenever a device buffer is reallocated
  double alloc_s_ = 0;             // cumulative host time in buffer growth
  uint64_t* slice_sums_ = nullptr;
  size_t slice_cap_ = 0;

  void* region_[2] = {nullptr, nullptr};
  size_t region_bytes_[2] = {0, 0};
  uint64_t* counters_ = nullptr;   // device, kNumCounters
  uint64_t* h_counters_ = nullptr; // pinned host mirror
  uint32_t* hist_ = nullptr;
  size_t hist_cap_ = 0;
  uint64_t* offsets_ = nullptr;
  size_t offsets_cap_ = 0;
  uint64_t* segs_[2] = {nullptr, nullptr};
  size_t segs_cap_[2] = {0, 0};
  uint64_t* partial_ = nullptr;
  size_t partial_cap_ = 0;
  uint64_t* run_start_ = nullptr;      // upper-bound rows: SCATTER run starts [D1][W]
  size_t run_start_cap_ = 0;
  uint32_t* run_len_ = nullptr;        // upper-bound rows: SCATTER run lengths [D1][W]
  size_t run_len_cap_ = 0;
  uint64_t* fan_ = nullptr;            // fanout sums U[W], caps [W], their prefix sums [W + 1]
  size_t fan_cap_ = 0;
  uint64_t* counters_rows_ = nullptr;  // per-row counters, (n + 1) * kNumCounters
  size_t counters_rows_cap_ = 0;
  uint64_t* seed_ = nullptr;           // seed frontier (SoA, seed_image layout)
  size_t seed_cap_ = 0;
  uint64_t* front_ = nullptr;          // materialised frontier (SoA), u64 units
  size_t front_cap_ = 0;
  uint64_t budget_ = 0;
  Legs occupied_{};
  uint32_t cu_count_ = 256;

  // profiling
  struct Pending {
    int kernel;
    double bytes;
    hipEvent_t a, b;
  };
  bool profiling_ = false;
  std::vector<hipEvent_t> event_pool_;
  std::vector<Pending> pending_;
  std::vector<std::string> kernel_names_;
  std::vector<KernelStat> kstats_;
  hipError_t mark_begin(hipEvent_t* ev);
  hipError_t mark_end(int kernel, double bytes, hipEvent_t begin);
  int kernel_id(const char* name);
  int resolve_pending(std::string* err);
};

}  // namespace peps
