// frontend.h -- host-side pieces shared by the C ABI entry points and the
// request server: option resolution, prefix states, prefix tasks, the
// frontier-size projection and the sliced launch of a group of tasks.
#pragma once

#include <stdint.h>

#include <functional>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/peps_mi355x.h"
#include "engine.h"
#include "host_math.h"

namespace peps {

// Copies `msg` into the caller's error buffer and returns `code`.
int fail(int code, const std::string& msg, char* err, size_t cap);
inline peps_u128 abi(u128 v) { return peps_u128{(uint64_t)v, (uint64_t)(v >> 64)}; }
// peak_rss_bytes (src/lib.rs:5786-5850): the process's maximum resident set.
uint64_t rss_bytes();

// How the tensor counters of a result are reported.
enum class TensorCounters {
  kCompiled,        // 17 / 17 for the compiled operator, 0 per layer (src/lib.rs:5323-5324)
  kSitewise,        // per layer: entries examined through the input index = applied entries
  kSitewiseLinear,  // per layer: every C entry scanned per partial-row visit (pre-index revision)
};

// ContractionResult (src/lib.rs:403-415) from an engine run.
void fill_result(unsigned n, const RunOut& ro, const std::vector<uint64_t>& primes, double elapsed,
                 peps_contraction_result* out, TensorCounters tc = TensorCounters::kCompiled,
                 uint64_t c_entries = 17);

// Share of one candidate buffer that the batched form and the server admit
// a launch's projected candidates into.  The rest keeps the margin that the
// fixed-capacity partition buckets need (about 1.8x the records of a row).
constexpr double kAdmitFraction = 0.55;

// peps_options after validation and defaults.
struct Resolved {
  std::vector<uint64_t> primes;
  uint32_t radix_bits = 8;
  uint32_t tile_states = 1024;
  uint32_t expand_bits = 5;
  uint32_t device = 0;
  uint64_t budget = 0;
  bool graph = false;
  uint32_t slice_group = 0;
  uint32_t merge_passes = 1;
  uint32_t merge_waves = 8;
  uint32_t site_counters = 0;
  int records = 0;                  // RecordFormat
  uint32_t packed_weight_bits = 52;
  uint32_t force = 0;               // ForcePath bits (record_format bits 16..19)
};

// Lane count: the reference plan (src/lib.rs:3859-3886) unless the caller
// asks for more lanes; fewer lanes than the plan would not certify exactness.
int resolve_options(unsigned n, const peps_options* o, Resolved* r, std::string* err);

// N > 42: the reference's text (src/lib.rs:5293); 22..42: the u64 key limit.
int check_board(unsigned n, std::string* err);

// Copies the resolved lanes and tuning into `spec`.
void apply_resolved(const Resolved& r, RunSpec* spec);

// Runs `spec` on the resolved device with the resolved tuning.
int run_engine(const Resolved& r, RunSpec& spec, RunOut* ro, std::string* err);

// Boundary after a k-row prefix; false if a queen attacks an earlier one.
bool prefix_state(unsigned n, const uint32_t* cols, unsigned k, uint64_t* key);

// One prefix task: the boundary after `depth` rows and its multiplicity.
struct Task {
  uint64_t key;
  u128 weight;
};

// build_wide_crt_tasks (src/lib.rs:3888-3956) for an explicit depth: expand
// the compiled relation row by row without merging, applying the orbit weight
// on row 0.
std::vector<Task> make_tasks(unsigned n, unsigned depth);

// Expands one boundary at row `from` to row `to` without merging and without
// the orbit filter (the rows above `from` are already placed).
std::vector<Task> expand_task(unsigned n, const Task& t, unsigned from, unsigned to);

// Number of ways to place `rows` more queens below a boundary: a cheap proxy
// (a few thousand nodes) for how large a task's frontier grows.
uint64_t extension_count(unsigned n, uint64_t key, unsigned rows);

// Candidates per proxy unit, learned per (N, start row) from the launches run
// so far (the maximum over rows of emitted candidates / Σ proxies).
class Projection {
 public:
  double proxy(unsigned n, unsigned depth, uint64_t key) const;
  double ratio(unsigned n, unsigned depth) const;  // 0: not yet calibrated
  void learn(unsigned n, unsigned depth, double proxy_sum, const RunOut& ro);
  // A launch of tasks with this proxy sum exceeded a candidate buffer of
  // `capacity` records: the ratio is at least capacity / proxy_sum.
  void exceeded(unsigned n, unsigned depth, double proxy_sum, double capacity);

 private:
  std::map<std::pair<unsigned, unsigned>, double> ratio_;
};

// Largest number of tasks one launch can carry as slices (ids above bit 3N).
size_t max_slices(unsigned n, uint32_t slice_group);

// The spec that runs tasks[a, b) as slices (slice id i - a above bit 3N).
RunSpec task_spec(unsigned n, unsigned depth, const std::vector<Task>& tasks, size_t a, size_t b);

// Runs the given task groups, each as the slices of one pipeline, through
// Engine::run_many: the first call with `key` records every group's size
// plan, later calls replay all groups as one HIP graph.
int run_task_groups(const Resolved& r, unsigned n, unsigned depth, const std::vector<Task>& tasks,
                    const std::vector<std::pair<size_t, size_t>>& groups, const std::string& key,
                    const std::function<void(size_t, size_t, const RunOut&)>& done, std::string* err);

// Runs tasks[g0, g1) as the slices of one launch starting at row `depth`; a
// launch that exceeds the HBM budget is split in halves.  `done` is called
// once per successful launch with its range and output; `slice_counts` of that
// output are per task of the range.
// `exceeded(a, b)` (may be empty) hears of every range that did not fit the
// HBM budget before it is split.
int run_task_range(const Resolved& r, unsigned n, unsigned depth, const std::vector<Task>& tasks,
                   size_t g0, size_t g1,
                   const std::function<void(size_t, size_t, const RunOut&)>& done,
                   std::string* err, const std::function<void(size_t, size_t)>& exceeded = nullptr);

}  // namespace peps
