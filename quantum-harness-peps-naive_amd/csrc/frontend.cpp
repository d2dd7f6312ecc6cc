// frontend.cpp -- option resolution, prefix tasks and sliced launches.
#include "frontend.h"

#include "split_retry.h"

#include <sys/resource.h>

#include <algorithm>
#include <cstring>

namespace peps {

int fail(int code, const std::string& msg, char* err, size_t cap) {
  if (err && cap) {
    const size_t n = std::min(cap - 1, msg.size());
    std::memcpy(err, msg.data(), n);
    err[n] = 0;
  }
  return code;
}

uint64_t rss_bytes() {
  struct rusage ru;
  if (getrusage(RUSAGE_SELF, &ru) != 0) return 0;
  return (uint64_t)ru.ru_maxrss * 1024ull;
}

void fill_result(unsigned n, const RunOut& ro, const std::vector<uint64_t>& primes, double elapsed,
                 peps_contraction_result* out, TensorCounters tc, uint64_t c_entries) {
  out->n = n;
  out->count = abi(ro.count);
  out->elapsed_s = elapsed;
  out->peak_states = ro.peak_states;
  u128 examined = tc == TensorCounters::kCompiled ? 17 : 0, matched = examined;
  out->row_operator_candidates = abi(ro.total_candidates);
  out->row_operator_matched = abi(ro.total_matched);
  out->layer_count = (uint32_t)std::min<size_t>(ro.layers.size(), PEPS_MAX_N);
  for (uint32_t i = 0; i < out->layer_count; ++i) {
    const LayerOut& l = ro.layers[i];
    peps_layer_metric& m = out->layers[i];
    m.row = l.row;
    m.input_states = l.input_states;
    u128 le = 0, lm = 0;
    if (tc != TensorCounters::kCompiled) {
      lm = l.site_matched;
      le = tc == TensorCounters::kSitewiseLinear ? l.site_visits * (u128)c_entries : l.site_matched;
    }
    examined += le;
    matched += lm;
    m.tensor_entries_examined = abi(le);
    m.tensor_entries_matched = abi(lm);
    m.row_operator_candidates = abi(l.candidates);
    m.row_operator_matched = abi(l.matched);
    m.completed_row_terms = abi(l.completed);
    m.output_states = l.output_states;
    m.output_weight = abi(l.output_weight);
    m.elapsed_s = l.elapsed_s;
    m.peak_rss_bytes = rss_bytes();
    m.device_bytes = l.device_bytes;
    m.partition_levels = l.levels;
    m.final_buckets = l.buckets;
  }
  out->tensor_entries_examined = abi(examined);
  out->tensor_entries_matched = abi(matched);
  out->crt_lanes = (uint32_t)primes.size();
  out->record_bytes = ro.record_bytes;
  for (size_t l = 0; l < primes.size() && l < PEPS_MAX_LANES; ++l) {
    out->primes[l] = primes[l];
    out->residues[l] = l < ro.residues.size() ? ro.residues[l] : 0;
  }
  out->device_peak_bytes = ro.device_peak_bytes;
  out->kernel_s = ro.kernel_s;
  out->alloc_s = ro.alloc_s;
  out->peak_rss_bytes = rss_bytes();
}

int resolve_options(unsigned n, const peps_options* o, Resolved* r, std::string* err) {
  u128 prod, fact;
  if (!crt_plan(n, &r->primes, &prod, &fact, err)) return PEPS_E_INVALID;
  if (!o) return PEPS_OK;
  if (o->crt_lanes) {
    if (o->crt_lanes > PEPS_MAX_LANES) {
      *err = "crt_lanes must be at most 4";
      return PEPS_E_INVALID;
    }
    if (o->crt_lanes < r->primes.size()) {
      *err = "crt_lanes below the certified plan (product of primes must exceed N!)";
      return PEPS_E_INVALID;
    }
    r->primes.assign(crt_primes(), crt_primes() + o->crt_lanes);
  }
  if (o->radix_bits) {
    if (o->radix_bits < 1 || o->radix_bits > 11) {
      *err = "radix_bits must be in 1..11";
      return PEPS_E_INVALID;
    }
    r->radix_bits = o->radix_bits;
  }
  if (o->tile_states) {
    if (o->tile_states < 256 || o->tile_states > 2560) {
      *err = "tile_states must be in 256..2560";
      return PEPS_E_INVALID;
    }
    r->tile_states = o->tile_states;
  }
  if (o->expand_bits) {
    if (o->expand_bits > 10) {
      *err = "expand_bits must be in 1..10";
      return PEPS_E_INVALID;
    }
    r->expand_bits = o->expand_bits;
  }
  r->device = o->device;
  r->budget = o->hbm_budget_bytes;
  r->graph = o->graph_mode != 0;
  r->slice_group = o->slice_group;
  if (o->merge_passes) {
    if (o->merge_passes > 8) {
      *err = "merge_passes must be in 1..8";
      return PEPS_E_INVALID;
    }
    r->merge_passes = o->merge_passes;
  }
  if (o->merge_waves) {
    if (o->merge_waves > 8) {
      *err = "merge_waves must be in 1..8";
      return PEPS_E_INVALID;
    }
    r->merge_waves = o->merge_waves;
  }
  if (o->site_counters > 1) {
    *err = "site_counters must be 0 (indexed) or 1 (linear scan)";
    return PEPS_E_INVALID;
  }
  r->site_counters = o->site_counters;
  const uint32_t fmt = o->record_format & 0xffu, wbits = (o->record_format >> 8) & 0xffu;
  if (fmt > PEPS_RECORDS_PACKED || (o->record_format >> 20)) {
    *err = "record_format must be 0 (auto), 1 (lane records) or 2 (packed), with a packed weight width in "
           "bits 8..15 and fallback-path bits in 16..19";
    return PEPS_E_INVALID;
  }
  r->force = (o->record_format >> 16) & 0xfu;
  r->records = (int)fmt;
  r->packed_weight_bits = wbits ? wbits : kPackedMaxWeightBits;
  return PEPS_OK;
}

int check_board(unsigned n, std::string* err) {
  if (n > 42) {
    *err = "the packed u128 virtual-boundary backend supports N <= 42";
    return PEPS_E_INVALID;
  }
  if (n > PEPS_MAX_DEVICE_N) {
    *err = "the MI355X boundary key (u64, 3N bits) supports N <= 21";
    return PEPS_E_INVALID;
  }
  return PEPS_OK;
}

void apply_resolved(const Resolved& r, RunSpec* spec) {
  spec->primes = r.primes;
  spec->radix_bits = r.radix_bits;
  spec->tile_states = r.tile_states;
  spec->expand_bits = r.expand_bits;
  spec->use_graph = r.graph;
  spec->merge_passes = r.merge_passes;
  spec->merge_waves = r.merge_waves;
  spec->budget_bytes = r.budget;
  spec->records = r.records;
  spec->packed_weight_bits = r.packed_weight_bits;
  spec->force = r.force;
}

int run_engine(const Resolved& r, RunSpec& spec, RunOut* ro, std::string* err) {
  Engine* eng = Engine::get((int)r.device, err);
  if (!eng) return PEPS_E_DEVICE;
  apply_resolved(r, &spec);
  std::lock_guard<std::mutex> lk(eng->mutex());
  return eng->run(spec, ro, err);
}

bool prefix_state(unsigned n, const uint32_t* cols, unsigned k, uint64_t* key) {
  const uint64_t bm = ((uint64_t)1 << n) - 1;
  uint64_t c = 0, dr = 0, dl = 0;
  for (unsigned i = 0; i < k; ++i) {
    const uint64_t b = (uint64_t)1 << cols[i];
    if ((c | dr | dl) & b) return false;
    c |= b;
    dr = ((dr | b) << 1) & bm;
    dl = (dl | b) >> 1;
  }
  *key = c | (dr << n) | (dl << (2 * n));
  return true;
}

namespace {

// One row of the compiled relation (src/lib.rs:594-638) applied to a task,
// with the row-0 orbit weights (src/lib.rs:721-749) when `orbit` is set.
void expand_row(unsigned n, const Task& t, bool orbit, std::vector<Task>* out) {
  const uint64_t bm = ((uint64_t)1 << n) - 1;
  const uint64_t c = t.key & bm, dr = (t.key >> n) & bm, dl = (t.key >> (2 * n)) & bm;
  uint64_t free = ~(c | dr | dl) & bm;
  while (free) {
    const unsigned col = (unsigned)__builtin_ctzll(free);
    free &= free - 1;
    const uint64_t b = (uint64_t)1 << col;
    u128 w = t.weight;
    if (orbit) {
      const unsigned mirror = n - 1 - col;
      if (col > mirror) continue;
      if (col != mirror) w *= 2;
    }
    const uint64_t nc = c | b, ndr = ((dr | b) << 1) & bm, ndl = (dl | b) >> 1;
    out->push_back(Task{nc | (ndr << n) | (ndl << (2 * n)), w});
  }
}

uint64_t extensions(unsigned n, uint64_t bm, uint64_t cols, uint64_t dr, uint64_t dl, unsigned rows) {
  if (rows == 0) return 1;
  uint64_t free = ~(cols | dr | dl) & bm, total = 0;
  while (free) {
    const uint64_t b = free & (~free + 1);
    free ^= b;
    total += extensions(n, bm, cols | b, ((dr | b) << 1) & bm, (dl | b) >> 1, rows - 1);
  }
  return total;
}

}  // namespace

std::vector<Task> make_tasks(unsigned n, unsigned depth) {
  std::vector<Task> tasks{Task{0, 1}};
  for (unsigned d = 0; d < depth && d < n; ++d) {
    std::vector<Task> next;
    for (const Task& t : tasks) expand_row(n, t, d == 0, &next);
    tasks.swap(next);
  }
  return tasks;
}

std::vector<Task> expand_task(unsigned n, const Task& t, unsigned from, unsigned to) {
  std::vector<Task> tasks{t};
  for (unsigned d = from; d < to && d < n; ++d) {
    std::vector<Task> next;
    for (const Task& x : tasks) expand_row(n, x, false, &next);
    tasks.swap(next);
  }
  return tasks;
}

uint64_t extension_count(unsigned n, uint64_t key, unsigned rows) {
  const uint64_t bm = ((uint64_t)1 << n) - 1;
  return extensions(n, bm, key & bm, (key >> n) & bm, (key >> (2 * n)) & bm, rows);
}

double Projection::proxy(unsigned n, unsigned depth, uint64_t key) const {
  const unsigned rows = std::min(3u, n > depth ? n - depth : 0u);
  return (double)std::max<uint64_t>(1, extension_count(n, key, rows));
}

double Projection::ratio(unsigned n, unsigned depth) const {
  auto it = ratio_.find({n, depth});
  return it == ratio_.end() ? 0.0 : it->second;
}

void Projection::learn(unsigned n, unsigned depth, double proxy_sum, const RunOut& ro) {
  // The largest candidate buffer the run needed: the exact per-row counts,
  // or the upper bounds a SCATTER wrote into.
  u128 cmax = ro.peak_write_bound;
  for (const auto& l : ro.layers) cmax = std::max(cmax, l.completed);
  if (proxy_sum <= 0) return;
  double& r = ratio_[{n, depth}];
  r = std::max(r, (double)cmax / proxy_sum);
}

void Projection::exceeded(unsigned n, unsigned depth, double proxy_sum, double capacity) {
  if (proxy_sum <= 0 || capacity <= 0) return;
  double& r = ratio_[{n, depth}];
  r = std::max(r, capacity / proxy_sum);
}

size_t max_slices(unsigned n, uint32_t slice_group) {
  const unsigned free_bits = std::min(63u - 3 * n, 20u);
  return std::min<size_t>((size_t)1 << free_bits, slice_group ? slice_group : SIZE_MAX);
}

RunSpec task_spec(unsigned n, unsigned depth, const std::vector<Task>& tasks, size_t a, size_t b) {
  RunSpec spec;
  spec.n = n;
  spec.first_row = depth;
  spec.orbit_first_row = depth == 0;
  unsigned bits = 0;
  while (((size_t)1 << bits) < b - a) ++bits;
  spec.slice_bits = bits;
  spec.per_slice = true;
  for (size_t i = a; i < b; ++i)
    spec.seeds.push_back(Seed{tasks[i].key | ((uint64_t)(i - a) << (3 * n)), tasks[i].weight});
  return spec;
}

int run_task_groups(const Resolved& r, unsigned n, unsigned depth, const std::vector<Task>& tasks,
                    const std::vector<std::pair<size_t, size_t>>& groups, const std::string& key,
                    const std::function<void(size_t, size_t, const RunOut&)>& done, std::string* err) {
  Engine* eng = Engine::get((int)r.device, err);
  if (!eng) return PEPS_E_DEVICE;
  std::vector<RunSpec> specs;
  for (const auto& [a, b] : groups) {
    specs.push_back(task_spec(n, depth, tasks, a, b));
    apply_resolved(r, &specs.back());
  }
  std::vector<RunOut> outs;
  {
    std::lock_guard<std::mutex> lk(eng->mutex());
    const int rc = eng->run_many(key, specs, &outs, err);
    if (rc) return rc;
  }
  for (size_t i = 0; i < groups.size(); ++i) done(groups[i].first, groups[i].second, outs[i]);
  return PEPS_OK;
}

int run_task_range(const Resolved& r, unsigned n, unsigned depth, const std::vector<Task>& tasks,
                   size_t g0, size_t g1,
                   const std::function<void(size_t, size_t, const RunOut&)>& done,
                   std::string* err, const std::function<void(size_t, size_t)>& exceeded) {
  auto run = [&](size_t a, size_t b) -> int {
    RunSpec spec = task_spec(n, depth, tasks, a, b);
    RunOut ro;
    std::string e;
    const int rc = run_engine(r, spec, &ro, &e);
    if (rc == PEPS_OK)
      done(a, b, ro);
    else
      *err = e;
    return rc;
  };
  return split_retry(g0, g1, run, exceeded);
}

}  // namespace peps
