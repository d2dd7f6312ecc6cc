// capi.cpp -- the C ABI of libpeps_mi355x.so (include/peps_mi355x.h).
//
// Mirrors the reference's library surface (src/lib.rs): contract_rows and its
// symmetry variants, known_count, peak_rss_bytes, the CRT plan and
// reconstruction, and the prefix-task builder; plus the prefix count and the
// batched (sliced) form.  Every contraction runs on the GPU engine; there is
// no CPU fallback: without a gfx950 device the calls fail with PEPS_E_DEVICE.
#include <sys/resource.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/peps_mi355x.h"
#include "engine.h"
#include "frontend.h"
#include "host_math.h"

using peps::u128;
using peps::abi;
using peps::fail;
using peps::fill_result;
using peps::rss_bytes;
using peps::Resolved;
using peps::check_board;
using peps::run_engine;

namespace {

peps_progress_fn g_progress = nullptr;
void* g_progress_user = nullptr;

}  // namespace

extern "C" {

int peps_contract_rows_with_modes(uint32_t n, uint32_t symmetry, const peps_options* opts,
                                  peps_contraction_result* out, char* err, size_t err_cap) {
  if (!out) return fail(PEPS_E_INVALID, "null result pointer", err, err_cap);
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  std::string e;
  if (symmetry > PEPS_SYMMETRY_TOP_ROW_ORBITS) return fail(PEPS_E_INVALID, "unknown symmetry mode", err, err_cap);
  if (n > 42) return fail(PEPS_E_INVALID, "the packed u128 virtual-boundary backend supports N <= 42", err, err_cap);
  if (n == 0) {  // src/lib.rs:5295-5307
    out->count = abi(1);
    out->peak_states = 1;
    out->peak_rss_bytes = rss_bytes();
    return PEPS_OK;
  }
  int rc = check_board(n, &e);
  if (rc) return fail(rc, e, err, err_cap);
  Resolved r;
  rc = peps::resolve_options(n, opts, &r, &e);
  if (rc) return fail(rc, e, err, err_cap);
  peps::RunSpec spec;
  spec.n = n;
  spec.first_row = 0;
  spec.orbit_first_row = symmetry == PEPS_SYMMETRY_TOP_ROW_ORBITS;
  spec.seeds.push_back(peps::Seed{0, 1});
  peps::RunOut ro;
  const auto t0 = std::chrono::steady_clock::now();
  rc = run_engine(r, spec, &ro, &e);
  if (rc) return fail(rc, e, err, err_cap);
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_result(n, ro, r.primes, el, out);
  return PEPS_OK;
}

int peps_contract_rows(uint32_t n, const peps_options* opts, peps_contraction_result* out,
                       char* err, size_t err_cap) {
  return peps_contract_rows_with_modes(n, PEPS_SYMMETRY_TOP_ROW_ORBITS, opts, out, err, err_cap);
}

int peps_count_prefix(uint32_t n, const uint32_t* cols, uint32_t k, const peps_options* opts,
                      peps_u128* count, char* err, size_t err_cap) {
  if (!count) return fail(PEPS_E_INVALID, "null count pointer", err, err_cap);
  *count = abi(0);
  std::string e;
  if (k > n) return fail(PEPS_E_INVALID, "prefix longer than the board", err, err_cap);
  if (k && !cols) return fail(PEPS_E_INVALID, "null prefix", err, err_cap);
  for (uint32_t i = 0; i < k; ++i)
    if (cols[i] >= n) return fail(PEPS_E_INVALID, "prefix column outside the board", err, err_cap);
  if (k == 0) {
    peps_contraction_result res;
    int rc = peps_contract_rows(n, opts, &res, err, err_cap);
    if (rc == PEPS_OK) *count = res.count;
    return rc;
  }
  int rc = check_board(n, &e);
  if (rc) return fail(rc, e, err, err_cap);
  uint64_t key;
  if (!peps::prefix_state(n, cols, k, &key)) return PEPS_OK;  // attacking prefix: C gives 0
  if (k == n) {
    *count = abi(1);
    return PEPS_OK;
  }
  Resolved r;
  rc = peps::resolve_options(n, opts, &r, &e);
  if (rc) return fail(rc, e, err, err_cap);
  peps::RunSpec spec;
  spec.n = n;
  spec.first_row = k;
  spec.seeds.push_back(peps::Seed{key, 1});
  peps::RunOut ro;
  rc = run_engine(r, spec, &ro, &e);
  if (rc) return fail(rc, e, err, err_cap);
  *count = abi(ro.count);
  return PEPS_OK;
}

int peps_count_prefix_no_alloc(uint32_t n, const uint32_t* cols, uint32_t k, const peps_options* opts,
                               peps_u128* count, char* err, size_t err_cap) {
  if (!count) return fail(PEPS_E_INVALID, "null count pointer", err, err_cap);
  *count = abi(0);
  std::string e;
  if (n > 42) return fail(PEPS_E_INVALID, "the packed u128 virtual-boundary backend supports N <= 42", err, err_cap);
  if (k > n) return fail(PEPS_E_INVALID, "prefix longer than the board", err, err_cap);
  if (k && !cols) return fail(PEPS_E_INVALID, "null prefix", err, err_cap);
  for (uint32_t i = 0; i < k; ++i)
    if (cols[i] >= n) return fail(PEPS_E_INVALID, "prefix column outside the board", err, err_cap);
  if (n == 0) {  // src/lib.rs:5295-5307
    *count = abi(1);
    return PEPS_OK;
  }
  int rc = check_board(n, &e);
  if (rc) return fail(rc, e, err, err_cap);
  uint64_t key = 0;
  if (k && !peps::prefix_state(n, cols, k, &key)) return PEPS_OK;  // attacking prefix: C gives 0
  if (k == n) {
    *count = abi(1);
    return PEPS_OK;
  }
  Resolved r;
  rc = peps::resolve_options(n, opts, &r, &e);
  if (rc) return fail(rc, e, err, err_cap);
  if (r.graph) return fail(PEPS_E_INVALID, "graph_mode is not available without allocation", err, err_cap);
  peps::RunSpec spec;
  spec.n = n;
  spec.first_row = k;
  spec.orbit_first_row = k == 0;  // an empty prefix is contract_rows
  spec.seeds.push_back(peps::Seed{key, 1});
  peps::apply_resolved(r, &spec);
  spec.use_graph = false;
  spec.merge_waves = 1;   // the scratch's histogram slices hold one wave of workgroups
  spec.budget_bytes = 0;  // the scratch is the budget
  peps::Engine* eng = peps::Engine::get_scratch((int)r.device, &e);
  if (!eng) return fail(PEPS_E_DEVICE, e, err, err_cap);
  peps::RunOut ro;
  {
    std::lock_guard<std::mutex> lk(eng->mutex());
    rc = eng->run(spec, &ro, &e);
  }
  if (rc) return fail(rc, e, err, err_cap);
  *count = abi(ro.count);
  return PEPS_OK;
}

int peps_count_batched(uint32_t n, uint32_t split_depth, const peps_options* opts,
                       peps_contraction_result* out, uint32_t* tasks_out, peps_u128* task_counts,
                       uint64_t task_cap, char* err, size_t err_cap) {
  if (!out) return fail(PEPS_E_INVALID, "null result pointer", err, err_cap);
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  std::string e;
  if (n > 42) return fail(PEPS_E_INVALID, "the packed u128 virtual-boundary backend supports N <= 42", err, err_cap);
  if (n == 0 || split_depth == 0 || split_depth >= n) {
    if (tasks_out) *tasks_out = 1;
    int rc = peps_contract_rows(n, opts, out, err, err_cap);
    if (rc == PEPS_OK && task_counts && task_cap) task_counts[0] = out->count;
    return rc;
  }
  int rc = check_board(n, &e);
  if (rc) return fail(rc, e, err, err_cap);
  Resolved r;
  rc = peps::resolve_options(n, opts, &r, &e);
  if (rc) return fail(rc, e, err, err_cap);
  const std::vector<peps::Task> tasks = peps::make_tasks(n, split_depth);
  if (tasks_out) *tasks_out = (uint32_t)tasks.size();
  const size_t max_group = peps::max_slices(n, r.slice_group);
  // Groups are packed by projected frontier size: each task's 3-row extension
  // count, scaled by the candidates-per-extension ratio measured on the
  // groups run so far (the first group is kept small to calibrate it).
  peps::Projection proj;
  std::vector<double> proxy(tasks.size());
  for (size_t i = 0; i < tasks.size(); ++i) proxy[i] = proj.proxy(n, split_depth, tasks[i].key);
  const auto t0 = std::chrono::steady_clock::now();
  peps::RunOut total;
  total.residues.assign(r.primes.size(), 0);
  total.peak_states = 0;
  std::vector<peps::LayerOut> layers;
  std::vector<u128> per_task(tasks.size(), 0);
  size_t next = 0;
  uint64_t done = 0;
  bool overflow = false;
  peps::Engine* eng = peps::Engine::get((int)r.device, &e);
  if (!eng) return fail(PEPS_E_DEVICE, e, err, err_cap);
  auto absorb = [&](size_t g0, size_t g1, const peps::RunOut& ro) {
    double psum = 0;
    for (size_t i = g0; i < g1; ++i) psum += proxy[i];
    proj.learn(n, split_depth, psum, ro);
    for (size_t i = g0; i < g1; ++i) per_task[i] = ro.slice_counts[i - g0];
    if (!peps::checked_add(total.count, ro.count, &total.count)) overflow = true;
    total.peak_states = std::max(total.peak_states, ro.peak_states);
    total.total_candidates += ro.total_candidates;
    total.total_matched += ro.total_matched;
    total.kernel_s += ro.kernel_s;
    total.alloc_s += ro.alloc_s;
    total.device_peak_bytes = std::max(total.device_peak_bytes, ro.device_peak_bytes);
    total.record_bytes = std::max(total.record_bytes, ro.record_bytes);  // widest format any group used
    if (layers.empty()) {
      layers = ro.layers;
    } else {
      for (size_t i = 0; i < layers.size() && i < ro.layers.size(); ++i) {
        layers[i].input_states += ro.layers[i].input_states;
        layers[i].candidates += ro.layers[i].candidates;
        layers[i].matched += ro.layers[i].matched;
        layers[i].completed += ro.layers[i].completed;
        layers[i].output_states += ro.layers[i].output_states;
        layers[i].output_weight += ro.layers[i].output_weight;
        layers[i].elapsed_s += ro.layers[i].elapsed_s;
      }
    }
    done += g1 - g0;
    if (g_progress) g_progress(g_progress_user, done, tasks.size());
  };
  // Graph mode: the grouping found by the first call is kept; the next call
  // records every group's size plan and later calls replay the whole batched
  // count as one HIP graph.
  static std::mutex cache_mu;
  static std::map<std::string, std::vector<std::pair<size_t, size_t>>> grouping;
  char keybuf[256];
  snprintf(keybuf, sizeof(keybuf), "batched/%u/%u/%zu/%u/%u/%u/%u/%u/%u/%llu/%u", n, split_depth, r.primes.size(),
           r.radix_bits, r.tile_states, r.expand_bits, r.merge_passes, r.merge_waves, r.slice_group,
           (unsigned long long)r.budget, r.device);
  const std::string key = keybuf;
  std::vector<std::pair<size_t, size_t>> ranges;
  bool cached = false;
  if (r.graph) {
    std::lock_guard<std::mutex> lk(cache_mu);
    auto it = grouping.find(key);
    if (it != grouping.end()) {
      ranges = it->second;
      cached = true;
    }
  }
  if (cached) {
    rc = peps::run_task_groups(r, n, split_depth, tasks, ranges, key, absorb, &e);
    if (rc) return fail(rc, e, err, err_cap);
    if (overflow) return fail(PEPS_E_OVERFLOW, "batched count overflow", err, err_cap);
    next = tasks.size();
  }
  auto record = [&](size_t g0, size_t g1, const peps::RunOut& ro) {
    ranges.emplace_back(g0, g1);
    absorb(g0, g1, ro);
  };
  while (next < tasks.size()) {
    // Admit tasks while their projected candidates fit kAdmitFraction of a buffer.
    const double ratio = proj.ratio(n, split_depth);
    const double cap = peps::kAdmitFraction * (double)eng->candidate_capacity((int)r.primes.size(), r.records);
    const size_t g0 = next;
    size_t g1 = next + 1;
    double load = proxy[g0] * ratio;
    const size_t first_cap = ratio > 0 ? max_group : std::min<size_t>(max_group, 8);
    while (g1 < tasks.size() && g1 - g0 < first_cap) {
      const double add = proxy[g1] * ratio;
      if (ratio > 0 && cap > 0 && load + add > cap) break;
      load += add;
      ++g1;
    }
    next = g1;
    auto exceeded = [&](size_t a, size_t b) {
      double psum = 0;
      for (size_t i = a; i < b; ++i) psum += proxy[i];
      proj.exceeded(n, split_depth, psum, cap / peps::kAdmitFraction);
    };
    rc = peps::run_task_range(r, n, split_depth, tasks, g0, g1, record, &e, exceeded);
    if (rc) return fail(rc, e, err, err_cap);
    if (overflow) return fail(PEPS_E_OVERFLOW, "batched count overflow", err, err_cap);
    if (next == tasks.size() && r.graph) {
      std::lock_guard<std::mutex> lk(cache_mu);
      grouping[key] = ranges;
    }
  }
  total.layers = layers;
  u128 fact;
  peps::factorial(n, &fact);
  if (total.count > fact) return fail(PEPS_E_OVERFLOW, "batched count exceeds N!", err, err_cap);
  for (size_t l = 0; l < r.primes.size(); ++l) total.residues[l] = (uint64_t)(total.count % r.primes[l]);
  for (uint64_t i = 0; task_counts && i < task_cap && i < tasks.size(); ++i) task_counts[i] = abi(per_task[i]);
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_result(n, total, r.primes, el, out);
  return PEPS_OK;
}

int peps_known_count(uint32_t n, peps_u128* out) {
  u128 v;
  if (!peps::known_count(n, &v)) return PEPS_E_UNKNOWN;
  if (out) *out = abi(v);
  return PEPS_OK;
}

uint64_t peps_peak_rss_bytes(void) { return rss_bytes(); }

int peps_crt_plan(uint32_t n, uint64_t* primes, uint32_t* count, peps_u128* modulus_product,
                  char* err, size_t err_cap) {
  std::vector<uint64_t> p;
  u128 prod, fact;
  std::string e;
  if (!peps::crt_plan(n, &p, &prod, &fact, &e)) return fail(PEPS_E_INVALID, e, err, err_cap);
  for (size_t i = 0; i < p.size(); ++i)
    if (primes) primes[i] = p[i];
  if (count) *count = (uint32_t)p.size();
  if (modulus_product) *modulus_product = abi(prod);
  return PEPS_OK;
}

int peps_crt_reconstruct(const uint64_t* residues, const uint64_t* primes, uint32_t count,
                         peps_u128* out, char* err, size_t err_cap) {
  if (!residues || !primes || !out || count > PEPS_MAX_LANES)
    return fail(PEPS_E_INVALID, "CRT residue/modulus length mismatch", err, err_cap);
  u128 v;
  std::string e;
  if (!peps::crt_reconstruct(residues, primes, count, &v, &e)) return fail(PEPS_E_OVERFLOW, e, err, err_cap);
  *out = abi(v);
  return PEPS_OK;
}

int peps_prefix_tasks(uint32_t n, uint32_t depth, uint64_t* tasks, uint64_t cap, uint64_t* count,
                      char* err, size_t err_cap) {
  if (n > PEPS_MAX_DEVICE_N)
    return fail(PEPS_E_INVALID, "prefix tasks support N <= 21", err, err_cap);
  std::vector<peps::Task> t = peps::make_tasks(n, depth);
  if (count) *count = t.size();
  const uint64_t bm = n ? ((uint64_t)1 << n) - 1 : 0;
  for (uint64_t i = 0; i < t.size() && i < cap && tasks; ++i) {
    tasks[4 * i + 0] = t[i].key & bm;
    tasks[4 * i + 1] = (t[i].key >> n) & bm;
    tasks[4 * i + 2] = (t[i].key >> (2 * n)) & bm;
    tasks[4 * i + 3] = (uint64_t)t[i].weight;
  }
  return PEPS_OK;
}

int peps_device_info(uint32_t device, char* name, size_t name_cap, uint32_t* cu_count,
                     uint64_t* hbm_bytes, uint32_t* lds_bytes, char* err, size_t err_cap) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(PEPS_E_DEVICE, "no HIP device available", err, err_cap);
  if ((int)device >= count) return fail(PEPS_E_INVALID, "HIP device ordinal out of range", err, err_cap);
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, (int)device) != hipSuccess)
    return fail(PEPS_E_DEVICE, "hipGetDeviceProperties failed", err, err_cap);
  if (name && name_cap) {
    std::string s = std::string(p.name) + " (" + p.gcnArchName + ")";
    size_t k = std::min(name_cap - 1, s.size());
    std::memcpy(name, s.data(), k);
    name[k] = 0;
  }
  if (cu_count) *cu_count = (uint32_t)p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (uint64_t)p.totalGlobalMem;
  if (lds_bytes) *lds_bytes = (uint32_t)p.sharedMemPerBlock;
  return PEPS_OK;
}

const char* peps_version(void) { return "peps_mi355x 0.1.0 (gfx950)"; }

int peps_set_progress(peps_progress_fn fn, void* user) {
  g_progress = fn;
  g_progress_user = user;
  return PEPS_OK;
}

int peps_profile_kernels(uint32_t device, uint32_t enable, char* err, size_t err_cap) {
  std::string e;
  peps::Engine* eng = peps::Engine::get((int)device, &e);
  if (!eng) return fail(PEPS_E_DEVICE, e, err, err_cap);
  std::lock_guard<std::mutex> lk(eng->mutex());
  eng->set_profiling(enable != 0);
  return PEPS_OK;
}

int peps_kernel_stats(uint32_t device, peps_kernel_stat* out, uint32_t cap, uint32_t* count,
                      char* err, size_t err_cap) {
  std::string e;
  peps::Engine* eng = peps::Engine::get((int)device, &e);
  if (!eng) return fail(PEPS_E_DEVICE, e, err, err_cap);
  std::lock_guard<std::mutex> lk(eng->mutex());
  auto v = eng->kernel_stats();
  if (count) *count = (uint32_t)v.size();
  for (uint32_t i = 0; i < v.size() && i < cap && out; ++i) {
    std::memset(out[i].name, 0, sizeof(out[i].name));
    std::strncpy(out[i].name, v[i].first.c_str(), sizeof(out[i].name) - 1);
    out[i].launches = v[i].second.launches;
    out[i].seconds = v[i].second.seconds;
    out[i].algorithmic_bytes = v[i].second.bytes;
  }
  return PEPS_OK;
}

}  // extern "C"
