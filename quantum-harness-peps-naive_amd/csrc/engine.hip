// engine.hip -- see engine.h.  Follows the reference's row loop
// (contract_rows_sort_reduce_with_modes, src/lib.rs:5287-5415): per row,
// expand every parent through the compiled operator, apply the row-0 orbit
// weights, merge equal boundaries, record the layer metric; after the last
// row keep the states whose columns are all set (v1) and sum them (v2 frees
// the diagonals).
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>

namespace peps {

namespace {

using Clock = std::chrono::steady_clock;

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      *err = std::string(#expr) + " failed: " + hipGetErrorString(e_);             \
      return 3;                                                                    \
    }                                                                              \
  } while (0)

unsigned ceil_log2(uint64_t v) {
  unsigned b = 0;
  while (b < 64 && ((uint64_t)1 << b) < v) ++b;
  return b;
}

uint64_t mask_bits(unsigned b) { return b >= 64 ? ~0ull : (((uint64_t)1 << b) - 1); }

std::mutex g_engines_mu;
std::map<int, std::unique_ptr<Engine>>& engines() {
  static std::map<int, std::unique_ptr<Engine>> m;
  return m;
}
std::map<int, std::unique_ptr<Engine>>& scratch_engines() {
  static std::map<int, std::unique_ptr<Engine>> m;
  return m;
}

}  // namespace

// The no-allocation engine's device memory: one array of kScratchBytes per
// device, part of the library's code object, held from the library's load
// on that device until the process ends.  Host code never passes the
// array's name to a kernel or a copy (on the host it is a stand-in symbol,
// not a device address): init_scratch looks up each device's copy with
// hipGetSymbolAddress, with that device current, and the engine uses that
// address only.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

Engine* Engine::get_scratch(int device, std::string* err) {
  std::lock_guard<std::mutex> lk(g_engines_mu);
  auto& m = scratch_engines();
  auto it = m.find(device);
  if (it != m.end()) return it->second.get();
  std::unique_ptr<Engine> e(new Engine(device));
  if (e->init_scratch(err) != 0) return nullptr;
  Engine* raw = e.get();
  m[device] = std::move(e);
  return raw;
}

int Engine::init_scratch(std::string* err) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    *err = "no HIP device available (the MI355X path needs a gfx950 GPU)";
    return 3;
  }
  if (device_ < 0 || device_ >= count) {
    *err = "HIP device ordinal out of range";
    return 1;
  }
  HIP_TRY(hipSetDevice(device_));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device_));
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) {
    *err = std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950";
    return 3;
  }
  cu_count_ = (uint32_t)std::max(1, prop.multiProcessorCount);
  void* base = nullptr;
  HIP_TRY(hipGetSymbolAddress(&base, HIP_SYMBOL(scratch)));
  if (!base) {
    *err = "scratch array has no device address";
    return 3;
  }
  fixed_ = true;
  stream_ = nullptr;  // the device's null stream: nothing to create
  h_counters_ = h_local_;
  // Fixed slices, 256-byte aligned.  The small buffers are sized for a
  // merge grid of one wave of resident workgroups (the no-allocation runs
  // use merge_waves = 1): COUNT histograms of 2^10 digits x 1024
  // workgroups; the two candidate regions share what is left.
  uint8_t* p = (uint8_t*)base;
  size_t left = kScratchBytes;
  auto take = [&](size_t bytes) -> void* {
    bytes = (bytes + 255) & ~(size_t)255;
    void* r = p;
    p += bytes;
    left -= bytes;
    return r;
  };
  constexpr size_t kCells = (size_t)1 << 20;
  counters_ = (uint64_t*)take(sizeof(uint64_t) * kNumCounters);
  counters_rows_cap_ = 64 * kNumCounters;
  counters_rows_ = (uint64_t*)take(8 * counters_rows_cap_);
  seed_cap_ = 8192;
  seed_ = (uint64_t*)take(8 * seed_cap_);
  slice_cap_ = 1024;
  slice_sums_ = (uint64_t*)take(8 * slice_cap_);
  hist_cap_ = kCells;
  hist_ = (uint32_t*)take(4 * hist_cap_);
  offsets_cap_ = kCells;
  offsets_ = (uint64_t*)take(8 * offsets_cap_);
  partial_cap_ = kCells / 4096 + 64;
  partial_ = (uint64_t*)take(8 * partial_cap_);
  for (int i = 0; i < 2; ++i) {
    segs_cap_[i] = (size_t)1 << 17;
    segs_[i] = (uint64_t*)take(8 * segs_cap_[i]);
  }
  run_start_cap_ = kCells;
  run_start_ = (uint64_t*)take(8 * run_start_cap_);
  run_len_cap_ = kCells;
  run_len_ = (uint32_t*)take(4 * run_len_cap_);
  fan_cap_ = kCells / 2 + 64;
  fan_ = (uint64_t*)take(8 * fan_cap_);
  const size_t half = (left / 2) & ~(size_t)255;
  for (int i = 0; i < 2; ++i) {
    region_bytes_[i] = half;
    region_[i] = take(half);
  }
  budget_ = arena_bytes();
  std::string cerr;
  if (!compile_row_operator(site_tensor_c(site_tensor_b()), &occupied_, &cerr)) {
    *err = cerr;
    return 1;
  }
  return 0;
}

Engine* Engine::get(int device, std::string* err) {
  std::lock_guard<std::mutex> lk(g_engines_mu);
  auto& m = engines();
  auto it = m.find(device);
  if (it != m.end()) return it->second.get();
  std::unique_ptr<Engine> e(new Engine(device));
  if (e->init(err) != 0) return nullptr;
  Engine* raw = e.get();
  m[device] = std::move(e);
  return raw;
}

int Engine::init(std::string* err) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    *err = "no HIP device available (the MI355X path needs a gfx950 GPU)";
    return 3;
  }
  if (device_ < 0 || device_ >= count) {
    *err = "HIP device ordinal out of range";
    return 1;
  }
  HIP_TRY(hipSetDevice(device_));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device_));
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) {
    *err = std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950";
    return 3;
  }
  cu_count_ = (uint32_t)std::max(1, prop.multiProcessorCount);
  HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&ev_start_));
  HIP_TRY(hipEventCreate(&ev_stop_));
  HIP_TRY(hipMalloc(&counters_, sizeof(uint64_t) * kNumCounters));
  HIP_TRY(hipHostMalloc((void**)&h_counters_, sizeof(uint64_t) * kNumCounters, hipHostMallocDefault));
  std::string cerr;
  if (!compile_row_operator(site_tensor_c(site_tensor_b()), &occupied_, &cerr)) {
    *err = cerr;
    return 1;
  }
  return 0;
}

uint64_t Engine::arena_bytes() const {
  return region_bytes_[0] + region_bytes_[1];
}

int Engine::ensure_region(int idx, size_t bytes, std::string* err) {
  if (region_bytes_[idx] >= bytes) return 0;
  if (fixed_) {
    *err = "the no-allocation scratch cannot hold this board: a row needs " + std::to_string(bytes) +
           " bytes per candidate buffer, the scratch has " + std::to_string(region_bytes_[idx]);
    return 4;
  }
  const auto ta = Clock::now();
  struct Tally {
    double* acc;
    Clock::time_point t;
    ~Tally() { *acc += std::chrono::duration<double>(Clock::now() - t).count(); }
  } tally{&alloc_s_, ta};
  // Freeing and re-allocating a region of tens of GB costs seconds, so a
  // region that must exceed 1 GiB is allocated once at its final size and
  // never regrown: half of the arena budget less a reserve for the
  // histogram, offset and segment buffers, so that both regions fit.  A row
  // that needs more than that fails with PEPS_E_MEMORY (the batched form and
  // the server then split their task group).
  size_t want = std::max<size_t>(bytes + bytes / 4, (size_t)1 << 20);
  if (bytes > ((size_t)1 << 30)) {
    want = (size_t)region_cap();
    if (bytes > want) {
      *err = "HBM budget exceeded: a row needs " + std::to_string(bytes) + " bytes per candidate buffer, " +
             "the arena allows " + std::to_string(want);
      return 4;
    }
  }
  want = (want + 4095) & ~(size_t)4095;
  if (region_[idx]) {
    HIP_TRY(hipStreamSynchronize(stream_));
    HIP_TRY(hipFree(region_[idx]));
    region_[idx] = nullptr;
    region_bytes_[idx] = 0;
  }
  uint64_t others = arena_bytes();
  if (others + want > budget_) {
    want = (bytes + 4095) & ~(size_t)4095;
    if (others + want > budget_) {
      *err = "HBM budget exceeded: row needs " + std::to_string(others + want) +
             " arena bytes, budget " + std::to_string(budget_);
      return 4;
    }
  }
  const hipError_t me = hipMalloc(&region_[idx], want);
  if (me == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    region_[idx] = nullptr;
    *err = "HBM allocation of " + std::to_string(want) + " bytes failed";
    return 4;
  }
  HIP_TRY(me);
  region_bytes_[idx] = want;
  ++generation_;
  return 0;
}

uint64_t Engine::region_cap() const { return (budget_ - budget_ / 32) / 2; }

uint64_t Engine::candidate_capacity(int lanes, int records) const {
  return budget_ ? region_cap() / record_bytes(records == kRecordsLanes ? lanes : kPackedRecords) : 0;
}

template <typename T>
int Engine::ensure_buffer(T** ptr, size_t* cap, size_t count, std::string* err) {
  if (*cap >= count) return 0;
  if (fixed_) {
    *err = "the no-allocation scratch cannot hold this board: a launch needs " + std::to_string(count) +
           " entries of a buffer whose scratch slice holds " + std::to_string(*cap);
    return 4;
  }
  struct Tally {
    double* acc;
    Clock::time_point t;
    ~Tally() { *acc += std::chrono::duration<double>(Clock::now() - t).count(); }
  } tally{&alloc_s_, Clock::now()};
  // Every growth synchronises, frees and reallocates (and invalidates
  // captured graphs); doubling keeps the number of growths logarithmic in
  // the largest launch seen.
  size_t want = std::max<size_t>(2 * count, 1024);
  if (*ptr) {
    HIP_TRY(hipStreamSynchronize(stream_));
    HIP_TRY(hipFree(*ptr));
    *ptr = nullptr;
  }
  HIP_TRY(hipMalloc((void**)ptr, want * sizeof(T)));
  *cap = want;
  ++generation_;
  return 0;
}

void Engine::set_profiling(bool on) {
  profiling_ = on;
  kstats_.assign(kernel_names_.size(), KernelStat{});
}

std::vector<std::pair<std::string, KernelStat>> Engine::kernel_stats() const {
  std::vector<std::pair<std::string, KernelStat>> v;
  for (size_t i = 0; i < kernel_names_.size() && i < kstats_.size(); ++i)
    if (kstats_[i].launches) v.emplace_back(kernel_names_[i], kstats_[i]);
  return v;
}

int Engine::kernel_id(const char* name) {
  for (size_t i = 0; i < kernel_names_.size(); ++i)
    if (kernel_names_[i] == name) return (int)i;
  kernel_names_.push_back(name);
  kstats_.resize(kernel_names_.size());
  return (int)kernel_names_.size() - 1;
}

hipError_t Engine::mark_begin(hipEvent_t* ev) {
  if (event_pool_.empty()) {
    hipError_t e = hipEventCreate(ev);
    if (e != hipSuccess) return e;
  } else {
    *ev = event_pool_.back();
    event_pool_.pop_back();
  }
  return hipEventRecord(*ev, stream_);
}

hipError_t Engine::mark_end(int kernel, double bytes, hipEvent_t begin) {
  hipEvent_t ev;
  if (event_pool_.empty()) {
    hipError_t e = hipEventCreate(&ev);
    if (e != hipSuccess) return e;
  } else {
    ev = event_pool_.back();
    event_pool_.pop_back();
  }
  pending_.push_back(Pending{kernel, bytes, begin, ev});
  return hipEventRecord(ev, stream_);
}

// Called after a stream synchronisation: every pending event has completed.
int Engine::resolve_pending(std::string* err) {
  for (const Pending& p : pending_) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
    KernelStat& k = kstats_[p.kernel];
    k.launches += 1;
    k.seconds += ms / 1000.0;
    k.bytes += p.bytes;
    event_pool_.push_back(p.a);
    event_pool_.push_back(p.b);
  }
  pending_.clear();
  return 0;
}

#define LAUNCH(name, bytes, call)                                       \
  do {                                                                  \
    hipEvent_t ev_ = nullptr;                                           \
    if (profiling_) HIP_TRY(mark_begin(&ev_));                          \
    HIP_TRY(call);                                                      \
    if (profiling_) HIP_TRY(mark_end(kernel_id(name), (double)(bytes), ev_)); \
  } while (0)

#define SYNC_AND_RESOLVE()                                              \
  do {                                                                  \
    HIP_TRY(hipStreamSynchronize(stream_));                             \
    if (!pending_.empty()) {                                            \
      int rc_ = resolve_pending(err);                                   \
      if (rc_) return rc_;                                              \
    }                                                                   \
  } while (0)

namespace {

// Internal return code: a packed record's weight outgrew its field.
constexpr int kRcPackedOverflow = 100;

// Return code of a failed consistency check (PEPS_E_INTERNAL): a device
// error flag or a size plan that does not match the device.  Callers that
// split a task group and retry do so for PEPS_E_MEMORY only, which only an
// allocation the arena refuses returns.
constexpr int kRcInternal = 8;

// The error-flag bits of a row's kError counter, in words.
std::string error_flag_text(uint64_t bits, const std::string& where) {
  std::string s;
  auto add = [&](const char* t) { s += (s.empty() ? "" : "; ") + std::string(t); };
  if (bits & 1) add("LDS merge table overflow");
  if (bits & 2) add("a kernel write reached past its buffer's planned capacity");
  if (bits & 8) add("sitewise partial-row limit exceeded");
  if (bits & 16) add("a capped SCATTER run overflowed");
  if (bits & 32) add("a fixed-capacity bucket overflowed");
  if (bits & ~59ull) add("unknown device error flag");
  return s + where;
}

// Rows with at least this many candidates use the upper-bound COUNT.
constexpr uint64_t kUpperBoundRows = (uint64_t)1 << 20;

bool packed_records(const RunSpec& s) { return s.records != kRecordsLanes && !s.materialize && !s.site; }

// Key and weight widths of the packed records of a run.
void packed_widths(const RunSpec& s, uint32_t* key_bits, uint32_t* weight_bits) {
  *key_bits = 3 * s.n + s.slice_bits;
  *weight_bits = std::min<uint32_t>(std::min<uint32_t>(96 - *key_bits, kPackedMaxWeightBits),
                                    std::max<uint32_t>(1, s.packed_weight_bits));
}

// u64 words of the device seed image.
size_t seed_words(const RunSpec& s) {
  const size_t S0 = s.seeds.size();
  return packed_records(s) ? 2 * S0 : (S0 * (8 + 4 * s.primes.size()) + 7) / 8;
}

// Device image of the seed frontier: keys[S0], then the exact u64 weights
// (packed records) or one u32 residue array per CRT prime (lane records).
int seed_image(const RunSpec& s, std::vector<uint64_t>* host) {
  const size_t S0 = s.seeds.size();
  host->assign(seed_words(s), 0);
  uint64_t* keys = host->data();
  if (packed_records(s)) {
    uint32_t kb, wb;
    packed_widths(s, &kb, &wb);
    for (size_t i = 0; i < S0; ++i) {
      keys[i] = s.seeds[i].key;
      if (s.seeds[i].weight >> wb) return kRcPackedOverflow;
      keys[S0 + i] = (uint64_t)s.seeds[i].weight;
    }
    return 0;
  }
  uint32_t* res = (uint32_t*)(keys + S0);
  for (size_t i = 0; i < S0; ++i) {
    keys[i] = s.seeds[i].key;
    for (size_t l = 0; l < s.primes.size(); ++l)
      res[l * S0 + i] = (uint32_t)(s.seeds[i].weight % s.primes[l]);
  }
  return 0;
}

// Identity of a run for the plan cache: everything that determines sizes.
std::string plan_key(const RunSpec& s) {
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the seeds
  for (const Seed& sd : s.seeds) {
    for (uint64_t v : {sd.key, (uint64_t)sd.weight, (uint64_t)(sd.weight >> 64)}) {
      for (int b = 0; b < 8; ++b) {
        h ^= (v >> (8 * b)) & 0xff;
        h *= 1099511628211ull;
      }
    }
  }
  char buf[256];
  snprintf(buf, sizeof(buf), "%u/%u/%d/%u/%zu/%u/%u/%u/%u/%u/%d/%zu/%llx/%d/%u/%u", s.n, s.first_row,
           (int)s.orbit_first_row, s.slice_bits, s.primes.size(), s.tile_states, s.radix_bits,
           s.expand_bits, s.merge_passes, s.merge_waves, (int)s.per_slice, s.seeds.size(),
           (unsigned long long)h, s.records, s.packed_weight_bits, s.force);
  return buf;
}

}  // namespace

int Engine::prepare(const RunSpec& spec, std::string* err) {
  HIP_TRY(hipSetDevice(device_));
  const unsigned n = spec.n;
  const int L = (int)spec.primes.size();
  if (n == 0 || n > 21 || L < 1 || L > kMaxLanes) {
    *err = "engine: unsupported board or lane count";
    return 1;
  }
  if (3 * n + spec.slice_bits > 63 || spec.slice_bits > 20) {
    *err = "engine: slice ids do not fit above the 3N-bit boundary key";
    return 1;
  }
  if (spec.first_row >= n || spec.seeds.empty()) {
    *err = "engine: nothing to contract";
    return 1;
  }
  if (!fixed_ && (budget_ == 0 || (spec.budget_bytes && spec.budget_bytes != budget_))) {
    if (spec.budget_bytes) {
      budget_ = spec.budget_bytes;
    } else {
      size_t free_b = 0, total_b = 0;
      HIP_TRY(hipMemGetInfo(&free_b, &total_b));
      budget_ = (uint64_t)((double)(free_b + arena_bytes()) * 0.85);
    }
  }
  return 0;
}

int Engine::run_many(const std::string& key, const std::vector<RunSpec>& specs, std::vector<RunOut>* outs,
                     std::string* err) {
  const int rc = run_many_once(key, specs, outs, err);
  if (rc != kRcPackedOverflow) return rc;
  for (const RunSpec& sp : specs)
    if (sp.records == kRecordsPacked) {
      *err = "packed record weight overflow";
      return 2;
    }
  std::vector<RunSpec> lanes = specs;
  for (RunSpec& sp : lanes) sp.records = kRecordsLanes;
  return run_many_once(key + "/lanes", lanes, outs, err);
}

int Engine::run_many_once(const std::string& key, const std::vector<RunSpec>& specs, std::vector<RunOut>* outs,
                          std::string* err) {
  outs->assign(specs.size(), RunOut{});
  for (const RunSpec& sp : specs) {
    if (sp.materialize || sp.site) {
      *err = "engine: grouped graph replay covers the fused row pipeline only";
      return 1;
    }
    const int rc = prepare(sp, err);
    if (rc) return rc;
  }
  auto it = batches_.find(key);
  if (it == batches_.end() || it->second.plans.size() != specs.size()) {
    // First call: eager runs, each recording its size plan.
    BatchGraph bg;
    bg.plans.resize(specs.size());
    for (size_t i = 0; i < specs.size(); ++i) {
      const auto t0 = Clock::now();
      const double a0 = alloc_s_;
      const int rc = pipeline(specs[i], &bg.plans[i], true, &(*outs)[i], err);
      if (rc) return rc;
      (*outs)[i].alloc_s = alloc_s_ - a0;
      (*outs)[i].elapsed_s = std::chrono::duration<double>(Clock::now() - t0).count();
    }
    batches_[key] = std::move(bg);
    return 0;
  }
  BatchGraph& bg = it->second;
  // Shared buffers at least as large as every recorded group needs.
  int rc = 0;
  for (const Plan& plan : bg.plans) {
    if (!rc) rc = ensure_region(0, plan.region_need, err);
    if (!rc) rc = ensure_region(1, plan.region_need, err);
    if (!rc) rc = ensure_buffer(&hist_, &hist_cap_, plan.hist_need, err);
    if (!rc) rc = ensure_buffer(&offsets_, &offsets_cap_, plan.offsets_need, err);
    if (!rc) rc = ensure_buffer(&partial_, &partial_cap_, plan.partial_need, err);
    if (!rc) rc = ensure_buffer(&segs_[0], &segs_cap_[0], plan.segs_need[0], err);
    if (!rc) rc = ensure_buffer(&segs_[1], &segs_cap_[1], plan.segs_need[1], err);
    if (!rc) rc = ensure_buffer(&run_start_, &run_start_cap_, plan.run_start_need, err);
    if (!rc) rc = ensure_buffer(&run_len_, &run_len_cap_, plan.run_len_need, err);
    if (!rc) rc = ensure_buffer(&fan_, &fan_cap_, plan.fan_need, err);
  }
  if (rc) return rc;
  // Per-group seeds, counters and slice sums, so that every group's pipeline
  // can sit in one graph.
  if (bg.bufs.empty()) {
    bg.bufs.resize(specs.size());
    for (size_t i = 0; i < specs.size(); ++i) {
      const RunSpec& sp = specs[i];
      GroupBufs& gb = bg.bufs[i];
      std::vector<uint64_t> host;
      if (seed_image(sp, &host)) return kRcPackedOverflow;
      gb.seed_words = host.size();
      gb.counter_words = (size_t)(sp.n + 1) * kNumCounters;
      gb.slice_words = sp.per_slice ? ((size_t)1 << sp.slice_bits) * sp.primes.size() : 0;
      HIP_TRY(hipMalloc((void**)&gb.seed, gb.seed_words * 8));
      HIP_TRY(hipMalloc((void**)&gb.counters, gb.counter_words * 8));
      if (gb.slice_words) HIP_TRY(hipMalloc((void**)&gb.slices, gb.slice_words * 8));
      HIP_TRY(hipMemcpy(gb.seed, host.data(), gb.seed_words * 8, hipMemcpyHostToDevice));
    }
  }
  struct Swap {
    Engine* e;
    uint64_t *seed, *counters, *slices;
    size_t seed_cap, counters_cap, slice_cap;
    Swap(Engine* en, const GroupBufs& gb)
        : e(en), seed(en->seed_), counters(en->counters_rows_), slices(en->slice_sums_),
          seed_cap(en->seed_cap_), counters_cap(en->counters_rows_cap_), slice_cap(en->slice_cap_) {
      e->seed_ = gb.seed;
      e->seed_cap_ = gb.seed_words;
      e->counters_rows_ = gb.counters;
      e->counters_rows_cap_ = gb.counter_words;
      e->slice_sums_ = gb.slices;
      e->slice_cap_ = gb.slice_words;
    }
    ~Swap() {
      e->seed_ = seed;
      e->seed_cap_ = seed_cap;
      e->counters_rows_ = counters;
      e->counters_rows_cap_ = counters_cap;
      e->slice_sums_ = slices;
      e->slice_cap_ = slice_cap;
    }
  };
  if (bg.exec && bg.generation != generation_) {
    HIP_TRY(hipGraphExecDestroy(bg.exec));
    HIP_TRY(hipGraphDestroy(bg.graph));
    bg.exec = nullptr;
    bg.graph = nullptr;
  }
  if (!bg.exec) {
    HIP_TRY(hipStreamSynchronize(stream_));
    const bool prof = profiling_;
    profiling_ = false;
    HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    for (size_t i = 0; i < specs.size() && !rc; ++i) {
      Swap sw(this, bg.bufs[i]);
      RunOut scratch;
      rc = pipeline(specs[i], &bg.plans[i], false, &scratch, err);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(stream_, &graph);
    profiling_ = prof;
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIP_TRY(ec);
    HIP_TRY(hipGraphInstantiate(&bg.exec, graph, nullptr, nullptr, 0));
    bg.graph = graph;
    bg.generation = generation_;
  }
  const auto t0 = Clock::now();
  HIP_TRY(hipEventRecord(ev_start_, stream_));
  HIP_TRY(hipGraphLaunch(bg.exec, stream_));
  HIP_TRY(hipEventRecord(ev_stop_, stream_));
  HIP_TRY(hipEventSynchronize(ev_stop_));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, ev_start_, ev_stop_));
  for (size_t i = 0; i < specs.size(); ++i) {
    Swap sw(this, bg.bufs[i]);
    rc = finish(specs[i], bg.plans[i], &(*outs)[i], err);
    if (rc) return rc;
    (*outs)[i].replayed = true;
  }
  (*outs)[0].kernel_s = ms / 1000.0;
  (*outs)[0].elapsed_s = std::chrono::duration<double>(Clock::now() - t0).count();
  return 0;
}

int Engine::run(const RunSpec& spec, RunOut* out, std::string* err) {
  if (fixed_ && (spec.use_graph || spec.materialize || spec.site || spec.per_slice || spec.slice_bits)) {
    *err = "the no-allocation engine runs the fused row pipeline of one prefix only";
    return 1;
  }
  if (fixed_ && profiling_) set_profiling(false);
  const int rc = run_once(spec, out, err);
  if (rc != kRcPackedOverflow) return rc;
  if (spec.records == kRecordsPacked) {
    *err = "packed record weight overflow";
    return 2;
  }
  RunSpec lanes = spec;
  lanes.records = kRecordsLanes;
  return run_once(lanes, out, err);
}

int Engine::run_once(const RunSpec& spec, RunOut* out, std::string* err) {
  const auto t0 = Clock::now();
  const double alloc0 = alloc_s_;
  struct AllocReport {
    RunOut* out;
    const double* acc;
    double base;
    ~AllocReport() { out->alloc_s = *acc - base; }
  } alloc_report{out, &alloc_s_, alloc0};
  HIP_TRY(hipSetDevice(device_));
  const unsigned n = spec.n;
  const int L = (int)spec.primes.size();
  if (n == 0 || n > 21 || L < 1 || L > kMaxLanes) {
    *err = "engine: unsupported board or lane count";
    return 1;
  }
  if (3 * n + spec.slice_bits > 63 || spec.slice_bits > 20) {
    *err = "engine: slice ids do not fit above the 3N-bit boundary key";
    return 1;
  }
  if (spec.first_row >= n || spec.seeds.empty()) {
    *err = "engine: nothing to contract";
    return 1;
  }
  if (!fixed_ && (budget_ == 0 || (spec.budget_bytes && spec.budget_bytes != budget_))) {
    if (spec.budget_bytes) {
      budget_ = spec.budget_bytes;
    } else {
      size_t free_b = 0, total_b = 0;
      HIP_TRY(hipMemGetInfo(&free_b, &total_b));
      budget_ = (uint64_t)((double)(free_b + arena_bytes()) * 0.85);
    }
  }
  out->replayed = false;
  int rc;
  if (spec.materialize || spec.site) {
    rc = materialized(spec, out, err);
    if (rc) return rc;
    out->elapsed_s = std::chrono::duration<double>(Clock::now() - t0).count();
    return 0;
  }
  if (spec.use_graph) {
    const std::string key = plan_key(spec);
    auto it = plans_.find(key);
    if (it != plans_.end()) {
      Plan& plan = it->second;
      // Buffers at least as large as the recorded run; any reallocation
      // invalidates a captured graph (it holds device pointers).
      std::vector<uint64_t> seed_host;
      if (seed_image(spec, &seed_host)) return kRcPackedOverflow;
      rc = ensure_region(0, plan.region_need, err);
      if (!rc) rc = ensure_region(1, plan.region_need, err);
      if (!rc) rc = ensure_buffer(&hist_, &hist_cap_, plan.hist_need, err);
      if (!rc) rc = ensure_buffer(&offsets_, &offsets_cap_, plan.offsets_need, err);
      if (!rc) rc = ensure_buffer(&partial_, &partial_cap_, plan.partial_need, err);
      if (!rc) rc = ensure_buffer(&segs_[0], &segs_cap_[0], plan.segs_need[0], err);
      if (!rc) rc = ensure_buffer(&segs_[1], &segs_cap_[1], plan.segs_need[1], err);
      if (!rc) rc = ensure_buffer(&run_start_, &run_start_cap_, plan.run_start_need, err);
      if (!rc) rc = ensure_buffer(&run_len_, &run_len_cap_, plan.run_len_need, err);
      if (!rc) rc = ensure_buffer(&fan_, &fan_cap_, plan.fan_need, err);
      if (!rc) rc = ensure_buffer(&counters_rows_, &counters_rows_cap_, (size_t)(n + 1) * kNumCounters, err);
      if (!rc) rc = ensure_buffer(&seed_, &seed_cap_, seed_host.size(), err);
      if (!rc && spec.per_slice)
        rc = ensure_buffer(&slice_sums_, &slice_cap_, ((size_t)1 << spec.slice_bits) * L, err);
      if (rc) return rc;
      if (plan.exec && plan.generation != generation_) {
        HIP_TRY(hipGraphExecDestroy(plan.exec));
        HIP_TRY(hipGraphDestroy(plan.graph));
        plan.exec = nullptr;
        plan.graph = nullptr;
      }
      if (!plan.exec) {
        HIP_TRY(hipStreamSynchronize(stream_));
        const bool prof = profiling_;
        profiling_ = false;  // no event records inside a captured graph
        HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        rc = pipeline(spec, &plan, false, out, err);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(stream_, &graph);
        profiling_ = prof;
        if (rc) {
          if (graph) (void)hipGraphDestroy(graph);
          return rc;
        }
        HIP_TRY(ec);
        HIP_TRY(hipGraphInstantiate(&plan.exec, graph, nullptr, nullptr, 0));
        plan.graph = graph;
        plan.generation = generation_;
      }
      // Seeds are uploaded outside the graph, then the graph replays.
      HIP_TRY(hipMemcpy(seed_, seed_host.data(), seed_host.size() * 8, hipMemcpyHostToDevice));
      HIP_TRY(hipEventRecord(ev_start_, stream_));
      HIP_TRY(hipGraphLaunch(plan.exec, stream_));
      HIP_TRY(hipEventRecord(ev_stop_, stream_));
      HIP_TRY(hipEventSynchronize(ev_stop_));
      rc = finish(spec, plan, out, err);
      if (rc) return rc;
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, ev_start_, ev_stop_));
      out->kernel_s = ms / 1000.0;
      out->replayed = true;
      out->elapsed_s = std::chrono::duration<double>(Clock::now() - t0).count();
      return 0;
    }
    Plan plan;
    rc = pipeline(spec, &plan, true, out, err);
    if (rc) return rc;
    plans_[key] = std::move(plan);
  } else {
    Plan plan;
    rc = pipeline(spec, &plan, true, out, err);
    if (rc) return rc;
  }
  out->elapsed_s = std::chrono::duration<double>(Clock::now() - t0).count();
  return 0;
}

// The row pipeline.  eager: sizes are read back from the device as the rows
// run (host sync per row), buffers grow on demand, layer statistics are
// collected per row, and every size is recorded into *plan.  !eager: the same
// launch sequence is enqueued from *plan without any synchronisation (used
// under HIP graph capture); statistics are read afterwards by finish().
int Engine::pipeline(const RunSpec& spec, Plan* plan, bool eager, RunOut* out, std::string* err) {
  const unsigned n = spec.n;
  const int L = (int)spec.primes.size();
  // Record layout of the candidate buffers: packed (key | exact weight, 12
  // bytes) or lane records (key + L residues).
  const bool packed = packed_records(spec);
  const int LK = packed ? kPackedRecords : L;
  const size_t rb = record_bytes(LK);
  const double R = packed ? 12.0 : 8.0 + 4.0 * L;  // algorithmic bytes per record
  Primes pr{};
  for (int l = 0; l < L; ++l) pr.p[l] = (uint32_t)spec.primes[l];
  pr.lanes = (uint32_t)L;
  u128 fact;
  factorial(n, &fact);

  RowGeom g{};
  g.n = n;
  g.board_mask = (uint32_t)mask_bits(n);
  g.state_mask = mask_bits(3 * n);
  g.orbit_mask = g.board_mask;
  g.col_in = occupied_.column_in ? g.board_mask : 0u;
  g.col_out = occupied_.column_out ? g.board_mask : 0u;
  g.dr_in = occupied_.diag_dr_in ? g.board_mask : 0u;
  g.dr_out = occupied_.diag_dr_out ? g.board_mask : 0u;
  g.dl_in = occupied_.diag_dl_in ? g.board_mask : 0u;
  g.dl_out = occupied_.diag_dl_out ? g.board_mask : 0u;
  packed_widths(spec, &g.key_bits, &g.weight_bits);
  const uint64_t final_mask = g.board_mask;

  const uint32_t rbits = std::max<uint32_t>(1, std::min<uint32_t>(spec.radix_bits, kMaxDigitBits));
  // LDS merge table: 2048 slots (more workgroups per CU) for small tiles,
  // 4096 slots otherwise; the bucket target stays within one merge pass.
  const uint32_t table_slots = spec.tile_states <= merge_cap(2048) ? 2048u : 4096u;
  // Buckets may span several LDS passes: the merge re-reads the bucket per
  // pass (from L2/MALL), which can be cheaper than one more partition level.
  const uint64_t target = (uint64_t)std::max<uint32_t>(256, std::min<uint32_t>(spec.tile_states, merge_cap(table_slots))) *
                          std::max<uint32_t>(1, std::min<uint32_t>(spec.merge_passes, 8));
  auto total_bits = [&](uint64_t C) -> unsigned {
    return C <= target ? 0u : ceil_log2((C + target - 1) / target);
  };
  // One full wave of resident SCATTER workgroups (times merge_waves), so the
  // bucket ranges finish together instead of leaving a partly occupied tail.
  const uint64_t merge_grid = (uint64_t)cu_count_ *
                              merge_resident_blocks(LK, table_slots, std::min<uint32_t>(spec.expand_bits, kFineBits)) *
                              std::max<uint32_t>(1, std::min<uint32_t>(spec.merge_waves, 8));
  auto split_bits = [&](unsigned B, unsigned d1) {
    std::vector<unsigned> d;
    unsigned used = d1;
    while (used < B) {
      const unsigned dk = std::min<unsigned>(rbits, B - used);
      d.push_back(dk);
      used += dk;
    }
    // A short trailing level would read and write every record once more
    // for a few bits: spread them over the levels before it instead, one
    // more bit each (from the last), when those levels can take them all.
    const unsigned cap = std::min<unsigned>(kMaxDigitBits, rbits + 1);
    if (d.size() >= 2) {
      unsigned rest = d.back(), room = 0;
      for (size_t i = 0; i + 1 < d.size(); ++i) room += d[i] < cap ? 1u : 0u;
      if (rest <= room) {
        d.pop_back();
        for (size_t i = d.size(); i-- > 0 && rest;)
          if (d[i] < cap) {
            d[i] += 1;
            rest -= 1;
          }
      }
    }
    return d;
  };
  int rc = 0;
  // Grow one candidate region; only ever called on a region that holds no
  // live records (the write target of the next kernel).
  auto grow = [&](int idx, size_t need_bytes) -> int {
    plan->region_need = std::max(plan->region_need, need_bytes);
    if (!eager) return 0;
    return ensure_region(idx, need_bytes, err);
  };
  auto need = [&](auto** ptr, size_t* cap, size_t* rec, size_t count) -> int {
    *rec = std::max(*rec, count);
    if (!eager) return 0;
    return ensure_buffer(ptr, cap, count, err);
  };

  // Per-row counters and slice sums, cleared with fill kernels.
  const size_t ncount = (size_t)(n + 1) * kNumCounters;
  size_t dummy = 0;
  rc = need(&counters_rows_, &counters_rows_cap_, &dummy, ncount);
  if (rc) return rc;
  auto row_counters = [&](unsigned row) { return counters_rows_ + (size_t)row * kNumCounters; };
  const size_t nslice = spec.per_slice ? ((size_t)1 << spec.slice_bits) * L : 0;
  if (nslice) {
    rc = need(&slice_sums_, &slice_cap_, &dummy, nslice);
    if (rc) return rc;
  }
  uint64_t* ssum = nslice ? slice_sums_ : nullptr;

  const uint64_t S0 = spec.seeds.size();
  if (eager) {
    std::vector<uint64_t> host;
    if (seed_image(spec, &host)) return kRcPackedOverflow;
    rc = ensure_buffer(&seed_, &seed_cap_, host.size(), err);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(stream_));
    HIP_TRY(hipMemcpy(seed_, host.data(), host.size() * 8, hipMemcpyHostToDevice));
    out->record_bytes = (uint32_t)rb;
    out->layers.clear();
    out->peak_states = S0;
    out->total_candidates = out->total_matched = 0;
    out->device_peak_bytes = 0;
    out->peak_write_bound = 0;
    out->residues.assign(L, 0);
    if (ev_start_) HIP_TRY(hipEventRecord(ev_start_, stream_));
  }
  const auto tk0 = Clock::now();
  LAUNCH("fill", 8.0 * ncount, launch_fill(counters_rows_, ncount, 0, stream_));
  if (nslice) LAUNCH("fill", 8.0 * nslice, launch_fill(ssum, nslice, 0, stream_));

  const unsigned f = spec.first_row;
  const bool orbit = spec.orbit_first_row && f == 0;
  const uint64_t* skeys = seed_;
  const uint32_t* sres = (const uint32_t*)(seed_ + S0);
  uint64_t* c0 = row_counters(f);
  if (orbit) {
    LAUNCH("count", 8.0 * S0, launch_count(skeys, S0, g, c0 + kMatchedSlot, stream_));
    g.orbit_row = 1;
    g.orbit_mask = (uint32_t)mask_bits((n - 1) / 2 + 1);  // columns c <= n-1-c
    g.centre_bit = (n % 2) ? (1u << (n / 2)) : 0u;
  }
  LAUNCH("count", 8.0 * S0, launch_count(skeys, S0, g, c0, stream_));
  uint64_t C;
  u128 matched;
  if (eager) {
    HIP_TRY(hipMemcpyAsync(h_counters_, c0, sizeof(uint64_t) * kNumCounters, hipMemcpyDeviceToHost, stream_));
    SYNC_AND_RESOLVE();
    C = h_counters_[kCandidates];
    matched = orbit ? (u128)h_counters_[kMatchedSlot] : (u128)C;
    plan->C_first = C;
    plan->matched_first = matched;
  } else {
    C = plan->C_first;
    matched = plan->matched_first;
  }

  int cur = 0;  // buffer holding the current row's final buckets
  int sidx = 0;
  Buckets bk{nullptr, 0, 1};  // the current row's final buckets (segs_[sidx])
  std::vector<uint64_t> final_res(L, 0);
  uint64_t S_in = S0;
  unsigned levels = 0;
  uint64_t* const spare = row_counters(0) + n * kNumCounters;

  // Geometry of the merge rows (no orbit weights) and the COUNT digit width.
  RowGeom gm = g;
  gm.orbit_row = 0;
  gm.orbit_mask = gm.board_mask;
  gm.centre_bit = 0;
  const uint32_t CB = std::min<uint32_t>(kFineBits, std::max(1u, spec.expand_bits));
  // Set when the last partition level of the current row summed its
  // records' fanout into fan_ (for a merge launch of fan_W workgroups).
  bool fan_ready = false;
  uint32_t fan_W = 0;

  // Reads one counter block back (eager runs only).
  auto read_block = [&](const uint64_t* dev) -> int {
    HIP_TRY(hipMemcpyAsync(h_counters_, dev, sizeof(uint64_t) * kNumCounters, hipMemcpyDeviceToHost, stream_));
    SYNC_AND_RESOLVE();
    return 0;
  };

  // Partition levels pd after the level-1 digit d1 of the buckets `in` in
  // region `from` (or, with `runs`, of the SCATTER's runs there).  A level
  // writes fixed-capacity buckets by per-stage claims when the region leaves
  // every bucket a wide margin over its mean fill; otherwise, or when a
  // claim overflowed (eager runs repeat the level), it takes the exact
  // histogram + scan path.  pfix/pcap record each level's capacity (0:
  // exact).  With `fan`, the last level also sums the records' fanout for
  // the merge of these buckets into fan_, and their total into
  // rowc[kFanout].
  auto partition_to_final = [&](int from, Buckets in, unsigned d1, const std::vector<unsigned>& pd,
                                std::vector<uint64_t>* pcap, uint64_t Crow, const Runs* runs, uint64_t* rowc,
                                bool fan) -> int {
    cur = from;
    bk = in;
    levels = d1 ? 1 : 0;
    unsigned used = d1;
    const Runs* rn = runs;  // the first level only
    fan_ready = false;
    if (eager) pcap->assign(pd.size(), 0);
    for (size_t li = 0; li < pd.size(); ++li) {
      const unsigned dk = pd[li];
      used += dk;
      const int to = cur ^ 1;
      const uint64_t nseg = bk.nseg;
      const uint64_t nout = nseg << dk;
      const bool last_level = li + 1 == pd.size();
      uint32_t chunks = partition_chunks(Crow, nseg, rn != nullptr);
      if (rn) chunks = std::max<uint32_t>(chunks, (rn->per_seg + kMaxChunkRuns - 1) / kMaxChunkRuns);
      // Twice the records where the arena allows, so that fixed-capacity
      // buckets keep their margin.
      int e = grow(to, std::max<uint64_t>(Crow * rb, std::min<uint64_t>(2 * Crow * rb, fixed_ ? 0 : region_cap())));
      if (!e) e = need(&segs_[sidx ^ 1], &segs_cap_[sidx ^ 1], &plan->segs_need[sidx ^ 1], nout + 1);
      if (e) return e;
      uint64_t out_cap = 0;
      if (eager) {
        // Fixed capacity when each bucket's share of the region exceeds its
        // mean fill by 10 standard deviations of a uniform hash plus 64.
        const uint64_t cap = (region_bytes_[to] / rb) / nout;
        const double mu = (double)Crow / (double)nout;
        if ((double)cap >= mu + 10.0 * std::sqrt(mu) + 64.0) out_cap = std::min<uint64_t>(cap, (uint64_t)1 << 40);
        if (spec.force & kForceExactLevels) out_cap = 0;
        if (out_cap && (spec.force & kForceLevelOverflow)) out_cap = std::max<uint64_t>(1, (uint64_t)(mu / 2));
      } else {
        out_cap = (*pcap)[li];
      }
      for (int attempt = 0; attempt < 2; ++attempt) {
        if (out_cap) {
          plan->region_need = std::max<size_t>(plan->region_need, nout * out_cap * rb);
          LAUNCH("fill", 8.0 * nout, launch_fill(segs_[sidx ^ 1], nout, 0, stream_));
          FanSum fs{};
          const bool with_fan = fan && last_level;
          if (with_fan) {
            const uint32_t W = (uint32_t)std::min<uint64_t>(nout, merge_grid);
            e = need(&fan_, &fan_cap_, &plan->fan_need, 3 * (size_t)W + 2);
            if (e) return e;
            LAUNCH("fill", 8.0 * W, launch_fill(fan_, W, 0, stream_));
            fs = FanSum{fan_, rowc + kFanout, W, (uint32_t)((nout + W - 1) / W), gm};
            fan_W = W;
          }
          LAUNCH(rn ? "part_scatter_runs" : "part_scatter", 2.0 * R * Crow,
                 launch_part_scatter(LK, region_[cur], region_[to], bk, chunks, 64 - used, dk, g.key_bits, rn,
                                     segs_[sidx ^ 1], out_cap, region_bytes_[to] / rb, spare,
                                     with_fan ? &fs : nullptr, stream_));
          if (eager) {
            e = read_block(spare);
            if (e) return e;
            if (h_counters_[kRetry] & 32) {
              // A bucket outgrew its capacity: clear the flag and repeat
              // the level exactly (its input is intact).
              LAUNCH("fill", 8.0, launch_fill(spare + kRetry, 1, 0, stream_));
              if (with_fan) LAUNCH("fill", 8.0, launch_fill(rowc + kFanout, 1, 0, stream_));
              out_cap = 0;
              continue;
            }
          }
          fan_ready = with_fan;
          break;
        }
        // Exact level: histogram, scan, scatter at the scanned offsets.
        const size_t cells = (size_t)nout * chunks;
        e = need(&hist_, &hist_cap_, &plan->hist_need, cells);
        if (!e) e = need(&offsets_, &offsets_cap_, &plan->offsets_need, cells);
        if (!e) e = need(&partial_, &partial_cap_, &plan->partial_need, cells / 4096 + 2);
        if (e) return e;
        LAUNCH(rn ? "part_hist_runs" : "part_hist", 8.0 * Crow,
               launch_part_hist(LK, region_[cur], bk, chunks, 64 - used, dk, g.key_bits, rn, hist_, stream_));
        LAUNCH("scan", 12.0 * cells,
               launch_scan(hist_, (uint32_t)nout, chunks, 1, offsets_, segs_[sidx ^ 1], partial_, spare,
                           stream_));
        LAUNCH(rn ? "part_scatter_runs" : "part_scatter", 2.0 * R * Crow,
               launch_part_scatter(LK, region_[cur], region_[to], bk, chunks, 64 - used, dk, g.key_bits, rn,
                                   offsets_, 0, region_bytes_[to] / rb, spare, nullptr, stream_));
        break;
      }
      if (eager) (*pcap)[li] = out_cap;
      cur = to;
      sidx ^= 1;
      bk = Buckets{segs_[sidx], out_cap, (uint32_t)nout};
      levels += 1;
      rn = nullptr;
    }
    return 0;
  };

  if (C) {
    const unsigned B = total_bits(C);
    const unsigned d1 = eager ? std::min<unsigned>(B, kFineBits) : plan->d1_first;
    const std::vector<unsigned> pd = eager ? split_bits(B, d1) : plan->pdigits_first;
    uint64_t tiles = plan->tiles, per_tile = plan->per_tile;
    if (eager) {
      tiles = std::min<uint64_t>(1024, std::max<uint64_t>(1, (S0 + 4095) / 4096));
      per_tile = (S0 + tiles - 1) / tiles;
      tiles = (S0 + per_tile - 1) / per_tile;
      plan->tiles = tiles;
      plan->per_tile = per_tile;
      plan->d1_first = d1;
      plan->pdigits_first = pd;
    }
    const uint32_t D1 = 1u << d1;
    rc = grow(0, C * rb);
    if (!rc) rc = need(&hist_, &hist_cap_, &plan->hist_need, tiles * D1);
    if (!rc) rc = need(&offsets_, &offsets_cap_, &plan->offsets_need, tiles * D1);
    if (!rc) rc = need(&partial_, &partial_cap_, &plan->partial_need, (tiles * D1) / 4096 + 2);
    if (!rc) rc = need(&segs_[0], &segs_cap_[0], &plan->segs_need[0], (size_t)D1 + 1);
    if (rc) return rc;
    sidx = 0;
    LAUNCH("expand_hist", 8.0 * S0,
           launch_expand_hist(skeys, S0, g, d1, (uint32_t)tiles, per_tile, hist_, stream_));
    LAUNCH("scan", 12.0 * tiles * D1,
           launch_scan(hist_, D1, (uint32_t)tiles, 1, offsets_, segs_[0], partial_, c0, stream_));
    LAUNCH("expand_scatter", R * (S0 + C),
           launch_expand_scatter(LK, skeys, sres, S0, S0, g, pr, d1, (uint32_t)tiles, per_tile, offsets_,
                                 region_[0], region_bytes_[0] / rb, spare, stream_));
    rc = partition_to_final(0, Buckets{segs_[0], 0, D1}, d1, pd, &plan->pcap_first, C, nullptr, c0,
                            f + 1 < n && C >= kUpperBoundRows);
    if (rc) return rc;
  }
  g.orbit_row = 0;
  g.orbit_mask = g.board_mask;
  g.centre_bit = 0;

  for (unsigned row = f; row < n; ++row) {
    const auto rt0 = Clock::now();
    if (eager) plan->rows.emplace_back();
    RowPlan& rp = plan->rows[row - f];
    LayerOut lo{};
    lo.row = row;
    lo.input_states = S_in;
    lo.candidates = (u128)n * S_in;
    lo.matched = matched;
    lo.completed = C;
    uint64_t* cr = row_counters(row);
    const bool last = row + 1 == n;
    uint64_t C_next = 0;
    if (C) {
      const uint32_t W = eager ? (uint32_t)std::min<uint64_t>(bk.nseg, merge_grid) : rp.W;
      rp.W = W;
      if (last) {
        LAUNCH("merge_last", R * C,
               launch_merge_expand(kMergeLast, LK, region_[cur], bk, W, g, pr, 0, nullptr, nullptr, nullptr,
                                   nullptr, 0, final_mask, ssum, nslice / L, cr, table_slots, stream_));
      } else {
        // How the SCATTER finds where to write (rows below 2^20 candidates
        // count the merged states' successors exactly; larger rows write
        // capped runs sized from the fanout sums, or, when those are not
        // available or a run overflowed, runs sized by an upper-bound count
        // of every record's successors).
        const uint32_t DF = 1u << CB;
        const int nsidx = sidx ^ 1;
        const int to = cur ^ 1;
        int mode = rp.mode;
        uint64_t fan_total = 0;
        if (eager) {
          mode = C < kUpperBoundRows ? RowPlan::kExact : (fan_ready && fan_W == W ? RowPlan::kCapped : RowPlan::kUpper);
          if (mode == RowPlan::kCapped && (spec.force & kForceUpperRuns)) mode = RowPlan::kUpper;
          if (mode == RowPlan::kCapped) {
            rc = read_block(cr);
            if (rc) return rc;
            fan_total = h_counters_[kFanout];
            if (total_bits(fan_total) < 2) mode = RowPlan::kExact;  // few or no successors
          }
        } else if (mode == RowPlan::kCapped && !(fan_ready && fan_W == W)) {
          *err = "internal: replayed fanout sums do not match row " + std::to_string(row + 1);
          return kRcInternal;
        }
        fan_ready = false;
        rc = need(&hist_, &hist_cap_, &plan->hist_need, (size_t)DF * W);
        if (!rc) rc = need(&offsets_, &offsets_cap_, &plan->offsets_need, (size_t)DF * W);
        if (!rc) rc = need(&partial_, &partial_cap_, &plan->partial_need, ((size_t)DF * W) / 4096 + 2);
        if (!rc) rc = need(&segs_[nsidx], &segs_cap_[nsidx], &plan->segs_need[nsidx], (size_t)DF + 1);
        if (!rc) rc = need(&run_start_, &run_start_cap_, &plan->run_start_need, (size_t)DF * W);
        if (!rc) rc = need(&run_len_, &run_len_cap_, &plan->run_len_need, (size_t)DF * W);
        if (rc) return rc;
        uint64_t* cn = row_counters(row + 1);
        unsigned d1 = rp.d1;
        uint64_t C_count = rp.C_count;  // records the SCATTER's layout spans
        for (int attempt = 0; attempt < 3; ++attempt) {
          if (mode == RowPlan::kCapped) {
            if (eager) {
              const unsigned B = total_bits(fan_total);
              d1 = std::min<unsigned>(B - 1, CB);
            }
            LAUNCH("caps", 8.0 * W,
                   launch_caps(fan_, W, d1, (spec.force & kForceCapOverflow) ? 6 : 0, fan_ + W, fan_ + 2 * W,
                               run_start_, cr, stream_));
            if (eager) {
              rc = read_block(cr);
              if (rc) return rc;
              C_count = h_counters_[kCapExtent];
              if (C_count * rb > std::max<uint64_t>(region_bytes_[to], region_cap())) {
                mode = RowPlan::kUpper;  // the caps do not fit: size the runs exactly
                continue;
              }
            }
          } else if (mode == RowPlan::kUpper) {
            LAUNCH("count_ub", 8.0 * C, launch_count_ub(LK, region_[cur], bk, W, g, CB, hist_, stream_));
            LAUNCH("scan", 12.0 * DF * W,
                   launch_scan(hist_, DF, W, 1, run_start_, segs_[nsidx], partial_, cn, stream_));
            if (eager) {
              rc = read_block(cn);
              if (rc) return rc;
              C_count = h_counters_[kCandidates];
              const unsigned B = total_bits(C_count);
              d1 = B ? std::min<unsigned>(B - 1, CB) : 0;
            }
            if (d1 && d1 < CB)
              LAUNCH("scan", 12.0 * DF * W,
                     launch_scan(hist_, 1u << d1, W, 1u << (CB - d1), run_start_, segs_[nsidx], partial_,
                                 cn, stream_));
          } else {
            LAUNCH("merge_count", 8.0 * C,
                   launch_merge_expand(kMergeCount, LK, region_[cur], bk, W, g, pr, CB, hist_, nullptr, nullptr,
                                       nullptr, 0, 0, nullptr, 0, cr, table_slots, stream_));
            LAUNCH("scan", 12.0 * DF * W,
                   launch_scan(hist_, DF, W, 1, offsets_, segs_[nsidx], partial_, cn, stream_));
            if (eager) {
              rc = read_block(cn);
              if (rc) return rc;
              C_count = h_counters_[kCandidates];
              d1 = std::min<unsigned>(total_bits(C_count), CB);
            }
            if (C_count && d1 < CB)
              LAUNCH("scan", 12.0 * DF * W,
                     launch_scan(hist_, 1u << d1, W, 1u << (CB - d1), offsets_, segs_[nsidx], partial_,
                                 cn, stream_));
          }
          if (eager) out->peak_write_bound = std::max(out->peak_write_bound, C_count);
          if (!C_count) break;
          const bool runs_out = mode != RowPlan::kExact;
          rc = grow(to, C_count * rb);
          if (rc) return rc;
          LAUNCH("merge_scatter", R * (C + C_count),
                 launch_merge_expand(kMergeScatter, LK, region_[cur], bk, W, g, pr, d1,
                                     runs_out ? run_len_ : nullptr, runs_out ? run_start_ : offsets_,
                                     mode == RowPlan::kCapped ? fan_ + W : nullptr, region_[to],
                                     region_bytes_[to] / rb, final_mask, ssum, nslice / L, cr, table_slots,
                                     stream_));
          if (eager) {
            HIP_TRY(hipMemcpyAsync(h_counters_, cr, sizeof(uint64_t) * kNumCounters, hipMemcpyDeviceToHost,
                                   stream_));
            HIP_TRY(hipStreamSynchronize(stream_));
            // Runs with gaps: the SCATTER's algorithmic bytes are its exact
            // writes, known only now.
            if (profiling_ && runs_out && !pending_.empty() && pending_.back().kernel == kernel_id("merge_scatter"))
              pending_.back().bytes = R * (double)(C + h_counters_[kWritten]);
            SYNC_AND_RESOLVE();
            if (h_counters_[kRetry] & 16) {
              // A capped run overflowed: forget this SCATTER's statistics
              // and repeat it with upper-bound offsets (its input is intact).
              LAUNCH("fill", 8.0 * kNumCounters, launch_fill(cr, kNumCounters, 0, stream_));
              mode = RowPlan::kUpper;
              continue;
            }
            C_next = runs_out ? h_counters_[kWritten] : C_count;
          } else {
            C_next = rp.C_next;
          }
          break;
        }
        rp.mode = mode;
        rp.d1 = d1;
        rp.C_count = C_count;
        if (C_count && C_next) {
          sidx = nsidx;
          std::vector<unsigned> pd = rp.pdigits;
          if (eager) {
            // Levels after the SCATTER's digit: runs with gaps need at least
            // one level to close them.
            const unsigned B = total_bits(C_next);
            pd = split_bits(mode != RowPlan::kExact ? std::max(B, d1 + 1) : B, d1);
            rp.pdigits = pd;
          }
          rp.C_next = C_next;
          const Runs runs{run_start_, run_len_, W};
          rc = partition_to_final(to, Buckets{segs_[sidx], 0, 1u << d1}, d1, pd, &rp.pcap, C_next,
                                  mode != RowPlan::kExact ? &runs : nullptr, cn,
                                  row + 2 < n && C_next >= kUpperBoundRows);
          if (rc) return rc;
        } else {
          C_next = 0;
          rp.C_next = 0;
          if (!C_count) {
            // No successor survives: merge for the statistics only.
            LAUNCH("merge_last", R * C,
                   launch_merge_expand(kMergeLast, LK, region_[cur], bk, W, g, pr, 0, nullptr, nullptr, nullptr,
                                       nullptr, 0, final_mask, ssum, nslice / L, cr, table_slots, stream_));
          }
        }
      }
      if (eager) {
        rc = read_block(cr);
        if (rc) return rc;
        if (h_counters_[kError] & 4) return kRcPackedOverflow;
        if (h_counters_[kError] || h_counters_[kRetry]) {
          *err = error_flag_text(h_counters_[kError] | h_counters_[kRetry], " in row " + std::to_string(row + 1));
          return kRcInternal;
        }
        std::vector<uint64_t> w(L);
        for (int l = 0; l < L; ++l) {
          w[l] = h_counters_[kWeight0 + l] % spec.primes[l];
          final_res[l] = h_counters_[kFinal0 + l] % spec.primes[l];
        }
        std::string cerr;
        if (!crt_reconstruct(w.data(), spec.primes.data(), L, &lo.output_weight, &cerr)) {
          *err = cerr;
          return 2;
        }
        if (lo.output_weight > fact) {
          *err = "CRT reconstruction exceeds certified N! count bound after row " + std::to_string(row + 1);
          return 2;
        }
        lo.output_states = h_counters_[kOutStates];
      }
    }
    if (eager) {
      lo.levels = C ? levels : 0;
      lo.buckets = C ? bk.nseg : 0;
      lo.device_bytes = arena_bytes() + hist_cap_ * 4 + offsets_cap_ * 8 + (segs_cap_[0] + segs_cap_[1]) * 8;
      lo.elapsed_s = std::chrono::duration<double>(Clock::now() - rt0).count();
      out->layers.push_back(lo);
      out->total_candidates += lo.candidates;
      out->total_matched += lo.matched;
      out->peak_states = std::max<uint64_t>(out->peak_states, lo.output_states);
      out->device_peak_bytes = std::max<uint64_t>(out->device_peak_bytes, lo.device_bytes);
      S_in = lo.output_states;
    }
    C = C_next;
    matched = C_next;
  }

  if (!eager) return 0;

  if (ev_start_) {
    HIP_TRY(hipEventRecord(ev_stop_, stream_));
    HIP_TRY(hipEventSynchronize(ev_stop_));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev_start_, ev_stop_));
    out->kernel_s = ms / 1000.0;
  } else {
    // no events (the no-allocation engine): host time of the row pipeline
    HIP_TRY(hipStreamSynchronize(stream_));
    out->kernel_s = std::chrono::duration<double>(Clock::now() - tk0).count();
  }
  // The partition and expansion kernels flag capacity violations in the
  // spare counter row n.
  std::vector<uint64_t> sp(kNumCounters);
  HIP_TRY(hipMemcpy(sp.data(), spare, sizeof(uint64_t) * kNumCounters, hipMemcpyDeviceToHost));
  if (sp[kError] & 4) return kRcPackedOverflow;
  if (sp[kError] || sp[kRetry]) {
    *err = error_flag_text(sp[kError] | sp[kRetry], " (candidate buffers)");
    return kRcInternal;
  }
  // Final v1/v2 contraction: the merges summed the residues of the states
  // whose columns are all set.
  u128 count = 0;
  const bool alive = !out->layers.empty() && out->layers.back().output_states > 0;
  if (alive) {
    std::string cerr;
    if (!crt_reconstruct(final_res.data(), spec.primes.data(), L, &count, &cerr)) {
      *err = cerr;
      return 2;
    }
    if (count > fact) {
      *err = "CRT reconstruction exceeds certified N! count bound";
      return 2;
    }
  }
  out->count = count;
  for (int l = 0; l < L; ++l) out->residues[l] = alive ? final_res[l] : 0;
  if (nslice) {
    std::vector<uint64_t> sums(nslice);
    HIP_TRY(hipMemcpy(sums.data(), slice_sums_, nslice * 8, hipMemcpyDeviceToHost));
    out->slice_counts.assign(nslice / L, 0);
    std::vector<uint64_t> r(L);
    for (size_t sl = 0; sl < nslice / L; ++sl) {
      for (int l = 0; l < L; ++l) r[l] = sums[sl * L + l] % spec.primes[l];
      std::string cerr;
      if (!crt_reconstruct(r.data(), spec.primes.data(), L, &out->slice_counts[sl], &cerr)) {
        *err = cerr;
        return 2;
      }
    }
  }
  return 0;
}

// Layer statistics of a replayed run, read from the per-row counters.
int Engine::finish(const RunSpec& spec, const Plan& plan, RunOut* out, std::string* err) {
  const unsigned n = spec.n, f = spec.first_row;
  const int L = (int)spec.primes.size();
  u128 fact;
  factorial(n, &fact);
  std::vector<uint64_t> cnt((size_t)(n + 1) * kNumCounters);
  HIP_TRY(hipMemcpy(cnt.data(), counters_rows_, cnt.size() * 8, hipMemcpyDeviceToHost));
  for (unsigned row = 0; row <= n; ++row)
    if (cnt[(size_t)row * kNumCounters + kError] & 4) return kRcPackedOverflow;
  if (cnt[(size_t)n * kNumCounters + kError] || cnt[(size_t)n * kNumCounters + kRetry]) {
    *err = error_flag_text(cnt[(size_t)n * kNumCounters + kError] | cnt[(size_t)n * kNumCounters + kRetry],
                           " (candidate buffers, replayed size plan)");
    return kRcInternal;
  }
  out->record_bytes = (uint32_t)record_bytes(packed_records(spec) ? kPackedRecords : L);
  out->layers.clear();
  out->peak_states = spec.seeds.size();
  out->total_candidates = out->total_matched = 0;
  out->residues.assign(L, 0);
  out->peak_write_bound = 0;
  for (const RowPlan& rp : plan.rows) out->peak_write_bound = std::max(out->peak_write_bound, rp.C_count);
  uint64_t S_in = spec.seeds.size(), C = plan.C_first;
  u128 matched = plan.matched_first;
  std::vector<uint64_t> fin(L, 0);
  for (unsigned row = f; row < n; ++row) {
    const uint64_t* c = &cnt[(size_t)row * kNumCounters];
    const RowPlan& rp = plan.rows[row - f];
    LayerOut lo{};
    lo.row = row;
    lo.input_states = S_in;
    lo.candidates = (u128)n * S_in;
    lo.matched = matched;
    lo.completed = C;
    if (c[kError] || c[kRetry]) {
      *err = error_flag_text(c[kError] | c[kRetry], " in row " + std::to_string(row + 1) + " (replayed size plan)");
      return kRcInternal;
    }
    if (row + 1 < n && C && rp.C_next && c[kWritten] != rp.C_next) {
      *err = "replayed size plan does not match the device counts";
      return kRcInternal;
    }
    if (C) {
      std::vector<uint64_t> w(L);
      for (int l = 0; l < L; ++l) {
        w[l] = c[kWeight0 + l] % spec.primes[l];
        fin[l] = c[kFinal0 + l] % spec.primes[l];
      }
      std::string cerr;
      if (!crt_reconstruct(w.data(), spec.primes.data(), L, &lo.output_weight, &cerr)) {
        *err = cerr;
        return 2;
      }
      lo.output_states = c[kOutStates];
    }
    out->layers.push_back(lo);
    out->total_candidates += lo.candidates;
    out->total_matched += lo.matched;
    out->peak_states = std::max<uint64_t>(out->peak_states, lo.output_states);
    S_in = lo.output_states;
    C = rp.C_next;
    matched = rp.C_next;
  }
  u128 count = 0;
  const bool alive = out->layers.back().output_states > 0;
  if (alive) {
    std::string cerr;
    if (!crt_reconstruct(fin.data(), spec.primes.data(), L, &count, &cerr)) {
      *err = cerr;
      return 2;
    }
    if (count > fact) {
      *err = "CRT reconstruction exceeds certified N! count bound";
      return 2;
    }
  }
  out->count = count;
  for (int l = 0; l < L; ++l) out->residues[l] = alive ? fin[l] : 0;
  out->device_peak_bytes = arena_bytes();
  if (spec.per_slice) {
    const size_t ns = ((size_t)1 << spec.slice_bits);
    std::vector<uint64_t> sums(ns * L);
    HIP_TRY(hipMemcpy(sums.data(), slice_sums_, sums.size() * 8, hipMemcpyDeviceToHost));
    out->slice_counts.assign(ns, 0);
    std::vector<uint64_t> r(L);
    for (size_t sl = 0; sl < ns; ++sl) {
      for (int l = 0; l < L; ++l) r[l] = sums[sl * L + l] % spec.primes[l];
      std::string cerr;
      if (!crt_reconstruct(r.data(), spec.primes.data(), L, &out->slice_counts[sl], &cerr)) {
        *err = cerr;
        return 2;
      }
    }
  }
  return 0;
}

// Materialised row loop: the merged frontier of every row is written to HBM
// (front_, SoA) and expanded from there.  Used for the literal sitewise C
// transfer (contract_rows_sitewise, src/lib.rs:5629-5750) and for layer
// traces; each row synchronises with the host.
int Engine::materialized(const RunSpec& spec, RunOut* out, std::string* err) {
  const unsigned n = spec.n;
  const int L = (int)spec.primes.size();
  const size_t rb = record_bytes(L);
  const size_t fb = 8 + 4 * (size_t)L;  // frontier bytes per state
  const double R = 8.0 + 4.0 * L;
  const bool sitewise = spec.site != nullptr;
  if (sitewise && spec.orbit_first_row) {
    *err = "the literal sitewise row transfer runs without row-0 orbit slicing";
    return 1;
  }
  if (spec.per_slice) {
    *err = "materialised rows do not report per-slice counts";
    return 1;
  }
  Primes pr{};
  for (int l = 0; l < L; ++l) pr.p[l] = (uint32_t)spec.primes[l];
  pr.lanes = (uint32_t)L;
  out->record_bytes = (uint32_t)record_bytes(L);
  u128 fact;
  factorial(n, &fact);

  RowGeom g0{};
  g0.n = n;
  g0.board_mask = (uint32_t)mask_bits(n);
  g0.state_mask = mask_bits(3 * n);
  g0.orbit_mask = g0.board_mask;
  g0.col_in = occupied_.column_in ? g0.board_mask : 0u;
  g0.col_out = occupied_.column_out ? g0.board_mask : 0u;
  g0.dr_in = occupied_.diag_dr_in ? g0.board_mask : 0u;
  g0.dr_out = occupied_.diag_dr_out ? g0.board_mask : 0u;
  g0.dl_in = occupied_.diag_dl_in ? g0.board_mask : 0u;
  g0.dl_out = occupied_.diag_dl_out ? g0.board_mask : 0u;
  const uint64_t final_mask = g0.board_mask;

  const uint32_t rbits = std::max<uint32_t>(1, std::min<uint32_t>(spec.radix_bits, kMaxDigitBits));
  const uint32_t table_slots = spec.tile_states <= merge_cap(2048) ? 2048u : 4096u;
  const uint64_t target = (uint64_t)std::max<uint32_t>(256, std::min<uint32_t>(spec.tile_states, merge_cap(table_slots)));
  const uint64_t merge_grid = (uint64_t)cu_count_ * merge_resident_blocks(L, table_slots, kFineBits) *
                              std::max<uint32_t>(1, std::min<uint32_t>(spec.merge_waves, 8));

  const size_t ncount = (size_t)(n + 1) * kNumCounters;
  int rc = ensure_buffer(&counters_rows_, &counters_rows_cap_, ncount, err);
  if (rc) return rc;
  uint64_t* spare = counters_rows_ + (size_t)n * kNumCounters;

  // Seed frontier, stride S0.
  const uint64_t S0 = spec.seeds.size();
  rc = ensure_buffer(&front_, &front_cap_, (S0 * fb + 7) / 8, err);
  if (rc) return rc;
  {
    std::vector<uint64_t> host((S0 * fb + 7) / 8, 0);
    uint32_t* res = (uint32_t*)(host.data() + S0);
    for (uint64_t i = 0; i < S0; ++i) {
      host[i] = spec.seeds[i].key;
      for (int l = 0; l < L; ++l) res[(size_t)l * S0 + i] = (uint32_t)(spec.seeds[i].weight % spec.primes[l]);
    }
    HIP_TRY(hipStreamSynchronize(stream_));
    HIP_TRY(hipMemcpy(front_, host.data(), S0 * fb, hipMemcpyHostToDevice));
  }
  out->layers.clear();
  out->peak_states = S0;
  out->total_candidates = out->total_matched = 0;
  out->device_peak_bytes = 0;
  out->residues.assign(L, 0);
  HIP_TRY(hipEventRecord(ev_start_, stream_));
  LAUNCH("fill", 8.0 * ncount, launch_fill(counters_rows_, ncount, 0, stream_));

  std::vector<uint64_t> h(kNumCounters);
  auto read_counters = [&](const uint64_t* dev) -> int {
    HIP_TRY(hipMemcpyAsync(h_counters_, dev, sizeof(uint64_t) * kNumCounters, hipMemcpyDeviceToHost, stream_));
    SYNC_AND_RESOLVE();
    std::copy(h_counters_, h_counters_ + kNumCounters, h.begin());
    return 0;
  };

  uint64_t S = S0, stride = S0;
  std::vector<uint64_t> final_res(L, 0);
  for (unsigned row = spec.first_row; row < n; ++row) {
    const auto rt0 = Clock::now();
    LayerOut lo{};
    lo.row = row;
    lo.input_states = S;
    uint64_t* cr = counters_rows_ + (size_t)row * kNumCounters;
    const bool orbit = !sitewise && spec.orbit_first_row && row == 0;
    RowGeom g = g0;
    const uint64_t* keys = front_;
    const uint32_t* res = (const uint32_t*)(front_ + stride);
    uint64_t C = 0;
    if (S) {
      if (sitewise) {
        LAUNCH("site_count", 8.0 * S, launch_site_count(keys, S, g, *spec.site, cr, stream_));
      } else {
        if (orbit) {
          LAUNCH("count", 8.0 * S, launch_count(keys, S, g, cr + kMatchedSlot, stream_));
          g.orbit_row = 1;
          g.orbit_mask = (uint32_t)mask_bits((n - 1) / 2 + 1);
          g.centre_bit = (n % 2) ? (1u << (n / 2)) : 0u;
        }
        LAUNCH("count", 8.0 * S, launch_count(keys, S, g, cr, stream_));
      }
      rc = read_counters(cr);
      if (rc) return rc;
      if (h[kError] & 8) {
        *err = "a parent needs more than " + std::to_string(kSitePartials) +
               " partial rows in the sitewise C contraction of row " + std::to_string(row + 1);
        return 1;
      }
      C = h[kCandidates];
      if (sitewise) {
        lo.site_visits = h[kSiteVisits];
        lo.site_matched = h[kSiteMatched];
      } else {
        lo.candidates = (u128)n * S;
        lo.matched = orbit ? (u128)h[kMatchedSlot] : (u128)C;
      }
    }
    lo.completed = C;
    uint64_t S_out = 0;
    if (C) {
      const unsigned B = C <= target ? 0u : ceil_log2((C + target - 1) / target);
      const unsigned d1 = std::min<unsigned>(B, kFineBits);
      std::vector<unsigned> pd;
      for (unsigned used = d1; used < B;) {
        const unsigned dk = std::min<unsigned>(rbits, B - used);
        pd.push_back(dk);
        used += dk;
      }
      const uint64_t per_unit = sitewise ? 1024 : 4096;
      uint64_t tiles = std::min<uint64_t>(sitewise ? 2048 : 1024, std::max<uint64_t>(1, (S + per_unit - 1) / per_unit));
      const uint64_t per_tile = (S + tiles - 1) / tiles;
      tiles = (S + per_tile - 1) / per_tile;
      const uint32_t D1 = 1u << d1;
      rc = ensure_region(0, C * rb, err);
      if (!rc) rc = ensure_buffer(&hist_, &hist_cap_, tiles * D1, err);
      if (!rc) rc = ensure_buffer(&offsets_, &offsets_cap_, tiles * D1, err);
      if (!rc) rc = ensure_buffer(&partial_, &partial_cap_, (tiles * D1) / 4096 + 2, err);
      if (!rc) rc = ensure_buffer(&segs_[0], &segs_cap_[0], (size_t)D1 + 1, err);
      if (rc) return rc;
      if (sitewise) {
        LAUNCH("site_hist", 8.0 * S,
               launch_site_hist(keys, S, g, *spec.site, d1, (uint32_t)tiles, per_tile, hist_, stream_));
      } else {
        LAUNCH("expand_hist", 8.0 * S,
               launch_expand_hist(keys, S, g, d1, (uint32_t)tiles, per_tile, hist_, stream_));
      }
      LAUNCH("scan", 12.0 * tiles * D1,
             launch_scan(hist_, D1, (uint32_t)tiles, 1, offsets_, segs_[0], partial_, cr, stream_));
      if (sitewise) {
        LAUNCH("site_scatter", R * (S + C),
               launch_site_scatter(L, keys, res, S, stride, g, *spec.site, pr, d1, (uint32_t)tiles, per_tile,
                                   offsets_, region_[0], region_bytes_[0] / rb, spare, stream_));
      } else {
        LAUNCH("expand_scatter", R * (S + C),
               launch_expand_scatter(L, keys, res, S, stride, g, pr, d1, (uint32_t)tiles, per_tile, offsets_,
                                     region_[0], region_bytes_[0] / rb, spare, stream_));
      }
      int cur = 0, sidx = 0;
      uint64_t nseg = (uint64_t)D1;
      unsigned used = d1;
      for (unsigned dk : pd) {
        used += dk;
        const int to = cur ^ 1;
        const uint32_t chunks = partition_chunks(C, nseg);
        const size_t cells = (size_t)(nseg << dk) * chunks;
        rc = ensure_region(to, C * rb, err);
        if (!rc) rc = ensure_buffer(&segs_[sidx ^ 1], &segs_cap_[sidx ^ 1], (size_t)(nseg << dk) + 1, err);
        if (!rc) rc = ensure_buffer(&hist_, &hist_cap_, cells, err);
        if (!rc) rc = ensure_buffer(&offsets_, &offsets_cap_, cells, err);
        if (!rc) rc = ensure_buffer(&partial_, &partial_cap_, cells / 4096 + 2, err);
        if (rc) return rc;
        LAUNCH("part_hist", 8.0 * C,
               launch_part_hist(L, region_[cur], Buckets{segs_[sidx], 0, (uint32_t)nseg}, chunks, 64 - used, dk, 0,
                                nullptr, hist_, stream_));
        LAUNCH("scan", 12.0 * cells,
               launch_scan(hist_, (uint32_t)(nseg << dk), chunks, 1, offsets_, segs_[sidx ^ 1], partial_, spare,
                           stream_));
        LAUNCH("part_scatter", 2.0 * R * C,
               launch_part_scatter(L, region_[cur], region_[to], Buckets{segs_[sidx], 0, (uint32_t)nseg}, chunks,
                                   64 - used, dk, 0, nullptr, offsets_, 0, region_bytes_[to] / rb, spare, nullptr,
                                   stream_));
        cur = to;
        sidx ^= 1;
        nseg <<= dk;
      }
      lo.levels = (d1 ? 1 : 0) + (uint32_t)pd.size();
      lo.buckets = (uint32_t)nseg;
      // The input frontier has been consumed: the merged states overwrite it.
      rc = ensure_buffer(&front_, &front_cap_, (C * fb + 7) / 8, err);
      if (rc) return rc;
      const uint32_t W = (uint32_t)std::min<uint64_t>(nseg, merge_grid);
      LAUNCH("merge_emit", R * 2.0 * C,
             launch_merge_expand(kMergeEmit, L, region_[cur], Buckets{segs_[sidx], 0, (uint32_t)nseg}, W, g0, pr, 0,
                                 nullptr, nullptr, nullptr, front_, C, final_mask, nullptr, 0, cr, table_slots,
                                 stream_));
      rc = read_counters(cr);
      if (rc) return rc;
      if (h[kError]) {
        *err = error_flag_text(h[kError], " in row " + std::to_string(row + 1));
        return kRcInternal;
      }
      S_out = h[kOutStates];
      if (h[kEmitted] != S_out) {
        *err = "merged state count disagrees with the states written in row " + std::to_string(row + 1);
        return kRcInternal;
      }
      std::vector<uint64_t> w(L);
      for (int l = 0; l < L; ++l) {
        w[l] = h[kWeight0 + l] % spec.primes[l];
        final_res[l] = h[kFinal0 + l] % spec.primes[l];
      }
      std::string cerr;
      if (!crt_reconstruct(w.data(), spec.primes.data(), L, &lo.output_weight, &cerr)) {
        *err = cerr;
        return 2;
      }
      if (spec.bound_check && lo.output_weight > fact) {
        *err = "CRT reconstruction exceeds certified N! count bound after row " + std::to_string(row + 1);
        return 2;
      }
      stride = C;
    } else {
      std::fill(final_res.begin(), final_res.end(), 0);
    }
    lo.output_states = S_out;
    if (spec.on_layer) {
      std::vector<uint64_t> hk(S_out);
      std::vector<uint32_t> hr((size_t)L * S_out);
      if (S_out) {
        HIP_TRY(hipMemcpy(hk.data(), front_, S_out * 8, hipMemcpyDeviceToHost));
        const uint32_t* dres = (const uint32_t*)(front_ + stride);
        for (int l = 0; l < L; ++l)
          HIP_TRY(hipMemcpy(hr.data() + (size_t)l * S_out, dres + (size_t)l * stride, S_out * 4,
                            hipMemcpyDeviceToHost));
      }
      rc = spec.on_layer(row, hk, hr, err);
      if (rc) return rc;
    }
    lo.device_bytes = arena_bytes() + front_cap_ * 8 + hist_cap_ * 4 + offsets_cap_ * 8;
    lo.elapsed_s = std::chrono::duration<double>(Clock::now() - rt0).count();
    out->layers.push_back(lo);
    out->total_candidates += lo.candidates;
    out->total_matched += lo.matched;
    out->peak_states = std::max<uint64_t>(out->peak_states, S_out);
    out->device_peak_bytes = std::max<uint64_t>(out->device_peak_bytes, lo.device_bytes);
    S = S_out;
  }
  HIP_TRY(hipEventRecord(ev_stop_, stream_));
  HIP_TRY(hipEventSynchronize(ev_stop_));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, ev_start_, ev_stop_));
  out->kernel_s = ms / 1000.0;
  std::vector<uint64_t> sp(kNumCounters);
  HIP_TRY(hipMemcpy(sp.data(), spare, sizeof(uint64_t) * kNumCounters, hipMemcpyDeviceToHost));
  if (sp[kError]) {
    *err = error_flag_text(sp[kError], " (candidate buffers)");
    return kRcInternal;
  }
  u128 count = 0;
  if (S) {
    std::string cerr;
    if (!crt_reconstruct(final_res.data(), spec.primes.data(), L, &count, &cerr)) {
      *err = cerr;
      return 2;
    }
    if (spec.bound_check && count > fact) {
      *err = "CRT reconstruction exceeds certified N! count bound";
      return 2;
    }
  }
  out->count = count;
  for (int l = 0; l < L; ++l) out->residues[l] = S ? final_res[l] : 0;
  return 0;
}

}  // namespace peps
