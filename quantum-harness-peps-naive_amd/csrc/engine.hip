// engine.hip -- see engine.h.  Follows the reference's row loop
// (contract_rows_sort_reduce_with_modes, src/lib.rs:5287-5415): per row,
// expand every parent through the compiled operator, apply the row-0 orbit
// weights, merge equal boundaries, record the layer metric; after the last
// row keep the states whose columns are all set (v1) and sum them (v2 frees
// the diagonals).
#include "engine.h"

#include <algorithm>
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

}  // namespace

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
  return region_bytes_[0] + region_bytes_[1] + region_bytes_[2];
}

int Engine::ensure_region(int idx, size_t bytes, std::string* err) {
  if (region_bytes_[idx] >= bytes) return 0;
  size_t want = std::max<size_t>(bytes + bytes / 4, (size_t)1 << 20);
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
  HIP_TRY(hipMalloc(&region_[idx], want));
  region_bytes_[idx] = want;
  return 0;
}

template <typename T>
int Engine::ensure_buffer(T** ptr, size_t* cap, size_t count, std::string* err) {
  if (*cap >= count) return 0;
  size_t want = std::max<size_t>(count + count / 4, 1024);
  if (*ptr) {
    HIP_TRY(hipStreamSynchronize(stream_));
    HIP_TRY(hipFree(*ptr));
    *ptr = nullptr;
  }
  HIP_TRY(hipMalloc((void**)ptr, want * sizeof(T)));
  *cap = want;
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

int Engine::run(const RunSpec& spec, RunOut* out, std::string* err) {
  const auto t0 = Clock::now();
  HIP_TRY(hipSetDevice(device_));
  const unsigned n = spec.n;
  const int L = (int)spec.primes.size();
  if (n == 0 || n > 21 || L < 1 || L > kMaxLanes) {
    *err = "engine: unsupported board or lane count";
    return 1;
  }
  if (3 * n + spec.slice_bits > 63) {
    *err = "engine: slice ids do not fit above the 3N-bit boundary key";
    return 1;
  }
  if (budget_ == 0 || (spec.budget_bytes && spec.budget_bytes != budget_)) {
    if (spec.budget_bytes) {
      budget_ = spec.budget_bytes;
    } else {
      size_t free_b = 0, total_b = 0;
      HIP_TRY(hipMemGetInfo(&free_b, &total_b));
      budget_ = (uint64_t)((double)(free_b + arena_bytes()) * 0.85);
    }
  }
  const size_t rb = record_bytes(L);
  const size_t soa = 8 + 4 * (size_t)L;
  Primes pr{};
  for (int l = 0; l < L; ++l) pr.p[l] = (uint32_t)spec.primes[l];

  RowGeom g{};
  g.n = n;
  g.board_mask = mask_bits(n);
  g.state_mask = mask_bits(3 * n);
  g.col_in = occupied_.column_in;
  g.col_out = occupied_.column_out;
  g.dr_in = occupied_.diag_dr_in;
  g.dr_out = occupied_.diag_dr_out;
  g.dl_in = occupied_.diag_dl_in;
  g.dl_out = occupied_.diag_dl_out;

  // Seed the frontier into region 0.
  int fi = 0;
  uint64_t S = spec.seeds.size();
  uint64_t stride = std::max<uint64_t>(S, 1);
  {
    int rc = ensure_region(fi, stride * soa, err);
    if (rc) return rc;
    std::vector<uint64_t> keys(S);
    std::vector<uint32_t> res((size_t)L * stride, 0);
    for (uint64_t i = 0; i < S; ++i) {
      keys[i] = spec.seeds[i].key;
      for (int l = 0; l < L; ++l) res[(size_t)l * stride + i] = (uint32_t)(spec.seeds[i].weight % spec.primes[l]);
    }
    char* base = (char*)region_[fi];
    if (S) {
      HIP_TRY(hipMemcpyAsync(base, keys.data(), S * 8, hipMemcpyHostToDevice, stream_));
      HIP_TRY(hipMemcpyAsync(base + stride * 8, res.data(), (size_t)L * stride * 4,
                             hipMemcpyHostToDevice, stream_));
    }
  }
  out->layers.clear();
  out->peak_states = std::max<uint64_t>(1, S);
  out->total_candidates = out->total_matched = 0;
  out->device_peak_bytes = 0;
  out->residues.assign(L, 0);
  u128 fact;
  factorial(n, &fact);
  HIP_TRY(hipEventRecord(ev_start_, stream_));

  std::vector<uint64_t> final_res(L, 0);
  for (unsigned row = spec.first_row; row < n; ++row) {
    const auto rt0 = Clock::now();
    const bool orbit = spec.orbit_first_row && row == 0;
    LayerOut lo{};
    lo.row = row;
    lo.input_states = S;
    lo.candidates = (u128)n * S;
    char* fbase = (char*)region_[fi];
    const uint64_t* fkeys = (const uint64_t*)fbase;
    const uint32_t* fres = (const uint32_t*)(fbase + stride * 8);

    g.orbit_row = orbit ? 1u : 0u;
    g.orbit_mask = g.board_mask;
    g.centre_bit = 0;
    uint64_t C = 0;
    HIP_TRY(hipMemsetAsync(counters_, 0, sizeof(uint64_t) * kNumCounters, stream_));
    if (S) {
      if (orbit) {
        LAUNCH("count", 8.0 * S, launch_count(fkeys, S, g, counters_ + kMatchedSlot, stream_));
        g.orbit_mask = mask_bits((n - 1) / 2 + 1);  // columns c <= n-1-c
        g.centre_bit = (n % 2) ? ((uint64_t)1 << (n / 2)) : 0;
      }
      LAUNCH("count", 8.0 * S, launch_count(fkeys, S, g, counters_, stream_));
      HIP_TRY(hipMemcpyAsync(h_counters_, counters_, sizeof(uint64_t) * kNumCounters,
                             hipMemcpyDeviceToHost, stream_));
      SYNC_AND_RESOLVE();
      C = h_counters_[kCandidates];
      lo.matched = orbit ? (u128)h_counters_[kMatchedSlot] : (u128)C;
    }
    lo.completed = C;
    if (S == 0 || C == 0) {
      lo.output_states = 0;
      lo.output_weight = 0;
      S = 0;
      lo.elapsed_s = std::chrono::duration<double>(Clock::now() - rt0).count();
      lo.device_bytes = arena_bytes();
      out->layers.push_back(lo);
      out->total_candidates += lo.candidates;
      out->total_matched += lo.matched;
      std::fill(final_res.begin(), final_res.end(), 0);
      continue;
    }

    // Plan the hash-radix levels so final buckets hold ~tile_states records.
    const uint32_t rbits = std::max<uint32_t>(1, std::min<uint32_t>(spec.radix_bits, 10));
    const uint64_t target = std::max<uint32_t>(256, std::min<uint32_t>(spec.tile_states, 2560));
    unsigned total_bits = C <= target ? 0 : ceil_log2((C + target - 1) / target);
    std::vector<unsigned> digits;
    if (total_bits) {
      const unsigned levels = (total_bits + rbits - 1) / rbits;
      unsigned left = total_bits;
      for (unsigned k = 0; k < levels; ++k) {
        const unsigned d = (left + (levels - k) - 1) / (levels - k);
        digits.push_back(d);
        left -= d;
      }
    } else {
      digits.push_back(0);
    }
    if ((C >> digits[0]) > 0xffffffffull) {
      *err = "row too large for a 32-bit level-1 bucket";
      return 4;
    }
    const int pa = (fi + 1) % 3, pb = (fi + 2) % 3;
    int rc = ensure_region(pa, C * rb, err);
    if (rc) return rc;
    rc = ensure_region(pb, C * rb, err);
    if (rc) return rc;

    // Level 1 fused with the expansion.
    const uint32_t D1 = 1u << digits[0];
    uint64_t tiles = std::min<uint64_t>(1024, std::max<uint64_t>(1, (S + 4095) / 4096));
    const uint64_t per_tile = (S + tiles - 1) / tiles;
    tiles = (S + per_tile - 1) / per_tile;
    rc = ensure_buffer(&hist_, &hist_cap_, tiles * D1, err);
    if (rc) return rc;
    rc = ensure_buffer(&offsets_, &offsets_cap_, tiles * D1, err);
    if (rc) return rc;
    uint64_t nseg_final = 1;
    for (unsigned d : digits) nseg_final <<= d;
    rc = ensure_buffer(&segs_[0], &segs_cap_[0], nseg_final + 1, err);
    if (rc) return rc;
    rc = ensure_buffer(&segs_[1], &segs_cap_[1], nseg_final + 1, err);
    if (rc) return rc;

    const double R = (double)soa;  // algorithmic bytes per record: key + L residues
    LAUNCH("expand_hist", 8.0 * S,
           launch_expand_hist(fkeys, S, g, digits[0], (uint32_t)tiles, per_tile, hist_, stream_));
    LAUNCH("scan_tiles", 12.0 * tiles * D1,
           launch_scan_tiles(hist_, (uint32_t)tiles, digits[0], offsets_, segs_[0], counters_, stream_));
    LAUNCH("expand_scatter", R * (S + C),
           launch_expand_scatter(L, fkeys, fres, S, stride, g, pr, digits[0], (uint32_t)tiles,
                                 per_tile, offsets_, region_[pa], stream_));
    int cur = pa, oth = pb, sidx = 0;
    uint64_t nseg = D1;
    unsigned used = digits[0];
    for (size_t k = 1; k < digits.size(); ++k) {
      used += digits[k];
      LAUNCH("partition", 2.0 * R * C,
             launch_partition(L, region_[cur], region_[oth], segs_[sidx], (uint32_t)nseg, 64 - used,
                              digits[k], segs_[sidx ^ 1], stream_));
      std::swap(cur, oth);
      sidx ^= 1;
      nseg <<= digits[k];
    }
    // Merge into the free partition buffer as the next SoA frontier.
    char* obase = (char*)region_[oth];
    const bool last = row + 1 == n;
    LAUNCH("merge", R * C,
           launch_merge(L, region_[cur], segs_[sidx], (uint32_t)nseg, pr, (uint64_t*)obase,
                        (uint32_t*)(obase + C * 8), C, last ? g.board_mask : 0, counters_, 4096,
                        stream_));
    const int merge_pending = profiling_ ? (int)pending_.size() - 1 : -1;
    HIP_TRY(hipMemcpyAsync(h_counters_, counters_, sizeof(uint64_t) * kNumCounters,
                           hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    // The merge's output size is known only now: add its written bytes.
    if (merge_pending >= 0) pending_[merge_pending].bytes += R * (double)h_counters_[kOutStates];
    if (!pending_.empty()) {
      int rc2 = resolve_pending(err);
      if (rc2) return rc2;
    }
    if (h_counters_[kError]) {
      *err = "LDS merge table overflow in row " + std::to_string(row + 1);
      return 4;
    }
    const uint64_t S_out = h_counters_[kOutStates];
    std::vector<uint64_t> w(L);
    for (int l = 0; l < L; ++l) {
      w[l] = h_counters_[kWeight0 + l] % spec.primes[l];
      final_res[l] = h_counters_[kFinal0 + l] % spec.primes[l];
    }
    u128 weight;
    std::string cerr;
    if (!crt_reconstruct(w.data(), spec.primes.data(), L, &weight, &cerr)) {
      *err = cerr;
      return 2;
    }
    if (weight > fact) {
      *err = "CRT reconstruction exceeds certified N! count bound after row " + std::to_string(row + 1);
      return 2;
    }
    lo.output_states = S_out;
    lo.output_weight = weight;
    lo.levels = (uint32_t)digits.size() - (digits[0] == 0 ? 1 : 0);
    lo.buckets = (uint32_t)nseg;
    lo.device_bytes = arena_bytes() + hist_cap_ * 4 + offsets_cap_ * 8 +
                      (segs_cap_[0] + segs_cap_[1]) * 8;
    lo.elapsed_s = std::chrono::duration<double>(Clock::now() - rt0).count();
    out->layers.push_back(lo);
    out->total_candidates += lo.candidates;
    out->total_matched += lo.matched;
    out->peak_states = std::max<uint64_t>(out->peak_states, S_out);
    out->device_peak_bytes = std::max<uint64_t>(out->device_peak_bytes, lo.device_bytes);
    S = S_out;
    stride = C;
    fi = oth;
  }
  HIP_TRY(hipEventRecord(ev_stop_, stream_));
  HIP_TRY(hipEventSynchronize(ev_stop_));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, ev_start_, ev_stop_));
  out->kernel_s = ms / 1000.0;

  // Final v1/v2 contraction: the merge of the last row summed the residues of
  // the states whose columns are all set.
  u128 count = 0;
  if (spec.first_row < n) {
    std::string cerr;
    if (S && !crt_reconstruct(final_res.data(), spec.primes.data(), L, &count, &cerr)) {
      *err = cerr;
      return 2;
    }
    if (count > fact) {
      *err = "CRT reconstruction exceeds certified N! count bound";
      return 2;
    }
  }
  out->count = count;
  for (int l = 0; l < L; ++l) out->residues[l] = S ? final_res[l] : 0;
  out->elapsed_s = std::chrono::duration<double>(Clock::now() - t0).count();
  return 0;
}

}  // namespace peps
