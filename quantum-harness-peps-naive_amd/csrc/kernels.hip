// kernels.hip -- gfx950 kernels of the explicit-C PEPS row contraction.
//
// The row transfer is the reference's compiled operator
// (contract_one_row_compiled, src/lib.rs:594-638): at every column whose three
// incoming legs match the occupied C entry's input legs, the occupied entry
// fires; its output legs are written and the diagonals shift to the next row.
// Duplicate successor boundaries are merged by exact modular addition of the
// CRT residues (the reference's checked u128 add, src/lib.rs:5354-5367).
//
// Merge strategy: a most-significant-digit radix partition on an invertible
// 64-bit mix of the key (so equal keys land in the same bucket and bucket
// sizes are uniform), finished per bucket by an LDS open-addressing table.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace peps {

namespace {

constexpr int kBlock = 256;           // threads per workgroup (4 wave64s)
constexpr int kWaves = kBlock / 64;
constexpr int kStage = 2048;          // LDS-staged records per flush
constexpr int kMaxDigitBits = 10;     // up to 1024 buckets per partition level
constexpr int kMaxDigits = 1 << kMaxDigitBits;

template <int L>
struct alignas(8) Rec {
  uint64_t key;
  uint32_t r[L];
};

// MurmurHash3's 64-bit finalizer (public domain): a bijection on u64, so
// partitioning by its bits never separates equal keys.
__device__ __forceinline__ uint64_t mix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

__device__ __forceinline__ uint32_t addmod(uint32_t a, uint32_t b, uint32_t p) {
  uint64_t s = (uint64_t)a + b;
  return (uint32_t)(s >= p ? s - p : s);
}

__device__ __forceinline__ uint64_t match_bits(uint64_t mask, uint32_t required, uint64_t board) {
  return required ? (mask & board) : (~mask & board);
}

// Legal positions of the occupied entry for a parent key (src/lib.rs:611-619,
// the sparse form src/lib.rs:656-663).
__device__ __forceinline__ uint64_t legal_positions(uint64_t key, const RowGeom& g) {
  const uint64_t bm = g.board_mask;
  const uint64_t cols = key & bm, dr = (key >> g.n) & bm, dl = (key >> (2 * g.n)) & bm;
  return match_bits(cols, g.col_in, bm) & match_bits(dr, g.dr_in, bm) & match_bits(dl, g.dl_in, bm) &
         g.orbit_mask;
}

// Successor boundary for the single set bit `b` (src/lib.rs:622-635).
__device__ __forceinline__ uint64_t successor(uint64_t key, uint64_t b, const RowGeom& g) {
  const uint64_t bm = g.board_mask;
  const uint64_t cols = key & bm, dr = (key >> g.n) & bm, dl = (key >> (2 * g.n)) & bm;
  const uint64_t ncols = g.col_out ? (cols | b) : (cols & ~b);
  const uint64_t ndr = ((g.dr_out ? (dr | b) : (dr & ~b)) << 1) & bm;
  const uint64_t ndl = (g.dl_out ? (dl | b) : (dl & ~b)) >> 1;
  return (key & ~g.state_mask) | ncols | (ndr << g.n) | (ndl << (2 * g.n));
}

// Exclusive scan of one u32 per thread over the workgroup.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total,
                                                         uint32_t* wave_tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wave_tot[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    uint32_t t = wave_tot[w];
    before += (w < wave) ? t : 0;
    all += t;
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* scratch) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  uint64_t s = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) s += scratch[w];
  __syncthreads();
  return s;
}

// Exclusive scan of counts[0..D) in LDS into starts[0..D) (D <= kMaxDigits).
__device__ __forceinline__ void scan_digits(const uint32_t* counts, uint32_t* starts, uint32_t D,
                                            uint32_t* wave_tot) {
  const uint32_t per = (D + kBlock - 1) / kBlock;
  const uint32_t lo = threadIdx.x * per;
  uint32_t local = 0;
  for (uint32_t i = 0; i < per; ++i)
    if (lo + i < D) local += counts[lo + i];
  uint32_t total;
  uint32_t run = block_exclusive_scan(local, &total, wave_tot);
  for (uint32_t i = 0; i < per; ++i) {
    if (lo + i < D) {
      starts[lo + i] = run;
      run += counts[lo + i];
    }
  }
  __syncthreads();
}

// LDS staging area shared by the expand-scatter and partition kernels: records
// are appended with a per-bucket rank, then written out grouped by bucket so
// consecutive threads store consecutive addresses of a bucket run.
template <int L>
struct Stage {
  Rec<L> rec[kStage];
  uint16_t bucket[kStage];
  uint16_t rank[kStage];
  uint16_t order[kStage];
  uint32_t count[kMaxDigits];
  uint32_t start[kMaxDigits];
  uint64_t gpos[kMaxDigits];
  uint32_t wave_tot[kWaves];
};

template <int L>
__device__ __forceinline__ void stage_put(Stage<L>& st, uint32_t slot, const Rec<L>& r,
                                          uint32_t b) {
  st.rec[slot] = r;
  st.bucket[slot] = (uint16_t)b;
  st.rank[slot] = (uint16_t)atomicAdd(&st.count[b], 1u);
}

// Write the `staged` records grouped by bucket, advance gpos, clear counts.
template <int L>
__device__ void stage_flush(Stage<L>& st, uint32_t staged, uint32_t D, Rec<L>* __restrict__ out) {
  __syncthreads();
  scan_digits(st.count, st.start, D, st.wave_tot);
  for (uint32_t i = threadIdx.x; i < staged; i += kBlock)
    st.order[st.start[st.bucket[i]] + st.rank[i]] = (uint16_t)i;
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < staged; j += kBlock) {
    const uint32_t i = st.order[j];
    const uint32_t b = st.bucket[i];
    out[st.gpos[b] + (j - st.start[b])] = st.rec[i];
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < D; b += kBlock) {
    st.gpos[b] += st.count[b];
    st.count[b] = 0;
  }
  __syncthreads();
}

// ------------------------------------------------------------------ count --

__global__ __launch_bounds__(kBlock) void k_count(const uint64_t* __restrict__ keys, uint64_t S,
                                                  RowGeom g, uint64_t* __restrict__ counters) {
  __shared__ uint64_t scratch[kWaves];
  uint64_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S;
       i += (uint64_t)gridDim.x * kBlock)
    c += __popcll(legal_positions(keys[i], g));
  c = block_sum_u64(c, scratch);
  if (threadIdx.x == 0 && c) atomicAdd((unsigned long long*)&counters[kCandidates], (unsigned long long)c);
}

// ------------------------------------------------------------ expand hist --

__global__ __launch_bounds__(kBlock) void k_expand_hist(const uint64_t* __restrict__ keys,
                                                        uint64_t S, RowGeom g, uint32_t digit_bits,
                                                        uint64_t per_tile,
                                                        uint32_t* __restrict__ hist) {
  __shared__ uint32_t lh[kMaxDigits];
  const uint32_t D = 1u << digit_bits;
  const uint32_t shift = 64 - digit_bits;
  for (uint32_t i = threadIdx.x; i < D; i += kBlock) lh[i] = 0;
  __syncthreads();
  const uint64_t lo = (uint64_t)blockIdx.x * per_tile;
  const uint64_t hi = lo + per_tile < S ? lo + per_tile : S;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
    const uint64_t key = keys[i];
    uint64_t free = legal_positions(key, g);
    while (free) {
      const uint64_t b = free & (~free + 1);
      free ^= b;
      const uint32_t d = digit_bits ? (uint32_t)(mix64(successor(key, b, g)) >> shift) : 0u;
      atomicAdd(&lh[d], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < D; i += kBlock) hist[(uint64_t)blockIdx.x * D + i] = lh[i];
}

// ------------------------------------------------------------- scan tiles --
// One workgroup of 1024 threads.  hist is [tiles][D]; offsets (same layout)
// receive each tile's first output position per digit in digit-major order;
// segs[0..D] receive the digit bucket boundaries; counters[kCandidates] the total.

constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const uint32_t* __restrict__ hist,
                                                           uint32_t tiles, uint32_t digit_bits,
                                                           uint64_t* __restrict__ offsets,
                                                           uint64_t* __restrict__ segs,
                                                           uint64_t* __restrict__ counters) {
  __shared__ uint64_t part[kScanBlock];
  __shared__ uint64_t dstart[kMaxDigits + 1];
  const uint32_t D = 1u << digit_bits;
  const uint32_t G = kScanBlock / D;  // tile groups
  const uint32_t d = threadIdx.x % D, gi = threadIdx.x / D;
  const uint32_t TG = (tiles + G - 1) / G;
  const uint32_t t0 = gi * TG, t1 = (t0 + TG < tiles) ? t0 + TG : tiles;
  uint64_t s = 0;
  if (gi < G)
    for (uint32_t t = t0; t < t1; ++t) s += hist[(uint64_t)t * D + d];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t run = 0;
    for (uint32_t dd = 0; dd < D; ++dd) {
      dstart[dd] = run;
      for (uint32_t gg = 0; gg < G; ++gg) run += part[gg * D + dd];
    }
    dstart[D] = run;
  }
  __syncthreads();
  if (gi < G) {
    uint64_t run = dstart[d];
    for (uint32_t gg = 0; gg < gi; ++gg) run += part[gg * D + d];
    for (uint32_t t = t0; t < t1; ++t) {
      offsets[(uint64_t)t * D + d] = run;
      run += hist[(uint64_t)t * D + d];
    }
  }
  for (uint32_t i = threadIdx.x; i <= D; i += kScanBlock) segs[i] = dstart[i];
  if (threadIdx.x == 0) counters[kCandidates] = dstart[D];
}

// --------------------------------------------------------- expand scatter --

template <int L>
__global__ __launch_bounds__(kBlock) void k_expand_scatter(
    const uint64_t* __restrict__ keys, const uint32_t* __restrict__ res, uint64_t S,
    uint64_t res_stride, RowGeom g, Primes pr, uint32_t digit_bits, uint64_t per_tile,
    const uint64_t* __restrict__ offsets, Rec<L>* __restrict__ out) {
  __shared__ Stage<L> st;
  const uint32_t D = 1u << digit_bits;
  const uint32_t shift = 64 - digit_bits;
  for (uint32_t b = threadIdx.x; b < D; b += kBlock) {
    st.count[b] = 0;
    st.gpos[b] = offsets[(uint64_t)blockIdx.x * D + b];
  }
  __syncthreads();
  const uint64_t lo = (uint64_t)blockIdx.x * per_tile;
  const uint64_t hi = lo + per_tile < S ? lo + per_tile : S;
  uint32_t staged = 0;
  for (uint64_t base = lo; base < hi; base += kBlock) {
    const uint64_t i = base + threadIdx.x;
    uint64_t key = 0, free = 0;
    uint32_t r[L];
    if (i < hi) {
      key = keys[i];
      free = legal_positions(key, g);
#pragma unroll
      for (int l = 0; l < L; ++l) r[l] = res[(uint64_t)l * res_stride + i];
    }
    const uint32_t cnt = (uint32_t)__popcll(free);
    uint32_t total;
    const uint32_t pos = block_exclusive_scan(cnt, &total, st.wave_tot);
    // Emit this round's successors through windows that fit the staging area.
    uint32_t win = 0;
    while (win < total) {
      if (staged == kStage) {
        stage_flush<L>(st, staged, D, out);
        staged = 0;
      }
      const uint32_t space = kStage - staged;
      const uint32_t wend = (total - win < space) ? total : win + space;
      // this thread's successors with global round index in [win, wend)
      if (cnt && pos < wend && pos + cnt > win) {
        uint64_t f = free;
        for (uint32_t j = 0; j < cnt; ++j) {
          const uint64_t b = f & (~f + 1);
          f ^= b;
          const uint32_t idx = pos + j;
          if (idx < win || idx >= wend) continue;
          Rec<L> rec;
          rec.key = successor(key, b, g);
          const bool twice = g.orbit_row && b != g.centre_bit;
#pragma unroll
          for (int l = 0; l < L; ++l) rec.r[l] = twice ? addmod(r[l], r[l], pr.p[l]) : r[l];
          const uint32_t d = digit_bits ? (uint32_t)(mix64(rec.key) >> shift) : 0u;
          stage_put<L>(st, staged + (idx - win), rec, d);
        }
      }
      staged += wend - win;
      win = wend;
      __syncthreads();
    }
  }
  if (staged) stage_flush<L>(st, staged, D, out);
}

// -------------------------------------------------------------- partition --
// One workgroup per segment (grid-stride).  Digit = bits [shift, shift+D) of
// the key mix.  subsegs[s*D + d] receives the start of digit d of segment s;
// subsegs[nseg*D] the end of the last segment.

template <int L>
__global__ __launch_bounds__(kBlock) void k_partition(const Rec<L>* __restrict__ in,
                                                      Rec<L>* __restrict__ out,
                                                      const uint64_t* __restrict__ segs,
                                                      uint32_t nseg, uint32_t shift,
                                                      uint32_t digit_bits,
                                                      uint64_t* __restrict__ subsegs) {
  __shared__ Stage<L> st;
  const uint32_t D = 1u << digit_bits;
  const uint64_t dmask = D - 1;
  for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
    const uint64_t lo = segs[s], hi = segs[s + 1];
    for (uint32_t b = threadIdx.x; b < D; b += kBlock) st.count[b] = 0;
    __syncthreads();
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kBlock)
      atomicAdd(&st.count[(uint32_t)((mix64(in[i].key) >> shift) & dmask)], 1u);
    __syncthreads();
    scan_digits(st.count, st.start, D, st.wave_tot);
    for (uint32_t b = threadIdx.x; b < D; b += kBlock) {
      st.gpos[b] = lo + st.start[b];
      subsegs[(uint64_t)s * D + b] = lo + st.start[b];
      st.count[b] = 0;
    }
    __syncthreads();
    for (uint64_t base = lo; base < hi; base += kStage) {
      const uint32_t n = (hi - base < (uint64_t)kStage) ? (uint32_t)(hi - base) : (uint32_t)kStage;
      for (uint32_t j = threadIdx.x; j < n; j += kBlock) {
        const Rec<L> rec = in[base + j];
        stage_put<L>(st, j, rec, (uint32_t)((mix64(rec.key) >> shift) & dmask));
      }
      stage_flush<L>(st, n, D, out);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) subsegs[(uint64_t)nseg * D] = segs[nseg];
}

// ------------------------------------------------------------------ merge --
// One workgroup per final bucket (grid-stride).  Records are inserted into an
// LDS open-addressing table (linear probing on the mix's low bits) and the
// residues of equal keys added modulo each prime.  A bucket larger than the
// table's safe load is processed in sub-passes that each take the keys whose
// mix falls in one residue class, so every pass sees disjoint keys.

template <int L, int T>
__global__ __launch_bounds__(kBlock) void k_merge(const Rec<L>* __restrict__ in,
                                                  const uint64_t* __restrict__ segs, uint32_t nseg,
                                                  Primes pr, uint64_t* __restrict__ out_keys,
                                                  uint32_t* __restrict__ out_res,
                                                  uint64_t out_stride, uint64_t final_mask,
                                                  uint64_t* __restrict__ counters) {
  __shared__ uint64_t tkey[T];
  __shared__ uint32_t tres[L][T];
  __shared__ uint32_t rcount[T / kBlock][kWaves];
  __shared__ uint64_t base_sh;
  __shared__ uint64_t scratch[kWaves];
  constexpr uint32_t kCap = (T * 5) / 8;
  constexpr uint32_t kRounds = T / kBlock;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t wsum[L], fsum[L];
#pragma unroll
  for (int l = 0; l < L; ++l) wsum[l] = fsum[l] = 0;
  uint32_t overflow = 0;

  for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
    const uint64_t lo = segs[s], hi = segs[s + 1];
    if (lo == hi) continue;
    const uint64_t m = hi - lo;
    const uint32_t passes = m <= kCap ? 1u : (uint32_t)((m + kCap - 1) / kCap);
    for (uint32_t pass = 0; pass < passes; ++pass) {
      for (uint32_t i = threadIdx.x; i < T; i += kBlock) {
        tkey[i] = kEmptyKey;
#pragma unroll
        for (int l = 0; l < L; ++l) tres[l][i] = 0;
      }
      __syncthreads();
      for (uint64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const Rec<L> rec = in[i];
        const uint64_t h = mix64(rec.key);
        if (passes > 1 && (uint32_t)((((h >> 24) & 0xffffffull) * passes) >> 24) != pass) continue;
        uint32_t slot = (uint32_t)h & (T - 1);
        bool placed = false;
        for (uint32_t probe = 0; probe < T; ++probe) {
          const unsigned long long old =
              atomicCAS((unsigned long long*)&tkey[slot], (unsigned long long)kEmptyKey,
                        (unsigned long long)rec.key);
          if (old == kEmptyKey || old == rec.key) {
            placed = true;
            break;
          }
          slot = (slot + 1) & (T - 1);
        }
        if (!placed) {
          overflow = 1;
          continue;
        }
#pragma unroll
        for (int l = 0; l < L; ++l) {
          uint32_t* a = &tres[l][slot];
          uint32_t old = *a, assumed;
          do {
            assumed = old;
            old = atomicCAS(a, assumed, addmod(assumed, rec.r[l], pr.p[l]));
          } while (old != assumed);
        }
      }
      __syncthreads();
      // Compact occupied slots: ballot ranks within each (round, wave).
      for (uint32_t rd = 0; rd < kRounds; ++rd) {
        const bool occ = tkey[rd * kBlock + threadIdx.x] != kEmptyKey;
        const uint64_t bal = __ballot(occ);
        if (lane == 0) rcount[rd][wave] = (uint32_t)__popcll(bal);
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t rd = 0; rd < kRounds; ++rd)
          for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = rcount[rd][w];
            rcount[rd][w] = run;
            run += c;
          }
        base_sh = run ? atomicAdd((unsigned long long*)&counters[kOutStates], (unsigned long long)run) : 0;
      }
      __syncthreads();
      const uint64_t base = base_sh;
      for (uint32_t rd = 0; rd < kRounds; ++rd) {
        const uint32_t slot = rd * kBlock + threadIdx.x;
        const uint64_t key = tkey[slot];
        const bool occ = key != kEmptyKey;
        const uint64_t bal = __ballot(occ);
        if (occ) {
          const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
          const uint64_t pos = base + rcount[rd][wave] + rank;
          out_keys[pos] = key;
          const bool fin = final_mask && (key & final_mask) == final_mask;
#pragma unroll
          for (int l = 0; l < L; ++l) {
            const uint32_t v = tres[l][slot];
            out_res[(uint64_t)l * out_stride + pos] = v;
            wsum[l] += v;
            if (fin) fsum[l] += v;
          }
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const uint64_t w = block_sum_u64(wsum[l] % pr.p[l], scratch);
    const uint64_t f = block_sum_u64(fsum[l] % pr.p[l], scratch);
    if (threadIdx.x == 0) {
      if (w) atomicAdd((unsigned long long*)&counters[kWeight0 + l], (unsigned long long)(w % pr.p[l]));
      if (f) atomicAdd((unsigned long long*)&counters[kFinal0 + l], (unsigned long long)(f % pr.p[l]));
    }
  }
  if (overflow) atomicOr((unsigned long long*)&counters[kError], 1ull);
}

uint32_t grid_for(uint64_t items, uint32_t per_block, uint32_t cap) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return g > cap ? cap : (uint32_t)g;
}

}  // namespace

size_t record_bytes(int lanes) {
  switch (lanes) {
    case 1: return sizeof(Rec<1>);
    case 2: return sizeof(Rec<2>);
    case 3: return sizeof(Rec<3>);
    default: return sizeof(Rec<4>);
  }
}

hipError_t launch_count(const uint64_t* keys, uint64_t S, const RowGeom& g, uint64_t* counters,
                        hipStream_t stream) {
  if (S == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count, dim3(grid_for(S, kBlock * 8, 4096)), dim3(kBlock), 0, stream, keys,
                     S, g, counters);
  return hipGetLastError();
}

hipError_t launch_expand_hist(const uint64_t* keys, uint64_t S, const RowGeom& g,
                              uint32_t digit_bits, uint32_t tiles, uint64_t per_tile,
                              uint32_t* hist, hipStream_t stream) {
  if (digit_bits > kMaxDigitBits || tiles == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_expand_hist, dim3(tiles), dim3(kBlock), 0, stream, keys, S, g, digit_bits,
                     per_tile, hist);
  return hipGetLastError();
}

hipError_t launch_scan_tiles(const uint32_t* hist, uint32_t tiles, uint32_t digit_bits,
                             uint64_t* offsets, uint64_t* segs, uint64_t* counters,
                             hipStream_t stream) {
  if (digit_bits > kMaxDigitBits || tiles == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanBlock), 0, stream, hist, tiles, digit_bits,
                     offsets, segs, counters);
  return hipGetLastError();
}

hipError_t launch_expand_scatter(int lanes, const uint64_t* keys, const uint32_t* res, uint64_t S,
                                 uint64_t res_stride, const RowGeom& g, const Primes& pr,
                                 uint32_t digit_bits, uint32_t tiles, uint64_t per_tile,
                                 const uint64_t* offsets, void* out, hipStream_t stream) {
  if (digit_bits > kMaxDigitBits || tiles == 0) return hipErrorInvalidValue;
  dim3 grid(tiles), block(kBlock);
  switch (lanes) {
    case 1: hipLaunchKernelGGL(k_expand_scatter<1>, grid, block, 0, stream, keys, res, S, res_stride, g, pr, digit_bits, per_tile, offsets, (Rec<1>*)out); break;
    case 2: hipLaunchKernelGGL(k_expand_scatter<2>, grid, block, 0, stream, keys, res, S, res_stride, g, pr, digit_bits, per_tile, offsets, (Rec<2>*)out); break;
    case 3: hipLaunchKernelGGL(k_expand_scatter<3>, grid, block, 0, stream, keys, res, S, res_stride, g, pr, digit_bits, per_tile, offsets, (Rec<3>*)out); break;
    case 4: hipLaunchKernelGGL(k_expand_scatter<4>, grid, block, 0, stream, keys, res, S, res_stride, g, pr, digit_bits, per_tile, offsets, (Rec<4>*)out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_partition(int lanes, const void* in, void* out, const uint64_t* segs,
                            uint32_t nseg, uint32_t shift, uint32_t digit_bits, uint64_t* subsegs,
                            hipStream_t stream) {
  if (digit_bits == 0 || digit_bits > kMaxDigitBits || nseg == 0) return hipErrorInvalidValue;
  dim3 grid(nseg < 4096 ? nseg : 4096), block(kBlock);
  switch (lanes) {
    case 1: hipLaunchKernelGGL(k_partition<1>, grid, block, 0, stream, (const Rec<1>*)in, (Rec<1>*)out, segs, nseg, shift, digit_bits, subsegs); break;
    case 2: hipLaunchKernelGGL(k_partition<2>, grid, block, 0, stream, (const Rec<2>*)in, (Rec<2>*)out, segs, nseg, shift, digit_bits, subsegs); break;
    case 3: hipLaunchKernelGGL(k_partition<3>, grid, block, 0, stream, (const Rec<3>*)in, (Rec<3>*)out, segs, nseg, shift, digit_bits, subsegs); break;
    case 4: hipLaunchKernelGGL(k_partition<4>, grid, block, 0, stream, (const Rec<4>*)in, (Rec<4>*)out, segs, nseg, shift, digit_bits, subsegs); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_merge(int lanes, const void* in, const uint64_t* segs, uint32_t nseg,
                        const Primes& pr, uint64_t* out_keys, uint32_t* out_res,
                        uint64_t out_stride, uint64_t final_mask, uint64_t* counters,
                        uint32_t table_slots, hipStream_t stream) {
  if (nseg == 0) return hipSuccess;
  if (table_slots != 4096) return hipErrorInvalidValue;
  dim3 grid(nseg < 8192 ? nseg : 8192), block(kBlock);
  switch (lanes) {
    case 1: hipLaunchKernelGGL((k_merge<1, 4096>), grid, block, 0, stream, (const Rec<1>*)in, segs, nseg, pr, out_keys, out_res, out_stride, final_mask, counters); break;
    case 2: hipLaunchKernelGGL((k_merge<2, 4096>), grid, block, 0, stream, (const Rec<2>*)in, segs, nseg, pr, out_keys, out_res, out_stride, final_mask, counters); break;
    case 3: hipLaunchKernelGGL((k_merge<3, 4096>), grid, block, 0, stream, (const Rec<3>*)in, segs, nseg, pr, out_keys, out_res, out_stride, final_mask, counters); break;
    case 4: hipLaunchKernelGGL((k_merge<4, 4096>), grid, block, 0, stream, (const Rec<4>*)in, segs, nseg, pr, out_keys, out_res, out_stride, final_mask, counters); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace peps
