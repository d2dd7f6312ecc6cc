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
// 64-bit mix of the key (so equal keys share a bucket and bucket sizes are
// uniform), finished per bucket by an LDS open-addressing table.  The merge of
// row r is fused with the expansion into row r+1, so the frontier itself is
// never written to HBM: candidates go straight from one row's buckets to the
// next row's buckets.
//
// Memory-level parallelism: each thread issues a batch of independent loads
// before using any of them, and workgroup barriers wait only for LDS traffic
// (lds_barrier), so global loads of the next batch stay in flight across
// them.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.h"

namespace peps {

namespace {

constexpr int kBlock = 256;           // threads per workgroup (4 wave64s)
constexpr int kWaves = kBlock / 64;
constexpr int kStage = 2048;          // LDS-staged records per flush
constexpr int kPerThread = kStage / kBlock;
constexpr int kScatterBlock = 512;    // SCATTER merge workgroup: 8 wave64s share one LDS table
constexpr int kXcds = 8;              // MI355X: accelerator complex dies, one L2 each

template <int L>
struct alignas(8) Rec {
  uint64_t key;
  uint32_t r[L];
};

// Packed record: 96 bits = key | weight << key_bits, as three u32 words
// (12 bytes, 4-byte aligned).  key_bits is in [3, 63] and the weight below
// 2^(96 - key_bits), so every shift below is in range.
template <>
struct alignas(4) Rec<kPackedRecords> {
  uint32_t w[3];
};
using Packed = Rec<kPackedRecords>;

__device__ __forceinline__ uint64_t packed_low(const Packed& r) {
  return (uint64_t)r.w[0] | ((uint64_t)r.w[1] << 32);
}
__device__ __forceinline__ uint64_t packed_key(const Packed& r, uint32_t kb) {
  return packed_low(r) & ((1ull << kb) - 1);
}
__device__ __forceinline__ uint64_t packed_weight(const Packed& r, uint32_t kb) {
  return (packed_low(r) >> kb) | ((uint64_t)r.w[2] << (64 - kb));
}
__device__ __forceinline__ Packed make_packed(uint64_t key, uint64_t weight, uint32_t kb) {
  const uint64_t lo = key | (weight << kb);
  Packed r;
  r.w[0] = (uint32_t)lo;
  r.w[1] = (uint32_t)(lo >> 32);
  r.w[2] = (uint32_t)(weight >> (64 - kb));
  return r;
}

// Key of either record kind (kb: packed key width).
template <int L>
__device__ __forceinline__ uint64_t rec_key(const Rec<L>& r, uint32_t kb) {
  if constexpr (L == kPackedRecords)
    return packed_key(r, kb);
  else
    return r.key;
}

// Streaming (non-temporal) loads for records that a pass reads once: they
// leave L2 capacity to the partly written lines of the scattered stores.
template <int L>
__device__ __forceinline__ Rec<L> load_stream(const Rec<L>* p) {
  Rec<L> r;
  if constexpr (L == kPackedRecords) {
#pragma unroll
    for (int j = 0; j < 3; ++j) r.w[j] = __builtin_nontemporal_load(&p->w[j]);
  } else {
    r.key = __builtin_nontemporal_load(&p->key);
#pragma unroll
    for (int l = 0; l < L; ++l) r.r[l] = __builtin_nontemporal_load(&p->r[l]);
  }
  return r;
}

template <int L>
__device__ __forceinline__ uint64_t load_key_stream(const Rec<L>* p, uint32_t kb) {
  if constexpr (L == kPackedRecords) {
    Packed r;
    r.w[0] = __builtin_nontemporal_load(&p->w[0]);
    r.w[1] = __builtin_nontemporal_load(&p->w[1]);
    r.w[2] = 0;
    return packed_key(r, kb);
  } else {
    return __builtin_nontemporal_load(&p->key);
  }
}

// Residue of an exact 128-bit sum (hi:lo) modulo p < 2^32.
__device__ __forceinline__ uint64_t mod_u128(uint64_t hi, uint64_t lo, uint32_t p) {
  const uint64_t t64 = (~0ull % p + 1) % p;  // 2^64 mod p
  return ((hi % p) * t64 % p + lo % p) % p;
}

// Workgroup barrier that waits for LDS operations only.  Outstanding global
// loads keep flying; their consumers are guarded by the compiler's own
// vmcnt waits.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Record range of bucket b of a level: contiguous buckets (cap == 0) are
// [segs[b], segs[b + 1]); fixed-capacity buckets are [b * cap, b * cap +
// min(segs[b], cap)), segs[b] being the bucket's fill cursor (a cursor past
// cap means the level overflowed, which its writer flagged).
__device__ __forceinline__ void bucket_range(const uint64_t* __restrict__ segs, uint64_t cap, uint64_t b,
                                             uint64_t* lo, uint64_t* hi) {
  if (cap) {
    const uint64_t n = segs[b];
    *lo = b * cap;
    *hi = *lo + (n < cap ? n : cap);
  } else {
    *lo = segs[b];
    *hi = segs[b + 1];
  }
}

// Bijective 64-bit mix (xorshift, odd multiply, xorshift): equal keys share
// every derived digit and slot; distinct keys spread uniformly over the top
// bits (radix digits), bits 12-35 (merge-pass class) and the low bits (LDS
// table slot).
__device__ __forceinline__ uint64_t mix64(uint64_t k) {
  k ^= k >> 31;
  k *= 0x9e3779b97f4a7c15ull;
  k ^= k >> 29;
  return k;
}

__device__ __forceinline__ uint32_t addmod(uint32_t a, uint32_t b, uint32_t p) {
  uint64_t s = (uint64_t)a + b;
  return (uint32_t)(s >= p ? s - p : s);
}

// A boundary key unpacked into its three N-bit signal masks.
struct Board {
  uint32_t cols, dr, dl;
};

__device__ __forceinline__ Board unpack(uint64_t key, const RowGeom& g) {
  return Board{(uint32_t)key & g.board_mask, (uint32_t)(key >> g.n) & g.board_mask,
               (uint32_t)(key >> (2 * g.n)) & g.board_mask};
}

// Columns whose three incoming legs match the occupied entry's input legs
// (src/lib.rs:611-619; the sparse form src/lib.rs:656-663).
__device__ __forceinline__ uint32_t legal_positions(const Board& s, const RowGeom& g) {
  return ~((s.cols ^ g.col_in) | (s.dr ^ g.dr_in) | (s.dl ^ g.dl_in)) & g.orbit_mask;
}

__device__ __forceinline__ uint32_t legal_positions(uint64_t key, const RowGeom& g) {
  return legal_positions(unpack(key, g), g);
}

// Successor boundary for the single column bit `b` (src/lib.rs:622-635): the
// output legs are written at b, diag_dr shifts right-down (<< 1, the bit
// leaving the board is contracted with v2), diag_dl shifts left-down (>> 1).
__device__ __forceinline__ uint64_t successor(uint64_t key, const Board& s, uint32_t b,
                                              const RowGeom& g) {
  const uint32_t c = (s.cols & ~b) | (b & g.col_out);
  const uint32_t dr = (((s.dr & ~b) | (b & g.dr_out)) << 1) & g.board_mask;
  const uint32_t dl = ((s.dl & ~b) | (b & g.dl_out)) >> 1;
  return (key & ~g.state_mask) | (uint64_t)c | ((uint64_t)dr << g.n) | ((uint64_t)dl << (2 * g.n));
}

// The same successor split into a per-state part and a per-column part:
// at a legal column b every incoming leg equals the occupied entry's input
// leg, so writing the output legs there flips exactly the bits of
// b & (in ^ out) in each signal, and the shifts distribute over the flips.
// successor(key, s, b, g) == successor_base(key, s, g) ^ successor_delta(b, g)
// for every b in legal_positions(s, g).
__device__ __forceinline__ uint64_t successor_base(uint64_t key, const Board& s, const RowGeom& g) {
  return (key & ~g.state_mask) | (uint64_t)s.cols | ((uint64_t)((s.dr << 1) & g.board_mask) << g.n) |
         ((uint64_t)(s.dl >> 1) << (2 * g.n));
}

__device__ __forceinline__ uint64_t successor_delta(uint32_t b, const RowGeom& g) {
  const uint32_t c = b & (g.col_in ^ g.col_out);
  const uint32_t dr = ((b & (g.dr_in ^ g.dr_out)) << 1) & g.board_mask;
  const uint32_t dl = (b & (g.dl_in ^ g.dl_out)) >> 1;
  return (uint64_t)c | ((uint64_t)dr << g.n) | ((uint64_t)dl << (2 * g.n));
}

// The top `bits` bits of mix64(key), for bits <= 29 (callers take at most
// kFineBits).  The final xorshift by 29 leaves the product's top 29 bits as
// they are, so only the product's high word is formed: with k = hi:lo and
// the multiplier ch:cl, it is umulhi(lo, cl) + lo * ch + hi * cl mod 2^32.
__device__ __forceinline__ uint32_t top_digit(uint64_t key, uint32_t bits) {
  if (bits == 0) return 0u;
  const uint64_t k = key ^ (key >> 31);
  const uint32_t lo = (uint32_t)k, hi = (uint32_t)(k >> 32);
  constexpr uint32_t ch = 0x9e3779b9u, cl = 0x7f4a7c15u;
  const uint32_t ph = __umulhi(lo, cl) + lo * ch + hi * cl;
  return ph >> (32 - bits);
}

// Exclusive scan of one u32 per thread over a workgroup of NW waves.
template <int NW = kWaves>
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
  lds_barrier();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    uint32_t t = wave_tot[w];
    before += (w < wave) ? t : 0;
    all += t;
  }
  lds_barrier();
  *total = all;
  return before + x - v;
}

__device__ __forceinline__ uint64_t block_exclusive_scan64(uint64_t v, uint64_t* total,
                                                           uint64_t* wave_tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wave_tot[wave] = x;
  lds_barrier();
  uint64_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    uint64_t t = wave_tot[w];
    before += (w < wave) ? t : 0;
    all += t;
  }
  lds_barrier();
  *total = all;
  return before + x - v;
}

template <int NW = kWaves>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* scratch) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) scratch[wave] = v;
  lds_barrier();
  uint64_t s = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) s += scratch[w];
  lds_barrier();
  return s;
}

// Exclusive scan of counts[0..D) in LDS into starts[0..D) by a workgroup of
// BT threads.
// (Counts may be u64 words whose low half is the count.)
template <int BT = kBlock, typename CT = uint32_t>
__device__ __forceinline__ void scan_digits(const CT* counts, uint32_t* starts, uint32_t D,
                                            uint32_t* wave_tot) {
  const uint32_t per = (D + BT - 1) / BT;
  const uint32_t lo = threadIdx.x * per;
  uint32_t local = 0;
  for (uint32_t i = 0; i < per; ++i)
    if (lo + i < D) local += (uint32_t)counts[lo + i];
  uint32_t total;
  uint32_t run = block_exclusive_scan<BT / 64>(local, &total, wave_tot);
  for (uint32_t i = 0; i < per; ++i) {
    if (lo + i < D) {
      starts[lo + i] = run;
      run += (uint32_t)counts[lo + i];
    }
  }
  lds_barrier();
}

// LDS staging area of the seed-row expand-scatter: records are appended
// with a per-bucket rank, then written out grouped by bucket so consecutive
// threads store consecutive addresses of a bucket run.  (A thread emits a
// variable number of successors there, so ranks cannot stay in registers as
// in the partition's PartStage.)
template <int L, int DMAX>
struct Stage {
  Rec<L> rec[kStage];
  std::conditional_t<(DMAX <= 256), uint8_t, uint16_t> bucket[kStage];
  uint16_t rank[kStage];
  uint16_t order[kStage];
  uint32_t count[DMAX];
  uint32_t start[DMAX];
  uint64_t gpos[DMAX];
  uint32_t wave_tot[kWaves];
};

template <int L, int DMAX>
__device__ __forceinline__ void stage_put(Stage<L, DMAX>& st, uint32_t slot, const Rec<L>& r,
                                          uint32_t b) {
  st.rec[slot] = r;
  st.bucket[slot] = (std::remove_reference_t<decltype(st.bucket[0])>)b;
  st.rank[slot] = (uint16_t)atomicAdd(&st.count[b], 1u);
}

// Write the `staged` records grouped by bucket, advance gpos, clear counts.
// A position at or beyond `cap` (a stale size plan) is skipped and flagged.
template <int L, int DMAX>
__device__ void stage_flush(Stage<L, DMAX>& st, uint32_t staged, uint32_t D, Rec<L>* __restrict__ out,
                            uint64_t cap, uint64_t* __restrict__ counters) {
  lds_barrier();
  scan_digits(st.count, st.start, D, st.wave_tot);
  for (uint32_t i = threadIdx.x; i < staged; i += kBlock)
    st.order[st.start[st.bucket[i]] + st.rank[i]] = (uint16_t)i;
  lds_barrier();
  for (uint32_t j = threadIdx.x; j < staged; j += kBlock) {
    const uint32_t i = st.order[j];
    const uint32_t b = st.bucket[i];
    const uint64_t pos = st.gpos[b] + (j - st.start[b]);
    if (pos < cap)
      out[pos] = st.rec[i];
    else
      atomicOr((unsigned long long*)&counters[kError], 2ull);
  }
  lds_barrier();
  for (uint32_t b = threadIdx.x; b < D; b += kBlock) {
    st.gpos[b] += st.count[b];
    st.count[b] = 0;
  }
  lds_barrier();
}

// ------------------------------------------------------------------ count --

__global__ __launch_bounds__(kBlock) void k_count(const uint64_t* __restrict__ keys, uint64_t S,
                                                  RowGeom g, uint64_t* __restrict__ counters) {
  __shared__ uint64_t scratch[kWaves];
  uint64_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < S;
       i += (uint64_t)gridDim.x * kBlock)
    c += __popc(legal_positions(keys[i], g));
  c = block_sum_u64(c, scratch);
  if (threadIdx.x == 0 && c) atomicAdd((unsigned long long*)&counters[kCandidates], (unsigned long long)c);
}

// Zero fill (used instead of hipMemsetAsync so the row pipeline can be
// captured into a HIP graph).
__global__ __launch_bounds__(kBlock) void k_fill_u64(uint64_t* __restrict__ p, uint64_t n, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
    p[i] = v;
}

// ------------------------------------------------------------ expand hist --

constexpr uint32_t kExpandDigits = 1u << kFineBits;

__global__ __launch_bounds__(kBlock) void k_expand_hist(const uint64_t* __restrict__ keys,
                                                        uint64_t S, RowGeom g, uint32_t digit_bits,
                                                        uint64_t per_tile,
                                                        uint32_t* __restrict__ hist) {
  __shared__ uint32_t lh[kExpandDigits];
  const uint32_t D = 1u << digit_bits;
  for (uint32_t i = threadIdx.x; i < D; i += kBlock) lh[i] = 0;
  lds_barrier();
  const uint64_t lo = (uint64_t)blockIdx.x * per_tile;
  const uint64_t hi = lo + per_tile < S ? lo + per_tile : S;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
    const uint64_t key = keys[i];
    const Board bs = unpack(key, g);
    uint32_t free = legal_positions(bs, g);
    while (free) {
      const uint32_t b = free & (~free + 1);
      free ^= b;
      atomicAdd(&lh[top_digit(successor(key, bs, b, g), digit_bits)], 1u);
    }
  }
  lds_barrier();
  for (uint32_t i = threadIdx.x; i < D; i += kBlock) hist[(uint64_t)i * gridDim.x + blockIdx.x] = lh[i];
}

// ------------------------------------------------------ upper-bound count --
// Successor digits of every record of row r's buckets, duplicates included,
// per merge workgroup (the same bucket assignment as the merges): an upper
// bound of what the SCATTER merge writes per (digit, workgroup), without
// the LDS table.  hist[d * W + w].
template <int L>
__global__ __launch_bounds__(kBlock) void k_count_ub(const Rec<L>* __restrict__ in,
                                                     const uint64_t* __restrict__ segs, uint64_t in_cap,
                                                     uint32_t nseg, uint32_t per, RowGeom g, uint32_t digit_bits,
                                                     uint32_t* __restrict__ hist) {
  __shared__ uint32_t lh[kExpandDigits];
  const uint32_t D = 1u << digit_bits;
  const uint32_t W = gridDim.x;
  for (uint32_t i = threadIdx.x; i < D; i += kBlock) lh[i] = 0;
  lds_barrier();
  const uint32_t b0 = blockIdx.x * per;
  const uint32_t b1 = (b0 + per < nseg) ? b0 + per : nseg;
  for (uint32_t b = b0; b < b1; ++b) {
    uint64_t lo, hi;
    bucket_range(segs, in_cap, b, &lo, &hi);
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
      const uint64_t key = load_key_stream(in + i, g.key_bits);
      const Board bs = unpack(key, g);
      uint32_t free = legal_positions(bs, g);
      while (free) {
        const uint32_t bit = free & (~free + 1);
        free ^= bit;
        atomicAdd(&lh[top_digit(successor(key, bs, bit, g), digit_bits)], 1u);
      }
    }
  }
  lds_barrier();
  for (uint32_t i = threadIdx.x; i < D; i += kBlock) hist[(uint64_t)i * W + blockIdx.x] = lh[i];
}

// ------------------------------------------------------------------ caps --
// Capped SCATTER runs: merge workgroup w's successors of digit d go to
// run_start[d * W + w] = d * T + P_w, at most cap_w of them, where
// cap_w = ceil(U_w / D) + 8 sqrt(U_w / D) + 32 from the fanout sum U_w of
// the records w will merge (an upper bound of its successors: duplicates
// count once per record) and P_w = cap_0 + .. + cap_{w-1}, T = P_W.  One
// workgroup scans the W caps; counters[kCapExtent] = D * T.
constexpr int kCapsBlock = 1024;
// shrink > 0 divides every cap by 2^shrink (a test that forces the overflow
// path).
__global__ __launch_bounds__(kCapsBlock) void k_caps_scan(const uint64_t* __restrict__ U, uint32_t W,
                                                          uint32_t digit_bits, uint32_t shrink,
                                                          uint64_t* __restrict__ capw, uint64_t* __restrict__ pre,
                                                          uint64_t* __restrict__ counters) {
  __shared__ uint64_t wave_tot[kCapsBlock / 64];
  const uint32_t per = (W + kCapsBlock - 1) / kCapsBlock;
  const uint32_t lo = threadIdx.x * per;
  const double D = (double)(1u << digit_bits);
  uint64_t local = 0;
  for (uint32_t i = 0; i < per; ++i) {
    const uint32_t w = lo + i;
    if (w < W) {
      const double m = (double)U[w] / D;
      const uint64_t c = ((uint64_t)__builtin_ceil(m) + (uint64_t)(8.0 * __builtin_sqrt(m)) + 32) >> shrink;
      capw[w] = c;
      local += c;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = local;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wave_tot[wave] = x;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int w = 0; w < kCapsBlock / 64; ++w) {
    before += (w < wave) ? wave_tot[w] : 0;
    all += wave_tot[w];
  }
  uint64_t run = before + x - local;
  for (uint32_t i = 0; i < per; ++i) {
    const uint32_t w = lo + i;
    if (w < W) {
      pre[w] = run;
      run += capw[w];
    }
  }
  if (threadIdx.x == 0) {
    pre[W] = all;
    counters[kCapExtent] = all << digit_bits;
  }
}

__global__ __launch_bounds__(kBlock) void k_caps_expand(const uint64_t* __restrict__ pre, uint32_t W,
                                                        uint32_t digit_bits, uint64_t* __restrict__ run_start) {
  const uint64_t T = pre[W];
  const uint64_t total = (uint64_t)W << digit_bits;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kBlock)
    run_start[i] = (i / W) * T + pre[i % W];
}

// --------------------------------------------------------- expand scatter --

template <int L>
__global__ __launch_bounds__(kBlock) void k_expand_scatter(
    const uint64_t* __restrict__ keys, const uint32_t* __restrict__ res, uint64_t S,
    uint64_t res_stride, RowGeom g, Primes pr, uint32_t digit_bits, uint64_t per_tile,
    const uint64_t* __restrict__ offsets, Rec<L>* __restrict__ out, uint64_t cap,
    uint64_t* __restrict__ counters) {
  __shared__ Stage<L, kExpandDigits> st;
  const uint32_t D = 1u << digit_bits;
  for (uint32_t b = threadIdx.x; b < D; b += kBlock) {
    st.count[b] = 0;
    st.gpos[b] = offsets[(uint64_t)b * gridDim.x + blockIdx.x];
  }
  lds_barrier();
  const uint64_t lo = (uint64_t)blockIdx.x * per_tile;
  const uint64_t hi = lo + per_tile < S ? lo + per_tile : S;
  uint32_t staged = 0;
  for (uint64_t base = lo; base < hi; base += kBlock) {
    const uint64_t i = base + threadIdx.x;
    uint64_t key = 0;
    uint32_t free = 0;
    Board bs{0, 0, 0};
    uint32_t r[L > 0 ? L : 1];
    uint64_t wt = 0;  // packed records: the exact seed weight
    if (i < hi) {
      key = keys[i];
      bs = unpack(key, g);
      free = legal_positions(bs, g);
      if constexpr (L == kPackedRecords) {
        wt = ((const uint64_t*)res)[i];
      } else {
#pragma unroll
        for (int l = 0; l < L; ++l) r[l] = res[(uint64_t)l * res_stride + i];
      }
    }
    const uint32_t cnt = (uint32_t)__popc(free);
    uint32_t total;
    const uint32_t pos = block_exclusive_scan(cnt, &total, st.wave_tot);
    // Emit this round's successors through windows that fit the staging area.
    uint32_t win = 0;
    while (win < total) {
      if (staged == kStage) {
        stage_flush<L, kExpandDigits>(st, staged, D, out, cap, counters);
        staged = 0;
      }
      const uint32_t space = kStage - staged;
      const uint32_t wend = (total - win < space) ? total : win + space;
      if (cnt && pos < wend && pos + cnt > win) {
        uint32_t f = free;
        for (uint32_t j = 0; j < cnt; ++j) {
          const uint32_t b = f & (~f + 1);
          f ^= b;
          const uint32_t idx = pos + j;
          if (idx < win || idx >= wend) continue;
          Rec<L> rec;
          const uint64_t sk = successor(key, bs, b, g);
          const bool twice = g.orbit_row && b != g.centre_bit;
          if constexpr (L == kPackedRecords) {
            const uint64_t w = twice ? 2 * wt : wt;  // seeds are below 2^52: no wrap
            if (w >> g.weight_bits) atomicOr((unsigned long long*)&counters[kError], 4ull);
            rec = make_packed(sk, w, g.key_bits);
          } else {
            rec.key = sk;
#pragma unroll
            for (int l = 0; l < L; ++l) rec.r[l] = twice ? addmod(r[l], r[l], pr.p[l]) : r[l];
          }
          stage_put<L, kExpandDigits>(st, staged + (idx - win), rec, top_digit(sk, digit_bits));
        }
      }
      staged += wend - win;
      win = wend;
      lds_barrier();
    }
  }
  if (staged) stage_flush<L, kExpandDigits>(st, staged, D, out, cap, counters);
}

// ---------------------------------------------------- literal sitewise C --
//
// contract_one_row_sitewise (src/lib.rs:513-592) applied to one parent per
// thread: the partial rows of the parent live in LDS (two generations,
// lane-interleaved so a wave's accesses hit distinct banks); at every site
// each partial row looks up the C entries matching (column, row signal,
// diag_dr, diag_dl) and forks once per entry, writing the entry's column
// output at the site, its diag_dr output one site to the right and its
// diag_dl output one site to the left (signals leaving the board are dropped:
// v2).  The left boundary is v0 (signal 0); after the last site only partial
// rows with signal 1 survive (v1).  Weights are multiplied by the entry value
// modulo every CRT prime.

constexpr int kSiteBlock = 64;

__device__ __forceinline__ uint32_t mulmod(uint32_t a, uint32_t b, uint32_t p) {
  return (uint32_t)(((uint64_t)a * b) % p);
}

template <int L, bool WEIGHTS, typename Emit>
__device__ __forceinline__ bool site_walk(uint64_t parent, const uint32_t* w_in, const RowGeom& g,
                                          const SiteTable& t, const Primes& pr, uint64_t* pk,
                                          uint32_t* pw, uint64_t& visits, uint64_t& matched,
                                          Emit&& emit) {
  constexpr uint64_t kSig = 1ull << 63;
  const int tid = threadIdx.x;
  auto K = [&](int gen, int j) -> uint64_t& { return pk[(gen * kSitePartials + j) * kSiteBlock + tid]; };
  auto Wt = [&](int gen, int j, int l) -> uint32_t& {
    return pw[((gen * kSitePartials + j) * L + l) * kSiteBlock + tid];
  };
  const unsigned n = g.n;
  int np = 1, cur = 0;
  bool ok = true;
  K(0, 0) = 0;
  if (WEIGHTS) {
#pragma unroll
    for (int l = 0; l < L; ++l) Wt(0, 0, l) = w_in[l];
  }
  for (unsigned c = 0; c < n; ++c) {
    const uint32_t col_in = (uint32_t)(parent >> c) & 1u;
    const uint32_t dr_in = (uint32_t)(parent >> (n + c)) & 1u;
    const uint32_t dl_in = (uint32_t)(parent >> (2 * n + c)) & 1u;
    const int nxt = cur ^ 1;
    int nn = 0;
    for (int j = 0; j < np; ++j) {
      const uint64_t k = K(cur, j);
      const uint32_t row_in = (uint32_t)(k >> 63);
      const uint32_t idx = col_in | (row_in << 1) | (dr_in << 2) | (dl_in << 3);
      const uint32_t e0 = t.first[idx], ec = t.count[idx];
      visits += 1;
      for (uint32_t e = e0; e < e0 + ec; ++e) {
        matched += 1;
        const uint32_t o = t.out[e];
        uint64_t sk = k & ~kSig;
        sk = (sk & ~(1ull << c)) | ((uint64_t)(o & 1u) << c);
        if (c + 1 < n) {
          const unsigned b = n + c + 1;
          sk = (sk & ~(1ull << b)) | ((uint64_t)((o >> 2) & 1u) << b);
        }
        if (c > 0) {
          const unsigned b = 2 * n + c - 1;
          sk = (sk & ~(1ull << b)) | ((uint64_t)((o >> 3) & 1u) << b);
        }
        sk |= (uint64_t)((o >> 1) & 1u) << 63;
        if (nn < kSitePartials) {
          K(nxt, nn) = sk;
          if (WEIGHTS) {
#pragma unroll
            for (int l = 0; l < L; ++l) Wt(nxt, nn, l) = mulmod(Wt(cur, j, l), t.value[e][l], pr.p[l]);
          }
          ++nn;
        } else {
          ok = false;
        }
      }
    }
    np = nn;
    cur = nxt;
  }
  const uint64_t slice = parent & ~g.state_mask;
  for (int j = 0; j < np; ++j) {
    const uint64_t k = K(cur, j);
    if (k >> 63) {
      uint32_t w[L];
#pragma unroll
      for (int l = 0; l < L; ++l) w[l] = WEIGHTS ? Wt(cur, j, l) : 0u;
      emit(slice | (k & ~kSig), w);
    }
  }
  return ok;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__global__ __launch_bounds__(kSiteBlock) void k_site_count(const uint64_t* __restrict__ keys, uint64_t S,
                                                           RowGeom g, SiteTable t,
                                                           uint64_t* __restrict__ counters) {
  __shared__ uint64_t pk[2 * kSitePartials * kSiteBlock];
  uint64_t emitted = 0, visits = 0, matched = 0;
  bool ok = true;
  Primes none{};
  for (uint64_t i = (uint64_t)blockIdx.x * kSiteBlock + threadIdx.x; i < S;
       i += (uint64_t)gridDim.x * kSiteBlock)
    ok &= site_walk<1, false>(keys[i], nullptr, g, t, none, pk, nullptr, visits, matched,
                              [&](uint64_t, const uint32_t*) { emitted += 1; });
  emitted = wave_sum_u64(emitted);
  visits = wave_sum_u64(visits);
  matched = wave_sum_u64(matched);
  if (threadIdx.x == 0) {
    if (emitted) atomicAdd((unsigned long long*)&counters[kCandidates], (unsigned long long)emitted);
    if (visits) atomicAdd((unsigned long long*)&counters[kSiteVisits], (unsigned long long)visits);
    if (matched) atomicAdd((unsigned long long*)&counters[kSiteMatched], (unsigned long long)matched);
  }
  if (!ok) atomicOr((unsigned long long*)&counters[kError], 8ull);
}

__global__ __launch_bounds__(kSiteBlock) void k_site_hist(const uint64_t* __restrict__ keys, uint64_t S,
                                                          RowGeom g, SiteTable t, uint32_t digit_bits,
                                                          uint64_t per_tile, uint32_t* __restrict__ hist) {
  __shared__ uint64_t pk[2 * kSitePartials * kSiteBlock];
  __shared__ uint32_t lh[kExpandDigits];
  const uint32_t D = 1u << digit_bits;
  for (uint32_t i = threadIdx.x; i < D; i += kSiteBlock) lh[i] = 0;
  __syncthreads();
  uint64_t visits = 0, matched = 0;
  Primes none{};
  const uint64_t lo = (uint64_t)blockIdx.x * per_tile;
  const uint64_t hi = lo + per_tile < S ? lo + per_tile : S;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += kSiteBlock)
    site_walk<1, false>(keys[i], nullptr, g, t, none, pk, nullptr, visits, matched,
                        [&](uint64_t sk, const uint32_t*) { atomicAdd(&lh[top_digit(sk, digit_bits)], 1u); });
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < D; i += kSiteBlock) hist[(uint64_t)i * gridDim.x + blockIdx.x] = lh[i];
}

template <int L>
__global__ __launch_bounds__(kSiteBlock) void k_site_scatter(
    const uint64_t* __restrict__ keys, const uint32_t* __restrict__ res, uint64_t S, uint64_t res_stride,
    RowGeom g, SiteTable t, Primes pr, uint32_t digit_bits, uint64_t per_tile,
    const uint64_t* __restrict__ offsets, Rec<L>* __restrict__ out, uint64_t cap,
    uint64_t* __restrict__ counters) {
  __shared__ uint64_t pk[2 * kSitePartials * kSiteBlock];
  __shared__ uint32_t pw[2 * kSitePartials * L * kSiteBlock];
  __shared__ uint32_t lcount[kExpandDigits];
  __shared__ uint64_t gpos[kExpandDigits];
  const uint32_t D = 1u << digit_bits;
  for (uint32_t d = threadIdx.x; d < D; d += kSiteBlock) {
    lcount[d] = 0;
    gpos[d] = offsets[(uint64_t)d * gridDim.x + blockIdx.x];
  }
  __syncthreads();
  uint64_t visits = 0, matched = 0;
  uint32_t overflow = 0;
  const uint64_t lo = (uint64_t)blockIdx.x * per_tile;
  const uint64_t hi = lo + per_tile < S ? lo + per_tile : S;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += kSiteBlock) {
    uint32_t w[L];
#pragma unroll
    for (int l = 0; l < L; ++l) w[l] = res[(uint64_t)l * res_stride + i];
    site_walk<L, true>(keys[i], w, g, t, pr, pk, pw, visits, matched,
                       [&](uint64_t sk, const uint32_t* v) {
                         const uint32_t d = top_digit(sk, digit_bits);
                         const uint64_t pos = gpos[d] + atomicAdd(&lcount[d], 1u);
                         if (pos < cap) {
                           Rec<L> rec;
                           rec.key = sk;
#pragma unroll
                           for (int l = 0; l < L; ++l) rec.r[l] = v[l];
                           out[pos] = rec;
                         } else {
                           overflow = 1;
                         }
                       });
  }
  if (overflow) atomicOr((unsigned long long*)&counters[kError], 2ull);
}

// ------------------------------------------------------------------- scan --

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t hist_value(const uint32_t* __restrict__ h, uint64_t i,
                                               uint32_t W, uint32_t group) {
  if (group == 1) return h[i];
  const uint64_t dc = i / W, w = i - dc * W;
  uint32_t s = 0;
  for (uint32_t k = 0; k < group; ++k) s += h[(dc * group + k) * W + w];
  return s;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ h, uint64_t M,
                                                        uint32_t W, uint32_t group,
                                                        uint64_t* __restrict__ partial) {
  __shared__ uint64_t scratch[kWaves];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
    if (i < M) s += hist_value(h, i, W, group);
  }
  s = block_sum_u64(s, scratch);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// One workgroup: exclusive scan of the per-tile partials; total into
// segs[D] and counters[kCandidates].
__global__ __launch_bounds__(kBlock) void k_scan_partials(uint64_t* __restrict__ partial,
                                                          uint32_t nb, uint64_t* __restrict__ segs,
                                                          uint32_t D, uint64_t* __restrict__ counters) {
  __shared__ uint64_t wave_tot[kWaves];
  uint64_t carry = 0;
  for (uint32_t base = 0; base < nb; base += kBlock) {
    const uint32_t i = base + threadIdx.x;
    const uint64_t v = i < nb ? partial[i] : 0;
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan64(v, &tot, wave_tot);
    if (i < nb) partial[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    segs[D] = carry;
    counters[kCandidates] = carry;
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* __restrict__ h, uint64_t M,
                                                       uint32_t W, uint32_t group,
                                                       const uint64_t* __restrict__ partial,
                                                       uint64_t* __restrict__ offsets,
                                                       uint64_t* __restrict__ segs) {
  __shared__ uint64_t wave_tot[kWaves];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint64_t local = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < M) ? hist_value(h, base + k, W, group) : 0u;
    local += v[k];
  }
  uint64_t tot;
  uint64_t run = partial[blockIdx.x] + block_exclusive_scan64(local, &tot, wave_tot);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint64_t i = base + k;
    if (i < M) {
      offsets[i] = run;
      if (i % W == 0) segs[i / W] = run;
    }
    run += v[k];
  }
}

// -------------------------------------------------------------- partition --
// One more hash-radix digit inside every segment.  Each segment is split into
// `chunks` near-equal ranges so a launch has nseg x chunks workgroups however
// few segments there are: k_part_hist counts the digits of every (segment,
// digit, chunk) at hist[(s * D + d) * chunks + p] -- which is the output
// order, so one exclusive scan gives every chunk's write position per digit
// and the new segment table -- and k_part_scatter writes the records through
// an LDS stage grouped by digit (the next chunk's loads in flight meanwhile).
// PB: widest digit of the launch; the per-digit LDS arrays are sized 2^PB.

__device__ __forceinline__ void chunk_range(const uint64_t* __restrict__ segs, uint64_t in_cap, uint32_t s,
                                            uint32_t p, uint32_t chunks, uint64_t* a, uint64_t* b) {
  uint64_t lo, hi;
  bucket_range(segs, in_cap, s, &lo, &hi);
  const uint64_t len = hi - lo;
  *a = lo + len * p / chunks;
  *b = lo + len * (p + 1) / chunks;
}

// Run mode: chunk p of segment s is the runs [per_seg * p / chunks,
// per_seg * (p + 1) / chunks) of that segment (at most kMaxChunkRuns), each a
// contiguous range of records.  The table holds their starts and the
// exclusive prefix sums of their lengths; pre[kMaxChunkRuns] is the total.
struct RunTable {
  uint64_t start[kMaxChunkRuns];
  uint32_t pre[kMaxChunkRuns + 1];
};

// Fills `t` with wave 0 (the caller then waits at a barrier); returns the
// number of runs of the chunk.
__device__ __forceinline__ uint32_t load_runs(const Runs& rn, uint32_t s, uint32_t p, uint32_t chunks,
                                              RunTable& t) {
  const uint32_t r0 = (uint32_t)((uint64_t)rn.per_seg * p / chunks);
  const uint32_t r1 = (uint32_t)((uint64_t)rn.per_seg * (p + 1) / chunks);
  const uint32_t nr = r1 - r0;
  if (threadIdx.x < 64) {
    const uint32_t lane = threadIdx.x;
    uint32_t len = 0;
    if (lane < nr) {
      const uint64_t idx = (uint64_t)s * rn.per_seg + r0 + lane;
      len = rn.len[idx];
      t.start[lane] = rn.start[idx];
    }
    uint32_t inc = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o, 64);
      if (lane >= (uint32_t)o) inc += y;
    }
    if (lane < nr) t.pre[lane] = inc - len;
    if (lane == 63) t.pre[kMaxChunkRuns] = inc;
  }
  return nr;
}

// Record index of the chunk's j-th record (j < total).  `r` is the thread's
// run cursor (0 at the chunk's start; a thread's j only grow): it moves to
// the last run whose prefix is <= j, which holds j (empty runs share their
// successor's prefix and are stepped over).
__device__ __forceinline__ uint64_t run_record(const RunTable& t, uint32_t nr, uint32_t j, uint32_t& r) {
  while (r + 1 < nr && t.pre[r + 1] <= j) ++r;
  return t.start[r] + (j - t.pre[r]);
}

template <int L, int PB, bool RUNS>
__global__ __launch_bounds__(kBlock) void k_part_hist(const Rec<L>* __restrict__ in,
                                                      const uint64_t* __restrict__ segs, uint64_t in_cap,
                                                      uint32_t nseg, uint32_t chunks, uint32_t shift,
                                                      uint32_t digit_bits, uint32_t kb, Runs rn,
                                                      uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[1 << PB];
  __shared__ RunTable rt[1];  // run mode only (unreferenced otherwise)
  const uint32_t D = 1u << digit_bits;
  const uint64_t dmask = D - 1;
  const uint64_t total = (uint64_t)nseg * chunks;
  for (uint64_t gi = blockIdx.x; gi < total; gi += gridDim.x) {
    const uint32_t s = (uint32_t)(gi / chunks), p = (uint32_t)(gi % chunks);
    uint64_t a = 0, b = 0;
    uint32_t nr = 0;
    if constexpr (RUNS)
      nr = load_runs(rn, s, p, chunks, rt[0]);
    else
      chunk_range(segs, in_cap, s, p, chunks, &a, &b);
    for (uint32_t d = threadIdx.x; d < D; d += kBlock) cnt[d] = 0;
    lds_barrier();
    // Records [a, b): record indices, or (run mode) indices into the
    // chunk's run table.
    if constexpr (RUNS) b = rt[0].pre[kMaxChunkRuns];
    uint32_t rcur = 0;
    for (uint64_t base = a; base < b; base += kStage) {
      uint64_t k[kPerThread];
#pragma unroll
      for (int j = 0; j < kPerThread; ++j) {
        const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
        k[j] = kEmptyKey;
        if (i < b) {
          uint64_t at = i;
          if constexpr (RUNS) at = run_record(rt[0], nr, (uint32_t)i, rcur);
          k[j] = load_key_stream(in + at, kb);
        }
      }
#pragma unroll
      for (int j = 0; j < kPerThread; ++j)
        if (k[j] != kEmptyKey) atomicAdd(&cnt[(uint32_t)((mix64(k[j]) >> shift) & dmask)], 1u);
    }
    lds_barrier();
    for (uint32_t d = threadIdx.x; d < D; d += kBlock) hist[((uint64_t)s * D + d) * chunks + p] = cnt[d];
    lds_barrier();
  }
}

// Per stage of PS records: every thread takes its records' digits and ranks
// (one LDS atomic each, the rank kept in a register), the digit counts are
// scanned, the records are placed in LDS already sorted by digit, and the
// workgroup writes them out linearly -- consecutive threads store consecutive
// addresses of one digit's run.
// The length of those runs bounds the kernel more than occupancy does: one
// 512-thread workgroup per CU holds a stage that fills the LDS (11264 packed
// records, 44 per 8-bit digit per flush), which moves records 1.3x faster
// than four 256-thread workgroups with 2048- or 2560-record stages.
constexpr int kPartBlock = 512;
constexpr int kPartLds = 152 * 1024;  // LDS budget of one partition workgroup
template <int L, int PB, bool FAN>
constexpr int part_stage() {
  // per staged record: the record and its digit; fixed: count, start and
  // gpos per digit, the run table and the wave totals, and (FAN) the
  // fanout counters of the merge workgroups one segment reaches
  constexpr int per = (int)sizeof(Rec<L>) + (PB <= 8 ? 1 : 2);
  constexpr int fixed = (1 << PB) * 16 + 1024 + (FAN ? ((1 << PB) + 2) * 4 + (1 << PB) * 6 + kPartBlock / 8 : 0);
  return (kPartLds - fixed) / per / kPartBlock * kPartBlock;
}

template <int L, int PB, bool FAN>
struct PartStage {
  Rec<L> rec[part_stage<L, PB, FAN>()];
  std::conditional_t<(PB <= 8), uint8_t, uint16_t> digit[part_stage<L, PB, FAN>()];
  // per digit: the stage's count (FAN: in the low word, the records' fanout
  // sum in the high word, both taken by one 64-bit LDS atomic)
  std::conditional_t<FAN, uint64_t, uint32_t> count[1 << PB];
  uint32_t start[1 << PB];
  uint64_t gpos[1 << PB];
  uint32_t wave_tot[kPartBlock / 64];
};

// FIXED: the output buckets have a fixed capacity `out_cap` (bucket (s, d)
// is [(s * D + d) * out_cap, ... + out_cap)) and a fill cursor each
// (cursor[s * D + d], zeroed by the caller); every stage claims each
// digit's run with one atomic add on its cursor, so no histogram pass is
// needed.  A claim past the capacity writes nothing beyond it and sets bit 5
// of counters[kError] (the engine then repeats the level with exact
// offsets).  Otherwise the stage writes at the exact `offsets` of a
// histogram pass and its scan.
// FAN: the last partition level of a row also adds, for every record, the
// number of its legal positions (successors, duplicates counted per record)
// to fan.U[w] of the merge workgroup w = (s * D + d) / fan.per that will own
// its final bucket, and to *fan.total.
template <int L, int PB, bool RUNS, bool FIXED, bool FAN>
__global__ __launch_bounds__(kPartBlock, 1) void k_part_scatter(const Rec<L>* __restrict__ in,
                                                         Rec<L>* __restrict__ out,
                                                         const uint64_t* __restrict__ segs, uint64_t in_cap,
                                                         uint32_t nseg, uint32_t chunks, uint32_t shift,
                                                         uint32_t digit_bits, uint32_t kb, Runs rn,
                                                         uint64_t* __restrict__ offsets, uint64_t out_cap,
                                                         uint64_t cap, uint64_t* __restrict__ counters, FanSum fan) {
  constexpr int PS = part_stage<L, PB, FAN>();
  constexpr int BT = kPartBlock;
  constexpr int PT = PS / BT;
  static_assert(PS % BT == 0, "stage is whole rounds of the workgroup");
  __shared__ PartStage<L, PB, FAN> st;
  __shared__ RunTable rt[1];  // run mode only (unreferenced otherwise)
  __shared__ uint32_t fl[FAN ? (1 << PB) + 2 : 1];
  __shared__ uint16_t wmap[FAN ? (1 << PB) : 1];
  __shared__ uint64_t fsum[FAN ? BT / 64 : 1];
  const uint32_t D = 1u << digit_bits;
  const uint64_t dmask = D - 1;
  const uint64_t total = (uint64_t)nseg * chunks;
  uint32_t overflow = 0;
  uint64_t fan_total = 0;
  // XCD-aware work order: workgroups are dispatched to the XCDs round robin,
  // so workgroup b runs on XCD b % kXcds.  Every chunk of a segment goes to
  // the same XCD, so the stages that append to one output bucket share an
  // L2, which joins their partly written lines before they reach HBM.
  const bool by_xcd = nseg >= (uint32_t)kXcds && gridDim.x % kXcds == 0;
  const uint32_t xcd = blockIdx.x % kXcds;
  const uint64_t items = by_xcd ? (uint64_t)((nseg - xcd + kXcds - 1) / kXcds) * chunks : total;
  const uint64_t first = by_xcd ? blockIdx.x / kXcds : blockIdx.x;
  const uint64_t stride = by_xcd ? gridDim.x / kXcds : gridDim.x;
  for (uint64_t gi = first; gi < items; gi += stride) {
    const uint32_t s = by_xcd ? xcd + (uint32_t)kXcds * (uint32_t)(gi / chunks) : (uint32_t)(gi / chunks);
    const uint32_t p = (uint32_t)(gi % chunks);
    uint64_t a = 0, b = 0;
    uint32_t nr = 0;
    if constexpr (RUNS)
      nr = load_runs(rn, s, p, chunks, rt[0]);
    else
      chunk_range(segs, in_cap, s, p, chunks, &a, &b);
    for (uint32_t d = threadIdx.x; d < D; d += BT) {
      st.count[d] = 0;
      if constexpr (!FIXED) st.gpos[d] = offsets[((uint64_t)s * D + d) * chunks + p];
    }
    uint32_t w0 = 0, span = 0;
    if constexpr (FAN) {
      // merge workgroup of each digit's final bucket, relative to the first
      w0 = (uint32_t)(((uint64_t)s * D) / fan.per);
      span = (uint32_t)(((uint64_t)s * D + D - 1) / fan.per) - w0 + 1;
      for (uint32_t i = threadIdx.x; i < span; i += BT) fl[i] = 0;
      for (uint32_t d = threadIdx.x; d < D; d += BT)
        wmap[d] = (uint16_t)((uint32_t)(((uint64_t)s * D + d) / fan.per) - w0);
    }
    lds_barrier();
    // Records [a, b): record indices, or (run mode) indices into the
    // chunk's run table.
    if constexpr (RUNS) b = rt[0].pre[kMaxChunkRuns];
    uint32_t rcur = 0;
    auto record_at = [&](uint64_t i) -> uint64_t {
      if constexpr (RUNS)
        return run_record(rt[0], nr, (uint32_t)i, rcur);
      else
        return i;
    };
    Rec<L> nx[PT];
#pragma unroll
    for (int j = 0; j < PT; ++j) {
      const uint64_t i = a + (uint64_t)j * BT + threadIdx.x;
      if (i < b) nx[j] = load_stream(in + record_at(i));
    }
    for (uint64_t base = a; base < b; base += PS) {
      const uint32_t nrec = (b - base < (uint64_t)PS) ? (uint32_t)(b - base) : (uint32_t)PS;
      Rec<L> cur[PT];
#pragma unroll
      for (int j = 0; j < PT; ++j) cur[j] = nx[j];
      const uint64_t nb = base + PS;
      if (nb < b) {
#pragma unroll
        for (int j = 0; j < PT; ++j) {
          const uint64_t i = nb + (uint64_t)j * BT + threadIdx.x;
          if (i < b) nx[j] = load_stream(in + record_at(i));
        }
      }
      // 1. digit and rank of every staged record (and its fanout)
      uint32_t dg[PT], rk[PT];
#pragma unroll
      for (int j = 0; j < PT; ++j) {
        const uint32_t slot = (uint32_t)j * BT + threadIdx.x;
        const uint64_t key = rec_key(cur[j], kb);
        dg[j] = (uint32_t)((mix64(key) >> shift) & dmask);
        rk[j] = 0;
        if (slot < nrec) {
          if constexpr (FAN) {
            const uint32_t f = (uint32_t)__popc(legal_positions(key, fan.g));
            rk[j] = (uint32_t)atomicAdd((unsigned long long*)&st.count[dg[j]], ((unsigned long long)f << 32) | 1ull);
            fan_total += f;
          } else {
            rk[j] = atomicAdd(&st.count[dg[j]], 1u);
          }
        }
      }
      lds_barrier();
      scan_digits<BT>(st.count, st.start, D, st.wave_tot);
      // FAN: the stage's fanout per digit, onto its merge workgroup's counter
      if constexpr (FAN) {
        for (uint32_t d = threadIdx.x; d < D; d += BT) {
          const uint32_t f = (uint32_t)(st.count[d] >> 32);
          if (f) atomicAdd(&fl[wmap[d]], f);
        }
      }
      // FIXED: claim every digit's run of this stage in its output bucket
      // (the claims' round trips overlap the placement below)
      constexpr int kClaims = FIXED ? ((1 << PB) + BT - 1) / BT : 1;
      uint64_t claim[kClaims];
      if constexpr (FIXED) {
#pragma unroll
        for (int c = 0; c < kClaims; ++c) {
          const uint32_t d = threadIdx.x + (uint32_t)c * BT;
          claim[c] = 0;
          const uint32_t cnt = d < D ? (uint32_t)st.count[d] : 0u;
          if (cnt) claim[c] = atomicAdd((unsigned long long*)&offsets[(uint64_t)s * D + d], (unsigned long long)cnt);
        }
      }
      // 2. place the records sorted by digit
#pragma unroll
      for (int j = 0; j < PT; ++j) {
        const uint32_t slot = (uint32_t)j * BT + threadIdx.x;
        if (slot < nrec) {
          const uint32_t at = st.start[dg[j]] + rk[j];
          st.rec[at] = cur[j];
          st.digit[at] = (std::remove_reference_t<decltype(st.digit[0])>)dg[j];
        }
      }
      if constexpr (FIXED) {
#pragma unroll
        for (int c = 0; c < kClaims; ++c) {
          const uint32_t d = threadIdx.x + (uint32_t)c * BT;
          if (d < D) st.gpos[d] = ((uint64_t)s * D + d) * out_cap + claim[c];
        }
      }
      lds_barrier();
      // 3. linear write-out of the digit runs
      for (uint32_t i = threadIdx.x; i < nrec; i += BT) {
        const uint32_t d = st.digit[i];
        const uint64_t pos = st.gpos[d] + (i - st.start[d]);
        bool ok = pos < cap;
        if constexpr (FIXED) ok = ok && pos < ((uint64_t)s * D + d + 1) * out_cap;
        if (ok)
          out[pos] = st.rec[i];
        else
          overflow |= FIXED && pos < cap ? 32u : 2u;
      }
      lds_barrier();
      for (uint32_t d = threadIdx.x; d < D; d += BT) {
        if constexpr (!FIXED) st.gpos[d] += (uint32_t)st.count[d];
        st.count[d] = 0;
      }
      lds_barrier();
    }
    if constexpr (FAN) {
      lds_barrier();
      for (uint32_t i = threadIdx.x; i < span; i += BT)
        if (fl[i]) atomicAdd((unsigned long long*)&fan.U[w0 + i], (unsigned long long)fl[i]);
    }
    // The run table (and the fanout counters) are rewritten by the next
    // chunk: every wave must be past its reads (an empty chunk runs no stage
    // and so no barrier above).
    if constexpr (RUNS || FAN) lds_barrier();
  }
  if constexpr (FAN) {
    const uint64_t t = block_sum_u64<BT / 64>(fan_total, fsum);
    if (threadIdx.x == 0 && t) atomicAdd((unsigned long long*)fan.total, (unsigned long long)t);
  }
  if (overflow & 2u) atomicOr((unsigned long long*)&counters[kError], 2ull);
  if (overflow & 32u) atomicOr((unsigned long long*)&counters[kRetry], 32ull);
}

// ----------------------------------------------------------- merge-expand --
// Workgroup w owns the buckets [w*per, (w+1)*per) of row r; the COUNT and
// SCATTER launches use the same assignment, so they see identical merged
// states.  Per bucket:
//   1. every record inserts its key into an LDS open-addressing table (linear
//      probing on the mix's low bits); the record that claims an empty slot
//      is the slot's owner and appends the slot to a list of unique slots
//      (one LDS atomic per wave);
//   2. owners store their residues into the slot (plain stores);
//   3. the rare duplicates add theirs modulo each prime;
//   4. the unique list is walked: layer statistics, then every merged state
//      is expanded through the next row and its successors are counted
//      (COUNT) or written (SCATTER) by next-row level-1 digit.
// A bucket above the table's safe load is merged in passes that each take
// the keys whose mix falls in one class (bits 12-35), so passes see disjoint
// keys.  Probe exhaustion sets kError; the call then fails.

// DB: widest next-row digit the kernel can write (LDS per-digit arrays are
// sized 2^DB; SCATTER at <= 8 bits keeps four workgroups per CU).
template <int L, int MODE, int T, int DB = kFineBits, int BT = kBlock>
__global__ __launch_bounds__(BT, (BT == kScatterBlock && L <= 2 && T == 2048 && DB <= kNarrowDigitBits ? 8 : 1)) void k_merge_expand(
    const Rec<L>* __restrict__ in, const uint64_t* __restrict__ segs, uint64_t in_cap, uint32_t nseg,
    uint32_t per, RowGeom g, Primes pr, uint32_t digit_bits, uint32_t* __restrict__ hist,
    const uint64_t* __restrict__ offsets, const uint64_t* __restrict__ capw, Rec<L>* __restrict__ out,
    uint64_t cap, uint64_t final_mask, uint64_t* __restrict__ slice_sums, uint64_t nslices,
    uint64_t* __restrict__ counters) {
  constexpr uint32_t kCap = (T * 5) / 8;             // records per pass
  constexpr int KB = (int)((kCap + BT - 1) / BT);  // records per thread per batch
  constexpr uint32_t kSpan = (uint32_t)KB * BT;  // records per batch (>= kCap)
  constexpr bool P = L == kPackedRecords;           // packed records: exact weights
  constexpr int NL = P ? 1 : L;                     // weight words per record
  constexpr int LL = (MODE == kMergeCount || P) ? 1 : L;  // residue rows of the table
  constexpr bool kExpand = MODE == kMergeCount || MODE == kMergeScatter;
  constexpr uint32_t DM = kExpand ? (1u << DB) : 1;
  static_assert(!(P && MODE == kMergeEmit), "EMIT writes lane records only");
  __shared__ uint64_t tkey[T];
  __shared__ uint32_t tres[LL][T];                            // lane records
  __shared__ uint64_t tw[(P && MODE != kMergeCount) ? T : 1];  // packed: exact weights
  __shared__ uint16_t ulist[T];
  __shared__ uint32_t lcount[MODE == kMergeScatter ? 1 : DM];  // COUNT: digit histogram
  // SCATTER: byte cursor in `out` of this workgroup's run of each digit (one
  // LDS atomic gives a successor its store address).
  __shared__ uint64_t gpos[MODE == kMergeScatter ? DM : 1];
  // successor_delta of every column (index: the column's bit position).
  __shared__ uint64_t dtab[kExpand ? 32 : 1];
  __shared__ uint64_t scratch[BT / 64];
  __shared__ uint32_t nuniq[2];
  // Last row: per-slice sums gathered in LDS, flushed with one global atomic
  // per (slice, lane) and workgroup.
  // Packed records keep one exact sum per slice (copied into every lane).
  constexpr uint32_t kLdsSlices = MODE == kMergeLast ? 1024u : 1u;
  __shared__ uint64_t ssl[kLdsSlices * NL];
  const bool lds_slices = MODE == kMergeLast && slice_sums && nslices <= kLdsSlices;
  const uint32_t D = kExpand ? (1u << digit_bits) : 1u;
  const uint32_t W = gridDim.x;
  const int lane = threadIdx.x & 63;

  for (uint32_t i = threadIdx.x; i < T; i += BT) {
    tkey[i] = kEmptyKey;
    if constexpr (P && MODE != kMergeCount) tw[i] = 0;
  }
  if (kExpand) {
    for (uint32_t d = threadIdx.x; d < D; d += BT) {
      if constexpr (MODE == kMergeScatter)
        gpos[d] = offsets[(uint64_t)d * W + blockIdx.x] * sizeof(Rec<L>);
      else
        lcount[d] = 0;
    }
    if (threadIdx.x < 32) dtab[threadIdx.x] = threadIdx.x < g.n ? successor_delta(1u << threadIdx.x, g) : 0;
  }
  if (threadIdx.x < 2) nuniq[threadIdx.x] = 0;
  if (lds_slices)
    for (uint32_t i = threadIdx.x; i < nslices * NL; i += BT) ssl[i] = 0;
  lds_barrier();

  uint32_t uniq = 0;  // merged states of this thread
  // Lane records: per-thread residue sums.  Packed: exact 128-bit sums
  // (index 0 low, 1 high) reduced to residues at the end.
  uint64_t wsum[P ? 2 : L], fsum[P ? 2 : L];
#pragma unroll
  for (int l = 0; l < (P ? 2 : L); ++l) wsum[l] = fsum[l] = 0;
  uint32_t overflow = 0;
  uint32_t parity = 0;

  const uint32_t b0 = blockIdx.x * per;
  const uint32_t b1 = (b0 + per < nseg) ? b0 + per : nseg;

  Rec<L> buf[KB];
  auto load_batch = [&](uint64_t base, uint64_t hi) {
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const uint64_t i = base + (uint64_t)k * BT + threadIdx.x;
      if (i < hi) {
        if constexpr (MODE == kMergeCount && P) {
          buf[k].w[0] = in[i].w[0];
          buf[k].w[1] = in[i].w[1];
        } else if constexpr (MODE == kMergeCount) {
          buf[k].key = in[i].key;
        } else {
          buf[k] = load_stream(in + i);
        }
      }
    }
  };
  uint64_t plo = 0, phi = 0;
  if (b0 < b1) {
    bucket_range(segs, in_cap, b0, &plo, &phi);
    if (phi - plo <= kCap) load_batch(plo, phi);
  }
  // Capped SCATTER runs: at most wcap successors per digit.  A run that
  // ends longer sets bit 4 of counters[kRetry] (checked once, after the last
  // bucket); its surplus stores have then landed in the ranges of later runs
  // of the same output region, and the engine repeats the whole SCATTER with
  // upper-bound offsets, which rewrites every position the next level reads.
  // Every store stays inside the output region.
  const uint64_t wcap = (MODE == kMergeScatter && capw) ? capw[blockIdx.x] : ~0ull;
  const uint64_t out_end = cap * sizeof(Rec<L>);  // bytes of the output region

  for (uint32_t b = b0; b < b1; ++b) {
    const uint64_t lo = plo, hi = phi;
    const uint64_t m = hi - lo;
    const uint32_t passes = m <= kCap ? 1u : (uint32_t)((m + kCap - 1) / kCap);
    uint64_t nlo = 0, nhi = 0;
    if (b + 1 < b1) bucket_range(segs, in_cap, b + 1, &nlo, &nhi);
    for (uint32_t pass = 0; pass < passes; ++pass) {
      uint32_t* const nu_ctr = &nuniq[parity];
      for (uint64_t base = lo; base < hi; base += kSpan) {
        if (passes > 1) load_batch(base, hi);
        // ---- 1. insert keys; owners append their slot to the unique list
        uint32_t own = 0, dup = 0;
        uint32_t sl[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          sl[k] = 0;
          const uint64_t i = base + (uint64_t)k * BT + threadIdx.x;
          if (i >= hi) continue;
          const uint64_t key = rec_key(buf[k], g.key_bits);
          const uint64_t h = mix64(key);
          if (passes > 1 && (uint32_t)((((h >> 12) & 0xffffffull) * passes) >> 24) != pass) continue;
          uint32_t slot = (uint32_t)h & (T - 1);
          bool placed = false;
          for (uint32_t probe = 0; probe < T; ++probe) {
            const unsigned long long old = atomicCAS((unsigned long long*)&tkey[slot],
                                                     (unsigned long long)kEmptyKey,
                                                     (unsigned long long)key);
            if (old == kEmptyKey) {
              own |= 1u << k;
              placed = true;
              break;
            }
            if (old == key) {
              dup |= 1u << k;
              placed = true;
              break;
            }
            slot = (slot + 1) & (T - 1);
          }
          if (!placed) overflow = 1;
          sl[k] = slot;
        }
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          const bool o = (own >> k) & 1u;
          const uint64_t bal = __ballot(o);
          if (bal == 0) continue;
          const int leader = __ffsll((long long)bal) - 1;
          uint32_t base_idx = 0;
          if (lane == leader) base_idx = atomicAdd(nu_ctr, (uint32_t)__popcll(bal));
          base_idx = __builtin_amdgcn_readlane(base_idx, leader);
          if (o) {
            const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            ulist[base_idx + r] = (uint16_t)sl[k];
          }
        }
        if constexpr (P && MODE != kMergeCount) {
          // ---- 2+3. packed: every record adds its exact weight into its
          // slot's sum (zero while the slot is free), owner or duplicate
          // alike, so no barrier has to order the owners' stores first (a
          // bucket's sum of weights below 2^52 stays below 2^64).
#pragma unroll
          for (int k = 0; k < KB; ++k)
            if (((own | dup) >> k) & 1u)
              atomicAdd((unsigned long long*)&tw[sl[k]], (unsigned long long)packed_weight(buf[k], g.key_bits));
        } else if constexpr (!P && MODE != kMergeCount) {
          lds_barrier();
          // ---- 2. owners store their residues
#pragma unroll
          for (int k = 0; k < KB; ++k)
            if ((own >> k) & 1u) {
#pragma unroll
              for (int l = 0; l < NL; ++l) tres[l < LL ? l : 0][sl[k]] = buf[k].r[l];
            }
          lds_barrier();
          // ---- 3. duplicates add modulo each prime
          if (dup) {
#pragma unroll
            for (int k = 0; k < KB; ++k)
              if ((dup >> k) & 1u) {
#pragma unroll
                for (int l = 0; l < NL; ++l) {
                  uint32_t* a = &tres[l < LL ? l : 0][sl[k]];
                  uint32_t old = *a, assumed;
                  do {
                    assumed = old;
                    old = atomicCAS(a, assumed, addmod(assumed, buf[k].r[l], pr.p[l]));
                  } while (old != assumed);
                }
              }
          }
        }
        lds_barrier();
      }
      // The batch registers are free: prefetch the next bucket.
      if (pass + 1 == passes && b + 1 < b1 && nhi - nlo <= kCap) load_batch(nlo, nhi);
      // ---- 4. statistics and expansion of the merged states
      const uint32_t nu = *nu_ctr;
      for (uint32_t u = threadIdx.x; u < nu; u += BT) {
        const uint32_t slot = ulist[u];
        const uint64_t key = tkey[slot];
        tkey[slot] = kEmptyKey;
        uniq += 1;
        uint32_t v[NL];
        uint64_t wv = 0;  // packed: the merged exact weight
        if (MODE != kMergeCount) {
          if constexpr (P) {
            wv = tw[slot];
            tw[slot] = 0;  // the slot is free again
            if (wv >> g.weight_bits) overflow |= 4u;
            wsum[0] += wv;
            wsum[1] += wsum[0] < wv;
          } else {
#pragma unroll
            for (int l = 0; l < L; ++l) {
              v[l] = tres[l < LL ? l : 0][slot];
              wsum[l] += v[l];
            }
          }
          // Complete boundaries (every column set) end here: the final v1/v2
          // contraction sums them, per slice when slices are tracked.  Only
          // the last row's merge (LAST, or EMIT on the materialised path)
          // can hold them: every row places one column.
          if (MODE != kMergeScatter && final_mask && (key & final_mask) == final_mask) {
            const uint64_t sl = (key & ~g.state_mask) >> (3 * g.n);
            if constexpr (P) {
              fsum[0] += wv;
              fsum[1] += fsum[0] < wv;
              if (slice_sums) {
                if (lds_slices)
                  atomicAdd((unsigned long long*)&ssl[sl], (unsigned long long)wv);
                else
                  for (uint32_t l = 0; l < pr.lanes; ++l)
                    atomicAdd((unsigned long long*)&slice_sums[sl * pr.lanes + l], (unsigned long long)wv);
              }
            } else {
#pragma unroll
              for (int l = 0; l < L; ++l) fsum[l] += v[l];
              if (slice_sums) {
#pragma unroll
                for (int l = 0; l < L; ++l) {
                  if (lds_slices)
                    atomicAdd((unsigned long long*)&ssl[sl * L + l], (unsigned long long)v[l]);
                  else
                    atomicAdd((unsigned long long*)&slice_sums[sl * L + l], (unsigned long long)v[l]);
                }
              }
            }
          }
        }
        if constexpr (MODE == kMergeEmit) {
          // One global atomic per wave: lanes take consecutive positions.
          uint64_t* ek = (uint64_t*)out;
          uint32_t* er = (uint32_t*)(ek + cap);
          const uint64_t act = __ballot(1);
          const int leader = __ffsll((long long)act) - 1;
          unsigned long long base_pos = 0;
          if (lane == leader)
            base_pos = atomicAdd((unsigned long long*)&counters[kEmitted], (unsigned long long)__popcll(act));
          base_pos = __shfl(base_pos, leader, 64);
          const uint64_t pos = base_pos + __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
          if (pos < cap) {
            ek[pos] = key;
#pragma unroll
            for (int l = 0; l < L; ++l) er[(uint64_t)l * cap + pos] = v[l];
          } else {
            overflow |= 2u;
          }
        }
        if (kExpand) {
          const Board bs = unpack(key, g);
          uint32_t free = legal_positions(bs, g);
          const uint64_t sbase = successor_base(key, bs, g);
          // Two successors per step: their table reads and run-cursor atomics
          // are in flight together.
          auto emit = [&](uint64_t sk, uint64_t at) {
            Rec<L> rec;
            if constexpr (P) {
              rec = make_packed(sk, wv, g.key_bits);
            } else {
              rec.key = sk;
#pragma unroll
              for (int l = 0; l < L; ++l) rec.r[l] = v[l];
            }
            if (at < out_end)
              *(Rec<L>*)((char*)out + at) = rec;
            else
              overflow |= 2u;
          };
          while (free) {
            const uint32_t c0 = (uint32_t)__builtin_ctz(free);
            free &= free - 1;
            const bool two = free != 0;
            const uint32_t c1 = two ? (uint32_t)__builtin_ctz(free) : c0;
            free &= free - 1;  // no-op when free is already 0
            const uint64_t sk0 = sbase ^ dtab[c0];
            const uint64_t sk1 = sbase ^ dtab[c1];
            const uint32_t d0 = top_digit(sk0, digit_bits);
            const uint32_t d1 = top_digit(sk1, digit_bits);
            if (MODE == kMergeCount) {
              atomicAdd(&lcount[d0], 1u);
              if (two) atomicAdd(&lcount[d1], 1u);
            } else {
              constexpr unsigned long long kStep = sizeof(Rec<L>);
              const uint64_t a0 = atomicAdd((unsigned long long*)&gpos[d0], kStep);
              uint64_t a1 = 0;
              if (two) a1 = atomicAdd((unsigned long long*)&gpos[d1], kStep);
              emit(sk0, a0);
              if (two) emit(sk1, a1);
            }
          }
        }
      }
      if (threadIdx.x == 0) nuniq[parity ^ 1] = 0;
      parity ^= 1;
      lds_barrier();
    }
    plo = nlo;
    phi = nhi;
  }

  // ---- flush statistics
  if constexpr (MODE == kMergeCount) {
    // the digit histogram
    for (uint32_t d = threadIdx.x; d < D; d += BT) hist[(uint64_t)d * W + blockIdx.x] = lcount[d];
  }
  if constexpr (MODE == kMergeScatter) {
    // Run lengths from the cursors; with upper-bound offsets (hist set) the
    // length of every run is written out (the rest of its range is a gap).
    uint64_t wr = 0;
    for (uint32_t d = threadIdx.x; d < D; d += BT) {
      const uint64_t len = gpos[d] / sizeof(Rec<L>) - offsets[(uint64_t)d * W + blockIdx.x];
      if (hist) hist[(uint64_t)d * W + blockIdx.x] = (uint32_t)len;
      if (len > wcap) overflow |= 16u;
      wr += len;
    }
    // A store past the region: with capped runs, the surplus of an
    // overflowing run (a retry); with exact or upper-bound offsets, an error.
    if (capw && (overflow & 2u)) overflow = (overflow & ~2u) | 16u;
    wr = block_sum_u64<BT / 64>(wr, scratch);
    if (threadIdx.x == 0 && wr) atomicAdd((unsigned long long*)&counters[kWritten], (unsigned long long)wr);
  }
  if (MODE != kMergeCount) {
    const uint64_t u = block_sum_u64<BT / 64>((uint64_t)uniq, scratch);
    if (threadIdx.x == 0 && u) atomicAdd((unsigned long long*)&counters[kOutStates], (unsigned long long)u);
    if constexpr (P) {
      // pr.lanes is uniform, so every thread takes the same block sums.
      for (uint32_t l = 0; l < pr.lanes; ++l) {
        const uint64_t w = block_sum_u64<BT / 64>(mod_u128(wsum[1], wsum[0], pr.p[l]), scratch);
        if (threadIdx.x == 0 && w) atomicAdd((unsigned long long*)&counters[kWeight0 + l], (unsigned long long)(w % pr.p[l]));
        const uint64_t f = block_sum_u64<BT / 64>(mod_u128(fsum[1], fsum[0], pr.p[l]), scratch);
        if (threadIdx.x == 0 && f) atomicAdd((unsigned long long*)&counters[kFinal0 + l], (unsigned long long)(f % pr.p[l]));
      }
    } else {
#pragma unroll
      for (int l = 0; l < L; ++l) {
        const uint64_t w = block_sum_u64<BT / 64>(wsum[l] % pr.p[l], scratch);
        if (threadIdx.x == 0 && w) atomicAdd((unsigned long long*)&counters[kWeight0 + l], (unsigned long long)(w % pr.p[l]));
        const uint64_t f = block_sum_u64<BT / 64>(fsum[l] % pr.p[l], scratch);
        if (threadIdx.x == 0 && f) atomicAdd((unsigned long long*)&counters[kFinal0 + l], (unsigned long long)(f % pr.p[l]));
      }
    }
  }
  if (lds_slices) {
    lds_barrier();
    if constexpr (P) {
      for (uint32_t i = threadIdx.x; i < nslices; i += BT)
        if (ssl[i])
          for (uint32_t l = 0; l < pr.lanes; ++l)
            atomicAdd((unsigned long long*)&slice_sums[i * pr.lanes + l], (unsigned long long)ssl[i]);
    } else {
      for (uint32_t i = threadIdx.x; i < nslices * L; i += BT)
        if (ssl[i]) atomicAdd((unsigned long long*)&slice_sums[i], (unsigned long long)ssl[i]);
    }
  }
  if (overflow & ~16u) atomicOr((unsigned long long*)&counters[kError], (unsigned long long)(overflow & ~16u));
  if (overflow & 16u) atomicOr((unsigned long long*)&counters[kRetry], 16ull);
}

uint32_t grid_for(uint64_t items, uint32_t per_block, uint32_t cap) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return g > cap ? cap : (uint32_t)g;
}

}  // namespace

size_t record_bytes(int lanes) {
  static_assert(sizeof(Packed) == 12, "packed records are 12 bytes");
  switch (lanes) {
    case kPackedRecords: return sizeof(Packed);
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
  if (digit_bits > kFineBits || tiles == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_expand_hist, dim3(tiles), dim3(kBlock), 0, stream, keys, S, g, digit_bits,
                     per_tile, hist);
  return hipGetLastError();
}

hipError_t launch_expand_scatter(int lanes, const uint64_t* keys, const uint32_t* res, uint64_t S,
                                 uint64_t res_stride, const RowGeom& g, const Primes& pr,
                                 uint32_t digit_bits, uint32_t tiles, uint64_t per_tile,
                                 const uint64_t* offsets, void* out, uint64_t cap,
                                 uint64_t* counters, hipStream_t stream) {
  if (digit_bits > kFineBits || tiles == 0) return hipErrorInvalidValue;
  dim3 grid(tiles), block(kBlock);
  switch (lanes) {
    case kPackedRecords: hipLaunchKernelGGL(k_expand_scatter<kPackedRecords>, grid, block, 0, stream, keys, res, S, res_stride, g, pr, digit_bits, per_tile, offsets, (Packed*)out, cap, counters); break;
    case 1: hipLaunchKernelGGL(k_expand_scatter<1>, grid, block, 0, stream, keys, res, S, res_stride, g, pr, digit_bits, per_tile, offsets, (Rec<1>*)out, cap, counters); break;
    case 2: hipLaunchKernelGGL(k_expand_scatter<2>, grid, block, 0, stream, keys, res, S, res_stride, g, pr, digit_bits, per_tile, offsets, (Rec<2>*)out, cap, counters); break;
    case 3: hipLaunchKernelGGL(k_expand_scatter<3>, grid, block, 0, stream, keys, res, S, res_stride, g, pr, digit_bits, per_tile, offsets, (Rec<3>*)out, cap, counters); break;
    case 4: hipLaunchKernelGGL(k_expand_scatter<4>, grid, block, 0, stream, keys, res, S, res_stride, g, pr, digit_bits, per_tile, offsets, (Rec<4>*)out, cap, counters); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_fill(uint64_t* p, uint64_t n, uint64_t v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, kBlock * 4, 1024)), dim3(kBlock), 0, stream, p, n, v);
  return hipGetLastError();
}

hipError_t launch_scan(const uint32_t* hist, uint32_t D, uint32_t W, uint32_t group,
                       uint64_t* offsets, uint64_t* segs, uint64_t* partial, uint64_t* counters,
                       hipStream_t stream) {
  const uint64_t M = (uint64_t)D * W;
  if (M == 0 || group == 0) return hipErrorInvalidValue;
  const uint32_t nb = (uint32_t)((M + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, stream, hist, M, W, group, partial);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kBlock), 0, stream, partial, nb, segs, D, counters);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, stream, hist, M, W, group,
                     (const uint64_t*)partial, offsets, segs);
  return hipGetLastError();
}

uint32_t partition_chunks(uint64_t records, uint64_t nseg, bool runs) {
  const uint64_t per = nseg * (runs ? kPartChunkRuns : kPartChunk);
  const uint64_t c = (records + per - 1) / per;
  return (uint32_t)(c < 1 ? 1 : (c > kMaxPartChunks ? kMaxPartChunks : c));
}

hipError_t launch_part_hist(int lanes, const void* in, const Buckets& bk, uint32_t chunks, uint32_t shift,
                            uint32_t digit_bits, uint32_t key_bits, const Runs* runs, uint32_t* hist,
                            hipStream_t stream) {
  const uint32_t nseg = bk.nseg;
  if (digit_bits == 0 || digit_bits > kMaxDigitBits || nseg == 0 || chunks == 0) return hipErrorInvalidValue;
  if (((uint64_t)nseg << digit_bits) > 0xffffffffull) return hipErrorInvalidValue;
  if (runs && (runs->per_seg == 0 || (runs->per_seg + chunks - 1) / chunks > kMaxChunkRuns))
    return hipErrorInvalidValue;
  const uint64_t total = (uint64_t)nseg * chunks;
  dim3 grid(total < 65536 ? (uint32_t)total : 65536u), block(kBlock);
  const bool narrow = digit_bits <= (uint32_t)kNarrowDigitBits;
  const bool mid = digit_bits <= (uint32_t)kMidDigitBits;
  const Runs rn = runs ? *runs : Runs{nullptr, nullptr, 0};
#define PEPS_PH_ONE(LN, PB, RUNS)                                                                          \
  hipLaunchKernelGGL((k_part_hist<LN, PB, RUNS>), grid, block, 0, stream, (const Rec<LN>*)in, bk.segs, bk.cap, \
                     nseg, chunks, shift, digit_bits, key_bits, rn, hist)
#define PEPS_PH(LN)                                                                      \
  if (narrow && runs)                                                                    \
    PEPS_PH_ONE(LN, kNarrowDigitBits, true);                                             \
  else if (narrow)                                                                       \
    PEPS_PH_ONE(LN, kNarrowDigitBits, false);                                            \
  else if (mid && runs)                                                                  \
    PEPS_PH_ONE(LN, kMidDigitBits, true);                                                \
  else if (mid)                                                                          \
    PEPS_PH_ONE(LN, kMidDigitBits, false);                                               \
  else if (runs)                                                                         \
    PEPS_PH_ONE(LN, kMaxDigitBits, true);                                                \
  else                                                                                   \
    PEPS_PH_ONE(LN, kMaxDigitBits, false);                                               \
  break
  switch (lanes) {
    case kPackedRecords: PEPS_PH(kPackedRecords);
    case 1: PEPS_PH(1);
    case 2: PEPS_PH(2);
    case 3: PEPS_PH(3);
    case 4: PEPS_PH(4);
    default: return hipErrorInvalidValue;
  }
#undef PEPS_PH
#undef PEPS_PH_ONE
  return hipGetLastError();
}

hipError_t launch_part_scatter(int lanes, const void* in, void* out, const Buckets& bk, uint32_t chunks,
                               uint32_t shift, uint32_t digit_bits, uint32_t key_bits, const Runs* runs,
                               uint64_t* offsets, uint64_t out_cap, uint64_t cap, uint64_t* counters,
                               const FanSum* fan, hipStream_t stream) {
  const uint32_t nseg = bk.nseg;
  if (digit_bits == 0 || digit_bits > kMaxDigitBits || nseg == 0 || chunks == 0) return hipErrorInvalidValue;
  if (runs && (runs->per_seg == 0 || (runs->per_seg + chunks - 1) / chunks > kMaxChunkRuns))
    return hipErrorInvalidValue;
  if (out_cap && ((uint64_t)nseg << digit_bits) > cap / out_cap) return hipErrorInvalidValue;
  if (fan && (out_cap == 0 || fan->U == nullptr || fan->total == nullptr || fan->per == 0 ||
              (uint64_t)fan->W * fan->per < ((uint64_t)nseg << digit_bits)))
    return hipErrorInvalidValue;
  const uint64_t total = (uint64_t)nseg * chunks;
  // a whole number of workgroups per XCD (the kernel's XCD-aware order)
  uint32_t g = total < 65536 ? (uint32_t)total : 65536u;
  if (nseg >= (uint32_t)kXcds && g >= (uint32_t)kXcds) g -= g % kXcds;
  dim3 grid(g);
  const bool narrow = digit_bits <= (uint32_t)kNarrowDigitBits;
  const bool mid = digit_bits <= (uint32_t)kMidDigitBits;
  const Runs rn = runs ? *runs : Runs{nullptr, nullptr, 0};
  const FanSum fs = fan ? *fan : FanSum{nullptr, nullptr, 0, 1, RowGeom{}};
  const bool fixed = out_cap != 0;
#define PEPS_PS_ONE(LN, PB, RUNS, FIXED, FAN)                                                                    \
  hipLaunchKernelGGL((k_part_scatter<LN, PB, RUNS, FIXED, FAN>), grid, dim3(kPartBlock), 0, stream,                \
                     (const Rec<LN>*)in, (Rec<LN>*)out, bk.segs, bk.cap, nseg, chunks, shift, digit_bits, key_bits, \
                     rn, offsets, out_cap, cap, counters, fs)
  // Run-mode levels read SCATTER's runs; fixed-capacity levels may fold in
  // the fanout sums; exact levels are the fallback and the materialised path.
#define PEPS_PS_PB(LN, PB)                            \
  if (runs && fixed && fan)                           \
    PEPS_PS_ONE(LN, PB, true, true, true);            \
  else if (runs && fixed)                             \
    PEPS_PS_ONE(LN, PB, true, true, false);           \
  else if (runs)                                      \
    PEPS_PS_ONE(LN, PB, true, false, false);          \
  else if (fixed && fan)                              \
    PEPS_PS_ONE(LN, PB, false, true, true);           \
  else if (fixed)                                     \
    PEPS_PS_ONE(LN, PB, false, true, false);          \
  else                                                \
    PEPS_PS_ONE(LN, PB, false, false, false)
#define PEPS_PS(LN)                                   \
  if (narrow)                                         \
    PEPS_PS_PB(LN, kNarrowDigitBits);                 \
  else if (mid)                                       \
    PEPS_PS_PB(LN, kMidDigitBits);                    \
  else                                                \
    PEPS_PS_PB(LN, kMaxDigitBits);                    \
  break
  switch (lanes) {
    case kPackedRecords: PEPS_PS(kPackedRecords);
    case 1: PEPS_PS(1);
    case 2: PEPS_PS(2);
    case 3: PEPS_PS(3);
    case 4: PEPS_PS(4);
    default: return hipErrorInvalidValue;
  }
#undef PEPS_PS
#undef PEPS_PS_PB
#undef PEPS_PS_ONE
  return hipGetLastError();
}

hipError_t launch_count_ub(int lanes, const void* in, const Buckets& bk, uint32_t W, const RowGeom& g,
                           uint32_t digit_bits, uint32_t* hist, hipStream_t stream) {
  const uint32_t nseg = bk.nseg;
  if (nseg == 0 || W == 0 || digit_bits > kFineBits) return hipErrorInvalidValue;
  const uint32_t per = (nseg + W - 1) / W;
#define PEPS_CU(LN)                                                                                              \
  hipLaunchKernelGGL(k_count_ub<LN>, dim3(W), dim3(kBlock), 0, stream, (const Rec<LN>*)in, bk.segs, bk.cap, nseg, \
                     per, g, digit_bits, hist)
  switch (lanes) {
    case kPackedRecords: PEPS_CU(kPackedRecords); break;
    case 1: PEPS_CU(1); break;
    case 2: PEPS_CU(2); break;
    case 3: PEPS_CU(3); break;
    case 4: PEPS_CU(4); break;
    default: return hipErrorInvalidValue;
  }
#undef PEPS_CU
  return hipGetLastError();
}

hipError_t launch_caps(const uint64_t* U, uint32_t W, uint32_t digit_bits, uint32_t shrink, uint64_t* capw,
                       uint64_t* pre, uint64_t* run_start, uint64_t* counters, hipStream_t stream) {
  if (W == 0 || digit_bits == 0 || digit_bits > kFineBits || shrink > 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_caps_scan, dim3(1), dim3(kCapsBlock), 0, stream, U, W, digit_bits, shrink, capw, pre,
                     counters);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint64_t total = (uint64_t)W << digit_bits;
  hipLaunchKernelGGL(k_caps_expand, dim3(grid_for(total, kBlock * 4, 4096)), dim3(kBlock), 0, stream, pre, W,
                     digit_bits, run_start);
  return hipGetLastError();
}

hipError_t launch_site_count(const uint64_t* keys, uint64_t S, const RowGeom& g, const SiteTable& t,
                             uint64_t* counters, hipStream_t stream) {
  if (S == 0) return hipSuccess;
  const uint32_t grid = grid_for(S, kSiteBlock, 8192);
  hipLaunchKernelGGL(k_site_count, dim3(grid), dim3(kSiteBlock), 0, stream, keys, S, g, t, counters);
  return hipGetLastError();
}

hipError_t launch_site_hist(const uint64_t* keys, uint64_t S, const RowGeom& g, const SiteTable& t,
                            uint32_t digit_bits, uint32_t tiles, uint64_t per_tile, uint32_t* hist,
                            hipStream_t stream) {
  if (digit_bits > kFineBits || tiles == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_site_hist, dim3(tiles), dim3(kSiteBlock), 0, stream, keys, S, g, t, digit_bits,
                     per_tile, hist);
  return hipGetLastError();
}

hipError_t launch_site_scatter(int lanes, const uint64_t* keys, const uint32_t* res, uint64_t S,
                               uint64_t res_stride, const RowGeom& g, const SiteTable& t, const Primes& pr,
                               uint32_t digit_bits, uint32_t tiles, uint64_t per_tile,
                               const uint64_t* offsets, void* out, uint64_t cap, uint64_t* counters,
                               hipStream_t stream) {
  if (digit_bits > kFineBits || tiles == 0) return hipErrorInvalidValue;
  dim3 grid(tiles), block(kSiteBlock);
  switch (lanes) {
    case 1: hipLaunchKernelGGL(k_site_scatter<1>, grid, block, 0, stream, keys, res, S, res_stride, g, t, pr, digit_bits, per_tile, offsets, (Rec<1>*)out, cap, counters); break;
    case 2: hipLaunchKernelGGL(k_site_scatter<2>, grid, block, 0, stream, keys, res, S, res_stride, g, t, pr, digit_bits, per_tile, offsets, (Rec<2>*)out, cap, counters); break;
    case 3: hipLaunchKernelGGL(k_site_scatter<3>, grid, block, 0, stream, keys, res, S, res_stride, g, t, pr, digit_bits, per_tile, offsets, (Rec<3>*)out, cap, counters); break;
    case 4: hipLaunchKernelGGL(k_site_scatter<4>, grid, block, 0, stream, keys, res, S, res_stride, g, t, pr, digit_bits, per_tile, offsets, (Rec<4>*)out, cap, counters); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int L, int T>
static hipError_t launch_me(int mode, dim3 grid, hipStream_t stream, const void* in,
                            const uint64_t* segs, uint64_t in_cap, uint32_t nseg, uint32_t per, const RowGeom& g,
                            const Primes& pr, uint32_t digit_bits, uint32_t* hist,
                            const uint64_t* offsets, const uint64_t* capw, void* out, uint64_t cap,
                            uint64_t final_mask, uint64_t* slice_sums, uint64_t nslices, uint64_t* counters) {
  const Rec<L>* i = (const Rec<L>*)in;
  Rec<L>* o = (Rec<L>*)out;
  switch (mode) {
    case kMergeCount:
      if (digit_bits <= kNarrowDigitBits)
        hipLaunchKernelGGL((k_merge_expand<L, kMergeCount, T, kNarrowDigitBits>), grid, dim3(kBlock), 0, stream, i, segs, in_cap, nseg, per, g, pr, digit_bits, hist, offsets, capw, o, cap, final_mask, slice_sums, nslices, counters);
      else
        hipLaunchKernelGGL((k_merge_expand<L, kMergeCount, T>), grid, dim3(kBlock), 0, stream, i, segs, in_cap, nseg, per, g, pr, digit_bits, hist, offsets, capw, o, cap, final_mask, slice_sums, nslices, counters);
      break;
    case kMergeScatter:
      if (digit_bits <= kNarrowDigitBits)
        hipLaunchKernelGGL((k_merge_expand<L, kMergeScatter, T, kNarrowDigitBits, kScatterBlock>), grid, dim3(kScatterBlock), 0, stream, i, segs, in_cap, nseg, per, g, pr, digit_bits, hist, offsets, capw, o, cap, final_mask, slice_sums, nslices, counters);
      else
        hipLaunchKernelGGL((k_merge_expand<L, kMergeScatter, T, kFineBits, kScatterBlock>), grid, dim3(kScatterBlock), 0, stream, i, segs, in_cap, nseg, per, g, pr, digit_bits, hist, offsets, capw, o, cap, final_mask, slice_sums, nslices, counters);
      break;
    case kMergeLast: hipLaunchKernelGGL((k_merge_expand<L, kMergeLast, T>), grid, dim3(kBlock), 0, stream, i, segs, in_cap, nseg, per, g, pr, digit_bits, hist, offsets, capw, o, cap, final_mask, slice_sums, nslices, counters); break;
    case kMergeEmit:
      if constexpr (L == kPackedRecords)
        return hipErrorInvalidValue;
      else
        hipLaunchKernelGGL((k_merge_expand<L, kMergeEmit, T>), grid, dim3(kBlock), 0, stream, i, segs, in_cap, nseg, per, g, pr, digit_bits, hist, offsets, capw, o, cap, final_mask, slice_sums, nslices, counters);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

uint32_t merge_cap(uint32_t table_slots) { return (table_slots * 5) / 8; }

uint32_t merge_resident_blocks(int lanes, uint32_t table_slots, uint32_t digit_bits) {
  int blocks = 0;
  const void* f = nullptr;
#define PEPS_KF(LN, TS)                                                                  \
  f = digit_bits <= kNarrowDigitBits ? (const void*)&k_merge_expand<LN, kMergeScatter, TS, kNarrowDigitBits, kScatterBlock> \
                                     : (const void*)&k_merge_expand<LN, kMergeScatter, TS, kFineBits, kScatterBlock>
  if (table_slots == 2048) {
    switch (lanes) {
      case kPackedRecords: PEPS_KF(kPackedRecords, 2048); break;
      case 1: PEPS_KF(1, 2048); break;
      case 2: PEPS_KF(2, 2048); break;
      case 3: PEPS_KF(3, 2048); break;
      default: PEPS_KF(4, 2048); break;
    }
  } else {
    switch (lanes) {
      case kPackedRecords: PEPS_KF(kPackedRecords, 4096); break;
      case 1: PEPS_KF(1, 4096); break;
      case 2: PEPS_KF(2, 4096); break;
      case 3: PEPS_KF(3, 4096); break;
      default: PEPS_KF(4, 4096); break;
    }
  }
#undef PEPS_KF
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, kScatterBlock, 0) != hipSuccess || blocks < 1) blocks = 1;
  return (uint32_t)blocks;
}

hipError_t launch_merge_expand(int mode, int lanes, const void* in, const Buckets& bk, uint32_t W,
                               const RowGeom& g, const Primes& pr, uint32_t digit_bits, uint32_t* hist,
                               const uint64_t* offsets, const uint64_t* capw, void* out, uint64_t cap,
                               uint64_t final_mask, uint64_t* slice_sums, uint64_t nslices, uint64_t* counters,
                               uint32_t table_slots, hipStream_t stream) {
  const uint32_t nseg = bk.nseg;
  if (nseg == 0 || W == 0 || digit_bits > kFineBits) return hipErrorInvalidValue;
  const uint32_t per = (nseg + W - 1) / W;
  dim3 grid(W);
#define PEPS_ME(LN, TS) return launch_me<LN, TS>(mode, grid, stream, in, bk.segs, bk.cap, nseg, per, g, pr, digit_bits, hist, offsets, capw, out, cap, final_mask, slice_sums, nslices, counters)
  if (table_slots == 2048) {
    switch (lanes) {
      case kPackedRecords: PEPS_ME(kPackedRecords, 2048);
      case 1: PEPS_ME(1, 2048);
      case 2: PEPS_ME(2, 2048);
      case 3: PEPS_ME(3, 2048);
      case 4: PEPS_ME(4, 2048);
      default: return hipErrorInvalidValue;
    }
  }
  if (table_slots == 4096) {
    switch (lanes) {
      case kPackedRecords: PEPS_ME(kPackedRecords, 4096);
      case 1: PEPS_ME(1, 4096);
      case 2: PEPS_ME(2, 4096);
      case 3: PEPS_ME(3, 4096);
      case 4: PEPS_ME(4, 4096);
      default: return hipErrorInvalidValue;
    }
  }
#undef PEPS_ME
  return hipErrorInvalidValue;
}

}  // namespace peps
