// kernels.h -- host-callable launchers for the gfx950 row-contraction kernels.
//
// One contracted row is:
//   count      : successors per parent (popcount of the legal positions)
//   expand     : parents -> successor records, hash-radix digit of level 1
//                (histogram pass + scan + LDS-staged scatter)
//   partition  : further hash-radix levels, one workgroup per segment
//   merge      : per final bucket, LDS hash table with modular adds of the
//                CRT residue lanes, compacted into the next frontier
// Keys are the reference's packed boundary (columns | diag_dr << n |
// diag_dl << 2n, src/lib.rs:357-363) in a u64, optionally with a slice id
// above bit 3n.  Residues are u32 values modulo the CRT primes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace peps {

constexpr int kMaxLanes = 4;
constexpr uint64_t kEmptyKey = ~0ull;

// Geometry of one row step (uniform across the launch).
struct RowGeom {
  uint32_t n;
  uint32_t orbit_row;    // 1: apply the row-0 vertical-reflection orbit weights
  uint64_t board_mask;   // (1 << n) - 1
  uint64_t state_mask;   // (1 << 3n) - 1; bits above are a slice id
  uint64_t orbit_mask;   // legal-position restriction (left half on the orbit row)
  uint64_t centre_bit;   // the odd-N fixed column on the orbit row (weight 1), else 0
  // Occupied C entry of the compiled operator (binary legs, src/lib.rs:425-485).
  uint32_t col_in, col_out, dr_in, dr_out, dl_in, dl_out;
};

struct Primes {
  uint32_t p[kMaxLanes];
};

// Device-side counters of one row (u64 each).
enum Counter : int {
  kCandidates = 0,  // successors emitted (completed row terms)
  kOutStates = 1,   // merged states written
  kWeight0 = 2,     // .. kWeight0 + L - 1: sum of residues of merged states (mod p per block)
  kFinal0 = 6,      // .. kFinal0 + L - 1: same restricted to columns == board mask
  kError = 10,      // nonzero: an LDS merge table overflowed
  kMatchedSlot = 11, // legal positions before the row-0 orbit filter
  kNumCounters = 12
};

// Record sizes for the AoS partition buffers: u64 key + L u32 residues.
size_t record_bytes(int lanes);

// All launchers run on `stream` and return hipSuccess or the launch error.
hipError_t launch_count(const uint64_t* keys, uint64_t S, const RowGeom& g, uint64_t* counters,
                        hipStream_t stream);

hipError_t launch_expand_hist(const uint64_t* keys, uint64_t S, const RowGeom& g, uint32_t digit_bits,
                              uint32_t tiles, uint64_t per_tile, uint32_t* hist, hipStream_t stream);

hipError_t launch_scan_tiles(const uint32_t* hist, uint32_t tiles, uint32_t digit_bits,
                             uint64_t* offsets, uint64_t* segs, uint64_t* counters,
                             hipStream_t stream);

hipError_t launch_expand_scatter(int lanes, const uint64_t* keys, const uint32_t* res,
                                 uint64_t S, uint64_t res_stride, const RowGeom& g, const Primes& pr,
                                 uint32_t digit_bits, uint32_t tiles, uint64_t per_tile,
                                 const uint64_t* offsets, void* out, hipStream_t stream);

hipError_t launch_partition(int lanes, const void* in, void* out, const uint64_t* segs,
                            uint32_t nseg, uint32_t shift, uint32_t digit_bits, uint64_t* subsegs,
                            hipStream_t stream);

hipError_t launch_merge(int lanes, const void* in, const uint64_t* segs, uint32_t nseg,
                        const Primes& pr, uint64_t* out_keys, uint32_t* out_res,
                        uint64_t out_stride, uint64_t final_mask, uint64_t* counters,
                        uint32_t table_slots, hipStream_t stream);

}  // namespace peps
