// kernels.h -- host-callable launchers for the gfx950 row-contraction kernels.
//
// Keys are the reference's packed boundary (columns | diag_dr << n |
// diag_dl << 2n, src/lib.rs:357-363) in a u64, optionally with a slice id
// above bit 3n.  A candidate record is either the key with one u32 residue
// per CRT prime ("lane records", 16 or 24 bytes) or a packed 12-byte record
// holding the key in its low key_bits and the exact integer weight above it
// (lanes == kPackedRecords).  Packed weights must stay below 2^weight_bits;
// a kernel that meets a larger one sets bit 2 of kError and the run is
// repeated with lane records.  Either way the layer statistics and final
// sums are reduced to residues modulo every CRT prime.
//
// Row pipeline (candidates of row r live in hash-radix buckets):
//   expand (seed rows only): frontier -> candidates, level-1 buckets
//   partition             : one more hash-radix level inside each bucket,
//                           into fixed-capacity buckets filled by per-stage
//                           claims (no histogram pass); the last level also
//                           sums every record's fanout per merge workgroup
//                           (an exact histogram + scan level is the fallback)
//   caps (large rows)     : per merge workgroup, a cap per digit on its
//                           SCATTER run from its fanout sum
//   merge_expand COUNT    : (smaller rows) per bucket, LDS-merge keys, count
//                           successors of the merged states per next-row
//                           level-1 digit, then a digit-major scan
//   merge_expand SCATTER  : merge with weights, layer statistics, write the
//                           successors into the next row's level-1 runs
//                           (capped runs have gaps, which the first
//                           partition level reads in run mode; a run over
//                           its cap makes the engine count exact upper
//                           bounds, count_ub, and repeat the SCATTER)
//   merge_expand LAST     : last row: merge, statistics, final v1/v2 sum
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace peps {

constexpr int kMaxLanes = 4;
constexpr uint64_t kEmptyKey = ~0ull;
constexpr uint32_t kFineBits = 10;       // histogram resolution of the COUNT pass
constexpr uint32_t kMaxDigitBits = 11;   // widest partition digit
constexpr int kNarrowDigitBits = 8;      // SCATTER digits up to this width use smaller LDS arrays
constexpr int kMidDigitBits = 9;         // partition digits of 9 bits get their own LDS sizing

// Geometry of one row step (uniform across the launch).  Board masks are
// 32-bit (N <= 21).  The occupied C entry's binary legs (src/lib.rs:425-485)
// are stored as masks: x_in = board mask if the incoming leg is 1 else 0, so
// "bit matches the leg" is ~(bit ^ x_in); x_out likewise for the outgoing leg.
struct RowGeom {
  uint32_t n;
  uint32_t orbit_row;    // 1: apply the row-0 vertical-reflection orbit weights
  uint32_t board_mask;   // (1 << n) - 1
  uint32_t orbit_mask;   // legal-position restriction (left half on the orbit row)
  uint32_t centre_bit;   // the odd-N fixed column on the orbit row (weight 1), else 0
  uint32_t col_in, col_out, dr_in, dr_out, dl_in, dl_out;  // leg masks
  uint32_t key_bits;     // packed records: 3n + slice-id bits (<= 63)
  uint32_t weight_bits;  // packed records: weights are < 2^weight_bits (<= 52)
  uint64_t state_mask;   // (1 << 3n) - 1; bits above are a slice id
};

struct Primes {
  uint32_t p[kMaxLanes];
  uint32_t lanes;        // CRT lanes of the statistics (1..kMaxLanes)
};

// `lanes` value of the launchers that selects packed 12-byte records.
constexpr int kPackedRecords = 0;
constexpr uint32_t kPackedMaxWeightBits = 52;  // bucket sums stay below 2^64

// Device-side counters of one row (u64 each).
enum Counter : int {
  kCandidates = 0,   // successors emitted (completed row terms of the next row)
  kOutStates = 1,    // merged states of this row
  kWeight0 = 2,      // .. kWeight0 + L - 1: residue sums of merged states (mod p per block)
  kFinal0 = 6,       // .. kFinal0 + L - 1: same restricted to columns == board mask
  kError = 10,       // bit 0: LDS merge table overflow; bit 1: capacity; bit 2: packed
                     // weight overflow; bit 3: sitewise partial rows
  kMatchedSlot = 11, // legal positions before the row-0 orbit filter
  kSiteVisits = 12,  // sitewise: partial-row visits (one per partial per site)
  kSiteMatched = 13, // sitewise: C entries applied (= entries examined through the input index)
  kEmitted = 14,     // EMIT merge: merged states written to the frontier
  kWritten = 15,     // SCATTER merge: successors written (the next row's exact candidates)
  kFanout = 16,      // last partition level: sum of every record's legal positions
  kCapExtent = 17,   // capped SCATTER: records its runs span (digits x summed caps)
  kRetry = 18,       // bit 4: a capped SCATTER run overflowed; bit 5: a fixed-capacity
                     // bucket overflowed (the engine repeats that step with exact offsets)
  kNumCounters = 19
};

// COUNT / SCATTER expand the merged states into the next row; LAST only sums;
// EMIT writes the merged states out as the next frontier (materialised mode).
enum MergeMode : int { kMergeCount = 0, kMergeScatter = 1, kMergeLast = 2, kMergeEmit = 3 };

// The explicit C tensor as the literal sitewise kernels read it.  Entries are
// grouped by input index col | row << 1 | dr << 2 | dl << 3 (the reference's
// input_key, src/lib.rs:157-162); entry e's output legs are packed the same
// way in out[e] and its value is held modulo every CRT prime.
constexpr int kSiteMaxEntries = 64;
constexpr int kSitePartials = 24;   // partial rows per parent (N + 1 for the Sec. VI tensor)
struct SiteTable {
  uint8_t first[16];
  uint8_t count[16];
  uint8_t out[kSiteMaxEntries];
  uint32_t value[kSiteMaxEntries][kMaxLanes];
};

size_t record_bytes(int lanes);

hipError_t launch_count(const uint64_t* keys, uint64_t S, const RowGeom& g, uint64_t* counters,
                        hipStream_t stream);

// Seed-row expansion: histogram ([D][tiles], digit-major) then scatter.
hipError_t launch_expand_hist(const uint64_t* keys, uint64_t S, const RowGeom& g, uint32_t digit_bits,
                              uint32_t tiles, uint64_t per_tile, uint32_t* hist, hipStream_t stream);
// `cap` is the output capacity in records: a write at or beyond it is skipped
// and sets bit 1 of counters[kError] (only a stale replayed size plan can
// cause it).  Packed records (lanes == kPackedRecords) read the seed weights
// as exact u64 values at (const uint64_t*)res (res_stride unused).
hipError_t launch_expand_scatter(int lanes, const uint64_t* keys, const uint32_t* res,
                                 uint64_t S, uint64_t res_stride, const RowGeom& g, const Primes& pr,
                                 uint32_t digit_bits, uint32_t tiles, uint64_t per_tile,
                                 const uint64_t* offsets, void* out, uint64_t cap,
                                 uint64_t* counters, hipStream_t stream);

// Literal sitewise row transfer (contract_one_row_sitewise, src/lib.rs:513-592)
// of a materialised frontier (keys[S], residues res[l * res_stride + i]):
// COUNT adds successors, partial-row visits and applied entries to counters
// (kCandidates, kSiteVisits, kSiteMatched; bit 3 of kError if a parent needs
// more than kSitePartials partial rows); HIST and SCATTER have the shapes of
// launch_expand_hist / launch_expand_scatter.
hipError_t launch_site_count(const uint64_t* keys, uint64_t S, const RowGeom& g, const SiteTable& t,
                             uint64_t* counters, hipStream_t stream);
hipError_t launch_site_hist(const uint64_t* keys, uint64_t S, const RowGeom& g, const SiteTable& t,
                            uint32_t digit_bits, uint32_t tiles, uint64_t per_tile, uint32_t* hist,
                            hipStream_t stream);
hipError_t launch_site_scatter(int lanes, const uint64_t* keys, const uint32_t* res, uint64_t S,
                               uint64_t res_stride, const RowGeom& g, const SiteTable& t, const Primes& pr,
                               uint32_t digit_bits, uint32_t tiles, uint64_t per_tile,
                               const uint64_t* offsets, void* out, uint64_t cap, uint64_t* counters,
                               hipStream_t stream);

// p[0..n) = v (a kernel, so it can be captured into a HIP graph).
hipError_t launch_fill(uint64_t* p, uint64_t n, uint64_t v, hipStream_t stream);

// Exclusive scan of a digit-major histogram hist[Dfine][W] coarsened by
// `group` fine digits per output digit: offsets[d*W + w] (u64), segs[0..D]
// the digit boundaries (segs[D] = total), counters[kCandidates] = total.
// `partial` needs ceil(D*W / 4096) + 1 u64 of scratch.
hipError_t launch_scan(const uint32_t* hist, uint32_t D, uint32_t W, uint32_t group,
                       uint64_t* offsets, uint64_t* segs, uint64_t* partial, uint64_t* counters,
                       hipStream_t stream);

// The buckets of one level of a row: contiguous (cap == 0: segs holds
// nseg + 1 bounds) or fixed capacity (cap > 0: bucket b is the records
// [b * cap, b * cap + min(segs[b], cap)), segs[b] its fill cursor).
struct Buckets {
  const uint64_t* segs;
  uint64_t cap;
  uint32_t nseg;
};

// One more hash-radix digit (bits [shift, shift + digit_bits) of the key mix)
// inside each of the nseg input buckets, written to `out`.  Every bucket is
// processed as `chunks` ranges (partition_chunks).
//  * Exact (out_cap == 0): launch_part_hist counts every (bucket, digit,
//    chunk) into hist, launch_scan(hist, nseg << digit_bits, chunks, 1,
//    offsets, subsegs, partial, counters) turns the counts into write
//    offsets and the next level's contiguous bounds, and launch_part_scatter
//    writes at those offsets.  hist and offsets need nseg * 2^digit_bits *
//    chunks entries, partial that / 4096 + 2.
//  * Fixed capacity (out_cap > 0): no histogram pass.  Output bucket
//    (s, d) is [(s * D + d) * out_cap, ... + out_cap) and `offsets` holds
//    its fill cursor (nseg * 2^digit_bits entries, zeroed by the caller);
//    each LDS stage claims its digit runs with one atomic add per digit.  A
//    bucket that would pass out_cap sets bit 5 of counters[kError] and
//    writes nothing beyond it (the engine repeats the level exactly).
// A write at or beyond `cap` (the output region's records) sets bit 1.
// Target records per partition work item, the fastest of a sweep at N = 18:
// contiguous segments (a row's last level) take 48 Ki records (24-96 Ki
// measured); levels that read SCATTER runs take 1 Mi (64 Ki - 1 Mi
// measured), since longer items append longer stretches to each output
// bucket (at least one item per kMaxChunkRuns runs keeps small rows spread
// over the GPU).
constexpr uint64_t kPartChunk = 49152;
constexpr uint64_t kPartChunkRuns = 1048576;
constexpr uint64_t kMaxPartChunks = 4096;
uint32_t partition_chunks(uint64_t records, uint64_t nseg, bool runs = false);
// Run mode of a partition level (its input is the SCATTER output with
// capped or upper-bound offsets): segment s is the runs w < per_seg, run
// (s, w) being the records [start[s * per_seg + w], start[...] +
// len[s * per_seg + w]).  Chunk p takes the runs [per_seg * p / chunks,
// per_seg * (p + 1) / chunks), at most kMaxChunkRuns of them, so
// chunks >= ceil(per_seg / kMaxChunkRuns).
constexpr uint32_t kMaxChunkRuns = 64;
struct Runs {
  const uint64_t* start;
  const uint32_t* len;
  uint32_t per_seg;
};
// Fanout sums folded into the last (fixed-capacity) partition level of a
// row: every record's number of legal positions in the next row (the
// successors it yields if it is a merged state; duplicates count once per
// record, so the sums bound the SCATTER's writes from above) is added to
// U[w] of the merge workgroup w = (s * 2^digit_bits + d) / per that will own
// its final bucket (a merge launch of W workgroups, workgroup w owning the
// final buckets [w * per, (w + 1) * per)), and to *total.  U is zeroed by
// the caller.
struct FanSum {
  uint64_t* U;
  uint64_t* total;
  uint32_t W, per;
  RowGeom g;
};
// Most merge workgroups whose buckets one segment's 2^digit_bits final
// buckets can touch.
inline uint32_t ub_span(uint32_t digit_bits, uint32_t per) { return ((1u << digit_bits) - 1) / per + 2; }
// key_bits: the packed records' key width (ignored for lane records).
// runs: run mode (bk.segs unused), or null.
hipError_t launch_part_hist(int lanes, const void* in, const Buckets& bk, uint32_t chunks, uint32_t shift,
                            uint32_t digit_bits, uint32_t key_bits, const Runs* runs, uint32_t* hist,
                            hipStream_t stream);
hipError_t launch_part_scatter(int lanes, const void* in, void* out, const Buckets& bk, uint32_t chunks,
                               uint32_t shift, uint32_t digit_bits, uint32_t key_bits, const Runs* runs,
                               uint64_t* offsets, uint64_t out_cap, uint64_t cap, uint64_t* counters,
                               const FanSum* fan, hipStream_t stream);

// Upper-bound COUNT: successor digits of every record of the buckets (no
// merge), per merge workgroup of a W-workgroup merge launch, hist[d * W + w].
hipError_t launch_count_ub(int lanes, const void* in, const Buckets& bk, uint32_t W, const RowGeom& g,
                           uint32_t digit_bits, uint32_t* hist, hipStream_t stream);

// Capped SCATTER runs from the fanout sums U[W] (FanSum): workgroup w may
// write cap_w = ceil(U_w / D) + 8 sqrt(U_w / D) + 32 successors per digit
// (D = 2^digit_bits) into the run starting at run_start[d * W + w] =
// d * T + (cap_0 + .. + cap_{w-1}), T = sum of the caps.  capw needs W and
// pre W + 1 entries, run_start D * W; counters[kCapExtent] = D * T.
// shrink > 0 divides the caps by 2^shrink (tests: forces the overflow path).
hipError_t launch_caps(const uint64_t* U, uint32_t W, uint32_t digit_bits, uint32_t shrink, uint64_t* capw,
                       uint64_t* pre, uint64_t* run_start, uint64_t* counters, hipStream_t stream);

// Fused merge of row r's buckets with the expansion into row r+1.
// COUNT: hist[(1<<kFineBits)][W] of successor digits.  SCATTER: successors
// written to out (capacity `cap` records) at offsets[(1<<digit_bits)][W]; if
// hist is given, the number written per (digit, workgroup) goes there (the
// run lengths when the offsets are capped or upper bounds).  capw (or null):
// workgroup w should write at most capw[w] successors per digit.  A longer
// run sets bit 4 of counters[kRetry] when the kernel ends; its surplus has
// been written into the ranges of later runs (never past `cap`), so the
// caller must repeat the SCATTER with other offsets, which rewrites every
// run the next level reads.
// LAST: no expansion.  EMIT (lane records only): the merged states are
// written to `out` as a frontier (keys[cap] then residues[l * cap + i]),
// positions counted in counters[kEmitted].  Merged states with (key & final_mask) == final_mask
// (all columns set) are summed into kFinal0 and, if slice_sums is given,
// into slice_sums[slice * L + l] (slice = key bits above 3N).
// table_slots selects the LDS table (2048 or 4096 slots); a bucket of up to
// merge_cap(table_slots) records is merged in one pass.
hipError_t launch_merge_expand(int mode, int lanes, const void* in, const Buckets& bk, uint32_t W,
                               const RowGeom& g, const Primes& pr, uint32_t digit_bits, uint32_t* hist,
                               const uint64_t* offsets, const uint64_t* capw, void* out, uint64_t cap,
                               uint64_t final_mask, uint64_t* slice_sums, uint64_t nslices, uint64_t* counters,
                               uint32_t table_slots, hipStream_t stream);
uint32_t merge_cap(uint32_t table_slots);
// Workgroups of the SCATTER merge kernel (writing digits of up to
// `digit_bits`) that fit on one CU at once.
uint32_t merge_resident_blocks(int lanes, uint32_t table_slots, uint32_t digit_bits);

}  // namespace peps
