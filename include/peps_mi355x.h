/*
 * peps_mi355x.h -- C ABI of libpeps_mi355x.so, the MI355X (gfx950) explicit-C
 * PEPS row-boundary contraction that counts N-Queens solutions.
 *
 * Every entry point is plain C: pointers, sizes and POD structs, no C++ or
 * torch types.  Functions return 0 on success; on failure they return a
 * nonzero PEPS_E_* code and copy a NUL-terminated message into `err`
 * (truncated to `err_cap` bytes; `err` may be NULL).  The caller owns every
 * buffer.  Results are exact: coefficients are carried as residues modulo the
 * reference's 32-bit CRT primes and reconstructed on the host into a checked
 * 128-bit integer.
 *
 * Reference interface each entry point replaces (reference crate
 * `nqueens-peps-naive`, paths relative to its root):
 *   peps_contract_rows            <- pub fn contract_rows            src/lib.rs:757-760
 *                                    (= contract_rows_d4_orbit_sort_reduce src/lib.rs:784-786,
 *                                     body contract_rows_sort_reduce_with_modes src/lib.rs:5287-5415)
 *   peps_contract_rows_with_modes <- contract_rows_sort_reduce      src/lib.rs:770-772 (symmetry None)
 *                                    contract_rows_d4_orbit_sort_reduce src/lib.rs:784-786 (orbits)
 *   peps_known_count              <- pub fn known_count             src/lib.rs:5752-5784
 *   peps_peak_rss_bytes           <- pub fn peak_rss_bytes          src/lib.rs:5786-5850
 *   peps_crt_plan                 <- pub fn wide_crt_plan           src/lib.rs:3859-3886
 *   peps_crt_reconstruct          <- fn reconstruct_crt             src/lib.rs:4134-4163
 *   peps_prefix_tasks             <- fn build_wide_crt_tasks        src/lib.rs:3888-3956
 *   peps_count_prefix             <- (no single reference function) the row contraction
 *                                    started from the boundary left by k placed rows;
 *                                    pinned by the DFS comparator src/dfs_bitmask.rs:67-109
 *   peps_count_batched            <- sliced form of contract_rows over the prefix tasks of
 *                                    build_wide_crt_tasks src/lib.rs:3888-3956
 *   peps_server_*                 <- (no counterpart) serving loop over peps_count_prefix
 */
#ifndef PEPS_MI355X_H
#define PEPS_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Largest board the reference's packed u128 boundary accepts (src/lib.rs:5292-5294). */
#define PEPS_MAX_N 42
/* Largest board whose 3N-bit boundary key fits the device's u64 key. */
#define PEPS_MAX_DEVICE_N 21
/* Largest number of CRT lanes (reference CRT_PRIME_CANDIDATES, src/lib.rs:3767). */
#define PEPS_MAX_LANES 4

enum {
  PEPS_OK = 0,
  PEPS_E_INVALID = 1,      /* bad argument (N too large, bad options, ...) */
  PEPS_E_OVERFLOW = 2,     /* exactness bound violated (CRT > N!, u128 overflow) */
  PEPS_E_DEVICE = 3,       /* HIP runtime error or no device */
  PEPS_E_MEMORY = 4,       /* HBM budget exceeded */
  PEPS_E_UNKNOWN = 5,      /* peps_known_count: N outside the table */
  PEPS_E_BUSY = 6,         /* serving: the request queue is full */
  PEPS_E_PENDING = 7,      /* serving: the request has not finished yet */
  PEPS_E_INTERNAL = 8      /* an internal consistency check failed (LDS merge table
                              overflow, a write past a buffer's planned capacity, a
                              replayed size plan that no longer matches the device);
                              never retried as a memory shortage */
};

/* Rust u128 as two little-endian 64-bit halves. */
typedef struct {
  uint64_t lo;
  uint64_t hi;
} peps_u128;

/* Mirrors LayerMetric (src/lib.rs:388-401) plus device extras. */
typedef struct {
  uint32_t row; /* 0-based row index, as LayerMetric.row */
  uint64_t input_states;
  peps_u128 tensor_entries_examined;
  peps_u128 tensor_entries_matched;
  peps_u128 row_operator_candidates;
  peps_u128 row_operator_matched;
  peps_u128 completed_row_terms;
  uint64_t output_states;
  peps_u128 output_weight;
  double elapsed_s;
  uint64_t peak_rss_bytes;
  /* device extras (0 on host-only paths) */
  uint64_t device_bytes;      /* device arena bytes live during this row */
  uint32_t partition_levels;  /* hash-radix partition levels used for the merge */
  uint32_t final_buckets;     /* LDS merge buckets used for the merge */
} peps_layer_metric;

/* Mirrors ContractionResult (src/lib.rs:403-415) plus device extras. */
typedef struct {
  uint32_t n;
  peps_u128 count;
  double elapsed_s;
  uint64_t peak_states;
  peps_u128 tensor_entries_examined;
  peps_u128 tensor_entries_matched;
  peps_u128 row_operator_candidates;
  peps_u128 row_operator_matched;
  uint64_t peak_rss_bytes;
  uint32_t layer_count;
  peps_layer_metric layers[PEPS_MAX_N];
  /* device extras */
  uint32_t crt_lanes;
  uint32_t record_bytes; /* candidate record size the solve used (12 packed, 16 or 24 lanes) */
  uint64_t primes[PEPS_MAX_LANES];
  uint64_t residues[PEPS_MAX_LANES];
  uint64_t device_peak_bytes;
  double kernel_s; /* device time between the first and last kernel of the solve */
  double alloc_s;  /* host time spent growing the device arena during the call */
} peps_contraction_result;

/* Symmetry mode (reference SymmetryMode, src/lib.rs:715-719). */
enum { PEPS_SYMMETRY_NONE = 0, PEPS_SYMMETRY_TOP_ROW_ORBITS = 1 };

/* Tuning and resource options.  Zero-initialised means "defaults". */
typedef struct {
  uint32_t crt_lanes;        /* 0: reference plan (shortest prime prefix with M > N!); else 1..4 */
  uint32_t radix_bits;       /* hash-radix digit width per partition level (0: auto; 1..11) */
  uint32_t tile_states;      /* target records per LDS merge bucket (0: auto; 256..2560) */
  uint32_t device;           /* HIP device ordinal */
  uint64_t hbm_budget_bytes; /* arena cap (0: auto, 85% of free HBM) */
  uint32_t collect_layers;   /* nonzero: fill per-layer metrics (needs a host sync per row) */
  uint32_t expand_bits;      /* max digit width the merge-expand scatter writes (0: auto; 1..10) */
  uint32_t graph_mode;       /* nonzero: record the run's size plan; repeats of the same
                                call replay the whole row pipeline as one HIP graph */
  uint32_t slice_group;      /* batched form: prefix tasks per device group (0: as many as
                                the slice-id bits above the 3N-bit key allow) */
  uint32_t merge_passes;     /* LDS merge passes per final bucket (0: auto; 1..8) */
  uint32_t merge_waves;      /* merge grid in waves of resident workgroups (0: 8; 1..8) */
  uint32_t site_counters;    /* sitewise tensor counters: 0 indexed (src/lib.rs:537-540),
                                1 every C entry per partial-row visit (benchmarks/n14_layers.csv) */
  uint32_t record_format;    /* low byte, PEPS_RECORDS_*: candidate record format of the fused
                                row pipeline; bits 8..15, if nonzero, narrow the packed weight
                                field to that many bits (exercises the overflow path); bits
                                16..19, PEPS_FORCE_* << 16, make the pipeline take its fallback
                                paths (tests; results never change) */
} peps_options;

/* Candidate records of the fused row pipeline.  AUTO: packed 12-byte records
 * (the 3N-bit key and slice id, then the exact integer weight in the
 * remaining bits, at most 52), repeated with LANES records if a weight
 * outgrows its field.  LANES: the key and one u32 residue per CRT prime (16
 * bytes for 1-2 primes, 24 for 3-4).  PACKED: packed only; a weight overflow
 * fails with PEPS_E_OVERFLOW.  Counts are identical in every format. */
enum { PEPS_RECORDS_AUTO = 0, PEPS_RECORDS_LANES = 1, PEPS_RECORDS_PACKED = 2 };

/* Fallback paths of the fused row pipeline (record_format bits 16..19).
 * The partition writes fixed-capacity buckets by claims and the SCATTER
 * writes capped runs sized from fanout sums; these bits force the exact
 * histogram levels, upper-bound-count runs, or capacities too small to hold
 * (so every such step overflows and is repeated by its fallback). */
enum {
  PEPS_FORCE_EXACT_LEVELS = 1,
  PEPS_FORCE_UPPER_RUNS = 2,
  PEPS_FORCE_LEVEL_OVERFLOW = 4,
  PEPS_FORCE_CAP_OVERFLOW = 8
};

/* ---------------------------------------------------------------- solve -- */

/* Reference contract_rows(n): first-row vertical-reflection orbit slicing, one
 * merged frontier, exact count.  N = 0 returns count 1 (src/lib.rs:5295-5307).
 * N > 42 fails with the reference text; 21 < N <= 42 fails with the device's
 * u64 key limit. */
int peps_contract_rows(uint32_t n, const peps_options* opts, peps_contraction_result* out,
                       char* err, size_t err_cap);

/* Same with an explicit symmetry mode (PEPS_SYMMETRY_*). */
int peps_contract_rows_with_modes(uint32_t n, uint32_t symmetry, const peps_options* opts,
                                  peps_contraction_result* out, char* err, size_t err_cap);

/* Count the completions of the k-row prefix cols[0..k) (queen of row i in
 * column cols[i]).  An attacking prefix counts 0.  k = 0 is peps_contract_rows. */
int peps_count_prefix(uint32_t n, const uint32_t* cols, uint32_t k, const peps_options* opts,
                      peps_u128* count, char* err, size_t err_cap);

/* The same count without allocating.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code that prepares HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * device buffer is a fixed slice of one static device array of 256 MiB per
 * GPU (PEPS_NO_ALLOC_SCRATCH_BYTES), part of the library's code object and
 * held from the library's load on that GPU until the process ends; kernels
 * run on the device's null stream and nothing is allocated on the device or
 * as pinned host memory (host bookkeeping uses the C++ heap).  A board whose
 * rows need more than the scratch holds (N = 14 fits; N = 15 and up only as
 * prefixes with small frontiers) fails with PEPS_E_MEMORY.  Calls on one GPU
 * share its array, which keeps the last call's contents: each call first
 * clears its counters with a kernel, and the library runs such calls one at
 * a time per GPU (a lock), so calls that would overlap wait.  graph_mode is
 * rejected; merge_waves is fixed at 1.  Same arguments, results and errors
 * as peps_count_prefix otherwise. */
#define PEPS_NO_ALLOC_SCRATCH_BYTES (256ull << 20)
int peps_count_prefix_no_alloc(uint32_t n, const uint32_t* cols, uint32_t k, const peps_options* opts,
                               peps_u128* count, char* err, size_t err_cap);

/* Batched form: all prefix tasks of depth `split_depth` (reference
 * build_wide_crt_tasks semantics: orbit weight on row 0) contracted as
 * independent slices (slice id above the 3N-bit key) in groups of
 * opts->slice_group tasks, each group one device pipeline (one HIP graph
 * launch when opts->graph_mode replays a recorded plan).  A group that does
 * not fit the HBM budget is split in two and retried.  `out` receives the
 * summed count and per-layer totals over the slices; `tasks_out` (may be
 * NULL) the number of tasks; task_counts[0..min(ntasks, task_cap)) (may be
 * NULL) each task's weighted count (orbit weight x completions). */
int peps_count_batched(uint32_t n, uint32_t split_depth, const peps_options* opts,
                       peps_contraction_result* out, uint32_t* tasks_out, peps_u128* task_counts,
                       uint64_t task_cap, char* err, size_t err_cap);

/* Progress of long calls (the batched form reports after every task group):
 * fn(user, tasks_done, tasks_total).  NULL disables.  Process-wide. */
typedef void (*peps_progress_fn)(void* user, uint64_t done, uint64_t total);
int peps_set_progress(peps_progress_fn fn, void* user);

/* ------------------------------------------------------------ utilities -- */

/* A000170 table Q(0..27); PEPS_E_UNKNOWN outside it. */
int peps_known_count(uint32_t n, peps_u128* out);

/* Peak resident set of this process in bytes (Linux getrusage). */
uint64_t peps_peak_rss_bytes(void);

/* CRT plan: the shortest prefix of the reference primes whose product exceeds
 * N!.  Writes up to 4 primes and their count; returns PEPS_E_INVALID for N > 34. */
int peps_crt_plan(uint32_t n, uint64_t* primes, uint32_t* count, peps_u128* modulus_product,
                  char* err, size_t err_cap);

/* Garner/Fermat reconstruction of residues[0..count) into a checked u128. */
int peps_crt_reconstruct(const uint64_t* residues, const uint64_t* primes, uint32_t count,
                         peps_u128* out, char* err, size_t err_cap);

/* Prefix tasks of depth d: writes up to `cap` tasks as (columns, diag_dr,
 * diag_dl, orbit_weight) quadruples into tasks[4*i..4*i+4) and the count into
 * *count (which may exceed cap; then only cap are written). */
int peps_prefix_tasks(uint32_t n, uint32_t depth, uint64_t* tasks, uint64_t cap, uint64_t* count,
                      char* err, size_t err_cap);

/* Device description: name (NUL-terminated, up to name_cap), CU count, HBM
 * bytes, LDS bytes per workgroup.  PEPS_E_DEVICE without a GPU. */
int peps_device_info(uint32_t device, char* name, size_t name_cap, uint32_t* cu_count,
                     uint64_t* hbm_bytes, uint32_t* lds_bytes, char* err, size_t err_cap);

/* Library version string. */
const char* peps_version(void);

/* ------------------------------------------------------------- profiling -- */

/* Per-kernel timing gathered with HIP events on the engine's stream. */
typedef struct {
  char name[32];
  uint64_t launches;
  double seconds;           /* summed event-to-event device time */
  double algorithmic_bytes; /* compulsory HBM bytes of those launches */
} peps_kernel_stat;

/* Enable (nonzero) or disable kernel timing on `device`; either resets the stats. */
int peps_profile_kernels(uint32_t device, uint32_t enable, char* err, size_t err_cap);

/* Copy up to `cap` kernel stats; *count receives how many kernels have stats. */
int peps_kernel_stats(uint32_t device, peps_kernel_stat* out, uint32_t cap, uint32_t* count,
                      char* err, size_t err_cap);

/* ------------------------------------------------------ tensor fidelity -- */

/* One entry of the explicit site tensor C: CEntry (src/lib.rs:88-92), legs
 * in the VirtualLegs order (src/lib.rs:19-36), value as u64. */
typedef struct {
  uint8_t column_in, column_out, row_in, row_out;
  uint8_t diag_dr_in, diag_dr_out, diag_dl_in, diag_dl_out;
  uint64_t value;
} peps_tensor_entry;

enum { PEPS_ROW_OPERATOR_COMPILED = 0, PEPS_ROW_OPERATOR_SITEWISE = 1 };

/* SiteTensorC::sec_vi() (src/lib.rs:102-139): the 17 entries in the
 * reference's legs_key order.  Writes up to cap entries; *count = 17. */
int peps_site_tensor_c(peps_tensor_entry* out, uint32_t cap, uint32_t* count);

/* CompiledRowOperator::compile (src/lib.rs:425-485): fails closed
 * (PEPS_E_INVALID with the reference's message) unless C is exactly the 16
 * identity pass-throughs plus one unit all-channel 0 -> 1 entry; on success
 * *occupied is that entry. */
int peps_compile_row_operator(const peps_tensor_entry* c, uint32_t count, peps_tensor_entry* occupied,
                              char* err, size_t err_cap);

/* contract_rows_sitewise (src/lib.rs:5629-5631 -> contract_rows_with_backend,
 * :5633-5750): every row applies the explicit C tensor literally, site by
 * site, to a frontier held in HBM (no row-0 orbit slicing).  Layer tensor
 * counters as the reference's sitewise backend (opts->site_counters selects
 * the pre-index revision's count); row_operator_* are 0. */
int peps_contract_rows_sitewise(uint32_t n, const peps_options* opts, peps_contraction_result* out,
                                char* err, size_t err_cap);

/* The same literal contraction with a caller-supplied C (up to 64 entries,
 * binary legs).  Counts are exact modulo the product of the CRT primes
 * (opts->crt_lanes, default the plan for N); no N! bound is applied, so a
 * C other than Sec. VI's may be contracted. */
int peps_contract_rows_with_tensor(uint32_t n, const peps_tensor_entry* c, uint32_t count,
                                   const peps_options* opts, peps_contraction_result* out, char* err,
                                   size_t err_cap);

/* One row of a layer trace: the merged boundary layer after `row`. */
typedef struct {
  uint32_t row;           /* 0-based, as LayerMetric.row */
  uint64_t support;       /* states in the layer */
  peps_u128 weight_sum;   /* exact sum of the weights */
  peps_u128 transitions;  /* successor terms that formed the layer (completed_row_terms) */
  uint8_t sha256[32];     /* SHA-256 of the canonical layer */
} peps_layer_digest;

/* Layer trace certificate: runs the materialised contraction and hashes each
 * merged layer in canonical form -- states sorted by (columns, diag_dr,
 * diag_dl, weight), each serialised as 40 little-endian bytes: columns u64,
 * diag_dr u64, diag_dl u64, weight u128 -- as proposed in
 * docs/sec_vi_row_automaton_equivalence.md:196-217.  row_operator:
 * PEPS_ROW_OPERATOR_COMPILED (C must compile; fails closed otherwise) or
 * PEPS_ROW_OPERATOR_SITEWISE (literal C).  c == NULL: Sec. VI's C.
 * symmetry: PEPS_SYMMETRY_* (orbit slicing only with the compiled operator).
 * Writes up to cap digests; *rows = N. */
int peps_trace_layers(uint32_t n, uint32_t row_operator, uint32_t symmetry, const peps_tensor_entry* c,
                      uint32_t c_count, const peps_options* opts, peps_layer_digest* out, uint32_t cap,
                      uint32_t* rows, char* err, size_t err_cap);

/* SHA-256 (FIPS 180-4) of a byte string, as used by the layer trace. */
int peps_sha256(const void* data, size_t len, uint8_t out[32]);

/* --------------------------------------------------------------- serving -- */

/* A request server: callers submit (N, prefix) counting requests from any
 * thread; one worker thread per server turns each request into prefix tasks
 * (rows 0..D placed, D = max(k, min(2, N-1)); an empty prefix carries the
 * row-0 orbit weights of build_wide_crt_tasks, src/lib.rs:3888-3956) and runs
 * them as the slices of engine launches.  Each launch takes tasks of one
 * (N, D) class from every active request of that class (continuous batching:
 * a request that arrives while a launch runs joins the next launch of its
 * class), admitted while the projected candidate bytes fit 55% of one HBM
 * candidate buffer.  Frontiers stay in HBM; a launch over budget is split.
 * The count of a request is the sum of its tasks' slice counts, so it is the
 * same whether it ran alone or in a mix.  There is no reference counterpart:
 * the reference computes one count per process (src/main.rs:25-70). */
typedef struct peps_server peps_server;

/* Which request's class the worker serves next: least projected remaining
 * work, oldest, or highest response ratio (wait + projected service time) /
 * projected service time -- short requests first, long ones aged in. */
enum { PEPS_POLICY_SHORTEST_FIRST = 0, PEPS_POLICY_FIFO = 1, PEPS_POLICY_RESPONSE_RATIO = 2 };

typedef struct {
  uint64_t ticket;
  uint32_t n;
  uint32_t prefix_len;
  int32_t status;        /* PEPS_OK, PEPS_E_PENDING, or the failure code */
  uint32_t tasks;        /* prefix tasks the request was split into */
  uint32_t launches;     /* engine launches that carried its tasks */
  uint32_t max_batch;    /* most requests that shared one launch with it (itself included) */
  peps_u128 count;       /* completions of the prefix (Q(N) for an empty prefix) */
  double submit_s;       /* seconds since the server was created */
  double start_s;        /* first launch carrying its tasks began */
  double finish_s;       /* result available */
  char error[160];
} peps_request_result;

typedef struct {
  uint64_t submitted, completed, failed, pending;
  uint64_t launches, tasks;
  uint64_t retries;      /* launches that exceeded the HBM budget and were re-run in halves */
  double busy_s;         /* worker time inside engine launches */
  double uptime_s;
} peps_server_stats;

/* opts: engine options as for peps_contract_rows (crt_lanes applies to every
 * N, so leave it 0 for mixed sizes).  policy: which class the worker serves
 * next (shortest projected remaining work first, or oldest request first).
 * max_pending: submit fails with PEPS_E_BUSY beyond this many unfinished
 * requests (0: no limit).  PEPS_E_DEVICE without a GPU. */
int peps_server_create(const peps_options* opts, uint32_t policy, uint32_t max_pending,
                       peps_server** out, char* err, size_t err_cap);

/* Queue the count of completions of prefix cols[0..k) on an N x N board.
 * Invalid input fails here (same texts as peps_count_prefix). */
int peps_server_submit(peps_server* s, uint32_t n, const uint32_t* cols, uint32_t k,
                       uint64_t* ticket, char* err, size_t err_cap);

/* PEPS_E_PENDING until the request is done, then its status; fills *out. */
int peps_server_poll(peps_server* s, uint64_t ticket, peps_request_result* out);

/* Wait up to timeout_s (negative: forever) for the request; as poll. */
int peps_server_wait(peps_server* s, uint64_t ticket, double timeout_s, peps_request_result* out);

int peps_server_stats_get(peps_server* s, peps_server_stats* out);

/* Finishes every queued request, stops the worker and frees the server. */
void peps_server_destroy(peps_server* s);

#ifdef __cplusplus
}
#endif

#endif /* PEPS_MI355X_H */
