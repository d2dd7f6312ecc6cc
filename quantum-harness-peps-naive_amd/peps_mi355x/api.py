"""ctypes bindings of libpeps_mi355x.so with the reference's function names.

Reference functions mirrored (crate nqueens-peps-naive, src/lib.rs):
  contract_rows                       src/lib.rs:757-760
  contract_rows_sort_reduce           src/lib.rs:770-772  (no orbit slicing)
  contract_rows_d4_orbit_sort_reduce  src/lib.rs:784-786
  known_count                         src/lib.rs:5752-5784
  peak_rss_bytes                      src/lib.rs:5786-5850
  wide_crt_plan / reconstruct_crt     src/lib.rs:3859-3886, 4134-4163
plus the prefix count, the batched (prefix-sliced) form and a request server
(Server), which the reference does not have.
Errors surface as PepsError carrying the library's message (the reference
returns Err(String) with the same texts).
"""
import ctypes
import dataclasses
import os

from .abi import (
    PEPS_E_INVALID,
    PEPS_E_PENDING,
    PEPS_POLICY_FIFO,
    PEPS_POLICY_RESPONSE_RATIO,
    PEPS_POLICY_SHORTEST_FIRST,
    PEPS_SYMMETRY_NONE,
    PEPS_SYMMETRY_TOP_ROW_ORBITS,
    PROGRESS_FN,
    U128,
    ContractionResult,
    PEPS_ROW_OPERATOR_COMPILED,
    PEPS_ROW_OPERATOR_SITEWISE,
    KernelStat,
    LayerDigest,
    Options,
    RequestResult,
    ServerStats,
    TensorEntry,
)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libpeps_mi355x.so"
_lib = None


class PepsError(RuntimeError):
    """A contraction or argument error reported by the library."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


@dataclasses.dataclass
class PepsOptions:
    crt_lanes: int = 0          # 0: reference CRT plan
    radix_bits: int = 0         # 0: default hash-radix digit width
    tile_states: int = 0        # 0: default LDS merge bucket size
    device: int = 0
    hbm_budget_bytes: int = 0
    expand_bits: int = 0        # 0: default merge-expand scatter fan-out
    graph_mode: int = 0         # 1: record size plans, replay repeats as one HIP graph
    slice_group: int = 0        # batched form: tasks per device group (0: max)
    merge_passes: int = 0       # LDS merge passes per final bucket (0: default)
    merge_waves: int = 0        # merge grid in waves of resident workgroups (0: default)
    site_counters: int = 0      # sitewise tensor counters: 0 indexed, 1 linear scan of C
    record_format: int = 0      # PEPS_RECORDS_*: 0 packed (lane fallback), 1 lane records, 2 packed only
    packed_weight_bits: int = 0  # nonzero: narrower packed weight field (exercises the fallback)
    force_paths: int = 0         # FORCE_* bits: take the pipeline's fallback paths (tests; same results)

    def to_c(self) -> Options:
        o = Options()
        o.crt_lanes = self.crt_lanes
        o.radix_bits = self.radix_bits
        o.tile_states = self.tile_states
        o.device = self.device
        o.hbm_budget_bytes = self.hbm_budget_bytes
        o.collect_layers = 1
        o.expand_bits = self.expand_bits
        o.graph_mode = self.graph_mode
        o.slice_group = self.slice_group
        o.merge_passes = self.merge_passes
        o.merge_waves = self.merge_waves
        o.site_counters = self.site_counters
        o.record_format = self.record_format | (self.packed_weight_bits << 8) | (self.force_paths << 16)
        return o


def lib_path() -> str:
    return os.environ.get("PEPS_MI355X_LIB", os.path.join(_HERE, _LIB_NAME))


def load():
    """Load the HIP library; raises PepsError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise PepsError(3, f"libpeps_mi355x.so is not built at {path}; run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    c = ctypes
    P = c.POINTER
    L.peps_contract_rows.argtypes = [c.c_uint32, P(Options), P(ContractionResult), c.c_char_p, c.c_size_t]
    L.peps_contract_rows_with_modes.argtypes = [c.c_uint32, c.c_uint32, P(Options), P(ContractionResult),
                                                c.c_char_p, c.c_size_t]
    L.peps_count_prefix.argtypes = [c.c_uint32, P(c.c_uint32), c.c_uint32, P(Options), P(U128),
                                    c.c_char_p, c.c_size_t]
    L.peps_count_prefix_no_alloc.argtypes = [c.c_uint32, P(c.c_uint32), c.c_uint32, P(Options), P(U128),
                                             c.c_char_p, c.c_size_t]
    L.peps_count_batched.argtypes = [c.c_uint32, c.c_uint32, P(Options), P(ContractionResult),
                                     P(c.c_uint32), P(U128), c.c_uint64, c.c_char_p, c.c_size_t]
    L.peps_known_count.argtypes = [c.c_uint32, P(U128)]
    L.peps_peak_rss_bytes.restype = c.c_uint64
    L.peps_crt_plan.argtypes = [c.c_uint32, P(c.c_uint64), P(c.c_uint32), P(U128), c.c_char_p, c.c_size_t]
    L.peps_crt_reconstruct.argtypes = [P(c.c_uint64), P(c.c_uint64), c.c_uint32, P(U128), c.c_char_p,
                                       c.c_size_t]
    L.peps_prefix_tasks.argtypes = [c.c_uint32, c.c_uint32, P(c.c_uint64), c.c_uint64, P(c.c_uint64),
                                    c.c_char_p, c.c_size_t]
    L.peps_device_info.argtypes = [c.c_uint32, c.c_char_p, c.c_size_t, P(c.c_uint32), P(c.c_uint64),
                                   P(c.c_uint32), c.c_char_p, c.c_size_t]
    L.peps_version.restype = c.c_char_p
    L.peps_profile_kernels.argtypes = [c.c_uint32, c.c_uint32, c.c_char_p, c.c_size_t]
    L.peps_kernel_stats.argtypes = [c.c_uint32, P(KernelStat), c.c_uint32, P(c.c_uint32), c.c_char_p,
                                    c.c_size_t]
    L.peps_set_progress.argtypes = [PROGRESS_FN, c.c_void_p]
    TE = TensorEntry
    L.peps_site_tensor_c.argtypes = [P(TE), c.c_uint32, P(c.c_uint32)]
    L.peps_compile_row_operator.argtypes = [P(TE), c.c_uint32, P(TE), c.c_char_p, c.c_size_t]
    L.peps_contract_rows_sitewise.argtypes = [c.c_uint32, P(Options), P(ContractionResult), c.c_char_p,
                                              c.c_size_t]
    L.peps_contract_rows_with_tensor.argtypes = [c.c_uint32, P(TE), c.c_uint32, P(Options),
                                                 P(ContractionResult), c.c_char_p, c.c_size_t]
    L.peps_trace_layers.argtypes = [c.c_uint32, c.c_uint32, c.c_uint32, P(TE), c.c_uint32, P(Options),
                                    P(LayerDigest), c.c_uint32, P(c.c_uint32), c.c_char_p, c.c_size_t]
    L.peps_sha256.argtypes = [c.c_char_p, c.c_size_t, P(c.c_uint8)]
    S = c.c_void_p
    L.peps_server_create.argtypes = [P(Options), c.c_uint32, c.c_uint32, P(S), c.c_char_p, c.c_size_t]
    L.peps_server_submit.argtypes = [S, c.c_uint32, P(c.c_uint32), c.c_uint32, P(c.c_uint64), c.c_char_p,
                                     c.c_size_t]
    L.peps_server_poll.argtypes = [S, c.c_uint64, P(RequestResult)]
    L.peps_server_wait.argtypes = [S, c.c_uint64, c.c_double, P(RequestResult)]
    L.peps_server_stats_get.argtypes = [S, P(ServerStats)]
    L.peps_server_destroy.argtypes = [S]
    L.peps_server_destroy.restype = None
    _lib = L
    return L


def _err():
    return ctypes.create_string_buffer(1024)


def _check(rc, e):
    if rc != 0:
        raise PepsError(rc, e.value.decode(errors="replace"))


def _opts(options):
    o = (options or PepsOptions()).to_c()
    return ctypes.byref(o), o


def contract_rows_with_modes(n: int, symmetry: int, options: PepsOptions = None) -> ContractionResult:
    out = ContractionResult()
    e = _err()
    ref, _keep = _opts(options)
    _check(load().peps_contract_rows_with_modes(n, symmetry, ref, ctypes.byref(out), e, len(e)), e)
    return out


def contract_rows(n: int, options: PepsOptions = None) -> ContractionResult:
    """Reference contract_rows(n): orbit-sliced exact row contraction."""
    return contract_rows_with_modes(n, PEPS_SYMMETRY_TOP_ROW_ORBITS, options)


def contract_rows_d4_orbit_sort_reduce(n: int, options: PepsOptions = None) -> ContractionResult:
    return contract_rows_with_modes(n, PEPS_SYMMETRY_TOP_ROW_ORBITS, options)


def contract_rows_sort_reduce(n: int, options: PepsOptions = None) -> ContractionResult:
    """Reference contract_rows_sort_reduce(n): no orbit slicing."""
    return contract_rows_with_modes(n, PEPS_SYMMETRY_NONE, options)


def count_prefix(n: int, cols, options: PepsOptions = None) -> int:
    cols = list(cols)
    arr = (ctypes.c_uint32 * max(1, len(cols)))(*cols)
    out = U128()
    e = _err()
    ref, _keep = _opts(options)
    _check(load().peps_count_prefix(n, arr, len(cols), ref, ctypes.byref(out), e, len(e)), e)
    return out.value


def count_prefix_no_alloc(n: int, cols, options: PepsOptions = None) -> int:
    """count_prefix through peps_count_prefix_no_alloc: no device allocation,
    every buffer in the library's 256 MiB static scratch (PEPS_E_MEMORY for
    boards whose rows do not fit it)."""
    cols = list(cols)
    arr = (ctypes.c_uint32 * max(1, len(cols)))(*cols)
    out = U128()
    e = _err()
    ref, _keep = _opts(options)
    _check(load().peps_count_prefix_no_alloc(n, arr, len(cols), ref, ctypes.byref(out), e, len(e)), e)
    return out.value


def count_batched(n: int, split_depth: int = 2, options: PepsOptions = None, task_counts: bool = False):
    """Prefix-sliced form.  Returns (result, ntasks) or, with task_counts,
    (result, ntasks, [weighted count of each task])."""
    out = ContractionResult()
    tasks = ctypes.c_uint32()
    e = _err()
    ref, _keep = _opts(options)
    cap = 4096 if task_counts else 0
    buf = (U128 * max(1, cap))()
    _check(load().peps_count_batched(n, split_depth, ref, ctypes.byref(out), ctypes.byref(tasks),
                                     buf if task_counts else None, cap, e, len(e)), e)
    if task_counts:
        return out, tasks.value, [buf[i].value for i in range(min(cap, tasks.value))]
    return out, tasks.value


def known_count(n: int):
    v = U128()
    return v.value if load().peps_known_count(n, ctypes.byref(v)) == 0 else None


def peak_rss_bytes() -> int:
    return int(load().peps_peak_rss_bytes())


def crt_plan(n: int):
    primes = (ctypes.c_uint64 * 4)()
    cnt = ctypes.c_uint32()
    prod = U128()
    e = _err()
    _check(load().peps_crt_plan(n, primes, ctypes.byref(cnt), ctypes.byref(prod), e, len(e)), e)
    return list(primes)[: cnt.value], prod.value


def crt_reconstruct(residues, primes) -> int:
    r = (ctypes.c_uint64 * max(1, len(residues)))(*residues)
    p = (ctypes.c_uint64 * max(1, len(primes)))(*primes)
    out = U128()
    e = _err()
    _check(load().peps_crt_reconstruct(r, p, len(residues), ctypes.byref(out), e, len(e)), e)
    return out.value


def prefix_tasks(n: int, depth: int):
    cnt = ctypes.c_uint64()
    e = _err()
    _check(load().peps_prefix_tasks(n, depth, None, 0, ctypes.byref(cnt), e, len(e)), e)
    buf = (ctypes.c_uint64 * (4 * max(1, cnt.value)))()
    _check(load().peps_prefix_tasks(n, depth, buf, cnt.value, ctypes.byref(cnt), e, len(e)), e)
    return [tuple(buf[4 * i: 4 * i + 4]) for i in range(cnt.value)]


def device_info(device: int = 0) -> dict:
    name = ctypes.create_string_buffer(256)
    cu, hbm, lds = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint32()
    e = _err()
    _check(load().peps_device_info(device, name, 256, ctypes.byref(cu), ctypes.byref(hbm), ctypes.byref(lds),
                                   e, len(e)), e)
    return {"name": name.value.decode(), "cu_count": cu.value, "hbm_bytes": hbm.value, "lds_bytes": lds.value}


def version() -> str:
    return load().peps_version().decode()


def profile_kernels(enable: bool, device: int = 0) -> None:
    """Turn HIP-event kernel timing on or off (either resets the stats)."""
    e = _err()
    _check(load().peps_profile_kernels(device, int(enable), e, len(e)), e)


def kernel_stats(device: int = 0) -> dict:
    """{kernel: {launches, seconds, algorithmic_bytes}} since the last reset."""
    cap = 32
    buf = (KernelStat * cap)()
    cnt = ctypes.c_uint32()
    e = _err()
    _check(load().peps_kernel_stats(device, buf, cap, ctypes.byref(cnt), e, len(e)), e)
    return {buf[i].name.decode(): {"launches": buf[i].launches, "seconds": buf[i].seconds,
                                   "algorithmic_bytes": buf[i].algorithmic_bytes}
            for i in range(min(cap, cnt.value))}


@dataclasses.dataclass
class RequestOutcome:
    """One finished (or pending) server request."""
    ticket: int
    n: int
    prefix_len: int
    status: int
    count: int
    tasks: int
    launches: int
    max_batch: int
    submit_s: float
    start_s: float
    finish_s: float
    error: str

    @property
    def latency_s(self) -> float:
        return self.finish_s - self.submit_s

    @staticmethod
    def of(r: "RequestResult") -> "RequestOutcome":
        return RequestOutcome(r.ticket, r.n, r.prefix_len, r.status, r.count.value, r.tasks, r.launches,
                              r.max_batch, r.submit_s, r.start_s, r.finish_s,
                              r.error.decode(errors="replace"))


class Server:
    """Request server over the engine (peps_server_* in the C ABI).

    ``submit(n, cols)`` queues the count of completions of the prefix ``cols``
    (rows 0..k-1) and returns a ticket; ``wait(ticket)`` returns its
    ``RequestOutcome``.  Requests of one N share launches (continuous
    batching); a result does not depend on what else was in flight.
    """

    def __init__(self, options: PepsOptions = None, policy: str = "response_ratio", max_pending: int = 0):
        pol = {"shortest": PEPS_POLICY_SHORTEST_FIRST, "fifo": PEPS_POLICY_FIFO,
               "response_ratio": PEPS_POLICY_RESPONSE_RATIO}[policy]
        self._h = ctypes.c_void_p()
        e = _err()
        ref, _keep = _opts(options)
        _check(load().peps_server_create(ref, pol, max_pending, ctypes.byref(self._h), e, len(e)), e)

    def submit(self, n: int, cols=()) -> int:
        cols = list(cols)
        arr = (ctypes.c_uint32 * max(1, len(cols)))(*cols)
        t = ctypes.c_uint64()
        e = _err()
        _check(load().peps_server_submit(self._h, n, arr, len(cols), ctypes.byref(t), e, len(e)), e)
        return t.value

    def poll(self, ticket: int):
        r = RequestResult()
        rc = load().peps_server_poll(self._h, ticket, ctypes.byref(r))
        if rc == PEPS_E_INVALID and r.ticket != ticket:
            raise PepsError(rc, f"unknown ticket {ticket}")
        return None if rc == PEPS_E_PENDING else RequestOutcome.of(r)

    def wait(self, ticket: int, timeout: float = None) -> RequestOutcome:
        r = RequestResult()
        rc = load().peps_server_wait(self._h, ticket, -1.0 if timeout is None else float(timeout),
                                     ctypes.byref(r))
        if rc == PEPS_E_PENDING:
            raise TimeoutError(f"request {ticket} not finished after {timeout} s")
        if r.ticket != ticket:
            raise PepsError(rc, f"unknown ticket {ticket}")
        return RequestOutcome.of(r)

    def stats(self) -> dict:
        s = ServerStats()
        load().peps_server_stats_get(self._h, ctypes.byref(s))
        return {name: getattr(s, name) for name, _ in ServerStats._fields_}

    def close(self) -> None:
        if self._h:
            load().peps_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ tensor fidelity --

def site_tensor_c():
    """SiteTensorC::sec_vi() (src/lib.rs:102-139): the 17 C entries, legs_key order."""
    arr = (TensorEntry * 64)()
    cnt = ctypes.c_uint32()
    load().peps_site_tensor_c(arr, 64, ctypes.byref(cnt))
    return [TensorEntry.from_buffer_copy(arr[i]) for i in range(cnt.value)]


def _tensor_array(entries):
    entries = list(entries)
    return (TensorEntry * max(1, len(entries)))(*entries), len(entries)


def compile_row_operator(entries) -> TensorEntry:
    """CompiledRowOperator::compile (src/lib.rs:425-485): the occupied entry, or
    PepsError with the reference's message when C is not exactly Sec. VI's."""
    arr, cnt = _tensor_array(entries)
    occ = TensorEntry()
    e = _err()
    _check(load().peps_compile_row_operator(arr, cnt, ctypes.byref(occ), e, len(e)), e)
    return occ


def contract_rows_sitewise(n: int, options: PepsOptions = None) -> ContractionResult:
    """Reference contract_rows_sitewise(n) (src/lib.rs:5629): the literal
    sitewise C contraction, run on the GPU over HBM-resident frontiers."""
    out = ContractionResult()
    e = _err()
    ref, _keep = _opts(options)
    _check(load().peps_contract_rows_sitewise(n, ref, ctypes.byref(out), e, len(e)), e)
    return out


def contract_rows_with_tensor(n: int, entries, options: PepsOptions = None) -> ContractionResult:
    """The literal sitewise contraction of a caller-supplied C (no N! bound)."""
    arr, cnt = _tensor_array(entries)
    out = ContractionResult()
    e = _err()
    ref, _keep = _opts(options)
    _check(load().peps_contract_rows_with_tensor(n, arr, cnt, ref, ctypes.byref(out), e, len(e)), e)
    return out


def contract_rows_hash_materialization(n: int, options: PepsOptions = None) -> ContractionResult:
    """Reference contract_rows_hash_materialization(n) (src/lib.rs:764-766): the
    compiled operator without orbit slicing; the merge structure does not change
    any reported field, so this is contract_rows_sort_reduce."""
    return contract_rows_sort_reduce(n, options)


def trace_layers(n: int, row_operator: str = "compiled", symmetry: bool = False, tensor=None,
                 options: PepsOptions = None):
    """Layer trace certificate: per row, the merged layer's support, exact
    weight sum, transitions and the SHA-256 of its canonical serialisation
    (states sorted by (columns, diag_dr, diag_dl, weight), 40 little-endian
    bytes each)."""
    op = {"compiled": PEPS_ROW_OPERATOR_COMPILED, "sitewise": PEPS_ROW_OPERATOR_SITEWISE}[row_operator]
    arr, cnt = (None, 0) if tensor is None else _tensor_array(tensor)
    out = (LayerDigest * 64)()
    rows = ctypes.c_uint32()
    e = _err()
    ref, _keep = _opts(options)
    sym = PEPS_SYMMETRY_TOP_ROW_ORBITS if symmetry else PEPS_SYMMETRY_NONE
    _check(load().peps_trace_layers(n, op, sym, arr, cnt, ref, out, 64, ctypes.byref(rows), e, len(e)), e)
    return [{"row": d.row, "support": d.support, "weight_sum": d.weight_sum.value,
             "transitions": d.transitions.value, "sha256": bytes(d.sha256).hex()}
            for d in out[: rows.value]]


def sha256(data: bytes) -> str:
    out = (ctypes.c_uint8 * 32)()
    load().peps_sha256(data, len(data), out)
    return bytes(out).hex()


def canonical_layer_bytes(states) -> bytes:
    """The trace's serialisation of a layer given as (columns, diag_dr, diag_dl, weight) tuples."""
    import struct
    return b"".join(struct.pack("<QQQQQ", c, dr, dl, w & ((1 << 64) - 1), w >> 64)
                    for c, dr, dl, w in sorted(states))


_progress_keep = None


def set_progress(callback) -> None:
    """callback(done, total) after every task group of the batched form (None: off)."""
    global _progress_keep
    fn = PROGRESS_FN(lambda _u, done, total: callback(done, total)) if callback else PROGRESS_FN()
    _progress_keep = fn
    load().peps_set_progress(fn, None)
