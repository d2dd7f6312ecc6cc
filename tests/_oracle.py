"""ctypes access to the CPU oracle (oracle/libpeps_oracle.so).

Test infrastructure only: the oracle is the checker, never the thing tested.
"""
import ctypes
import os

from peps_mi355x.abi import U128, ContractionResult, TensorEntry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "libpeps_oracle.so")

SORT_REDUCE, HASH, SITEWISE, PARALLEL = 0, 1, 2, 3

_lib = None
LAYER_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64,
                            ctypes.POINTER(ctypes.c_uint64))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        c = ctypes
        L.oracle_contract_rows.argtypes = [c.c_uint32] * 6 + [c.POINTER(ContractionResult), c.c_char_p, c.c_size_t]
        L.oracle_contract_prefix.argtypes = [c.c_uint32, c.POINTER(c.c_uint32), c.c_uint32, c.c_uint32,
                                             c.POINTER(ContractionResult), c.c_char_p, c.c_size_t]
        L.oracle_known_count.argtypes = [c.c_uint32, c.POINTER(U128)]
        L.oracle_crt_plan.argtypes = [c.c_uint32, c.POINTER(c.c_uint64), c.POINTER(c.c_uint32),
                                      c.POINTER(U128), c.POINTER(U128), c.c_char_p, c.c_size_t]
        L.oracle_crt_reconstruct.argtypes = [c.POINTER(c.c_uint64), c.POINTER(c.c_uint64), c.c_uint32,
                                             c.POINTER(U128), c.c_char_p, c.c_size_t]
        L.oracle_is_prime.argtypes = [c.c_uint64]
        L.oracle_prefix_tasks.argtypes = [c.c_uint32, c.c_uint32, c.POINTER(c.c_uint64), c.c_uint64,
                                          c.POINTER(c.c_uint64)]
        L.oracle_dfs.argtypes = [c.c_uint32, c.c_uint32, c.c_uint32, c.POINTER(U128), c.POINTER(U128),
                                 c.POINTER(U128), c.POINTER(c.c_uint32), c.POINTER(c.c_uint64),
                                 c.c_char_p, c.c_size_t]
        L.oracle_count_prefix_dfs.argtypes = [c.c_uint32, c.POINTER(c.c_uint32), c.c_uint32,
                                              c.POINTER(U128), c.c_char_p, c.c_size_t]
        L.oracle_site_tensors.argtypes = [c.POINTER(c.c_uint8), c.POINTER(c.c_uint8), c.POINTER(c.c_uint32)]
        L.oracle_row_transfer.argtypes = [c.c_uint32, c.c_uint64, c.c_uint64, c.c_uint64, c.c_uint32,
                                          c.POINTER(c.c_uint64), c.c_uint32, c.POINTER(c.c_uint32)]
        L.oracle_depth_counts.argtypes = [c.c_uint32, c.c_uint32, c.c_uint32, c.POINTER(U128), c.c_char_p,
                                          c.c_size_t]
        L.oracle_count_prefix_parallel.argtypes = [c.c_uint32, c.POINTER(c.c_uint32), c.c_uint32, c.c_uint32,
                                                   c.POINTER(U128), c.c_char_p, c.c_size_t]
        L.oracle_frontier_census.argtypes = [c.c_uint32, c.c_uint32, c.c_uint32, c.POINTER(c.c_uint64),
                                             c.c_char_p, c.c_size_t]
        L.oracle_peak_rss_bytes.restype = c.c_uint64
        L.oracle_contract_rows_tensor.argtypes = [c.c_uint32, c.c_uint32, c.c_uint32, c.POINTER(TensorEntry),
                                                  c.c_uint32, c.c_uint32, LAYER_FN, c.c_void_p,
                                                  c.POINTER(ContractionResult), c.c_char_p, c.c_size_t]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _err():
    return ctypes.create_string_buffer(512)


def contract_rows(n, symmetry=True, sparse=False, backend=SORT_REDUCE, threads=1,
                  linear_scan_counters=False) -> ContractionResult:
    out = ContractionResult()
    e = _err()
    rc = lib().oracle_contract_rows(n, int(symmetry), int(sparse), backend, threads,
                                    int(linear_scan_counters), ctypes.byref(out), e, 512)
    if rc:
        raise OracleError(e.value.decode())
    return out


def contract_prefix(n, cols, threads=1) -> ContractionResult:
    arr = (ctypes.c_uint32 * max(1, len(cols)))(*cols)
    out = ContractionResult()
    e = _err()
    rc = lib().oracle_contract_prefix(n, arr, len(cols), threads, ctypes.byref(out), e, 512)
    if rc:
        raise OracleError(e.value.decode())
    return out


def known_count(n):
    v = U128()
    return v.value if lib().oracle_known_count(n, ctypes.byref(v)) == 0 else None


def crt_plan(n):
    primes = (ctypes.c_uint64 * 4)()
    cnt = ctypes.c_uint32()
    prod, fact = U128(), U128()
    e = _err()
    rc = lib().oracle_crt_plan(n, primes, ctypes.byref(cnt), ctypes.byref(prod), ctypes.byref(fact), e, 512)
    if rc:
        raise OracleError(e.value.decode())
    return list(primes)[: cnt.value], prod.value, fact.value


def crt_reconstruct(residues, primes):
    r = (ctypes.c_uint64 * len(residues))(*residues)
    p = (ctypes.c_uint64 * len(primes))(*primes)
    out = U128()
    e = _err()
    rc = lib().oracle_crt_reconstruct(r, p, len(residues), ctypes.byref(out), e, 512)
    if rc:
        raise OracleError(e.value.decode())
    return out.value


def is_prime(v):
    return bool(lib().oracle_is_prime(v))


def prefix_tasks(n, depth):
    cnt = ctypes.c_uint64()
    lib().oracle_prefix_tasks(n, depth, None, 0, ctypes.byref(cnt))
    buf = (ctypes.c_uint64 * (4 * max(1, cnt.value)))()
    lib().oracle_prefix_tasks(n, depth, buf, cnt.value, ctypes.byref(cnt))
    return [tuple(buf[4 * i: 4 * i + 4]) for i in range(cnt.value)]


def dfs(n, threads=1, track=False):
    c, nodes, cands = U128(), U128(), U128()
    sd, nt = ctypes.c_uint32(), ctypes.c_uint64()
    e = _err()
    rc = lib().oracle_dfs(n, threads, int(track), ctypes.byref(c), ctypes.byref(nodes), ctypes.byref(cands),
                          ctypes.byref(sd), ctypes.byref(nt), e, 512)
    if rc:
        raise OracleError(e.value.decode())
    return {"count": c.value, "recursive_nodes": nodes.value, "candidate_placements": cands.value,
            "split_depth": sd.value, "tasks": nt.value}


def count_prefix_dfs(n, cols):
    arr = (ctypes.c_uint32 * max(1, len(cols)))(*cols)
    out = U128()
    e = _err()
    rc = lib().oracle_count_prefix_dfs(n, arr, len(cols), ctypes.byref(out), e, 512)
    if rc:
        raise OracleError(e.value.decode())
    return out.value


def count_prefix_parallel(n, cols, threads=8):
    """count_prefix_dfs on `threads` threads (layer_census.cpp)."""
    arr = (ctypes.c_uint32 * max(1, len(cols)))(*cols)
    out = U128()
    e = _err()
    rc = lib().oracle_count_prefix_parallel(n, arr, len(cols), threads, ctypes.byref(out), e, 512)
    if rc:
        raise OracleError(e.value.decode())
    return out.value


def site_tensors():
    b = (ctypes.c_uint8 * 170)()
    c = (ctypes.c_uint8 * 153)()
    buckets = (ctypes.c_uint32 * 16)()
    nc = lib().oracle_site_tensors(b, c, buckets)
    return ([tuple(b[10 * i: 10 * i + 10]) for i in range(17)],
            [tuple(c[9 * i: 9 * i + 9]) for i in range(nc)], list(buckets))


def row_transfer(n, cols, dr, dl, sitewise):
    cap = 64
    buf = (ctypes.c_uint64 * (4 * cap))()
    cnt = ctypes.c_uint32()
    rc = lib().oracle_row_transfer(n, cols, dr, dl, int(sitewise), buf, cap, ctypes.byref(cnt))
    if rc:
        raise OracleError("row transfer failed")
    return [tuple(buf[4 * i: 4 * i + 4]) for i in range(cnt.value)]


def contract_rows_tensor(n, entries=None, symmetry=False, backend=SITEWISE, linear_scan_counters=False,
                         layers=False):
    """Contraction with a caller-supplied C (sitewise backend) or Sec. VI's.

    Returns (result, {row: sorted [(columns, diag_dr, diag_dl, weight)]}) when
    layers is set, else the result."""
    arr = None
    if entries is not None:
        arr = (TensorEntry * max(1, len(entries)))(*entries)
    dump = {}

    def cb(_user, row, count, states):
        flat = states[: 5 * count]
        dump[row] = sorted((flat[5 * i], flat[5 * i + 1], flat[5 * i + 2],
                            flat[5 * i + 3] | (flat[5 * i + 4] << 64)) for i in range(count))

    fn = LAYER_FN(cb) if layers else LAYER_FN()
    out = ContractionResult()
    e = _err()
    rc = lib().oracle_contract_rows_tensor(n, int(symmetry), backend, arr, len(entries or ()),
                                           int(linear_scan_counters), fn, None, ctypes.byref(out), e, 512)
    if rc:
        raise OracleError(e.value.decode())
    return (out, dump) if layers else out


def depth_counts(n, orbits=True, threads=8):
    """[c_0 .. c_n]: c_r = the number of non-attacking r-row prefixes, each
    weighted by the row-0 orbit weight when `orbits` (layer_census.cpp)."""
    buf = (U128 * (n + 1))()
    e = _err()
    rc = lib().oracle_depth_counts(n, int(orbits), threads, buf, e, 512)
    if rc:
        raise OracleError(e.value.decode())
    return [buf[i].value for i in range(n + 1)]


CENSUS_KEYS = ("row", "input_states", "row_operator_matched", "completed_row_terms", "output_states")


def frontier_census(n, orbits=True, threads=8):
    """Per row: {row, input_states, row_operator_matched, completed_row_terms,
    output_states} of the default row loop, by the grouped breadth-first pass
    (layer_census.cpp)."""
    buf = (ctypes.c_uint64 * (5 * n))()
    e = _err()
    rc = lib().oracle_frontier_census(n, int(orbits), threads, buf, e, 512)
    if rc:
        raise OracleError(e.value.decode())
    return [dict(zip(CENSUS_KEYS, buf[5 * i: 5 * i + 5])) for i in range(n)]
