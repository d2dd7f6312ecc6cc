"""Autotuner for the contraction's launch shape.

Sweeps the engine's tunables -- hash-radix digit width per partition level
(``radix_bits``), final bucket size (``tile_states``; up to 1280 records use a
2048-slot LDS table, more a 4096-slot table), the digit width the merge-expand
scatter writes (``expand_bits``), the merge grid (``merge_waves``) and, for
the batched form, the slice grouping (``slice_group``) -- times each shape on
the real workload, rejects any shape whose count differs from the known
count, and caches the fastest per (device, library version and build, N, form,
depth).

    best, table = autotune(18)                 # plain form
    best, table = autotune(18, form="batched")
    opts = tuned_options(18)                   # cached best, or None
"""
import itertools
import json
import os
import time

from . import api

DEFAULT_GRID = {
    "radix_bits": (4, 8, 11),
    "tile_states": (512, 1024, 2048),
    "expand_bits": (5, 6, 7),
    "merge_waves": (1, 4),
}
BATCHED_GRID = {"slice_group": (0, 8, 32)}


def cache_path() -> str:
    return os.environ.get("PEPS_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "peps_mi355x", "autotune.json"))


def build_id() -> str:
    """SHA-256 prefix of the loaded library's bytes: a rebuilt library (new
    kernels, new fusion rules) never reuses a shape cached for an older one."""
    import hashlib

    h = hashlib.sha256()
    with open(api.lib_path(), "rb") as f:
        for block in iter(lambda: f.read(1 << 20), b""):
            h.update(block)
    return h.hexdigest()[:16]


def _key(n, form, depth):
    info = api.device_info(0)
    return (f"{info['name']}|{api.version()}|build={build_id()}|n={n}|{form}"
            f"|depth={depth if form == 'batched' else 0}")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def shapes(grid):
    names = sorted(grid)
    for values in itertools.product(*(grid[k] for k in names)):
        yield dict(zip(names, values))


def _runner(n, form, depth):
    def run(opts):
        if form == "plain":
            r = api.contract_rows(n, opts)
        else:
            r, _ = api.count_batched(n, depth, opts)
        return r.count.value
    return run


def autotune(n, form="plain", depth=2, grid=None, repeats=2, budget_s=None, run=None, save=True,
             progress=None):
    """Time every shape of `grid` (default: DEFAULT_GRID, plus BATCHED_GRID for
    the batched form); returns (best PepsOptions, [(shape, seconds)]).

    `run(options) -> count` may replace the engine call (tests).  A shape whose
    count differs from the known count, or that fails, is never chosen.
    `budget_s` stops the sweep once exceeded (the shapes timed so far count)."""
    if grid is None:
        grid = dict(DEFAULT_GRID)
        if form == "batched":
            grid.update(BATCHED_GRID)
    run = run or _runner(n, form, depth)
    want = api.known_count(n)
    table = []
    t_start = time.perf_counter()
    for shape in shapes(grid):
        if budget_s is not None and table and time.perf_counter() - t_start > budget_s:
            break
        opts = api.PepsOptions(**shape)
        try:
            first = run(opts)  # warm: allocation, size plan
            best = None
            for _ in range(repeats):
                t0 = time.perf_counter()
                c = run(opts)
                dt = time.perf_counter() - t0
                if c != first:
                    raise api.PepsError(2, "count changed between repeats")
                best = dt if best is None else min(best, dt)
            ok = want is None or first == want
        except api.PepsError as e:
            ok, best = False, None
            shape = dict(shape, error=str(e))
        table.append((shape, best if ok else None))
        if progress:
            progress(shape, best if ok else None)
    timed = [(s, t) for s, t in table if t is not None]
    if not timed:
        raise api.PepsError(1, "autotune: no shape produced the known count")
    best_shape = min(timed, key=lambda st: st[1])[0]
    if save:
        path = cache_path()
        data = _load(path)
        data[_key(n, form, depth)] = {"shape": best_shape, "seconds": min(t for _, t in timed)}
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    return api.PepsOptions(**best_shape), table


def tuned_options(n, form="plain", depth=2):
    """The cached best shape for this device and library, or None."""
    entry = _load(cache_path()).get(_key(n, form, depth))
    return api.PepsOptions(**entry["shape"]) if entry else None
