"""Autotuner: shape selection and rejection on the CPU (with a stand-in runner
and clock), and a real sweep on the GPU whose every shape must give the known
count."""
import pytest

import peps_mi355x as pm
from peps_mi355x import autotune as at


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


def test_grid_enumerates_every_shape():
    g = {"radix_bits": (4, 8, 11), "tile_states": (512, 1024)}
    s = list(at.shapes(g))
    assert len(s) == 6 and {"radix_bits": 11, "tile_states": 512} in s


def test_fastest_correct_shape_wins(monkeypatch):
    clock = FakeClock()
    monkeypatch.setattr(at.time, "perf_counter", clock)
    cost = {4: 3.0, 8: 1.0, 11: 0.5}

    def run(opts):
        clock.t += cost[opts.radix_bits]
        # the fastest shape returns a wrong count and must lose
        return 92 + (1 if opts.radix_bits == 11 else 0)

    best, table = at.autotune(8, grid={"radix_bits": (4, 8, 11)}, run=run, save=False)
    assert best.radix_bits == 8
    assert [t for _, t in table] == [3.0, 1.0, None]


def test_failing_shapes_are_skipped_and_all_wrong_raises(monkeypatch):
    def run(opts):
        if opts.tile_states == 512:
            raise pm.PepsError(4, "HBM budget exceeded")
        return 92

    best, table = at.autotune(8, grid={"tile_states": (512, 1024)}, run=run, save=False)
    assert best.tile_states == 1024 and "error" in table[0][0]
    with pytest.raises(pm.PepsError, match="no shape"):
        at.autotune(8, grid={"tile_states": (512,)}, run=lambda o: 91, save=False)


def test_cache_round_trip(monkeypatch, tmp_path):
    monkeypatch.setenv("PEPS_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(at, "_key", lambda n, form, depth: f"dev|{n}|{form}|{depth}")
    at.autotune(8, grid={"radix_bits": (4, 8)}, run=lambda o: 92)
    got = at.tuned_options(8)
    assert got is not None and got.radix_bits in (4, 8)
    assert at.tuned_options(9) is None


@pytest.mark.gpu
def test_gpu_sweep_all_shapes_exact(monkeypatch, tmp_path):
    monkeypatch.setenv("PEPS_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    grid = {"radix_bits": (4, 11), "tile_states": (256, 2048), "merge_waves": (1, 4)}
    best, table = at.autotune(13, grid=grid, repeats=1)
    assert all(t is not None for _, t in table), table
    assert pm.contract_rows(13, best).count.value == 73712
    assert at.tuned_options(13).tile_states == best.tile_states
    bb, tb = at.autotune(13, form="batched", depth=2, grid={"slice_group": (0, 4)}, repeats=1)
    assert all(t is not None for _, t in tb)
    b, _ = pm.count_batched(13, 2, bb)
    assert b.count.value == 73712


def test_cache_key_follows_the_library_build(tmp_path, monkeypatch):
    # A rebuilt library must not reuse shapes cached for an older build.
    lib = tmp_path / "libpeps_mi355x.so"
    monkeypatch.setattr(pm.api, "lib_path", lambda: str(lib))
    lib.write_bytes(b"build one")
    a = at.build_id()
    lib.write_bytes(b"build two")
    b = at.build_id()
    lib.write_bytes(b"build one")
    assert a != b and at.build_id() == a and len(a) == 16
