"""GPU parity: libpeps_mi355x.so (HIP, gfx950) against the CPU oracle and the
reference's published counters.  Integer work: every comparison is exact."""
import json
import os
import random

import pytest

import _oracle as o
import peps_mi355x as pm

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
KEYS = ("input_states", "row_operator_candidates", "row_operator_matched", "completed_row_terms",
        "output_states", "output_weight")


def tuples(res):
    return [tuple(l[k] for k in KEYS) for l in res.layer_list()]


def show(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("symmetry", [True, False])
def test_layers_match_oracle_through_n12(symmetry):
    for n in range(1, 13):
        show("layers", n, symmetry)
        g = pm.contract_rows(n) if symmetry else pm.contract_rows_sort_reduce(n)
        r = o.contract_rows(n, symmetry=symmetry)
        assert g.count.value == r.count.value == o.known_count(n), n
        assert tuples(g) == tuples(r), n
        assert g.peak_states == r.peak_states
        assert g.row_operator_candidates.value == r.row_operator_candidates.value
        assert g.row_operator_matched.value == r.row_operator_matched.value
        assert g.tensor_entries_examined.value == 17 and g.tensor_entries_matched.value == 17


def test_published_totals_n13_to_n16():
    for row in FIX["e12_totals"]:
        if row["mode"] not in ("orbit", "dense"):
            continue
        show("e12", row["mode"], row["n"])
        g = pm.contract_rows(row["n"]) if row["mode"] == "orbit" else pm.contract_rows_sort_reduce(row["n"])
        assert (g.count.value, g.peak_states, g.row_operator_candidates.value,
                g.row_operator_matched.value) == (row["count"], row["peak_states"],
                                                  row["row_operator_candidates"],
                                                  row["row_operator_matched"]), row["src"]
    for row in FIX["orbit_n16"]:
        g = pm.contract_rows(16)
        assert g.count.value == row["count"] and g.peak_states == row["peak_states"], row["src"]
        assert g.row_operator_matched.value == row["row_operator_matched"], row["src"]


def test_orbit_layer_supports_match_published_audit():
    for n_str, layers in FIX["orbit_layer_supports"].items():
        g = pm.contract_rows(int(n_str))
        got = [1] + [l["output_states"] for l in g.layer_list()]
        want = [l["reachable_states"] for l in sorted(layers, key=lambda x: x["row"])]
        assert got == want, n_str


def test_n14_n15_layers_match_oracle():
    for n in (14, 15):
        show("layers", n)
        g = pm.contract_rows(n)
        r = o.contract_rows(n, backend=o.PARALLEL, threads=8)
        assert tuples(g) == tuples(r)


def test_known_counts_n17():
    g = pm.contract_rows(17)
    assert g.count.value == o.known_count(17) == 95815104
    assert g.residues[0] == 95815104 % 4294967291


@pytest.mark.parametrize("lanes", [2, 3, 4])
def test_lane_counts_agree(lanes):
    base = pm.contract_rows(14)
    g = pm.contract_rows(14, pm.PepsOptions(crt_lanes=lanes))
    assert g.crt_lanes == lanes
    assert tuples(g) == tuples(base) and g.count.value == base.count.value
    for l in range(lanes):
        assert g.residues[l] == 365596 % g.primes[l]


def test_lanes_below_plan_rejected():
    with pytest.raises(pm.PepsError, match="certified plan"):
        pm.contract_rows(14, pm.PepsOptions(crt_lanes=1))


@pytest.mark.parametrize("radix_bits,tile", [(4, 256), (10, 2560), (2, 512), (7, 1024)])
def test_partition_shapes_agree(radix_bits, tile):
    # Small tiles and digits force several hash-radix levels per row.
    base = o.contract_rows(13, backend=o.PARALLEL, threads=8)
    g = pm.contract_rows(13, pm.PepsOptions(radix_bits=radix_bits, tile_states=tile))
    assert tuples(g) == tuples(base) and g.count.value == base.count.value
    assert max(l["partition_levels"] for l in g.layer_list()) >= 1


def test_prefix_counts_match_dfs():
    rng = random.Random(11)
    for n in range(4, 13):
        for k in (1, 2, 3):
            for _ in range(3):
                cols = [rng.randrange(n) for _ in range(k)]
                show("prefix", n, cols)
                assert pm.count_prefix(n, cols) == o.count_prefix_dfs(n, cols), (n, cols)


def test_prefix_sums_reconstruct_q():
    for n in (10, 13):
        assert sum(pm.count_prefix(n, [c]) for c in range(n)) == o.known_count(n)
        assert pm.count_prefix(n, []) == o.known_count(n)


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_batched_matches_plain(depth):
    for n in (8, 11, 14):
        show("batched", n, depth)
        b, tasks = pm.count_batched(n, depth)
        assert tasks == len(o.prefix_tasks(n, depth))
        assert b.count.value == o.known_count(n)
        plain = pm.contract_rows(n)
        # per-row total weight counts prefixes, so it is slicing-invariant
        pw = [l["output_weight"] for l in plain.layer_list()][depth:]
        bw = [l["output_weight"] for l in b.layer_list()]
        assert bw == pw


def test_batched_n16_two_row_tasks():
    b, tasks = pm.count_batched(16, 2)
    assert b.count.value == 14772512 and tasks == len(o.prefix_tasks(16, 2))


@pytest.mark.parametrize("force", [0, 1, 2, 4, 8, 3])
def test_fallback_paths_agree_layer_by_layer(force):
    # The partition writes fixed-capacity buckets by claims and large rows'
    # SCATTER writes runs capped from fanout sums; the force bits make the
    # engine take the exact histogram levels (1), upper-bound-count runs (2),
    # or capacities too small so that every level (4) or capped SCATTER (8)
    # overflows and is repeated by its fallback.  Every path gives the same
    # layers as the oracle's census at N = 17.
    from peps_mi355x.abi import (PEPS_FORCE_CAP_OVERFLOW, PEPS_FORCE_EXACT_LEVELS, PEPS_FORCE_LEVEL_OVERFLOW,
                                 PEPS_FORCE_UPPER_RUNS)
    want = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "orbit_layers_n17_n18.json")))
    want = [tuple(l[k] for k in KEYS) for l in want["boards"]["17"]["layers"]]
    pm.profile_kernels(True)
    try:
        r = pm.contract_rows(17, pm.PepsOptions(force_paths=force))
        ks = pm.kernel_stats()
    finally:
        pm.profile_kernels(False)
    show("force", force, sorted(ks))
    assert tuples(r) == want and r.count.value == o.known_count(17)
    if force == 0:
        assert "caps" in ks and "count_ub" not in ks
    if force & (PEPS_FORCE_EXACT_LEVELS | PEPS_FORCE_LEVEL_OVERFLOW):
        assert "part_hist" in ks or "part_hist_runs" in ks
    if force & (PEPS_FORCE_UPPER_RUNS | PEPS_FORCE_CAP_OVERFLOW | PEPS_FORCE_EXACT_LEVELS |
                PEPS_FORCE_LEVEL_OVERFLOW):
        assert "count_ub" in ks
    if force == PEPS_FORCE_CAP_OVERFLOW:
        assert "caps" in ks
    # Graph replay records the path the eager run settled on.
    for rep in range(3):
        g = pm.contract_rows(17, pm.PepsOptions(force_paths=force, graph_mode=1))
        assert tuples(g) == want, rep
    if force in (0, 4):
        eager_b = pm.count_batched(17, 2, pm.PepsOptions(force_paths=force), task_counts=True)
        for rep in range(3):
            b = pm.count_batched(17, 2, pm.PepsOptions(force_paths=force, graph_mode=1), task_counts=True)
            assert b[2] == eager_b[2] and b[0].count.value == o.known_count(17), rep


def test_repeated_runs_identical():
    a = pm.contract_rows(12)
    b = pm.contract_rows(12)
    assert tuples(a) == tuples(b) and a.count.value == b.count.value


def test_device_is_gfx950():
    info = pm.device_info(0)
    assert "gfx950" in info["name"] and info["cu_count"] >= 200


@pytest.mark.parametrize("n", [8, 12, 15])
def test_graph_replay_matches_eager(n):
    # The first call records the size plan (eager), later calls replay the
    # captured HIP graph; all must agree with the oracle layer by layer.
    o_ = pm.PepsOptions(graph_mode=1)
    ref = o.contract_rows(n, backend=o.PARALLEL, threads=8)
    for rep in range(3):
        show("graph", n, rep)
        g = pm.contract_rows(n, o_)
        assert g.count.value == ref.count.value
        assert tuples(g) == tuples(ref), rep
        assert g.peak_states == ref.peak_states


def test_graph_replay_interleaved_with_other_sizes():
    # Replays after other calls that reallocate buffers must recapture.
    o_ = pm.PepsOptions(graph_mode=1)
    for n in (14, 16, 14, 10, 14):
        g = pm.contract_rows(n, o_)
        assert g.count.value == o.known_count(n)


def test_batched_task_counts_match_dfs():
    n, depth = 11, 2
    b, ntasks, counts = pm.count_batched(n, depth, task_counts=True)
    assert ntasks == len(counts) == len(o.prefix_tasks(n, depth))
    tasks = [(c0, c1) for c0 in range((n + 1) // 2) for c1 in range(n)
             if c1 != c0 and abs(c1 - c0) != 1]
    assert len(tasks) == ntasks
    for (c0, c1), cnt in zip(tasks, counts):
        w = 1 if c0 == n - 1 - c0 else 2
        assert cnt == w * o.count_prefix_dfs(n, [c0, c1]), (c0, c1)
    assert sum(counts) == b.count.value == o.known_count(n)


@pytest.mark.parametrize("group", [1, 5, 0])
def test_batched_grouping_and_graph(group):
    # graph_mode: call 1 finds the grouping, call 2 records every group's size
    # plan, calls 3+ replay all groups as one HIP graph.
    ref = None
    for graph in (0, 1):
        for rep in range(4 if graph else 2):
            show("batched graph", group, graph, rep)
            b, ntasks, counts = pm.count_batched(14, 2, pm.PepsOptions(slice_group=group, graph_mode=graph),
                                                 task_counts=True)
            assert b.count.value == 365596 and sum(counts) == 365596
            layers = [tuple(l[k] for k in KEYS) for l in b.layer_list()]
            if ref is None:
                ref = (counts, layers)
            assert (counts, layers) == ref, (graph, rep)


def test_batched_graph_replay_interleaved():
    # Replays of one batched shape between other calls that reallocate.
    o_ = pm.PepsOptions(graph_mode=1)
    want = [pm.count_batched(13, 2, task_counts=True)[2], pm.count_batched(15, 2, task_counts=True)[2]]
    for rep in range(3):
        for i, n in enumerate((13, 15)):
            show("batched interleave", n, rep)
            b, _, counts = pm.count_batched(n, 2, o_, task_counts=True)
            assert counts == want[i] and b.count.value == o.known_count(n)
        pm.contract_rows(16)
