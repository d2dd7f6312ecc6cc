"""GPU parity at the benchmark's own sizes: the N = 16 layers against the
oracle's parallel sort-reduce restatement, every N = 17 and N = 18 layer
against the oracle's census and prefix-DFS figures (the committed fixture
made by tests/golden/make_layer_fixtures.py), the batched N = 18 count under
single-graph replay, the N = 20 count with three CRT lanes, and served
requests at N = 17..20.  Integer work: every comparison is exact."""
import json
import os

import pytest

import _oracle as o
import peps_mi355x as pm
from peps_mi355x.abi import PEPS_RECORDS_LANES

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
FIX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))
LAYERS = json.load(open(os.path.join(HERE, "golden", "orbit_layers_n17_n18.json")))["boards"]
KEYS = ("input_states", "row_operator_candidates", "row_operator_matched", "completed_row_terms",
        "output_states", "output_weight")
THREADS = 16  # the GPU host's CPU share


def tuples(layers):
    return [tuple(l[k] for k in KEYS) for l in layers]


def show(*a):
    print(*a, flush=True)


def test_n16_layers_match_oracle_parallel_sort_reduce():
    # contract_rows_d4_orbit_parallel_sort_reduce, src/lib.rs:5469-5626
    g = pm.contract_rows(16)
    show("n16 gpu done")
    r = o.contract_rows(16, backend=o.PARALLEL, threads=THREADS)
    assert tuples(g.layer_list()) == tuples(r.layer_list())
    assert g.peak_states == r.peak_states == 105530452
    assert g.row_operator_matched.value == r.row_operator_matched.value == 477265568
    assert g.count.value == r.count.value == 14772512


@pytest.mark.parametrize("n", [17, 18])
def test_every_layer_matches_the_census_and_prefix_counts(n):
    want = LAYERS[str(n)]
    g = pm.contract_rows(n)
    show("layers", n, g.peak_states, g.row_operator_matched.value)
    assert tuples(g.layer_list()) == tuples(want["layers"])
    assert g.peak_states == want["peak_states"]
    # the benchmark's transitions numerator
    assert g.row_operator_matched.value == want["row_operator_matched"]
    assert g.row_operator_candidates.value == want["row_operator_candidates"]
    assert g.count.value == want["count"] == o.known_count(n)


def test_n18_lane_records_match_packed_records_layer_by_layer():
    p = pm.contract_rows(18)
    q = pm.contract_rows(18, pm.PepsOptions(record_format=PEPS_RECORDS_LANES))
    assert (p.record_bytes, q.record_bytes) == (12, 16)
    assert tuples(p.layer_list()) == tuples(q.layer_list()) == tuples(LAYERS["18"]["layers"])
    assert list(p.residues)[:2] == list(q.residues)[:2]


def test_batched_n18_single_graph_replay():
    # Config 3: the 136 two-row prefix tasks of build_wide_crt_tasks
    # (src/lib.rs:3888-3956).  graph_mode: call 1 fixes the task grouping,
    # call 2 records every group's size plan, calls 3+ replay one HIP graph.
    n = 18
    row = [r for r in FIX["crt_release"] if r["n"] == n][0]
    opts = pm.PepsOptions(graph_mode=1)
    first = None
    for call in range(4):
        b, ntasks, counts = pm.count_batched(n, 2, opts, task_counts=True)
        show("batched n18 call", call, b.count.value, b.kernel_s)
        assert ntasks == len(counts) == 136 == len(o.prefix_tasks(n, 2))
        assert b.count.value == sum(counts) == row["count"], row["src"]
        assert list(b.primes)[:b.crt_lanes] == row["primes"] and list(b.residues)[:b.crt_lanes] == row["residues"]
        # the per-row totals of the slices equal the plain count's from row 2 on
        assert [l["output_weight"] for l in b.layer_list()] == \
               [l["output_weight"] for l in LAYERS["18"]["layers"]][2:]
        if first is None:
            first = counts
        assert counts == first, call
    # Every first-row column's tasks sum to twice (orbit weight) the DFS
    # completions of that one-queen prefix.
    tasks = [(c0, c1) for c0 in range((n + 1) // 2) for c1 in range(n) if c1 != c0 and abs(c1 - c0) != 1]
    for c0 in range((n + 1) // 2):
        got = sum(c for (a, _), c in zip(tasks, first) if a == c0)
        assert got == 2 * o.count_prefix_parallel(n, [c0], threads=THREADS), c0
    # and three single tasks against their own DFS
    for i in (0, 67, 135):
        c0, c1 = tasks[i]
        assert first[i] == 2 * o.count_prefix_parallel(n, [c0, c1], threads=THREADS), (c0, c1)


def test_n20_three_lane_count():
    # Config 5: N = 20 as 171 two-row tasks with three CRT lanes.  Q(20)
    # (src/lib.rs:5752-5784; benchmarks/issue34_same_node_algorithm_comparison_scnet.csv:66).
    n = 20
    b, ntasks, counts = pm.count_batched(n, 2, pm.PepsOptions(crt_lanes=3), task_counts=True)
    show("n20", b.count.value, b.kernel_s, b.record_bytes)
    assert ntasks == 171 == len(counts)
    assert b.count.value == sum(counts) == o.known_count(n) == 39029188884
    assert b.crt_lanes == 3
    assert list(b.residues)[:3] == [374483265, 374483373, 374483805]
    assert list(b.residues)[:3] == [o.known_count(n) % p for p in list(b.primes)[:3]]


def test_served_requests_n17_to_n20():
    # Config 4's large end: served prefix counts at N = 17..20.  N <= 18 is
    # pinned by the prefix DFS; N = 19 by the sum over the first-row column
    # (Q(19)) and mirror symmetry; N = 20 by mirror symmetry and, for the
    # two-row prefixes, the same request run alone.
    reqs = [(17, [3]), (17, [8]), (17, [0, 5]), (17, [16, 2]), (17, [4, 11]),
            (18, [0]), (18, [7]), (18, [1, 9]), (18, [12, 4]), (18, [17, 0])]
    reqs += [(19, [c]) for c in range(19)]
    reqs += [(20, [2]), (20, [17]), (20, [1, 5]), (20, [18, 14]), (20, [6, 9]), (20, [13, 10])]
    with pm.Server() as s:
        tickets = [s.submit(n, cols) for n, cols in reqs]
        outs = [s.wait(t, 600) for t in tickets]
        st = s.stats()
    assert st["failed"] == 0 and st["completed"] == len(reqs)
    got = {}
    for (n, cols), r in zip(reqs, outs):
        show("served", n, cols, r.count, r.tasks, r.launches, r.max_batch)
        assert r.status == 0, r.error
        got[(n, tuple(cols))] = r.count
    for (n, cols) in reqs:
        if n <= 18:
            assert got[(n, tuple(cols))] == o.count_prefix_parallel(n, cols, threads=THREADS), (n, cols)
            assert got[(n, tuple(cols))] == pm.count_prefix(n, cols), (n, cols)
    assert sum(got[(19, (c,))] for c in range(19)) == o.known_count(19) == 4968057848
    for c in range(19):
        assert got[(19, (c,))] == got[(19, (18 - c,))], c
    assert got[(20, (2,))] == got[(20, (17,))] > 0
    assert got[(20, (1, 5))] == got[(20, (18, 14))] > 0
    assert got[(20, (6, 9))] == got[(20, (13, 10))] > 0
    for cols in ([1, 5], [6, 9]):
        assert got[(20, tuple(cols))] == pm.count_prefix(20, cols), cols
