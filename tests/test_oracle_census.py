"""The oracle's two independent routes to the per-row figures at large N
(oracle/layer_census.cpp): the per-depth prefix DFS (every row's
output_weight) and the column-set-grouped breadth-first census (every row's
input_states, row_operator_matched, completed_row_terms and output_states).
Pinned against the reference's published layer supports and totals, against
the sort-reduce restatement, and against brute force."""
import itertools
import json
import os

import pytest

import _oracle as o

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
CK = o.CENSUS_KEYS[1:]


def brute_prefixes(n, r):
    """Non-attacking r-row prefixes of an n x n board, by enumeration."""
    total = 0
    for cols in itertools.permutations(range(n), r):
        if all(abs(cols[i] - cols[j]) != i - j for i in range(r) for j in range(i)):
            total += 1
    return total


def test_depth_counts_match_brute_force():
    for n in range(1, 8):
        want = [brute_prefixes(n, r) for r in range(n + 1)]
        assert o.depth_counts(n, orbits=False, threads=2) == want, n
        # the orbit weights (2 per mirror pair, 1 for the centre) preserve
        # every depth's total by the board's mirror symmetry
        assert o.depth_counts(n, orbits=True, threads=3) == want, n


def test_depth_counts_end_at_known_counts_and_match_the_row_loop():
    for n in range(1, 17):
        d = o.depth_counts(n, orbits=True, threads=8)
        assert d[-1] == o.known_count(n) and d[0] == 1 and d[1] == n, n
        if n <= 13:
            for orb in (True, False):
                w = [l["output_weight"] for l in o.contract_rows(n, symmetry=orb).layer_list()]
                assert w == d[1:], (n, orb)


def test_depth_counts_match_published_n14_layer_weights():
    # benchmarks/n14_layers.csv: output_weight of every row of the N = 14 count
    rows = sorted(FIX["n14_sitewise_layers"], key=lambda r: r["row"])
    assert [r["output_weight"] for r in rows] == o.depth_counts(14, orbits=False)[1:]


def test_census_matches_sort_reduce_restatement():
    for n in range(1, 14):
        for orb in (True, False):
            cen = o.frontier_census(n, orbits=orb, threads=4)
            lay = o.contract_rows(n, symmetry=orb).layer_list()
            assert [tuple(c[k] for k in CK) for c in cen] == [tuple(l[k] for k in CK) for l in lay], (n, orb)


def test_census_matches_published_orbit_layer_supports():
    for n_str, layers in FIX["orbit_layer_supports"].items():
        cen = o.frontier_census(int(n_str), threads=8)
        got = [1] + [c["output_states"] for c in cen]
        want = [l["reachable_states"] for l in sorted(layers, key=lambda x: x["row"])]
        assert got == want, layers[0]["src"]


@pytest.mark.parametrize("mode", ["orbit", "dense"])
def test_census_matches_published_totals_n13_n15(mode):
    for row in FIX["e12_totals"]:
        if row["mode"] != mode:
            continue
        cen = o.frontier_census(row["n"], orbits=mode == "orbit", threads=8)
        assert max(c["output_states"] for c in cen) == row["peak_states"], row["src"]
        assert sum(c["row_operator_matched"] for c in cen) == row["row_operator_matched"], row["src"]
        assert sum(row["n"] * c["input_states"] for c in cen) == row["row_operator_candidates"], row["src"]


def test_census_matches_published_n16_peak_and_transitions():
    cen = o.frontier_census(16, threads=8)
    for row in FIX["orbit_n16"]:
        assert max(c["output_states"] for c in cen) == row["peak_states"], row["src"]
        assert sum(c["row_operator_matched"] for c in cen) == row["row_operator_matched"], row["src"]
    # unsymmetrised N = 16 peak (benchmarks/issue34_same_node_algorithm_comparison_scnet.csv:59)
    peak16 = [r for r in FIX["naive_scnet"] if r["n"] == 16][0]
    dense = o.frontier_census(16, orbits=False, threads=8)
    assert max(c["output_states"] for c in dense) == peak16["peak_states"], peak16["src"]


def test_parallel_prefix_count_matches_the_serial_dfs():
    import random
    rng = random.Random(5)
    for n in (1, 6, 9, 12, 14):
        for k in range(0, 4):
            for _ in range(3):
                cols = [rng.randrange(n) for _ in range(min(k, n))]
                assert o.count_prefix_parallel(n, cols, threads=4) == o.count_prefix_dfs(n, cols), (n, cols)
    # the first-row prefixes of N = 15 sum to Q(15)
    assert sum(o.count_prefix_parallel(15, [c]) for c in range(15)) == o.known_count(15)


def test_committed_n17_fixture_reproduces():
    # tests/golden/orbit_layers_n17_n18.json was written by
    # tests/golden/make_layer_fixtures.py; its N = 17 board is rebuilt here
    # (N = 18 needs ~40 GB and minutes, so only its totals are checked
    # against each other).
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "orbit_layers_n17_n18.json")))
    b = fix["boards"]["17"]
    cen = o.frontier_census(17, threads=8)
    dep = o.depth_counts(17, threads=8)
    assert [tuple(c[k] for k in CK) for c in cen] == [tuple(l[k] for k in CK) for l in b["layers"]]
    assert [l["output_weight"] for l in b["layers"]] == dep[1:]
    for nb in fix["boards"].values():
        n, L = nb["n"], nb["layers"]
        assert nb["count"] == L[-1]["output_weight"] == o.known_count(n)
        assert nb["peak_states"] == max(l["output_states"] for l in L)
        assert nb["row_operator_matched"] == sum(l["row_operator_matched"] for l in L)
        assert all(l["row_operator_candidates"] == n * l["input_states"] for l in L)
        assert all(L[i + 1]["input_states"] == L[i]["output_states"] for i in range(n - 1))
        assert L[0]["output_weight"] == n and L[0]["completed_row_terms"] == (n + 1) // 2
