"""Pins the CPU oracle before it is trusted as the checker.

Each check is against something the oracle's own code does not restate:
the reference's published counters (tests/golden/reference_fixtures.json,
extracted from the reference's benchmark CSVs), the reference's own test
assertions (src/lib.rs:6139-6975, src/dfs_bitmask.rs:354-391), brute force
on small boards, and closed-form invariants.
"""
import itertools
import json
import os
import random

import pytest

import _oracle as o

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))

P1, P2, P3, P4 = 4294967291, 4294967279, 4294967231, 4294967197


# ---------------------------------------------------------------- tensors --

def test_b_has_the_seventeen_sec_vi_entries():
    # src/lib.rs:6188-6233
    b, c, buckets = o.site_tensors()
    assert len(b) == 17
    alpha0 = [e for e in b if e[0] == 0]
    alpha1 = [e for e in b if e[0] == 1]
    assert len(alpha0) == 16 and len(alpha1) == 1
    for e in alpha0:
        _, ci, co, ri, ro, dri, dro, dli, dlo, v = e
        assert (ci, ri, dri, dli) == (co, ro, dro, dlo) and v == 1
    assert alpha1[0][1:] == (0, 1, 0, 1, 0, 1, 0, 1, 1)
    # all 16 incoming signatures appear exactly once among alpha=0
    assert sorted((e[1], e[3], e[5], e[7]) for e in alpha0) == sorted(itertools.product((0, 1), repeat=4))


def test_c_has_seventeen_unit_entries_and_truth_table():
    # src/lib.rs:6284-6297
    _, c, buckets = o.site_tensors()
    assert len(c) == 17 and all(e[8] == 1 for e in c)
    key = lambda ci, ri, dri, dli: ci | (ri << 1) | (dri << 2) | (dli << 3)
    assert buckets[key(0, 0, 0, 0)] == 2
    assert buckets[key(1, 0, 0, 0)] == 1
    assert buckets[key(0, 1, 0, 0)] == 1
    assert buckets[key(1, 1, 1, 1)] == 1
    # index == scan (src/lib.rs:6300-6326)
    for sig in range(16):
        ci, ri, dri, dli = sig & 1, (sig >> 1) & 1, (sig >> 2) & 1, (sig >> 3) & 1
        scanned = [e for e in c if (e[0], e[2], e[4], e[6]) == (ci, ri, dri, dli)]
        assert len(scanned) == buckets[sig]
    # C rows are sorted by legs_key
    lk = [sum(bit << i for i, bit in enumerate(e[:8])) for e in c]
    assert lk == sorted(lk)


def _line_amplitude(b, occupations, end):
    # v0 = (1,0) at the start of a line of B tensors along the column channel.
    vec = [1, 0]
    for alpha in occupations:
        nxt = [0, 0]
        seen = {}
        for e in b:
            if e[0] != alpha:
                continue
            i, j = e[1], e[2]
            if (i, j) in seen:
                assert seen[(i, j)] == e[9]
                continue
            seen[(i, j)] = e[9]
            nxt[j] += vec[i] * e[9]
        vec = nxt
    return vec[0] * end[0] + vec[1] * end[1]


def test_v0_v1_v2_line_boundaries_enforce_exactly_and_at_most_one():
    # src/lib.rs:6264-6281
    b, _, _ = o.site_tensors()
    for bits in range(16):
        occ = [(bits >> s) & 1 for s in range(4)]
        q = sum(occ)
        assert _line_amplitude(b, occ, (0, 1)) == (1 if q == 1 else 0)
        assert _line_amplitude(b, occ, (1, 1)) == (1 if q <= 1 else 0)


# ------------------------------------------------------------ brute force --

def brute_force(n):
    # src/lib.rs:6161-6185, restated independently: place row by row.
    if n == 0:
        return 1
    count = 0

    def place(row, cols):
        nonlocal count
        if row == n:
            count += 1
            return
        for c in range(n):
            if all(c != oc and abs(row - r) != abs(c - oc) for r, oc in enumerate(cols)):
                place(row + 1, cols + [c])

    place(0, [])
    return count


def test_contraction_matches_brute_force_and_known_counts():
    # src/lib.rs:6329-6334 and :6707-6715
    for n in range(0, 10):
        assert o.contract_rows(n).count.value == brute_force(n) == o.known_count(n)
    assert o.contract_rows(10).count.value == o.known_count(10) == 724


def test_known_count_table_edges():
    assert o.known_count(27) == 234907967154122528
    assert o.known_count(28) is None


def test_final_layer_weight_is_q8():
    # src/lib.rs:6718-6722
    r = o.contract_rows(8)
    assert r.layers[7].output_weight.value == 92 and r.count.value == 92


def _prefix_counts(n, rows):
    """Number of non-attacking placements in the first `rows` rows (brute force)."""
    total = 0
    for cols in itertools.product(range(n), repeat=rows):
        if all(cols[i] != cols[j] and abs(i - j) != abs(cols[i] - cols[j])
               for i in range(rows) for j in range(i)):
            total += 1
    return total


def test_layer_weights_count_row_prefixes():
    # Invariant: orbit weighting keeps the total, so sum(output_weight) after
    # row r = number of non-attacking r+1-row prefixes.
    for n in (5, 6, 7):
        r = o.contract_rows(n)
        for layer in range(n):
            assert r.layers[layer].output_weight.value == _prefix_counts(n, layer + 1), (n, layer)


def test_rejects_boards_wider_than_u128_layout():
    # src/lib.rs:6725-6727
    with pytest.raises(o.OracleError, match="supports N <= 42"):
        o.contract_rows(43)
    with pytest.raises(o.OracleError, match="at least one thread"):
        o.contract_rows(8, backend=o.PARALLEL, threads=0)


# ---------------------------------------------------- operator equivalence --

def _reachable(n):
    """All boundary states reachable row by row (independent placement rule)."""
    mask = (1 << n) - 1
    layers = [{(0, 0, 0)}]
    for _ in range(n):
        nxt = set()
        for cols, dr, dl in layers[-1]:
            for c in range(n):
                b = 1 << c
                if (cols | dr | dl) & b:
                    continue
                nxt.add((cols | b, ((dr | b) << 1) & mask, (dl | b) >> 1))
        layers.append(nxt)
    return layers


def test_compiled_operator_matches_sitewise_for_every_reachable_parent_through_n8():
    # src/lib.rs:6345-6390
    for n in range(1, 9):
        for layer in _reachable(n)[:-1]:
            for (cols, dr, dl) in layer:
                a = sorted(o.row_transfer(n, cols, dr, dl, sitewise=True))
                b = sorted(o.row_transfer(n, cols, dr, dl, sitewise=False))
                assert a == b, (n, cols, dr, dl)


def test_invariants_on_boundaries():
    # bit 0 of diag_dr and bit n-1 of diag_dl are never set; popcount(cols) = rows
    n = 9
    for r, layer in enumerate(_reachable(n)):
        for cols, dr, dl in layer:
            assert bin(cols).count("1") == r
            assert dr & 1 == 0 and (dl >> (n - 1)) & 1 == 0


# ---------------------------------------------------- backend cross-checks --

LAYER_KEYS = ("input_states", "completed_row_terms", "output_states", "output_weight")


def _tuples(res):
    return [tuple(l[k] for k in LAYER_KEYS) for l in res.layer_list()]


def test_sort_reduce_matches_hash_materialization_through_n10():
    # src/lib.rs:6607-6650
    for n in range(0, 11):
        a = o.contract_rows(n, symmetry=False)
        b = o.contract_rows(n, symmetry=False, backend=o.HASH)
        assert _tuples(a) == _tuples(b) and a.count.value == b.count.value
        assert a.peak_states == b.peak_states


def test_sitewise_and_compiled_contractions_agree_through_n10():
    # src/lib.rs:6596-6604
    for n in range(0, 11):
        a = o.contract_rows(n, symmetry=False, backend=o.SITEWISE)
        b = o.contract_rows(n, symmetry=False, backend=o.HASH)
        assert _tuples(a) == _tuples(b)


def test_parallel_matches_serial_through_n10():
    # src/lib.rs:6653-6704 and :6904-6920
    for n in range(0, 11):
        for sym in (False, True):
            serial = o.contract_rows(n, symmetry=sym)
            for t in (1, 2, 4):
                par = o.contract_rows(n, symmetry=sym, backend=o.PARALLEL, threads=t)
                assert _tuples(par) == _tuples(serial) and par.count.value == serial.count.value
                assert par.row_operator_matched.value == serial.row_operator_matched.value


def test_sparse_matches_dense_counts():
    for n in range(0, 11):
        d = o.contract_rows(n, sparse=False)
        s = o.contract_rows(n, sparse=True)
        assert _tuples(d) == _tuples(s)
        assert s.row_operator_candidates.value == s.row_operator_matched.value == d.row_operator_matched.value


def test_orbit_slicing_is_exact_and_reduces_support_through_n11():
    # src/lib.rs:6883-6901
    for n in range(2, 12):
        full = o.contract_rows(n, symmetry=False)
        orb = o.contract_rows(n, symmetry=True)
        assert full.count.value == orb.count.value
        if n >= 4:
            assert orb.peak_states < full.peak_states


def test_top_row_orbit_weights():
    # src/lib.rs:6862-6880: weights sum to N, exactly one weight-1 term iff N odd
    for n in range(1, 12):
        tasks = o.prefix_tasks(n, 1)
        ws = [t[3] for t in tasks]
        assert sum(ws) == n
        assert ws.count(1) == (1 if n % 2 else 0)
        assert len(tasks) == (n + 1) // 2


# -------------------------------------------------- published fixtures ----

def test_orbit_layer_supports_match_published_audit():
    # benchmarks/e35_frontier_audit_layers.csv (N=10..14)
    for n_str, layers in FIX["orbit_layer_supports"].items():
        n = int(n_str)
        res = o.contract_rows(n, backend=o.PARALLEL, threads=8)
        got = [1] + [l["output_states"] for l in res.layer_list()]
        want = [l["reachable_states"] for l in sorted(layers, key=lambda x: x["row"])]
        assert got == want, n


def test_orbit_audit_totals():
    # benchmarks/e35_frontier_audit_release.csv: forward transitions = completed terms
    for row in FIX["orbit_audit_totals"]:
        res = o.contract_rows(row["n"], backend=o.PARALLEL, threads=8)
        assert res.count.value == row["count"]
        assert res.peak_states == row["peak_states"]
        assert sum(l["completed_row_terms"] for l in res.layer_list()) == row["completed_row_terms"]


@pytest.mark.parametrize("mode", ["orbit", "dense", "orbit_sparse", "dense_sparse"])
def test_e12_totals_n13_n15(mode):
    # benchmarks/e12_d4_orbit_slicing_release.csv (serial rows), run with the
    # parallel restatement (same counters, tested equal above).
    for row in FIX["e12_totals"]:
        if row["mode"] != mode or (row["n"] == 15 and mode != "orbit"):
            continue
        res = o.contract_rows(row["n"], symmetry=mode.startswith("orbit"),
                              sparse=mode.endswith("sparse"), backend=o.PARALLEL, threads=8)
        assert res.count.value == row["count"]
        assert res.peak_states == row["peak_states"]
        assert res.row_operator_candidates.value == row["row_operator_candidates"]
        assert res.row_operator_matched.value == row["row_operator_matched"]


def test_dense_totals_n10_n14():
    # benchmarks/current_scaling_peps_release.csv (serial_sort_reduce)
    for row in FIX["dense_totals"]:
        if row["n"] > 14:
            continue
        res = o.contract_rows(row["n"], symmetry=False, backend=o.PARALLEL, threads=8)
        assert (res.count.value, res.peak_states, res.row_operator_candidates.value,
                res.row_operator_matched.value) == (row["count"], row["peak_states"],
                                                    row["row_operator_candidates"],
                                                    row["row_operator_matched"])


def test_naive_sitewise_counters_n4_n11():
    # benchmarks/naive_release.csv: the pre-index sitewise HashMap revision
    # scanned all 17 C entries per partial row per site.
    for row in FIX["naive_totals"]:
        if row["n"] > 11:
            continue
        res = o.contract_rows(row["n"], symmetry=False, backend=o.SITEWISE, linear_scan_counters=True)
        assert res.count.value == row["count"]
        assert res.peak_states == row["peak_states"]
        assert res.tensor_entries_examined.value == row["tensor_entries_examined"]
        assert res.tensor_entries_matched.value == row["tensor_entries_matched"]


def test_naive_scnet_peaks_n1_n12():
    for row in FIX["naive_scnet"]:
        if row["n"] > 12:
            continue
        res = o.contract_rows(row["n"], symmetry=False, backend=o.SITEWISE)
        assert res.count.value == row["count"] and res.peak_states == row["peak_states"]
        assert res.tensor_entries_matched.value == row["tensor_entries_matched"]


# -------------------------------------------------------------------- CRT --

def test_crt_plans_match_published_products():
    # benchmarks/e45_wide_crt_plan_n22_n28.csv
    for row in FIX["crt_plans"]:
        primes, prod, fact = o.crt_plan(row["n"])
        assert primes == row["primes"] and prod == row["modulus_product"]
        assert fact == row["factorial_bound"] and prod > fact
    assert o.crt_plan(12)[0] == [P1]
    assert o.crt_plan(13)[0] == [P1, P2]
    assert o.crt_plan(20)[0] == [P1, P2]
    assert o.crt_plan(21)[0] == [P1, P2, P3]
    with pytest.raises(o.OracleError):
        o.crt_plan(35)


def test_crt_reconstructs_published_residues():
    # benchmarks/e45_wide_crt_release.csv: Q(19) residues 673090557|673090569
    for row in FIX["crt_release"]:
        assert o.crt_reconstruct(row["residues"], row["primes"]) == row["count"]
        assert [row["count"] % p for p in row["primes"]] == row["residues"]


def test_crt_round_trips_and_primes():
    # src/lib.rs:6530-6545: Q(27) through 4 residues; every candidate prime is prime
    for p in (P1, P2, P3, P4):
        assert o.is_prime(p)
    assert not o.is_prime(P1 + 2) or (P1 + 2) in (P1, P2, P3, P4)
    assert not o.is_prime(4294967297)  # 641 * 6700417
    ps = [P1, P2, P3, P4]
    q27 = o.known_count(27)
    assert o.crt_reconstruct([q27 % p for p in ps], ps) == q27
    q20 = o.known_count(20)
    assert [q20 % p for p in ps[:3]] == [374483265, 374483373, 374483805]
    rng = random.Random(7)
    for _ in range(200):
        v = rng.randrange(0, P1 * P2 * P3)
        assert o.crt_reconstruct([v % p for p in ps[:3]], ps[:3]) == v


# -------------------------------------------------------------------- DFS --

def test_dfs_matches_known_counts_through_fourteen():
    # src/dfs_bitmask.rs:360-365 (through 16 there; 14 keeps the CPU suite short)
    for n in range(0, 15):
        assert o.dfs(n)["count"] == o.known_count(n)


def test_dfs_parallel_and_profile_agree():
    # src/dfs_bitmask.rs:368-384
    for n in range(8, 14):
        assert o.dfs(n, 1)["count"] == o.dfs(n, 4)["count"]
    s, p = o.dfs(12, 1, True), o.dfs(12, 4, True)
    assert s["recursive_nodes"] == p["recursive_nodes"]
    assert s["candidate_placements"] == p["candidate_placements"]
    with pytest.raises(o.OracleError):
        o.dfs(8, 0)
    with pytest.raises(o.OracleError):
        o.dfs(28, 1)


# ----------------------------------------------------------- prefixes ----

def test_prefix_task_counts():
    assert len(o.prefix_tasks(18, 2)) == 136
    assert len(o.prefix_tasks(20, 2)) == 171
    for n in (8, 11, 14):
        tasks = o.prefix_tasks(n, 2)
        # weights of depth-2 tasks sum to the number of 2-row prefixes
        assert sum(t[3] for t in tasks) == _prefix_counts(n, 2)


def test_prefix_contraction_matches_dfs():
    rng = random.Random(0)
    for n in range(4, 12):
        for k in (1, 2, 3):
            for _ in range(4):
                cols = [rng.randrange(n) for _ in range(k)]
                got = o.contract_prefix(n, cols).count.value
                assert got == o.count_prefix_dfs(n, cols), (n, cols)


def test_prefix_sums_reconstruct_the_count():
    for n in (6, 8, 9):
        assert sum(o.contract_prefix(n, [c]).count.value for c in range(n)) == o.known_count(n)
        assert sum(o.count_prefix_dfs(n, [c]) for c in range(n)) == o.known_count(n)


def test_attacking_prefix_counts_zero():
    assert o.contract_prefix(8, [0, 1]).count.value == 0
    assert o.count_prefix_dfs(8, [3, 3]) == 0
    assert o.count_prefix_dfs(8, [0, 4, 7, 5, 2, 6, 1, 3]) == 1  # src/lib.rs:6759
