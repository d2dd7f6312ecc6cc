"""Candidate record formats of the fused row pipeline.  Packed 12-byte
records (the key and the exact integer weight in 96 bits) and lane records
(the key and one u32 residue per CRT prime) must give identical counts,
layer statistics and residues.  A packed weight that outgrows its field must
fall back to lane records, or fail under PEPS_RECORDS_PACKED.  The host
narrows the field to force that case.  Every comparison is exact."""
import pytest

import _oracle as o
import peps_mi355x as pm
from peps_mi355x.abi import PEPS_RECORDS_AUTO, PEPS_RECORDS_LANES, PEPS_RECORDS_PACKED

pytestmark = pytest.mark.gpu

KEYS = ("input_states", "row_operator_candidates", "row_operator_matched", "completed_row_terms",
        "output_states", "output_weight", "partition_levels", "final_buckets")


def opts(fmt, wbits=0, **kw):
    return pm.PepsOptions(record_format=fmt, packed_weight_bits=wbits, **kw)


def tuples(res):
    return [tuple(l[k] for k in KEYS) for l in res.layer_list()]


def show(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("symmetry", [True, False])
def test_packed_and_lane_records_agree_layer_by_layer(symmetry):
    run = pm.contract_rows if symmetry else pm.contract_rows_sort_reduce
    for n in range(1, 17):
        show("records", n, symmetry)
        p = run(n, opts(PEPS_RECORDS_PACKED))
        q = run(n, opts(PEPS_RECORDS_LANES))
        assert p.record_bytes == 12 and q.record_bytes == 16, n
        assert p.count.value == q.count.value == o.known_count(n), n
        assert tuples(p) == tuples(q), n
        assert list(p.residues)[:p.crt_lanes] == list(q.residues)[:q.crt_lanes]
        assert p.peak_states == q.peak_states
        assert p.row_operator_matched.value == q.row_operator_matched.value


def test_packed_records_against_the_oracle_n13():
    g = pm.contract_rows(13)
    r = o.contract_rows(13)
    assert g.record_bytes == 12
    assert g.count.value == r.count.value == 73712
    assert [tuple(l[k] for k in KEYS[:6]) for l in g.layer_list()] == \
           [tuple(l[k] for k in KEYS[:6]) for l in r.layer_list()]


def test_three_lane_records_match_packed():
    # Three CRT lanes: 24-byte lane records against 12-byte packed records.
    p = pm.contract_rows(14, opts(PEPS_RECORDS_AUTO, crt_lanes=3))
    q = pm.contract_rows(14, opts(PEPS_RECORDS_LANES, crt_lanes=3))
    assert (p.record_bytes, q.record_bytes) == (12, 24)
    assert p.count.value == q.count.value == 365596
    assert list(p.residues)[:3] == list(q.residues)[:3] == [365596 % x for x in list(p.primes)[:3]]
    assert tuples(p) == tuples(q)


@pytest.mark.parametrize("wbits", [1, 2])
def test_narrow_packed_weights_fall_back_to_lane_records(wbits):
    # wbits = 1: the doubled orbit weight of the seed expansion overflows;
    # wbits = 2: a merged weight of 4 or more overflows.  N = 10 ends with 724.
    show("fallback", wbits)
    r = pm.contract_rows(10, opts(PEPS_RECORDS_AUTO, wbits))
    assert r.count.value == 724 and r.record_bytes == 16
    assert tuples(r) == tuples(pm.contract_rows(10, opts(PEPS_RECORDS_LANES)))
    with pytest.raises(pm.PepsError, match="overflow"):
        pm.contract_rows(10, opts(PEPS_RECORDS_PACKED, wbits))
    assert pm.contract_rows(10, opts(PEPS_RECORDS_AUTO, 40)).record_bytes == 12


def test_packed_field_width_sweep_is_exact_and_monotone():
    # Whatever width a run needs, the count is exact; once a width suffices,
    # every wider field stays packed.
    for n in (11, 13):
        sizes = []
        for wbits in range(1, 13):
            r = pm.contract_rows(n, opts(PEPS_RECORDS_AUTO, wbits))
            assert r.count.value == o.known_count(n), (n, wbits)
            sizes.append(r.record_bytes)
        show("widths", n, sizes)
        assert set(sizes) <= {12, 16} and sizes[0] == 16
        first = sizes.index(12) if 12 in sizes else len(sizes)
        assert all(x == 12 for x in sizes[first:]), sizes


def test_batched_and_graph_replay_in_both_formats():
    ref = None
    for fmt in (PEPS_RECORDS_AUTO, PEPS_RECORDS_LANES):
        show("batched", fmt)
        b, tasks, per_task = pm.count_batched(13, 2, opts(fmt), task_counts=True)
        assert b.count.value == 73712 and sum(per_task) == 73712
        if ref is None:
            ref = (b.row_operator_matched.value, tasks, per_task)
        assert (b.row_operator_matched.value, tasks, per_task) == ref
        for _ in range(2):
            g = pm.contract_rows(12, opts(fmt, graph_mode=1))
            assert g.count.value == 14200 and g.record_bytes == (12 if fmt == PEPS_RECORDS_AUTO else 16)
    # the forced fallback under graph replay: every repeat stays exact
    for _ in range(3):
        g = pm.contract_rows(11, opts(PEPS_RECORDS_AUTO, 2, graph_mode=1))
        assert g.count.value == 2680 and g.record_bytes == 16
    b, _ = pm.count_batched(12, 2, opts(PEPS_RECORDS_AUTO, 2, graph_mode=1, slice_group=8))
    assert b.count.value == 14200


def test_prefix_counts_packed_match_dfs():
    for cols in ([0], [1, 3], [2, 0, 3]):
        assert pm.count_prefix(12, cols, opts(PEPS_RECORDS_PACKED)) == o.count_prefix_dfs(12, cols), cols


def test_n21_prefixes_at_the_widest_key():
    # N = 21 is the widest board of the u64 key (3N = 63 bits, no slice
    # bits), where the packed weight field is 96 - 63 = 33 bits.  Three-row
    # prefixes (18 rows of work): packed (default) and lane records against
    # the oracle's parallel prefix DFS, and the mirror prefix (c -> 20 - c),
    # whose count the board's reflection fixes.
    for cols in ([0, 2, 4], [10, 3, 17]):
        show("n21 prefix", cols)
        want = o.count_prefix_parallel(21, cols, threads=16)
        got = pm.count_prefix(21, cols)
        lanes = pm.count_prefix(21, cols, opts(PEPS_RECORDS_LANES))
        mirror = pm.count_prefix(21, [20 - c for c in cols])
        assert got == lanes == mirror == want, (cols, got, lanes, mirror, want)


def test_published_crt_release_n18_n19():
    # benchmarks/e45_wide_crt_release.csv rows for N = 18 and 19: counts and
    # both CRT residues.  N = 18 through the default (packed) plain count;
    # N = 19's plain frontier (1.25e10 candidates) does not fit one GPU's
    # two candidate buffers, so it runs as 2-row prefix slices.
    import json
    import os
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
    rows = {r["n"]: r for r in fix["crt_release"]}
    for n in (18, 19):
        show("crt release", n)
        g = pm.contract_rows(n) if n == 18 else pm.count_batched(n, 2)[0]
        row = rows[n]
        assert g.record_bytes == 12
        assert g.count.value == row["count"], row["src"]
        assert list(g.primes)[:g.crt_lanes] == row["primes"], row["src"]
        assert list(g.residues)[:g.crt_lanes] == row["residues"], row["src"]


@pytest.mark.parametrize("lanes,record_bytes", [(2, 16), (3, 24)])
def test_lane_records_through_capped_runs_and_fixed_buckets(lanes, record_bytes):
    # N = 17's large rows write capped SCATTER runs and fixed-capacity
    # partition buckets; lane records take their own instantiations of them.
    p = pm.contract_rows(17, opts(PEPS_RECORDS_AUTO, crt_lanes=lanes))
    pm.profile_kernels(True)
    try:
        q = pm.contract_rows(17, opts(PEPS_RECORDS_LANES, crt_lanes=lanes))
        ks = pm.kernel_stats()
    finally:
        pm.profile_kernels(False)
    show("lanes", lanes, sorted(ks))
    assert (p.record_bytes, q.record_bytes) == (12, record_bytes)
    assert "caps" in ks
    assert p.count.value == q.count.value == o.known_count(17)
    assert list(q.residues)[:lanes] == [o.known_count(17) % x for x in list(q.primes)[:lanes]]
    assert tuples(p) == tuples(q)
