"""Tensor fidelity: the explicit C tensor, its fail-closed compilation, the
literal sitewise contraction on the GPU and the layer trace certificate.

CPU tests pin the product's C tensor and compile against the oracle and the
reference's messages, SHA-256 against hashlib, and the oracle's
custom-tensor path against closed forms.  GPU tests compare the literal
sitewise kernels with the oracle layer by layer (bit-exact), with the
published per-layer site counters (benchmarks/n14_layers.csv), and the trace
digests with hashlib over the oracle's canonical layers."""
import hashlib
import json
import os
import random

import pytest

import _oracle as o
import peps_mi355x as pm
from peps_mi355x.abi import TensorEntry

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
KEYS = ("input_states", "tensor_entries_examined", "tensor_entries_matched", "row_operator_candidates",
        "row_operator_matched", "completed_row_terms", "output_states", "output_weight")


def tuples(res):
    return [tuple(l[k] for k in KEYS) for l in res.layer_list()]


def show(*a):
    print(*a, flush=True)


def entry(legs, value=1):
    return TensorEntry(*legs, value)


def sec_vi():
    return pm.site_tensor_c()


def mutations():
    """Mutations named in docs/sec_vi_row_automaton_equivalence.md:213-217 that
    stay inside binary legs: each must be rejected by the compiler."""
    c = sec_vi()
    occ_i = next(i for i, e in enumerate(c) if e.legs() == (0, 1, 0, 1, 0, 1, 0, 1))
    empty_i = next(i for i, e in enumerate(c) if e.legs() == (0,) * 8)
    out = {}
    m = list(c)
    m[occ_i] = entry(c[occ_i].legs(), 2)
    out["coefficient"] = (m, "requires unit Sec. VI coefficients")
    m = [e for i, e in enumerate(c) if i != empty_i]
    out["drop_empty_passthrough"] = (m, "missing a pass-through signature")
    m = list(c)
    legs = list(c[occ_i].legs())
    legs[2], legs[3] = 1, 0  # the occupied entry's row legs swapped (in 1, out 0)
    m[occ_i] = entry(tuple(legs))
    out["permuted_leg"] = (m, "outside the Sec. VI row automaton")
    m = list(c)
    legs = list(c[occ_i].legs())
    legs[7] = 0  # only the diag_dl outgoing endpoint reversed
    m[occ_i] = entry(tuple(legs))
    out["one_diagonal_endpoint"] = (m, "outside the Sec. VI row automaton")
    m = list(c) + [entry((0,) * 8)]
    out["duplicate_passthrough"] = (m, "duplicate C pass-through signature 0000")
    m = [e for i, e in enumerate(c) if i != occ_i]
    out["no_occupied"] = (m, "missing the occupied entry")
    m = list(c) + [entry(c[occ_i].legs())]
    out["two_occupied"] = (m, "multiple occupied C entries")
    return out


# ---------------------------------------------------------------- CPU side --

def test_site_tensor_c_matches_oracle():
    _, oc, buckets = o.site_tensors()
    got = [e.legs() + (e.value,) for e in sec_vi()]
    assert got == oc and len(got) == 17
    assert buckets == [2] + [1] * 15


def test_compile_accepts_sec_vi_and_rejects_every_mutation():
    occ = pm.compile_row_operator(sec_vi())
    assert occ.legs() == (0, 1, 0, 1, 0, 1, 0, 1) and occ.value == 1
    shuffled = sec_vi()
    random.Random(5).shuffle(shuffled)
    assert pm.compile_row_operator(shuffled).legs() == occ.legs()
    for name, (m, msg) in mutations().items():
        with pytest.raises(pm.PepsError, match=msg):
            pm.compile_row_operator(m)
        show("rejected", name)


def test_sha256_known_answers():
    for msg in (b"", b"abc", b"a" * 55, b"a" * 56, b"a" * 64, bytes(range(256)) * 40):
        assert pm.sha256(msg) == hashlib.sha256(msg).hexdigest()


def test_canonical_layer_serialisation():
    b = pm.canonical_layer_bytes([(3, 0, 1, 5), (1, 2, 0, (1 << 70) + 9)])
    assert len(b) == 80
    # sorted by columns first; little-endian u64 fields; u128 weight lo then hi
    assert b[:8] == (1).to_bytes(8, "little") and b[24:40] == ((1 << 70) + 9).to_bytes(16, "little")
    assert b[40:48] == (3).to_bytes(8, "little")


def test_oracle_custom_tensor_closed_forms():
    # Sec. VI's C in any entry order gives the sitewise result unchanged.
    c = sec_vi()
    random.Random(1).shuffle(c)
    for n in range(1, 9):
        a = o.contract_rows_tensor(n, c)
        b = o.contract_rows(n, symmetry=False, backend=o.SITEWISE)
        assert tuples(a) == tuples(b) and a.count.value == o.known_count(n)
    # Occupied coefficient 2: every solution has exactly N occupied sites, so
    # the count is 2^N Q(N); after row r every weight carries 2^(r+1).
    m, _ = mutations()["coefficient"]
    for n in range(1, 9):
        r = o.contract_rows_tensor(n, m)
        assert r.count.value == (2 ** n) * o.known_count(n)
        plain = o.contract_rows(n, symmetry=False, backend=o.SITEWISE)
        for lm, lp in zip(r.layer_list(), plain.layer_list()):
            assert lm["output_states"] == lp["output_states"]
            assert lm["output_weight"] == lp["output_weight"] * 2 ** (lm["row"] + 1)


def test_oracle_sitewise_n14_matches_published_site_counters():
    # benchmarks/n14_layers.csv (pre-index revision: 17 entries scanned per
    # partial row per site).
    r = o.contract_rows(14, symmetry=False, backend=o.SITEWISE, linear_scan_counters=True)
    for got, want in zip(r.layer_list(), FIX["n14_sitewise_layers"]):
        assert got["row"] + 1 == want["row"], want["src"]
        for k in ("input_states", "tensor_entries_examined", "tensor_entries_matched",
                  "completed_row_terms", "output_states", "output_weight"):
            assert got[k] == want[k], (k, want["src"])


# ---------------------------------------------------------------- GPU side --

@pytest.mark.gpu
@pytest.mark.parametrize("counters", [0, 1])
def test_sitewise_gpu_matches_oracle_through_n11(counters):
    for n in range(1, 12):
        show("sitewise", n, counters)
        g = pm.contract_rows_sitewise(n, pm.PepsOptions(site_counters=counters))
        r = o.contract_rows(n, symmetry=False, backend=o.SITEWISE, linear_scan_counters=bool(counters))
        assert tuples(g) == tuples(r), n
        assert g.count.value == r.count.value == o.known_count(n)
        assert g.peak_states == r.peak_states
        assert g.tensor_entries_examined.value == r.tensor_entries_examined.value
        assert g.tensor_entries_matched.value == r.tensor_entries_matched.value
        assert g.row_operator_candidates.value == 0 and g.row_operator_matched.value == 0


@pytest.mark.gpu
def test_sitewise_gpu_n14_matches_published_site_counters():
    g = pm.contract_rows_sitewise(14, pm.PepsOptions(site_counters=1))
    assert g.count.value == 365596
    for got, want in zip(g.layer_list(), FIX["n14_sitewise_layers"]):
        for k in ("input_states", "tensor_entries_examined", "tensor_entries_matched",
                  "completed_row_terms", "output_states", "output_weight"):
            assert got[k] == want[k], (k, want["src"])


@pytest.mark.gpu
def test_sitewise_gpu_agrees_with_compiled_path():
    for n in (12, 13, 15):
        s = pm.contract_rows_sitewise(n)
        c = pm.contract_rows_sort_reduce(n)
        keys = ("input_states", "completed_row_terms", "output_states", "output_weight")
        assert [tuple(l[k] for k in keys) for l in s.layer_list()] == \
               [tuple(l[k] for k in keys) for l in c.layer_list()]
        assert s.count.value == c.count.value


@pytest.mark.gpu
def test_custom_tensors_match_oracle():
    shuffled = sec_vi()
    random.Random(9).shuffle(shuffled)
    # A duplicated pass-through doubles the partial rows at every empty site,
    # so it is contracted only while they fit the kernel's partial capacity.
    cases = [("shuffled", shuffled)] + [(k, v[0]) for k, v in mutations().items()]
    for name, m in cases:
        for n in ((2, 3) if name == "duplicate_passthrough" else (5, 8)):
            show("tensor", name, n)
            g = pm.contract_rows_with_tensor(n, m, pm.PepsOptions(crt_lanes=4))
            r = o.contract_rows_tensor(n, m)
            assert tuples(g) == tuples(r), (name, n)
            assert g.count.value == r.count.value, (name, n)
    g = pm.contract_rows_with_tensor(9, mutations()["coefficient"][0], pm.PepsOptions(crt_lanes=4))
    assert g.count.value == 2 ** 9 * 352
    # 2^k + 2^k - 1 partial rows after k empty sites: N = 5 exceeds the
    # capacity of 24 and fails cleanly.
    with pytest.raises(pm.PepsError, match="partial rows"):
        pm.contract_rows_with_tensor(5, mutations()["duplicate_passthrough"][0])


def _oracle_digests(n, symmetry, backend, entries=None):
    res, layers = o.contract_rows_tensor(n, entries, symmetry=symmetry, backend=backend, layers=True)
    return [{"row": row, "support": len(st), "weight_sum": sum(w for *_, w in st),
             "sha256": hashlib.sha256(pm.canonical_layer_bytes(st)).hexdigest()}
            for row, st in sorted(layers.items())], res


@pytest.mark.gpu
def test_trace_digests_match_oracle_layers():
    for n in (6, 9, 11):
        for op, sym, backend in (("compiled", False, o.SORT_REDUCE), ("compiled", True, o.SORT_REDUCE),
                                 ("sitewise", False, o.SITEWISE)):
            show("trace", n, op, sym)
            got = pm.trace_layers(n, op, symmetry=sym)
            want, res = _oracle_digests(n, sym, backend)
            assert [{k: d[k] for k in ("row", "support", "weight_sum", "sha256")} for d in got] == want
            assert [d["transitions"] for d in got] == [l["completed_row_terms"] for l in res.layer_list()]


@pytest.mark.gpu
def test_trace_literal_and_compiled_agree_and_mutations_show():
    n = 10
    a = pm.trace_layers(n, "compiled")
    b = pm.trace_layers(n, "sitewise")
    assert a == b
    for name, (m, msg) in mutations().items():
        with pytest.raises(pm.PepsError, match=msg):
            pm.trace_layers(n, "compiled", tensor=m)
        k = 3 if name == "duplicate_passthrough" else n
        t = pm.trace_layers(k, "sitewise", tensor=m, options=pm.PepsOptions(crt_lanes=4))
        base = a if k == n else pm.trace_layers(k, "compiled")
        assert [d["sha256"] for d in t] != [d["sha256"] for d in base], name


@pytest.mark.gpu
def test_materialised_orbit_trace_matches_fused_layers():
    for n in (12, 14):
        t = pm.trace_layers(n, "compiled", symmetry=True)
        f = pm.contract_rows(n)
        assert [(d["support"], d["weight_sum"], d["transitions"]) for d in t] == \
               [(l["output_states"], l["output_weight"], l["completed_row_terms"]) for l in f.layer_list()]
