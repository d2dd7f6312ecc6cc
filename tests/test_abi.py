"""CPU-side checks of libpeps_mi355x.so: it loads, exports every entry point
include/peps_mi355x.h declares, its host-only utilities agree with the
oracle, the CLI keeps the reference's grammar, and contractions fail loudly
without a GPU (no CPU fallback)."""
import ctypes
import os
import re
import subprocess

import pytest

import _oracle as o
import peps_mi355x as pm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "peps_mi355x.h")
CLI = os.path.join(ROOT, "quantum-harness-peps-naive_amd", "nqueens-peps")


def _has_gpu():
    try:
        pm.device_info(0)
        return True
    except pm.PepsError:
        return False


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|uint64_t|void|const char\*)\s+(peps_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = pm.load()
    syms = declared_symbols()
    assert len(syms) >= 11
    for s in syms:
        assert hasattr(lib, s), s


def test_abi_struct_sizes_match_header_layout():
    # peps_u128 is two u64; the result struct holds PEPS_MAX_N layers.
    from peps_mi355x.abi import ContractionResult, LayerMetric, U128
    assert ctypes.sizeof(U128) == 16
    assert ctypes.sizeof(ContractionResult) > 42 * ctypes.sizeof(LayerMetric)


def test_known_count_matches_oracle():
    for n in range(0, 30):
        assert pm.known_count(n) == o.known_count(n)


def test_crt_plan_and_reconstruct_match_oracle():
    for n in range(0, 35):
        primes, prod = pm.crt_plan(n)
        oprimes, oprod, _ = o.crt_plan(n)
        assert (primes, prod) == (oprimes, oprod)
    with pytest.raises(pm.PepsError):
        pm.crt_plan(35)
    import random
    rng = random.Random(3)
    ps = [4294967291, 4294967279, 4294967231, 4294967197]
    for k in (1, 2, 3, 4):
        for _ in range(50):
            v = rng.randrange(0, 1 << (32 * k - 1))
            res = [v % p for p in ps[:k]]
            assert pm.crt_reconstruct(res, ps[:k]) == o.crt_reconstruct(res, ps[:k]) == v


def test_prefix_tasks_match_oracle():
    for n in (1, 2, 5, 8, 11, 18, 20):
        for d in (0, 1, 2, 3):
            assert pm.prefix_tasks(n, d) == o.prefix_tasks(n, d), (n, d)


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU failure path")
def test_contraction_fails_loudly_without_gpu():
    with pytest.raises(pm.PepsError) as ei:
        pm.contract_rows(8)
    assert ei.value.code == 3 and "HIP device" in str(ei.value)


def test_host_side_errors_before_device():
    # These never reach the device: N=0, N > 42, N > 21.
    r = pm.contract_rows(0)
    assert r.count.value == 1 and r.peak_states == 1 and r.layer_count == 0
    with pytest.raises(pm.PepsError, match="supports N <= 42"):
        pm.contract_rows(43)
    with pytest.raises(pm.PepsError, match="N <= 21"):
        pm.contract_rows(22)
    with pytest.raises(pm.PepsError, match="outside the board"):
        pm.count_prefix(8, [8])
    assert pm.count_prefix(8, [0, 1]) == 0            # attacking prefix: no device call
    assert pm.count_prefix(4, [1, 3, 0, 2]) == 1      # complete prefix


def _cli(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_cli_grammar_and_errors():
    assert os.path.exists(CLI)
    r = _cli()
    assert r.returncode == 1 and r.stderr.endswith("error: missing or invalid command\n")
    assert "solve N [--layers]" in r.stderr
    r = _cli("solve", "43")
    assert r.returncode == 1 and r.stderr == "error: the packed u128 virtual-boundary backend supports N <= 42\n"
    r = _cli("solve", "x")
    assert r.returncode == 1 and r.stderr == "error: invalid N\n"
    r = _cli("bench", "4", "--repeats", "0")
    assert r.returncode == 1 and r.stderr == "error: --repeats must be positive\n"
    r = _cli("bench", "4", "--min", "5")
    assert r.returncode == 1 and r.stderr == "error: --min must not exceed MAX_N\n"
    r = _cli("bench", "4", "--bogus")
    assert r.returncode == 1 and r.stderr == "error: unknown option: --bogus\n"
    r = _cli("solve", "0")
    assert r.returncode == 0
    assert r.stdout.startswith("N=0 Q(N)=1 elapsed_s=0.000000 peak_rss_bytes=")
    assert r.stdout.rstrip().endswith(
        "peak_states=1 tensor_entries_examined=0 tensor_entries_matched=0 "
        "row_operator_candidates=0 row_operator_matched=0 verified=true")


def test_ctypes_layouts_match_the_c_header(tmp_path):
    """Every field offset of the ctypes mirrors equals the C compiler's
    offsetof for the same struct in include/peps_mi355x.h."""
    from peps_mi355x.abi import ContractionResult, LayerMetric, Options, RequestResult, ServerStats
    structs = {"peps_options": Options, "peps_contraction_result": ContractionResult,
               "peps_layer_metric": LayerMetric, "peps_server_stats": ServerStats,
               "peps_request_result": RequestResult}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "peps_mi355x.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        cname, fname, v = line.split()
        got[(cname, fname)] = int(v)
    for cname, cls in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
