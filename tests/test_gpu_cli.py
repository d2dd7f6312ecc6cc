"""Output-format parity of the nqueens-peps CLI with the reference CLI
(src/main.rs:25-152): the solve line, the --layers CSV and the bench CSV are
byte-for-byte in the reference's formats, with every counter equal to the
oracle's.  The bench CSV is read with the columns the reference's analyzer
uses (experiments/e51_truncated_boundary_mps/scripts/analyze_algorithm_comparison.py:225-252)."""
import csv
import io
import os
import re
import subprocess

import pytest

import _oracle as o

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "quantum-harness-peps-naive_amd",
                   "nqueens-peps")
LAYER_HEADER = ("row,input_states,tensor_entries_examined,tensor_entries_matched,row_operator_candidates,"
                "row_operator_matched,completed_row_terms,output_states,output_weight,elapsed_s,peak_rss_bytes")
BENCH_HEADER = ("N,count,known_count,verified,median_elapsed_s,min_elapsed_s,peak_rss_bytes,peak_states,"
                "tensor_entries_examined,tensor_entries_matched,row_operator_candidates,row_operator_matched,"
                "repeats")


def run(*args):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_solve_line_and_layers_match_reference_format():
    for n in (8, 12):
        out = run("solve", str(n), "--layers").splitlines()
        ref = o.contract_rows(n)
        m = re.fullmatch(r"N=(\d+) Q\(N\)=(\d+) elapsed_s=\d+\.\d{6} peak_rss_bytes=\d+ peak_states=(\d+) "
                         r"tensor_entries_examined=(\d+) tensor_entries_matched=(\d+) "
                         r"row_operator_candidates=(\d+) row_operator_matched=(\d+) verified=(true|false)", out[0])
        assert m, out[0]
        assert tuple(int(x) for x in m.groups()[:7]) == (
            n, ref.count.value, ref.peak_states, ref.tensor_entries_examined.value,
            ref.tensor_entries_matched.value, ref.row_operator_candidates.value, ref.row_operator_matched.value)
        assert m.group(8) == "true"
        assert out[1] == LAYER_HEADER
        rows = out[2:]
        assert len(rows) == n
        for line, l in zip(rows, ref.layer_list()):
            f = line.split(",")
            assert len(f) == 11 and re.fullmatch(r"\d+\.\d{6}", f[9]), line
            assert [int(x) for x in f[:9]] == [l["row"] + 1, l["input_states"], l["tensor_entries_examined"],
                                               l["tensor_entries_matched"], l["row_operator_candidates"],
                                               l["row_operator_matched"], l["completed_row_terms"],
                                               l["output_states"], l["output_weight"]]


def test_bench_csv_matches_reference_format_and_analyzer_columns():
    out = run("bench", "11", "--min", "8", "--repeats", "2", "--csv")
    lines = out.splitlines()
    assert lines[0] == BENCH_HEADER
    rows = list(csv.DictReader(io.StringIO(out)))
    assert [int(r["N"]) for r in rows] == [8, 9, 10, 11]
    for r in rows:
        n = int(r["N"])
        ref = o.contract_rows(n)
        assert int(r["count"]) == int(r["known_count"]) == ref.count.value
        assert r["verified"] == "true" and r["repeats"] == "2"
        assert re.fullmatch(r"\d+\.\d{9}", r["median_elapsed_s"]) and re.fullmatch(r"\d+\.\d{9}", r["min_elapsed_s"])
        assert float(r["min_elapsed_s"]) <= float(r["median_elapsed_s"])
        assert int(r["peak_states"]) == ref.peak_states
        assert (int(r["tensor_entries_examined"]), int(r["tensor_entries_matched"])) == (17, 17)
        assert int(r["row_operator_candidates"]) == ref.row_operator_candidates.value
        assert int(r["row_operator_matched"]) == ref.row_operator_matched.value
        assert int(r["peak_rss_bytes"]) > 0


def test_bench_text_lines_match_reference_format():
    out = run("bench", "9", "--min", "8", "--repeats", "1").splitlines()
    # "N={n:>2} Q(N)={:<10} median={:.6}s peak_rss={:.1}MiB peak_states={:<8} verified={}"
    assert re.fullmatch(r"N= 8 Q\(N\)=92         median=\d+\.\d{6}s peak_rss=\d+\.\dMiB peak_states=\d+ +"
                        r"verified=true", out[0]), out[0]
    assert re.fullmatch(r"N= 9 Q\(N\)=352        median=\d+\.\d{6}s peak_rss=\d+\.\dMiB peak_states=\d+ +"
                        r"verified=true", out[1]), out[1]
    assert len(out[0].split("peak_states=")[1].split(" verified")[0]) >= 8
