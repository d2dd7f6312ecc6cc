"""Writes tests/golden/orbit_layers_n17_n18.json: the per-row figures of the
reference's default row loop (contract_rows, src/lib.rs:757-760 ->
contract_rows_sort_reduce_with_modes, :5287-5415, counters :5370-5390) for
N = 17 and 18, where the sort-reduce restatement needs more memory than a
test can spend.

Each row's input_states, row_operator_matched, completed_row_terms and
output_states come from the oracle's column-set-grouped census, and its
output_weight from the oracle's per-depth prefix DFS (oracle/layer_census.cpp);
row_operator_candidates is N x input_states (src/lib.rs:5373).  Both routes
are pinned in tests/test_oracle_census.py against the reference's published
layer supports and totals through N = 16.

Usage: python tests/golden/make_layer_fixtures.py [--threads T] [--n 17 18]
(N = 18 needs about 40 GB of host memory and a few minutes on 8 threads.)
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "quantum-harness-peps-naive_amd"))

import _oracle as o  # noqa: E402

OUT = os.path.join(HERE, "orbit_layers_n17_n18.json")


def layers(n, threads):
    t0 = time.time()
    cen = o.frontier_census(n, orbits=True, threads=threads)
    t1 = time.time()
    dep = o.depth_counts(n, orbits=True, threads=threads)
    t2 = time.time()
    print(f"N={n}: census {t1 - t0:.1f}s, depth counts {t2 - t1:.1f}s", flush=True)
    rows = []
    for c in cen:
        r = c["row"]
        rows.append({"row": r, "input_states": c["input_states"],
                     "row_operator_candidates": n * c["input_states"],
                     "row_operator_matched": c["row_operator_matched"],
                     "completed_row_terms": c["completed_row_terms"],
                     "output_states": c["output_states"], "output_weight": dep[r + 1]})
    return {"n": n, "count": dep[n], "peak_states": max(c["output_states"] for c in cen),
            "row_operator_candidates": sum(x["row_operator_candidates"] for x in rows),
            "row_operator_matched": sum(x["row_operator_matched"] for x in rows), "layers": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--n", type=int, nargs="+", default=[17, 18])
    args = ap.parse_args()
    data = {"source": ("oracle/layer_census.cpp: frontier_census (input_states, row_operator_matched, "
                       "completed_row_terms, output_states) and depth_counts (output_weight), row-0 orbit "
                       "slicing; row_operator_candidates = N x input_states; written by "
                       "tests/golden/make_layer_fixtures.py"),
            "boards": {}}
    if os.path.exists(OUT):
        data["boards"].update(json.load(open(OUT))["boards"])
    for n in args.n:
        b = layers(n, args.threads)
        assert b["count"] == o.known_count(n), n
        data["boards"][str(n)] = b
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
            f.write("\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
