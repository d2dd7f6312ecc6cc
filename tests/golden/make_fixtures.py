"""Extract the reference's published per-N and per-layer numbers into
tests/golden/reference_fixtures.json.

Run in the build container (it reads the read-only reference checkout):
    python tests/golden/make_fixtures.py /root/reference

Only data is copied: counters and counts from the reference's benchmark CSVs,
each value tagged with the CSV file and line it came from.  Nothing from the
reference is imported or executed.
"""
import csv
import json
import os
import sys


def rows(path):
    with open(path, newline="") as f:
        reader = csv.DictReader(f)
        for lineno, row in enumerate(reader, start=2):
            yield lineno, row


def main(ref):
    bench = os.path.join(ref, "benchmarks")
    out = {"source": "reference benchmarks/*.csv (data only)"}

    # Per-layer reachable supports of the orbit-sliced default path (N=10..14).
    layers = {}
    for ln, r in rows(os.path.join(bench, "e35_frontier_audit_layers.csv")):
        layers.setdefault(r["N"], []).append(
            {"row": int(r["row"]), "reachable_states": int(r["reachable_states"]),
             "src": f"benchmarks/e35_frontier_audit_layers.csv:{ln}"})
    out["orbit_layer_supports"] = layers

    # Totals of the orbit path (forward transitions = completed terms) N=10..14.
    out["orbit_audit_totals"] = [
        {"n": int(r["N"]), "count": int(r["count"]),
         "peak_states": int(r["peak_reachable_states"]),
         "completed_row_terms": int(r["forward_transitions"]),
         "src": f"benchmarks/e35_frontier_audit_release.csv:{ln}"}
        for ln, r in rows(os.path.join(bench, "e35_frontier_audit_release.csv"))]

    # Default-path and dense (no orbit) totals N=13..15 (serial rows only).
    keep = {"d4_orbit_serial_sort_reduce": "orbit", "dense_serial_sort_reduce": "dense",
            "d4_sparse_serial_sort_reduce": "orbit_sparse",
            "sparse_serial_sort_reduce": "dense_sparse"}
    e12 = []
    for ln, r in rows(os.path.join(bench, "e12_d4_orbit_slicing_release.csv")):
        if r["backend"] in keep:
            e12.append({"mode": keep[r["backend"]], "n": int(r["N"]), "count": int(r["count"]),
                        "peak_states": int(r["peak_support"]),
                        "row_operator_candidates": int(r["row_operator_candidates"]),
                        "row_operator_matched": int(r["row_operator_matched"]),
                        "src": f"benchmarks/e12_d4_orbit_slicing_release.csv:{ln}"})
    out["e12_totals"] = e12

    # Dense compiled (no orbit) counters N=10..15.
    out["dense_totals"] = [
        {"n": int(r["N"]), "count": int(r["count"]), "peak_states": int(r["peak_support"]),
         "row_operator_candidates": int(r["row_operator_candidates"]),
         "row_operator_matched": int(r["row_operator_matched"]),
         "src": f"benchmarks/current_scaling_peps_release.csv:{ln}"}
        for ln, r in rows(os.path.join(bench, "current_scaling_peps_release.csv"))
        if r["backend"] == "serial_sort_reduce"]

    # N=16 orbit path: peak support and transitions.
    n16 = []
    for name in ("e32_u64_promotion_release.csv", "e36_joint_u64_release.csv"):
        for ln, r in rows(os.path.join(bench, name)):
            if r["N"] == "16":
                n16.append({"n": 16, "count": int(r["count"]),
                            "peak_states": int(r["peak_support"]),
                            "row_operator_matched": int(r["row_operator_matched"]),
                            "src": f"benchmarks/{name}:{ln}"})
    out["orbit_n16"] = n16

    # Unsymmetrized sitewise per-layer tuple at N=14 (pre-index revision counters).
    out["n14_sitewise_layers"] = [
        {"row": int(r["row"]), "input_states": int(r["input_states"]),
         "tensor_entries_examined": int(r["tensor_entries_examined"]),
         "tensor_entries_matched": int(r["tensor_entries_matched"]),
         "completed_row_terms": int(r["completed_row_terms"]),
         "output_states": int(r["output_states"]), "output_weight": int(r["output_weight"]),
         "src": f"benchmarks/n14_layers.csv:{ln}"}
        for ln, r in rows(os.path.join(bench, "n14_layers.csv"))]

    # Naive sitewise totals N=4..14 (pre-index revision counters).
    out["naive_totals"] = [
        {"n": int(r["N"]), "count": int(r["count"]), "peak_states": int(r["peak_states"]),
         "tensor_entries_examined": int(r["tensor_entries_examined"]),
         "tensor_entries_matched": int(r["tensor_entries_matched"]),
         "src": f"benchmarks/naive_release.csv:{ln}"}
        for ln, r in rows(os.path.join(bench, "naive_release.csv"))]

    # Naive HashMap peaks and matched entries N=1..16 (SCNet same-node runs).
    out["naive_scnet"] = [
        {"n": int(r["N"]), "count": int(r["count"]),
         "peak_states": int(r["peak_sparse_support"]),
         "tensor_entries_matched": int(r["local_tensor_entries_accepted"]),
         "src": f"benchmarks/issue34_same_node_algorithm_comparison_scnet.csv:{ln}"}
        for ln, r in rows(os.path.join(bench, "issue34_same_node_algorithm_comparison_scnet.csv"))
        if r["algorithm_key"] == "naive_peps"]

    # CRT: residues of Q(N) and the prime plans.
    out["crt_release"] = [
        {"n": int(r["N"]), "count": int(r["count"]), "primes": [int(p) for p in r["primes"].split("|")],
         "modulus_product": int(r["modulus_product"]), "factorial_bound": int(r["factorial_bound"]),
         "residues": [int(x) for x in r["residues"].split("|")],
         "src": f"benchmarks/e45_wide_crt_release.csv:{ln}"}
        for ln, r in rows(os.path.join(bench, "e45_wide_crt_release.csv"))]
    out["crt_plans"] = [
        {"n": int(r["N"]), "factorial_bound": int(r["factorial_bound"]),
         "primes": [int(p) for p in r["primes"].split("|")],
         "modulus_product": int(r["modulus_product"]),
         "src": f"benchmarks/e45_wide_crt_plan_n22_n28.csv:{ln}"}
        for ln, r in rows(os.path.join(bench, "e45_wide_crt_plan_n22_n28.csv"))]

    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
