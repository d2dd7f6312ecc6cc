"""Summarise rocprofv3 CSV output of the contraction into a short text file.

    python -m peps_mi355x.profile_summary --title T --command C \\
        --kernel-stats kt_kernel_stats.csv \\
        [--fetch pmc_fetch.csv --write pmc_write.csv --copy-fetch c1.csv --copy-write c2.csv \\
         --copy-bytes B] > summary.txt

Kernel names are shortened to their template heads.  When the FETCH_SIZE /
WRITE_SIZE runs are given, each kernel's per-dispatch HBM traffic is reported
after scaling by the factor the same counter reads for the copy probe's known
bytes.  Lines starting with "traffic " are read back by bench.py.
"""
import argparse
import collections
import csv
import re


def short(name: str) -> str:
    name = name.replace("void ", "").replace("peps::(anonymous namespace)::", "")
    m = re.match(r"([A-Za-z_0-9]+(?:<[^>]*>)?)", name)
    return m.group(1) if m else name[:40]


def kernel_stats(path):
    rows = []
    for r in csv.DictReader(open(path)):
        rows.append((short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"])))
    return rows


def counter_totals(path, counter):
    tot = collections.defaultdict(float)
    calls = collections.defaultdict(set)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        k = short(r["Kernel_Name"])
        tot[k] += float(r["Counter_Value"])
        calls[k].add(r["Dispatch_Id"])
    return {k: (tot[k], len(calls[k])) for k in tot}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--title", required=True)
    ap.add_argument("--command", required=True)
    ap.add_argument("--kernel-stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=float)
    a = ap.parse_args()

    print(f"# {a.title}")
    print(f"# command: {a.command}")
    print("#")
    print(f"# {'kernel':40s} {'calls':>6s} {'total_ms':>10s} {'avg_us':>10s}")
    for k, calls, tot, avg in sorted(kernel_stats(a.kernel_stats), key=lambda x: -x[2]):
        print(f"  {k:40s} {calls:6d} {tot / 1e6:10.3f} {avg / 1e3:10.2f}")
    if a.fetch and a.write and a.copy_fetch and a.copy_write and a.copy_bytes:
        cf = counter_totals(a.copy_fetch, "FETCH_SIZE")
        cw = counter_totals(a.copy_write, "WRITE_SIZE")
        kf = [v for k, v in cf.items() if "k_copy" in k][0]
        kw = [v for k, v in cw.items() if "k_copy" in k][0]
        # counters are in KiB; factor = counted bytes / true bytes of the copy
        f_fac = kf[0] * 1024.0 / kf[1] / a.copy_bytes
        w_fac = kw[0] * 1024.0 / kw[1] / a.copy_bytes
        print("#")
        print(f"# copy probe calibration: FETCH_SIZE reads {f_fac:.4f} x true bytes, "
              f"WRITE_SIZE {w_fac:.4f} x true bytes")
        print(f"# {'kernel':40s} {'dispatches':>10s} {'fetch_MB/disp':>14s} {'write_MB/disp':>14s}")
        fe = counter_totals(a.fetch, "FETCH_SIZE")
        wr = counter_totals(a.write, "WRITE_SIZE")
        for k in sorted(set(fe) | set(wr)):
            f = fe.get(k, (0.0, 1))
            w = wr.get(k, (0.0, 1))
            fb = f[0] * 1024.0 / f_fac / max(1, f[1])
            wb = w[0] * 1024.0 / w_fac / max(1, w[1])
            print(f"  {k:40s} {max(f[1], w[1]):10d} {fb / 1e6:14.2f} {wb / 1e6:14.2f}")
            print(f"traffic {k} {fb + wb:.6e}")


if __name__ == "__main__":
    main()
