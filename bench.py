"""Benchmark: exact N-Queens count by the explicit-C PEPS row contraction on
one MI355X.

A step is one full count at N (default 18): every row of the boundary
contraction from the empty board to the final v1/v2 sum, with the reference's
row-0 orbit slicing, on device-resident frontiers.  `value` is
state-transitions per second, where transitions are the reference's
`row_operator_matched` counter for that count (src/lib.rs:5370-5379) and the
time is the wall time per count with the device warmed.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 18]
                       [--form plain|batched] [--no-cpu-baseline] [--no-serving]
       python bench.py --form serving [--requests 2000] [--rates 50,200,800]
                       [--serve-n-min 10] [--serve-n-max 16] [--policy response_ratio]
       python bench.py --form cpu [--n 18] [--task-range a:b]   (CPU restatement only)

The plain form (the default) also streams a bounded mixed-N request set
(300 requests, N uniform in [10, 18], 0..2 placed rows, saturating arrivals)
through the request server after the timed region and adds its p50/p99
latency, requests/s and the ratio to back-to-back plain counts as extra keys;
then a second stream of 2000 small requests (N uniform in [10, 16]), whose
figures carry the suffix _small.

The serving form streams requests (N, random non-attacking prefix of 0..2
rows) with Poisson arrivals at each rate through the request server and
reports requests/s and p50/p90/p99 latency per rate.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "quantum-harness-peps-naive_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure

# rocprofv3 kernel names of the engine's profiling labels, as patterns over
# every instantiation a label covers (L: CRT lanes of the record, 0 for
# packed 12-byte records; the digit-width and table arguments vary by row).
ROCPROF_PATTERNS = {
    "merge_count": r"k_merge_expand<{L}, 0, \d+, \d+, \d+>",
    "merge_scatter": r"k_merge_expand<{L}, 1, \d+, \d+, \d+>",
    "merge_last": r"k_merge_expand<{L}, 2, \d+, \d+, \d+>",
    "count_ub": r"k_count_ub<{L}>",
    "caps": r"k_caps_(scan|expand)",
    "part_hist": r"k_part_hist<{L}, \d+, false>",
    "part_hist_runs": r"k_part_hist<{L}, \d+, true>",
    "part_scatter": r"k_part_scatter<{L}, \d+, false, (true|false), (true|false)>",
    "part_scatter_runs": r"k_part_scatter<{L}, \d+, true, (true|false), (true|false)>",
}


def record_tag(result):
    """Template argument L of the record type a run used."""
    return 0 if result.record_bytes == 12 else result.crt_lanes


def measured_traffic(n, form, kernel, lanes):
    """HBM bytes per launch of the engine label `kernel` from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the same workload: the
    call-weighted mean over every instantiation the label covers (None if
    the summary or the label is absent)."""
    import re

    path = os.path.join(ROOT, "profiles", f"kernels_n{n}_{form}.txt")
    pat = ROCPROF_PATTERNS.get(kernel)
    if not pat or not os.path.exists(path):
        return None
    rx = re.compile(pat.format(L=lanes) + "$")
    calls, traffic = {}, {}
    row = re.compile(r"^\s+(\S.*?)\s+(\d+)\s+([\d.]+)\s+([\d.]+)\s*$")
    for line in open(path):
        if line.startswith("traffic "):
            k, v = line[len("traffic "):].rstrip("\n").rsplit(" ", 1)
            traffic[k] = float(v)
        elif not line.startswith("#"):
            m = row.match(line)
            if m and m.group(1) not in calls:
                calls[m.group(1)] = int(m.group(2))
    hit = [(calls[k], traffic[k]) for k in traffic if rx.match(k) and calls.get(k)]
    if not hit:
        return None
    return sum(c * t for c, t in hit) / sum(c for c, _ in hit)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(n, budget_s):
    """The oracle's restatement of the reference's rayon path
    (contract_rows_d4_orbit_parallel_sort_reduce, src/lib.rs:5469-5626) run on
    a bounded sample: whole 2-row prefix tasks of the same board, stratified
    over the first-row column (centre and edges first, then the columns
    between), timed until the budget is spent.  Rate = transitions
    (row_operator_matched) / s over the sampled tasks; the whole CPU count's
    time is estimated as the sum of every task's transitions at that rate."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle

    threads = min(host_threads(), 64)
    # the 2-row prefix tasks of the reference's task builder (row 0 left half)
    tasks = [(c0, c1) for c0 in range((n + 1) // 2) for c1 in range(n)
             if c1 != c0 and abs(c1 - c0) != 1]
    half = (n + 1) // 2
    # first-row columns in van der Corput order over [0, half): the sample
    # spreads from the centre and the edge towards the columns in between
    def vdc(i):
        x, f = 0.0, 0.5
        while i:
            x += f * (i & 1)
            i >>= 1
            f /= 2
        return x
    c0_order = [half - 1, 0]
    for i in range(4 * half):
        c = min(half - 1, int(round(vdc(i) * (half - 1)))) if i else half - 1
        if c not in c0_order:
            c0_order.append(c)
    c0_order += [c for c in range(half) if c not in c0_order]
    order = []
    for c0 in c0_order:
        legal = [t for t in tasks if t[0] == c0]
        order.append(legal[len(legal) // 2])
    done, matched, t_total = [], 0, 0.0
    for cand in order:
        t0 = time.perf_counter()
        r = _oracle.contract_prefix(n, list(cand), threads=threads)
        dt = time.perf_counter() - t0
        t_total += dt
        matched += r.row_operator_matched.value
        done.append({"prefix": list(cand), "seconds": round(dt, 3),
                     "transitions": r.row_operator_matched.value})
        log(f"cpu baseline task {cand}: {dt:.2f}s matched={r.row_operator_matched.value}")
        if t_total >= budget_s:
            break
    return {
        "value": matched / t_total if t_total else None,
        "unit": "transitions/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{len(done)} of {len(tasks)} two-row prefix tasks of N={n}, one per first-row column "
                   f"in the order {[d['prefix'][0] for d in done]} (the median-index second-row column "
                   f"of each), through the oracle's restatement of the reference's rayon sort-reduce "
                   f"(checked u128 records, {threads} threads, {host_cpu_model()}); rate = their "
                   f"transitions / their time"),
        "tasks": done,
        "seconds": t_total,
    }


def cpu_full(args):
    """The oracle's restatement of the reference's rayon path over every
    2-row prefix task of N (or the range --task-range a:b of them), in task
    order, timed per task: the full CPU count that the sampled cpu_baseline
    extrapolates.  Prints one JSON line."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle

    n = args.n
    threads = min(host_threads(), 64)
    tasks = [(c0, c1) for c0 in range((n + 1) // 2) for c1 in range(n)
             if c1 != c0 and abs(c1 - c0) != 1]
    a, b = 0, len(tasks)
    if args.task_range:
        a, b = (int(x) for x in args.task_range.split(":"))
    count, matched, secs, per = 0, 0, 0.0, []
    for i in range(a, b):
        t0 = time.perf_counter()
        r = _oracle.contract_prefix(n, list(tasks[i]), threads=threads)
        dt = time.perf_counter() - t0
        w = 1 if tasks[i][0] == n - 1 - tasks[i][0] else 2
        count += w * r.count.value
        matched += r.row_operator_matched.value
        secs += dt
        per.append(round(dt, 3))
        log(f"cpu task {i} {tasks[i]}: {dt:.2f}s (total {secs:.1f}s)")
    print(json.dumps({"form": "cpu", "n": n, "tasks": [a, b], "of_tasks": len(tasks), "seconds": secs,
                      "weighted_count": count, "transitions": matched, "threads": threads,
                      "cpu": host_cpu_model(), "per_task_s": per}), flush=True)


def serving_leg(pm, requests, n_min, n_max, seed, rate):
    """A bounded mixed-N request stream through the request server after the
    timed region: open-loop arrivals at `rate` (saturating by default: the
    throughput), the same requests again paced at half the throughput that
    stream reached (the latency a caller sees below capacity), then the same
    requests as back-to-back plain GPU counts.  Every served count must equal
    its back-to-back count (and known_count for an empty prefix)."""
    from peps_mi355x import serving as sv

    reqs = sv.make_requests(seed, requests, rate, n_min, n_max)
    with pm.Server() as server:
        sv.run_open_loop(server, sv.make_requests(seed + 1, 40, 1e6, n_min, min(n_max, 14)))  # warm
        st0 = server.stats()
        outs = sv.run_open_loop(server, reqs, timeout_s=600,
                                progress=lambda stage, d, t: log(f"serving: {stage} {d}/{t}"))
        st1 = server.stats()
        summ = sv.summarize(outs)
        paced_rate = 0.5 * summ["requests_per_s"]
        paced_reqs = sv.make_requests(seed, requests, paced_rate, n_min, n_max)
        paced_outs = sv.run_open_loop(server, paced_reqs, timeout_s=600,
                                      progress=lambda stage, d, t: log(f"serving (paced): {stage} {d}/{t}"))
        paced = sv.summarize(paced_outs)
    t0 = time.perf_counter()
    alone = []
    for i, (_, n, cols) in enumerate(reqs):
        alone.append(pm.count_prefix(n, cols))
        if (i + 1) % 100 == 0:
            log(f"back-to-back: {i + 1}/{len(reqs)}")
    b2b_s = time.perf_counter() - t0
    bad = [(n, c) for (_, n, c), o, p, a in zip(reqs, outs, paced_outs, alone)
           if o.status != 0 or p.status != 0 or o.count != a or p.count != a
           or (not c and a != pm.known_count(n))]
    if bad:
        raise SystemExit(f"serving: wrong or failed requests: {bad[:5]}")
    b2b_rate = len(reqs) / b2b_s
    return {
        "requests": len(reqs),
        "n_range": [n_min, n_max],
        "prefix_rows": [0, 2],
        "offered_rate": rate,
        "seed": seed,
        "requests_per_s": summ["requests_per_s"],
        "p50_s": summ["p50_s"],
        "p90_s": summ["p90_s"],
        "p99_s": summ["p99_s"],
        "max_s": summ["max_s"],
        "percentile_rule": "sorted[(len - 1) * q] (src/bin/dfs_bitmask.rs:70-73)",
        "launches": st1["launches"] - st0["launches"],
        "paced": {"offered_rate": paced_rate, "requests_per_s": paced["requests_per_s"],
                  "p50_s": paced["p50_s"], "p90_s": paced["p90_s"], "p99_s": paced["p99_s"],
                  "max_s": paced["max_s"]},
        "back_to_back_requests_per_s": b2b_rate,
        "vs_back_to_back": summ["requests_per_s"] / b2b_rate,
        "checked": "every count equals the same request as a plain GPU count; empty prefixes equal known_count",
    }


def serve(args, pm):
    """Open-loop request streams through the request server, one per rate."""
    from peps_mi355x import serving as sv
    from peps_mi355x.abi import PEPS_E_MEMORY

    info = pm.device_info(0)
    log(f"device: {info['name']} CUs={info['cu_count']}")
    sweep = []
    with pm.Server(policy=args.policy) as server:
        # warm the engine (arena, code objects) on a short stream
        warm = sv.make_requests(args.seed + 1, 50, 1e6, args.serve_n_min, args.serve_n_max)
        sv.run_open_loop(server, warm)
        for rate in [float(x) for x in args.rates.split(",")]:
            reqs = sv.make_requests(args.seed, args.requests, rate, args.serve_n_min, args.serve_n_max)
            st0 = server.stats()
            outs = sv.run_open_loop(server, reqs, timeout_s=1800, every=args.progress_every,
                                    progress=lambda stage, d, t: log(f"rate {rate:g}: {stage} {d}/{t}"))
            st1 = server.stats()
            bad = [(n, c, o.count) for (_, n, c), o in zip(reqs, outs)
                   if o.status != 0 or (not c and o.count != pm.known_count(n))]
            if bad:
                raise SystemExit(f"serving: wrong or failed requests: {bad[:5]}")
            summ = sv.summarize(outs)
            summ.update({"offered_rate": rate, "launches": st1["launches"] - st0["launches"],
                         "busy_s": st1["busy_s"] - st0["busy_s"]})
            log(json.dumps(summ))
            sweep.append(summ)
        # The same requests back to back, one plain GPU count each (no batching).
        reqs = sv.make_requests(args.seed, args.requests, 1.0, args.serve_n_min, args.serve_n_max)

        def plain(n, cols):
            # A prefix whose frontier outgrows the arena is counted as the sum
            # of its one-row extensions, each a plain count of its own.
            try:
                return pm.count_prefix(n, cols)
            except pm.PepsError as e:
                if e.code != PEPS_E_MEMORY:
                    raise
            taken = set(cols)
            r = len(cols)
            return sum(plain(n, cols + [c]) for c in range(n)
                       if c not in taken and all(abs(c - p) != r - i for i, p in enumerate(cols)))

        t0 = time.perf_counter()
        for i, (_, n, cols) in enumerate(reqs):
            if not cols and n >= 19:
                pm.count_batched(n, 2)  # the empty board's plain frontier does not fit one GPU
            else:
                plain(n, list(cols))
            if (i + 1) % args.progress_every == 0 or n >= 19:
                log(f"back-to-back: {i + 1}/{len(reqs)} (N={n})")
        b2b_s = time.perf_counter() - t0
    best = max(sweep, key=lambda x: x["requests_per_s"])
    b2b_rate = args.requests / b2b_s
    out = {
        "metric": "requests/s",
        "value": best["requests_per_s"],
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": len(sweep),
        "warmup": 1,
        "ms_per_step": 1000.0 * best["span_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": (f"synthetic: {args.requests} requests per rate from splitmix64 seed {args.seed}, "
                 f"Poisson arrivals, N uniform in [{args.serve_n_min}, {args.serve_n_max}], "
                 f"random non-attacking prefix of 0..2 rows"),
        "config": {"workload": "exact prefix-completion counts through the request server",
                   "requests": args.requests, "rates": args.rates, "policy": args.policy,
                   "n_range": [args.serve_n_min, args.serve_n_max]},
        "p50_s": best["p50_s"],
        "p99_s": best["p99_s"],
        "back_to_back_requests_per_s": b2b_rate,
        "vs_back_to_back": best["requests_per_s"] / b2b_rate,
        "sweep": sweep,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=18)
    ap.add_argument("--form", choices=("plain", "batched", "serving", "cpu"), default="plain")
    ap.add_argument("--task-range", default="", help="--form cpu: tasks a:b of the 2-row prefix tasks")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rates", default="100,400,1600,100000")
    ap.add_argument("--serve-n-min", type=int, default=10)
    ap.add_argument("--serve-n-max", type=int, default=16)
    ap.add_argument("--progress-every", type=int, default=250,
                    help="serving: report every this many submitted / finished requests")
    ap.add_argument("--policy", default="response_ratio", choices=("shortest", "fifo", "response_ratio"),
                    help="request server scheduling policy")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=0, help="CRT lanes (0: the plan for N)")
    ap.add_argument("--graph", action="store_true",
                    help="record the size plan on the warmup and replay each timed step as HIP graphs")
    ap.add_argument("--autotune", action="store_true",
                    help="sweep the launch shape first (bounded by --autotune-seconds) and time the best")
    ap.add_argument("--autotune-seconds", type=float, default=240.0)
    ap.add_argument("--depth", type=int, default=2)
    ap.add_argument("--options", default="",
                    help='engine options as JSON, e.g. \'{"merge_passes": 2}\' (PepsOptions fields)')
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--profile-steps", type=int, default=1,
                    help="untimed steps with per-kernel HIP-event timing after the timed region")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the mixed-N serving leg of the plain form")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-small-requests", type=int, default=2000,
                    help="plain form: a second serving leg of small requests (0: off)")
    ap.add_argument("--serve-small-n-max", type=int, default=16)
    ap.add_argument("--serve-leg-n-max", type=int, default=18)
    ap.add_argument("--serve-rate", type=float, default=1e5,
                    help="offered request rate of the serving leg (default: saturating)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: the contraction runs on one MI355X "
              f"and only the one-GPU throughput is measured",
              file=sys.stderr)
        sys.exit(2)

    if args.form == "cpu":
        return cpu_full(args)

    import peps_mi355x as pm

    if args.form == "serving":
        return serve(args, pm)
    n = args.n
    info = pm.device_info(0)
    log(f"device: {info['name']} CUs={info['cu_count']} HBM={info['hbm_bytes'] / 2**30:.0f} GiB")

    opts = None
    if args.autotune:
        from peps_mi355x import autotune as at
        opts, table = at.autotune(n, form=args.form, depth=args.depth, repeats=1,
                                  budget_s=args.autotune_seconds,
                                  progress=lambda sh, t: log(f"autotune {sh}: {t}"))
        log(f"autotune best: {opts}")
    if args.options:
        opts = opts or pm.PepsOptions()
        for k, v in json.loads(args.options).items():
            if not hasattr(opts, k):
                raise SystemExit(f"unknown engine option {k}")
            setattr(opts, k, v)
    if args.lanes or args.graph:
        opts = opts or pm.PepsOptions()
        opts.crt_lanes = args.lanes
        opts.graph_mode = 1 if args.graph else 0

    if args.form == "batched":
        # Progress per task group on stderr (long counts stay visibly alive).
        pm.set_progress(lambda done, total: log(f"batched: {done}/{total} tasks"))

    def step():
        if args.form == "plain":
            r = pm.contract_rows(n, opts)
        else:
            r, _ = pm.count_batched(n, args.depth, opts)
        if r.count.value != pm.known_count(n):
            raise SystemExit(f"count mismatch at N={n}: {r.count.value}")
        return r

    for i in range(args.warmup):
        t0 = time.perf_counter()
        r = step()
        log(f"warmup {i}: {time.perf_counter() - t0:.3f}s count={r.count.value}")

    # Transitions are the reference's counter for the plain (merged) path.
    plain_matched = None
    if args.form == "plain":
        plain_matched = r.row_operator_matched.value if args.warmup else None

    # Timed steps run without kernel timing (no HIP events between launches).
    times, last = [], None
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        last = step()
        times.append(time.perf_counter() - t0)
        log(f"step {i}: {times[-1]:.3f}s kernel={last.kernel_s:.3f}s")
    t_all = time.perf_counter() - t_all
    # One more step with HIP-event timing of every launch, for the per-kernel
    # figures and the roofline (outside the timed region).
    pm.profile_kernels(True)
    prof_steps = max(1, args.profile_steps)
    for _ in range(prof_steps):
        step()
    stats = pm.kernel_stats()
    pm.profile_kernels(False)
    if plain_matched is None:
        plain_matched = last.row_operator_matched.value if args.form == "plain" else None
    transitions_source = "plain count (the reference's contract_rows counter)"
    if plain_matched is None:
        try:
            plain_matched = pm.contract_rows(n).row_operator_matched.value
        except pm.PepsError as e:
            # The single merged frontier does not fit one GPU (N >= 19): use
            # the sliced count's own transitions and say so.
            log(f"plain count unavailable ({e}); transitions from the {args.form} form")
            plain_matched = last.row_operator_matched.value
            transitions_source = f"{args.form} count (slices are not merged with each other)"

    ms_per_step = 1000.0 * t_all / args.steps
    value = plain_matched / (t_all / args.steps)

    per_kernel = {}
    for name, s in stats.items():
        gbs = s["algorithmic_bytes"] / s["seconds"] / 1e9 if s["seconds"] else 0.0
        per_kernel[name] = {"launches_per_step": s["launches"] / prof_steps,
                            "ms_per_step": 1000.0 * s["seconds"] / prof_steps,
                            "avg_launch_us": 1e6 * s["seconds"] / max(1, s["launches"]),
                            "algorithmic_GBps": gbs}
    dom = max(stats.items(), key=lambda kv: kv[1]["seconds"]) if stats else None
    roofline = None
    if dom:
        name, s = dom
        achieved = s["algorithmic_bytes"] / s["seconds"] / 1e9
        roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(n, args.form, name, record_tag(last)),
                    "avg_launch_us": round(1e6 * s["seconds"] / s["launches"], 3),
                    "algorithmic_bytes_per_launch": s["algorithmic_bytes"] / s["launches"]}

    serving = None
    if args.form == "plain" and not args.no_serving:
        serving = serving_leg(pm, args.serve_requests, args.serve_n_min, args.serve_leg_n_max, args.seed,
                              args.serve_rate)
        log(f"serving leg: {json.dumps(serving)}")
        if args.serve_small_requests:
            # Requests that are small against the GPU (N <= 16): the regime in
            # which continuous batching overlaps many requests per launch.
            serving["small_requests"] = serving_leg(pm, args.serve_small_requests, args.serve_n_min,
                                                    args.serve_small_n_max, args.seed, args.serve_rate)
            log(f"serving leg (small requests): {json.dumps(serving['small_requests'])}")

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(n, args.cpu_seconds)
        except Exception as e:  # report, do not hide
            cpu = {"value": None, "error": str(e)}

    out = {
        "metric": "state-transitions/s",
        "value": value,
        "unit": "transitions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64" if last.record_bytes == 12 else "u32",
        "data": "synthetic (deterministic: the empty N x N board)",
        "config": {"workload": f"exact N-Queens count, N={n}, "
                               f"{'single merged frontier (plain drop-in)' if args.form == 'plain' else f'{args.depth}-row prefix slices (batched)'}"
                               ", row-0 orbit slicing, "
                               + ("exact u64 weights in packed 12-byte records" if last.record_bytes == 12
                                  else "u32 CRT residue records"),
                   "n": n, "form": args.form, "crt_lanes": last.crt_lanes,
                   "record_bytes": last.record_bytes,
                   "options": ({k: v for k, v in vars(opts).items() if v} if opts else "defaults"),
                   "transitions_per_count": plain_matched, "transitions_source": transitions_source},
        "time_to_count_s": min(times),
        "time_to_count_median_s": sorted(times)[len(times) // 2],
        "kernel_s": last.kernel_s,
        "count": last.count.value,
        "peak_states": last.peak_states,
        "device_peak_bytes": last.device_peak_bytes,
        "per_kernel": per_kernel,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if serving:
        # mixed-N serving after the timed region (extra keys; `value` is the count rate above)
        out.update({"p50_s": serving["p50_s"], "p99_s": serving["p99_s"],
                    "p50_s_paced": serving["paced"]["p50_s"], "p99_s_paced": serving["paced"]["p99_s"],
                    "requests_per_s": serving["requests_per_s"],
                    "vs_back_to_back": serving["vs_back_to_back"], "serving": serving})
        if "small_requests" in serving:
            sm = serving["small_requests"]
            out.update({"requests_per_s_small": sm["requests_per_s"], "p50_s_small": sm["p50_s"],
                        "p99_s_small": sm["p99_s"], "vs_back_to_back_small": sm["vs_back_to_back"]})
    if cpu and cpu.get("value"):
        # The CPU path counts the board as its 2-row prefix tasks, whose
        # frontiers do not merge with each other: the whole job is the sum of
        # the tasks' transitions (the batched GPU form's counter), at the
        # sampled rate.
        try:
            rb, ntasks = pm.count_batched(n, 2)
            job = rb.row_operator_matched.value
            cpu["job_transitions"] = job
            cpu["job_tasks"] = ntasks
            out["cpu_time_to_count_est_s"] = job / cpu["value"]
        except pm.PepsError as e:
            log(f"task transitions unavailable ({e})")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
