"""Open-loop request stream for the request server, and its latency report.

Workload: ``count`` requests from a splitmix64 stream seeded with ``seed``;
Poisson arrivals at ``rate`` requests/s (exponential gaps from the same
stream); N uniform in [n_min, n_max]; prefix length k uniform in [0, k_max];
the prefix places one queen per row 0..k-1, each uniform over the columns the
earlier queens leave free (a row with no free column ends the prefix early).

Percentiles follow the reference's rule ``sorted[(len - 1) * num // den]``
(src/bin/dfs_bitmask.rs:70-73).
"""
import math
import time

MASK64 = (1 << 64) - 1


class SplitMix64:
    """Vigna's splitmix64: state += 0x9e3779b97f4a7c15, then two xor-shift-multiply rounds."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, bound: int) -> int:
        """Uniform integer in [0, bound) (multiply-high; bias below 2^-40 here)."""
        return (self.next() * bound) >> 64

    def unit(self) -> float:
        """Uniform float in (0, 1]."""
        return ((self.next() >> 11) + 1) / float(1 << 53)


def random_prefix(rng: SplitMix64, n: int, k: int):
    cols, dr, dl = 0, 0, 0
    mask = (1 << n) - 1
    out = []
    for _ in range(k):
        free = ~(cols | dr | dl) & mask
        options = [c for c in range(n) if (free >> c) & 1]
        if not options:
            break
        c = options[rng.below(len(options))]
        b = 1 << c
        out.append(c)
        cols |= b
        dr = ((dr | b) << 1) & mask
        dl = (dl | b) >> 1
    return out


def make_requests(seed: int = 0, count: int = 2000, rate: float = 100.0, n_min: int = 10,
                  n_max: int = 20, k_max: int = 2):
    """[(arrival_s, n, prefix_columns)] in arrival order."""
    rng = SplitMix64(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += -math.log(rng.unit()) / rate
        n = n_min + rng.below(n_max - n_min + 1)
        k = rng.below(k_max + 1)
        out.append((t, n, random_prefix(rng, n, k)))
    return out


def percentile(sorted_values, num: int, den: int):
    return sorted_values[(len(sorted_values) - 1) * num // den]


def run_open_loop(server, requests, timeout_s: float = 600.0, progress=None, every: int = 250):
    """Submit each request at its arrival time (relative to now); return the outcomes in order.

    ``progress(stage, done, total)`` is called every ``every`` submissions,
    and whenever requests finish or the server completes more engine
    launches (at most twice a second), so a long stream keeps reporting."""
    tickets = [None] * len(requests)
    t0 = time.perf_counter()
    for i, (arrival, n, cols) in enumerate(requests):
        delay = arrival - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tickets[i] = server.submit(n, cols)
        if progress and (i + 1) % every == 0:
            progress("submitted", i + 1, len(requests))
    deadline = time.perf_counter() + timeout_s
    out = [None] * len(tickets)
    left = set(range(len(tickets)))
    reported, launches = 0, None
    while left:
        for i in sorted(left):
            r = server.poll(tickets[i])
            if r is not None:
                out[i] = r
                left.discard(i)
        done = len(tickets) - len(left)
        if progress:
            st = server.stats()["launches"]
            if done // every > reported // every or (st != launches and done < len(tickets)):
                progress(f"finished (engine launches {st})", done, len(tickets))
                reported, launches = done, st
        if not left:
            break
        if time.perf_counter() > deadline:
            raise TimeoutError(f"{len(left)} requests not finished after {timeout_s} s")
        time.sleep(0.5)
    return out


def summarize(outcomes):
    lat = sorted(o.latency_s for o in outcomes)
    first = min(o.submit_s for o in outcomes)
    last = max(o.finish_s for o in outcomes)
    span = max(last - first, 1e-9)
    return {
        "requests": len(outcomes),
        "failed": sum(1 for o in outcomes if o.status != 0),
        "span_s": span,
        "requests_per_s": len(outcomes) / span,
        "p50_s": percentile(lat, 1, 2),
        "p90_s": percentile(lat, 9, 10),
        "p99_s": percentile(lat, 99, 100),
        "max_s": lat[-1],
        "mean_batch": sum(o.max_batch for o in outcomes) / len(outcomes),
    }

