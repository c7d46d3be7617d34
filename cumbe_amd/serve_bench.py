"""Serving benchmark: Poisson arrivals against the serving loop (mbe_server_*).

Requests are drawn from a pool of small graphs (the batched workload's size
distribution, given as generator seeds) and, with probability `large_frac`,
from a pool of large graphs. Submitting threads release each request at its
arrival time; every result is compared with the same graph's result when it
was served alone. Reported per rate: completed requests, bicliques/s over the
run (first arrival to last completion), latency p50/p99/p99.9 per class and
overall, and the isolated (alone) latencies of the same requests.

Usage: python -m cumbe_amd.serve_bench --seeds-csv tests/golden/batch_c3.csv
           [--rates 200,1000,2000] [--duration 60] [--large-frac 0.01]
"""
from __future__ import annotations

import argparse
import json
import math
import random
import threading
import time
from typing import Dict, List, Optional, Sequence

from . import Graph, Server


def pct(xs: Sequence[float], q: float) -> Optional[float]:
    if not xs:
        return None
    s = sorted(xs)
    return s[min(len(s) - 1, max(0, math.ceil(q / 100.0 * len(s)) - 1))]


def summary(xs: Sequence[float]) -> dict:
    return {"n": len(xs), "p50_ms": pct(xs, 50), "p99_ms": pct(xs, 99), "p999_ms": pct(xs, 99.9)}


def isolated_results(srv: Server, graphs: Dict[str, Graph]) -> Dict[str, object]:
    """Serve each graph alone (nothing else in flight)."""
    out = {}
    for k, g in graphs.items():
        out[k] = srv.poll(srv.submit(g), wait_ms=-1)
    return out


def poisson_run(srv: Server, small: Dict[str, Graph], large: Dict[str, Graph], rate: float,
                duration: float, large_frac: float, seed: int, submitters: int = 4) -> dict:
    rng = random.Random(seed)
    small_keys, large_keys = sorted(small), sorted(large)
    arrivals = []
    t = 0.0
    while True:
        t += rng.expovariate(rate)
        if t >= duration:
            break
        if large_keys and rng.random() < large_frac:
            arrivals.append((t, "large", rng.choice(large_keys)))
        else:
            arrivals.append((t, "small", rng.choice(small_keys)))
    tickets: List[Optional[int]] = [None] * len(arrivals)
    late: List[float] = [0.0] * len(arrivals)
    t0 = time.perf_counter() + 0.05

    def feed(k: int) -> None:
        for i in range(k, len(arrivals), submitters):
            at, cls, key = arrivals[i]
            wait = t0 + at - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            late[i] = max(0.0, time.perf_counter() - (t0 + at))
            tickets[i] = srv.submit(small[key] if cls == "small" else large[key])

    th = [threading.Thread(target=feed, args=(k,)) for k in range(submitters)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    submitted_s = time.perf_counter() - t0
    results = [srv.poll(t, wait_ms=-1) for t in tickets]
    wall = time.perf_counter() - t0
    return {"arrivals": arrivals, "results": results, "wall_s": wall,
            "submit_s": submitted_s, "max_submit_lag_ms": 1000.0 * max(late or [0.0])}


def run(small: Dict[str, Graph], large: Dict[str, Graph], rates: Sequence[float],
        duration: float, large_frac: float, seed: int = 7,
        server_kw: Optional[dict] = None, progress=lambda msg: None) -> dict:
    out = {"duration_s": duration, "large_frac": large_frac, "rates": {}}
    with Server(**(server_kw or {})) as srv:
        alone = isolated_results(srv, {**{("s", k): g for k, g in small.items()},
                                       **{("l", k): g for k, g in large.items()}})
        progress(f"serving: {len(alone)} graphs served alone")
        for rate in rates:
            r = poisson_run(srv, small, large, rate, duration, large_frac, seed)
            progress(f"serving: rate {rate} done in {r['wall_s']:.1f} s")
            lat = {"small": [], "large": []}
            iso = {"small": [], "large": []}
            mismatches = 0
            bic = 0
            for (at, cls, key), res in zip(r["arrivals"], r["results"]):
                a = alone[("s" if cls == "small" else "l", key)]
                if (res.status, res.count, res.hash) != (a.status, a.count, a.hash):
                    mismatches += 1
                lat[cls].append(res.latency_ms)
                iso[cls].append(a.latency_ms)
                bic += res.count
            all_lat = lat["small"] + lat["large"]
            all_iso = iso["small"] + iso["large"]
            out["rates"][str(rate)] = {
                "requests": len(r["results"]),
                "bicliques_per_s": bic / r["wall_s"],
                "wall_s": r["wall_s"],
                "max_submit_lag_ms": r["max_submit_lag_ms"],
                "latency": summary(all_lat),
                "isolated_latency": summary(all_iso),
                "p99_vs_isolated_p99": (pct(all_lat, 99) / pct(all_iso, 99)) if all_iso else None,
                "small": summary(lat["small"]),
                "large": summary(lat["large"]),
                "results_equal_alone": mismatches == 0,
                "mismatches": mismatches,
            }
        out["server_stats"] = srv.stats()
    return out


def seeds_from_csv(path: str, max_work: Optional[int] = None) -> List[int]:
    """Seeds whose oracle result is known (the golden file's finished rows)."""
    seeds = []
    for line in open(path):
        if line.startswith("#") or line.startswith("seed"):
            continue
        f = line.strip().split(",")
        if f[4] != "-" and (max_work is None or int(f[6]) <= max_work):
            seeds.append(int(f[0]))
    return seeds


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds-csv", required=True)
    ap.add_argument("--pool", type=int, default=512, help="distinct small graphs")
    ap.add_argument("--rates", default="200,1000,2000")
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--large-frac", type=float, default=0.01)
    ap.add_argument("--large", default="200000,100000,1500000,2.0,128",
                    help="n_u,n_v,edges,gamma,max_deg of the large class")
    ap.add_argument("--large-seeds", default="11,12")
    args = ap.parse_args()
    seeds = seeds_from_csv(args.seeds_csv)
    rng = random.Random(1)
    pick = sorted(rng.sample(seeds, min(args.pool, len(seeds))))
    small = {s: Graph.batch_graph(s) for s in pick}
    nu, nv, m, gam, cap = args.large.split(",")
    large = {s: Graph.chung_lu(int(nu), int(nv), int(m), float(gam), float(gam), int(cap), s)
             for s in map(int, args.large_seeds.split(",")) if args.large_frac > 0}
    res = run(small, large, [float(x) for x in args.rates.split(",")], args.duration,
              args.large_frac)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
