"""Autotuner for the enumeration launch parameters.

Picks, per workload key, the per-worker LDS budget together with the workers
per CU (pairs that fill a CU's LDS), the donation interval of the
work-stealing deque, the two list-versus-bitmap cutoffs -- sparse_ratio (a
root whose bit matrices would take more than this many words per local
adjacency entry is localised as sorted lists) and list_cutoff (|P|+|Q| at or
below which a node tests candidate rows instead of AND/OR-ing L' columns) --
and the re-localisation ratio. Workers are single waves, so the workgroup
size is not a search axis. Coordinate descent from a start point: each trial is
the mean of two runs and a move needs a 5 % gain (single runs of one
configuration vary by several percent). Every trial runs under the engine's
watchdog (a time limit), and a trial that hits it is discarded. Results are cached in tuning_cache.json next to this file and
returned by cached(key).

Usage: python -m cumbe_amd.autotune <workload> [--budget-s S] [--cache PATH]
  workloads: c1_roots400 (the bench workload), batch_sample (64 graphs of the
  batched workload's distribution whose oracle result is known; the batched
  path -- BatchPlan, batch_solve and the benchmark's batch leg -- launches
  with this entry when no tuning is given)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from typing import Callable, Dict, List, Optional

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuning_cache.json")

# LDS words per worker and workers per CU are searched together: each pair
# fills a CU's 160 KiB of LDS (more workers hide more latency, more LDS keeps
# more local problems out of HBM)
OCCUPANCY = [(6144, 3), (5120, 4), (4096, 4), (3328, 6), (3072, 6), (2560, 8), (2048, 10),
             (1536, 13)]
GRID = {
    "occupancy": OCCUPANCY,
    "donate_interval": [16, 64, 256, 1024],
    "list_cutoff": [0, 16, 64, 256],
    "compact_ratio": [2, 4, 0xFFFFFFFF],
    "sparse_ratio": [8, 32, 0xFFFFFFFF],
}
START = {"occupancy": (3072, 6), "donate_interval": 64, "list_cutoff": 0, "compact_ratio": 2,
         "sparse_ratio": 0xFFFFFFFF}


def expand(point: dict) -> dict:
    """A search point as engine tuning (occupancy -> lds_words, waves_per_cu)."""
    t = {k: v for k, v in point.items() if k != "occupancy"}
    if "occupancy" in point:
        t["lds_words"], t["waves_per_cu"] = point["occupancy"]
    return t


def cached(key: str) -> Optional[dict]:
    if not os.path.exists(CACHE):
        return None
    with open(CACHE) as f:
        return json.load(f).get(key, {}).get("tuning")


def _store(key: str, tuning: dict, ms: float, trials: List[dict]) -> None:
    data = {}
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            data = json.load(f)
    data[key] = {"tuning": tuning, "kernel_ms": ms, "trials": trials,
                 "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    with open(CACHE, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def tune(key: str, run: Callable[[dict], Optional[float]], budget_s: float,
         start: Optional[dict] = None, repeats: int = 2, gain: float = 0.05) -> dict:
    """Coordinate descent over GRID: run(tuning) returns kernel ms (None if it
    failed), averaged over `repeats` runs; a neighbour replaces the current
    point when it is faster by more than `gain`. The cached and returned
    tuning is in engine terms (expand)."""
    one = run

    def run(t):  # noqa: F811  (mean of repeated runs)
        ms = [one(t) for _ in range(repeats)]
        return None if any(m is None for m in ms) else sum(ms) / len(ms)

    best = dict(start or START)
    best_ms = run(expand(best))
    trials = [{"tuning": expand(best), "ms": best_ms}]
    t0 = time.time()
    improved = True
    while improved and time.time() - t0 < budget_s:
        improved = False
        for name, values in GRID.items():
            for v in values:
                if v == best[name] or time.time() - t0 > budget_s:
                    continue
                cand = dict(best, **{name: v})
                ms = run(expand(cand))
                trials.append({"tuning": expand(cand), "ms": ms})
                print(f"autotune {key}: {name}={v} -> {ms}", flush=True)
                if ms is not None and (best_ms is None or ms < best_ms * (1.0 - gain)):
                    best, best_ms, improved = cand, ms, True
    _store(key, expand(best), best_ms, trials)
    return expand(best)


def batch_sample_seeds(n: int = 64) -> List[int]:
    """Every 54th seed of the batched workload whose oracle result is known
    (tests/golden/batch_c3.csv), n of them."""
    path = os.path.join(os.path.dirname(HERE), "tests", "golden", "batch_c3.csv")
    seeds = []
    for line in open(path):
        if line.startswith("#") or line.startswith("seed"):
            continue
        f = line.strip().split(",")
        if f[4] != "-":
            seeds.append(int(f[0]))
    return seeds[::54][:n]


def _workload(key: str):
    sys.path.insert(0, os.path.dirname(HERE))
    import cumbe_amd as M

    if key == "c1_roots400":
        g = M.Graph.chung_lu(94000, 30000, 293000, 2.0, 1.8, 7600, 2)

        def run(t):
            r = M.solve(g, "cuMBE", dict(t, root_begin=400, time_limit_ms=60000),
                        allow_partial=True)
            return None if r.partial else r.kernel_ms
        return run
    if key == "batch_sample":
        gs = [M.Graph.batch_graph(s) for s in batch_sample_seeds()]

        def run(t):
            try:
                per, tot = M.batch_solve(gs, "cuMBE", dict(t, time_limit_ms=60000))
            except M.MbeError:
                return None
            return tot.kernel_ms
        return run
    raise SystemExit(f"unknown workload {key}")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("workload")
    ap.add_argument("--budget-s", type=float, default=240)
    ap.add_argument("--cache", help="cache file to update (default: the package's)")
    args = ap.parse_args()
    if args.cache:
        global CACHE
        if not os.path.exists(args.cache) and os.path.exists(CACHE):
            with open(CACHE) as f, open(args.cache, "w") as g:
                g.write(f.read())
        CACHE = args.cache
    best = tune(args.workload, _workload(args.workload), args.budget_s)
    print(json.dumps({"workload": args.workload, "tuning": best}))


if __name__ == "__main__":
    main()
