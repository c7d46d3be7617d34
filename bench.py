#!/usr/bin/env python3
"""Benchmark: maximal bicliques per second, bit-exact, on one MI355X.

Workload (one "step" = one enumeration pass over it): the one-graph-per-launch
Chung-Lu graph |U| = 94,000, |V| = 30,000, ~293,000 edges, exponents 2.0 (U)
and 1.8 (V), degrees capped at 7,600, seed 2 -- restricted to the root tasks
of rank >= 400 (19,834 of its 20,234 roots). The 400 highest-degree roots span
the graph's near-complete hub core, whose maximal bicliques number well beyond
10^9 (their enumeration does not finish in ten minutes on the GPU); the root
subset is a deterministic, oracle-checkable slice of the same graph and the
same kernel path.

Reported per JSON line (rank 0, one GPU):
  value       bicliques / kernel seconds (HIP events around the enumeration
              launches on the engine's stream; the graph is already in HBM)
  roofline    algorithmic bytes / launch time against 8 TB/s HBM: 8 bytes per
              bit-matrix operand word the visited nodes read plus 8 bytes per
              adjacency entry read while localising roots (each entry: one
              4-byte id and one 4-byte position lookup)
  cpu_baseline the CPU oracle (the reference algorithm restated serially,
              root-parallel over host threads) on a stride-64 sample of the
              roots of rank >= 1000; the GPU solves the same sample and must
              match its count and hash bit-exactly
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import cumbe_amd as M  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(n_u=94000, n_v=30000, m=293000, gamma_u=2.0, gamma_v=1.8, max_deg=7600, seed=2)
ROOT_BEGIN = 400
SAMPLE = dict(begin=1000, end=0, stride=64)
TUNING_KEY = "c1_roots400"
TUNING = dict(lds_words=4096, waves_per_cu=4, donate_interval=64)


def traffic_from_profiles():
    """HBM bytes per launch measured by rocprofv3 --pmc (committed summary)."""
    path = os.path.join(ROOT, "profiles", "mbe_persistent_c1_roots400.pmc.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^\s*hbm_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(g, gpu_tuning):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O  # the oracle: checker and CPU baseline only

    n_u, n_v, off, adj = g.csr()
    threads = max(1, min(16, os.cpu_count() or 1))
    r = O.solve_roots(n_u, n_v, off, adj, begin=SAMPLE["begin"], end=SAMPLE["end"],
                      stride=SAMPLE["stride"], threads=threads)
    tun = dict(gpu_tuning, root_begin=SAMPLE["begin"], root_end=SAMPLE["end"],
               root_stride=SAMPLE["stride"])
    gr = M.solve(g, "cuMBE", tun)
    parity = (gr.count, gr.hash) == (r.count, r.hash)
    return {
        "value": r.count / max(r.seconds, 1e-9),
        "unit": "bicliques/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle (serial reference algorithm, root-parallel) on the roots of rank "
                   f">= {SAMPLE['begin']} with stride {SAMPLE['stride']} of the same graph: "
                   f"{r.count} bicliques in {r.seconds:.2f} s; GPU on the same sample "
                   f"{'matches' if parity else 'DIFFERS'} (hash {r.hash:#018x})"),
        "parity": parity,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--variant", default="cuMBE", choices=list(M.VARIANTS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--root-begin", type=int, default=ROOT_BEGIN)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine "
              f"shards a graph by root interleaving, see mbe_tuning.root_stride, but this "
              f"benchmark runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    g = M.Graph.chung_lu(WORKLOAD["n_u"], WORKLOAD["n_v"], WORKLOAD["m"], WORKLOAD["gamma_u"],
                         WORKLOAD["gamma_v"], WORKLOAD["max_deg"], WORKLOAD["seed"])
    n_u, n_v, n_e = g.info()
    from cumbe_amd import autotune
    base = autotune.cached(TUNING_KEY) or TUNING
    tun = dict(base, root_begin=args.root_begin)
    # engine initialisation (device, code objects) outside any step
    M.solve(M.Graph.chung_lu(200, 100, 600, 2.1, 2.1, 0, 1), args.variant)

    for _ in range(args.warmup):
        M.solve(g, args.variant, tun)
    count = hsh = nodes = words = build = 0
    kernel_ms = wall_ms = 0.0
    launches = 0
    first = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = M.solve(g, args.variant, tun)
        first = first or (r.count, r.hash)
        if (r.count, r.hash) != first:
            raise SystemExit("non-deterministic result across steps")
        count += r.count
        hsh = r.hash
        nodes += r.nodes
        words += r.words
        build += r.build_words
        kernel_ms += r.kernel_ms
        launches += r.launches
    wall_ms = (time.perf_counter() - t0) * 1000.0

    value = count / (kernel_ms / 1000.0)
    alg_bytes = 8.0 * words + 8.0 * build
    launch_ms = kernel_ms / max(launches, 1)
    achieved = alg_bytes / max(launches, 1) / (launch_ms / 1000.0) / 1e9
    traffic = traffic_from_profiles()
    line = {
        "metric": "maximal bicliques/s (bit-exact)",
        "value": value,
        "unit": "bicliques/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": kernel_ms / args.steps,
        "wall_ms_per_step": wall_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": (f"Chung-Lu {n_u}x{n_v}, {n_e} edges, gamma 2.0/1.8, max degree 7600, "
                         f"seed 2; root tasks of rank >= {args.root_begin}"),
            "variant": args.variant,
            "tuning": tun,
        },
        "count": first[0],
        "hash": f"{first[1]:#018x}",
        "nodes_per_step": nodes // args.steps,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(g, base),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
