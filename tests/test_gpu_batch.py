"""GPU parity of the batched, planned and sharded paths and of the
configuration graphs' root subsets, against the oracle's golden vectors
(tests/golden/) and against each other.

Golden sources: tests/golden/batch_c3.csv (oracle per seed of the batched
workload, written by tests/golden/golden_scan.c), tests/golden/configs.json
(oracle on root subsets of the larger configuration graphs) and
tests/golden/c0.json.
"""
import json
import os
import sys

import pytest


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
pytestmark = pytest.mark.gpu
M = pytest.importorskip("cumbe_amd")
VARIANTS = ("cuMBE", "noRS", "noES", "noWS")


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def c3_rows():
    rows = []
    for line in open(os.path.join(GOLD, "batch_c3.csv")):
        if line.startswith("#") or line.startswith("seed"):
            continue
        s, nu, nv, ne, cnt, h, w = line.strip().split(",")
        if cnt == "-":
            rows.append((int(s), None, None, None))
        else:
            rows.append((int(s), int(cnt), int(h, 16), int(w)))
    return rows


def configs():
    return json.load(open(os.path.join(GOLD, "configs.json")))


def test_batch_plan_replays_match_goldens():
    """One captured persistent launch over 1,500 light graphs of the batched
    workload; two replays, every graph bit-exact against its golden."""
    rows = [r for r in c3_rows() if r[1] is not None and r[3] <= 20_000_000][:1500]
    assert len(rows) == 1500
    gs = [M.Graph.batch_graph(s) for s, _, _, _ in rows]
    say(f"plan over {len(gs)} graphs")
    plan = M.BatchPlan(gs, "cuMBE")
    for rep in range(2):
        say(f"replay {rep}")
        per, tot = plan.run()
        bad = [(s, r.count, c) for (s, c, h, _), r in zip(rows, per) if (r.count, r.hash) != (c, h)]
        assert not bad, bad[:5]
        assert tot.launches == 1
    plan.close()


def test_batch_plan_equals_per_graph_solves():
    """Batched (shared pool) and one-launch-per-graph results are identical
    and equal the oracle's on light and heavy graphs of the batched
    workload."""
    rows = c3_rows()
    light = [r for r in rows if r[1] is not None and r[3] <= 5_000_000][:40]
    heavy = [r for r in rows if r[1] is not None and r[3] > 200_000_000][:3]
    seeds = [r[0] for r in light + heavy]
    gs = [M.Graph.batch_graph(s) for s in seeds]
    say(f"plan vs per-graph on {len(gs)} graphs")
    per, _ = M.BatchPlan(gs, "cuMBE").run()
    for s, g, r in zip(seeds, gs, per):
        one = M.solve(g, "cuMBE")
        assert (one.count, one.hash) == (r.count, r.hash), s
    for (s, c, h, _), r in zip(light + heavy, per):
        assert (r.count, r.hash) == (c, h), s


def test_batch_plan_recaptures_after_overflow_pass():
    """A plan whose roots overflow the arena finishes them in follow-up
    launches; the next replay re-captures and stays exact."""
    rows = [r for r in c3_rows() if r[1] is not None and 1_000_000 <= r[3] <= 20_000_000][:12]
    gs = [M.Graph.batch_graph(s) for s, _, _, _ in rows]
    plan = M.BatchPlan(gs, "cuMBE", {"arena_words": 1024, "lds_words": 512, "workers": 64})
    for rep in range(2):
        say(f"small-arena replay {rep}")
        per, tot = plan.run()
        assert tot.launches >= 2
        for (s, c, h, _), r in zip(rows, per):
            assert (r.count, r.hash) == (c, h), s
    plan.close()


@pytest.mark.parametrize("variant", VARIANTS)
def test_config1_root_sample(variant):
    """Configuration-1 graph, sampled roots (rank >= 1000, stride 64)."""
    c = configs()["c1"]
    g = M.Graph.chung_lu(**c["graph"])
    for sub in c["subsets"]:
        say(f"c1 {variant} {sub['begin']}:{sub['end']}:{sub['stride']}")
        r = M.solve(g, variant, {"root_begin": sub["begin"], "root_end": sub["end"],
                                 "root_stride": sub["stride"]})
        assert (r.count, f"{r.hash:#018x}") == (sub["count"], sub["hash"])


@pytest.mark.parametrize("variant", VARIANTS)
def test_config2_planted_subsets(variant):
    """Configuration-2 graph (ER background + 60 planted bicliques): the 17
    deepest tractable planted roots (2^17 - 1 bicliques, recursion depth 16)
    and every non-planted root."""
    c = configs()["c2"]
    g = M.Graph.er_planted(**c["graph"])
    assert g.info()[2] == c["edges"]
    for sub in c["subsets"]:
        say(f"c2 {variant} {sub['begin']}:{sub['end']}")
        r = M.solve(g, variant, {"root_begin": sub["begin"], "root_end": sub["end"],
                                 "root_stride": sub["stride"]})
        assert (r.count, f"{r.hash:#018x}") == (sub["count"], sub["hash"])


def test_config2_deep_roots_spill_and_steal():
    """The planted subset with a small LDS budget (frames spill to HBM) and
    few workers (subtrees move through the deque)."""
    c = configs()["c2"]
    g = M.Graph.er_planted(**c["graph"])
    sub = c["subsets"][0]
    r = M.solve(g, "cuMBE", {"root_begin": 0, "root_end": 17, "lds_words": 600, "workers": 32})
    assert (r.count, f"{r.hash:#018x}") == (sub["count"], sub["hash"])
    assert r.donated > 0


def test_root_shards_sum_to_whole_on_gpu():
    gold = json.load(open(os.path.join(GOLD, "c0.json")))
    g = M.Graph.chung_lu(**gold["graph"])
    parts = [M.solve(g, "cuMBE", {"root_begin": k, "root_stride": 5}) for k in range(5)]
    assert sum(p.count for p in parts) == gold["count"]
    assert f"{sum(p.hash for p in parts) & ((1 << 64) - 1):#018x}" == gold["hash"]


def test_graphs_the_oracle_did_not_finish():
    """Graphs of the batched workload that the oracle did not finish in 20 s
    (tests/golden/c3_gpu_only.json, recorded on the GPU): one batched launch
    over the first 96 that the GPU finished, one launch per graph, and the
    other three variants on 8 of them reproduce the recorded results."""
    gold = json.load(open(os.path.join(GOLD, "c3_gpu_only.json")))
    rows = gold["finished_graphs"][:96]
    gs = [M.Graph.batch_graph(r["seed"]) for r in rows]
    say(f"plan over {len(gs)} graphs the oracle did not finish")
    per, tot = M.BatchPlan(gs, "cuMBE").run()
    for r, p in zip(rows, per):
        assert (p.count, f"{p.hash:#018x}") == (r["count"], r["hash"]), r["seed"]
    for r, g in list(zip(rows, gs))[:24]:
        one = M.solve(g, "cuMBE")
        assert (one.count, f"{one.hash:#018x}") == (r["count"], r["hash"]), r["seed"]
    for r in gold["variants_checked"][:8]:
        g = M.Graph.batch_graph(r["seed"])
        for v in VARIANTS[1:]:
            say(f"seed {r['seed']} {v}")
            x = M.solve(g, v)
            assert (x.count, f"{x.hash:#018x}") == (r["count"], r["hash"]), (r["seed"], v)
