"""GPU tests of the serving loop (mbe_server_*): every request's count and
hash is the same alone and in a mix of concurrent requests and equals the
oracle's golden vector; requests the serving kernel does not take (other
variants, local problems beyond its arenas) come back exact from the engine
path; kernel generations rotate when the request table fills; admission
rejects what the HBM budget cannot hold.

Goldens: tests/golden/batch_c3.csv (the oracle per seed of the batched
workload's size distribution).
"""
import os
import sys
import threading
import time

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu
M = pytest.importorskip("cumbe_amd")


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def golden_rows(max_work, n, min_work=0):
    rows = []
    for line in open(os.path.join(ROOT, "tests", "golden", "batch_c3.csv")):
        if line.startswith("#") or line.startswith("seed"):
            continue
        s, nu, nv, ne, cnt, h, w = line.strip().split(",")
        if cnt != "-" and min_work <= int(w) <= max_work:
            rows.append((int(s), int(cnt), int(h, 16)))
    return rows[:n]


def collect(srv, tickets, timeout_s=120):
    out = {}
    t0 = time.time()
    for k, t in tickets.items():
        r = srv.poll(t, wait_ms=int(max(1, timeout_s - (time.time() - t0)) * 1000))
        assert r is not None, f"request {k} did not finish"
        out[k] = r
    return out


def test_alone_equals_mix_equals_golden():
    rows = golden_rows(50_000_000, 48) + golden_rows(2_000_000_000, 4, min_work=200_000_000)
    graphs = {s: M.Graph.batch_graph(s) for s, _, _ in rows}
    with M.Server() as srv:
        alone = {}
        for s, _, _ in rows:
            say(f"alone {s}")
            alone[s] = collect(srv, {s: srv.submit(graphs[s])})[s]
        # the mix: two submitting threads, interleaved, while earlier
        # requests are still on the GPU
        tickets = {}
        lock = threading.Lock()

        def feed(part):
            for s, _, _ in part:
                t = srv.submit(graphs[s])
                with lock:
                    tickets[s] = t
                time.sleep(0.0005)

        th = [threading.Thread(target=feed, args=(rows[i::2],)) for i in range(2)]
        say("mix")
        for x in th:
            x.start()
        for x in th:
            x.join()
        mix = collect(srv, tickets)
        st = srv.stats()
    for s, c, h in rows:
        a, m = alone[s], mix[s]
        assert a.status == 0 and m.status == 0, s
        assert (a.count, a.hash) == (c, h), s
        assert (m.count, m.hash) == (c, h), s
    assert st["completed"] == 2 * len(rows) and st["in_flight"] == 0
    assert st["bytes_in_use"] == 0


def test_engine_path_variants_and_rotation():
    rows = golden_rows(20_000_000, 12, min_work=1_000_000)
    graphs = [(s, c, h, M.Graph.batch_graph(s)) for s, c, h in rows]
    # a table of 5 requests per generation forces several generations; the
    # small arena sends the larger requests to the engine path
    with M.Server(table_cap=5, tuning={"arena_words": 1 << 14}) as srv:
        tickets = [(s, c, h, v, srv.submit(g, v)) for (s, c, h, g) in graphs
                   for v in ("cuMBE", "noES")]
        say("engine path + variants + rotation")
        routes = set()
        for s, c, h, v, t in tickets:
            r = srv.poll(t, wait_ms=120_000)
            assert r is not None and r.status == 0, (s, v)
            assert (r.count, r.hash) == (c, h), (s, v)
            if v == "noES":
                assert r.route == 1
            routes.add(r.route)
        st = srv.stats()
    assert routes == {0, 1}
    assert st["launches"] >= 2 and st["completed"] == len(tickets)


def footprint(g):
    """Admission footprint as the server defines it: packed CSR bytes plus
    the largest local problem (u64 words) any root can build, with the rows
    of every candidate and the kept-Q list that localisation builds at the
    tail of the arena."""
    import numpy as np
    n_u, n_v, off, adj = g.csr()
    du = np.diff(off).astype(np.int64)
    dv = np.bincount(adj, minlength=n_v).astype(np.int64)
    row = np.repeat(np.arange(n_u), du)
    if n_u <= n_v:
        nR, nL, dR, dL = n_u, n_v, du, dv
        s = np.bincount(row, weights=dv[adj] - 1, minlength=n_u)
    else:
        nR, nL, dR, dL = n_v, n_u, dv, du
        s = np.bincount(adj, weights=du[row] - 1, minlength=n_v)
    d = dR[dR > 0]
    nc = np.minimum(s[dR > 0].astype(np.int64), nR - 1)
    W, Wc = (d + 63) // 64, (nc + 63) // 64
    words = (d + nc + nc * W + d * Wc + (nc + 1) // 2 + 1 + (d + 1) // 2 + 1 +
             (np.minimum(d, nc) + 2) * (4 + W + 3 * Wc) +
             (d + 2) // 2 + np.maximum(d, (nc + 1) // 2) + 1 +
             nc * W + (nc + 1) // 2 + 1)
    slot = lambda b: (b + 8 + 255) & ~255
    E = int(off[-1])
    packed = (slot((nR + 1) * 4) + slot(E * 4) + slot((nL + 1) * 4) + slot(E * 4) +
              slot(nR * 8) + slot(nL * 8))
    return packed + 8 * int(words.max())


def test_admission_budget():
    small = M.Graph.batch_graph(golden_rows(100_000, 1)[0][0])
    big = M.Graph.chung_lu(20000, 8000, 200000, 2.0, 2.0, 0, 5)
    fs, fb = footprint(small), footprint(big)
    assert fs < fb
    budget = (fs + fb) // 2
    with M.Server(hbm_budget=budget) as srv:
        with pytest.raises(M.MbeError) as e:
            srv.submit(big)
        assert e.value.status == -5
        r = srv.poll(srv.submit(small), wait_ms=60_000)
        assert r is not None and r.status == 0 and r.footprint == fs
        assert srv.stats()["rejected"] == 1


def test_large_class_exact_alone_and_served():
    """The serving workload's large class (200,000 x 100,000, 1.5M edges,
    degrees capped at 128) against the oracle, through mbe_solve and through
    the server together with small requests."""
    import json
    c = json.load(open(os.path.join(ROOT, "tests", "golden", "configs.json")))["serving_large"]
    graphs = {e["seed"]: M.Graph.chung_lu(**c["graph"], seed=e["seed"]) for e in c["seeds"]}
    gold = {e["seed"]: (e["count"], int(e["hash"], 16)) for e in c["seeds"]}
    for s, g in graphs.items():
        say(f"large {s} alone")
        r = M.solve(g)
        assert (r.count, r.hash) == gold[s], s
    small = golden_rows(5_000_000, 20)
    with M.Server() as srv:
        t_large = {s: srv.submit(g) for s, g in graphs.items()}
        t_small = {s: srv.submit(M.Graph.batch_graph(s)) for s, _, _ in small}
        say("large + small served together")
        for s, t in t_large.items():
            r = srv.poll(t, wait_ms=120_000)
            assert r is not None and r.status == 0 and (r.count, r.hash) == gold[s], s
        for s, c0, h0 in small:
            r = srv.poll(t_small[s], wait_ms=120_000)
            assert r is not None and (r.count, r.hash) == (c0, h0), s


@pytest.mark.parametrize("large_edges", [1, 0, 0xFFFFFFFF])
def test_claim_lanes(large_edges):
    """Small and large requests in the two claim lanes (every request large,
    the default split, a single lane): each result equals the oracle's."""
    import json
    c = json.load(open(os.path.join(ROOT, "tests", "golden", "configs.json")))["serving_large"]
    large = [(e["seed"], e["count"], int(e["hash"], 16)) for e in c["seeds"]]
    small = golden_rows(5_000_000, 24)
    with M.Server(large_edges=large_edges) as srv:
        say(f"claim lanes, large_edges {large_edges}")
        tickets = []
        for i, (s, cnt, h) in enumerate(small):
            if i % 12 == 0:
                ls, lc, lh = large[(i // 12) % len(large)]
                tickets.append((ls, lc, lh, srv.submit(M.Graph.chung_lu(**c["graph"], seed=ls))))
            tickets.append((s, cnt, h, srv.submit(M.Graph.batch_graph(s))))
        for s, cnt, h, t in tickets:
            r = srv.poll(t, wait_ms=120_000)
            assert r is not None and r.status == 0, s
            assert (r.count, r.hash) == (cnt, h), s
