"""The engine's kernels on the CPU: cumbe_amd/csrc/mbe_kernel.hip compiled
under the wave64 emulator (tests/emu/, one fiber per lane, waves on OS
threads) and checked against the oracle.

These tests exercise the device code's logic without a GPU: root
localisation with Q reduction, candidate selection, the reverse and forward
checks, twin pruning, compaction into nested problems, donation through the
deque and the LDS spill. The GPU tests check the same source on gfx950.
"""
import json
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU_DIR = os.path.join(ROOT, "tests", "emu")
EMU = os.path.join(EMU_DIR, "emu_run")
GOLD = os.path.join(ROOT, "tests", "golden")

import oracle_lib as O  # noqa: E402
from graphs import crown, half, random_graph  # noqa: E402

M = pytest.importorskip("cumbe_amd")


@pytest.fixture(scope="module", autouse=True)
def built():
    subprocess.run(["make", "-C", EMU_DIR], check=True, capture_output=True)


def emu(*args, timeout=600):
    out = subprocess.run([EMU, *map(str, args)], capture_output=True, text=True,
                         timeout=timeout)
    assert out.returncode == 0, out.stderr
    m = re.search(r"count (\d+) hash (0x[0-9a-f]+) nodes (\d+) donated (\d+) stolen (\d+)",
                  out.stdout)
    return {"count": int(m.group(1)), "hash": int(m.group(2), 16), "nodes": int(m.group(3)),
            "donated": int(m.group(4)), "stolen": int(m.group(5))}


def bi_file(tmp_path, name, g):
    n_u, n_v, off, adj = g
    path = str(tmp_path / f"{name}.bi")
    M.Graph.from_csr(n_u, n_v, off, adj).write_bi(path)
    return path


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_smoke_graph_all_variants(variant):
    g = M.Graph.chung_lu(400, 300, 2500, 2.1, 2.1, 0, 11)
    ref = O.solve(*g.csr())
    r = emu("-v", variant, "cl", 400, 300, 2500, 2.1, 2.1, 0, 11)
    assert (r["count"], r["hash"]) == (ref.count, ref.hash)


def test_small_suite(tmp_path):
    """Closed-form and random graphs; every variant on the first, the
    reverse-scan and forward forms on the rest."""
    cases = {"crown7": crown(7), "half9": half(9)}
    for s in range(3):
        cases[f"rand{s}"] = random_graph(30 + 7 * s, 25 + 5 * s, 0.12 + 0.05 * s, s)
    for i, (name, g) in enumerate(cases.items()):
        ref = O.solve(*g)
        path = bi_file(tmp_path, name, g)
        for v in ((0, 1, 2, 3) if i == 0 else (0, 1)):
            r = emu("-v", v, "-w", 3, "bi", path)
            assert (r["count"], r["hash"]) == (ref.count, ref.hash), (name, v)


def test_config0():
    gold = json.load(open(os.path.join(GOLD, "c0.json")))
    a = [gold["graph"][k] for k in ("n_u", "n_v", "m", "gamma_u", "gamma_v", "max_deg", "seed")]
    r = emu("cl", *a)
    assert (r["count"], f"{r['hash']:#018x}") == (gold["count"], gold["hash"])


def test_reference_tree_and_no_compaction():
    """Twin pruning off (the reference's tree) and re-localisation off give
    the same bicliques."""
    g = M.Graph.chung_lu(600, 400, 3500, 2.1, 2.1, 0, 5)
    ref = O.solve(*g.csr())
    for extra in (("-t", 0), ("-r", 4294967295)):
        r = emu(*extra, "cl", 600, 400, 3500, 2.1, 2.1, 0, 5)
        assert (r["count"], r["hash"]) == (ref.count, ref.hash), extra


def test_deep_planted_roots_spill_and_steal():
    """Planted-biclique roots (exponentially many bicliques each) with a tiny
    LDS budget (frames spill to the arena) and a short donation interval:
    subtrees move through the deque and the sums stay exact."""
    args = (6500, 1900, 0.008, 60, 8, 40, 0.05, 3)
    g = M.Graph.er_planted(*args)
    ref = O.solve_roots(*g.csr(), begin=0, end=12, threads=4)
    r = emu("-w", 8, "-e", 12, "-l", 600, "-i", 8, "er", *args)
    assert (r["count"], r["hash"]) == (ref.count, ref.hash)
    assert r["stolen"] > 0


def test_config1_sample_large_q_sets():
    """Configuration-1 roots of rank >= 1000, stride 64: thousands of Q
    candidates per root, reduced before the local problem is laid out."""
    c = json.load(open(os.path.join(GOLD, "configs.json")))["c1"]
    sub = c["subsets"][0]
    gr = c["graph"]
    r = emu("-w", 8, "-b", sub["begin"], "-k", sub["stride"], "cl", gr["n_u"], gr["n_v"], gr["m"],
            gr["gamma_u"], gr["gamma_v"], gr["max_deg"], gr["seed"])
    assert (r["count"], f"{r['hash']:#018x}") == (sub["count"], sub["hash"])


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_list_mode_forced(variant):
    """Every root in list mode (sparse ratio 1): sorted adjacency lists, the
    list check, list-mode compaction; same tree and results as bit matrices."""
    g = M.Graph.chung_lu(400, 300, 2500, 2.1, 2.1, 0, 11)
    ref = O.solve(*g.csr())
    a = emu("-v", variant, "cl", 400, 300, 2500, 2.1, 2.1, 0, 11)
    s = emu("-p", 1, "-v", variant, "cl", 400, 300, 2500, 2.1, 2.1, 0, 11)
    assert (s["count"], s["hash"]) == (ref.count, ref.hash)
    assert s["nodes"] == a["nodes"]


def test_list_mode_with_stealing():
    """List-mode roots donate levels as compacted bit-matrix snapshots."""
    args = (6500, 1900, 0.008, 60, 8, 40, 0.05, 3)
    g = M.Graph.er_planted(*args)
    ref = O.solve_roots(*g.csr(), begin=0, end=12, threads=4)
    r = emu("-p", 1, "-w", 8, "-e", 12, "-l", 600, "-i", 8, "er", *args)
    assert (r["count"], r["hash"]) == (ref.count, ref.hash)
    assert r["stolen"] > 0


def test_wide_rows_twin_pruning(tmp_path):
    """A root whose local rows span 66 words (the wide branch of twin
    pruning) with twin candidates: pruned, reference tree and list mode."""
    from graphs import wide_twins
    g = wide_twins()
    ref = O.solve(*g)
    path = bi_file(tmp_path, "wide", g)
    a = emu("bi", path)
    b = emu("-t", 0, "bi", path)
    s = emu("-p", 1, "bi", path)
    for r in (a, b, s):
        assert (r["count"], r["hash"]) == (ref.count, ref.hash)
    assert a["nodes"] < b["nodes"] and s["nodes"] == a["nodes"]


@pytest.mark.parametrize("large_edges", [5000, 1 << 18])
def test_serving_kernel(large_edges):
    """The serving kernel: a host thread publishes requests to the running
    kernel's mailbox (staggered), a scaled-down large request among small
    ones, in two claim lanes or one; every request equals the oracle."""
    rows = []
    for line in open(os.path.join(GOLD, "batch_c3.csv")):
        if line[0] in "#s":
            continue
        s, _, _, _, cnt, h, w = line.strip().split(",")
        if cnt != "-" and int(w) <= 5_000_000:
            rows.append((f"b{s}", int(cnt), int(h, 16)))
    rows = rows[:12]
    big = (2000, 1000, 6000, 2.0, 2.0, 64, 11)
    ref = O.solve(*M.Graph.chung_lu(*big).csr())
    reqs = rows[:4] + [("c" + ",".join(map(str, big)), ref.count, ref.hash)] + rows[4:]
    args = ["-w", 8, "-l", 4096, "-x", 256, "-L", large_edges, "-u", 200, "serve"]
    out = subprocess.run([EMU, *map(str, args), *[r[0] for r in reqs]], capture_output=True,
                         text=True, timeout=600)
    assert out.returncode == 0, out.stderr + out.stdout
    got = {int(m.group(1)): (int(m.group(2)), int(m.group(3), 16))
           for m in re.finditer(r"req (\d+) count (\d+) hash (0x[0-9a-f]+) err 0 done 1",
                                out.stdout)}
    assert got == {i: (c, h) for i, (_, c, h) in enumerate(reqs)}


def test_random_tunings():
    """Seeded random graphs under random launch parameters -- tiny arenas
    (overflow passes), tiny LDS (spill), donation every node, small deques,
    forced list mode, early splitting, every variant -- with the uncleared
    device buffers filled with random words; each result equals the oracle."""
    import random
    rng = random.Random(20261203)
    for case in range(6):
        ga = [rng.randint(60, 400), rng.randint(30, 250), 0, round(rng.uniform(1.8, 2.5), 2),
              round(rng.uniform(1.8, 2.5), 2), rng.choice([0, 40]), rng.randint(1, 10**6)]
        ga[2] = rng.randint(ga[0], ga[0] * 4)
        ref = O.solve(*M.Graph.chung_lu(*ga).csr())
        opts = ["-v", case % 4, "-w", rng.choice([1, 3, 8]), "-l", rng.choice([420, 1024, 6144]),
                "-a", rng.choice([2048, 16384, 1 << 18]), "-s", rng.choice([4, 64]),
                "-c", rng.choice([0, 8]), "-r", rng.choice([0, 2, 8]), "-i", rng.choice([1, 64]),
                "-p", rng.choice([0, 1, 8]), "-x", rng.choice([0, 16, 4096]),
                "-G", rng.randint(1, 10**6)]
        r = emu(*opts, "cl", *ga)
        assert (r["count"], r["hash"]) == (ref.count, ref.hash), (case, opts, ga)
