"""GPU tests of the `mbe` command line (the reference's `./bin/mbe <g.bi>
<algo>`, src/main.cu:121-234) and of the section / per-worker timing
instrumentation (the reference's DESECTION and DEVARIANCE builds,
src/header.cuh:25-32, src/mbe_noES.cuh:49-55)."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cumbe_amd", "mbe")
pytestmark = pytest.mark.gpu
M = pytest.importorskip("cumbe_amd")


@pytest.fixture(scope="module")
def c0(tmp_path_factory):
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "c0.json")))
    d = tmp_path_factory.mktemp("cli")
    path = d / "c0.bi"
    M.Graph.chung_lu(**gold["graph"]).write_bi(str(path))
    return d, path, gold


def run(d, args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([EXE] + args, cwd=d, capture_output=True, text=True, timeout=300, env=e)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


@pytest.mark.parametrize("algo", ["cuMBE", "noRS", "noES", "noWS", "bogus"])
def test_cli_lines_and_result_file(c0, algo):
    d, path, gold = c0
    out = run(d, [str(path), algo])
    name = algo if algo != "bogus" else "cuMBE"  # unknown names fall back
    assert out[1] == f"algorithm: {name}"
    assert out[3] == "|R|: 1000, |L|: 2000, |E|: 10000"
    assert f"maximal bicliques: {gold['count']}" in out
    assert f"hash: {gold['hash']}" in out
    assert any(x.startswith("runtime (s): ") for x in out)
    saved = (d / "result" / f"c0_{name}").read_text().splitlines()
    assert saved[1:] == out[1:]


def test_cli_section_percentages(c0):
    d, path, gold = c0
    out = run(d, [str(path), "cuMBE"], {"MBE_SECTIONS": "1"})
    line = [x for x in out if x.startswith("time percentage:")]
    assert len(line) == 1
    vals = [float(v) for v in line[0].split(":")[1].split()]
    assert len(vals) == 11 and abs(sum(vals) - 100.0) < 0.1


# utils/plot_section.py:7-8 of the reference: the sections its plots draw and
# the time-percentage slots each one sums
PLOT_SECTIONS = ["L' Construction", "Maximality Checking", "Maximality Expansion",
                 "Candidate Selection", "Subtree Fetching", "Work-stealing"]
PLOT_INGREDIENTS = [[3], [4], [5], [7], [8], [2, 9]]


def test_desection_alias_through_the_plot_table(c0):
    """`mbe_desection` (the name the reference's harness runs, utils/run.py:55)
    prints the time-percentage line without MBE_SECTIONS; mapped through the
    plot script's ingredient table, selection time lands under Candidate
    Selection and the solve's other sections under theirs."""
    d, path, gold = c0
    r = subprocess.run([os.path.join(ROOT, "cumbe_amd", "mbe_desection"), str(path), "cuMBE"],
                       cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    line = [x for x in r.stdout.splitlines() if x.startswith("time percentage:")]
    assert len(line) == 1
    vals = [float(v) for v in line[0].replace("time percentage:", "").split()]
    plot = {s: sum(vals[i] for i in ing) for s, ing in zip(PLOT_SECTIONS, PLOT_INGREDIENTS)}
    for s in ("L' Construction", "Maximality Checking", "Candidate Selection", "Subtree Fetching"):
        assert plot[s] > 0, (s, vals)
    # the same slots through the ABI: slot 7 is the selection clock
    g = M.Graph.load_bi(str(path))
    res = M.solve(g, "cuMBE", {"flags": M.FLAG_SECTIONS})
    assert M.SECTION_NAMES[7] == "select" and res.sections[7] > 0
    assert f"maximal bicliques: {gold['count']}" in r.stdout


def test_devariance_alias(c0):
    d, path, gold = c0
    r = subprocess.run([os.path.join(ROOT, "cumbe_amd", "mbe_devariance"), str(path), "noWS"],
                       cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    workers = int(out[4].split(",")[0].split(":")[1])
    assert len(out[5:-2]) == workers
    assert out[-2] == f"maximal bicliques: {gold['count']}"


def test_cli_worker_variance(c0):
    d, path, gold = c0
    out = run(d, [str(path), "noWS"], {"MBE_VARIANCE": "1"})
    workers = int(out[4].split(",")[0].split(":")[1])
    body = out[5:-2]
    assert len(body) == workers and all(int(x) >= 0 for x in body)
    assert out[-2] == f"maximal bicliques: {gold['count']}"
    assert out[-1].startswith("runtime (s): ")


def test_variance_flag_through_the_abi(c0):
    d, path, gold = c0
    g = M.Graph.load_bi(str(path))
    r = M.solve(g, "cuMBE", {"flags": M.FLAG_VARIANCE, "workers": 128})
    busy = M.last_worker_busy()
    assert len(busy) == 128 and max(busy) > 0
    assert (r.count, f"{r.hash:#018x}") == (gold["count"], gold["hash"])
    assert sum(busy) <= sum(r.sections)


def test_cli_list(c0):
    d, path, gold = c0
    out = run(d, ["list", str(path), str(d / "c0.txt"), "noES"])
    assert out == [f"maximal bicliques: {gold['count']}", f"hash: {gold['hash']}"]
    lines = (d / "c0.txt").read_text().splitlines()
    assert len(lines) == gold["count"]
    a, b = lines[0].split(" | ")
    assert a.startswith("A:") and b.startswith("B:")
    # each line is a biclique of the file: every A row is adjacent to every B column
    g = M.Graph.load_bi(str(path))
    n_u, n_v, off, adj = g.csr()
    for line in lines[:200]:
        a, b = line.split(" | ")
        rows = [int(x) for x in a.split()[1:]]
        cols = set(int(x) for x in b.split()[1:])
        for u in rows:
            assert cols <= set(adj[off[u]:off[u + 1]].tolist())
