"""Host-side checks that need no GPU: the C ABI exports, graph I/O and the
.bi format, the synthetic generators, the command-line tool's data paths,
the benchmark's single-GPU contract and the autotuner's search logic.

The .bi format follows the reference reader and writer (src/main.cu:38-48,
data/gen_bi.cpp:72-79): `NUM_R NUM_L NUM_EDGES`, a blank line, NUM_R lines of
`start length`, then NUM_EDGES entries of `col 1`.
"""
import ctypes as C
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import graphs as G
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
M = pytest.importorskip("cumbe_amd")


def header_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mbe_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    names = header_functions(os.path.join(ROOT, "include", "mbe.h"))
    assert "mbe_solve" in names and "mbe_batch_plan_run" in names
    lib = M.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_and_errors_without_device():
    lib = M.lib()
    for code in range(0, -11, -1):
        assert lib.mbe_status_str(code) != b"unknown status"
    with pytest.raises(M.MbeError) as e:
        M.Graph.from_csr(2, 2, [0, 1, 2], [0, 5])  # id 5 out of range
    assert e.value.status == -1
    # argument checks of the listing entry point come before any device call
    used = C.c_uint64(0)
    assert lib.mbe_solve_list(None, 0, None, None, 0, C.byref(used), None) == -1
    g = M.Graph.from_csr(2, 2, [0, 1, 2], [0, 1])
    assert lib.mbe_solve_list(g.handle, 0, None, None, 5, C.byref(used), None) == -1
    assert lib.mbe_solve_list(g.handle, 7, None, None, 0, C.byref(used), None) == -1
    assert lib.mbe_solve_list(g.handle, 0, None, None, 0, None, None) == -1


def test_bi_round_trip(tmp_path):
    g = M.Graph.chung_lu(300, 120, 900, 2.1, 2.1, 0, 5)
    p = tmp_path / "g.bi"
    g.write_bi(str(p))
    lines = p.read_text().split("\n")
    n_r, n_l, n_e = map(int, lines[0].split())
    assert (n_r, n_l, n_e) == g.info()
    assert lines[1] == ""
    h = M.Graph.load_bi(str(p))
    a, b = g.csr(), h.csr()
    assert a[:2] == b[:2]
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


def test_bi_reader_accepts_unsorted_lists(tmp_path):
    # gen_bi writes lists in hash-set order; the reader sorts and dedupes
    p = tmp_path / "u.bi"
    p.write_text("2 4 5\n\n0 3\n3 2\n3 1 1 1 0 1\n2 1 0 1\n")
    g = M.Graph.load_bi(str(p))
    n_u, n_v, off, adj = g.csr()
    assert (n_u, n_v) == (2, 4)
    assert off.tolist() == [0, 3, 5]
    assert adj.tolist() == [0, 1, 3, 0, 2]


def test_bi_reader_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.bi"
    p.write_text("2 4 3\n\n0 2\n2 2\n0 1 9 1 1 1\n")  # truncated / out of range
    with pytest.raises(M.MbeError):
        M.Graph.load_bi(str(p))
    with pytest.raises(M.MbeError) as e:
        M.Graph.load_bi(str(tmp_path / "missing.bi"))
    assert e.value.status == -6


def test_edge_list_conversion(tmp_path):
    # first-seen relabelling of both sides and duplicate removal
    # (data/gen_bi.cpp:34-51); header words and per-edge extras skipped
    p = tmp_path / "out.edges"
    p.write_text("% bip unweighted\n% 6 3 4\n10 7 1\n12 7 1\n10 7 1\n11 9 1\n12 8 1\n10 8 1\n")
    g = M.Graph.from_edge_list(str(p), skip_words=7, skip_per_edge=1)
    n_u, n_v, off, adj = g.csr()
    assert (n_u, n_v) == (3, 3)  # U: 10,12,11 -> 0,1,2; V: 7,9,8 -> 0,1,2
    rows = [sorted(adj[off[i]:off[i + 1]].tolist()) for i in range(n_u)]
    assert rows == [[0, 2], [0, 2], [1]]


def test_generators_are_deterministic_and_sized():
    a = M.Graph.chung_lu(2000, 1000, 10000, 2.1, 2.1, 0, 1)
    b = M.Graph.chung_lu(2000, 1000, 10000, 2.1, 2.1, 0, 1)
    assert a.info() == (2000, 1000, 10000)
    assert all(np.array_equal(x, y) for x, y in zip(a.csr()[2:], b.csr()[2:]))
    c = M.Graph.chung_lu(2000, 1000, 10000, 2.1, 2.1, 0, 2)
    assert not np.array_equal(a.csr()[3], c.csr()[3])
    e = M.Graph.er_planted(650, 190, 0.02, 6, 8, 12, 0.05, 3)
    n_u, n_v, off, adj = e.csr()
    assert (n_u, n_v) == (650, 190)
    assert abs(e.info()[2] - 0.02 * 650 * 190) < 0.02 * 650 * 190 * 0.5 + 6 * 144


@pytest.mark.parametrize("seed", [1000, 1001, 1500, 2048, 5095])
def test_batch_graph_distribution(seed):
    n_u, n_v, m = M.Graph.batch_graph(seed).info()
    assert 50 <= n_u <= 5000 and 20 <= n_v <= 2000
    assert m <= max(200, n_u * n_v // 4) and m <= 50000


def test_capped_generator_respects_degree_cap():
    g = M.Graph.chung_lu(9400, 3000, 29300, 2.0, 1.8, 760, 2)
    n_u, n_v, off, adj = g.csr()
    assert np.diff(off).max() <= 760
    assert np.bincount(adj, minlength=n_v).max() <= 760


def test_cli_convert_and_gen(tmp_path):
    exe = os.path.join(ROOT, "cumbe_amd", "mbe")
    if not os.path.exists(exe):
        pytest.skip("mbe command-line tool not built")
    edges = tmp_path / "e.txt"
    edges.write_text("1 2\n1 3\n4 2\n")
    out = tmp_path / "e.bi"
    subprocess.run([exe, "convert", str(edges), str(out), "0", "0"], check=True, capture_output=True)
    assert out.read_text().split("\n")[0].split() == ["2", "2", "3"]
    gen = tmp_path / "c0.bi"
    subprocess.run([exe, "gen", "chung-lu", str(gen), "2000", "1000", "10000", "2.1", "2.1",
                    "0", "1"], check=True, capture_output=True)
    assert M.Graph.load_bi(str(gen)).info() == (2000, 1000, 10000)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_autotune_coordinate_descent(tmp_path, monkeypatch):
    from cumbe_amd import autotune as A
    monkeypatch.setattr(A, "CACHE", str(tmp_path / "cache.json"))
    best_point = {"occupancy": (2560, 8), "donate_interval": 16, "list_cutoff": 64,
                  "compact_ratio": 4, "sparse_ratio": 32}

    def cost(t):  # separable bowl with its minimum on the grid, steep enough
        # that every step towards it beats the 3% acceptance threshold
        p = dict(t)
        p["occupancy"] = (p.pop("lds_words"), p.pop("waves_per_cu"))
        return 100.0 * 1.5 ** sum(abs(A.GRID[k].index(p[k]) - A.GRID[k].index(v))
                                  for k, v in best_point.items())

    got = A.tune("toy", cost, budget_s=60)
    assert got == A.expand(best_point)
    assert got["lds_words"] == 2560 and got["waves_per_cu"] == 8
    assert A.cached("toy") == got
    data = json.load(open(tmp_path / "cache.json"))
    assert data["toy"]["kernel_ms"] == 100.0


def test_golden_fixture_c0():
    """The configuration-0 graph's golden count and hash (the oracle's, on
    the same generator output) are reproduced by the oracle."""
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "c0.json")))
    g = M.Graph.chung_lu(**gold["graph"])
    n_u, n_v, off, adj = g.csr()
    assert [n_u, n_v, int(off[-1])] == gold["shape"]
    r = O.solve(n_u, n_v, off, adj)
    assert (r.count, f"{r.hash:#018x}") == (gold["count"], gold["hash"])


def test_root_shards_add_up():
    """Interleaved root shards (the multi-device split) sum to the whole."""
    g = M.Graph.chung_lu(2000, 1000, 10000, 2.1, 2.1, 0, 1)
    n_u, n_v, off, adj = g.csr()
    whole = O.solve(n_u, n_v, off, adj)
    parts = [O.solve_roots(n_u, n_v, off, adj, begin=k, end=0, stride=4) for k in range(4)]
    assert sum(p.count for p in parts) == whole.count
    assert sum(p.hash for p in parts) & ((1 << 64) - 1) == whole.hash


def test_edge_list_conversion_matches_reference_gen_bi(tmp_path):
    """`mbe convert` against the reference's own converter (data/gen_bi.cpp,
    compiled from its source into oracle/_ref by `make -C oracle ref`): the
    same edge file gives the same .bi graph (gen_bi writes each row in hash-set
    order, so lists are compared sorted)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "gen_bi")
    if not os.path.exists(ref):
        pytest.skip("oracle/_ref/gen_bi not built (the reference is not present)")
    rng = np.random.default_rng(5)
    lines = ["% bip", "% 1 2 3"]
    for _ in range(3000):  # duplicates, sparse ids, an extra word per edge
        lines.append(f"{int(rng.integers(0, 400)) * 7 + 3} {int(rng.integers(0, 150)) * 11} "
                     f"{int(rng.integers(1, 9))}")
    src = tmp_path / "edges.txt"
    src.write_text("\n".join(lines) + "\n")
    theirs = tmp_path / "ref.bi"
    subprocess.run([ref, str(src), str(theirs), "6", "1"], check=True, capture_output=True,
                   timeout=120)
    ours = tmp_path / "ours.bi"
    M.Graph.from_edge_list(str(src), skip_words=6, skip_per_edge=1).write_bi(str(ours))
    assert theirs.read_text().split("\n")[0] == ours.read_text().split("\n")[0]
    a, b = M.Graph.load_bi(str(theirs)).csr(), M.Graph.load_bi(str(ours)).csr()
    assert a[:2] == b[:2]
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


def test_serve_bench_helpers(tmp_path):
    from cumbe_amd import serve_bench as SB
    xs = [float(i) for i in range(1, 101)]
    assert SB.pct(xs, 50) == 50.0 and SB.pct(xs, 99) == 99.0 and SB.pct(xs, 99.9) == 100.0
    assert SB.pct([], 50) is None
    assert SB.summary([3.0, 1.0, 2.0]) == {"n": 3, "p50_ms": 2.0, "p99_ms": 3.0, "p999_ms": 3.0}
    seeds = SB.seeds_from_csv(os.path.join(ROOT, "tests", "golden", "batch_c3.csv"))
    assert len(seeds) == 3471 and seeds[0] == 1000 and 1006 not in seeds
    light = SB.seeds_from_csv(os.path.join(ROOT, "tests", "golden", "batch_c3.csv"), max_work=10**6)
    assert set(light) < set(seeds)


def test_section_slots_follow_the_reference_plot_table():
    """mbe_result.sections uses the reference's DESECTION slots as its plot
    script groups them (utils/plot_section.py:7-8): 3 L', 4 checking,
    5 expansion, 7 selection, 8 subtree fetching, 2 + 9 work stealing."""
    names = M.SECTION_NAMES
    assert len(names) == 11
    assert (names[3], names[4], names[5], names[7], names[8]) == (
        "lprime", "check", "expand", "select", "root_build")
    assert {names[2], names[9]} == {"donate", "fetch"}
    for alias in ("mbe_desection", "mbe_devariance"):
        p = os.path.join(ROOT, "cumbe_amd", alias)
        assert os.path.exists(p) and os.path.realpath(p) == os.path.realpath(
            os.path.join(ROOT, "cumbe_amd", "mbe"))


def test_batched_path_reads_the_tuning_cache(tmp_path, monkeypatch):
    """BatchPlan and batch_solve launch with the autotuner's batched entry
    when no tuning is given, and with the caller's tuning otherwise."""
    from cumbe_amd import autotune as A
    monkeypatch.setattr(A, "CACHE", str(tmp_path / "cache.json"))
    assert M.batch_tuning(None) is None
    A._store("batch_sample", {"lds_words": 4096, "waves_per_cu": 4}, 10.0, [])
    assert M.batch_tuning(None) == {"lds_words": 4096, "waves_per_cu": 4}
    assert M.batch_tuning({"workers": 8}) == {"workers": 8}
    seeds = A.batch_sample_seeds()
    assert len(seeds) == 64 and len(set(seeds)) == 64
