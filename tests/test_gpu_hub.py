"""GPU parity on the roots the headline leaves out and on the wide-row
corner cases: the configuration-1 hub core (the 17 highest-degree roots,
2^17 - 1 bicliques) in list mode (sorted adjacency lists, the default for
such sparse roots) and in the bit-matrix form, the serving workload's large
class without its degree cap (an oracle-checked root sample), and twin
pruning on a root whose local rows span more than 64 words.

Goldens: tests/golden/configs.json (the oracle on the same root subsets).
"""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu
M = pytest.importorskip("cumbe_amd")
import oracle_lib as O  # noqa: E402
from graphs import wide_twins  # noqa: E402

VARIANTS = ("cuMBE", "noRS", "noES", "noWS")
BITMAP_ONLY = 0xFFFFFFFF


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def configs():
    return json.load(open(os.path.join(ROOT, "tests", "golden", "configs.json")))


@pytest.mark.parametrize("variant", VARIANTS)
def test_hub_core(variant):
    """The 17 highest-degree roots: the first few are localised as lists
    (huge, sparse local problems with at most 8 P candidates), the rest as
    bit matrices."""
    c = configs()["c1"]
    g = M.Graph.chung_lu(**c["graph"])
    for sub in c["hub_subsets"]:
        say(f"hub roots {sub['begin']}:{sub['end']} {variant}")
        r = M.solve(g, variant, {"root_begin": sub["begin"], "root_end": sub["end"]})
        assert (r.count, f"{r.hash:#018x}") == (sub["count"], sub["hash"])
        b = M.solve(g, variant, {"root_begin": sub["begin"], "root_end": sub["end"],
                                 "sparse_ratio": BITMAP_ONLY})
        assert (b.count, f"{b.hash:#018x}") == (sub["count"], sub["hash"])


def test_list_versus_bitmap_on_hub_roots():
    """Roots 0-5 of the serving workload's uncapped large class (degrees up
    to 45,910, about 95,000 candidates each, at most 5 of them in P) are
    localised as lists by default; both forms give the oracle's result, and
    the kernel times of the two are printed."""
    c = configs()["serving_large_uncapped"]
    g = M.Graph.chung_lu(**c["graph"], seed=c["hub_roots"]["seed"])
    for sub in c["hub_roots"]["subsets"]:
        t = {"root_begin": sub["begin"], "root_end": sub["end"]}
        lst = M.solve(g, "cuMBE", t)
        bmp = M.solve(g, "cuMBE", dict(t, sparse_ratio=BITMAP_ONLY))
        say(f"uncapped roots {sub['begin']}:{sub['end']}: lists {lst.kernel_ms:.1f} ms "
            f"({lst.launches} launches), bit matrices {bmp.kernel_ms:.1f} ms "
            f"({bmp.launches} launches)")
        for r in (lst, bmp):
            assert (r.count, f"{r.hash:#018x}") == (sub["count"], sub["hash"])
        assert lst.nodes == bmp.nodes


def test_uncapped_large_class_sample():
    c = configs()["serving_large_uncapped"]
    for e in c["seeds"]:
        g = M.Graph.chung_lu(**c["graph"], seed=e["seed"])
        say(f"uncapped large class seed {e['seed']}")
        r = M.solve(g, "cuMBE", {"root_begin": e["begin"], "root_end": e["end"],
                                 "root_stride": e["stride"]})
        assert (r.count, f"{r.hash:#018x}") == (e["count"], e["hash"])


def test_twin_pruning_on_wide_rows():
    """A root with 4,200 L vertices (66-word rows) and twin candidates: the
    pruned tree, the reference tree, list mode and the oracle agree, and
    pruning visits fewer nodes."""
    n_u, n_v, off, adj = wide_twins()
    g = M.Graph.from_csr(n_u, n_v, off, adj)
    ref = O.solve(n_u, n_v, off, adj)
    tw, tw_nodes = O.solve_twins(n_u, n_v, off, adj, 1)
    full, full_nodes = O.solve_twins(n_u, n_v, off, adj, 0)
    assert (tw.count, tw.hash) == (ref.count, ref.hash) == (full.count, full.hash)
    assert tw_nodes < full_nodes
    for v in VARIANTS:
        a = M.solve(g, v)
        b = M.solve(g, v, {"flags": M.FLAG_REFERENCE_TREE})
        s = M.solve(g, v, {"sparse_ratio": 1})
        for r in (a, b, s):
            assert (r.count, r.hash) == (ref.count, ref.hash), v
        assert a.nodes < b.nodes and s.nodes == a.nodes
