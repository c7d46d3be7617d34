"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bit-exact count and hash on the same inputs, for all four variants.
"""
import os
import sys

import numpy as np
import pytest

import graphs as G
import oracle_lib as O

import cumbe_amd as M

pytestmark = pytest.mark.gpu
VARIANTS = ["cuMBE", "noRS", "noES", "noWS"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_graph(g):
    n_u, n_v, off, adj = g
    return M.Graph.from_csr(n_u, n_v, off, adj)


SUITE = G.small_suite()


@pytest.mark.parametrize("name,g", SUITE, ids=[n for n, _ in SUITE])
def test_small_suite(name, g):
    ref = O.solve(*g)
    gg = gpu_graph(g)
    for v in VARIANTS:
        say(f"case {name} variant {v}")
        r = M.solve(gg, v)
        assert (r.count, r.hash) == (ref.count, ref.hash), (name, v)


@pytest.mark.parametrize("seed", range(8))
def test_random_medium(seed):
    rng = np.random.default_rng(seed)
    n_u, n_v = int(rng.integers(30, 200)), int(rng.integers(30, 200))
    p = float(rng.uniform(0.03, 0.25))
    g = G.random_graph(n_u, n_v, p, 1000 + seed)
    ref = O.solve(*g)
    gg = gpu_graph(g)
    for v in VARIANTS:
        say(f"random seed {seed} {n_u}x{n_v} p={p:.3f} variant {v}")
        r = M.solve(gg, v)
        assert (r.count, r.hash) == (ref.count, ref.hash), (seed, v)


def test_side_choice_invariance():
    g = G.random_graph(80, 60, 0.15, 77)
    ref = O.solve(*g)
    gg = gpu_graph(g)
    for side in (-1, 0, 1):
        say(f"side {side}")
        r = M.solve(gg, "cuMBE", {"side": side})
        assert (r.count, r.hash) == (ref.count, ref.hash)


def test_small_arena_defers_roots():
    """Roots whose local problem exceeds the arena are deferred to a retry pass."""
    g = G.random_graph(120, 90, 0.2, 5)
    ref = O.solve(*g)
    r = M.solve(gpu_graph(g), "cuMBE", {"arena_words": 300, "lds_words": 420, "workers": 16})
    assert (r.count, r.hash) == (ref.count, ref.hash)
    assert r.launches >= 2


def test_tiny_lds_spills_stack_to_hbm():
    g = G.random_graph(40, 40, 0.5, 9)
    ref = O.solve(*g)
    for lds in (8, 64, 4096):
        say(f"lds_words {lds}")
        r = M.solve(gpu_graph(g), "cuMBE", {"lds_words": lds})
        assert (r.count, r.hash) == (ref.count, ref.hash)


def test_config0_chung_lu():
    g = M.Graph.chung_lu(2000, 1000, 10000, 2.1, 2.1, 0, 1)
    n_u, n_v, off, adj = g.csr()
    ref = O.solve(n_u, n_v, off, adj)
    for v in VARIANTS:
        say(f"config0 variant {v}")
        r = M.solve(g, v)
        assert (r.count, r.hash) == (ref.count, ref.hash), v


def test_batch_matches_per_graph_oracle():
    gs = [G.random_graph(20 + 7 * i, 15 + 5 * i, 0.2 + 0.02 * i, 300 + i) for i in range(12)]
    gs.append(G.csr_from_edges(4, 4, []))
    refs = [O.solve(*g) for g in gs]
    per, tot = M.batch_solve([gpu_graph(g) for g in gs], "cuMBE")
    for i, (r, ref) in enumerate(zip(per, refs)):
        assert (r.count, r.hash) == (ref.count, ref.hash), i
    assert tot.count == sum(r.count for r in refs)


def test_work_stealing_donates():
    """Subtrees move through the HBM deque; results stay bit-exact."""
    g = G.crown(14)
    ref = O.solve(*g)
    assert ref.count == 2 ** 14 - 2
    for v in ("cuMBE", "noRS", "noES"):
        say(f"stealing variant {v}")
        r = M.solve(gpu_graph(g), v, {"workers": 8})
        assert (r.count, r.hash) == (ref.count, ref.hash), v
        assert r.donated > 0, v
    r = M.solve(gpu_graph(g), "noWS", {"workers": 8})
    assert (r.count, r.hash) == (ref.count, ref.hash)
    assert r.donated == 0


def test_watchdog_stops_cleanly():
    """A node budget stops the kernel early with a partial result; the next
    solve on the same engine is complete and exact."""
    g = G.crown(16)
    r = M.solve(gpu_graph(g), "cuMBE", {"max_nodes": 2000}, allow_partial=True)
    assert r.partial and r.count < 2 ** 16 - 2
    with pytest.raises(M.MbeError):
        M.solve(gpu_graph(g), "cuMBE", {"max_nodes": 2000})
    ref = O.solve(*G.crown(10))
    r = M.solve(gpu_graph(G.crown(10)), "cuMBE")
    assert not r.partial and (r.count, r.hash) == (ref.count, ref.hash)


def test_reference_tree_flag_same_results_more_nodes():
    """Twin pruning (default) and the reference's own tree (FLAG_REFERENCE_TREE)
    report the same bicliques; the reference tree visits at least as many
    nodes."""
    g = G.random_graph(150, 120, 0.12, 9090)
    ref = O.solve(*g)
    gg = gpu_graph(g)
    for v in VARIANTS:
        say(f"reference tree variant {v}")
        a = M.solve(gg, v)
        b = M.solve(gg, v, {"flags": M.FLAG_REFERENCE_TREE, "side": -1})
        assert (a.count, a.hash) == (ref.count, ref.hash), v
        assert (b.count, b.hash) == (ref.count, ref.hash), v
        assert b.nodes >= a.nodes, (v, a.nodes, b.nodes)
