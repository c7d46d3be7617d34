"""GPU parity of the materialised bicliques (mbe_solve_list) with the oracle.

The oracle enumerates every maximal biclique as sets (A rows, B columns);
the GPU list must hold exactly the same sets, sorted, for every variant and
for launches that spill, steal and overflow. Counts and hashes come along.
"""
import ctypes as C
import json
import os
import sys

import numpy as np
import pytest

import graphs as G
import oracle_lib as O

import cumbe_amd as M

pytestmark = pytest.mark.gpu
VARIANTS = ["cuMBE", "noRS", "noES", "noWS"]
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def as_sets(lst):
    return {(frozenset(a), frozenset(b)) for a, b in lst}


def check_list(lst, ref_sets):
    assert lst == sorted(lst), "records not sorted"
    assert all(list(a) == sorted(a) and list(b) == sorted(b) for a, b in lst)
    assert len(set(lst)) == len(lst), "duplicate records"
    assert as_sets(lst) == ref_sets


SUITE = G.small_suite()


@pytest.mark.parametrize("name,g", SUITE, ids=[n for n, _ in SUITE])
def test_list_small_suite(name, g):
    ref, rr = O.enumerate_sets(*g)
    gg = M.Graph.from_csr(*g)
    for v in VARIANTS:
        say(f"list {name} variant {v}")
        r, lst = M.solve_list(gg, v)
        assert (r.count, r.hash) == (rr.count, rr.hash), (name, v)
        check_list(lst, set(ref))


@pytest.mark.parametrize("seed", range(3))
def test_list_random_medium_all_variants_equal(seed):
    rng = np.random.default_rng(50 + seed)
    n_u, n_v = int(rng.integers(40, 160)), int(rng.integers(40, 160))
    g = G.random_graph(n_u, n_v, float(rng.uniform(0.05, 0.2)), 2000 + seed)
    ref, rr = O.enumerate_sets(*g)
    gg = M.Graph.from_csr(*g)
    lists = []
    for v in VARIANTS:
        say(f"list random seed {seed} {n_u}x{n_v} variant {v}")
        r, lst = M.solve_list(gg, v)
        assert (r.count, r.hash) == (rr.count, rr.hash)
        check_list(lst, set(ref))
        lists.append(lst)
    assert all(x == lists[0] for x in lists)


def test_list_c0_matches_oracle_and_golden():
    gold = json.load(open(os.path.join(GOLD, "c0.json")))
    g = M.Graph.chung_lu(**gold["graph"])
    say("list c0")
    r, lst = M.solve_list(g, "cuMBE")
    assert (r.count, f"{r.hash:#018x}") == (gold["count"], gold["hash"])
    ref, _ = O.enumerate_sets(*g.csr())
    check_list(lst, set(ref))
    # the hash of the listed sets, recomputed one biclique at a time
    h = sum(O.biclique_hash(a, b) for a, b in lst) & ((1 << 64) - 1)
    assert h == r.hash


def test_list_spill_steal_overflow_and_sides():
    g = G.random_graph(220, 180, 0.08, 4242)
    ref, rr = O.enumerate_sets(*g)
    gg = M.Graph.from_csr(*g)
    for tun in ({"lds_words": 600, "workers": 32, "donate_interval": 1},
                {"arena_words": 1024, "workers": 64},
                {"side": 0}, {"side": 1}):
        say(f"list tuning {tun}")
        t = dict(tun)
        t.setdefault("side", -1)
        r, lst = M.solve_list(gg, "cuMBE", t)
        assert (r.count, r.hash) == (rr.count, rr.hash), tun
        check_list(lst, set(ref))


def test_list_size_query_and_short_buffer():
    g = G.crown(5)  # 30 bicliques
    gg = M.Graph.from_csr(*g)
    L = M.lib()
    used = C.c_uint64(0)
    r = M.Result()
    say("list size query")
    assert L.mbe_solve_list(gg.handle, 0, None, None, 0, C.byref(used), C.byref(r)) == 0
    assert r.count == 30 and used.value == sum(2 + len(a) + len(b) for a, b in
                                               O.enumerate_sets(*g)[0])
    buf = (C.c_uint32 * used.value)()
    short = C.c_uint64(0)
    say("list short buffer")
    st = L.mbe_solve_list(gg.handle, 0, None, buf, used.value - 1, C.byref(short), C.byref(r))
    assert st == -9 and short.value == used.value
    assert L.mbe_solve_list(gg.handle, 0, None, buf, used.value, C.byref(short), C.byref(r)) == 0
