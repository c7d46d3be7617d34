"""Small deterministic graphs shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np


def csr_from_edges(n_u, n_v, edges):
    e = np.asarray(sorted(set(map(tuple, edges))), dtype=np.int64).reshape(-1, 2)
    off = np.zeros(n_u + 1, dtype=np.uint32)
    if len(e):
        np.add.at(off, e[:, 0] + 1, 1)
    off = np.cumsum(off).astype(np.uint32)
    adj = e[:, 1].astype(np.uint32) if len(e) else np.zeros(0, np.uint32)
    return n_u, n_v, off, adj


def random_graph(n_u, n_v, p, seed):
    rng = np.random.default_rng(seed)
    m = rng.random((n_u, n_v)) < p
    return csr_from_edges(n_u, n_v, np.argwhere(m))


def complete(a, b):
    return csr_from_edges(a, b, [(i, j) for i in range(a) for j in range(b)])


def star(n):
    return csr_from_edges(1, n, [(0, j) for j in range(n)])


def matching(n):
    return csr_from_edges(n, n, [(i, i) for i in range(n)])


def crown(n):
    """K_{n,n} minus a perfect matching: 2^n - 2 maximal bicliques."""
    return csr_from_edges(n, n, [(i, j) for i in range(n) for j in range(n) if i != j])


def half(n):
    """u_i ~ v_j iff i <= j: n maximal bicliques."""
    return csr_from_edges(n, n, [(i, j) for i in range(n) for j in range(n) if i <= j])


def disjoint_union(g1, g2):
    n_u1, n_v1, off1, adj1 = g1
    n_u2, n_v2, off2, adj2 = g2
    edges = []
    for u in range(n_u1):
        edges += [(u, int(v)) for v in adj1[off1[u]:off1[u + 1]]]
    for u in range(n_u2):
        edges += [(n_u1 + u, n_v1 + int(v)) for v in adj2[off2[u]:off2[u + 1]]]
    return csr_from_edges(n_u1 + n_u2, n_v1 + n_v2, edges)


def transpose(g):
    n_u, n_v, off, adj = g
    edges = []
    for u in range(n_u):
        edges += [(int(v), u) for v in adj[off[u]:off[u + 1]]]
    return csr_from_edges(n_v, n_u, edges)


def small_suite():
    """(name, graph) pairs covering the edge cases and random densities."""
    out = [
        ("empty", csr_from_edges(3, 4, [])),
        ("no_vertices", csr_from_edges(0, 0, [])),
        ("one_edge", csr_from_edges(1, 1, [(0, 0)])),
        ("k34", complete(3, 4)),
        ("k43", complete(4, 3)),
        ("star7", star(7)),
        ("matching5", matching(5)),
        ("crown6", crown(6)),
        ("half7", half(7)),
        ("isolated_mix", csr_from_edges(6, 5, [(0, 1), (0, 2), (3, 1), (3, 2), (5, 4)])),
    ]
    seed = 100
    for (n_u, n_v, p) in [(8, 6, 0.5), (12, 9, 0.4), (15, 15, 0.3), (20, 12, 0.5),
                          (30, 20, 0.2), (10, 40, 0.3), (40, 10, 0.6), (25, 25, 0.7),
                          (60, 50, 0.1), (16, 16, 0.9)]:
        out.append((f"rand_{n_u}x{n_v}_p{p}", random_graph(n_u, n_v, p, seed)))
        seed += 1
    return out


def wide_twins(seed=5, d=4200, n_l=6000, n_c=6):
    """One enumerated vertex x whose neighbourhood spans d > 4096 L vertices
    (local rows of more than 64 words), and n_c higher-degree candidates
    whose neighbourhoods meet N(x) in large random sets, two pairs of them
    identical (twins). Returns a CSR over U = the enumerated side."""
    rng = np.random.default_rng(seed)
    rows = [list(range(d))]
    sets = []
    for i in range(n_c):
        if i in (1, 3):
            sets.append(sets[-1])
            continue
        k = int(rng.integers(d + 100, n_l))
        sets.append(sorted(rng.choice(n_l, size=k, replace=False).tolist()))
    rows += sets
    edges = [(u, v) for u, r in enumerate(rows) for v in r]
    return csr_from_edges(len(rows), n_l, edges)
