"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only the tests, the smoke check and the benchmark's CPU-baseline leg load this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")


class OracleResult(C.Structure):
    _fields_ = [("count", C.c_uint64), ("hash", C.c_uint64), ("work", C.c_uint64),
                ("seconds", C.c_double)]


_lib = None
EMIT = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_uint32), C.c_uint32,
                   C.POINTER(C.c_uint32), C.c_uint32)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        L = C.CDLL(ORACLE_SO)
        P = C.POINTER(C.c_uint32)
        L.mbe_oracle_solve.argtypes = [C.c_uint32, C.c_uint32, P, P, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.POINTER(OracleResult)]
        L.mbe_oracle_solve.restype = C.c_int
        L.mbe_oracle_solve_emit.argtypes = [C.c_uint32, C.c_uint32, P, P, C.c_int, C.c_int,
                                            C.c_int, EMIT, C.c_void_p, C.POINTER(OracleResult)]
        L.mbe_oracle_solve_emit.restype = C.c_int
        L.mbe_oracle_solve_roots.argtypes = [C.c_uint32, C.c_uint32, P, P, C.c_int, C.c_int,
                                             C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64,
                                             C.POINTER(OracleResult)]
        L.mbe_oracle_solve_roots.restype = C.c_int
        L.mbe_oracle_solve_twins.argtypes = [C.c_uint32, C.c_uint32, P, P, C.c_int,
                                             C.POINTER(OracleResult), C.POINTER(C.c_uint64)]
        L.mbe_oracle_solve_twins.restype = C.c_int
        L.mbe_oracle_brute.argtypes = [C.c_uint32, C.c_uint32, P, P, C.POINTER(OracleResult)]
        L.mbe_oracle_brute.restype = C.c_int
        L.mbe_oracle_biclique_hash.argtypes = [P, C.c_uint32, P, C.c_uint32]
        L.mbe_oracle_biclique_hash.restype = C.c_uint64
        _lib = L
    return _lib


def _arrs(off, adj):
    off = np.ascontiguousarray(off, dtype=np.uint32)
    adj = np.ascontiguousarray(adj if len(adj) else np.zeros(1, np.uint32), dtype=np.uint32)
    P = C.POINTER(C.c_uint32)
    return off, adj, off.ctypes.data_as(P), adj.ctypes.data_as(P)


def solve(n_u, n_v, off, adj, rs=1, es=1, side=-1, threads=1) -> OracleResult:
    off, adj, po, pa = _arrs(off, adj)
    r = OracleResult()
    rc = lib().mbe_oracle_solve(n_u, n_v, po, pa, rs, es, side, threads, C.byref(r))
    if rc:
        raise ValueError(f"oracle rejected input ({rc})")
    return r


def solve_roots(n_u, n_v, off, adj, begin=0, end=0, stride=1, rs=1, es=1, side=-1,
                threads=1) -> OracleResult:
    """Oracle restricted to the roots of rank in [begin, end) step stride."""
    off, adj, po, pa = _arrs(off, adj)
    r = OracleResult()
    rc = lib().mbe_oracle_solve_roots(n_u, n_v, po, pa, rs, es, side, threads, begin, end,
                                      stride, C.byref(r))
    if rc:
        raise ValueError(f"oracle rejected input ({rc})")
    return r


def enumerate_sets(n_u, n_v, off, adj, rs=1, es=1, side=-1):
    """All maximal bicliques as a set of (frozenset(A), frozenset(B)), plus the result."""
    off, adj, po, pa = _arrs(off, adj)
    out = []

    def cb(_user, a, na, b, nb):
        out.append((frozenset(a[i] for i in range(na)), frozenset(b[i] for i in range(nb))))

    f = EMIT(cb)
    r = OracleResult()
    rc = lib().mbe_oracle_solve_emit(n_u, n_v, po, pa, rs, es, side, f, None, C.byref(r))
    if rc:
        raise ValueError(f"oracle rejected input ({rc})")
    return out, r


def solve_twins(n_u, n_v, off, adj, prune):
    """(result, tree nodes) with or without iMBEA's twin removal."""
    off, adj, po, pa = _arrs(off, adj)
    r = OracleResult()
    nodes = C.c_uint64(0)
    rc = lib().mbe_oracle_solve_twins(n_u, n_v, po, pa, prune, C.byref(r), C.byref(nodes))
    if rc:
        raise ValueError(f"oracle rejected input ({rc})")
    return r, nodes.value


def brute(n_u, n_v, off, adj) -> OracleResult:
    off, adj, po, pa = _arrs(off, adj)
    r = OracleResult()
    rc = lib().mbe_oracle_brute(n_u, n_v, po, pa, C.byref(r))
    if rc:
        raise ValueError(f"brute force rejected input ({rc})")
    return r


def biclique_hash(a, b) -> int:
    na, nb = len(a), len(b)
    a = np.ascontiguousarray(sorted(a) or [0], dtype=np.uint32)
    b = np.ascontiguousarray(sorted(b) or [0], dtype=np.uint32)
    P = C.POINTER(C.c_uint32)
    return lib().mbe_oracle_biclique_hash(a.ctypes.data_as(P), na, b.ctypes.data_as(P), nb)
