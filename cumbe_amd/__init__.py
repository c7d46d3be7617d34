"""Python access to libmbe.so, the MI355X maximal biclique enumeration engine.

This is a thin ctypes mirror of the C ABI in include/mbe.h. The reference's
only interface is its command line (``mbe <graph.bi> <cuMBE|noRS|noES|noWS>``,
src/main.cu:121-122); the same entry point is the ``mbe`` binary built next to
the library, and :func:`solve` is its in-process form.

There is no CPU fallback: if libmbe.so is missing, importing the solver
functions raises, and a solve without a HIP device returns MBE_EHIP.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Iterable, List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmbe.so")
CLI_PATH = os.path.join(_HERE, "mbe")

VARIANTS = {"cuMBE": 0, "noRS": 1, "noES": 2, "noWS": 3}

STATUS = {
    0: "ok", -1: "invalid argument", -2: "out of memory", -3: "HIP error",
    -4: "timeout", -5: "over admission budget", -6: "I/O error", -7: "busy",
    -8: "not ready", -9: "output buffer too small", -10: "internal self-check failed",
}


class MbeError(RuntimeError):
    def __init__(self, status: int, detail: str = ""):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}: {detail}")


class Tuning(C.Structure):
    _fields_ = [
        ("workers", C.c_uint32),
        ("waves_per_cu", C.c_uint32),
        ("lds_words", C.c_uint32),
        ("arena_words", C.c_uint32),
        ("list_cutoff", C.c_uint32),
        ("donate_depth", C.c_uint32),
        ("compact_ratio", C.c_uint32),
        ("donate_interval", C.c_uint32),
        ("side", C.c_int32),
        ("flags", C.c_uint32),
        ("root_begin", C.c_uint32),
        ("root_end", C.c_uint32),
        ("root_stride", C.c_uint32),
        ("max_nodes", C.c_uint64),
        ("time_limit_ms", C.c_double),
        ("sparse_ratio", C.c_uint32),
        ("split_nodes", C.c_uint32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("count", C.c_uint64),
        ("hash", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("nodes", C.c_uint64),
        ("donated", C.c_uint64),
        ("launches", C.c_uint32),
        ("workers", C.c_uint32),
        ("sections", C.c_uint64 * 11),
        ("words", C.c_uint64),
        ("build_words", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["sections"] = list(self.sections)
        return d


class PackedGraph(C.Structure):
    """mbe_packed_graph: layout of a graph packed by mbe_graph_pack."""
    _fields_ = [("n_r", C.c_uint32), ("n_l", C.c_uint32), ("n_roots", C.c_uint32),
                ("r_is_u", C.c_uint32), ("off", C.c_uint64 * 6), ("bytes", C.c_uint64)]


# mbe_result.sections, in the slots of the reference's DESECTION clocks
# (src/header.cuh:25-32, grouped by utils/plot_section.py:7-8)
SECTION_NAMES = ["other", "compact", "donate", "lprime", "check", "expand", "report",
                 "select", "root_build", "fetch", "wait"]
FLAG_SECTIONS = 1
FLAG_VARIANCE = 2
FLAG_REFERENCE_TREE = 4   # walk exactly the reference tree (no twin pruning)


_lib = None


def lib() -> C.CDLL:
    """Load libmbe.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    G = C.c_void_p
    L.mbe_status_str.restype = C.c_char_p
    L.mbe_last_error.restype = C.c_char_p
    L.mbe_graph_load_bi.argtypes = [C.c_char_p, P(G)]
    L.mbe_graph_write_bi.argtypes = [G, C.c_char_p]
    L.mbe_graph_from_edge_list.argtypes = [C.c_char_p, C.c_int, C.c_int, P(G)]
    L.mbe_graph_from_csr.argtypes = [C.c_uint32, C.c_uint32, P(C.c_uint32), P(C.c_uint32), P(G)]
    L.mbe_graph_info.argtypes = [G, P(C.c_uint32), P(C.c_uint32), P(C.c_uint64)]
    L.mbe_graph_csr.argtypes = [G, P(P(C.c_uint32)), P(P(C.c_uint32))]
    L.mbe_graph_free.argtypes = [G]
    L.mbe_graph_free.restype = None
    L.mbe_gen_chung_lu.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_double, C.c_double,
                                   C.c_uint32, C.c_uint64, P(G)]
    L.mbe_gen_er_planted.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32,
                                     C.c_uint32, C.c_double, C.c_uint64, P(G)]
    L.mbe_gen_batch_graph.argtypes = [C.c_uint64, P(G)]
    L.mbe_solve.argtypes = [G, C.c_int, P(Tuning), P(Result)]
    L.mbe_solve_list.argtypes = [G, C.c_int, P(Tuning), P(C.c_uint32), C.c_uint64,
                                 P(C.c_uint64), P(Result)]
    L.mbe_init.argtypes = []
    L.mbe_graph_pack.argtypes = [G, C.c_int, C.c_void_p, C.c_uint64, P(PackedGraph)]
    L.mbe_solve_noalloc.argtypes = [C.c_void_p, P(PackedGraph), C.c_int, P(Tuning), P(Result)]
    L.mbe_batch_solve.argtypes = [P(G), C.c_size_t, C.c_int, P(Tuning), P(Result), P(Result)]
    L.mbe_batch_plan_create.argtypes = [P(G), C.c_size_t, C.c_int, P(Tuning), P(G)]
    L.mbe_batch_plan_run.argtypes = [G, P(Result), P(Result)]
    L.mbe_batch_plan_free.argtypes = [G]
    L.mbe_batch_plan_free.restype = None
    L.mbe_last_worker_busy.argtypes = [P(C.c_uint64), C.c_uint32]
    L.mbe_last_worker_busy.restype = C.c_uint32
    L.mbe_server_create.argtypes = [P(ServerConfig), P(G)]
    L.mbe_server_submit.argtypes = [G, G, C.c_int, P(C.c_uint64)]
    L.mbe_server_poll.argtypes = [G, C.c_uint64, C.c_int, P(RequestResult)]
    L.mbe_server_get_stats.argtypes = [G, P(ServerStats)]
    L.mbe_server_destroy.argtypes = [G]
    L.mbe_server_destroy.restype = None
    for name in ("mbe_graph_load_bi", "mbe_graph_write_bi", "mbe_graph_from_edge_list",
                 "mbe_graph_from_csr", "mbe_graph_info", "mbe_graph_csr", "mbe_gen_chung_lu",
                 "mbe_gen_er_planted", "mbe_gen_batch_graph", "mbe_solve", "mbe_solve_list",
                 "mbe_init", "mbe_graph_pack", "mbe_solve_noalloc",
                 "mbe_batch_solve",
                 "mbe_batch_plan_create", "mbe_batch_plan_run", "mbe_server_create",
                 "mbe_server_submit", "mbe_server_poll", "mbe_server_get_stats"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(status: int) -> None:
    if status != 0:
        raise MbeError(status, lib().mbe_last_error().decode(errors="replace"))


class Graph:
    """A host CSR over U (rows of the .bi file), sorted and deduplicated."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.mbe_graph_free(h)
            self._h = None

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @classmethod
    def _make(cls, fn, *args) -> "Graph":
        h = C.c_void_p()
        _check(fn(*args, C.byref(h)))
        return cls(h.value)

    @classmethod
    def load_bi(cls, path: str) -> "Graph":
        return cls._make(lib().mbe_graph_load_bi, path.encode())

    @classmethod
    def from_edge_list(cls, path: str, skip_words: int = 0, skip_per_edge: int = 0) -> "Graph":
        return cls._make(lib().mbe_graph_from_edge_list, path.encode(), skip_words, skip_per_edge)

    @classmethod
    def from_csr(cls, n_u: int, n_v: int, off, adj) -> "Graph":
        off = np.ascontiguousarray(off, dtype=np.uint32)
        adj = np.ascontiguousarray(adj, dtype=np.uint32)
        if len(off) != n_u + 1:
            raise ValueError("offsets must have n_u + 1 entries")
        P = C.POINTER(C.c_uint32)
        adj_p = adj.ctypes.data_as(P) if adj.size else C.cast(C.c_void_p(0), P)
        return cls._make(lib().mbe_graph_from_csr, n_u, n_v, off.ctypes.data_as(P), adj_p)

    @classmethod
    def from_edges(cls, n_u: int, n_v: int, edges: Iterable[Sequence[int]]) -> "Graph":
        e = np.asarray(list(edges), dtype=np.int64).reshape(-1, 2)
        order = np.lexsort((e[:, 1], e[:, 0])) if len(e) else np.zeros(0, dtype=np.int64)
        e = e[order]
        off = np.zeros(n_u + 1, dtype=np.uint32)
        if len(e):
            np.add.at(off, e[:, 0] + 1, 1)
        off = np.cumsum(off).astype(np.uint32)
        return cls.from_csr(n_u, n_v, off, e[:, 1].astype(np.uint32))

    @classmethod
    def chung_lu(cls, n_u, n_v, m, gamma_u, gamma_v, max_deg, seed) -> "Graph":
        return cls._make(lib().mbe_gen_chung_lu, n_u, n_v, m, gamma_u, gamma_v, max_deg, seed)

    @classmethod
    def er_planted(cls, n_u, n_v, p, n_planted, lo, hi, top_frac, seed) -> "Graph":
        return cls._make(lib().mbe_gen_er_planted, n_u, n_v, p, n_planted, lo, hi, top_frac, seed)

    @classmethod
    def batch_graph(cls, seed: int) -> "Graph":
        return cls._make(lib().mbe_gen_batch_graph, seed)

    def write_bi(self, path: str) -> None:
        _check(lib().mbe_graph_write_bi(self._h, path.encode()))

    def info(self):
        nu, nv, ne = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(lib().mbe_graph_info(self._h, C.byref(nu), C.byref(nv), C.byref(ne)))
        return nu.value, nv.value, ne.value

    def pack(self, side: int = -1):
        """(bytes as a numpy uint8 array, PackedGraph): the graph prepared and
        packed for mbe_solve_noalloc (mbe_graph_pack)."""
        info = PackedGraph()
        _check(lib().mbe_graph_pack(self._h, side, None, 0, C.byref(info)))
        buf = np.zeros(max(info.bytes, 1), dtype=np.uint8)
        _check(lib().mbe_graph_pack(self._h, side, buf.ctypes.data, info.bytes, C.byref(info)))
        return buf[:info.bytes], info

    def csr(self):
        """(n_u, n_v, off, adj) as numpy copies."""
        nu, nv, ne = self.info()
        off_p = C.POINTER(C.c_uint32)()
        adj_p = C.POINTER(C.c_uint32)()
        _check(lib().mbe_graph_csr(self._h, C.byref(off_p), C.byref(adj_p)))
        off = np.ctypeslib.as_array(off_p, shape=(nu + 1,)).copy()
        adj = np.ctypeslib.as_array(adj_p, shape=(max(ne, 1),))[:ne].copy() if ne else np.zeros(0, np.uint32)
        return nu, nv, off, adj


def _tuning(tuning: Optional[dict]) -> Optional[Tuning]:
    if not tuning:
        return None
    t = Tuning()
    t.side = -1
    for k, v in tuning.items():
        setattr(t, k, v)
    return t


def solve(graph: Graph, variant: str = "cuMBE", tuning: Optional[dict] = None,
          allow_partial: bool = False) -> Result:
    """Enumerate all maximal bicliques of `graph` on the GPU (one launch per graph).

    With a watchdog in `tuning` (max_nodes / time_limit_ms) and allow_partial,
    a run the watchdog stopped returns its partial result with r.partial set.
    """
    r = Result()
    t = _tuning(tuning)
    st = lib().mbe_solve(graph.handle, VARIANTS[variant], C.byref(t) if t else None, C.byref(r))
    r.partial = st == -4
    if not (allow_partial and r.partial):
        _check(st)
    return r


def init() -> None:
    """mbe_init: set the engine up on device 0 (needed before solve_noalloc)."""
    _check(lib().mbe_init())


def solve_noalloc(dev_graph: int, info: PackedGraph, variant: str = "cuMBE",
                  tuning: Optional[dict] = None) -> Result:
    """mbe_solve_noalloc on a packed graph at device address dev_graph."""
    r = Result()
    t = _tuning(tuning)
    _check(lib().mbe_solve_noalloc(C.c_void_p(dev_graph), C.byref(info), VARIANTS[variant],
                                   C.byref(t) if t else None, C.byref(r)))
    return r


def solve_list(graph: Graph, variant: str = "cuMBE", tuning: Optional[dict] = None):
    """Every maximal biclique of `graph`: returns (Result, [(A, B), ...]) with A
    the row ids and B the column ids of each biclique, as sorted tuples, and
    the list sorted (mbe_solve_list: a size query, then the listing run)."""
    L = lib()
    t = _tuning(tuning)
    tp = C.byref(t) if t else None
    r = Result()
    used = C.c_uint64(0)
    _check(L.mbe_solve_list(graph.handle, VARIANTS[variant], tp, None, 0, C.byref(used), C.byref(r)))
    n = used.value
    buf = np.zeros(max(n, 1), dtype=np.uint32)
    _check(L.mbe_solve_list(graph.handle, VARIANTS[variant], tp,
                            buf.ctypes.data_as(C.POINTER(C.c_uint32)), n, C.byref(used), C.byref(r)))
    out, i = [], 0
    words = buf[:n].tolist()
    while i < n:
        a, b = words[i], words[i + 1]
        out.append((tuple(words[i + 2:i + 2 + a]), tuple(words[i + 2 + a:i + 2 + a + b])))
        i += 2 + a + b
    return r, out


def last_worker_busy() -> List[int]:
    """Per-worker busy ticks (10 ns) of this thread's last FLAG_VARIANCE solve."""
    n = lib().mbe_last_worker_busy(None, 0)
    buf = (C.c_uint64 * max(n, 1))()
    lib().mbe_last_worker_busy(buf, n)
    return list(buf)[:n]


def batch_tuning(tuning: Optional[dict]) -> Optional[dict]:
    """The batched path's launch parameters: the caller's, else the
    autotuner's cached entry for the batched workload, else the library's
    defaults (None)."""
    if tuning is not None:
        return tuning
    from . import autotune
    return autotune.cached("batch_sample")


def batch_solve(graphs: List[Graph], variant: str = "cuMBE", tuning: Optional[dict] = None,
                allow_partial: bool = False):
    """All graphs in one persistent launch; returns (per-graph results, total).
    Without a tuning the autotuner's batched entry is used (batch_tuning).

    With a watchdog in `tuning` and allow_partial, a stopped run returns its
    partial results with tot.partial set."""
    tuning = batch_tuning(tuning)
    n = len(graphs)
    arr = (C.c_void_p * max(n, 1))(*[g.handle.value for g in graphs])
    per = (Result * max(n, 1))()
    tot = Result()
    t = _tuning(tuning)
    st = lib().mbe_batch_solve(arr, n, VARIANTS[variant], C.byref(t) if t else None, per, C.byref(tot))
    tot.partial = st == -4
    if not (allow_partial and tot.partial):
        _check(st)
    return list(per)[:n], tot


class BatchPlan:
    """Graphs resident on the device with their launch captured in a HIP graph;
    run() replays it (mbe_batch_plan_*)."""

    def __init__(self, graphs: List[Graph], variant: str = "cuMBE",
                 tuning: Optional[dict] = None):
        tuning = batch_tuning(tuning)
        self.tuning = tuning
        self.n = len(graphs)
        arr = (C.c_void_p * max(self.n, 1))(*[g.handle.value for g in graphs])
        h = C.c_void_p()
        t = _tuning(tuning)
        _check(lib().mbe_batch_plan_create(arr, self.n, VARIANTS[variant],
                                           C.byref(t) if t else None, C.byref(h)))
        self._h = h

    def run(self, allow_partial: bool = False):
        per = (Result * max(self.n, 1))()
        tot = Result()
        st = lib().mbe_batch_plan_run(self._h, per, C.byref(tot))
        tot.partial = st == -4
        if not (allow_partial and tot.partial):
            _check(st)
        return list(per)[:self.n], tot

    def close(self) -> None:
        if self._h:
            lib().mbe_batch_plan_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerConfig(C.Structure):
    _fields_ = [("table_cap", C.c_uint32), ("queue_cap", C.c_uint32), ("max_nr", C.c_uint32),
                ("large_edges", C.c_uint32), ("hbm_budget", C.c_uint64), ("data_bytes", C.c_uint64),
                ("staging_bytes", C.c_uint64), ("tuning", Tuning)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("route", C.c_uint32), ("count", C.c_uint64),
                ("hash", C.c_uint64), ("footprint", C.c_uint64), ("queued_ms", C.c_double),
                ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "completed", "rejected", "launches",
                                          "queued", "in_flight", "bytes_in_use", "bicliques")]


class Server:
    """Serving loop (mbe_server_*): submit graphs, poll their results."""

    def __init__(self, hbm_budget: int = 0, table_cap: int = 0, data_bytes: int = 0,
                 staging_bytes: int = 0, tuning: Optional[dict] = None, large_edges: int = 0):
        cfg = ServerConfig()
        cfg.hbm_budget, cfg.table_cap = hbm_budget, table_cap
        cfg.large_edges = large_edges
        cfg.data_bytes, cfg.staging_bytes = data_bytes, staging_bytes
        cfg.tuning.side = -1
        for k, v in (tuning or {}).items():
            setattr(cfg.tuning, k, v)
        h = C.c_void_p()
        _check(lib().mbe_server_create(C.byref(cfg), C.byref(h)))
        self._h = h

    def submit(self, graph: Graph, variant: str = "cuMBE") -> int:
        t = C.c_uint64()
        _check(lib().mbe_server_submit(self._h, graph.handle, VARIANTS[variant], C.byref(t)))
        return t.value

    def poll(self, ticket: int, wait_ms: int = 0) -> Optional[RequestResult]:
        """The request's result, or None if it has not finished."""
        r = RequestResult()
        st = lib().mbe_server_poll(self._h, ticket, wait_ms, C.byref(r))
        if st == -8:
            return None
        _check(st)
        return r

    def stats(self) -> dict:
        s = ServerStats()
        _check(lib().mbe_server_get_stats(self._h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in ServerStats._fields_}

    def close(self) -> None:
        if self._h:
            lib().mbe_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
