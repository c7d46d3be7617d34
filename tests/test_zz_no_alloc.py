"""The allocation-free solve (mbe_solve_noalloc) against mbe_solve, through
the C ABI: two calls in a row on one device, the second with a smaller
graph, then every variant, a root subset, and the size limit.

Device memory for the packed graphs comes from hipMalloc through ctypes:
the caller owns it, the library only reads it.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
pytestmark = pytest.mark.gpu
M = pytest.importorskip("cumbe_amd")


class DevCopy:
    """A host buffer copied into device memory (hipMalloc / hipMemcpy)."""

    def __init__(self, host: np.ndarray):
        self.hip = C.CDLL("libamdhip64.so")
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(max(host.nbytes, 1))) == 0
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p),
                                  C.c_size_t(host.nbytes), 1) == 0  # host to device

    def close(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def check(g, variant="cuMBE", tuning=None):
    buf, info = g.pack()
    dev = DevCopy(buf)
    try:
        r = M.solve_noalloc(dev.ptr.value, info, variant, tuning)
    finally:
        dev.close()
    ref = M.solve(g, variant, tuning)
    assert (r.count, r.hash) == (ref.count, ref.hash), variant
    return r


def test_noalloc_two_calls_then_smaller():
    M.init()
    gold = json.load(open(os.path.join(GOLD, "c0.json")))
    big = M.Graph.chung_lu(**gold["graph"])
    small = M.Graph.chung_lu(400, 300, 2500, 2.1, 2.1, 0, 11)
    say("noalloc: configuration 0, then a smaller graph")
    r1 = check(big)
    assert (r1.count, f"{r1.hash:#018x}") == (gold["count"], gold["hash"])
    r2 = check(small)
    assert r2.count < r1.count


def test_noalloc_variants_subset_and_limits():
    M.init()
    g = M.Graph.chung_lu(400, 300, 2500, 2.1, 2.1, 0, 11)
    for v in ("cuMBE", "noRS", "noES", "noWS"):
        say(f"noalloc variant {v}")
        check(g, v)
    say("noalloc root subset and small arenas")
    check(g, "cuMBE", {"root_begin": 3, "root_stride": 4})
    r = check(g, "cuMBE", {"arena_words": 256, "workers": 32})
    assert r.launches >= 2  # roots that did not fit ran with larger arenas
    buf, info = g.pack()
    dev = DevCopy(buf)
    try:
        with pytest.raises(M.MbeError) as e:
            M.solve_noalloc(dev.ptr.value, info, "cuMBE", {"workers": 1 << 20})
        assert e.value.status == -9
    finally:
        dev.close()
