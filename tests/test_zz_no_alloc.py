"""GPU parity of blest_bfs_noalloc (the single-source BFS whose traversal state lives in
the library's per-GPU scratch array, with no device allocation) against blest_bfs on the
same resident graphs, called through the C ABI: two calls in a row, the second on a
smaller graph, levels and statistics equal (and equal to the oracle's levels).  Also the
argument checks: host or undersized output arrays and graphs whose traversal state does
not fit the scratch are rejected before anything runs."""
import ctypes as C
import sys

import numpy as np
import pytest
from conftest import progress

pytestmark = pytest.mark.gpu

HIP_D2H = 2


def say(*a):
    progress(*a)


class DeviceBuffer:
    """Device memory allocated by the test (hipMalloc through libamdhip64), not by the
    entry point under test."""

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def to_numpy(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, HIP_D2H) == 0
        return out

    def free(self):
        if self.ptr.value:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def noalloc_bfs(blest, dg, src, buf):
    lib = blest._lib.load()
    st = blest._lib.BfsStats()
    blest.check(lib.blest_bfs_noalloc(dg._h, src, C.cast(buf.ptr, blest._lib.I32P), C.byref(st), None))
    return buf.to_numpy(np.int32, dg.n), st


def free_hbm(blest):
    lib = blest._lib.load()
    f, t = C.c_uint64(), C.c_uint64()
    blest.check(lib.blest_device_memory_info(0, C.byref(f), C.byref(t)))
    return f.value


CASES = [
    # (name, graph factory, reorder): the first is the larger one; the second call of
    # the pair runs on a smaller graph and keeps the first call's scratch contents
    ("rmat14_jaccard", lambda b: b.CSC.rmat(14, 16, seed=21), "JACCARD"),
    ("rmat11_natural", lambda b: b.CSC.rmat(11, 8, seed=22), None),
]


@pytest.mark.parametrize("direction", [0, 1])
def test_noalloc_matches_blest_bfs(blest, oracle, direction):
    graphs = []
    for name, make, reorder in CASES:
        g = make(blest)
        cp0, rows0 = g.arrays()
        inv = None if reorder is None else g.reorder(getattr(blest.Reorder, reorder), window=1024)
        bv = blest.BVSS(8, 4).construct_from_csc(g)
        dg = blest.DeviceGraph(bv, g, inv, direction=direction)
        graphs.append((name, g, cp0, rows0, dg))
    for name, g, cp0, rows0, dg in graphs:
        say(f"case noalloc {name} direction={direction} n={g.n}")
        buf = DeviceBuffer(4 * g.n)
        k = blest.BVSSBFSKernel(dg)
        try:
            for src in [int(r) for r in g.pick_roots(3, seed=5)]:
                want = k.run(src)
                got, st = noalloc_bfs(blest, dg, src, buf)
                np.testing.assert_array_equal(got, want.levels, err_msg=f"{name} src={src}")
                np.testing.assert_array_equal(got, oracle.bfs(g.n, cp0, rows0, src), err_msg=f"{name} src={src}")
                assert st.total_levels == want.total_levels and st.no_visited == want.no_visited
                assert st.edges_traversed_component == want.edges_in_component
                assert st.bottom_up_mask == want.bottom_up_mask
                assert st.kernel_ms > 0
        finally:
            buf.free()


def test_noalloc_allocates_nothing(blest, oracle):
    g = blest.CSC.rmat(12, 16, seed=23)
    cp0, rows0 = g.arrays()
    inv = g.reorder(blest.Reorder.JACCARD, window=1024)
    dg = blest.DeviceGraph(blest.BVSS(8, 4).construct_from_csc(g), g, inv)
    buf = DeviceBuffer(4 * g.n)
    try:
        roots = [int(r) for r in g.pick_roots(4, seed=6)]
        noalloc_bfs(blest, dg, roots[0], buf)  # set-up of this device on the first call
        before = free_hbm(blest)
        for src in roots:
            got, _ = noalloc_bfs(blest, dg, src, buf)
            np.testing.assert_array_equal(got, oracle.bfs(g.n, cp0, rows0, src))
        assert free_hbm(blest) == before
    finally:
        buf.free()


def test_noalloc_rejects(blest):
    lib = blest._lib.load()
    st = blest._lib.BfsStats()
    g = blest.CSC.rmat(10, 8, seed=24)
    dg = blest.DeviceGraph(blest.BVSS(8, 4).construct_from_csc(g), g, None)
    host = np.zeros(g.n, dtype=np.int32)
    assert lib.blest_bfs_noalloc(dg._h, 0, host.ctypes.data_as(blest._lib.I32P), C.byref(st), None) == 1
    small = DeviceBuffer(4 * (g.n - 1))
    try:
        assert lib.blest_bfs_noalloc(dg._h, 0, C.cast(small.ptr, blest._lib.I32P), C.byref(st), None) == 1
    finally:
        small.free()
    ok = DeviceBuffer(4 * g.n)
    try:
        assert lib.blest_bfs_noalloc(dg._h, g.n, C.cast(ok.ptr, blest._lib.I32P), C.byref(st), None) == 1
    finally:
        ok.free()
    # an edgeless graph of 17 M vertices with a permutation: its reordered levels alone
    # (68 MB) exceed the 64 MiB scratch
    n = 17_000_000
    big = blest.CSC.from_arrays(n, np.zeros(n + 1, dtype=np.uint32), np.zeros(0, dtype=np.uint32))
    dgb = blest.DeviceGraph(blest.BVSS(8, 4).construct_from_csc(big, full_padding=False), big,
                            np.arange(n, dtype=np.uint32))
    out = DeviceBuffer(4 * n)
    try:
        assert lib.blest_bfs_noalloc(dgb._h, 0, C.cast(out.ptr, blest._lib.I32P), C.byref(st), None) == 1
        assert b"scratch" in lib.blest_last_error()
    finally:
        out.free()
