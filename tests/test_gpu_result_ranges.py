"""Result destinations in device memory: every entry point that copies results (BFS,
multi-source BFS rows and blocks, CC labels, the graph batch) checks that a device
destination holds the whole result before anything runs, and returns
BLEST_ERR_INVALID_ARGUMENT for one that is too small; a large enough device buffer
receives the same result as host memory."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HIP_D2H = 2


class DeviceBuffer:
    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), nbytes) == 0

    def i32(self):
        return C.cast(self.ptr, C.POINTER(C.c_int32))

    def u32(self):
        return C.cast(self.ptr, C.POINTER(C.c_uint32))

    def numpy(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, HIP_D2H) == 0
        return out

    def __del__(self):
        if self.ptr.value:
            self.hip.hipFree(self.ptr)


def test_result_ranges(blest, oracle):
    print("case result ranges", flush=True)
    lib = blest._lib.load()
    g = blest.CSC.rmat(11, 8, seed=31)
    cp, rows = g.arrays()
    dg = blest.DeviceGraph(blest.BVSS(8, 4).construct_from_csc(g), g, None)
    n = g.n
    q = blest._Query(dg, 8)
    st = blest._lib.BfsStats()
    r = int(g.pick_roots(1, seed=2)[0])
    small, big = DeviceBuffer(4 * n - 4), DeviceBuffer(4 * n * 8)
    # single-source BFS
    assert lib.blest_bfs(q._h, r, small.i32(), C.byref(st)) == 1
    blest.check(lib.blest_bfs(q._h, r, big.i32(), C.byref(st)))
    np.testing.assert_array_equal(big.numpy(np.int32, n), oracle.bfs(n, cp, rows, r))
    # multi-source block: k x n
    roots = np.array([int(x) for x in g.pick_roots(8, seed=3)], dtype=np.uint32)
    o = blest._lib.MbfsOpts(0, 64)
    assert lib.blest_mbfs(q._h, blest._u32p(roots), 8, small.i32(), C.byref(o), C.byref(st)) == 1
    mid = DeviceBuffer(4 * n * 8 - 4)
    assert lib.blest_mbfs(q._h, blest._u32p(roots), 8, mid.i32(), C.byref(o), C.byref(st)) == 1
    blest.check(lib.blest_mbfs(q._h, blest._u32p(roots), 8, big.i32(), C.byref(o), C.byref(st)))
    got = big.numpy(np.int32, 8 * n).reshape(8, n)
    for j, rt in enumerate(roots):
        np.testing.assert_array_equal(got[j], oracle.bfs(n, cp, rows, int(rt)))
    # CC labels
    cs = blest._lib.CcStats()
    assert lib.blest_cc(q._h, 0, C.cast(small.ptr, blest._lib.U32P), C.byref(cs)) == 1
    blest.check(lib.blest_cc(q._h, 0, big.u32(), C.byref(cs)))
    labels, count = oracle.cc(n, cp, rows)
    np.testing.assert_array_equal(big.numpy(np.uint32, n), labels)
    assert cs.no_components == count
    # graph batch: the concatenation of every graph's levels
    graphs = [blest.CSC.erdos_renyi(2000 + 500 * i, 6.0, seed=40 + i) for i in range(3)]
    gb = blest.GraphBatch(graphs)
    total = sum(x.n for x in graphs)
    broots = np.array([int(x.pick_roots(1, seed=1)[0]) for x in graphs], dtype=np.uint32)
    ms = C.c_double()
    short = DeviceBuffer(4 * total - 4)
    assert lib.blest_graph_batch_bfs(gb._h, blest._u32p(broots), short.i32(), C.byref(ms)) == 1
    whole = DeviceBuffer(4 * total)
    blest.check(lib.blest_graph_batch_bfs(gb._h, blest._u32p(broots), whole.i32(), C.byref(ms)))
    got = whole.numpy(np.int32, total)
    off = 0
    for x, rt in zip(graphs, broots):
        xc, xr = x.arrays()
        np.testing.assert_array_equal(got[off:off + x.n], oracle.bfs(x.n, xc, xr, int(rt)))
        off += x.n
