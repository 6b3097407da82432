"""C-ABI entry points that validate their arguments before touching a device: defaults,
null handles and bad sizes must come back as status codes (no GPU needed)."""
import ctypes as C

import numpy as np
import pytest

INVALID = 1


@pytest.fixture(scope="module")
def lib(blest):
    return blest._lib.load()


def test_device_opts_defaults(blest):
    o = blest.device_opts()
    assert (o.blocks_per_cu, o.alpha, o.beta, o.direction, o.keep_csr) == (4, 15.0, 18.0, 0, 1)
    assert o.use_graphs == 0  # captured level chunks replayed once a traversal was deep


def test_server_opts_defaults(blest, lib):
    o = blest._lib.ServerOpts()
    lib.blest_server_opts_default(C.byref(o))
    assert (o.workers, o.max_batch, o.hbm_budget, o.capacity, o.min_multi) == (4, 64, 0, 65536, 8)


def test_server_entry_points_reject_null(blest, lib):
    out = C.c_void_p()
    assert lib.blest_server_create(None, 1, None, C.byref(out)) == INVALID
    arr = (C.c_void_p * 1)(None)
    assert lib.blest_server_create(arr, 0, None, C.byref(out)) == INVALID
    assert lib.blest_server_create(arr, 1, None, None) == INVALID
    t = C.c_uint64()
    assert lib.blest_server_submit(None, 0, 0, 0, None, C.byref(t)) == INVALID
    assert lib.blest_server_wait(None, 1, 0, None) == INVALID
    assert lib.blest_server_drain(None) == INVALID
    lib.blest_server_free(None)
    assert "null" in lib.blest_last_error().decode()


def test_server_rejects_bad_opts_before_device_work(blest, lib):
    o = blest._lib.ServerOpts()
    lib.blest_server_opts_default(C.byref(o))
    fake = (C.c_void_p * 1)(None)  # null graph: rejected, never dereferenced
    out = C.c_void_p()
    for field, value in (("workers", 0), ("workers", 65), ("max_batch", 0), ("max_batch", 129), ("capacity", 0)):
        bad = blest._lib.ServerOpts()
        C.memmove(C.byref(bad), C.byref(o), C.sizeof(o))
        setattr(bad, field, value)
        assert lib.blest_server_create(fake, 1, C.byref(bad), C.byref(out)) == INVALID, field


def test_tuning_and_tracing_reject_null(blest, lib):
    assert lib.blest_device_graph_set_thresholds(None, 15.0, 18.0, 0) == INVALID
    best = blest._lib.DeviceOpts()
    roots = np.zeros(1, dtype=np.uint32)
    assert lib.blest_autotune_bfs(None, None, None, roots.ctypes.data_as(blest._lib.U32P), 1, None,
                                  C.byref(best), None, None) == INVALID
    g = blest.CSC.rmat(6, 4, seed=1)
    b = blest.BVSS(8, 4).construct_from_csc(g)
    assert lib.blest_autotune_bfs(b._h, g._h, None, roots.ctypes.data_as(blest._lib.U32P), 0, None,
                                  C.byref(best), None, None) == INVALID  # no roots
    cnt = C.c_uint32()
    assert lib.blest_bfs_trace(None, 0, None, None, None, 0, C.byref(cnt)) == INVALID
    q = C.c_uint64()
    assert lib.blest_query_footprint(None, 1, C.byref(q)) == INVALID
    assert lib.blest_host_alloc(16, None) == INVALID
    lib.blest_host_free(None)


def test_server_wrapper_validates_graph_index(blest):
    # Server.submit checks the index before calling into the library
    srv = object.__new__(blest.Server)
    srv._n = [10]
    srv._lib = blest._lib.load()
    srv._h = C.c_void_p(0)
    srv._out = {}
    with pytest.raises(blest.BlestError):
        srv.submit(3, "bfs", 0)
