"""GPU parity of the request server: every served result (alone, in a multi-source
batch, under eviction pressure) equals the oracle's BFS levels / canonical CC labels."""
import sys

import numpy as np
import pytest
from conftest import progress

pytestmark = pytest.mark.gpu


def say(*a):
    progress(*a)


def make_graphs(blest):
    out = []
    for name, make, reorder in [("rmat12", lambda: blest.CSC.rmat(12, 16, seed=4), blest.Reorder.JACCARD),
                                ("grid64", lambda: blest.CSC.grid(64, 0.9, 0.01, 2), blest.Reorder.RCM),
                                ("er", lambda: blest.CSC.erdos_renyi(6000, 3.0, 7), None)]:
        g = make()
        cp, rows = g.arrays()
        inv = None if reorder is None else g.reorder(reorder, window=1024)
        b = blest.BVSS(8, 4).construct_from_csc(g)
        out.append((name, cp, rows, blest.DeviceGraph(b, g, inv)))
    return out


def check(oracle, graphs, reqs, results):
    batched = 0
    for (gi, kind, root), (out, res) in zip(reqs, results):
        name, cp, rows, _ = graphs[gi]
        n = len(cp) - 1
        assert res["status"] == 0
        if kind == "bfs":
            want = oracle.bfs(n, cp, rows, root)
            np.testing.assert_array_equal(out, want, err_msg=f"{name} root={root} batch={res['batch_size']}")
            assert res["total_levels"] == want.max() + 1 and res["no_visited"] == int((want >= 0).sum())
        else:
            labels, count = oracle.cc(n, cp, rows)
            np.testing.assert_array_equal(out, labels, err_msg=f"{name} cc")
            assert res["no_visited"] == count
        batched += res["batch_size"] > 1
    return batched


@pytest.mark.parametrize("workers,max_batch,min_multi", [(1, 1, 8), (4, 64, 8), (2, 128, 2), (3, 64, 1)])
def test_server_mixed_stream(blest, oracle, workers, max_batch, min_multi):
    say(f"case server workers={workers} max_batch={max_batch} min_multi={min_multi}")
    graphs = make_graphs(blest)
    srv = blest.Server([g[3] for g in graphs], workers=workers, max_batch=max_batch, min_multi=min_multi)
    rng = np.random.default_rng(workers * 131 + max_batch)
    reqs = []
    for i in range(300):
        gi = int(rng.integers(len(graphs)))
        n = len(graphs[gi][1]) - 1
        kind = "cc" if i % 37 == 5 else "bfs"
        reqs.append((gi, kind, int(rng.integers(n)) if kind == "bfs" else 0))
    tickets = [srv.submit(gi, kind, root) for gi, kind, root in reqs]
    results = [srv.wait(t) for t in tickets]
    srv.drain()
    batched = check(oracle, graphs, reqs, results)
    if max_batch == 1:
        assert batched == 0
    say(f"  batched requests: {batched}")


def test_server_eviction_budget(blest, oracle):
    """A budget that holds one batch query at a time forces LRU eviction between graphs."""
    say("case server eviction")
    graphs = make_graphs(blest)
    lib = blest._lib.load()
    import ctypes as C
    need = []
    for g in graphs:
        b = C.c_uint64()
        blest._lib.check(lib.blest_query_footprint(g[3]._h, 16, C.byref(b)))
        need.append(b.value)
    srv = blest.Server([g[3] for g in graphs], workers=3, max_batch=16, hbm_budget=max(need), min_multi=2)
    rng = np.random.default_rng(9)
    reqs = [(i % 3, "bfs", int(rng.integers(len(graphs[i % 3][1]) - 1))) for i in range(120)]
    tickets = [srv.submit(*r) for r in reqs]
    results = [srv.wait(t) for t in tickets]
    check(oracle, graphs, reqs, results)


def test_server_errors(blest):
    say("case server errors")
    graphs = make_graphs(blest)
    dgs = [g[3] for g in graphs]
    with pytest.raises(blest.BlestError):
        blest.Server(dgs, workers=0)
    with pytest.raises(blest.BlestError):
        blest.Server(dgs, max_batch=129)
    srv = blest.Server(dgs, workers=1, hbm_budget=1024)  # no query fits
    t = srv.submit(0, "bfs", 0)
    with pytest.raises(blest.BlestError) as e:
        srv.wait(t)
    assert e.value.status == 6  # BLEST_ERR_OUT_OF_DEVICE_MEMORY
    with pytest.raises(blest.BlestError):
        srv.submit(0, "bfs", 1 << 30)  # root out of range
    with pytest.raises(blest.BlestError):
        srv.submit(5, "bfs", 0)
    with pytest.raises(blest.BlestError):
        srv.wait(123456)  # unknown ticket
