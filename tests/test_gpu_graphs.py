"""Replayed level chunks (captured HIP graphs): every replay must give the oracle's
levels, for shallow and deep traversals, single- and multi-source, and across
threshold changes (which recapture)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def build(blest, g, reorder=None, **dev):
    cp, rows = g.arrays()
    inv = None if reorder is None else g.reorder(reorder, window=1024)
    b = blest.BVSS(8, 4).construct_from_csc(g)
    return cp, rows, inv, blest.DeviceGraph(b, g, inv, **dev)


@pytest.mark.parametrize("name,make,reorder", [
    ("rmat12", lambda B: B.CSC.rmat(12, 16, seed=11), None),
    ("grid_path", lambda B: B.CSC.grid(160, 1.0, 0.0, 1), None),   # diameter 318
    ("grid_rcm", lambda B: B.CSC.grid(96, 0.9, 0.0, 5), "RCM"),
])
@pytest.mark.parametrize("use_graphs", [1, 0])
def test_bfs_graph_replay(blest, oracle, name, make, reorder, use_graphs):
    print(f"case graphs bfs {name} use_graphs={use_graphs}", flush=True)
    g = make(blest)
    cp, rows, inv, dg = build(blest, g, getattr(blest.Reorder, reorder) if reorder else None, use_graphs=use_graphs)
    k = blest.BVSSBFSKernel(dg)
    n = g.n
    for src in [0, n - 1, n // 2, 7, 0]:
        res = k.run(int(src))
        want = oracle.bfs(n, cp, rows, int(src))
        np.testing.assert_array_equal(res.levels, want, err_msg=f"src={src}")
        assert res.total_levels == want.max() + 1


def test_bfs_graph_recapture_on_threshold_change(blest, oracle):
    print("case graphs recapture", flush=True)
    g = blest.CSC.grid(128, 1.0, 0.0, 2)
    cp, rows, inv, dg = build(blest, g, None, use_graphs=1)
    k = blest.BVSSBFSKernel(dg)
    for direction in (0, 1, 2, 0):
        dg.set_thresholds(15, 18, direction)
        for src in (3, g.n - 5):
            np.testing.assert_array_equal(k.run(src).levels, oracle.bfs(g.n, cp, rows, src))


@pytest.mark.parametrize("k", [5, 64, 100])
def test_mbfs_graph_replay(blest, oracle, k):
    print(f"case graphs mbfs k={k}", flush=True)
    g = blest.CSC.grid(80, 1.0, 0.0, 3)
    cp, rows, inv, dg = build(blest, g, None, use_graphs=1)
    m = blest.BVSSMBFSKernel(dg, batch=128)
    srcs = list(range(0, g.n, max(1, g.n // k)))[:k]
    for _ in range(2):
        res = m.run(srcs)
        for j, s in enumerate(srcs):
            np.testing.assert_array_equal(res.level_arrays[j], oracle.bfs(g.n, cp, rows, s), err_msg=f"root {s}")
