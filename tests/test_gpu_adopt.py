"""GPU parity through the drop-in path a reference plugin takes: a BVSS's five arrays
adopted as the reference's accessors return them (BVSS.cuh:54-68) and the bottom-up CSR
as BVSS::getCSR() holds it (the transpose of the CSC, BVSS.cuh:187), instead of the
library's own construction from the CSC.  BFS levels and canonical CC labels are
compared with the oracle on the original graph."""
import sys

import numpy as np
import pytest
from conftest import progress

pytestmark = pytest.mark.gpu


def say(*a):
    progress(*a)


def transpose_arrays(n, cp, rows):
    """Column i of the result lists the j of every edge j -> i (sorted)."""
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(cp.astype(np.int64)))
    order = np.lexsort((cols, rows.astype(np.int64)))
    t_rows = cols[order].astype(np.uint32)
    t_cp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(t_cp, rows.astype(np.int64) + 1, 1)
    return np.cumsum(t_cp).astype(np.uint32), t_rows


def adopted_graph(blest, g, inv, full_padding=None, **dev):
    b = blest.BVSS(8, 4).construct_from_csc(g, full_padding=full_padding)
    arr, info = b.arrays(), b.info()
    a = blest.BVSS.adopt(g.n, 8, 4, arr["slice_set_ptrs"], arr["virtual_to_real"], arr["real_ptrs"],
                         arr["row_ids"], arr["masks"], info["update_divergence"], bool(info["full_padding"]))
    cp, rows = g.arrays()
    csr = blest.CSC.from_arrays(g.n, *transpose_arrays(g.n, cp, rows))
    return blest.DeviceGraph(a, inv_perm=inv, csr=csr, **dev)


CASES = {
    "rmat13_jaccard": (lambda b: b.CSC.rmat(13, 16, seed=8), "JACCARD"),
    "grid96_rcm": (lambda b: b.CSC.grid(96, 0.9, 0.01, 4), "RCM"),
    "er_natural": (lambda b: b.CSC.erdos_renyi(30000, 6.0, 3), None),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_adopted_bfs_and_cc(blest, oracle, name, direction):
    say(f"case adopt {name} direction={direction}")
    make, reorder = CASES[name]
    g = make(blest)
    cp0, rows0 = g.arrays()
    inv = None if reorder is None else g.reorder(getattr(blest.Reorder, reorder), window=1024)
    dg = adopted_graph(blest, g, inv, direction=direction)
    assert dg.info()["symmetric"] == 1 and dg.info()["has_csr"] == 1
    k = blest.BVSSBFSKernel(dg)
    for src in [int(r) for r in g.pick_roots(4, seed=2)]:
        np.testing.assert_array_equal(k.run(src).levels, oracle.bfs(g.n, cp0, rows0, src), err_msg=f"src={src}")
    cc = blest.BVSSCCKernel(dg).run()
    labels, count = oracle.cc(g.n, cp0, rows0)
    np.testing.assert_array_equal(cc.components, labels)
    assert cc.no_components == count
    m = blest.BVSSMBFSKernel(dg, batch=64).run(g.pick_roots(64, seed=9))
    for j, r in enumerate(g.pick_roots(64, seed=9)[::13]):
        np.testing.assert_array_equal(m.level_arrays[13 * j], oracle.bfs(g.n, cp0, rows0, int(r)))


@pytest.mark.parametrize("direction", [1, 2])
def test_adopted_directed_graph(blest, oracle, direction):
    """A directed graph: the CSR handed over is the transpose, so bottom-up pulls
    in-neighbours exactly as the library's own derivation does."""
    say(f"case adopt directed direction={direction}")
    rng = np.random.default_rng(5)
    n = 6000
    src, dst = rng.integers(0, n, 40000), rng.integers(0, n, 40000)
    g = blest.CSC.from_edges(n, src, dst, undirected=False)
    cp0, rows0 = g.arrays()
    dg = adopted_graph(blest, g, None, direction=direction)
    assert dg.info()["symmetric"] == 0
    k = blest.BVSSBFSKernel(dg)
    for s in [int(r) for r in g.pick_roots(5, seed=1)]:
        np.testing.assert_array_equal(k.run(s).levels, oracle.bfs(n, cp0, rows0, s), err_msg=f"src={s}")
