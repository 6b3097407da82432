"""GPU parity: single-source BFS and CC through the C ABI against the oracle.
Levels must be bit-exact (distances, -1 unreachable) in the ORIGINAL vertex numbering;
CC labels must equal the smallest original vertex ID of each component."""
import sys

import numpy as np
import pytest
from conftest import progress

pytestmark = pytest.mark.gpu


def say(*a):
    progress(*a)


def pipeline(blest, make, reorder=None, window=1024, full_padding=None, **dev):
    """Reference pipeline: CSC -> (reorder) -> BVSS -> device graph.  Returns the original
    pattern (for the oracle), the inverse permutation and the kernels' device graph."""
    g = make()
    cp0, rows0 = g.arrays()
    inv = None
    if reorder is not None:
        inv = g.reorder(reorder, window=window)
    b = blest.BVSS(8, 4).construct_from_csc(g, full_padding=full_padding)
    dg = blest.DeviceGraph(b, g, inv, **dev)
    return g, cp0, rows0, inv, dg


GRAPHS = {
    "rmat12": lambda b: b.CSC.rmat(12, 16, seed=1),
    "rmat16": lambda b: b.CSC.rmat(16, 16, seed=1),
    "grid64": lambda b: b.CSC.grid(64, 0.9, 0.01, 1),
    "grid_path": lambda b: b.CSC.grid(200, 1.0, 0.0, 3),
    "er": lambda b: b.CSC.erdos_renyi(20000, 8.0, 2),
    "sparse_er": lambda b: b.CSC.erdos_renyi(5000, 1.2, 5),
}


@pytest.mark.parametrize("name,reorder", [("rmat12", None), ("rmat16", None), ("rmat12", "JACCARD"),
                                          ("rmat12", "GORDER_JACCARD"), ("grid64", "RCM"), ("grid_path", None),
                                          ("er", "JACCARD"), ("sparse_er", "RCM")])
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_bfs_parity(blest, oracle, name, reorder, direction):
    say(f"case bfs {name} reorder={reorder} direction={direction}")
    kind = None if reorder is None else getattr(blest.Reorder, reorder)
    g, cp0, rows0, inv, dg = pipeline(blest, lambda: GRAPHS[name](blest), kind, direction=direction)
    k = blest.BVSSBFSKernel(dg)
    n = g.n
    roots = list(g.pick_roots(4, seed=3)) + [0, n - 1]
    deg0 = np.diff(cp0.astype(np.int64))
    for src in roots:
        res = k.run(int(src))
        want = oracle.bfs(n, cp0, rows0, int(src))
        np.testing.assert_array_equal(res.levels, want, err_msg=f"src={src}")
        assert res.total_levels == want.max() + 1
        assert res.no_visited == int((want >= 0).sum())
        if direction != 1:
            assert res.edges_in_component == deg0[want >= 0].sum() // 2


@pytest.mark.parametrize("full_padding", [False, True])
def test_bfs_parity_both_paddings(blest, oracle, full_padding):
    say(f"case bfs padding={full_padding}")
    g, cp0, rows0, inv, dg = pipeline(blest, lambda: blest.CSC.rmat(13, 8, seed=9), blest.Reorder.RCM,
                                      full_padding=full_padding)
    k = blest.BVSSBFSKernel(dg)
    for src in g.pick_roots(5, seed=1):
        np.testing.assert_array_equal(k.run(int(src)).levels, oracle.bfs(g.n, cp0, rows0, int(src)))


def test_bfs_tiny_and_edge_cases(blest, oracle):
    say("case bfs tiny")
    cases = [
        (1, [], []),                                   # single vertex, no edges
        (2, [], []),                                   # two isolated vertices
        (3, [0, 1], [1, 2]),                           # path
        (65, list(range(64)), list(range(1, 65))),     # path across a 64-bit word boundary
        (130, [0] * 129, list(range(1, 130))),         # star with > 128 leaves (complete vset)
        (9, [0, 2, 4, 6], [1, 3, 5, 7]),               # four components + isolated vertex 8
    ]
    for n, s, d in cases:
        g = blest.CSC.from_edges(n, s, d)
        cp, rows = g.arrays()
        for fp in (False, True):
            b = blest.BVSS().construct_from_csc(g, full_padding=fp)
            k = blest.BVSSBFSKernel(b, g)
            for src in sorted({0, n // 2, n - 1}):
                res = k.run(src)
                np.testing.assert_array_equal(res.levels, oracle.bfs(n, cp, rows, src), err_msg=f"n={n} src={src}")


def test_bfs_directed(blest, oracle):
    say("case bfs directed")
    rng = np.random.default_rng(4)
    n = 3000
    e = rng.integers(0, n, size=(4 * n, 2))
    g = blest.CSC.from_edges(n, e[:, 0], e[:, 1], undirected=False)
    assert not g.check_symmetry()
    cp, rows = g.arrays()
    b = blest.BVSS().construct_from_csc(g, full_padding=False)
    for direction in (0, 1, 2):
        k = blest.BVSSBFSKernel(b, g, direction=direction, alpha=2.0)
        for src in (0, 5, 2999):
            np.testing.assert_array_equal(k.run(src).levels, oracle.bfs(n, cp, rows, src))


def test_bfs_repeated_launches_and_shapes(blest, oracle):
    """The persistent kernel's barrier state carries across launches; shapes vary."""
    say("case bfs repeated")
    g = blest.CSC.rmat(14, 16, seed=2)
    cp, rows = g.arrays()
    b = blest.BVSS().construct_from_csc(g)
    roots = g.pick_roots(6, seed=2)
    want = {int(s): oracle.bfs(g.n, cp, rows, int(s)) for s in roots}
    for block, bpc in ((256, 1), (512, 2), (1024, 1), (64, 4)):
        k = blest.BVSSBFSKernel(b, g, block_size=block, blocks_per_cu=bpc)
        for _ in range(2):
            for s in roots:
                np.testing.assert_array_equal(k.run(int(s)).levels, want[int(s)])


def test_bfs_rejects_bad_source(blest):
    g = blest.CSC.rmat(8, 8, seed=1)
    k = blest.BVSSBFSKernel(blest.BVSS().construct_from_csc(g), g)
    with pytest.raises(blest.BlestError):
        k.run(g.n)


@pytest.mark.parametrize("name,reorder", [("rmat12", "JACCARD"), ("grid64", "RCM"), ("sparse_er", None),
                                          ("er", "GORDER")])
@pytest.mark.parametrize("algo", [0, 1])
def test_cc_parity(blest, oracle, name, reorder, algo):
    say(f"case cc {name} {reorder} algo={algo}")
    kind = None if reorder is None else getattr(blest.Reorder, reorder)
    g, cp0, rows0, inv, dg = pipeline(blest, lambda: GRAPHS[name](blest), kind)
    res = blest.BVSSCCKernel(dg, algo=algo).run()
    want, k = oracle.cc(g.n, cp0, rows0)
    np.testing.assert_array_equal(res.components, want)
    assert res.no_components == k


def test_cc_tiny(blest, oracle):
    say("case cc tiny")
    for n, s, d in [(1, [], []), (5, [0, 3], [4, 1]), (70, list(range(0, 68, 2)), list(range(1, 69, 2)))]:
        g = blest.CSC.from_edges(n, s, d)
        cp, rows = g.arrays()
        for algo in (0, 1):
            res = blest.BVSSCCKernel(blest.BVSS().construct_from_csc(g), g, algo=algo).run()
            np.testing.assert_array_equal(res.components, oracle.cc(n, cp, rows)[0])
