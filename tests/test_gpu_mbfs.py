"""GPU parity of the multi-source BFS (each root's levels equal a single-source oracle
BFS) and of the many-graphs batch (each graph's levels equal the oracle on it)."""
import sys

import numpy as np
import pytest
from conftest import progress

pytestmark = pytest.mark.gpu


def say(*a):
    progress(*a)


@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("k", [1, 5, 64, 100, 128])
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_mbfs_parity(blest, oracle, k, direction, algo):
    """algo 0: bit-parallel VALU product; algo 1: the same product as an i8 MFMA."""
    say(f"case mbfs k={k} direction={direction} algo={algo}")
    g = blest.CSC.rmat(12, 16, seed=5)
    cp, rows = g.arrays()
    inv = g.reorder(blest.Reorder.JACCARD, window=1024)
    b = blest.BVSS().construct_from_csc(g)
    rng = np.random.default_rng(k)
    roots = rng.integers(0, g.n, size=k).astype(np.uint32)  # duplicates and isolated roots allowed
    kern = blest.BVSSMBFSKernel(b, g, inv, batch=128 if k > 64 else 64, algo=algo, direction=direction)
    res = kern.run(roots)
    assert len(res.level_arrays) == k
    for j, r in enumerate(roots):
        np.testing.assert_array_equal(res.level_arrays[j], oracle.bfs(g.n, cp, rows, int(r)), err_msg=f"root {r}")


@pytest.mark.parametrize("algo", [0, 1])
def test_mbfs_grid_and_many_batches(blest, oracle, algo):
    say(f"case mbfs grid algo={algo}")
    g = blest.CSC.grid(80, 0.9, 0.01, 4)
    cp, rows = g.arrays()
    inv = g.reorder(blest.Reorder.RCM)
    b = blest.BVSS().construct_from_csc(g)
    roots = g.pick_roots(150, seed=2)  # two launches: 128 + 22
    res = blest.BVSSMBFSKernel(b, g, inv, batch=128, algo=algo).run(roots)
    for j in (0, 1, 77, 127, 128, 149):
        np.testing.assert_array_equal(res.level_arrays[j], oracle.bfs(g.n, cp, rows, int(roots[j])))


@pytest.mark.parametrize("algo", [0, 1])
def test_mbfs_directed(blest, oracle, algo):
    say(f"case mbfs directed algo={algo}")
    rng = np.random.default_rng(8)
    n = 2000
    e = rng.integers(0, n, size=(5 * n, 2))
    g = blest.CSC.from_edges(n, e[:, 0], e[:, 1], undirected=False)
    cp, rows = g.arrays()
    b = blest.BVSS().construct_from_csc(g, full_padding=False)
    roots = rng.integers(0, n, size=70).astype(np.uint32)
    res = blest.BVSSMBFSKernel(b, g, batch=128, algo=algo, alpha=2.0).run(roots)
    for j, r in enumerate(roots):
        np.testing.assert_array_equal(res.level_arrays[j], oracle.bfs(n, cp, rows, int(r)))


def test_mbfs_rejects_bad_input(blest):
    g = blest.CSC.rmat(8, 8, seed=1)
    k = blest.BVSSMBFSKernel(blest.BVSS().construct_from_csc(g), g, batch=64)
    with pytest.raises(blest.BlestError):
        k.run(np.array([g.n], dtype=np.uint32))
    k2 = blest.BVSSMBFSKernel(blest.BVSS().construct_from_csc(g), g, batch=64, algo=2)
    with pytest.raises(blest.BlestError):
        k2.run(np.array([0], dtype=np.uint32))


def test_graph_batch_parity(blest, oracle):
    say("case graph batch")
    rng = np.random.default_rng(3)
    sizes = [1, 2, 63, 64, 65, 1000] + list(rng.integers(1024, 20000, size=40))
    graphs = [blest.CSC.erdos_renyi(int(n), 8.0 if n > 64 else 1.5, seed=i + 1) for i, n in enumerate(sizes)]
    roots = np.array([rng.integers(0, g.n) for g in graphs], dtype=np.uint32)
    gb = blest.GraphBatch(graphs)
    levels, secs = gb.bfs(roots)
    assert secs > 0
    for g, r, lv in zip(graphs, roots, levels):
        cp, rows = g.arrays()
        np.testing.assert_array_equal(lv, oracle.bfs(g.n, cp, rows, int(r)))
    levels2, _ = gb.bfs(roots[::-1].copy() % np.array([g.n for g in graphs], dtype=np.uint32))
    for g, r, lv in zip(graphs, roots[::-1] % np.array([g.n for g in graphs]), levels2):
        cp, rows = g.arrays()
        np.testing.assert_array_equal(lv, oracle.bfs(g.n, cp, rows, int(r)))
