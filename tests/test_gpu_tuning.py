"""Autotuner and threshold changes: the chosen configuration must give the oracle's
levels (tuning changes time only)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_then_parity(blest, oracle):
    print("case autotune rmat14 jaccard", flush=True)
    g = blest.CSC.rmat(14, 16, seed=6)
    cp, rows = g.arrays()
    inv = g.reorder(blest.Reorder.JACCARD, window=4096)
    b = blest.BVSS(8, 4).construct_from_csc(g)
    roots = g.pick_roots(4, seed=2)  # pick_roots on the reordered graph: map back to original IDs
    perm_new_to_old = np.argsort(inv)
    roots_orig = perm_new_to_old[roots]
    best, ms, trials = blest.autotune_bfs(b, g, inv, roots_orig)
    assert trials >= 4 and ms > 0
    assert best["blocks_per_cu"] in (1, 2, 4, 8)
    assert best["alpha"] in (4, 8, 15, 30, 60, 120) and best["beta"] in (6, 12, 18, 24, 48, 96)
    dg = blest.DeviceGraph(b, g, inv, **{k: best[k] for k in ("blocks_per_cu", "alpha", "beta")})
    k = blest.BVSSBFSKernel(dg)
    for r in roots_orig:
        np.testing.assert_array_equal(k.run(int(r)).levels, oracle.bfs(g.n, cp, rows, int(r)))


@pytest.mark.parametrize("alpha,beta,direction", [(1, 1, 0), (1000, 1000, 0), (15, 18, 1), (15, 18, 2)])
def test_set_thresholds(blest, oracle, alpha, beta, direction):
    print(f"case thresholds {alpha} {beta} {direction}", flush=True)
    g = blest.CSC.rmat(12, 8, seed=8)
    cp, rows = g.arrays()
    b = blest.BVSS(8, 4).construct_from_csc(g)
    dg = blest.DeviceGraph(b, g, None)
    dg.set_thresholds(alpha, beta, direction)
    k = blest.BVSSBFSKernel(dg)
    for r in g.pick_roots(3, seed=1):
        np.testing.assert_array_equal(k.run(int(r)).levels, oracle.bfs(g.n, cp, rows, int(r)))


def test_set_thresholds_rejects(blest):
    g = blest.CSC.rmat(10, 8, seed=8)
    b = blest.BVSS(8, 4).construct_from_csc(g)
    dg = blest.DeviceGraph(b, None, None)  # no CSR: bottom-up unavailable
    with pytest.raises(blest.BlestError):
        dg.set_thresholds(15, 18, 0)
    with pytest.raises(blest.BlestError):
        dg.set_thresholds(-1, 18, 1)
