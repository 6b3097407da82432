"""Gorder with its priority updates spread over threads (a partitioned queue whose links
carry stamps from one global sequence) returns exactly the single-thread order: the
greedy's ties depend on the order of the updates, which each partition keeps.  Graphs of
at least 2^16 vertices take the multi-threaded path."""
import numpy as np
import pytest


def _directed(blest, n, m, seed):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = (src + rng.zipf(1.6, m)) % n
    return blest.CSC.from_edges(n, src, dst, undirected=False)


CASES = [
    ("rmat16", lambda b: b.CSC.rmat(16, 16, seed=4, permute=True), 65536),
    ("rmat16_w5", lambda b: b.CSC.rmat(16, 16, seed=5, permute=True), 5),
    ("directed", lambda b: _directed(b, 70000, 700000, 3), 4096),
    ("grid", lambda b: b.CSC.grid(300, 0.9, 0.01, seed=2), 1024),
]


@pytest.mark.parametrize("name,make,window", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("threads", [2, 3, 8])
def test_gorder_threads_same_order(blest, name, make, window, threads):
    one = make(blest).reorder(blest.Reorder.GORDER, window=window, threads=1)
    many = make(blest).reorder(blest.Reorder.GORDER, window=window, threads=threads)
    assert np.array_equal(one, many)
    assert np.array_equal(np.sort(one), np.arange(len(one), dtype=one.dtype))


def test_gorder_jaccard_threads_same_order(blest):
    g1 = blest.CSC.rmat(16, 16, seed=6, permute=True)
    g8 = blest.CSC.rmat(16, 16, seed=6, permute=True)
    a = g1.reorder(blest.Reorder.GORDER_JACCARD, window=65536, threads=1)
    b = g8.reorder(blest.Reorder.GORDER_JACCARD, window=65536, threads=8)
    assert np.array_equal(a, b)
