"""GPU parity at sizes where every producer segment holds several bitmap words.

A device graph cuts its 64-bit bitmap words into at most 4 x CUs segments, so graphs
with more than 65,536 vertices give each segment several words (chunked list writers,
wave word ranges, multi-chunk segment lists) and, with n not a multiple of 64, a
partial last word.  Smaller graphs have one word per segment.  Every result is compared
bit-exactly with the oracle: BFS levels in original IDs (-1 unreachable), each
multi-source row against its own single-source BFS, canonical CC labels."""
import sys

import numpy as np
import pytest
from conftest import progress

pytestmark = pytest.mark.gpu


def say(*a):
    progress(*a)


# name -> (generator, reordering or None, window)
MAKERS = {
    "rmat18_jaccard": (lambda b: b.CSC.rmat(18, 16, seed=11), "JACCARD", 8192),
    "rmat18": (lambda b: b.CSC.rmat(18, 16, seed=12), None, 0),
    "grid512_rcm": (lambda b: b.CSC.grid(512, 0.9, 0.01, 3), "RCM", 0),
    "er200003": (lambda b: b.CSC.erdos_renyi(200003, 4.0, 9), None, 0),  # a partial last word
}
_cache = {}


def prepared(blest, name):
    """(original col_ptrs, rows, inverse permutation, BVSS, reordered CSC); built once."""
    if name not in _cache:
        make, reorder, window = MAKERS[name]
        g = make(blest)
        cp, rows = g.arrays()  # original IDs, copied before the in-place reordering
        inv = None if reorder is None else g.reorder(getattr(blest.Reorder, reorder), window=window or 65536)
        b = blest.BVSS(8, 4).construct_from_csc(g)
        _cache[name] = (cp, rows, inv, b, g)
    return _cache[name]


def test_segments_hold_several_words(blest):
    cp, rows, inv, b, g = prepared(blest, "rmat18")
    dg = blest.DeviceGraph(b, g, inv)
    info = dg.info()
    words = (info["n"] + 63) // 64
    assert words > info["grid_blocks"], "the graph must give each segment more than one word"


@pytest.mark.parametrize("name", list(MAKERS))
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_bfs_parity_multiword(blest, oracle, name, direction):
    say(f"case scale bfs {name} direction={direction}")
    cp, rows, inv, b, g = prepared(blest, name)
    n = len(cp) - 1
    dg = blest.DeviceGraph(b, g, inv, direction=direction)
    k = blest.BVSSBFSKernel(dg)
    roots = [int(r) for r in blest.CSC.from_arrays(n, cp, rows).pick_roots(3, seed=5)] + [n - 1]
    deg = np.diff(cp.astype(np.int64))
    for src in roots:
        res = k.run(src)
        want = oracle.bfs(n, cp, rows, src)
        np.testing.assert_array_equal(res.levels, want, err_msg=f"{name} src={src}")
        assert res.no_visited == int((want >= 0).sum())
        if direction != 1:
            assert res.edges_in_component == deg[want >= 0].sum() // 2


@pytest.mark.parametrize("name", ["rmat18_jaccard", "grid512_rcm", "er200003"])
@pytest.mark.parametrize("k,direction", [(9, 0), (64, 1), (128, 0), (100, 2)])
def test_mbfs_parity_multiword(blest, oracle, name, k, direction):
    say(f"case scale mbfs {name} k={k} direction={direction}")
    cp, rows, inv, b, g = prepared(blest, name)
    n = len(cp) - 1
    rng = np.random.default_rng(k + 7 * direction)
    roots = rng.integers(0, n, size=k).astype(np.uint32)
    kern = blest.BVSSMBFSKernel(b, g, inv, batch=128 if k > 64 else 64, direction=direction)
    res = kern.run(roots)
    assert len(res.level_arrays) == k
    for j, r in enumerate(roots):
        np.testing.assert_array_equal(res.level_arrays[j], oracle.bfs(n, cp, rows, int(r)),
                                      err_msg=f"{name} row {j} root {r}")


@pytest.mark.parametrize("name", list(MAKERS))
@pytest.mark.parametrize("algo", [0, 1])
def test_cc_parity_multiword(blest, oracle, name, algo):
    say(f"case scale cc {name} algo={algo}")
    cp, rows, inv, b, g = prepared(blest, name)
    n = len(cp) - 1
    res = blest.BVSSCCKernel(b, g, inv, algo=algo).run()
    labels, count = oracle.cc(n, cp, rows)
    np.testing.assert_array_equal(res.components, labels, err_msg=name)
    assert res.no_components == count


@pytest.mark.parametrize("name", ["rmat18_jaccard", "rmat18"])
def test_bfs_hub_root_chunked_lists(blest, oracle, name):
    """Roots in rsets of more than 16 vsets: the source's own rset is listed as chunk
    entries by the init kernel, and its neighbours' hub rsets by the update kernel."""
    say(f"case scale bfs hub roots {name}")
    cp, rows, inv, b, g = prepared(blest, name)
    n = len(cp) - 1
    rp = b.arrays()["real_ptrs"].astype(np.int64)
    vsets = np.diff(rp)
    new_of_old = np.arange(n) if inv is None else np.asarray(inv)
    deg = np.diff(cp.astype(np.int64))
    hubs = [int(v) for v in np.argsort(deg)[::-1][:64] if vsets[new_of_old[v] >> 3] > 16][:4]
    assert hubs, "the graph must have roots in rsets of more than 16 vsets"
    k = blest.BVSSBFSKernel(blest.DeviceGraph(b, g, inv))
    for src in hubs:
        res = k.run(src)
        np.testing.assert_array_equal(res.levels, oracle.bfs(n, cp, rows, src), err_msg=f"{name} hub src={src}")
