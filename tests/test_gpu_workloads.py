"""GPU parity on the benchmark's own workloads, at their full sizes.

Each workload is prepared exactly as the benchmark prepares it (generator, seed,
reordering, window, BVSS padding, Benchmark.cuh:149-229) and the GPU results are compared
bit-exactly with the oracle on the original graph: BFS levels in original IDs (-1
unreachable) for the benchmark's own roots, canonical CC labels, and multi-source rows.
R-MAT 22 under Gorder+Jaccard needs its cached ordering (computing it takes far longer
than a test may); without the cache those cases are skipped."""
import sys
import time

import numpy as np
import pytest
from conftest import progress

pytestmark = pytest.mark.gpu

_prepared = {}


def say(*a):
    progress(*a)


def workload(name):
    from blest_amd import workloads
    if name not in _prepared:
        w = workloads.WORKLOADS[name]
        if w.reorder.name == "GORDER_JACCARD":
            stem = f"{w.name}_{int(w.reorder)}_{w.window}_v{workloads.ORDERING_VERSION}_"
            if not any(workloads.CACHE.glob(stem + "*_ordering.bin")):
                pytest.skip(f"{name}: no cached ordering under {workloads.CACHE}")
        _prepared[name] = workloads.prepare(name, keep_original=True, log=say)
    return _prepared[name]


def device_graph(blest, p, **opts):
    return blest.DeviceGraph(p.bvss, p.csc, p.inv_perm, **opts)


# workload -> roots checked against the oracle (all of the benchmark's roots)
BFS_CASES = [("rmat16_natural", 16), ("rmat22_gorder", 64), ("grid4096_rcm", 64)]


@pytest.mark.parametrize("name,count", BFS_CASES)
def test_bfs_benchmark_roots(blest, oracle, name, count):
    say(f"case workload bfs {name}")
    p = workload(name)
    n = p.csc.n
    cp, rows = p.orig_col_ptrs, p.orig_rows
    deg = np.diff(cp.astype(np.int64))
    k = blest.BVSSBFSKernel(device_graph(blest, p))
    roots = [int(r) for r in p.roots[:count]]
    for src, want in oracle.bfs_many(n, cp, rows, roots, workers=8):
        say(f"  root {src}")
        res = k.run(src)
        np.testing.assert_array_equal(res.levels, want, err_msg=f"{name} src={src}")
        assert res.no_visited == int((want >= 0).sum())
        assert res.edges_in_component == deg[want >= 0].sum() // 2


@pytest.mark.parametrize("name", ["rmat16_natural", "rmat22_gorder", "grid4096_rcm"])
def test_cc_benchmark_graph(blest, oracle, name):
    say(f"case workload cc {name}")
    p = workload(name)
    res = blest.BVSSCCKernel(p.bvss, p.csc, p.inv_perm).run()
    labels, count = oracle.cc(p.csc.n, p.orig_col_ptrs, p.orig_rows)
    np.testing.assert_array_equal(res.components, labels, err_msg=name)
    assert res.no_components == count


def test_mbfs_128_roots_rmat22(blest, oracle):
    """One 128-root launch on R-MAT 22: every row equals the single-source GPU BFS of its
    root, and 16 rows (spread over both 64-bit halves of the batch) equal the oracle."""
    say("case workload mbfs rmat22_gorder 128 roots")
    p = workload("rmat22_gorder")
    n = p.csc.n
    roots = blest.CSC.from_arrays(n, p.orig_col_ptrs, p.orig_rows).pick_roots(128, seed=3)
    res = blest.BVSSMBFSKernel(p.bvss, p.csc, p.inv_perm, batch=128).run(roots)
    assert len(res.level_arrays) == 128
    single = blest.BVSSBFSKernel(device_graph(blest, p))
    for j, r in enumerate(roots):
        np.testing.assert_array_equal(res.level_arrays[j], single.run(int(r)).levels,
                                      err_msg=f"row {j} root {r}: multi-source vs single-source")
    # the same launch through the i8 MFMA product: every row equal
    mf = blest.BVSSMBFSKernel(p.bvss, p.csc, p.inv_perm, batch=128, algo=1).run(roots)
    for j in range(128):
        np.testing.assert_array_equal(mf.level_arrays[j], res.level_arrays[j], err_msg=f"row {j}: MFMA vs VALU")
    del mf
    rows_checked = [0, 9, 18, 31, 40, 52, 63, 64, 71, 80, 95, 100, 111, 119, 126, 127]
    for j, (r, want) in zip(rows_checked, oracle.bfs_many(n, p.orig_col_ptrs, p.orig_rows,
                                                          [int(roots[j]) for j in rows_checked])):
        np.testing.assert_array_equal(res.level_arrays[j], want, err_msg=f"row {j} root {r}: multi-source vs oracle")


def test_rmat26_bfs_cc_and_1024_root_mbfs(blest, oracle):
    """R-MAT 26 (67 M vertices, 2.1 B directed edges), natural order, as benchmarked:
    BFS from 4 of the benchmark's roots and CC against the oracle, then the benchmark's
    1,024 roots as 8 launches of 128: in every launch one row of each 64-bit half equals
    the oracle (16 rows in all), and every row reaches exactly its root's component
    (vertex count against the oracle's component sizes)."""
    say("case workload rmat26 bfs / cc / 1024-root mbfs")
    t0 = time.time()
    p = workload("rmat26_mbfs")
    n = p.csc.n
    cp, rows = p.orig_col_ptrs, p.orig_rows
    say(f"  prepared in {time.time() - t0:.0f}s")
    dg = device_graph(blest, p)
    single = blest.BVSSBFSKernel(dg)
    roots = [int(r) for r in p.roots]
    assert len(roots) == 1024
    for src, want in oracle.bfs_many(n, cp, rows, roots[:4], workers=4):
        say(f"  bfs root {src}")
        np.testing.assert_array_equal(single.run(src).levels, want, err_msg=f"src={src}")
    del single
    say(f"  cc ({time.time() - t0:.0f}s)")
    res = blest.BVSSCCKernel(dg).run()
    labels, count = oracle.cc(n, cp, rows)
    np.testing.assert_array_equal(res.components, labels)
    assert res.no_components == count
    del res
    comp_size = np.bincount(labels, minlength=n)
    mb = blest.BVSSMBFSKernel(dg, batch=128)
    for b in range(8):
        batch = np.asarray(roots[128 * b:128 * (b + 1)], dtype=np.uint32)
        lv = mb.run(batch).level_arrays
        picks = [(7 * b + 3) % 64, 64 + (11 * b + 5) % 64]
        for j, (r, want) in zip(picks, oracle.bfs_many(n, cp, rows, [int(batch[j]) for j in picks], workers=2)):
            np.testing.assert_array_equal(lv[j], want, err_msg=f"launch {b} row {j} root {r}")
        for j, r in enumerate(batch):
            assert np.count_nonzero(lv[j] >= 0) == comp_size[labels[r]], f"launch {b} row {j} root {r}"
        say(f"  mbfs launch {b}: rows {picks} equal the oracle ({time.time() - t0:.0f}s)")
        del lv


def test_graph_batch_er4096(blest, oracle):
    """The many-small-graphs workload at full size: 4,096 Erdos-Renyi graphs, n ~ U[1024,
    65536], average degree 8, one root each, all in one launch; every graph's levels
    equal the oracle's."""
    from blest_amd import workloads
    say("case workload er_batch 4096 graphs")
    t0 = time.time()
    graphs, roots = workloads.er_batch(4096, 3)
    gb = blest.GraphBatch(graphs)
    levels, secs = gb.bfs(roots)
    assert secs > 0 and len(levels) == 4096
    say(f"  one launch of {sum(g.n for g in graphs)} vertices ({time.time() - t0:.0f}s)")
    from concurrent.futures import ThreadPoolExecutor

    def check(i):
        cp, rows = graphs[i].arrays()
        return np.array_equal(levels[i], oracle.bfs(graphs[i].n, cp, rows, int(roots[i])))

    with ThreadPoolExecutor(8) as ex:
        ok = list(ex.map(check, range(len(graphs))))
    bad = [i for i, x in enumerate(ok) if not x]
    assert not bad, f"graphs {bad[:10]} differ from the oracle"
    say(f"  4096 graphs equal the oracle ({time.time() - t0:.0f}s)")
