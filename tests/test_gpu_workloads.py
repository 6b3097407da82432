"""GPU parity on the benchmark's own workloads, at their full sizes.

Each workload is prepared exactly as the benchmark prepares it (generator, seed,
reordering, window, BVSS padding, Benchmark.cuh:149-229) and the GPU results are compared
bit-exactly with the oracle on the original graph: BFS levels in original IDs (-1
unreachable) for the benchmark's own roots, canonical CC labels, and multi-source rows.
R-MAT 22 under Gorder+Jaccard needs its cached ordering (computing it takes far longer
than a test may); without the cache those cases are skipped."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_prepared = {}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def workload(name):
    from blest_amd import workloads
    if name not in _prepared:
        w = workloads.WORKLOADS[name]
        if w.reorder.name == "GORDER_JACCARD":
            stem = f"{w.name}_{int(w.reorder)}_{w.window}_"
            if not any(workloads.CACHE.glob(stem + "*_ordering.bin")):
                pytest.skip(f"{name}: no cached ordering under {workloads.CACHE}")
        _prepared[name] = workloads.prepare(name, keep_original=True, log=say)
    return _prepared[name]


def device_graph(blest, p, **opts):
    return blest.DeviceGraph(p.bvss, p.csc, p.inv_perm, **opts)


# workload -> roots checked against the oracle (the first of the benchmark's roots)
BFS_CASES = [("rmat16_natural", 16), ("rmat22_gorder", 6), ("grid4096_rcm", 4)]


@pytest.mark.parametrize("name,count", BFS_CASES)
def test_bfs_benchmark_roots(blest, oracle, name, count):
    say(f"case workload bfs {name}")
    p = workload(name)
    n = p.csc.n
    cp, rows = p.orig_col_ptrs, p.orig_rows
    deg = np.diff(cp.astype(np.int64))
    k = blest.BVSSBFSKernel(device_graph(blest, p))
    for src in [int(r) for r in p.roots[:count]]:
        say(f"  root {src}")
        res = k.run(src)
        want = oracle.bfs(n, cp, rows, src)
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
    root, and 6 rows (spread over both 64-bit halves of the batch) equal the oracle."""
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
    for j in (0, 31, 63, 64, 100, 127):
        np.testing.assert_array_equal(res.level_arrays[j], oracle.bfs(n, p.orig_col_ptrs, p.orig_rows, int(roots[j])),
                                      err_msg=f"row {j} root {roots[j]}: multi-source vs oracle")


@pytest.mark.skipif(os.environ.get("BLEST_LARGE_TESTS") != "1",
                    reason="R-MAT 26 (67 M vertices, 2.1 B edges) takes minutes; set BLEST_LARGE_TESTS=1")
def test_rmat26_bfs_mbfs_cc(blest, oracle):
    """R-MAT 26, natural order, as benchmarked: BFS from 2 of the benchmark's roots and CC
    against the oracle, and a 128-root multi-source launch (holding those 2 roots) whose
    every row equals the single-source GPU BFS of its root."""
    say("case workload rmat26 (large)")
    p = workload("rmat26")
    n = p.csc.n
    cp, rows = p.orig_col_ptrs, p.orig_rows
    single = blest.BVSSBFSKernel(device_graph(blest, p))
    roots = [int(r) for r in p.roots[:64]]
    for src in roots[:2]:
        say(f"  bfs root {src}")
        np.testing.assert_array_equal(single.run(src).levels, oracle.bfs(n, cp, rows, src), err_msg=f"src={src}")
    say("  cc")
    res = blest.BVSSCCKernel(p.bvss, p.csc, p.inv_perm).run()
    labels, count = oracle.cc(n, cp, rows)
    np.testing.assert_array_equal(res.components, labels)
    assert res.no_components == count
    del labels, res
    # 128 roots, as benchmarked: the benchmark's 64 and 64 more (two words per vertex)
    extra = [int(r) for r in blest.CSC.from_arrays(n, cp, rows).pick_roots(192, seed=5) if int(r) not in set(roots)]
    batch = roots + extra[:128 - len(roots)]
    say(f"  mbfs {len(batch)} roots")
    m = blest.BVSSMBFSKernel(p.bvss, p.csc, p.inv_perm, batch=128).run(np.asarray(batch, dtype=np.uint32))
    assert len(m.level_arrays) == 128
    for j, r in enumerate(batch):
        np.testing.assert_array_equal(m.level_arrays[j], single.run(r).levels, err_msg=f"row {j} root {r}")
