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
    assert best["alpha"] in (4, 8, 15, 30, 60, 120, 250, 500, 1000)
    assert best["beta"] in (6, 12, 18, 24, 48, 96, 200, 400, 800)
    assert best["td_quads"] in (1, 2, 4) and best["bu_group"] in (4, 8)
    assert best["fuse_quads"] in (512, 1024, 2048, 4096)
    dg = blest.DeviceGraph(b, g, inv, **{k: best[k] for k in ("blocks_per_cu", "td_quads", "bu_group", "alpha",
                                                                "beta", "fuse_quads")})
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


def test_bfs_trace_matches_plain_run(blest, oracle):
    """The traced run gives the same levels and one timing entry per level."""
    import ctypes as C
    print("case bfs trace", flush=True)
    g = blest.CSC.rmat(13, 16, seed=12)
    cp, rows = g.arrays()
    b = blest.BVSS(8, 4).construct_from_csc(g)
    dg = blest.DeviceGraph(b, g, None)
    q = blest._Query(dg, 1)
    lib = blest._lib.load()
    for src in g.pick_roots(3, seed=4):
        lv = np.empty(g.n, dtype=np.int32)
        st = blest._lib.BfsStats()
        tr = (blest._lib.LevelTrace * 64)()
        cnt = C.c_uint32()
        blest._lib.check(lib.blest_bfs_trace(q._h, int(src), lv.ctypes.data_as(blest._lib.I32P), C.byref(st), tr,
                                             64, C.byref(cnt)))
        want = oracle.bfs(g.n, cp, rows, int(src))
        np.testing.assert_array_equal(lv, want)
        assert cnt.value == st.total_levels
        for i in range(cnt.value):
            assert tr[i].level == i + 1 and tr[i].expand_ms >= 0 and tr[i].update_ms >= 0
            assert tr[i].bottom_up == ((st.bottom_up_mask >> i) & 1)


@pytest.mark.parametrize("td_quads", [1, 2, 4])
@pytest.mark.parametrize("bu_group", [4, 8])
@pytest.mark.parametrize("direction", [0, 2])
def test_tile_knobs(blest, oracle, td_quads, bu_group, direction):
    """Every top-down slice tile and bottom-up staging depth gives the oracle's levels,
    at creation and when changed in place."""
    print(f"case tiles td_quads={td_quads} bu_group={bu_group} direction={direction}", flush=True)
    g = blest.CSC.rmat(13, 16, seed=14)
    cp, rows = g.arrays()
    inv = g.reorder(blest.Reorder.JACCARD, window=2048)
    b = blest.BVSS(8, 4).construct_from_csc(g)
    dg = blest.DeviceGraph(b, g, inv, td_quads=td_quads, bu_group=bu_group, direction=direction)
    k = blest.BVSSBFSKernel(dg)
    roots = [int(r) for r in np.argsort(inv)[g.pick_roots(3, seed=3)]]
    for r in roots:
        np.testing.assert_array_equal(k.run(r).levels, oracle.bfs(g.n, cp, rows, r))
    lib = blest._lib.load()
    blest._lib.check(lib.blest_device_graph_set_tiles(dg._h, 4 if td_quads != 4 else 1, 12 - bu_group))
    for r in roots:
        np.testing.assert_array_equal(k.run(r).levels, oracle.bfs(g.n, cp, rows, r))


def test_tile_knobs_rejected(blest):
    g = blest.CSC.rmat(10, 8, seed=8)
    b = blest.BVSS(8, 4).construct_from_csc(g)
    with pytest.raises(blest.BlestError):
        blest.DeviceGraph(b, g, None, td_quads=3)
    with pytest.raises(blest.BlestError):
        blest.DeviceGraph(b, g, None, bu_group=16)
    dg = blest.DeviceGraph(b, g, None)
    lib = blest._lib.load()
    assert lib.blest_device_graph_set_tiles(dg._h, 8, 8) == 1
    assert lib.blest_device_graph_set_tiles(dg._h, 2, 5) == 1


FUSE_GRAPHS = [
    ("rmat13", lambda B: B.CSC.rmat(13, 8, seed=21), "JACCARD"),
    ("grid_sparse", lambda B: B.CSC.grid(72, 0.62, 0.0, 9), None),  # many components, thin paths
    ("grid_deep", lambda B: B.CSC.grid(120, 1.0, 0.0, 4), "RCM"),    # ~240 levels, all small
]


@pytest.mark.parametrize("name,make,reorder", FUSE_GRAPHS, ids=[f[0] for f in FUSE_GRAPHS])
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_fused_small_levels(blest, oracle, name, make, reorder, direction):
    """Small levels run fused in one workgroup (at the default bound, at the largest one,
    at a tiny one that hands back after every first level, and never): every setting
    gives the oracle's levels and the same level count, reach and traversed edges."""
    print(f"case fused levels {name} direction={direction}", flush=True)
    g = make(blest)
    cp, rows = g.arrays()
    inv = g.reorder(getattr(blest.Reorder, reorder), window=4096) if reorder else None
    b = blest.BVSS(8, 4).construct_from_csc(g)
    dg = blest.DeviceGraph(b, g, inv, direction=direction)
    k = blest.BVSSBFSKernel(dg)
    roots = g.pick_roots(3, seed=5)
    if inv is not None:
        roots = np.argsort(inv)[roots]
    for r in [int(x) for x in roots]:
        want = oracle.bfs(g.n, cp, rows, r)
        got = {}
        for fuse in (-1, 0, 32, 65536):
            dg.set_fuse(fuse)
            for rep in range(2):  # the second run starts from the first one's level hint
                res = k.run(r)
                np.testing.assert_array_equal(res.levels, want, err_msg=f"fuse={fuse} rep={rep}")
            got[fuse] = (res.total_levels, res.no_visited, res.edges_in_component)
        assert len(set(got.values())) == 1, got
        assert got[-1][0] == want.max() + 1


def test_set_fuse_rejected(blest):
    g = blest.CSC.rmat(10, 8, seed=8)
    b = blest.BVSS(8, 4).construct_from_csc(g)
    with pytest.raises(blest.BlestError):
        blest.DeviceGraph(b, g, None, fuse_quads=65537)
    dg = blest.DeviceGraph(b, g, None)
    with pytest.raises(blest.BlestError):
        dg.set_fuse(1 << 20)


@pytest.mark.parametrize("bu_probe", [0, -1])
@pytest.mark.parametrize("direction", [0, 2])
def test_bottom_up_probe(blest, oracle, bu_probe, direction):
    """Bottom-up sweeps give the oracle's levels whether each vertex first tests its
    highest-degree in-neighbour or its first list entry, on a symmetric and a directed
    graph (the directed one's sweeps read the transpose)."""
    print(f"case bottom-up probe {bu_probe} direction={direction}", flush=True)
    rng = np.random.default_rng(5)
    n = 6000
    src, dst = rng.integers(0, n, 60000), rng.integers(0, n, 60000)
    for g in (blest.CSC.rmat(13, 16, seed=31), blest.CSC.from_edges(n, src, dst, undirected=False)):
        cp, rows = g.arrays()
        b = blest.BVSS(8, 4).construct_from_csc(g)
        dg = blest.DeviceGraph(b, g, None, direction=direction, bu_probe=bu_probe)
        k = blest.BVSSBFSKernel(dg)
        for r in g.pick_roots(3, seed=2):
            np.testing.assert_array_equal(k.run(int(r)).levels, oracle.bfs(g.n, cp, rows, int(r)))
