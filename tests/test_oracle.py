"""Pins the oracle (CPU restatement of the reference path) against independent truths:
scipy graph routines, brute-force emulation, hand-computed fixtures and closed forms.
No GPU."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import breadth_first_order, connected_components, shortest_path

PAD = 0xFFFFFFFF


def csc_from_edges(n, edges, undirected=True):
    src = [a for a, b in edges]
    dst = [b for a, b in edges]
    if undirected:
        src, dst = src + dst, dst + src
    m = sp.coo_matrix((np.ones(len(src)), (np.array(dst, dtype=np.int64), np.array(src, dtype=np.int64))),
                      shape=(n, n)).tocsc()
    m.sum_duplicates()
    m.sort_indices()
    return m.indptr.astype(np.uint32), m.indices.astype(np.uint32)


def random_graph(n, m, seed, undirected=True):
    rng = np.random.default_rng(seed)
    e = rng.integers(0, n, size=(m, 2))
    e = e[e[:, 0] != e[:, 1]]
    return csc_from_edges(n, [tuple(x) for x in e], undirected)


def scipy_levels(n, cp, rows, src):
    # A[i, j] = 1 for edge j -> i (column j); scipy follows row->col, so use the transpose.
    a = sp.csc_matrix((np.ones(len(rows)), rows, cp), shape=(n, n)).T.tocsr()
    d = shortest_path(a, unweighted=True, directed=True, indices=src)
    return np.where(np.isinf(d), -1, d).astype(np.int32)


# --------------------------------------------------------------------- textbook BFS
@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("undirected", [True, False])
def test_bfs_matches_scipy(oracle, seed, undirected):
    n = 300 + 37 * seed
    cp, rows = random_graph(n, 2 * n, seed, undirected)
    for src in (0, n // 2, n - 1):
        np.testing.assert_array_equal(oracle.bfs(n, cp, rows, src), scipy_levels(n, cp, rows, src))


def test_bfs_hand_graphs(oracle):
    cp, rows = csc_from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4)])  # path
    np.testing.assert_array_equal(oracle.bfs(5, cp, rows, 0), [0, 1, 2, 3, 4])
    np.testing.assert_array_equal(oracle.bfs(5, cp, rows, 2), [2, 1, 0, 1, 2])
    cp, rows = csc_from_edges(6, [(0, 1), (0, 2), (0, 3)])  # star + isolated 4, 5
    np.testing.assert_array_equal(oracle.bfs(6, cp, rows, 1), [1, 0, 2, 2, -1, -1])
    np.testing.assert_array_equal(oracle.bfs(6, cp, rows, 5), [-1, -1, -1, -1, -1, 0])
    cp, rows = csc_from_edges(3, [(0, 1), (1, 2)], undirected=False)  # directed chain
    np.testing.assert_array_equal(oracle.bfs(3, cp, rows, 0), [0, 1, 2])
    np.testing.assert_array_equal(oracle.bfs(3, cp, rows, 2), [-1, -1, 0])


# ------------------------------------------------------------------------ union-find
@pytest.mark.parametrize("seed", range(5))
def test_cc_matches_scipy(oracle, seed):
    n = 500
    cp, rows = random_graph(n, n // 2 + 50 * seed, seed)
    lab, k = oracle.cc(n, cp, rows)
    a = sp.csc_matrix((np.ones(len(rows)), rows, cp), shape=(n, n))
    k2, comp = connected_components(a, directed=False)
    assert k == k2
    # canonical: smallest vertex of each scipy component
    smallest = np.full(k2, n)
    np.minimum.at(smallest, comp, np.arange(n))
    np.testing.assert_array_equal(lab, smallest[comp])


# --------------------------------------------------------------- BVSS construction
def fixture_graph():
    # columns (sources) -> rows (targets); directed so each slice is easy to read off
    cols = {0: [1, 5], 1: [1], 3: [9, 12], 7: [5, 15], 8: [0], 9: [0, 3], 15: [2]}
    edges = [(j, i) for j, rs in cols.items() for i in rs]
    return csc_from_edges(16, edges, undirected=False)


def test_bvss_fixture_variable_padding(oracle):
    cp, rows = fixture_graph()
    b = oracle.bvss_build(16, cp, rows, 8, 4, full_padding=False)
    np.testing.assert_array_equal(b["slice_set_ptrs"], [0, 8, 12])
    np.testing.assert_array_equal(b["virtual_to_real"], [0, 1])
    np.testing.assert_array_equal(b["real_ptrs"], [0, 1, 2])
    np.testing.assert_array_equal(b["row_ids"], [1, 5, 9, 12, 15, PAD, PAD, PAD, 0, 2, 3, PAD])
    np.testing.assert_array_equal(b["masks"], [0x08088103, 0x00000080, 0x00028003])
    assert b["update_divergence"] == pytest.approx(0.875)


def test_bvss_fixture_full_padding(oracle):
    cp, rows = fixture_graph()
    b = oracle.bvss_build(16, cp, rows, 8, 4, full_padding=True)
    np.testing.assert_array_equal(b["slice_set_ptrs"], [0, 128, 256])
    rid = b["row_ids"].reshape(2, 32, 4)
    msk = b["masks"].reshape(2, 32)
    np.testing.assert_array_equal(rid[0, :5, 0], [1, 5, 9, 12, 15])
    np.testing.assert_array_equal(msk[0, :5], [0x03, 0x81, 0x08, 0x08, 0x80])
    np.testing.assert_array_equal(rid[1, :3, 0], [0, 2, 3])
    np.testing.assert_array_equal(msk[1, :3], [0x03, 0x80, 0x02])
    assert (rid[0, 5:, :] == PAD).all() and (rid[0, :, 1:] == PAD).all()
    assert (msk[0, 5:] == 0).all() and (msk[1, 3:] == 0).all()


@pytest.mark.parametrize("full_padding", [False, True])
def test_bvss_complete_vset_lane_map(oracle, full_padding):
    # one column with 130 rows -> one complete vset of 128 + a leftover of 2
    n = 256
    edges = [(0, 100 + k) for k in range(130)]
    cp, rows = csc_from_edges(n, edges, undirected=False)
    b = oracle.bvss_build(n, cp, rows, 8, 4, full_padding)
    rid = b["row_ids"]
    for t in range(32):
        for p in range(4):
            assert rid[4 * t + p] == 100 + p * 32 + t
    assert (b["masks"][:32] == 0x01010101).all()
    if full_padding:
        assert b["slice_set_ptrs"][2] == 256
        assert rid[128] == 228 and rid[132] == 229 and rid[129] == PAD
    else:
        assert b["slice_set_ptrs"][2] == 132
        np.testing.assert_array_equal(rid[128:132], [228, 229, PAD, PAD])
        assert b["masks"][32] == 0x00000101


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("full_padding", [False, True])
def test_bvss_decodes_to_csc(oracle, seed, full_padding):
    n = 700 + seed * 61
    cp, rows = random_graph(n, 5 * n, seed)
    b = oracle.bvss_build(n, cp, rows, 8, 4, full_padding)
    ssp, v2r, rid, msk = b["slice_set_ptrs"], b["virtual_to_real"], b["row_ids"], b["masks"]
    got = set()
    pop = 0
    for v in range(len(v2r)):
        r = int(v2r[v])
        for k in range(int(ssp[v]), int(ssp[v + 1])):
            m = (int(msk[k // 4]) >> (8 * (k % 4))) & 0xFF
            pop += bin(m).count("1")
            if m == 0:
                assert rid[k] == PAD or not full_padding or True
            for bit in range(8):
                if m >> bit & 1:
                    got.add((8 * r + bit, int(rid[k])))
    want = {(j, int(i)) for j in range(n) for i in rows[cp[j]:cp[j + 1]]}
    assert got == want
    assert pop == len(rows)  # BVSS.cuh:507-538 "Check"
    # rows ascend inside each rset when read in source order, prefix sums are monotone
    assert (np.diff(ssp.astype(np.int64)) > 0).all()
    assert (np.diff(b["real_ptrs"].astype(np.int64)) >= 0).all()
    assert (np.diff(ssp.astype(np.int64)) % 4 == 0).all()


def test_bvss_rejects_bad_slice_size(oracle):
    cp, rows = fixture_graph()
    with pytest.raises(ValueError):
        oracle.bvss_build(16, cp, rows, 8, 2)
    with pytest.raises(ValueError):
        oracle.bvss_build(16, cp, rows, 3, 10)


# ------------------------------------------------ m8n8k128 .b1 AND+POPC (A11)
def mma_b1_and_popc(frag_a, frag_b):
    """Brute-force m8n8k128 with the PTX ISA fragment layout of .b1 (row.col):
    lane l holds A[l>>2][32*(l&3) : +32], B[32*(l&3) : +32][l>>2], and receives
    C[l>>2][2*(l&3)], C[l>>2][2*(l&3)+1]."""
    A = np.zeros((8, 128), dtype=np.uint8)
    B = np.zeros((128, 8), dtype=np.uint8)
    for l in range(32):
        for k in range(32):
            A[l >> 2, 32 * (l & 3) + k] = (frag_a[l] >> k) & 1
            B[32 * (l & 3) + k, l >> 2] = (frag_b[l] >> k) & 1
    Cm = (A.astype(np.int64) @ B.astype(np.int64))  # AND then popcount-sum over k
    return [(Cm[l >> 2, 2 * (l & 3)], Cm[l >> 2, 2 * (l & 3) + 1]) for l in range(32)]


@pytest.mark.parametrize("seed", range(20))
def test_m8n8k128_reduces_to_per_lane_and(seed):
    rng = np.random.default_rng(seed)
    masks = rng.integers(0, 2**32, size=32, dtype=np.uint64)
    f = int(rng.integers(0, 256))
    # reference packing (BFS/BVSSBFSKernel.cuh:151-170)
    frag_b = [f if l % 9 == 0 else (f << 8) if l % 9 == 4 else 0 for l in range(32)]
    lo = mma_b1_and_popc([int(m) & 0xFFFF for m in masks], frag_b)
    hi = mma_b1_and_popc([(int(m) >> 16) & 0xFFFF for m in masks], frag_b)
    for l in range(32):
        m = int(masks[l])
        want = [bin(((m >> (8 * s)) & 0xFF) & f).count("1") for s in range(4)]
        assert [lo[l][0], lo[l][1], hi[l][0], hi[l][1]] == want


# ------------------------------------------------------ level replay and byte model
def test_bvss_bfs_byte_model_path(oracle):
    cp, rows = csc_from_edges(3, [(0, 1), (1, 2)])
    b = oracle.bvss_build(3, cp, rows, 8, 4, False)
    lv, nlev, per = oracle.bvss_bfs(3, b, cp, rows, 0, allow_bottom_up=False)
    np.testing.assert_array_equal(lv, [0, 1, 2])
    assert nlev == 3
    assert [p["bytes"] for p in per] == [65, 65, 57]  # hand-computed B_td
    lv, nlev, per = oracle.bvss_bfs(3, b, cp, rows, 0, schedule=[0, 1])
    np.testing.assert_array_equal(lv, [0, 1, 2])
    assert per[1]["bottom_up"] == 1 and per[1]["unvisited"] == 1 and per[1]["scanned"] == 1
    assert per[1]["bytes"] == 4 + 1 + 12 + 12 + 4  # hand-computed B_bu


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("full_padding", [False, True])
def test_bvss_bfs_equals_textbook(oracle, seed, full_padding):
    n = 1000 + 113 * seed
    cp, rows = random_graph(n, 4 * n, seed)
    b = oracle.bvss_build(n, cp, rows, 8, 4, full_padding)
    rng = np.random.default_rng(seed)
    for src in rng.integers(0, n, size=3):
        want = oracle.bfs(n, cp, rows, int(src))
        for kw in ({}, {"allow_bottom_up": False}, {"schedule": [0, 1, 1, 1, 0, 1]}, {"threads": 4}):
            lv, nlev, _ = oracle.bvss_bfs(n, b, cp, rows, int(src), **kw)
            np.testing.assert_array_equal(lv, want)
            assert nlev == want.max() + 1


def test_bvss_bfs_directed(oracle):
    n = 400
    cp, rows = random_graph(n, 3 * n, 9, undirected=False)
    a = sp.csc_matrix((np.ones(len(rows)), rows, cp), shape=(n, n)).tocsr()  # CSR = in-neighbours
    b = oracle.bvss_build(n, cp, rows, 8, 4, False)
    for src in (0, 17, 399):
        lv, _, _ = oracle.bvss_bfs(n, b, a.indptr.astype(np.uint32), a.indices.astype(np.uint32), src,
                                   schedule=[0, 1, 1, 0, 1])
        np.testing.assert_array_equal(lv, oracle.bfs(n, cp, rows, src))


@pytest.mark.parametrize("seed", range(4))
def test_bvss_cc_equals_union_find(oracle, seed):
    n = 900
    cp, rows = random_graph(n, n // 2 + 100 * seed, seed)
    b = oracle.bvss_build(n, cp, rows, 8, 4, seed % 2 == 1)
    lab, it = oracle.bvss_cc(n, b)
    want, _ = oracle.cc(n, cp, rows)
    np.testing.assert_array_equal(lab, want)
    assert it >= 1


def test_canonical_labels(oracle):
    # components {0,2} and {1,3} in reordered IDs; inv_perm[old] = new
    labels_new = np.array([0, 1, 0, 1], dtype=np.uint32)
    inv = np.array([2, 3, 0, 1], dtype=np.uint32)  # old 0 -> new 2 ...
    # old vertices 0,2 sit at new 2,0 (component 0) -> min old = 0; old 1,3 -> min old = 1
    np.testing.assert_array_equal(oracle.canonical_labels(labels_new, inv), [0, 1, 0, 1])


# ----------------------------------------------------------------- social classifier
def test_social_closed_forms():
    from oracle.social import social_degree_distribution as sdd
    assert not sdd([4] * 1000)                       # regular: no heavy tail, no fit points
    assert sdd([999] + [1] * 999)                    # star: top 1% holds half the degree mass
    ks = np.arange(5, 200)
    counts = np.floor(1e6 * ks ** -2.5).astype(int)  # power law, gamma = 2.5
    deg = np.repeat(ks, counts)
    assert sdd(np.concatenate([deg, np.ones(50 * len(deg), dtype=int)]))  # only the fit fires
    assert not sdd([])
    assert not sdd([0, 0, 0])


def test_bench_byte_model_matches_oracle_replay(oracle, blest):
    """bench.reference_level_bytes (numpy, from a traversal's own levels and direction
    mask) equals the oracle's level replay byte counts, level by level, for top-down,
    bottom-up and mixed schedules on social, grid and sparse graphs."""
    import bench
    rng = np.random.default_rng(0)
    for make, reorder in [(lambda: blest.CSC.rmat(11, 16, seed=1), blest.Reorder.JACCARD),
                          (lambda: blest.CSC.grid(48, 0.9, 0.01, 2), blest.Reorder.RCM),
                          (lambda: blest.CSC.erdos_renyi(3001, 3.0, 4), None)]:
        g = make()
        if reorder is not None:
            g.reorder(reorder, window=1024)
        bv = blest.BVSS(8, 4).construct_from_csc(g, full_padding=not g.is_social_network()).arrays()
        cp, rows = g.arrays()
        for src in g.pick_roots(2, seed=2):
            for mask in (0, 0b110, 0b11110, int(rng.integers(0, 1 << 20))):
                sched = [(mask >> i) & 1 for i in range(64)]
                lv, _, per = oracle.bvss_bfs(g.n, bv, cp, rows, int(src), schedule=sched, threads=2)
                assert bench.reference_level_bytes(g.n, cp, rows, bv, lv, mask) == [p["bytes"] for p in per]


# ------------------------------------------------------------- generator distributions
# The generator restatement is bit-exact with libblest's generators (test_host.py); these
# checks pin the restatement itself against the distributions the workloads specify.
def _binomial_ok(count, m, p, sigmas=6.0):
    return abs(count - m * p) <= sigmas * np.sqrt(m * p * (1 - p)) + 1


@pytest.mark.parametrize("a,b,c", [(0.57, 0.19, 0.19), (0.25, 0.25, 0.25), (0.45, 0.15, 0.3)])
def test_rmat_quadrant_frequencies(a, b, c):
    """Every recursion level draws quadrant a/b/c/d with the configured probabilities, and
    consecutive levels (the two halves of one hash) are independent."""
    from oracle import gen
    scale, ef = 12, 16
    _, _, quads = gen.rmat_samples(scale, ef, a, b, c, seed=11)
    m = ef << scale
    probs = [a, b, c, 1 - a - b - c]
    for lvl, q in enumerate(quads):
        counts = np.bincount(q, minlength=4)
        for k in range(4):
            assert _binomial_ok(counts[k], m, probs[k]), (lvl, k, counts[k], m * probs[k])
    for lvl in range(scale - 1):
        joint = np.bincount(quads[lvl].astype(np.int64) * 4 + quads[lvl + 1], minlength=16)
        for i in range(4):
            for j in range(4):
                assert _binomial_ok(joint[4 * i + j], m, probs[i] * probs[j]), (lvl, i, j)


def test_rmat_degree_skew_matches_quadrant_mass():
    """Closed form: the expected share of raw samples whose source lies in the lower half
    of the ID range is a + b at every level; with a = 0.57, b = 0.19 that is 0.76."""
    from oracle import gen
    su, sv, _ = gen.rmat_samples(14, 16, 0.57, 0.19, 0.19, seed=3)
    m = su.size
    low_u = int((su < (1 << 13)).sum())
    low_v = int((sv < (1 << 13)).sum())
    assert _binomial_ok(low_u, m, 0.76) and _binomial_ok(low_v, m, 0.76)
    # vertex 0 is the most likely endpoint: P(u = 0) = (a + b)^scale
    assert (su == 0).sum() >= 0.5 * m * 0.76 ** 14


def test_grid_edge_keep_rate_and_shortcuts():
    from oracle import gen
    side, keep = 300, 0.9
    n, cp, rows = gen.grid(side, keep, 0.0, seed=5)
    deg = np.diff(cp.astype(np.int64))
    lattice = 2 * side * (side - 1)  # undirected grid edges
    kept = int(deg.sum()) // 2
    assert _binomial_ok(kept, lattice, keep)
    n2, cp2, _ = gen.grid(side, keep, 0.01, seed=5)
    extra = int(np.diff(cp2.astype(np.int64)).sum()) // 2 - kept
    cnt = round(0.01 * n)
    # shortcuts: cnt draws minus self-loops and coincidences with lattice or other shortcuts
    assert 0.95 * cnt <= extra <= cnt


def test_erdos_renyi_uniform_endpoints():
    from oracle import gen
    n, d = 4000, 8.0
    _, cp, rows = gen.erdos_renyi(n, d, seed=9)
    deg = np.diff(cp.astype(np.int64))
    m = round(n * d / 2)
    # undirected edges after dropping self-loops and duplicates: close to m
    assert 0.99 * m <= deg.sum() // 2 <= m
    # degrees follow Binomial(2m, 1/n) closely: mean d, variance about d
    assert abs(deg.mean() - d) < 0.1
    assert 0.8 * d < deg.var() < 1.2 * d


def _csc_of(adj, n):
    cp, rows = [0], []
    for j in range(n):
        rows += sorted(adj.get(j, []))
        cp.append(len(rows))
    return np.array(cp, np.uint32), np.array(rows, np.uint32)


def test_jaccard_window_hand_orders(oracle):
    """CSC::jackardWithWindow restated, pinned by orders worked out by hand from the
    reference's rules (CSC.cuh:717-920)."""
    # two triangles' worth of ties: an 8-cycle 3-4-5-6-7 closed by 3-7 and a triangle
    # 0-1-2, all degree 2, slots of 4 in one window of 8.  Slot 0: seed = the LAST
    # unplaced column of largest degree (7); candidates 4 and 5 tie at 1/3 and the first
    # in queue order wins (4), the last queue entry moving into its place; then 5 (1/4
    # vs 6), then 6 (1/5 vs 3).  Slot 1: seed 3 (last of 0..3), its rows reach only
    # placed columns, so the seed rule picks again (2), then 1 (first of the tie with 0),
    # then 0.
    adj = {0: [1, 2], 1: [0, 2], 2: [0, 1], 3: [4, 7], 4: [3, 5], 5: [4, 6], 6: [5, 7], 7: [3, 6]}
    cp, rows = _csc_of(adj, 8)
    np.testing.assert_array_equal(oracle.jaccard_window(8, cp, rows, 8, 4), [7, 6, 5, 4, 1, 2, 3, 0])
    # two windows of 4: window 0 seeds with column 3 (degree 4), then 0 (ratio 3/4 over
    # 1/2, 1/2), then 2 before 1 (tie; 2 moved to the front of the queue when 0 left);
    # window 1 seeds with 6 (last of the degree-3 columns), then 7 (2/3 over 1/2), 4, 5.
    adj = {0: [4, 5, 6], 1: [4, 5], 2: [6, 7], 3: [4, 5, 6, 7], 4: [0, 1, 3], 5: [0, 1, 3], 6: [0, 2, 3],
           7: [2, 3]}
    cp, rows = _csc_of(adj, 8)
    np.testing.assert_array_equal(oracle.jaccard_window(8, cp, rows, 4, 4), [1, 3, 2, 0, 6, 7, 4, 5])


def test_jaccard_window_is_a_windowed_bijection(oracle):
    rng = np.random.default_rng(5)
    n = 1000
    e = rng.integers(0, n, size=(6000, 2))
    g = {j: [] for j in range(n)}
    for a, b in e:
        if a != b:
            g[int(a)].append(int(b))
            g[int(b)].append(int(a))
    g = {j: sorted(set(v)) for j, v in g.items()}
    cp, rows = _csc_of(g, n)
    inv = oracle.jaccard_window(n, cp, rows, 128, 8)
    assert sorted(inv.tolist()) == list(range(n))
    old = np.arange(n)
    assert np.array_equal(inv // 128, old // 128)  # every column stays inside its window
