"""Host side of libblest through the C ABI: symbol exports, generators, CSC utilities,
reorderings, BVSS construction, caches and error paths.  No GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import gen, social

ROOT = Path(__file__).resolve().parent.parent
PAD = 0xFFFFFFFF


def declared_functions():
    text = (ROOT / "include" / "blest.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(blest_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(blest):
    from blest_amd import _lib
    lib = _lib.load()
    names = declared_functions()
    assert len(names) > 40
    for name in names:
        assert hasattr(lib, name), name
    assert set(_lib.PROTOTYPES) <= set(names)
    assert lib.blest_version().decode().startswith("blest")
    assert lib.blest_status_str(0).decode() == "ok"


# ------------------------------------------------------------------- generators
@pytest.mark.parametrize("scale,seed,permute", [(6, 1, True), (8, 7, True), (9, 3, False), (10, 11, True)])
def test_rmat_matches_numpy_restatement(blest, scale, seed, permute):
    g = blest.CSC.rmat(scale, 16, 0.57, 0.19, 0.19, seed=seed, permute=permute)
    cp, rows = g.arrays()
    n, cp2, rows2 = gen.rmat(scale, 16, 0.57, 0.19, 0.19, seed, permute)
    assert g.n == n
    np.testing.assert_array_equal(cp, cp2)
    np.testing.assert_array_equal(rows, rows2)
    assert g.check_symmetry()


@pytest.mark.parametrize("side,keep,frac,seed", [(16, 0.9, 0.01, 1), (33, 0.9, 0.01, 5), (20, 1.0, 0.0, 2)])
def test_grid_matches_numpy_restatement(blest, side, keep, frac, seed):
    g = blest.CSC.grid(side, keep, frac, seed)
    cp, rows = g.arrays()
    n, cp2, rows2 = gen.grid(side, keep, frac, seed)
    np.testing.assert_array_equal(cp, cp2)
    np.testing.assert_array_equal(rows, rows2)


def test_full_grid_structure(blest):
    g = blest.CSC.grid(10, 1.0, 0.0, 1)
    cp, rows = g.arrays()
    deg = np.diff(cp.astype(np.int64))
    assert g.nnz == 2 * 2 * 10 * 9          # 180 undirected edges
    assert deg[0] == 2 and deg[5] == 3 and deg[55] == 4


@pytest.mark.parametrize("n,d,seed", [(1000, 8.0, 1), (4097, 8.0, 2), (50, 3.0, 3)])
def test_er_matches_numpy_restatement(blest, n, d, seed):
    g = blest.CSC.erdos_renyi(n, d, seed)
    cp, rows = g.arrays()
    _, cp2, rows2 = gen.erdos_renyi(n, d, seed)
    np.testing.assert_array_equal(cp, cp2)
    np.testing.assert_array_equal(rows, rows2)


def test_pick_roots_matches_restatement(blest):
    g = blest.CSC.rmat(10, 16, seed=4)
    cp, _ = g.arrays()
    r = g.pick_roots(16, seed=9)
    np.testing.assert_array_equal(r, gen.pick_roots(g.n, cp, 16, 9))
    assert len(set(r.tolist())) == 16 and (np.diff(cp.astype(np.int64))[r] > 0).all()


def test_generator_seed_and_thread_independence(blest):
    a = blest.CSC.rmat(11, 16, seed=5).arrays()
    b = blest.CSC.rmat(11, 16, seed=5).arrays()
    c = blest.CSC.rmat(11, 16, seed=6).arrays()
    np.testing.assert_array_equal(a[1], b[1])
    assert len(a[1]) != len(c[1]) or not np.array_equal(a[1], c[1])


# ------------------------------------------------------------------ CSC utilities
def test_csc_create_normalises(blest):
    # duplicates, self loop and unsorted column are cleaned
    g = blest.CSC.from_arrays(4, [0, 4, 5, 5, 6], [3, 1, 3, 0, 2, 3])
    cp, rows = g.arrays()
    np.testing.assert_array_equal(cp, [0, 2, 3, 3, 3])
    np.testing.assert_array_equal(rows, [1, 3, 2])


def test_csc_create_rejects_bad_input(blest):
    with pytest.raises(blest.BlestError):
        blest.CSC.from_arrays(3, [0, 1, 2, 3], [0, 1, 5])
    with pytest.raises(blest.BlestError):
        blest.CSC.from_arrays(3, [0, 2, 1, 3], [0, 1, 2])


def test_social_classifier_matches_restatement(blest):
    for g in (blest.CSC.rmat(12, 16, seed=1), blest.CSC.grid(64, 0.9, 0.01, 1), blest.CSC.erdos_renyi(5000, 8, 1),
              blest.CSC.rmat(10, 4, seed=2)):
        cp, rows = g.arrays()
        assert g.is_social_network() == social.is_social(g.n, cp, rows)
    assert blest.CSC.rmat(12, 16, seed=1).is_social_network()
    assert not blest.CSC.grid(64, 0.9, 0.01, 1).is_social_network()


def test_metrics(blest):
    g = blest.CSC.grid(10, 1.0, 0.0, 1)
    m = g.metrics()
    assert m["max_degree"] == 4 and m["max_bandwidth"] == 10 and m["avg_degree"] == pytest.approx(3.6)


# -------------------------------------------------------------------- MTX + caches
def test_mtx_reader(blest, tmp_path):
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n% comment\n4 4 4\n1 2\n2 3\n3 3\n4 1\n")
    g = blest.CSC.from_mtx(str(p), undirected=False, binary=True)
    cp, rows = g.arrays()
    # line "j i" is the edge j -> i (column j-1 holds row i-1); self loop 3 3 dropped
    np.testing.assert_array_equal(cp, [0, 1, 2, 2, 3])
    np.testing.assert_array_equal(rows, [1, 2, 0])
    q = tmp_path / "h.mtx"
    q.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 2\n2 1 0.5\n3 1 1.5\n")
    h = blest.CSC.from_mtx(str(q), undirected=True, binary=False)
    cp, rows = h.arrays()
    np.testing.assert_array_equal(cp, [0, 2, 3, 4])
    np.testing.assert_array_equal(rows, [1, 2, 0, 0])
    with pytest.raises(blest.BlestError) as e:
        blest.CSC.from_mtx(str(tmp_path / "missing.mtx"), True)
    assert e.value.status == 4


def test_binary_caches_roundtrip(blest, tmp_path):
    g = blest.CSC.rmat(9, 8, seed=3)
    g.save_binary(str(tmp_path / "g_csc.bin"))
    h = blest.CSC.load_binary(str(tmp_path / "g_csc.bin"))
    for a, b in zip(g.arrays(), h.arrays()):
        np.testing.assert_array_equal(a, b)
    raw = (tmp_path / "g_csc.bin").read_bytes()
    n, nnz = np.frombuffer(raw[:8], dtype=np.uint32)
    assert n == g.n and nnz == g.nnz and raw[8] == int(g.is_social_network())
    assert len(raw) == 9 + 4 * (n + 1) + 4 * nnz  # CSC.cuh:225-266 layout
    b = blest.BVSS(8, 4).construct_from_csc(g)
    b.save_binary(str(tmp_path / "g_bvss.bin"))
    b2 = blest.BVSS.load_binary(str(tmp_path / "g_bvss.bin"))
    for k, v in b.arrays().items():
        np.testing.assert_array_equal(v, b2.arrays()[k])
    assert b.info()["update_divergence"] == b2.info()["update_divergence"]
    inv = g.reorder(blest.Reorder.RANDOM, seed=4)
    lib = blest._lib.load()
    blest.check(lib.blest_ordering_save(str(tmp_path / "o.bin").encode(), inv.ctypes.data_as(blest._lib.U32P), len(inv)))
    back = np.empty_like(inv)
    blest.check(lib.blest_ordering_load(str(tmp_path / "o.bin").encode(), back.ctypes.data_as(blest._lib.U32P), len(inv)))
    np.testing.assert_array_equal(inv, back)


# --------------------------------------------------------------------- reordering
def relabel(cp, rows, inv):
    n = len(cp) - 1
    cols = np.repeat(np.arange(n), np.diff(cp.astype(np.int64)))
    return {(int(inv[c]), int(inv[r])) for c, r in zip(cols, rows)}


@pytest.mark.parametrize("kind", ["NATURAL", "RCM", "JACCARD", "GORDER", "GORDER_JACCARD", "RANDOM", "DEGREE"])
def test_reorder_is_a_relabelling(blest, kind):
    g = blest.CSC.rmat(10, 8, seed=2)
    cp, rows = g.arrays()
    inv = g.reorder(getattr(blest.Reorder, kind), window=256)
    assert sorted(inv.tolist()) == list(range(g.n))
    cp2, rows2 = g.arrays()
    assert relabel(cp, rows, inv) == relabel(cp2, rows2, np.arange(g.n))
    for j in range(g.n):  # columns stay sorted
        col = rows2[cp2[j]:cp2[j + 1]]
        assert (np.diff(col.astype(np.int64)) > 0).all()


def test_jaccard_window_confinement(blest):
    g = blest.CSC.rmat(11, 8, seed=3)
    inv = g.reorder(blest.Reorder.JACCARD, window=512)
    old = np.arange(g.n)
    assert (old // 512 == inv // 512).all()  # columns never leave their window


def test_jaccard_window_must_divide(blest):
    g = blest.CSC.rmat(8, 8, seed=3)
    with pytest.raises(blest.BlestError):
        g.reorder(blest.Reorder.JACCARD, window=100)


def test_jaccard_packs_twins(blest):
    # columns with identical neighbourhoods end up in the same slice set
    n = 64
    edges = []
    for a, b in [(3, 40), (17, 50), (9, 33)]:  # (twin group anchor, shared rows base)
        for r in range(b, b + 6):
            edges += [(a, r), (a + 20, r)]
    g = blest.CSC.from_edges(n, [e[0] for e in edges], [e[1] for e in edges], undirected=False)
    inv = g.reorder(blest.Reorder.JACCARD, window=64)
    for a in (3, 17, 9):
        assert inv[a] // 8 == inv[a + 20] // 8


def test_rcm_reduces_bandwidth_on_shuffled_grid(blest):
    g = blest.CSC.grid(40, 1.0, 0.0, 1)
    g.reorder(blest.Reorder.RANDOM, seed=3)
    before = g.metrics()["max_bandwidth"]
    g.reorder(blest.Reorder.RCM)
    after = g.metrics()["max_bandwidth"]
    assert after < before / 5 and after <= 100


def test_auto_dispatch(blest):
    soc = blest.CSC.rmat(10, 16, seed=1)
    assert soc.is_social_network()
    inv = soc.reorder(blest.Reorder.AUTO, jaccard=False)
    np.testing.assert_array_equal(inv, np.arange(soc.n))  # social & !jaccard -> natural
    grid = blest.CSC.grid(30, 0.9, 0.01, 1)
    a = blest.CSC.grid(30, 0.9, 0.01, 1)
    np.testing.assert_array_equal(grid.reorder(blest.Reorder.AUTO), a.reorder(blest.Reorder.RCM))


def locality_score(cp, rows, order, w):
    """Gorder's objective F = sum over pairs within w positions of
    |in(u) & in(v)| + [u->v] + [v->u] (symmetric graphs: in = out)."""
    n = len(cp) - 1
    nb = [set(rows[cp[v]:cp[v + 1]].tolist()) for v in range(n)]
    seq = np.argsort(order)
    f = 0
    for i in range(n):
        u = seq[i]
        for j in range(max(0, i - w), i):
            v = seq[j]
            f += len(nb[u] & nb[v]) + (v in nb[u]) + (u in nb[v])
    return f


def test_gorder_locality_close_to_reference(blest, oracle):
    if not oracle.gorder_ref_available():
        pytest.skip("reference Gorder not built (oracle/_ref)")
    g = blest.CSC.rmat(9, 8, seed=5)
    cp, rows = g.arrays()
    w = 5
    ref = oracle.gorder_ref(g.n, cp, rows, w, transform=False)
    assert sorted(ref.tolist()) == list(range(g.n))
    h = blest.CSC.rmat(9, 8, seed=5)
    mine = h.reorder(blest.Reorder.GORDER, window=w)
    nat = locality_score(cp, rows, np.arange(g.n), w)
    f_ref = locality_score(cp, rows, ref, w)
    f_mine = locality_score(cp, rows, mine, w)
    assert f_ref > nat and f_mine > nat
    assert f_mine >= 0.9 * f_ref


# --------------------------------------------------------------------------- BVSS
@pytest.mark.parametrize("maker", ["rmat", "grid", "er", "tiny"])
@pytest.mark.parametrize("full_padding", [False, True])
def test_bvss_matches_oracle(blest, oracle, maker, full_padding):
    g = {"rmat": lambda: blest.CSC.rmat(11, 16, seed=3), "grid": lambda: blest.CSC.grid(45, 0.9, 0.01, 2),
         "er": lambda: blest.CSC.erdos_renyi(3001, 8, 4),
         "tiny": lambda: blest.CSC.from_arrays(5, [0, 1, 2, 2, 3, 3], [1, 0, 4, 0])}[maker]()
    cp, rows = g.arrays()
    b = blest.BVSS(8, 4).construct_from_csc(g, full_padding=full_padding)
    mine = b.arrays()
    ref = oracle.bvss_build(g.n, cp, rows, 8, 4, full_padding)
    for k in mine:
        np.testing.assert_array_equal(mine[k], ref[k], err_msg=k)
    info = b.info()
    assert info["update_divergence"] == ref["update_divergence"]
    assert info["edges_check"] == g.nnz
    assert info["no_slices"] == info["no_unpadded_slices"] + info["no_padded_slices"]


@pytest.mark.parametrize("slice_size,no_masks", [(4, 8), (16, 2), (32, 1), (2, 16)])
def test_bvss_other_slice_sizes_match_oracle(blest, oracle, slice_size, no_masks):
    g = blest.CSC.rmat(9, 8, seed=1)
    cp, rows = g.arrays()
    b = blest.BVSS(slice_size, no_masks).construct_from_csc(g, full_padding=False)
    ref = oracle.bvss_build(g.n, cp, rows, slice_size, no_masks, False)
    for k, v in b.arrays().items():
        np.testing.assert_array_equal(v, ref[k], err_msg=k)


def test_bvss_invalid_shape(blest):
    g = blest.CSC.rmat(6, 4, seed=1)
    for s, m in [(8, 2), (3, 10), (64, 0)]:
        with pytest.raises(blest.BlestError) as e:
            blest.BVSS(s, m).construct_from_csc(g)
        assert e.value.status == 2


def test_bvss_full_padding_default_follows_social(blest):
    soc = blest.CSC.rmat(10, 16, seed=1)
    assert blest.BVSS().construct_from_csc(soc).info()["full_padding"] == 0
    grid = blest.CSC.grid(32, 0.9, 0.01, 1)
    assert blest.BVSS().construct_from_csc(grid).info()["full_padding"] == 1


def test_bvss_csv(blest):
    b = blest.BVSS().construct_from_csc(blest.CSC.rmat(8, 8, seed=1))
    head, row = b.csv()
    assert head.split("\t")[0] == "N" and len(head.split("\t")) == 16
    cells = row.rstrip("\t").split("\t")
    assert len(cells) == 16 and all(c.startswith("'") for c in cells)


def test_permutation_check(blest):
    g = blest.CSC.rmat(6, 4, seed=1)
    bad = np.zeros(g.n, dtype=np.uint32)
    with pytest.raises(blest.BlestError) as e:
        g.apply_permutation(bad)
    assert e.value.status == 3


# --------------------------------------------------- adopting a BVSS built elsewhere
@pytest.mark.parametrize("full_padding", [False, True])
def test_bvss_adopt_roundtrip(blest, full_padding):
    """The five arrays of a BVSS (as the reference's accessors return them) adopted back
    give the same arrays and statistics."""
    g = blest.CSC.rmat(10, 8, seed=4)
    b = blest.BVSS(8, 4).construct_from_csc(g, full_padding=full_padding)
    arr, info = b.arrays(), b.info()
    a = blest.BVSS.adopt(g.n, 8, 4, arr["slice_set_ptrs"], arr["virtual_to_real"], arr["real_ptrs"],
                         arr["row_ids"], arr["masks"], info["update_divergence"], full_padding)
    for k, v in arr.items():
        np.testing.assert_array_equal(v, a.arrays()[k], err_msg=k)
    ai = a.info()
    for k in ("no_slices", "no_unpadded_slices", "no_padded_slices", "edges_check", "no_real_slice_sets",
              "no_virtual_slice_sets", "update_divergence", "full_padding"):
        assert ai[k] == info[k], k
    assert ai["edges_check"] == g.nnz


def _adopt(blest, n, arr):
    return blest.BVSS.adopt(n, 8, 4, arr["slice_set_ptrs"], arr["virtual_to_real"], arr["real_ptrs"],
                            arr["row_ids"], arr["masks"])


def test_bvss_adopt_rejects_malformed_arrays(blest):
    """Each planted fault in otherwise valid arrays is refused with a message."""
    g = blest.CSC.rmat(9, 8, seed=2)
    good = blest.BVSS(8, 4).construct_from_csc(g).arrays()
    _adopt(blest, g.n, good)  # accepted

    def planted(key, fn):
        a = {k: v.copy() for k, v in good.items()}
        fn(a[key])
        with pytest.raises(blest.BlestError) as e:
            _adopt(blest, g.n, a)
        assert e.value.status == 1 and "BVSS arrays" in str(e.value)

    big = int(np.argmax(np.diff(good["slice_set_ptrs"].astype(np.int64))))
    planted("row_ids", lambda x: x.__setitem__(int(np.flatnonzero(x != 0xFFFFFFFF)[0]), g.n))  # row out of range
    planted("virtual_to_real", lambda x: x.__setitem__(len(x) - 1, x[-1] - 1))  # wrong owner
    planted("real_ptrs", lambda x: x.__setitem__(1, x[2] + 1))  # not monotone
    planted("slice_set_ptrs", lambda x: x.__setitem__(big + 1, x[big + 1] - 2))  # vset not on a mask word
    with pytest.raises(blest.BlestError):  # R != ceil(n / 8)
        _adopt(blest, g.n + 64, good)
    with pytest.raises(blest.BlestError):  # invalid slice size (BVSS.cuh:103-110)
        blest.BVSS.adopt(g.n, 8, 3, *good.values())


def _hub_graph(blest, n=5000):
    src, dst = [], []
    for h in range(4):
        for v in range(n):
            if v != h and (h == 0 or v % (h + 1) == 0):
                src.append(h)
                dst.append(v)
    return blest.CSC.from_edges(n, np.array(src), np.array(dst), undirected=True)


JACCARD_EXACT = {
    "rmat14_w4096": (lambda b: b.CSC.rmat(14, 16, seed=3), 4096),
    "rmat14_w1024": (lambda b: b.CSC.rmat(14, 16, seed=3), 1024),
    "rmat12_w65536": (lambda b: b.CSC.rmat(12, 16, seed=4), 65536),
    "rmat13_natural_w256": (lambda b: b.CSC.rmat(13, 16, seed=5, permute=False), 256),
    "grid100_w2048": (lambda b: b.CSC.grid(100, 0.9, 0.01, 3), 2048),
    "hubs_w1024": (_hub_graph, 1024),  # rows of thousands of columns per window
    "er20000_w8192": (lambda b: b.CSC.erdos_renyi(20000, 12.0, seed=2), 8192),
}


@pytest.mark.parametrize("case", list(JACCARD_EXACT))
def test_jaccard_order_equals_oracle(blest, oracle, case):
    """The product's Jaccard window produces exactly the order of the plain restatement
    of CSC::jackardWithWindow (CSC.cuh:717-920), tie rules included."""
    make, window = JACCARD_EXACT[case]
    g = make(blest)
    cp, rows = g.arrays()
    want = oracle.jaccard_window(g.n, cp, rows, window)
    np.testing.assert_array_equal(g.reorder(blest.Reorder.JACCARD, window=window), want)


def test_jaccard_order_equals_oracle_directed(blest, oracle):
    rng = np.random.default_rng(1)
    n = 3000
    e = rng.integers(0, n, size=(20000, 2))
    g = blest.CSC.from_edges(n, e[:, 0], e[:, 1], undirected=False)
    cp, rows = g.arrays()
    np.testing.assert_array_equal(g.reorder(blest.Reorder.JACCARD, window=512),
                                  oracle.jaccard_window(n, cp, rows, 512))


def test_gorder_jaccard_chain_equals_oracle_jaccard(blest, oracle):
    """GORDER_JACCARD = the Jaccard window applied to the Gorder-relabelled CSC
    (CSC.cuh:932-948): its Jaccard stage equals the restatement on that CSC."""
    g = blest.CSC.rmat(12, 16, seed=9)
    h = blest.CSC.from_arrays(g.n, *g.arrays())
    inv_g = h.reorder(blest.Reorder.GORDER, window=4096)
    cp, rows = h.arrays()
    inv_j = oracle.jaccard_window(g.n, cp, rows, 4096)
    chained = inv_j[inv_g]
    np.testing.assert_array_equal(g.reorder(blest.Reorder.GORDER_JACCARD, window=4096), chained)
