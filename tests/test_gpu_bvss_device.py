"""GPU-side BVSS construction (blest_bvss_build_device): every array, the update
divergence and the statistics equal the host construction's (which the oracle pins,
tests/test_host.py) bit for bit, for both padding rules, on social, grid, directed,
sparse and empty graphs, and on the R-MAT 22 benchmark graph as benchmarked."""
import sys
import time

import numpy as np
import pytest
from conftest import progress

pytestmark = pytest.mark.gpu


def say(*a):
    progress(*a)


def assert_same(host, dev, name):
    hi, di = host.info(), dev.info()
    for k, v in hi.items():
        assert di[k] == v, f"{name}: info {k}: host {v}, device {di[k]}"
    ha, da = host.arrays(), dev.arrays()
    for k in ha:
        np.testing.assert_array_equal(da[k], ha[k], err_msg=f"{name}: {k}")


def sparse_graph(blest, n=5000):
    rng = np.random.default_rng(9)
    e = rng.integers(0, n, size=(800, 2))  # most vertices isolated
    return blest.CSC.from_edges(n, e[:, 0], e[:, 1], undirected=True)


def directed_graph(blest, n=3000):
    rng = np.random.default_rng(4)
    e = rng.integers(0, n, size=(25000, 2))
    return blest.CSC.from_edges(n, e[:, 0], e[:, 1], undirected=False)


CASES = {
    "rmat12": lambda b: b.CSC.rmat(12, 16, seed=2),
    "rmat14_jaccard": lambda b: b.CSC.rmat(14, 16, seed=3),
    "grid120": lambda b: b.CSC.grid(120, 0.9, 0.01, 5),
    "er_odd_n": lambda b: b.CSC.erdos_renyi(20001, 7.0, seed=6),
    "sparse": sparse_graph,
    "directed": directed_graph,
    "edgeless": lambda b: b.CSC.from_arrays(77, np.zeros(78, np.uint32), np.zeros(0, np.uint32)),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("full_padding", [None, True, False])
def test_device_build_equals_host(blest, name, full_padding):
    say(f"case bvss device {name} full_padding={full_padding}")
    g = CASES[name](blest)
    if name == "rmat14_jaccard":
        g.reorder(blest.Reorder.JACCARD, window=2048)
    host = blest.BVSS(8, 4).construct_from_csc(g, full_padding=full_padding)
    dev = blest.BVSS(8, 4).construct_on_device(g, full_padding=full_padding)
    assert_same(host, dev, name)


def test_device_build_rejects_other_shapes(blest):
    g = blest.CSC.rmat(8, 8, seed=1)
    with pytest.raises(blest.BlestError):
        blest.BVSS(16, 2).construct_on_device(g)
    with pytest.raises(blest.BlestError):
        blest.BVSS(8, 3).construct_on_device(g)


def test_device_build_benchmark_graph(blest):
    """R-MAT 22 under Gorder+Jaccard (its cached ordering), as the benchmark builds it."""
    from blest_amd import workloads
    w = workloads.WORKLOADS["rmat22_gorder"]
    stem = f"{w.name}_{int(w.reorder)}_{w.window}_v{workloads.ORDERING_VERSION}_"
    if not any(workloads.CACHE.glob(stem + "*_ordering.bin")):
        pytest.skip("no cached ordering for R-MAT 22")
    say("case bvss device rmat22_gorder")
    p = workloads.prepare("rmat22_gorder", log=say)
    dev = blest.BVSS(8, 4).construct_on_device(p.csc, full_padding=not p.social)  # warm-up (code load)
    t0 = time.perf_counter()
    dev = blest.BVSS(8, 4).construct_on_device(p.csc, full_padding=not p.social)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    host = blest.BVSS(8, 4).construct_from_csc(p.csc, full_padding=not p.social)
    t_host = time.perf_counter() - t0
    say(f"  R-MAT 22 BVSS: device {t_dev:.3f} s, host {t_host:.3f} s")
    assert_same(host, dev, "rmat22_gorder")
