"""Named synthetic workloads and the preprocessing pipeline that prepares them.

The pipeline is the reference's Benchmark::run (Benchmark.cuh:149-229): generate the
CSC, classify it (FULL_PADDING = !social), reorder (with an ordering cache in the
reference's `_ordering.bin` format, CSC.cuh:335-356), build the BVSS.

Workloads (sizes as named; every seed fixed):
  rmat16_natural   R-MAT scale 16, edge factor 16, natural order, 16 roots
  rmat22_gorder    R-MAT scale 22, edge factor 16, Gorder+Jaccard window 65,536, 64 roots
  grid4096_rcm     4096x4096 grid, keep 0.9, 1% shortcuts, RCM, 64 roots
  rmat26           R-MAT scale 26, edge factor 16, natural order, 64 roots
  rmat26_mbfs      R-MAT scale 26, edge factor 16, natural order, 1,024 roots (MBFS)
  er_batch         4,096 Erdos-Renyi graphs, n ~ U[1024, 65536], degree 8, a root each
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import BVSS, CSC, Reorder

ROOT = Path(__file__).resolve().parent.parent
CACHE = Path(os.environ.get("BLEST_CACHE", ROOT / "bench_cache"))
# Version of the ordering algorithms (reorder.cpp) a cached ordering was computed with:
# raised whenever a change could alter an order, so that no stale ordering is reused.
ORDERING_VERSION = 2


@dataclass
class Workload:
    name: str
    kind: str               # rmat | grid | er
    scale: int = 0
    edge_factor: int = 16
    side: int = 0
    keep: float = 0.9
    shortcut_frac: float = 0.01
    n: int = 0
    avg_degree: float = 8.0
    seed: int = 1
    reorder: Reorder = Reorder.NATURAL
    window: int = 65536
    roots: int = 16
    root_seed: int = 7


WORKLOADS = {
    "rmat16_natural": Workload("rmat16_natural", "rmat", scale=16, reorder=Reorder.NATURAL, roots=16),
    "rmat22_gorder": Workload("rmat22_gorder", "rmat", scale=22, reorder=Reorder.GORDER_JACCARD, window=65536,
                              roots=64),
    "rmat22_jaccard": Workload("rmat22_jaccard", "rmat", scale=22, reorder=Reorder.JACCARD, window=65536,
                               roots=64),
    "grid4096_rcm": Workload("grid4096_rcm", "grid", side=4096, reorder=Reorder.RCM, roots=64),
    "grid4096_plain_rcm": Workload("grid4096_plain_rcm", "grid", side=4096, keep=1.0, shortcut_frac=0.0,
                                   reorder=Reorder.RCM, roots=64),
    "rmat26": Workload("rmat26", "rmat", scale=26, reorder=Reorder.NATURAL, roots=64),
    "rmat26_mbfs": Workload("rmat26_mbfs", "rmat", scale=26, reorder=Reorder.NATURAL, roots=1024),
}


def generate(w: Workload) -> CSC:
    if w.kind == "rmat":
        return CSC.rmat(w.scale, w.edge_factor, 0.57, 0.19, 0.19, seed=w.seed, permute=True)
    if w.kind == "grid":
        return CSC.grid(w.side, w.keep, w.shortcut_frac, seed=w.seed)
    if w.kind == "er":
        return CSC.erdos_renyi(w.n, w.avg_degree, seed=w.seed)
    raise ValueError(w.kind)


def csc_digest(g: CSC) -> str:
    cp, rows = g.arrays()
    h = hashlib.blake2b(digest_size=16)
    h.update(np.uint64(g.n).tobytes())
    h.update(cp.tobytes())
    h.update(rows.tobytes())
    return h.hexdigest()


def _cache_paths(w: Workload, digest: str) -> tuple[Path, Path]:
    stem = f"{w.name}_{int(w.reorder)}_{w.window}_v{ORDERING_VERSION}_{digest[:16]}"
    return CACHE / f"{stem}_ordering.bin", CACHE / f"{stem}_ordering.json"


def reorder_cached(g: CSC, w: Workload, log=print) -> np.ndarray:
    """Reorder g in place; reuse a cached inverse permutation for this exact graph."""
    if w.reorder == Reorder.NATURAL:
        return np.arange(g.n, dtype=np.uint32)
    digest = csc_digest(g)
    binp, meta = _cache_paths(w, digest)
    if binp.exists() and meta.exists():
        info = json.loads(meta.read_text())
        if info.get("digest") == digest and info.get("n") == g.n and info.get("version") == ORDERING_VERSION:
            inv = np.fromfile(binp, dtype=np.uint32)
            if inv.size == g.n:
                g.apply_permutation(inv)
                log(f"[prep] ordering loaded from cache {binp.name}")
                return inv
    t = time.time()
    log(f"[prep] reordering {w.name} with {w.reorder.name} window={w.window} (no cache)")
    inv = g.reorder(w.reorder, window=w.window)
    CACHE.mkdir(parents=True, exist_ok=True)
    inv.astype(np.uint32).tofile(binp)
    meta.write_text(json.dumps({"digest": digest, "version": ORDERING_VERSION, "n": g.n, "nnz": g.nnz, "workload": w.name,
                                "reorder": w.reorder.name, "window": w.window, "seconds": time.time() - t}))
    log(f"[prep] reordered in {time.time() - t:.1f}s, cached as {binp.name}")
    return inv


def er_batch(count: int = 4096, seed: int = 3, lo: int = 1024, hi: int = 65536, avg_degree: float = 8.0):
    """The many-small-graphs workload: `count` Erdos-Renyi graphs with n drawn uniformly
    from [lo, hi] (seeded), average degree avg_degree, one seeded root each (a vertex of
    degree >= 1).  Returns (graphs, roots as uint32)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, size=count)
    graphs = [CSC.erdos_renyi(int(n), avg_degree, seed=10_000 + i) for i, n in enumerate(sizes)]
    roots = np.array([int(g.pick_roots(1, seed=i)[0]) for i, g in enumerate(graphs)], dtype=np.uint32)
    return graphs, roots


@dataclass
class Prepared:
    workload: Workload
    csc: CSC                  # reordered
    inv_perm: np.ndarray      # old -> new
    bvss: BVSS
    orig_col_ptrs: np.ndarray | None
    orig_rows: np.ndarray | None
    roots: np.ndarray         # original IDs
    social: bool
    seconds: dict


def prepare(name: str, keep_original: bool = False, log=print, device_bvss: bool = False) -> Prepared:
    """device_bvss: build the BVSS on the GPU (blest_bvss_build_device, the same arrays
    as the host construction) instead of on the host."""
    w = WORKLOADS[name]
    secs = {}
    t = time.time()
    g = generate(w)
    secs["generate"] = time.time() - t
    log(f"[prep] {name}: n={g.n} nnz={g.nnz} ({secs['generate']:.1f}s)")
    social = g.is_social_network()
    roots = g.pick_roots(w.roots, seed=w.root_seed)
    cp0 = rows0 = None
    if keep_original:
        cp0, rows0 = g.arrays()
    t = time.time()
    inv = reorder_cached(g, w, log)
    secs["reorder"] = time.time() - t
    t = time.time()
    if device_bvss:
        b = BVSS(8, 4).construct_on_device(g, full_padding=not social)
    else:
        b = BVSS(8, 4).construct_from_csc(g, full_padding=not social)  # Benchmark.cuh:187-194
    secs["bvss"] = time.time() - t
    log(f"[prep] bvss ({'device' if device_bvss else 'host'}): {b.info()['no_slices']} slices, "
        f"{b.info()['no_virtual_slice_sets']} vsets ({secs['bvss']:.1f}s)")
    return Prepared(w, g, inv, b, cp0, rows0, roots, social, secs)
