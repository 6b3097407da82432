"""Oracle: CPU restatement of the reference path.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
package, as the checker (or the timed baseline).  The product (blest_amd, libblest.so)
never imports, links or calls it.

Contents and what pins each function (see tests/test_oracle.py):
  bfs            textbook FIFO BFS            pinned by scipy.sparse.csgraph BFS + hand graphs
  bfs_many       `bfs` from many sources on host threads (same routine)
  cc             union-find, min-ID labels    pinned by scipy connected_components + hand graphs
  bvss_build     BVSS.cuh:185-467             pinned by decode == CSC, popcount == nnz, hand fixtures
  bvss_bfs       BFS/BVSSBFSKernel.cuh level replay + byte model
                                              pinned by equality with `bfs` on every graph tested
  bvss_cc        CC/BVSSCCKernel.cuh label propagation
                                              pinned by equality with `cc`
  jaccard_window CSC.cuh:717-920 restated     pinned by hand-computed orders (tests/test_oracle.py)
  gen.*          numpy restatement of the generators (pins blest's C++ generators bit-exactly)
  social.*       Common.cuh:100-250 restated  pinned by closed-form degree sequences
  gorder_ref     the reference's own Gorder compiled into oracle/_ref (when built)
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_GORDER = HERE / "_ref" / "libgorder_ref.so"

_lib = None
_gref = None

U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
I32P = C.POINTER(C.c_int32)


class Level(C.Structure):
    _fields_ = [("bottom_up", C.c_int32), ("frontier_vsets", C.c_uint32), ("new_vertices", C.c_uint32),
                ("next_vsets", C.c_uint32), ("slices", C.c_uint64), ("unvisited", C.c_uint64),
                ("scanned", C.c_uint64), ("bytes", C.c_uint64)]


def build() -> None:
    import subprocess
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        lib.oracle_bfs.restype = C.c_uint32
        lib.oracle_bfs.argtypes = [C.c_uint32, U32P, U32P, C.c_uint32, I32P]
        lib.oracle_cc.restype = C.c_uint32
        lib.oracle_cc.argtypes = [C.c_uint32, U32P, U32P, U32P]
        lib.oracle_bvss_build.restype = C.c_int
        lib.oracle_bvss_build.argtypes = [C.c_uint32, U32P, U32P, C.c_uint32, C.c_uint32, C.c_int, U32P, U64P,
                                          U64P, U32P, U32P, U32P, U32P, C.POINTER(C.c_double)]
        lib.oracle_bvss_bfs.restype = C.c_uint32
        lib.oracle_bvss_bfs.argtypes = [C.c_uint32, C.c_uint32, U64P, U32P, U32P, U32P, U32P, U32P, U32P,
                                        C.c_uint32, C.c_int, I32P, C.c_uint32, I32P, C.POINTER(Level),
                                        C.c_uint32, C.c_int]
        lib.oracle_bvss_cc.restype = C.c_uint32
        lib.oracle_bvss_cc.argtypes = [C.c_uint32, C.c_uint32, U64P, U32P, U32P, U32P, U32P, U32P]
        lib.oracle_max_threads.restype = C.c_int
        lib.oracle_jaccard_window.restype = None
        lib.oracle_jaccard_window.argtypes = [C.c_uint32, U32P, U32P, C.c_uint32, C.c_uint32, U32P]
        _lib = lib
    return _lib


def _p32(a):
    return a.ctypes.data_as(U32P)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def bfs(n: int, col_ptrs, rows, src: int) -> np.ndarray:
    cp, rw = _u32(col_ptrs), _u32(rows)
    lv = np.empty(n, dtype=np.int32)
    load().oracle_bfs(n, _p32(cp), _p32(rw) if rw.size else _p32(np.zeros(1, np.uint32)), src,
                      lv.ctypes.data_as(I32P))
    return lv


def bfs_many(n: int, col_ptrs, rows, srcs, workers: int = 8):
    """`bfs` from every source, `workers` at a time on host threads (the C routine runs
    without the GIL).  Yields (src, levels) in the order of srcs."""
    from concurrent.futures import ThreadPoolExecutor
    cp, rw = _u32(col_ptrs), _u32(rows)
    srcs = [int(s) for s in srcs]
    workers = max(1, workers)
    with ThreadPoolExecutor(workers) as ex:
        for i in range(0, len(srcs), workers):  # at most `workers` level arrays held at once
            part = srcs[i:i + workers]
            futs = [ex.submit(bfs, n, cp, rw, s) for s in part]
            for s, f in zip(part, futs):
                yield s, f.result()


def cc(n: int, col_ptrs, rows) -> tuple[np.ndarray, int]:
    cp, rw = _u32(col_ptrs), _u32(rows)
    lab = np.empty(n, dtype=np.uint32)
    k = load().oracle_cc(n, _p32(cp), _p32(rw) if rw.size else _p32(np.zeros(1, np.uint32)), _p32(lab))
    return lab, int(k)


def bvss_build(n: int, col_ptrs, rows, slice_size: int = 8, no_masks: int = 4, full_padding: bool = False) -> dict:
    lib = load()
    cp, rw = _u32(col_ptrs), _u32(rows)
    rwp = _p32(rw) if rw.size else _p32(np.zeros(1, np.uint32))
    V, S = C.c_uint32(), C.c_uint64()
    if lib.oracle_bvss_build(n, _p32(cp), rwp, slice_size, no_masks, int(full_padding), C.byref(V), C.byref(S),
                             None, None, None, None, None, None) != 0:
        raise ValueError("invalid slice size / noMasks")
    R = (n + slice_size - 1) // slice_size
    ssp = np.zeros(V.value + 1, dtype=np.uint64)
    v2r = np.zeros(V.value, dtype=np.uint32)
    rp = np.zeros(R + 1, dtype=np.uint32)
    rid = np.zeros(S.value, dtype=np.uint32)
    msk = np.zeros(S.value // no_masks, dtype=np.uint32)
    div = C.c_double()
    e = np.zeros(1, np.uint32)
    lib.oracle_bvss_build(n, _p32(cp), rwp, slice_size, no_masks, int(full_padding), C.byref(V), C.byref(S),
                          ssp.ctypes.data_as(U64P), _p32(v2r) if v2r.size else _p32(e), _p32(rp),
                          _p32(rid) if rid.size else _p32(e), _p32(msk) if msk.size else _p32(e), C.byref(div))
    return {"slice_set_ptrs": ssp, "virtual_to_real": v2r, "real_ptrs": rp, "row_ids": rid, "masks": msk,
            "update_divergence": div.value}


def bvss_bfs(n: int, bv: dict, csr_ptr, csr_idx, src: int, allow_bottom_up: bool = True, schedule=None,
             threads: int = 1, max_levels: int = 1 << 16):
    lib = load()
    lv = np.empty(n, dtype=np.int32)
    per = (Level * max_levels)()
    cp = None if csr_ptr is None else _u32(csr_ptr)
    ci = None if csr_idx is None else _u32(csr_idx)
    sch = None if schedule is None else np.ascontiguousarray(schedule, dtype=np.int32)
    e = np.zeros(1, np.uint32)
    nlev = lib.oracle_bvss_bfs(
        n, len(bv["virtual_to_real"]), bv["slice_set_ptrs"].ctypes.data_as(U64P),
        _p32(bv["virtual_to_real"]) if len(bv["virtual_to_real"]) else _p32(e), _p32(bv["real_ptrs"]),
        _p32(bv["row_ids"]) if len(bv["row_ids"]) else _p32(e), _p32(bv["masks"]) if len(bv["masks"]) else _p32(e),
        _p32(cp) if cp is not None else None, (_p32(ci) if ci.size else _p32(e)) if ci is not None else None,
        src, int(allow_bottom_up), sch.ctypes.data_as(I32P) if sch is not None else None,
        0 if sch is None else len(sch), lv.ctypes.data_as(I32P), per, max_levels, threads)
    levels = [{k: getattr(per[i], k) for k, _ in Level._fields_} for i in range(min(nlev, max_levels))]
    return lv, int(nlev), levels


def bvss_cc(n: int, bv: dict) -> tuple[np.ndarray, int]:
    lib = load()
    lab = np.empty(n, dtype=np.uint32)
    e = np.zeros(1, np.uint32)
    it = lib.oracle_bvss_cc(n, len(bv["virtual_to_real"]), bv["slice_set_ptrs"].ctypes.data_as(U64P),
                            _p32(bv["virtual_to_real"]) if len(bv["virtual_to_real"]) else _p32(e),
                            _p32(bv["real_ptrs"]), _p32(bv["row_ids"]) if len(bv["row_ids"]) else _p32(e),
                            _p32(bv["masks"]) if len(bv["masks"]) else _p32(e), _p32(lab))
    return lab, int(it)


def jaccard_window(n: int, col_ptrs, rows, window: int, slice_size: int = 8) -> np.ndarray:
    """CSC::jackardWithWindow restated: inv_perm[old] = new."""
    cp, rw = _u32(col_ptrs), _u32(rows)
    inv = np.empty(n, dtype=np.uint32)
    load().oracle_jaccard_window(n, _p32(cp), _p32(rw) if rw.size else _p32(np.zeros(1, np.uint32)), slice_size,
                                 window, _p32(inv))
    return inv


def max_threads() -> int:
    return int(load().oracle_max_threads())


def gorder_ref_available() -> bool:
    return REF_GORDER.exists()


def gorder_ref(n: int, col_ptrs, rows, window: int, transform: bool = False) -> np.ndarray:
    """The reference's own GorderGreedy (oracle/_ref).  Returns order[v] = new position."""
    global _gref
    if _gref is None:
        _gref = C.CDLL(str(REF_GORDER))
        _gref.gorder_ref.restype = C.c_int
        _gref.gorder_ref.argtypes = [C.c_uint32, C.c_uint32, U32P, U32P, C.c_longlong, C.c_int,
                                     C.POINTER(C.c_int64)]
    cp, rw = _u32(col_ptrs), _u32(rows)
    out = np.empty(n, dtype=np.int64)
    _gref.gorder_ref(n, int(rw.size), _p32(cp), _p32(rw), window, int(transform), out.ctypes.data_as(C.POINTER(C.c_int64)))
    return out


def canonical_labels(labels_new: np.ndarray, inv_perm: np.ndarray | None) -> np.ndarray:
    """Labels in reordered IDs -> labels in original IDs, each component named by its
    smallest ORIGINAL vertex ID.  inv_perm[old] = new."""
    n = len(labels_new)
    if inv_perm is None:
        inv_perm = np.arange(n, dtype=np.uint32)
    inv_perm = np.asarray(inv_perm, dtype=np.int64)
    comp_of_old = np.asarray(labels_new, dtype=np.int64)[inv_perm]   # component id of each old vertex
    old = np.arange(n, dtype=np.int64)
    best = np.full(n, np.iinfo(np.int64).max, dtype=np.int64)
    np.minimum.at(best, comp_of_old, old)
    return best[comp_of_old].astype(np.uint32)
