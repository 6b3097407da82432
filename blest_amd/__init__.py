"""blest_amd — MI355X-native BVSS bit-matrix traversal (BFS, multi-source BFS, CC).

Python mirror of the reference's host interface, over the C ABI of libblest.so:

    CSC            DataStructures/CSC.cuh:35-96       (graph, MTX reader, reorderings)
    BVSS           DataStructures/BVSS.cuh:28-95      (BitMatrix subclass)
    BVSSBFSKernel  BFS/BFSKernel.cuh:30-56, BFS/BVSSBFSKernel.cuh:1264-1283
    BVSSCCKernel   CC/CCKernel.cuh:28-54, CC/BVSSCCKernel.cuh
    BVSSMBFSKernel MBFS/MBFSKernel.cuh:27-53
    BFSResult / CCResult / MBFSResult  BFS/BFSKernel.cuh:21-28, CC/CCKernel.cuh:21-26,
                                       MBFS/MBFSKernel.cuh:21-25

All computation happens in libblest.so (host C++ and HIP kernels); this module only
moves numpy arrays across the boundary.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np

from . import _lib
from ._lib import BlestError, check

__all__ = ["CSC", "BVSS", "BitMatrix", "DeviceGraph", "BVSSBFSKernel", "BVSSCCKernel", "BVSSMBFSKernel",
           "BFSResult", "CCResult", "MBFSResult", "Reorder", "BlestError", "device_opts", "GraphBatch", "Server", "pinned_empty", "autotune_bfs"]


def _u32p(a: np.ndarray):
    return a.ctypes.data_as(_lib.U32P)


def _as_u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


class Reorder(IntEnum):
    AUTO = 0
    NATURAL = 1
    RCM = 2
    JACCARD = 3
    GORDER = 4
    GORDER_JACCARD = 5
    RANDOM = 6
    DEGREE = 7


class CSC:
    """Column-compressed adjacency: column j holds the rows i of the edges j -> i."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        self._lib = _lib.load()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.blest_csc_free(h)
            self._h = C.c_void_p(0)

    # ---- constructors
    @classmethod
    def from_arrays(cls, n: int, col_ptrs, rows) -> "CSC":
        lib = _lib.load()
        cp, rw = _as_u32(col_ptrs), _as_u32(rows)
        h = C.c_void_p()
        check(lib.blest_csc_create(n, _u32p(cp), _u32p(rw) if rw.size else None, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, n: int, src, dst, undirected: bool = True) -> "CSC":
        """Edges src[k] -> dst[k]; both directions when undirected."""
        src, dst = np.asarray(src, dtype=np.int64), np.asarray(dst, dtype=np.int64)
        if undirected:
            src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
        order = np.lexsort((dst, src))
        src, dst = src[order], dst[order]
        col_ptrs = np.zeros(n + 1, dtype=np.uint32)
        np.add.at(col_ptrs, src + 1, 1)
        return cls.from_arrays(n, np.cumsum(col_ptrs, dtype=np.uint64).astype(np.uint32), dst.astype(np.uint32))

    @classmethod
    def from_mtx(cls, path: str, undirected: bool, binary: bool = True) -> "CSC":
        lib = _lib.load()
        h = C.c_void_p()
        check(lib.blest_csc_from_mtx(str(path).encode(), int(undirected), int(binary), C.byref(h)))
        return cls(h.value)

    @classmethod
    def load_binary(cls, path: str) -> "CSC":
        lib = _lib.load()
        h = C.c_void_p()
        check(lib.blest_csc_load_binary(str(path).encode(), C.byref(h)))
        return cls(h.value)

    def save_binary(self, path: str) -> None:
        check(self._lib.blest_csc_save_binary(self._h, str(path).encode()))

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1, permute: bool = True) -> "CSC":
        lib = _lib.load()
        h = C.c_void_p()
        check(lib.blest_csc_generate_rmat(scale, edge_factor, a, b, c, seed, int(permute), C.byref(h)))
        return cls(h.value)

    @classmethod
    def grid(cls, side: int, keep: float = 0.9, shortcut_frac: float = 0.01, seed: int = 1) -> "CSC":
        lib = _lib.load()
        h = C.c_void_p()
        check(lib.blest_csc_generate_grid(side, keep, shortcut_frac, seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def erdos_renyi(cls, n: int, avg_degree: float = 8.0, seed: int = 1) -> "CSC":
        lib = _lib.load()
        h = C.c_void_p()
        check(lib.blest_csc_generate_er(n, avg_degree, seed, C.byref(h)))
        return cls(h.value)

    # ---- queries
    def _info(self):
        n, nnz, soc = C.c_uint32(), C.c_uint64(), C.c_int()
        check(self._lib.blest_csc_info(self._h, C.byref(n), C.byref(nnz), C.byref(soc)))
        return n.value, nnz.value, bool(soc.value)

    @property
    def n(self) -> int:
        return self._info()[0]

    @property
    def nnz(self) -> int:
        return self._info()[1]

    def is_social_network(self) -> bool:
        return self._info()[2]

    def check_symmetry(self) -> bool:
        s = C.c_int()
        check(self._lib.blest_csc_is_symmetric(self._h, C.byref(s)))
        return bool(s.value)

    def arrays(self) -> tuple[np.ndarray, np.ndarray]:
        n, nnz, _ = self._info()
        cp = np.empty(n + 1, dtype=np.uint32)
        rw = np.empty(nnz, dtype=np.uint32)
        check(self._lib.blest_csc_export(self._h, _u32p(cp), _u32p(rw) if nnz else None))
        return cp, rw

    def metrics(self) -> dict:
        m = _lib.CscMetrics()
        check(self._lib.blest_csc_metrics_get(self._h, C.byref(m)))
        return {k: getattr(m, k) for k, _ in m._fields_ if k != "_pad"}

    def pick_roots(self, count: int, seed: int = 1) -> np.ndarray:
        out = np.empty(count, dtype=np.uint32)
        check(self._lib.blest_csc_pick_roots(self._h, count, seed, _u32p(out)))
        return out

    # ---- reordering (CSC::reorder, CSC.cuh:358-386)
    def reorder(self, kind: Reorder = Reorder.AUTO, slice_size: int = 8, window: int = 65536,
                jaccard: bool = True, gorder: bool = False, threads: int = 0, seed: int = 1) -> np.ndarray:
        o = _lib.ReorderOpts()
        self._lib.blest_reorder_opts_default(C.byref(o))
        o.kind, o.slice_size, o.window = int(kind), slice_size, window
        o.jaccard, o.gorder, o.threads, o.seed = int(jaccard), int(gorder), threads, seed
        inv = np.empty(self.n, dtype=np.uint32)
        check(self._lib.blest_csc_reorder(self._h, C.byref(o), _u32p(inv)))
        return inv

    def apply_permutation(self, inv_perm) -> None:
        inv = _as_u32(inv_perm)
        check(self._lib.blest_csc_apply_permutation(self._h, _u32p(inv)))


class BitMatrix:
    """Abstract base of the bit-matrix formats (DataStructures/BitMatrix.cuh:21-30)."""


class BVSS(BitMatrix):
    """BVSS(sliceSize, noMasks) + constructFromCSCMatrix (BVSS.cuh:97-113, 185-418)."""

    def __init__(self, slice_size: int = 8, no_masks: int = 4):
        self.slice_size, self.no_masks = slice_size, no_masks
        self._h = C.c_void_p(0)
        self._lib = _lib.load()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.blest_bvss_free(h)
            self._h = C.c_void_p(0)

    def construct_from_csc(self, csc: CSC, full_padding: bool | None = None) -> "BVSS":
        h = C.c_void_p()
        fp = -1 if full_padding is None else int(full_padding)
        check(self._lib.blest_bvss_build(csc._h, self.slice_size, self.no_masks, fp, C.byref(h)))
        if self._h.value:
            self._lib.blest_bvss_free(self._h)
        self._h = h
        return self

    def construct_on_device(self, csc: CSC, full_padding: bool | None = None, device: int = 0) -> "BVSS":
        """The same arrays built on the GPU (blest_bvss_build_device)."""
        h = C.c_void_p()
        fp = -1 if full_padding is None else int(full_padding)
        check(self._lib.blest_bvss_build_device(csc._h, self.slice_size, self.no_masks, fp, device, C.byref(h)))
        if self._h.value:
            self._lib.blest_bvss_free(self._h)
        self._h = h
        return self

    @classmethod
    def load_binary(cls, path: str) -> "BVSS":
        b = cls()
        h = C.c_void_p()
        check(b._lib.blest_bvss_load_binary(str(path).encode(), C.byref(h)))
        b._h = h
        i = b.info()
        b.slice_size, b.no_masks = i["slice_size"], i["no_masks"]
        return b

    @classmethod
    def adopt(cls, n: int, slice_size: int, no_masks: int, slice_set_ptrs, virtual_to_real, real_ptrs, row_ids,
              masks, update_divergence: float = 0.0, full_padding: bool = False) -> "BVSS":
        """A BVSS from the five arrays of a BVSS built elsewhere (the reference's
        getSliceSetPtrs / getVirtualToReal / getRealPtrs / getRowIds / getMasks,
        BVSS.cuh:54-68); copied and validated by the library."""
        b = cls(slice_size, no_masks)
        ssp = np.ascontiguousarray(slice_set_ptrs, dtype=np.uint64)
        v2r, rp = _as_u32(virtual_to_real), _as_u32(real_ptrs)
        rows, msk = _as_u32(row_ids), _as_u32(masks)
        h = C.c_void_p()
        check(b._lib.blest_bvss_adopt(n, slice_size, no_masks, rp.size - 1, ssp.size - 1, rows.size,
                                      ssp.ctypes.data_as(_lib.U64P), _u32p(v2r), _u32p(rp), _u32p(rows), _u32p(msk),
                                      float(update_divergence), int(full_padding), C.byref(h)))
        b._h = h
        return b

    def save_binary(self, path: str) -> None:
        check(self._lib.blest_bvss_save_binary(self._h, str(path).encode()))

    def info(self) -> dict:
        i = _lib.BvssInfo()
        check(self._lib.blest_bvss_info_get(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    def arrays(self) -> dict:
        i = self.info()
        V, R, S = i["no_virtual_slice_sets"], i["no_real_slice_sets"], i["no_slices"]
        ssp = np.empty(V + 1, dtype=np.uint64)
        v2r = np.empty(V, dtype=np.uint32)
        rp = np.empty(R + 1, dtype=np.uint32)
        rows = np.empty(S, dtype=np.uint32)
        masks = np.empty(S // i["no_masks"], dtype=np.uint32)
        check(self._lib.blest_bvss_export(self._h, ssp.ctypes.data_as(_lib.U64P), _u32p(v2r), _u32p(rp),
                                          _u32p(rows), _u32p(masks)))
        return {"slice_set_ptrs": ssp, "virtual_to_real": v2r, "real_ptrs": rp, "row_ids": rows, "masks": masks}

    def csv(self) -> tuple[str, str]:
        buf = C.create_string_buffer(4096)
        check(self._lib.blest_bvss_csv_header(buf, 4096))
        head = buf.value.decode()
        check(self._lib.blest_bvss_csv_row(self._h, buf, 4096))
        return head, buf.value.decode()


def device_opts(**kw) -> _lib.DeviceOpts:
    o = _lib.DeviceOpts()
    _lib.load().blest_device_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class DeviceGraph:
    """The BVSS (and bottom-up CSR) resident in HBM; queries reuse it."""

    def __init__(self, bvss: BVSS, csc: CSC | None = None, inv_perm=None, csr: CSC | None = None, **opts):
        """csc: the CSC the BVSS was built from; or csr: its transpose as the reference's
        BVSS keeps it (BVSS::getCSR(), BVSS.cuh:54,187)."""
        self._lib = _lib.load()
        self._h = C.c_void_p(0)
        if csc is not None and csr is not None:
            raise ValueError("pass the CSC or its transpose (csr), not both")
        o = device_opts(**opts)
        g = csr if csr is not None else csc
        if g is None:
            o.keep_csr = 0
        inv = None if inv_perm is None else _as_u32(inv_perm)
        h = C.c_void_p()
        create = self._lib.blest_device_graph_create_csr if csr is not None else self._lib.blest_device_graph_create
        check(create(bvss._h, g._h if g is not None else None, _u32p(inv) if inv is not None else None, C.byref(o),
                     C.byref(h)))
        self._h = h
        self._keep = (bvss, g)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.blest_device_graph_free(h)
            self._h = C.c_void_p(0)

    def info(self) -> dict:
        i = _lib.DeviceGraphInfo()
        check(self._lib.blest_device_graph_info_get(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    @property
    def n(self) -> int:
        return self.info()["n"]

    def set_thresholds(self, alpha: float, beta: float, direction: int = 0) -> None:
        check(self._lib.blest_device_graph_set_thresholds(self._h, alpha, beta, direction))

    def set_fuse(self, fuse_quads: int) -> None:
        """Quads a small top-down level may stream inside one workgroup (levels fused with no
        kernel boundary between them); 0 = default (2048), < 0 = never."""
        check(self._lib.blest_device_graph_set_fuse(self._h, fuse_quads))


def autotune_bfs(bvss: BVSS, csc: CSC | None, inv_perm, roots, **base) -> tuple[dict, float, int]:
    """Native autotuner: returns (best device opts, their mean BFS ms, candidates timed)."""
    lib = _lib.load()
    o = device_opts(**base)
    best = _lib.DeviceOpts()
    ms = C.c_double()
    trials = C.c_uint32()
    inv = None if inv_perm is None else _as_u32(inv_perm)
    r = _as_u32(roots)
    check(lib.blest_autotune_bfs(bvss._h, csc._h if csc is not None else None,
                                 _u32p(inv) if inv is not None else None, _u32p(r), len(r), C.byref(o),
                                 C.byref(best), C.byref(ms), C.byref(trials)))
    return {k: getattr(best, k) for k, _ in best._fields_}, ms.value, trials.value


class _Query:
    def __init__(self, dg: DeviceGraph, max_sources: int = 1):
        self._lib = _lib.load()
        self.dg = dg
        h = C.c_void_p()
        check(self._lib.blest_query_create(dg._h, max_sources, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.blest_query_free(h)
            self._h = C.c_void_p(0)


@dataclass
class BFSResult:
    """BFS/BFSKernel.cuh:21-28 (levels as int32, -1 = unreachable)."""
    time: float
    source_vertex: int
    levels: np.ndarray
    total_levels: int
    no_visited: int
    td_levels: int = 0
    bu_levels: int = 0
    edges_in_component: int = 0
    bottom_up_mask: int = 0


@dataclass
class CCResult:
    """CC/CCKernel.cuh:21-26; components = smallest original vertex ID per component."""
    time: float
    components: np.ndarray
    no_components: int
    iterations: int = 0


@dataclass
class MBFSResult:
    """MBFS/MBFSKernel.cuh:21-25; level_arrays[r] = levels of root r."""
    time: float
    level_arrays: list = field(default_factory=list)


class BFSKernel:
    """BFS/BFSKernel.cuh:30-56: run(src) forwards to hostCode(src)."""

    def __init__(self, matrix: BitMatrix):
        self.matrix = matrix

    def host_code(self, source_vertex: int) -> BFSResult:
        raise NotImplementedError

    def run(self, source_vertex: int) -> BFSResult:
        return self.host_code(source_vertex)


class BVSSBFSKernel(BFSKernel):
    """Single-source BFS on the GPU (BFS/BVSSBFSKernel.cuh:1283-1608).  Unlike the
    reference the graph is uploaded once (at construction) rather than per source."""

    def __init__(self, matrix: BVSS | DeviceGraph, csc: CSC | None = None, inv_perm=None, **opts):
        super().__init__(matrix)
        self.dg = matrix if isinstance(matrix, DeviceGraph) else DeviceGraph(matrix, csc, inv_perm, **opts)
        self.q = _Query(self.dg, 1)

    def host_code(self, source_vertex: int) -> BFSResult:
        n = self.dg.n
        levels = np.empty(n, dtype=np.int32)
        st = _lib.BfsStats()
        check(_lib.load().blest_bfs(self.q._h, int(source_vertex), levels.ctypes.data_as(_lib.I32P),
                                    C.byref(st)))
        return BFSResult(st.kernel_ms / 1e3, int(source_vertex), levels, st.total_levels, st.no_visited,
                         st.td_levels, st.bu_levels, st.edges_traversed_component, st.bottom_up_mask)


class CCKernel:
    """CC/CCKernel.cuh:28-54."""

    def __init__(self, matrix: BitMatrix):
        self.matrix = matrix

    def host_code(self) -> CCResult:
        raise NotImplementedError

    def run(self) -> CCResult:
        return self.host_code()


class BVSSCCKernel(CCKernel):
    """Connected components on the GPU; algo 0 = hook/compress, 1 = label propagation."""

    def __init__(self, matrix: BVSS | DeviceGraph, csc: CSC | None = None, inv_perm=None, algo: int = 0, **opts):
        super().__init__(matrix)
        self.dg = matrix if isinstance(matrix, DeviceGraph) else DeviceGraph(matrix, csc, inv_perm, **opts)
        self.q = _Query(self.dg, 1)
        self.algo = algo

    def host_code(self) -> CCResult:
        n = self.dg.n
        labels = np.empty(n, dtype=np.uint32)
        st = _lib.CcStats()
        check(_lib.load().blest_cc(self.q._h, self.algo, _u32p(labels), C.byref(st)))
        return CCResult(st.kernel_ms / 1e3, labels, st.no_components, st.iterations)


class MBFSKernel:
    """MBFS/MBFSKernel.cuh:27-53 (the reference's run() takes no sources; here the
    sources are given at construction or per run)."""

    def __init__(self, matrix: BitMatrix):
        self.matrix = matrix

    def host_code(self, sources) -> MBFSResult:
        raise NotImplementedError

    def run(self, sources=None) -> MBFSResult:
        return self.host_code(sources)


class BVSSMBFSKernel(MBFSKernel):
    def __init__(self, matrix: BVSS | DeviceGraph, csc: CSC | None = None, inv_perm=None, sources=None,
                 algo: int = 0, batch: int = 64, **opts):
        super().__init__(matrix)
        self.dg = matrix if isinstance(matrix, DeviceGraph) else DeviceGraph(matrix, csc, inv_perm, **opts)
        self.sources = None if sources is None else _as_u32(sources)
        self.algo, self.batch = algo, batch
        self.q = _Query(self.dg, batch)

    def host_code(self, sources) -> MBFSResult:
        srcs = self.sources if sources is None else _as_u32(sources)
        n = self.dg.n
        out, total = [], 0.0
        o = _lib.MbfsOpts(self.algo, self.batch)
        for b in range(0, len(srcs), self.batch):
            chunk = np.ascontiguousarray(srcs[b:b + self.batch])
            lv = np.empty((len(chunk), n), dtype=np.int32)
            st = _lib.BfsStats()
            check(_lib.load().blest_mbfs(self.q._h, _u32p(chunk), len(chunk), lv.ctypes.data_as(_lib.I32P),
                                         C.byref(o), C.byref(st)))
            total += st.kernel_ms / 1e3
            out.extend(list(lv))
        return MBFSResult(total, out)


class GraphBatch:
    """Many small graphs traversed by one persistent launch (one root per graph)."""

    def __init__(self, graphs: list[CSC], device: int = 0):
        self._lib = _lib.load()
        arr = (C.c_void_p * len(graphs))(*[g._h.value for g in graphs])
        h = C.c_void_p()
        check(self._lib.blest_graph_batch_create(arr, len(graphs), device, C.byref(h)))
        self._h = h
        self.sizes = [g.n for g in graphs]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.blest_graph_batch_free(h)
            self._h = C.c_void_p(0)

    def bfs(self, roots) -> tuple[list[np.ndarray], float]:
        r = _as_u32(roots)
        total = int(sum(self.sizes))
        lv = np.empty(total, dtype=np.int32)
        ms = C.c_double()
        check(self._lib.blest_graph_batch_bfs(self._h, _u32p(r), lv.ctypes.data_as(_lib.I32P), C.byref(ms)))
        out, off = [], 0
        for s in self.sizes:
            out.append(lv[off:off + s])
            off += s
        return out, ms.value / 1e3


class Server:
    """Request server with continuous batching over resident device graphs.

    submit() queues a single-root BFS ("bfs") or a connected-components ("cc") request
    and returns a ticket; wait(ticket) returns (levels or labels, result dict).
    """

    KINDS = {"bfs": 0, "cc": 1}

    def __init__(self, graphs: list[DeviceGraph], workers: int = 4, max_batch: int = 64,
                 hbm_budget: int = 0, capacity: int = 65536, min_multi: int = 8):
        self._lib = _lib.load()
        o = _lib.ServerOpts()
        self._lib.blest_server_opts_default(C.byref(o))
        o.workers, o.max_batch, o.hbm_budget, o.capacity = workers, max_batch, hbm_budget, capacity
        o.min_multi = min_multi
        arr = (C.c_void_p * len(graphs))(*[g._h.value for g in graphs])
        h = C.c_void_p()
        check(self._lib.blest_server_create(arr, len(graphs), C.byref(o), C.byref(h)))
        self._h = h
        self._graphs = list(graphs)
        self._n = [g.n for g in graphs]
        self._out: dict[int, np.ndarray] = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.blest_server_free(h)
            self._h = C.c_void_p(0)

    def submit(self, graph: int, kind: str = "bfs", root: int = 0, out: np.ndarray | None = None) -> int:
        """out: optional caller array of n int32 (BFS) / uint32 (CC), e.g. pinned_empty()."""
        k = self.KINDS[kind]
        if not 0 <= graph < len(self._n):
            raise BlestError(1, "graph index out of range")
        want = np.int32 if k == 0 else np.uint32
        if out is None:
            out = np.empty(self._n[graph], dtype=want)
        elif out.dtype != want or out.size < self._n[graph] or not out.flags["C_CONTIGUOUS"]:
            raise ValueError(f"out must be a contiguous {np.dtype(want).name} array of at least n entries")
        t = C.c_uint64()
        check(self._lib.blest_server_submit(self._h, graph, k, root, out.ctypes.data_as(C.c_void_p), C.byref(t)))
        self._out[t.value] = out
        return t.value

    def wait(self, ticket: int, timeout_ms: int = 0xFFFFFFFF) -> tuple[np.ndarray, dict]:
        r = _lib.RequestResult()
        st = self._lib.blest_server_wait(self._h, ticket, timeout_ms, C.byref(r))
        if st != 8:  # the ticket stays live on a timeout
            out = self._out.pop(ticket, None)
        check(st)
        return out, {k: getattr(r, k) for k, _ in r._fields_}

    def drain(self) -> None:
        check(self._lib.blest_server_drain(self._h))


class _Pinned:
    def __init__(self, nbytes: int):
        self._lib = _lib.load()
        p = C.c_void_p()
        check(self._lib.blest_host_alloc(nbytes, C.byref(p)))
        self.ptr = p

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            self._lib.blest_host_free(self.ptr)
            self.ptr = None


def pinned_empty(n: int, dtype=np.int32) -> np.ndarray:
    """A numpy array in page-locked host memory (freed with the array)."""
    dt = np.dtype(dtype)
    holder = _Pinned(max(n, 1) * dt.itemsize)
    buf = (C.c_char * (max(n, 1) * dt.itemsize)).from_address(holder.ptr.value)
    arr = np.frombuffer(buf, dtype=dt, count=n)
    buf._holder = holder  # the ctypes buffer (referenced by the array) keeps the allocation alive
    return arr
