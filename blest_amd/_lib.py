"""ctypes declarations of libblest.so (include/blest.h).  Loading fails loudly when the
library has not been built: there is no pure-Python fallback for any entry point."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("BLEST_LIB", _PKG / "libblest.so"))


class BlestError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


# ---- structs (field order = include/blest.h)
class ReorderOpts(C.Structure):
    _fields_ = [("kind", C.c_int), ("slice_size", C.c_uint32), ("window", C.c_uint32), ("jaccard", C.c_int),
                ("gorder", C.c_int), ("threads", C.c_uint32), ("seed", C.c_uint64)]


class CscMetrics(C.Structure):
    _fields_ = [("max_bandwidth", C.c_uint32), ("max_profile", C.c_uint32), ("max_degree", C.c_uint32),
                ("_pad", C.c_uint32), ("avg_bandwidth", C.c_double), ("avg_profile", C.c_double),
                ("avg_degree", C.c_double)]


class BvssInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("slice_size", C.c_uint32), ("no_masks", C.c_uint32),
                ("no_real_slice_sets", C.c_uint32), ("no_virtual_slice_sets", C.c_uint32),
                ("full_padding", C.c_int), ("no_slices", C.c_uint64), ("no_unpadded_slices", C.c_uint64),
                ("no_padded_slices", C.c_uint64), ("edges_check", C.c_uint64),
                ("update_divergence", C.c_double), ("avg_slices_per_set", C.c_double),
                ("std_slices_per_set", C.c_double), ("min_slices_per_set", C.c_uint32),
                ("max_slices_per_set", C.c_uint32), ("compression_ratio", C.c_double)]


class DeviceOpts(C.Structure):
    _fields_ = [("device", C.c_int), ("block_size", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("alpha", C.c_double), ("beta", C.c_double), ("direction", C.c_int), ("keep_csr", C.c_int),
                ("timeout_ms", C.c_uint32), ("use_graphs", C.c_int), ("td_quads", C.c_uint32),
                ("bu_group", C.c_uint32), ("fuse_quads", C.c_int32), ("bu_probe", C.c_int32)]


class DeviceGraphInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("rsets", C.c_uint32), ("vsets", C.c_uint32), ("grid_blocks", C.c_uint32),
                ("block_size", C.c_uint32), ("cus", C.c_uint32), ("symmetric", C.c_uint32),
                ("has_csr", C.c_uint32), ("slices", C.c_uint64), ("nnz", C.c_uint64),
                ("device_bytes", C.c_uint64), ("device", C.c_int32), ("stream_quads", C.c_uint64)]


class BfsStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_levels", C.c_uint32), ("no_visited", C.c_uint32),
                ("td_levels", C.c_uint32), ("bu_levels", C.c_uint32),
                ("edges_traversed_component", C.c_uint64), ("bottom_up_mask", C.c_uint64)]


class LevelTrace(C.Structure):
    _fields_ = [("level", C.c_uint32), ("bottom_up", C.c_uint32), ("expand_ms", C.c_float),
                ("update_ms", C.c_float)]


class MbfsOpts(C.Structure):
    _fields_ = [("algo", C.c_int), ("batch", C.c_uint32)]


class CcStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("no_components", C.c_uint32), ("iterations", C.c_uint32)]


class ServerOpts(C.Structure):
    _fields_ = [("workers", C.c_uint32), ("max_batch", C.c_uint32), ("hbm_budget", C.c_uint64),
                ("capacity", C.c_uint32), ("min_multi", C.c_uint32)]


class RequestResult(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("status", C.c_int32), ("batch_size", C.c_uint32),
                ("queued_ms", C.c_double), ("service_ms", C.c_double), ("kernel_ms", C.c_double),
                ("total_levels", C.c_uint32), ("no_visited", C.c_uint32),
                ("edges_traversed_component", C.c_uint64)]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
I32P = C.POINTER(C.c_int32)
DP = C.POINTER(C.c_double)

# name: (restype, argtypes)
PROTOTYPES = {
    "blest_status_str": (C.c_char_p, [C.c_int]),
    "blest_last_error": (C.c_char_p, []),
    "blest_version": (C.c_char_p, []),
    "blest_csc_create": (C.c_int, [C.c_uint32, U32P, U32P, PP]),
    "blest_csc_from_mtx": (C.c_int, [C.c_char_p, C.c_int, C.c_int, PP]),
    "blest_csc_load_binary": (C.c_int, [C.c_char_p, PP]),
    "blest_csc_save_binary": (C.c_int, [P, C.c_char_p]),
    "blest_csc_generate_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                          C.c_uint64, C.c_int, PP]),
    "blest_csc_generate_grid": (C.c_int, [C.c_uint32, C.c_double, C.c_double, C.c_uint64, PP]),
    "blest_csc_generate_er": (C.c_int, [C.c_uint32, C.c_double, C.c_uint64, PP]),
    "blest_csc_free": (None, [P]),
    "blest_csc_info": (C.c_int, [P, U32P, U64P, C.POINTER(C.c_int)]),
    "blest_csc_export": (C.c_int, [P, U32P, U32P]),
    "blest_csc_is_social": (C.c_int, [P, C.POINTER(C.c_int)]),
    "blest_csc_is_symmetric": (C.c_int, [P, C.POINTER(C.c_int)]),
    "blest_csc_metrics_get": (C.c_int, [P, C.POINTER(CscMetrics)]),
    "blest_csc_pick_roots": (C.c_int, [P, C.c_uint32, C.c_uint64, U32P]),
    "blest_reorder_opts_default": (None, [C.POINTER(ReorderOpts)]),
    "blest_csc_reorder": (C.c_int, [P, C.POINTER(ReorderOpts), U32P]),
    "blest_csc_apply_permutation": (C.c_int, [P, U32P]),
    "blest_ordering_save": (C.c_int, [C.c_char_p, U32P, C.c_uint32]),
    "blest_ordering_load": (C.c_int, [C.c_char_p, U32P, C.c_uint32]),
    "blest_bvss_build": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_int, PP]),
    "blest_bvss_build_device": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_int, C.c_int, PP]),
    "blest_bvss_free": (None, [P]),
    "blest_bvss_info_get": (C.c_int, [P, C.POINTER(BvssInfo)]),
    "blest_bvss_export": (C.c_int, [P, U64P, U32P, U32P, U32P, U32P]),
    "blest_bvss_save_binary": (C.c_int, [P, C.c_char_p]),
    "blest_bvss_load_binary": (C.c_int, [C.c_char_p, PP]),
    "blest_bvss_adopt": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, U64P,
                                   U32P, U32P, U32P, U32P, C.c_double, C.c_int, PP]),
    "blest_bvss_csv_header": (C.c_int, [C.c_char_p, C.c_size_t]),
    "blest_bvss_csv_row": (C.c_int, [P, C.c_char_p, C.c_size_t]),
    "blest_device_opts_default": (None, [C.POINTER(DeviceOpts)]),
    "blest_device_graph_create": (C.c_int, [P, P, U32P, C.POINTER(DeviceOpts), PP]),
    "blest_device_graph_create_csr": (C.c_int, [P, P, U32P, C.POINTER(DeviceOpts), PP]),
    "blest_device_graph_free": (None, [P]),
    "blest_device_graph_info_get": (C.c_int, [P, C.POINTER(DeviceGraphInfo)]),
    "blest_query_create": (C.c_int, [P, C.c_uint32, PP]),
    "blest_query_free": (None, [P]),
    "blest_bfs": (C.c_int, [P, C.c_uint32, I32P, C.POINTER(BfsStats)]),
    "blest_bfs_async": (C.c_int, [P, C.c_uint32, P]),
    "blest_query_sync": (C.c_int, [P, C.POINTER(BfsStats)]),
    "blest_query_levels_device": (P, [P]),
    "blest_bfs_noalloc": (C.c_int, [P, C.c_uint32, I32P, C.POINTER(BfsStats), P]),
    "blest_bfs_trace": (C.c_int, [P, C.c_uint32, I32P, C.POINTER(BfsStats), C.POINTER(LevelTrace), C.c_uint32,
                                  U32P]),
    "blest_mbfs": (C.c_int, [P, U32P, C.c_uint32, I32P, C.POINTER(MbfsOpts), C.POINTER(BfsStats)]),
    "blest_cc": (C.c_int, [P, C.c_int, U32P, C.POINTER(CcStats)]),
    "blest_device_graph_set_thresholds": (C.c_int, [P, C.c_double, C.c_double, C.c_int]),
    "blest_device_graph_set_tiles": (C.c_int, [P, C.c_uint32, C.c_uint32]),
    "blest_device_graph_set_fuse": (C.c_int, [P, C.c_int32]),
    "blest_autotune_bfs": (C.c_int, [P, P, U32P, U32P, C.c_uint32, C.POINTER(DeviceOpts), C.POINTER(DeviceOpts),
                                     DP, U32P]),
    "blest_host_alloc": (C.c_int, [C.c_size_t, PP]),
    "blest_host_free": (None, [P]),
    "blest_device_memory_info": (C.c_int, [C.c_int, U64P, U64P]),
    "blest_query_footprint": (C.c_int, [P, C.c_uint32, U64P]),
    "blest_server_opts_default": (None, [C.POINTER(ServerOpts)]),
    "blest_server_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(ServerOpts), PP]),
    "blest_server_submit": (C.c_int, [P, C.c_uint32, C.c_int, C.c_uint32, P, U64P]),
    "blest_server_wait": (C.c_int, [P, C.c_uint64, C.c_uint32, C.POINTER(RequestResult)]),
    "blest_server_drain": (C.c_int, [P]),
    "blest_server_free": (None, [P]),
    "blest_graph_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_int, PP]),
    "blest_graph_batch_free": (None, [P]),
    "blest_graph_batch_bfs": (C.c_int, [P, U32P, I32P, DP]),
}

_lib = None


def load() -> C.CDLL:
    """Load libblest.so once; raises if it is missing (build it with blest_amd.build)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run `python -m blest_amd.build` first")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status: int) -> None:
    if status != 0:
        lib = load()
        msg = lib.blest_last_error().decode() or lib.blest_status_str(status).decode()
        raise BlestError(status, msg)
