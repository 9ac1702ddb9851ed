"""ctypes binding of libdygrass.so (the C ABI in include/dygrass.h).

The shared object is built in-tree by ``dygrass_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libdygrass.so")
SYNTH_PATH = os.path.join(_PKG, "lib", "libdygrass_synth.so")
KAPPA_PATH = os.path.join(_PKG, "lib", "libdygrass_kappa.so")
BIN_DIR = os.path.join(_PKG, "bin")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)
i8p = C.POINTER(C.c_int8)
f32p = C.POINTER(C.c_float)
vp = C.c_void_p

DYG_WALK_NO_PRUNE = 1
DYG_WALK_PER_EDGE = 2
DYG_RUN_HOST_EDITS = 4
DYG_WALK_DECISION_ONLY = 8

STATUS = {
    0: "DYG_OK", -1: "DYG_ERR_ARG", -2: "DYG_ERR_HIP", -3: "DYG_ERR_NOMEM", -4: "DYG_ERR_IO",
    -5: "DYG_ERR_FORMAT", -6: "DYG_ERR_STATE", -7: "DYG_ERR_ADMISSION",
}


class WalkParams(C.Structure):
    _fields_ = [("walkers", C.c_uint32), ("max_steps", C.c_uint32), ("threshold", C.c_float),
                ("flags", C.c_uint32), ("seed", C.c_uint64), ("block_edges", C.c_uint32),
                ("reserved", C.c_uint32)]


class WalkStats(C.Structure):
    _fields_ = [("hops", C.c_uint64), ("redirects", C.c_uint64), ("leaf_stops", C.c_uint64),
                ("kernel_ms", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [("updates", C.c_uint64), ("added", C.c_uint64), ("batches", C.c_uint64),
                ("hops", C.c_uint64), ("redirects", C.c_uint64), ("leaf_stops", C.c_uint64),
                ("walk_ms", C.c_double), ("update_ms", C.c_double), ("total_ms", C.c_double)]


class StreamParams(C.Structure):
    _fields_ = [("ins", WalkParams), ("del_", WalkParams), ("heur_walkers", C.c_uint32),
                ("max_batch", C.c_uint32), ("max_inserts", C.c_uint64), ("slack", C.c_uint32),
                ("flags", C.c_uint32)]


class StreamStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("updates", "weight_updates", "found", "added",
                                          "skipped", "g_only", "recovered", "heuristic",
                                          "invalid", "candidates", "hops")] + [
        ("walk_ms", C.c_double), ("total_ms", C.c_double), ("graph", C.c_uint64),
        ("redirects", C.c_uint64)]


class ServeResult(C.Structure):
    _fields_ = [("n", C.c_uint32), ("queue_ms", C.c_double), ("latency_ms", C.c_double),
                ("hops", C.c_uint64), ("redirects", C.c_uint64), ("walk_ms", C.c_double)]


class MtxInfo(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("has_header", "has_dimensions", "weighted", "laplacian",
                                       "triangular", "base")] + [("entries", C.c_uint64)]


DYG_STREAM_INSERT, DYG_STREAM_DELETE = 0, 1
DYG_STREAM_PLAIN = 1
DYG_STREAM_GENERAL_ONLY = 2
DYG_STREAM_DENSE_NO_DESC = 4
DYG_STREAM_GRAPH = 8
DYG_STREAM_SMALL_1024 = 16
DYG_NOALLOC_MAX_EDGES = 1048576
OUTCOMES = {1: "weight", 2: "found", 3: "added", 4: "skip", 5: "g_only", 6: "recovered",
            7: "heuristic", 8: "invalid"}


class DygError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


# name -> (restype, argtypes)
_SIGS = {
    "dyg_version": (C.c_char_p, []),
    "dyg_ctx_create": (C.c_int32, [C.c_int, C.POINTER(vp)]),
    "dyg_ctx_destroy": (None, [vp]),
    "dyg_last_error": (C.c_char_p, [vp]),
    "dyg_ctx_stream": (vp, [vp]),
    "dyg_ctx_sync": (C.c_int32, [vp]),
    "dyg_dev_alloc": (C.c_int32, [vp, C.c_size_t, C.POINTER(vp)]),
    "dyg_dev_free": (C.c_int32, [vp, vp]),
    "dyg_copy_to_device": (C.c_int32, [vp, vp, vp, C.c_size_t]),
    "dyg_copy_to_host": (C.c_int32, [vp, vp, vp, C.c_size_t]),
    "dyg_copy_on_device": (C.c_int32, [vp, vp, vp, C.c_size_t]),
    "dyg_graph_check": (C.c_int32, [vp, u64p]),
    "dyg_graph_upload": (C.c_int32, [vp, C.c_uint32, u64p, u32p, f32p, C.c_uint32, C.c_int,
                                     C.POINTER(vp)]),
    "dyg_graph_info": (C.c_int32, [vp, u32p, u64p, u64p]),
    "dyg_graph_download": (C.c_int32, [vp, u64p, u32p, f32p]),
    "dyg_graph_insert": (C.c_int32, [vp, u32p, u32p, f32p, C.c_uint32]),
    "dyg_graph_delete": (C.c_int32, [vp, u32p, u32p, C.c_uint32, i8p]),
    "dyg_graph_free": (None, [vp]),
    "dyg_walk_incremental": (C.c_int32, [vp, u32p, u32p, f32p, u64p, C.c_uint32,
                                         C.POINTER(WalkParams), i32p, f32p, u32p, u32p,
                                         C.POINTER(WalkStats)]),
    "dyg_walk_incremental_noalloc": (C.c_int32, [vp, u32p, u32p, f32p, u64p, C.c_uint32,
                                                 C.POINTER(WalkParams), i32p, f32p, u32p, u32p,
                                                 C.POINTER(WalkStats)]),
    "dyg_walk_decremental": (C.c_int32, [vp, u32p, u32p, u64p, C.c_uint32,
                                         C.POINTER(WalkParams), i32p, f32p, u32p, u32p, u32p,
                                         C.POINTER(WalkStats)]),
    "dyg_walk_heuristic": (C.c_int32, [vp, u32p, u64p, C.c_uint32, C.c_uint32, C.c_uint64,
                                       u32p]),
    "dyg_walk_incremental_dev": (C.c_int32, [vp, vp, vp, vp, vp, C.c_uint32,
                                             C.POINTER(WalkParams), vp, vp, vp, vp, vp]),
    "dyg_walk_decremental_dev": (C.c_int32, [vp, vp, vp, vp, C.c_uint32,
                                             C.POINTER(WalkParams), vp, vp, vp, vp, vp, vp]),
    "dyg_incremental_update": (C.c_int32, [vp, u32p, u32p, f32p, C.c_uint32, C.c_uint32,
                                           C.POINTER(WalkParams), i32p, C.POINTER(RunStats)]),
    "dyg_incremental_update_dev": (C.c_int32, [vp, vp, vp, vp, C.c_uint32, C.c_uint32,
                                               C.POINTER(WalkParams), vp, vp,
                                               C.POINTER(RunStats)]),
    "dyg_graph_reserve": (C.c_int32, [vp, C.c_uint64, C.c_uint32]),
    "dyg_dual_create": (C.c_int32, [vp, C.c_uint32, u64p, u32p, f32p, u32p, u32p, C.c_uint64,
                                    C.POINTER(vp)]),
    "dyg_dual_key_multiplier": (C.c_int64, [vp]),
    "dyg_dual_remove": (C.c_int32, [vp, u32p, u32p, u64p, C.c_uint32, i8p, u32p]),
    "dyg_dual_walk": (C.c_int32, [vp, C.POINTER(WalkParams), u32p, u64p, C.POINTER(WalkStats)]),
    "dyg_dual_heuristic": (C.c_int32, [vp, C.c_uint32, C.c_uint64]),
    "dyg_dual_sparse_size": (C.c_uint64, [vp]),
    "dyg_dual_added_count": (C.c_uint64, [vp]),
    "dyg_dual_added_edges": (C.c_int32, [vp, u32p, u32p]),
    "dyg_dual_dense": (vp, [vp]),
    "dyg_dual_free": (None, [vp]),
    "dyg_stream_create": (C.c_int32, [vp, C.c_uint32, u64p, u32p, f32p, u64p, u32p, f32p,
                                      C.POINTER(StreamParams), C.POINTER(vp)]),
    # the per-batch calls take raw addresses (addr()): a typed ctypes pointer
    # costs microseconds to build, which a batch of a few updates notices
    "dyg_stream_apply": (C.c_int32, [vp, C.c_int, vp, vp, vp, C.c_uint32, vp,
                                     C.POINTER(StreamStats)]),
    "dyg_stream_apply_dev": (C.c_int32, [vp, C.c_int, vp, vp, vp, C.c_uint32, vp,
                                         C.POINTER(StreamStats)]),
    "dyg_stream_apply_many_dev": (C.c_int32, [vp, C.c_uint32, vp, vp, vp, vp, vp, vp,
                                              C.POINTER(StreamStats)]),
    "dyg_stream_graph": (vp, [vp, C.c_int]),
    "dyg_stream_position": (C.c_uint64, [vp]),
    "dyg_stream_free": (None, [vp]),
    "dyg_serve_open": (C.c_int32, [vp, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(vp)]),
    "dyg_serve_submit": (C.c_int32, [vp, C.c_int, vp, vp, vp, C.c_uint32,
                                     C.POINTER(C.c_uint64)]),
    "dyg_serve_submit_dev": (C.c_int32, [vp, C.c_int, vp, vp, vp, C.c_uint32,
                                         C.POINTER(C.c_uint64)]),
    "dyg_serve_wait": (C.c_int32, [vp, C.c_uint64, vp, C.POINTER(ServeResult)]),
    "dyg_serve_close": (None, [vp]),
    "dyg_serve_fail_next_copy": (C.c_int32, [vp]),
    "dyg_mtx_read_graph": (C.c_int32, [C.c_char_p, C.c_int, C.c_int, C.POINTER(vp)]),
    "dyg_mtx_read_edges": (C.c_int32, [C.c_char_p, C.c_int, C.POINTER(vp)]),
    "dyg_mtx_read_auto": (C.c_int32, [C.c_char_p, C.POINTER(MtxInfo), C.POINTER(vp)]),
    "dyg_mtx_nv": (C.c_uint32, [vp]),
    "dyg_mtx_nlines": (C.c_uint64, [vp]),
    "dyg_mtx_row_ptr": (u64p, [vp]),
    "dyg_mtx_col": (u32p, [vp]),
    "dyg_mtx_w": (f32p, [vp]),
    "dyg_mtx_line_u": (u32p, [vp]),
    "dyg_mtx_line_v": (u32p, [vp]),
    "dyg_mtx_line_w": (f32p, [vp]),
    "dyg_mtx_free": (None, [vp]),
    "dyg_mtx_write_weighted": (C.c_int32, [C.c_char_p, u32p, u32p, f32p, C.c_uint64, C.c_int]),
}

_SYNTH_SIGS = {
    "dyg_synth_free": (None, [vp]),
    "dyg_synth_nv": (C.c_uint32, [vp]),
    "dyg_synth_ne": (C.c_uint64, [vp]),
    "dyg_synth_u": (u32p, [vp]),
    "dyg_synth_v": (u32p, [vp]),
    "dyg_synth_w": (f32p, [vp]),
    "dyg_synth_xy": (f32p, [vp]),
    "dyg_synth_grid2d": (vp, [C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_uint64]),
    "dyg_synth_rgg_knn": (vp, [C.c_uint32, C.c_uint32, C.c_uint64]),
    "dyg_synth_trimesh": (vp, [C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_uint64]),
    "dyg_synth_grid3d_sc": (vp, [C.c_uint32, C.c_uint64, C.c_float, C.c_float, C.c_uint64]),
    "dyg_synth_sparsifier": (vp, [C.c_uint32, u32p, u32p, f32p, C.c_uint64, C.c_uint64,
                                  C.c_uint64]),
    "dyg_synth_insertions": (vp, [C.c_uint32, u32p, u32p, C.c_uint64, f32p, C.c_uint64,
                                  C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_uint64]),
    "dyg_synth_insertions_walk": (vp, [C.c_uint32, u32p, u32p, C.c_uint64, C.c_uint64,
                                       C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                       C.c_uint64]),
    "dyg_synth_deletions": (vp, [C.c_uint32, u32p, u32p, f32p, C.c_uint64, C.c_uint64,
                                 C.c_uint64]),
    "dyg_synth_write_mtx": (C.c_int, [C.c_char_p, u32p, u32p, f32p, C.c_uint64, C.c_int]),
}

_lib = None
_synth = None


def _bind(path: str, sigs: dict) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: the HIP library has not been built "
            "(run `python -m dygrass_amd.build`); there is no fallback path")
    lib = C.CDLL(path)
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = _bind(LIB_PATH, _SIGS)
    return _lib


def synth_lib() -> C.CDLL:
    global _synth
    if _synth is None:
        _synth = _bind(SYNTH_PATH, _SYNTH_SIGS)
    return _synth


_KAPPA_SIGS = {
    "dyg_kappa_spmv": (None, [C.c_int64, vp, vp, vp, vp, vp, C.c_int]),
    "dyg_kappa_pcg": (C.c_int, [C.c_int64, vp, vp, vp, vp, vp, C.c_double, C.c_int, C.c_int]),
}
_kappa = None


def kappa_lib() -> C.CDLL:
    """The host FP64 kernels of the kappa check (libdygrass_kappa.so)."""
    global _kappa
    if _kappa is None:
        _kappa = _bind(KAPPA_PATH, _KAPPA_SIGS)
    return _kappa


def ptr(a: np.ndarray | None, ct):
    """Pointer to a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(ct)


def addr(a: np.ndarray | None):
    """Address of a contiguous numpy array, for void-pointer arguments (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data


def check(status: int, ctx=None) -> None:
    if status != 0:
        msg = lib().dyg_last_error(ctx)
        raise DygError(status, msg.decode() if msg else "")
