"""ctypes binding of libxtric.so (the C ABI declared in include/xtric.h).

The library is the product: every count runs in its HIP kernels.  If the
shared object is missing this module raises at load time; nothing falls back
to a CPU path.
"""
import ctypes
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libxtric.so")
HEADER = os.path.join(ROOT, "include", "xtric.h")

XT_OK = 0
XT_ERR_ARG = -1
XT_ERR_HIP = -2
XT_ERR_NOMEM = -3
XT_ERR_INPUT = -4
XT_ERR_CAPACITY = -5
XT_ERR_NOTFOUND = -6
XT_ERR_STATE = -7
XT_ERR_INTERNAL = -8
XT_PENDING = 1

XT_ORIENT_DEGREE = 0x0
XT_ORIENT_ID = 0x1
XT_PER_VERTEX = 0x2
XT_VALIDATE = 0x4
XT_CHECK_SYMMETRY = 0x8
XT_CLEAN = 0x10

XT_BATCH_PERSISTENT = 0
XT_BATCH_HIPGRAPH = 1
XT_BATCH_PER_GRAPH = 2

XT_H2D = 1
XT_D2H = 2
XT_D2D = 3
XT_D2D_KERNEL = 4

XT_RGG_STRICT = 0x1
XT_RGG_CONNECT = 0x2

_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_int = ctypes.c_int
_dbl = ctypes.c_double
_p = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)


class XtTiming(ctypes.Structure):
    _fields_ = [("orient_ms", _dbl), ("count_ms", _dbl), ("total_ms", _dbl),
                ("light_ms", _dbl), ("items_ms", _dbl), ("group_ms", _dbl),
                ("ne_plus", _i64), ("n_heavy", _i64), ("n_tiles", _i64), ("n_wave_tiles", _i64)]


class XtGraphStats(ctypes.Structure):
    _fields_ = [("alg_bytes", _u64), ("sum_min", _u64), ("ne_plus", _i64),
                ("sum_min_heavy", _u64), ("heavy_edges", _i64), ("group_bytes", _u64)]


class XtBatchTiming(ctypes.Structure):
    _fields_ = [("upload_ms", _dbl), ("run_ms", _dbl), ("n_graphs", _i64),
                ("n_vertices", _i64), ("n_edges", _i64), ("n_tiles", _i64)]


class XtServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", _i64), ("max_batch_graphs", _i64), ("max_inflight", _i64),
                ("max_queue", _i64), ("max_wait_us", _dbl)]


class XtServeStats(ctypes.Structure):
    _fields_ = [("submitted", _i64), ("completed", _i64), ("rejected", _i64), ("batches", _i64),
                ("p50_us", _dbl), ("p99_us", _dbl), ("mean_us", _dbl), ("max_us", _dbl),
                ("graphs_per_s", _dbl), ("peak_resident_bytes", _i64), ("isolated_batches", _i64),
                ("admit_s", _dbl), ("stage_s", _dbl), ("launch_s", _dbl), ("complete_s", _dbl),
                ("device_s", _dbl)]


_SIGS = {
    "xt_version": (ctypes.c_char_p, []),
    "xt_ctx_create": (_int, [_int, _pp]),
    "xt_ctx_destroy": (_int, [_p]),
    "xt_last_error": (ctypes.c_char_p, [_p]),
    "xt_last_timing": (_int, [_p, ctypes.POINTER(XtTiming)]),
    "xt_ctx_tune": (_int, [_p, _int, _int, _int]),
    "xt_ctx_tune_tiles": (_int, [_p, _int, _int]),
    "xt_ctx_tune_waves": (_int, [_p, _int]),
    "xt_ctx_tune_hash": (_int, [_p, _int]),
    "xt_ctx_force_wide": (_int, [_p, _int]),
    "xt_ctx_fail_allocs": (_int, [_p, _int]),
    "xt_dev_alloc": (_int, [_p, ctypes.c_size_t, _pp]),
    "xt_dev_free": (_int, [_p, _p]),
    "xt_memcpy": (_int, [_p, _p, _p, ctypes.c_size_t, _int]),
    "xt_dev_sync": (_int, [_p]),
    "xt_mem_ceiling": (_int, [_p, _int, _i64, _int, _i64, _int, _p]),
    "xt_count_csr": (_int, [_p, _i64, _i64, _p, _p, _u32, _p, _p]),
    "xt_single_gpu_count": (_int, [_p, _i64, _p, _p, _p, _p]),
    "xt_single_gpu_count_noalloc": (_int, [_p, _i64, _p, _p, _p, _p]),
    "xt_single_gpu_count_masked": (_int, [_p, _i64, _p, _p, _p, _p, _p]),
    "xt_local_count_gpu_inner": (_int, [_p, _i64, _i64, _i64, _i64, _p, _p, _p]),
    "xt_graph_from_csr": (_int, [_p, _i64, _i64, _p, _p, _int, _u32, _pp]),
    "xt_graph_rmat": (_int, [_p, _int, _int, _dbl, _dbl, _dbl, _u64, _int, _pp]),
    "xt_graph_er": (_int, [_p, _i64, _u64, _u64, _pp]),
    "xt_graph_from_edges": (_int, [_p, _i64, _i64, _p, _p, _pp]),
    "xt_graph_from_vite": (_int, [_p, ctypes.c_char_p, _int, _u32, _pp]),
    "xt_count_edges": (_int, [_p, _i64, _i64, _p, _p, _u32, _p, _p]),
    "xt_graph_info": (_int, [_p, _p, _p]),
    "xt_graph_download": (_int, [_p, _p, _p, _p]),
    "xt_graph_count": (_int, [_p, _p, _u32, _p, _p]),
    "xt_graph_slice": (_int, [_p, _p, _i64, _i64, _pp]),
    "xt_graph_count_owned": (_int, [_p, _p, _i64, _i64, _u32, _p, _p, _p]),
    "xt_graph_alg_bytes": (_int, [_p, _p, _u32, _p, _p, _p]),
    "xt_graph_stats_get": (_int, [_p, _p, _u32, ctypes.POINTER(XtGraphStats)]),
    "xt_graph_destroy": (_int, [_p, _p]),
    "xt_rgg_generate": (_int, [_i64, _u32, _dbl, _p, _p, _i64, _p]),
    "xt_rgg_generate_ranks": (_int, [_i64, _int, _u32, _dbl, _p, _p, _i64, _p]),
    "xt_rgg_radius": (_dbl, [_i64]),
    "xt_rgg_generate_opts": (_int, [_i64, _int, _u32, _dbl, _u32, _p, _p, _i64, _p]),
    "xt_rgg_radius_connect": (_dbl, [_i64]),
    "xt_vite_read": (_int, [ctypes.c_char_p, _int, _p, _p, _p, _p]),
    "xt_vite_write": (_int, [ctypes.c_char_p, _int, _i64, _p, _p]),
    "xt_edgelist_to_csr": (_int, [ctypes.c_char_p, _p, _p, _p, _p]),
    "xt_partition_vertices": (_int, [_i64, _p, _i64, _int, _p]),
    "xt_batch_create": (_int, [_p, _i64, _i64, _i64, _pp]),
    "xt_batch_add": (_int, [_p, _i64, _i64, _p, _p, _p]),
    "xt_batch_add_generated": (_int, [_p, _u64, _i64, _i64, _i64, _i64]),
    "xt_batch_run": (_int, [_p, _int, _u32]),
    "xt_batch_results": (_int, [_p, _p, _i64]),
    "xt_batch_timing_get": (_int, [_p, ctypes.POINTER(XtBatchTiming)]),
    "xt_batch_stats_get": (_int, [_p, _u32, ctypes.POINTER(XtGraphStats)]),
    "xt_batch_clear": (_int, [_p]),
    "xt_batch_destroy": (_int, [_p]),
    "xt_batch_params": (_int, [_u64, _i64, _i64, _i64, _p, _p, _p, _p]),
    "xt_batch_graph_csr": (_int, [_u64, _i64, _i64, _i64, _p, _p, _p, _p]),
    "xt_serve_create": (_int, [_p, ctypes.POINTER(XtServeConfig), _pp]),
    "xt_serve_submit": (_int, [_p, _i64, _i64, _p, _p, _dbl, _p]),
    "xt_serve_submit_borrowed": (_int, [_p, _i64, _i64, _p, _p, _dbl, _p]),
    "xt_serve_poll": (_int, [_p, _u64, _p, _p]),
    "xt_serve_drain": (_int, [_p, _dbl]),
    "xt_serve_replay": (_int, [_p, _i64, _p, _p, _p, _p, _p, _i64, _p, _p, _dbl, _p, _p]),
    "xt_serve_stats_get": (_int, [_p, ctypes.POINTER(XtServeStats)]),
    "xt_serve_now_us": (_dbl, [_p]),
    "xt_serve_ticket_state": (_int, [_p, _u64]),
    "xt_serve_pause": (_int, [_p, _int]),
    "xt_serve_destroy": (_int, [_p]),
}

_lib = None


def header_functions():
    """Names of the functions include/xtric.h declares."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(xt_[a-z0-9_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libxtric.so is not built (run python -m xxtric_amd.build); "
                               "there is no fallback path")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:  # reported by the export test; calling it raises AttributeError
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data


class XtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("xtric error %d: %s" % (code, msg))
        self.code = code


def csr_arrays(off, tails, tails_dtype=np.int64):
    off = np.ascontiguousarray(off, dtype=np.int64)
    tails = np.ascontiguousarray(tails, dtype=tails_dtype)
    return off, tails
