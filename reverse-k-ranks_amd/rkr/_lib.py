"""ctypes binding of librkr.so (the C ABI declared in include/rkr.h).

The shared library is built in-tree (reverse-k-ranks_amd/librkr.so) by
``make -C reverse-k-ranks_amd``.  There is no fallback: when the library is
missing or no GPU is present, every call fails loudly.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "librkr.so")

RKR_OK = 0
RKR_ERR_INVALID = -1
RKR_ERR_STATE = -2
RKR_ERR_NOMEM = -3
RKR_ERR_HIP = -4
RKR_ERR_IO = -5
RKR_ERR_NODEVICE = -6

RKR_FORM_PLAIN = 0
RKR_FORM_SCREENED = 1
RKR_OUTPUT_EXACT = 0
RKR_OUTPUT_REFERENCE = 1
RKR_OPT_FORM = 1
RKR_OPT_OUTPUT = 2
RKR_OPT_QUERY_BATCH = 3
RKR_OPT_BUILD_BATCH = 4
RKR_OPT_SCREEN_CHUNK = 5
RKR_OPT_DEDUP = 6
RKR_OPT_SCREEN_ROWS = 7
RKR_OPT_DENSE_TOP = 8
RKR_OPT_COARSE = 9
RKR_OPT_QUERY_SLOTS = 10

STATUS_NAMES = {
    RKR_OK: "RKR_OK", RKR_ERR_INVALID: "RKR_ERR_INVALID", RKR_ERR_STATE: "RKR_ERR_STATE",
    RKR_ERR_NOMEM: "RKR_ERR_NOMEM", RKR_ERR_HIP: "RKR_ERR_HIP", RKR_ERR_IO: "RKR_ERR_IO",
    RKR_ERR_NODEVICE: "RKR_ERR_NODEVICE",
}

# Every symbol include/rkr.h declares.
EXPORTS = (
    "rkr_create", "rkr_destroy", "rkr_last_error", "rkr_set_option", "rkr_get_option", "rkr_sync",
    "rkr_version", "rkr_normalize_users", "rkr_set_data", "rkr_uniform_sample_ranks", "rkr_build_table",
    "rkr_set_table", "rkr_get_table", "rkr_get_ranks", "rkr_save_table", "rkr_load_table", "rkr_sorted_scores",
    "rkr_sorted_scores_noalloc", "rkr_save_score_table", "rkr_sample_items", "rkr_query_rank_distribution", "rkr_sort_kth_rank",
    "rkr_find_optimal_sample", "rkr_query",
    "rkr_index_size_bytes", "rkr_query_table_cols", "rkr_insert_items", "rkr_delete_items", "rkr_insert_users", "rkr_delete_users",
    "rkr_get_counts", "rkr_reserve", "rkr_now", "rkr_server_start", "rkr_server_submit_query", "rkr_server_submit_update",
    "rkr_server_wait", "rkr_server_poll", "rkr_server_counters", "rkr_server_rebuilds", "rkr_server_stop",
    "rkr_qsrp_build", "rkr_qsrp_info", "rkr_qsrp_screen", "rkr_qsrp_phi", "rkr_qsrp_fit_line",
)

RKR_OPT_QSRP = 11

RKR_SAMPLE_OPTIMAL_PART = 0
RKR_SAMPLE_OPTIMAL_UNIFORM = 1

RKR_UPDATE_INSERT_USERS = 1
RKR_UPDATE_INSERT_ITEMS = 2
RKR_UPDATE_DELETE_USERS = 3
RKR_UPDATE_DELETE_ITEMS = 4


class QueryStats(C.Structure):
    """rkr_query_stats (mirrors SingleQueryPerformance, src/include/util/FileIO.hpp:150-173)."""
    _fields_ = [("query_id", C.c_int32), ("n_prune_user", C.c_int32), ("n_result_user", C.c_int32),
                ("n_refine_user", C.c_int32), ("n_candidate", C.c_int32), ("reserved", C.c_int32),
                ("ip_cost", C.c_int64), ("n_rechecked", C.c_int64), ("n_screened", C.c_int64)]


# numpy view of rkr_query_stats (same field order, sizes and offsets)
QUERY_STATS_DTYPE = np.dtype([(f, np.int32 if t is C.c_int32 else np.int64) for f, t in QueryStats._fields_],
                             align=True)
assert QUERY_STATS_DTYPE.itemsize == C.sizeof(QueryStats)


class BatchTiming(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("ip_ms", C.c_double), ("bound_ms", C.c_double),
                ("prune_ms", C.c_double), ("refine_ms", C.c_double), ("recheck_ms", C.c_double),
                ("topk_ms", C.c_double),
                ("h2d_ms", C.c_double), ("d2h_ms", C.c_double), ("n_batches", C.c_int64),
                ("n_refined", C.c_int64), ("n_rechecked", C.c_int64), ("n_screened", C.c_int64),
                ("max_screen_err", C.c_double), ("n_listed", C.c_int64), ("n_fallback", C.c_int64),
                ("n_coarse_refined", C.c_int64)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("n_slots", C.c_int32), ("update_mode", C.c_int32),
                ("update_cus", C.c_int32), ("rebuild_pending", C.c_int32)]


RKR_SERVE_BARRIER = 0
RKR_SERVE_SNAPSHOT = 1


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("batch_size", C.c_int32), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double), ("stats", QueryStats), ("error", C.c_char * 256)]


class RkrError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_lib = None


def load():
    """Loads librkr.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR}` "
                           "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f32p, i32p = C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p
    L.rkr_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]
    L.rkr_destroy.argtypes = [vp]
    L.rkr_destroy.restype = None
    L.rkr_last_error.argtypes = [vp]
    L.rkr_last_error.restype = C.c_char_p
    L.rkr_set_option.argtypes = [vp, C.c_int, i64]
    L.rkr_get_option.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
    L.rkr_sync.argtypes = [vp]
    L.rkr_version.restype = C.c_char_p
    L.rkr_normalize_users.argtypes = [f32p, i64, i32]
    L.rkr_set_data.argtypes = [vp, f32p, i64, f32p, i64, i32, C.c_int]
    L.rkr_uniform_sample_ranks.argtypes = [i64, i32, i32p]
    L.rkr_build_table.argtypes = [vp, i32p, i32, C.POINTER(C.c_double)]
    L.rkr_set_table.argtypes = [vp, i32p, i32, f32p, C.c_float]
    L.rkr_get_table.argtypes = [vp, i64, i64, f32p]
    L.rkr_get_ranks.argtypes = [vp, i32p, C.POINTER(C.c_int32)]
    L.rkr_save_table.argtypes = [vp, C.c_char_p]
    L.rkr_load_table.argtypes = [vp, C.c_char_p, C.c_float]
    L.rkr_sorted_scores.argtypes = [vp, i64, i64, f32p]
    L.rkr_sorted_scores_noalloc.argtypes = [vp, i64, i64, C.c_void_p]
    L.rkr_save_score_table.argtypes = [vp, C.c_char_p, i64, C.c_void_p]
    L.rkr_sample_items.argtypes = [i64, i32, i32p]
    L.rkr_query_rank_distribution.argtypes = [vp, i32p, i32, i32p, C.POINTER(C.c_double)]
    L.rkr_sort_kth_rank.argtypes = [i32p, i64, i32, i32, i32p, i32p]
    L.rkr_find_optimal_sample.argtypes = [vp, i64, i64, i32, i32p, i32p, i32, i32, i32p, C.POINTER(C.c_int64)]
    L.rkr_query.argtypes = [vp, f32p, i32, i32p, i32, i32p, i32p, f32p, C.POINTER(QueryStats),
                            C.POINTER(BatchTiming)]
    L.rkr_insert_items.argtypes = [vp, f32p, i64]
    L.rkr_delete_items.argtypes = [vp, C.c_void_p, i64]
    L.rkr_insert_users.argtypes = [vp, f32p, i64]
    L.rkr_delete_users.argtypes = [vp, C.c_void_p, i64]
    L.rkr_get_counts.argtypes = [vp] + [C.POINTER(C.c_int64)] * 4
    L.rkr_reserve.argtypes = [vp, i64, i64]
    L.rkr_now.restype = C.c_double
    L.rkr_server_start.argtypes = [vp, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]
    L.rkr_server_submit_query.argtypes = [vp, f32p, i32, C.POINTER(C.c_uint64)]
    L.rkr_server_submit_update.argtypes = [vp, C.c_int, f32p, C.c_void_p, i64, C.POINTER(C.c_uint64)]
    L.rkr_server_wait.argtypes = [vp, C.c_uint64, i32p, i32p, f32p, C.POINTER(RequestInfo)]
    L.rkr_server_poll.argtypes = [vp, C.c_uint64, C.POINTER(C.c_int)]
    L.rkr_server_counters.argtypes = [vp] + [C.POINTER(C.c_int64)] * 3
    L.rkr_server_stop.argtypes = [vp]
    L.rkr_server_rebuilds.argtypes = [vp, C.POINTER(C.c_int64)]
    L.rkr_qsrp_build.argtypes = [vp, C.c_float, C.POINTER(C.c_double)]
    L.rkr_qsrp_info.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
    L.rkr_qsrp_screen.argtypes = [vp, f32p, i32, C.c_void_p]
    L.rkr_qsrp_phi.argtypes = [vp, f32p, i64, f32p]
    L.rkr_qsrp_fit_line.argtypes = [C.c_void_p, C.c_void_p, i32] + [C.POINTER(C.c_double)] * 3
    L.rkr_index_size_bytes.argtypes = [vp]
    L.rkr_index_size_bytes.restype = C.c_int64
    L.rkr_query_table_cols.argtypes = [vp]
    L.rkr_query_table_cols.restype = C.c_int32
    for name in EXPORTS:
        if name not in ("rkr_destroy", "rkr_last_error", "rkr_version", "rkr_index_size_bytes", "rkr_now",
                        "rkr_query_table_cols"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L
