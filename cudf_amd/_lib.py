"""ctypes binding of libcudf_amd.so (the C ABI declared in include/mgb_groupby.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C
cudf_amd/csrc``). Loading fails loudly when it is missing: there is no CPU or
PyTorch fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os
import threading

# MGB_LIB_PATH: another build of the library (the host sanitizer build)
LIB_PATH = os.environ.get("MGB_LIB_PATH") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libcudf_amd.so")

MGB_SUCCESS = 0
MGB_ERROR_LOGIC = 1
MGB_ERROR_UNSUPPORTED = 2
MGB_ERROR_HIP = 3
MGB_ERROR_ALLOC = 4


class mgb_column(ctypes.Structure):
    _fields_ = [
        ("type_id", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("size", ctypes.c_int64),
        ("data", ctypes.c_void_p),
        ("null_mask", ctypes.c_void_p),
        ("offset", ctypes.c_int64),
        ("null_count", ctypes.c_int64),
    ]


class mgb_aggregation(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("ddof", ctypes.c_int32), ("null_policy", ctypes.c_int32),
                ("interpolation", ctypes.c_int32), ("quantile", ctypes.c_double)]


class mgb_request(ctypes.Structure):
    _fields_ = [
        ("values", mgb_column),
        ("aggs", ctypes.POINTER(mgb_aggregation)),
        ("num_aggs", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)
DEALLOC_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)


ALLOC_MANY_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_size_t),
                                 ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p)


class mgb_allocator(ctypes.Structure):
    pass


mgb_allocator._fields_ = [("allocate", ALLOC_FN), ("deallocate", DEALLOC_FN), ("ctx", ctypes.c_void_p),
                          ("allocate_many", ALLOC_MANY_FN), ("temporaries", ctypes.POINTER(mgb_allocator))]


class mgb_pool_stats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "limit_bytes", "held_bytes", "in_use_bytes", "peak_in_use_bytes", "chunks", "stream_waits")]


class mgb_out_column(ctypes.Structure):
    _fields_ = [
        ("type_id", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("size", ctypes.c_int64),
        ("data", ctypes.c_void_p),
        ("null_mask", ctypes.c_void_p),
        ("null_count", ctypes.c_int64),
        ("data_bytes", ctypes.c_size_t),
        ("mask_bytes", ctypes.c_size_t),
        ("child_type_id", ctypes.c_int32),
        ("child_reserved", ctypes.c_int32),
        ("child_size", ctypes.c_int64),
        ("child_data", ctypes.c_void_p),
        ("child_null_mask", ctypes.c_void_p),
        ("child_null_count", ctypes.c_int64),
        ("child_data_bytes", ctypes.c_size_t),
        ("child_mask_bytes", ctypes.c_size_t),
    ]


class mgb_batch_options(ctypes.Structure):
    _fields_ = [
        ("force_mode", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("direct_item_rows", ctypes.c_int64),
        ("estimate_scale", ctypes.c_double),
        ("estimates", ctypes.POINTER(ctypes.c_int64)),
        ("alloc", ctypes.POINTER(mgb_allocator)),
    ]


class mgb_batch_stats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "num_queries", "direct_queries", "global_queries", "partitioned_queries", "num_items",
        "output_rows_reserved", "device_bytes", "reruns")]


class mgb_server_options(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_int64),
        ("max_inflight", ctypes.c_int32),
        ("max_batch_queries", ctypes.c_int32),
        ("batching", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("reserve_bytes", ctypes.c_int64),
    ]


class mgb_server_request(ctypes.Structure):
    _fields_ = [
        ("column", ctypes.c_int32),
        ("num_aggs", ctypes.c_int32),
        ("aggs", ctypes.POINTER(mgb_aggregation)),
    ]


class mgb_server_stats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "submitted", "completed", "failed", "batches", "deferred", "reruns", "peak_held_bytes", "resident_bytes",
        "held_bytes", "pending")] + [(name, ctypes.c_double) for name in (
            "mean_batch_queries", "host_plan_ms", "host_finish_ms")] + [("device_bytes", ctypes.c_int64)]


# Every symbol include/mgb_groupby.h declares, with its ctypes signature.
SIGNATURES = {
    "mgb_version": (ctypes.c_int32, []),
    "mgb_last_error_message": (ctypes.c_char_p, []),
    "mgb_target_type": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32]),
    "mgb_groupby_aggregate": (
        ctypes.c_int,
        [
            ctypes.POINTER(mgb_column), ctypes.c_int32, ctypes.c_int32,
            ctypes.POINTER(mgb_request), ctypes.c_int32,
            ctypes.POINTER(mgb_out_column), ctypes.POINTER(mgb_out_column),
            ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.POINTER(mgb_allocator),
        ],
    ),
    "mgb_groupby_aggregate_noalloc": (
        ctypes.c_int,
        [
            ctypes.POINTER(mgb_column), ctypes.c_int32, ctypes.c_int32,
            ctypes.POINTER(mgb_request), ctypes.c_int32,
            ctypes.POINTER(mgb_out_column), ctypes.POINTER(mgb_out_column),
            ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p,
        ],
    ),
    "mgb_estimate_groups": (
        ctypes.c_int,
        [ctypes.POINTER(mgb_column), ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(mgb_allocator),
         ctypes.POINTER(ctypes.c_int64)],
    ),
    "mgb_murmurhash3_x86_32": (
        ctypes.c_int,
        [ctypes.POINTER(mgb_column), ctypes.c_int32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "mgb_spark_murmurhash3_x86_32": (
        ctypes.c_int,
        [ctypes.POINTER(mgb_column), ctypes.c_int32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "mgb_profiling_enable": (ctypes.c_int, [ctypes.c_int32]),
    "mgb_profiling_read": (
        ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)],
    ),
    "mgb_release_workspace": (ctypes.c_int, [ctypes.c_void_p]),
    "mgb_pool_create": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]),
    "mgb_pool_allocator": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(mgb_allocator)]),
    "mgb_pool_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(mgb_pool_stats)]),
    "mgb_pool_trim": (ctypes.c_int, [ctypes.c_void_p]),
    "mgb_pool_destroy": (None, [ctypes.c_void_p]),
    "mgb_batch_create": (
        ctypes.c_int,
        [
            ctypes.POINTER(mgb_column), ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(mgb_request), ctypes.c_int32,
            ctypes.c_int32, ctypes.POINTER(mgb_batch_options), ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
        ],
    ),
    "mgb_batch_execute": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "mgb_batch_finish": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "mgb_batch_result": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(mgb_column), ctypes.POINTER(mgb_column),
         ctypes.POINTER(ctypes.c_int64)],
    ),
    "mgb_batch_copy_result": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(mgb_out_column), ctypes.POINTER(mgb_out_column),
         ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.POINTER(mgb_allocator)],
    ),
    "mgb_batch_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(mgb_batch_stats)]),
    "mgb_batch_query_mode": (
        ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)],
    ),
    "mgb_batch_destroy": (None, [ctypes.c_void_p]),
    "mgb_server_create": (ctypes.c_int, [ctypes.POINTER(mgb_server_options), ctypes.POINTER(ctypes.c_void_p)]),
    "mgb_server_add_table": (
        ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(mgb_column), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
    ),
    "mgb_server_submit": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, ctypes.c_int32,
         ctypes.POINTER(mgb_server_request), ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)],
    ),
    "mgb_server_wait": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_int64),
         ctypes.POINTER(ctypes.c_double)],
    ),
    "mgb_server_take_result": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(mgb_out_column), ctypes.POINTER(mgb_out_column),
         ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.POINTER(mgb_allocator)],
    ),
    "mgb_server_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "mgb_server_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(mgb_server_stats)]),
    "mgb_server_destroy": (None, [ctypes.c_void_p]),
}

_lib = None
_lock = threading.Lock()


class LibraryNotBuilt(RuntimeError):
    pass


def load():
    """Returns the loaded library (raises LibraryNotBuilt if it was never built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise LibraryNotBuilt(
                    f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                    "(make -C cudf_amd/csrc); there is no fallback implementation"
                )
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
        return _lib


class CudfError(Exception):
    """Base of the errors raised from library status codes."""


class LogicError(CudfError, ValueError):
    """cudf::logic_error: invalid argument or type/aggregation pair."""


class HipError(CudfError, RuntimeError):
    """cudf::cuda_error: a HIP runtime call failed."""


class AllocError(CudfError, MemoryError):
    """rmm::bad_alloc: an allocation failed."""


def check(status: int) -> None:
    if status == MGB_SUCCESS:
        return
    msg = load().mgb_last_error_message().decode()
    if status == MGB_ERROR_HIP:
        raise HipError(msg)
    if status == MGB_ERROR_ALLOC:
        raise AllocError(msg)
    raise LogicError(msg)
