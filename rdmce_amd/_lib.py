"""ctypes binding of librdmce.so (the C ABI declared in include/rdmce.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
rdmce_amd/csrc``).  Nothing here falls back to another implementation: if the
library is missing, importing the solver raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_int32, c_uint32,
                    c_uint64, c_void_p)

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# RDMCE_LIB names another build of the library (e.g. to compare two builds).
LIB_PATH = os.environ.get("RDMCE_LIB") or os.path.join(PKG_DIR, "librdmce.so")

MCE_OK = 0
MCE_ERR_ARG = 1
MCE_ERR_IO = 2
MCE_ERR_NOMEM = 3
MCE_ERR_DEVICE = 4
MCE_ERR_UNSUPPORTED = 5
MCE_ERR_INTERNAL = 6

ORDER_UNCHANGED, ORDER_ASC, ORDER_DES, ORDER_DEG = 0, 1, 2, 3
ROOT_VERTEX, ROOT_EDGE = 1, 2
SCHED_PLAIN, SCHED_CUSTOM = 0, 1


class MceOptions(Structure):
    _fields_ = [
        ("device", c_int32),
        ("root_mode", c_int32),
        ("waves_per_cu", c_int32),
        ("share_bound", c_int32),
        ("want_list", c_int32),
        ("want_root_stats", c_int32),
        ("schedule", c_int32),
        ("lds_tier", c_int32),
        ("list_capacity", c_uint64),
        ("slot_bytes", c_uint64),
        ("roots", POINTER(c_uint64)),
        ("n_roots", c_uint64),
        ("use_graph", c_int32),
        ("lds_bytes", c_uint32),
        ("eager", POINTER(ctypes.c_uint8)),
        ("word_path", c_int32),
        ("pack", c_int32),
        ("pack_slots", c_uint32),
        ("pack_waves_per_cu", c_uint32),
        ("runs", POINTER(c_uint64)),
        ("n_runs", c_uint64),
        ("reserved", c_uint64 * 2),
    ]


class MceBatchResult(Structure):
    _fields_ = [
        ("n_graphs", c_uint64),
        ("hist_stride", c_uint32),
        ("count", POINTER(c_uint64)),
        ("hash", POINTER(c_uint64)),
        ("hist", POINTER(c_uint64)),
        ("total_count", c_uint64),
        ("total_nodes", c_uint64),
        ("kernel_ms", c_double),
        ("total_ms", c_double),
        ("graph_replay", c_int32),
        ("passes", c_int32),
        ("reserved", c_uint64 * 6),
    ]


class MceTuneResult(Structure):
    _fields_ = [
        ("waves_per_cu", c_int32),
        ("share_bound", c_int32),
        ("lds_tier", c_int32),
        ("trials", c_int32),
        ("best_ms", c_double),
        ("base_ms", c_double),
        ("pack_slots", c_uint32),
        ("pack_waves_per_cu", c_uint32),
        ("reserved", c_uint64 * 3),
    ]


class MceResult(Structure):
    _fields_ = [
        ("count", c_uint64),
        ("hash", c_uint64),
        ("hist_len", c_uint32),
        ("hist", POINTER(c_uint64)),
        ("list_len", c_uint64),
        ("list", POINTER(c_uint32)),
        ("n_roots", c_uint64),
        ("root_nodes", POINTER(c_uint64)),
        ("root_cycles", POINTER(c_uint64)),
        ("total_nodes", c_uint64),
        ("deferred_roots", c_uint64),
        ("shared_tasks", c_uint64),
        ("kernel_ms", c_double),
        ("total_ms", c_double),
        ("passes", c_int32),
        ("waves", c_int32),
        ("search_bytes", c_uint64),
        ("build_bytes", c_uint64),
        ("published_levels", c_uint64),
        ("owner_tasks", c_uint64),
        ("reserved", c_uint64 * 4),
    ]


class MceGraphInfo(Structure):
    _fields_ = [("n", c_uint64), ("m", c_uint64), ("max_degree", c_uint64), ("degeneracy", c_uint64)]


class MceError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"rdmce error {status}: {message}")
        self.status = status


_lib = None


def lib() -> ctypes.CDLL:
    """Loads librdmce.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not found: build it with __graft_entry__.build() "
                           "or `make -C rdmce_amd/csrc`")
    L = ctypes.CDLL(LIB_PATH)
    G = c_void_p
    PG = POINTER(c_void_p)
    sigs = {
        "mce_last_error": ([], c_char_p),
        "mce_version": ([], c_char_p),
        "mce_graph_load_coo_text": ([c_char_p, PG], c_int32),
        "mce_graph_load_bin": ([c_char_p, PG], c_int32),
        "mce_graph_store_bin": ([G, c_char_p], c_int32),
        "mce_graph_bin_path": ([c_char_p, ctypes.c_char_p, c_uint64], c_int32),
        "mce_graph_from_pairs": ([POINTER(c_uint32), POINTER(c_uint32), c_uint64, PG], c_int32),
        "mce_graph_from_csr": ([c_uint64, POINTER(c_uint64), POINTER(c_uint32), POINTER(c_uint32), c_uint64, PG],
                               c_int32),
        "mce_graph_order": ([G, c_int32], c_int32),
        "mce_graph_get_info": ([G, POINTER(MceGraphInfo)], c_int32),
        "mce_graph_get_csr": ([G, POINTER(POINTER(c_uint64)), POINTER(POINTER(c_uint32)),
                               POINTER(POINTER(c_uint32))], c_int32),
        "mce_graph_get_core": ([G, POINTER(POINTER(c_uint32))], c_int32),
        "mce_graph_write_coo_text": ([G, c_char_p], c_int32),
        "mce_graph_free": ([G], None),
        "mce_gen_er": ([c_uint32, c_double, c_uint64, PG], c_int32),
        "mce_gen_rmat": ([c_int32, c_uint64, c_double, c_double, c_double, c_uint64, PG], c_int32),
        "mce_gen_chung_lu_planted": ([c_uint64, c_double, c_double, c_uint32, c_uint32, c_uint32, c_double,
                                      c_uint64, PG], c_int32),
        "mce_gen_batch_params": ([c_uint64, c_uint64, c_uint32, c_uint32, c_double, c_double, POINTER(c_uint32),
                                  POINTER(c_double), POINTER(c_uint64)], c_int32),
        "mce_options_init": ([POINTER(MceOptions)], None),
        "mce_solve": ([G, POINTER(MceOptions), POINTER(POINTER(MceResult))], c_int32),
        "mce_solve_csr": ([c_uint64, POINTER(c_uint64), POINTER(c_uint32), POINTER(c_uint32), c_uint64,
                           POINTER(MceOptions), POINTER(POINTER(MceResult))], c_int32),
        "mce_result_free": ([POINTER(MceResult)], None),
        "mce_plan_create": ([G, POINTER(MceOptions), PG], c_int32),
        "mce_plan_run": ([c_void_p, POINTER(POINTER(MceResult))], c_int32),
        "mce_plan_roots": ([G, c_int32, POINTER(c_uint64), c_uint64, POINTER(c_uint64)], c_int32),
        "mce_plan_free": ([c_void_p], None),
        "mce_device_count": ([], c_int32),
        "mce_device_sync": ([c_int32], c_int32),
        "mce_batch_create": ([POINTER(c_void_p), c_uint64, POINTER(MceOptions), PG], c_int32),
        "mce_batch_run": ([c_void_p, POINTER(POINTER(MceBatchResult))], c_int32),
        "mce_batch_result_free": ([POINTER(MceBatchResult)], None),
        "mce_batch_free": ([c_void_p], None),
        "mce_plan_autotune": ([c_void_p, c_uint64, POINTER(MceTuneResult)], c_int32),
        "mce_batch_autotune": ([c_void_p, c_uint64, POINTER(MceTuneResult)], c_int32),
        "mce_copy_probe": ([c_int32, c_uint64, c_int32, POINTER(c_double)], c_int32),
    }
    for name, (args, ret) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ret
    _lib = L
    return L


def check(status: int) -> None:
    if status != MCE_OK:
        msg = lib().mce_last_error()
        raise MceError(status, msg.decode() if msg else "")
