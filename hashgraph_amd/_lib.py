"""ctypes binding of libhashgraph_amd.so (the C ABI of include/hashgraph.h).

Device buffers are torch tensors on a ROCm device; their data_ptr() is
passed through the C ABI.  There is no fallback: if the library is missing
or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhashgraph_amd.so")

HG_OK = 0
STATUS = {0: "HG_OK", 1: "HG_ERR_INVALID_ARG", 2: "HG_ERR_OOM", 3: "HG_ERR_DEVICE", 4: "HG_ERR_CAPACITY",
          5: "HG_ERR_MISMATCH", 6: "HG_ERR_UNSUPPORTED"}
HG_HASH_MURMUR3 = 0
HG_HASH_RMI = 1

# Every symbol include/hashgraph.h declares (checked by the CPU tests).
EXPORTS = [
    "hg_ctx_create", "hg_ctx_destroy", "hg_ctx_set_stream", "hg_ctx_reserve", "hg_scratch_bytes", "hg_last_error",
    "hg_version", "hg_ctx_profile", "hg_ctx_profile_read", "hg_build", "hg_intersect", "hg_probe", "hg_intersect_rle", "hg_find_splits",
    "hg_build_partitioned", "hg_rmi_model_bytes", "hg_rmi_train", "hg_rmi_train_batch", "hg_batch_build", "hg_batch_probe", "hg_batch_probe_emit",
    "hg_exclusive_scan", "hg_hash_bins", "hg_table_bins", "hg_batch_plan_create", "hg_batch_plan_run",
    "hg_batch_plan_query", "hg_batch_plan_destroy", "hg_build_noalloc", "hg_build_indexed",
    "hg_build_partitioned_indexed", "hg_build_staged", "hg_build_partitioned_staged", "hg_find_splits_staged",
    "hg_intersect_staged", "hg_intersect_indexed",
]


class HGError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS.get(status, status)}: {msg}")
        self.status = status


class TableDesc(ctypes.Structure):
    _fields_ = [("key_bytes", ctypes.c_uint32), ("offset_bytes", ctypes.c_uint32), ("n_keys", ctypes.c_uint64),
                ("n_bins", ctypes.c_uint64), ("hash_kind", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("d_rmi_model", ctypes.c_void_p)]


class RmiReport(ctypes.Structure):
    _fields_ = [("sample_bins", ctypes.c_uint64), ("max_bin_rmi", ctypes.c_uint64),
                ("max_bin_murmur", ctypes.c_uint64), ("mean_probe_rmi", ctypes.c_double),
                ("mean_probe_murmur", ctypes.c_double)]


HG_PLAN_AUTOTUNE = 1
HG_PLAN_TRAIN = 2


class BatchPlanInfo(ctypes.Structure):
    _fields_ = [("tile_keys", ctypes.c_uint32), ("fill", ctypes.c_float), ("candidates", ctypes.c_uint32),
                ("tuned_ms", ctypes.c_float), ("learned_tables", ctypes.c_uint32)]


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double)]


class BatchDesc(ctypes.Structure):
    _fields_ = [("n_tables", ctypes.c_uint32), ("key_bytes", ctypes.c_uint32),
                ("h_key_off", ctypes.POINTER(ctypes.c_uint64)), ("h_n_bins", ctypes.POINTER(ctypes.c_uint64)),
                ("h_off_off", ctypes.POINTER(ctypes.c_uint64)), ("h_hash_kind", ctypes.POINTER(ctypes.c_uint32)),
                ("d_rmi_models", ctypes.c_void_p), ("rmi_leaves", ctypes.c_uint32)]


_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH):
    """Load the shared library (raises if it is missing: no CPU fallback)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(f"libhashgraph_amd.so not built ({path}); run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        P = ctypes.POINTER
        sig = {
            "hg_ctx_create": (i32, [i32, P(vp)]),
            "hg_ctx_destroy": (i32, [vp]),
            "hg_ctx_set_stream": (i32, [vp, vp]),
            "hg_ctx_reserve": (i32, [vp, ctypes.c_size_t]),
            "hg_scratch_bytes": (i32, [P(TableDesc), i32, P(ctypes.c_size_t)]),
            "hg_last_error": (ctypes.c_char_p, []),
            "hg_version": (ctypes.c_char_p, []),
            "hg_ctx_profile": (i32, [vp, i32]),
            "hg_ctx_profile_read": (i32, [vp, P(KernelTime), u32, P(u32), i32]),
            "hg_build": (i32, [vp, P(TableDesc), vp, vp, vp]),
            "hg_build_noalloc": (i32, [vp, P(TableDesc), vp, vp, vp]),
            "hg_intersect": (i32, [vp, P(TableDesc), vp, vp, P(TableDesc), vp, vp, vp, u64, P(u64)]),
            "hg_probe": (i32, [vp, P(TableDesc), vp, vp, vp, u64, vp, u64, P(u64)]),
            "hg_intersect_rle": (i32, [vp, P(TableDesc), vp, vp, P(TableDesc), vp, vp, vp, vp, vp, u64, P(u64),
                                       P(u64)]),
            "hg_find_splits": (i32, [vp, P(TableDesc), vp, u64, u32, P(ctypes.c_int64)]),
            "hg_build_partitioned": (i32, [vp, P(TableDesc), vp, P(ctypes.c_int64), u32, vp, vp, P(u64)]),
            "hg_build_indexed": (i32, [vp, P(TableDesc), vp, vp, vp, vp]),
            "hg_intersect_indexed": (i32, [vp, P(TableDesc), vp, vp, vp, P(TableDesc), vp, vp, vp, vp, vp, u64,
                                           P(u64)]),
            "hg_build_partitioned_indexed": (i32, [vp, P(TableDesc), vp, P(ctypes.c_int64), u32, vp, vp, vp, P(u64)]),
            "hg_build_staged": (i32, [vp, P(TableDesc), vp, vp, vp, ctypes.c_size_t]),
            "hg_build_partitioned_staged": (i32, [vp, P(TableDesc), vp, P(ctypes.c_int64), u32, vp, vp, P(u64),
                                                  ctypes.c_size_t]),
            "hg_find_splits_staged": (i32, [vp, P(TableDesc), vp, u64, u32, P(ctypes.c_int64), ctypes.c_size_t]),
            "hg_intersect_staged": (i32, [vp, P(TableDesc), vp, vp, P(TableDesc), vp, vp, vp, u64, P(u64),
                                          ctypes.c_size_t]),
            "hg_rmi_model_bytes": (ctypes.c_size_t, [u32]),
            "hg_rmi_train": (i32, [vp, vp, u64, u32, u64, u64, u64, u32, vp, P(u32), P(RmiReport)]),
            "hg_rmi_train_batch": (i32, [vp, u32, u32, vp, P(u64), P(u64), P(u64), P(u64), u32, vp, P(u32),
                                         P(RmiReport)]),
            "hg_batch_build": (i32, [vp, P(BatchDesc), vp, vp, vp]),
            "hg_batch_probe": (i32, [vp, P(BatchDesc), vp, vp, vp, P(u64), vp, P(u64)]),
            "hg_batch_probe_emit": (i32, [vp, P(BatchDesc), vp, vp, vp, P(u64), P(u64), vp, vp, P(u64)]),
            "hg_batch_plan_create": (i32, [vp, P(BatchDesc), vp, vp, vp, vp, P(u64), vp, u32, P(vp)]),
            "hg_batch_plan_run": (i32, [vp]),
            "hg_batch_plan_query": (i32, [vp, P(BatchPlanInfo)]),
            "hg_batch_plan_destroy": (i32, [vp]),
            "hg_exclusive_scan": (i32, [vp, vp, vp, u64, u32]),
            "hg_hash_bins": (i32, [vp, vp, u64, u32, u64, vp]),
            "hg_table_bins": (i32, [vp, P(TableDesc), vp, u64, vp]),
        }
        for name, (res, args) in sig.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _lib = L
        return L


def exported_symbols(path: str = LIB_PATH):
    """Names of EXPORTS that the library actually exports."""
    L = ctypes.CDLL(path)
    return [n for n in EXPORTS if hasattr(L, n)]


def check(status: int, where: str):
    if status != HG_OK:
        msg = load().hg_last_error()
        raise HGError(status, where, msg.decode() if msg else "")
