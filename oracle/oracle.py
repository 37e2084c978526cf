"""numpy front end of the CPU oracle (``hg_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``, always to check results (or as the
timed CPU port of the reference), never on the product path.

Each wrapper names the reference file:line its C function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhgoracle.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.hgo_murmur3_x86_32.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32]
        L.hgo_murmur3_x86_32.restype = ctypes.c_uint32
        L.hgo_hash_key.argtypes = [ctypes.c_uint64]
        L.hgo_hash_key.restype = ctypes.c_uint32
        L.hgo_bins.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64, _vp]
        L.hgo_build_csr.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64, _vp, _vp]
        L.hgo_build_lrb.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.c_int, _vp, _vp]
        L.hgo_canon.argtypes = [_vp, _vp, ctypes.c_uint64, ctypes.c_int]
        L.hgo_sort_keys.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int]
        L.hgo_intersect.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_uint64, ctypes.c_int, _vp, ctypes.c_uint64]
        L.hgo_intersect.restype = ctypes.c_uint64
        L.hgo_intersect_par.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_uint64, ctypes.c_int, _vp, ctypes.c_uint64,
                                        ctypes.c_int]
        L.hgo_intersect_par.restype = ctypes.c_uint64
        L.hgo_intersect_rle.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_uint64, ctypes.c_int, _vp, _vp, _vp,
                                        ctypes.c_uint64, _vp]
        L.hgo_intersect_rle.restype = ctypes.c_uint64
        L.hgo_bag_intersect_sorted.argtypes = [_vp, ctypes.c_uint64, _vp, ctypes.c_uint64, ctypes.c_int]
        L.hgo_bag_intersect_sorted.restype = ctypes.c_uint64
        L.hgo_find_splits.argtypes = [_vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                      _vp, _vp]
        L.hgo_partition.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64, _vp, ctypes.c_uint64,
                                    _vp, _vp]
        L.hgo_build_multi.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64, _vp, ctypes.c_uint64,
                                      _vp, _vp]
        L.hgo_build_csr_bins.argtypes = [_vp, _vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64, _vp, _vp]
        L.hgo_rmi_leaf_of.argtypes = [_vp, ctypes.c_uint64]
        L.hgo_rmi_leaf_of.restype = ctypes.c_uint32
        L.hgo_rmi_bin.argtypes = [_vp, ctypes.c_uint64, ctypes.c_uint64]
        L.hgo_rmi_bin.restype = ctypes.c_uint32
        L.hgo_bins_rmi.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int, _vp, ctypes.c_uint64, _vp]
        L.hgo_splitmix.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.hgo_splitmix.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _kb(keys: np.ndarray) -> int:
    if keys.dtype == np.uint32:
        return 4
    if keys.dtype == np.uint64:
        return 8
    raise TypeError(f"keys must be uint32 or uint64, got {keys.dtype}")


def murmur3_x86_32(data: bytes, seed: int = 0) -> int:
    """MurmurHash3_x86_32 of a byte string (general form of
    include/SingleHashGraphOperators.cuh:31-69)."""
    return int(lib().hgo_murmur3_x86_32(data, len(data), seed & 0xFFFFFFFF))


def hash_key(key: int) -> int:
    """murmur3 of the low 32 bits, seed 0 (include/MultiHashGraphDeviceOperators.cuh:32-73)."""
    return int(lib().hgo_hash_key(key & 0xFFFFFFFFFFFFFFFF))


def bins(keys: np.ndarray, M: int) -> np.ndarray:
    """bin = murmur(low32(k)) % M (include/SingleHashGraphOperators.cuh:71-79)."""
    keys = np.ascontiguousarray(keys)
    out = np.empty(len(keys), dtype=np.uint32)
    lib().hgo_bins(_ptr(keys), len(keys), _kb(keys), M, _ptr(out))
    return out


def build_csr(keys: np.ndarray, M: int):
    """Canonical CSR (counting sort, stable in a bin); offsets uint64[M+1]
    (include/MultiHashGraphHostOperators.cuh:316-343)."""
    keys = np.ascontiguousarray(keys)
    off = np.empty(M + 1, dtype=np.uint64)
    out = np.empty_like(keys)
    rc = lib().hgo_build_csr(_ptr(keys), len(keys), _kb(keys), M, _ptr(off), _ptr(out))
    if rc:
        raise ValueError(f"hgo_build_csr failed ({rc})")
    return off, out


def build_lrb(keys: np.ndarray, M: int, lrb_bins: int = 16000, threads: int = 0):
    """The reference's LRB data flow with CPU threads (src/SingleHashGraph.cu:129-232)."""
    keys = np.ascontiguousarray(keys)
    off = np.empty(M + 1, dtype=np.uint64)
    out = np.empty_like(keys)
    rc = lib().hgo_build_lrb(_ptr(keys), len(keys), _kb(keys), M, lrb_bins, threads, _ptr(off), _ptr(out))
    if rc:
        raise ValueError(f"hgo_build_lrb failed ({rc})")
    return off, out


def build_csr_bins(keys: np.ndarray, bins_arr: np.ndarray, M: int):
    """Canonical CSR from given bins (any hash); offsets uint64[M+1]."""
    keys = np.ascontiguousarray(keys)
    bins_arr = np.ascontiguousarray(bins_arr, dtype=np.uint32)
    off = np.empty(M + 1, dtype=np.uint64)
    out = np.empty_like(keys)
    rc = lib().hgo_build_csr_bins(_ptr(keys), _ptr(bins_arr), len(keys), _kb(keys), M, _ptr(off), _ptr(out))
    if rc:
        raise ValueError(f"hgo_build_csr_bins failed ({rc})")
    return off, out


def rmi_leaf(model: np.ndarray, key: int) -> int:
    """Root stage of the learned model (model = its bytes as uint8)."""
    model = np.ascontiguousarray(model, dtype=np.uint8)
    return int(lib().hgo_rmi_leaf_of(_ptr(model), key & 0xFFFFFFFFFFFFFFFF))


def rmi_bin(model: np.ndarray, key: int, M: int) -> int:
    model = np.ascontiguousarray(model, dtype=np.uint8)
    return int(lib().hgo_rmi_bin(_ptr(model), key & 0xFFFFFFFFFFFFFFFF, M))


def bins_rmi(keys: np.ndarray, model: np.ndarray, M: int) -> np.ndarray:
    """Bins of keys under the learned model (model = its bytes as uint8)."""
    keys = np.ascontiguousarray(keys)
    model = np.ascontiguousarray(model, dtype=np.uint8)
    out = np.empty(len(keys), dtype=np.uint32)
    lib().hgo_bins_rmi(_ptr(keys), len(keys), _kb(keys), _ptr(model), M, _ptr(out))
    return out


def rmi_model(key_min: int, root_slope: float, root_icpt: float, leaves, M: int, root_kind: int = 0) -> np.ndarray:
    """Bytes of a learned model: leaves = [(x0, slope, anchor), ...]."""
    L = len(leaves)
    hdr = np.zeros(1, dtype=[("key_min", "<u8"), ("rs", "<f4"), ("ri", "<f4"), ("L", "<u4"), ("M", "<u4"),
                             ("kind", "<u4"), ("pad", "<u4")])
    hdr["key_min"] = key_min
    hdr["rs"] = root_slope
    hdr["ri"] = root_icpt
    hdr["L"] = L
    hdr["M"] = M
    hdr["kind"] = root_kind
    lv = np.zeros(L, dtype=[("x0", "<u8"), ("slope", "<f4"), ("anchor", "<u4")])
    for i, (x0, sl, b) in enumerate(leaves):
        lv[i] = (x0, sl, b)
    return np.concatenate([hdr.view(np.uint8), lv.view(np.uint8)])


def canon(off: np.ndarray, keys: np.ndarray) -> np.ndarray:
    """Sort the keys inside every bin in place (src/MultiHashGraph.cu:826-841)."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    assert keys.flags.c_contiguous
    lib().hgo_canon(_ptr(off), _ptr(keys), len(off) - 1, _kb(keys))
    return keys


def intersect(offA, keysA, offB, keysB, emit: bool = True):
    """Bin-wise bag-product join (src/MultiHashGraph.cu:631-740,
    include/MultiHashGraphDeviceOperators.cuh:513-543). Returns (count, list|None)."""
    offA = np.ascontiguousarray(offA, dtype=np.uint64)
    offB = np.ascontiguousarray(offB, dtype=np.uint64)
    kb = _kb(keysA)
    assert keysB.dtype == keysA.dtype and len(offA) == len(offB)
    M = len(offA) - 1
    cnt = int(lib().hgo_intersect(_ptr(offA), _ptr(keysA), _ptr(offB), _ptr(keysB), M, kb, None, 0))
    if not emit:
        return cnt, None
    out = np.empty(cnt, dtype=keysA.dtype)
    lib().hgo_intersect(_ptr(offA), _ptr(keysA), _ptr(offB), _ptr(keysB), M, kb, _ptr(out), cnt)
    return cnt, out


def intersect_par(offA, keysA, offB, keysB, threads: int, out=None):
    """Threaded bin join (CPU baseline); same count and list as intersect()."""
    offA = np.ascontiguousarray(offA, dtype=np.uint64)
    offB = np.ascontiguousarray(offB, dtype=np.uint64)
    M = len(offA) - 1
    return int(lib().hgo_intersect_par(_ptr(offA), _ptr(keysA), _ptr(offB), _ptr(keysB), M, _kb(keysA),
                                       _ptr(out) if out is not None else None, len(out) if out is not None else 0,
                                       threads))


def intersect_rle(offA, keysA, offB, keysB):
    """Run-length join over canonical bins: records (key, cA, cB); count = sum cA*cB."""
    offA = np.ascontiguousarray(offA, dtype=np.uint64)
    offB = np.ascontiguousarray(offB, dtype=np.uint64)
    kb = _kb(keysA)
    M = len(offA) - 1
    n = ctypes.c_uint64(0)
    lib().hgo_intersect_rle(_ptr(offA), _ptr(keysA), _ptr(offB), _ptr(keysB), M, kb, None, None, None, 0,
                            ctypes.byref(n))
    k = np.empty(n.value, dtype=np.uint64)
    ca = np.empty(n.value, dtype=np.uint64)
    cb = np.empty(n.value, dtype=np.uint64)
    cnt = int(lib().hgo_intersect_rle(_ptr(offA), _ptr(keysA), _ptr(offB), _ptr(keysB), M, kb, _ptr(k), _ptr(ca),
                                      _ptr(cb), n.value, ctypes.byref(n)))
    return cnt, k, ca, cb


def bag_intersect_count(A: np.ndarray, B: np.ndarray) -> int:
    """Hash-free answer of the reference driver (test/MultiHashGraphTest.cu:369-397)."""
    A = np.array(A, copy=True)
    B = np.array(B, copy=True)
    return int(lib().hgo_bag_intersect_sorted(_ptr(A), len(A), _ptr(B), len(B), _kb(A)))


def find_splits(bins_arr: np.ndarray, table_size: int, bin_count: int = 16000, gpu_count: int = 1):
    """Key-balanced split points (include/MultiHashGraphHostOperators.cuh:8-108)."""
    bins_arr = np.ascontiguousarray(bins_arr, dtype=np.uint32)
    splits = np.zeros(gpu_count + 1, dtype=np.int64)
    ps = np.zeros(bin_count + 1, dtype=np.int64)
    rc = lib().hgo_find_splits(_ptr(bins_arr), len(bins_arr), table_size, bin_count, gpu_count, _ptr(splits),
                               _ptr(ps))
    if rc:
        raise ValueError("hgo_find_splits: invalid arguments")
    return splits, ps


def partition(keys: np.ndarray, M: int, splits: np.ndarray):
    """Local partition by split range (include/MultiHashGraphDeviceOperators.cuh:297-345)."""
    keys = np.ascontiguousarray(keys)
    splits = np.ascontiguousarray(splits, dtype=np.int64)
    P = len(splits) - 1
    out = np.empty_like(keys)
    po = np.empty(P + 1, dtype=np.uint64)
    rc = lib().hgo_partition(_ptr(keys), len(keys), _kb(keys), M, _ptr(splits), P, _ptr(out), _ptr(po))
    if rc:
        raise ValueError("hgo_partition failed")
    return out[: int(po[P])], po


def build_multi(keys: np.ndarray, M: int, splits: np.ndarray):
    """Per-partition CSRs (src/MultiHashGraph.cu:408-629). Offsets of partition p
    live at off[splits[p]+p : splits[p+1]+p+1]."""
    keys = np.ascontiguousarray(keys)
    splits = np.ascontiguousarray(splits, dtype=np.int64)
    P = len(splits) - 1
    off = np.zeros(int(splits[P]) + P, dtype=np.uint64)
    out = np.empty_like(keys)
    rc = lib().hgo_build_multi(_ptr(keys), len(keys), _kb(keys), M, _ptr(splits), P, _ptr(off), _ptr(out))
    if rc:
        raise ValueError("hgo_build_multi failed")
    return off, out


def splitmix(seed: int, i: int) -> int:
    return int(lib().hgo_splitmix(seed, i))
