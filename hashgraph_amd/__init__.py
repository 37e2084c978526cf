"""MI355X-native HashGraph: hash tables stored as CSR graphs, built and
probed by hand-written gfx950 kernels (libhashgraph_amd.so).

This module is a thin torch front end over the C ABI (include/hashgraph.h)
for tests and benchmarks: torch tensors provide device memory and the stream;
all work happens in the library.  The C++ drop-in classes SingleHashGraph /
MultiHashGraph (include/SingleHashGraph.hpp, include/MultiHashGraph.hpp) are
the reference-compatible host API.

Keys are uint32 or uint64 carried in int32 / int64 tensors (same bits).
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import HG_HASH_MURMUR3, HG_HASH_RMI, HGError, check

__all__ = ["Context", "build", "probe", "intersect", "intersect_rle", "exclusive_scan", "hash_bins", "find_splits",
           "build_partitioned", "rmi_train", "rmi_train_batch", "rmi_model_bytes", "batch_build", "batch_probe", "batch_probe_emit", "HGError",
           "HG_HASH_MURMUR3", "HG_HASH_RMI", "version", "table_bins", "BatchPlan", "build_staged",
           "build_partitioned_staged", "find_splits_staged", "intersect_staged", "build_indexed",
           "intersect_indexed"]

_ctx = {}


def _torch():
    import torch
    return torch


def version() -> str:
    return _lib.load().hg_version().decode()


class Context:
    """One hg_ctx per device (scratch arena + stream binding)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.device = device
        h = ctypes.c_void_p()
        check(self.lib.hg_ctx_create(device, ctypes.byref(h)), "hg_ctx_create")
        self.h = h

    def bind_stream(self):
        torch = _torch()
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(self.lib.hg_ctx_set_stream(self.h, ctypes.c_void_p(s)), "hg_ctx_set_stream")
        return self.h

    def profile(self, enable: bool = True):
        check(self.lib.hg_ctx_profile(self.h, 1 if enable else 0), "hg_ctx_profile")

    def profile_read(self, reset: bool = True):
        """{kernel name: (launches, total ms)} since the last reset (synchronises)."""
        buf = (_lib.KernelTime * 64)()
        n = ctypes.c_uint32(0)
        check(self.lib.hg_ctx_profile_read(self.h, buf, 64, ctypes.byref(n), 1 if reset else 0),
              "hg_ctx_profile_read")
        return {buf[i].name.decode(): (int(buf[i].launches), float(buf[i].total_ms)) for i in range(n.value)}

    def reserve(self, nbytes: int):
        check(self.lib.hg_ctx_reserve(self.h, nbytes), "hg_ctx_reserve")

    def close(self):
        if self.h:
            self.lib.hg_ctx_destroy(self.h)
            self.h = None


def context(device=None) -> Context:
    torch = _torch()
    if device is None:
        device = torch.cuda.current_device()
    if device not in _ctx:
        _ctx[device] = Context(device)
    return _ctx[device]


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _key_bytes(t) -> int:
    torch = _torch()
    if t.dtype in (torch.int32, torch.uint32):
        return 4
    if t.dtype in (torch.int64, torch.uint64):
        return 8
    raise TypeError(f"keys must be a 32- or 64-bit integer tensor, got {t.dtype}")


def _desc(n_keys, n_bins, key_bytes, offset_bytes=4, hash_kind=HG_HASH_MURMUR3, rmi_model=None):
    d = _lib.TableDesc()
    d.key_bytes = key_bytes
    d.offset_bytes = offset_bytes
    d.n_keys = n_keys
    d.n_bins = n_bins
    d.hash_kind = hash_kind
    d.d_rmi_model = rmi_model.data_ptr() if rmi_model is not None else None
    return d


def _check_dev(*ts):
    for t in ts:
        if t is not None and (not t.is_cuda or not t.is_contiguous()):
            raise ValueError("tensors must be contiguous device tensors")


def build(keys, n_bins: int, offset_bytes: int = 4, hash_kind: int = HG_HASH_MURMUR3, rmi_model=None, out=None):
    """Build the CSR of `keys` over n_bins bins.  Returns (offsets, csr_keys):
    offsets int32/int64[n_bins+1], csr_keys like keys."""
    torch = _torch()
    _check_dev(keys)
    kb = _key_bytes(keys)
    c = context(keys.device.index)
    if out is None:
        off = torch.empty(n_bins + 1, dtype=torch.int32 if offset_bytes == 4 else torch.int64, device=keys.device)
        csr = torch.empty_like(keys)
    else:
        off, csr = out
    d = _desc(keys.numel(), n_bins, kb, offset_bytes, hash_kind, rmi_model)
    check(c.lib.hg_build(c.bind_stream(), ctypes.byref(d), _p(keys), _p(off), _p(csr)), "hg_build")
    return off, csr


def probe(offsets, csr_keys, n_bins: int, probe_keys, emit: bool = True, capacity: int | None = None,
          hash_kind: int = HG_HASH_MURMUR3, rmi_model=None, out=None):
    """Probe raw keys against a built table.  Returns (count, matched keys or None)."""
    torch = _torch()
    _check_dev(offsets, csr_keys, probe_keys)
    kb = _key_bytes(csr_keys)
    ob = offsets.element_size()
    c = context(csr_keys.device.index)
    d = _desc(csr_keys.numel(), n_bins, kb, ob, hash_kind, rmi_model)
    cnt = ctypes.c_uint64(0)
    if not emit:
        check(c.lib.hg_probe(c.bind_stream(), ctypes.byref(d), _p(offsets), _p(csr_keys), _p(probe_keys),
                             probe_keys.numel(), None, 0, ctypes.byref(cnt)), "hg_probe")
        return cnt.value, None
    if out is None:
        if capacity is None:
            check(c.lib.hg_probe(c.bind_stream(), ctypes.byref(d), _p(offsets), _p(csr_keys), _p(probe_keys),
                                 probe_keys.numel(), None, 0, ctypes.byref(cnt)), "hg_probe")
            capacity = cnt.value
        out = torch.empty(max(capacity, 1), dtype=csr_keys.dtype, device=csr_keys.device)
    else:
        capacity = out.numel()
    check(c.lib.hg_probe(c.bind_stream(), ctypes.byref(d), _p(offsets), _p(csr_keys), _p(probe_keys),
                         probe_keys.numel(), _p(out), capacity, ctypes.byref(cnt)), "hg_probe")
    return cnt.value, out[: cnt.value]


def intersect(offA, keysA, offB, keysB, n_bins: int, emit: bool = True, hash_kind: int = HG_HASH_MURMUR3,
              rmi_model=None):
    """Bag-product join of two built tables (MultiHashGraph::intersect).
    Returns (count, matched keys or None)."""
    torch = _torch()
    _check_dev(offA, keysA, offB, keysB)
    kb = _key_bytes(keysA)
    ob = offA.element_size()
    c = context(keysA.device.index)
    dA = _desc(keysA.numel(), n_bins, kb, ob, hash_kind, rmi_model)
    dB = _desc(keysB.numel(), n_bins, kb, ob, hash_kind, rmi_model)
    cnt = ctypes.c_uint64(0)
    check(c.lib.hg_intersect(c.bind_stream(), ctypes.byref(dA), _p(offA), _p(keysA), ctypes.byref(dB), _p(offB),
                             _p(keysB), None, 0, ctypes.byref(cnt)), "hg_intersect")
    if not emit:
        return cnt.value, None
    out = torch.empty(max(cnt.value, 1), dtype=keysA.dtype, device=keysA.device)
    check(c.lib.hg_intersect(c.bind_stream(), ctypes.byref(dA), _p(offA), _p(keysA), ctypes.byref(dB), _p(offB),
                             _p(keysB), _p(out), cnt.value, ctypes.byref(cnt)), "hg_intersect")
    return cnt.value, out[: cnt.value]


def intersect_rle(offA, keysA, offB, keysB, n_bins: int, hash_kind: int = HG_HASH_MURMUR3, rmi_model=None,
                  capacity: int | None = None):
    """Run-length join: returns (count, keys, countA, countB) of every value in both tables."""
    torch = _torch()
    _check_dev(offA, keysA, offB, keysB)
    kb = _key_bytes(keysA)
    ob = offA.element_size()
    c = context(keysA.device.index)
    dA = _desc(keysA.numel(), n_bins, kb, ob, hash_kind, rmi_model)
    dB = _desc(keysB.numel(), n_bins, kb, ob, hash_kind, rmi_model)
    cnt, nrec = ctypes.c_uint64(0), ctypes.c_uint64(0)
    if capacity is None:
        check(c.lib.hg_intersect_rle(c.bind_stream(), ctypes.byref(dA), _p(offA), _p(keysA), ctypes.byref(dB),
                                     _p(offB), _p(keysB), None, None, None, 0, ctypes.byref(nrec), ctypes.byref(cnt)),
              "hg_intersect_rle")
        capacity = nrec.value
    rk = torch.empty(max(capacity, 1), dtype=keysA.dtype, device=keysA.device)
    ca = torch.empty(max(capacity, 1), dtype=torch.int64, device=keysA.device)
    cb = torch.empty(max(capacity, 1), dtype=torch.int64, device=keysA.device)
    check(c.lib.hg_intersect_rle(c.bind_stream(), ctypes.byref(dA), _p(offA), _p(keysA), ctypes.byref(dB), _p(offB),
                                 _p(keysB), _p(rk), _p(ca), _p(cb), capacity, ctypes.byref(nrec), ctypes.byref(cnt)),
          "hg_intersect_rle")
    n = nrec.value
    return cnt.value, rk[:n], ca[:n], cb[:n]


def build_indexed(keys, n_bins: int, offset_bytes: int = 4):
    """hg_build plus every CSR key's input position (the INDEX_TRACK entries).
    Returns (offsets, csr_keys, csr_index int64)."""
    torch = _torch()
    _check_dev(keys)
    c = context(keys.device.index)
    off = torch.empty(n_bins + 1, dtype=torch.int32 if offset_bytes == 4 else torch.int64, device=keys.device)
    csr = torch.empty_like(keys)
    ind = torch.empty(max(keys.numel(), 1), dtype=torch.int64, device=keys.device)
    d = _desc(keys.numel(), n_bins, _key_bytes(keys), offset_bytes)
    check(c.lib.hg_build_indexed(c.bind_stream(), ctypes.byref(d), _p(keys), _p(off), _p(csr), _p(ind)),
          "hg_build_indexed")
    return off, csr, ind[: keys.numel()]


def intersect_indexed(offA, keysA, indA, offB, keysB, indB, n_bins: int, emit: bool = True,
                      capacity: int | None = None):
    """Row-id join of two index-carrying tables: for every matching pair, A's
    and B's input positions.  Returns (count, pos_a, pos_b), the positions
    None for count only."""
    torch = _torch()
    _check_dev(offA, keysA, indA, offB, keysB, indB)
    kb = _key_bytes(keysA)
    ob = offA.element_size()
    c = context(keysA.device.index)
    dA = _desc(keysA.numel(), n_bins, kb, ob)
    dB = _desc(keysB.numel(), n_bins, kb, ob)
    cnt = ctypes.c_uint64(0)
    if capacity is None:
        check(c.lib.hg_intersect_indexed(c.bind_stream(), ctypes.byref(dA), _p(offA), _p(keysA), _p(indA),
                                         ctypes.byref(dB), _p(offB), _p(keysB), _p(indB), None, None, 0,
                                         ctypes.byref(cnt)), "hg_intersect_indexed")
        if not emit:
            return cnt.value, None, None
        capacity = cnt.value
    pa = torch.empty(max(capacity, 1), dtype=torch.int64, device=keysA.device)
    pb = torch.empty(max(capacity, 1), dtype=torch.int64, device=keysA.device)
    check(c.lib.hg_intersect_indexed(c.bind_stream(), ctypes.byref(dA), _p(offA), _p(keysA), _p(indA),
                                     ctypes.byref(dB), _p(offB), _p(keysB), _p(indB), _p(pa), _p(pb), capacity,
                                     ctypes.byref(cnt)), "hg_intersect_indexed")
    return cnt.value, pa[: cnt.value], pb[: cnt.value]


def exclusive_scan(x):
    """Exclusive prefix sum of a 32/64-bit integer tensor; returns n+1 values (total last)."""
    torch = _torch()
    _check_dev(x)
    c = context(x.device.index)
    out = torch.empty(x.numel() + 1, dtype=x.dtype, device=x.device)
    check(c.lib.hg_exclusive_scan(c.bind_stream(), _p(x), _p(out), x.numel(), x.element_size()),
          "hg_exclusive_scan")
    return out


def hash_bins(keys, n_bins: int):
    """murmur3(low32(key)) mod n_bins for every key (int32 tensor holding uint32)."""
    torch = _torch()
    _check_dev(keys)
    c = context(keys.device.index)
    out = torch.empty(keys.numel(), dtype=torch.int32, device=keys.device)
    check(c.lib.hg_hash_bins(c.bind_stream(), _p(keys), keys.numel(), _key_bytes(keys), n_bins, _p(out)),
          "hg_hash_bins")
    return out


def find_splits(keys, n_bins: int, bin_count: int = 16000, n_parts: int = 16, hash_kind: int = HG_HASH_MURMUR3,
                rmi_model=None):
    """Key-balanced partition split points over the bin range (list of n_parts+1 ints)."""
    _check_dev(keys)
    c = context(keys.device.index)
    d = _desc(keys.numel(), n_bins, _key_bytes(keys), 4, hash_kind, rmi_model)
    sp = (ctypes.c_int64 * (n_parts + 1))()
    check(c.lib.hg_find_splits(c.bind_stream(), ctypes.byref(d), _p(keys), bin_count, n_parts, sp),
          "hg_find_splits")
    return list(sp)


def build_partitioned(keys, n_bins: int, splits, offset_bytes: int = 4, hash_kind: int = HG_HASH_MURMUR3,
                      rmi_model=None):
    """Per-partition CSRs.  Returns (offsets, csr_keys, part_offsets): partition p's
    offsets are offsets[splits[p]+p : splits[p+1]+p+1], its keys
    csr_keys[part_offsets[p]:part_offsets[p+1]]."""
    torch = _torch()
    _check_dev(keys)
    n_parts = len(splits) - 1
    c = context(keys.device.index)
    d = _desc(keys.numel(), n_bins, _key_bytes(keys), offset_bytes, hash_kind, rmi_model)
    off = torch.empty(int(splits[-1]) + n_parts, dtype=torch.int32 if offset_bytes == 4 else torch.int64,
                      device=keys.device)
    csr = torch.empty_like(keys)
    sp = (ctypes.c_int64 * (n_parts + 1))(*[int(x) for x in splits])
    po = (ctypes.c_uint64 * (n_parts + 1))()
    check(c.lib.hg_build_partitioned(c.bind_stream(), ctypes.byref(d), _p(keys), sp, n_parts, _p(off), _p(csr), po),
          "hg_build_partitioned")
    return off, csr, list(po)


# ---- staged tables: keys, offsets and CSR keys in host memory -------------

def _hp(a):
    """Host address of a contiguous numpy array or CPU tensor (pageable or pinned)."""
    if a is None:
        return ctypes.c_void_p(0)
    if hasattr(a, "ctypes"):
        if not a.flags.c_contiguous:
            raise ValueError("host arrays must be contiguous")
        return ctypes.c_void_p(a.ctypes.data)
    if a.is_cuda or not a.is_contiguous():
        raise ValueError("host tensors must be contiguous CPU tensors")
    return ctypes.c_void_p(a.data_ptr())


def _hbytes(a) -> int:
    kb = a.dtype.itemsize if hasattr(a, "ctypes") else a.element_size()
    if kb not in (4, 8):
        raise TypeError("keys must be 32- or 64-bit integers")
    return kb


def _hempty(like, n, nbytes):
    """Uninitialised host array of n elements of nbytes each, of like's kind."""
    if hasattr(like, "ctypes"):
        import numpy as np
        return np.empty(n, dtype=np.uint32 if nbytes == 4 else np.uint64)
    torch = _torch()
    return torch.empty(n, dtype=torch.int32 if nbytes == 4 else torch.int64, pin_memory=like.is_pinned())


def build_staged(keys, n_bins: int, offset_bytes: int = 4, hash_kind: int = HG_HASH_MURMUR3, rmi_model=None,
                 device_budget: int = 0, device=None, out=None):
    """hg_build over host-resident keys (numpy array or CPU tensor), for tables
    larger than the device: returns host (offsets, csr_keys)."""
    kb = _hbytes(keys)
    n = len(keys)
    c = context(device)
    if out is None:
        off, csr = _hempty(keys, n_bins + 1, offset_bytes), _hempty(keys, max(n, 1), kb)
    else:
        off, csr = out
    d = _desc(n, n_bins, kb, offset_bytes, hash_kind, rmi_model)
    check(c.lib.hg_build_staged(c.bind_stream(), ctypes.byref(d), _hp(keys), _hp(off), _hp(csr), device_budget),
          "hg_build_staged")
    return off, csr[:n]


def build_partitioned_staged(keys, n_bins: int, splits, offset_bytes: int = 4, hash_kind: int = HG_HASH_MURMUR3,
                             rmi_model=None, device_budget: int = 0, device=None):
    """build_partitioned over host-resident keys: returns host (offsets, csr_keys, part_offsets)."""
    kb = _hbytes(keys)
    n = len(keys)
    n_parts = len(splits) - 1
    c = context(device)
    off = _hempty(keys, int(splits[-1]) + n_parts, offset_bytes)
    csr = _hempty(keys, max(n, 1), kb)
    d = _desc(n, n_bins, kb, offset_bytes, hash_kind, rmi_model)
    sp = (ctypes.c_int64 * (n_parts + 1))(*[int(x) for x in splits])
    po = (ctypes.c_uint64 * (n_parts + 1))()
    check(c.lib.hg_build_partitioned_staged(c.bind_stream(), ctypes.byref(d), _hp(keys), sp, n_parts, _hp(off),
                                            _hp(csr), po, device_budget), "hg_build_partitioned_staged")
    return off, csr[:n], list(po)


def find_splits_staged(keys, n_bins: int, bin_count: int = 16000, n_parts: int = 16,
                       hash_kind: int = HG_HASH_MURMUR3, rmi_model=None, device_budget: int = 0, device=None):
    """find_splits over host-resident keys."""
    c = context(device)
    d = _desc(len(keys), n_bins, _hbytes(keys), 4, hash_kind, rmi_model)
    sp = (ctypes.c_int64 * (n_parts + 1))()
    check(c.lib.hg_find_splits_staged(c.bind_stream(), ctypes.byref(d), _hp(keys), bin_count, n_parts, sp,
                                      device_budget), "hg_find_splits_staged")
    return list(sp)


def intersect_staged(offA, keysA, offB, keysB, n_bins: int, emit: bool = True, hash_kind: int = HG_HASH_MURMUR3,
                     rmi_model=None, device_budget: int = 0, device=None, capacity: int | None = None, out=None):
    """intersect over two host-resident tables: returns (count, host matched keys or None).
    `out` (a host array of `capacity` keys or more) receives the list."""
    kb = _hbytes(keysA)
    ob = _hbytes(offA)
    c = context(device)
    dA = _desc(len(keysA), n_bins, kb, ob, hash_kind, rmi_model)
    dB = _desc(len(keysB), n_bins, kb, ob, hash_kind, rmi_model)
    cnt = ctypes.c_uint64(0)
    if out is not None and capacity is None:
        capacity = len(out)
    if capacity is None:
        check(c.lib.hg_intersect_staged(c.bind_stream(), ctypes.byref(dA), _hp(offA), _hp(keysA), ctypes.byref(dB),
                                        _hp(offB), _hp(keysB), None, 0, ctypes.byref(cnt), device_budget),
              "hg_intersect_staged")
        if not emit:
            return cnt.value, None
        capacity = cnt.value
    if out is None:
        out = _hempty(keysA, max(capacity, 1), kb)
    check(c.lib.hg_intersect_staged(c.bind_stream(), ctypes.byref(dA), _hp(offA), _hp(keysA), ctypes.byref(dB),
                                    _hp(offB), _hp(keysB), _hp(out), capacity, ctypes.byref(cnt), device_budget),
          "hg_intersect_staged")
    return cnt.value, out[: cnt.value]


def rmi_model_bytes(n_leaves: int = 4096) -> int:
    return int(_lib.load().hg_rmi_model_bytes(n_leaves))


def rmi_train(keys, n_bins: int, n_sample: int | None = None, seed: int = 7, n_leaves: int = 4096,
              report: bool = False):
    """Fit the learned CDF model on a seeded sample of `keys` (default 1%).
    Returns (model tensor, use_rmi) or, with report=True, (model, use_rmi, dict)
    where the dict holds the sampled max-bin / mean-probe figures of both hashes."""
    torch = _torch()
    _check_dev(keys)
    c = context(keys.device.index)
    n = keys.numel()
    if n_sample is None:
        n_sample = max(1, n // 100)
    model = torch.empty(rmi_model_bytes(n_leaves), dtype=torch.uint8, device=keys.device)
    use = ctypes.c_uint32(0)
    rep = _lib.RmiReport()
    check(c.lib.hg_rmi_train(c.bind_stream(), _p(keys), n, _key_bytes(keys), n_sample, seed, n_bins, n_leaves,
                             _p(model), ctypes.byref(use), ctypes.byref(rep)), "hg_rmi_train")
    if not report:
        return model, bool(use.value)
    return model, bool(use.value), {
        "sample_bins": rep.sample_bins, "max_bin_rmi": rep.max_bin_rmi, "max_bin_murmur": rep.max_bin_murmur,
        "mean_probe_rmi": rep.mean_probe_rmi, "mean_probe_murmur": rep.mean_probe_murmur}


def rmi_train_batch(keys, key_off, n_bins, n_sample, seeds, n_leaves: int = 4096, report: bool = False):
    """One learned model per table (table t = keys[key_off[t]:key_off[t+1]]),
    the models rmi_train would fit for (n_sample[t], seeds[t], n_bins[t]).
    Returns (models uint8 tensor, model t at t * rmi_model_bytes(n_leaves),
    list of use flags) or, with report=True, also the list of report dicts."""
    torch = _torch()
    _check_dev(keys)
    c = context(keys.device.index)
    T = len(n_bins)
    mb = rmi_model_bytes(n_leaves)
    models = torch.zeros(max(T * mb, 1), dtype=torch.uint8, device=keys.device)
    ko = (ctypes.c_uint64 * (T + 1))(*[int(x) for x in key_off])
    nb = (ctypes.c_uint64 * T)(*[int(x) for x in n_bins])
    ns = (ctypes.c_uint64 * T)(*[int(x) for x in n_sample])
    sd = (ctypes.c_uint64 * T)(*[int(x) for x in seeds])
    use = (ctypes.c_uint32 * T)()
    rep = (_lib.RmiReport * T)()
    check(c.lib.hg_rmi_train_batch(c.bind_stream(), T, _key_bytes(keys), _p(keys), ko, nb, ns, sd, n_leaves,
                                   _p(models), use, rep), "hg_rmi_train_batch")
    flags = [bool(use[t]) for t in range(T)]
    if not report:
        return models, flags
    reps = [{"sample_bins": r.sample_bins, "max_bin_rmi": r.max_bin_rmi, "max_bin_murmur": r.max_bin_murmur,
             "mean_probe_rmi": r.mean_probe_rmi, "mean_probe_murmur": r.mean_probe_murmur} for r in rep]
    return models, flags, reps


def table_bins(keys, n_bins: int, hash_kind: int = HG_HASH_MURMUR3, rmi_model=None):
    """Bins of `keys` under a table's hash (murmur3 or the learned model), int32 tensor of uint32."""
    torch = _torch()
    _check_dev(keys)
    c = context(keys.device.index)
    out = torch.empty(keys.numel(), dtype=torch.int32, device=keys.device)
    d = _desc(keys.numel(), n_bins, _key_bytes(keys), 4, hash_kind, rmi_model)
    check(c.lib.hg_table_bins(c.bind_stream(), ctypes.byref(d), _p(keys), keys.numel(), _p(out)), "hg_table_bins")
    return out


def _batch_desc(key_off, n_bins, off_off, key_bytes, hash_kinds=None, rmi_models=None, rmi_leaves=0):
    bd = _lib.BatchDesc()
    T = len(n_bins)
    bd.n_tables = T
    bd.key_bytes = key_bytes
    ko = (ctypes.c_uint64 * (T + 1))(*[int(x) for x in key_off])
    nb = (ctypes.c_uint64 * T)(*[int(x) for x in n_bins])
    oo = (ctypes.c_uint64 * T)(*[int(x) for x in off_off])
    bd.h_key_off = ko
    bd.h_n_bins = nb
    bd.h_off_off = oo
    hk = None
    if hash_kinds is not None:
        hk = (ctypes.c_uint32 * T)(*[int(x) for x in hash_kinds])
        bd.h_hash_kind = hk
    bd.d_rmi_models = rmi_models.data_ptr() if rmi_models is not None else None
    bd.rmi_leaves = rmi_leaves
    bd._keep = (ko, nb, oo, hk)
    return bd


def batch_build(keys, key_off, n_bins, hash_kinds=None, rmi_models=None, rmi_leaves=0):
    """Build len(n_bins) tables at once.  Table t's keys are keys[key_off[t]:key_off[t+1]].
    Returns (offsets int32, csr_keys, off_off) with table t's offsets at
    offsets[off_off[t] : off_off[t]+n_bins[t]+1]."""
    torch = _torch()
    _check_dev(keys)
    off_off, run = [], 0
    for m in n_bins:
        off_off.append(run)
        run += int(m) + 1
    bd = _batch_desc(key_off, n_bins, off_off, _key_bytes(keys), hash_kinds, rmi_models, rmi_leaves)
    c = context(keys.device.index)
    off = torch.empty(max(run, 1), dtype=torch.int32, device=keys.device)
    csr = torch.empty_like(keys)
    check(c.lib.hg_batch_build(c.bind_stream(), ctypes.byref(bd), _p(keys), _p(off), _p(csr)), "hg_batch_build")
    return off, csr, off_off


def batch_probe(offsets, csr_keys, key_off, n_bins, off_off, probe, probe_off, hash_kinds=None, rmi_models=None,
                rmi_leaves=0):
    """Per-table match counts of probe[probe_off[t]:probe_off[t+1]] against table t."""
    torch = _torch()
    _check_dev(offsets, csr_keys, probe)
    bd = _batch_desc(key_off, n_bins, off_off, _key_bytes(csr_keys), hash_kinds, rmi_models, rmi_leaves)
    c = context(csr_keys.device.index)
    T = len(n_bins)
    counts = torch.zeros(T, dtype=torch.int64, device=csr_keys.device)
    po = (ctypes.c_uint64 * (T + 1))(*[int(x) for x in probe_off])
    total = ctypes.c_uint64(0)
    check(c.lib.hg_batch_probe(c.bind_stream(), ctypes.byref(bd), _p(offsets), _p(csr_keys), _p(probe), po,
                               _p(counts), ctypes.byref(total)), "hg_batch_probe")
    return total.value, counts


def batch_probe_emit(offsets, csr_keys, key_off, n_bins, off_off, probe, probe_off, out_off, out=None,
                     hash_kinds=None, rmi_models=None, rmi_leaves=0):
    """batch_probe with the matches written: table t's matched probe keys go to
    out[out_off[t]:out_off[t+1]] (out allocated here when None).  Returns
    (total, counts, out); raises HGError (HG_ERR_CAPACITY) when a table's
    matches exceed its region."""
    torch = _torch()
    _check_dev(offsets, csr_keys, probe)
    bd = _batch_desc(key_off, n_bins, off_off, _key_bytes(csr_keys), hash_kinds, rmi_models, rmi_leaves)
    c = context(csr_keys.device.index)
    T = len(n_bins)
    oo = [int(x) for x in out_off]
    if len(oo) != T + 1:
        raise ValueError("out_off needs n_tables + 1 entries")
    if out is None:
        out = torch.empty(max(oo[-1], 1), dtype=csr_keys.dtype, device=csr_keys.device)
    counts = torch.zeros(T, dtype=torch.int64, device=csr_keys.device)
    po = (ctypes.c_uint64 * (T + 1))(*[int(x) for x in probe_off])
    ho = (ctypes.c_uint64 * (T + 1))(*oo)
    total = ctypes.c_uint64(0)
    check(c.lib.hg_batch_probe_emit(c.bind_stream(), ctypes.byref(bd), _p(offsets), _p(csr_keys), _p(probe), po, ho,
                                    _p(out), _p(counts), ctypes.byref(total)), "hg_batch_probe_emit")
    return total.value, counts, out


class BatchPlan:
    """A captured batch (hg_batch_plan): build of every table, then the
    per-table probe counts when probe keys are given, replayed by run().
    train=True trains every table's learned model inside the graph on each
    replay (rmi_models: the models' buffer, rmi_leaves leaves each; the
    sampled check picks learned or murmur3 per table, hash_kinds is ignored)."""

    def __init__(self, keys, key_off, n_bins, probe=None, probe_off=None, hash_kinds=None, rmi_models=None,
                 rmi_leaves=0, autotune=True, train=False):
        torch = _torch()
        _check_dev(keys)
        self.off_off, run = [], 0
        for m in n_bins:
            self.off_off.append(run)
            run += int(m) + 1
        T = len(n_bins)
        self._bd = _batch_desc(key_off, n_bins, self.off_off, _key_bytes(keys), hash_kinds, rmi_models, rmi_leaves)
        self._c = context(keys.device.index)
        self.offsets = torch.empty(max(run, 1), dtype=torch.int32, device=keys.device)
        self.csr = torch.empty_like(keys)
        self.counts = torch.zeros(T, dtype=torch.int64, device=keys.device)
        self._keep = (keys, probe, rmi_models)
        po = None
        if probe is not None:
            po = (ctypes.c_uint64 * (T + 1))(*[int(x) for x in probe_off])
        self._plan = ctypes.c_void_p()
        check(self._c.lib.hg_batch_plan_create(self._c.bind_stream(), ctypes.byref(self._bd), _p(keys),
                                               _p(self.offsets), _p(self.csr),
                                               _p(probe) if probe is not None else None, po,
                                               _p(self.counts) if probe is not None else None,
                                               (_lib.HG_PLAN_AUTOTUNE if autotune else 0) |
                                               (_lib.HG_PLAN_TRAIN if train else 0),
                                               ctypes.byref(self._plan)), "hg_batch_plan_create")

    def run(self):
        self._c.bind_stream()
        check(self._c.lib.hg_batch_plan_run(self._plan), "hg_batch_plan_run")

    def info(self):
        i = _lib.BatchPlanInfo()
        check(self._c.lib.hg_batch_plan_query(self._plan, ctypes.byref(i)), "hg_batch_plan_query")
        return {"tile_keys": i.tile_keys, "fill": i.fill, "candidates": i.candidates, "tuned_ms": i.tuned_ms,
                "learned_tables": i.learned_tables}

    def close(self):
        if self._plan:
            self._c.lib.hg_batch_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
