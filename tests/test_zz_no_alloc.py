"""hg_build_noalloc, the build that allocates nothing (its temporaries live
in the library's static device area), called through the C ABI: offsets
identical to hg_build's and bins equal as sorted multisets, on two calls in a
row (the second with a smaller table, over the area the first one left), for
both key widths, and a table beyond the area is refused with a clear error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from hashgraph_amd import workloads as wl
from oracle import oracle as o


def _to_dev(a, dev):
    import torch
    return torch.from_numpy(a.view(np.int32 if a.dtype == np.uint32 else np.int64).copy()).to(dev)


def _noalloc(dk, M, kb, ob):
    import torch
    import hashgraph_amd as hg
    c = hg.context(dk.device.index)
    off = torch.full((M + 1,), -1, dtype=torch.int32 if ob == 4 else torch.int64, device=dk.device)
    csr = torch.zeros_like(dk)
    d = hg._desc(dk.numel(), M, kb, ob)
    st = c.lib.hg_build_noalloc(c.bind_stream(), ctypes.byref(d), hg._p(dk), hg._p(off), hg._p(csr))
    return st, off, csr


def _same_table(off_a, csr_a, off_b, csr_b, dtype):
    oa = off_a.cpu().numpy().astype(np.uint64)
    ob_ = off_b.cpu().numpy().astype(np.uint64)
    assert np.array_equal(oa, ob_), "offsets differ"
    a = csr_a.cpu().numpy().view(dtype).copy()
    b = csr_b.cpu().numpy().view(dtype).copy()
    o.canon(oa, a)
    o.canon(oa, b)
    assert np.array_equal(a, b), "bin contents differ"


@pytest.mark.parametrize("kb,ob", [(4, 4), (8, 8), (8, 4)])
def test_noalloc_matches_build_twice(dev, kb, ob):
    import torch
    import hashgraph_amd as hg
    dtype = np.uint32 if kb == 4 else np.uint64
    for n, M, seed in ((1 << 22, 1 << 22, 61), (100003, 1 << 15, 62)):
        print(f"noalloc kb={kb} ob={ob} n={n} M={M}", flush=True)
        keys = wl.uniform_keys(seed, n, kb)
        dk = _to_dev(keys, dev)
        st, off, csr = _noalloc(dk, M, kb, ob)
        assert st == 0, hg._lib.load().hg_last_error()
        ref_off, ref_csr = hg.build(dk, M, offset_bytes=ob)
        torch.cuda.synchronize()
        _same_table(off, csr, ref_off, ref_csr, dtype)
        eoff, _ = o.build_csr(keys, M)
        assert np.array_equal(off.cpu().numpy().astype(np.uint64), eoff)


def test_noalloc_skewed_and_empty(dev):
    """A hot key (oversized buckets through the multi-workgroup path) and an
    empty table, after a larger call."""
    import torch
    import hashgraph_amd as hg
    keys = np.concatenate([np.full(200000, 7, dtype=np.uint32), wl.uniform_keys(63, 300000, 4)])
    dk = _to_dev(keys, dev)
    st, off, csr = _noalloc(dk, 1 << 16, 4, 4)
    assert st == 0
    ref_off, ref_csr = hg.build(dk, 1 << 16)
    torch.cuda.synchronize()
    _same_table(off, csr, ref_off, ref_csr, np.uint32)
    empty = torch.empty(0, dtype=torch.int32, device=dev)
    st, off, _ = _noalloc(empty, 100, 4, 4)
    assert st == 0 and int(off.abs().sum()) == 0


def test_noalloc_refuses_oversized_table(dev):
    import hashgraph_amd as hg
    dk = _to_dev(wl.uniform_keys(64, 1 << 23, 8), dev)
    st, _, _ = _noalloc(dk, 1 << 23, 8, 8)
    assert st == 6  # HG_ERR_UNSUPPORTED
    assert b"static area" in hg._lib.load().hg_last_error()
