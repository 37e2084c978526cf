"""GPU build parity: hg_build (HIP kernels) against the CPU oracle.

Offsets must be identical; bins must hold the same keys (compared sorted,
since the reference fills bins with atomics)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from hashgraph_amd import workloads as wl
from oracle import oracle as o


def _to_dev(a, dev):
    import torch
    t = torch.from_numpy(a.view(np.int32 if a.dtype == np.uint32 else np.int64).copy())
    return t.to(dev)


def _host(t, dtype):
    return t.cpu().numpy().view(dtype)


def _check_build(keys, M, dev, ob=4):
    import hashgraph_amd as hg
    print(f"build n={len(keys)} M={M} kb={keys.itemsize} ob={ob}", flush=True)
    dk = _to_dev(keys, dev)
    off, csr = hg.build(dk, M, offset_bytes=ob)
    goff = _host(off, np.uint32 if ob == 4 else np.uint64).astype(np.uint64)
    gcsr = _host(csr, keys.dtype).copy()
    eoff, ecsr = o.build_csr(keys, M)
    assert np.array_equal(goff, eoff), "offsets differ"
    o.canon(goff, gcsr)
    o.canon(eoff, ecsr)
    assert np.array_equal(gcsr, ecsr), "bin contents differ"


def test_c1_shape_u32(dev):
    # 2^20 uint32 keys, seed 1, 2^20 bins
    keys = wl.uniform_keys(1, 1 << 20, 4)
    _check_build(keys, 1 << 20, dev)


@pytest.mark.parametrize("n,M,kb,ob", [
    (1, 1, 4, 4), (2, 1, 4, 4), (1000, 1, 4, 4), (100, 100000, 4, 4), (77777, 3, 4, 4), (5000, 97, 8, 4),
    (300000, 299993, 4, 8), (1 << 18, 1 << 22, 4, 4), (1 << 18, 1 << 22, 8, 8), (123457, 65536, 8, 4),
    (200000, 1000, 4, 4), (1 << 20, 1 << 16, 4, 4),
])
def test_shapes(n, M, kb, ob, dev):
    keys = wl.uniform_keys(n + M, n, kb)
    _check_build(keys, M, dev, ob)


def test_empty_table(dev):
    import torch
    import hashgraph_amd as hg
    keys = torch.empty(0, dtype=torch.int32, device=dev)
    off, csr = hg.build(keys, 50)
    assert off.cpu().numpy().tolist() == [0] * 51


def test_oversized_buckets_single_bin(dev):
    # M = 1: every key in one bin, far beyond one workgroup's VGPR capacity
    keys = wl.uniform_keys(5, 150000, 4)
    _check_build(keys, 1, dev)


def test_heavy_duplicates(dev):
    # 200k keys over 50 distinct values: bins with thousands of equal keys
    base = wl.uniform_keys(6, 50, 8)
    idx = (wl.stream_np(7, 200000) % np.uint64(50)).astype(np.int64)
    keys = base[idx]
    _check_build(keys, 1 << 16, dev)


def test_many_oversized_buckets(dev):
    # 300 hot values x 60k copies plus uniform keys: about 300 buckets beyond a
    # workgroup's capacity, more than the multi-workgroup path takes in one
    # launch (256); the rest take the single-workgroup fallback
    hot = wl.uniform_keys(8, 300, 4)
    keys = np.concatenate([np.repeat(hot, 60000), wl.uniform_keys(9, 2000000, 4)])
    keys = keys[np.argsort(wl.stream_np(10, len(keys)), kind="stable")]
    _check_build(keys, 1 << 22, dev)


def test_giant_bin_u64(dev):
    # one value with 3M copies (a Zipf head) among 1M uniform uint64 keys
    keys = np.concatenate([np.full(3000000, 0x123456789ABCDEF, dtype=np.uint64), wl.uniform_keys(11, 1000000, 8)])
    keys = keys[np.argsort(wl.stream_np(12, len(keys)), kind="stable")]
    _check_build(keys, 1 << 20, dev, ob=8)


def test_sequential_keys(dev):
    keys = wl.sequential_keys(1 << 20, 8)
    _check_build(keys, 1 << 20, dev, ob=8)


def test_u64_low_word_collisions(dev):
    lo = wl.uniform_keys(8, 5000, 4).astype(np.uint64)
    keys = np.concatenate([lo, lo | np.uint64(1 << 33), lo | np.uint64(0xFFFF << 40)])
    _check_build(keys, 4096, dev)


def test_repeat_build_is_stable(dev):
    import hashgraph_amd as hg
    keys = wl.uniform_keys(10, 1 << 19, 4)
    dk = _to_dev(keys, dev)
    a = hg.build(dk, 1 << 19)
    b = hg.build(dk, 1 << 19)
    assert np.array_equal(a[0].cpu().numpy(), b[0].cpu().numpy())


def test_full_size_2e28_u32_properties(dev):
    """2^28 uint32 keys, 2^28 bins: size-independent checks on the device."""
    import torch
    import hashgraph_amd as hg
    n = M = 1 << 28
    keys = wl.uniform_keys(1, n, 4, device=dev)
    off, csr = hg.build(keys, M)
    torch.cuda.synchronize()
    offl = off.to(torch.int64)
    deg = offl[1:] - offl[:-1]
    assert int(offl[0]) == 0 and int(offl[-1]) == n
    assert bool((deg >= 0).all())
    bins = hg.hash_bins(csr, M).to(torch.int64)
    # the bin of every CSR key is the bin whose range holds it
    expect = torch.repeat_interleave(torch.arange(M, device=dev), deg, output_size=n)
    assert bool((bins == expect).all())
    del expect, bins
    # same multiset of keys
    assert torch.equal(torch.sort(csr.view(torch.int32))[0], torch.sort(keys)[0])
    # offsets equal the bincount of the input's bins
    cnt = torch.bincount(hg.hash_bins(keys, M).to(torch.int64), minlength=M)
    assert torch.equal(cnt, deg)


def test_hash_bins_matches_oracle(dev):
    import hashgraph_amd as hg
    for kb in (4, 8):
        keys = wl.uniform_keys(12, 100000, kb)
        for M in (1, 1000, 1 << 20, 4294967291):
            g = hg.hash_bins(_to_dev(keys, dev), M).cpu().numpy().view(np.uint32)
            assert np.array_equal(g, o.bins(keys, M))


@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 1000003, 1 << 24])
@pytest.mark.parametrize("width", [4, 8])
def test_exclusive_scan(n, width, dev):
    import torch
    import hashgraph_amd as hg
    dt = np.uint32 if width == 4 else np.uint64
    x = (wl.stream_np(n, n) % np.uint64(1000)).astype(dt)
    t = torch.from_numpy(x.view(np.int32 if width == 4 else np.int64).copy()).to(dev)
    out = hg.exclusive_scan(t).cpu().numpy().view(dt)
    exp = np.zeros(n + 1, dtype=dt)
    exp[1:] = np.cumsum(x, dtype=dt)
    assert np.array_equal(out, exp)


def test_exclusive_scan_u64_beyond_2e62(dev):
    """uint64 sums beyond 2^62 and 2^63 (and wrapping mod 2^64): the scan's
    64-bit aggregate and prefix words carry them exactly."""
    import torch
    import hashgraph_amd as hg
    n = 100000
    x = (wl.stream_np(77, n) >> np.uint64(9)).astype(np.uint64)  # ~2^54 each: the sum passes 2^62, 2^63, 2^64
    x[n // 2] = np.uint64(2**63 + 12345)
    t = torch.from_numpy(x.view(np.int64).copy()).to(dev)
    out = hg.exclusive_scan(t).cpu().numpy().view(np.uint64)
    exp = np.zeros(n + 1, dtype=np.uint64)
    exp[1:] = np.cumsum(x, dtype=np.uint64)
    assert exp[-1] < exp[n // 2]  # wrapped at least once
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("kb", [4, 8])
def test_build_indexed_carries_positions(kb, dev):
    """hg_build_indexed (the reference's INDEX_TRACK entries {key, ind}): the
    offsets equal hg_build's, every CSR key equals the input key at its
    position, and the positions are a permutation of the input."""
    import ctypes
    import torch
    import hashgraph_amd as hg
    n, M = 300000, 1 << 17
    keys = wl.uniform_keys(71, n, kb)
    keys[:5000] = keys[5000:10000]  # duplicates
    dk = _to_dev(keys, dev)
    c = hg.context(0)
    off = torch.empty(M + 1, dtype=torch.int32, device=dev)
    csr = torch.empty_like(dk)
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    d = hg._desc(n, M, kb)
    st = c.lib.hg_build_indexed(c.bind_stream(), ctypes.byref(d), hg._p(dk), hg._p(off), hg._p(csr), hg._p(idx))
    assert st == 0, hg._lib.load().hg_last_error()
    ref_off, _ = hg.build(dk, M)
    torch.cuda.synchronize()
    assert torch.equal(off, ref_off)
    assert torch.equal(csr, dk[idx])
    assert torch.equal(torch.sort(idx)[0], torch.arange(n, device=dev))
    eoff, _ = o.build_csr(keys, M)
    assert np.array_equal(off.cpu().numpy().astype(np.uint64), eoff)


@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("shape", ["three", "cut", "one"])
def test_build_partitioned_indexed_carries_positions(kb, shape, dev):
    """hg_build_partitioned_indexed: offsets and partition offsets equal
    hg_build_partitioned's, every bin holds the same keys, every CSR key equals
    the input key at its position, and the positions are exactly those of the
    keys whose bins lie in the splits' range (all keys, or, with cut splits,
    the keys of the first seven eighths of the bins; one partition over every
    bin is built as a plain table)."""
    import ctypes
    import torch
    import hashgraph_amd as hg
    n, M = 300000, 1 << 17
    keys = wl.uniform_keys(73, n, kb)
    keys[:5000] = keys[5000:10000]  # duplicates
    dk = _to_dev(keys, dev)
    splits = {"three": [0, M // 3, M // 2, M], "cut": [0, M // 2, 7 * M // 8], "one": [0, M]}[shape]
    P = len(splits) - 1
    print(f"partitioned indexed kb={kb} splits={splits}", flush=True)
    ref_off, ref_csr, ref_po = hg.build_partitioned(dk, M, splits)
    c = hg.context(0)
    off = torch.empty(splits[-1] + P, dtype=torch.int32, device=dev)
    csr = torch.empty_like(dk)
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    d = hg._desc(n, M, kb)
    sp = (ctypes.c_int64 * (P + 1))(*splits)
    po = (ctypes.c_uint64 * (P + 1))()
    st = c.lib.hg_build_partitioned_indexed(c.bind_stream(), ctypes.byref(d), hg._p(dk), sp, P, hg._p(off),
                                            hg._p(csr), hg._p(idx), po)
    assert st == 0, hg._lib.load().hg_last_error()
    torch.cuda.synchronize()
    kept = int(po[P])
    assert list(po) == ref_po
    assert torch.equal(off, ref_off)
    assert torch.equal(csr[:kept], dk[idx[:kept]])
    bins = hg.table_bins(dk, M).to(torch.int64) & 0xFFFFFFFF
    inr = (bins >= splits[0]) & (bins < splits[-1])
    assert torch.equal(torch.sort(idx[:kept])[0], torch.nonzero(inr).flatten())
    goff = off.cpu().numpy()
    gcsr, ecsr = csr.cpu().numpy().view(keys.dtype), ref_csr.cpu().numpy().view(keys.dtype)
    for p in range(P):
        loc = goff[splits[p] + p: splits[p + 1] + p + 1]
        a = gcsr[ref_po[p]: ref_po[p + 1]].copy()
        b = ecsr[ref_po[p]: ref_po[p + 1]].copy()
        o.canon(loc, a)
        o.canon(loc, b)
        assert np.array_equal(a, b), f"partition {p} bins differ"
