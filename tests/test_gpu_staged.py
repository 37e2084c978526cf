"""GPU parity of the staged tables: keys, offsets and CSR keys in host memory,
passed through HBM in chunks and bin-range slabs (the replacement of the
reference's MANAGED_MEM mode, src/MultiHashGraph.cu:229-247,
include/MultiHashGraphHostOperators.cuh:255-280).  Results must equal the
oracle's (and the in-HBM calls') on the same inputs; the device budgets are
small enough that every call runs in many chunks and slabs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from hashgraph_amd import workloads as wl
from oracle import oracle as o

MiB = 1 << 20


def _u(a, nbytes):
    """numpy view of a host result (numpy array or CPU tensor) as unsigned."""
    a = a if hasattr(a, "ctypes") else a.numpy()
    return a.view(np.uint32 if nbytes == 4 else np.uint64)


def _host_keys(keys, pinned):
    if not pinned:
        return keys
    import torch
    return torch.from_numpy(keys.view(np.int32 if keys.itemsize == 4 else np.int64).copy()).pin_memory()


def _check_build(keys, M, ob, budget, pinned=False):
    import hashgraph_amd as hg
    print(f"staged build n={len(keys)} M={M} kb={keys.itemsize} ob={ob} budget={budget} pinned={pinned}",
          flush=True)
    off, csr = hg.build_staged(_host_keys(keys, pinned), M, offset_bytes=ob, device_budget=budget, device=0)
    goff = _u(off, ob).astype(np.uint64)
    gcsr = _u(csr, keys.itemsize).copy()
    eoff, ecsr = o.build_csr(keys, M)
    assert np.array_equal(goff, eoff), "offsets differ"
    o.canon(eoff, gcsr)
    o.canon(eoff, ecsr)
    assert np.array_equal(gcsr, ecsr), "bin contents differ"


@pytest.mark.parametrize("n,M,kb,ob,budget,pinned", [
    (1 << 20, 1 << 20, 4, 4, 4 * MiB, False),
    (1 << 20, 1 << 18, 8, 8, 4 * MiB, True),
    (300000, 7, 4, 4, 4 * MiB, False),
    (1 << 20, 3000017, 4, 8, 4 * MiB, True),
    (77777, 1, 8, 4, 8 * MiB, False),
    (1 << 20, 1 << 20, 4, 4, 64 * MiB, False),
])
def test_build_shapes(n, M, kb, ob, budget, pinned, dev):
    keys = wl.uniform_keys(n + 5, n, kb)
    _check_build(keys, M, ob, budget, pinned)


def test_build_zipf_u64(dev):
    # top key ~10% of the keys: its bin fills most of one slab
    keys = wl.zipf_keys(3, 1 << 20, 1.1)
    _check_build(keys, 1 << 20, 8, 16 * MiB)


def test_build_empty(dev):
    import hashgraph_amd as hg
    off, csr = hg.build_staged(np.empty(0, dtype=np.uint32), 100, device_budget=4 * MiB, device=0)
    assert _u(off, 4).tolist() == [0] * 101 and len(csr) == 0


def test_bin_beyond_budget_is_refused(dev):
    # 2^20 copies of one key: one bin of 4 MiB of keys, more than a 4 MiB
    # budget's slab; the count pass finds it and the call refuses it
    import hashgraph_amd as hg
    keys = np.full(1 << 20, 12345, dtype=np.uint32)
    with pytest.raises(hg.HGError) as e:
        hg.build_staged(keys, 1 << 16, device_budget=4 * MiB, device=0)
    assert e.value.status == 6
    # a budget whose slab holds the bin builds it
    _check_build(keys, 1 << 16, 4, 64 * MiB)


def test_build_learned_hash_equals_in_hbm_build(dev):
    import torch
    import hashgraph_amd as hg
    keys = wl.lognormal_keys(5, 1 << 20)
    M = 1 << 20
    dk = torch.from_numpy(keys.view(np.int64).copy()).to(dev)
    model, _ = hg.rmi_train(dk, M)
    eoff, ecsr = hg.build(dk, M, hash_kind=hg.HG_HASH_RMI, rmi_model=model)
    off, csr = hg.build_staged(keys, M, hash_kind=hg.HG_HASH_RMI, rmi_model=model, device_budget=4 * MiB,
                               device=0)
    eoff = eoff.cpu().numpy().view(np.uint32).astype(np.uint64)
    ecsr = ecsr.cpu().numpy().view(np.uint64).copy()
    goff = _u(off, 4).astype(np.uint64)
    gcsr = _u(csr, 8).copy()
    assert np.array_equal(goff, eoff)
    o.canon(eoff, gcsr)
    o.canon(eoff, ecsr)
    assert np.array_equal(gcsr, ecsr)


@pytest.mark.parametrize("n,M,P,kb", [(1 << 20, 1 << 20, 16, 4), (200000, 1 << 18, 1, 8), (5000, 5000, 3, 4)])
def test_find_splits_staged(n, M, P, kb, dev):
    import hashgraph_amd as hg
    keys = wl.uniform_keys(n, n, kb)
    got = hg.find_splits_staged(keys, M, 16000, P, device_budget=4 * MiB, device=0)
    exp, _ = o.find_splits(o.bins(keys, M), M, 16000, P)
    assert got == exp.tolist()


def _check_partitioned(keys, M, splits, ob, budget):
    import hashgraph_amd as hg
    P = len(splits) - 1
    print(f"staged partitioned n={len(keys)} M={M} P={P} kb={keys.itemsize} ob={ob} budget={budget}", flush=True)
    off, csr, po = hg.build_partitioned_staged(keys, M, splits, offset_bytes=ob, device_budget=budget, device=0)
    eoff, ecsr = o.build_multi(keys, M, np.array(splits, dtype=np.int64))
    goff = _u(off, ob).astype(np.uint64)
    assert np.array_equal(goff, eoff), "per-partition offsets differ"
    gcsr = _u(csr, keys.itemsize).copy()
    # keys kept = the sum of every partition's last offset (the rest are dropped)
    assert po[P] == sum(int(eoff[splits[p + 1] + p]) for p in range(P))
    for p in range(P):
        lo, hi = splits[p], splits[p + 1]
        loc = eoff[lo + p: hi + p + 1]
        a = gcsr[po[p]: po[p + 1]].copy()
        b = ecsr[po[p]: po[p + 1]].copy()
        o.canon(loc, a)
        o.canon(loc, b)
        assert np.array_equal(a, b), f"partition {p} bins differ"


def test_partitioned_zipf_u64(dev):
    import hashgraph_amd as hg
    keys = wl.zipf_keys(3, 1 << 20, 1.1)
    M = 1 << 20
    splits = hg.find_splits_staged(keys, M, 16000, 16, device_budget=16 * MiB, device=0)
    _check_partitioned(keys, M, splits, 8, 16 * MiB)


def test_partitioned_empty_and_dropped_ranges(dev):
    # empty partitions, and keys above the last split dropped, as in the reference
    keys = wl.uniform_keys(9, 400000, 4)
    M = 1 << 18
    _check_partitioned(keys, M, [0, 1000, 1000, 70000, 70000, 200000], 4, 4 * MiB)


def test_intersect_staged_equals_oracle(dev):
    import hashgraph_amd as hg
    M = 1 << 20
    A = wl.uniform_keys(31, 1 << 20, 4)
    B = wl.probe_keys(A, 32, 1 << 20, 0.25, 4)
    offA, kA = hg.build_staged(A, M, device_budget=4 * MiB, device=0)
    offB, kB = hg.build_staged(B, M, device_budget=4 * MiB, device=0)
    cnt, out = hg.intersect_staged(offA, kA, offB, kB, M, device_budget=4 * MiB, device=0)
    eoA, ekA = o.build_csr(A, M)
    eoB, ekB = o.build_csr(B, M)
    ecnt, eout = o.intersect(eoA, ekA, eoB, ekB)
    assert cnt == ecnt == o.bag_intersect_count(A, B)
    assert np.array_equal(np.sort(_u(out, 4)), np.sort(eout))


def test_intersect_staged_output_in_halves(dev):
    # 2^20 keys per side over 2^17 values (~8 copies a side): ~8.4 M matches,
    # far more than a 4 MiB budget stages at once, in bins of ~64 matches
    import hashgraph_amd as hg
    M = 1 << 17
    A = (wl.uniform_keys(41, 1 << 20, 8) % np.uint64(1 << 17)).astype(np.uint64)
    B = (wl.uniform_keys(42, 1 << 20, 8) % np.uint64(1 << 17)).astype(np.uint64)
    offA, kA = hg.build_staged(A, M, offset_bytes=8, device_budget=4 * MiB, device=0)
    offB, kB = hg.build_staged(B, M, offset_bytes=8, device_budget=4 * MiB, device=0)
    cnt, out = hg.intersect_staged(offA, kA, offB, kB, M, device_budget=4 * MiB, device=0)
    assert cnt == o.bag_intersect_count(A, B)
    eoA, ekA = o.build_csr(A, M)
    eoB, ekB = o.build_csr(B, M)
    _, eout = o.intersect(eoA, ekA, eoB, ekB)
    assert np.array_equal(np.sort(_u(out, 8)), np.sort(eout))
    # a capacity below the count is refused with the count reported
    with pytest.raises(hg.HGError) as e:
        hg.intersect_staged(offA, kA, offB, kB, M, device_budget=4 * MiB, device=0, capacity=cnt - 1)
    assert e.value.status == 4


def test_intersect_staged_count_zipf_equals_in_hbm(dev):
    import torch
    import hashgraph_amd as hg
    M = 1 << 18
    A = wl.zipf_keys(7, 1 << 18, 1.1)
    B = wl.zipf_keys(8, 1 << 18, 1.1)
    offA, kA = hg.build_staged(A, M, offset_bytes=8, device_budget=8 * MiB, device=0)
    offB, kB = hg.build_staged(B, M, offset_bytes=8, device_budget=8 * MiB, device=0)
    cnt, _ = hg.intersect_staged(offA, kA, offB, kB, M, emit=False, device_budget=8 * MiB, device=0)
    t = lambda a: torch.from_numpy(a.view(np.int64).copy()).to(dev)
    dcnt, _ = hg.intersect(t(offA), t(kA), t(offB), t(kB), M, emit=False)
    assert cnt == dcnt == o.bag_intersect_count(A, B)


def test_c2_shape_staged_equals_in_hbm_build(dev):
    # 2^28 uint32 keys and bins, host-resident, through a 256 MiB budget (the
    # table and its CSR are 2 GiB): offsets identical to hg_build's, bins equal
    # as multisets (compared on the device by sorting (bin, key) pairs)
    import torch
    import hashgraph_amd as hg
    n = M = 1 << 28
    dk = wl.uniform_keys(1, n, 4, device=dev)
    hk = dk.cpu().pin_memory()
    print("staged C2 shape: build", flush=True)
    off, csr = hg.build_staged(hk, M, device_budget=256 * MiB, device=0)
    eoff, ecsr = hg.build(dk, M)
    goff = off.to(dev)
    assert torch.equal(goff, eoff)
    del dk
    counts = (eoff[1:] - eoff[:-1]).to(torch.int64)
    binid = torch.repeat_interleave(torch.arange(M, device=dev, dtype=torch.int64), counts)
    del counts
    gc = csr.to(dev).to(torch.int64) & 0xFFFFFFFF
    ec = ecsr.to(torch.int64) & 0xFFFFFFFF
    a = torch.sort((binid << 32) | gc).values
    del gc
    b = torch.sort((binid << 32) | ec).values
    assert torch.equal(a, b)


def test_staged_build_of_2e32_keys(dev):
    # 2^32 uint32 keys (16 GiB, pinned host memory; key i = i * 0x9E3779B1 mod
    # 2^32, every 32-bit value once), 2^24 bins, 8-byte offsets, through an
    # 8 GiB budget: more keys than one in-HBM call takes.
    # Checked exactly on the device, chunk by chunk: offsets against the
    # histogram of the input's bins; every CSR key hashes to the bin its
    # position lies in; per-bin sums of keys and of squared keys (mod 2^64)
    # equal the input's.
    import torch
    import hashgraph_amd as hg
    n, M, G = 1 << 32, 1 << 24, 1 << 28
    c = 0x9E3779B1

    def keys_of(a, b):
        # (i * c) mod 2^32 from 16-bit halves, inside int64 range
        i = torch.arange(a, b, dtype=torch.int64, device=dev)
        lo, hi = i & 0xFFFF, (i >> 16) & 0xFFFF
        return (lo * c + (((hi * c) & 0xFFFF) << 16)) & 0xFFFFFFFF

    def as_i32(k):
        return (k - ((k >> 31) << 32)).to(torch.int32)

    print("2^32 keys: generate", flush=True)
    hk = torch.empty(n, dtype=torch.int32, pin_memory=True)
    cnt, s1, s2 = (torch.zeros(M, dtype=torch.int64, device=dev) for _ in range(3))
    for a in range(0, n, G):
        b = min(n, a + G)
        k = keys_of(a, b)
        k32 = as_i32(k)
        hk[a:b].copy_(k32)
        bins = hg.hash_bins(k32, M).to(torch.int64)
        cnt += torch.bincount(bins, minlength=M)
        s1.index_add_(0, bins, k)
        s2.index_add_(0, bins, k * k)
    print("2^32 keys: staged build", flush=True)
    off, csr = hg.build_staged(hk, M, offset_bytes=8, device_budget=8 << 30, device=0)
    del hk
    goff = off.to(dev)
    assert int(goff[0]) == 0 and int(goff[-1]) == n
    assert torch.equal(goff[1:] - goff[:-1], cnt)
    print("2^32 keys: check the CSR", flush=True)
    t1, t2 = (torch.zeros(M, dtype=torch.int64, device=dev) for _ in range(2))
    for a in range(0, n, G):
        b = min(n, a + G)
        k32 = csr[a:b].to(dev)
        bins = hg.hash_bins(k32, M).to(torch.int64)
        pos = torch.arange(a, b, dtype=torch.int64, device=dev)
        assert torch.equal(bins, torch.searchsorted(goff, pos, right=True) - 1)
        k = k32.to(torch.int64) & 0xFFFFFFFF
        t1.index_add_(0, bins, k)
        t2.index_add_(0, bins, k * k)
    assert torch.equal(t1, s1) and torch.equal(t2, s2)


def test_device_offsets_are_refused(dev):
    # the staged calls read offsets (the join) and write empty partitions'
    # entries (the builds) from the host: device memory there is an argument
    # error, not a crash
    import ctypes
    import torch
    import hashgraph_amd as hg
    c = hg.context(0)
    keys = wl.uniform_keys(3, 1000, 4)
    off, csr = hg.build_staged(keys, 1000, device_budget=4 * MiB, device=0)
    doff = torch.from_numpy(off.view(np.int32).copy()).to(dev)
    d = hg._desc(len(keys), 1000, 4, 4)
    cnt = ctypes.c_uint64(0)
    st = c.lib.hg_intersect_staged(c.bind_stream(), ctypes.byref(d), ctypes.c_void_p(doff.data_ptr()), hg._hp(csr),
                                   ctypes.byref(d), hg._hp(off), hg._hp(csr), None, 0, ctypes.byref(cnt), 4 * MiB)
    assert st == 1
    dout = torch.empty(1001, dtype=torch.int32, device=dev)
    st = c.lib.hg_build_staged(c.bind_stream(), ctypes.byref(d), hg._hp(keys), ctypes.c_void_p(dout.data_ptr()),
                               hg._hp(csr), 4 * MiB)
    assert st == 1
