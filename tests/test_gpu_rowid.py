"""GPU parity of the row-id join (hg_intersect_indexed): the index-pair output
of MultiHashGraph::intersect over INDEX_TRACK tables (keyval_ind entries,
include/MultiHashGraph.cuh:82-87; pair output
include/MultiHashGraphDeviceOperators.cuh:535-537).  The pairs must equal,
as a set, {(i, j) : A[i] == B[j]} over the input positions: a closed form that
no hash or bin layout enters, computed here by sorting."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from hashgraph_amd import workloads as wl


def _to_dev(a, dev):
    import torch
    return torch.from_numpy(a.view(np.int32 if a.dtype == np.uint32 else np.int64).copy()).to(dev)


def _expected_pairs(A, B):
    """Every (i, j) with A[i] == B[j], packed as i << 32 | j, sorted."""
    ia, ib = np.argsort(A, kind="stable"), np.argsort(B, kind="stable")
    va, sa, na = np.unique(A[ia], return_index=True, return_counts=True)
    vb, sb, nb = np.unique(B[ib], return_index=True, return_counts=True)
    _, xa, xb = np.intersect1d(va, vb, assume_unique=True, return_indices=True)
    a0, ac, b0, bc = sa[xa], na[xa], sb[xb], nb[xb]
    per = ac.astype(np.int64) * bc
    total = int(per.sum())
    if total == 0:
        return np.empty(0, dtype=np.uint64)
    k = np.repeat(np.arange(len(per)), per)
    local = np.arange(total, dtype=np.int64) - np.repeat(np.cumsum(per) - per, per)
    i = ia[a0[k] + local // bc[k]].astype(np.uint64)
    j = ib[b0[k] + local % bc[k]].astype(np.uint64)
    return np.sort((i << np.uint64(32)) | j)


def _check(A, B, M, dev, ob=4):
    import hashgraph_amd as hg
    print(f"row-id join nA={len(A)} nB={len(B)} M={M} kb={A.itemsize} ob={ob}", flush=True)
    oa, ka, ia = hg.build_indexed(_to_dev(A, dev), M, offset_bytes=ob)
    ob_, kb, ib = hg.build_indexed(_to_dev(B, dev), M, offset_bytes=ob)
    cnt, pa, pb = hg.intersect_indexed(oa, ka, ia, ob_, kb, ib, M)
    exp = _expected_pairs(A, B)
    assert cnt == len(exp)
    got = np.sort((pa.cpu().numpy().astype(np.uint64) << np.uint64(32)) | pb.cpu().numpy().astype(np.uint64))
    assert np.array_equal(got, exp)
    return cnt


@pytest.mark.parametrize("frac", [0.25, 0.5, 1.0])
def test_uniform_u32_light_bins(frac, dev):
    # one key a bin on average: a workgroup's 2048 bins hold ~2048 * frac
    # matches, so 0.5 puts ranges either side of the 1024 matches a workgroup
    # buffers and 1.0 sends every range through the re-walk
    A = wl.uniform_keys(1, 1 << 18, 4)
    B = wl.probe_keys(A, 2, 1 << 18, frac, 4)
    assert _check(A, B, 1 << 18, dev) > (1 << 18) * frac * 0.9


@pytest.mark.parametrize("M", [64, 4])
def test_dense_duplicates_take_the_tiled_path(M, dev):
    # 2^16 keys a side over 2^12 values (~16 copies each): every bin holds
    # ~2^16 / M keys a side, far above 4096 key pairs; with M = 4 a bin spans
    # 16 x 16 tiles of 1024 x 1024 entries
    A = (wl.uniform_keys(5, 1 << 16, 4) % np.uint32(1 << 12)).astype(np.uint32)
    B = (wl.uniform_keys(6, 1 << 16, 4) % np.uint32(1 << 12)).astype(np.uint32)
    assert _check(A, B, M, dev) > (1 << 19)


def test_u64_keys_mixed_bins_and_offsets_u64(dev):
    # mostly unique keys plus one hot key (3000 copies a side): light bins and
    # one tiled bin, 8-byte offsets
    A = wl.uniform_keys(7, 200000, 8)
    B = wl.probe_keys(A, 8, 150000, 0.3, 8)
    A[:3000] = np.uint64(12345678901)
    B[-3000:] = np.uint64(12345678901)
    _check(A, B, 1 << 17, dev, ob=8)


@pytest.mark.parametrize("nA,nB,M", [(0, 1000, 512), (1000, 0, 512), (0, 0, 1), (300, 200, 1)])
def test_empty_and_single_bin_tables(nA, nB, M, dev):
    # an empty side joins to nothing; one bin holding both sides is one
    # nested loop (300 x 200 key pairs, below the tiled path)
    A = (wl.uniform_keys(17, nA, 4) % np.uint32(50)).astype(np.uint32)
    B = (wl.uniform_keys(18, nB, 4) % np.uint32(50)).astype(np.uint32)
    cnt = _check(A, B, M, dev)
    assert (cnt == 0) == (nA == 0 or nB == 0)


def test_short_capacity_and_compare_limit(dev):
    import hashgraph_amd as hg
    A = wl.uniform_keys(9, 1 << 16, 4)
    B = wl.probe_keys(A, 10, 1 << 16, 0.5, 4)
    oa, ka, ia = hg.build_indexed(_to_dev(A, dev), 1 << 16)
    ob, kb, ib = hg.build_indexed(_to_dev(B, dev), 1 << 16)
    cnt, _, _ = hg.intersect_indexed(oa, ka, ia, ob, kb, ib, 1 << 16, emit=False)
    assert cnt > 1
    with pytest.raises(hg.HGError) as e:
        hg.intersect_indexed(oa, ka, ia, ob, kb, ib, 1 << 16, capacity=cnt - 1)
    assert e.value.status == 4
    # one bin of 2^19 keys a side: 2^38 key pairs to compare, refused before
    # any tile runs; the count is still answered
    A = wl.uniform_keys(11, 1 << 19, 4)
    B = wl.probe_keys(A, 12, 1 << 19, 0.5, 4)
    oa, ka, ia = hg.build_indexed(_to_dev(A, dev), 1)
    ob, kb, ib = hg.build_indexed(_to_dev(B, dev), 1)
    cnt, _, _ = hg.intersect_indexed(oa, ka, ia, ob, kb, ib, 1, emit=False)
    assert cnt == len(_expected_pairs(A, B))
    with pytest.raises(hg.HGError) as e:
        hg.intersect_indexed(oa, ka, ia, ob, kb, ib, 1)
    assert e.value.status == 6


def test_short_capacity_with_tiled_bins_reports_exact_count(dev):
    # bins above 4096 key pairs go through the tiled compare; a buffer one pair
    # short still gets the exact count with HG_ERR_CAPACITY, and a buffer of
    # the exact count the full pair set
    import hashgraph_amd as hg
    A = (wl.uniform_keys(15, 1 << 15, 4) % np.uint32(1 << 10)).astype(np.uint32)
    B = (wl.uniform_keys(16, 1 << 15, 4) % np.uint32(1 << 10)).astype(np.uint32)
    M = 16
    oa, ka, ia = hg.build_indexed(_to_dev(A, dev), M)
    ob, kb, ib = hg.build_indexed(_to_dev(B, dev), M)
    exp = _expected_pairs(A, B)
    import ctypes
    import torch
    c = hg.context(0)
    dA, dB = hg._desc(ka.numel(), M, 4, 4), hg._desc(kb.numel(), M, 4, 4)
    short = len(exp) - 1
    pa = torch.empty(short, dtype=torch.int64, device=dev)
    pb = torch.empty(short, dtype=torch.int64, device=dev)
    got_cnt = ctypes.c_uint64(0)
    st = c.lib.hg_intersect_indexed(c.bind_stream(), ctypes.byref(dA), hg._p(oa), hg._p(ka), hg._p(ia), ctypes.byref(dB),
                                    hg._p(ob), hg._p(kb), hg._p(ib), hg._p(pa), hg._p(pb), short,
                                    ctypes.byref(got_cnt))
    assert st == 4 and got_cnt.value == len(exp)
    cnt, pa, pb = hg.intersect_indexed(oa, ka, ia, ob, kb, ib, M, capacity=len(exp))
    assert cnt == len(exp)
    got = np.sort((pa.cpu().numpy().astype(np.uint64) << np.uint64(32)) | pb.cpu().numpy().astype(np.uint64))
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("kb,n", [(4, 1 << 28), (8, 1 << 26)])
def test_large_shape_properties(kb, n, dev):
    """The row-id join at its benchmark shape (2^28 + 2^28 uint32 keys, 2^28
    bins, B 25% drawn from A; uint64 keys at 2^26), checked on the device
    without hashing: every pair's two keys are equal, no pair repeats, and
    the count equals the sum over values of cA(v) * cB(v) by sort-merge.
    Together these fix the pair set to {(i, j) : A[i] == B[j]}."""
    import torch
    import hashgraph_amd as hg
    M = n
    print(f"row-id join kb={kb} n={n} M={M}", flush=True)
    A = wl.uniform_keys(1, n, kb, device=dev)
    B = wl.probe_keys(A, 2, n, 0.25, kb, device=dev)
    ta = hg.build_indexed(A, M)
    tb = hg.build_indexed(B, M)
    cnt, pa, pb = hg.intersect_indexed(*ta, *tb, M)
    del ta, tb
    assert torch.equal(A[pa], B[pb]), "a pair joins different keys"
    ks = torch.sort(pa * n + pb)[0]
    assert not bool((ks[1:] == ks[:-1]).any()), "a pair repeats"
    del ks, pa, pb
    # uint32 keys as non-negative int64; uint64 keys keep their int64 view
    # (equality is all the merge needs)
    wa = A.to(torch.int64) & 0xFFFFFFFF if kb == 4 else A
    wb = B.to(torch.int64) & 0xFFFFFFFF if kb == 4 else B
    ua, ca = torch.unique(wa, return_counts=True)
    ub, cb = torch.unique(wb, return_counts=True)
    del wa, wb
    pos = torch.searchsorted(ua, ub).clamp(max=ua.numel() - 1)
    hit = ua[pos] == ub
    expect = int((ca[pos][hit].to(torch.int64) * cb[hit].to(torch.int64)).sum())
    assert cnt == expect and cnt > n // 5
    del A, B, ua, ca, ub, cb, pos, hit
    torch.cuda.empty_cache()
