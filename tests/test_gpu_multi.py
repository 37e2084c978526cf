"""GPU parity of the one-GPU MultiHashGraph path: split finding, partitioned
build and the run-length join, against the oracle (the reference's
buildSingle / CPU-intersect checks, src/MultiHashGraph.cu:742-862,
test/MultiHashGraphTest.cu:340-425)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from hashgraph_amd import workloads as wl
from oracle import oracle as o


def _to_dev(a, dev):
    import torch
    return torch.from_numpy(a.view(np.int32 if a.dtype == np.uint32 else np.int64).copy()).to(dev)


@pytest.mark.parametrize("n,M,P,kb", [(100000, 80000, 16, 8), (1 << 20, 1 << 20, 16, 4), (5000, 5000, 3, 4),
                                      (200000, 1 << 18, 1, 8)])
def test_find_splits(n, M, P, kb, dev):
    import hashgraph_amd as hg
    keys = wl.uniform_keys(n, n, kb)
    got = hg.find_splits(_to_dev(keys, dev), M, 16000, P)
    exp, _ = o.find_splits(o.bins(keys, M), M, 16000, P)
    assert got == exp.tolist()


def _check_partitioned(keys, M, P, dev, ob=4):
    import hashgraph_amd as hg
    print(f"partitioned n={len(keys)} M={M} P={P} kb={keys.itemsize} ob={ob}", flush=True)
    dk = _to_dev(keys, dev)
    splits = hg.find_splits(dk, M, 16000, P)
    off, csr, po = hg.build_partitioned(dk, M, splits, offset_bytes=ob)
    eoff, ecsr = o.build_multi(keys, M, np.array(splits, dtype=np.int64))
    goff = off.cpu().numpy().view(np.uint32 if ob == 4 else np.uint64).astype(np.uint64)
    assert np.array_equal(goff, eoff), "per-partition offsets differ"
    gcsr = csr.cpu().numpy().view(keys.dtype).copy()
    # canonicalise each partition's bins and compare
    for p in range(P):
        lo, hi = splits[p], splits[p + 1]
        loc = eoff[lo + p: hi + p + 1]
        a = gcsr[po[p]: po[p + 1]].copy()
        b = ecsr[po[p]: po[p + 1]].copy()
        o.canon(loc, a)
        o.canon(loc, b)
        assert np.array_equal(a, b), f"partition {p} bins differ"
    return splits


def test_partitioned_uniform(dev):
    keys = wl.uniform_keys(21, 1 << 20, 8)
    _check_partitioned(keys, 1 << 20, 16, dev)


def test_partitioned_zipf(dev):
    keys = wl.zipf_keys(3, 1 << 20, 1.1)
    _check_partitioned(keys, 1 << 20, 16, dev, ob=8)


def test_partitioned_u32_few_parts(dev):
    keys = wl.uniform_keys(22, 300000, 4)
    _check_partitioned(keys, 250000, 3, dev)


@pytest.mark.parametrize("kb", [4, 8])
def test_partitioned_single_partition(kb, dev):
    # one partition over every bin keeps every key: built as a plain table
    keys = wl.uniform_keys(23, 300000, kb)
    _check_partitioned(keys, 250000, 1, dev)


def _rle_expected(A, B, M):
    oa, ka = o.build_csr(A, M)
    ob, kb = o.build_csr(B, M)
    o.canon(oa, ka)
    o.canon(ob, kb)
    return o.intersect_rle(oa, ka, ob, kb)


def _check_rle(A, B, M, dev):
    import hashgraph_amd as hg
    print(f"rle nA={len(A)} nB={len(B)} M={M} kb={A.itemsize}", flush=True)
    offA, csrA = hg.build(_to_dev(A, dev), M)
    offB, csrB = hg.build(_to_dev(B, dev), M)
    cnt, rk, ca, cb = hg.intersect_rle(offA, csrA, offB, csrB, M)
    ecnt, ek, eca, ecb = _rle_expected(A, B, M)
    assert cnt == ecnt
    rk = rk.cpu().numpy().view(A.dtype).astype(np.uint64)
    order = np.argsort(rk, kind="stable")
    eorder = np.argsort(ek, kind="stable")
    assert np.array_equal(rk[order], ek[eorder])
    assert np.array_equal(ca.cpu().numpy()[order].astype(np.uint64), eca[eorder])
    assert np.array_equal(cb.cpu().numpy()[order].astype(np.uint64), ecb[eorder])


def test_rle_uniform(dev):
    A = wl.uniform_keys(1, 1 << 18, 4)
    B = wl.probe_keys(A, 2, 1 << 18, 0.5, 4)
    _check_rle(A, B, 1 << 18, dev)


def test_rle_dense_bins_in_sub_ranges_u64(dev):
    # ~2.3 keys per bin: a chunk's keys exceed the staging buffer and are
    # walked in staged sub-ranges of its bins
    A = wl.uniform_keys(41, 300000, 8)
    B = wl.probe_keys(A, 42, 250000, 0.5, 8)
    _check_rle(A, B, 1 << 17, dev)


def test_rle_very_dense_bins_from_global_u64(dev):
    # ~6 keys per bin: no sub-range fits the staging buffer, keys are read
    # from global memory
    A = wl.uniform_keys(43, 300000, 8)
    B = wl.probe_keys(A, 44, 300000, 0.7, 8)
    _check_rle(A, B, 50000, dev)


def test_rle_zipf_heavy(dev):
    # C3's shape at 2^20: the top key holds ~10% of the keys (heavy bins)
    A = wl.zipf_keys(3, 1 << 20, 1.1)
    B = wl.zipf_keys(4, 1 << 19, 1.1)
    _check_rle(A, B, 1 << 20, dev)


def test_rle_all_ones_key(dev):
    # the all-ones key is the heavy table's sentinel: counted apart
    A = np.concatenate([np.full(5000, 2**64 - 1, dtype=np.uint64), wl.uniform_keys(5, 3000, 8)])
    B = np.concatenate([np.full(300, 2**64 - 1, dtype=np.uint64), A[5000:5100]])
    _check_rle(A, B, 1024, dev)


def test_rle_count_matches_bag_count(dev):
    A = wl.zipf_keys(7, 300000, 1.2)
    B = wl.zipf_keys(8, 200000, 1.2)
    import hashgraph_amd as hg
    M = 1 << 17
    offA, csrA = hg.build(_to_dev(A, dev), M)
    offB, csrB = hg.build(_to_dev(B, dev), M)
    cnt, *_ = hg.intersect_rle(offA, csrA, offB, csrB, M)
    assert cnt == o.bag_intersect_count(A, B)


def test_rle_giant_single_value_bin(dev):
    # one key holds 300000 of A's and 280000 of B's keys: its bin exceeds
    # 2^18 keys and is folded by the multi-workgroup path
    rng = np.random.default_rng(31)
    hot = np.uint64(0x1234_5678_9ABC_DEF0)
    rest = wl.uniform_keys(32, 100000, 8)
    A = np.concatenate([np.full(300000, hot, dtype=np.uint64), rest])
    B = np.concatenate([np.full(280000, hot, dtype=np.uint64), rest[rng.integers(0, len(rest), 50000)]])
    rng.shuffle(A)
    rng.shuffle(B)
    _check_rle(A, B, 1 << 16, dev)


def test_rle_giant_bins_many_values(dev):
    # two bins of ~400000 keys each, ~600 distinct values (with the all-ones
    # key) spread over many chunks: per-chunk tables merged across chunks
    rng = np.random.default_rng(33)
    vals = np.concatenate([wl.uniform_keys(34, 600, 8), np.array([2**64 - 1], dtype=np.uint64)])
    A = vals[rng.integers(0, len(vals), 800000)]
    B = vals[rng.integers(0, len(vals) - 100, 500000)]
    _check_rle(A, B, 2, dev)


def test_rle_giant_u32(dev):
    rng = np.random.default_rng(35)
    vals = wl.uniform_keys(36, 300, 4)
    A = vals[rng.integers(0, len(vals), 600000)]
    B = np.concatenate([vals[rng.integers(0, len(vals), 200000)], wl.uniform_keys(37, 1000, 4)])
    _check_rle(A, B, 3, dev)


def test_rle_value_rich_heavy_bins_u64(dev):
    """Two bins of ~6000 keys of A with ~5000 distinct values each: more
    than the heavy-bin LDS table holds, so they take the global-table path;
    records equal the oracle's, one per value."""
    vals = wl.uniform_keys(51, 10000, 8)
    A = np.concatenate([vals, vals[:2000], np.full(20, 2**64 - 1, dtype=np.uint64)])
    B = np.concatenate([vals[::2], vals[:100], wl.uniform_keys(52, 3000, 8), np.full(7, 2**64 - 1, dtype=np.uint64)])
    _check_rle(A, B, 2, dev)


def test_rle_value_rich_giant_bin(dev):
    """A giant bin (over 2^18 keys) with 120000 distinct values: its chunk
    tables overflow, so the bin is redone in a global table; beside it a
    giant single-value bin stays on the giant path."""
    rng = np.random.default_rng(53)
    vals = wl.uniform_keys(54, 120000, 4)
    A = vals[rng.integers(0, len(vals), 400000)]
    B = np.concatenate([vals[rng.integers(0, len(vals), 150000)], wl.uniform_keys(55, 20000, 4)])
    _check_rle(A, B, 1, dev)
    hot = np.uint32(77)
    A2 = np.concatenate([A, np.full(300000, hot, dtype=np.uint32)])
    B2 = np.concatenate([B, np.full(1000, hot, dtype=np.uint32)])
    _check_rle(A2, B2, 3, dev)


def test_rle_short_capacity_reports_exact_counts(dev):
    """A run-length join with one record less than it needs (the single-pass
    form, which packs records after writing them at padded places): the call
    returns HG_ERR_CAPACITY with the exact record and pair counts, and every
    record it wrote is one of the true records."""
    import ctypes
    import torch
    import hashgraph_amd as hg
    M = 1 << 18
    A = wl.zipf_keys(5, 1 << 18, 1.1)
    B = wl.zipf_keys(6, 1 << 17, 1.1)
    oa, ka = hg.build(_to_dev(A, dev), M)
    ob, kb = hg.build(_to_dev(B, dev), M)
    cnt, rk, ca, cb = hg.intersect_rle(oa, ka, ob, kb, M)
    n = rk.numel()
    assert n > 1 and cnt == o.bag_intersect_count(A, B)
    c = hg.context(0)
    dA = hg._desc(ka.numel(), M, 8, 4)
    dB = hg._desc(kb.numel(), M, 8, 4)
    cap = n - 1
    rk2 = torch.zeros(cap, dtype=torch.int64, device=dev)
    ca2 = torch.zeros(cap, dtype=torch.int64, device=dev)
    cb2 = torch.zeros(cap, dtype=torch.int64, device=dev)
    nrec, tot = ctypes.c_uint64(0), ctypes.c_uint64(0)
    st = c.lib.hg_intersect_rle(c.bind_stream(), ctypes.byref(dA), hg._p(oa), hg._p(ka), ctypes.byref(dB), hg._p(ob),
                                hg._p(kb), hg._p(rk2), hg._p(ca2), hg._p(cb2), cap, ctypes.byref(nrec),
                                ctypes.byref(tot))
    assert st == 4
    assert nrec.value == n and tot.value == cnt
    true = set(zip(rk.tolist(), ca.tolist(), cb.tolist()))
    got = set(zip(rk2.tolist(), ca2.tolist(), cb2.tolist()))
    assert got <= true and len(got) == cap  # every slot holds a distinct true record
