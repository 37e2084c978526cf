"""Pins of the CPU oracle (oracle/hg_oracle.c) against published vectors,
closed forms, brute force and the reference's own invariants.  CPU only."""
import struct

import numpy as np
import pytest

from oracle import oracle as o
from hashgraph_amd import workloads as wl

# Published MurmurHash3_x86_32 test vectors (SMHasher lineage).
MURMUR_VECTORS = [
    (b"", 0, 0x00000000),
    (b"", 1, 0x514E28B7),
    (b"", 0xFFFFFFFF, 0x81F16F39),
    (b"\x00\x00\x00\x00", 0, 0x2362F9DE),
    (b"\xff\xff\xff\xff", 0, 0x76293B50),
    (b"\x21\x43\x65\x87", 0, 0xF55B516B),
    (b"\x21\x43\x65\x87", 0x5082EDEE, 0x2362F9DE),
    (b"\x21\x43\x65", 0, 0x7E4A8634),
    (b"\x21\x43", 0, 0xA0F7B07A),
    (b"\x21", 0, 0x72661CF4),
    (b"aaaa", 0x9747B28C, 0x5A97808A),
    (b"aaa", 0x9747B28C, 0x283E0130),
    (b"aa", 0x9747B28C, 0x5D211726),
    (b"a", 0x9747B28C, 0x7FA09EA6),
    (b"abcd", 0x9747B28C, 0xF0478627),
    (b"Hello, world!", 0x9747B28C, 0x24884CBA),
    (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD),
]


@pytest.mark.parametrize("data,seed,expect", MURMUR_VECTORS)
def test_murmur_published_vectors(data, seed, expect):
    assert o.murmur3_x86_32(data, seed) == expect


def test_hash_key_is_murmur_of_low_four_bytes():
    # the reference hashes sizeof(int) bytes of the key object, seed 0
    rng = np.random.default_rng(5)
    for k in rng.integers(0, 2**63, size=200, dtype=np.uint64).tolist() + [0, 1, 2**32 - 1, 2**32, 2**64 - 1]:
        lo = k & 0xFFFFFFFF
        assert o.hash_key(k) == o.murmur3_x86_32(struct.pack("<I", lo), 0)
        assert o.hash_key(k) == o.hash_key(lo)  # high 32 bits never enter the hash


def test_bins_vectorised_match_scalar():
    keys = wl.uniform_keys(3, 5000, 8)
    for M in (1, 7, 1000, 2**20, 2**32 - 5):
        b = o.bins(keys, M)
        exp = np.array([o.hash_key(int(k)) % M for k in keys[:300]], dtype=np.uint32)
        assert np.array_equal(b[:300], exp)
        assert int(b.max()) < M


def _canonical(keys, M):
    """CSR from an independent formulation: numpy bincount + lexsort by (bin, key)."""
    bins = o.bins(keys, M).astype(np.int64)
    off = np.zeros(M + 1, dtype=np.uint64)
    off[1:] = np.cumsum(np.bincount(bins, minlength=M))
    order = np.lexsort((keys, bins))
    return off, keys[order]


@pytest.mark.parametrize("n,M,kb", [(0, 5, 4), (1, 1, 4), (1000, 1, 4), (5000, 97, 4), (20000, 20000, 8),
                                    (3000, 100000, 4), (2**16, 2**12, 8)])
def test_build_csr_closed_form(n, M, kb):
    keys = wl.uniform_keys(11, n, kb)
    off, csr = o.build_csr(keys, M)
    eoff, ecsr = _canonical(keys, M)
    assert np.array_equal(off, eoff)
    assert off[M] == n
    o.canon(off, csr)
    assert np.array_equal(csr, ecsr)


def test_build_csr_brute_force_tiny():
    keys = np.array([5, 9, 5, 0, 123456789, 9, 2**32 - 1], dtype=np.uint32)
    M = 3
    off, csr = o.build_csr(keys, M)
    for b in range(M):
        mine = sorted(int(k) for k in keys if o.hash_key(int(k)) % M == b)
        assert off[b + 1] - off[b] == len(mine)
        assert sorted(csr[off[b]:off[b + 1]].tolist()) == mine


@pytest.mark.parametrize("lrb_bins,threads", [(16000, 4), (1, 1), (7, 3), (100000, 2)])
def test_lrb_build_equals_plain_build(lrb_bins, threads):
    # the LRB two-level flow (src/SingleHashGraph.cu:129-181) gives the same CSR
    # as the plain build for any number of coarse bins
    keys = wl.uniform_keys(4, 30000, 4)
    M = 25000
    off0, csr0 = o.build_csr(keys, M)
    off1, csr1 = o.build_lrb(keys, M, lrb_bins, threads)
    assert np.array_equal(off0, off1)
    o.canon(off0, csr0)
    o.canon(off1, csr1)
    assert np.array_equal(csr0, csr1)


def _brute_intersect(A, B):
    eq = A[:, None] == B[None, :]
    i, j = np.nonzero(eq)
    return int(eq.sum()), np.sort(B[j])


@pytest.mark.parametrize("kb", [4, 8])
def test_intersect_brute_force(kb):
    rng = np.random.default_rng(1)
    dt = np.uint32 if kb == 4 else np.uint64
    A = rng.integers(0, 60, size=700).astype(dt)
    B = rng.integers(20, 90, size=500).astype(dt)
    if kb == 8:  # same low 32 bits, different high bits: same bin, never equal
        A[:50] |= dt(1 << 40)
    M = 37
    oa, ka = o.build_csr(A, M)
    ob, kb_ = o.build_csr(B, M)
    cnt, lst = o.intersect(oa, ka, ob, kb_)
    bc, bl = _brute_intersect(A, B)
    assert cnt == bc
    assert np.array_equal(np.sort(lst), bl)
    assert o.bag_intersect_count(A, B) == bc


@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_intersect_par_equals_serial_and_brute_force(kb, threads):
    """The threaded join of the CPU baseline: same count as the brute force,
    and the same list, in the same order, as the serial join (bins split
    across threads, each thread writing at its prefix)."""
    rng = np.random.default_rng(7 + threads)
    dt = np.uint32 if kb == 4 else np.uint64
    A = rng.integers(0, 400, size=3000).astype(dt)
    B = rng.integers(100, 600, size=2500).astype(dt)
    M = 101
    oa, ka = o.build_lrb(A, M, 16000, threads)
    ob, kb_ = o.build_lrb(B, M, 16000, threads)
    cnt, lst = o.intersect(oa, ka, ob, kb_)
    assert o.intersect_par(oa, ka, ob, kb_, threads) == cnt
    out = np.empty(cnt, dtype=dt)
    assert o.intersect_par(oa, ka, ob, kb_, threads, out) == cnt
    assert np.array_equal(out, lst)
    bc, bl = _brute_intersect(A, B)
    assert cnt == bc
    assert np.array_equal(np.sort(out), bl)
    # a capacity below the count: the count is still exact, the list is cut
    short = np.zeros(cnt // 2, dtype=dt)
    assert o.intersect_par(oa, ka, ob, kb_, threads, short) == cnt
    assert np.array_equal(short, lst[:cnt // 2])


def test_intersect_sequential_closed_form():
    # the reference driver's inputs: A = 0..nA-1, B = 0..nB-1 -> {0..min-1}
    nA, nB = 5000, 3000
    A = wl.sequential_keys(nA, 8)
    B = wl.sequential_keys(nB, 8)
    M = nA
    oa, ka = o.build_csr(A, M)
    ob, kb_ = o.build_csr(B, M)
    cnt, lst = o.intersect(oa, ka, ob, kb_)
    assert cnt == min(nA, nB)
    assert np.array_equal(np.sort(lst), np.arange(min(nA, nB), dtype=np.uint64))


def test_intersect_rle_counts():
    rng = np.random.default_rng(2)
    A = rng.integers(0, 200, size=4000).astype(np.uint64)
    B = rng.integers(100, 300, size=3000).astype(np.uint64)
    M = 64
    oa, ka = o.build_csr(A, M)
    ob, kb_ = o.build_csr(B, M)
    o.canon(oa, ka)
    o.canon(ob, kb_)
    cnt, keys, ca, cb = o.intersect_rle(oa, ka, ob, kb_)
    ua, na = np.unique(A, return_counts=True)
    ub, nb = np.unique(B, return_counts=True)
    common, ia, ib = np.intersect1d(ua, ub, return_indices=True)
    assert cnt == int((na[ia] * nb[ib]).sum())
    order = np.argsort(keys)
    assert np.array_equal(keys[order], common)
    assert np.array_equal(ca[order], na[ia])
    assert np.array_equal(cb[order], nb[ib])
    assert cnt == o.intersect(oa, ka, ob, kb_, emit=False)[0]


def test_find_splits_hand_example():
    # M = 100 bins, binCount = 10 -> binRange = 10; bins chosen by hand
    bins = np.array([0, 1, 2, 15, 16, 33, 35, 36, 37, 38, 71, 99], dtype=np.uint32)
    # coarse counts: [3,2,0,5,0,0,0,1,0,1]; prefix ps = [0,3,5,5,10,10,10,10,11,11,12]
    # avg = ceil(12/3) = 4: upper 4 -> search 4 in ps -> insertion index 2 -> 20
    #                       upper 8 -> index 4 -> 40;  upper 12 -> ps[10] == 12 -> 100
    splits, ps = o.find_splits(bins, 100, 10, 3)
    assert ps.tolist() == [0, 3, 5, 5, 10, 10, 10, 10, 11, 11, 12]
    assert splits.tolist() == [0, 20, 40, 100]


def test_find_splits_invariants():
    keys = wl.uniform_keys(7, 50000, 8)
    M = 40000
    b = o.bins(keys, M)
    for P in (1, 2, 5, 16):
        splits, ps = o.find_splits(b, M, 16000, P)
        rng_ = int(np.ceil(np.float32(M) / np.float32(16000)))
        assert splits[0] == 0 and np.all(np.diff(splits) >= 0)
        assert int(b.max()) < splits[-1]  # every key lands in some partition
        for s in splits[1:]:
            assert s == M or s % rng_ == 0
        avg = int(np.ceil(np.float32(len(keys)) / np.float32(P)))
        for t in range(1, P + 1):  # each split covers at least avg*t keys
            assert int((b < splits[t]).sum()) >= min(avg * t, len(keys))


@pytest.mark.parametrize("P", [1, 3, 16])
def test_multi_build_equals_single(P):
    # the reference's buildSingle check (src/MultiHashGraph.cu:796-857)
    keys = wl.uniform_keys(9, 40000, 4)
    M = 30000
    b = o.bins(keys, M)
    splits, _ = o.find_splits(b, M, 16000, P)
    off, csr = o.build_multi(keys, M, splits)
    goff, gcsr = o.build_csr(keys, M)
    o.canon(goff, gcsr)
    kept = 0
    for p in range(P):
        lo, hi = int(splits[p]), int(splits[p + 1])
        loc = off[lo + p: hi + p + 1].astype(np.int64)
        assert loc[0] == 0
        part_keys = csr[kept: kept + int(loc[-1])]
        for j in range(hi - lo):
            g0, g1 = int(goff[lo + j]), int(goff[lo + j + 1])
            assert loc[j + 1] - loc[j] == g1 - g0
            assert np.array_equal(np.sort(part_keys[loc[j]:loc[j + 1]]), gcsr[g0:g1])
        kept += int(loc[-1])
    assert kept == len(keys)
