"""GPU probe / join parity: hg_probe and hg_intersect against the oracle's
bin-wise bag-product join (src/MultiHashGraph.cu:631-740) and the reference
driver's hash-free check (test/MultiHashGraphTest.cu:369-397)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from hashgraph_amd import workloads as wl
from oracle import oracle as o


def _to_dev(a, dev):
    import torch
    return torch.from_numpy(a.view(np.int32 if a.dtype == np.uint32 else np.int64).copy()).to(dev)


def _expected(A, B, M):
    oa, ka = o.build_csr(A, M)
    ob, kb = o.build_csr(B, M)
    cnt, lst = o.intersect(oa, ka, ob, kb)
    return cnt, np.sort(lst)


def _check_probe(A, B, M, dev, ob=4):
    import hashgraph_amd as hg
    print(f"probe nA={len(A)} nB={len(B)} M={M} kb={A.itemsize} ob={ob}", flush=True)
    dA = _to_dev(A, dev)
    off, csr = hg.build(dA, M, offset_bytes=ob)
    dB = _to_dev(B, dev)
    cnt_only, _ = hg.probe(off, csr, M, dB, emit=False)
    cnt, out = hg.probe(off, csr, M, dB)
    ecnt, elst = _expected(A, B, M)
    assert cnt_only == ecnt
    assert cnt == ecnt
    assert np.array_equal(np.sort(out.cpu().numpy().view(A.dtype)), elst)
    assert cnt == o.bag_intersect_count(A, B)


def _check_join(A, B, M, dev, ob=4):
    import hashgraph_amd as hg
    print(f"join nA={len(A)} nB={len(B)} M={M} kb={A.itemsize} ob={ob}", flush=True)
    offA, csrA = hg.build(_to_dev(A, dev), M, offset_bytes=ob)
    offB, csrB = hg.build(_to_dev(B, dev), M, offset_bytes=ob)
    cnt, out = hg.intersect(offA, csrA, offB, csrB, M)
    ecnt, elst = _expected(A, B, M)
    assert cnt == ecnt
    assert np.array_equal(np.sort(out.cpu().numpy().view(A.dtype)), elst)


def test_c1_probe(dev):
    # A = 2^20 uint32 (seed 1), B = 2^20 with 50% drawn from A (seed 2), 2^20 bins
    A = wl.uniform_keys(1, 1 << 20, 4)
    B = wl.probe_keys(A, 2, 1 << 20, 0.5, 4)
    _check_probe(A, B, 1 << 20, dev)


def test_c1_join(dev):
    A = wl.uniform_keys(1, 1 << 20, 4)
    B = wl.probe_keys(A, 2, 1 << 20, 0.5, 4)
    _check_join(A, B, 1 << 20, dev)


@pytest.mark.parametrize("nA,nB,M,kb,ob", [
    (1, 1, 1, 4, 4), (1000, 5000, 97, 4, 4), (50000, 20000, 1 << 20, 8, 8), (300000, 100000, 299993, 4, 4),
    (1 << 18, 1 << 19, 1 << 18, 8, 4), (20000, 200000, 5000, 4, 8),
])
def test_probe_shapes(nA, nB, M, kb, ob, dev):
    A = wl.uniform_keys(nA + 1, nA, kb)
    B = wl.probe_keys(A, nB + 2, nB, 0.3, kb)
    _check_probe(A, B, M, dev, ob)
    _check_join(A, B, M, dev, ob)


@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("dup", [1, 2])
def test_self_join_match_buffer(kb, dup, dev):
    # A joined with itself over one key a bin: ~2048 matches in every 2048-bin
    # workgroup range, within the 3072 an emitting workgroup buffers in LDS;
    # with every value twice on each side, ~4096, so every range takes the
    # counting re-walk
    vals = wl.uniform_keys(13, (1 << 18) // dup, kb)
    A = np.repeat(vals, dup)
    _check_join(A, A[::-1].copy(), 1 << 18, dev)


def test_duplicates_on_both_sides(dev):
    # bag product: a value with cA copies in A and cB in B contributes cA*cB
    vals = wl.uniform_keys(3, 3000, 4)
    A = vals[(wl.stream_np(4, 20000) % np.uint64(3000)).astype(np.int64)]
    B = vals[(wl.stream_np(5, 15000) % np.uint64(6000)).astype(np.int64) % 3000]
    _check_probe(A, B, 1 << 12, dev)
    _check_join(A, B, 1 << 12, dev)


def test_sequential_closed_form(dev):
    import hashgraph_amd as hg
    nA, nB = 1 << 20, 1 << 19
    A = wl.sequential_keys(nA, 8)
    B = wl.sequential_keys(nB, 8)
    off, csr = hg.build(_to_dev(A, dev), nA, offset_bytes=8)
    cnt, out = hg.probe(off, csr, nA, _to_dev(B, dev))
    assert cnt == nB
    assert np.array_equal(np.sort(out.cpu().numpy().view(np.uint64)), np.arange(nB, dtype=np.uint64))


def test_empty_inputs(dev):
    import torch
    import hashgraph_amd as hg
    A = wl.uniform_keys(1, 1000, 4)
    off, csr = hg.build(_to_dev(A, dev), 512)
    empty = torch.empty(0, dtype=torch.int32, device=dev)
    assert hg.probe(off, csr, 512, empty)[0] == 0
    eoff, ecsr = hg.build(empty, 512)
    assert hg.probe(eoff, ecsr, 512, _to_dev(A, dev))[0] == 0
    assert hg.intersect(off, csr, eoff, ecsr, 512)[0] == 0


def test_heavy_bins_are_answered(dev):
    """Bins beyond the walk limit (one value repeated 100k times) are answered
    from a per-bin value table: count and list equal the bag join."""
    import hashgraph_amd as hg
    rest = wl.uniform_keys(5, 20000, 4)
    A = np.concatenate([np.full(100000, 12345, dtype=np.uint32), rest])
    B = np.concatenate([np.full(1000, 12345, dtype=np.uint32), rest[:3000], wl.uniform_keys(6, 3000, 4)])
    off, csr = hg.build(_to_dev(A, dev), 1024)
    cnt, _ = hg.probe(off, csr, 1024, _to_dev(B, dev), emit=False)
    assert cnt == o.bag_intersect_count(A, B)
    # emission, small enough to list
    A2 = np.concatenate([np.full(5000, 7, dtype=np.uint32), np.full(4500, 0xFFFFFFFF, dtype=np.uint32), rest[:100]])
    B2 = np.concatenate([np.full(3, 7, dtype=np.uint32), np.full(2, 0xFFFFFFFF, dtype=np.uint32), rest[:50]])
    _check_probe(A2, B2, 64, dev)


def test_many_long_bins_answered_in_groups(dev):
    """12000 values with 3 copies each in 512 bins: every bin is beyond the
    walk limit, and the bucket's long bins hold more distinct values than one
    value table, so they are answered in several groups."""
    rng = np.random.default_rng(41)
    vals = wl.uniform_keys(42, 12000, 4)
    A = np.repeat(vals, 3)
    rng.shuffle(A)
    B = np.concatenate([vals[rng.integers(0, len(vals), 5000)], wl.uniform_keys(43, 2000, 4)])
    _check_probe(A, B, 512, dev)


def test_oversized_bucket_of_long_and_mid_bins(dev):
    """One oversized bucket (M = 2^20, so a bucket spans 65535 bins) holding
    1500 long bins of 40 keys among its first 8192 bins -- more long bins than
    the heavy kernel lists at once -- and 600 bins of 20 copies further on,
    which the oversized-bucket walk reads in groups (B's side of the bucket
    is oversized too); the rest of the table is sparse.  Count and list
    equal the oracle's bag join."""
    M = 1 << 20
    cand = wl.uniform_keys(61, 400000, 4)
    bins = o.bins(cand, M)
    _, first = np.unique(bins, return_index=True)  # one value per bin
    cand, bins = cand[first], bins[first]
    long_v = cand[bins < 8192][:1500]
    mid_v = cand[(bins >= 8192) & (bins < 65535)][:600]
    assert len(long_v) == 1500 and len(mid_v) == 600
    rest = cand[bins >= 65535][:20000]
    A = np.concatenate([np.repeat(long_v, 40), np.repeat(mid_v, 20), rest])
    rng = np.random.default_rng(62)
    rng.shuffle(A)
    # B's part of the bucket (38000 keys) is beyond one probe workgroup's capacity
    B = np.concatenate([long_v[rng.integers(0, 1500, 30000)], mid_v[rng.integers(0, 600, 8000)], rest[:5000],
                        wl.uniform_keys(63, 4000, 4)])
    _check_probe(A, B, M, dev)


def test_heavy_bin_with_many_values(dev):
    """One bin (M = 1) of 10000 distinct values, some repeated: beyond the
    walk limit and beyond one value table, so the bin is folded over several
    table rounds; probe and join equal the oracle's bag join."""
    base = wl.uniform_keys(1, 10000, 4)
    A = np.concatenate([base, base[:700], np.full(40, 0xFFFFFFFF, dtype=np.uint32)])
    B = np.concatenate([base[:1000], base[5000:5600], wl.uniform_keys(2, 800, 4),
                        np.full(3, 0xFFFFFFFF, dtype=np.uint32)])
    _check_probe(A, B, 1, dev)
    _check_join(A, B, 1, dev)


def test_value_rich_bins_u64(dev):
    """Bins of ~6000 keys with ~5000 distinct values each (M = 4): every join
    bin pair is above the nested loop's pair limit and every probe bin above
    the walk limit; both take several table rounds per bin."""
    vals = wl.uniform_keys(3, 20000, 8)
    A = np.concatenate([vals, vals[:4000]])
    B = np.concatenate([vals[::3], wl.uniform_keys(4, 5000, 8)])
    _check_probe(A, B, 4, dev, ob=8)
    _check_join(A, B, 4, dev, ob=8)


def test_join_bin_above_2e22_pairs(dev):
    """A hot value with 3000 copies in A and 2000 in B (6e6 pairs in one bin,
    above 2^22): the join answers it from a value table; count equals the
    bag count, and a smaller hot bin's emitted list equals the oracle's."""
    import hashgraph_amd as hg
    rest = wl.uniform_keys(5, 30000, 4)
    A = np.concatenate([np.full(3000, 987654321, dtype=np.uint32), rest])
    B = np.concatenate([np.full(2000, 987654321, dtype=np.uint32), rest[:10000], wl.uniform_keys(6, 5000, 4)])
    M = 1 << 14
    offA, csrA = hg.build(_to_dev(A, dev), M)
    offB, csrB = hg.build(_to_dev(B, dev), M)
    cnt, _ = hg.intersect(offA, csrA, offB, csrB, M, emit=False)
    assert cnt == o.bag_intersect_count(A, B)
    A2 = np.concatenate([np.full(150, 42, dtype=np.uint32), rest[:2000]])
    B2 = np.concatenate([np.full(90, 42, dtype=np.uint32), rest[:500]])
    _check_join(A2, B2, 1 << 10, dev)


def test_mismatched_tables_rejected(dev):
    import hashgraph_amd as hg
    A = wl.uniform_keys(1, 1000, 4)
    offA, csrA = hg.build(_to_dev(A, dev), 512)
    offB, csrB = hg.build(_to_dev(A, dev), 256)
    with pytest.raises(hg.HGError) as e:
        import ctypes
        from hashgraph_amd import _lib
        c = hg.context(0)
        dA = hg._desc(1000, 512, 4)
        dB = hg._desc(1000, 256, 4)
        cnt = ctypes.c_uint64()
        _lib.check(c.lib.hg_intersect(c.bind_stream(), ctypes.byref(dA), hg._p(offA), hg._p(csrA), ctypes.byref(dB),
                                      hg._p(offB), hg._p(csrB), None, 0, ctypes.byref(cnt)), "hg_intersect")
    assert e.value.status == 5


def test_full_size_2e28_probe_properties(dev):
    """2^28 uint32 A, B with 25% from A: count equals a device-side sort-merge
    count, every emitted key is in A and in B."""
    import torch
    import hashgraph_amd as hg
    n = M = 1 << 28
    A = wl.uniform_keys(1, n, 4, device=dev)
    B = wl.probe_keys(A, 2, n, 0.25, 4, device=dev)
    off, csr = hg.build(A, M)
    cnt, out = hg.probe(off, csr, M, B)
    del off, csr
    sa = torch.sort(A)[0]
    del A
    sb = torch.sort(B)[0]
    del B
    # hash-free count: for every value, count in A times count in B
    ua, ca = torch.unique_consecutive(sa, return_counts=True)
    ub, cb = torch.unique_consecutive(sb, return_counts=True)
    pos = torch.searchsorted(ua, ub)
    pos = pos.clamp(max=ua.numel() - 1)
    hit = ua[pos] == ub
    expect = int((ca[pos][hit].to(torch.int64) * cb[hit].to(torch.int64)).sum())
    assert cnt == expect
    so = torch.sort(out)[0]
    # emitted multiset: value v appears cA(v)*cB(v) times
    uo, co = torch.unique_consecutive(so, return_counts=True)
    assert torch.equal(uo, ub[hit])
    assert torch.equal(co.to(torch.int64), (ca[pos][hit].to(torch.int64) * cb[hit].to(torch.int64)))


def test_two_level_probe_2e28_u64_properties(dev):
    """2^28 uint64 keys probe a 2^28-bin table: more buckets than one
    partition level holds, so B is partitioned in two levels.  The count
    equals a device-side sort-merge count and the emitted multiset holds
    cA(v) * cB(v) copies of each value."""
    import torch
    import hashgraph_amd as hg
    n = M = 1 << 28
    print("two-level probe 2^28 u64", flush=True)
    A = wl.uniform_keys(11, n, 8, device=dev)
    B = wl.probe_keys(A, 12, n, 0.25, 8, device=dev)
    off, csr = hg.build(A, M)
    cnt, out = hg.probe(off, csr, M, B)
    del off, csr
    sa = torch.sort(A)[0]
    del A
    sb = torch.sort(B)[0]
    del B
    ua, ca = torch.unique_consecutive(sa, return_counts=True)
    ub, cb = torch.unique_consecutive(sb, return_counts=True)
    del sa, sb
    pos = torch.searchsorted(ua, ub).clamp(max=ua.numel() - 1)
    hit = ua[pos] == ub
    want = ca[pos][hit].to(torch.int64) * cb[hit].to(torch.int64)
    assert cnt == int(want.sum())
    uo, co = torch.unique_consecutive(torch.sort(out)[0], return_counts=True)
    assert torch.equal(uo, ub[hit])
    assert torch.equal(co.to(torch.int64), want)


def test_join_2e28_sequential_every_key_matches(dev):
    """The key join at the reference experiment's shape (A = B = 0 .. 2^28-1
    as uint32, 2^28 bins): every key matches once, so every 2048-bin range
    emits about 2048 keys from its LDS match buffer.  Closed form: the count
    is 2^28 and the sorted matched keys are 0 .. 2^28-1."""
    import torch
    import hashgraph_amd as hg
    n = M = 1 << 28
    print(f"sequential join n={n} M={M}", flush=True)
    A = wl.sequential_keys(n, 4, device=dev)
    offA, csrA = hg.build(A, M)
    offB, csrB = hg.build(A.clone(), M)
    cnt, _ = hg.intersect(offA, csrA, offB, csrB, M, emit=False)
    assert cnt == n
    cnt, out = hg.intersect(offA, csrA, offB, csrB, M)
    assert cnt == n
    del offA, csrA, offB, csrB
    assert torch.equal(torch.sort(out)[0], torch.arange(n, dtype=torch.int32, device=dev))
    del out, A
    torch.cuda.empty_cache()
