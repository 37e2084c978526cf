"""Every configuration at its full shape on the GPU, checked through
size-independent properties computed on the device (the oracle cannot run
at these sizes): every CSR key lies in the bin whose offset range holds it,
the key multiset is preserved, degrees equal the bincount of the keys' bins,
join counts equal a hash-free sort-merge count, and run-length records equal
cA(v) * cB(v) per value.  The 2^28 uint32 build and probe are covered in
test_gpu_build.py / test_gpu_probe.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from hashgraph_amd import workloads as wl


def _value_counts(x):
    import torch
    s = torch.sort(x)[0]
    return torch.unique_consecutive(s, return_counts=True)


def _join_expect(A, B):
    """(values present on both sides, cA, cB) by sort-merge, no hashing."""
    import torch
    ua, ca = _value_counts(A)
    ub, cb = _value_counts(B)
    pos = torch.searchsorted(ua, ub).clamp(max=ua.numel() - 1)
    hit = ua[pos] == ub
    return ub[hit], ca[pos][hit].to(torch.int64), cb[hit].to(torch.int64)


def _check_csr(keys, off, csr, bins_of_keys, bins_of_csr, M):
    """bins_of_* are int64 tensors of global bins; off has M+1 entries."""
    import torch
    deg = (off[1:].to(torch.int64) - off[:-1].to(torch.int64))
    assert int(off[0]) == 0 and int(off[-1]) == keys.numel()
    assert torch.equal(deg, torch.bincount(bins_of_keys, minlength=M))
    owner = torch.repeat_interleave(torch.arange(M, device=off.device), deg)
    assert torch.equal(owner, bins_of_csr), "a key sits outside its bin's offset range"
    del owner
    assert torch.equal(torch.sort(csr)[0], torch.sort(keys)[0]), "key multiset changed"


def _bins(keys, M, hk=0, model=None):
    import hashgraph_amd as hg
    import torch
    b = hg.table_bins(keys, M, hash_kind=hk, rmi_model=model)
    return b.to(torch.int64) & 0xFFFFFFFF


def test_c2_u64_2e27_build_and_probe(dev):
    """C2 uint64: A, B = 2^27 keys uniform over [0, 2^64), 25% overlap, 2^27 bins."""
    import torch
    import hashgraph_amd as hg
    n = M = 1 << 27
    print("full shape c2-u64 2^27", flush=True)
    A = wl.uniform_keys(1, n, 8, device=dev)
    B = wl.probe_keys(A, 2, n, 0.25, 8, device=dev)
    off, csr = hg.build(A, M)
    _check_csr(A, off, csr, _bins(A, M), _bins(csr, M), M)
    cnt, out = hg.probe(off, csr, M, B)
    v, ca, cb = _join_expect(A, B)
    want = ca * cb
    assert cnt == int(want.sum())
    uo, co = _value_counts(out)
    assert torch.equal(uo, v) and torch.equal(co.to(torch.int64), want)
    # the join of two built tables gives the same count
    offB, csrB = hg.build(B, M)
    cnt2, _ = hg.intersect(off, csr, offB, csrB, M, emit=False)
    assert cnt2 == cnt


def test_c3_2e30_zipf_partitioned_and_rle(dev):
    """C3: A = 2^30 uint64 Zipf(1.1) keys, 2^30 bins, 16 partitions (hot-key
    bins of ~10^8 keys through the oversized-bucket path); B = 2^29 with A's
    splits; per-partition run-length join against the sort-merge records."""
    import torch
    import hashgraph_amd as hg
    nA, nB, M, P = 1 << 30, 1 << 29, 1 << 30, 16
    print("full shape c3 2^30 zipf", flush=True)
    A = wl.zipf_keys(3, nA, 1.1, device=dev)
    sp = hg.find_splits(A, M, 16000, P)
    offA, csrA, poA = hg.build_partitioned(A, M, sp)
    assert poA[-1] == nA
    # global bins of every key and of every CSR position
    bA = _bins(A, M)
    bC = _bins(csrA, M)
    for p in range(P):
        lo, hi = sp[p], sp[p + 1]
        o = offA[lo + p: hi + p + 1].to(torch.int64)
        keys_p = csrA[poA[p]: poA[p + 1]]
        assert int(o[0]) == 0 and int(o[-1]) == keys_p.numel()
        deg = o[1:] - o[:-1]
        sel = (bA >= lo) & (bA < hi)
        assert torch.equal(deg, torch.bincount(bA[sel] - lo, minlength=hi - lo))
        owner = torch.repeat_interleave(torch.arange(lo, hi, device=dev), deg)
        assert torch.equal(owner, bC[poA[p]: poA[p + 1]])
        del owner, sel
    del bA, bC
    assert torch.equal(torch.sort(csrA)[0], torch.sort(A)[0])
    del A
    torch.cuda.empty_cache()
    B = wl.zipf_keys(4, nB, 1.1, device=dev)
    offB, csrB, poB = hg.build_partitioned(B, M, sp)
    total, keys, cas, cbs = 0, [], [], []
    for p in range(P):
        lo, hi = sp[p], sp[p + 1]
        if hi == lo:
            continue
        cnt, rk, ca, cb = hg.intersect_rle(offA[lo + p: hi + p + 1], csrA[poA[p]: poA[p + 1]],
                                           offB[lo + p: hi + p + 1], csrB[poB[p]: poB[p + 1]], hi - lo)
        total += cnt
        keys.append(rk)
        cas.append(ca)
        cbs.append(cb)
    rk = torch.cat(keys)
    ca = torch.cat(cas)
    cb = torch.cat(cbs)
    order = torch.argsort(rk)
    v, eca, ecb = _join_expect(csrA, B)
    assert torch.equal(rk[order], v)
    assert torch.equal(ca[order].to(torch.int64), eca) and torch.equal(cb[order].to(torch.int64), ecb)
    assert total == int((eca * ecb).sum())


@pytest.mark.parametrize("kind", ["lognormal", "clusters"])
def test_c4_2e27_learned_equals_murmur(dev, kind):
    """C4: 2^27 uint64 skewed keys; the learned-hash table's CSR is a valid
    table under its model, and its probe gives murmur's count and list."""
    import torch
    import hashgraph_amd as hg
    n = M = 1 << 27
    print(f"full shape c4 {kind} 2^27", flush=True)
    gen = wl.lognormal_keys if kind == "lognormal" else wl.cluster_keys
    A = gen(21, n, device=dev)
    B = wl.mixed_probe(A, gen(22, n, device=dev), 23, 0.5)
    model, use = hg.rmi_train(A, M, n_sample=n // 100, seed=24)
    off, csr = hg.build(A, M, hash_kind=hg.HG_HASH_RMI, rmi_model=model)
    _check_csr(A, off, csr, _bins(A, M, 1, model), _bins(csr, M, 1, model), M)
    cnt, out = hg.probe(off, csr, M, B, hash_kind=hg.HG_HASH_RMI, rmi_model=model)
    del off, csr
    offm, csrm = hg.build(A, M)
    cntm, outm = hg.probe(offm, csrm, M, B)
    assert cnt == cntm
    assert torch.equal(torch.sort(out)[0], torch.sort(outm)[0])
    v, ca, cb = _join_expect(A, B)
    assert cnt == int((ca * cb).sum())


def test_c5_4096_tables_plan_equals_separate_calls(dev):
    """C5: 4096 tables (2^8..2^20 uint32 keys, half Zipf(1.2)), learned hash
    where its sampled check wins, one autotuned captured plan: every table's
    CSR is valid and the per-table counts equal separate single-table calls."""
    import torch
    import hashgraph_amd as hg
    T, L = 4096, 256
    print("full shape c5 4096 tables", flush=True)
    keys, ko, probe, po = wl.batch_tables(5, T, 8.0, 20.0, 1.2, 0.3, device=dev)
    n_bins = [max(int(ko[t + 1] - ko[t]), 1) for t in range(T)]
    mb = hg.rmi_model_bytes(L)
    models = torch.zeros(T * mb, dtype=torch.uint8, device=dev)
    kinds = []
    for t in range(T):
        n = int(ko[t + 1] - ko[t])
        m, use = hg.rmi_train(keys[int(ko[t]):int(ko[t + 1])], n_bins[t], n_sample=max(n // 100, min(n, 1024)),
                              seed=t, n_leaves=L)
        models[t * mb:(t + 1) * mb] = m
        kinds.append(hg.HG_HASH_RMI if use else hg.HG_HASH_MURMUR3)
    assert 0 < sum(kinds) < T
    plan = hg.BatchPlan(keys, ko, n_bins, probe=probe, probe_off=po, hash_kinds=kinds, rmi_models=models,
                        rmi_leaves=L, autotune=True)
    plan.run()
    plan.run()
    torch.cuda.synchronize()
    counts = plan.counts.cpu().tolist()
    offs = plan.offsets
    for t in range(T):
        a, b = int(ko[t]), int(ko[t + 1])
        mdl = models[t * mb:(t + 1) * mb] if kinds[t] else None
        off, csr = hg.build(keys[a:b], n_bins[t], hash_kind=kinds[t], rmi_model=mdl)
        k, _ = hg.probe(off, csr, n_bins[t], probe[a:b], emit=False, hash_kind=kinds[t], rmi_model=mdl)
        assert counts[t] == k, f"table {t}"
        o0 = plan.off_off[t]
        assert torch.equal(offs[o0: o0 + n_bins[t] + 1], off), f"table {t} offsets"
        if t % 64 == 0:
            bk = _bins(keys[a:b], n_bins[t], kinds[t], mdl)
            bc = _bins(plan.csr[a:b], n_bins[t], kinds[t], mdl)
            _check_csr(keys[a:b], off, plan.csr[a:b], bk, bc, n_bins[t])
    plan.close()
