"""Learned hash (RMI) on the GPU: model evaluation parity with the oracle,
training against a CPU restatement of its sampling and np.polyfit (a library
least-squares routine), the sampled fallback figures against exact CPU
counts, and the CSR / intersection of learned-hash tables (the intersection
must equal murmur's: the join does not depend on the hash)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from hashgraph_amd import workloads as wl
from oracle import oracle as o

HDR = np.dtype([("key_min", "<u8"), ("rs", "<f4"), ("ri", "<f4"), ("L", "<u4"), ("M", "<u4"), ("kind", "<u4"),
                ("pad", "<u4")])
LEAF = np.dtype([("x0", "<u8"), ("slope", "<f4"), ("anchor", "<u4")])


def _to_dev(a, dev):
    import torch
    return torch.from_numpy(a.view(np.int32 if a.dtype == np.uint32 else np.int64).copy()).to(dev)


def _model_dev(m, dev):
    import torch
    return torch.from_numpy(m.copy()).to(dev)


def _parse(model_bytes):
    h = model_bytes[:32].view(HDR)[0]
    lv = model_bytes[32:].view(LEAF)
    return h, lv


def _sample_cpu(keys, n_sample, seed):
    pos = wl._np_hi64(wl.stream_np(seed, n_sample), len(keys)).astype(np.int64)
    return keys[pos]


def test_table_bins_match_oracle_u64(dev):
    import hashgraph_amd as hg
    rng = np.random.default_rng(5)
    L, M = 64, 1 << 20
    x0 = np.sort(rng.integers(0, 1 << 63, L, dtype=np.uint64))
    base = np.sort(rng.integers(0, M, L))
    leaves = [(int(x0[i]), float(rng.random() * 1e-12), int(base[i])) for i in range(L)]
    m = o.rmi_model(int(x0[0]), float(L / 2.0 ** 63), 0.25, leaves, M)
    mlog = o.rmi_model(int(x0[0]), 2.0 ** -23 * 3.1, -400.3, leaves, M, root_kind=1)
    keys = rng.integers(0, 1 << 64, 1 << 18, dtype=np.uint64)
    keys[:8] = [0, 1, int(x0[0]) - 1, int(x0[0]), (1 << 64) - 1, (1 << 63), int(x0[5]), int(x0[5]) + 1]
    print("table_bins rmi u64", flush=True)
    got = hg.table_bins(_to_dev(keys, dev), M, hg.HG_HASH_RMI, _model_dev(m, dev)).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, o.bins_rmi(keys, m, M))
    got = hg.table_bins(_to_dev(keys, dev), M, hg.HG_HASH_RMI, _model_dev(mlog, dev)).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, o.bins_rmi(keys, mlog, M))


def test_table_bins_extremes_match_oracle(dev):
    """Model values far outside their ranges: huge and negative slopes,
    offsets beyond 2^32, anchors at M-1, a negative root slope and a root that
    overflows to infinity; the evaluation's clamps must equal the oracle's."""
    import hashgraph_amd as hg
    rng = np.random.default_rng(8)
    L, M = 16, (1 << 32) - 5
    x0 = np.sort(rng.integers(0, 1 << 64, L, dtype=np.uint64))
    slopes = [1e30, -3.0, 1e-3, 2.5e9, 0.0, 1e-19, 7.0, -1e30]
    leaves = [(int(x0[i]), slopes[i % 8], int(rng.integers(0, M))) for i in range(L)]
    leaves[3] = (int(x0[3]), 1e30, M - 1)
    leaves[5] = (int(x0[5]), 1e30, 0)
    models = [o.rmi_model(int(x0[0]), -L / 2.0 ** 64, L + 0.5, leaves, M),
              o.rmi_model(0, 3e38, -1e38, leaves, M, root_kind=1),
              o.rmi_model(int(x0[2]), 2.0 ** -60, -1.5, leaves, M)]
    keys = np.concatenate([rng.integers(0, 1 << 64, 1 << 16, dtype=np.uint64), x0, x0 - np.uint64(1),
                           x0 + np.uint64(1), np.array([0, (1 << 64) - 1], dtype=np.uint64)])
    print("table_bins rmi extremes", flush=True)
    for m in models:
        got = hg.table_bins(_to_dev(keys, dev), M, hg.HG_HASH_RMI, _model_dev(m, dev)).cpu().numpy().view(np.uint32)
        assert np.array_equal(got, o.bins_rmi(keys, m, M))


def test_table_bins_match_oracle_u32(dev):
    import hashgraph_amd as hg
    M = 1000
    leaves = [(1024 * l, 0.25, 250 * l) for l in range(4)]
    m = o.rmi_model(0, 1.0 / 1024, 0.0, leaves, M)
    keys = np.arange(0, 1 << 16, dtype=np.uint32)
    got = hg.table_bins(_to_dev(keys, dev), M, hg.HG_HASH_RMI, _model_dev(m, dev)).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, o.bins_rmi(keys, m, M))


@pytest.mark.parametrize("kind", ["uniform64", "lognormal", "clusters", "uniform32"])
def test_train_against_cpu_fit(dev, kind):
    import hashgraph_amd as hg
    n, ns, M, L, seed = 1 << 20, 1 << 15, 1 << 20, 1024, 9
    if kind == "uniform64":
        A = wl.uniform_keys(3, n, 8)
    elif kind == "uniform32":
        A = wl.uniform_keys(3, n, 4)
    elif kind == "lognormal":
        A = wl.lognormal_keys(3, n)
    else:
        A = wl.cluster_keys(3, n)
    print(f"rmi train {kind}", flush=True)
    model, use, rep = hg.rmi_train(_to_dev(A, dev), M, n_sample=ns, seed=seed, n_leaves=L, report=True)
    mb = model.cpu().numpy()
    h, lv = _parse(mb)
    S = np.sort(_sample_cpu(A, ns, seed)).astype(np.uint64)
    assert h["key_min"] == S[0] and h["L"] == L and h["M"] == M
    # root: four candidates (raw / log feature, least squares / range fit),
    # the one with the smallest sum of squared leaf loads is kept
    y = (np.arange(ns, dtype=np.float64) + 0.5) * (L / ns)
    raw = (S - S[0]).astype(np.float32)
    feats = [raw, (raw + np.float32(1.0)).view(np.uint32).astype(np.float32)]
    cands = []
    for f, F in enumerate(feats):
        Fd = F.astype(np.float64)
        b1, b0 = np.polyfit(Fd, y, 1)
        cands.append((f, max(b1, 0.0), b0 if b1 > 0 else y.mean()))
        sr = L * (1 - 1e-6) / (Fd.max() - Fd.min())
        cands.append((f, sr, -sr * Fd.min()))
    scores = []
    for f, a_, b_ in cands:
        t = np.float32(np.float32(a_) * feats[f]) + np.float32(b_)
        lf = np.clip(np.floor(t), 0, L - 1).astype(np.int64)
        scores.append((np.bincount(lf, minlength=L).astype(np.float64) ** 2).sum())
    ok = [c for c, sc in zip(cands, scores) if sc <= min(scores) * (1 + 1e-3)]
    assert any(h["kind"] == f and np.isclose(h["rs"], a_, rtol=1e-4) and abs(h["ri"] - b_) <= 1e-3 * L
               for f, a_, b_ in ok), (h, cands, scores)
    # routing of the sorted sample under the stored FP32 root
    leaf = np.array([o.rmi_leaf(mb, int(k)) for k in S])
    assert (np.diff(leaf) >= 0).all()
    start = np.searchsorted(leaf, np.arange(L + 1), side="left")
    yb = (np.arange(ns, dtype=np.float64) + 0.5) * (M / ns)
    checked = 0
    for l in range(L):
        a, b = start[l], start[l + 1]
        x0 = S[min(a, ns - 1)]
        assert lv[l]["x0"] == x0
        sparse = b - a < 8 or S[b - 1] == x0
        lo, hi = (max(a - 1, 0) if sparse else a), (min(b + 1, ns) if sparse else b)
        xs = S[lo:hi]
        d = np.where(xs >= x0, (xs - x0).astype(np.float64), -(x0 - xs).astype(np.float64))
        if hi - lo >= 2 and d.max() > d.min():
            s1, alpha = np.polyfit(d, yb[lo:hi], 1)
            if s1 <= 0:
                s1, alpha = 0.0, yb[lo:hi].mean()
        else:
            s1, alpha = 0.0, (yb[lo:hi].mean() if hi > lo else (a + 0.5) * (M / ns))
        if s1 > 0:
            assert np.isclose(lv[l]["slope"], s1, rtol=1e-4, atol=0), l
            checked += 1
        else:
            assert lv[l]["slope"] == 0
        want = min(max(np.floor(alpha + 0.5), 0), M - 1)
        assert abs(int(lv[l]["anchor"]) - int(want)) <= 1, l
    assert checked > 0
    # the fallback figures, recomputed exactly on the CPU
    Sm = _sample_cpu(A, ns, seed)
    Mp = max(1, min(M, int(M * ns / n + 0.5)))
    br = o.bins_rmi(Sm, mb, M).astype(np.uint64) * np.uint64(Mp) // np.uint64(M)
    bm = o.bins(Sm, M).astype(np.uint64) * np.uint64(Mp) // np.uint64(M)
    cr = np.bincount(br.astype(np.int64), minlength=Mp)
    cm = np.bincount(bm.astype(np.int64), minlength=Mp)
    assert rep["sample_bins"] == Mp
    assert rep["max_bin_rmi"] == cr.max() and rep["max_bin_murmur"] == cm.max()
    assert np.isclose(rep["mean_probe_rmi"], (cr.astype(np.float64) ** 2).sum() / ns)
    assert np.isclose(rep["mean_probe_murmur"], (cm.astype(np.float64) ** 2).sum() / ns)
    assert use == (cr.max() <= cm.max() and (cr ** 2).sum() <= (cm ** 2).sum())


@pytest.mark.parametrize("kind,kb", [("clusters", 8), ("lognormal", 8), ("uniform", 4), ("uniform", 8)])
def test_build_and_probe_learned(dev, kind, kb):
    """CSR under the learned hash equals the oracle's CSR from the same bins,
    and the probe result equals murmur's and the oracle's bag join."""
    import hashgraph_amd as hg
    n, M = 1 << 20, 1 << 20
    if kind == "clusters":
        A = wl.cluster_keys(4, n)
        B = wl.mixed_probe(A, wl.cluster_keys(40, n), 5, 0.5)
    elif kind == "lognormal":
        A = wl.lognormal_keys(4, n)
        B = wl.mixed_probe(A, wl.lognormal_keys(40, n), 5, 0.5)
    else:
        A = wl.uniform_keys(4, n, kb)
        B = wl.probe_keys(A, 5, n, 0.5, kb)
    print(f"rmi build/probe {kind} kb={kb}", flush=True)
    dA, dB = _to_dev(A, dev), _to_dev(B, dev)
    model, use = hg.rmi_train(dA, M, n_sample=n // 100, seed=1)
    mb = model.cpu().numpy()
    off, csr = hg.build(dA, M, hash_kind=hg.HG_HASH_RMI, rmi_model=model)
    eoff, ekeys = o.build_csr_bins(A, o.bins_rmi(A, mb, M), M)
    offn = off.cpu().numpy().astype(np.uint64)
    assert np.array_equal(offn, eoff)
    got = csr.cpu().numpy().view(A.dtype).copy()
    o.canon(eoff, got)
    o.canon(eoff, ekeys)
    assert np.array_equal(got, ekeys)
    cnt, out = hg.probe(off, csr, M, dB, hash_kind=hg.HG_HASH_RMI, rmi_model=model)
    offm, csrm = hg.build(dA, M)
    cntm, outm = hg.probe(offm, csrm, M, dB)
    assert cnt == cntm == o.bag_intersect_count(A, B)
    assert np.array_equal(np.sort(out.cpu().numpy().view(A.dtype)), np.sort(outm.cpu().numpy().view(A.dtype)))


@pytest.mark.parametrize("kb", [4, 8])
def test_batch_training_equals_single_training(dev, kb):
    """hg_rmi_train_batch (one workgroup per table, everything in LDS) fits
    the models hg_rmi_train fits, byte for byte, with the same fallback
    decisions and sampled figures: uniform, Zipf and clustered tables, tiny
    tables, an empty table, a leaf count that leaves most leaves empty, and a
    sample too large for one workgroup (trained by the multi-kernel path)."""
    import torch
    import hashgraph_amd as hg
    rng = np.random.default_rng(90 + kb)
    parts, ns = [], []
    sizes = [1, 3, 0, 500, 5000, 40000, 200000, 1 << 20]
    for i, n in enumerate(sizes):
        if i % 3 == 0:
            a = wl.uniform_keys(300 + i, n, kb)
        elif i % 3 == 1:
            a = (wl.zipf_keys(300 + i, n, 1.2) & np.uint64(0xFFFFFFFF if kb == 4 else (1 << 64) - 1)).astype(
                np.uint32 if kb == 4 else np.uint64)
        else:
            c = rng.integers(0, 2 ** (8 * kb - 1), 8, dtype=np.uint64)
            a = (c[rng.integers(0, 8, n)] + rng.integers(0, 1 << 20, n, dtype=np.uint64)).astype(
                np.uint32 if kb == 4 else np.uint64)
        parts.append(a)
        ns.append(max(1, min(n, max(n // 100, 1024))) if n else 1)
    ns[-1] = 20000  # beyond one workgroup for either key width
    keys = np.concatenate(parts)
    ko = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n_bins = [max(n, 1) for n in sizes]
    seeds = [7 + 13 * t for t in range(len(sizes))]
    L = 256
    dk = _to_dev(keys, dev)
    print(f"rmi batch kb={kb} tables={len(sizes)}", flush=True)
    models, use, reps = hg.rmi_train_batch(dk, ko, n_bins, ns, seeds, n_leaves=L, report=True)
    mb = hg.rmi_model_bytes(L)
    got = models.cpu().numpy()
    for t, n in enumerate(sizes):
        if n == 0:
            assert not use[t]
            continue
        m, u, r = hg.rmi_train(dk[int(ko[t]):int(ko[t + 1])], n_bins[t], n_sample=ns[t], seed=seeds[t], n_leaves=L,
                               report=True)
        assert np.array_equal(got[t * mb:(t + 1) * mb], m.cpu().numpy()), f"table {t} model"
        assert use[t] == u, f"table {t} decision"
        assert reps[t] == r, f"table {t} report"
    torch.cuda.synchronize()
