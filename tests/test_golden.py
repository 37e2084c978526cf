"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must still reproduce them (CPU), and the GPU path must match them
(offsets identical, bins equal as sorted multisets, counts and sorted lists
identical, run-length records identical)."""
import os

import numpy as np
import pytest

from oracle import oracle as o

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(G, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _to_dev(a, dev):
    import torch
    return torch.from_numpy(a.view(np.int32 if a.dtype == np.uint32 else np.int64).copy()).to(dev)


@pytest.mark.parametrize("name", ["c1_u32.npz", "c2_u64.npz"])
def test_oracle_reproduces_join_fixture(name):
    f = _load(name)
    M = int(f["M"])
    off, csr = o.build_csr(f["A"], M)
    o.canon(off, csr)
    assert np.array_equal(off, f["offA"]) and np.array_equal(csr, f["csrA"])
    assert o.bag_intersect_count(f["A"], f["B"]) == int(f["count"])
    offB, csrB = o.build_csr(f["B"], M)
    cnt, lst = o.intersect(off, csr, offB, csrB)
    assert cnt == int(f["count"]) and np.array_equal(np.sort(lst), f["matches"])


def test_oracle_reproduces_multi_fixture():
    f = _load("multi.npz")
    M, P = int(f["M"]), int(f["P"])
    splits, _ = o.find_splits(o.bins(f["keys"], M), M, 16000, P)
    assert np.array_equal(splits, f["splits"])
    off, _ = o.build_multi(f["keys"], M, splits)
    assert np.array_equal(off, f["off"])


def test_oracle_reproduces_rle_fixture():
    f = _load("rle.npz")
    M = int(f["M"])
    offA, csrA = o.build_csr(f["A"], M)
    offB, csrB = o.build_csr(f["B"], M)
    o.canon(offA, csrA)
    o.canon(offB, csrB)
    cnt, rk, ca, cb = o.intersect_rle(offA, csrA, offB, csrB)
    order = np.argsort(rk, kind="stable")
    assert cnt == int(f["count"])
    assert np.array_equal(rk[order], f["rk"]) and np.array_equal(ca[order], f["ca"]) and \
        np.array_equal(cb[order], f["cb"])


@pytest.mark.gpu
@pytest.mark.parametrize("name,ob", [("c1_u32.npz", 4), ("c2_u64.npz", 4), ("c2_u64.npz", 8)])
def test_gpu_matches_join_fixture(dev, name, ob):
    import hashgraph_amd as hg
    f = _load(name)
    M = int(f["M"])
    A, B = f["A"], f["B"]
    print(f"golden {name} ob={ob}", flush=True)
    off, csr = hg.build(_to_dev(A, dev), M, offset_bytes=ob)
    goff = off.cpu().numpy().view(np.uint32 if ob == 4 else np.uint64).astype(np.uint64)
    assert np.array_equal(goff, f["offA"])
    g = csr.cpu().numpy().view(A.dtype).copy()
    o.canon(goff, g)
    assert np.array_equal(g, f["csrA"])
    cnt, out = hg.probe(off, csr, M, _to_dev(B, dev))
    assert cnt == int(f["count"])
    assert np.array_equal(np.sort(out.cpu().numpy().view(A.dtype)), f["matches"])
    offB, csrB = hg.build(_to_dev(B, dev), M, offset_bytes=ob)
    cnt2, out2 = hg.intersect(off, csr, offB, csrB, M)
    assert cnt2 == int(f["count"])
    assert np.array_equal(np.sort(out2.cpu().numpy().view(A.dtype)), f["matches"])


@pytest.mark.gpu
def test_gpu_matches_multi_fixture(dev):
    import hashgraph_amd as hg
    f = _load("multi.npz")
    M, P = int(f["M"]), int(f["P"])
    dk = _to_dev(f["keys"], dev)
    splits = hg.find_splits(dk, M, 16000, P)
    assert splits == f["splits"].tolist()
    off, csr, po = hg.build_partitioned(dk, M, splits, offset_bytes=8)
    goff = off.cpu().numpy().view(np.uint64)
    assert np.array_equal(goff, f["off"])
    g = csr.cpu().numpy().view(np.uint64).copy()
    for p in range(P):
        lo, hi = splits[p], splits[p + 1]
        loc = goff[lo + p: hi + p + 1]
        seg = g[po[p]: po[p + 1]].copy()
        o.canon(loc, seg)
        assert np.array_equal(seg, f["csr"][po[p]: po[p + 1]]), p


@pytest.mark.gpu
def test_gpu_matches_rle_fixture(dev):
    import hashgraph_amd as hg
    f = _load("rle.npz")
    M = int(f["M"])
    offA, csrA = hg.build(_to_dev(f["A"], dev), M)
    offB, csrB = hg.build(_to_dev(f["B"], dev), M)
    cnt, rk, ca, cb = hg.intersect_rle(offA, csrA, offB, csrB, M)
    assert cnt == int(f["count"])
    rk = rk.cpu().numpy().view(np.uint64)
    order = np.argsort(rk, kind="stable")
    assert np.array_equal(rk[order], f["rk"])
    assert np.array_equal(ca.cpu().numpy()[order].astype(np.uint64), f["ca"].astype(np.uint64))
    assert np.array_equal(cb.cpu().numpy()[order].astype(np.uint64), f["cb"].astype(np.uint64))
