"""Batched tables on the GPU: every table of a batch (eager, and replayed
from a captured plan) equals the oracle's single-table CSR, and the per-table
probe counts equal the oracle's bag joins.  The batch has no counterpart in
the reference; each table is pinned by the same oracle as a single build."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from hashgraph_amd import workloads as wl
from oracle import oracle as o


def _to_dev(a, dev):
    import torch
    return torch.from_numpy(a.view(np.int32 if a.dtype == np.uint32 else np.int64).copy()).to(dev)


def _batch(seed, T, lo, hi):
    keys, ko, probe, po = wl.batch_tables(seed, T, lo, hi)
    # a few edge tables: empty, single key, one bin
    return keys, ko, probe, po


def _check_tables(keys, ko, n_bins, off_off, off, csr, bins_fn=None):
    offn = off.cpu().numpy().view(np.uint32).astype(np.uint64)
    csrn = csr.cpu().numpy().view(np.uint32)
    for t in range(len(n_bins)):
        a = keys[ko[t]:ko[t + 1]]
        M = int(n_bins[t])
        if bins_fn is None:
            eoff, ekeys = o.build_csr(a, M)
        else:
            eoff, ekeys = o.build_csr_bins(a, bins_fn(t, a, M), M)
        goff = offn[off_off[t]: off_off[t] + M + 1]
        assert np.array_equal(goff, eoff), f"table {t} offsets"
        g = csrn[ko[t]:ko[t + 1]].copy()
        o.canon(eoff, g)
        o.canon(eoff, ekeys)
        assert np.array_equal(g, ekeys), f"table {t} keys"


def _expected_counts(keys, ko, probe, po):
    return [o.bag_intersect_count(keys[ko[t]:ko[t + 1]], probe[po[t]:po[t + 1]]) if ko[t + 1] > ko[t] and
            po[t + 1] > po[t] else 0 for t in range(len(ko) - 1)]


@pytest.mark.parametrize("T,lo,hi", [(64, 4, 14), (300, 0, 10)])
def test_batch_build_and_probe(dev, T, lo, hi):
    import hashgraph_amd as hg
    keys, ko, probe, po = _batch(11, T, lo, hi)
    n_bins = np.maximum(np.diff(ko), 1)
    print(f"batch T={T} keys={len(keys)}", flush=True)
    dk, dp = _to_dev(keys, dev), _to_dev(probe, dev)
    off, csr, off_off = hg.batch_build(dk, ko, n_bins)
    _check_tables(keys, ko, n_bins, off_off, off, csr)
    total, counts = hg.batch_probe(off, csr, ko, n_bins, off_off, dp, po)
    exp = _expected_counts(keys, ko, probe, po)
    assert counts.cpu().numpy().tolist() == exp
    assert total == sum(exp)


def test_batch_with_empty_and_tiny_tables(dev):
    import hashgraph_amd as hg
    rng = np.random.default_rng(2)
    sizes = [0, 1, 2, 0, 5000, 1, 70000, 0]
    parts = [wl.uniform_keys(100 + i, s, 4) for i, s in enumerate(sizes)]
    keys = np.concatenate(parts)
    ko = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n_bins = np.array([1, 1, 7, 3, 4096, 1, 65536, 2])
    probe = keys[rng.integers(0, len(keys), len(keys))]
    po = ko.copy()
    dk, dp = _to_dev(keys, dev), _to_dev(probe, dev)
    off, csr, off_off = hg.batch_build(dk, ko, n_bins)
    _check_tables(keys, ko, n_bins, off_off, off, csr)
    total, counts = hg.batch_probe(off, csr, ko, n_bins, off_off, dp, po)
    assert counts.cpu().numpy().tolist() == _expected_counts(keys, ko, probe, po)


def test_plan_replays_match_eager(dev):
    """Captured plan (autotuned): several replays, with eager library work in
    between, each equal to the oracle."""
    import torch
    import hashgraph_amd as hg
    keys, ko, probe, po = _batch(12, 200, 6, 14)
    n_bins = np.maximum(np.diff(ko), 1)
    dk, dp = _to_dev(keys, dev), _to_dev(probe, dev)
    exp = _expected_counts(keys, ko, probe, po)
    plan = hg.BatchPlan(dk, ko, n_bins, probe=dp, probe_off=po, autotune=True)
    info = plan.info()
    assert info["candidates"] == 6 and info["tile_keys"] in (8192, 32768, 131072)
    for r in range(3):
        plan.run()
        torch.cuda.synchronize()
        _check_tables(keys, ko, n_bins, plan.off_off, plan.offsets, plan.csr)
        assert plan.counts.cpu().numpy().tolist() == exp, f"replay {r}"
        # unrelated eager work between replays
        hg.build(dk[:1000], 512)
        plan.offsets.fill_(-1)
        plan.counts.fill_(-7)
        torch.cuda.synchronize()
    plan.close()


def test_batch_learned_tables(dev):
    """Tables with the learned hash (trained per table) beside murmur tables:
    CSRs equal the oracle's from the same bins, counts equal murmur's."""
    import torch
    import hashgraph_amd as hg
    keys, ko, probe, po = _batch(13, 32, 8, 14)
    n_bins = np.maximum(np.diff(ko), 1)
    T = len(n_bins)
    dk, dp = _to_dev(keys, dev), _to_dev(probe, dev)
    L = 256
    mb = hg.rmi_model_bytes(L)
    models = torch.zeros(T * mb, dtype=torch.uint8, device=dev)
    kinds = []
    for t in range(T):
        n = int(ko[t + 1] - ko[t])
        m, use = hg.rmi_train(dk[int(ko[t]):int(ko[t + 1])], int(n_bins[t]), n_sample=max(n // 100, min(n, 1024)),
                              seed=t, n_leaves=L)
        models[t * mb:(t + 1) * mb] = m
        kinds.append(hg.HG_HASH_RMI if (use or t % 3 == 0) else hg.HG_HASH_MURMUR3)
    mh = models.cpu().numpy()

    def bins_fn(t, a, M):
        if kinds[t] == hg.HG_HASH_RMI:
            return o.bins_rmi(a, mh[t * mb:(t + 1) * mb], M)
        return o.bins(a, M)

    off, csr, off_off = hg.batch_build(dk, ko, n_bins, hash_kinds=kinds, rmi_models=models, rmi_leaves=L)
    _check_tables(keys, ko, n_bins, off_off, off, csr, bins_fn)
    total, counts = hg.batch_probe(off, csr, ko, n_bins, off_off, dp, po, hash_kinds=kinds, rmi_models=models,
                                   rmi_leaves=L)
    assert counts.cpu().numpy().tolist() == _expected_counts(keys, ko, probe, po)


def test_batch_u64_tables(dev):
    """uint64 tables (uniform and Zipf(1.1) keys, sizes 0..20000): offsets and
    canonical keys equal the oracle's, per-table counts equal the bag joins."""
    import hashgraph_amd as hg
    rng = np.random.default_rng(17)
    T = 40
    sizes = [int(x) for x in rng.integers(0, 20000, T)]
    sizes[3] = 0
    parts = [wl.uniform_keys(200 + t, s, 8) if t % 2 == 0 else wl.zipf_keys(300 + t, s, 1.1) for t, s in enumerate(sizes)]
    keys = np.concatenate(parts) if sum(sizes) else np.zeros(0, dtype=np.uint64)
    ko = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    probe_parts = []
    for t, s in enumerate(sizes):
        a = parts[t]
        half = a[rng.integers(0, max(len(a), 1), s // 2)] if len(a) else np.zeros(0, dtype=np.uint64)
        probe_parts.append(np.concatenate([half, wl.uniform_keys(400 + t, s - s // 2, 8)]))
    probe = np.concatenate(probe_parts)
    po = ko.copy()
    n_bins = np.maximum(np.array(sizes), 1)
    print(f"batch u64 T={T} keys={len(keys)}", flush=True)
    dk, dp = _to_dev(keys, dev), _to_dev(probe, dev)
    off, csr, off_off = hg.batch_build(dk, ko, n_bins)
    offn = off.cpu().numpy().view(np.uint32).astype(np.uint64)
    csrn = csr.cpu().numpy().view(np.uint64)
    for t in range(T):
        a = keys[ko[t]:ko[t + 1]]
        M = int(n_bins[t])
        eoff, ekeys = o.build_csr(a, M)
        goff = offn[off_off[t]: off_off[t] + M + 1]
        assert np.array_equal(goff, eoff), f"table {t} offsets"
        g = csrn[ko[t]:ko[t + 1]].copy()
        o.canon(eoff, g)
        o.canon(eoff, ekeys)
        assert np.array_equal(g, ekeys), f"table {t} keys"
    total, counts = hg.batch_probe(off, csr, ko, n_bins, off_off, dp, po)
    exp = _expected_counts(keys, ko, probe, po)
    assert counts.cpu().numpy().tolist() == exp
    assert total == sum(exp)


def test_plan_replays_value_rich_bins(dev):
    """A batch whose tables have long bins with more distinct values than one
    probe value table (a 5000-value table with 2 bins, an 8000-value table
    with 1 bin): every plan replay's counts equal the oracle's."""
    import torch
    import hashgraph_amd as hg
    sizes = [5000, 300, 8000, 2000]
    parts = [wl.uniform_keys(200 + i, s, 4) for i, s in enumerate(sizes)]
    keys = np.concatenate(parts)
    ko = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n_bins = np.array([2, 300, 1, 2048])
    rng = np.random.default_rng(5)
    probe = np.concatenate([np.concatenate([p[rng.integers(0, len(p), len(p) // 2)],
                                            wl.uniform_keys(300 + i, len(p) - len(p) // 2, 4)])
                            for i, p in enumerate(parts)])
    po = ko.copy()
    dk, dp = _to_dev(keys, dev), _to_dev(probe, dev)
    exp = _expected_counts(keys, ko, probe, po)
    off, csr, off_off = hg.batch_build(dk, ko, n_bins)
    total, counts = hg.batch_probe(off, csr, ko, n_bins, off_off, dp, po)
    assert counts.cpu().numpy().tolist() == exp
    plan = hg.BatchPlan(dk, ko, n_bins, probe=dp, probe_off=po, autotune=False)
    for r in range(3):
        plan.run()
        torch.cuda.synchronize()
        assert plan.counts.cpu().numpy().tolist() == exp, f"replay {r}"
        plan.counts.fill_(-1)
        torch.cuda.synchronize()
    plan.close()


def test_batch_probe_emit_lists(dev):
    """The batch probe with its matches written to per-table regions: every
    table's list equals the oracle's bin join (sorted), including Zipf tables
    (long bins answered from the value table) and a table whose probe set
    piles 40000 copies of one key into an oversized bucket (the bucket is cut
    into tasks over the grid); regions shorter than the counts give
    HG_ERR_CAPACITY with exact counts."""
    import hashgraph_amd as hg
    keys, ko, probe, po = _batch(17, 48, 4, 14)
    hot = np.uint32(123456789)
    ea = np.concatenate([np.full(20, hot, dtype=np.uint32), wl.uniform_keys(71, 20000, 4)])
    eb = np.concatenate([np.full(40000, hot, dtype=np.uint32), ea[5000:9000], wl.uniform_keys(72, 3000, 4)])
    keys = np.concatenate([keys, ea])
    probe = np.concatenate([probe, eb])
    ko = np.append(ko, ko[-1] + len(ea)).astype(np.uint64)
    po = np.append(po, po[-1] + len(eb)).astype(np.uint64)
    T = len(ko) - 1
    n_bins = np.maximum(np.diff(ko), 1)
    print(f"batch emit T={T} keys={len(keys)} probe={len(probe)}", flush=True)
    dk, dp = _to_dev(keys, dev), _to_dev(probe, dev)
    off, csr, off_off = hg.batch_build(dk, ko, n_bins)
    total, counts = hg.batch_probe(off, csr, ko, n_bins, off_off, dp, po)
    cnt = counts.cpu().numpy().astype(np.int64)
    out_off = np.concatenate([[0], np.cumsum(cnt)])
    total2, counts2, out = hg.batch_probe_emit(off, csr, ko, n_bins, off_off, dp, po, out_off)
    assert total2 == total and np.array_equal(counts2.cpu().numpy(), cnt)
    outn = out.cpu().numpy().view(np.uint32)
    for t in range(T):
        a, b, M = keys[ko[t]:ko[t + 1]], probe[po[t]:po[t + 1]], int(n_bins[t])
        oa, ka = o.build_csr(a, M)
        ob, kb = o.build_csr(b, M)
        c, lst = o.intersect(oa, ka, ob, kb)
        assert c == cnt[t], f"table {t} count"
        assert np.array_equal(np.sort(outn[out_off[t]:out_off[t + 1]]), np.sort(lst)), f"table {t} list"
    short = np.concatenate([[0], np.cumsum(cnt // 2)])
    with pytest.raises(hg.HGError):
        hg.batch_probe_emit(off, csr, ko, n_bins, off_off, dp, po, short)
    # the counts of the refused call are still exact
    _, counts3 = hg.batch_probe(off, csr, ko, n_bins, off_off, dp, po)
    assert np.array_equal(counts3.cpu().numpy(), cnt)


def test_plan_trains_models_in_graph(dev):
    """A plan with HG_PLAN_TRAIN fits every table's model inside the graph on
    each replay: after a replay the models equal the batch trainer's (seed t,
    sample max(n/100, min(n, 1024))), each table's CSR equals the oracle's
    under the kind its check chose, and the counts equal the bag joins; after
    the keys change in place, the next replay follows them."""
    import torch
    import hashgraph_amd as hg
    keys, ko, probe, po = _batch(19, 40, 4, 14)
    n_bins = np.maximum(np.diff(ko), 1)
    T = len(n_bins)
    L = 256
    mb = hg.rmi_model_bytes(L)
    dk, dp = _to_dev(keys, dev), _to_dev(probe, dev)
    models = torch.zeros(T * mb, dtype=torch.uint8, device=dev)
    plan = hg.BatchPlan(dk, ko, n_bins, probe=dp, probe_off=po, rmi_models=models, rmi_leaves=L, autotune=False,
                        train=True)
    ns = [max(int(n) // 100, min(int(n), 1024)) for n in np.diff(ko)]
    for rep in range(2):
        if rep == 1:  # new key contents in the same buffers
            keys = keys[::-1].copy() ^ np.uint32(0x5A5A5A5A)
            probe = probe ^ np.uint32(0x5A5A5A5A)
            dk.copy_(_to_dev(keys, dev))
            dp.copy_(_to_dev(probe, dev))
        plan.run()
        torch.cuda.synchronize()
        ref, use = hg.rmi_train_batch(dk, ko, n_bins, ns, list(range(T)), n_leaves=L)
        got = models.cpu().numpy()
        exp = ref.cpu().numpy()
        for t in range(T):
            if ko[t + 1] > ko[t]:
                assert np.array_equal(got[t * mb:(t + 1) * mb], exp[t * mb:(t + 1) * mb]), f"replay {rep} table {t}"
        assert plan.info()["learned_tables"] == sum(use)

        def bins_fn(t, a, M):
            if use[t]:
                return o.bins_rmi(a, got[t * mb:(t + 1) * mb], M)
            return o.bins(a, M)

        _check_tables(keys, ko, n_bins, plan.off_off, plan.offsets, plan.csr, bins_fn)
        assert plan.counts.cpu().numpy().tolist() == _expected_counts(keys, ko, probe, po)
    print(f"learned on {sum(use)} of {T} tables", flush=True)
    plan.close()
