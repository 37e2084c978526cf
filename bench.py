#!/usr/bin/env python3
"""HashGraph build + intersect throughput on one MI355X.

Headline step = build the CSR of A (2^28 uint32 keys, 2^28 bins, murmur3) and
intersect B (2^28 uint32 keys, 25% drawn from A) against it with the matched
keys written out (hg_build + hg_probe).  Inputs are generated on the device
before the timed region.  Prints one JSON line: the headline figures at the
top level, then one sub-result per other workload:
  c2_u64   2^27 + 2^27 uint64 keys, 2^27 bins
  c3       one GPU's share of MultiHashGraph: 2^30 uint64 Zipf(1.1) keys,
           16 logical partitions, B = 2^29, run-length join
  c5       4096 table/probe pairs (2^8..2^20 uint32 keys) in one captured plan
  learned  the learned hash beside murmur3 on 2^27 skewed uint64 sets
  rowid    the C2 uint32 shape built with input positions (INDEX_TRACK) and
           the row-id join writing every matching pair's positions
  staged   the C2 uint32 step with the tables in pinned host memory (the
           managed-memory replacement), bound by the host link
Each carries a `roofline` (the dominant kernel's algorithmic bytes over its
average launch time, from HIP events the library records around each kernel
on its stream, hg_ctx_profile) and a `cpu_baseline` (the oracle's threaded
port of the reference algorithm on the host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "Gkeys/s build & intersect at 2^28 u32 keys, %HBM peak; learned/murmur ratio"
# host threads the CPU baselines use: at most 16, one GPU's share of a node's cores
CPU_THREADS = min(16, os.cpu_count() or 1)

CONFIGS = {
    # name: (nA, nB, M, key_bytes, probe fraction)
    "c2-u32": (1 << 28, 1 << 28, 1 << 28, 4, 0.25),
    "c2-u64": (1 << 27, 1 << 27, 1 << 27, 8, 0.25),
    "c1": (1 << 20, 1 << 20, 1 << 20, 4, 0.5),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(nA, nB, M, s, K):
    """Algorithmic bytes (recomputable temporaries removed):
    build = read keys, write+read one partitioned copy, write CSR keys, write
    offsets = 4 s N + 4 (M+1); intersect (probe) = partition B (read+write),
    read partitioned B, read A's CSR keys and offsets, write K keys, count."""
    build = 4 * s * nA + 4 * (M + 1)
    inter = 3 * s * nB + s * nA + 4 * (M + 1) + s * K + 8
    return build, inter


def c2_kernel_bytes(name, nA, nB, M, s, K):
    """Algorithmic bytes per launch of each kernel of a C2 step.  part_hist and
    part_scatter run once for A (build) and once for B (probe); nA == nB in
    every configuration here."""
    n = nA
    return {
        "part_hist": s * n,                      # read the keys
        "part_scatter": 2 * s * n,               # read keys, write the partitioned copy
        "final": 2 * s * nA + 4 * (M + 1),       # read partitioned A, write CSR keys + offsets
        "probe": s * nB + s * nA + 4 * (M + 1) + s * K,  # partitioned B, A's CSR, matches
    }.get(name)


def load_traffic():
    """Measured HBM bytes per launch from the committed rocprofv3 PMC summary."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def roofline_from(kt, bytes_total, traffic=None, steps=1):
    """Dominant kernel (largest summed time) of a profile {name: (launches,
    ms)} and its achieved algorithmic bandwidth: bytes_total[name] (all of its
    launches) over its summed time."""
    known = {k: v for k, v in kt.items() if bytes_total.get(k)}
    dom = max(kt.items(), key=lambda kv: kv[1][1])[0]
    if dom not in known:  # a kernel without algorithmic bytes: the largest one with them
        dom = max(known.items(), key=lambda kv: kv[1][1])[0] if known else dom
    launches, ms = kt[dom]
    b = bytes_total.get(dom)
    achieved = b / (ms / 1e3) / 1e9 if b else None
    return {
        "kernel": dom, "bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": traffic,
        "alg_bytes_per_launch": int(b / launches) if b else None, "avg_launch_ms": round(ms / launches, 5),
        "launches_per_step": round(launches / steps, 2),
    }


def kernel_table(kt, steps, step_ms, bytes_total=None):
    out = {}
    for name, (launches, tot) in kt.items():
        row = {"launches_per_step": round(launches / steps, 2), "avg_ms": round(tot / max(launches, 1), 5),
               "share": round(tot / steps / step_ms, 4)}
        b = (bytes_total or {}).get(name)
        if b:
            row["alg_GBps"] = round(b / (tot / 1e3) / 1e9, 1)
        out[name] = row
    return out


# ---------------------------------------------------------------------------
# CPU baselines: the oracle's port of the reference algorithm (LRB build with
# per-thread coarse histograms and atomic fine counters, bin join) on the
# host cores, on bounded samples of each workload.
# ---------------------------------------------------------------------------
def _cpu_join_c2(A, B, M, threads):
    import numpy as np
    from oracle import oracle as o
    o.build_lrb(A[: 1 << 18], 1 << 18, 16000, threads)  # warm-up: page in, thread pool
    t0 = time.perf_counter()
    oa, ka = o.build_lrb(A, M, 16000, threads)
    ob, kbk = o.build_lrb(B, M, 16000, threads)
    cnt = o.intersect_par(oa, ka, ob, kbk, threads)
    out = np.empty(max(cnt, 1), dtype=A.dtype)
    o.intersect_par(oa, ka, ob, kbk, threads, out)
    return time.perf_counter() - t0, cnt


def cpu_baseline_c2(A_dev, B_dev, M, threads, one_thread_keys):
    """Full workload at `threads` threads, and a 1-thread figure on the first
    one_thread_keys keys of each side (M scaled with them)."""
    import numpy as np
    kb = A_dev.element_size()
    dt = np.uint32 if kb == 4 else np.uint64
    A = A_dev.cpu().numpy().view(dt)
    B = B_dev.cpu().numpy().view(dt)
    n = len(A)
    t, _ = _cpu_join_c2(A, B, M, threads)
    m1 = max(1, M * one_thread_keys // n)
    t1, _ = _cpu_join_c2(A[:one_thread_keys], B[:one_thread_keys], m1, 1)
    w = "uint32" if kb == 4 else "uint64"
    return {
        "value": round(2 * n / t / 1e9, 4), "unit": "Gkeys/s", "cores": threads, "kind": "port",
        "sample": f"the full workload (A, B = {n} {w} keys, {M} bins): LRB build of A and of B (16000 coarse bins) + "
                  f"bin join with emitted list, {threads} threads, {t:.2f} s",
        "one_thread": {"value": round(2 * one_thread_keys / t1 / 1e9, 4), "unit": "Gkeys/s", "cores": 1,
                       "sample": f"first {one_thread_keys} keys of A and of B, {m1} bins, 1 thread, {t1:.2f} s"},
    }


def cpu_baseline_c3(threads, nA=1 << 24, nB=1 << 23):
    """C3 sample: Zipf(1.1) uint64 keys (same generator, seeds 3 and 4), LRB
    builds of A and B over nA bins, canonical bins and the run-length join."""
    from hashgraph_amd import workloads as wl
    from oracle import oracle as o
    A = wl.zipf_keys(3, nA, 1.1)
    B = wl.zipf_keys(4, nB, 1.1)
    M = nA
    t0 = time.perf_counter()
    oa, ka = o.build_lrb(A, M, 16000, threads)
    ob, kbk = o.build_lrb(B, M, 16000, threads)
    o.canon(oa, ka)
    o.canon(ob, kbk)
    cnt, rk, _, _ = o.intersect_rle(oa, ka, ob, kbk)
    t = time.perf_counter() - t0
    return {"value": round((nA + nB) / t / 1e9, 4), "unit": "Gkeys/s", "cores": threads, "kind": "port",
            "sample": f"A = {nA}, B = {nB} uint64 Zipf(1.1) keys (the workload's generator, seeds 3 and 4), {M} bins: "
                      f"LRB builds, per-bin sort and run-length join ({len(rk)} records, count {cnt}), {threads} "
                      f"threads, {t:.2f} s"}


def cpu_baseline_c5(threads, budget_s=10.0, T=4096):
    """C5 sample: the batch's table/probe pairs in order (same generator),
    each built and joined on its own (the reference has no batched form),
    until budget_s seconds of CPU work have passed."""
    from hashgraph_amd import workloads as wl
    from oracle import oracle as o
    sizes = wl.batch_sizes(5, T, 8.0, 20.0)
    total, nk, done, t_work = 0, 0, 0, 0.0
    for t in range(T):
        a, b = wl.batch_pair(5, t, int(sizes[t]), 1.2, 0.3)
        m = max(len(a), 1)
        t0 = time.perf_counter()
        oa, ka = o.build_lrb(a, m, 16000, threads)
        ob, kbk = o.build_lrb(b, m, 16000, threads)
        total += o.intersect_par(oa, ka, ob, kbk, threads)
        t_work += time.perf_counter() - t0
        nk += len(a) + len(b)
        done += 1
        if t_work >= budget_s:
            break
    return {"value": round(nk / t_work / 1e9, 4), "unit": "Gkeys/s", "cores": threads, "kind": "port",
            "sample": f"the first {done} of the {T} pairs, in order ({nk} keys over both sides): per pair LRB builds "
                      f"of table and probe set + bin-join count ({total} matches; the Zipf tables' hot bins make "
                      f"the reference's nested join quadratic), {threads} threads, {t_work:.2f} s"}


# ---------------------------------------------------------------------------
# GPU workloads
# ---------------------------------------------------------------------------
def run_c2(cfg_name, dev, steps, warmup, cpu=True, one_thread_keys=1 << 25):
    import torch
    import hashgraph_amd as hg
    from hashgraph_amd import workloads as wl
    nA, nB, M, kb, frac = CONFIGS[cfg_name]
    s = kb
    log(f"{cfg_name}: nA={nA} nB={nB} M={M} key_bytes={kb}")
    A = wl.uniform_keys(1, nA, kb, device=dev)
    B = wl.probe_keys(A, 2, nB, frac, kb, device=dev)
    torch.cuda.synchronize()
    off = torch.empty(M + 1, dtype=torch.int32, device=dev)
    csr = torch.empty_like(A)
    hg.build(A, M, out=(off, csr))
    K, _ = hg.probe(off, csr, M, B, emit=False)
    matches = torch.empty(max(K, 1), dtype=A.dtype, device=dev)
    ctx = hg.context(0)

    def step():
        hg.build(A, M, out=(off, csr))
        cnt, _ = hg.probe(off, csr, M, B, out=matches)
        return cnt

    for _ in range(warmup):
        assert step() == K
    # timed region: `steps` steps bracketed by synchronisation
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    kt = ctx.profile_read(reset=True)
    ctx.profile(False)
    # build-only and probe-only phases (separate loops, for the split figures)
    e0.record()
    for _ in range(steps):
        hg.build(A, M, out=(off, csr))
    e1.record()
    torch.cuda.synchronize()
    ms_build = e0.elapsed_time(e1) / steps
    e0.record()
    for _ in range(steps):
        hg.probe(off, csr, M, B, out=matches)
    e1.record()
    torch.cuda.synchronize()
    ms_probe = e0.elapsed_time(e1) / steps
    b_bytes, i_bytes = alg_bytes(nA, nB, M, s, K)
    per_launch = {k: c2_kernel_bytes(k, nA, nB, M, s, K) for k in kt}
    bytes_total = {k: (v * kt[k][0] if v else None) for k, v in per_launch.items()}
    traffic = None
    roof = roofline_from(kt, bytes_total, steps=steps)
    tr = load_traffic().get(cfg_name, {}).get(roof["kernel"])
    roof["traffic"] = tr
    step_gbs = (b_bytes + i_bytes) / (ms / 1e3) / 1e9
    res = {
        "value": round((nA + nB) / (ms / 1e3) / 1e9, 3), "unit": "Gkeys/s", "ms_per_step": round(ms, 4),
        "dtype": "u32" if kb == 4 else "u64", "data": "synthetic (SplitMix64 uniform keys generated on the device)",
        "config": {
            "workload": f"build A + intersect B: A = {nA} {'uint32' if kb == 4 else 'uint64'} uniform keys, {M} bins, "
                        f"murmur3; B = {nB} keys, {int(frac * 100)}% drawn from A; matched keys emitted",
            "n_a": nA, "n_b": nB, "n_bins": M, "matches": K, "parallelism": "single GPU"},
        "build_gkeys_s": round(nA / (ms_build / 1e3) / 1e9, 3),
        "intersect_gkeys_s": round(nB / (ms_probe / 1e3) / 1e9, 3),
        "ms_build": round(ms_build, 4), "ms_intersect": round(ms_probe, 4),
        "step_alg_GBps": round(step_gbs, 1), "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "roofline": roof, "kernels": kernel_table(kt, steps, ms, bytes_total),
    }
    del off, csr, matches, traffic
    if cpu:
        log(f"{cfg_name}: cpu baseline ({CPU_THREADS} threads, then 1 thread)")
        res["cpu_baseline"] = cpu_baseline_c2(A, B, M, CPU_THREADS, min(one_thread_keys, nA))
    else:
        res["cpu_baseline"] = None
    del A, B
    torch.cuda.empty_cache()
    return res


def bin_stats(off):
    """max bin, mean length of non-empty bins, mean probe length (sum c^2 / n)."""
    import torch
    c = (off[1:] - off[:-1]).to(torch.int64)
    n = int(c.sum().item())
    nz = int((c > 0).sum().item())
    return {"max_bin": int(c.max().item()), "mean_nonempty_bin": round(n / max(nz, 1), 4),
            "mean_probe_len": round(float((c.double() ** 2).sum().item()) / max(n, 1), 4)}


def _time_table(A, B, M, hk, mdl, steps, dev):
    import torch
    import hashgraph_amd as hg
    n = A.numel()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    off = torch.empty(M + 1, dtype=torch.int32, device=dev)
    csr = torch.empty_like(A)
    hg.build(A, M, out=(off, csr), hash_kind=hk, rmi_model=mdl)
    K, _ = hg.probe(off, csr, M, B, emit=False, hash_kind=hk, rmi_model=mdl)
    out = torch.empty(max(K, 1), dtype=A.dtype, device=dev)
    for _ in range(2):
        hg.probe(off, csr, M, B, out=out, hash_kind=hk, rmi_model=mdl)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        hg.build(A, M, out=(off, csr), hash_kind=hk, rmi_model=mdl)
    e1.record()
    torch.cuda.synchronize()
    ms_b = e0.elapsed_time(e1) / steps
    e0.record()
    for _ in range(steps):
        hg.probe(off, csr, M, B, out=out, hash_kind=hk, rmi_model=mdl)
    e1.record()
    torch.cuda.synchronize()
    ms_p = e0.elapsed_time(e1) / steps
    # the probe's phases (partition of B, then the bucket kernels)
    ctx = hg.context(dev.index if dev.index is not None else 0)
    ctx.profile(True)
    ctx.profile_read(reset=True)
    hg.probe(off, csr, M, B, out=out, hash_kind=hk, rmi_model=mdl)
    kt = ctx.profile_read(reset=True)
    ctx.profile(False)
    st = bin_stats(off)
    st.update({"build_gkeys_s": round(n / ms_b / 1e6, 3), "probe_gkeys_s": round(B.numel() / ms_p / 1e6, 3),
               "ms_build": round(ms_b, 4), "ms_probe": round(ms_p, 4), "matches": K,
               "probe_phase_ms": {k: round(v[1], 4) for k, v in kt.items()}})
    return st, torch.sort(out[:K])[0]


def learned_block(dev, steps, n=1 << 27, frac=0.5):
    """Learned hash beside murmur on the skewed u64 sets (lognormal(40, 2) and
    64 clusters of width 2^24): probe throughput of both, bin statistics, the
    fallback decision and equality of the intersections."""
    import torch
    import hashgraph_amd as hg
    from hashgraph_amd import workloads as wl
    M = n
    res = {}
    for name in ("lognormal", "clusters"):
        gen = wl.lognormal_keys if name == "lognormal" else wl.cluster_keys
        A = gen(21, n, device=dev)
        B = wl.mixed_probe(A, gen(22, n, device=dev), 23, frac)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        model, use, rep = hg.rmi_train(A, M, n_sample=n // 100, seed=24, report=True)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        row = {"fallback_to_murmur": not use, "train_ms": round(t_train * 1e3, 3), "sample_check": rep}
        lists = {}
        row["murmur"], lists["murmur"] = _time_table(A, B, M, hg.HG_HASH_MURMUR3, None, steps, dev)
        if use:
            row["learned"], lists["learned"] = _time_table(A, B, M, hg.HG_HASH_RMI, model, steps, dev)
            row["identical_intersection"] = bool(lists["murmur"].numel() == lists["learned"].numel() and
                                                 torch.equal(lists["murmur"], lists["learned"]))
            row["probe_ratio_learned_over_murmur"] = round(row["learned"]["probe_gkeys_s"] /
                                                           row["murmur"]["probe_gkeys_s"], 4)
            # what any hash could gain here: only the bucket kernels walk bins;
            # even if all of their time scaled with the probe length and the
            # learned evaluation cost nothing, murmur's probe would shrink by
            # t_bucket * (1 - L_learned / L_murmur)
            mu, le = row["murmur"], row["learned"]
            t_bucket = mu["probe_phase_ms"].get("probe", 0.0)
            saving = max(0.0, 1.0 - le["mean_probe_len"] / mu["mean_probe_len"])
            row["probe_ratio_upper_bound"] = {
                "value": round(mu["ms_probe"] / (mu["ms_probe"] - t_bucket * saving), 4),
                "how": "murmur probe ms / (murmur probe ms - murmur bucket-kernel ms * (1 - learned/murmur mean "
                       "probe length)): the partition of B does not depend on the probe length"}
        else:
            # the sampled check kept murmur3: the learned form's tables are murmur's
            row["learned"] = {"fallback": "murmur3 (the model lost the sampled check); figures as murmur"}
            row["probe_ratio_learned_over_murmur"] = 1.0
        res[name] = row
        del A, B, model, lists
        torch.cuda.empty_cache()
    return {"workload": f"A = {n} uint64 keys, {M} bins; B = {n} keys, {int(frac * 100)}% drawn from A, the rest "
                        f"fresh from the same distribution; RMI with 4096 leaves trained on a 1% sample",
            **res}


def c3_kernel_bytes(nA, nB, M, P, s):
    """Algorithmic bytes of a C3 step's kernels, summed over their launches.
    Each side is partitioned twice (by split range, then by bucket inside each
    partition), then built; the run-length join needs one read of both sides'
    keys and 4-byte offsets (its count and emit passes read them twice; the
    measured traffic shows what they move)."""
    n2 = 2 * (nA + nB)
    return {
        "part_hist": s * n2,
        "part_scatter": 2 * s * n2,
        "final": 2 * s * (nA + nB) + 4 * 2 * (M + P),
        "rle_small": s * (nA + nB) + 8 * (M + P),
    }


def run_c3(dev, steps, warmup, cpu=True):
    """One GPU's share of MultiHashGraph: 2^30 uint64 Zipf(1.1) keys over 2^32
    ranks, 2^30 bins, 16 logical partitions (key-balanced splits over a
    16000-range histogram), per-partition CSR; B = 2^29 Zipf keys built with
    A's splits and joined partition by partition with the run-length join
    (the count, ~10^16, needs 64 bits and a run-length output)."""
    import torch
    import hashgraph_amd as hg
    from hashgraph_amd import workloads as wl
    nA, nB, M, P = 1 << 30, 1 << 29, 1 << 30, 16
    log("c3: generating keys")
    A = wl.zipf_keys(3, nA, 1.1, device=dev)
    B = wl.zipf_keys(4, nB, 1.1, device=dev)
    torch.cuda.synchronize()
    ctx = hg.context(0)

    def build_side(keys, splits=None):
        sp = hg.find_splits(keys, M, 16000, P) if splits is None else splits
        off, csr, po = hg.build_partitioned(keys, M, sp)
        return sp, off, csr, po

    def join(sa, ta, tb):
        _, offA, csrA, poA = ta
        _, offB, csrB, poB = tb
        total, recs = 0, 0
        for p in range(P):
            lo, hi = sa[p], sa[p + 1]
            if hi == lo:  # an empty bin range (repeated split point)
                continue
            oa = offA[lo + p: hi + p + 1]
            ob = offB[lo + p: hi + p + 1]
            ka = csrA[poA[p]: poA[p + 1]]
            kb = csrB[poB[p]: poB[p + 1]]
            try:  # one pass when the records fit the guess, else count first
                cnt, rk, ca, cb = hg.intersect_rle(oa, ka, ob, kb, hi - lo, capacity=1 << 23)
            except hg.HGError:
                cnt, rk, ca, cb = hg.intersect_rle(oa, ka, ob, kb, hi - lo)
            total += cnt
            recs += rk.numel()
        return total, recs

    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    for _ in range(warmup):
        ta = build_side(A)
        tb = build_side(B, ta[0])
        cnt, recs = join(ta[0], ta, tb)
        del ta, tb
    torch.cuda.synchronize()
    tb_ms, ti_ms = 0.0, 0.0
    ctx.profile(True)
    ctx.profile_read(reset=True)
    for _ in range(steps):
        e0.record()
        ta = build_side(A)
        e1.record()
        torch.cuda.synchronize()
        tb_ms += e0.elapsed_time(e1)
        e0.record()
        tb = build_side(B, ta[0])
        cnt, recs = join(ta[0], ta, tb)
        e1.record()
        torch.cuda.synchronize()
        ti_ms += e0.elapsed_time(e1)
        del ta, tb
    kt = ctx.profile_read(reset=True)
    ctx.profile(False)
    tb_ms /= steps
    ti_ms /= steps
    ms = tb_ms + ti_ms
    s = 8
    build_bytes = 4 * s * nA + 4 * (M + 1)
    inter_bytes = 3 * s * nB + s * nA + 4 * (M + 1)  # + the run-length records
    # kt sums `steps` steps: the per-step bytes scaled to match
    bt = {k: v * steps for k, v in c3_kernel_bytes(nA, nB, M, P, s).items()}
    res = {
        "value": round((nA + nB) / (ms / 1e3) / 1e9, 3), "unit": "Gkeys/s", "ms_per_step": round(ms, 3),
        "dtype": "u64", "data": "synthetic (Zipf(1.1) by rejection-inversion on the device)",
        "config": {"workload": "one GPU's share of MultiHashGraph: A = 2^30 uint64 Zipf(1.1) keys over 2^32 ranks, "
                               "2^30 bins, 16 logical partitions; B = 2^29 Zipf keys with A's splits; run-length join "
                               "per partition", "parallelism": "single GPU"},
        "build_gkeys_s": round(nA / tb_ms / 1e6, 3), "intersect_gkeys_s": round(nB / ti_ms / 1e6, 3),
        "ms_build": round(tb_ms, 3), "ms_intersect": round(ti_ms, 3),
        "build_hbm_frac": round(build_bytes / (tb_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "intersect_hbm_frac": round(inter_bytes / (ti_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "count": cnt, "rle_records": recs,
        "roofline": roofline_from(kt, bt, steps=steps), "kernels": kernel_table(kt, steps, ms, bt),
    }
    res["roofline"]["traffic"] = load_traffic().get("c3", {}).get(res["roofline"]["kernel"])
    del A, B
    torch.cuda.empty_cache()
    if cpu:
        log(f"c3: cpu baseline ({CPU_THREADS} threads)")
        res["cpu_baseline"] = cpu_baseline_c3(CPU_THREADS)
    else:
        res["cpu_baseline"] = None
    return res


def run_c5(dev, steps, warmup, T=4096, cpu=True, plan_only=False):
    """Batched form: T table/probe pairs, table sizes round(2^U[8,20]) uint32
    keys (even tables uniform, odd tables Zipf(1.2) ranks), probe sets of
    equal size with 30% drawn from their table, M_t = n_t; the learned hash
    per table where its sampled check wins.  The batch is one captured,
    autotuned plan (build + per-table probe counts); it is timed against T
    separate single-table build + probe calls."""
    import numpy as np
    import torch
    import hashgraph_amd as hg
    from hashgraph_amd import workloads as wl
    log(f"c5: generating {T} tables")
    keys, ko, probe, po = wl.batch_tables(5, T, 8.0, 20.0, 1.2, 0.3, device=dev)
    n_bins = [max(int(ko[t + 1] - ko[t]), 1) for t in range(T)]
    torch.cuda.synchronize()
    L = 256
    mb = hg.rmi_model_bytes(L)
    ns = [max(int(ko[t + 1] - ko[t]) // 100, min(int(ko[t + 1] - ko[t]), 1024)) for t in range(T)]
    hg.rmi_train_batch(keys, ko, n_bins, ns, list(range(T)), n_leaves=L)  # warm-up (first launch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    models, flags = hg.rmi_train_batch(keys, ko, n_bins, ns, list(range(T)), n_leaves=L)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    kinds = [hg.HG_HASH_RMI if u else hg.HG_HASH_MURMUR3 for u in flags]
    log(f"c5: trained {T} models in {t_train * 1e3:.1f} ms (batched trainer), learned on {sum(flags)} tables")
    if plan_only:
        # for counter runs: the plan's kernels alone, at the configuration the
        # autotuner picks here (32 K-key tiles, fill 0.875, its default)
        plan = hg.BatchPlan(keys, ko, n_bins, probe=probe, probe_off=po, hash_kinds=kinds, rmi_models=models,
                            rmi_leaves=L, autotune=False)
        for _ in range(warmup + steps):
            plan.run()
        torch.cuda.synchronize()
        out = {"replays": warmup + steps, "info": plan.info(), "count": int(plan.counts.sum().item())}
        plan.close()
        return out
    t0 = time.perf_counter()
    plan = hg.BatchPlan(keys, ko, n_bins, probe=probe, probe_off=po, hash_kinds=kinds, rmi_models=models,
                        rmi_leaves=L, autotune=True)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    for _ in range(warmup):
        plan.run()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        plan.run()
    e1.record()
    torch.cuda.synchronize()
    ms_batch = e0.elapsed_time(e1) / steps
    total = int(plan.counts.sum().item())
    info = plan.info()
    plan.close()
    # the same batch with murmur3 on every table: what the learned models
    # cost or save in throughput (the counts must be the same)
    plan_m = hg.BatchPlan(keys, ko, n_bins, probe=probe, probe_off=po, hash_kinds=[hg.HG_HASH_MURMUR3] * T,
                          autotune=True)
    for _ in range(warmup):
        plan_m.run()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        plan_m.run()
    e1.record()
    torch.cuda.synchronize()
    ms_murmur = e0.elapsed_time(e1) / steps
    assert int(plan_m.counts.sum().item()) == total
    murmur_only = {"ms_per_step": round(ms_murmur, 3), "info": plan_m.info()}
    plan_m.close()
    # the same batch with every model trained inside the graph on each replay
    models_t = torch.zeros(T * mb, dtype=torch.uint8, device=dev)
    plan_t = hg.BatchPlan(keys, ko, n_bins, probe=probe, probe_off=po, rmi_models=models_t, rmi_leaves=L,
                          autotune=True, train=True)
    for _ in range(warmup):
        plan_t.run()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        plan_t.run()
    e1.record()
    torch.cuda.synchronize()
    ms_trained = e0.elapsed_time(e1) / steps
    trained = {"ms_per_step": round(ms_trained, 3), "info": plan_t.info(),
               "count": int(plan_t.counts.sum().item()),
               "models_equal_pretrained": bool(torch.equal(models_t, models))}
    plan_t.close()
    del models_t
    # kernel breakdown of the same work, eager (the plan replays one graph)
    ctx = hg.context(0)
    off, csr, off_off = hg.batch_build(keys, ko, n_bins, hash_kinds=kinds, rmi_models=models, rmi_leaves=L)
    hg.batch_probe(off, csr, ko, n_bins, off_off, probe, po, hash_kinds=kinds, rmi_models=models, rmi_leaves=L)
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    off, csr, off_off = hg.batch_build(keys, ko, n_bins, hash_kinds=kinds, rmi_models=models, rmi_leaves=L)
    tot_e, _ = hg.batch_probe(off, csr, ko, n_bins, off_off, probe, po, hash_kinds=kinds, rmi_models=models,
                              rmi_leaves=L)
    kt = ctx.profile_read(reset=True)
    ctx.profile(False)
    del off, csr
    # the same work as separate single-table calls
    t_each = []
    cnt_each = 0
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c = 0
        for t in range(T):
            a, b = int(ko[t]), int(ko[t + 1])
            mdl = models[t * mb:(t + 1) * mb] if kinds[t] == hg.HG_HASH_RMI else None
            off, csr = hg.build(keys[a:b], n_bins[t], hash_kind=kinds[t], rmi_model=mdl)
            k, _ = hg.probe(off, csr, n_bins[t], probe[a:b], emit=False, hash_kind=kinds[t], rmi_model=mdl)
            c += k
        torch.cuda.synchronize()
        t_each.append(time.perf_counter() - t0)
        cnt_each = c
    ms_each = min(t_each) * 1e3
    nk = int(ko[T]) + int(po[T])
    NA, NB, SM = int(ko[T]), int(po[T]), int(np.sum(n_bins))
    s = 4
    # small-table bytes (every table's keys + CSR fit 64 MiB): build reads the
    # keys and writes CSR keys and offsets; the probe reads the probe keys,
    # the table's keys and offsets, and writes one count per table
    step_bytes = (2 * s * NA + 4 * (SM + T)) + (s * NB + s * NA + 4 * (SM + T) + 8 * T)
    ach = step_bytes / (ms_batch / 1e3) / 1e9
    bt = {"final": 2 * s * NA + 4 * (SM + T), "probe": s * NB + s * NA + 4 * (SM + T) + 8 * T,
          "part_hist": s * (NA + NB), "part_scatter": 2 * s * (NA + NB)}
    ms_eager = sum(v[1] for v in kt.values())
    res = {
        "value": round(nk / (ms_batch / 1e3) / 1e9, 3), "unit": "Gkeys/s", "ms_per_step": round(ms_batch, 3),
        "dtype": "u32", "data": "synthetic (SplitMix64 / Zipf(1.2) on the device)",
        "config": {"workload": f"{T} table/probe pairs, sizes round(2^U[8,20]) uint32 keys, half uniform, half "
                               f"Zipf(1.2); probe sets of equal size, 30% drawn from their table; bins = table size; "
                               f"one captured autotuned plan: build + per-table probe counts",
                   "parallelism": "single GPU"},
        "keys": NA, "probe_keys": NB, "count": total, "count_eager": tot_e, "count_separate_calls": cnt_each,
        "learned_tables": sum(k == 1 for k in kinds), "plan": info, "ms_separate_calls": round(ms_each, 3),
        "speedup_vs_separate_calls": round(ms_each / ms_batch, 2),
        "separate_calls_gkeys_s": round(nk / (ms_each / 1e3) / 1e9, 3),
        "train_ms": round(t_train * 1e3, 2), "plan_build_s": round(t_plan, 2),
        "plan_murmur_only": {**murmur_only, "value": round(nk / (ms_murmur / 1e3) / 1e9, 3), "unit": "Gkeys/s",
                             "note": "murmur3 on every table, same inputs and counts"},
        "plan_training_in_graph": {**trained, "value": round(nk / (trained["ms_per_step"] / 1e3) / 1e9, 3),
                                   "unit": "Gkeys/s",
                                   "note": "every model trained inside the graph on each replay (batched trainer, "
                                           "one workgroup per table), then build + probe counts"},
        "roofline": {"kernel": "plan replay (one graph: every kernel of the batch)", "bound": "hbm",
                     "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": load_traffic().get("c5", {}).get("plan replay"),
                     "alg_bytes_per_launch": step_bytes,
                     "avg_launch_ms": round(ms_batch, 4)},
        "kernels_eager": kernel_table(kt, 1, ms_eager, bt),
    }
    del keys, probe, models
    torch.cuda.empty_cache()
    if cpu:
        log(f"c5: cpu baseline ({CPU_THREADS} threads)")
        res["cpu_baseline"] = cpu_baseline_c5(CPU_THREADS)
    else:
        res["cpu_baseline"] = None
    return res


def run_rowid(dev, steps, warmup):
    """Row-id join at the C2 uint32 shape: A and B built with every entry's
    input position (the INDEX_TRACK entries: 8-byte records through the
    partition), then the input positions of every matching pair written."""
    import torch
    import hashgraph_amd as hg
    from hashgraph_amd import workloads as wl
    nA, nB, M, kb, frac = CONFIGS["c2-u32"]
    log(f"rowid: nA={nA} nB={nB} M={M} indexed builds + row-id join")
    A = wl.uniform_keys(1, nA, kb, device=dev)
    B = wl.probe_keys(A, 2, nB, frac, kb, device=dev)
    ta = hg.build_indexed(A, M)
    tb = hg.build_indexed(B, M)
    K, _, _ = hg.intersect_indexed(*ta, *tb, M, emit=False)
    ctx = hg.context(0)

    def step():
        a = hg.build_indexed(A, M)
        b = hg.build_indexed(B, M)
        return hg.intersect_indexed(*a, *b, M, capacity=K)[0]

    for _ in range(warmup):
        assert step() == K
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    kt = ctx.profile_read(reset=True)
    ctx.profile(False)
    # the partition counts the keys and scatters them as 8-byte records (input
    # position << 32 | key); the final stage stores each record as its key and
    # its position; the join reads both tables' keys, positions and offsets
    # and writes 16 bytes per pair
    bt = {"part_hist": kb * (nA + nB) * steps, "part_scatter": (kb + 8) * (nA + nB) * steps,
          "final": ((8 + kb + 8) * (nA + nB) + 4 * 2 * (M + 1)) * steps,
          "join_pairs": ((kb + 8) * (nA + nB) + 8 * (M + 1) + 16 * K) * steps}
    roof = roofline_from(kt, bt, steps=steps)
    roof["traffic"] = load_traffic().get("rowid", {}).get(roof["kernel"])
    res = {
        "value": round((nA + nB) / (ms / 1e3) / 1e9, 3), "unit": "Gkeys/s", "ms_per_step": round(ms, 3),
        "config": {"workload": f"indexed builds of A and B ({nA} uint32 keys each, {M} bins, 25% overlap) and the "
                               f"row-id join writing the input positions of all {K} matching pairs"},
        "matches": K, "roofline": roof, "kernels": kernel_table(kt, steps, ms, bt),
    }
    del A, B, ta, tb
    torch.cuda.empty_cache()
    return res


def run_staged(dev, steps, warmup, budget_mib=256):
    """The C2 uint32 step with both tables in pinned host memory (the
    MANAGED_MEM replacement): A and B built through a device budget of
    budget_mib MiB (about an eighth of one table with its CSR), then the staged
    join with its matched keys copied out.  Its bound is the host link, whose
    pinned copy rates are measured here in both directions."""
    import torch
    import hashgraph_amd as hg
    from hashgraph_amd import workloads as wl
    nA, nB, M, kb, frac = CONFIGS["c2-u32"]
    budget = budget_mib << 20
    log(f"staged: nA={nA} nB={nB} M={M} in pinned host memory, device budget {budget_mib} MiB")
    A = wl.uniform_keys(1, nA, kb, device=dev)
    B = wl.probe_keys(A, 2, nB, frac, kb, device=dev)
    off = torch.empty(M + 1, dtype=torch.int32, device=dev)
    csr = torch.empty_like(A)
    hg.build(A, M, out=(off, csr))
    K, _ = hg.probe(off, csr, M, B, emit=False)
    del off, csr
    hA, hB = A.cpu().pin_memory(), B.cpu().pin_memory()
    del A, B
    torch.cuda.empty_cache()

    def pin(n):
        return torch.empty(n, dtype=torch.int32, pin_memory=True)

    offA, csrA, offB, csrB, matches = pin(M + 1), pin(nA), pin(M + 1), pin(nB), pin(max(K, 1))
    x = torch.empty(nA, dtype=torch.int32, device=dev)

    def rate(f, reps=3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        return reps * nA * kb / (time.perf_counter() - t0) / 1e9

    h2d = rate(lambda: x.copy_(hA, non_blocking=True))
    d2h = rate(lambda: csrA.copy_(x, non_blocking=True))
    del x
    torch.cuda.empty_cache()

    def build(h, o, c):
        hg.build_staged(h, M, device_budget=budget, device=0, out=(o, c))

    def step():
        build(hA, offA, csrA)
        build(hB, offB, csrB)
        cnt, _ = hg.intersect_staged(offA, csrA, offB, csrB, M, device_budget=budget, device=0, out=matches)
        return cnt

    for _ in range(warmup):
        assert step() == K
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        build(hA, offA, csrA)
    ms_build = (time.perf_counter() - t0) * 1e3 / steps
    # host-link bytes: a staged build moves its keys up three times and down
    # twice plus its offsets down; the join moves both tables up and the
    # matches down
    up_b, down_b = 3 * nA * kb, 2 * nA * kb + 4 * (M + 1)
    up_j, down_j = (nA + nB) * kb + 2 * 4 * (M + 1), K * kb
    bound_build = (up_b / h2d + down_b / d2h) / 1e6
    bound_step = (2 * up_b / h2d + 2 * down_b / d2h + up_j / h2d + down_j / d2h) / 1e6
    res = {
        "value": round((nA + nB) / (ms / 1e3) / 1e9, 3), "unit": "Gkeys/s", "ms_per_step": round(ms, 2),
        "config": {"workload": f"C2 uint32 step with the tables in pinned host memory: build A and B "
                               f"({nA} keys each, {M} bins) through a {budget_mib} MiB device budget, staged "
                               f"join with the matched keys copied out", "device_budget_bytes": budget,
                   "matches": K},
        "build_gkeys_s": round(nA / (ms_build / 1e3) / 1e9, 3), "ms_build": round(ms_build, 2),
        "link_GBps": {"h2d_pinned": round(h2d, 1), "d2h_pinned": round(d2h, 1)},
        "link_roofline": {"bound": "host link", "build_bound_ms": round(bound_build, 2),
                          "build_frac": round(bound_build / ms_build, 3), "step_bound_ms": round(bound_step, 2),
                          "step_frac": round(bound_step / ms, 3)},
    }
    del hA, hB, offA, csrA, offB, csrB, matches
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="all",
                    choices=["all"] + sorted(CONFIGS) + ["c3", "c5", "learned", "staged", "rowid"],
                    help="all: the C2 u32 headline plus every other workload as sub-results")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--plan-only", action="store_true",
                    help="with --config c5: replay only the untuned plan warmup + steps times (counter runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the path would shard "
              f"by bin range across GPUs, which is not built here", file=sys.stderr)
        sys.exit(2)

    import torch

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    dev = torch.device("cuda", 0)
    if args.copy_calibration:
        n = args.copy_calibration << 20
        a = torch.ones(n // 4, dtype=torch.int32, device=dev)
        b = torch.empty_like(a)
        for _ in range(max(args.steps, 1)):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": max(args.steps, 1)}))
        return
    cpu = not args.no_cpu
    common = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None}
    if args.config in ("c3", "c5", "learned", "staged", "rowid") or args.config in CONFIGS and args.config != "c2-u32":
        if args.config == "rowid":
            r = {"rowid": run_rowid(dev, max(1, min(args.steps, 5)), min(args.warmup, 1))}
        elif args.config == "staged":
            r = {"staged": run_staged(dev, max(1, min(args.steps, 3)), min(args.warmup, 1))}
        elif args.config == "c3":
            r = run_c3(dev, args.steps, args.warmup, cpu)
        elif args.config == "c5":
            r = run_c5(dev, args.steps, args.warmup, cpu=cpu, plan_only=args.plan_only)
        elif args.config == "learned":
            r = {"learned": learned_block(dev, max(2, min(args.steps, 5)))}
        else:
            r = run_c2(args.config, dev, args.steps, args.warmup, cpu)
        print(json.dumps({**common, **r}))
        return
    res = {**common, **run_c2("c2-u32", dev, args.steps, args.warmup, cpu)}
    if args.config == "all":
        sub_steps, sub_warm = max(1, min(args.steps, 5)), max(1, min(args.warmup, 2))
        res["c2_u64"] = run_c2("c2-u64", dev, args.steps, args.warmup, cpu, one_thread_keys=1 << 24)
        log("learned-hash comparison")
        res["learned"] = learned_block(dev, max(2, min(args.steps, 5)))
        ratios = [res["learned"][k].get("probe_ratio_learned_over_murmur") for k in ("lognormal", "clusters")]
        res["learned_murmur_probe_ratio"] = {k: r for k, r in zip(("lognormal", "clusters"), ratios)}
        res["c3"] = run_c3(dev, sub_steps, sub_warm, cpu)
        res["c5"] = run_c5(dev, sub_steps, sub_warm, cpu=cpu)
        log("row-id join")
        res["rowid"] = run_rowid(dev, sub_steps, sub_warm)
        log("staged tables (host memory)")
        res["staged"] = run_staged(dev, max(1, min(args.steps, 3)), 1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
