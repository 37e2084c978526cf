#!/usr/bin/env python3
"""HashGraph build + intersect throughput on one MI355X.

One step = build the CSR of A (2^28 uint32 keys, 2^28 bins, murmur3) and
intersect B (2^28 uint32 keys, 25% drawn from A) against it with the matched
keys written out (hg_build + hg_probe).  Inputs are generated on the device
before the timed region.  Prints one JSON line.

The per-kernel roofline uses HIP events recorded by the library around each
kernel on its stream (hg_ctx_profile) during the timed steps; the algorithmic
bytes per launch of each kernel are defined in ALG_BYTES below.  The CPU
baseline is the oracle's threaded port of the reference's LRB build and bin
join on a smaller sample of the same workload.
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


def kernel_alg_bytes(name, nA, nB, M, s, K):
    """Algorithmic bytes per launch of each kernel of a step.  part_hist and
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


def cpu_baseline(cfg, threads):
    import numpy as np
    from hashgraph_amd import workloads as wl
    from oracle import oracle as o
    nA, nB, M, kb, frac = cfg
    scale = max(1, nA // (1 << 26))
    n, m = nA // scale, M // scale
    A = wl.uniform_keys(1, n, kb)
    B = wl.probe_keys(A, 2, n, frac, kb)
    # warm-up on a small slice (page in, thread pool)
    o.build_lrb(A[: 1 << 20], 1 << 20, 16000, threads)
    t0 = time.perf_counter()
    oa, ka = o.build_lrb(A, m, 16000, threads)
    ob, kbk = o.build_lrb(B, m, 16000, threads)
    cnt = o.intersect_par(oa, ka, ob, kbk, threads)
    out = np.empty(max(cnt, 1), dtype=A.dtype)
    o.intersect_par(oa, ka, ob, kbk, threads, out)
    t = time.perf_counter() - t0
    return {
        "value": round(2 * n / t / 1e9, 4),
        "unit": "Gkeys/s",
        "cores": threads,
        "kind": "port",
        "sample": f"A,B = 2^{n.bit_length() - 1} keys ({'uint32' if kb == 4 else 'uint64'}, same generators), "
                  f"2^{m.bit_length() - 1} bins: LRB build of A and of B (16000 coarse bins) + bin join with "
                  f"emitted list, {threads} threads, {t:.2f} s",
    }


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
    st = bin_stats(off)
    st.update({"build_gkeys_s": round(n / ms_b / 1e6, 3), "probe_gkeys_s": round(B.numel() / ms_p / 1e6, 3),
               "ms_build": round(ms_b, 4), "ms_probe": round(ms_p, 4), "matches": K})
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
        for hk, tag in ((hg.HG_HASH_MURMUR3, "murmur"), (hg.HG_HASH_RMI, "learned")):
            mdl = model if hk == hg.HG_HASH_RMI else None
            try:
                row[tag], lists[tag] = _time_table(A, B, M, hk, mdl, steps, dev)
            except hg.HGError as e:
                row[tag] = {"error": str(e)}
        if "learned" not in lists or "murmur" not in lists:
            res[name] = row
            continue
        row["identical_intersection"] = bool(lists["murmur"].numel() == lists["learned"].numel() and
                                             torch.equal(lists["murmur"], lists["learned"]))
        row["probe_ratio_learned_over_murmur"] = round(row["learned"]["probe_gkeys_s"] /
                                                       row["murmur"]["probe_gkeys_s"], 4)
        res[name] = row
        del A, B, model, lists
        torch.cuda.empty_cache()
    return {"workload": f"A = {n} uint64 keys, {M} bins; B = {n} keys, {int(frac * 100)}% drawn from A, the rest "
                        f"fresh from the same distribution; RMI with 4096 leaves trained on a 1% sample",
            **res}


def c3_run(dev, steps, warmup):
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
    tb_ms /= steps
    ti_ms /= steps
    return {"nA": nA, "nB": nB, "M": M, "P": P, "ms_build": tb_ms, "ms_intersect": ti_ms, "count": cnt,
            "records": recs}


def c5_run(dev, steps, warmup, T=4096):
    """Batched form: T table/probe pairs, table sizes round(2^U[8,20]) uint32
    keys (even tables uniform, odd tables Zipf(1.2) ranks), probe sets of
    equal size with 30% drawn from their table, M_t = n_t; the learned hash
    per table where its sampled check wins.  The batch is one captured,
    autotuned plan (build + per-table probe counts); it is timed against T
    separate single-table build + probe calls."""
    import torch
    import hashgraph_amd as hg
    from hashgraph_amd import workloads as wl
    log(f"c5: generating {T} tables")
    keys, ko, probe, po = wl.batch_tables(5, T, 8.0, 20.0, 1.2, 0.3, device=dev)
    n_bins = [max(int(ko[t + 1] - ko[t]), 1) for t in range(T)]
    torch.cuda.synchronize()
    L = 256
    mb = hg.rmi_model_bytes(L)
    models = torch.zeros(T * mb, dtype=torch.uint8, device=dev)
    kinds = []
    t0 = time.perf_counter()
    for t in range(T):
        n = int(ko[t + 1] - ko[t])
        m, use = hg.rmi_train(keys[int(ko[t]):int(ko[t + 1])], n_bins[t], n_sample=max(n // 100, min(n, 1024)),
                              seed=t, n_leaves=L)
        models[t * mb:(t + 1) * mb] = m
        kinds.append(hg.HG_HASH_RMI if use else hg.HG_HASH_MURMUR3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    log(f"c5: trained {T} models in {t_train:.1f} s, learned on {sum(k == 1 for k in kinds)} tables")
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
    info = plan.info()
    plan.close()
    return {"T": T, "keys": int(ko[T]), "probe_keys": int(po[T]), "ms_batch": ms_batch, "ms_separate": ms_each,
            "count": total, "count_separate": cnt_each, "learned_tables": sum(k == 1 for k in kinds),
            "plan": info, "train_s": t_train, "plan_s": t_plan, "nk": nk}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2-u32", choices=sorted(CONFIGS) + ["c3", "c5"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-hash comparison")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the path would shard "
              f"by bin range across GPUs, which is not built here", file=sys.stderr)
        sys.exit(2)

    import torch
    import hashgraph_amd as hg
    from hashgraph_amd import workloads as wl

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
    if args.config == "c3":
        r = c3_run(dev, args.steps, args.warmup)
        ms = r["ms_build"] + r["ms_intersect"]
        build_bytes = 36 * r["nA"]  # 4 s N + 4 (M + 1) with s = 8, N = M
        inter_bytes = 24 * r["nB"] + 8 * r["nA"] + 4 * r["M"]
        out = {
            "metric": "Gkeys/s build & intersect at 2^28 u32 keys, %HBM peak; learned/murmur ratio",
            "value": round((r["nA"] + r["nB"]) / (ms / 1e3) / 1e9, 3), "unit": "Gkeys/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic (Zipf(1.1) on the device)",
            "config": {"workload": "one GPU's share of MultiHashGraph: A = 2^30 uint64 Zipf(1.1) keys over 2^32 "
                                   "ranks, 2^30 bins, 16 logical partitions; B = 2^29 Zipf keys with A's splits; "
                                   "run-length join per partition", "parallelism": "single GPU"},
            "build_gkeys_s": round(r["nA"] / r["ms_build"] / 1e6, 3),
            "intersect_gkeys_s": round(r["nB"] / r["ms_intersect"] / 1e6, 3),
            "ms_build": round(r["ms_build"], 3), "ms_intersect": round(r["ms_intersect"], 3),
            "build_hbm_frac": round(build_bytes / (r["ms_build"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "intersect_hbm_frac": round(inter_bytes / (r["ms_intersect"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "count": r["count"], "rle_records": r["records"], "roofline": None, "cpu_baseline": None,
        }
        print(json.dumps(out))
        return
    if args.config == "c5":
        r = c5_run(dev, args.steps, args.warmup)
        out = {
            "metric": "Gkeys/s build & intersect at 2^28 u32 keys, %HBM peak; learned/murmur ratio",
            "value": round(r["nk"] / (r["ms_batch"] / 1e3) / 1e9, 3), "unit": "Gkeys/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r["ms_batch"], 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic (SplitMix64 / Zipf(1.2) on the device)",
            "config": {"workload": f"{r['T']} table/probe pairs, sizes round(2^U[8,20]) uint32 keys, half uniform, "
                                   f"half Zipf(1.2); probe sets of equal size, 30% drawn from their table; bins = "
                                   f"table size; one captured autotuned plan: build + per-table probe counts",
                       "parallelism": "single GPU"},
            "keys": r["keys"], "probe_keys": r["probe_keys"], "count": r["count"],
            "count_separate_calls": r["count_separate"], "learned_tables": r["learned_tables"],
            "plan": r["plan"], "ms_separate_calls": round(r["ms_separate"], 3),
            "speedup_vs_separate_calls": round(r["ms_separate"] / r["ms_batch"], 2),
            "separate_calls_gkeys_s": round(r["nk"] / (r["ms_separate"] / 1e3) / 1e9, 3),
            "train_s": round(r["train_s"], 2), "plan_build_s": round(r["plan_s"], 2),
            "roofline": None, "cpu_baseline": None,
        }
        print(json.dumps(out))
        return
    nA, nB, M, kb, frac = CONFIGS[args.config]
    s = kb
    log(f"config {args.config}: nA={nA} nB={nB} M={M} key_bytes={kb}")
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

    for _ in range(args.warmup):
        cnt = step()
        assert cnt == K
    # timed region: K steps bracketed by synchronisation
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    kt = ctx.profile_read(reset=True)
    ctx.profile(False)
    # build-only and probe-only phases (separate loops, for the split figures)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        hg.build(A, M, out=(off, csr))
    e1.record()
    torch.cuda.synchronize()
    ms_build = e0.elapsed_time(e1) / args.steps
    e0.record()
    for _ in range(args.steps):
        hg.probe(off, csr, M, B, out=matches)
    e1.record()
    torch.cuda.synchronize()
    ms_probe = e0.elapsed_time(e1) / args.steps

    b_bytes, i_bytes = alg_bytes(nA, nB, M, s, K)
    per_kernel = {}
    for name, (launches, tot) in kt.items():
        avg = tot / max(launches, 1)
        kb_ = kernel_alg_bytes(name, nA, nB, M, s, K)
        per_kernel[name] = {"launches_per_step": launches / args.steps, "avg_ms": round(avg, 5),
                            "share": round(tot / args.steps / ms, 4)}
        if kb_:
            per_kernel[name]["alg_GBps"] = round(kb_ / (avg / 1e3) / 1e9, 1)
    dom = max(kt.items(), key=lambda kv: kv[1][1])[0]
    traffic = load_traffic().get(args.config, {}).get(dom)
    dom_bytes = kernel_alg_bytes(dom, nA, nB, M, s, K)
    dom_avg = kt[dom][1] / kt[dom][0]
    achieved = dom_bytes / (dom_avg / 1e3) / 1e9 if dom_bytes else None
    roofline = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": round(achieved, 1) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
        "traffic": traffic,
        "alg_bytes_per_launch": dom_bytes,
    }
    step_gbs = (b_bytes + i_bytes) / (ms / 1e3) / 1e9
    res = {
        "metric": "Gkeys/s build & intersect at 2^28 u32 keys, %HBM peak; learned/murmur ratio",
        "value": round((nA + nB) / (ms / 1e3) / 1e9, 3),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32" if kb == 4 else "u64",
        "data": "synthetic (SplitMix64 uniform keys generated on the device)",
        "config": {
            "workload": f"build A + intersect B: A = {nA} {'uint32' if kb == 4 else 'uint64'} uniform keys, "
                        f"{M} bins, murmur3; B = {nB} keys, {int(frac * 100)}% drawn from A; matched keys emitted",
            "n_a": nA, "n_b": nB, "n_bins": M, "matches": K, "parallelism": "single GPU",
        },
        "build_gkeys_s": round(nA / (ms_build / 1e3) / 1e9, 3),
        "intersect_gkeys_s": round(nB / (ms_probe / 1e3) / 1e9, 3),
        "ms_build": round(ms_build, 4),
        "ms_intersect": round(ms_probe, 4),
        "step_alg_GBps": round(step_gbs, 1),
        "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "kernels": per_kernel,
    }
    del A, B, off, csr, matches
    torch.cuda.empty_cache()
    if not args.no_learned and args.config.startswith("c2"):
        log("learned-hash comparison")
        res["learned"] = learned_block(dev, max(2, min(args.steps, 5)))
        ratios = [res["learned"][k].get("probe_ratio_learned_over_murmur") for k in ("lognormal", "clusters")]
        res["learned_murmur_probe_ratio"] = {k: r for k, r in zip(("lognormal", "clusters"), ratios)}
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline ({threads} threads)")
        res["cpu_baseline"] = cpu_baseline(CONFIGS[args.config], threads)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res))


if __name__ == "__main__":
    main()
