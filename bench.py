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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2-u32", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
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
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline ({threads} threads)")
        res["cpu_baseline"] = cpu_baseline(CONFIGS[args.config], threads)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res))


if __name__ == "__main__":
    main()
