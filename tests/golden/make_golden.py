"""Writes the golden fixtures of the GPU parity tests (tests/golden/*.npz).

The reference repository commits no golden vectors for this path, so the
fixtures are produced by the oracle (whose functions are pinned by the
published MurmurHash3 vectors, closed forms and brute force in
tests/test_oracle.py) on small seeded inputs of every workload shape:

  c1_u32   : A = 4096 uniform uint32 keys (SplitMix64 seed 1), B = 4096 keys,
             50% drawn from A (seed 2), M = 4096: CSR offsets, canonical CSR
             keys, intersection count and sorted matched list
  c2_u64   : the same at 3000 uint64 keys, M = 2999 (not a power of two)
  multi    : 20000 Zipf(1.1) uint64 keys, M = 16384, 16 partitions: splits,
             per-partition offsets and canonical keys
  rle      : run-length join of two Zipf(1.1) tables (8000 and 6000 keys,
             M = 4096): count and records (key, countA, countB)

Run: python tests/golden/make_golden.py   (rewrites the .npz files)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from hashgraph_amd import workloads as wl  # noqa: E402
from oracle import oracle as o  # noqa: E402


def canon_csr(keys, M):
    off, csr = o.build_csr(keys, M)
    o.canon(off, csr)
    return off, csr


def join_case(A, B, M):
    offA, csrA = canon_csr(A, M)
    offB, csrB = canon_csr(B, M)
    cnt, lst = o.intersect(offA, csrA, offB, csrB)
    return dict(A=A, B=B, M=np.array(M), offA=offA, csrA=csrA, count=np.array(cnt), matches=np.sort(lst))


def main():
    A = wl.uniform_keys(1, 4096, 4)
    B = wl.probe_keys(A, 2, 4096, 0.5, 4)
    np.savez_compressed(os.path.join(HERE, "c1_u32.npz"), **join_case(A, B, 4096))

    A = wl.uniform_keys(3, 3000, 8)
    B = wl.probe_keys(A, 4, 3000, 0.3, 8)
    np.savez_compressed(os.path.join(HERE, "c2_u64.npz"), **join_case(A, B, 2999))

    K = wl.zipf_keys(5, 20000, 1.1)
    M, P = 16384, 16
    splits, _ = o.find_splits(o.bins(K, M), M, 16000, P)
    off, csr = o.build_multi(K, M, splits)
    for p in range(P):
        lo, hi = int(splits[p]), int(splits[p + 1])
        loc = off[lo + p: hi + p + 1]
        base = int(sum(off[int(splits[q + 1]) + q] for q in range(p)))
        seg = csr[base: base + int(loc[-1])]
        o.canon(loc, seg)
        csr[base: base + int(loc[-1])] = seg
    np.savez_compressed(os.path.join(HERE, "multi.npz"), keys=K, M=np.array(M), P=np.array(P), splits=splits,
                        off=off, csr=csr)

    A = wl.zipf_keys(6, 8000, 1.1)
    B = wl.zipf_keys(7, 6000, 1.1)
    M = 4096
    offA, csrA = canon_csr(A, M)
    offB, csrB = canon_csr(B, M)
    cnt, rk, ca, cb = o.intersect_rle(offA, csrA, offB, csrB)
    order = np.argsort(rk, kind="stable")
    np.savez_compressed(os.path.join(HERE, "rle.npz"), A=A, B=B, M=np.array(M), count=np.array(cnt),
                        rk=rk[order], ca=ca[order], cb=cb[order])
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
