"""Builds libhashgraph_amd.so (gfx950) in-tree with hipcc.

The library holds the HIP kernels, the C ABI of include/hashgraph.h and the
C++ drop-in classes (SingleHashGraph / MultiHashGraph).  The driver programs
that mirror the reference's executables are linked against it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libhashgraph_amd.so")

ARCH = os.environ.get("HG_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# -packed-fp32-ops: no packed FP32 instructions are generated; -ffp-contract=off:
# float products and sums stay separately rounded (the learned model's
# evaluation must equal its CPU restatement bit for bit)
DEVFLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off"]
CFLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"--offload-arch={ARCH}",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]

SOURCES = ["hg_kernels.hip", "hg_rmi.hip", "hg_engine.cpp", "SingleHashGraph.cpp", "MultiHashGraph.cpp"]
# driver programs: name -> (source, extra flags).  multi-hash-32,
# multi-hash-idx and the -managed ones are the reference's B32, INDEX_TRACK
# and MANAGED_MEM builds of the same driver (include/MultiHashGraph.cuh:62-74
# switches them at compile time).
DRIVERS = {"sing-hash": ("drivers/sing_hash.cpp", []), "multi-hash": ("drivers/multi_hash.cpp", []),
           "multi-hash-32": ("drivers/multi_hash.cpp", ["-DB32"]),
           "multi-hash-idx": ("drivers/multi_hash.cpp", ["-DINDEX_TRACK"]),
           "multi-hash-managed": ("drivers/multi_hash.cpp", ["-DMANAGED_MEM"]),
           "multi-hash-32-managed": ("drivers/multi_hash.cpp", ["-DB32", "-DMANAGED_MEM"]),
           "multi-hash-idx-managed": ("drivers/multi_hash.cpp", ["-DINDEX_TRACK", "-DMANAGED_MEM"])}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJDIR, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [HIPCC] + CFLAGS + DEVFLAGS + ["-c", "-x", "hip", path, "-o", obj]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(len(srcs), 6)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for name, (src, extra) in DRIVERS.items():
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        _run([HIPCC] + CFLAGS + extra + ["-x", "hip", path, "-o", os.path.join(BINDIR, name), "-L", LIBDIR,
                                         "-lhashgraph_amd", "-Wl,-rpath,$ORIGIN/../lib"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
