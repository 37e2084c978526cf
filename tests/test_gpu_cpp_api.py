"""The C++ drop-in classes through the driver programs that mirror the
reference's executables (test/SingleHashGraphTest.cu, test/MultiHashGraphTest.cu).
multi-hash with "check" runs buildSingle() (per-bin degree and sorted-list
equality with a plain build, src/MultiHashGraph.cu:796-857) and the
reference's CPU intersect answer (test/MultiHashGraphTest.cu:340-425)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hashgraph_amd", "bin")


def _run(args, timeout=120):
    print("run", args, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    return r


def test_sing_hash(dev):
    r = _run([os.path.join(BIN, "sing-hash"), str(1 << 22), str(1 << 22), "16000"])
    assert r.returncode == 0
    assert "single buildTable() time:" in r.stdout


@pytest.mark.parametrize("gpus", [1, 4, 16])
def test_multi_hash_build_check(gpus, dev):
    r = _run([os.path.join(BIN, "multi-hash"), str(1 << 20), str(1 << 20), "16000", str(gpus), "16000", "check",
              str(1 << 19), "build"])
    assert r.returncode == 0
    assert "multi buildTable() time:" in r.stdout
    assert "error" not in r.stderr.lower()


@pytest.mark.parametrize("gpus", [1, 16])
def test_multi_hash_intersect_check(gpus, dev):
    nA, nB = 1 << 20, 1 << 19
    r = _run([os.path.join(BIN, "multi-hash"), str(nA), str(nA), "16000", str(gpus), "16000", "check", str(nB), "x"])
    assert r.returncode == 0
    assert "intersect check ok" in r.stdout
    assert f"intersect count: {min(nA, nB)}" in r.stdout  # sequential keys: {0..min-1}
    assert "error" not in r.stderr.lower()


@pytest.mark.parametrize("exe", ["multi-hash-32", "multi-hash-idx"])
@pytest.mark.parametrize("gpus", [1, 16])
def test_multi_hash_b32_and_index_modes(exe, gpus, dev):
    """The reference's B32 (uint32 keys) and INDEX_TRACK ({key, ind} entries)
    builds of the driver: buildSingle's checks (degrees, sorted lists and, with
    INDEX_TRACK, every entry's position pointing at its key) and the CPU
    intersect answer."""
    nA, nB = 1 << 20, 1 << 19
    r = _run([os.path.join(BIN, exe), str(nA), str(nA), "16000", str(gpus), "16000", "check", str(nB), "x"])
    assert r.returncode == 0
    assert "intersect check ok" in r.stdout
    assert f"intersect count: {min(nA, nB)}" in r.stdout
    assert "error" not in r.stderr.lower()
    r = _run([os.path.join(BIN, exe), str(nA), str(nA), "16000", str(gpus), "16000", "check", str(nB), "build"])
    assert r.returncode == 0 and "error" not in r.stderr.lower()


@pytest.mark.parametrize("exe", ["multi-hash-managed", "multi-hash-32-managed", "multi-hash-idx-managed"])
@pytest.mark.parametrize("gpus", [1, 16])
def test_multi_hash_managed_modes(exe, gpus, dev):
    """The reference's MANAGED_MEM builds (src/MultiHashGraph.cu:229-247): the
    tables live in managed memory and stream through a 4 MiB device budget
    (HG_STAGED_BUDGET), so every build and join runs in many slabs; the same
    buildSingle and CPU intersect checks as the device-memory builds."""
    nA, nB = 1 << 20, 1 << 19
    env = dict(os.environ, HG_STAGED_BUDGET=str(4 << 20))
    for mode in ("x", "build"):
        args = [os.path.join(BIN, exe), str(nA), str(nA), "16000", str(gpus), "16000", "check", str(nB), mode]
        print("run", args, flush=True)
        r = subprocess.run(args, capture_output=True, text=True, timeout=120, env=env)
        print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
        assert r.returncode == 0 and "error" not in r.stderr.lower()
        if mode == "x":
            assert "intersect check ok" in r.stdout
            assert f"intersect count: {min(nA, nB)}" in r.stdout


@pytest.mark.parametrize("name", ["c1_u32", "c2_u64"])
def test_single_hash_graph_class_against_golden(name, dev, tmp_path):
    """SingleHashGraph as a class (caller keys, build, CSR accessors,
    intersect) against the golden fixtures: offsets identical, every bin the
    same multiset of keys, intersection count and sorted key list identical
    (keys are int64 in the class; the uint32 fixture's values widen)."""
    import numpy as np
    g = np.load(os.path.join(ROOT, "tests", "golden", name + ".npz"))
    A, B, M = g["A"].astype(np.int64), g["B"].astype(np.int64), int(g["M"])
    A.tofile(tmp_path / "A.bin")
    B.tofile(tmp_path / "B.bin")
    r = _run([os.path.join(BIN, "sing-hash"), "--files", str(tmp_path / "A.bin"), str(tmp_path / "B.bin"), str(M),
              str(tmp_path)])
    assert r.returncode == 0
    assert f"intersect count: {int(g['count'])}" in r.stdout
    off = np.fromfile(tmp_path / "off.bin", dtype=np.int32).astype(np.int64)
    csr = np.fromfile(tmp_path / "csr.bin", dtype=np.int64)
    out = np.fromfile(tmp_path / "out.bin", dtype=np.int64)
    assert np.array_equal(off, g["offA"].astype(np.int64))
    csr, out = csr.view(np.uint64), out.view(np.uint64)  # the fixtures' order is unsigned
    for b in range(M):  # the fixture's bins are sorted (canonical order)
        lo, hi = int(off[b]), int(off[b + 1])
        assert np.array_equal(np.sort(csr[lo:hi]), g["csrA"][lo:hi].astype(np.uint64))
    assert np.array_equal(np.sort(out), np.sort(g["matches"].astype(np.uint64)))
