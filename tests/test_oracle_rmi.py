"""Pins of the oracle's learned-hash (RMI) evaluation.

The model has no counterpart in the reference.  Its evaluation is pinned by
closed forms on hand-built models (where every float step is exact), by an
exact integer restatement of the u64 -> f32 round-to-nearest-even
conversion, and by monotonicity of the map.
"""
import numpy as np
import pytest

from oracle import oracle as o

M32 = (1 << 32) - 1


def exact_f32(x: int) -> float:
    """Round-to-nearest-even of a non-negative integer to binary32, in integers."""
    if x == 0:
        return 0.0
    e = x.bit_length()
    if e <= 24:
        return float(x)
    sh = e - 24
    q, r = x >> sh, x & ((1 << sh) - 1)
    half = 1 << (sh - 1)
    if r > half or (r == half and (q & 1)):
        q += 1
    return float(q) * float(1 << sh)


def test_exact_f32_known_values():
    # 2^24 + 1 ties to even (2^24), 2^24 + 3 ties up (2^24 + 4)
    assert exact_f32((1 << 24) + 1) == float(1 << 24)
    assert exact_f32((1 << 24) + 3) == float((1 << 24) + 4)
    assert exact_f32((1 << 64) - 1) == 2.0 ** 64
    assert np.float32(exact_f32(123456789)) == np.float32(123456789.0)


def test_identity_leaf_is_truncation():
    """One leaf, slope 1, base 0: bin = min(f32(key - x0), M-1) exactly."""
    M = 1 << 20
    m = o.rmi_model(0, 0.0, 0.0, [(0, 1.0, 0)], M)
    for k in [0, 1, 7, 1000, M - 2, M - 1, M, 1 << 22, (1 << 24) - 1]:
        assert o.rmi_bin(m, k, M) == min(k, M - 1)


def test_power_of_two_slope():
    M = 1 << 20
    m = o.rmi_model(0, 0.0, 0.0, [(0, 1.0 / 16, 0)], M)
    keys = np.random.default_rng(3).integers(0, 1 << 24, 4000)
    for k in keys:
        assert o.rmi_bin(m, int(k), M) == min(int(k) >> 4, M - 1)


def test_piecewise_model_closed_form():
    """Root k >> 10 picks one of 4 leaves; leaf l maps to l*256 + (k - 1024 l)/4,
    so the whole map is k >> 2 on [0, 4096)."""
    leaves = [(1024 * l, 0.25, 256 * l) for l in range(4)]
    m = o.rmi_model(0, 1.0 / 1024, 0.0, leaves, 1024)
    for k in range(4096):
        assert o.rmi_leaf(m, k) == k >> 10
        assert o.rmi_bin(m, k, 1024) == k >> 2


def test_key_min_and_x0_anchors():
    """Keys below key_min go to leaf 0; keys below x0 take the leaf's base."""
    leaves = [(5000, 1.0, 10), (9000, 1.0, 5000)]
    m = o.rmi_model(4000, 1.0 / 4000, 0.0, leaves, 1 << 16)
    assert o.rmi_leaf(m, 0) == 0 and o.rmi_bin(m, 0, 1 << 16) == 10
    assert o.rmi_bin(m, 4999, 1 << 16) == 10
    assert o.rmi_bin(m, 5003, 1 << 16) == 13
    # leaf 0 is clamped by leaf 1's base
    assert o.rmi_bin(m, 7999, 1 << 16) == min(10 + 2999, 5000)
    assert o.rmi_leaf(m, 8000) == 1 and o.rmi_bin(m, 8000, 1 << 16) == 5000  # 8000 < x0 of leaf 1
    assert o.rmi_bin(m, 9100, 1 << 16) == 5100


def test_nan_and_negative_root_go_to_leaf_0():
    m = o.rmi_model(0, float("nan"), 0.0, [(0, 0.0, 3), (0, 0.0, 9)], 64)
    assert o.rmi_leaf(m, 12345) == 0
    m = o.rmi_model(0, 1.0, -1e30, [(0, 0.0, 3), (0, 0.0, 9)], 64)
    assert o.rmi_leaf(m, 12345) == 0
    m = o.rmi_model(0, 1.0, 1e30, [(0, 0.0, 3), (0, 0.0, 9)], 64)
    assert o.rmi_leaf(m, 0) == 1  # saturates at L - 1


@pytest.mark.parametrize("d", [(1 << 25) + 1, (1 << 25) + 2, (1 << 25) + 6, (1 << 31) + 12345, 4294967295])
def test_u64_to_f32_rounding(d):
    """slope 1 exposes the rounding of the key distance to binary32."""
    M = M32
    x0 = (1 << 63) + 17
    m = o.rmi_model(0, 0.0, 0.0, [(x0, 1.0, 0)], M)
    want = min(int(exact_f32(d)), 0xFFFFFF00, M - 1)
    assert o.rmi_bin(m, x0 + d, M) == want


def test_root_rounding_of_large_keys():
    """The root converts key - key_min of 64-bit keys: 2^64-1 rounds to 2^64."""
    L = 4096
    m = o.rmi_model(0, float(np.float32(L / 2.0 ** 64)), 0.0, [(0, 0.0, l) for l in range(L)], 1 << 20)
    assert o.rmi_leaf(m, (1 << 64) - 1) == L - 1
    for k in [0, 1 << 52, (1 << 63) + (1 << 40), 3 << 62]:
        t = np.float32(np.float32(L / 2.0 ** 64) * np.float32(exact_f32(k)))
        assert o.rmi_leaf(m, k) == min(int(t), L - 1)


def test_monotone_on_random_models():
    rng = np.random.default_rng(11)
    for trial in range(5):
        L = 64
        M = 1 << 16
        x0 = np.sort(rng.integers(0, 1 << 62, L, dtype=np.uint64))
        base = np.sort(rng.integers(0, M, L))
        slopes = rng.random(L) * 1e-12
        m = o.rmi_model(int(x0[0]), float(L / 2.0 ** 62), 0.0,
                        [(int(x0[i]), float(slopes[i]), int(base[i])) for i in range(L)], M)
        keys = np.sort(rng.integers(0, 1 << 62, 20000, dtype=np.uint64))
        b = o.bins_rmi(keys, m, M)
        assert (np.diff(b.astype(np.int64)) >= 0).all()
        assert b.max() < M


def test_build_csr_bins_matches_murmur_build():
    """The CSR from given bins equals the murmur build when given murmur's bins."""
    from hashgraph_amd import workloads as wl
    A = wl.uniform_keys(1, 5000, 4)
    off1, k1 = o.build_csr(A, 4096)
    off2, k2 = o.build_csr_bins(A, o.bins(A, 4096), 4096)
    assert (off1 == off2).all() and (k1 == k2).all()
