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
    """One leaf, slope 1, anchor 0: bin = f32(key - x0) exactly while that is
    a bin; keys past M - 1 are beyond the model and take murmur3's bin."""
    M = 1 << 20
    m = o.rmi_model(0, 0.0, 0.0, [(0, 1.0, 0)], M)
    for k in [0, 1, 7, 1000, M - 2, M - 1]:
        assert o.rmi_bin(m, k, M) == k
    for k in [M, M + 1, 1 << 22, (1 << 24) - 1]:
        assert o.rmi_bin(m, k, M) == o.hash_key(k) % M


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


def test_signed_extrapolation_below_x0():
    """Keys below a leaf's first key go down from its anchor (truncated
    distance), clamped at bin 0."""
    m = o.rmi_model(0, 0.0, 0.0, [(5000, 0.5, 100)], 1 << 16)
    assert o.rmi_bin(m, 5000, 1 << 16) == 100
    assert o.rmi_bin(m, 5011, 1 << 16) == 105
    assert o.rmi_bin(m, 4989, 1 << 16) == 95      # 100 - trunc(5.5)
    assert o.rmi_bin(m, 4800, 1 << 16) == 0       # 100 - 100
    assert o.rmi_bin(m, 4798, 1 << 16) == o.hash_key(4798) % (1 << 16)  # 100 - 101 < 0: murmur3
    assert o.rmi_bin(m, 0, 1 << 16) == o.hash_key(0) % (1 << 16)  # line -2400: murmur3
    assert o.rmi_bin(m, 5000 + 10 ** 6, 1 << 16) == o.hash_key(5000 + 10 ** 6) % (1 << 16)


def test_out_of_range_keys_take_murmur3():
    """The edges of the line: the last in-range value on each side is kept,
    one past it is murmur3 (published vectors pin hash_key); the rule
    depends on the key alone, so a table and its probe keys agree."""
    M = 1000
    m = o.rmi_model(0, 0.0, 0.0, [(10000, 1.0, 500)], M)
    assert o.rmi_bin(m, 10000 + 499, M) == 999       # line 999: the last bin
    assert o.rmi_bin(m, 10000 + 500, M) == o.hash_key(10500) % M
    assert o.rmi_bin(m, 10000 - 500, M) == 0         # line 0
    assert o.rmi_bin(m, 10000 - 501, M) == o.hash_key(9499) % M
    # murmur3 of the four low bytes, seed 0 (published vector of 0x87654321)
    assert o.hash_key(0x87654321) == 0xF55B516B
    m2 = o.rmi_model(0, 0.0, 0.0, [(0, 1.0, 0)], 1 << 30)
    assert o.rmi_bin(m2, (1 << 32) | 0x87654321, 1 << 30) == 0xF55B516B % (1 << 30)


def test_key_min_anchor_of_root():
    leaves = [(5000, 1.0, 10), (9000, 1.0, 5000)]
    m = o.rmi_model(4000, 1.0 / 4000, 0.0, leaves, 1 << 16)
    assert o.rmi_leaf(m, 0) == 0 and o.rmi_leaf(m, 4000) == 0 and o.rmi_leaf(m, 7999) == 0
    assert o.rmi_leaf(m, 8000) == 1
    assert o.rmi_bin(m, 8000, 1 << 16) == 5000 - 1000


def test_log_feature_root():
    """root_kind 1: the feature is the bits of f32(key - key_min) + 1 as an
    integer; with slope 2^-23 and icpt -127 the leaf is floor(log2(d + 1))."""
    L = 64
    m = o.rmi_model(0, 2.0 ** -23, -127.0, [(0, 0.0, l) for l in range(L)], 1 << 10, root_kind=1)
    for e in range(0, 60):
        k = 1 << e
        want = min(int(np.floor(np.log2(exact_f32(k) + 1.0))) if e < 24 else e, L - 1)
        assert o.rmi_leaf(m, k) == want, e
        assert o.rmi_bin(m, k, 1 << 10) == want


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
    # and below x0, from an anchor at the top
    m2 = o.rmi_model(0, 0.0, 0.0, [(x0, 1.0, M - 1)], M)
    assert o.rmi_bin(m2, x0 - d, M) == max(M - 1 - min(int(exact_f32(d)), 0xFFFFFF00), 0)


def test_root_rounding_of_large_keys():
    """The root converts key - key_min of 64-bit keys: 2^64-1 rounds to 2^64."""
    L = 4096
    m = o.rmi_model(0, float(np.float32(L / 2.0 ** 64)), 0.0, [(0, 0.0, l) for l in range(L)], 1 << 20)
    assert o.rmi_leaf(m, (1 << 64) - 1) == L - 1
    for k in [0, 1 << 52, (1 << 63) + (1 << 40), 3 << 62]:
        t = np.float32(np.float32(L / 2.0 ** 64) * np.float32(exact_f32(k)))
        assert o.rmi_leaf(m, k) == min(int(t), L - 1)


def test_monotone_when_leaves_are_monotone():
    """Leaves whose lines do not cross (anchors and slopes chosen so each
    leaf ends below the next one's start) give a monotone map."""
    rng = np.random.default_rng(11)
    L, M = 64, 1 << 16
    span = 1 << 56
    leaves = [(l * span, (M / L) / span * 0.9, l * (M // L)) for l in range(L)]
    m = o.rmi_model(0, float(L / 2.0 ** 62), 0.0, leaves, M)
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
