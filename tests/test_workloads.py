"""The synthetic key generators: SplitMix64 published outputs, and the
numpy (CPU) and torch (device) paths agree bit for bit."""
import numpy as np

from hashgraph_amd import workloads as wl
from oracle import oracle as o


def test_splitmix64_published_sequence():
    # splitmix64.c (Vigna) seeded with x = 1234567: first outputs
    expect = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
              16408922859458223821]
    x = 1234567
    got = [wl.mix64_int(x + (i + 1) * wl.GAMMA) for i in range(5)]
    assert got == expect


def test_stream_matches_oracle_and_counter_form():
    s = wl.stream_np(42, 1000)
    s0 = wl.mix64_int(42)
    assert int(s[0]) == wl.mix64_int(s0 + wl.GAMMA)
    assert all(int(s[i]) == o.splitmix(42, i) for i in (0, 1, 7, 999))
    assert np.array_equal(wl.stream_np(42, 10, start=990), s[990:])


def test_torch_cpu_matches_numpy():
    import torch
    for kb in (4, 8):
        a = wl.uniform_keys(3, 5000, kb)
        t = wl.uniform_keys(3, 5000, kb, device=torch.device("cpu"))
        assert np.array_equal(t.numpy().view(a.dtype), a)
        b = wl.probe_keys(a, 8, 4000, 0.25, kb)
        tb = wl.probe_keys(t, 8, 4000, 0.25, kb, device=torch.device("cpu"))
        assert np.array_equal(tb.numpy().view(a.dtype), b)


def test_probe_fraction():
    a = wl.uniform_keys(1, 1 << 16, 4)
    b = wl.probe_keys(a, 2, 1 << 16, 0.5, 4)
    frac = np.isin(b, a).mean()
    assert 0.49 < frac < 0.52


def test_zipf_matches_pmf_and_torch():
    import torch
    from scipy.special import zeta
    s, n = 1.1, 1 << 32
    z = wl.zipf_keys(3, 1 << 18, s)
    Z = zeta(s) - n ** (1 - s) / (s - 1)  # sum_{k<=n} k^-s (Euler-Maclaurin tail)
    for r in range(1, 8):
        p = r ** -s / Z
        f = float((z == r - 1).mean())
        sd = (p * (1 - p) / len(z)) ** 0.5
        assert abs(f - p) < 5 * sd, (r, f, p)
    assert int(z.max()) < n
    zt = wl.zipf_keys(3, 1 << 14, s, device=torch.device("cpu"))
    assert np.array_equal(zt.numpy().view(np.uint64), wl.zipf_keys(3, 1 << 14, s))
