"""Seeded synthetic key sets of the benchmark configurations.

The reference's inputs come from Thrust's minstd engine or a plain sequence
(test/MultiHashGraphTest.cu:81-128); these configurations define their own:
SplitMix64 (Steele, Lea & Flood, 2014) in counter form,
    x_i = mix64(s0 + (i + 1) * 0x9E3779B97F4A7C15),  s0 = mix64(seed),
uint32 keys = the top 32 bits of x_i, uint64 keys = x_i.  The same integer
arithmetic runs in numpy (CPU) and in torch int64 (on the GPU, wrapping
multiply, logical shifts by masking), so both produce identical bytes.
"""
from __future__ import annotations

import numpy as np

GAMMA = 0x9E3779B97F4A7C15
M1 = 0xBF58476D1CE4E5B9
M2 = 0x94D049BB133111EB
MASK64 = (1 << 64) - 1


def mix64_int(z: int) -> int:
    z &= MASK64
    z = ((z ^ (z >> 30)) * M1) & MASK64
    z = ((z ^ (z >> 27)) * M2) & MASK64
    return z ^ (z >> 31)


def _np_mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(M1)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(M2)
    return z ^ (z >> np.uint64(31))


def stream_np(seed: int, n: int, start: int = 0) -> np.ndarray:
    """x_start .. x_{start+n-1} of stream `seed` (uint64)."""
    s0 = np.uint64(mix64_int(seed))
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        return _np_mix64(s0 + i * np.uint64(GAMMA))


def _t_shr(x, k):
    import torch
    return (x >> k) & ((1 << (64 - k)) - 1)


def _t_mix64(z):
    z = (z ^ _t_shr(z, 30)) * _s64(M1)
    z = (z ^ _t_shr(z, 27)) * _s64(M2)
    return z ^ _t_shr(z, 31)


def _s64(v: int) -> int:
    """uint64 constant as the int64 with the same bits."""
    v &= MASK64
    return v - (1 << 64) if v >= (1 << 63) else v


def stream_torch(seed: int, n: int, device, start: int = 0, chunk: int = 1 << 26):
    """Same values as stream_np, computed on `device` as int64 bit patterns."""
    import torch
    out = torch.empty(n, dtype=torch.int64, device=device)
    s0 = _s64(mix64_int(seed))
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        i = torch.arange(start + c0 + 1, start + c1 + 1, dtype=torch.int64, device=device)
        out[c0:c1] = _t_mix64(i * _s64(GAMMA) + s0)
    return out


def uniform_keys(seed: int, n: int, key_bytes: int, device=None):
    """n uniform keys: uint32 = top 32 bits, uint64 = all 64 bits.
    numpy uint32/uint64 array when device is None, else an int32/int64 tensor."""
    if device is None:
        x = stream_np(seed, n)
        return (x >> np.uint64(32)).astype(np.uint32) if key_bytes == 4 else x
    import torch
    x = stream_torch(seed, n, device)
    if key_bytes == 8:
        return x
    return _t_shr(x, 32).to(torch.int32)  # wraps: same low 32 bits


def probe_keys(A, seed: int, n: int, frac: float, key_bytes: int, device=None):
    """n probe keys, each drawn from A with probability `frac` (position
    A[(r >> 32) * len(A) >> 32]) and fresh uniform otherwise (stream seed+1).
    r is element i of stream `seed`; the draw uses its low 16 bits."""
    thr = int(round(frac * 65536))
    nA = len(A) if device is None else A.numel()
    if device is None:
        r = stream_np(seed, n)
        take = (r & np.uint64(0xFFFF)) < np.uint64(thr)
        idx = (((r >> np.uint64(32)) * np.uint64(nA)) >> np.uint64(32)).astype(np.int64)
        fresh = uniform_keys(seed + 1, n, key_bytes)
        return np.where(take, A[idx], fresh)
    import torch
    r = stream_torch(seed, n, device)
    take = (r & 0xFFFF) < thr
    idx = (_t_shr(r, 32) * nA) >> 32  # < 2^32 * nA < 2^63 for nA < 2^31
    fresh = uniform_keys(seed + 1, n, key_bytes, device)
    return torch.where(take, A[idx], fresh)


def sequential_keys(n: int, key_bytes: int, device=None):
    """0 .. n-1 (the reference driver's default input, test/MultiHashGraphTest.cu:106-110)."""
    if device is None:
        return np.arange(n, dtype=np.uint32 if key_bytes == 4 else np.uint64)
    import torch
    return torch.arange(n, dtype=torch.int32 if key_bytes == 4 else torch.int64, device=device)


# ---------------------------------------------------------------------------
# Zipf(s) over ranks 1..n by rejection-inversion (Hormann & Derflinger,
# "Rejection-inversion to generate variates from monotone discrete
# distributions", 1996).  Attempt t of element i draws its uniform from
# element i*64 + t of the stream (53-bit doubles); an element that needs
# more than 64 attempts (probability far below 1e-12) keeps rank 1.
# Keys are rank - 1.
# ---------------------------------------------------------------------------
def _zipf_consts(s: float, n: int):
    import math
    def h_int(x):
        lx = math.log(x)
        t = (1.0 - s) * lx
        return (math.expm1(t) / t if abs(t) > 1e-8 else 1.0 + t / 2.0) * lx
    def h_inv(x):
        t = x * (1.0 - s)
        if t < -1.0:
            t = -1.0
        return math.exp((math.log1p(t) / t if abs(t) > 1e-8 else 1.0 - t / 2.0) * x)
    hx1 = h_int(1.5) - 1.0
    hn = h_int(n + 0.5)
    sp = 2.0 - h_inv(h_int(2.5) - math.exp(-s * math.log(2.0)))
    return hx1, hn, sp


def _zipf_np(u, s, n, hx1, hn, sp):
    """One rejection-inversion attempt for a vector of uniforms; returns (k, accepted)."""
    x_u = hn + u * (hx1 - hn)
    t = x_u * (1.0 - s)
    t = np.maximum(t, -1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        f1 = np.where(np.abs(t) > 1e-8, np.log1p(t) / t, 1.0 - t / 2.0)
    x = np.exp(f1 * x_u)
    k = np.floor(x + 0.5)
    k = np.clip(k, 1.0, float(n))
    lk = np.log(k + 0.5)
    t2 = (1.0 - s) * lk
    with np.errstate(divide="ignore", invalid="ignore"):
        f2 = np.where(np.abs(t2) > 1e-8, np.expm1(t2) / t2, 1.0 + t2 / 2.0)
    h_int_k = f2 * lk
    hk = np.exp(-s * np.log(k))
    acc = (k - x <= sp) | (x_u >= h_int_k - hk)
    return k, acc


def zipf_keys(seed: int, count: int, s: float = 1.1, universe: int = 1 << 32, device=None, chunk: int = 1 << 24):
    """count keys = rank - 1 with rank ~ Zipf(s) over 1..universe (uint64 numpy, or int64 on device)."""
    hx1, hn, sp = _zipf_consts(s, universe)
    if device is None:
        out = np.ones(count, dtype=np.uint64)
        for c0 in range(0, count, chunk):
            c1 = min(count, c0 + chunk)
            idx = np.arange(c0, c1, dtype=np.uint64)
            k_res = np.ones(c1 - c0)
            todo = np.ones(c1 - c0, dtype=bool)
            for t in range(64):
                if not todo.any():
                    break
                ii = np.nonzero(todo)[0]
                r = _np_mix64(np.uint64(mix64_int(seed)) + (idx[ii] * np.uint64(64) + np.uint64(t + 1)) *
                              np.uint64(GAMMA))
                u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
                k, acc = _zipf_np(u, s, universe, hx1, hn, sp)
                k_res[ii[acc]] = k[acc]
                todo[ii[acc]] = False
            out[c0:c1] = (k_res - 1.0).astype(np.uint64)
        return out
    import torch
    out = torch.empty(count, dtype=torch.int64, device=device)
    s0 = _s64(mix64_int(seed))
    for c0 in range(0, count, chunk):
        c1 = min(count, c0 + chunk)
        idx = torch.arange(c0, c1, dtype=torch.int64, device=device)
        k_res = torch.ones(c1 - c0, dtype=torch.float64, device=device)
        todo = torch.ones(c1 - c0, dtype=torch.bool, device=device)
        for t in range(64):
            ii = torch.nonzero(todo).flatten()
            if ii.numel() == 0:
                break
            r = _t_mix64((idx[ii] * 64 + (t + 1)) * _s64(GAMMA) + s0)
            u = _t_shr(r, 11).to(torch.float64) * (1.0 / 9007199254740992.0)
            x_u = hn + u * (hx1 - hn)
            tt = torch.clamp(x_u * (1.0 - s), min=-1.0)
            f1 = torch.where(tt.abs() > 1e-8, torch.log1p(tt) / tt, 1.0 - tt / 2.0)
            x = torch.exp(f1 * x_u)
            k = torch.clamp(torch.floor(x + 0.5), 1.0, float(universe))
            lk = torch.log(k + 0.5)
            t2 = (1.0 - s) * lk
            f2 = torch.where(t2.abs() > 1e-8, torch.expm1(t2) / t2, 1.0 + t2 / 2.0)
            acc = (k - x <= sp) | (x_u >= f2 * lk - torch.exp(-s * torch.log(k)))
            k_res[ii[acc]] = k[acc]
            todo[ii[acc]] = False
        out[c0:c1] = (k_res - 1.0).to(torch.int64)
    return out


# ---------------------------------------------------------------------------
# Skewed u64 sets of the learned-hash comparison.
#   lognormal: key = floor(exp(mu + sigma * z)), z standard normal by
#     Box-Muller (cos branch) from the 53-bit doubles u1 = (r_a >> 11 + 1)/2^53,
#     u2 = (r_b >> 11)/2^53 with r_a, r_b = elements 2j, 2j+1 of the stream,
#     j = i*64 + t for attempt t; a value >= 2^64 is redrawn (attempt t+1),
#     and after 64 attempts the key is 2^64 - 1.  numpy and torch agree to
#     the last bit of libm's exp/log/cos only approximately, so each side's
#     keys are used on their own.
#   clusters: 64 centres c_j = hi64(x_j * (2^64 - 2^24)), x_j = element j of
#     stream `seed`; key i = c[r >> 58] + (r & (2^24 - 1)), r = element i of
#     stream seed + 1.  Integer-only: numpy and torch give identical keys.
# ---------------------------------------------------------------------------
def _np_hi64(x, m: int):
    """hi64(x * m) for a uint64 array x and 0 < m < 2^64 (exact, by 32-bit limbs)."""
    x = x.astype(np.uint64)
    xl, xh = x & np.uint64(0xFFFFFFFF), x >> np.uint64(32)
    ml, mh = np.uint64(m & 0xFFFFFFFF), np.uint64(m >> 32)
    with np.errstate(over="ignore"):
        ll = xl * ml
        lh = xl * mh
        hl = xh * ml
        hh = xh * mh
        mid = (ll >> np.uint64(32)) + (lh & np.uint64(0xFFFFFFFF)) + (hl & np.uint64(0xFFFFFFFF))
        return hh + (lh >> np.uint64(32)) + (hl >> np.uint64(32)) + (mid >> np.uint64(32))


def cluster_keys(seed: int, n: int, n_clusters: int = 64, width_bits: int = 24, device=None):
    span = (1 << 64) - (1 << width_bits)
    cx = stream_np(seed, n_clusters)
    centres = _np_hi64(cx, span)
    cbits = n_clusters.bit_length() - 1
    assert 1 << cbits == n_clusters
    if device is None:
        r = stream_np(seed + 1, n)
        idx = (r >> np.uint64(64 - cbits)).astype(np.int64)
        with np.errstate(over="ignore"):
            return centres[idx] + (r & np.uint64((1 << width_bits) - 1))
    import torch
    ct = torch.tensor(centres.view(np.int64), device=device)
    r = stream_torch(seed + 1, n, device)
    idx = _t_shr(r, 64 - cbits)
    return ct[idx] + (r & ((1 << width_bits) - 1))


def lognormal_keys(seed: int, n: int, mu: float = 40.0, sigma: float = 2.0, device=None, chunk: int = 1 << 24):
    two64 = 18446744073709551616.0
    if device is None:
        out = np.empty(n, dtype=np.uint64)
        for c0 in range(0, n, chunk):
            c1 = min(n, c0 + chunk)
            idx = np.arange(c0, c1, dtype=np.uint64)
            res = np.full(c1 - c0, two64)
            todo = np.ones(c1 - c0, dtype=bool)
            s0 = np.uint64(mix64_int(seed))
            for t in range(64):
                ii = np.nonzero(todo)[0]
                if len(ii) == 0:
                    break
                j = idx[ii] * np.uint64(64) + np.uint64(t)
                with np.errstate(over="ignore"):
                    ra = _np_mix64(s0 + (j * np.uint64(2) + np.uint64(1)) * np.uint64(GAMMA))
                    rb = _np_mix64(s0 + (j * np.uint64(2) + np.uint64(2)) * np.uint64(GAMMA))
                u1 = ((ra >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
                u2 = (rb >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
                z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
                v = np.floor(np.exp(mu + sigma * z))
                ok = v < two64
                res[ii[ok]] = v[ok]
                todo[ii[ok]] = False
            o = np.empty(c1 - c0, dtype=np.uint64)
            big = res >= two64
            o[~big] = res[~big].astype(np.uint64)
            o[big] = np.uint64(0xFFFFFFFFFFFFFFFF)
            out[c0:c1] = o
        return out
    import torch
    out = torch.empty(n, dtype=torch.int64, device=device)
    s0 = _s64(mix64_int(seed))
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        idx = torch.arange(c0, c1, dtype=torch.int64, device=device)
        res = torch.full((c1 - c0,), two64, dtype=torch.float64, device=device)
        todo = torch.ones(c1 - c0, dtype=torch.bool, device=device)
        for t in range(64):
            ii = torch.nonzero(todo).flatten()
            if ii.numel() == 0:
                break
            j = idx[ii] * 64 + t
            ra = _t_mix64((j * 2 + 1) * _s64(GAMMA) + s0)
            rb = _t_mix64((j * 2 + 2) * _s64(GAMMA) + s0)
            u1 = (_t_shr(ra, 11).to(torch.float64) + 1.0) * (1.0 / 9007199254740992.0)
            u2 = _t_shr(rb, 11).to(torch.float64) * (1.0 / 9007199254740992.0)
            z = torch.sqrt(-2.0 * torch.log(u1)) * torch.cos(2.0 * 3.141592653589793 * u2)
            v = torch.floor(torch.exp(mu + sigma * z))
            ok = v < two64
            res[ii[ok]] = v[ok]
            todo[ii[ok]] = False
        # float64 in [2^63, 2^64) -> int64 bit pattern of the uint64 value
        hi = res >= 9223372036854775808.0
        lo_part = torch.where(hi, res - 9223372036854775808.0, res)
        o = lo_part.to(torch.int64)
        o = torch.where(hi, o + (-9223372036854775807 - 1), o)
        o = torch.where(res >= two64, torch.full_like(o, -1), o)
        out[c0:c1] = o
    return out


def mixed_probe(A, fresh, seed: int, frac: float):
    """Probe set: element i is A[(r >> 32) * len(A) >> 32] when the low 16 bits
    of r (element i of stream `seed`) fall below frac * 65536, else fresh[i]."""
    thr = int(round(frac * 65536))
    if isinstance(A, np.ndarray):
        n = len(fresh)
        r = stream_np(seed, n)
        take = (r & np.uint64(0xFFFF)) < np.uint64(thr)
        idx = (((r >> np.uint64(32)) * np.uint64(len(A))) >> np.uint64(32)).astype(np.int64)
        return np.where(take, A[idx], fresh)
    import torch
    n = fresh.numel()
    r = stream_torch(seed, n, A.device)
    take = (r & 0xFFFF) < thr
    idx = (_t_shr(r, 32) * A.numel()) >> 32
    return torch.where(take, A[idx], fresh)


# ---------------------------------------------------------------------------
# Batched tables: size of table t = round(2^u), u uniform on [lo_bits, hi_bits]
# from element t of stream `seed` (53-bit double); even tables hold uniform
# uint32 keys (stream seed*4096 + 2t+1), odd tables Zipf(s) ranks over 2^32
# (stream seed*4096 + 2t+2).  Probe set t has the same size, `frac` of it
# drawn from table t (probe_keys rule, stream seed*4096 + 2t+1 + 8192*4096).
# ---------------------------------------------------------------------------
def batch_sizes(seed: int, T: int, lo_bits: float = 8.0, hi_bits: float = 20.0):
    r = stream_np(seed, T)
    u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.rint(np.exp2(lo_bits + u * (hi_bits - lo_bits))).astype(np.int64)


def batch_pair(seed: int, t: int, n: int, zipf_s: float = 1.2, frac: float = 0.3, device=None):
    """Table t of a batch (n keys: uniform for even t, Zipf ranks' low 32 bits
    for odd t) and its probe set (n keys, frac drawn from the table)."""
    base = seed * 4096 + 2 * t + 1
    if t % 2 == 0:
        a = uniform_keys(base, n, 4, device)
    elif device is None:
        a = (zipf_keys(base + 1, n, zipf_s) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    else:
        z = zipf_keys(base + 1, n, zipf_s, device=device)
        a = (z - ((z >> 31) << 32)).to(__import__("torch").int32)  # low 32 bits
    return a, (probe_keys(a, base + 8192 * 4096, n, frac, 4, device) if n else a)


def batch_tables(seed: int, T: int, lo_bits: float = 8.0, hi_bits: float = 20.0, zipf_s: float = 1.2,
                 frac: float = 0.3, device=None):
    """Keys and probe sets of a batch: (keys, key_off, probe, probe_off); numpy
    uint32 arrays, or int32 tensors on `device` (same values)."""
    sizes = batch_sizes(seed, T, lo_bits, hi_bits)
    keys, probes = [], []
    for t, n in enumerate(sizes.tolist()):
        a, b = batch_pair(seed, t, n, zipf_s, frac, device)
        keys.append(a)
        probes.append(b)
    ko = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    if device is None:
        return np.concatenate(keys), ko, np.concatenate(probes), ko.copy()
    import torch
    return torch.cat(keys), ko, torch.cat(probes), ko.copy()
