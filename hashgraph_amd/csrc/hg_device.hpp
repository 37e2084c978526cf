// hg_device.hpp -- device helpers for the gfx950 kernels (wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "hg_common.hpp"

namespace hg {

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

// MurmurHash3_x86_32 of the four little-endian bytes of k, seed 0: the
// reference's hash_murmur with len = sizeof(int)
// (include/SingleHashGraphOperators.cuh:31-69): one body block, no tail,
// h ^= 4, fmix32.
__device__ __forceinline__ uint32_t murmur4(uint32_t k) {
  k *= 0xcc9e2d51u;
  k = rotl32(k, 15);
  k *= 0x1b873593u;
  uint32_t h = rotl32(k, 13) * 5u + 0xe6546b64u;
  h ^= 4u;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// Raw buffer loads: one resource (base, size in bytes) per tile or bucket;
// the hardware returns 0 past the end, so slot j of thread t needs neither a
// clamp nor a branch.  The range check covers the VGPR offset (plus the
// instruction's immediate) but not the SGPR offset, so both parts of the
// offset go into the VGPR operand.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename K>
__device__ __forceinline__ K buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t off2);
template <>
__device__ __forceinline__ uint32_t buf_load<uint32_t>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t off2) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, (int)(voff + off2), 0, 0);
}
template <>
__device__ __forceinline__ uint64_t buf_load<uint64_t>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t off2) {
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(voff + off2), 0, 0);
  return (uint64_t)v[0] | ((uint64_t)v[1] << 32);
}

// a / d and a mod d with the invariant divider (layout in hg_common.hpp).
// The mode is uniform, so the branches are scalar and only one path runs.
__device__ __forceinline__ uint32_t divq(uint32_t a, const Div32 &v) {
  if (v.mode == 0) return a >> v.s;
  const uint32_t t = __umulhi(a, v.m);
  if (v.mode == 1) return t >> v.s;
  return (((a - t) >> 1) + t) >> v.s;
}
__device__ __forceinline__ uint32_t modq(uint32_t a, const Div32 &v) {
  if (v.mode == 0) return a & (v.d - 1);
  return a - divq(a, v) * v.d;
}

// dst[0..n) = src[0..n) through the workgroup: every thread issues PER
// bounds-checked buffer loads before it stores any of them, so a chunk of
// BLOCK*PER elements costs one memory latency (a plain strided loop waits
// for each load before the next).
template <typename T, int BLOCK, int PER>
__device__ __forceinline__ void lds_fill(T *dst, const T *src, uint32_t n) {
  for (uint32_t c0 = 0; c0 < n; c0 += BLOCK * PER) {
    const uint32_t m = min(n - c0, (uint32_t)(BLOCK * PER));
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(src + c0, m * (uint32_t)sizeof(T));
    T v[PER];
#pragma unroll
    for (int j = 0; j < PER; j++)
      v[j] = buf_load<T>(rs, threadIdx.x * (uint32_t)sizeof(T), (uint32_t)(j * BLOCK * sizeof(T)));
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const uint32_t q = j * BLOCK + threadIdx.x;
      if (q < m) dst[c0 + q] = v[j];
    }
  }
}

// Fills dk[0..nk) from sk and dof[0..no) with so[i] - base as 16-bit values
// (the caller guarantees they fit), with one batch of PER loads of each array
// per thread in flight before any store.
template <typename T, typename OFF, int BLOCK, int PER>
__device__ __forceinline__ void lds_fill_keys_offs(T *dk, const T *sk, uint32_t nk, uint16_t *dof, const OFF *so,
                                                   uint32_t no, OFF base) {
  const uint32_t n = max(nk, no);
  for (uint32_t c0 = 0; c0 < n; c0 += BLOCK * PER) {
    const uint32_t mk = c0 < nk ? min(nk - c0, (uint32_t)(BLOCK * PER)) : 0u;
    const uint32_t mo = c0 < no ? min(no - c0, (uint32_t)(BLOCK * PER)) : 0u;
    const __amdgpu_buffer_rsrc_t rk = make_rsrc(sk + c0, mk * (uint32_t)sizeof(T));
    const __amdgpu_buffer_rsrc_t ro = make_rsrc(so + c0, mo * (uint32_t)sizeof(OFF));
    T v[PER];
    OFF w[PER];
#pragma unroll
    for (int j = 0; j < PER; j++)
      v[j] = buf_load<T>(rk, threadIdx.x * (uint32_t)sizeof(T), (uint32_t)(j * BLOCK * sizeof(T)));
#pragma unroll
    for (int j = 0; j < PER; j++)
      w[j] = buf_load<OFF>(ro, threadIdx.x * (uint32_t)sizeof(OFF), (uint32_t)(j * BLOCK * sizeof(OFF)));
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const uint32_t q = j * BLOCK + threadIdx.x;
      if (q < mk) dk[c0 + q] = v[j];
      if (q < mo) dof[c0 + q] = (uint16_t)(w[j] - base);
    }
  }
}

// Learned CDF model (layout and rounding rules in hg_common.hpp).  The
// products and sums must not fuse into FMAs (HIP contracts by default).
__device__ __forceinline__ float rmi_feature(uint64_t key, uint64_t key_min, uint32_t kind) {
#pragma clang fp contract(off)
  const float xr = key > key_min ? (float)(key - key_min) : 0.0f;
  if (kind == 0) return xr;
  return (float)__float_as_uint(__fadd_rn(xr, 1.0f));
}

// v_cvt_u32_f32 saturates: NaN and values <= 0 give 0, values >= 2^32 give
// 0xFFFFFFFF, the rest truncate.
__device__ __forceinline__ uint32_t cvt_u32_sat(float x) {
  uint32_t r;
  asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// Leaf of a root value t: 0 for NaN and t <= 0, L-1 from L-1 on, else
// trunc(t) (the same as clamping in float while L-1 < 2^24).
__device__ __forceinline__ uint32_t rmi_leaf_at(float feature, float slope, float icpt, uint32_t L) {
#pragma clang fp contract(off)
  const float t = __fadd_rn(__fmul_rn(slope, feature), icpt);
  return min(cvt_u32_sat(t), L - 1);
}

__device__ __forceinline__ uint32_t rmi_leaf(uint64_t key, const RmiHeader &h) {
  return rmi_leaf_at(rmi_feature(key, h.key_min, h.root_kind), h.root_slope, h.root_icpt, h.n_leaves);
}

// Bin of a key routed to leaf l (M bins).  A key whose line value lies
// outside [0, M-1] (beyond the range the model was trained on) takes
// murmur3's bin instead of piling up in an edge bin.
__device__ __forceinline__ uint32_t rmi_leaf_bin(uint64_t key, const RmiLeaf &l, uint32_t M) {
#pragma clang fp contract(off)
  const bool up = key >= l.x0;
  const float d = up ? (float)(key - l.x0) : (float)(l.x0 - key);
  const float y = __fmul_rn(l.slope, d);
  // offset: 0 for NaN and y <= 0, trunc(y) up to 0xFFFFFF00 (the largest
  // float below 2^32); the line value anchor +- offset
  const uint32_t off = min(cvt_u32_sat(y), 0xFFFFFF00u);
  const int64_t b = up ? (int64_t)l.anchor + off : (int64_t)l.anchor - off;
  if (b >= 0 && b < (int64_t)M) return (uint32_t)b;
  return murmur4((uint32_t)key) % M;
}

__device__ __forceinline__ uint32_t rmi_bin(uint64_t key, const void *model, uint32_t M) {
  const RmiHeader *h = (const RmiHeader *)model;
  return rmi_leaf_bin(key, ((const RmiLeaf *)(h + 1))[rmi_leaf(key, *h)], M);
}

// Learned-model bins from LDS windows of leaves, for kernels whose keys sit
// in registers (kr[0..KPT), slots below nv valid).  A window holds leaves
// [w0, w0 + n); rmi_window_bins calls f(j, bin) for the slots whose leaf is
// in it.  One code path per slot whatever the model's size: a workgroup walks
// as many windows as its keys' leaf range needs (one on most workgroups).
template <typename K, int KPT, typename F>
__device__ __forceinline__ void rmi_window_bins(const K *kr, uint32_t nv, const RmiHeader &h, const RmiLeaf *s_leaf,
                                                uint32_t w0, uint32_t n, uint32_t M, F &&f) {
#pragma unroll
  for (int j = 0; j < KPT; j++)
    if (j < (int)nv) {
      uint64_t k = (uint64_t)kr[j];
      asm volatile("" : "+v"(k));  // recomputed per window: nothing is hoisted into registers across windows
      const uint32_t leaf = rmi_leaf(k, h);
      if (leaf - w0 < n) f(j, rmi_leaf_bin(k, s_leaf[leaf - w0], M));
    }
}

// Copies leaves [w0, w0 + n) of a model to s_leaf (no barrier).
template <int BLOCK>
__device__ __forceinline__ void rmi_copy_leaves(const void *model, RmiLeaf *s_leaf, uint32_t w0, uint32_t n) {
  const uint4 *src = (const uint4 *)((const RmiHeader *)model + 1) + w0;
  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) ((uint4 *)s_leaf)[i] = src[i];
}

// Leaf range of a workgroup's keys (kr[0..KPT), slots below nv valid): the
// root is monotone in the key, so every key's leaf lies between the leaves of
// the smallest and the largest key.  Loads the header into h; [l0, l1] is
// empty (l0 > l1) when no slot is valid.  Ends with a barrier.
template <typename K, int BLOCK, int KPT>
__device__ __forceinline__ void rmi_leaf_range(const K *kr, uint32_t nv, const void *model, RmiHeader &h,
                                               uint32_t &l0, uint32_t &l1) {
  __shared__ uint64_t s_mn[BLOCK / 64], s_mx[BLOCK / 64];
  __shared__ uint32_t s_rng[2];
  h = *(const RmiHeader *)model;
  uint64_t mn = ~0ull, mx = 0;
#pragma unroll
  for (int j = 0; j < KPT; j++)
    if (j < (int)nv) {
      const uint64_t k = (uint64_t)kr[j];
      mn = k < mn ? k : mn;
      mx = k > mx ? k : mx;
    }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const uint64_t a = __shfl_xor(mn, d, 64), b = __shfl_xor(mx, d, 64);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  if ((threadIdx.x & 63) == 0) {
    s_mn[threadIdx.x >> 6] = mn;
    s_mx[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLOCK / 64; w++) {
      mn = s_mn[w] < mn ? s_mn[w] : mn;
      mx = s_mx[w] > mx ? s_mx[w] : mx;
    }
    uint32_t a = 1, b = 0;  // empty
    if (mn <= mx) {
      const uint32_t x = rmi_leaf(mn, h), y = rmi_leaf(mx, h);
      a = x < y ? x : y;
      b = x < y ? y : x;
    }
    s_rng[0] = a;
    s_rng[1] = b;
  }
  __syncthreads();
  l0 = s_rng[0];
  l1 = s_rng[1];
}

// Local bins (bin - bin_lo, 16 bits, OR-ed into the packed lb) of a
// workgroup's keys under a learned model, through LDS windows of cap leaves.
template <typename K, int BLOCK, int KPT>
__device__ __forceinline__ void rmi_local_bins(const K *kr, uint32_t nv, uint32_t *lb, const void *model,
                                               RmiLeaf *s_leaf, uint32_t cap, uint32_t bin_lo, uint32_t M) {
  RmiHeader h;
  uint32_t l0, l1;
  rmi_leaf_range<K, BLOCK, KPT>(kr, nv, model, h, l0, l1);
  for (uint32_t w0 = l0; w0 <= l1; w0 += cap) {
    const uint32_t n = min(cap, l1 - w0 + 1);
    rmi_copy_leaves<BLOCK>(model, s_leaf, w0, n);
    __syncthreads();
    rmi_window_bins<K, KPT>(kr, nv, h, s_leaf, w0, n, M, [&](int j, uint32_t bin) {
      lb[j >> 1] |= ((bin - bin_lo) & 0xFFFFu) << ((j & 1) * 16);
    });
    __syncthreads();
    if (l1 - w0 < cap) break;  // last window (w0 + cap may wrap)
  }
}

// Hash kinds a kernel is compiled for: only the selected path is inlined
// into every unrolled slot (keeps the kernels' code inside the I-cache).
enum : int { HK_MURMUR = 0, HK_RMI = 1, HK_ANY = 2 };

template <typename K, int HK = HK_ANY>
__device__ __forceinline__ uint32_t bin_of(K key, const BinMap &m) {
  if (HK == HK_MURMUR || (HK == HK_ANY && m.kind == 0)) {
    return modq(murmur4((uint32_t)key), m.mod);
  }
  return rmi_bin((uint64_t)key, m.rmi, m.M);
}

// Bucket kinds: SK_RANGE = (bin - bin_lo) / width, SK_SPLIT = partition by
// split points, SK_ANY = chosen per segment at run time.
enum : int { SK_RANGE = 0, SK_SPLIT = 1, SK_ANY = 2 };

// Split-point bucket: the number of splits spl[1..n] (sorted) that are <= bin,
// i.e. the first p with bin < spl[p+1] (n: past the last split).  A
// branch-free binary search (uniform step, clamped reads).
__device__ __forceinline__ uint32_t split_bucket(uint32_t bin, const int64_t *spl, uint32_t n) {
  uint32_t p = 0;
  for (uint32_t step = n ? 1u << (31 - __clz(n)) : 0u; step; step >>= 1) {
    const uint32_t q = p + step;
    const int64_t v = spl[q <= n ? q : n];
    p = (q <= n && v <= (int64_t)bin) ? q : p;
  }
  return p;
}

// Bucket of a bin; spl: the segment's splits (an LDS copy, or sd.splits).
template <int SK = SK_ANY>
__device__ __forceinline__ uint32_t bucket_of_bin(uint32_t bin, const SegDesc &sd, const int64_t *spl) {
  if (SK == SK_RANGE || (SK == SK_ANY && sd.kind == 0)) return divq(bin - (uint32_t)sd.bin_lo, sd.div);
  return split_bucket(bin, spl, sd.nsplits);
}

template <int SK = SK_ANY>
__device__ __forceinline__ uint32_t bucket_of_bin(uint32_t bin, const SegDesc &sd) {
  return bucket_of_bin<SK>(bin, sd, sd.splits);
}

// Copies a split-point segment's splits to LDS (s_spl holds MAX_SPLITS + 1)
// and returns the pointer to use: s_spl, or sd.splits for a range segment.
// Ends with a barrier.
template <int SK, int BLOCK>
__device__ __forceinline__ const int64_t *stage_splits(const SegDesc &sd, int64_t *s_spl) {
  if (SK == SK_RANGE) {
    __syncthreads();  // the barrier callers rely on (e.g. after zeroing counters)
    return sd.splits;
  }
  const bool split = SK == SK_SPLIT || sd.kind == 1;
  if (split)
    for (uint32_t i = threadIdx.x; i <= sd.nsplits; i += BLOCK) s_spl[i] = sd.splits[i];
  __syncthreads();
  return split ? s_spl : sd.splits;
}

template <typename K, int HK = HK_ANY, int SK = SK_ANY>
__device__ __forceinline__ uint32_t bucket_of(K key, const SegDesc &sd) {
  return bucket_of_bin<SK>(bin_of<K, HK>(key, sd.map), sd);
}

template <typename K, int HK, int SK>
__device__ __forceinline__ uint32_t bucket_of(K key, const SegDesc &sd, const int64_t *spl) {
  return bucket_of_bin<SK>(bin_of<K, HK>(key, sd.map), sd, spl);
}

// Inclusive wave64 scan of a 32/64-bit value.
template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// A value the same in every lane, moved to scalar registers (frees VGPRs
// in kernels whose block-uniform values come from LDS).
__device__ __forceinline__ uint32_t uni32(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
  return ((uint64_t)uni32((uint32_t)(x >> 32)) << 32) | uni32((uint32_t)x);
}

// Exclusive block-wide sum of one value per thread.  s_tmp needs
// BLOCK/64 + 1 elements of T.  Ends with a barrier so s_tmp can be reused.
template <int BLOCK, typename T>
__device__ __forceinline__ T block_exclusive_sum(T v, T *s_tmp, T &total) {
  constexpr int NW = BLOCK / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  T x = wave_inclusive_sum<T>(v);
  if (lane == 63) s_tmp[wid] = x;
  __syncthreads();
  if (wid == 0) {
    T w = lane < NW ? s_tmp[lane] : T(0);
    T wx = wave_inclusive_sum<T>(w);
    if (lane < NW) s_tmp[lane] = wx - w;
    if (lane == NW - 1) s_tmp[NW] = wx;
  }
  __syncthreads();
  total = s_tmp[NW];
  T r = s_tmp[wid] + x - v;
  __syncthreads();
  return r;
}

template <int BLOCK, typename T>
__device__ __forceinline__ T block_sum(T v, T *s_tmp) {
  T tot;
  (void)block_exclusive_sum<BLOCK, T>(v, s_tmp, tot);
  return tot;
}

// Exclusive scan of a[0..n) in LDS (n <= a few times BLOCK*64).  Each thread
// scans a contiguous chunk.  Returns the total; ends with a barrier.
template <int BLOCK>
__device__ __forceinline__ uint32_t lds_exclusive_scan(uint32_t *a, uint32_t n, uint32_t *s_tmp) {
  const uint32_t per = (n + BLOCK - 1) / BLOCK;
  const uint32_t lo = min(threadIdx.x * per, n), hi = min(lo + per, n);
  uint32_t s = 0;
  for (uint32_t i = lo; i < hi; i++) s += a[i];
  uint32_t tot;
  uint32_t ex = block_exclusive_sum<BLOCK, uint32_t>(s, s_tmp, tot);
  for (uint32_t i = lo; i < hi; i++) {
    uint32_t v = a[i];
    a[i] = ex;
    ex += v;
  }
  __syncthreads();
  return tot;
}

}  // namespace hg
