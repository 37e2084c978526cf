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

__device__ __forceinline__ uint32_t rmi_leaf_at(float feature, float slope, float icpt, uint32_t L) {
#pragma clang fp contract(off)
  const float t = __fadd_rn(__fmul_rn(slope, feature), icpt);
  return !(t > 0.0f) ? 0u : (t >= (float)(L - 1) ? L - 1 : (uint32_t)t);
}

__device__ __forceinline__ uint32_t rmi_leaf(uint64_t key, const RmiHeader &h) {
  return rmi_leaf_at(rmi_feature(key, h.key_min, h.root_kind), h.root_slope, h.root_icpt, h.n_leaves);
}

__device__ __forceinline__ uint32_t rmi_bin(uint64_t key, const void *model, uint32_t M) {
#pragma clang fp contract(off)
  const RmiHeader *h = (const RmiHeader *)model;
  const RmiLeaf *lv = (const RmiLeaf *)(h + 1);
  const RmiLeaf l = lv[rmi_leaf(key, *h)];
  const bool up = key >= l.x0;
  const float d = up ? (float)(key - l.x0) : (float)(l.x0 - key);
  const float y = __fmul_rn(l.slope, d);
  const int64_t off = !(y > 0.0f) ? 0 : (y >= 4294967040.0f ? 0xFFFFFF00ll : (int64_t)(uint32_t)y);
  int64_t b = (int64_t)l.anchor + (up ? off : -off);
  b = b < 0 ? 0 : (b > (int64_t)M - 1 ? (int64_t)M - 1 : b);
  return (uint32_t)b;
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

template <typename K, int HK = HK_ANY, int SK = SK_ANY>
__device__ __forceinline__ uint32_t bucket_of(K key, const SegDesc &sd) {
  uint32_t bin = bin_of<K, HK>(key, sd.map);
  if (SK == SK_RANGE || (SK == SK_ANY && sd.kind == 0)) return divq(bin - (uint32_t)sd.bin_lo, sd.div);
  uint32_t p = 0;
  while (p < sd.nsplits && (int64_t)bin >= sd.splits[p + 1]) p++;
  return p;  // p == nsplits: past the last split
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
