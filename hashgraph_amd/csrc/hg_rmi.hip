// hg_rmi.hip -- training of the learned CDF model (RMI) on gfx950.
//
// No counterpart in the reference.  Steps, all on the device:
//   1. seeded sample of the table's keys (with replacement),
//   2. LSD radix sort of the sample (8-bit digits, stable wave-ballot ranks),
//   3. four root candidates (raw or log feature; least squares or range
//      fit, in FP64, stored as FP32), the one with the most even leaf loads
//      on the sample is kept,
//   4. leaf ranges of the sorted sample under the stored FP32 root,
//   5. one least-squares fit per leaf (key -> global rank position * M / n),
//   6. the fallback check: both hashes bin the sample at the table's load
//      factor; max bin and sum of squared bin sizes are compared.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "hg_common.hpp"
#include "hg_device.hpp"
#include "hg_kernels.hpp"

namespace hg {

constexpr int SORT_BLOCK = 256;
constexpr int SORT_IPT = 16;
constexpr uint32_t SORT_TILE = SORT_BLOCK * SORT_IPT;
constexpr uint64_t SPLITMIX_GAMMA = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// sample[i] = keys[hi64(x_i * n)], x_i = mix64(s0 + (i + 1) * gamma).
template <typename K>
__global__ __launch_bounds__(256) void k_sample(const K *__restrict__ keys, uint64_t n, uint64_t s0, uint64_t ns,
                                                K *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < ns; i += (uint64_t)gridDim.x * 256) {
    const uint64_t x = mix64(s0 + (i + 1) * SPLITMIX_GAMMA);
    out[i] = keys[__umul64hi(x, n)];
  }
}

// Digit histogram of one tile, digit-major: counts[d * ntiles + tile].
template <typename K>
__global__ __launch_bounds__(SORT_BLOCK) void k_sort_hist(const K *__restrict__ in, uint64_t n, int shift,
                                                          uint32_t *__restrict__ counts, uint32_t ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
#pragma unroll
  for (int r = 0; r < SORT_IPT; r++) {
    const uint64_t i = base + r * SORT_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(in[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile by digit.  offs = exclusive scan of counts.
// Round r takes tile elements [256 r, 256 r + 256); within a wave the rank
// among equal digits comes from 8 ballots, across waves from LDS counts.
template <typename K>
__global__ __launch_bounds__(SORT_BLOCK) void k_sort_scatter(const K *__restrict__ in, K *__restrict__ out, uint64_t n,
                                                             int shift, const uint32_t *__restrict__ offs,
                                                             uint32_t ntiles) {
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_w[SORT_BLOCK / 64][256];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t below = (1ull << lane) - 1ull;
  s_base[threadIdx.x] = offs[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
  const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
  for (int r = 0; r < SORT_IPT; r++) {
#pragma unroll
    for (int q = 0; q < SORT_BLOCK / 64; q++) s_w[q][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t i = base + r * SORT_BLOCK + threadIdx.x;
    const bool valid = i < n;
    const K k = valid ? in[i] : (K)0;
    const uint32_t d = (uint32_t)(k >> shift) & 255u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(m & below);
    if (valid && rank == 0) s_w[w][d] = (uint32_t)__popcll(m);
    __syncthreads();
    if (valid) {
      uint32_t pos = s_base[d] + rank;
      for (uint32_t q = 0; q < w; q++) pos += s_w[q][d];
      out[pos] = k;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int q = 0; q < SORT_BLOCK / 64; q++) add += s_w[q][threadIdx.x];
    s_base[threadIdx.x] += add;
    __syncthreads();
  }
}

// Deterministic block sum of NV doubles (fixed tree order).
template <int NV>
__device__ __forceinline__ void block_sum_d(double (&v)[NV], double *s_red /* 256 * NV */) {
#pragma unroll
  for (int j = 0; j < NV; j++) s_red[j * 256 + threadIdx.x] = v[j];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < (uint32_t)st)
#pragma unroll
      for (int j = 0; j < NV; j++) s_red[j * 256 + threadIdx.x] += s_red[j * 256 + threadIdx.x + st];
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < NV; j++) v[j] = s_red[j * 256];
  __syncthreads();
}

// block_sum_d's 256-thread tree computed by one wave: lane i holds virtual
// threads i + 64 q (q = 0..3) in a[q]; every level adds the same two values
// as the shared-memory tree (st = 128 and 64 inside a lane, 32..1 across
// lanes), so the total (returned in every lane) is block_sum_d's, bit for bit.
template <int NV>
__device__ __forceinline__ void wave_tree_256(double (&a)[4][NV]) {
  const uint32_t lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < NV; j++) {
    a[0][j] += a[2][j];
    a[1][j] += a[3][j];
    a[0][j] += a[1][j];
  }
#pragma unroll
  for (int st = 32; st >= 1; st >>= 1)
#pragma unroll
    for (int j = 0; j < NV; j++) {
      const double o = __shfl_down(a[0][j], st, 64);
      if (lane < (uint32_t)st) a[0][j] += o;
    }
#pragma unroll
  for (int j = 0; j < NV; j++) a[0][j] = __shfl(a[0][j], 0, 64);
}

// Root candidates.  The root is linear in one of two features of the key
// (raw distance from the smallest key, or the bits of its float: a
// piecewise-linear log); for each feature it is fitted either by least
// squares of (i + 1/2) L / n on the feature, or through the feature's range
// (min -> 0, max -> L).  The candidate whose leaf loads on the sample have
// the smallest sum of squares is kept.
struct RootCand {
  uint32_t kind;
  float slope, icpt;
  uint32_t pad;
};

// Sums of x, y, x^2, x*y for both features (8 doubles) of moment block blk
// of nblk (elements blk*256 + t, stepping nblk*256), in v of every thread
// (256 threads; fixed tree order).  Shared by the multi-kernel and the
// one-workgroup trainers, so both give the same model bit for bit.
// One (virtual) thread t's moment sums of block blk: elements blk*256 + t,
// stepping nblk*256.
template <typename K>
__device__ __forceinline__ void root_moments_acc(const K *sorted, uint64_t n, double yscale, uint64_t blk,
                                                 uint64_t nblk, uint32_t t, double (&v)[8]) {
  const uint64_t kmin = (uint64_t)sorted[0];
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = 0.0;
  for (uint64_t i = blk * 256 + t; i < n; i += nblk * 256) {
    const double y = ((double)i + 0.5) * yscale;
#pragma unroll
    for (uint32_t f = 0; f < 2; f++) {
      const double x = (double)rmi_feature((uint64_t)sorted[i], kmin, f);
      v[4 * f + 0] += x;
      v[4 * f + 1] += y;
      v[4 * f + 2] += x * x;
      v[4 * f + 3] += x * y;
    }
  }
}

template <typename K>
__device__ __forceinline__ void root_moments_block(const K *sorted, uint64_t n, double yscale, uint64_t blk,
                                                   uint64_t nblk, double (&v)[8], double *s_red) {
  root_moments_acc<K>(sorted, n, yscale, blk, nblk, threadIdx.x, v);
  block_sum_d<8>(v, s_red);
}

// The same sums with one wave per block (wave_tree_256): every lane gets v.
template <typename K>
__device__ __forceinline__ void root_moments_block_wave(const K *sorted, uint64_t n, double yscale, uint64_t blk,
                                                        uint64_t nblk, double (&v)[8]) {
  double a[4][8];
#pragma unroll
  for (int q = 0; q < 4; q++) root_moments_acc<K>(sorted, n, yscale, blk, nblk, (threadIdx.x & 63) + 64 * q, a[q]);
  wave_tree_256<8>(a);
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = a[0][j];
}

// Per block: sums of x, y, x^2, x*y for both features (8 doubles).
template <typename K>
__global__ __launch_bounds__(256) void k_rmi_root_moments(const K *__restrict__ sorted, uint64_t n, double yscale,
                                                          double *__restrict__ part) {
  __shared__ double s_red[8 * 256];
  double v[8];
  root_moments_block<K>(sorted, n, yscale, blockIdx.x, gridDim.x, v, s_red);
  if (threadIdx.x < 8) part[(uint64_t)blockIdx.x * 8 + threadIdx.x] = v[threadIdx.x];
}

// Closed-form least squares y = alpha + beta x from the partial sums; a
// non-positive or undefined slope becomes 0 (the root stays monotone).
__device__ __forceinline__ void lsq_line(double n, double sx, double sy, double sxx, double sxy, double &alpha,
                                         double &beta) {
  const double mx = sx / n, my = sy / n;
  const double vxx = sxx - sx * mx, vxy = sxy - sx * my;
  beta = vxx > 0.0 ? vxy / vxx : 0.0;
  if (!(beta > 0.0) || !(beta < 3.0e38)) beta = 0.0;
  alpha = my - beta * mx;
}

// The four root candidates and the header (key range, sizes) from the moment
// blocks' partial sums (256 threads; thread 0 writes cand and h).
template <typename K>
__device__ __forceinline__ void root_fit_from(const K *sorted, uint64_t n, const double *part, uint32_t nblk,
                                              uint32_t L, uint32_t M, RmiHeader *h, RootCand *cand, double *s_red) {
  double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t b = threadIdx.x; b < nblk; b += 256)
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] += part[(uint64_t)b * 8 + j];
  block_sum_d<8>(v, s_red);
  if (threadIdx.x == 0) {
    const uint64_t kmin = (uint64_t)sorted[0], kmax = (uint64_t)sorted[n - 1];
    for (uint32_t f = 0; f < 2; f++) {
      double alpha, beta;
      lsq_line((double)n, v[4 * f], v[4 * f + 1], v[4 * f + 2], v[4 * f + 3], alpha, beta);
      RootCand c{f, (float)beta, (float)alpha, 0};
      cand[2 * f] = c;
      const double lo = (double)rmi_feature(kmin, kmin, f), hi = (double)rmi_feature(kmax, kmin, f);
      const double s = hi > lo ? (double)L * (1.0 - 1e-6) / (hi - lo) : 0.0;
      RootCand r{f, (float)s, (float)(-s * lo), 0};
      cand[2 * f + 1] = r;
    }
    RmiHeader hd{};
    hd.key_min = kmin;
    hd.n_leaves = L;
    hd.n_bins = M;
    *h = hd;
  }
}

template <typename K>
__global__ __launch_bounds__(256) void k_rmi_root_fit(const K *__restrict__ sorted, uint64_t n, const double *part,
                                                      uint32_t nblk, uint32_t L, uint32_t M, RmiHeader *__restrict__ h,
                                                      RootCand *__restrict__ cand) {
  __shared__ double s_red[8 * 256];
  root_fit_from<K>(sorted, n, part, nblk, L, M, h, cand, s_red);
}

// Leaf loads of the sample under every candidate: cnt[c * L + leaf].
template <typename K>
__global__ __launch_bounds__(256) void k_rmi_root_score(const K *__restrict__ sorted, uint64_t n,
                                                        const RootCand *__restrict__ cand, uint32_t L,
                                                        uint32_t *__restrict__ cnt) {
  const uint64_t kmin = (uint64_t)sorted[0];
  RootCand c[4];
#pragma unroll
  for (int j = 0; j < 4; j++) c[j] = cand[j];
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint64_t k = (uint64_t)sorted[i];
    const float fr = rmi_feature(k, kmin, 0), fl = rmi_feature(k, kmin, 1);
#pragma unroll
    for (int j = 0; j < 4; j++)
      atomicAdd(&cnt[(uint32_t)j * L + rmi_leaf_at(c[j].kind ? fl : fr, c[j].slope, c[j].icpt, L)], 1u);
  }
}

// Keeps the candidate with the smallest sum of squared leaf loads.
__global__ __launch_bounds__(1024) void k_rmi_root_pick(const uint32_t *__restrict__ cnt, uint32_t L,
                                                        const RootCand *__restrict__ cand, RmiHeader *__restrict__ h) {
  __shared__ unsigned long long s_t[1024 / 64 + 1];
  __shared__ unsigned long long s_sc[4];
  for (int j = 0; j < 4; j++) {
    unsigned long long q = 0;
    for (uint32_t l = threadIdx.x; l < L; l += 1024) {
      const unsigned long long c = cnt[(uint32_t)j * L + l];
      q += c * c;
    }
    q = block_sum<1024, unsigned long long>(q, s_t);
    if (threadIdx.x == 0) s_sc[j] = q;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int best = 0;
    for (int j = 1; j < 4; j++)
      if (s_sc[j] < s_sc[best]) best = j;
    h->root_kind = cand[best].kind;
    h->root_slope = cand[best].slope;
    h->root_icpt = cand[best].icpt;
  }
}

// start[l] = first sample position routed to leaf >= l (start[L] = n).
template <typename K>
__global__ __launch_bounds__(256) void k_rmi_bounds(const K *__restrict__ sorted, uint64_t n,
                                                    const RmiHeader *__restrict__ hp, uint32_t *__restrict__ start) {
  const RmiHeader h = *hp;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * 256) {
    const int64_t cur = i < n ? (int64_t)rmi_leaf((uint64_t)sorted[i], h) : (int64_t)h.n_leaves;
    const int64_t prev = i > 0 ? (int64_t)rmi_leaf((uint64_t)sorted[i - 1], h) : -1;
    for (int64_t l = prev + 1; l <= cur; l++) start[l] = (uint32_t)i;
  }
}

// One workgroup per leaf: least squares of the rank position (i + 1/2) M / n
// on the signed distance from the leaf's first key.  A sparse leaf (fewer
// than 8 keys or a single value) also takes the sample keys on either side,
// so that its line follows its neighbourhood.
// A leaf's fitting range: the sample positions [a, b) routed to it, widened
// by one key on each side when the leaf is sparse (fewer than 8 keys or a
// single value), and its first key x0.
template <typename K>
__device__ __forceinline__ void leaf_range(const K *sorted, uint64_t n, uint32_t a, uint32_t b, uint64_t &x0,
                                           uint32_t &lo, uint32_t &hi) {
  x0 = (uint64_t)sorted[a < n ? a : n - 1];
  const bool sparse = b - a < 8 || (uint64_t)sorted[b - 1] == x0;
  lo = sparse && a > 0 ? a - 1 : a;
  hi = sparse && b < n ? b + 1 : b;
}

// One (virtual) thread t's sums of x, y, x^2, x*y over the leaf's range.
template <typename K>
__device__ __forceinline__ void leaf_acc(const K *sorted, uint64_t x0, uint32_t lo, uint32_t hi, double yscale,
                                         uint32_t t, double (&v)[4]) {
  v[0] = v[1] = v[2] = v[3] = 0.0;
  for (uint32_t i = lo + t; i < hi; i += 256) {
    const uint64_t k = (uint64_t)sorted[i];
    const double x = k >= x0 ? (double)(k - x0) : -(double)(x0 - k);
    const double y = ((double)i + 0.5) * yscale;
    v[0] += x;
    v[1] += y;
    v[2] += x * x;
    v[3] += x * y;
  }
}

// The leaf's line from its sums: least squares of the rank position on the
// signed distance from x0, the anchor rounded and clamped to [0, M-1].
__device__ __forceinline__ RmiLeaf leaf_line(const double (&v)[4], uint64_t n, uint32_t a, uint32_t lo, uint32_t hi,
                                             uint64_t x0, uint32_t M) {
  const double yscale = (double)M / (double)n;
  double alpha, beta;
  if (hi > lo) {
    lsq_line((double)(hi - lo), v[0], v[1], v[2], v[3], alpha, beta);
  } else {
    alpha = ((double)a + 0.5) * yscale;
    beta = 0.0;
  }
  double an = floor(alpha + 0.5);
  an = an < 0.0 ? 0.0 : (an > (double)(M - 1) ? (double)(M - 1) : an);
  RmiLeaf lf;
  lf.x0 = x0;
  lf.slope = (float)beta;
  lf.anchor = (uint32_t)an;
  return lf;
}

template <typename K>
__device__ __forceinline__ void leaf_fit_one(const K *sorted, uint64_t n, uint32_t a, uint32_t b, uint32_t M,
                                             RmiLeaf *leaf, double *s_red) {
  const double yscale = (double)M / (double)n;
  uint64_t x0;
  uint32_t lo, hi;
  leaf_range<K>(sorted, n, a, b, x0, lo, hi);
  double v[4];
  leaf_acc<K>(sorted, x0, lo, hi, yscale, threadIdx.x, v);
  block_sum_d<4>(v, s_red);
  if (threadIdx.x == 0) *leaf = leaf_line(v, n, a, lo, hi, x0, M);
}

// The same fit by one wave (wave_tree_256); lane 0 writes the leaf.
template <typename K>
__device__ __forceinline__ void leaf_fit_one_wave(const K *sorted, uint64_t n, uint32_t a, uint32_t b, uint32_t M,
                                                  RmiLeaf *leaf) {
  const double yscale = (double)M / (double)n;
  uint64_t x0;
  uint32_t lo, hi;
  leaf_range<K>(sorted, n, a, b, x0, lo, hi);
  double acc[4][4];
#pragma unroll
  for (int q = 0; q < 4; q++) leaf_acc<K>(sorted, x0, lo, hi, yscale, (threadIdx.x & 63) + 64 * q, acc[q]);
  wave_tree_256<4>(acc);
  if ((threadIdx.x & 63) == 0) *leaf = leaf_line(acc[0], n, a, lo, hi, x0, M);
}

template <typename K>
__global__ __launch_bounds__(256) void k_rmi_leaf_fit(const K *__restrict__ sorted, uint64_t n,
                                                      const uint32_t *__restrict__ start, uint32_t M,
                                                      RmiLeaf *__restrict__ leaves) {
  __shared__ double s_red[4 * 256];
  const uint32_t l = blockIdx.x;
  leaf_fit_one<K>(sorted, n, start[l], start[l + 1], M, &leaves[l], s_red);
}

// Both hashes bin the sample; bins are scaled to Mp = M * n_sample / n_keys
// so that the sample sees the table's load factor.
template <typename K>
__global__ __launch_bounds__(256) void k_rmi_check(const K *__restrict__ keys, uint64_t n, const void *model,
                                                   BinMap mm, uint32_t Mp, uint32_t *__restrict__ cnt_r,
                                                   uint32_t *__restrict__ cnt_m) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const K k = keys[i];
    const uint64_t br = rmi_bin((uint64_t)k, model, mm.M);
    const uint64_t bm = bin_of<K, HK_MURMUR>(k, mm);
    atomicAdd(&cnt_r[br * Mp / mm.M], 1u);
    atomicAdd(&cnt_m[bm * Mp / mm.M], 1u);
  }
}

// out[0] = max count, out[1] = sum of squares (learned); out[2], out[3] murmur.
__global__ __launch_bounds__(256) void k_rmi_check_reduce(const uint32_t *__restrict__ cnt_r,
                                                          const uint32_t *__restrict__ cnt_m, uint32_t Mp,
                                                          unsigned long long *__restrict__ out) {
  __shared__ unsigned long long s_t[8];
  unsigned long long mr = 0, sr = 0, mm = 0, sm = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < Mp; i += (uint64_t)gridDim.x * 256) {
    const unsigned long long a = cnt_r[i], b = cnt_m[i];
    mr = a > mr ? a : mr;
    mm = b > mm ? b : mm;
    sr += a * a;
    sm += b * b;
  }
  sr = wave_sum(sr);
  sm = wave_sum(sm);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long x = __shfl_xor(mr, d, 64), y = __shfl_xor(mm, d, 64);
    mr = x > mr ? x : mr;
    mm = y > mm ? y : mm;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&out[0], mr);
    atomicAdd(&out[1], sr);
    atomicMax(&out[2], mm);
    atomicAdd(&out[3], sm);
  }
  (void)s_t;
}

// ---------------------------------------------------------------------------
// Batched training: one 256-thread workgroup per table, the whole pipeline
// in LDS (sample, bitonic sort, root moments and candidates, leaf loads and
// pick, leaf bounds, leaf fits, sampled check).  The floating-point steps
// are the multi-kernel trainer's own device functions in the same order
// (moment blocks of 256 keys, fixed tree sums), so a table's model is the
// one hg_rmi_train fits, bit for bit.
// ---------------------------------------------------------------------------
template <typename K>
__global__ __launch_bounds__(256) void k_rmi_train_one(const K *__restrict__ keys,
                                                       const RmiBatchItem *__restrict__ items, uint32_t L,
                                                       uint32_t keys_cap, uint32_t cnt_cap, uint8_t *__restrict__ models,
                                                       uint32_t *__restrict__ use,
                                                       unsigned long long *__restrict__ report) {
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ double s_red[8 * 256];
  __shared__ double s_part[RMI_ONE_MAX_BLOCKS * 8];
  __shared__ RootCand s_cand[4];
  __shared__ RmiHeader s_h;
  __shared__ unsigned long long s_t[256 / 64 + 1];
  __shared__ unsigned long long s_sc[4];
  const RmiBatchItem it = items[blockIdx.x];
  K *s_keys = (K *)smem;                                             // keys_cap
  RmiLeaf *s_leaf = (RmiLeaf *)(smem + (size_t)keys_cap * sizeof(K));  // L
  uint32_t *s_start = (uint32_t *)(s_leaf + L);                      // L + 1 (rounded to 4)
  uint32_t *s_cnt = s_start + ((L + 4) & ~3u);                       // cnt_cap
  const uint32_t ns = it.ns, M = it.map.M;
  // 1. sample (as k_sample) padded to a power of two with the largest key
  const K *kp = keys + it.key_begin;
  for (uint32_t i = threadIdx.x; i < it.np2; i += 256) {
    K v = (K)~(K)0;
    if (i < ns) v = kp[__umul64hi(mix64(it.s0 + (uint64_t)(i + 1) * SPLITMIX_GAMMA), it.n_keys)];
    s_keys[i] = v;
  }
  __syncthreads();
  // 2. bitonic sort (the sorted sample is unique whatever the method)
  for (uint32_t k = 2; k <= it.np2; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < it.np2; i += 256) {
        const uint32_t p = i ^ j;
        if (p > i) {
          const K a = s_keys[i], b = s_keys[p];
          if ((a > b) == ((i & k) == 0)) {
            s_keys[i] = b;
            s_keys[p] = a;
          }
        }
      }
      __syncthreads();
    }
  // 3. root candidates (k_rmi_root_moments + k_rmi_root_fit)
  const uint32_t nblk = (ns + 255) / 256;
  const double yscale = (double)L / (double)ns;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t b = wave; b < nblk; b += 4) {  // one wave per moment block
    double v[8];
    root_moments_block_wave<K>(s_keys, ns, yscale, b, nblk, v);
    if (lane < 8) s_part[b * 8 + lane] = v[lane];
  }
  __syncthreads();
  root_fit_from<K>(s_keys, ns, s_part, nblk, L, M, &s_h, s_cand, s_red);
  __syncthreads();
  // 4. leaf loads of the candidates and the pick (k_rmi_root_score + pick)
  for (uint32_t i = threadIdx.x; i < 4 * L; i += 256) s_cnt[i] = 0;
  __syncthreads();
  {
    const uint64_t kmin = (uint64_t)s_keys[0];
    for (uint32_t i = threadIdx.x; i < ns; i += 256) {
      const uint64_t k = (uint64_t)s_keys[i];
      const float fr = rmi_feature(k, kmin, 0), fl = rmi_feature(k, kmin, 1);
#pragma unroll
      for (int j = 0; j < 4; j++)
        atomicAdd(&s_cnt[(uint32_t)j * L + rmi_leaf_at(s_cand[j].kind ? fl : fr, s_cand[j].slope, s_cand[j].icpt, L)],
                  1u);
    }
  }
  __syncthreads();
  for (int j = 0; j < 4; j++) {
    unsigned long long q = 0;
    for (uint32_t l = threadIdx.x; l < L; l += 256) {
      const unsigned long long c = s_cnt[(uint32_t)j * L + l];
      q += c * c;
    }
    q = block_sum<256, unsigned long long>(q, s_t);
    if (threadIdx.x == 0) s_sc[j] = q;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int best = 0;
    for (int j = 1; j < 4; j++)
      if (s_sc[j] < s_sc[best]) best = j;
    s_h.root_kind = s_cand[best].kind;
    s_h.root_slope = s_cand[best].slope;
    s_h.root_icpt = s_cand[best].icpt;
  }
  __syncthreads();
  // 5. leaf bounds of the sorted sample (k_rmi_bounds)
  {
    const RmiHeader h = s_h;
    for (uint32_t i = threadIdx.x; i <= ns; i += 256) {
      const int64_t cur = i < ns ? (int64_t)rmi_leaf((uint64_t)s_keys[i], h) : (int64_t)h.n_leaves;
      const int64_t prev = i > 0 ? (int64_t)rmi_leaf((uint64_t)s_keys[i - 1], h) : -1;
      for (int64_t l = prev + 1; l <= cur; l++) s_start[l] = i;
    }
  }
  __syncthreads();
  // 6. leaf fits (k_rmi_leaf_fit's sums, one wave per leaf)
  for (uint32_t l = wave; l < L; l += 4) leaf_fit_one_wave<K>(s_keys, ns, s_start[l], s_start[l + 1], M, &s_leaf[l]);
  __syncthreads();
  // 7. the model to global memory
  {
    uint32_t *dst = (uint32_t *)(models + it.model_off);
    const uint32_t *hw = (const uint32_t *)&s_h, *lw = (const uint32_t *)s_leaf;
    const uint32_t hn = (uint32_t)(sizeof(RmiHeader) / 4), ln = L * (uint32_t)(sizeof(RmiLeaf) / 4);
    for (uint32_t i = threadIdx.x; i < hn + ln; i += 256) dst[i] = i < hn ? hw[i] : lw[i - hn];
  }
  // 8. the sampled check (k_rmi_check + reduce): learned bins, then murmur3
  unsigned long long res[4];
  for (int pass = 0; pass < 2; pass++) {
    for (uint32_t i = threadIdx.x; i < it.Mp; i += 256) s_cnt[i] = 0;
    __syncthreads();
    const RmiHeader h = s_h;
    for (uint32_t i = threadIdx.x; i < ns; i += 256) {
      const K k = s_keys[i];
      const uint64_t b = pass == 0 ? (uint64_t)rmi_leaf_bin((uint64_t)k, s_leaf[rmi_leaf((uint64_t)k, h)], M)
                                   : (uint64_t)bin_of<K, HK_MURMUR>(k, it.map);
      atomicAdd(&s_cnt[b * it.Mp / M], 1u);
    }
    __syncthreads();
    unsigned long long mx = 0, sq = 0;
    for (uint32_t i = threadIdx.x; i < it.Mp; i += 256) {
      const unsigned long long c = s_cnt[i];
      mx = c > mx ? c : mx;
      sq += c * c;
    }
    sq = block_sum<256, unsigned long long>(sq, s_t);
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      const unsigned long long x = __shfl_xor(mx, d, 64);
      mx = x > mx ? x : mx;
    }
    if ((threadIdx.x & 63) == 0) s_t[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = s_t[0];
    for (int w = 1; w < 4; w++) mx = s_t[w] > mx ? s_t[w] : mx;
    res[2 * pass] = mx;
    res[2 * pass + 1] = sq;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    use[it.table] = (res[0] <= res[2] && res[1] <= res[3]) ? 1u : 0u;
    if (report)
      for (int j = 0; j < 4; j++) report[(uint64_t)it.table * 4 + j] = res[j];
  }
  (void)cnt_cap;
}

static uint32_t grid_for(uint64_t n, uint32_t cap) {
  const uint64_t g = (n + 255) / 256;
  return (uint32_t)(g == 0 ? 1 : (g < cap ? g : cap));
}

template <typename K>
static hipError_t launch_rmi_train_t(const LaunchRmi &r, hipStream_t st) {
  const uint64_t ns = r.n_sample;
  K *a = (K *)r.buf0, *b = (K *)r.buf1;
  hipLaunchKernelGGL((k_sample<K>), dim3(grid_for(ns, 4096)), dim3(256), 0, st, (const K *)r.keys, r.n_keys, r.s0, ns,
                     a);
  const uint32_t ntiles = (uint32_t)((ns + SORT_TILE - 1) / SORT_TILE);
  for (int pass = 0; pass < (int)sizeof(K); pass++) {
    const int shift = pass * 8;
    hipLaunchKernelGGL((k_sort_hist<K>), dim3(ntiles), dim3(SORT_BLOCK), 0, st, (const K *)a, ns, shift, r.counts,
                       ntiles);
    hipError_t e = launch_scan(r.counts, r.counts, (uint64_t)256 * ntiles, 4, r.scan_status, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_sort_scatter<K>), dim3(ntiles), dim3(SORT_BLOCK), 0, st, (const K *)a, b, ns, shift,
                       (const uint32_t *)r.counts, ntiles);
    K *t = a;
    a = b;
    b = t;
  }
  // sizeof(K) passes (even): the sorted sample is back in buf0
  RmiHeader *h = (RmiHeader *)r.model;
  RmiLeaf *leaves = (RmiLeaf *)(h + 1);
  RootCand *cand = (RootCand *)r.cand;
  hipLaunchKernelGGL((k_rmi_root_moments<K>), dim3(r.nblk), dim3(256), 0, st, (const K *)a, ns,
                     (double)r.n_leaves / (double)ns, r.partials);
  hipLaunchKernelGGL((k_rmi_root_fit<K>), dim3(1), dim3(256), 0, st, (const K *)a, ns, (const double *)r.partials, r.nblk,
                     r.n_leaves, r.map.M, h, cand);
  hipError_t e0 = launch_fill_u32(r.leafcnt, 0u, (uint64_t)4 * r.n_leaves, st);
  if (e0 != hipSuccess) return e0;
  hipLaunchKernelGGL((k_rmi_root_score<K>), dim3(grid_for(ns, 2048)), dim3(256), 0, st, (const K *)a, ns,
                     (const RootCand *)cand, r.n_leaves, r.leafcnt);
  hipLaunchKernelGGL(k_rmi_root_pick, dim3(1), dim3(1024), 0, st, (const uint32_t *)r.leafcnt, r.n_leaves,
                     (const RootCand *)cand, h);
  hipLaunchKernelGGL((k_rmi_bounds<K>), dim3(grid_for(ns + 1, 4096)), dim3(256), 0, st, (const K *)a, ns,
                     (const RmiHeader *)h, r.start);
  hipLaunchKernelGGL((k_rmi_leaf_fit<K>), dim3(r.n_leaves), dim3(256), 0, st, (const K *)a, ns,
                     (const uint32_t *)r.start, r.map.M, leaves);
  // fallback check on the sample
  hipError_t e = launch_fill_u32(r.cnt_r, 0u, r.Mp, st);
  if (e == hipSuccess) e = launch_fill_u32(r.cnt_m, 0u, r.Mp, st);
  if (e == hipSuccess) e = launch_fill_u64(r.out, 0ull, 4, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_rmi_check<K>), dim3(grid_for(ns, 4096)), dim3(256), 0, st, (const K *)a, ns,
                     (const void *)r.model, r.map, r.Mp, r.cnt_r, r.cnt_m);
  hipLaunchKernelGGL(k_rmi_check_reduce, dim3(grid_for(r.Mp, 2048)), dim3(256), 0, st, (const uint32_t *)r.cnt_r,
                     (const uint32_t *)r.cnt_m, r.Mp, r.out);
  return hipGetLastError();
}

hipError_t launch_rmi_train(const LaunchRmi &r, hipStream_t st) {
  return r.key_bytes == 4 ? launch_rmi_train_t<uint32_t>(r, st) : launch_rmi_train_t<uint64_t>(r, st);
}

size_t rmi_one_lds(uint32_t keys_cap, int key_bytes, uint32_t L, uint32_t cnt_cap) {
  return (size_t)keys_cap * key_bytes + (size_t)L * sizeof(RmiLeaf) + (size_t)((L + 4) & ~3u) * 4 + (size_t)cnt_cap * 4;
}

hipError_t launch_rmi_train_batch(const LaunchRmiBatch &b, hipStream_t st) {
  if (!b.n_items) return hipSuccess;
  const size_t lds = rmi_one_lds(b.keys_cap, b.key_bytes, b.n_leaves, b.cnt_cap);
  if (b.key_bytes == 4)
    hipLaunchKernelGGL((k_rmi_train_one<uint32_t>), dim3(b.n_items), dim3(256), lds, st, (const uint32_t *)b.keys, b.items,
                       b.n_leaves, b.keys_cap, b.cnt_cap, (uint8_t *)b.models, b.use, b.report);
  else
    hipLaunchKernelGGL((k_rmi_train_one<uint64_t>), dim3(b.n_items), dim3(256), lds, st, (const uint64_t *)b.keys, b.items,
                       b.n_leaves, b.keys_cap, b.cnt_cap, (uint8_t *)b.models, b.use, b.report);
  return hipGetLastError();
}

hipError_t set_rmi_kernel_attributes() {
  for (const void *fn : {(const void *)&k_rmi_train_one<uint32_t>, (const void *)&k_rmi_train_one<uint64_t>}) {
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, fn);
    if (e != hipSuccess) return e;
    int dev = 0, lim = 0;
    (void)hipGetDevice(&dev);
    e = hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    const int room = lim - (int)fa.sharedSizeBytes;
    if (room > 0) e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, room);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

uint32_t rmi_moment_blocks(uint64_t n_sample) { return grid_for(n_sample, 1024); }
uint32_t rmi_sort_tiles(uint64_t n_sample) { return (uint32_t)((n_sample + SORT_TILE - 1) / SORT_TILE); }

}  // namespace hg
