// hg_kernels.hip -- gfx950 kernels of the HashGraph build / probe / join.
//
// Build = partition pass (hash -> coarse bucket, per-tile LDS histogram,
// exact count matrix, column scan, LDS-staged coalesced scatter) followed by
// the final stage (one workgroup per bucket: keys held in VGPRs, packed u16
// LDS counters give the bucket's CSR offsets and every key's slot, keys are
// staged through LDS and written coalesced).  Replaces the reference's LRB
// build (src/SingleHashGraph.cu:129-181, kernels
// include/SingleHashGraphOperators.cuh:71-154), whose random global atomics
// are replaced by LDS work and deterministic positions.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "hg_common.hpp"
#include "hg_device.hpp"
#include "hg_kernels.hpp"

namespace hg {

// Bins longer than this are not walked by the probe kernels but answered from
// a value table by k_probe_heavy (duplicates: a walk would cost
// |A_bin| x |B_bin| comparisons).
constexpr uint64_t PROBE_BIN_LIMIT = 32;
// Leaves of a learned model a bucket-local kernel stages in LDS (the range
// between its smallest and largest key's leaves; wider ranges are read from
// global memory).
constexpr uint32_t RANGE_LEAVES = 512;
// A join bin pair with more key pairs than this is answered from a value
// table (k_join_heavy) instead of the nested loop.
constexpr uint64_t JOIN_PAIR_LIMIT = 4096;
// bins per k_join workgroup: with their offsets and keys (78 KB of LDS) two
// 1024-thread workgroups fit a CU
constexpr int JOIN_BINS = 2048;
// matches an emitting k_join workgroup buffers in LDS to write them out
// contiguously (a range with more re-walks its bins): the row-id join keeps a
// pair of 16-bit LDS slots per match, the key join B's slot only
constexpr uint32_t JOIN_EMIT_PAIRS = 1024, JOIN_EMIT_KEYS = 3072;
// what a k_join instance does with its matches: count them, write the matched
// keys, or write the pairs of input positions (row-id join)
constexpr int JOIN_COUNT = 0, JOIN_KEYS = 1, JOIN_PAIRS = 2;

// Keys per thread that the final stage and the probe keep in VGPRs (u32
// keys; u64 keys use 32).  1024 threads x KPT is the bucket capacity.
#ifndef HG_KPT_U32
#define HG_KPT_U32 32
#endif
constexpr int KPT_U32 = HG_KPT_U32;

// Probe staging: bins per LDS sub-range (by offset width) and the LDS bytes
// of one probe workgroup (two workgroups fit a CU's 160 KB).
#ifndef HG_PROBE_LDS
#define HG_PROBE_LDS (78 * 1024)
#endif
template <typename OFF> constexpr int probe_sub() { return sizeof(OFF) == 4 ? 8192 : 4096; }
constexpr size_t PROBE_LDS = HG_PROBE_LDS;

// Oversized buckets handled by the multi-workgroup path (per launch), their
// chunk size in keys and the bins per LDS pass.
constexpr uint32_t BIG_CAP = 16384;
constexpr uint32_t BIG_CHUNK = 65536;
constexpr uint32_t BIG_BINS = 32768;
constexpr int BIG_UB = 8;  // keys per thread whose loads the oversized-bucket kernels keep in flight

// Workgroup b of a partition pass runs on XCD b % 8 (round-robin dispatch).
// Tiles are handed out so that each XCD walks a contiguous range of tiles:
// the runs that concurrently running tiles write for one bucket are then
// adjacent and written from one XCD, so they merge in that XCD's L2 instead
// of reaching HBM as partial lines.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nt) {
  const uint32_t q = nt >> 3, r = nt & 7, x = b & 7, i = b >> 3;
  return x * q + min(x, r) + i;
}

// ---------------------------------------------------------------------------
// Partition pass: histogram.  One workgroup per tile; tile t of segment s
// writes row t of the segment's count matrix (ncols u32).
// ---------------------------------------------------------------------------
template <typename K, int BLOCK, int KPT, int HK, int SK>
__global__ __launch_bounds__(BLOCK) void k_part_hist(const K *__restrict__ keys, const TileDesc *__restrict__ tiles,
                                                     const SegDesc *__restrict__ segs, uint32_t *__restrict__ cmat,
                                                     uint32_t leaf_cap) {
  extern __shared__ __align__(16) uint32_t s_cnt[];
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const TileDesc td = tiles[tile];
  const SegDesc sd = segs[td.seg];  // a copy: uniform fields stay in SGPRs
  const uint32_t ncols = sd.ncols;
  __shared__ int64_t s_spl[SK == SK_RANGE ? 1 : MAX_SPLITS + 1];
  for (uint32_t i = threadIdx.x; i < ncols; i += BLOCK) s_cnt[i] = 0;
  const int64_t *spl = stage_splits<SK, BLOCK>(sd, s_spl);  // ends with a barrier
  // a tile is processed in rounds of BLOCK*KPT keys (slot j of a round holds
  // key j*BLOCK + tid of the round)
  if constexpr (HK == HK_RMI) {
    // learned model: its leaves are read from LDS, in windows of leaf_cap
    // (one window for models up to leaf_cap leaves; each further window
    // reads the tile again)
    RmiLeaf *s_leaf = (RmiLeaf *)(s_cnt + ((ncols + 3) & ~3u));
    const RmiHeader h = *(const RmiHeader *)sd.map.rmi;
    for (uint32_t w0 = 0; w0 < h.n_leaves; w0 += leaf_cap) {
      const uint32_t n = min(leaf_cap, h.n_leaves - w0);
      if (w0) __syncthreads();  // the previous window's readers are done
      rmi_copy_leaves<BLOCK>(sd.map.rmi, s_leaf, w0, n);
      __syncthreads();
      for (uint32_t base = 0; base < td.count; base += BLOCK * KPT) {
        const uint32_t cnt = min((uint32_t)(BLOCK * KPT), td.count - base);
        const K *kp = keys + td.key_start + base;
        const uint32_t nv = cnt > threadIdx.x ? (cnt - threadIdx.x + BLOCK - 1) / BLOCK : 0;
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(kp, cnt * (uint32_t)sizeof(K));
        K kr[KPT];
#pragma unroll
        for (int j = 0; j < KPT; j++)
          kr[j] = buf_load<K>(rs, threadIdx.x * (uint32_t)sizeof(K), j * BLOCK * (uint32_t)sizeof(K));
        rmi_window_bins<K, KPT>(kr, nv, h, s_leaf, w0, n, sd.map.M, [&](int, uint32_t bin) {
          atomicAdd(&s_cnt[bucket_of_bin<SK>(bin, sd, spl)], 1u);
        });
      }
    }
  } else {
    for (uint32_t base = 0; base < td.count; base += BLOCK * KPT) {
      const uint32_t cnt = min((uint32_t)(BLOCK * KPT), td.count - base);
      const K *kp = keys + td.key_start + base;
      const uint32_t nv = cnt > threadIdx.x ? (cnt - threadIdx.x + BLOCK - 1) / BLOCK : 0;
      const __amdgpu_buffer_rsrc_t rs = make_rsrc(kp, cnt * (uint32_t)sizeof(K));
      K kr[KPT];
#pragma unroll
      for (int j = 0; j < KPT; j++)
        kr[j] = buf_load<K>(rs, threadIdx.x * (uint32_t)sizeof(K), j * BLOCK * (uint32_t)sizeof(K));
#pragma unroll
      for (int j = 0; j < KPT; j++)
        if (j < nv) atomicAdd(&s_cnt[bucket_of<K, HK, SK>(kr[j], sd, spl)], 1u);
    }
  }
  __syncthreads();
  uint32_t *row = cmat + sd.row_base + (uint64_t)(tile - sd.tile_begin) * ncols;
  for (uint32_t i = threadIdx.x; i < ncols; i += BLOCK) row[i] = s_cnt[i];
}

// ---------------------------------------------------------------------------
// Column scan of the count matrix (bucket-major order inside a segment):
//   pos[t][b] = bucket_start[b] + sum_{t' < t} C[t'][b]
// in three steps over chunks of CHUNK rows.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_col_chunksum(const uint32_t *__restrict__ cmat, const ColItem *__restrict__ items,
                                                      const SegDesc *__restrict__ segs, const uint64_t *__restrict__ psum_base,
                                                      uint32_t *__restrict__ psum) {
  const ColItem it = items[blockIdx.x];
  const SegDesc &sd = segs[it.seg];
  const uint32_t b = it.col_lo + threadIdx.x;
  if (b >= sd.ncols) return;
  const uint32_t *p = cmat + sd.row_base + (uint64_t)it.row_lo * sd.ncols + b;
  uint32_t s = 0;
  // rows in groups of 16: the group's loads are all in flight before the sums
  for (uint32_t r = it.row_lo; r < it.row_hi; r += 16, p += 16 * (uint64_t)sd.ncols) {
    const uint32_t n = min(16u, it.row_hi - r);
    uint32_t v[16];
#pragma unroll
    for (uint32_t u = 0; u < 16; u++) v[u] = u < n ? p[(uint64_t)u * sd.ncols] : 0u;
#pragma unroll
    for (uint32_t u = 0; u < 16; u++) s += v[u];
  }
  psum[psum_base[it.seg] + (uint64_t)it.chunk * sd.ncols + b] = s;
}

// Per column block of a segment: exclusive scan of the chunk sums down each
// column; the column total goes to coltot[bucket_base + b].
__global__ __launch_bounds__(256) void k_col_chunkscan(const ColItem *__restrict__ items, const SegDesc *__restrict__ segs,
                                                       const uint64_t *__restrict__ psum_base,
                                                       const uint32_t *__restrict__ nchunks, uint32_t *__restrict__ psum,
                                                       uint32_t *__restrict__ coltot) {
  const ColItem it = items[blockIdx.x];
  const SegDesc &sd = segs[it.seg];
  const uint32_t b = it.col_lo + threadIdx.x;
  if (b >= sd.ncols) return;
  const uint32_t nch = nchunks[it.seg];
  uint32_t *p = psum + psum_base[it.seg] + b;
  uint32_t run = 0;
  // chunks in groups of 16: loads first, then the stores
  for (uint32_t c = 0; c < nch; c += 16, p += 16 * (uint64_t)sd.ncols) {
    const uint32_t n = min(16u, nch - c);
    uint32_t v[16];
#pragma unroll
    for (uint32_t u = 0; u < 16; u++) v[u] = u < n ? p[(uint64_t)u * sd.ncols] : 0u;
#pragma unroll
    for (uint32_t u = 0; u < 16; u++)
      if (u < n) {
        p[(uint64_t)u * sd.ncols] = run;
        run += v[u];
      }
  }
  coltot[sd.bucket_base + b] = run;
}

// One workgroup per segment: bucket starts = key_begin + exclusive scan of
// the column totals (the drop column of a split segment gets no start).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_seg_bucketscan(const SegDesc *__restrict__ segs, const uint32_t *__restrict__ coltot,
                                                          uint32_t *__restrict__ bstart) {
  __shared__ uint32_t s_tmp[BLOCK / 64 + 1];
  const SegDesc &sd = segs[blockIdx.x];
  const uint32_t ncols = sd.ncols, nb = sd.nbuckets;
  const uint32_t per = (ncols + BLOCK - 1) / BLOCK;
  const uint32_t c0 = min(threadIdx.x * per, ncols), c1 = min(c0 + per, ncols);
  uint32_t local = 0;
  for (uint32_t b = c0; b < c1; b++)
    if (b < nb) local += coltot[sd.bucket_base + b];
  uint32_t tot;
  uint32_t base = (uint32_t)sd.key_begin + block_exclusive_sum<BLOCK, uint32_t>(local, s_tmp, tot);
  for (uint32_t b = c0; b < c1; b++) {
    if (b < nb) {
      const uint32_t v = coltot[sd.bucket_base + b];
      bstart[sd.bucket_base + b] = base;
      base += v;
    } else {
      bstart[sd.bucket_base + b] = 0xFFFFFFFFu;
    }
  }
}

__global__ __launch_bounds__(256) void k_col_apply(uint32_t *__restrict__ cmat, const ColItem *__restrict__ items,
                                                   const SegDesc *__restrict__ segs, const uint64_t *__restrict__ psum_base,
                                                   const uint32_t *__restrict__ psum, const uint32_t *__restrict__ bstart) {
  const ColItem it = items[blockIdx.x];
  const SegDesc &sd = segs[it.seg];
  const uint32_t b = it.col_lo + threadIdx.x;
  if (b >= sd.ncols) return;
  uint32_t base = bstart[sd.bucket_base + b] + psum[psum_base[it.seg] + (uint64_t)it.chunk * sd.ncols + b];
  uint32_t *p = cmat + sd.row_base + (uint64_t)it.row_lo * sd.ncols + b;
  // rows in groups of 16: loads first, then the stores of the positions
  for (uint32_t r = it.row_lo; r < it.row_hi; r += 16, p += 16 * (uint64_t)sd.ncols) {
    const uint32_t n = min(16u, it.row_hi - r);
    uint32_t v[16];
#pragma unroll
    for (uint32_t u = 0; u < 16; u++) v[u] = u < n ? p[(uint64_t)u * sd.ncols] : 0u;
#pragma unroll
    for (uint32_t u = 0; u < 16; u++)
      if (u < n) {
        p[(uint64_t)u * sd.ncols] = base;
        base += v[u];
      }
  }
}

// ---------------------------------------------------------------------------
// Partition pass: scatter.  Recounts the tile in LDS (rank of every key in
// its bucket), sorts the tile by bucket through an LDS stage and writes each
// (tile, bucket) run contiguously at its count-matrix position.
// Requires tile count <= 65536 and ncols <= 65536.
// ---------------------------------------------------------------------------
template <typename K, int BLOCK, int KPT, int HK, int SK, bool PACK = false>
__global__ __launch_bounds__(BLOCK) void k_part_scatter(const K *__restrict__ keys, const TileDesc *__restrict__ tiles,
                                                        const SegDesc *__restrict__ segs, const uint32_t *__restrict__ cpos,
                                                        K *__restrict__ out, uint32_t stage_cap) {
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ uint32_t s_tmp[BLOCK / 64 + 1];
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const TileDesc td = tiles[tile];
  const SegDesc sd = segs[td.seg];  // a copy: uniform fields stay in SGPRs
  const uint32_t ncols = sd.ncols;
  // LDS: s_cnt (round counts, then round offsets), s_adj (global position of
  // the tile's next key of each bucket, minus the round offset while a round
  // is written), then the staging area
  const uint32_t cnt_words = (ncols + 3) & ~3u;
  uint32_t *s_cnt = (uint32_t *)smem;
  uint32_t *s_adj = s_cnt + cnt_words;
  K *s_key = (K *)(s_adj + cnt_words);
  uint16_t *s_bkt = (uint16_t *)(s_key + stage_cap);
  const uint32_t *row = cpos + sd.row_base + (uint64_t)(tile - sd.tile_begin) * ncols;
  __shared__ int64_t s_spl[SK == SK_RANGE ? 1 : MAX_SPLITS + 1];
  const int64_t *spl = stage_splits<SK, BLOCK>(sd, s_spl);
  lds_fill<uint32_t, BLOCK, 4>(s_adj, row, ncols);
  const uint32_t nb = sd.nbuckets;
  for (uint32_t base = 0; base < td.count; base += BLOCK * KPT) {
    const uint32_t cnt = min((uint32_t)(BLOCK * KPT), td.count - base);
    for (uint32_t i = threadIdx.x; i < ncols; i += BLOCK) s_cnt[i] = 0;
    __syncthreads();
    const uint32_t nv = cnt > threadIdx.x ? (cnt - threadIdx.x + BLOCK - 1) / BLOCK : 0;
    K kr[KPT];
    uint32_t pk[KPT];
    if constexpr (PACK) {
      // 32-bit input keys made into 8-byte records: input position << 32 | key
      const uint32_t *kp = (const uint32_t *)keys + td.key_start + base;
      const __amdgpu_buffer_rsrc_t rs = make_rsrc(kp, cnt * 4u);
      const uint64_t i0 = td.key_start + base + threadIdx.x;
#pragma unroll
      for (int j = 0; j < KPT; j++)
        kr[j] = ((i0 + (uint64_t)j * BLOCK) << 32) | buf_load<uint32_t>(rs, threadIdx.x * 4u, j * BLOCK * 4u);
    } else {
      const K *kp = keys + td.key_start + base;
      const __amdgpu_buffer_rsrc_t rs = make_rsrc(kp, cnt * (uint32_t)sizeof(K));
#pragma unroll
      for (int j = 0; j < KPT; j++)
        kr[j] = buf_load<K>(rs, threadIdx.x * (uint32_t)sizeof(K), j * BLOCK * (uint32_t)sizeof(K));
    }
    if constexpr (sizeof(K) == 4) {
#pragma unroll
      for (int j = 0; j < KPT; j++) pk[j] = j < (int)nv ? bucket_of<K, HK, SK>(kr[j], sd, spl) : 0u;
      // ranks in groups of RG atomics issued before their results are used
      // (one LDS round trip per group instead of one per key)
      constexpr int RG = 4;
#pragma unroll
      for (int g = 0; g < KPT; g += RG) {
        uint32_t r[RG];
#pragma unroll
        for (int u = 0; u < RG; u++) {
          r[u] = 0;
          if (g + u < (int)nv) r[u] = atomicAdd(&s_cnt[pk[g + u]], 1u);
        }
#pragma unroll
        for (int u = 0; u < RG; u++) pk[g + u] = (pk[g + u] << 16) | r[u];
      }
    } else {  // 64-bit keys fill the registers: one key at a time
#pragma unroll
      for (int j = 0; j < KPT; j++) {
        if (j < (int)nv) {
          const uint32_t b = bucket_of<K, HK, SK>(kr[j], sd, spl);
          const uint32_t r = atomicAdd(&s_cnt[b], 1u);
          pk[j] = (b << 16) | r;
        }
      }
    }
    __syncthreads();
    const uint32_t tot = lds_exclusive_scan<BLOCK>(s_cnt, ncols, s_tmp);
#pragma unroll
    for (int j = 0; j < KPT; j++) {
      if (j < nv) {
        const uint32_t b = pk[j] >> 16;
        pk[j] = (b << 16) | (s_cnt[b] + (pk[j] & 0xFFFFu));
      }
    }
    for (uint32_t b = threadIdx.x; b < ncols; b += BLOCK) s_adj[b] -= s_cnt[b];
    __syncthreads();
    for (uint32_t r0 = 0; r0 < cnt; r0 += stage_cap) {
#pragma unroll
      for (int j = 0; j < KPT; j++) {
        if (j < nv) {
          const uint32_t q = (pk[j] & 0xFFFFu) - r0;
          if (q < stage_cap) {
            s_key[q] = kr[j];
            s_bkt[q] = (uint16_t)(pk[j] >> 16);
          }
        }
      }
      __syncthreads();
      const uint32_t n = min(stage_cap, cnt - r0);
      // runs out (the LDS reads of a thread's staged keys are independent)
      for (uint32_t q = threadIdx.x; q < n; q += BLOCK) {
        const uint32_t b = s_bkt[q];
        if (b < nb) out[(uint32_t)(s_adj[b] + r0 + q)] = s_key[q];
      }
      __syncthreads();
    }
    // next round: s_adj[b] = global position after this round's keys of b
    for (uint32_t b = threadIdx.x; b < ncols; b += BLOCK) s_adj[b] += b + 1 < ncols ? s_cnt[b + 1] : tot;
    __syncthreads();
  }
}


// Final-stage store of an index-carrying build of 32-bit keys (SPLIT): the
// 8-byte record (input position << 32 | key) goes to the CSR key array as its
// low half and to the position array as its high half.
__device__ __forceinline__ void split_store(uint32_t *keys, uint64_t *pos, uint64_t i, uint64_t rec) {
  keys[i] = (uint32_t)rec;
  pos[i] = rec >> 32;
}

// ---------------------------------------------------------------------------
// Final stage: one workgroup per bucket (count <= BLOCK*KPT, count <= 65535,
// nbins <= 65535).  Keys stay in VGPRs; the bucket-local bin (16 bit) is
// computed once and kept packed.  LDS holds nbins u16 counters (two per
// word); their exclusive scan gives the offsets, a second atomic pass gives
// every key's slot, then LDS is reused to stage the bin-ordered keys.
// ---------------------------------------------------------------------------
template <typename K, typename OFF, int BLOCK, int KPT, int HK, bool SPLIT = false>
__global__ __launch_bounds__(BLOCK) void k_final(const K *__restrict__ part, K *__restrict__ csr, OFF *__restrict__ off,
                                                 const FinItem *__restrict__ items, const BinMap *__restrict__ maps,
                                                 uint32_t lds_bytes, unsigned int *__restrict__ big_seen,
                                                 uint64_t *__restrict__ csr_index) {
  extern __shared__ __align__(16) uint32_t smem32[];
  __shared__ uint32_t s_tmp[BLOCK / 64 + 1];
  const FinItem it = items[blockIdx.x];
  if (it.count > BLOCK * KPT || it.count > 65535u) {  // the oversized-bucket kernels' item: they run
    if (threadIdx.x == 0) atomicOr(big_seen, 1u);
    return;
  }
  const BinMap mp = maps[it.table];
  const uint32_t W = it.nbins, cnt = it.count;
  const uint32_t nw = (W + 1) >> 1;
  for (uint32_t i = threadIdx.x; i < nw; i += BLOCK) smem32[i] = 0;
  __syncthreads();
  const K *kp = part + it.key_start;
  const uint32_t nv = cnt > threadIdx.x ? (cnt - threadIdx.x + BLOCK - 1) / BLOCK : 0;
  // slots below nfull are valid in every lane: a uniform (scalar) branch, so
  // no per-slot exec mask has to stay live across the unrolled loops
  const uint32_t nfull = cnt / BLOCK;
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(kp, cnt * (uint32_t)sizeof(K));
  K kr[KPT];
  uint32_t lb[(KPT + 1) / 2];
#pragma unroll
  for (int j = 0; j < KPT; j++) kr[j] = buf_load<K>(rs, threadIdx.x * (uint32_t)sizeof(K), j * BLOCK * (uint32_t)sizeof(K));
#pragma unroll
  for (int j = 0; j < (KPT + 1) / 2; j++) lb[j] = 0;
  // a learned model: bins from the leaves this bucket's keys reach, copied
  // to LDS (one or a few leaves on most buckets)
  __shared__ RmiLeaf s_leaf[HK == HK_RMI ? RANGE_LEAVES : 1];
  if (HK == HK_RMI)
    rmi_local_bins<K, BLOCK, KPT>(kr, nv, lb, mp.rmi, s_leaf, RANGE_LEAVES, (uint32_t)it.bin_lo, mp.M);
  auto count_slot = [&](int j) {
    uint32_t l;
    if (HK == HK_RMI) {
      l = (lb[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu;
    } else {
      l = bin_of<K, HK>(kr[j], mp) - (uint32_t)it.bin_lo;
      lb[j >> 1] |= l << ((j & 1) * 16);
    }
    atomicAdd(&smem32[l >> 1], 1u << ((l & 1) * 16));
  };
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    if (__builtin_amdgcn_readfirstlane(j < (int)nfull)) count_slot(j);
    else if (j < nv) count_slot(j);
  }
  __syncthreads();
  // exclusive scan of the u16 counts (bucket total <= 65535, so every
  // prefix fits 16 bits and the packed halves never carry into each other)
  {
    const uint32_t per = (nw + BLOCK - 1) / BLOCK;
    const uint32_t w0 = min(threadIdx.x * per, nw), w1 = min(w0 + per, nw);
    uint32_t s = 0;
    for (uint32_t w = w0; w < w1; w++) {
      const uint32_t v = smem32[w];
      s += (v & 0xFFFFu) + (v >> 16);
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_sum<BLOCK, uint32_t>(s, s_tmp, tot);
    for (uint32_t w = w0; w < w1; w++) {
      const uint32_t v = smem32[w];
      const uint32_t lo = v & 0xFFFFu;
      smem32[w] = ex | ((ex + lo) << 16);
      ex += lo + (v >> 16);
    }
  }
  __syncthreads();
  const uint64_t obase = it.key_start - it.key_rel;
  for (uint32_t l = threadIdx.x; l < W; l += BLOCK) {
    const uint32_t pre = (smem32[l >> 1] >> ((l & 1) * 16)) & 0xFFFFu;
    off[it.off_pos + l] = (OFF)(obase + pre);
  }
  if (it.end_pos != ~0ull && threadIdx.x == 0) off[it.end_pos] = (OFF)(obase + cnt);
  __syncthreads();
  auto place_slot = [&](int j) {
    const uint32_t sh = (j & 1) * 16;
    const uint32_t l = (lb[j >> 1] >> sh) & 0xFFFFu;
    const uint32_t old = atomicAdd(&smem32[l >> 1], 1u << ((l & 1) * 16));
    const uint32_t pos = (old >> ((l & 1) * 16)) & 0xFFFFu;
    lb[j >> 1] = (lb[j >> 1] & ~(0xFFFFu << sh)) | (pos << sh);
  };
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    if (__builtin_amdgcn_readfirstlane(j < (int)nfull)) place_slot(j);
    else if (j < nv) place_slot(j);
  }
  __syncthreads();
  K *st = (K *)smem32;
  const uint32_t cap = lds_bytes / sizeof(K);
  K *dst = csr + it.key_start;
  for (uint32_t r0 = 0; r0 < cnt; r0 += cap) {
    auto stage_slot = [&](int j) {
      const uint32_t q = ((lb[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu) - r0;
      if (q < cap) st[q] = kr[j];
    };
#pragma unroll
    for (int j = 0; j < KPT; j++) {
      if (__builtin_amdgcn_readfirstlane(j < (int)nfull)) stage_slot(j);
      else if (j < nv) stage_slot(j);
    }
    __syncthreads();
    const uint32_t n = min(cap, cnt - r0);
    if constexpr (SPLIT) {
      uint32_t *dk = (uint32_t *)csr + it.key_start;
      uint64_t *di = csr_index + it.key_start;
      for (uint32_t q = threadIdx.x; q < n; q += BLOCK) split_store(dk, di, r0 + q, st[q]);
    } else {
      for (uint32_t q = threadIdx.x; q < n; q += BLOCK) dst[r0 + q] = st[q];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Final stage for buckets too large for k_final (skewed keys).  One
// workgroup streams the bucket: per sub-range of SUB bins, count in LDS,
// scan, write offsets, stream again placing keys with LDS atomics.
// ---------------------------------------------------------------------------
template <typename K, typename OFF, int BLOCK, int SUB, bool SPLIT = false>
__global__ __launch_bounds__(BLOCK) void k_final_big(const K *__restrict__ part, K *__restrict__ csr, OFF *__restrict__ off,
                                                     const FinItem *__restrict__ items, const BinMap *__restrict__ maps,
                                                     uint32_t small_cap, const uint32_t *__restrict__ big_slot,
                                                     const unsigned int *__restrict__ big_seen,
                                                     uint64_t *__restrict__ csr_index) {
  __shared__ uint32_t s_cnt[SUB];
  __shared__ uint32_t s_tmp[BLOCK / 64 + 1];
  if (*big_seen == 0) return;  // no oversized item
  const FinItem it = items[blockIdx.x];
  if (it.count <= small_cap) return;               // k_final's item
  if (big_slot[blockIdx.x] != ~0u) return;         // the multi-workgroup path's item
  const BinMap mp = maps[it.table];
  const uint64_t cnt = it.count;
  const K *kp = part + it.key_start;
  K *dst = csr + it.key_start;
  const uint64_t obase = it.key_start - it.key_rel;
  uint64_t done = 0;
  for (uint32_t s0 = 0; s0 < it.nbins; s0 += SUB) {
    const uint32_t n = min((uint32_t)SUB, it.nbins - s0);
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) s_cnt[i] = 0;
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < cnt; i += BLOCK) {
      const uint32_t l = bin_of<K>(kp[i], mp) - it.bin_lo - s0;
      if (l < n) atomicAdd(&s_cnt[l], 1u);
    }
    __syncthreads();
    const uint32_t tot = lds_exclusive_scan<BLOCK>(s_cnt, n, s_tmp);
    for (uint32_t l = threadIdx.x; l < n; l += BLOCK) off[it.off_pos + s0 + l] = (OFF)(obase + done + s_cnt[l]);
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < cnt; i += BLOCK) {
      const K k = kp[i];
      const uint32_t l = bin_of<K>(k, mp) - it.bin_lo - s0;
      if (l < n) {
        const uint64_t w = done + atomicAdd(&s_cnt[l], 1u);
        if constexpr (SPLIT)
          split_store((uint32_t *)csr + it.key_start, csr_index + it.key_start, w, k);
        else
          dst[w] = k;
      }
    }
    __syncthreads();
    done += tot;
  }
  if (it.end_pos != ~0ull && threadIdx.x == 0) off[it.end_pos] = (OFF)(obase + cnt);
}

// ---------------------------------------------------------------------------
// One-pass probe, for buckets whose keys of A and offsets fit in LDS (all of
// them on uniform keys).  B's keys stay in VGPRs with their 16-bit local
// bins; the bucket's keys of A and its W+1 offsets (relative to the bucket's
// first key, so 16 bits each) are staged in LDS together, with both arrays'
// loads in flight at once; each B key reads its bin's two offsets and walks
// the bin in LDS, counting matches in 16-bit halves.  The first 4 keys of a
// bin are compared without a branch.  Emission as in k_probe.  Heavy bins
// are left to k_probe_heavy.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t probe1_off_bytes(uint32_t W) { return ((W + 1) * 2u + 15u) & ~15u; }
constexpr uint32_t PROBE1_LEAF_BYTES = 1024;  // room for 64 leaves of a learned model

// Whether k_probe1 (LDS budget lds_bytes, kpb keys per workgroup) takes an
// item; k_probe takes the others of its size class.
__device__ __forceinline__ bool probe1_takes(uint32_t count, uint64_t na, uint32_t W, uint32_t key_bytes, uint32_t kpb,
                                             uint32_t lds_bytes) {
  return count <= kpb && count <= 65535u &&
         na * key_bytes + 16 + probe1_off_bytes(W) + PROBE1_LEAF_BYTES <= lds_bytes;
}

// Output region of table t's matches: the whole buffer (out_off null), or
// out[out_off[t] .. out_off[t+1]) when a batch gives every table its own.
template <typename K>
__device__ __forceinline__ K *out_region(K *out, uint64_t capacity, const uint64_t *out_off, uint32_t t,
                                         uint64_t &cap) {
  cap = capacity;
  if (out == nullptr || out_off == nullptr) return out;
  cap = out_off[t + 1] - out_off[t];
  return out + out_off[t];
}

template <typename K, typename OFF, int BLOCK, int KPT, int HK>
__global__ __launch_bounds__(BLOCK) void k_probe1(const K *__restrict__ bpart, const FinItem *__restrict__ items,
                                                  const BinMap *__restrict__ maps, const OFF *__restrict__ offA,
                                                  const K *__restrict__ keysA, K *__restrict__ out, uint64_t capacity,
                                                  unsigned long long *__restrict__ cursor, uint32_t lds_bytes,
                                                  unsigned int *__restrict__ hmark, unsigned int *__restrict__ gate,
                                                  const uint64_t *__restrict__ out_off) {
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ uint32_t s_tmp[BLOCK / 64 + 1];
  __shared__ unsigned long long s_base;
  const FinItem it = items[blockIdx.x];
  const uint32_t W = it.nbins;
  const OFF *ob = offA + it.off_pos;
  const uint64_t a_lo = (uint64_t)ob[0], a_hi = (uint64_t)ob[W];
  if (!probe1_takes(it.count, a_hi - a_lo, W, (uint32_t)sizeof(K), BLOCK * KPT, lds_bytes)) {
    if (threadIdx.x == 0) atomicOr(&gate[0], 1u);  // k_probe's or k_probe_big's item: they run
    return;
  }
  const uint32_t na = (uint32_t)(a_hi - a_lo);
  uint16_t *s_off = (uint16_t *)smem;
  K *s_ak = (K *)(smem + probe1_off_bytes(W));
  const BinMap mp = maps[it.table];
  const uint32_t cnt = it.count;
  const K *kp = bpart + it.key_start;
  const K *ka = keysA + it.key_rel;
  const uint32_t nv = cnt > threadIdx.x ? (cnt - threadIdx.x + BLOCK - 1) / BLOCK : 0;
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(kp, cnt * (uint32_t)sizeof(K));
  K kr[KPT];
  uint32_t lb[(KPT + 1) / 2], mc[(KPT + 1) / 2];
#pragma unroll
  for (int j = 0; j < KPT; j++) kr[j] = buf_load<K>(rs, threadIdx.x * (uint32_t)sizeof(K), j * BLOCK * (uint32_t)sizeof(K));
  lds_fill_keys_offs<K, OFF, BLOCK, sizeof(K) == 4 ? 12 : 4>(s_ak, ka + a_lo, na, s_off, ob, W + 1, (OFF)a_lo);
#pragma unroll
  for (int j = 0; j < (KPT + 1) / 2; j++) {
    lb[j] = 0;
    mc[j] = 0;
  }
  if (HK == HK_RMI) {
    // a learned model: bins from the leaves B's keys reach, copied to the
    // LDS after A's keys (probe1_takes leaves at least PROBE1_LEAF_BYTES)
    const uint32_t used = probe1_off_bytes(W) + ((na * (uint32_t)sizeof(K) + 15u) & ~15u);
    RmiLeaf *s_leaf = (RmiLeaf *)(smem + used);
    const uint32_t cap = min((lds_bytes - used) / (uint32_t)sizeof(RmiLeaf), RANGE_LEAVES);
    rmi_local_bins<K, BLOCK, KPT>(kr, nv, lb, mp.rmi, s_leaf, cap, (uint32_t)it.bin_lo, mp.M);
  } else {
#pragma unroll
    for (int j = 0; j < KPT; j++)
      if (j < nv) lb[j >> 1] |= ((bin_of<K, HK>(kr[j], mp) - it.bin_lo) & 0xFFFFu) << ((j & 1) * 16);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    const uint32_t l = (lb[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu;  // < W: the key's bin is in this bucket
    if (j >= nv || l >= W) continue;
    const uint32_t b0 = s_off[l];
    uint32_t len = s_off[l + 1] - b0;
    if (len > PROBE_BIN_LIMIT) {
      hmark[blockIdx.x] = 1u;  // k_probe_heavy answers this bin
      gate[1] = 1u;
      len = 0;
    }
    uint32_t m = 0;
#pragma unroll
    for (uint32_t q = 0; q < 4; q++) {
      const K v = s_ak[q < len ? b0 + q : 0u];
      m += (q < len && v == kr[j]) ? 1u : 0u;
    }
    for (uint32_t q = 4; q < len; q++) m += s_ak[b0 + q] == kr[j];
    mc[j >> 1] += m << ((j & 1) * 16);  // m <= PROBE_BIN_LIMIT: no carry into the other half
  }
  uint32_t mine = 0;
#pragma unroll
  for (int j = 0; j < (KPT + 1) / 2; j++) mine += (mc[j] & 0xFFFFu) + (mc[j] >> 16);
  uint32_t tot;
  const uint32_t ex = block_exclusive_sum<BLOCK, uint32_t>(mine, s_tmp, tot);
  if (threadIdx.x == 0) s_base = tot ? atomicAdd(cursor + it.table, (unsigned long long)tot) : 0ull;
  __syncthreads();
  if (out == nullptr || !mine) return;
  uint64_t cap;
  K *o = out_region(out, capacity, out_off, it.table, cap);
  unsigned long long w = s_base + ex;
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    const uint32_t m = (mc[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu;
    for (uint32_t r = 0; r < m; r++, w++)
      if (w < cap) o[w] = kr[j];
  }
}

// ---------------------------------------------------------------------------
// Probe: B's keys of one bucket stay in VGPRs with their 16-bit local bins.
// The bucket's bins of A are visited in sub-ranges of at most SUB bins and at
// most akey_cap keys; each sub-range's offsets and keys are staged in LDS
// and every B key in it walks its A bin there, accumulating a 16-bit match
// count.  A match is always the probe key itself, so the emission needs no
// second walk: after the last sub-range one block scan and one atomic
// reservation per workgroup place each key's copies (output grouped by
// bucket).  A bin longer than PROBE_BIN_LIMIT is not walked: the item is
// marked in hmark and k_probe_heavy answers the keys of such bins.
// ---------------------------------------------------------------------------
template <typename K, typename OFF, int BLOCK, int KPT, int SUB, int HK>
__global__ __launch_bounds__(BLOCK) void k_probe(const K *__restrict__ bpart, const FinItem *__restrict__ items,
                                                 const BinMap *__restrict__ maps, const OFF *__restrict__ offA,
                                                 const K *__restrict__ keysA, K *__restrict__ out, uint64_t capacity,
                                                 unsigned long long *__restrict__ cursor, uint32_t akey_cap,
                                                 unsigned int *__restrict__ hmark, uint32_t one_cnt,
                                                 uint32_t one_lds, unsigned int *__restrict__ gate,
                                                 const uint64_t *__restrict__ out_off) {
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ uint32_t s_tmp[BLOCK / 64 + 1];
  __shared__ uint32_t s_n;
  __shared__ unsigned long long s_base;
  OFF *s_off = (OFF *)smem;
  K *s_ak = (K *)(smem + ((SUB + 2) * sizeof(OFF) + 15) / 16 * 16);
  if (gate[0] == 0) return;  // k_probe1 took every item
  const FinItem it = items[blockIdx.x];
  if (it.count > BLOCK * KPT || it.count > 65535u) return;  // k_probe_big's item
  if (probe1_takes(it.count, (uint64_t)(offA[it.off_pos + it.nbins] - offA[it.off_pos]), it.nbins,
                   (uint32_t)sizeof(K), one_cnt, one_lds))
    return;  // k_probe1's item
  const BinMap mp = maps[it.table];
  const uint32_t cnt = it.count, W = it.nbins;
  const K *kp = bpart + it.key_start;
  const K *ka = keysA + it.key_rel;  // A's keys of this table (batched tables)
  const uint32_t nv = cnt > threadIdx.x ? (cnt - threadIdx.x + BLOCK - 1) / BLOCK : 0;
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(kp, cnt * (uint32_t)sizeof(K));
  K kr[KPT];
  uint32_t lb[(KPT + 1) / 2], mc[(KPT + 1) / 2];
#pragma unroll
  for (int j = 0; j < KPT; j++) kr[j] = buf_load<K>(rs, threadIdx.x * (uint32_t)sizeof(K), j * BLOCK * (uint32_t)sizeof(K));
#pragma unroll
  for (int j = 0; j < (KPT + 1) / 2; j++) {
    lb[j] = 0;
    mc[j] = 0;
  }
#pragma unroll
  for (int j = 0; j < KPT; j++)
    if (j < nv) lb[j >> 1] |= ((bin_of<K, HK>(kr[j], mp) - it.bin_lo) & 0xFFFFu) << ((j & 1) * 16);
  for (uint32_t s0 = 0; s0 < W;) {
    uint32_t n = min((uint32_t)SUB, W - s0);
    lds_fill<OFF, BLOCK, (sizeof(K) == 4 && sizeof(OFF) == 4) ? 8 : 4>(s_off, offA + it.off_pos + s0, n + 1);
    __syncthreads();
    const OFF a0 = s_off[0];
    if ((uint64_t)(s_off[n] - a0) > akey_cap) {
      // shrink to the longest prefix of bins whose keys fit (at least one bin)
      if (threadIdx.x == 0) {
        uint32_t lo = 1, hi = n;  // s_off[lo] fits or lo == 1
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if ((uint64_t)(s_off[mid] - a0) <= akey_cap) lo = mid;
          else hi = mid - 1;
        }
        s_n = lo;
      }
      __syncthreads();
      n = s_n;
    }
    const uint32_t na = (uint32_t)min((uint64_t)(s_off[n] - a0), (uint64_t)akey_cap);
    lds_fill<K, BLOCK, sizeof(K) == 4 ? 8 : 4>(s_ak, ka + a0, na);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < KPT; j++) {
      const uint32_t sh = (j & 1) * 16;
      const uint32_t l = ((lb[j >> 1] >> sh) & 0xFFFFu) - s0;
      if (j < nv && l < n) {
        const uint32_t b0 = (uint32_t)(s_off[l] - a0), e = (uint32_t)(s_off[l + 1] - a0);
        uint32_t m = 0;
        if (e - b0 > PROBE_BIN_LIMIT) {  // k_probe_heavy answers this bin
          hmark[blockIdx.x] = 1u;
          gate[1] = 1u;
        }
        else
          for (uint32_t q = b0; q < e; q++) m += s_ak[q] == kr[j];
        mc[j >> 1] += m << sh;  // m <= PROBE_BIN_LIMIT: no carry into the other half
      }
    }
    s0 += n;
    __syncthreads();
  }
  uint32_t mine = 0;
#pragma unroll
  for (int j = 0; j < (KPT + 1) / 2; j++) mine += (mc[j] & 0xFFFFu) + (mc[j] >> 16);
  uint32_t tot;
  const uint32_t ex = block_exclusive_sum<BLOCK, uint32_t>(mine, s_tmp, tot);
  if (threadIdx.x == 0) s_base = tot ? atomicAdd(cursor + it.table, (unsigned long long)tot) : 0ull;
  __syncthreads();
  if (out == nullptr || !mine) return;
  uint64_t cap;
  K *o = out_region(out, capacity, out_off, it.table, cap);
  unsigned long long w = s_base + ex;
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    const uint32_t m = (mc[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu;
    for (uint32_t r = 0; r < m; r++, w++)
      if (w < cap) o[w] = kr[j];
  }
}

// ---------------------------------------------------------------------------
// Join of two built tables over a range of bins (MultiHashGraph::intersect,
// kernel simpleIntersect include/MultiHashGraphDeviceOperators.cuh:513-543):
// every (a in A_b, x in B_b) with a == x counts and emits x (MODE: count,
// keys, or the pair of the entries' input positions for the row-id join).
// One workgroup per JB bins; A's and B's keys of the range are staged in LDS
// when they fit.  A bin pair above JOIN_PAIR_LIMIT key pairs is listed for
// k_join_heavy (keys) or the tiled pair compare (row-id join).
// ---------------------------------------------------------------------------
template <typename K, typename OFF, int BLOCK, int JB, int MODE>
__global__ __launch_bounds__(BLOCK) void k_join(const OFF *__restrict__ offA, const K *__restrict__ keysA,
                                                const OFF *__restrict__ offB, const K *__restrict__ keysB, uint64_t M,
                                                K *__restrict__ out, uint64_t capacity,
                                                unsigned long long *__restrict__ cursor, uint32_t key_cap,
                                                uint32_t *__restrict__ hlist, unsigned int *__restrict__ hlist_n,
                                                uint32_t hlist_cap, const uint64_t *__restrict__ indA,
                                                const uint64_t *__restrict__ indB, uint64_t *__restrict__ pa,
                                                uint64_t *__restrict__ pb) {
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ OFF s_oa[JB + 1], s_ob[JB + 1];
  __shared__ unsigned long long s_base;
  // the row-id join stages the entries' input positions beside their keys,
  // and an emitting join buffers its matches (s_m), so that the output is
  // written from LDS in one contiguous run per workgroup
  constexpr bool emitting = MODE != JOIN_COUNT, pairs = MODE == JOIN_PAIRS;
  constexpr uint32_t EBUF = pairs ? JOIN_EMIT_PAIRS : emitting ? JOIN_EMIT_KEYS : 0u;
  const uint32_t icap = pairs ? key_cap : 0u;
  uint64_t *s_ia = (uint64_t *)smem, *s_ib = s_ia + icap;
  uint32_t *s_m = (uint32_t *)(s_ib + icap);  // pairs: (A slot << 16) | B slot
  uint16_t *s_m16 = (uint16_t *)s_m;          // keys: B slot
  K *s_a = (K *)((uint8_t *)s_m + EBUF * (pairs ? 4u : 2u));
  K *s_b = s_a + key_cap;
  __shared__ uint32_t s_mn;
  const uint64_t b0 = (uint64_t)blockIdx.x * JB;
  const uint32_t n = (uint32_t)min((uint64_t)JB, M - b0);
  if (threadIdx.x == 0) s_mn = 0;
  // staging: every thread issues all its loads of a round before its LDS
  // stores, so a workgroup keeps U * BLOCK loads per array in flight
  constexpr uint32_t U = (JB + BLOCK) / BLOCK;
  {
    OFF va[U], vb[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t q = u * BLOCK + threadIdx.x;
      if (q <= n) {
        va[u] = offA[b0 + q];
        vb[u] = offB[b0 + q];
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t q = u * BLOCK + threadIdx.x;
      if (q <= n) {
        s_oa[q] = va[u];
        s_ob[q] = vb[u];
      }
    }
  }
  __syncthreads();
  const OFF a0 = s_oa[0], x0 = s_ob[0];
  const uint64_t na = (uint64_t)(s_oa[n] - a0), nbk = (uint64_t)(s_ob[n] - x0);
  const bool a_lds = na <= key_cap, b_lds = nbk <= key_cap;
  const uint32_t sa = a_lds ? (uint32_t)na : 0u, sb = b_lds ? (uint32_t)nbk : 0u;
  for (uint32_t r0 = 0; r0 < max(sa, sb); r0 += U * BLOCK) {
    K ka[U], kx[U];
    uint64_t ia_[U], ix_[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t q = r0 + u * BLOCK + threadIdx.x;
      if (q < sa) {
        ka[u] = keysA[a0 + q];
        if (pairs) ia_[u] = indA[a0 + q];
      }
      if (q < sb) {
        kx[u] = keysB[x0 + q];
        if (pairs) ix_[u] = indB[x0 + q];
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t q = r0 + u * BLOCK + threadIdx.x;
      if (q < sa) {
        s_a[q] = ka[u];
        if (pairs) s_ia[q] = ia_[u];
      }
      if (q < sb) {
        s_b[q] = kx[u];
        if (pairs) s_ib[q] = ix_[u];
      }
    }
  }
  __syncthreads();
  // an emitting join with both sides in LDS: every match appends its two LDS
  // slots (below key_cap < 2^16) to s_m during the counting walk
  const bool buffered = emitting && a_lds && b_lds;
  uint64_t mine = 0;
  for (uint32_t l = threadIdx.x; l < n; l += BLOCK) {
    const uint64_t ea = (uint64_t)(s_oa[l + 1] - a0), eb = (uint64_t)(s_ob[l + 1] - x0);
    if ((ea - (uint64_t)(s_oa[l] - a0)) * (eb - (uint64_t)(s_ob[l] - x0)) > JOIN_PAIR_LIMIT) {
      const uint32_t i = atomicAdd(hlist_n, 1u);  // k_join_heavy answers this bin
      if (i < hlist_cap) hlist[i] = (uint32_t)(b0 + l);
      continue;
    }
    if (buffered) {
      for (uint32_t ia = (uint32_t)(s_oa[l] - a0); ia < (uint32_t)ea; ia++) {
        const K a = s_a[ia];
        for (uint32_t ib = (uint32_t)(s_ob[l] - x0); ib < (uint32_t)eb; ib++)
          if (s_b[ib] == a) {
            mine++;
            const uint32_t slot = atomicAdd(&s_mn, 1u);
            if (slot < EBUF) {
              if (pairs)
                s_m[slot] = (ia << 16) | ib;
              else
                s_m16[slot] = (uint16_t)ib;
            }
          }
      }
      continue;
    }
    for (uint64_t ia = (uint64_t)(s_oa[l] - a0); ia < ea; ia++) {
      const K a = a_lds ? s_a[ia] : keysA[a0 + ia];
      for (uint64_t ib = (uint64_t)(s_ob[l] - x0); ib < eb; ib++) mine += (b_lds ? s_b[ib] : keysB[x0 + ib]) == a;
    }
  }
  if (buffered) {
    __syncthreads();
    const uint32_t nm = s_mn;
    if (nm <= EBUF) {  // one reservation, then contiguous stores
      if (threadIdx.x == 0) s_base = nm ? atomicAdd(cursor, (unsigned long long)nm) : 0ull;
      __syncthreads();
      for (uint32_t q = threadIdx.x; q < nm; q += BLOCK) {
        const unsigned long long w = s_base + q;
        if (w < capacity) {
          if (pairs) {
            const uint32_t v = s_m[q];
            pa[w] = s_ia[v >> 16];
            pb[w] = s_ib[v & 0xFFFFu];
          } else {
            out[w] = s_b[s_m16[q]];
          }
        }
      }
      return;
    }
  }
  if (!emitting) {
    __shared__ unsigned long long s_t64[BLOCK / 64 + 1];
    unsigned long long t = block_sum<BLOCK, unsigned long long>(mine, s_t64);
    if (threadIdx.x == 0 && t) atomicAdd(cursor, t);
    return;
  }
  __shared__ unsigned long long s_t64[BLOCK / 64 + 1];
  unsigned long long tot;
  unsigned long long ex = block_exclusive_sum<BLOCK, unsigned long long>(mine, s_t64, tot);
  if (threadIdx.x == 0) s_base = tot ? atomicAdd(cursor, tot) : 0ull;
  __syncthreads();
  unsigned long long w = s_base + ex;
  if (!mine) return;
  for (uint32_t l = threadIdx.x; l < n; l += BLOCK) {
    const uint64_t ea = (uint64_t)(s_oa[l + 1] - a0), eb = (uint64_t)(s_ob[l + 1] - x0);
    if ((ea - (uint64_t)(s_oa[l] - a0)) * (eb - (uint64_t)(s_ob[l] - x0)) > JOIN_PAIR_LIMIT) continue;
    for (uint64_t ia = (uint64_t)(s_oa[l] - a0); ia < ea; ia++) {
      const K a = a_lds ? s_a[ia] : keysA[a0 + ia];
      for (uint64_t ib = (uint64_t)(s_ob[l] - x0); ib < eb; ib++) {
        const K x = b_lds ? s_b[ib] : keysB[x0 + ib];
        if (x == a) {
          if (w < capacity) {
            if (pairs) {  // row-id join: the two entries' input positions
              pa[w] = a_lds ? s_ia[ia] : indA[a0 + ia];
              pb[w] = b_lds ? s_ib[ib] : indB[x0 + ib];
            } else {
              out[w] = x;
            }
          }
          w++;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Single-pass exclusive scan with decoupled look-back (Merrill & Garland
// 2016).  Tiles take ids from an atomic ticket so a tile only waits on tiles
// that started before it.  Per tile: a status flag (0 none, 1 aggregate,
// 2 inclusive prefix) and two 64-bit words, the aggregate and the inclusive
// prefix, each written once before the flag that publishes it (release) and
// read after the flag (acquire), so sums use all 64 bits.
// ---------------------------------------------------------------------------
template <typename T, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void k_scan_lookback(const T *__restrict__ in, T *__restrict__ out, uint64_t n,
                                                         unsigned int *__restrict__ flags,
                                                         unsigned long long *__restrict__ agg_w,
                                                         unsigned long long *__restrict__ inc_w,
                                                         unsigned int *__restrict__ ticket) {
  __shared__ unsigned int s_tile;
  __shared__ unsigned long long s_excl;
  __shared__ T s_tmp[BLOCK / 64 + 1];
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint64_t tile = s_tile;
  const uint64_t base = tile * (uint64_t)(BLOCK * IPT) + (uint64_t)threadIdx.x * IPT;
  T v[IPT];
  T sum = 0;
#pragma unroll
  for (int j = 0; j < IPT; j++) {
    v[j] = base + j < n ? in[base + j] : T(0);
    sum += v[j];
  }
  T agg;
  T ex = block_exclusive_sum<BLOCK, T>(sum, s_tmp, agg);
  if (threadIdx.x < 64) {
    unsigned long long prefix = 0;
    if (tile == 0) {
      if (threadIdx.x == 0) {
        __hip_atomic_store(&inc_w[0], (unsigned long long)agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&flags[0], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else {
      if (threadIdx.x == 0) {
        __hip_atomic_store(&agg_w[tile], (unsigned long long)agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&flags[tile], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
      // wave-wide look-back: lane i inspects tile (p - i)
      int64_t p = (int64_t)tile - 1;
      const int lane = threadIdx.x;
      while (true) {
        const int64_t q = p - lane;
        const unsigned int flag =
            q >= 0 ? __hip_atomic_load(&flags[q], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) : 2u;
        // lanes up to and including the first inclusive prefix must all be ready
        const unsigned long long inc_mask = __ballot(flag == 2);
        const unsigned long long zero_mask = __ballot(flag == 0);
        const int first_inc = inc_mask ? __ffsll((long long)inc_mask) - 1 : 64;
        const unsigned long long need = first_inc >= 63 ? ~0ull : ((2ull << first_inc) - 1);
        if (zero_mask & need) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        unsigned long long contrib = 0;
        if (lane <= first_inc && q >= 0)
          contrib = flag == 2 ? __hip_atomic_load(&inc_w[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : __hip_atomic_load(&agg_w[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        contrib = wave_sum<unsigned long long>(contrib);
        prefix += contrib;
        if (first_inc < 64) break;
        p -= 64;
      }
      if (threadIdx.x == 0) {
        __hip_atomic_store(&inc_w[tile], prefix + (unsigned long long)agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&flags[tile], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        s_excl = prefix;
      }
    }
    if (tile == 0 && threadIdx.x == 0) s_excl = 0;
  }
  __syncthreads();
  T run = (T)s_excl + ex;
#pragma unroll
  for (int j = 0; j < IPT; j++) {
    if (base + j < n) {
      T x = v[j];
      out[base + j] = run;
      run += x;
    }
  }
  const uint64_t last = (uint64_t)(n - 1) / (BLOCK * IPT);
  if (tile == last && threadIdx.x == BLOCK - 1) out[n] = (T)s_excl + agg;
}

template <typename T>
__global__ void k_fill(T *__restrict__ p, T v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// p[i] += d (mod 2^bits of T): rebases a slab's offsets onto table positions
template <typename T>
__global__ void k_add(T *__restrict__ p, T d, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] += d;
}

// SplitMix64 counter form x_i = mix64(s0 + (i+1)*gamma); key = x_i reduced
// to [0, range) by the high half of x_i * range (range 0: no reduction;
// 32-bit keys take the top 32 bits).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
template <typename K>
__global__ void k_gen_uniform(K *__restrict__ out, uint64_t n, uint64_t s0, uint64_t range) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = mix64(s0 + (i + 1) * 0x9E3779B97F4A7C15ull);
    uint64_t k = range ? __umul64hi(x, range) : x;
    if (sizeof(K) == 4 && !range) k >>= 32;
    out[i] = (K)k;
  }
}

template <typename K>
__global__ void k_bins(const K *__restrict__ keys, uint64_t n, BinMap mp, uint32_t *__restrict__ bins) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    bins[i] = bin_of<K>(keys[i], mp);
}

// Histogram of bin / bin_range over bin_count ranges (countBinSizes,
// include/MultiHashGraphDeviceOperators.cuh:153-164), LDS-privatised.
template <typename K>
__global__ __launch_bounds__(1024) void k_range_hist(const K *__restrict__ keys, uint64_t n, BinMap mp, uint32_t range,
                                                     uint32_t nranges, unsigned long long *__restrict__ hist,
                                                     unsigned int *__restrict__ bad) {
  extern __shared__ uint32_t s_h[];
  for (uint32_t i = threadIdx.x; i < nranges; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t c = bin_of<K>(keys[i], mp) / range;
    if (c < nranges) atomicAdd(&s_h[c], 1u);
    else atomicOr(bad, 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nranges; i += blockDim.x)
    if (s_h[i]) atomicAdd(&hist[i], (unsigned long long)s_h[i]);
}

// ---------------------------------------------------------------------------
// Run-length join of two built tables: one record (key, cA, cB) per value
// present in both tables.  Small bins (|A_b|, |B_b| <= RLE_SMALL) are done by
// one thread each (quadratic in the bin, in bin order); a bin above that is
// appended to a heavy list and handled by k_rle_heavy.
// A workgroup takes a chunk of RLE_CHUNK bins: both sides' offsets, and their
// keys when the chunk holds at most RLE_KCAP keys per side (no heavy bin),
// are read coalesced into LDS first; thread t then walks bins t*PER.. there.
// Pass 0 counts records per chunk, pass 1 writes them at the chunk's
// exclusive prefix.
// ---------------------------------------------------------------------------
constexpr uint32_t RLE_SMALL = 64;
constexpr uint32_t RLE_CHUNK = 1024;
constexpr uint32_t RLE_KCAP = 768;
constexpr uint32_t RLE_TINY = 8;  // bins above this on a side (up to RLE_SMALL) are walked by a wave

// A medium bin (A's keys a0.., a of them; B's x0.., b of them; a, b <= 64)
// walked by one wave: lane i takes A's i-th key and counts its copies on both
// sides by broadcast reads; it holds a record if the key occurs first at i
// and B has it.
template <typename K, typename GA, typename GB>
__device__ __forceinline__ bool rle_medium(uint64_t a0, uint32_t a, uint64_t x0, uint32_t b, GA getA, GB getB,
                                           uint32_t lane, K &v, uint32_t &ca, uint32_t &cb) {
  const bool mine = lane < a;
  v = getA(a0 + (mine ? lane : 0u));
  ca = 0;
  cb = 0;
  bool first = true;
  for (uint32_t j = 0; j < a; j++) {
    const bool eq = getA(a0 + j) == v;
    ca += eq ? 1u : 0u;
    first &= !(eq && j < lane);
  }
  for (uint32_t j = 0; j < b; j++) cb += getB(x0 + j) == v ? 1u : 0u;
  return mine && first && cb > 0;
}

// A thread's walk over its PER bins of the chunk.  WRITE == false counts the
// tiny bins' records and pairs, the heavy bins, and marks the medium bins in
// medmask; WRITE == true lists the heavy bins from hw on and writes the tiny
// bins' records from w on.  Without emit only the heavy bins are looked at.
template <typename K, int PER, bool WRITE, typename OFF, typename GA, typename GB>
__device__ __forceinline__ void rle_walk(const OFF *s_oa, const OFF *s_ob, uint32_t lo, uint32_t hi, bool emit,
                                         GA getA, GB getB,
                                         uint64_t b0, uint32_t &nrec, unsigned long long &npair, uint32_t &nheavy,
                                         uint32_t &medmask, unsigned long long w, uint64_t hw, uint32_t *heavy,
                                         uint64_t heavy_cap, K *rk, unsigned long long *rca, unsigned long long *rcb,
                                         uint64_t capacity) {
  for (int q = 0; q < PER; q++) {
    const uint32_t l = threadIdx.x * PER + q;
    if (l >= hi) break;
    if (l < lo) continue;
    const uint64_t a0 = s_oa[l], a1 = s_oa[l + 1], x0 = s_ob[l], x1 = s_ob[l + 1];
    if (a1 == a0 || x1 == x0) continue;
    if (a1 - a0 > RLE_SMALL || x1 - x0 > RLE_SMALL) {
      if (WRITE) {
        if (hw < heavy_cap) heavy[hw] = (uint32_t)(b0 + l);
        hw++;
      } else {
        nheavy++;
      }
      continue;
    }
    if (!emit) continue;
    if (a1 - a0 > RLE_TINY || x1 - x0 > RLE_TINY) {  // a wave takes it (rle_medium)
      if (!WRITE) medmask |= 1u << q;
      continue;
    }
    for (uint64_t i = a0; i < a1; i++) {
      const K v = getA(i);
      bool first = true;
      for (uint64_t j = a0; j < i; j++) first &= getA(j) != v;
      if (!first) continue;
      unsigned long long ca = 0, cb = 0;
      for (uint64_t j = i; j < a1; j++) ca += getA(j) == v;
      for (uint64_t j = x0; j < x1; j++) cb += getB(j) == v;
      if (!cb) continue;
      if (!WRITE) {
        nrec++;
        npair += ca * cb;
      } else {
        if (w < capacity) {
          rk[w] = v;
          rca[w] = ca;
          rcb[w] = cb;
        }
        w++;
      }
    }
  }
}

// Pass 0 stores per chunk its record count, heavy-bin count and pair count
// (plain stores: one atomic per chunk on a shared counter would serialise the
// whole grid); pass 1, after the host scanned the first two, lists the heavy
// bins at the chunk's heavy prefix and, when emit, writes the records at the
// chunk's record prefix.  Pass 1 without emit only lists the heavy bins.
template <typename K, typename OFF, int BLOCK>
__global__ __launch_bounds__(BLOCK, 6) void k_rle_small(const OFF *__restrict__ offA, const K *__restrict__ keysA,
                                                     const OFF *__restrict__ offB, const K *__restrict__ keysB, uint64_t M,
                                                     int pass, unsigned long long *__restrict__ chunk_cnt,
                                                     const unsigned long long *__restrict__ chunk_base,
                                                     uint32_t *__restrict__ chunk_heavy,
                                                     const uint32_t *__restrict__ heavy_base,
                                                     unsigned long long *__restrict__ chunk_pairs, K *__restrict__ rk,
                                                     unsigned long long *__restrict__ rca,
                                                     unsigned long long *__restrict__ rcb, uint64_t capacity,
                                                     uint32_t *__restrict__ heavy, uint64_t heavy_cap) {
  constexpr int PER = RLE_CHUNK / BLOCK;
  constexpr int KPT = RLE_KCAP / BLOCK;
  __shared__ unsigned long long s_t64[BLOCK / 64 + 1];
  __shared__ OFF s_oa[RLE_CHUNK + 1], s_ob[RLE_CHUNK + 1];
  __shared__ K s_ka[RLE_KCAP], s_kb[RLE_KCAP];
  const uint64_t b0 = (uint64_t)blockIdx.x * RLE_CHUNK;
  const uint32_t nb = (uint32_t)min((uint64_t)RLE_CHUNK, M - b0);
  const bool emit = pass != 1 || capacity != 0;
  {
    // all offset loads of the thread in flight together (one round trip)
    OFF va[PER + 1], vb[PER + 1];
#pragma unroll
    for (int j = 0; j <= PER; j++) {
      const uint32_t i = threadIdx.x + j * BLOCK;
      if (i <= nb) {
        va[j] = offA[b0 + i];
        vb[j] = offB[b0 + i];
      }
    }
#pragma unroll
    for (int j = 0; j <= PER; j++) {
      const uint32_t i = threadIdx.x + j * BLOCK;
      if (i <= nb) {
        s_oa[i] = va[j];
        s_ob[i] = vb[j];
      }
    }
  }
  __syncthreads();
  // Sub-ranges: a chunk whose keys exceed RLE_KCAP on a side is walked in
  // 2 or 4 sub-ranges of bins, each staged on its own (dense tables); when
  // none fits, in one range reading the keys from global memory.
  uint32_t nsub = 1;
  bool staged = false;
  if (emit) {
    for (uint32_t ns = 1; ns <= 4 && !staged; ns <<= 1) {
      const uint32_t span = RLE_CHUNK / ns;
      bool fits = true;
      for (uint32_t r = 0; r < ns && r * span < nb; r++) {
        const uint32_t lo = r * span, hi = min(lo + span, nb);
        fits &= (uint64_t)s_oa[hi] - (uint64_t)s_oa[lo] <= RLE_KCAP && (uint64_t)s_ob[hi] - (uint64_t)s_ob[lo] <= RLE_KCAP;
      }
      if (fits) {
        nsub = ns;
        staged = true;
      }
    }
  }
  // list of the medium bins (thread order: deterministic)
  __shared__ uint16_t s_med[RLE_CHUNK];
  __shared__ uint16_t s_mcnt[RLE_CHUNK];  // records of a medium bin, then its base in the sub-range (< 2^16)
  __shared__ unsigned long long s_mbase;
  __shared__ uint32_t s_mtot;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr uint32_t NW = BLOCK / 64;
  // pass 0: the chunk's totals (records | heavy << 32, pairs); passes 1
  // and 3: the records and heavy bins written so far (pass 3 also sums the
  // pairs)
  unsigned long long acc = 0, accp = 0;
  const unsigned long long cbase = pass == 0 ? 0ull : chunk_base[blockIdx.x];
  const uint64_t hbase = pass == 0 ? 0ull : (uint64_t)heavy_base[blockIdx.x];
  const uint32_t span = RLE_CHUNK / nsub;
  for (uint32_t r = 0; r < nsub; r++) {
    const uint32_t lo = r * span, hi = min(lo + span, nb);
    if (lo >= nb) break;  // uniform
    const uint64_t ka0 = uni64(s_oa[lo]), kb0 = uni64(s_ob[lo]);
    const uint64_t na = uni64((uint64_t)s_oa[hi] - ka0), nx = uni64((uint64_t)s_ob[hi] - kb0);
    if (staged) {
      if (r) __syncthreads();  // the previous sub-range is done with s_ka / s_kb
      K ta[KPT], tb[KPT];
#pragma unroll
      for (int j = 0; j < KPT; j++) {
        const uint32_t i = threadIdx.x + j * BLOCK;
        if (i < na) ta[j] = keysA[ka0 + i];
        if (i < nx) tb[j] = keysB[kb0 + i];
      }
#pragma unroll
      for (int j = 0; j < KPT; j++) {
        const uint32_t i = threadIdx.x + j * BLOCK;
        if (i < na) s_ka[i] = ta[j];
        if (i < nx) s_kb[i] = tb[j];
      }
      __syncthreads();
    }
    auto lA = [&](uint64_t i) { return s_ka[i - ka0]; };
    auto lB = [&](uint64_t i) { return s_kb[i - kb0]; };
    auto gA = [&](uint64_t i) { return keysA[i]; };
    auto gB = [&](uint64_t i) { return keysB[i]; };
    uint32_t nrec = 0, nheavy = 0, medmask = 0;
    unsigned long long npair = 0;
    if (staged)
      rle_walk<K, PER, false>(s_oa, s_ob, lo, hi, emit, lA, lB, b0, nrec, npair, nheavy, medmask, 0ull, 0ull, heavy, 0,
                              rk, rca, rcb, 0);
    else
      rle_walk<K, PER, false>(s_oa, s_ob, lo, hi, emit, gA, gB, b0, nrec, npair, nheavy, medmask, 0ull, 0ull, heavy, 0,
                              rk, rca, rcb, 0);
    unsigned long long nmed_tot;
    const unsigned long long mex = block_exclusive_sum<BLOCK, unsigned long long>(
        (unsigned long long)__popc(medmask), s_t64, nmed_tot);
    {
      uint32_t k = (uint32_t)mex;
      for (int q = 0; q < PER; q++)
        if (medmask & (1u << q)) s_med[k++] = (uint16_t)(threadIdx.x * PER + q);
    }
    __syncthreads();
    const uint32_t nmed = uni32((uint32_t)nmed_tot);
    // medium bins: one wave each
    uint32_t mrec = 0;
    unsigned long long mpair = 0;
    if (pass == 0)
      for (uint32_t k = wv; k < nmed; k += NW) {
        const uint32_t l = s_med[k];
        const uint64_t a0 = s_oa[l], x0 = s_ob[l];
        const uint32_t a = (uint32_t)((uint64_t)s_oa[l + 1] - a0), b = (uint32_t)((uint64_t)s_ob[l + 1] - x0);
        K v;
        uint32_t ca, cb;
        const bool rec = staged ? rle_medium<K>(a0, a, x0, b, lA, lB, lane, v, ca, cb)
                                : rle_medium<K>(a0, a, x0, b, gA, gB, lane, v, ca, cb);
        mrec += rec ? 1u : 0u;
        mpair += rec ? (unsigned long long)ca * cb : 0ull;
      }
    // records in the low word, heavy bins in the high word (a chunk holds at
    // most 64 K records and 1024 heavy bins)
    unsigned long long tot;
    const unsigned long long ex = block_exclusive_sum<BLOCK, unsigned long long>(
        (unsigned long long)(nrec + mrec) | ((unsigned long long)nheavy << 32), s_t64, tot);
    if (pass == 0) {
      acc += uni64(tot);
      accp += uni64(block_sum<BLOCK, unsigned long long>(npair + mpair, s_t64));
      continue;
    }
    if (pass == 3) accp += uni64(block_sum<BLOCK, unsigned long long>(npair, s_t64));  // the tiny bins' pairs
    // pass 1: the tiny bins' records at the sub-range's base plus the threads
    // before, then the medium bins' records after all of them
    const unsigned long long rbase = cbase + (acc & 0xFFFFFFFFull);
    const uint64_t hw = hbase + (acc >> 32) + (ex >> 32);
    if (staged)
      rle_walk<K, PER, true>(s_oa, s_ob, lo, hi, emit, lA, lB, b0, nrec, npair, nheavy, medmask,
                             rbase + (ex & 0xFFFFFFFFull), hw, heavy, heavy_cap, rk, rca, rcb, capacity);
    else
      rle_walk<K, PER, true>(s_oa, s_ob, lo, hi, emit, gA, gB, b0, nrec, npair, nheavy, medmask,
                             rbase + (ex & 0xFFFFFFFFull), hw, heavy, heavy_cap, rk, rca, rcb, capacity);
    acc += uni64(tot);  // tiny records and heavy bins of the sub-range
    if (nmed == 0) continue;  // uniform
    // medium bins: count each bin's records, then the bins' bases in bin order
    for (uint32_t k = wv; k < nmed; k += NW) {
      const uint32_t l = s_med[k];
      const uint64_t a0 = s_oa[l], x0 = s_ob[l];
      const uint32_t a = (uint32_t)((uint64_t)s_oa[l + 1] - a0), b = (uint32_t)((uint64_t)s_ob[l + 1] - x0);
      K v;
      uint32_t ca, cb;
      const bool rec = staged ? rle_medium<K>(a0, a, x0, b, lA, lB, lane, v, ca, cb)
                              : rle_medium<K>(a0, a, x0, b, gA, gB, lane, v, ca, cb);
      const uint64_t bal = __ballot(rec);
      if (lane == 0) s_mcnt[k] = (uint16_t)__popcll(bal);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t m = 0;
      for (uint32_t k = 0; k < nmed; k++) {  // record base of each medium bin past the tiny bins' records
        const uint32_t c = s_mcnt[k];
        s_mcnt[k] = (uint16_t)m;
        m += c;
      }
      s_mbase = cbase + (acc & 0xFFFFFFFFull);
      s_mtot = m;
    }
    __syncthreads();
    const unsigned long long mb = uni64(s_mbase);
    acc += uni32(s_mtot);  // the medium bins' records
    unsigned long long mp = 0;  // pass 3: the medium bins' pairs
    for (uint32_t k = wv; k < nmed; k += NW) {
      const uint32_t l = s_med[k];
      const uint64_t a0 = s_oa[l], x0 = s_ob[l];
      const uint32_t a = (uint32_t)((uint64_t)s_oa[l + 1] - a0), b = (uint32_t)((uint64_t)s_ob[l + 1] - x0);
      K v;
      uint32_t ca, cb;
      const bool rec = staged ? rle_medium<K>(a0, a, x0, b, lA, lB, lane, v, ca, cb)
                              : rle_medium<K>(a0, a, x0, b, gA, gB, lane, v, ca, cb);
      const uint64_t bal = __ballot(rec);
      if (rec) {
        const unsigned long long p = mb + s_mcnt[k] + (unsigned long long)__popcll(bal & ((1ull << lane) - 1ull));
        if (p < capacity) {
          rk[p] = v;
          rca[p] = ca;
          rcb[p] = cb;
        }
        mp += (unsigned long long)ca * cb;
      }
    }
    if (pass == 3) accp += uni64(block_sum<BLOCK, unsigned long long>(mp, s_t64));
  }
  if ((pass == 0 || pass == 3) && threadIdx.x == 0) {
    chunk_cnt[blockIdx.x] = acc & 0xFFFFFFFFull;
    chunk_heavy[blockIdx.x] = (uint32_t)(acc >> 32);
    chunk_pairs[blockIdx.x] = accp;
  }
}

// Heavy bins: one workgroup per listed bin builds an LDS table value ->
// (count in A, count in B) by streaming both sides, then emits the values
// found on both sides.  Slots are claimed by compare-and-swap of the key
// itself against the all-ones sentinel (no lane ever waits on another);
// keys equal to the sentinel are counted apart.  A bin whose A side holds
// more than HT_SLOTS/2 distinct values sets the overflow flag.
constexpr uint32_t HT_SLOTS = 4096;
// records a k_rle_heavy workgroup keeps in its private region
constexpr uint32_t RLE_PRIV_CAP = 1024;
// heavy bins of at most RLE_HEAVY_SPLIT keys of A: 256-thread workgroups with
// RLE_HEAVY_SMALL_TS-slot tables (several per CU); larger ones: 1024 threads
// and HT_SLOTS slots.  Grids of each form; their private regions follow each
// other.
constexpr uint32_t RLE_HEAVY_SPLIT = 1024;
constexpr uint32_t RLE_HEAVY_SMALL_TS = 2048;
constexpr uint32_t RLE_HEAVY_GRID_S = 2048;
constexpr uint32_t RLE_HEAVY_GRID = 1024;
constexpr int HEAVY_RU = 4;  // keys per thread whose loads k_rle_heavy keeps in flight

__device__ __forceinline__ uint32_t lds_cas(uint32_t *p, uint32_t cmp, uint32_t v) { return atomicCAS(p, cmp, v); }
__device__ __forceinline__ uint64_t lds_cas(uint64_t *p, uint64_t cmp, uint64_t v) {
  return atomicCAS((unsigned long long *)p, (unsigned long long)cmp, (unsigned long long)v);
}

// Wave aggregation for the value tables: when every active lane of the wave
// holds the same value (a hot key's bin), one lane acts for all of them with
// the lane count; otherwise every lane acts for itself.  Returns the number
// of values this lane stands for (0: another lane does).
template <typename K>
__device__ __forceinline__ uint32_t wave_same_value(K v) {
  const uint64_t act = __ballot(1);
  const uint32_t first = (uint32_t)__ffsll((unsigned long long)act) - 1;
  const K v0 = __shfl(v, first, 64);
  const uint64_t same = __ballot(v == v0);
  if (same != act) return 1u;
  return (threadIdx.x & 63) == first ? (uint32_t)__popcll(act) : 0u;
}

// Adds c copies of v to an LDS value -> count table (compare-and-swap of the
// key against the EMPTY sentinel; no lane waits on another).
template <typename K, typename C>
__device__ __forceinline__ void vt_insert(K *s_key, C *s_cnt, uint32_t *s_fill, K v, uint32_t c,
                                          uint32_t mask = HT_SLOTS - 1) {
  constexpr K EMPTY = (K)~(K)0;
  uint32_t slot = murmur4((uint32_t)v ^ (uint32_t)((uint64_t)v >> 32)) & mask;
  for (uint32_t probe = 0; probe <= mask; probe++, slot = (slot + 1) & mask) {
    const K old = lds_cas(&s_key[slot], EMPTY, v);
    if (old == EMPTY || old == v) {
      atomicAdd(&s_cnt[slot], (C)c);
      if (old == EMPTY) atomicAdd(s_fill, 1u);
      return;
    }
  }
}

template <typename K, typename C>
__device__ __forceinline__ void vt_add_existing(const K *s_key, C *s_cnt, K v, uint32_t c,
                                                uint32_t mask = HT_SLOTS - 1) {
  constexpr K EMPTY = (K)~(K)0;
  uint32_t slot = murmur4((uint32_t)v ^ (uint32_t)((uint64_t)v >> 32)) & mask;
  for (uint32_t probe = 0; probe <= mask; probe++, slot = (slot + 1) & mask) {
    const K cur = s_key[slot];
    if (cur == EMPTY) return;
    if (cur == v) {
      atomicAdd(&s_cnt[slot], (C)c);
      return;
    }
  }
}

// ---------------------------------------------------------------------------
// Value table of the probe's and the join's long bins: HT_SLOTS values of A
// with u32 copy counts in LDS; keys equal to the all-ones sentinel are
// counted apart.  A bin of A is folded in rounds: before each chunk of BLOCK
// keys a table more than half full is answered (the probe keys look their
// values up) and cleared, so a chunk's inserts (at most BLOCK new values)
// always find a free slot.  Copies of one value folded in different rounds
// are answered in each round; their counts and emitted copies add up, so the
// result equals one answer over the whole bin.
// ---------------------------------------------------------------------------
template <typename K, int BLOCK>
__device__ __forceinline__ void vt_clear(K *s_key, uint32_t *s_cnt, uint32_t *s_fill, uint32_t *s_ecnt) {
  for (uint32_t i = threadIdx.x; i < HT_SLOTS; i += BLOCK) {
    s_key[i] = (K)~(K)0;
    s_cnt[i] = 0;
  }
  if (threadIdx.x == 0) {
    *s_fill = 0;
    *s_ecnt = 0;
  }
  __syncthreads();
}

// Every key of q[0..n) looks its value up in the table: the matches are
// added to *cursor and, with out, each key is written once per copy (one
// reservation per workgroup and pass).  Ends with a barrier.
template <typename K, int BLOCK>
__device__ void vt_answer(const K *s_key, const uint32_t *s_cnt, uint32_t ecnt, const K *__restrict__ q, uint64_t n,
                          K *__restrict__ out, uint64_t capacity, unsigned long long *cursor,
                          unsigned long long *s_t64, unsigned long long *s_base) {
  constexpr K EMPTY = (K)~(K)0;
  constexpr int U = 8;  // keys per thread per pass (loads in flight together)
  for (uint64_t c0 = 0; c0 < n; c0 += (uint64_t)BLOCK * U) {
    K kk[U];
    uint32_t mm[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t i = c0 + (uint64_t)u * BLOCK + threadIdx.x;
      kk[u] = i < n ? q[i] : (K)0;
    }
    unsigned long long mine = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t i = c0 + (uint64_t)u * BLOCK + threadIdx.x;
      const K k = kk[u];
      uint32_t m = 0;
      if (i < n) {
        if (k == EMPTY) {
          m = ecnt;
        } else {
          uint32_t slot = murmur4((uint32_t)k ^ (uint32_t)((uint64_t)k >> 32)) & (HT_SLOTS - 1);
          for (uint32_t pr = 0; pr < HT_SLOTS; pr++, slot = (slot + 1) & (HT_SLOTS - 1)) {
            const K cur = s_key[slot];
            if (cur == EMPTY) break;
            if (cur == k) {
              m = s_cnt[slot];
              break;
            }
          }
        }
      }
      mm[u] = m;
      mine += m;
    }
    if (out == nullptr) {
      const unsigned long long t = block_sum<BLOCK, unsigned long long>(mine, s_t64);
      if (threadIdx.x == 0 && t) atomicAdd(cursor, t);
      continue;
    }
    unsigned long long tot;
    const unsigned long long ex = block_exclusive_sum<BLOCK, unsigned long long>(mine, s_t64, tot);
    if (threadIdx.x == 0) *s_base = tot ? atomicAdd(cursor, tot) : 0ull;
    __syncthreads();
    unsigned long long w = *s_base + ex;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const unsigned long long m = mm[u];
      if (m && w < capacity) {
        const unsigned long long lim = min(m, (unsigned long long)capacity - w);
        for (unsigned long long r = 0; r < lim; r++) out[w + r] = kk[u];
      }
      w += m;
    }
    __syncthreads();
  }
}

// Folds the sequence key_at(0..n) into the table.  The keys are loaded
// FOLD_U * BLOCK at a time (all loads in flight together) and inserted
// BLOCK at a time; before each insert step a table more than half full is
// answered and cleared, so a step's inserts (at most BLOCK new values) always
// find a free slot.  Must start after a barrier; ends with one.  pending:
// values folded since the last answer.
constexpr int FOLD_U = 4;
template <typename K, int BLOCK, typename AT, typename ANSWER>
__device__ __forceinline__ void vt_fold_seq(uint64_t n, AT &&key_at, K *s_key, uint32_t *s_cnt, uint32_t *s_fill,
                                            uint32_t *s_ecnt, ANSWER &&answer, bool &pending) {
  constexpr K EMPTY = (K)~(K)0;
  for (uint64_t c0 = 0; c0 < n; c0 += (uint64_t)BLOCK * FOLD_U) {
    K v[FOLD_U];
#pragma unroll
    for (int u = 0; u < FOLD_U; u++) {
      const uint64_t i = c0 + (uint64_t)u * BLOCK + threadIdx.x;
      v[u] = i < n ? key_at(i) : EMPTY;
    }
#pragma unroll
    for (int u = 0; u < FOLD_U; u++) {
      if (c0 + (uint64_t)u * BLOCK >= n) break;  // uniform
      const bool full = *s_fill > HT_SLOTS / 2;  // uniform: read after a barrier
      __syncthreads();                           // every thread has read it before it changes
      if (full) {
        answer();
        vt_clear<K, BLOCK>(s_key, s_cnt, s_fill, s_ecnt);
      }
      if (c0 + (uint64_t)u * BLOCK + threadIdx.x < n) {
        const uint32_t c = wave_same_value(v[u]);
        if (c) {
          if (v[u] == EMPTY) atomicAdd(s_ecnt, c);
          else vt_insert(s_key, s_cnt, s_fill, v[u], c);
        }
      }
      __syncthreads();
      pending = true;
    }
  }
}

// Folds keys[a0..a1) into the table (vt_fold_seq over one bin).
template <typename K, int BLOCK, typename ANSWER>
__device__ __forceinline__ void vt_fold(const K *__restrict__ keys, uint64_t a0, uint64_t a1, K *s_key, uint32_t *s_cnt,
                                        uint32_t *s_fill, uint32_t *s_ecnt, ANSWER &&answer, bool &pending) {
  const K *kp = keys + a0;
  vt_fold_seq<K, BLOCK>(a1 - a0, [&](uint64_t i) { return kp[i]; }, s_key, s_cnt, s_fill, s_ecnt, answer, pending);
}

// ---------------------------------------------------------------------------
// Giant heavy bins of the run-length join (a hot key's bin: up to 10^8 keys
// per side).  Up to GIANT_CAP of them are split into chunks of GIANT_CHUNK
// keys over the whole grid: each chunk folds its keys into an LDS value table
// (wave-aggregated), then merges the table into the bin's global table of
// GT_SLOTS values (compare-and-swap on the key, atomic counts).  A's chunks
// insert, B's chunks only add to values present in A, and one workgroup per
// giant bin emits the records.  k_rle_heavy skips the bins taken here.
// ---------------------------------------------------------------------------
constexpr uint32_t GIANT_MIN = 1u << 18;
constexpr uint32_t GIANT_CAP = 256;
constexpr uint32_t GIANT_CHUNK = 1u << 16;
constexpr uint32_t GT_SLOTS = 4096;

struct GiantCtl {
  uint32_t n, totalA, totalB, pad;
  uint32_t nS, nL, pad2[2];  // k_rle_heavy's lists: bins of the small-table and HT_SLOTS forms
  uint32_t hidx[GIANT_CAP];
  uint32_t bad[GIANT_CAP];   // the bin has more distinct values than the tables hold (k_rle_ovf_* redo it)
  uint32_t chA[GIANT_CAP + 1], chB[GIANT_CAP + 1];
  unsigned long long eA[GIANT_CAP], eB[GIANT_CAP];
};

// Appends v to list at a wave-aggregated position (one atomic per wave).
__device__ __forceinline__ void wave_append(bool p, uint32_t *ctr, uint32_t *list, uint32_t v) {
  const uint64_t m = __ballot(p);
  if (m == 0) return;  // uniform over the wave
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  if (p) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
}

// Sorts the heavy bins: giant ones (up to GIANT_CAP) to ctl->hidx, the rest
// to the lists of k_rle_heavy's two forms (lists[0..), lists[heavy_cap..)).
template <typename OFF>
__global__ __launch_bounds__(256) void k_giant_mark(const OFF *__restrict__ offA, const OFF *__restrict__ offB,
                                                    const uint32_t *__restrict__ heavy, uint64_t nheavy,
                                                    uint32_t *__restrict__ lists, uint64_t heavy_cap,
                                                    GiantCtl *__restrict__ ctl) {
  for (uint64_t h = (uint64_t)blockIdx.x * 256 + threadIdx.x; h < nheavy; h += (uint64_t)gridDim.x * 256) {
    const uint32_t b = heavy[h];
    const uint64_t na = (uint64_t)(offA[b + 1] - offA[b]);
    const uint64_t sz = na + (uint64_t)(offB[b + 1] - offB[b]);
    bool giant = false;
    if (sz > GIANT_MIN) {
      const uint32_t s = atomicAdd(&ctl->n, 1u);
      if (s < GIANT_CAP) {
        giant = true;
        ctl->hidx[s] = (uint32_t)h;
      }
    }
    wave_append(!giant && na <= RLE_HEAVY_SPLIT, &ctl->nS, lists, b);
    wave_append(!giant && na > RLE_HEAVY_SPLIT, &ctl->nL, lists + heavy_cap, b);
  }
}

template <typename OFF>
__global__ __launch_bounds__(64) void k_giant_prefix(const OFF *__restrict__ offA, const OFF *__restrict__ offB,
                                                     const uint32_t *__restrict__ heavy, GiantCtl *__restrict__ ctl) {
  if (threadIdx.x != 0) return;
  const uint32_t n = min(ctl->n, GIANT_CAP);
  uint32_t a = 0, b = 0;
  for (uint32_t s = 0; s < n; s++) {
    const uint64_t bin = heavy[ctl->hidx[s]];
    ctl->chA[s] = a;
    ctl->chB[s] = b;
    a += (uint32_t)(((uint64_t)(offA[bin + 1] - offA[bin]) + GIANT_CHUNK - 1) / GIANT_CHUNK);
    b += (uint32_t)(((uint64_t)(offB[bin + 1] - offB[bin]) + GIANT_CHUNK - 1) / GIANT_CHUNK);
    ctl->eA[s] = 0;
    ctl->eB[s] = 0;
    ctl->bad[s] = 0;
  }
  ctl->chA[n] = a;
  ctl->chB[n] = b;
  ctl->totalA = a;
  ctl->totalB = b;
  ctl->n = n;
}

template <typename K>
__global__ __launch_bounds__(256) void k_giant_clear(const GiantCtl *__restrict__ ctl, K *__restrict__ gkey,
                                                     unsigned long long *__restrict__ gca,
                                                     unsigned long long *__restrict__ gcb) {
  const uint64_t n = (uint64_t)ctl->n * GT_SLOTS;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    gkey[i] = (K)~(K)0;
    gca[i] = 0;
    gcb[i] = 0;
  }
}

__device__ __forceinline__ uint32_t giant_slot_of(const uint32_t *ch, uint32_t n, uint32_t g) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ch[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// SIDE 0: A's chunks insert into the global tables; SIDE 1: B's chunks add
// to the values already there.
template <typename K, typename OFF, int SIDE>
__global__ __launch_bounds__(1024) void k_giant_fold(const OFF *__restrict__ off, const K *__restrict__ keys,
                                                     const uint32_t *__restrict__ heavy, GiantCtl *__restrict__ ctl,
                                                     K *__restrict__ gkey, unsigned long long *__restrict__ gcnt) {
  constexpr K EMPTY = (K)~(K)0;
  __shared__ K s_key[HT_SLOTS];
  __shared__ uint32_t s_cnt[HT_SLOTS];
  __shared__ uint32_t s_fill, s_e;
  const uint32_t n = ctl->n;
  const uint32_t *ch = SIDE == 0 ? ctl->chA : ctl->chB;
  const uint32_t total = SIDE == 0 ? ctl->totalA : ctl->totalB;
  for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
    const uint32_t slot = giant_slot_of(ch, n, g);
    const uint64_t bin = heavy[ctl->hidx[slot]];
    const uint64_t b0 = (uint64_t)off[bin], b1 = (uint64_t)off[bin + 1];
    const uint64_t c0 = b0 + (uint64_t)(g - ch[slot]) * GIANT_CHUNK;
    const uint64_t c1 = min(b1, c0 + GIANT_CHUNK);
    for (uint32_t i = threadIdx.x; i < HT_SLOTS; i += 1024) {
      s_key[i] = EMPTY;
      s_cnt[i] = 0;
    }
    if (threadIdx.x == 0) {
      s_fill = 0;
      s_e = 0;
    }
    __syncthreads();
    // BIG_UB keys per thread per step, their loads in flight together
    for (uint64_t i0 = c0; i0 < c1; i0 += 1024 * BIG_UB) {
      K vv[BIG_UB];
#pragma unroll
      for (int u = 0; u < BIG_UB; u++) {
        const uint64_t i = i0 + (uint64_t)u * 1024 + threadIdx.x;
        vv[u] = i < c1 ? keys[i] : EMPTY;
      }
#pragma unroll
      for (int u = 0; u < BIG_UB; u++) {
        if (i0 + (uint64_t)u * 1024 + threadIdx.x >= c1) continue;
        const K v = vv[u];
        const uint32_t c = wave_same_value(v);
        if (c == 0) continue;
        if (v == EMPTY) atomicAdd(&s_e, c);
        else vt_insert(s_key, s_cnt, &s_fill, v, c);
      }
    }
    __syncthreads();
    const bool ovf = s_fill > HT_SLOTS / 2;
    K *tk = gkey + (uint64_t)slot * GT_SLOTS;
    unsigned long long *tc = gcnt + (uint64_t)slot * GT_SLOTS;
    if (ovf) {
      if (threadIdx.x == 0) atomicOr(&ctl->bad[slot], 1u);
    } else {
      for (uint32_t i = threadIdx.x; i < HT_SLOTS; i += 1024) {
        const K v = s_key[i];
        if (v == EMPTY) continue;
        uint32_t q = murmur4((uint32_t)v ^ (uint32_t)((uint64_t)v >> 32)) & (GT_SLOTS - 1);
        bool done = false;
        for (uint32_t p = 0; p < GT_SLOTS && !done; p++, q = (q + 1) & (GT_SLOTS - 1)) {
          if (SIDE == 0) {
            const K old = atomicCAS((typename std::conditional<sizeof(K) == 4, unsigned int, unsigned long long>::type *)&tk[q],
                                    (typename std::conditional<sizeof(K) == 4, unsigned int, unsigned long long>::type)EMPTY,
                                    (typename std::conditional<sizeof(K) == 4, unsigned int, unsigned long long>::type)v);
            if (old == EMPTY || old == v) {
              atomicAdd(&tc[q], (unsigned long long)s_cnt[i]);
              done = true;
            }
          } else {
            const K cur = tk[q];
            if (cur == EMPTY) break;
            if (cur == v) {
              atomicAdd(&tc[q], (unsigned long long)s_cnt[i]);
              done = true;
            }
          }
        }
        if (SIDE == 0 && !done) atomicOr(&ctl->bad[slot], 1u);
      }
      if (threadIdx.x == 0 && s_e) atomicAdd(SIDE == 0 ? &ctl->eA[slot] : &ctl->eB[slot], (unsigned long long)s_e);
    }
    __syncthreads();
  }
}

// Records of the giant bins; a bin whose tables overflowed is listed for
// the global-table path instead (ovf_list).
template <typename K>
__global__ __launch_bounds__(1024) void k_giant_emit(const GiantCtl *__restrict__ ctl, const K *__restrict__ gkey,
                                                     const unsigned long long *__restrict__ gca,
                                                     const unsigned long long *__restrict__ gcb, K *__restrict__ rk,
                                                     unsigned long long *__restrict__ rca,
                                                     unsigned long long *__restrict__ rcb, uint64_t capacity,
                                                     unsigned long long *__restrict__ cursor,
                                                     unsigned long long *__restrict__ pairs,
                                                     const uint32_t *__restrict__ heavy, uint32_t *__restrict__ ovf_list,
                                                     unsigned int *__restrict__ ovf_n, uint64_t ovf_cap) {
  constexpr K EMPTY = (K)~(K)0;
  __shared__ unsigned long long s_t64[1024 / 64 + 1], s_base;
  for (uint32_t slot = blockIdx.x; slot < ctl->n; slot += gridDim.x) {
    if (ctl->bad[slot]) {  // uniform
      if (threadIdx.x == 0) {
        const uint32_t i = atomicAdd(ovf_n, 1u);
        if (i < ovf_cap) ovf_list[i] = heavy[ctl->hidx[slot]];
      }
      continue;
    }
    const K *tk = gkey + (uint64_t)slot * GT_SLOTS;
    const unsigned long long *ta = gca + (uint64_t)slot * GT_SLOTS, *tb = gcb + (uint64_t)slot * GT_SLOTS;
    for (uint32_t i0 = 0; i0 <= GT_SLOTS; i0 += 1024) {
      const uint32_t i = i0 + threadIdx.x;
      K v = EMPTY;
      unsigned long long a = 0, b = 0;
      if (i < GT_SLOTS) {
        v = tk[i];
        a = ta[i];
        b = tb[i];
        if (v == EMPTY) a = b = 0;
      } else if (i == GT_SLOTS) {  // keys equal to the sentinel, counted apart
        a = ctl->eA[slot];
        b = ctl->eB[slot];
      }
      const bool hit = a && b;
      unsigned long long tot;
      const unsigned long long ex = block_exclusive_sum<1024, unsigned long long>(hit ? 1ull : 0ull, s_t64, tot);
      const unsigned long long pt = block_sum<1024, unsigned long long>(hit ? a * b : 0ull, s_t64);
      if (threadIdx.x == 0) {
        s_base = tot ? atomicAdd(cursor, tot) : 0ull;
        if (pt) atomicAdd(pairs, pt);
      }
      __syncthreads();
      if (hit) {
        const unsigned long long w = s_base + ex;
        if (w < capacity) {
          rk[w] = v;
          rca[w] = a;
          rcb[w] = b;
        }
      }
      __syncthreads();
    }
  }
}

template <typename K, typename OFF, int BLOCK, uint32_t TS>
__global__ __launch_bounds__(BLOCK) void k_rle_heavy(const OFF *__restrict__ offA, const K *__restrict__ keysA,
                                                     const OFF *__restrict__ offB, const K *__restrict__ keysB,
                                                     K *__restrict__ rk, unsigned long long *__restrict__ rca,
                                                     unsigned long long *__restrict__ rcb, uint64_t capacity,
                                                     unsigned long long *__restrict__ cursor,
                                                     unsigned long long *__restrict__ pairs,
                                                     unsigned int *__restrict__ overflow,
                                                     const uint32_t *__restrict__ list,
                                                     const uint32_t *__restrict__ list_n, K *__restrict__ pk,
                                                     unsigned long long *__restrict__ pca,
                                                     unsigned long long *__restrict__ pcb,
                                                     uint32_t *__restrict__ priv_n, uint32_t region0,
                                                     uint32_t *__restrict__ ovf_list, unsigned int *__restrict__ ovf_n,
                                                     uint64_t ovf_cap) {
  constexpr K EMPTY = (K)~(K)0;
  __shared__ K s_key[TS];
  // 32-bit counts: a bin with 2^32 or more keys on a side is refused (overflow flag)
  __shared__ uint32_t s_ca[TS], s_cb[TS];
  __shared__ unsigned int s_fill, s_pn, s_priv;
  __shared__ unsigned long long s_eca, s_ecb, s_base, s_t64[BLOCK / 64 + 1];
  unsigned long long wpairs = 0;  // thread 0: the workgroup's pairs
  if (threadIdx.x == 0) s_pn = 0;
  const uint64_t pbase = (uint64_t)(region0 + blockIdx.x) * RLE_PRIV_CAP;
  // this form's bins (k_giant_mark's list: giant bins are k_giant_fold's)
  const uint32_t nlist = *list_n;
  for (uint32_t h = blockIdx.x; h < nlist; h += gridDim.x) {
    const uint64_t b = list[h];
    const uint64_t a0 = offA[b], a1 = offA[b + 1], x0 = offB[b], x1 = offB[b + 1];
    if (a1 - a0 > 0xFFFFFFFFull || x1 - x0 > 0xFFFFFFFFull) {
      if (threadIdx.x == 0) atomicOr(overflow, 1u);
      continue;
    }
    // a table of at least twice A's keys in the bin (its distinct values),
    // 64 to TS slots: small bins clear and scan a small table
    uint32_t slots = 64;
    while (slots < TS && (uint64_t)slots < 2 * (a1 - a0)) slots <<= 1;
    const uint32_t mask = slots - 1;
    for (uint32_t i = threadIdx.x; i < slots; i += BLOCK) {
      s_key[i] = EMPTY;
      s_ca[i] = 0;
      s_cb[i] = 0;
    }
    if (threadIdx.x == 0) {
      s_fill = 0;
      s_eca = 0;
      s_ecb = 0;
    }
    __syncthreads();
    // HEAVY_RU keys per thread per step, their loads in flight together
    for (uint64_t i0 = a0; i0 < a1; i0 += (uint64_t)BLOCK * HEAVY_RU) {
      K vv[HEAVY_RU];
#pragma unroll
      for (int u = 0; u < HEAVY_RU; u++) {
        const uint64_t i = i0 + (uint64_t)u * BLOCK + threadIdx.x;
        vv[u] = i < a1 ? keysA[i] : EMPTY;
      }
#pragma unroll
      for (int u = 0; u < HEAVY_RU; u++) {
        if (i0 + (uint64_t)u * BLOCK + threadIdx.x >= a1) continue;
        const K v = vv[u];
        const uint32_t c = wave_same_value(v);
        if (c == 0) continue;
        if (v == EMPTY) atomicAdd(&s_eca, (unsigned long long)c);
        else vt_insert(s_key, s_ca, &s_fill, v, c, mask);
      }
    }
    __syncthreads();
    if (s_fill > slots / 2) {  // too many distinct values for the table (only when slots == HT_SLOTS)
      if (threadIdx.x == 0) {  // the global-table path takes the bin
        const uint32_t i = atomicAdd(ovf_n, 1u);
        if (i < ovf_cap) ovf_list[i] = (uint32_t)b;
      }
      __syncthreads();
      continue;
    }
    for (uint64_t i0 = x0; i0 < x1; i0 += (uint64_t)BLOCK * HEAVY_RU) {
      K vv[HEAVY_RU];
#pragma unroll
      for (int u = 0; u < HEAVY_RU; u++) {
        const uint64_t i = i0 + (uint64_t)u * BLOCK + threadIdx.x;
        vv[u] = i < x1 ? keysB[i] : EMPTY;
      }
#pragma unroll
      for (int u = 0; u < HEAVY_RU; u++) {
        if (i0 + (uint64_t)u * BLOCK + threadIdx.x >= x1) continue;
        const K v = vv[u];
        const uint32_t c = wave_same_value(v);
        if (c == 0) continue;
        if (v == EMPTY) atomicAdd(&s_ecb, (unsigned long long)c);
        else vt_add_existing(s_key, s_cb, v, c, mask);
      }
    }
    __syncthreads();
    unsigned long long mine = 0, mp = 0;
    for (uint32_t i = threadIdx.x; i < slots; i += BLOCK)
      if (s_cb[i]) {
        mine++;
        mp += (unsigned long long)s_ca[i] * s_cb[i];
      }
    const bool ehit = threadIdx.x == 0 && s_eca && s_ecb;
    if (ehit) {
      mine++;
      mp += s_eca * s_ecb;
    }
    unsigned long long tot;
    unsigned long long ex = block_exclusive_sum<BLOCK, unsigned long long>(mine, s_t64, tot);
    const unsigned long long pt = block_sum<BLOCK, unsigned long long>(mp, s_t64);
    if (threadIdx.x == 0) {
      wpairs += pt;
      // the records go to the workgroup's private region while it has room
      // (k_rle_compact appends them), else straight to the output
      if ((uint64_t)s_pn + tot <= RLE_PRIV_CAP) {
        s_priv = 1;
        s_base = s_pn;
        s_pn += (uint32_t)tot;
      } else {
        s_priv = 0;
        s_base = tot ? atomicAdd(cursor, tot) : 0ull;
      }
    }
    __syncthreads();
    const bool priv = s_priv != 0;
    K *ok = priv ? pk + pbase : rk;
    unsigned long long *oca = priv ? pca + pbase : rca, *ocb = priv ? pcb + pbase : rcb;
    const uint64_t ocap = priv ? RLE_PRIV_CAP : capacity;
    unsigned long long w = s_base + ex;
    for (uint32_t i = threadIdx.x; i < slots; i += BLOCK)
      if (s_cb[i]) {
        if (w < ocap) {
          ok[w] = s_key[i];
          oca[w] = s_ca[i];
          ocb[w] = s_cb[i];
        }
        w++;
      }
    if (ehit) {
      if (w < ocap) {
        ok[w] = EMPTY;
        oca[w] = s_eca;
        ocb[w] = s_ecb;
      }
      w++;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    priv_n[region0 + blockIdx.x] = s_pn;
    if (wpairs) atomicAdd(pairs, wpairs);
  }
}

// ---------------------------------------------------------------------------
// Heavy bins of the run-length join with more distinct values than an LDS or
// giant-bin table holds (listed in ovf_list by k_rle_heavy / k_giant_emit).
// The host lays them out in batches of global-memory tables, one per bin,
// of at least twice the bin's keys of A (so every insert finds a slot):
// k_ovf_fold<0> inserts A's keys (compare-and-swap on the key, 32-bit
// counts), k_ovf_fold<1> adds B's keys to the values present, k_ovf_emit
// writes a record per value found on both sides.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ovf_sizes(const uint32_t *__restrict__ list, uint32_t n,
                                                   const void *__restrict__ offA, const void *__restrict__ offB,
                                                   int off_bytes, unsigned long long *__restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t b = list[i];
  unsigned long long a0, a1, x0, x1;
  if (off_bytes == 4) {
    a0 = ((const uint32_t *)offA)[b];
    a1 = ((const uint32_t *)offA)[b + 1];
    x0 = ((const uint32_t *)offB)[b];
    x1 = ((const uint32_t *)offB)[b + 1];
  } else {
    a0 = ((const unsigned long long *)offA)[b];
    a1 = ((const unsigned long long *)offA)[b + 1];
    x0 = ((const unsigned long long *)offB)[b];
    x1 = ((const unsigned long long *)offB)[b + 1];
  }
  out[4 * (uint64_t)i + 0] = a0;
  out[4 * (uint64_t)i + 1] = a1;
  out[4 * (uint64_t)i + 2] = x0;
  out[4 * (uint64_t)i + 3] = x1;
}

template <typename K, int SIDE>
__global__ __launch_bounds__(1024) void k_ovf_fold(const K *__restrict__ keys, const OvfChunk *__restrict__ chunks,
                                                   uint32_t nchunks, const OvfTable *__restrict__ tabs,
                                                   K *__restrict__ tkey, uint32_t *__restrict__ tcnt,
                                                   unsigned long long *__restrict__ ecnt) {
  using U = typename std::conditional<sizeof(K) == 4, unsigned int, unsigned long long>::type;
  constexpr K EMPTY = (K)~(K)0;
  for (uint32_t g = blockIdx.x; g < nchunks; g += gridDim.x) {
    const OvfChunk ch = chunks[g];
    const OvfTable tb = tabs[ch.tab];
    for (uint64_t i = ch.begin + threadIdx.x; i < ch.end; i += 1024) {
      const K v = keys[i];
      const uint32_t c = wave_same_value(v);
      if (c == 0) continue;
      if (v == EMPTY) {
        atomicAdd(&ecnt[ch.tab], (unsigned long long)c);
        continue;
      }
      uint64_t slot = murmur4((uint32_t)v ^ (uint32_t)((uint64_t)v >> 32)) & tb.mask;
      for (uint64_t p = 0; p <= tb.mask; p++, slot = (slot + 1) & tb.mask) {
        if (SIDE == 0) {
          const K old = (K)atomicCAS((U *)&tkey[tb.base + slot], (U)EMPTY, (U)v);
          if (old == EMPTY || old == v) {
            atomicAdd(&tcnt[tb.base + slot], c);
            break;
          }
        } else {
          const K cur = tkey[tb.base + slot];
          if (cur == EMPTY) break;
          if (cur == v) {
            atomicAdd(&tcnt[tb.base + slot], c);
            break;
          }
        }
      }
    }
  }
}

// One record per table slot holding a value present on both sides, then
// (last workgroup-sized pass) the sentinel-valued keys of each bin.
template <typename K>
__global__ __launch_bounds__(1024) void k_ovf_emit(const K *__restrict__ tkey, const uint32_t *__restrict__ tca,
                                                   const uint32_t *__restrict__ tcb, uint64_t nslots,
                                                   const unsigned long long *__restrict__ eA,
                                                   const unsigned long long *__restrict__ eB, uint32_t ntabs,
                                                   K *__restrict__ rk, unsigned long long *__restrict__ rca,
                                                   unsigned long long *__restrict__ rcb, uint64_t capacity,
                                                   unsigned long long *__restrict__ cursor,
                                                   unsigned long long *__restrict__ pairs) {
  constexpr K EMPTY = (K)~(K)0;
  __shared__ unsigned long long s_t64[1024 / 64 + 1], s_base;
  const uint64_t total = nslots + ntabs;  // table slots, then one sentinel entry per table
  for (uint64_t i0 = (uint64_t)blockIdx.x * 1024; i0 < total; i0 += (uint64_t)gridDim.x * 1024) {
    const uint64_t i = i0 + threadIdx.x;
    K v = EMPTY;
    unsigned long long a = 0, b = 0;
    if (i < nslots) {
      v = tkey[i];
      if (v != EMPTY) {
        a = tca[i];
        b = tcb[i];
      }
    } else if (i < total) {
      a = eA[i - nslots];
      b = eB[i - nslots];
    }
    const bool hit = a && b;
    unsigned long long tot;
    const unsigned long long ex = block_exclusive_sum<1024, unsigned long long>(hit ? 1ull : 0ull, s_t64, tot);
    const unsigned long long pt = block_sum<1024, unsigned long long>(hit ? a * b : 0ull, s_t64);
    if (threadIdx.x == 0) {
      s_base = tot ? atomicAdd(cursor, tot) : 0ull;
      if (pt) atomicAdd(pairs, pt);
    }
    __syncthreads();
    if (hit) {
      const unsigned long long w = s_base + ex;
      if (w < capacity) {
        rk[w] = v;
        rca[w] = a;
        rcb[w] = b;
      }
    }
    __syncthreads();
  }
}

// Appends the private regions of k_rle_heavy's workgroups after the records
// written so far (workgroup order), then k_rle_finish moves the cursor past
// them.
template <typename K>
__global__ __launch_bounds__(256) void k_rle_compact(const uint32_t *__restrict__ priv_n, const K *__restrict__ pk,
                                                     const unsigned long long *__restrict__ pca,
                                                     const unsigned long long *__restrict__ pcb, K *__restrict__ rk,
                                                     unsigned long long *__restrict__ rca,
                                                     unsigned long long *__restrict__ rcb, uint64_t capacity,
                                                     const unsigned long long *__restrict__ cursor) {
  __shared__ unsigned long long s_t64[256 / 64 + 1];
  unsigned long long before = 0;
  for (uint32_t i = threadIdx.x; i < blockIdx.x; i += 256) before += priv_n[i];
  before = block_sum<256, unsigned long long>(before, s_t64);
  const uint32_t n = priv_n[blockIdx.x];
  const unsigned long long base = *cursor + before;
  const uint64_t pbase = (uint64_t)blockIdx.x * RLE_PRIV_CAP;
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    const unsigned long long w = base + i;
    if (w < capacity) {
      rk[w] = pk[pbase + i];
      rca[w] = pca[pbase + i];
      rcb[w] = pcb[pbase + i];
    }
  }
}

__global__ __launch_bounds__(1024) void k_rle_finish(const uint32_t *__restrict__ priv_n, uint32_t nblk,
                                                     unsigned long long *__restrict__ cursor) {
  __shared__ unsigned long long s_t64[1024 / 64 + 1];
  unsigned long long v = 0;
  for (uint32_t i = threadIdx.x; i < nblk; i += 1024) v += priv_n[i];
  const unsigned long long tot = block_sum<1024, unsigned long long>(v, s_t64);
  if (threadIdx.x == 0) *cursor += tot;
}

// acc += sum of in[0..n) (one workgroup)
__global__ __launch_bounds__(1024) void k_sum_add(const unsigned long long *__restrict__ in, uint64_t n,
                                                  unsigned long long *__restrict__ acc) {
  __shared__ unsigned long long s_t64[1024 / 64 + 1];
  unsigned long long s = 0;
  for (uint64_t i = threadIdx.x; i < n; i += 1024) s += in[i];
  s = block_sum<1024, unsigned long long>(s, s_t64);
  if (threadIdx.x == 0 && s) atomicAdd(acc, s);
}

// ---------------------------------------------------------------------------
// Descriptors of a one-segment partition level written on the device (the
// no-allocation build uploads nothing from the host): tiles of tile_keys
// keys, the column-scan items (column blocks of 256, chunks of `chunk`
// rows) in the order the host's level layout uses, the segment and its map.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_level_setup(SegDesc sd, BinMap map, uint32_t tile_keys, uint32_t chunk,
                                                     TileDesc *__restrict__ tiles, SegDesc *__restrict__ segs,
                                                     ColItem *__restrict__ items, ColItem *__restrict__ items2,
                                                     uint64_t *__restrict__ psum_base, uint32_t *__restrict__ nchunks,
                                                     BinMap *__restrict__ maps) {
  const uint32_t ntiles = sd.ntiles;
  const uint32_t nch = (ntiles + chunk - 1) / chunk;
  const uint32_t ncb = (sd.ncols + 255) / 256;
  const uint32_t nitems = ncb * nch;
  const uint32_t total = ntiles + nitems + ncb + 1;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256) {
    if (i < ntiles) {
      TileDesc t;
      t.key_start = sd.key_begin + (uint64_t)i * tile_keys;
      t.count = (uint32_t)min((uint64_t)tile_keys, sd.key_end - t.key_start);
      t.seg = 0;
      tiles[i] = t;
    } else if (i < ntiles + nitems) {
      const uint32_t j = i - ntiles, cb = j / nch, ch = j % nch;
      ColItem x{};
      x.seg = 0;
      x.chunk = ch;
      x.row_lo = ch * chunk;
      x.row_hi = min(ntiles, (ch + 1) * chunk);
      x.col_lo = cb * 256;
      items[j] = x;
    } else if (i < ntiles + nitems + ncb) {
      ColItem x{};
      x.seg = 0;
      x.col_lo = (i - ntiles - nitems) * 256;
      items2[i - ntiles - nitems] = x;
    } else {
      segs[0] = sd;
      maps[0] = map;
      psum_base[0] = 0;
      nchunks[0] = nch;
    }
  }
}

hipError_t launch_level_setup(const SegDesc &sd, const BinMap &map, uint32_t tile_keys, uint32_t chunk, TileDesc *tiles,
                              SegDesc *segs, ColItem *items, ColItem *items2, uint64_t *psum_base, uint32_t *nchunks,
                              BinMap *maps, hipStream_t st) {
  hipLaunchKernelGGL(k_level_setup, dim3(64), dim3(256), 0, st, sd, map, tile_keys, chunk, tiles, segs, items, items2,
                     psum_base, nchunks, maps);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Work items of the final stage / probe from the partition pass results: one
// item per column (bucket) of every segment.
// ---------------------------------------------------------------------------
__global__ void k_make_items(const SegDesc *__restrict__ segs, uint32_t nsegs, const uint32_t *__restrict__ bstart,
                             const uint32_t *__restrict__ coltot, FinItem *__restrict__ items, uint32_t ncols_total) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncols_total) return;
  uint32_t lo = 0, hi = nsegs;  // last segment with bucket_base <= g
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (segs[mid].bucket_base <= g) lo = mid; else hi = mid;
  }
  const SegDesc &sd = segs[lo];
  const uint32_t b = g - (uint32_t)sd.bucket_base;
  FinItem it;
  it.table = sd.table;
  it.key_rel = sd.fin_key_rel;
  if (b >= sd.nbuckets || sd.kind != 0) {  // drop column
    it.key_start = 0; it.count = 0; it.nbins = 0; it.bin_lo = 0; it.off_pos = 0; it.end_pos = ~0ull;
  } else {
    const uint64_t blo = (uint64_t)b * sd.div.d;
    it.key_start = bstart[g];
    it.count = coltot[g];
    it.bin_lo = (uint32_t)(sd.bin_lo + blo);
    it.nbins = (uint32_t)min((uint64_t)sd.div.d, sd.seg_nbins - blo);
    it.off_pos = sd.fin_off_base + blo;
    it.end_pos = (b + 1 == sd.nbuckets) ? sd.fin_end_pos : ~0ull;
  }
  items[g] = it;
}

// ---------------------------------------------------------------------------
// Oversized buckets (skewed keys: a hot key's bin can hold 10^8 keys).  Up to
// BIG_CAP of them are split into chunks of BIG_CHUNK keys and spread over the
// whole grid: per-chunk LDS bin counts are added into global per-bin
// counters, one workgroup per bucket scans them into the offsets, and each
// chunk then reserves its places bin by bin (one global atomic per chunk and
// non-empty bin) and writes its keys.  A wave whose keys all fall into one
// bin adds them with one LDS atomic.  Bins are handled in passes of
// BIG_BINS.  Buckets beyond BIG_CAP take k_final_big.
// ---------------------------------------------------------------------------
struct BigCtl {
  uint32_t nbig, total_chunks, total_bins;
  uint32_t seen;  // k_final left an oversized item (zeroed before k_final)
  uint32_t item[BIG_CAP];
  uint32_t chunk_start[BIG_CAP + 1];
  uint32_t bin_base[BIG_CAP + 1];
};

// Takes the oversized items (count > small_cap) in order while their bins
// fit big_cap counters (and at most BIG_CAP of them); prefix sums give chunk
// (chunk_keys keys each) and bin bases.  *seen == 0: no oversized item.
__global__ __launch_bounds__(1024) void k_big_list(const FinItem *__restrict__ items, uint32_t nitems, uint32_t small_cap,
                                                   uint32_t big_cap, uint32_t chunk_keys, const unsigned int *seen,
                                                   uint32_t *__restrict__ big_slot, BigCtl *__restrict__ ctl) {
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  __shared__ uint32_t s_slot, s_ch, s_bin, s_nacc, s_acc_ch, s_acc_bin;
  if (*seen == 0) {  // no oversized item: empty lists (the big-item kernels read no big_slot)
    if (threadIdx.x == 0) {
      ctl->chunk_start[0] = 0;
      ctl->bin_base[0] = 0;
      ctl->nbig = 0;
      ctl->total_chunks = 0;
      ctl->total_bins = 0;
    }
    return;
  }
  if (threadIdx.x == 0) s_slot = s_ch = s_bin = s_nacc = s_acc_ch = s_acc_bin = 0;
  __syncthreads();
  for (uint32_t i0 = 0; i0 < nitems; i0 += 1024) {
    const uint32_t i = i0 + threadIdx.x;
    FinItem it{};
    if (i < nitems) it = items[i];
    const bool big = i < nitems && it.count > small_cap;
    const uint32_t nch = big ? (uint32_t)((it.count + chunk_keys - 1) / chunk_keys) : 0u;
    const uint32_t nbn = big ? it.nbins : 0u;
    uint32_t t_slot, t_ch, t_bin;
    const uint32_t e_slot = block_exclusive_sum<1024, uint32_t>(big ? 1u : 0u, s_tmp, t_slot);
    const uint32_t e_ch = block_exclusive_sum<1024, uint32_t>(nch, s_tmp, t_ch);
    const uint32_t e_bin = block_exclusive_sum<1024, uint32_t>(nbn, s_tmp, t_bin);
    const uint32_t slot = s_slot + e_slot, cb = s_ch + e_ch, bb = s_bin + e_bin;
    const bool acc = big && slot < BIG_CAP && (uint64_t)bb + nbn <= big_cap;
    if (i < nitems) big_slot[i] = acc ? slot : ~0u;
    if (acc) {
      ctl->item[slot] = i;
      ctl->chunk_start[slot] = cb;
      ctl->bin_base[slot] = bb;
    }
    uint32_t a_tot, c_tot, b_tot;
    (void)block_exclusive_sum<1024, uint32_t>(acc ? 1u : 0u, s_tmp, a_tot);
    (void)block_exclusive_sum<1024, uint32_t>(acc ? nch : 0u, s_tmp, c_tot);
    (void)block_exclusive_sum<1024, uint32_t>(acc ? nbn : 0u, s_tmp, b_tot);
    if (threadIdx.x == 0) {
      s_slot += t_slot;
      s_ch += t_ch;
      s_bin += t_bin;
      s_nacc += a_tot;
      s_acc_ch += c_tot;
      s_acc_bin += b_tot;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint32_t nbig = s_nacc;
    ctl->chunk_start[nbig] = s_acc_ch;
    ctl->bin_base[nbig] = s_acc_bin;
    ctl->nbig = nbig;
    ctl->total_chunks = s_acc_ch;
    ctl->total_bins = s_acc_bin;
  }
}

__global__ __launch_bounds__(256) void k_big_zero(const BigCtl *__restrict__ ctl, uint32_t *__restrict__ gcnt) {
  const uint32_t n = ctl->total_bins;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) gcnt[i] = 0;
}

// slot of chunk g: last s with chunk_start[s] <= g
__device__ __forceinline__ uint32_t big_slot_of(const BigCtl *ctl, uint32_t g) {
  uint32_t lo = 0, hi = ctl->nbig;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ctl->chunk_start[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// LDS add of 1 at l for every valid lane; a wave whose valid lanes share one
// l adds them with one atomic.  Returns each lane's old value + its rank
// among the lanes with that l when RET (positions), else 0.
template <bool RET>
__device__ __forceinline__ uint32_t lds_add_agg(uint32_t *cnt, uint32_t l, bool valid) {
  const uint64_t act = __ballot(valid);
  if (act == 0) return 0;
  const uint32_t first = (uint32_t)__ffsll((unsigned long long)act) - 1;
  const uint32_t l0 = __shfl(l, first, 64);
  const uint64_t same = __ballot(valid && l == l0);
  const uint32_t lane = threadIdx.x & 63;
  if (same == act) {
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(&cnt[l0], (uint32_t)__popcll(act));
    if (!RET) return 0;
    base = __shfl(base, first, 64);
    return base + (uint32_t)__popcll(act & ((1ull << lane) - 1ull));
  }
  uint32_t r = 0;
  if (valid) r = atomicAdd(&cnt[l], 1u);
  return r;
}

template <typename K, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_big_count(const K *__restrict__ part, const FinItem *__restrict__ items,
                                                     const BinMap *__restrict__ maps, const BigCtl *__restrict__ ctl,
                                                     uint32_t *__restrict__ gcnt) {
  __shared__ uint32_t s_cnt[BIG_BINS];
  const uint32_t total = ctl->total_chunks;
  for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
    const uint32_t slot = big_slot_of(ctl, g);
    const FinItem it = items[ctl->item[slot]];
    const BinMap mp = maps[it.table];
    const uint64_t c0 = (uint64_t)(g - ctl->chunk_start[slot]) * BIG_CHUNK;
    const uint32_t n = (uint32_t)min((uint64_t)BIG_CHUNK, it.count - c0);
    const K *kp = part + it.key_start + c0;
    for (uint32_t b0 = 0; b0 < it.nbins; b0 += BIG_BINS) {
      const uint32_t nb = min((uint32_t)BIG_BINS, it.nbins - b0);
      for (uint32_t i = threadIdx.x; i < nb; i += BLOCK) s_cnt[i] = 0;
      __syncthreads();
      // BIG_UB keys per thread per step, their loads in flight together
      for (uint32_t i0 = 0; i0 < n; i0 += BLOCK * BIG_UB) {
        K kk[BIG_UB];
#pragma unroll
        for (int u = 0; u < BIG_UB; u++) {
          const uint32_t i = i0 + u * BLOCK + threadIdx.x;
          kk[u] = i < n ? kp[i] : (K)0;
        }
#pragma unroll
        for (int u = 0; u < BIG_UB; u++) {
          const uint32_t i = i0 + u * BLOCK + threadIdx.x;
          uint32_t l = 0;
          if (i < n) l = bin_of<K>(kk[u], mp) - it.bin_lo - b0;
          lds_add_agg<false>(s_cnt, l, i < n && l < nb);
        }
      }
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < nb; i += BLOCK)
        if (s_cnt[i]) atomicAdd(&gcnt[ctl->bin_base[slot] + b0 + i], s_cnt[i]);
      __syncthreads();
    }
  }
}

// One workgroup per oversized bucket: offsets from the global bin counts;
// gcnt becomes each bin's next free place (relative to the bucket).
template <typename OFF>
__global__ __launch_bounds__(1024) void k_big_scan(const FinItem *__restrict__ items, const BigCtl *__restrict__ ctl,
                                                   uint32_t *__restrict__ gcnt, OFF *__restrict__ off) {
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  for (uint32_t slot = blockIdx.x; slot < ctl->nbig; slot += gridDim.x) {
    const FinItem it = items[ctl->item[slot]];
    uint32_t *g = gcnt + ctl->bin_base[slot];
    const uint64_t obase = it.key_start - it.key_rel;
    uint32_t run = 0;
    for (uint32_t l0 = 0; l0 < it.nbins; l0 += 1024) {
      const uint32_t l = l0 + threadIdx.x;
      const uint32_t v = l < it.nbins ? g[l] : 0u;
      uint32_t tot;
      const uint32_t ex = block_exclusive_sum<1024, uint32_t>(v, s_tmp, tot);
      if (l < it.nbins) {
        g[l] = run + ex;
        off[it.off_pos + l] = (OFF)(obase + run + ex);
      }
      run += tot;
    }
    if (it.end_pos != ~0ull && threadIdx.x == 0) off[it.end_pos] = (OFF)(obase + it.count);
    __syncthreads();
  }
}

template <typename K, int BLOCK, bool SPLIT = false>
__global__ __launch_bounds__(BLOCK) void k_big_place(const K *__restrict__ part, K *__restrict__ csr,
                                                     const FinItem *__restrict__ items, const BinMap *__restrict__ maps,
                                                     const BigCtl *__restrict__ ctl, uint32_t *__restrict__ gcur,
                                                     uint64_t *__restrict__ csr_index) {
  __shared__ uint32_t s_cnt[BIG_BINS];
  const uint32_t total = ctl->total_chunks;
  for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
    const uint32_t slot = big_slot_of(ctl, g);
    const FinItem it = items[ctl->item[slot]];
    const BinMap mp = maps[it.table];
    const uint64_t c0 = (uint64_t)(g - ctl->chunk_start[slot]) * BIG_CHUNK;
    const uint32_t n = (uint32_t)min((uint64_t)BIG_CHUNK, it.count - c0);
    const K *kp = part + it.key_start + c0;
    K *dst = csr + it.key_start;
    for (uint32_t b0 = 0; b0 < it.nbins; b0 += BIG_BINS) {
      const uint32_t nb = min((uint32_t)BIG_BINS, it.nbins - b0);
      for (uint32_t i = threadIdx.x; i < nb; i += BLOCK) s_cnt[i] = 0;
      __syncthreads();
      for (uint32_t i0 = 0; i0 < n; i0 += BLOCK * BIG_UB) {
        K kk[BIG_UB];
#pragma unroll
        for (int u = 0; u < BIG_UB; u++) {
          const uint32_t i = i0 + u * BLOCK + threadIdx.x;
          kk[u] = i < n ? kp[i] : (K)0;
        }
#pragma unroll
        for (int u = 0; u < BIG_UB; u++) {
          const uint32_t i = i0 + u * BLOCK + threadIdx.x;
          uint32_t l = 0;
          if (i < n) l = bin_of<K>(kk[u], mp) - it.bin_lo - b0;
          lds_add_agg<false>(s_cnt, l, i < n && l < nb);
        }
      }
      __syncthreads();
      // reserve this chunk's places, bin by bin (RU returning atomics of a
      // thread in flight together)
      {
        constexpr int RU = 8;
        uint32_t *gb = gcur + ctl->bin_base[slot] + b0;
        for (uint32_t i0 = threadIdx.x; i0 < nb; i0 += BLOCK * RU) {
          uint32_t c[RU], r[RU];
#pragma unroll
          for (int j = 0; j < RU; j++) {
            const uint32_t i = i0 + j * BLOCK;
            c[j] = i < nb ? s_cnt[i] : 0u;
          }
#pragma unroll
          for (int j = 0; j < RU; j++) r[j] = c[j] ? atomicAdd(&gb[i0 + j * BLOCK], c[j]) : 0u;
#pragma unroll
          for (int j = 0; j < RU; j++)
            if (c[j]) s_cnt[i0 + j * BLOCK] = r[j];
        }
      }
      __syncthreads();
      for (uint32_t i0 = 0; i0 < n; i0 += BLOCK * BIG_UB) {
        K kk[BIG_UB];
#pragma unroll
        for (int u = 0; u < BIG_UB; u++) {
          const uint32_t i = i0 + u * BLOCK + threadIdx.x;
          kk[u] = i < n ? kp[i] : (K)0;
        }
#pragma unroll
        for (int u = 0; u < BIG_UB; u++) {
          const uint32_t i = i0 + u * BLOCK + threadIdx.x;
          uint32_t l = 0;
          if (i < n) l = bin_of<K>(kk[u], mp) - it.bin_lo - b0;
          const bool v = i < n && l < nb;
          const uint32_t pos = lds_add_agg<true>(s_cnt, l, v);
          if (v) {
            if constexpr (SPLIT)
              split_store((uint32_t *)csr + it.key_start, csr_index + it.key_start, pos, kk[u]);
            else
              dst[pos] = kk[u];
          }
        }
      }
      __syncthreads();
    }
  }
}

// Probe of buckets too large for k_probe: B's keys of the bucket are
// streamed and each walks its bin of A in global memory.  probe_big_range
// handles keys [k0, k1) of item `ii` (a workgroup-uniform range).
constexpr uint32_t PROBE_BIG_CHUNK = 32768;  // keys of an oversized bucket per workgroup task
constexpr uint32_t PROBE_BIG_GRID = 1024;
template <typename K, typename OFF, int BLOCK>
__device__ void probe_big_range(const K *__restrict__ bpart, const FinItem &it, uint32_t ii, const BinMap &mp,
                                const OFF *__restrict__ offA, const K *__restrict__ keysA, K *__restrict__ out,
                                uint64_t capacity, unsigned long long *__restrict__ cursor,
                                unsigned int *__restrict__ hmark, unsigned int *__restrict__ gate, uint64_t k0,
                                uint64_t k1, unsigned long long *s_t64, unsigned long long *s_base,
                                const uint64_t *__restrict__ out_off) {
  const K *kp = bpart + it.key_start;
  const K *ka = keysA + it.key_rel;
  uint64_t cap;
  K *o = out_region(out, capacity, out_off, it.table, cap);
  // U keys per thread per pass: every load of a pass is in flight at once
  constexpr int U = 8;
  for (uint64_t c0 = k0; c0 < k1; c0 += (uint64_t)BLOCK * U) {
    K kk[U];
    uint64_t ss[U], ee[U];
    uint32_t mm[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t i = c0 + (uint64_t)u * BLOCK + threadIdx.x;
      kk[u] = i < k1 ? kp[i] : (K)0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t i = c0 + (uint64_t)u * BLOCK + threadIdx.x;
      ss[u] = ee[u] = 0;
      if (i < k1) {
        const uint64_t l = bin_of<K>(kk[u], mp) - it.bin_lo;
        ss[u] = (uint64_t)offA[it.off_pos + l];
        ee[u] = (uint64_t)offA[it.off_pos + l + 1];
      }
    }
    unsigned long long mine = 0;
    uint32_t maxlen = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (ee[u] - ss[u] > PROBE_BIN_LIMIT) {
        hmark[ii] = 1u;  // k_probe_heavy answers this bin
        gate[1] = 1u;
        ee[u] = ss[u];
      }
      mm[u] = 0;
      maxlen = max(maxlen, (uint32_t)(ee[u] - ss[u]));
    }
    // the walks of the thread's U keys advance together, WG keys of every
    // walk per step: all U * WG loads of a step are in flight at once
    constexpr int WG = 4;
    for (uint32_t g = 0; g < maxlen; g += WG) {
      K v[U][WG];
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int i = 0; i < WG; i++) {
          const uint64_t q = ss[u] + g + i;
          v[u][i] = q < ee[u] ? ka[q] : (K)~kk[u];
        }
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int i = 0; i < WG; i++) mm[u] += v[u][i] == kk[u];
    }
#pragma unroll
    for (int u = 0; u < U; u++) mine += mm[u];
    if (out == nullptr) {
      const unsigned long long t = block_sum<BLOCK, unsigned long long>(mine, s_t64);
      if (threadIdx.x == 0 && t) atomicAdd(cursor + it.table, t);
      continue;
    }
    unsigned long long tot;
    const unsigned long long ex = block_exclusive_sum<BLOCK, unsigned long long>(mine, s_t64, tot);
    if (threadIdx.x == 0) *s_base = tot ? atomicAdd(cursor + it.table, tot) : 0ull;
    __syncthreads();
    unsigned long long w = *s_base + ex;
#pragma unroll
    for (int u = 0; u < U; u++)
      for (uint64_t q = ss[u]; q < ee[u] && mm[u]; q++)
        if (ka[q] == kk[u]) {
          if (w < cap) o[w] = kk[u];
          w++;
        }
    __syncthreads();
  }
}

// Oversized buckets that k_big_list did not list (big_slot ~0: beyond
// BIG_CAP of them): one workgroup per bucket.
template <typename K, typename OFF, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_probe_big(const K *__restrict__ bpart, const FinItem *__restrict__ items,
                                                     const BinMap *__restrict__ maps, const OFF *__restrict__ offA,
                                                     const K *__restrict__ keysA, K *__restrict__ out, uint64_t capacity,
                                                     unsigned long long *__restrict__ cursor, uint32_t small_cap,
                                                     const uint32_t *__restrict__ big_slot,
                                                     unsigned int *__restrict__ hmark,
                                                     unsigned int *__restrict__ gate,
                                                     const uint64_t *__restrict__ out_off) {
  __shared__ unsigned long long s_t64[BLOCK / 64 + 1];
  __shared__ unsigned long long s_base;
  if (gate[0] == 0) return;  // k_probe1 took every item
  const FinItem it = items[blockIdx.x];
  if (it.count <= small_cap || big_slot[blockIdx.x] != ~0u) return;
  const BinMap mp = maps[it.table];
  probe_big_range<K, OFF, BLOCK>(bpart, it, blockIdx.x, mp, offA, keysA, out, capacity, cursor, hmark, gate, 0,
                                 it.count, s_t64, &s_base, out_off);
}

// The listed oversized buckets, cut into tasks of PROBE_BIG_CHUNK keys that
// the grid shares (a bucket of a hot or out-of-distribution range of B keys
// no longer runs on one workgroup).
template <typename K, typename OFF, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_probe_big_chunks(const K *__restrict__ bpart, const FinItem *__restrict__ items,
                                                            const BinMap *__restrict__ maps, const OFF *__restrict__ offA,
                                                            const K *__restrict__ keysA, K *__restrict__ out,
                                                            uint64_t capacity, unsigned long long *__restrict__ cursor,
                                                            const BigCtl *__restrict__ ctl,
                                                            unsigned int *__restrict__ hmark,
                                                            unsigned int *__restrict__ gate,
                                                            const uint64_t *__restrict__ out_off) {
  __shared__ unsigned long long s_t64[BLOCK / 64 + 1];
  __shared__ unsigned long long s_base;
  if (gate[0] == 0) return;  // k_probe1 took every item (the list was not written)
  const uint32_t total = ctl->total_chunks;
  for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
    const uint32_t slot = big_slot_of(ctl, g);
    const uint32_t ii = ctl->item[slot];
    const FinItem it = items[ii];
    const BinMap mp = maps[it.table];
    const uint64_t k0 = (uint64_t)(g - ctl->chunk_start[slot]) * PROBE_BIG_CHUNK;
    const uint64_t k1 = min((uint64_t)it.count, k0 + PROBE_BIG_CHUNK);
    probe_big_range<K, OFF, BLOCK>(bpart, it, ii, mp, offA, keysA, out, capacity, cursor, hmark, gate, k0, k1, s_t64,
                                   &s_base, out_off);
  }
}

// Bins longer than PROBE_BIN_LIMIT (duplicates) of the items k_probe1, k_probe
// and k_probe_big marked: one workgroup per marked item folds the bucket's
// long bins of A into one LDS value table (a value lives in one bin, so the
// bins share the table).  The bucket's offsets are read HEAVY_UB bins per
// thread at a time; the long bins found are listed in LDS (HEAVY_LIST at a
// time, with the exclusive prefix of their lengths) and their keys are
// folded as one concatenated sequence (vt_fold_seq: a key's bin is found by
// a binary search of the prefix), so short long bins share fold steps.  Each
// answer of the table looks every probe key of the bucket up (a value found
// is in a long bin; the walking kernels skipped those).  Bins with more
// distinct values than the table holds are folded over several rounds.
constexpr int HEAVY_UB = 8;
constexpr uint32_t HEAVY_LIST = 1024;
template <typename K, typename OFF, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_probe_heavy(const K *__restrict__ bpart, const FinItem *__restrict__ items,
                                                       const OFF *__restrict__ offA, const K *__restrict__ keysA,
                                                       K *__restrict__ out, uint64_t capacity,
                                                       unsigned long long *__restrict__ cursor,
                                                       const unsigned int *__restrict__ hmark,
                                                       const unsigned int *__restrict__ gate,
                                                       const uint64_t *__restrict__ out_off) {
  static_assert(HEAVY_LIST == BLOCK, "one listed bin per thread in the prefix step");
  __shared__ K s_key[HT_SLOTS];
  __shared__ uint32_t s_cnt[HT_SLOTS];
  __shared__ uint64_t s_hs[HEAVY_LIST];      // first key (index into A's bucket) of each listed bin
  __shared__ uint64_t s_hx[HEAVY_LIST + 1];  // exclusive prefix of the listed bins' lengths
  __shared__ uint32_t s_fill, s_ecnt;
  __shared__ uint32_t s_t32[BLOCK / 64 + 1];
  __shared__ unsigned long long s_t64[BLOCK / 64 + 1], s_base;
  if (gate[1] == 0 || !hmark[blockIdx.x]) return;
  const FinItem it = items[blockIdx.x];
  const K *kp = bpart + it.key_start;
  const K *ka = keysA + it.key_rel;
  const OFF *of = offA + it.off_pos;  // of[0..nbins] are the bucket's offsets
  uint64_t cap;
  K *o = out_region(out, capacity, out_off, it.table, cap);
  auto answer = [&]() {
    vt_answer<K, BLOCK>(s_key, s_cnt, s_ecnt, kp, it.count, o, cap, cursor + it.table, s_t64, &s_base);
  };
  vt_clear<K, BLOCK>(s_key, s_cnt, &s_fill, &s_ecnt);
  bool pending = false;  // values folded since the last answer (uniform)
  for (uint64_t w0 = 0; w0 < it.nbins; w0 += (uint64_t)BLOCK * HEAVY_UB) {
    // this thread's bins: l0 .. l0 + HEAVY_UB - 1
    const uint64_t l0 = w0 + (uint64_t)threadIdx.x * HEAVY_UB;
    uint64_t o[HEAVY_UB + 1];
#pragma unroll
    for (int i = 0; i <= HEAVY_UB; i++) o[i] = l0 + i <= it.nbins ? (uint64_t)of[l0 + i] : 0ull;
    uint32_t hm = 0;
#pragma unroll
    for (int i = 0; i < HEAVY_UB; i++)
      if (l0 + i < it.nbins && o[i + 1] - o[i] > PROBE_BIN_LIMIT) hm |= 1u << i;
    uint32_t nlong;
    const uint32_t ex = block_exclusive_sum<BLOCK, uint32_t>((uint32_t)__popc(hm), s_t32, nlong);
    for (uint32_t p0 = 0; p0 < nlong; p0 += HEAVY_LIST) {
      const uint32_t nh = min(HEAVY_LIST, nlong - p0);
      uint32_t pos = ex;
#pragma unroll
      for (int i = 0; i < HEAVY_UB; i++) {
        if ((hm >> i) & 1u) {
          if (pos >= p0 && pos < p0 + nh) {
            s_hs[pos - p0] = o[i];
            s_hx[pos - p0] = o[i + 1] - o[i];  // length; the prefix replaces it below
          }
          pos++;
        }
      }
      __syncthreads();
      const uint64_t len = threadIdx.x < nh ? s_hx[threadIdx.x] : 0ull;
      unsigned long long tot;
      const unsigned long long lx = block_exclusive_sum<BLOCK, unsigned long long>(len, s_t64, tot);
      if (threadIdx.x < nh) s_hx[threadIdx.x] = lx;
      if (threadIdx.x == 0) s_hx[nh] = tot;
      __syncthreads();
      // key i of the sequence: bin j with s_hx[j] <= i < s_hx[j + 1]
      auto key_at = [&](uint64_t i) {
        uint32_t lo = 0, hi = nh;  // s_hx[lo] <= i < s_hx[hi]
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (s_hx[mid] <= i) lo = mid;
          else hi = mid;
        }
        return ka[s_hs[lo] + (i - s_hx[lo])];
      };
      vt_fold_seq<K, BLOCK>(tot, key_at, s_key, s_cnt, &s_fill, &s_ecnt, answer, pending);
    }
  }
  if (pending) answer();
}

// Join bins of two built tables above JOIN_PAIR_LIMIT key pairs, listed by
// k_join: one workgroup per bin folds A's keys of the bin into the value
// table (in rounds) and answers it with B's keys of the bin, so each B key x
// is counted and emitted cA(x) times, as the nested loop would.
template <typename K, typename OFF, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_join_heavy(const OFF *__restrict__ offA, const K *__restrict__ keysA,
                                                      const OFF *__restrict__ offB, const K *__restrict__ keysB,
                                                      K *__restrict__ out, uint64_t capacity,
                                                      unsigned long long *__restrict__ cursor,
                                                      const uint32_t *__restrict__ list,
                                                      const unsigned int *__restrict__ list_n, uint32_t list_cap) {
  __shared__ K s_key[HT_SLOTS];
  __shared__ uint32_t s_cnt[HT_SLOTS];
  __shared__ uint32_t s_fill, s_ecnt;
  __shared__ unsigned long long s_t64[BLOCK / 64 + 1], s_base;
  const uint32_t n = min(*list_n, list_cap);
  for (uint32_t h = blockIdx.x; h < n; h += gridDim.x) {
    const uint64_t b = list[h];
    const uint64_t a0 = (uint64_t)offA[b], a1 = (uint64_t)offA[b + 1];
    const uint64_t x0 = (uint64_t)offB[b], x1 = (uint64_t)offB[b + 1];
    auto answer = [&]() {
      vt_answer<K, BLOCK>(s_key, s_cnt, s_ecnt, keysB + x0, x1 - x0, out, capacity, cursor, s_t64, &s_base);
    };
    vt_clear<K, BLOCK>(s_key, s_cnt, &s_fill, &s_ecnt);
    bool pending = false;
    vt_fold<K, BLOCK>(keysA, a0, a1, s_key, s_cnt, &s_fill, &s_ecnt, answer, pending);
    if (pending) answer();
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int V> using IC = std::integral_constant<int, V>;

// Calls f(IC<hash kind>) for the specialisation of hash kind hk.
template <typename F> static hipError_t dispatch_hk(int hk, F &&f) {
  if (hk == HK_MURMUR) return f(IC<HK_MURMUR>{});
  if (hk == HK_RMI) return f(IC<HK_RMI>{});
  return f(IC<HK_ANY>{});
}
// Calls f(IC<hash kind>, IC<bucket kind>); a mixed bucket kind takes the
// fully general kernel.
template <typename F> static hipError_t dispatch_part(int hk, int sk, F &&f) {
  if (sk != SK_RANGE && sk != SK_SPLIT) return f(IC<HK_ANY>{}, IC<SK_ANY>{});
  return dispatch_hk(hk, [&](auto h) { return sk == SK_SPLIT ? f(h, IC<SK_SPLIT>{}) : f(h, IC<SK_RANGE>{}); });
}

// keys per thread of a scatter round (round = 1024 * 32 keys, the engine's
// tiling unit)
constexpr int SCATTER_KPT = 32;
// The histogram walks a tile in rounds of 1024 * 16 keys: fewer registers,
// so two workgroups share a CU.
constexpr int HIST_KPT = 16;

// Leaves of a learned model the histogram stages in LDS (64 KB; larger
// models are read from global memory).
constexpr uint32_t HIST_LEAF_CAP = 4096;

template <typename K>
static hipError_t launch_part_hist_t(const LaunchPart &lp, hipStream_t st) {
  const uint32_t leaf_cap = lp.hk == HK_RMI ? HIST_LEAF_CAP : 0u;
  const size_t lds = (size_t)((lp.max_cols + 3) & ~3u) * 4 + (size_t)leaf_cap * sizeof(RmiLeaf);
  return dispatch_part(lp.hk, lp.sk, [&](auto h, auto b) {
    hipLaunchKernelGGL((k_part_hist<K, 1024, HIST_KPT, decltype(h)::value, decltype(b)::value>), dim3(lp.ntiles),
                       dim3(1024), lds, st, (const K *)lp.keys, lp.tiles, lp.segs, lp.cmat, leaf_cap);
    return hipGetLastError();
  });
}

template <typename K>
static hipError_t launch_part_scatter_t(const LaunchPart &lp, hipStream_t st) {
  const size_t cnt_bytes = (size_t)((lp.max_cols + 3) & ~3u) * 4;
  const size_t lds = 2 * cnt_bytes + (size_t)lp.stage_cap * (sizeof(K) + 2);
  return dispatch_part(lp.hk, lp.sk, [&](auto h, auto b) {
    constexpr int H = decltype(h)::value, S = decltype(b)::value;
    // the packing instance exists for murmur3 levels of 8-byte records only
    // (launch_part_scatter checks)
    auto kern = &k_part_scatter<K, 1024, SCATTER_KPT, H, S>;
    if constexpr (sizeof(K) == 8 && H == HK_MURMUR) {
      if (lp.pack) kern = &k_part_scatter<K, 1024, SCATTER_KPT, H, S, true>;
    }
    hipLaunchKernelGGL(kern, dim3(lp.ntiles), dim3(1024), lds, st, (const K *)lp.keys, lp.tiles, lp.segs, lp.cmat,
                       (K *)lp.out, lp.stage_cap);
    return hipGetLastError();
  });
}

// Keys staged per LDS sub-round of the scatter: what fits beside the round
// counters, the positions, the static scan scratch and (split kinds) the
// splits in 160 KB of LDS, in multiples of 1024 (0 when fewer than 1024 fit).
uint32_t part_stage_cap(uint32_t max_cols, int key_bytes, int sk) {
  const size_t used = 2 * (size_t)((max_cols + 3) & ~3u) * 4 + 256 + (sk != SK_RANGE ? (MAX_SPLITS + 1) * 8 : 0);
  const size_t avail = used < 160 * 1024 ? 160 * 1024 - used : 0;
  size_t cap = avail / (size_t)(key_bytes + 2) / 1024 * 1024;
  cap = cap > 1024u * SCATTER_KPT ? 1024u * SCATTER_KPT : cap;
  return (uint32_t)cap;
}

hipError_t launch_part_hist(const LaunchPart &lp, hipStream_t st) {
  // packing levels count the 32-bit keys themselves (their records fall in
  // the same buckets: murmur3 bins depend on the low 32 bits only)
  return lp.key_bytes == 4 || lp.pack ? launch_part_hist_t<uint32_t>(lp, st) : launch_part_hist_t<uint64_t>(lp, st);
}
hipError_t launch_part_scatter(const LaunchPart &lp, hipStream_t st) {
  if (lp.pack && (lp.key_bytes != 8 || lp.hk != HK_MURMUR)) return hipErrorInvalidValue;
  return lp.key_bytes == 4 ? launch_part_scatter_t<uint32_t>(lp, st) : launch_part_scatter_t<uint64_t>(lp, st);
}

hipError_t launch_colscan(const LaunchCol &lc, hipStream_t st) {
  if (lc.nitems)
    hipLaunchKernelGGL(k_col_chunksum, dim3(lc.nitems), dim3(256), 0, st, lc.cmat, lc.items, lc.segs, lc.psum_base, lc.psum);
  if (lc.nitems2)
    hipLaunchKernelGGL(k_col_chunkscan, dim3(lc.nitems2), dim3(256), 0, st, lc.items2, lc.segs, lc.psum_base, lc.nchunks,
                       lc.psum, lc.coltot);
  if (lc.nsegs < 64)  // few wide segments: more threads per segment
    hipLaunchKernelGGL((k_seg_bucketscan<1024>), dim3(lc.nsegs), dim3(1024), 0, st, lc.segs, lc.coltot, lc.bstart);
  else
    hipLaunchKernelGGL((k_seg_bucketscan<256>), dim3(lc.nsegs), dim3(256), 0, st, lc.segs, lc.coltot, lc.bstart);
  if (lc.nitems)
    hipLaunchKernelGGL(k_col_apply, dim3(lc.nitems), dim3(256), 0, st, lc.cmat, lc.items, lc.segs, lc.psum_base, lc.psum,
                       lc.bstart);
  return hipGetLastError();
}

template <typename K, typename OFF, bool SPLIT>
static hipError_t launch_final_t(const LaunchFinal &lf, hipStream_t st) {
  if (lf.nitems) hipLaunchKernelGGL(k_fill<uint32_t>, dim3(1), dim3(64), 0, st, &((BigCtl *)lf.big_ctl)->seen, 0u, (uint64_t)1);
  if (lf.nitems) {
    constexpr int KPT = sizeof(K) == 4 ? KPT_U32 : 32;
    const hipError_t e = dispatch_hk(lf.hk, [&](auto h) {
      constexpr int H = decltype(h)::value;
      // the split store exists for murmur3 tables only (launch_final checks)
      auto kern = &k_final<K, OFF, 1024, KPT, H>;
      if constexpr (SPLIT && H == HK_MURMUR) kern = &k_final<K, OFF, 1024, KPT, H, true>;
      hipLaunchKernelGGL(kern, dim3(lf.nitems), dim3(1024), lf.lds_bytes, st, (const K *)lf.part, (K *)lf.csr,
                         (OFF *)lf.off, lf.items, lf.maps, lf.lds_bytes, &((BigCtl *)lf.big_ctl)->seen, lf.csr_index);
      return hipGetLastError();
    });
    if (e != hipSuccess) return e;
  }
  if (lf.nitems) {
    constexpr int KPT = sizeof(K) == 4 ? KPT_U32 : 32;
    const uint32_t small_cap = std::min(1024u * KPT, 65535u);
    BigCtl *ctl = (BigCtl *)lf.big_ctl;
    hipLaunchKernelGGL(k_big_list, dim3(1), dim3(1024), 0, st, lf.items, lf.nitems, small_cap, lf.big_cap, BIG_CHUNK,
                       (const unsigned int *)&ctl->seen, lf.big_slot, ctl);
    hipLaunchKernelGGL(k_big_zero, dim3(256), dim3(256), 0, st, (const BigCtl *)ctl, lf.big_cnt);
    hipLaunchKernelGGL((k_big_count<K, 1024>), dim3(2048), dim3(1024), 0, st, (const K *)lf.part, lf.items, lf.maps,
                       (const BigCtl *)ctl, lf.big_cnt);
    hipLaunchKernelGGL((k_big_scan<OFF>), dim3(256), dim3(1024), 0, st, lf.items, (const BigCtl *)ctl, lf.big_cnt,
                       (OFF *)lf.off);
    hipLaunchKernelGGL((k_big_place<K, 1024, SPLIT>), dim3(2048), dim3(1024), 0, st, (const K *)lf.part, (K *)lf.csr,
                       lf.items, lf.maps, (const BigCtl *)ctl, lf.big_cnt, lf.csr_index);
    hipLaunchKernelGGL((k_final_big<K, OFF, 1024, 16384, SPLIT>), dim3(lf.nitems), dim3(1024), 0, st, (const K *)lf.part,
                       (K *)lf.csr, (OFF *)lf.off, lf.items, lf.maps, small_cap, (const uint32_t *)lf.big_slot,
                       (const unsigned int *)&ctl->seen, lf.csr_index);
  }
  return hipGetLastError();
}

hipError_t launch_final(const LaunchFinal &lf, hipStream_t st) {
  if (lf.csr_index) {  // index-carrying build of 32-bit keys: 8-byte murmur3 records split on store
    if (lf.key_bytes != 8 || lf.hk != HK_MURMUR) return hipErrorInvalidValue;
    return lf.off_bytes == 4 ? launch_final_t<uint64_t, uint32_t, true>(lf, st)
                             : launch_final_t<uint64_t, uint64_t, true>(lf, st);
  }
  if (lf.key_bytes == 4)
    return lf.off_bytes == 4 ? launch_final_t<uint32_t, uint32_t, false>(lf, st)
                             : launch_final_t<uint32_t, uint64_t, false>(lf, st);
  return lf.off_bytes == 4 ? launch_final_t<uint64_t, uint32_t, false>(lf, st)
                           : launch_final_t<uint64_t, uint64_t, false>(lf, st);
}

int final_kpt(int key_bytes) { return key_bytes == 4 ? KPT_U32 : 32; }
uint32_t final_big_items() { return BIG_CAP; }
size_t final_big_ctl_bytes() { return sizeof(BigCtl); }

// One-pass probe's bucket limits: keys of B per workgroup, keys of A in LDS.
template <typename K> constexpr int probe1_kpt() { return sizeof(K) == 4 ? 32 : 16; }
constexpr uint32_t PROBE1_LDS = 156 * 1024;  // keys of A + 16-bit offsets

template <typename K, typename OFF>
static hipError_t launch_probe_t(const LaunchProbe &p, hipStream_t st) {
  constexpr int KPT = sizeof(K) == 4 ? KPT_U32 : 32;
  constexpr int KPT1 = probe1_kpt<K>();
  constexpr int SUB = probe_sub<OFF>();
  if (p.akey_cap < PROBE_BIN_LIMIT) return hipErrorInvalidValue;  // a walkable bin must fit in LDS
  const size_t lds = ((SUB + 2) * sizeof(OFF) + 15) / 16 * 16 + (size_t)p.akey_cap * sizeof(K);
  const hipError_t e = dispatch_hk(p.hk, [&](auto h) {
    hipLaunchKernelGGL((k_probe1<K, OFF, 1024, KPT1, decltype(h)::value>), dim3(p.nitems), dim3(1024),
                       (size_t)PROBE1_LDS, st, (const K *)p.bpart, p.items, p.maps, (const OFF *)p.offA,
                       (const K *)p.keysA, (K *)p.out, p.capacity, p.cursor, PROBE1_LDS, p.hmark, p.gate, p.out_off);
    hipLaunchKernelGGL((k_probe<K, OFF, 1024, KPT, SUB, decltype(h)::value>), dim3(p.nitems), dim3(1024), lds, st,
                       (const K *)p.bpart, p.items, p.maps, (const OFF *)p.offA, (const K *)p.keysA, (K *)p.out,
                       p.capacity, p.cursor, p.akey_cap, p.hmark, (uint32_t)(1024 * KPT1), PROBE1_LDS, p.gate,
                       p.out_off);
    return hipGetLastError();
  });
  if (e != hipSuccess) return e;
  const uint32_t small_cap = std::min(1024u * KPT, 65535u);
  // oversized buckets: up to BIG_CAP of them listed and cut into tasks of
  // PROBE_BIG_CHUNK keys for the whole grid, the rest one workgroup each
  BigCtl *ctl = (BigCtl *)p.big_ctl;
  hipLaunchKernelGGL(k_big_list, dim3(1), dim3(1024), 0, st, p.items, p.nitems, small_cap, 0xFFFFFFFFu, PROBE_BIG_CHUNK,
                     (const unsigned int *)p.gate, p.big_slot, ctl);
  hipLaunchKernelGGL((k_probe_big<K, OFF, 1024>), dim3(p.nitems), dim3(1024), 0, st, (const K *)p.bpart, p.items, p.maps,
                     (const OFF *)p.offA, (const K *)p.keysA, (K *)p.out, p.capacity, p.cursor, small_cap,
                     (const uint32_t *)p.big_slot, p.hmark, p.gate, p.out_off);
  hipLaunchKernelGGL((k_probe_big_chunks<K, OFF, 1024>), dim3(PROBE_BIG_GRID), dim3(1024), 0, st, (const K *)p.bpart,
                     p.items, p.maps, (const OFF *)p.offA, (const K *)p.keysA, (K *)p.out, p.capacity, p.cursor,
                     (const BigCtl *)ctl, p.hmark, p.gate, p.out_off);
  hipLaunchKernelGGL((k_probe_heavy<K, OFF, 1024>), dim3(p.nitems), dim3(1024), 0, st, (const K *)p.bpart, p.items,
                     (const OFF *)p.offA, (const K *)p.keysA, (K *)p.out, p.capacity, p.cursor,
                     (const unsigned int *)p.hmark, (const unsigned int *)p.gate, p.out_off);
  return hipGetLastError();
}

hipError_t launch_probe(const LaunchProbe &p, hipStream_t st) {
  if (!p.nitems) return hipSuccess;
  if (p.key_bytes == 4)
    return p.off_bytes == 4 ? launch_probe_t<uint32_t, uint32_t>(p, st) : launch_probe_t<uint32_t, uint64_t>(p, st);
  return p.off_bytes == 4 ? launch_probe_t<uint64_t, uint32_t>(p, st) : launch_probe_t<uint64_t, uint64_t>(p, st);
}

uint32_t probe_akey_cap(int key_bytes, int off_bytes) {
  const size_t sub = off_bytes == 4 ? probe_sub<uint32_t>() : probe_sub<uint64_t>();
  const size_t used = ((sub + 2) * (size_t)off_bytes + 15) / 16 * 16;
  return (uint32_t)((PROBE_LDS - used) / key_bytes);
}

// Row-id join of the bins above JOIN_PAIR_LIMIT key pairs: every pair of
// entries is compared, in tiles of PAIR_TILE x PAIR_TILE entries spread over
// the grid.  k_pair_tiles gives every listed bin its tile count and adds its
// key pairs (each capped at 2^37, so the sum of up to 2^26 bins cannot wrap)
// to *pairs.
constexpr uint32_t PAIR_TILE = 1024;
template <typename OFF>
__global__ void k_pair_tiles(const OFF *__restrict__ offA, const OFF *__restrict__ offB,
                             const uint32_t *__restrict__ hlist, uint32_t n, unsigned long long *__restrict__ tiles,
                             unsigned long long *__restrict__ pairs) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t b = hlist[i];
    const uint64_t a = (uint64_t)offA[b + 1] - (uint64_t)offA[b], x = (uint64_t)offB[b + 1] - (uint64_t)offB[b];
    tiles[i] = ((a + PAIR_TILE - 1) / PAIR_TILE) * ((x + PAIR_TILE - 1) / PAIR_TILE);
    const unsigned long long ax = (a >> 31) || (x >> 31) ? (1ull << 37) : (unsigned long long)(a * x);
    atomicAdd(pairs, ax < (1ull << 37) ? ax : (1ull << 37));
  }
}

// One workgroup per tile (grid-stride): the tile's entries of A and of B in
// LDS, its matches counted, their places reserved with one atomic, then
// written as (A's input position, B's input position).
template <typename K, typename OFF, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_join_pairs_heavy(
    const OFF *__restrict__ offA, const K *__restrict__ keysA, const uint64_t *__restrict__ indA,
    const OFF *__restrict__ offB, const K *__restrict__ keysB, const uint64_t *__restrict__ indB,
    const uint32_t *__restrict__ hlist, uint32_t nh, const unsigned long long *__restrict__ tile_base, uint64_t ntiles,
    uint64_t *__restrict__ pa, uint64_t *__restrict__ pb, uint64_t capacity, unsigned long long *__restrict__ cursor) {
  __shared__ K s_a[PAIR_TILE], s_b[PAIR_TILE];
  __shared__ unsigned long long s_t64[BLOCK / 64 + 1];
  __shared__ unsigned long long s_base;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    // the listed bin of tile t: tile_base[h] <= t < tile_base[h + 1]
    uint32_t lo = 0, hi = nh;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (tile_base[mid] <= t) lo = mid;
      else hi = mid;
    }
    const uint32_t bin = hlist[lo];
    const uint64_t a0 = offA[bin], a1 = offA[bin + 1], x0 = offB[bin], x1 = offB[bin + 1];
    const uint64_t nbt = (x1 - x0 + PAIR_TILE - 1) / PAIR_TILE;
    const uint64_t lt = t - tile_base[lo];
    const uint64_t ia0 = a0 + (lt / nbt) * PAIR_TILE, ix0 = x0 + (lt % nbt) * PAIR_TILE;
    const uint32_t na = (uint32_t)min((uint64_t)PAIR_TILE, a1 - ia0), nb = (uint32_t)min((uint64_t)PAIR_TILE, x1 - ix0);
    for (uint32_t q = threadIdx.x; q < na; q += BLOCK) s_a[q] = keysA[ia0 + q];
    for (uint32_t q = threadIdx.x; q < nb; q += BLOCK) s_b[q] = keysB[ix0 + q];
    __syncthreads();
    unsigned long long c = 0;
    for (uint32_t i = threadIdx.x; i < na; i += BLOCK) {
      const K a = s_a[i];
      for (uint32_t j = 0; j < nb; j++) c += s_b[j] == a;
    }
    unsigned long long tot;
    const unsigned long long ex = block_exclusive_sum<BLOCK, unsigned long long>(c, s_t64, tot);
    if (threadIdx.x == 0) s_base = tot ? atomicAdd(cursor, tot) : 0ull;
    __syncthreads();
    unsigned long long w = s_base + ex;
    if (c)
      for (uint32_t i = threadIdx.x; i < na; i += BLOCK) {
        const K a = s_a[i];
        for (uint32_t j = 0; j < nb; j++)
          if (s_b[j] == a) {
            if (w < capacity) {
              pa[w] = indA[ia0 + i];
              pb[w] = indB[ix0 + j];
            }
            w++;
          }
      }
    __syncthreads();  // the next tile overwrites s_a, s_b
  }
}

hipError_t launch_pair_tiles(const void *offA, const void *offB, int off_bytes, const uint32_t *hlist, uint32_t n,
                             unsigned long long *tiles, unsigned long long *pairs, hipStream_t st) {
  if (!n) return hipSuccess;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  if (off_bytes == 4)
    hipLaunchKernelGGL(k_pair_tiles<uint32_t>, dim3(grid), dim3(256), 0, st, (const uint32_t *)offA,
                       (const uint32_t *)offB, hlist, n, tiles, pairs);
  else
    hipLaunchKernelGGL(k_pair_tiles<uint64_t>, dim3(grid), dim3(256), 0, st, (const uint64_t *)offA,
                       (const uint64_t *)offB, hlist, n, tiles, pairs);
  return hipGetLastError();
}

template <typename K, typename OFF>
static hipError_t launch_join_pairs_heavy_t(const LaunchJoinPairs &p, hipStream_t st) {
  const uint32_t grid = (uint32_t)std::min<uint64_t>(p.ntiles, 8192);
  hipLaunchKernelGGL((k_join_pairs_heavy<K, OFF, 256>), dim3(grid), dim3(256), 0, st, (const OFF *)p.offA,
                     (const K *)p.keysA, p.indA, (const OFF *)p.offB, (const K *)p.keysB, p.indB, p.hlist, p.nh,
                     p.tile_base, p.ntiles, p.pa, p.pb, p.capacity, p.cursor);
  return hipGetLastError();
}

hipError_t launch_join_pairs_heavy(const LaunchJoinPairs &p, hipStream_t st) {
  if (!p.ntiles) return hipSuccess;
  if (p.key_bytes == 4)
    return p.off_bytes == 4 ? launch_join_pairs_heavy_t<uint32_t, uint32_t>(p, st)
                            : launch_join_pairs_heavy_t<uint32_t, uint64_t>(p, st);
  return p.off_bytes == 4 ? launch_join_pairs_heavy_t<uint64_t, uint32_t>(p, st)
                          : launch_join_pairs_heavy_t<uint64_t, uint64_t>(p, st);
}

template <typename K, typename OFF>
static hipError_t launch_join_t(const LaunchJoin &j, hipStream_t st) {
  constexpr int JB = JOIN_BINS;
  const uint64_t nblk = (j.M + JB - 1) / JB;
  const size_t lds = 2 * (size_t)j.key_cap * (sizeof(K) + (j.pa ? 8 : 0)) +
                     (j.pa ? JOIN_EMIT_PAIRS * 4 : j.out ? JOIN_EMIT_KEYS * 2 : 0);
  auto kern = j.pa ? &k_join<K, OFF, 1024, JB, JOIN_PAIRS>
                  : j.out ? &k_join<K, OFF, 1024, JB, JOIN_KEYS> : &k_join<K, OFF, 1024, JB, JOIN_COUNT>;
  hipLaunchKernelGGL(kern, dim3((uint32_t)nblk), dim3(1024), lds, st, (const OFF *)j.offA, (const K *)j.keysA,
                     (const OFF *)j.offB, (const K *)j.keysB, j.M, (K *)j.out, j.capacity, j.cursor, j.key_cap, j.hlist,
                     j.hlist_n, j.hlist_cap, j.indA, j.indB, j.pa, j.pb);
  if (j.pa) return hipGetLastError();  // the row-id join answers its listed bins itself
  // the heavy bins (a grid that returns at once when k_join listed none)
  const uint32_t gh = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(j.hlist_cap, 1), 1024);
  hipLaunchKernelGGL((k_join_heavy<K, OFF, 1024>), dim3(gh), dim3(1024), 0, st, (const OFF *)j.offA, (const K *)j.keysA,
                     (const OFF *)j.offB, (const K *)j.keysB, (K *)j.out, j.capacity, j.cursor,
                     (const uint32_t *)j.hlist, (const unsigned int *)j.hlist_n, j.hlist_cap);
  return hipGetLastError();
}

hipError_t launch_join(const LaunchJoin &j, hipStream_t st) {
  if (j.M == 0) return hipSuccess;
  if (j.key_bytes == 4)
    return j.off_bytes == 4 ? launch_join_t<uint32_t, uint32_t>(j, st) : launch_join_t<uint32_t, uint64_t>(j, st);
  return j.off_bytes == 4 ? launch_join_t<uint64_t, uint32_t>(j, st) : launch_join_t<uint64_t, uint64_t>(j, st);
}

uint32_t join_key_cap(int key_bytes, int off_bytes, bool emit, bool with_index) {
  // two key arrays (for the row-id join with their 8-byte positions; for an
  // emitting join beside the match buffer) next to the 2*(JB+1) offsets held
  // in static LDS; the cap stays below 2^16 (buffered slots are 16-bit)
  const size_t budget = 78 * 1024 - 2 * (JOIN_BINS + 1) * (size_t)off_bytes -
                        (with_index ? JOIN_EMIT_PAIRS * 4 : emit ? JOIN_EMIT_KEYS * 2 : 0);
  return (uint32_t)(budget / (2 * ((size_t)key_bytes + (with_index ? 8 : 0))));
}

// Status area of a scan over n elements: flags (u32 per tile), aggregates
// and inclusive prefixes (u64 per tile each), the ticket.
struct ScanStatus {
  unsigned int *flags;
  unsigned long long *agg, *inc;
  unsigned int *ticket;
};
static ScanStatus scan_carve(void *status, uint64_t ntiles) {
  char *p = (char *)status;
  ScanStatus s;
  const size_t fb = (ntiles * 4 + 255) & ~(size_t)255;
  s.flags = (unsigned int *)p;
  s.agg = (unsigned long long *)(p + fb);
  s.inc = s.agg + ntiles;
  s.ticket = (unsigned int *)(p + fb + ntiles * 16);
  return s;
}
size_t scan_status_bytes(uint64_t n) {
  const uint64_t nt = scan_tiles(n);
  return ((nt * 4 + 255) & ~(size_t)255) + nt * 16 + 256;
}

hipError_t launch_scan(const void *in, void *out, uint64_t n, int elem_bytes, void *status, hipStream_t st) {
  constexpr int B = 256, IPT = 16;
  const uint64_t ntiles = scan_tiles(n);
  if (!ntiles) return hipSuccess;
  const ScanStatus ss = scan_carve(status, ntiles);
  hipLaunchKernelGGL((k_fill<unsigned int>), dim3(256), dim3(256), 0, st, ss.flags, 0u, ntiles);
  hipLaunchKernelGGL((k_fill<unsigned int>), dim3(1), dim3(64), 0, st, ss.ticket, 0u, (uint64_t)1);
  if (elem_bytes == 4)
    hipLaunchKernelGGL((k_scan_lookback<uint32_t, B, IPT>), dim3((uint32_t)ntiles), dim3(B), 0, st, (const uint32_t *)in,
                       (uint32_t *)out, n, ss.flags, ss.agg, ss.inc, ss.ticket);
  else
    hipLaunchKernelGGL((k_scan_lookback<uint64_t, B, IPT>), dim3((uint32_t)ntiles), dim3(B), 0, st, (const uint64_t *)in,
                       (uint64_t *)out, n, ss.flags, ss.agg, ss.inc, ss.ticket);
  return hipGetLastError();
}
uint64_t scan_tiles(uint64_t n) { return (n + 256 * 16 - 1) / (256 * 16); }

// Bounds of every chunk of the run-length join from the offsets alone: its
// records (values present on both sides) number at most min(A keys, B keys),
// its heavy bins (more than RLE_SMALL keys on a side) at most
// A keys / (RLE_SMALL + 1) + B keys / (RLE_SMALL + 1).
template <typename OFF>
__global__ void k_rle_caps(const OFF *__restrict__ offA, const OFF *__restrict__ offB, uint64_t M, uint32_t nch,
                           unsigned long long *__restrict__ cap_rec, uint32_t *__restrict__ cap_h) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nch; j += gridDim.x * blockDim.x) {
    const uint64_t lo = (uint64_t)j * RLE_CHUNK, hi = min(lo + (uint64_t)RLE_CHUNK, M);
    const uint64_t na = (uint64_t)offA[hi] - (uint64_t)offA[lo], nb = (uint64_t)offB[hi] - (uint64_t)offB[lo];
    cap_rec[j] = min(na, nb);
    cap_h[j] = (uint32_t)min(hi - lo, na / (RLE_SMALL + 1) + nb / (RLE_SMALL + 1));
  }
}

// One workgroup per chunk (grid-stride): its records and heavy bins from the
// padded places to the compact ones; records at or beyond capacity are
// dropped, as when written directly.
template <typename K>
__global__ __launch_bounds__(256) void k_rle_gather(LaunchRleGather g) {
  const K *srk = (const K *)g.srk;
  K *rk = (K *)g.rk;
  for (uint32_t j = blockIdx.x; j < g.nch; j += gridDim.x) {
    const uint64_t s = g.pbase[j], n = g.cnt[j], d = g.cbase[j];
    const uint32_t hs = g.phbase[j], hn = g.hcnt[j], hd = g.hbase[j];
    if (n > g.pbase[j + 1] - s || hn > g.phbase[j + 1] - hs) {  // uniform
      if (threadIdx.x == 0) atomicOr(g.bad, 1u);
      continue;
    }
    for (uint64_t i = threadIdx.x; i < n; i += 256)
      if (d + i < g.capacity) {
        rk[d + i] = srk[s + i];
        g.rca[d + i] = g.srca[s + i];
        g.rcb[d + i] = g.srcb[s + i];
      }
    for (uint32_t i = threadIdx.x; i < hn; i += 256)
      if ((uint64_t)hd + i < g.heavy_cap) g.heavy[hd + i] = g.sheavy[hs + i];
  }
}

hipError_t launch_rle_caps(const void *offA, const void *offB, int off_bytes, uint64_t M,
                           unsigned long long *cap_rec, uint32_t *cap_h, hipStream_t st) {
  const uint32_t nch = (uint32_t)((M + RLE_CHUNK - 1) / RLE_CHUNK);
  if (!nch) return hipSuccess;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((nch + 255) / 256, 4096);
  if (off_bytes == 4)
    hipLaunchKernelGGL(k_rle_caps<uint32_t>, dim3(grid), dim3(256), 0, st, (const uint32_t *)offA,
                       (const uint32_t *)offB, M, nch, cap_rec, cap_h);
  else
    hipLaunchKernelGGL(k_rle_caps<uint64_t>, dim3(grid), dim3(256), 0, st, (const uint64_t *)offA,
                       (const uint64_t *)offB, M, nch, cap_rec, cap_h);
  return hipGetLastError();
}

hipError_t launch_rle_gather(const LaunchRleGather &g, hipStream_t st) {
  if (!g.nch) return hipSuccess;
  const uint32_t grid = std::min<uint32_t>(g.nch, 8192);
  if (g.key_bytes == 4)
    hipLaunchKernelGGL(k_rle_gather<uint32_t>, dim3(grid), dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL(k_rle_gather<uint64_t>, dim3(grid), dim3(256), 0, st, g);
  return hipGetLastError();
}

template <typename K, typename OFF>
static hipError_t launch_rle_t(const LaunchRle &r, hipStream_t st) {
  const uint32_t nchunks = (uint32_t)((r.M + RLE_CHUNK - 1) / RLE_CHUNK);
  if ((r.pass <= 1 || r.pass == 3) && nchunks) {
    hipLaunchKernelGGL((k_rle_small<K, OFF, 256>), dim3(nchunks), dim3(256), 0, st, (const OFF *)r.offA,
                       (const K *)r.keysA, (const OFF *)r.offB, (const K *)r.keysB, r.M, r.pass, r.chunk_cnt,
                       r.chunk_base, r.chunk_heavy, r.heavy_base, r.chunk_pairs, (K *)r.rk, r.rca, r.rcb, r.capacity,
                       r.heavy, r.heavy_cap);
    if (r.pass == 0 || r.pass == 3)
      hipLaunchKernelGGL(k_sum_add, dim3(1), dim3(1024), 0, st, (const unsigned long long *)r.chunk_pairs,
                         (uint64_t)nchunks, r.pairs);
  }
  if (r.pass == 2 && r.nheavy) {
    GiantCtl *ctl = (GiantCtl *)r.gctl;
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(1), dim3(64), 0, st, (uint32_t *)&ctl->n, 0u, (uint64_t)8);
    hipLaunchKernelGGL((k_giant_mark<OFF>), dim3((uint32_t)std::min<uint64_t>((r.nheavy + 255) / 256, 4096)), dim3(256), 0,
                       st, (const OFF *)r.offA, (const OFF *)r.offB, (const uint32_t *)r.heavy, r.nheavy, r.gmark,
                       r.heavy_cap, ctl);
    hipLaunchKernelGGL((k_giant_prefix<OFF>), dim3(1), dim3(64), 0, st, (const OFF *)r.offA, (const OFF *)r.offB,
                       (const uint32_t *)r.heavy, ctl);
    hipLaunchKernelGGL((k_giant_clear<K>), dim3(1024), dim3(256), 0, st, (const GiantCtl *)ctl, (K *)r.gkey, r.gca, r.gcb);
    hipLaunchKernelGGL((k_giant_fold<K, OFF, 0>), dim3(1024), dim3(1024), 0, st, (const OFF *)r.offA, (const K *)r.keysA,
                       (const uint32_t *)r.heavy, ctl, (K *)r.gkey, r.gca);
    hipLaunchKernelGGL((k_giant_fold<K, OFF, 1>), dim3(1024), dim3(1024), 0, st, (const OFF *)r.offB, (const K *)r.keysB,
                       (const uint32_t *)r.heavy, ctl, (K *)r.gkey, r.gcb);
    hipLaunchKernelGGL((k_giant_emit<K>), dim3(GIANT_CAP), dim3(1024), 0, st, (const GiantCtl *)ctl, (const K *)r.gkey,
                       (const unsigned long long *)r.gca, (const unsigned long long *)r.gcb, (K *)r.rk, r.rca, r.rcb,
                       r.capacity, r.cursor, r.pairs, (const uint32_t *)r.heavy, r.ovf_list, r.ovf_n, r.heavy_cap);
    const uint32_t gs = (uint32_t)std::min<uint64_t>(r.nheavy, RLE_HEAVY_GRID_S);
    const uint32_t gb = (uint32_t)std::min<uint64_t>(r.nheavy, RLE_HEAVY_GRID);
    hipLaunchKernelGGL((k_rle_heavy<K, OFF, 256, RLE_HEAVY_SMALL_TS>), dim3(gs), dim3(256), 0, st, (const OFF *)r.offA,
                       (const K *)r.keysA, (const OFF *)r.offB, (const K *)r.keysB, (K *)r.rk, r.rca,
                       r.rcb, r.capacity, r.cursor, r.pairs, r.overflow, (const uint32_t *)r.gmark,
                       (const uint32_t *)&ctl->nS, (K *)r.pkey, r.pca, r.pcb, r.priv_n, 0u, r.ovf_list, r.ovf_n,
                       r.heavy_cap);
    hipLaunchKernelGGL((k_rle_heavy<K, OFF, 1024, HT_SLOTS>), dim3(gb), dim3(1024), 0, st, (const OFF *)r.offA,
                       (const K *)r.keysA, (const OFF *)r.offB, (const K *)r.keysB, (K *)r.rk, r.rca,
                       r.rcb, r.capacity, r.cursor, r.pairs, r.overflow, (const uint32_t *)(r.gmark + r.heavy_cap),
                       (const uint32_t *)&ctl->nL, (K *)r.pkey, r.pca, r.pcb, r.priv_n, gs, r.ovf_list, r.ovf_n,
                       r.heavy_cap);
    const uint32_t g = gs + gb;
    hipLaunchKernelGGL((k_rle_compact<K>), dim3(g), dim3(256), 0, st, (const uint32_t *)r.priv_n, (const K *)r.pkey,
                       (const unsigned long long *)r.pca, (const unsigned long long *)r.pcb, (K *)r.rk, r.rca, r.rcb,
                       r.capacity, (const unsigned long long *)r.cursor);
    hipLaunchKernelGGL(k_rle_finish, dim3(1), dim3(1024), 0, st, (const uint32_t *)r.priv_n, g, r.cursor);
  }
  return hipGetLastError();
}

template <typename K>
static hipError_t launch_rle_ovf_t(const LaunchRleOvf &o, hipStream_t st) {
  if (o.nchunksA) {
    const uint32_t g = (uint32_t)std::min<uint64_t>(o.nchunksA, 2048);
    hipLaunchKernelGGL((k_ovf_fold<K, 0>), dim3(g), dim3(1024), 0, st, (const K *)o.keysA, o.chunksA, o.nchunksA, o.tabs,
                       (K *)o.tkey, o.tca, o.eA);
  }
  if (o.nchunksB) {
    const uint32_t g = (uint32_t)std::min<uint64_t>(o.nchunksB, 2048);
    hipLaunchKernelGGL((k_ovf_fold<K, 1>), dim3(g), dim3(1024), 0, st, (const K *)o.keysB, o.chunksB, o.nchunksB, o.tabs,
                       (K *)o.tkey, o.tcb, o.eB);
  }
  const uint32_t ge = (uint32_t)std::min<uint64_t>((o.nslots + o.ntabs + 1023) / 1024, 4096);
  hipLaunchKernelGGL((k_ovf_emit<K>), dim3(ge), dim3(1024), 0, st, (const K *)o.tkey, (const uint32_t *)o.tca,
                     (const uint32_t *)o.tcb, o.nslots, (const unsigned long long *)o.eA, (const unsigned long long *)o.eB,
                     o.ntabs, (K *)o.rk, o.rca, o.rcb, o.capacity, o.cursor, o.pairs);
  return hipGetLastError();
}

hipError_t launch_rle_ovf(const LaunchRleOvf &o, hipStream_t st) {
  return o.key_bytes == 4 ? launch_rle_ovf_t<uint32_t>(o, st) : launch_rle_ovf_t<uint64_t>(o, st);
}

hipError_t launch_ovf_sizes(const uint32_t *list, uint32_t n, const void *offA, const void *offB, int off_bytes,
                            unsigned long long *out, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_ovf_sizes, dim3((n + 255) / 256), dim3(256), 0, st, list, n, offA, offB, off_bytes, out);
  return hipGetLastError();
}

hipError_t launch_rle(const LaunchRle &r, hipStream_t st) {
  if (r.key_bytes == 4)
    return r.off_bytes == 4 ? launch_rle_t<uint32_t, uint32_t>(r, st) : launch_rle_t<uint32_t, uint64_t>(r, st);
  return r.off_bytes == 4 ? launch_rle_t<uint64_t, uint32_t>(r, st) : launch_rle_t<uint64_t, uint64_t>(r, st);
}
size_t rle_giant_slots() { return (size_t)GIANT_CAP * GT_SLOTS; }
size_t rle_priv_records() { return (size_t)(RLE_HEAVY_GRID_S + RLE_HEAVY_GRID) * RLE_PRIV_CAP; }
uint32_t rle_heavy_grid() { return RLE_HEAVY_GRID_S + RLE_HEAVY_GRID; }
size_t rle_giant_ctl_bytes() { return sizeof(GiantCtl); }
size_t rle_giant_bytes(int key_bytes) { return sizeof(GiantCtl) + rle_giant_slots() * (key_bytes + 16) + 1024; }

uint64_t rle_chunks(uint64_t M) { return (M + RLE_CHUNK - 1) / RLE_CHUNK; }

hipError_t launch_make_items(const SegDesc *segs, uint32_t nsegs, const uint32_t *bstart, const uint32_t *coltot,
                             FinItem *items, uint32_t ncols_total, hipStream_t st) {
  if (!ncols_total) return hipSuccess;
  hipLaunchKernelGGL(k_make_items, dim3((ncols_total + 255) / 256), dim3(256), 0, st, segs, nsegs, bstart, coltot, items,
                     ncols_total);
  return hipGetLastError();
}

// Raise a kernel's dynamic-LDS limit to what the device allows beside the
// kernel's static LDS.
template <typename F>
static hipError_t set_lds(F *fn) {
  hipFuncAttributes fa;
  hipError_t e = hipFuncGetAttributes(&fa, (const void *)fn);
  if (e != hipSuccess) return e;
  int dev = 0, lim = 0;
  (void)hipGetDevice(&dev);
  e = hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e != hipSuccess) return e;
  const int room = lim - (int)fa.sharedSizeBytes;
  if (room <= 0) return hipSuccess;
  return hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, room);
}


template <int MODE>
static hipError_t join_set_lds() {
  hipError_t r = set_lds(&k_join<uint32_t, uint32_t, 1024, JOIN_BINS, MODE>);
  if (r == hipSuccess) r = set_lds(&k_join<uint32_t, uint64_t, 1024, JOIN_BINS, MODE>);
  if (r == hipSuccess) r = set_lds(&k_join<uint64_t, uint32_t, 1024, JOIN_BINS, MODE>);
  if (r == hipSuccess) r = set_lds(&k_join<uint64_t, uint64_t, 1024, JOIN_BINS, MODE>);
  return r;
}

hipError_t set_kernel_attributes() {
  hipError_t e = hipSuccess;
#define HG_SET(...) do { hipError_t e_ = set_lds(__VA_ARGS__); if (e_ != hipSuccess) e = e_; } while (0)
#define HG_SET_D(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) e = e_; } while (0)
  for (int hk = 0; hk < 3; hk++)
    for (int sk = 0; sk < 3; sk++) {
      HG_SET_D(dispatch_part(hk, sk, [&](auto h, auto b) {
        constexpr int H = decltype(h)::value, S = decltype(b)::value;
        hipError_t r = set_lds(&k_part_hist<uint32_t, 1024, HIST_KPT, H, S>);
        if (r == hipSuccess) r = set_lds(&k_part_hist<uint64_t, 1024, HIST_KPT, H, S>);
        if (r == hipSuccess) r = set_lds(&k_part_scatter<uint32_t, 1024, SCATTER_KPT, H, S>);
        if (r == hipSuccess) r = set_lds(&k_part_scatter<uint64_t, 1024, SCATTER_KPT, H, S>);
        if constexpr (H == HK_MURMUR) {
          if (r == hipSuccess) r = set_lds(&k_part_scatter<uint64_t, 1024, SCATTER_KPT, H, S, true>);
        }
        return r;
      }));
    }
  for (int hk = 0; hk < 3; hk++)
    HG_SET_D(dispatch_hk(hk, [&](auto h) {
      constexpr int H = decltype(h)::value;
      hipError_t r = set_lds(&k_final<uint32_t, uint32_t, 1024, KPT_U32, H>);
      if (r == hipSuccess) r = set_lds(&k_final<uint32_t, uint64_t, 1024, KPT_U32, H>);
      if (r == hipSuccess) r = set_lds(&k_final<uint64_t, uint32_t, 1024, 32, H>);
      if (r == hipSuccess) r = set_lds(&k_final<uint64_t, uint64_t, 1024, 32, H>);
      if constexpr (H == HK_MURMUR) {
        if (r == hipSuccess) r = set_lds(&k_final<uint64_t, uint32_t, 1024, 32, H, true>);
        if (r == hipSuccess) r = set_lds(&k_final<uint64_t, uint64_t, 1024, 32, H, true>);
      }
      if (r == hipSuccess) r = set_lds(&k_probe1<uint32_t, uint32_t, 1024, probe1_kpt<uint32_t>(), H>);
      if (r == hipSuccess) r = set_lds(&k_probe1<uint32_t, uint64_t, 1024, probe1_kpt<uint32_t>(), H>);
      if (r == hipSuccess) r = set_lds(&k_probe1<uint64_t, uint32_t, 1024, probe1_kpt<uint64_t>(), H>);
      if (r == hipSuccess) r = set_lds(&k_probe1<uint64_t, uint64_t, 1024, probe1_kpt<uint64_t>(), H>);
      if (r == hipSuccess) r = set_lds(&k_probe<uint32_t, uint32_t, 1024, KPT_U32, probe_sub<uint32_t>(), H>);
      if (r == hipSuccess) r = set_lds(&k_probe<uint32_t, uint64_t, 1024, KPT_U32, probe_sub<uint64_t>(), H>);
      if (r == hipSuccess) r = set_lds(&k_probe<uint64_t, uint32_t, 1024, 32, probe_sub<uint32_t>(), H>);
      if (r == hipSuccess) r = set_lds(&k_probe<uint64_t, uint64_t, 1024, 32, probe_sub<uint64_t>(), H>);
      return r;
    }));
  HG_SET_D(join_set_lds<JOIN_COUNT>());
  HG_SET_D(join_set_lds<JOIN_KEYS>());
  HG_SET_D(join_set_lds<JOIN_PAIRS>());
  HG_SET(&k_range_hist<uint32_t>);
  HG_SET(&k_range_hist<uint64_t>);
  HG_SET_D(set_rmi_kernel_attributes());
#undef HG_SET
#undef HG_SET_D
  return e;
}

// Bin maps of a batch whose models were trained on the device: table t's
// maps take the learned kind when use[t] is set, murmur3 otherwise.
__global__ __launch_bounds__(256) void k_set_kinds(const uint32_t *__restrict__ use, SegDesc *__restrict__ segs,
                                                   uint32_t nsegs, BinMap *__restrict__ maps, uint32_t nmaps) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < nsegs) segs[i].map.kind = use[segs[i].table] ? 1u : 0u;
  if (i < nmaps) maps[i].kind = use[i] ? 1u : 0u;
}

hipError_t launch_set_kinds(const uint32_t *use, SegDesc *segs, uint32_t nsegs, BinMap *maps, uint32_t nmaps,
                            hipStream_t st) {
  const uint32_t n = nsegs > nmaps ? nsegs : nmaps;
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_set_kinds, dim3((n + 255) / 256), dim3(256), 0, st, use, segs, nsegs, maps, nmaps);
  return hipGetLastError();
}

// Records of the index-carrying builds: the key's input position above the
// low 32 bits of the key (murmur3 bins depend on those bits only), and their
// unpacking into CSR keys (a 32-bit key from the record, a 64-bit one gathered
// from the input by position) and positions.
template <typename K>
__global__ void k_pack_index(const K *__restrict__ keys, uint64_t n, uint64_t *__restrict__ rec) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    rec[i] = (i << 32) | (uint32_t)keys[i];
}
template <typename K>
__global__ void k_unpack_index(const uint64_t *__restrict__ rec, uint64_t n, const K *__restrict__ keys,
                               K *__restrict__ csr_keys, uint64_t *__restrict__ csr_index) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = rec[j], i = r >> 32;
    csr_index[j] = i;
    // a 32-bit key is whole in the record's low half; a 64-bit one is read back
    if constexpr (sizeof(K) == 4)
      csr_keys[j] = (K)r;
    else
      csr_keys[j] = keys[i];
  }
}

hipError_t launch_pack_index(const void *keys, uint64_t n, int key_bytes, uint64_t *rec, hipStream_t st) {
  if (!n) return hipSuccess;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 16384);
  if (key_bytes == 4)
    hipLaunchKernelGGL((k_pack_index<uint32_t>), dim3(grid), dim3(256), 0, st, (const uint32_t *)keys, n, rec);
  else
    hipLaunchKernelGGL((k_pack_index<uint64_t>), dim3(grid), dim3(256), 0, st, (const uint64_t *)keys, n, rec);
  return hipGetLastError();
}

hipError_t launch_unpack_index(const uint64_t *rec, uint64_t n, const void *keys, int key_bytes, void *csr_keys,
                               uint64_t *csr_index, hipStream_t st) {
  if (!n) return hipSuccess;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 16384);
  if (key_bytes == 4)
    hipLaunchKernelGGL((k_unpack_index<uint32_t>), dim3(grid), dim3(256), 0, st, rec, n, (const uint32_t *)keys,
                       (uint32_t *)csr_keys, csr_index);
  else
    hipLaunchKernelGGL((k_unpack_index<uint64_t>), dim3(grid), dim3(256), 0, st, rec, n, (const uint64_t *)keys,
                       (uint64_t *)csr_keys, csr_index);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(64), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL((k_fill<uint32_t>), dim3(256), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
hipError_t launch_offsets_add(void *off, int off_bytes, uint64_t n, uint64_t delta, hipStream_t st) {
  if (!n) return hipSuccess;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  if (off_bytes == 4)
    hipLaunchKernelGGL((k_add<uint32_t>), dim3(grid), dim3(256), 0, st, (uint32_t *)off, (uint32_t)delta, n);
  else
    hipLaunchKernelGGL((k_add<uint64_t>), dim3(grid), dim3(256), 0, st, (uint64_t *)off, delta, n);
  return hipGetLastError();
}

hipError_t launch_gen_uniform(void *out, uint64_t n, int key_bytes, uint64_t s0, uint64_t range, hipStream_t st) {
  if (!n) return hipSuccess;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 16384);
  if (key_bytes == 4)
    hipLaunchKernelGGL((k_gen_uniform<uint32_t>), dim3(grid), dim3(256), 0, st, (uint32_t *)out, n, s0, range);
  else
    hipLaunchKernelGGL((k_gen_uniform<uint64_t>), dim3(grid), dim3(256), 0, st, (uint64_t *)out, n, s0, range);
  return hipGetLastError();
}

hipError_t launch_bins(const void *keys, uint64_t n, int key_bytes, const BinMap &mp, uint32_t *bins, hipStream_t st) {
  if (!n) return hipSuccess;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 8192);
  if (key_bytes == 4)
    hipLaunchKernelGGL((k_bins<uint32_t>), dim3(grid), dim3(256), 0, st, (const uint32_t *)keys, n, mp, bins);
  else
    hipLaunchKernelGGL((k_bins<uint64_t>), dim3(grid), dim3(256), 0, st, (const uint64_t *)keys, n, mp, bins);
  return hipGetLastError();
}

hipError_t launch_range_hist(const void *keys, uint64_t n, int key_bytes, const BinMap &mp, uint32_t range,
                             uint32_t nranges, unsigned long long *hist, unsigned int *bad, hipStream_t st) {
  if (!n) return hipSuccess;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 1023) / 1024, 1024);
  const size_t lds = (size_t)nranges * 4;
  if (key_bytes == 4)
    hipLaunchKernelGGL((k_range_hist<uint32_t>), dim3(grid), dim3(1024), lds, st, (const uint32_t *)keys, n, mp, range,
                       nranges, hist, bad);
  else
    hipLaunchKernelGGL((k_range_hist<uint64_t>), dim3(grid), dim3(1024), lds, st, (const uint64_t *)keys, n, mp, range,
                       nranges, hist, bad);
  return hipGetLastError();
}

}  // namespace hg
