/*
 * hg_oracle.c -- CPU restatement of the HashGraph reference algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only the tests, the smoke check and the benchmark's CPU
 * baseline leg use it, and only to check results / as the CPU port of
 * the reference.  It is plain C (gcc, optional OpenMP), integer arithmetic
 * everywhere except where the reference itself uses float (split points,
 * LRB bucket width) and the learned-hash model, which restates the FP32
 * model evaluation of the product with -ffp-contract=off.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the reference repository root).
 *
 * Pins (what checks each function against something other than itself) are
 * listed in the tests; functions with no independent pin are marked
 * "parity unpinned" below.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* MurmurHash3_x86_32 (Appleby, public domain algorithm).                    */
/* General form with length and seed, used to pin the published test         */
/* vectors; the reference calls the len=4, seed=0 specialisation             */
/* (include/SingleHashGraphOperators.cuh:18-69,                               */
/*  include/MultiHashGraphDeviceOperators.cuh:21-73).                        */
/* ------------------------------------------------------------------------ */
static inline uint32_t rotl32_(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static inline uint32_t fmix32_(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

uint32_t hgo_murmur3_x86_32(const uint8_t *data, uint64_t len, uint32_t seed) {
  const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
  uint32_t h1 = seed;
  uint64_t nblocks = len / 4;
  for (uint64_t i = 0; i < nblocks; i++) {
    uint32_t k1 = (uint32_t)data[4 * i] | ((uint32_t)data[4 * i + 1] << 8) |
                  ((uint32_t)data[4 * i + 2] << 16) | ((uint32_t)data[4 * i + 3] << 24);
    k1 *= c1;
    k1 = rotl32_(k1, 15);
    k1 *= c2;
    h1 ^= k1;
    h1 = rotl32_(h1, 13);
    h1 = h1 * 5 + 0xe6546b64u;
  }
  const uint8_t *tail = data + nblocks * 4;
  uint32_t k1 = 0;
  switch (len & 3) {
    case 3: k1 ^= (uint32_t)tail[2] << 16; /* fall through */
    case 2: k1 ^= (uint32_t)tail[1] << 8;  /* fall through */
    case 1:
      k1 ^= tail[0];
      k1 *= c1;
      k1 = rotl32_(k1, 15);
      k1 *= c2;
      h1 ^= k1;
  }
  h1 ^= (uint32_t)len;
  return fmix32_(h1);
}

/* The reference hashes the first 4 bytes of the key object with seed 0
 * (len = sizeof(int), include/SingleHashGraphOperators.cuh:35; for 64-bit keys
 * include/MultiHashGraphDeviceOperators.cuh:38-49 reads the low 32 bits on a
 * little-endian machine; the single-GPU path converts int64 -> uint32 at the
 * call, include/SingleHashGraphOperators.cuh:76). */
uint32_t hgo_hash_key(uint64_t key) {
  uint8_t b[4];
  uint32_t lo = (uint32_t)key;
  b[0] = (uint8_t)lo; b[1] = (uint8_t)(lo >> 8); b[2] = (uint8_t)(lo >> 16); b[3] = (uint8_t)(lo >> 24);
  return hgo_murmur3_x86_32(b, 4, 0);
}

static inline uint64_t load_key(const void *keys, int kb, uint64_t i) {
  return kb == 4 ? (uint64_t)((const uint32_t *)keys)[i] : ((const uint64_t *)keys)[i];
}
static inline void store_key(void *keys, int kb, uint64_t i, uint64_t v) {
  if (kb == 4) ((uint32_t *)keys)[i] = (uint32_t)v; else ((uint64_t *)keys)[i] = v;
}

/* bin = murmur(low32(key)) % M  (include/SingleHashGraphOperators.cuh:71-79,
 * include/MultiHashGraphDeviceOperators.cuh:76-83). */
void hgo_bins(const void *keys, uint64_t n, int kb, uint64_t M, uint32_t *bins) {
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) bins[i] = (uint32_t)((uint64_t)hgo_hash_key(load_key(keys, kb, (uint64_t)i)) % M);
}

/* ------------------------------------------------------------------------ */
/* Canonical CSR build: counting sort by bin, stable within a bin.            */
/* Same CSR as the reference's plain buildTable<>                             */
/* (include/MultiHashGraphHostOperators.cuh:316-343: hash, count, exclusive   */
/* scan, offsets[M] = N, scatter) with the nondeterministic atomic order      */
/* inside a bin replaced by input order.                                       */
/* ------------------------------------------------------------------------ */
int hgo_build_csr_bins(const void *keys, const uint32_t *bins, uint64_t n, int kb, uint64_t M, uint64_t *off,
                       void *out) {
  if (M == 0) return -1;
  uint64_t *cur = (uint64_t *)calloc(M + 1, 8);
  if (!cur) return -2;
  memset(off, 0, (M + 1) * 8);
  for (uint64_t i = 0; i < n; i++) {
    if (bins[i] >= M) { free(cur); return -3; }
    off[bins[i]]++;
  }
  uint64_t run = 0;
  for (uint64_t b = 0; b < M; b++) { uint64_t c = off[b]; off[b] = run; cur[b] = run; run += c; }
  off[M] = n;  /* src/SingleHashGraph.cu:173 writes offsets[M] = N */
  for (uint64_t i = 0; i < n; i++) store_key(out, kb, cur[bins[i]]++, load_key(keys, kb, i));
  free(cur);
  return 0;
}

int hgo_build_csr(const void *keys, uint64_t n, int kb, uint64_t M, uint64_t *off, void *out) {
  if (M == 0) return -1;
  uint32_t *bins = (uint32_t *)malloc(n ? n * 4 : 4);
  if (!bins) return -2;
  hgo_bins(keys, n, kb, M, bins);
  int rc = hgo_build_csr_bins(keys, bins, n, kb, M, off, out);
  free(bins);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Learned CDF model (no counterpart in the reference): evaluation of the    */
/* FP32 two-stage model, byte layout as the product's (32-byte header: u64   */
/* key_min, f32 root_slope, f32 root_icpt, u32 n_leaves, u32 n_bins, u32      */
/* root_kind, 4 pad bytes; then n_leaves leaves of {u64 x0, f32 slope, u32    */
/* anchor}).  Every product and sum is rounded on its own                     */
/* (-ffp-contract=off), u64 -> f32 conversions round to nearest even, float   */
/* -> u32 truncates.  A key whose leaf line lands outside [0, M-1] takes     */
/* murmur3's bin.  Pinned by closed forms on hand-built models, an exact      */
/* integer restatement of the u64 -> f32 rounding in the tests, the murmur3   */
/* vectors for the out-of-range keys, and the intersection being independent  */
/* of the hash.                                                               */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t key_min; float root_slope, root_icpt; uint32_t n_leaves, n_bins, root_kind, pad; }
    hgo_rmi_header;
typedef struct { uint64_t x0; float slope; uint32_t anchor; } hgo_rmi_leaf;

static float rmi_feature_(uint64_t key, uint64_t key_min, uint32_t kind) {
  float xr = key > key_min ? (float)(key - key_min) : 0.0f;
  if (kind == 0) return xr;
  float x1 = xr + 1.0f;
  uint32_t bits;
  memcpy(&bits, &x1, 4);
  return (float)bits;
}

uint32_t hgo_rmi_leaf_of(const void *model, uint64_t key) {
  const hgo_rmi_header *h = (const hgo_rmi_header *)model;
  float f = rmi_feature_(key, h->key_min, h->root_kind);
  float p = h->root_slope * f;
  float t = p + h->root_icpt;
  uint32_t L = h->n_leaves;
  if (!(t > 0.0f)) return 0;
  if (t >= (float)(L - 1)) return L - 1;
  return (uint32_t)t;
}

uint32_t hgo_rmi_bin(const void *model, uint64_t key, uint64_t M) {
  const hgo_rmi_header *h = (const hgo_rmi_header *)model;
  const hgo_rmi_leaf *lv = (const hgo_rmi_leaf *)(h + 1);
  hgo_rmi_leaf l = lv[hgo_rmi_leaf_of(model, key)];
  int up = key >= l.x0;
  float d = up ? (float)(key - l.x0) : (float)(l.x0 - key);
  float y = l.slope * d;
  int64_t off = !(y > 0.0f) ? 0 : (y >= 4294967040.0f ? (int64_t)0xFFFFFF00u : (int64_t)(uint32_t)y);
  int64_t b = (int64_t)l.anchor + (up ? off : -off);
  /* a line value outside [0, M-1] (a key beyond the trained range): murmur3 */
  if (b < 0 || b > (int64_t)M - 1) return (uint32_t)(hgo_hash_key(key) % M);
  return (uint32_t)b;
}

void hgo_bins_rmi(const void *keys, uint64_t n, int kb, const void *model, uint64_t M, uint32_t *bins) {
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) bins[i] = hgo_rmi_bin(model, load_key(keys, kb, (uint64_t)i), M);
}

/* ------------------------------------------------------------------------ */
/* LRB build: the reference's two-level data flow, restated with CPU threads  */
/* and atomics (src/SingleHashGraph.cu:129-181 and :216-232;                  */
/* kernels include/SingleHashGraphOperators.cuh:71-79,102-154).  Used as the  */
/* CPU baseline ("port").  The order inside a bin is atomic-order, as in the  */
/* reference; compare with hgo_canon() first.                                 */
/* ------------------------------------------------------------------------ */
int hgo_build_lrb(const void *keys, uint64_t n, int kb, uint64_t M, uint64_t lrbBins, int nthreads,
                  uint64_t *off, void *out) {
  if (M == 0 || lrbBins == 0) return -1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  /* src/SingleHashGraph.cu:219: lrbBinSize = ceil(tableSize / (float)lrbBins) */
  uint64_t lrbBinSize = (uint64_t)ceilf((float)M / (float)lrbBins);
  if (lrbBinSize == 0) return -1;
  uint64_t ncoarse = lrbBins + 2;  /* src/SingleHashGraph.cu:97-102 allocates lrbBins + 2 */
  uint64_t maxc = (M - 1) / lrbBinSize;
  if (maxc + 1 > ncoarse) ncoarse = maxc + 2;
  uint32_t *hash = (uint32_t *)malloc(n ? n * 4 : 4);
  uint64_t *lcnt = (uint64_t *)calloc(ncoarse + 1, 8);
  uint64_t *lpre = (uint64_t *)calloc(ncoarse + 1, 8);
  void *lrb = malloc(n ? n * (uint64_t)kb : 8);
  uint64_t *cnt = (uint64_t *)calloc(M + 1, 8);
  if (!hash || !lcnt || !lpre || !lrb || !cnt) { free(hash); free(lcnt); free(lpre); free(lrb); free(cnt); return -2; }
  /* hashValuesD (include/SingleHashGraphOperators.cuh:71-79) */
  hgo_bins(keys, n, kb, M, hash);
  /* lrbCountHashD (include/SingleHashGraphOperators.cuh:102-111) */
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) __atomic_fetch_add(&lcnt[hash[i] / lrbBinSize], 1, __ATOMIC_RELAXED);
  /* ExclusiveSum over the coarse counters (src/SingleHashGraph.cu:147-152) */
  uint64_t run = 0;
  for (uint64_t c = 0; c < ncoarse; c++) { lpre[c] = run; run += lcnt[c]; lcnt[c] = 0; }
  /* lrbRehashD (include/SingleHashGraphOperators.cuh:113-128) */
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint64_t c = hash[i] / lrbBinSize;
    uint64_t pos = __atomic_fetch_add(&lcnt[c], 1, __ATOMIC_RELAXED) + lpre[c];
    store_key(lrb, kb, pos, load_key(keys, kb, (uint64_t)i));
  }
  /* lrbCountHashGlobalD (include/SingleHashGraphOperators.cuh:130-140): rehash the reordered keys */
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint64_t b = (uint64_t)hgo_hash_key(load_key(lrb, kb, (uint64_t)i)) % M;
    __atomic_fetch_add(&cnt[b], 1, __ATOMIC_RELAXED);
  }
  /* ExclusiveSum over M + offsets[M] = N (src/SingleHashGraph.cu:166-177) */
  run = 0;
  for (uint64_t b = 0; b < M; b++) { off[b] = run; run += cnt[b]; cnt[b] = 0; }
  off[M] = n;
  /* lrbCopyToGraphD (include/SingleHashGraphOperators.cuh:142-154) */
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint64_t k = load_key(lrb, kb, (uint64_t)i);
    uint64_t b = (uint64_t)hgo_hash_key(k) % M;
    uint64_t pos = __atomic_fetch_add(&cnt[b], 1, __ATOMIC_RELAXED) + off[b];
    store_key(out, kb, pos, k);
  }
  free(hash); free(lcnt); free(lpre); free(lrb); free(cnt);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Canonical bin contents: sort the keys inside each bin (the reference's     */
/* own comparison criterion, src/MultiHashGraph.cu:826-841).                  */
/* ------------------------------------------------------------------------ */
static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}
static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return (x > y) - (x < y);
}
static void small_sort(void *base, uint64_t cnt, int kb) {
  if (cnt < 2) return;
  if (cnt <= 16) {  /* insertion sort for the common tiny bins */
    if (kb == 4) {
      uint32_t *a = (uint32_t *)base;
      for (uint64_t i = 1; i < cnt; i++) { uint32_t v = a[i]; uint64_t j = i; while (j > 0 && a[j - 1] > v) { a[j] = a[j - 1]; j--; } a[j] = v; }
    } else {
      uint64_t *a = (uint64_t *)base;
      for (uint64_t i = 1; i < cnt; i++) { uint64_t v = a[i]; uint64_t j = i; while (j > 0 && a[j - 1] > v) { a[j] = a[j - 1]; j--; } a[j] = v; }
    }
    return;
  }
  qsort(base, cnt, kb, kb == 4 ? cmp_u32 : cmp_u64);
}
void hgo_canon(const uint64_t *off, void *keys, uint64_t M, int kb) {
  #pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t b = 0; b < (int64_t)M; b++)
    small_sort((uint8_t *)keys + off[b] * (uint64_t)kb, off[b + 1] - off[b], kb);
}
void hgo_sort_keys(void *keys, uint64_t n, int kb) { qsort(keys, n, kb, kb == 4 ? cmp_u32 : cmp_u64); }

/* ------------------------------------------------------------------------ */
/* Intersection of two HashGraphs over the same bin space                     */
/* (src/MultiHashGraph.cu:631-740, kernel simpleIntersect at                  */
/* include/MultiHashGraphDeviceOperators.cuh:513-543):                         */
/*   for each bin, for a in A_bin, for x in B_bin: if a == x emit x.          */
/* Returns the bag-product count; writes up to cap keys (B's key, bin order,  */
/* a-major inside a bin) when out != NULL.                                     */
/* ------------------------------------------------------------------------ */
uint64_t hgo_intersect(const uint64_t *offA, const void *keysA, const uint64_t *offB, const void *keysB,
                       uint64_t M, int kb, void *out, uint64_t cap) {
  uint64_t count = 0;
  for (uint64_t b = 0; b < M; b++) {
    for (uint64_t ia = offA[b]; ia < offA[b + 1]; ia++) {
      uint64_t a = load_key(keysA, kb, ia);
      for (uint64_t ib = offB[b]; ib < offB[b + 1]; ib++) {
        uint64_t x = load_key(keysB, kb, ib);
        if (a == x) {
          if (out && count < cap) store_key(out, kb, count, x);
          count++;
        }
      }
    }
  }
  return count;
}

/* Threaded form of hgo_intersect (CPU baseline): bins are split into one
 * contiguous chunk per thread; counts first, then each thread emits at its
 * prefix.  Same output as hgo_intersect (bin order, a-major). */
uint64_t hgo_intersect_par(const uint64_t *offA, const void *keysA, const uint64_t *offB, const void *keysB,
                           uint64_t M, int kb, void *out, uint64_t cap, int nthreads) {
  int T = nthreads > 0 ? nthreads : 1;
  uint64_t *cnt = (uint64_t *)calloc((size_t)T + 1, 8);
  if (!cnt) return 0;
  #pragma omp parallel num_threads(T)
  {
    int t = 0, nt = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    uint64_t b0 = M * (uint64_t)t / (uint64_t)nt, b1 = M * (uint64_t)(t + 1) / (uint64_t)nt, c = 0;
    for (uint64_t b = b0; b < b1; b++)
      for (uint64_t ia = offA[b]; ia < offA[b + 1]; ia++) {
        uint64_t a = load_key(keysA, kb, ia);
        for (uint64_t ib = offB[b]; ib < offB[b + 1]; ib++) c += load_key(keysB, kb, ib) == a;
      }
    cnt[t + 1] = c;
    #pragma omp barrier
    #pragma omp single
    for (int i = 0; i < nt; i++) cnt[i + 1] += cnt[i];
    if (out) {
      uint64_t w = cnt[t];
      for (uint64_t b = b0; b < b1; b++)
        for (uint64_t ia = offA[b]; ia < offA[b + 1]; ia++) {
          uint64_t a = load_key(keysA, kb, ia);
          for (uint64_t ib = offB[b]; ib < offB[b + 1]; ib++) {
            uint64_t x = load_key(keysB, kb, ib);
            if (x == a) {
              if (w < cap) store_key(out, kb, w, x);
              w++;
            }
          }
        }
    }
  }
  uint64_t total = 0;
  for (int i = 0; i <= T; i++) total = cnt[i] > total ? cnt[i] : total;
  free(cnt);
  return total;
}

/* Same result, run-length form: one record (key, cA, cB) per distinct key   */
/* present in both tables, bins in order, keys ascending inside a bin.  The   */
/* bins must be canonical (hgo_canon).  count = sum cA*cB.  Heavy Zipf bins   */
/* make the nested loop quadratic; this merge is linear per bin.              */
uint64_t hgo_intersect_rle(const uint64_t *offA, const void *keysA, const uint64_t *offB, const void *keysB,
                           uint64_t M, int kb, uint64_t *rec_key, uint64_t *rec_ca, uint64_t *rec_cb,
                           uint64_t cap, uint64_t *nrec_out) {
  uint64_t count = 0, nrec = 0;
  for (uint64_t b = 0; b < M; b++) {
    uint64_t ia = offA[b], ea = offA[b + 1], ib = offB[b], eb = offB[b + 1];
    while (ia < ea && ib < eb) {
      uint64_t a = load_key(keysA, kb, ia), x = load_key(keysB, kb, ib);
      if (a < x) ia++;
      else if (x < a) ib++;
      else {
        uint64_t ca = 0, cb = 0;
        while (ia < ea && load_key(keysA, kb, ia) == a) { ia++; ca++; }
        while (ib < eb && load_key(keysB, kb, ib) == a) { ib++; cb++; }
        if (rec_key && nrec < cap) { rec_key[nrec] = a; rec_ca[nrec] = ca; rec_cb[nrec] = cb; }
        nrec++;
        count += ca * cb;
      }
    }
  }
  if (nrec_out) *nrec_out = nrec;
  return count;
}

/* The reference driver's hash-free CPU answer (test/MultiHashGraphTest.cu:369-397):
 * sort A and B, for every a count the equal elements of B.  Inputs are sorted in place. */
uint64_t hgo_bag_intersect_sorted(void *A, uint64_t nA, void *B, uint64_t nB, int kb) {
  hgo_sort_keys(A, nA, kb);
  hgo_sort_keys(B, nB, kb);
  uint64_t count = 0, j = 0;
  for (uint64_t i = 0; i < nA; i++) {
    uint64_t a = load_key(A, kb, i);
    while (j < nB && load_key(B, kb, j) < a) j++;
    uint64_t k = j;
    while (k < nB && load_key(B, kb, k) == a) { k++; count++; }
  }
  return count;
}

/* ------------------------------------------------------------------------ */
/* Split points (one GPU per range in the reference; logical partitions here) */
/* binarySearch: include/MultiHashGraphHostOperators.cuh:8-27 (size_t         */
/* recursion, returns some index holding x, else the insertion point).        */
/* ------------------------------------------------------------------------ */
static uint64_t ref_binary_search(const int64_t *bins, uint64_t l, uint64_t r, uint64_t x) {
  for (;;) {
    if (r >= l && r != (uint64_t)-1) {
      uint64_t mid = l + (r - l) / 2;
      if ((uint64_t)bins[mid] == x) return mid;
      if ((uint64_t)bins[mid] > x) { if (mid == 0) return l; r = mid - 1; continue; }
      l = mid + 1;
      continue;
    }
    return l;
  }
}

/* countBinSizes (include/MultiHashGraphHostOperators.cuh:29-108 with kernel
 * include/MultiHashGraphDeviceOperators.cuh:153-164), all logical partitions at
 * once (one loop iteration per "tid").  Inputs are the global bins of all keys.
 *   binRange   = ceil(tableSize / (float)binCount)          src/MultiHashGraph.cu:410
 *   histogram of bin / binRange over binCount entries
 *   ps[0]=0, ps[i]=ps[i-1]+h[i-1], ps[binCount]=countSize  HostOperators.cuh:55-59
 *   avg = ceil(countSize / (float)gpuCount)                 HostOperators.cuh:83
 *   splits[t+1] = min(binarySearch(ps, 0, binCount, avg*(t+1)) * binRange, tableSize)
 * splits[0] = 0 (zero-initialised array, src/MultiHashGraph.cu:127).
 * Writes binCount+1 prefix entries to ps_out when non-NULL. Returns 0, or -1
 * when a bin index falls outside the histogram (float rounding made binRange
 * too small; the reference would write out of bounds there). */
int hgo_find_splits(const uint32_t *bins, uint64_t n, uint64_t tableSize, uint64_t binCount, uint64_t gpuCount,
                    int64_t *splits, int64_t *ps_out) {
  if (binCount == 0 || gpuCount == 0 || tableSize == 0) return -1;
  uint64_t binRange = (uint64_t)ceilf((float)tableSize / (float)binCount);
  if (binRange == 0) return -1;
  int64_t *h = (int64_t *)calloc(binCount, 8);
  int64_t *ps = (int64_t *)calloc(binCount + 1, 8);
  if (!h || !ps) { free(h); free(ps); return -2; }
  int bad = 0;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t c = bins[i] / binRange;
    if (c >= binCount) { bad = 1; break; }
    h[c]++;
  }
  if (bad) { free(h); free(ps); return -1; }
  ps[0] = 0;
  for (uint64_t i = 1; i < binCount; i++) ps[i] = ps[i - 1] + h[i - 1];
  ps[binCount] = (int64_t)n;
  int64_t avg = (int64_t)ceilf((float)n / (float)gpuCount);
  splits[0] = 0;
  for (uint64_t t = 0; t < gpuCount; t++) {
    uint64_t upperVal = (uint64_t)(avg * (int64_t)(t + 1));
    uint64_t idx = ref_binary_search(ps, 0, binCount, upperVal);
    uint64_t minRange = idx * binRange;
    splits[t + 1] = (int64_t)(minRange < tableSize ? minRange : tableSize);
  }
  if (ps_out) memcpy(ps_out, ps, (binCount + 1) * 8);
  free(h); free(ps);
  return 0;
}

/* Local partition of keys by split ranges (countKeyBuffSizes + populateKeyBuffs /
 * binHashValues: include/MultiHashGraphHostOperators.cuh:110-200,
 * include/MultiHashGraphDeviceOperators.cuh:220-246,297-345): partition p
 * receives the keys whose global bin is < splits[p+1] and not in an earlier
 * partition.  Stable (input order) here; the reference's order is atomic. */
int hgo_partition(const void *keys, uint64_t n, int kb, uint64_t M, const int64_t *splits, uint64_t P,
                  void *out, uint64_t *part_off) {
  uint32_t *bins = (uint32_t *)malloc(n ? n * 4 : 4);
  uint32_t *part = (uint32_t *)malloc(n ? n * 4 : 4);
  uint64_t *cur = (uint64_t *)calloc(P + 1, 8);
  if (!bins || !part || !cur) { free(bins); free(part); free(cur); return -2; }
  hgo_bins(keys, n, kb, M, bins);
  memset(part_off, 0, (P + 1) * 8);
  for (uint64_t i = 0; i < n; i++) {
    uint64_t p = P;  /* keys beyond the last split are dropped, as in the reference loop */
    for (uint64_t j = 0; j < P; j++) if ((int64_t)bins[i] < splits[j + 1]) { p = j; break; }
    part[i] = (uint32_t)p;
    if (p < P) part_off[p]++;
  }
  uint64_t run = 0;
  for (uint64_t p = 0; p < P; p++) { uint64_t c = part_off[p]; part_off[p] = run; cur[p] = run; run += c; }
  part_off[P] = run;
  for (uint64_t i = 0; i < n; i++) if (part[i] < P) store_key(out, kb, cur[part[i]]++, load_key(keys, kb, i));
  free(bins); free(part); free(cur);
  return 0;
}

/* Per-partition CSR over the local bin range [splits[p], splits[p+1]):
 * local bin = global bin - splits[p] (decrHash, include/MultiHashGraphDeviceOperators.cuh:371-380;
 * lrbCountHashGlobalD/lrbCopyToGraphD :430-458).  Offsets are written with partition p
 * at off + splits[p] + p (R_p + 1 entries each), keys of all partitions contiguously
 * in partition order.  Canonical (stable) order inside bins. */
int hgo_build_multi(const void *keys, uint64_t n, int kb, uint64_t M, const int64_t *splits, uint64_t P,
                    uint64_t *off, void *out) {
  uint64_t *poff = (uint64_t *)calloc(P + 1, 8);
  void *pk = malloc(n ? n * (uint64_t)kb : 8);
  if (!poff || !pk) { free(poff); free(pk); return -2; }
  int rc = hgo_partition(keys, n, kb, M, splits, P, pk, poff);
  if (rc) { free(poff); free(pk); return rc; }
  for (uint64_t p = 0; p < P; p++) {
    uint64_t lo = (uint64_t)splits[p], hi = (uint64_t)splits[p + 1];
    uint64_t R = hi > lo ? hi - lo : 0;
    uint64_t *o = off + lo + p;
    uint64_t cntp = poff[p + 1] - poff[p];
    memset(o, 0, (R + 1) * 8);
    uint8_t *src = (uint8_t *)pk + poff[p] * (uint64_t)kb;
    uint8_t *dst = (uint8_t *)out + poff[p] * (uint64_t)kb;
    uint64_t *cur = (uint64_t *)calloc(R + 1, 8);
    if (!cur) { free(poff); free(pk); return -2; }
    for (uint64_t i = 0; i < cntp; i++) {
      uint64_t b = (uint64_t)hgo_hash_key(load_key(src, kb, i)) % M - lo;
      o[b]++;
    }
    uint64_t run = 0;
    for (uint64_t b = 0; b < R; b++) { uint64_t c = o[b]; o[b] = run; cur[b] = run; run += c; }
    o[R] = cntp;
    for (uint64_t i = 0; i < cntp; i++) {
      uint64_t k = load_key(src, kb, i);
      uint64_t b = (uint64_t)hgo_hash_key(k) % M - lo;
      store_key(dst, kb, cur[b]++, k);
    }
    free(cur);
  }
  free(poff); free(pk);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Workload generators (not in the reference, whose keys come from Thrust's  */
/* minstd engine or a sequence).  SplitMix64 (Steele, Lea & Flood, 2014),    */
/* counter form: x_i = mix64(state0 + (i+1)*gamma) with state0 = mix64(seed). */
/* ------------------------------------------------------------------------ */
static inline uint64_t mix64_(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
#define GOLDEN_GAMMA 0x9E3779B97F4A7C15ull
uint64_t hgo_splitmix(uint64_t seed, uint64_t i) { return mix64_(mix64_(seed) + (i + 1) * GOLDEN_GAMMA); }
