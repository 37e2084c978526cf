// hg_engine.cpp -- host side of the C ABI (include/hashgraph.h): scratch
// arena, descriptor upload and the multi-level build / probe drivers.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hashgraph.h"
#include "hg_common.hpp"
#include "hg_kernels.hpp"

using namespace hg;

namespace {

thread_local std::string g_err;

hg_status fail(hg_status s, const std::string &m) {
  g_err = m;
  return s;
}

#define HG_HIP(call)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) return fail(HG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define HG_TRY(call)                 \
  do {                               \
    hg_status s_ = (call);           \
    if (s_ != HG_OK) return s_;      \
  } while (0)

constexpr uint32_t kFMax = 16384;     // buckets per segment of one partition pass
constexpr uint32_t kTileKpt = 32;     // keys per thread of a partition tile
constexpr uint32_t kTile = 1024 * kTileKpt;
constexpr uint32_t kChunk = 64;       // count-matrix rows per column-scan chunk

uint64_t splitmix_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
  z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
  return z ^ (z >> 31);
}

// Invariant-divider constants (the derivation is in hg_common.hpp).
Div32 make_div(uint32_t d) {
  Div32 v{0, 0, d, 0};
  if (d == 0) return v;
  const uint32_t l = 31 - (uint32_t)__builtin_clz(d);
  if ((d & (d - 1)) == 0) {
    v.s = l;
    return v;
  }
  const unsigned __int128 p = (unsigned __int128)1 << (32 + l);
  const uint64_t m1 = (uint64_t)(p / d) + 1;  // ceil(p / d): p % d != 0 for d not a power of two
  const uint64_t e = (uint64_t)((unsigned __int128)m1 * d - p);
  if (e <= (1ull << l)) {
    v.mode = 1;
    v.m = (uint32_t)m1;
    v.s = l;
    return v;
  }
  const unsigned __int128 p2 = (unsigned __int128)1 << (33 + l);
  const unsigned __int128 m2 = p2 / d + 1;  // in (2^32, 2^33)
  v.mode = 2;
  v.m = (uint32_t)(m2 - ((unsigned __int128)1 << 32));
  v.s = l;
  return v;
}

uint32_t final_cap(int kb) { return std::min<uint32_t>(1024u * (uint32_t)final_kpt(kb), 65535u); }

// Bucket fill of the build (HG_BUILD_FILL overrides it for tuning: above 1
// buckets exceed the final stage's VGPR capacity and take k_final_big).
double build_fill() {
  if (const char *e = getenv("HG_BUILD_FILL")) {
    const double v = atof(e);
    if (v > 0.01 && v < 64.0) return v;
  }
  return 0.875;
}

// Super-buckets of a two-level partition (0: one level whenever the buckets
// fit one pass).  HG_L1_FANOUT sets a first-level fan-out for tuning: a table
// with more than twice that many buckets is then partitioned in two
// low-fan-out passes.
uint64_t l1_fanout() {
  if (const char *e = getenv("HG_L1_FANOUT")) {
    const long v = atol(e);
    if (v >= 2 && v <= (long)kFMax) return (uint64_t)v;
  }
  return 0;
}
// Buckets per super-bucket of a table of F buckets (1: one level).
uint64_t level1_group(uint64_t F) {
  uint64_t k = F > kFMax ? (F + kFMax - 1) / kFMax : 1;
  const uint64_t f1 = l1_fanout();
  if (f1 && F > 2 * f1) k = std::max<uint64_t>(k, (F + f1 - 1) / f1);
  return k;
}

// Bucket fill of the probe: a bucket's keys of A and its 16-bit offsets must
// fit the one-pass probe's 156 KB of LDS (W bins at load factor 1 take 6 W
// bytes for uint32 keys, 10 W for uint64: 0.72 and 0.43 of 32 K bins leave
// ~10% headroom), and its keys of B its VGPRs (32 K uint32 or 16 K uint64).
double probe_fill(int kb) { return kb == 4 ? 0.72 : 0.43; }

// Bucket width W (bins) of a table so that a bucket holds ~7/8 of what the
// final stage keeps in VGPRs; W <= 65535 (u16 LDS counters).
uint32_t choose_width(uint64_t n, uint64_t M, int kb, double fill = 0.875) {
  const double target = fill * final_cap(kb);
  double w = n ? target * (double)M / (double)n : 65535.0;
  w = std::floor(w);
  if (w < 1) w = 1;
  if (w > 65535) w = 65535;
  return (uint32_t)w;
}

}  // namespace

struct hg_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // arena 0 holds an operation's large buffers (partitioned copies), arena 1
  // one partition level's descriptors and matrices; arena 1 may be regrown
  // between levels without touching arena 0.
  // arena 2 holds the records of the index-carrying builds, arena 3 the
  // chunk and slab buffers of the staged (host-resident) tables
  char *arena[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t arena_cap[4] = {0, 0, 0, 0};
  static constexpr int kSlots = 4;
  char *hslot[kSlots] = {nullptr, nullptr, nullptr, nullptr};
  size_t hslot_cap[kSlots] = {0, 0, 0, 0};
  hipEvent_t hev[kSlots] = {nullptr, nullptr, nullptr, nullptr};
  bool hev_used[kSlots] = {false, false, false, false};
  int next_slot = 0;
  unsigned long long *h_small = nullptr;  // pinned read-back area
  // staged tables: upload and copy-back streams and their events
  hipStream_t st_in = nullptr, st_out = nullptr;
  hipEvent_t st_ev[6] = {};
  // per-kernel timing (hg_ctx_profile)
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Rec {
    const char *name;
    size_t a, b;
  };
  std::vector<Rec> recs;
};

namespace {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// Brackets the launches of one kernel (or kernel pair) with events when the
// context profiles.
struct Prof {
  hg_ctx c;
  const char *name;
  size_t a = 0;
  bool on;
  hipEvent_t get(size_t &idx) {
    if (c->ev_used == c->ev_pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      c->ev_pool.push_back(e);
    }
    idx = c->ev_used++;
    return c->ev_pool[idx];
  }
  Prof(hg_ctx c_, const char *n) : c(c_), name(n), on(c_->prof) {
    if (on) {
      hipEvent_t e = get(a);
      if (e) (void)hipEventRecord(e, c->stream);
      else on = false;
    }
  }
  ~Prof() {
    if (!on) return;
    size_t b;
    hipEvent_t e = get(b);
    if (!e) return;
    (void)hipEventRecord(e, c->stream);
    c->recs.push_back({name, a, b});
  }
};

hg_status arena_ensure(hg_ctx c, int a, size_t bytes) {
  if (bytes <= c->arena_cap[a]) return HG_OK;
  HG_HIP(hipStreamSynchronize(c->stream));
  size_t want = std::max(bytes, c->arena_cap[a] * 5 / 4);
  want = (want + (1 << 20) - 1) & ~(size_t)((1 << 20) - 1);
  if (c->arena[a]) HG_HIP(hipFree(c->arena[a]));
  c->arena[a] = nullptr;
  c->arena_cap[a] = 0;
  if (hipMalloc(&c->arena[a], want) != hipSuccess) {
    (void)hipGetLastError();
    return fail(HG_ERR_OOM, "scratch arena: cannot allocate " + std::to_string(want) + " bytes");
  }
  c->arena_cap[a] = want;
  return HG_OK;
}

// Sizes of the arena pieces of one call, carved in order.
struct Plan {
  std::vector<size_t> sizes;
  size_t add(size_t bytes) {
    sizes.push_back((bytes + 255) & ~(size_t)255);
    return sizes.size() - 1;
  }
  size_t total() const {
    size_t t = 0;
    for (size_t s : sizes) t += s;
    return t;
  }
};

struct Carve {
  std::vector<char *> ptr;
  void init(hg_ctx c, int a, const Plan &p) {
    ptr.clear();
    size_t off = 0;
    for (size_t s : p.sizes) {
      ptr.push_back(c->arena[a] + off);
      off += s;
    }
  }
  template <typename T>
  T *get(size_t i) const { return (T *)ptr[i]; }
};

// Copies the host blob to the device through a pinned slot; the slot is not
// rewritten before the previous copy out of it has finished.
hg_status upload(hg_ctx c, void *dst, const void *src, size_t bytes) {
  if (!bytes) return HG_OK;
  const int s = c->next_slot;
  c->next_slot = (s + 1) % hg_ctx_s::kSlots;
  if (c->hev_used[s]) HG_HIP(hipEventSynchronize(c->hev[s]));
  if (c->hslot_cap[s] < bytes) {
    if (c->hslot[s]) HG_HIP(hipHostFree(c->hslot[s]));
    c->hslot[s] = nullptr;
    size_t want = std::max<size_t>(bytes, 1 << 16);
    HG_HIP(hipHostMalloc((void **)&c->hslot[s], want, hipHostMallocDefault));
    c->hslot_cap[s] = want;
  }
  memcpy(c->hslot[s], src, bytes);
  HG_HIP(hipMemcpyAsync(dst, c->hslot[s], bytes, hipMemcpyHostToDevice, c->stream));
  HG_HIP(hipEventRecord(c->hev[s], c->stream));
  c->hev_used[s] = true;
  return HG_OK;
}

// Host blob builder: appends arrays with 256-byte alignment.
struct Blob {
  std::vector<char> data;
  template <typename T>
  size_t put(const std::vector<T> &v) {
    size_t off = (data.size() + 255) & ~(size_t)255;
    data.resize(off + std::max<size_t>(v.size() * sizeof(T), 1));
    if (!v.empty()) memcpy(data.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

BinMap make_map(uint64_t M, uint32_t kind, const void *rmi) {
  BinMap m;
  m.M = (uint32_t)M;
  m.mod = make_div((uint32_t)M);
  m.kind = kind;
  m.rmi = rmi;
  return m;
}

hg_status check_desc(const hg_table_desc *d, bool need_keys = true) {
  if (!d) return fail(HG_ERR_INVALID_ARG, "null table description");
  if (d->key_bytes != 4 && d->key_bytes != 8) return fail(HG_ERR_INVALID_ARG, "key_bytes must be 4 or 8");
  if (d->offset_bytes != 4 && d->offset_bytes != 8) return fail(HG_ERR_INVALID_ARG, "offset_bytes must be 4 or 8");
  if (d->n_bins == 0 || d->n_bins >= (1ull << 32)) return fail(HG_ERR_INVALID_ARG, "n_bins must be in [1, 2^32)");
  if (d->hash_kind != HG_HASH_MURMUR3 && d->hash_kind != HG_HASH_RMI)
    return fail(HG_ERR_INVALID_ARG, "unknown hash_kind");
  if (d->hash_kind == HG_HASH_RMI && !d->d_rmi_model) return fail(HG_ERR_INVALID_ARG, "HG_HASH_RMI needs d_rmi_model");
  if (need_keys && d->n_keys >= (1ull << 32))
    return fail(HG_ERR_UNSUPPORTED, "n_keys >= 2^32 is not supported by this build");
  if (d->offset_bytes == 4 && d->n_keys >= (1ull << 32))
    return fail(HG_ERR_INVALID_ARG, "offset_bytes 4 needs n_keys < 2^32");
  return HG_OK;
}

// One partition pass over host-described segments.  On the device it
// produces positions in cmat, bucket starts (bstart) and sizes (coltot) for
// every global column.  Its descriptors and matrices live in arena 1.
struct Level {
  std::vector<SegDesc> segs;   // tile_begin/ntiles/row_base/bucket_base filled by prepare()
  std::vector<TileDesc> tiles;
  std::vector<ColItem> items, items2;
  std::vector<uint64_t> psum_base;
  std::vector<uint32_t> nchunks;
  std::vector<BinMap> maps;
  std::vector<int64_t> splits;  // kind-1 segments point into this array
  uint32_t ncols_total = 0, max_cols = 0;
  uint64_t cmat_elems = 0, psum_elems = 0;
  uint32_t tile_keys = kTile;   // keys per partition tile
  uint32_t big_cap = 1;         // per-bin counters of the final stage's oversized buckets

  bool any_kind = false;        // kinds change on the device (trained in a plan): general kernels
  // One kind shared by all segments selects a specialised kernel.
  int hash_kind() const {
    if (any_kind) return KIND_MIXED;
    int k = -1;
    for (const BinMap &m : maps) k = (k < 0 || k == (int)m.kind) ? (int)m.kind : KIND_MIXED;
    for (const SegDesc &sd : segs) k = (k < 0 || k == (int)sd.map.kind) ? (int)sd.map.kind : KIND_MIXED;
    return k < 0 ? 0 : k;
  }
  int bucket_kind() const {
    int k = -1;
    for (const SegDesc &sd : segs) k = (k < 0 || k == (int)sd.kind) ? (int)sd.kind : KIND_MIXED;
    return k < 0 ? 0 : k;
  }

  // Keys per tile: rounds of 32 K keys.  More rounds give longer (tile,
  // bucket) runs and a smaller count matrix but fewer, longer workgroups; on
  // 2^28 keys (F = 9363) one round took 1.24 + 0.31 ms (scatter + column
  // scan), two 1.31 + 0.16 ms, four 1.42 + 0.09 ms, ten 1.91 + 0.08 ms.  Two
  // rounds are used from 2^25 keys on (>= 512 tiles remain).  HG_TILE_ROUNDS
  // sets the round count (1..64) for tuning.
  void auto_tile(uint64_t n) { tile_keys = tile_keys_for(n); }
  static uint32_t tile_keys_for(uint64_t n) {
    uint64_t r = n >= (1ull << 25) ? 2 : 1;
    if (const char *e = getenv("HG_TILE_ROUNDS")) {
      const long v = atol(e);
      if (v >= 1 && v <= 64) r = (uint64_t)v;
    }
    return (uint32_t)(kTile * r);
  }

  void prepare() {
    tiles.clear();
    items.clear();
    items2.clear();
    psum_base.clear();
    nchunks.clear();
    uint64_t row = 0, ps = 0;
    uint32_t bb = 0;
    max_cols = 1;
    for (uint32_t s = 0; s < segs.size(); s++) {
      SegDesc &sd = segs[s];
      sd.tile_begin = (uint32_t)tiles.size();
      for (uint64_t k = sd.key_begin; k < sd.key_end; k += tile_keys) {
        TileDesc t;
        t.key_start = k;
        t.count = (uint32_t)std::min<uint64_t>(tile_keys, sd.key_end - k);
        t.seg = s;
        tiles.push_back(t);
      }
      sd.ntiles = (uint32_t)tiles.size() - sd.tile_begin;
      sd.row_base = row;
      sd.bucket_base = bb;
      row += (uint64_t)sd.ntiles * sd.ncols;
      bb += sd.ncols;
      max_cols = std::max(max_cols, sd.ncols);
      const uint32_t nch = (sd.ntiles + kChunk - 1) / kChunk;
      psum_base.push_back(ps);
      nchunks.push_back(nch);
      ps += (uint64_t)nch * sd.ncols;
      for (uint32_t cb = 0; cb < sd.ncols; cb += 256) {
        ColItem ci{};
        ci.seg = s;
        ci.col_lo = cb;
        items2.push_back(ci);
        for (uint32_t ch = 0; ch < nch; ch++) {
          ColItem x{};
          x.seg = s;
          x.chunk = ch;
          x.row_lo = ch * kChunk;
          x.row_hi = std::min(sd.ntiles, (ch + 1) * kChunk);
          x.col_lo = cb;
          items.push_back(x);
        }
      }
    }
    ncols_total = bb;
    cmat_elems = row;
    psum_elems = ps;
  }
};

struct LevelDev {
  TileDesc *tiles;
  SegDesc *segs;
  ColItem *items, *items2;
  uint64_t *psum_base;
  uint32_t *nchunks;
  BinMap *maps;
  int64_t *splits;
  uint32_t *cmat, *psum, *coltot, *bstart;
  FinItem *fin;
  uint32_t *big_slot, *big_cnt;
  void *big_ctl;
};

// Lays out a level in arena 1 (growing it if needed; the stream must be
// idle of earlier users of arena 1, which holds because every level is
// planned either at the start of an operation or after a read-back), patches
// device pointers into the descriptors and uploads them.
hg_status setup_level(hg_ctx c, Level &L, LevelDev &d) {
  Blob b0;
  const size_t o_tiles = b0.put(L.tiles), o_segs = b0.put(L.segs), o_items = b0.put(L.items), o_items2 = b0.put(L.items2),
               o_psb = b0.put(L.psum_base), o_nch = b0.put(L.nchunks), o_maps = b0.put(L.maps), o_spl = b0.put(L.splits);
  Plan P;
  const size_t i_blob = P.add(b0.data.size());
  const size_t i_cmat = P.add(std::max<uint64_t>(L.cmat_elems, 1) * 4);
  const size_t i_psum = P.add(std::max<uint64_t>(L.psum_elems, 1) * 4);
  const size_t i_coltot = P.add(std::max<uint32_t>(L.ncols_total, 1) * 4);
  const size_t i_bstart = P.add(std::max<uint32_t>(L.ncols_total, 1) * 4);
  const size_t i_fin = P.add((size_t)std::max<uint32_t>(L.ncols_total, 1) * sizeof(FinItem));
  // per-bin counters of the oversized buckets: 256 M counters at most (1 GB),
  // fewer when the level has fewer bins
  uint64_t nbins_total = 0;
  for (const SegDesc &sd : L.segs) nbins_total += sd.seg_nbins;
  L.big_cap = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(nbins_total, 1ull << 28));
  const size_t i_bslot = P.add((size_t)std::max<uint32_t>(L.ncols_total, 1) * 4);
  const size_t i_bctl = P.add(final_big_ctl_bytes());
  const size_t i_bcnt = P.add((size_t)L.big_cap * 4);
  HG_TRY(arena_ensure(c, 1, P.total()));
  Carve cv;
  cv.init(c, 1, P);
  char *blob = cv.get<char>(i_blob);
  d.tiles = (TileDesc *)(blob + o_tiles);
  d.segs = (SegDesc *)(blob + o_segs);
  d.items = (ColItem *)(blob + o_items);
  d.items2 = (ColItem *)(blob + o_items2);
  d.psum_base = (uint64_t *)(blob + o_psb);
  d.nchunks = (uint32_t *)(blob + o_nch);
  d.maps = (BinMap *)(blob + o_maps);
  d.splits = (int64_t *)(blob + o_spl);
  d.cmat = cv.get<uint32_t>(i_cmat);
  d.psum = cv.get<uint32_t>(i_psum);
  d.coltot = cv.get<uint32_t>(i_coltot);
  d.bstart = cv.get<uint32_t>(i_bstart);
  d.fin = cv.get<FinItem>(i_fin);
  d.big_slot = cv.get<uint32_t>(i_bslot);
  d.big_ctl = cv.get<void>(i_bctl);
  d.big_cnt = cv.get<uint32_t>(i_bcnt);
  for (SegDesc &sd : L.segs)
    if (sd.kind == 1) sd.splits = d.splits;
  Blob b1;
  b1.put(L.tiles);
  b1.put(L.segs);
  b1.put(L.items);
  b1.put(L.items2);
  b1.put(L.psum_base);
  b1.put(L.nchunks);
  b1.put(L.maps);
  b1.put(L.splits);
  return upload(c, blob, b1.data.data(), b1.data.size());
}

// scatter = false stops after the column scan (bucket sizes only).
hg_status run_level(hg_ctx c, const Level &L, const LevelDev &d, int kb, const void *in, void *out,
                    bool scatter = true, bool pack = false) {
  LaunchPart p{};
  p.key_bytes = kb;
  p.pack = pack;
  p.hk = L.hash_kind();
  p.sk = L.bucket_kind();
  p.ntiles = (uint32_t)L.tiles.size();
  p.max_cols = L.max_cols;
  p.stage_cap = part_stage_cap(L.max_cols, kb, p.sk);
  if (L.tiles.size() && p.stage_cap < 1024) return fail(HG_ERR_UNSUPPORTED, "partition level too wide for LDS");
  p.keys = in;
  p.out = out;
  p.tiles = d.tiles;
  p.segs = d.segs;
  p.cmat = d.cmat;
  if (p.ntiles) {
    Prof pf(c, "part_hist");
    HG_HIP(launch_part_hist(p, c->stream));
  }
  LaunchCol lc{};
  lc.nitems = (uint32_t)L.items.size();
  lc.nitems2 = (uint32_t)L.items2.size();
  lc.nsegs = (uint32_t)L.segs.size();
  lc.cmat = d.cmat;
  lc.items = d.items;
  lc.items2 = d.items2;
  lc.segs = d.segs;
  lc.psum_base = d.psum_base;
  lc.nchunks = d.nchunks;
  lc.psum = d.psum;
  lc.coltot = d.coltot;
  lc.bstart = d.bstart;
  if (lc.nsegs) {
    Prof pf(c, "col_scan");
    HG_HIP(launch_colscan(lc, c->stream));
  }
  if (p.ntiles && scatter) {
    Prof pf(c, "part_scatter");
    HG_HIP(launch_part_scatter(p, c->stream));
  }
  return HG_OK;
}

SegDesc range_seg(uint64_t kbeg, uint64_t kend, uint64_t bin_lo, uint64_t nbins, uint32_t W, const BinMap &m) {
  SegDesc s{};
  s.key_begin = kbeg;
  s.key_end = kend;
  s.bin_lo = bin_lo;
  s.map = m;
  s.div = make_div(W);
  s.kind = 0;
  s.nbuckets = (uint32_t)((nbins + W - 1) / W);
  s.ncols = s.nbuckets;
  s.seg_nbins = nbins;
  s.fin_end_pos = ~0ull;
  return s;
}

// Final stage over the columns of a level.
hg_status run_final(hg_ctx c, const Level &L, const LevelDev &d, int kb, int ob, const void *part, void *csr, void *off,
                    uint64_t *csr_index = nullptr) {
  HG_HIP(launch_make_items(d.segs, (uint32_t)L.segs.size(), d.bstart, d.coltot, d.fin, L.ncols_total, c->stream));
  LaunchFinal f{};
  f.key_bytes = kb;
  f.off_bytes = ob;
  f.hk = L.hash_kind();
  f.nitems = L.ncols_total;
  f.lds_bytes = 128 * 1024;
  f.part = part;
  f.csr = csr;
  f.off = off;
  f.items = d.fin;
  f.maps = d.maps;
  f.big_slot = d.big_slot;
  f.big_ctl = d.big_ctl;
  f.big_cnt = d.big_cnt;
  f.big_cap = L.big_cap;
  f.csr_index = csr_index;
  {
    Prof pf(c, "final");
    HG_HIP(launch_final(f, c->stream));
  }
  return HG_OK;
}

hg_status sync_read(hg_ctx c, void *dst, const void *src, size_t bytes) {
  if (bytes > (1 << 20)) return fail(HG_ERR_INVALID_ARG, "read-back too large");
  HG_HIP(hipMemcpyAsync(c->h_small, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HG_HIP(hipStreamSynchronize(c->stream));
  memcpy(dst, c->h_small, bytes);
  return HG_OK;
}

hg_status fill_offsets(hg_ctx c, void *off, int ob, uint64_t n, uint64_t v) {
  if (!n) return HG_OK;
  if (ob == 4) HG_HIP(launch_fill_u32((uint32_t *)off, (uint32_t)v, n, c->stream));
  else HG_HIP(launch_fill_u64((unsigned long long *)off, (unsigned long long)v, n, c->stream));
  return HG_OK;
}

// Second partition level over the buckets of a first level whose sizes were
// read back: each first-level bucket becomes a range segment.
struct Coarse {
  uint64_t key_begin, count, bin_lo, nbins;
  uint64_t off_base, key_rel, end_pos;
};

hg_status finish_two_level(hg_ctx c, const std::vector<Coarse> &cs, const BinMap &map, int kb, int ob, const void *part1,
                           void *part2, void *csr, void *off, uint64_t *csr_index = nullptr) {
  Level L2;
  for (const Coarse &q : cs) {
    if (q.nbins == 0) continue;
    const uint32_t W = choose_width(q.count, q.nbins, kb);
    SegDesc s = range_seg(q.key_begin, q.key_begin + q.count, q.bin_lo, q.nbins, W, map);
    if (s.nbuckets > kFMax) return fail(HG_ERR_UNSUPPORTED, "table too large or too skewed for two partition levels");
    s.fin_off_base = q.off_base;
    s.fin_key_rel = q.key_rel;
    s.fin_end_pos = q.end_pos;
    s.table = 0;
    L2.segs.push_back(s);
  }
  if (L2.segs.empty()) return HG_OK;
  L2.maps.push_back(map);
  uint64_t n2 = 0;
  for (const SegDesc &sd : L2.segs) n2 += sd.key_end - sd.key_begin;
  L2.auto_tile(n2);
  L2.prepare();
  LevelDev d2;
  HG_TRY(setup_level(c, L2, d2));
  HG_TRY(run_level(c, L2, d2, kb, part1, part2));
  return run_final(c, L2, d2, kb, ob, part2, csr, off, csr_index);
}

// CSR of n keys whose bins all lie in [lo, lo + R): d_off[0..R] with values
// 0..n (positions in d_csr) and d_csr.  hg_build is the range [0, M); a
// staged build runs it once per slab of bins.
hg_status build_range(hg_ctx c, const BinMap &map, int kb, int ob, uint64_t n, uint64_t lo, uint64_t R,
                      const void *d_keys, void *d_off, void *d_csr, uint64_t *d_csr_index = nullptr,
                      bool pack_in = false) {
  if (n == 0) return fill_offsets(c, d_off, ob, R + 1, 0);
  const uint32_t W = choose_width(n, R, kb, build_fill());
  const uint64_t F = (R + W - 1) / W;
  const uint64_t k = level1_group(F);
  if (k == 1) {
    Level L;
    SegDesc s = range_seg(0, n, lo, R, W, map);
    s.fin_off_base = 0;
    s.fin_key_rel = 0;
    s.fin_end_pos = R;
    s.table = 0;
    L.segs.push_back(s);
    L.maps.push_back(map);
    L.auto_tile(n);
    L.prepare();
    Plan P0;
    const size_t i_part = P0.add(n * kb);
    HG_TRY(arena_ensure(c, 0, P0.total()));
    Carve cv0;
    cv0.init(c, 0, P0);
    LevelDev d;
    HG_TRY(setup_level(c, L, d));
    void *part = cv0.get<void>(i_part);
    HG_TRY(run_level(c, L, d, kb, d_keys, part, true, pack_in));
    return run_final(c, L, d, kb, ob, part, d_csr, d_off, d_csr_index);
  }
  // Two levels: coarse super-buckets of k*W bins, their sizes read back, then
  // every super-bucket is a segment of the second pass.
  const uint64_t W1 = (uint64_t)W * k;
  if (W1 >= (1ull << 32)) return fail(HG_ERR_UNSUPPORTED, "table geometry too large");
  Level L1;
  L1.segs.push_back(range_seg(0, n, lo, R, (uint32_t)W1, map));
  L1.maps.push_back(map);
  L1.auto_tile(n);
  L1.prepare();
  Plan P0;
  const size_t i_part1 = P0.add(n * kb), i_part2 = P0.add(n * kb);
  HG_TRY(arena_ensure(c, 0, P0.total()));
  Carve cv0;
  cv0.init(c, 0, P0);
  LevelDev d1;
  HG_TRY(setup_level(c, L1, d1));
  HG_TRY(run_level(c, L1, d1, kb, d_keys, cv0.get<void>(i_part1), true, pack_in));
  std::vector<uint32_t> cnt(L1.ncols_total), st(L1.ncols_total);
  HG_TRY(sync_read(c, cnt.data(), d1.coltot, cnt.size() * 4));
  HG_TRY(sync_read(c, st.data(), d1.bstart, st.size() * 4));
  std::vector<Coarse> cs;
  for (uint32_t g = 0; g < L1.ncols_total; g++) {
    Coarse q;
    q.key_begin = st[g];
    q.count = cnt[g];
    q.off_base = (uint64_t)g * W1;
    q.bin_lo = lo + q.off_base;
    q.nbins = std::min<uint64_t>(W1, R - q.off_base);
    q.key_rel = 0;
    q.end_pos = (g + 1 == L1.ncols_total) ? R : ~0ull;
    cs.push_back(q);
  }
  return finish_two_level(c, cs, map, kb, ob, cv0.get<void>(i_part1), cv0.get<void>(i_part2), d_csr, d_off,
                          d_csr_index);
}

// Driver of hg_build.
hg_status build_table(hg_ctx c, const hg_table_desc *desc, const void *d_keys, void *d_off, void *d_csr) {
  const BinMap map = make_map(desc->n_bins, desc->hash_kind, desc->d_rmi_model);
  return build_range(c, map, (int)desc->key_bytes, (int)desc->offset_bytes, desc->n_keys, 0, desc->n_bins, d_keys,
                     d_off, d_csr);
}

// Heavy bins of the run-length join with more distinct values than the LDS
// and giant-bin tables hold: one global-memory table per bin of at least
// twice its keys of A (4096 slots at least), laid out in batches of at most
// kOvfBatchSlots slots (a larger bin takes a batch of its own).  Runs after
// the rest of the join; its records follow the others.
constexpr uint64_t kOvfBatchSlots = 1ull << 26;
// Key pairs of the row-id join's bins above 4096 pairs, compared one by one,
// beyond which the call is refused (~2^36 compares take about a second).
constexpr uint64_t kPairCompareLimit = 1ull << 36;
// Padded records of the run-length join's single pass (record + heavy-list
// bytes) beyond which the two-pass form is used.
constexpr uint64_t kRleOnePassBytes = 4ull << 30;
constexpr uint64_t kOvfChunk = 1ull << 16;

hg_status rle_value_rich(hg_ctx c, const LaunchRle &r, const uint32_t *d_list, uint32_t n) {
  const int kb = r.key_bytes;
  // (a0, a1, x0, x1) of each listed bin, read back in pieces
  std::vector<unsigned long long> sz((size_t)n * 4);
  {
    Plan P;
    const size_t i_sz = P.add((size_t)n * 32);
    HG_TRY(arena_ensure(c, 0, P.total()));
    Carve cv;
    cv.init(c, 0, P);
    unsigned long long *d_sz = cv.get<unsigned long long>(i_sz);
    HG_HIP(launch_ovf_sizes(d_list, n, r.offA, r.offB, r.off_bytes, d_sz, c->stream));
    for (size_t i = 0; i < sz.size(); i += 1u << 17) {
      const size_t m = std::min<size_t>(1u << 17, sz.size() - i);
      HG_TRY(sync_read(c, sz.data() + i, d_sz + i, m * 8));
    }
  }
  uint32_t i0 = 0;
  while (i0 < n) {
    // one batch: bins i0 .. i1
    std::vector<OvfTable> tabs;
    std::vector<OvfChunk> ca, cb;
    uint64_t slots = 0;
    uint32_t i1 = i0;
    while (i1 < n) {
      const uint64_t a0 = sz[4 * i1], a1 = sz[4 * i1 + 1], x0 = sz[4 * i1 + 2], x1 = sz[4 * i1 + 3];
      uint64_t t = 4096;
      while (t < 2 * (a1 - a0)) t <<= 1;
      if (i1 > i0 && slots + t > kOvfBatchSlots) break;
      OvfTable tb;
      tb.base = slots;
      tb.mask = t - 1;
      const uint32_t ti = (uint32_t)tabs.size();
      tabs.push_back(tb);
      slots += t;
      for (uint64_t k = a0; k < a1; k += kOvfChunk) ca.push_back({k, std::min(a1, k + kOvfChunk), ti, 0});
      for (uint64_t k = x0; k < x1; k += kOvfChunk) cb.push_back({k, std::min(x1, k + kOvfChunk), ti, 0});
      i1++;
    }
    Blob bl;
    const size_t o_tabs = bl.put(tabs), o_ca = bl.put(ca), o_cb = bl.put(cb);
    Plan P;
    const size_t i_key = P.add(slots * kb), i_ca = P.add(slots * 4), i_cb = P.add(slots * 4),
                 i_e = P.add(tabs.size() * 16), i_blob = P.add(bl.data.size());
    HG_TRY(arena_ensure(c, 0, P.total()));
    Carve cv;
    cv.init(c, 0, P);
    char *blob = cv.get<char>(i_blob);
    HG_TRY(upload(c, blob, bl.data.data(), bl.data.size()));
    LaunchRleOvf o{};
    o.key_bytes = kb;
    o.keysA = r.keysA;
    o.keysB = r.keysB;
    o.tabs = (const OvfTable *)(blob + o_tabs);
    o.ntabs = (uint32_t)tabs.size();
    o.chunksA = (const OvfChunk *)(blob + o_ca);
    o.chunksB = (const OvfChunk *)(blob + o_cb);
    o.nchunksA = (uint32_t)ca.size();
    o.nchunksB = (uint32_t)cb.size();
    o.nslots = slots;
    o.tkey = cv.get<void>(i_key);
    o.tca = cv.get<uint32_t>(i_ca);
    o.tcb = cv.get<uint32_t>(i_cb);
    o.eA = cv.get<unsigned long long>(i_e);
    o.eB = o.eA + tabs.size();
    o.rk = r.rk;
    o.rca = r.rca;
    o.rcb = r.rcb;
    o.capacity = r.capacity;
    o.cursor = r.cursor;
    o.pairs = r.pairs;
    if (kb == 4) HG_HIP(launch_fill_u32((uint32_t *)o.tkey, 0xFFFFFFFFu, slots, c->stream));
    else HG_HIP(launch_fill_u64((unsigned long long *)o.tkey, ~0ull, slots, c->stream));
    HG_HIP(launch_fill_u32(o.tca, 0u, slots, c->stream));
    HG_HIP(launch_fill_u32(o.tcb, 0u, slots, c->stream));
    HG_HIP(launch_fill_u64(o.eA, 0ull, 2 * tabs.size(), c->stream));
    {
      Prof pf(c, "rle_value_rich");
      HG_HIP(launch_rle_ovf(o, c->stream));
    }
    i0 = i1;
  }
  return HG_OK;
}

// ---------------------------------------------------------------------------
// Static scratch of hg_build_noalloc: the only device memory that entry point
// uses beyond the caller's buffers (partitioned copy, level descriptors,
// count matrix, final-stage items and oversized-bucket counters).  Each
// device holds its own copy from module load to process end.  In host code
// `scratch` is only a name for the symbol: its device address on each device
// comes from hipGetSymbolAddress, taken once per device when a context is
// created with that device current, and kept in g_scratch.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = 40ull << 20;
static __device__ __align__(256) char scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
char *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

hg_status scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(HG_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return HG_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(HG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = (char *)p;
  return HG_OK;
}

uint64_t ref_binary_search(const std::vector<int64_t> &bins, uint64_t l, uint64_t r, uint64_t x) {
  // include/MultiHashGraphHostOperators.cuh:8-27 (recursive, size_t indices)
  for (;;) {
    if (r >= l && r != (uint64_t)-1) {
      const uint64_t mid = l + (r - l) / 2;
      if ((uint64_t)bins[mid] == x) return mid;
      if ((uint64_t)bins[mid] > x) {
        if (mid == 0) return l;
        r = mid - 1;
        continue;
      }
      l = mid + 1;
      continue;
    }
    return l;
  }
}

hg_status check_partitioned(hg_ctx c, const hg_table_desc *desc, const void *d_keys, const int64_t *h_splits,
                            uint32_t n_parts, void *d_offsets, void *d_csr_keys, uint64_t *h_part_offsets);
hg_status index_records(hg_ctx c, uint64_t n, uint64_t *&rin, uint64_t *&rout);
hg_status build_partitioned(hg_ctx c, const hg_table_desc *desc, const void *d_keys, const int64_t *h_splits,
                            uint32_t n_parts, void *d_offsets, void *d_csr_keys, uint64_t *h_part_offsets,
                            uint64_t *d_csr_index = nullptr, bool pack_in = false);
void splits_from_hist(const std::vector<unsigned long long> &h, uint64_t bin_count, uint64_t n, uint32_t n_parts,
                      uint64_t range, uint64_t M, int64_t *h_splits);

// Launches of the batched trainer: items sorted by sample size and split at
// 2048 keys, so that the many small samples run with small LDS (several
// workgroups per CU) and only the large ones take a CU each.  The items must
// be uploaded (in the returned order) to d_items.
std::vector<LaunchRmiBatch> rmi_launches(std::vector<RmiBatchItem> &items, int kb, uint32_t L) {
  std::stable_sort(items.begin(), items.end(),
                   [](const RmiBatchItem &x, const RmiBatchItem &y) { return x.np2 < y.np2; });
  std::vector<LaunchRmiBatch> out;
  size_t i = 0;
  while (i < items.size()) {
    const bool small = items[i].np2 <= 2048;
    size_t j = i;
    LaunchRmiBatch b{};
    b.key_bytes = kb;
    b.n_leaves = L;
    b.keys_cap = 16;
    b.cnt_cap = 4 * L;
    while (j < items.size() && (items[j].np2 <= 2048) == small) {
      b.keys_cap = std::max<uint32_t>(b.keys_cap, items[j].np2);
      b.cnt_cap = std::max<uint32_t>(b.cnt_cap, items[j].Mp);
      j++;
    }
    b.n_items = (uint32_t)(j - i);
    b.items = (const RmiBatchItem *)(uintptr_t)i;  // the item index until the upload
    out.push_back(b);
    i = j;
  }
  return out;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *hg_last_error(void) { return g_err.c_str(); }
const char *hg_version(void) { return "hashgraph_amd 0.1 (gfx950)"; }

hg_status hg_ctx_create(int device, hg_ctx *out) {
  if (!out) return fail(HG_ERR_INVALID_ARG, "null out");
  int n = 0;
  HG_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(HG_ERR_INVALID_ARG, "no such device");
  DeviceGuard g(device);
  hg_ctx c = new hg_ctx_s();
  c->device = device;
  for (int i = 0; i < hg_ctx_s::kSlots; i++) {
    if (hipEventCreateWithFlags(&c->hev[i], hipEventDisableTiming) != hipSuccess) {
      hg_ctx_destroy(c);
      return fail(HG_ERR_DEVICE, "hipEventCreate failed");
    }
  }
  if (hipHostMalloc((void **)&c->h_small, 1 << 20, hipHostMallocDefault) != hipSuccess) {
    hg_ctx_destroy(c);
    return fail(HG_ERR_DEVICE, "hipHostMalloc failed");
  }
  {
    const hg_status st = scratch_setup(device);  // device is current (DeviceGuard)
    if (st != HG_OK) {
      hg_ctx_destroy(c);
      return st;
    }
  }
  static bool attrs_done[64] = {false};
  if (device < 64 && !attrs_done[device]) {
    // best effort: a kernel whose limit cannot be raised reports its own
    // launch error if it ever needs more
    if (set_kernel_attributes() != hipSuccess) (void)hipGetLastError();
    attrs_done[device] = true;
  }
  *out = c;
  return HG_OK;
}

hg_status hg_ctx_destroy(hg_ctx c) {
  if (!c) return HG_OK;
  DeviceGuard g(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (int a = 0; a < 4; a++)
    if (c->arena[a]) (void)hipFree(c->arena[a]);
  for (int i = 0; i < hg_ctx_s::kSlots; i++) {
    if (c->hslot[i]) (void)hipHostFree(c->hslot[i]);
    if (c->hev[i]) (void)hipEventDestroy(c->hev[i]);
  }
  if (c->h_small) (void)hipHostFree(c->h_small);
  if (c->st_in) (void)hipStreamDestroy(c->st_in);
  if (c->st_out) (void)hipStreamDestroy(c->st_out);
  for (hipEvent_t e : c->st_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  delete c;
  return HG_OK;
}

hg_status hg_ctx_set_stream(hg_ctx c, void *stream) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  if ((hipStream_t)stream != c->stream) {
    // work queued on the old stream may still read the arena
    DeviceGuard g(c->device);
    HG_HIP(hipStreamSynchronize(c->stream));
    c->stream = (hipStream_t)stream;
  }
  return HG_OK;
}

hg_status hg_ctx_profile(hg_ctx c, int enable) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  c->prof = enable != 0;
  return HG_OK;
}

hg_status hg_ctx_profile_read(hg_ctx c, hg_kernel_time *out, uint32_t capacity, uint32_t *h_count, int reset) {
  if (!c || !h_count) return fail(HG_ERR_INVALID_ARG, "null argument");
  DeviceGuard g(c->device);
  HG_HIP(hipStreamSynchronize(c->stream));
  std::vector<hg_kernel_time> agg;
  for (const auto &r : c->recs) {
    float ms = 0;
    HG_HIP(hipEventElapsedTime(&ms, c->ev_pool[r.a], c->ev_pool[r.b]));
    size_t i = 0;
    while (i < agg.size() && strcmp(agg[i].name, r.name) != 0) i++;
    if (i == agg.size()) {
      hg_kernel_time t{};
      strncpy(t.name, r.name, sizeof(t.name) - 1);
      agg.push_back(t);
    }
    agg[i].launches++;
    agg[i].total_ms += ms;
  }
  *h_count = (uint32_t)agg.size();
  for (uint32_t i = 0; i < agg.size() && i < capacity && out; i++) out[i] = agg[i];
  if (reset) {
    c->recs.clear();
    c->ev_used = 0;
  }
  return HG_OK;
}

hg_status hg_ctx_reserve(hg_ctx c, size_t bytes) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  DeviceGuard g(c->device);
  HG_TRY(arena_ensure(c, 0, bytes / 2));
  return arena_ensure(c, 1, bytes - bytes / 2);
}

hg_status hg_scratch_bytes(const hg_table_desc *desc, int op, size_t *bytes) {
  HG_TRY(check_desc(desc));
  if (!bytes) return fail(HG_ERR_INVALID_ARG, "null bytes");
  const uint64_t n = desc->n_keys, M = desc->n_bins;
  const int kb = (int)desc->key_bytes;
  const uint32_t W = choose_width(n, M, kb);
  const uint64_t F = std::min<uint64_t>((M + W - 1) / W, kFMax);
  const uint64_t tiles = (n + kTile - 1) / kTile + 1;
  (void)op;
  *bytes = (size_t)n * kb * 2 + (size_t)(tiles * (F + 1) * 4) * 2 + (size_t)F * 64 + (size_t)tiles * 64 + (1 << 20);
  return HG_OK;
}

hg_status hg_build(hg_ctx c, const hg_table_desc *desc, const void *d_keys, void *d_offsets, void *d_csr_keys) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(desc));
  if (!d_offsets || (desc->n_keys && (!d_keys || !d_csr_keys))) return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  return build_table(c, desc, d_keys, d_offsets, d_csr_keys);
}

hg_status hg_build_noalloc(hg_ctx c, const hg_table_desc *desc, const void *d_keys, void *d_offsets, void *d_csr_keys) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(desc));
  if (!d_offsets || (desc->n_keys && (!d_keys || !d_csr_keys))) return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  char *const base = c->device >= 0 && c->device < kMaxDevices ? g_scratch[c->device] : nullptr;
  if (!base) return fail(HG_ERR_DEVICE, "hg_build_noalloc: no static scratch address for this device");
  const uint64_t n = desc->n_keys, M = desc->n_bins;
  const int kb = (int)desc->key_bytes, ob = (int)desc->offset_bytes;
  const BinMap map = make_map(M, desc->hash_kind, desc->d_rmi_model);
  if (n == 0) return fill_offsets(c, d_offsets, ob, M + 1, 0);
  const uint32_t W = choose_width(n, M, kb, build_fill());
  const uint64_t F = (M + W - 1) / W;
  const uint32_t tile_keys = Level::tile_keys_for(n);
  const uint64_t ntiles = (n + tile_keys - 1) / tile_keys;
  const uint64_t nch = (ntiles + kChunk - 1) / kChunk, ncb = (F + 255) / 256;
  const uint64_t big_cap = std::min<uint64_t>(M, 1ull << 20);
  // scratch layout (256-byte aligned pieces), checked against the static size
  size_t off = 0;
  auto take = [&](uint64_t bytes) {
    const size_t o = off;
    off += (size_t)((bytes + 255) & ~255ull);
    return o;
  };
  const size_t o_part = take(n * kb), o_tiles = take(ntiles * sizeof(TileDesc)), o_segs = take(sizeof(SegDesc)),
               o_items = take(ncb * nch * sizeof(ColItem)), o_items2 = take(ncb * sizeof(ColItem)), o_psb = take(8),
               o_nch = take(4), o_maps = take(sizeof(BinMap)), o_cmat = take(ntiles * F * 4),
               o_psum = take(nch * F * 4), o_coltot = take(F * 4), o_bstart = take(F * 4),
               o_fin = take(F * sizeof(FinItem)), o_bslot = take(F * 4), o_bctl = take(final_big_ctl_bytes()),
               o_bcnt = take(big_cap * 4);
  if (F > kFMax || off > kNoAllocBytes)
    return fail(HG_ERR_UNSUPPORTED, "hg_build_noalloc: the table needs " + std::to_string(off) +
                                        " bytes of scratch, the static area holds " + std::to_string(kNoAllocBytes) +
                                        " (up to 2^22 keys of either width and 2^28 bins fit)");
  SegDesc sd = range_seg(0, n, 0, M, W, map);
  sd.fin_off_base = 0;
  sd.fin_key_rel = 0;
  sd.fin_end_pos = M;
  sd.table = 0;
  sd.tile_begin = 0;
  sd.ntiles = (uint32_t)ntiles;
  sd.row_base = 0;
  sd.bucket_base = 0;
  LevelDev d;
  d.tiles = (TileDesc *)(base + o_tiles);
  d.segs = (SegDesc *)(base + o_segs);
  d.items = (ColItem *)(base + o_items);
  d.items2 = (ColItem *)(base + o_items2);
  d.psum_base = (uint64_t *)(base + o_psb);
  d.nchunks = (uint32_t *)(base + o_nch);
  d.maps = (BinMap *)(base + o_maps);
  d.splits = nullptr;
  d.cmat = (uint32_t *)(base + o_cmat);
  d.psum = (uint32_t *)(base + o_psum);
  d.coltot = (uint32_t *)(base + o_coltot);
  d.bstart = (uint32_t *)(base + o_bstart);
  d.fin = (FinItem *)(base + o_fin);
  d.big_slot = (uint32_t *)(base + o_bslot);
  d.big_ctl = base + o_bctl;
  d.big_cnt = (uint32_t *)(base + o_bcnt);
  void *part = base + o_part;
  HG_HIP(launch_level_setup(sd, map, tile_keys, kChunk, d.tiles, d.segs, d.items, d.items2, d.psum_base, d.nchunks,
                            d.maps, c->stream));
  LaunchPart p{};
  p.key_bytes = kb;
  p.hk = (int)map.kind;
  p.sk = 0;
  p.ntiles = (uint32_t)ntiles;
  p.max_cols = (uint32_t)F;
  p.stage_cap = part_stage_cap((uint32_t)F, kb, 0);
  if (p.stage_cap < 1024) return fail(HG_ERR_UNSUPPORTED, "partition level too wide for LDS");
  p.keys = d_keys;
  p.out = part;
  p.tiles = d.tiles;
  p.segs = d.segs;
  p.cmat = d.cmat;
  HG_HIP(launch_part_hist(p, c->stream));
  LaunchCol lc{};
  lc.nitems = (uint32_t)(ncb * nch);
  lc.nitems2 = (uint32_t)ncb;
  lc.nsegs = 1;
  lc.cmat = d.cmat;
  lc.items = d.items;
  lc.items2 = d.items2;
  lc.segs = d.segs;
  lc.psum_base = d.psum_base;
  lc.nchunks = d.nchunks;
  lc.psum = d.psum;
  lc.coltot = d.coltot;
  lc.bstart = d.bstart;
  HG_HIP(launch_colscan(lc, c->stream));
  HG_HIP(launch_part_scatter(p, c->stream));
  HG_HIP(launch_make_items(d.segs, 1, d.bstart, d.coltot, d.fin, (uint32_t)F, c->stream));
  LaunchFinal f{};
  f.key_bytes = kb;
  f.off_bytes = ob;
  f.hk = (int)map.kind;
  f.nitems = (uint32_t)F;
  f.lds_bytes = 128 * 1024;
  f.part = part;
  f.csr = d_csr_keys;
  f.off = d_offsets;
  f.items = d.fin;
  f.maps = d.maps;
  f.big_slot = d.big_slot;
  f.big_ctl = d.big_ctl;
  f.big_cnt = d.big_cnt;
  f.big_cap = (uint32_t)big_cap;
  HG_HIP(launch_final(f, c->stream));
  return HG_OK;
}

hg_status hg_table_bins(hg_ctx c, const hg_table_desc *desc, const void *d_keys, uint64_t n, uint32_t *d_bins) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(desc));
  if (n && (!d_keys || !d_bins)) return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  HG_HIP(launch_bins(d_keys, n, (int)desc->key_bytes, make_map(desc->n_bins, desc->hash_kind, desc->d_rmi_model), d_bins,
                     c->stream));
  return HG_OK;
}

size_t hg_rmi_model_bytes(uint32_t n_leaves) {
  if (n_leaves == 0 || n_leaves > 65536) return 0;
  return sizeof(RmiHeader) + (size_t)n_leaves * sizeof(RmiLeaf);
}

hg_status hg_rmi_train(hg_ctx c, const void *d_keys, uint64_t n_keys, uint32_t key_bytes, uint64_t n_sample,
                       uint64_t seed, uint64_t n_bins, uint32_t n_leaves, void *d_model, uint32_t *h_use_rmi,
                       hg_rmi_report *h_report) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  if (key_bytes != 4 && key_bytes != 8) return fail(HG_ERR_INVALID_ARG, "key_bytes must be 4 or 8");
  if (n_bins == 0 || n_bins >= (1ull << 32)) return fail(HG_ERR_INVALID_ARG, "n_bins must be in [1, 2^32)");
  if (n_leaves == 0 || n_leaves > 65536) return fail(HG_ERR_INVALID_ARG, "n_leaves must be in [1, 65536]");
  if (n_keys == 0 || n_sample == 0 || n_sample >= (1ull << 31))
    return fail(HG_ERR_INVALID_ARG, "need n_keys >= 1 and n_sample in [1, 2^31)");
  if (!d_keys || !d_model || !h_use_rmi) return fail(HG_ERR_INVALID_ARG, "null argument");
  DeviceGuard g(c->device);
  const uint32_t ntiles = rmi_sort_tiles(n_sample);
  const uint32_t nblk = rmi_moment_blocks(n_sample);
  // sample bins at the table's load factor
  long double mp = (long double)n_bins * (long double)n_sample / (long double)n_keys;
  uint64_t Mp = (uint64_t)(mp + 0.5L);
  Mp = std::max<uint64_t>(1, std::min<uint64_t>(Mp, n_bins));
  Plan P;
  const size_t i_b0 = P.add(n_sample * key_bytes), i_b1 = P.add(n_sample * key_bytes);
  const size_t i_cnt = P.add(((size_t)256 * ntiles + 1) * 4);
  const size_t i_st = P.add(scan_status_bytes((uint64_t)256 * ntiles));
  const size_t i_part = P.add((size_t)nblk * 8 * sizeof(double));
  const size_t i_cand = P.add(64), i_lcnt = P.add((size_t)4 * n_leaves * 4);
  const size_t i_start = P.add(((size_t)n_leaves + 1) * 4);
  const size_t i_cr = P.add(Mp * 4), i_cm = P.add(Mp * 4), i_out = P.add(64);
  HG_TRY(arena_ensure(c, 1, P.total()));
  Carve cv;
  cv.init(c, 1, P);
  LaunchRmi r{};
  r.key_bytes = (int)key_bytes;
  r.keys = d_keys;
  r.n_keys = n_keys;
  r.n_sample = n_sample;
  r.s0 = splitmix_mix(seed);
  r.buf0 = cv.get<void>(i_b0);
  r.buf1 = cv.get<void>(i_b1);
  r.counts = cv.get<uint32_t>(i_cnt);
  r.scan_status = cv.get<void>(i_st);
  r.partials = cv.get<double>(i_part);
  r.cand = cv.get<void>(i_cand);
  r.leafcnt = cv.get<uint32_t>(i_lcnt);
  r.nblk = nblk;
  r.n_leaves = n_leaves;
  r.map = make_map(n_bins, 0, nullptr);
  r.model = d_model;
  r.start = cv.get<uint32_t>(i_start);
  r.Mp = (uint32_t)Mp;
  r.cnt_r = cv.get<uint32_t>(i_cr);
  r.cnt_m = cv.get<uint32_t>(i_cm);
  r.out = cv.get<unsigned long long>(i_out);
  {
    Prof pf(c, "rmi_train");
    HG_HIP(launch_rmi_train(r, c->stream));
  }
  unsigned long long res[4];
  HG_TRY(sync_read(c, res, r.out, sizeof(res)));
  *h_use_rmi = (res[0] <= res[2] && res[1] <= res[3]) ? 1u : 0u;
  if (h_report) {
    h_report->sample_bins = Mp;
    h_report->max_bin_rmi = res[0];
    h_report->max_bin_murmur = res[2];
    h_report->mean_probe_rmi = (double)res[1] / (double)n_sample;
    h_report->mean_probe_murmur = (double)res[3] / (double)n_sample;
  }
  return HG_OK;
}

hg_status hg_rmi_train_batch(hg_ctx c, uint32_t n_tables, uint32_t key_bytes, const void *d_keys,
                             const uint64_t *h_key_off, const uint64_t *h_n_bins, const uint64_t *h_n_sample,
                             const uint64_t *h_seed, uint32_t n_leaves, void *d_models, uint32_t *h_use,
                             hg_rmi_report *h_report) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  if (key_bytes != 4 && key_bytes != 8) return fail(HG_ERR_INVALID_ARG, "key_bytes must be 4 or 8");
  if (n_leaves == 0 || n_leaves > 65536) return fail(HG_ERR_INVALID_ARG, "n_leaves must be in [1, 65536]");
  if (!n_tables) return HG_OK;
  if (!h_key_off || !h_n_bins || !h_n_sample || !h_seed || !d_models || !h_use || !d_keys)
    return fail(HG_ERR_INVALID_ARG, "null argument");
  for (uint32_t t = 0; t < n_tables; t++) {
    if (h_key_off[t + 1] < h_key_off[t]) return fail(HG_ERR_INVALID_ARG, "key offsets must not decrease");
    if (h_n_bins[t] == 0 || h_n_bins[t] >= (1ull << 32)) return fail(HG_ERR_INVALID_ARG, "n_bins must be in [1, 2^32)");
    if (h_n_sample[t] == 0 || h_n_sample[t] >= (1ull << 31))
      return fail(HG_ERR_INVALID_ARG, "n_sample must be in [1, 2^31)");
  }
  DeviceGuard g(c->device);
  const size_t mb = hg_rmi_model_bytes(n_leaves);
  const uint32_t kcap = RMI_ONE_MAX_KEY_BYTES / key_bytes;
  std::vector<RmiBatchItem> items;
  std::vector<uint32_t> big;  // tables for the multi-kernel trainer
  for (uint32_t t = 0; t < n_tables; t++) {
    h_use[t] = 0;
    if (h_report) h_report[t] = hg_rmi_report{};
    const uint64_t n = h_key_off[t + 1] - h_key_off[t], ns = h_n_sample[t];
    if (n == 0) continue;  // no keys: murmur3 (the model is not written)
    long double mp = (long double)h_n_bins[t] * (long double)ns / (long double)n;
    uint64_t Mp = (uint64_t)(mp + 0.5L);
    Mp = std::max<uint64_t>(1, std::min<uint64_t>(Mp, h_n_bins[t]));
    uint64_t np2 = 16;
    while (np2 < ns) np2 <<= 1;
    if (np2 > kcap || Mp > RMI_ONE_MAX_CHECK_BINS || n_leaves > RMI_ONE_MAX_LEAVES) {
      big.push_back(t);
      continue;
    }
    RmiBatchItem it{};
    it.key_begin = h_key_off[t];
    it.n_keys = n;
    it.s0 = splitmix_mix(h_seed[t]);
    it.model_off = (uint64_t)t * mb;
    it.map = make_map(h_n_bins[t], 0, nullptr);
    it.ns = (uint32_t)ns;
    it.np2 = (uint32_t)np2;
    it.Mp = (uint32_t)Mp;
    it.table = t;
    items.push_back(it);
  }
  if (!items.empty()) {
    std::vector<LaunchRmiBatch> lb = rmi_launches(items, (int)key_bytes, n_leaves);
    Blob bl;
    const size_t o_items = bl.put(items);
    Plan P;
    const size_t i_blob = P.add(bl.data.size()), i_use = P.add((size_t)n_tables * 4),
                 i_rep = P.add((size_t)n_tables * 32);
    HG_TRY(arena_ensure(c, 2, P.total()));
    Carve cv;
    cv.init(c, 2, P);
    char *blob = cv.get<char>(i_blob);
    HG_TRY(upload(c, blob, bl.data.data(), bl.data.size()));
    uint32_t *d_use = cv.get<uint32_t>(i_use);
    unsigned long long *d_rep = cv.get<unsigned long long>(i_rep);
    for (LaunchRmiBatch &b : lb) {
      b.keys = d_keys;
      b.items = (const RmiBatchItem *)(blob + o_items) + (uintptr_t)b.items;
      b.models = d_models;
      b.use = d_use;
      b.report = d_rep;
      Prof pf(c, "rmi_train");
      HG_HIP(launch_rmi_train_batch(b, c->stream));
    }
    std::vector<uint32_t> use(n_tables);
    std::vector<unsigned long long> rep((size_t)n_tables * 4);
    HG_TRY(sync_read(c, use.data(), d_use, (size_t)n_tables * 4));
    if (h_report) HG_TRY(sync_read(c, rep.data(), d_rep, (size_t)n_tables * 32));
    for (const RmiBatchItem &it : items) {
      h_use[it.table] = use[it.table];
      if (h_report) {
        hg_rmi_report &r = h_report[it.table];
        r.sample_bins = it.Mp;
        r.max_bin_rmi = rep[(size_t)it.table * 4];
        r.max_bin_murmur = rep[(size_t)it.table * 4 + 2];
        r.mean_probe_rmi = (double)rep[(size_t)it.table * 4 + 1] / (double)it.ns;
        r.mean_probe_murmur = (double)rep[(size_t)it.table * 4 + 3] / (double)it.ns;
      }
    }
  }
  // samples beyond one workgroup's LDS: the multi-kernel trainer, table by table
  for (uint32_t t : big)
    HG_TRY(hg_rmi_train(c, (const char *)d_keys + h_key_off[t] * key_bytes, h_key_off[t + 1] - h_key_off[t], key_bytes,
                        h_n_sample[t], h_seed[t], h_n_bins[t], n_leaves, (char *)d_models + (size_t)t * mb, &h_use[t],
                        h_report ? &h_report[t] : nullptr));
  return HG_OK;
}

hg_status hg_hash_bins(hg_ctx c, const void *d_keys, uint64_t n, uint32_t key_bytes, uint64_t n_bins, uint32_t *d_bins) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  if (key_bytes != 4 && key_bytes != 8) return fail(HG_ERR_INVALID_ARG, "key_bytes must be 4 or 8");
  if (n_bins == 0 || n_bins >= (1ull << 32)) return fail(HG_ERR_INVALID_ARG, "n_bins must be in [1, 2^32)");
  if (n && (!d_keys || !d_bins)) return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  HG_HIP(launch_bins(d_keys, n, (int)key_bytes, make_map(n_bins, 0, nullptr), d_bins, c->stream));
  return HG_OK;
}

hg_status hg_exclusive_scan(hg_ctx c, const void *d_in, void *d_out, uint64_t n, uint32_t elem_bytes) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  if (elem_bytes != 4 && elem_bytes != 8) return fail(HG_ERR_INVALID_ARG, "elem_bytes must be 4 or 8");
  if (!d_out || (n && !d_in)) return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  if (n == 0) return fill_offsets(c, d_out, (int)elem_bytes, 1, 0);
  Plan P;
  const size_t i_status = P.add(scan_status_bytes(n));
  HG_TRY(arena_ensure(c, 1, P.total()));
  Carve cv;
  cv.init(c, 1, P);
  HG_HIP(launch_scan(d_in, d_out, n, (int)elem_bytes, cv.get<void>(i_status), c->stream));
  return HG_OK;
}

hg_status hg_probe(hg_ctx c, const hg_table_desc *descA, const void *d_offA, const void *d_keysA, const void *d_probe,
                   uint64_t n_probe, void *d_out, uint64_t capacity, uint64_t *h_count) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(descA));
  if (!h_count) return fail(HG_ERR_INVALID_ARG, "null h_count");
  if (!d_offA || (descA->n_keys && !d_keysA) || (n_probe && !d_probe)) return fail(HG_ERR_INVALID_ARG, "null buffer");
  if (n_probe >= (1ull << 32)) return fail(HG_ERR_UNSUPPORTED, "n_probe >= 2^32 is not supported by this build");
  DeviceGuard g(c->device);
  *h_count = 0;
  if (n_probe == 0 || descA->n_keys == 0) return HG_OK;
  const uint64_t M = descA->n_bins;
  const int kb = (int)descA->key_bytes, ob = (int)descA->offset_bytes;
  const BinMap map = make_map(M, descA->hash_kind, descA->d_rmi_model);
  const uint32_t W = choose_width(std::max<uint64_t>(n_probe, descA->n_keys), M, kb, probe_fill(kb));
  const uint64_t F = (M + W - 1) / W;
  // Beyond kFMax buckets, B is partitioned twice: into super-buckets of k
  // buckets, whose sizes are read back, then each super-bucket into its
  // buckets (as the build's second level).
  const uint64_t k = level1_group(F);
  const bool two = k > 1;
  if ((uint64_t)W * k >= (1ull << 32)) return fail(HG_ERR_UNSUPPORTED, "probe geometry too large");
  Plan P0;
  const size_t i_part = P0.add(n_probe * kb);
  const size_t i_part2 = P0.add(two ? n_probe * kb : 1);
  const size_t i_small = P0.add(256);
  const size_t i_mark = P0.add((size_t)(F + kFMax + 1) * 4);  // >= the buckets of the last level
  HG_TRY(arena_ensure(c, 0, P0.total()));
  Carve cv0;
  cv0.init(c, 0, P0);
  void *part = cv0.get<void>(i_part);
  unsigned long long *small = cv0.get<unsigned long long>(i_small);
  unsigned int *hmark = cv0.get<unsigned int>(i_mark);
  Level L;
  LevelDev d;
  if (!two) {
    SegDesc s = range_seg(0, n_probe, 0, M, W, map);
    s.fin_off_base = 0;
    s.fin_key_rel = 0;
    s.table = 0;
    L.segs.push_back(s);
    L.maps.push_back(map);
    L.auto_tile(n_probe);
    L.prepare();
    HG_TRY(setup_level(c, L, d));
    HG_TRY(run_level(c, L, d, kb, d_probe, part));
  } else {
    const uint64_t W1 = (uint64_t)W * k;
    Level L1;
    L1.segs.push_back(range_seg(0, n_probe, 0, M, (uint32_t)W1, map));
    L1.maps.push_back(map);
    L1.auto_tile(n_probe);
    L1.prepare();
    LevelDev d1;
    HG_TRY(setup_level(c, L1, d1));
    HG_TRY(run_level(c, L1, d1, kb, d_probe, part));
    std::vector<uint32_t> cnt(L1.ncols_total), st(L1.ncols_total);
    HG_TRY(sync_read(c, cnt.data(), d1.coltot, cnt.size() * 4));
    HG_TRY(sync_read(c, st.data(), d1.bstart, st.size() * 4));
    for (uint32_t g = 0; g < L1.ncols_total; g++) {
      const uint64_t lo = (uint64_t)g * W1, nb = std::min<uint64_t>(W1, M - lo);
      SegDesc s = range_seg(st[g], (uint64_t)st[g] + cnt[g], lo, nb, W, map);
      s.fin_off_base = lo;
      s.fin_key_rel = 0;
      s.table = 0;
      L.segs.push_back(s);
    }
    L.maps.push_back(map);
    L.auto_tile(n_probe);
    L.prepare();
    HG_TRY(setup_level(c, L, d));
    void *part2 = cv0.get<void>(i_part2);
    HG_TRY(run_level(c, L, d, kb, part, part2));
    part = part2;
  }
  HG_HIP(launch_fill_u64(small, 0ull, 4, c->stream));
  HG_HIP(launch_fill_u32(hmark, 0u, L.ncols_total, c->stream));
  HG_HIP(launch_make_items(d.segs, (uint32_t)L.segs.size(), d.bstart, d.coltot, d.fin, L.ncols_total, c->stream));
  LaunchProbe pr{};
  pr.key_bytes = kb;
  pr.off_bytes = ob;
  pr.hk = L.hash_kind();
  pr.nitems = L.ncols_total;
  pr.akey_cap = probe_akey_cap(kb, ob);
  pr.bpart = part;
  pr.items = d.fin;
  pr.maps = d.maps;
  pr.offA = d_offA;
  pr.keysA = d_keysA;
  pr.out = d_out;
  pr.capacity = d_out ? capacity : 0;
  pr.cursor = small;
  pr.gate = (unsigned int *)(small + 2);
  pr.hmark = hmark;
  pr.big_slot = d.big_slot;
  pr.big_ctl = d.big_ctl;
  {
    Prof pf(c, "probe");
    HG_HIP(launch_probe(pr, c->stream));
  }
  unsigned long long cnt = 0;
  HG_TRY(sync_read(c, &cnt, small, 8));
  *h_count = cnt;
  if (d_out && cnt > capacity) return fail(HG_ERR_CAPACITY, "output capacity smaller than the match count");
  return HG_OK;
}

hg_status hg_intersect(hg_ctx c, const hg_table_desc *descA, const void *d_offA, const void *d_keysA,
                       const hg_table_desc *descB, const void *d_offB, const void *d_keysB, void *d_out,
                       uint64_t capacity, uint64_t *h_count) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(descA));
  HG_TRY(check_desc(descB));
  if (!h_count) return fail(HG_ERR_INVALID_ARG, "null h_count");
  if (descA->n_bins != descB->n_bins || descA->key_bytes != descB->key_bytes || descA->hash_kind != descB->hash_kind ||
      descA->offset_bytes != descB->offset_bytes)
    return fail(HG_ERR_MISMATCH, "tables do not share a bin space");
  if (!d_offA || !d_offB || (descA->n_keys && !d_keysA) || (descB->n_keys && !d_keysB))
    return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  *h_count = 0;
  // bins above the nested loop's pair limit hold more than 128 keys of the
  // two tables together, so at most (N_A + N_B) / 128 of them exist
  const uint64_t hcap = std::min<uint64_t>(descA->n_bins, (descA->n_keys + descB->n_keys) / 128 + 1);
  Plan P;
  const size_t i_small = P.add(256);
  const size_t i_hlist = P.add(hcap * 4);
  HG_TRY(arena_ensure(c, 1, P.total()));
  Carve cv;
  cv.init(c, 1, P);
  unsigned long long *small = cv.get<unsigned long long>(i_small);
  HG_HIP(launch_fill_u64(small, 0ull, 4, c->stream));
  LaunchJoin j{};
  j.key_bytes = (int)descA->key_bytes;
  j.off_bytes = (int)descA->offset_bytes;
  j.M = descA->n_bins;
  j.key_cap = join_key_cap(j.key_bytes, j.off_bytes, d_out != nullptr);
  j.offA = d_offA;
  j.keysA = d_keysA;
  j.offB = d_offB;
  j.keysB = d_keysB;
  j.out = d_out;
  j.capacity = d_out ? capacity : 0;
  j.cursor = small;
  j.hlist = cv.get<uint32_t>(i_hlist);
  j.hlist_n = (unsigned int *)(small + 1);
  j.hlist_cap = (uint32_t)hcap;
  {
    Prof pf(c, "join");
    HG_HIP(launch_join(j, c->stream));
  }
  unsigned long long res[2];
  HG_TRY(sync_read(c, res, small, 16));
  *h_count = res[0];
  if ((res[1] & 0xFFFFFFFFull) > hcap) return fail(HG_ERR_DEVICE, "heavy join bins exceed their bound");
  if (d_out && res[0] > capacity) return fail(HG_ERR_CAPACITY, "output capacity smaller than the match count");
  return HG_OK;
}

hg_status hg_intersect_indexed(hg_ctx c, const hg_table_desc *descA, const void *d_offA, const void *d_keysA,
                               const uint64_t *d_indA, const hg_table_desc *descB, const void *d_offB,
                               const void *d_keysB, const uint64_t *d_indB, uint64_t *d_pos_a, uint64_t *d_pos_b,
                               uint64_t capacity, uint64_t *h_count) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  if (!h_count) return fail(HG_ERR_INVALID_ARG, "null h_count");
  if ((d_pos_a == nullptr) != (d_pos_b == nullptr)) return fail(HG_ERR_INVALID_ARG, "give both d_pos_a and d_pos_b");
  if (d_pos_a && ((descA && descA->n_keys && !d_indA) || (descB && descB->n_keys && !d_indB)))
    return fail(HG_ERR_INVALID_ARG, "null input positions");
  // count only: the key join's count.  With an output, the pair pass counts
  // as it writes (places past capacity are dropped), and its cursor is the count.
  if (!d_pos_a) return hg_intersect(c, descA, d_offA, d_keysA, descB, d_offB, d_keysB, nullptr, 0, h_count);
  HG_TRY(check_desc(descA));
  HG_TRY(check_desc(descB));
  if (descA->n_bins != descB->n_bins || descA->key_bytes != descB->key_bytes || descA->hash_kind != descB->hash_kind ||
      descA->offset_bytes != descB->offset_bytes)
    return fail(HG_ERR_MISMATCH, "tables do not share a bin space");
  if (!d_offA || !d_offB || (descA->n_keys && !d_keysA) || (descB->n_keys && !d_keysB))
    return fail(HG_ERR_INVALID_ARG, "null buffer");
  *h_count = 0;
  DeviceGuard g(c->device);
  const uint64_t hcap = std::min<uint64_t>(descA->n_bins, (descA->n_keys + descB->n_keys) / 128 + 1);
  Plan P;
  const size_t i_small = P.add(256), i_hlist = P.add(hcap * 4), i_tiles = P.add((hcap + 1) * 8),
               i_tbase = P.add((hcap + 1) * 8), i_status = P.add(scan_status_bytes(hcap));
  HG_TRY(arena_ensure(c, 1, P.total()));
  Carve cv;
  cv.init(c, 1, P);
  unsigned long long *small = cv.get<unsigned long long>(i_small);  // [cursor, listed bins, pairs]
  HG_HIP(launch_fill_u64(small, 0ull, 4, c->stream));
  LaunchJoin j{};
  j.key_bytes = (int)descA->key_bytes;
  j.off_bytes = (int)descA->offset_bytes;
  j.M = descA->n_bins;
  j.key_cap = join_key_cap(j.key_bytes, j.off_bytes, true, true);
  j.offA = d_offA;
  j.keysA = d_keysA;
  j.offB = d_offB;
  j.keysB = d_keysB;
  j.out = nullptr;
  j.capacity = capacity;
  j.cursor = small;
  j.hlist = cv.get<uint32_t>(i_hlist);
  j.hlist_n = (unsigned int *)(small + 1);
  j.hlist_cap = (uint32_t)hcap;
  j.indA = d_indA;
  j.indB = d_indB;
  j.pa = d_pos_a;
  j.pb = d_pos_b;
  {
    Prof pf(c, "join_pairs");
    HG_HIP(launch_join(j, c->stream));
  }
  unsigned long long res[2];
  HG_TRY(sync_read(c, res, small, 16));
  const uint64_t nh = res[1] & 0xFFFFFFFFull;
  if (nh > hcap) return fail(HG_ERR_DEVICE, "heavy join bins exceed their bound");
  if (nh) {
    unsigned long long *tiles = cv.get<unsigned long long>(i_tiles), *tbase = cv.get<unsigned long long>(i_tbase);
    HG_HIP(launch_pair_tiles(d_offA, d_offB, j.off_bytes, j.hlist, (uint32_t)nh, tiles, small + 2, c->stream));
    HG_HIP(launch_scan(tiles, (void *)tbase, nh, 8, cv.get<void>(i_status), c->stream));
    unsigned long long pairs = 0, ntiles = 0;
    HG_TRY(sync_read(c, &pairs, small + 2, 8));
    HG_TRY(sync_read(c, &ntiles, tbase + nh, 8));
    if (pairs > kPairCompareLimit) {
      HG_TRY(hg_intersect(c, descA, d_offA, d_keysA, descB, d_offB, d_keysB, nullptr, 0, h_count));
      return fail(HG_ERR_UNSUPPORTED, "bins with more than 4096 key pairs hold " + std::to_string(pairs) +
                                          " key pairs together; the row-id join compares them one by one and "
                                          "takes at most 2^36");
    }
    LaunchJoinPairs p{};
    p.key_bytes = j.key_bytes;
    p.off_bytes = j.off_bytes;
    p.offA = d_offA;
    p.keysA = d_keysA;
    p.offB = d_offB;
    p.keysB = d_keysB;
    p.indA = d_indA;
    p.indB = d_indB;
    p.hlist = j.hlist;
    p.nh = (uint32_t)nh;
    p.tile_base = tbase;
    p.ntiles = ntiles;
    p.pa = d_pos_a;
    p.pb = d_pos_b;
    p.capacity = capacity;
    p.cursor = small;
    Prof pf(c, "join_pairs_heavy");
    HG_HIP(launch_join_pairs_heavy(p, c->stream));
  }
  unsigned long long total = 0;
  HG_TRY(sync_read(c, &total, small, 8));
  *h_count = total;
  if (total > capacity) return fail(HG_ERR_CAPACITY, "output capacity smaller than the match count");
  return HG_OK;
}

hg_status hg_intersect_rle(hg_ctx c, const hg_table_desc *descA, const void *d_offA, const void *d_keysA,
                           const hg_table_desc *descB, const void *d_offB, const void *d_keysB, void *d_rec_keys,
                           uint64_t *d_rec_ca, uint64_t *d_rec_cb, uint64_t capacity, uint64_t *h_records,
                           uint64_t *h_count) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(descA));
  HG_TRY(check_desc(descB));
  if (!h_records || !h_count) return fail(HG_ERR_INVALID_ARG, "null h_records / h_count");
  if (descA->n_bins != descB->n_bins || descA->key_bytes != descB->key_bytes || descA->hash_kind != descB->hash_kind ||
      descA->offset_bytes != descB->offset_bytes)
    return fail(HG_ERR_MISMATCH, "tables do not share a bin space");
  if (!d_offA || !d_offB || (descA->n_keys && !d_keysA) || (descB->n_keys && !d_keysB))
    return fail(HG_ERR_INVALID_ARG, "null buffer");
  if (capacity && (!d_rec_keys || !d_rec_ca || !d_rec_cb)) return fail(HG_ERR_INVALID_ARG, "null record buffers");
  DeviceGuard g(c->device);
  *h_records = 0;
  *h_count = 0;
  const uint64_t M = descA->n_bins, nch = rle_chunks(M);
  // a heavy bin holds more than 64 keys on a side, so at most
  // N_A / 65 + N_B / 65 of them exist
  const uint64_t heavy_cap = std::min<uint64_t>(M, descA->n_keys / 65 + descB->n_keys / 65 + 1);
  Plan P;
  const size_t i_cnt = P.add((nch + 1) * 8), i_base = P.add((nch + 1) * 8), i_hcnt = P.add((nch + 1) * 4),
               i_hbase = P.add((nch + 1) * 4), i_pairs = P.add((nch + 1) * 8), i_small = P.add(256),
               i_heavy = P.add(heavy_cap * 4), i_ovf = P.add(heavy_cap * 4), i_status = P.add(scan_status_bytes(nch)),
               i_gmark = P.add(heavy_cap * 8), i_giant = P.add(rle_giant_bytes((int)descA->key_bytes)),
               i_pkey = P.add(rle_priv_records() * descA->key_bytes), i_pca = P.add(rle_priv_records() * 8),
               i_pcb = P.add(rle_priv_records() * 8), i_pn = P.add((size_t)rle_heavy_grid() * 4),
               i_caprec = P.add((nch + 1) * 8), i_pbase = P.add((nch + 1) * 8), i_caph = P.add((nch + 1) * 4),
               i_phbase = P.add((nch + 1) * 4);
  HG_TRY(arena_ensure(c, 1, P.total()));
  Carve cv;
  cv.init(c, 1, P);
  unsigned long long *small = cv.get<unsigned long long>(i_small);  // [-, pairs, cursor, overflow, ovf_n]
  HG_HIP(launch_fill_u64(small, 0ull, 8, c->stream));
  LaunchRle r{};
  r.key_bytes = (int)descA->key_bytes;
  r.off_bytes = (int)descA->offset_bytes;
  r.M = M;
  r.offA = d_offA;
  r.keysA = d_keysA;
  r.offB = d_offB;
  r.keysB = d_keysB;
  r.chunk_cnt = cv.get<unsigned long long>(i_cnt);
  r.chunk_base = cv.get<unsigned long long>(i_base);
  r.chunk_heavy = cv.get<uint32_t>(i_hcnt);
  r.heavy_base = cv.get<uint32_t>(i_hbase);
  r.chunk_pairs = cv.get<unsigned long long>(i_pairs);
  r.pkey = cv.get<char>(i_pkey);
  r.pca = cv.get<unsigned long long>(i_pca);
  r.pcb = cv.get<unsigned long long>(i_pcb);
  r.priv_n = cv.get<uint32_t>(i_pn);
  r.rk = d_rec_keys;
  r.rca = (unsigned long long *)d_rec_ca;
  r.rcb = (unsigned long long *)d_rec_cb;
  r.capacity = capacity;
  r.heavy = cv.get<uint32_t>(i_heavy);
  r.heavy_cap = heavy_cap;
  r.pairs = small + 1;
  r.cursor = small + 2;
  r.overflow = (unsigned int *)(small + 3);
  r.ovf_list = cv.get<uint32_t>(i_ovf);
  r.ovf_n = (unsigned int *)(small + 4);
  {
    // giant-bin scratch: control block, then keys, counts of A, counts of B
    char *gb = cv.get<char>(i_giant);
    const size_t ctl_bytes = (rle_giant_ctl_bytes() + 255) & ~(size_t)255;
    r.gmark = cv.get<uint32_t>(i_gmark);
    r.gctl = gb;
    r.gkey = gb + ctl_bytes;
    r.gca = (unsigned long long *)(gb + ctl_bytes + (((size_t)rle_giant_slots() * descA->key_bytes + 255) & ~(size_t)255));
    r.gcb = r.gca + rle_giant_slots();
  }
  void *status = cv.get<void>(i_status);
  unsigned long long nsmall = 0;
  uint32_t nheavy = 0;
  // With an output, one pass of the small-bin kernel counts and writes the
  // records and heavy bins at padded places bounded from the offsets alone,
  // and a gather packs them; without one (or when the padded records would
  // take more than kRleOnePassBytes) a count pass sizes the output first.
  bool one_pass = false;
  unsigned long long *pbase = cv.get<unsigned long long>(i_pbase);
  uint32_t *phbase = cv.get<uint32_t>(i_phbase);
  unsigned long long s_rec = 0;
  uint32_t s_heavy = 0;
  if (capacity) {
    HG_HIP(launch_rle_caps(d_offA, d_offB, r.off_bytes, M, cv.get<unsigned long long>(i_caprec),
                           cv.get<uint32_t>(i_caph), c->stream));
    HG_HIP(launch_scan(cv.get<unsigned long long>(i_caprec), (void *)pbase, nch, 8, status, c->stream));
    HG_HIP(launch_scan(cv.get<uint32_t>(i_caph), (void *)phbase, nch, 4, status, c->stream));
    HG_TRY(sync_read(c, &s_rec, pbase + nch, 8));
    HG_TRY(sync_read(c, &s_heavy, phbase + nch, 4));
    one_pass = (double)s_rec * (descA->key_bytes + 16) + (double)s_heavy * 4 <= (double)kRleOnePassBytes;
  }
  if (one_pass) {
    Plan P0;
    const size_t j_rk = P0.add(std::max<uint64_t>(s_rec, 1) * descA->key_bytes),
                 j_ca = P0.add(std::max<uint64_t>(s_rec, 1) * 8), j_cb = P0.add(std::max<uint64_t>(s_rec, 1) * 8),
                 j_h = P0.add(((uint64_t)s_heavy + 1) * 4);
    HG_TRY(arena_ensure(c, 0, P0.total()));
    Carve cv0;
    cv0.init(c, 0, P0);
    LaunchRle r3 = r;
    r3.pass = 3;
    r3.chunk_base = pbase;
    r3.heavy_base = phbase;
    r3.rk = cv0.get<void>(j_rk);
    r3.rca = cv0.get<unsigned long long>(j_ca);
    r3.rcb = cv0.get<unsigned long long>(j_cb);
    r3.capacity = s_rec;
    r3.heavy = cv0.get<uint32_t>(j_h);
    r3.heavy_cap = s_heavy;
    {
      Prof pf(c, "rle_small");
      HG_HIP(launch_rle(r3, c->stream));
    }
    HG_HIP(launch_scan(r.chunk_cnt, (void *)r.chunk_base, nch, 8, status, c->stream));
    HG_HIP(launch_scan(r.chunk_heavy, (void *)r.heavy_base, nch, 4, status, c->stream));
    LaunchRleGather gt{};
    gt.key_bytes = r.key_bytes;
    gt.nch = (uint32_t)nch;
    gt.pbase = pbase;
    gt.cbase = r.chunk_base;
    gt.cnt = r.chunk_cnt;
    gt.srk = r3.rk;
    gt.srca = r3.rca;
    gt.srcb = r3.rcb;
    gt.rk = r.rk;
    gt.rca = r.rca;
    gt.rcb = r.rcb;
    gt.capacity = capacity;
    gt.phbase = phbase;
    gt.hbase = r.heavy_base;
    gt.hcnt = r.chunk_heavy;
    gt.sheavy = r3.heavy;
    gt.heavy = r.heavy;
    gt.heavy_cap = heavy_cap;
    gt.bad = (unsigned int *)(small + 5);
    {
      Prof pf(c, "rle_gather");
      HG_HIP(launch_rle_gather(gt, c->stream));
    }
    HG_TRY(sync_read(c, &nsmall, r.chunk_base + nch, 8));
    HG_TRY(sync_read(c, &nheavy, r.heavy_base + nch, 4));
    unsigned int bad = 0;
    HG_TRY(sync_read(c, &bad, small + 5, 4));
    if (bad) return fail(HG_ERR_DEVICE, "a chunk of the run-length join exceeded its bound");
    if (nheavy > heavy_cap) return fail(HG_ERR_DEVICE, "heavy bins exceed their bound");
  } else {
    r.pass = 0;
    {
      Prof pf(c, "rle_small");
      HG_HIP(launch_rle(r, c->stream));
    }
    // record positions of the small bins and heavy-list positions: exclusive
    // scans of the chunk counts
    HG_HIP(launch_scan(r.chunk_cnt, (void *)r.chunk_base, nch, 8, status, c->stream));
    HG_HIP(launch_scan(r.chunk_heavy, (void *)r.heavy_base, nch, 4, status, c->stream));
    HG_TRY(sync_read(c, &nsmall, r.chunk_base + nch, 8));
    HG_TRY(sync_read(c, &nheavy, r.heavy_base + nch, 4));
    if (nheavy > heavy_cap) return fail(HG_ERR_DEVICE, "heavy bins exceed their bound");
    // pass 1 lists the heavy bins and, with a capacity, writes the records
    r.pass = 1;
    {
      Prof pf(c, "rle_small");
      HG_HIP(launch_rle(r, c->stream));
    }
  }
  // heavy bins append after the small bins' records
  HG_HIP(launch_fill_u64(small + 2, nsmall, 1, c->stream));
  r.pass = 2;
  r.nheavy = nheavy;
  {
    Prof pf(c, "rle_heavy");
    HG_HIP(launch_rle(r, c->stream));
  }
  unsigned long long res[5];
  HG_TRY(sync_read(c, res, small, 40));
  if (res[3] & 0xFFFFFFFFull) return fail(HG_ERR_UNSUPPORTED, "a heavy bin holds 2^32 or more keys on a side");
  const uint64_t novf = res[4] & 0xFFFFFFFFull;
  if (novf > heavy_cap) return fail(HG_ERR_DEVICE, "value-rich bins exceed their bound");
  if (novf) {
    HG_TRY(rle_value_rich(c, r, r.ovf_list, (uint32_t)novf));
    HG_TRY(sync_read(c, res, small, 40));
  }
  *h_records = res[2];
  *h_count = res[1];
  if (res[2] > capacity && capacity) return fail(HG_ERR_CAPACITY, "record capacity smaller than the record count");
  return HG_OK;
}

hg_status hg_find_splits(hg_ctx c, const hg_table_desc *desc, const void *d_keys, uint64_t bin_count, uint32_t n_parts,
                         int64_t *h_splits) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(desc));
  if (!h_splits || n_parts == 0 || bin_count == 0 || bin_count > 32768)
    return fail(HG_ERR_INVALID_ARG, "bad split arguments (bin_count must be in [1, 32768])");
  if (desc->n_keys && !d_keys) return fail(HG_ERR_INVALID_ARG, "null keys");
  DeviceGuard g(c->device);
  const uint64_t M = desc->n_bins, n = desc->n_keys;
  // binRange = ceil(tableSize / (float)binCount)  (src/MultiHashGraph.cu:410)
  const uint64_t range = (uint64_t)std::ceil((float)M / (float)bin_count);
  if (range == 0) return fail(HG_ERR_INVALID_ARG, "bin range is zero");
  Plan P;
  const size_t i_hist = P.add(bin_count * 8 + 64);
  HG_TRY(arena_ensure(c, 1, P.total()));
  Carve cv;
  cv.init(c, 1, P);
  unsigned long long *hist = cv.get<unsigned long long>(i_hist);
  unsigned int *bad = (unsigned int *)(hist + bin_count);
  HG_HIP(launch_fill_u64(hist, 0ull, bin_count + 1, c->stream));
  HG_HIP(launch_range_hist(d_keys, n, (int)desc->key_bytes, make_map(M, desc->hash_kind, desc->d_rmi_model),
                           (uint32_t)range, (uint32_t)bin_count, hist, bad, c->stream));
  std::vector<unsigned long long> h(bin_count + 1);
  HG_TRY(sync_read(c, h.data(), hist, (bin_count + 1) * 8));
  if (h[bin_count]) return fail(HG_ERR_INVALID_ARG, "bin range too small for bin_count (float rounding)");
  splits_from_hist(h, bin_count, n, n_parts, range, M, h_splits);
  return HG_OK;
}

hg_status hg_build_partitioned(hg_ctx c, const hg_table_desc *desc, const void *d_keys, const int64_t *h_splits,
                               uint32_t n_parts, void *d_offsets, void *d_csr_keys, uint64_t *h_part_offsets) {
  HG_TRY(check_partitioned(c, desc, d_keys, h_splits, n_parts, d_offsets, d_csr_keys, h_part_offsets));
  DeviceGuard g(c->device);
  return build_partitioned(c, desc, d_keys, h_splits, n_parts, d_offsets, d_csr_keys, h_part_offsets);
}

hg_status hg_build_indexed(hg_ctx c, const hg_table_desc *desc, const void *d_keys, void *d_offsets, void *d_csr_keys,
                           uint64_t *d_csr_index) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(desc));
  if (!d_offsets || (desc->n_keys && (!d_keys || !d_csr_keys || !d_csr_index)))
    return fail(HG_ERR_INVALID_ARG, "null buffer");
  if (desc->hash_kind != HG_HASH_MURMUR3)
    return fail(HG_ERR_UNSUPPORTED, "index-carrying builds take murmur3 tables (bins of the low 32 key bits)");
  DeviceGuard g(c->device);
  const uint64_t n = desc->n_keys;
  if (n == 0) return fill_offsets(c, d_offsets, (int)desc->offset_bytes, desc->n_bins + 1, 0);
  hg_table_desc d8 = *desc;
  d8.key_bytes = 8;
  if (desc->key_bytes == 4) {
    // the first partition pass forms the records from the keys, the final
    // stage stores each record as its key and its position
    const BinMap map = make_map(d8.n_bins, d8.hash_kind, d8.d_rmi_model);
    return build_range(c, map, 8, (int)d8.offset_bytes, n, 0, d8.n_bins, d_keys, d_offsets, d_csr_keys, d_csr_index,
                       true);
  }
  uint64_t *rin = nullptr, *rout = nullptr;
  HG_TRY(index_records(c, n, rin, rout));
  {
    Prof pf(c, "pack_index");
    HG_HIP(launch_pack_index(d_keys, n, (int)desc->key_bytes, rin, c->stream));
  }
  HG_TRY(build_table(c, &d8, rin, d_offsets, rout));
  Prof pf(c, "unpack_index");
  HG_HIP(launch_unpack_index(rout, n, d_keys, (int)desc->key_bytes, d_csr_keys, d_csr_index, c->stream));
  return HG_OK;
}

hg_status hg_build_partitioned_indexed(hg_ctx c, const hg_table_desc *desc, const void *d_keys, const int64_t *h_splits,
                                       uint32_t n_parts, void *d_offsets, void *d_csr_keys, uint64_t *d_csr_index,
                                       uint64_t *h_part_offsets) {
  HG_TRY(check_partitioned(c, desc, d_keys, h_splits, n_parts, d_offsets, d_csr_keys, h_part_offsets));
  if (desc->n_keys && !d_csr_index) return fail(HG_ERR_INVALID_ARG, "null buffer");
  if (desc->hash_kind != HG_HASH_MURMUR3)
    return fail(HG_ERR_UNSUPPORTED, "index-carrying builds take murmur3 tables (bins of the low 32 key bits)");
  DeviceGuard g(c->device);
  const uint64_t n = desc->n_keys;
  hg_table_desc d8 = *desc;
  d8.key_bytes = 8;
  if (desc->key_bytes == 4)  // records formed by the first pass, stored back as key and position
    return build_partitioned(c, &d8, d_keys, h_splits, n_parts, d_offsets, d_csr_keys, h_part_offsets, d_csr_index,
                             true);
  uint64_t *rin = nullptr, *rout = nullptr;
  HG_TRY(index_records(c, std::max<uint64_t>(n, 1), rin, rout));
  {
    Prof pf(c, "pack_index");
    HG_HIP(launch_pack_index(d_keys, n, (int)desc->key_bytes, rin, c->stream));
  }
  HG_TRY(build_partitioned(c, &d8, rin, h_splits, n_parts, d_offsets, rout, h_part_offsets));
  Prof pf(c, "unpack_index");
  HG_HIP(launch_unpack_index(rout, h_part_offsets[n_parts], d_keys, (int)desc->key_bytes, d_csr_keys, d_csr_index,
                             c->stream));
  return HG_OK;
}

}  // extern "C"

namespace {

hg_status check_partitioned(hg_ctx c, const hg_table_desc *desc, const void *d_keys, const int64_t *h_splits,
                            uint32_t n_parts, void *d_offsets, void *d_csr_keys, uint64_t *h_part_offsets) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(desc));
  if (!h_splits || n_parts == 0 || n_parts > 4096 || !d_offsets || !h_part_offsets)
    return fail(HG_ERR_INVALID_ARG, "bad partition arguments");
  const uint64_t M = desc->n_bins, n = desc->n_keys;
  if (h_splits[0] != 0) return fail(HG_ERR_INVALID_ARG, "splits[0] must be 0");
  for (uint32_t p = 0; p < n_parts; p++)
    if (h_splits[p + 1] < h_splits[p] || (uint64_t)h_splits[p + 1] > M)
      return fail(HG_ERR_INVALID_ARG, "splits must be non-decreasing within [0, M]");
  if (n && (!d_keys || !d_csr_keys)) return fail(HG_ERR_INVALID_ARG, "null buffer");
  return HG_OK;
}

// The two record arrays (input, CSR) of an index-carrying build, in arena 2.
// Record buffers of the index-carrying builds of 64-bit keys: the packed
// input and the built records (unpacked afterwards).
hg_status index_records(hg_ctx c, uint64_t n, uint64_t *&rin, uint64_t *&rout) {
  Plan P;
  const size_t i_in = P.add(n * 8), i_out = P.add(n * 8);
  HG_TRY(arena_ensure(c, 2, P.total()));
  Carve cv;
  cv.init(c, 2, P);
  rin = cv.get<uint64_t>(i_in);
  rout = cv.get<uint64_t>(i_out);
  return HG_OK;
}

hg_status build_partitioned(hg_ctx c, const hg_table_desc *desc, const void *d_keys, const int64_t *h_splits,
                            uint32_t n_parts, void *d_offsets, void *d_csr_keys, uint64_t *h_part_offsets,
                            uint64_t *d_csr_index, bool pack_in) {
  const uint64_t M = desc->n_bins, n = desc->n_keys;
  const int kb = (int)desc->key_bytes, ob = (int)desc->offset_bytes;
  const BinMap map = make_map(M, desc->hash_kind, desc->d_rmi_model);
  if (n_parts == 1 && (uint64_t)h_splits[1] == M) {
    // one partition over every bin keeps every key: the plain build, whose
    // offsets are the partition's
    h_part_offsets[0] = 0;
    h_part_offsets[1] = n;
    return build_range(c, map, kb, ob, n, 0, M, d_keys, d_offsets, d_csr_keys, d_csr_index, pack_in);
  }
  Plan P0;
  const size_t i_part1 = P0.add(std::max<uint64_t>(n, 1) * kb), i_part2 = P0.add(std::max<uint64_t>(n, 1) * kb);
  HG_TRY(arena_ensure(c, 0, P0.total()));
  Carve cv0;
  cv0.init(c, 0, P0);
  void *part1 = cv0.get<void>(i_part1), *part2 = cv0.get<void>(i_part2);
  // level 1: the local partition by split ranges (a kind-1 segment)
  Level L1;
  SegDesc s1{};
  s1.key_begin = 0;
  s1.key_end = n;
  s1.map = map;
  s1.div = make_div(1);
  s1.kind = 1;
  s1.nbuckets = n_parts;
  s1.ncols = n_parts + 1;
  s1.nsplits = n_parts;
  s1.fin_end_pos = ~0ull;
  L1.segs.push_back(s1);
  L1.maps.push_back(map);
  L1.splits.assign(h_splits, h_splits + n_parts + 1);
  L1.auto_tile(n);
  L1.prepare();
  LevelDev d1;
  HG_TRY(setup_level(c, L1, d1));
  HG_TRY(run_level(c, L1, d1, kb, d_keys, part1, true, pack_in));
  std::vector<uint32_t> cnt(L1.ncols_total), st(L1.ncols_total);
  HG_TRY(sync_read(c, cnt.data(), d1.coltot, cnt.size() * 4));
  HG_TRY(sync_read(c, st.data(), d1.bstart, st.size() * 4));
  std::vector<Coarse> cs;
  uint64_t kept = 0;
  for (uint32_t p = 0; p < n_parts; p++) {
    const uint64_t lo = (uint64_t)h_splits[p], R = (uint64_t)h_splits[p + 1] - lo;
    h_part_offsets[p] = kept;
    if (R == 0) {  // empty range: its single offset entry is 0
      HG_TRY(fill_offsets(c, (char *)d_offsets + (lo + p) * ob, ob, 1, 0));
      continue;
    }
    Coarse q;
    q.key_begin = st[p];
    q.count = cnt[p];
    q.bin_lo = lo;
    q.nbins = R;
    q.off_base = lo + p;
    q.key_rel = st[p];
    q.end_pos = lo + R + p;
    cs.push_back(q);
    kept += cnt[p];
  }
  h_part_offsets[n_parts] = kept;
  return finish_two_level(c, cs, map, kb, ob, part1, part2, d_csr_keys, d_offsets, d_csr_index);
}

}  // namespace

// ---------------------------------------------------------------------------
// Staged tables: keys, offsets and CSR keys stay in host memory (pageable,
// pinned, or managed memory that the device cannot hold) and pass through
// HBM in chunks and bin-range slabs.  They replace the reference's MANAGED_MEM
// mode (src/MultiHashGraph.cu:229-247, include/MultiHashGraphHostOperators.cuh:
// 255-280, src/MultiHashGraph.cu:664-670), which puts a table in one
// cudaMallocManaged blob and lets page faults move it: on a device without
// demand paging, sequential DMA of whole slabs is how a table larger than
// HBM is built.
//   pass 1  each chunk of keys is uploaded and counted into pieces (bin
//           ranges of at most wf bins inside one partition, <= MAX_SPLITS);
//   pass 2  each chunk is uploaded again, partitioned by piece, and every
//           slab's share is copied to the slab's final place in the caller's
//           CSR keys (slab = consecutive pieces of one partition whose keys
//           and bins fit the budget);
//   pass 3  each slab is uploaded, built as a bin range in HBM, and its CSR
//           keys and rebased offsets are copied back over its region.
// Host-link traffic: 3 N s up, 2 N s down, offsets once down.  Uploads run on
// one stream, copies back on another and kernels on the context's stream,
// with two buffers of each kind, so that both link directions and the
// kernels overlap.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kStageMinBudget = 4ull << 20;
constexpr uint64_t kStageMaxChunk = 1ull << 31;  // 32-bit key positions inside the kernels

// Device memory per staged key and per staged bin, for the budget:
//  chunk key (passes 1, 2): two upload and two partition buffers plus the
//    piece level's matrices (~0.5 byte);
//  slab key (pass 3): two upload and two CSR buffers, the build's two
//    partitioned copies and its matrices (~1 byte);
//  slab bin: two offsets buffers and the final stage's per-bin counters.
struct StageGeom {
  double cap = 0;  // bytes usable for buffers (0.9 of the budget)
  double key_cost = 0, bin_cost = 0;
  uint64_t chunk = 0;
  bool fits(uint64_t keys, uint64_t bins) const {
    return keys <= kStageMaxChunk && (double)keys * key_cost + (double)bins * bin_cost <= cap;
  }
};

hg_status stage_geometry(hg_ctx c, int kb, int ob, size_t &budget, StageGeom &g) {
  (void)c;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    HG_HIP(hipMemGetInfo(&fr, &tot));
    budget = fr / 2;
  }
  if (budget < kStageMinBudget) return fail(HG_ERR_INVALID_ARG, "device_budget below 4 MiB");
  g.cap = 0.9 * (double)budget;
  g.key_cost = 6.0 * kb + 1.0;
  g.bin_cost = 2.0 * ob + 4.0;
  g.chunk = std::min<uint64_t>((uint64_t)(g.cap / (4.0 * kb + 0.5)), kStageMaxChunk);
  return HG_OK;
}

// The upload stream, the copy-back stream and their events (created on
// first use, default priority).
struct StageStreams {
  hipStream_t in = nullptr, out = nullptr;
  hipEvent_t ev_in[2] = {}, ev_done[2] = {}, ev_out[2] = {};
};

hg_status stage_copy(hipStream_t s, void *dst, const void *src, size_t bytes) {
  if (bytes) HG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
  return HG_OK;
}

struct Pieces {
  std::vector<int64_t> bounds;  // piece p holds bins [bounds[p], bounds[p+1])
  std::vector<uint32_t> part;   // its partition
};

hg_status make_pieces(const int64_t *splits, uint32_t n_parts, const StageGeom &g, Pieces &pc) {
  const uint64_t top = (uint64_t)splits[n_parts];
  const uint64_t avail = MAX_SPLITS - n_parts;
  const uint64_t wf = std::max<uint64_t>(1, (top + avail - 1) / avail);
  if ((double)wf * g.bin_cost > 0.5 * g.cap)
    return fail(HG_ERR_UNSUPPORTED, "device_budget too small for the table's bin count");
  for (uint32_t p = 0; p < n_parts; p++)
    for (uint64_t b = (uint64_t)splits[p]; b < (uint64_t)splits[p + 1]; b += wf) {
      pc.bounds.push_back((int64_t)b);
      pc.part.push_back(p);
    }
  pc.bounds.push_back((int64_t)top);
  return HG_OK;
}

// One kind-1 level over n keys: bucket p = piece p, the last column drops
// bins at or above the last bound.
void piece_level(Level &L, const BinMap &map, const Pieces &pc, uint64_t n) {
  const uint32_t np = (uint32_t)pc.part.size();
  SegDesc s{};
  s.key_begin = 0;
  s.key_end = n;
  s.map = map;
  s.div = make_div(1);
  s.kind = 1;
  s.nbuckets = np;
  s.ncols = np + 1;
  s.nsplits = np;
  s.fin_end_pos = ~0ull;
  L.segs.push_back(s);
  L.maps.push_back(map);
  L.splits = pc.bounds;
  L.auto_tile(n);
  L.prepare();
}

struct Slab {
  uint32_t p0 = 0, p1 = 0, part = 0;
  uint64_t lo = 0, R = 0, count = 0, pos = 0;
};

// Waits for every stream of a staged call on all exits, so that no copy is
// left in flight over the context's buffers.
struct StageDrain {
  hg_ctx c;
  const StageStreams &s;
  ~StageDrain() {
    (void)hipStreamSynchronize(s.in);
    (void)hipStreamSynchronize(s.out);
    (void)hipStreamSynchronize(c->stream);
  }
};

// Device-only memory (hipMalloc) the host cannot read or write: the staged
// calls read offsets and write empty partitions' entries from the host.
// Pageable (unknown to HIP), pinned and managed memory pass.
bool device_only(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

hg_status stage_streams(hg_ctx c, StageStreams &s) {
  if (!c->st_in) HG_HIP(hipStreamCreateWithFlags(&c->st_in, hipStreamNonBlocking));
  if (!c->st_out) HG_HIP(hipStreamCreateWithFlags(&c->st_out, hipStreamNonBlocking));
  for (hipEvent_t &e : c->st_ev)
    if (!e) HG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  s.in = c->st_in;
  s.out = c->st_out;
  for (int b = 0; b < 2; b++) {
    s.ev_in[b] = c->st_ev[b];
    s.ev_done[b] = c->st_ev[2 + b];
    s.ev_out[b] = c->st_ev[4 + b];
  }
  return HG_OK;
}

// splits over [0, M) of n_parts partitions; h_part_offsets gets the key
// start of every partition (n_parts + 1 entries)
hg_status staged_partitioned(hg_ctx c, const hg_table_desc *desc, const void *h_keys, const int64_t *splits,
                             uint32_t n_parts, void *h_off, void *h_csr, uint64_t *h_part_offsets, size_t budget) {
  const uint64_t n = desc->n_keys;
  const int kb = (int)desc->key_bytes, ob = (int)desc->offset_bytes;
  const BinMap map = make_map(desc->n_bins, desc->hash_kind, desc->d_rmi_model);
  StageGeom g;
  HG_TRY(stage_geometry(c, kb, ob, budget, g));
  Pieces pc;
  HG_TRY(make_pieces(splits, n_parts, g, pc));
  const uint32_t np = (uint32_t)pc.part.size();
  StageStreams sst;
  HG_TRY(stage_streams(c, sst));
  const StageStreams *ss = &sst;
  StageDrain drain{c, sst};
  const uint64_t C = std::max<uint64_t>(std::min(g.chunk, n), 1);
  const uint64_t nch = (n + C - 1) / C;
  // passes 1 and 2: arena 3 = two upload buffers and two partition buffers
  Plan P3;
  const size_t i_in0 = P3.add(C * kb), i_in1 = P3.add(C * kb), i_pt0 = P3.add(C * kb), i_pt1 = P3.add(C * kb);
  HG_TRY(arena_ensure(c, 3, P3.total()));
  Carve cv3;
  cv3.init(c, 3, P3);
  char *d_in[2] = {cv3.get<char>(i_in0), cv3.get<char>(i_in1)};
  char *d_pt[2] = {cv3.get<char>(i_pt0), cv3.get<char>(i_pt1)};
  // chunk j goes to buffer j & 1 once chunk j - 2's kernels are done with it
  auto upload = [&](uint64_t j) -> hg_status {
    const int b = (int)(j & 1);
    if (j >= 2) HG_HIP(hipStreamWaitEvent(ss->in, ss->ev_done[b], 0));
    HG_TRY(stage_copy(ss->in, d_in[b], (const char *)h_keys + j * C * kb, std::min(C, n - j * C) * kb));
    HG_HIP(hipEventRecord(ss->ev_in[b], ss->in));
    return HG_OK;
  };
  std::vector<uint32_t> cnt(np + 1), st(np + 1);
  // pass 1: keys per piece
  std::vector<uint64_t> T(np, 0);
  if (nch) HG_TRY(upload(0));
  for (uint64_t j = 0; j < nch; j++) {
    const int b = (int)(j & 1);
    if (j + 1 < nch) HG_TRY(upload(j + 1));
    HG_HIP(hipStreamWaitEvent(c->stream, ss->ev_in[b], 0));
    Level L;
    piece_level(L, map, pc, std::min(C, n - j * C));
    LevelDev d;
    HG_TRY(setup_level(c, L, d));
    HG_TRY(run_level(c, L, d, kb, d_in[b], nullptr, false));
    HG_HIP(hipEventRecord(ss->ev_done[b], c->stream));
    HG_TRY(sync_read(c, cnt.data(), d.coltot, (np + 1) * 4));
    for (uint32_t p = 0; p < np; p++) T[p] += cnt[p];
  }
  // slabs: consecutive pieces of one partition within the budget
  std::vector<Slab> slabs;
  std::vector<uint64_t> po(n_parts + 1, 0);
  uint64_t pos = 0, smax = 1, rmax = 1;
  for (uint32_t p = 0; p < np;) {
    Slab sl;
    sl.p0 = sl.p1 = p;
    sl.part = pc.part[p];
    sl.lo = (uint64_t)pc.bounds[p];
    sl.pos = pos;
    while (p < np && pc.part[p] == sl.part) {
      const uint64_t pb = (uint64_t)(pc.bounds[p + 1] - pc.bounds[p]);
      if (!g.fits(T[p], pb))
        return fail(HG_ERR_UNSUPPORTED, "bins [" + std::to_string(pc.bounds[p]) + ", " +
                                            std::to_string(pc.bounds[p + 1]) + ") hold " + std::to_string(T[p]) +
                                            " keys, more than a slab of this device_budget holds");
      if (sl.p1 > sl.p0 && !g.fits(sl.count + T[p], sl.R + pb)) break;
      sl.count += T[p];
      sl.R += pb;
      sl.p1 = ++p;
    }
    pos += sl.count;
    smax = std::max(smax, sl.count);
    rmax = std::max(rmax, sl.R);
    slabs.push_back(sl);
  }
  for (const Slab &sl : slabs) po[sl.part + 1] += sl.count;
  for (uint32_t p = 0; p < n_parts; p++) po[p + 1] += po[p];
  // pass 2: every chunk's share of a slab goes to the slab's final region
  std::vector<uint64_t> cur(slabs.size(), 0);
  if (nch) HG_TRY(upload(0));
  for (uint64_t j = 0; j < nch; j++) {
    const int b = (int)(j & 1);
    if (j + 1 < nch) HG_TRY(upload(j + 1));
    HG_HIP(hipStreamWaitEvent(c->stream, ss->ev_in[b], 0));
    if (j >= 2) HG_HIP(hipStreamWaitEvent(c->stream, ss->ev_out[b], 0));  // chunk j - 2's copies are out
    Level L;
    piece_level(L, map, pc, std::min(C, n - j * C));
    LevelDev d;
    HG_TRY(setup_level(c, L, d));
    HG_TRY(run_level(c, L, d, kb, d_in[b], d_pt[b]));
    HG_HIP(hipEventRecord(ss->ev_done[b], c->stream));
    HG_TRY(sync_read(c, cnt.data(), d.coltot, (np + 1) * 4));
    HG_TRY(sync_read(c, st.data(), d.bstart, (np + 1) * 4));
    HG_HIP(hipStreamWaitEvent(ss->out, ss->ev_done[b], 0));
    for (size_t s = 0; s < slabs.size(); s++) {
      const Slab &sl = slabs[s];
      const uint64_t a = st[sl.p0], e = (uint64_t)st[sl.p1 - 1] + cnt[sl.p1 - 1];
      if (cur[s] + (e - a) > sl.count) return fail(HG_ERR_DEVICE, "staged build: keys changed between passes");
      HG_TRY(stage_copy(ss->out, (char *)h_csr + (sl.pos + cur[s]) * kb, d_pt[b] + a * kb, (e - a) * kb));
      cur[s] += e - a;
    }
    HG_HIP(hipEventRecord(ss->ev_out[b], ss->out));
  }
  for (size_t s = 0; s < slabs.size(); s++)
    if (cur[s] != slabs[s].count) return fail(HG_ERR_DEVICE, "staged build: keys changed between passes");
  HG_HIP(hipStreamSynchronize(ss->out));
  HG_HIP(hipStreamSynchronize(ss->in));
  // pass 3: arena 3 = two upload, two CSR and two offsets buffers
  Plan Q3;
  const size_t j_in0 = Q3.add(smax * kb), j_in1 = Q3.add(smax * kb), j_out0 = Q3.add(smax * kb),
               j_out1 = Q3.add(smax * kb), j_off0 = Q3.add((rmax + 1) * ob), j_off1 = Q3.add((rmax + 1) * ob);
  HG_TRY(arena_ensure(c, 3, Q3.total()));
  HG_TRY(arena_ensure(c, 0, 2 * smax * kb + 512));
  cv3.init(c, 3, Q3);
  char *s_in[2] = {cv3.get<char>(j_in0), cv3.get<char>(j_in1)};
  char *s_out[2] = {cv3.get<char>(j_out0), cv3.get<char>(j_out1)};
  void *s_off[2] = {cv3.get<void>(j_off0), cv3.get<void>(j_off1)};
  auto upload_slab = [&](size_t s) -> hg_status {
    const int b = (int)(s & 1);
    if (s >= 2) HG_HIP(hipStreamWaitEvent(ss->in, ss->ev_done[b], 0));
    HG_TRY(stage_copy(ss->in, s_in[b], (const char *)h_csr + slabs[s].pos * kb, slabs[s].count * kb));
    HG_HIP(hipEventRecord(ss->ev_in[b], ss->in));
    return HG_OK;
  };
  if (!slabs.empty()) HG_TRY(upload_slab(0));
  for (size_t s = 0; s < slabs.size(); s++) {
    const Slab &sl = slabs[s];
    const int b = (int)(s & 1);
    if (s + 1 < slabs.size()) HG_TRY(upload_slab(s + 1));
    HG_HIP(hipStreamWaitEvent(c->stream, ss->ev_in[b], 0));
    if (s >= 2) HG_HIP(hipStreamWaitEvent(c->stream, ss->ev_out[b], 0));  // slab s - 2's copies are out
    HG_TRY(build_range(c, map, kb, ob, sl.count, sl.lo, sl.R, s_in[b], s_off[b], s_out[b]));
    // offsets of a partitioned table count from the partition's first key
    HG_HIP(launch_offsets_add(s_off[b], ob, sl.R + 1, sl.pos - po[sl.part], c->stream));
    HG_HIP(hipEventRecord(ss->ev_done[b], c->stream));
    HG_HIP(hipStreamWaitEvent(ss->out, ss->ev_done[b], 0));
    HG_TRY(stage_copy(ss->out, (char *)h_csr + sl.pos * kb, s_out[b], sl.count * kb));
    HG_TRY(stage_copy(ss->out, (char *)h_off + (sl.lo + sl.part) * ob, s_off[b], (sl.R + 1) * ob));
    HG_HIP(hipEventRecord(ss->ev_out[b], ss->out));
  }
  HG_HIP(hipStreamSynchronize(ss->out));
  HG_HIP(hipStreamSynchronize(c->stream));
  // a partition of no bins has the single offset entry 0
  for (uint32_t p = 0; p < n_parts; p++)
    if (splits[p + 1] == splits[p]) {
      char *q = (char *)h_off + ((uint64_t)splits[p] + p) * ob;
      if (ob == 4) *(uint32_t *)q = 0;
      else *(uint64_t *)q = 0;
    }
  for (uint32_t p = 0; p <= n_parts; p++) h_part_offsets[p] = po[p];
  return HG_OK;
}

// include/MultiHashGraphHostOperators.cuh:55-59,83-88: key-balanced split
// points from the bin_count-range histogram h (h[bin_count] = keys beyond)
void splits_from_hist(const std::vector<unsigned long long> &h, uint64_t bin_count, uint64_t n, uint32_t n_parts,
                      uint64_t range, uint64_t M, int64_t *h_splits) {
  std::vector<int64_t> ps(bin_count + 1);
  ps[0] = 0;
  for (uint64_t i = 1; i < bin_count; i++) ps[i] = ps[i - 1] + (int64_t)h[i - 1];
  ps[bin_count] = (int64_t)n;
  const int64_t avg = (int64_t)std::ceil((float)n / (float)n_parts);
  h_splits[0] = 0;
  for (uint32_t t = 0; t < n_parts; t++) {
    const uint64_t upper = (uint64_t)(avg * (int64_t)(t + 1));
    const uint64_t idx = ref_binary_search(ps, 0, bin_count, upper);
    h_splits[t + 1] = (int64_t)std::min<uint64_t>(idx * range, M);
  }
}

uint64_t off_at(const void *off, int ob, uint64_t i) {
  return ob == 4 ? (uint64_t)((const uint32_t *)off)[i] : ((const uint64_t *)off)[i];
}

// One slab of bins [lo, hi) of two host tables resident in HBM: offsets
// rebased to 0, keys from the slab's first key.
struct JoinSlab {
  uint64_t lo = 0, hi = 0;
  uint64_t a0 = 0, b0 = 0;  // first key of A and of B in the host tables
  void *offA = nullptr, *offB = nullptr;
  char *keysA = nullptr, *keysB = nullptr;
};

// The matches' staging buffers of a staged join: emissions alternate between
// two, each copied back on the copy-back stream.
struct JoinStage {
  void *buf[2] = {nullptr, nullptr};
  uint64_t cap = 0;
  int next = 0;
  uint32_t used = 0;  // emissions so far (a buffer is waited for from its second use on)
};

// Matches of bins [x, y) of the slab: counted, then emitted into a stage
// buffer and copied to the host list when they fit it (else the range is
// halved).
hg_status join_range(hg_ctx c, const StageStreams &ss, const hg_table_desc *descA, const void *h_offA,
                     const void *h_offB, const JoinSlab &js, uint64_t x, uint64_t y, JoinStage &stg, void *h_out,
                     uint64_t capacity, uint64_t &written, uint64_t &total, bool &full) {
  const int kb = (int)descA->key_bytes, ob = (int)descA->offset_bytes;
  hg_table_desc dA = *descA, dB = *descA;
  dA.n_bins = dB.n_bins = y - x;
  dA.n_keys = off_at(h_offA, ob, y) - off_at(h_offA, ob, x);
  dB.n_keys = off_at(h_offB, ob, y) - off_at(h_offB, ob, x);
  const void *oa = (const char *)js.offA + (x - js.lo) * ob, *obp = (const char *)js.offB + (x - js.lo) * ob;
  uint64_t cnt = 0;
  HG_TRY(hg_intersect(c, &dA, oa, js.keysA, &dB, obp, js.keysB, nullptr, 0, &cnt));
  total += cnt;
  if (!h_out || full || cnt == 0) return HG_OK;
  if (written + cnt > capacity) {
    full = true;
    return HG_OK;
  }
  if (cnt > stg.cap) {
    if (y - x == 1)
      return fail(HG_ERR_UNSUPPORTED, "bin " + std::to_string(x) + " has " + std::to_string(cnt) +
                                          " matches, more than this device_budget stages (" +
                                          std::to_string(stg.cap) + ")");
    total -= cnt;
    const uint64_t m = x + (y - x) / 2;
    HG_TRY(join_range(c, ss, descA, h_offA, h_offB, js, x, m, stg, h_out, capacity, written, total, full));
    return join_range(c, ss, descA, h_offA, h_offB, js, m, y, stg, h_out, capacity, written, total, full);
  }
  const int e = stg.next;
  if (stg.used >= 2) HG_HIP(hipStreamWaitEvent(c->stream, ss.ev_out[e], 0));  // its last copy back is done
  uint64_t got = 0;
  HG_TRY(hg_intersect(c, &dA, oa, js.keysA, &dB, obp, js.keysB, stg.buf[e], stg.cap, &got));
  if (got != cnt) return fail(HG_ERR_DEVICE, "staged intersect: count changed between passes");
  // hg_intersect has synchronised the context's stream: the matches are in place
  HG_TRY(stage_copy(ss.out, (char *)h_out + written * kb, stg.buf[e], cnt * kb));
  HG_HIP(hipEventRecord(ss.ev_out[e], ss.out));
  stg.next = 1 - e;
  stg.used++;
  written += cnt;
  return HG_OK;
}

}  // namespace

extern "C" {

hg_status hg_build_staged(hg_ctx c, const hg_table_desc *desc, const void *h_keys, void *h_offsets,
                          void *h_csr_keys, size_t device_budget) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(desc, false));
  if (!h_offsets || (desc->n_keys && (!h_keys || !h_csr_keys))) return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  if (device_only(h_offsets)) return fail(HG_ERR_INVALID_ARG, "h_offsets must be host memory (pageable, pinned or managed)");
  const int64_t splits[2] = {0, (int64_t)desc->n_bins};
  uint64_t po[2];
  return staged_partitioned(c, desc, h_keys, splits, 1, h_offsets, h_csr_keys, po, device_budget);
}

hg_status hg_build_partitioned_staged(hg_ctx c, const hg_table_desc *desc, const void *h_keys,
                                      const int64_t *h_splits, uint32_t n_parts, void *h_offsets, void *h_csr_keys,
                                      uint64_t *h_part_offsets, size_t device_budget) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(desc, false));
  if (!h_splits || n_parts == 0 || n_parts > MAX_SPLITS / 2 || !h_offsets || !h_part_offsets)
    return fail(HG_ERR_INVALID_ARG, "bad partition arguments (n_parts must be in [1, 2048])");
  if (h_splits[0] != 0) return fail(HG_ERR_INVALID_ARG, "splits[0] must be 0");
  for (uint32_t p = 0; p < n_parts; p++)
    if (h_splits[p + 1] < h_splits[p] || (uint64_t)h_splits[p + 1] > desc->n_bins)
      return fail(HG_ERR_INVALID_ARG, "splits must be non-decreasing within [0, M]");
  if (desc->n_keys && (!h_keys || !h_csr_keys)) return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  if (device_only(h_offsets)) return fail(HG_ERR_INVALID_ARG, "h_offsets must be host memory (pageable, pinned or managed)");
  return staged_partitioned(c, desc, h_keys, h_splits, n_parts, h_offsets, h_csr_keys, h_part_offsets,
                            device_budget);
}

hg_status hg_find_splits_staged(hg_ctx c, const hg_table_desc *desc, const void *h_keys, uint64_t bin_count,
                                uint32_t n_parts, int64_t *h_splits, size_t device_budget) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(desc, false));
  if (!h_splits || n_parts == 0 || bin_count == 0 || bin_count > 32768)
    return fail(HG_ERR_INVALID_ARG, "bad split arguments (bin_count must be in [1, 32768])");
  if (desc->n_keys && !h_keys) return fail(HG_ERR_INVALID_ARG, "null keys");
  DeviceGuard g(c->device);
  const uint64_t M = desc->n_bins, n = desc->n_keys;
  const int kb = (int)desc->key_bytes;
  const uint64_t range = (uint64_t)std::ceil((float)M / (float)bin_count);
  if (range == 0) return fail(HG_ERR_INVALID_ARG, "bin range is zero");
  StageGeom geo;
  HG_TRY(stage_geometry(c, kb, (int)desc->offset_bytes, device_budget, geo));
  Plan P3;
  const size_t i_in = P3.add(std::max<uint64_t>(std::min(geo.chunk, n), 1) * kb);
  HG_TRY(arena_ensure(c, 3, P3.total()));
  Carve cv3;
  cv3.init(c, 3, P3);
  char *d_in = cv3.get<char>(i_in);
  Plan P;
  const size_t i_hist = P.add(bin_count * 8 + 64);
  HG_TRY(arena_ensure(c, 1, P.total()));
  Carve cv;
  cv.init(c, 1, P);
  unsigned long long *hist = cv.get<unsigned long long>(i_hist);
  unsigned int *bad = (unsigned int *)(hist + bin_count);
  HG_HIP(launch_fill_u64(hist, 0ull, bin_count + 1, c->stream));
  const BinMap map = make_map(M, desc->hash_kind, desc->d_rmi_model);
  for (uint64_t k0 = 0; k0 < n; k0 += geo.chunk) {
    const uint64_t cn = std::min(geo.chunk, n - k0);
    HG_TRY(stage_copy(c->stream, d_in, (const char *)h_keys + k0 * kb, cn * kb));
    HG_HIP(launch_range_hist(d_in, cn, kb, map, (uint32_t)range, (uint32_t)bin_count, hist, bad, c->stream));
  }
  std::vector<unsigned long long> h(bin_count + 1);
  HG_TRY(sync_read(c, h.data(), hist, (bin_count + 1) * 8));
  if (h[bin_count]) return fail(HG_ERR_INVALID_ARG, "bin range too small for bin_count (float rounding)");
  splits_from_hist(h, bin_count, n, n_parts, range, M, h_splits);
  return HG_OK;
}

hg_status hg_intersect_staged(hg_ctx c, const hg_table_desc *descA, const void *h_offA, const void *h_keysA,
                              const hg_table_desc *descB, const void *h_offB, const void *h_keysB, void *h_out,
                              uint64_t capacity, uint64_t *h_count, size_t device_budget) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(check_desc(descA, false));
  HG_TRY(check_desc(descB, false));
  if (!h_count) return fail(HG_ERR_INVALID_ARG, "null h_count");
  if (descA->n_bins != descB->n_bins || descA->key_bytes != descB->key_bytes ||
      descA->hash_kind != descB->hash_kind || descA->offset_bytes != descB->offset_bytes)
    return fail(HG_ERR_MISMATCH, "tables do not share a bin space");
  if (!h_offA || !h_offB || (descA->n_keys && !h_keysA) || (descB->n_keys && !h_keysB))
    return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  *h_count = 0;
  if (device_only(h_offA) || device_only(h_offB))
    return fail(HG_ERR_INVALID_ARG, "offsets must be host memory (pageable, pinned or managed)");
  const uint64_t M = descA->n_bins;
  const int kb = (int)descA->key_bytes, ob = (int)descA->offset_bytes;
  if (off_at(h_offA, ob, M) != descA->n_keys || off_at(h_offB, ob, M) != descB->n_keys)
    return fail(HG_ERR_INVALID_ARG, "offsets[n_bins] differs from n_keys");
  if (device_budget == 0) {
    size_t fr = 0, tot = 0;
    HG_HIP(hipMemGetInfo(&fr, &tot));
    device_budget = fr / 2;
  }
  if (device_budget < kStageMinBudget) return fail(HG_ERR_INVALID_ARG, "device_budget below 4 MiB");
  // two slab buffers (each a quarter of the budget for keys of A and B and a
  // twentieth for each table's offsets) and two stage buffers (an eighth each)
  const uint64_t kmax = std::min<uint64_t>((uint64_t)(0.25 * (double)device_budget / kb), kStageMaxChunk);
  const uint64_t rmax = std::min<uint64_t>((uint64_t)(0.05 * (double)device_budget / ob), (1ull << 32) - 2);
  JoinStage stg;
  stg.cap = (uint64_t)(0.125 * (double)device_budget / kb);
  Plan P3;
  size_t i_k[2], i_oa[2], i_ob[2], i_st[2];
  for (int b = 0; b < 2; b++) {
    i_k[b] = P3.add(kmax * kb);
    i_oa[b] = P3.add((rmax + 1) * ob);
    i_ob[b] = P3.add((rmax + 1) * ob);
    i_st[b] = P3.add(stg.cap * kb);
  }
  StageStreams ss;
  HG_TRY(stage_streams(c, ss));
  StageDrain drain{c, ss};
  HG_TRY(arena_ensure(c, 3, P3.total()));
  Carve cv3;
  cv3.init(c, 3, P3);
  for (int b = 0; b < 2; b++) stg.buf[b] = cv3.get<void>(i_st[b]);
  auto keys_in = [&](uint64_t lo, uint64_t hi) {
    return (off_at(h_offA, ob, hi) - off_at(h_offA, ob, lo)) + (off_at(h_offB, ob, hi) - off_at(h_offB, ob, lo));
  };
  // slabs: the largest hi with the keys within kmax and the bins within rmax
  std::vector<JoinSlab> slabs;
  for (uint64_t lo = 0; lo < M;) {
    uint64_t top = std::min<uint64_t>(M, lo + rmax);
    if (keys_in(lo, lo + 1) > kmax)
      return fail(HG_ERR_UNSUPPORTED, "bin " + std::to_string(lo) + " holds " + std::to_string(keys_in(lo, lo + 1)) +
                                          " keys of A and B, more than this device_budget stages (" +
                                          std::to_string(kmax) + ")");
    uint64_t good = lo + 1;
    while (good < top) {
      const uint64_t mid = good + (top - good + 1) / 2;
      if (keys_in(lo, mid) <= kmax) good = mid;
      else top = mid - 1;
    }
    JoinSlab js;
    js.lo = lo;
    js.hi = good;
    js.a0 = off_at(h_offA, ob, lo);
    js.b0 = off_at(h_offB, ob, lo);
    slabs.push_back(js);
    lo = good;
  }
  // slab s uploads into buffer s & 1 (its offsets rebased on the upload
  // stream) once slab s - 2's joins are done with it
  auto upload = [&](size_t s) -> hg_status {
    const int b = (int)(s & 1);
    JoinSlab &js = slabs[s];
    const uint64_t na = off_at(h_offA, ob, js.hi) - js.a0, nb = off_at(h_offB, ob, js.hi) - js.b0;
    js.keysA = cv3.get<char>(i_k[b]);
    js.keysB = js.keysA + na * kb;
    js.offA = cv3.get<void>(i_oa[b]);
    js.offB = cv3.get<void>(i_ob[b]);
    if (s >= 2) HG_HIP(hipStreamWaitEvent(ss.in, ss.ev_done[b], 0));
    HG_TRY(stage_copy(ss.in, js.keysA, (const char *)h_keysA + js.a0 * kb, na * kb));
    HG_TRY(stage_copy(ss.in, js.keysB, (const char *)h_keysB + js.b0 * kb, nb * kb));
    HG_TRY(stage_copy(ss.in, js.offA, (const char *)h_offA + js.lo * ob, (js.hi - js.lo + 1) * ob));
    HG_TRY(stage_copy(ss.in, js.offB, (const char *)h_offB + js.lo * ob, (js.hi - js.lo + 1) * ob));
    HG_HIP(launch_offsets_add(js.offA, ob, js.hi - js.lo + 1, (uint64_t)0 - js.a0, ss.in));
    HG_HIP(launch_offsets_add(js.offB, ob, js.hi - js.lo + 1, (uint64_t)0 - js.b0, ss.in));
    HG_HIP(hipEventRecord(ss.ev_in[b], ss.in));
    return HG_OK;
  };
  uint64_t written = 0, total = 0;
  bool full = false;
  if (!slabs.empty()) HG_TRY(upload(0));
  for (size_t s = 0; s < slabs.size(); s++) {
    const int b = (int)(s & 1);
    if (s + 1 < slabs.size()) HG_TRY(upload(s + 1));
    HG_HIP(hipStreamWaitEvent(c->stream, ss.ev_in[b], 0));
    HG_TRY(join_range(c, ss, descA, h_offA, h_offB, slabs[s], slabs[s].lo, slabs[s].hi, stg, h_out, capacity,
                      written, total, full));
    HG_HIP(hipEventRecord(ss.ev_done[b], c->stream));
  }
  HG_HIP(hipStreamSynchronize(ss.out));
  HG_HIP(hipStreamSynchronize(c->stream));
  *h_count = total;
  if (h_out && total > capacity) return fail(HG_ERR_CAPACITY, "output capacity smaller than the match count");
  return HG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched tables (no counterpart in the reference).  Every table is a range
// segment of one partition level, so one set of launches builds (or probes)
// all of them; a plan keeps the descriptors resident and replays the launches
// as one captured hipGraph.
// ---------------------------------------------------------------------------
namespace {

hg_status batch_validate(const hg_batch_desc *bd) {
  if (!bd || bd->n_tables == 0) return fail(HG_ERR_INVALID_ARG, "empty batch");
  if (bd->key_bytes != 4 && bd->key_bytes != 8) return fail(HG_ERR_INVALID_ARG, "key_bytes must be 4 or 8");
  if (!bd->h_key_off || !bd->h_n_bins || !bd->h_off_off) return fail(HG_ERR_INVALID_ARG, "null batch array");
  for (uint32_t t = 0; t < bd->n_tables; t++) {
    if (bd->h_key_off[t + 1] < bd->h_key_off[t]) return fail(HG_ERR_INVALID_ARG, "key offsets must not decrease");
    if (bd->h_key_off[t + 1] - bd->h_key_off[t] >= (1ull << 32))
      return fail(HG_ERR_UNSUPPORTED, "a batched table needs fewer than 2^32 keys (uint32 offsets)");
    if (bd->h_n_bins[t] == 0 || bd->h_n_bins[t] >= (1ull << 32))
      return fail(HG_ERR_INVALID_ARG, "n_bins must be in [1, 2^32)");
    const uint32_t k = bd->h_hash_kind ? bd->h_hash_kind[t] : HG_HASH_MURMUR3;
    if (k != HG_HASH_MURMUR3 && k != HG_HASH_RMI) return fail(HG_ERR_INVALID_ARG, "unknown hash kind");
    if (k == HG_HASH_RMI && (!bd->d_rmi_models || hg_rmi_model_bytes(bd->rmi_leaves) == 0))
      return fail(HG_ERR_INVALID_ARG, "learned tables need d_rmi_models and rmi_leaves");
  }
  return HG_OK;
}

BinMap batch_map(const hg_batch_desc *bd, uint32_t t) {
  const uint32_t k = bd->h_hash_kind ? bd->h_hash_kind[t] : HG_HASH_MURMUR3;
  const void *m = k == HG_HASH_RMI ? (const char *)bd->d_rmi_models + (size_t)t * hg_rmi_model_bytes(bd->rmi_leaves)
                                   : nullptr;
  return make_map(bd->h_n_bins[t], k, m);
}

struct BatchCfg {
  uint32_t tile_keys = kTile;
  double fill = 0.875;
  bool train = false;  // models trained inside the plan: kinds set on the device
};

// One side of a batch: the level over all tables, its device descriptors and
// the partitioned copy (in arena 0 of the context that owns the side).
struct BatchSide {
  Level L;
  LevelDev d;
  void *part = nullptr;
  unsigned int *flag = nullptr;
  unsigned int *hmark = nullptr;
  uint64_t nkeys = 0;
};

// Lays out the level of a batch side.  probe_off == nullptr: the build side
// over the tables' keys; else the probe side over the probe sets.
hg_status batch_prepare(hg_ctx c, const hg_batch_desc *bd, const uint64_t *probe_off, const BatchCfg &cfg,
                        BatchSide &S) {
  const int kb = (int)bd->key_bytes;
  S.L = Level();
  const uint64_t *ko = bd->h_key_off;
  const uint64_t *src = probe_off ? probe_off : ko;
  for (uint32_t t = 0; t < bd->n_tables; t++) {
    const uint64_t n = ko[t + 1] - ko[t], np = src[t + 1] - src[t], M = bd->h_n_bins[t];
    const BinMap map = batch_map(bd, t);
    const uint32_t W = choose_width(probe_off ? std::max(n, np) : n, M, kb,
                                    probe_off ? std::min(cfg.fill, probe_fill(kb)) : cfg.fill);
    SegDesc s = range_seg(src[t], src[t + 1], 0, M, W, map);
    if (s.nbuckets > kFMax) return fail(HG_ERR_UNSUPPORTED, "a batched table is too large for one partition level");
    s.fin_off_base = bd->h_off_off[t];
    s.fin_key_rel = ko[t];
    s.fin_end_pos = probe_off ? ~0ull : bd->h_off_off[t] + M;
    s.table = t;
    S.L.segs.push_back(s);
    S.L.maps.push_back(map);
  }
  S.L.tile_keys = cfg.tile_keys;
  S.L.any_kind = cfg.train;
  S.L.prepare();
  S.nkeys = src[bd->n_tables] - src[0];
  Plan P0;
  const size_t i_part = P0.add(std::max<uint64_t>(src[bd->n_tables], 1) * kb);
  const size_t i_flag = P0.add(64);
  const size_t i_mark = P0.add((size_t)std::max<uint32_t>(S.L.ncols_total, 1) * 4);
  HG_TRY(arena_ensure(c, 0, P0.total()));
  Carve cv;
  cv.init(c, 0, P0);
  S.part = cv.get<void>(i_part);
  S.flag = cv.get<unsigned int>(i_flag);
  S.hmark = cv.get<unsigned int>(i_mark);
  return setup_level(c, S.L, S.d);
}

// Kernel launches only (no allocation, upload or host sync): capturable.
hg_status batch_launch_build(hg_ctx c, const hg_batch_desc *bd, BatchSide &S, const void *d_keys, uint32_t *d_offsets,
                             void *d_csr) {
  HG_TRY(run_level(c, S.L, S.d, (int)bd->key_bytes, d_keys, S.part));
  return run_final(c, S.L, S.d, (int)bd->key_bytes, 4, S.part, d_csr, d_offsets);
}

hg_status batch_launch_probe(hg_ctx c, const hg_batch_desc *bd, BatchSide &S, const void *d_probe,
                             const uint32_t *d_offsets, const void *d_csr, uint64_t *d_counts, void *d_out = nullptr,
                             const uint64_t *d_out_off = nullptr) {
  HG_HIP(launch_fill_u64((unsigned long long *)d_counts, 0ull, bd->n_tables, c->stream));
  HG_HIP(launch_fill_u32(S.flag, 0u, 3, c->stream));  // [1], [2]: the probe gates
  HG_HIP(launch_fill_u32(S.hmark, 0u, S.L.ncols_total, c->stream));
  HG_TRY(run_level(c, S.L, S.d, (int)bd->key_bytes, d_probe, S.part));
  HG_HIP(launch_make_items(S.d.segs, (uint32_t)S.L.segs.size(), S.d.bstart, S.d.coltot, S.d.fin, S.L.ncols_total,
                           c->stream));
  LaunchProbe pr{};
  pr.key_bytes = (int)bd->key_bytes;
  pr.off_bytes = 4;
  pr.hk = S.L.hash_kind();
  pr.nitems = S.L.ncols_total;
  pr.akey_cap = probe_akey_cap(pr.key_bytes, 4);
  pr.bpart = S.part;
  pr.items = S.d.fin;
  pr.maps = S.d.maps;
  pr.offA = d_offsets;
  pr.keysA = d_csr;
  pr.out = d_out;  // with per-table regions d_out_off (null: counts only)
  pr.out_off = d_out ? d_out_off : nullptr;
  pr.capacity = 0;
  pr.cursor = (unsigned long long *)d_counts;
  pr.gate = S.flag + 1;
  pr.hmark = S.hmark;
  pr.big_slot = S.d.big_slot;
  pr.big_ctl = S.d.big_ctl;
  Prof pf(c, "probe");
  HG_HIP(launch_probe(pr, c->stream));
  return HG_OK;
}

}  // namespace

struct hg_batch_plan_s {
  int device = 0;
  hg_ctx owner = nullptr;             // caller's context (stream to order against)
  hg_ctx cb = nullptr, cp = nullptr;  // private contexts of the build and probe sides
  hipStream_t stream = nullptr;       // private stream the graph runs on
  hipEvent_t ev_in = nullptr, ev_out = nullptr, t0 = nullptr, t1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hg_batch_desc bd{};
  std::vector<uint64_t> key_off, n_bins, off_off, probe_off;
  std::vector<uint32_t> kinds;
  const void *d_keys = nullptr, *d_probe = nullptr;
  uint32_t *d_offsets = nullptr;
  void *d_csr = nullptr;
  uint64_t *d_counts = nullptr;
  BatchSide sb, sp;
  BatchCfg cfg;
  hg_batch_plan_info info{};
  // models trained inside the plan (HG_PLAN_TRAIN)
  bool train = false;
  std::vector<LaunchRmiBatch> rb;
  uint32_t *use = nullptr;
};

namespace {

void plan_drop_graph(hg_batch_plan p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

// In-plan training: table t's model (at d_rmi_models + t * model bytes) is
// fitted on max(n/100, min(n, 1024)) sampled keys with seed t, on every
// replay; its bin maps point at the model and take the kind its sampled
// check chooses.  Samples must fit the one-workgroup trainer.
hg_status plan_setup_training(hg_batch_plan p) {
  const hg_batch_desc &bd = p->bd;
  const uint32_t T = bd.n_tables;
  const uint32_t L = bd.rmi_leaves;
  if (!bd.d_rmi_models || L == 0 || L > RMI_ONE_MAX_LEAVES)
    return fail(HG_ERR_INVALID_ARG, "in-plan training needs d_rmi_models and rmi_leaves in [1, 1024]");
  const int kb = (int)bd.key_bytes;
  const size_t mb = hg_rmi_model_bytes(L);
  std::vector<RmiBatchItem> items;
  for (uint32_t t = 0; t < T; t++) {
    const uint64_t n = p->key_off[t + 1] - p->key_off[t];
    if (n == 0) continue;
    const uint64_t ns = std::max<uint64_t>(n / 100, std::min<uint64_t>(n, 1024));
    long double mp = (long double)p->n_bins[t] * (long double)ns / (long double)n;
    uint64_t Mp = (uint64_t)(mp + 0.5L);
    Mp = std::max<uint64_t>(1, std::min<uint64_t>(Mp, p->n_bins[t]));
    uint64_t np2 = 16;
    while (np2 < ns) np2 <<= 1;
    if (np2 * kb > RMI_ONE_MAX_KEY_BYTES || Mp > RMI_ONE_MAX_CHECK_BINS)
      return fail(HG_ERR_UNSUPPORTED, "in-plan training takes tables whose 1% sample fits one workgroup (up to "
                                      "16384 uint32 or 8192 uint64 keys)");
    RmiBatchItem it{};
    it.key_begin = p->key_off[t];
    it.n_keys = n;
    it.s0 = splitmix_mix(t);
    it.model_off = (uint64_t)t * mb;
    it.map = make_map(p->n_bins[t], 0, nullptr);
    it.ns = (uint32_t)ns;
    it.np2 = (uint32_t)np2;
    it.Mp = (uint32_t)Mp;
    it.table = t;
    items.push_back(it);
  }
  p->rb = rmi_launches(items, kb, L);
  // every table's maps point at its model slot; kinds follow the check
  p->kinds.assign(T, HG_HASH_RMI);
  p->bd.h_hash_kind = p->kinds.data();
  Blob bl;
  const size_t o_items = bl.put(items);
  Plan P;
  const size_t i_blob = P.add(std::max<size_t>(bl.data.size(), 16)), i_use = P.add((size_t)T * 4);
  HG_TRY(arena_ensure(p->cb, 2, P.total()));
  Carve cv;
  cv.init(p->cb, 2, P);
  char *blob = cv.get<char>(i_blob);
  if (!bl.data.empty()) HG_TRY(upload(p->cb, blob, bl.data.data(), bl.data.size()));
  p->use = cv.get<uint32_t>(i_use);
  HG_HIP(launch_fill_u32(p->use, 0u, T, p->stream));  // empty tables stay murmur3
  for (LaunchRmiBatch &b : p->rb) {
    b.keys = p->d_keys;
    b.items = (const RmiBatchItem *)(blob + o_items) + (uintptr_t)b.items;
    b.models = const_cast<void *>(bd.d_rmi_models);
    b.use = p->use;
    b.report = nullptr;
  }
  p->train = true;
  return HG_OK;
}

// Lays out both sides for cfg, then captures their launches into a graph.
hg_status plan_capture(hg_batch_plan p, const BatchCfg &cfg) {
  plan_drop_graph(p);
  HG_TRY(batch_prepare(p->cb, &p->bd, nullptr, cfg, p->sb));
  if (p->d_probe) HG_TRY(batch_prepare(p->cp, &p->bd, p->probe_off.data(), cfg, p->sp));
  HG_HIP(hipStreamSynchronize(p->stream));  // uploads done before capture
  HG_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  hg_status st = HG_OK;
  if (p->train) {
    // every table's model and its sampled check, then each side's bin maps
    // take the kind the check chose
    hipError_t e = hipSuccess;
    for (const LaunchRmiBatch &b : p->rb)
      if (e == hipSuccess) e = launch_rmi_train_batch(b, p->stream);
    if (e == hipSuccess)
      e = launch_set_kinds(p->use, p->sb.d.segs, (uint32_t)p->sb.L.segs.size(), p->sb.d.maps,
                           (uint32_t)p->sb.L.maps.size(), p->stream);
    if (e == hipSuccess && p->d_probe)
      e = launch_set_kinds(p->use, p->sp.d.segs, (uint32_t)p->sp.L.segs.size(), p->sp.d.maps,
                           (uint32_t)p->sp.L.maps.size(), p->stream);
    if (e != hipSuccess) st = fail(HG_ERR_DEVICE, std::string("plan training launch: ") + hipGetErrorString(e));
  }
  if (st == HG_OK) st = batch_launch_build(p->cb, &p->bd, p->sb, p->d_keys, p->d_offsets, p->d_csr);
  if (st == HG_OK && p->d_probe)
    st = batch_launch_probe(p->cp, &p->bd, p->sp, p->d_probe, p->d_offsets, p->d_csr, p->d_counts);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(p->stream, &g);
  if (st != HG_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return fail(HG_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
  p->graph = g;
  HG_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  p->cfg = cfg;
  return HG_OK;
}

hg_status plan_replay(hg_batch_plan p) {
  HG_HIP(hipGraphLaunch(p->exec, p->stream));
  return HG_OK;
}

}  // namespace

extern "C" {

hg_status hg_batch_build(hg_ctx c, const hg_batch_desc *bd, const void *d_keys, uint32_t *d_offsets, void *d_csr_keys) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(batch_validate(bd));
  if (!d_offsets || (bd->h_key_off[bd->n_tables] > bd->h_key_off[0] && (!d_keys || !d_csr_keys)))
    return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  BatchSide S;
  HG_TRY(batch_prepare(c, bd, nullptr, BatchCfg(), S));
  return batch_launch_build(c, bd, S, d_keys, d_offsets, d_csr_keys);
}

hg_status hg_batch_probe(hg_ctx c, const hg_batch_desc *bd, const uint32_t *d_offsets, const void *d_csr_keys,
                         const void *d_probe, const uint64_t *h_probe_off, uint64_t *d_counts, uint64_t *h_total) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(batch_validate(bd));
  if (!h_probe_off || !d_counts || !d_offsets) return fail(HG_ERR_INVALID_ARG, "null argument");
  for (uint32_t t = 0; t < bd->n_tables; t++)
    if (h_probe_off[t + 1] < h_probe_off[t]) return fail(HG_ERR_INVALID_ARG, "probe offsets must not decrease");
  DeviceGuard g(c->device);
  BatchSide S;
  HG_TRY(batch_prepare(c, bd, h_probe_off, BatchCfg(), S));
  HG_TRY(batch_launch_probe(c, bd, S, d_probe, d_offsets, d_csr_keys, d_counts));
  if (h_total) {
    std::vector<uint64_t> cnt(bd->n_tables);
    for (uint32_t t0 = 0; t0 < bd->n_tables; t0 += 65536) {
      const uint32_t k = std::min<uint32_t>(65536, bd->n_tables - t0);
      HG_TRY(sync_read(c, cnt.data() + t0, d_counts + t0, (size_t)k * 8));
    }
    uint64_t s = 0;
    for (uint64_t v : cnt) s += v;
    *h_total = s;
  }
  return HG_OK;
}

hg_status hg_batch_probe_emit(hg_ctx c, const hg_batch_desc *bd, const uint32_t *d_offsets, const void *d_csr_keys,
                              const void *d_probe, const uint64_t *h_probe_off, const uint64_t *h_out_off, void *d_out,
                              uint64_t *d_counts, uint64_t *h_total) {
  if (!c) return fail(HG_ERR_INVALID_ARG, "null ctx");
  HG_TRY(batch_validate(bd));
  if (!h_probe_off || !h_out_off || !d_counts || !d_offsets) return fail(HG_ERR_INVALID_ARG, "null argument");
  const uint32_t T = bd->n_tables;
  for (uint32_t t = 0; t < T; t++) {
    if (h_probe_off[t + 1] < h_probe_off[t]) return fail(HG_ERR_INVALID_ARG, "probe offsets must not decrease");
    if (h_out_off[t + 1] < h_out_off[t]) return fail(HG_ERR_INVALID_ARG, "output offsets must not decrease");
  }
  if (h_out_off[T] > h_out_off[0] && !d_out) return fail(HG_ERR_INVALID_ARG, "null output buffer");
  DeviceGuard g(c->device);
  BatchSide S;
  HG_TRY(batch_prepare(c, bd, h_probe_off, BatchCfg(), S));
  // the regions' offsets on the device (arena 2: unused by the batch's levels)
  Plan P;
  const size_t i_oo = P.add(((size_t)T + 1) * 8);
  HG_TRY(arena_ensure(c, 2, P.total()));
  Carve cv;
  cv.init(c, 2, P);
  uint64_t *d_oo = cv.get<uint64_t>(i_oo);
  HG_TRY(upload(c, d_oo, h_out_off, ((size_t)T + 1) * 8));
  // a dummy non-null buffer keeps the emitting kernels when every region is empty
  void *out = d_out ? d_out : (void *)d_oo;
  HG_TRY(batch_launch_probe(c, bd, S, d_probe, d_offsets, d_csr_keys, d_counts, out, d_oo));
  std::vector<uint64_t> cnt(T);
  for (uint32_t t0 = 0; t0 < T; t0 += 65536) {
    const uint32_t k = std::min<uint32_t>(65536, T - t0);
    HG_TRY(sync_read(c, cnt.data() + t0, d_counts + t0, (size_t)k * 8));
  }
  uint64_t total = 0;
  bool short_region = false;
  for (uint32_t t = 0; t < T; t++) {
    total += cnt[t];
    short_region |= cnt[t] > h_out_off[t + 1] - h_out_off[t];
  }
  if (h_total) *h_total = total;
  if (short_region) return fail(HG_ERR_CAPACITY, "a table's matches exceed its output region");
  return HG_OK;
}

hg_status hg_batch_plan_create(hg_ctx c, const hg_batch_desc *bd, const void *d_keys, uint32_t *d_offsets,
                               void *d_csr_keys, const void *d_probe, const uint64_t *h_probe_off, uint64_t *d_counts,
                               uint32_t autotune, hg_batch_plan *out) {
  if (!c || !out) return fail(HG_ERR_INVALID_ARG, "null argument");
  HG_TRY(batch_validate(bd));
  if (!d_offsets || (d_probe && (!h_probe_off || !d_counts))) return fail(HG_ERR_INVALID_ARG, "null buffer");
  DeviceGuard g(c->device);
  hg_batch_plan p = new hg_batch_plan_s();
  auto bail = [&](hg_status st) {
    hg_batch_plan_destroy(p);
    return st;
  };
  p->device = c->device;
  p->owner = c;
  const uint32_t T = bd->n_tables;
  p->key_off.assign(bd->h_key_off, bd->h_key_off + T + 1);
  p->n_bins.assign(bd->h_n_bins, bd->h_n_bins + T);
  p->off_off.assign(bd->h_off_off, bd->h_off_off + T);
  if (bd->h_hash_kind) p->kinds.assign(bd->h_hash_kind, bd->h_hash_kind + T);
  if (d_probe) p->probe_off.assign(h_probe_off, h_probe_off + T + 1);
  p->bd = *bd;
  p->bd.h_key_off = p->key_off.data();
  p->bd.h_n_bins = p->n_bins.data();
  p->bd.h_off_off = p->off_off.data();
  p->bd.h_hash_kind = bd->h_hash_kind ? p->kinds.data() : nullptr;
  p->d_keys = d_keys;
  p->d_offsets = d_offsets;
  p->d_csr = d_csr_keys;
  p->d_probe = d_probe;
  p->d_counts = d_counts;
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&p->t0) != hipSuccess || hipEventCreate(&p->t1) != hipSuccess)
    return bail(fail(HG_ERR_DEVICE, "plan stream/event creation failed"));
  hg_status st = hg_ctx_create(c->device, &p->cb);
  if (st == HG_OK) st = hg_ctx_create(c->device, &p->cp);
  if (st != HG_OK) return bail(st);
  p->cb->stream = p->stream;
  p->cp->stream = p->stream;
  // inputs are produced on the caller's stream
  if (hipEventRecord(p->ev_in, c->stream) != hipSuccess || hipStreamWaitEvent(p->stream, p->ev_in, 0) != hipSuccess)
    return bail(fail(HG_ERR_DEVICE, "stream ordering failed"));
  if (autotune & HG_PLAN_TRAIN) {
    st = plan_setup_training(p);
    if (st != HG_OK) return bail(st);
  }
  std::vector<BatchCfg> cands;
  if (autotune & HG_PLAN_AUTOTUNE) {
    for (double f : {0.875, 0.5})
      for (uint32_t tk : {kTile / 4, kTile, kTile * 4}) {
        BatchCfg x;
        x.fill = f;
        x.tile_keys = tk;
        x.train = p->train;
        cands.push_back(x);
      }
  } else {
    BatchCfg x;
    x.train = p->train;
    cands.push_back(x);
  }
  BatchCfg best = cands[0];
  float best_ms = -1.0f;
  if (cands.size() > 1) {
    for (const BatchCfg &x : cands) {
      st = plan_capture(p, x);
      if (st != HG_OK) return bail(st);
      float ms = 0.0f;
      if (plan_replay(p) != HG_OK || hipEventRecord(p->t0, p->stream) != hipSuccess) return bail(HG_ERR_DEVICE);
      for (int r = 0; r < 3; r++)
        if (plan_replay(p) != HG_OK) return bail(HG_ERR_DEVICE);
      if (hipEventRecord(p->t1, p->stream) != hipSuccess || hipEventSynchronize(p->t1) != hipSuccess ||
          hipEventElapsedTime(&ms, p->t0, p->t1) != hipSuccess)
        return bail(fail(HG_ERR_DEVICE, "timing a candidate failed"));
      ms /= 3.0f;
      if (best_ms < 0.0f || ms < best_ms) {
        best_ms = ms;
        best = x;
      }
    }
  }
  st = plan_capture(p, best);
  if (st != HG_OK) return bail(st);
  p->info.tile_keys = best.tile_keys;
  p->info.fill = (float)best.fill;
  p->info.candidates = (uint32_t)cands.size();
  p->info.tuned_ms = best_ms;
  *out = p;
  return HG_OK;
}

hg_status hg_batch_plan_run(hg_batch_plan p) {
  if (!p) return fail(HG_ERR_INVALID_ARG, "null plan");
  DeviceGuard g(p->device);
  HG_HIP(hipEventRecord(p->ev_in, p->owner->stream));
  HG_HIP(hipStreamWaitEvent(p->stream, p->ev_in, 0));
  HG_TRY(plan_replay(p));
  HG_HIP(hipEventRecord(p->ev_out, p->stream));
  HG_HIP(hipStreamWaitEvent(p->owner->stream, p->ev_out, 0));
  return HG_OK;
}

hg_status hg_batch_plan_query(hg_batch_plan p, hg_batch_plan_info *info) {
  if (!p || !info) return fail(HG_ERR_INVALID_ARG, "null argument");
  *info = p->info;
  info->learned_tables = 0;
  if (p->train) {  // the last replay's decisions
    DeviceGuard g(p->device);
    std::vector<uint32_t> use(p->bd.n_tables);
    HG_HIP(hipStreamSynchronize(p->stream));
    HG_HIP(hipMemcpy(use.data(), p->use, use.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t u : use) info->learned_tables += u ? 1 : 0;
  } else if (p->bd.h_hash_kind) {
    for (uint32_t k : p->kinds) info->learned_tables += k == HG_HASH_RMI ? 1 : 0;
  }
  return HG_OK;
}

hg_status hg_batch_plan_destroy(hg_batch_plan p) {
  if (!p) return HG_OK;
  DeviceGuard g(p->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  plan_drop_graph(p);
  if (p->cb) {
    p->cb->stream = nullptr;
    hg_ctx_destroy(p->cb);
  }
  if (p->cp) {
    p->cp->stream = nullptr;
    hg_ctx_destroy(p->cp);
  }
  for (hipEvent_t e : {p->ev_in, p->ev_out, p->t0, p->t1})
    if (e) (void)hipEventDestroy(e);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return HG_OK;
}

}  // extern "C"
