// hg_kernels.hpp -- host-side launch interface of the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include "hg_common.hpp"

namespace hg {

// Kernel specialisations: hash kind (0 murmur3, 1 learned, 2 mixed) and
// bucket kind (0 range, 1 split, 2 mixed) of every segment of a launch.
enum : int { KIND_MIXED = 2 };

struct LaunchPart {
  int key_bytes;
  int hk, sk;            // hash kind, bucket kind
  uint32_t ntiles;
  uint32_t max_cols;     // largest ncols over the segments
  uint32_t stage_cap;    // keys staged per LDS round in the scatter (part_stage_cap)
  const void *keys;
  void *out;
  const TileDesc *tiles;
  const SegDesc *segs;
  uint32_t *cmat;        // counts after hist, positions after the column scan
  // keys are 32-bit and the level writes 8-byte records (input position << 32
  // | key) of them (key_bytes 8, murmur3 only): index-carrying builds
  bool pack = false;
};

struct LaunchCol {
  uint32_t nitems;       // chunk items (seg, chunk, column block)
  uint32_t nitems2;      // column-block items (seg, column block)
  uint32_t nsegs;
  uint32_t *cmat;
  const ColItem *items;
  const ColItem *items2;
  const SegDesc *segs;
  const uint64_t *psum_base;
  const uint32_t *nchunks;
  uint32_t *psum;
  uint32_t *coltot;
  uint32_t *bstart;
};

struct LaunchFinal {
  int key_bytes, off_bytes;
  int hk;
  uint32_t nitems;
  uint32_t lds_bytes;
  const void *part;
  void *csr;
  void *off;
  const FinItem *items;
  const BinMap *maps;
  // oversized buckets: per-item slot (nitems), control block, big_cap per-bin
  // counters (buckets whose bins do not fit take the single-workgroup path)
  uint32_t *big_slot;
  void *big_ctl;
  uint32_t *big_cnt;
  uint32_t big_cap;
  // index-carrying build of 32-bit keys: part holds 8-byte records (input
  // position << 32 | key); csr then receives the 32-bit keys and csr_index the
  // positions (murmur3 tables only)
  uint64_t *csr_index = nullptr;
};

struct LaunchProbe {
  int key_bytes, off_bytes;
  int hk;
  uint32_t nitems;
  uint32_t akey_cap;
  const void *bpart;
  const FinItem *items;
  const BinMap *maps;
  const void *offA;
  const void *keysA;
  void *out;
  uint64_t capacity;
  unsigned long long *cursor;  // match counters, one per table (FinItem.table)
  unsigned int *hmark;         // nitems, zeroed: items with a bin beyond the walk limit
  unsigned int *gate;          // 2 zeroed words: [0] k_probe1 left items, [1] a bin was marked heavy
  uint32_t *big_slot;          // nitems words and a final_big_ctl_bytes() area: the list of the
  void *big_ctl;               // oversized buckets (written by the probe itself)
  const uint64_t *out_off;     // null: out[0 .. capacity) for every table; else table t's
                               // matches go to out[out_off[t] .. out_off[t+1]) (device array)
};

struct LaunchJoin {
  int key_bytes, off_bytes;
  uint64_t M;
  uint32_t key_cap;
  const void *offA, *keysA, *offB, *keysB;
  void *out;
  uint64_t capacity;
  unsigned long long *cursor;
  // bins above the nested loop's pair limit: list (hlist_cap entries) and its
  // zeroed counter
  uint32_t *hlist;
  unsigned int *hlist_n;
  uint32_t hlist_cap;
  // row-id join (pa != null): the entries' input positions, and the pairs
  // (A position, B position) written instead of keys; listed bins are left
  // to launch_join_pairs_heavy
  const uint64_t *indA = nullptr, *indB = nullptr;
  uint64_t *pa = nullptr, *pb = nullptr;
};

// The listed bins of a row-id join, compared pair by pair in tiles.
struct LaunchJoinPairs {
  int key_bytes, off_bytes;
  const void *offA, *keysA, *offB, *keysB;
  const uint64_t *indA, *indB;
  const uint32_t *hlist;
  uint32_t nh;
  const unsigned long long *tile_base;  // exclusive scan of the bins' tiles (nh + 1)
  uint64_t ntiles;
  uint64_t *pa, *pb;
  uint64_t capacity;
  unsigned long long *cursor;
};

// Moves the run-length join's records and heavy bins of every chunk from
// the padded places a single pass wrote them to (scans of their bounds) to
// the compact places (scans of their counts).
struct LaunchRleGather {
  int key_bytes;
  uint32_t nch;
  const unsigned long long *pbase, *cbase, *cnt;  // padded base, compact base, records per chunk
  const void *srk;
  const unsigned long long *srca, *srcb;
  void *rk;
  unsigned long long *rca, *rcb;
  uint64_t capacity;
  const uint32_t *phbase, *hbase, *hcnt;  // the same for the heavy bins
  const uint32_t *sheavy;
  uint32_t *heavy;
  uint64_t heavy_cap;
  unsigned int *bad;  // set when a chunk exceeded its bound
};

struct LaunchRle {
  int key_bytes, off_bytes;
  int pass;  // 0 count small bins, 1 emit small bins, 2 heavy bins,
             // 3 count and emit small bins at padded bases (chunk_base, heavy_base)
  uint64_t M;
  const void *offA, *keysA, *offB, *keysB;
  unsigned long long *chunk_cnt;         // per chunk: records (pass 0)
  const unsigned long long *chunk_base;  // its exclusive scan
  uint32_t *chunk_heavy;                 // per chunk: heavy bins (pass 0)
  const uint32_t *heavy_base;            // its exclusive scan
  unsigned long long *chunk_pairs;       // per chunk: pairs (pass 0)
  void *rk;
  unsigned long long *rca, *rcb;
  uint64_t capacity;
  uint32_t *heavy;
  uint64_t heavy_cap, nheavy;
  unsigned long long *cursor, *pairs;
  unsigned int *overflow;
  // giant heavy bins: per-heavy-bin slot (heavy_cap), control block, tables
  uint32_t *gmark;  // 2 * heavy_cap: the heavy bins of k_rle_heavy's two forms
  void *gctl;
  void *gkey;
  unsigned long long *gca, *gcb;
  // k_rle_heavy's private record regions (rle_priv_records each) and counts (rle_heavy_grid)
  void *pkey;
  unsigned long long *pca, *pcb;
  uint32_t *priv_n;
  // heavy bins with more distinct values than the LDS / giant tables hold:
  // list (heavy_cap entries) and its zeroed counter
  uint32_t *ovf_list;
  unsigned int *ovf_n;
};

// One batch of the run-length join's global-table path (value-rich bins).
struct LaunchRleOvf {
  int key_bytes;
  const void *keysA, *keysB;
  const OvfChunk *chunksA, *chunksB;
  uint32_t nchunksA, nchunksB;
  const OvfTable *tabs;
  uint32_t ntabs;
  uint64_t nslots;           // table slots of the batch
  void *tkey;                // nslots keys (cleared to all ones)
  uint32_t *tca, *tcb;       // nslots counts each (zeroed)
  unsigned long long *eA, *eB;  // ntabs sentinel-key counts each (zeroed)
  void *rk;
  unsigned long long *rca, *rcb;
  uint64_t capacity;
  unsigned long long *cursor, *pairs;
};

struct LaunchRmi {
  int key_bytes;
  const void *keys;
  uint64_t n_keys, n_sample;
  uint64_t s0;               // mix64(seed)
  void *buf0, *buf1;         // n_sample keys each
  uint32_t *counts;          // 256 * sort tiles + 1
  void *scan_status;         // scan_status_bytes(256 * sort tiles)
  double *partials;          // 8 * nblk
  void *cand;                // 4 root candidates (16 bytes each)
  uint32_t *leafcnt;         // 4 * n_leaves
  uint32_t nblk;
  uint32_t n_leaves;
  BinMap map;                // murmur3 map of the table's bins
  void *model;
  uint32_t *start;           // n_leaves + 1
  uint32_t Mp;               // bins of the sampled check
  uint32_t *cnt_r, *cnt_m;   // Mp each
  unsigned long long *out;   // 4
};

hipError_t launch_rmi_train(const LaunchRmi &r, hipStream_t st);

// Batched training, one workgroup per table with everything in LDS: samples
// up to RMI_ONE_MAX_KEY_BYTES of keys, up to RMI_ONE_MAX_LEAVES leaves and
// RMI_ONE_MAX_CHECK_BINS bins in the sampled check.
constexpr uint32_t RMI_ONE_MAX_KEY_BYTES = 64 * 1024;
constexpr uint32_t RMI_ONE_MAX_LEAVES = 1024;
constexpr uint32_t RMI_ONE_MAX_CHECK_BINS = 12288;
constexpr uint32_t RMI_ONE_MAX_BLOCKS = 64;  // moment blocks of 256 keys
struct RmiBatchItem {
  uint64_t key_begin;  // the table's first key
  uint64_t n_keys;
  uint64_t s0;         // mix64(seed)
  uint64_t model_off;  // byte offset of the table's model
  BinMap map;          // murmur3 map of the table's bins
  uint32_t ns;         // sample keys
  uint32_t np2;        // ns rounded up to a power of two
  uint32_t Mp;         // bins of the sampled check
  uint32_t table;      // index of use[] and report[]
};
struct LaunchRmiBatch {
  int key_bytes;
  uint32_t n_items, n_leaves;
  uint32_t keys_cap;   // max np2 of the items
  uint32_t cnt_cap;    // max(4 * n_leaves, max Mp)
  const void *keys;
  const RmiBatchItem *items;
  void *models;
  uint32_t *use;                  // per table: 1 = the model passed the check
  unsigned long long *report;     // per table: 4 figures, or null
};
size_t rmi_one_lds(uint32_t keys_cap, int key_bytes, uint32_t L, uint32_t cnt_cap);
hipError_t launch_rmi_train_batch(const LaunchRmiBatch &b, hipStream_t st);
hipError_t set_rmi_kernel_attributes();
// maps[t].kind and each segment's map kind from use[] (learned when set)
hipError_t launch_set_kinds(const uint32_t *use, SegDesc *segs, uint32_t nsegs, BinMap *maps, uint32_t nmaps,
                            hipStream_t st);
uint32_t rmi_moment_blocks(uint64_t n_sample);
uint32_t rmi_sort_tiles(uint64_t n_sample);
hipError_t launch_rle(const LaunchRle &r, hipStream_t st);
// per chunk of the run-length join: records <= min(keys of A, keys of B),
// heavy bins <= keys of A / 65 + keys of B / 65 (from the offsets alone)
hipError_t launch_rle_caps(const void *offA, const void *offB, int off_bytes, uint64_t M,
                           unsigned long long *cap_rec, uint32_t *cap_h, hipStream_t st);
hipError_t launch_rle_gather(const LaunchRleGather &g, hipStream_t st);
hipError_t launch_rle_ovf(const LaunchRleOvf &o, hipStream_t st);
// (a0, a1, x0, x1) of each listed bin -> out (4 entries per bin)
hipError_t launch_ovf_sizes(const uint32_t *list, uint32_t n, const void *offA, const void *offB, int off_bytes,
                            unsigned long long *out, hipStream_t st);
uint64_t rle_chunks(uint64_t M);
// scratch bytes of the giant-bin path of the run-length join
size_t rle_giant_bytes(int key_bytes);
size_t rle_giant_slots();
size_t rle_priv_records();
uint32_t rle_heavy_grid();
size_t rle_giant_ctl_bytes();
hipError_t launch_part_hist(const LaunchPart &lp, hipStream_t st);
hipError_t launch_part_scatter(const LaunchPart &lp, hipStream_t st);
hipError_t launch_colscan(const LaunchCol &lc, hipStream_t st);
hipError_t launch_final(const LaunchFinal &lf, hipStream_t st);
hipError_t launch_probe(const LaunchProbe &p, hipStream_t st);
hipError_t launch_join(const LaunchJoin &j, hipStream_t st);
// tiles (1024 x 1024 entries) and key pairs (summed, each capped at 2^37) of listed bins
hipError_t launch_pair_tiles(const void *offA, const void *offB, int off_bytes, const uint32_t *hlist, uint32_t n,
                             unsigned long long *tiles, unsigned long long *pairs, hipStream_t st);
hipError_t launch_join_pairs_heavy(const LaunchJoinPairs &p, hipStream_t st);
// exclusive scan of n elements (total at out[n]); status: scan_status_bytes(n)
hipError_t launch_scan(const void *in, void *out, uint64_t n, int elem_bytes, void *status, hipStream_t st);
size_t scan_status_bytes(uint64_t n);
hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, uint64_t n, hipStream_t st);
// index-carrying builds: rec[i] = (i << 32) | low32(key i); unpack gathers
// csr_keys[j] = keys[rec[j] >> 32] and csr_index[j] = rec[j] >> 32
hipError_t launch_pack_index(const void *keys, uint64_t n, int key_bytes, uint64_t *rec, hipStream_t st);
hipError_t launch_unpack_index(const uint64_t *rec, uint64_t n, const void *keys, int key_bytes, void *csr_keys,
                               uint64_t *csr_index, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st);
// off[i] += delta modulo 2^(8 off_bytes), i < n
hipError_t launch_offsets_add(void *off, int off_bytes, uint64_t n, uint64_t delta, hipStream_t st);
hipError_t launch_bins(const void *keys, uint64_t n, int key_bytes, const BinMap &mp, uint32_t *bins, hipStream_t st);
hipError_t launch_range_hist(const void *keys, uint64_t n, int key_bytes, const BinMap &mp, uint32_t range,
                             uint32_t nranges, unsigned long long *hist, unsigned int *bad, hipStream_t st);
hipError_t set_kernel_attributes();
hipError_t launch_gen_uniform(void *out, uint64_t n, int key_bytes, uint64_t s0, uint64_t range, hipStream_t st);
// descriptors of a one-segment level (sd.ntiles tiles) written by a kernel
hipError_t launch_level_setup(const SegDesc &sd, const BinMap &map, uint32_t tile_keys, uint32_t chunk, TileDesc *tiles,
                              SegDesc *segs, ColItem *items, ColItem *items2, uint64_t *psum_base, uint32_t *nchunks,
                              BinMap *maps, hipStream_t st);
hipError_t launch_make_items(const SegDesc *segs, uint32_t nsegs, const uint32_t *bstart, const uint32_t *coltot,
                             FinItem *items, uint32_t ncols_total, hipStream_t st);

int final_kpt(int key_bytes);
uint32_t final_big_items();
size_t final_big_ctl_bytes();
// keys staged per LDS round of the scatter for a level whose widest segment
// has max_cols columns (0: the columns do not leave room for staging)
uint32_t part_stage_cap(uint32_t max_cols, int key_bytes, int sk);
uint32_t probe_akey_cap(int key_bytes, int off_bytes);
uint32_t join_key_cap(int key_bytes, int off_bytes, bool emit, bool with_index = false);
uint64_t scan_tiles(uint64_t n);

}  // namespace hg
