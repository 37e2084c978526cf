/*
 * hashgraph.h -- C ABI of the MI355X HashGraph library (libhashgraph_amd.so).
 *
 * A HashGraph is a hash table stored as a CSR graph: bin b's keys are
 * keys[offsets[b] .. offsets[b+1]), offsets[M] = N.  bin(k) is
 * murmur3_x86_32(low 32 bits of k, seed 0) mod M, as in the reference
 * (include/SingleHashGraphOperators.cuh:71-79,
 *  include/MultiHashGraphDeviceOperators.cuh:76-83), or the learned CDF model.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Pointers named d_* are device pointers
 *    on the context's device; h_* are host pointers.
 *  - Every call is ordered on the context's stream (hg_ctx_set_stream; the
 *    null stream by default).  Calls that return a host value (h_count ...)
 *    synchronise that stream before returning.
 *  - The caller allocates inputs and outputs.  Temporary device memory lives
 *    in a per-context arena that grows on demand (hg_ctx_reserve pre-sizes it).
 *  - No call exits or throws; each returns hg_status, and hg_last_error()
 *    describes the last failure on the calling thread.
 *  - Bin order inside a bin is unspecified (the reference fills bins with
 *    atomics); offsets are exact.
 */
#ifndef HASHGRAPH_H
#define HASHGRAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  HG_OK = 0,
  HG_ERR_INVALID_ARG = 1,  /* bad sizes, null pointers, mismatched tables   */
  HG_ERR_OOM = 2,          /* scratch arena could not grow                  */
  HG_ERR_DEVICE = 3,       /* HIP runtime error                             */
  HG_ERR_CAPACITY = 4,     /* output buffer smaller than the result         */
  HG_ERR_MISMATCH = 5,     /* two tables do not share a bin space           */
  HG_ERR_UNSUPPORTED = 6   /* size or mode outside what this build handles */
} hg_status;

typedef enum { HG_HASH_MURMUR3 = 0, HG_HASH_RMI = 1 } hg_hash_kind;

typedef struct hg_ctx_s *hg_ctx;

/* Table description shared by every call.
 *  key_bytes    4 (uint32 keys) or 8 (uint64 keys; bins use the low 32 bits
 *               under murmur, as the reference does)
 *  offset_bytes 4 (uint32 offsets) or 8 (uint64 offsets)
 *  n_keys       N, number of keys
 *  n_bins       M, number of bins, 1 <= M < 2^32
 *  hash_kind    HG_HASH_MURMUR3 or HG_HASH_RMI
 *  d_rmi_model  device pointer to an hg_rmi_model (HG_HASH_RMI only)      */
typedef struct {
  uint32_t key_bytes;
  uint32_t offset_bytes;
  uint64_t n_keys;
  uint64_t n_bins;
  uint32_t hash_kind;
  uint32_t reserved;
  const void *d_rmi_model;
} hg_table_desc;

/* ---- context ----------------------------------------------------------- */
hg_status hg_ctx_create(int device, hg_ctx *out);
hg_status hg_ctx_destroy(hg_ctx ctx);
/* stream is a hipStream_t (NULL = the null stream). */
hg_status hg_ctx_set_stream(hg_ctx ctx, void *stream);
/* Grow the scratch arena to at least bytes now (no allocation later for
 * calls whose scratch fits). */
hg_status hg_ctx_reserve(hg_ctx ctx, size_t bytes);
/* Device bytes of scratch a build / probe of this description needs. */
hg_status hg_scratch_bytes(const hg_table_desc *desc, int op, size_t *bytes);
const char *hg_last_error(void);
const char *hg_version(void);

/* Per-kernel timing (no counterpart in the reference, whose drivers time
 * whole calls with cudaEvents).  When enabled, every kernel launch of the
 * context is bracketed by hipEvents on the context's stream; read returns,
 * per kernel name, the number of launches and the summed event time since
 * the last reset (it synchronises the stream).                            */
typedef struct {
  char name[32];
  uint64_t launches;
  double total_ms;
} hg_kernel_time;
hg_status hg_ctx_profile(hg_ctx ctx, int enable);
hg_status hg_ctx_profile_read(hg_ctx ctx, hg_kernel_time *out, uint32_t capacity, uint32_t *h_count, int reset);

/* ---- build --------------------------------------------------------------
 * Replaces SingleHashGraph::build (src/SingleHashGraph.cu:216-232, the LRB
 * build at :129-181) and the per-GPU CSR build of MultiHashGraph::build
 * (src/MultiHashGraph.cu:303-375).
 *  d_keys     N keys
 *  d_offsets  M+1 offsets (offset_bytes each); offsets[M] = N
 *  d_csr_keys N keys grouped by bin                                        */
hg_status hg_build(hg_ctx ctx, const hg_table_desc *desc, const void *d_keys, void *d_offsets,
                   void *d_csr_keys);

/* ---- build without allocation ---------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same result as hg_build, but every temporary lives
 * in a static device area of 40 MiB that the library holds on each device
 * from load to process end; the call allocates nothing (no device or pinned
 * memory, no upload from the host).  Tables needing more scratch than the
 * area holds are refused with HG_ERR_UNSUPPORTED: up to 2^22 keys of either
 * width with up to 2^28 bins always fit.  All calls on one device share the
 * area, so calls that overlap on one device race (order them on one stream). */
hg_status hg_build_noalloc(hg_ctx ctx, const hg_table_desc *desc, const void *d_keys, void *d_offsets,
                           void *d_csr_keys);

/* ---- intersection of two built tables -----------------------------------
 * Replaces MultiHashGraph::intersect (src/MultiHashGraph.cu:631-740, kernel
 * simpleIntersect include/MultiHashGraphDeviceOperators.cuh:513-543): for
 * every bin, every pair (a in A_bin, x in B_bin) with a == x counts once and
 * emits x.  Count = sum over values v of cA(v)*cB(v).  Both tables must
 * share key_bytes, hash_kind and n_bins.
 *  h_count    total matches (host, exact 64-bit)
 *  d_out      NULL for count only; else capacity keys: the same multiset as
 *             the reference's list, grouped by bin ranges (the order inside
 *             the list is unspecified, as bin contents are)
 *  Every bin is answered, however heavy: a bin pair above 4096 key pairs is
 *  joined through a value table of A's bin instead of the nested loop.
 *  Returns HG_ERR_CAPACITY (with *h_count set) when capacity < count.     */
hg_status hg_intersect(hg_ctx ctx, const hg_table_desc *descA, const void *d_offA, const void *d_keysA,
                       const hg_table_desc *descB, const void *d_offB, const void *d_keysB, void *d_out,
                       uint64_t capacity, uint64_t *h_count);

/* ---- probe: raw keys against a built table -------------------------------
 * Same result as building d_probe into a table B over A's bins and calling
 * hg_intersect(A, B) (same count, same multiset of emitted keys: each probe
 * key x is emitted cA(x) times), without materialising B's CSR.  Output order
 * is grouped by buckets of A's bins.  Any table hg_build accepts can be
 * probed (B is partitioned in two levels for tables beyond ~2^28 bins).    */
hg_status hg_probe(hg_ctx ctx, const hg_table_desc *descA, const void *d_offA, const void *d_keysA,
                   const void *d_probe, uint64_t n_probe, void *d_out, uint64_t capacity, uint64_t *h_count);

/* ---- run-length intersection ----------------------------------------------
 * Heavy duplicates make the emitted list of hg_intersect infeasible (a value
 * with 10^8 copies in both tables yields 10^16 pairs).  This form writes one
 * record per distinct value present in both tables: key (key_bytes),
 * count in A and count in B (uint64 each).  The records of bins with at most
 * 64 keys per side come first, in bin order; those of heavier bins follow.
 *  h_count    sum cA*cB (exact below 2^64)
 *  h_records  number of records                                           */
hg_status hg_intersect_rle(hg_ctx ctx, const hg_table_desc *descA, const void *d_offA, const void *d_keysA,
                           const hg_table_desc *descB, const void *d_offB, const void *d_keysB, void *d_rec_keys,
                           uint64_t *d_rec_ca, uint64_t *d_rec_cb, uint64_t capacity, uint64_t *h_records,
                           uint64_t *h_count);

/* ---- logical partitions (one GPU's share of MultiHashGraph) ---------------
 * hg_find_splits replaces countBinSizes + binarySearch
 * (include/MultiHashGraphHostOperators.cuh:8-108): a histogram of
 * bin / binRange over bin_count ranges (binRange = ceil(M / (float)bin_count)),
 * key-balanced split points h_splits[0..n_parts] (int64, host).            */
hg_status hg_find_splits(hg_ctx ctx, const hg_table_desc *desc, const void *d_keys, uint64_t bin_count,
                         uint32_t n_parts, int64_t *h_splits);

/* hg_build_partitioned replaces the local partition + per-partition build of
 * MultiHashGraph::build (src/MultiHashGraph.cu:408-629 with
 * include/MultiHashGraphDeviceOperators.cuh:220-246,297-345,371-458):
 * partition p holds the keys whose bin lies in [h_splits[p], h_splits[p+1])
 * and gets its own CSR over local bins bin - h_splits[p].
 *  d_offsets      partition p's R_p+1 offsets start at element
 *                 h_splits[p] + p (R_p = h_splits[p+1] - h_splits[p]); each
 *                 partition's offsets start at 0
 *  d_csr_keys     all partitions' keys, partition after partition
 *  h_part_offsets n_parts+1 host entries: partition p's keys are
 *                 d_csr_keys[h_part_offsets[p] .. h_part_offsets[p+1])
 * Keys whose bin is >= h_splits[n_parts] are dropped, as in the reference. */
hg_status hg_build_partitioned(hg_ctx ctx, const hg_table_desc *desc, const void *d_keys, const int64_t *h_splits,
                               uint32_t n_parts, void *d_offsets, void *d_csr_keys, uint64_t *h_part_offsets);

/* ---- builds that carry each key's input position (INDEX_TRACK) -------------
 * Replace the reference's INDEX_TRACK mode, whose CSR entries are
 * keyval_ind {key, ind} (include/MultiHashGraph.cuh:62,82-87,103) with
 * ind = the key's position in the input (binHashValues,
 * include/MultiHashGraphDeviceOperators.cuh:297-345).  Same results as
 * hg_build / hg_build_partitioned, plus d_csr_index (uint64, one per CSR key):
 * d_csr_keys[j] == d_keys[d_csr_index[j]].  Murmur3 tables only (their bins
 * depend on the low 32 key bits, which the build carries beside the
 * position); n_keys < 2^32.                                               */
hg_status hg_build_indexed(hg_ctx ctx, const hg_table_desc *desc, const void *d_keys, void *d_offsets,
                           void *d_csr_keys, uint64_t *d_csr_index);
hg_status hg_build_partitioned_indexed(hg_ctx ctx, const hg_table_desc *desc, const void *d_keys,
                                       const int64_t *h_splits, uint32_t n_parts, void *d_offsets, void *d_csr_keys,
                                       uint64_t *d_csr_index, uint64_t *h_part_offsets);

/* ---- row-id join of two index-carrying tables (INDEX_TRACK) --------------
 * The join of MultiHashGraph::intersect (src/MultiHashGraph.cu:631-740) over
 * tables whose entries carry their input positions (keyval_ind,
 * include/MultiHashGraph.cuh:82-87; built by hg_build_indexed): for every
 * matching pair (a in A_bin, x in B_bin, a == x), the pair of their input
 * positions (d_pos_a[w], d_pos_b[w]), the index-pair output the reference
 * wrote for its simpleIntersect (include/MultiHashGraphDeviceOperators.cuh:
 * 535-537, commented out there).  Same count as hg_intersect; the pairs form
 * the same multiset whatever the bins' order, grouped by bin ranges for bins
 * up to 4096 key pairs.
 *  d_indA, d_indB    the tables' d_csr_index (uint64, one per CSR key)
 *  d_pos_a, d_pos_b  NULL for count only; else capacity uint64 each
 * Bins above 4096 key pairs are compared pair by pair in 1024 x 1024 tiles
 * over the grid; more than 2^36 such pairs in all is HG_ERR_UNSUPPORTED
 * (*h_count still set).  HG_ERR_CAPACITY (with *h_count set) when
 * capacity < count.                                                       */
hg_status hg_intersect_indexed(hg_ctx ctx, const hg_table_desc *descA, const void *d_offA, const void *d_keysA,
                               const uint64_t *d_indA, const hg_table_desc *descB, const void *d_offB,
                               const void *d_keysB, const uint64_t *d_indB, uint64_t *d_pos_a, uint64_t *d_pos_b,
                               uint64_t capacity, uint64_t *h_count);

/* ---- staged tables: larger than HBM ----------------------------------------
 * Replace the reference's MANAGED_MEM mode (src/MultiHashGraph.cu:229-247,
 * include/MultiHashGraphHostOperators.cuh:255-280; intersect
 * src/MultiHashGraph.cu:664-670), which keeps a table in one
 * cudaMallocManaged blob and lets page faults move it.  Here keys, offsets,
 * CSR keys and matched lists stay in host memory (pageable, pinned, or
 * managed) and pass through HBM in chunks and bin-range slabs by sequential
 * copies; results are those of hg_build / hg_build_partitioned /
 * hg_find_splits / hg_intersect on the same inputs.
 *  device_budget  bytes of device memory the call's buffers may take
 *                 (0 = half of the free memory); at least 4 MiB.  Buffers
 *                 stay in the context's scratch until hg_ctx_destroy.
 *  n_keys may reach 2^32 and beyond (offset_bytes 8).  A build slab of k keys
 *  and r bins takes about k (6 key_bytes + 1) + r (2 offset_bytes + 4) bytes
 *  of 0.9 of the budget; a join slab up to budget / (4 key_bytes) keys of A
 *  and B together.  A run of bins with more keys than one slab holds is
 *  refused with HG_ERR_UNSUPPORTED, as is a single bin with more matches than
 *  budget / (8 key_bytes).  Offsets buffers must be host-accessible (device-
 *  only memory is HG_ERR_INVALID_ARG).  The call returns after its last copy.
 * hg_build_staged       offsets (n_bins+1) and CSR keys in host memory; the
 *                       CSR keys buffer also stages the keys between passes
 * hg_build_partitioned_staged   hg_build_partitioned's layout; n_parts <= 2048
 * hg_find_splits_staged         hg_find_splits over host keys
 * hg_intersect_staged   two host tables of one bin space; h_out (host, may be
 *                       NULL for count only) gets the matched keys, bin range
 *                       after bin range; HG_ERR_CAPACITY (with *h_count set)
 *                       when capacity < count, the list then incomplete.   */
hg_status hg_build_staged(hg_ctx ctx, const hg_table_desc *desc, const void *h_keys, void *h_offsets,
                          void *h_csr_keys, size_t device_budget);
hg_status hg_build_partitioned_staged(hg_ctx ctx, const hg_table_desc *desc, const void *h_keys,
                                      const int64_t *h_splits, uint32_t n_parts, void *h_offsets, void *h_csr_keys,
                                      uint64_t *h_part_offsets, size_t device_budget);
hg_status hg_find_splits_staged(hg_ctx ctx, const hg_table_desc *desc, const void *h_keys, uint64_t bin_count,
                                uint32_t n_parts, int64_t *h_splits, size_t device_budget);
hg_status hg_intersect_staged(hg_ctx ctx, const hg_table_desc *descA, const void *h_offA, const void *h_keysA,
                              const hg_table_desc *descB, const void *h_offB, const void *h_keysB, void *h_out,
                              uint64_t capacity, uint64_t *h_count, size_t device_budget);

/* ---- learned hash (RMI) ---------------------------------------------------
 * No counterpart in the reference.  A two-stage piecewise-linear CDF model in
 * FP32: a linear root picks one of n_leaves leaves, each leaf's line maps the
 * key's distance from the leaf's first key to a bin; bases are monotone, so
 * the bin order follows the key order.  A key whose line value falls outside
 * [0, n_bins - 1] (beyond the range the model was trained on) takes murmur3's
 * bin, so out-of-range keys spread over the table instead of filling an edge
 * bin.  Layout: 32-byte header, then
 * n_leaves leaves of 16 bytes.
 *
 * hg_rmi_train draws n_sample keys of d_keys (with replacement, SplitMix64
 * stream `seed`), sorts them, fits root and leaves by least squares in FP64
 * on the device, stores the model in FP32 at d_model, then bins the sample
 * with the model and with murmur3 at the table's load factor (n_bins *
 * n_sample / n_keys bins).  *h_use_rmi = 1 when the model's max bin and its
 * sum of squared bin sizes (the mean probe length times n_sample) are both
 * no larger than murmur's; 0 means murmur is the better hash (the fallback).
 * h_report (optional) receives the four figures.  n_leaves in [1, 65536],
 * n_sample in [1, 2^31).  d_model needs hg_rmi_model_bytes(n_leaves).     */
typedef struct {
  uint64_t sample_bins;     /* bins the sample was binned into */
  uint64_t max_bin_rmi, max_bin_murmur;
  double mean_probe_rmi, mean_probe_murmur; /* sum of squared bin sizes / n_sample */
} hg_rmi_report;

size_t hg_rmi_model_bytes(uint32_t n_leaves);
hg_status hg_rmi_train(hg_ctx ctx, const void *d_keys, uint64_t n_keys, uint32_t key_bytes, uint64_t n_sample,
                       uint64_t seed, uint64_t n_bins, uint32_t n_leaves, void *d_model, uint32_t *h_use_rmi,
                       hg_rmi_report *h_report);
/* One model per table of a batch: table t's keys are
 * d_keys[h_key_off[t] .. h_key_off[t+1]), its model (hg_rmi_model_bytes(n_leaves)
 * bytes) goes to d_models + t * hg_rmi_model_bytes(n_leaves), and h_use[t] /
 * h_report[t] (optional) are hg_rmi_train's outputs for (h_n_sample[t],
 * h_seed[t], h_n_bins[t]).  The models equal hg_rmi_train's bit for bit.
 * Tables whose sample fits one workgroup's LDS (up to 16384 uint32 or 8192
 * uint64 keys, up to 1024 leaves, up to 12288 check bins) train together in
 * one launch; larger ones table by table.  An empty table gets h_use[t] = 0
 * and no model.                                                           */
hg_status hg_rmi_train_batch(hg_ctx ctx, uint32_t n_tables, uint32_t key_bytes, const void *d_keys,
                             const uint64_t *h_key_off, const uint64_t *h_n_bins, const uint64_t *h_n_sample,
                             const uint64_t *h_seed, uint32_t n_leaves, void *d_models, uint32_t *h_use,
                             hg_rmi_report *h_report);

/* ---- batched form -----------------------------------------------------------
 * No counterpart in the reference.  Builds n_tables independent tables (and
 * optionally probes one key set against each) in one set of launches.
 * Table t: keys d_keys[key_off[t] .. key_off[t+1]), n_bins[t] bins, offsets
 * at d_offsets[off_off[t] .. off_off[t] + n_bins[t]], CSR keys at
 * d_csr_keys[key_off[t] ..].  Arrays h_* are host arrays of n_tables(+1).   */
typedef struct {
  uint32_t n_tables;
  uint32_t key_bytes;
  const uint64_t *h_key_off;   /* n_tables + 1 */
  const uint64_t *h_n_bins;    /* n_tables     */
  const uint64_t *h_off_off;   /* n_tables     */
  const uint32_t *h_hash_kind; /* n_tables, NULL = all murmur */
  const void *d_rmi_models;    /* n_tables models of rmi_leaves leaves each, or NULL */
  uint32_t rmi_leaves;
} hg_batch_desc;

hg_status hg_batch_build(hg_ctx ctx, const hg_batch_desc *bd, const void *d_keys, uint32_t *d_offsets,
                         void *d_csr_keys);
/* Probe set of table t: d_probe[probe_off[t] .. probe_off[t+1]).  Writes the
 * per-table match counts to d_counts (uint64, n_tables) and their sum to
 * *h_total. */
hg_status hg_batch_probe(hg_ctx ctx, const hg_batch_desc *bd, const uint32_t *d_offsets, const void *d_csr_keys,
                         const void *d_probe, const uint64_t *h_probe_off, uint64_t *d_counts, uint64_t *h_total);
/* The same probe with the matches written (hg_probe's output, per table):
 * table t's matched probe keys (each once per equal key of the table) go to
 * d_out[h_out_off[t] .. h_out_off[t+1]) in no particular order.  d_counts
 * and *h_total as above.  A table with more matches than its region gets the
 * region filled and the call returns HG_ERR_CAPACITY (counts stay exact);
 * regions from the exclusive prefix sum of a counting call always fit.
 * h_out_off: host array of n_tables + 1 non-decreasing entries.            */
hg_status hg_batch_probe_emit(hg_ctx ctx, const hg_batch_desc *bd, const uint32_t *d_offsets, const void *d_csr_keys,
                              const void *d_probe, const uint64_t *h_probe_off, const uint64_t *h_out_off, void *d_out,
                              uint64_t *d_counts, uint64_t *h_total);

/* Plans: the batch's descriptors stay resident on the device and its launches
 * (build, then the probe counts when d_probe is given) are captured once into
 * a hipGraph; hg_batch_plan_run replays it, ordered after the work already
 * queued on the context's stream and before work queued later.  With
 * autotune & HG_PLAN_AUTOTUNE the plan times candidate tile sizes (8 K, 32 K, 128 K keys per
 * partition workgroup) and bucket fills (bins per final-stage workgroup) by replaying
 * each candidate's graph, and keeps the fastest.  Buffers passed here must
 * stay valid for the plan's life.  A plan has its own stream and scratch;
 * destroy it before its context. */
typedef struct hg_batch_plan_s *hg_batch_plan;
typedef struct {
  uint32_t tile_keys;   /* keys per partition tile */
  float fill;           /* bucket fill target of the final stage */
  uint32_t candidates;  /* configurations timed */
  float tuned_ms;       /* replay time of the chosen one (-1 without tuning) */
  uint32_t learned_tables; /* tables on their learned model (with HG_PLAN_TRAIN: after the last replay) */
} hg_batch_plan_info;
/* Flags of hg_batch_plan_create's `autotune` argument. */
#define HG_PLAN_AUTOTUNE 1u
/* Train every table's learned model inside the plan, on every replay (the
 * plan then follows changing key contents): table t's model is fitted on
 * max(n/100, min(n, 1024)) sampled keys with seed t into
 * bd->d_rmi_models + t * hg_rmi_model_bytes(bd->rmi_leaves) (rmi_leaves in
 * [1, 1024]), and the table uses it when its sampled check wins, murmur3
 * otherwise (bd->h_hash_kind is ignored).  Samples must fit one workgroup
 * (up to 16384 uint32 or 8192 uint64 keys): HG_ERR_UNSUPPORTED otherwise. */
#define HG_PLAN_TRAIN 2u

hg_status hg_batch_plan_create(hg_ctx ctx, const hg_batch_desc *bd, const void *d_keys, uint32_t *d_offsets,
                               void *d_csr_keys, const void *d_probe, const uint64_t *h_probe_off, uint64_t *d_counts,
                               uint32_t autotune, hg_batch_plan *out);
hg_status hg_batch_plan_run(hg_batch_plan plan);
hg_status hg_batch_plan_query(hg_batch_plan plan, hg_batch_plan_info *info);
hg_status hg_batch_plan_destroy(hg_batch_plan plan);

/* ---- primitives ----------------------------------------------------------
 * Exclusive prefix sum (single-pass decoupled look-back), replaces
 * cub::DeviceScan::ExclusiveSum at src/SingleHashGraph.cu:147-152,166-172 and
 * src/MultiHashGraph.cu:327-336,353-365,706-715.  elem_bytes 4 or 8 (sums
 * modulo 2^32 / 2^64, every bit kept); d_out may alias d_in; d_out has n + 1
 * elements (the total goes last).                                         */
hg_status hg_exclusive_scan(hg_ctx ctx, const void *d_in, void *d_out, uint64_t n, uint32_t elem_bytes);
/* bins of n keys (murmur3, seed 0, low 32 bits, mod M) -> uint32. */
hg_status hg_hash_bins(hg_ctx ctx, const void *d_keys, uint64_t n, uint32_t key_bytes, uint64_t n_bins,
                       uint32_t *d_bins);
/* bins of n keys under a table's hash (murmur3 or its learned model). */
hg_status hg_table_bins(hg_ctx ctx, const hg_table_desc *desc, const void *d_keys, uint64_t n, uint32_t *d_bins);

#ifdef __cplusplus
}
#endif
#endif /* HASHGRAPH_H */
