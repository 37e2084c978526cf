// hg_common.hpp -- types shared by the host engine and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace hg {

// Most partitions a split-point segment may have (its splits are staged in
// LDS by the partition kernels).
constexpr uint32_t MAX_SPLITS = 4096;

// Division of a 32-bit value by an invariant d (1 <= d < 2^32) with one
// 32-bit multiply-high (Granlund & Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994), l = floor(log2 d):
//   mode 0, d = 2^s:  q = a >> s
//   mode 1:           q = mulhi(a, m) >> s,  m = ceil(2^(32+l) / d), s = l,
//                     used when m*d - 2^(32+l) <= 2^l (then exact for all a)
//   mode 2:           t = mulhi(a, m); q = (((a - t) >> 1) + t) >> s,
//                     m = ceil(2^(33+l) / d) - 2^32, s = l (always exact)
struct Div32 {
  uint32_t m;
  uint32_t s;
  uint32_t d;
  uint32_t mode;
};

// Bin map of one table: bin = murmur3_x86_32(low32(key), seed 0) mod M
// (include/SingleHashGraphOperators.cuh:71-79), or the learned CDF model.
struct BinMap {
  Div32 mod;          // division by M (the remainder is h - q*M)
  uint32_t M;         // number of bins
  uint32_t kind;      // 0 murmur3, 1 learned (RMI)
  const void *rmi;    // device pointer to the RMI model (kind == 1)
};

// One segment of a partition pass: a contiguous key range that is split
// into buckets by one bucket function.
//   kind 0 (range): bucket = (bin - bin_lo) / width, for bin in
//                   [bin_lo, bin_lo + nbuckets*width)
//   kind 1 (split): bucket = first p with bin < splits[p+1]; bucket
//                   nbuckets collects keys past the last split (dropped)
struct SegDesc {
  uint64_t key_begin;   // first key of the segment (input and output)
  uint64_t key_end;
  uint64_t bin_lo;
  BinMap map;
  Div32 div;
  uint32_t kind;
  uint32_t nbuckets;    // F_s (plus one drop bucket for kind 1)
  uint64_t bucket_base; // global id of bucket 0 of this segment
  uint64_t row_base;    // first count-matrix element of this segment
  uint32_t tile_begin;
  uint32_t ntiles;
  uint32_t ncols;       // row length of the count matrix (F_s, +1 for kind 1)
  uint32_t nsplits;     // kind 1: number of partitions P
  const int64_t *splits;// kind 1: device pointer, P+1 entries
  // used when this segment feeds the final stage / the probe:
  uint64_t seg_nbins;   // bins of the segment's range (kind 0)
  uint64_t fin_off_base;// offsets index of local bin 0 (bin_lo)
  uint64_t fin_key_rel; // key index subtracted from offset values
  uint64_t fin_end_pos; // offsets index of the end entry, ~0 = none
  uint32_t table;       // BinMap index
  uint32_t pad2;
};

struct TileDesc {
  uint64_t key_start;
  uint32_t count;
  uint32_t seg;
};

// Work item of the column scan: rows [row_lo, row_hi) of one segment,
// columns [col_lo, col_lo + 256).
struct ColItem {
  uint32_t seg;
  uint32_t chunk;       // chunk index inside the segment
  uint32_t row_lo, row_hi;  // tile indices relative to the segment
  uint32_t col_lo;
  uint32_t pad;
};

// One bucket of the final stage: keys part[key_start .. +count) all have a
// bin in [bin_lo, bin_lo + nbins) of table `map`; they are written as CSR
// keys at csr[key_start ..] (same range) and the offsets of those bins at
// off[off_pos .. off_pos + nbins) with value key_start - key_rel + prefix.
// When end_pos != ~0 the item also writes off[end_pos] = key_start + count
// - key_rel (the table's final offset).
struct FinItem {
  uint64_t key_start;
  uint64_t key_rel;
  uint64_t off_pos;
  uint64_t end_pos;
  uint32_t count;
  uint32_t nbins;
  uint32_t bin_lo;
  uint32_t table;   // index into the BinMap array
};

// Global-memory value table of one value-rich heavy bin of the run-length
// join: slots [base, base + mask + 1) of the batch's table arrays.
struct OvfTable {
  uint64_t base;
  uint64_t mask;
};
// A chunk of one side's keys of such a bin: keys[begin .. end), table tab.
struct OvfChunk {
  uint64_t begin, end;
  uint32_t tab;
  uint32_t pad;
};

// Learned CDF model (no counterpart in the reference).  Two FP32 linear
// stages over integer anchors:
//   feature = f(key - key_min)                     (root_kind 0), or
//             the bits of f(key - key_min) + 1 read as an unsigned integer
//             and converted to f32 (root_kind 1: a piecewise-linear log)
//   leaf    = clamp(floor(root_slope * feature + root_icpt), 0, L-1)
//   line    = anchor[leaf] +- floor(slope[leaf] * f(|key - x0[leaf]|))
//                                  (+ for key >= x0, - below it)
//   bin     = line when 0 <= line <= M-1, else murmur3(key) mod M (keys
//             beyond the trained range are hashed, not piled into an edge bin)
// f() is the round-to-nearest u64 -> f32 conversion, every product and sum is
// rounded separately (no fused multiply-add), floor saturates at 2^32-256.
// x0 (the leaf's first training key) keeps FP32 precision where it is
// needed: the float only holds a distance from the smallest key or from x0.
// The header is followed by L leaves.
struct RmiHeader {
  uint64_t key_min;
  float root_slope;
  float root_icpt;
  uint32_t n_leaves;
  uint32_t n_bins;
  uint32_t root_kind;
  uint32_t pad;
};
struct RmiLeaf {
  uint64_t x0;      // first training key routed to the leaf
  float slope;      // bins per key unit
  uint32_t anchor;  // bin of x0
};
static_assert(sizeof(RmiHeader) == 32 && sizeof(RmiLeaf) == 16, "RMI layout");

}  // namespace hg
