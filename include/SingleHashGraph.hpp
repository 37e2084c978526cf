// SingleHashGraph.hpp -- drop-in for the reference's SingleHashGraph
// (include/SingleHashGraph.cuh:89-116, src/SingleHashGraph.cu:79-232),
// implemented over the C ABI of hashgraph.h on one MI355X.
//
// Kept from the reference: the two signatures
//   SingleHashGraph(countSize, maxkey, tableSize, lrbBins)   (keys generated
//     on the device, uniform in [0, maxkey])
//   void build(countSize, tableSize)
// and the types hkey_t = int64_t (keys), index_t = int32_t (offsets),
// HashKey = uint32_t.  Added (the reference has none of them): a
// constructor taking caller keys, CSR accessors and intersect() with
// MultiHashGraph::intersect semantics.  Deviations: errors throw
// hashgraph::error instead of exit(0); the destructor frees what the object
// allocated; lrbBins is accepted and ignored (the bucket geometry is chosen
// by the library, the CSR does not depend on it).
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "hashgraph.h"

namespace hashgraph {
struct error : std::runtime_error {
  hg_status status;
  error(hg_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};
}  // namespace hashgraph

class SingleHashGraph {
 public:
  using hkey_t = int64_t;
  using index_t = int32_t;
  using HashKey = uint32_t;

  // Reference constructor: countSize keys uniform in [0, maxkey] (seeded
  // SplitMix64 instead of Thrust's minstd engine), tableSize bins.
  SingleHashGraph(int64_t countSize, int64_t maxkey, int64_t tableSize, int64_t lrbBins);
  // Keys supplied by the caller (device pointer, copied).
  SingleHashGraph(const hkey_t *d_keys, int64_t countSize, int64_t tableSize, int64_t lrbBins = 16000);
  ~SingleHashGraph();
  SingleHashGraph(const SingleHashGraph &) = delete;
  SingleHashGraph &operator=(const SingleHashGraph &) = delete;

  void build(int64_t countSize, int64_t tableSize);

  // CSR of the last build (device pointers owned by the object).
  const index_t *offsets() const { return d_offset; }  // tableSize + 1
  const hkey_t *keys() const { return d_edges; }       // countSize
  const hkey_t *input() const { return d_vals; }
  int64_t countSize() const { return count_; }
  int64_t tableSize() const { return table_; }

  // Bag-product join of two built graphs over the same bins: every pair of
  // equal keys counts once and emits the key.  *d_out is allocated with
  // hipMalloc (caller frees) unless the count is 0.  Returns the count.
  static uint64_t intersect(const SingleHashGraph &A, const SingleHashGraph &B, hkey_t **d_out);

 private:
  hkey_t *d_vals = nullptr;
  index_t *d_offset = nullptr;
  hkey_t *d_edges = nullptr;
  int64_t count_ = 0, table_ = 0, lrbBins_ = 16000;
  hg_ctx ctx_ = nullptr;
  void alloc(int64_t countSize, int64_t tableSize);
};
