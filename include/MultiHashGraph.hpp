// MultiHashGraph.hpp -- drop-in for the reference's MultiHashGraph
// (include/MultiHashGraph.cuh:150-293, src/MultiHashGraph.cu:61-862) for one
// MI355X: gpuCount is the number of logical partitions of the bin range held
// by this GPU ("one GPU's share"); there is no exchange between devices.
//
// Kept: the constructor, build(findSplits, tid), the static intersect(A, B,
// h_Common, h_dOutput, tid), buildSingle(), destroyMulti(), the public split
// fields (h_binSplits, h_dBinSplits: a second graph may take the first one's
// splits as in test/MultiHashGraphTest.cu:298-299), h_vals, and the byte
// offsets into h_dFinalKeys[tid] (h_offsetOff, h_edgesOff) through which
// partition tid's CSR is read.  Types: hkey_t = int64_t, index_t = int64_t,
// keys stored without the index payload (keyval_key, 8 bytes: the
// reference's INDEX_TRACK mode is not built), offsets index_t.
// Ownership as in the reference: build() frees the caller's
// h_dVals[i].d_keys / d_hash; intersect() hipMallocs h_dOutput[tid] (caller
// frees).  Errors throw hashgraph::error (the reference prints and exits).
// build() may be called once per tid from an OpenMP region or once from a
// single thread; the first call builds every partition, later calls for the
// same round return.
#pragma once
#include <stdint.h>

#include <mutex>

#include "SingleHashGraph.hpp"  // hashgraph::error
#include "hashgraph.h"

class MultiHashGraph {
 public:
  using hkey_t = int64_t;
  using index_t = int64_t;
  using HashKey = int64_t;
  struct keyval_key {
    hkey_t key;
  };
  struct keypair {
    index_t right;
  };
  struct inputData {
    hkey_t *d_keys;
    HashKey *d_hash;
    index_t len;
  };
  using keyval = keyval_key;

  MultiHashGraph(inputData *h_dVals, index_t countSize, index_t maxkey, index_t tableSize, index_t binCount,
                 index_t lrbBins, index_t gpuCount);
  ~MultiHashGraph();
  MultiHashGraph(const MultiHashGraph &) = delete;
  MultiHashGraph &operator=(const MultiHashGraph &) = delete;

  void build(bool findSplits, index_t tid);
  void buildSingle();  // self-check against a plain build; mismatches go to stderr
  static void intersect(MultiHashGraph &mgA, MultiHashGraph &mgB, index_t *h_Common, keypair **h_dOutput,
                        index_t tid);
  void destroyMulti();

  // partition tid's CSR: offsets (index_t, R+1 entries, local bins) at
  // h_dFinalKeys[tid] + h_offsetOff[tid], keys at h_dFinalKeys[tid] + h_edgesOff[tid]
  char **h_dFinalKeys = nullptr;
  index_t *h_hashOff = nullptr;
  index_t *h_counterOff = nullptr;
  index_t *h_offsetOff = nullptr;
  index_t *h_edgesOff = nullptr;
  index_t *h_lrbOff = nullptr;

  index_t *h_binSplits = nullptr;   // gpuCount + 1 entries
  index_t **h_dBinSplits = nullptr;

  index_t countSize = 0;
  index_t tableSize = 0;
  index_t gpuCount = 0;

  hkey_t *h_vals = nullptr;  // host copy of all keys (for correctness checks)

  // number of keys of partition tid after the last build
  index_t partitionKeys(index_t tid) const { return h_partOff[tid + 1] - h_partOff[tid]; }

 private:
  inputData *h_dVals = nullptr;
  index_t maxkey = 0, binCount = 0, lrbBins = 0;
  hg_ctx ctx_ = nullptr;
  char *blob_ = nullptr;  // [ CSR keys (countSize) | offsets (tableSize + gpuCount) ]
  hkey_t *d_keys_ = nullptr;
  index_t *h_ownSplits = nullptr;
  index_t *h_partOff = nullptr;
  std::mutex mu_;
  bool built_ = false;
  int round_calls_ = 0;
};
