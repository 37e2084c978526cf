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
// partition tid's CSR is read.
// Types follow the reference's compile-time switches
// (include/MultiHashGraph.cuh:62-111):
//   B32          hkey_t = HashKey = uint32_t (else int64_t); index_t = int64_t
//   INDEX_TRACK  CSR entries are keyval_ind {key, ind} (ind = the key's
//                position in the concatenated input), else keyval_key {key}
// The reference's header defines INDEX_TRACK itself; here both switches are
// opt-in compiler definitions (-DB32, -DINDEX_TRACK), so the default is the
// 64-bit, index-free form of its headline experiments.  Offsets are index_t.
// MANAGED_MEM (include/MultiHashGraph.cuh:63, src/MultiHashGraph.cu:229-247)
// puts the table's storage (the key copy, CSR entries and offsets, and the
// intersect's output) in hipMallocManaged memory prefetched to the device, and
// builds and joins through the staged calls of hashgraph.h, which stream it
// through HBM in slabs of at most stagedBudget bytes of device memory: the
// table may be larger than the GPU's memory.
// Ownership as in the reference: build() frees the caller's
// h_dVals[i].d_keys / d_hash; intersect() hipMallocs h_dOutput[tid] (caller
// frees).  Errors throw hashgraph::error (the reference prints and exits).
// build() may be called once per tid from an OpenMP region or once from a
// single thread; the first call builds every partition, later calls for the
// same round return.
#pragma once
#include <stdint.h>

#include <mutex>
#include <type_traits>

#include "SingleHashGraph.hpp"  // hashgraph::error
#include "hashgraph.h"

namespace hashgraph {

template <typename KeyT, bool IndexTrack, bool Managed = false>
class BasicMultiHashGraph {
 public:
  using hkey_t = KeyT;
  using index_t = int64_t;
  using HashKey = KeyT;
  struct keyval_key {
    hkey_t key;
  };
  struct keyval_ind {
    hkey_t key;
    index_t ind;
  };
  struct keypair {
    index_t right;
  };
  struct inputData {
    hkey_t *d_keys;
    HashKey *d_hash;
    index_t len;
  };
  using keyval = typename std::conditional<IndexTrack, keyval_ind, keyval_key>::type;

  BasicMultiHashGraph(inputData *h_dVals, index_t countSize, index_t maxkey, index_t tableSize, index_t binCount,
                      index_t lrbBins, index_t gpuCount);
  ~BasicMultiHashGraph();
  BasicMultiHashGraph(const BasicMultiHashGraph &) = delete;
  BasicMultiHashGraph &operator=(const BasicMultiHashGraph &) = delete;

  void build(bool findSplits, index_t tid);
  void buildSingle();  // self-check against a plain build; mismatches go to stderr
  // output: B's key of every matching pair, as in the reference (keypair.right)
  static void intersect(BasicMultiHashGraph &mgA, BasicMultiHashGraph &mgB, index_t *h_Common, keypair **h_dOutput,
                        index_t tid);
  void destroyMulti();

  // partition tid's CSR: offsets (index_t, R+1 entries, local bins) at
  // h_dFinalKeys[tid] + h_offsetOff[tid], keyval entries at
  // h_dFinalKeys[tid] + h_edgesOff[tid]
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

  // MANAGED_MEM: device memory each staged call may take (0 = half of the
  // free memory)
  size_t stagedBudget = 0;

  // number of keys of partition tid after the last build
  index_t partitionKeys(index_t tid) const { return h_partOff[tid + 1] - h_partOff[tid]; }

 private:
  inputData *h_dVals = nullptr;
  index_t maxkey = 0, binCount = 0, lrbBins = 0;
  hg_ctx ctx_ = nullptr;
  char *blob_ = nullptr;  // [ CSR entries (countSize) | offsets (tableSize + gpuCount) ]
  hkey_t *d_keys_ = nullptr;
  hkey_t *d_csrk_ = nullptr;     // CSR keys alone (the joins read them; == the entries without INDEX_TRACK)
  uint64_t *d_csri_ = nullptr;   // CSR input positions (INDEX_TRACK)
  index_t *h_ownSplits = nullptr;
  index_t *h_partOff = nullptr;
  std::mutex mu_;
  bool built_ = false;
};

}  // namespace hashgraph

#ifdef B32
#define HASHGRAPH_KEY_T uint32_t
#else
#define HASHGRAPH_KEY_T int64_t
#endif
#ifdef INDEX_TRACK
#define HASHGRAPH_INDEX_TRACK true
#else
#define HASHGRAPH_INDEX_TRACK false
#endif
#ifdef MANAGED_MEM
#define HASHGRAPH_MANAGED true
#else
#define HASHGRAPH_MANAGED false
#endif
using MultiHashGraph = hashgraph::BasicMultiHashGraph<HASHGRAPH_KEY_T, HASHGRAPH_INDEX_TRACK, HASHGRAPH_MANAGED>;
