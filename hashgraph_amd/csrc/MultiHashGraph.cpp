// MultiHashGraph.cpp -- the reference's multi-GPU class for one GPU's share,
// over the C ABI: split finding, local partition and per-partition CSRs run
// on this device; gpuCount is the number of logical partitions.
#include "MultiHashGraph.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <iostream>
#include <vector>

namespace {

void ok(hg_status s, const char *where) {
  if (s != HG_OK) throw hashgraph::error(s, std::string(where) + ": " + hg_last_error());
}

void hip_ok(hipError_t e, const char *where) {
  if (e != hipSuccess) throw hashgraph::error(HG_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

hg_table_desc desc_of(uint64_t n, uint64_t M) {
  hg_table_desc d{};
  d.key_bytes = 8;
  d.offset_bytes = 8;
  d.n_keys = n;
  d.n_bins = M;
  d.hash_kind = HG_HASH_MURMUR3;
  return d;
}

}  // namespace

// src/MultiHashGraph.cu:61-251: copies every input key to the host (h_vals)
// and allocates the per-partition structures.
MultiHashGraph::MultiHashGraph(inputData *h_dVals_, index_t countSize_, index_t maxkey_, index_t tableSize_,
                               index_t binCount_, index_t lrbBins_, index_t gpuCount_) {
  if (countSize_ < 0 || tableSize_ <= 0 || binCount_ <= 0 || gpuCount_ <= 0 || !h_dVals_)
    throw hashgraph::error(HG_ERR_INVALID_ARG, "MultiHashGraph: bad arguments");
  h_dVals = h_dVals_;
  countSize = countSize_;
  maxkey = maxkey_;
  tableSize = tableSize_;
  binCount = binCount_;
  lrbBins = lrbBins_;
  gpuCount = gpuCount_;
  int dev = 0;
  hip_ok(hipGetDevice(&dev), "hipGetDevice");
  ok(hg_ctx_create(dev, &ctx_), "hg_ctx_create");
  h_vals = new hkey_t[std::max<index_t>(countSize, 1)]();
  hip_ok(hipMalloc(&d_keys_, (size_t)std::max<index_t>(countSize, 1) * sizeof(hkey_t)), "hipMalloc keys");
  index_t at = 0;
  for (index_t i = 0; i < gpuCount; i++) {
    const index_t len = h_dVals[i].len;
    if (len < 0 || at + len > countSize) throw hashgraph::error(HG_ERR_INVALID_ARG, "MultiHashGraph: input lengths exceed countSize");
    if (len) {
      hip_ok(hipMemcpy(d_keys_ + at, h_dVals[i].d_keys, (size_t)len * sizeof(hkey_t), hipMemcpyDeviceToDevice), "copy keys");
      hip_ok(hipMemcpy(h_vals + at, h_dVals[i].d_keys, (size_t)len * sizeof(hkey_t), hipMemcpyDeviceToHost), "copy keys");
    }
    at += len;
  }
  if (at != countSize) throw hashgraph::error(HG_ERR_INVALID_ARG, "MultiHashGraph: input lengths do not sum to countSize");
  h_ownSplits = new index_t[gpuCount + 1]();
  h_binSplits = h_ownSplits;
  h_dBinSplits = new index_t *[gpuCount]();
  index_t *d_spl = nullptr;
  hip_ok(hipMalloc(&d_spl, (size_t)(gpuCount + 1) * sizeof(index_t)), "hipMalloc splits");
  for (index_t i = 0; i < gpuCount; i++) h_dBinSplits[i] = d_spl;
  h_partOff = new index_t[gpuCount + 1]();
  h_dFinalKeys = new char *[gpuCount]();
  h_hashOff = new index_t[gpuCount]();
  h_counterOff = new index_t[gpuCount]();
  h_offsetOff = new index_t[gpuCount]();
  h_edgesOff = new index_t[gpuCount]();
  h_lrbOff = new index_t[gpuCount]();
  const size_t bytes = (size_t)countSize * sizeof(hkey_t) + (size_t)(tableSize + gpuCount) * sizeof(index_t);
  hip_ok(hipMalloc(&blob_, bytes), "hipMalloc CSR");
}

MultiHashGraph::~MultiHashGraph() {
  if (ctx_) hg_ctx_destroy(ctx_);
  (void)hipFree(blob_);
  (void)hipFree(d_keys_);
  if (h_dBinSplits) (void)hipFree(h_dBinSplits[0]);
  delete[] h_dBinSplits;
  delete[] h_ownSplits;
  delete[] h_partOff;
  delete[] h_dFinalKeys;
  delete[] h_hashOff;
  delete[] h_counterOff;
  delete[] h_offsetOff;
  delete[] h_edgesOff;
  delete[] h_lrbOff;
  delete[] h_vals;
}

// The reference's destroyMulti is compiled out (src/MultiHashGraph.cu:262-296).
void MultiHashGraph::destroyMulti() {}

// src/MultiHashGraph.cu:408-629 for every partition of this GPU: splits
// (countBinSizes) when findSplits, local partition, per-partition CSR.
void MultiHashGraph::build(bool findSplits, index_t tid) {
  if (tid < 0 || tid >= gpuCount) throw hashgraph::error(HG_ERR_INVALID_ARG, "build: tid out of range");
  std::lock_guard<std::mutex> lk(mu_);
  if (tid != 0 && gpuCount > 1) return;  // partition 0's call builds them all
  const hg_table_desc d = desc_of((uint64_t)countSize, (uint64_t)tableSize);
  if (findSplits) {
    std::vector<int64_t> sp(gpuCount + 1);
    ok(hg_find_splits(ctx_, &d, d_keys_, (uint64_t)binCount, (uint32_t)gpuCount, sp.data()), "hg_find_splits");
    for (index_t i = 0; i <= gpuCount; i++) h_binSplits[i] = sp[i];
  }
  hip_ok(hipMemcpy(h_dBinSplits[0], h_binSplits, (size_t)(gpuCount + 1) * sizeof(index_t), hipMemcpyHostToDevice),
         "copy splits");
  std::vector<int64_t> sp(h_binSplits, h_binSplits + gpuCount + 1);
  std::vector<uint64_t> po(gpuCount + 1);
  hkey_t *csr = (hkey_t *)blob_;
  index_t *off = (index_t *)(blob_ + (size_t)countSize * sizeof(hkey_t));
  ok(hg_build_partitioned(ctx_, &d, d_keys_, sp.data(), (uint32_t)gpuCount, off, csr, po.data()),
     "hg_build_partitioned");
  hip_ok(hipStreamSynchronize(nullptr), "build");
  for (index_t t = 0; t < gpuCount; t++) {
    h_partOff[t] = (index_t)po[t];
    h_dFinalKeys[t] = blob_;
    h_edgesOff[t] = (index_t)(po[t] * sizeof(hkey_t));
    h_offsetOff[t] = (index_t)((size_t)countSize * sizeof(hkey_t) + (size_t)(sp[t] + t) * sizeof(index_t));
    h_hashOff[t] = h_counterOff[t] = h_lrbOff[t] = -1;  // not materialised
  }
  h_partOff[gpuCount] = (index_t)po[gpuCount];
  // the reference frees the caller's inputs here (src/MultiHashGraph.cu:498-503)
  for (index_t i = 0; i < gpuCount; i++) {
    if (h_dVals[i].d_keys) (void)hipFree(h_dVals[i].d_keys);
    if (h_dVals[i].d_hash) (void)hipFree(h_dVals[i].d_hash);
    h_dVals[i].d_keys = nullptr;
    h_dVals[i].d_hash = nullptr;
  }
  built_ = true;
}

// src/MultiHashGraph.cu:631-740 for partition tid.
void MultiHashGraph::intersect(MultiHashGraph &A, MultiHashGraph &B, index_t *h_Common, keypair **h_dOutput,
                               index_t tid) {
  if (tid < 0 || tid >= A.gpuCount || tid >= B.gpuCount) throw hashgraph::error(HG_ERR_INVALID_ARG, "intersect: tid");
  const index_t R = A.h_binSplits[tid + 1] - A.h_binSplits[tid];
  if (R != B.h_binSplits[tid + 1] - B.h_binSplits[tid])
    throw hashgraph::error(HG_ERR_MISMATCH, "intersect: table size not the same between the two graphs");
  h_Common[tid] = 0;
  h_dOutput[tid] = nullptr;
  if (R == 0) return;
  const hg_table_desc dA = desc_of((uint64_t)A.partitionKeys(tid), (uint64_t)R);
  const hg_table_desc dB = desc_of((uint64_t)B.partitionKeys(tid), (uint64_t)R);
  const void *oa = A.h_dFinalKeys[tid] + A.h_offsetOff[tid], *ka = A.h_dFinalKeys[tid] + A.h_edgesOff[tid];
  const void *ob = B.h_dFinalKeys[tid] + B.h_offsetOff[tid], *kb = B.h_dFinalKeys[tid] + B.h_edgesOff[tid];
  uint64_t cnt = 0;
  ok(hg_intersect(A.ctx_, &dA, oa, ka, &dB, ob, kb, nullptr, 0, &cnt), "hg_intersect");
  h_Common[tid] = (index_t)cnt;
  if (cnt) {
    hip_ok(hipMalloc(&h_dOutput[tid], cnt * sizeof(keypair)), "hipMalloc output");
    ok(hg_intersect(A.ctx_, &dA, oa, ka, &dB, ob, kb, h_dOutput[tid], cnt, &cnt), "hg_intersect");
  }
}

// src/MultiHashGraph.cu:742-862: plain build of all keys on this device,
// then, for every partition and local bin, the degree and the sorted key
// list must equal those of global bin j + splits[i]; errors go to stderr.
void MultiHashGraph::buildSingle() {
  if (!built_) throw hashgraph::error(HG_ERR_INVALID_ARG, "buildSingle: build() first");
  hkey_t *d_k = nullptr, *d_csr = nullptr;
  index_t *d_off = nullptr;
  hip_ok(hipMalloc(&d_k, (size_t)std::max<index_t>(countSize, 1) * sizeof(hkey_t)), "hipMalloc");
  hip_ok(hipMalloc(&d_csr, (size_t)std::max<index_t>(countSize, 1) * sizeof(hkey_t)), "hipMalloc");
  hip_ok(hipMalloc(&d_off, (size_t)(tableSize + 1) * sizeof(index_t)), "hipMalloc");
  hip_ok(hipMemcpy(d_k, h_vals, (size_t)countSize * sizeof(hkey_t), hipMemcpyHostToDevice), "copy");
  const hg_table_desc d = desc_of((uint64_t)countSize, (uint64_t)tableSize);
  ok(hg_build(ctx_, &d, d_k, d_off, d_csr), "hg_build");
  std::vector<index_t> off(tableSize + 1);
  std::vector<hkey_t> edges(countSize);
  hip_ok(hipMemcpy(off.data(), d_off, off.size() * sizeof(index_t), hipMemcpyDeviceToHost), "copy");
  hip_ok(hipMemcpy(edges.data(), d_csr, edges.size() * sizeof(hkey_t), hipMemcpyDeviceToHost), "copy");
  std::vector<hkey_t> mk(countSize);
  std::vector<index_t> moff(tableSize + gpuCount);
  hip_ok(hipMemcpy(mk.data(), blob_, mk.size() * sizeof(hkey_t), hipMemcpyDeviceToHost), "copy");
  hip_ok(hipMemcpy(moff.data(), blob_ + (size_t)countSize * sizeof(hkey_t), moff.size() * sizeof(index_t),
                   hipMemcpyDeviceToHost), "copy");
  for (index_t i = 0; i < gpuCount; i++) {
    const index_t lo = h_binSplits[i], R = h_binSplits[i + 1] - lo;
    const index_t *po = moff.data() + lo + i;
    const hkey_t *pk = mk.data() + h_partOff[i];
    for (index_t j = 0; j < R; j++) {
      const index_t hash = j + lo;
      const index_t md = po[j + 1] - po[j], sd = off[hash + 1] - off[hash];
      if (md != sd) std::cerr << "Degree error hash: " << hash << " multi: " << md << " single: " << sd << "\n";
      std::vector<hkey_t> a(pk + po[j], pk + po[j + 1]), b(edges.begin() + off[hash], edges.begin() + off[hash + 1]);
      std::sort(a.begin(), a.end());
      std::sort(b.begin(), b.end());
      if (a != b) std::cerr << "List error hash: " << hash << "\n";
    }
  }
  (void)hipFree(d_k);
  (void)hipFree(d_csr);
  (void)hipFree(d_off);
}
