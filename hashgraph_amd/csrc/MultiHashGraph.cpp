// MultiHashGraph.cpp -- the reference's multi-GPU class for one GPU's share,
// over the C ABI: split finding, local partition and per-partition CSRs run
// on this device; gpuCount is the number of logical partitions.  Compiled for
// the reference's four type modes (B32 x INDEX_TRACK).
#include "MultiHashGraph.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <iostream>
#include <vector>

namespace hashgraph {
namespace {

void ok(hg_status s, const char *where) {
  if (s != HG_OK) throw hashgraph::error(s, std::string(where) + ": " + hg_last_error());
}

void hip_ok(hipError_t e, const char *where) {
  if (e != hipSuccess) throw hashgraph::error(HG_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

template <typename K>
hg_table_desc desc_of(uint64_t n, uint64_t M) {
  hg_table_desc d{};
  d.key_bytes = sizeof(K);
  d.offset_bytes = 8;
  d.n_keys = n;
  d.n_bins = M;
  d.hash_kind = HG_HASH_MURMUR3;
  return d;
}

// keyval_ind entries from the CSR keys and their input positions
template <typename K, typename E>
__global__ void k_pack_entries(const K *__restrict__ keys, const uint64_t *__restrict__ pos, uint64_t n,
                               E *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    E e{};
    e.key = keys[i];
    e.ind = (int64_t)pos[i];
    out[i] = e;
  }
}

// keypair.right (int64) from 32-bit matched keys
__global__ void k_widen(const uint32_t *__restrict__ in, uint64_t n, int64_t *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (int64_t)in[i];
}

uint32_t grid_of(uint64_t n) { return (uint32_t)std::min<uint64_t>((n + 255) / 256, 16384); }

// Storage of a table: device memory, or with MANAGED_MEM managed memory
// prefetched to the device (src/MultiHashGraph.cu:229-247).  Where the device
// has no demand paging, managed memory is host memory, and the prefetch a
// hint that leaves it there; the staged calls stream it through HBM.
template <bool MM, typename T>
void table_alloc(T **p, size_t bytes, const char *what) {
  if constexpr (MM) {
    hip_ok(hipMallocManaged((void **)p, bytes), what);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipMemPrefetchAsync(*p, bytes, dev, nullptr) != hipSuccess)
      (void)hipGetLastError();
  } else {
    hip_ok(hipMalloc((void **)p, bytes), what);
  }
}

}  // namespace

// src/MultiHashGraph.cu:61-251: copies every input key to the host (h_vals)
// and allocates the per-partition structures.
template <typename K, bool IT, bool MM>
BasicMultiHashGraph<K, IT, MM>::BasicMultiHashGraph(inputData *h_dVals_, index_t countSize_, index_t maxkey_,
                                                index_t tableSize_, index_t binCount_, index_t lrbBins_,
                                                index_t gpuCount_) {
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
  const size_t nk = (size_t)std::max<index_t>(countSize, 1);
  h_vals = new hkey_t[nk]();
  table_alloc<MM>(&d_keys_, nk * sizeof(hkey_t), "allocate keys");
  index_t at = 0;
  for (index_t i = 0; i < gpuCount; i++) {
    const index_t len = h_dVals[i].len;
    if (len < 0 || at + len > countSize)
      throw hashgraph::error(HG_ERR_INVALID_ARG, "MultiHashGraph: input lengths exceed countSize");
    if (len) {
      hip_ok(hipMemcpy(d_keys_ + at, h_dVals[i].d_keys, (size_t)len * sizeof(hkey_t), hipMemcpyDefault),
             "copy keys");
      hip_ok(hipMemcpy(h_vals + at, h_dVals[i].d_keys, (size_t)len * sizeof(hkey_t), hipMemcpyDefault),
             "copy keys");
    }
    at += len;
  }
  if (at != countSize)
    throw hashgraph::error(HG_ERR_INVALID_ARG, "MultiHashGraph: input lengths do not sum to countSize");
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
  const size_t bytes = nk * sizeof(keyval) + (size_t)(tableSize + gpuCount) * sizeof(index_t);
  table_alloc<MM>(&blob_, bytes, "allocate CSR");
  if (IT) {
    table_alloc<MM>(&d_csrk_, nk * sizeof(hkey_t), "allocate CSR keys");
    table_alloc<MM>(&d_csri_, nk * sizeof(uint64_t), "allocate CSR positions");
  } else {
    d_csrk_ = (hkey_t *)blob_;  // the entries are the keys
  }
}

template <typename K, bool IT, bool MM>
BasicMultiHashGraph<K, IT, MM>::~BasicMultiHashGraph() {
  if (ctx_) hg_ctx_destroy(ctx_);
  if (IT) {
    (void)hipFree(d_csrk_);
    (void)hipFree(d_csri_);
  }
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
template <typename K, bool IT, bool MM>
void BasicMultiHashGraph<K, IT, MM>::destroyMulti() {}

// src/MultiHashGraph.cu:408-629 for every partition of this GPU: splits
// (countBinSizes) when findSplits, local partition, per-partition CSR; with
// INDEX_TRACK the entries carry the key's input position.
template <typename K, bool IT, bool MM>
void BasicMultiHashGraph<K, IT, MM>::build(bool findSplits, index_t tid) {
  if (tid < 0 || tid >= gpuCount) throw hashgraph::error(HG_ERR_INVALID_ARG, "build: tid out of range");
  std::lock_guard<std::mutex> lk(mu_);
  if (tid != 0 && gpuCount > 1) return;  // partition 0's call builds them all
  const hg_table_desc d = desc_of<K>((uint64_t)countSize, (uint64_t)tableSize);
  if (findSplits) {
    std::vector<int64_t> sp(gpuCount + 1);
    if constexpr (MM)
      ok(hg_find_splits_staged(ctx_, &d, d_keys_, (uint64_t)binCount, (uint32_t)gpuCount, sp.data(), stagedBudget),
         "hg_find_splits_staged");
    else
      ok(hg_find_splits(ctx_, &d, d_keys_, (uint64_t)binCount, (uint32_t)gpuCount, sp.data()), "hg_find_splits");
    for (index_t i = 0; i <= gpuCount; i++) h_binSplits[i] = sp[i];
  }
  hip_ok(hipMemcpy(h_dBinSplits[0], h_binSplits, (size_t)(gpuCount + 1) * sizeof(index_t), hipMemcpyHostToDevice),
         "copy splits");
  std::vector<int64_t> sp(h_binSplits, h_binSplits + gpuCount + 1);
  std::vector<uint64_t> po(gpuCount + 1);
  index_t *off = (index_t *)(blob_ + (size_t)std::max<index_t>(countSize, 1) * sizeof(keyval));
  if constexpr (MM && IT) {
    // records (input position << 32 | low 32 key bits) fall in their keys'
    // bins; the entries are then filled from the host copy of the keys
    const uint64_t n = (uint64_t)countSize;
    if (n >= (1ull << 32)) throw hashgraph::error(HG_ERR_UNSUPPORTED, "INDEX_TRACK needs fewer than 2^32 keys");
    std::vector<uint64_t> rec(std::max<uint64_t>(n, 1)), out(std::max<uint64_t>(n, 1));
    for (uint64_t i = 0; i < n; i++) rec[i] = (i << 32) | (uint32_t)h_vals[i];
    hg_table_desc d8 = d;
    d8.key_bytes = 8;
    ok(hg_build_partitioned_staged(ctx_, &d8, rec.data(), sp.data(), (uint32_t)gpuCount, off, out.data(), po.data(),
                                   stagedBudget),
       "hg_build_partitioned_staged");
    keyval_ind *ent = (keyval_ind *)blob_;
    for (uint64_t j = 0; j < po[gpuCount]; j++) {
      const uint64_t i = out[j] >> 32;
      d_csrk_[j] = h_vals[i];
      d_csri_[j] = i;
      ent[j].key = h_vals[i];
      ent[j].ind = (index_t)i;
    }
  } else if constexpr (MM) {
    ok(hg_build_partitioned_staged(ctx_, &d, d_keys_, sp.data(), (uint32_t)gpuCount, off, d_csrk_, po.data(),
                                   stagedBudget),
       "hg_build_partitioned_staged");
  } else if constexpr (IT) {
    ok(hg_build_partitioned_indexed(ctx_, &d, d_keys_, sp.data(), (uint32_t)gpuCount, off, d_csrk_, d_csri_,
                                    po.data()),
       "hg_build_partitioned_indexed");
    const uint64_t kept = po[gpuCount];
    if (kept)
      hipLaunchKernelGGL((k_pack_entries<hkey_t, keyval_ind>), dim3(grid_of(kept)), dim3(256), 0, nullptr,
                         (const hkey_t *)d_csrk_, (const uint64_t *)d_csri_, kept, (keyval_ind *)blob_);
    hip_ok(hipGetLastError(), "pack entries");
  } else {
    ok(hg_build_partitioned(ctx_, &d, d_keys_, sp.data(), (uint32_t)gpuCount, off, d_csrk_, po.data()),
       "hg_build_partitioned");
  }
  hip_ok(hipStreamSynchronize(nullptr), "build");
  for (index_t t = 0; t < gpuCount; t++) {
    h_partOff[t] = (index_t)po[t];
    h_dFinalKeys[t] = blob_;
    h_edgesOff[t] = (index_t)(po[t] * sizeof(keyval));
    h_offsetOff[t] =
        (index_t)((size_t)std::max<index_t>(countSize, 1) * sizeof(keyval) + (size_t)(sp[t] + t) * sizeof(index_t));
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

// src/MultiHashGraph.cu:631-740 for partition tid: the matched B keys, as
// keypair.right (int64).
template <typename K, bool IT, bool MM>
void BasicMultiHashGraph<K, IT, MM>::intersect(BasicMultiHashGraph &A, BasicMultiHashGraph &B, index_t *h_Common,
                                           keypair **h_dOutput, index_t tid) {
  if (tid < 0 || tid >= A.gpuCount || tid >= B.gpuCount) throw hashgraph::error(HG_ERR_INVALID_ARG, "intersect: tid");
  const index_t R = A.h_binSplits[tid + 1] - A.h_binSplits[tid];
  if (R != B.h_binSplits[tid + 1] - B.h_binSplits[tid])
    throw hashgraph::error(HG_ERR_MISMATCH, "intersect: table size not the same between the two graphs");
  h_Common[tid] = 0;
  h_dOutput[tid] = nullptr;
  if (R == 0) return;
  const hg_table_desc dA = desc_of<K>((uint64_t)A.partitionKeys(tid), (uint64_t)R);
  const hg_table_desc dB = desc_of<K>((uint64_t)B.partitionKeys(tid), (uint64_t)R);
  const void *oa = A.h_dFinalKeys[tid] + A.h_offsetOff[tid], *ka = A.d_csrk_ + A.h_partOff[tid];
  const void *ob = B.h_dFinalKeys[tid] + B.h_offsetOff[tid], *kb = B.d_csrk_ + B.h_partOff[tid];
  uint64_t cnt = 0;
  if constexpr (MM) {
    // src/MultiHashGraph.cu:664-670: managed memory for the join's arrays;
    // the tables stream through HBM slab by slab
    ok(hg_intersect_staged(A.ctx_, &dA, oa, ka, &dB, ob, kb, nullptr, 0, &cnt, A.stagedBudget),
       "hg_intersect_staged");
    h_Common[tid] = (index_t)cnt;
    if (!cnt) return;
    table_alloc<MM>(&h_dOutput[tid], cnt * sizeof(keypair), "allocate output");
    if (sizeof(K) == 8) {
      ok(hg_intersect_staged(A.ctx_, &dA, oa, ka, &dB, ob, kb, h_dOutput[tid], cnt, &cnt, A.stagedBudget),
         "hg_intersect_staged");
      return;
    }
    std::vector<uint32_t> t(cnt);
    ok(hg_intersect_staged(A.ctx_, &dA, oa, ka, &dB, ob, kb, t.data(), cnt, &cnt, A.stagedBudget),
       "hg_intersect_staged");
    for (uint64_t i = 0; i < cnt; i++) h_dOutput[tid][i].right = (index_t)t[i];
    return;
  }
  ok(hg_intersect(A.ctx_, &dA, oa, ka, &dB, ob, kb, nullptr, 0, &cnt), "hg_intersect");
  h_Common[tid] = (index_t)cnt;
  if (!cnt) return;
  hip_ok(hipMalloc(&h_dOutput[tid], cnt * sizeof(keypair)), "hipMalloc output");
  if (sizeof(K) == 8) {
    ok(hg_intersect(A.ctx_, &dA, oa, ka, &dB, ob, kb, h_dOutput[tid], cnt, &cnt), "hg_intersect");
    return;
  }
  // 32-bit keys: matched keys first, then widened into keypair.right
  uint32_t *tmp = nullptr;
  hip_ok(hipMalloc(&tmp, cnt * sizeof(uint32_t)), "hipMalloc");
  const hg_status st = hg_intersect(A.ctx_, &dA, oa, ka, &dB, ob, kb, tmp, cnt, &cnt);
  if (st == HG_OK)
    hipLaunchKernelGGL(k_widen, dim3(grid_of(cnt)), dim3(256), 0, nullptr, (const uint32_t *)tmp, cnt,
                       (int64_t *)h_dOutput[tid]);
  const hipError_t e = hipDeviceSynchronize();
  (void)hipFree(tmp);
  ok(st, "hg_intersect");
  hip_ok(e, "widen");
}

// src/MultiHashGraph.cu:742-862: plain build of all keys on this device,
// then, for every partition and local bin, the degree and the sorted key
// list must equal those of global bin j + splits[i]; errors go to stderr.
// With INDEX_TRACK every entry's position must also point at its key.
template <typename K, bool IT, bool MM>
void BasicMultiHashGraph<K, IT, MM>::buildSingle() {
  if (!built_) throw hashgraph::error(HG_ERR_INVALID_ARG, "buildSingle: build() first");
  hkey_t *d_k = nullptr, *d_csr = nullptr;
  index_t *d_off = nullptr;
  hip_ok(hipMalloc(&d_k, (size_t)std::max<index_t>(countSize, 1) * sizeof(hkey_t)), "hipMalloc");
  hip_ok(hipMalloc(&d_csr, (size_t)std::max<index_t>(countSize, 1) * sizeof(hkey_t)), "hipMalloc");
  hip_ok(hipMalloc(&d_off, (size_t)(tableSize + 1) * sizeof(index_t)), "hipMalloc");
  hip_ok(hipMemcpy(d_k, h_vals, (size_t)countSize * sizeof(hkey_t), hipMemcpyHostToDevice), "copy");
  const hg_table_desc d = desc_of<K>((uint64_t)countSize, (uint64_t)tableSize);
  ok(hg_build(ctx_, &d, d_k, d_off, d_csr), "hg_build");
  std::vector<index_t> off(tableSize + 1);
  std::vector<hkey_t> edges(countSize);
  hip_ok(hipMemcpy(off.data(), d_off, off.size() * sizeof(index_t), hipMemcpyDeviceToHost), "copy");
  hip_ok(hipMemcpy(edges.data(), d_csr, edges.size() * sizeof(hkey_t), hipMemcpyDeviceToHost), "copy");
  std::vector<keyval> mk(countSize);
  std::vector<index_t> moff(tableSize + gpuCount);
  hip_ok(hipMemcpy(mk.data(), blob_, mk.size() * sizeof(keyval), hipMemcpyDefault), "copy");
  hip_ok(hipMemcpy(moff.data(), blob_ + (size_t)std::max<index_t>(countSize, 1) * sizeof(keyval),
                   moff.size() * sizeof(index_t), hipMemcpyDefault),
         "copy");
  for (index_t i = 0; i < gpuCount; i++) {
    const index_t lo = h_binSplits[i], R = h_binSplits[i + 1] - lo;
    const index_t *po = moff.data() + lo + i;
    const keyval *pk = mk.data() + h_partOff[i];
    for (index_t j = 0; j < R; j++) {
      const index_t hash = j + lo;
      const index_t md = po[j + 1] - po[j], sd = off[hash + 1] - off[hash];
      if (md != sd) std::cerr << "Degree error hash: " << hash << " multi: " << md << " single: " << sd << "\n";
      std::vector<hkey_t> a, b(edges.begin() + off[hash], edges.begin() + off[hash + 1]);
      for (index_t q = po[j]; q < po[j + 1]; q++) {
        a.push_back(pk[q].key);
        if constexpr (IT) {
          if (pk[q].ind < 0 || pk[q].ind >= countSize || h_vals[pk[q].ind] != pk[q].key)
            std::cerr << "Index error hash: " << hash << "\n";
        }
      }
      std::sort(a.begin(), a.end());
      std::sort(b.begin(), b.end());
      if (a != b) std::cerr << "List error hash: " << hash << "\n";
    }
  }
  (void)hipFree(d_k);
  (void)hipFree(d_csr);
  (void)hipFree(d_off);
}

template class BasicMultiHashGraph<int64_t, false, false>;
template class BasicMultiHashGraph<uint32_t, false, false>;
template class BasicMultiHashGraph<int64_t, true, false>;
template class BasicMultiHashGraph<uint32_t, true, false>;
template class BasicMultiHashGraph<int64_t, false, true>;
template class BasicMultiHashGraph<uint32_t, false, true>;
template class BasicMultiHashGraph<int64_t, true, true>;
template class BasicMultiHashGraph<uint32_t, true, true>;

}  // namespace hashgraph
