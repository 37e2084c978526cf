// SingleHashGraph.cpp -- the reference's single-GPU class over the C ABI.
#include "SingleHashGraph.hpp"

#include <hip/hip_runtime.h>

#include "hg_kernels.hpp"

namespace {

void ok(hg_status s, const char *where) {
  if (s != HG_OK) throw hashgraph::error(s, std::string(where) + ": " + hg_last_error());
}

void hip_ok(hipError_t e, const char *where) {
  if (e != hipSuccess) throw hashgraph::error(HG_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

hg_table_desc desc_of(int64_t n, int64_t M) {
  hg_table_desc d{};
  d.key_bytes = 8;
  d.offset_bytes = 4;
  d.n_keys = (uint64_t)n;
  d.n_bins = (uint64_t)M;
  d.hash_kind = HG_HASH_MURMUR3;
  return d;
}

}  // namespace

void SingleHashGraph::alloc(int64_t countSize, int64_t tableSize) {
  if (countSize < 0 || tableSize <= 0) throw hashgraph::error(HG_ERR_INVALID_ARG, "SingleHashGraph: bad sizes");
  count_ = countSize;
  table_ = tableSize;
  int dev = 0;
  hip_ok(hipGetDevice(&dev), "hipGetDevice");
  ok(hg_ctx_create(dev, &ctx_), "hg_ctx_create");
  hip_ok(hipMalloc(&d_vals, (size_t)std::max<int64_t>(countSize, 1) * sizeof(hkey_t)), "hipMalloc keys");
  hip_ok(hipMalloc(&d_edges, (size_t)std::max<int64_t>(countSize, 1) * sizeof(hkey_t)), "hipMalloc edges");
  hip_ok(hipMalloc(&d_offset, (size_t)(tableSize + 1) * sizeof(index_t)), "hipMalloc offsets");
}

// src/SingleHashGraph.cu:79-124: allocates and fills the keys on the device.
SingleHashGraph::SingleHashGraph(int64_t countSize, int64_t maxkey, int64_t tableSize, int64_t lrbBins) {
  lrbBins_ = lrbBins;
  alloc(countSize, tableSize);
  if (maxkey < 0) throw hashgraph::error(HG_ERR_INVALID_ARG, "SingleHashGraph: maxkey < 0");
  // uniform in [0, maxkey] (prg(0, maxkey) draws from the closed range)
  const uint64_t range = (uint64_t)maxkey + 1;  // 0 when maxkey = 2^64-1: full range
  hip_ok(hg::launch_gen_uniform(d_vals, (uint64_t)countSize, 8, mix64(0), range, nullptr), "key generation");
}

SingleHashGraph::SingleHashGraph(const hkey_t *d_keys, int64_t countSize, int64_t tableSize, int64_t lrbBins) {
  lrbBins_ = lrbBins;
  alloc(countSize, tableSize);
  if (countSize)
    hip_ok(hipMemcpy(d_vals, d_keys, (size_t)countSize * sizeof(hkey_t), hipMemcpyDeviceToDevice), "copy keys");
}

// The reference's destructor is empty (src/SingleHashGraph.cu:126-127); this one frees.
SingleHashGraph::~SingleHashGraph() {
  if (ctx_) hg_ctx_destroy(ctx_);
  (void)hipFree(d_vals);
  (void)hipFree(d_edges);
  (void)hipFree(d_offset);
}

// src/SingleHashGraph.cu:216-232 (LRB build of countSize keys over tableSize bins)
void SingleHashGraph::build(int64_t countSize, int64_t tableSize) {
  if (countSize != count_ || tableSize != table_)
    throw hashgraph::error(HG_ERR_INVALID_ARG, "SingleHashGraph::build: sizes differ from the constructor's");
  const hg_table_desc d = desc_of(countSize, tableSize);
  ok(hg_build(ctx_, &d, d_vals, d_offset, d_edges), "hg_build");
}

// MultiHashGraph::intersect semantics (src/MultiHashGraph.cu:631-740) for one table.
uint64_t SingleHashGraph::intersect(const SingleHashGraph &A, const SingleHashGraph &B, hkey_t **d_out) {
  if (A.table_ != B.table_) throw hashgraph::error(HG_ERR_MISMATCH, "intersect: tables have different bin counts");
  const hg_table_desc dA = desc_of(A.count_, A.table_), dB = desc_of(B.count_, B.table_);
  uint64_t cnt = 0;
  ok(hg_intersect(A.ctx_, &dA, A.d_offset, A.d_edges, &dB, B.d_offset, B.d_edges, nullptr, 0, &cnt), "hg_intersect");
  if (d_out) {
    *d_out = nullptr;
    if (cnt) {
      hip_ok(hipMalloc(d_out, cnt * sizeof(hkey_t)), "hipMalloc output");
      ok(hg_intersect(A.ctx_, &dA, A.d_offset, A.d_edges, &dB, B.d_offset, B.d_edges, *d_out, cnt, &cnt),
         "hg_intersect");
    }
  }
  return cnt;
}
