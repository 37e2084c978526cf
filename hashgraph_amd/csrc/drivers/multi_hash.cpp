// multi-hash: the reference's multi-GPU driver (test/MultiHashGraphTest.cu)
// for one GPU's share (built three times: multi-hash with 64-bit keys,
// multi-hash-32 with -DB32, multi-hash-idx with -DINDEX_TRACK):
//   multi-hash countA maxkey binCount gpuCount lrbBins check|x countB build|x
// gpuCount = logical partitions on this GPU.  Inputs are the reference's
// sequential keys (A = 0..countA-1, B = 0..countB-1, test/MultiHashGraphTest.cu:81-128).
// Build mode times MultiHashGraph::build; intersect mode times build A +
// build B (sharing A's splits) + intersect, as the reference does
// (test/MultiHashGraphTest.cu:288-338).  "check" runs buildSingle() and, for
// intersect, the reference's CPU answer (sort both, count equal pairs).
// Built with -DMANAGED_MEM (multi-hash-managed, -32-managed, -idx-managed) the
// tables live in managed memory and stream through HBM; HG_STAGED_BUDGET
// (bytes) then caps the device memory of each staged call.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <iostream>
#include <vector>

#include "MultiHashGraph.hpp"

using index_t = MultiHashGraph::index_t;
using hkey_t = MultiHashGraph::hkey_t;

static void generateInput(MultiHashGraph::inputData *h_dVals, index_t countSize, uint32_t gpuCount) {
  const index_t avg = (index_t)((countSize + gpuCount - 1) / gpuCount);
  for (uint32_t i = 0; i < gpuCount; i++) {
    const index_t lo = std::min<index_t>(avg * i, countSize), hi = std::min<index_t>(avg * (i + 1), countSize);
    const index_t n = hi - lo;
    std::vector<hkey_t> h(n);
    for (index_t j = 0; j < n; j++) h[j] = lo + j;
    hipMalloc(&h_dVals[i].d_keys, std::max<index_t>(n, 1) * sizeof(hkey_t));
    hipMalloc(&h_dVals[i].d_hash, std::max<index_t>(n, 1) * sizeof(MultiHashGraph::HashKey));
    if (n) hipMemcpy(h_dVals[i].d_keys, h.data(), n * sizeof(hkey_t), hipMemcpyHostToDevice);
    h_dVals[i].len = n;
  }
}

static void set_budget(MultiHashGraph &g) {
#ifdef MANAGED_MEM
  if (const char *e = getenv("HG_STAGED_BUDGET")) g.stagedBudget = (size_t)strtoull(e, nullptr, 0);
#else
  (void)g;
#endif
}

int main(int argc, char **argv) {
  index_t countSizeA = 1L << 24, maxkey = 1L << 26, countSizeB = 1L << 22, lrbBins = -1;
  uint32_t binCount = 16000, gpuCount = 4;
  bool check = false, buildTest = false;
  if (argc >= 2 && argc < 9) {
    std::cerr << "Please specify all arguments.\n";
    return 1;
  }
  if (argc >= 9) {
    countSizeA = strtoll(argv[1], nullptr, 0);
    maxkey = strtoll(argv[2], nullptr, 0);
    binCount = (uint32_t)atoi(argv[3]);
    gpuCount = (uint32_t)atoi(argv[4]);
    lrbBins = strtoll(argv[5], nullptr, 0);
    check = !strcmp(argv[6], "check");
    countSizeB = strtoll(argv[7], nullptr, 0);
    buildTest = !strcmp(argv[8], "build");
  }
  const index_t tableSize = maxkey;
  std::cout << "countSizeA: " << countSizeA << "\nmaxkey: " << maxkey << std::endl;
  hipEvent_t start, stop;
  hipEventCreate(&start);
  hipEventCreate(&stop);
  float ms = 0;
  try {
    if (buildTest) {
      std::vector<MultiHashGraph::inputData> in(gpuCount);
      generateInput(in.data(), countSizeA, gpuCount);
      MultiHashGraph mhg(in.data(), countSizeA, maxkey, tableSize, binCount, lrbBins, gpuCount);
      set_budget(mhg);
      hipEventRecord(start);
      for (uint32_t t = 0; t < gpuCount; t++) mhg.build(true, t);
      hipEventRecord(stop);
      hipEventSynchronize(stop);
      hipEventElapsedTime(&ms, start, stop);
      std::cout << "multi buildTable() time: " << (ms / 1000.0) << "\n";
      if (check) mhg.buildSingle();
    } else {
      std::vector<MultiHashGraph::inputData> inA(gpuCount), inB(gpuCount);
      generateInput(inA.data(), countSizeA, gpuCount);
      generateInput(inB.data(), countSizeB, gpuCount);
      MultiHashGraph A(inA.data(), countSizeA, maxkey, tableSize, binCount, lrbBins, gpuCount);
      MultiHashGraph B(inB.data(), countSizeB, maxkey, tableSize, binCount, lrbBins, gpuCount);
      set_budget(A);
      set_budget(B);
      std::vector<MultiHashGraph::keypair *> out(gpuCount, nullptr);
      std::vector<index_t> common(gpuCount, 0);
      hipEventRecord(start);
      for (uint32_t t = 0; t < gpuCount; t++) A.build(true, t);
      for (uint32_t i = 0; i <= gpuCount; i++) B.h_binSplits[i] = A.h_binSplits[i];  // share A's splits
      for (uint32_t t = 0; t < gpuCount; t++) B.build(false, t);
      for (uint32_t t = 0; t < gpuCount; t++) MultiHashGraph::intersect(A, B, common.data(), out.data(), t);
      hipEventRecord(stop);
      hipEventSynchronize(stop);
      hipEventElapsedTime(&ms, start, stop);
      std::cout << "multi intersect() time: " << (ms / 1000.0) << "\n";
      index_t total = 0;
      for (index_t c : common) total += c;
      std::cout << "intersect count: " << total << "\n";
      if (check) {
        A.buildSingle();
        B.buildSingle();
        std::vector<hkey_t> result;
        for (uint32_t t = 0; t < gpuCount; t++) {
          std::vector<MultiHashGraph::keypair> part(common[t]);
          if (common[t])
            hipMemcpy(part.data(), out[t], common[t] * sizeof(MultiHashGraph::keypair), hipMemcpyDefault);
          for (const auto &kp : part) result.push_back((hkey_t)kp.right);
        }
        std::sort(A.h_vals, A.h_vals + countSizeA);
        std::sort(B.h_vals, B.h_vals + countSizeB);
        std::vector<hkey_t> ans;
        for (index_t i = 0; i < countSizeA; i++) {
          auto r = std::equal_range(B.h_vals, B.h_vals + countSizeB, A.h_vals[i]);
          for (auto p = r.first; p != r.second; ++p) ans.push_back(A.h_vals[i]);
        }
        if ((index_t)ans.size() != total) {
          std::cerr << "ERROR: INTERSECT OUTPUT HAS INCORRECT SIZE\nansSize: " << ans.size() << " outputSize: " << total
                    << std::endl;
          return 3;
        }
        std::sort(result.begin(), result.end());
        std::sort(ans.begin(), ans.end());
        if (result != ans) {
          std::cerr << "ERROR: INTERSECT OUTPUT HAS INCORRECT CONTENT" << std::endl;
          return 3;
        }
        std::cout << "intersect check ok" << std::endl;
      }
      for (auto p : out) (void)hipFree(p);
    }
  } catch (const std::exception &e) {
    std::cerr << "ERROR: " << e.what() << "\n";
    return 2;
  }
  return 0;
}
