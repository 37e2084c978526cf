// sing-hash: the reference's single-GPU driver (test/SingleHashGraphTest.cu):
//   sing-hash [countSize maxkey lrbBins]
// builds countSize keys drawn uniformly from [0, maxkey] into tableSize =
// maxkey bins (the reference leaves tableSize at its 2^26 default whatever
// argv says, test/SingleHashGraphTest.cu:21-22,34; here it follows maxkey)
// and prints the build time the way the reference does.
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include <iostream>

#include "SingleHashGraph.hpp"

int main(int argc, char **argv) {
  int64_t countSize = 1L << 24, maxkey = 1L << 26, lrbBins = 16000;
  if (argc >= 3 && argc < 4) {
    std::cerr << "Please specify all arguments.\n";
    return 1;
  }
  if (argc >= 4) {
    countSize = strtoll(argv[1], nullptr, 0);
    maxkey = strtoll(argv[2], nullptr, 0);
    lrbBins = strtoll(argv[3], nullptr, 0);
  }
  const int64_t tableSize = maxkey;
  try {
    hipEvent_t start, stop;
    hipEventCreate(&start);
    hipEventCreate(&stop);
    SingleHashGraph shg(countSize, maxkey, tableSize, lrbBins);
    shg.build(countSize, tableSize);  // warm-up (scratch allocation)
    hipDeviceSynchronize();
    hipEventRecord(start);
    shg.build(countSize, tableSize);
    hipEventRecord(stop);
    hipEventSynchronize(stop);
    float ms = 0;
    hipEventElapsedTime(&ms, start, stop);
    std::cout << "single buildTable() time: " << (ms / 1000.0) << "\n";
  } catch (const std::exception &e) {
    std::cerr << "ERROR: " << e.what() << "\n";
    return 2;
  }
  return 0;
}
