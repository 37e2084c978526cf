// sing-hash: the reference's single-GPU driver (test/SingleHashGraphTest.cu):
//   sing-hash [countSize maxkey lrbBins]
// builds countSize keys drawn uniformly from [0, maxkey] into tableSize =
// maxkey bins (the reference leaves tableSize at its 2^26 default whatever
// argv says, test/SingleHashGraphTest.cu:21-22,34; here it follows maxkey)
// and prints the build time the way the reference does.
//   sing-hash --files A.bin B.bin tableSize outdir
// (an addition, for the class-level parity tests) builds two graphs from the
// caller's int64 keys (raw little-endian files), intersects them and writes
// A's offsets (off.bin, int32), A's CSR keys (csr.bin, int64) and the
// intersection's keys (out.bin, int64) to outdir.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <iostream>
#include <string>
#include <vector>

#include "SingleHashGraph.hpp"

static std::vector<int64_t> read_keys(const char *path) {
  std::vector<int64_t> v;
  FILE *f = fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  int64_t x;
  while (fread(&x, sizeof(x), 1, f) == 1) v.push_back(x);
  fclose(f);
  return v;
}

template <typename T>
static void write_dev(const std::string &path, const T *d, size_t n) {
  std::vector<T> h(n);
  if (n && hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
    throw std::runtime_error("copy to host failed");
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  if (n) fwrite(h.data(), sizeof(T), n, f);
  fclose(f);
}

static int files_mode(char **argv) {
  const std::vector<int64_t> A = read_keys(argv[2]), B = read_keys(argv[3]);
  const int64_t M = strtoll(argv[4], nullptr, 0);
  const std::string dir = argv[5];
  int64_t *dA = nullptr, *dB = nullptr;
  if (hipMalloc(&dA, std::max<size_t>(A.size(), 1) * 8) != hipSuccess ||
      hipMalloc(&dB, std::max<size_t>(B.size(), 1) * 8) != hipSuccess)
    throw std::runtime_error("hipMalloc failed");
  if (A.size()) (void)hipMemcpy(dA, A.data(), A.size() * 8, hipMemcpyHostToDevice);
  if (B.size()) (void)hipMemcpy(dB, B.data(), B.size() * 8, hipMemcpyHostToDevice);
  SingleHashGraph ga(dA, (int64_t)A.size(), M), gb(dB, (int64_t)B.size(), M);
  ga.build((int64_t)A.size(), M);
  gb.build((int64_t)B.size(), M);
  SingleHashGraph::hkey_t *dout = nullptr;
  const uint64_t cnt = SingleHashGraph::intersect(ga, gb, &dout);
  write_dev(dir + "/off.bin", ga.offsets(), (size_t)M + 1);
  write_dev(dir + "/csr.bin", ga.keys(), A.size());
  write_dev(dir + "/out.bin", dout, cnt);
  std::cout << "intersect count: " << cnt << "\n";
  (void)hipFree(dout);
  (void)hipFree(dA);
  (void)hipFree(dB);
  return 0;
}

int main(int argc, char **argv) {
  if (argc == 6 && !strcmp(argv[1], "--files")) {
    try {
      return files_mode(argv);
    } catch (const std::exception &e) {
      std::cerr << "ERROR: " << e.what() << "\n";
      return 2;
    }
  }
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
