// Compares the kernels' cbrtf_ref (turbo-metrics_amd/csrc/ssim2_math.h, built
// here as host code) with the platform libm cbrtf for every float in [lo, hi].
// Prints the number of floats checked and the number that differ.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "ssim2_math.h"

int main(int argc, char** argv) {
  const float lo = argc > 1 ? std::strtof(argv[1], nullptr) : 0.0f;
  const float hi = argc > 2 ? std::strtof(argv[2], nullptr) : 2.0f;
  const int nt = argc > 3 ? std::atoi(argv[3]) : 8;
  uint32_t a, b;
  std::memcpy(&a, &lo, 4);
  std::memcpy(&b, &hi, 4);
  std::vector<uint64_t> bad(nt, 0);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t]() {
      for (uint64_t u = a + t; u <= b; u += nt) {
        float x;
        const uint32_t v = (uint32_t)u;
        std::memcpy(&x, &v, 4);
        const float r = ssim2::cbrtf_ref(x), c = cbrtf(x);
        if (std::memcmp(&r, &c, 4) != 0) ++bad[t];
      }
    });
  for (auto& th : pool) th.join();
  uint64_t nbad = 0;
  for (auto v : bad) nbad += v;
  std::printf("%llu %llu\n", (unsigned long long)(uint64_t(b) - a + 1), (unsigned long long)nbad);
  return 0;
}
