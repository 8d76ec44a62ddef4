// Compares the kernels' cube roots (turbo-metrics_amd/csrc/ssim2_math.h, built
// here as host code: cbrtf_ref, and cbrtf_nonneg with its scale table for
// x >= 0) with the platform libm cbrtf for every float in [lo, hi]. Prints the
// number of floats checked and the number where either differs.
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
  std::vector<double> scale(ssim2::kCbrtScales);
  for (int i = 0; i < ssim2::kCbrtScales; ++i) scale[i] = ssim2::cbrt_scale(i + ssim2::kCbrtExpMin);
  std::vector<uint64_t> bad(nt, 0);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t]() {
      for (uint64_t u = a + t; u <= b; u += nt) {
        float x;
        const uint32_t v = (uint32_t)u;
        std::memcpy(&x, &v, 4);
        const float r = ssim2::cbrtf_ref(x), c = cbrtf(x);
        const float n = x >= 0.0f ? ssim2::cbrtf_nonneg<true>(x, scale) : c;
        if (std::memcmp(&r, &c, 4) != 0 || std::memcmp(&n, &c, 4) != 0) ++bad[t];
      }
    });
  for (auto& th : pool) th.join();
  uint64_t nbad = 0;
  for (auto v : bad) nbad += v;
  std::printf("%llu %llu\n", (unsigned long long)(uint64_t(b) - a + 1), (unsigned long long)nbad);
  return 0;
}
