// Finalisation shared by the exact path and the fast form, in one launch:
// per-pair, per-(scale, channel) reduction of the partial-sum records, then
// the 108-feature score (crates/ssimulacra2-cuda/examples/cpu.rs:633-634,
// 676-679 for the averages, 840-868 for the weighted sum and transform), in
// f64 on the device. One block per (pair, scale*3 + channel) unit,
// persistent: 256 threads stride over the records, then a fixed-order tree
// (wave shuffles, waves in sequence), so the order does not depend on which
// other pairs share the launch. Each unit publishes its sums and counts
// itself in the pair's result record; the block that counts a pair's last
// unit evaluates the 108 features in parallel, forms the weighted sum in
// cpu.rs's order and resets the counter to zero for the next launch.
#include <algorithm>

#include "ssim2_internal.h"
#include "ssim2_plan.h"
#include "ssim2_tables.h"

namespace ssim2 {

namespace {

__constant__ double c_weight_fin[108] = SSIM2_WEIGHTS_INIT;

constexpr int kGroups = kMaxScales * 3;

__device__ __forceinline__ double wave_sum_f(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace

__global__ __launch_bounds__(kThreads) void k_finalize(const PairDesc* __restrict__ pairs,
                                                       const int* __restrict__ meta, int npairs_static) {
  __shared__ double red[kThreads / 64][kNumSums];
  __shared__ double feat[kSumsPerPair];
  __shared__ int last;
  const int npairs = meta ? meta[0] : npairs_static;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int u = blockIdx.x; u < npairs * kGroups; u += gridDim.x) {
    const PairDesc& P = pairs[u / kGroups];
    const int g = u % kGroups, s = g / 3, c = g % 3;
    double v[kNumSums] = {0, 0, 0, 0, 0, 0};
    if (s < P.nscales) {
      const ScaleDesc& S = P.sc[s];
      const double* p = P.partials + (long long)(S.part + c * S.colblocks) * kNumSums;
      for (int b = t; b < S.colblocks; b += kThreads)
#pragma unroll
        for (int q = 0; q < kNumSums; ++q) v[q] += p[(long long)b * kNumSums + q];
    }
    lds_sync();  // previous unit's red[] / feat[] reads are done
#pragma unroll
    for (int q = 0; q < kNumSums; ++q) {
      const double r = wave_sum_f(v[q]);
      if (lane == 0) red[wv][q] = r;
    }
    lds_sync();
    if (t < kNumSums) {
      double r = 0.0;
      for (int w = 0; w < kThreads / 64; ++w) r += red[w][t];
      P.result[1 + g * kNumSums + t] = r;
    }
    if (g == 0 && t < 2 * kMaxScales) {
      const int ss = t / 2;
      P.result[1 + kSumsPerPair + t] = ss < P.nscales ? (double)(t & 1 ? P.sc[ss].h : P.sc[ss].w) : 0.0;
    }
    // publish, then count this unit; the pair's last unit goes on to score it
    __threadfence();
    lds_sync();
    unsigned* counter = reinterpret_cast<unsigned*>(P.result + kResultCounter);
    if (t == 0) last = atomicAdd(counter, 1u) == (unsigned)(kGroups - 1);
    lds_sync();
    if (!last) continue;
    __threadfence();  // the other units' sums are visible from here
    const int ns = P.nscales, nf = 3 * ns * 6;
    // feature k in cpu.rs order: channel, present scale, {L1 of the three maps, L4 of the three}
    if (t < nf) {
      const int fc = t / (ns * 6), r = t % (ns * 6), fs = r / 6, j = r % 6;
      const double opp = 1.0 / (double)((long long)P.sc[fs].w * P.sc[fs].h);
      const double* q = P.result + 1 + (fs * 3 + fc) * kNumSums;
      feat[t] = j < 3 ? fabs(opp * q[2 * j]) : fabs(sqrt(sqrt(opp * q[2 * (j - 3) + 1])));
    }
    lds_sync();
    if (t == 0) {
      double ssim = 0.0;
      for (int k = 0; k < nf; ++k) ssim = fma(c_weight_fin[k], feat[k], ssim);
      ssim *= 0.9562382616834844;
      ssim = fma(6.248496625763138e-5 * ssim * ssim, ssim,
                 fma(2.326765642916932, ssim, -0.020884521182843837 * ssim * ssim));
      ssim = ssim > 0.0 ? fma(pow(ssim, 0.6276336467831387), -10.0, 100.0) : 100.0;
      P.result[0] = ssim;
      atomicExch(counter, 0u);
    }
  }
}

hipError_t launch_finalize_pairs(const PairDesc* pairs, const int* meta, int npairs_static, int grid_units,
                                 int max_pairs, hipStream_t stream) {
  (void)max_pairs;
  hipLaunchKernelGGL(k_finalize, dim3(grid_units), dim3(kThreads), 0, stream, pairs, meta, npairs_static);
  return hipGetLastError();
}

}  // namespace ssim2
