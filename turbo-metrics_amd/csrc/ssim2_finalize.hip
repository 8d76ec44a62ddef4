// Finalisation shared by the exact path and the fast form: per-pair,
// per-(scale, channel) reduction of the partial-sum records, then the
// 108-feature score (crates/ssimulacra2-cuda/examples/cpu.rs:633-634, 676-679
// for the averages, 840-868 for the weighted sum and transform), in f64 on the
// device. Two launches:
//   k_reduce_groups  one block per (pair, scale*3 + channel) unit, persistent:
//                    256 threads stride over the records, then a fixed-order
//                    tree (wave shuffles, waves in sequence); the order does
//                    not depend on which other pairs share the launch.
//   k_score          one wave per pair: the 108 features in parallel, then the
//                    weighted sum in cpu.rs's order on one lane.
#include <algorithm>

#include "ssim2_internal.h"
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

__global__ __launch_bounds__(kThreads) void k_reduce_groups(const PairDesc* __restrict__ pairs,
                                                            const int* __restrict__ meta, int npairs_static) {
  __shared__ double red[kThreads / 64][kNumSums];
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
    lds_sync();  // previous unit's red[] reads are done
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
  }
}

__global__ __launch_bounds__(64) void k_score(const PairDesc* __restrict__ pairs, const int* __restrict__ meta,
                                              int npairs_static) {
  __shared__ double feat[kSumsPerPair];
  const int npairs = meta ? meta[0] : npairs_static;
  const int lane = threadIdx.x;
  for (int pi = blockIdx.x; pi < npairs; pi += gridDim.x) {
    const PairDesc& P = pairs[pi];
    const int ns = P.nscales, nf = 3 * ns * 6;
    // feature k in cpu.rs order: channel, present scale, {L1 of the three maps, L4 of the three}
    for (int k = lane; k < nf; k += 64) {
      const int c = k / (ns * 6), r = k % (ns * 6), s = r / 6, j = r % 6;
      const double opp = 1.0 / (double)((long long)P.sc[s].w * P.sc[s].h);
      const double* q = P.result + 1 + (s * 3 + c) * kNumSums;
      feat[k] = j < 3 ? fabs(opp * q[2 * j]) : fabs(sqrt(sqrt(opp * q[2 * (j - 3) + 1])));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    if (lane == 0) {
      double ssim = 0.0;
      for (int k = 0; k < nf; ++k) ssim = fma(c_weight_fin[k], feat[k], ssim);
      ssim *= 0.9562382616834844;
      ssim = fma(6.248496625763138e-5 * ssim * ssim, ssim,
                 fma(2.326765642916932, ssim, -0.020884521182843837 * ssim * ssim));
      ssim = ssim > 0.0 ? fma(pow(ssim, 0.6276336467831387), -10.0, 100.0) : 100.0;
      P.result[0] = ssim;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  }
}

hipError_t launch_finalize_pairs(const PairDesc* pairs, const int* meta, int npairs_static, int grid_units,
                                 int max_pairs, hipStream_t stream) {
  hipLaunchKernelGGL(k_reduce_groups, dim3(grid_units), dim3(kThreads), 0, stream, pairs, meta, npairs_static);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_score, dim3(std::max(1, std::min(max_pairs, 1024))), dim3(64), 0, stream, pairs, meta,
                     npairs_static);
  return hipGetLastError();
}

}  // namespace ssim2
