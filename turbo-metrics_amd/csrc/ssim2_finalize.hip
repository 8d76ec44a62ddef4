// Finalisation shared by the exact path and the fast form: per-pair,
// per-(scale, channel) reduction of the partial-sum records, then the
// 108-feature score (crates/ssimulacra2-cuda/examples/cpu.rs:633-634, 676-679
// for the averages, 840-868 for the weighted sum and transform), in f64 on the
// device, in one launch. One block per pair (persistent over pairs): its waves
// take the 18 (scale, channel) groups in turn, each lane striding over the
// group's records and the wave summing its lanes with a fixed shuffle tree;
// after a block barrier the 108 features are formed in parallel and one lane
// forms the weighted sum in cpu.rs's order. Every order is fixed and depends on
// the pair alone, so a pair scores the same in any batch. (No cross-block
// step: a device-scope fence inside a kernel writes back and invalidates each
// XCD's L2.)
#include <algorithm>

#include "ssim2_internal.h"
#include "ssim2_tables.h"

namespace ssim2 {

namespace {

__constant__ double c_weight_fin[108] = SSIM2_WEIGHTS_INIT;

constexpr int kGroups = kMaxScales * 3;

__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace

constexpr int kFinThreads = 512;

__global__ __launch_bounds__(kFinThreads) void k_finalize(const PairDesc* __restrict__ pairs,
                                                          const int* __restrict__ meta, int npairs_static) {
  __shared__ double sums[kGroups][kNumSums];
  __shared__ double feat[kSumsPerPair];
  __shared__ double wts[kSumsPerPair];
  const int npairs = meta ? meta[0] : npairs_static;
  const int t = threadIdx.x, lane = t & 63;
  // the weights are read once per block, all in one round trip; the first
  // barrier below publishes them
  if (t < kSumsPerPair) wts[t] = c_weight_fin[t];
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  for (int pi = blockIdx.x; pi < npairs; pi += gridDim.x) {
    const PairDesc& P = pairs[pi];
    for (int g = wv; g < kGroups; g += kFinThreads / 64) {
      const int s = g / 3, c = g % 3;
      double v[kNumSums] = {0, 0, 0, 0, 0, 0};
      if (s < P.nscales) {
        const ScaleDesc& S = P.sc[s];
        const double* p = P.partials + (long long)(S.part + c * S.colblocks) * kNumSums;
        // records lane, lane + 64, ... in that order; four at a time so that
        // their loads are in flight together
        int b = lane;
        for (; b + 192 < S.colblocks; b += 256) {
          double r4[4][kNumSums];
#pragma unroll
          for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int q = 0; q < kNumSums; ++q) r4[k][q] = p[(long long)(b + 64 * k) * kNumSums + q];
#pragma unroll
          for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int q = 0; q < kNumSums; ++q) v[q] += r4[k][q];
        }
        for (; b < S.colblocks; b += 64)
#pragma unroll
          for (int q = 0; q < kNumSums; ++q) v[q] += p[(long long)b * kNumSums + q];
      }
      // the six sums step by step together, so that each step's shuffles
      // are in flight at once (the same tree per sum as a sum by sum)
      for (int off = 32; off > 0; off >>= 1) {
        double o[kNumSums];
#pragma unroll
        for (int q = 0; q < kNumSums; ++q) o[q] = __shfl_down(v[q], off, 64);
#pragma unroll
        for (int q = 0; q < kNumSums; ++q) v[q] += o[q];
      }
      if (lane == 0) {
#pragma unroll
        for (int q = 0; q < kNumSums; ++q) sums[g][q] = v[q];
      }
    }
    lds_sync();
    if (t < kSumsPerPair) P.result[1 + t] = sums[t / kNumSums][t % kNumSums];
    if (t >= kSumsPerPair && t < kSumsPerPair + 2 * kMaxScales) {
      const int k = t - kSumsPerPair, ss = k / 2;
      P.result[1 + kSumsPerPair + k] = ss < P.nscales ? (double)(k & 1 ? P.sc[ss].h : P.sc[ss].w) : 0.0;
    }
    const int ns = P.nscales, nf = 3 * ns * 6;
    // feature k in cpu.rs order: channel, present scale, {L1 of the three maps, L4 of the three}
    if (t < nf) {
      const int fc = t / (ns * 6), r = t % (ns * 6), fs = r / 6, j = r % 6;
      const double opp = 1.0 / (double)((long long)P.sc[fs].w * P.sc[fs].h);
      const double* q = sums[fs * 3 + fc];
      feat[t] = j < 3 ? fabs(opp * q[2 * j]) : fabs(sqrt(sqrt(opp * q[2 * (j - 3) + 1])));
    }
    lds_sync();
    if (t == 0) {
      // the weighted sum in cpu.rs's order; nf is a multiple of 18, and each
      // group of 18 features and weights is loaded before its fma chain (one
      // load latency per group instead of one per term)
      double ssim = 0.0;
      for (int k0 = 0; k0 < nf; k0 += 18) {
        double fw[18], ww[18];
#pragma unroll
        for (int k = 0; k < 18; ++k) {
          fw[k] = feat[k0 + k];
          ww[k] = wts[k0 + k];
        }
#pragma unroll
        for (int k = 0; k < 18; ++k) ssim = fma(ww[k], fw[k], ssim);
      }
      ssim *= 0.9562382616834844;
      ssim = fma(6.248496625763138e-5 * ssim * ssim, ssim,
                 fma(2.326765642916932, ssim, -0.020884521182843837 * ssim * ssim));
      ssim = ssim > 0.0 ? fma(pow(ssim, 0.6276336467831387), -10.0, 100.0) : 100.0;
      P.result[0] = ssim;
    }
    lds_sync();  // sums[] and feat[] are free for the next pair
  }
}

hipError_t launch_finalize_pairs(const PairDesc* pairs, const int* meta, int npairs_static, int grid_units,
                                 int max_pairs, hipStream_t stream) {
  (void)grid_units;
  const int grid = std::max(1, std::min(max_pairs, 1024));
  hipLaunchKernelGGL(k_finalize, dim3(grid), dim3(kFinThreads), 0, stream, pairs, meta, npairs_static);
  return hipGetLastError();
}

}  // namespace ssim2
