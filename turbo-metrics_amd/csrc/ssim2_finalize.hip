// Finalisation shared by the exact path and the fast form, one launch: per
// pair, the per-(scale, channel) reduction of the partial-sum records and the
// 108-feature score (crates/ssimulacra2-cuda/examples/cpu.rs:633-634, 676-679
// for the averages, 840-868 for the weighted sum and transform), in f64 on the
// device. One block per pair (persistent over the pairs of a batch): the
// pair's partial records pass through LDS and thread (group, sum) adds that
// sum over the group's records in record order, so the order never depends on
// which other pairs share the launch; the 108 features
// are evaluated in parallel and one thread forms the weighted sum in cpu.rs's
// order.
#include <algorithm>

#include "ssim2_internal.h"
#include "ssim2_tables.h"

namespace ssim2 {

namespace {

__constant__ double c_weight_fin[108] = SSIM2_WEIGHTS_INIT;

constexpr int kGroups = kMaxScales * 3;
constexpr int kStageRecords = 512;  // partial records staged in LDS at a time (24 KB)

__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace

__global__ __launch_bounds__(kThreads) void k_finalize(const PairDesc* __restrict__ pairs,
                                                       const int* __restrict__ meta, int npairs_static) {
  __shared__ double stage[kStageRecords * kNumSums];
  __shared__ double sums[kSumsPerPair];
  __shared__ double feat[kSumsPerPair];
  const int npairs = meta ? meta[0] : npairs_static;
  const int t = threadIdx.x;
  for (int pi = blockIdx.x; pi < npairs; pi += gridDim.x) {
    const PairDesc& P = pairs[pi];
    const int ns = P.nscales;
    // The pair's records are one contiguous run (scale by scale, channel by
    // channel, block by block): stream it through LDS in coalesced chunks and
    // let thread (group, sum) add its group's records serially, in record order.
    const int g = t / kNumSums, q = t % kNumSums, gs = g / 3, gc = g % 3;
    const bool own = t < kGroups * kNumSums && gs < ns;
    const int rec0 = P.sc[0].part;
    const int rec_end = P.sc[ns - 1].part + 3 * P.sc[ns - 1].colblocks;
    const int my0 = own ? P.sc[gs].part + gc * P.sc[gs].colblocks : 0;
    const int my1 = own ? my0 + P.sc[gs].colblocks : 0;
    double v = 0.0;
    for (int c0 = rec0; c0 < rec_end; c0 += kStageRecords) {
      const int c1 = min(c0 + kStageRecords, rec_end);
      const double* src = P.partials + (long long)c0 * kNumSums;
      for (int i = t; i < (c1 - c0) * kNumSums; i += kThreads) stage[i] = src[i];
      lds_sync();
      for (int r = max(my0, c0); r < min(my1, c1); ++r) v += stage[(r - c0) * kNumSums + q];
      lds_sync();
    }
    if (t < kGroups * kNumSums) {
      sums[t] = v;
      P.result[1 + t] = v;
    }
    if (t < 2 * kMaxScales) {
      const int ss = t / 2;
      P.result[1 + kSumsPerPair + t] = ss < ns ? (double)(t & 1 ? P.sc[ss].h : P.sc[ss].w) : 0.0;
    }
    lds_sync();
    // feature k in cpu.rs order: channel, present scale, {L1 of the three maps, L4 of the three}
    if (t < 3 * ns * 6) {
      const int c = t / (ns * 6), r = t % (ns * 6), s = r / 6, j = r % 6;
      const double opp = 1.0 / (double)((long long)P.sc[s].w * P.sc[s].h);
      const double* q = sums + (s * 3 + c) * kNumSums;
      feat[t] = j < 3 ? fabs(opp * q[2 * j]) : fabs(sqrt(sqrt(opp * q[2 * (j - 3) + 1])));
    }
    lds_sync();
    if (t == 0) {
      double ssim = 0.0;
      for (int k = 0; k < 3 * ns * 6; ++k) ssim = fma(c_weight_fin[k], feat[k], ssim);
      ssim *= 0.9562382616834844;
      ssim = fma(6.248496625763138e-5 * ssim * ssim, ssim,
                 fma(2.326765642916932, ssim, -0.020884521182843837 * ssim * ssim));
      ssim = ssim > 0.0 ? fma(pow(ssim, 0.6276336467831387), -10.0, 100.0) : 100.0;
      P.result[0] = ssim;
    }
    lds_sync();  // sums/feat are reused by the next pair
  }
}

hipError_t launch_finalize_pairs(const PairDesc* pairs, const int* meta, int npairs_static, int grid_units,
                                 int max_pairs, hipStream_t stream) {
  const int grid = std::max(1, std::min(grid_units, max_pairs));
  hipLaunchKernelGGL(k_finalize, dim3(grid), dim3(kThreads), 0, stream, pairs, meta, npairs_static);
  return hipGetLastError();
}

}  // namespace ssim2
