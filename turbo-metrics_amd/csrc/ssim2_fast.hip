// ssimulacra2 fast form for gfx950: batched, FP32/FP16 storage.
//
// Same metric as the exact path, different arithmetic, chosen for throughput:
//  * the blur is the 9-tap FIR that the reference's recursive Gaussian equals
//    in exact arithmetic (crates/ssimulacra2-cuda-kernel/build.rs:28-140; the
//    identity is checked in tests/test_oracle.py), evaluated as a separable
//    tile convolution in LDS instead of a sequential scan;
//  * the blurred quantities are formed in shifted coordinates
//    x' = x - shift(channel); the statistics are rebuilt with the exact
//    corrections for the zero-padded border (b1 = blur of ones, e = 1 - b1):
//        var = blur(x'^2) - blur(x')^2 + e*c*(2*blur(x') + c*b1)
//        cov = blur(r'd') - blur(r')*blur(d') + e*c*(blur(r') + blur(d') + c*b1)
//        x - mu = (x' - blur(x')) + c*e,  mu1 - mu2 = blur(r') - blur(d')
//    which removes most f32 cancellation in sigma - mu^2; the maps use the
//    two variances only as var1 + var2, so r'^2 + d'^2 is blurred as one
//    quantity (four blurs per channel: r'^2 + d'^2, r'd', r', d');
//  * the cube root is exp2(log2(x)/3) on the hardware log2/exp2, the error
//    maps are f32 with hardware reciprocals for their two quotients, tile sums
//    are f32 within a wave and f64 across waves and tiles;
//  * the XYB planes of every scale are stored in HBM as (value - shift) in
//    the storage type (f32 or IEEE half) by the pyramid kernel; the tile
//    kernel reads them and never the input. (bfloat16 storage was measured
//    0.30-0.35 from the FP64 score on 4K frames, seven times the 0.05 bar:
//    its 8-bit mantissa quantises flat regions into false variance.)
//
// Launches (each block takes a contiguous range of a device-resident item
// list, so one captured hipGraph serves every batch; the grids that follow
// the batch's size are set on the graph's nodes per batch):
//   k_fast_pyramid  one 32x32 scale-0 tile per block: shifted XYB of scale 0
//                   from each thread's 2x2 input block, scale-1 pixels from
//                   the same blocks, linear scales >= 1 in LDS, shifted XYB
//                   planes of every scale.
//   k_fast_tile     one TXxTY output tile of one scale of one pair per item:
//                   XYB tile + 4 px halo -> products -> horizontal FIR ->
//                   vertical FIR -> maps -> f64 block sums.
//   finalisation    k_finalize, one block per pair (ssim2_finalize.hip).
#include <type_traits>

#include "ssim2_input.h"
#include "ssim2_internal.h"
#include "ssim2_math.h"
#include "ssim2_tables.h"

namespace ssim2 {

namespace {

__constant__ float c_lut_fast[256] = SSIM2_SRGB8_LUT_INIT;

// 9-tap FIR (centre first), the double taps of the recursive Gaussian rounded to f32.
constexpr float kH0 = 0x1.0ef8d6p-2f, kH1 = 0x1.b413e8p-3f, kH2 = 0x1.bfd672p-4f,
                kH3 = 0x1.2700c8p-5f, kH4 = 0x1.347d70p-7f;

// Channel shifts: centre of the sRGB gamut in positive XYB (X 0.205..0.813,
// Y 0.010..0.855, B 0.272..0.938), exactly representable.
__device__ __forceinline__ float xyb_shift(int c) { return c == 0 ? 0.5f : (c == 1 ? 0.4375f : 0.625f); }

// Cube root as exp2(log2(x) / 3) on the hardware log2/exp2 (about 1 ulp
// each): a few ulp in all, far below the binary16 storage quantum and the
// FIR's own rounding, and two quarter-rate operations instead of three.
// x <= 0 gives 0 (log2(0) = -inf, exp2(-inf) = 0).
__device__ __forceinline__ float cbrt_fast(float x) {
  return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(fmaxf(x, 0.0f)) * 0.333333343f);
}

// Shifted XYB (value minus xyb_shift) of linear RGB: cpu.rs:458-496 and
// make_positive_xyb (cpu.rs:509-515) with the constant terms regrouped. With
// cr, cg, cb the cube roots of the mixed channels, the positive XYB is
// X = 7 (cr - cg) + 0.42, Y = (cr + cg) / 2 - cbrt(bias) + 0.01 and
// B = cb - (cr + cg) / 2 + 0.55: the bias root cancels in X and B.
__device__ __forceinline__ void xyb_fast_shifted(float r, float g, float b, float& X, float& Y, float& B) {
  const float cr = cbrt_fast(fmaf(kM00, r, fmaf(kM01, g, fmaf(kM02, b, kB0))));
  const float cg = cbrt_fast(fmaf(kM10, r, fmaf(kM11, g, fmaf(kM12, b, kB0))));
  const float cb = cbrt_fast(fmaf(kM20, r, fmaf(kM21, g, fmaf(kM22, b, kB0))));
  constexpr float kX0 = 0.42f - 0.5f, kY0 = (0.01f - 0.4375f) - kB0Root, kBs = 0.55f - 0.625f;
  const float sum = cr + cg;
  X = fmaf(cr - cg, 7.0f, kX0);
  Y = fmaf(sum, 0.5f, kY0);
  B = fmaf(sum, -0.5f, cb) + kBs;
}

// Storage codecs.
template <typename T>
struct Codec;
template <>
struct Codec<float> {
  __device__ static float enc(float v) { return v; }
  __device__ static float dec(float v) { return v; }
};
template <>
struct Codec<_Float16> {
  __device__ static _Float16 enc(float v) { return (_Float16)v; }
  __device__ static float dec(_Float16 v) { return (float)v; }
};

__device__ __forceinline__ void fast_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Global address space (see ssim2_kernels.hip: generic accesses force full
// waits on every outstanding load).
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* fast_plane(const PairDesc& P, int s, int img, int c) {
  const ScaleDesc& S = P.sc[s];
  if (img == 0 && P.ref_planes)  // a resident reference pyramid (uniform per pair)
    return (__attribute__((address_space(1))) T*)(reinterpret_cast<T*>(P.ref_planes + S.cache) + (long long)c * S.plane);
  return (__attribute__((address_space(1))) T*)(reinterpret_cast<T*>(reinterpret_cast<unsigned char*>(P.arena) + S.xyb) +
                                               (long long)(img * 3 + c) * S.plane);
}
typedef __attribute__((address_space(1))) const unsigned char gcu8;
template <typename T>
using gT = __attribute__((address_space(1))) T;

// Two adjacent elements as one store; q must be aligned to two elements.
template <typename T>
__device__ __forceinline__ void store_pair(gT<T>* q, T a, T b) {
  typedef T V2 __attribute__((ext_vector_type(2)));
  V2 v;
  v.x = a;
  v.y = b;
  *(gT<V2>*)q = v;
}

// Sum of the in-range taps of the 1-D FIR at position g of a line of n (the
// 1-D blur of ones); exactly 1 away from the border.
__device__ __forceinline__ float tap_sum(int g, int n) {
  if (g >= 4 && g < n - 4) return 1.0f;
  const float h[5] = {kH0, kH1, kH2, kH3, kH4};
  float acc = 0.0f;
#pragma unroll
  for (int k = 4; k >= 1; --k) {
    acc += (g - k >= 0) ? h[k] : 0.0f;
    acc += (g + k < n) ? h[k] : 0.0f;
  }
  return acc + h[0];
}

// Index of the pair that owns flat item idx: prefix[p] is the first item of
// pair p (prefix[0] = 0, nondecreasing).
__device__ __forceinline__ int find_pair(const int* __restrict__ prefix, int npairs, int idx) {
  int lo = 0, hi = npairs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= idx)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Workgroups go to the 8 XCDs round-robin by id. This renumbering gives the
// blocks of each XCD one contiguous run of ids (a bijection on [0, grid)),
// so neighbouring items, whose halo rows and output lines overlap, meet in
// the same L2 instead of being fetched (and partially written) by several.
__device__ __forceinline__ int xcd_grouped_block() {
  constexpr int kXcds = 8;
  const int g = (int)gridDim.x, b = (int)blockIdx.x;
  const int x = b % kXcds, q = g / kXcds, r = g % kXcds;
  return x * q + min(x, r) + b / kXcds;
}

// A contiguous range of the flat item list per block (equal-work items), and
// the owning pair tracked by a cursor that only moves forward: one prefix load
// per pair boundary instead of a binary search per item.
struct ItemRange {
  const int* prefix;
  int first, end, pair, next;
  __device__ ItemRange(const int* __restrict__ pre, int npairs, int n) : prefix(pre) {
    const int per = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    first = min(n, xcd_grouped_block() * per);
    end = min(n, first + per);
    pair = first < n ? find_pair(pre, npairs, first) : 0;
    next = first < n ? pre[pair + 1] : 0;
  }
  __device__ int pair_of(int idx) {
    while (idx >= next) {
      ++pair;
      next = prefix[pair + 1];
    }
    return pair;
  }
};

// The 18 per-lane sums of a tile, each summed over the 64 lanes of the wave
// in a fixed order, by halving butterflies: at each step a lane pair keeps
// one value of each pair of values (the lane's side bit picks which) and
// adds its partner's copy of it, so the 18 values take 9 + 4 + 2 + 1 + 1
// exchanges instead of 18 full reductions. Partners: lane ^ 15 within a row
// (side bit 3), lane ^ 7 within 8 (bit 2), lane ^ 2 (bit 1), lane ^ 1 (bit 0),
// lane ^ 16 (bit 4); an odd value out at a step is summed over the pair in
// both lanes. Lane L < 16 ends with sum 8 b0 + 4 b1 + 2 b2 + b3 (bits of
// L), lane 16 with sum 16 and lane 24 with sum 17; *which is -1 for the
// other lanes.
template <int CTRL>
__device__ __forceinline__ float dpp_pair(float keep, float send) {
  return keep + __builtin_amdgcn_update_dpp(0.0f, send, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_half(float u, float v, bool side) {
  return dpp_pair<CTRL>(side ? v : u, side ? u : v);
}
__device__ __forceinline__ float wave_sums18(const float (&v)[18], int lane, int* which) {
  constexpr int kMirror = 0x140, kHalfMirror = 0x141, kXor2 = 0x4E, kXor1 = 0xB1;
  const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4, b3 = lane & 8, b4 = lane & 16;
  float s1[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) s1[i] = dpp_half<kMirror>(v[2 * i], v[2 * i + 1], b3);
  float s2[5];
#pragma unroll
  for (int k = 0; k < 4; ++k) s2[k] = dpp_half<kHalfMirror>(s1[2 * k], s1[2 * k + 1], b2);
  s2[4] = dpp_pair<kHalfMirror>(s1[8], s1[8]);
  const float s30 = dpp_half<kXor2>(s2[0], s2[1], b1), s31 = dpp_half<kXor2>(s2[2], s2[3], b1);
  const float s32 = dpp_pair<kXor2>(s2[4], s2[4]);
  const float s40 = dpp_half<kXor1>(s30, s31, b0), s41 = dpp_pair<kXor1>(s32, s32);
  // lane ^ 16 (swizzle, bit-mask mode: and 0x1f, xor 0x10), then lane ^ 32
  const float keep = b4 ? s41 : s40, send = b4 ? s40 : s41;
  const float s5 = keep + __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(send), 0x401F));
  const float t = s5 + __int_as_float(__builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, __float_as_int(s5)));
  *which = lane < 16 ? 8 * (lane & 1) + 4 * ((lane >> 1) & 1) + 2 * ((lane >> 2) & 1) + ((lane >> 3) & 1)
                     : (lane == 16 ? 16 : (lane == 24 ? 17 : -1));
  return t;
}

}  // namespace

// ---------------------------------------------------------------- pyramid

// sRGB bytes of the 2x2 scale-0 block behind one scale-1 pixel, for both
// images, held as loaded (combined only when converted, so that the loads
// stay in flight meanwhile).
struct TileBytes {
  unsigned short lo[2][4];  // bytes 0, 1
  unsigned char hi[2][4];   // byte 2
};

// One 32x32 scale-0 tile per item. Scale 0 never goes through LDS: each
// thread loads the 2x2 input block of its scale-1 pixel (clamped at the image
// edge, as the downscale clamps), stores the shifted XYB of the block's
// in-image pixels (scale 0 of every plane is computed here once, not per tile
// halo in the tile kernel) and forms the scale-1 mean directly, in the order
// ((x0,y0) + (x1,y0)) + (x0,y1)) + (x1,y1) of cpu.rs:545-579; only scales
// >= 1 are kept in LDS.
template <typename T>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(7))) void k_fast_pyramid(const PairDesc* __restrict__ pairs,
                                                           const int* __restrict__ prefix,
                                                           const int* __restrict__ meta) {
  constexpr int T0 = kPyrTile, T1 = T0 / 2;
  static_assert(T1 * T1 == kThreads, "one scale-1 pixel per thread");
  __shared__ float bufB[2][6][T1][T1 + 1];       // scales 1, 3, 5 of two items in flight
  __shared__ float bufA[6][T1 / 2][T1 / 2 + 1];  // scales 2, 4
  __shared__ float lut_s[256];  // divergent lookups: LDS, not constant memory
  const int npairs = meta[0], n = prefix[npairs];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int sx = tid % T1, sy = tid / T1;  // this thread's scale-1 pixel in the tile
  lds_cfloat* lut = (lds_cfloat*)lut_s;
  int nb = 0;  // items with scale 1 so far: selects the scale-1 buffer and the level wave
  // Scales 2.. (at most 8x8 pixels) of an item run on one wave after the
  // next item's barrier, rotating over the items: a wave's LDS accesses take
  // effect in order, so the levels need no block barrier of their own, and
  // the other waves meanwhile work on the next item. Scale 2 takes a lane per
  // pixel with both images side by side; scales 3.. (at most 16 pixels per
  // image) a lane per (image, pixel).
  auto level_pixel = [&](const PairDesc& Q, int s, int img, int gx, int gy, const float (&rgb)[3]) {
    float X, Y, B;
    xyb_fast_shifted(rgb[0], rgb[1], rgb[2], X, Y, B);
    const long long off = (long long)gy * Q.sc[s].pitch + gx;
    fast_plane<T>(Q, s, img, 0)[off] = Codec<T>::enc(X);
    fast_plane<T>(Q, s, img, 1)[off] = Codec<T>::enc(Y);
    fast_plane<T>(Q, s, img, 2)[off] = Codec<T>::enc(B);
  };
  auto levels = [&](const PairDesc& Q, int tx, int ty, float (*B1)[T1][T1 + 1]) {
    {
      // scale 2 from the scale-1 buffer
      constexpr int Ts = T0 >> 2;
      const int pw = Q.sc[1].w, ph = Q.sc[1].h;
      const int lx = lane % Ts, ly = lane / Ts;
      const int gx = tx * Ts + lx, gy = ty * Ts + ly;
      if (gx < Q.sc[2].w && gy < Q.sc[2].h) {
        const int x0 = min(2 * gx, pw - 1) - 2 * tx * Ts, x1 = min(2 * gx + 1, pw - 1) - 2 * tx * Ts;
        const int y0 = min(2 * gy, ph - 1) - 2 * ty * Ts, y1 = min(2 * gy + 1, ph - 1) - 2 * ty * Ts;
        float rgb[2][3];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const float v = (((B1[k][y0][x0] + B1[k][y0][x1]) + B1[k][y1][x0]) + B1[k][y1][x1]) * 0.25f;
          rgb[k / 3][k % 3] = v;
          bufA[k][ly][lx] = v;
        }
#pragma unroll
        for (int img = 0; img < 2; ++img)
          if (!((Q.skip_img >> img) & 1)) level_pixel(Q, 2, img, gx, gy, rgb[img]);
      }
    }
    for (int s = 3; s < Q.nscales; ++s) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const int Ts = T0 >> s, np = Ts * Ts;
      const int pw = Q.sc[s - 1].w, ph = Q.sc[s - 1].h;
      const int ox0 = tx * Ts, oy0 = ty * Ts;
      const int img = lane / np, li = lane % np;
      const int lx = li % Ts, ly = li / Ts;
      const int gx = ox0 + lx, gy = oy0 + ly;
      if (img >= 2 || ((Q.skip_img >> img) & 1) || gx >= Q.sc[s].w || gy >= Q.sc[s].h) continue;
      const int x0 = min(2 * gx, pw - 1) - 2 * ox0, x1 = min(2 * gx + 1, pw - 1) - 2 * ox0;
      const int y0 = min(2 * gy, ph - 1) - 2 * oy0, y1 = min(2 * gy + 1, ph - 1) - 2 * oy0;
      float rgb[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int k = img * 3 + c;
        float sum;
        if (s & 1)
          sum = ((bufA[k][y0][x0] + bufA[k][y0][x1]) + bufA[k][y1][x0]) + bufA[k][y1][x1];
        else
          sum = ((B1[k][y0][x0] + B1[k][y0][x1]) + B1[k][y1][x0]) + B1[k][y1][x1];
        rgb[c] = sum * 0.25f;
        if (s & 1)
          B1[k][ly][lx] = rgb[c];
        else
          bufA[k][ly][lx] = rgb[c];
      }
      level_pixel(Q, s, img, gx, gy, rgb);
    }
  };
  int pend = -1, pend_tx = 0, pend_ty = 0, pend_nb = 0;  // item whose levels are due
  ItemRange items(prefix, npairs, n);
  // Scale-0 coordinates of the 2x2 block, clamped into the image.
  auto block_xy = [&](const PairDesc& Q, int local, int k, int& gx, int& gy) {
    const int ntx = (Q.w + T0 - 1) / T0;
    const int bx = (local % ntx) * T0, by = (local / ntx) * T0;
    gx = min(bx + 2 * sx + (k & 1), Q.w - 1);
    gy = min(by + 2 * sy + (k >> 1), Q.h - 1);
  };
  // Input bytes of an item's sRGB u8 blocks, fetched before its barrier (and
  // the previous item's levels) so that their latency overlaps them.
  // The loads are issued whatever the input kind (other kinds read the
  // first bytes of the descriptor array, unused), so the count of loads in
  // flight is the same on every path and later waits can count them.
  auto fetch = [&](int j, int pj, TileBytes& tb) {
    const PairDesc& Q = pairs[pj];
    const bool u8 = Q.in_kind == kInputSrgbU8;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int gx, gy;
      block_xy(Q, j - prefix[pj], k, gx, gy);
#pragma unroll
      for (int img = 0; img < 2; ++img) {
        const bool used = u8 && !((Q.skip_img >> img) & 1);
        gin_u8* px = used ? (gin_u8*)(img ? Q.in_dis : Q.in_ref) + gy * (img ? Q.in_dis_pitch : Q.in_ref_pitch) + gx * 3
                          : (gin_u8*)pairs;
        __builtin_memcpy(&tb.lo[img][k], (const void*)px, 2);
        tb.hi[img][k] = px[2];
      }
    }
  };
  TileBytes tb;
  if (items.first >= items.end) return;  // uniform: a block beyond the last item
  // The table entry is loaded first and the first item's bytes after it, so
  // storing the entry waits for its load only and the item's loads stay in
  // flight across the barrier (which waits for LDS stores).
  static_assert(kThreads == 256, "one table entry per thread");
  const float lut_entry = c_lut_fast[tid];
  int pi = items.pair_of(items.first);
  fetch(items.first, pi, tb);
  lut_s[tid] = lut_entry;
  fast_barrier();  // the table is complete
  for (int idx = items.first; idx < items.end; ++idx) {
    if (idx > items.first) {
      pi = items.pair_of(idx);
      fetch(idx, pi, tb);  // in flight across the barrier and the levels
      // The one barrier of a later item: the previous item's scale 1 is in
      // LDS, and every wave is done with the scale-1 buffer of the item
      // before that.
      fast_barrier();
      if (pend >= 0) {
        if (wv == (pend_nb & 3)) levels(pairs[pend], pend_tx, pend_ty, bufB[pend_nb & 1]);
        pend = -1;
      }
    }
    const PairDesc& P = pairs[pi];
    const int local = idx - prefix[pi];
    const int ntx = (P.w + T0 - 1) / T0;
    const int tx = local % ntx, ty = local / ntx;
    const bool u8 = P.in_kind == kInputSrgbU8;  // uniform over the block
    // Linear RGB of this thread's 2x2 block in both images.
    float px[2][4][3];
    if (u8) {
      // the fetched bytes are the last loads issued; stores issued before
      // them may still be in flight, and only a full wait proves them done
      __builtin_amdgcn_s_waitcnt(vmcnt_only<0>());
#pragma unroll
      for (int img = 0; img < 2; ++img)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const unsigned lo = tb.lo[img][k];
          px[img][k][0] = lut[lo & 255u];
          px[img][k][1] = lut[lo >> 8];
          px[img][k][2] = lut[tb.hi[img][k]];
        }
    } else {
#pragma unroll
      for (int img = 0; img < 2; ++img)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          int gx, gy;
          block_xy(P, local, k, gx, gy);
          if ((P.skip_img >> img) & 1)
            px[img][k][0] = px[img][k][1] = px[img][k][2] = 0.0f;
          else
            read_linear(input_of(P, img), gx, gy, lut, px[img][k]);
        }
    }
    // Scale 0: shifted XYB of the block's in-image rows, each row of the
    // block as one two-element store. ux is even and so is the pitch, so the
    // pair is aligned and stays inside its plane row; at an odd width its
    // second element lands in the row's padding column, which no reader
    // takes as a pixel.
    const int ux = tx * T0 + 2 * sx, uy = ty * T0 + 2 * sy;
    const bool in0 = ux < P.w;
#pragma unroll
    for (int img = 0; img < 2; ++img) {
      if ((P.skip_img >> img) & 1) continue;  // uniform: a resident or not-needed pyramid
      float v[4][3];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        xyb_fast_shifted(px[img][k][0], px[img][k][1], px[img][k][2], v[k][0], v[k][1], v[k][2]);
      }
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        if (!in0 || uy + r >= P.h) continue;
        const long long off = (long long)(uy + r) * P.sc[0].pitch + ux;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          store_pair(fast_plane<T>(P, 0, img, c) + off, Codec<T>::enc(v[2 * r][c]), Codec<T>::enc(v[2 * r + 1][c]));
        }
      }
    }
    // scale-1 mean in the order of cpu.rs:545-579
    float lin[2][3];
#pragma unroll
    for (int img = 0; img < 2; ++img)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        lin[img][c] = (((px[img][0][c] + px[img][1][c]) + px[img][2][c]) + px[img][3][c]) * 0.25f;
    if (P.nscales < 2) continue;  // uniform: the whole block has this item
    // Scale 1, one pixel a thread, into the scale-1 buffer of this item's
    // parity (the other may be in use by the previous item's levels).
    float (*B1)[T1][T1 + 1] = bufB[nb & 1];
    const int gx = tx * T1 + sx, gy = ty * T1 + sy;
#pragma unroll
    for (int img = 0; img < 2; ++img) {
      if ((P.skip_img >> img) & 1) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) B1[img * 3 + c][sy][sx] = lin[img][c];
      if (gx < P.sc[1].w && gy < P.sc[1].h) {
        float X, Y, B;
        xyb_fast_shifted(lin[img][0], lin[img][1], lin[img][2], X, Y, B);
        const long long off = (long long)gy * P.sc[1].pitch + gx;
        fast_plane<T>(P, 1, img, 0)[off] = Codec<T>::enc(X);
        fast_plane<T>(P, 1, img, 1)[off] = Codec<T>::enc(Y);
        fast_plane<T>(P, 1, img, 2)[off] = Codec<T>::enc(B);
      }
    }
    if (P.nscales > 2) {
      pend = pi;
      pend_tx = tx;
      pend_ty = ty;
      pend_nb = nb;
    }
    ++nb;
  }
  if (pend >= 0) {  // uniform: the last item's levels
    fast_barrier();
    if (wv == (pend_nb & 3)) levels(pairs[pend], pend_tx, pend_ty, bufB[pend_nb & 1]);
  }
}

// ---------------------------------------------------------------- fused tile

// RU = output rows per vertical-pass unit (TY / RU units per column).
// One channel at a time is in LDS (a quarter of the block's LDS for all
// three); the next channel's halo is loaded into registers, four elements
// a load, while the current one is filtered.
// Waves per SIMD the registers of a tile shape are sized for (1: no bound):
// 6 (80 VGPRs) for 32x16 with 16-bit planes; f32 planes need a few more
// registers (their 4-wide loads are twice as wide) and would spill at 80.
template <typename T, int TX, int TY>
constexpr int tile_waves() {
  return TX == 32 && TY == 16 && sizeof(T) == 2 ? 6 : 1;
}

template <typename T, int TX, int TY, int RU>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(tile_waves<T, TX, TY>())))
void k_fast_tile(const PairDesc* __restrict__ pairs,
                                                        const int* __restrict__ prefix,
                                                        const int* __restrict__ meta) {
  static_assert(TX % 4 == 0 && TY % RU == 0, "tile shape");
  constexpr int IW = TX + 8, IH = TY + 8;
  constexpr int XS = IW + 4, HS = TX + 4;  // LDS row strides (16-byte rows)
  constexpr int CW = IW / 4, NCH = IH * CW;  // 4-element chunks of one halo plane
  constexpr int NCT = (NCH + kThreads - 1) / kThreads;
  // Blurred quantities: r'^2 + d'^2, r'd', r', d'. The maps use the two
  // variances only through their sum, so one blur serves both.
  constexpr int kQ = 4;
  typedef T V4 __attribute__((ext_vector_type(4)));
  typedef float F4 __attribute__((ext_vector_type(4)));
  __shared__ F4 xs4[2][IH][XS / 4];  // shifted XYB of one channel, zero outside the image
  __shared__ F4 hq4[kQ][IH][HS / 4];  // horizontal pass of that channel
  __shared__ double red[kThreads / 64][18];
  float (*xs)[IH][XS] = reinterpret_cast<float (*)[IH][XS]>(&xs4[0][0][0]);
  float (*hq)[IH][HS] = reinterpret_cast<float (*)[IH][HS]>(&hq4[0][0][0]);
  const int npairs = meta[0], n = prefix[npairs];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  ItemRange items(prefix, npairs, n);
  for (int idx = items.first; idx < items.end; ++idx) {
    const int pi = items.pair_of(idx);
    const PairDesc& P = pairs[pi];
    int s = 0, tile = idx - prefix[pi];
    while (s + 1 < P.nscales && tile >= P.sc[s + 1].tile0) ++s;
    tile -= P.sc[s].tile0;
    const ScaleDesc& S = P.sc[s];
    const int w = S.w, h = S.h;
    const int tx = tile % S.tiles_x, ty = tile / S.tiles_x;
    const int x0 = tx * TX, y0 = ty * TY;
    const bool interior = x0 >= 4 && x0 + TX <= w - 4 && y0 >= 4 && y0 + TY <= h - 4;

    // Halo chunks of channel c into registers. A chunk starts at a multiple
    // of 4 (x0 is one) and so does every plane row (the pitch is a multiple
    // of 16), so a chunk is one aligned 4-element load; a chunk that starts
    // inside the image ends inside its row, the columns past w being the
    // row's padding, zeroed when the chunk is written to LDS.
    V4 raw[NCT][2];
    auto load = [&](int c) {
#pragma unroll
      for (int t = 0; t < NCT; ++t) {
        const int e = tid + t * kThreads;
        const int iy = e / CW, j = e % CW;
        const int gx = x0 - 4 + 4 * j, gy = y0 - 4 + iy;
#pragma unroll
        for (int img = 0; img < 2; ++img) {
          if (e < NCH && gx >= 0 && gx < w && gy >= 0 && gy < h)
            raw[t][img] = *(const gT<V4>*)(fast_plane<T>(P, s, img, c) + (long long)gy * S.pitch + gx);
          else
            raw[t][img] = V4{};
        }
      }
    };
    auto to_lds = [&]() {
#pragma unroll
      for (int t = 0; t < NCT; ++t) {
        const int e = tid + t * kThreads;
        if (e >= NCH) continue;
        const int iy = e / CW, j = e % CW;
        const int gx = x0 - 4 + 4 * j;
#pragma unroll
        for (int img = 0; img < 2; ++img) {
          F4 f;
#pragma unroll
          for (int k = 0; k < 4; ++k) f[k] = gx + k < w ? Codec<T>::dec(raw[t][img][k]) : 0.0f;
          xs4[img][iy][j] = f;
        }
      }
    };
    load(0);
    float acc[3][6];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < 6; ++j) acc[c][j] = 0.0f;

    for (int c = 0; c < 3; ++c) {
      fast_barrier();  // the previous channel's (or item's) LDS reads are done
      to_lds();
      if (c < 2) load(c + 1);
      fast_barrier();
      // Horizontal FIR of the four quantities, 4 outputs per unit.
      for (int e = tid; e < IH * (TX / 4); e += kThreads) {
        const int iy = e / (TX / 4), j0 = e % (TX / 4);
        float a[12], d[12];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          const F4 va = xs4[0][iy][j0 + m], vd = xs4[1][iy][j0 + m];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            a[4 * m + k] = va[k];
            d[4 * m + k] = vd[k];
          }
        }
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
          float qv[12];
#pragma unroll
          for (int k = 0; k < 12; ++k)
            qv[k] = q == 0 ? a[k] * a[k] + d[k] * d[k] : (q == 1 ? a[k] * d[k] : (q == 2 ? a[k] : d[k]));
          F4 r4;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            float r = kH4 * (qv[j] + qv[j + 8]);
            r = fmaf(kH3, qv[j + 1] + qv[j + 7], r);
            r = fmaf(kH2, qv[j + 2] + qv[j + 6], r);
            r = fmaf(kH1, qv[j + 3] + qv[j + 5], r);
            r = fmaf(kH0, qv[j + 4], r);
            r4[j] = r;
          }
          hq4[q][iy][j0] = r4;
        }
      }
      fast_barrier();
      // Vertical FIR (RU rows per unit), statistics, error maps.
      // A tile whose outputs all lie 4 px or more inside the image has
      // b1 = 1 and e = 0 everywhere, so its statistics need none of the
      // zero-padding terms (the choice is uniform over the block).
      const float sh = xyb_shift(c);
      auto maps = [&](auto edge_tag) {
        constexpr bool kEdge = decltype(edge_tag)::value;
        for (int e = tid; e < (TY / RU) * TX; e += kThreads) {
          const int oy0 = (e / TX) * RU, ox = e % TX;
          float bq[kQ][RU];
#pragma unroll
          for (int q = 0; q < kQ; ++q) {
            float col[RU + 8];
#pragma unroll
            for (int k = 0; k < RU + 8; ++k) col[k] = hq[q][oy0 + k][ox];
#pragma unroll
            for (int j = 0; j < RU; ++j) {
              float r = kH4 * (col[j] + col[j + 8]);
              r = fmaf(kH3, col[j + 1] + col[j + 7], r);
              r = fmaf(kH2, col[j + 2] + col[j + 6], r);
              r = fmaf(kH1, col[j + 3] + col[j + 5], r);
              r = fmaf(kH0, col[j + 4], r);
              bq[q][j] = r;
            }
          }
          const int gx = x0 + ox;
          const float b1x = kEdge ? tap_sum(gx, w) : 1.0f;
#pragma unroll
          for (int j = 0; j < RU; ++j) {
            const float ssum = bq[0][j], s12 = bq[1][j], m1 = bq[2][j], m2 = bq[3][j];
            // var1 + var2 and cov (plus, near the border, the zero-padding
            // terms e*c*(2*(m1 + m2) + 2*c*b1) and e*c*((m1 + m2) + c*b1)).
            // Products rounded on their own: for identical images r'^2 + d'^2
            // is exactly 2*r'd' (and so are its blur and m1^2 + m2^2), which
            // makes var1 + var2 exactly 2*cov and the quotient exactly 1.
            float vsum = ssum - (m1 * m1 + m2 * m2), cov = s12 - m1 * m2;
            float ec = 0.0f;
            if constexpr (kEdge) {
              const int gy = y0 + oy0 + j;
              if (gx >= w || gy >= h) continue;
              const float b1 = b1x * tap_sum(gy, h);
              ec = (1.0f - b1) * sh;
              const float t = (m1 + m2) + sh * b1;
              vsum = fmaf(2.0f * ec, t, vsum);
              cov = fmaf(ec, t, cov);
            }
            const float md = m1 - m2;
            const float num_m = fmaf(md, -md, 1.0f);
            const float num_s = fmaf(2.0f, cov, kC2);
            const float den = vsum + kC2;
            // quotients by hardware reciprocal; equal operands (identical
            // images) keep their exact quotient 1, so identical pairs still
            // score 100
            const float nm = num_m * num_s;
            const float dd = fmaxf(1.0f - (nm == den ? 1.0f : nm * __builtin_amdgcn_rcpf(den)), 0.0f);
            const float d2 = dd * dd;
            acc[c][0] += dd;
            acc[c][1] = fmaf(d2, d2, acc[c][1]);
            const float src = xs[0][oy0 + j + 4][ox + 4], dis = xs[1][oy0 + j + 4][ox + 4];
            float rn = dis - m2, rd = src - m1;
            if constexpr (kEdge) {
              rn += ec;
              rd += ec;
            }
            const float en = 1.0f + fabsf(rn), ed = 1.0f + fabsf(rd);
            const float d1 = en == ed ? 0.0f : fmaf(en, __builtin_amdgcn_rcpf(ed), -1.0f);
            const float art = fmaxf(d1, 0.0f), det = fmaxf(-d1, 0.0f);
            const float a2 = art * art, e2 = det * det;
            acc[c][2] += art;
            acc[c][3] = fmaf(a2, a2, acc[c][3]);
            acc[c][4] += det;
            acc[c][5] = fmaf(e2, e2, acc[c][5]);
          }
        }
      };
      if (interior)
        maps(std::false_type{});
      else
        maps(std::true_type{});
    }
    // Block sums in a fixed order: f32 within each wave, then f64 over the
    // waves in sequence.
    {
      float v18[18];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < 6; ++j) v18[c * 6 + j] = acc[c][j];
      int which;
      const float v = wave_sums18(v18, lane, &which);
      if (which >= 0) red[wv][which] = (double)v;
    }
    fast_barrier();
    if (tid < 18) {
      double v = 0.0;
      for (int q = 0; q < kThreads / 64; ++q) v += red[q][tid];
      const int c = tid / 6, j = tid % 6;
      ((__attribute__((address_space(1))) double*)P.partials)[(long long)(S.part + c * S.colblocks + tile) * kNumSums + j] = v;
    }
  }
}

// ---------------------------------------------------------------- fused tile on the matrix cores

// The separable 9-tap FIR of one 16-column strip is a banded matrix product,
// so the f16-storage form can run both passes on the matrix cores
// (v_mfma_f32_16x16x32_f16: products of binary16 operands are exact, sums in
// f32). A wave owns 16 output columns x 32 output rows of one channel; a
// block is four such waves side by side (a 64x32 tile). Lane (g, l) below
// means lane 16 g + l.
//
// Horizontal pass, for each of three row blocks b (input rows
// y0 - 4 + 16 b + l): A = 16 rows x 32 input columns (lane (g, l) holds the 8
// columns xw - 8 + 8 g .. + 7 of row l, one aligned 16-byte load; the outer 4
// columns on each side only widen the aligned window and carry weight 0),
// B = the band matrix W[k][n] = tap(|k - 8 - n|), D = 16 rows x 16 output
// columns, lane (g, n) holding rows 4 g .. 4 g + 3 of column n.
// Vertical pass, transposed so that a lane ends with four neighbouring
// columns of one output row: D' = T^T x Wv, where A = T^T takes row blocks b
// and b + 1 of the horizontal results as they sit in the registers (lane
// (g, n): elements 0-3 = block b rows 4 g + i, elements 4-7 = block b + 1
// rows 4 g + i; A and B share the K order, so any order of the T rows works
// as long as the weights follow it) and Wv[k][r] = tap(|rho(k) - r - 4|).
// D' lane (g, r) holds output row r, columns 4 g .. 4 g + 3.
// Precision: binary16 operands carry 11 significant bits, too few for the
// blurred squares (their f16 rounding alone moves scores by 0.3), so every
// operand that is not already a binary16 value is split: taps as hi + lo
// (hi = f16(tap), lo = f16(tap - hi)); r'^2, d'^2 and r'd' of the stored
// binary16 planes as hi + lo exactly (hi = f16(x y); lo = fma(x, y, -hi) in
// binary16 is the exact remainder), so blur(r'^2 + d'^2) is the sum of
// hi.Whi + hi.Wlo + lo.Whi over both squares; the horizontal results T as
// hi + lo before the vertical pass (hi rounded to nearest, the exact
// remainder rounded toward zero: T - hi - lo below 2^-21 relative). The
// lo.lo products (below 2^-21 relative) are dropped.
typedef _Float16 mx_h8 __attribute__((ext_vector_type(8)));
typedef _Float16 mx_h4 __attribute__((ext_vector_type(4)));
typedef float mx_f4 __attribute__((ext_vector_type(4)));
constexpr int kMxCols = 16, kMxRows = 32, kMxWaves = 4;
static_assert(kMxWaves * kMxCols == 64 && kMxRows == 32, "the 64x32 tile of tile_cfg 4");

__device__ __forceinline__ mx_f4 mx_mfma(mx_h8 a, mx_h8 b, mx_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Power-of-two operand scales (exact in binary16 and f32) that keep the lo
// parts of products, taps and horizontal results out of the binary16
// subnormal range, where they would lose bits: the planes x4 before they are
// squared, the taps x16 (x4 for the horizontal pass of the means). The
// vertical results are then 64 times the blurred means and 4096 = 64^2 times
// the blurred squares, so the statistics come out in one common scale: var
// and cov x4096, the means x64. The squares stay finite in binary16 for
// |shifted XYB| up to about 11 (linear inputs in [0, 1] give at most 0.45).
constexpr float kMxTapScale = 16.0f, kMxMeanTapScale = 4.0f;
constexpr float kMxMeanScale = 64.0f, kMxVarScale = kMxMeanScale * kMxMeanScale;

// Tap |t| of the FIR (0 beyond 4) times `scale`, split hi/lo in binary16.
__device__ __forceinline__ void mx_tap(int t, float scale, _Float16& hi, _Float16& lo) {
  t = t < 0 ? -t : t;
  const float v = scale * (t == 0 ? kH0 : t == 1 ? kH1 : t == 2 ? kH2 : t == 3 ? kH3 : t == 4 ? kH4 : 0.0f);
  hi = (_Float16)v;
  lo = (_Float16)(v - (float)hi);
}

// Constant operands of a lane: horizontal band W (lane (g, n = l): k = 8 g + j)
// and vertical band Wv (lane (g, r = l): rho = 4 g + j, or 16 + 4 g + j - 4).
struct MxTaps {
  mx_h8 h_hi, h_lo, m_hi, m_lo, v_hi, v_lo;  // horizontal (squares), horizontal (means), vertical
  // -1 and 0 as values the compiler cannot see through (mx_split)
  float n1;
  unsigned z;
};

// One channel's inputs of a wave: the horizontal A operands of row blocks
// 0..2 for both images, and the pixels of the wave's maps (output row
// y0 + 16 ob + l, columns xw + 4 g .. + 3).
struct MxIn {
  mx_h8 a[3], d[3];
  mx_h4 src[2], dis[2];
};

// Tile geometry of a wave: its first output column xw, the tile's first row
// y0, the scale's size and pitch; near the border also the per-item edge
// data: the 1-D blurs of ones of the lane's map columns and rows.
struct MxGeo {
  int xw, y0, w, h, g, l;
  long long pitch;
  float b1x[4], b1y[2];
};

// Per-item edge data of a lane (see MxGeo).
__device__ __forceinline__ void mx_edge_setup(MxGeo& G) {
#pragma unroll
  for (int i = 0; i < 4; ++i) G.b1x[i] = tap_sum(G.xw + 4 * G.g + i, G.w);
#pragma unroll
  for (int ob = 0; ob < 2; ++ob) G.b1y[ob] = tap_sum(G.y0 + 16 * ob + G.l, G.h);
}

// Two f32 values rounded to binary16, packed (one v_cvt_pk_f16_f32).
__device__ __forceinline__ unsigned mx_pk(float x, float y) {
  typedef _Float16 h2 __attribute__((ext_vector_type(2)));
  const h2 v = {(_Float16)x, (_Float16)y};
  return __builtin_bit_cast(unsigned, v);
}

// The remainders x - hi.lo and y - hi.hi in f32 (exact: hi is x, y rounded
// to binary16), one mixed-precision fma each (v_fma_mix_f32, the binary16
// operand read in place; n1 = -1), packed to binary16 rounding toward zero.
__device__ __forceinline__ unsigned mx_rem(unsigned hi, float x, float y, float n1) {
  typedef _Float16 h2 __attribute__((ext_vector_type(2)));
  const h2 h = __builtin_bit_cast(h2, hi);
  return __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(fmaf((float)h.x, n1, x), fmaf((float)h.y, n1, y)));
}

// hi and lo binary16 parts of the f32 values u[0..3], v[0..3] as one operand.
// Plain code, no inline assembly: the compiler must see every instruction
// that reads or writes a register near the matrix products, since it alone
// inserts the wait states they require (a matrix product reads its
// accumulator over several passes; inline assembly is not checked against
// that). The opaque z (0) keeps each packed hi word whole, so the remainders
// read its halves in place instead of converting x and y a second time; the
// opaque n1 (-1) keeps each remainder an fma rather than a subtraction,
// which has no mixed-precision form.
__device__ __forceinline__ void mx_split(mx_f4 u, mx_f4 v, float n1, unsigned z, mx_h8& hi, mx_h8& lo) {
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  u4 h, l;
  h[0] = mx_pk(u[0], u[1]) ^ z;
  h[1] = mx_pk(u[2], u[3]) ^ z;
  h[2] = mx_pk(v[0], v[1]) ^ z;
  h[3] = mx_pk(v[2], v[3]) ^ z;
  l[0] = mx_rem(h[0], u[0], u[1], n1);
  l[1] = mx_rem(h[1], u[2], u[3], n1);
  l[2] = mx_rem(h[2], v[0], v[1], n1);
  l[3] = mx_rem(h[3], v[2], v[3], n1);
  hi = __builtin_bit_cast(mx_h8, h);
  lo = __builtin_bit_cast(mx_h8, l);
}

// Horizontal pass of row block b: T of r', d', Q = r'^2 + d'^2 - 2 r'd', r'd'.
__device__ __forceinline__ void mx_hpass(const mx_h8& a, const mx_h8& d, const MxTaps& K, mx_f4 (&H)[4]) {
  const mx_f4 z4 = {0.0f, 0.0f, 0.0f, 0.0f};
  const mx_h8 four = {4.0f16, 4.0f16, 4.0f16, 4.0f16, 4.0f16, 4.0f16, 4.0f16, 4.0f16};
  const mx_h8 a4 = a * four, d4 = d * four;
  const mx_h8 a2h = a4 * a4, d2h = d4 * d4, adh = a4 * d4;
  const mx_h8 a2l = __builtin_elementwise_fma(a4, a4, -a2h);
  const mx_h8 d2l = __builtin_elementwise_fma(d4, d4, -d2h);
  const mx_h8 adl = __builtin_elementwise_fma(a4, d4, -adh);
  H[0] = mx_mfma(a, K.m_hi, mx_mfma(a, K.m_lo, z4));
  H[1] = mx_mfma(d, K.m_hi, mx_mfma(d, K.m_lo, z4));
  // r'^2 and d'^2 in chains of their own, in the order of r'd''s, and kept as
  // Q = (r'^2 + d'^2) - 2 r'd' (blurred): for identical images both chains
  // equal r'd''s bit for bit, Q is exactly 0, and the vertical pass rebuilds
  // blur(r'^2 + d'^2) as exactly twice blur(r'd'), so var1 + var2 is exactly
  // 2 cov (the quotient 1, the score 100)
  const mx_f4 sa = mx_mfma(a2h, K.h_hi, mx_mfma(a2h, K.h_lo, mx_mfma(a2l, K.h_hi, z4)));
  const mx_f4 sd = mx_mfma(d2h, K.h_hi, mx_mfma(d2h, K.h_lo, mx_mfma(d2l, K.h_hi, z4)));
  H[3] = mx_mfma(adh, K.h_hi, mx_mfma(adh, K.h_lo, mx_mfma(adl, K.h_hi, z4)));
#pragma unroll
  for (int i = 0; i < 4; ++i) H[2][i] = fmaf(-2.0f, H[3][i], sa[i] + sd[i]);
}

// Vertical pass of one quantity from row blocks b (upper) and b + 1 (lower).
__device__ __forceinline__ mx_f4 mx_vpass(mx_f4 upper, mx_f4 lower, const MxTaps& K) {
  const mx_f4 z4 = {0.0f, 0.0f, 0.0f, 0.0f};
  mx_h8 thi, tlo;
  mx_split(upper, lower, K.n1, K.z, thi, tlo);
  return mx_mfma(thi, K.v_hi, mx_mfma(thi, K.v_lo, mx_mfma(tlo, K.v_hi, z4)));
}

// k x - m for a binary16 x (half `hi` of the pair), in one mixed-precision fma.
template <int hi>
__device__ __forceinline__ float mx_scaled_diff(unsigned pair, float k, float m) {
  typedef _Float16 h2 __attribute__((ext_vector_type(2)));
  const h2 h = __builtin_bit_cast(h2, pair);
  return fmaf((float)(hi ? h.y : h.x), k, -m);
}

// Error maps of output block ob (row y0 + 16 ob + l, columns xw + 4 g + i)
// from the blurred quantities V (means x64, Q and r'd' x4096) and the
// pixels, the formulas of k_fast_tile rearranged so that equal operands give
// exact zeros without comparisons:
//   var1 + var2 - 2 cov = Q - (mu1 - mu2)^2, so den = num_s + (Q - md^2) and
//   1 - ssim = (den - num_m num_s) / den;
//   (1 + |dis - mu2|) / (1 + |src - mu1|) - 1 = (|e2| - |e1|) / (64 + |e1|)
//   with e = 64 (x - mu);
// near the border cov takes its zero-padding term (var1 + var2 - 2 cov has
// none) and e its shift term. The six sums are kept as
// {sum d, sum d^4, sum |d1|, sum d1, sum d1^4, sum d1^3 |d1|}: artifact and
// detail loss are their half sums and differences (mx_finish_sums).
template <bool kEdge>
__device__ __forceinline__ void mx_maps(const mx_f4 (&V)[4], mx_h4 src, mx_h4 dis, float sh, const MxGeo& G, int ob,
                                        float (&a6)[6]) {
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  const u2 sp = __builtin_bit_cast(u2, src), dp = __builtin_bit_cast(u2, dis);
  const float k64 = kMxMeanScale;
  const int gy = G.y0 + 16 * ob + G.l, gx0 = G.xw + 4 * G.g;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float m1 = V[0][i], m2 = V[1][i], q = V[2][i], s12 = V[3][i];
    float cov = s12 - m1 * m2;
    float ec = 0.0f;
    if constexpr (kEdge) {
      if (gx0 + i >= G.w || gy >= G.h) continue;
      const float b1 = G.b1x[i] * G.b1y[ob];
      ec = (1.0f - b1) * (kMxMeanScale * sh);
      cov = fmaf(ec, (m1 + m2) + (kMxMeanScale * sh) * b1, cov);
    }
    const float md = m1 - m2, md2 = md * md;
    const float num_m = fmaf(md2, -1.0f / kMxVarScale, 1.0f);
    const float num_s = fmaf(2.0f, cov, kMxVarScale * kC2);
    const float den = num_s + (q - md2);
    const float dd = fmaxf((den - num_m * num_s) * __builtin_amdgcn_rcpf(den), 0.0f);
    const float d2 = dd * dd;
    a6[0] += dd;
    a6[1] = fmaf(d2, d2, a6[1]);
    const unsigned sw = i < 2 ? sp[0] : sp[1], dw = i < 2 ? dp[0] : dp[1];
    float e2 = (i & 1) ? mx_scaled_diff<1>(dw, k64, m2) : mx_scaled_diff<0>(dw, k64, m2);
    float e1 = (i & 1) ? mx_scaled_diff<1>(sw, k64, m1) : mx_scaled_diff<0>(sw, k64, m1);
    if constexpr (kEdge) {
      e2 += ec;
      e1 += ec;
    }
    const float d1 = (fabsf(e2) - fabsf(e1)) * __builtin_amdgcn_rcpf(k64 + fabsf(e1));
    const float d1s = d1 * d1;
    a6[2] += fabsf(d1);
    a6[3] += d1;
    a6[4] = fmaf(d1s, d1s, a6[4]);
    a6[5] = fmaf(d1s, d1 * fabsf(d1), a6[5]);
  }
}

// The kept sums of mx_maps in the order of the other tile kernels: ssim
// d and d^4, artifact and its fourth powers, detail loss and its.
__device__ __forceinline__ void mx_finish_sums(float (&a6)[6]) {
  const float abs1 = a6[2], sgn1 = a6[3], p4 = a6[4], sgn4 = a6[5];
  a6[2] = 0.5f * (abs1 + sgn1);
  a6[3] = 0.5f * (p4 + sgn4);
  a6[4] = 0.5f * (abs1 - sgn1);
  a6[5] = 0.5f * (p4 - sgn4);
}

// One channel of a wave: row blocks 0 and 1, output block 0, then row block
// 2 (in block 0's registers) and output block 1.
template <bool kEdge>
__device__ __forceinline__ void mx_channel(const MxIn& in, const MxTaps& K, float sh, const MxGeo& G,
                                           float (&a6)[6]) {
  mx_f4 H0[4], H1[4], V[4];
  mx_hpass(in.a[0], in.d[0], K, H0);
  mx_hpass(in.a[1], in.d[1], K, H1);
#pragma unroll
  for (int q = 0; q < 4; ++q) V[q] = mx_vpass(H0[q], H1[q], K);
  mx_maps<kEdge>(V, in.src[0], in.dis[0], sh, G, 0, a6);
  mx_hpass(in.a[2], in.d[2], K, H0);
#pragma unroll
  for (int q = 0; q < 4; ++q) V[q] = mx_vpass(H1[q], H0[q], K);
  mx_maps<kEdge>(V, in.src[1], in.dis[1], sh, G, 1, a6);
}

// ---- inputs through LDS by DMA
// A wave's inputs of one channel are its six horizontal A operands (row
// blocks 0..2, both images), each one 16-byte load per lane; the wave loads
// them straight into its LDS buffer (global_load_lds_dwordx4: lane i's 16
// bytes land in slot i of the operand's 1 KB). A step reads its operands
// from the buffer into registers and then issues the next step's DMA, which
// lands while the step is filtered: no register holds inputs in flight, and
// the next item's first channel loads during the current item's last. (The
// compiler's wait insertion makes any LDS read wait for every DMA in flight,
// so the reads come before the next DMA.) The addresses are clamped into the
// plane; what lies outside the image is cleared when an operand is read
// (edge tiles only: an interior tile's used rows and columns are all inside,
// and its unused rows past y0 + 35 carry weight 0). The map pixels are read
// from the operands that contain them.
constexpr int kMxOps = 6;
typedef unsigned mx_u4 __attribute__((ext_vector_type(4)));
typedef unsigned mx_u2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) mx_u4 lds_u4;
typedef __attribute__((address_space(3))) const mx_u4 lds_cu4;
typedef __attribute__((address_space(3))) const mx_u2 lds_cu2;
typedef __attribute__((address_space(1))) const void gcvoid;
typedef __attribute__((address_space(4))) const PairDesc CPairDesc;

// Issues the six DMA loads of one channel step into `dst` (kMxOps x 64 slots).
__device__ __forceinline__ void mx_dma(lds_u4* dst, const gT<_Float16>* pr, const gT<_Float16>* pd, const MxGeo& Gs) {
  const int xc = min(max(Gs.xw - 8 + 8 * Gs.g, 0), (int)Gs.pitch - 8);
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    // rows past y0 + 35 carry weight 0: their slots keep the zeros written
    // at kernel start
    if (b == 2 && Gs.l >= 8) break;
    const int yr = min(max(Gs.y0 - 4 + 16 * b + Gs.l, 0), Gs.h - 1);
    // sides are at most kMaxDim = 32768: a plane holds fewer than 2^32 elements
    const unsigned o = (unsigned)yr * (unsigned)Gs.pitch + (unsigned)xc;
    __builtin_amdgcn_global_load_lds((gcvoid*)(pr + o), (void __attribute__((address_space(3)))*)(dst + (2 * b) * 64), 16, 0, 0);
    __builtin_amdgcn_global_load_lds((gcvoid*)(pd + o), (void __attribute__((address_space(3)))*)(dst + (2 * b + 1) * 64), 16, 0, 0);
  }
}

// Map pixels of output block ob (row y0 + 16 ob + l, columns xw + 4 g .. + 3)
// of image img, from the operand slot that holds them.
__device__ __forceinline__ mx_h4 mx_px(lds_cu4* src, int ob, int img, const MxGeo& G) {
  const int b = G.l < 12 ? ob : ob + 1, lp = G.l < 12 ? G.l + 4 : G.l - 12;
  lds_cu2* p = (lds_cu2*)(src + (2 * b + img) * 64 + 16 * (1 + (G.g >> 1)) + lp);
  return __builtin_bit_cast(mx_h4, p[G.g & 1]);
}

// One channel step's inputs, read from its operand slots; kEdge: what lies
// outside the image is zero (the lane's window columns past the image by one
// dword mask per step, rows outside the image as a whole).
template <bool kEdge>
__device__ __forceinline__ void mx_read(MxIn& in, lds_cu4* src, const MxGeo& G, int lane) {
  unsigned cm[4];
  if constexpr (kEdge) {
    const int xc = G.xw - 8 + 8 * G.g;
    const int nv = xc >= 0 ? G.w - xc : 0;  // columns of the window inside the image
#pragma unroll
    for (int k = 0; k < 4; ++k) cm[k] = 2 * k + 1 < nv ? 0xFFFFFFFFu : (2 * k < nv ? 0x0000FFFFu : 0u);
  }
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    mx_u4 ua = src[(2 * b) * 64 + lane], ud = src[(2 * b + 1) * 64 + lane];
    if constexpr (kEdge) {
      const int yr = G.y0 - 4 + 16 * b + G.l;
      const bool rv = yr >= 0 && yr < G.h;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned m = rv ? cm[k] : 0u;
        ua[k] &= m;
        ud[k] &= m;
      }
    }
    in.a[b] = __builtin_bit_cast(mx_h8, ua);
    in.d[b] = __builtin_bit_cast(mx_h8, ud);
  }
#pragma unroll
  for (int ob = 0; ob < 2; ++ob) {
    in.src[ob] = mx_px(src, ob, 0, G);
    in.dis[ob] = mx_px(src, ob, 1, G);
  }
}

// s_waitcnt immediate for lgkmcnt(0) only (vmcnt and expcnt at their maxima).
constexpr int kLgkmZero = 15 | (7 << 4) | (0 << 8) | (3 << 14);

// Workgroup barrier for data exchanged through ds_write / ds_read only: it
// waits for the wave's LDS accesses but not for its DMA loads in flight
// (which count in vmcnt; a release fence on LDS would wait for them too).
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(kLgkmZero);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Sum over the four waves of block sum j, in wave order. The restrict
// pointer tells the compiler's wait insertion that these reads do not touch
// the DMA buffers.
__device__ __forceinline__ double mx_red_sum(const __attribute__((address_space(3))) double* __restrict__ red, int j) {
  double t = 0.0;
#pragma unroll
  for (int q = 0; q < kMxWaves; ++q) t += red[q * 18 + j];
  return t;
}

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_fast_tile_mfma(const PairDesc* __restrict__ pairs,
                                                             const int* __restrict__ prefix,
                                                             const int* __restrict__ meta) {
  static_assert(kThreads == 64 * kMxWaves, "four waves per block");
  __shared__ double red[kMxWaves][18];
  __shared__ mx_u4 lbuf[kMxWaves][kMxOps * 64];
  const int npairs = meta[0], n = prefix[npairs];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  MxGeo G;
  G.g = lane >> 4;
  G.l = lane & 15;
  MxTaps K;
  K.n1 = -1.0f;
  K.z = 0u;
  asm volatile("" : "+s"(K.n1), "+s"(K.z));
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    _Float16 hi, lo;
    mx_tap(8 * G.g + j - 8 - G.l, kMxTapScale, hi, lo);
    K.h_hi[j] = hi;
    K.h_lo[j] = lo;
    mx_tap(8 * G.g + j - 8 - G.l, kMxMeanTapScale, hi, lo);
    K.m_hi[j] = hi;
    K.m_lo[j] = lo;
    const int rho = j < 4 ? 4 * G.g + j : 16 + 4 * G.g + (j - 4);
    mx_tap(rho - G.l - 4, kMxTapScale, hi, lo);
    K.v_hi[j] = hi;
    K.v_lo[j] = lo;
  }
  ItemRange items(prefix, npairs, n);
  if (items.first >= items.end) return;  // uniform: a block beyond the last item
  // A channel step: pair, scale, tile in the scale and the tile's geometry;
  // interior = every input window inside the image (every output 4 px or
  // more from the border: b1 = 1, e = 0), uniform over the block. Steps are
  // kept in plain variables picked by selects, so they stay in scalar
  // registers and the descriptors they address are read by scalar loads.
  struct Step {
    int pair, s, tile, w, h, y0, xw;
    long long pitch;
    int interior;
  };
  auto locate = [&](int idx) {
    Step t;
    t.pair = items.pair_of(idx);
    // descriptors are constant over the launch: scalar loads
    const CPairDesc& Q = ((const CPairDesc*)pairs)[t.pair];
    int s = 0, tile = idx - prefix[t.pair];
    while (s + 1 < Q.nscales && tile >= Q.sc[s + 1].tile0) ++s;
    t.s = s;
    t.tile = tile - Q.sc[s].tile0;
    const int tiles_x = Q.sc[s].tiles_x;
    t.w = Q.sc[s].w;
    t.h = Q.sc[s].h;
    t.pitch = Q.sc[s].pitch;
    const int x0 = (t.tile % tiles_x) * (kMxWaves * kMxCols);
    t.y0 = (t.tile / tiles_x) * kMxRows;
    t.xw = x0 + kMxCols * wv;
    t.interior = x0 >= 8 && x0 + kMxWaves * kMxCols + 8 <= t.w && t.y0 >= 4 && t.y0 + kMxRows + 4 <= t.h;
    return t;
  };
  auto geo = [](MxGeo& g, const Step& t) {
    g.w = t.w;
    g.h = t.h;
    g.pitch = t.pitch;
    g.y0 = t.y0;
    g.xw = t.xw;
  };
  lds_u4* const buf = (lds_u4*)&lbuf[wv][0];
  // Step k is channel k % 3 of item first + k / 3. Its operands are read
  // from the buffer into registers, then the DMA of step k + 1 refills the
  // buffer while step k is filtered.
  const int nsteps = 3 * (items.end - items.first);
  int ld_idx = items.first, ld_c = 0;
  Step ld = locate(ld_idx);
  if (G.l >= 8) {  // the slots of row block 2 that no DMA writes (LDS starts undefined)
    const mx_u4 z = {0u, 0u, 0u, 0u};
    buf[4 * 64 + lane] = z;
    buf[5 * 64 + lane] = z;
  }
  auto issue = [&]() {
    // the step (ld_idx, ld_c) into the buffer; returns it and advances
    const PairDesc& Q = pairs[ld.pair];
    MxGeo gd = G;
    geo(gd, ld);
    mx_dma(buf, fast_plane<_Float16>(Q, ld.s, 0, ld_c), fast_plane<_Float16>(Q, ld.s, 1, ld_c), gd);
    const Step t = ld;
    if (++ld_c == 3 && ld_idx + 1 < items.end) {
      ld_c = 0;
      ld = locate(++ld_idx);
    }
    return t;
  };
  Step nxt = issue();
  float acc[3][6];
  for (int k = 0; k < nsteps; ++k) {
    const Step cur = nxt;
    const int c = k % 3;
    if (c == 0) {
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
#pragma unroll
        for (int j = 0; j < 6; ++j) acc[cc][j] = 0.0f;
    }
    geo(G, cur);
    if (c == 0 && !cur.interior) mx_edge_setup(G);  // kept for the item's three channels
    MxIn in;
    if (cur.interior)
      mx_read<false>(in, buf, G, lane);
    else
      mx_read<true>(in, buf, G, lane);
    // every read of the buffer has returned before the DMA refills it
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(kLgkmZero);
    if (k + 1 < nsteps) nxt = issue();
    auto channel = [&](auto c_tag) {
      constexpr int C = decltype(c_tag)::value;
      if (cur.interior)
        mx_channel<false>(in, K, xyb_shift(C), G, acc[C]);
      else if (G.xw < G.w)  // uniform per wave: a wave past the right edge adds zeros
        mx_channel<true>(in, K, xyb_shift(C), G, acc[C]);
    };
    if (c == 0)
      channel(std::integral_constant<int, 0>{});
    else if (c == 1)
      channel(std::integral_constant<int, 1>{});
    else
      channel(std::integral_constant<int, 2>{});
    if (c == 2) {
      const CPairDesc& P = ((const CPairDesc*)pairs)[cur.pair];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) mx_finish_sums(acc[cc]);
      // Block sums in a fixed order: f32 within each wave, then f64 over the
      // four waves in sequence.
      float v18[18];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
#pragma unroll
        for (int j = 0; j < 6; ++j) v18[cc * 6 + j] = acc[cc][j];
      // lane-derived shuffle addresses are formed here, not held across the loop
      int which, ln = lane;
      asm volatile("" : "+v"(ln));
      const float vs = wave_sums18(v18, ln, &which);
      lds_barrier();  // the previous item's reads of red are done
      if (which >= 0) red[wv][which] = (double)vs;
      lds_barrier();
      if (tid < 18) {
        // indexes of this lane's sum, derived here rather than held in
        // registers across the loop
        int i = tid;
        asm volatile("" : "+v"(i));
        const double t = mx_red_sum((const __attribute__((address_space(3))) double*)&red[0][0], i);
        const int cc = i / 6, j = i % 6;
        ((__attribute__((address_space(1))) double*)P.partials)[(long long)(P.sc[cur.s].part +
                                                                            cc * P.sc[cur.s].colblocks + cur.tile) *
                                                                    kNumSums +
                                                                j] = t;
      }
    }
  }
}

// ---------------------------------------------------------------- launchers

// One block per tile: blocks resident together start at different times, so
// their phases (loads, cube roots, stores) overlap instead of running in step
// as a persistent grid's do. A smaller grid still covers every item (each
// block takes a contiguous range).
const void* fast_pyramid_kernel(int storage) {
  switch (storage) {
    case kStorageF32:
      return reinterpret_cast<const void*>(&k_fast_pyramid<float>);
    default:
      return reinterpret_cast<const void*>(&k_fast_pyramid<_Float16>);
  }
}

hipError_t launch_fast_pyramid(int storage, const PairDesc* pairs, const int* prefix, const int* meta, int grid,
                               hipStream_t stream) {
  grid = grid < 1 ? 1 : grid;
  switch (storage) {
    case kStorageF32:
      hipLaunchKernelGGL(k_fast_pyramid<float>, dim3(grid), dim3(kThreads), 0, stream, pairs, prefix, meta);
      break;
    default:
      hipLaunchKernelGGL(k_fast_pyramid<_Float16>, dim3(grid), dim3(kThreads), 0, stream, pairs, prefix, meta);
      break;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_tile_t(int tile_cfg, const PairDesc* pairs, const int* prefix, const int* meta,
                                int grid, hipStream_t stream) {
  switch (tile_cfg) {
    case 0:
      hipLaunchKernelGGL((k_fast_tile<T, 32, 16, 2>), dim3(grid), dim3(kThreads), 0, stream, pairs, prefix, meta);
      break;
    case 1:
      hipLaunchKernelGGL((k_fast_tile<T, 64, 16, 2>), dim3(grid), dim3(kThreads), 0, stream, pairs, prefix, meta);
      break;
    case 2:
      hipLaunchKernelGGL((k_fast_tile<T, 32, 32, 2>), dim3(grid), dim3(kThreads), 0, stream, pairs, prefix, meta);
      break;
    case 3:
      hipLaunchKernelGGL((k_fast_tile<T, 32, 24, 3>), dim3(grid), dim3(kThreads), 0, stream, pairs, prefix, meta);
      break;
    default:  // 64x32: the matrix-core kernel for binary16 planes
      if constexpr (std::is_same<T, _Float16>::value)
        hipLaunchKernelGGL(k_fast_tile_mfma, dim3(grid), dim3(kThreads), 0, stream, pairs, prefix, meta);
      else
        hipLaunchKernelGGL((k_fast_tile<T, 64, 32, 2>), dim3(grid), dim3(kThreads), 0, stream, pairs, prefix, meta);
      break;
  }
  return hipGetLastError();
}

template <typename T>
static const void* tile_kernel_t(int tile_cfg) {
  switch (tile_cfg) {
    case 0: return reinterpret_cast<const void*>(&k_fast_tile<T, 32, 16, 2>);
    case 1: return reinterpret_cast<const void*>(&k_fast_tile<T, 64, 16, 2>);
    case 2: return reinterpret_cast<const void*>(&k_fast_tile<T, 32, 32, 2>);
    case 3: return reinterpret_cast<const void*>(&k_fast_tile<T, 32, 24, 3>);
    default:
      if constexpr (std::is_same<T, _Float16>::value)
        return reinterpret_cast<const void*>(&k_fast_tile_mfma);
      else
        return reinterpret_cast<const void*>(&k_fast_tile<T, 64, 32, 2>);
  }
}

const void* fast_tile_kernel(int storage, int tile_cfg) {
  switch (storage) {
    case kStorageF32: return tile_kernel_t<float>(tile_cfg);
    default: return tile_kernel_t<_Float16>(tile_cfg);
  }
}

hipError_t launch_fast_tile(int storage, int tile_cfg, const PairDesc* pairs, const int* prefix,
                            const int* meta, int grid, hipStream_t stream) {
  grid = grid < 1 ? 1 : grid;
  switch (storage) {
    case kStorageF32: return launch_tile_t<float>(tile_cfg, pairs, prefix, meta, grid, stream);
    default: return launch_tile_t<_Float16>(tile_cfg, pairs, prefix, meta, grid, stream);
  }
}

hipError_t launch_fast_finalize(const PairDesc* pairs, const int* meta, int grid, int max_pairs,
                                hipStream_t stream) {
  return launch_finalize_pairs(pairs, meta, 0, grid, max_pairs, stream);
}

}  // namespace ssim2
