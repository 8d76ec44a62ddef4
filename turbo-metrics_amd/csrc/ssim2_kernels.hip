// ssimulacra2 HIP kernels for gfx950 (MI355X): the exact f32 path.
//
// The per-pair work is four launches, each a flat list of blocks (WorkItem)
// so one launch can cover every scale of every pair in a batch:
//
//   k_pyramid  one block per 32x32 scale-0 tile: input -> linear (LUT or f32)
//              -> XYB planes of every scale. The linear pyramid never leaves
//              LDS: a 32x32 tile at scale 0 owns exactly its 16x16 .. 1x1
//              sub-tiles at scales 1..5, including the ceil(n/2) clamped edge.
//   k_rows     one wave per (scale, channel, quantity, band of 21 rows): the
//              row-direction recursive Gaussian of cpu.rs (row form), one lane
//              per (row, cosine term), inputs streamed from the XYB planes
//              through a register prefetch queue.
//   k_cols     one block per (scale, channel, 16-column strip): the column
//              pass (column form), then SSIM / edge-difference error maps and
//              f64 L1/L4 partial sums, all in the same block.
//   finalize   (ssim2_finalize.hip) fixed-order reduction of the partial sums
//              and the 108-feature weighted score, in f64 on the device.
//
// Arithmetic follows crates/ssimulacra2-cuda/examples/cpu.rs exactly (see
// ssim2_math.h), so every f32 plane equals the CPU's bit for bit and the score
// differs only by the order of the f64 sums.
#include "ssim2_input.h"
#include "ssim2_internal.h"
#include "ssim2_math.h"
#include "ssim2_tables.h"

#include <algorithm>

namespace ssim2 {

__constant__ float c_lut[256] = SSIM2_SRGB8_LUT_INIT;
__constant__ double c_weight[108] = SSIM2_WEIGHTS_INIT;
__constant__ double c_cbrt_factor[5] = SSIM2_CBRT_FACTORS_INIT;

namespace {

template <int SIDES>
struct SideTraits;
// Quantity index q: 0 ref^2, 1 dis^2, 2 ref*dis, 3 ref, 4 dis.
template <>
struct SideTraits<kSidesBoth> {
  static constexpr int NQ = 5;
  static constexpr int NI = 2;
  __device__ static int q(int i) { return i; }
};
template <>
struct SideTraits<kSidesRef> {
  static constexpr int NQ = 2;
  static constexpr int NI = 1;
  __device__ static int q(int i) { return i == 0 ? 0 : 3; }
};
template <>
struct SideTraits<kSidesDis> {
  static constexpr int NQ = 3;
  static constexpr int NI = 1;
  __device__ static int q(int i) { return i == 0 ? 1 : (i == 1 ? 2 : 4); }
};

// Global-address-space views of plane pointers. Pointers read out of a
// descriptor in memory are generic, and generic (flat) accesses may complete
// out of order, which makes the compiler wait for every outstanding load
// before using any one of them. Global loads complete in order, so a
// register prefetch queue only waits for the oldest entry.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const float gcfloat;
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) vf4 gfloat4;
typedef __attribute__((address_space(1))) const vf4 gcfloat4;
__device__ __forceinline__ float4 to_f4(vf4 v) { return make_float4(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ gfloat* gbl(float* p) { return (gfloat*)p; }

// Workgroup barrier that orders LDS only. __syncthreads() also waits for every
// outstanding global load and store (vmcnt(0)), which would drain the
// register prefetch queues at each barrier; LDS visibility needs only
// lgkmcnt(0) before s_barrier.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;
constexpr unsigned kRsrcWord3 = 0x00020000u;    // raw buffer, 32-bit data format
constexpr unsigned kOffNone = 0xC0000000u;      // beyond every plane (< 2^30 floats)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const float* p, long long bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), 0, (int)bytes, kRsrcWord3);
}

__device__ __forceinline__ float4 u2f(u32x4 v) {
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}

// XYB plane of (side image, channel) at scale s for the given launch sides.
// Both: ref -> xyb[c], dis -> xyb[3+c]. Ref: ref -> cache[c]. Dis: dis -> xyb[3+c].
__device__ __forceinline__ float* xyb_plane(const PairDesc& P, int s, int img, int c, int sides) {
  const ScaleDesc& S = P.sc[s];
  if (img == 0 && sides != kSidesBoth) return P.arena + S.cache + (long long)c * S.plane;
  return P.arena + S.xyb + (long long)(img * 3 + c) * S.plane;
}

}  // namespace

// ---------------------------------------------------------------- pyramid

// 2x2 mean of the previous scale (cpu.rs:545-579): source coordinates clamped
// to its last column and row, summed (x0,y0), (x1,y0), (x0,y1), (x1,y1);
// at(x, y) reads the previous scale at global coordinates.
template <typename At>
__device__ __forceinline__ float mean2x2(int gx, int gy, int pw, int ph, At at) {
  const int x0 = min(2 * gx, pw - 1), x1 = min(2 * gx + 1, pw - 1);
  const int y0 = min(2 * gy, ph - 1), y1 = min(2 * gy + 1, ph - 1);
  float sum = 0.0f;
  sum += at(x0, y0);
  sum += at(x1, y0);
  sum += at(x0, y1);
  sum += at(x1, y1);
  return sum * 0.25f;
}

template <int SIDES>
__global__ __launch_bounds__(kThreads) void k_pyramid(const PairDesc* __restrict__ pairs,
                                                      const WorkItem* __restrict__ items,
                                                      InputArgs in) {
  constexpr int NI = SideTraits<SIDES>::NI;
  constexpr int T0 = kPyrTile, T1 = T0 / 2, T2 = T0 / 4, T3 = T0 / 8;
  static_assert(T1 * T1 == kThreads, "one scale-1 pixel per thread");
  static_assert(NI * (64 + 21) <= kThreads, "scale-2 threads and the scale 3-5 threads fit the block");
  __shared__ float lin1[NI * 3][T1][T1 + 1];  // scale 1, linear
  __shared__ float lin2[NI * 3][T2][T2 + 1];  // scale 2
  __shared__ float lin3[NI * 3][T3][T3 + 1];  // scale 3
  __shared__ float lut_s[256];
  __shared__ double cbrt_scale_s[kCbrtScales];
  const WorkItem it = items[blockIdx.x];
  const PairDesc& P = pairs[it.pair];
  const int tid = threadIdx.x;
  const int tx = it.a, ty = it.b;
  const int w = P.w, h = P.h;
  if (!in.use) {
    in.ref = P.in_ref;
    in.dis = P.in_dis;
    in.ref_pitch = P.in_ref_pitch;
    in.dis_pitch = P.in_dis_pitch;
    in.lin_ref_out = P.lin_ref_out;
    in.lin_dis_out = P.lin_dis_out;
    in.lin_ref_pitch = P.lin_ref_pitch;
    in.lin_dis_pitch = P.lin_dis_pitch;
    in.ref_c0 = P.in_ref_c0;
    in.ref_c1 = P.in_ref_c1;
    in.dis_c0 = P.in_dis_c0;
    in.dis_c1 = P.in_dis_c1;
    in.ref_cpitch = P.in_ref_cpitch;
    in.dis_cpitch = P.in_dis_cpitch;
    in.kind = P.in_kind;
  }

  // Scales 0 and 1 without a scale-0 tile: each thread takes the 2x2 scale-0
  // block behind its scale-1 pixel (clamped at the image edge, as the
  // downscale clamps), converts and stores the block, and forms the scale-1
  // mean from the same registers. sRGB u8 input: every input byte of the
  // block is requested before the tables are staged.
  const int sx = tid % T1, sy = tid / T1;
  const bool u8 = in.kind == kInputSrgbU8;
  unsigned short blo[NI][4];  // bytes 0, 1 as loaded
  unsigned char bhi[NI][4];   // byte 2
  if (u8) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int gx = min(tx * T0 + 2 * sx + (k & 1), w - 1), gy = min(ty * T0 + 2 * sy + (k >> 1), h - 1);
#pragma unroll
      for (int im = 0; im < NI; ++im) {
        const int img = SIDES == kSidesDis ? 1 : im;
        gin_u8* px = (gin_u8*)(img ? in.dis : in.ref) + gy * (img ? in.dis_pitch : in.ref_pitch) + gx * 3;
        __builtin_memcpy(&blo[im][k], (const void*)px, 2);
        bhi[im][k] = px[2];
      }
    }
  }

  // The sRGB table in LDS (the lookups are divergent, so constant memory
  // would serve them as one global load each) and the cube root's scale table.
  if (u8)
    for (int i = tid; i < 256; i += kThreads) lut_s[i] = c_lut[i];
  for (int i = tid; i < kCbrtScales; i += kThreads) cbrt_scale_s[i] = cbrt_scale(i + kCbrtExpMin);
  lds_barrier();
  lds_cfloat* lut = (lds_cfloat*)lut_s;
  const lds_cdouble* cbrt_scale_t = (const lds_cdouble*)cbrt_scale_s;

  // XYB of pixel `off` of a scale's planes (global stores: the row pass
  // reads these planes next, and measured 10 % slower after buffer stores).
  auto store_xyb = [&](int s, int img, bool inside, unsigned off, float X, float Y, float B) {
    if (!inside) return;
    const ScaleDesc& S = P.sc[s];
    float* base = xyb_plane(P, s, img, 0, SIDES);
    gbl(base)[off] = X;
    gbl(base + S.plane)[off] = Y;
    gbl(base + 2 * S.plane)[off] = B;
  };

  float sum1[NI][3];
#pragma unroll
  for (int im = 0; im < NI; ++im)
#pragma unroll
    for (int c = 0; c < 3; ++c) sum1[im][c] = 0.0f;
  // Block pixel k = (x0 + (k & 1), y0 + (k >> 1)): the summation order of
  // cpu.rs:545-579 is (x0,y0), (x1,y0), (x0,y1), (x1,y1).
  const int x0 = tx * T0 + 2 * sx;
#pragma unroll
  for (int row = 0; row < 2; ++row) {
    const int uy = ty * T0 + 2 * sy + row;
    const int gy = min(uy, h - 1);
    const unsigned off = (unsigned)(gy * P.sc[0].pitch + x0);
#pragma unroll
    for (int im = 0; im < NI; ++im) {
      const int img = SIDES == kSidesDis ? 1 : im;
#pragma unroll
      for (int kx = 0; kx < 2; ++kx) {
        const int k = 2 * row + kx;
        const int ux = x0 + kx;
        const int gx = min(ux, w - 1);
        float rgb[3];
        if (u8) {
          const unsigned lo = blo[im][k];
          rgb[0] = lut[lo & 255u];
          rgb[1] = lut[lo >> 8];
          rgb[2] = lut[bhi[im][k]];
        } else {
          read_linear(input_of(in, img), gx, gy, lut, rgb);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) sum1[im][c] += rgb[c];
        if (in.kind != kInputLinearF32 && ux < w && uy < h) {
          float* lo = img == 0 ? in.lin_ref_out : in.lin_dis_out;
          if (lo) {
            const long long lp = img == 0 ? in.lin_ref_pitch : in.lin_dis_pitch;
            gfloat* q = gbl(reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(lo) + gy * lp)) + gx * 3;
            q[0] = rgb[0];
            q[1] = rgb[1];
            q[2] = rgb[2];
          }
        }
        float X, Y, B;
        linear_to_xyb(rgb[0], rgb[1], rgb[2], X, Y, B, cbrt_scale_t);
        store_xyb(0, img, ux < w && uy < h, off + kx, X, Y, B);
      }
    }
  }
  if (P.nscales < 2) return;
  {
    const int gx = tx * T1 + sx, gy = ty * T1 + sy;
    const bool inside = gx < P.sc[1].w && gy < P.sc[1].h;
    const unsigned off = (unsigned)(gy * P.sc[1].pitch + gx);
#pragma unroll
    for (int im = 0; im < NI; ++im) {
      const int img = SIDES == kSidesDis ? 1 : im;
      float rgb[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        rgb[c] = sum1[im][c] * 0.25f;
        lin1[im * 3 + c][sy][sx] = rgb[c];
      }
      float X, Y, B;
      linear_to_xyb(rgb[0], rgb[1], rgb[2], X, Y, B, cbrt_scale_t);
      store_xyb(1, img, inside, off, X, Y, B);
    }
  }
  if (P.nscales < 3) return;

  // Scales 2..5 (8x8, 4x4, 2x2 and 1x1 pixels of the tile): the means level
  // by level with a barrier each, then every pixel's XYB at once. Threads
  // 0 .. 64 NI - 1 take a scale-2 pixel of one image; the next 21 NI threads
  // a scale-3, -4 or -5 pixel (16 + 4 + 1 per image).
  lds_barrier();
  const int ox1 = tx * T1, oy1 = ty * T1, ox2 = tx * T2, oy2 = ty * T2, ox3 = tx * T3, oy3 = ty * T3;
  const bool s2_thread = tid < 64 * NI;
  const int t = tid - 64 * NI;
  const bool tail_thread = t >= 0 && t < 21 * NI;
  const int im = s2_thread ? tid / 64 : (tail_thread ? t / 21 : 0);
  const int j = tail_thread ? t % 21 : 0;
  const int s = s2_thread ? 2 : (j < 16 ? 3 : (j < 20 ? 4 : 5));
  const int T = T0 >> s;
  const int li = s2_thread ? tid % 64 : (j < 16 ? j : (j < 20 ? j - 16 : 0));
  const int gx = tx * T + li % T, gy = ty * T + li / T;
  // a pixel of a scale the image has (the scale descriptors past the last
  // one are not read)
  const bool live = (s2_thread || tail_thread) && s < P.nscales;
  float rgb[3] = {0.0f, 0.0f, 0.0f};
  if (s2_thread) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rgb[c] = mean2x2(gx, gy, P.sc[1].w, P.sc[1].h,
                       [&](int x, int y) { return lin1[im * 3 + c][y - oy1][x - ox1]; });
      lin2[im * 3 + c][gy - oy2][gx - ox2] = rgb[c];
    }
  }
  lds_barrier();
  if (live && s == 3) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rgb[c] = mean2x2(gx, gy, P.sc[2].w, P.sc[2].h,
                       [&](int x, int y) { return lin2[im * 3 + c][y - oy2][x - ox2]; });
      lin3[im * 3 + c][gy - oy3][gx - ox3] = rgb[c];
    }
  }
  lds_barrier();
  if (live && s >= 4) {
    const int w3 = P.sc[3].w, h3 = P.sc[3].h, w4 = P.sc[4].w, h4 = P.sc[4].h;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      auto at3 = [&](int x, int y) { return lin3[im * 3 + c][y - oy3][x - ox3]; };
      if (s == 4)
        rgb[c] = mean2x2(gx, gy, w3, h3, at3);
      else
        rgb[c] = mean2x2(gx, gy, w4, h4, [&](int x, int y) { return mean2x2(x, y, w3, h3, at3); });
    }
  }
  if (live) {
    const int img = SIDES == kSidesDis ? 1 : im;
    float X, Y, B;
    linear_to_xyb(rgb[0], rgb[1], rgb[2], X, Y, B, cbrt_scale_t);
    store_xyb(s, img, gx < P.sc[s].w && gy < P.sc[s].h, (unsigned)(gy * P.sc[s].pitch + gx), X, Y, B);
  }
}

// ---------------------------------------------------------------- row pass

// One wave per (scale, channel, quantity, band of 20 rows). The scan along a
// row is sequential (cpu.rs:967-1022, row form), so its time is the number of
// instructions a wave issues per column: each row is spread over three lanes,
// one per cosine term, which issue one recurrence each per column; the terms
// are summed in the store phase. That leaves more waves than SIMDs at 1080p.
//
// Memory is staged through LDS in chunks of 20 rows x 32 columns so that every
// global access is coalesced (8 rows x 128 B per instruction, instead of one
// row per lane): buffer loads with scalar column offsets fill a padded LDS
// tile, each lane reads its row back, and the outputs leave the same way. The
// next chunk's loads are in flight while the current one is scanned.
//
// Every plane has at least 8 zero columns of padding (pitch >= w + 8), which
// supply the zero inputs right of the row; the left ones start as zeros in
// the ring. The buffer range check turns reads of rows past the image into 0
// and drops stores outside the band's rows.
constexpr int kChunk = 32;                      // columns per staged chunk (8 groups of 4)
constexpr int kRowPrefetch = 2;                 // chunks of row-pass input in flight per wave
// LDS tiles: rows of 8 float4 (one chunk) with segment g of row r in slot
// g ^ (r & 7), and the three tiles 3 float4 apart beyond their rows (the
// first of those is the slot where lanes past a chunk park). LDS serves 16
// lanes x 16 B a cycle; with this layout the staging and store pattern (two
// rows of 8 segments per 16 lanes), the scan's reads (five rows, one
// segment) and most of the term writes fall on distinct banks.
constexpr int kTileRowF4 = kChunk / 4;
constexpr int kTileF4 = kRowsPerWave * kTileRowF4 + 3;
__device__ __forceinline__ int tile_slot(int r, int g) { return r * kTileRowF4 + (g ^ (r & 7)); }

// The row pass exchanges data between the lanes of one wave through LDS. A
// wave's LDS operations execute in issue order, so only the compiler's
// reordering has to be stopped between the phases.
__device__ __forceinline__ void wave_lds_order() { asm volatile("" ::: "memory"); }

template <int Q>
__device__ __forceinline__ float4 row_quantity(float4 x, float4 y) {
  if (Q == 0) return make_float4(x.x * x.x, x.y * x.y, x.z * x.z, x.w * x.w);
  if (Q == 1) return make_float4(y.x * y.x, y.y * y.y, y.z * y.z, y.w * y.w);
  if (Q == 2) return make_float4(x.x * y.x, x.y * y.y, x.z * y.z, x.w * y.w);
  if (Q == 3) return x;
  return y;
}

// Step n (cpu.rs runs n = -4 .. w-1) adds x[n-6] and x[n+4]. Output group g
// covers n = 4g .. 4g+3: its right inputs are x-group g+1 (columns 4g+4..4g+7)
// and its left inputs the upper half of x-group g-2 and the lower half of
// x-group g-1. Chunk k computes output groups 8k..8k+7 (columns 32k..32k+31)
// from x-groups 8k+1..8k+8, i.e. columns 32k+4..32k+35, staged in LDS; the
// ring keeps the quantities of the last four x-groups (ring[g & 3] = group g).
//
// Lanes: each 16-lane row of the wave holds five rows of the band as lane
// triples (lane 15 idle); lane 3t+k runs cosine term k of its row, so a lane
// issues one recurrence (4 instructions) per column. The quantity (a square,
// a product or a plane itself) is formed once per element in the coalesced
// layout and staged in one tile. Each lane writes its term's outputs to the
// LDS tile of that term; the coalesced store phase reads the three tiles and
// forms (o1 + o3) + o5 (cpu.rs's order) once per output, one chunk later, so
// the LDS round trip of the outputs overlaps the next chunk's scan.
//
// The buffer range check applies to the per-lane offset only, so rows past
// the image carry an out-of-range per-lane offset and the column is the
// scalar one.
typedef __attribute__((address_space(3))) vf4 lds_vf4;
__device__ __forceinline__ vf4 to_vf4(float4 v) { return vf4{v.x, v.y, v.z, v.w}; }

template <int Q>
__device__ __forceinline__ void row_scan(__amdgpu_buffer_rsrc_t ra, __amdgpu_buffer_rsrc_t rb,
                                         __amdgpu_buffer_rsrc_t ro, int w, int pitch, int nrows,
                                         lds_vf4* tile_x, lds_vf4* tile_o) {
  constexpr bool NA = Q != 1 && Q != 4;
  constexpr bool NB = Q == 1 || Q == 2 || Q == 4;
  constexpr int NE = kRowsPerWave * 8;       // float4 of one chunk: rows x 8 segments
  constexpr int NI = (NE + 63) / 64;         // coalesced instructions per chunk and plane
  const int lane = threadIdx.x;
  const int dl = lane & 15, tr = min(dl / 3, 4), k = dl - 3 * (dl / 3);
  const int r = 5 * (lane >> 4) + tr;        // this lane's row in the band
  const bool active = dl < 15;               // lane 15 of each 16 duplicates lane 12
  lds_vf4* tile_k = tile_o + k * kTileF4;    // this lane's term outputs
  lds_vf4* tile_o1 = tile_o;
  lds_vf4* tile_o3 = tile_o + kTileF4;
  lds_vf4* tile_o5 = tile_o + 2 * kTileF4;
  const float mi = k == 0 ? kMulIn1 : (k == 1 ? kMulIn3 : kMulIn5);
  const float mp = k == 0 ? kMulPrev1 : (k == 1 ? kMulPrev3 : kMulPrev5);
  const u32x4 z = {0u, 0u, 0u, 0u};
  const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
  // coalesced layout: element e = 64 i + lane is row e / 8, segment e % 8;
  // lanes past the chunk use the padding slot of tile row 0 and move nothing
  unsigned voff[NI];
  int eidx[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int e = 64 * i + lane, er = e >> 3, es = e & 7;
    eidx[i] = e < NE ? tile_slot(er, es) : kRowsPerWave * kTileRowF4;
    voff[i] = e < NE && er < nrows ? (unsigned)(er * pitch * 4 + es * 16) : kOffNone;
  }
  int rslot[8];  // this lane's row, segment j
#pragma unroll
  for (int j = 0; j < 8; ++j) rslot[j] = tile_slot(r, j);
  const int glast = (w - 1) / 4;  // last output group
  const int nchunks = glast / 8 + 1;
  float4 ring[4];
  {
    // x-group 0 (loaded before the chunks: it is needed first), then the
    // steps n = -4 .. -1 (no outputs): right inputs x-group 0, left inputs
    // the zeros of x-groups -3 and -2.
    const unsigned vrow = r < nrows ? (unsigned)(r * pitch * 4) : kOffNone;
    const float4 a = NA ? u2f(__builtin_amdgcn_raw_buffer_load_b128(ra, vrow, 0, 0)) : z4;
    const float4 b = NB ? u2f(__builtin_amdgcn_raw_buffer_load_b128(rb, vrow, 0, 0)) : z4;
    ring[0] = row_quantity<Q>(a, b);
    ring[1] = ring[2] = ring[3] = z4;
  }
  __builtin_amdgcn_sched_barrier(0);
  // kRowPrefetch chunks of input in flight: one wave per SIMD cannot hide an
  // HBM round trip behind one chunk of scan. Each buffer keeps its own
  // registers (the chunk loop is unrolled by kRowPrefetch), since copying a
  // register that a load is still filling would wait for the load.
  u32x4 qa[kRowPrefetch][NI], qb[kRowPrefetch][NI];
  auto load_chunk = [&](int kc, u32x4(&da)[NI], u32x4(&db)[NI]) {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      da[i] = NA ? __builtin_amdgcn_raw_buffer_load_b128(ra, voff[i], kc * 128 + 16, 0) : z;
      db[i] = NB ? __builtin_amdgcn_raw_buffer_load_b128(rb, voff[i], kc * 128 + 16, 0) : z;
    }
  };
  // Every load and chunk is unconditional, so the loads' wait counts stay
  // exact: the chunk count is padded to a multiple of kRowPrefetch, loads
  // past the last chunk re-read it, and the padded chunks' stores carry an
  // out-of-range offset (8 kc > glast), so they write nothing.
  const int nchunks_pad = (nchunks + kRowPrefetch - 1) / kRowPrefetch * kRowPrefetch;
  // in chunk order (the first chunk's data is needed first)
#pragma unroll
  for (int d = 0; d < kRowPrefetch; ++d) {
    load_chunk(min(d, nchunks - 1), qa[d], qb[d]);
    __builtin_amdgcn_sched_barrier(0);
  }
  float p = 0.f, q = 0.f;
  // one step of this lane's term (cpu.rs:992-1010 for term k)
  auto step = [&](float left, float right) {
    const float sum = left + right;
    float o = sum * mi;
    o = fmaf(-1.0f, q, o);
    q = p;
    o = fmaf(mp, p, o);
    p = o;
    return o;
  };
  step(0.f, ring[0].x);
  step(0.f, ring[0].y);
  step(0.f, ring[0].z);
  step(0.f, ring[0].w);
  // stores of chunk kc from the term tiles (kc < 0: nothing)
  auto store_chunk = [&](int kc) {
    float4 v[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const float4 o1 = to_f4(tile_o1[eidx[i]]), o3 = to_f4(tile_o3[eidx[i]]), o5 = to_f4(tile_o5[eidx[i]]);
      v[i] = make_float4((o1.x + o3.x) + o5.x, (o1.y + o3.y) + o5.y, (o1.z + o3.z) + o5.z, (o1.w + o3.w) + o5.w);
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const bool seg_ok = kc >= 0 && 8 * kc + ((64 * i + lane) & 7) <= glast;
      const u32x4 u = {__float_as_uint(v[i].x), __float_as_uint(v[i].y), __float_as_uint(v[i].z),
                       __float_as_uint(v[i].w)};
      // Column offset in the per-lane offset and an immediate scalar offset of
      // 0: a 16-byte buffer store reads its data registers after issue, and
      // the compiler only guards them against the next instructions (wait
      // states) when the scalar offset is not a register. The range check
      // still drops kOffNone + kc*128 (< 2^32).
      __builtin_amdgcn_raw_buffer_store_b128(u, ro, seg_ok ? voff[i] + (unsigned)(kc * 128) : kOffNone, 0, 0);
    }
  };
  auto chunk = [&](int kc, u32x4(&da)[NI], u32x4(&db)[NI]) {
    // This chunk's loads are done once at most the later chunks' loads are
    // outstanding: loads complete in issue order, but stores (some of them
    // wholly out of range) may complete before earlier loads, so the wait
    // must not count on the stores issued in between.
    __builtin_amdgcn_s_waitcnt(vmcnt_only<(kRowPrefetch - 1) * NI * (int(NA) + int(NB))>());
    // stage the quantity of x columns 32k+4 .. 32k+35 of the band's rows
#pragma unroll
    for (int i = 0; i < NI; ++i) tile_x[eidx[i]] = to_vf4(row_quantity<Q>(u2f(da[i]), u2f(db[i])));
    wave_lds_order();
    load_chunk(min(kc + kRowPrefetch, nchunks - 1), da, db);
    // the previous chunk's outputs leave while this chunk's inputs come in
    store_chunk(kc - 1);
    float4 x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = to_f4(tile_x[rslot[j]]);
    __builtin_amdgcn_sched_barrier(0);
    float4 out[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      ring[(j + 1) & 3] = x[j];  // x-group 8k+j+1
      const float4 rr = ring[(j + 1) & 3], m2 = ring[(j + 2) & 3], m1 = ring[(j + 3) & 3];
      out[j].x = step(m2.z, rr.x);
      out[j].y = step(m2.w, rr.y);
      out[j].z = step(m1.x, rr.z);
      out[j].w = step(m1.y, rr.w);
    }
    wave_lds_order();
    if (active) {
#pragma unroll
      for (int j = 0; j < 8; ++j) tile_k[rslot[j]] = to_vf4(out[j]);
    }
    wave_lds_order();
  };
  for (int kc = 0; kc < nchunks_pad; kc += kRowPrefetch) {
#pragma unroll
    for (int d = 0; d < kRowPrefetch; ++d) chunk(kc + d, qa[d], qb[d]);
  }
  store_chunk(nchunks_pad - 1);
}

template <int SIDES>
__global__ __launch_bounds__(64) void k_rows(const PairDesc* __restrict__ pairs,
                                             const WorkItem* __restrict__ items) {
  // the staged quantity and the three terms' outputs
  __shared__ float4 tiles[4 * kTileF4];
  lds_vf4* tile_x = (lds_vf4*)tiles;
  lds_vf4* tile_o = (lds_vf4*)(tiles + kTileF4);
  const WorkItem it = items[blockIdx.x];
  const PairDesc& P = pairs[it.pair];
  const int s = it.a, c = it.b / kNumQ, q = it.b % kNumQ;
  const ScaleDesc& S = P.sc[s];
  const int row0 = it.c * kRowsPerWave;
  const long long off = (long long)row0 * S.pitch;
  const int nrows = min(kRowsPerWave, S.h - row0);
  const long long bytes = (long long)nrows * S.pitch * 4;
  const __amdgpu_buffer_rsrc_t ra = buffer_rsrc(xyb_plane(P, s, 0, c, SIDES) + off, bytes);
  const __amdgpu_buffer_rsrc_t rb = buffer_rsrc(xyb_plane(P, s, 1, c, SIDES) + off, bytes);
  const __amdgpu_buffer_rsrc_t ro =
      buffer_rsrc(P.arena + S.hrow + (long long)(c * kNumQ + q) * S.plane + off, bytes);
  switch (q) {
    case 0: row_scan<0>(ra, rb, ro, S.w, S.pitch, nrows, tile_x, tile_o); break;
    case 1: row_scan<1>(ra, rb, ro, S.w, S.pitch, nrows, tile_x, tile_o); break;
    case 2: row_scan<2>(ra, rb, ro, S.w, S.pitch, nrows, tile_x, tile_o); break;
    case 3: row_scan<3>(ra, rb, ro, S.w, S.pitch, nrows, tile_x, tile_o); break;
    default: row_scan<4>(ra, rb, ro, S.w, S.pitch, nrows, tile_x, tile_o); break;
  }
}

// ---------------------------------------------------------------- column pass + maps

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Waves of a column-pass block: scan waves (two quantities each), then map waves.
template <int SIDES>
struct ColWaves {
  static constexpr int scan = (SideTraits<SIDES>::NQ + 1) / 2;
  static constexpr int map = 2 * scan;  // 1x: 58.8 us at 1080p, 2x: 53.7, 3x: 52.3
  static constexpr int threads = 64 * (scan + map);
};

// One block per (scale, channel, strip of 32 columns) with two kinds of waves.
// Scan waves (ceil(NQ / 2)): each scans two of the blurred quantities (lanes
// 0-31 one, lanes 32-63 the next); lane x of quantity slot j scans column x
// with the three cosine terms of cpu.rs:1054-1115 (column form) in registers
// and forms (o1 + o3) + o5 itself. Rows go in chunks of kColChunk: the bottom
// inputs (row n+4 for step n) of the next chunk are loaded while the current
// one is scanned, and the top inputs (row n-6) are earlier bottom inputs kept
// in registers. Loads are buffer loads with the row as a scalar offset,
// clamped to the 4 zero rows under every plane, so no lane computes a row
// address or a mask. Map waves (twice as many) evaluate the SSIM and
// edge-difference maps (cpu.rs:581-683) with f64 L1/L4 sums. The two run one
// chunk apart on a double-buffered LDS tile, separated by one barrier per
// phase: in phase k the scan waves produce chunk k while the map waves
// consume chunk k-1, so the maps are off the column scan's sequential path.
template <int SIDES>
__global__ __launch_bounds__(ColWaves<SIDES>::threads) void k_cols(
    const PairDesc* __restrict__ pairs, const WorkItem* __restrict__ items) {
  constexpr int NQ = SideTraits<SIDES>::NQ;
  constexpr int NW = ColWaves<SIDES>::scan, MW = ColWaves<SIDES>::map;
  constexpr int RY = kColChunk, CB = kColBlock, NT = 64 * MW;  // NT: map threads
  static_assert(CB == 32, "two quantities of 32 columns per wave");
  static_assert(RY >= 2 * kRadius, "top inputs come from the previous chunk");
  __shared__ float O[2][NQ][RY][CB];
  __shared__ double red[MW][kNumSums];
  const WorkItem it = items[blockIdx.x];
  const PairDesc& P = pairs[it.pair];
  const int s = it.a, c = it.b, xb = it.c * CB;
  const ScaleDesc& S = P.sc[s];
  const int w = S.w, h = S.h, pitch = S.pitch;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nch = (h + kRadius - 1 + RY - 1) / RY;  // chunks: rows -kRadius+1 .. h-1
  // The scale-0 strips are the longest scans and set the pass's length: their
  // waves issue first on a SIMD they share with a smaller scale's block.
  if (s == 0) __builtin_amdgcn_s_setprio(2);
  const int y00 = -kRadius + 1;

  if (wv < NW) {
    // ---------------------------------------------------------------- scan waves
    const int x = lane & 31;
    const int slot = min(2 * wv + (lane >> 5), NQ - 1);  // lanes past NQ rescan the last one, unused
    const int q0 = SideTraits<SIDES>::q(2 * wv), q = SideTraits<SIDES>::q(slot);
    // the wave's first plane as the buffer base; the second one by its
    // distance (at most 3 planes, < 2^32 bytes); rows are clamped, so no range
    // check is needed
    const __amdgpu_buffer_rsrc_t rq =
        __builtin_amdgcn_make_buffer_rsrc(P.arena + S.hrow + (long long)(c * kNumQ + q0) * S.plane, 0, -1, kRsrcWord3);
    const unsigned vcol = (unsigned)((long long)(q - q0) * S.plane * 4 + (xb + x) * 4);
    const int row_bytes = pitch * 4;
    auto load_chunk = [&](int y0, float* B) {
#pragma unroll
      for (int i = 0; i < RY; ++i) {
        const int r = min(y0 + i + kRadius - 1, h + kPadRows - 1);  // rows h .. h+3 are zero
        B[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rq, vcol, r * row_bytes, 0));
      }
    };
    float Bprev[RY], Bcur[RY], Bnext[RY];
#pragma unroll
    for (int i = 0; i < RY; ++i) Bprev[i] = 0.0f;
    load_chunk(y00, Bcur);
    float p1 = 0.f, p3 = 0.f, p5 = 0.f, q1 = 0.f, q3 = 0.f, q5 = 0.f;
    for (int k = 0; k <= nch; ++k) {
      if (k < nch) {
        const int y0 = y00 + k * RY;
        load_chunk(y0 + RY, Bnext);
        const int buf = k & 1;
#pragma unroll
        for (int i = 0; i < RY; ++i) {
          // top input row y0+i-6 = bottom input of step i-10
          const float top = i >= 2 * kRadius ? Bcur[i - 2 * kRadius] : Bprev[i + RY - 2 * kRadius];
          const float sum = top + Bcur[i];
          const float t1 = fmaf(p1, kVertMulPrev1, q1), t3 = fmaf(p3, kVertMulPrev3, q3),
                      t5 = fmaf(p5, kVertMulPrev5, q5);
          const float o1 = fmaf(sum, kMulIn1, -t1), o3 = fmaf(sum, kMulIn3, -t3), o5 = fmaf(sum, kMulIn5, -t5);
          q1 = p1;
          q3 = p3;
          q5 = p5;
          p1 = o1;
          p3 = o3;
          p5 = o5;
          O[buf][slot][i][x] = (o1 + o3) + o5;
        }
#pragma unroll
        for (int i = 0; i < RY; ++i) {
          Bprev[i] = Bcur[i];
          Bcur[i] = Bnext[i];
        }
      }
      lds_barrier();
    }
    if (SIDES != kSidesRef) lds_barrier();  // the map waves' reduction barrier
    return;
  }

  // ------------------------------------------------------------------ map waves
  const int mt = tid - 64 * NW;  // thread index among the map waves
  gcfloat* srcp = gbl(xyb_plane(P, s, 0, c, SIDES));
  gcfloat* disp = gbl(xyb_plane(P, s, 1, c, SIDES));
  gfloat* cache_s11 = gbl(P.arena + S.cache + (long long)(3 + c) * S.plane);
  gfloat* cache_mu1 = gbl(P.arena + S.cache + (long long)(6 + c) * S.plane);
  // Map inputs (source and distorted XYB, and the cached reference statistics
  // when only the distorted side runs) of this thread's pixels, loaded one
  // chunk ahead: the maps would otherwise wait out an HBM latency per chunk.
  constexpr int NM = (RY * CB + NT - 1) / NT;  // map pixels per thread per chunk
  constexpr int NV = SIDES == kSidesDis ? 4 : (SIDES == kSidesBoth ? 2 : 0);
  static_assert(NT % CB == 0, "a map thread keeps its column in every chunk");
  float mv[NM][NV > 0 ? NV : 1], mv_next[NM][NV > 0 ? NV : 1];
  // buffer loads: the planes as resources, one 32-bit offset per pixel
  // (row clamped to the image, column clamped once per thread)
  const long long plane_bytes = S.plane * 4;
  const __amdgpu_buffer_rsrc_t r_src = buffer_rsrc((const float*)srcp, plane_bytes);
  const __amdgpu_buffer_rsrc_t r_dis = buffer_rsrc((const float*)disp, plane_bytes);
  const __amdgpu_buffer_rsrc_t r_s11 = buffer_rsrc((const float*)cache_s11, plane_bytes);
  const __amdgpu_buffer_rsrc_t r_mu1 = buffer_rsrc((const float*)cache_mu1, plane_bytes);
  const int mrow = mt / CB, mcol4 = min(xb + mt % CB, w - 1) * 4, pitch4 = pitch * 4;
  auto load_maps = [&](int y0c, float (*dst)[NV > 0 ? NV : 1]) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const int n = min(max(y0c + mrow + (NT / CB) * m, 0), h - 1);
      const unsigned off = (unsigned)(n * pitch4 + mcol4);
      if (NV >= 2) {
        dst[m][0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r_src, off, 0, 0));
        dst[m][1] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r_dis, off, 0, 0));
      }
      if (NV == 4) {
        dst[m][2] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r_s11, off, 0, 0));
        dst[m][3] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r_mu1, off, 0, 0));
      }
    }
  };
  load_maps(y00, mv);
  double acc[kNumSums] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k <= nch; ++k) {
    if (k >= 1) {
      // chunk k-1, produced in the previous phase
      const int y0 = y00 + (k - 1) * RY, buf = (k - 1) & 1;
      load_maps(y0 + RY, mv_next);
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const int px = mt + NT * m;
        const int i = px / CB, col = xb + px % CB, n = y0 + i;
        if ((RY * CB) % NT == 0 || px < RY * CB) {
          if (n >= 0 && n < h && col < w) {
            float bq[NQ];
#pragma unroll
            for (int q2 = 0; q2 < NQ; ++q2) bq[q2] = O[buf][q2][i][px % CB];
            const long long off = (long long)n * pitch + col;
            if (SIDES == kSidesRef) {
              cache_s11[off] = bq[0];
              cache_mu1[off] = bq[1];
            } else {
              const float* v = mv[m];
              float s11, s22, s12, mu1, mu2;
              if (SIDES == kSidesBoth) {
                s11 = bq[0];
                s22 = bq[1];
                s12 = bq[2];
                mu1 = bq[3];
                mu2 = bq[4];
              } else {
                s11 = v[2];
                mu1 = v[3];
                s22 = bq[0];
                s12 = bq[1];
                mu2 = bq[2];
              }
              // per-pixel terms as cpu.rs forms them; the fourth powers
              // are accumulated with one rounding (the sums' order differs
              // from cpu.rs's anyway)
              const double d = ssim_error(mu1, mu2, s11, s22, s12);
              const double d2 = d * d;
              acc[0] += d;
              acc[1] = fma(d2, d2, acc[1]);
              const double d1 = edge_ratio(v[0], mu1, v[1], mu2);
              const double art = d1 > 0.0 ? d1 : 0.0;
              const double det = -d1 > 0.0 ? -d1 : 0.0;
              const double a2 = art * art, e2 = det * det;
              acc[2] += art;
              acc[3] = fma(a2, a2, acc[3]);
              acc[4] += det;
              acc[5] = fma(e2, e2, acc[5]);
            }
          }
        }
      }
#pragma unroll
      for (int m = 0; m < NM; ++m)
#pragma unroll
        for (int q2 = 0; q2 < (NV > 0 ? NV : 1); ++q2) mv[m][q2] = mv_next[m][q2];
    }
    lds_barrier();
  }
  if (SIDES == kSidesRef) return;
  // Map-wave reduction in a fixed order: wave shuffles, then the map waves in
  // sequence (the scan waves wait at the same barrier).
#pragma unroll
  for (int k = 0; k < kNumSums; ++k) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) red[wv - NW][k] = v;
  }
  lds_barrier();
  if (mt < kNumSums) {
    double v = 0.0;
    for (int k = 0; k < MW; ++k) v += red[k][mt];
    P.partials[(long long)(S.part + c * S.colblocks + it.c) * kNumSums + mt] = v;
  }
}

// ---------------------------------------------------------------- utilities

// crates/ssimulacra2-cuda-kernel/src/srgb.rs:58-70 (Kernel::srgb_to_linear):
// packed u8 RGB -> packed f32 linear RGB through the 256-entry table.
__global__ __launch_bounds__(kThreads) void k_srgb_to_linear(const unsigned char* __restrict__ in,
                                                             long long in_pitch, float* __restrict__ out,
                                                             long long out_pitch, int w3, int h) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= w3 || y >= h) return;
  float* o = reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(out) + y * out_pitch);
  o[x] = c_lut[in[y * in_pitch + x]];
}

// Plain streaming copy, 16 bytes per lane per iteration: the known-size
// reference kernel that hardware byte counters are calibrated against.
__global__ __launch_bounds__(kThreads) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                   long long n16) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// crates/cuda-colorspace-kernel/src/biplanar.rs:8-70 (biplanaryuv420_to_linearrgb):
// YUV 4:2:0 -> packed f32 linear RGB, one thread per pixel.
__global__ __launch_bounds__(kThreads) void k_yuv_to_linear(InputView in, float* __restrict__ out,
                                                            long long out_pitch, int w, int h) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= w || y >= h) return;
  float rgb[3];
  read_linear(in, x, y, c_lut, rgb);
  gfloat* o = gbl(reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(out) + y * out_pitch)) + x * 3;
  o[0] = rgb[0];
  o[1] = rgb[1];
  o[2] = rgb[2];
}

// ---------------------------------------------------------------- launchers

#define SSIM2_DISPATCH_SIDES(KERNEL, GRID, BLOCK, STREAM, ...)                         \
  do {                                                                                 \
    if (sides == kSidesBoth)                                                           \
      hipLaunchKernelGGL(KERNEL<kSidesBoth>, GRID, BLOCK, 0, STREAM, __VA_ARGS__);     \
    else if (sides == kSidesRef)                                                       \
      hipLaunchKernelGGL(KERNEL<kSidesRef>, GRID, BLOCK, 0, STREAM, __VA_ARGS__);      \
    else                                                                               \
      hipLaunchKernelGGL(KERNEL<kSidesDis>, GRID, BLOCK, 0, STREAM, __VA_ARGS__);      \
  } while (0)

hipError_t launch_pyramid(const PairDesc* pairs, const WorkItem* items, int nitems, int sides,
                          const InputArgs& in, hipStream_t stream) {
  if (nitems <= 0) return hipSuccess;
  SSIM2_DISPATCH_SIDES(k_pyramid, dim3(nitems), dim3(kThreads), stream, pairs, items, in);
  return hipGetLastError();
}

hipError_t launch_rows(const PairDesc* pairs, const WorkItem* items, int nitems, int sides,
                       hipStream_t stream) {
  if (nitems <= 0) return hipSuccess;
  SSIM2_DISPATCH_SIDES(k_rows, dim3(nitems), dim3(64), stream, pairs, items);
  return hipGetLastError();
}

hipError_t launch_cols(const PairDesc* pairs, const WorkItem* items, int nitems, int sides,
                       hipStream_t stream) {
  if (nitems <= 0) return hipSuccess;
  if (sides == kSidesBoth)
    hipLaunchKernelGGL(k_cols<kSidesBoth>, dim3(nitems), dim3(ColWaves<kSidesBoth>::threads), 0, stream, pairs,
                       items);
  else if (sides == kSidesRef)
    hipLaunchKernelGGL(k_cols<kSidesRef>, dim3(nitems), dim3(ColWaves<kSidesRef>::threads), 0, stream, pairs,
                       items);
  else
    hipLaunchKernelGGL(k_cols<kSidesDis>, dim3(nitems), dim3(ColWaves<kSidesDis>::threads), 0, stream, pairs,
                       items);
  return hipGetLastError();
}

hipError_t launch_finalize(const PairDesc* pairs, int npairs, hipStream_t stream) {
  if (npairs <= 0) return hipSuccess;
  return launch_finalize_pairs(pairs, nullptr, npairs, npairs * kMaxScales * 3, npairs, stream);
}

// Partners of the swept operand in the division checks (both signs, zero,
// subnormal, tiny, the workloads' magnitudes and large values).
__constant__ float c_check_partners[64] = {
    0.0f,      0x1p-149f,  0x1p-126f,   0x1p-110f,   0x1p-100f,   0x1p-90f,    0x1p-60f,   1e-20f,
    1e-10f,    1e-7f,      1e-5f,       1e-4f,       0.0009f,     0.001f,      0.0123f,    0.03125f,
    0.1f,      0.1234567f, 0.25f,       0.3333333f,  0.5f,        0.70710677f, 0.9f,       0.99999994f,
    1.0f,      1.0000001f, 1.2599211f,  1.5f,        1.9999999f,  2.0f,        3.0f,       3.1415927f,
    4.0f,      7.389056f,  10.0f,       100.0f,      1e4f,        1e8f,        1e20f,      0x1p100f,
    -0x1p-149f, -1e-20f,   -1e-7f,      -0.0009f,    -0.001f,     -0.0123f,    -0.1f,      -0.25f,
    -0.5f,     -0.9f,      -1.0f,       -1.5f,       -2.0f,       -3.0f,       -10.0f,     -100.0f,
    0.0008999f, 0.00090001f, 0.6180340f, 0.8414710f, 1.7320508f,  2.7182817f,  5.0f,       0.4999999f};

// Diagnostics of the reduced arithmetic against the IEEE forms, on the
// device, for every float x whose bits lie in [lo, hi]; differing results are
// counted in *mismatches. which 0: the kernels' cube root of x >= 0 (scale
// table, reciprocal division) against the reference form with the IEEE
// division, which equals the platform cbrtf on [0, 2]. which 1: the SSIM
// map's 1 - q for q = a / x in f32, a over the partners (x is the
// denominator, >= C2 minus rounding in the maps). which 2: the edge ratio with
// |dis - mu2| = x and |src - mu1| over the partners' magnitudes.
__global__ __launch_bounds__(kThreads) void k_check_arithmetic(int which, unsigned lo, unsigned hi,
                                                               unsigned long long* __restrict__ mismatches) {
  __shared__ double factor_s[5];
  __shared__ double scale_s[kCbrtScales];
  if (threadIdx.x < 5) factor_s[threadIdx.x] = c_cbrt_factor[threadIdx.x];
  for (int i = threadIdx.x; i < kCbrtScales; i += kThreads) scale_s[i] = cbrt_scale(i + kCbrtExpMin);
  lds_barrier();
  const lds_cdouble* factor = (const lds_cdouble*)factor_s;
  const lds_cdouble* scale = (const lds_cdouble*)scale_s;
  unsigned long long bad = 0;
  const unsigned long long n = (unsigned long long)hi - lo + 1;
  for (unsigned long long i = (unsigned long long)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * kThreads) {
    const float x = __uint_as_float(lo + (unsigned)i);
    if (which == 0) {
      bad += __float_as_uint(cbrtf_nonneg<false>(x, scale)) != __float_as_uint(cbrtf_ref<true>(x, factor));
    } else if (which == 1) {
      for (int k = 0; k < 64; ++k) {
        const float a = c_check_partners[k];
        const double f = 1.0 - (double)div_rn_f32<false>(a, x), g = 1.0 - (double)div_rn_f32<true>(a, x);
        bad += __double_as_longlong(f) != __double_as_longlong(g);
      }
    } else {
      for (int k = 0; k < 64; ++k) {
        const float b = fabsf(c_check_partners[k]);
        const double f = edge_ratio<false>(b, 0.0f, x, 0.0f), g = edge_ratio<true>(b, 0.0f, x, 0.0f);
        bad += __double_as_longlong(f) != __double_as_longlong(g);
      }
    }
  }
  if (bad) atomicAdd(mismatches, bad);
}

hipError_t launch_check_arithmetic(int which, unsigned lo, unsigned hi, unsigned long long* mismatches,
                                   hipStream_t stream) {
  hipLaunchKernelGGL(k_check_arithmetic, dim3(8192), dim3(kThreads), 0, stream, which, lo, hi, mismatches);
  return hipGetLastError();
}

// Kernel::downscale_by_2 (crates/ssimulacra2-cuda-kernel/src/downscale.rs:4-35,
// cpu.rs:545-579): one thread per output pixel, the 2x2 mean of packed f32
// RGB with the source coordinates clamped to the last row and column, summed
// in the order (x0,y0), (x1,y0), (x0,y1), (x1,y1).
__global__ __launch_bounds__(kThreads) void k_downscale_by_2(const float* __restrict__ src, long long src_pitch,
                                                             int sw, int sh, float* __restrict__ dst,
                                                             long long dst_pitch, int dw, int dh) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= dw || y >= dh) return;
  const int x0 = min(2 * x, sw - 1), x1 = min(2 * x + 1, sw - 1);
  const int y0 = min(2 * y, sh - 1), y1 = min(2 * y + 1, sh - 1);
  gcfloat* r0 = gbl(const_cast<float*>(reinterpret_cast<const float*>(reinterpret_cast<const unsigned char*>(src) + y0 * src_pitch)));
  gcfloat* r1 = gbl(const_cast<float*>(reinterpret_cast<const float*>(reinterpret_cast<const unsigned char*>(src) + y1 * src_pitch)));
  gfloat* o = gbl(reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(dst) + y * dst_pitch)) + 3 * x;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float sum = 0.0f;
    sum += r0[3 * x0 + c];
    sum += r0[3 * x1 + c];
    sum += r1[3 * x0 + c];
    sum += r1[3 * x1 + c];
    o[c] = sum * 0.25f;
  }
}

// Kernel::linear_to_xyb (crates/ssimulacra2-cuda-kernel/src/xyb.rs:81-102,
// cpu.rs:458-496): packed f32 linear RGB to packed positive XYB, one thread
// per pixel, the pyramid's arithmetic (cube root with its scale table in LDS).
__global__ __launch_bounds__(kThreads) void k_linear_to_xyb(const float* __restrict__ src, long long src_pitch,
                                                            float* __restrict__ dst, long long dst_pitch, int w,
                                                            int h) {
  __shared__ double scale_s[kCbrtScales];
  for (int i = threadIdx.x; i < kCbrtScales; i += kThreads) scale_s[i] = cbrt_scale(i + kCbrtExpMin);
  lds_barrier();
  const lds_cdouble* scale = (const lds_cdouble*)scale_s;
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  gcfloat* in = gbl(const_cast<float*>(reinterpret_cast<const float*>(reinterpret_cast<const unsigned char*>(src) + y * src_pitch))) + 3 * x;
  gfloat* o = gbl(reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(dst) + y * dst_pitch)) + 3 * x;
  float X, Y, B;
  linear_to_xyb(in[0], in[1], in[2], X, Y, B, scale);
  o[0] = X;
  o[1] = Y;
  o[2] = B;
}

hipError_t launch_downscale_by_2(const float* src, long long src_pitch, int sw, int sh, float* dst,
                                 long long dst_pitch, int dw, int dh, hipStream_t stream) {
  hipLaunchKernelGGL(k_downscale_by_2, dim3((dw + 63) / 64, (dh + 3) / 4), dim3(kThreads), 0, stream, src, src_pitch,
                     sw, sh, dst, dst_pitch, dw, dh);
  return hipGetLastError();
}

hipError_t launch_linear_to_xyb(const float* src, long long src_pitch, float* dst, long long dst_pitch, int w, int h,
                                hipStream_t stream) {
  hipLaunchKernelGGL(k_linear_to_xyb, dim3((w + 63) / 64, (h + 3) / 4), dim3(kThreads), 0, stream, src, src_pitch,
                     dst, dst_pitch, w, h);
  return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, long long bytes, hipStream_t stream) {
  const long long n16 = bytes / 16;
  if (n16 <= 0) return hipSuccess;
  const long long blocks = std::min<long long>((n16 + kThreads - 1) / kThreads, 8192);
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(kThreads), 0, stream,
                     static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
  return hipGetLastError();
}

hipError_t launch_yuv_to_linear(const InputView& in, float* out, long long out_pitch, int w, int h,
                                hipStream_t stream) {
  if (w <= 0 || h <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_yuv_to_linear, dim3((w + kThreads - 1) / kThreads, h), dim3(kThreads), 0, stream, in, out,
                     out_pitch, w, h);
  return hipGetLastError();
}

hipError_t launch_srgb_to_linear(const unsigned char* in, long long in_pitch, float* out,
                                 long long out_pitch, int w, int h, hipStream_t stream) {
  if (w <= 0 || h <= 0) return hipSuccess;
  const int w3 = w * 3;
  hipLaunchKernelGGL(k_srgb_to_linear, dim3((w3 + kThreads - 1) / kThreads, h), dim3(kThreads), 0,
                     stream, in, in_pitch, out, out_pitch, w3, h);
  return hipGetLastError();
}

}  // namespace ssim2
