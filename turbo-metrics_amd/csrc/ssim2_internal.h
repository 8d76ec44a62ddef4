// Internal device-side descriptors and kernel launchers of the ssimulacra2
// HIP engine. Not part of the public C ABI (include/ssim2hip.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace ssim2 {

constexpr int kMaxScales = 6;
// Quantities blurred per channel: ref^2, dis^2, ref*dis, ref, dis.
constexpr int kNumQ = 5;
// Per (scale, channel) error-map sums: ssim S1, ssim S4, artifact S1,
// artifact S4, detail-loss S1, detail-loss S4 (S4 = sum of d^4).
constexpr int kNumSums = 6;
constexpr int kSumsPerPair = kMaxScales * 3 * kNumSums;  // 108

// Tile / block geometry (compile-time; the launch builders use the same values).
constexpr int kPyrTile = 32;      // scale-0 tile edge of the pyramid kernel
constexpr int kRowsPerWave = 20;  // rows per row-pass wave (five lane triples per 16-lane row)
constexpr int kColBlock = 32;     // columns per column-pass block
constexpr int kColChunk = 24;     // rows per column-pass chunk
constexpr int kPadRows = 4;       // zero rows under every exact-path plane (blur radius - 1)
constexpr int kThreads = 256;

enum Storage : int {
  kStorageF32 = 0,
  kStorageF16 = 1,
};

// Output tile shapes of the fast form (k_fast_tile instantiations).
constexpr int kNumTileCfgs = 5;
constexpr int kTileCfgW[kNumTileCfgs] = {32, 64, 32, 32, 64};
constexpr int kTileCfgH[kNumTileCfgs] = {16, 16, 32, 24, 32};

enum InputKind : int {
  kInputSrgbU8 = 0,     // packed RGB u8, pitched
  kInputLinearF32 = 1,  // packed RGB f32 linear, pitched
  kInputYuvP10 = 2,     // planar 4:2:0, u16 10-bit LSB-aligned: Y, Cb (c0), Cr (c1)
  kInputYuvP016 = 3,    // biplanar 4:2:0, u16 MSB-aligned: Y, interleaved CbCr (c0)
};

// Which sides a launch processes.
enum Sides : int {
  kSidesBoth = 0,  // reference and distorted
  kSidesRef = 1,   // reference only (cached-reference preparation)
  kSidesDis = 2,   // distorted only, reference state read from the cache
};

struct ScaleDesc {
  int w, h;          // scale dimensions
  int pitch;         // plane row pitch in elements (multiple of 64)
  int colblocks;     // ceil(w / kColBlock)
  long long plane;   // elements per plane (pitch * h)
  long long xyb;     // element offset of the 6 XYB planes (ref X,Y,B, dis X,Y,B)
  long long hrow;    // element offset of the 15 row-pass planes (c*5 + q)
  long long cache;   // element offset of the 9 cached-reference planes: ref XYB c, blurred
                     // ref^2 (sigma11) c, blurred ref (mu1) c, for c = 0..2; fast form:
                     // byte offset of this scale's planes in PairDesc::ref_planes
  int part;          // index of the first partial-sum record of this scale
  int tiles_x;       // fast form: output tiles per row of tiles
  int tile0;         // fast form: index of this scale's first tile within the pair
  int pad;
};

struct PairDesc {
  const unsigned char* in_ref;  // device input, layout per in_kind
  const unsigned char* in_dis;
  long long in_ref_pitch;       // bytes
  long long in_dis_pitch;
  float* lin_ref_out;           // optional packed f32 linear copy of the scale-0 input
  float* lin_dis_out;
  long long lin_ref_pitch, lin_dis_pitch;
  const unsigned char* in_ref_c0;  // YUV chroma planes (see InputKind)
  const unsigned char* in_ref_c1;
  const unsigned char* in_dis_c0;
  const unsigned char* in_dis_c1;
  long long in_ref_cpitch, in_dis_cpitch;
  int in_kind;
  int nscales;
  int w, h;
  float* arena;                 // all planes of this pair
  unsigned char* ref_planes;    // fast form: a resident reference pyramid (its X,Y,B planes of
                                // scale s at byte sc[s].cache), read in place of image 0's planes
  int skip_img;                 // fast form: bit i set = the pyramid kernel leaves image i alone
  int pad_desc;
  double* partials;             // partial-sum records, kNumSums doubles each
  double* result;               // [0] score, [1..108] sums (scale*18 + c*6 + j), [109..120] dims
  ScaleDesc sc[kMaxScales];
};

// Per-launch input override for single-pair launches: the pyramid kernel reads
// the images named here instead of the descriptor's, so a context's device
// descriptor never changes between calls.
struct InputArgs {
  const unsigned char* ref;
  const unsigned char* dis;
  long long ref_pitch, dis_pitch;
  float* lin_ref_out;
  float* lin_dis_out;
  long long lin_ref_pitch, lin_dis_pitch;
  const unsigned char* ref_c0;
  const unsigned char* ref_c1;
  const unsigned char* dis_c0;
  const unsigned char* dis_c1;
  long long ref_cpitch, dis_cpitch;
  int kind;
  int use;  // 0: take inputs from the PairDesc
};

// Linear RGB of input pixel (gx, gy) of one image, for every InputKind.
struct InputView {
  const unsigned char* p;   // packed RGB, or the Y plane
  const unsigned char* c0;  // Cb plane, or interleaved CbCr
  const unsigned char* c1;  // Cr plane
  long long pitch, cpitch;
  int kind;
};

// One work item per block: {pair, a, b, c}; the meaning depends on the kernel.
// s_waitcnt immediate that waits for vmcnt <= N only (gfx9 encoding: vmcnt
// bits 3:0 and 15:14, expcnt 6:4 and lgkmcnt 11:8 left at their maxima).
// Loads complete in issue order, stores may complete before earlier loads:
// to know a load done, wait for N = the loads issued after it, whatever
// stores came in between.
template <int N>
constexpr int vmcnt_only() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  return (N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14);
}

struct WorkItem {
  int pair, a, b, c;
};

// Records the message returned by ssim2_last_error() on this thread.
void set_last_error(const std::string& msg);

// Launchers (kernels in ssim2_kernels.hip). `pairs` and `items` are device
// pointers; `nitems` blocks are launched.
hipError_t launch_pyramid(const PairDesc* pairs, const WorkItem* items, int nitems, int sides,
                          const InputArgs& in, hipStream_t stream);
hipError_t launch_rows(const PairDesc* pairs, const WorkItem* items, int nitems, int sides,
                       hipStream_t stream);
hipError_t launch_cols(const PairDesc* pairs, const WorkItem* items, int nitems, int sides,
                       hipStream_t stream);
hipError_t launch_finalize(const PairDesc* pairs, int npairs, hipStream_t stream);
// Fast form (ssim2_fast.hip). `prefix` holds npairs+1 first-item indices,
// `meta[0]` the pair count; every grid covers all items of a batch whatever
// its size (a block takes a contiguous range of items).
hipError_t launch_fast_pyramid(int storage, const PairDesc* pairs, const int* prefix, const int* meta, int grid,
                               hipStream_t stream);  // grid: best one block per tile
const void* fast_pyramid_kernel(int storage);     // the kernel launch_fast_pyramid runs
hipError_t launch_fast_tile(int storage, int tile_cfg, const PairDesc* pairs, const int* prefix,
                            const int* meta, int grid, hipStream_t stream);
const void* fast_tile_kernel(int storage, int tile_cfg);  // the kernel launch_fast_tile runs
hipError_t launch_fast_finalize(const PairDesc* pairs, const int* meta, int grid, int max_pairs,
                                hipStream_t stream);
// Shared finalisation (ssim2_finalize.hip): per-(pair, scale, channel)
// reduction then the score; `meta` (device, may be null) holds the pair count,
// else `npairs_static` is used.
hipError_t launch_finalize_pairs(const PairDesc* pairs, const int* meta, int npairs_static, int grid_units,
                                 int max_pairs, hipStream_t stream);
hipError_t launch_yuv_to_linear(const InputView& in, float* out, long long out_pitch, int w, int h,
                                hipStream_t stream);
hipError_t launch_copy(const void* src, void* dst, long long bytes, hipStream_t stream);
hipError_t launch_downscale_by_2(const float* src, long long src_pitch, int sw, int sh, float* dst,
                                 long long dst_pitch, int dw, int dh, hipStream_t stream);
hipError_t launch_linear_to_xyb(const float* src, long long src_pitch, float* dst, long long dst_pitch, int w, int h,
                                hipStream_t stream);
hipError_t launch_check_arithmetic(int which, unsigned lo, unsigned hi, unsigned long long* mismatches,
                                   hipStream_t stream);
hipError_t launch_srgb_to_linear(const unsigned char* in, long long in_pitch, float* out,
                                 long long out_pitch, int w, int h, hipStream_t stream);

}  // namespace ssim2
