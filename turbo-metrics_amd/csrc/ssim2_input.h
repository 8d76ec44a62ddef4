// Device-side readers of the scale-0 input, for every InputKind.
#pragma once

#include "ssim2_internal.h"
#include "ssim2_math.h"

namespace ssim2 {

typedef __attribute__((address_space(1))) const unsigned char gin_u8;
typedef __attribute__((address_space(1))) const unsigned short gin_u16;
typedef __attribute__((address_space(1))) const float gin_f32;

typedef __attribute__((address_space(3))) const float lds_cfloat;
typedef __attribute__((address_space(3))) const double lds_cdouble;

// Linear RGB of pixel (gx, gy). `lut` is the sRGB u8 table (constant memory or
// an LDS copy). Chroma of 4:2:0 inputs is taken from the sample of the pixel's
// 2x2 quad (nearest), as crates/cuda-colorspace-kernel/src/biplanar.rs:31-35.
template <typename Lut>
__device__ __forceinline__ void read_linear(const InputView& v, int gx, int gy, Lut lut, float* rgb) {
  switch (v.kind) {
    case kInputSrgbU8: {
      gin_u8* px = (gin_u8*)v.p + gy * v.pitch + gx * 3;
      rgb[0] = lut[px[0]];
      rgb[1] = lut[px[1]];
      rgb[2] = lut[px[2]];
      break;
    }
    case kInputLinearF32: {
      gin_f32* px = (gin_f32*)(v.p + gy * v.pitch) + gx * 3;
      rgb[0] = px[0];
      rgb[1] = px[1];
      rgb[2] = px[2];
      break;
    }
    case kInputYuvP10: {
      const unsigned y = ((gin_u16*)(v.p + gy * v.pitch))[gx];
      const unsigned u = ((gin_u16*)(v.c0 + (gy >> 1) * v.cpitch))[gx >> 1];
      const unsigned w = ((gin_u16*)(v.c1 + (gy >> 1) * v.cpitch))[gx >> 1];
      yuv16_to_linear(y << 6, u << 6, w << 6, rgb);
      break;
    }
    default: {  // kInputYuvP016
      const unsigned y = ((gin_u16*)(v.p + gy * v.pitch))[gx];
      gin_u16* uv = (gin_u16*)(v.c0 + (gy >> 1) * v.cpitch) + 2 * (gx >> 1);
      yuv16_to_linear(y, uv[0], uv[1], rgb);
      break;
    }
  }
}

__device__ __forceinline__ InputView input_of(const PairDesc& P, int img) {
  InputView v;
  v.p = img ? P.in_dis : P.in_ref;
  v.c0 = img ? P.in_dis_c0 : P.in_ref_c0;
  v.c1 = img ? P.in_dis_c1 : P.in_ref_c1;
  v.pitch = img ? P.in_dis_pitch : P.in_ref_pitch;
  v.cpitch = img ? P.in_dis_cpitch : P.in_ref_cpitch;
  v.kind = P.in_kind;
  return v;
}

__device__ __forceinline__ InputView input_of(const InputArgs& in, int img) {
  InputView v;
  v.p = img ? in.dis : in.ref;
  v.c0 = img ? in.dis_c0 : in.ref_c0;
  v.c1 = img ? in.dis_c1 : in.ref_c1;
  v.pitch = img ? in.dis_pitch : in.ref_pitch;
  v.cpitch = img ? in.dis_cpitch : in.ref_cpitch;
  v.kind = in.kind;
  return v;
}

}  // namespace ssim2
