// Scalar math shared by the ssimulacra2 HIP kernels (host+device inline).
//
// Every function here reproduces the arithmetic of the reference CPU path,
// crates/ssimulacra2-cuda/examples/cpu.rs, operation for operation, so that
// the GPU's f32 planes equal the CPU's bit for bit. The whole library is
// compiled with -ffp-contract=off: a multiply-add is fused only where an
// explicit fmaf/fma appears, which is exactly where cpu.rs calls mul_add.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define SSIM2_HD __host__ __device__ __forceinline__

namespace ssim2 {

// ---- opsin absorbance / XYB constants (cpu.rs:420-434, f32) ----
constexpr float kM02 = 0.078f;
constexpr float kM00 = 0.30f;
constexpr float kM01 = (1.0f - kM02) - kM00;
constexpr float kM12 = 0.078f;
constexpr float kM10 = 0.23f;
constexpr float kM11 = (1.0f - kM12) - kM10;
constexpr float kM20 = 0.24342269f;
constexpr float kM21 = 0.20476745f;
constexpr float kM22 = (1.0f - kM20) - kM21;
constexpr float kB0 = 0.0037930734f;
constexpr float kB0Root = 0.1559542025327239180319220163705f;

// ---- recursive Gaussian, sigma 1.5 (cpu.rs:931-948) ----
// Row form:    o = MUL_PREV*prev + (-prev2 + MUL_IN*sum)      (two fmas)
// Column form: o = VERT_MUL_IN*sum - (VERT_MUL_PREV*prev + prev2)
constexpr int kRadius = 5;
constexpr float kMulIn1 = 0x1.c4fa84p-5f, kMulIn3 = -0x1.e1fd7ap-5f, kMulIn5 = 0x1.a8894ap-7f;
constexpr float kMulPrev1 = 0x1.e6f0e2p+0f, kMulPrev3 = 0x1.2cf230p+0f, kMulPrev5 = 0x1.1a6264p-53f;
constexpr float kVertMulPrev1 = -0x1.e6f0e2p+0f, kVertMulPrev3 = -0x1.2cf230p+0f,
                kVertMulPrev5 = -0x1.1a6264p-53f;

constexpr float kC2 = 0.0009f;

// ---- cube root -----------------------------------------------------------
// The reference evaluates f32::cbrt, i.e. the platform libm cbrtf (glibc on
// Linux). Its result differs from the correctly rounded cube root in about
// one input in ten, and the score is sensitive to that last bit through the
// marginally stable IIR. This restates glibc's float cube root
// (sysdeps/ieee754/flt-32/s_cbrtf.c, Ulrich Drepper 1997, part of the GNU C
// Library, LGPL-2.1-or-later; its polynomial and factor constants are used
// here): split x = m * 2^e with m in [0.5, 1), a quadratic first guess u in
// double rounded to float, one Halley step u (u^3 + 2m) / (2u^3 + m) in
// double, times 2^(r/3) for r = e % 3 (C remainder), rounded to float and
// scaled by 2^(e/3). Equal to the host cbrtf for every float in [0, 2] (the
// XYB argument range is [0.0038, 1.0038]).

// glibc's factor table 2^(r/3), r = -2 .. 2, at index r + 2
#define SSIM2_CBRT_FACTORS_INIT \
  { 1.0 / 1.5874010519681994748, 1.0 / 1.2599210498948731648, 1.0, 1.2599210498948731648, 1.5874010519681994748 }

// a / b rounded to nearest, for positive normal a and b whose quotient is
// normal. On the device this is the IEEE double division sequence of the
// compiler without its range scaling and special-case fix-up (identities for
// such operands): hardware reciprocal, two Newton steps, the quotient and one
// remainder correction. kIeeeDiv selects the plain division (the form the
// reciprocal one is checked against on the GPU).
template <bool kIeeeDiv = false>
SSIM2_HD double div_rn(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (!kIeeeDiv) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
  }
#endif
  return a / b;
}

// `factor` indexes like SSIM2_CBRT_FACTORS_INIT (an array in any address
// space: the kernels keep it in LDS).
template <bool kIeeeDiv = false, typename Factors>
SSIM2_HD float cbrtf_ref(float x, const Factors& factor) {
  int e;
  const float m = frexpf(fabsf(x), &e);
  const double md = m;
  const float u = (float)(0.492659620528969547 + (0.697570460207922770 - 0.191502161678719066 * md) * md);
  const float t2 = u * u * u;
  const double ud = u, t2d = t2;
  // (t2 + 2m) and (2 t2 + m): the doublings are exact, so one fma rounds
  // each sum exactly as the sum of the doubled term does
  const double q = div_rn<kIeeeDiv>(ud * fma(2.0, md, t2d), fma(2.0, t2d, md));
  const float y = (float)(q * factor[e % 3 + 2]);
  const float res = ldexpf(x > 0.0f ? y : -y, e / 3);
  // zero, infinity and NaN: x + x (frexpf leaves them with exponent 0)
  return (x == 0.0f || !(x == x) || isinf(x)) ? x + x : res;
}

SSIM2_HD float cbrtf_ref(float x) {
  const double factor[5] = SSIM2_CBRT_FACTORS_INIT;
  return cbrtf_ref(x, factor);
}

// The same cube root for x >= 0 with the exponent handling folded into one
// table: for x = m * 2^e, cbrtf_ref rounds q * factor[e % 3 + 2] to float and
// scales by 2^(e / 3); scaling by a power of two commutes with both roundings
// (the double product stays normal, the float result is at least
// cbrt(2^-149) > 2^-50), so y = (float)(q * scale[e]) with
// scale[e] = factor[e % 3 + 2] * 2^(e / 3) is the same float. Exponents of
// positive finite floats run from -148 to 128.
constexpr int kCbrtExpMin = -148;
constexpr int kCbrtScales = 129 - kCbrtExpMin;  // e = -148 .. 128

SSIM2_HD double cbrt_scale(int e) {
  const int r = e % 3, k = e / 3;  // C remainder and quotient, as cbrtf_ref
  const double f = r == -2   ? 1.0 / 1.5874010519681994748
                   : r == -1 ? 1.0 / 1.2599210498948731648
                   : r == 0  ? 1.0
                   : r == 1  ? 1.2599210498948731648
                             : 1.5874010519681994748;
  return ldexp(f, k);
}

// `scale` indexes like cbrt_scale(e) at e - kCbrtExpMin (an array in any
// address space: the kernels keep it in LDS). On the device the quotient is
// the hardware reciprocal with one Newton step and one remainder correction;
// its equality to the IEEE form for every float in [0, 2] is checked on the
// device. Zero and +infinity return themselves (x + x in cbrtf_ref).
template <bool kIeeeDiv = false, typename Scales>
SSIM2_HD float cbrtf_nonneg(float x, const Scales& scale) {
  int e;
  const float m = frexpf(x, &e);
  const double md = m;
  const float u = (float)(0.492659620528969547 + (0.697570460207922770 - 0.191502161678719066 * md) * md);
  const float t2 = u * u * u;
  const double ud = u, t2d = t2;
  const double num = ud * fma(2.0, md, t2d), den = fma(2.0, t2d, md);
  double q;
#if defined(__HIP_DEVICE_COMPILE__)
  if (!kIeeeDiv) {
    double r = __builtin_amdgcn_rcp(den);
    r = fma(r, fma(-den, r, 1.0), r);
    q = num * r;
    q = fma(fma(-den, q, num), r, q);
  } else {
    q = num / den;
  }
#else
  q = num / den;
#endif
  const float y = (float)(q * scale[e - kCbrtExpMin]);
  return (x == 0.0f || isinf(x)) ? x : y;
}

// cpu.rs:458-496 px_linear_rgb_to_xyb followed by the positive shift;
// `scale` is the cube root's table (cbrt_scale).
template <typename Scales>
SSIM2_HD void linear_to_xyb(float r, float g, float b, float& X, float& Y, float& B, const Scales& scale) {
  float rg = fmaf(kM00, r, fmaf(kM01, g, fmaf(kM02, b, kB0)));
  float gr = fmaf(kM10, r, fmaf(kM11, g, fmaf(kM12, b, kB0)));
  float bb = fmaf(kM20, r, fmaf(kM21, g, fmaf(kM22, b, kB0)));
  rg = cbrtf_nonneg(fmaxf(rg, 0.0f), scale) - kB0Root;
  gr = cbrtf_nonneg(fmaxf(gr, 0.0f), scale) - kB0Root;
  bb = cbrtf_nonneg(fmaxf(bb, 0.0f), scale) - kB0Root;
  float x = 0.5f * (rg - gr);
  float y = 0.5f * (rg + gr);
  X = fmaf(x, 14.0f, 0.42f);
  Y = y + 0.01f;
  B = (bb - y) + 0.55f;
}

// a / b rounded to nearest in f32 for finite b with |b| in [2^-20, 2^100]
// and |a / b| >= 2^-100 (smaller quotients come out within 2^-90 of
// zero, which is all ssim_error needs of them): on the device the compiler's
// IEEE f32 division sequence without its range scaling and fix-up (hardware
// reciprocal, one Newton step, the quotient and two remainder corrections).
template <bool kIeeeDiv = false>
SSIM2_HD float div_rn_f32(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (!kIeeeDiv) {
    float r = __builtin_amdgcn_rcpf(b);
    r = fmaf(fmaf(-b, r, 1.0f), r, r);
    float q = a * r;
    q = fmaf(fmaf(-b, q, a), r, q);
    return fmaf(fmaf(-b, q, a), r, q);
  }
#endif
  return a / b;
}

// cpu.rs:605-630: 1 - SSIM' for one pixel, f32 ratio, f64 "1 -", clamped at 0.
// den >= C2 minus the rounding of the variances; a quotient below 2^-90 in
// magnitude leaves 1 - q equal to 1.0 in double either way.
template <bool kIeeeDiv = false>
SSIM2_HD double ssim_error(float mu1, float mu2, float s11, float s22, float s12) {
  float mu11 = mu1 * mu1, mu22 = mu2 * mu2, mu12 = mu1 * mu2;
  float md = mu1 - mu2;
  float num_m = fmaf(md, -md, 1.0f);
  float num_s = fmaf(2.0f, s12 - mu12, kC2);
  float den = ((s11 - mu11) + (s22 - mu22)) + kC2;
  double d = 1.0 - (double)div_rn_f32<kIeeeDiv>(num_m * num_s, den);
  return d > 0.0 ? d : 0.0;
}

// cpu.rs:658-674: edge-difference ratio in f64 (f32 differences). Both
// operands lie in [1, 2^128) for finite images.
template <bool kIeeeDiv = false>
SSIM2_HD double edge_ratio(float src, float mu1, float dis, float mu2) {
  return div_rn<kIeeeDiv>(1.0 + (double)fabsf(dis - mu2), 1.0 + (double)fabsf(src - mu1)) - 1.0;
}

// ---- YUV 4:2:0 BT.709 limited range -> linear RGB ------------------------
// crates/cuda-colorspace-kernel/src/biplanar.rs:8-70 with lib.rs:205-246:
// 16-bit-range samples (10-bit values shifted left by 6), luma floor 16<<8,
// chroma neutral 1<<15, Kr/Kb derived from the BT.709 primaries and D65, one
// chroma sample per 2x2 quad, BT.709 inverse OETF, clamped to [0, 1]. The
// coefficients are the f32 values of that derivation
// (y = 1/56064, r = 2(1-Kr)/57344, b = 2(1-Kb)/57344,
//  g1 = -2(1-Kb)Kb/Kg/57344, g2 = -2(1-Kr)Kr/Kg/57344).
constexpr float kYuvY = 0x1.2b404ap-16f, kYuvR = 0x1.ccb7fap-16f, kYuvB = 0x1.0f7334p-15f;
constexpr float kYuvG1 = -0x1.b66c0ap-19f, kYuvG2 = -0x1.11f7e2p-17f;
constexpr float kBt709AlphaM1 = 0x1.96b840p-4f, kBt709Alpha = 0x1.196b84p+0f;
constexpr float kBt709Threshold = 0x1.4cc550p-4f, kBt709InvGamma = 0x1.1c71c8p+1f;

// The power is 2^(log2(x) / 0.45) with the hardware base-2 logarithm and
// exponential (v_log_f32 / v_exp_f32, about 1 ulp each), as the reference's
// powf_fast is exp2(y log2 x) on its hardware. x = (v + 0.099) / 1.099 lies in
// [0.16, 1.2], so |log2 x| <= 2.6 and the result is within about 4e-7
// (relative) of the exact power.
SSIM2_HD float bt709_pow(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * kBt709InvGamma);
#else
  return exp2f(log2f(x) * kBt709InvGamma);
#endif
}

SSIM2_HD float bt709_eotf(float v) {
  return v >= kBt709Threshold ? bt709_pow((v + kBt709AlphaM1) / kBt709Alpha) : v / 4.5f;
}

SSIM2_HD void yuv16_to_linear(unsigned y, unsigned u, unsigned v, float* rgb) {
  const float cb = (float)((int)u - 32768), cr = (float)((int)v - 32768);
  const float r_ = kYuvR * cr;
  const float g_ = fmaf(kYuvG1, cb, kYuvG2 * cr);
  const float b_ = kYuvB * cb;
  const float luma = (float)((y > 4096u ? y : 4096u) - 4096u) * kYuvY;
  rgb[0] = fminf(fmaxf(bt709_eotf(luma + r_), 0.0f), 1.0f);
  rgb[1] = fminf(fmaxf(bt709_eotf(luma + g_), 0.0f), 1.0f);
  rgb[2] = fminf(fmaxf(bt709_eotf(luma + b_), 0.0f), 1.0f);
}

}  // namespace ssim2
