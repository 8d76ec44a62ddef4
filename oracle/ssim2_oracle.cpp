// ssimulacra2 CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the checker / the timed CPU baseline. The product
// (turbo-metrics_amd/) never links or calls it.
//
// It restates the reference's in-tree CPU ssimulacra2,
// crates/ssimulacra2-cuda/examples/cpu.rs, in three arithmetic modes:
//
//   MODE_F32_MIRROR (0)  the same f32/f64 mix and operation order as cpu.rs,
//                        fused multiply-adds exactly where cpu.rs calls
//                        mul_add, f64 accumulation per pixel. Built with
//                        -ffp-contract=off so no other contraction happens.
//   MODE_F64        (1)  every stage in double; the blur is the exact
//                        zero-padded 9-tap FIR the recursive Gaussian equals
//                        (double taps below). This is the FP64 oracle.
//   MODE_F32_FIR    (2)  f32 everywhere like the mirror, but the blur is the
//                        9-tap FIR in f32 (diagnostic: the spread between
//                        blur evaluation orders).
//   MODE_FAST       (3)  the arithmetic of the fast (batched / serving) GPU
//                        form: correctly rounded f32 cube root, 9-tap FIR in
//                        f32, error maps in f32, f64 sums.
//   MODE_FAST_F16   (4)  MODE_FAST with the XYB planes of scales >= 1 stored
//                        as IEEE binary16 of (value - channel shift), as the
//                        FP16-storage form keeps them in HBM.
//   MODE_FAST_F16_ALL (5) as 4, for every scale including scale 0.
//   MODE_FAST_LIN16 (6)  diagnostic: linear pyramid of scales >= 1 in binary16.
//   MODE_FAST_SHIFT (7)  MODE_FAST, but the five quantities are blurred in
//                        shifted coordinates x' = x - shift(channel) and the
//                        statistics are rebuilt with the exact zero-padding
//                        corrections (b1 = blur of ones, e = 1 - b1):
//                          var = (blur(x'^2) - blur(x')^2) + e*c*(2*blur(x') + c*b1)
//                          cov = (blur(r'd') - blur(r')blur(d')) + e*c*(blur(r') + blur(d') + c*b1)
//                          mu  = blur(x') + c*b1,   x - mu = (x' - blur(x')) + c*e
//   MODE_FAST_SHIFT_F16 (8) MODE_FAST_SHIFT with the XYB planes of scales >= 1
//                        stored in binary16 as in MODE_FAST_F16.
//   MODE_FAST_SHIFT_F16_ALL (9) MODE_FAST_SHIFT with the XYB planes of every
//                        scale, scale 0 included, stored in binary16 (the
//                        FP16-storage form keeps all scales in HBM).
//
// Parity pins (each checked in tests/):
//   * LUT and WEIGHT tables: equal, bit for bit, to the tables extracted from
//     cpu.rs:20-277 and cpu.rs:729-838 (tests/golden/reference_tables.json).
//   * IIR constants: re-derived in oracle/derive_iir.py and equal to
//     cpu.rs:931-948; the double IIR equals the 9-tap FIR used by MODE_F64.
//   * XYB: closed form of cpu.rs:458-496, compared with an independent double
//     evaluation; downscale against brute force on odd sizes.
//   * error maps: hand-worked pixels with closed-form values (equal means:
//     d from the variances only; src = mu1: pure artifact; dis = mu2: pure
//     detail loss), for the mirror's and the fast form's maps.
//   * score: each of the 108 features alone against cpu.rs:840-868's
//     transform with the fixture's weights, over six and over four scales.
//   * fast modes: round_f16 equal to IEEE binary16 (numpy) on every float
//     below 65520 of either sign; MODE_FAST_SHIFT within 1e-3 of MODE_F64 on
//     tiny and odd images (the shifted-statistics identities are exact);
//     MODE_FAST_SHIFT_F16_ALL within 0.05 of both. Each of these pins fails
//     with the matching defect planted (oracle_plant).
//   * whole pipeline: identical inputs score exactly 100 (cpu.rs:862-868),
//     JPEG corpus scores of the published ssimulacra2 0.5.1 crate (loose),
//     monotone in distortion strength, FP32 mirror vs FP64 within bounds.
//   * YUV 4:2:0 -> linear (video input): the reference holds no vectors for
//     it; pinned by the BT.709 Kr/Kb of Rec. ITU-R BT.709 (0.2126, 0.0722),
//     neutral chroma giving grey, the luma floor/ceiling giving 0/1, the
//     inverse OETF's continuity at the threshold, and a forward-encode ->
//     decode round trip (tests/test_oracle_yuv.py).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "ref_tables.h"

namespace {

// Planted defects (oracle_plant): 0 everywhere except in the CPU tests that
// show each pin of the oracle fails on a plausible mistake.
//   1  shifted statistics without the zero-padding corrections (e = 0)
//   2  binary16 rounding by truncation instead of to nearest even
//   3  edge maps on the signed difference (the absolute value dropped)
//   4  SSIM map without the luminance term num_m
//   5  score weights indexed as if all six scales were present
int g_plant = 0;

const float kLut[256] = ORACLE_SRGB8_LUT_INIT;
const double kWeight[108] = ORACLE_WEIGHTS_INIT;

// ---- XYB constants, cpu.rs:420-434 (f32, const-evaluated in f32) ----
const float K_M02 = 0.078f;
const float K_M00 = 0.30f;
const float K_M01 = (1.0f - K_M02) - K_M00;
const float K_M12 = 0.078f;
const float K_M10 = 0.23f;
const float K_M11 = (1.0f - K_M12) - K_M10;
const float K_M20 = 0.24342269f;
const float K_M21 = 0.20476745f;
const float K_M22 = (1.0f - K_M20) - K_M21;
const float K_B0 = 0.0037930734f;
const float K_B0_ROOT = 0.1559542025327239180319220163705f;

// ---- recursive Gaussian, cpu.rs:931-948 (values re-derived by derive_iir.py) ----
const int RADIUS = 5;
const float MUL_IN[3] = {0x1.c4fa84p-5f, -0x1.e1fd7ap-5f, 0x1.a8894ap-7f};
const float MUL_PREV[3] = {0x1.e6f0e2p+0f, 0x1.2cf230p+0f, 0x1.1a6264p-53f};
const float MUL_PREV2[3] = {-1.0f, -1.0f, -1.0f};
const float VERT_MUL_IN[3] = {0x1.c4fa84p-5f, -0x1.e1fd7ap-5f, 0x1.a8894ap-7f};
const float VERT_MUL_PREV[3] = {-0x1.e6f0e2p+0f, -0x1.2cf230p+0f, -0x1.1a6264p-53f};

// Centre-first taps of the 9-tap FIR equal to the double IIR (derive_iir.fir_taps_f64).
const double FIR64[5] = {0x1.0ef8d5c79cdacp-2, 0x1.b413e7d88b729p-3, 0x1.bfd672a600a7cp-4,
                         0x1.2700c7f8017a7p-5, 0x1.347d70ed6fe2cp-7};

const int NUM_SCALES = 6;

// ------------------------------------------------------------------ f32 mirror

// cpu.rs:458-496 px_linear_rgb_to_xyb (f32, FMA chain as written).
inline void xyb_f32(float r, float g, float b, float* X, float* Y, float* B) {
  float rg = std::fma(K_M00, r, std::fma(K_M01, g, std::fma(K_M02, b, K_B0)));
  float gr = std::fma(K_M10, r, std::fma(K_M11, g, std::fma(K_M12, b, K_B0)));
  float bb = std::fma(K_M20, r, std::fma(K_M21, g, std::fma(K_M22, b, K_B0)));
  rg = std::cbrt(std::max(rg, 0.0f)) - K_B0_ROOT;
  gr = std::cbrt(std::max(gr, 0.0f)) - K_B0_ROOT;
  bb = std::cbrt(std::max(bb, 0.0f)) - K_B0_ROOT;
  float x = 0.5f * (rg - gr);
  float y = 0.5f * (rg + gr);
  *X = std::fma(x, 14.0f, 0.42f);
  *Y = y + 0.01f;
  *B = (bb - y) + 0.55f;
}

// cpu.rs:545-579: ceil(in/2), 2x2 mean with coordinates clamped to in-1.
template <typename T>
void downscale_packed(const T* in, int w, int h, T* out, int ow, int oh) {
  const T norm = T(1) / T(4);
  for (int oy = 0; oy < oh; ++oy)
    for (int ox = 0; ox < ow; ++ox)
      for (int c = 0; c < 3; ++c) {
        T sum = 0;
        for (int iy = 0; iy < 2; ++iy)
          for (int ix = 0; ix < 2; ++ix) {
            int x = std::min(ox * 2 + ix, w - 1);
            int y = std::min(oy * 2 + iy, h - 1);
            sum += in[(size_t(y) * w + x) * 3 + c];
          }
        out[(size_t(oy) * ow + ox) * 3 + c] = sum * norm;
      }
}

// cpu.rs:967-1022 horizontal_row (row form).
void iir_row_f32(const float* in, float* out, int w) {
  float prev[3] = {0, 0, 0}, prev2[3] = {0, 0, 0};
  for (int n = -RADIUS + 1; n < w; ++n) {
    int left = n - RADIUS - 1, right = n + RADIUS - 1;
    float lv = left >= 0 ? in[left] : 0.0f;
    float rv = right < w ? in[right] : 0.0f;
    float sum = lv + rv;
    float o[3];
    for (int k = 0; k < 3; ++k) {
      o[k] = sum * MUL_IN[k];
      o[k] = std::fma(MUL_PREV2[k], prev2[k], o[k]);
      prev2[k] = prev[k];
      o[k] = std::fma(MUL_PREV[k], prev[k], o[k]);
      prev[k] = o[k];
    }
    if (n >= 0) out[n] = (o[0] + o[1]) + o[2];
  }
}

// cpu.rs:1054-1115 vertical_pass (column form), applied to every column.
void iir_cols_f32(const float* in, float* out, int w, int h) {
  std::vector<float> prev(3 * size_t(w), 0.0f), prev2(3 * size_t(w), 0.0f);
  for (int n = -RADIUS + 1; n < h; ++n) {
    int top = n - RADIUS - 1, bottom = n + RADIUS - 1;
    const float* trow = top >= 0 ? in + size_t(top) * w : nullptr;
    const float* brow = bottom < h ? in + size_t(bottom) * w : nullptr;
    for (int i = 0; i < w; ++i) {
      float sum = (trow ? trow[i] : 0.0f) + (brow ? brow[i] : 0.0f);
      float o[3];
      for (int k = 0; k < 3; ++k) {
        float t = std::fma(prev[k * size_t(w) + i], VERT_MUL_PREV[k], prev2[k * size_t(w) + i]);
        o[k] = std::fma(sum, VERT_MUL_IN[k], -t);
      }
      for (int k = 0; k < 3; ++k) {
        prev2[k * size_t(w) + i] = prev[k * size_t(w) + i];
        prev[k * size_t(w) + i] = o[k];
      }
      if (n >= 0) out[size_t(n) * w + i] = (o[0] + o[1]) + o[2];
    }
  }
}

// cpu.rs:913-918 blur_plane: rows then columns.
void blur_iir_f32(const float* in, float* out, int w, int h, std::vector<float>& tmp) {
  tmp.resize(size_t(w) * h);
  for (int y = 0; y < h; ++y) iir_row_f32(in + size_t(y) * w, tmp.data() + size_t(y) * w, w);
  iir_cols_f32(tmp.data(), out, w, h);
}

// 9-tap zero-padded FIR, rows then columns, outer taps first.
template <typename T>
void blur_fir(const T* in, T* out, int w, int h, std::vector<T>& tmp) {
  tmp.resize(size_t(w) * h);
  T taps[5];
  for (int k = 0; k < 5; ++k) taps[k] = T(FIR64[k]);
  auto at_row = [&](const T* row, int x) -> T { return (x >= 0 && x < w) ? row[x] : T(0); };
  for (int y = 0; y < h; ++y) {
    const T* row = in + size_t(y) * w;
    T* o = tmp.data() + size_t(y) * w;
    for (int x = 0; x < w; ++x) {
      T acc = taps[4] * (at_row(row, x - 4) + at_row(row, x + 4));
      for (int k = 3; k >= 1; --k) acc = std::fma(taps[k], at_row(row, x - k) + at_row(row, x + k), acc);
      o[x] = std::fma(taps[0], row[x], acc);
    }
  }
  auto at_col = [&](int x, int y) -> T { return (y >= 0 && y < h) ? tmp[size_t(y) * w + x] : T(0); };
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      T acc = taps[4] * (at_col(x, y - 4) + at_col(x, y + 4));
      for (int k = 3; k >= 1; --k) acc = std::fma(taps[k], at_col(x, y - k) + at_col(x, y + k), acc);
      out[size_t(y) * w + x] = std::fma(taps[0], tmp[size_t(y) * w + x], acc);
    }
}

// Per-scale sums in the layout sums[c*6 + j], j = {ssim S1, ssim S4, art S1,
// art S4, det S1, det S4}; S1 = sum of d, S4 = sum of d^4 (f64).
//
// cpu.rs:581-638 ssim_map (f32 formula, f64 "1 -" and accumulation).
void ssim_sums_f32(int n, const float* m1, const float* m2, const float* s11, const float* s22,
                   const float* s12, double* s1, double* s4) {
  const float C2 = 0.0009f;
  double a = 0, b = 0;
  for (int i = 0; i < n; ++i) {
    float mu1 = m1[i], mu2 = m2[i];
    float mu11 = mu1 * mu1, mu22 = mu2 * mu2, mu12 = mu1 * mu2;
    float mu_diff = mu1 - mu2;
    float num_m = g_plant == 4 ? 1.0f : std::fma(mu_diff, -mu_diff, 1.0f);
    float num_s = std::fma(2.0f, s12[i] - mu12, C2);
    float denom_s = ((s11[i] - mu11) + (s22[i] - mu22)) + C2;
    double d = 1.0 - double((num_m * num_s) / denom_s);
    d = std::max(d, 0.0);
    a += d;
    double d2 = d * d;
    b += d2 * d2;
  }
  *s1 = a;
  *s4 = b;
}

// cpu.rs:640-683 edge_diff_map (f32 differences, f64 ratio).
void edge_sums_f32(int n, const float* i1, const float* m1, const float* i2, const float* m2,
                   double* out4) {
  double s[4] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const float e2d = i2[i] - m2[i], e1d = i1[i] - m1[i];
    double d1 = g_plant == 3 ? (1.0 + double(e2d)) / (1.0 + double(e1d)) - 1.0
                             : (1.0 + double(std::fabs(e2d))) / (1.0 + double(std::fabs(e1d))) - 1.0;
    double art = std::max(d1, 0.0);
    double a2 = art * art;
    s[0] += art;
    s[1] += a2 * a2;
    double det = std::max(-d1, 0.0);
    double e2 = det * det;
    s[2] += det;
    s[3] += e2 * e2;
  }
  for (int k = 0; k < 4; ++k) out4[k] = s[k];
}

// ------------------------------------------------------------------ fast form

// IEEE binary16 round trip (round to nearest even) for finite |f| < 65520.
inline float round_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = x & 0x80000000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax < 0x38800000u) {  // below 2^-14: binary16 subnormal, quantum 2^-24
    float a;
    std::memcpy(&a, &ax, 4);
    const float q = 0x1.0p-24f;
    const float r = (g_plant == 2 ? std::trunc(a / q) : std::nearbyint(a / q)) * q;
    return sign ? -r : r;
  }
  const uint32_t lsb = (ax >> 13) & 1u;
  if (g_plant != 2) ax += 0xfffu + lsb;
  ax &= ~0x1fffu;
  const uint32_t y = sign | ax;
  float r;
  std::memcpy(&r, &y, 4);
  return r;
}

// Per-channel centre of the sRGB gamut in positive XYB (X 0.205..0.813,
// Y 0.010..0.855, B 0.272..0.938): binary16 stores value - shift, so the
// absolute rounding error is smallest where values cluster.
const float kXybShift[3] = {0.5f, 0.4375f, 0.625f};

inline void xyb_fast(float r, float g, float b, float* X, float* Y, float* B) {
  float rg = std::fma(K_M00, r, std::fma(K_M01, g, std::fma(K_M02, b, K_B0)));
  float gr = std::fma(K_M10, r, std::fma(K_M11, g, std::fma(K_M12, b, K_B0)));
  float bb = std::fma(K_M20, r, std::fma(K_M21, g, std::fma(K_M22, b, K_B0)));
  rg = (float)std::cbrt((double)std::max(rg, 0.0f)) - K_B0_ROOT;
  gr = (float)std::cbrt((double)std::max(gr, 0.0f)) - K_B0_ROOT;
  bb = (float)std::cbrt((double)std::max(bb, 0.0f)) - K_B0_ROOT;
  float x = 0.5f * (rg - gr);
  float y = 0.5f * (rg + gr);
  *X = std::fma(x, 14.0f, 0.42f);
  *Y = y + 0.01f;
  *B = (bb - y) + 0.55f;
}

void map_sums_fast(int n, const float* src, const float* dis, const float* m1, const float* m2,
                   const float* s11, const float* s22, const float* s12, double* out6) {
  const float C2 = 0.0009f;
  double a[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    float md = m1[i] - m2[i];
    float num_m = g_plant == 4 ? 1.0f : std::fma(md, -md, 1.0f);
    float num_s = std::fma(2.0f, s12[i] - m1[i] * m2[i], C2);
    float den = ((s11[i] - m1[i] * m1[i]) + (s22[i] - m2[i] * m2[i])) + C2;
    float d = std::max(1.0f - (num_m * num_s) / den, 0.0f);
    float d2 = d * d;
    a[0] += d;
    a[1] += d2 * d2;
    const float e2d = dis[i] - m2[i], e1d = src[i] - m1[i];
    float d1 = g_plant == 3 ? (1.0f + e2d) / (1.0f + e1d) - 1.0f
                            : (1.0f + std::fabs(e2d)) / (1.0f + std::fabs(e1d)) - 1.0f;
    float art = std::max(d1, 0.0f), det = std::max(-d1, 0.0f);
    float a2 = art * art, e2 = det * det;
    a[2] += art;
    a[3] += a2 * a2;
    a[4] += det;
    a[5] += e2 * e2;
  }
  for (int k = 0; k < 6; ++k) out6[k] = a[k];
}

// MODE_FAST_SHIFT statistics for one scale (see the header for the identities).
void shifted_scale_sums(const std::vector<float>* p1, const std::vector<float>* p2, int w, int h,
                        double* ssc) {
  const int n = w * h;
  std::vector<float> ones(n, 1.0f), b1(n), a(n), b(n), q(n), m1(n), m2(n), s11(n), s22(n), s12(n), tmp;
  blur_fir<float>(ones.data(), b1.data(), w, h, tmp);
  const float C2 = 0.0009f;
  for (int c = 0; c < 3; ++c) {
    const float sh = kXybShift[c];
    for (int i = 0; i < n; ++i) {
      a[i] = p1[c][i] - sh;
      b[i] = p2[c][i] - sh;
    }
    blur_fir<float>(a.data(), m1.data(), w, h, tmp);
    blur_fir<float>(b.data(), m2.data(), w, h, tmp);
    for (int i = 0; i < n; ++i) q[i] = a[i] * a[i];
    blur_fir<float>(q.data(), s11.data(), w, h, tmp);
    for (int i = 0; i < n; ++i) q[i] = b[i] * b[i];
    blur_fir<float>(q.data(), s22.data(), w, h, tmp);
    for (int i = 0; i < n; ++i) q[i] = a[i] * b[i];
    blur_fir<float>(q.data(), s12.data(), w, h, tmp);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
      const float e = g_plant == 1 ? 0.0f : 1.0f - b1[i];
      const float cb = sh * b1[i];
      const float var1 = (s11[i] - m1[i] * m1[i]) + e * sh * (2.0f * m1[i] + cb);
      const float var2 = (s22[i] - m2[i] * m2[i]) + e * sh * (2.0f * m2[i] + cb);
      const float cov = (s12[i] - m1[i] * m2[i]) + e * sh * ((m1[i] + m2[i]) + cb);
      const float md = m1[i] - m2[i];
      const float num_m = std::fma(md, -md, 1.0f);
      const float num_s = std::fma(2.0f, cov, C2);
      const float den = (var1 + var2) + C2;
      const float d = std::max(1.0f - (num_m * num_s) / den, 0.0f);
      const float d2 = d * d;
      acc[0] += d;
      acc[1] += d2 * d2;
      const float ce = sh * e;
      const float d1 = (1.0f + std::fabs((b[i] - m2[i]) + ce)) / (1.0f + std::fabs((a[i] - m1[i]) + ce)) - 1.0f;
      const float art = std::max(d1, 0.0f), det = std::max(-d1, 0.0f);
      const float a2 = art * art, e2 = det * det;
      acc[2] += art;
      acc[3] += a2 * a2;
      acc[4] += det;
      acc[5] += e2 * e2;
    }
    for (int k = 0; k < 6; ++k) ssc[c * 6 + k] = acc[k];
  }
}

// ------------------------------------------------------------------ f64 oracle

inline void xyb_f64(double r, double g, double b, double* X, double* Y, double* B) {
  const double m00 = 0.30, m02 = 0.078, m01 = 1.0 - m02 - m00;
  const double m10 = 0.23, m12 = 0.078, m11 = 1.0 - m12 - m10;
  const double m20 = 0.24342269, m21 = 0.20476745, m22 = 1.0 - m20 - m21;
  const double bias = 0.0037930734;
  const double root = 0.1559542025327239180319220163705;
  double rg = m00 * r + m01 * g + m02 * b + bias;
  double gr = m10 * r + m11 * g + m12 * b + bias;
  double bb = m20 * r + m21 * g + m22 * b + bias;
  rg = std::cbrt(std::max(rg, 0.0)) - root;
  gr = std::cbrt(std::max(gr, 0.0)) - root;
  bb = std::cbrt(std::max(bb, 0.0)) - root;
  double x = 0.5 * (rg - gr), y = 0.5 * (rg + gr);
  *X = x * 14.0 + 0.42;
  *Y = y + 0.01;
  *B = bb - y + 0.55;
}

void ssim_sums_f64(int n, const double* m1, const double* m2, const double* s11, const double* s22,
                   const double* s12, double* s1, double* s4) {
  const double C2 = 0.0009;
  double a = 0, b = 0;
  for (int i = 0; i < n; ++i) {
    double md = m1[i] - m2[i];
    double num_m = 1.0 - md * md;
    double num_s = 2.0 * (s12[i] - m1[i] * m2[i]) + C2;
    double den = (s11[i] - m1[i] * m1[i]) + (s22[i] - m2[i] * m2[i]) + C2;
    double d = std::max(1.0 - num_m * num_s / den, 0.0);
    a += d;
    b += (d * d) * (d * d);
  }
  *s1 = a;
  *s4 = b;
}

void edge_sums_f64(int n, const double* i1, const double* m1, const double* i2, const double* m2,
                   double* out4) {
  double s[4] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    double d1 = (1.0 + std::fabs(i2[i] - m2[i])) / (1.0 + std::fabs(i1[i] - m1[i])) - 1.0;
    double art = std::max(d1, 0.0), det = std::max(-d1, 0.0);
    s[0] += art;
    s[1] += (art * art) * (art * art);
    s[2] += det;
    s[3] += (det * det) * (det * det);
  }
  for (int k = 0; k < 4; ++k) out4[k] = s[k];
}

// ------------------------------------------------------------------ pipeline

// cpu.rs:342-410 compute_frame_ssimulacra2. Inputs are packed linear RGB in
// the working type T. Writes per-scale sums (18 per scale) and dims.
template <typename T>
int pipeline(const T* ref_lin, const T* dis_lin, int w0, int h0, int mode, double* sums,
             int* dims) {
  std::vector<T> img1(ref_lin, ref_lin + size_t(w0) * h0 * 3);
  std::vector<T> img2(dis_lin, dis_lin + size_t(w0) * h0 * 3);
  int w = w0, h = h0, nscales = 0;
  std::vector<T> p1[3], p2[3], mul, s11, s22, s12, mu1[3], mu2[3], tmp, t1, t2;
  for (int scale = 0; scale < NUM_SCALES; ++scale) {
    if (w < 8 || h < 8) break;
    if (scale > 0) {
      int ow = (w + 1) / 2, oh = (h + 1) / 2;
      t1.assign(size_t(ow) * oh * 3, T(0));
      t2.assign(size_t(ow) * oh * 3, T(0));
      downscale_packed(img1.data(), w, h, t1.data(), ow, oh);
      downscale_packed(img2.data(), w, h, t2.data(), ow, oh);
      img1.swap(t1);
      img2.swap(t2);
      w = ow;
      h = oh;
    }
    const int n = w * h;
    for (int c = 0; c < 3; ++c) {
      p1[c].resize(n);
      p2[c].resize(n);
    }
    for (int i = 0; i < n; ++i) {
      T a[3], b[3];
      if constexpr (sizeof(T) == 4) {
        if (mode >= 3) {
          float l1[3], l2[3];
          for (int c = 0; c < 3; ++c) {
            l1[c] = img1[i * 3 + c];
            l2[c] = img2[i * 3 + c];
            if (mode == 6 && scale > 0) {
              l1[c] = round_f16(l1[c]);
              l2[c] = round_f16(l2[c]);
            }
          }
          xyb_fast(l1[0], l1[1], l1[2], &a[0], &a[1], &a[2]);
          xyb_fast(l2[0], l2[1], l2[2], &b[0], &b[1], &b[2]);
          if (((mode == 4 || mode == 8) && scale > 0) || mode == 5 || mode == 9) {
            for (int c = 0; c < 3; ++c) {
              a[c] = round_f16(a[c] - kXybShift[c]) + kXybShift[c];
              b[c] = round_f16(b[c] - kXybShift[c]) + kXybShift[c];
            }
          }
        } else {
          xyb_f32(img1[i * 3], img1[i * 3 + 1], img1[i * 3 + 2], &a[0], &a[1], &a[2]);
          xyb_f32(img2[i * 3], img2[i * 3 + 1], img2[i * 3 + 2], &b[0], &b[1], &b[2]);
        }
      } else {
        xyb_f64(img1[i * 3], img1[i * 3 + 1], img1[i * 3 + 2], &a[0], &a[1], &a[2]);
        xyb_f64(img2[i * 3], img2[i * 3 + 1], img2[i * 3 + 2], &b[0], &b[1], &b[2]);
      }
      for (int c = 0; c < 3; ++c) {
        p1[c][i] = a[c];
        p2[c][i] = b[c];
      }
    }
    auto blur = [&](const std::vector<T>& in, std::vector<T>& out) {
      out.resize(n);
      if constexpr (sizeof(T) == 4) {
        if (mode == 0)
          blur_iir_f32(in.data(), out.data(), w, h, tmp);
        else
          blur_fir<float>(in.data(), out.data(), w, h, tmp);
      } else {
        blur_fir<double>(in.data(), out.data(), w, h, tmp);
      }
    };
    double* ssc = sums + size_t(scale) * 18;
    if constexpr (sizeof(T) == 4) {
      if (mode == 7 || mode == 8 || mode == 9) {
        shifted_scale_sums(p1, p2, w, h, ssc);
        dims[scale * 2] = w;
        dims[scale * 2 + 1] = h;
        ++nscales;
        continue;
      }
    }
    mul.resize(n);
    for (int c = 0; c < 3; ++c) {
      for (int i = 0; i < n; ++i) mul[i] = p1[c][i] * p1[c][i];
      blur(mul, s11);
      for (int i = 0; i < n; ++i) mul[i] = p2[c][i] * p2[c][i];
      blur(mul, s22);
      for (int i = 0; i < n; ++i) mul[i] = p1[c][i] * p2[c][i];
      blur(mul, s12);
      blur(p1[c], mu1[c]);
      blur(p2[c], mu2[c]);
      double e[4];
      if constexpr (sizeof(T) == 4) {
        if (mode >= 3) {
          double o6[6];
          map_sums_fast(n, p1[c].data(), p2[c].data(), mu1[c].data(), mu2[c].data(), s11.data(),
                        s22.data(), s12.data(), o6);
          ssc[c * 6 + 0] = o6[0];
          ssc[c * 6 + 1] = o6[1];
          for (int k = 0; k < 4; ++k) e[k] = o6[2 + k];
        } else {
          ssim_sums_f32(n, mu1[c].data(), mu2[c].data(), s11.data(), s22.data(), s12.data(),
                        &ssc[c * 6 + 0], &ssc[c * 6 + 1]);
          edge_sums_f32(n, p1[c].data(), mu1[c].data(), p2[c].data(), mu2[c].data(), e);
        }
      } else {
        ssim_sums_f64(n, mu1[c].data(), mu2[c].data(), s11.data(), s22.data(), s12.data(),
                      &ssc[c * 6 + 0], &ssc[c * 6 + 1]);
        edge_sums_f64(n, p1[c].data(), mu1[c].data(), p2[c].data(), mu2[c].data(), e);
      }
      for (int k = 0; k < 4; ++k) ssc[c * 6 + 2 + k] = e[k];
    }
    dims[scale * 2] = w;
    dims[scale * 2 + 1] = h;
    ++nscales;
  }
  return nscales;
}

// ---- YUV 4:2:0 BT.709 limited range -> linear RGB, in double ----
// crates/cuda-colorspace-kernel/src/biplanar.rs:8-70: 16-bit-range samples,
// luma floor 16<<8 over a range of (235-16)<<8, chroma neutral 1<<15 over
// (240-16)<<8, one chroma sample per 2x2 quad; lib.rs:185-201 derives the
// matrix from Kr/Kb, lib.rs:205-223 (constants_from_primaries) gets Kr/Kb from
// the BT.709 primaries and D65 white, lib.rs:206-223 the inverse OETF.
struct YuvCoeffs {
  double y, r, b, g1, g2;
};

YuvCoeffs bt709_coeffs() {
  struct V3 {
    double x, y, z;
  };
  auto xyz = [](double x, double y) { return V3{x / y, 1.0, (1.0 - x - y) / y}; };
  auto cross = [](V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; };
  auto dot = [](V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
  const V3 r = xyz(0.640, 0.330), g = xyz(0.300, 0.600), b = xyz(0.150, 0.060), w = xyz(0.3127, 0.3290);
  const V3 xr{r.x, g.x, b.x}, yr{r.y, g.y, b.y}, zr{r.z, g.z, b.z};
  const double mul = 1.0 / dot(xr, cross(yr, zr));
  const double kr = dot(w, cross(g, b)) * mul, kb = dot(w, cross(r, g)) * mul;
  const double kg = 1.0 - kr - kb;
  const double luma_range = double((235 - 16) << 8), chroma_range = double((240 - 16) << 8);
  return {1.0 / luma_range, 2.0 * (1.0 - kr) / chroma_range, 2.0 * (1.0 - kb) / chroma_range,
          -2.0 * (1.0 - kb) * kb / kg / chroma_range, -2.0 * (1.0 - kr) * kr / kg / chroma_range};
}

double bt709_eotf_f64(double v) {
  const double beta = 0.018053968510807, alpha = 1.0 + 5.5 * beta;
  const double threshold = alpha * std::pow(beta, 0.45) - (alpha - 1.0);  // OETF(beta)
  return v >= threshold ? std::pow((v + (alpha - 1.0)) / alpha, 1.0 / 0.45) : v / 4.5;
}

}  // namespace

extern "C" {

double oracle_score_from_sums(const double* sums, const int* dims, int nscales);

// Kr, Kb and the five matrix coefficients the oracle derives (for the pins).
void oracle_bt709_coeffs(double* out7) {
  const YuvCoeffs c = bt709_coeffs();
  const double m[7] = {c.y, c.r, c.b, c.g1, c.g2, 0, 0};
  std::memcpy(out7, m, sizeof(m));
  // Kr and Kb back from r and b: r = 2(1-Kr)/range
  const double chroma_range = double((240 - 16) << 8);
  out7[5] = 1.0 - c.r * chroma_range / 2.0;
  out7[6] = 1.0 - c.b * chroma_range / 2.0;
}

double oracle_bt709_eotf(double v) { return bt709_eotf_f64(v); }

// Planar 10-bit (LSB-aligned u16) 4:2:0 frame -> packed linear RGB double,
// w*h*3. Chroma planes are ceil(w/2) x ceil(h/2), packed.
void oracle_yuv420p10_to_linear_f64(const uint16_t* y, const uint16_t* cb, const uint16_t* cr, int w, int h,
                                    double* out) {
  const YuvCoeffs c = bt709_coeffs();
  const int cw = (w + 1) / 2;
  for (int py = 0; py < h; ++py)
    for (int px = 0; px < w; ++px) {
      const size_t ci = size_t(py / 2) * cw + px / 2;
      const double u = double(int(cb[ci] << 6) - 32768), v = double(int(cr[ci] << 6) - 32768);
      const int yy = int(y[size_t(py) * w + px]) << 6;
      const double luma = double(std::max(yy, 4096) - 4096) * c.y;
      const double rgb[3] = {luma + c.r * v, luma + c.g1 * u + c.g2 * v, luma + c.b * u};
      for (int k = 0; k < 3; ++k)
        out[(size_t(py) * w + px) * 3 + k] = std::min(std::max(bt709_eotf_f64(rgb[k]), 0.0), 1.0);
    }
}

// FP64 score of a pair of packed linear RGB double images (MODE_F64 pipeline).
int oracle_score_linear_f64(const double* ref, const double* dis, int w, int h, double* score) {
  if (w <= 0 || h <= 0) return -1;
  double sums[108] = {0};
  int dims[12] = {0};
  const int ns = pipeline<double>(ref, dis, w, h, 1, sums, dims);
  *score = oracle_score_from_sums(sums, dims, ns);
  return ns;
}

// FP64 score of a 10-bit 4:2:0 frame pair (conversion and pipeline in double).
int oracle_score_yuv420p10_f64(const uint16_t* yr, const uint16_t* ur, const uint16_t* vr, const uint16_t* yd,
                               const uint16_t* ud, const uint16_t* vd, int w, int h, double* score) {
  std::vector<double> a(size_t(w) * h * 3), b(size_t(w) * h * 3);
  oracle_yuv420p10_to_linear_f64(yr, ur, vr, w, h, a.data());
  oracle_yuv420p10_to_linear_f64(yd, ud, vd, w, h, b.data());
  double sums[108] = {0};
  int dims[12] = {0};
  const int ns = pipeline<double>(a.data(), b.data(), w, h, 1, sums, dims);
  *score = oracle_score_from_sums(sums, dims, ns);
  return ns;
}

// ---- per-stage entry points (for stage parity tests) ----

void oracle_srgb_lut(float* out256) { std::memcpy(out256, kLut, sizeof(kLut)); }

void oracle_weights(double* out108) { std::memcpy(out108, kWeight, sizeof(kWeight)); }

void oracle_srgb_to_linear(const uint8_t* in, size_t n, float* out) {
  for (size_t i = 0; i < n; ++i) out[i] = kLut[in[i]];
}

// Packed RGB -> packed XYB (positive-shifted), f32 mirror.
void oracle_linear_to_xyb_f32(const float* rgb, size_t npix, float* xyb) {
  for (size_t i = 0; i < npix; ++i)
    xyb_f32(rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2], &xyb[i * 3], &xyb[i * 3 + 1], &xyb[i * 3 + 2]);
}

void oracle_linear_to_xyb_f64(const double* rgb, size_t npix, double* xyb) {
  for (size_t i = 0; i < npix; ++i)
    xyb_f64(rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2], &xyb[i * 3], &xyb[i * 3 + 1], &xyb[i * 3 + 2]);
}

void oracle_downscale_f32(const float* in, int w, int h, float* out) {
  downscale_packed(in, w, h, out, (w + 1) / 2, (h + 1) / 2);
}

void oracle_blur_f32(const float* in, int w, int h, float* out) {
  std::vector<float> tmp;
  blur_iir_f32(in, out, w, h, tmp);
}

void oracle_blur_rows_f32(const float* in, int w, int h, float* out) {
  for (int y = 0; y < h; ++y) iir_row_f32(in + size_t(y) * w, out + size_t(y) * w, w);
}

void oracle_blur_cols_f32(const float* in, int w, int h, float* out) { iir_cols_f32(in, out, w, h); }

void oracle_blur_fir_f32(const float* in, int w, int h, float* out) {
  std::vector<float> tmp;
  blur_fir<float>(in, out, w, h, tmp);
}

void oracle_blur_f64(const double* in, int w, int h, double* out) {
  std::vector<double> tmp;
  blur_fir<double>(in, out, w, h, tmp);
}

// Per-pixel f32 error-map sums for one channel (cpu.rs:581-683); out6 in the
// sums layout {ssim S1, S4, art S1, S4, det S1, S4}.
void oracle_map_sums_f32(int n, const float* src, const float* dis, const float* mu1,
                         const float* mu2, const float* s11, const float* s22, const float* s12,
                         double* out6) {
  ssim_sums_f32(n, mu1, mu2, s11, s22, s12, &out6[0], &out6[1]);
  edge_sums_f32(n, src, mu1, dis, mu2, out6 + 2);
}

// Fast-form f32 error-map sums for one channel (the arithmetic of modes 3-6);
// out6 as oracle_map_sums_f32.
void oracle_map_sums_fast(int n, const float* src, const float* dis, const float* mu1, const float* mu2,
                          const float* s11, const float* s22, const float* s12, double* out6) {
  map_sums_fast(n, src, dis, mu1, mu2, s11, s22, s12, out6);
}

// The binary16 round trip the FP16-storage modes apply, element-wise.
void oracle_round_f16(const float* in, size_t n, float* out) {
  for (size_t i = 0; i < n; ++i) out[i] = round_f16(in[i]);
}

// Selects a planted defect (see g_plant); 0 restores the oracle.
void oracle_plant(int which) { g_plant = which; }

// Scale count and sizes per cpu.rs:358-368 (the first sub-8 scale is scored).
int oracle_scale_dims(int w, int h, int* dims) {
  int n = 0;
  for (int s = 0; s < NUM_SCALES; ++s) {
    if (w < 8 || h < 8) break;
    if (s > 0) {
      w = (w + 1) / 2;
      h = (h + 1) / 2;
    }
    dims[2 * s] = w;
    dims[2 * s + 1] = h;
    ++n;
  }
  return n;
}

// cpu.rs:633-634, 676-679 (averages) and cpu.rs:840-868 (score).
double oracle_score_from_sums(const double* sums, const int* dims, int nscales) {
  double ssim = 0.0;
  int i = 0;
  for (int c = 0; c < 3; ++c)
    for (int s = 0; s < nscales; ++s) {
      if (g_plant == 5) i = c * 36 + s * 6;
      const double opp = 1.0 / double(size_t(dims[2 * s]) * size_t(dims[2 * s + 1]));
      const double* q = sums + size_t(s) * 18 + c * 6;
      const double ssim_avg[2] = {opp * q[0], std::sqrt(std::sqrt(opp * q[1]))};
      const double art[2] = {opp * q[2], std::sqrt(std::sqrt(opp * q[3]))};
      const double det[2] = {opp * q[4], std::sqrt(std::sqrt(opp * q[5]))};
      for (int nrm = 0; nrm < 2; ++nrm) {
        ssim = std::fma(kWeight[i++], std::fabs(ssim_avg[nrm]), ssim);
        ssim = std::fma(kWeight[i++], std::fabs(art[nrm]), ssim);
        ssim = std::fma(kWeight[i++], std::fabs(det[nrm]), ssim);
      }
    }
  ssim *= 0.9562382616834844;
  ssim = std::fma(6.248496625763138e-5 * ssim * ssim, ssim,
                  std::fma(2.326765642916932, ssim, -0.020884521182843837 * ssim * ssim));
  if (ssim > 0.0)
    ssim = std::fma(std::pow(ssim, 0.6276336467831387), -10.0, 100.0);
  else
    ssim = 100.0;
  return ssim;
}

// Packed linear RGB f32 pair -> score. sums (108) / dims (12) may be null.
int oracle_score_linear(const float* ref, const float* dis, int w, int h, int mode, double* score,
                        double* sums_out, int* dims_out) {
  if (w <= 0 || h <= 0) return -1;
  double sums[108] = {0};
  int dims[12] = {0};
  int ns;
  if (mode == 1) {
    std::vector<double> a(ref, ref + size_t(w) * h * 3), b(dis, dis + size_t(w) * h * 3);
    ns = pipeline<double>(a.data(), b.data(), w, h, mode, sums, dims);
  } else {
    ns = pipeline<float>(ref, dis, w, h, mode, sums, dims);
  }
  *score = oracle_score_from_sums(sums, dims, ns);
  if (sums_out) std::memcpy(sums_out, sums, sizeof(sums));
  if (dims_out) std::memcpy(dims_out, dims, sizeof(dims));
  return ns;
}

// Packed sRGB u8 pair -> score (cpu.rs CpuImg::from_srgb then the pipeline).
int oracle_score_srgb(const uint8_t* ref, const uint8_t* dis, int w, int h, int mode, double* score,
                      double* sums_out, int* dims_out) {
  const size_t n = size_t(w) * h * 3;
  std::vector<float> a(n), b(n);
  oracle_srgb_to_linear(ref, n, a.data());
  oracle_srgb_to_linear(dis, n, b.data());
  return oracle_score_linear(a.data(), b.data(), w, h, mode, score, sums_out, dims_out);
}

// CPU throughput baseline: score `npairs` independent pairs on `nthreads`
// std::threads (mode as above). Pairs are packed sRGB u8.
int oracle_score_srgb_many(const uint8_t* const* refs, const uint8_t* const* diss, const int* ws,
                           const int* hs, int npairs, int mode, int nthreads, double* scores) {
  if (nthreads < 1) nthreads = 1;
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; ++t)
    pool.emplace_back([=]() {
      for (int i = t; i < npairs; i += nthreads)
        oracle_score_srgb(refs[i], diss[i], ws[i], hs[i], mode, &scores[i], nullptr, nullptr);
    });
  for (auto& th : pool) th.join();
  return 0;
}

}  // extern "C"
