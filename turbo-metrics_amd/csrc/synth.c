/* Deterministic synthetic image-pair generators for the ssimulacra2 workloads.
 *
 * Host-only C (no GPU code). Built with -ffp-contract=off and without
 * -ffast-math so the bytes are identical on every x86-64 host with the same
 * libm; tests pin SHA-256 digests of small outputs.
 *
 * Randomness: SplitMix64 derives stream seeds, xoshiro256** drives noise and
 * choices, and a stateless 64-bit mix hashes lattice coordinates for value
 * noise.
 *
 * Workloads (w, h arbitrary >= 1):
 *   SYNTH_NOISE_GRAD_BLOCKMEAN  reference: 4 octaves of value noise (periods
 *       64/32/16/8 px, amplitudes 1, 1/2, 1/4, 1/8, normalised) mixed 0.7/0.3
 *       with the gradient (x+y)/(w+h). Distorted: reference + N(0, 4^2) per
 *       sample, then 10% of 8x8 blocks replaced by their per-channel mean.
 *   SYNTH_NOISE_RECT_DCT        reference: value noise with periods
 *       256/64/16/4 px plus 5 random-colour axis-aligned hard-edged
 *       rectangles. Distorted: per 8x8 block and channel, orthonormal 2-D
 *       DCT-II, coefficients quantised round(c/q)*q with AC step `param` and DC
 *       step param/2, inverse DCT. Partial edge blocks are completed by edge
 *       replication before the transform.
 *   SYNTH_NOISE_GRAD_DCT        reference as SYNTH_NOISE_GRAD_BLOCKMEAN,
 *       distortion as SYNTH_NOISE_RECT_DCT.
 * All outputs are packed RGB u8, rounded half away from zero and clamped.
 *
 * synth_yuv420_pair: 10-bit BT.709 limited-range 4:2:0 frames (u16 samples,
 *   LSB aligned). Reference: value noise sampled at (x + 2t, y) for frame t,
 *   mapped to Y in [64, 940] and Cb/Cr in [64, 960]. Distorted: luma
 *   requantised to 6-bit steps, Y' = 64 + round((Y-64)/16)*16 clamped to
 *   [64, 940]; chroma 3x3 box blur with edge clamp.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum {
  SYNTH_NOISE_GRAD_BLOCKMEAN = 1,
  SYNTH_NOISE_RECT_DCT = 2,
  SYNTH_NOISE_GRAD_DCT = 4,
};

static uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

typedef struct {
  uint64_t s[4];
} xoshiro;

static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static void xo_seed(xoshiro* r, uint64_t seed) {
  uint64_t sm = seed;
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&sm);
}

static uint64_t xo_next(xoshiro* r) {
  uint64_t* s = r->s;
  const uint64_t result = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

/* uniform [0,1) with 53 random bits */
static double xo_unif(xoshiro* r) { return (double)(xo_next(r) >> 11) * 0x1.0p-53; }

static double xo_gauss(xoshiro* r) {
  /* Box-Muller, one output per call (the second is discarded for simplicity). */
  double u1 = xo_unif(r), u2 = xo_unif(r);
  if (u1 < 0x1.0p-60) u1 = 0x1.0p-60;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 33)) * 0xFF51AFD7ED558CCDull;
  z = (z ^ (z >> 33)) * 0xC4CEB9FE1A85EC53ull;
  return z ^ (z >> 33);
}

static double lattice(uint64_t seed, int64_t ix, int64_t iy, int oct, int ch) {
  uint64_t h = mix64(seed ^ mix64((uint64_t)ix * 0x9E3779B97F4A7C15ull ^
                                  mix64((uint64_t)iy * 0xD1B54A32D192ED03ull ^
                                        ((uint64_t)(oct * 8 + ch) << 56))));
  return (double)(h >> 11) * 0x1.0p-53;
}

static double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

/* Value noise in [0,1) at (x, y): octaves with the given periods and
 * amplitudes 1, 1/2, 1/4, ..., normalised by the amplitude sum. */
static double value_noise(uint64_t seed, double x, double y, int ch, const int* periods, int nocts) {
  double acc = 0.0, amp = 1.0, norm = 0.0;
  for (int o = 0; o < nocts; ++o) {
    double fx = x / periods[o], fy = y / periods[o];
    double flx = floor(fx), fly = floor(fy);
    int64_t ix = (int64_t)flx, iy = (int64_t)fly;
    double tx = smooth(fx - flx), ty = smooth(fy - fly);
    double v00 = lattice(seed, ix, iy, o, ch), v10 = lattice(seed, ix + 1, iy, o, ch);
    double v01 = lattice(seed, ix, iy + 1, o, ch), v11 = lattice(seed, ix + 1, iy + 1, o, ch);
    double a = v00 + (v10 - v00) * tx;
    double b = v01 + (v11 - v01) * tx;
    acc += amp * (a + (b - a) * ty);
    norm += amp;
    amp *= 0.5;
  }
  return acc / norm;
}

static uint8_t to_u8(double v) {
  double r = v < 0.0 ? 0.0 : (v > 255.0 ? 255.0 : v);
  return (uint8_t)floor(r + 0.5);
}

static void ref_noise_grad(uint64_t seed, int w, int h, double* out) {
  static const int periods[4] = {64, 32, 16, 8};
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      double g = (double)(x + y) / (double)(w + h);
      for (int c = 0; c < 3; ++c) {
        double n = value_noise(seed, x, y, c, periods, 4);
        out[((size_t)y * w + x) * 3 + c] = (0.7 * n + 0.3 * g) * 255.0;
      }
    }
}

static void ref_noise_rects(uint64_t seed, int w, int h, double* out) {
  static const int periods[4] = {256, 64, 16, 4};
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x)
      for (int c = 0; c < 3; ++c)
        out[((size_t)y * w + x) * 3 + c] = value_noise(seed, x, y, c, periods, 4) * 255.0;
  xoshiro r;
  xo_seed(&r, seed ^ 0x5245435453ull);
  for (int k = 0; k < 5; ++k) {
    int x0 = (int)(xo_unif(&r) * w), y0 = (int)(xo_unif(&r) * h);
    int rw = 1 + (int)(xo_unif(&r) * (w / 3 + 1)), rh = 1 + (int)(xo_unif(&r) * (h / 3 + 1));
    double col[3];
    for (int c = 0; c < 3; ++c) col[c] = floor(xo_unif(&r) * 256.0);
    for (int y = y0; y < y0 + rh && y < h; ++y)
      for (int x = x0; x < x0 + rw && x < w; ++x)
        for (int c = 0; c < 3; ++c) out[((size_t)y * w + x) * 3 + c] = col[c];
  }
}

static void quantize_u8(const double* in, size_t n, uint8_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = to_u8(in[i]);
}

/* Orthonormal 8-point DCT-II basis, B[k][n] = a_k cos(pi (2n+1) k / 16). */
static void dct_basis(double B[8][8]) {
  for (int k = 0; k < 8; ++k)
    for (int n = 0; n < 8; ++n) {
      double a = k == 0 ? sqrt(0.125) : 0.5;
      B[k][n] = a * cos(3.141592653589793 * (2 * n + 1) * k / 16.0);
    }
}

static double round_half_away(double v) { return v < 0 ? -floor(-v + 0.5) : floor(v + 0.5); }

/* Distort packed u8 RGB by block-DCT quantisation (AC step q, DC step q/2). */
static void distort_dct(const uint8_t* ref, int w, int h, double q, uint8_t* dis) {
  double B[8][8];
  dct_basis(B);
  const double qdc = q * 0.5;
  for (int by = 0; by < h; by += 8)
    for (int bx = 0; bx < w; bx += 8)
      for (int c = 0; c < 3; ++c) {
        double blk[8][8], tmp[8][8], coef[8][8];
        for (int y = 0; y < 8; ++y)
          for (int x = 0; x < 8; ++x) {
            int sy = by + y < h ? by + y : h - 1, sx = bx + x < w ? bx + x : w - 1;
            blk[y][x] = ref[((size_t)sy * w + sx) * 3 + c];
          }
        for (int y = 0; y < 8; ++y)
          for (int k = 0; k < 8; ++k) {
            double s = 0;
            for (int n = 0; n < 8; ++n) s += B[k][n] * blk[y][n];
            tmp[y][k] = s;
          }
        for (int l = 0; l < 8; ++l)
          for (int k = 0; k < 8; ++k) {
            double s = 0;
            for (int n = 0; n < 8; ++n) s += B[l][n] * tmp[n][k];
            double step = (l == 0 && k == 0) ? qdc : q;
            coef[l][k] = round_half_away(s / step) * step;
          }
        for (int n = 0; n < 8; ++n)
          for (int k = 0; k < 8; ++k) {
            double s = 0;
            for (int l = 0; l < 8; ++l) s += B[l][n] * coef[l][k];
            tmp[n][k] = s;
          }
        for (int y = 0; y < 8; ++y)
          for (int x = 0; x < 8; ++x) {
            if (by + y >= h || bx + x >= w) continue;
            double s = 0;
            for (int k = 0; k < 8; ++k) s += B[k][x] * tmp[y][k];
            dis[((size_t)(by + y) * w + bx + x) * 3 + c] = to_u8(s);
          }
      }
}

static void distort_noise_blockmean(const double* refd, int w, int h, uint64_t seed, uint8_t* dis) {
  size_t n = (size_t)w * h * 3;
  double* v = (double*)malloc(n * sizeof(double));
  xoshiro r;
  xo_seed(&r, seed ^ 0x4E4F495345ull);
  for (size_t i = 0; i < n; ++i) v[i] = floor(refd[i] < 0 ? 0 : (refd[i] > 255 ? 255 : refd[i]) + 0.5) + 4.0 * xo_gauss(&r);
  xoshiro rb;
  xo_seed(&rb, seed ^ 0x424C4F434Bull);
  for (int by = 0; by < h; by += 8)
    for (int bx = 0; bx < w; bx += 8) {
      if (xo_unif(&rb) >= 0.1) continue;
      int ey = by + 8 < h ? by + 8 : h, ex = bx + 8 < w ? bx + 8 : w;
      for (int c = 0; c < 3; ++c) {
        double s = 0;
        for (int y = by; y < ey; ++y)
          for (int x = bx; x < ex; ++x) s += v[((size_t)y * w + x) * 3 + c];
        s /= (double)((ey - by) * (ex - bx));
        for (int y = by; y < ey; ++y)
          for (int x = bx; x < ex; ++x) v[((size_t)y * w + x) * 3 + c] = s;
      }
    }
  quantize_u8(v, n, dis);
  free(v);
}

/* Generate one packed-RGB u8 pair. Returns 0 on success, -1 on bad arguments. */
int synth_pair(int kind, uint64_t seed, int w, int h, double param, uint8_t* ref, uint8_t* dis) {
  if (w < 1 || h < 1 || !ref || !dis) return -1;
  size_t n = (size_t)w * h * 3;
  double* refd = (double*)malloc(n * sizeof(double));
  if (!refd) return -1;
  switch (kind) {
    case SYNTH_NOISE_GRAD_BLOCKMEAN:
      ref_noise_grad(seed, w, h, refd);
      quantize_u8(refd, n, ref);
      distort_noise_blockmean(refd, w, h, seed, dis);
      break;
    case SYNTH_NOISE_RECT_DCT:
      ref_noise_rects(seed, w, h, refd);
      quantize_u8(refd, n, ref);
      distort_dct(ref, w, h, param > 0 ? param : 16.0, dis);
      break;
    case SYNTH_NOISE_GRAD_DCT:
      ref_noise_grad(seed, w, h, refd);
      quantize_u8(refd, n, ref);
      distort_dct(ref, w, h, param > 0 ? param : 16.0, dis);
      break;
    default:
      free(refd);
      return -1;
  }
  free(refd);
  return 0;
}

/* Seeded helpers for the batched and serving workloads. */
uint64_t synth_mix_seed(uint64_t seed, uint64_t index) {
  uint64_t s = seed ^ mix64(index + 0x632BE59BD9B4E019ull);
  return splitmix64(&s);
}

/* Log-uniform integer in [lo, hi] from (seed, index, salt). */
int synth_log_uniform(uint64_t seed, uint64_t index, uint64_t salt, int lo, int hi) {
  xoshiro r;
  xo_seed(&r, synth_mix_seed(seed, index) ^ (salt * 0x9E3779B97F4A7C15ull));
  double u = xo_unif(&r);
  double v = exp(log((double)lo) + u * (log((double)hi + 1.0) - log((double)lo)));
  int iv = (int)floor(v);
  return iv < lo ? lo : (iv > hi ? hi : iv);
}

/* Uniform double in [0,1) from (seed, index, salt). */
double synth_uniform(uint64_t seed, uint64_t index, uint64_t salt) {
  xoshiro r;
  xo_seed(&r, synth_mix_seed(seed, index) ^ (salt * 0xD1B54A32D192ED03ull));
  return xo_unif(&r);
}

/* One 10-bit BT.709 limited-range 4:2:0 pair, frame t. Planes: Y w*h, Cb and
 * Cr ceil(w/2)*ceil(h/2). */
int synth_yuv420_pair(uint64_t seed, int w, int h, int t, uint16_t* y_ref, uint16_t* u_ref,
                      uint16_t* v_ref, uint16_t* y_dis, uint16_t* u_dis, uint16_t* v_dis) {
  static const int periods[4] = {128, 32, 8, 4};
  if (w < 2 || h < 2) return -1;
  int cw = (w + 1) / 2, ch = (h + 1) / 2;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      double n = value_noise(seed, x + 2.0 * t, y, 0, periods, 4);
      double Y = floor(64.0 + n * 876.0 + 0.5);
      uint16_t yv = (uint16_t)Y;
      y_ref[(size_t)y * w + x] = yv;
      double q = 64.0 + round_half_away((Y - 64.0) / 16.0) * 16.0;
      if (q > 940.0) q = 940.0;
      if (q < 64.0) q = 64.0;
      y_dis[(size_t)y * w + x] = (uint16_t)q;
    }
  for (int y = 0; y < ch; ++y)
    for (int x = 0; x < cw; ++x) {
      double nu = value_noise(seed, 2.0 * x + 2.0 * t, 2.0 * y, 1, periods, 4);
      double nv = value_noise(seed, 2.0 * x + 2.0 * t, 2.0 * y, 2, periods, 4);
      u_ref[(size_t)y * cw + x] = (uint16_t)floor(64.0 + nu * 896.0 + 0.5);
      v_ref[(size_t)y * cw + x] = (uint16_t)floor(64.0 + nv * 896.0 + 0.5);
    }
  for (int y = 0; y < ch; ++y)
    for (int x = 0; x < cw; ++x) {
      unsigned su = 0, sv = 0;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int yy = y + dy < 0 ? 0 : (y + dy >= ch ? ch - 1 : y + dy);
          int xx = x + dx < 0 ? 0 : (x + dx >= cw ? cw - 1 : x + dx);
          su += u_ref[(size_t)yy * cw + xx];
          sv += v_ref[(size_t)yy * cw + xx];
        }
      u_dis[(size_t)y * cw + x] = (uint16_t)((su + 4) / 9);
      v_dis[(size_t)y * cw + x] = (uint16_t)((sv + 4) / 9);
    }
  return 0;
}
