// oracle_math.h -- scalar arithmetic primitives of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product may include, link or call
// this file; it exists so that tests/, the smoke check and the CPU-baseline leg
// of bench.py can check the HIP path.
//
// The reference draws its random numbers from curand XORWOW seeded with
// clock64() (HPM.cu:579), which makes every run different.  The rebuild replaces
// it with the counter-based Philox4x32-10 generator of Salmon, Moraes, Dror and
// Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11), keyed per pixel.
// The reference's float transcendentals come from CUDA's fast-math library
// (CMakeLists.txt:10, --use_fast_math), which cannot be reproduced; the oracle
// instead defines its own expf / sinf / cosf / exp (Cody-Waite argument
// reduction followed by a Taylor polynomial evaluated with fused multiply-adds,
// as in J.-M. Muller, "Elementary Functions", ch. 11), so that the same
// sequence of IEEE operations can be evaluated on the CPU and on the GPU.
// Every function here is pinned by tests/test_oracle_math.py against libm /
// numpy error bounds and, for Philox, against published known-answer vectors.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

// ---------------------------------------------------------------- Philox4x32-10
struct U4 { uint32_t v[4]; };

inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

inline U4 philox4x32_10(const uint32_t ctr[4], const uint32_t key[2]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    philox_round(c, k);
  }
  U4 o; std::memcpy(o.v, c, sizeof(c)); return o;
}

// Per-pixel stream: key = the 64-bit run seed, counter = (block, pixel, stage, 0).
// `stage` separates the initialisation kernel (0) from PatchMatch iteration it
// (1 + it); `block` advances every four draws.  Uniforms are (0, 1], like
// curand_uniform: u = ((x >> 8) + 1) * 2^-24, exact in float.
struct PixelRng {
  uint32_t key[2];
  uint32_t pixel, stage, block;
  uint32_t buf[4];
  int pos;
  PixelRng(uint64_t seed, uint32_t pixel_, uint32_t stage_)
      : pixel(pixel_), stage(stage_), block(0), pos(4) {
    key[0] = (uint32_t)seed; key[1] = (uint32_t)(seed >> 32);
  }
  uint32_t next_u32() {
    if (pos == 4) {
      const uint32_t ctr[4] = {block, pixel, stage, 0u};
      U4 o = philox4x32_10(ctr, key);
      std::memcpy(buf, o.v, sizeof(buf));
      ++block; pos = 0;
    }
    return buf[pos++];
  }
  float uniform() { return (float)((next_u32() >> 8) + 1u) * 5.9604644775390625e-8f; }
};

// ---------------------------------------------------------------- bit helpers
inline float pow2f(int k) {  // 2^k for k in [-126, 127]
  uint32_t b = (uint32_t)(k + 127) << 23; float f; std::memcpy(&f, &b, 4); return f;
}
inline double pow2d(int k) {  // 2^k for k in [-1022, 1023]
  uint64_t b = (uint64_t)(k + 1023) << 52; double d; std::memcpy(&d, &b, 8); return d;
}

// ---------------------------------------------------------------- float exp
// e^x = 2^k * e^r, k = rint(x / ln2), r = x - k*ln2 (two-part ln2), |r| <= ln2/2,
// e^r by its degree-7 Taylor polynomial (truncation < 6e-9 relative).
inline float expf_hpm(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return INFINITY;
  if (x < -103.97208404541015625f) return 0.0f;
  const float k = std::rint(x * 1.44269502162933349609375f);
  float r = std::fmaf(k, -0.693145751953125f, x);
  r = std::fmaf(k, -1.428606765330187045e-06f, r);
  float p = 1.98412698412698412698e-4f;
  p = std::fmaf(p, r, 1.38888888888888888889e-3f);
  p = std::fmaf(p, r, 8.33333333333333333333e-3f);
  p = std::fmaf(p, r, 4.16666666666666666667e-2f);
  p = std::fmaf(p, r, 1.66666666666666666667e-1f);
  p = std::fmaf(p, r, 0.5f);
  p = std::fmaf(p, r, 1.0f);
  p = std::fmaf(p, r, 1.0f);
  const int ki = (int)k;
  if (ki > 127) return (p * pow2f(127)) * pow2f(ki - 127);
  if (ki < -126) return (p * pow2f(ki + 100)) * pow2f(-100);
  return p * pow2f(ki);
}

// ---------------------------------------------------------------- double exp
// Same scheme in double: fdlibm's two-part ln2, degree-13 Taylor polynomial.
inline double exp_hpm(double x) {
  if (x != x) return x;
  if (x > 709.782712893383973096) return INFINITY;
  if (x < -745.133219101941108420) return 0.0;
  const double k = std::rint(x * 1.44269504088896338700e+00);
  double r = std::fma(k, -6.93147180369123816490e-01, x);
  r = std::fma(k, -1.90821492927058770002e-10, r);
  static const double c[14] = {
      1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
      1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,      1.0 / 720.0,
      1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,         0.5,
      1.0,                1.0};
  double p = c[0];
  for (int i = 1; i < 14; ++i) p = std::fma(p, r, c[i]);
  const int ki = (int)k;
  if (ki > 1023) return (p * pow2d(1023)) * pow2d(ki - 1023);
  if (ki < -1022) return (p * pow2d(ki + 600)) * pow2d(-600);
  return p * pow2d(ki);
}

// ---------------------------------------------------------------- float sin/cos
// Reduction by pi/2 (three-part Cody-Waite constant), then Taylor polynomials on
// |r| <= pi/4 (sin to x^11, cos to x^12; truncation < 2e-9).
inline float sin_poly(float r) {
  const float r2 = r * r;
  float p = -2.50521083854417187751e-8f;
  p = std::fmaf(p, r2, 2.75573192239858906526e-6f);
  p = std::fmaf(p, r2, -1.98412698412698412698e-4f);
  p = std::fmaf(p, r2, 8.33333333333333333333e-3f);
  p = std::fmaf(p, r2, -1.66666666666666666667e-1f);
  return std::fmaf(p * r2, r, r);
}
inline float cos_poly(float r) {
  const float r2 = r * r;
  float p = 2.08767569878680989792e-9f;
  p = std::fmaf(p, r2, -2.75573192239858906526e-7f);
  p = std::fmaf(p, r2, 2.48015873015873015873e-5f);
  p = std::fmaf(p, r2, -1.38888888888888888889e-3f);
  p = std::fmaf(p, r2, 4.16666666666666666667e-2f);
  p = std::fmaf(p, r2, -0.5f);
  return std::fmaf(p, r2, 1.0f);
}
inline int reduce_half_pi(float x, float* r) {
  const float k = std::rint(x * 0.636619772367581343076f);
  float t = std::fmaf(k, -1.5703125f, x);                  // pi/2 high part (8 bits)
  t = std::fmaf(k, -4.838705062866211e-4f, t);             // middle part (12 bits)
  t = std::fmaf(k, 4.371138828673793e-08f, t);             // low part: pi/2 - high - middle
  *r = t;
  return ((int)k) & 3;
}
inline float sinf_hpm(float x) {
  if (!(std::fabs(x) < 1.0e5f)) return x - x;  // NaN / inf / out of supported range
  float r; const int q = reduce_half_pi(x, &r);
  switch (q) {
    case 0: return sin_poly(r);
    case 1: return cos_poly(r);
    case 2: return -sin_poly(r);
    default: return -cos_poly(r);
  }
}
inline float cosf_hpm(float x) {
  if (!(std::fabs(x) < 1.0e5f)) return x - x;
  float r; const int q = reduce_half_pi(x, &r);
  switch (q) {
    case 0: return cos_poly(r);
    case 1: return -sin_poly(r);
    case 2: return -cos_poly(r);
    default: return sin_poly(r);
  }
}

// acos of a float, evaluated in double and rounded once (GetAngle's acosf,
// HPM.cpp:281-290).  asin(z) = z + sum_n c_n z^(2n+1), c_n = C(2n,n) / (4^n (2n+1))
// (the Taylor series of arcsin), truncated after 18 terms for |z| <= 1/2, where
// the first omitted term is below 2e-13 relative.  |x| <= 1/2 uses
// acos x = pi/2 - asin x; otherwise acos |x| = 2 asin(sqrt((1 - |x|) / 2)) and
// acos(-|x|) = pi - acos |x|.  Out-of-domain and NaN inputs give NaN.
inline double asin_series(double z) {
  const double c[18] = {0.16666666666666666, 0.075, 0.044642857142857144, 0.030381944444444444, 0.022372159090909092,
      0.017352764423076924, 0.01396484375, 0.011551800896139705, 0.009761609529194078,
      0.008390335809616815, 0.0073125258735988454, 0.006447210311889649, 0.005740037670841924,
      0.005153309682319905, 0.004660143486915096, 0.004240907093679363, 0.003880964558837669,
      0.0035692053938259347};
  const double z2 = z * z;
  double p = c[17];
  for (int k = 16; k >= 0; --k) p = std::fma(p, z2, c[k]);
  return std::fma(z * z2, p, z);
}
inline float acosf_hpm(float xf) {
  const double x = (double)xf;
  const double ax = x < 0.0 ? -x : x;
  if (!(ax <= 1.0)) return std::nanf("");
  const double pio2 = 1.5707963267948966, pi = 3.141592653589793;
  if (ax <= 0.5) return (float)(pio2 - asin_series(x));
  const double a = 2.0 * asin_series(std::sqrt((1.0 - ax) * 0.5));
  return (float)(x > 0.0 ? a : pi - a);
}

}  // namespace orc
