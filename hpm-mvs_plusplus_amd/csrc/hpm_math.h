// hpm_math.h -- scalar primitives shared by the HIP kernels and their host-side
// set-up code: the per-pixel Philox4x32-10 stream and the float/double
// exponential and float sine/cosine used by PatchMatch.
//
// Why own transcendentals: the reference evaluates exp/sin/cos with CUDA's
// fast-math library (CMakeLists.txt:10), whose results cannot be reproduced.
// These are written from the textbook method (Cody-Waite reduction by ln2 or
// pi/2, Taylor polynomial in Horner form with fused multiply-adds; J.-M.
// Muller, "Elementary Functions", ch. 11) as a fixed sequence of IEEE
// operations, so the exact FP32 path gives the same bits wherever it runs.
// The random generator replaces curand XORWOW seeded from clock64()
// (HPM.cu:579) with Philox4x32-10 (Salmon et al., SC'11): counter
// (block, pixel, stage, 0), key = 64-bit run seed.  Draws are (0, 1] like
// curand_uniform: ((x >> 8) + 1) * 2^-24.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HPM_HD __host__ __device__ __forceinline__

namespace hpm {

HPM_HD float bits_to_float(uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __uint_as_float(b);
#else
  float f; __builtin_memcpy(&f, &b, 4); return f;
#endif
}
HPM_HD double bits_to_double(uint64_t b) {
  double d; __builtin_memcpy(&d, &b, 8); return d;
}
HPM_HD float ffma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
HPM_HD double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
HPM_HD float frint(float x) { return __builtin_rintf(x); }
HPM_HD double drint(double x) { return __builtin_rint(x); }

// ------------------------------------------------------------------ Philox
struct Philox {
  uint32_t k0, k1, pixel, stage, block;
  uint32_t b0, b1, b2, b3;
  int pos;

  HPM_HD Philox(uint64_t seed, uint32_t pixel_, uint32_t stage_)
      : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), pixel(pixel_), stage(stage_), block(0),
        b0(0), b1(0), b2(0), b3(0), pos(4) {}

  HPM_HD static void rounds(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
      const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
      const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
      const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
      const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
      c1 = (uint32_t)p1;
      c3 = (uint32_t)p0;
      c0 = n0;
      c2 = n2;
    }
  }
  HPM_HD uint32_t next() {
    if (pos == 4) {
      b0 = block; b1 = pixel; b2 = stage; b3 = 0u;
      rounds(b0, b1, b2, b3, k0, k1);
      ++block;
      pos = 0;
    }
    const uint32_t v = pos == 0 ? b0 : (pos == 1 ? b1 : (pos == 2 ? b2 : b3));
    ++pos;
    return v;
  }
  HPM_HD float uniform() { return (float)((next() >> 8) + 1u) * 5.9604644775390625e-8f; }
};

// ------------------------------------------------------------------ exp (float)
HPM_HD float exp2i_f(int k) { return bits_to_float((uint32_t)(k + 127) << 23); }

HPM_HD float expf_hpm(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return __builtin_huge_valf();
  if (x < -103.97208404541015625f) return 0.0f;
  const float k = frint(x * 1.44269502162933349609375f);
  float r = ffma(k, -0.693145751953125f, x);
  r = ffma(k, -1.428606765330187045e-06f, r);
  float p = 1.98412698412698412698e-4f;
  p = ffma(p, r, 1.38888888888888888889e-3f);
  p = ffma(p, r, 8.33333333333333333333e-3f);
  p = ffma(p, r, 4.16666666666666666667e-2f);
  p = ffma(p, r, 1.66666666666666666667e-1f);
  p = ffma(p, r, 0.5f);
  p = ffma(p, r, 1.0f);
  p = ffma(p, r, 1.0f);
  const int ki = (int)k;
  if (ki > 127) return (p * exp2i_f(127)) * exp2i_f(ki - 127);
  if (ki < -126) return (p * exp2i_f(ki + 100)) * exp2i_f(-100);
  return p * exp2i_f(ki);
}

// ------------------------------------------------------------------ exp (double)
HPM_HD double exp2i_d(int k) { return bits_to_double((uint64_t)(k + 1023) << 52); }

HPM_HD double exp_hpm(double x) {
  if (x != x) return x;
  if (x > 709.782712893383973096) return __builtin_huge_val();
  if (x < -745.133219101941108420) return 0.0;
  const double k = drint(x * 1.44269504088896338700e+00);
  double r = dfma(k, -6.93147180369123816490e-01, x);
  r = dfma(k, -1.90821492927058770002e-10, r);
  double p = 1.0 / 6227020800.0;
  p = dfma(p, r, 1.0 / 479001600.0);
  p = dfma(p, r, 1.0 / 39916800.0);
  p = dfma(p, r, 1.0 / 3628800.0);
  p = dfma(p, r, 1.0 / 362880.0);
  p = dfma(p, r, 1.0 / 40320.0);
  p = dfma(p, r, 1.0 / 5040.0);
  p = dfma(p, r, 1.0 / 720.0);
  p = dfma(p, r, 1.0 / 120.0);
  p = dfma(p, r, 1.0 / 24.0);
  p = dfma(p, r, 1.0 / 6.0);
  p = dfma(p, r, 0.5);
  p = dfma(p, r, 1.0);
  p = dfma(p, r, 1.0);
  const int ki = (int)k;
  if (ki > 1023) return (p * exp2i_d(1023)) * exp2i_d(ki - 1023);
  if (ki < -1022) return (p * exp2i_d(ki + 600)) * exp2i_d(-600);
  return p * exp2i_d(ki);
}

// ------------------------------------------------------------------ sin / cos (float)
HPM_HD float sin_kernel(float r) {
  const float r2 = r * r;
  float p = -2.50521083854417187751e-8f;
  p = ffma(p, r2, 2.75573192239858906526e-6f);
  p = ffma(p, r2, -1.98412698412698412698e-4f);
  p = ffma(p, r2, 8.33333333333333333333e-3f);
  p = ffma(p, r2, -1.66666666666666666667e-1f);
  return ffma(p * r2, r, r);
}
HPM_HD float cos_kernel(float r) {
  const float r2 = r * r;
  float p = 2.08767569878680989792e-9f;
  p = ffma(p, r2, -2.75573192239858906526e-7f);
  p = ffma(p, r2, 2.48015873015873015873e-5f);
  p = ffma(p, r2, -1.38888888888888888889e-3f);
  p = ffma(p, r2, 4.16666666666666666667e-2f);
  p = ffma(p, r2, -0.5f);
  return ffma(p, r2, 1.0f);
}
HPM_HD int quadrant(float x, float* r) {
  const float k = frint(x * 0.636619772367581343076f);
  float t = ffma(k, -1.5703125f, x);
  t = ffma(k, -4.838705062866211e-4f, t);
  t = ffma(k, 4.371138828673793e-08f, t);
  *r = t;
  return ((int)k) & 3;
}
HPM_HD float sinf_hpm(float x) {
  if (!(__builtin_fabsf(x) < 1.0e5f)) return x - x;
  float r;
  const int q = quadrant(x, &r);
  const float s = sin_kernel(r), c = cos_kernel(r);
  return q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
}
HPM_HD float cosf_hpm(float x) {
  if (!(__builtin_fabsf(x) < 1.0e5f)) return x - x;
  float r;
  const int q = quadrant(x, &r);
  const float s = sin_kernel(r), c = cos_kernel(r);
  return q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}

// acos of a float, evaluated in double and rounded once (GetAngle's acosf,
// HPM.cpp:281-290).  asin(z) = z + sum_n c_n z^(2n+1), c_n = C(2n,n) / (4^n (2n+1))
// (the Taylor series of arcsin), truncated after 18 terms for |z| <= 1/2, where
// the first omitted term is below 2e-13 relative.  |x| <= 1/2 uses
// acos x = pi/2 - asin x; otherwise acos |x| = 2 asin(sqrt((1 - |x|) / 2)) and
// acos(-|x|) = pi - acos |x|.  Out-of-domain and NaN inputs give NaN.
HPM_HD double asin_series(double z) {
  const double c[18] = {0.16666666666666666, 0.075, 0.044642857142857144, 0.030381944444444444, 0.022372159090909092,
      0.017352764423076924, 0.01396484375, 0.011551800896139705, 0.009761609529194078,
      0.008390335809616815, 0.0073125258735988454, 0.006447210311889649, 0.005740037670841924,
      0.005153309682319905, 0.004660143486915096, 0.004240907093679363, 0.003880964558837669,
      0.0035692053938259347};
  const double z2 = z * z;
  double p = c[17];
  for (int k = 16; k >= 0; --k) p = dfma(p, z2, c[k]);
  return dfma(z * z2, p, z);
}
HPM_HD float acosf_hpm(float xf) {
  const double x = (double)xf;
  const double ax = x < 0.0 ? -x : x;
  if (!(ax <= 1.0)) return __builtin_nanf("");
  const double pio2 = 1.5707963267948966, pi = 3.141592653589793;
  if (ax <= 0.5) return (float)(pio2 - asin_series(x));
  const double a = 2.0 * asin_series(__builtin_sqrt((1.0 - ax) * 0.5));
  return (float)(x > 0.0 ? a : pi - a);
}

}  // namespace hpm
