// hpm_fast.h -- mixed-precision cost engine: FP16 / BF16 / FP32 image storage
// with FP32 accumulation, plugged into the same PatchMatch control flow as the
// exact engine (same random draws, same decisions given the costs).
//
// What changes against the reference's ComputeBilateralNCC (HPM.cu:375-448):
//  * Homography.  H = K_s (R_rel - t_rel n^T / d) K_r^-1 = A - b m^T with the
//    view-pair constants A = K_s R_rel K_r^-1, b = K_s t_rel built once per run
//    in FP64 on the host, and m = K_r^-T n / d per plane.  The centre
//    projection (u_c, v_c) is evaluated in FP64 once per (pixel, plane, view).
//  * Tap positions are offsets from the centre, u = u_c + (i a_u + j b_u) /
//    (1 + i g6 + j g7), in FP32: small magnitudes keep ~1e-6 px accuracy where
//    absolute FP32 pixel coordinates carry ~6e-5 px of rounding at 1600 px.
//  * Texels.  A column-pair image holds, for every element (x, y) of the
//    (w+2) x (h+1) grid starting at texel (-1, -1), the texel column segment
//    (t(x, y), t(x, y+1) - t(x, y)) in grid coordinates, with the reference's
//    border clamp applied; the bilinear cell at (x, y) is elements x and
//    x + 1, one 8-byte
//    load (FP16 / BF16) or one 16-byte load (FP32) at element alignment.  The
//    sample is t0 + ay dv0 + ax ((t1 + ay dv1) - (t0 + ay dv0)): two fused
//    multiply-adds reading the FP16 halves directly (v_fma_mix), a subtraction
//    and one more fma.  For 8-bit images every component is an integer of
//    magnitude <= 255, exact in FP16 and BF16.  At 4 B per texel for FP16
//    (half the 8 B of a 2x2 cell per texel), the 32 lanes of a wave row that
//    gather at source columns two apart touch 256 B of a row instead of 512 B.
//  * Tap addresses: the clamped cell offset (dy, dx) is combined as the float
//    dy * (w + 2) + dx, exact below 2^24 elements per view (checked on the host).
//  * Sums use shifted data: the reference window is centred on its centre
//    texel, the source window on the bilinear sample at the centre projection
//    (one more gather per pixel and view), and where that still leaves the
//    window mean far from zero the window is summed again about its mean, so
//    E[x^2] - E[x]^2 never cancels large magnitudes.  The reference-side
//    statistics are accumulated in FP64 once per pixel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hpm_device.h"


namespace hpm {

// columns of taps whose gathers are issued ahead of the column being summed
constexpr int kPipeCols = 1;

// Column-pair images.  Element (x, y) of the (w + 2) x (h + 1) grid holds
// texel column x - 1 at rows y - 1 and y, border-clamped, as (t, t_below - t).
// The bilinear cell at grid position (x, y), whose top-left texel is
// (x - 1, y - 1), is the two adjacent elements x and x + 1: one 8-byte (FP16,
// BF16) or 16-byte (FP32) load at element alignment.  For 8-bit images both
// components are integers of magnitude <= 255, exact in FP16 and BF16.  The
// sample is  l = t0 + ay dv0,  r = t1 + ay dv1,  l + ax (r - l).
__device__ __forceinline__ uint32_t f16_bits(float f) {
  const _Float16 x = (_Float16)f;  // round to nearest even
  uint16_t b;
  __builtin_memcpy(&b, &x, 2);
  return b;
}
__device__ __forceinline__ uint32_t bf16_bits(float f) {  // round to nearest even (finite inputs)
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

struct PairF16 {
  using E = uint32_t;  // one element
  using Q = uint2;     // a cell: elements x and x + 1
  static constexpr int kind = 1;    // HPM_PREC_FP16
  static constexpr int kShift = 2;  // log2 sizeof(E)
  __device__ __forceinline__ static E pack(float top, float bot) {
    const _Float16 t = (_Float16)top;
    return f16_bits(top) | (f16_bits(bot - (float)t) << 16);
  }
  __device__ __forceinline__ static Q load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    return make_uint2(v[0], v[1]);
  }
  // two v_fma_mix_f32 reading the halves directly, then one lerp along x
  __device__ __forceinline__ static float bilinear(Q q, float ax, float ay) {
    const _Float16* c0 = reinterpret_cast<const _Float16*>(&q.x);
    const _Float16* c1 = reinterpret_cast<const _Float16*>(&q.y);
    const float l = ffma(ay, (float)c0[1], (float)c0[0]);
    const float r = ffma(ay, (float)c1[1], (float)c1[0]);
    return ffma(ax, r - l, l);
  }
};
struct PairBF16 {
  using E = uint32_t;
  using Q = uint2;
  static constexpr int kind = 2;  // HPM_PREC_BF16
  static constexpr int kShift = 2;
  __device__ __forceinline__ static E pack(float top, float bot) {
    const uint32_t t = bf16_bits(top);
    return t | (bf16_bits(bot - __uint_as_float(t << 16)) << 16);
  }
  __device__ __forceinline__ static Q load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    return make_uint2(v[0], v[1]);
  }
  __device__ __forceinline__ static float bilinear(Q q, float ax, float ay) {
    const float l = ffma(ay, __uint_as_float(q.x & 0xffff0000u), __uint_as_float(q.x << 16));
    const float r = ffma(ay, __uint_as_float(q.y & 0xffff0000u), __uint_as_float(q.y << 16));
    return ffma(ax, r - l, l);
  }
};
struct PairF32 {
  using E = float2;
  using Q = float4;
  static constexpr int kind = 3;  // HPM_PREC_FP32_FAST
  static constexpr int kShift = 3;
  __device__ __forceinline__ static E pack(float top, float bot) { return make_float2(top, bot - top); }
  __device__ __forceinline__ static Q load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
  }
  __device__ __forceinline__ static float bilinear(Q q, float ax, float ay) {
    const float l = ffma(ay, q.y, q.x);
    const float r = ffma(ay, q.w, q.z);
    return ffma(ax, r - l, l);
  }
};

// 8-bit column pairs: the storage of the FP16 precision for view sets whose
// FP16 column-pair images would not stay in the 256 MB Infinity Cache (the
// host chooses it; every texel must be an integer in [0, 255]).  Element
// (x, y) holds the texels (t, t_below) as two bytes, 2 B per element, half the
// FP16 footprint; a cell is one 4-byte load at element (2-byte) alignment.
// The sample is l = t0 + ay (b0 - t0), r = t1 + ay (b1 - t1), l + ax (r - l):
// the same fused multiply-adds of the same exact operands as the FP16 column
// pair's (t, b - t) halves, so costs are bit-identical to FP16 storage.  It
// costs six more VALU instructions per tap (byte extraction and conversion):
// slower where the images stay in the caches, faster where the FP16 images
// spill from them.
constexpr int kStorageU8 = 4;  // storage kind beside HPM_PREC_FP16 / BF16 / FP32_FAST
struct PairU8 {
  using E = uint16_t;
  using Q = uint32_t;
  static constexpr int kind = kStorageU8;
  static constexpr int kShift = 1;
  __device__ __forceinline__ static E pack(float top, float bot) {
    return (E)((uint32_t)top | ((uint32_t)bot << 8));
  }
  __device__ __forceinline__ static Q load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
  }
  __device__ __forceinline__ static float bilinear(Q q, float ax, float ay) {
    const float t0 = (float)(q & 0xffu), b0 = (float)((q >> 8) & 0xffu);
    const float t1 = (float)((q >> 16) & 0xffu), b1 = (float)(q >> 24);
    const float l = ffma(ay, b0 - t0, t0);
    const float r = ffma(ay, b1 - t1, t1);
    return ffma(ax, r - l, l);
  }
};

struct FastWin {
  float w[kTaps];    // bilateral weights (HPM.cu:342-347)
  float wr[kTaps];   // w * (r - r_centre)
  float inv_w;       // 1 / sum w
  float m_r;         // weighted mean of r - r_centre
  float var_r;       // weighted variance of the reference window
};

// Bilateral weight of tap (i, j) with reference texel r around centre texel c
// (HPM.cu:342-347).
__device__ __forceinline__ float fast_weight(int i, int j, float r, float c, float den_s, float den_c) {
  const float spatial = sqrtf((float)(i * i + j * j));
  return expf_hpm(-spatial / den_s - fabsf(r - c) / den_c);
}

template <class S>
struct FastEngine {
  static constexpr int kLanes = 1;
  using Win = FastWin;

  __device__ __forceinline__ static void build(const RunConsts& C, int px, int py, Win& fw) {
    const float* ref = C.img[0];
    const int W = C.cams[0].width, H = C.cams[0].height;
    const float c = tex_at(ref, W, H, px, py);
    const float den_s = 2.0f * C.sigma_spatial * C.sigma_spatial;
    const float den_c = 2.0f * C.sigma_color * C.sigma_color;
    double sw = 0.0, sr = 0.0, srr = 0.0;
#pragma unroll
    for (int ii = 0; ii < kTapsPerAxis; ++ii) {
#pragma unroll
      for (int jj = 0; jj < kTapsPerAxis; ++jj) {
        const int i = -5 + 2 * ii, j = -5 + 2 * jj;
        const float r = tex_at(ref, W, H, px + i, py + j);
        const float w = fast_weight(i, j, r, c, den_s, den_c);
        const float rs = r - c;  // exact: integer-valued texels
        fw.w[ii * kTapsPerAxis + jj] = w;
        fw.wr[ii * kTapsPerAxis + jj] = w * rs;
        sw += (double)w;
        sr += (double)w * (double)rs;
        srr += (double)w * (double)rs * (double)rs;
      }
    }
    const double m = sr / sw;
    fw.inv_w = (float)(1.0 / sw);
    fw.m_r = (float)m;
    fw.var_r = (float)(srr / sw - m * m);
  }

  // 1 / x in double: the hardware estimate refined by two Newton steps
  // (each doubles the correct bits), in place of the IEEE division sequence.
  __device__ __forceinline__ static double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = dfma(r, dfma(-x, r, 1.0), r);
    return dfma(r, dfma(-x, r, 1.0), r);
  }

  __device__ __forceinline__ static float ncc(const RunConsts& C, int v, int px, int py, float4 pl, const Win& fw) {
    // m = K_r^-T n / d (per plane), H = A - b m^T (per view pair), in double
    const double inv_d = recip((double)pl.w);
    const double m0 = (double)pl.x * C.kinv[0] * inv_d;
    const double m1 = (double)pl.y * C.kinv[1] * inv_d;
    const double m2 = dfma(-(double)pl.y, C.kinv[3], dfma(-(double)pl.x, C.kinv[2], (double)pl.z)) * inv_d;
    const double* A = C.A[v - 1];
    const double* b = C.bvec[v - 1];
    const double h0 = dfma(-b[0], m0, A[0]), h1 = dfma(-b[0], m1, A[1]), h2 = dfma(-b[0], m2, A[2]);
    const double h3 = dfma(-b[1], m0, A[3]), h4 = dfma(-b[1], m1, A[4]), h5 = dfma(-b[1], m2, A[5]);
    const double h6 = dfma(-b[2], m0, A[6]), h7 = dfma(-b[2], m1, A[7]), h8 = dfma(-b[2], m2, A[8]);
    const double X = dfma(h0, px, dfma(h1, py, h2));
    const double Y = dfma(h3, px, dfma(h4, py, h5));
    const double Z = dfma(h6, px, dfma(h7, py, h8));
    const double rz = recip(Z);
    const double uc = X * rz, vc = Y * rz;
    const DevCam& sc = C.cams[v];
    const int sw = sc.width, sh = sc.height;
    if (uc >= (double)sw || uc < 0.0 || vc >= (double)sh || vc < 0.0) return 2.0f;
    if (fw.var_r < 1e-5f) return 2.0f;
    const double g6 = h6 * rz, g7 = h7 * rz;
    const float fg6 = (float)g6, fg7 = (float)g7;
    const float au = (float)dfma(-uc, g6, h0 * rz), bu = (float)dfma(-uc, g7, h1 * rz);
    const float av = (float)dfma(-vc, g6, h3 * rz), bv = (float)dfma(-vc, g7, h4 * rz);
    const double flu = floor(uc), flv = floor(vc);
    const int iu = (int)flu + 1, iv = (int)flv + 1;  // +1: the pair grid starts at (-1, -1)
    const float fu = (float)(uc - flu), fv = (float)(vc - flv);
    // clamping the cell offset in float keeps the cell inside [0, w] x [0, h]
    // (the reference's border clamp); a NaN offset converts to 0, also inside.
    const float lox = (float)(-iu), hix = (float)(sw - iu);
    const float loy = (float)(-iv), hiy = (float)(sh - iv);
    const int qp = sw + 2;  // elements per row of the pair image
    const float qpf = (float)qp;
    const unsigned base = (unsigned)(iv * qp + iu);
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<void*>(C.pairs[v]), (short)0, (qp * (sh + 1)) << S::kShift, 0x00020000);
    // Tap (i, j) in the source cell frame: x = (fu + i cu + j bx) / D,
    // y = (fv + i cv + j by) / D, D = 1 + i g6 + j g7 (the projective map of
    // the window offset around the centre, with the centre's fraction folded in).
    const float cu = ffma(fu, fg6, au), bx = ffma(fu, fg7, bu);
    const float cv = ffma(fv, fg6, av), by = ffma(fv, fg7, bv);
    auto tap = [&](int ii, int jj, unsigned& off, float& ax, float& ay) {
      const float i = (float)(-5 + 2 * ii), j = (float)(-5 + 2 * jj);
      const float den = ffma(j, fg7, ffma(i, fg6, 1.0f));
      const float r = __builtin_amdgcn_rcpf(den);
      const float x = ffma(j, bx, ffma(i, cu, fu)) * r;
      const float y = ffma(j, by, ffma(i, cv, fv)) * r;
      const float fx = floorf(x), fy = floorf(y);
      ax = x - fx;
      ay = y - fy;
      const float cxl = __builtin_amdgcn_fmed3f(fx, lox, hix);
      const float cyl = __builtin_amdgcn_fmed3f(fy, loy, hiy);
      const int rel = (int)ffma(cyl, qpf, cxl);  // exact: |rel| < (w + 1) (h + 1) < 2^24
      off = ((unsigned)rel + base) << S::kShift;
    };
    // The window sums run on shifted samples.  The shift is the window-centre
    // sample (cell (iu, iv) at fractions (fu, fv), always inside the pair
    // grid), which lies near the window mean, so S2 / W - m_s^2 does not
    // cancel: shifted by the first tap, a corner sample 40-70 levels from the
    // mean, the FP32 sums lost up to 4e-4 of cost in windows of variance 10-80.
    const typename S::Q qc = S::load(rsrc, base << S::kShift);
    // Software pipeline over the six columns: the loads of columns
    // ii + 1 .. ii + kPipeCols are in flight while column ii is consumed.
    constexpr int L = kPipeCols;
    constexpr int R = L + 1;  // ring of column buffers
    typename S::Q q[R][kTapsPerAxis];
    float axr[R][kTapsPerAxis], ayr[R][kTapsPerAxis];
#pragma unroll
    for (int c = 0; c < L && c < kTapsPerAxis; ++c) {
#pragma unroll
      for (int jj = 0; jj < kTapsPerAxis; ++jj) {
        unsigned off;
        tap(c, jj, off, axr[c % R][jj], ayr[c % R][jj]);
        q[c % R][jj] = S::load(rsrc, off);
      }
    }
    float s0 = S::bilinear(qc, fu, fv);
    float S1 = 0.0f, S2 = 0.0f, S3 = 0.0f;
#pragma unroll
    for (int ii = 0; ii < kTapsPerAxis; ++ii) {
      if (ii + L < kTapsPerAxis) {
#pragma unroll
        for (int jj = 0; jj < kTapsPerAxis; ++jj) {
          unsigned off;
          tap(ii + L, jj, off, axr[(ii + L) % R][jj], ayr[(ii + L) % R][jj]);
          q[(ii + L) % R][jj] = S::load(rsrc, off);
        }
      }
#pragma unroll
      for (int jj = 0; jj < kTapsPerAxis; ++jj) {
        const float sp = S::bilinear(q[ii % R][jj], axr[ii % R][jj], ayr[ii % R][jj]) - s0;
        const float ws = fw.w[ii * kTapsPerAxis + jj] * sp;
        S1 += ws;
        S2 = ffma(ws, sp, S2);
        S3 = ffma(fw.wr[ii * kTapsPerAxis + jj], sp, S3);
      }
    }
    float m_s = S1 * fw.inv_w;
    float var_s = ffma(-m_s, m_s, S2 * fw.inv_w);
    if (__builtin_expect(m_s * m_s > 32.0f * var_s, 0)) {
      // The centre sample far from the window mean (m_s^2 > 32 var_s: windows
      // cut by the source border and windows whose centre is an outlier, from
      // a few in a million evaluations at 1600 x 1200 to 1.5e-4 on small views
      // of wide-baseline scenes): the window is summed again about the first
      // pass's mean, one column of six taps per iteration, each column's
      // weights taken from the window's registers through a uniform switch.
      // Measured alternatives: a loop of single taps that recomputed the
      // weights from the reference image ran so slowly that the few waves
      // taking it doubled short launches (768 x 512); unrolled, or with the
      // column loads and recomputed weights, it doubled the common path's
      // spills; running the whole first pass again in a loop cost 4.5 % at
      // 1600 x 1200.  At 32 var_s the largest cost error stays below 6e-6 on
      // the 1600 x 1200 plane scene and the full-resolution orbit views.
      s0 += m_s;
      S1 = 0.0f;
      S2 = 0.0f;
      S3 = 0.0f;
#pragma nounroll
      for (int ii = 0; ii < kTapsPerAxis; ++ii) {
        float wc[kTapsPerAxis], wrc[kTapsPerAxis];
        switch (ii) {
// case n copies column n's weights (compile-time register indices)
#define HPM_COL(n) case n: for (int jj = 0; jj < kTapsPerAxis; ++jj) { wc[jj] = fw.w[n * kTapsPerAxis + jj]; wrc[jj] = fw.wr[n * kTapsPerAxis + jj]; } break;
          HPM_COL(0) HPM_COL(1) HPM_COL(2) HPM_COL(3) HPM_COL(4) default: HPM_COL(5)
#undef HPM_COL
        }
        typename S::Q qs[kTapsPerAxis];
        float ax[kTapsPerAxis], ay[kTapsPerAxis];
#pragma unroll
        for (int jj = 0; jj < kTapsPerAxis; ++jj) {
          unsigned off;
          tap(ii, jj, off, ax[jj], ay[jj]);
          qs[jj] = S::load(rsrc, off);
        }
#pragma unroll
        for (int jj = 0; jj < kTapsPerAxis; ++jj) {
          const float sp = S::bilinear(qs[jj], ax[jj], ay[jj]) - s0;
          const float ws = wc[jj] * sp;
          S1 += ws;
          S2 = ffma(ws, sp, S2);
          S3 = ffma(wrc[jj], sp, S3);
        }
      }
      m_s = S1 * fw.inv_w;
      var_s = ffma(-m_s, m_s, S2 * fw.inv_w);
    }
    if (var_s < 1e-5f) return 2.0f;
    const float cov = ffma(-fw.m_r, m_s, S3 * fw.inv_w);
    const float c = 1.0f - cov * __builtin_amdgcn_rsqf(fw.var_r * var_s);
    return fmaxf(0.0f, fminf(2.0f, c));
  }
};

// Builds the column-pair image of one view from its FP32 texels.  Thread
// (x, y) of the (w+2) x (h+1) grid writes the element of texel column x-1,
// rows y-1 and y.
template <class S>
__global__ __launch_bounds__(256) void k_build_pairs(const float* __restrict__ img, int w, int h, void* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x > w + 1 || y > h) return;
  typename S::E* e = (typename S::E*)out;
  e[y * (w + 2) + x] = S::pack(tex_at(img, w, h, x - 1, y - 1), tex_at(img, w, h, x - 1, y));
}

// Bytes of the column-pair image of a w x h view.
template <class S> __host__ __device__ constexpr size_t pair_image_bytes(int w, int h) {
  return ((size_t)(w + 2) * (size_t)(h + 1)) << S::kShift;
}

}  // namespace hpm


