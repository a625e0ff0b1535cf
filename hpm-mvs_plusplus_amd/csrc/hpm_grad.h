// hpm_grad.h -- backward pass of the multi-view bilateral NCC with respect to
// each pixel's plane, and the refinements built on it.
//
// The reference has no gradient.  The definition (shared with the FP64 oracle):
//   plane of a pixel from p = (d, t, f): n = (sin t cos f, sin t sin f, -cos t),
//   w = -n . X(d), X(d) = d q, q = K_r^-1 (x, y, 1) (Get3DPoint, HPM.cu:114-119);
//   C = sum_v a_v C_v / sum_v a_v over the views with weight a_v > 0 (weights
//   frozen), C_v the bilateral NCC of HPM.cu:375-448 on the mixed-precision
//   engine's FP32 texels; dC_v/dp = 0 where C_v is 2 (centre outside, small
//   variance) or clamped.
// Chain: the homography is H_v = A_v - b_v m^T with m = K_r^-T n / w, the same m
// for every view, so a tap at reference pixel (x, y) maps to u = X / Z with
// du/dm = (b2 u - b0) (x, y, 1) / Z (and v alike).  Per view the taps
// accumulate G = sum_k dC/ds_k (s_u du/dm + s_v dv/dm), i.e. three sums
// g_k * (x, y, 1), and dC_v/dp = G . dm/dp with dm/dd = -m / d,
// dm/dn_e = L e_e / w - m q_e / (n . q) (L = K_r^-T).  The bilateral weights
// are recomputed per pixel (36 exp), not stored; the taps are gathered twice
// (sums first, then the per-tap derivative terms).
// Gauss-Newton: 1 - NCC = |r^ - s^|^2 / 2 with r^, s^ the weighted, centred,
// unit-variance windows, so the residuals e_k = sqrt(w_k / W) (r^_k - s^_k)
// give J^T J = (S_DD - Dbar Dbar^T - E E^T) / var_s with D_k = ds_k/dp,
// S_DD = sum (w_k / W) D_k D_k^T, Dbar = sum (w_k / W) D_k,
// E = sum (w_k / W) s^_k D_k: three per-tap sums, no per-tap storage.
#pragma once
#include "hpm_fast.h"

namespace hpm {

// The per-(pixel, plane, view) frame of the 36 taps (FastEngine::ncc's).
struct TapFrame {
  float fu, fv, fg6, fg7, cu, bx, cv, by, lox, hix, loy, hiy, qpf;
  unsigned base;
  double flu, flv, Zc;
  __amdgpu_buffer_rsrc_t rsrc;
};

__device__ __forceinline__ double grad_recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = dfma(r, dfma(-x, r, 1.0), r);
  return dfma(r, dfma(-x, r, 1.0), r);
}

// false when the centre maps outside the source image
__device__ __forceinline__ bool tap_frame(const RunConsts& C, int v, int px, int py, const double m[3], TapFrame& f) {
  const double* A = C.A[v - 1];
  const double* b = C.bvec[v - 1];
  const double h0 = dfma(-b[0], m[0], A[0]), h1 = dfma(-b[0], m[1], A[1]), h2 = dfma(-b[0], m[2], A[2]);
  const double h3 = dfma(-b[1], m[0], A[3]), h4 = dfma(-b[1], m[1], A[4]), h5 = dfma(-b[1], m[2], A[5]);
  const double h6 = dfma(-b[2], m[0], A[6]), h7 = dfma(-b[2], m[1], A[7]), h8 = dfma(-b[2], m[2], A[8]);
  const double X = dfma(h0, px, dfma(h1, py, h2));
  const double Y = dfma(h3, px, dfma(h4, py, h5));
  const double Z = dfma(h6, px, dfma(h7, py, h8));
  const double rz = grad_recip(Z);
  const double uc = X * rz, vc = Y * rz;
  const int sw = C.cams[v].width, sh = C.cams[v].height;
  if (!(uc < (double)sw && uc >= 0.0 && vc < (double)sh && vc >= 0.0)) return false;
  const double g6 = h6 * rz, g7 = h7 * rz;
  f.fg6 = (float)g6;
  f.fg7 = (float)g7;
  const float au = (float)dfma(-uc, g6, h0 * rz), bu = (float)dfma(-uc, g7, h1 * rz);
  const float av = (float)dfma(-vc, g6, h3 * rz), bv = (float)dfma(-vc, g7, h4 * rz);
  f.flu = floor(uc);
  f.flv = floor(vc);
  const int iu = (int)f.flu + 1, iv = (int)f.flv + 1;
  f.fu = (float)(uc - f.flu);
  f.fv = (float)(vc - f.flv);
  f.lox = (float)(-iu); f.hix = (float)(sw - iu);
  f.loy = (float)(-iv); f.hiy = (float)(sh - iv);
  const int qp = sw + 2;  // elements per row of the pair image
  f.qpf = (float)qp;
  f.base = (unsigned)(iv * qp + iu);
  f.cu = ffma(f.fu, f.fg6, au);
  f.bx = ffma(f.fu, f.fg7, bu);
  f.cv = ffma(f.fv, f.fg6, av);
  f.by = ffma(f.fv, f.fg7, bv);
  f.Zc = Z;
  f.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(C.pairs[v]), (short)0,
                                             (qp * (sh + 1)) << PairF32::kShift, 0x00020000);
  return true;
}

// Tap (i, j): sample s, its derivatives (su, sv), the cell-frame position (x, y)
// and the ratio D = Z_k / Z_centre.
__device__ __forceinline__ void grad_tap(const TapFrame& f, float i, float j, float& s, float& su, float& sv, float& x,
                                         float& y, float& den) {
  den = ffma(j, f.fg7, ffma(i, f.fg6, 1.0f));
  const float r = __builtin_amdgcn_rcpf(den);
  x = ffma(j, f.bx, ffma(i, f.cu, f.fu)) * r;
  y = ffma(j, f.by, ffma(i, f.cv, f.fv)) * r;
  const float fx = floorf(x), fy = floorf(y);
  const float ax = x - fx, ay = y - fy;
  const float cxl = __builtin_amdgcn_fmed3f(fx, f.lox, f.hix);
  const float cyl = __builtin_amdgcn_fmed3f(fy, f.loy, f.hiy);
  const int rel = (int)ffma(cyl, f.qpf, cxl);
  const PairF32::Q q = PairF32::load(f.rsrc, ((unsigned)rel + f.base) << PairF32::kShift);
  // columns (t0, dv0) and (t1, dv1): l = t0 + ay dv0, r = t1 + ay dv1,
  // s = l + ax (r - l); ds/dax = r - l, ds/day = dv0 + ax (dv1 - dv0)
  const float l = ffma(ay, q.y, q.x);
  const float rt = ffma(ay, q.w, q.z);
  su = rt - l;
  sv = ffma(ax, q.w - q.y, q.y);
  s = ffma(ax, su, l);
}

// One thread per pixel.  params: (d, t, f, -) per pixel; weights: V per pixel.
// kGrad = false computes the cost only.
template <bool kGrad>
__global__ __launch_bounds__(256) void k_cost_grad(const RunConsts* __restrict__ Cp, const float4* __restrict__ params,
                                                   const float* __restrict__ weights, float* __restrict__ out_cost,
                                                   float* __restrict__ out_grad, float* __restrict__ out_gn) {
  const RunConsts& C = *Cp;
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= C.W || py >= C.H) return;
  const size_t p = (size_t)py * C.W + px;
  const int V = C.V;
  const float4 pp = params[p];
  const float d = pp.x, th = pp.y, ph = pp.z;
  // plane from (d, t, f); all derivative bookkeeping in double
  const double st = sin((double)th), ct = cos((double)th), sp = sin((double)ph), cp = cos((double)ph);
  const double n[3] = {st * cp, st * sp, -ct};
  const double dn_t[3] = {ct * cp, ct * sp, st};
  const double dn_f[3] = {-st * sp, st * cp, 0.0};
  const double q[3] = {((double)px - C.cams[0].K[2]) / C.cams[0].K[0], ((double)py - C.cams[0].K[5]) / C.cams[0].K[4], 1.0};
  const double nq = n[0] * q[0] + n[1] * q[1] + n[2] * q[2];
  const double w = -(double)d * nq;
  const double iw = grad_recip(w);
  double m[3];
  m[0] = n[0] * C.kinv[0] * iw;
  m[1] = n[1] * C.kinv[1] * iw;
  m[2] = dfma(-n[1], C.kinv[3], dfma(-n[0], C.kinv[2], n[2])) * iw;
  // dm/dp columns
  auto Ln = [&](const double e[3], double out[3]) {
    out[0] = e[0] * C.kinv[0];
    out[1] = e[1] * C.kinv[1];
    out[2] = e[2] - e[0] * C.kinv[2] - e[1] * C.kinv[3];
  };
  double Td[3], Tt[3], Tf[3], Lt[3], Lf[3];
  Ln(dn_t, Lt);
  Ln(dn_f, Lf);
  const double qt = q[0] * dn_t[0] + q[1] * dn_t[1] + q[2] * dn_t[2];
  const double qf = q[0] * dn_f[0] + q[1] * dn_f[1] + q[2] * dn_f[2];
  for (int e = 0; e < 3; ++e) {
    Td[e] = -m[e] / (double)d;
    Tt[e] = Lt[e] * iw - m[e] * qt / nq;
    Tf[e] = Lf[e] * iw - m[e] * qf / nq;
  }
  const float Tdf[3] = {(float)Td[0], (float)Td[1], (float)Td[2]};
  const float Ttf[3] = {(float)Tt[0], (float)Tt[1], (float)Tt[2]};
  const float Tff[3] = {(float)Tf[0], (float)Tf[1], (float)Tf[2]};
  FastWin fw;
  FastEngine<PairF32>::build(C, px, py, fw);
  const float W = 1.0f / fw.inv_w;
  float wsum = 0.0f, csum = 0.0f;
  double g[3] = {0.0, 0.0, 0.0}, Hm[6] = {0, 0, 0, 0, 0, 0};
  for (int v = 1; v <= V; ++v) {
    const float av = weights[p * V + (v - 1)];
    if (!(av > 0.0f)) continue;
    wsum += av;
    TapFrame f;
    float cost = 2.0f;
    bool live = tap_frame(C, v, px, py, m, f) && !(fw.var_r < 1e-5f);
    float s0 = 0.0f, S1 = 0.0f, S2 = 0.0f, S3 = 0.0f;
    float m_s = 0.0f, var_s = 0.0f, cov = 0.0f, den = 1.0f;
    if (live) {
      {  // shift: the sample at the window centre, near the window mean
        float su, sv, x, y, dd;
        grad_tap(f, 0.0f, 0.0f, s0, su, sv, x, y, dd);
      }
#pragma unroll
      for (int ii = 0; ii < kTapsPerAxis; ++ii)
#pragma unroll
        for (int jj = 0; jj < kTapsPerAxis; ++jj) {
          float s, su, sv, x, y, dd;
          grad_tap(f, (float)(-5 + 2 * ii), (float)(-5 + 2 * jj), s, su, sv, x, y, dd);
          const float spv = s - s0;
          const float ws = fw.w[ii * kTapsPerAxis + jj] * spv;
          S1 += ws;
          S2 = ffma(ws, spv, S2);
          S3 = ffma(fw.wr[ii * kTapsPerAxis + jj], spv, S3);
        }
      m_s = S1 * fw.inv_w;
      var_s = ffma(-m_s, m_s, S2 * fw.inv_w);
      if (var_s < 1e-5f) {
        live = false;
      } else {
        cov = ffma(-fw.m_r, m_s, S3 * fw.inv_w);
        den = sqrtf(fw.var_r * var_s);
        const float raw = 1.0f - cov / den;
        cost = fmaxf(0.0f, fminf(2.0f, raw));
        live = raw > 0.0f && raw < 2.0f;
      }
    }
    csum += av * cost;
    if (!kGrad || !live) continue;
    // second pass: dC/ds_k = -(1 / (W den)) [w_k (r_k - c - m_r) - (cov / var_s) w_k (s_k - s0 - m_s)]
    const float kc = -1.0f / (W * den), rho = cov / var_s;
    const double* b = C.bvec[v - 1];
    const float b0 = (float)b[0], b1 = (float)b[1], b2 = (float)b[2];
    const float Zc = (float)f.Zc, flu = (float)f.flu, flv = (float)f.flv;
    const bool want_gn = out_gn != nullptr;
    float G0 = 0.0f, G1 = 0.0f, G2 = 0.0f;
    // Gauss-Newton terms of the normalised-window residuals (see k_cost_grad's
    // comment): sums of w_k D_k D_k^T, w_k D_k and w_k (s_k - m_s) D_k
    float M[6] = {0, 0, 0, 0, 0, 0}, Db[3] = {0, 0, 0}, Eb[3] = {0, 0, 0};
#pragma unroll
    for (int ii = 0; ii < kTapsPerAxis; ++ii)
#pragma unroll
      for (int jj = 0; jj < kTapsPerAxis; ++jj) {
        const int k = ii * kTapsPerAxis + jj;
        float s, su, sv, x, y, dd;
        grad_tap(f, (float)(-5 + 2 * ii), (float)(-5 + 2 * jj), s, su, sv, x, y, dd);
        const float sc = s - s0 - m_s;
        const float dcds = kc * (fw.wr[k] - fw.w[k] * fw.m_r - rho * fw.w[k] * sc);
        const float Zk = Zc * dd;
        const float u = flu + x, vv = flv + y;
        const float beta = (su * (b2 * u - b0) + sv * (b2 * vv - b1)) / Zk;  // ds_k/dm = beta (x, y, 1)
        const float gk = dcds * beta;
        const float qx = (float)(px - 5 + 2 * ii), qy = (float)(py - 5 + 2 * jj);
        G0 = ffma(gk, qx, G0);
        G1 = ffma(gk, qy, G1);
        G2 += gk;
        if (want_gn) {
          const float D0 = beta * ffma(qx, Tdf[0], ffma(qy, Tdf[1], Tdf[2]));
          const float D1 = beta * ffma(qx, Ttf[0], ffma(qy, Ttf[1], Ttf[2]));
          const float D2 = beta * ffma(qx, Tff[0], ffma(qy, Tff[1], Tff[2]));
          const float wk = fw.w[k];
          const float w0 = wk * D0, w1 = wk * D1, w2 = wk * D2;
          M[0] = ffma(w0, D0, M[0]); M[1] = ffma(w0, D1, M[1]); M[2] = ffma(w0, D2, M[2]);
          M[3] = ffma(w1, D1, M[3]); M[4] = ffma(w1, D2, M[4]); M[5] = ffma(w2, D2, M[5]);
          Db[0] += w0; Db[1] += w1; Db[2] += w2;
          Eb[0] = ffma(w0, sc, Eb[0]); Eb[1] = ffma(w1, sc, Eb[1]); Eb[2] = ffma(w2, sc, Eb[2]);
        }
      }
    const double gd = G0 * Td[0] + G1 * Td[1] + G2 * Td[2];
    const double gt = G0 * Tt[0] + G1 * Tt[1] + G2 * Tt[2];
    const double gf = G0 * Tf[0] + G1 * Tf[1] + G2 * Tf[2];
    g[0] += av * gd;
    g[1] += av * gt;
    g[2] += av * gf;
    if (want_gn) {
      // H_v = (M / W - Db Db^T / W^2 - E E^T / (W^2 var_s)) / var_s
      const double iW = 1.0 / (double)W, ivs = 1.0 / (double)var_s;
      const double d0 = Db[0] * iW, d1 = Db[1] * iW, d2 = Db[2] * iW;
      const double e0 = Eb[0] * iW, e1 = Eb[1] * iW, e2 = Eb[2] * iW;
      const double hv[6] = {M[0] * iW - d0 * d0 - e0 * e0 * ivs, M[1] * iW - d0 * d1 - e0 * e1 * ivs,
                            M[2] * iW - d0 * d2 - e0 * e2 * ivs, M[3] * iW - d1 * d1 - e1 * e1 * ivs,
                            M[4] * iW - d1 * d2 - e1 * e2 * ivs, M[5] * iW - d2 * d2 - e2 * e2 * ivs};
      for (int e = 0; e < 6; ++e) Hm[e] += av * hv[e] * ivs;
    }
  }
  const float iws = wsum > 0.0f ? 1.0f / wsum : 0.0f;
  out_cost[p] = wsum > 0.0f ? csum * iws : 2.0f;
  if (kGrad) {
    out_grad[3 * p + 0] = (float)(g[0] * iws);
    out_grad[3 * p + 1] = (float)(g[1] * iws);
    out_grad[3 * p + 2] = (float)(g[2] * iws);
    if (out_gn)
      for (int e = 0; e < 6; ++e) out_gn[6 * p + e] = (float)(Hm[e] * iws);
  }
}

// Levenberg-Marquardt step of the Gauss-Newton system of the normalised-window
// residuals: (H + lambda diag(H) + eps I) delta = -grad; the candidate is clamped to
// the depth range and to normals facing the camera (t in [0, 85 deg]).
__global__ __launch_bounds__(256) void k_gn_step(int P, const float4* __restrict__ params, const float* __restrict__ grad,
                                                 const float* __restrict__ gn, const float* __restrict__ lambda,
                                                 float dmin, float dmax, float4* __restrict__ cand) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float* h = gn + 6 * (size_t)p;
  const float l = lambda[p];
  const double a00 = h[0] * (1.0 + l) + 1e-12, a01 = h[1], a02 = h[2];
  const double a11 = h[3] * (1.0 + l) + 1e-12, a12 = h[4], a22 = h[5] * (1.0 + l) + 1e-12;
  const double r0 = -grad[3 * p], r1 = -grad[3 * p + 1], r2 = -grad[3 * p + 2];
  // Cramer's rule on the symmetric 3 x 3 system
  const double c00 = a11 * a22 - a12 * a12, c01 = a02 * a12 - a01 * a22, c02 = a01 * a12 - a02 * a11;
  const double det = a00 * c00 + a01 * c01 + a02 * c02;
  float4 x = params[p];
  if (det > 0.0 && isfinite(det)) {
    const double c11 = a00 * a22 - a02 * a02, c12 = a01 * a02 - a00 * a12, c22 = a00 * a11 - a01 * a01;
    const double inv = 1.0 / det;
    x.x += (float)((c00 * r0 + c01 * r1 + c02 * r2) * inv);
    x.y += (float)((c01 * r0 + c11 * r1 + c12 * r2) * inv);
    x.z += (float)((c02 * r0 + c12 * r1 + c22 * r2) * inv);
  }
  x.x = fminf(fmaxf(x.x, dmin), dmax);
  x.y = fminf(fmaxf(x.y, 0.0f), 1.4835299f);  // 85 degrees
  cand[p] = x;
}

__global__ __launch_bounds__(256) void k_fill_f32(float* __restrict__ x, int n, float v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// Keeps the candidate where its cost is lower; adapts the damping.
__global__ __launch_bounds__(256) void k_accept(int P, float4* __restrict__ params, float* __restrict__ cost,
                                                const float4* __restrict__ cand, const float* __restrict__ cand_cost,
                                                float* __restrict__ lambda) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float c = cand_cost[p];
  if (c < cost[p]) {
    params[p] = cand[p];
    cost[p] = c;
    if (lambda) lambda[p] = fmaxf(lambda[p] * 0.3f, 1e-4f);
  } else if (lambda) {
    lambda[p] = fminf(lambda[p] * 4.0f, 1e6f);
  }
}

// Random-perturbation candidate (the search PlaneHypothesisRefinement does,
// HPM.cu:765-841, on the (d, t, f) parameters): d + U(-1, 1) sd, t and f +
// U(-1, 1) sa, clamped like k_gn_step.
__global__ __launch_bounds__(256) void k_random_step(int P, uint64_t seed, uint32_t stage, const float4* __restrict__ params,
                                                     float sd_rel, float sa, float dmin, float dmax,
                                                     float4* __restrict__ cand) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  Philox rng(seed, (uint32_t)p, stage);
  float4 x = params[p];
  const float u0 = 2.0f * rng.uniform() - 1.0f, u1 = 2.0f * rng.uniform() - 1.0f, u2 = 2.0f * rng.uniform() - 1.0f;
  x.x = fminf(fmaxf(x.x * (1.0f + sd_rel * u0), dmin), dmax);
  x.y = fminf(fmaxf(x.y + sa * u1, 0.0f), 1.4835299f);
  x.z = x.z + sa * u2;
  cand[p] = x;
}

// Separable Gaussian smoothing of an FP32 image (border replicated), for the
// smoothed first stage of the coarse-to-fine Gauss-Newton refinement.
constexpr int kBlurMaxRadius = 16;
struct BlurTaps {
  float w[2 * kBlurMaxRadius + 1];
  int r;
};

template <bool kRows>
__global__ __launch_bounds__(256) void k_blur(const float* __restrict__ in, float* __restrict__ out, int w, int h,
                                              BlurTaps t) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  float acc = 0.0f;
  for (int k = -t.r; k <= t.r; ++k) {
    int xx = kRows ? x + k : x, yy = kRows ? y : y + k;
    xx = xx < 0 ? 0 : (xx >= w ? w - 1 : xx);
    yy = yy < 0 ? 0 : (yy >= h ? h - 1 : yy);
    acc = fmaf(t.w[k + t.r], in[(size_t)yy * w + xx], acc);
  }
  out[(size_t)y * w + x] = acc;
}

}  // namespace hpm
