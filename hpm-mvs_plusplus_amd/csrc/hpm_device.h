// hpm_device.h -- device-side data layout and per-pixel building blocks of the
// PatchMatch kernels (exact FP32 path).
//
// Layout in HBM (one reference view):
//   images      FP32 row-major, one buffer per view (source sizes may differ)
//   planes      float4 per pixel: during PatchMatch (n_ref, d) with n.X + d = 0
//               (HPM.cu:135-140); after the output kernel (n_world, depth)
//               (HPM.cu:2840-2841)
//   costs, pre_costs, confidences, texture   float per pixel
//   selected    uint32 view mask per pixel
//   prior       float4 per pixel + uint32 mask (prior passes only)
//   src depths  FP32 per source view (geometric passes only)
// Everything else (cameras, the per-view-pair relative pose, thresholds) is
// one RunConsts block in device memory that every thread reads with scalar
// loads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hpm_math.h"

namespace hpm {

constexpr int kMaxImages = 33;
constexpr int kMaxSrc = 32;
constexpr int kTapsPerAxis = 6;            // window 11 at step 2 (HPM.cu:378,397,405)
constexpr int kTaps = kTapsPerAxis * kTapsPerAxis;

struct DevCam {  // struct Camera, main.h:39-47
  float K[9];
  float R[9];
  float t[3];
  int height;
  int width;
  float depth_min;
  float depth_max;
};

// View-pair part of ComputeHomography (HPM.cu:253-279), computed once per run
// on the host with the reference's FP32 operations.
struct PairBasis {
  float R_rel[9];
  float t_rel[3];
};

struct RunConsts {
  int W, H, V;                 // reference size, number of source views
  int storage;                 // image storage of the mixed-precision engine (HPM_PREC_*)
  int geom, prior, hierarchy, upsample, mand;
  int num_iterations;
  float depth_min, depth_max;  // cameras[0] range * (0.6, 1.2), HPM.cpp:685-686
  float sigma_spatial, sigma_color;
  int top_k;
  float scaled_cols, scaled_rows;
  uint64_t seed;
  float judge_thr[3][3];       // JudgeExtend good threshold (HPM.cu:1326-1327), float
  float vs_thr[3];             // view-selection threshold (HPM.cu:1706)
  float vs_fb[3];              // expf(thr^2 / -0.32) (HPM.cu:1724)
  float depth_sigma, two_dss, angle_sigma, two_ass, pert_refine, pert_prior_init;
  DevCam cams[kMaxImages];
  PairBasis pair[kMaxSrc];
  const float* img[kMaxImages];
  const float* dimg[kMaxImages];
  // mixed-precision engine: column-pair texel images and the FP64 homography basis
  const void* pairs[kMaxImages];  // (w+2) x (h+1) elements, origin at texel (-1, -1)
  double A[kMaxSrc][9];          // K_s R_rel K_r^-1
  double bvec[kMaxSrc][3];       // K_s t_rel
  double kinv[4];                // 1/fx, 1/fy, cx/fx, cy/fy of the reference camera
};

struct DevState {
  float4* planes;
  float* costs;
  float* pre_costs;
  uint32_t* sel;
  const float4* prior;
  const uint32_t* masks;
  float* conf;
  const uint32_t* canny;
  float* texture;
  const float4* scaled;
  unsigned long long* counters;  // [0] init NCC, [1] propagate NCC, [2] reference-count NCC
};

// ------------------------------------------------------------------ texel access
__device__ __forceinline__ float tex_at(const float* __restrict__ img, int w, int h, int x, int y) {
  x = x < 0 ? 0 : (x >= w ? w - 1 : x);
  y = y < 0 ? 0 : (y >= h ? h - 1 : y);
  return img[y * w + x];
}

// Bilinear sample at continuous position (u, v), texel centres at integers,
// border clamp (tex2D(src, u + 0.5, v + 0.5), HPM.cu:409).
__device__ __forceinline__ float bilinear(const float* __restrict__ img, int w, int h, float u, float v) {
  const float fu = floorf(u), fv = floorf(v);
  const float a = u - fu, b = v - fv;
  const float cu = fminf(fmaxf(fu, -1.0f), (float)w);
  const float cv = fminf(fmaxf(fv, -1.0f), (float)h);
  const int x0 = (int)cu, y0 = (int)cv;
  const int xa = x0 < 0 ? 0 : (x0 >= w ? w - 1 : x0);
  const int xb = x0 + 1 < 0 ? 0 : (x0 + 1 >= w ? w - 1 : x0 + 1);
  const int ya = y0 < 0 ? 0 : (y0 >= h ? h - 1 : y0);
  const int yb = y0 + 1 < 0 ? 0 : (y0 + 1 >= h ? h - 1 : y0 + 1);
  const float* ra = img + ya * w;
  const float* rb = img + yb * w;
  const float i00 = ra[xa], i10 = ra[xb], i01 = rb[xa], i11 = rb[xb];
  const float top = ffma(a, i10 - i00, i00);
  const float bot = ffma(a, i11 - i01, i01);
  return ffma(b, bot - top, top);
}

// ------------------------------------------------------------------ camera math
__device__ __forceinline__ void get3d(const DevCam& c, int px, int py, float depth, float X[3]) {  // HPM.cu:114-119
  X[0] = depth * ((float)px - c.K[2]) / c.K[0];
  X[1] = depth * ((float)py - c.K[5]) / c.K[4];
  X[2] = depth;
}
__device__ __forceinline__ float4 view_dir(const DevCam& c, int px, int py, float depth) {  // HPM.cu:121-133
  float X[3];
  get3d(c, px, py, depth, X);
  const float norm = sqrtf(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
  return make_float4(X[0] / norm, X[1] / norm, X[2] / norm, 0.0f);
}
__device__ __forceinline__ float dist_to_origin(const DevCam& c, int px, int py, float depth, float4 n) {  // HPM.cu:135-140
  float X[3];
  get3d(c, px, py, depth, X);
  return -(n.x * X[0] + n.y * X[1] + n.z * X[2]);
}
__device__ __forceinline__ float depth_from_plane(const DevCam& c, float4 pl, int px, int py) {  // HPM.cu:154-157
  const float K0 = c.K[0], K2 = c.K[2], K4 = c.K[4], K5 = c.K[5];
  return -pl.w * K0 / (((float)px - K2) * pl.x + (K0 / K4) * ((float)py - K5) * pl.y + K0 * pl.z);
}
__device__ __forceinline__ void normalize3(float4& v) {  // HPM.cu:89-96
  const float ns = v.x * v.x + v.y * v.y + v.z * v.z;
  const float inv = 1.0f / sqrtf(ns);
  v.x *= inv; v.y *= inv; v.z *= inv;
}
__device__ __forceinline__ float4 to_world(const DevCam& c, float4 p) {  // TransformNormal, HPM.cu:322-330
  const float* R = c.R;
  return make_float4(R[0] * p.x + R[3] * p.y + R[6] * p.z, R[1] * p.x + R[4] * p.y + R[7] * p.z,
                     R[2] * p.x + R[5] * p.y + R[8] * p.z, p.w);
}
__device__ __forceinline__ float4 to_ref(const DevCam& c, float4 p) {  // TransformNormal2RefCam, HPM.cu:332-340
  const float* R = c.R;
  return make_float4(R[0] * p.x + R[1] * p.y + R[2] * p.z, R[3] * p.x + R[4] * p.y + R[5] * p.z,
                     R[6] * p.x + R[7] * p.y + R[8] * p.z, p.w);
}

// Plane-dependent part of ComputeHomography (HPM.cu:281-310).
__device__ __forceinline__ void homography(const RunConsts& C, int v, float4 pl, float H[9]) {
  const PairBasis& B = C.pair[v - 1];
  const DevCam& rc = C.cams[0];
  const DevCam& sc = C.cams[v];
  const float n[3] = {pl.x, pl.y, pl.z};
  float Hm[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) Hm[3 * a + b] = B.R_rel[3 * a + b] - B.t_rel[a] * n[b] / pl.w;
  const float rK0 = rc.K[0], rK2 = rc.K[2], rK4 = rc.K[4], rK5 = rc.K[5];
  float tmp[9];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    tmp[3 * a + 0] = Hm[3 * a + 0] / rK0;
    tmp[3 * a + 1] = Hm[3 * a + 1] / rK4;
    tmp[3 * a + 2] = -Hm[3 * a + 0] * rK2 / rK0 - Hm[3 * a + 1] * rK5 / rK4 + Hm[3 * a + 2];
  }
  const float sK0 = sc.K[0], sK2 = sc.K[2], sK4 = sc.K[4], sK5 = sc.K[5], sK8 = sc.K[8];
  H[0] = sK0 * tmp[0] + sK2 * tmp[6];
  H[1] = sK0 * tmp[1] + sK2 * tmp[7];
  H[2] = sK0 * tmp[2] + sK2 * tmp[8];
  H[3] = sK4 * tmp[3] + sK5 * tmp[6];
  H[4] = sK4 * tmp[4] + sK5 * tmp[7];
  H[5] = sK4 * tmp[5] + sK5 * tmp[8];
  H[6] = sK8 * tmp[6];
  H[7] = sK8 * tmp[7];
  H[8] = sK8 * tmp[8];
}

// ------------------------------------------------------------------ reference window
// Everything in ComputeBilateralNCC that depends only on the reference image:
// the 36 bilateral weights (HPM.cu:342-347), w*ref per tap, and the weighted
// reference mean/variance with the reference's two-level summation order
// (HPM.cu:397-436).  Identical in every call for a given pixel.
struct RefWin {
  float w[kTaps];
  float wr[kTaps];
  float inv_wsum;
  float m_ref;
  float var_ref;
};

__device__ __forceinline__ void build_refwin(const RunConsts& C, int px, int py, RefWin& rw) {
  const float* ref = C.img[0];
  const int W = C.cams[0].width, H = C.cams[0].height;
  const float center = tex_at(ref, W, H, px, py);
  const float den_s = 2.0f * C.sigma_spatial * C.sigma_spatial;
  const float den_c = 2.0f * C.sigma_color * C.sigma_color;
  float s_ref = 0.0f, s_rr = 0.0f, s_w = 0.0f;
#pragma unroll
  for (int ii = 0; ii < kTapsPerAxis; ++ii) {
    const int i = -5 + 2 * ii;
    float r_ref = 0.0f, r_rr = 0.0f, r_w = 0.0f;
#pragma unroll
    for (int jj = 0; jj < kTapsPerAxis; ++jj) {
      const int j = -5 + 2 * jj;
      const float rp = tex_at(ref, W, H, px + i, py + j);
      const float spatial = sqrtf((float)i * (float)i + (float)j * (float)j);
      const float color = fabsf(rp - center);
      const float w = expf_hpm(-spatial / den_s - color / den_c);
      const float wr = w * rp;
      rw.w[ii * kTapsPerAxis + jj] = w;
      rw.wr[ii * kTapsPerAxis + jj] = wr;
      r_ref += wr;
      r_rr = ffma(wr, rp, r_rr);
      r_w += w;
    }
    s_ref += r_ref;
    s_rr += r_rr;
    s_w += r_w;
  }
  rw.inv_wsum = 1.0f / s_w;
  rw.m_ref = s_ref * rw.inv_wsum;
  const float m_rr = s_rr * rw.inv_wsum;
  rw.var_ref = ffma(-rw.m_ref, rw.m_ref, m_rr);
}

// ComputeBilateralNCC (HPM.cu:375-448), exact FP32.
__device__ __forceinline__ float ncc_exact(const RunConsts& C, int v, int px, int py, float4 pl, const RefWin& rw) {
  float H[9];
  homography(C, v, pl, H);
  const DevCam& sc = C.cams[v];
  {
    const float x = (float)px, y = (float)py;
    const float X = ffma(H[0], x, ffma(H[1], y, H[2]));
    const float Y = ffma(H[3], x, ffma(H[4], y, H[5]));
    const float Z = ffma(H[6], x, ffma(H[7], y, H[8]));
    const float u = X / Z, vv = Y / Z;
    if (u >= (float)sc.width || u < 0.0f || vv >= (float)sc.height || vv < 0.0f) return 2.0f;
  }
  const float* src = C.img[v];
  const int sw = sc.width, sh = sc.height;
  float s_src = 0.0f, s_ss = 0.0f, s_rs = 0.0f;
#pragma unroll
  for (int ii = 0; ii < kTapsPerAxis; ++ii) {
    const float x = (float)(px - 5 + 2 * ii);
    float r_src = 0.0f, r_ss = 0.0f, r_rs = 0.0f;
#pragma unroll
    for (int jj = 0; jj < kTapsPerAxis; ++jj) {
      const float y = (float)(py - 5 + 2 * jj);
      const float X = ffma(H[0], x, ffma(H[1], y, H[2]));
      const float Y = ffma(H[3], x, ffma(H[4], y, H[5]));
      const float Z = ffma(H[6], x, ffma(H[7], y, H[8]));
      const float s = bilinear(src, sw, sh, X / Z, Y / Z);
      const float ws = rw.w[ii * kTapsPerAxis + jj] * s;
      r_src += ws;
      r_ss = ffma(ws, s, r_ss);
      r_rs = ffma(rw.wr[ii * kTapsPerAxis + jj], s, r_rs);
    }
    s_src += r_src;
    s_ss += r_ss;
    s_rs += r_rs;
  }
  const float m_s = s_src * rw.inv_wsum;
  const float m_ss = s_ss * rw.inv_wsum;
  const float m_rs = s_rs * rw.inv_wsum;
  const float var_s = ffma(-m_s, m_s, m_ss);
  if (rw.var_ref < 1e-5f || var_s < 1e-5f) return 2.0f;
  const float cov = ffma(-rw.m_ref, m_s, m_rs);
  const float den = sqrtf(rw.var_ref * var_s);
  return fmaxf(0.0f, fminf(2.0f, 1.0f - cov / den));
}

// The exact FP32 cost engine: the reference's arithmetic, bit for bit with the
// FP32 oracle.
struct ExactEngine {
  static constexpr int kLanes = 1;
  using Win = RefWin;
  __device__ __forceinline__ static void build(const RunConsts& C, int px, int py, Win& w) { build_refwin(C, px, py, w); }
  __device__ __forceinline__ static float ncc(const RunConsts& C, int v, int px, int py, float4 pl, const Win& w) {
    return ncc_exact(C, v, px, py, pl, w);
  }
};

// ------------------------------------------------------------------ random planes
__device__ __forceinline__ float4 random_normal(const DevCam& cam, int px, int py, Philox& rng, float depth) {  // HPM.cu:159-185
  float q1 = 1.0f, q2 = 1.0f, s = 2.0f;
  while (s >= 1.0f) {
    q1 = 2.0f * rng.uniform() - 1.0f;
    q2 = 2.0f * rng.uniform() - 1.0f;
    s = q1 * q1 + q2 * q2;
  }
  const float sq = sqrtf(1.0f - s);
  float4 n = make_float4(2.0f * q1 * sq, 2.0f * q2 * sq, 1.0f - 2.0f * s, 0.0f);
  const float4 vd = view_dir(cam, px, py, depth);
  const float dot = n.x * vd.x + n.y * vd.y + n.z * vd.z;
  if (dot > 0.0f) { n.x = -n.x; n.y = -n.y; n.z = -n.z; }
  normalize3(n);
  return n;
}

__device__ __forceinline__ float4 perturbed_normal(const DevCam& cam, int px, int py, float4 normal, Philox& rng, float pert) {  // HPM.cu:187-222
  const float4 vd = view_dir(cam, px, py, 1.0f);
  const float a1 = (rng.uniform() - 0.5f) * pert;
  const float a2 = (rng.uniform() - 0.5f) * pert;
  const float a3 = (rng.uniform() - 0.5f) * pert;
  const float s1 = sinf_hpm(a1), s2 = sinf_hpm(a2), s3 = sinf_hpm(a3);
  const float c1 = cosf_hpm(a1), c2 = cosf_hpm(a2), c3 = cosf_hpm(a3);
  float R[9];
  R[0] = c2 * c3;
  R[1] = c3 * s1 * s2 - c1 * s3;
  R[2] = s1 * s3 + c1 * c3 * s2;
  R[3] = c2 * s3;
  R[4] = c1 * c3 + s1 * s2 * s3;
  R[5] = c1 * s2 * s3 - c3 * s1;
  R[6] = -s2;
  R[7] = c2 * s1;
  R[8] = c1 * c2;
  float4 np = make_float4(R[0] * normal.x + R[1] * normal.y + R[2] * normal.z,
                          R[3] * normal.x + R[4] * normal.y + R[5] * normal.z,
                          R[6] * normal.x + R[7] * normal.y + R[8] * normal.z, normal.w);
  if (np.x * vd.x + np.y * vd.y + np.z * vd.z >= 0.0f) np = normal;
  normalize3(np);
  return np;
}

// ------------------------------------------------------------------ geometric cost
__device__ __forceinline__ float3 to_world_point(float x, float y, float depth, const DevCam& cam) {  // HPM.cu:501-525
  const float* K = cam.K; const float* R = cam.R; const float* t = cam.t;
  const float X0 = depth * (x - K[2]) / K[0];
  const float X1 = depth * (y - K[5]) / K[4];
  const float X2 = depth;
  const float tx = R[0] * X0 + R[3] * X1 + R[6] * X2;
  const float ty = R[1] * X0 + R[4] * X1 + R[7] * X2;
  const float tz = R[2] * X0 + R[5] * X1 + R[8] * X2;
  const float Cx = -(R[0] * t[0] + R[3] * t[1] + R[6] * t[2]);
  const float Cy = -(R[1] * t[0] + R[4] * t[1] + R[7] * t[2]);
  const float Cz = -(R[2] * t[0] + R[5] * t[1] + R[8] * t[2]);
  return make_float3(tx + Cx, ty + Cy, tz + Cz);
}
__device__ __forceinline__ void project(float3 X, const DevCam& cam, float& u, float& v, float& depth) {  // HPM.cu:527-537
  const float* K = cam.K; const float* R = cam.R; const float* t = cam.t;
  const float tx = R[0] * X.x + R[1] * X.y + R[2] * X.z + t[0];
  const float ty = R[3] * X.x + R[4] * X.y + R[5] * X.z + t[1];
  const float tz = R[6] * X.x + R[7] * X.y + R[8] * X.z + t[2];
  depth = K[6] * tx + K[7] * ty + K[8] * tz;
  u = (K[0] * tx + K[1] * ty + K[2] * tz) / depth;
  v = (K[3] * tx + K[4] * ty + K[5] * tz) / depth;
}
__device__ __forceinline__ int trunc_texel(float v, int n) {  // (int)v of HPM.cu:549, defined off-image
  if (!(v > -1.0f)) return 0;
  if (v >= (float)n) return n - 1;
  return (int)v;
}
__device__ __forceinline__ float geom_cost(const RunConsts& C, int v, float4 pl, int px, int py) {  // HPM.cu:539-564
  const DevCam& rc = C.cams[0];
  const DevCam& sc = C.cams[v];
  const float depth = depth_from_plane(rc, pl, px, py);
  const float3 fwd = to_world_point((float)px, (float)py, depth, rc);
  float su, sv, sd;
  project(fwd, sc, su, sv, sd);
  const float src_depth = tex_at(C.dimg[v], sc.width, sc.height, trunc_texel(su, sc.width), trunc_texel(sv, sc.height));
  if (src_depth == 0.0f) return 3.0f;
  const float3 s3 = to_world_point(su, sv, src_depth, sc);
  float bu, bv, bd;
  project(s3, rc, bu, bv, bd);
  const float dx = (float)px - bu, dy = (float)py - bv;
  return fminf(3.0f, sqrtf(dx * dx + dy * dy));
}

// ------------------------------------------------------------------ misc
__device__ __forceinline__ float texture_factor(const RunConsts& C, const uint32_t* canny, int px, int py) {  // HPM.cu:349-373
  const int W = C.W, H = C.H;
  int total = 0, edges = 0;
#pragma unroll
  for (int i = -5; i < 6; i += 2)
#pragma unroll
    for (int j = -5; j < 6; j += 2) {
      const int x = px + j, y = py + i;
      if (x >= 0 && x < W && y >= 0 && y < H) {
        ++total;
        if (canny[y * W + x] >= 1u) ++edges;
      }
    }
  const float e = (float)edges * 1.0f / (float)total;
  return (float)(((double)e + 0.00005) / ((double)e + 0.0001));
}

__device__ __forceinline__ float norm_diff(float4 a, float4 b) {  // HPM.cu:497-499
  return fabsf(a.x - b.x) + fabsf(a.y - b.y) + fabsf(a.z - b.z);
}
__device__ __forceinline__ float prior_term(float dd, float nd, float texture, float two_dss, float two_ass) {  // HPM.cu:824
  const double a = (double)0.5f * (1.2 - 0.2 * (double)texture);
  const float b = expf_hpm(-dd * dd / two_dss) * expf_hpm(-nd * nd / two_ass);
  return (float)(a + (double)b);
}
__device__ __forceinline__ float restricted_cost(float cost, float beta, float texture, float prior) {  // HPM.cu:825
  const float q = -cost * cost / beta;
  return (float)(exp_hpm((double)q * (1 + 0.2 * (double)texture)) * (double)prior);
}

}  // namespace hpm
