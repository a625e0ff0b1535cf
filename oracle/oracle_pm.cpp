// oracle_pm.cpp -- CPU restatement of HPM-MVS++'s PatchMatch depth/normal path.
//
// TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg are the only users.  The product never includes, links or calls
// anything under oracle/.
//
// Each function follows the reference function cited beside it (HPM.cu,
// HPM.cpp).  It is templated on the arithmetic type T:
//   T = float  : the FP32 oracle.  Operations happen in the order the reference
//                writes them, with fused multiply-adds at exactly these sites
//                (written FMA(...) below): the homography projection
//                fma(H0, x, fma(H1, y, H2)), the two bilinear lerps and the
//                vertical lerp, the per-tap weighted accumulations, and the
//                variance/covariance differences; everything else is a plain
//                IEEE operation (build with -ffp-contract=off), division and
//                square root are IEEE, and exp/sin/cos are the oracle's own
//                (oracle_math.h).  The GPU's exact FP32 path performs the same
//                operations, so the two agree bit for bit.
//   T = double : the FP64 oracle (error reference).  Same algorithm, same random
//                draws, libm transcendentals.
// Resolutions of reference undefined behaviour and quirks:
//   Q1  plane_hypotheses_now starts as the pixel's current plane.
//   Q3  invalid-direction cost rows keep the reference's aggregate-initialiser
//       contents (row 0 starts {2, 0, 0, ...}, other rows 0).
//   Q6  hierarchy reload leaves xyz = 0 (the zero pages a fresh new[] returns).
//   Q7  bilinear sampling clamps to the border and is not quantised.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "oracle_math.h"

extern "C" {
// Mirrors Camera (main.h:39-47).
struct orc_camera {
  float K[9];
  float R[9];
  float t[3];
  int height;
  int width;
  float depth_min;
  float depth_max;
};

// Live fields of PatchMatchParams (HPM.h:32-55) plus the per-pass mode flags.
struct orc_params {
  int patch_size;        // 11
  int radius_increment;  // 2
  float sigma_spatial;   // 5
  float sigma_color;     // 3
  int top_k;             // 4
  int num_images;        // 1 + number of source views
  float depth_min;       // camera range * 0.6 (HPM.cpp:685)
  float depth_max;       // camera range * 1.2 (HPM.cpp:686)
  float scaled_cols;     // coarse map size in hierarchy/upsample mode (HPM.cpp:871-872)
  float scaled_rows;
  int geom_consistency;
  int prior_consistency;
  int hierarchy;
  int upsample;
  int mand_consistency;
  int num_iterations;    // 3 (HPM.cu:3013)
};

// Test probes (orc_probe_pixel): a cost callback that replaces the bilateral
// NCC or the geometric-consistency cost of one (view, plane), and what one
// pixel's update computed along the way.
typedef float (*orc_cost_fn)(void* user, int view, int px, int py, const float* plane);
struct orc_probe_out {
  int draws_used;          // uniforms the update consumed
  int positions[8];        // neighbour search result per direction (pixel index)
  int flags[8];            // direction valid
  float rows[8][32];       // per-direction cost rows as the update reads them
  int n_costvec;           // cost-vector evaluations of the neighbour search and extension
  int costvec_pos[16];     // ... and the pixel whose plane each one evaluated
  float pdf[32];           // view-selection probabilities times the neighbours' prior, before the CDF
  float weights[32];       // view weights (draw counts)
  uint32_t sel;            // view mask of the draws
  float weight_norm;
  float fc[8];             // weighted direction costs
  float rfc[8];            // restricted costs (prior / mandatory-consistency updates)
  int n_cand;              // refinement candidates evaluated
  float cand[5][4];        // their planes (normal, d)
  float cand_depth[5];     // the depth each candidate was built from
  float cand_cost[5];      // their aggregated costs
};
}

namespace {

template <class T> struct V4 { T x, y, z, w; };
template <class T> struct V2 { T x, y; };

inline float FMA(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double FMA(double a, double b, double c) { return std::fma(a, b, c); }
inline float EXP(float x) { return orc::expf_hpm(x); }
inline double EXP(double x) { return std::exp(x); }
inline float SIN(float x) { return orc::sinf_hpm(x); }
inline double SIN(double x) { return std::sin(x); }
inline float COS(float x) { return orc::cosf_hpm(x); }
inline double COS(double x) { return std::cos(x); }
// exp of an expression the reference evaluates in double (Q10): own double exp
// for the FP32 oracle (the GPU evaluates the same function), libm for FP64.
template <class T> inline double DEXP(double x);
template <> inline double DEXP<float>(double x) { return orc::exp_hpm(x); }
template <> inline double DEXP<double>(double x) { return std::exp(x); }
// fminf/fmaxf semantics (a NaN operand yields the other operand).
template <class T> inline T MINF(T a, T b) { return (a != a) ? b : ((b != b) ? a : (b < a ? b : a)); }
template <class T> inline T MAXF(T a, T b) { return (a != a) ? b : ((b != b) ? a : (a < b ? b : a)); }

// ---------------------------------------------------------------- images
struct Image {
  const float* data;
  int w, h;
  // Texel fetch at an integer position; out-of-range positions clamp to the
  // border (cudaAddressModeWrap with unnormalised coordinates, HPM.cpp:760).
  float at(int x, int y) const {
    x = x < 0 ? 0 : (x >= w ? w - 1 : x);
    y = y < 0 ? 0 : (y >= h ? h - 1 : y);
    return data[(size_t)y * w + x];
  }
};

// Bilinear sample at continuous pixel position (u, v), texel centres at
// integers: tex2D(tex, u + 0.5f, v + 0.5f) of HPM.cu:409 without the fixed-point
// fraction of the texture unit (Q7).
template <class T> T Bilinear(const Image& im, T u, T v) {
  const T fu = std::floor(u), fv = std::floor(v);
  const T a = u - fu, b = v - fv;
  const T cu = MINF(MAXF(fu, (T)-1), (T)im.w);  // keeps the int conversion defined
  const T cv = MINF(MAXF(fv, (T)-1), (T)im.h);
  const int x0 = (int)cu, y0 = (int)cv;
  const T i00 = (T)im.at(x0, y0), i10 = (T)im.at(x0 + 1, y0);
  const T i01 = (T)im.at(x0, y0 + 1), i11 = (T)im.at(x0 + 1, y0 + 1);
  const T top = FMA(a, i10 - i00, i00);
  const T bot = FMA(a, i11 - i01, i01);
  return FMA(b, bot - top, top);
}

// ---------------------------------------------------------------- camera math
template <class T> void Get3DPoint(const orc_camera& c, int px, int py, T depth, T X[3]) {  // HPM.cu:114-119
  X[0] = depth * ((T)px - (T)c.K[2]) / (T)c.K[0];
  X[1] = depth * ((T)py - (T)c.K[5]) / (T)c.K[4];
  X[2] = depth;
}
template <class T> V4<T> GetViewDirection(const orc_camera& c, int px, int py, T depth) {  // HPM.cu:121-133
  T X[3];
  Get3DPoint(c, px, py, depth, X);
  const T norm = std::sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
  return V4<T>{X[0] / norm, X[1] / norm, X[2] / norm, (T)0};
}
template <class T> T GetDistance2Origin(const orc_camera& c, int px, int py, T depth, const V4<T>& n) {  // HPM.cu:135-140
  T X[3];
  Get3DPoint(c, px, py, depth, X);
  return -(n.x * X[0] + n.y * X[1] + n.z * X[2]);
}
template <class T> T ComputeDepthfromPlaneHypothesis(const orc_camera& c, const V4<T>& pl, int px, int py) {  // HPM.cu:154-157
  const T K0 = c.K[0], K2 = c.K[2], K4 = c.K[4], K5 = c.K[5];
  return -pl.w * K0 / (((T)px - K2) * pl.x + (K0 / K4) * ((T)py - K5) * pl.y + K0 * pl.z);
}
template <class T> void NormalizeVec3(V4<T>* v) {  // HPM.cu:89-96 (rsqrtf -> 1/sqrt, both IEEE)
  const T ns = v->x * v->x + v->y * v->y + v->z * v->z;
  const T inv = (T)1 / std::sqrt(ns);
  v->x *= inv; v->y *= inv; v->z *= inv;
}
template <class T> V4<T> TransformNormal(const orc_camera& c, const V4<T>& p) {  // HPM.cu:322-330
  const float* R = c.R;
  return V4<T>{R[0] * p.x + R[3] * p.y + R[6] * p.z, R[1] * p.x + R[4] * p.y + R[7] * p.z,
               R[2] * p.x + R[5] * p.y + R[8] * p.z, p.w};
}
template <class T> V4<T> TransformNormal2RefCam(const orc_camera& c, const V4<T>& p) {  // HPM.cu:332-340
  const float* R = c.R;
  return V4<T>{R[0] * p.x + R[1] * p.y + R[2] * p.z, R[3] * p.x + R[4] * p.y + R[5] * p.z,
               R[6] * p.x + R[7] * p.y + R[8] * p.z, p.w};
}

// Plane-induced homography, HPM.cu:251-311.  The camera-centre/relative-pose
// part depends only on the view pair; it is recomputed here per call like the
// reference does (same operations, same result every call).
template <class T> void ComputeHomography(const orc_camera& rc, const orc_camera& sc, const V4<T>& pl, T H[9]) {
  const T* dummy = nullptr; (void)dummy;
  T rR[9], sR[9], rt[3], st[3];
  for (int i = 0; i < 9; ++i) { rR[i] = rc.R[i]; sR[i] = sc.R[i]; }
  for (int i = 0; i < 3; ++i) { rt[i] = rc.t[i]; st[i] = sc.t[i]; }
  T ref_C[3], src_C[3];
  for (int r = 0; r < 3; ++r) {
    ref_C[r] = -(rR[0 + r] * rt[0] + rR[3 + r] * rt[1] + rR[6 + r] * rt[2]);
    src_C[r] = -(sR[0 + r] * st[0] + sR[3 + r] * st[1] + sR[6 + r] * st[2]);
  }
  T R_rel[9], C_rel[3], t_rel[3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      R_rel[3 * a + b] = sR[3 * a + 0] * rR[3 * b + 0] + sR[3 * a + 1] * rR[3 * b + 1] + sR[3 * a + 2] * rR[3 * b + 2];
  for (int r = 0; r < 3; ++r) C_rel[r] = ref_C[r] - src_C[r];
  for (int a = 0; a < 3; ++a) t_rel[a] = sR[3 * a + 0] * C_rel[0] + sR[3 * a + 1] * C_rel[1] + sR[3 * a + 2] * C_rel[2];
  const T n[3] = {pl.x, pl.y, pl.z};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) H[3 * a + b] = R_rel[3 * a + b] - t_rel[a] * n[b] / pl.w;
  const T rK0 = rc.K[0], rK2 = rc.K[2], rK4 = rc.K[4], rK5 = rc.K[5];
  T tmp[9];
  for (int a = 0; a < 3; ++a) {
    tmp[3 * a + 0] = H[3 * a + 0] / rK0;
    tmp[3 * a + 1] = H[3 * a + 1] / rK4;
    tmp[3 * a + 2] = -H[3 * a + 0] * rK2 / rK0 - H[3 * a + 1] * rK5 / rK4 + H[3 * a + 2];
  }
  const T sK0 = sc.K[0], sK2 = sc.K[2], sK4 = sc.K[4], sK5 = sc.K[5], sK8 = sc.K[8];
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

template <class T> V2<T> ComputeCorrespondingPoint(const T H[9], int px, int py) {  // HPM.cu:313-320
  const T x = (T)px, y = (T)py;
  const T X = FMA(H[0], x, FMA(H[1], y, H[2]));
  const T Y = FMA(H[3], x, FMA(H[4], y, H[5]));
  const T Z = FMA(H[6], x, FMA(H[7], y, H[8]));
  return V2<T>{X / Z, Y / Z};
}

template <class T> T ComputeBilateralWeight(T xd, T yd, T pix, T cpix, T ss, T sc) {  // HPM.cu:342-347
  const T spatial = std::sqrt(xd * xd + yd * yd);
  const T color = std::fabs(pix - cpix);
  return EXP(-spatial / ((T)2 * ss * ss) - color / ((T)2 * sc * sc));
}

// ---------------------------------------------------------------- probes
// Injected into one pixel's update by orc_probe_pixel; a normal run has none.
struct Probe {
  const float* draws = nullptr;  // the uniforms the update draws, in order
  int n_draws = 0;
  orc_cost_fn ncc = nullptr, geom = nullptr;
  void* user = nullptr;
  orc_probe_out* out = nullptr;
};

// The pixel's uniform stream: Philox (orc::PixelRng), or a probe's fixed list.
struct Draws {
  orc::PixelRng prng;
  Probe* probe;
  Draws(uint64_t seed, uint32_t pixel, uint32_t stage, Probe* p) : prng(seed, pixel, stage), probe(p) {}
  float uniform() {
    if (probe && probe->draws) {
      const int k = probe->out->draws_used++;
      return k < probe->n_draws ? probe->draws[k] : 0.5f;  // overrun: the caller sees draws_used > n_draws
    }
    return prng.uniform();
  }
};

// ---------------------------------------------------------------- context
template <class T> struct Ctx {
  int W = 0, H = 0, nv = 0;           // nv = num_images (ref + sources)
  const orc_camera* cams = nullptr;
  std::vector<Image> imgs;            // nv images
  std::vector<Image> dimgs;           // nv depth maps (geom mode), index 0 unused
  orc_params prm{};
  uint64_t seed = 0;
  V4<T>* planes = nullptr;
  T* costs = nullptr;
  T* pre_costs = nullptr;
  uint32_t* sel = nullptr;
  const V4<T>* prior = nullptr;
  const uint32_t* masks = nullptr;
  T* conf = nullptr;
  const uint32_t* canny = nullptr;
  T* texture = nullptr;
  const V4<T>* scaled = nullptr;
  float judge_thr[3][3] = {};         // JudgeExtend good threshold per (iter, ext round) (HPM.cu:1326-1327)
  T vs_thr[3] = {};                   // view-selection cost threshold per iter (HPM.cu:1706)
  T vs_fallback[3] = {};              // expf(thr^2 / -0.32) per iter (HPM.cu:1724)
  int row_begin = 0, row_end = 0;     // pixel rows updated (CPU-baseline sampling)
  Probe* probe = nullptr;             // test probe of a single pixel (orc_probe_pixel)
};

template <class T> float* ProbePlane(const V4<T>& pl, float out[4]) {
  out[0] = (float)pl.x; out[1] = (float)pl.y; out[2] = (float)pl.z; out[3] = (float)pl.w;
  return out;
}

template <class T> T ComputeTexture(const Ctx<T>& c, int px, int py) {  // HPM.cu:349-373
  int total = 0, edges = 0;
  for (int i = -5; i < 6; i += 2)
    for (int j = -5; j < 6; j += 2) {
      const int x = px + j, y = py + i;
      if (x >= 0 && x < c.W && y >= 0 && y < c.H) {
        ++total;
        if (c.canny[(size_t)y * c.W + x] >= 1) ++edges;
      }
    }
  const T e = (T)edges * (T)1 / (T)total;
  return (T)(((double)e + 0.00005) / ((double)e + 0.0001));
}

// HPM.cu:375-448.  Outer loop i = x offset, inner j = y offset, per-column
// partial sums added into the totals (HPM.cu:397-427).
template <class T> T ComputeBilateralNCC(const Ctx<T>& c, int v, int px, int py, const V4<T>& pl) {
  if (c.probe && c.probe->ncc) {
    float q[4];
    return (T)c.probe->ncc(c.probe->user, v, px, py, ProbePlane(pl, q));
  }
  const T cost_max = 2;
  const int radius = c.prm.patch_size / 2;
  const orc_camera& rc = c.cams[0];
  const orc_camera& sc = c.cams[v];
  T H[9];
  ComputeHomography(rc, sc, pl, H);
  const V2<T> pt = ComputeCorrespondingPoint(H, px, py);
  if (pt.x >= (T)sc.width || pt.x < (T)0 || pt.y >= (T)sc.height || pt.y < (T)0) return cost_max;
  const Image& ref = c.imgs[0];
  const Image& src = c.imgs[v];
  const T ss = c.prm.sigma_spatial, scol = c.prm.sigma_color;
  T sum_ref = 0, sum_ref_ref = 0, sum_src = 0, sum_src_src = 0, sum_ref_src = 0, wsum = 0;
  const T ref_center = (T)ref.at(px, py);
  for (int i = -radius; i < radius + 1; i += c.prm.radius_increment) {
    T r_ref = 0, r_src = 0, r_rr = 0, r_ss = 0, r_rs = 0, r_w = 0;
    for (int j = -radius; j < radius + 1; j += c.prm.radius_increment) {
      const int qx = px + i, qy = py + j;
      const T ref_pix = (T)ref.at(qx, qy);
      const V2<T> sp = ComputeCorrespondingPoint(H, qx, qy);
      const T src_pix = Bilinear(src, sp.x, sp.y);
      const T w = ComputeBilateralWeight((T)i, (T)j, ref_pix, ref_center, ss, scol);
      const T wr = w * ref_pix, ws = w * src_pix;
      r_ref += wr;
      r_rr = FMA(wr, ref_pix, r_rr);
      r_src += ws;
      r_ss = FMA(ws, src_pix, r_ss);
      r_rs = FMA(wr, src_pix, r_rs);
      r_w += w;
    }
    sum_ref += r_ref; sum_ref_ref += r_rr; sum_src += r_src;
    sum_src_src += r_ss; sum_ref_src += r_rs; wsum += r_w;
  }
  const T inv = (T)1 / wsum;
  sum_ref *= inv; sum_ref_ref *= inv; sum_src *= inv; sum_src_src *= inv; sum_ref_src *= inv;
  const T var_ref = FMA(-sum_ref, sum_ref, sum_ref_ref);
  const T var_src = FMA(-sum_src, sum_src, sum_src_src);
  const T kMinVar = (T)1e-5f;
  if (var_ref < kMinVar || var_src < kMinVar) return cost_max;
  const T covar = FMA(-sum_ref, sum_src, sum_ref_src);
  const T den = std::sqrt(var_ref * var_src);
  return MAXF((T)0, MINF(cost_max, (T)1 - covar / den));
}

template <class T> void InsertionSort(T* d, int n) {  // HPM.cu:15-24
  for (int i = 1; i < n; i++) {
    const T tmp = d[i];
    int j;
    for (j = i; j >= 1 && tmp < d[j - 1]; j--) d[j] = d[j - 1];
    d[j] = tmp;
  }
}

// HPM.cu:450-488
template <class T> T ComputeMultiViewInitialCostandSelectedViews(const Ctx<T>& c, int px, int py, const V4<T>& pl, uint32_t* selected) {
  const T cost_max = 2;
  T cv[32], cvc[32];
  int count = 0, valid = 0;
  for (int i = 1; i < c.nv; ++i) {
    const T x = ComputeBilateralNCC(c, i, px, py, pl);
    cv[i - 1] = x; cvc[i - 1] = x; ++count;
    if (x < cost_max) ++valid;
  }
  InsertionSort(cv, count);
  *selected = 0;
  const int top_k = std::min(valid, c.prm.top_k);
  if (top_k > 0) {
    T cost = 0;
    for (int i = 0; i < top_k; ++i) cost += cv[i];
    const T thr = cv[top_k - 1];
    for (int i = 0; i < c.nv - 1; ++i)
      if (cvc[i] <= thr) *selected |= (1u << i);
    return cost / (T)top_k;
  }
  return cost_max;
}

template <class T> void ComputeMultiViewCostVector(const Ctx<T>& c, int px, int py, const V4<T>& pl, T* out) {  // HPM.cu:490-495
  for (int i = 1; i < c.nv; ++i) out[i - 1] = ComputeBilateralNCC(c, i, px, py, pl);
}

// ---------------------------------------------------------------- random planes
template <class T> V4<T> GenerateRandomNormal(const orc_camera& cam, int px, int py, Draws& rng, T depth) {  // HPM.cu:159-185
  T q1 = 1, q2 = 1, s = 2;
  while (s >= (T)1) {
    q1 = (T)2 * (T)rng.uniform() - (T)1;
    q2 = (T)2 * (T)rng.uniform() - (T)1;
    s = q1 * q1 + q2 * q2;
  }
  const T sq = std::sqrt((T)1 - s);
  V4<T> n{(T)2 * q1 * sq, (T)2 * q2 * sq, (T)1 - (T)2 * s, (T)0};
  const V4<T> vd = GetViewDirection(cam, px, py, depth);
  const T dot = n.x * vd.x + n.y * vd.y + n.z * vd.z;
  if (dot > (T)0) { n.x = -n.x; n.y = -n.y; n.z = -n.z; }
  NormalizeVec3(&n);
  return n;
}

template <class T> V4<T> GeneratePerturbedNormal(const orc_camera& cam, int px, int py, const V4<T>& normal, Draws& rng, T pert) {  // HPM.cu:187-222
  const V4<T> vd = GetViewDirection(cam, px, py, (T)1);
  const T a1 = ((T)rng.uniform() - (T)0.5f) * pert;
  const T a2 = ((T)rng.uniform() - (T)0.5f) * pert;
  const T a3 = ((T)rng.uniform() - (T)0.5f) * pert;
  const T s1 = SIN(a1), s2 = SIN(a2), s3 = SIN(a3);
  const T c1 = COS(a1), c2 = COS(a2), c3 = COS(a3);
  T R[9];
  R[0] = c2 * c3;
  R[1] = c3 * s1 * s2 - c1 * s3;
  R[2] = s1 * s3 + c1 * c3 * s2;
  R[3] = c2 * s3;
  R[4] = c1 * c3 + s1 * s2 * s3;
  R[5] = c1 * s2 * s3 - c3 * s1;
  R[6] = -s2;
  R[7] = c2 * s1;
  R[8] = c1 * c2;
  V4<T> np{R[0] * normal.x + R[1] * normal.y + R[2] * normal.z, R[3] * normal.x + R[4] * normal.y + R[5] * normal.z,
           R[6] * normal.x + R[7] * normal.y + R[8] * normal.z, normal.w};
  if (np.x * vd.x + np.y * vd.y + np.z * vd.z >= (T)0) np = normal;
  NormalizeVec3(&np);
  return np;
}

template <class T> V4<T> GenerateRandomPlaneHypothesis(const orc_camera& cam, int px, int py, Draws& rng, T dmin, T dmax) {  // HPM.cu:224-230
  const T depth = (T)rng.uniform() * (dmax - dmin) + dmin;
  V4<T> pl = GenerateRandomNormal(cam, px, py, rng, depth);
  pl.w = GetDistance2Origin(cam, px, py, depth, pl);
  return pl;
}

// ---------------------------------------------------------------- geometric cost
template <class T> V4<T> Get3DPointonWorld_cu(T x, T y, T depth, const orc_camera& cam) {  // HPM.cu:501-525
  const float* K = cam.K; const float* R = cam.R; const float* t = cam.t;
  const T X0 = depth * (x - (T)K[2]) / (T)K[0];
  const T X1 = depth * (y - (T)K[5]) / (T)K[4];
  const T X2 = depth;
  const T tx = R[0] * X0 + R[3] * X1 + R[6] * X2;
  const T ty = R[1] * X0 + R[4] * X1 + R[7] * X2;
  const T tz = R[2] * X0 + R[5] * X1 + R[8] * X2;
  const T Cx = -((T)R[0] * t[0] + (T)R[3] * t[1] + (T)R[6] * t[2]);
  const T Cy = -((T)R[1] * t[0] + (T)R[4] * t[1] + (T)R[7] * t[2]);
  const T Cz = -((T)R[2] * t[0] + (T)R[5] * t[1] + (T)R[8] * t[2]);
  return V4<T>{tx + Cx, ty + Cy, tz + Cz, (T)0};
}
template <class T> void ProjectonCamera_cu(const V4<T>& X, const orc_camera& cam, V2<T>* pt, T* depth) {  // HPM.cu:527-537
  const float* K = cam.K; const float* R = cam.R; const float* t = cam.t;
  const T tx = R[0] * X.x + R[1] * X.y + R[2] * X.z + t[0];
  const T ty = R[3] * X.x + R[4] * X.y + R[5] * X.z + t[1];
  const T tz = R[6] * X.x + R[7] * X.y + R[8] * X.z + t[2];
  *depth = K[6] * tx + K[7] * ty + K[8] * tz;
  pt->x = (K[0] * tx + K[1] * ty + K[2] * tz) / *depth;
  pt->y = (K[3] * tx + K[4] * ty + K[5] * tz) / *depth;
}
inline int TruncToTexel(double v, int n) {  // (int)v of HPM.cu:549 with a defined result off-image
  if (!(v > -1.0)) return 0;       // also NaN
  if (v >= (double)n) return n - 1;
  return (int)v;
}
template <class T> T ComputeGeomConsistencyCost(const Ctx<T>& c, int v, const V4<T>& pl, int px, int py) {  // HPM.cu:539-564
  if (c.probe && c.probe->geom) {
    float q[4];
    return (T)c.probe->geom(c.probe->user, v, px, py, ProbePlane(pl, q));
  }
  const T max_cost = 3;
  const orc_camera& rc = c.cams[0];
  const orc_camera& sc = c.cams[v];
  const T depth = ComputeDepthfromPlaneHypothesis(rc, pl, px, py);
  const V4<T> fwd = Get3DPointonWorld_cu((T)px, (T)py, depth, rc);
  V2<T> sp; T sd;
  ProjectonCamera_cu(fwd, sc, &sp, &sd);
  const Image& dimg = c.dimgs[v];
  const T src_depth = (T)dimg.at(TruncToTexel((double)sp.x, dimg.w), TruncToTexel((double)sp.y, dimg.h));
  if (src_depth == (T)0) return max_cost;
  const V4<T> s3 = Get3DPointonWorld_cu(sp.x, sp.y, src_depth, sc);
  V2<T> bp; T rd;
  ProjectonCamera_cu(s3, rc, &bp, &rd);
  const T dx = (T)px - bp.x, dy = (T)py - bp.y;
  return MINF(max_cost, std::sqrt(dx * dx + dy * dy));
}

// ---------------------------------------------------------------- initialisation
// JBU of the coarse normal/cost map inside RandomInitialization, HPM.cu:633-698.
template <class T> void UpsampleInit(Ctx<T>& c, int px, int py) {
  const size_t center = (size_t)py * c.W + px;
  const float sc_cols = c.prm.scaled_cols, sc_rows = c.prm.scaled_rows;
  const float scale = (float)(1.0 * sc_cols / c.W);
  const float sigmad = 0.50f, sigmar = 25.5f;
  const int Imagescale = (int)std::max((float)c.W / sc_cols, (float)c.H / sc_rows);
  const int WinWidth = Imagescale * Imagescale + 1;
  const int nn = WinWidth / 2;
  const float o_y = (float)py * scale, o_x = (float)px * scale;
  const Image& ref = c.imgs[0];
  const T refPix = (T)ref.at(px, py);
  T nf = 0, ctot = 0;
  V4<T> ntot{0, 0, 0, 0};
  for (int j = -nn; j <= nn; ++j) {
    int r_y = (int)(o_y + (float)j);
    r_y = (r_y > 0 ? ((float)r_y < sc_rows ? r_y : (int)(sc_rows - 1)) : 0);
    const int r_ys = py + j;
    for (int i = -nn; i <= nn; ++i) {
      int r_x = (int)(o_x + (float)i);
      r_x = (r_x > 0 ? ((float)r_x < sc_cols ? r_x : (int)(sc_cols - 1)) : 0);
      const int s_center = (int)((float)r_y * sc_cols + (float)r_x);
      const V4<T> sn = c.scaled[s_center];
      const T srcPix = sn.w;
      const T nb = (T)ref.at(px + i, r_ys);
      // SpatialGauss / RangeGauss (HPM.cu:142-152): float square, double exp
      const float dxs = o_x - (float)r_x, dys = o_y - (float)r_y;
      const float dis = dxs * dxs + dys * dys - 0.0f;
      const T sg = (T)DEXP<T>(-1.0 * (double)dis / (double)(2 * sigmad * sigmad));
      const T xr = std::fabs(refPix - nb);
      const T rg = (T)DEXP<T>(-1.0 * (double)(xr * xr) / (double)(2 * sigmar * sigmar));
      const T tg = sg * rg;
      nf += tg;
      ctot += srcPix * tg;
      ntot.x = ntot.x + sn.x * tg;
      ntot.y = ntot.y + sn.y * tg;
      ntot.z = ntot.z + sn.z * tg;
    }
  }
  c.costs[center] = ctot / nf;
  ntot.x = ntot.x / nf; ntot.y = ntot.y / nf; ntot.z = ntot.z / nf;
  NormalizeVec3(&ntot);
  c.costs[center] = ComputeMultiViewInitialCostandSelectedViews(c, px, py, c.planes[center], &c.sel[center]);
  c.pre_costs[center] = c.costs[center];
  V4<T> pl = TransformNormal2RefCam(c.cams[0], V4<T>{ntot.x, ntot.y, ntot.z, ntot.w});
  const T depth = c.planes[center].w;
  pl.w = GetDistance2Origin(c.cams[0], px, py, depth, pl);
  c.planes[center] = pl;
  c.costs[center] = ComputeMultiViewInitialCostandSelectedViews(c, px, py, pl, &c.sel[center]);
}

template <class T> void ReloadPlane(Ctx<T>& c, int px, int py) {  // HPM.cu:749-760 (and the mand copies)
  const size_t center = (size_t)py * c.W + px;
  V4<T> pl = c.prm.hierarchy ? c.scaled[center] : c.planes[center];
  pl = TransformNormal2RefCam(c.cams[0], pl);
  const T depth = pl.w;
  pl.w = GetDistance2Origin(c.cams[0], px, py, depth, pl);
  c.planes[center] = pl;
  c.costs[center] = ComputeMultiViewInitialCostandSelectedViews(c, px, py, pl, &c.sel[center]);
}

template <class T> void RandomInitialization(Ctx<T>& c, int px, int py) {  // HPM.cu:567-763
  const size_t center = (size_t)py * c.W + px;
  Draws rng(c.seed, (uint32_t)center, 0u, c.probe);
  const orc_params& P = c.prm;
  const orc_camera& cam = c.cams[0];
  if (!P.prior_consistency && !P.mand_consistency) c.texture[center] = ComputeTexture(c, px, py);
  if (!P.geom_consistency && !P.hierarchy && !P.prior_consistency) {
    c.planes[center] = GenerateRandomPlaneHypothesis(cam, px, py, rng, (T)P.depth_min, (T)P.depth_max);
    c.costs[center] = ComputeMultiViewInitialCostandSelectedViews(c, px, py, c.planes[center], &c.sel[center]);
  } else if (P.prior_consistency) {
    if (c.masks[center] > 0) {
      if (c.conf[center] > (T)0.3) {
        V4<T> pl = c.planes[center];
        const T depth = pl.w;
        pl.w = GetDistance2Origin(cam, px, py, depth, pl);
        c.planes[center] = pl;
        c.costs[center] = (T)0.1;
      } else if (c.costs[center] > (T)0.1f) {
        const float perturbation = 0.02f;
        V4<T> pl = c.prior[center];
        T dp = pl.w;
        const T dminp = (T)(1 - 3 * perturbation) * dp;
        const T dmaxp = (T)(1 + 3 * perturbation) * dp;
        dp = (T)rng.uniform() * (dmaxp - dminp) + dminp;
        const T pert = (T)(float)(3 * perturbation * 3.14159265358979323846);
        V4<T> pp = GeneratePerturbedNormal(cam, px, py, pl, rng, pert);
        pp.w = dp;  // the reference stores the depth in w here (HPM.cu:609)
        c.planes[center] = pp;
        c.costs[center] = ComputeMultiViewInitialCostandSelectedViews(c, px, py, pp, &c.sel[center]);
      } else {
        V4<T> pl = c.planes[center];
        const T depth = pl.w;
        pl.w = GetDistance2Origin(cam, px, py, depth, pl);
        c.planes[center] = pl;
        c.costs[center] = ComputeMultiViewInitialCostandSelectedViews(c, px, py, pl, &c.sel[center]);
      }
    } else {
      V4<T> pl = c.planes[center];
      const T depth = pl.w;
      pl.w = GetDistance2Origin(cam, px, py, depth, pl);
      c.planes[center] = pl;
      if (c.conf[center] > (T)0.3) c.costs[center] = (T)0.1;
    }
  } else {
    if (P.upsample) {
      UpsampleInit(c, px, py);
    } else if (P.mand_consistency) {
      if (c.conf[center] >= (T)0.3) {
        ReloadPlane(c, px, py);
      } else {
        const T tex = ComputeTexture(c, px, py);
        if (tex > (T)0.5) c.conf[center] = (T)0.1;
        ReloadPlane(c, px, py);
      }
    } else {
      ReloadPlane(c, px, py);
    }
  }
}

// ---------------------------------------------------------------- refinement
template <class T> T NormDiff(const V4<T>& a, const V4<T>& b) {  // HPM.cu:497-499
  return std::fabs(a.x - b.x) + std::fabs(a.y - b.y) + std::fabs(a.z - b.z);
}

// prior term of HPM.cu:824 / 1808: gamma*(1.2-0.2t) in double, two float exps.
template <class T> T PriorTerm(T depth_diff, T norm1_diff, T texture, T two_dss, T two_ass) {
  const float gamma = 0.5f;
  const double a = (double)gamma * (1.2 - 0.2 * (double)texture);
  const T b = EXP(-depth_diff * depth_diff / two_dss) * EXP(-norm1_diff * norm1_diff / two_ass);
  return (T)(a + (double)b);
}
template <class T> T RestrictedCost(T cost, T beta, T texture, T prior) {  // HPM.cu:825 / 1809
  const T q = -cost * cost / beta;
  return (T)(DEXP<T>((double)q * (1 + 0.2 * (double)texture)) * (double)prior);
}

template <class T> struct Consts {
  T depth_sigma, two_dss, angle_sigma, two_ass, pert_refine, beta;
};
template <class T> Consts<T> MakeConsts(const orc_params& P) {  // HPM.cu:767-775
  Consts<T> k;
  k.depth_sigma = ((T)P.depth_max - (T)P.depth_min) / (T)64.0f;
  k.two_dss = (T)2 * k.depth_sigma * k.depth_sigma;
  k.angle_sigma = (T)(float)(3.14159265358979323846 * (5.0f / 180.0f));
  k.two_ass = (T)2 * k.angle_sigma * k.angle_sigma;
  k.pert_refine = (T)(float)(0.02f * 3.14159265358979323846);
  k.beta = (T)0.18f;
  return k;
}

// HPM.cu:765-841.  `views` limits the cost evaluation to the views the weights
// select; unselected views' costs are never read by the reference either.
template <class T> void PlaneHypothesisRefinement(const Ctx<T>& c, V4<T>* plane, T* depth, T* cost, Draws& rng,
                                                  const T* vw, T weight_norm, T* restricted_cost, int px, int py, T texture) {
  const orc_params& P = c.prm;
  const orc_camera& cam = c.cams[0];
  const size_t center = (size_t)py * c.W + px;
  const Consts<T> k = MakeConsts<T>(P);
  const float perturbation = 0.02f;
  T depth_prior = 0, depth_rand;
  V4<T> plane_rand;
  const bool use_prior = P.prior_consistency && c.masks[center] > 0;
  if (use_prior) {
    depth_prior = ComputeDepthfromPlaneHypothesis(cam, c.prior[center], px, py);
    depth_rand = (T)rng.uniform() * (T)6 * k.depth_sigma + (depth_prior - (T)3 * k.depth_sigma);
    plane_rand = GeneratePerturbedNormal(cam, px, py, c.prior[center], rng, k.angle_sigma);
  } else {
    depth_rand = (T)rng.uniform() * ((T)P.depth_max - (T)P.depth_min) + (T)P.depth_min;
    plane_rand = GenerateRandomNormal(cam, px, py, rng, *depth);
  }
  T depth_perturbed = *depth;
  const T dminp = (T)(1 - perturbation) * depth_perturbed;
  const T dmaxp = (T)(1 + perturbation) * depth_perturbed;
  depth_perturbed = (T)rng.uniform() * (dmaxp - dminp) + dminp;  // the do/while of HPM.cu:792-794 runs once
  const V4<T> plane_pert = GeneratePerturbedNormal(cam, px, py, *plane, rng, k.pert_refine);
  const T depths[5] = {depth_rand, *depth, depth_rand, *depth, depth_perturbed};
  const V4<T> normals[5] = {*plane, plane_rand, plane_rand, plane_pert, *plane};
  for (int i = 0; i < 5; ++i) {
    V4<T> tp = normals[i];
    tp.w = GetDistance2Origin(cam, px, py, depths[i], tp);
    T tc = 0;
    for (int j = 0; j < c.nv - 1; ++j) {
      if (vw[j] > (T)0) {
        const T cj = ComputeBilateralNCC(c, j + 1, px, py, tp);
        if (P.geom_consistency)
          tc += vw[j] * (cj + (T)0.2f * ComputeGeomConsistencyCost(c, j + 1, tp, px, py));
        else
          tc += vw[j] * cj;
      }
    }
    tc /= weight_norm;
    if (c.probe) {
      orc_probe_out* o = c.probe->out;
      ProbePlane(tp, o->cand[i]);
      o->cand_depth[i] = (float)depths[i];
      o->cand_cost[i] = (float)tc;
      o->n_cand = i + 1;
    }
    const T depth_before = ComputeDepthfromPlaneHypothesis(cam, tp, px, py);
    if (use_prior) {
      const T dd = depths[i] - depth_prior;
      const T nd = NormDiff(c.prior[center], tp);
      const T prior = PriorTerm(dd, nd, texture, k.two_dss, k.two_ass);
      const T rtc = RestrictedCost(tc, k.beta, texture, prior);
      if (depth_before >= (T)P.depth_min && depth_before <= (T)P.depth_max && rtc > *restricted_cost) {
        *depth = depth_before; *plane = tp; *cost = tc; *restricted_cost = rtc;
      }
    } else {
      if (depth_before >= (T)P.depth_min && depth_before <= (T)P.depth_max && tc < *cost) {
        *depth = depth_before; *plane = tp; *cost = tc;
      }
    }
  }
}

// ---------------------------------------------------------------- propagation
template <class T> struct Sampled {
  T cost_array[8][32];
  bool flag[8];
  int positions[8];
};

// Neighbour search, HPM.cu:1385-1608 (identical in both mandatory-consistency
// branches, HPM.cu:1905-2101 and 2323-2527), then left-up extension
// HPM.cu:1609-1680 with ExtendedLeftUpPropagation HPM.cu:1063-1126 and
// JudgeExtend HPM.cu:1322-1345.  Only direction 0 ever extends (Q2).
template <class T> void SampleNeighbours(const Ctx<T>& c, int px, int py, int iter, Sampled<T>* s) {
  const int W = c.W, Hh = c.H;
  const int center = py * W + px;
  const int nvs = c.nv - 1;
  for (int d = 0; d < 8; ++d) { for (int v = 0; v < 32; ++v) s->cost_array[d][v] = 0; s->flag[d] = false; }
  s->cost_array[0][0] = 2;  // `float cost_array[8][32] = { 2.0f }` (Q3)
  int left_up = center - 5 - 6 * W, right_up = center + 6 - 5 * W;
  int right_down = center + 5 + 6 * W, left_down = center - 6 + 5 * W;
  int left_far = center - 5, right_far = center + 5, up_far = center - W * 5, down_far = center + W * 5;
  const T* costs = c.costs;
  // ComputeMultiViewCostVector of the plane at pixel q, seen from this pixel
  auto costvec = [&](int q, T* out) {
    if (c.probe) {
      orc_probe_out* o = c.probe->out;
      if (o->n_costvec < 16) o->costvec_pos[o->n_costvec] = q;
      ++o->n_costvec;
    }
    ComputeMultiViewCostVector(c, px, py, c.planes[q], out);
  };
  if (py > 4) {
    s->flag[1] = true;
    T m = costs[up_far]; int mp = up_far;
    for (int i = 1; i < 5; ++i) {
      if (py > 4 + 2 * i) { const int pt = up_far - 2 * i * W; if (costs[pt] < m) { m = costs[pt]; mp = pt; } } else break;
    }
    up_far = mp;
    costvec(up_far, s->cost_array[1]);
  }
  if (py < Hh - 5) {
    s->flag[3] = true;
    T m = costs[down_far]; int mp = down_far;
    for (int i = 1; i < 5; ++i) {
      if (py < Hh - 5 - 2 * i) { const int pt = down_far + 2 * i * W; if (costs[pt] < m) { m = costs[pt]; mp = pt; } } else break;
    }
    down_far = mp;
    costvec(down_far, s->cost_array[3]);
  }
  if (px > 4) {
    s->flag[5] = true;
    T m = costs[left_far]; int mp = left_far;
    for (int i = 1; i < 5; ++i) {
      if (px > 4 + 2 * i) { const int pt = left_far - 2 * i; if (costs[pt] < m) { m = costs[pt]; mp = pt; } } else break;
    }
    left_far = mp;
    costvec(left_far, s->cost_array[5]);
  }
  if (px < W - 5) {
    s->flag[7] = true;
    T m = costs[right_far]; int mp = right_far;
    for (int i = 1; i < 5; ++i) {  // keeps the maximum (Q4, HPM.cu:1464)
      if (px < W - 5 - 2 * i) { const int pt = right_far + 2 * i; if (m < costs[pt]) { m = costs[pt]; mp = pt; } } else break;
    }
    right_far = mp;
    costvec(right_far, s->cost_array[7]);
  }
  // Diagonal zig-zag searches: start (sx, sy), even steps move along `ex`, odd along `oy`.
  auto diag = [&](int dir, bool ok, int start, int sx, int sy, bool even_moves_x, int dx, int dy, int* out) {
    if (!ok) return;
    s->flag[dir] = true;
    T m = costs[start]; int mp = start;
    int x = sx, y = sy;
    for (int i = 0; i < 7; i++) {
      const bool even = (i % 2 == 0);
      if (even == even_moves_x) x += dx; else y += dy;
      if (x >= 0 && x < W && y < Hh && y >= 0) {
        const int pt = y * W + x;
        if (costs[pt] < m) { m = costs[pt]; mp = pt; }
      } else break;
    }
    *out = mp;
    costvec(mp, s->cost_array[dir]);
  };
  diag(0, px > 4 && py > 5, left_up, px - 5, py - 6, true, -2, -2, &left_up);
  diag(2, px < W - 6 && py > 4, right_up, px + 6, py - 5, false, +2, -2, &right_up);
  diag(4, px < W - 5 && py < Hh - 6, right_down, px + 5, py + 6, true, +2, +2, &right_down);
  diag(6, px > 5 && py < Hh - 5, left_down, px - 6, py + 5, false, -2, +2, &left_down);
  int pos[8] = {left_up, up_far, right_up, down_far, right_down, left_far, left_down, right_far};
  // extension of direction 0 only
  bool active = true;
  for (int it = 0; it < 3; ++it) {
    if (!active) continue;
    bool extend = false;
    if (s->flag[0]) {
      const T good = (T)c.judge_thr[iter][it];
      const T bad = (T)1.2f;
      int gs = 0, bs = 0;
      for (int v = 0; v < nvs; v++) {
        if (s->cost_array[0][v] < good) gs++;
        if (s->cost_array[0][v] > bad) bs++;
      }
      extend = !(gs >= 1 && bs <= 2);
    }
    if (!extend) { active = false; continue; }
    // ExtendedLeftUpPropagation, HPM.cu:1063-1126
    const int lu = center - 5 - 6 * W - 8 * it - 8 * W * it;
    T tmp[32];
    int newpos;
    if (px > 4 + 8 * it && py > 5 + 8 * it) {
      bool symbol = true;
      T m = costs[lu]; int mp = lu;
      int x = px - 5 - 8 * it, y = py - 6 - 8 * it;
      for (int i = 0; i < 7; i++) {
        if (i % 2 == 0) x -= 2; else y -= 2;
        if (x >= 0 && x < W && y < Hh && y >= 0) {
          const int pt = y * W + x;
          if (costs[pt] < m) { m = costs[pt]; mp = pt; }
        } else break;
      }
      if (m > costs[pos[0]]) { m = costs[pos[0]]; mp = pos[0]; symbol = false; }
      newpos = mp;
      if (symbol) costvec(newpos, tmp);
      else for (int v = 0; v < nvs; ++v) tmp[v] = s->cost_array[0][v];
    } else {
      newpos = pos[0];
      for (int v = 0; v < nvs; ++v) tmp[v] = s->cost_array[0][v];
    }
    pos[0] = newpos;
    for (int v = 0; v < nvs; ++v) s->cost_array[0][v] = tmp[v];
  }
  for (int d = 0; d < 8; ++d) s->positions[d] = pos[d];
  if (c.probe) {
    orc_probe_out* o = c.probe->out;
    for (int d = 0; d < 8; ++d) {
      o->positions[d] = pos[d];
      o->flags[d] = s->flag[d];
      for (int v = 0; v < 32; ++v) o->rows[d][v] = (float)s->cost_array[d][v];
    }
  }
}

// Multi-hypothesis joint view selection, HPM.cu:1686-1751.
template <class T> void ViewSelection(const Ctx<T>& c, int px, int py, int iter, const Sampled<T>& s, Draws& rng,
                                      T* vw, uint32_t* sel_out, T* weight_norm_out, int* nsel_out) {
  const int W = c.W;
  const int center = py * W + px;
  const int nvs = c.nv - 1;
  for (int v = 0; v < 32; ++v) vw[v] = 0;
  T prior[32] = {};
  const int nbr[4] = {center - W * 5, center + W * 5, center - 5, center + 5};
  for (int i = 0; i < 4; ++i) {
    if (s.flag[2 * i + 1]) {
      for (int j = 0; j < nvs; ++j) {
        if ((c.sel[nbr[i]] >> j) & 1u) prior[j] += (T)0.9f; else prior[j] += (T)0.1f;
      }
    }
  }
  T probs[32] = {};
  const T thr = c.vs_thr[iter];
  for (int i = 0; i < nvs; i++) {
    T count = 0; int count_false = 0; T tmpw = 0;
    for (int j = 0; j < 8; j++) {
      const T x = s.cost_array[j][i];
      if (x < thr) { tmpw += EXP(x * x / (T)(-0.18f)); count++; }
      if (x > (T)1.2f) count_false++;
    }
    if (count > 2 && count_false < 3) probs[i] = tmpw / count;
    else if (count_false < 3) probs[i] = c.vs_fallback[iter];
    probs[i] = probs[i] * prior[i];
  }
  if (c.probe)
    for (int i = 0; i < 32; ++i) c.probe->out->pdf[i] = (float)probs[i];
  // TransformPDFToCDF, HPM.cu:98-112
  T sum = 0;
  for (int i = 0; i < nvs; ++i) sum += probs[i];
  const T inv = (T)1 / sum;
  T cum = 0;
  for (int i = 0; i < nvs; ++i) { cum += probs[i] * inv; probs[i] = cum; }
  for (int k = 0; k < 15; ++k) {
    const T r = (T)(rng.uniform() - FLT_EPSILON);
    for (int v = 0; v < nvs; ++v) {
      if (probs[v] > r) { vw[v] += 1; break; }
    }
  }
  uint32_t sel = 0; T wn = 0; int ns = 0;
  for (int i = 0; i < nvs; ++i) if (vw[i] > (T)0) { sel |= (1u << i); wn += vw[i]; ns++; }
  *sel_out = sel; *weight_norm_out = wn; *nsel_out = ns;
  if (c.probe) {
    orc_probe_out* o = c.probe->out;
    for (int i = 0; i < 32; ++i) o->weights[i] = (float)vw[i];
    o->sel = sel;
    o->weight_norm = (float)wn;
  }
}

// Weighted direction costs, HPM.cu:1753-1771.
template <class T> void FinalCosts(const Ctx<T>& c, int px, int py, const Sampled<T>& s, const T* vw, T wn, T* fc) {
  const int nvs = c.nv - 1;
  for (int i = 0; i < 8; ++i) {
    fc[i] = 0;
    for (int j = 0; j < nvs; ++j) {
      if (vw[j] > (T)0) {
        if (c.prm.geom_consistency) {
          if (s.flag[i])
            fc[i] += vw[j] * (s.cost_array[i][j] + (T)0.2f * ComputeGeomConsistencyCost(c, j + 1, c.planes[s.positions[i]], px, py));
          else
            fc[i] += vw[j] * (s.cost_array[i][j] + (T)(0.1f * 3.0f));
        } else {
          fc[i] += vw[j] * s.cost_array[i][j];
        }
      }
    }
    fc[i] /= wn;
  }
  if (c.probe)
    for (int i = 0; i < 8; ++i) c.probe->out->fc[i] = (float)fc[i];
}

// Cost of the current plane with the selected view weights, HPM.cu:1775-1786.
template <class T> T CurrentCost(const Ctx<T>& c, int px, int py, const V4<T>& pl, const T* vw, T wn) {
  const int nvs = c.nv - 1;
  T cn = 0;
  for (int i = 0; i < nvs; ++i) {
    if (!(vw[i] > (T)0)) continue;  // zero weights add exactly +0 in the reference
    const T ci = ComputeBilateralNCC(c, i + 1, px, py, pl);
    if (c.prm.geom_consistency)
      cn += vw[i] * (ci + (T)0.2f * ComputeGeomConsistencyCost(c, i + 1, pl, px, py));
    else
      cn += vw[i] * ci;
  }
  return cn / wn;
}

template <class T> int FindMinCostIndex(const T* x, int n) {  // HPM.cu:41-52 (last minimum)
  T m = x[0]; int idx = 0;
  for (int i = 1; i < n; ++i) if (x[i] <= m) { m = x[i]; idx = i; }
  return idx;
}
template <class T> int FindMaxCostIndex(const T* x, int n) {  // HPM.cu:54-65 (last maximum)
  T m = x[0]; int idx = 0;
  for (int i = 1; i < n; ++i) if (x[i] >= m) { m = x[i]; idx = i; }
  return idx;
}

// HPM.cu:1347-1873
template <class T> void CheckerboardPropagation(Ctx<T>& c, int px, int py, int iter) {
  const orc_params& P = c.prm;
  const orc_camera& cam = c.cams[0];
  const int center = py * c.W + px;
  if (P.prior_consistency) {
    if (c.conf[center] > (T)0.3) return;
    if (c.masks[center] == 0) return;
  }
  Draws rng(c.seed, (uint32_t)center, (uint32_t)(1 + iter), c.probe);
  Sampled<T> s;
  SampleNeighbours(c, px, py, iter, &s);
  T vw[32]; uint32_t tsel; T wn; int nsel;
  ViewSelection(c, px, py, iter, s, rng, vw, &tsel, &wn, &nsel);
  T fc[8];
  FinalCosts(c, px, py, s, vw, wn, fc);
  const int min_idx = FindMinCostIndex(fc, 8);
  T cost_now = CurrentCost(c, px, py, c.planes[center], vw, wn);
  c.costs[center] = cost_now;
  T depth_now = ComputeDepthfromPlaneHypothesis(cam, c.planes[center], px, py);
  T restricted_cost = 0;
  T texture = 0;
  if (P.prior_consistency) {
    const Consts<T> k = MakeConsts<T>(P);
    const T depth_prior = ComputeDepthfromPlaneHypothesis(cam, c.prior[center], px, py);
    texture = ComputeTexture(c, px, py);
    if (c.masks[center] > 0) {
      T rfc[8] = {};
      for (int i = 0; i < 8; i++) {
        if (s.flag[i]) {
          const V4<T>& q = c.planes[s.positions[i]];
          const T dn = ComputeDepthfromPlaneHypothesis(cam, q, px, py);
          const T prior = PriorTerm(dn - depth_prior, NormDiff(c.prior[center], q), texture, k.two_dss, k.two_ass);
          rfc[i] = RestrictedCost(fc[i], k.beta, texture, prior);
        }
      }
      if (c.probe)
        for (int i = 0; i < 8; ++i) c.probe->out->rfc[i] = (float)rfc[i];
      const int max_idx = FindMaxCostIndex(rfc, 8);
      const T dn = ComputeDepthfromPlaneHypothesis(cam, c.planes[center], px, py);
      const T prior = PriorTerm(dn - depth_prior, NormDiff(c.prior[center], c.planes[center]), texture, k.two_dss, k.two_ass);
      const T rc_now = RestrictedCost(cost_now, k.beta, texture, prior);
      if (s.flag[max_idx]) {
        const T db = ComputeDepthfromPlaneHypothesis(cam, c.planes[s.positions[max_idx]], px, py);
        if (db >= (T)P.depth_min && db <= (T)P.depth_max && rfc[max_idx] > rc_now) {
          depth_now = db;
          c.planes[center] = c.planes[s.positions[max_idx]];
          c.costs[center] = fc[max_idx];
          restricted_cost = rfc[max_idx];
          c.sel[center] = tsel;
        }
      }
    } else if (s.flag[min_idx]) {
      const T db = ComputeDepthfromPlaneHypothesis(cam, c.planes[s.positions[min_idx]], px, py);
      if (db >= (T)P.depth_min && db <= (T)P.depth_max && fc[min_idx] < cost_now) {
        depth_now = db;
        c.planes[center] = c.planes[s.positions[min_idx]];
        c.costs[center] = fc[min_idx];
      }
    }
  }
  V4<T> plane_now = c.planes[center];  // Q1
  if (!P.prior_consistency && s.flag[min_idx]) {
    const T db = ComputeDepthfromPlaneHypothesis(cam, c.planes[s.positions[min_idx]], px, py);
    if (db >= (T)P.depth_min && db <= (T)P.depth_max && fc[min_idx] < cost_now) {
      depth_now = db;
      plane_now = c.planes[s.positions[min_idx]];
      cost_now = fc[min_idx];
      c.sel[center] = tsel;
    }
    if (c.costs[center] != c.costs[center]) {
      depth_now = db;
      plane_now = c.planes[s.positions[min_idx]];
      cost_now = fc[min_idx];
      c.sel[center] = tsel;
    }
  }
  PlaneHypothesisRefinement(c, &plane_now, &depth_now, &cost_now, rng, vw, wn, &restricted_cost, px, py, texture);
  if (P.hierarchy) {
    if (cost_now < c.pre_costs[center] - (T)0.1f) {
      c.costs[center] = cost_now;
      c.planes[center] = plane_now;
    }
  } else {
    c.costs[center] = cost_now;
    c.planes[center] = plane_now;
  }
}

// HPM.cu:1875-2791
template <class T> void CheckerboardPropagationMand(Ctx<T>& c, int px, int py, int iter) {
  const orc_params& P = c.prm;
  const orc_camera& cam = c.cams[0];
  const int W = c.W;
  const int center = py * W + px;
  Draws rng(c.seed, (uint32_t)center, (uint32_t)(1 + iter), c.probe);
  Sampled<T> s;
  SampleNeighbours(c, px, py, iter, &s);
  T vw[32]; uint32_t tsel; T wn; int nsel;
  ViewSelection(c, px, py, iter, s, rng, vw, &tsel, &wn, &nsel);
  if (c.conf[center] >= (T)0.3) {
    // HPM.cu:1903-2321: the plain update without prior or hierarchy
    T fc[8];
    FinalCosts(c, px, py, s, vw, wn, fc);
    const int min_idx = FindMinCostIndex(fc, 8);
    T cost_now = CurrentCost(c, px, py, c.planes[center], vw, wn);
    c.costs[center] = cost_now;
    T depth_now = ComputeDepthfromPlaneHypothesis(cam, c.planes[center], px, py);
    V4<T> plane_now = c.planes[center];  // Q1
    if (s.flag[min_idx]) {
      const T db = ComputeDepthfromPlaneHypothesis(cam, c.planes[s.positions[min_idx]], px, py);
      if (db >= (T)P.depth_min && db <= (T)P.depth_max && fc[min_idx] < cost_now) {
        depth_now = db;
        plane_now = c.planes[s.positions[min_idx]];
        cost_now = fc[min_idx];
        c.sel[center] = tsel;
      }
      if (c.costs[center] != c.costs[center]) {
        plane_now = c.planes[s.positions[min_idx]];
        cost_now = fc[min_idx];
        c.sel[center] = tsel;
      }
    }
    T unused_rc = 0;
    PlaneHypothesisRefinement(c, &plane_now, &depth_now, &cost_now, rng, vw, wn, &unused_rc, px, py, (T)0);
    c.costs[center] = cost_now;
    c.planes[center] = plane_now;
    return;
  }
  // HPM.cu:2322-2790: confidence-encoded adoption, no refinement
  const int nvs = c.nv - 1;
  T fc[8], rfc[8] = {};
  for (int i = 0; i < 8; i++) {
    fc[i] = 0;
    for (int j = 0; j < nvs; ++j) {
      if (vw[j] > (T)0) {
        if (P.geom_consistency) {
          if (s.flag[i])
            fc[i] += vw[j] * (s.cost_array[i][j] + (T)0.2f * ComputeGeomConsistencyCost(c, j + 1, c.planes[s.positions[i]], px, py));
          else
            fc[i] += vw[j] * (s.cost_array[i][j] + (T)(0.1f * 3.0f));
        } else {
          fc[i] += vw[j] * s.cost_array[i][j];
        }
      }
    }
    fc[i] /= wn;
    if (s.flag[i]) {
      const T beta = (T)0.18f;
      const T gce = EXP(-fc[i] * fc[i] / beta);
      const T mmc = (T)nsel * (T)0.5f;
      const T cp = c.conf[s.positions[i]];
      const T dlt = MAXF((T)0, mmc - cp);
      const T mce = (T)(0.5 + DEXP<T>((double)(-dlt * dlt) / ((double)(mmc * mmc) * 1.0)));
      rfc[i] = (cp > (T)0) ? MAXF((T)3.0f, mce) * gce : mce * gce;
    }
    if (s.flag[i]) {
      const T cp = c.conf[s.positions[i]];
      if (cp > (T)0) fc[i] = (T)((double)fc[i] * std::fmin((double)EXP(-cp * (T)2), 0.35));
    }
  }
  if (c.probe)
    for (int i = 0; i < 8; ++i) { c.probe->out->fc[i] = (float)fc[i]; c.probe->out->rfc[i] = (float)rfc[i]; }
  const int max_idx = FindMaxCostIndex(rfc, 8);
  T cost_now = 0;
  for (int i = 0; i < nvs; ++i) {
    if (!(vw[i] > (T)0)) continue;
    const T ci = ComputeBilateralNCC(c, i + 1, px, py, c.planes[center]);
    if (P.geom_consistency)
      cost_now += vw[i] * (ci + (T)0.2f * ComputeGeomConsistencyCost(c, i + 1, c.planes[center], px, py));
    else
      cost_now += vw[i] * ci;
  }
  cost_now /= wn;
  c.costs[center] = cost_now;
  const T beta = (T)0.18f;
  const T gce = EXP(-cost_now * cost_now / beta);
  const T mmc = (T)nsel * (T)0.5f;
  const T cc = c.conf[center];
  const T dlt = MAXF((T)0, mmc - cc);
  const T mce = (T)(0.5 + DEXP<T>((double)(-dlt * dlt) / ((double)(mmc * mmc) * 1.0)));
  const T rc_now = (cc > (T)0) ? MAXF((T)3.0f, mce) * gce : mce * gce;
  if (cc < (T)0.3) cost_now = (T)((double)cost_now * DEXP<T>(0.2 + (double)(cc * (T)2)));
  else cost_now = cost_now * EXP(-cc * (T)2);
  V4<T> plane_now = c.planes[center];  // Q1
  if (s.flag[max_idx]) {
    const int q = s.positions[max_idx];
    const T db = ComputeDepthfromPlaneHypothesis(cam, c.planes[q], px, py);
    if (db >= (T)P.depth_min && db <= (T)P.depth_max && rfc[max_idx] > rc_now) {
      plane_now = c.planes[q];
      cost_now = fc[max_idx];
      c.sel[center] = tsel;
      c.conf[center] = (T)(0.5 * (double)c.conf[q]);
    }
    if (c.costs[center] != c.costs[center]) {
      plane_now = c.planes[q];
      cost_now = fc[max_idx];
      c.sel[center] = tsel;
      c.conf[center] = (T)(0.5 * (double)c.conf[q]);
    }
  }
  c.costs[center] = cost_now;
  c.planes[center] = plane_now;
}

template <class T> void GetDepthandNormal(Ctx<T>& c, int px, int py) {  // HPM.cu:2829-2842
  const size_t center = (size_t)py * c.W + px;
  V4<T> pl = c.planes[center];
  pl.w = ComputeDepthfromPlaneHypothesis(c.cams[0], pl, px, py);
  c.planes[center] = TransformNormal(c.cams[0], pl);
}

template <class T> void SetupThresholds(Ctx<T>& c) {
  for (int it = 0; it < 3; ++it) {
    for (int e = 0; e < 3; ++e) {
      const int iter_tmp = 3 - e;
      c.judge_thr[it][e] = (float)(0.8 * orc::exp_hpm(-it * it * iter_tmp / 90.0));  // HPM.cu:1326-1327
    }
    // HPM.cu:1706 and 1724; float expf evaluated with the oracle's expf for both T
    const float thr = (float)(0.8 * (double)orc::expf_hpm((float)(it * it) / (-90.0f)));
    c.vs_thr[it] = (T)thr;
    c.vs_fallback[it] = (T)orc::expf_hpm(thr * thr / (-0.32f));
  }
}

// Sweep order inside a colour: 0 = rows top-down, columns left to right;
// 1 = rows bottom-up, columns right to left.  Every pixel of a colour reads
// only pixels of the other colour, so the order cannot change any result;
// the reversed sweep exists to test exactly that.
static int g_reverse_sweep = 0;

// Red/black pass: colour 0 = black = (x + y) even (HPM.cu:2793-2809), 1 = red.
template <class T> void ColourPass(Ctx<T>& c, int colour, int iter) {
  const int H2 = (c.H / 2) * 2;  // an odd last row is never visited (Q9, HPM.cu:3006)
  const int r0 = std::max(0, c.row_begin), r1 = std::min(H2, c.row_end);
  const bool rev = g_reverse_sweep != 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = r0; k < r1; ++k) {
    const int y = rev ? r1 - 1 - (k - r0) : k;
    const int x0 = (y + colour) & 1, n = (c.W - x0 + 1) / 2;
    for (int i = 0; i < n; ++i) {
      const int x = x0 + 2 * (rev ? n - 1 - i : i);
      if (c.prm.mand_consistency) CheckerboardPropagationMand(c, x, y, iter);
      else CheckerboardPropagation(c, x, y, iter);
    }
  }
}

template <class T> int RunPatchMatch(Ctx<T>& c, int n_halfpasses, int do_output) {  // HPM.cu:2987-3035
  SetupThresholds(c);
  const int r0 = std::max(0, c.row_begin), r1 = std::min(c.H, c.row_end);
  // a partial band (CPU-baseline sampling) also initialises the rows its
  // propagation reads: up to 29 rows above and 13 below
  const int i0 = (r0 == 0) ? 0 : std::max(0, r0 - 32), i1 = (r1 == c.H) ? c.H : std::min(c.H, r1 + 16);
#pragma omp parallel for schedule(dynamic, 1)
  for (int y = i0; y < i1; ++y)
    for (int x = 0; x < c.W; ++x) RandomInitialization(c, x, y);
  const int iters = c.prm.num_iterations;
  int done = 0;
  for (int i = 0; i < iters; ++i) {
    if (n_halfpasses >= 0 && done >= n_halfpasses) break;
    ColourPass(c, 0, i); ++done;
    if (n_halfpasses >= 0 && done >= n_halfpasses) break;
    ColourPass(c, 1, i); ++done;
  }
  if (do_output) {
#pragma omp parallel for schedule(static)
    for (int y = r0; y < r1; ++y)
      for (int x = 0; x < c.W; ++x) GetDepthandNormal(c, x, y);
  }
  return 0;
}

template <class T> int RunImpl(const orc_camera* cams, int num_images, const float* const* images,
                               const float* const* depth_images, const orc_params* params, uint64_t seed,
                               T* planes, T* costs, T* pre_costs, uint32_t* selected, const T* prior_planes,
                               const uint32_t* masks, T* confidences, const uint32_t* canny, T* texture,
                               const T* scaled_planes, int n_halfpasses, int do_output, int row_begin, int row_end) {
  if (num_images < 2 || num_images > 33) return -1;
  Ctx<T> c;
  c.W = cams[0].width; c.H = cams[0].height; c.nv = num_images; c.cams = cams; c.prm = *params; c.seed = seed;
  c.prm.num_images = num_images;
  for (int i = 0; i < num_images; ++i) c.imgs.push_back(Image{images[i], cams[i].width, cams[i].height});
  if (params->geom_consistency) {
    for (int i = 0; i < num_images; ++i)
      c.dimgs.push_back(Image{depth_images ? depth_images[i] : nullptr, cams[i].width, cams[i].height});
  }
  c.planes = reinterpret_cast<V4<T>*>(planes);
  c.costs = costs; c.pre_costs = pre_costs; c.sel = selected;
  c.prior = reinterpret_cast<const V4<T>*>(prior_planes);
  c.masks = masks; c.conf = confidences; c.canny = canny; c.texture = texture;
  c.scaled = reinterpret_cast<const V4<T>*>(scaled_planes);
  c.row_begin = row_begin < 0 ? 0 : row_begin;
  c.row_end = row_end < 0 ? c.H : row_end;
  return RunPatchMatch(c, n_halfpasses, do_output);
}

// ------------------------------------------------------------ plane gradient
// The aggregated cost of a pixel's plane and its gradient with respect to the
// plane's depth and two normal angles.  The reference has no gradient; this is
// the definition the GPU backward kernel is checked against:
//   C = sum_v w_v C_v / sum_v w_v over the views with w_v > 0 (weights frozen),
//   C_v = ComputeBilateralNCC (HPM.cu:375-448) of the plane
//   n = (sin t cos p, sin t sin p, -cos t), w = -n . X(d), X(d) = Get3DPoint at
//   depth d (HPM.cu:114-140),
// differentiated through H = K_s (R_rel - t_rel n^T / w) K_r^-1 (HPM.cu:251-311):
// dC_v/dH accumulated over the 36 taps (bilinear sample derivative times the
// projective map's derivative), then dH/d(d, t, p).  Zero where C_v is 2
// (centre outside, small variance) or clamped.  FP64 throughout.
namespace grad {

// K_s M K_r^-1 with the operations of ComputeHomography.
void MapHomography(const orc_camera& rc, const orc_camera& sc, const double M[9], double H[9]) {
  const double rK0 = rc.K[0], rK2 = rc.K[2], rK4 = rc.K[4], rK5 = rc.K[5];
  double tmp[9];
  for (int a = 0; a < 3; ++a) {
    tmp[3 * a + 0] = M[3 * a + 0] / rK0;
    tmp[3 * a + 1] = M[3 * a + 1] / rK4;
    tmp[3 * a + 2] = -M[3 * a + 0] * rK2 / rK0 - M[3 * a + 1] * rK5 / rK4 + M[3 * a + 2];
  }
  const double sK0 = sc.K[0], sK2 = sc.K[2], sK4 = sc.K[4], sK5 = sc.K[5], sK8 = sc.K[8];
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

void RelativePose(const orc_camera& rc, const orc_camera& sc, double R_rel[9], double t_rel[3]) {
  double ref_C[3], src_C[3], C_rel[3];
  for (int r = 0; r < 3; ++r) {
    ref_C[r] = -((double)rc.R[0 + r] * rc.t[0] + (double)rc.R[3 + r] * rc.t[1] + (double)rc.R[6 + r] * rc.t[2]);
    src_C[r] = -((double)sc.R[0 + r] * sc.t[0] + (double)sc.R[3 + r] * sc.t[1] + (double)sc.R[6 + r] * sc.t[2]);
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      R_rel[3 * a + b] = (double)sc.R[3 * a] * rc.R[3 * b] + (double)sc.R[3 * a + 1] * rc.R[3 * b + 1] +
                         (double)sc.R[3 * a + 2] * rc.R[3 * b + 2];
  for (int r = 0; r < 3; ++r) C_rel[r] = ref_C[r] - src_C[r];
  for (int a = 0; a < 3; ++a)
    t_rel[a] = (double)sc.R[3 * a] * C_rel[0] + (double)sc.R[3 * a + 1] * C_rel[1] + (double)sc.R[3 * a + 2] * C_rel[2];
}

// bilinear sample and its derivatives along u and v (Bilinear's cell and clamp)
double BilinearGrad(const Image& im, double u, double v, double& su, double& sv) {
  const double fu = std::floor(u), fv = std::floor(v);
  const double a = u - fu, b = v - fv;
  const double cu = std::min(std::max(fu, -1.0), (double)im.w), cv = std::min(std::max(fv, -1.0), (double)im.h);
  const int x0 = (int)cu, y0 = (int)cv;
  const double i00 = im.at(x0, y0), i10 = im.at(x0 + 1, y0), i01 = im.at(x0, y0 + 1), i11 = im.at(x0 + 1, y0 + 1);
  const double top = i00 + a * (i10 - i00), bot = i01 + a * (i11 - i01);
  su = (1.0 - b) * (i10 - i00) + b * (i11 - i01);
  sv = bot - top;
  return top + b * (bot - top);
}

// C_v and dC_v/dH for one view (returns false on the constant-cost branches).
bool CostAndDH(const Ctx<double>& c, int v, int px, int py, const double H[9], double& cost, double dH[9]) {
  const orc_camera& sc = c.cams[v];
  for (int k = 0; k < 9; ++k) dH[k] = 0;
  cost = 2.0;
  {
    const double X = H[0] * px + H[1] * py + H[2], Y = H[3] * px + H[4] * py + H[5], Z = H[6] * px + H[7] * py + H[8];
    const double u = X / Z, vv = Y / Z;
    if (u >= sc.width || u < 0 || vv >= sc.height || vv < 0) return false;
  }
  const Image& ref = c.imgs[0];
  const Image& src = c.imgs[v];
  const double ss = c.prm.sigma_spatial, scol = c.prm.sigma_color;
  const double rc0 = ref.at(px, py);
  double w[36], r[36], s[36], su[36], sv[36], uu[36], vvv[36], zz[36];
  int qx[36], qy[36];
  int k = 0;
  for (int i = -5; i <= 5; i += 2)
    for (int j = -5; j <= 5; j += 2, ++k) {
      qx[k] = px + i;
      qy[k] = py + j;
      r[k] = ref.at(qx[k], qy[k]);
      w[k] = std::exp(-std::sqrt((double)(i * i + j * j)) / (2 * ss * ss) - std::fabs(r[k] - rc0) / (2 * scol * scol));
      const double X = H[0] * qx[k] + H[1] * qy[k] + H[2], Y = H[3] * qx[k] + H[4] * qy[k] + H[5];
      zz[k] = H[6] * qx[k] + H[7] * qy[k] + H[8];
      uu[k] = X / zz[k];
      vvv[k] = Y / zz[k];
      s[k] = BilinearGrad(src, uu[k], vvv[k], su[k], sv[k]);
    }
  double W = 0, mr = 0, ms = 0;
  for (k = 0; k < 36; ++k) { W += w[k]; mr += w[k] * r[k]; ms += w[k] * s[k]; }
  mr /= W; ms /= W;
  double vr = 0, vs = 0, cv = 0;
  for (k = 0; k < 36; ++k) {
    vr += w[k] * (r[k] - mr) * (r[k] - mr);
    vs += w[k] * (s[k] - ms) * (s[k] - ms);
    cv += w[k] * (r[k] - mr) * (s[k] - ms);
  }
  vr /= W; vs /= W; cv /= W;
  if (vr < 1e-5 || vs < 1e-5) return false;
  const double den = std::sqrt(vr * vs);
  const double raw = 1.0 - cv / den;
  cost = std::max(0.0, std::min(2.0, raw));
  if (raw <= 0.0 || raw >= 2.0) return false;
  for (k = 0; k < 36; ++k) {
    const double dc_ds = -(w[k] / W) / den * ((r[k] - mr) - (cv / vs) * (s[k] - ms));
    const double gu = dc_ds * su[k] / zz[k], gv = dc_ds * sv[k] / zz[k];
    const double p[3] = {(double)qx[k], (double)qy[k], 1.0};
    for (int b = 0; b < 3; ++b) {
      dH[b] += gu * p[b];
      dH[3 + b] += gv * p[b];
      dH[6 + b] -= (gu * uu[k] + gv * vvv[k]) * p[b];
    }
  }
  return true;
}

}  // namespace grad

template <class T> int CostFixedImpl(const orc_camera* cams, int num_images, const float* const* images,
                                     const orc_params* params, const T* planes, T* out_costs, int row_begin, int row_end) {
  Ctx<T> c;
  c.W = cams[0].width; c.H = cams[0].height; c.nv = num_images; c.cams = cams; c.prm = *params;
  for (int i = 0; i < num_images; ++i) c.imgs.push_back(Image{images[i], cams[i].width, cams[i].height});
  const V4<T>* pl = reinterpret_cast<const V4<T>*>(planes);
  const int nvs = num_images - 1;
  const int r0 = row_begin < 0 ? 0 : row_begin, r1 = row_end < 0 ? c.H : std::min(c.H, row_end);
#pragma omp parallel for schedule(dynamic, 1)
  for (int y = r0; y < r1; ++y)
    for (int x = 0; x < c.W; ++x) {
      const size_t p = (size_t)y * c.W + x;
      for (int v = 1; v <= nvs; ++v) out_costs[p * nvs + (v - 1)] = ComputeBilateralNCC(c, v, x, y, pl[p]);
    }
  return 0;
}

template <class T> int InitCostImpl(const orc_camera* cams, int num_images, const float* const* images,
                                    const orc_params* params, const T* planes, T* out_cost, uint32_t* out_sel) {
  Ctx<T> c;
  c.W = cams[0].width; c.H = cams[0].height; c.nv = num_images; c.cams = cams; c.prm = *params;
  for (int i = 0; i < num_images; ++i) c.imgs.push_back(Image{images[i], cams[i].width, cams[i].height});
  const V4<T>* pl = reinterpret_cast<const V4<T>*>(planes);
#pragma omp parallel for schedule(dynamic, 1)
  for (int y = 0; y < c.H; ++y)
    for (int x = 0; x < c.W; ++x) {
      const size_t p = (size_t)y * c.W + x;
      out_cost[p] = ComputeMultiViewInitialCostandSelectedViews(c, x, y, pl[p], &out_sel[p]);
    }
  return 0;
}

template <class T> int GeomCostImpl(const orc_camera* cams, int num_images, const float* const* depth_images,
                                    const T* planes, T* out) {
  Ctx<T> c;
  c.W = cams[0].width; c.H = cams[0].height; c.nv = num_images; c.cams = cams;
  for (int i = 0; i < num_images; ++i) c.dimgs.push_back(Image{depth_images[i], cams[i].width, cams[i].height});
  const V4<T>* pl = reinterpret_cast<const V4<T>*>(planes);
  const int nvs = num_images - 1;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < c.H; ++y)
    for (int x = 0; x < c.W; ++x) {
      const size_t p = (size_t)y * c.W + x;
      for (int v = 1; v <= nvs; ++v) out[p * nvs + (v - 1)] = ComputeGeomConsistencyCost(c, v, pl[p], x, y);
    }
  return 0;
}

}  // namespace

extern "C" {

void orc_set_reverse_sweep(int reverse) { g_reverse_sweep = reverse; }

// Threads of the oracle's OpenMP loops from now on (the CPU-baseline timing
// runs them at the host's thread count and at 1).
void orc_set_threads(int n) { omp_set_num_threads(n < 1 ? 1 : n); }


int orc_run_patchmatch_f32(const orc_camera* cams, int num_images, const float* const* images,
                           const float* const* depth_images, const orc_params* params, uint64_t seed, float* planes,
                           float* costs, float* pre_costs, uint32_t* selected, const float* prior_planes,
                           const uint32_t* masks, float* confidences, const uint32_t* canny, float* texture,
                           const float* scaled_planes, int n_halfpasses, int do_output, int row_begin, int row_end) {
  return RunImpl<float>(cams, num_images, images, depth_images, params, seed, planes, costs, pre_costs, selected,
                        prior_planes, masks, confidences, canny, texture, scaled_planes, n_halfpasses, do_output,
                        row_begin, row_end);
}

int orc_run_patchmatch_f64(const orc_camera* cams, int num_images, const float* const* images,
                           const float* const* depth_images, const orc_params* params, uint64_t seed, double* planes,
                           double* costs, double* pre_costs, uint32_t* selected, const double* prior_planes,
                           const uint32_t* masks, double* confidences, const uint32_t* canny, double* texture,
                           const double* scaled_planes, int n_halfpasses, int do_output, int row_begin, int row_end) {
  return RunImpl<double>(cams, num_images, images, depth_images, params, seed, planes, costs, pre_costs, selected,
                         prior_planes, masks, confidences, canny, texture, scaled_planes, n_halfpasses, do_output,
                         row_begin, row_end);
}

int orc_cost_fixed_f32(const orc_camera* cams, int num_images, const float* const* images, const orc_params* params,
                       const float* planes, float* out, int row_begin, int row_end) {
  return CostFixedImpl<float>(cams, num_images, images, params, planes, out, row_begin, row_end);
}
int orc_cost_fixed_f64(const orc_camera* cams, int num_images, const float* const* images, const orc_params* params,
                       const double* planes, double* out, int row_begin, int row_end) {
  return CostFixedImpl<double>(cams, num_images, images, params, planes, out, row_begin, row_end);
}
int orc_init_cost_f32(const orc_camera* cams, int num_images, const float* const* images, const orc_params* params,
                      const float* planes, float* out_cost, uint32_t* out_sel) {
  return InitCostImpl<float>(cams, num_images, images, params, planes, out_cost, out_sel);
}
int orc_geom_cost_f32(const orc_camera* cams, int num_images, const float* const* depth_images, const float* planes,
                      float* out) {
  return GeomCostImpl<float>(cams, num_images, depth_images, planes, out);
}

// Scalar primitives, exported for the pinning tests.
// Aggregated cost and gradient per pixel (see namespace grad): params holds
// (d, theta, phi, unused) per pixel, weights (num_images - 1) per pixel;
// out_cost one double per pixel, out_grad (dC/dd, dC/dtheta, dC/dphi).
int orc_cost_grad_f64(const orc_camera* cams, int num_images, const float* const* images, const orc_params* params,
                      const float* pparams, const float* weights, double* out_cost, double* out_grad) {
  Ctx<double> c;
  c.W = cams[0].width; c.H = cams[0].height; c.nv = num_images; c.cams = cams; c.prm = *params;
  for (int i = 0; i < num_images; ++i) c.imgs.push_back(Image{images[i], cams[i].width, cams[i].height});
  const int nvs = num_images - 1;
  std::vector<double> Rr(9 * nvs), tr(3 * nvs);
  for (int v = 1; v <= nvs; ++v) grad::RelativePose(cams[0], cams[v], &Rr[9 * (v - 1)], &tr[3 * (v - 1)]);
#pragma omp parallel for schedule(dynamic, 1)
  for (int y = 0; y < c.H; ++y)
    for (int x = 0; x < c.W; ++x) {
      const size_t p = (size_t)y * c.W + x;
      const double d = pparams[4 * p], th = pparams[4 * p + 1], ph = pparams[4 * p + 2];
      const double n[3] = {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), -std::cos(th)};
      const double dn_dt[3] = {std::cos(th) * std::cos(ph), std::cos(th) * std::sin(ph), std::sin(th)};
      const double dn_dp[3] = {-std::sin(th) * std::sin(ph), std::sin(th) * std::cos(ph), 0.0};
      double X[3];
      Get3DPoint<double>(cams[0], x, y, d, X);
      const double wpl = -(n[0] * X[0] + n[1] * X[1] + n[2] * X[2]);
      const double q[3] = {X[0] / d, X[1] / d, X[2] / d};
      double wsum = 0, C = 0, g[3] = {0, 0, 0};
      for (int v = 1; v <= nvs; ++v) {
        const double wv = weights[p * nvs + (v - 1)];
        if (!(wv > 0)) continue;
        const double* R = &Rr[9 * (v - 1)];
        const double* t = &tr[3 * (v - 1)];
        double M[9], H[9], dMn[3][9], dMw[9], dHn[3][9], dHw[9];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            M[3 * a + b] = R[3 * a + b] - t[a] * n[b] / wpl;
            dMw[3 * a + b] = t[a] * n[b] / (wpl * wpl);
            for (int e = 0; e < 3; ++e) dMn[e][3 * a + b] = (b == e) ? -t[a] / wpl : 0.0;
          }
        grad::MapHomography(cams[0], cams[v], M, H);
        grad::MapHomography(cams[0], cams[v], dMw, dHw);
        for (int e = 0; e < 3; ++e) {
          // MapHomography is linear in M; the constant part of the K maps is 0
          grad::MapHomography(cams[0], cams[v], dMn[e], dHn[e]);
        }
        double cv, dH[9];
        const bool live = grad::CostAndDH(c, v, x, y, H, cv, dH);
        wsum += wv;
        C += wv * cv;
        if (!live) continue;
        // dC_v/dd = dH . dH/dw * dw/dd, dw/dd = -n . q
        double dHdw = 0, dHdn[3] = {0, 0, 0};
        for (int k = 0; k < 9; ++k) {
          dHdw += dH[k] * dHw[k];
          for (int e = 0; e < 3; ++e) dHdn[e] += dH[k] * dHn[e][k];
        }
        // total derivative with respect to n_e includes w(n): dw/dn_e = -X_e
        double dCdn[3];
        for (int e = 0; e < 3; ++e) dCdn[e] = dHdn[e] + dHdw * (-X[e]);
        const double nq = n[0] * q[0] + n[1] * q[1] + n[2] * q[2];
        g[0] += wv * dHdw * (-nq);
        g[1] += wv * (dCdn[0] * dn_dt[0] + dCdn[1] * dn_dt[1] + dCdn[2] * dn_dt[2]);
        g[2] += wv * (dCdn[0] * dn_dp[0] + dCdn[1] * dn_dp[1] + dCdn[2] * dn_dp[2]);
      }
      out_cost[p] = wsum > 0 ? C / wsum : 2.0;
      for (int e = 0; e < 3; ++e) out_grad[3 * p + e] = wsum > 0 ? g[e] / wsum : 0.0;
    }
  return 0;
}

// Aggregated cost only, at the same parameterisation (for finite differences).
int orc_cost_params_f64(const orc_camera* cams, int num_images, const float* const* images,
                        const orc_params* params, const double* pparams, const float* weights, double* out_cost) {
  Ctx<double> c;
  c.W = cams[0].width; c.H = cams[0].height; c.nv = num_images; c.cams = cams; c.prm = *params;
  for (int i = 0; i < num_images; ++i) c.imgs.push_back(Image{images[i], cams[i].width, cams[i].height});
  const int nvs = num_images - 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int y = 0; y < c.H; ++y)
    for (int x = 0; x < c.W; ++x) {
      const size_t p = (size_t)y * c.W + x;
      const double d = pparams[4 * p], th = pparams[4 * p + 1], ph = pparams[4 * p + 2];
      const double n[3] = {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), -std::cos(th)};
      double X[3];
      Get3DPoint<double>(cams[0], x, y, d, X);
      const V4<double> pl{n[0], n[1], n[2], -(n[0] * X[0] + n[1] * X[1] + n[2] * X[2])};
      double wsum = 0, C = 0;
      for (int v = 1; v <= nvs; ++v) {
        const double wv = weights[p * nvs + (v - 1)];
        if (!(wv > 0)) continue;
        wsum += wv;
        C += wv * ComputeBilateralNCC(c, v, x, y, pl);
      }
      out_cost[p] = wsum > 0 ? C / wsum : 2.0;
    }
  return 0;
}

// One pixel's RandomInitialization (op 0) or red/black update at iteration
// `iter` (op 1: CheckerboardPropagation, or the mandatory-consistency update
// when prm->mand_consistency), FP32, on the caller's maps (the size of
// cams[0]), with the pixel's uniforms taken from `draws` when it is not null,
// and the bilateral NCC and geometric cost of (view, plane) given by `ncc` and
// `geom` when they are not null.  `out` receives what the update computed.
// ref_image (cams[0] size) is read by the texture-free paths that sample the
// reference image (hierarchy upsampling); it may be null otherwise.
int orc_probe_pixel(int op, const orc_camera* cams, int num_images, const float* ref_image, const orc_params* prm,
                    float* planes, float* costs, float* pre_costs, uint32_t* selected, const float* prior_planes,
                    const uint32_t* masks, float* confidences, const uint32_t* canny, float* texture,
                    const float* scaled_planes, int px, int py, int iter, const float* draws, int n_draws,
                    orc_cost_fn ncc, orc_cost_fn geom, void* user, orc_probe_out* out) {
  if (num_images < 2 || num_images > 33 || !out || iter < 0 || iter > 2) return -1;
  Ctx<float> c;
  c.W = cams[0].width; c.H = cams[0].height; c.nv = num_images; c.cams = cams; c.prm = *prm; c.seed = 0;
  c.prm.num_images = num_images;
  if (px < 0 || py < 0 || px >= c.W || py >= c.H) return -1;
  c.imgs.push_back(Image{ref_image, cams[0].width, cams[0].height});
  for (int i = 1; i < num_images; ++i) c.imgs.push_back(Image{nullptr, cams[i].width, cams[i].height});
  if (prm->geom_consistency)
    for (int i = 0; i < num_images; ++i) c.dimgs.push_back(Image{nullptr, cams[i].width, cams[i].height});
  c.planes = reinterpret_cast<V4<float>*>(planes);
  c.costs = costs; c.pre_costs = pre_costs; c.sel = selected;
  c.prior = reinterpret_cast<const V4<float>*>(prior_planes);
  c.masks = masks; c.conf = confidences; c.canny = canny; c.texture = texture;
  c.scaled = reinterpret_cast<const V4<float>*>(scaled_planes);
  c.row_begin = 0; c.row_end = c.H;
  SetupThresholds(c);
  std::memset(out, 0, sizeof(*out));
  Probe pr;
  pr.draws = draws; pr.n_draws = n_draws; pr.ncc = ncc; pr.geom = geom; pr.user = user; pr.out = out;
  c.probe = &pr;
  if (op == 0) RandomInitialization(c, px, py);
  else if (c.prm.mand_consistency) CheckerboardPropagationMand(c, px, py, iter);
  else CheckerboardPropagation(c, px, py, iter);
  return 0;
}

float orc_expf(float x) { return orc::expf_hpm(x); }
double orc_exp(double x) { return orc::exp_hpm(x); }
float orc_sinf(float x) { return orc::sinf_hpm(x); }
float orc_cosf(float x) { return orc::cosf_hpm(x); }
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  orc::U4 o = orc::philox4x32_10(ctr, key);
  std::memcpy(out, o.v, 16);
}
void orc_uniforms(uint64_t seed, uint32_t pixel, uint32_t stage, int n, float* out) {
  orc::PixelRng r(seed, pixel, stage);
  for (int i = 0; i < n; ++i) out[i] = r.uniform();
}
float orc_bilinear(const float* img, int w, int h, float u, float v) { return Bilinear<float>(Image{img, w, h}, u, v); }
double orc_bilinear_f64(const float* img, int w, int h, double u, double v) { return Bilinear<double>(Image{img, w, h}, u, v); }
void orc_homography_f64(const orc_camera* rc, const orc_camera* sc, const double* plane, double* H) {
  ComputeHomography<double>(*rc, *sc, V4<double>{plane[0], plane[1], plane[2], plane[3]}, H);
}
void orc_random_plane_f32(const orc_camera* cam, int px, int py, uint64_t seed, uint32_t pixel, float dmin, float dmax,
                          float* out) {
  Draws rng(seed, pixel, 0u, nullptr);
  V4<float> p = GenerateRandomPlaneHypothesis<float>(*cam, px, py, rng, dmin, dmax);
  out[0] = p.x; out[1] = p.y; out[2] = p.z; out[3] = p.w;
}
float orc_depth_from_plane_f64(const orc_camera* cam, const double* plane, int px, int py) {
  return (float)ComputeDepthfromPlaneHypothesis<double>(*cam, V4<double>{plane[0], plane[1], plane[2], plane[3]}, px, py);
}
double orc_distance_to_origin_f64(const orc_camera* cam, int px, int py, double depth, const double* n) {
  return GetDistance2Origin<double>(*cam, px, py, depth, V4<double>{n[0], n[1], n[2], 0});
}

}  // extern "C"
