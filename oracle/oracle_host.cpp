// oracle_host.cpp -- CPU restatement of the passes around PatchMatch that the
// GPU library also implements: joint bilateral upsampling, the multi-view
// consistency check of fusion / confidence evaluation, and the planar prior
// (support points, triangle planes, raster, depth test, the coarse prior of
// distances 1 and 2 and its per-pixel reload).
//
// TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg are the only users.  The product never includes, links or calls
// anything under oracle/.
//
// Arithmetic: plain IEEE float operations in the order the reference writes
// them (built with -ffp-contract=off, no fused multiply-adds).  Calls the
// reference makes with float arguments use float overloads (exp(float) is the
// oracle's expf, acosf the oracle's acosf); expressions the reference promotes
// to double are evaluated in double: pow(float, int) returns double in C++11
// (main.cpp is host C++11), so the reprojection error
// sqrt(pow(c - u, 2) + pow(r - v, 2)) is double until its assignment to
// float, and double literals are evaluated with the oracle's own double exp.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_math.h"

namespace {

struct Cam {  // the layout of Camera (main.h:39-47)
  float K[9];
  float R[9];
  float t[3];
  int height;
  int width;
  float depth_min;
  float depth_max;
};

inline float Texel(const float* img, int w, int h, int x, int y) {  // clamp-to-edge texture read
  x = x < 0 ? 0 : (x >= w ? w - 1 : x);
  y = y < 0 ? 0 : (y >= h ? h - 1 : y);
  return img[(size_t)y * w + x];
}

// SpatialGauss, RangeGauss (HPM.cu:142-152): the -1.0 literal makes the
// exponent a double expression.
inline float SpatialGauss(float x1, float y1, float x2, float y2, float sigma) {
  const float dis = (x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) - 0.0f;
  return (float)orc::exp_hpm(-1.0 * (double)dis / (double)(2 * sigma * sigma));
}
inline float RangeGauss(float x, float sigma) {
  const float x_p = x - 0.0f;
  return (float)orc::exp_hpm(-1.0 * (double)(x_p * x_p) / (double)(2 * sigma * sigma));
}

struct Jbu {
  int height, width, s_height, s_width, scale;
};

Jbu MakeJbu(int w, int h, int cw, int ch) {
  // Imagescale = max(rows / s_rows, cols / s_cols), HPM.cpp:1270
  const int a = h / ch, b = w / cw;
  return Jbu{h, w, ch, cw, a > b ? a : b};
}

// JBU_cu, HPM.cu:3037-3095.
void JbuDepthPixel(const Jbu& jp, const float* fine, const float* coarse, int px, int py, float* out) {
  const float scale = (float)(1.0 * jp.s_width / jp.width);
  const float sigmad = 0.50f, sigmar = 25.5f;
  const int WinWidth = jp.scale * jp.scale + 1;
  const int num_neighbors = WinWidth / 2;
  const float o_y = py * scale, o_x = px * scale;
  const float refPix = Texel(fine, jp.width, jp.height, px, py);
  float total_val = 0.0f, normalizing_factor = 0.0f;
  for (int j = -num_neighbors; j <= num_neighbors; ++j) {
    int r_y = (int)(o_y + j);
    r_y = (r_y > 0 ? (r_y < jp.s_height ? r_y : jp.s_height - 1) : 0);
    int r_ys = py + j;
    r_ys = (r_ys > 0 ? (r_ys < jp.height ? r_ys : jp.height - 1) : 0);
    for (int i = -num_neighbors; i <= num_neighbors; ++i) {
      int r_x = (int)(o_x + i);
      r_x = (r_x > 0 ? (r_x < jp.s_width ? r_x : jp.s_width - 1) : 0);
      const float srcPix = Texel(coarse, jp.s_width, jp.s_height, r_x, r_y);
      int r_xs = px + i;
      r_xs = (r_xs > 0 ? (r_xs < jp.width ? r_xs : jp.width - 1) : 0);
      const float neighborPix = Texel(fine, jp.width, jp.height, r_xs, r_ys);
      const float sgauss = SpatialGauss(o_x, o_y, (float)r_x, (float)r_y, sigmad);
      const float rgauss = RangeGauss(std::fabs(refPix - neighborPix), sigmar);
      const float totalgauss = sgauss * rgauss;
      normalizing_factor += totalgauss;
      total_val += srcPix * totalgauss;
    }
  }
  out[(size_t)py * jp.width + px] = total_val / normalizing_factor;
}

// JBU_cu_prior, HPM.cu:3130-3197: the fine-image neighbour coordinates are not
// clamped (the texture clamps them); the normal sum is divided by the weight
// sum and normalised with 1 / sqrt (NormalizeVec3, HPM.cu:89-96); w stays 0.
void JbuPriorPixel(const Jbu& jp, const float* fine, const float* coarse, const float* coarse_n, int px, int py,
                   float* out_d, float* out_n) {
  const float scale = (float)(1.0 * jp.s_width / jp.width);
  const float sigmad = 0.50f, sigmar = 25.5f;
  const int WinWidth = jp.scale * jp.scale + 1;
  const int num_neighbors = WinWidth / 2;
  const float o_y = py * scale, o_x = px * scale;
  const float refPix = Texel(fine, jp.width, jp.height, px, py);
  float c_total_val = 0.0f, normalizing_factor = 0.0f;
  float nx = 0.0f, ny = 0.0f, nz = 0.0f;
  for (int j = -num_neighbors; j <= num_neighbors; ++j) {
    int r_y = (int)(o_y + j);
    r_y = (r_y > 0 ? (r_y < jp.s_height ? r_y : jp.s_height - 1) : 0);
    const int r_ys = py + j;
    for (int i = -num_neighbors; i <= num_neighbors; ++i) {
      int r_x = (int)(o_x + i);
      r_x = (r_x > 0 ? (r_x < jp.s_width ? r_x : jp.s_width - 1) : 0);
      const size_t s_center = (size_t)r_y * jp.s_width + r_x;
      const float srcPix = Texel(coarse, jp.s_width, jp.s_height, r_x, r_y);
      const float* sn = coarse_n + 4 * s_center;
      const int r_xs = px + i;
      const float neighborPix = Texel(fine, jp.width, jp.height, r_xs, r_ys);
      const float sgauss = SpatialGauss(o_x, o_y, (float)r_x, (float)r_y, sigmad);
      const float rgauss = RangeGauss(std::fabs(refPix - neighborPix), sigmar);
      const float totalgauss = sgauss * rgauss;
      normalizing_factor += totalgauss;
      c_total_val += srcPix * totalgauss;
      nx = nx + sn[0] * totalgauss;
      ny = ny + sn[1] * totalgauss;
      nz = nz + sn[2] * totalgauss;
    }
  }
  nx = nx / normalizing_factor;
  ny = ny / normalizing_factor;
  nz = nz / normalizing_factor;
  const float ns = nx * nx + ny * ny + nz * nz;
  const float inv = 1.0f / std::sqrt(ns);
  const size_t c = (size_t)py * jp.width + px;
  out_d[c] = c_total_val / normalizing_factor;
  out_n[4 * c + 0] = nx * inv;
  out_n[4 * c + 1] = ny * inv;
  out_n[4 * c + 2] = nz * inv;
  out_n[4 * c + 3] = 0.0f;
}

// Get3DPointonWorld, HPM.cpp:221-245.
inline void PointOnWorld(int x, int y, float depth, const Cam& c, float X[3]) {
  const float px = depth * ((float)x - c.K[2]) / c.K[0];
  const float py = depth * ((float)y - c.K[5]) / c.K[4];
  const float pz = depth;
  const float tx = c.R[0] * px + c.R[3] * py + c.R[6] * pz;
  const float ty = c.R[1] * px + c.R[4] * py + c.R[7] * pz;
  const float tz = c.R[2] * px + c.R[5] * py + c.R[8] * pz;
  const float Cx = -(c.R[0] * c.t[0] + c.R[3] * c.t[1] + c.R[6] * c.t[2]);
  const float Cy = -(c.R[1] * c.t[0] + c.R[4] * c.t[1] + c.R[7] * c.t[2]);
  const float Cz = -(c.R[2] * c.t[0] + c.R[5] * c.t[1] + c.R[8] * c.t[2]);
  X[0] = tx + Cx;
  X[1] = ty + Cy;
  X[2] = tz + Cz;
}

// ProjectonCamera, HPM.cpp:269-279.
inline void ProjectOnCamera(const float X[3], const Cam& c, float& u, float& v, float& depth) {
  const float tx = c.R[0] * X[0] + c.R[1] * X[1] + c.R[2] * X[2] + c.t[0];
  const float ty = c.R[3] * X[0] + c.R[4] * X[1] + c.R[5] * X[2] + c.t[1];
  const float tz = c.R[6] * X[0] + c.R[7] * X[1] + c.R[8] * X[2] + c.t[2];
  depth = c.K[6] * tx + c.K[7] * ty + c.K[8] * tz;
  u = (c.K[0] * tx + c.K[1] * ty + c.K[2] * tz) / depth;
  v = (c.K[3] * tx + c.K[4] * ty + c.K[5] * tz) / depth;
}

// (int)(v + 0.5f) of main.cpp:908-909,1046-1047 with a defined result when v
// is NaN or beyond the int range (x86's conversion gives INT_MIN there, which
// every bounds test rejects): -1.
inline int round_index(float v) {
  const float x = v + 0.5f;
  return (x > -2147483648.0f && x < 2147483648.0f) ? (int)x : -1;
}

// The reprojection error sqrt(pow(c - u, 2) + pow(r - v, 2)) of
// main.cpp:778,922,1064: c - u is a float, pow(float, int) is double in C++11,
// so the squares, the sum and sqrt are double; the result is stored as float.
inline float ReprojError(int c, int r, float u, float v) {
  const float dx = (float)c - u, dy = (float)r - v;
  return (float)std::sqrt((double)dx * (double)dx + (double)dy * (double)dy);
}

// GetAngle, HPM.cpp:281-290.
inline float GetAngle(const float* a, const float* b) {
  const float dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  const float angle = orc::acosf_hpm(dot);
  if (angle != angle) return 0.0f;
  return angle;
}

// ------------------------------------------------------------ planar prior
// GetSupportPoints_Classify_Check (HPM.cpp:1026-1073): per 5 x 5 block,
// column-major over the blocks, the pixel minimising cost - confidence and
// the one minimising cost + 0.2 - confidence (a double expression stored as
// float), among pixels with cost < 2; each kept when below 0.1 (double).
// The texture flag the reference computes is never read.
// The maps have `stride` columns; the blocks cover the first w x h of them
// (the coarse prior scans (int)(width * factor) columns of maps resized to
// round(width * factor), HPM.cpp:1030-1031).
int SupportPoints(const float* costs, const float* conf, int stride, int w, int h, int32_t* out_xy, int cap) {
  int n = 0;
  for (int col = 0; col < w; col += 5)
    for (int row = 0; row < h; row += 5) {
      float best_plain = 2.0f, best_shift = 2.0f;
      int plain_x = 0, plain_y = 0, shift_x = 0, shift_y = 0;
      for (int c = col; c < std::min(w, col + 5); ++c)
        for (int r = row; r < std::min(h, row + 5); ++r) {
          const float cost = costs[(size_t)r * stride + c], cf = conf[(size_t)r * stride + c];
          if (!(cost < 2.0f)) continue;
          const float plain = cost - cf;
          const float shifted = (float)(cost + 0.2 - cf);
          if (best_plain > plain) { best_plain = plain; plain_x = c; plain_y = r; }
          if (best_shift > shifted) { best_shift = shifted; shift_x = c; shift_y = r; }
        }
      // the shifted winner is pushed first (HPM.cpp:1065-1070)
      if (best_shift < 0.1) { if (n < cap) { out_xy[2 * n] = shift_x; out_xy[2 * n + 1] = shift_y; } ++n; }
      if (best_plain < 0.1) { if (n < cap) { out_xy[2 * n] = plain_x; out_xy[2 * n + 1] = plain_y; } ++n; }
    }
  return n;
}

// Get3DPointonRefCam (HPM.cpp:209-219): the pixel back-projected at `depth`.
inline void RefCamPoint(const Cam& c, int x, int y, float depth, float X[3]) {
  X[0] = depth * ((float)x - c.K[2]) / c.K[0];
  X[1] = depth * ((float)y - c.K[5]) / c.K[4];
  X[2] = depth;
}

// Get3DPointonRefCam_factor (HPM.cpp:247-256): the same with the intrinsics
// scaled by `factor` inside the expression.
inline void RefCamPointFactor(const Cam& c, int x, int y, float depth, float factor, float X[3]) {
  X[0] = depth * ((float)x - c.K[2] * factor) / (c.K[0] * factor);
  X[1] = depth * ((float)y - c.K[5] * factor) / (c.K[4] * factor);
  X[2] = depth;
}

// GetPriorPlaneParams (HPM.cpp:1134-1167): the unit null vector (n, w) of the
// rows [X_k 1] -- here n = (X2 - X1) x (X3 - X1), w = -n . X1 in double,
// scaled to unit length and stored as float as cv::SVD::solveZ returns it --
// then divided by |n| (a double sqrt of float squares), with the sign that
// makes w >= 0.
void PriorPlane(const float X1[3], const float X2[3], const float X3[3], float plane[4]) {
  const double a[3] = {(double)X2[0] - X1[0], (double)X2[1] - X1[1], (double)X2[2] - X1[2]};
  const double b[3] = {(double)X3[0] - X1[0], (double)X3[1] - X1[1], (double)X3[2] - X1[2]};
  double v[4] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0], 0.0};
  v[3] = -(v[0] * X1[0] + v[1] * X1[1] + v[2] * X1[2]);
  const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
  float n4[4];
  for (int k = 0; k < 4; ++k) n4[k] = (float)(v[k] / len);
  float norm2 = (float)std::sqrt((double)n4[0] * n4[0] + (double)n4[1] * n4[1] + (double)n4[2] * n4[2]);
  if (n4[3] < 0) norm2 *= -1;
  for (int k = 0; k < 4; ++k) plane[k] = n4[k] / norm2;
}

// GetDepthFromPlaneParam (HPM.cpp:1169-1172).
inline float PlaneDepth(const Cam& c, const float pl[4], int x, int y) {
  return -pl[3] * c.K[0] / (((float)x - c.K[2]) * pl[0] + (c.K[0] / c.K[4]) * ((float)y - c.K[5]) * pl[1] + c.K[0] * pl[2]);
}

// GetDepthFromPlaneParam_factor (HPM.cpp:1174-1177).
inline float PlaneDepthFactor(const Cam& c, const float pl[4], int x, int y, float factor) {
  return -pl[3] * (c.K[0] * factor) /
         (((float)x - c.K[2] * factor) * pl[0] + (c.K[0] / c.K[4]) * ((float)y - c.K[5] * factor) * pl[1] +
          c.K[0] * factor * pl[2]);
}

// The triangles with all three corners inside the w x h grid, in order, get
// index k = 1, 2, ...; each is rasterised by the barycentric walk of
// main.cpp:396-404 / 499-513 (later triangles overwrite earlier ones) into
// `mask`, and its plane is the plane through its corners back-projected by
// `point` (x, y, depth at the corner, X).
template <class PointFn>
void RasterTrianglePlanes(int w, int h, const float* depths, const int32_t* tris, int ntri, PointFn point,
                          std::vector<float>& mask, std::vector<float>& planes) {
  mask.assign((size_t)w * h, 0.0f);
  planes.clear();
  uint32_t idx = 0;
  for (int t = 0; t < ntri; ++t) {
    const int32_t* T = tris + 6 * t;
    bool inside = true;
    for (int k = 0; k < 3; ++k) inside = inside && T[2 * k] >= 0 && T[2 * k] < w && T[2 * k + 1] >= 0 && T[2 * k + 1] < h;
    if (!inside) continue;
    const int x1 = T[0], y1 = T[1], x2 = T[2], y2 = T[3], x3 = T[4], y3 = T[5];
    // float L = sqrt(pow(int, 2) + pow(int, 2)): the squares and sqrt in double
    const float L01 = (float)std::sqrt(std::pow((double)(x1 - x2), 2) + std::pow((double)(y1 - y2), 2));
    const float L02 = (float)std::sqrt(std::pow((double)(x1 - x3), 2) + std::pow((double)(y1 - y3), 2));
    const float L12 = (float)std::sqrt(std::pow((double)(x2 - x3), 2) + std::pow((double)(y2 - y3), 2));
    const float longest = std::max(L01, std::max(L02, L12));
    const float step = (float)(1.0 / longest);
    for (float p = 0; p < 1.0; p += step)
      for (float q = 0; q < 1.0 - p; q += step) {
        // p * x1 + q * x2 are float products and a float sum; (1.0 - p - q) * x3 is double
        const float fx = p * (float)x1 + q * (float)x2, fy = p * (float)y1 + q * (float)y2;
        const int x = (int)((double)fx + (1.0 - p - q) * x3);
        const int y = (int)((double)fy + (1.0 - p - q) * y3);
        if (x >= 0 && x < w && y >= 0 && y < h) mask[(size_t)y * w + x] = (float)(idx + 1.0);
      }
    float X1[3], X2[3], X3[3], pl[4];
    point(x1, y1, depths[(size_t)y1 * w + x1], X1);
    point(x2, y2, depths[(size_t)y2 * w + x2], X2);
    point(x3, y3, depths[(size_t)y3 * w + x3], X3);
    PriorPlane(X1, X2, X3, pl);
    planes.insert(planes.end(), pl, pl + 4);
    ++idx;
  }
}

// The distance-0 prior of ProcessProblem (main.cpp:385-431) and
// CudaPlanarPriorInitialization (HPM.cpp:964-984): triangles with all three
// corners inside the image, in order, get index k = 1, 2, ...; each is
// rasterised by the barycentric walk of main.cpp:396-404 (later triangles
// overwrite earlier ones) and gets the plane through its back-projected
// corners; a pixel keeps its triangle only when that plane's depth there is
// within the pass's depth range.  Prior planes of pixels without a triangle
// stay 0.
void PriorAtLevel(const Cam& cam, const float* depths, const int32_t* tris, int ntri, float* prior, uint32_t* masks) {
  const int w = cam.width, h = cam.height;
  std::vector<float> mask, planes;
  RasterTrianglePlanes(w, h, depths, tris, ntri,
                       [&](int x, int y, float d, float X[3]) { RefCamPoint(cam, x, y, d, X); }, mask, planes);
  const float dmin = cam.depth_min * 0.6f, dmax = cam.depth_max * 1.2f;  // GetMinDepth / GetMaxDepth
  for (int x = 0; x < w; ++x)
    for (int y = 0; y < h; ++y) {
      const size_t i = (size_t)y * w + x;
      uint32_t m = (uint32_t)mask[i];
      if (m > 0) {
        const float d = PlaneDepth(cam, &planes[4 * (m - 1)], x, y);
        if (!(d <= dmax && d >= dmin)) m = 0;
      }
      masks[i] = m;
      for (int k = 0; k < 4; ++k) prior[4 * i + k] = m > 0 ? planes[4 * (m - 1) + k] : 0.0f;
    }
}

// The coarse planar prior of ProcessProblem at hpm_scale_distance 1 or 2
// (main.cpp:443-541), on the maps resized to hw x hh = round(factor x the
// view's size): triangle planes back-projected with the scaled intrinsics
// (GetPriorPlaneParams_factor, HPM.cpp:1099-1132), the raster, and per
// pixel with a triangle the plane's depth (GetDepthFromPlaneParam_factor)
// when it lies within [GetMinDepth * 0.6, GetMaxDepth * 1.2], with the
// plane's normal turned to the world frame (TransformNormal,
// HPM.cpp:1179-1187); other pixels get depth 0 and normal 0.  prior_n holds
// 4 floats per pixel (normal, 0).
void PriorCoarse(const Cam& cam, const float* depths, int hw, int hh, const int32_t* tris, int ntri, float factor,
                 float* prior_d, float* prior_n) {
  std::vector<float> mask, planes;
  RasterTrianglePlanes(hw, hh, depths, tris, ntri,
                       [&](int x, int y, float d, float X[3]) { RefCamPointFactor(cam, x, y, d, factor, X); }, mask,
                       planes);
  const float dmin = cam.depth_min * 0.6f, dmax = cam.depth_max * 1.2f;  // GetMinDepth / GetMaxDepth
  for (int x = 0; x < hw; ++x)
    for (int y = 0; y < hh; ++y) {
      const size_t i = (size_t)y * hw + x;
      prior_d[i] = 0.0f;
      for (int k = 0; k < 4; ++k) prior_n[4 * i + k] = 0.0f;
      const uint32_t m = (uint32_t)mask[i];
      if (m == 0) continue;
      const float* pl = &planes[4 * (m - 1)];
      const float d = PlaneDepthFactor(cam, pl, x, y, factor);
      if (d <= dmax * 1.2f && d >= dmin * 0.6f) {
        prior_d[i] = d;
        const float* R = cam.R;
        prior_n[4 * i + 0] = R[0] * pl[0] + R[3] * pl[1] + R[6] * pl[2];
        prior_n[4 * i + 1] = R[1] * pl[0] + R[4] * pl[1] + R[7] * pl[2];
        prior_n[4 * i + 2] = R[2] * pl[0] + R[5] * pl[1] + R[8] * pl[2];
      }
    }
}

// The per-pixel reload of the upsampled coarse prior (main.cpp:553-574):
// where the upsampled depth lies within [GetMinDepth, GetMaxDepth] the mask
// is 1 and the prior plane is the upsampled world normal turned to the
// camera frame (TransformNormal2RefCam, HPM.cpp:1189-1197) with w =
// GetDistance2Origin (HPM.cpp:1199-1206); elsewhere mask 0, plane 0.
void PriorReload(const Cam& cam, const float* up_d, const float* up_n, float* prior, uint32_t* masks) {
  const int w = cam.width, h = cam.height;
  const float dmin = cam.depth_min * 0.6f, dmax = cam.depth_max * 1.2f;
  const float* R = cam.R;
  for (int x = 0; x < w; ++x)
    for (int y = 0; y < h; ++y) {
      const size_t c = (size_t)y * w + x;
      const float d = up_d[c];
      for (int k = 0; k < 4; ++k) prior[4 * c + k] = 0.0f;
      masks[c] = 0;
      if (!(d <= dmax && d >= dmin)) continue;
      masks[c] = 1;
      const float n0 = up_n[4 * c], n1 = up_n[4 * c + 1], n2 = up_n[4 * c + 2];
      const float t0 = R[0] * n0 + R[1] * n1 + R[2] * n2;
      const float t1 = R[3] * n0 + R[4] * n1 + R[5] * n2;
      const float t2 = R[6] * n0 + R[7] * n1 + R[8] * n2;
      const float X0 = d * ((float)x - cam.K[2]) / cam.K[0];
      const float X1 = d * ((float)y - cam.K[5]) / cam.K[4];
      const float X2 = d;
      prior[4 * c + 0] = t0;
      prior[4 * c + 1] = t1;
      prior[4 * c + 2] = t2;
      prior[4 * c + 3] = -(t0 * X0 + t1 * X1 + t2 * X2);
    }
}

}  // namespace

extern "C" {

int orc_support_points(const float* costs, const float* conf, int w, int h, int32_t* out_xy, int cap, int* count) {
  *count = SupportPoints(costs, conf, w, w, h, out_xy, cap);
  return 0;
}

// Support points over the first w x h of maps with `stride` columns.
int orc_support_points_scan(const float* costs, const float* conf, int stride, int w, int h, int32_t* out_xy, int cap,
                            int* count) {
  *count = SupportPoints(costs, conf, stride, w, h, out_xy, cap);
  return 0;
}

int orc_prior_coarse(const Cam* cam, const float* depths, int hw, int hh, const int32_t* tris, int ntri, float factor,
                     float* prior_d, float* prior_n) {
  PriorCoarse(*cam, depths, hw, hh, tris, ntri, factor, prior_d, prior_n);
  return 0;
}

int orc_prior_reload(const Cam* cam, const float* up_d, const float* up_n, float* prior, uint32_t* masks) {
  PriorReload(*cam, up_d, up_n, prior, masks);
  return 0;
}

void orc_prior_plane(const float* X1, const float* X2, const float* X3, float* plane) { PriorPlane(X1, X2, X3, plane); }

int orc_prior_level(const Cam* cam, const float* depths, const int32_t* tris, int ntri, float* prior, uint32_t* masks) {
  PriorAtLevel(*cam, depths, tris, ntri, prior, masks);
  return 0;
}

void orc_jbu_depth(const float* fine, int w, int h, const float* coarse, int cw, int ch, float* out) {
  const Jbu jp = MakeJbu(w, h, cw, ch);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) JbuDepthPixel(jp, fine, coarse, x, y, out);
}

void orc_jbu_prior(const float* fine, int w, int h, const float* coarse, const float* coarse_normals, int cw, int ch,
                   float* out_depth, float* out_normals) {
  const Jbu jp = MakeJbu(w, h, cw, ch);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) JbuPriorPixel(jp, fine, coarse, coarse_normals, x, y, out_depth, out_normals);
}

float orc_acosf(float x) { return orc::acosf_hpm(x); }

void orc_reproj_error(const int* c, const int* r, const float* u, const float* v, int n, float* out) {
  for (int i = 0; i < n; ++i) out[i] = ReprojError(c[i], r[i], u[i], v[i]);
}

// RunFusion (mode 0, main.cpp:823-957) and ConfidenceEvaluation (mode 1,
// main.cpp:959-1109), in the reference's order: images in order, pixels
// row-major, source views in pair order; masks greedy across images.  The
// `used_list` of an image is not reset between pixels (main.cpp:883,1018), so
// an accepted pixel also masks the source pixels that earlier, rejected pixels
// of the same image matched.
//   cams[n], depths[i] (w*h), normals[i] (3 floats per pixel, cv::Vec3f);
//   src_ids[src_offsets[i] .. src_offsets[i+1]) are image i's sources.
// mode 0: out_pairs receives (image, pixel) of every fused point, in order, up
//   to capacity; *count is the total.  out_points (6 floats per point: world
//   position, reference normal) may be null.
// mode 1: out_conf[i] (w*h floats) receives image i's confidence.dmb: the
//   in-memory consistency map as it stands when image i is done
//   (writeDepthDmb inside the image loop, main.cpp:1095-1100).  Images after
//   i that list it as a source keep writing into the in-memory map
//   (main.cpp:1086); those writes never reach the file.
// mode 2: as mode 0 with RunFusion_Sky_Strict's float threshold (main.cpp:659-821).
int orc_consistency(int mode, int n, const Cam* cams, const float* const* depths, const float* const* normals,
                    const int* src_offsets, const int* src_ids, float* const* out_conf, int32_t* out_pairs,
                    float* out_points, int64_t capacity, int64_t* count) {
  std::vector<std::vector<uint8_t>> masks(n);
  for (int i = 0; i < n; ++i) masks[i].assign((size_t)cams[i].width * cams[i].height, 0);
  std::vector<std::vector<float>> consistency(mode == 1 ? n : 0);  // cv::Mat::zeros, main.cpp:1005
  for (size_t i = 0; i < consistency.size(); ++i) consistency[i].assign((size_t)cams[i].width * cams[i].height, 0.0f);
  int64_t npts = 0;
  for (int i = 0; i < n; ++i) {
    const int cols = cams[i].width, rows = cams[i].height;
    const int num_ngb = src_offsets[i + 1] - src_offsets[i];
    const int* srcs = src_ids + src_offsets[i];
    std::vector<int> used_x(num_ngb, -1), used_y(num_ngb, -1);
    for (int r = 0; r < rows; ++r) {
      for (int c = 0; c < cols; ++c) {
        const size_t center = (size_t)r * cols + c;
        if (masks[i][center] == 1) continue;
        const float ref_depth = depths[i][center];
        const float* ref_normal = normals[i] + 3 * center;
        if (ref_depth <= 0.0) continue;
        if (mode == 1 && (ref_depth < cams[i].depth_min || ref_depth > cams[i].depth_max)) continue;
        float PointX[3];
        PointOnWorld(c, r, ref_depth, cams[i], PointX);
        int num_consistent = 0;
        float dynamic_consistency = 0;
        for (int j = 0; j < num_ngb; ++j) {
          const int src_id = srcs[j];
          const int src_cols = cams[src_id].width, src_rows = cams[src_id].height;
          float pu, pv, proj_depth;
          ProjectOnCamera(PointX, cams[src_id], pu, pv, proj_depth);
          const int src_r = round_index(pv);
          const int src_c = round_index(pu);
          if (src_c >= 0 && src_c < src_cols && src_r >= 0 && src_r < src_rows) {
            const size_t sc = (size_t)src_r * src_cols + src_c;
            if (masks[src_id][sc] == 1) continue;
            const float src_depth = depths[src_id][sc];
            const float* src_normal = normals[src_id] + 3 * sc;
            if (src_depth <= 0.0) continue;
            if (mode == 1 && (src_depth < cams[i].depth_min || src_depth > cams[i].depth_max)) continue;
            float tmp_X[3];
            PointOnWorld(src_c, src_r, src_depth, cams[src_id], tmp_X);
            float tu, tv;
            ProjectOnCamera(tmp_X, cams[i], tu, tv, proj_depth);
            const float reproj_error = ReprojError(c, r, tu, tv);
            const float relative_depth_diff = std::fabs(proj_depth - ref_depth) / ref_depth;
            const float angle = GetAngle(ref_normal, src_normal);
            if (reproj_error < 2.0f && relative_depth_diff < 0.01f && angle < 0.174533f) {
              used_x[j] = src_c;
              used_y[j] = src_r;
              const float tmp_index = reproj_error + 200 * relative_depth_diff + angle * 10;
              dynamic_consistency += orc::expf_hpm(-tmp_index);
              num_consistent++;
            }
          }
        }
        // RunFusion compares with the double 0.3 * n; RunFusion_Sky_Strict
        // (main.cpp:793-797) with `float factor = 0.3`, i.e. in float
        const bool pass = mode == 2 ? dynamic_consistency > 0.3f * (float)num_consistent
                                    : (double)dynamic_consistency > 0.3 * num_consistent;
        if (num_consistent >= 1 && pass) {
          if (mode != 1) {
            if (npts < capacity) {
              out_pairs[2 * npts] = i;
              out_pairs[2 * npts + 1] = (int32_t)center;
              if (out_points) {
                float* p = out_points + 6 * npts;
                p[0] = PointX[0]; p[1] = PointX[1]; p[2] = PointX[2];
                p[3] = ref_normal[0]; p[4] = ref_normal[1]; p[5] = ref_normal[2];
              }
            }
            ++npts;
          } else {
            consistency[i][center] = dynamic_consistency;
          }
          for (int j = 0; j < num_ngb; ++j) {
            if (used_x[j] == -1) continue;
            const int s = srcs[j];
            const size_t sc = (size_t)used_y[j] * cams[s].width + used_x[j];
            masks[s][sc] = 1;
            if (mode == 1) consistency[s][sc] = dynamic_consistency;
          }
        }
      }
    }
    if (mode == 1) std::memcpy(out_conf[i], consistency[i].data(), consistency[i].size() * sizeof(float));
  }
  *count = npts;
  return 0;
}

}  // extern "C"
