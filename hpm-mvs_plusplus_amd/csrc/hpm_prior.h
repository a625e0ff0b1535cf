// hpm_prior.h -- planar-prior construction on the device, from the maps a
// context already holds (reloaded depths and costs, confidences):
//
//   k_prior_downsample  the depth, cost and confidence maps resized by
//                       1 / (2 distance) for a prior built on a coarser grid
//                       (cv::resize INTER_LINEAR of float maps, main.cpp:452-454)
//   k_support_points    GetSupportPoints_Classify_Check (HPM.cpp:1026-1073):
//                       per 5 x 5 block the first minimum of cost - conf and
//                       of cost + 0.2 - conf, each kept below 0.1
//   k_tri_planes        GetPriorPlaneParams[_factor] (HPM.cpp:1099-1167): the
//                       unit null vector of [X 1] through the three corners
//   k_raster            the barycentric raster of main.cpp:389-409 / 493-514:
//                       later triangles overwrite earlier ones, so each pixel
//                       keeps the largest index of the triangles that reach
//                       it (an atomic max gives the same as the ordered loop)
//   k_prior_level       distance 0: the depth range test on the plane's depth
//                       (main.cpp:411-431), then the per-pixel prior planes and
//                       triangle masks of CudaPlanarPriorInitialization
//                       (HPM.cpp:964-984)
//   k_prior_coarse      distance 1, 2: the coarse prior depth and world normal
//                       (main.cpp:516-546), upsampled by k_jbu_prior
//   k_prior_reload      the per-pixel reload of main.cpp:548-576 and
//                       ReloadPlanarPriorInitialization (HPM.cpp:1433-1453)
//   k_split_planes,     (not the prior) the output planes as a depth map and
//   k_join_planes       a normal map and back, for hpm_get_depth_normals and
//                       hpm_reload_hypotheses_maps
//   k_widen_edges       (not the prior) 8-bit edge map to 32 bits
//
// The Delaunay triangulation between the support points and the triangles
// stays on the host: cv::Subdiv2D inserts points one at a time and its result
// on co-circular points depends on that order.
//
// Arithmetic follows the host statements operation by operation (IEEE float
// and double, no contraction): the results equal the host construction's bit
// for bit.
#pragma once
#include "hpm_device.h"

namespace hpm {

// Map element (x, y) of a float map stored every `es` floats (es = 4 reads the
// depth in the w of the reloaded float4 hypotheses).
struct FloatMap {
  const float* base;
  int width, height, es;
  __device__ __forceinline__ float at(int x, int y) const { return base[((size_t)y * width + x) * es]; }
};

// One triangle as the raster and plane kernels read it, corners inside the
// map (the host checks), with its barycentric step 1 / longest edge computed
// on the host as the raster loop does.  Its raster index is its position.
struct PriorTri {
  int x1, y1, x2, y2, x3, y3;
  float step;
  int pad;
};

// cv::resize INTER_LINEAR of a float map as the host's ResizeLinear computes
// it: source position (d + 0.5) s - 0.5 in double, weights in float, border
// clamp.
__device__ __forceinline__ void resize_axis(int d, int sn, double s, int& i0, float& a) {
  const double f = ((double)d + 0.5) * s - 0.5;
  int i = (int)floor(f);
  float w = (float)(f - (double)i);
  if (i < 0) { i = 0; w = 0.0f; }
  if (i >= sn - 1) { i = sn - 1; w = 0.0f; }
  i0 = i;
  a = w;
}

__device__ __forceinline__ float resize_sample(const FloatMap& m, int xa, int ya, float ax, float ay) {
  const int xb = min(xa + 1, m.width - 1), yb = min(ya + 1, m.height - 1);
  const float top = m.at(xa, ya) + ax * (m.at(xb, ya) - m.at(xa, ya));
  const float bot = m.at(xa, yb) + ax * (m.at(xb, yb) - m.at(xa, yb));
  return top + ay * (bot - top);
}

__global__ __launch_bounds__(256) void k_prior_downsample(FloatMap depth, FloatMap cost, FloatMap conf, int dw, int dh,
                                                          double sx, double sy, float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= dw * dh) return;
  const int y = p / dw, x = p - y * dw;
  int xa, ya;
  float ax, ay;
  resize_axis(x, depth.width, sx, xa, ax);
  resize_axis(y, depth.height, sy, ya, ay);
  const size_t n = (size_t)dw * dh;
  out[p] = resize_sample(depth, xa, ya, ax, ay);
  out[n + p] = resize_sample(cost, xa, ya, ax, ay);
  out[2 * n + p] = resize_sample(conf, xa, ya, ax, ay);
}

// One thread per 5 x 5 block, blocks numbered column-major as the reference
// visits them; out[b] = (x, y) of the textured candidate, then (x, y) of the
// untextured one, -1 where a candidate is not kept.
__global__ __launch_bounds__(256) void k_support_points(FloatMap cost, FloatMap conf, int scan_w, int scan_h,
                                                        int4* __restrict__ out) {
  const int step = 5;
  const int nrb = (scan_h + step - 1) / step, ncb = (scan_w + step - 1) / step;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nrb * ncb) return;
  const int col = (b / nrb) * step, row = (b % nrb) * step;
  float min_no_texture = 2.0f, min_texture = 2.0f;
  int nx = -1, ny = -1, tx = -1, ty = -1;
  const int c_bound = min(scan_w, col + step), r_bound = min(scan_h, row + step);
  for (int c = col; c < c_bound; ++c) {
    for (int r = row; r < r_bound; ++r) {
      const float cst = cost.at(c, r), cf = conf.at(c, r);
      if (cst < 2.0f) {
        const float f_no_texture = cst - cf;
        const float f_texture = (float)((double)cst + 0.2 - (double)cf);
        if (min_no_texture > f_no_texture) { nx = c; ny = r; min_no_texture = f_no_texture; }
        if (min_texture > f_texture) { tx = c; ty = r; min_texture = f_texture; }
      }
    }
  }
  int4 o = make_int4(-1, -1, -1, -1);
  if ((double)min_texture < 0.1) { o.x = tx; o.y = ty; }
  if ((double)min_no_texture < 0.1) { o.z = nx; o.w = ny; }
  out[b] = o;
}

// Get3DPointonRefCam[_factor] (HPM.cpp:247-256) with factor 1 written as the
// unscaled form, so that the distance-0 prior performs exactly its operations.
__device__ __forceinline__ void corner_point(int x, int y, float depth, const float* K, bool scaled, float factor,
                                            double X[3]) {
  float px, py;
  if (scaled) {
    px = depth * ((float)x - K[2] * factor) / (K[0] * factor);
    py = depth * ((float)y - K[5] * factor) / (K[4] * factor);
  } else {
    px = depth * ((float)x - K[2]) / K[0];
    py = depth * ((float)y - K[5]) / K[4];
  }
  X[0] = px; X[1] = py; X[2] = depth;
}

// cv::SVD::solveZ of the 3 x 4 matrix [X 1]: its null vector is the
// generalised cross product of the rows (component k = (-1)^k times the minor
// without column k), in double, scaled to unit length; then the plane is
// scaled so that |n| = 1 and w >= 0 (HPM.cpp:1120-1131).
__device__ __forceinline__ float4 plane_through(const double A[3][4]) {
  double n[4];
  for (int k = 0; k < 4; ++k) {
    int c[3], m = 0;
    for (int j = 0; j < 4; ++j)
      if (j != k) c[m++] = j;
    const double det = A[0][c[0]] * (A[1][c[1]] * A[2][c[2]] - A[1][c[2]] * A[2][c[1]]) -
                       A[0][c[1]] * (A[1][c[0]] * A[2][c[2]] - A[1][c[2]] * A[2][c[0]]) +
                       A[0][c[2]] * (A[1][c[0]] * A[2][c[1]] - A[1][c[1]] * A[2][c[0]]);
    n[k] = (k % 2 ? -det : det);
  }
  const double s = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] + n[3] * n[3]);
  float n4[4];
  for (int k = 0; k < 4; ++k) n4[k] = (float)(n[k] / s);
  float norm2 = (float)sqrt((double)n4[0] * n4[0] + (double)n4[1] * n4[1] + (double)n4[2] * n4[2]);
  if (n4[3] < 0) norm2 *= -1;
  return make_float4(n4[0] / norm2, n4[1] / norm2, n4[2] / norm2, n4[3] / norm2);
}

__global__ __launch_bounds__(256) void k_tri_planes(const PriorTri* __restrict__ tris, int n, FloatMap depth, DevCam cam,
                                                    int scaled, float factor, float4* __restrict__ planes) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const PriorTri T = tris[t];
  double A[3][4];
  const int xs[3] = {T.x1, T.x2, T.x3}, ys[3] = {T.y1, T.y2, T.y3};
  for (int k = 0; k < 3; ++k) {
    corner_point(xs[k], ys[k], depth.at(xs[k], ys[k]), cam.K, scaled != 0, factor, A[k]);
    A[k][3] = 1.0;
  }
  planes[t] = plane_through(A);
}

// One thread per triangle; mask holds (raster index + 1), zero-initialised.
__global__ __launch_bounds__(256) void k_raster(const PriorTri* __restrict__ tris, int n, int w, int h,
                                                unsigned int* __restrict__ mask) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const PriorTri T = tris[t];
  const unsigned int v = (unsigned int)t + 1u;
  const float step = T.step;
  for (float p = 0; p < 1.0; p += step) {
    for (float q = 0; (double)q < 1.0 - p; q += step) {
      const int x = (int)((double)(p * (float)T.x1 + q * (float)T.x2) + (1.0 - p - q) * (double)T.x3);
      const int y = (int)((double)(p * (float)T.y1 + q * (float)T.y2) + (1.0 - p - q) * (double)T.y3);
      if (x >= 0 && x < w && y >= 0 && y < h) atomicMax(&mask[(size_t)y * w + x], v);
    }
  }
}

// The raster stores the index in a float map ((float)(idx + 1.0)) that is read
// back through (size_t) and (uint32_t) casts.
__device__ __forceinline__ unsigned int mask_value(unsigned int raw) { return (unsigned int)(float)(double)raw; }

// Distance 0: planes of the rastered triangles, kept where the plane's depth
// at the pixel lies in [depth_min, depth_max] (GetDepthFromPlaneParam,
// HPM.cpp:1169-1172).
__global__ __launch_bounds__(256) void k_prior_level(const float4* __restrict__ planes, int w, int h, DevCam cam,
                                                     float dmin, float dmax, unsigned int* __restrict__ mask,
                                                     float4* __restrict__ prior) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w * h) return;
  const int y = p / w, x = p - y * w;
  unsigned int m = mask_value(mask[p]);
  float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (m > 0) {
    const float4 pl = planes[m - 1];
    const float* K = cam.K;
    const float d = -pl.w * K[0] / (((float)x - K[2]) * pl.x + (K[0] / K[4]) * ((float)y - K[5]) * pl.y + K[0] * pl.z);
    if (!(d <= dmax && d >= dmin)) m = 0;
    else out = pl;
  }
  mask[p] = m;
  prior[p] = out;
}

// Distances 1, 2 on the coarse grid: the plane's depth there
// (GetDepthFromPlaneParam_factor, HPM.cpp:1174-1177), kept in
// [0.6 depth_min', 1.2 depth_max'] where depth_min' / depth_max' are the
// widened range, with the world normal (TransformNormal, HPM.cpp:1179-1187).
__global__ __launch_bounds__(256) void k_prior_coarse(const float4* __restrict__ planes, const unsigned int* __restrict__ mask,
                                                      int w, int h, DevCam cam, float factor, float dmin, float dmax,
                                                      float* __restrict__ prior_d, float4* __restrict__ prior_n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w * h) return;
  const int y = p / w, x = p - y * w;
  const unsigned int m = mask_value(mask[p]);
  float d_out = 0.0f;
  float4 n_out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (m > 0) {
    const float4 pl = planes[m - 1];
    const float* K = cam.K;
    const float d = -pl.w * (K[0] * factor) /
                    (((float)x - K[2] * factor) * pl.x + (K[0] / K[4]) * ((float)y - K[5] * factor) * pl.y +
                     K[0] * factor * pl.z);
    if (d <= dmax * 1.2f && d >= dmin * 0.6f) {
      const float* R = cam.R;
      d_out = d;
      n_out = make_float4(R[0] * pl.x + R[3] * pl.y + R[6] * pl.z, R[1] * pl.x + R[4] * pl.y + R[7] * pl.z,
                          R[2] * pl.x + R[5] * pl.y + R[8] * pl.z, d);
    }
  }
  prior_d[p] = d_out;
  prior_n[p] = n_out;
}

// The upsampled depth and world normal back to a reference-camera plane
// (TransformNormal2RefCam, GetDistance2Origin, HPM.cpp:1189-1206) where the
// depth lies in [depth_min, depth_max].  up_n may alias prior (same pixel).
__global__ __launch_bounds__(256) void k_prior_reload(const float* __restrict__ up_d, const float4* up_n, int w, int h,
                                                      DevCam cam, float dmin, float dmax, unsigned int* __restrict__ mask,
                                                      float4* prior) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w * h) return;
  const int y = p / w, x = p - y * w;
  const float d = up_d[p];
  float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  unsigned int m = 0;
  if (d <= dmax && d >= dmin) {
    m = 1;
    const float4 in = up_n[p];
    const float* R = cam.R;
    out.x = R[0] * in.x + R[1] * in.y + R[2] * in.z;
    out.y = R[3] * in.x + R[4] * in.y + R[5] * in.z;
    out.z = R[6] * in.x + R[7] * in.y + R[8] * in.z;
    const float* K = cam.K;
    const float X0 = d * ((float)x - K[2]) / K[0];
    const float X1 = d * ((float)y - K[5]) / K[4];
    out.w = -(out.x * X0 + out.y * X1 + out.z * d);
  }
  mask[p] = m;
  prior[p] = out;
}

// The planes after the output kernel (world normal xyz, depth w) as the
// host's depth map and 3-channel normal map (main.cpp:311-322).
__global__ __launch_bounds__(256) void k_split_planes(const float4* __restrict__ planes, size_t n,
                                                      float* __restrict__ depth, float* __restrict__ normals) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float4 v = planes[p];
  depth[p] = v.w;
  normals[3 * p] = v.x;
  normals[3 * p + 1] = v.y;
  normals[3 * p + 2] = v.z;
}

// The inverse, for hpm_reload_hypotheses_maps: (normal xyz, depth) planes
// from a depth map and a 3-channel normal map (HPM.cpp:942-962).
__global__ __launch_bounds__(256) void k_join_planes(const float* __restrict__ depth, const float* __restrict__ normals,
                                                     size_t n, float4* __restrict__ planes) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  planes[p] = make_float4(normals[3 * p], normals[3 * p + 1], normals[3 * p + 2], depth[p]);
}

// An 8-bit edge map widened to the 32-bit map the kernels read, for
// hpm_set_canny_u8.
__global__ __launch_bounds__(256) void k_widen_edges(const uint8_t* __restrict__ in, size_t n, uint32_t* __restrict__ out) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) out[p] = in[p];
}

}  // namespace hpm
