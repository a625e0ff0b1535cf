// oracle_host.cpp -- CPU restatement of the passes around PatchMatch that the
// GPU library also implements: joint bilateral upsampling and the multi-view
// consistency check of fusion / confidence evaluation.
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

}  // namespace

extern "C" {

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
