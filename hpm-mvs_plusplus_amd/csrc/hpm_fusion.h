// hpm_fusion.h -- multi-view consistency check of fusion and confidence
// evaluation (RunFusion main.cpp:823-957, ConfidenceEvaluation
// main.cpp:959-1109).
//
// The reference walks images in order and pixels row-major, and every accepted
// pixel masks the source pixels it matched, so later pixels skip them.  The
// expensive part does not depend on the masks: for each (reference pixel,
// source view) it is two projections, a reprojection, a normal angle and an
// exp.  k_consistency computes it for all pixels of one image at once and
// stores, per pixel, the sources that confirm it (slot j, matched source
// pixel, score exp(-(e + 200 dd + 10 angle))) in slot order, and their count.
// The greedy, ordered part -- masks, the score sum over unmasked matches, the
// acceptance test and the mask updates -- is a short sequential pass on the
// host over those lists, with the reference's float operations in the
// reference's order; a pixel no source confirms changes nothing and is
// skipped without reading its list.
//
// Arithmetic: plain IEEE float operations in the reference's order (no fused
// multiply-adds; the library is built with -ffp-contract=off); exp of a float
// is expf_hpm, acosf is acosf_hpm.  The reprojection error
// sqrt(pow(c - u, 2) + pow(r - v, 2)) (main.cpp:922,1064) is a double
// expression: main.cpp is host C++11, where pow(float, int) returns double, so
// the squares, their sum and the square root are double, rounded to float on
// assignment.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hpm_device.h"

namespace hpm {

// Get3DPointonWorld, HPM.cpp:221-245.
HPM_HD void point_on_world(int x, int y, float depth, const DevCam& c, float X[3]) {
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
HPM_HD void project_on_camera(const float X[3], const DevCam& c, float& u, float& v, float& depth) {
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
HPM_HD int round_index(float v) {
  const float x = v + 0.5f;
  return (x > -2147483648.0f && x < 2147483648.0f) ? (int)x : -1;
}

struct FusionImage {
  const float* depth;   // w * h
  const float* normal;  // 3 floats per pixel
};

constexpr int kFusionMaxImages = 1024;

// A source that confirms a pixel: its slot j (top 5 bits) and the matched
// source pixel m (low 27 bits), and the score.
constexpr int kMatchPixelBits = 27;
constexpr int kFusionMaxSources = 32;
struct MatchTerm {
  unsigned int jm;
  float term;
};

// One thread per pixel p0 + i (i < count) of reference image `ref`.  Its row
// out[i * num_ngb ...] holds the sources that confirm it, in slot order, and
// n_match[i] their number (pixel-major, so the host's ordered pass reads a
// pixel's matches contiguously and skips pixels without one).
__global__ __launch_bounds__(256) void k_consistency(const DevCam* __restrict__ cams, const FusionImage* __restrict__ imgs,
                                                     const int* __restrict__ srcs, int num_ngb, int ref, int mode,
                                                     int p0, int count, MatchTerm* __restrict__ out,
                                                     unsigned char* __restrict__ n_match) {
  const DevCam& rc = cams[ref];
  const int cols = rc.width;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int p = p0 + i;
  const int r = p / cols, c = p - r * cols;
  const float ref_depth = imgs[ref].depth[p];
  const float* rn = imgs[ref].normal + 3 * (size_t)p;
  bool ref_ok = ref_depth > 0.0f;
  if (mode == 1 && (ref_depth < rc.depth_min || ref_depth > rc.depth_max)) ref_ok = false;
  float X[3];
  point_on_world(c, r, ref_depth, rc, X);
  int n = 0;
  for (int j = 0; j < num_ngb && ref_ok; ++j) {
    const int s = srcs[j];
    const DevCam& sc = cams[s];
    float pu, pv, proj_depth;
    project_on_camera(X, sc, pu, pv, proj_depth);
    const int src_r = round_index(pv);
    const int src_c = round_index(pu);
    if (src_c >= 0 && src_c < sc.width && src_r >= 0 && src_r < sc.height) {
      const int sp = src_r * sc.width + src_c;
      const float src_depth = imgs[s].depth[sp];
      const float* sn = imgs[s].normal + 3 * (size_t)sp;
      bool ok = src_depth > 0.0f;
      if (mode == 1 && (src_depth < rc.depth_min || src_depth > rc.depth_max)) ok = false;
      if (ok) {
        float tX[3];
        point_on_world(src_c, src_r, src_depth, sc, tX);
        float tu, tv;
        project_on_camera(tX, rc, tu, tv, proj_depth);
        const float dx = (float)c - tu, dy = (float)r - tv;
        const float reproj_error = (float)sqrt((double)dx * (double)dx + (double)dy * (double)dy);
        const float relative_depth_diff = fabsf(proj_depth - ref_depth) / ref_depth;
        const float dot = rn[0] * sn[0] + rn[1] * sn[1] + rn[2] * sn[2];
        float angle = acosf_hpm(dot);  // GetAngle, HPM.cpp:281-290
        if (angle != angle) angle = 0.0f;
        if (reproj_error < 2.0f && relative_depth_diff < 0.01f && angle < 0.174533f) {
          const float tmp_index = reproj_error + 200 * relative_depth_diff + angle * 10;
          out[(size_t)i * num_ngb + n] = MatchTerm{((unsigned int)j << kMatchPixelBits) | (unsigned int)sp,
                                                   expf_hpm(-tmp_index)};
          ++n;
        }
      }
    }
  }
  n_match[i] = (unsigned char)n;
}

}  // namespace hpm

extern "C" {

int hpm_consistency_check(int mode, int num_images, const hpm_camera* cameras, const float* const* depths,
                          const float* const* normals, const int32_t* src_offsets, const int32_t* src_ids,
                          float* const* out_confidence, int32_t* out_pairs, float* out_points, int64_t capacity,
                          int64_t* out_count) {
  return hpm_consistency_check_cb(mode, num_images, cameras, depths, normals, src_offsets, src_ids, out_confidence,
                                  out_pairs, out_points, capacity, out_count, nullptr, nullptr);
}

int hpm_consistency_check_cb(int mode, int num_images, const hpm_camera* cameras, const float* const* depths,
                             const float* const* normals, const int32_t* src_offsets, const int32_t* src_ids,
                             float* const* out_confidence, int32_t* out_pairs, float* out_points, int64_t capacity,
                             int64_t* out_count, void (*image_done)(int image, void* user), void* user) {
  using namespace hpm;
  if ((mode != HPM_CHECK_FUSION && mode != HPM_CHECK_CONFIDENCE && mode != HPM_CHECK_FUSION_MASK) || num_images < 1 ||
      num_images > kFusionMaxImages ||
      !cameras || !depths || !normals || !src_offsets || !src_ids || !out_count)
    return HPM_E_INVALID;
  if (mode == HPM_CHECK_CONFIDENCE) {
    if (!out_confidence) return HPM_E_INVALID;
    for (int i = 0; i < num_images; ++i)
      if (!out_confidence[i]) return HPM_E_INVALID;
  }
  const bool fusion = mode != HPM_CHECK_CONFIDENCE;
  if (fusion && capacity > 0 && !out_pairs) return HPM_E_INVALID;
  if (src_offsets[0] != 0) return HPM_E_INVALID;
  size_t max_p = 0;
  int max_ngb = 0;
  for (int i = 0; i < num_images; ++i) {
    if (cameras[i].width <= 0 || cameras[i].height <= 0 || !depths[i] || !normals[i]) return HPM_E_INVALID;
    if ((size_t)cameras[i].width * cameras[i].height > ((size_t)1 << kMatchPixelBits)) return HPM_E_INVALID;
    const int n = src_offsets[i + 1] - src_offsets[i];
    if (n < 0 || n > kFusionMaxSources) return HPM_E_INVALID;
    for (int j = src_offsets[i]; j < src_offsets[i + 1]; ++j)
      if (src_ids[j] < 0 || src_ids[j] >= num_images) return HPM_E_INVALID;
    max_p = std::max(max_p, (size_t)cameras[i].width * cameras[i].height);
    max_ngb = std::max(max_ngb, n);
  }
  std::vector<DevCam> hc(num_images);
  for (int i = 0; i < num_images; ++i) std::memcpy(&hc[i], &cameras[i], sizeof(DevCam));

  DevCam* d_cams = nullptr;
  FusionImage* d_imgs = nullptr;
  int* d_srcs = nullptr;
  std::vector<float*> d_buf(2 * num_images, nullptr);
  int rc = HPM_OK;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && rc == HPM_OK) rc = (e == hipErrorOutOfMemory ? HPM_E_NOMEM : HPM_E_DEVICE);
    return rc == HPM_OK;
  };
  std::vector<FusionImage> himgs(num_images);
  const int total_src = src_offsets[num_images];
  // The per-pixel tests run on the GPU in chunks of rows, in the order the
  // ordered pass consumes them; each chunk is copied into one of a few
  // pinned host buffers while the host resolves an earlier one.
  const size_t chunk_px = std::min<size_t>(max_p, (size_t)1 << 20);
  const size_t chunk_elems = chunk_px * (size_t)std::max(max_ngb, 1);
  MatchTerm* d_out = nullptr;
  // kChunkBuffers chunks in flight: the GPU runs ahead of the ordered host
  // pass, also while other work keeps it busy
  constexpr int kChunkBuffers = 4;
  MatchTerm* h_out[kChunkBuffers] = {};
  unsigned char* d_cnt = nullptr;
  unsigned char* h_cnt[kChunkBuffers] = {};
  hipStream_t st = nullptr;
  hipEvent_t ev[kChunkBuffers] = {};
  if (ok(hipMalloc((void**)&d_cams, sizeof(DevCam) * num_images)) &&
      ok(hipMalloc((void**)&d_imgs, sizeof(FusionImage) * num_images)) &&
      ok(hipMalloc((void**)&d_srcs, sizeof(int) * std::max(total_src, 1))) &&
      ok(hipMalloc((void**)&d_out, sizeof(MatchTerm) * chunk_elems)) &&
      ok(hipMalloc((void**)&d_cnt, chunk_px)) &&
      ok(hipStreamCreate(&st))) {
    for (int b = 0; b < kChunkBuffers; ++b)
      if (!(ok(hipHostMalloc((void**)&h_out[b], sizeof(MatchTerm) * chunk_elems, hipHostMallocDefault)) &&
            ok(hipHostMalloc((void**)&h_cnt[b], chunk_px, hipHostMallocDefault)) &&
            ok(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming))))
        break;
    for (int i = 0; i < num_images && rc == HPM_OK; ++i) {
      const size_t P = (size_t)cameras[i].width * cameras[i].height;
      if (ok(hipMalloc((void**)&d_buf[2 * i], P * 4)) && ok(hipMalloc((void**)&d_buf[2 * i + 1], P * 12)) &&
          ok(hipMemcpyAsync(d_buf[2 * i], depths[i], P * 4, hipMemcpyHostToDevice, st)) &&
          ok(hipMemcpyAsync(d_buf[2 * i + 1], normals[i], P * 12, hipMemcpyHostToDevice, st)))
        himgs[i] = FusionImage{d_buf[2 * i], d_buf[2 * i + 1]};
    }
  }
  if (rc == HPM_OK && ok(hipMemcpyAsync(d_cams, hc.data(), sizeof(DevCam) * num_images, hipMemcpyHostToDevice, st)) &&
      ok(hipMemcpyAsync(d_imgs, himgs.data(), sizeof(FusionImage) * num_images, hipMemcpyHostToDevice, st)) &&
      total_src > 0)
    ok(hipMemcpyAsync(d_srcs, src_ids, sizeof(int) * total_src, hipMemcpyHostToDevice, st));
  if (rc == HPM_OK) ok(hipStreamSynchronize(st));  // the pageable sources are only read synchronously

  // chunks in consumption order: (image, first pixel, pixels)
  struct Chunk { int img; size_t p0, n; };
  std::vector<Chunk> chunks;
  for (int i = 0; i < num_images; ++i) {
    if (src_offsets[i + 1] - src_offsets[i] <= 0) continue;
    const size_t P = (size_t)cameras[i].width * cameras[i].height;
    for (size_t p0 = 0; p0 < P; p0 += chunk_px) chunks.push_back({i, p0, std::min(chunk_px, P - p0)});
  }
  auto issue = [&](size_t k) {
    const Chunk& c = chunks[k];
    const int num_ngb = src_offsets[c.img + 1] - src_offsets[c.img];
    hipLaunchKernelGGL(k_consistency, dim3((unsigned)((c.n + 255) / 256)), dim3(256), 0, st, d_cams, d_imgs,
                       d_srcs + src_offsets[c.img], num_ngb, c.img, mode, (int)c.p0, (int)c.n, d_out, d_cnt);
    return ok(hipGetLastError()) &&
           ok(hipMemcpyAsync(h_out[k % kChunkBuffers], d_out, sizeof(MatchTerm) * c.n * num_ngb,
                             hipMemcpyDeviceToHost, st)) &&
           ok(hipMemcpyAsync(h_cnt[k % kChunkBuffers], d_cnt, c.n, hipMemcpyDeviceToHost, st)) &&
           ok(hipEventRecord(ev[k % kChunkBuffers], st));
  };

  std::vector<std::vector<uint8_t>> masks(num_images);
  // ConfidenceEvaluation's in-memory consistency maps (cv::Mat::zeros,
  // main.cpp:1005).  out_confidence[i] receives image i's confidence.dmb:
  // the map as it stands when image i is done (main.cpp:1095-1100).  Later
  // images still write into the in-memory map (main.cpp:1086), not the file.
  std::vector<std::vector<float>> consistency(mode == HPM_CHECK_CONFIDENCE ? num_images : 0);
  int64_t npts = 0;
  for (int i = 0; i < num_images && rc == HPM_OK; ++i) {
    masks[i].assign((size_t)cameras[i].width * cameras[i].height, 0);
    if (mode == HPM_CHECK_CONFIDENCE) consistency[i].assign(masks[i].size(), 0.0f);
  }
  for (size_t k = 0; k < chunks.size() && k < (size_t)kChunkBuffers && rc == HPM_OK; ++k) issue(k);
  size_t k = 0;
  std::vector<int> used;
  // HPM_VERBOSE=1: time waiting for chunks vs resolving them, match counts
  const char* verbose_env = std::getenv("HPM_VERBOSE");
  const bool verbose = verbose_env && std::atoi(verbose_env) > 0;
  double t_wait = 0.0;
  int64_t n_pixels = 0, n_matched_pixels = 0, n_matches = 0;
  const auto t_begin = std::chrono::steady_clock::now();
  for (int i = 0; i < num_images && rc == HPM_OK; ++i) {
    const int cols = cameras[i].width, rows = cameras[i].height;
    const size_t P = (size_t)cols * rows;
    const int num_ngb = src_offsets[i + 1] - src_offsets[i];
    const int* srcs = src_ids + src_offsets[i];
    // The ordered greedy pass (main.cpp:884-952 / 1019-1088); used_list lives
    // for the whole image, as in the reference.
    used.assign(num_ngb, -1);
    size_t p_begin = 0, p_end = 0;
    const MatchTerm* mt = nullptr;
    const unsigned char* nm = nullptr;
    // chunk k consumed: its pinned buffer takes chunk k + kChunkBuffers
    auto release = [&]() {
      return !(k + kChunkBuffers < chunks.size() && !issue(k + kChunkBuffers)) && (++k, true);
    };
    for (size_t p = 0; p < P; ++p) {
      if (num_ngb > 0 && p == p_end) {  // next chunk of this image
        if (mt && !release()) break;
        const auto tw = std::chrono::steady_clock::now();
        if (!ok(hipEventSynchronize(ev[k % kChunkBuffers]))) break;
        t_wait += std::chrono::duration<double>(std::chrono::steady_clock::now() - tw).count();
        mt = h_out[k % kChunkBuffers];
        nm = h_cnt[k % kChunkBuffers];
        p_begin = chunks[k].p0;
        p_end = p_begin + chunks[k].n;
      }
      // a pixel no source confirms changes nothing (num_consistent stays 0)
      const int n_row = num_ngb > 0 ? nm[p - p_begin] : 0;
      if (verbose) {
        ++n_pixels;
        n_matched_pixels += n_row > 0;
        n_matches += n_row;
      }
      if (n_row == 0) continue;
      if (masks[i][p] == 1) continue;
      int num_consistent = 0;
      float dynamic_consistency = 0;
      const MatchTerm* row = mt + (p - p_begin) * num_ngb;
      for (int e = 0; e < n_row; ++e) {  // slots in increasing order, as the reference sums them
        const int j = (int)(row[e].jm >> kMatchPixelBits);
        const int m = (int)(row[e].jm & ((1u << kMatchPixelBits) - 1u));
        if (masks[srcs[j]][m] == 1) continue;
        used[j] = m;
        dynamic_consistency += row[e].term;
        num_consistent++;
      }
      // RunFusion compares with the double 0.3 * n; RunFusion_Sky_Strict
      // (main.cpp:793-797) with `float factor = 0.3`, i.e. in float
      const bool pass = mode == HPM_CHECK_FUSION_MASK ? dynamic_consistency > 0.3f * (float)num_consistent
                                                      : (double)dynamic_consistency > 0.3 * num_consistent;
      if (num_consistent >= 1 && pass) {
        if (fusion) {
          if (npts < capacity) {
            out_pairs[2 * npts] = i;
            out_pairs[2 * npts + 1] = (int32_t)p;
            if (out_points) {
              const int r = (int)(p / cols), c = (int)(p - (size_t)r * cols);
              float X[3];
              point_on_world(c, r, depths[i][p], hc[i], X);
              float* q = out_points + 6 * npts;
              q[0] = X[0]; q[1] = X[1]; q[2] = X[2];
              q[3] = normals[i][3 * p]; q[4] = normals[i][3 * p + 1]; q[5] = normals[i][3 * p + 2];
            }
          }
          ++npts;
        } else {
          consistency[i][p] = dynamic_consistency;
        }
        for (int j = 0; j < num_ngb; ++j) {
          if (used[j] < 0) continue;
          masks[srcs[j]][used[j]] = 1;
          if (mode == HPM_CHECK_CONFIDENCE) consistency[srcs[j]][used[j]] = dynamic_consistency;
        }
      }
    }
    if (mt && rc == HPM_OK) release();
    if (rc == HPM_OK && mode == HPM_CHECK_CONFIDENCE)
      std::memcpy(out_confidence[i], consistency[i].data(), P * sizeof(float));
    if (rc == HPM_OK && image_done) image_done(i, user);
  }
  if (verbose) {
    const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();
    std::fprintf(stderr,
                 "[hpm] consistency check: %.1f ms (waiting for chunks %.1f ms), %lld pixels, %.3f with a match, "
                 "%.2f matches per pixel\n",
                 1e3 * total, 1e3 * t_wait, (long long)n_pixels,
                 n_pixels ? (double)n_matched_pixels / n_pixels : 0.0, n_pixels ? (double)n_matches / n_pixels : 0.0);
    std::fflush(stderr);
  }
  if (st) hipStreamSynchronize(st);
  for (MatchTerm* h : h_out)
    if (h) hipHostFree(h);
  for (unsigned char* h : h_cnt)
    if (h) hipHostFree(h);
  hipFree(d_cnt);
  for (hipEvent_t e : ev)
    if (e) hipEventDestroy(e);
  if (st) hipStreamDestroy(st);
  hipFree(d_out);
  for (float* b : d_buf) hipFree(b);
  hipFree(d_cams); hipFree(d_imgs); hipFree(d_srcs);
  *out_count = npts;
  return rc;
}

}  // extern "C"
