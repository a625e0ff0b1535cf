// hpm_host_capi.cpp -- the C ABI of include/hpm_pipeline.h over hpm_host.
#include <cstring>
#include <string>

#include "../../include/hpm_pipeline.h"
#include "hpm_host.hpp"

namespace {
thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_error.clear();
    return HPM_OK;
  } catch (const hpmhost::Error& e) {
    g_error = e.what();
    return HPM_E_INVALID;
  } catch (const std::exception& e) {
    g_error = e.what();
    return HPM_E_DEVICE;
  }
}

// An argument check that fails before any work: HPM_E_INVALID, with the
// function's name as the message.
int bad_args(const char* fn) {
  g_error = std::string(fn) + ": invalid argument";
  return HPM_E_INVALID;
}
}  // namespace

extern "C" {

const char* hpm_host_last_error(void) { return g_error.c_str(); }

int hpm_host_process_problem(const char* dense_folder, int idx, int geom, int prior, int mand, int multi_geometry,
                             int precision, uint64_t seed) {
  if (!dense_folder) return bad_args(__func__);
  return guarded([&] {
    const auto problems = hpmhost::ReadPairFile(std::string(dense_folder) + "/pair.txt");
    if (idx < 0 || idx >= (int)problems.size()) throw hpmhost::Error("problem index out of range");
    hpmhost::PassFlags f;
    f.geom_consistency = geom != 0;
    f.prior_consistency = prior != 0;
    f.mand_consistency = mand != 0;
    f.multi_geometry = multi_geometry != 0;
    hpmhost::ProcessProblem(dense_folder, problems, idx, f, precision, seed);
  });
}

int hpm_host_confidence(const char* dense_folder, int geom) {
  if (!dense_folder) return bad_args(__func__);
  return guarded([&] {
    hpmhost::ConfidenceEvaluation(dense_folder, hpmhost::ReadPairFile(std::string(dense_folder) + "/pair.txt"),
                                  geom != 0);
  });
}

int hpm_host_fusion(const char* dense_folder, int geom, int64_t* num_points) {
  if (!dense_folder || !num_points) return bad_args(__func__);
  return guarded([&] {
    *num_points = (int64_t)hpmhost::RunFusion(dense_folder,
                                               hpmhost::ReadPairFile(std::string(dense_folder) + "/pair.txt"),
                                               geom != 0).size();
  });
}

int hpm_host_run_pipeline(const char* dense_folder, int precision, uint64_t seed, int64_t* num_points) {
  if (!dense_folder || !num_points) return bad_args(__func__);
  return guarded([&] { *num_points = (int64_t)hpmhost::RunPipeline(dense_folder, precision, seed).size(); });
}

int hpm_host_run_pyramid_pipeline(const char* dense_folder, int initial_image_scale, int size_bound, int precision,
                                  uint64_t seed, int mask_fusion, int64_t* num_points) {
  if (!dense_folder || !num_points || initial_image_scale < 1 || size_bound < 1) return bad_args(__func__);
  return guarded([&] {
    *num_points = (int64_t)hpmhost::RunPyramidPipeline(dense_folder, initial_image_scale, size_bound, precision, seed,
                                                       mask_fusion != 0)
                      .size();
  });
}

int hpm_host_fusion_mask(const char* dense_folder, int geom, int64_t* num_points) {
  if (!dense_folder || !num_points) return bad_args(__func__);
  return guarded([&] {
    *num_points = (int64_t)hpmhost::RunFusionSkyStrict(
                      dense_folder, hpmhost::ReadPairFile(std::string(dense_folder) + "/pair.txt"), geom != 0)
                      .size();
  });
}

int hpm_host_stage_times(double* seconds, int64_t* calls, int capacity) {
  const hpmhost::StageTimes t = hpmhost::GetStageTimes();
  for (int i = 0; i < capacity && i < hpmhost::kStageCount; ++i) {
    if (seconds) seconds[i] = t.seconds[i];
    if (calls) calls[i] = t.calls[i];
  }
  return hpmhost::kStageCount;
}

void hpm_host_reset_stage_times(void) { hpmhost::ResetStageTimes(); }

void hpm_host_set_verbose(int level) { hpmhost::SetVerbose(level); }

void hpm_host_set_threads(int n) { hpmhost::SetPassThreads(n); }

void hpm_host_set_prior_on_device(int enable) { hpmhost::SetPriorOnDevice(enable); }

void hpm_host_set_fused_min_pixels(int64_t n) { hpmhost::SetFusedMinPixels(n); }

void hpm_host_set_confidence_overlap(int enable) { hpmhost::SetConfidenceOverlap(enable); }

int hpm_host_decode_jpeg(const uint8_t* data, int64_t size, int* w, int* h, int* channels, uint8_t* out,
                         int64_t capacity) {
  if (!data || size <= 0 || !w || !h || !channels) return bad_args(__func__);
  return guarded([&] {
    hpmhost::MapU8 img;
    std::string err;
    if (!hpmhost::DecodeJpeg(std::vector<uint8_t>(data, data + size), img, &err)) throw hpmhost::Error(err);
    *w = img.width;
    *h = img.height;
    *channels = img.channels;
    if (!out) return;  // size query
    if (capacity < (int64_t)img.data.size()) throw hpmhost::Error("output buffer smaller than w * h * channels");
    std::memcpy(out, img.data.data(), img.data.size());
  });
}

int hpm_host_support_points(const float* costs, const float* confidences, int w, int h, int32_t* out_xy,
                            int capacity, int* count) {
  if (!costs || !confidences || w <= 0 || h <= 0 || !count || (capacity > 0 && !out_xy)) return bad_args(__func__);
  return guarded([&] {
    hpmhost::MapF c(w, h), f(w, h);
    std::memcpy(c.data.data(), costs, sizeof(float) * w * h);
    std::memcpy(f.data.data(), confidences, sizeof(float) * w * h);
    const auto pts = hpmhost::GetSupportPoints(c, f, w, h);
    *count = (int)pts.size();
    for (int i = 0; i < (int)pts.size() && i < capacity; ++i) {
      out_xy[2 * i] = pts[i].x;
      out_xy[2 * i + 1] = pts[i].y;
    }
  });
}

int hpm_host_delaunay(int w, int h, const int32_t* xy, int n, int32_t* out_tris, int capacity, int* count) {
  if (w <= 0 || h <= 0 || n < 0 || (n > 0 && !xy) || !count || (capacity > 0 && !out_tris)) return bad_args(__func__);
  if (w > (1 << 20) || h > (1 << 20)) return bad_args(__func__);
  return guarded([&] {
    std::vector<hpmhost::Point2i> pts(n);
    for (int i = 0; i < n; ++i) pts[i] = {xy[2 * i], xy[2 * i + 1]};
    const auto tris = hpmhost::DelaunayTriangulation(w, h, pts);
    *count = (int)tris.size();
    for (int i = 0; i < (int)tris.size() && i < capacity; ++i) {
      const auto& t = tris[i];
      const int32_t v[6] = {t.pt1.x, t.pt1.y, t.pt2.x, t.pt2.y, t.pt3.x, t.pt3.y};
      std::memcpy(out_tris + 6 * i, v, sizeof(v));
    }
  });
}

int hpm_host_raster(const int32_t* tris, int n, int w, int h, float* mask) {
  if (w <= 0 || h <= 0 || n < 0 || (n > 0 && !tris) || !mask) return bad_args(__func__);
  return guarded([&] {
    std::vector<hpmhost::Triangle> t(n);
    for (int i = 0; i < n; ++i)
      t[i] = {{tris[6 * i], tris[6 * i + 1]}, {tris[6 * i + 2], tris[6 * i + 3]}, {tris[6 * i + 4], tris[6 * i + 5]}};
    hpmhost::MapF m(w, h);
    hpmhost::RasterTriangles(t, m);
    std::memcpy(mask, m.data.data(), sizeof(float) * w * h);
  });
}

int hpm_host_plane_through_points(const float* X1, const float* X2, const float* X3, float* plane) {
  if (!X1 || !X2 || !X3 || !plane) return bad_args(__func__);
  return guarded([&] { hpmhost::PlaneThroughPoints(X1, X2, X3, plane); });
}

int hpm_host_canny(const uint8_t* image, int w, int h, int channels, float low, float high, uint8_t* out) {
  if (!image || !out || w <= 0 || h <= 0 || (channels != 1 && channels != 3)) return bad_args(__func__);
  return guarded([&] {
    hpmhost::MapU8 g(w, h, channels);
    std::memcpy(g.data.data(), image, (size_t)w * h * channels);
    const auto e = hpmhost::CannyEdges(g, low, high);
    std::memcpy(out, e.data.data(), (size_t)w * h);
  });
}

}  // extern "C"
