// hpm_host.hpp -- C++ host side of the MI355X PatchMatch library: the dense
// folder formats, the planar-prior construction, an edge detector, the `HPM`
// facade with the reference class's method names, and the per-view pass
// driver with its single-level schedule.  Everything that touches the GPU
// goes through the C ABI of include/hpm_mi355x.h.
#pragma once
#include <sys/mman.h>

#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <mutex>
#include <new>
#include <stdexcept>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hpm_mi355x.h"

namespace hpmhost {

// ---------------------------------------------------------------- formats --

struct Problem {  // main.h:49-58 (the fields the pass driver uses)
  int ref_image_id = 0;
  std::vector<int> src_image_ids;
};

// Storage of the per-pixel maps.  Blocks of 4 MiB and more (a 3072 x 2048
// float map is 24 MiB, and every pass allocates several) are 2 MiB aligned
// and advised as transparent huge pages, so that a fresh map faults in a few
// huge pages instead of thousands of small ones.
template <class T>
struct MapAllocator {
  using value_type = T;
  static constexpr size_t kHuge = (size_t)2 << 20;
  MapAllocator() = default;
  template <class U>
  MapAllocator(const MapAllocator<U>&) {}
  T* allocate(size_t n) {
    const size_t bytes = n * sizeof(T);
    if (bytes < 2 * kHuge) return static_cast<T*>(::operator new(bytes));
    const size_t rounded = (bytes + kHuge - 1) / kHuge * kHuge;
    void* p = std::aligned_alloc(kHuge, rounded);
    if (!p) throw std::bad_alloc();
    madvise(p, rounded, MADV_HUGEPAGE);  // advice only: ignored where huge pages are off
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t n) {
    if (n * sizeof(T) < 2 * kHuge) ::operator delete(p);
    else std::free(p);
  }
  template <class U>
  bool operator==(const MapAllocator<U>&) const { return true; }
  template <class U>
  bool operator!=(const MapAllocator<U>&) const { return false; }
};

template <class T>
struct Map {  // row-major image / per-pixel map
  int width = 0, height = 0, channels = 1;
  std::vector<T, MapAllocator<T>> data;
  Map() = default;
  Map(int w, int h, int c = 1, T v = T()) : width(w), height(h), channels(c), data((size_t)w * h * c, v) {}
  T& at(int x, int y, int c = 0) { return data[((size_t)y * width + x) * channels + c]; }
  const T& at(int x, int y, int c = 0) const { return data[((size_t)y * width + x) * channels + c]; }
  bool empty() const { return data.empty(); }
};
using MapF = Map<float>;
using MapU8 = Map<uint8_t>;

// GenerateSampleList, main.cpp:171-200: sources with score <= 0 are dropped.
std::vector<Problem> ReadPairFile(const std::string& path);
// ReadCamera, HPM.cpp:159-184 ("extrinsic" 4x4 of which 3 rows are read,
// "intrinsic" 3x3, then depth_min interval depth_num depth_max).  Width and
// height are not in the file; they stay 0 until the image is known.
hpm_camera ReadCamera(const std::string& path);
// readDepthDmb / readNormalDmb (HPM.cpp:291-380): int32 type (1), h, w, nb,
// then h*w*nb floats.  Return false on a missing file or another type.
bool ReadDmb(const std::string& path, MapF& out);
// writeDepthDmb / writeNormalDmb (HPM.cpp:323-407).
bool WriteDmb(const std::string& path, const MapF& m);

// Maps of a pipeline run.  While a MapStoreScope is alive, SaveMap keeps the
// map in memory for LoadMap and queues its file write to a background
// thread; Finish() waits for the files (and reports a failed write).  With
// no scope alive, LoadMap reads the file and SaveMap writes it at once.
// Runs in one process share the store of the first scope; when that scope
// ends (its files written), the others go on with the files directly.
using MapRef = std::shared_ptr<const MapF>;
MapRef LoadMap(const std::string& path);
void SaveMap(const std::string& path, MapF&& m);
class MapStoreScope {
 public:
  MapStoreScope();
  ~MapStoreScope();  // writes what is still queued
  MapStoreScope(const MapStoreScope&) = delete;
  MapStoreScope& operator=(const MapStoreScope&) = delete;
  void Finish();

 private:
  bool owner_ = false;  // false inside another scope, which keeps the store
};
// Binary PGM (P5) or PPM (P6), 8 bit.  The reference decodes JPEG with
// OpenCV; this reader takes the uncompressed Netpbm formats.
bool ReadNetpbm(const std::string& path, MapU8& out);
// JPEG (hpm_jpeg.cpp), Huffman-coded sequential or progressive: the pixels
// libjpeg's default decoding gives (accurate integer IDCT, fancy upsampling,
// BT.601 YCbCr -> RGB), as grey (1 channel) or R G B (3 channels).  False with
// a message for arithmetic-coded, lossless or 12-bit files and for corrupt data.
bool DecodeJpeg(const std::vector<uint8_t>& data, MapU8& out, std::string* err = nullptr);
bool ReadJpeg(const std::string& path, MapU8& out, std::string* err = nullptr);
bool ReadJpegSize(const std::string& path, int& w, int& h);
// Grey float image of an 8-bit image: grey as is; R G B with the ITU-R BT.601
// weights in Q14, (4899 R + 9617 G + 1868 B + 2^13) >> 14 -- the fixed point
// cv::cvtColor(COLOR_BGR2GRAY) uses for 8-bit images (InuputInitialization,
// HPM.cpp:660-661).
MapF ToGrayFloat(const MapU8& img);

// cv::resize(src, dst, size, 0, 0, INTER_LINEAR): each destination pixel
// samples the source at ((x + 0.5) sx - 0.5, (y + 0.5) sy - 0.5), sx = sw / dw,
// bilinearly, with coordinates clamped to the border.  8-bit results are
// rounded to the nearest integer (OpenCV interpolates 8-bit images in fixed
// point; results can differ by one level).
MapU8 ResizeLinear(const MapU8& src, int w, int h);
MapF ResizeLinear(const MapF& src, int w, int h);

// The dense folder's images at their original size and the scaled copies of
// the current pyramid level (PrepareScaleImages, main.cpp:107-161, kept in
// memory instead of temporary PNG files).
//
// What every pass of a level derives from its images is derived once per
// level: the grey float image, the edge map, the camera scaled to the level
// size and, after BuildPool, one device copy of every image shared by the
// passes' contexts.  The results are those of deriving them in every pass.
class SceneImages {
 public:
  explicit SceneImages(std::string dense_folder) : folder_(std::move(dense_folder)) {}
  ~SceneImages();
  SceneImages(const SceneImages&) = delete;
  SceneImages& operator=(const SceneImages&) = delete;
  const MapU8& Original(int id);
  // scales image `id` so that its larger side is at most target_max_side
  void PrepareLevel(int id, int target_max_side);
  // the current level's image (the original when no level was prepared)
  const MapU8& Current(int id);
  // its grey float image (cv::cvtColor + convertTo, main.cpp:645-648)
  const MapF& CurrentGray(int id);
  // its edge map (cv::Canny 50 / 150, main.cpp:287)
  const MapU8& CurrentEdges(int id);
  // camera `id` of cams/%08d_cam.txt with K scaled to the current size
  // (HPM.cpp:668-674)
  hpm_camera CurrentCamera(int id);
  // uploads the current grey images of `ids` to one device pool (replacing
  // the previous level's); Pool() is null until then
  void BuildPool(const std::vector<int>& ids, int device = 0);
  hpm_pool* Pool() const { return pool_; }
  // the image's index in the pool, -1 when it is not there
  int PoolIndex(int id) const;
 private:
  std::string folder_;
  std::vector<MapU8> orig_, cur_, edges_;
  std::vector<MapF> gray_;
  std::vector<bool> have_orig_, have_cur_, have_gray_, have_edges_;
  hpm_pool* pool_ = nullptr;
  std::vector<int> pool_index_;
  void ensure(int id);
  void drop_pool();
};

struct FusedPoint {  // PointList, main.h:92-96
  float coord[3];
  float normal[3];
  float color[3];  // B, G, R as the reference stores them
};
// ExportPointCloud (HPM.cpp:445-498): x y z + uchar r g b, in point order.
bool ExportPointCloud(const std::string& path, const std::vector<FusedPoint>& pts);
// ExportFusedPointCloud (HPM.cpp:409-443): finite points with a non-zero
// normal; x y z, unit normal, uchar r g b.
bool ExportFusedPointCloud(const std::string& path, const std::vector<FusedPoint>& pts);

// ------------------------------------------------------------ planar prior --

struct Point2i {
  int x = 0, y = 0;
  bool operator==(const Point2i& o) const { return x == o.x && y == o.y; }
};
struct Triangle {  // main.h:87-90
  Point2i pt1, pt2, pt3;
};
// The value types of the reference interface: CUDA's float4 / int2 and
// cv::Rect (x, y, width, height; contains() is half-open).
struct Float4 {
  float x = 0, y = 0, z = 0, w = 0;
};
struct Int2 {
  int x = 0, y = 0;
};
struct Rect {
  int x = 0, y = 0, width = 0, height = 0;
  bool contains(const Point2i& p) const { return p.x >= x && p.x < x + width && p.y >= y && p.y < y + height; }
};

// HPM::GetSupportPoints_Classify_Check, HPM.cpp:1026-1073.
std::vector<Point2i> GetSupportPoints(const MapF& costs, const MapF& confidences, int width, int height);
// HPM::DelaunayTriangulation, HPM.cpp:1075-1097: Delaunay triangles of the
// points inside [0, w) x [0, h) (duplicates ignored), grown inside the same
// bounding triangle as cv::Subdiv2D, whose vertices are then dropped.  Exact
// integer predicates; among co-circular configurations any valid
// triangulation.
std::vector<Triangle> DelaunayTriangulation(int width, int height, const std::vector<Point2i>& points);
// Pixels of each triangle in the reference's barycentric raster
// (main.cpp:389-404): mask(x, y) = index + 1 of the last triangle that wrote it.
void RasterTriangles(const std::vector<Triangle>& tris, MapF& mask);
// HPM::GetPriorPlaneParams, HPM.cpp:1134-1167: the plane n.X + w = 0 through
// the three back-projected corners, |n| = 1, w >= 0.
void PlaneThroughPoints(const float X1[3], const float X2[3], const float X3[3], float plane[4]);

// -------------------------------------------------------------- edge map --

// Edge map for the texture term (main.cpp:287 calls cv::Canny(image, 50, 150)):
// 3x3 Sobel gradients, L1 magnitude, non-maximum suppression along the
// quantised gradient direction and hysteresis between the two thresholds
// (Canny 1986).  255 = edge, 0 = none.
MapU8 CannyEdges(const MapU8& gray, float low, float high);

// ------------------------------------------------------------ HPM facade --

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// The reference's `class HPM` (HPM.h:57-140) over the C ABI: same method names,
// same meaning; failures throw hpmhost::Error instead of exiting.
class HPM {
 public:
  HPM();
  ~HPM();
  HPM(const HPM&) = delete;
  HPM& operator=(const HPM&) = delete;

  void SetGeomConsistencyParams(bool multi_geometry);
  void SetPlanarPriorParams();
  void SetHierarchyParams();
  void SetMandConsistencyParams(bool flag);
  void SetPrecision(int precision);  // HPM_PREC_*; the reference has FP32 only

  // InuputInitialization (HPM.cpp:644-733): grey images and cameras of the
  // problem's reference and sources from `images/%08d.{pgm,ppm}` and
  // `cams/%08d_cam.txt`; source depth maps in geometric mode.
  void InuputInitialization(const std::string& dense_folder, const std::vector<Problem>& problems, int idx,
                            SceneImages* images = nullptr);
  // CudaSpaceInitialization (HPM.cpp:741-906): uploads; reloads the
  // reference's own previous maps in geometric mode and the coarse level's
  // maps in hierarchy mode.
  void CudaSpaceInitialization(const std::string& dense_folder, const Problem& problem);
  void CudaCannyInitialization(const MapU8& canny);
  void CudaConfidenceInitialization(const std::string& dense_folder, const std::vector<Problem>& problems, int idx);
  void TextureInformationInitialization() {}
  void CudaHypothesesReload(const MapF& depths, const MapF& costs, const MapF& normals);
  // The planar prior of a prior pass at `distance` 0, 1 or 2 built on the
  // device from the reloaded hypotheses and confidences (after
  // CudaHypothesesReload and CudaConfidenceInitialization): support points
  // and everything after the triangulation run there, the Delaunay
  // triangulation on the host.
  void BuildPlanarPrior(int distance);
  void CudaPlanarPriorInitialization(const std::vector<float>& plane_params, const MapF& masks);
  void ReloadPlanarPriorInitialization(const MapF& masks, const std::vector<float>& prior_planes);
  void RunPatchMatch(uint64_t seed);
  // RunPatchMatch() (HPM.cu:2987-3035) with the seed of SetSeed (0 by
  // default); the reference seeds curand with clock64(), this library
  // seeds its counter-based generator with the seed.
  void RunPatchMatch() { RunPatchMatch(seed_); }
  void SetSeed(uint64_t seed) { seed_ = seed; }

  int GetReferenceImageWidth() const { return cams_.empty() ? 0 : cams_[0].width; }
  int GetReferenceImageHeight() const { return cams_.empty() ? 0 : cams_[0].height; }
  const MapF& GetReferenceImage() const { return *images_.at(0); }
  const hpm_camera& GetCamera(int i) const { return cams_.at(i); }
  float GetMinDepth() const;
  float GetMaxDepth() const;
  // Bulk getters (GetPlaneHypothesis / GetCost / GetTexture for every pixel).
  void GetPlanes(std::vector<float>& planes) const;  // (world normal, depth)
  // depth map and world normal map of the result (main.cpp:311-322)
  void GetDepthNormals(MapF& depths, MapF& normals) const;
  void GetCosts(MapF& costs) const;
  void GetTexture(MapF& texture) const;

  // GetPriorPlaneParams (HPM.cpp:1134-1167): plane of a triangle's corners in
  // the reference camera frame.
  void GetPriorPlaneParams(const Triangle& t, const MapF& depths, float plane[4]) const;
  // the _factor variants (HPM.cpp:1099-1132, 1174-1177): the reference camera
  // scaled by `factor`
  void GetPriorPlaneParamsFactor(const Triangle& t, const MapF& depths, float factor, float plane[4]) const;
  float GetDepthFromPlaneParamFactor(const float plane[4], int x, int y, float factor) const;
  float GetDepthFromPlaneParam(const float plane[4], int x, int y) const;
  void TransformNormal(const float in[4], float out[4]) const;
  void TransformNormal2RefCam(const float in[4], float out[4]) const;
  float GetDistance2Origin(int x, int y, float depth, const float normal[4]) const;

  // The reference's signatures of the same (HPM.h:74-101), for ProcessProblem
  // (main.cpp:257-631) unchanged.  The per-pixel getters read host copies of
  // the planes, costs and texture taken on the first call after a run
  // (RunPatchMatch's D2H copies, HPM.cu:3029-3033).
  Float4 GetPlaneHypothesis(int index);  // (world normal, depth)
  float GetCost(int index);
  float GetTexture(int index);
  // support points of a (downsampled) map whose size is the reference size x
  // hpm_factor, truncated (HPM.cpp:1026-1073); texture is not used there
  void GetSupportPoints_Classify_Check(std::vector<Point2i>& support2DPoints, const MapF& costs,
                                       const MapF& confidences, const MapF& texture, float hpm_factor);
  std::vector<Triangle> DelaunayTriangulation(const Rect& boundRC, const std::vector<Point2i>& points);
  Float4 GetPriorPlaneParams(const Triangle& triangle, const MapF& depths) const;
  Float4 GetPriorPlaneParams_factor(const Triangle& triangle, const MapF& depths, float factor) const;
  float GetDepthFromPlaneParam(const Float4& plane, int x, int y) const;
  float GetDepthFromPlaneParam_factor(const Float4& plane, int x, int y, float factor) const;
  Float4 TransformNormal(const Float4& plane) const;
  Float4 TransformNormal2RefCam(const Float4& plane) const;
  float GetDistance2Origin(const Int2& p, float depth, const Float4& normal) const;
  void CudaPlanarPriorInitialization(const std::vector<Float4>& plane_params, const MapF& masks);
  void ReloadPlanarPriorInitialization(const MapF& masks, const Float4* prior_plane_parameters);
  // JointBilateralUpsampling_prior (HPM.cpp:1323-1392): the coarse prior
  // depth and 3-channel normal maps upsampled to the reference image's size,
  // guided by it (JBU_cu_prior, HPM.cu:3130-3197).
  void JointBilateralUpsampling_prior(const MapF& scaled_image_float, const MapF& src_depthmap,
                                      MapF& upsample_depthmap, const MapF& src_normal, MapF& upsample_normal);
  // CudaPlanarPriorRelease (HPM.cpp:589-597), CudaSpaceRelease
  // (HPM.cpp:599-633): device buffers freed; the getters keep working from
  // their host copies.  ReleaseProblemHostMemory (HPM.cpp:635-642): images,
  // cameras and source depths dropped.
  void CudaPlanarPriorRelease();
  void CudaSpaceRelease(bool geom_consistency);
  void ReleaseProblemHostMemory();

 private:
  void check(int rc, const char* what) const;
  hpm_ctx* ctx_ = nullptr;
  hpm_batch* fused_ = nullptr;  // one-view fused launch of large views, made on first use
  hpm_params params_{};
  bool multi_geometry_ = false;
  std::vector<hpm_camera> cams_;
  std::vector<const MapF*> images_;  // into owned_images_ or a SceneImages cache
  std::vector<MapF> owned_images_;
  std::vector<MapRef> depths_;
  hpm_pool* pool_ = nullptr;  // the level's device images, when every view is there
  std::vector<int32_t> pool_ids_;
  uint64_t seed_ = 0;
  // host copies for the per-pixel getters, valid from the first getter call
  // after a run until the next run
  void FetchResults();
  bool have_results_ = false;
  bool ran_ = false;  // a run's results are on the device (until CudaSpaceRelease)
  std::vector<float> planes_host_, costs_host_, texture_host_;
};

// ---------------------------------------------------------- pass driver --

struct PassFlags {
  bool geom_consistency = false;
  bool prior_consistency = false;
  bool mand_consistency = false;
  bool hierarchy = false;
  bool multi_geometry = false;
  int image_scale = 0;          // the pyramid level (texture{k}.dmb index)
  int hpm_scale_distance = 0;   // prior passes: 0 at the level, 1 or 2 levels coarser
};

// ProcessProblem (main.cpp:237-634): one pass of one reference view at the
// current pyramid level, reading and writing the reference's files under
// <dense_folder>/HPM_MVS_plusplus/2333_%08d/.  `images` holds the level's
// images (nullptr: the originals from disk).
void ProcessProblem(const std::string& dense_folder, const std::vector<Problem>& problems, int idx,
                    const PassFlags& flags, int precision, uint64_t seed, SceneImages* images = nullptr);
// JointBilateralUpsampling (main.cpp:636-657) + RunJBU (HPM.cpp:1266-1319): the
// coarser level's depths_geom.dmb upsampled to the current level's size,
// guided by the reference image, written as depths.dmb.
void JointBilateralUpsampling(const std::string& dense_folder, const Problem& problem, SceneImages& images);
// ConfidenceEvaluation (main.cpp:959-1109) -> confidence.dmb per view.
// Views whose confidence map a running ConfidenceEvaluation has finished
// (saved with SaveMap): a pass started alongside it waits per view.
class ReadyViews {
 public:
  explicit ReadyViews(int n);
  void Mark(int i);
  void Fail(const std::string& msg);  // wakes the waiters, which throw
  void Wait(int i);

 private:
  mutable std::mutex mu_;
  std::condition_variable cv_;
  std::vector<char> ready_;
  std::string error_;
};
void ConfidenceEvaluation(const std::string& dense_folder, const std::vector<Problem>& problems, bool geom,
                          ReadyViews* ready = nullptr);
// RunFusion (main.cpp:823-957) -> HPM_MVS_plusplus.ply and fused.ply; returns
// the fused points in the reference's order.
std::vector<FusedPoint> RunFusion(const std::string& dense_folder, const std::vector<Problem>& problems, bool geom);
// RunFusion_Sky_Strict (main.cpp:659-821, --mask): the same consistency pass,
// then a fused point is dropped when its pixel in masks/%08d.{ppm,pgm}
// (resized to the depth map) has B == 234 or G == 235 or R == 55; the sources
// are masked either way.  Writes HPM_MVS_plusplus_mask.ply and fused.ply.
std::vector<FusedPoint> RunFusionSkyStrict(const std::string& dense_folder, const std::vector<Problem>& problems,
                                           bool geom);
// The single-level schedule of main.cpp:1249-1291: photometric pass, confidence,
// prior pass, three geometric passes (the last two with mandatory
// consistency after a confidence evaluation), fusion.
std::vector<FusedPoint> RunPipeline(const std::string& dense_folder, int precision, uint64_t seed);
// The pyramid schedule of main.cpp:1183-1352: images scaled by
// 1 / initial_image_scale, halved until the larger side is <= size_bound
// (ComputeMultiScaleSettings, main.cpp:202-235; the reference's bound is 1000);
// coarsest level as RunPipeline, finer levels JBU -> hierarchy pass -> prior
// passes at distances max..0 -> three geometric passes; then fusion.
std::vector<FusedPoint> RunPyramidPipeline(const std::string& dense_folder, int initial_image_scale, int size_bound,
                                           int precision, uint64_t seed, bool mask_fusion = false);

std::string ResultFolder(const std::string& dense_folder, int ref_image_id);

// Wall time spent in each stage of the passes above, summed since the last
// reset (process-wide).  Stages: setup (image and camera loading, device
// upload, map loading), edges (Canny), prior (support points, Delaunay,
// planes, prior JBU), patchmatch (RunPatchMatch and read-back), write (.dmb
// output), jbu (JointBilateralUpsampling), confidence, fusion.
enum Stage { kSetup, kEdges, kPrior, kPatchMatch, kWrite, kJbu, kConfidence, kFusion, kStageCount };
struct StageTimes {
  double seconds[kStageCount];
  long long calls[kStageCount];
};
StageTimes GetStageTimes();
void ResetStageTimes();
// 0: quiet; 1: one line per ProcessProblem on stderr (view, pass, stage times)
void SetVerbose(int level);
// Host threads that run the views of an order-free pass concurrently (each
// with its own context and stream; 0: min(8, hardware threads), 1: in view
// order).  Every pass except the geometric passes over depths_geom.dmb is
// order-free: its views read no file that another view of the pass writes,
// so the files it produces do not depend on the thread count.
void SetPassThreads(int n);
// Where the prior passes build the planar prior: 1 (default) on the device
// (hpm_prior_support_points / hpm_prior_build; only the Delaunay
// triangulation runs on the host), 0 on the host.  The files are identical.
void SetPriorOnDevice(int enable);
// Reference views of at least n pixels run as a one-view fused launch with
// 16-column tiles (mixed-precision engine only); default 4,000,000.
void SetFusedMinPixels(int64_t n);
// Whether the pipeline runs a pass alongside the confidence evaluation
// before it, each view starting when its confidence map is written (1,
// default), or after it (0).  The files are identical.
void SetConfidenceOverlap(int enable);

}  // namespace hpmhost
