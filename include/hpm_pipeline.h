/* hpm_pipeline.h -- C ABI of the host-side passes around PatchMatch: the dense
 * folder pipeline (ProcessProblem, ConfidenceEvaluation, RunFusion and the
 * single-level schedule of main.cpp) and the planar-prior construction steps
 * of class HPM.  Library: hpm-mvs_plusplus_amd/lib/libhpm_host.so (links
 * libhpm_mi355x.so).  Every function returns HPM_OK (0) or a negative HPM_E*
 * status; hpm_host_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Dense folder layout (the reference's): images/%08d.pgm|.ppm (binary
 * Netpbm; the reference reads JPEG through OpenCV), cams/%08d_cam.txt,
 * pair.txt; results in HPM_MVS_plusplus/2333_%08d/{depths,depths_geom,normals,
 * costs,confidence,texture0}.dmb and HPM_MVS_plusplus/{HPM_MVS_plusplus,fused}.ply.
 */
#ifndef HPM_PIPELINE_H_
#define HPM_PIPELINE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* hpm_host_last_error(void);

/* ProcessProblem (main.cpp:237-634) for problem `idx` of pair.txt at one
 * pyramid level (prior passes at distance 0). */
int hpm_host_process_problem(const char* dense_folder, int idx, int geom_consistency, int prior_consistency,
                             int mand_consistency, int multi_geometry, int precision, uint64_t seed);
/* ConfidenceEvaluation (main.cpp:959-1109): writes confidence.dmb per view. */
int hpm_host_confidence(const char* dense_folder, int geom_consistency);
/* RunFusion (main.cpp:823-957): writes the two PLY files; *num_points = fused points. */
int hpm_host_fusion(const char* dense_folder, int geom_consistency, int64_t* num_points);
/* RunFusion_Sky_Strict (main.cpp:659-821, the --mask option): reads
 * masks/%08d.ppm (or .pgm) per view, drops fused points whose mask pixel has
 * B == 234 or G == 235 or R == 55; writes HPM_MVS_plusplus_mask.ply and
 * fused.ply.  Also the reference's --fusion-only with --mask; hpm_host_fusion
 * is --fusion-only without it (main.cpp:1196-1206). */
int hpm_host_fusion_mask(const char* dense_folder, int geom_consistency, int64_t* num_points);
/* The single-level schedule of main.cpp:1249-1291 followed by fusion. */
int hpm_host_run_pipeline(const char* dense_folder, int precision, uint64_t seed, int64_t* num_points);
/* The pyramid schedule of main.cpp:1183-1352 followed by fusion
 * (main.cpp:1353-1359 without --mask): images scaled by 1/initial_image_scale
 * (the reference's --scale, default 4), halved until the larger side is
 * <= size_bound (ComputeMultiScaleSettings, main.cpp:202-235, where the bound
 * is fixed at 1000).  Coarsest level: photometric, prior and geometric
 * passes; each finer level: JBU of the coarser depths_geom, a hierarchy pass,
 * prior passes from 2 and 1 levels coarser down to the level itself, and three
 * geometric passes.  mask_fusion != 0 ends with hpm_host_fusion_mask's
 * fusion (--mask, main.cpp:1353-1356) instead of RunFusion. */
int hpm_host_run_pyramid_pipeline(const char* dense_folder, int initial_image_scale, int size_bound, int precision,
                                  uint64_t seed, int mask_fusion, int64_t* num_points);

/* Wall time of the pipeline stages, summed process-wide since the last reset
 * (no counterpart in the reference, which prints progress lines).  Stage order:
 * 0 setup (image/camera loading, device upload, map loading), 1 edges (Canny),
 * 2 prior (support points, Delaunay, planes, prior JBU), 3 patchmatch
 * (RunPatchMatch and read-back), 4 write (.dmb output), 5 jbu, 6 confidence,
 * 7 fusion.  Fills up to `capacity` entries; returns the number of stages. */
int hpm_host_stage_times(double* seconds, int64_t* calls, int capacity);
void hpm_host_reset_stage_times(void);
/* 0 quiet; 1: one line per ProcessProblem on stderr (view, pass, stage times). */
void hpm_host_set_verbose(int level);
/* Host threads running the views of a pass concurrently, each view with its
 * own context and stream (0: min(8, hardware threads); 1: view order).  Only
 * passes whose views read no file another view of the pass writes run
 * concurrently (all but the geometric passes over depths_geom.dmb, which the
 * reference runs in view order); the output files do not depend on it. */
void hpm_host_set_threads(int n);

/* Where prior passes build the planar prior (main.cpp:324-597): 1 (default)
 * on the device with hpm_prior_support_points / hpm_prior_build, only the
 * Delaunay triangulation on the host; 0 entirely on the host.  The output
 * files are identical. */
void hpm_host_set_prior_on_device(int enable);

/* The size from which a pass runs a reference view as a one-view fused launch
 * (hpm_batch) with 16-column colour tiles and 4-tile XCD strips instead of the
 * context's own launch: reference views of at least `n` pixels, default
 * 4,000,000, mixed-precision engine only (the exact engine has one tile
 * width).  At 3072 x 2048 with 10 sources this takes 7-20 % less time per
 * RunPatchMatch.  The results are identical either way.  No reference
 * counterpart: the reference has one launch shape (HPM.cu:2987-3035). */
void hpm_host_set_fused_min_pixels(int64_t n);

/* Whether the pipelines run each pass that follows a confidence evaluation
 * alongside it, each view starting once its own confidence map is written
 * (1, default), or after it (0).  The output files are identical. */
void hpm_host_set_confidence_overlap(int enable);

/* JPEG decoding as the pipeline reads images/%08d.jpg, which the reference
 * reads with cv::imread (main.cpp:125, main.cpp:213): Huffman-coded
 * sequential or progressive, 8-bit, the pixels of libjpeg's default decoding
 * (accurate integer IDCT, fancy upsampling, BT.601 YCbCr -> RGB).  *w, *h,
 * *channels (1 grey, 3 R G B) are set; `out` (NULL to query the size)
 * receives w*h*channels bytes and must hold that many (`capacity`).
 * HPM_E_INVALID for data it cannot decode (arithmetic coding, lossless,
 * 12-bit, corrupt) or a short buffer, with the reason in
 * hpm_host_last_error(). */
int hpm_host_decode_jpeg(const uint8_t* data, int64_t size, int* w, int* h, int* channels, uint8_t* out,
                         int64_t capacity);

/* GetSupportPoints_Classify_Check (HPM.cpp:1026-1073) at factor 1: costs and
 * confidences w*h; out_xy receives up to `capacity` (x, y) pairs. */
int hpm_host_support_points(const float* costs, const float* confidences, int w, int h, int32_t* out_xy,
                            int capacity, int* count);
/* DelaunayTriangulation (HPM.cpp:1075-1097): n points (x, y) inside w x h;
 * out_tris receives up to `capacity` triangles (x1 y1 x2 y2 x3 y3). */
int hpm_host_delaunay(int w, int h, const int32_t* xy, int n, int32_t* out_tris, int capacity, int* count);
/* The barycentric raster of main.cpp:389-404: mask (w*h floats, zero-filled
 * here) receives triangle index + 1. */
int hpm_host_raster(const int32_t* tris, int n, int w, int h, float* mask);
/* GetPriorPlaneParams' plane through three camera-frame points
 * (HPM.cpp:1134-1167): plane = (n, w), |n| = 1, w >= 0. */
int hpm_host_plane_through_points(const float* X1, const float* X2, const float* X3, float* plane);
/* Edge map standing in for cv::Canny (main.cpp:287): 255 = edge.  `image` is
 * w*h*channels bytes, channels 1 (grey) or 3 (R G B, as a PPM holds it); a
 * colour pixel takes the gradient of its channel of largest |dx| + |dy|. */
int hpm_host_canny(const uint8_t* image, int w, int h, int channels, float low, float high, uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif /* HPM_PIPELINE_H_ */
