/* hpm_mi355x.h -- C ABI of the MI355X PatchMatch depth/normal estimator.
 *
 * Drop-in boundary for HPM-MVS++'s `class HPM` (HPM.h:57-140) and the JBU
 * helpers (HPM.h:158-202, RunJBU HPM.h:20).  The reference is a C++ class used
 * in-process by ProcessProblem (main.cpp:237-634); here the same operations are
 * plain C functions on an opaque context so that the C++ facade
 * (hpm-mvs_plusplus_amd/host/hpm_host.hpp, class hpmhost::HPM), Python
 * (ctypes) or any FFI can call them.
 *
 * Conventions
 *   - Every function returns HPM_OK (0) or a negative HPM_E* status; it never
 *     exits the process (the reference's CudaSafeCall exits, HPM.cpp:68-76).
 *     hpm_last_error() gives the message of the last failure on a context.
 *   - Images and per-pixel maps are row-major, width = cameras[i].width.
 *   - "planes" are float4 per pixel (x, y, z, w).  What x,y,z,w hold is stated
 *     per function, following the reference.
 *   - One context owns one HIP stream on one device; calls on a context are not
 *     thread-safe (the reference is single-threaded too, main.cpp:250).
 *   - A call that fails leaves no half-built state on the context: a failed
 *     hpm_set_views / hpm_set_views_pooled that got past its argument checks
 *     leaves the context without views, a failed hpm_set_src_depths without
 *     source depths, a failed hpm_prior_support_points without support
 *     points (later calls that need them return HPM_E_STATE), and a failed
 *     run or refinement rebuilds what it needs when called again.
 *   - Fault injection for tests: with HPM_FAULT_ALLOC_AFTER=n in the
 *     environment when hpm_create runs, that context's device allocation
 *     number n (counted from 0 over the buffers it allocates after creation)
 *     fails once with HPM_E_NOMEM, as if the device were out of memory.
 */
#ifndef HPM_MI355X_H_
#define HPM_MI355X_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPM_OK 0
#define HPM_E_INVALID (-1)      /* bad argument / unsupported configuration */
#define HPM_E_STATE (-2)        /* call out of order (e.g. run before set_views) */
#define HPM_E_DEVICE (-3)       /* HIP runtime error */
#define HPM_E_NOMEM (-4)        /* device allocation failed */
#define HPM_E_TOO_LARGE (-5)    /* input larger than an entry point's fixed scratch */

#define HPM_MAX_IMAGES 33       /* reference + 32 sources: view masks are uint32 (HPM.cu:67-75) */

/* Precision of the image storage used by the cost.  All accumulation is FP32.
 * FP32_EXACT reproduces the FP32 oracle bit for bit.  The mixed-precision
 * storages (FP16, BF16, FP32_FAST) take views whose (w + 2) x (h + 1)
 * column-pair grid holds fewer than 2^24 elements (4095 x 4094 holds
 * 2^24 - 1); a larger view makes the run or cost call return
 * HPM_E_INVALID before any launch, and FP32_EXACT takes it. */
#define HPM_PREC_FP32_EXACT 0
#define HPM_PREC_FP16 1
#define HPM_PREC_BF16 2
#define HPM_PREC_FP32_FAST 3


/* Mirrors struct Camera (main.h:39-47): x_cam = R * X + t. 100 bytes. */
typedef struct hpm_camera {
  float K[9];
  float R[9];
  float t[3];
  int32_t height;
  int32_t width;
  float depth_min;
  float depth_max;
} hpm_camera;

/* The live fields of PatchMatchParams (HPM.h:32-55) plus the pass mode that
 * SetGeomConsistencyParams / SetPlanarPriorParams / SetHierarchyParams /
 * SetMandConsistencyParams select (HPM.cpp:565-587). */
typedef struct hpm_params {
  int32_t patch_size;        /* 11 (only 11 with radius_increment 2 is supported) */
  int32_t radius_increment;  /* 2 */
  float sigma_spatial;       /* 5 */
  float sigma_color;         /* 3 */
  int32_t top_k;             /* 4 */
  int32_t geom_consistency;
  int32_t prior_consistency;
  int32_t hierarchy;
  int32_t mand_consistency;
  int32_t num_iterations;    /* 3 (HPM.cu:3013 hard-codes 3) */
  int32_t precision;         /* HPM_PREC_* */
} hpm_params;

typedef struct hpm_ctx hpm_ctx;

/* Fills the reference defaults (HPM.h:32-55, 3 iterations, FP32 exact). */
void hpm_default_params(hpm_params* p);

/* HPM::HPM (HPM.cpp:103-113): creates a context on HIP device `device`. */
int hpm_create(hpm_ctx** out, int device);
/* HPM::~HPM + CudaSpaceRelease (HPM.cpp:115-157, 599-633). NULL is a no-op. */
void hpm_destroy(hpm_ctx* ctx);
const char* hpm_last_error(const hpm_ctx* ctx);

/* Set* mode calls (HPM.cpp:565-587) in one struct. */
int hpm_set_params(hpm_ctx* ctx, const hpm_params* params);

/* Storage of HPM_PREC_FP16 for large view sets (no counterpart in the
 * reference, which samples FP32 textures): when the context's views' FP16
 * column-pair images would take at least min_bytes (default 192 MiB, about
 * the Infinity Cache's 256 MB with the run's other buffers; -1 = never, 0 =
 * always) and every texel is an integer in [0, 255] (checked on the device
 * once per set of views), the images are stored as 8-bit column pairs, half
 * the bytes.  Costs and planes are bit-identical to FP16 storage.  Runs use
 * them only when they start from random planes (no geometric, prior,
 * hierarchy or mandatory-consistency mode): their first launches scatter the
 * gathers and gain (3072 x 2048 with 10 sources: 12 % less time per launch),
 * while runs from converged planes and views whose images fit the caches are
 * bound by VALU work, which the 8-bit pairs raise.  hpm_cost_fixed uses them
 * by footprint alone.  A fused batch uses 8-bit pairs only if every view
 * qualifies. */
int hpm_set_u8_pairs_min_bytes(hpm_ctx* ctx, long long min_bytes);

/* InuputInitialization (HPM.cpp:644-733) + the image/camera half of
 * CudaSpaceInitialization (HPM.cpp:741-772): num_images = 1 reference + sources,
 * grey float images (HPM.cpp:660-661), cameras already scaled to the image size
 * (HPM.cpp:668-676).  The depth range is cameras[0].depth_min*0.6 ..
 * depth_max*1.2 (HPM.cpp:685-686).  Copies everything to the device and
 * allocates the per-pixel state. */
int hpm_set_views(hpm_ctx* ctx, int num_images, const hpm_camera* cameras, const float* const* images);

/* Image pool: the images and cameras of a whole scene, uploaded once to
 * `device` and shared by the contexts of its reference views (hpm_set_views
 * would copy the reference and source images into every context).  The
 * column-pair images of the mixed-precision engine are built once per
 * precision on first use.  No
 * counterpart in the reference, which uploads the images of every problem
 * into every HPM object (HPM.cpp:741-772).  The pool must outlive the
 * contexts that use it; contexts may run concurrently on it. */
typedef struct hpm_pool hpm_pool;
int hpm_pool_create(hpm_pool** out, int device, int num_images, const hpm_camera* cameras,
                    const float* const* images);
void hpm_pool_destroy(hpm_pool* pool);
/* hpm_set_views with pool images ids[0] (reference), ids[1..num_images-1]
 * (sources), without copying them. */
int hpm_set_views_pooled(hpm_ctx* ctx, hpm_pool* pool, int num_images, const int32_t* ids);

/* Geometric-consistency depth textures (HPM.cpp:693-732, 786-811): one depth
 * map per image, index 0 (the reference's own) is not read by the cost. */
int hpm_set_src_depths(hpm_ctx* ctx, const float* const* depths);

/* CudaCannyInitialization (HPM.cpp:908-920): edge map of the reference image,
 * nonzero = edge (cv::Canny output, main.cpp:287). */
int hpm_set_canny(hpm_ctx* ctx, const uint32_t* canny);
/* The same from an 8-bit edge map (cv::Canny's own output type). */
int hpm_set_canny_u8(hpm_ctx* ctx, const uint8_t* canny);

/* CudaConfidenceInitialization (HPM.cpp:922-940). */
int hpm_set_confidence(hpm_ctx* ctx, const float* confidences);

/* CudaHypothesesReload (HPM.cpp:942-962) and the geometric-pass reload
 * (HPM.cpp:813-846): planes = (world normal xyz, depth w) per pixel, costs. */
int hpm_reload_hypotheses(hpm_ctx* ctx, const float* planes, const float* costs);
/* The same from the maps as ProcessProblem keeps them (depths.dmb /
 * normals.dmb / costs.dmb): a depth map, a 3-channel world normal map and
 * the costs; the planes are assembled on the device. */
int hpm_reload_hypotheses_maps(hpm_ctx* ctx, const float* depth, const float* normals, const float* costs);

/* CudaPlanarPriorInitialization / ReloadPlanarPriorInitialization
 * (HPM.cpp:964-984, 1433-1453): prior plane per pixel and triangle mask
 * (0 = no prior, else triangle index + 1). */
int hpm_set_prior(hpm_ctx* ctx, const float* prior_planes, const uint32_t* masks);
/* CudaPlanarPriorRelease (HPM.cpp:589-597): frees the prior planes, the
 * triangle masks and the edge map (freed in the order of the context's
 * stream). */
int hpm_release_prior(hpm_ctx* ctx);
/* CudaSpaceRelease (HPM.cpp:599-633): frees every per-view device buffer
 * (images, hypotheses, costs, masks, source depths, prior); the context keeps
 * its parameters and takes new views with hpm_set_views. */
int hpm_release(hpm_ctx* ctx);

/* The planar prior built on the device from the maps the context holds:
 * reloaded hypotheses (hpm_reload_hypotheses: the depth in w) and costs, and
 * the confidences (hpm_set_confidence), as ProcessProblem's prior branch
 * builds it from depths/costs/confidence.dmb (main.cpp:324-597).
 *
 * hpm_prior_support_points: GetSupportPoints_Classify_Check (HPM.cpp:1026-1073)
 * at `distance` 0 (the reference grid) or 1, 2 (the maps resized by
 * 1 / (2 distance), main.cpp:443-454).  Writes up to `capacity` (x, y) pairs
 * in the reference's order and sets *count to their number.
 *
 * hpm_prior_build: the rest, from the Delaunay triangles of those points
 * (DelaunayTriangulation, HPM.cpp:1075-1097; 6 int32 per triangle, corners on
 * the same grid): per-triangle planes (GetPriorPlaneParams[_factor],
 * HPM.cpp:1099-1167), the barycentric raster (main.cpp:389-409, 493-514),
 * the depth range test and, at distance 0, the prior planes and triangle
 * masks of CudaPlanarPriorInitialization (HPM.cpp:964-984); at distance 1, 2
 * the coarse depth and world normal upsampled by JBU_cu_prior and reloaded
 * per pixel (main.cpp:516-576, ReloadPlanarPriorInitialization,
 * HPM.cpp:1433-1453).  Replaces hpm_set_prior; bit-identical to building the
 * prior on the host with the same steps.
 *
 * hpm_get_prior: the context's prior planes and masks (as hpm_set_prior takes
 * them). */
int hpm_prior_support_points(hpm_ctx* ctx, int distance, int32_t* out_xy, int capacity, int* count);
int hpm_prior_build(hpm_ctx* ctx, const int32_t* triangles, int num_triangles);
int hpm_get_prior(hpm_ctx* ctx, float* prior_planes, uint32_t* masks);

/* The edge map ProcessProblem uploads with CudaCannyInitialization:
 * cv::Canny(image, low, high) with a 3 x 3 aperture and the L1 gradient
 * (main.cpp:280-289) of a w x h image of 1 (grey) or 3 (R G B) channels;
 * out_edges = 255 on edges, else 0.  Runs on `device` on a stream of its own
 * and returns when the map is on the host. */
int hpm_canny(int device, const uint8_t* image, int w, int h, int channels, float low, float high, uint8_t* out_edges);
/* hpm_canny for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  It allocates nothing: image_dev
 * (w*h*channels bytes) and out_edges_dev (w*h bytes) are device buffers on the
 * current device, the gradient, magnitude and class planes live in a static
 * device array of 7 bytes per pixel for up to HPM_CANNY_NOALLOC_MAX_PIXELS
 * pixels (44 MiB, held by every GPU the library runs on until the process
 * ends).  w*h above the limit returns HPM_E_TOO_LARGE before any launch.  Runs
 * on the calling thread's default stream (hipStreamPerThread) and returns when
 * the edges are written; the same edges as hpm_canny.  Calls share their
 * device's array, which keeps the last call's contents: calls that overlap on
 * one device race. */
#define HPM_CANNY_NOALLOC_MAX_PIXELS (3072 * 2048)
int hpm_canny_noalloc(const uint8_t* image_dev, int w, int h, int channels, float low, float high,
                      uint8_t* out_edges_dev);

/* Hierarchy reload (HPM.cpp:849-905): the coarse level's maps (cw x ch) and the
 * JBU-upsampled fine depth.  coarse = (world normal xyz, w) where w is the
 * coarse cost when the sizes differ (`upsample`) else the coarse depth.  The
 * reference's size test compares width with rows (HPM.cpp:869); that test is
 * kept.  The fine planes get xyz = 0 and w = fine depth (HPM.cpp:894-901). */
int hpm_set_coarse(hpm_ctx* ctx, const float* coarse_planes, int cw, int ch, const float* fine_depth);

/* RunPatchMatch (HPM.cu:2987-3035): initialisation, num_iterations red/black
 * sweeps and the plane -> (world normal, depth) conversion.  `seed` keys the
 * per-pixel random streams.  Asynchronous on the context stream. */
int hpm_run_patchmatch(hpm_ctx* ctx, uint64_t seed);

/* Partial schedule for tests: run initialisation plus `n_halfpasses` colour
 * passes (-1 = all) and the output conversion if do_output != 0. */
int hpm_run_patchmatch_partial(hpm_ctx* ctx, uint64_t seed, int n_halfpasses, int do_output);

/* Enables (1) or disables (0) hipGraph replay: the kernel sequence of
 * RunPatchMatch is captured into a graph on the first run and replayed on later
 * runs until a buffer, size, mode or precision changes (then recaptured).
 * No counterpart in the reference, which launches and synchronises each kernel
 * (HPM.cu:3015-3027).  Timing (hpm_set_timing) runs without the graph. */
int hpm_set_graph(hpm_ctx* ctx, int enable);

/* Batched RunPatchMatch (HPM.cu:2987-3035 for n reference views): every
 * context's schedule is issued on its own stream, so that views of the same or
 * different sizes run concurrently on the GPU; returns when all are done.
 * seeds[i] keys context i's random streams. */
int hpm_run_patchmatch_batch(hpm_ctx* const* ctxs, int n, const uint64_t* seeds);

/* Fused multi-view RunPatchMatch.  The reference runs one view per
 * ProcessProblem (main.cpp:237-634, RunPatchMatch HPM.cu:2987-3035); a batch
 * runs the schedules of n contexts -- views of any size and source count, e.g.
 * all reference views of a pass or several pyramid levels -- as one launch per
 * stage (init, each colour pass, output) over the concatenated tile grids of
 * all views, optionally replayed as one hipGraph.  Every context keeps its own
 * parameters, modes and buffers; all must share one device, one precision and
 * one iteration count.  A context's results equal hpm_run_patchmatch's with the
 * same seed.  The batch holds the context pointers: keep the contexts alive
 * while it exists (their views and modes may change between runs). */
typedef struct hpm_batch hpm_batch;
int hpm_batch_create(hpm_batch** out, hpm_ctx* const* ctxs, int n);
void hpm_batch_destroy(hpm_batch* batch);
const char* hpm_batch_last_error(const hpm_batch* batch);
/* Runs every context's RunPatchMatch; seeds[i] keys context i's random streams.
 * Returns when all views are done. */
int hpm_batch_run(hpm_batch* batch, const uint64_t* seeds);
int hpm_batch_set_graph(hpm_batch* batch, int enable);
/* Launch shape: colour tile width in same-colour columns (16, 32 or 64; 0 =
 * 32; the exact FP32 path takes 32 only) and the height, in tiles, of the
 * strips each XCD walks (1..8).  Results do not depend on it. */
int hpm_batch_set_shape(hpm_batch* batch, int tile_w, int strip);
int hpm_batch_get_shape(const hpm_batch* batch, int* tile_w, int* strip);
/* Pixels (threads) per colour-pass workgroup: 256 or 128 (0 = 256; the exact
 * FP32 path takes 256 only).  The colour tile is tile_w same-colour columns
 * by threads / tile_w rows.  Results do not depend on it.  No counterpart in
 * the reference, whose checkerboard kernels use BLOCK_W x BLOCK_W / 2 blocks
 * (HPM.cu:2992-3011). */
int hpm_batch_set_workgroup(hpm_batch* batch, int threads);
int hpm_batch_get_workgroup(const hpm_batch* batch, int* threads);
/* Device time of the last run's schedule, milliseconds (HIP events). */
float hpm_batch_last_ms(const hpm_batch* batch);
/* Runs the batch once per launch shape (threads per workgroup 256, 128 x tile
 * widths x strips 1, 2, 4), keeps the fastest and runs it once more, so the
 * views hold its result; *best_ms = its schedule time.  Photometric passes
 * only (their init overwrites all state, so reruns are identical); otherwise
 * HPM_E_STATE. */
int hpm_batch_autotune(hpm_batch* batch, const uint64_t* seeds, float* best_ms);

/* Waits for the context stream. */
int hpm_synchronize(hpm_ctx* ctx);

/* Bulk getters replacing GetPlaneHypothesis / GetCost / GetTexture
 * (HPM.cpp:1001-1014) and the D2H copies of HPM.cu:3029-3033.  Each copies
 * width*height elements (planes: 4 floats per pixel). */
int hpm_get_planes(hpm_ctx* ctx, float* out_planes);
/* The planes as ProcessProblem stores them (main.cpp:311-322): the depth map
 * (width*height floats) and the world normal map (3 floats per pixel). */
int hpm_get_depth_normals(hpm_ctx* ctx, float* out_depth, float* out_normals);
int hpm_get_costs(hpm_ctx* ctx, float* out_costs);
int hpm_get_texture(hpm_ctx* ctx, float* out_texture);
int hpm_get_selected_views(hpm_ctx* ctx, uint32_t* out_masks);
int hpm_get_confidence(hpm_ctx* ctx, float* out_confidences);
int hpm_get_size(const hpm_ctx* ctx, int* width, int* height, int* num_images);

/* Work counters of the last run: NCC evaluations per kernel class. */
typedef struct hpm_stats {
  uint64_t ncc_init;        /* NCC evaluations in the initialisation kernel */
  uint64_t ncc_propagate;   /* NCC evaluations in the red/black kernels */
  uint64_t ncc_reference;   /* evaluations the reference would perform: V at init,
                               (14 + E) * V per updated pixel and iteration */
  double ms_init;           /* device time of the init kernel (HIP events) */
  double ms_propagate;      /* summed device time of the red/black kernels */
  double ms_output;
  int32_t propagate_launches;
} hpm_stats;
int hpm_get_stats(hpm_ctx* ctx, hpm_stats* out);
/* Enables HIP-event timing of each kernel class (adds no work to the kernels). */
int hpm_set_timing(hpm_ctx* ctx, int enable);

/* Cost microbenchmark at fixed hypotheses (ComputeMultiViewCostVector,
 * HPM.cu:490-495 over every pixel): planes = (ref-camera normal, d) with
 * n.X + d = 0; out = width*height*(num_images-1) costs, pixel-major.
 * Host pointers. */
int hpm_cost_fixed(hpm_ctx* ctx, const float* planes, float* out_costs, int precision);
/* Same on device pointers already in HBM (for timing), asynchronous. */
int hpm_cost_fixed_device(hpm_ctx* ctx, const void* planes_dev, void* out_costs_dev, int precision);

/* ComputeMultiViewInitialCostandSelectedViews (HPM.cu:450-488) at fixed planes. */
int hpm_init_cost(hpm_ctx* ctx, const float* planes, float* out_cost, uint32_t* out_masks);

/* ComputeGeomConsistencyCost (HPM.cu:539-564) for every (pixel, source view);
 * needs hpm_set_src_depths.  out = width*height*(num_images-1). */
int hpm_geom_cost(hpm_ctx* ctx, const float* planes, float* out_costs);

/* JBU::CudaRun / RunJBU (HPM.cu:3037-3128, HPM.cpp:1266-1319): joint bilateral
 * upsampling of a coarse depth map (cw x ch) guided by the fine grey image
 * (w x h).  No NaN patching (RunJBU does that on the host, HPM.cpp:1297-1300). */
int hpm_jbu_depth(const float* fine_gray, int w, int h, const float* coarse_depth, int cw, int ch, float* out_depth);

/* JBU_prior::CudaRun_prior (HPM.cu:3130-3235): upsamples depth and normal
 * (coarse_normals = 4 floats per coarse pixel, xyz used). out_normals gets 4
 * floats per fine pixel. */
int hpm_jbu_prior(const float* fine_gray, int w, int h, const float* coarse_depth, const float* coarse_normals, int cw,
                  int ch, float* out_depth, float* out_normals);

/* Backward pass of the multi-view bilateral NCC (no counterpart in the
 * reference, which has no gradient).  Per pixel, params = (d, theta, phi, -):
 * plane depth d and normal n = (sin theta cos phi, sin theta sin phi,
 * -cos theta) in the reference camera frame; weights = num_images-1 frozen view
 * weights per pixel (views with weight <= 0 are skipped).  Outputs
 * out_cost = sum_v a_v C_v / sum_v a_v (C_v the NCC cost of HPM.cu:375-448 on
 * FP32 texels), out_grad = dC/d(d, theta, phi) (3 per pixel; may be NULL), and
 * out_gn (6 per pixel, upper triangle of the Gauss-Newton matrix
 * sum_v a_v J_v^T J_v / sum_v a_v of the residuals sqrt(w_k / W) (r^_k - s^_k)
 * between the normalised reference and source windows, whose squared norm / 2
 * is C_v; may be NULL).  The gradient is 0 on the clamp and small-variance
 * branches. */
int hpm_cost_grad(hpm_ctx* ctx, const float* params, const float* weights, float* out_cost, float* out_grad,
                  float* out_gn);
/* Same on device pointers already in HBM (e.g. torch tensors), asynchronous on
 * the context stream. */
int hpm_cost_grad_device(hpm_ctx* ctx, const void* params_dev, const void* weights_dev, void* out_cost_dev,
                         void* out_grad_dev, void* out_gn_dev);
/* Gauss-Newton (Levenberg-Marquardt) refinement of every pixel's plane:
 * `iterations` steps on the normalised-window residuals, each kept only where
 * it lowers the cost.  params (host, 4 floats per pixel) are updated in place;
 * out_cost (may be NULL) receives the final cost.  Beside the reference's
 * random-perturbation refinement (PlaneHypothesisRefinement, HPM.cu:765-841). */
int hpm_refine_gn(hpm_ctx* ctx, float* params, const float* weights, int iterations, float* out_cost);
/* The random-perturbation counterpart on the same parameters and cost: per
 * step a candidate d (1 + 0.02 s u0), theta + 5 deg s u1, phi + 5 deg s u2
 * (u uniform in [-1, 1], s = 0.8^step), kept where it lowers the cost. */
int hpm_refine_random(hpm_ctx* ctx, float* params, const float* weights, int iterations, uint64_t seed,
                      float* out_cost);
/* Both in turn: per step the random candidate above, then a Gauss-Newton step
 * from the result, each kept where it lowers the cost (the random candidate
 * reaches the basin of the optimum when the start lies outside the few
 * pixels of disparity where the GN model holds; GN then converges in it). */
int hpm_refine_hybrid(hpm_ctx* ctx, float* params, const float* weights, int iterations, uint64_t seed,
                      float* out_cost);
/* Coarse-to-fine Gauss-Newton: the first smoothed_iterations steps minimise
 * the cost of the views smoothed by a Gaussian of `sigma` px (0 < sigma <= 5,
 * radius ceil(3 sigma)), whose basin of convergence is about as wide as the
 * blur, the remaining steps the views' own cost (the sharp cost's basin is
 * about a texel of the finest texture: at full resolution a 2 % depth error
 * lies outside it).  out_cost receives the sharp views' cost. */
int hpm_refine_gn_smoothed(hpm_ctx* ctx, float* params, const float* weights, int iterations,
                           int smoothed_iterations, float sigma, float* out_cost);

/* Multi-view consistency check of RunFusion (mode HPM_CHECK_FUSION,
 * main.cpp:823-957) and ConfidenceEvaluation (mode HPM_CHECK_CONFIDENCE,
 * main.cpp:959-1109), with the reference's greedy masks: images in order,
 * pixels row-major, sources in pair order.
 *   cameras[i]  camera of image i, already rescaled to its depth map
 *               (RescaleImageAndCamera, HPM.cpp:186-207);
 *   depths[i]   width*height depths; normals[i] 3 floats per pixel (world
 *               normal, as normals.dmb holds);
 *   src_ids[src_offsets[i] .. src_offsets[i+1]) are image i's source images
 *               (Problem::src_image_ids), src_offsets has num_images+1 entries.
 * Limits: at most 1024 images, 32 sources per image (the reference's
 * per-pixel arrays hold 32) and 2^27 pixels per image.
 * HPM_CHECK_FUSION: out_pairs receives (image, pixel) of every fused point in
 *   the reference's PointCloud order, out_points (may be NULL) 6 floats per
 *   point (world position, normal), both up to `capacity` points;
 *   *out_count is the total number of fused points.
 * HPM_CHECK_CONFIDENCE: out_confidence[i] (width*height floats each) receives
 *   image i's confidence.dmb: the consistency map as it stands when image i is
 *   done, which is when the reference writes the file (main.cpp:1095-1100);
 *   writes by later images that list i as a source (main.cpp:1086) change
 *   their masks but not this map.  Zeros where not consistent; the depth-range
 *   filter uses cameras[i].depth_min/max for reference and source depths, as the
 *   reference does (main.cpp:1032,1057).
 * Runs on the current HIP device; blocking. */
#define HPM_CHECK_FUSION 0
#define HPM_CHECK_CONFIDENCE 1
/* RunFusion_Sky_Strict (main.cpp:659-821, --mask): as HPM_CHECK_FUSION with
 * that function's float acceptance threshold (`float factor = 0.3`); the
 * mask colour test on the fused points is the caller's. */
#define HPM_CHECK_FUSION_MASK 2
int hpm_consistency_check(int mode, int num_images, const hpm_camera* cameras, const float* const* depths,
                          const float* const* normals, const int32_t* src_offsets, const int32_t* src_ids,
                          float* const* out_confidence, int32_t* out_pairs, float* out_points, int64_t capacity,
                          int64_t* out_count);
/* The same, calling image_done(i, user) on the calling thread as soon as the
 * ordered pass has finished image i (images in order).  out_confidence[i] is
 * then complete and never written again: a caller can hand it out while the
 * check goes on (no counterpart in the reference; for overlapping the next
 * pass with the check). */
int hpm_consistency_check_cb(int mode, int num_images, const hpm_camera* cameras, const float* const* depths,
                             const float* const* normals, const int32_t* src_offsets, const int32_t* src_ids,
                             float* const* out_confidence, int32_t* out_pairs, float* out_points, int64_t capacity,
                             int64_t* out_count, void (*image_done)(int image, void* user), void* user);

#ifdef __cplusplus
}
#endif
#endif /* HPM_MI355X_H_ */
