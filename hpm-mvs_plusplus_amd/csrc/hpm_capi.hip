// hpm_capi.hip -- the C ABI of include/hpm_mi355x.h: context, device buffers,
// per-run constant block and kernel launches on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cstdint>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hpm_mi355x.h"
#include "hpm_fast.h"
#include "hpm_fusion.h"
#include "hpm_internal.h"
#include "hpm_grad.h"
#include "hpm_jbu.h"
#include "hpm_prior.h"
#include "hpm_canny.h"
#include "hpm_kernels.h"

using namespace hpm;

namespace {

int fail(hpm_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_OK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return fail(ctx, HPM_E_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

// True for the allocation that HPM_FAULT_ALLOC_AFTER names (counted per
// context over the device allocations it makes after hpm_create); that one
// fails instead of reaching the device.
bool injected_fault(hpm_ctx* ctx) {
  if (ctx->fault_alloc < 0) return false;
  return ctx->fault_alloc-- == 0;
}

hipError_t ctx_malloc_async(hpm_ctx* ctx, void** p, size_t bytes) {
  if (injected_fault(ctx)) return hipErrorOutOfMemory;
  return hipMallocAsync(p, bytes, ctx->stream);
}

// Context buffers are allocated and freed in the order of the context's
// stream: unlike hipFree, hipFreeAsync does not wait for the whole device,
// so contexts driven from other host threads keep their work running.
template <class T> int dalloc(hpm_ctx* ctx, T** p, size_t n) {
  if (*p) { hipFreeAsync(*p, ctx->stream); *p = nullptr; }
  if (n == 0) return HPM_OK;
  hipError_t e = ctx_malloc_async(ctx, (void**)p, n * sizeof(T));
  if (e != hipSuccess) { *p = nullptr; return fail(ctx, HPM_E_NOMEM, std::string("hipMallocAsync: ") + hipGetErrorString(e)); }
  return HPM_OK;
}

// The device's default memory pool keeps what contexts free for the next
// ones instead of returning it to the driver at every synchronisation.
void keep_pool_memory(int device) {
  static std::mutex mu;
  static std::vector<int> done;
  std::lock_guard<std::mutex> lock(mu);
  if (std::find(done.begin(), done.end(), device) != done.end()) return;
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
    uint64_t threshold = UINT64_MAX;
    hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
  }
  done.push_back(device);
}

// Frees a buffer from dalloc in the order of the context's stream.
template <class T> void dfree(hpm_ctx* ctx, T* p) {
  if (p) hipFreeAsync(p, ctx->stream);
}

void free_views(hpm_ctx* c) {
  if (!c->pool) {
    for (float* p : c->d_img) if (p) hipFreeAsync(p, c->stream);
    for (void* p : c->d_img_lp) if (p) hipFreeAsync(p, c->stream);
  }
  c->pool = nullptr;
  c->pool_ids.clear();
  for (float* p : c->d_dimg) if (p) hipFreeAsync(p, c->stream);
  c->d_img.clear(); c->d_dimg.clear(); c->d_img_lp.clear(); c->lp_kind = -1; c->u8_ok = -1;
  void* bufs[] = {c->d_planes, c->d_costs, c->d_pre_costs, c->d_conf, c->d_texture, c->d_sel,
                  c->d_masks,  c->d_canny, c->d_prior,     c->d_scaled};
  for (void* p : bufs) if (p) hipFreeAsync(p, c->stream);
  c->d_planes = nullptr; c->d_costs = nullptr; c->d_pre_costs = nullptr; c->d_conf = nullptr;
  c->d_texture = nullptr; c->d_sel = nullptr; c->d_masks = nullptr; c->d_canny = nullptr;
  c->d_prior = nullptr; c->d_scaled = nullptr;
  if (c->d_prior_ds) hipFreeAsync(c->d_prior_ds, c->stream);
  if (c->d_scratch) hipFreeAsync(c->d_scratch, c->stream);
  c->d_prior_ds = nullptr; c->d_scratch = nullptr; c->scratch_bytes = 0;
  c->prior_dist = -1; c->prior_w = c->prior_h = 0;
  c->nimg = 0; c->W = c->H = 0;
}

// The reference's FP32 view-pair basis (HPM.cu:253-279).
void pair_basis(const hpm_camera& rc, const hpm_camera& sc, PairBasis& B) {
  float ref_C[3], src_C[3];
  for (int r = 0; r < 3; ++r) {
    ref_C[r] = -(rc.R[0 + r] * rc.t[0] + rc.R[3 + r] * rc.t[1] + rc.R[6 + r] * rc.t[2]);
    src_C[r] = -(sc.R[0 + r] * sc.t[0] + sc.R[3 + r] * sc.t[1] + sc.R[6 + r] * sc.t[2]);
  }
  float C_rel[3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      B.R_rel[3 * a + b] = sc.R[3 * a + 0] * rc.R[3 * b + 0] + sc.R[3 * a + 1] * rc.R[3 * b + 1] + sc.R[3 * a + 2] * rc.R[3 * b + 2];
  for (int r = 0; r < 3; ++r) C_rel[r] = ref_C[r] - src_C[r];
  for (int a = 0; a < 3; ++a) B.t_rel[a] = sc.R[3 * a + 0] * C_rel[0] + sc.R[3 * a + 1] * C_rel[1] + sc.R[3 * a + 2] * C_rel[2];
}

int build_consts(hpm_ctx* ctx, RunConsts& C) {
  std::memset(&C, 0, sizeof(C));
  const hpm_params& P = ctx->params;
  C.W = ctx->W; C.H = ctx->H; C.V = ctx->nimg - 1;
  C.geom = P.geom_consistency; C.prior = P.prior_consistency; C.hierarchy = P.hierarchy;
  C.mand = P.mand_consistency; C.upsample = (P.hierarchy && ctx->upsample) ? 1 : 0;
  C.num_iterations = P.num_iterations;
  C.depth_min = ctx->cams[0].depth_min * 0.6f;  // HPM.cpp:685
  C.depth_max = ctx->cams[0].depth_max * 1.2f;  // HPM.cpp:686
  C.sigma_spatial = P.sigma_spatial; C.sigma_color = P.sigma_color; C.top_k = P.top_k;
  C.scaled_cols = (float)ctx->scaled_w; C.scaled_rows = (float)ctx->scaled_h;
  for (int it = 0; it < 3; ++it) {
    for (int e = 0; e < 3; ++e) C.judge_thr[it][e] = (float)(0.8 * exp_hpm(-it * it * (3 - e) / 90.0));
    const float thr = (float)(0.8 * (double)expf_hpm((float)(it * it) / (-90.0f)));
    C.vs_thr[it] = thr;
    C.vs_fb[it] = expf_hpm(thr * thr / (-0.32f));
  }
  C.depth_sigma = (C.depth_max - C.depth_min) / 64.0f;
  C.two_dss = 2 * C.depth_sigma * C.depth_sigma;
  C.angle_sigma = (float)(3.14159265358979323846 * (5.0f / 180.0f));
  C.two_ass = 2 * C.angle_sigma * C.angle_sigma;
  C.pert_refine = (float)(0.02f * 3.14159265358979323846);
  C.pert_prior_init = (float)(3 * 0.02f * 3.14159265358979323846);
  for (int i = 0; i < ctx->nimg; ++i) {
    std::memcpy(&C.cams[i], &ctx->cams[i], sizeof(DevCam));
    C.img[i] = ctx->d_img[i];
    C.dimg[i] = i < (int)ctx->d_dimg.size() ? ctx->d_dimg[i] : nullptr;
  }
  for (int v = 1; v < ctx->nimg; ++v) pair_basis(ctx->cams[0], ctx->cams[v], C.pair[v - 1]);
  // FP64 view-pair basis of the mixed-precision engine: H = A - b m^T with
  // A = K_s R_rel K_r^-1, b = K_s t_rel, m = K_r^-T n / d (zero skew and
  // K_r[8] = 1, as HPM.cu:292-310 assumes).
  const hpm_camera& rc = ctx->cams[0];
  const double fx = rc.K[0], fy = rc.K[4], cx = rc.K[2], cy = rc.K[5];
  C.kinv[0] = 1.0 / fx; C.kinv[1] = 1.0 / fy; C.kinv[2] = cx / fx; C.kinv[3] = cy / fy;
  const double Kri[9] = {1.0 / fx, 0.0, -cx / fx, 0.0, 1.0 / fy, -cy / fy, 0.0, 0.0, 1.0};
  for (int v = 1; v < ctx->nimg; ++v) {
    const hpm_camera& sc = ctx->cams[v];
    double rR[9], sR[9], rC[3], sC[3], Rrel[9], trel[3];
    for (int i = 0; i < 9; ++i) { rR[i] = rc.R[i]; sR[i] = sc.R[i]; }
    for (int r = 0; r < 3; ++r) {
      rC[r] = -(rR[0 + r] * rc.t[0] + rR[3 + r] * rc.t[1] + rR[6 + r] * rc.t[2]);
      sC[r] = -(sR[0 + r] * sc.t[0] + sR[3 + r] * sc.t[1] + sR[6 + r] * sc.t[2]);
    }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        Rrel[3 * a + b] = sR[3 * a] * rR[3 * b] + sR[3 * a + 1] * rR[3 * b + 1] + sR[3 * a + 2] * rR[3 * b + 2];
    for (int a = 0; a < 3; ++a)
      trel[a] = sR[3 * a] * (rC[0] - sC[0]) + sR[3 * a + 1] * (rC[1] - sC[1]) + sR[3 * a + 2] * (rC[2] - sC[2]);
    const double Ks[9] = {sc.K[0], 0.0, sc.K[2], 0.0, sc.K[4], sc.K[5], 0.0, 0.0, sc.K[8]};
    double KR[9];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) KR[3 * a + b] = Ks[3 * a] * Rrel[b] + Ks[3 * a + 1] * Rrel[3 + b] + Ks[3 * a + 2] * Rrel[6 + b];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) C.A[v - 1][3 * a + b] = KR[3 * a] * Kri[b] + KR[3 * a + 1] * Kri[3 + b] + KR[3 * a + 2] * Kri[6 + b];
    for (int a = 0; a < 3; ++a) C.bvec[v - 1][a] = Ks[3 * a] * trel[0] + Ks[3 * a + 1] * trel[1] + Ks[3 * a + 2] * trel[2];
  }
  C.storage = ctx->lp_kind;
  for (int i = 0; i < ctx->nimg && i < (int)ctx->d_img_lp.size(); ++i) C.pairs[i] = ctx->d_img_lp[i];
  return HPM_OK;
}

// Column-pair images of every view for the mixed-precision engine (rebuilt
// when the storage precision changes).
template <class S> int build_pair_images(hpm_ctx* ctx) {
  for (int i = 0; i < ctx->nimg; ++i)  // tap addresses are formed exactly in float
    if ((size_t)(ctx->cams[i].width + 2) * (ctx->cams[i].height + 1) >= ((size_t)1 << 24))
      return fail(ctx, HPM_E_INVALID, "the mixed-precision path takes views below 2^24 texels (e.g. 4094 x 4095); "
                                      "use HPM_PREC_FP32_EXACT for larger images");
  for (void* p : ctx->d_img_lp) if (p) hipFreeAsync(p, ctx->stream);
  ctx->d_img_lp.assign(ctx->nimg, nullptr);
  ctx->lp_kind = -1;  // until every view's images are built
  for (int i = 0; i < ctx->nimg; ++i) {
    const int w = ctx->cams[i].width, h = ctx->cams[i].height;
    hipError_t e = ctx_malloc_async(ctx, &ctx->d_img_lp[i], pair_image_bytes<S>(w, h));
    if (e != hipSuccess) { ctx->d_img_lp[i] = nullptr; return fail(ctx, HPM_E_NOMEM, "pair image allocation failed"); }
    const dim3 b(16, 16), g((w + 2 + 15) / 16, (h + 1 + 15) / 16);
    hipLaunchKernelGGL(k_build_pairs<S>, g, b, 0, ctx->stream, ctx->d_img[i], w, h, ctx->d_img_lp[i]);
    if (hipGetLastError() != hipSuccess) return fail(ctx, HPM_E_DEVICE, "k_build_pairs launch failed");
  }
  ctx->lp_kind = S::kind;
  return HPM_OK;
}

template <class S> int build_pool_pair_images(hpm_ctx* ctx) {
  hpm_pool* pool = ctx->pool;
  std::lock_guard<std::mutex> lock(pool->pairs_mu);
  std::vector<void*>& q = pool->pairs[S::kind];
  if (q.size() != pool->d_img.size()) {
    for (size_t i = 0; i < pool->cams.size(); ++i)
      if ((size_t)(pool->cams[i].width + 2) * (pool->cams[i].height + 1) >= ((size_t)1 << 24))
        return fail(ctx, HPM_E_INVALID, "the mixed-precision path takes views below 2^24 texels");
    q.assign(pool->d_img.size(), nullptr);
    for (size_t i = 0; i < pool->d_img.size(); ++i) {
      const int w = pool->cams[i].width, h = pool->cams[i].height;
      if (injected_fault(ctx) || hipMalloc(&q[i], pair_image_bytes<S>(w, h)) != hipSuccess) {
        // no partial set stays behind: the next run builds them all again
        HIP_OK(hipStreamSynchronize(ctx->stream));
        for (void* p : q) if (p) hipFree(p);
        q.clear();
        return fail(ctx, HPM_E_NOMEM, "pair image allocation failed");
      }
      const dim3 b(16, 16), g((w + 2 + 15) / 16, (h + 1 + 15) / 16);
      hipLaunchKernelGGL(k_build_pairs<S>, g, b, 0, ctx->stream, pool->d_img[i], w, h, q[i]);
      if (hipGetLastError() != hipSuccess) return fail(ctx, HPM_E_DEVICE, "k_build_pairs launch failed");
    }
    // other contexts of the pool may run on other streams
    HIP_OK(hipStreamSynchronize(ctx->stream));
  }
  ctx->d_img_lp.assign(ctx->nimg, nullptr);
  for (int i = 0; i < ctx->nimg; ++i) ctx->d_img_lp[i] = q[ctx->pool_ids[i]];
  ctx->lp_kind = S::kind;
  return HPM_OK;
}

// Sets *bad to 1 if a texel of the n-texel image is not an integer in [0, 255].
__global__ __launch_bounds__(256) void k_check_u8(const float* __restrict__ img, int n, int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = img[i];
  if (!(v >= 0.0f && v <= 255.0f && v == rintf(v))) atomicOr(bad, 1);
}

// Whether every image of the context is integral in [0, 255] (8-bit column
// pairs hold it exactly); checked on the device once per set of views, or once
// per pool.
int images_are_u8(hpm_ctx* ctx, bool& ok) {
  // a pool's answer is shared by its contexts, which may ask from several host threads
  std::unique_lock<std::mutex> lock;
  if (ctx->pool) lock = std::unique_lock<std::mutex>(ctx->pool->pairs_mu);
  int& known = ctx->pool ? ctx->pool->u8_ok : ctx->u8_ok;
  if (known < 0) {
    const std::vector<float*>& imgs = ctx->pool ? ctx->pool->d_img : ctx->d_img;
    const std::vector<hpm_camera>& cams = ctx->pool ? ctx->pool->cams : ctx->cams;
    int* d_bad = nullptr;
    HIP_OK(hipMallocAsync((void**)&d_bad, sizeof(int), ctx->stream));
    HIP_OK(hipMemsetAsync(d_bad, 0, sizeof(int), ctx->stream));
    for (size_t i = 0; i < imgs.size(); ++i) {
      const int n = cams[i].width * cams[i].height;
      hipLaunchKernelGGL(k_check_u8, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, imgs[i], n, d_bad);
    }
    int bad = 1;
    HIP_OK(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipFreeAsync(d_bad, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    known = bad ? 0 : 1;
  }
  ok = known == 1;
  return HPM_OK;
}

// Storage of the column-pair images for a mixed precision: the precision
// itself, except FP16 on 8-bit column pairs (bit-identical costs) when the
// views' FP16 pairs would take at least u8_min_bytes and the images are 8-bit.
int pair_storage(hpm_ctx* ctx, int precision, bool allow_u8, int& kind) {
  kind = precision;
  if (precision != HPM_PREC_FP16 || !allow_u8 || ctx->u8_min_bytes < 0) return HPM_OK;
  size_t bytes = 0;
  for (int i = 0; i < ctx->nimg; ++i) bytes += pair_image_bytes<PairF16>(ctx->cams[i].width, ctx->cams[i].height);
  if (bytes < (size_t)ctx->u8_min_bytes) return HPM_OK;
  bool ok = false;
  const int rc = images_are_u8(ctx, ok);
  if (rc) return rc;
  if (ok) kind = kStorageU8;
  return HPM_OK;
}

int ensure_pair_images(hpm_ctx* ctx, int precision, bool allow_u8 = true) {
  int kind;
  int rc = pair_storage(ctx, precision, allow_u8, kind);
  if (rc) return rc;
  if (ctx->lp_kind == kind && (int)ctx->d_img_lp.size() == ctx->nimg) return HPM_OK;
  if (ctx->pool) {
    switch (kind) {
      case HPM_PREC_FP16: return build_pool_pair_images<PairF16>(ctx);
      case HPM_PREC_BF16: return build_pool_pair_images<PairBF16>(ctx);
      case HPM_PREC_FP32_FAST: return build_pool_pair_images<PairF32>(ctx);
      case kStorageU8: return build_pool_pair_images<PairU8>(ctx);
      default: return fail(ctx, HPM_E_INVALID, "unknown precision");
    }
  }
  switch (kind) {
    case HPM_PREC_FP16: return build_pair_images<PairF16>(ctx);
    case HPM_PREC_BF16: return build_pair_images<PairBF16>(ctx);
    case HPM_PREC_FP32_FAST: return build_pair_images<PairF32>(ctx);
    case kStorageU8: return build_pair_images<PairU8>(ctx);
    default: return fail(ctx, HPM_E_INVALID, "unknown precision");
  }
}

DevState make_state(hpm_ctx* c) {
  DevState S;
  S.planes = c->d_planes; S.costs = c->d_costs; S.pre_costs = c->d_pre_costs; S.sel = c->d_sel;
  S.prior = c->d_prior; S.masks = c->d_masks; S.conf = c->d_conf; S.canny = c->d_canny;
  S.texture = c->d_texture; S.scaled = c->d_scaled; S.counters = c->d_counters;
  return S;
}

int check_params(hpm_ctx* ctx, const hpm_params& p) {
  if (p.patch_size != 11 || p.radius_increment != 2)
    return fail(ctx, HPM_E_INVALID, "only patch_size 11 with radius_increment 2 (36 taps) is supported");
  if (p.num_iterations < 0 || p.num_iterations > 3)
    return fail(ctx, HPM_E_INVALID, "num_iterations must be in [0, 3]");
  if (p.top_k < 1) return fail(ctx, HPM_E_INVALID, "top_k must be >= 1");
  if (p.precision < 0 || p.precision > 3) return fail(ctx, HPM_E_INVALID, "unknown precision");
  if (p.prior_consistency && p.mand_consistency)
    return fail(ctx, HPM_E_INVALID, "prior and mandatory consistency are separate passes");
  return HPM_OK;
}

size_t npix(const hpm_ctx* c) { return (size_t)c->W * c->H; }

}  // namespace

// Launch sequence of RunPatchMatch (HPM.cu:2987-3035) on the context stream:
// init, num_iterations x {black, red}, output.  The reference synchronises the
// device after every launch; stream order is enough here.
template <class E> int launch_schedule(hpm_ctx* ctx, int n_halfpasses, int do_output) {
  const DevState S = make_state(ctx);
  using T = Tiles<E>;
  const dim3 b256(256);
  const dim3 gi(((ctx->W + T::sq - 1) / T::sq) * ((ctx->H + T::sq - 1) / T::sq));
  const dim3 b16(16, 16), g16((ctx->W + 15) / 16, (ctx->H + 15) / 16);
  const int H2 = (ctx->H / 2) * 2;
  const dim3 gp((((ctx->W + 1) / 2 + T::cw - 1) / T::cw) * ((H2 + T::ch - 1) / T::ch));
  hipStream_t st = ctx->stream;
  const int V = ctx->nimg - 1;
  const size_t lds_bytes = (size_t)8 * V * kRowStride * sizeof(float);
  if (ctx->timing) HIP_OK(hipEventRecord(ctx->ev[0], st));
  hipLaunchKernelGGL(k_init<E>, gi, b256, 0, st, ctx->d_consts, S);
  HIP_OK(hipGetLastError());
  if (ctx->timing) HIP_OK(hipEventRecord(ctx->ev[1], st));
  int done = 0;
  for (int it = 0; it < ctx->params.num_iterations; ++it) {
    for (int colour = 0; colour < 2; ++colour) {
      if (n_halfpasses >= 0 && done >= n_halfpasses) break;
      if (V <= kLdsMaxSrc)
        hipLaunchKernelGGL((k_propagate<E, true>), gp, b256, lds_bytes, st, ctx->d_consts, S, colour, it, H2);
      else
        hipLaunchKernelGGL((k_propagate<E, false>), gp, b256, 0, st, ctx->d_consts, S, colour, it, H2);
      HIP_OK(hipGetLastError());
      ++done;
    }
  }
  if (ctx->timing) HIP_OK(hipEventRecord(ctx->ev[2], st));
  if (do_output) {
    hipLaunchKernelGGL(k_output, g16, b16, 0, st, ctx->d_consts, ctx->d_planes);
    HIP_OK(hipGetLastError());
  }
  if (ctx->timing) HIP_OK(hipEventRecord(ctx->ev[3], st));
  ctx->stats.propagate_launches = done;
  return HPM_OK;
}

// Per-kernel attributes the schedule needs (set outside any stream capture).
template <class E> int prepare_schedule(hpm_ctx* ctx) {
  const int V = ctx->nimg - 1;
  const size_t lds_bytes = (size_t)8 * V * kRowStride * sizeof(float);
  if (V <= kLdsMaxSrc)  // the rows need up to 160 KiB of dynamic LDS
    HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_propagate<E, true>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
  return HPM_OK;
}

// Runs the schedule directly, or through the context's hipGraph: captured on
// first use (and whenever the kernel arguments would change: buffers, sizes,
// modes, precision, partial schedule), then replayed.  The per-run values
// (seed, thresholds) live in the RunConsts buffer in device memory, written on
// the stream before the launch, so one graph serves every run.
template <class E> int run_schedule(hpm_ctx* ctx, int n_halfpasses, int do_output) {
  int rc = prepare_schedule<E>(ctx);
  if (rc) return rc;
  if (!ctx->use_graph || ctx->timing) return launch_schedule<E>(ctx, n_halfpasses, do_output);
  struct Key {
    DevState S;
    int W, H, nimg, precision, storage, geom, prior, hier, mand, iters, nh, out;
  } k;
  std::memset(&k, 0, sizeof(k));
  k.S = make_state(ctx);
  k.W = ctx->W; k.H = ctx->H; k.nimg = ctx->nimg; k.precision = ctx->params.precision; k.storage = ctx->lp_kind;
  k.geom = ctx->params.geom_consistency; k.prior = ctx->params.prior_consistency;
  k.hier = ctx->params.hierarchy; k.mand = ctx->params.mand_consistency;
  k.iters = ctx->params.num_iterations; k.nh = n_halfpasses; k.out = do_output;
  const unsigned char* kb = reinterpret_cast<const unsigned char*>(&k);
  if (!ctx->gexec || ctx->gkey.size() != sizeof(k) || std::memcmp(ctx->gkey.data(), kb, sizeof(k)) != 0) {
    if (ctx->gexec) { HIP_OK(hipGraphExecDestroy(ctx->gexec)); ctx->gexec = nullptr; }
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = launch_schedule<E>(ctx, n_halfpasses, do_output);
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(ctx, HPM_E_DEVICE, std::string("stream capture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) { ctx->gexec = nullptr; return fail(ctx, HPM_E_DEVICE, "hipGraphInstantiate failed"); }
    ctx->gkey.assign(kb, kb + sizeof(k));
    ctx->g_launches = ctx->stats.propagate_launches;
  }
  ctx->stats.propagate_launches = ctx->g_launches;
  HIP_OK(hipGraphLaunch(ctx->gexec, ctx->stream));
  return HPM_OK;
}

template <class E> int launch_cost_fixed(hpm_ctx* ctx, const void* planes_dev, void* out_dev) {
  using T = Tiles<E>;
  const dim3 gi(((ctx->W + T::sq - 1) / T::sq) * ((ctx->H + T::sq - 1) / T::sq));
  hipLaunchKernelGGL(k_cost_fixed<E>, gi, dim3(256), 0, ctx->stream, ctx->d_consts, (const float4*)planes_dev, (float*)out_dev);
  HIP_OK(hipGetLastError());
  return HPM_OK;
}

extern "C" {

void hpm_default_params(hpm_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->patch_size = 11;
  p->radius_increment = 2;
  p->sigma_spatial = 5.0f;
  p->sigma_color = 3.0f;
  p->top_k = 4;
  p->num_iterations = 3;
  p->precision = HPM_PREC_FP32_EXACT;
}

int hpm_create(hpm_ctx** out, int device) {
  if (!out) return HPM_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return HPM_E_DEVICE;
  hpm_ctx* ctx = new hpm_ctx();
  ctx->device = device;
  hpm_default_params(&ctx->params);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    return HPM_E_DEVICE;
  }
  keep_pool_memory(device);
  if (const char* f = std::getenv("HPM_FAULT_ALLOC_AFTER")) {  // a non-negative integer, else ignored
    char* end = nullptr;
    const long long n = std::strtoll(f, &end, 10);
    if (end != f && *end == '\0' && n >= 0) ctx->fault_alloc = n;
  }
  if (hipMallocAsync((void**)&ctx->d_consts, sizeof(RunConsts), ctx->stream) != hipSuccess ||
      hipMallocAsync((void**)&ctx->d_counters, 8 * sizeof(unsigned long long), ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    if (ctx->d_consts) hipFreeAsync(ctx->d_consts, ctx->stream);
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return HPM_E_NOMEM;
  }
  for (auto& e : ctx->ev) hipEventCreate(&e);
  *out = ctx;
  return HPM_OK;
}

void hpm_destroy(hpm_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  if (ctx->gexec) hipGraphExecDestroy(ctx->gexec);
  free_views(ctx);
  hipFreeAsync(ctx->d_consts, ctx->stream);
  hipFreeAsync(ctx->d_counters, ctx->stream);
  hipStreamSynchronize(ctx->stream);  // the frees above; no other stream is waited for
  for (auto& e : ctx->ev) hipEventDestroy(e);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* hpm_last_error(const hpm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int hpm_set_params(hpm_ctx* ctx, const hpm_params* params) {
  if (!ctx || !params) return HPM_E_INVALID;
  int rc = check_params(ctx, *params);
  if (rc) return rc;
  ctx->params = *params;
  return HPM_OK;
}

static int alloc_pixel_state(hpm_ctx* ctx);

static int set_views(hpm_ctx* ctx, int num_images, const hpm_camera* cameras, const float* const* images,
                     bool& dropped);
static int set_views_pooled(hpm_ctx* ctx, hpm_pool* pool, int num_images, const int32_t* ids, bool& dropped);

// A call refused for its arguments keeps the views the context had.  One
// that fails after the old views were dropped leaves the context without
// views (runs then return HPM_E_STATE), never with a part of them.
int hpm_set_views(hpm_ctx* ctx, int num_images, const hpm_camera* cameras, const float* const* images) {
  if (!ctx) return HPM_E_INVALID;
  bool dropped = false;
  const int rc = set_views(ctx, num_images, cameras, images, dropped);
  if (rc != HPM_OK && dropped) free_views(ctx);
  return rc;
}

int hpm_set_views_pooled(hpm_ctx* ctx, hpm_pool* pool, int num_images, const int32_t* ids) {
  if (!ctx) return HPM_E_INVALID;
  bool dropped = false;
  const int rc = set_views_pooled(ctx, pool, num_images, ids, dropped);
  if (rc != HPM_OK && dropped) free_views(ctx);
  return rc;
}

static int set_views(hpm_ctx* ctx, int num_images, const hpm_camera* cameras, const float* const* images,
                     bool& dropped) {
  if (!cameras || !images) return fail(ctx, HPM_E_INVALID, "null cameras or images");
  if (num_images < 2 || num_images > HPM_MAX_IMAGES)
    return fail(ctx, HPM_E_INVALID, "num_images must be in [2, 33] (view masks are 32-bit)");
  for (int i = 0; i < num_images; ++i) {
    if (!images[i]) return fail(ctx, HPM_E_INVALID, "null image pointer");
    if (cameras[i].width <= 0 || cameras[i].height <= 0 || (size_t)cameras[i].width * cameras[i].height > (1u << 28))
      return fail(ctx, HPM_E_INVALID, "bad image size");
  }
  if (cameras[0].width < 16 || cameras[0].height < 16)
    return fail(ctx, HPM_E_INVALID, "reference image must be at least 16x16");
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  free_views(ctx);
  dropped = true;
  ctx->nimg = num_images;
  ctx->cams.assign(cameras, cameras + num_images);
  ctx->W = cameras[0].width;
  ctx->H = cameras[0].height;
  int rc;
  ctx->d_img.assign(num_images, nullptr);
  for (int i = 0; i < num_images; ++i) {
    const size_t n = (size_t)cameras[i].width * cameras[i].height;
    if ((rc = dalloc(ctx, &ctx->d_img[i], n))) return rc;
    HIP_OK(hipMemcpyAsync(ctx->d_img[i], images[i], n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  }
  return alloc_pixel_state(ctx);
}

int hpm_pool_create(hpm_pool** out, int device, int num_images, const hpm_camera* cameras,
                    const float* const* images) {
  if (!out || !cameras || !images || num_images < 1) return HPM_E_INVALID;
  *out = nullptr;
  for (int i = 0; i < num_images; ++i)
    if (!images[i] || cameras[i].width <= 0 || cameras[i].height <= 0 ||
        (size_t)cameras[i].width * cameras[i].height > (1u << 28))
      return HPM_E_INVALID;
  if (hipSetDevice(device) != hipSuccess) return HPM_E_DEVICE;
  hpm_pool* pool = new hpm_pool();
  pool->device = device;
  pool->cams.assign(cameras, cameras + num_images);
  pool->d_img.assign(num_images, nullptr);
  for (int i = 0; i < num_images; ++i) {
    const size_t n = (size_t)cameras[i].width * cameras[i].height;
    if (hipMalloc((void**)&pool->d_img[i], n * sizeof(float)) != hipSuccess) {
      hpm_pool_destroy(pool);
      return HPM_E_NOMEM;
    }
    if (hipMemcpy(pool->d_img[i], images[i], n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
      hpm_pool_destroy(pool);
      return HPM_E_DEVICE;
    }
  }
  *out = pool;
  return HPM_OK;
}

void hpm_pool_destroy(hpm_pool* pool) {
  if (!pool) return;
  hipSetDevice(pool->device);
  hipDeviceSynchronize();
  for (float* p : pool->d_img) hipFree(p);
  for (auto& q : pool->pairs)
    for (void* p : q) hipFree(p);
  delete pool;
}

static int set_views_pooled(hpm_ctx* ctx, hpm_pool* pool, int num_images, const int32_t* ids, bool& dropped) {
  if (!pool || !ids) return fail(ctx, HPM_E_INVALID, "null pool or ids");
  if (pool->device != ctx->device) return fail(ctx, HPM_E_INVALID, "pool and context are on different devices");
  if (num_images < 2 || num_images > HPM_MAX_IMAGES)
    return fail(ctx, HPM_E_INVALID, "num_images must be in [2, 33] (view masks are 32-bit)");
  for (int i = 0; i < num_images; ++i)
    if (ids[i] < 0 || ids[i] >= (int)pool->d_img.size()) return fail(ctx, HPM_E_INVALID, "image id outside the pool");
  const hpm_camera& rc0 = pool->cams[ids[0]];
  if (rc0.width < 16 || rc0.height < 16) return fail(ctx, HPM_E_INVALID, "reference image must be at least 16x16");
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  free_views(ctx);
  dropped = true;
  ctx->pool = pool;
  ctx->pool_ids.assign(ids, ids + num_images);
  ctx->nimg = num_images;
  ctx->cams.clear();
  ctx->d_img.clear();
  for (int i = 0; i < num_images; ++i) {
    ctx->cams.push_back(pool->cams[ids[i]]);
    ctx->d_img.push_back(pool->d_img[ids[i]]);
  }
  ctx->W = rc0.width;
  ctx->H = rc0.height;
  return alloc_pixel_state(ctx);
}

static int alloc_pixel_state(hpm_ctx* ctx) {
  int rc;
  const size_t P = npix(ctx);
  if ((rc = dalloc(ctx, &ctx->d_planes, P)) || (rc = dalloc(ctx, &ctx->d_costs, P)) ||
      (rc = dalloc(ctx, &ctx->d_pre_costs, P)) || (rc = dalloc(ctx, &ctx->d_conf, P)) ||
      (rc = dalloc(ctx, &ctx->d_texture, P)) || (rc = dalloc(ctx, &ctx->d_sel, P)) ||
      (rc = dalloc(ctx, &ctx->d_canny, P)))
    return rc;
  HIP_OK(hipMemsetAsync(ctx->d_planes, 0, P * sizeof(float4), ctx->stream));
  HIP_OK(hipMemsetAsync(ctx->d_costs, 0, P * sizeof(float), ctx->stream));
  HIP_OK(hipMemsetAsync(ctx->d_pre_costs, 0, P * sizeof(float), ctx->stream));
  HIP_OK(hipMemsetAsync(ctx->d_conf, 0, P * sizeof(float), ctx->stream));
  HIP_OK(hipMemsetAsync(ctx->d_texture, 0, P * sizeof(float), ctx->stream));
  HIP_OK(hipMemsetAsync(ctx->d_sel, 0, P * sizeof(uint32_t), ctx->stream));
  HIP_OK(hipMemsetAsync(ctx->d_canny, 0, P * sizeof(uint32_t), ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return HPM_OK;
}

static int set_src_depths(hpm_ctx* ctx, const float* const* depths);

// A failed hpm_set_src_depths leaves no source depths on the context
// (geometric runs then return HPM_E_STATE).
int hpm_set_src_depths(hpm_ctx* ctx, const float* const* depths) {
  if (!ctx || !depths) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  const int rc = set_src_depths(ctx, depths);
  if (rc != HPM_OK) {
    for (float* p : ctx->d_dimg) if (p) hipFreeAsync(p, ctx->stream);
    ctx->d_dimg.clear();
  }
  return rc;
}

static int set_src_depths(hpm_ctx* ctx, const float* const* depths) {
  HIP_OK(hipSetDevice(ctx->device));
  for (float* p : ctx->d_dimg) if (p) hipFreeAsync(p, ctx->stream);
  ctx->d_dimg.assign(ctx->nimg, nullptr);
  for (int i = 0; i < ctx->nimg; ++i) {
    if (!depths[i]) { if (i == 0) continue; return fail(ctx, HPM_E_INVALID, "null source depth map"); }
    const size_t n = (size_t)ctx->cams[i].width * ctx->cams[i].height;
    int rc = dalloc(ctx, &ctx->d_dimg[i], n);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(ctx->d_dimg[i], depths[i], n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  }
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return HPM_OK;
}

static int upload(hpm_ctx* ctx, void* dst, const void* src, size_t bytes) {
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return HPM_OK;
}

int hpm_set_canny(hpm_ctx* ctx, const uint32_t* canny) {
  if (!ctx || !canny) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  return upload(ctx, ctx->d_canny, canny, npix(ctx) * sizeof(uint32_t));
}

int hpm_set_confidence(hpm_ctx* ctx, const float* conf) {
  if (!ctx || !conf) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  return upload(ctx, ctx->d_conf, conf, npix(ctx) * sizeof(float));
}

int hpm_reload_hypotheses(hpm_ctx* ctx, const float* planes, const float* costs) {
  if (!ctx || !planes || !costs) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  int rc = upload(ctx, ctx->d_planes, planes, npix(ctx) * sizeof(float4));
  if (rc) return rc;
  return upload(ctx, ctx->d_costs, costs, npix(ctx) * sizeof(float));
}

int hpm_set_prior(hpm_ctx* ctx, const float* prior_planes, const uint32_t* masks) {
  if (!ctx || !prior_planes || !masks) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  int rc;
  if ((rc = dalloc(ctx, &ctx->d_prior, npix(ctx))) || (rc = dalloc(ctx, &ctx->d_masks, npix(ctx)))) return rc;
  if ((rc = upload(ctx, ctx->d_prior, prior_planes, npix(ctx) * sizeof(float4)))) return rc;
  return upload(ctx, ctx->d_masks, masks, npix(ctx) * sizeof(uint32_t));
}

int hpm_release_prior(hpm_ctx* ctx) {
  if (!ctx) return HPM_E_INVALID;
  HIP_OK(hipSetDevice(ctx->device));
  void* bufs[] = {ctx->d_prior, ctx->d_masks, ctx->d_canny, ctx->d_prior_ds};
  for (void* p : bufs)
    if (p) hipFreeAsync(p, ctx->stream);
  ctx->d_prior = nullptr; ctx->d_masks = nullptr; ctx->d_canny = nullptr; ctx->d_prior_ds = nullptr;
  ctx->prior_dist = -1; ctx->prior_w = ctx->prior_h = 0;
  return HPM_OK;
}

int hpm_release(hpm_ctx* ctx) {
  if (!ctx) return HPM_E_INVALID;
  HIP_OK(hipSetDevice(ctx->device));
  if (ctx->gexec) {
    HIP_OK(hipStreamSynchronize(ctx->stream));
    hipGraphExecDestroy(ctx->gexec);
    ctx->gexec = nullptr;
    ctx->gkey.clear();
  }
  free_views(ctx);
  return HPM_OK;
}

}  // extern "C"

namespace {

DevCam dev_cam(const hpm_camera& c) {
  static_assert(sizeof(DevCam) == sizeof(hpm_camera), "camera layouts");
  DevCam d;
  std::memcpy(&d, &c, sizeof(d));
  return d;
}

unsigned grid1(size_t n) { return (unsigned)((n + 255) / 256); }

// Grow-only device scratch of the context (prior construction, map
// splitting), carved into 256-byte aligned pieces.
int ctx_scratch(hpm_ctx* ctx, const std::vector<size_t>& sizes, std::vector<char*>& out) {
  std::vector<size_t> off;
  size_t total = 0;
  for (size_t b : sizes) {
    off.push_back(total);
    total += (b + 255) & ~(size_t)255;
  }
  if (total > ctx->scratch_bytes) {
    if (ctx->d_scratch) HIP_OK(hipFreeAsync(ctx->d_scratch, ctx->stream));
    ctx->d_scratch = nullptr;
    ctx->scratch_bytes = 0;
    if (ctx_malloc_async(ctx, (void**)&ctx->d_scratch, total) != hipSuccess) {
      ctx->d_scratch = nullptr;
      return fail(ctx, HPM_E_NOMEM, "hipMallocAsync (scratch)");
    }
    ctx->scratch_bytes = total;
  }
  out.clear();
  for (size_t o : off) out.push_back((char*)ctx->d_scratch + o);
  return HPM_OK;
}

}  // namespace

extern "C" {

int hpm_prior_support_points(hpm_ctx* ctx, int distance, int32_t* out_xy, int capacity, int* count) {
  if (!ctx || !count || distance < 0 || distance > 2 || capacity < 0 || (capacity > 0 && !out_xy)) return HPM_E_INVALID;
  *count = 0;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  ctx->prior_dist = -1;  // set again once the support points are found: a failed call leaves nothing to build on
  HIP_OK(hipSetDevice(ctx->device));
  const int W = ctx->W, H = ctx->H;
  FloatMap depth{(const float*)ctx->d_planes + 3, W, H, 4}, cost{ctx->d_costs, W, H, 1}, conf{ctx->d_conf, W, H, 1};
  int scan_w = W, scan_h = H, mw = W, mh = H;
  if (distance > 0) {
    // the maps resized by 1 / (2 distance) (main.cpp:448-454); the support
    // points scan (int)(width * factor) columns (HPM.cpp:1030-1031)
    const float factor = (float)(1.0 / (distance * 2));
    const int hw = (int)std::round(W * factor), hh = (int)std::round(H * factor);
    if (hw < 1 || hh < 1) return fail(ctx, HPM_E_INVALID, "reference image too small for a prior at this distance");
    const size_t n = (size_t)hw * hh;
    int rc = dalloc(ctx, &ctx->d_prior_ds, 3 * n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_prior_downsample, dim3(grid1(n)), dim3(256), 0, ctx->stream, depth, cost, conf, hw, hh,
                       (double)W / hw, (double)H / hh, ctx->d_prior_ds);
    HIP_OK(hipGetLastError());
    depth = FloatMap{ctx->d_prior_ds, hw, hh, 1};
    cost = FloatMap{ctx->d_prior_ds + n, hw, hh, 1};
    conf = FloatMap{ctx->d_prior_ds + 2 * n, hw, hh, 1};
    scan_w = std::min((int)(W * factor), hw);
    scan_h = std::min((int)(H * factor), hh);
    mw = hw;
    mh = hh;
  }
  const size_t nb = (size_t)((scan_w + 4) / 5) * ((scan_h + 4) / 5);
  std::vector<char*> buf;
  int rc = ctx_scratch(ctx, {nb * sizeof(int4)}, buf);
  if (rc) return rc;
  int4* d_sp = (int4*)buf[0];
  hipLaunchKernelGGL(k_support_points, dim3(grid1(nb)), dim3(256), 0, ctx->stream, cost, conf, scan_w, scan_h, d_sp);
  HIP_OK(hipGetLastError());
  std::vector<int4> h(nb);
  HIP_OK(hipMemcpyAsync(h.data(), d_sp, nb * sizeof(int4), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  int n = 0;
  auto push = [&](int x, int y) {
    if (n < capacity) { out_xy[2 * n] = x; out_xy[2 * n + 1] = y; }
    ++n;
  };
  for (const int4& o : h) {  // the textured candidate first (HPM.cpp:1067-1070)
    if (o.x >= 0) push(o.x, o.y);
    if (o.z >= 0) push(o.z, o.w);
  }
  *count = n;
  ctx->prior_dist = distance;
  ctx->prior_w = mw;
  ctx->prior_h = mh;
  return HPM_OK;
}

int hpm_prior_build(hpm_ctx* ctx, const int32_t* tris, int n) {
  if (!ctx || n < 0 || (n > 0 && !tris)) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  if (ctx->prior_dist < 0) return fail(ctx, HPM_E_STATE, "hpm_prior_support_points first");
  // the raster stores triangle index + 1 in a float map: exact below 2^24
  if (n >= (1 << 24) - 1) return fail(ctx, HPM_E_INVALID, "too many triangles");
  const int W = ctx->W, H = ctx->H, mw = ctx->prior_w, mh = ctx->prior_h, distance = ctx->prior_dist;
  std::vector<PriorTri> ht((size_t)n);
  for (int t = 0; t < n; ++t) {
    const int32_t* v = tris + 6 * (size_t)t;
    for (int k = 0; k < 3; ++k)
      if (v[2 * k] < 0 || v[2 * k] >= mw || v[2 * k + 1] < 0 || v[2 * k + 1] >= mh)
        return fail(ctx, HPM_E_INVALID, "triangle corner outside the prior grid");
    // step = 1 / longest edge, as the raster loop computes it (main.cpp:393-398)
    auto len = [](int ax, int ay, int bx, int by) {
      const double dx = ax - bx, dy = ay - by;
      return (float)std::sqrt(dx * dx + dy * dy);
    };
    const float max_edge = std::max(len(v[0], v[1], v[2], v[3]), std::max(len(v[0], v[1], v[4], v[5]), len(v[2], v[3], v[4], v[5])));
    ht[t] = PriorTri{v[0], v[1], v[2], v[3], v[4], v[5], (float)(1.0 / max_edge), 0};
  }
  HIP_OK(hipSetDevice(ctx->device));
  const size_t P = npix(ctx), Pc = (size_t)mw * mh;
  int rc;
  if ((!ctx->d_prior && (rc = dalloc(ctx, &ctx->d_prior, P))) || (!ctx->d_masks && (rc = dalloc(ctx, &ctx->d_masks, P))))
    return rc;
  std::vector<char*> buf;
  const size_t nt = std::max(n, 1);
  if (distance == 0) rc = ctx_scratch(ctx, {nt * sizeof(PriorTri), nt * sizeof(float4)}, buf);
  else
    rc = ctx_scratch(ctx, {nt * sizeof(PriorTri), nt * sizeof(float4), Pc * 4, Pc * 4, Pc * sizeof(float4), P * 4}, buf);
  if (rc) return rc;
  PriorTri* d_tris = (PriorTri*)buf[0];
  float4* d_tplanes = (float4*)buf[1];
  if (n > 0) HIP_OK(hipMemcpyAsync(d_tris, ht.data(), (size_t)n * sizeof(PriorTri), hipMemcpyHostToDevice, ctx->stream));
  const DevCam cam = dev_cam(ctx->cams[0]);
  const float dmin = ctx->cams[0].depth_min * 0.6f, dmax = ctx->cams[0].depth_max * 1.2f;  // HPM.cpp:685-686
  const float factor = distance > 0 ? (float)(1.0 / (distance * 2)) : 1.0f;
  const FloatMap depth = distance == 0 ? FloatMap{(const float*)ctx->d_planes + 3, W, H, 4}
                                       : FloatMap{ctx->d_prior_ds, mw, mh, 1};
  unsigned int* d_raster = distance == 0 ? ctx->d_masks : (unsigned int*)buf[2];
  HIP_OK(hipMemsetAsync(d_raster, 0, Pc * sizeof(unsigned int), ctx->stream));
  if (n > 0) {
    hipLaunchKernelGGL(k_tri_planes, dim3(grid1(n)), dim3(256), 0, ctx->stream, d_tris, n, depth, cam, distance > 0 ? 1 : 0,
                       factor, d_tplanes);
    hipLaunchKernelGGL(k_raster, dim3(grid1(n)), dim3(256), 0, ctx->stream, d_tris, n, mw, mh, d_raster);
    HIP_OK(hipGetLastError());
  }
  if (distance == 0) {
    hipLaunchKernelGGL(k_prior_level, dim3(grid1(P)), dim3(256), 0, ctx->stream, d_tplanes, W, H, cam, dmin, dmax,
                       ctx->d_masks, ctx->d_prior);
  } else {
    float* d_cd = (float*)buf[3];
    float4* d_cn = (float4*)buf[4];
    float* d_up = (float*)buf[5];
    hipLaunchKernelGGL(k_prior_coarse, dim3(grid1(Pc)), dim3(256), 0, ctx->stream, d_tplanes, d_raster, mw, mh, cam,
                       factor, dmin, dmax, d_cd, d_cn);
    // JointBilateralUpsampling_prior (HPM.cpp:1323-1392) to the reference
    // size; with an integer scale of 1 nothing is upsampled and the upsampled
    // maps stay zero (main.cpp:548-576 then keeps no prior)
    JbuDims jd;
    jd.height = H; jd.width = W; jd.s_height = mh; jd.s_width = mw;
    jd.imagescale = std::max(H / mh, W / mw);
    if (jd.imagescale != 1) {
      hipLaunchKernelGGL(k_jbu_prior, dim3((W + 15) / 16, (H + 15) / 16), dim3(16, 16), 0, ctx->stream, jd,
                         (const float*)ctx->d_img[0], d_cd, d_cn, d_up, ctx->d_prior);
    } else {
      HIP_OK(hipMemsetAsync(d_up, 0, P * 4, ctx->stream));
      HIP_OK(hipMemsetAsync(ctx->d_prior, 0, P * sizeof(float4), ctx->stream));
    }
    hipLaunchKernelGGL(k_prior_reload, dim3(grid1(P)), dim3(256), 0, ctx->stream, d_up, ctx->d_prior, W, H, cam, dmin,
                       dmax, ctx->d_masks, ctx->d_prior);
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return HPM_OK;
}

int hpm_canny(int device, const uint8_t* image, int w, int h, int channels, float low, float high, uint8_t* out_edges) {
  if (!image || !out_edges || w <= 0 || h <= 0 || (channels != 1 && channels != 3)) return HPM_E_INVALID;
  if ((size_t)w * h >= ((size_t)1 << 31)) return HPM_E_INVALID;
  if (hipSetDevice(device) != hipSuccess) return HPM_E_DEVICE;
  const size_t P = (size_t)w * h;
  // image, gx, gy, mag, cls, edge, flag; 256-byte aligned pieces
  const size_t sizes[7] = {P * channels, 2 * P, 2 * P, 2 * P, P, P, 4};
  size_t off[7], total = 0;
  for (int k = 0; k < 7; ++k) { off[k] = total; total += (sizes[k] + 255) & ~(size_t)255; }
  hipStream_t st = nullptr;
  char* d = nullptr;
  int rc = HPM_OK;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && rc == HPM_OK) rc = (e == hipErrorOutOfMemory ? HPM_E_NOMEM : HPM_E_DEVICE);
    return rc == HPM_OK;
  };
  // a stream of its own and stream-ordered memory: concurrent passes on other
  // streams are neither waited for nor made to wait
  if (ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && ok(hipMallocAsync((void**)&d, total, st)) &&
      ok(hipMemcpyAsync(d + off[0], image, sizes[0], hipMemcpyHostToDevice, st))) {
    const uint8_t* d_img = (const uint8_t*)(d + off[0]);
    int16_t* gx = (int16_t*)(d + off[1]);
    int16_t* gy = (int16_t*)(d + off[2]);
    int16_t* mag = (int16_t*)(d + off[3]);
    uint8_t* cls = (uint8_t*)(d + off[4]);
    uint8_t* edge = (uint8_t*)(d + off[5]);
    int* flag = (int*)(d + off[6]);
    const dim3 g8((w + 31) / 32, (h + 7) / 8), g32((w + 31) / 32, (h + 31) / 32);
    hipLaunchKernelGGL(k_canny_grad, g8, dim3(256), 0, st, d_img, w, h, channels, gx, gy, mag);
    hipLaunchKernelGGL(k_canny_nms, g8, dim3(256), 0, st, gx, gy, mag, w, h, low, high, cls, edge);
    ok(hipGetLastError());
    // every pass that goes on adds at least one pixel: at most P + 1 passes
    for (size_t it = 0; it <= P && rc == HPM_OK; ++it) {
      int changed = 0;
      if (!ok(hipMemsetAsync(flag, 0, sizeof(int), st))) break;
      hipLaunchKernelGGL(k_canny_hyst, g32, dim3(256), 0, st, cls, w, h, edge, flag);
      if (!ok(hipGetLastError()) || !ok(hipMemcpyAsync(&changed, flag, sizeof(int), hipMemcpyDeviceToHost, st)) ||
          !ok(hipStreamSynchronize(st)) || !changed)
        break;
    }
    if (rc == HPM_OK) ok(hipMemcpyAsync(out_edges, edge, P, hipMemcpyDeviceToHost, st));
  }
  if (d) hipFreeAsync(d, st);
  if (st) {
    ok(hipStreamSynchronize(st));
    hipStreamDestroy(st);
  }
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// hpm_canny_noalloc's scratch: the gradient, magnitude and class planes of up
// to HPM_CANNY_NOALLOC_MAX_PIXELS pixels (7 bytes each) and the hysteresis
// flag, in a file-scope device array.  Every GPU the library runs on holds its
// own copy until the process ends.  In host code `scratch` only names the
// symbol: its device address on each device comes from hipGetSymbolAddress,
// taken once per device (with that device current) on the first call there.
namespace {
constexpr size_t kCannyScratchPlanes = (size_t)HPM_CANNY_NOALLOC_MAX_PIXELS * 7;
constexpr size_t kCannyScratchBytes = ((kCannyScratchPlanes + 255) & ~(size_t)255) + 256;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace
static __device__ __align__(256) char scratch[kCannyScratchBytes];

static int canny_scratch(int device, char** out) {
  if (device < 0 || device >= kMaxDevices) return HPM_E_DEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    char* ptr = nullptr;
    if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return HPM_E_DEVICE;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return HPM_OK;
}

extern "C" {

int hpm_canny_noalloc(const uint8_t* image_dev, int w, int h, int channels, float low, float high,
                      uint8_t* out_edges_dev) {
  if (!image_dev || !out_edges_dev || w <= 0 || h <= 0 || (channels != 1 && channels != 3)) return HPM_E_INVALID;
  if ((size_t)w * h > (size_t)HPM_CANNY_NOALLOC_MAX_PIXELS) return HPM_E_TOO_LARGE;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return HPM_E_DEVICE;
  char* base = nullptr;
  int rc = canny_scratch(device, &base);
  if (rc) return rc;
  const size_t P = (size_t)w * h;
  // gx, gy, mag (int16), cls (uint8), flag (int): 256-byte aligned pieces
  const size_t sizes[5] = {2 * P, 2 * P, 2 * P, P, 4};
  size_t off[5], total = 0;
  for (int k = 0; k < 5; ++k) { off[k] = total; total += (sizes[k] + 255) & ~(size_t)255; }
  if (total > kCannyScratchBytes) return HPM_E_TOO_LARGE;
  int16_t* gx = (int16_t*)(base + off[0]);
  int16_t* gy = (int16_t*)(base + off[1]);
  int16_t* mag = (int16_t*)(base + off[2]);
  uint8_t* cls = (uint8_t*)(base + off[3]);
  int* flag = (int*)(base + off[4]);
  const hipStream_t st = hipStreamPerThread;
  const dim3 g8((w + 31) / 32, (h + 7) / 8), g32((w + 31) / 32, (h + 31) / 32);
  // every plane is written in full by the kernel that produces it; only the
  // flag must start at zero, and it is cleared before every hysteresis pass
  hipLaunchKernelGGL(k_canny_grad, g8, dim3(256), 0, st, image_dev, w, h, channels, gx, gy, mag);
  hipLaunchKernelGGL(k_canny_nms, g8, dim3(256), 0, st, gx, gy, mag, w, h, low, high, cls, out_edges_dev);
  if (hipGetLastError() != hipSuccess) return HPM_E_DEVICE;
  // every pass that goes on adds at least one pixel: at most P + 1 passes
  for (size_t it = 0; it <= P; ++it) {
    int changed = 0;
    if (hipMemsetAsync(flag, 0, sizeof(int), st) != hipSuccess) return HPM_E_DEVICE;
    hipLaunchKernelGGL(k_canny_hyst, g32, dim3(256), 0, st, cls, w, h, out_edges_dev, flag);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&changed, flag, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return HPM_E_DEVICE;
    if (!changed) break;
  }
  return HPM_OK;
}

int hpm_get_prior(hpm_ctx* ctx, float* prior_planes, uint32_t* masks) {
  if (!ctx || !prior_planes || !masks) return HPM_E_INVALID;
  if (!ctx->d_prior || !ctx->d_masks) return fail(ctx, HPM_E_STATE, "no prior on the context");
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipMemcpyAsync(prior_planes, ctx->d_prior, npix(ctx) * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipMemcpyAsync(masks, ctx->d_masks, npix(ctx) * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return HPM_OK;
}

int hpm_set_coarse(hpm_ctx* ctx, const float* coarse, int cw, int ch, const float* fine_depth) {
  if (!ctx || !coarse || !fine_depth || cw <= 0 || ch <= 0) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  int rc = dalloc(ctx, &ctx->d_scaled, (size_t)cw * ch);
  if (rc) return rc;
  if ((rc = upload(ctx, ctx->d_scaled, coarse, (size_t)cw * ch * sizeof(float4)))) return rc;
  ctx->scaled_w = cw;
  ctx->scaled_h = ch;
  // HPM.cpp:869 compares the coarse width with the fine rows (kept as is)
  ctx->upsample = (cw != ctx->H || ch != ctx->W);
  if (!ctx->upsample && (cw != ctx->W || ch != ctx->H))
    return fail(ctx, HPM_E_INVALID, "coarse map must match the reference size when not upsampling");
  std::vector<float4> fine(npix(ctx));
  for (size_t i = 0; i < fine.size(); ++i) fine[i] = make_float4(0.0f, 0.0f, 0.0f, fine_depth[i]);
  return upload(ctx, ctx->d_planes, fine.data(), fine.size() * sizeof(float4));
}

int hpm_set_timing(hpm_ctx* ctx, int enable) {
  if (!ctx) return HPM_E_INVALID;
  ctx->timing = enable != 0;
  return HPM_OK;
}

}  // extern "C"

namespace hpm_detail {

int prepare_run(hpm_ctx* ctx, uint64_t seed, RunConsts& C, bool allow_u8) {
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  const hpm_params& P = ctx->params;
  if (P.geom_consistency && (int)ctx->d_dimg.size() != ctx->nimg)
    return fail(ctx, HPM_E_STATE, "geometric consistency needs hpm_set_src_depths");
  if (P.prior_consistency && (!ctx->d_prior || !ctx->d_masks))
    return fail(ctx, HPM_E_STATE, "prior consistency needs hpm_set_prior");
  if (P.hierarchy && !ctx->d_scaled) return fail(ctx, HPM_E_STATE, "hierarchy needs hpm_set_coarse");
  HIP_OK(hipSetDevice(ctx->device));
  int rc;
  // 8-bit column pairs pay off only in runs from random planes, whose first
  // launches scatter their gathers (3072 x 2048, 10 sources: 20.1 ms per launch
  // against 22.6); runs from converged planes (hierarchy, prior, geometric
  // passes) are bound by VALU work there and run slower on them.
  const bool random_init = !P.geom_consistency && !P.prior_consistency && !P.hierarchy && !P.mand_consistency;
  if (P.precision != HPM_PREC_FP32_EXACT && (rc = ensure_pair_images(ctx, P.precision, allow_u8 && random_init)))
    return rc;
  build_consts(ctx, C);
  C.seed = seed;
  return HPM_OK;
}

DevState state_of(hpm_ctx* ctx) { return make_state(ctx); }

int record_error(hpm_ctx* ctx, int code, const std::string& msg) { return fail(ctx, code, msg); }

}  // namespace hpm_detail

extern "C" {

int hpm_run_patchmatch_partial(hpm_ctx* ctx, uint64_t seed, int n_halfpasses, int do_output) {
  if (!ctx) return HPM_E_INVALID;
  RunConsts C;
  int rc = hpm_detail::prepare_run(ctx, seed, C);
  if (rc) return rc;
  const hpm_params& P = ctx->params;
  HIP_OK(hipMemcpyAsync(ctx->d_consts, &C, sizeof(C), hipMemcpyHostToDevice, ctx->stream));
  HIP_OK(hipMemsetAsync(ctx->d_counters, 0, 8 * sizeof(unsigned long long), ctx->stream));
  switch (P.precision) {
    case HPM_PREC_FP16:
      if (ctx->lp_kind == kStorageU8) return run_schedule<FastFor<PairU8>>(ctx, n_halfpasses, do_output);
      return run_schedule<FastFor<PairF16>>(ctx, n_halfpasses, do_output);
    case HPM_PREC_BF16: return run_schedule<FastFor<PairBF16>>(ctx, n_halfpasses, do_output);
    case HPM_PREC_FP32_FAST: return run_schedule<FastFor<PairF32>>(ctx, n_halfpasses, do_output);
    default: return run_schedule<ExactEngine>(ctx, n_halfpasses, do_output);
  }
}

int hpm_set_u8_pairs_min_bytes(hpm_ctx* ctx, long long min_bytes) {
  if (!ctx) return HPM_E_INVALID;
  ctx->u8_min_bytes = min_bytes < 0 ? -1 : min_bytes;
  return HPM_OK;
}

int hpm_set_graph(hpm_ctx* ctx, int enable) {
  if (!ctx) return HPM_E_INVALID;
  ctx->use_graph = enable != 0;
  return HPM_OK;
}

int hpm_run_patchmatch_batch(hpm_ctx* const* ctxs, int n, const uint64_t* seeds) {
  if (!ctxs || !seeds || n < 0) return HPM_E_INVALID;
  for (int i = 0; i < n; ++i)
    if (!ctxs[i]) return HPM_E_INVALID;
  for (int i = 0; i < n; ++i) {
    const int rc = hpm_run_patchmatch_partial(ctxs[i], seeds[i], -1, 1);
    if (rc) return rc;
  }
  for (int i = 0; i < n; ++i) {
    const int rc = hpm_synchronize(ctxs[i]);
    if (rc) return rc;
  }
  return HPM_OK;
}

int hpm_run_patchmatch(hpm_ctx* ctx, uint64_t seed) { return hpm_run_patchmatch_partial(ctx, seed, -1, 1); }

int hpm_synchronize(hpm_ctx* ctx) {
  if (!ctx) return HPM_E_INVALID;
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return HPM_OK;
}

int hpm_get_stats(hpm_ctx* ctx, hpm_stats* out) {
  if (!ctx || !out) return HPM_E_INVALID;
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  unsigned long long cnt[8];
  HIP_OK(hipMemcpy(cnt, ctx->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
  hpm_stats s = ctx->stats;
  s.ncc_init = cnt[0];
  s.ncc_propagate = cnt[1];
  s.ncc_reference = cnt[2] + (uint64_t)ctx->W * ctx->H * (ctx->nimg - 1);
  if (ctx->timing) {
    float a = 0, b = 0, c = 0;
    HIP_OK(hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]));
    HIP_OK(hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]));
    HIP_OK(hipEventElapsedTime(&c, ctx->ev[2], ctx->ev[3]));
    s.ms_init = a; s.ms_propagate = b; s.ms_output = c;
  }
  *out = s;
  return HPM_OK;
}

static int download(hpm_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || !dst) return HPM_E_INVALID;
  if (ctx->nimg == 0 || !src) return fail(ctx, HPM_E_STATE, "nothing to read yet");
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return HPM_OK;
}

int hpm_get_planes(hpm_ctx* ctx, float* out) { return download(ctx, out, ctx ? ctx->d_planes : nullptr, ctx ? npix(ctx) * 16 : 0); }
int hpm_get_costs(hpm_ctx* ctx, float* out) { return download(ctx, out, ctx ? ctx->d_costs : nullptr, ctx ? npix(ctx) * 4 : 0); }
int hpm_get_texture(hpm_ctx* ctx, float* out) { return download(ctx, out, ctx ? ctx->d_texture : nullptr, ctx ? npix(ctx) * 4 : 0); }
int hpm_get_selected_views(hpm_ctx* ctx, uint32_t* out) { return download(ctx, out, ctx ? ctx->d_sel : nullptr, ctx ? npix(ctx) * 4 : 0); }
int hpm_get_confidence(hpm_ctx* ctx, float* out) { return download(ctx, out, ctx ? ctx->d_conf : nullptr, ctx ? npix(ctx) * 4 : 0); }

int hpm_set_canny_u8(hpm_ctx* ctx, const uint8_t* canny) {
  if (!ctx || !canny) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  HIP_OK(hipSetDevice(ctx->device));
  const size_t P = npix(ctx);
  std::vector<char*> buf;
  int rc = ctx_scratch(ctx, {P}, buf);
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(buf[0], canny, P, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_widen_edges, dim3(grid1(P)), dim3(256), 0, ctx->stream, (const uint8_t*)buf[0], P, ctx->d_canny);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return HPM_OK;
}

int hpm_reload_hypotheses_maps(hpm_ctx* ctx, const float* depth, const float* normals, const float* costs) {
  if (!ctx || !depth || !normals || !costs) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  HIP_OK(hipSetDevice(ctx->device));
  const size_t P = npix(ctx);
  std::vector<char*> buf;
  int rc = ctx_scratch(ctx, {P * sizeof(float), 3 * P * sizeof(float)}, buf);
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(buf[0], depth, P * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  HIP_OK(hipMemcpyAsync(buf[1], normals, 3 * P * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  HIP_OK(hipMemcpyAsync(ctx->d_costs, costs, P * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_join_planes, dim3(grid1(P)), dim3(256), 0, ctx->stream, (const float*)buf[0],
                     (const float*)buf[1], P, ctx->d_planes);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return HPM_OK;
}

int hpm_get_depth_normals(hpm_ctx* ctx, float* out_depth, float* out_normals) {
  if (!ctx || !out_depth || !out_normals) return HPM_E_INVALID;
  if (ctx->nimg == 0 || !ctx->d_planes) return fail(ctx, HPM_E_STATE, "nothing to read yet");
  HIP_OK(hipSetDevice(ctx->device));
  const size_t P = npix(ctx);
  std::vector<char*> buf;
  int rc = ctx_scratch(ctx, {P * sizeof(float), 3 * P * sizeof(float)}, buf);
  if (rc) return rc;
  hipLaunchKernelGGL(k_split_planes, dim3(grid1(P)), dim3(256), 0, ctx->stream, ctx->d_planes, P, (float*)buf[0],
                     (float*)buf[1]);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out_depth, buf[0], P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipMemcpyAsync(out_normals, buf[1], 3 * P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return HPM_OK;
}

int hpm_get_size(const hpm_ctx* ctx, int* w, int* h, int* n) {
  if (!ctx) return HPM_E_INVALID;
  if (w) *w = ctx->W;
  if (h) *h = ctx->H;
  if (n) *n = ctx->nimg;
  return HPM_OK;
}

int hpm_cost_fixed_device(hpm_ctx* ctx, const void* planes_dev, void* out_dev, int precision) {
  if (!ctx || !planes_dev || !out_dev) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  if (precision < 0 || precision > 3) return fail(ctx, HPM_E_INVALID, "unknown precision");
  HIP_OK(hipSetDevice(ctx->device));
  int rc;
  if (precision != HPM_PREC_FP32_EXACT && (rc = ensure_pair_images(ctx, precision))) return rc;
  RunConsts C;
  build_consts(ctx, C);
  HIP_OK(hipMemcpyAsync(ctx->d_consts, &C, sizeof(C), hipMemcpyHostToDevice, ctx->stream));
  switch (precision) {
    case HPM_PREC_FP16:
      if (ctx->lp_kind == kStorageU8) return launch_cost_fixed<FastFor<PairU8>>(ctx, planes_dev, out_dev);
      return launch_cost_fixed<FastFor<PairF16>>(ctx, planes_dev, out_dev);
    case HPM_PREC_BF16: return launch_cost_fixed<FastFor<PairBF16>>(ctx, planes_dev, out_dev);
    case HPM_PREC_FP32_FAST: return launch_cost_fixed<FastFor<PairF32>>(ctx, planes_dev, out_dev);
    default: return launch_cost_fixed<ExactEngine>(ctx, planes_dev, out_dev);
  }
}

int hpm_cost_fixed(hpm_ctx* ctx, const float* planes, float* out, int precision) {
  if (!ctx || !planes || !out) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  HIP_OK(hipSetDevice(ctx->device));
  const size_t P = npix(ctx), V = ctx->nimg - 1;
  float4* dp = nullptr;
  float* dout = nullptr;
  int rc;
  if ((rc = dalloc(ctx, &dp, P)) || (rc = dalloc(ctx, &dout, P * V))) { dfree(ctx, dp); return rc; }
  rc = upload(ctx, dp, planes, P * sizeof(float4));
  if (!rc) rc = hpm_cost_fixed_device(ctx, dp, dout, precision);
  if (!rc) rc = download(ctx, out, dout, P * V * sizeof(float));
  dfree(ctx, dp);
  dfree(ctx, dout);
  return rc;
}

int hpm_init_cost(hpm_ctx* ctx, const float* planes, float* out_cost, uint32_t* out_masks) {
  if (!ctx || !planes || !out_cost || !out_masks) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  HIP_OK(hipSetDevice(ctx->device));
  const size_t P = npix(ctx);
  float4* dp = nullptr;
  float* dc = nullptr;
  uint32_t* dm = nullptr;
  int rc;
  if ((rc = dalloc(ctx, &dp, P)) || (rc = dalloc(ctx, &dc, P)) || (rc = dalloc(ctx, &dm, P))) {
    dfree(ctx, dp); dfree(ctx, dc); return rc;
  }
  RunConsts C;
  build_consts(ctx, C);
  rc = upload(ctx, ctx->d_consts, &C, sizeof(C));
  if (!rc) rc = upload(ctx, dp, planes, P * sizeof(float4));
  if (!rc) {
    const dim3 b16(16, 16), g16((ctx->W + 15) / 16, (ctx->H + 15) / 16);
    hipLaunchKernelGGL(k_init_cost<ExactEngine>, g16, b16, 0, ctx->stream, ctx->d_consts, dp, dc, dm);
    if (hipGetLastError() != hipSuccess) rc = fail(ctx, HPM_E_DEVICE, "k_init_cost launch failed");
  }
  if (!rc) rc = download(ctx, out_cost, dc, P * sizeof(float));
  if (!rc) rc = download(ctx, out_masks, dm, P * sizeof(uint32_t));
  dfree(ctx, dp); dfree(ctx, dc); dfree(ctx, dm);
  return rc;
}

int hpm_geom_cost(hpm_ctx* ctx, const float* planes, float* out) {
  if (!ctx || !planes || !out) return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  if ((int)ctx->d_dimg.size() != ctx->nimg) return fail(ctx, HPM_E_STATE, "hpm_set_src_depths first");
  HIP_OK(hipSetDevice(ctx->device));
  const size_t P = npix(ctx), V = ctx->nimg - 1;
  float4* dp = nullptr;
  float* dout = nullptr;
  int rc;
  if ((rc = dalloc(ctx, &dp, P)) || (rc = dalloc(ctx, &dout, P * V))) { dfree(ctx, dp); return rc; }
  RunConsts C;
  build_consts(ctx, C);
  rc = upload(ctx, ctx->d_consts, &C, sizeof(C));
  if (!rc) rc = upload(ctx, dp, planes, P * sizeof(float4));
  if (!rc) {
    const dim3 b16(16, 16), g16((ctx->W + 15) / 16, (ctx->H + 15) / 16);
    hipLaunchKernelGGL(k_geom_cost, g16, b16, 0, ctx->stream, ctx->d_consts, dp, dout);
    if (hipGetLastError() != hipSuccess) rc = fail(ctx, HPM_E_DEVICE, "k_geom_cost launch failed");
  }
  if (!rc) rc = download(ctx, out, dout, P * V * sizeof(float));
  dfree(ctx, dp); dfree(ctx, dout);
  return rc;
}


// ------------------------------------------------------------------ gradient

static int grad_setup(hpm_ctx* ctx) {
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  HIP_OK(hipSetDevice(ctx->device));
  int rc = ensure_pair_images(ctx, HPM_PREC_FP32_FAST);
  if (rc) return rc;
  RunConsts C;
  build_consts(ctx, C);
  HIP_OK(hipMemcpyAsync(ctx->d_consts, &C, sizeof(C), hipMemcpyHostToDevice, ctx->stream));
  return HPM_OK;
}

static int grad_launch(hpm_ctx* ctx, const void* params, const void* weights, void* cost, void* grad, void* gn,
                       const RunConsts* consts = nullptr) {
  const dim3 b(16, 16), g((ctx->W + 15) / 16, (ctx->H + 15) / 16);
  const RunConsts* C = consts ? consts : ctx->d_consts;
  if (grad)
    hipLaunchKernelGGL(k_cost_grad<true>, g, b, 0, ctx->stream, C, (const float4*)params, (const float*)weights,
                       (float*)cost, (float*)grad, (float*)gn);
  else
    hipLaunchKernelGGL(k_cost_grad<false>, g, b, 0, ctx->stream, C, (const float4*)params, (const float*)weights,
                       (float*)cost, nullptr, nullptr);
  HIP_OK(hipGetLastError());
  return HPM_OK;
}

// The views smoothed by a Gaussian of `sigma` px (radius ceil(3 sigma), at
// most 16), as FP32 images and FP32 column-pair images, and a constant block
// reading them:
// the cost of the smoothed images has a basin of convergence as wide as the
// blur, where the sharp images' is about a texel of their finest texture.
struct SmoothedViews {
  std::vector<float*> img;
  std::vector<void*> pairs;
  RunConsts* d_consts = nullptr;
  ~SmoothedViews() {
    for (float* p : img) hipFree(p);
    for (void* p : pairs) hipFree(p);
    hipFree(d_consts);
  }
};

static int build_smoothed(hpm_ctx* ctx, float sigma, SmoothedViews& S) {
  BlurTaps t{};
  t.r = std::min(kBlurMaxRadius, std::max(1, (int)std::ceil(3.0f * sigma)));
  double sum = 0.0;
  for (int k = -t.r; k <= t.r; ++k) sum += std::exp(-0.5 * (double)k * k / ((double)sigma * sigma));
  for (int k = -t.r; k <= t.r; ++k) t.w[k + t.r] = (float)(std::exp(-0.5 * (double)k * k / ((double)sigma * sigma)) / sum);
  float* tmp = nullptr;
  size_t max_px = 0;
  for (int i = 0; i < ctx->nimg; ++i) max_px = std::max(max_px, (size_t)ctx->cams[i].width * ctx->cams[i].height);
  int rc = dalloc(ctx, &tmp, max_px);
  if (rc) return rc;
  S.img.assign(ctx->nimg, nullptr);
  S.pairs.assign(ctx->nimg, nullptr);
  for (int i = 0; i < ctx->nimg && rc == HPM_OK; ++i) {
    const int w = ctx->cams[i].width, h = ctx->cams[i].height;
    if ((rc = dalloc(ctx, &S.img[i], (size_t)w * h))) break;
    if (injected_fault(ctx) || hipMalloc(&S.pairs[i], pair_image_bytes<PairF32>(w, h)) != hipSuccess) {
      S.pairs[i] = nullptr;
      rc = fail(ctx, HPM_E_NOMEM, "smoothed pair image allocation failed");
      break;
    }
    const dim3 b(16, 16), g((w + 15) / 16, (h + 15) / 16);
    hipLaunchKernelGGL(k_blur<true>, g, b, 0, ctx->stream, ctx->d_img[i], tmp, w, h, t);
    hipLaunchKernelGGL(k_blur<false>, g, b, 0, ctx->stream, tmp, S.img[i], w, h, t);
    const dim3 gq((w + 2 + 15) / 16, (h + 1 + 15) / 16);
    hipLaunchKernelGGL(k_build_pairs<PairF32>, gq, b, 0, ctx->stream, S.img[i], w, h, S.pairs[i]);
    if (hipGetLastError() != hipSuccess) rc = fail(ctx, HPM_E_DEVICE, "smoothing launch failed");
  }
  if (rc == HPM_OK) {
    RunConsts C;
    build_consts(ctx, C);
    for (int i = 0; i < ctx->nimg; ++i) {
      C.img[i] = S.img[i];
      C.pairs[i] = S.pairs[i];
    }
    C.storage = HPM_PREC_FP32_FAST;
    if ((rc = dalloc(ctx, &S.d_consts, 1)) == HPM_OK &&
        hipMemcpyAsync(S.d_consts, &C, sizeof(C), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
      rc = fail(ctx, HPM_E_DEVICE, "constant upload failed");
    if (rc == HPM_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, HPM_E_DEVICE, "smoothing failed");
  }
  hipFree(tmp);
  return rc;
}

int hpm_cost_grad_device(hpm_ctx* ctx, const void* params_dev, const void* weights_dev, void* out_cost_dev,
                         void* out_grad_dev, void* out_gn_dev) {
  if (!ctx || !params_dev || !weights_dev || !out_cost_dev) return HPM_E_INVALID;
  int rc = grad_setup(ctx);
  if (rc) return rc;
  return grad_launch(ctx, params_dev, weights_dev, out_cost_dev, out_grad_dev, out_grad_dev ? out_gn_dev : nullptr);
}

// Device buffers of the host-pointer gradient entry points.
struct GradBuffers {
  float4* params = nullptr;
  float* weights = nullptr;
  float* cost = nullptr;
  float* grad = nullptr;
  float* gn = nullptr;
  float* lambda = nullptr;
  float4* cand = nullptr;
  float* cand_cost = nullptr;
  ~GradBuffers() {
    hipFree(params); hipFree(weights); hipFree(cost); hipFree(grad); hipFree(gn); hipFree(lambda);
    hipFree(cand); hipFree(cand_cost);
  }
};

static int grad_alloc(hpm_ctx* ctx, GradBuffers& B, const float* params, const float* weights, bool refine) {
  const size_t P = npix(ctx), V = ctx->nimg - 1;
  int rc;
  if ((rc = dalloc(ctx, &B.params, P)) || (rc = dalloc(ctx, &B.weights, P * V)) || (rc = dalloc(ctx, &B.cost, P)) ||
      (rc = dalloc(ctx, &B.grad, 3 * P)) || (rc = dalloc(ctx, &B.gn, 6 * P)))
    return rc;
  if (refine && ((rc = dalloc(ctx, &B.lambda, P)) || (rc = dalloc(ctx, &B.cand, P)) ||
                 (rc = dalloc(ctx, &B.cand_cost, P))))
    return rc;
  if ((rc = upload(ctx, B.params, params, P * sizeof(float4))) || (rc = upload(ctx, B.weights, weights, P * V * 4)))
    return rc;
  return HPM_OK;
}

int hpm_cost_grad(hpm_ctx* ctx, const float* params, const float* weights, float* out_cost, float* out_grad,
                  float* out_gn) {
  if (!ctx || !params || !weights || !out_cost) return HPM_E_INVALID;
  int rc = grad_setup(ctx);
  if (rc) return rc;
  GradBuffers B;
  const size_t P = npix(ctx);
  if ((rc = grad_alloc(ctx, B, params, weights, false))) return rc;
  if ((rc = grad_launch(ctx, B.params, B.weights, B.cost, out_grad ? B.grad : nullptr, out_gn ? B.gn : nullptr)))
    return rc;
  if ((rc = download(ctx, out_cost, B.cost, P * 4))) return rc;
  if (out_grad && (rc = download(ctx, out_grad, B.grad, 3 * P * 4))) return rc;
  if (out_grad && out_gn && (rc = download(ctx, out_gn, B.gn, 6 * P * 4))) return rc;
  return HPM_OK;
}

// Refinement loop: per iteration a random-perturbation proposal (kRandom),
// a Gauss-Newton / Levenberg-Marquardt step (kGaussNewton), or both in turn
// (kHybrid: the random proposal reaches the basin, the GN step converges in
// it); every proposal is kept only where it lowers the cost.
enum RefineMode { kRandom = 0, kGaussNewton = 1, kHybrid = 2 };
static int refine_common(hpm_ctx* ctx, float* params, const float* weights, int iterations, float* out_cost,
                         int mode, uint64_t seed, int smoothed_iterations = 0, const RunConsts* smoothed = nullptr) {
  if (!ctx || !params || !weights || iterations < 0) return HPM_E_INVALID;
  int rc = grad_setup(ctx);
  if (rc) return rc;
  GradBuffers B;
  const int P = (int)npix(ctx);
  if ((rc = grad_alloc(ctx, B, params, weights, true))) return rc;
  const float dmin = ctx->cams[0].depth_min * 0.6f, dmax = ctx->cams[0].depth_max * 1.2f;
  const dim3 b1(256), g1((P + 255) / 256);
  hipLaunchKernelGGL(k_fill_f32, g1, b1, 0, ctx->stream, B.lambda, P, 1e-2f);
  HIP_OK(hipGetLastError());
  // the first smoothed_iterations steps run on the smoothed views' cost
  const RunConsts* cur = (smoothed && smoothed_iterations > 0) ? smoothed : nullptr;
  const bool gn_first = mode == kGaussNewton;
  if ((rc = grad_launch(ctx, B.params, B.weights, B.cost, gn_first ? B.grad : nullptr, gn_first ? B.gn : nullptr,
                        cur)))
    return rc;
  auto propose_and_accept = [&](bool gn_step) -> int {
    if ((rc = grad_launch(ctx, B.cand, B.weights, B.cand_cost, nullptr, nullptr, cur))) return rc;
    hipLaunchKernelGGL(k_accept, g1, b1, 0, ctx->stream, P, B.params, B.cost, B.cand, B.cand_cost,
                       gn_step ? B.lambda : nullptr);
    HIP_OK(hipGetLastError());
    return HPM_OK;
  };
  for (int it = 0; it < iterations; ++it) {
    if (cur && it == smoothed_iterations) {  // to the sharp views: their cost, their gradient, a fresh damping
      cur = nullptr;
      hipLaunchKernelGGL(k_fill_f32, g1, b1, 0, ctx->stream, B.lambda, P, 1e-2f);
      HIP_OK(hipGetLastError());
      if ((rc = grad_launch(ctx, B.params, B.weights, B.cost, gn_first ? B.grad : nullptr, gn_first ? B.gn : nullptr)))
        return rc;
    }
    const bool fresh = cur == nullptr && smoothed && it == smoothed_iterations;
    if (mode != kGaussNewton) {
      const float decay = powf(0.8f, (float)it);
      hipLaunchKernelGGL(k_random_step, g1, b1, 0, ctx->stream, P, seed, (uint32_t)(1000 + it), B.params,
                         0.02f * decay, 0.0872665f * decay, dmin, dmax, B.cand);
      HIP_OK(hipGetLastError());
      if ((rc = propose_and_accept(false))) return rc;
    }
    if (mode != kRandom) {
      if ((it > 0 || mode == kHybrid) && !fresh && (rc = grad_launch(ctx, B.params, B.weights, B.cost, B.grad, B.gn, cur)))
        return rc;
      hipLaunchKernelGGL(k_gn_step, g1, b1, 0, ctx->stream, P, B.params, B.grad, B.gn, B.lambda, dmin, dmax, B.cand);
      HIP_OK(hipGetLastError());
      if ((rc = propose_and_accept(true))) return rc;
    }
  }
  if (cur) {  // every step ran on the smoothed views: report the sharp views' cost
    if ((rc = grad_launch(ctx, B.params, B.weights, B.cost, nullptr, nullptr))) return rc;
  }
  if ((rc = download(ctx, params, B.params, (size_t)P * sizeof(float4)))) return rc;
  if (out_cost && (rc = download(ctx, out_cost, B.cost, (size_t)P * 4))) return rc;
  return HPM_OK;
}

int hpm_refine_gn(hpm_ctx* ctx, float* params, const float* weights, int iterations, float* out_cost) {
  return refine_common(ctx, params, weights, iterations, out_cost, kGaussNewton, 0);
}

int hpm_refine_random(hpm_ctx* ctx, float* params, const float* weights, int iterations, uint64_t seed,
                      float* out_cost) {
  return refine_common(ctx, params, weights, iterations, out_cost, kRandom, seed);
}

int hpm_refine_hybrid(hpm_ctx* ctx, float* params, const float* weights, int iterations, uint64_t seed,
                      float* out_cost) {
  return refine_common(ctx, params, weights, iterations, out_cost, kHybrid, seed);
}

int hpm_refine_gn_smoothed(hpm_ctx* ctx, float* params, const float* weights, int iterations,
                           int smoothed_iterations, float sigma, float* out_cost) {
  if (!ctx || !params || !weights || iterations < 0 || smoothed_iterations < 0 || !(sigma > 0.0f) || sigma > 5.0f)
    return HPM_E_INVALID;
  if (ctx->nimg == 0) return fail(ctx, HPM_E_STATE, "hpm_set_views first");
  HIP_OK(hipSetDevice(ctx->device));
  SmoothedViews S;
  int rc = build_smoothed(ctx, sigma, S);
  if (rc) return rc;
  return refine_common(ctx, params, weights, iterations, out_cost, kGaussNewton, 0,
                       std::min(smoothed_iterations, iterations), S.d_consts);
}

}  // extern "C"
