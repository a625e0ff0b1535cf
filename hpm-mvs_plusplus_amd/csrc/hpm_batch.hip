// hpm_batch.hip -- fused multi-view RunPatchMatch: the schedules of many
// contexts (reference views of any size and source count) run as one launch
// per stage -- init, every colour pass, output -- over the concatenated tile
// grids of all views, optionally replayed as one hipGraph.  The launch shape
// (colour tile width, XCD strip height, pixels per colour workgroup) is a
// run-time choice that hpm_batch_autotune picks by timing.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "hpm_batch_impl.h"

using namespace hpm;
using namespace hpm_batch_detail;

int hpm_batch_detail::bfail(hpm_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

namespace {

// The layout's device tables are allocated, uploaded and freed in the order
// of the batch's own stream (stream-ordered allocator): a batch made and
// dropped per RunPatchMatch, as the host pipeline does for large views, then
// never waits for the kernels other host threads run on the device.
void free_layout(hpm_batch* b) {
  void* bufs[] = {b->d_consts, b->d_states, b->d_views, b->d_wg[0], b->d_wg[1], b->d_wg[2]};
  for (void* p : bufs)
    if (p) hipFreeAsync(p, b->stream);
  b->d_consts = nullptr;
  b->d_states = nullptr;
  b->d_views = nullptr;
  for (auto& p : b->d_wg) p = nullptr;
}

int colour_tiles(int tile_w, int threads, int W, int H2) {
  const int ch = threads / tile_w;
  return (((W + 1) / 2 + tile_w - 1) / tile_w) * ((H2 + ch - 1) / ch);
}

int round8(int n) { return (n + 7) & ~7; }

// (Re)builds the slots and workgroup tables when a view's buffers, size or
// source count, or the colour tile shape, changed since the last run.
int ensure_layout(hpm_batch* b) {
  std::vector<unsigned char> key;
  append(key, b->tile_w);
  append(key, b->threads);
  for (hpm_ctx* c : b->ctxs) {
    append(key, hpm_detail::state_of(c));
    append(key, c->W);
    append(key, c->H);
    append(key, c->nimg);
  }
  if (key == b->layout_key && b->d_views) return HPM_OK;
  const int n = (int)b->ctxs.size();
  BHIP_OK(hipStreamSynchronize(b->stream));  // the previous layout's uploads have read their host sources
  std::vector<MultiView>& views = b->h_views;
  std::vector<DevState>& states = b->h_states;
  std::vector<uint16_t>* wg = b->h_wg;
  views.assign(n, MultiView());
  states.assign(n, DevState());
  for (int s = 0; s < 3; ++s) wg[s].clear();
  b->max_v = 0;
  for (int i = 0; i < n; ++i) {
    hpm_ctx* c = b->ctxs[i];
    MultiView& m = views[i];
    std::memset(&m, 0, sizeof(m));
    states[i] = hpm_detail::state_of(c);
    m.H2 = (c->H / 2) * 2;
    m.count[kStageInit] = tiles_2d<ExactEngine>(c->W, c->H);  // 16 x 16 tiles for every one-lane engine
    m.count[kStageColour] = colour_tiles(b->tile_w, b->threads, c->W, m.H2);
    m.count[kStageOutput] = (int)(((size_t)c->W * c->H + 255) / 256);
    for (int s = 0; s < 3; ++s) {
      m.first[s] = (int)wg[s].size();
      wg[s].insert(wg[s].end(), (size_t)round8(m.count[s]), (uint16_t)i);
    }
    if (c->nimg - 1 > b->max_v) b->max_v = c->nimg - 1;
  }
  free_layout(b);
  b->layout_key.clear();  // set again only once every table is in place
  hipStream_t st = b->stream;
  BHIP_OK(hipMallocAsync((void**)&b->d_consts, sizeof(RunConsts) * n, st));
  BHIP_OK(hipMallocAsync((void**)&b->d_states, sizeof(DevState) * n, st));
  BHIP_OK(hipMallocAsync((void**)&b->d_views, sizeof(MultiView) * n, st));
  BHIP_OK(hipMemcpyAsync(b->d_states, states.data(), sizeof(DevState) * n, hipMemcpyHostToDevice, st));
  BHIP_OK(hipMemcpyAsync(b->d_views, views.data(), sizeof(MultiView) * n, hipMemcpyHostToDevice, st));
  for (int s = 0; s < 3; ++s) {
    if (wg[s].size() > (size_t)0x7fffffff) return bfail(b, HPM_E_INVALID, "batch grid too large");
    b->grid[s] = (int)wg[s].size();
    BHIP_OK(hipMallocAsync((void**)&b->d_wg[s], sizeof(uint16_t) * wg[s].size(), st));
    BHIP_OK(hipMemcpyAsync(b->d_wg[s], wg[s].data(), sizeof(uint16_t) * wg[s].size(), hipMemcpyHostToDevice, st));
  }
  b->layout_key = key;
  return HPM_OK;
}

bool shape_for(int precision, int tile_w, int threads) {
  if (tile_w == 32 && threads == 256) return true;
  return precision != HPM_PREC_FP32_EXACT && (tile_w == 16 || tile_w == 32 || tile_w == 64) &&
         (threads == 128 || threads == 256);
}

int run_once(hpm_batch* b, const uint64_t* seeds) {
  const int n = (int)b->ctxs.size();
  b->consts.resize(n);
  int precision = -1, iters = -1;
  for (int i = 0; i < n; ++i) {
    hpm_ctx* c = b->ctxs[i];
    if (c->device != b->device) return bfail(b, HPM_E_INVALID, "contexts on different devices");
    if (i == 0) { precision = c->params.precision; iters = c->params.num_iterations; }
    if (c->params.precision != precision || c->params.num_iterations != iters)
      return bfail(b, HPM_E_INVALID, "a fused batch needs one precision and one iteration count");
  }
  // one storage for the whole launch: 8-bit column pairs only if every view
  // takes them, else FP16 pairs for all
  for (bool allow_u8 : {true, false}) {
    bool all_u8 = true, any_u8 = false;
    for (int i = 0; i < n; ++i) {
      hpm_ctx* c = b->ctxs[i];
      const int rc = hpm_detail::prepare_run(c, seeds[i], b->consts[i], allow_u8);
      if (rc) return bfail(b, rc, "view " + std::to_string(i) + ": " + c->err);
      all_u8 = all_u8 && c->lp_kind == kStorageU8;
      any_u8 = any_u8 || c->lp_kind == kStorageU8;
    }
    b->storage = all_u8 ? kStorageU8 : precision;
    if (all_u8 || !any_u8) break;
  }
  BHIP_OK(hipSetDevice(b->device));
  for (hpm_ctx* c : b->ctxs) BHIP_OK(hipStreamSynchronize(c->stream));  // pair-image builds, uploads
  if (!shape_for(precision, b->tile_w, b->threads))
    return bfail(b, HPM_E_INVALID, "launch shape not available for this precision");
  int rc = ensure_layout(b);
  if (rc) return rc;
  BHIP_OK(hipMemcpyAsync(b->d_consts, b->consts.data(), sizeof(RunConsts) * n, hipMemcpyHostToDevice, b->stream));
  for (int i = 0; i < n; ++i)
    BHIP_OK(hipMemsetAsync(b->ctxs[i]->d_counters, 0, 8 * sizeof(unsigned long long), b->stream));
  BHIP_OK(hipEventRecord(b->ev[0], b->stream));
  switch (precision) {
    case HPM_PREC_FP16:
      rc = b->storage == kStorageU8 ? run_engine<FastFor<PairU8>, true>(b, precision, iters)
                                    : run_engine<FastFor<PairF16>, true>(b, precision, iters);
      break;
    case HPM_PREC_BF16: rc = run_engine<FastFor<PairBF16>, true>(b, precision, iters); break;
    case HPM_PREC_FP32_FAST: rc = run_engine<FastFor<PairF32>, true>(b, precision, iters); break;
    default: rc = run_engine<ExactEngine, false>(b, precision, iters); break;
  }
  if (rc) return rc;
  BHIP_OK(hipEventRecord(b->ev[1], b->stream));
  BHIP_OK(hipStreamSynchronize(b->stream));
  BHIP_OK(hipEventElapsedTime(&b->last_ms, b->ev[0], b->ev[1]));
  for (hpm_ctx* c : b->ctxs) c->stats.propagate_launches = 2 * iters;
  return HPM_OK;
}

}  // namespace

extern "C" {

int hpm_batch_create(hpm_batch** out, hpm_ctx* const* ctxs, int n) {
  if (!out) return HPM_E_INVALID;
  *out = nullptr;
  if (!ctxs || n < 1 || n > 65535) return HPM_E_INVALID;
  for (int i = 0; i < n; ++i)
    if (!ctxs[i] || ctxs[i]->device != ctxs[0]->device) return HPM_E_INVALID;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j)
      if (ctxs[i] == ctxs[j]) return HPM_E_INVALID;  // one view's state cannot be two slots
  hpm_batch* b = new hpm_batch();
  b->device = ctxs[0]->device;
  b->ctxs.assign(ctxs, ctxs + n);
  if (hipSetDevice(b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess ||
      hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess) {
    hpm_batch_destroy(b);
    return HPM_E_DEVICE;
  }
  *out = b;
  return HPM_OK;
}

void hpm_batch_destroy(hpm_batch* b) {
  if (!b) return;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->stream) {
    free_layout(b);
    hipStreamSynchronize(b->stream);
  }
  for (auto e : b->ev) if (e) hipEventDestroy(e);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

const char* hpm_batch_last_error(const hpm_batch* b) { return b ? b->err.c_str() : "null batch"; }

int hpm_batch_set_graph(hpm_batch* b, int enable) {
  if (!b) return HPM_E_INVALID;
  b->use_graph = enable != 0;
  return HPM_OK;
}

int hpm_batch_set_shape(hpm_batch* b, int tile_w, int strip) {
  if (!b) return HPM_E_INVALID;
  if (tile_w == 0) tile_w = HPM_COLOUR_TILE_W;
  if (tile_w != 16 && tile_w != 32 && tile_w != 64) return bfail(b, HPM_E_INVALID, "tile width must be 16, 32 or 64");
  if (strip < 1 || strip > 8) return bfail(b, HPM_E_INVALID, "strip must be in [1, 8]");
  b->tile_w = tile_w;
  b->strip = strip;
  return HPM_OK;
}

int hpm_batch_get_shape(const hpm_batch* b, int* tile_w, int* strip) {
  if (!b || !tile_w || !strip) return HPM_E_INVALID;
  *tile_w = b->tile_w;
  *strip = b->strip;
  return HPM_OK;
}

int hpm_batch_set_workgroup(hpm_batch* b, int threads) {
  if (!b) return HPM_E_INVALID;
  if (threads == 0) threads = 256;
  if (threads != 128 && threads != 256) return bfail(b, HPM_E_INVALID, "threads per workgroup must be 128 or 256");
  b->threads = threads;
  return HPM_OK;
}

int hpm_batch_get_workgroup(const hpm_batch* b, int* threads) {
  if (!b || !threads) return HPM_E_INVALID;
  *threads = b->threads;
  return HPM_OK;
}

int hpm_batch_run(hpm_batch* b, const uint64_t* seeds) {
  if (!b || !seeds) return HPM_E_INVALID;
  return run_once(b, seeds);
}

float hpm_batch_last_ms(const hpm_batch* b) { return b ? b->last_ms : 0.0f; }

int hpm_batch_autotune(hpm_batch* b, const uint64_t* seeds, float* best_ms) {
  if (!b || !seeds) return HPM_E_INVALID;
  for (hpm_ctx* c : b->ctxs) {
    const hpm_params& P = c->params;
    if (P.geom_consistency || P.prior_consistency || P.hierarchy || P.mand_consistency)
      return bfail(b, HPM_E_STATE, "autotuning reruns the schedule: photometric passes only (init overwrites the state)");
  }
  const int precision = b->ctxs[0]->params.precision;
  const int widths[3] = {16, 32, 64}, strips[3] = {1, 2, 4}, threads[2] = {256, 128};
  int best_w = b->tile_w, best_s = b->strip, best_t = b->threads;
  float best = -1.0f;
  for (int t : threads)
    for (int w : widths) {
      if (!shape_for(precision, w, t)) continue;
      for (int s : strips) {
        b->tile_w = w;
        b->strip = s;
        b->threads = t;
        const int rc = run_once(b, seeds);
        if (rc) return rc;
        if (best < 0.0f || b->last_ms < best) { best = b->last_ms; best_w = w; best_s = s; best_t = t; }
      }
    }
  b->tile_w = best_w;
  b->strip = best_s;
  b->threads = best_t;
  // leave the views with the result of the chosen shape
  const int rc = run_once(b, seeds);
  if (rc) return rc;
  if (best_ms) *best_ms = best;
  return HPM_OK;
}

}  // extern "C"
