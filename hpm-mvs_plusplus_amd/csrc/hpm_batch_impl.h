// hpm_batch_impl.h -- the batch plan behind the opaque hpm_batch handle and
// the per-engine launch code of the fused multi-view schedule.  The engines'
// instantiations are spread over two translation units so that they compile
// in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "hpm_internal.h"
#include "hpm_multi.h"

struct hpm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<hpm_ctx*> ctxs;
  std::vector<hpm::RunConsts> consts;  // host sources of the per-run constant uploads
  // layout: per-view slots and the per-workgroup view tables of the three grids
  std::vector<unsigned char> layout_key;
  hpm::RunConsts* d_consts = nullptr;  // one constant block per view, uploaded per run
  hpm::DevState* d_states = nullptr;
  hpm::MultiView* d_views = nullptr;
  uint16_t* d_wg[3] = {nullptr, nullptr, nullptr};
  // host sources of the layout uploads, kept until the next layout (the
  // uploads are asynchronous on `stream`)
  std::vector<hpm::MultiView> h_views;
  std::vector<hpm::DevState> h_states;
  std::vector<uint16_t> h_wg[3];
  int grid[3] = {0, 0, 0};
  int max_v = 0;
  int storage = -1;  // column-pair storage of the last run (precision, or kStorageU8)
  // launch shape
  int tile_w = HPM_COLOUR_TILE_W;
  int strip = HPM_XCD_STRIP;
  int threads = 256;  // threads (pixels) per colour-pass workgroup: 256 or 128
  // graph
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> gkey;
  hipEvent_t ev[2] = {nullptr, nullptr};
  float last_ms = 0.0f;
};

namespace hpm_batch_detail {

int bfail(hpm_batch* b, int code, const std::string& msg);

#define BHIP_OK(call)                                                                                     \
  do {                                                                                                    \
    hipError_t e_ = (call);                                                                               \
    if (e_ != hipSuccess) return bfail(b, HPM_E_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T> void append(std::vector<unsigned char>& k, const T& v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  k.insert(k.end(), p, p + sizeof(T));
}

using namespace hpm;

inline size_t lds_bytes(const hpm_batch* b) { return (size_t)8 * b->max_v * b->threads * sizeof(float); }

template <class E, int CW, int NT> int set_lds(hpm_batch* b) {
  if (b->max_v <= kLdsMaxSrc)
    BHIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_propagate_multi<E, true, CW, NT>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes(b)));
  return HPM_OK;
}

template <class E, int CW, int NT> int launch_stages(hpm_batch* b, int iters) {
  hipStream_t st = b->stream;
  const dim3 blk(256), blk_colour(NT);
  hipLaunchKernelGGL(k_init_multi<E>, dim3(b->grid[kStageInit]), blk, 0, st, b->d_consts, b->d_states, b->d_views,
                     b->d_wg[kStageInit], b->strip);
  BHIP_OK(hipGetLastError());
  for (int it = 0; it < iters; ++it)
    for (int colour = 0; colour < 2; ++colour) {
      if (b->max_v <= kLdsMaxSrc)
        hipLaunchKernelGGL((k_propagate_multi<E, true, CW, NT>), dim3(b->grid[kStageColour]), blk_colour, lds_bytes(b),
                           st, b->d_consts, b->d_states, b->d_views, b->d_wg[kStageColour], colour, it, b->strip);
      else
        hipLaunchKernelGGL((k_propagate_multi<E, false, CW, NT>), dim3(b->grid[kStageColour]), blk_colour, 0, st,
                           b->d_consts, b->d_states, b->d_views, b->d_wg[kStageColour], colour, it, b->strip);
      BHIP_OK(hipGetLastError());
    }
  hipLaunchKernelGGL(k_output_multi, dim3(b->grid[kStageOutput]), blk, 0, st, b->d_consts, b->d_states, b->d_views,
                     b->d_wg[kStageOutput]);
  BHIP_OK(hipGetLastError());
  return HPM_OK;
}

template <class E, int CW, int NT> int run_stages(hpm_batch* b, int precision, int iters) {
  int rc = set_lds<E, CW, NT>(b);
  if (rc) return rc;
  if (!b->use_graph) return launch_stages<E, CW, NT>(b, iters);
  std::vector<unsigned char> key = b->layout_key;
  append(key, precision);
  append(key, b->storage);
  append(key, iters);
  append(key, b->strip);
  append(key, b->threads);
  append(key, b->grid);
  append(key, b->max_v);
  if (!b->gexec || key != b->gkey) {
    if (b->gexec) { BHIP_OK(hipGraphExecDestroy(b->gexec)); b->gexec = nullptr; }
    hipGraph_t g = nullptr;
    BHIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = launch_stages<E, CW, NT>(b, iters);
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return bfail(b, HPM_E_DEVICE, std::string("stream capture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) { b->gexec = nullptr; return bfail(b, HPM_E_DEVICE, "hipGraphInstantiate failed"); }
    b->gkey = key;
  }
  BHIP_OK(hipGraphLaunch(b->gexec, b->stream));
  return HPM_OK;
}

// Launch shapes instantiated per engine: colour tile widths 16, 32 and 64 and
// 256 or 128 threads per workgroup for the mixed-precision engines; the exact
// engine keeps the default width and 256 threads.
template <class E, bool kAllShapes> int run_engine(hpm_batch* b, int precision, int iters) {
  if (b->tile_w == 32 && b->threads == 256) return run_stages<E, 32, 256>(b, precision, iters);
  if constexpr (kAllShapes) {
    if (b->threads == 256) {
      if (b->tile_w == 16) return run_stages<E, 16, 256>(b, precision, iters);
      if (b->tile_w == 64) return run_stages<E, 64, 256>(b, precision, iters);
    } else if (b->threads == 128) {
      if (b->tile_w == 16) return run_stages<E, 16, 128>(b, precision, iters);
      if (b->tile_w == 32) return run_stages<E, 32, 128>(b, precision, iters);
      if (b->tile_w == 64) return run_stages<E, 64, 128>(b, precision, iters);
    }
  }
  return bfail(b, HPM_E_INVALID, "launch shape not available for this precision");
}

}  // namespace hpm_batch_detail

// run_engine of the FP16 and BF16 engines is instantiated in hpm_batch_lp.hip
extern template int hpm_batch_detail::run_engine<FastFor<hpm::PairF16>, true>(hpm_batch*, int, int);
extern template int hpm_batch_detail::run_engine<FastFor<hpm::PairU8>, true>(hpm_batch*, int, int);
extern template int hpm_batch_detail::run_engine<FastFor<hpm::PairBF16>, true>(hpm_batch*, int, int);
