// hpm_internal.h -- the context and image-pool structures behind the opaque
// handles of include/hpm_mi355x.h, shared by the translation units of the
// library (single-view schedule and C ABI; fused multi-view schedule).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hpm_mi355x.h"
#include "hpm_fast.h"

// The mixed-precision engine of storage type S (one lane per pixel).
template <class S> using FastFor = hpm::FastEngine<S>;

// Images of a whole scene on one device, shared by the contexts of its
// reference views: FP32 images plus, per storage precision, the column-pair images of
// the mixed-precision engine (built on first use).
struct hpm_pool {
  int device = 0;
  std::vector<hpm_camera> cams;
  std::vector<float*> d_img;
  std::vector<void*> pairs[5];  // indexed by storage kind: HPM_PREC_* or kStorageU8 (0 unused)
  int u8_ok = -1;               // every image integral in [0, 255]: 1 yes, 0 no, -1 not checked
  std::mutex pairs_mu;          // contexts on several host threads may build them first
};

struct hpm_ctx {
  int device = 0;
  hpm_pool* pool = nullptr;        // images borrowed from a pool (not owned)
  std::vector<int> pool_ids;
  hipStream_t stream = nullptr;
  std::string err;
  hpm_params params{};
  int nimg = 0, W = 0, H = 0;
  std::vector<hpm_camera> cams;
  std::vector<float*> d_img;
  std::vector<float*> d_dimg;
  std::vector<void*> d_img_lp;     // column-pair images of the mixed-precision engine
  int lp_kind = -1;                // their storage kind (HPM_PREC_FP16 / BF16 / FP32_FAST, kStorageU8)
  // FP16 precision: 8-bit column pairs when the views' FP16 column pairs
  // would take at least this many bytes (-1: never) and every texel is an
  // integer in [0, 255] (u8_ok, checked once per set of views)
  long long u8_min_bytes = 192ll << 20;
  int u8_ok = -1;
  // Allocation fault injection: with HPM_FAULT_ALLOC_AFTER=n in the
  // environment at hpm_create, the context's allocation number n (from 0,
  // after creation) fails as if the device were out of memory, once; -1 = off.
  long long fault_alloc = -1;
  float4* d_planes = nullptr;
  float* d_costs = nullptr;
  float* d_pre_costs = nullptr;
  float* d_conf = nullptr;
  float* d_texture = nullptr;
  uint32_t* d_sel = nullptr;
  uint32_t* d_masks = nullptr;
  uint32_t* d_canny = nullptr;
  float4* d_prior = nullptr;
  float4* d_scaled = nullptr;
  int scaled_w = 0, scaled_h = 0;
  // planar prior built on the device: the distance and grid of the support
  // points last computed and the maps resized to that grid (distance > 0)
  int prior_dist = -1, prior_w = 0, prior_h = 0;
  float* d_prior_ds = nullptr;
  // grow-only scratch: the prior's triangles and intermediate maps, split maps
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  bool upsample = false;
  hpm::RunConsts* d_consts = nullptr;
  unsigned long long* d_counters = nullptr;
  bool timing = false;
  hipEvent_t ev[8] = {};
  hpm_stats stats{};
  // hipGraph of the kernel sequence of one RunPatchMatch, replayed while the
  // kernel arguments it was captured with stay valid
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> gkey;  // what the captured arguments depend on
  int g_launches = 0;
};

namespace hpm_detail {
// Validates the context for a run, builds the column-pair images of its storage
// precision when needed and fills the run's constant block (seed included);
// nothing is launched.  Returns HPM_OK or an error recorded on the context.
// allow_u8 = false keeps the FP16 precision on FP16 column pairs (a fused
// batch whose views do not all qualify for 8-bit pairs).
int prepare_run(hpm_ctx* ctx, uint64_t seed, hpm::RunConsts& C, bool allow_u8 = true);
hpm::DevState state_of(hpm_ctx* ctx);
int record_error(hpm_ctx* ctx, int code, const std::string& msg);
}  // namespace hpm_detail
