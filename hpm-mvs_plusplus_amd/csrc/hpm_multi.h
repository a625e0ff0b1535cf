// hpm_multi.h -- the PatchMatch schedule of many reference views in one launch
// per stage.  Every view keeps its own constant block and pixel state; a
// launch's grid is the concatenation of the views' tile grids, each view
// starting at a multiple of 8 workgroups so that the XCD-aware tile walk of
// the single-view kernels applies inside it unchanged.  A workgroup finds its
// view through a per-workgroup table (one 16-bit load) and runs exactly the
// single-view body on that view's tile, so results are those of the
// single-view kernels.
//
// The views' constant blocks and states are arrays passed as read-only,
// non-aliasing kernel arguments: indexed by the wave-uniform view number,
// their fields load through the scalar cache as the single-view kernels'
// arguments do (a constant block reached through a pointer read from memory
// would load through the vector path and leave the image pointers, which
// feed the buffer resources, in vector registers).
#pragma once
#include "hpm_kernels.h"

namespace hpm {

enum { kStageInit = 0, kStageColour = 1, kStageOutput = 2 };

struct MultiView {
  int H2;        // rows the colour passes visit, (H / 2) * 2
  int first[3];  // first workgroup of the view in the init / colour / output grids
  int count[3];  // the view's workgroups in each grid
};

// The host sizes every view's init grid in 16 x 16 tiles: one-lane engines.
template <class E>
__global__ HPM_PROP_ATTR void k_init_multi(const RunConsts* __restrict__ Cs, const DevState* __restrict__ Ss,
                                           const MultiView* __restrict__ views, const uint16_t* __restrict__ wg,
                                           int strip) {
  static_assert(E::kLanes == 1, "multi-view launches take one-lane engines");
  const int v = __builtin_amdgcn_readfirstlane((int)wg[blockIdx.x]);
  const int id = (int)blockIdx.x - views[v].first[kStageInit];
  if (id >= views[v].count[kStageInit]) return;  // padding up to the next multiple of 8
  init_body<E>(Cs[v], Ss[v], id, strip);
}

// NT threads per workgroup (256 or 128), CW same-colour columns per tile.
template <class E, bool kLds, int CW, int NT>
__global__ HPM_PROP_ATTR_NT(NT) void k_propagate_multi(const RunConsts* __restrict__ Cs, const DevState* __restrict__ Ss,
                                                       const MultiView* __restrict__ views,
                                                       const uint16_t* __restrict__ wg, int colour, int iter, int strip) {
  static_assert(E::kLanes == 1, "multi-view launches take one-lane engines");
  const int v = __builtin_amdgcn_readfirstlane((int)wg[blockIdx.x]);
  const int id = (int)blockIdx.x - views[v].first[kStageColour];
  if (id >= views[v].count[kStageColour]) return;
  propagate_body<E, kLds, CW, NT>(Cs[v], Ss[v], colour, iter, views[v].H2, id, strip);
}

// 256 consecutive pixels (row-major) per workgroup
static __global__ __launch_bounds__(256) void k_output_multi(const RunConsts* __restrict__ Cs,
                                                             const DevState* __restrict__ Ss,
                                                             const MultiView* __restrict__ views,
                                                             const uint16_t* __restrict__ wg) {
  const int v = __builtin_amdgcn_readfirstlane((int)wg[blockIdx.x]);
  const int id = (int)blockIdx.x - views[v].first[kStageOutput];
  if (id >= views[v].count[kStageOutput]) return;
  const RunConsts& C = Cs[v];
  const int p = id * 256 + (int)threadIdx.x;
  if (p >= C.W * C.H) return;
  output_pixel(C, Ss[v].planes, p % C.W, p / C.W);
}

}  // namespace hpm
