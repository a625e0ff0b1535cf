// hpm_multi.h -- the PatchMatch schedule of many reference views in one launch
// per stage.  Every view keeps its own constant block and pixel state; a
// launch's grid is the concatenation of the views' tile grids, each view
// starting at a multiple of 8 workgroups so that the XCD-aware tile walk of
// the single-view kernels applies inside it unchanged.  A workgroup finds its
// view through a per-workgroup table (one 16-bit load) and runs exactly the
// single-view body on that view's tile, so results are those of the
// single-view kernels.
#pragma once
#include "hpm_kernels.h"

namespace hpm {

enum { kStageInit = 0, kStageColour = 1, kStageOutput = 2 };

struct MultiSlot {
  const RunConsts* Cp;  // the view's constant block (device memory)
  DevState S;           // the view's pixel state
  int H2;               // rows the colour passes visit, (H / 2) * 2
  int first[3];         // first workgroup of the view in the init / colour / output grids
  int count[3];         // the view's workgroups in each grid
};

// The host sizes every view's init grid in 16 x 16 tiles: one-lane engines.
template <class E>
__global__ HPM_PROP_ATTR void k_init_multi(const MultiSlot* __restrict__ slots, const uint16_t* __restrict__ wg, int strip) {
  static_assert(E::kLanes == 1, "multi-view launches take one-lane engines");
  const MultiSlot& m = slots[wg[blockIdx.x]];
  const int id = (int)blockIdx.x - m.first[kStageInit];
  if (id >= m.count[kStageInit]) return;  // padding up to the next multiple of 8
  const DevState S = m.S;
  init_body<E>(*m.Cp, S, id, strip);
}

template <class E, bool kLds, int CW>
__global__ HPM_PROP_ATTR void k_propagate_multi(const MultiSlot* __restrict__ slots, const uint16_t* __restrict__ wg,
                                                int colour, int iter, int strip) {
  static_assert(E::kLanes == 1, "multi-view launches take one-lane engines");
  const MultiSlot& m = slots[wg[blockIdx.x]];
  const int id = (int)blockIdx.x - m.first[kStageColour];
  if (id >= m.count[kStageColour]) return;
  const DevState S = m.S;
  propagate_body<E, kLds, CW>(*m.Cp, S, colour, iter, m.H2, id, strip);
}

// 256 consecutive pixels (row-major) per workgroup
static __global__ __launch_bounds__(256) void k_output_multi(const MultiSlot* __restrict__ slots, const uint16_t* __restrict__ wg) {
  const MultiSlot& m = slots[wg[blockIdx.x]];
  const int id = (int)blockIdx.x - m.first[kStageOutput];
  if (id >= m.count[kStageOutput]) return;
  const RunConsts& C = *m.Cp;
  const int p = id * 256 + (int)threadIdx.x;
  if (p >= C.W * C.H) return;
  output_pixel(C, m.S.planes, p % C.W, p / C.W);
}

}  // namespace hpm
