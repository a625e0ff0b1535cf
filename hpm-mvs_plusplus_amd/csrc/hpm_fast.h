// hpm_fast.h -- mixed-precision (FP16 / BF16 image storage, FP32 accumulation)
// PatchMatch path.  Declarations only; the kernels live in hpm_fast.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct hpm_ctx;

namespace hpm {
struct FastConsts {
  int placeholder;
};
}  // namespace hpm

int hpm_fast_run(hpm_ctx* ctx, uint64_t seed, int n_halfpasses, int do_output);
int hpm_fast_cost_fixed(hpm_ctx* ctx, const void* planes_dev, void* out_dev, int precision);
