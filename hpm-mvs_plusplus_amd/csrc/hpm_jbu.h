// hpm_jbu.h -- joint bilateral upsampling kernels.
//
//   k_jbu_depth  JBU_cu        (HPM.cu:3037-3095), driven by RunJBU (HPM.cpp:1266-1319)
//   k_jbu_prior  JBU_cu_prior  (HPM.cu:3130-3197), driven by
//                JointBilateralUpsampling_prior (HPM.cpp:1323-1392)
//
// One thread per fine pixel, 16 x 16 blocks.  The coarse texture reads of the
// reference are at integer texel centres (tex2D(r + 0.5)), i.e. plain clamped
// loads; SpatialGauss / RangeGauss (HPM.cu:142-152) square in float and take
// exp in double, which uses the project's own double exp.
#pragma once
#include <initializer_list>
#include <vector>

#include "hpm_device.h"

namespace hpm {

struct JbuDims {
  int height, width;      // fine
  int s_height, s_width;  // coarse
  int imagescale;         // max(rows / s_rows, cols / s_cols) (HPM.cpp:1270)
};

__device__ __forceinline__ float jbu_weight(float o_x, float o_y, int r_x, int r_y, float refPix, float nbPix) {
  const float sigmad = 0.50f, sigmar = 25.5f;
  const float dx = o_x - (float)r_x, dy = o_y - (float)r_y;
  const float dis = dx * dx + dy * dy - 0.0f;
  const float sg = (float)exp_hpm(-1.0 * (double)dis / (double)(2 * sigmad * sigmad));
  const float xr = fabsf(refPix - nbPix) - 0.0f;
  const float rg = (float)exp_hpm(-1.0 * (double)(xr * xr) / (double)(2 * sigmar * sigmar));
  return sg * rg;
}

__global__ __launch_bounds__(256) void k_jbu_depth(JbuDims d, const float* __restrict__ fine, const float* __restrict__ coarse,
                                                   float* __restrict__ out) {
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= d.width || py >= d.height) return;
  const float scale = (float)(1.0 * d.s_width / d.width);
  const int nn = (d.imagescale * d.imagescale + 1) / 2;
  const float o_y = (float)py * scale, o_x = (float)px * scale;
  const float refPix = tex_at(fine, d.width, d.height, px, py);
  float total = 0.0f, norm = 0.0f;
  for (int j = -nn; j <= nn; ++j) {
    int r_y = (int)(o_y + (float)j);
    r_y = (r_y > 0 ? (r_y < d.s_height ? r_y : d.s_height - 1) : 0);
    int r_ys = py + j;
    r_ys = (r_ys > 0 ? (r_ys < d.height ? r_ys : d.height - 1) : 0);
    for (int i = -nn; i <= nn; ++i) {
      int r_x = (int)(o_x + (float)i);
      r_x = (r_x > 0 ? (r_x < d.s_width ? r_x : d.s_width - 1) : 0);
      const float srcPix = coarse[r_y * d.s_width + r_x];
      int r_xs = px + i;
      r_xs = (r_xs > 0 ? (r_xs < d.width ? r_xs : d.width - 1) : 0);
      const float nb = fine[r_ys * d.width + r_xs];
      const float tg = jbu_weight(o_x, o_y, r_x, r_y, refPix, nb);
      norm += tg;
      total += srcPix * tg;
    }
  }
  out[py * d.width + px] = total / norm;
}

__global__ __launch_bounds__(256) void k_jbu_prior(JbuDims d, const float* __restrict__ fine, const float* __restrict__ coarse,
                                                   const float4* __restrict__ coarse_n, float* __restrict__ out_depth,
                                                   float4* __restrict__ out_n) {
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= d.width || py >= d.height) return;
  const float scale = (float)(1.0 * d.s_width / d.width);
  const int nn = (d.imagescale * d.imagescale + 1) / 2;
  const float o_y = (float)py * scale, o_x = (float)px * scale;
  const float refPix = tex_at(fine, d.width, d.height, px, py);
  float ctot = 0.0f, norm = 0.0f;
  float4 nt = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int j = -nn; j <= nn; ++j) {
    int r_y = (int)(o_y + (float)j);
    r_y = (r_y > 0 ? (r_y < d.s_height ? r_y : d.s_height - 1) : 0);
    const int r_ys = py + j;  // not clamped in JBU_cu_prior; the texture clamps
    for (int i = -nn; i <= nn; ++i) {
      int r_x = (int)(o_x + (float)i);
      r_x = (r_x > 0 ? (r_x < d.s_width ? r_x : d.s_width - 1) : 0);
      const int s_center = r_y * d.s_width + r_x;
      const float srcPix = coarse[s_center];
      const float4 sn = coarse_n[s_center];
      const float nb = tex_at(fine, d.width, d.height, px + i, r_ys);
      const float tg = jbu_weight(o_x, o_y, r_x, r_y, refPix, nb);
      norm += tg;
      ctot += srcPix * tg;
      nt.x = nt.x + sn.x * tg;
      nt.y = nt.y + sn.y * tg;
      nt.z = nt.z + sn.z * tg;
    }
  }
  nt.x = nt.x / norm; nt.y = nt.y / norm; nt.z = nt.z / norm;
  normalize3(nt);
  out_depth[py * d.width + px] = ctot / norm;
  out_n[py * d.width + px] = nt;
}

}  // namespace hpm

extern "C" {

static int jbu_common(const float* fine_gray, int w, int h, const float* coarse_depth, int cw, int ch, hpm::JbuDims* d) {
  if (!fine_gray || !coarse_depth || w <= 0 || h <= 0 || cw <= 0 || ch <= 0) return -1;
  d->height = h; d->width = w; d->s_height = ch; d->s_width = cw;
  const int a = h / ch, b = w / cw;
  d->imagescale = a > b ? a : b;
  return 0;
}

// Both run on a non-blocking stream of their own with stream-ordered
// memory, so that they neither wait for nor hold up the contexts other host
// threads drive; they return when the result is on the host.
namespace {
struct JbuRun {
  hipStream_t st = nullptr;
  char* d = nullptr;
  int rc = 0;
  bool ok(hipError_t e) {
    if (e != hipSuccess && rc == 0) rc = e == hipErrorOutOfMemory ? -4 : -3;
    return rc == 0;
  }
  // one device block holding pieces of the given sizes, 256-byte aligned
  std::vector<char*> alloc(std::initializer_list<size_t> sizes) {
    size_t total = 0;
    std::vector<size_t> off;
    for (size_t b : sizes) { off.push_back(total); total += (b + 255) & ~(size_t)255; }
    std::vector<char*> out;
    if (ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && ok(hipMallocAsync((void**)&d, total, st)))
      for (size_t o : off) out.push_back(d + o);
    return out;
  }
  ~JbuRun() {
    if (d) hipFreeAsync(d, st);
    if (st) {
      hipStreamSynchronize(st);
      hipStreamDestroy(st);
    }
  }
};
}  // namespace

int hpm_jbu_depth(const float* fine_gray, int w, int h, const float* coarse_depth, int cw, int ch, float* out_depth) {
  hpm::JbuDims d;
  if (!out_depth || jbu_common(fine_gray, w, h, coarse_depth, cw, ch, &d)) return -1;
  const size_t nf = (size_t)w * h, nc = (size_t)cw * ch;
  JbuRun r;
  const std::vector<char*> b = r.alloc({nf * 4, nc * 4, nf * 4});
  if (r.rc) return r.rc;
  float *df = (float*)b[0], *dc = (float*)b[1], *dout = (float*)b[2];
  if (r.ok(hipMemcpyAsync(df, fine_gray, nf * 4, hipMemcpyHostToDevice, r.st)) &&
      r.ok(hipMemcpyAsync(dc, coarse_depth, nc * 4, hipMemcpyHostToDevice, r.st))) {
    hipLaunchKernelGGL(hpm::k_jbu_depth, dim3((w + 15) / 16, (h + 15) / 16), dim3(16, 16), 0, r.st, d, df, dc, dout);
    if (r.ok(hipGetLastError()) && r.ok(hipMemcpyAsync(out_depth, dout, nf * 4, hipMemcpyDeviceToHost, r.st)))
      r.ok(hipStreamSynchronize(r.st));
  }
  return r.rc;
}

int hpm_jbu_prior(const float* fine_gray, int w, int h, const float* coarse_depth, const float* coarse_normals, int cw,
                  int ch, float* out_depth, float* out_normals) {
  hpm::JbuDims d;
  if (!out_depth || !out_normals || !coarse_normals || jbu_common(fine_gray, w, h, coarse_depth, cw, ch, &d)) return -1;
  const size_t nf = (size_t)w * h, nc = (size_t)cw * ch;
  JbuRun r;
  const std::vector<char*> b = r.alloc({nf * 4, nc * 4, nf * 4, nc * 16, nf * 16});
  if (r.rc) return r.rc;
  float *df = (float*)b[0], *dc = (float*)b[1], *dout = (float*)b[2];
  float4 *dn = (float4*)b[3], *don = (float4*)b[4];
  if (r.ok(hipMemcpyAsync(df, fine_gray, nf * 4, hipMemcpyHostToDevice, r.st)) &&
      r.ok(hipMemcpyAsync(dc, coarse_depth, nc * 4, hipMemcpyHostToDevice, r.st)) &&
      r.ok(hipMemcpyAsync(dn, coarse_normals, nc * 16, hipMemcpyHostToDevice, r.st))) {
    hipLaunchKernelGGL(hpm::k_jbu_prior, dim3((w + 15) / 16, (h + 15) / 16), dim3(16, 16), 0, r.st, d, df, dc, dn, dout,
                       don);
    if (r.ok(hipGetLastError()) && r.ok(hipMemcpyAsync(out_depth, dout, nf * 4, hipMemcpyDeviceToHost, r.st)) &&
        r.ok(hipMemcpyAsync(out_normals, don, nf * 16, hipMemcpyDeviceToHost, r.st)))
      r.ok(hipStreamSynchronize(r.st));
  }
  return r.rc;
}

}  // extern "C"
