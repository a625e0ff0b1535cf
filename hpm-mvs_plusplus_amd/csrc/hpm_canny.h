// hpm_canny.h -- the edge map of ProcessProblem (cv::Canny(image, 50, 150),
// main.cpp:280-289) on the device, with the semantics of the host CannyEdges:
//
//   k_canny_grad  3 x 3 Sobel with replicated borders; a colour image keeps,
//                 per pixel, the channel of largest |dx| + |dy| (channels
//                 visited B, G, R as in an OpenCV BGR image, first wins ties)
//   k_canny_nms   non-maximum suppression along the gradient sector
//                 (tan 22.5 / tan 67.5 degree tests in double), then weak
//                 (> low) and strong (> high) classes; strong pixels are edges
//   k_canny_hyst  hysteresis: weak pixels 8-connected to an edge through weak
//                 pixels become edges.  The set is order-free, so it is grown
//                 tile by tile in LDS until no tile changes, which gives the
//                 pixels the host's depth-first stack reaches.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hpm {

__device__ __forceinline__ int canny_px(const uint8_t* __restrict__ img, int w, int h, int c, int C, int x, int y) {
  x = x < 0 ? 0 : (x >= w ? w - 1 : x);
  y = y < 0 ? 0 : (y >= h ? h - 1 : y);
  return (int)img[((size_t)y * w + x) * C + c];
}

__global__ __launch_bounds__(256) void k_canny_grad(const uint8_t* __restrict__ img, int w, int h, int C,
                                                    int16_t* __restrict__ gx, int16_t* __restrict__ gy,
                                                    int16_t* __restrict__ mag) {
  const int x = blockIdx.x * 32 + (threadIdx.x & 31), y = blockIdx.y * 8 + (threadIdx.x >> 5);
  if (x >= w || y >= h) return;
  int best = -1, bx = 0, by = 0;
  for (int k = 0; k < C; ++k) {
    const int c = C - 1 - k;
    const int dx = (canny_px(img, w, h, c, C, x + 1, y - 1) + 2 * canny_px(img, w, h, c, C, x + 1, y) +
                    canny_px(img, w, h, c, C, x + 1, y + 1)) -
                   (canny_px(img, w, h, c, C, x - 1, y - 1) + 2 * canny_px(img, w, h, c, C, x - 1, y) +
                    canny_px(img, w, h, c, C, x - 1, y + 1));
    const int dy = (canny_px(img, w, h, c, C, x - 1, y + 1) + 2 * canny_px(img, w, h, c, C, x, y + 1) +
                    canny_px(img, w, h, c, C, x + 1, y + 1)) -
                   (canny_px(img, w, h, c, C, x - 1, y - 1) + 2 * canny_px(img, w, h, c, C, x, y - 1) +
                    canny_px(img, w, h, c, C, x + 1, y - 1));
    const int v = abs(dx) + abs(dy);
    if (v > best) { best = v; bx = dx; by = dy; }
  }
  const size_t i = (size_t)y * w + x;
  gx[i] = (int16_t)bx;
  gy[i] = (int16_t)by;
  mag[i] = (int16_t)best;
}

// cls: 0 none, 1 weak, 2 strong; edge = 255 on strong pixels, else 0.
__global__ __launch_bounds__(256) void k_canny_nms(const int16_t* __restrict__ gx, const int16_t* __restrict__ gy,
                                                   const int16_t* __restrict__ mag, int w, int h, float low, float high,
                                                   uint8_t* __restrict__ cls, uint8_t* __restrict__ edge) {
  const int x = blockIdx.x * 32 + (threadIdx.x & 31), y = blockIdx.y * 8 + (threadIdx.x >> 5);
  if (x >= w || y >= h) return;
  const size_t i = (size_t)y * w + x;
  auto m = [&](int xx, int yy) { return (xx < 0 || yy < 0 || xx >= w || yy >= h) ? 0 : (int)mag[(size_t)yy * w + xx]; };
  const int v = mag[i];
  uint8_t c = 0;
  if (!((float)v <= low)) {
    const int ix = gx[i], iy = gy[i];
    const double ax = abs(ix), ay = abs(iy);
    const double t22 = 0.41421356237309503, t67 = 2.414213562373095;  // tan 22.5, tan 67.5 degrees
    bool peak;
    if (ay <= ax * t22) peak = v > m(x - 1, y) && v >= m(x + 1, y);
    else if (ay >= ax * t67) peak = v > m(x, y - 1) && v >= m(x, y + 1);
    else if ((ix ^ iy) < 0) peak = v > m(x + 1, y - 1) && v >= m(x - 1, y + 1);
    else peak = v > m(x - 1, y - 1) && v >= m(x + 1, y + 1);
    if (peak) c = (float)v > high ? 2 : 1;
  }
  cls[i] = c;
  edge[i] = c == 2 ? 255 : 0;
}

// One 32 x 32 tile per workgroup of 256 lanes (32 x 8, four rows each), with
// a one-pixel halo whose edge state is read, not written.  Grows the edge set
// inside the tile until it is stable; *changed = 1 if the tile added a pixel.
__global__ __launch_bounds__(256) void k_canny_hyst(const uint8_t* __restrict__ cls, int w, int h,
                                                    uint8_t* __restrict__ edge, int* __restrict__ changed) {
  __shared__ uint8_t s_cls[34][34];
  __shared__ uint8_t s_edge[34][34];
  const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
  for (int k = threadIdx.x; k < 34 * 34; k += 256) {
    const int ly = k / 34, lx = k - ly * 34;
    const int x = x0 + lx - 1, y = y0 + ly - 1;
    const bool in = x >= 0 && y >= 0 && x < w && y < h;
    s_cls[ly][lx] = in ? cls[(size_t)y * w + x] : 0;
    s_edge[ly][lx] = in ? edge[(size_t)y * w + x] : 0;
  }
  __syncthreads();
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  int grew = 0;
  for (;;) {
    int step = 0;
    for (int r = 0; r < 4; ++r) {
      const int lx = tx + 1, ly = ty + 8 * r + 1;
      if (s_cls[ly][lx] == 1 && !s_edge[ly][lx]) {
        const bool nb = s_edge[ly - 1][lx - 1] | s_edge[ly - 1][lx] | s_edge[ly - 1][lx + 1] | s_edge[ly][lx - 1] |
                        s_edge[ly][lx + 1] | s_edge[ly + 1][lx - 1] | s_edge[ly + 1][lx] | s_edge[ly + 1][lx + 1];
        if (nb) {
          s_edge[ly][lx] = 255;
          step = 1;
        }
      }
    }
    // every lane reaches this barrier in every round; the loop ends when no
    // lane of the tile added a pixel (each round that goes on adds one)
    if (!__syncthreads_or(step)) break;
    grew = 1;
  }
  if (grew) {
    for (int r = 0; r < 4; ++r) {
      const int x = x0 + tx, y = y0 + ty + 8 * r;
      if (x < w && y < h) edge[(size_t)y * w + x] = s_edge[ty + 8 * r + 1][tx + 1];
    }
    if (threadIdx.x == 0) atomicOr(changed, 1);
  }
}

}  // namespace hpm
