// hpm_kernels.h -- the PatchMatch kernels of the exact FP32 path.
//
//   k_init        RandomInitialization (HPM.cu:567-763), one thread per pixel
//   k_propagate   BlackPixelUpdate / RedPixelUpdate (HPM.cu:2793-2827) with
//                 CheckerboardPropagation (HPM.cu:1347-1873) and its
//                 mandatory-consistency variant (HPM.cu:1875-2791)
//   k_output      GetDepthandNormal (HPM.cu:2829-2842)
//
// Red/black scheme: a colour pass updates the pixels with (x + y) even
// (black, colour 0) or odd (red, colour 1).  Every neighbour a pixel reads
// (far/diagonal samples HPM.cu:1367-1607, the extended left-up samples
// HPM.cu:1070-1100, the view-selection neighbours HPM.cu:1690) lies at an odd
// offset, i.e. on the other colour, so a pass has no read/write races and its
// result does not depend on the order pixels are visited in.
#pragma once
#include "hpm_device.h"

// Occupancy target of the PatchMatch kernels (waves per SIMD); the register
// budget follows from it (512 / waves VGPRs per lane).
#ifndef HPM_PROP_WAVES
#define HPM_PROP_WAVES 2
#endif
#define HPM_PROP_ATTR_NT(NT) __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(HPM_PROP_WAVES)))
#define HPM_PROP_ATTR HPM_PROP_ATTR_NT(256)

namespace hpm {

// Cost rows of the 8 sampled directions, direction-major (row d, view v at
// d * V + v).  They live in LDS, thread-interleaved (element k of a thread at
// ca[k * threads per block]) so that a wave's accesses are bank-conflict
// free; views beyond kLdsMaxSrc fall back to a private array.
constexpr int kRowStride = 256;  // threads per block of the single-view colour passes
constexpr int kLdsMaxSrc = 20;   // 8 * 20 * 256 * 4 B = 160 KiB
struct Sample {
  float* ca;
  int cs;  // element stride: threads per block in LDS, 1 in a private array
  int pos[8];
  bool flag[8];
};

// Pixel tiles of an NT-thread (1D) block, NT = 256 unless a fused launch
// picks 128.  Engines with one lane per pixel cover 16 x 16 pixels (2D
// kernels) or CW same-colour columns x NT / CW rows (colour passes, CW =
// HPM_COLOUR_TILE_W unless a launch picks another width); four-lane engines
// cover 8 x 8 and 8 x 8.
#ifndef HPM_COLOUR_TILE_W
#define HPM_COLOUR_TILE_W 32
#endif
template <class E, int CW = (E::kLanes == 1 ? HPM_COLOUR_TILE_W : 8), int NT = 256> struct Tiles {
  static constexpr int px_per_block = NT / E::kLanes;
  static constexpr int sq = E::kLanes == 1 ? 16 : 8;  // 2D tile side
  static constexpr int cw = CW;                       // colour tile width (same-colour columns)
  static constexpr int ch = px_per_block / cw;        // colour tile height
  static_assert(px_per_block % CW == 0, "colour tile width must divide the pixels of a block");
};

#ifndef HPM_XCD_STRIP
#define HPM_XCD_STRIP 2
#endif
// Workgroup -> tile, XCD-aware.  The dispatcher deals workgroups round-robin to
// the 8 XCDs (XCD = id % 8), each with its own 4 MB L2.  Every XCD gets one
// contiguous range of tiles, walked column-major inside horizontal strips of
// `strip` tiles, so the workgroups an XCD runs at the same time cover a
// compact area and the source texels they gather stay in its L2.  `id` is the
// workgroup's index among the nx * ny of its view; a multi-view launch starts
// every view at a multiple of 8 so that id % 8 is still the XCD.
__device__ __forceinline__ void xcd_tile(int id, int nx, int ny, int strip, int& tx, int& ty) {
  const int N = nx * ny;
  const int xcd = id & 7, k = id >> 3;
  const int base = N >> 3, rem = N & 7;
  const int t = xcd * base + (xcd < rem ? xcd : rem) + k;  // XCD x owns base + (x < rem) tiles
  const int S = strip;
  const int band = t / (nx * S);
  const int h = (ny - band * S) < S ? (ny - band * S) : S;
  const int r = t - band * nx * S;
  tx = r / h;
  ty = band * S + r % h;
}

template <class E> __host__ __device__ constexpr int tiles_2d(int W, int H) {
  return ((W + Tiles<E>::sq - 1) / Tiles<E>::sq) * ((H + Tiles<E>::sq - 1) / Tiles<E>::sq);
}
template <class E, int CW> __host__ __device__ constexpr int tiles_colour(int W, int H2) {
  return (((W + 1) / 2 + Tiles<E, CW>::cw - 1) / Tiles<E, CW>::cw) * ((H2 + Tiles<E, CW>::ch - 1) / Tiles<E, CW>::ch);
}
template <class E> __device__ __forceinline__ void tile_pixel(int id, int strip, int& px, int& py, int W, int H) {
  const int pix = (int)threadIdx.x / E::kLanes;
  int tx, ty;
  xcd_tile(id, (W + Tiles<E>::sq - 1) / Tiles<E>::sq, (H + Tiles<E>::sq - 1) / Tiles<E>::sq, strip, tx, ty);
  px = tx * Tiles<E>::sq + pix % Tiles<E>::sq;
  py = ty * Tiles<E>::sq + pix / Tiles<E>::sq;
}
template <class E, int CW, int NT> __device__ __forceinline__ void colour_pixel(int id, int strip, int& cx, int& py, int W, int H2) {
  using T = Tiles<E, CW, NT>;
  const int pix = (int)threadIdx.x / E::kLanes;
  int tx, ty;
  xcd_tile(id, ((W + 1) / 2 + T::cw - 1) / T::cw, (H2 + T::ch - 1) / T::ch, strip, tx, ty);
  cx = tx * T::cw + pix % T::cw;
  py = ty * T::ch + pix / T::cw;
}
// counts once per pixel: only the first lane of a multi-lane pixel contributes
template <class E> __device__ __forceinline__ unsigned first_lane(unsigned x) {
  return (E::kLanes == 1 || (threadIdx.x % E::kLanes) == 0) ? x : 0u;
}

__device__ __forceinline__ void wave_count_add(unsigned long long* counter, unsigned long long x) {
  // one atomic per wave
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0) atomicAdd(counter, x);
}

// ComputeMultiViewInitialCostandSelectedViews, HPM.cu:450-488
template <class E>
__device__ __forceinline__ float init_cost(const RunConsts& C, int px, int py, float4 pl, const typename E::Win& rw, uint32_t& sel, unsigned& nncc) {
  const int V = C.V;
  float cv[kMaxSrc], cvc[kMaxSrc];
  int valid = 0;
  for (int i = 0; i < V; ++i) {
    const float x = E::ncc(C, i + 1, px, py, pl, rw);
    cv[i] = x;
    cvc[i] = x;
    if (x < 2.0f) ++valid;
  }
  nncc += V;
  for (int i = 1; i < V; i++) {  // insertion sort, HPM.cu:15-24
    const float tmp = cv[i];
    int j;
    for (j = i; j >= 1 && tmp < cv[j - 1]; j--) cv[j] = cv[j - 1];
    cv[j] = tmp;
  }
  sel = 0u;
  const int top_k = valid < C.top_k ? valid : C.top_k;
  if (top_k > 0) {
    float cost = 0.0f;
    for (int i = 0; i < top_k; ++i) cost += cv[i];
    const float thr = cv[top_k - 1];
    for (int i = 0; i < V; ++i)
      if (cvc[i] <= thr) sel |= (1u << i);
    return cost / (float)top_k;
  }
  return 2.0f;
}

template <class E>
__device__ __forceinline__ void reload_plane(const RunConsts& C, const DevState& S, int px, int py, const typename E::Win& rw, unsigned& nncc) {  // HPM.cu:749-760
  const int center = py * C.W + px;
  float4 pl = C.hierarchy ? S.scaled[center] : S.planes[center];
  pl = to_ref(C.cams[0], pl);
  const float depth = pl.w;
  pl.w = dist_to_origin(C.cams[0], px, py, depth, pl);
  S.planes[center] = pl;
  uint32_t sel;
  S.costs[center] = init_cost<E>(C, px, py, pl, rw, sel, nncc);
  S.sel[center] = sel;
}

// JBU of the coarse normal/cost map inside RandomInitialization, HPM.cu:633-698.
template <class E>
__device__ __forceinline__ void upsample_init(const RunConsts& C, const DevState& S, int px, int py, const typename E::Win& rw, unsigned& nncc) {
  const int W = C.W, H = C.H;
  const int center = py * W + px;
  const float sc_cols = C.scaled_cols, sc_rows = C.scaled_rows;
  const float scale = (float)(1.0 * sc_cols / W);
  const float sigmad = 0.50f, sigmar = 25.5f;
  const float fa = (float)W / sc_cols, fb = (float)H / sc_rows;
  const int Imagescale = (int)(fa > fb ? fa : fb);
  const int WinWidth = Imagescale * Imagescale + 1;
  const int nn = WinWidth / 2;
  const float o_y = (float)py * scale, o_x = (float)px * scale;
  const float* ref = C.img[0];
  const float refPix = tex_at(ref, W, H, px, py);
  float nf = 0.0f, ctot = 0.0f;
  float4 nt = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int j = -nn; j <= nn; ++j) {
    int r_y = (int)(o_y + (float)j);
    r_y = (r_y > 0 ? ((float)r_y < sc_rows ? r_y : (int)(sc_rows - 1)) : 0);
    const int r_ys = py + j;
    for (int i = -nn; i <= nn; ++i) {
      int r_x = (int)(o_x + (float)i);
      r_x = (r_x > 0 ? ((float)r_x < sc_cols ? r_x : (int)(sc_cols - 1)) : 0);
      const int s_center = (int)((float)r_y * sc_cols + (float)r_x);
      const float4 sn = S.scaled[s_center];
      const float nb = tex_at(ref, W, H, px + i, r_ys);
      const float dxs = o_x - (float)r_x, dys = o_y - (float)r_y;
      const float dis = dxs * dxs + dys * dys - 0.0f;
      const float sg = (float)exp_hpm(-1.0 * (double)dis / (double)(2 * sigmad * sigmad));
      const float xr = fabsf(refPix - nb);
      const float rg = (float)exp_hpm(-1.0 * (double)(xr * xr) / (double)(2 * sigmar * sigmar));
      const float tg = sg * rg;
      nf += tg;
      ctot += sn.w * tg;
      nt.x = nt.x + sn.x * tg;
      nt.y = nt.y + sn.y * tg;
      nt.z = nt.z + sn.z * tg;
    }
  }
  S.costs[center] = ctot / nf;
  nt.x = nt.x / nf; nt.y = nt.y / nf; nt.z = nt.z / nf;
  normalize3(nt);
  uint32_t sel;
  float c0 = init_cost<E>(C, px, py, S.planes[center], rw, sel, nncc);
  S.pre_costs[center] = c0;
  float4 pl = to_ref(C.cams[0], nt);
  const float depth = S.planes[center].w;
  pl.w = dist_to_origin(C.cams[0], px, py, depth, pl);
  S.planes[center] = pl;
  S.costs[center] = init_cost<E>(C, px, py, pl, rw, sel, nncc);
  S.sel[center] = sel;
}

template <class E>
__device__ __forceinline__ void init_body(const RunConsts& C, const DevState& S, int id, int strip) {
  int px, py;
  tile_pixel<E>(id, strip, px, py, C.W, C.H);
  unsigned nncc = 0;
  if (px < C.W && py < C.H) {
    const int center = py * C.W + px;
    Philox rng(C.seed, (uint32_t)center, 0u);
    const DevCam& cam = C.cams[0];
    if (!C.prior && !C.mand) S.texture[center] = texture_factor(C, S.canny, px, py);
    bool need_rw = !(C.prior && !(S.masks[center] > 0u && !(S.conf[center] > 0.3f)));
    typename E::Win rw;
    if (need_rw) E::build(C, px, py, rw);
    if (!C.geom && !C.hierarchy && !C.prior) {
      // GenerateRandomPlaneHypothesis, HPM.cu:224-230
      const float depth = rng.uniform() * (C.depth_max - C.depth_min) + C.depth_min;
      float4 pl = random_normal(cam, px, py, rng, depth);
      pl.w = dist_to_origin(cam, px, py, depth, pl);
      S.planes[center] = pl;
      uint32_t sel;
      S.costs[center] = init_cost<E>(C, px, py, pl, rw, sel, nncc);
      S.sel[center] = sel;
    } else if (C.prior) {
      if (S.masks[center] > 0u) {
        if (S.conf[center] > 0.3f) {
          float4 pl = S.planes[center];
          const float depth = pl.w;
          pl.w = dist_to_origin(cam, px, py, depth, pl);
          S.planes[center] = pl;
          S.costs[center] = 0.1f;
        } else if (S.costs[center] > 0.1f) {
          const float perturbation = 0.02f;
          const float4 pl = S.prior[center];
          float dp = pl.w;
          const float dminp = (1 - 3 * perturbation) * dp;
          const float dmaxp = (1 + 3 * perturbation) * dp;
          dp = rng.uniform() * (dmaxp - dminp) + dminp;
          float4 pp = perturbed_normal(cam, px, py, pl, rng, C.pert_prior_init);
          pp.w = dp;  // the depth goes into w here (HPM.cu:609)
          S.planes[center] = pp;
          uint32_t sel;
          S.costs[center] = init_cost<E>(C, px, py, pp, rw, sel, nncc);
          S.sel[center] = sel;
        } else {
          float4 pl = S.planes[center];
          const float depth = pl.w;
          pl.w = dist_to_origin(cam, px, py, depth, pl);
          S.planes[center] = pl;
          uint32_t sel;
          S.costs[center] = init_cost<E>(C, px, py, pl, rw, sel, nncc);
          S.sel[center] = sel;
        }
      } else {
        float4 pl = S.planes[center];
        const float depth = pl.w;
        pl.w = dist_to_origin(cam, px, py, depth, pl);
        S.planes[center] = pl;
        if (S.conf[center] > 0.3f) S.costs[center] = 0.1f;
      }
    } else {
      if (C.upsample) {
        upsample_init<E>(C, S, px, py, rw, nncc);
      } else if (C.mand) {
        if (!(S.conf[center] >= 0.3f)) {
          const float tex = texture_factor(C, S.canny, px, py);
          if (tex > 0.5f) S.conf[center] = 0.1f;
        }
        reload_plane<E>(C, S, px, py, rw, nncc);
      } else {
        reload_plane<E>(C, S, px, py, rw, nncc);
      }
    }
  }
  wave_count_add(&S.counters[0], first_lane<E>(nncc));
}

template <class E>
__global__ HPM_PROP_ATTR void k_init(const RunConsts* __restrict__ Cp, DevState S) {
  init_body<E>(*Cp, S, (int)blockIdx.x, HPM_XCD_STRIP);
}

// ------------------------------------------------------------------ propagation
// Neighbour search (HPM.cu:1385-1608) and left-up extension (HPM.cu:1609-1680,
// 1063-1126, 1322-1345).  Only direction 0 extends: the reference's
// `bool symbol_eight_orientations[10] = { true }` enables index 0 only.
template <class E>
__device__ __forceinline__ void sample_neighbours(const RunConsts& C, const DevState& S, int px, int py, int iter, const typename E::Win& rw,
                                  Sample& s, unsigned& nncc, unsigned& n_ext) {
  const int W = C.W, H = C.H, V = C.V;
  const int center = py * W + px;
  const float* __restrict__ costs = S.costs;
  for (int d = 0; d < 8; ++d) s.flag[d] = false;
  for (int k = 0; k < 8 * V; ++k) s.ca[s.cs * (k)] = 0.0f;
  s.ca[s.cs * (0)] = 2.0f;  // `float cost_array[8][32] = { 2.0f }`
  int pos[8] = {center - 5 - 6 * W, center - W * 5, center + 6 - 5 * W, center + W * 5,
                center + 5 + 6 * W, center - 5,     center - 6 + 5 * W, center + 5};
  // axis directions: 1 up, 3 down, 5 left, 7 right (7 keeps the maximum)
  if (py > 4) {
    s.flag[1] = true;
    float m = costs[pos[1]]; int mp = pos[1];
    for (int i = 1; i < 5; ++i) {
      if (py > 4 + 2 * i) { const int pt = pos[1] - 2 * i * W; const float c = costs[pt]; if (c < m) { m = c; mp = pt; } } else break;
    }
    pos[1] = mp;
  }
  if (py < H - 5) {
    s.flag[3] = true;
    float m = costs[pos[3]]; int mp = pos[3];
    for (int i = 1; i < 5; ++i) {
      if (py < H - 5 - 2 * i) { const int pt = pos[3] + 2 * i * W; const float c = costs[pt]; if (c < m) { m = c; mp = pt; } } else break;
    }
    pos[3] = mp;
  }
  if (px > 4) {
    s.flag[5] = true;
    float m = costs[pos[5]]; int mp = pos[5];
    for (int i = 1; i < 5; ++i) {
      if (px > 4 + 2 * i) { const int pt = pos[5] - 2 * i; const float c = costs[pt]; if (c < m) { m = c; mp = pt; } } else break;
    }
    pos[5] = mp;
  }
  if (px < W - 5) {
    s.flag[7] = true;
    float m = costs[pos[7]]; int mp = pos[7];
    for (int i = 1; i < 5; ++i) {
      if (px < W - 5 - 2 * i) { const int pt = pos[7] + 2 * i; const float c = costs[pt]; if (m < c) { m = c; mp = pt; } } else break;
    }
    pos[7] = mp;
  }
  // diagonal zig-zags (HPM.cu:1480-1608)
  const bool dok[4] = {px > 4 && py > 5, px < W - 6 && py > 4, px < W - 5 && py < H - 6, px > 5 && py < H - 5};
  const int sx[4] = {px - 5, px + 6, px + 5, px - 6};
  const int sy[4] = {py - 6, py - 5, py + 6, py + 5};
  const bool even_x[4] = {true, false, true, false};
  const int ddx[4] = {-2, 2, 2, -2};
  const int ddy[4] = {-2, -2, 2, 2};
  for (int k = 0; k < 4; ++k) {
    if (!dok[k]) continue;
    const int d = 2 * k;
    s.flag[d] = true;
    float m = costs[pos[d]]; int mp = pos[d];
    int x = sx[k], y = sy[k];
    for (int i = 0; i < 7; i++) {
      if (((i & 1) == 0) == even_x[k]) x += ddx[k]; else y += ddy[k];
      if (x >= 0 && x < W && y < H && y >= 0) {
        const int pt = y * W + x;
        const float c = costs[pt];
        if (c < m) { m = c; mp = pt; }
      } else break;
    }
    pos[d] = mp;
  }
  // cost rows of the valid directions
  for (int d = 0; d < 8; ++d) {
    if (!s.flag[d]) continue;
    const float4 pl = S.planes[pos[d]];
    for (int v = 0; v < V; ++v) s.ca[s.cs * (d * V + v)] = E::ncc(C, v + 1, px, py, pl, rw);
    nncc += V;
  }
  // extension of direction 0
  for (int it = 0; it < 3; ++it) {
    if (!s.flag[0]) break;
    const float good = C.judge_thr[iter][it];
    int gs = 0, bs = 0;
    for (int v = 0; v < V; v++) {
      const float x = s.ca[s.cs * (v)];
      if (x < good) gs++;
      if (x > 1.2f) bs++;
    }
    if (gs >= 1 && bs <= 2) break;
    const int lu = center - 5 - 6 * W - 8 * it - 8 * W * it;
    if (px > 4 + 8 * it && py > 5 + 8 * it) {
      float m = costs[lu]; int mp = lu;
      int x = px - 5 - 8 * it, y = py - 6 - 8 * it;
      for (int i = 0; i < 7; i++) {
        if ((i & 1) == 0) x -= 2; else y -= 2;
        if (x >= 0 && x < W && y < H && y >= 0) {
          const int pt = y * W + x;
          const float c = costs[pt];
          if (c < m) { m = c; mp = pt; }
        } else break;
      }
      if (m > costs[pos[0]]) {
        // keeps the previous sample and its costs
      } else {
        pos[0] = mp;
        const float4 pl = S.planes[mp];
        for (int v = 0; v < V; ++v) s.ca[s.cs * (v)] = E::ncc(C, v + 1, px, py, pl, rw);
        nncc += V;
        n_ext += 1;
      }
    }
  }
  for (int d = 0; d < 8; ++d) s.pos[d] = pos[d];
}

// View weights of the joint view selection: counts 0..15 (15 draws) packed
// four bits per view, views 0-15 in `lo`, 16-31 in `hi`, so that no
// per-thread array (and no scratch memory) is needed.
struct ViewWeights {
  uint64_t lo = 0, hi = 0;
  __device__ __forceinline__ int get(int v) const { return (int)(((v < 16 ? lo : hi) >> (4 * (v & 15))) & 15ull); }
  __device__ __forceinline__ bool on(int v) const { return get(v) > 0; }
  __device__ __forceinline__ float w(int v) const { return (float)get(v); }
  __device__ __forceinline__ void inc(int v) {
    if (v < 16) lo += 1ull << (4 * v); else hi += 1ull << (4 * (v - 16));
  }
};

// Sampling weight of view i before normalisation: the view's probability from
// the 8 direction costs times its prior from the neighbours' masks
// (HPM.cu:1689-1727).
__device__ __forceinline__ float view_prob(const RunConsts& C, const Sample& s, int i, int V, float thr, float fb,
                                           const uint32_t* nm, int nflags) {
  float prior = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if ((nflags >> k) & 1) prior += ((nm[k] >> i) & 1u) ? 0.9f : 0.1f;
  float count = 0.0f, tmpw = 0.0f;
  int count_false = 0;
  for (int j = 0; j < 8; j++) {
    const float x = s.ca[s.cs * (j * V + i)];
    if (x < thr) { tmpw += expf_hpm(x * x / (-0.18f)); count++; }
    if (x > 1.2f) count_false++;
  }
  float p = 0.0f;
  if (count > 2 && count_false < 3) p = tmpw / count;
  else if (count_false < 3) p = fb;
  return p * prior;
}

// Joint view selection, HPM.cu:1686-1751.  The 15 draws are taken first (the
// scans between them draw nothing) and resolved in one sweep over the CDF:
// draw k goes to the first view whose cumulative probability exceeds it.
__device__ __forceinline__ void view_selection(const RunConsts& C, const DevState& S, int px, int py, int iter, const Sample& s, Philox& rng,
                               ViewWeights& vw, uint32_t& tsel, float& wn, int& nsel) {
  const int W = C.W, V = C.V;
  const int center = py * W + px;
  const int nbr[4] = {center - W * 5, center + W * 5, center - 5, center + 5};
  uint32_t nm[4];
  int nflags = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    nm[i] = 0u;
    if (s.flag[2 * i + 1]) { nm[i] = S.sel[nbr[i]]; nflags |= 1 << i; }
  }
  const float thr = C.vs_thr[iter], fb = C.vs_fb[iter];
  float sum = 0.0f;
  for (int i = 0; i < V; ++i) sum += view_prob(C, s, i, V, thr, fb, nm, nflags);
  const float inv = 1.0f / sum;
  float r[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) r[k] = rng.uniform() - 1.19209290e-07f;
  vw = ViewWeights();
  uint32_t assigned = 0u;
  float cum = 0.0f;
  for (int v = 0; v < V; ++v) {
    cum += view_prob(C, s, v, V, thr, fb, nm, nflags) * inv;
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      if (!((assigned >> k) & 1u) && cum > r[k]) { assigned |= 1u << k; vw.inc(v); }
    }
  }
  tsel = 0u; wn = 0.0f; nsel = 0;
  for (int i = 0; i < V; ++i)
    if (vw.on(i)) { tsel |= (1u << i); wn += vw.w(i); nsel++; }
}

__device__ __forceinline__ void final_costs(const RunConsts& C, const DevState& S, int px, int py, const Sample& s,
                                            const ViewWeights& vw, float wn, float* fc) {  // HPM.cu:1753-1771
  const int V = C.V;
  for (int i = 0; i < 8; ++i) {
    float f = 0.0f;
    for (int j = 0; j < V; ++j) {
      if (vw.on(j)) {
        const float c = s.ca[s.cs * (i * V + j)];
        if (C.geom) {
          if (s.flag[i]) f += vw.w(j) * (c + 0.2f * geom_cost(C, j + 1, S.planes[s.pos[i]], px, py));
          else f += vw.w(j) * (c + 0.1f * 3.0f);
        } else {
          f += vw.w(j) * c;
        }
      }
    }
    fc[i] = f / wn;
  }
}

template <class E>
__device__ __forceinline__ float current_cost(const RunConsts& C, int px, int py, float4 pl, const ViewWeights& vw, float wn,
                                              const typename E::Win& rw, unsigned& nncc) {  // HPM.cu:1775-1786
  const int V = C.V;
  float cn = 0.0f;
  for (int i = 0; i < V; ++i) {
    if (!vw.on(i)) continue;
    const float ci = E::ncc(C, i + 1, px, py, pl, rw);
    ++nncc;
    if (C.geom) cn += vw.w(i) * (ci + 0.2f * geom_cost(C, i + 1, pl, px, py));
    else cn += vw.w(i) * ci;
  }
  return cn / wn;
}

__device__ __forceinline__ int find_min_idx(const float* x) {  // HPM.cu:41-52
  float m = x[0]; int idx = 0;
  for (int i = 1; i < 8; ++i) if (x[i] <= m) { m = x[i]; idx = i; }
  return idx;
}
__device__ __forceinline__ int find_max_idx(const float* x) {  // HPM.cu:54-65
  float m = x[0]; int idx = 0;
  for (int i = 1; i < 8; ++i) if (x[i] >= m) { m = x[i]; idx = i; }
  return idx;
}

// PlaneHypothesisRefinement, HPM.cu:765-841
template <class E>
__device__ __forceinline__ void refine(const RunConsts& C, const DevState& S, float4& plane, float& depth, float& cost, Philox& rng,
                       const ViewWeights& vw, float wn, float& rcost, int px, int py, float texture, const typename E::Win& rw,
                       unsigned& nncc) {
  const DevCam& cam = C.cams[0];
  const int center = py * C.W + px;
  const float perturbation = 0.02f;
  float depth_prior = 0.0f, depth_rand;
  float4 plane_rand;
  const bool use_prior = C.prior && S.masks[center] > 0u;
  if (use_prior) {
    const float4 pp = S.prior[center];
    depth_prior = depth_from_plane(cam, pp, px, py);
    depth_rand = rng.uniform() * 6 * C.depth_sigma + (depth_prior - 3 * C.depth_sigma);
    plane_rand = perturbed_normal(cam, px, py, pp, rng, C.angle_sigma);
  } else {
    depth_rand = rng.uniform() * (C.depth_max - C.depth_min) + C.depth_min;
    plane_rand = random_normal(cam, px, py, rng, depth);
  }
  float depth_pert = depth;
  const float dminp = (1 - perturbation) * depth_pert;
  const float dmaxp = (1 + perturbation) * depth_pert;
  depth_pert = rng.uniform() * (dmaxp - dminp) + dminp;
  const float4 plane_pert = perturbed_normal(cam, px, py, plane, rng, C.pert_refine);
  // candidates (depth, normal): (rand, cur) (cur, rand) (rand, rand) (cur, pert) (pert, cur), HPM.cu:797-799
  const float depth0 = depth;
  const float4 plane0 = plane;
  const int V = C.V;
  for (int i = 0; i < 5; ++i) {
    const float di = (i == 0 || i == 2) ? depth_rand : (i == 4 ? depth_pert : depth0);
    float4 tp = (i == 0 || i == 4) ? plane0 : (i == 3 ? plane_pert : plane_rand);
    tp.w = dist_to_origin(cam, px, py, di, tp);
    float tc = 0.0f;
    for (int j = 0; j < V; ++j) {
      if (!vw.on(j)) continue;
      const float cj = E::ncc(C, j + 1, px, py, tp, rw);
      ++nncc;
      if (C.geom) tc += vw.w(j) * (cj + 0.2f * geom_cost(C, j + 1, tp, px, py));
      else tc += vw.w(j) * cj;
    }
    tc /= wn;
    const float db = depth_from_plane(cam, tp, px, py);
    if (use_prior) {
      const float dd = di - depth_prior;
      const float nd = norm_diff(S.prior[center], tp);
      const float pr = prior_term(dd, nd, texture, C.two_dss, C.two_ass);
      const float rtc = restricted_cost(tc, 0.18f, texture, pr);
      if (db >= C.depth_min && db <= C.depth_max && rtc > rcost) { depth = db; plane = tp; cost = tc; rcost = rtc; }
    } else {
      if (db >= C.depth_min && db <= C.depth_max && tc < cost) { depth = db; plane = tp; cost = tc; }
    }
  }
}

// CheckerboardPropagation, HPM.cu:1347-1873
template <class E>
__device__ __forceinline__ void propagate_pixel(const RunConsts& C, const DevState& S, int px, int py, int iter, const typename E::Win& rw,
                                Sample& s, unsigned& nncc, unsigned& n_ext) {
  const DevCam& cam = C.cams[0];
  const int center = py * C.W + px;
  Philox rng(C.seed, (uint32_t)center, (uint32_t)(1 + iter));
  sample_neighbours<E>(C, S, px, py, iter, rw, s, nncc, n_ext);
  ViewWeights vw; uint32_t tsel; float wn; int nsel;
  view_selection(C, S, px, py, iter, s, rng, vw, tsel, wn, nsel);
  float fc[8];
  final_costs(C, S, px, py, s, vw, wn, fc);
  const int min_idx = find_min_idx(fc);
  float cost_now = current_cost<E>(C, px, py, S.planes[center], vw, wn, rw, nncc);
  S.costs[center] = cost_now;
  float depth_now = depth_from_plane(cam, S.planes[center], px, py);
  float rcost = 0.0f, texture = 0.0f;
  if (C.prior) {
    const float4 pc = S.prior[center];
    const float depth_prior = depth_from_plane(cam, pc, px, py);
    texture = texture_factor(C, S.canny, px, py);
    if (S.masks[center] > 0u) {
      float rfc[8];
      for (int i = 0; i < 8; i++) {
        rfc[i] = 0.0f;
        if (s.flag[i]) {
          const float4 q = S.planes[s.pos[i]];
          const float dn = depth_from_plane(cam, q, px, py);
          const float pr = prior_term(dn - depth_prior, norm_diff(pc, q), texture, C.two_dss, C.two_ass);
          rfc[i] = restricted_cost(fc[i], 0.18f, texture, pr);
        }
      }
      const int max_idx = find_max_idx(rfc);
      const float4 cur = S.planes[center];
      const float dn = depth_from_plane(cam, cur, px, py);
      const float pr = prior_term(dn - depth_prior, norm_diff(pc, cur), texture, C.two_dss, C.two_ass);
      const float rc_now = restricted_cost(cost_now, 0.18f, texture, pr);
      if (s.flag[max_idx]) {
        const float4 q = S.planes[s.pos[max_idx]];
        const float db = depth_from_plane(cam, q, px, py);
        if (db >= C.depth_min && db <= C.depth_max && rfc[max_idx] > rc_now) {
          depth_now = db;
          S.planes[center] = q;
          S.costs[center] = fc[max_idx];
          rcost = rfc[max_idx];
          S.sel[center] = tsel;
        }
      }
    } else if (s.flag[min_idx]) {
      const float4 q = S.planes[s.pos[min_idx]];
      const float db = depth_from_plane(cam, q, px, py);
      if (db >= C.depth_min && db <= C.depth_max && fc[min_idx] < cost_now) {
        depth_now = db;
        S.planes[center] = q;
        S.costs[center] = fc[min_idx];
      }
    }
  }
  float4 plane_now = S.planes[center];  // the reference leaves this uninitialised
  if (!C.prior && s.flag[min_idx]) {
    const float4 q = S.planes[s.pos[min_idx]];
    const float db = depth_from_plane(cam, q, px, py);
    if (db >= C.depth_min && db <= C.depth_max && fc[min_idx] < cost_now) {
      depth_now = db; plane_now = q; cost_now = fc[min_idx]; S.sel[center] = tsel;
    }
    const float cc = S.costs[center];
    if (cc != cc) {
      depth_now = db; plane_now = q; cost_now = fc[min_idx]; S.sel[center] = tsel;
    }
  }
  refine<E>(C, S, plane_now, depth_now, cost_now, rng, vw, wn, rcost, px, py, texture, rw, nncc);
  if (C.hierarchy) {
    if (cost_now < S.pre_costs[center] - 0.1f) {
      S.costs[center] = cost_now;
      S.planes[center] = plane_now;
    }
  } else {
    S.costs[center] = cost_now;
    S.planes[center] = plane_now;
  }
}

// CheckerboardPropagation_MandatoryConsistency, HPM.cu:1875-2791
template <class E>
__device__ __forceinline__ void propagate_pixel_mand(const RunConsts& C, const DevState& S, int px, int py, int iter, const typename E::Win& rw,
                                     Sample& s, unsigned& nncc, unsigned& n_ext) {
  const DevCam& cam = C.cams[0];
  const int center = py * C.W + px;
  Philox rng(C.seed, (uint32_t)center, (uint32_t)(1 + iter));
  sample_neighbours<E>(C, S, px, py, iter, rw, s, nncc, n_ext);
  ViewWeights vw; uint32_t tsel; float wn; int nsel;
  view_selection(C, S, px, py, iter, s, rng, vw, tsel, wn, nsel);
  // Both branches of the reference compute the 8 direction costs (the low-
  // confidence branch then rescales them) and the current plane's cost:
  // computed once, before the branch, so that a wave holding pixels of both
  // kinds does not evaluate these NCCs twice.  The operations are the same.
  float fc[8];
  final_costs(C, S, px, py, s, vw, wn, fc);
  float cost_now = current_cost<E>(C, px, py, S.planes[center], vw, wn, rw, nncc);
  S.costs[center] = cost_now;
  if (S.conf[center] >= 0.3f) {
    const int min_idx = find_min_idx(fc);
    float depth_now = depth_from_plane(cam, S.planes[center], px, py);
    float4 plane_now = S.planes[center];
    if (s.flag[min_idx]) {
      const float4 q = S.planes[s.pos[min_idx]];
      const float db = depth_from_plane(cam, q, px, py);
      if (db >= C.depth_min && db <= C.depth_max && fc[min_idx] < cost_now) {
        depth_now = db; plane_now = q; cost_now = fc[min_idx]; S.sel[center] = tsel;
      }
      const float cc = S.costs[center];
      if (cc != cc) { plane_now = q; cost_now = fc[min_idx]; S.sel[center] = tsel; }
    }
    float unused = 0.0f;
    refine<E>(C, S, plane_now, depth_now, cost_now, rng, vw, wn, unused, px, py, 0.0f, rw, nncc);
    S.costs[center] = cost_now;
    S.planes[center] = plane_now;
    return;
  }
  float rfc[8];
  for (int i = 0; i < 8; i++) {
    float f = fc[i];
    rfc[i] = 0.0f;
    if (s.flag[i]) {
      const float gce = expf_hpm(-f * f / 0.18f);
      const float mmc = (float)nsel * 0.5f;
      const float cp = S.conf[s.pos[i]];
      const float dlt = fmaxf(0.0f, mmc - cp);
      const float mce = (float)(0.5 + exp_hpm((double)(-dlt * dlt) / ((double)(mmc * mmc) * 1.0)));
      rfc[i] = (cp > 0.0f) ? fmaxf(3.0f, mce) * gce : mce * gce;
      if (cp > 0.0f) f = (float)((double)f * fmin((double)expf_hpm(-cp * 2.0f), 0.35));
    }
    fc[i] = f;
  }
  const int max_idx = find_max_idx(rfc);
  const float gce = expf_hpm(-cost_now * cost_now / 0.18f);
  const float mmc = (float)nsel * 0.5f;
  const float cc = S.conf[center];
  const float dlt = fmaxf(0.0f, mmc - cc);
  const float mce = (float)(0.5 + exp_hpm((double)(-dlt * dlt) / ((double)(mmc * mmc) * 1.0)));
  const float rc_now = (cc > 0.0f) ? fmaxf(3.0f, mce) * gce : mce * gce;
  if (cc < 0.3f) cost_now = (float)((double)cost_now * exp_hpm(0.2 + (double)(cc * 2.0f)));
  else cost_now = cost_now * expf_hpm(-cc * 2.0f);
  float4 plane_now = S.planes[center];
  if (s.flag[max_idx]) {
    const int q = s.pos[max_idx];
    const float4 qp = S.planes[q];
    const float db = depth_from_plane(cam, qp, px, py);
    if (db >= C.depth_min && db <= C.depth_max && rfc[max_idx] > rc_now) {
      plane_now = qp; cost_now = fc[max_idx]; S.sel[center] = tsel;
      S.conf[center] = (float)(0.5 * (double)S.conf[q]);
    }
    const float c2 = S.costs[center];
    if (c2 != c2) {
      plane_now = qp; cost_now = fc[max_idx]; S.sel[center] = tsel;
      S.conf[center] = (float)(0.5 * (double)S.conf[q]);
    }
  }
  S.costs[center] = cost_now;
  S.planes[center] = plane_now;
}

// One colour pass.  Thread (tx, y) of a 32 x 8 block handles the tx-th pixel of
// the pass's colour in row y: x = 2 tx + ((y + colour) & 1).
template <class E, bool kLds, int CW, int NT = 256>
__device__ __forceinline__ void propagate_body(const RunConsts& C, const DevState& S, int colour, int iter, int H2, int id,
                                               int strip) {
  extern __shared__ float lds_rows[];
  float private_rows[kLds ? 1 : 8 * kMaxSrc];
  int tx, py;
  colour_pixel<E, CW, NT>(id, strip, tx, py, C.W, H2);
  const int px = 2 * tx + ((py + colour) & 1);
  unsigned nncc = 0, n_ext = 0;
  bool active = px < C.W && py < H2;
  if (active && C.prior) {
    const int center = py * C.W + px;
    if (S.conf[center] > 0.3f || S.masks[center] == 0u) active = false;
  }
  if (active) {
    typename E::Win rw;
    E::build(C, px, py, rw);
    Sample s;
    s.ca = kLds ? lds_rows + threadIdx.x : private_rows;
    s.cs = kLds ? NT : 1;
    if (C.mand) propagate_pixel_mand<E>(C, S, px, py, iter, rw, s, nncc, n_ext);
    else propagate_pixel<E>(C, S, px, py, iter, rw, s, nncc, n_ext);
  }
  wave_count_add(&S.counters[1], first_lane<E>(nncc));
  // evaluations the reference performs for this pixel: (8 + E + 1 + 5) * V
  const unsigned ref_count = active ? (unsigned)((14 + n_ext) * C.V) : 0u;
  wave_count_add(&S.counters[2], first_lane<E>(ref_count));
}

template <class E, bool kLds>
__global__ HPM_PROP_ATTR void k_propagate(const RunConsts* __restrict__ Cp, DevState S, int colour, int iter, int H2) {
  propagate_body<E, kLds, Tiles<E>::cw>(*Cp, S, colour, iter, H2, (int)blockIdx.x, HPM_XCD_STRIP);
}

// GetDepthandNormal, HPM.cu:2829-2842
__device__ __forceinline__ void output_pixel(const RunConsts& C, float4* planes, int px, int py) {
  const int center = py * C.W + px;
  float4 pl = planes[center];
  pl.w = depth_from_plane(C.cams[0], pl, px, py);
  planes[center] = to_world(C.cams[0], pl);
}

static __global__ __launch_bounds__(256) void k_output(const RunConsts* __restrict__ Cp, float4* planes) {
  const RunConsts& C = *Cp;
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= C.W || py >= C.H) return;
  output_pixel(C, planes, px, py);
}

// ComputeMultiViewCostVector over all pixels at fixed planes (cost microbench).
template <class E>
__global__ __launch_bounds__(256) void k_cost_fixed(const RunConsts* __restrict__ Cp, const float4* __restrict__ planes,
                                                    float* __restrict__ out) {
  const RunConsts& C = *Cp;
  int px, py;
  tile_pixel<E>((int)blockIdx.x, HPM_XCD_STRIP, px, py, C.W, C.H);
  if (px >= C.W || py >= C.H) return;
  const int p = py * C.W + px;
  typename E::Win rw;
  E::build(C, px, py, rw);
  const float4 pl = planes[p];
  for (int v = 0; v < C.V; ++v) {
    const float c = E::ncc(C, v + 1, px, py, pl, rw);
    if (first_lane<E>(1u)) out[(size_t)p * C.V + v] = c;
  }
}

template <class E>
__global__ __launch_bounds__(256) void k_init_cost(const RunConsts* __restrict__ Cp, const float4* __restrict__ planes,
                                                   float* __restrict__ out, uint32_t* __restrict__ sel) {
  const RunConsts& C = *Cp;
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= C.W || py >= C.H) return;
  const int p = py * C.W + px;
  typename E::Win rw;
  E::build(C, px, py, rw);
  unsigned n = 0;
  uint32_t m;
  out[p] = init_cost<E>(C, px, py, planes[p], rw, m, n);
  sel[p] = m;
}

static __global__ __launch_bounds__(256) void k_geom_cost(const RunConsts* __restrict__ Cp, const float4* __restrict__ planes,
                                                   float* __restrict__ out) {
  const RunConsts& C = *Cp;
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= C.W || py >= C.H) return;
  const int p = py * C.W + px;
  const float4 pl = planes[p];
  for (int v = 0; v < C.V; ++v) out[(size_t)p * C.V + v] = geom_cost(C, v + 1, pl, px, py);
}

}  // namespace hpm
