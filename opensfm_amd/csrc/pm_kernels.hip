// HIP kernels of the PatchMatch depthmap engine for gfx950 (MI355X).
//
// Mapping: one lane per reference pixel, 64-lane wavefronts over 2D pixel
// blocks (16 x 4 pixels per wave, 256-thread workgroups) so that the 49
// bilinear gathers of a wave land in a compact region of the neighbour view
// and hit the same L1/L2 lines.  The hypothesis-invariant reference-side
// patch data (bilateral weights, w*x, the x moments) lives in registers for
// the whole kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "pm_device.h"
#include "pm_kernels.h"

namespace pm {

// Adds each lane's count to *c with one atomic per wavefront.  Every lane of
// the wave must reach this call.
__device__ __forceinline__ void wave_count(unsigned long long* c, unsigned v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(c, (unsigned long long)v);
}

// Tile order.  The MI355X deals the workgroups of a launch round-robin to its
// 8 XCDs (each with its own L2), so consecutive row-major tiles land on
// different XCDs.  A static renumbering that gives each XCD one contiguous
// band was measured 4 % slower on the 1600x1200 workload (and 9 % slower for
// its 64-plane sweep: 47.3 against 43.3 ms): the per-pixel
// work varies across the image (masked pixels, skipped duplicate
// candidates), so fixed bands leave XCDs unevenly loaded.  The one-shot
// kernels (init, sweep) therefore use plain row-major order on a 1-D grid;
// the red/black kernel, which dominates, bands dynamically with band
// switching at the end (k_pm_half).
struct Tile {
  int x, y;
  bool ok;
};

__device__ __forceinline__ Tile row_tile(int tiles_x, int tiles_y) {
  const int t = blockIdx.x;
  Tile r;
  r.ok = t < tiles_x * tiles_y;
  r.x = t % tiles_x;
  r.y = t / tiles_x;
  return r;
}

static inline dim3 tile_grid(int tiles_x, int tiles_y) {
  return dim3(tiles_x * tiles_y);
}

// Index of the job whose tiles (first index at offset `field` of PmJob or
// ClJob) contain tile t; t0 of job 0 is 0.
template <typename J>
__device__ __forceinline__ int find_job(const J* __restrict__ jobs, int njobs,
                                        int t, int J::*first) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].*first <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}


// ---------------------------------------------------------------------------
// Zero-bordered quad texture from a u8 image (see pm_device.h).

// With it, the same quads as FP16 bilinear coefficients (pm_device.h,
// CoefFetch).
__global__ void k_build_quad(const uint8_t* __restrict__ img,
                             uint32_t* __restrict__ quad,
                             uint2* __restrict__ quadc, int W, int H) {
  const int tx = blockIdx.x * blockDim.x + threadIdx.x;
  const int ty = blockIdx.y * blockDim.y + threadIdx.y;
  if (tx > W || ty > H) return;
  const int ix = tx - 1, iy = ty - 1;
  uint32_t v = 0;
  int a = 0, b = 0, c = 0, d = 0;
  if (ix >= 0 && iy >= 0 && ix <= W - 2 && iy <= H - 2) {
    const uint8_t* r0 = img + (size_t)iy * W + ix;
    const uint8_t* r1 = r0 + W;
    a = r0[0];
    b = r0[1];
    c = r1[0];
    d = r1[1];
    v = (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16) | ((uint32_t)d << 24);
  }
  const size_t o = (size_t)ty * (W + 1) + tx;
  quad[o] = v;
  pm_half2 ab, cd;
  ab[0] = (_Float16)a;
  ab[1] = (_Float16)(b - a);
  cd[0] = (_Float16)(c - a);
  cd[1] = (_Float16)(d - c - b + a);
  quadc[o] = make_uint2(__builtin_bit_cast(uint32_t, ab), __builtin_bit_cast(uint32_t, cd));
}

// The same quads as 16-bit values: uint2 {(a, b), (c, d)} with a, b the top
// row and c, d the bottom row of the 2x2 neighbourhood, FP16 or BF16 (u8
// values are exact in both).
template <bool kBF16>
__global__ void k_build_quad16(const uint8_t* __restrict__ img,
                               uint2* __restrict__ quad, int W, int H) {
  const int tx = blockIdx.x * blockDim.x + threadIdx.x;
  const int ty = blockIdx.y * blockDim.y + threadIdx.y;
  if (tx > W || ty > H) return;
  const int ix = tx - 1, iy = ty - 1;
  float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (ix >= 0 && iy >= 0 && ix <= W - 2 && iy <= H - 2) {
    const uint8_t* r0 = img + (size_t)iy * W + ix;
    const uint8_t* r1 = r0 + W;
    v[0] = r0[0];
    v[1] = r0[1];
    v[2] = r1[0];
    v[3] = r1[1];
  }
  uint2 q;
  if constexpr (kBF16) {
    pm_bf162 top, bot;
    top[0] = (__bf16)v[0];
    top[1] = (__bf16)v[1];
    bot[0] = (__bf16)v[2];
    bot[1] = (__bf16)v[3];
    q.x = __builtin_bit_cast(uint32_t, top);
    q.y = __builtin_bit_cast(uint32_t, bot);
  } else {
    pm_half2 top, bot;
    top[0] = (_Float16)v[0];
    top[1] = (_Float16)v[1];
    bot[0] = (_Float16)v[2];
    bot[1] = (_Float16)v[3];
    q.x = __builtin_bit_cast(uint32_t, top);
    q.y = __builtin_bit_cast(uint32_t, bot);
  }
  quad[(size_t)ty * (W + 1) + tx] = q;
}

// ---------------------------------------------------------------------------
// Plane sweep = ComputeBruteForce (depthmap.cc:184-200), exact FP32 order.
// Htab[(d * nviews + v) * 9] is the FP32 homography of sweep plane d into
// view v, composed on the host in FP64 exactly as the reference does;
// planes[3 d] is the plane vector.  Scores start at 0 (AssignMatrices), so a
// pixel whose best NCC is <= 0 keeps depth 0.

template <int P>
__device__ __forceinline__ unsigned sweep_pixel(const PmArgs& a,
                                                const float* __restrict__ Htab,
                                                const float* __restrict__ planes,
                                                int nplanes) {
  constexpr int hpz = (P - 1) / 2;
  const Tile t = row_tile((a.W + 15) >> 4, (a.H + 15) >> 4);
  if (!t.ok) return 0;
  const int j = t.x * 16 + (threadIdx.x & 15);
  const int i = t.y * 16 + (threadIdx.x >> 4);
  if (i < hpz || i >= a.H - hpz || j < hpz || j >= a.W - hpz) return 0;
  RefPatch<P> rp;
  if (!load_ref_patch<P>(rp, a.ref_img, a.W, i, j, a.wlut)) return 0;

  float best_s = 0.0f;
  int best_n = 0;
  int best_d = -1;
  for (int d = 0; d < nplanes; ++d) {
    float s = -1.0f;
    int n = 0;
    for (int v = 1; v < a.nviews; ++v) {
      const PmView& V = a.views[v];
      const float* H = Htab + ((size_t)d * a.nviews + v) * 9;
      float Hr[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) Hr[k] = H[k];
      const QuadTex tex = make_tex(V.quad, V.width, V.height);
      const float sc = score_exact<P>(rp, Hr, i, j, tex);
      if (sc > s) {
        s = sc;
        n = v;
      }
    }
    if (s > best_s) {
      best_s = s;
      best_n = n;
      best_d = d;
    }
  }
  const size_t idx = (size_t)i * a.W + j;
  if (best_d >= 0) {
    const float v0 = planes[3 * best_d], v1 = planes[3 * best_d + 1],
                v2 = planes[3 * best_d + 2];
    const float depth = depth_of_plane_f64(a.Kinv0, v0, v1, v2, (double)j, (double)i);
    reinterpret_cast<float4*>(a.state)[idx] = make_float4(v0, v1, v2, depth);
    a.score[idx] = best_s;
    a.nghbr[idx] = best_n;
  }
  return (unsigned)(nplanes * (a.nviews - 1));
}

template <int P>
__global__ __launch_bounds__(256) void k_sweep(PmArgs a,
                                               const float* __restrict__ Htab,
                                               const float* __restrict__ planes,
                                               int nplanes) {
  const unsigned n = sweep_pixel<P>(a, Htab, planes, nplanes);
  if (a.counter) wave_count(a.counter, n);
}

// ---------------------------------------------------------------------------
// Fixed-hypothesis scoring: ComputePlaneImageScore for a list of
// (pixel, homography) pairs against one neighbour view; the host composes
// each FP32 homography in FP64 with the reference's order.  out[k] is the
// FP32 NCC score.  The host groups hypotheses by view (one launch each).

template <int P>
__global__ __launch_bounds__(256) void k_score_hyp(PmArgs a, int view, int n,
                                                   const int* __restrict__ ij,
                                                   const float* __restrict__ Hs,
                                                   float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = ij[2 * k], j = ij[2 * k + 1];
  RefPatch<P> rp;
  load_ref_patch<P>(rp, a.ref_img, a.W, i, j, a.wlut);
  const PmView& V = a.views[view];
  float H[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) H[m] = Hs[9 * (size_t)k + m];
  const QuadTex tex = make_tex(V.quad, V.width, V.height);
  out[k] = score_exact<P>(rp, H, i, j, tex);
}

// ---------------------------------------------------------------------------
// PatchMatch (fast path).

struct PixelRay {
  float x, y, z;  // K_0^-1 (j, i, 1)
};

__device__ __forceinline__ PixelRay pixel_ray(const PmArgs& a, int i, int j) {
  const float fj = (float)j, fi = (float)i;
  PixelRay r;
  r.x = __builtin_fmaf(a.Kinv0f[0], fj, __builtin_fmaf(a.Kinv0f[1], fi, a.Kinv0f[2]));
  r.y = __builtin_fmaf(a.Kinv0f[3], fj, __builtin_fmaf(a.Kinv0f[4], fi, a.Kinv0f[5]));
  r.z = __builtin_fmaf(a.Kinv0f[6], fj, __builtin_fmaf(a.Kinv0f[7], fi, a.Kinv0f[8]));
  return r;
}

// PlaneFromDepthAndNormal (depthmap.cc:120-125) in FP32.
__device__ __forceinline__ void plane_from_depth_normal(const PixelRay& r,
                                                        float depth, float n0,
                                                        float n1, float n2,
                                                        float* v) {
  const float px = depth * r.x, py = depth * r.y, pz = depth * r.z;
  const float denom =
      -__builtin_fmaf(n0, px, __builtin_fmaf(n1, py, n2 * pz));
  const float inv = __frcp_rn(fmaxf(1e-6f, denom));
  v[0] = n0 * inv;
  v[1] = n1 * inv;
  v[2] = n2 * inv;
}

// A plane v in the reference camera frame, prepared for scoring at one
// pixel: g = K_0^-T v (first two components) and gq = v . K_0^-1 (j, i, 1).
// Its depth at the pixel is 1 / max(1e-6, -gq) (DepthOfPlaneBackprojection,
// depthmap.cc:114-118).
struct PlaneAtPixel {
  float g0, g1, gq;
};

__device__ __forceinline__ PlaneAtPixel plane_at_pixel(const PmArgs& a,
                                                       const PixelRay& r,
                                                       float v0, float v1,
                                                       float v2) {
  const float* Ki = a.Kinv0f;
  PlaneAtPixel p;
  p.g0 = __builtin_fmaf(v0, Ki[0], __builtin_fmaf(v1, Ki[3], v2 * Ki[6]));
  p.g1 = __builtin_fmaf(v0, Ki[1], __builtin_fmaf(v1, Ki[4], v2 * Ki[7]));
  p.gq = __builtin_fmaf(v0, r.x, __builtin_fmaf(v1, r.y, v2 * r.z));
  return p;
}

// Local-affine warp of pixel (j, i) into view V under the plane:
// H = M_v + b_v g^T (the FP32 form of PlaneInducedHomographyBaked), then
// the image point and Jacobian of ComputePlaneImageScore
// (depthmap.cc:448-462).  False when w == 0 or the warp is not finite
// (the scorer then returns -1, as the reference does for w == 0).
__device__ __forceinline__ bool warp_to_view(const PmView& V,
                                             const PlaneAtPixel& p, float fj,
                                             float fi, Warp* wp) {
  const float* M = V.M;
  const float* b = V.b;
  const float u = __builtin_fmaf(b[0], p.gq, __builtin_fmaf(M[0], fj, __builtin_fmaf(M[1], fi, M[2])));
  const float vv = __builtin_fmaf(b[1], p.gq, __builtin_fmaf(M[3], fj, __builtin_fmaf(M[4], fi, M[5])));
  const float w = __builtin_fmaf(b[2], p.gq, __builtin_fmaf(M[6], fj, __builtin_fmaf(M[7], fi, M[8])));
  const float H00 = __builtin_fmaf(b[0], p.g0, M[0]);
  const float H01 = __builtin_fmaf(b[0], p.g1, M[1]);
  const float H10 = __builtin_fmaf(b[1], p.g0, M[3]);
  const float H11 = __builtin_fmaf(b[1], p.g1, M[4]);
  const float H20 = __builtin_fmaf(b[2], p.g0, M[6]);
  const float H21 = __builtin_fmaf(b[2], p.g1, M[7]);
  const float rw = __frcp_rn(w);
  wp->Hx0 = u * rw;
  wp->Hy0 = vv * rw;
  // (H00 w - H20 u) / w^2 = (H00 - H20 u/w) / w
  wp->dfdx_x = __builtin_fmaf(-H20, wp->Hx0, H00) * rw;
  wp->dfdx_y = __builtin_fmaf(-H20, wp->Hy0, H10) * rw;
  wp->dfdy_x = __builtin_fmaf(-H21, wp->Hx0, H01) * rw;
  wp->dfdy_y = __builtin_fmaf(-H21, wp->Hy0, H11) * rw;
  const float s = wp->Hx0 + wp->Hy0 + wp->dfdx_x + wp->dfdx_y + wp->dfdy_x + wp->dfdy_y;
  return w != 0.0f && __builtin_isfinite(s);
}

// Fast NCC of one view under a warp, with the view's texture in the given
// sample precision (0 FP32 from FP16 coefficients, 1 FP16, 2 BF16,
// kPrecFp32Quads FP32 from u8 quads).
// kLaneView: the view may differ between lanes (global loads through a
// LaneTex instead of a wave-uniform buffer descriptor; same samples).
template <int P, int Prec, int kPipe = 1, bool kLaneView = false>
__device__ __forceinline__ float score_view_fast(const RefPatch<P>& rp,
                                                 const Warp& wp,
                                                 const PmView& V) {
  if constexpr (kLaneView) {
    static_assert(Prec == 0 || Prec == kPrecFp32Quads, "per-lane views: FP32 samples");
    if constexpr (Prec == 0) {
      CoefLaneFetch g;
      g.tex = make_lane_tex<2>(V.quadc, V.width, V.height);
      return score_fast_affine<P, kPipe>(rp, wp, g);
    } else {
      QuadLaneFetch g;
      g.tex = make_lane_tex<1>(V.quad, V.width, V.height);
      return score_fast_affine<P, kPipe>(rp, wp, g);
    }
  } else if constexpr (Prec == 0) {
    CoefFetch g;
    g.tex = make_tex16(V.quadc, V.width, V.height);
    return score_fast_affine<P, kPipe>(rp, wp, g);
  } else if constexpr (Prec == kPrecFp32Quads) {
    QuadFetch g;
    g.tex = make_tex(V.quad, V.width, V.height);
    return score_fast_affine<P, kPipe>(rp, wp, g);
  } else {
    Global16Fetch<Prec == 2> g;
    g.tex = make_tex16(V.quad16, V.width, V.height);
    return score_fast_affine<P, kPipe>(rp, wp, g);
  }
}

// Score a plane against every neighbour view from global memory: max NCC
// with the lowest view index winning ties (ComputePlaneScore,
// depthmap.cc:408-419).  Returns the score; *nb gets the view and *depth the
// plane's depth at the pixel.  With G > 1 the G adjacent lanes of a group
// score the same plane of the same pixel, lane `sub` the views 1 + sub,
// 1 + sub + G, ..., and the group's results are merged by the same rule (max,
// then the lowest view; -1 and view 0 when no view scores above -1), so every
// lane of the group returns the bits of the one-lane loop.
template <int P, int Prec, int kPipe = 1, int G = 1>
__device__ __forceinline__ float score_plane_all(const PmArgs& a,
                                                 const RefPatch<P>& rp,
                                                 const PixelRay& r, int i,
                                                 int j, float v0, float v1,
                                                 float v2, int* nb,
                                                 float* depth, int sub = 0) {
  const PlaneAtPixel p = plane_at_pixel(a, r, v0, v1, v2);
  *depth = __frcp_rn(fmaxf(1e-6f, -p.gq));
  const float fj = (float)j, fi = (float)i;
  float best = -1.0f;
  int bn = 0;
  for (int v = 1 + sub; v < a.nviews; v += G) {
    const PmView& V = a.views[v];
    Warp wp;
    float s = -1.0f;
    if (warp_to_view(V, p, fj, fi, &wp)) s = score_view_fast<P, Prec, kPipe, (G > 1)>(rp, wp, V);
    if (s > best) {
      best = s;
      bn = v;
    }
  }
  if constexpr (G > 1) {
    // equal scores are either both -1 with view 0 or both found, so the
    // lower view index decides
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
      const float ob = __shfl_xor(best, o, G);
      const int on = __shfl_xor(bn, o, G);
      if (ob > best || (ob == best && on < bn)) {
        best = ob;
        bn = on;
      }
    }
  }
  *nb = bn;
  return best;
}

// Score a plane against one view per lane (ComputePlaneImageScore for the
// sample-mode PatchMatch, depthmap.cc:388-396), where the view may differ
// between lanes.  The wave loops over the distinct views its pending lanes
// ask for, so every texture access uses a wave-uniform view.  Lanes with
// pend == false skip.  Returns the score (-1 if not scored) and the plane's
// depth at the pixel in *depth.
template <int P, int Prec>
__device__ __forceinline__ float score_plane_view(const PmArgs& a,
                                                  const RefPatch<P>& rp,
                                                  const PixelRay& r, int i,
                                                  int j, float v0, float v1,
                                                  float v2, int view,
                                                  bool pend, float* depth) {
  const PlaneAtPixel p = plane_at_pixel(a, r, v0, v1, v2);
  *depth = __frcp_rn(fmaxf(1e-6f, -p.gq));
  const float fj = (float)j, fi = (float)i;
  float s = -1.0f;
  while (true) {
    const unsigned long long m = __ballot(pend);
    if (m == 0) break;
    const int vv = __shfl(view, __ffsll((long long)m) - 1);
    if (pend && view == vv) {
      const PmView& V = a.views[vv];
      Warp wp;
      if (warp_to_view(V, p, fj, fi, &wp)) s = score_view_fast<P, Prec>(rp, wp, V);
      pend = false;
    }
  }
  return s;
}

// The PatchMatch scorer for fixed (pixel, plane) hypotheses against one
// view: the device homography M_v + b_v g^T and the fast NCC in the
// sample precision Prec.  -1 where the warp is degenerate or the patch
// untextured, like the scorer inside PatchMatch.
template <int P, int Prec>
__global__ __launch_bounds__(256) void k_score_hyp_fast(
    PmArgs a, int view, int n, const int* __restrict__ ij,
    const float* __restrict__ planes, float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = ij[2 * k], j = ij[2 * k + 1];
  RefPatch<P> rp;
  float s = -1.0f;
  if (load_ref_patch<P, true>(rp, a.ref_img, a.W, i, j, a.wlut)) {
    const PixelRay r = pixel_ray(a, i, j);
    const PlaneAtPixel p =
        plane_at_pixel(a, r, planes[3 * k], planes[3 * k + 1], planes[3 * k + 2]);
    Warp wp;
    const PmView& V = a.views[view];
    if (warp_to_view(V, p, (float)j, (float)i, &wp)) s = score_view_fast<P, Prec>(rp, wp, V);
  }
  out[k] = s;
}

__device__ __forceinline__ bool is_interior(const PmArgs& a, int i, int j) {
  return i >= a.hpz && i < a.H - a.hpz && j >= a.hpz && j < a.W - a.hpz;
}

// Random initialisation + ignore mask (RandomInitialization and
// ComputeIgnoreMask, depthmap.cc:237-281).  Masked and low-variance pixels
// keep the zero state (the reference scores them and then zeroes them).
// Channel ch of the learned initialisation map (PmArgs::init_map) blended
// bilinearly at pixel (i, j)'s position in the coarse grid.
__device__ __forceinline__ float init_map_at(const PmArgs& a, int i, int j, int ch) {
  const float inv = 1.0f / (float)a.init_scale;
  const float gx = fminf(fmaxf(((float)j + 0.5f) * inv - 0.5f, 0.0f), (float)(a.init_w - 1));
  const float gy = fminf(fmaxf(((float)i + 0.5f) * inv - 0.5f, 0.0f), (float)(a.init_h - 1));
  const int x0 = min((int)gx, a.init_w - 2 < 0 ? 0 : a.init_w - 2);
  const int y0 = min((int)gy, a.init_h - 2 < 0 ? 0 : a.init_h - 2);
  const int x1 = min(x0 + 1, a.init_w - 1), y1 = min(y0 + 1, a.init_h - 1);
  const float fx = gx - (float)x0, fy = gy - (float)y0;
  const float* q = a.init_map + ch;
  const size_t st = (size_t)a.init_stride;
  const float v00 = q[((size_t)y0 * a.init_w + x0) * st];
  const float v01 = q[((size_t)y0 * a.init_w + x1) * st];
  const float v10 = q[((size_t)y1 * a.init_w + x0) * st];
  const float v11 = q[((size_t)y1 * a.init_w + x1) * st];
  const float top = v00 + fx * (v01 - v00), bot = v10 + fx * (v11 - v10);
  return top + fy * (bot - top);
}

template <int P, bool kSample, int Prec>
__device__ __forceinline__ unsigned init_pixel(const PmArgs& a, int tile_x,
                                               int tile_y) {
  constexpr int hpz = (P - 1) / 2;
  const int j = tile_x * 16 + (threadIdx.x & 15);
  const int i = tile_y * 16 + (threadIdx.x >> 4);
  if (!is_interior(a, i, j)) return 0;
  const size_t idx = (size_t)i * a.W + j;
  if (a.ref_mask[idx] == 0) return 0;
  // PatchVariance / Variance (depthmap.cc:32-44, 271-281): two-pass, FP32
  float sum = 0;
#pragma unroll
  for (int u = -hpz; u <= hpz; ++u)
#pragma unroll
    for (int v = -hpz; v <= hpz; ++v) sum += (float)a.ref_img[(i + u) * a.W + j + v];
  const float mean = __fdiv_rn(sum, (float)(P * P));
  float sum2 = 0;
#pragma unroll
  for (int u = -hpz; u <= hpz; ++u)
#pragma unroll
    for (int v = -hpz; v <= hpz; ++v) {
      const float dd = (float)a.ref_img[(i + u) * a.W + j + v] - mean;
      sum2 += dd * dd;
    }
  if (__fdiv_rn(sum2, (float)(P * P)) < a.min_patch_var) return 0;

  const u4 r = philox4x32_10(a.seed_lo ^ (uint32_t)idx, a.seed_hi,
                             u4{0xFFFFFFFFu, 0u, 0u, 0u});
  float depth = expf(a.log_min + (a.log_max - a.log_min) * u01(r.x));
  float n0 = -1.0f + 2.0f * u01(r.y);
  float n1 = -1.0f + 2.0f * u01(r.z);
  bool seeded = false;
  if (a.init_map) {
    // Learned initialisation (not in the reference): the map's
    // (t, sx, sy, c) blended bilinearly at this pixel's position in the
    // coarse grid; confident pixels (logit c >= 0) start from depth
    // exp(log_min + t (log_max - log_min)) and normal (sx, sy, -1), the
    // others from the random plane above.
    float m[4];
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) m[ch] = init_map_at(a, i, j, ch);
    if (m[3] >= 0.0f) {
      // the map's plane, perturbed per pixel by Gaussians of spread
      // init_sigma (log depth) and init_sigma_slope, so that neighbouring
      // pixels hold distinct nearby hypotheses for propagation to choose
      // among (0: the map's plane itself)
      float g0, g1, g2, g3;
      box_muller(r.x, r.y, &g0, &g1);
      box_muller(r.z, r.w, &g2, &g3);
      const float t = fminf(fmaxf(m[0], 0.0f), 1.0f);
      depth = expf(a.log_min + (a.log_max - a.log_min) * t + a.init_sigma * g0);
      n0 = fminf(fmaxf(m[1] + a.init_sigma_slope * g1, -50.0f), 50.0f);
      n1 = fminf(fmaxf(m[2] + a.init_sigma_slope * g2, -50.0f), 50.0f);
      seeded = true;
    }
  }
  const PixelRay ray = pixel_ray(a, i, j);
  float v[3];
  plane_from_depth_normal(ray, depth, n0, n1, -1.0f, v);
  RefPatch<P> rp;
  float s = -1.0f;
  int nb = 0;
  unsigned count = 0;
  const bool textured = load_ref_patch<P, true>(rp, a.ref_img, a.W, i, j, a.wlut);
  float dtmp;
  if constexpr (kSample) {
    // one random neighbour view (uni_(rng_), depthmap.cc:247-249)
    nb = 1 + (int)(r.w % (uint32_t)(a.nviews - 1));
    s = score_plane_view<P, Prec>(a, rp, ray, i, j, v[0], v[1], v[2], nb, textured, &dtmp);
    count = textured ? 1 : 0;
  } else if (textured) {
    s = score_plane_all<P, Prec>(a, rp, ray, i, j, v[0], v[1], v[2], &nb, &dtmp);
    count = a.nviews - 1;
    // A seeded pixel also searches depth along the seed's normal, coarse
    // then fine: the seed's depth scaled by exp(+-0.03), exp(+-0.06)
    // (candidates 2..5 of init_search), then the best so far scaled by
    // exp(+-0.01), exp(+-0.02), ... (the rest).  The network's depth is
    // accurate to a few percent; the first refinement steps reach 2 %.
    if (seeded) {
      const int n_coarse = min(a.init_search, 5);
      float best_depth = depth, centre = depth;
      for (int k = 1; k < a.init_search; ++k) {
        const bool fine = k >= n_coarse;
        if (k == n_coarse) centre = best_depth;  // fine stage around the winner
        const int m = fine ? k - n_coarse : k - 1;
        const float mag = (fine ? 0.01f : 0.03f) * (float)(m / 2 + 1);
        const float dk = centre * expf((m & 1) ? -mag : mag);
        float vk[3];
        plane_from_depth_normal(ray, dk, n0, n1, -1.0f, vk);
        int nbk;
        const float sk = score_plane_all<P, Prec>(a, rp, ray, i, j, vk[0], vk[1], vk[2], &nbk, &dtmp);
        count += a.nviews - 1;
        if (sk > s) {
          s = sk;
          nb = nbk;
          v[0] = vk[0];
          v[1] = vk[1];
          v[2] = vk[2];
          best_depth = dk;
        }
      }
      depth = best_depth;
    }
  }
  reinterpret_cast<float4*>(a.state)[idx] = make_float4(v[0], v[1], v[2], depth);
  a.score[idx] = s;
  a.nghbr[idx] = nb;
  return count;
}

template <int P, bool kSample, int Prec>
__global__ __launch_bounds__(256) void k_pm_init(PmArgs a) {
  const Tile t = row_tile((a.W + 15) >> 4, (a.H + 15) >> 4);
  const unsigned n = t.ok ? init_pixel<P, kSample, Prec>(a, t.x, t.y) : 0u;
  if (a.counter) wave_count(a.counter, n);
}

// ---------------------------------------------------------------------------
// Inputs of the learned initialiser (not in the reference).

// Box-filtered (mode 0: mean, rounded; mode 1: min) downsampling of a u8
// image by an integer factor: out[i][j] over in[s i .. s i + s - 1][s j ..].
__global__ void k_downsample_u8(const uint8_t* __restrict__ in, int W, int H,
                                int sc, int mode, uint8_t* __restrict__ out,
                                int Wl, int Hl) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= Hl || j >= Wl) return;
  int acc = mode ? 255 : 0;
  for (int u = 0; u < sc; ++u)
    for (int v = 0; v < sc; ++v) {
      const int val = in[(size_t)(i * sc + u) * W + (j * sc + v)];
      acc = mode ? min(acc, val) : acc + val;
    }
  out[(size_t)i * Wl + j] =
      (uint8_t)(mode ? acc : (acc + (sc * sc) / 2) / (sc * sc));
}

// The network input at one coarse pixel, NHWC BF16 with 40 channels:
// channel 0 the image (x - 128) / 64, channels 1..nplanes the max-over-views
// NCC of fronto-parallel planes at depths uniform in inverse depth over
// [min, max] (the plane sweep of ComputeBruteForce, depthmap.cc:184-200,
// scored by the PatchMatch scorer with the given sample precision; -1 for
// untextured or border pixels), the rest 0.  Pixels of the padded grid
// beyond the image are all 0.
template <int P, int Prec>
__global__ __launch_bounds__(256) void k_init_input(PmArgs a, int nplanes,
                                                    float inv_min,
                                                    float inv_step,
                                                    __bf16* __restrict__ out,
                                                    int ow, int oh) {
  const int j = blockIdx.x * 16 + (threadIdx.x & 15);
  const int i = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= oh || j >= ow) return;
  constexpr int C = 40;
  __bf16* o = out + ((size_t)i * ow + j) * C;
  const bool inside = i < a.H && j < a.W;
  const bool interior = inside && is_interior(a, i, j);
  RefPatch<P> rp;
  const bool textured =
      interior && load_ref_patch<P, true>(rp, a.ref_img, a.W, i, j, a.wlut);
  o[0] = (__bf16)(inside ? ((float)a.ref_img[(size_t)i * a.W + j] - 128.0f) / 64.0f : 0.0f);
  const PixelRay ray = pixel_ray(a, i, j);
  for (int d = 0; d < nplanes; ++d) {
    float sc = inside ? -1.0f : 0.0f;
    if (textured) {
      const float depth = 1.0f / (inv_min + (float)d * inv_step);
      int nb;
      float dep;
      sc = score_plane_all<P, Prec>(a, rp, ray, i, j, 0.0f, 0.0f, -1.0f / depth, &nb, &dep);
      sc = sc == sc ? sc : -1.0f;
    }
    o[1 + d] = (__bf16)sc;
  }
  for (int ch = 1 + nplanes; ch < C; ++ch) o[ch] = (__bf16)0.0f;
}

// ---------------------------------------------------------------------------
// Red/black half-iteration.
//
// Every interior pixel with (i + j) % 2 == colour tests the planes of its
// four direct neighbours, the best-scoring pixel among n_far odd-distance
// probes in each direction, n_explore exploration planes (below), and
// n_refine random refinements of its current plane with the reference's
// schedule (depth * exp(0.02 * 0.3^k N), slopes + 0.5 * 0.8^k N,
// depthmap.cc:333-360), each scored over all views.  Only opposite-colour
// pixels are read, so the result does not depend on the order in which
// lanes run.
//
// Texels are gathered straight from global memory.  Staging each view's
// footprint in LDS was measured slower on the 1600x1200 workload (per
// converged iteration: 11.7 ms direct, 13.2 ms with per-wave LDS tiles,
// 16.2 ms with per-workgroup tiles behind barriers): once neighbouring
// pixels hold similar planes, a wave's gathers stay within a few L1 lines,
// and building the footprint box costs more than the LDS reads save.

// Pixels scoring this or more skip exploration.  Measured (single view at 5
// iterations / 16-view street batch at 3): 0.95 gives 34.9 / 119.3 Mpix/s at
// accuracy 0.9609 / 0.7091, 0.97 34.8 / 117.8 at 0.9610 / 0.7134, 0.99
// 34.2 / 115.7 at 0.9612 / 0.7194, exploring everywhere 31.5 / 111.6 at
// 0.9612 / 0.7241.
constexpr float kExploreBelow = 0.97f;

// Tile (tile_x, tile_y) of TX x TY lanes covers columns [2 TX tile_x,
// 2 TX (tile_x + 1)) and rows [TY tile_y, TY (tile_y + 1)); each lane takes
// the pixel of the given colour in its pair of columns.
// kLean (the single-view kernel): the propagation candidates seen so far in
// LDS (hist[m * 256 + threadIdx.x]) instead of registers.  The batched
// kernel keeps them in registers: with the history in LDS it ran at 3 waves
// per SIMD and 26-38 % slower.  Patch 7 on the FP16 coefficients scores with
// two patch rows of gathers in flight (kPipe: converged launches 3-4 %
// faster; three rows were 1 % slower than two, on the u8 quads of the first
// iteration it was slower, and at patch 5 its registers cost the batched
// kernel its third wave per SIMD).
// G > 1 (not in sample mode): G adjacent lanes per pixel, the tile being TX x
// TY pixels of G * TX * TY lanes; every lane of a group runs the pixel's
// candidate loop (same state reads, same random numbers) and scores its share
// of the views (score_plane_all), and lane 0 of the group writes the result.
// The maps are bit-identical for every G; more lanes per pixel fill the chip
// when a launch has few pixels.
template <int P, bool kSample, int TX, int TY, int Prec, bool kLean = false, int G = 1>
__device__ __forceinline__ unsigned half_pixel(const PmArgs& a, int colour,
                                               int iter, int tile_x,
                                               int tile_y, float4* __restrict__ hist) {
  static_assert(G == 1 || !kSample, "lane groups score all views of a candidate");
  const int sub = (int)(threadIdx.x % G);
  const int cx = tile_x * TX + (int)((threadIdx.x / G) % TX);
  const int i = tile_y * TY + (int)((threadIdx.x / G) / TX);
  const int j = 2 * cx + ((i + colour) & 1);
  if (!is_interior(a, i, j)) return 0;
  const size_t idx = (size_t)i * a.W + j;
  const float4* st = reinterpret_cast<const float4*>(a.state);
  float4 cur = st[idx];
  if (cur.w == 0.0f) return 0;
  RefPatch<P> rp;
  if (!load_ref_patch<P, true>(rp, a.ref_img, a.W, i, j, a.wlut)) return 0;
  const PixelRay ray = pixel_ray(a, i, j);
  float best_s = a.score[idx];
  int best_n = a.nghbr[idx];
  const bool explore_all =
      iter >= 1 || (a.init_map != nullptr && init_map_at(a, i, j, 3) >= 0.0f);

  // One loop over 8 propagation, n_explore exploration and n_refine
  // refinement candidates (+ in sample mode the current plane against a
  // random other view), so the scorer is inlined once.  In sample mode every candidate is scored against one view
  // only (depthmap.cc:325-373): a propagated plane against its source
  // pixel's view, a refinement against the pixel's view after propagation.
  constexpr int kPipe = (Prec == 0 && P == 7) ? 2 : 1;
  const int ncand = 8 + a.n_explore + a.n_refine + (kSample ? 1 : 0);
  unsigned tries = 0;
  float depth_range = 0.02f, normal_range = 0.5f;
  float tx[8], ty[8], tz[8];  // propagation planes seen so far (not kLean)
  int tvw[8];                 // and their views (sample mode)
  int cur_view = best_n;      // current_nghbr of depthmap.cc:336
  for (int c = 0; c < ncand; ++c) {
    float v0, v1, v2;
    bool go;
    int view = 0;
    if (c == 8) cur_view = best_n;
    if (c < 8) {
      const int k = c & 3;
      const int di = k == 0 ? -1 : (k == 1 ? 1 : 0);
      const int dj = k == 2 ? -1 : (k == 3 ? 1 : 0);
      size_t q = (size_t)(i + di) * a.W + (j + dj);
      bool ok = true;
      if (c >= 4) {
        float bs = -2.0f;
        q = idx;
        for (int m = 0; m < a.n_far; ++m) {
          const int dist = 3 + 2 * m;
          const int ii = i + di * dist, jj = j + dj * dist;
          if (!is_interior(a, ii, jj)) break;
          const size_t p = (size_t)ii * a.W + jj;
          const float s = a.score[p];
          if (s > bs) {
            bs = s;
            q = p;
          }
        }
        ok = bs > -2.0f;
      }
      const float4 nb = st[q];
      v0 = nb.x;
      v1 = nb.y;
      v2 = nb.z;
      if constexpr (kSample) view = a.nghbr[q];
      // A candidate bit-equal to the current one or to an earlier candidate
      // (plane, and in sample mode view) scores the same and cannot win a
      // strict comparison: skip it.
      bool dup = v0 == cur.x && v1 == cur.y && v2 == cur.z &&
                 (!kSample || view == best_n);
      if constexpr (kLean) {
        for (int m = 0; m < c; ++m) {
          const float4 h = hist[m * 256 + threadIdx.x];
          dup |= h.x == v0 && h.y == v1 && h.z == v2 &&
                 (!kSample || __float_as_int(h.w) == view);
        }
        hist[c * 256 + threadIdx.x] = make_float4(v0, v1, v2, __int_as_float(view));
      } else {
#pragma unroll
        for (int m = 0; m < 7; ++m)
          dup |= m < c && tx[m] == v0 && ty[m] == v1 && tz[m] == v2 &&
                 (!kSample || tvw[m] == view);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          if (m == c) {
            tx[m] = v0;
            ty[m] = v1;
            tz[m] = v2;
            tvw[m] = view;
          }
        }
      }
      go = ok && nb.w != 0.0f && !dup;
    } else if (kSample && c == ncand - 1) {
      // the current plane against a random view other than current_nghbr
      // (depthmap.cc:362-373), only with at least two neighbour views
      const u4 r = philox4x32_10(a.seed_lo ^ (uint32_t)idx, a.seed_hi,
                                 u4{(uint32_t)iter, 14u, 2u, 0u});
      v0 = cur.x;
      v1 = cur.y;
      v2 = cur.z;
      go = a.nviews > 2 && cur_view >= 1;
      int other = 1 + (int)(r.x % (uint32_t)max(1, a.nviews - 2));
      if (other >= cur_view) ++other;
      view = other;
    } else if (c < 8 + a.n_explore) {
      // Exploration (not in the reference).  Strongly slanted surfaces
      // (slopes of 5-50 at grazing angles) are out of reach of the
      // 0.5 * 0.8^k slope steps within a few red/black updates; the
      // reference reaches them by chaining refinements along its sequential
      // sweep.  Samples 0 and 1 are the planes through the 3-D points of
      // the four neighbours at distance 1 and 3 (their depths are often
      // right before their normals are); the others keep the current depth
      // with a normal drawn uniformly over the cone of half-angle 87 degrees
      // around the optical axis.  Pixels already scoring kExploreBelow or
      // more skip them.  In the first iteration
      // pixels that started from a random plane skip samples 0 and 1: their
      // neighbours' points are random too (measured: the 1600x1200 scene
      // at 5 iterations 7 % faster with accuracy 0.9609 against 0.9605; the
      // 16-view street batch at 3 iterations 6 % faster with accuracy
      // 0.7091 against 0.7089 -- skipping sample 2 as well lost 1.3 points
      // there, since its normals reach the grazing surfaces).  Pixels seeded
      // by a confident initialisation map take all samples at once.
      const int k = c - 8;
      float v[3] = {0.0f, 0.0f, 0.0f};
      if (k < 2) {
        // the four neighbours at distance 1 (sample 0) or 3 (sample 1),
        // both of the opposite colour
        const int D = k == 0 ? 1 : 3;
        go = i >= D && i + D < a.H && j >= D && j + D < a.W;
        const size_t dW = (size_t)D * a.W;
        const float dl = go ? st[idx - D].w : 0.0f;
        const float dr = go ? st[idx + D].w : 0.0f;
        const float du = go ? st[idx - dW].w : 0.0f;
        const float dd = go ? st[idx + dW].w : 0.0f;
        // X(i, j +- D) = d (ray +- D c0), X(i +- D, j) = d (ray +- D c1)
        // with c0, c1 the first two columns of K_0^-1
        const float fD = (float)D;
        const float c0x = fD * a.Kinv0f[0], c0y = fD * a.Kinv0f[3],
                    c0z = fD * a.Kinv0f[6];
        const float c1x = fD * a.Kinv0f[1], c1y = fD * a.Kinv0f[4],
                    c1z = fD * a.Kinv0f[7];
        const float sh = dr - dl, sv = dd - du, ah = dr + dl, av = dd + du;
        const float ex = sh * ray.x + ah * c0x, ey = sh * ray.y + ah * c0y,
                    ez = sh * ray.z + ah * c0z;  // X(j+D) - X(j-D)
        const float fx = sv * ray.x + av * c1x, fy = sv * ray.y + av * c1y,
                    fz = sv * ray.z + av * c1z;  // X(i+D) - X(i-D)
        // normal f x e faces the camera for a surface seen from the front
        const float nx = fy * ez - fz * ey, ny = fz * ex - fx * ez,
                    nz = fx * ey - fy * ex;
        const float s4 = ah + av;  // 4 * centroid = s4 ray + sh c0 + sv c1
        const float cx = s4 * ray.x + sh * c0x + sv * c1x,
                    cy = s4 * ray.y + sh * c0y + sv * c1y,
                    cz = s4 * ray.z + sh * c0z + sv * c1z;
        const float nc = 0.25f * (nx * cx + ny * cy + nz * cz);
        go = go && dl > 0.0f && dr > 0.0f && du > 0.0f && dd > 0.0f && nc < 0.0f;
        if (go) {
          const float inv = -1.0f / nc;
          v[0] = nx * inv;
          v[1] = ny * inv;
          v[2] = nz * inv;
        }
      } else {
        const u4 r = philox4x32_10(a.seed_lo ^ (uint32_t)idx, a.seed_hi,
                                   u4{(uint32_t)iter, (uint32_t)k, 3u, 0u});
        const float cos_t = 1.0f - u01(r.x) * (1.0f - 0.05233596f);
        const float tan_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t)) / cos_t;
        float sn, cs;
        __sincosf(6.28318530717958647692f * u01(r.y), &sn, &cs);
        plane_from_depth_normal(ray, cur.w, tan_t * cs, tan_t * sn, -1.0f, v);
        go = true;
      }
      v0 = v[0];
      v1 = v[1];
      v2 = v[2];
      view = cur_view;
      go = go && best_s < kExploreBelow && (explore_all || k >= 2);
    } else {
      const int k = c - 8 - a.n_explore;
      if (k > 0 && k % 6 == 0) {  // the 6-step schedule restarts
        depth_range = 0.02f;
        normal_range = 0.5f;
      }
      const u4 r = philox4x32_10(a.seed_lo ^ (uint32_t)idx, a.seed_hi,
                                 u4{(uint32_t)iter, (uint32_t)k, 1u, 0u});
      float g0, g1, g2, g3;
      box_muller(r.x, r.y, &g0, &g1);
      box_muller(r.z, r.w, &g2, &g3);
      const float depth = cur.w * expf(depth_range * g0);
      go = cur.z != 0.0f;
      float v[3] = {0.0f, 0.0f, 0.0f};
      if (go) {
        const float nx = -cur.x / cur.z + normal_range * g1;
        const float ny = -cur.y / cur.z + normal_range * g2;
        plane_from_depth_normal(ray, depth, nx, ny, -1.0f, v);
        depth_range *= 0.3f;
        normal_range *= 0.8f;
      }
      v0 = v[0];
      v1 = v[1];
      v2 = v[2];
      view = cur_view;
    }
    if constexpr (kSample) {
      float dep;
      const float s = score_plane_view<P, Prec>(a, rp, ray, i, j, v0, v1, v2, view, go, &dep);
      if (go) {
        tries += 1;
        if (s > best_s) {
          best_s = s;
          best_n = view;
          cur = make_float4(v0, v1, v2, dep);
        }
      }
    } else if (go) {
      int nb;
      float dep;
      const float s =
          score_plane_all<P, Prec, kPipe, G>(a, rp, ray, i, j, v0, v1, v2, &nb, &dep, sub);
      tries += sub == 0 ? (unsigned)(a.nviews - 1) : 0u;
      if (s > best_s) {
        best_s = s;
        best_n = nb;
        cur = make_float4(v0, v1, v2, dep);
      }
    }
  }
  if (sub == 0) {
    reinterpret_cast<float4*>(a.state)[idx] = cur;
    a.score[idx] = best_s;
    a.nghbr[idx] = best_n;
  }
  return tries;
}

// ---------------------------------------------------------------------------
// Output packing with the optional PostProcess (depthmap.cc:489-502):
// 5x5 median with replicate border, then depth = 0 where d == 0 or
// |d - m| / d > 0.05.  Plane/score/nghbr are copied unchanged.

__device__ __forceinline__ void cswap(float& x, float& y) {
  const float lo = fminf(x, y), hi = fmaxf(x, y);
  x = lo;
  y = hi;
}

// Median of the 5x5 window of depths around (i, j) with replicate border
// (cv::medianBlur(depth, 5) of a float map).
__device__ __forceinline__ float median5_depth(const float4* __restrict__ st, int W, int H,
                                               int i, int j) {
  float win[25];
#pragma unroll
  for (int u = 0; u < 5; ++u) {
    const int ii = min(max(i + u - 2, 0), H - 1);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const int jj = min(max(j + v - 2, 0), W - 1);
      win[u * 5 + v] = st[(size_t)ii * W + jj].w;
    }
  }
  // odd-even transposition sort (25 rounds) -> exact median win[12]
#pragma unroll
  for (int r = 0; r < 25; ++r) {
#pragma unroll
    for (int k = (r & 1); k + 1 < 25; k += 2) cswap(win[k], win[k + 1]);
  }
  return win[12];
}

// PostProcess's rejection test (depthmap.cc:497-498): true when depth d is
// kept against window median m.
__device__ __forceinline__ bool median_keeps(float d, float m) {
  return d != 0.0f && !(__fdiv_rn(fabsf(d - m), d) > 0.05);
}

// ---------------------------------------------------------------------------
// A level's PatchMatch state as an initialisation map for the next finer
// level (PmArgs::init_map, 4 floats per pixel): t = (log depth - log_min) /
// (log_max - log_min), the slopes (sx, sy) of the plane's normal (sx, sy, -1)
// (plane v: sx = -v0 / v2, sy = -v1 / v2) and a confidence logit, +10 where
// the depth is live (and, with `median`, PostProcess's test keeps it), else
// -10 (the finer level starts those pixels from random planes).

__global__ __launch_bounds__(256) void k_state_to_map(const float4* __restrict__ st, int W,
                                                      int H, float log_min, float log_max,
                                                      int median, float4* __restrict__ map) {
  const int j = blockIdx.x * 16 + (threadIdx.x & 15);
  const int i = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= H || j >= W) return;
  const size_t idx = (size_t)i * W + j;
  const float4 s = st[idx];
  bool ok = s.w > 0.0f && s.z != 0.0f;
  if (ok && median) ok = median_keeps(s.w, median5_depth(st, W, H, i, j));
  float4 m = make_float4(0.0f, 0.0f, 0.0f, -10.0f);
  if (ok) {
    m.x = (logf(s.w) - log_min) / (log_max - log_min);
    m.y = -s.x / s.z;
    m.z = -s.y / s.z;
    m.w = 10.0f;
  }
  map[idx] = m;
}

// The same over every job of a batch: the map of job k lies at the same
// pixel offset as its state (one workgroup per 16 x 16 pixel tile).
__global__ __launch_bounds__(256) void k_state_to_map_batch(const PmJob* __restrict__ jobs,
                                                            int njobs,
                                                            const float4* __restrict__ state0,
                                                            int median,
                                                            float4* __restrict__ map0) {
  const int t = blockIdx.x;
  const PmJob& J = jobs[find_job(jobs, njobs, t, &PmJob::pix_t0)];
  const int lt = t - J.pix_t0;
  const int i = (lt / J.pix_tx) * 16 + (threadIdx.x >> 4);
  const int j = (lt % J.pix_tx) * 16 + (threadIdx.x & 15);
  const int W = J.a.W, H = J.a.H;
  if (i >= H || j >= W) return;
  const float4* st = reinterpret_cast<const float4*>(J.a.state);
  const size_t idx = (size_t)i * W + j;
  const float4 s = st[idx];
  bool ok = s.w > 0.0f && s.z != 0.0f;
  if (ok && median) ok = median_keeps(s.w, median5_depth(st, W, H, i, j));
  float4 m = make_float4(0.0f, 0.0f, 0.0f, -10.0f);
  if (ok) {
    m.x = (logf(s.w) - J.a.log_min) / (J.a.log_max - J.a.log_min);
    m.y = -s.x / s.z;
    m.z = -s.y / s.z;
    m.w = 10.0f;
  }
  map0[(st - state0) + idx] = m;
}

__device__ __forceinline__ void pack_pixel(const PmArgs& a, int median, int i,
                                           int j, float* __restrict__ out_depth,
                                           float* __restrict__ out_plane,
                                           float* __restrict__ out_score,
                                           int* __restrict__ out_nghbr) {
  if (i >= a.H || j >= a.W) return;
  const size_t idx = (size_t)i * a.W + j;
  const float4* st = reinterpret_cast<const float4*>(a.state);
  const float4 s = st[idx];
  float d = s.w;
  if (median && !median_keeps(d, median5_depth(st, a.W, a.H, i, j))) d = 0.0f;
  out_depth[idx] = d;
  out_plane[3 * idx + 0] = s.x;
  out_plane[3 * idx + 1] = s.y;
  out_plane[3 * idx + 2] = s.z;
  out_score[idx] = a.score[idx];
  out_nghbr[idx] = a.nghbr[idx];
}

__global__ __launch_bounds__(256) void k_pack(PmArgs a, int median,
                                              float* __restrict__ out_depth,
                                              float* __restrict__ out_plane,
                                              float* __restrict__ out_score,
                                              int* __restrict__ out_nghbr) {
  pack_pixel(a, median, blockIdx.y * 16 + (threadIdx.x >> 4),
             blockIdx.x * 16 + (threadIdx.x & 15), out_depth, out_plane,
             out_score, out_nghbr);
}

// ---------------------------------------------------------------------------
// Median-consistent re-selection (not in the reference).  A live pixel whose
// depth PostProcess would reject (more than 5 % from the median of its 5x5
// window, depthmap.cc:489-502) re-selects its plane among the planes of that
// window's other pixels whose depth at this pixel passes the same test
// against the median: the best NCC wins (strict >, window order), then the
// reference's refinement schedule around it, each step kept only while it
// passes the test too.  Such pixels hold a plane that fits the patch better
// than the surrounding surface does -- next to occlusion edges, a plane
// bridging the two sides -- which PostProcess would zero.  The re-selection
// is kept when its NCC reaches res.min_score and its depth lies within 1 %
// of the depths of at least res.min_support of the window's other pixels; otherwise, or without a consistent candidate, the pixel keeps its
// state (and PostProcess zeroes it).
//
// Three launches: k_consist_mark lists the rejected pixels (with their
// window medians), k_consist_select computes each listed pixel's new state
// from the unmodified maps into res_*, one lane per listed pixel, and
// k_consist_apply writes them back, so the result does not depend on the
// order of lanes.

// Appends pixel (i, j) of the view described by `a` to the list when
// PostProcess would reject its depth (all lanes of the wave must call this).
__device__ __forceinline__ void consist_mark_pixel(const PmArgs& a, int i, int j, int job,
                                                   unsigned* __restrict__ count,
                                                   unsigned* __restrict__ list_idx,
                                                   float* __restrict__ list_med,
                                                   int* __restrict__ list_job) {
  const float4* st = reinterpret_cast<const float4*>(a.state);
  bool want = false;
  float m = 0.0f;
  size_t idx = 0;
  if (i < a.H && j < a.W && is_interior(a, i, j)) {
    idx = (size_t)i * a.W + j;
    const float d = st[idx].w;
    if (d != 0.0f) {
      m = median5_depth(st, a.W, a.H, i, j);
      want = m != 0.0f && !median_keeps(d, m);
    }
  }
  const unsigned long long mask = __ballot(want);
  if (mask == 0) return;
  const int lane = (int)(threadIdx.x & 63);
  const int leader = __ffsll((long long)mask) - 1;
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(mask));
  base = __shfl(base, leader);
  if (want) {
    const unsigned k = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
    list_idx[k] = (unsigned)idx;
    list_med[k] = m;
    if (list_job) list_job[k] = job;
  }
}

__global__ __launch_bounds__(256) void k_consist_mark(PmArgs a, unsigned* __restrict__ count,
                                                      unsigned* __restrict__ list_idx,
                                                      float* __restrict__ list_med) {
  consist_mark_pixel(a, blockIdx.y * 16 + (threadIdx.x >> 4), blockIdx.x * 16 + (threadIdx.x & 15),
                     0, count, list_idx, list_med, nullptr);
}

__global__ __launch_bounds__(256) void k_consist_mark_batch(const PmJob* __restrict__ jobs,
                                                            int njobs,
                                                            unsigned* __restrict__ count,
                                                            unsigned* __restrict__ list_idx,
                                                            float* __restrict__ list_med,
                                                            int* __restrict__ list_job) {
  const int t = blockIdx.x;
  const int k = find_job(jobs, njobs, t, &PmJob::pix_t0);
  const PmJob& J = jobs[k];
  const int lt = t - J.pix_t0;
  consist_mark_pixel(J.a, (lt / J.pix_tx) * 16 + (threadIdx.x >> 4),
                     (lt % J.pix_tx) * 16 + (threadIdx.x & 15), k, count, list_idx, list_med,
                     list_job);
}

// Window pixel w (0..24, row-major 5x5 around (i, j), replicate border).
__device__ __forceinline__ size_t consist_window(const PmArgs& a, int i, int j, int w) {
  const int ii = min(max(i + w / 5 - 2, 0), a.H - 1);
  const int jj = min(max(j + w % 5 - 2, 0), a.W - 1);
  return (size_t)ii * a.W + jj;
}

// Refinement of the re-selected plane bp (score bs, view bn) with the
// reference's schedule (depthmap.cc:333-360): step k perturbs the best plane
// so far; a step whose depth fails the median test against m is not scored.
// Each lane runs on to its next step that passes before the wave scores, so
// the lanes of a wave score together.
template <int P, bool kSample, int Prec>
__device__ __forceinline__ unsigned consist_refine(const PmArgs& a, const RefPatch<P>& rp,
                                                   const PixelRay& ray, int i, int j,
                                                   size_t idx, float m, unsigned salt,
                                                   float& bs, int& bn, float4& bp) {
  unsigned tries = 0;
  float depth_range = 0.02f, normal_range = 0.5f;
  int k = 0;
  for (;;) {
    float v[3] = {0.0f, 0.0f, 0.0f};
    bool go = false;
    for (; k < a.n_refine; ++k) {
      if (k > 0 && k % 6 == 0) {
        depth_range = 0.02f;
        normal_range = 0.5f;
      }
      const u4 r = philox4x32_10(a.seed_lo ^ (uint32_t)idx, a.seed_hi,
                                 u4{0xFFFFFFFEu - salt, (uint32_t)k, 4u, 0u});
      float g0, g1, g2, g3;
      box_muller(r.x, r.y, &g0, &g1);
      box_muller(r.z, r.w, &g2, &g3);
      go = bp.z != 0.0f;
      if (go) {
        const float depth = bp.w * expf(depth_range * g0);
        const float nx = -bp.x / bp.z + normal_range * g1;
        const float ny = -bp.y / bp.z + normal_range * g2;
        plane_from_depth_normal(ray, depth, nx, ny, -1.0f, v);
        depth_range *= 0.3f;
        normal_range *= 0.8f;
      }
      const PlaneAtPixel pp = plane_at_pixel(a, ray, v[0], v[1], v[2]);
      go = go && median_keeps(__frcp_rn(fmaxf(1e-6f, -pp.gq)), m);
      if (go) break;
    }
    if (!go) break;
    ++k;
    int nb = bn;
    float d2 = 0.0f, sc;
    if constexpr (kSample) {
      sc = score_plane_view<P, Prec>(a, rp, ray, i, j, v[0], v[1], v[2], bn, true, &d2);
      tries += 1;
    } else {
      sc = score_plane_all<P, Prec>(a, rp, ray, i, j, v[0], v[1], v[2], &nb, &d2);
      tries += (unsigned)(a.nviews - 1);
    }
    if (sc > bs) {
      bs = sc;
      bn = nb;
      bp = make_float4(v[0], v[1], v[2], d2);
    }
  }
  return tries;
}

// The window phase of one listed pixel in plain mode: the window positions
// whose plane passes the median test at this pixel, except a plane equal to
// an earlier window position's (it scores the same, or fails the same test,
// and cannot win a strict comparison).  Bit w stands for window pixel w.
__device__ __forceinline__ unsigned consist_todo(const PmArgs& a, size_t idx, float m) {
  const int i = (int)(idx / (size_t)a.W), j = (int)(idx % (size_t)a.W);
  const float4* st = reinterpret_cast<const float4*>(a.state);
  const PixelRay ray = pixel_ray(a, i, j);
  float px[25], py[25], pz[25];
  unsigned todo = 0;
#pragma unroll
  for (int w = 0; w < 25; ++w) {
    if (w == 12) continue;
    const float4 c = st[consist_window(a, i, j, w)];
    px[w] = c.x;
    py[w] = c.y;
    pz[w] = c.z;
    const PlaneAtPixel pp = plane_at_pixel(a, ray, c.x, c.y, c.z);
    bool go = c.w != 0.0f && median_keeps(__frcp_rn(fmaxf(1e-6f, -pp.gq)), m);
#pragma unroll
    for (int u = 0; u < w; ++u) {
      if (u == 12) continue;
      go = go && !(px[u] == c.x && py[u] == c.y && pz[u] == c.z);
    }
    todo |= go ? (1u << w) : 0u;
  }
  return todo;
}

// Allocates the items of listed entry e (todo from consist_todo; e >= n:
// none) with one atomic per wave, and writes them in window order.  Every
// lane of the wave must call this.
__device__ __forceinline__ void consist_alloc(const ConsistItems& it, unsigned e, unsigned n,
                                              unsigned todo) {
  const unsigned c = e < n ? (unsigned)__popc(todo) : 0u;
  const int lane = (int)(threadIdx.x & 63);
  unsigned incl = c;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned t = __shfl_up(incl, off);
    if (lane >= off) incl += t;
  }
  const unsigned total = __shfl(incl, 63);
  unsigned long long base = 0;
  if (lane == 0 && total) base = atomicAdd(it.count, (unsigned long long)total);
  base = __shfl(base, 0);
  if (e >= n) return;
  const unsigned long long first = base + incl - c;
  const bool fits = first + c <= (unsigned long long)it.cap;
  it.first[e] = fits ? (unsigned)first : 0u;
  it.todo[e] = fits ? todo : (todo | 0x80000000u);
  unsigned long long k = first;
  for (unsigned t = todo; t; t &= t - 1u, ++k) {
    if (k >= (unsigned long long)it.cap) break;
    it.entry[k] = fits ? e : 0xFFFFFFFFu;
    it.win[k] = (unsigned char)__builtin_ctz(t);
  }
}

// Scores window item k (window position w of the pixel at idx) over all
// views.
template <int P, int Prec>
__device__ __forceinline__ unsigned consist_score_item(const PmArgs& a, size_t idx, int w,
                                                       const ConsistItems& it, unsigned k) {
  const int i = (int)(idx / (size_t)a.W), j = (int)(idx % (size_t)a.W);
  const float4* st = reinterpret_cast<const float4*>(a.state);
  RefPatch<P> rp;
  float sc = -1.0f, d2 = 0.0f;
  int nb = 0;
  unsigned tries = 0;
  if (load_ref_patch<P, true>(rp, a.ref_img, a.W, i, j, a.wlut)) {
    const PixelRay ray = pixel_ray(a, i, j);
    const float4 c = st[consist_window(a, i, j, w)];
    sc = score_plane_all<P, Prec>(a, rp, ray, i, j, c.x, c.y, c.z, &nb, &d2);
    tries = (unsigned)(a.nviews - 1);
  }
  it.score[k] = sc;
  it.nghbr[k] = nb;
  it.depth[k] = d2;
  return tries;
}

// Plain mode reads the window phase's scores from the items (in window
// order, strict >, so the first of equal scores wins as in a sequential
// loop); sample mode scores each window plane against its pixel's view here,
// the lanes taking the window positions in step.
template <int P, bool kSample, int Prec>
__device__ __forceinline__ unsigned consist_select(const PmArgs& a, size_t idx, float m,
                                                   const ConsistOut& res, unsigned e,
                                                   unsigned salt, const ConsistItems& items) {
  const int i = (int)(idx / (size_t)a.W), j = (int)(idx % (size_t)a.W);
  const float4* st = reinterpret_cast<const float4*>(a.state);
  float4 cur = st[idx];
  float best_s = a.score[idx];
  int best_n = a.nghbr[idx];
  unsigned tries = 0;
  RefPatch<P> rp;
  if (load_ref_patch<P, true>(rp, a.ref_img, a.W, i, j, a.wlut)) {
    const PixelRay ray = pixel_ray(a, i, j);
    float bs = -2.0f;
    int bn = 0;
    float4 bp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (kSample) {
      // a plane equal to one of the last four scored planes scores the same
      // and cannot win a strict comparison: skipped
      float px[4] = {0.0f, 0.0f, 0.0f, 0.0f}, py[4] = {0.0f, 0.0f, 0.0f, 0.0f},
            pz[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // the last scored planes
      int nseen = 0;
      for (int w = 0; w < 25; ++w) {
        if (w == 12) continue;
        const size_t q = consist_window(a, i, j, w);
        const float4 c = st[q];
        bool go = c.w != 0.0f;
        const PlaneAtPixel pp = plane_at_pixel(a, ray, c.x, c.y, c.z);
        go = go && median_keeps(__frcp_rn(fmaxf(1e-6f, -pp.gq)), m);
        bool dup = false;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          dup |= k < nseen && px[k] == c.x && py[k] == c.y && pz[k] == c.z;
        go = go && !dup;
        const int nb = a.nghbr[q];
        float d2 = 0.0f;
        const float sc = score_plane_view<P, Prec>(a, rp, ray, i, j, c.x, c.y, c.z, nb, go, &d2);
        if (go) {
          tries += 1u;
          const int slot = nseen & 3;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            if (k == slot) {
              px[k] = c.x;
              py[k] = c.y;
              pz[k] = c.z;
            }
          }
          ++nseen;
          if (sc > bs) {
            bs = sc;
            bn = nb;
            bp = make_float4(c.x, c.y, c.z, d2);
          }
        }
      }
    } else {
      const unsigned todo = items.todo[e];
      if (!(todo >> 31)) {
        unsigned k = items.first[e];
        for (unsigned t = todo; t; t &= t - 1u, ++k) {
          const float sc = items.score[k];
          if (sc > bs) {
            const float4 c = st[consist_window(a, i, j, __builtin_ctz(t))];
            bs = sc;
            bn = items.nghbr[k];
            bp = make_float4(c.x, c.y, c.z, items.depth[k]);
          }
        }
      } else {
        // items did not fit: score the window planes here
        for (unsigned t = todo & 0x7FFFFFFFu; t; t &= t - 1u) {
          const float4 c = st[consist_window(a, i, j, __builtin_ctz(t))];
          int nb = 0;
          float d2 = 0.0f;
          const float sc = score_plane_all<P, Prec>(a, rp, ray, i, j, c.x, c.y, c.z, &nb, &d2);
          tries += (unsigned)(a.nviews - 1);
          if (sc > bs) {
            bs = sc;
            bn = nb;
            bp = make_float4(c.x, c.y, c.z, d2);
          }
        }
      }
    }
    if (bs > -2.0f) {
      tries += consist_refine<P, kSample, Prec>(a, rp, ray, i, j, idx, m, salt, bs, bn, bp);
      // support: window pixels whose depth is within 1 % of the new depth
      int support = 0;
      for (int w = 0; w < 25; ++w) {
        if (w == 12) continue;
        const float dq = st[consist_window(a, i, j, w)].w;
        support += (dq > 0.0f && fabsf(dq - bp.w) < 0.01f * bp.w) ? 1 : 0;
      }
      if (bs >= res.min_score && support >= res.min_support) {
        cur = bp;
        best_s = bs;
        best_n = bn;
      }
    }
  }
  res.plane[3 * (size_t)e] = cur.x;
  res.plane[3 * (size_t)e + 1] = cur.y;
  res.plane[3 * (size_t)e + 2] = cur.z;
  res.depth[e] = cur.w;
  res.score[e] = best_s;
  res.nghbr[e] = best_n;
  return tries;
}

__global__ __launch_bounds__(256) void k_consist_plan(PmArgs a, const unsigned* __restrict__ count,
                                                      const unsigned* __restrict__ list_idx,
                                                      const float* __restrict__ list_med,
                                                      ConsistItems it) {
  const unsigned n = *count;
  for (unsigned base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
    const unsigned e = base + threadIdx.x;
    const unsigned todo = e < n ? consist_todo(a, list_idx[e], list_med[e]) : 0u;
    consist_alloc(it, e, n, todo);
  }
}

__global__ __launch_bounds__(256) void k_consist_plan_batch(const PmJob* __restrict__ jobs,
                                                            const unsigned* __restrict__ count,
                                                            const unsigned* __restrict__ list_idx,
                                                            const float* __restrict__ list_med,
                                                            const int* __restrict__ list_job,
                                                            ConsistItems it) {
  const unsigned n = *count;
  for (unsigned base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
    const unsigned e = base + threadIdx.x;
    const unsigned todo =
        e < n ? consist_todo(jobs[list_job[e]].a, list_idx[e], list_med[e]) : 0u;
    consist_alloc(it, e, n, todo);
  }
}

template <int P, int Prec>
__global__ __launch_bounds__(256) void k_consist_score(PmArgs a,
                                                       const unsigned* __restrict__ list_idx,
                                                       ConsistItems it) {
  const unsigned long long total = *it.count;
  const unsigned n = (unsigned)(total < it.cap ? total : it.cap);
  unsigned tries = 0;
  for (unsigned k = blockIdx.x * 256u + threadIdx.x; k < n; k += gridDim.x * 256u) {
    const unsigned e = it.entry[k];
    if (e != 0xFFFFFFFFu) tries += consist_score_item<P, Prec>(a, list_idx[e], it.win[k], it, k);
  }
  if (a.counter) wave_count(a.counter, tries);
}

// The texture descriptors must be wave-uniform, so the wave takes the jobs
// of its items one at a time (items of one job are mostly contiguous).
template <int P, int Prec>
__global__ __launch_bounds__(256) void k_consist_score_batch(
    const PmJob* __restrict__ jobs, const unsigned* __restrict__ list_idx,
    const int* __restrict__ list_job, ConsistItems it, unsigned long long* __restrict__ counter) {
  const unsigned long long total = *it.count;
  const unsigned n = (unsigned)(total < it.cap ? total : it.cap);
  unsigned tries = 0;
  for (unsigned base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
    const unsigned k = base + threadIdx.x;
    const unsigned e = k < n ? it.entry[k] : 0xFFFFFFFFu;
    bool pend = e != 0xFFFFFFFFu;
    const int job = pend ? list_job[e] : -1;
    for (;;) {
      const unsigned long long m = __ballot(pend);
      if (m == 0) break;
      const int ju = __builtin_amdgcn_readfirstlane(__shfl(job, __ffsll((long long)m) - 1));
      if (pend && job == ju) {
        tries += consist_score_item<P, Prec>(jobs[ju].a, list_idx[e], it.win[k], it, k);
        pend = false;
      }
    }
  }
  if (counter) wave_count(counter, tries);
}

template <int P, bool kSample, int Prec>
__global__ __launch_bounds__(256) void k_consist_select(PmArgs a, const unsigned* __restrict__ count,
                                                        const unsigned* __restrict__ list_idx,
                                                        const float* __restrict__ list_med,
                                                        ConsistOut res, unsigned salt,
                                                        ConsistItems items) {
  const unsigned n = *count;
  unsigned tries = 0;
  for (unsigned e = blockIdx.x * 256u + threadIdx.x; e < n; e += gridDim.x * 256u)
    tries += consist_select<P, kSample, Prec>(a, list_idx[e], list_med[e], res, e, salt, items);
  if (a.counter) wave_count(a.counter, tries);
}

template <int P, bool kSample, int Prec>
__global__ __launch_bounds__(256) void k_consist_select_batch(
    const PmJob* __restrict__ jobs, const unsigned* __restrict__ count,
    const unsigned* __restrict__ list_idx, const float* __restrict__ list_med,
    const int* __restrict__ list_job, ConsistOut res, unsigned salt,
    unsigned long long* __restrict__ counter, ConsistItems items) {
  // The views' texture descriptors must be wave-uniform, so the wave takes
  // the jobs of its entries one at a time (entries of one job are mostly
  // contiguous in the list).
  const unsigned n = *count;
  unsigned tries = 0;
  for (unsigned base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
    const unsigned e = base + threadIdx.x;
    bool pend = e < n;
    const int job = pend ? list_job[e] : -1;
    for (;;) {
      const unsigned long long m = __ballot(pend);
      if (m == 0) break;
      const int ju = __builtin_amdgcn_readfirstlane(__shfl(job, __ffsll((long long)m) - 1));
      if (pend && job == ju) {
        tries += consist_select<P, kSample, Prec>(jobs[ju].a, list_idx[e], list_med[e], res, e,
                                                  salt, items);
        pend = false;
      }
    }
  }
  if (counter) wave_count(counter, tries);
}

__device__ __forceinline__ void consist_apply(const PmArgs& a, unsigned idx, const ConsistOut& res,
                                              unsigned e) {
  reinterpret_cast<float4*>(a.state)[idx] =
      make_float4(res.plane[3 * (size_t)e], res.plane[3 * (size_t)e + 1],
                  res.plane[3 * (size_t)e + 2], res.depth[e]);
  a.score[idx] = res.score[e];
  a.nghbr[idx] = res.nghbr[e];
}

__global__ __launch_bounds__(256) void k_consist_apply(PmArgs a, const unsigned* __restrict__ count,
                                                       const unsigned* __restrict__ list_idx,
                                                       ConsistOut res) {
  const unsigned n = *count;
  for (unsigned e = blockIdx.x * 256u + threadIdx.x; e < n; e += gridDim.x * 256u)
    consist_apply(a, list_idx[e], res, e);
}

__global__ __launch_bounds__(256) void k_consist_apply_batch(const PmJob* __restrict__ jobs,
                                                             const unsigned* __restrict__ count,
                                                             const unsigned* __restrict__ list_idx,
                                                             const int* __restrict__ list_job,
                                                             ConsistOut res) {
  const unsigned n = *count;
  for (unsigned e = blockIdx.x * 256u + threadIdx.x; e < n; e += gridDim.x * 256u)
    consist_apply(jobs[list_job[e]].a, list_idx[e], res, e);
}

// ---------------------------------------------------------------------------
// DepthmapCleaner::Clean (depthmap.cc:524-555): FP64 back-projection and
// projection, FP32 bilinear of the other depthmaps, vote counting from 1.

__device__ __forceinline__ float lerp_depth(const float* __restrict__ img,
                                            int W, int H, float y, float x) {
  if (!(x >= 0.0f) || x >= W - 1 || !(y >= 0.0f) || y >= H - 1) return 0.0f;
  const int ix = (int)x, iy = (int)y;
  const float dx = x - ix, dy = y - iy;
  const float* r0 = img + (size_t)iy * W + ix;
  const float* r1 = r0 + W;
  const float im0 = (1 - dx) * r0[0] + dx * r0[1];
  const float im1 = (1 - dx) * r1[0] + dx * r1[1];
  return (1 - dy) * im0 + dy * im1;
}

__device__ __forceinline__ void clean_pixel(const ClView* __restrict__ views,
                                            const ClRef* __restrict__ ref,
                                            int n, float thr, int min_views,
                                            int i, int j,
                                            float* __restrict__ out) {
  const ClView& V0 = views[0];
  if (i >= V0.height || j >= V0.width) return;
  const size_t idx = (size_t)i * V0.width + j;
  const float depth = V0.depth[idx];
  // Backproject (depthmap.cc:109-112): R^T (depth K^-1 (x, y, 1) - t)
  double S[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) S[k] = ref->Kinv[k] * (double)depth;
  const double q[3] = {(double)j, (double)i, 1.0};
  double X[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double s = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += S[r * 3 + k] * q[k];
    X[r] = s - V0.t[r];
  }
  double Y[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double s = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += ref->Rt[r * 3 + k] * X[k];
    Y[r] = s;
  }
  const float p0 = (float)Y[0], p1 = (float)Y[1], p2 = (float)Y[2];
  int consistent = 1;
  for (int o = 1; o < n; ++o) {
    const ClView& Vo = views[o];
    const double pd[3] = {(double)p0, (double)p1, (double)p2};
    double Z[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double s = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) s += Vo.R[r * 3 + k] * pd[k];
      Z[r] = s + Vo.t[r];
    }
    double U[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double s = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) s += Vo.K[r * 3 + k] * Z[k];
      U[r] = s;
    }
    const float r0 = (float)U[0], r1 = (float)U[1], r2 = (float)U[2];
    if (r2 < 1e-8 || isnan(r2)) continue;
    const float u = __fdiv_rn(r0, r2);
    const float v = __fdiv_rn(r1, r2);
    const float dr = lerp_depth(Vo.depth, Vo.width, Vo.height, v, u);
    if (fabsf(dr - r2) < r2 * thr) consistent++;
  }
  out[idx] = consistent >= min_views ? depth : 0.0f;
}

__global__ __launch_bounds__(256) void k_clean(const ClView* __restrict__ views,
                                               const ClRef* __restrict__ ref,
                                               int n, float thr, int min_views,
                                               float* __restrict__ out) {
  clean_pixel(views, ref, n, thr, min_views,
              blockIdx.y * 16 + (threadIdx.x >> 4),
              blockIdx.x * 16 + (threadIdx.x & 15), out);
}

// ---------------------------------------------------------------------------
// Batched launches: many reference views (jobs) of different sizes and view
// counts in one launch.  The tiles of all jobs form one flat index space in
// job order; a tile finds its job by binary search over the jobs' first
// tile indices (wave-uniform, so the job's arguments come in through scalar
// loads).  The red/black and init kernels are persistent: a grid sized to
// the resident workgroups pulls tiles from an atomic queue until it runs
// dry, so large and small views balance across the CUs.  Each launch has
// its own queue slot, cleared by k_fill_u32 at the start of a run (a kernel,
// not a memset, so that the run can be captured in a graph).

__global__ void k_fill_u32(unsigned* __restrict__ p, size_t n, unsigned v) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    p[k] = v;
}

// Next tile of the workgroup from the queue (wave-uniform result).
__device__ __forceinline__ int next_tile(unsigned* __restrict__ queue) {
  __shared__ int s_tile;
  if (threadIdx.x == 0) s_tile = (int)atomicAdd(queue, 1u);
  __syncthreads();
  const int t = __builtin_amdgcn_readfirstlane(s_tile);
  __syncthreads();
  return t;
}

template <int P, bool kSample, int Prec>
__global__ __launch_bounds__(256) void k_pm_init_batch(
    const PmJob* __restrict__ jobs, int njobs, int total,
    unsigned* __restrict__ queue, unsigned long long* __restrict__ count) {
  unsigned n = 0;
  for (;;) {
    const int t = next_tile(queue);
    if (t >= total) break;
    const PmJob& J = jobs[find_job(jobs, njobs, t, &PmJob::pix_t0)];
    const int lt = t - J.pix_t0;
    n += init_pixel<P, kSample, Prec>(J.a, lt % J.pix_tx, lt / J.pix_tx);
  }
  if (count) wave_count(count, n);
}

// Single-view red/black half-iteration, persistent and XCD-banded.  The
// tile rows are split into 8 bands; workgroup b starts on band b % 8 -- the
// XCD the hardware deals it to -- and pulls that band's tiles from the
// band's queue, then moves on to the other bands' queues until all are dry.
// Within a band the tiles go in column blocks of kBandBlockCols tiles, each
// block row by row, so an XCD's active tiles cover a compact region and its
// L2 holds that region's neighbour-view footprint; the band switching at the
// end balances the XCDs' loads.  The order changes no result (a half-
// iteration reads only the other colour).  queues[0..7] start at zero.
constexpr int kBandBlockCols = 20;  // 640 image columns; 12-25 measured alike

template <int P, bool kSample, int Prec>
__global__ __launch_bounds__(256) void k_pm_half(PmArgs a, int colour, int iter,
                                                 unsigned* __restrict__ queues) {
  const int tiles_x = ((a.W + 1) / 2 + 15) >> 4;
  const int tiles_y = (a.H + 15) >> 4;
  const int home = (int)(blockIdx.x & 7u);
  __shared__ float4 hist[8 * 256];
  unsigned n = 0;
  for (int s = 0; s < 8; ++s) {
    const int band = (home + s) & 7;
    const int r0 = band * tiles_y / 8, rows = (band + 1) * tiles_y / 8 - r0;
    const int per_block = kBandBlockCols * rows;
    for (;;) {
      const int u = next_tile(queues + band);
      if (u >= rows * tiles_x) break;
      const int blk = u / per_block, rem = u - blk * per_block;
      const int bw = min(kBandBlockCols, tiles_x - blk * kBandBlockCols);
      n += half_pixel<P, kSample, 16, 16, Prec, true>(a, colour, iter,
                                                      blk * kBandBlockCols + rem % bw,
                                                      r0 + rem / bw, hist);
    }
  }
  if (a.counter) wave_count(a.counter, n);
}

// Single-view red/black half-iteration with G lanes per pixel, for launches
// with few pixels (the coarse levels of a pyramid), where one lane per pixel
// leaves most of the chip idle and each wave waits on its own serial loop
// over the views.  One workgroup per tile of 8 x (32 / (G / 8)) pixels; maps
// bit-identical to k_pm_half's.
template <int P, int Prec, int G>
__global__ __launch_bounds__(256) void k_pm_half_group(PmArgs a, int colour, int iter) {
  constexpr int TX = 8, TY = 256 / G / TX;
  __shared__ float4 hist[8 * 256];
  const int tiles_x = ((a.W + 1) / 2 + TX - 1) / TX;
  const int t = (int)blockIdx.x;
  const unsigned n = half_pixel<P, false, TX, TY, Prec, true, G>(a, colour, iter, t % tiles_x,
                                                                 t / tiles_x, hist);
  if (a.counter) wave_count(a.counter, n);
}

// Batched red/black half-iteration: the flat tile space of all jobs split
// into 8 bands, pulled XCD by XCD as in k_pm_half.  queues[0..7] start at
// zero.
template <int P, bool kSample, int TX, int TY, int Prec>
__global__ __launch_bounds__(256) void k_pm_half_batch(
    const PmJob* __restrict__ jobs, int njobs, int total,
    unsigned* __restrict__ queues, int colour, int iter,
    unsigned long long* __restrict__ count) {
  const int home = (int)(blockIdx.x & 7u);
  unsigned n = 0;
  for (int s = 0; s < 8; ++s) {
    const int band = (home + s) & 7;
    const int t0 = (int)((long long)band * total / 8);
    const int t1 = (int)((long long)(band + 1) * total / 8);
    for (;;) {
      const int t = t0 + next_tile(queues + band);
      if (t >= t1) break;
      const PmJob& J = jobs[find_job(jobs, njobs, t, &PmJob::half_t0)];
      // a job's tiles in column blocks of 512 image columns, each block row
      // by row (compact active regions, as in k_pm_half; no result changes)
      constexpr int B = 256 / TX;
      const int lt = t - J.half_t0;
      const int per = B * J.half_ty;
      const int blk = lt / per, rem = lt - blk * per;
      const int bw = min(B, J.half_tx - blk * B);
      n += half_pixel<P, kSample, TX, TY, Prec>(J.a, colour, iter, blk * B + rem % bw,
                                                rem / bw, nullptr);
    }
  }
  if (count) wave_count(count, n);
}

// One workgroup per 16 x 16 pixel tile (memory-bound, no queue needed).
__global__ __launch_bounds__(256) void k_pack_batch(const PmJob* __restrict__ jobs,
                                                    int njobs, int median) {
  const int t = blockIdx.x;
  const PmJob& J = jobs[find_job(jobs, njobs, t, &PmJob::pix_t0)];
  const int lt = t - J.pix_t0;
  pack_pixel(J.a, median, (lt / J.pix_tx) * 16 + (threadIdx.x >> 4),
             (lt % J.pix_tx) * 16 + (threadIdx.x & 15), J.out_depth,
             J.out_plane, J.out_score, J.out_nghbr);
}

__global__ __launch_bounds__(256) void k_clean_batch(
    const ClJob* __restrict__ jobs, int njobs, const ClView* __restrict__ views,
    float thr, int min_views) {
  const int t = blockIdx.x;
  const ClJob& J = jobs[find_job(jobs, njobs, t, &ClJob::pix_t0)];
  const int lt = t - J.pix_t0;
  clean_pixel(views + J.view0, &J.ref, J.nviews, thr, min_views,
              (lt / J.pix_tx) * 16 + (threadIdx.x >> 4),
              (lt % J.pix_tx) * 16 + (threadIdx.x & 15), J.out);
}

// ---------------------------------------------------------------------------
// GPUDepthmapCleaner::Clean (cleaner.cc:56-179): forward depth check with
// space carving and a backward reprojection check, FP32 throughout, in the
// operation order of its kernel (opencl_kernels.h:2155-2273).  Neighbour
// depths come resized to the reference size (cv::resize INTER_LINEAR,
// cleaner.cc:114-117) while their K stay as given, as the reference does.

// cv::resize(INTER_LINEAR) of a float map: source coordinate
// (x + 0.5) * sw / w - 0.5, floor and fraction, clamped at the borders.
__global__ void k_resize_linear(const float* __restrict__ src, int sw, int sh,
                                float* __restrict__ dst, int w, int h) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  auto axis = [](int d, int n_dst, int n_src, int* i0, int* i1, float* a) {
    const float f = (float)(((double)d + 0.5) * ((double)n_src / n_dst) - 0.5);
    int i = (int)floorf(f);
    float fr = f - (float)i;
    if (i < 0) {
      fr = 0.0f;
      i = 0;
    }
    if (i >= n_src - 1) {
      fr = 0.0f;
      i = n_src - 1;
    }
    *i0 = i;
    *i1 = min(i + 1, n_src - 1);
    *a = fr;
  };
  int x0, x1, y0, y1;
  float ax, ay;
  axis(x, w, sw, &x0, &x1, &ax);
  axis(y, h, sh, &y0, &y1, &ay);
  const float* r0 = src + (size_t)y0 * sw;
  const float* r1 = src + (size_t)y1 * sw;
  const float top = r0[x0] * (1.0f - ax) + r0[x1] * ax;
  const float bot = r1[x0] * (1.0f - ax) + r1[x1] * ax;
  dst[(size_t)y * w + x] = top * (1.0f - ay) + bot * ay;
}

__global__ __launch_bounds__(256) void k_gpu_clean(
    const float* __restrict__ ref_depth, const float* const* __restrict__ src,
    const GcCamera* __restrict__ cams, int w, int h, int nsrc,
    float thr, int min_views, float carving, int max_carved,
    float* __restrict__ out) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= w || y >= h) return;
  const size_t idx = (size_t)y * w + x;
  const float d = ref_depth[idx];
  if (!(d > 0.0f)) {
    out[idx] = 0.0f;
    return;
  }
  const GcCamera& rc = cams[0];
  const float fx = rc.K[0], cx = rc.K[2], fy = rc.K[4], cy = rc.K[5];
  const float px = (float)x, py = (float)y;
  const float pcx = __fdiv_rn(px - cx, fx) * d - rc.t[0];
  const float pcy = __fdiv_rn(py - cy, fy) * d - rc.t[1];
  const float pcz = d - rc.t[2];
  const float wx = rc.R[0] * pcx + rc.R[3] * pcy + rc.R[6] * pcz;
  const float wy = rc.R[1] * pcx + rc.R[4] * pcy + rc.R[7] * pcz;
  const float wz = rc.R[2] * pcx + rc.R[5] * pcy + rc.R[8] * pcz;
  int consistent = 1, carved = 0;
  for (int v = 0; v < nsrc; ++v) {
    const GcCamera& sc = cams[v + 1];
    const float sx = sc.R[0] * wx + sc.R[1] * wy + sc.R[2] * wz + sc.t[0];
    const float sy = sc.R[3] * wx + sc.R[4] * wy + sc.R[5] * wz + sc.t[1];
    const float sz = sc.R[6] * wx + sc.R[7] * wy + sc.R[8] * wz + sc.t[2];
    if (!(sz > 1e-6f)) continue;
    const float su = __fdiv_rn(sc.K[0] * sx, sz) + sc.K[2];
    const float sv = __fdiv_rn(sc.K[4] * sy, sz) + sc.K[5];
    if (!(su >= 0.0f) || !(sv >= 0.0f) || su >= (float)(w - 1) || sv >= (float)(h - 1)) continue;
    // bilinear depth of the (resized) neighbour map at (su, sv)
    const int x0 = (int)su, y0 = (int)sv;
    const float a = su - (float)x0, b = sv - (float)y0;
    const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
    const float* m = src[v];
    const float r0 = m[(size_t)y0 * w + x0] * (1.0f - a) + m[(size_t)y0 * w + x1] * a;
    const float r1 = m[(size_t)y1 * w + x0] * (1.0f - a) + m[(size_t)y1 * w + x1] * a;
    const float sd = r0 * (1.0f - b) + r1 * b;
    if (!(sd > 0.0f)) continue;
    if (sd > sz * (1.0f + carving)) {  // the neighbour sees through the point
      carved++;
      continue;
    }
    if (!(fabsf(sd - sz) < sz * thr)) continue;
    // back from the neighbour's depth into the reference view
    const float qx = __fdiv_rn(su - sc.K[2], sc.K[0]) * sd - sc.t[0];
    const float qy = __fdiv_rn(sv - sc.K[5], sc.K[4]) * sd - sc.t[1];
    const float qz = sd - sc.t[2];
    const float w2x = sc.R[0] * qx + sc.R[3] * qy + sc.R[6] * qz;
    const float w2y = sc.R[1] * qx + sc.R[4] * qy + sc.R[7] * qz;
    const float w2z = sc.R[2] * qx + sc.R[5] * qy + sc.R[8] * qz;
    const float rx = rc.R[0] * w2x + rc.R[1] * w2y + rc.R[2] * w2z + rc.t[0];
    const float ry = rc.R[3] * w2x + rc.R[4] * w2y + rc.R[5] * w2z + rc.t[1];
    const float rz = rc.R[6] * w2x + rc.R[7] * w2y + rc.R[8] * w2z + rc.t[2];
    if (!(rz > 1e-6f)) continue;
    const float ru = __fdiv_rn(fx * rx, rz) + cx;
    const float rv = __fdiv_rn(fy * ry, rz) + cy;
    if (fabsf(ru - px) + fabsf(rv - py) > 2.0f) continue;
    if (!(fabsf(rz - d) < d * thr)) continue;
    consistent++;
  }
  out[idx] = (consistent >= min_views && carved < max_carved) ? d : 0.0f;
}

// ---------------------------------------------------------------------------
// DepthmapPruner::Prune (depthmap.cc:574-636): a pixel of view 0 is dropped
// when another view sees, at the (rounded) projection, a depth beyond
// (1 - threshold) times the point's depth with a larger "area"
// -n_z / depth * f.  Types as the reference: cv::normalize of a Vec3f
// (FP64 norm and scale, FP32 result), FP64 Backproject/Project, FP32 area.
// Two passes give the kept pixels in row-major order: counts per
// 256-pixel block, then a scatter at the blocks' exclusive prefix offsets.

__device__ __forceinline__ void normalize3f(const float* v, float* out) {
  double s = 0.0;
  s += (double)v[0] * (double)v[0];
  s += (double)v[1] * (double)v[1];
  s += (double)v[2] * (double)v[2];
  const double nv = __dsqrt_rn(s);
  const double a = nv != 0.0 ? __ddiv_rn(1.0, nv) : 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) out[k] = (float)((double)v[k] * a);
}

// x / z + 0.5 is rounded to std::int64_t and used through int arguments
// (IsInsideImage, Mat::at): the low 32 bits.  Outside int64 or NaN the
// conversion is undefined in C++; x86 gives INT64_MIN (low bits 0), and so
// does this.
__device__ __forceinline__ int ref_pixel_index(double u) {
  if (!(fabs(u) < 9.223372036854775808e18)) return 0;
  return (int)(unsigned int)(unsigned long long)(long long)u;
}

__device__ __forceinline__ void mat3_vec_d(const double* A, const double* x,
                                           double* y) {
  double t[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += A[3 * r + k] * x[k];
    t[r] = s;
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) y[r] = t[r];
}

// Whether pixel p of view 0 is kept; its point and normal when it is.
__device__ __forceinline__ bool prune_pixel(const PrView* __restrict__ views,
                                            int n, const PrRef& ref, float thr,
                                            int p, float* point,
                                            float* normal) {
  const PrView& V0 = views[0];
  const int i = p / V0.width, j = p - (p / V0.width) * V0.width;
  const float depth = V0.depth[p];
  if (depth <= 0.0f) return false;  // a NaN depth goes on, as in the reference
  normalize3f(V0.plane + 3 * (size_t)p, normal);
  const float area = (float)((double)__fdiv_rn(-normal[2], depth) * V0.K[0]);
  double S[9], X[3];
#pragma unroll
  for (int k = 0; k < 9; ++k) S[k] = (double)depth * ref.Kinv[k];
  const double q[3] = {(double)j, (double)i, 1.0};
  mat3_vec_d(S, q, X);
#pragma unroll
  for (int k = 0; k < 3; ++k) X[k] = X[k] - V0.t[k];
  mat3_vec_d(ref.Rt, X, X);
#pragma unroll
  for (int k = 0; k < 3; ++k) point[k] = (float)X[k];
  for (int o = 1; o < n; ++o) {
    const PrView& Vo = views[o];
    const double pd[3] = {point[0], point[1], point[2]};
    double y[3];
    mat3_vec_d(Vo.R, pd, y);
#pragma unroll
    for (int k = 0; k < 3; ++k) y[k] = y[k] + Vo.t[k];
    mat3_vec_d(Vo.K, y, y);
    if (y[2] < 1e-8 || isnan(y[2])) continue;
    const int iu = ref_pixel_index(__ddiv_rn(y[0], y[2]) + 0.5);
    const int iv = ref_pixel_index(__ddiv_rn(y[1], y[2]) + 0.5);
    if (!(iv >= 0 && iv < Vo.height && iu >= 0 && iu < Vo.width)) continue;
    const size_t po = (size_t)iv * Vo.width + (size_t)iu;
    const float d_o = Vo.depth[po];
    if ((double)d_o > (double)(1.0f - thr) * y[2]) {
      float no[3];
      normalize3f(Vo.plane + 3 * po, no);
      if (d_o == 0.0f || (double)__fdiv_rn(-no[2], d_o) * Vo.K[0] > (double)area) return false;
    }
  }
  return true;
}

__global__ __launch_bounds__(256) void k_prune_count(
    const PrView* __restrict__ views, int n, PrRef ref, float thr, int npix,
    int* __restrict__ counts) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  float pt[3], nm[3];
  const bool keep = p < npix && prune_pixel(views, n, ref, thr, p, pt, nm);
  const int c = __syncthreads_count(keep);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void k_prune_write(
    const PrView* __restrict__ views, int n, PrRef ref, float thr, int npix,
    const long long* __restrict__ offsets, const uint8_t* __restrict__ color,
    const uint8_t* __restrict__ label, float* __restrict__ points,
    float* __restrict__ normals, uint8_t* __restrict__ out_color,
    uint8_t* __restrict__ out_label) {
  __shared__ int wave_base[4];
  const int p = blockIdx.x * 256 + threadIdx.x;
  float pt[3], nm[3];
  const bool keep = p < npix && prune_pixel(views, n, ref, thr, p, pt, nm);
  const unsigned long long m = __ballot(keep);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rank = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wave_base[wave] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < 4; ++w) {
      const int c = wave_base[w];
      wave_base[w] = acc;
      acc += c;
    }
  }
  __syncthreads();
  if (!keep) return;
  const long long o = offsets[blockIdx.x] + wave_base[wave] + rank;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += ref.Rinv[3 * r + k] * nm[k];
    points[3 * o + r] = pt[r];
    normals[3 * o + r] = s;
    out_color[3 * o + r] = color[3 * (size_t)p + r];
  }
  out_label[o] = label[p];
}

// ---------------------------------------------------------------------------
// Launch wrappers

static inline dim3 grid16(int w, int h) {
  return dim3((w + 15) / 16, (h + 15) / 16);
}

hipError_t launch_build_quad16(const uint8_t* img, uint32_t* quad16, int W,
                               int H, int precision, hipStream_t s) {
  dim3 blk(16, 16);
  dim3 grd((W + 1 + 15) / 16, (H + 1 + 15) / 16);
  uint2* q = reinterpret_cast<uint2*>(quad16);
  if (precision == 2)
    hipLaunchKernelGGL(k_build_quad16<true>, grd, blk, 0, s, img, q, W, H);
  else if (precision == 1)
    hipLaunchKernelGGL(k_build_quad16<false>, grd, blk, 0, s, img, q, W, H);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_build_quad(const uint8_t* img, uint32_t* quad, uint32_t* quadc,
                             int W, int H, hipStream_t s) {
  dim3 blk(16, 16);
  dim3 grd((W + 1 + 15) / 16, (H + 1 + 15) / 16);
  hipLaunchKernelGGL(k_build_quad, grd, blk, 0, s, img, quad,
                     reinterpret_cast<uint2*>(quadc), W, H);
  return hipGetLastError();
}

#define PM_DISPATCH_P(P_, ...)          \
  switch (P_) {                         \
    case 3: { constexpr int P = 3; __VA_ARGS__; } break; \
    case 5: { constexpr int P = 5; __VA_ARGS__; } break; \
    case 7: { constexpr int P = 7; __VA_ARGS__; } break; \
    case 9: { constexpr int P = 9; __VA_ARGS__; } break; \
    case 11: { constexpr int P = 11; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue; \
  }

hipError_t launch_sweep(const PmArgs& a, const float* Htab, const float* planes,
                        int nplanes, hipStream_t s) {
  const int P = 2 * a.hpz + 1;
  PM_DISPATCH_P(P, hipLaunchKernelGGL(k_sweep<P>, tile_grid((a.W + 15) / 16, (a.H + 15) / 16), dim3(256),
                                      0, s, a, Htab, planes, nplanes));
  return hipGetLastError();
}

hipError_t launch_score_hyp(const PmArgs& a, int view, int n, const int* ij,
                            const float* Hs, float* out, hipStream_t s) {
  const int P = 2 * a.hpz + 1;
  PM_DISPATCH_P(P, hipLaunchKernelGGL(k_score_hyp<P>, dim3((n + 255) / 256),
                                      dim3(256), 0, s, a, view, n, ij, Hs, out));
  return hipGetLastError();
}

// Reduced-precision samples are built for patch sizes 5 and 7 only.
#define PM_DISPATCH_P57(P_, ...)        \
  switch (P_) {                         \
    case 5: { constexpr int P = 5; __VA_ARGS__; } break; \
    case 7: { constexpr int P = 7; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue; \
  }

#define PM_DISPATCH_PREC(PREC_, P_, ...)                                 \
  switch (PREC_) {                                                       \
    case 0: { constexpr int Prec = 0; PM_DISPATCH_P(P_, __VA_ARGS__); } break;   \
    case 1: { constexpr int Prec = 1; PM_DISPATCH_P57(P_, __VA_ARGS__); } break; \
    case 2: { constexpr int Prec = 2; PM_DISPATCH_P57(P_, __VA_ARGS__); } break; \
    case kPrecFp32Quads: {                                               \
      constexpr int Prec = kPrecFp32Quads; PM_DISPATCH_P(P_, __VA_ARGS__); } break; \
    default: return hipErrorInvalidValue;                                \
  }

bool precision_supported(int precision, int patch) {
  return precision == 0 || ((precision == 1 || precision == 2) &&
                            (patch == 5 || patch == 7));
}

hipError_t launch_pm_init(const PmArgs& a, bool sample, hipStream_t s) {
  const int P_ = 2 * a.hpz + 1;
  const dim3 grid = tile_grid((a.W + 15) / 16, (a.H + 15) / 16);
  if (sample) {
    PM_DISPATCH_PREC(a.precision, P_,
                     hipLaunchKernelGGL((k_pm_init<P, true, Prec>), grid, dim3(256), 0, s, a));
  } else {
    PM_DISPATCH_PREC(a.precision, P_,
                     hipLaunchKernelGGL((k_pm_init<P, false, Prec>), grid, dim3(256), 0, s, a));
  }
  return hipGetLastError();
}

template <int P, bool kSample, int Prec>
static hipError_t half_banded(const PmArgs& a, int colour, int iter,
                              unsigned* queues, hipStream_t s) {
  auto f = k_pm_half<P, kSample, Prec>;
  int grid = 0;
  hipError_t e = persistent_grid(f, 256, &grid);
  if (e != hipSuccess) return e;
  const int ntiles = (((a.W + 1) / 2 + 15) / 16) * ((a.H + 15) / 16);
  // whole multiples of 8 keep workgroup b on XCD b % 8
  grid = std::max(8, std::min(grid, (ntiles + 7) / 8 * 8));
  hipLaunchKernelGGL(f, dim3(grid), dim3(256), 0, s, a, colour, iter, queues);
  return hipGetLastError();
}

template <int P, int Prec>
static hipError_t half_group(const PmArgs& a, int colour, int iter, hipStream_t s) {
  constexpr int G = kLanesPerPixelSmall, TX = 8, TY = 256 / G / TX;
  const int ntiles = (((a.W + 1) / 2 + TX - 1) / TX) * ((a.H + TY - 1) / TY);
  hipLaunchKernelGGL((k_pm_half_group<P, Prec, G>), dim3(ntiles), dim3(256), 0, s, a, colour,
                     iter);
  return hipGetLastError();
}

bool half_uses_groups(const PmArgs& a, int lanes_per_pixel) {
  if (lanes_per_pixel == 1) return false;
  if (a.precision != 0 && a.precision != kPrecFp32Quads) return false;
  return lanes_per_pixel == kLanesPerPixelSmall ||
         (size_t)a.W * a.H / 2 < kGroupBelowPixels;
}

hipError_t launch_pm_half(const PmArgs& a, bool sample, int colour, int iter,
                          unsigned* queues, hipStream_t s, int lanes_per_pixel) {
  const int P_ = 2 * a.hpz + 1;
  if (!sample && half_uses_groups(a, lanes_per_pixel)) {
    if (a.precision == 0) {
      constexpr int Prec = 0;
      PM_DISPATCH_P(P_, return (half_group<P, Prec>(a, colour, iter, s)));
    } else {
      constexpr int Prec = kPrecFp32Quads;
      PM_DISPATCH_P(P_, return (half_group<P, Prec>(a, colour, iter, s)));
    }
    return hipErrorInvalidValue;
  }
  if (sample) {
    PM_DISPATCH_PREC(a.precision, P_,
                     return (half_banded<P, true, Prec>(a, colour, iter, queues, s)));
  }
  PM_DISPATCH_PREC(a.precision, P_,
                   return (half_banded<P, false, Prec>(a, colour, iter, queues, s)));
  return hipSuccess;
}

hipError_t launch_state_to_map(const float* state, int W, int H, float log_min, float log_max,
                               int median, float* map, hipStream_t s) {
  hipLaunchKernelGGL(k_state_to_map, grid16(W, H), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(state), W, H, log_min, log_max, median,
                     reinterpret_cast<float4*>(map));
  return hipGetLastError();
}

hipError_t launch_state_to_map_batch(const PmJob* jobs, int njobs, int pix_total,
                                     const float* state0, int median, float* map0,
                                     hipStream_t s) {
  if (pix_total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_state_to_map_batch, dim3(pix_total), dim3(256), 0, s, jobs, njobs,
                     reinterpret_cast<const float4*>(state0), median,
                     reinterpret_cast<float4*>(map0));
  return hipGetLastError();
}

hipError_t launch_downsample_u8(const uint8_t* in, int W, int H, int sc,
                                int mode, uint8_t* out, int Wl, int Hl,
                                hipStream_t s) {
  dim3 blk(16, 16);
  dim3 grd((Wl + 15) / 16, (Hl + 15) / 16);
  hipLaunchKernelGGL(k_downsample_u8, grd, blk, 0, s, in, W, H, sc, mode, out, Wl, Hl);
  return hipGetLastError();
}

hipError_t launch_init_input(const PmArgs& a, int nplanes, float inv_min,
                             float inv_step, void* out, int ow, int oh,
                             hipStream_t s) {
  const int P_ = 2 * a.hpz + 1;
  dim3 grd((ow + 15) / 16, (oh + 15) / 16);
  __bf16* o = static_cast<__bf16*>(out);
  PM_DISPATCH_P57(P_, hipLaunchKernelGGL((k_init_input<P, 1>), grd, dim3(256), 0, s, a,
                                         nplanes, inv_min, inv_step, o, ow, oh));
  return hipGetLastError();
}

hipError_t launch_score_hyp_fast(const PmArgs& a, int view, int n,
                                 const int* ij, const float* planes, float* out,
                                 hipStream_t s) {
  const int P_ = 2 * a.hpz + 1;
  if (n <= 0) return hipSuccess;
  PM_DISPATCH_PREC(a.precision, P_,
                   hipLaunchKernelGGL((k_score_hyp_fast<P, Prec>), dim3((n + 255) / 256),
                                      dim3(256), 0, s, a, view, n, ij, planes, out));
  return hipGetLastError();
}

hipError_t launch_pm_consist(const PmArgs& a, bool sample, unsigned* count, unsigned* list_idx,
                             float* list_med, const ConsistOut& res, unsigned salt,
                             const ConsistItems& items, hipStream_t s) {
  const int P_ = 2 * a.hpz + 1;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, count, (size_t)1, 0u);
  hipLaunchKernelGGL(k_consist_mark, grid16(a.W, a.H), dim3(256), 0, s, a, count, list_idx,
                     list_med);
  // fixed grids striding over the list and the items (their lengths are on
  // the device)
  const dim3 grid(1024);
  if (sample) {
    PM_DISPATCH_PREC(a.precision, P_,
                     hipLaunchKernelGGL((k_consist_select<P, true, Prec>), grid, dim3(256), 0, s,
                                        a, count, list_idx, list_med, res, salt, items));
  } else {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s,
                       reinterpret_cast<unsigned*>(items.count), (size_t)2, 0u);
    hipLaunchKernelGGL(k_consist_plan, grid, dim3(256), 0, s, a, count, list_idx, list_med, items);
    PM_DISPATCH_PREC(a.precision, P_,
                     hipLaunchKernelGGL((k_consist_score<P, Prec>), dim3(2048), dim3(256), 0, s,
                                        a, list_idx, items));
    PM_DISPATCH_PREC(a.precision, P_,
                     hipLaunchKernelGGL((k_consist_select<P, false, Prec>), grid, dim3(256), 0, s,
                                        a, count, list_idx, list_med, res, salt, items));
  }
  hipLaunchKernelGGL(k_consist_apply, grid, dim3(256), 0, s, a, count, list_idx, res);
  return hipGetLastError();
}

hipError_t launch_pm_consist_batch(const PmJob* jobs, int njobs, int pix_total, int patch,
                                   int precision, bool sample, unsigned* count,
                                   unsigned* list_idx, float* list_med, int* list_job,
                                   const ConsistOut& res, unsigned salt,
                                   const ConsistItems& items, unsigned long long* counter,
                                   hipStream_t s) {
  const int P_ = patch;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, count, (size_t)1, 0u);
  if (pix_total <= 0) return hipGetLastError();
  hipLaunchKernelGGL(k_consist_mark_batch, dim3(pix_total), dim3(256), 0, s, jobs, njobs, count,
                     list_idx, list_med, list_job);
  const dim3 grid(2048);
  if (sample) {
    PM_DISPATCH_PREC(precision, P_,
                     hipLaunchKernelGGL((k_consist_select_batch<P, true, Prec>), grid, dim3(256),
                                        0, s, jobs, count, list_idx, list_med, list_job, res,
                                        salt, counter, items));
  } else {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s,
                       reinterpret_cast<unsigned*>(items.count), (size_t)2, 0u);
    hipLaunchKernelGGL(k_consist_plan_batch, grid, dim3(256), 0, s, jobs, count, list_idx,
                       list_med, list_job, items);
    PM_DISPATCH_PREC(precision, P_,
                     hipLaunchKernelGGL((k_consist_score_batch<P, Prec>), dim3(4096), dim3(256),
                                        0, s, jobs, list_idx, list_job, items, counter));
    PM_DISPATCH_PREC(precision, P_,
                     hipLaunchKernelGGL((k_consist_select_batch<P, false, Prec>), grid, dim3(256),
                                        0, s, jobs, count, list_idx, list_med, list_job, res,
                                        salt, counter, items));
  }
  hipLaunchKernelGGL(k_consist_apply_batch, grid, dim3(256), 0, s, jobs, count, list_idx, list_job,
                     res);
  return hipGetLastError();
}

hipError_t launch_pack(const PmArgs& a, int median, float* depth, float* plane,
                       float* score, int* nghbr, hipStream_t s) {
  hipLaunchKernelGGL(k_pack, grid16(a.W, a.H), dim3(256), 0, s, a, median,
                     depth, plane, score, nghbr);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<size_t>((n + 255) / 256, 65536)), dim3(256),
                     0, s, p, n, v);
  return hipGetLastError();
}

// Resident workgroups of kernel `f` with `threads` lanes on the current
// device (the persistent grid), cached per (device, kernel) so that launches
// under graph capture make no device queries after the first eager run.
template <typename F>
static hipError_t persistent_grid(F f, int threads, int* grid) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, int> cache;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const auto key = std::make_pair(dev, reinterpret_cast<const void*>(f));
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      *grid = it->second;
      return hipSuccess;
    }
  }
  int per_cu = 0, cus = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, threads, 0);
  if (e != hipSuccess) return e;
  *grid = std::max(1, per_cu) * cus;
  std::lock_guard<std::mutex> lock(mu);
  cache[key] = *grid;
  return hipSuccess;
}

template <int P, bool kSample, int Prec>
static hipError_t init_batch(const PmJob* jobs, int njobs, int total,
                             unsigned* queue, unsigned long long* count,
                             hipStream_t s) {
  auto f = k_pm_init_batch<P, kSample, Prec>;
  int grid = 0;
  hipError_t e = persistent_grid(f, 256, &grid);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(f, dim3(std::min(grid, total)), dim3(256), 0, s, jobs,
                     njobs, total, queue, count);
  return hipGetLastError();
}

hipError_t launch_pm_init_batch(const PmJob* jobs, int njobs, int total,
                                int patch, int precision, bool sample,
                                unsigned* queue, unsigned long long* count,
                                hipStream_t s) {
  if (total <= 0) return hipSuccess;
  if (sample) {
    PM_DISPATCH_PREC(precision, patch,
                     return (init_batch<P, true, Prec>(jobs, njobs, total, queue, count, s)));
  }
  PM_DISPATCH_PREC(precision, patch,
                   return (init_batch<P, false, Prec>(jobs, njobs, total, queue, count, s)));
  return hipSuccess;
}

template <int P, bool kSample, int Prec>
static hipError_t half_batch_tiled(int tile, const PmJob* jobs, int njobs,
                                   int total, unsigned* queue, int colour,
                                   int iter, unsigned long long* count,
                                   hipStream_t s) {
  int grid = 0;
  hipError_t e = hipSuccess;
  switch (tile) {
#define PM_TILE_CASE(ID, TX, TY)                                              \
  case ID: {                                                                  \
    auto f = k_pm_half_batch<P, kSample, TX, TY, Prec>;                       \
    e = persistent_grid(f, TX * TY, &grid);                                   \
    if (e != hipSuccess) return e;                                            \
    grid = std::max(8, std::min(grid, (total + 7) / 8 * 8));                  \
    hipLaunchKernelGGL(f, dim3(grid), dim3(TX * TY), 0, s,                    \
                       jobs, njobs, total, queue, colour, iter, count);       \
  } break;
    PM_TILE_CASE(0, 8, 8)
    PM_TILE_CASE(1, 16, 16)
    PM_TILE_CASE(2, 32, 8)
#undef PM_TILE_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

void tile_shape(int tile, int* tx, int* ty) {
  static const int shapes[3][2] = {{8, 8}, {16, 16}, {32, 8}};
  const int k = (tile >= 0 && tile < 3) ? tile : 1;
  *tx = shapes[k][0];
  *ty = shapes[k][1];
}

hipError_t launch_pm_half_batch(const PmJob* jobs, int njobs, int total,
                                int patch, int precision, bool sample, int tile,
                                int colour, int iter, unsigned* queue,
                                unsigned long long* count, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  if (sample) {
    PM_DISPATCH_PREC(precision, patch,
                     return (half_batch_tiled<P, true, Prec>(tile, jobs, njobs, total, queue,
                                                             colour, iter, count, s)));
  }
  PM_DISPATCH_PREC(precision, patch,
                   return (half_batch_tiled<P, false, Prec>(tile, jobs, njobs, total, queue,
                                                            colour, iter, count, s)));
  return hipSuccess;
}

hipError_t launch_pack_batch(const PmJob* jobs, int njobs, int total,
                             int median, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_batch, dim3(total), dim3(256), 0, s, jobs, njobs,
                     median);
  return hipGetLastError();
}

// The cluster estimator's result layout (DepthmapResult, estimator.h:33-38)
// from a PatchMatch result, with the readback rules of estimator.cc:960-993:
// a pixel with a non-finite plane gets depth 0, normal 0 and cost 0; a
// depth outside [dmin, dmax] is zeroed with its normal; every other pixel
// keeps its depth and the unit normal plane / |plane| (camera frame, facing
// the camera; 0 when |plane| <= 1e-6).  cost = 1 - score (the NCC cost,
// opencl_kernels.h:460) wherever it is finite, else 0, whatever the depth;
// confidence = 1 / (1 + exp((cost - 0.5) * 5)) where depth > 0, else 0
// (estimator.cc:1022-1040).
__global__ __launch_bounds__(256) void k_cluster_out(const float* __restrict__ depth,
                                                     const float* __restrict__ plane,
                                                     const float* __restrict__ score, int n,
                                                     float dmin, float dmax,
                                                     float* __restrict__ out_depth,
                                                     float* __restrict__ normal,
                                                     float* __restrict__ cost,
                                                     float* __restrict__ conf) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float p0 = plane[3 * (size_t)i], p1 = plane[3 * (size_t)i + 1], p2 = plane[3 * (size_t)i + 2];
  const bool finite = __builtin_isfinite(p0) && __builtin_isfinite(p1) && __builtin_isfinite(p2);
  float d = depth[i];
  if (!finite || !__builtin_isfinite(d) || !(d > 0.0f && d >= dmin && d <= dmax)) d = 0.0f;
  const float nrm = sqrtf(p0 * p0 + p1 * p1 + p2 * p2);
  const bool unit = d > 0.0f && nrm > 1e-6f;
  out_depth[i] = d;
  normal[3 * (size_t)i] = unit ? p0 / nrm : 0.0f;
  normal[3 * (size_t)i + 1] = unit ? p1 / nrm : 0.0f;
  normal[3 * (size_t)i + 2] = unit ? p2 / nrm : 0.0f;
  const float c = 1.0f - score[i];
  const float cf = finite && __builtin_isfinite(c) ? c : 0.0f;
  cost[i] = cf;
  conf[i] = d > 0.0f ? 1.0f / (1.0f + expf((cf - 0.5f) * 5.0f)) : 0.0f;
}

hipError_t launch_cluster_out(const float* depth, const float* plane, const float* score, int n,
                              float dmin, float dmax, float* out_depth, float* normal,
                              float* cost, float* conf, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cluster_out, dim3((n + 255) / 256), dim3(256), 0, s, depth, plane, score,
                     n, dmin, dmax, out_depth, normal, cost, conf);
  return hipGetLastError();
}

hipError_t launch_clean_batch(const ClJob* jobs, int njobs, int total,
                              const ClView* views, float thr, int min_views,
                              hipStream_t s) {
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_clean_batch, dim3(total), dim3(256), 0, s, jobs, njobs,
                     views, thr, min_views);
  return hipGetLastError();
}

hipError_t launch_resize_linear(const float* src, int sw, int sh, float* dst,
                                int w, int h, hipStream_t s) {
  dim3 blk(16, 16);
  dim3 grd((w + 15) / 16, (h + 15) / 16);
  hipLaunchKernelGGL(k_resize_linear, grd, blk, 0, s, src, sw, sh, dst, w, h);
  return hipGetLastError();
}

hipError_t launch_gpu_clean(const float* ref_depth, const float* const* src,
                            const GcCamera* cams, int w, int h, int nsrc,
                            float thr, int min_views, float carving,
                            int max_carved, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_gpu_clean, grid16(w, h), dim3(256), 0, s, ref_depth, src, cams, w, h,
                     nsrc, thr, min_views, carving, max_carved, out);
  return hipGetLastError();
}

hipError_t launch_prune_count(const PrView* views, int n, const PrRef& ref,
                              float thr, int npix, int* counts, hipStream_t s) {
  hipLaunchKernelGGL(k_prune_count, dim3((npix + 255) / 256), dim3(256), 0, s, views, n, ref,
                     thr, npix, counts);
  return hipGetLastError();
}

hipError_t launch_prune_write(const PrView* views, int n, const PrRef& ref,
                              float thr, int npix, const long long* offsets,
                              const uint8_t* color, const uint8_t* label,
                              float* points, float* normals, uint8_t* out_color,
                              uint8_t* out_label, hipStream_t s) {
  hipLaunchKernelGGL(k_prune_write, dim3((npix + 255) / 256), dim3(256), 0, s, views, n, ref,
                     thr, npix, offsets, color, label, points, normals, out_color, out_label);
  return hipGetLastError();
}

hipError_t launch_clean(const ClView* views, const ClRef* ref, int n, int W,
                        int H, float thr, int min_views, float* out,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_clean, grid16(W, H), dim3(256), 0, s, views, ref, n, thr,
                     min_views, out);
  return hipGetLastError();
}

}  // namespace pm
