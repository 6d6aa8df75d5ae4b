// Device-side building blocks of the PatchMatch depthmap engine (gfx950).
//
// The whole library is compiled with -ffp-contract=off, so plain C++
// arithmetic below follows the written operation order exactly (this is what
// the "exact" scorer relies on for parity with the reference's FP32 code);
// the "fast" scorer asks for FMAs explicitly with __builtin_fmaf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_types.h"

namespace pm {

// ---------------------------------------------------------------------------
// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Each pixel owns a
// key; the counter encodes (iteration, half-iteration, step).

struct u4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u4 philox4x32_10(uint32_t k0, uint32_t k1, u4 c) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t m0 = 0xD2511F53u, m1 = 0xCD9E8D57u;
    uint32_t hi0 = __umulhi(m0, c.x), lo0 = m0 * c.x;
    uint32_t hi1 = __umulhi(m1, c.z), lo1 = m1 * c.z;
    c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// [0, 1) with 24 random bits
__device__ __forceinline__ float u01(uint32_t x) {
  return (float)(x >> 8) * (1.0f / 16777216.0f);
}
// (0, 1]
__device__ __forceinline__ float u01_open0(uint32_t x) {
  return ((float)(x >> 8) + 1.0f) * (1.0f / 16777216.0f);
}

// Two standard normals from two uniforms (Box-Muller).
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float* n0,
                                           float* n1) {
  float r = sqrtf(-2.0f * __logf(u01_open0(a)));
  float s, c;
  __sincosf(6.28318530717958647692f * u01(b), &s, &c);
  *n0 = r * c;
  *n1 = r * s;
}

// ---------------------------------------------------------------------------
// Packed bilinear texels.  View images are stored as "quads": texel (ix, iy)
// packs the four u8 neighbours im[iy][ix], im[iy][ix+1], im[iy+1][ix],
// im[iy+1][ix+1] into one u32, so one 4-byte gather feeds one bilinear
// sample.  Texels exist for ix in [-1, W-1] and iy in [-1, H-1] (pitch W+1,
// one zero column/row in front); only 0 <= ix <= W-2, 0 <= iy <= H-2 hold
// image data, all others are zero, so a clamped out-of-image sample reads 0.

__device__ __forceinline__ float ub0(uint32_t v) { return (float)(v & 0xffu); }
__device__ __forceinline__ float ub1(uint32_t v) { return (float)((v >> 8) & 0xffu); }
__device__ __forceinline__ float ub2(uint32_t v) { return (float)((v >> 16) & 0xffu); }
__device__ __forceinline__ float ub3(uint32_t v) { return (float)(v >> 24); }

// ---------------------------------------------------------------------------
// DepthOfPlaneBackprojection (depthmap.cc:114-118) in FP64 with the
// reference's order: r = plane^T K^-1, s = r . (x, y, 1).
__device__ __forceinline__ float depth_of_plane_f64(const double* Kinv,
                                                    float p0, float p1,
                                                    float p2, double x,
                                                    double y) {
  const double p[3] = {(double)p0, (double)p1, (double)p2};
  double r[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double s = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += p[k] * Kinv[k * 3 + j];
    r[j] = s;
  }
  double s = 0;
  s += r[0] * x;
  s += r[1] * y;
  s += r[2] * 1.0;
  float denom = (float)(-s);
  float m = (1e-6f < denom) ? denom : 1e-6f;
  return __fdiv_rn(1.0f, m);
}

// ---------------------------------------------------------------------------
// Reference-side patch data: hypothesis-invariant per pixel.  w[k] is the
// bilateral weight of sample k (from the host-tabulated table, bit-identical
// to expf of the reference's argument) and wx[k] = w[k] * x[k], the same
// FP32 product the reference forms in NCCEstimator::Push.  The reference
// accumulates Sum w x, Sum w x x and Sum w in sample order; those sums do
// not depend on the hypothesis and are formed once.

//
// The fast scorer's values are centred instead: x relative to c, the
// weighted mean of the reference patch (load_ref_patch<P, true>), and the
// samples y of each hypothesis relative to its first sample, so that NCC's
// un-centred FP32 moments (Sum w y y / Sum w - mean_y^2) do not cancel on
// bright, nearly flat patches.  With the reference's raw values such a patch
// can score well above 1 (1.08 measured on street facades), a rounding
// artefact the search then prefers.  NCC is invariant to both shifts.

template <int P>
struct RefPatch {
  static constexpr int N = P * P;
  float w[N];
  float wx[N];              // w[k] * x[k], the FP32 product the reference forms
                            // (centred: w[k] * (x[k] - c))
  float sumw, meanx, varx;  // exact-order FP32 (centred: of x - c)
  float rsumw;              // 1 / sumw for the fast path
  float c;                  // centre of the fast path's x values (0: raw)
};

// Bilateral weight table: wlut[|dcolor| * nr2 + (dx*dx + dy*dy)], nr2 =
// 2*hpz*hpz + 1.
template <int P, bool kCentred = false>
__device__ __forceinline__ bool load_ref_patch(RefPatch<P>& rp,
                                               const uint8_t* __restrict__ img,
                                               int W, int i, int j,
                                               const float* __restrict__ wlut) {
  constexpr int hpz = (P - 1) / 2;
  constexpr int nr2 = 2 * hpz * hpz + 1;
  const int c = img[i * W + j];
  if constexpr (kCentred) {
    float sw = 0, swx = 0;
#pragma unroll
    for (int dy = -hpz; dy <= hpz; ++dy) {
#pragma unroll
      for (int dx = -hpz; dx <= hpz; ++dx) {
        const int k = (dy + hpz) * P + (dx + hpz);
        const int x = img[(i + dy) * W + (j + dx)];
        const int dc = x - c;
        const float w = wlut[(dc < 0 ? -dc : dc) * nr2 + dx * dx + dy * dy];
        rp.w[k] = w;
        sw += w;
        swx = __builtin_fmaf(w, (float)x, swx);
      }
    }
    const float rsw = __frcp_rn(sw);
    const float cen = swx * rsw;
    float sx = 0, sxx = 0;
#pragma unroll
    for (int dy = -hpz; dy <= hpz; ++dy) {
#pragma unroll
      for (int dx = -hpz; dx <= hpz; ++dx) {
        const int k = (dy + hpz) * P + (dx + hpz);
        const float xs = (float)img[(i + dy) * W + (j + dx)] - cen;
        const float wx = rp.w[k] * xs;
        rp.wx[k] = wx;
        sx += wx;
        sxx = __builtin_fmaf(wx, xs, sxx);
      }
    }
    rp.sumw = sw;
    rp.rsumw = rsw;
    rp.c = cen;
    rp.meanx = sx * rsw;
    rp.varx = __builtin_fmaf(-rp.meanx, rp.meanx, sxx * rsw);
    // NCCEstimator::Get returns -1 for every hypothesis when var_x < 0.1.
    return !(rp.varx < 0.1f);
  }
  float sumx = 0, sumxx = 0, sumw = 0;
#pragma unroll
  for (int dy = -hpz; dy <= hpz; ++dy) {
#pragma unroll
    for (int dx = -hpz; dx <= hpz; ++dx) {
      const int k = (dy + hpz) * P + (dx + hpz);
      const int x = img[(i + dy) * W + (j + dx)];
      const int dc = x - c;
      const float w = wlut[(dc < 0 ? -dc : dc) * nr2 + dx * dx + dy * dy];
      const float xf = (float)x;
      rp.w[k] = w;
      const float wx = w * xf;
      rp.wx[k] = wx;
      sumx += wx;
      sumxx += wx * xf;
      sumw += w;
    }
  }
  rp.sumw = sumw;
  rp.rsumw = __frcp_rn(sumw);
  const float meanx = __fdiv_rn(sumx, sumw);
  const float meanxx = __fdiv_rn(sumxx, sumw);
  rp.meanx = meanx;
  rp.varx = meanxx - meanx * meanx;
  rp.c = 0.0f;
  // NCCEstimator::Get returns -1 for every hypothesis when var_x < 0.1.
  return !(rp.varx < 0.1);
}

// ---------------------------------------------------------------------------
// Correctly rounded FP32 square root for normal positive x: the hardware
// estimate (within 1 ulp) corrected to the neighbouring float by the sign of
// the FMA residuals x - s_down * s and x - s_up * s.

__device__ __forceinline__ float sqrt_rn(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sdn = __int_as_float(__float_as_int(s) - 1);
  const float sup = __int_as_float(__float_as_int(s) + 1);
  const float vp = __builtin_fmaf(-sdn, s, x);
  const float vs = __builtin_fmaf(-sup, s, x);
  float r = (vp <= 0.0f) ? sdn : s;
  r = (vs > 0.0f) ? sup : r;
  return r;
}

// ---------------------------------------------------------------------------
// Texel gathers through a structured buffer resource: the record index (one
// VGPR per address) is scaled by the texel size in the address unit, and
// records outside [0, num_records) read as zero (the hardware range check, a
// second line of defence behind the coordinate clamps).  Against a byte
// offset this saves the shift of every gather's address (measured 0.7 % on
// the red/black kernel).

typedef unsigned int pm_u32x2 __attribute__((ext_vector_type(2)));
// The LLVM buffer intrinsics with an index operand (buffer_load ... idxen),
// for which clang has no builtin.
__device__ unsigned int pm_struct_load_b32(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset,
                                           int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.i32");
__device__ pm_u32x2 pm_struct_load_b64(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset,
                                       int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v2i32");

struct QuadTex {
  __amdgpu_buffer_rsrc_t rsrc;  // base = texel (0, 0)
  int pitch;                    // texels per row (W + 1)
  float xmax, ymax;             // W - 1, H - 1: largest texel index (zero)
};

// Wave-uniform copies (scalar registers) of values the caller knows are
// uniform, e.g. fields of a view selected by a uniform loop index.
__device__ __forceinline__ int uniform(int x) {
  return __builtin_amdgcn_readfirstlane(x);
}
__device__ __forceinline__ float uniform(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
template <typename T>
__device__ __forceinline__ T* uniform(T* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (T*)(((uint64_t)hi << 32) | lo);
}

// Texture of a view whose index is uniform across the wavefront.
__device__ __forceinline__ QuadTex make_tex(const uint32_t* quad, int W, int H) {
  QuadTex t;
  quad = uniform(quad);
  W = uniform(W);
  H = uniform(H);
  t.pitch = W + 1;
  // records from texel (0, 0) to the end; texel (-1, -1) would be index
  // -(pitch + 1), an out-of-range record that reads 0 like the border
  const int records = t.pitch * (H + 1) - (t.pitch + 1);
  t.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)(quad + t.pitch + 1),
                                             (short)4, records, 0x00020000);
  t.xmax = (float)(W - 1);
  t.ymax = (float)(H - 1);
  return t;
}

// Texel whose top-left corner is (floor x, floor y), clamped into the zero
// border: a sample outside [0, W-1) x [0, H-1) reads four zeros.  Texel
// (ix, iy) is record iy * pitch + ix; row -1 gives negative (out-of-range)
// record indices, which the buffer unit returns as 0.
__device__ __forceinline__ uint32_t fetch_quad(const QuadTex& t, float flx,
                                               float fly) {
  const int ix = (int)__builtin_amdgcn_fmed3f(flx, -1.0f, t.xmax);
  const int iy = (int)__builtin_amdgcn_fmed3f(fly, -1.0f, t.ymax);
  const int rec = __mul24(iy, t.pitch) + ix;
  return pm_struct_load_b32(t.rsrc, rec, 0, 0, 0);
}

// floor(x) as an int in one instruction (x within the int range).
__device__ __forceinline__ int cvt_floor(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// The texel of fetch_quad from unfloored coordinates: clamping x into
// [-1, xmax] and then flooring equals flooring and then clamping, since the
// bounds are integers.
__device__ __forceinline__ uint32_t fetch_quad_xy(const QuadTex& t, float x, float y) {
  const int ix = cvt_floor(__builtin_amdgcn_fmed3f(x, -1.0f, t.xmax));
  const int iy = cvt_floor(__builtin_amdgcn_fmed3f(y, -1.0f, t.ymax));
  const int rec = __mul24(iy, t.pitch) + ix;
  return pm_struct_load_b32(t.rsrc, rec, 0, 0, 0);
}

// Texture of 8-byte quads (FP16 or BF16 values) of a wave-uniform view.
__device__ __forceinline__ QuadTex make_tex16(const uint32_t* quad16, int W,
                                              int H) {
  QuadTex t;
  quad16 = uniform(quad16);
  W = uniform(W);
  H = uniform(H);
  t.pitch = W + 1;
  const int records = t.pitch * (H + 1) - (t.pitch + 1);
  t.rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(quad16 + 2 * (t.pitch + 1)), (short)8, records, 0x00020000);
  t.xmax = (float)(W - 1);
  t.ymax = (float)(H - 1);
  return t;
}

// The same for the 8-byte texels of a make_tex16 texture.
__device__ __forceinline__ uint2 fetch_quad16_xy(const QuadTex& t, float x, float y) {
  const int ix = cvt_floor(__builtin_amdgcn_fmed3f(x, -1.0f, t.xmax));
  const int iy = cvt_floor(__builtin_amdgcn_fmed3f(y, -1.0f, t.ymax));
  const int rec = __mul24(iy, t.pitch) + ix;
  const pm_u32x2 v = pm_struct_load_b64(t.rsrc, rec, 0, 0, 0);
  return make_uint2(v[0], v[1]);
}

// A texture whose view may differ between the lanes of a wave (the lane
// groups of the red/black kernel, where lane k of a group scores its own
// views): per-lane base pointer and size, read with global loads.  Texel
// (ix, iy) is the record of make_tex / make_tex16 at the same clamped
// indices, and a negative record -- row -1, or column -1 of row 0, which the
// buffer unit returns as 0 -- reads 0 here too; every non-negative record lies
// inside the texture (at most (H - 1) pitch + W - 1 < pitch (H + 1) -
// (pitch + 1) records).
struct LaneTex {
  const char* base;  // texel (0, 0)
  int pitch;
  float xmax, ymax;
};

template <int kWords>
__device__ __forceinline__ LaneTex make_lane_tex(const uint32_t* quad, int W, int H) {
  LaneTex t;
  t.pitch = W + 1;
  t.base = reinterpret_cast<const char*>(quad + kWords * (t.pitch + 1));
  t.xmax = (float)(W - 1);
  t.ymax = (float)(H - 1);
  return t;
}

template <typename T>
__device__ __forceinline__ T fetch_lane_xy(const LaneTex& t, float x, float y) {
  const int ix = cvt_floor(__builtin_amdgcn_fmed3f(x, -1.0f, t.xmax));
  const int iy = cvt_floor(__builtin_amdgcn_fmed3f(y, -1.0f, t.ymax));
  const int rec = __mul24(iy, t.pitch) + ix;
  const T v = *reinterpret_cast<const T*>(t.base + (size_t)max(rec, 0) * sizeof(T));
  return rec < 0 ? T{} : v;
}

// Texel sources for the fast scorer: the texel holding (floor x, floor y)
// from the unfloored coordinates, and its bilinear blend at fractions
// (fx, fy).

// FP32 from the u8 quads, blended in the coefficient form of CoefFetch
// (below) with the coefficients formed exactly in integers: the same
// samples, bit for bit, from 4-byte texels.
struct QuadFetch {
  QuadTex tex;
  __device__ __forceinline__ uint32_t operator()(float x, float y) const {
    return fetch_quad_xy(tex, x, y);
  }
  static __device__ __forceinline__ float bilerp(uint32_t t, float fx, float fy) {
    const int a = (int)(t & 0xffu), b = (int)((t >> 8) & 0xffu);
    const int c = (int)((t >> 16) & 0xffu), d = (int)(t >> 24);
    const float top = __builtin_fmaf(fx, (float)(b - a), (float)a);
    const float dif = __builtin_fmaf(fx, (float)(d - c - b + a), (float)(c - a));
    return __builtin_fmaf(fy, dif, top);
  }
};

typedef _Float16 pm_half2 __attribute__((ext_vector_type(2)));
typedef __bf16 pm_bf162 __attribute__((ext_vector_type(2)));

// FP32 from bilinear coefficients: the quad (a, b; c, d) is stored as
// A = a, B = b - a, C = c - a, D = d - c - b + a in FP16 (integers of at most
// 510 in magnitude, exact), so that the blend
//   A + fx B + fy (C + fx D)
// is two mixed-precision FMAs (v_fma_mix_f32 reads the FP16 halves
// directly) and one FP32 FMA.  A + fx B is the top row's blend with the
// same single rounding as fma(fx, b - a, a); C + fx D is the difference of
// the two rows' blends with one rounding instead of three.
struct CoefFetch {
  QuadTex tex;
  __device__ __forceinline__ uint2 operator()(float x, float y) const {
    return fetch_quad16_xy(tex, x, y);
  }
  static __device__ __forceinline__ float bilerp(uint2 t, float fx, float fy) {
    const pm_half2 ab = __builtin_bit_cast(pm_half2, t.x);
    const pm_half2 cd = __builtin_bit_cast(pm_half2, t.y);
    const float top = __builtin_fmaf(fx, (float)ab[1], (float)ab[0]);
    const float dif = __builtin_fmaf(fx, (float)cd[1], (float)cd[0]);
    return __builtin_fmaf(fy, dif, top);
  }
};

// CoefFetch and QuadFetch over a LaneTex (per-lane views).
struct CoefLaneFetch {
  LaneTex tex;
  __device__ __forceinline__ uint2 operator()(float x, float y) const {
    return fetch_lane_xy<uint2>(tex, x, y);
  }
  static __device__ __forceinline__ float bilerp(uint2 t, float fx, float fy) {
    return CoefFetch::bilerp(t, fx, fy);
  }
};
struct QuadLaneFetch {
  LaneTex tex;
  __device__ __forceinline__ uint32_t operator()(float x, float y) const {
    return fetch_lane_xy<uint32_t>(tex, x, y);
  }
  static __device__ __forceinline__ float bilerp(uint32_t t, float fx, float fy) {
    return QuadFetch::bilerp(t, fx, fy);
  }
};

// FP16 / BF16: the quad holds (a, b) and (c, d) as 16-bit pairs; each row
// is blended by one dot product with the weights (1 - fx, fx) rounded to
// the same 16-bit format and accumulated in FP32 (v_dot2_f32_f16 /
// v_dot2_f32_bf16), the column blend is FP32.
template <bool kBF16>
struct Global16Fetch {
  QuadTex tex;
  __device__ __forceinline__ uint2 operator()(float x, float y) const {
    return fetch_quad16_xy(tex, x, y);
  }
  static __device__ __forceinline__ float bilerp(uint2 t, float fx, float fy) {
    float im0, im1;
    if constexpr (kBF16) {
      pm_bf162 w;
      w[0] = (__bf16)(1.0f - fx);
      w[1] = (__bf16)fx;
      im0 = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(pm_bf162, t.x), w, 0.0f, false);
      im1 = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(pm_bf162, t.y), w, 0.0f, false);
    } else {
      const pm_half2 w =
          __builtin_bit_cast(pm_half2, __builtin_amdgcn_cvt_pkrtz(1.0f - fx, fx));
      im0 = __builtin_amdgcn_fdot2(__builtin_bit_cast(pm_half2, t.x), w, 0.0f, false);
      im1 = __builtin_amdgcn_fdot2(__builtin_bit_cast(pm_half2, t.y), w, 0.0f, false);
    }
    return __builtin_fmaf(fy, im1 - im0, im0);
  }
};

// Local-affine warp of a reference pixel into one view: the pixel's image
// (Hx0, Hy0) and the Jacobian columns (dfdx_*, dfdy_*).
struct Warp {
  float Hx0, Hy0, dfdx_x, dfdx_y, dfdy_x, dfdy_y;
};

// One patch row of gathered texels and their bilinear fractions.
template <int P>
struct RowFetch {
  uint32_t t[P];
  float fx[P], fy[P];
};

// ---------------------------------------------------------------------------
// Exact scorer: ComputePlaneImageScore (depthmap.cc:441-478) with the
// reference's FP32 operation order, given the FP32 homography H.
// LinearInterpolation (depthmap.cc:14-30) returns 0 outside
// [0, W-1) x [0, H-1): here the clamped fetch reads a zero texel there, the
// bilinear blend of zeros is +0, and a non-finite coordinate (NaN blend) is
// mapped to 0 as well.  Inside, floor(x) == (int)x and x - floor(x) is the
// reference's dx bit for bit.

template <int P>
__device__ __forceinline__ float score_exact(const RefPatch<P>& rp,
                                             const float* H, int i, int j,
                                             const QuadTex& tex) {
  constexpr int hpz = (P - 1) / 2;
  const float fj = (float)j, fi = (float)i;
  const float u = H[0] * fj + H[1] * fi + H[2];
  const float v = H[3] * fj + H[4] * fi + H[5];
  const float w = H[6] * fj + H[7] * fi + H[8];
  if (w == 0.0f) return -1.0f;
  const float ww = w * w;
  const float dfdx_x = __fdiv_rn(H[0] * w - H[6] * u, ww);
  const float dfdx_y = __fdiv_rn(H[3] * w - H[6] * v, ww);
  const float dfdy_x = __fdiv_rn(H[1] * w - H[7] * u, ww);
  const float dfdy_y = __fdiv_rn(H[4] * w - H[7] * v, ww);
  const float Hx0 = __fdiv_rn(u, w);
  const float Hy0 = __fdiv_rn(v, w);
  // A non-finite warp (a NaN or infinite plane, or an overflow) scores -1;
  // the reference's LinearInterpolation reads all zeros for infinite
  // coordinates (-1 again, var_y < 0.1) and is undefined for NaN ones.
  if (!__builtin_isfinite(Hx0 + Hy0 + dfdx_x + dfdx_y + dfdy_x + dfdy_y)) return -1.0f;

  float cx[P], cy[P];
#pragma unroll
  for (int dx = -hpz; dx <= hpz; ++dx) {
    cx[dx + hpz] = Hx0 + dfdx_x * (float)dx;
    cy[dx + hpz] = Hy0 + dfdx_y * (float)dx;
  }
  // Two-stage row pipeline: the gathers of patch row r+1 are issued before
  // row r is blended and accumulated (two rows of texels and fractions
  // live), so one row's loads are in flight while the previous computes.
  RowFetch<P> buf[2];
  auto fetch = [&](RowFetch<P>& f, int dy) {
    const float rx = dfdy_x * (float)dy;
    const float ry = dfdy_y * (float)dy;
#pragma unroll
    for (int dx = -hpz; dx <= hpz; ++dx) {
      const float x2 = cx[dx + hpz] + rx;
      const float y2 = cy[dx + hpz] + ry;
      const float flx = floorf(x2), fly = floorf(y2);
      f.t[dx + hpz] = fetch_quad(tex, flx, fly);
      f.fx[dx + hpz] = x2 - flx;
      f.fy[dx + hpz] = y2 - fly;
    }
  };
  float sumy = 0, sumyy = 0, sumxy = 0;
  fetch(buf[0], -hpz);
#pragma unroll
  for (int r = 0; r < P; ++r) {
    if (r + 1 < P) fetch(buf[(r + 1) & 1], r + 1 - hpz);
    __builtin_amdgcn_sched_barrier(0);
    const RowFetch<P>& f = buf[r & 1];
#pragma unroll
    for (int c = 0; c < P; ++c) {
      const int k = r * P + c;
      const uint32_t t = f.t[c];
      const float fx = f.fx[c], fy = f.fy[c];
      const float im0 = (1.0f - fx) * ub0(t) + fx * ub1(t);
      const float im1 = (1.0f - fx) * ub2(t) + fx * ub3(t);
      float y = (1.0f - fy) * im0 + fy * im1;
      y = (y == y) ? y : 0.0f;
      const float wy = rp.w[k] * y;
      sumy += wy;
      sumyy += wy * y;
      sumxy += rp.wx[k] * y;
    }
    // the row's sums materialised here (otherwise the optimiser sinks the
    // accumulation chains past the last row and keeps every sample live)
    asm volatile("" : "+v"(sumy), "+v"(sumyy), "+v"(sumxy));
    __builtin_amdgcn_sched_barrier(0);
  }
  // NCCEstimator::Get (depthmap.cc:58-74)
  const float meany = __fdiv_rn(sumy, rp.sumw);
  const float meanyy = __fdiv_rn(sumyy, rp.sumw);
  const float meanxy = __fdiv_rn(sumxy, rp.sumw);
  const float vary = meanyy - meany * meany;
  if (rp.varx < 0.1 || vary < 0.1) return -1.0f;
  // var_x, var_y >= 0.1 here, so the product is a normal float
  return __fdiv_rn(meanxy - rp.meanx * meany, sqrt_rn(rp.varx * vary));
}

// ---------------------------------------------------------------------------
// Fast scorer for PatchMatch: the same local-affine bilateral NCC, with FMAs
// and one-ulp reciprocals.  Takes the pixel's warp and a texel source
// (CoefFetch, QuadFetch or Global16Fetch: the texel of floor coordinates and its
// bilinear blend); the NCC sums are FP32 for every texel format.
// A non-finite coordinate yields a NaN score,
// which never wins a strict comparison.

template <int P, class Fetch>
__device__ __forceinline__ float score_fast_loop(const RefPatch<P>& rp, const Warp& wp,
                                                 const Fetch& fetch_texel) {
  constexpr int hpz = (P - 1) / 2;
  const float Hx0 = wp.Hx0, Hy0 = wp.Hy0;
  const float dfdx_x = wp.dfdx_x, dfdx_y = wp.dfdx_y;
  const float dfdy_x = wp.dfdy_x, dfdy_y = wp.dfdy_y;
  // One patch row of gathers in flight at a time (score_fast_pipe keeps two,
  // for kernels with the registers to spare).
  float sumy = 0, sumyy = 0, sumxy = 0, shift = 0;
#pragma unroll
  for (int dy = -hpz; dy <= hpz; ++dy) {
    const float rx = __builtin_fmaf(dfdy_x, (float)dy, Hx0);
    const float ry = __builtin_fmaf(dfdy_y, (float)dy, Hy0);
#pragma unroll
    for (int dx = -hpz; dx <= hpz; ++dx) {
      const int k = (dy + hpz) * P + (dx + hpz);
      const float x2 = __builtin_fmaf(dfdx_x, (float)dx, rx);
      const float y2 = __builtin_fmaf(dfdx_y, (float)dx, ry);
      const auto t = fetch_texel(x2, y2);
      const float yr = Fetch::bilerp(t, __builtin_amdgcn_fractf(x2), __builtin_amdgcn_fractf(y2));
      if (k == 0) shift = yr;  // the first sample centres the others
      const float y = yr - shift;
      const float wy = rp.w[k] * y;
      sumy += wy;
      sumyy = __builtin_fmaf(wy, y, sumyy);
      sumxy = __builtin_fmaf(rp.wx[k], y, sumxy);
    }
    // the row's sums materialised here (otherwise the optimiser sinks the
    // accumulation chains past the last row and keeps every sample live)
    asm volatile("" : "+v"(sumy), "+v"(sumyy), "+v"(sumxy));
    __builtin_amdgcn_sched_barrier(0);
  }
  const float meany = sumy * rp.rsumw;
  const float meanyy = sumyy * rp.rsumw;
  const float meanxy = sumxy * rp.rsumw;
  const float vary = __builtin_fmaf(-meany, meany, meanyy);
  if (vary < 0.1f) return -1.0f;
  const float cov = __builtin_fmaf(-rp.meanx, meany, meanxy);
  return cov * rsqrtf(rp.varx * vary);
}

// D rows of gathers in flight: the gathers of patch row r + D - 1 are issued
// before row r is blended (the texels of the rows ahead, and the sample
// coordinates their fractions come from, stay live: about 4 P registers per
// extra row).  The same arithmetic in the same order, so the same score bit
// for bit.
template <int P, int D, class Fetch>
__device__ __forceinline__ float score_fast_pipe(const RefPatch<P>& rp, const Warp& wp,
                                                 const Fetch& fetch_texel) {
  constexpr int hpz = (P - 1) / 2;
  using Tex = decltype(fetch_texel(0.0f, 0.0f));
  struct Row {
    Tex t[P];
    float x[P], y[P];
  };
  const float Hx0 = wp.Hx0, Hy0 = wp.Hy0;
  const float dfdx_x = wp.dfdx_x, dfdx_y = wp.dfdx_y;
  const float dfdy_x = wp.dfdy_x, dfdy_y = wp.dfdy_y;
  Row buf[D];
  auto fetch = [&](Row& f, int dy) {
    const float rx = __builtin_fmaf(dfdy_x, (float)dy, Hx0);
    const float ry = __builtin_fmaf(dfdy_y, (float)dy, Hy0);
#pragma unroll
    for (int dx = -hpz; dx <= hpz; ++dx) {
      const float x2 = __builtin_fmaf(dfdx_x, (float)dx, rx);
      const float y2 = __builtin_fmaf(dfdx_y, (float)dx, ry);
      f.t[dx + hpz] = fetch_texel(x2, y2);
      f.x[dx + hpz] = x2;
      f.y[dx + hpz] = y2;
    }
  };
  float sumy = 0, sumyy = 0, sumxy = 0, shift = 0;
#pragma unroll
  for (int r = 0; r < D - 1; ++r) fetch(buf[r], r - hpz);
#pragma unroll
  for (int r = 0; r < P; ++r) {
    if (r + D - 1 < P) fetch(buf[(r + D - 1) % D], r + D - 1 - hpz);
    __builtin_amdgcn_sched_barrier(0);
    const Row& f = buf[r % D];
#pragma unroll
    for (int c = 0; c < P; ++c) {
      const int k = r * P + c;
      const float yr = Fetch::bilerp(f.t[c], __builtin_amdgcn_fractf(f.x[c]),
                                     __builtin_amdgcn_fractf(f.y[c]));
      if (k == 0) shift = yr;
      const float y = yr - shift;
      const float wy = rp.w[k] * y;
      sumy += wy;
      sumyy = __builtin_fmaf(wy, y, sumyy);
      sumxy = __builtin_fmaf(rp.wx[k], y, sumxy);
    }
    asm volatile("" : "+v"(sumy), "+v"(sumyy), "+v"(sumxy));
    __builtin_amdgcn_sched_barrier(0);
  }
  const float meany = sumy * rp.rsumw;
  const float meanyy = sumyy * rp.rsumw;
  const float meanxy = sumxy * rp.rsumw;
  const float vary = __builtin_fmaf(-meany, meany, meanyy);
  if (vary < 0.1f) return -1.0f;
  const float cov = __builtin_fmaf(-rp.meanx, meany, meanxy);
  return cov * rsqrtf(rp.varx * vary);
}

// kPipe: rows of gathers in flight (<= 1: one row at a time).
template <int P, int kPipe = 1, class Fetch>
__device__ __forceinline__ float score_fast_affine(const RefPatch<P>& rp,
                                                   const Warp& wp,
                                                   const Fetch& fetch_texel) {
  if constexpr (kPipe >= 2) return score_fast_pipe<P, kPipe>(rp, wp, fetch_texel);
  return score_fast_loop<P>(rp, wp, fetch_texel);
}

}  // namespace pm
