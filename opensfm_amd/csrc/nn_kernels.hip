// BF16 kernels of the learned PatchMatch initialiser (a small U-Net over the
// reference image and a coarse plane-sweep cost volume), forward and
// backward, for gfx950.
//
// Layouts: activations NHWC (channels innermost) with a channel stride, so
// that a skip connection and an upsampled tensor share one concatenation
// buffer without copies.  Convolution weights are [Cout][Kp] with
// k = tap * Cin + ci (tap = 3 * (dy + 1) + (dx + 1)), zero padded to Kp, a
// multiple of 32.  Activations are BF16, gradients FP32, weights FP32
// masters with BF16 copies.
//
// Every 3x3 convolution is an implicit GEMM on the MFMA unit
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation): M = pixels, N = output
// channels, K = 9 * input channels.  The input gradient is the same
// convolution of the output gradient with the weights flipped in space and
// transposed in channels; the weight gradient is a GEMM over pixels
// (M = Cout, N = K, K = pixels) reduced across workgroups with atomics.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>

#include "nn_kernels.h"

namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int e = 0; e < 8; ++e) z[e] = (__bf16)0.0f;
  return z;
}

// 8 consecutive channels as BF16 (16-byte aligned: channel strides and
// offsets are multiples of 8).
__device__ __forceinline__ bf16x8 load8(const __bf16* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}
__device__ __forceinline__ bf16x8 load8(const float* p) {
  const float4 u = reinterpret_cast<const float4*>(p)[0];
  const float4 v = reinterpret_cast<const float4*>(p)[1];
  bf16x8 r;
  r[0] = (__bf16)u.x;
  r[1] = (__bf16)u.y;
  r[2] = (__bf16)u.z;
  r[3] = (__bf16)u.w;
  r[4] = (__bf16)v.x;
  r[5] = (__bf16)v.y;
  r[6] = (__bf16)v.z;
  r[7] = (__bf16)v.w;
  return r;
}

__device__ __forceinline__ float lrelu(float v) { return v > 0.0f ? v : 0.1f * v; }

// ---------------------------------------------------------------------------
// 3x3 convolution, stride 1, zero padding 1.  One wave computes 16 pixels x
// (NT * 16) output channels; a workgroup holds 4 waves (64 pixels).
//   x: [npix][xs] (TI = BF16 activations, or FP32 gradients converted on
//      load), channels [0, cin) read; images of H x W stacked.
//   w: [cout rows][kp] BF16.
//   y: [npix][ys] at channel offset yo, BF16 or FP32 (y_f32); with
//      accumulate (FP32 only) the result is added to y.
//   act: optional BF16 lrelu(y + bias) at [npix][as] offset ao.
template <typename TI, int NT>
__global__ __launch_bounds__(256) void k_conv3x3(
    const TI* __restrict__ x, int npix, int H, int W, int cin, int xs,
    const __bf16* __restrict__ w, int cout, int kp,
    const float* __restrict__ bias, void* __restrict__ y, int y_f32, int ys,
    int yo, int accumulate, __bf16* __restrict__ act, int as, int ao) {
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile * 16 >= npix) return;  // whole wave
  const int g = lane >> 4, c = lane & 15;
  const int co0 = blockIdx.y * NT * 16;
  const int p = tile * 16 + c;  // this lane's pixel in the A operand
  const bool pv = p < npix;
  const int hw = H * W;
  const int n = pv ? p / hw : 0;
  const int rem = p - n * hw;
  const int py = rem / W, px = rem - (rem / W) * W;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < kp; k0 += 32) {
    const int k = k0 + 8 * g;
    const int tap = k / cin;
    const int ci = k - tap * cin;
    bf16x8 a = zero8();
    if (pv && tap < 9) {
      const int yy = py + tap / 3 - 1, xx = px + tap % 3 - 1;
      if (yy >= 0 && yy < H && xx >= 0 && xx < W)
        a = load8(x + (size_t)(n * hw + yy * W + xx) * xs + ci);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int co = co0 + t * 16 + c;
      const bf16x8 b = co < cout ? load8(w + (size_t)co * kp + k) : zero8();
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
  }
  // D layout: lane (g, c) holds rows 4g..4g+3 (pixels) of column c (channel)
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = co0 + t * 16 + c;
    if (co >= cout) continue;
    const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = tile * 16 + 4 * g + r;
      if (q >= npix) continue;
      const float v = acc[t][r] + bv;
      const size_t o = (size_t)q * ys + yo + co;
      if (y) {
        if (y_f32) {
          float* yf = static_cast<float*>(y);
          yf[o] = accumulate ? yf[o] + v : v;
        } else {
          static_cast<__bf16*>(y)[o] = (__bf16)v;
        }
      }
      if (act) act[(size_t)q * as + ao + co] = (__bf16)lrelu(v);
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient: dw[co][k] += sum_p dz[p][co] * A[p][k], A the im2col of
// x.  Wave = one 16 x 16 (co, k) tile over a chunk of `chunk` pixels
// (multiple of 32); workgroup = 4 k tiles; grid (k tiles / 4, co tiles,
// pixel chunks).
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(
    const __bf16* __restrict__ x, int npix, int H, int W, int cin, int xs,
    const float* __restrict__ dz, int cout, int dzs, float* __restrict__ dw,
    int kp, int chunk) {
  const int lane = threadIdx.x & 63;
  const int kt = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (kt * 16 >= kp) return;  // whole wave
  const int g = lane >> 4, c = lane & 15;
  const int co0 = blockIdx.y * 16;
  const int k = kt * 16 + c;  // this lane's column (B operand)
  const int tap = k / cin;
  const int ci = k - tap * cin;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const int coa = co0 + c;  // this lane's row (A operand)
  const int hw = H * W;
  const int p_begin = blockIdx.z * chunk;
  const int p_end = min(npix, p_begin + chunk);
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int p0 = p_begin; p0 < p_end; p0 += 32) {
    bf16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int p = p0 + 8 * g + e;
      float av = 0.0f, bv = 0.0f;
      if (p < p_end) {
        if (coa < cout) av = dz[(size_t)p * dzs + coa];
        if (tap < 9) {
          const int n = p / hw;
          const int rem = p - n * hw;
          const int yy = rem / W + dy, xx = rem - (rem / W) * W + dx;
          if (yy >= 0 && yy < H && xx >= 0 && xx < W)
            bv = (float)x[(size_t)(n * hw + yy * W + xx) * xs + ci];
        }
      }
      a[e] = (__bf16)av;
      b[e] = (__bf16)bv;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int co = co0 + 4 * g + r;
    if (co < cout && k < kp) atomicAdd(dw + (size_t)co * kp + k, acc[r]);
  }
}

// Bias gradient: db[co] += sum_p dz[p][co].
__global__ __launch_bounds__(256) void k_bias_grad(const float* __restrict__ dz,
                                                   int npix, int cout, int dzs,
                                                   float* __restrict__ db) {
  const int co = blockIdx.y;
  float s = 0.0f;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += gridDim.x * blockDim.x)
    s += dz[(size_t)p * dzs + co];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(db + co, s);
}

// Flipped, channel-transposed weights for the input gradient:
// wt[ci][tap * cout + co] = w[co][(8 - tap) * cin + ci], zero padded to kpt.
__global__ void k_weight_flip(const __bf16* __restrict__ w, int cout, int cin,
                              int kp, __bf16* __restrict__ wt, int kpt) {
  const int ci = blockIdx.y;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kpt;
       k += gridDim.x * blockDim.x) {
    __bf16 v = (__bf16)0.0f;
    if (k < 9 * cout && ci < cin) {
      const int tap = k / cout, co = k - tap * cout;
      v = w[(size_t)co * kp + (8 - tap) * cin + ci];
    }
    wt[(size_t)ci * kpt + k] = v;
  }
}

// ---------------------------------------------------------------------------
// Elementwise and resampling kernels (one thread per element).

// dz = da * lrelu'(.) where the sign is read from the activation a.
__global__ void k_lrelu_backward(const float* __restrict__ da, int das, int dao,
                                 const __bf16* __restrict__ a, int as, int ao,
                                 float* __restrict__ dz, int npix, int c) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)npix * c) return;
  const size_t p = i / c, k = i - p * c;
  const float g = da[p * das + dao + k];
  dz[i] = (float)a[p * as + ao + k] > 0.0f ? g : 0.1f * g;
}

// 2x2 max pool of x ([N][H][W] at stride xs, offset xo) into y ([N][H/2][W/2]
// x c, contiguous); idx keeps the winning child (0..3, first maximum).
__global__ void k_maxpool2(const __bf16* __restrict__ x, int xs, int xo, int N,
                           int H, int W, int c, __bf16* __restrict__ y,
                           uint8_t* __restrict__ idx) {
  const int Ho = H / 2, Wo = W / 2;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * Ho * Wo * c) return;
  const size_t q = i / c;
  const int k = (int)(i - q * c);
  const int n = (int)(q / (Ho * Wo));
  const int rem = (int)(q - (size_t)n * Ho * Wo);
  const int oy = rem / Wo, ox = rem - (rem / Wo) * Wo;
  float best = 0.0f;
  int bi = 0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int yy = 2 * oy + (s >> 1), xx = 2 * ox + (s & 1);
    const float v = (float)x[((size_t)n * H * W + (size_t)yy * W + xx) * xs + xo + k];
    if (s == 0 || v > best) {
      best = v;
      bi = s;
    }
  }
  y[i] = (__bf16)best;
  idx[i] = (uint8_t)bi;
}

// dx[winner] += dy  (dx at stride dxs, offset dxo; each child is written by
// exactly one pooled element, so no atomics are needed).
__global__ void k_maxpool2_backward(const float* __restrict__ dy,
                                    const uint8_t* __restrict__ idx, int N,
                                    int H, int W, int c, float* __restrict__ dx,
                                    int dxs, int dxo) {
  const int Ho = H / 2, Wo = W / 2;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * Ho * Wo * c) return;
  const size_t q = i / c;
  const int k = (int)(i - q * c);
  const int n = (int)(q / (Ho * Wo));
  const int rem = (int)(q - (size_t)n * Ho * Wo);
  const int oy = rem / Wo, ox = rem - (rem / Wo) * Wo;
  const int s = idx[i];
  const int yy = 2 * oy + (s >> 1), xx = 2 * ox + (s & 1);
  dx[((size_t)n * H * W + (size_t)yy * W + xx) * dxs + dxo + k] += dy[i];
}

// Nearest 2x upsampling of x ([N][h][w] x c contiguous) into y
// ([N][2h][2w] at stride ys, offset yo).
__global__ void k_upsample2(const __bf16* __restrict__ x, int N, int h, int w,
                            int c, __bf16* __restrict__ y, int ys, int yo) {
  const int H = 2 * h, W = 2 * w;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * H * W * c) return;
  const size_t q = i / c;
  const int k = (int)(i - q * c);
  const int n = (int)(q / ((size_t)H * W));
  const int rem = (int)(q - (size_t)n * H * W);
  const int yy = rem / W, xx = rem - (rem / W) * W;
  y[q * ys + yo + k] = x[((size_t)n * h * w + (size_t)(yy / 2) * w + xx / 2) * c + k];
}

// dx[low] = sum of the 4 children of dy ([N][2h][2w] at stride dys, offset
// dyo); dx contiguous.
__global__ void k_upsample2_backward(const float* __restrict__ dy, int dys,
                                     int dyo, int N, int h, int w, int c,
                                     float* __restrict__ dx) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * h * w * c) return;
  const size_t q = i / c;
  const int k = (int)(i - q * c);
  const int n = (int)(q / ((size_t)h * w));
  const int rem = (int)(q - (size_t)n * h * w);
  const int oy = rem / w, ox = rem - (rem / w) * w;
  const int W = 2 * w;
  float s = 0.0f;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int yy = 2 * oy + (t >> 1), xx = 2 * ox + (t & 1);
    s += dy[((size_t)n * 4 * h * w + (size_t)yy * W + xx) * dys + dyo + k];
  }
  dx[i] = s;
}

// Loss of the 4 output channels (t = normalised log depth, slopes sx, sy,
// confidence logit) against targets (t*, sx*, sy*, valid):
//   L = sum_valid [huber_0.02(t - t*) + wn (huber_0.5(sx - sx*) +
//       huber_0.5(sy - sy*))] * inv_valid + wc * sum_all BCE(c, valid) * inv_all
// Writes dL/dout and adds the three terms to sums[0..2].
__device__ __forceinline__ float huber(float d, float delta, float* g) {
  const float a = fabsf(d);
  if (a <= delta) {
    *g = d / delta;
    return 0.5f * d * d / delta;
  }
  *g = d > 0.0f ? 1.0f : -1.0f;
  return a - 0.5f * delta;
}

__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ out,
                                              int os,
                                              const float* __restrict__ tgt,
                                              int npix, float wn, float wc,
                                              float inv_valid, float inv_all,
                                              float* __restrict__ dout,
                                              float* __restrict__ sums) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f;
  if (p < npix) {
    const float* o = out + (size_t)p * os;
    const float* t = tgt + (size_t)p * 4;
    const bool valid = t[3] > 0.5f;
    float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
    if (valid) {
      float h;
      l0 = huber(o[0] - t[0], 0.02f, &h) * inv_valid;
      g0 = h * inv_valid;
      l1 = wn * huber(o[1] - t[1], 0.5f, &h) * inv_valid;
      g1 = wn * h * inv_valid;
      l1 += wn * huber(o[2] - t[2], 0.5f, &h) * inv_valid;
      g2 = wn * h * inv_valid;
    }
    // BCE with logits: log(1 + e^z) - y z, gradient sigmoid(z) - y
    const float z = o[3], yv = valid ? 1.0f : 0.0f;
    const float sp = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));
    l2 = wc * (sp - yv * z) * inv_all;
    const float sg = 1.0f / (1.0f + expf(-z));
    float* d = dout + (size_t)p * os;
    d[0] = g0;
    d[1] = g1;
    d[2] = g2;
    d[3] = wc * (sg - yv) * inv_all;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    l0 += __shfl_xor(l0, off);
    l1 += __shfl_xor(l1, off);
    l2 += __shfl_xor(l2, off);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(sums + 0, l0);
    atomicAdd(sums + 1, l1);
    atomicAdd(sums + 2, l2);
  }
}

// Adam on FP32 masters, refreshing the BF16 copy.
__global__ void k_adam(float* __restrict__ w, __bf16* __restrict__ w16,
                       float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  if (w16) w16[i] = (__bf16)wi;
}

// FP32 -> BF16 copy (weights after initialisation or loading).
__global__ void k_to_bf16(const float* __restrict__ x, __bf16* __restrict__ y,
                          int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

// ---------------------------------------------------------------------------
// Launchers

static inline unsigned blocks(size_t n, int t = 256) {
  return (unsigned)((n + t - 1) / t);
}

template <typename TI>
static hipError_t conv_nt(const TI* x, int npix, int H, int W, int cin, int xs,
                          const __bf16* w, int cout, int kp, const float* bias,
                          void* y, int y_f32, int ys, int yo, int acc,
                          __bf16* act, int as, int ao, hipStream_t s) {
  const unsigned gx = blocks((size_t)(npix + 15) / 16, 4);
  if (cout <= 16) {
    hipLaunchKernelGGL((k_conv3x3<TI, 1>), dim3(gx, 1), dim3(256), 0, s, x, npix, H, W,
                       cin, xs, w, cout, kp, bias, y, y_f32, ys, yo, acc, act, as, ao);
  } else if (cout <= 32) {
    hipLaunchKernelGGL((k_conv3x3<TI, 2>), dim3(gx, 1), dim3(256), 0, s, x, npix, H, W,
                       cin, xs, w, cout, kp, bias, y, y_f32, ys, yo, acc, act, as, ao);
  } else {
    hipLaunchKernelGGL((k_conv3x3<TI, 4>), dim3(gx, (cout + 63) / 64), dim3(256), 0, s, x,
                       npix, H, W, cin, xs, w, cout, kp, bias, y, y_f32, ys, yo, acc, act,
                       as, ao);
  }
  return hipGetLastError();
}

hipError_t conv3x3(const void* x, int x_f32, int npix, int H, int W, int cin,
                   int xs, const void* w, int cout, int kp, const float* bias,
                   void* y, int y_f32, int ys, int yo, int accumulate, void* act,
                   int as, int ao, hipStream_t s) {
  const __bf16* wb = static_cast<const __bf16*>(w);
  __bf16* ab = static_cast<__bf16*>(act);
  if (x_f32)
    return conv_nt(static_cast<const float*>(x), npix, H, W, cin, xs, wb, cout, kp, bias, y,
                   y_f32, ys, yo, accumulate, ab, as, ao, s);
  return conv_nt(static_cast<const __bf16*>(x), npix, H, W, cin, xs, wb, cout, kp, bias, y,
                 y_f32, ys, yo, accumulate, ab, as, ao, s);
}

hipError_t conv3x3_wgrad(const void* x, int npix, int H, int W, int cin, int xs,
                         const float* dz, int cout, int dzs, float* dw, int kp,
                         float* db, hipStream_t s) {
  const int chunk = 1024;
  dim3 grid(blocks((size_t)(kp + 15) / 16, 4), (cout + 15) / 16, blocks(npix, chunk));
  hipLaunchKernelGGL(k_conv3x3_wgrad, grid, dim3(256), 0, s,
                     static_cast<const __bf16*>(x), npix, H, W, cin, xs, dz, cout, dzs,
                     dw, kp, chunk);
  if (db) {
    hipLaunchKernelGGL(k_bias_grad, dim3(std::min(64u, blocks(npix)), cout), dim3(256), 0, s,
                       dz, npix, cout, dzs, db);
  }
  return hipGetLastError();
}

hipError_t weight_flip(const void* w, int cout, int cin, int kp, void* wt,
                       int kpt, hipStream_t s) {
  hipLaunchKernelGGL(k_weight_flip, dim3(blocks(kpt), cin), dim3(256), 0, s,
                     static_cast<const __bf16*>(w), cout, cin, kp,
                     static_cast<__bf16*>(wt), kpt);
  return hipGetLastError();
}

hipError_t lrelu_backward(const float* da, int das, int dao, const void* a,
                          int as, int ao, float* dz, int npix, int c,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_lrelu_backward, dim3(blocks((size_t)npix * c)), dim3(256), 0, s, da,
                     das, dao, static_cast<const __bf16*>(a), as, ao, dz, npix, c);
  return hipGetLastError();
}

hipError_t maxpool2(const void* x, int xs, int xo, int N, int H, int W, int c,
                    void* y, uint8_t* idx, hipStream_t s) {
  hipLaunchKernelGGL(k_maxpool2, dim3(blocks((size_t)N * (H / 2) * (W / 2) * c)), dim3(256),
                     0, s, static_cast<const __bf16*>(x), xs, xo, N, H, W, c,
                     static_cast<__bf16*>(y), idx);
  return hipGetLastError();
}

hipError_t maxpool2_backward(const float* dy, const uint8_t* idx, int N, int H,
                             int W, int c, float* dx, int dxs, int dxo,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_maxpool2_backward, dim3(blocks((size_t)N * (H / 2) * (W / 2) * c)),
                     dim3(256), 0, s, dy, idx, N, H, W, c, dx, dxs, dxo);
  return hipGetLastError();
}

hipError_t upsample2(const void* x, int N, int h, int w, int c, void* y, int ys,
                     int yo, hipStream_t s) {
  hipLaunchKernelGGL(k_upsample2, dim3(blocks((size_t)N * 4 * h * w * c)), dim3(256), 0, s,
                     static_cast<const __bf16*>(x), N, h, w, c, static_cast<__bf16*>(y), ys,
                     yo);
  return hipGetLastError();
}

hipError_t upsample2_backward(const float* dy, int dys, int dyo, int N, int h,
                              int w, int c, float* dx, hipStream_t s) {
  hipLaunchKernelGGL(k_upsample2_backward, dim3(blocks((size_t)N * h * w * c)), dim3(256),
                     0, s, dy, dys, dyo, N, h, w, c, dx);
  return hipGetLastError();
}

hipError_t loss(const float* out, int os, const float* tgt, int npix, float wn,
                float wc, float inv_valid, float inv_all, float* dout,
                float* sums, hipStream_t s) {
  hipLaunchKernelGGL(k_loss, dim3(blocks(npix)), dim3(256), 0, s, out, os, tgt, npix, wn, wc,
                     inv_valid, inv_all, dout, sums);
  return hipGetLastError();
}

hipError_t adam(float* w, void* w16, float* m, float* v, const float* g, int n,
                float lr, float b1, float b2, float eps, int step,
                hipStream_t s) {
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
  const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
  hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, s, w,
                     static_cast<__bf16*>(w16), m, v, g, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t to_bf16(const float* x, void* y, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n)), dim3(256), 0, s, x,
                     static_cast<__bf16*>(y), n);
  return hipGetLastError();
}

}  // namespace nn
