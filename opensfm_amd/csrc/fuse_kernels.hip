// Depthmap fusion (dense::DepthmapFuser::Fuse, opensfm/src/lib/dense/fuser.cc:113-480)
// on gfx950.
//
// The reference walks the primary views one after another and, inside a
// view, the pixels in row-major order; each fused point marks the pixels it
// used, so a later pixel cannot use them again.  That greedy order is kept
// exactly by rounds of claim / decide over the pixels still undecided:
//
//   claim   every undecided pixel p gathers the neighbour pixels S(p) that
//           pass the geometric tests and are not yet fused; if 1 + |S(p)| is
//           below the minimum it is decided (no point; S only shrinks).
//           Otherwise it claims its own pixel and S(p) with a key that
//           orders by (round, lowest pixel index first).
//   decide  p is decided when it holds the claim on every pixel it asked
//           for: no earlier undecided pixel can still take one of them, and
//           later pixels decided before it never touched S(p) (they would
//           have lost the claim to p).  It emits the medians and marks its
//           pixels fused.
//   compact the undecided pixels form the next round's list.
//
// The lowest undecided pixel always wins all its claims, so every round
// decides at least one pixel; in practice most pixels decide in the first
// rounds.  The fused points are then written in row-major order of the
// view by a count / scan / scatter.
//
// Float arithmetic follows the reference's Eigen fixed-size expressions
// (3-term sums as c0 + (c1 + c2), (s * A) * v as sum_k (s * A_ik) * v_k,
// divisions and square roots correctly rounded -- hipcc's default for
// sqrtf and /, whereas __fsqrt_rn is the native approximation), compiled with
// -ffp-contract=off, so the result is the oracle's to the bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "fuse_kernels.h"

namespace fu {

namespace {

__device__ __forceinline__ float dot3(const float* a, const float* b) {
  return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

__device__ __forceinline__ void mv3(const float* M, const float* x, float* y) {
  float t[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] = dot3(M + 3 * i, x);
#pragma unroll
  for (int i = 0; i < 3; ++i) y[i] = t[i];
}

// Rinv * (d * Kinv * (x, y, 1) - t)  (fuser.cc:183-184, 276-278)
__device__ __forceinline__ void backproject(const Camera& C, float d, float x, float y,
                                            float* out) {
  float w[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float a = d * C.Kinv[3 * i], b = d * C.Kinv[3 * i + 1], c = d * C.Kinv[3 * i + 2];
    w[i] = a * x + (b * y + c * 1.0f);
    w[i] = w[i] - C.t[i];
  }
  mv3(C.Rinv, w, out);
}

// K * (R * X + t)  (fuser.cc:225, 279)
__device__ __forceinline__ void project(const Camera& C, const float* X, float* out) {
  float a[3];
  mv3(C.R, X, a);
#pragma unroll
  for (int i = 0; i < 3; ++i) a[i] = a[i] + C.t[i];
  mv3(C.K, a, out);
}

__device__ __forceinline__ float norm3(const float* v) {
  return sqrtf(v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]));  // correctly rounded
}

// static_cast<int>(std::round(x)); outside int (or NaN) x86 gives INT_MIN,
// which every bounds test rejects.
__device__ __forceinline__ int round_int(float x) {
  const float r = roundf(x);
  if (!(fabsf(r) < 2147483648.0f)) return INT_MIN;
  return (int)r;
}

// BilinearDepth (fuser.cc:95-111) over foundation::BilinearInterpolation.
__device__ __forceinline__ float bilinear_depth(const float* __restrict__ D, int W, int H,
                                                float u, float v) {
  const int x0 = (int)floorf(u), y0 = (int)floorf(v);
  const int x1 = x0 + 1, y1 = y0 + 1;
  if (x0 < 0 || y0 < 0 || x1 >= W || y1 >= H) return -1.0f;
  const float d00 = D[(size_t)y0 * W + x0], d01 = D[(size_t)y0 * W + x1];
  const float d10 = D[(size_t)y1 * W + x0], d11 = D[(size_t)y1 * W + x1];
  if (d00 <= 0.0f || d01 <= 0.0f || d10 <= 0.0f || d11 <= 0.0f) return -1.0f;
  if (u < 0.0f || u >= (float)(W - 1) || v < 0.0f || v >= (float)(H - 1)) return -1.0f;
  const float dx = u - (float)x0, dy = v - (float)y0;  // (int)u == floor(u) for u >= 0
  const float a = (1.0f - dx) * d00 + dx * d01;
  const float b = (1.0f - dx) * d10 + dx * d11;
  return (1.0f - dy) * a + dy * b;
}

// The neighbour pixel the reference point P lands on (rounded), or -1 when
// it is behind the camera or outside the border margin (fuser.cc:225-237).
__device__ __forceinline__ long long land(const View& S, const float* P, int bm, float* proj,
                                          float* su, float* sv, int* sc, int* sr) {
  project(S.cam, P, proj);
  if (proj[2] < 1e-6f) return -1;
  *su = __fdiv_rn(proj[0], proj[2]);
  *sv = __fdiv_rn(proj[1], proj[2]);
  *sc = round_int(*su);
  *sr = round_int(*sv);
  if (*sc < bm || *sc >= S.width - bm || *sr < bm || *sr >= S.height - bm) return -1;
  return (long long)(*sr) * S.width + *sc;
}

// The tests after the landing pixel (fuser.cc:240-318): mask, depth
// (asymmetric), forward-backward reprojection, normal angle.  Fills the
// neighbour's point and world normal when it passes.  The fused flag is the
// caller's.
__device__ __forceinline__ bool consistent(const View& A, const View& S, const Params& prm,
                                           int c, int r, const float* nw, const float* proj,
                                           float su, float sv, int sc, int sr, long long sp,
                                           float* SP, float* snw) {
  if (S.mask && S.mask[sp] == 0) return false;
  float sd = bilinear_depth(S.depth, S.width, S.height, su, sv);
  if (sd <= 0.0f) {
    sd = S.depth[sp];
    if (sd <= 0.0f) return false;
  }
  const float dp = proj[2];
  const float front = prm.max_depth_error * dp;
  const float back = front * prm.behind;
  const float diff = sd - dp;
  if (diff > front || diff < -back) return false;
  backproject(S.cam, sd, (float)sc, (float)sr, SP);
  float rp[3];
  project(A.cam, SP, rp);
  if (rp[2] < 1e-6f) return false;
  const float e0 = __fdiv_rn(rp[0], rp[2]) - (float)c;
  const float e1 = __fdiv_rn(rp[1], rp[2]) - (float)r;
  const float err = e0 * e0 + e1 * e1;
  if (err > prm.max_reproj_sq) return false;
  float sn[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) sn[k] = S.normal[3 * sp + k];
  const float snlen = norm3(sn);
  if (snlen < 1e-6f) return false;
#pragma unroll
  for (int k = 0; k < 3; ++k) sn[k] = __fdiv_rn(sn[k], snlen);
  mv3(S.cam.Rinv, sn, snw);
  return !(dot3(nw, snw) < prm.min_cos);
}

// The reference pixel's own observation; false when the pixel is skipped
// (fuser.cc:168-193).
__device__ __forceinline__ bool own(const View& A, int p, int c, int r, float* P, float* nw) {
  if (A.mask && A.mask[p] == 0) return false;
  const float d = A.depth[p];
  if (d <= 0.0f) return false;
  backproject(A.cam, d, (float)c, (float)r, P);
  float nc[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) nc[k] = A.normal[3 * (size_t)p + k];
  const float nlen = norm3(nc);
  if (nlen < 1e-6f) return false;
#pragma unroll
  for (int k = 0; k < 3; ++k) nc[k] = __fdiv_rn(nc[k], nlen);
  mv3(A.cam.Rinv, nc, nw);
  return true;
}

__device__ __forceinline__ unsigned long long claim_key(unsigned int round, int p) {
  return ((unsigned long long)round << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)p);
}

// (n/2)-th smallest (std::nth_element then v[n/2]); ties by position.
template <typename T>
__device__ __forceinline__ T median(const T* v, int n) {
  const int k = n / 2;
  T out = v[0];
  for (int i = 0; i < n; ++i) {
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += (v[j] < v[i] || (v[j] == v[i] && j < i)) ? 1 : 0;
    if (rank == k) out = v[i];
  }
  return out;
}

__global__ __launch_bounds__(256) void k_fuse_init(int W, int H, int bm, int* __restrict__ active,
                                                   int* __restrict__ count,
                                                   uint8_t* __restrict__ st,
                                                   uint8_t* __restrict__ ok) {
  const int bw = W - 2 * bm, bh = H - 2 * bm;
  const long long nbox = (bw > 0 && bh > 0) ? (long long)bw * bh : 0;
  const long long npix = (long long)W * H;
  const long long stride = (long long)gridDim.x * 256;
  for (long long i = blockIdx.x * 256 + threadIdx.x; i < npix; i += stride) {
    st[i] = 0;
    ok[i] = 0;
    if (i < nbox) {
      const int r = bm + (int)(i / bw), c = bm + (int)(i % bw);
      active[i] = r * W + c;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = (int)nbox;
}

__global__ __launch_bounds__(256) void k_fuse_claim(
    const View* __restrict__ views, Params prm, const int* __restrict__ active,
    const int* __restrict__ count, int* __restrict__ next_count, unsigned int round,
    unsigned long long* __restrict__ masks, uint8_t* __restrict__ st) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0;
  const int n = *count;
  const View& A = views[prm.ref];
  for (int idx = blockIdx.x * 256 + threadIdx.x; idx < n; idx += gridDim.x * 256) {
    const int p = active[idx];
    const int r = p / A.width, c = p - (p / A.width) * A.width;
    float P[3], nw[3];
    if (A.fused[p] || !own(A, p, c, r, P, nw)) {
      st[p] = 1;
      continue;
    }
    unsigned long long m = 0;
    int cnt = 0;
    for (int s = 0; s < prm.n_nb; ++s) {
      const View& S = views[prm.nb[s]];
      float proj[3], su, sv, SP[3], snw[3];
      int sc, sr;
      const long long sp = land(S, P, prm.bm, proj, &su, &sv, &sc, &sr);
      if (sp < 0 || S.fused[sp]) continue;
      if (!consistent(A, S, prm, c, r, nw, proj, su, sv, sc, sr, sp, SP, snw)) continue;
      m |= 1ull << s;
      ++cnt;
    }
    if (1 + cnt < prm.min_consistent) {
      st[p] = 1;
      continue;
    }
    masks[idx] = m;
    const unsigned long long key = claim_key(round, p);
    atomicMax(&A.claim[p], key);
    for (unsigned long long b = m; b; b &= b - 1) {
      const int s = __ffsll((long long)b) - 1;
      const View& S = views[prm.nb[s]];
      float proj[3], su, sv;
      int sc, sr;
      const long long sp = land(S, P, prm.bm, proj, &su, &sv, &sc, &sr);
      atomicMax(&S.claim[sp], key);
    }
  }
}

// (n/2)-th smallest of the first n of N register-held values, ties by
// position: entries from n on are padded with `pad` (above every value)
// and an odd-even transposition sort -- stable, as it swaps only strictly
// greater neighbours -- orders the array with static indices, so it stays
// in registers.
template <int N, typename T>
__device__ __forceinline__ T median_n(const T (&in)[N], int n, T pad) {
  T v[N];
#pragma unroll
  for (int t = 0; t < N; ++t) v[t] = t < n ? in[t] : pad;
#pragma unroll
  for (int pass = 0; pass < N; ++pass) {
#pragma unroll
    for (int t = pass & 1; t + 1 < N; t += 2) {
      const T a = v[t], b = v[t + 1];
      const bool sw = a > b;
      v[t] = sw ? b : a;
      v[t + 1] = sw ? a : b;
    }
  }
  const int k = n / 2;
  T out = v[0];
#pragma unroll
  for (int t = 0; t < N; ++t)
    if (t == k) out = v[t];
  return out;
}

// A decided pixel: gathers its observations (itself, then the passing
// neighbours in slot order; fuser.cc:195-318), marks the neighbour pixels
// fused and writes the medians (fuser.cc:320-357).
//
// Up to kSmall observations: three passes (point, normal, colour), each
// recomputing the observations -- the computation is deterministic -- and
// keeping only that pass's three components in registers (static indices),
// so the kernel needs no private memory.  More observations: one pass into
// private memory (a separate kernel, launched only for views with that many
// neighbours).
constexpr int kSmall = 12;

template <int kPass>
__device__ __forceinline__ void fuse_pass(const View* __restrict__ views, const Params& prm,
                                          const View& A, int p, int c, int r, const float* P,
                                          const float* nw, unsigned long long m, float* out) {
  float v[3][kSmall];
  uint32_t col[kSmall];
  if constexpr (kPass == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k][0] = P[k];
  } else if constexpr (kPass == 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k][0] = nw[k];
  } else {
    col[0] = A.color[3 * (size_t)p] | (A.color[3 * (size_t)p + 1] << 8) |
             (A.color[3 * (size_t)p + 2] << 16);
  }
  int no = 1;
  for (unsigned long long b = m; b; b &= b - 1) {
    const int s = __ffsll((long long)b) - 1;
    const View& S = views[prm.nb[s]];
    float proj[3], su, sv, SP[3], snw[3];
    int sc, sr;
    const long long sp = land(S, P, prm.bm, proj, &su, &sv, &sc, &sr);
    if constexpr (kPass == 2) {
      const uint32_t cc =
          S.color[3 * sp] | (S.color[3 * sp + 1] << 8) | (S.color[3 * sp + 2] << 16);
#pragma unroll
      for (int t = 1; t < kSmall; ++t)
        if (t == no) col[t] = cc;
      S.fused[sp] = 1;
    } else {
      consistent(A, S, prm, c, r, nw, proj, su, sv, sc, sr, sp, SP, snw);
#pragma unroll
      for (int t = 1; t < kSmall; ++t)
        if (t == no) {
#pragma unroll
          for (int k = 0; k < 3; ++k) v[k][t] = kPass == 0 ? SP[k] : snw[k];
        }
    }
    ++no;
  }
  if constexpr (kPass == 2) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      uint32_t ch[kSmall];
#pragma unroll
      for (int t = 0; t < kSmall; ++t) ch[t] = (col[t] >> (8 * k)) & 255u;
      out[k] = (float)median_n<kSmall>(ch, no, 256u);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      out[k] = median_n<kSmall>(v[k], no, INFINITY);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <bool kLarge>
__device__ __forceinline__ void fuse_emit(const View* __restrict__ views, const Params& prm,
                                          const View& A, int p, int c, int r, const float* P,
                                          const float* nw, unsigned long long m,
                                          float* __restrict__ res,
                                          uint8_t* __restrict__ res_clr) {
  float mp[3], mn[3], mc[3];
  if constexpr (kLarge) {
    constexpr int N = kMaxNeighbours + 1;
    float ob[6][N];
    uint8_t oc[3][N];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ob[k][0] = P[k];
      ob[3 + k][0] = nw[k];
      oc[k][0] = A.color[3 * (size_t)p + k];
    }
    int no = 1;
    for (unsigned long long b = m; b; b &= b - 1) {
      const int s = __ffsll((long long)b) - 1;
      const View& S = views[prm.nb[s]];
      float proj[3], su, sv, SP[3], snw[3];
      int sc, sr;
      const long long sp = land(S, P, prm.bm, proj, &su, &sv, &sc, &sr);
      consistent(A, S, prm, c, r, nw, proj, su, sv, sc, sr, sp, SP, snw);
      for (int k = 0; k < 3; ++k) {
        ob[k][no] = SP[k];
        ob[3 + k][no] = snw[k];
        oc[k][no] = S.color[3 * sp + k];
      }
      ++no;
      S.fused[sp] = 1;
    }
    for (int k = 0; k < 3; ++k) {
      mp[k] = median(ob[k], no);
      mn[k] = median(ob[3 + k], no);
      mc[k] = (float)median(oc[k], no);
    }
  } else {
    fuse_pass<0>(views, prm, A, p, c, r, P, nw, m, mp);
    fuse_pass<1>(views, prm, A, p, c, r, P, nw, m, mn);
    fuse_pass<2>(views, prm, A, p, c, r, P, nw, m, mc);
  }
  const float mlen = norm3(mn);
  if (mlen > 1e-6f) {
#pragma unroll
    for (int k = 0; k < 3; ++k) mn[k] = __fdiv_rn(mn[k], mlen);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    res[6 * (size_t)p + k] = mp[k];
    res[6 * (size_t)p + 3 + k] = mn[k];
    res_clr[3 * (size_t)p + k] = (uint8_t)mc[k];
  }
}

// Decides the pixels that hold all their claims and have at most kSmall
// observations (kLarge: the ones with more), writes their points and marks
// their pixels fused.

template <bool kLarge>
__global__ __launch_bounds__(256) void k_fuse_decide(
    const View* __restrict__ views, Params prm, const int* __restrict__ active,
    const int* __restrict__ count, unsigned int round,
    const unsigned long long* __restrict__ masks, uint8_t* __restrict__ st,
    float* __restrict__ res, uint8_t* __restrict__ res_clr, uint8_t* __restrict__ ok) {
  const int n = *count;
  const View& A = views[prm.ref];
  for (int idx = blockIdx.x * 256 + threadIdx.x; idx < n; idx += gridDim.x * 256) {
    const int p = active[idx];
    if (st[p]) continue;
    const unsigned long long m = masks[idx];
    if ((1 + __popcll(m) > kSmall) != kLarge) continue;
    const unsigned long long key = claim_key(round, p);
    if (A.claim[p] != key) continue;
    const int r = p / A.width, c = p - (p / A.width) * A.width;
    float P[3], nw[3];
    own(A, p, c, r, P, nw);
    bool mine = true;
    for (unsigned long long b = m; b && mine; b &= b - 1) {
      const int s = __ffsll((long long)b) - 1;
      const View& S = views[prm.nb[s]];
      float proj[3], su, sv;
      int sc, sr;
      const long long sp = land(S, P, prm.bm, proj, &su, &sv, &sc, &sr);
      mine = S.claim[sp] == key;
    }
    if (!mine) continue;
    fuse_emit<kLarge>(views, prm, A, p, c, r, P, nw, m, res, res_clr);
    ok[p] = 1;
    st[p] = 1;
    A.fused[p] = 1;
  }
}

__global__ __launch_bounds__(256) void k_fuse_compact(const int* __restrict__ active,
                                                      const int* __restrict__ count,
                                                      const uint8_t* __restrict__ st,
                                                      int* __restrict__ next,
                                                      int* __restrict__ next_count) {
  const int n = *count;
  const int lane = threadIdx.x & 63;
  const int stride = gridDim.x * 256;
  for (int base = blockIdx.x * 256 + (threadIdx.x & ~63); base < n; base += stride) {
    const int idx = base + lane;
    const int p = idx < n ? active[idx] : 0;
    const bool keep = idx < n && st[p] == 0;
    const unsigned long long b = __ballot(keep);
    int off = 0;
    if (lane == 0 && b) off = atomicAdd(next_count, __popcll(b));
    off = __shfl(off, 0);
    if (keep) next[off + __popcll(b & ((1ull << lane) - 1ull))] = p;
  }
}

__global__ __launch_bounds__(256) void k_fuse_count(const uint8_t* __restrict__ ok, int npix,
                                                    int* __restrict__ counts) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int c = __syncthreads_count(p < npix && ok[p]);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void k_fuse_scatter(
    const uint8_t* __restrict__ ok, int npix, const long long* __restrict__ offsets,
    const float* __restrict__ res, const uint8_t* __restrict__ res_clr,
    float* __restrict__ points, float* __restrict__ normals, uint8_t* __restrict__ colors) {
  __shared__ int wave_base[4];
  const int p = blockIdx.x * 256 + threadIdx.x;
  const bool keep = p < npix && ok[p];
  const unsigned long long m = __ballot(keep);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
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
  const long long o = offsets[blockIdx.x] + wave_base[wave] + __popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    points[3 * o + k] = res[6 * (size_t)p + k];
    normals[3 * o + k] = res[6 * (size_t)p + 3 + k];
    colors[3 * o + k] = res_clr[3 * (size_t)p + k];
  }
}

// ---------------------------------------------------------------------------
// Statistical outlier removal (fuser.cc:369-478).
//
// Points are quantised to 21 bits per axis over their bounding cube and
// sorted by Morton code; a cell of level L is a run of equal code >> 3L.
// Each level from the first searched one up has a hash table code -> run.
// A query searches the 3 x 3 x 3 cells around it at level L and stops when
// the k-th nearest distance is below the distance to the searched block's
// boundary (no point outside can be nearer); otherwise it restarts one
// level up.  The top level's block holds every point, so the search is
// exact.  Distances are vlfeat's float L2 ((dx^2 + dy^2) + dz^2, the query
// minus the point) and the k nearest others are summed as square roots in
// ascending order, as the reference's k-d forest returns them.

__device__ __forceinline__ unsigned long long spread3(unsigned int x) {
  unsigned long long v = x & 0x1FFFFFu;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

__device__ __forceinline__ unsigned long long morton3(unsigned int x, unsigned int y,
                                                      unsigned int z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

__device__ __forceinline__ unsigned int hash_slot(unsigned long long key, unsigned int mask) {
  return (unsigned int)((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
}

constexpr unsigned int kCells = 1u << 21;

__device__ __forceinline__ unsigned int quantise(float x, float lo, float quantum) {
  const float f = __fdiv_rn(x - lo, quantum);
  if (!(f > 0.0f)) return 0;
  const unsigned int q = f >= (float)kCells ? kCells - 1 : (unsigned int)f;
  return q < kCells ? q : kCells - 1;
}

__global__ __launch_bounds__(256) void k_bbox(const float* __restrict__ pts, long long n,
                                              float* __restrict__ lohi) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (long long i = blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float v = pts[3 * i + k];
      lo[k] = fminf(lo[k], v);
      hi[k] = fmaxf(hi[k], v);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    for (int o = 32; o > 0; o >>= 1) {
      lo[k] = fminf(lo[k], __shfl_xor(lo[k], o));
      hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o));
    }
  }
  if ((threadIdx.x & 63) == 0) {
    // order-preserving integer images of the floats for atomic min / max
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int a = __float_as_int(lo[k]), b = __float_as_int(hi[k]);
      atomicMin((int*)&lohi[k], a >= 0 ? a : (int)(0x80000000u ^ ~(unsigned)a));
      atomicMax((int*)&lohi[3 + k], b >= 0 ? b : (int)(0x80000000u ^ ~(unsigned)b));
    }
  }
}

__global__ void k_bbox_finish(float* lohi) {
  const int k = threadIdx.x;
  if (k < 6) {
    const int a = __float_as_int(lohi[k]);
    lohi[k] = __int_as_float(a >= 0 ? a : (int)~(0x80000000u ^ (unsigned)a));
  }
}

__global__ void k_bbox_init(float* lohi) {
  const int k = threadIdx.x;
  if (k < 3) ((int*)lohi)[k] = 0x7FFFFFFF;
  else if (k < 6) ((int*)lohi)[k] = (int)0x80000000;
}

__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ pts, long long n,
                                                const float* __restrict__ lo, float quantum,
                                                unsigned long long* __restrict__ codes,
                                                unsigned int* __restrict__ idx) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const unsigned int x = quantise(pts[3 * i], lo[0], quantum);
  const unsigned int y = quantise(pts[3 * i + 1], lo[1], quantum);
  const unsigned int z = quantise(pts[3 * i + 2], lo[2], quantum);
  codes[i] = morton3(x, y, z);
  idx[i] = (unsigned int)i;
}

__global__ __launch_bounds__(256) void k_gather_sorted(const float* __restrict__ pts,
                                                       const unsigned int* __restrict__ idx,
                                                       long long n, float* __restrict__ out) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const unsigned int j = idx[i];
#pragma unroll
  for (int k = 0; k < 3; ++k) out[3 * i + k] = pts[3 * (size_t)j + k];
}

// Cells of every level at once: one pass over the sorted codes.
__global__ __launch_bounds__(256) void k_level_count_all(
    const unsigned long long* __restrict__ codes, long long n, unsigned int* __restrict__ n_cells) {
  __shared__ unsigned int blk[22];
  if (threadIdx.x < 22) blk[threadIdx.x] = 0;
  __syncthreads();
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  const unsigned long long cur = i < n ? codes[i] : 0ull;
  const unsigned long long prev = (i > 0 && i < n) ? codes[i - 1] : 0ull;
  const unsigned long long diff = cur ^ prev;
#pragma unroll
  for (int L = 0; L < 22; ++L) {
    const bool start = i < n && (i == 0 || (diff >> (3 * L)) != 0ull);
    const unsigned long long b = __ballot(start);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&blk[L], (unsigned int)__popcll(b));
  }
  __syncthreads();
  if (threadIdx.x < 22 && blk[threadIdx.x]) atomicAdd(&n_cells[threadIdx.x], blk[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_level_insert(const unsigned long long* __restrict__ codes,
                                                      long long n, int level,
                                                      unsigned long long* __restrict__ keys,
                                                      unsigned int* __restrict__ start,
                                                      unsigned int mask) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const unsigned long long pre = codes[i] >> (3 * level);
  if (i > 0 && (codes[i - 1] >> (3 * level)) == pre) return;
  const unsigned long long key = pre + 1;
  unsigned int s = hash_slot(key, mask);
  while (true) {
    const unsigned long long old = atomicCAS(&keys[s], 0ull, key);
    if (old == 0ull) {
      start[s] = (unsigned int)i;
      return;
    }
    s = (s + 1) & mask;
  }
}

__device__ __forceinline__ int lookup(const unsigned long long* __restrict__ keys,
                                      unsigned int mask, unsigned long long key) {
  unsigned int s = hash_slot(key, mask);
  while (true) {
    const unsigned long long k = keys[s];
    if (k == key) return (int)s;
    if (k == 0ull) return -1;
    s = (s + 1) & mask;
  }
}

__global__ __launch_bounds__(256) void k_level_finish(const unsigned long long* __restrict__ codes,
                                                      long long n, int level,
                                                      const unsigned long long* __restrict__ keys,
                                                      const unsigned int* __restrict__ start,
                                                      unsigned int* __restrict__ cnt,
                                                      unsigned int mask) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const unsigned long long pre = codes[i] >> (3 * level);
  if (i + 1 < n && (codes[i + 1] >> (3 * level)) == pre) return;
  const int s = lookup(keys, mask, pre + 1);
  if (s >= 0) cnt[s] = (unsigned int)(i + 1) - start[s];
}

__global__ __launch_bounds__(256) void k_knn_mean(Grid g, const float* __restrict__ sorted,
                                                  const unsigned int* __restrict__ idx,
                                                  long long n, int k,
                                                  float* __restrict__ mean_knn) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const float q[3] = {sorted[3 * i], sorted[3 * i + 1], sorted[3 * i + 2]};
  float fq[3];
  unsigned int iq[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    fq[a] = __fdiv_rn(q[a] - g.lo[a], g.quantum);
    iq[a] = quantise(q[a], g.lo[a], g.quantum);
  }
  float bd[kMaxKnn];
  int found = 0;
  for (int L = g.level0; L < g.level0 + g.levels; ++L) {
    const int li = L - g.level0;
    const unsigned int top = (kCells >> L) - 1;  // largest cell coordinate at L
    const unsigned int cx = iq[0] >> L, cy = iq[1] >> L, cz = iq[2] >> L;
    found = 0;
    for (int dz = -1; dz <= 1; ++dz) {
      if ((dz < 0 && cz == 0) || (dz > 0 && cz == top)) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        if ((dy < 0 && cy == 0) || (dy > 0 && cy == top)) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          if ((dx < 0 && cx == 0) || (dx > 0 && cx == top)) continue;
          const unsigned long long key = morton3(cx + dx, cy + dy, cz + dz) + 1;
          const int s = lookup(g.keys[li], g.mask[li], key);
          if (s < 0) continue;
          const unsigned int b = g.start[li][s], e = b + g.cnt[li][s];
          for (unsigned int j = b; j < e; ++j) {
            if ((long long)j == i) continue;
            const float d0 = q[0] - sorted[3 * (size_t)j];
            const float d1 = q[1] - sorted[3 * (size_t)j + 1];
            const float d2 = q[2] - sorted[3 * (size_t)j + 2];
            float acc = 0.0f;
            acc += d0 * d0;
            acc += d1 * d1;
            acc += d2 * d2;
            float worst = 0.0f;
#pragma unroll
            for (int t = 0; t < kMaxKnn; ++t)
              if (t == k - 1) worst = bd[t];
            if (found == k && !(acc < worst)) continue;
            // the new distance goes to the end of the ascending list (over
            // the worst when full) and bubbles down to its place
            const int last = found < k ? found : k - 1;
#pragma unroll
            for (int t = 0; t < kMaxKnn; ++t)
              if (t == last) bd[t] = acc;
#pragma unroll
            for (int t = kMaxKnn - 1; t > 0; --t) {
              if (t <= last && bd[t - 1] > bd[t]) {
                const float x = bd[t - 1];
                bd[t - 1] = bd[t];
                bd[t] = x;
              }
            }
            if (found < k) ++found;
          }
        }
      }
    }
    // distance from the query to the searched block's boundary, in cell
    // units of level 0 (sides at the grid's edge hold no points beyond)
    float gq = INFINITY;
    const unsigned int c3[3] = {cx, cy, cz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float step = (float)(1u << L);
      if (c3[a] > 0) gq = fminf(gq, fq[a] - (float)(c3[a] - 1) * step);
      if (c3[a] < top) gq = fminf(gq, (float)(c3[a] + 2) * step - fq[a]);
    }
    if (gq == INFINITY) break;  // the block holds every point
    if (found == k) {
      float worst = 0.0f;
#pragma unroll
      for (int t = 0; t < kMaxKnn; ++t)
        if (t == k - 1) worst = bd[t];
      const float gr = gq * g.quantum * (1.0f - 1e-5f) - g.quantum * 2.0f;
      if (gr > 0.0f && worst < gr * gr) break;
    }
  }
  float sum = 0.0f;
#pragma unroll
  for (int t = 0; t < kMaxKnn; ++t)
    if (t < found) sum += sqrtf(bd[t]);
  mean_knn[idx[i]] = found > 0 ? __fdiv_rn(sum, (float)found) : 1e30f;
}

// The k nearest other points of every point as (index, squared distance),
// ascending by (distance, index): ties at the k-th place go to the lower
// index.  Fewer than k others leave index -1.
__device__ __forceinline__ bool knn_before(float d, unsigned int i, float e, unsigned int j) {
  return d < e || (d == e && i < j);
}

__global__ __launch_bounds__(256) void k_knn_list(Grid g, const float* __restrict__ sorted,
                                                  const unsigned int* __restrict__ idx,
                                                  long long n, int k, int* __restrict__ out_idx,
                                                  float* __restrict__ out_d2) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const float q[3] = {sorted[3 * i], sorted[3 * i + 1], sorted[3 * i + 2]};
  float fq[3];
  unsigned int iq[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    fq[a] = __fdiv_rn(q[a] - g.lo[a], g.quantum);
    iq[a] = quantise(q[a], g.lo[a], g.quantum);
  }
  float bd[kMaxKnn];
  unsigned int bi[kMaxKnn];
  int found = 0;
  for (int L = g.level0; L < g.level0 + g.levels; ++L) {
    const int li = L - g.level0;
    const unsigned int top = (kCells >> L) - 1;
    const unsigned int cx = iq[0] >> L, cy = iq[1] >> L, cz = iq[2] >> L;
    found = 0;
    for (int dz = -1; dz <= 1; ++dz) {
      if ((dz < 0 && cz == 0) || (dz > 0 && cz == top)) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        if ((dy < 0 && cy == 0) || (dy > 0 && cy == top)) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          if ((dx < 0 && cx == 0) || (dx > 0 && cx == top)) continue;
          const unsigned long long key = morton3(cx + dx, cy + dy, cz + dz) + 1;
          const int s = lookup(g.keys[li], g.mask[li], key);
          if (s < 0) continue;
          const unsigned int b = g.start[li][s], e = b + g.cnt[li][s];
          for (unsigned int j = b; j < e; ++j) {
            if ((long long)j == i) continue;
            const float d0 = q[0] - sorted[3 * (size_t)j];
            const float d1 = q[1] - sorted[3 * (size_t)j + 1];
            const float d2 = q[2] - sorted[3 * (size_t)j + 2];
            float acc = 0.0f;
            acc += d0 * d0;
            acc += d1 * d1;
            acc += d2 * d2;
            const unsigned int oj = idx[j];
            float wd = 0.0f;
            unsigned int wi = 0;
#pragma unroll
            for (int t = 0; t < kMaxKnn; ++t)
              if (t == k - 1) {
                wd = bd[t];
                wi = bi[t];
              }
            if (found == k && !knn_before(acc, oj, wd, wi)) continue;
            const int last = found < k ? found : k - 1;
#pragma unroll
            for (int t = 0; t < kMaxKnn; ++t)
              if (t == last) {
                bd[t] = acc;
                bi[t] = oj;
              }
#pragma unroll
            for (int t = kMaxKnn - 1; t > 0; --t) {
              if (t <= last && knn_before(bd[t], bi[t], bd[t - 1], bi[t - 1])) {
                const float x = bd[t - 1];
                const unsigned int y = bi[t - 1];
                bd[t - 1] = bd[t];
                bi[t - 1] = bi[t];
                bd[t] = x;
                bi[t] = y;
              }
            }
            if (found < k) ++found;
          }
        }
      }
    }
    float gq = INFINITY;
    const unsigned int c3[3] = {cx, cy, cz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float step = (float)(1u << L);
      if (c3[a] > 0) gq = fminf(gq, fq[a] - (float)(c3[a] - 1) * step);
      if (c3[a] < top) gq = fminf(gq, (float)(c3[a] + 2) * step - fq[a]);
    }
    if (gq == INFINITY) break;  // the block holds every point
    if (found == k) {
      float worst = 0.0f;
#pragma unroll
      for (int t = 0; t < kMaxKnn; ++t)
        if (t == k - 1) worst = bd[t];
      const float gr = gq * g.quantum * (1.0f - 1e-5f) - g.quantum * 2.0f;
      if (gr > 0.0f && worst < gr * gr) break;
    }
  }
  const size_t o = (size_t)idx[i] * k;
#pragma unroll
  for (int t = 0; t < kMaxKnn; ++t) {
    if (t < k) {
      out_idx[o + t] = t < found ? (int)bi[t] : -1;
      out_d2[o + t] = t < found ? bd[t] : 0.0f;
    }
  }
}

__global__ __launch_bounds__(256) void k_sor_count(const float* __restrict__ mean_knn,
                                                   long long n, float thr,
                                                   int* __restrict__ counts) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  const int c = __syncthreads_count(i < n && mean_knn[i] <= thr);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void k_sor_scatter(
    const float* __restrict__ mean_knn, long long n, float thr,
    const long long* __restrict__ offsets, const float* __restrict__ pts,
    const float* __restrict__ nrm, const uint8_t* __restrict__ clr, float* __restrict__ opts,
    float* __restrict__ onrm, uint8_t* __restrict__ oclr) {
  __shared__ int wave_base[4];
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  const bool keep = i < n && mean_knn[i] <= thr;
  const unsigned long long m = __ballot(keep);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
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
  const long long o = offsets[blockIdx.x] + wave_base[wave] + __popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    opts[3 * o + k] = pts[3 * i + k];
    onrm[3 * o + k] = nrm[3 * i + k];
    oclr[3 * o + k] = clr[3 * i + k];
  }
}

inline int blocks(long long n) { return (int)((n + 255) / 256); }

}  // namespace

hipError_t launch_fuse_init(int W, int H, int bm, int* active, int* count, uint8_t* st,
                            uint8_t* ok, hipStream_t s) {
  const long long npix = (long long)W * H;
  const int grid = (int)std::min<long long>(blocks(npix), 8192);
  hipLaunchKernelGGL(k_fuse_init, dim3(std::max(grid, 1)), dim3(256), 0, s, W, H, bm, active,
                     count, st, ok);
  return hipGetLastError();
}

hipError_t launch_fuse_round(const View* views, const Params& prm, const int* active,
                             int* next_active, const int* count, int* next_count,
                             unsigned int round, unsigned long long* masks, uint8_t* st,
                             float* res, uint8_t* res_clr, uint8_t* ok, int grid,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_fuse_claim, dim3(grid), dim3(256), 0, s, views, prm, active, count,
                     next_count, round, masks, st);
  hipLaunchKernelGGL(k_fuse_decide<false>, dim3(grid), dim3(256), 0, s, views, prm, active,
                     count, round, masks, st, res, res_clr, ok);
  if (prm.n_nb + 1 > kSmall)
    hipLaunchKernelGGL(k_fuse_decide<true>, dim3(grid), dim3(256), 0, s, views, prm, active,
                       count, round, masks, st, res, res_clr, ok);
  hipLaunchKernelGGL(k_fuse_compact, dim3(grid), dim3(256), 0, s, active, count, st,
                     next_active, next_count);
  return hipGetLastError();
}

hipError_t launch_fuse_count(const uint8_t* ok, int npix, int* counts, hipStream_t s) {
  hipLaunchKernelGGL(k_fuse_count, dim3(blocks(npix)), dim3(256), 0, s, ok, npix, counts);
  return hipGetLastError();
}

hipError_t launch_fuse_scatter(const uint8_t* ok, int npix, const long long* offsets,
                               const float* res, const uint8_t* res_clr, float* points,
                               float* normals, uint8_t* colors, hipStream_t s) {
  hipLaunchKernelGGL(k_fuse_scatter, dim3(blocks(npix)), dim3(256), 0, s, ok, npix, offsets,
                     res, res_clr, points, normals, colors);
  return hipGetLastError();
}

hipError_t launch_bbox(const float* points, long long n, float* lohi, hipStream_t s) {
  hipLaunchKernelGGL(k_bbox_init, dim3(1), dim3(64), 0, s, lohi);
  const int grid = (int)std::min<long long>(std::max(blocks(n), 1), 2048);
  hipLaunchKernelGGL(k_bbox, dim3(grid), dim3(256), 0, s, points, n, lohi);
  hipLaunchKernelGGL(k_bbox_finish, dim3(1), dim3(64), 0, s, lohi);
  return hipGetLastError();
}

hipError_t launch_morton(const float* points, long long n, const float* lo, float quantum,
                         unsigned long long* codes, unsigned int* idx, hipStream_t s) {
  hipLaunchKernelGGL(k_morton, dim3(blocks(n)), dim3(256), 0, s, points, n, lo, quantum, codes,
                     idx);
  return hipGetLastError();
}

hipError_t launch_gather_sorted(const float* points, const unsigned int* idx, long long n,
                                float* sorted, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_sorted, dim3(blocks(n)), dim3(256), 0, s, points, idx, n,
                     sorted);
  return hipGetLastError();
}

hipError_t launch_level_count_all(const unsigned long long* codes, long long n,
                                  unsigned int* n_cells, hipStream_t s) {
  hipLaunchKernelGGL(k_level_count_all, dim3(blocks(n)), dim3(256), 0, s, codes, n, n_cells);
  return hipGetLastError();
}

hipError_t launch_level_build(const unsigned long long* codes, long long n, int level,
                              unsigned long long* keys, unsigned int* start,
                              unsigned int* cnt, unsigned int mask, hipStream_t s) {
  hipLaunchKernelGGL(k_level_insert, dim3(blocks(n)), dim3(256), 0, s, codes, n, level, keys,
                     start, mask);
  hipLaunchKernelGGL(k_level_finish, dim3(blocks(n)), dim3(256), 0, s, codes, n, level, keys,
                     start, cnt, mask);
  return hipGetLastError();
}

hipError_t launch_knn_mean(const Grid& g, const float* sorted, const unsigned int* idx,
                           long long n, int k, float* mean_knn, hipStream_t s) {
  hipLaunchKernelGGL(k_knn_mean, dim3(blocks(n)), dim3(256), 0, s, g, sorted, idx, n, k,
                     mean_knn);
  return hipGetLastError();
}

hipError_t launch_knn_list(const Grid& g, const float* sorted, const unsigned int* idx,
                           long long n, int k, int* out_idx, float* out_d2, hipStream_t s) {
  hipLaunchKernelGGL(k_knn_list, dim3(blocks(n)), dim3(256), 0, s, g, sorted, idx, n, k, out_idx,
                     out_d2);
  return hipGetLastError();
}

hipError_t launch_sor_count(const float* mean_knn, long long n, float threshold, int* counts,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_sor_count, dim3(blocks(n)), dim3(256), 0, s, mean_knn, n, threshold,
                     counts);
  return hipGetLastError();
}

hipError_t launch_sor_scatter(const float* mean_knn, long long n, float threshold,
                              const long long* offsets, const float* points,
                              const float* normals, const uint8_t* colors, float* out_points,
                              float* out_normals, uint8_t* out_colors, hipStream_t s) {
  hipLaunchKernelGGL(k_sor_scatter, dim3(blocks(n)), dim3(256), 0, s, mean_knn, n, threshold,
                     offsets, points, normals, colors, out_points, out_normals, out_colors);
  return hipGetLastError();
}

hipError_t sort_codes(unsigned long long* keys_in, unsigned long long* keys_out,
                      unsigned int* vals_in, unsigned int* vals_out, int n, void* temp,
                      size_t* temp_bytes, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, keys_in, keys_out, vals_in,
                                            vals_out, n, 0, 63, s);
}

}  // namespace fu
