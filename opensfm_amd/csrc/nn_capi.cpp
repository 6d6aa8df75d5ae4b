// C ABI of the initialiser network's kernels (pm_nn_* in
// include/opensfm_amd.h): shape and alignment checks, then one launch on the
// caller's stream.  Pointers are device pointers owned by the caller.
#include <hip/hip_runtime.h>

#include <string>

#include "nn_kernels.h"
#include "pm_host.h"

using namespace pmhost;

namespace {

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

int launched(hipError_t e) {
  if (e != hipSuccess)
    return fail(PM_ERR_RUNTIME, std::string("HIP error '") + hipGetErrorString(e) + "'");
  return PM_OK;
}

}  // namespace

extern "C" {

int pm_nn_conv_kpad(int cin) { return (9 * cin + 31) / 32 * 32; }

int pm_nn_conv3x3(const void* x, int x_f32, int npix, int H, int W, int cin,
                  int xs, const void* w, int cout, int kp, const float* bias,
                  void* y, int y_f32, int ys, int yo, int accumulate, void* act,
                  int as, int ao, void* stream) {
  if (!x || !w || (!y && !act)) return fail(PM_ERR_INVALID, "null tensor");
  if (npix <= 0 || H <= 0 || W <= 0 || npix % (H * W) != 0)
    return fail(PM_ERR_INVALID, "npix must be a positive multiple of H * W");
  if (cin <= 0 || cin % 8 || xs < cin || xs % 8)
    return fail(PM_ERR_INVALID, "input channels and stride must be multiples of 8");
  if (kp != pm_nn_conv_kpad(cin)) return fail(PM_ERR_INVALID, "kp must be pm_nn_conv_kpad(cin)");
  if (cout <= 0) return fail(PM_ERR_INVALID, "cout must be positive");
  if (y && (yo < 0 || ys < yo + cout)) return fail(PM_ERR_INVALID, "output stride/offset");
  if (act && (ao < 0 || as < ao + cout)) return fail(PM_ERR_INVALID, "activation stride/offset");
  if (accumulate && (!y || !y_f32)) return fail(PM_ERR_INVALID, "accumulate needs an FP32 output");
  if (!aligned16(x) || !aligned16(w)) return fail(PM_ERR_INVALID, "x and w must be 16-byte aligned");
  return launched(nn::conv3x3(x, x_f32, npix, H, W, cin, xs, w, cout, kp, bias, y, y_f32, ys,
                              yo, accumulate, act, as, ao, (hipStream_t)stream));
}

int pm_nn_conv3x3_wgrad(const void* x, int npix, int H, int W, int cin, int xs,
                        const float* dz, int cout, int dzs, float* dw, int kp,
                        float* db, void* stream) {
  if (!x || !dz || !dw) return fail(PM_ERR_INVALID, "null tensor");
  if (npix <= 0 || H <= 0 || W <= 0 || npix % (H * W) != 0)
    return fail(PM_ERR_INVALID, "npix must be a positive multiple of H * W");
  if (cin <= 0 || cin % 8 || xs < cin) return fail(PM_ERR_INVALID, "input channels");
  if (kp != pm_nn_conv_kpad(cin)) return fail(PM_ERR_INVALID, "kp must be pm_nn_conv_kpad(cin)");
  if (cout <= 0 || dzs < cout) return fail(PM_ERR_INVALID, "gradient channels");
  return launched(nn::conv3x3_wgrad(x, npix, H, W, cin, xs, dz, cout, dzs, dw, kp, db,
                                    (hipStream_t)stream));
}

int pm_nn_weight_flip(const void* w, int cout, int cin, int kp, void* wt,
                      int kpt, void* stream) {
  if (!w || !wt) return fail(PM_ERR_INVALID, "null tensor");
  if (kp != pm_nn_conv_kpad(cin) || kpt != pm_nn_conv_kpad(cout) || cout % 8)
    return fail(PM_ERR_INVALID, "flip needs kp = kpad(cin), kpt = kpad(cout), cout % 8 == 0");
  return launched(nn::weight_flip(w, cout, cin, kp, wt, kpt, (hipStream_t)stream));
}

int pm_nn_lrelu_backward(const float* da, int das, int dao, const void* a,
                         int as, int ao, float* dz, int npix, int c,
                         void* stream) {
  if (!da || !a || !dz) return fail(PM_ERR_INVALID, "null tensor");
  if (npix <= 0 || c <= 0 || das < dao + c || as < ao + c)
    return fail(PM_ERR_INVALID, "shape");
  return launched(nn::lrelu_backward(da, das, dao, a, as, ao, dz, npix, c, (hipStream_t)stream));
}

int pm_nn_maxpool2(const void* x, int xs, int xo, int N, int H, int W, int c,
                   void* y, uint8_t* idx, void* stream) {
  if (!x || !y || !idx) return fail(PM_ERR_INVALID, "null tensor");
  if (N <= 0 || H <= 0 || W <= 0 || H % 2 || W % 2 || xs < xo + c)
    return fail(PM_ERR_INVALID, "max pool needs even H, W");
  return launched(nn::maxpool2(x, xs, xo, N, H, W, c, y, idx, (hipStream_t)stream));
}

int pm_nn_maxpool2_backward(const float* dy, const uint8_t* idx, int N, int H,
                            int W, int c, float* dx, int dxs, int dxo,
                            void* stream) {
  if (!dy || !idx || !dx) return fail(PM_ERR_INVALID, "null tensor");
  if (N <= 0 || H % 2 || W % 2 || dxs < dxo + c) return fail(PM_ERR_INVALID, "shape");
  return launched(nn::maxpool2_backward(dy, idx, N, H, W, c, dx, dxs, dxo, (hipStream_t)stream));
}

int pm_nn_upsample2(const void* x, int N, int h, int w, int c, void* y, int ys,
                    int yo, void* stream) {
  if (!x || !y) return fail(PM_ERR_INVALID, "null tensor");
  if (N <= 0 || h <= 0 || w <= 0 || ys < yo + c) return fail(PM_ERR_INVALID, "shape");
  return launched(nn::upsample2(x, N, h, w, c, y, ys, yo, (hipStream_t)stream));
}

int pm_nn_upsample2_backward(const float* dy, int dys, int dyo, int N, int h,
                             int w, int c, float* dx, void* stream) {
  if (!dy || !dx) return fail(PM_ERR_INVALID, "null tensor");
  if (N <= 0 || h <= 0 || w <= 0 || dys < dyo + c) return fail(PM_ERR_INVALID, "shape");
  return launched(nn::upsample2_backward(dy, dys, dyo, N, h, w, c, dx, (hipStream_t)stream));
}

int pm_nn_loss(const float* out, int os, const float* target, int npix,
               int n_valid, float w_normal, float w_conf, float* dout,
               float* sums, void* stream) {
  if (!out || !target || !dout || !sums) return fail(PM_ERR_INVALID, "null tensor");
  if (npix <= 0 || os < 4) return fail(PM_ERR_INVALID, "shape");
  const float inv_valid = n_valid > 0 ? 1.0f / (float)n_valid : 0.0f;
  return launched(nn::loss(out, os, target, npix, w_normal, w_conf, inv_valid,
                           1.0f / (float)npix, dout, sums, (hipStream_t)stream));
}

int pm_nn_adam(float* w, void* w16, float* m, float* v, const float* g, int n,
               float lr, float beta1, float beta2, float eps, int step,
               void* stream) {
  if (!w || !m || !v || !g) return fail(PM_ERR_INVALID, "null tensor");
  if (n <= 0 || step < 1) return fail(PM_ERR_INVALID, "n > 0 and step >= 1");
  return launched(nn::adam(w, w16, m, v, g, n, lr, beta1, beta2, eps, step, (hipStream_t)stream));
}

int pm_nn_to_bf16(const float* x, void* y, int n, void* stream) {
  if (!x || !y) return fail(PM_ERR_INVALID, "null tensor");
  if (n <= 0) return fail(PM_ERR_INVALID, "n > 0");
  return launched(nn::to_bf16(x, y, n, (hipStream_t)stream));
}

}  // extern "C"
