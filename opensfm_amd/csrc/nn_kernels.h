// Launchers of the initialiser network's kernels (nn_kernels.hip).  All
// pointers are device pointers; shapes are checked by the C ABI (nn_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nn {

hipError_t conv3x3(const void* x, int x_f32, int npix, int H, int W, int cin,
                   int xs, const void* w, int cout, int kp, const float* bias,
                   void* y, int y_f32, int ys, int yo, int accumulate, void* act,
                   int as, int ao, hipStream_t s);
hipError_t conv3x3_wgrad(const void* x, int npix, int H, int W, int cin, int xs,
                         const float* dz, int cout, int dzs, float* dw, int kp,
                         float* db, hipStream_t s);
hipError_t weight_flip(const void* w, int cout, int cin, int kp, void* wt,
                       int kpt, hipStream_t s);
hipError_t lrelu_backward(const float* da, int das, int dao, const void* a,
                          int as, int ao, float* dz, int npix, int c,
                          hipStream_t s);
hipError_t maxpool2(const void* x, int xs, int xo, int N, int H, int W, int c,
                    void* y, uint8_t* idx, hipStream_t s);
hipError_t maxpool2_backward(const float* dy, const uint8_t* idx, int N, int H,
                             int W, int c, float* dx, int dxs, int dxo,
                             hipStream_t s);
hipError_t upsample2(const void* x, int N, int h, int w, int c, void* y, int ys,
                     int yo, hipStream_t s);
hipError_t upsample2_backward(const float* dy, int dys, int dyo, int N, int h,
                              int w, int c, float* dx, hipStream_t s);
hipError_t loss(const float* out, int os, const float* tgt, int npix, float wn,
                float wc, float inv_valid, float inv_all, float* dout,
                float* sums, hipStream_t s);
hipError_t adam(float* w, void* w16, float* m, float* v, const float* g, int n,
                float lr, float b1, float b2, float eps, int step, hipStream_t s);
hipError_t to_bf16(const float* x, void* y, int n, hipStream_t s);

}  // namespace nn
