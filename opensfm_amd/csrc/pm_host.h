// Host-side helpers shared by the estimator, cleaner and batch entry points
// that touch the HIP runtime: error mapping, device buffers, device checks.
// The HIP-free part (error strings, FP64 3x3 algebra in the reference's
// order, the bilateral weight table, view descriptions) is pm_math.h.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <cmath>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/opensfm_amd.h"
#include "pm_math.h"
#include "pm_types.h"

namespace pmhost {


struct HipError {
  hipError_t err;
  const char* what;
};

#define HIP_CHECK(expr)                                   \
  do {                                                    \
    hipError_t e_ = (expr);                               \
    if (e_ != hipSuccess) throw HipError{e_, #expr};      \
  } while (0)

// Runs `fn`, mapping HIP failures and C++ exceptions to status codes.
template <typename F>
int guarded(F&& fn) {
  try {
    g_last_error.clear();
    return fn();
  } catch (const HipError& e) {
    return fail(PM_ERR_RUNTIME, std::string("HIP error '") +
                                    hipGetErrorString(e.err) + "' in " + e.what);
  } catch (const std::bad_alloc&) {
    return fail(PM_ERR_RUNTIME, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(PM_ERR_RUNTIME, e.what());
  }
}

// ---- device memory RAII ---------------------------------------------------

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t n) {
    if (n <= bytes && p) return;
    release();
    HIP_CHECK(hipMalloc(&p, n ? n : 1));
    bytes = n;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Host-to-device upload of a width x height u8 image with row stride
// `stride` on `stream`.  A pitched copy from pageable memory goes row by row
// (about 8 ms for 480 rows), so contiguous rows go as one transfer and
// strided ones are packed on the host first.  The caller synchronises the
// stream before the host image may change.
inline void upload_image(void* dst, const uint8_t* src, int width, int height, int stride,
                         hipStream_t stream) {
  const size_t n = (size_t)width * height;
  if (stride == width) {
    HIP_CHECK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, stream));
    return;
  }
  std::vector<uint8_t> packed(n);
  for (int y = 0; y < height; ++y)
    memcpy(packed.data() + (size_t)y * width, src + (size_t)y * stride, (size_t)width);
  HIP_CHECK(hipMemcpyAsync(dst, packed.data(), n, hipMemcpyHostToDevice, stream));
  HIP_CHECK(hipStreamSynchronize(stream));  // before `packed` goes away
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    HIP_CHECK(hipGetDevice(&prev));
    if (prev != dev) HIP_CHECK(hipSetDevice(dev));
  }
  ~DeviceGuard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0)
      hipSetDevice(prev);
  }
};

inline int check_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(PM_ERR_RUNTIME, "no HIP device available");
  if (device < 0 || device >= n)
    return fail(PM_ERR_RANGE, "device index " + std::to_string(device) +
                                  " out of range [0, " + std::to_string(n) + ")");
  return PM_OK;
}


}  // namespace pmhost
