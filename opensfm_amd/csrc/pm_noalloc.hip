// DepthmapCleaner::Clean over depthmaps that are already in device memory,
// with no device allocation: the per-view table the clean kernel reads
// lives in a file-scope device array instead of library-allocated memory.
// For callers that cannot allocate (code under graph capture, serving
// loops, harnesses that forbid hipMalloc).  No counterpart in the
// reference; the arithmetic is the clean kernel of DepthmapCleaner::Clean
// (depthmap.cc:524-555), shared with pm_cleaner_clean.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/opensfm_amd.h"
#include "pm_host.h"
#include "pm_kernels.h"
#include "pm_types.h"

using namespace pmhost;

// Layout of the scratch array: the inverse pose of view 0, then the view
// table (views[0] is the map being cleaned).
constexpr size_t kRefOffset = 0;
constexpr size_t kViewsOffset = 256;
constexpr size_t kScratchBytes = kViewsOffset + sizeof(pm::ClView) * PM_NOALLOC_MAX_VIEWS;
static_assert(sizeof(pm::ClRef) <= kViewsOffset, "ClRef must fit before the view table");
static_assert(kViewsOffset % alignof(pm::ClView) == 0, "view table alignment");

// One copy per device, held until the process ends (12 KiB for 64 views).  In host code
// the name only stands in for the symbol: the device address comes from
// hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
void* g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current),
// looked up once per device.
hipError_t scratch_ptr(int device, void** out) {
  std::call_once(g_once[device], [device] {
    void* p = nullptr;
    g_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_ptr[device] = g_err[device] == hipSuccess ? p : nullptr;
  });
  *out = g_ptr[device];
  return g_err[device];
}

}  // namespace

extern "C" int pm_cleaner_clean_noalloc(int device, int n_views, const double* K,
                                        const double* R, const double* t,
                                        const float* const* depths, const int32_t* widths,
                                        const int32_t* heights, float same_depth_threshold,
                                        int min_consistent_views, float* clean_depth,
                                        void* stream) {
  if (n_views < 1) return fail(PM_ERR_INVALID, "need at least one view");
  if (n_views > PM_NOALLOC_MAX_VIEWS)
    return fail(PM_ERR_RANGE, "pm_cleaner_clean_noalloc: " + std::to_string(n_views) +
                                  " views exceed the maximum of " +
                                  std::to_string(PM_NOALLOC_MAX_VIEWS));
  if (!K || !R || !t || !depths || !widths || !heights || !clean_depth)
    return fail(PM_ERR_INVALID, "null argument");
  if (device < 0 || device >= kMaxDevices) return fail(PM_ERR_RANGE, "device index out of range");
  int rc = check_device(device);
  if (rc) return rc;
  for (int v = 0; v < n_views; ++v) {
    if (!depths[v]) return fail(PM_ERR_INVALID, "null depthmap pointer");
    if (widths[v] < 1 || heights[v] < 1) return fail(PM_ERR_INVALID, "empty depthmap");
    if (!all_finite(K + 9 * v, 9) || !all_finite(R + 9 * v, 9) || !all_finite(t + 3 * v, 3))
      return fail(PM_ERR_INVALID, "K, R and t must be finite");
  }
  return guarded([&] {
    DeviceGuard dg(device);
    void* base = nullptr;
    HIP_CHECK(scratch_ptr(device, &base));
    pm::ClView tab[PM_NOALLOC_MAX_VIEWS];
    for (int v = 0; v < n_views; ++v) {
      tab[v].depth = depths[v];
      tab[v].width = widths[v];
      tab[v].height = heights[v];
      memcpy(tab[v].K, K + 9 * v, sizeof(tab[v].K));
      memcpy(tab[v].R, R + 9 * v, sizeof(tab[v].R));
      memcpy(tab[v].t, t + 3 * v, sizeof(tab[v].t));
    }
    pm::ClRef ref;
    if (!inv3(K, ref.Kinv)) return fail(PM_ERR_INVALID, "K of view 0 is singular");
    mat3_transpose(R, ref.Rt);
    auto* dref = reinterpret_cast<pm::ClRef*>(static_cast<unsigned char*>(base) + kRefOffset);
    auto* dtab = reinterpret_cast<pm::ClView*>(static_cast<unsigned char*>(base) + kViewsOffset);
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_CHECK(hipMemcpyAsync(dref, &ref, sizeof(ref), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(dtab, tab, sizeof(pm::ClView) * n_views, hipMemcpyHostToDevice, s));
    HIP_CHECK(pm::launch_clean(dtab, dref, n_views, widths[0], heights[0], same_depth_threshold,
                               min_consistent_views, clean_depth, s));
    // the table copies read host memory that goes away on return
    HIP_CHECK(hipStreamSynchronize(s));
    return PM_OK;
  });
}
