// Host side of the depthmap fuser's C ABI (include/opensfm_amd.h):
// dense::DepthmapFuser (opensfm/src/lib/dense/fuser.h:28-75,
// src/fuser.cc:37-480) with its views resident on the device.
//
// Built with -ffp-contract=off: the float cameras (K^-1 by Eigen's
// cofactor form) are made here with the reference's operation order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "../../include/opensfm_amd.h"
#include "fuse_kernels.h"
#include "knn_grid.h"
#include "pm_host.h"

using namespace pmhost;

namespace {

float cofactor(const float* m, int i, int j) {
  const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return m[3 * i1 + j1] * m[3 * i2 + j2] - m[3 * i1 + j2] * m[3 * i2 + j1];
}

// Eigen's fixed-size 3x3 inverse: cofactors of column 0, det = c00 m00 +
// (c10 m10 + c20 m20), inverse(i, j) = cofactor(j, i) * (1 / det).
void inverse3f(const float* m, float* r) {
  const float c00 = cofactor(m, 0, 0), c10 = cofactor(m, 1, 0), c20 = cofactor(m, 2, 0);
  const float det = c00 * m[0] + (c10 * m[3] + c20 * m[6]);
  const float invdet = 1.0f / det;
  r[3] = cofactor(m, 0, 1) * invdet;
  r[4] = cofactor(m, 1, 1) * invdet;
  r[6] = cofactor(m, 0, 2) * invdet;
  r[5] = cofactor(m, 2, 1) * invdet;
  r[7] = cofactor(m, 1, 2) * invdet;
  r[8] = cofactor(m, 2, 2) * invdet;
  r[0] = c00 * invdet;
  r[1] = c10 * invdet;
  r[2] = c20 * invdet;
}

void make_camera(const double* K, const double* R, const double* t, fu::Camera& c) {
  for (int k = 0; k < 9; ++k) {
    c.K[k] = (float)K[k];
    c.R[k] = (float)R[k];
  }
  for (int k = 0; k < 3; ++k) c.t[k] = (float)t[k];
  inverse3f(c.K, c.Kinv);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c.Rinv[3 * i + j] = c.R[3 * j + i];
}

// Grows b to at least `need` bytes keeping its first `used` bytes.
void grow(DevBuf& b, size_t need, size_t used, hipStream_t s) {
  if (b.p && need <= b.bytes) return;
  const size_t cap = std::max(need, b.bytes * 2);
  void* np = nullptr;
  HIP_CHECK(hipMalloc(&np, cap ? cap : 1));
  if (used) {
    HIP_CHECK(hipMemcpyAsync(np, b.p, used, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
  b.release();
  b.p = np;
  b.bytes = cap;
}

void swap_buf(DevBuf& a, DevBuf& b) {
  std::swap(a.p, b.p);
  std::swap(a.bytes, b.bytes);
}

struct FuHostView {
  int width = 0, height = 0;
  fu::Camera cam;
  bool has_mask = false, primary = true;
  std::vector<int> neighbors;
  DevBuf depth, normal, color, mask, fused, claim;
};

}  // namespace

struct pm_fuser {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  // defaults of fuser.cc:37-46
  int min_num_consistent = 3;
  float max_reproj_error_sq = 2.0f * 2.0f;
  float max_depth_error = 0.01f;
  float min_cos_normal_error = cosf(10.0f * (float)3.14159265358979323846 / 180.0f);
  int border_margin = 5;
  int num_threads = 4;
  int sor_knn = 12;
  float sor_stddev_factor = 2.5f;
  float behind_depth_factor = 0.3f;
  std::vector<std::unique_ptr<FuHostView>> views;
  unsigned int round = 0;
  // work buffers
  DevBuf table, active[2], counts, masks, st, ok, res, res_clr, blk, offs;
  DevBuf points, normals, colors;        // fused points (before / after SOR)
  DevBuf knn_pts, knn_nrm, knn_clr, mean_knn;
  fu::KnnGrid grid;
  int64_t count = -1, fused_before_sor = 0;
  int64_t rounds = 0;
  float fuse_ms = 0, sor_ms = 0;
  ~pm_fuser() {
    for (auto& e : ev)
      if (e) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
  }
};

namespace {

// Count per 256-item block, host exclusive scan, device offsets.
template <typename CountFn>
long long block_offsets(pm_fuser* f, long long n, CountFn launch_count) {
  const long long nblk = (n + 255) / 256;
  f->blk.ensure(std::max<long long>(nblk, 1) * sizeof(int));
  f->offs.ensure(std::max<long long>(nblk, 1) * sizeof(long long));
  HIP_CHECK(launch_count(f->blk.as<int>()));
  std::vector<int> counts(nblk);
  if (nblk)
    HIP_CHECK(hipMemcpyAsync(counts.data(), f->blk.p, nblk * sizeof(int), hipMemcpyDeviceToHost,
                             f->stream));
  HIP_CHECK(hipStreamSynchronize(f->stream));
  std::vector<long long> offsets(nblk);
  long long total = 0;
  for (long long b = 0; b < nblk; ++b) {
    offsets[b] = total;
    total += counts[b];
  }
  if (nblk)
    HIP_CHECK(hipMemcpyAsync(f->offs.p, offsets.data(), nblk * sizeof(long long),
                             hipMemcpyHostToDevice, f->stream));
  return total;
}

// One primary view: rounds of the parallel greedy, then its fused points
// appended in row-major order.  Returns the new total.
long long fuse_view(pm_fuser* f, int vi, long long total) {
  const FuHostView& A = *f->views[vi];
  const int W = A.width, H = A.height;
  const int npix = W * H;
  fu::Params prm;
  memset(&prm, 0, sizeof(prm));
  prm.ref = vi;
  prm.bm = std::max(0, f->border_margin);
  prm.min_consistent = f->min_num_consistent;
  prm.max_reproj_sq = f->max_reproj_error_sq;
  prm.max_depth_error = f->max_depth_error;
  prm.min_cos = f->min_cos_normal_error;
  prm.behind = f->behind_depth_factor;
  const int nv = (int)f->views.size();
  for (int id : A.neighbors)
    if (id >= 0 && id < nv) prm.nb[prm.n_nb++] = id;  // fuser.cc:214-216
  const int bw = W - 2 * prm.bm, bh = H - 2 * prm.bm;
  long long live = (bw > 0 && bh > 0) ? (long long)bw * bh : 0;
  HIP_CHECK(fu::launch_fuse_init(W, H, prm.bm, f->active[0].as<int>(), f->counts.as<int>(),
                                 f->st.as<uint8_t>(), f->ok.as<uint8_t>(), f->stream));
  int cur = 0;
  int host_count = 0;
  int batch = 2;  // rounds between reads of the live count; grows while rounds continue
  while (live > 0) {
    const int grid = (int)std::min<long long>((live + 255) / 256, 4096);
    for (int j = 0; j < batch; ++j) {
      ++f->round;
      ++f->rounds;
      HIP_CHECK(fu::launch_fuse_round(
          f->table.as<fu::View>(), prm, f->active[cur].as<int>(), f->active[cur ^ 1].as<int>(),
          f->counts.as<int>() + cur, f->counts.as<int>() + (cur ^ 1), f->round,
          f->masks.as<unsigned long long>(), f->st.as<uint8_t>(), f->res.as<float>(),
          f->res_clr.as<uint8_t>(), f->ok.as<uint8_t>(), grid, f->stream));
      cur ^= 1;
    }
    HIP_CHECK(hipMemcpyAsync(&host_count, f->counts.as<int>() + cur, sizeof(int),
                             hipMemcpyDeviceToHost, f->stream));
    HIP_CHECK(hipStreamSynchronize(f->stream));
    live = host_count;
    batch = std::min(batch * 2, 16);
  }
  const uint8_t* ok = f->ok.as<uint8_t>();
  const long long added = block_offsets(f, npix, [&](int* counts) {
    return fu::launch_fuse_count(ok, npix, counts, f->stream);
  });
  if (added == 0) return total;
  // offsets are relative to this view's first point
  const size_t need = (size_t)(total + added);
  grow(f->points, need * 3 * sizeof(float), (size_t)total * 3 * sizeof(float), f->stream);
  grow(f->normals, need * 3 * sizeof(float), (size_t)total * 3 * sizeof(float), f->stream);
  grow(f->colors, need * 3, (size_t)total * 3, f->stream);
  HIP_CHECK(fu::launch_fuse_scatter(ok, npix, f->offs.as<long long>(), f->res.as<float>(),
                                    f->res_clr.as<uint8_t>(), f->points.as<float>() + 3 * total,
                                    f->normals.as<float>() + 3 * total,
                                    f->colors.as<uint8_t>() + 3 * total, f->stream));
  return total + added;
}

// Statistical outlier removal (fuser.cc:369-478).  Returns the kept count;
// the kept points are in knn_pts / knn_nrm / knn_clr.
long long remove_outliers(pm_fuser* f, long long n) {
  const int K = f->sor_knn;
  hipStream_t s = f->stream;
  f->grid.build(f->points.as<float>(), n, s);
  const unsigned int* order = f->grid.order();
  f->mean_knn.ensure(n * sizeof(float));
  HIP_CHECK(fu::launch_knn_mean(f->grid.g, f->grid.sorted_points(), order, n, K,
                                f->mean_knn.as<float>(), s));
  std::vector<float> mk(n);
  HIP_CHECK(hipMemcpyAsync(mk.data(), f->mean_knn.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  // fuser.cc:452-462, in index order in double
  double sum_d = 0, sum_d2 = 0;
  for (long long i = 0; i < n; ++i) {
    const double d = mk[i];
    if (d < 1e20f) {
      sum_d += d;
      sum_d2 += d * d;
    }
  }
  const double mean_d = sum_d / n;
  const double var_d = sum_d2 / n - mean_d * mean_d;
  const double std_d = sqrt(std::max(0.0, var_d));
  const float thr = (float)(mean_d + f->sor_stddev_factor * std_d);
  const float* mkd = f->mean_knn.as<float>();
  const long long kept = block_offsets(f, n, [&](int* counts) {
    return fu::launch_sor_count(mkd, n, thr, counts, s);
  });
  f->knn_pts.ensure(std::max<long long>(kept, 1) * 3 * sizeof(float));
  f->knn_nrm.ensure(std::max<long long>(kept, 1) * 3 * sizeof(float));
  f->knn_clr.ensure(std::max<long long>(kept, 1) * 3);
  HIP_CHECK(fu::launch_sor_scatter(mkd, n, thr, f->offs.as<long long>(), f->points.as<float>(),
                                   f->normals.as<float>(), f->colors.as<uint8_t>(),
                                   f->knn_pts.as<float>(), f->knn_nrm.as<float>(),
                                   f->knn_clr.as<uint8_t>(), s));
  HIP_CHECK(hipStreamSynchronize(s));
  return kept;
}

}  // namespace

extern "C" {

int pm_fuser_create(int device, pm_fuser** out) {
  if (!out) return fail(PM_ERR_INVALID, "null output handle");
  *out = nullptr;
  int rc = check_device(device);
  if (rc) return rc;
  return guarded([&] {
    std::unique_ptr<pm_fuser> f(new pm_fuser());
    f->device = device;
    DeviceGuard dg(device);
    HIP_CHECK(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
    for (auto& e : f->ev) HIP_CHECK(hipEventCreate(&e));
    *out = f.release();
    return PM_OK;
  });
}

int pm_fuser_destroy(pm_fuser* f) {
  if (!f) return PM_OK;
  return guarded([&] {
    DeviceGuard dg(f->device);
    delete f;
    return PM_OK;
  });
}

int pm_fuser_set_min_num_consistent(pm_fuser* f, int n) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  f->min_num_consistent = n;
  return PM_OK;
}

int pm_fuser_set_max_reproj_error(pm_fuser* f, float px) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  f->max_reproj_error_sq = px * px;
  return PM_OK;
}

int pm_fuser_set_max_depth_error(pm_fuser* f, float ratio) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  f->max_depth_error = ratio;
  return PM_OK;
}

int pm_fuser_set_max_normal_error(pm_fuser* f, float degrees) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  f->min_cos_normal_error = cosf(degrees * (float)3.14159265358979323846 / 180.0f);
  return PM_OK;
}

int pm_fuser_set_border_margin(pm_fuser* f, int px) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  f->border_margin = std::max(0, px);
  return PM_OK;
}

int pm_fuser_set_num_threads(pm_fuser* f, int n) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  f->num_threads = std::max(1, n);
  return PM_OK;
}

int pm_fuser_set_sor_params(pm_fuser* f, int knn, float stddev_factor) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  f->sor_knn = std::max(0, knn);
  f->sor_stddev_factor = stddev_factor;
  return PM_OK;
}

int pm_fuser_set_behind_depth_factor(pm_fuser* f, float v) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  f->behind_depth_factor = std::min(std::max(v, 0.0f), 1.0f);
  return PM_OK;
}

int pm_fuser_add_view(pm_fuser* f, const double* K, const double* R, const double* t,
                      const float* depth, const float* normal, const uint8_t* color,
                      const uint8_t* mask, int width, int height, const int32_t* neighbor_ids,
                      int n_neighbors, int primary) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  if (!K || !R || !t || !depth || !normal || !color)
    return fail(PM_ERR_INVALID, "null K, R, t, depth, normal or color");
  if (width < 1 || height < 1) return fail(PM_ERR_INVALID, "empty depth map");
  if ((int64_t)width * height > (int64_t)INT32_MAX / 8)
    return fail(PM_ERR_INVALID, "depth map larger than 2^28 pixels");
  if (n_neighbors < 0 || (n_neighbors > 0 && !neighbor_ids))
    return fail(PM_ERR_INVALID, "neighbour list");
  return guarded([&] {
    std::unique_ptr<FuHostView> v(new FuHostView());
    v->width = width;
    v->height = height;
    make_camera(K, R, t, v->cam);
    v->primary = primary != 0;
    v->has_mask = mask != nullptr;
    v->neighbors.assign(neighbor_ids, neighbor_ids + n_neighbors);
    DeviceGuard dg(f->device);
    const size_t np = (size_t)width * height;
    hipStream_t s = f->stream;
    v->depth.ensure(np * sizeof(float));
    v->normal.ensure(np * 3 * sizeof(float));
    v->color.ensure(np * 3);
    v->fused.ensure(np);
    v->claim.ensure(np * sizeof(unsigned long long));
    HIP_CHECK(hipMemcpyAsync(v->depth.p, depth, np * sizeof(float), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(v->normal.p, normal, np * 3 * sizeof(float), hipMemcpyHostToDevice,
                             s));
    HIP_CHECK(hipMemcpyAsync(v->color.p, color, np * 3, hipMemcpyHostToDevice, s));
    if (mask) {
      v->mask.ensure(np);
      HIP_CHECK(hipMemcpyAsync(v->mask.p, mask, np, hipMemcpyHostToDevice, s));
    }
    HIP_CHECK(hipMemsetAsync(v->fused.p, 0, np, s));
    HIP_CHECK(hipMemsetAsync(v->claim.p, 0, np * sizeof(unsigned long long), s));
    HIP_CHECK(hipStreamSynchronize(s));
    f->views.push_back(std::move(v));
    return PM_OK;
  });
}

int pm_fuser_fuse(pm_fuser* f, int64_t* count) {
  if (!f || !count) return fail(PM_ERR_INVALID, "null argument");
  const int nv = (int)f->views.size();
  for (int v = 0; v < nv; ++v) {
    int valid = 0;
    for (int id : f->views[v]->neighbors) valid += (id >= 0 && id < nv);
    if (f->views[v]->primary && valid > fu::kMaxNeighbours)
      return fail(PM_ERR_INVALID, "view " + std::to_string(v) + " has " + std::to_string(valid) +
                                      " neighbours; at most " +
                                      std::to_string(fu::kMaxNeighbours) + " are supported");
  }
  if (f->sor_knn > fu::kMaxKnn)
    return fail(PM_ERR_INVALID, "outlier removal with k above " + std::to_string(fu::kMaxKnn) +
                                    " is not supported");
  return guarded([&] {
    DeviceGuard dg(f->device);
    hipStream_t s = f->stream;
    f->count = 0;
    f->fused_before_sor = 0;
    f->fuse_ms = f->sor_ms = 0;
    f->rounds = 0;
    *count = 0;
    if (nv == 0) return PM_OK;
    std::vector<fu::View> table(nv);
    int max_pix = 1;
    bool any_primary = false;
    for (int v = 0; v < nv; ++v) {
      FuHostView& H = *f->views[v];
      fu::View& T = table[v];
      T.depth = H.depth.as<float>();
      T.normal = H.normal.as<float>();
      T.color = H.color.as<uint8_t>();
      T.mask = H.has_mask ? H.mask.as<uint8_t>() : nullptr;
      T.fused = H.fused.as<uint8_t>();
      T.claim = H.claim.as<unsigned long long>();
      T.width = H.width;
      T.height = H.height;
      T.cam = H.cam;
      if (H.primary) {
        any_primary = true;
        max_pix = std::max(max_pix, H.width * H.height);
      }
    }
    if (!any_primary) return PM_OK;  // fuser.cc:143-145: nothing, no outlier removal
    f->table.ensure(nv * sizeof(fu::View));
    HIP_CHECK(hipMemcpyAsync(f->table.p, table.data(), nv * sizeof(fu::View),
                             hipMemcpyHostToDevice, s));
    for (auto& a : f->active) a.ensure((size_t)max_pix * sizeof(int));
    f->counts.ensure(2 * sizeof(int));
    f->masks.ensure((size_t)max_pix * sizeof(unsigned long long));
    f->st.ensure(max_pix);
    f->ok.ensure(max_pix);
    f->res.ensure((size_t)max_pix * 6 * sizeof(float));
    f->res_clr.ensure((size_t)max_pix * 3);
    HIP_CHECK(hipEventRecord(f->ev[0], s));
    long long total = 0;
    for (int v = 0; v < nv; ++v)
      if (f->views[v]->primary) total = fuse_view(f, v, total);
    HIP_CHECK(hipEventRecord(f->ev[1], s));
    f->fused_before_sor = total;
    long long kept = total;
    const bool sor = f->sor_knn > 0 && total > (long long)f->sor_knn + 1;
    if (sor) kept = remove_outliers(f, total);
    HIP_CHECK(hipEventRecord(f->ev[2], s));
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipEventElapsedTime(&f->fuse_ms, f->ev[0], f->ev[1]));
    if (sor) HIP_CHECK(hipEventElapsedTime(&f->sor_ms, f->ev[1], f->ev[2]));
    if (sor) {
      swap_buf(f->points, f->knn_pts);
      swap_buf(f->normals, f->knn_nrm);
      swap_buf(f->colors, f->knn_clr);
    }
    f->count = kept;
    *count = kept;
    return PM_OK;
  });
}

int pm_fuser_download(pm_fuser* f, float* points, float* normals, uint8_t* colors) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  if (f->count < 0) return fail(PM_ERR_INVALID, "download before fuse");
  if (f->count == 0) return PM_OK;
  if (!points || !normals || !colors) return fail(PM_ERR_INVALID, "null output");
  return guarded([&] {
    DeviceGuard dg(f->device);
    const size_t m = (size_t)f->count;
    HIP_CHECK(hipMemcpyAsync(points, f->points.p, m * 3 * sizeof(float), hipMemcpyDeviceToHost,
                             f->stream));
    HIP_CHECK(hipMemcpyAsync(normals, f->normals.p, m * 3 * sizeof(float),
                             hipMemcpyDeviceToHost, f->stream));
    HIP_CHECK(hipMemcpyAsync(colors, f->colors.p, m * 3, hipMemcpyDeviceToHost, f->stream));
    HIP_CHECK(hipStreamSynchronize(f->stream));
    return PM_OK;
  });
}

int pm_fuser_stats(pm_fuser* f, int64_t* fused_before_sor, int64_t* rounds, float* fuse_ms,
                   float* sor_ms) {
  if (!f) return fail(PM_ERR_INVALID, "null fuser");
  if (fused_before_sor) *fused_before_sor = f->fused_before_sor;
  if (rounds) *rounds = f->rounds;
  if (fuse_ms) *fuse_ms = f->fuse_ms;
  if (sor_ms) *sor_ms = f->sor_ms;
  return PM_OK;
}

}  // extern "C"
