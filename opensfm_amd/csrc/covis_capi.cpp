// Host side of the neighbour selection's C ABI (include/opensfm_amd.h):
// sfm::dense_helpers::BuildCovisibilityGraph and SelectNeighbors
// (opensfm/src/lib/sfm/dense_helpers.h:40-67, src/dense_helpers.cc) over
// plain arrays, computed on the device.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "../../include/opensfm_amd.h"
#include "covis_kernels.h"
#include "fuse_kernels.h"
#include "knn_grid.h"
#include "pm_host.h"

using namespace pmhost;

struct pm_covis {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // graph
  DevBuf coords, obs_off, obs_shot, flag, blk, offs, keep, fc, kidx, kd2, kdist, parent, changed,
      rank, sp[2], member[2], counts, point_sp, moff, members, sp_coord, obs_cnt, obs_pos,
      keys[2], index, vis, vcounts, voff, table, edges, weights, sort_tmp;
  fu::KnnGrid grid;
  int64_t n_points = 0, n_sp = -1, n_vis = 0, n_edges = 0;
  int n_shots = 0;
  float scale = 0.0f;
  // neighbours
  DevBuf origins, nsp_coord, nvoff, nvis, ntracks, pcount, pscore, psum, best, best_n, all_n,
      all_off, all;
  int nb_shots = -1, nb_num = 0;
  int64_t n_best = 0, n_all = 0;
  float last_ms = 0;
  ~pm_covis() {
    for (auto& e : ev)
      if (e) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
  }
};

namespace {

// Block counts of flags, host exclusive scan, device offsets; returns the
// number flagged.
long long flag_offsets(pm_covis* c, const uint8_t* flag, long long n) {
  const long long nblk = (n + 255) / 256;
  c->blk.ensure(std::max<long long>(nblk, 1) * sizeof(int));
  c->offs.ensure(std::max<long long>(nblk, 1) * sizeof(long long));
  if (nblk == 0) return 0;
  HIP_CHECK(cv::launch_flag_count(flag, n, c->blk.as<int>(), c->stream));
  std::vector<int> counts(nblk);
  HIP_CHECK(hipMemcpyAsync(counts.data(), c->blk.p, nblk * sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  std::vector<long long> offsets(nblk);
  long long total = 0;
  for (long long b = 0; b < nblk; ++b) {
    offsets[b] = total;
    total += counts[b];
  }
  HIP_CHECK(hipMemcpyAsync(c->offs.p, offsets.data(), nblk * sizeof(long long),
                           hipMemcpyHostToDevice, c->stream));
  return total;
}

// Exclusive scan of device counts (n entries of T) into device offsets
// (n + 1 entries); returns the total.
template <typename T>
long long scan_counts(pm_covis* c, const T* dcounts, long long n, DevBuf& off) {
  std::vector<T> h(std::max<long long>(n, 1));
  if (n)
    HIP_CHECK(hipMemcpyAsync(h.data(), dcounts, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  std::vector<long long> o(n + 1);
  long long t = 0;
  for (long long i = 0; i < n; ++i) {
    o[i] = t;
    t += (long long)h[i];
  }
  o[n] = t;
  off.ensure((n + 1) * sizeof(long long));
  HIP_CHECK(hipMemcpyAsync(off.p, o.data(), (n + 1) * sizeof(long long), hipMemcpyHostToDevice,
                           c->stream));
  return t;
}

}  // namespace

extern "C" {

int pm_covis_create(int device, pm_covis** out) {
  if (!out) return fail(PM_ERR_INVALID, "null output handle");
  *out = nullptr;
  int rc = check_device(device);
  if (rc) return rc;
  return guarded([&] {
    std::unique_ptr<pm_covis> c(new pm_covis());
    c->device = device;
    DeviceGuard dg(device);
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIP_CHECK(hipEventCreate(&e));
    *out = c.release();
    return PM_OK;
  });
}

int pm_covis_destroy(pm_covis* c) {
  if (!c) return PM_OK;
  return guarded([&] {
    DeviceGuard dg(c->device);
    delete c;
    return PM_OK;
  });
}

int pm_covis_build(pm_covis* c, int64_t n_points, const double* coords, const int64_t* obs_off,
                   const int32_t* obs_shot, int n_shots, int fuse_knn,
                   double fuse_radius_factor, int64_t* n_super_points, int64_t* n_vis,
                   int64_t* n_edges) {
  if (!c || !n_super_points || !n_vis || !n_edges) return fail(PM_ERR_INVALID, "null argument");
  if (n_points < 0 || (n_points > 0 && (!coords || !obs_off || !obs_shot)))
    return fail(PM_ERR_INVALID, "null points or observations");
  if (n_points >= (int64_t)INT32_MAX) return fail(PM_ERR_INVALID, "too many points");
  if (n_shots < 0 || n_shots > cv::kMaxShots)
    return fail(PM_ERR_INVALID, "n_shots must be in [0, " + std::to_string(cv::kMaxShots) + "]");
  const int knn = std::max(1, fuse_knn);
  if (std::min(knn, (int)std::max<int64_t>(n_points - 1, 1)) > fu::kMaxKnn)
    return fail(PM_ERR_INVALID, "fuse_knn above " + std::to_string(fu::kMaxKnn) +
                                    " is not supported");
  return guarded([&] {
    DeviceGuard dg(c->device);
    hipStream_t s = c->stream;
    const long long n = n_points;
    c->n_points = n;
    c->n_shots = n_shots;
    c->n_sp = 0;
    c->n_vis = c->n_edges = 0;
    c->scale = 0.0f;
    *n_super_points = *n_vis = *n_edges = 0;
    c->point_sp.ensure(std::max<long long>(n, 1) * sizeof(int));
    if (n == 0) return PM_OK;
    HIP_CHECK(hipEventRecord(c->ev[0], s));
    const long long nobs = obs_off[n];
    c->coords.ensure(n * 3 * sizeof(double));
    c->obs_off.ensure((n + 1) * sizeof(long long));
    c->obs_shot.ensure(std::max<long long>(nobs, 1) * sizeof(int));
    HIP_CHECK(hipMemcpyAsync(c->coords.p, coords, n * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->obs_off.p, obs_off, (n + 1) * sizeof(long long),
                             hipMemcpyHostToDevice, s));
    if (nobs)
      HIP_CHECK(hipMemcpyAsync(c->obs_shot.p, obs_shot, nobs * sizeof(int), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemsetAsync(c->point_sp.p, 0xFF, n * sizeof(int), s));
    // 1. points seen by >= 2 processable shots, in input order
    c->flag.ensure(n);
    HIP_CHECK(cv::launch_valid(c->obs_off.as<long long>(), c->obs_shot.as<int>(), n, n_shots,
                               c->flag.as<uint8_t>(), s));
    const long long m = flag_offsets(c, c->flag.as<uint8_t>(), n);
    if (m == 0) {
      HIP_CHECK(hipStreamSynchronize(s));
      return PM_OK;
    }
    c->keep.ensure(m * sizeof(int));
    HIP_CHECK(cv::launch_flag_rank(c->flag.as<uint8_t>(), n, c->offs.as<long long>(), nullptr,
                                   c->keep.as<int>(), s));
    // 2. k nearest (float coordinates), scale
    const int k = std::max(1, std::min(fuse_knn, (int)(m - 1)));
    c->fc.ensure(m * 3 * sizeof(float));
    HIP_CHECK(cv::launch_gather_points(c->coords.as<double>(), c->keep.as<int>(), m,
                                       c->fc.as<float>(), s));
    c->kidx.ensure(m * k * sizeof(int));
    c->kd2.ensure(m * k * sizeof(float));
    c->kdist.ensure(m * sizeof(float));
    c->grid.build(c->fc.as<float>(), m, s);
    HIP_CHECK(fu::launch_knn_list(c->grid.g, c->grid.sorted_points(), c->grid.order(), m, k,
                                  c->kidx.as<int>(), c->kd2.as<float>(), s));
    HIP_CHECK(cv::launch_kdist(c->kidx.as<int>(), c->kd2.as<float>(), m, k, c->kdist.as<float>(),
                               s));
    std::vector<float> kd(m);
    HIP_CHECK(hipMemcpyAsync(kd.data(), c->kdist.p, m * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::nth_element(kd.begin(), kd.begin() + m / 2, kd.end());
    const double scale = (double)kd[m / 2];
    const double radius = fuse_radius_factor * scale;
    const float r2 = (float)(radius * radius);
    c->scale = (float)scale;
    // 3. union-find over the merge pairs
    c->parent.ensure(m * sizeof(int));
    std::vector<int> iota(m);
    for (long long j = 0; j < m; ++j) iota[j] = (int)j;
    HIP_CHECK(hipMemcpyAsync(c->parent.p, iota.data(), m * sizeof(int), hipMemcpyHostToDevice, s));
    c->changed.ensure(sizeof(int));
    for (int it = 0;; ++it) {
      int changed = 0;
      HIP_CHECK(hipMemsetAsync(c->changed.p, 0, sizeof(int), s));
      HIP_CHECK(cv::launch_union_round(c->kidx.as<int>(), c->kd2.as<float>(), m, k, r2,
                                       c->parent.as<int>(), c->changed.as<int>(), s));
      HIP_CHECK(hipMemcpyAsync(&changed, c->changed.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      if (!changed) break;
      if (it > 4 * 64) throw std::runtime_error("union-find did not converge");
    }
    // 4. super-points numbered by smallest member; members grouped
    c->flag.ensure(std::max<long long>(n, m));
    HIP_CHECK(cv::launch_is_root(c->parent.as<int>(), m, c->flag.as<uint8_t>(), s));
    const long long nsp = flag_offsets(c, c->flag.as<uint8_t>(), m);
    c->rank.ensure(m * sizeof(int));
    HIP_CHECK(cv::launch_flag_rank(c->flag.as<uint8_t>(), m, c->offs.as<long long>(),
                                   c->rank.as<int>(), nullptr, s));
    for (int b = 0; b < 2; ++b) {
      c->sp[b].ensure(m * sizeof(unsigned int));
      c->member[b].ensure(m * sizeof(unsigned int));
    }
    c->counts.ensure(nsp * sizeof(unsigned int));
    HIP_CHECK(hipMemsetAsync(c->counts.p, 0, nsp * sizeof(unsigned int), s));
    HIP_CHECK(cv::launch_sp_of(c->parent.as<int>(), c->rank.as<int>(), m, c->keep.as<int>(),
                               c->sp[0].as<unsigned int>(), c->member[0].as<unsigned int>(),
                               c->point_sp.as<int>(), c->counts.as<unsigned int>(), s));
    int bits = 1;
    while (bits < 32 && (1ll << bits) <= nsp) ++bits;
    size_t tmp = 0;
    HIP_CHECK(cv::sort_u32_pairs(c->sp[0].as<unsigned int>(), c->sp[1].as<unsigned int>(),
                                 c->member[0].as<unsigned int>(), c->member[1].as<unsigned int>(),
                                 (int)m, bits, nullptr, &tmp, s));
    c->sort_tmp.ensure(tmp);
    HIP_CHECK(cv::sort_u32_pairs(c->sp[0].as<unsigned int>(), c->sp[1].as<unsigned int>(),
                                 c->member[0].as<unsigned int>(), c->member[1].as<unsigned int>(),
                                 (int)m, bits, c->sort_tmp.p, &tmp, s));
    scan_counts(c, c->counts.as<unsigned int>(), nsp, c->moff);
    c->members.ensure(m * sizeof(int));
    c->sp_coord.ensure(nsp * 3 * sizeof(double));
    HIP_CHECK(cv::launch_centroid(c->coords.as<double>(), c->keep.as<int>(),
                                  c->member[1].as<unsigned int>(), c->moff.as<long long>(), nsp,
                                  c->sp_coord.as<double>(), c->members.as<int>(), s));
    // 5. visibility: sorted unique (super-point, shot) keys
    c->obs_cnt.ensure(m * sizeof(long long));
    HIP_CHECK(cv::launch_obs_count(c->obs_off.as<long long>(), c->obs_shot.as<int>(),
                                   c->keep.as<int>(), m, n_shots, c->obs_cnt.as<long long>(), s));
    const long long tobs = scan_counts(c, c->obs_cnt.as<long long>(), m, c->obs_pos);
    for (int b = 0; b < 2; ++b) c->keys[b].ensure(tobs * sizeof(unsigned long long));
    HIP_CHECK(cv::launch_obs_keys(c->obs_off.as<long long>(), c->obs_shot.as<int>(),
                                  c->keep.as<int>(), c->sp[0].as<unsigned int>(),
                                  c->obs_pos.as<long long>(), m, n_shots,
                                  c->keys[0].as<unsigned long long>(), s));
    const unsigned long long maxkey = (unsigned long long)nsp * (unsigned long long)n_shots;
    int kbits = 1;
    while (kbits < 64 && (1ull << kbits) <= maxkey) ++kbits;
    tmp = 0;
    HIP_CHECK(cv::sort_u64_keys(c->keys[0].as<unsigned long long>(),
                                c->keys[1].as<unsigned long long>(), (int)tobs, kbits, nullptr,
                                &tmp, s));
    c->sort_tmp.ensure(tmp);
    HIP_CHECK(cv::sort_u64_keys(c->keys[0].as<unsigned long long>(),
                                c->keys[1].as<unsigned long long>(), (int)tobs, kbits,
                                c->sort_tmp.p, &tmp, s));
    c->flag.ensure(std::max<long long>({n, m, tobs}));
    HIP_CHECK(cv::launch_unique(c->keys[1].as<unsigned long long>(), tobs, c->flag.as<uint8_t>(),
                                s));
    const long long nv = flag_offsets(c, c->flag.as<uint8_t>(), tobs);
    c->index.ensure(std::max<long long>(nv, 1) * sizeof(int));
    HIP_CHECK(cv::launch_flag_rank(c->flag.as<uint8_t>(), tobs, c->offs.as<long long>(), nullptr,
                                   c->index.as<int>(), s));
    c->vis.ensure(std::max<long long>(nv, 1) * sizeof(int));
    c->vcounts.ensure(nsp * sizeof(unsigned int));
    HIP_CHECK(hipMemsetAsync(c->vcounts.p, 0, nsp * sizeof(unsigned int), s));
    HIP_CHECK(cv::launch_vis_write(c->keys[1].as<unsigned long long>(), c->index.as<int>(), nv,
                                   n_shots, c->vis.as<int>(), c->vcounts.as<unsigned int>(), s));
    scan_counts(c, c->vcounts.as<unsigned int>(), nsp, c->voff);
    // 6. shot pairs sharing a super-point
    const long long nn = (long long)n_shots * n_shots;
    long long ne = 0;
    if (nn > 0) {
      c->table.ensure(nn * sizeof(unsigned int));
      HIP_CHECK(hipMemsetAsync(c->table.p, 0, nn * sizeof(unsigned int), s));
      HIP_CHECK(cv::launch_pair_count(c->voff.as<long long>(), c->vis.as<int>(), nsp, n_shots,
                                      c->table.as<unsigned int>(), s));
      c->flag.ensure(std::max<long long>({n, m, tobs, nn}));
      HIP_CHECK(cv::launch_edge_flag(c->table.as<unsigned int>(), n_shots, c->flag.as<uint8_t>(),
                                     s));
      ne = flag_offsets(c, c->flag.as<uint8_t>(), nn);
      c->index.ensure(std::max<long long>({nv, ne, 1}) * sizeof(int));
      HIP_CHECK(cv::launch_flag_rank(c->flag.as<uint8_t>(), nn, c->offs.as<long long>(), nullptr,
                                     c->index.as<int>(), s));
      c->edges.ensure(std::max<long long>(ne, 1) * 2 * sizeof(int));
      c->weights.ensure(std::max<long long>(ne, 1) * sizeof(int));
      HIP_CHECK(cv::launch_edge_write(c->table.as<unsigned int>(), c->index.as<int>(), ne,
                                      n_shots, c->edges.as<int>(), c->weights.as<int>(), s));
    }
    HIP_CHECK(hipEventRecord(c->ev[1], s));
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms, c->ev[0], c->ev[1]));
    c->n_sp = nsp;
    c->n_vis = nv;
    c->n_edges = ne;
    *n_super_points = nsp;
    *n_vis = nv;
    *n_edges = ne;
    return PM_OK;
  });
}

int pm_covis_download(pm_covis* c, int32_t* point_sp, double* sp_coords, int64_t* sp_member_off,
                      int32_t* sp_members, int64_t* sp_vis_off, int32_t* sp_vis, int32_t* edges,
                      int32_t* weights, float* scale) {
  if (!c) return fail(PM_ERR_INVALID, "null handle");
  if (c->n_sp < 0) return fail(PM_ERR_INVALID, "download before build");
  return guarded([&] {
    DeviceGuard dg(c->device);
    hipStream_t s = c->stream;
    auto get = [&](void* dst, const DevBuf& src, size_t bytes) {
      if (dst && bytes) HIP_CHECK(hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, s));
    };
    get(point_sp, c->point_sp, c->n_points * sizeof(int));
    if (c->n_sp > 0) {
      get(sp_coords, c->sp_coord, c->n_sp * 3 * sizeof(double));
      get(sp_member_off, c->moff, (c->n_sp + 1) * sizeof(long long));
      long long nm = 0;
      HIP_CHECK(hipMemcpyAsync(&nm, c->moff.as<long long>() + c->n_sp, sizeof(long long),
                               hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      get(sp_members, c->members, nm * sizeof(int));
      get(sp_vis_off, c->voff, (c->n_sp + 1) * sizeof(long long));
      get(sp_vis, c->vis, c->n_vis * sizeof(int));
      get(edges, c->edges, c->n_edges * 2 * sizeof(int));
      get(weights, c->weights, c->n_edges * sizeof(int));
    } else if (sp_member_off) {
      sp_member_off[0] = 0;
      if (sp_vis_off) sp_vis_off[0] = 0;
    }
    if (scale) *scale = c->scale;
    HIP_CHECK(hipStreamSynchronize(s));
    return PM_OK;
  });
}

int pm_covis_select_neighbors(pm_covis* c, int n_shots, const double* origins, int64_t n_sp,
                              const double* sp_coords, const int64_t* sp_vis_off,
                              const int32_t* sp_vis, const int32_t* sp_ntracks,
                              int num_neighbors, int min_point_best, int min_point_all,
                              double theta_min_deg, double theta_max_deg, int64_t* n_best,
                              int64_t* n_all) {
  if (!c || !n_best || !n_all) return fail(PM_ERR_INVALID, "null argument");
  if (n_shots < 0 || n_shots > cv::kMaxShots)
    return fail(PM_ERR_INVALID, "n_shots must be in [0, " + std::to_string(cv::kMaxShots) + "]");
  if (n_shots > 0 && !origins) return fail(PM_ERR_INVALID, "null origins");
  if (n_sp < 0 || (n_sp > 0 && (!sp_coords || !sp_vis_off || !sp_vis || !sp_ntracks)))
    return fail(PM_ERR_INVALID, "null super-points");
  const int nnb = std::max(0, num_neighbors);
  if (nnb > cv::kMaxNeighbours)
    return fail(PM_ERR_INVALID, "num_neighbors above " + std::to_string(cv::kMaxNeighbours) +
                                    " is not supported");
  return guarded([&] {
    DeviceGuard dg(c->device);
    hipStream_t s = c->stream;
    c->nb_shots = n_shots;
    c->nb_num = nnb;
    c->n_best = c->n_all = 0;
    *n_best = *n_all = 0;
    if (n_shots == 0) return PM_OK;
    HIP_CHECK(hipEventRecord(c->ev[0], s));
    const double pi = 3.14159265358979323846;
    const double cos_lo = cos(theta_max_deg * pi / 180.0);  // dense_helpers.cc:344-347
    const double cos_hi = cos(theta_min_deg * pi / 180.0);
    c->origins.ensure(n_shots * 3 * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(c->origins.p, origins, n_shots * 3 * sizeof(double),
                             hipMemcpyHostToDevice, s));
    const long long nn = (long long)n_shots * n_shots;
    c->pcount.ensure(nn * sizeof(unsigned int));
    c->pscore.ensure(nn * sizeof(unsigned int));
    c->psum.ensure(nn * sizeof(long long));
    HIP_CHECK(hipMemsetAsync(c->pcount.p, 0, nn * sizeof(unsigned int), s));
    HIP_CHECK(hipMemsetAsync(c->pscore.p, 0, nn * sizeof(unsigned int), s));
    HIP_CHECK(hipMemsetAsync(c->psum.p, 0, nn * sizeof(long long), s));
    if (n_sp > 0) {
      const long long nvis = sp_vis_off[n_sp];
      c->nsp_coord.ensure(n_sp * 3 * sizeof(double));
      c->nvoff.ensure((n_sp + 1) * sizeof(long long));
      c->nvis.ensure(std::max<long long>(nvis, 1) * sizeof(int));
      c->ntracks.ensure(n_sp * sizeof(int));
      HIP_CHECK(hipMemcpyAsync(c->nsp_coord.p, sp_coords, n_sp * 3 * sizeof(double),
                               hipMemcpyHostToDevice, s));
      HIP_CHECK(hipMemcpyAsync(c->nvoff.p, sp_vis_off, (n_sp + 1) * sizeof(long long),
                               hipMemcpyHostToDevice, s));
      if (nvis)
        HIP_CHECK(hipMemcpyAsync(c->nvis.p, sp_vis, nvis * sizeof(int), hipMemcpyHostToDevice, s));
      HIP_CHECK(hipMemcpyAsync(c->ntracks.p, sp_ntracks, n_sp * sizeof(int),
                               hipMemcpyHostToDevice, s));
      HIP_CHECK(cv::launch_pair_score(c->origins.as<double>(), n_shots, c->nsp_coord.as<double>(),
                                      c->nvoff.as<long long>(), c->nvis.as<int>(),
                                      c->ntracks.as<int>(), n_sp, cos_lo, cos_hi,
                                      c->pcount.as<unsigned int>(), c->pscore.as<unsigned int>(),
                                      c->psum.as<long long>(), s));
    }
    c->best.ensure((size_t)n_shots * (nnb + 1) * sizeof(int));
    c->best_n.ensure(n_shots * sizeof(int));
    c->all_n.ensure(n_shots * sizeof(long long));
    HIP_CHECK(cv::launch_select(c->pcount.as<unsigned int>(), c->pscore.as<unsigned int>(),
                                c->psum.as<long long>(), n_shots, nnb, min_point_best,
                                min_point_all, cos_lo, cos_hi, c->best.as<int>(),
                                c->best_n.as<int>(), c->all_n.as<long long>(), s));
    const long long total_all = scan_counts(c, c->all_n.as<long long>(), n_shots, c->all_off);
    c->all.ensure(std::max<long long>(total_all, 1) * sizeof(int));
    HIP_CHECK(cv::launch_select_all(c->pcount.as<unsigned int>(), c->pscore.as<unsigned int>(),
                                    c->psum.as<long long>(), n_shots, nnb, min_point_all,
                                    c->best.as<int>(), c->best_n.as<int>(),
                                    c->all_off.as<long long>(), c->all.as<int>(), s));
    std::vector<int> bn(n_shots);
    HIP_CHECK(hipMemcpyAsync(bn.data(), c->best_n.p, n_shots * sizeof(int), hipMemcpyDeviceToHost,
                             s));
    HIP_CHECK(hipEventRecord(c->ev[1], s));
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms, c->ev[0], c->ev[1]));
    long long tb = 0;
    for (int v : bn) tb += v;
    c->n_best = tb;
    c->n_all = total_all;
    *n_best = tb;
    *n_all = total_all;
    return PM_OK;
  });
}

int pm_covis_download_neighbors(pm_covis* c, int64_t* best_off, int32_t* best, int64_t* all_off,
                                int32_t* all) {
  if (!c) return fail(PM_ERR_INVALID, "null handle");
  if (c->nb_shots < 0) return fail(PM_ERR_INVALID, "download before select_neighbors");
  if (!best_off || !best || !all_off || !all) return fail(PM_ERR_INVALID, "null output");
  return guarded([&] {
    DeviceGuard dg(c->device);
    hipStream_t s = c->stream;
    const int n = c->nb_shots;
    best_off[0] = 0;
    all_off[0] = 0;
    if (n == 0) return PM_OK;
    std::vector<int> bn(n), bl((size_t)n * (c->nb_num + 1));
    HIP_CHECK(hipMemcpyAsync(bn.data(), c->best_n.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(bl.data(), c->best.p, bl.size() * sizeof(int), hipMemcpyDeviceToHost,
                             s));
    HIP_CHECK(hipMemcpyAsync(all_off, c->all_off.p, (n + 1) * sizeof(long long),
                             hipMemcpyDeviceToHost, s));
    if (c->n_all)
      HIP_CHECK(hipMemcpyAsync(all, c->all.p, c->n_all * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    long long w = 0;
    for (int i = 0; i < n; ++i) {
      for (int t = 0; t < bn[i]; ++t) best[w + t] = bl[(size_t)i * (c->nb_num + 1) + t];
      w += bn[i];
      best_off[i + 1] = w;
    }
    return PM_OK;
  });
}

int pm_covis_last_ms(pm_covis* c, float* ms) {
  if (!c || !ms) return fail(PM_ERR_INVALID, "null argument");
  *ms = c->last_ms;
  return PM_OK;
}

}  // extern "C"
