// Host side of the C ABI (include/opensfm_amd.h): handle objects, FP64
// camera algebra, device buffers, launches.
//
// Built with -ffp-contract=off like the kernels: the host composes the plane
// sweep's homographies and tabulates the bilateral weights with the
// reference's operation order (depthmap.cc:96-102, 120-125, 184-200,
// 480-487).
#include <hip/hip_runtime.h>
#include <math.h>

#include <cmath>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/opensfm_amd.h"
#include "pm_host.h"
#include "pm_kernels.h"
#include "pm_types.h"

using namespace pmhost;

// ===========================================================================
// DepthmapEstimator

struct HostView {
  double K[9], R[9], t[3], Kinv[9], Q[9], a[3];
  int width, height;
  DevBuf img, quad;
  DevBuf quadc;         // FP16 bilinear coefficients of the quads
  DevBuf quad16;        // 16-bit quads, built on first use
  int quad16_kind = 0;  // precision they were built for (0: none)
};

struct pm_estimator {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // parameters (defaults of DepthmapEstimator::DepthmapEstimator,
  // depthmap.cc:134-144)
  int patch_size = 7;
  double min_depth = 0, max_depth = 0;
  int num_depth_planes = 50;
  int iterations = 3;
  float min_patch_variance = 5 * 5;
  uint64_t seed = 0x5EEDull;
  int n_far = 6;
  int n_refine = 6;
  int n_explore = 3;
  int precision = PM_PRECISION_FP32;
  int post_process = 1;
  int consistency = 1;
  std::vector<std::unique_ptr<HostView>> views;
  DevBuf mask0;
  // per-compute device buffers
  DevBuf viewtab, wlut, state, score, nghbr, counter, queues;
  // host copy of the view table: it feeds an asynchronous upload, so it lives
  // until the compute's final synchronisation
  std::vector<pm::PmView> tab_host;
  DevBuf cons_count, cons_idx, cons_med;  // median-consistent re-selection list
  // and its window-phase work items (pm::ConsistItems)
  DevBuf item_count, item_entry, item_win, item_score, item_nghbr, item_depth, item_first,
      item_todo;
  DevBuf out_depth, out_plane, out_score, out_nghbr;
  DevBuf sweep_H, sweep_planes;
  bool have_result = false;
  pm_stats stats{};
  int wlut_patch = -1;

  std::vector<hipEvent_t> half_ev;  // around the red/black launches
  // learned initialisation map (caller's device memory) and the coarse
  // copies of the views its input is computed on
  const float* init_map = nullptr;
  int init_w = 0, init_h = 0, init_stride = 0, init_scale = 1, init_search = 1;
  float init_sigma = 0.05f, init_sigma_slope = 0.3f;
  int coarse_scale = 0;
  std::vector<std::unique_ptr<HostView>> coarse;
  DevBuf coarse_viewtab;
  // half-resolution level (set_pyramid): its own estimator over the views
  // box-averaged by 2, run first; its state seeds this one's initialisation.
  // With pyramid_levels > 2 that estimator has a half level of its own, and
  // so on: every coarser level runs pyramid_iters iterations.
  int pyramid_iters = 0;
  int pyramid_levels = 2;
  // explicit iterations per coarser level, coarsest first (set_pyramid_
  // schedule; empty: pyramid_iters on each of the pyramid_levels - 1 levels)
  std::vector<int> level_iters;
  // lanes per pixel of the red/black launches (0: chosen by size)
  int lanes_per_pixel = 0;
  std::unique_ptr<pm_estimator> half;
  DevBuf half_map;  // the half level's state as an initialisation map
  bool own_stream = true;  // a pyramid level runs on its parent's stream
  int top_consistency = -1;  // a coarser level: the full level's setting
  int pyramid_consistency = 1;  // the re-selection at the coarser levels too
  // a compute enqueued and not yet waited for (run with wait = false): what
  // its statistics need once the stream has drained
  struct {
    int mode = 0, nl = 0, launches = 0;
    bool pyramid = false;
    float sweep_ms = 0.0f;
  } pending;

  ~pm_estimator() {
    for (auto& e : ev)
      if (e) hipEventDestroy(e);
    for (auto& e : half_ev) hipEventDestroy(e);
    if (stream && own_stream) hipStreamDestroy(stream);
  }
};

namespace {


// PlaneFromDepthAndNormal (depthmap.cc:120-125) in the reference's order.
void plane_from_depth_and_normal(float x, float y, const double* Kinv,
                                 float depth, const float* normal,
                                 float* plane) {
  double S[9];
  for (int k = 0; k < 9; ++k) S[k] = Kinv[k] * depth;
  const double q[3] = {x, y, 1};
  double pd[3];
  mat3_vec(S, q, pd);
  const float point[3] = {(float)pd[0], (float)pd[1], (float)pd[2]};
  float s = 0;
  for (int k = 0; k < 3; ++k) s += normal[k] * point[k];
  const float denom = -s;
  const float m = (1e-6f < denom) ? denom : 1e-6f;
  const float inv = 1.f / m;
  for (int k = 0; k < 3; ++k) plane[k] = normal[k] * inv;
}

// PlaneInducedHomographyBaked (depthmap.cc:96-102).
void homography_baked(const double* K1inv, const double* Q2, const double* a2,
                      const double* K2, const float* vf, float* H) {
  const double v[3] = {vf[0], vf[1], vf[2]};
  double S[9], T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      s += a2[i] * v[j];
      S[i * 3 + j] = Q2[i * 3 + j] + s;
    }
  mat3_mul(K2, S, T);
  mat3_mul(T, K1inv, T);
  for (int k = 0; k < 9; ++k) H[k] = (float)T[k];
}

// Iterations of the coarser levels, coarsest first (empty: no pyramid).
std::vector<int> pyramid_schedule(const pm_estimator* e) {
  if (!e->level_iters.empty()) return e->level_iters;
  if (e->pyramid_iters <= 0) return {};
  return std::vector<int>(e->pyramid_levels - 1, e->pyramid_iters);
}

// with_pyramid: the compute runs the pyramid when one is set
int validate_for_compute(pm_estimator* e, bool with_pyramid = false) {
  if (e->views.size() < 2)
    return fail(PM_ERR_INVALID, "need a reference view and at least one neighbour view (got " +
                                    std::to_string(e->views.size()) + ")");
  if ((int)e->views.size() > pm::kMaxViews)
    return fail(PM_ERR_INVALID, "at most " + std::to_string(pm::kMaxViews) + " views");
  if (!(e->min_depth > 0) || !(e->max_depth > e->min_depth) ||
      !std::isfinite(e->max_depth))
    return fail(PM_ERR_INVALID, "depth range must satisfy 0 < min_depth < max_depth (call set_depth_range)");
  const HostView& r = *e->views[0];
  const int hpz = (e->patch_size - 1) / 2;
  if (r.width < 2 * hpz + 1 || r.height < 2 * hpz + 1)
    return fail(PM_ERR_INVALID, "reference view smaller than one patch");
  const int sh = (int)pyramid_schedule(e).size();
  if (with_pyramid && sh > 0 && e->init_map && e->init_scale % (1 << sh) != 0)
    return fail(PM_ERR_INVALID, "an initialisation map seeding a pyramid of " +
                                    std::to_string(sh + 1) + " levels needs a scale divisible by " +
                                    std::to_string(1 << sh));
  if (with_pyramid && sh > 0) {
    // the coarsest level: every view box-averaged by 2^(levels - 1), the
    // reference at least one patch
    bool ok = (r.width >> sh) >= 2 * hpz + 1 && (r.height >> sh) >= 2 * hpz + 1;
    for (const auto& V : e->views) ok = ok && (V->width >> sh) >= 2 && (V->height >> sh) >= 2;
    if (!ok)
      return fail(PM_ERR_INVALID, "views too small for a pyramid of " +
                                      std::to_string(sh + 1) +
                                      " levels (the coarsest reference must hold one patch)");
  }
  return PM_OK;
}

// Fills the device view table and the argument block for a compute.
pm::PmArgs prepare(pm_estimator* e) {
  const int n = (int)e->views.size();
  const HostView& r = *e->views[0];
  std::vector<pm::PmView>& tab = e->tab_host;
  tab.assign(n, pm::PmView{});
  for (int v = 0; v < n; ++v) {
    HostView& V = *e->views[v];
    if (e->precision != PM_PRECISION_FP32 && V.quad16_kind != e->precision) {
      V.quad16.ensure((size_t)(V.width + 1) * (V.height + 1) * 8);
      HIP_CHECK(pm::launch_build_quad16(V.img.as<uint8_t>(), V.quad16.as<uint32_t>(),
                                        V.width, V.height, e->precision, e->stream));
      V.quad16_kind = e->precision;
    }
    fill_view(tab[v], V.quad.as<uint32_t>(), V.width, V.height, V.K, V.Q, V.a,
              r.Kinv);
    tab[v].quad16 = V.quad16.as<uint32_t>();
    tab[v].quadc = V.quadc.as<uint32_t>();
  }
  e->viewtab.ensure(sizeof(pm::PmView) * n);
  HIP_CHECK(hipMemcpyAsync(e->viewtab.p, tab.data(), sizeof(pm::PmView) * n,
                           hipMemcpyHostToDevice, e->stream));
  const int hpz = (e->patch_size - 1) / 2;
  if (e->wlut_patch != e->patch_size) {
    std::vector<float> t = bilateral_table(hpz);
    e->wlut.ensure(t.size() * sizeof(float));
    HIP_CHECK(hipMemcpyAsync(e->wlut.p, t.data(), t.size() * sizeof(float),
                             hipMemcpyHostToDevice, e->stream));
    HIP_CHECK(hipStreamSynchronize(e->stream));
    e->wlut_patch = e->patch_size;
  }
  const size_t np = (size_t)r.width * r.height;
  e->state.ensure(np * 4 * sizeof(float));
  e->score.ensure(np * sizeof(float));
  e->nghbr.ensure(np * sizeof(int));
  e->counter.ensure(2 * sizeof(unsigned long long));
  HIP_CHECK(hipMemsetAsync(e->state.p, 0, np * 4 * sizeof(float), e->stream));
  HIP_CHECK(hipMemsetAsync(e->score.p, 0, np * sizeof(float), e->stream));
  HIP_CHECK(hipMemsetAsync(e->nghbr.p, 0, np * sizeof(int), e->stream));
  HIP_CHECK(hipMemsetAsync(e->counter.p, 0, 2 * sizeof(unsigned long long), e->stream));

  pm::PmArgs a;
  memset(&a, 0, sizeof(a));
  a.W = r.width;
  a.H = r.height;
  a.nviews = n;
  a.hpz = hpz;
  a.views = e->viewtab.as<pm::PmView>();
  a.ref_img = r.img.as<uint8_t>();
  a.ref_mask = e->mask0.as<uint8_t>();
  a.wlut = e->wlut.as<float>();
  a.state = e->state.as<float>();
  a.score = e->score.as<float>();
  a.nghbr = e->nghbr.as<int>();
  for (int k = 0; k < 9; ++k) {
    a.Kinv0f[k] = (float)r.Kinv[k];
    a.Kinv0[k] = r.Kinv[k];
  }
  a.log_min = (float)log(e->min_depth);
  a.log_max = (float)log(e->max_depth);
  a.min_patch_var = e->min_patch_variance;
  a.seed_lo = (uint32_t)(e->seed & 0xffffffffu);
  a.seed_hi = (uint32_t)(e->seed >> 32);
  a.n_far = e->n_far;
  a.n_refine = e->n_refine;
  a.n_explore = e->n_explore;
  a.precision = e->precision;
  a.init_map = e->init_map;
  a.init_w = e->init_w;
  a.init_h = e->init_h;
  a.init_stride = e->init_stride;
  a.init_scale = e->init_scale;
  a.init_search = e->init_search;
  a.init_sigma = e->init_sigma;
  a.init_sigma_slope = e->init_sigma_slope;
  a.counter = e->counter.as<unsigned long long>();
  return a;
}

// One median-consistent re-selection pass over the estimator's maps.
hipError_t consistency_pass(pm_estimator* e, const pm::PmArgs& a0, bool sample, unsigned salt,
                            bool final_pass) {
  pm::PmArgs a = a0;
  const size_t np = (size_t)a.W * a.H;
  e->cons_count.ensure(sizeof(unsigned));
  e->cons_idx.ensure(np * sizeof(unsigned));
  e->cons_med.ensure(np * sizeof(float));
  // the output maps hold the listed pixels' new states until the pack
  e->out_depth.ensure(np * sizeof(float));
  e->out_plane.ensure(np * 3 * sizeof(float));
  e->out_score.ensure(np * sizeof(float));
  e->out_nghbr.ensure(np * sizeof(int));
  a.counter = e->counter.as<unsigned long long>();  // slot 0: not the dominant kernel
  // the last pass keeps only re-selections that agree with enough of the
  // window; earlier ones keep all, so that the following iterations can
  // improve them
  const pm::ConsistOut res{e->out_plane.as<float>(), e->out_depth.as<float>(),
                           e->out_score.as<float>(), e->out_nghbr.as<int>(),
                           pm::kConsistMinScore, final_pass ? pm::kConsistFinalSupport : 0};
  // window items: room for two per pixel (more are scored inline)
  const size_t cap = std::min<size_t>(2 * np, (size_t)1 << 26);
  e->item_count.ensure(sizeof(unsigned long long));
  e->item_entry.ensure(cap * sizeof(unsigned));
  e->item_win.ensure(cap);
  e->item_score.ensure(cap * sizeof(float));
  e->item_nghbr.ensure(cap * sizeof(int));
  e->item_depth.ensure(cap * sizeof(float));
  e->item_first.ensure(np * sizeof(unsigned));
  e->item_todo.ensure(np * sizeof(unsigned));
  const pm::ConsistItems items{e->item_count.as<unsigned long long>(), (unsigned)cap,
                               e->item_entry.as<unsigned>(), e->item_win.as<unsigned char>(),
                               e->item_score.as<float>(), e->item_nghbr.as<int>(),
                               e->item_depth.as<float>(), e->item_first.as<unsigned>(),
                               e->item_todo.as<unsigned>()};
  return pm::launch_pm_consist(a, sample, e->cons_count.as<unsigned>(),
                               e->cons_idx.as<unsigned>(), e->cons_med.as<float>(), res, salt,
                               items, e->stream);
}

// The half-resolution level's estimator: every view box-averaged by 2
// (K scaled about pixel centres, x_half = (x + 0.5) / 2 - 0.5; poses
// unchanged), the reference's mask the minimum of each 2 x 2 block.  Built on
// the device from the uploaded views, and rebuilt when views are added.
void ensure_half(pm_estimator* e) {
  if (e->half && e->half->views.size() == e->views.size()) return;
  std::unique_ptr<pm_estimator> h(new pm_estimator());
  h->device = e->device;
  h->stream = e->stream;  // in stream order with this level: no host wait between levels
  h->own_stream = false;
  for (auto& ev : h->ev) HIP_CHECK(hipEventCreate(&ev));
  for (size_t k = 0; k < e->views.size(); ++k) {
    const HostView& V = *e->views[k];
    std::unique_ptr<HostView> c(new HostView());
    memcpy(c->R, V.R, sizeof(c->R));
    memcpy(c->t, V.t, sizeof(c->t));
    memcpy(c->Q, V.Q, sizeof(c->Q));
    memcpy(c->a, V.a, sizeof(c->a));
    memcpy(c->K, V.K, sizeof(c->K));
    for (int r = 0; r < 2; ++r) {
      for (int col = 0; col < 3; ++col) c->K[r * 3 + col] *= 0.5;
      c->K[r * 3 + 2] -= 0.25 * V.K[8];
    }
    if (!inv3(c->K, c->Kinv)) throw std::runtime_error("half-level K is singular");
    c->width = V.width / 2;
    c->height = V.height / 2;
    if (c->width < 2 || c->height < 2) throw std::runtime_error("view too small for a half level");
    const size_t np = (size_t)c->width * c->height;
    c->img.ensure(np);
    HIP_CHECK(pm::launch_downsample_u8(V.img.as<uint8_t>(), V.width, V.height, 2, 0,
                                       c->img.as<uint8_t>(), c->width, c->height, h->stream));
    c->quad.ensure((size_t)(c->width + 1) * (c->height + 1) * sizeof(uint32_t));
    c->quadc.ensure((size_t)(c->width + 1) * (c->height + 1) * 8);
    HIP_CHECK(pm::launch_build_quad(c->img.as<uint8_t>(), c->quad.as<uint32_t>(),
                                    c->quadc.as<uint32_t>(), c->width, c->height, h->stream));
    if (k == 0) {
      h->mask0.ensure(np);
      HIP_CHECK(pm::launch_downsample_u8(e->mask0.as<uint8_t>(), V.width, V.height, 2, 1,
                                         h->mask0.as<uint8_t>(), c->width, c->height,
                                         h->stream));
    }
    h->views.push_back(std::move(c));
  }
  HIP_CHECK(hipStreamSynchronize(h->stream));
  e->half = std::move(h);
}

int run(pm_estimator* e, int mode, bool wait = true);

// Runs the half level (same settings, the last entry of the schedule as its
// iterations and the rest as its own pyramid, its own
// seed) and turns its state into e->half_map: per half-level pixel the
// normalised log depth, the normal's slopes and a confidence logit (+10 where
// the depth is live and PostProcess's median test keeps it, else -10).
int run_half(pm_estimator* e, int mode) {
  ensure_half(e);
  pm_estimator* h = e->half.get();
  h->patch_size = e->patch_size;
  h->min_depth = e->min_depth;
  h->max_depth = e->max_depth;
  h->num_depth_planes = e->num_depth_planes;
  std::vector<int> sched = pyramid_schedule(e);
  h->iterations = sched.back();
  sched.pop_back();
  h->level_iters = sched;  // empty: the coarsest level
  h->pyramid_iters = 0;
  h->pyramid_levels = (int)sched.size() + 1;
  h->lanes_per_pixel = e->lanes_per_pixel;
  h->min_patch_variance = e->min_patch_variance;
  h->seed = e->seed + pm::kHalfLevelSeed;
  h->n_far = e->n_far;
  h->n_refine = e->n_refine;
  h->n_explore = e->n_explore;
  h->precision = e->precision;
  h->post_process = e->post_process;
  // the re-selection at a coarser level: the full level's setting unless
  // set_pyramid_consistency(0)
  const int top = e->top_consistency >= 0 ? e->top_consistency : e->consistency;
  h->top_consistency = top;
  h->pyramid_consistency = e->pyramid_consistency;
  h->consistency = top && e->pyramid_consistency;
  // a caller's initialisation map seeds the coarsest level (its cells cover
  // half as many pixels of each coarser level)
  h->init_map = e->init_map;
  h->init_w = e->init_map ? e->init_w : 0;
  h->init_h = e->init_map ? e->init_h : 0;
  h->init_stride = e->init_map ? e->init_stride : 0;
  h->init_scale = e->init_map ? e->init_scale / 2 : 1;
  h->init_search = e->init_search;
  h->init_sigma = e->init_sigma;
  h->init_sigma_slope = e->init_sigma_slope;
  int rc = run(h, mode, false);  // enqueued on this level's stream
  if (rc) return rc;
  const HostView& r = *h->views[0];
  e->half_map.ensure((size_t)r.width * r.height * 4 * sizeof(float));
  HIP_CHECK(pm::launch_state_to_map(h->state.as<float>(), r.width, r.height,
                                    (float)std::log(e->min_depth), (float)std::log(e->max_depth),
                                    e->post_process, e->half_map.as<float>(), e->stream));
  return PM_OK;
}

// Statistics of a compute whose stream has drained (its levels' too).
void collect_stats(pm_estimator* e) {
  const auto& pd = e->pending;
  if (pd.pyramid) collect_stats(e->half.get());
  const HostView& r = *e->views[0];
  float total_ms = 0, dom_ms = pd.sweep_ms;
  HIP_CHECK(hipEventElapsedTime(&total_ms, e->ev[0], e->ev[1]));
  for (int l = 0; l < pd.nl; ++l) {
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e->half_ev[2 * l], e->half_ev[2 * l + 1]));
    dom_ms += ms;
  }
  unsigned long long cnt[2] = {0, 0};
  HIP_CHECK(hipMemcpy(cnt, e->counter.p, sizeof(cnt), hipMemcpyDeviceToHost));
  e->stats.scorings = (int64_t)(cnt[0] + cnt[1]);
  e->stats.dominant_scorings =
      (int64_t)(pd.mode == PM_MODE_PLANE_SWEEP ? cnt[0] : cnt[1]);
  e->stats.patch_size = e->patch_size;
  e->stats.width = r.width;
  e->stats.height = r.height;
  e->stats.launches = pd.launches;
  e->stats.dominant_ms = dom_ms;
  e->stats.total_ms = total_ms;
  if (pd.pyramid) {
    // the whole compute: the coarser levels' launches of the same kernels
    // and their device time count too
    const pm_stats& hs = e->half->stats;
    e->stats.scorings += hs.scorings;
    e->stats.dominant_scorings += hs.dominant_scorings;
    e->stats.launches += hs.launches;
    e->stats.dominant_ms += hs.dominant_ms;
    e->stats.total_ms += hs.total_ms;
  }
}

// One compute, and with a pyramid its coarser levels first, all on e's
// stream; wait = false leaves it enqueued (collect_stats after a wait).
int run(pm_estimator* e, int mode, bool wait) {
  int rc = validate_for_compute(e, mode != PM_MODE_PLANE_SWEEP);
  if (rc) return rc;
  if (mode != PM_MODE_PLANE_SWEEP && mode != PM_MODE_PATCHMATCH &&
      mode != PM_MODE_PATCHMATCH_SAMPLE)
    return fail(PM_ERR_INVALID, "unknown compute mode " + std::to_string(mode));
  if (mode != PM_MODE_PLANE_SWEEP && !pm::precision_supported(e->precision, e->patch_size))
    return fail(PM_ERR_INVALID, "FP16/BF16 samples need patch size 5 or 7");
  DeviceGuard dg(e->device);
  e->have_result = false;
  // the half level seeds plain PatchMatch when no initialisation map is set
  const bool pyramid = mode != PM_MODE_PLANE_SWEEP && !pyramid_schedule(e).empty();
  if (pyramid) {
    rc = run_half(e, mode);
    if (rc) return rc;
  }
  pm::PmArgs a = prepare(e);
  if (pyramid) {
    const HostView& hr = *e->half->views[0];
    a.init_map = e->half_map.as<float>();
    a.init_w = hr.width;
    a.init_h = hr.height;
    a.init_stride = 4;
    a.init_scale = 2;
    a.init_search = 1;
    a.init_sigma = 0.0f;
    a.init_sigma_slope = 0.0f;
  }
  const HostView& r = *e->views[0];
  const int n = (int)e->views.size();
  float dom_ms = 0.0f;
  int launches = 0, nl = 0;  // red/black launches timed by half_ev
  HIP_CHECK(hipEventRecord(e->ev[0], e->stream));
  if (mode == PM_MODE_PLANE_SWEEP) {
    const int D = e->num_depth_planes;
    if (D < 2) return fail(PM_ERR_INVALID, "plane sweep needs num_depth_planes >= 2");
    // The sweep plane v_d = n / (d K^-1 q)_z is pixel independent when the
    // last row of K^-1 is (0, 0, k); every pinhole K satisfies this.
    if (r.Kinv[6] != 0.0 || r.Kinv[7] != 0.0)
      return fail(PM_ERR_INVALID, "plane sweep needs a pinhole K (last row 0, 0, k)");
    std::vector<float> planes(3 * (size_t)D), H(9 * (size_t)D * n, 0.0f);
    for (int d = 0; d < D; ++d) {
      // ComputeBruteForce (depthmap.cc:190-195)
      const float depth = (float)(1 / (1 / e->min_depth +
                                       d * (1 / e->max_depth - 1 / e->min_depth) /
                                           (D - 1)));
      const float normal[3] = {0, 0, -1};
      plane_from_depth_and_normal(0.0f, 0.0f, r.Kinv, depth, normal, &planes[3 * d]);
      for (int v = 1; v < n; ++v) {
        const HostView& V = *e->views[v];
        homography_baked(r.Kinv, V.Q, V.a, V.K, &planes[3 * d],
                         &H[((size_t)d * n + v) * 9]);
      }
    }
    e->sweep_H.ensure(H.size() * sizeof(float));
    e->sweep_planes.ensure(planes.size() * sizeof(float));
    HIP_CHECK(hipMemcpyAsync(e->sweep_H.p, H.data(), H.size() * sizeof(float),
                             hipMemcpyHostToDevice, e->stream));
    HIP_CHECK(hipMemcpyAsync(e->sweep_planes.p, planes.data(),
                             planes.size() * sizeof(float),
                             hipMemcpyHostToDevice, e->stream));
    HIP_CHECK(hipEventRecord(e->ev[2], e->stream));
    HIP_CHECK(pm::launch_sweep(a, e->sweep_H.as<float>(),
                               e->sweep_planes.as<float>(), D, e->stream));
    HIP_CHECK(hipEventRecord(e->ev[3], e->stream));
    HIP_CHECK(hipEventSynchronize(e->ev[3]));
    HIP_CHECK(hipEventElapsedTime(&dom_ms, e->ev[2], e->ev[3]));
    launches = 1;
  } else {
    const bool sample = mode == PM_MODE_PATCHMATCH_SAMPLE;
    // unseeded FP32 runs score random planes in the initialisation and the
    // first iteration: those launches read the u8 quads
    const bool quads_first = a.precision == PM_PRECISION_FP32 && !a.init_map;
    pm::PmArgs ai = a;
    if (quads_first) ai.precision = pm::kPrecFp32Quads;
    HIP_CHECK(pm::launch_pm_init(ai, sample, e->stream));
    pm::PmArgs ah = a;  // the red/black launches count into slot 1
    ah.counter = a.counter + 1;
    // an event before and after each red/black launch (the consistency
    // passes run between them); no host wait inside the loop
    nl = 2 * e->iterations;
    while ((int)e->half_ev.size() < 2 * nl) {
      hipEvent_t ev;
      HIP_CHECK(hipEventCreate(&ev));
      e->half_ev.push_back(ev);
    }
    // 8 band queues per red/black launch, cleared once per compute
    e->queues.ensure((size_t)8 * std::max(nl, 1) * sizeof(unsigned));
    HIP_CHECK(pm::launch_fill_u32(e->queues.as<unsigned>(), 8 * nl, 0u, e->stream));
    for (int it = 0; it < e->iterations; ++it) {
      ah.precision = (quads_first && it == 0) ? pm::kPrecFp32Quads : a.precision;
      for (int colour = 0; colour < 2; ++colour) {
        HIP_CHECK(hipEventRecord(e->half_ev[2 * launches], e->stream));
        HIP_CHECK(pm::launch_pm_half(ah, sample, colour, it,
                                     e->queues.as<unsigned>() + 8 * launches, e->stream,
                                     e->lanes_per_pixel));
        HIP_CHECK(hipEventRecord(e->half_ev[2 * launches + 1], e->stream));
        ++launches;
      }
      // median-consistent re-selection after every iteration but the
      // first (whose planes are still mostly random), and after the last
      if (e->consistency && (it >= 1 || it + 1 == e->iterations))
        HIP_CHECK(consistency_pass(e, ah, sample, (unsigned)it, it + 1 == e->iterations));
    }
  }
  const size_t np = (size_t)r.width * r.height;
  e->out_depth.ensure(np * sizeof(float));
  e->out_plane.ensure(np * 3 * sizeof(float));
  e->out_score.ensure(np * sizeof(float));
  e->out_nghbr.ensure(np * sizeof(int));
  HIP_CHECK(pm::launch_pack(a, mode == PM_MODE_PLANE_SWEEP ? 0 : e->post_process,
                            e->out_depth.as<float>(), e->out_plane.as<float>(),
                            e->out_score.as<float>(), e->out_nghbr.as<int>(),
                            e->stream));
  HIP_CHECK(hipEventRecord(e->ev[1], e->stream));
  e->pending.mode = mode;
  e->pending.nl = nl;
  e->pending.launches = launches;
  e->pending.pyramid = pyramid;
  e->pending.sweep_ms = dom_ms;
  if (!wait) return PM_OK;
  HIP_CHECK(hipEventSynchronize(e->ev[1]));  // the one host wait of a compute
  collect_stats(e);
  e->have_result = true;
  return PM_OK;
}

int download(pm_estimator* e, float* depth, float* plane, float* score,
             int32_t* nghbr) {
  if (!e->have_result) return fail(PM_ERR_INVALID, "no computed maps to download");
  if (!depth || !plane || !score || !nghbr)
    return fail(PM_ERR_INVALID, "null output buffer");
  DeviceGuard dg(e->device);
  const HostView& r = *e->views[0];
  const size_t np = (size_t)r.width * r.height;
  HIP_CHECK(hipMemcpyAsync(depth, e->out_depth.p, np * sizeof(float),
                           hipMemcpyDeviceToHost, e->stream));
  HIP_CHECK(hipMemcpyAsync(plane, e->out_plane.p, np * 3 * sizeof(float),
                           hipMemcpyDeviceToHost, e->stream));
  HIP_CHECK(hipMemcpyAsync(score, e->out_score.p, np * sizeof(float),
                           hipMemcpyDeviceToHost, e->stream));
  HIP_CHECK(hipMemcpyAsync(nghbr, e->out_nghbr.p, np * sizeof(int32_t),
                           hipMemcpyDeviceToHost, e->stream));
  HIP_CHECK(hipStreamSynchronize(e->stream));
  return PM_OK;
}

// Coarse copies of the views at 1/scale (box mean of the image; K scaled
// about pixel centres, poses unchanged), with FP16 quads for the
// initialiser's cost volume.  Built once per scale and
// extended when views are added.
void build_coarse(pm_estimator* e, int scale) {
  if (e->coarse_scale != scale) {
    e->coarse.clear();
    e->coarse_scale = scale;
  }
  for (size_t v = e->coarse.size(); v < e->views.size(); ++v) {
    const HostView& V = *e->views[v];
    std::unique_ptr<HostView> c(new HostView());
    memcpy(c->K, V.K, sizeof(c->K));
    memcpy(c->R, V.R, sizeof(c->R));
    memcpy(c->t, V.t, sizeof(c->t));
    memcpy(c->Q, V.Q, sizeof(c->Q));
    memcpy(c->a, V.a, sizeof(c->a));
    // x_coarse = (x + 0.5) / scale - 0.5
    const double inv = 1.0 / scale;
    for (int k = 0; k < 3; ++k) {
      c->K[k] = V.K[k] * inv;
      c->K[3 + k] = V.K[3 + k] * inv;
    }
    c->K[2] += (inv - 1.0) * 0.5 * V.K[8];
    c->K[5] += (inv - 1.0) * 0.5 * V.K[8];
    if (!inv3(c->K, c->Kinv)) throw std::runtime_error("coarse K is singular");
    c->width = V.width / scale;
    c->height = V.height / scale;
    const size_t np = (size_t)c->width * c->height;
    c->img.ensure(np);
    HIP_CHECK(pm::launch_downsample_u8(V.img.as<uint8_t>(), V.width, V.height, scale, 0,
                                       c->img.as<uint8_t>(), c->width, c->height, e->stream));
    c->quad16.ensure((size_t)(c->width + 1) * (c->height + 1) * 8);
    HIP_CHECK(pm::launch_build_quad16(c->img.as<uint8_t>(), c->quad16.as<uint32_t>(), c->width,
                                      c->height, PM_PRECISION_FP16, e->stream));
    c->quad16_kind = PM_PRECISION_FP16;
    e->coarse.push_back(std::move(c));
  }
  HIP_CHECK(hipStreamSynchronize(e->stream));
}

// ComputePlaneImageScore of fixed hypotheses: the exact FP32 scorer with
// host-composed homographies, or (fast) the PatchMatch scorer in the
// estimator's sample precision.
int score_hypotheses(pm_estimator* e, int n, const int32_t* ij,
                     const float* planes, const int32_t* views, float* out,
                     bool fast) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (n < 0) return fail(PM_ERR_INVALID, "negative hypothesis count");
  if (n == 0) return PM_OK;
  if (!ij || !planes || !views || !out) return fail(PM_ERR_INVALID, "null buffer");
  if (e->views.size() < 2)
    return fail(PM_ERR_INVALID, "need a reference view and at least one neighbour view");
  const HostView& r = *e->views[0];
  const int hpz = (e->patch_size - 1) / 2;
  for (int k = 0; k < n; ++k) {
    const int i = ij[2 * k], j = ij[2 * k + 1];
    if (i < hpz || i >= r.height - hpz || j < hpz || j >= r.width - hpz)
      return fail(PM_ERR_RANGE, "hypothesis " + std::to_string(k) +
                                    ": pixel closer than (patch-1)/2 to the border");
    if (views[k] < 1 || views[k] >= (int)e->views.size())
      return fail(PM_ERR_RANGE, "hypothesis " + std::to_string(k) + ": view index out of range");
  }
  if (fast && !pm::precision_supported(e->precision, e->patch_size))
    return fail(PM_ERR_INVALID, "FP16/BF16 samples need patch size 5 or 7");
  return guarded([&] {
    DeviceGuard dg(e->device);
    pm::PmArgs a = prepare(e);
    // group by view: one launch per view keeps the view index uniform
    const int nv = (int)e->views.size();
    std::vector<int> order(n), start(nv + 1, 0);
    for (int k = 0; k < n; ++k) start[views[k] + 1]++;
    for (int v = 0; v < nv; ++v) start[v + 1] += start[v];
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int k = 0; k < n; ++k) order[fill[views[k]]++] = k;
    std::vector<int32_t> s_ij(2 * (size_t)n);
    // FP64 composition on the host, FP32 homography per hypothesis
    // (PlaneInducedHomographyBaked, depthmap.cc:96-102)
    // exact: the FP32 homography per hypothesis; fast: the plane itself
    const int per = fast ? 3 : 9;
    std::vector<float> s_H(per * (size_t)n), s_out(n);
    for (int m = 0; m < n; ++m) {
      const int k = order[m];
      s_ij[2 * m] = ij[2 * k];
      s_ij[2 * m + 1] = ij[2 * k + 1];
      const HostView& V = *e->views[views[k]];
      if (fast)
        memcpy(&s_H[3 * (size_t)m], planes + 3 * (size_t)k, 3 * sizeof(float));
      else
        homography_baked(r.Kinv, V.Q, V.a, V.K, planes + 3 * (size_t)k, &s_H[9 * (size_t)m]);
    }
    DevBuf d_ij, d_H, d_out;
    d_ij.ensure(sizeof(int32_t) * 2 * n);
    d_H.ensure(sizeof(float) * per * n);
    d_out.ensure(sizeof(float) * n);
    HIP_CHECK(hipMemcpyAsync(d_ij.p, s_ij.data(), sizeof(int32_t) * 2 * n, hipMemcpyHostToDevice, e->stream));
    HIP_CHECK(hipMemcpyAsync(d_H.p, s_H.data(), sizeof(float) * per * n, hipMemcpyHostToDevice, e->stream));
    for (int v = 1; v < nv; ++v) {
      const int cnt = start[v + 1] - start[v];
      if (cnt == 0) continue;
      const int* dij = d_ij.as<int>() + 2 * start[v];
      const float* dh = d_H.as<float>() + per * (size_t)start[v];
      float* dout = d_out.as<float>() + start[v];
      if (fast)
        HIP_CHECK(pm::launch_score_hyp_fast(a, v, cnt, dij, dh, dout, e->stream));
      else
        HIP_CHECK(pm::launch_score_hyp(a, v, cnt, dij, dh, dout, e->stream));
    }
    HIP_CHECK(hipMemcpyAsync(s_out.data(), d_out.p, sizeof(float) * n, hipMemcpyDeviceToHost, e->stream));
    HIP_CHECK(hipStreamSynchronize(e->stream));
    for (int m = 0; m < n; ++m) out[order[m]] = s_out[m];
    return PM_OK;
  });
}


}  // namespace

// ===========================================================================
// GPUDepthmapCleaner (opensfm/src/lib/dense/cleaner.h:16-63)

struct GcView {
  double K[9], R[9], t[3];
  int width, height;
  DevBuf depth;
};

struct pm_gpu_cleaner {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // defaults of cleaner.h:47-50
  float same_depth_threshold = 0.01f;
  int min_consistent_views = 3;
  float carving_threshold = 0.2f;
  int max_carved_views = 1;
  std::vector<std::unique_ptr<GcView>> views;
  DevBuf resized, ptrs, cams, out;
  float last_ms = 0;
  ~pm_gpu_cleaner() {
    for (auto& e : ev)
      if (e) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
  }
};

extern "C" {

const char* pm_last_error(void) { return g_last_error.c_str(); }

int pm_device_count(int* count) {
  if (!count) return fail(PM_ERR_INVALID, "null count");
  int n = 0;
  hipError_t err = hipGetDeviceCount(&n);
  if (err != hipSuccess) n = 0;
  *count = n;
  g_last_error.clear();
  return PM_OK;
}

namespace {
// Reserved bytes per device (the dense scheduler's bookkeeping).
std::mutex g_reserve_mu;
std::vector<uint64_t> g_reserved;

int device_props(int device, hipDeviceProp_t* p) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  if (device < 0 || device >= n)
    return fail(PM_ERR_RANGE, "device " + std::to_string(device) + " out of range (" +
                                  std::to_string(n) + " HIP devices)");
  if (hipGetDeviceProperties(p, device) != hipSuccess)
    return fail(PM_ERR_RUNTIME, "hipGetDeviceProperties failed");
  return PM_OK;
}
}  // namespace

int pm_device_info(int device, char* name, int name_capacity, uint64_t* total_bytes,
                   uint64_t* free_bytes) {
  hipDeviceProp_t p;
  int rc = device_props(device, &p);
  if (rc) return rc;
  if (name && name_capacity > 0) {
    const size_t n = std::min(strlen(p.name), (size_t)name_capacity - 1);
    memcpy(name, p.name, n);
    name[n] = 0;
  }
  if (total_bytes) *total_bytes = (uint64_t)p.totalGlobalMem;
  if (free_bytes) {
    size_t fr = 0, tot = 0;
    DeviceGuard dg(device);
    if (hipMemGetInfo(&fr, &tot) != hipSuccess)
      return fail(PM_ERR_RUNTIME, "hipMemGetInfo failed");
    *free_bytes = (uint64_t)fr;
  }
  g_last_error.clear();
  return PM_OK;
}

int pm_device_reserve(int device, int64_t bytes) {
  hipDeviceProp_t p;
  int rc = device_props(device, &p);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_reserve_mu);
  if ((int)g_reserved.size() <= device) g_reserved.resize(device + 1, 0);
  uint64_t& r = g_reserved[device];
  if (bytes >= 0) {
    r += (uint64_t)bytes;
  } else {
    const uint64_t rel = (uint64_t)(-(bytes + 1)) + 1;
    r = rel > r ? 0 : r - rel;
  }
  g_last_error.clear();
  return PM_OK;
}

int pm_device_available(int device, uint64_t* available) {
  if (!available) return fail(PM_ERR_INVALID, "null output");
  hipDeviceProp_t p;
  int rc = device_props(device, &p);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_reserve_mu);
  const uint64_t total = (uint64_t)p.totalGlobalMem;
  const uint64_t r = (int)g_reserved.size() > device ? g_reserved[device] : 0;
  *available = r >= total ? 0 : total - r;
  g_last_error.clear();
  return PM_OK;
}

int pm_estimator_create(int device, pm_estimator** out) {
  if (!out) return fail(PM_ERR_INVALID, "null handle pointer");
  *out = nullptr;
  int rc = check_device(device);
  if (rc) return rc;
  return guarded([&] {
    std::unique_ptr<pm_estimator> e(new pm_estimator());
    e->device = device;
    DeviceGuard dg(device);
    HIP_CHECK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    for (auto& ev : e->ev) HIP_CHECK(hipEventCreate(&ev));
    *out = e.release();
    return PM_OK;
  });
}

int pm_estimator_destroy(pm_estimator* est) {
  if (!est) return PM_OK;
  return guarded([&] {
    DeviceGuard dg(est->device);
    hipStreamSynchronize(est->stream);
    delete est;
    return PM_OK;
  });
}

int pm_estimator_add_view(pm_estimator* e, const double* K, const double* R,
                          const double* t, const uint8_t* image,
                          const uint8_t* mask, int width, int height,
                          int stride) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (!K || !R || !t || !image || !mask)
    return fail(PM_ERR_INVALID, "null K, R, t, image or mask");
  if (width < 2 || height < 2)
    return fail(PM_ERR_INVALID, "image must be at least 2x2");
  if (stride < width) return fail(PM_ERR_INVALID, "stride smaller than width");
  if (!all_finite(K, 9) || !all_finite(R, 9) || !all_finite(t, 3))
    return fail(PM_ERR_INVALID, "K, R and t must be finite");
  if ((int)e->views.size() >= pm::kMaxViews)
    return fail(PM_ERR_INVALID, "at most " + std::to_string(pm::kMaxViews) + " views");
  return guarded([&] {
    std::unique_ptr<HostView> v(new HostView());
    memcpy(v->K, K, sizeof(v->K));
    memcpy(v->R, R, sizeof(v->R));
    memcpy(v->t, t, sizeof(v->t));
    if (!inv3(v->K, v->Kinv)) return fail(PM_ERR_INVALID, "K is singular");
    const HostView* ref = e->views.empty() ? v.get() : e->views[0].get();
    relative_pose(ref->R, ref->t, v->R, v->t, v->Q, v->a);
    v->width = width;
    v->height = height;
    DeviceGuard dg(e->device);
    const size_t np = (size_t)width * height;
    v->img.ensure(np);
    upload_image(v->img.p, image, width, height, stride, e->stream);
    v->quad.ensure((size_t)(width + 1) * (height + 1) * sizeof(uint32_t));
    v->quadc.ensure((size_t)(width + 1) * (height + 1) * 8);
    HIP_CHECK(pm::launch_build_quad(v->img.as<uint8_t>(), v->quad.as<uint32_t>(),
                                    v->quadc.as<uint32_t>(), width, height, e->stream));
    if (e->views.empty()) {
      e->mask0.ensure(np);
      upload_image(e->mask0.p, mask, width, height, stride, e->stream);
    }
    HIP_CHECK(hipStreamSynchronize(e->stream));
    e->views.push_back(std::move(v));
    e->have_result = false;
    return PM_OK;
  });
}

int pm_estimator_set_depth_range(pm_estimator* e, double min_depth,
                                 double max_depth, int num_depth_planes) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  e->min_depth = min_depth;
  e->max_depth = max_depth;
  e->num_depth_planes = num_depth_planes;
  g_last_error.clear();
  return PM_OK;
}

int pm_estimator_set_patchmatch_iterations(pm_estimator* e, int n) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (n < 0) return fail(PM_ERR_INVALID, "iterations must be >= 0");
  e->iterations = n;
  return PM_OK;
}

int pm_estimator_set_patch_size(pm_estimator* e, int size) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (size < 3 || size > 11 || size % 2 == 0)
    return fail(PM_ERR_INVALID, "patch size must be odd and in [3, 11], got " +
                                    std::to_string(size));
  e->patch_size = size;
  return PM_OK;
}

int pm_estimator_set_min_patch_sd(pm_estimator* e, float sd) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  e->min_patch_variance = sd * sd;
  return PM_OK;
}

int pm_estimator_set_seed(pm_estimator* e, uint64_t seed) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  e->seed = seed;
  return PM_OK;
}

int pm_estimator_set_refinement_steps(pm_estimator* e, int n) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (n < 0 || n > 48) return fail(PM_ERR_INVALID, "refinement steps must be in [0, 48]");
  e->n_refine = n;
  return PM_OK;
}

int pm_estimator_set_exploration(pm_estimator* e, int n) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (n < 0 || n > 8) return fail(PM_ERR_INVALID, "exploration samples must be in [0, 8]");
  e->n_explore = n;
  return PM_OK;
}

int pm_estimator_set_precision(pm_estimator* e, int precision) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (precision < PM_PRECISION_FP32 || precision > PM_PRECISION_BF16)
    return fail(PM_ERR_INVALID, "unknown precision " + std::to_string(precision));
  e->precision = precision;
  return PM_OK;
}

int pm_estimator_set_pyramid(pm_estimator* e, int half_iterations) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (half_iterations < 0 || half_iterations > 1000)
    return fail(PM_ERR_INVALID, "half-level iterations must be in [0, 1000]");
  e->pyramid_iters = half_iterations;
  e->level_iters.clear();
  return PM_OK;
}

int pm_estimator_set_pyramid_schedule(pm_estimator* e, const int32_t* iterations, int n) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (n < 1 || n > 7) return fail(PM_ERR_INVALID, "a schedule has 1 to 7 coarser levels");
  if (!iterations) return fail(PM_ERR_INVALID, "null schedule");
  for (int k = 0; k < n; ++k)
    if (iterations[k] < 1 || iterations[k] > 1000)
      return fail(PM_ERR_INVALID, "level iterations must be in [1, 1000]");
  e->level_iters.assign(iterations, iterations + n);
  e->pyramid_levels = n + 1;
  return PM_OK;
}

int pm_estimator_set_lanes_per_pixel(pm_estimator* e, int lanes) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (lanes != 0 && lanes != 1 && lanes != pm::kLanesPerPixelSmall)
    return fail(PM_ERR_INVALID, "lanes per pixel must be 0, 1 or " +
                                    std::to_string(pm::kLanesPerPixelSmall));
  e->lanes_per_pixel = lanes;
  return PM_OK;
}

int pm_estimator_set_pyramid_consistency(pm_estimator* e, int enable) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  e->pyramid_consistency = enable ? 1 : 0;
  return PM_OK;
}

int pm_estimator_set_pyramid_levels(pm_estimator* e, int levels) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (levels < 2 || levels > 8) return fail(PM_ERR_INVALID, "pyramid levels must be in [2, 8]");
  e->pyramid_levels = levels;
  e->level_iters.clear();
  return PM_OK;
}

int pm_estimator_set_far_probes(pm_estimator* e, int n) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (n < 0 || n > 64) return fail(PM_ERR_INVALID, "far probes must be in [0, 64]");
  e->n_far = n;
  return PM_OK;
}

int pm_estimator_set_consistency(pm_estimator* e, int enable) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  e->consistency = enable ? 1 : 0;
  return PM_OK;
}

int pm_estimator_set_post_process(pm_estimator* e, int enable) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  e->post_process = enable ? 1 : 0;
  return PM_OK;
}

int pm_estimator_run(pm_estimator* e, int mode) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  return guarded([&] { return run(e, mode); });
}

int pm_estimator_download(pm_estimator* e, float* depth, float* plane,
                          float* score, int32_t* nghbr) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  return guarded([&] { return download(e, depth, plane, score, nghbr); });
}

int pm_estimator_compute(pm_estimator* e, int mode, float* depth, float* plane,
                         float* score, int32_t* nghbr) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  return guarded([&] {
    int rc = run(e, mode);
    if (rc) return rc;
    return download(e, depth, plane, score, nghbr);
  });
}

int pm_estimator_score_hypotheses(pm_estimator* e, int n, const int32_t* ij,
                                  const float* planes, const int32_t* views,
                                  float* out) {
  return score_hypotheses(e, n, ij, planes, views, out, false);
}

int pm_estimator_score_hypotheses_fast(pm_estimator* e, int n, const int32_t* ij,
                                       const float* planes, const int32_t* views,
                                       float* out) {
  return score_hypotheses(e, n, ij, planes, views, out, true);
}

int pm_estimator_initializer_input(pm_estimator* e, int scale, int nplanes,
                                   void* out, int out_w, int out_h) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  int rc = validate_for_compute(e);
  if (rc) return rc;
  if (!out) return fail(PM_ERR_INVALID, "null output");
  if (scale < 1 || scale > 16) return fail(PM_ERR_INVALID, "scale must be in [1, 16]");
  if (nplanes < 2 || nplanes > 39) return fail(PM_ERR_INVALID, "nplanes must be in [2, 39]");
  if (e->patch_size != 5 && e->patch_size != 7)
    return fail(PM_ERR_INVALID, "the initialiser input needs patch size 5 or 7");
  const HostView& r = *e->views[0];
  const int wl = r.width / scale, hl = r.height / scale;
  if (wl < e->patch_size || hl < e->patch_size)
    return fail(PM_ERR_INVALID, "reference view too small for this scale");
  if (out_w < wl || out_h < hl)
    return fail(PM_ERR_INVALID, "output grid smaller than the coarse view (" +
                                    std::to_string(wl) + "x" + std::to_string(hl) + ")");
  return guarded([&] {
    DeviceGuard dg(e->device);
    build_coarse(e, scale);
    const int n = (int)e->views.size();
    const HostView& c0 = *e->coarse[0];
    std::vector<pm::PmView> tab(n);
    for (int v = 0; v < n; ++v) {
      const HostView& V = *e->coarse[v];
      fill_view(tab[v], nullptr, V.width, V.height, V.K, V.Q, V.a, c0.Kinv);
      tab[v].quad16 = V.quad16.as<uint32_t>();
    }
    e->coarse_viewtab.ensure(sizeof(pm::PmView) * n);
    HIP_CHECK(hipMemcpyAsync(e->coarse_viewtab.p, tab.data(), sizeof(pm::PmView) * n,
                             hipMemcpyHostToDevice, e->stream));
    const int hpz = (e->patch_size - 1) / 2;
    if (e->wlut_patch != e->patch_size) {
      std::vector<float> t = bilateral_table(hpz);
      e->wlut.ensure(t.size() * sizeof(float));
      HIP_CHECK(hipMemcpyAsync(e->wlut.p, t.data(), t.size() * sizeof(float),
                               hipMemcpyHostToDevice, e->stream));
      HIP_CHECK(hipStreamSynchronize(e->stream));
      e->wlut_patch = e->patch_size;
    }
    pm::PmArgs a;
    memset(&a, 0, sizeof(a));
    a.W = c0.width;
    a.H = c0.height;
    a.nviews = n;
    a.hpz = hpz;
    a.views = e->coarse_viewtab.as<pm::PmView>();
    a.ref_img = c0.img.as<uint8_t>();
    a.ref_mask = nullptr;
    a.wlut = e->wlut.as<float>();
    for (int k = 0; k < 9; ++k) {
      a.Kinv0f[k] = (float)c0.Kinv[k];
      a.Kinv0[k] = c0.Kinv[k];
    }
    a.precision = PM_PRECISION_FP16;
    const float inv_min = (float)(1.0 / e->min_depth);
    const float inv_step = (float)((1.0 / e->max_depth - 1.0 / e->min_depth) / (nplanes - 1));
    HIP_CHECK(pm::launch_init_input(a, nplanes, inv_min, inv_step, out, out_w, out_h, e->stream));
    HIP_CHECK(hipStreamSynchronize(e->stream));
    return PM_OK;
  });
}

int pm_estimator_set_init_spread(pm_estimator* e, float sigma_log_depth,
                                 float sigma_slope) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (!(sigma_log_depth >= 0) || !(sigma_slope >= 0) || sigma_log_depth > 1 || sigma_slope > 10)
    return fail(PM_ERR_INVALID, "spreads must be in [0, 1] and [0, 10]");
  e->init_sigma = sigma_log_depth;
  e->init_sigma_slope = sigma_slope;
  return PM_OK;
}

int pm_estimator_set_init_search(pm_estimator* e, int n) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (n < 1 || n > 15) return fail(PM_ERR_INVALID, "init search must be in [1, 15]");
  e->init_search = n;
  return PM_OK;
}

int pm_estimator_set_init_map(pm_estimator* e, const float* map, int w, int h,
                              int stride, int scale) {
  if (!e) return fail(PM_ERR_INVALID, "null estimator");
  if (map && (w < 1 || h < 1 || stride < 4 || scale < 1))
    return fail(PM_ERR_INVALID, "init map needs w, h >= 1, stride >= 4, scale >= 1");
  e->init_map = map;
  e->init_w = map ? w : 0;
  e->init_h = map ? h : 0;
  e->init_stride = map ? stride : 0;
  e->init_scale = map ? scale : 1;
  return PM_OK;
}

int pm_estimator_stats(pm_estimator* e, pm_stats* out) {
  if (!e || !out) return fail(PM_ERR_INVALID, "null argument");
  *out = e->stats;
  return PM_OK;
}

}  // extern "C"

// ===========================================================================
// DepthmapCleaner

struct CleanerView {
  double K[9], R[9], t[3];
  int width, height;
  DevBuf depth;
};

struct pm_cleaner {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  float same_depth_threshold = 0.01f;
  int min_consistent_views = 2;
  std::vector<std::unique_ptr<CleanerView>> views;
  DevBuf table, ref, out;
  float last_ms = 0;
  ~pm_cleaner() {
    for (auto& e : ev)
      if (e) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
  }
};

extern "C" {

int pm_cleaner_create(int device, pm_cleaner** out) {
  if (!out) return fail(PM_ERR_INVALID, "null handle pointer");
  *out = nullptr;
  int rc = check_device(device);
  if (rc) return rc;
  return guarded([&] {
    std::unique_ptr<pm_cleaner> c(new pm_cleaner());
    c->device = device;
    DeviceGuard dg(device);
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& ev : c->ev) HIP_CHECK(hipEventCreate(&ev));
    *out = c.release();
    return PM_OK;
  });
}

int pm_cleaner_destroy(pm_cleaner* c) {
  if (!c) return PM_OK;
  return guarded([&] {
    DeviceGuard dg(c->device);
    hipStreamSynchronize(c->stream);
    delete c;
    return PM_OK;
  });
}

int pm_cleaner_set_same_depth_threshold(pm_cleaner* c, float t) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  c->same_depth_threshold = t;
  return PM_OK;
}

int pm_cleaner_set_min_consistent_views(pm_cleaner* c, int n) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  c->min_consistent_views = n;
  return PM_OK;
}

int pm_cleaner_add_view(pm_cleaner* c, const double* K, const double* R,
                        const double* t, const float* depth, int width,
                        int height) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  if (!K || !R || !t || !depth) return fail(PM_ERR_INVALID, "null K, R, t or depth");
  if (width < 1 || height < 1) return fail(PM_ERR_INVALID, "empty depthmap");
  if (!all_finite(K, 9) || !all_finite(R, 9) || !all_finite(t, 3))
    return fail(PM_ERR_INVALID, "K, R and t must be finite");
  if ((int)c->views.size() >= 1024) return fail(PM_ERR_INVALID, "too many views");
  return guarded([&] {
    std::unique_ptr<CleanerView> v(new CleanerView());
    memcpy(v->K, K, sizeof(v->K));
    memcpy(v->R, R, sizeof(v->R));
    memcpy(v->t, t, sizeof(v->t));
    v->width = width;
    v->height = height;
    DeviceGuard dg(c->device);
    const size_t bytes = (size_t)width * height * sizeof(float);
    v->depth.ensure(bytes);
    HIP_CHECK(hipMemcpyAsync(v->depth.p, depth, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->views.push_back(std::move(v));
    return PM_OK;
  });
}

int pm_cleaner_clean(pm_cleaner* c, float* clean_depth) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  if (!clean_depth) return fail(PM_ERR_INVALID, "null output");
  if (c->views.empty()) return fail(PM_ERR_INVALID, "no views added");
  return guarded([&] {
    DeviceGuard dg(c->device);
    const int n = (int)c->views.size();
    std::vector<pm::ClView> tab(n);
    for (int v = 0; v < n; ++v) {
      const CleanerView& V = *c->views[v];
      tab[v].depth = V.depth.as<float>();
      tab[v].width = V.width;
      tab[v].height = V.height;
      memcpy(tab[v].K, V.K, sizeof(V.K));
      memcpy(tab[v].R, V.R, sizeof(V.R));
      memcpy(tab[v].t, V.t, sizeof(V.t));
    }
    pm::ClRef ref;
    if (!inv3(c->views[0]->K, ref.Kinv)) return fail(PM_ERR_INVALID, "K of view 0 is singular");
    mat3_transpose(c->views[0]->R, ref.Rt);
    c->table.ensure(sizeof(pm::ClView) * n);
    c->ref.ensure(sizeof(pm::ClRef));
    const CleanerView& V0 = *c->views[0];
    const size_t np = (size_t)V0.width * V0.height;
    c->out.ensure(np * sizeof(float));
    HIP_CHECK(hipMemcpyAsync(c->table.p, tab.data(), sizeof(pm::ClView) * n,
                             hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(c->ref.p, &ref, sizeof(ref), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipEventRecord(c->ev[0], c->stream));
    HIP_CHECK(pm::launch_clean(c->table.as<pm::ClView>(), c->ref.as<pm::ClRef>(), n,
                               V0.width, V0.height, c->same_depth_threshold,
                               c->min_consistent_views, c->out.as<float>(), c->stream));
    HIP_CHECK(hipEventRecord(c->ev[1], c->stream));
    HIP_CHECK(hipMemcpyAsync(clean_depth, c->out.p, np * sizeof(float),
                             hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms, c->ev[0], c->ev[1]));
    return PM_OK;
  });
}

int pm_cleaner_last_ms(pm_cleaner* c, float* ms) {
  if (!c || !ms) return fail(PM_ERR_INVALID, "null argument");
  *ms = c->last_ms;
  return PM_OK;
}


int pm_gpu_cleaner_create(int device, pm_gpu_cleaner** out) {
  if (!out) return fail(PM_ERR_INVALID, "null output handle");
  *out = nullptr;
  int rc = check_device(device);
  if (rc) return rc;
  return guarded([&] {
    std::unique_ptr<pm_gpu_cleaner> c(new pm_gpu_cleaner());
    c->device = device;
    DeviceGuard dg(device);
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIP_CHECK(hipEventCreate(&e));
    *out = c.release();
    return PM_OK;
  });
}

int pm_gpu_cleaner_destroy(pm_gpu_cleaner* c) {
  if (!c) return PM_OK;
  return guarded([&] {
    DeviceGuard dg(c->device);
    delete c;
    return PM_OK;
  });
}

int pm_gpu_cleaner_set_same_depth_threshold(pm_gpu_cleaner* c, float t) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  c->same_depth_threshold = t;
  return PM_OK;
}

int pm_gpu_cleaner_set_min_consistent_views(pm_gpu_cleaner* c, int n) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  c->min_consistent_views = n;
  return PM_OK;
}

int pm_gpu_cleaner_set_carving_threshold(pm_gpu_cleaner* c, float t) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  c->carving_threshold = t > 0.0f ? t : 0.0f;  // cleaner.cc:30-32
  return PM_OK;
}

int pm_gpu_cleaner_set_max_carved_views(pm_gpu_cleaner* c, int n) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  c->max_carved_views = n > 0 ? n : 0;  // cleaner.cc:34-36
  return PM_OK;
}

int pm_gpu_cleaner_add_view(pm_gpu_cleaner* c, const double* K, const double* R,
                            const double* t, const float* depth, int width,
                            int height, int* index) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  if (!K || !R || !t || !depth) return fail(PM_ERR_INVALID, "null K, R, t or depth");
  if (width < 2 || height < 2) return fail(PM_ERR_INVALID, "depth map must be at least 2x2");
  return guarded([&] {
    std::unique_ptr<GcView> v(new GcView());
    memcpy(v->K, K, sizeof(v->K));
    memcpy(v->R, R, sizeof(v->R));
    memcpy(v->t, t, sizeof(v->t));
    v->width = width;
    v->height = height;
    DeviceGuard dg(c->device);
    const size_t bytes = (size_t)width * height * sizeof(float);
    v->depth.ensure(bytes);
    HIP_CHECK(hipMemcpyAsync(v->depth.p, depth, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if (index) *index = (int)c->views.size();
    c->views.push_back(std::move(v));
    return PM_OK;
  });
}

int pm_gpu_cleaner_clean(pm_gpu_cleaner* c, int ref_idx, const int32_t* neighbor_ids,
                         int n, float* clean_depth) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  if (!clean_depth) return fail(PM_ERR_INVALID, "null output");
  if (n < 0 || (n > 0 && !neighbor_ids)) return fail(PM_ERR_INVALID, "neighbour list");
  const int nv = (int)c->views.size();
  // cleaner.cc:68-81
  if (ref_idx < 0 || ref_idx >= nv)
    return fail(PM_ERR_RANGE, "Clean: ref_idx " + std::to_string(ref_idx) +
                                  " out of range [0, " + std::to_string(nv) + ")");
  for (int i = 0; i < n; ++i)
    if (neighbor_ids[i] < 0 || neighbor_ids[i] >= nv)
      return fail(PM_ERR_RANGE, "Clean: neighbor_ids[" + std::to_string(i) + "] = " +
                                    std::to_string(neighbor_ids[i]) + " out of range [0, " +
                                    std::to_string(nv) + ")");
  return guarded([&] {
    DeviceGuard dg(c->device);
    const GcView& R = *c->views[ref_idx];
    const int w = R.width, h = R.height;
    const size_t np = (size_t)w * h;
    const int nsrc = std::min(n, 16);  // MAX_CLEAN_SOURCES, opencl_kernels.h:2103
    c->resized.ensure(np * sizeof(float) * std::max(1, nsrc));
    std::vector<const float*> ptrs(std::max(1, nsrc), nullptr);
    std::vector<pm::GcCamera> cams(1 + nsrc);
    auto cam = [](const GcView& V, pm::GcCamera& g) {
      for (int k = 0; k < 9; ++k) {
        g.K[k] = (float)V.K[k];
        g.R[k] = (float)V.R[k];
      }
      for (int k = 0; k < 3; ++k) g.t[k] = (float)V.t[k];
    };
    cam(R, cams[0]);
    for (int v = 0; v < nsrc; ++v) {
      const GcView& S = *c->views[neighbor_ids[v]];
      float* dst = c->resized.as<float>() + (size_t)v * np;
      if (S.width == w && S.height == h) {
        HIP_CHECK(hipMemcpyAsync(dst, S.depth.p, np * sizeof(float), hipMemcpyDeviceToDevice,
                                 c->stream));
      } else {
        HIP_CHECK(pm::launch_resize_linear(S.depth.as<float>(), S.width, S.height, dst, w, h,
                                           c->stream));
      }
      ptrs[v] = dst;
      cam(S, cams[1 + v]);
    }
    c->ptrs.ensure(ptrs.size() * sizeof(float*));
    c->cams.ensure(cams.size() * sizeof(pm::GcCamera));
    c->out.ensure(np * sizeof(float));
    HIP_CHECK(hipMemcpyAsync(c->ptrs.p, ptrs.data(), ptrs.size() * sizeof(float*),
                             hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(c->cams.p, cams.data(), cams.size() * sizeof(pm::GcCamera),
                             hipMemcpyHostToDevice, c->stream));
    const int eff_min = std::min(c->min_consistent_views, 1 + n);  // cleaner.cc:152
    HIP_CHECK(hipEventRecord(c->ev[0], c->stream));
    HIP_CHECK(pm::launch_gpu_clean(R.depth.as<float>(), c->ptrs.as<const float*>(),
                                   c->cams.as<pm::GcCamera>(), w, h, nsrc,
                                   c->same_depth_threshold, eff_min, c->carving_threshold,
                                   c->max_carved_views, c->out.as<float>(), c->stream));
    HIP_CHECK(hipEventRecord(c->ev[1], c->stream));
    HIP_CHECK(hipMemcpyAsync(clean_depth, c->out.p, np * sizeof(float), hipMemcpyDeviceToHost,
                             c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms, c->ev[0], c->ev[1]));
    return PM_OK;
  });
}

int pm_gpu_cleaner_clear(pm_gpu_cleaner* c) {
  if (!c) return fail(PM_ERR_INVALID, "null cleaner");
  return guarded([&] {
    DeviceGuard dg(c->device);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->views.clear();
    c->resized.release();
    c->ptrs.release();
    c->cams.release();
    c->out.release();
    return PM_OK;
  });
}

int pm_gpu_cleaner_last_ms(pm_gpu_cleaner* c, float* ms) {
  if (!c || !ms) return fail(PM_ERR_INVALID, "null argument");
  *ms = c->last_ms;
  return PM_OK;
}

}  // extern "C"

// ===========================================================================
// DepthmapPruner (depthmap.h:131-153, depthmap.cc:557-637)

struct PrHostView {
  double K[9], R[9], t[3];
  int width, height;
  DevBuf depth, plane;
};

struct pm_pruner {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  float same_depth_threshold = 0.01f;  // depthmap.cc:557
  std::vector<std::unique_ptr<PrHostView>> views;
  DevBuf color, label;  // of view 0, the only one the merge reads
  DevBuf table, counts, offsets, points, normals, out_color, out_label;
  int64_t count = -1;  // kept points of the last Prune, -1 before one
  float last_ms = 0;
  ~pm_pruner() {
    for (auto& e : ev)
      if (e) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
  }
};

extern "C" {

int pm_pruner_create(int device, pm_pruner** out) {
  if (!out) return fail(PM_ERR_INVALID, "null output handle");
  *out = nullptr;
  int rc = check_device(device);
  if (rc) return rc;
  return guarded([&] {
    std::unique_ptr<pm_pruner> c(new pm_pruner());
    c->device = device;
    DeviceGuard dg(device);
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIP_CHECK(hipEventCreate(&e));
    *out = c.release();
    return PM_OK;
  });
}

int pm_pruner_destroy(pm_pruner* c) {
  if (!c) return PM_OK;
  return guarded([&] {
    DeviceGuard dg(c->device);
    delete c;
    return PM_OK;
  });
}

int pm_pruner_set_same_depth_threshold(pm_pruner* c, float t) {
  if (!c) return fail(PM_ERR_INVALID, "null pruner");
  c->same_depth_threshold = t;
  return PM_OK;
}

int pm_pruner_add_view(pm_pruner* c, const double* K, const double* R, const double* t,
                       const float* depth, const float* plane, const uint8_t* color,
                       const uint8_t* label, int width, int height) {
  if (!c) return fail(PM_ERR_INVALID, "null pruner");
  if (!K || !R || !t || !depth || !plane || !color || !label)
    return fail(PM_ERR_INVALID, "null K, R, t, depth, plane, color or label");
  if (width < 1 || height < 1) return fail(PM_ERR_INVALID, "empty depth map");
  if ((int64_t)width * height > (int64_t)INT32_MAX)
    return fail(PM_ERR_INVALID, "depth map larger than 2^31 pixels");
  if (c->views.empty()) {
    double Kinv[9];
    if (!inv3(K, Kinv)) return fail(PM_ERR_INVALID, "singular K of the first view");
  }
  return guarded([&] {
    std::unique_ptr<PrHostView> v(new PrHostView());
    memcpy(v->K, K, sizeof(v->K));
    memcpy(v->R, R, sizeof(v->R));
    memcpy(v->t, t, sizeof(v->t));
    v->width = width;
    v->height = height;
    DeviceGuard dg(c->device);
    const size_t np = (size_t)width * height;
    v->depth.ensure(np * sizeof(float));
    v->plane.ensure(np * 3 * sizeof(float));
    HIP_CHECK(hipMemcpyAsync(v->depth.p, depth, np * sizeof(float), hipMemcpyHostToDevice,
                             c->stream));
    HIP_CHECK(hipMemcpyAsync(v->plane.p, plane, np * 3 * sizeof(float), hipMemcpyHostToDevice,
                             c->stream));
    if (c->views.empty()) {
      c->color.ensure(np * 3);
      c->label.ensure(np);
      HIP_CHECK(hipMemcpyAsync(c->color.p, color, np * 3, hipMemcpyHostToDevice, c->stream));
      HIP_CHECK(hipMemcpyAsync(c->label.p, label, np, hipMemcpyHostToDevice, c->stream));
    }
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->views.push_back(std::move(v));
    c->count = -1;
    return PM_OK;
  });
}

int pm_pruner_prune(pm_pruner* c, int64_t* count) {
  if (!c || !count) return fail(PM_ERR_INVALID, "null argument");
  if (c->views.empty()) return fail(PM_ERR_INVALID, "Prune needs at least one view");
  return guarded([&] {
    DeviceGuard dg(c->device);
    const int n = (int)c->views.size();
    const PrHostView& V0 = *c->views[0];
    const int npix = V0.width * V0.height;
    const int nblk = (npix + 255) / 256;
    std::vector<pm::PrView> table(n);
    for (int v = 0; v < n; ++v) {
      const PrHostView& H = *c->views[v];
      pm::PrView& T = table[v];
      T.depth = H.depth.as<float>();
      T.plane = H.plane.as<float>();
      T.width = H.width;
      T.height = H.height;
      memcpy(T.K, H.K, sizeof(T.K));
      memcpy(T.R, H.R, sizeof(T.R));
      memcpy(T.t, H.t, sizeof(T.t));
    }
    pm::PrRef ref;
    inv3(V0.K, ref.Kinv);
    mat3_transpose(V0.R, ref.Rt);
    for (int k = 0; k < 9; ++k) ref.Rinv[k] = (float)ref.Rt[k];
    c->table.ensure(n * sizeof(pm::PrView));
    c->counts.ensure(nblk * sizeof(int));
    c->offsets.ensure(nblk * sizeof(long long));
    HIP_CHECK(hipMemcpyAsync(c->table.p, table.data(), n * sizeof(pm::PrView),
                             hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipEventRecord(c->ev[0], c->stream));
    HIP_CHECK(pm::launch_prune_count(c->table.as<pm::PrView>(), n, ref, c->same_depth_threshold,
                                     npix, c->counts.as<int>(), c->stream));
    // The block offsets are an exclusive scan of at most ~2^23 counts, done
    // on the host: the total sizes the output buffers anyway.
    std::vector<int> counts(nblk);
    HIP_CHECK(hipMemcpyAsync(counts.data(), c->counts.p, nblk * sizeof(int),
                             hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    std::vector<long long> offsets(nblk);
    long long total = 0;
    for (int b = 0; b < nblk; ++b) {
      offsets[b] = total;
      total += counts[b];
    }
    const size_t cap = (size_t)std::max<long long>(total, 1);
    c->points.ensure(cap * 3 * sizeof(float));
    c->normals.ensure(cap * 3 * sizeof(float));
    c->out_color.ensure(cap * 3);
    c->out_label.ensure(cap);
    HIP_CHECK(hipMemcpyAsync(c->offsets.p, offsets.data(), nblk * sizeof(long long),
                             hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(pm::launch_prune_write(
        c->table.as<pm::PrView>(), n, ref, c->same_depth_threshold, npix,
        c->offsets.as<long long>(), c->color.as<uint8_t>(), c->label.as<uint8_t>(),
        c->points.as<float>(), c->normals.as<float>(), c->out_color.as<uint8_t>(),
        c->out_label.as<uint8_t>(), c->stream));
    HIP_CHECK(hipEventRecord(c->ev[1], c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms, c->ev[0], c->ev[1]));
    c->count = total;
    *count = total;
    return PM_OK;
  });
}

int pm_pruner_download(pm_pruner* c, float* points, float* normals, uint8_t* colors,
                       uint8_t* labels) {
  if (!c) return fail(PM_ERR_INVALID, "null pruner");
  if (c->count < 0) return fail(PM_ERR_INVALID, "download before prune");
  if (c->count == 0) return PM_OK;
  if (!points || !normals || !colors || !labels) return fail(PM_ERR_INVALID, "null output");
  return guarded([&] {
    DeviceGuard dg(c->device);
    const size_t m = (size_t)c->count;
    HIP_CHECK(hipMemcpyAsync(points, c->points.p, m * 3 * sizeof(float), hipMemcpyDeviceToHost,
                             c->stream));
    HIP_CHECK(hipMemcpyAsync(normals, c->normals.p, m * 3 * sizeof(float),
                             hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(colors, c->out_color.p, m * 3, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(labels, c->out_label.p, m, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return PM_OK;
  });
}

int pm_pruner_last_ms(pm_pruner* c, float* ms) {
  if (!c || !ms) return fail(PM_ERR_INVALID, "null argument");
  *ms = c->last_ms;
  return PM_OK;
}

}  // extern "C"
