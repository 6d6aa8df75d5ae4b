// Host side of the batched C ABI (pm_batch_* in include/opensfm_amd.h).
//
// A batch holds a pool of views (camera + image, uploaded once) and a list
// of jobs, each a reference view with its neighbour views and depth range:
// job k computes what one DepthmapEstimator with add_view(ref),
// add_view(neighbours...) and set_seed(seed + k) computes (depthmap.cc:
// 146-227), bit for bit.  All jobs run together: per-pixel state lives in
// flat device buffers at per-job offsets, and each launch walks the tiles of
// every job (PmJob table, pm_types.h).  A run -- clears, random init,
// 2 x iterations red/black launches, output packing -- is captured once in a
// hipGraph and replayed while the configuration is unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pm_host.h"
#include "pm_kernels.h"
#include "pm_types.h"

using namespace pmhost;

namespace {

struct PoolView {
  double K[9], R[9], t[3], Kinv[9];
  int width = 0, height = 0;
  DevBuf img, mask, quad;
  DevBuf quadc;        // FP16 bilinear coefficients of the quads
  DevBuf quad16;       // 16-bit quads, built on first use
  int quad16_kind = 0;  // precision they were built for (0: none)
};

// A pool view at 1/scale for the learned initialiser's input (box mean of
// the image, K scaled about pixel centres, FP16 quads), as the estimator's
// coarse views.
struct CoarseView {
  double K[9], Kinv[9];
  int width = 0, height = 0;
  DevBuf img, quad16;
};

struct BatchJob {
  int ref = 0;
  std::vector<int> nb;
  double min_depth = 0, max_depth = 0;
  size_t off = 0;  // first pixel in the flat per-pixel buffers
  // learned initialisation map of this job (device memory; null: random)
  const float* init_map = nullptr;
  int init_w = 0, init_h = 0, init_stride = 0, init_scale = 1;
};

}  // namespace

struct pm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // DepthmapEstimator defaults (depthmap.cc:134-144) plus the engine's own
  int patch_size = 7;
  int iterations = 3;
  float min_patch_variance = 5 * 5;
  uint64_t seed = 0x5EEDull;
  int n_far = 6, n_refine = 6, n_explore = 3;
  int init_search = 1;
  float init_sigma = 0.05f, init_sigma_slope = 0.3f;
  int post_process = 1;
  int consistency = 1;
  int tile = 1;
  int use_graph = 1;
  int precision = PM_PRECISION_FP32;
  // half-resolution level (set_pyramid): a batch over the pool views
  // box-averaged by 2 with the same jobs, run first on this batch's stream;
  // its planes seed plain-mode runs here through per-job initialisation maps;
  // with pyramid_levels > 2 that batch has a half level of its own, and so
  // on, every coarser level running pyramid_iters iterations
  int pyramid_iters = 0;
  int pyramid_levels = 2;
  std::vector<int> level_iters;  // explicit schedule, coarsest first
  std::unique_ptr<pm_batch> half;
  DevBuf half_map;  // the half level's state as initialisation maps
  DevBuf jobtab_pyr;  // the job table with those maps (plain mode)
  bool own_stream = true;
  int top_consistency = -1;  // a coarser level: the full level's setting
  int pyramid_consistency = 1;  // the re-selection at the coarser levels too
  std::vector<std::unique_ptr<PoolView>> views;
  std::vector<BatchJob> jobs;
  // device state of the prepared configuration
  DevBuf viewtab, jobtab, wlut, state, score, nghbr;
  DevBuf out_depth, out_plane, out_score, out_nghbr, queue, counter;
  DevBuf cons_count, cons_idx, cons_med, cons_job;  // median-consistent re-selection list
  // and its window-phase work items (pm::ConsistItems)
  DevBuf item_count, item_entry, item_win, item_score, item_nghbr, item_depth, item_first,
      item_todo;
  unsigned item_cap = 0;
  // coarse pool views of the learned initialiser's input
  int coarse_scale = 0;
  std::vector<std::unique_ptr<CoarseView>> coarse;
  DevBuf coarse_viewtab, coarse_wlut;
  int coarse_wlut_patch = -1;
  DevBuf cluster_out;  // staging of pm_batch_download_cluster
  DevBuf cl_views, cl_jobs, cl_out;
  int wlut_patch = -1;
  int prepared_jobs = -1;  // jobs in the uploaded table (-1: stale)
  size_t total_px = 0;
  int half_total = 0, pix_total = 0;
  hipGraphExec_t exec = nullptr;
  int graph_mode = -1;  // mode the graph was captured for (-1: none)
  bool have_result = false, have_clean = false;
  pm_stats stats{};
  float clean_ms = 0;

  void invalidate() {
    prepared_jobs = -1;
    drop_graph();
    have_result = have_clean = false;
  }
  void drop_graph() {
    if (exec) hipGraphExecDestroy(exec);
    exec = nullptr;
    graph_mode = -1;
  }
  ~pm_batch() {
    drop_graph();
    for (auto& e : ev)
      if (e) hipEventDestroy(e);
    if (stream && own_stream) hipStreamDestroy(stream);
  }
};

namespace {

// Iterations of the coarser levels, coarsest first (empty: no pyramid).
std::vector<int> pyramid_schedule(const pm_batch* b) {
  if (!b->level_iters.empty()) return b->level_iters;
  if (b->pyramid_iters <= 0) return {};
  return std::vector<int>(b->pyramid_levels - 1, b->pyramid_iters);
}

int validate(pm_batch* b) {
  if (b->jobs.empty()) return fail(PM_ERR_INVALID, "batch has no jobs");
  if (!pm::precision_supported(b->precision, b->patch_size))
    return fail(PM_ERR_INVALID, "FP16/BF16 samples need patch size 5 or 7");
  const int hpz = (b->patch_size - 1) / 2;
  // with the pyramid, the coarsest level's views (box-averaged by
  // 2^(levels - 1)): references at least one patch, the others 2 x 2
  const int sh = (int)pyramid_schedule(b).size();
  for (size_t k = 0; k < b->jobs.size(); ++k) {
    const PoolView& r = *b->views[b->jobs[k].ref];
    if ((r.width >> sh) < 2 * hpz + 1 || (r.height >> sh) < 2 * hpz + 1)
      return fail(PM_ERR_INVALID, "job " + std::to_string(k) +
                                      (sh ? ": reference view smaller than one patch at the "
                                            "pyramid's coarsest level"
                                          : ": reference view smaller than one patch"));
  }
  if (sh) {
    for (const auto& V : b->views)
      if ((V->width >> sh) < 2 || (V->height >> sh) < 2)
        return fail(PM_ERR_INVALID, "a view is too small for the pyramid's coarsest level");
    for (size_t k = 0; k < b->jobs.size(); ++k)
      if (b->jobs[k].init_map && b->jobs[k].init_scale % (1 << sh) != 0)
        return fail(PM_ERR_INVALID, "job " + std::to_string(k) +
                                        ": an initialisation map seeding the pyramid needs a "
                                        "scale divisible by " + std::to_string(1 << sh));
  }
  return PM_OK;
}

int ceil_div(int a, int b) { return (a + b - 1) / b; }

void prepare(pm_batch* b, int count);

// The half-resolution level of b: every pool view box-averaged by 2 (K
// scaled about pixel centres, x_half = (x + 0.5) / 2 - 0.5; poses unchanged;
// masks the minimum of each 2 x 2 block), the same jobs, b's settings with
// pyramid_iters iterations and seed + kHalfLevelSeed, on b's stream (so that
// a captured run holds both levels).  Rebuilt when views or jobs are added.
void ensure_half(pm_batch* b, int count) {
  if (!b->half || b->half->views.size() != b->views.size() ||
      b->half->jobs.size() != b->jobs.size()) {
    std::unique_ptr<pm_batch> h(new pm_batch());
    h->device = b->device;
    h->stream = b->stream;
    h->own_stream = false;
    for (auto& e : h->ev) HIP_CHECK(hipEventCreate(&e));
    for (const auto& Vp : b->views) {
      const PoolView& V = *Vp;
      std::unique_ptr<PoolView> c(new PoolView());
      memcpy(c->R, V.R, sizeof(c->R));
      memcpy(c->t, V.t, sizeof(c->t));
      memcpy(c->K, V.K, sizeof(c->K));
      for (int r = 0; r < 2; ++r) {
        for (int col = 0; col < 3; ++col) c->K[r * 3 + col] *= 0.5;
        c->K[r * 3 + 2] -= 0.25 * V.K[8];
      }
      if (!inv3(c->K, c->Kinv)) throw std::runtime_error("half-level K is singular");
      c->width = V.width / 2;
      c->height = V.height / 2;
      if (c->width < 2 || c->height < 2)
        throw std::runtime_error("view too small for a half level");
      const size_t np = (size_t)c->width * c->height;
      c->img.ensure(np);
      c->mask.ensure(np);
      HIP_CHECK(pm::launch_downsample_u8(V.img.as<uint8_t>(), V.width, V.height, 2, 0,
                                         c->img.as<uint8_t>(), c->width, c->height, b->stream));
      HIP_CHECK(pm::launch_downsample_u8(V.mask.as<uint8_t>(), V.width, V.height, 2, 1,
                                         c->mask.as<uint8_t>(), c->width, c->height,
                                         b->stream));
      c->quad.ensure((size_t)(c->width + 1) * (c->height + 1) * sizeof(uint32_t));
      c->quadc.ensure((size_t)(c->width + 1) * (c->height + 1) * 8);
      HIP_CHECK(pm::launch_build_quad(c->img.as<uint8_t>(), c->quad.as<uint32_t>(),
                                      c->quadc.as<uint32_t>(), c->width, c->height, b->stream));
      h->views.push_back(std::move(c));
    }
    for (const BatchJob& J : b->jobs) {
      BatchJob c;
      c.ref = J.ref;
      c.nb = J.nb;
      c.min_depth = J.min_depth;
      c.max_depth = J.max_depth;
      h->jobs.push_back(std::move(c));
    }
    HIP_CHECK(hipStreamSynchronize(b->stream));
    b->half = std::move(h);
  }
  pm_batch* h = b->half.get();
  h->patch_size = b->patch_size;
  // a job's own initialisation map seeds the coarsest level (its cells cover
  // half as many pixels of each coarser level)
  for (size_t k = 0; k < b->jobs.size(); ++k) {
    const BatchJob& J = b->jobs[k];
    BatchJob& c = h->jobs[k];
    c.init_map = J.init_map;
    c.init_w = J.init_w;
    c.init_h = J.init_h;
    c.init_stride = J.init_stride;
    c.init_scale = J.init_map ? J.init_scale / 2 : 1;
  }
  h->init_search = b->init_search;
  h->init_sigma = b->init_sigma;
  h->init_sigma_slope = b->init_sigma_slope;
  std::vector<int> sched = pyramid_schedule(b);
  h->iterations = sched.back();
  sched.pop_back();
  h->level_iters = sched;  // empty: the coarsest level
  h->pyramid_iters = 0;
  h->pyramid_levels = (int)sched.size() + 1;
  h->min_patch_variance = b->min_patch_variance;
  const int top = b->top_consistency >= 0 ? b->top_consistency : b->consistency;
  h->top_consistency = top;
  h->pyramid_consistency = b->pyramid_consistency;
  h->seed = b->seed + pm::kHalfLevelSeed;
  h->n_far = b->n_far;
  h->n_refine = b->n_refine;
  h->n_explore = b->n_explore;
  h->post_process = b->post_process;
  h->consistency = top && b->pyramid_consistency;  // estimator: run_half
  h->tile = b->tile;
  h->precision = b->precision;
  h->use_graph = 0;
  prepare(h, count);
}

// Uploads the view and job tables for the first `count` jobs and sizes the
// per-pixel buffers.  Host-synchronous.
void prepare(pm_batch* b, int count) {
  const int hpz = (b->patch_size - 1) / 2;
  if (b->wlut_patch != b->patch_size) {
    std::vector<float> t = bilateral_table(hpz);
    b->wlut.ensure(t.size() * sizeof(float));
    HIP_CHECK(hipMemcpy(b->wlut.p, t.data(), t.size() * sizeof(float),
                        hipMemcpyHostToDevice));
    b->wlut_patch = b->patch_size;
  }
  int tx, ty;
  pm::tile_shape(b->tile, &tx, &ty);
  if (b->precision != PM_PRECISION_FP32) {
    for (int k = 0; k < count; ++k) {
      std::vector<int> used(b->jobs[k].nb);
      used.push_back(b->jobs[k].ref);
      for (int v : used) {
        PoolView& V = *b->views[v];
        if (V.quad16_kind == b->precision) continue;
        V.quad16.ensure((size_t)(V.width + 1) * (V.height + 1) * 8);
        HIP_CHECK(pm::launch_build_quad16(V.img.as<uint8_t>(), V.quad16.as<uint32_t>(),
                                          V.width, V.height, b->precision, b->stream));
        V.quad16_kind = b->precision;
      }
    }
    HIP_CHECK(hipStreamSynchronize(b->stream));
  }
  std::vector<pm::PmView> vt;
  std::vector<int> view0(count);
  size_t px = 0;
  for (int k = 0; k < count; ++k) {
    BatchJob& J = b->jobs[k];
    const PoolView& R = *b->views[J.ref];
    J.off = px;
    px += (size_t)R.width * R.height;
    view0[k] = (int)vt.size();
    // the reference itself first, with its pose relative to itself as
    // AddView computes it for view 0
    double Q[9], a[3];
    relative_pose(R.R, R.t, R.R, R.t, Q, a);
    vt.emplace_back();
    fill_view(vt.back(), R.quad.as<uint32_t>(), R.width, R.height, R.K, Q, a,
              R.Kinv);
    vt.back().quad16 = R.quad16.as<uint32_t>();
    vt.back().quadc = R.quadc.as<uint32_t>();
    for (int v : J.nb) {
      const PoolView& V = *b->views[v];
      relative_pose(R.R, R.t, V.R, V.t, Q, a);
      vt.emplace_back();
      fill_view(vt.back(), V.quad.as<uint32_t>(), V.width, V.height, V.K, Q,
                a, R.Kinv);
      vt.back().quad16 = V.quad16.as<uint32_t>();
      vt.back().quadc = V.quadc.as<uint32_t>();
    }
  }
  b->total_px = px;
  b->viewtab.ensure(vt.size() * sizeof(pm::PmView));
  HIP_CHECK(hipMemcpy(b->viewtab.p, vt.data(), vt.size() * sizeof(pm::PmView),
                      hipMemcpyHostToDevice));
  b->state.ensure(px * 4 * sizeof(float));
  b->score.ensure(px * sizeof(float));
  b->nghbr.ensure(px * sizeof(int));
  b->out_depth.ensure(px * sizeof(float));
  b->out_plane.ensure(px * 3 * sizeof(float));
  b->out_score.ensure(px * sizeof(float));
  b->out_nghbr.ensure(px * sizeof(int));
  if (b->consistency) {
    b->cons_count.ensure(sizeof(unsigned));
    b->cons_idx.ensure(px * sizeof(unsigned));
    b->cons_med.ensure(px * sizeof(float));
    b->cons_job.ensure(px * sizeof(int));
    // window items: room for two per pixel, at most 2^26 (more are scored
    // inline)
    const size_t cap = std::min<size_t>(2 * px, (size_t)1 << 26);
    b->item_count.ensure(sizeof(unsigned long long));
    b->item_entry.ensure(cap * sizeof(unsigned));
    b->item_win.ensure(cap);
    b->item_score.ensure(cap * sizeof(float));
    b->item_nghbr.ensure(cap * sizeof(int));
    b->item_depth.ensure(cap * sizeof(float));
    b->item_first.ensure(px * sizeof(unsigned));
    b->item_todo.ensure(px * sizeof(unsigned));
    b->item_cap = (unsigned)cap;
  }
  // one init queue, then 8 band queues per red/black launch
  b->queue.ensure((1 + 16 * (size_t)std::max(b->iterations, 1)) * sizeof(unsigned));
  b->counter.ensure(2 * sizeof(unsigned long long));

  std::vector<pm::PmJob> jt(count);
  int half_t = 0, pix_t = 0;
  for (int k = 0; k < count; ++k) {
    const BatchJob& J = b->jobs[k];
    const PoolView& R = *b->views[J.ref];
    pm::PmJob& T = jt[k];
    memset(&T, 0, sizeof(T));
    pm::PmArgs& a = T.a;
    a.W = R.width;
    a.H = R.height;
    a.nviews = 1 + (int)J.nb.size();
    a.hpz = hpz;
    a.views = b->viewtab.as<pm::PmView>() + view0[k];
    a.ref_img = R.img.as<uint8_t>();
    a.ref_mask = R.mask.as<uint8_t>();
    a.wlut = b->wlut.as<float>();
    a.state = b->state.as<float>() + 4 * J.off;
    a.score = b->score.as<float>() + J.off;
    a.nghbr = b->nghbr.as<int>() + J.off;
    for (int m = 0; m < 9; ++m) {
      a.Kinv0f[m] = (float)R.Kinv[m];
      a.Kinv0[m] = R.Kinv[m];
    }
    a.log_min = (float)log(J.min_depth);
    a.log_max = (float)log(J.max_depth);
    a.min_patch_var = b->min_patch_variance;
    const uint64_t s = b->seed + (uint64_t)k;
    a.seed_lo = (uint32_t)(s & 0xffffffffu);
    a.seed_hi = (uint32_t)(s >> 32);
    a.n_far = b->n_far;
    a.n_refine = b->n_refine;
    a.n_explore = b->n_explore;
    a.precision = b->precision;
    a.init_map = J.init_map;
    a.init_w = J.init_w;
    a.init_h = J.init_h;
    a.init_stride = J.init_stride;
    a.init_scale = J.init_scale;
    a.init_search = b->init_search;
    a.init_sigma = b->init_sigma;
    a.init_sigma_slope = b->init_sigma_slope;
    a.counter = nullptr;
    T.half_tx = ceil_div((R.width + 1) / 2, tx);
    T.half_ty = ceil_div(R.height, ty);
    T.half_t0 = half_t;
    half_t += T.half_tx * T.half_ty;
    T.pix_tx = ceil_div(R.width, 16);
    T.pix_ty = ceil_div(R.height, 16);
    T.pix_t0 = pix_t;
    pix_t += T.pix_tx * T.pix_ty;
    T.out_depth = b->out_depth.as<float>() + J.off;
    T.out_plane = b->out_plane.as<float>() + 3 * J.off;
    T.out_score = b->out_score.as<float>() + J.off;
    T.out_nghbr = b->out_nghbr.as<int>() + J.off;
  }
  b->half_total = half_t;
  b->pix_total = pix_t;
  b->jobtab.ensure(jt.size() * sizeof(pm::PmJob));
  HIP_CHECK(hipMemcpy(b->jobtab.p, jt.data(), jt.size() * sizeof(pm::PmJob),
                      hipMemcpyHostToDevice));
  if (!pyramid_schedule(b).empty()) {
    // every job starts from the half level's map, sampled bilinearly with
    // no spread (estimator: run_half); a caller's map seeds the coarsest
    // level instead
    ensure_half(b, count);
    const pm_batch& h = *b->half;
    b->half_map.ensure(h.total_px * 4 * sizeof(float));
    for (int k = 0; k < count; ++k) {
      const PoolView& HR = *h.views[h.jobs[k].ref];
      pm::PmArgs& a = jt[k].a;
      a.init_map = b->half_map.as<float>() + 4 * h.jobs[k].off;
      a.init_w = HR.width;
      a.init_h = HR.height;
      a.init_stride = 4;
      a.init_scale = 2;
      a.init_search = 1;
      a.init_sigma = 0.0f;
      a.init_sigma_slope = 0.0f;
    }
    b->jobtab_pyr.ensure(jt.size() * sizeof(pm::PmJob));
    HIP_CHECK(hipMemcpy(b->jobtab_pyr.p, jt.data(), jt.size() * sizeof(pm::PmJob),
                        hipMemcpyHostToDevice));
  }
  b->prepared_jobs = count;
}

// Enqueues one whole run on b->stream.  Only kernel launches (no memsets or
// host work), so that it can be captured in a graph.  half_ev, if given,
// receives 4 * iterations events: one before and one after each red/black
// launch.
void enqueue(pm_batch* b, int mode, hipEvent_t* half_ev) {
  const bool sample = mode == PM_MODE_PATCHMATCH_SAMPLE;
  const int nj = b->prepared_jobs;
  const bool pyramid = !pyramid_schedule(b).empty();
  const pm::PmJob* jt = (pyramid ? b->jobtab_pyr : b->jobtab).as<pm::PmJob>();
  if (pyramid) {
    // the coarser level in the same mode (sample mode's levels sample too)
    pm_batch* h = b->half.get();
    enqueue(h, mode, nullptr);
    HIP_CHECK(pm::launch_state_to_map_batch(h->jobtab.as<pm::PmJob>(), h->prepared_jobs,
                                            h->pix_total, h->state.as<float>(),
                                            b->post_process, b->half_map.as<float>(),
                                            b->stream));
  }
  unsigned* q = b->queue.as<unsigned>();
  unsigned long long* cnt = b->counter.as<unsigned long long>();
  const int nq = 1 + 16 * b->iterations;
  HIP_CHECK(pm::launch_fill_u32(q, nq, 0u, b->stream));
  HIP_CHECK(pm::launch_fill_u32(reinterpret_cast<unsigned*>(cnt), 4, 0u, b->stream));
  // every pixel starts at the zero state (border, masked and low-variance
  // pixels keep it)
  HIP_CHECK(pm::launch_fill_u32(b->state.as<unsigned>(), 4 * b->total_px, 0u, b->stream));
  HIP_CHECK(pm::launch_fill_u32(b->score.as<unsigned>(), b->total_px, 0u, b->stream));
  HIP_CHECK(pm::launch_fill_u32(b->nghbr.as<unsigned>(), b->total_px, 0u, b->stream));
  // FP32: the random planes of the initialisation and the first iteration
  // are scored from the u8 quads (pm_types.h, kPrecFp32Quads); seeded runs
  // (a pyramid level's map, or any job's own map) start from mostly right
  // planes, which the coefficients score faster, as the estimator does
  bool seeded = pyramid;
  for (int k = 0; k < nj && !seeded; ++k) seeded = b->jobs[k].init_map != nullptr;
  const int first_prec =
      b->precision == PM_PRECISION_FP32 && !seeded ? pm::kPrecFp32Quads : b->precision;
  HIP_CHECK(pm::launch_pm_init_batch(jt, nj, b->pix_total, b->patch_size,
                                     first_prec, sample, q, cnt, b->stream));
  int l = 0;
  for (int it = 0; it < b->iterations; ++it) {
    for (int colour = 0; colour < 2; ++colour) {
      if (half_ev) HIP_CHECK(hipEventRecord(half_ev[2 * l], b->stream));
      HIP_CHECK(pm::launch_pm_half_batch(jt, nj, b->half_total, b->patch_size,
                                         it == 0 ? first_prec : b->precision, sample,
                                         b->tile, colour, it,
                                         q + 1 + 8 * l, cnt + 1, b->stream));
      if (half_ev) HIP_CHECK(hipEventRecord(half_ev[2 * l + 1], b->stream));
      ++l;
    }
    // median-consistent re-selection, as the single-view estimator does it
    // (after every iteration but the first, and after the last)
    if (b->consistency && (it >= 1 || it + 1 == b->iterations)) {
      const bool last = it + 1 == b->iterations;
      const pm::ConsistOut res{b->out_plane.as<float>(), b->out_depth.as<float>(),
                               b->out_score.as<float>(), b->out_nghbr.as<int>(),
                               pm::kConsistMinScore, last ? pm::kConsistFinalSupport : 0};
      const pm::ConsistItems items{
          b->item_count.as<unsigned long long>(), b->item_cap, b->item_entry.as<unsigned>(),
          b->item_win.as<unsigned char>(), b->item_score.as<float>(), b->item_nghbr.as<int>(),
          b->item_depth.as<float>(), b->item_first.as<unsigned>(), b->item_todo.as<unsigned>()};
      HIP_CHECK(pm::launch_pm_consist_batch(
          jt, nj, b->pix_total, b->patch_size, b->precision, sample,
          b->cons_count.as<unsigned>(), b->cons_idx.as<unsigned>(), b->cons_med.as<float>(),
          b->cons_job.as<int>(), res, (unsigned)it, items, cnt, b->stream));
    }
  }
  HIP_CHECK(pm::launch_pack_batch(jt, nj, b->pix_total, b->post_process, b->stream));
}

int run(pm_batch* b, int mode) {
  if (mode != PM_MODE_PATCHMATCH && mode != PM_MODE_PATCHMATCH_SAMPLE)
    return fail(PM_ERR_INVALID, "batched runs support PM_MODE_PATCHMATCH and "
                                "PM_MODE_PATCHMATCH_SAMPLE");
  int rc = validate(b);
  if (rc) return rc;
  DeviceGuard dg(b->device);
  if (b->prepared_jobs != (int)b->jobs.size()) {
    b->drop_graph();
    prepare(b, (int)b->jobs.size());
  }
  if (b->total_px > 0x7fffffffu)
    return fail(PM_ERR_INVALID, "batch too large: at most 2147483647 reference pixels");
  b->have_result = b->have_clean = false;
  HIP_CHECK(hipEventRecord(b->ev[0], b->stream));
  if (b->use_graph) {
    if (!b->exec || b->graph_mode != mode) {
      b->drop_graph();
      // one eager run first: it is the result of this call, and it resolves
      // the persistent grid sizes outside the capture
      enqueue(b, mode, nullptr);
      hipGraph_t g = nullptr;
      HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(b, mode, nullptr);
      } catch (...) {
        hipStreamEndCapture(b->stream, &g);
        if (g) hipGraphDestroy(g);
        throw;
      }
      HIP_CHECK(hipStreamEndCapture(b->stream, &g));
      hipError_t e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      HIP_CHECK(e);
      b->graph_mode = mode;
    } else {
      HIP_CHECK(hipGraphLaunch(b->exec, b->stream));
    }
  } else {
    enqueue(b, mode, nullptr);
  }
  HIP_CHECK(hipEventRecord(b->ev[1], b->stream));
  HIP_CHECK(hipEventSynchronize(b->ev[1]));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  unsigned long long cnt[2] = {0, 0};
  HIP_CHECK(hipMemcpy(cnt, b->counter.p, sizeof(cnt), hipMemcpyDeviceToHost));
  b->stats = pm_stats{};
  b->stats.scorings = (int64_t)(cnt[0] + cnt[1]);
  b->stats.dominant_scorings = (int64_t)cnt[1];
  b->stats.patch_size = b->patch_size;
  b->stats.width = 0;
  b->stats.height = 0;
  b->stats.launches = 2 * b->iterations;
  // the coarser levels ran in the same sequence: their launches of the same
  // kernels count too (the device time above already holds them)
  for (const pm_batch* c = b; !pyramid_schedule(c).empty(); c = c->half.get()) {
    const pm_batch* h = c->half.get();
    unsigned long long hc[2] = {0, 0};
    HIP_CHECK(hipMemcpy(hc, h->counter.p, sizeof(hc), hipMemcpyDeviceToHost));
    b->stats.scorings += (int64_t)(hc[0] + hc[1]);
    b->stats.dominant_scorings += (int64_t)hc[1];
    b->stats.launches += 2 * h->iterations;
  }
  b->stats.dominant_ms = ms;
  b->stats.total_ms = ms;
  b->have_result = true;
  return PM_OK;
}

// Times the second red/black iteration (converged planes, the sample source
// most of a run reads) of a 2-iteration run over the first jobs (up to about
// 8 Mpx) for one (patch, tile) candidate; milliseconds per Mpx.
float probe(pm_batch* b, int mode) {
  size_t px = 0;
  int count = 0;
  while (count < (int)b->jobs.size() && (count == 0 || px < (size_t)8 << 20)) {
    const PoolView& R = *b->views[b->jobs[count].ref];
    px += (size_t)R.width * R.height;
    ++count;
  }
  const int iters = b->iterations;
  const int pyr = b->pyramid_iters;  // the tile is timed on plain runs
  const std::vector<int> lv = b->level_iters;
  b->iterations = 2;
  b->pyramid_iters = 0;
  b->level_iters.clear();
  std::vector<hipEvent_t> ev(8, nullptr);
  float ms = 0;
  try {
    for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
    prepare(b, count);
    enqueue(b, mode, ev.data());
    HIP_CHECK(hipEventSynchronize(ev[7]));
    for (int l = 2; l < 4; ++l) {
      float t = 0;
      HIP_CHECK(hipEventElapsedTime(&t, ev[2 * l], ev[2 * l + 1]));
      ms += t;
    }
  } catch (...) {
    for (auto& e : ev)
      if (e) hipEventDestroy(e);
    b->iterations = iters;
    b->pyramid_iters = pyr;
    b->level_iters = lv;
    b->prepared_jobs = -1;
    throw;
  }
  for (auto& e : ev) hipEventDestroy(e);
  b->iterations = iters;
  b->pyramid_iters = pyr;
  b->level_iters = lv;
  b->prepared_jobs = -1;
  return ms / (float)(px / 1e6);
}

}  // namespace

extern "C" {

int pm_batch_create(int device, pm_batch** out) {
  if (!out) return fail(PM_ERR_INVALID, "null output handle");
  *out = nullptr;
  int rc = check_device(device);
  if (rc) return rc;
  return guarded([&] {
    std::unique_ptr<pm_batch> b(new pm_batch());
    b->device = device;
    DeviceGuard dg(device);
    HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (auto& e : b->ev) HIP_CHECK(hipEventCreate(&e));
    *out = b.release();
    return PM_OK;
  });
}

int pm_batch_destroy(pm_batch* b) {
  if (!b) return PM_OK;
  return guarded([&] {
    DeviceGuard dg(b->device);
    delete b;
    return PM_OK;
  });
}

int pm_batch_add_view(pm_batch* b, const double* K, const double* R,
                      const double* t, const uint8_t* image, const uint8_t* mask,
                      int width, int height, int stride, int* view_id) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (!K || !R || !t || !image || !mask)
    return fail(PM_ERR_INVALID, "null K, R, t, image or mask");
  if (width < 2 || height < 2) return fail(PM_ERR_INVALID, "image must be at least 2x2");
  if (stride < width) return fail(PM_ERR_INVALID, "stride smaller than width");
  if (!all_finite(K, 9) || !all_finite(R, 9) || !all_finite(t, 3))
    return fail(PM_ERR_INVALID, "K, R and t must be finite");
  return guarded([&] {
    std::unique_ptr<PoolView> v(new PoolView());
    memcpy(v->K, K, sizeof(v->K));
    memcpy(v->R, R, sizeof(v->R));
    memcpy(v->t, t, sizeof(v->t));
    if (!inv3(v->K, v->Kinv)) return fail(PM_ERR_INVALID, "K is singular");
    v->width = width;
    v->height = height;
    DeviceGuard dg(b->device);
    const size_t np = (size_t)width * height;
    v->img.ensure(np);
    v->mask.ensure(np);
    upload_image(v->img.p, image, width, height, stride, b->stream);
    upload_image(v->mask.p, mask, width, height, stride, b->stream);
    v->quad.ensure((size_t)(width + 1) * (height + 1) * sizeof(uint32_t));
    v->quadc.ensure((size_t)(width + 1) * (height + 1) * 8);
    HIP_CHECK(pm::launch_build_quad(v->img.as<uint8_t>(), v->quad.as<uint32_t>(),
                                    v->quadc.as<uint32_t>(), width, height, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    if (view_id) *view_id = (int)b->views.size();
    b->views.push_back(std::move(v));
    return PM_OK;
  });
}

int pm_batch_add_job(pm_batch* b, int ref_view, const int* neighbours,
                     int n_neighbours, double min_depth, double max_depth,
                     int* job_id) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  const int nv = (int)b->views.size();
  if (ref_view < 0 || ref_view >= nv)
    return fail(PM_ERR_RANGE, "reference view " + std::to_string(ref_view) +
                                  " out of range [0, " + std::to_string(nv) + ")");
  if (n_neighbours < 1 || !neighbours)
    return fail(PM_ERR_INVALID, "a job needs at least one neighbour view");
  if (n_neighbours > pm::kMaxViews - 1)
    return fail(PM_ERR_INVALID, "at most " + std::to_string(pm::kMaxViews - 1) +
                                    " neighbour views per job");
  for (int k = 0; k < n_neighbours; ++k)
    if (neighbours[k] < 0 || neighbours[k] >= nv)
      return fail(PM_ERR_RANGE, "neighbour view " + std::to_string(neighbours[k]) +
                                    " out of range [0, " + std::to_string(nv) + ")");
  if (!(min_depth > 0) || !(max_depth > min_depth) || !std::isfinite(max_depth))
    return fail(PM_ERR_INVALID, "depth range must satisfy 0 < min_depth < max_depth");
  BatchJob J;
  J.ref = ref_view;
  J.nb.assign(neighbours, neighbours + n_neighbours);
  J.min_depth = min_depth;
  J.max_depth = max_depth;
  if (job_id) *job_id = (int)b->jobs.size();
  b->jobs.push_back(std::move(J));
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_pyramid_consistency(pm_batch* b, int enable) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  b->pyramid_consistency = enable ? 1 : 0;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_pyramid_levels(pm_batch* b, int levels) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (levels < 2 || levels > 8) return fail(PM_ERR_INVALID, "pyramid levels must be in [2, 8]");
  b->pyramid_levels = levels;
  b->level_iters.clear();
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_pyramid_schedule(pm_batch* b, const int32_t* iterations, int n) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (n < 1 || n > 7) return fail(PM_ERR_INVALID, "a schedule has 1 to 7 coarser levels");
  if (!iterations) return fail(PM_ERR_INVALID, "null schedule");
  for (int k = 0; k < n; ++k)
    if (iterations[k] < 1 || iterations[k] > 1000)
      return fail(PM_ERR_INVALID, "level iterations must be in [1, 1000]");
  b->level_iters.assign(iterations, iterations + n);
  b->pyramid_levels = n + 1;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_pyramid(pm_batch* b, int half_iterations) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (half_iterations < 0 || half_iterations > 1000)
    return fail(PM_ERR_INVALID, "half-level iterations must be in [0, 1000]");
  b->pyramid_iters = half_iterations;
  b->level_iters.clear();
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_patchmatch_iterations(pm_batch* b, int n) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (n < 0 || n > 1000) return fail(PM_ERR_INVALID, "iterations must be in [0, 1000]");
  b->iterations = n;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_patch_size(pm_batch* b, int size) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (size < 3 || size > 11 || size % 2 == 0)
    return fail(PM_ERR_INVALID, "patch size must be odd and in [3, 11]");
  b->patch_size = size;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_min_patch_sd(pm_batch* b, float sd) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  b->min_patch_variance = sd * sd;
  b->invalidate();
  return PM_OK;
}

int pm_batch_initializer_input(pm_batch* b, int job, int scale, int nplanes, void* out,
                               int out_w, int out_h) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (job < 0 || job >= (int)b->jobs.size()) return fail(PM_ERR_RANGE, "job index out of range");
  if (!out) return fail(PM_ERR_INVALID, "null output");
  if (scale < 1 || scale > 16) return fail(PM_ERR_INVALID, "scale must be in [1, 16]");
  if (nplanes < 2 || nplanes > 39) return fail(PM_ERR_INVALID, "nplanes must be in [2, 39]");
  if (b->patch_size != 5 && b->patch_size != 7)
    return fail(PM_ERR_INVALID, "the initialiser input needs patch size 5 or 7");
  const BatchJob& J = b->jobs[job];
  const PoolView& R0 = *b->views[J.ref];
  const int wl = R0.width / scale, hl = R0.height / scale;
  if (wl < b->patch_size || hl < b->patch_size)
    return fail(PM_ERR_INVALID, "reference view too small for this scale");
  if (out_w < wl || out_h < hl)
    return fail(PM_ERR_INVALID, "output grid smaller than the coarse view (" +
                                    std::to_string(wl) + "x" + std::to_string(hl) + ")");
  return guarded([&] {
    DeviceGuard dg(b->device);
    if (b->coarse_scale != scale) {
      b->coarse.clear();
      b->coarse_scale = scale;
    }
    // coarse copies of the pool views, built once per scale (the estimator's
    // build_coarse, per pool view)
    for (size_t v = b->coarse.size(); v < b->views.size(); ++v) {
      const PoolView& V = *b->views[v];
      std::unique_ptr<CoarseView> c(new CoarseView());
      memcpy(c->K, V.K, sizeof(c->K));
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
      c->img.ensure((size_t)c->width * c->height);
      HIP_CHECK(pm::launch_downsample_u8(V.img.as<uint8_t>(), V.width, V.height, scale, 0,
                                         c->img.as<uint8_t>(), c->width, c->height, b->stream));
      c->quad16.ensure((size_t)(c->width + 1) * (c->height + 1) * 8);
      HIP_CHECK(pm::launch_build_quad16(c->img.as<uint8_t>(), c->quad16.as<uint32_t>(),
                                        c->width, c->height, PM_PRECISION_FP16, b->stream));
      b->coarse.push_back(std::move(c));
    }
    std::vector<int> vs(1, J.ref);
    vs.insert(vs.end(), J.nb.begin(), J.nb.end());
    const int n = (int)vs.size();
    const CoarseView& c0 = *b->coarse[J.ref];
    std::vector<pm::PmView> tab(n);
    for (int k = 0; k < n; ++k) {
      const PoolView& V = *b->views[vs[k]];
      const CoarseView& C = *b->coarse[vs[k]];
      double Q[9], a[3];
      relative_pose(R0.R, R0.t, V.R, V.t, Q, a);
      fill_view(tab[k], nullptr, C.width, C.height, C.K, Q, a, c0.Kinv);
      tab[k].quad16 = C.quad16.as<uint32_t>();
    }
    b->coarse_viewtab.ensure(sizeof(pm::PmView) * n);
    HIP_CHECK(hipMemcpyAsync(b->coarse_viewtab.p, tab.data(), sizeof(pm::PmView) * n,
                             hipMemcpyHostToDevice, b->stream));
    const int hpz = (b->patch_size - 1) / 2;
    if (b->coarse_wlut_patch != b->patch_size) {
      std::vector<float> t = bilateral_table(hpz);
      b->coarse_wlut.ensure(t.size() * sizeof(float));
      HIP_CHECK(hipMemcpyAsync(b->coarse_wlut.p, t.data(), t.size() * sizeof(float),
                               hipMemcpyHostToDevice, b->stream));
      b->coarse_wlut_patch = b->patch_size;
    }
    pm::PmArgs a;
    memset(&a, 0, sizeof(a));
    a.W = c0.width;
    a.H = c0.height;
    a.nviews = n;
    a.hpz = hpz;
    a.views = b->coarse_viewtab.as<pm::PmView>();
    a.ref_img = c0.img.as<uint8_t>();
    a.wlut = b->coarse_wlut.as<float>();
    for (int k = 0; k < 9; ++k) {
      a.Kinv0f[k] = (float)c0.Kinv[k];
      a.Kinv0[k] = c0.Kinv[k];
    }
    a.precision = PM_PRECISION_FP16;
    const float inv_min = (float)(1.0 / J.min_depth);
    const float inv_step = (float)((1.0 / J.max_depth - 1.0 / J.min_depth) / (nplanes - 1));
    HIP_CHECK(pm::launch_init_input(a, nplanes, inv_min, inv_step, out, out_w, out_h, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    return PM_OK;
  });
}

int pm_batch_set_init_map(pm_batch* b, int job, const float* map, int w, int h, int stride,
                           int scale) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (job < 0 || job >= (int)b->jobs.size()) return fail(PM_ERR_RANGE, "job index out of range");
  if (map && (w < 1 || h < 1 || stride < 4 || scale < 1))
    return fail(PM_ERR_INVALID, "init map needs w, h >= 1, stride >= 4, scale >= 1");
  BatchJob& J = b->jobs[job];
  J.init_map = map;
  J.init_w = map ? w : 0;
  J.init_h = map ? h : 0;
  J.init_stride = map ? stride : 0;
  J.init_scale = map ? scale : 1;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_init_spread(pm_batch* b, float sigma_log_depth, float sigma_slope) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (!(sigma_log_depth >= 0) || !(sigma_slope >= 0) || sigma_log_depth > 1 || sigma_slope > 10)
    return fail(PM_ERR_INVALID, "spreads must be in [0, 1] and [0, 10]");
  b->init_sigma = sigma_log_depth;
  b->init_sigma_slope = sigma_slope;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_init_search(pm_batch* b, int n) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (n < 1 || n > 15) return fail(PM_ERR_INVALID, "init search must be in [1, 15]");
  b->init_search = n;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_seed(pm_batch* b, uint64_t seed) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  b->seed = seed;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_exploration(pm_batch* b, int n) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (n < 0 || n > 8) return fail(PM_ERR_INVALID, "exploration samples must be in [0, 8]");
  b->n_explore = n;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_consistency(pm_batch* b, int enable) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  b->consistency = enable ? 1 : 0;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_post_process(pm_batch* b, int enable) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  b->post_process = enable ? 1 : 0;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_tile(pm_batch* b, int tile) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (tile < PM_TILE_8x8 || tile > PM_TILE_32x8)
    return fail(PM_ERR_INVALID, "unknown tile " + std::to_string(tile));
  b->tile = tile;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_precision(pm_batch* b, int precision) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (precision < PM_PRECISION_FP32 || precision > PM_PRECISION_BF16)
    return fail(PM_ERR_INVALID, "unknown precision " + std::to_string(precision));
  b->precision = precision;
  b->invalidate();
  return PM_OK;
}

int pm_batch_set_graph(pm_batch* b, int enable) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  b->use_graph = enable ? 1 : 0;
  b->drop_graph();
  return PM_OK;
}

int pm_batch_autotune(pm_batch* b, int mode, const int* patch_sizes, int n,
                      int* patch_out, int* tile_out) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (mode != PM_MODE_PATCHMATCH && mode != PM_MODE_PATCHMATCH_SAMPLE)
    return fail(PM_ERR_INVALID, "autotune supports PM_MODE_PATCHMATCH and "
                                "PM_MODE_PATCHMATCH_SAMPLE");
  std::vector<int> patches;
  if (n > 0 && !patch_sizes) return fail(PM_ERR_INVALID, "null patch size list");
  for (int k = 0; k < n; ++k) {
    const int p = patch_sizes[k];
    if (p < 3 || p > 11 || p % 2 == 0)
      return fail(PM_ERR_INVALID, "patch sizes must be odd and in [3, 11]");
    patches.push_back(p);
  }
  if (patches.empty()) patches.push_back(b->patch_size);
  int rc = validate(b);
  if (rc) return rc;
  return guarded([&] {
    DeviceGuard dg(b->device);
    b->invalidate();
    const int keep_patch = b->patch_size, keep_tile = b->tile;
    float best = 0;
    int best_p = keep_patch, best_t = keep_tile;
    bool first = true;
    try {
      for (int p : patches) {
        for (int t = PM_TILE_8x8; t <= PM_TILE_32x8; ++t) {
          b->patch_size = p;
          b->tile = t;
          // smaller views than the patch cannot run at this size
          if (validate(b)) continue;
          const float ms = probe(b, mode);
          if (first || ms < best) {
            best = ms;
            best_p = p;
            best_t = t;
            first = false;
          }
        }
      }
    } catch (...) {
      b->patch_size = keep_patch;
      b->tile = keep_tile;
      throw;
    }
    if (first) return fail(PM_ERR_INVALID, "no candidate patch size fits every job");
    b->patch_size = best_p;
    b->tile = best_t;
    b->invalidate();
    if (patch_out) *patch_out = best_p;
    if (tile_out) *tile_out = best_t;
    return PM_OK;
  });
}

int pm_batch_run(pm_batch* b, int mode) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  return guarded([&] { return run(b, mode); });
}

int pm_batch_download(pm_batch* b, int job, float* depth, float* plane,
                      float* score, int32_t* nghbr) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (!b->have_result) return fail(PM_ERR_INVALID, "no computed maps to download");
  if (job < 0 || job >= (int)b->jobs.size())
    return fail(PM_ERR_RANGE, "job " + std::to_string(job) + " out of range");
  if (!depth || !plane || !score || !nghbr) return fail(PM_ERR_INVALID, "null output buffer");
  return guarded([&] {
    DeviceGuard dg(b->device);
    const BatchJob& J = b->jobs[job];
    const PoolView& R = *b->views[J.ref];
    const size_t np = (size_t)R.width * R.height;
    HIP_CHECK(hipMemcpyAsync(depth, b->out_depth.as<float>() + J.off, np * sizeof(float),
                             hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipMemcpyAsync(plane, b->out_plane.as<float>() + 3 * J.off,
                             3 * np * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipMemcpyAsync(score, b->out_score.as<float>() + J.off, np * sizeof(float),
                             hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipMemcpyAsync(nghbr, b->out_nghbr.as<int>() + J.off, np * sizeof(int32_t),
                             hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    return PM_OK;
  });
}

// The cluster estimator's result of one job (DepthmapClusterEstimator::Run,
// depthmap_bind.h:53-73): depth, unit normal, cost, confidence.
int pm_batch_download_cluster(pm_batch* b, int job, float* depth, float* normal, float* cost,
                              float* confidence) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (!b->have_result) return fail(PM_ERR_INVALID, "no computed maps to download");
  if (job < 0 || job >= (int)b->jobs.size())
    return fail(PM_ERR_RANGE, "job " + std::to_string(job) + " out of range");
  if (!depth || !normal || !cost || !confidence) return fail(PM_ERR_INVALID, "null output buffer");
  return guarded([&] {
    DeviceGuard dg(b->device);
    const BatchJob& J = b->jobs[job];
    const PoolView& R = *b->views[J.ref];
    const size_t np = (size_t)R.width * R.height;
    b->cluster_out.ensure(np * 6 * sizeof(float));
    float* nrm = b->cluster_out.as<float>();
    float* cst = nrm + 3 * np;
    float* cnf = cst + np;
    float* dep = cnf + np;
    HIP_CHECK(pm::launch_cluster_out(b->out_depth.as<float>() + J.off,
                                     b->out_plane.as<float>() + 3 * J.off,
                                     b->out_score.as<float>() + J.off, (int)np,
                                     (float)J.min_depth, (float)J.max_depth, dep, nrm, cst, cnf,
                                     b->stream));
    HIP_CHECK(hipMemcpyAsync(depth, dep, np * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipMemcpyAsync(normal, nrm, 3 * np * sizeof(float), hipMemcpyDeviceToHost,
                             b->stream));
    HIP_CHECK(hipMemcpyAsync(cost, cst, np * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipMemcpyAsync(confidence, cnf, np * sizeof(float), hipMemcpyDeviceToHost,
                             b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    return PM_OK;
  });
}

// DepthmapCleaner::Clean (depthmap.cc:524-555) of every job's depthmap
// against the depthmaps of its neighbour views that are themselves the
// reference view of a job (their packed, post-processed depth).
int pm_batch_clean(pm_batch* b, float same_depth_threshold, int min_consistent_views) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (!b->have_result) return fail(PM_ERR_INVALID, "run the batch before cleaning");
  return guarded([&] {
    DeviceGuard dg(b->device);
    const int nj = (int)b->jobs.size();
    std::vector<int> job_of(b->views.size(), -1);
    for (int k = 0; k < nj; ++k)
      if (job_of[b->jobs[k].ref] < 0) job_of[b->jobs[k].ref] = k;
    std::vector<pm::ClView> cv;
    std::vector<pm::ClJob> cj(nj);
    int pix_t = 0;
    for (int k = 0; k < nj; ++k) {
      const BatchJob& J = b->jobs[k];
      pm::ClJob& T = cj[k];
      memset(&T, 0, sizeof(T));
      T.view0 = (int)cv.size();
      auto push = [&](int view, int job) {
        const PoolView& V = *b->views[view];
        pm::ClView c;
        memset(&c, 0, sizeof(c));
        c.depth = b->out_depth.as<float>() + b->jobs[job].off;
        c.width = V.width;
        c.height = V.height;
        memcpy(c.K, V.K, sizeof(c.K));
        memcpy(c.R, V.R, sizeof(c.R));
        memcpy(c.t, V.t, sizeof(c.t));
        cv.push_back(c);
      };
      push(J.ref, k);
      for (int v : J.nb)
        if (job_of[v] >= 0) push(v, job_of[v]);
      T.nviews = (int)cv.size() - T.view0;
      const PoolView& R = *b->views[J.ref];
      memcpy(T.ref.Kinv, R.Kinv, sizeof(T.ref.Kinv));
      mat3_transpose(R.R, T.ref.Rt);
      T.pix_tx = ceil_div(R.width, 16);
      T.pix_t0 = pix_t;
      pix_t += T.pix_tx * ceil_div(R.height, 16);
      T.out = nullptr;  // set below once cl_out is sized
    }
    b->cl_out.ensure(b->total_px * sizeof(float));
    for (int k = 0; k < nj; ++k) cj[k].out = b->cl_out.as<float>() + b->jobs[k].off;
    b->cl_views.ensure(cv.size() * sizeof(pm::ClView));
    b->cl_jobs.ensure(cj.size() * sizeof(pm::ClJob));
    HIP_CHECK(hipMemcpyAsync(b->cl_views.p, cv.data(), cv.size() * sizeof(pm::ClView),
                             hipMemcpyHostToDevice, b->stream));
    HIP_CHECK(hipMemcpyAsync(b->cl_jobs.p, cj.data(), cj.size() * sizeof(pm::ClJob),
                             hipMemcpyHostToDevice, b->stream));
    HIP_CHECK(hipEventRecord(b->ev[2], b->stream));
    HIP_CHECK(pm::launch_clean_batch(b->cl_jobs.as<pm::ClJob>(), nj, pix_t,
                                     b->cl_views.as<pm::ClView>(), same_depth_threshold,
                                     min_consistent_views, b->stream));
    HIP_CHECK(hipEventRecord(b->ev[3], b->stream));
    HIP_CHECK(hipEventSynchronize(b->ev[3]));
    HIP_CHECK(hipEventElapsedTime(&b->clean_ms, b->ev[2], b->ev[3]));
    b->have_clean = true;
    return PM_OK;
  });
}

int pm_batch_download_clean(pm_batch* b, int job, float* depth) {
  if (!b) return fail(PM_ERR_INVALID, "null batch");
  if (!b->have_clean) return fail(PM_ERR_INVALID, "no cleaned maps to download");
  if (job < 0 || job >= (int)b->jobs.size())
    return fail(PM_ERR_RANGE, "job " + std::to_string(job) + " out of range");
  if (!depth) return fail(PM_ERR_INVALID, "null output buffer");
  return guarded([&] {
    DeviceGuard dg(b->device);
    const BatchJob& J = b->jobs[job];
    const PoolView& R = *b->views[J.ref];
    HIP_CHECK(hipMemcpyAsync(depth, b->cl_out.as<float>() + J.off,
                             (size_t)R.width * R.height * sizeof(float),
                             hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    return PM_OK;
  });
}

int pm_batch_stats(pm_batch* b, pm_stats* out) {
  if (!b || !out) return fail(PM_ERR_INVALID, "null argument");
  *out = b->stats;
  return PM_OK;
}

int pm_batch_clean_ms(pm_batch* b, float* ms) {
  if (!b || !ms) return fail(PM_ERR_INVALID, "null argument");
  *ms = b->clean_ms;
  return PM_OK;
}

}  // extern "C"
