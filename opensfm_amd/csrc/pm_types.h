// Plain-data structures shared by the host code and the kernels.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace pm {

constexpr int kMaxViews = 64;  // reference view + up to 63 neighbours

// Sample source of the FP32 PatchMatch launches that score random planes
// (the initialisation and the first red/black iteration of an unseeded run):
// the 4-byte u8 quads instead of the 8-byte FP16 coefficients.  Random
// planes scatter the gathers over whole epipolar segments, so those
// launches are bound by the cache lines they fetch; once planes have
// converged the gathers stay within a few lines and the coefficients'
// shorter blend wins.
constexpr int kPrecFp32Quads = 3;

// Median-consistent re-selection: minimum NCC of a re-selected plane, and
// the minimum number of the 24 other pixels of the 5x5 window whose depth
// lies within 1 % of the re-selected depth in the pass after the last
// iteration (measured on the 1600x1200 and 320x240 scenes: below 9 of 24
// fewer than one re-selection in ten is within 1 % of the true depth).
constexpr float kConsistMinScore = 0.5f;
constexpr int kConsistFinalSupport = 10;

// Seed offset of the half-resolution level: the level of an estimator with
// seed s runs with seed s + kHalfLevelSeed, so that batched job k (seed
// s + k) and an estimator with seed s + k draw the same numbers at both
// levels.
constexpr uint64_t kHalfLevelSeed = 0x9E3779B97F4A7C15ull;
// Single-view red/black launches with fewer pixels of one colour than this
// score each pixel on kLanesPerPixelSmall lanes, one share of the views each
// (k_pm_half_group).  Measured with 8 neighbours, patch 7: 15 k pixels
// (200 x 150) 0.15 ms per launch against 0.42 on one lane; 60 k pixels
// (400 x 300) 0.57 ms against 0.40, since every lane of a group repeats the
// candidate loop and the reference patch once the chip is full (2 or 4 lanes
// per pixel were no faster there).
constexpr int kLanesPerPixelSmall = 8;
constexpr size_t kGroupBelowPixels = 32768;

// One view as the kernels see it.  For neighbour view v (v >= 1) the plane-
// induced homography of plane v_pi (in the reference camera frame) is
//   H_v = K_v (Q_v + a_v v_pi^T) K_0^-1 = M_v + b_v (K_0^-T v_pi)^T
// with Q_v = R_v R_0^T, a_v = Q_v t_0 - t_v (depthmap.cc:153-155),
// M_v = K_v Q_v K_0^-1 and b_v = K_v a_v composed in FP64 on the host.
struct PmView {
  const uint32_t* quad;  // zero-bordered 2x2 texel quads, pitch width + 1
  const uint32_t* quad16;  // the same quads as 4 FP16 or BF16 values (8 B
                           // per texel), for reduced-precision PatchMatch
  const uint32_t* quadc;   // bilinear coefficients of the quads as 4 FP16
                           // values (8 B per texel), for FP32 PatchMatch
  int width, height;
  float M[9];
  float b[3];
  double K[9], Q[9], a[3];  // FP64 inputs of the exact homography
};

// Arguments common to the estimator kernels (passed by value).
struct PmArgs {
  int W, H;            // reference view size
  int nviews;          // views including the reference
  int hpz;             // (patch_size - 1) / 2
  const PmView* views;  // device table [nviews]
  const uint8_t* ref_img;
  const uint8_t* ref_mask;
  const float* wlut;   // bilateral weight table, 256 x (2 hpz^2 + 1)
  float* state;        // 4 floats per pixel: plane v (3) and depth
  float* score;
  int* nghbr;
  float Kinv0f[9];     // K_0^-1 in FP32 (fast path)
  double Kinv0[9];     // K_0^-1 in FP64 (exact path)
  float log_min, log_max;
  float min_patch_var;
  uint32_t seed_lo, seed_hi;
  int n_far;           // far propagation probes per direction
  int n_refine;        // random refinements per pixel update (6-step schedule)
  int n_explore;       // exploration samples per pixel update (0: none)
  int precision;       // PatchMatch bilinear samples: 0 FP32, 1 FP16, 2 BF16,
                       // kPrecFp32Quads (FP32 from the u8 quads)
  // Learned initialisation map (nullptr: random initialisation only):
  // init_h x init_w cells of init_stride floats (t, sx, sy, c logit, ...),
  // cell (gx, gy) centred on pixel ((gx + 0.5) init_scale - 0.5, ...).
  const float* init_map;
  int init_w, init_h, init_stride, init_scale;
  int init_search;     // depth hypotheses tried at a seeded pixel (>= 1)
  float init_sigma, init_sigma_slope;  // per-pixel spread around the map
  unsigned long long* counter;  // scorings executed (nullptr: not counted)
};

// One median-consistent re-selection pass: the new states of the listed
// pixels, one entry per listed pixel (plane, 3 floats; depth; score; view),
// and its acceptance rule -- a re-selected plane is kept when its NCC is at
// least min_score and its depth lies within 1 % of the depths of at least
// min_support of the 24 other pixels of the 5x5 window.
struct ConsistOut {
  float* plane;
  float* depth;
  float* score;
  int* nghbr;
  float min_score;
  int min_support;
};

// The window phase of a re-selection pass as a flat list of work items, one
// per (listed pixel, window position) to score, so that the lanes of a wave
// score the candidates of several pixels at once (k_consist_plan allocates
// them, k_consist_score scores them, k_consist_select reads the scores).
// Per listed entry e: todo[e] holds the window positions it scores (bit w)
// and first[e] its first item; bit 31 of todo[e] marks an entry whose items
// did not fit in `cap` and which k_consist_select scores itself.  Items of
// such an entry below `cap` carry entry = 0xFFFFFFFF and are skipped.
struct ConsistItems {
  unsigned long long* count;  // items allocated (may exceed cap)
  unsigned cap;               // capacity of the item arrays
  unsigned* entry;            // listed entry of item k
  unsigned char* win;         // window position of item k
  float* score;               // the window plane's score over all views
  int* nghbr;                 // the view of that score
  float* depth;               // the plane's depth at the pixel
  unsigned* first;            // per listed entry
  unsigned* todo;             // per listed entry
};

// One reference view of a batch: its estimator arguments plus where its
// tiles sit in a launch's flat tile space.  Red/black tiles are TX x TY lanes
// over (columns of one colour) x rows; init/pack/clean tiles are 16 x 16
// pixels.
struct PmJob {
  PmArgs a;                      // a.counter unused (the batch counts per launch)
  int half_tx, half_ty, half_t0;  // red/black tiles across, down, first index
  int pix_tx, pix_ty, pix_t0;     // 16 x 16 pixel tiles across, down, first
  float* out_depth;               // packed outputs (H*W, H*W*3, H*W, H*W)
  float* out_plane;
  float* out_score;
  int* out_nghbr;
};

// Cleaner view (DepthmapCleaner::AddView, depthmap.cc:515-522).
struct ClView {
  const float* depth;
  int width, height;
  double K[9], R[9], t[3];
};

// Inverse pose of the view being cleaned: K_0^-1 and R_0^T.
struct ClRef {
  double Kinv[9], Rt[9];
};

// Camera of the carving cleaner (GPUDepthmapCleaner): FP32 row-major K, R
// and t, as the reference hands them to its kernel (opencl_utils.h:230-252).
struct GcCamera {
  float K[9], R[9], t[3];
};

// A view of the pruner (DepthmapPruner::AddView, depthmap.cc:563-575).
struct PrView {
  const float* depth;
  const float* plane;  // H x W x 3
  int width, height;
  double K[9], R[9], t[3];
};

// View 0 of the pruner: K^-1, R^T (FP64) and R^T in FP32 (the normals'
// rotation, a cv::Matx33f in the reference).
struct PrRef {
  double Kinv[9], Rt[9];
  float Rinv[9];
};

// One depthmap of a batched clean: views[view0] is the map being cleaned,
// views[view0 + 1 .. view0 + nviews - 1] the maps it is checked against.
struct ClJob {
  ClRef ref;
  int view0, nviews;
  int pix_t0, pix_tx;  // 16 x 16 pixel tiles: first index, tiles across
  float* out;
};

}  // namespace pm
