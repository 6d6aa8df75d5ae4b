// Device-side types and launchers of the depthmap fuser
// (dense::DepthmapFuser, opensfm/src/lib/dense/fuser.h:28-75).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fu {

constexpr int kMaxNeighbours = 64;  // one 64-bit mask of passing neighbours
constexpr int kMaxKnn = 32;         // outlier removal's k

// Float camera of a view as the reference casts it (fuser.cc:158-162):
// K, R, t rounded to float, K^-1 by the cofactor form, R^T.
struct Camera {
  float K[9], R[9], t[3], Kinv[9], Rinv[9];
};

struct View {
  const float* depth;          // H x W
  const float* normal;         // H x W x 3 (camera frame)
  const uint8_t* color;        // H x W x 3
  const uint8_t* mask;         // H x W, or null (no mask)
  uint8_t* fused;              // H x W, 1 once a fused point used the pixel
  unsigned long long* claim;   // H x W, round-tagged claims of the parallel greedy
  int width, height;
  Camera cam;
};

struct Params {
  int ref, n_nb, bm, min_consistent;
  float max_reproj_sq, max_depth_error, min_cos, behind;
  int nb[kMaxNeighbours];
};

// Fusion of reference view prm.ref, one round of the parallel greedy:
// claim, decide, compact.  counts[0] is the active count of this round,
// counts[1] receives the next one.
hipError_t launch_fuse_round(const View* views, const Params& prm, const int* active,
                             int* next_active, const int* count, int* next_count,
                             unsigned int round, unsigned long long* masks, uint8_t* st,
                             float* res, uint8_t* res_clr, uint8_t* ok, int grid,
                             hipStream_t s);
hipError_t launch_fuse_init(int W, int H, int bm, int* active, int* count, uint8_t* st,
                            uint8_t* ok, hipStream_t s);
// Kept-pixel counts per 256-pixel block of view W x H, then the scatter of
// the fused points in row-major order from offset `base`.
hipError_t launch_fuse_count(const uint8_t* ok, int npix, int* counts, hipStream_t s);
hipError_t launch_fuse_scatter(const uint8_t* ok, int npix, const long long* offsets,
                               const float* res, const uint8_t* res_clr, float* points,
                               float* normals, uint8_t* colors, hipStream_t s);

// Statistical outlier removal: mean distance of each point to its k
// nearest others, over a Morton-ordered multi-level cell hash.
struct Grid {
  float lo[3];
  float quantum;               // cell edge at level 0
  int level0, levels;          // first level searched, levels with tables
  unsigned long long* keys[22];   // per level: hash table of cell codes
  unsigned int* start[22];
  unsigned int* cnt[22];
  unsigned int mask[22];       // table size - 1
};
hipError_t launch_bbox(const float* points, long long n, float* lohi, hipStream_t s);
hipError_t launch_morton(const float* points, long long n, const float* lo, float quantum,
                         unsigned long long* codes, unsigned int* idx, hipStream_t s);
hipError_t launch_gather_sorted(const float* points, const unsigned int* idx, long long n,
                                float* sorted, hipStream_t s);
// The cell counts of all 22 levels in one pass (n_cells[22], zeroed).
hipError_t launch_level_count_all(const unsigned long long* codes, long long n,
                                  unsigned int* n_cells, hipStream_t s);
hipError_t launch_level_build(const unsigned long long* codes, long long n, int level,
                              unsigned long long* keys, unsigned int* start,
                              unsigned int* cnt, unsigned int mask, hipStream_t s);
hipError_t launch_knn_mean(const Grid& g, const float* sorted, const unsigned int* idx,
                           long long n, int k, float* mean_knn, hipStream_t s);
// The k nearest others of each point: out_idx / out_d2 (n x k, by original
// index) ascending by (squared distance, index); -1 where fewer exist.
hipError_t launch_knn_list(const Grid& g, const float* sorted, const unsigned int* idx,
                           long long n, int k, int* out_idx, float* out_d2, hipStream_t s);
// hipCUB radix sort of (code, index) pairs over bits 0..62; with temp null
// only *temp_bytes is set.
hipError_t sort_codes(unsigned long long* keys_in, unsigned long long* keys_out,
                      unsigned int* vals_in, unsigned int* vals_out, int n, void* temp,
                      size_t* temp_bytes, hipStream_t s);
hipError_t launch_sor_count(const float* mean_knn, long long n, float threshold, int* counts,
                            hipStream_t s);
hipError_t launch_sor_scatter(const float* mean_knn, long long n, float threshold,
                              const long long* offsets, const float* points,
                              const float* normals, const uint8_t* colors, float* out_points,
                              float* out_normals, uint8_t* out_colors, hipStream_t s);

}  // namespace fu
