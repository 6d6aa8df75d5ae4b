// Device launchers of the neighbour selection before the depthmaps
// (sfm::dense_helpers, opensfm/src/lib/sfm/src/dense_helpers.cc).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cv {

constexpr int kMaxShots = 16384;       // dense shot-pair tables
constexpr int kMaxNeighbours = 64;     // best neighbours per shot

hipError_t launch_valid(const long long* off, const int* shots, long long n, int n_shots,
                        uint8_t* flag, hipStream_t s);
// Per 256-entry block count of flags, then each flagged entry's rank
// (row-major) and the rank -> entry index.
hipError_t launch_flag_count(const uint8_t* flag, long long n, int* counts, hipStream_t s);
hipError_t launch_flag_rank(const uint8_t* flag, long long n, const long long* offsets, int* rank,
                            int* index, hipStream_t s);
hipError_t launch_gather_points(const double* coords, const int* keep, long long m, float* fc,
                                hipStream_t s);
hipError_t launch_kdist(const int* kidx, const float* kd2, long long m, int k, float* kdist,
                        hipStream_t s);
// One hook + compress round of the union-find; *changed set when a hook
// happened.
hipError_t launch_union_round(const int* kidx, const float* kd2, long long m, int k, float r2,
                              int* parent, int* changed, hipStream_t s);
hipError_t launch_is_root(const int* parent, long long m, uint8_t* flag, hipStream_t s);
hipError_t launch_sp_of(const int* parent, const int* root_rank, long long m, const int* keep,
                        unsigned int* sp, unsigned int* member, int* point_sp,
                        unsigned int* counts, hipStream_t s);
hipError_t launch_centroid(const double* coords, const int* keep, const unsigned int* members,
                           const long long* moff, long long nsp, double* out, int* out_members,
                           hipStream_t s);
hipError_t launch_obs_count(const long long* off, const int* shots, const int* keep, long long m,
                            int n_shots, long long* cnt, hipStream_t s);
hipError_t launch_obs_keys(const long long* off, const int* shots, const int* keep,
                           const unsigned int* sp, const long long* pos, long long m, int n_shots,
                           unsigned long long* keys, hipStream_t s);
hipError_t launch_unique(const unsigned long long* keys, long long t, uint8_t* flag,
                         hipStream_t s);
hipError_t launch_vis_write(const unsigned long long* keys, const int* index, long long nv,
                            int n_shots, int* vis, unsigned int* counts, hipStream_t s);
hipError_t launch_pair_count(const long long* voff, const int* vis, long long nsp, int n_shots,
                             unsigned int* table, hipStream_t s);
hipError_t launch_edge_flag(const unsigned int* table, int n_shots, uint8_t* flag,
                            hipStream_t s);
hipError_t launch_edge_write(const unsigned int* table, const int* index, long long ne,
                             int n_shots, int* edges, int* weights, hipStream_t s);
hipError_t launch_pair_score(const double* origins, int n_shots, const double* sp,
                             const long long* voff, const int* vis, const int* ntracks,
                             long long nsp, double cos_lo, double cos_hi, unsigned int* count,
                             unsigned int* score, long long* sumfx, hipStream_t s);
hipError_t launch_select(const unsigned int* count, const unsigned int* score,
                         const long long* sumfx, int n_shots, int num_neighbors,
                         long long min_best, long long min_all, double cos_lo, double cos_hi,
                         int* best, int* best_n, long long* all_n, hipStream_t s);
hipError_t launch_select_all(const unsigned int* count, const unsigned int* score,
                             const long long* sumfx, int n_shots, int num_neighbors,
                             long long min_all, const int* best, const int* best_n,
                             const long long* all_off, int* all, hipStream_t s);
// hipCUB radix sorts; with temp null only *temp_bytes is set.
hipError_t sort_u32_pairs(unsigned int* keys_in, unsigned int* keys_out, unsigned int* vals_in,
                          unsigned int* vals_out, int n, int end_bit, void* temp,
                          size_t* temp_bytes, hipStream_t s);
hipError_t sort_u64_keys(unsigned long long* keys_in, unsigned long long* keys_out, int n,
                         int end_bit, void* temp, size_t* temp_bytes, hipStream_t s);

}  // namespace cv
