// Launch wrappers of the engine's HIP kernels (defined in pm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "pm_types.h"

namespace pm {

// u8 quads (exact scorer) and their FP16 bilinear coefficients (8 bytes
// per texel; fast FP32 scorer) of a u8 image.
hipError_t launch_build_quad(const uint8_t* img, uint32_t* quad, uint32_t* quadc,
                             int W, int H, hipStream_t s);
// 8-byte quads of FP16 (precision 1) or BF16 (precision 2) values.
hipError_t launch_build_quad16(const uint8_t* img, uint32_t* quad16, int W,
                               int H, int precision, hipStream_t s);
// Whether PatchMatch kernels exist for this (sample precision, patch size).
bool precision_supported(int precision, int patch);
// The PatchMatch scorer (device homography, sample precision a.precision)
// for planes[3k..3k+2] at pixels ij[2k], ij[2k+1] against one view.
hipError_t launch_score_hyp_fast(const PmArgs& a, int view, int n,
                                 const int* ij, const float* planes, float* out,
                                 hipStream_t s);
hipError_t launch_sweep(const PmArgs& a, const float* Htab, const float* planes,
                        int nplanes, hipStream_t s);
hipError_t launch_score_hyp(const PmArgs& a, int view, int n, const int* ij,
                            const float* Hs, float* out, hipStream_t s);
hipError_t launch_pm_init(const PmArgs& a, bool sample, hipStream_t s);
// queues: 8 zeroed counters (the band queues of one launch).
// A level's state (W x H float4) as a 4-float-per-pixel initialisation map
// for the next finer level (k_state_to_map).
hipError_t launch_state_to_map(const float* state, int W, int H, float log_min, float log_max,
                               int median, float* map, hipStream_t s);
// The same for every job of a batch (map0 laid out like the state, 4 floats
// per pixel at the jobs' pixel offsets; depth ranges from the job table).
hipError_t launch_state_to_map_batch(const PmJob* jobs, int njobs, int pix_total,
                                     const float* state0, int median, float* map0,
                                     hipStream_t s);
// lanes_per_pixel: 0 chooses (kLanesPerPixelSmall lanes per pixel below
// kGroupBelowPixels pixels of one colour, FP32 samples, not in sample mode),
// 1 forces one lane per pixel, kLanesPerPixelSmall forces the groups (FP32).
hipError_t launch_pm_half(const PmArgs& a, bool sample, int colour, int iter,
                          unsigned* queues, hipStream_t s, int lanes_per_pixel = 0);
bool half_uses_groups(const PmArgs& a, int lanes_per_pixel);
// Box downsampling (mode 0 mean, 1 min) by an integer factor.
hipError_t launch_downsample_u8(const uint8_t* in, int W, int H, int sc,
                                int mode, uint8_t* out, int Wl, int Hl,
                                hipStream_t s);
// Learned-initialiser input (image + nplanes-plane NCC volume, BF16 NHWC,
// 40 channels) over an ow x oh grid for the (coarse) views of `a`; patch
// sizes 5 and 7.
hipError_t launch_init_input(const PmArgs& a, int nplanes, float inv_min,
                             float inv_step, void* out, int ow, int oh,
                             hipStream_t s);
// Median-consistent re-selection of a's maps, in place: count (1 unsigned)
// and list_idx / list_med (W*H entries each) are scratch, res holds W*H
// entries, items (plain mode) holds the window phase's work items (first /
// todo W*H entries, the item arrays items.cap entries).  salt varies the
// refinement draws between passes.
hipError_t launch_pm_consist(const PmArgs& a, bool sample, unsigned* count, unsigned* list_idx,
                             float* list_med, const ConsistOut& res, unsigned salt,
                             const ConsistItems& items, hipStream_t s);
// The same over every job of a batch (list_job: W*H entries over all jobs).
hipError_t launch_pm_consist_batch(const PmJob* jobs, int njobs, int pix_total, int patch,
                                   int precision, bool sample, unsigned* count,
                                   unsigned* list_idx, float* list_med, int* list_job,
                                   const ConsistOut& res, unsigned salt,
                                   const ConsistItems& items, unsigned long long* counter,
                                   hipStream_t s);
hipError_t launch_pack(const PmArgs& a, int median, float* depth, float* plane,
                       float* score, int* nghbr, hipStream_t s);
// Batched launches over the jobs of a PmJob table (see pm_types.h).  `tile`
// selects the red/black lane layout: 0 = 8x8, 1 = 16x16, 2 = 32x8 lanes.
hipError_t launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s);
void tile_shape(int tile, int* tx, int* ty);
hipError_t launch_pm_init_batch(const PmJob* jobs, int njobs, int total,
                                int patch, int precision, bool sample,
                                unsigned* queue, unsigned long long* count,
                                hipStream_t s);
// launch_pm_half_batch: queue = 8 zeroed band counters.
hipError_t launch_pm_half_batch(const PmJob* jobs, int njobs, int total,
                                int patch, int precision, bool sample, int tile,
                                int colour, int iter, unsigned* queue,
                                unsigned long long* count, hipStream_t s);
hipError_t launch_pack_batch(const PmJob* jobs, int njobs, int total,
                             int median, hipStream_t s);
// Cluster-estimator outputs (depth within [dmin, dmax], unit normal, cost,
// confidence) of n pixels.
hipError_t launch_cluster_out(const float* depth, const float* plane, const float* score, int n,
                              float dmin, float dmax, float* out_depth, float* normal,
                              float* cost, float* conf, hipStream_t s);
hipError_t launch_clean_batch(const ClJob* jobs, int njobs, int total,
                              const ClView* views, float thr, int min_views,
                              hipStream_t s);
// GPUDepthmapCleaner: cv::resize(INTER_LINEAR) of a float map, and the
// carving/backward-check clean of ref_depth against nsrc resized maps.
hipError_t launch_resize_linear(const float* src, int sw, int sh, float* dst,
                                int w, int h, hipStream_t s);
hipError_t launch_gpu_clean(const float* ref_depth, const float* const* src,
                            const GcCamera* cams, int w, int h, int nsrc,
                            float thr, int min_views, float carving,
                            int max_carved, float* out, hipStream_t s);
// DepthmapPruner: kept-pixel counts per 256-pixel block, then the kept
// pixels written at the blocks' offsets (row-major order).
hipError_t launch_prune_count(const PrView* views, int n, const PrRef& ref,
                              float thr, int npix, int* counts, hipStream_t s);
hipError_t launch_prune_write(const PrView* views, int n, const PrRef& ref,
                              float thr, int npix, const long long* offsets,
                              const uint8_t* color, const uint8_t* label,
                              float* points, float* normals, uint8_t* out_color,
                              uint8_t* out_label, hipStream_t s);
hipError_t launch_clean(const ClView* views, const ClRef* ref, int n, int W,
                        int H, float thr, int min_views, float* out,
                        hipStream_t s);

}  // namespace pm
