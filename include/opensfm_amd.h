/*
 * opensfm_amd.h -- C ABI of the MI355X PatchMatch depthmap engine.
 *
 * Drop-in boundary for OpenSfM's legacy dense path.  Each entry point
 * replaces one method of dense::DepthmapEstimator / dense::DepthmapCleaner
 * (opensfm/src/lib/dense/depthmap.h:57-129, depthmap.cc:134-555), which the
 * reference exposes to Python as pydense.DepthmapEstimator /
 * pydense.DepthmapCleaner (opensfm/src/lib/dense/pydense.pyi:28-60).
 *
 * Conventions
 *   - Plain pointers and sizes only.  Matrices are row-major double[9]
 *     (cv::Matx33d(const double*) reads row-major, depthmap.cc:150-152).
 *   - Images are uint8 grayscale, `stride` bytes between rows.
 *   - Inputs are copied on the call (the reference deep-copies in AddView,
 *     depthmap.cc:156-157); output buffers are caller-owned host memory.
 *   - Every function returns PM_OK or a negative status; pm_last_error()
 *     gives the message of the calling thread's last failure.  No C++
 *     exception crosses this boundary.
 *   - A handle is not thread-safe; use one per thread (the reference's
 *     callers do the same, opensfm/dense.py:191-238).  Each handle owns one
 *     HIP stream on its device.
 */
#ifndef OPENSFM_AMD_H_
#define OPENSFM_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the Python layer maps them to ValueError / RuntimeError /
 * IndexError like pybind11 maps std::invalid_argument / runtime_error /
 * out_of_range, opensfm/src/lib/dense/depthmap_bind.h:253-259) */
#define PM_OK 0
#define PM_ERR_INVALID (-1)
#define PM_ERR_RUNTIME (-2)
#define PM_ERR_RANGE (-3)

/* compute modes */
#define PM_MODE_PLANE_SWEEP 0      /* ComputeBruteForce, depthmap.cc:184-200 */
#define PM_MODE_PATCHMATCH 1       /* ComputePatchMatch, depthmap.cc:202-213 */
#define PM_MODE_PATCHMATCH_SAMPLE 2 /* ComputePatchMatchSample, :215-227 */

/* sample precision of the PatchMatch scorer (the plane sweep and
 * pm_estimator_score_hypotheses are always the reference's FP32) */
#define PM_PRECISION_FP32 0 /* FP32 samples: random-plane launches blend the
                              u8 2x2 quads, converged launches the same
                              quads stored as FP16 bilinear coefficients
                              (integers <= 510, exact), FP32 arithmetic */
#define PM_PRECISION_FP16 1 /* FP16 texels and weights, FP32 dot products */
#define PM_PRECISION_BF16 2 /* BF16 texels and weights, FP32 dot products */

typedef struct pm_estimator pm_estimator;
typedef struct pm_cleaner pm_cleaner;

/* Work and timing of the last compute on a handle. */
typedef struct pm_stats {
  int64_t scorings;      /* (hypothesis, view) NCC scorings executed */
  int64_t dominant_scorings; /* of which in the dominant kernel's launches */
  int32_t patch_size;
  int32_t width, height; /* reference view */
  int32_t launches;      /* launches of the dominant kernel (the plane
                            sweep, or the red/black PatchMatch halves,
                            both levels' with set_pyramid) */
  float dominant_ms;     /* summed HIP-event time of those launches */
  float total_ms;        /* HIP-event time of the whole device compute
                            (both levels' with set_pyramid) */
} pm_stats;

/* Message of the calling thread's last failed call ("" if none). */
const char* pm_last_error(void);

/* Number of visible HIP devices.
 * DepthmapClusterEstimator::NumDevices  depthmap_bind.h:85 */
int pm_device_count(int* count);

/* Name (NUL-terminated, truncated to name_capacity - 1 bytes), global memory
 * and currently free memory of HIP device `device`.
 * DeviceName / DeviceMemoryBytes  depthmap_bind.h:87-93 (CLDevice name and
 * global memory size, opencl_utils.h:34-125); free_bytes is the runtime's
 * live figure (hipMemGetInfo), which the reference does not report. */
int pm_device_info(int device, char* name, int name_capacity, uint64_t* total_bytes,
                   uint64_t* free_bytes);

/* Memory-reservation bookkeeping of the dense scheduler: reserve (bytes > 0)
 * or release (bytes < 0) on device `device`; releasing more than is
 * reserved leaves zero.  *available = global memory - reserved bytes (0 if
 * more is reserved than exists).  Nothing is allocated.
 * ReserveDeviceMemory / ReleaseDeviceMemory / DeviceAvailableMemory
 * depthmap_bind.h:99-108, CLDevice::ReserveMemory / ReleaseMemory /
 * AvailableMemory opencl_utils.h:68-76. */
int pm_device_reserve(int device, int64_t bytes);
int pm_device_available(int device, uint64_t* available);

/* ---- DepthmapEstimator ------------------------------------------------ */

/* DepthmapEstimator::DepthmapEstimator  depthmap.cc:134-144
 * (defaults: patch 7, 50 planes, 3 iterations, min patch SD 5). */
int pm_estimator_create(int device, pm_estimator** out);
int pm_estimator_destroy(pm_estimator* est);

/* DepthmapEstimator::AddView  depthmap.cc:146-162.  View 0 is the
 * reference view; K, R, t are world->camera (x_c = R X + t). */
int pm_estimator_add_view(pm_estimator* est, const double* K, const double* R,
                          const double* t, const uint8_t* image,
                          const uint8_t* mask, int width, int height,
                          int stride);

/* DepthmapEstimator::SetDepthRange  depthmap.cc:164-169 */
int pm_estimator_set_depth_range(pm_estimator* est, double min_depth,
                                 double max_depth, int num_depth_planes);
/* DepthmapEstimator::SetPatchMatchIterations  depthmap.cc:171-173 */
int pm_estimator_set_patchmatch_iterations(pm_estimator* est, int n);
/* DepthmapEstimator::SetPatchSize  depthmap.cc:175-178 (odd, 3..11) */
int pm_estimator_set_patch_size(pm_estimator* est, int size);
/* DepthmapEstimator::SetMinPatchSD  depthmap.cc:180-182 */
int pm_estimator_set_min_patch_sd(pm_estimator* est, float sd);
/* Seed of the counter-based generator that replaces rand() / mt19937
 * (depthmap.cc:127-132, 141).  Same seed => bit-identical maps. */
int pm_estimator_set_seed(pm_estimator* est, uint64_t seed);
/* Far-propagation probes per direction of the red/black kernel (default 6;
 * 0 = direct neighbours only). */
int pm_estimator_set_far_probes(pm_estimator* est, int n);
/* Half-resolution level (not in the reference's legacy estimator; the
 * reference's OpenCL engine runs a pyramid, estimator.cc:561-658): with
 * half_iterations > 0, plain PatchMatch first runs half_iterations
 * iterations on the views box-averaged by 2 (K scaled about pixel centres,
 * same settings, its own seed), and its planes -- where the depth is live and
 * the median test keeps it -- seed this level's initialisation (bilinear in
 * depth and slopes, as an initialisation map would); the rest start from
 * random planes.  Then set_patchmatch_iterations iterations run here.
 * Default 0 (off).  With an initialisation map set, the map seeds the
 * coarsest level (its scale must be divisible by 2 per coarser level) and
 * each level seeds the next as above.  In sample
 * mode every level runs in sample mode. */
int pm_estimator_set_pyramid(pm_estimator* est, int half_iterations);
/* Levels of that pyramid, the full resolution included (default 2; in
 * [2, 8]; the reference's OpenCL engine's hierarchy_levels, estimator.h:28,
 * scale 1 / 2^(N - 1 - level) at estimator.cc:168-172): with N > 2 the half
 * level is itself seeded by a quarter level, and so on; every level coarser
 * than the full one runs half_iterations iterations. */
int pm_estimator_set_pyramid_levels(pm_estimator* est, int levels);
/* The pyramid with its own iteration count per coarser level: n coarser
 * levels (1 to 7; levels = n + 1), iterations[0] for the coarsest (views
 * box-averaged by 2^n) up to iterations[n - 1] for the half level, each at
 * least 1; set_patchmatch_iterations still counts the full level's.
 * set_pyramid and set_pyramid_levels replace the schedule by a uniform one. */
int pm_estimator_set_pyramid_schedule(pm_estimator* est, const int32_t* iterations, int n);
/* Median-consistent re-selection at the pyramid's coarser levels (default
 * 1: as set_consistency; 0: at the full level only).  Measured: on the
 * 1600x1200 view with 2, 2, 1 + 2 iterations off is 3 % faster at the same
 * accuracy, but a learned map seeding the coarsest level needs it on. */
int pm_estimator_set_pyramid_consistency(pm_estimator* est, int enable);
/* Lanes per pixel of the red/black PatchMatch launches (no counterpart in
 * the reference, whose CPU loop scores the views of a candidate one after
 * another, depthmap.cc:408-419): 0 (default) scores each pixel on 8 lanes,
 * one share of the views each, when a launch has fewer than 32768 pixels of
 * one colour (the coarse levels of a pyramid) and FP32 samples are used, else
 * on one lane; 1 always one lane; 8 always 8 lanes (FP32 samples).  The maps
 * are bit-identical for every setting; sample mode always uses one lane. */
int pm_estimator_set_lanes_per_pixel(pm_estimator* est, int lanes);
/* Random refinements per pixel update (default 6, the reference's count,
 * depthmap.cc:337; the 0.02 * 0.3^k / 0.5 * 0.8^k schedule restarts every
 * 6 steps). */
int pm_estimator_set_refinement_steps(pm_estimator* est, int n);
/* Exploration samples per pixel update (0..8, default 3; 0 = the
 * reference's candidate list, depthmap.cc:305-374).  Not in the reference:
 * sample 0 and 1 are the planes through the 3-D points of the four
 * neighbours at distance 1 and 3, further samples keep the current depth
 * with a normal drawn within 87 degrees of the optical axis.  Scored after
 * propagation and before the refinements, only at pixels whose best score
 * is below 0.97.  They reach strongly slanted surfaces that the reference
 * reaches by chaining refinements along its sequential sweep. */
int pm_estimator_set_exploration(pm_estimator* est, int n);
/* Sample precision of PatchMatch (PM_PRECISION_*, default FP32; FP16 and
 * BF16 need patch size 5 or 7).  Image samples are held as FP16/BF16 2x2
 * quads; each bilinear row blend is one FP16/BF16 dot product accumulated
 * in FP32, and the NCC sums and plane/depth state stay FP32.  Not in the
 * reference (all FP32, depthmap.cc:441-478). */
int pm_estimator_set_precision(pm_estimator* est, int precision);
/* Median-consistent re-selection (default 1; 0 = the reference's search).
 * Not in the reference.  After every red/black iteration but the first, and
 * after the last, each live pixel whose depth PostProcess (depthmap.cc:
 * 489-502) would reject -- more than 5 % from the median of its 5x5 window --
 * re-selects its plane among the planes of that window's other pixels whose
 * depth at this pixel passes the same test: best NCC over the views, then
 * the reference's 6 refinements around it, each kept only while it passes
 * the test; the new plane is kept when its NCC is at least 0.5 and, after
 * the last iteration, when its depth lies within 1 % of the depths of at
 * least 10 of the 24 window pixels.  Such pixels hold planes that fit their
 * patch better than the surrounding surface does (next to occlusion edges,
 * planes bridging both sides). */
int pm_estimator_set_consistency(pm_estimator* est, int enable);
/* Run PostProcess (5x5 median, depthmap.cc:489-502) after PatchMatch
 * (default 1, as the reference always does; 0 returns the raw depth). */
int pm_estimator_set_post_process(pm_estimator* est, int enable);

/* ComputeBruteForce / ComputePatchMatch / ComputePatchMatchSample
 * (depthmap.cc:184-227) into caller buffers: depth H*W f32, plane H*W*3 f32
 * (the plane vector v of v^T X + 1 = 0 in the reference camera frame),
 * score H*W f32, nghbr H*W i32 (DepthmapEstimatorResult, depthmap.h:51-56).
 * Equivalent to pm_estimator_run followed by pm_estimator_download. */
int pm_estimator_compute(pm_estimator* est, int mode, float* depth,
                         float* plane, float* score, int32_t* nghbr);

/* Device-only compute: leaves the four maps in HBM (for timing). */
int pm_estimator_run(pm_estimator* est, int mode);
/* Copies the maps of the last run to host buffers. */
int pm_estimator_download(pm_estimator* est, float* depth, float* plane,
                          float* score, int32_t* nghbr);

/* DepthmapEstimator::ComputePlaneImageScore  depthmap.cc:441-478 for n
 * fixed hypotheses: pixel (ij[2k], ij[2k+1]) = (row, col), plane
 * planes[3k..3k+2], neighbour view views[k] (1..N-1); homographies composed
 * in FP64 as PlaneInducedHomographyBaked (depthmap.cc:96-102), NCC in the
 * reference's FP32 order.  Pixels must be at least (patch-1)/2 from the
 * border. */
int pm_estimator_score_hypotheses(pm_estimator* est, int n, const int32_t* ij,
                                  const float* planes, const int32_t* views,
                                  float* out);

/* The same hypotheses scored by the PatchMatch scorer: FP32 homography
 * M_v + b_v (K_0^-T v)^T composed on the device, FMA-contracted NCC, samples
 * in the estimator's precision.  For measuring the scorer's error against
 * an FP64 recomputation. */
int pm_estimator_score_hypotheses_fast(pm_estimator* est, int n,
                                       const int32_t* ij, const float* planes,
                                       const int32_t* views, float* out);

/* Learned initialisation (not in the reference; RandomInitialization,
 * depthmap.cc:237-256, is what it replaces where it is confident).
 *
 * pm_estimator_initializer_input writes the initialiser network's input for
 * the current views into device memory `out` (BF16, out_h x out_w x 40,
 * NHWC; out_w >= W / scale, out_h >= H / scale): channel 0 the reference
 * image box-averaged by `scale`, (x - 128) / 64; channels 1..nplanes the
 * plane-sweep cost volume at that scale -- max-over-views NCC of nplanes
 * fronto-parallel planes uniform in inverse depth over the depth range,
 * scored with FP16 samples (-1 where untextured); other channels and the
 * padding 0.  Needs patch size 5 or 7.  Synchronous.
 *
 * pm_estimator_set_init_map gives PatchMatch's initialisation a coarse map
 * (device memory, h x w cells of `stride` floats: t, sx, sy, c) at 1/scale:
 * each pixel blends the map bilinearly, and where the confidence logit
 * c >= 0 starts from depth exp(log_min + t (log_max - log_min)) and normal
 * (sx, sy, -1) instead of a random plane (and takes every exploration
 * sample from the first iteration on).  The map is read by the
 * initialisation and the first red/black iteration, so it must stay valid
 * through every compute made while it is set; NULL restores random
 * initialisation.  With set_pyramid the map seeds the coarsest level, at
 * scale / 2^(levels - 1) of that level's pixels. */
int pm_estimator_initializer_input(pm_estimator* est, int scale, int nplanes,
                                   void* out, int out_w, int out_h);
int pm_estimator_set_init_map(pm_estimator* est, const float* map, int w,
                              int h, int stride, int scale);
/* Depth hypotheses tried at a seeded, confident pixel (1..15),
 * all with the map's normal: the map's depth, then scaled by exp(+-0.03)
 * and exp(+-0.06) (hypotheses 2..5), then the best so far scaled by
 * exp(+-0.01), exp(+-0.02), ... (the rest); the best-scoring one starts
 * PatchMatch.  Default 1 (no search). */
int pm_estimator_set_init_search(pm_estimator* est, int n);
/* Per-pixel Gaussian spread around the map's plane at confident pixels
 * (default 0.05 in log depth, 0.3 in slope; 0, 0: the map's plane itself).
 * Distinct nearby hypotheses at neighbouring pixels keep PatchMatch's
 * propagation a search. */
int pm_estimator_set_init_spread(pm_estimator* est, float sigma_log_depth,
                                 float sigma_slope);

int pm_estimator_stats(pm_estimator* est, pm_stats* out);

/* ---- DepthmapCleaner -------------------------------------------------- */

/* DepthmapCleaner::DepthmapCleaner  depthmap.cc:504-505
 * (defaults: threshold 0.01, 2 consistent views). */
int pm_cleaner_create(int device, pm_cleaner** out);
int pm_cleaner_destroy(pm_cleaner* cl);
/* DepthmapCleaner::SetSameDepthThreshold  depthmap.cc:507-509 */
int pm_cleaner_set_same_depth_threshold(pm_cleaner* cl, float t);
/* DepthmapCleaner::SetMinConsistentViews  depthmap.cc:511-513 */
int pm_cleaner_set_min_consistent_views(pm_cleaner* cl, int n);
/* DepthmapCleaner::AddView  depthmap.cc:515-522 (view 0 is cleaned) */
int pm_cleaner_add_view(pm_cleaner* cl, const double* K, const double* R,
                        const double* t, const float* depth, int width,
                        int height);
/* DepthmapCleaner::Clean  depthmap.cc:524-555 into clean_depth
 * (view 0's H*W f32). */
int pm_cleaner_clean(pm_cleaner* cl, float* clean_depth);
/* HIP-event time of the last clean kernel, milliseconds. */
int pm_cleaner_last_ms(pm_cleaner* cl, float* ms);

/* Maximum number of views of pm_cleaner_clean_noalloc. */
#define PM_NOALLOC_MAX_VIEWS 64
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same clean as pm_cleaner_clean (DepthmapCleaner::
 * Clean's arithmetic, bit for bit) over n_views depthmaps already in device
 * memory, written to the caller's device buffer, with no device allocation.
 * K, R (n_views x 9, row-major) and t (n_views x 3) are host arrays as in
 * pm_cleaner_add_view; depths[v] is a device pointer to view v's
 * heights[v] x widths[v] f32 map; view 0 is cleaned into clean_depth
 * (device, heights[0] x widths[0] f32).  stream is a hipStream_t (NULL: the
 * default stream); the call returns after the clean has finished on it.
 * The per-view table lives in a fixed device array of this library (about
 * 12 KiB per device, held until the process ends); more than
 * PM_NOALLOC_MAX_VIEWS views is PM_ERR_RANGE.  Every call on one device
 * shares that array, so calls that overlap on one device race. */
int pm_cleaner_clean_noalloc(int device, int n_views, const double* K,
                             const double* R, const double* t,
                             const float* const* depths,
                             const int32_t* widths, const int32_t* heights,
                             float same_depth_threshold,
                             int min_consistent_views, float* clean_depth,
                             void* stream);

/* ---- GPUDepthmapCleaner ----------------------------------------------- */

/* The live pipeline's cleaner (opensfm/src/lib/dense/cleaner.h:16-63,
 * cleaner.cc:56-179, kernel acmmp_clean_depthmap in opencl_kernels.h:
 * 2155-2273): a pixel of view ref_idx keeps its depth when at least
 * min(min_consistent_views, 1 + n) views (itself included) agree -- the
 * neighbour's bilinear depth at the projection within the threshold of the
 * projected depth, and the point back-projected from that depth landing
 * within 2 px (L1) and the threshold in depth in the reference view -- and
 * fewer than max_carved_views neighbours see more than (1 + carving
 * threshold) times further along the ray.  FP32 throughout.  At most 16
 * neighbours are used; a neighbour of another size is resized to the
 * reference's (cv::resize INTER_LINEAR) with its K unchanged, as the
 * reference does.  Defaults: threshold 0.01, 3 views, carving 0.2, 1
 * carved view. */
typedef struct pm_gpu_cleaner pm_gpu_cleaner;
int pm_gpu_cleaner_create(int device, pm_gpu_cleaner** out);
int pm_gpu_cleaner_destroy(pm_gpu_cleaner* c);
int pm_gpu_cleaner_set_same_depth_threshold(pm_gpu_cleaner* c, float t);
int pm_gpu_cleaner_set_min_consistent_views(pm_gpu_cleaner* c, int n);
/* negative values clamp to 0 (cleaner.cc:30-36) */
int pm_gpu_cleaner_set_carving_threshold(pm_gpu_cleaner* c, float t);
int pm_gpu_cleaner_set_max_carved_views(pm_gpu_cleaner* c, int n);
/* GPUDepthmapCleaner::AddView: *index receives the view's index. */
int pm_gpu_cleaner_add_view(pm_gpu_cleaner* c, const double* K,
                            const double* R, const double* t,
                            const float* depth, int width, int height,
                            int* index);
/* GPUDepthmapCleaner::Clean into clean_depth (view ref_idx's H*W f32);
 * indices out of range give PM_ERR_RANGE. */
int pm_gpu_cleaner_clean(pm_gpu_cleaner* c, int ref_idx,
                         const int32_t* neighbor_ids, int n,
                         float* clean_depth);
/* GPUDepthmapCleaner::Clear */
int pm_gpu_cleaner_clear(pm_gpu_cleaner* c);
int pm_gpu_cleaner_last_ms(pm_gpu_cleaner* c, float* ms);

/* ---- DepthmapPruner (depthmap.h:131-153, depthmap.cc:557-637) --------- */

/* Merges view 0's depthmap into points, dropping each pixel that another
 * view sees, at the rounded projection, with a depth beyond (1 - threshold)
 * times the point's and a larger -n_z / depth * f (a finer sample of the
 * surface).  Kept pixels come out in row-major order with the reference's
 * types: FP64 back-projection and projection, cv::normalize of the plane
 * (FP64 norm), FP32 area and normal rotation R0^T.  Default threshold
 * 0.01.  Only view 0's colour and label are read (the merge outputs them). */
typedef struct pm_pruner pm_pruner;
int pm_pruner_create(int device, pm_pruner** out);
int pm_pruner_destroy(pm_pruner* c);
/* DepthmapPruner::SetSameDepthThreshold (depthmap.cc:559-561) */
int pm_pruner_set_same_depth_threshold(pm_pruner* c, float t);
/* DepthmapPruner::AddView (depthmap.cc:563-575): K, R row-major 3x3, t 3;
 * depth H*W f32, plane H*W*3 f32, color H*W*3 u8, label H*W u8. */
int pm_pruner_add_view(pm_pruner* c, const double* K, const double* R,
                       const double* t, const float* depth,
                       const float* plane, const uint8_t* color,
                       const uint8_t* label, int width, int height);
/* DepthmapPruner::Prune (depthmap.cc:577-636): runs the merge; *count
 * receives the number of kept points, which stay on the device until
 * pm_pruner_download copies count*3 floats of points and of normals,
 * count*3 bytes of colours and count bytes of labels. */
int pm_pruner_prune(pm_pruner* c, int64_t* count);
int pm_pruner_download(pm_pruner* c, float* points, float* normals,
                       uint8_t* colors, uint8_t* labels);
int pm_pruner_last_ms(pm_pruner* c, float* ms);

/* ---- DepthmapFuser (fuser.h:28-75, src/fuser.cc:37-480) -------------- */

/* Fuses the primary views' depthmaps into a point cloud: a pixel whose
 * point is seen consistently (depth within max_depth_error in front and
 * behind_depth_factor times that behind, forward-backward reprojection
 * within max_reproj_error px, normals within max_normal_error degrees) by
 * at least min_num_consistent observations (itself included) among its
 * neighbour views' not-yet-fused pixels emits the component-wise medians,
 * and all pixels it used are marked fused.  Then statistical outlier
 * removal (mean distance to the sor_knn nearest points above mean +
 * factor * std is dropped; knn 0 disables it).  Float arithmetic in the
 * reference's order.  Views are fused in index order, pixels in row-major
 * order: the result of the reference with one thread (its multi-threaded
 * runs share the fused flags between threads without order, so their
 * result depends on the schedule); set_num_threads is accepted and has no
 * effect on the result.  Fused flags persist across pm_fuser_fuse calls, as
 * in the reference.  Limits: 64 valid neighbours per primary view, k <= 32.
 * Defaults: 3 views, 2 px, 0.01, 10 degrees, margin 5, k 12, factor 2.5,
 * behind 0.3. */
typedef struct pm_fuser pm_fuser;
int pm_fuser_create(int device, pm_fuser** out);
int pm_fuser_destroy(pm_fuser* f);
int pm_fuser_set_min_num_consistent(pm_fuser* f, int n);
int pm_fuser_set_max_reproj_error(pm_fuser* f, float px);
int pm_fuser_set_max_depth_error(pm_fuser* f, float ratio);
int pm_fuser_set_max_normal_error(pm_fuser* f, float degrees);
/* negative margins clamp to 0 (fuser.cc:62-64) */
int pm_fuser_set_border_margin(pm_fuser* f, int px);
int pm_fuser_set_num_threads(pm_fuser* f, int n);
/* negative k clamps to 0 (fuser.cc:68-71) */
int pm_fuser_set_sor_params(pm_fuser* f, int knn, float stddev_factor);
/* clamped to [0, 1] (fuser.cc:73-75) */
int pm_fuser_set_behind_depth_factor(pm_fuser* f, float v);
/* DepthmapFuser::AddView (fuser.cc:77-93): K, R row-major 3x3, t 3; depth
 * H*W f32, normal H*W*3 f32 (camera frame), color H*W*3 u8, mask H*W u8 or
 * NULL for none; neighbour view indices (entries outside [0, views) at fuse
 * time are skipped); primary 0 makes the view a source only. */
int pm_fuser_add_view(pm_fuser* f, const double* K, const double* R,
                      const double* t, const float* depth,
                      const float* normal, const uint8_t* color,
                      const uint8_t* mask, int width, int height,
                      const int32_t* neighbor_ids, int n_neighbors,
                      int primary);
/* DepthmapFuser::Fuse: *count receives the number of points, which stay on
 * the device until pm_fuser_download copies count*3 floats of points and of
 * normals and count*3 bytes of colours. */
int pm_fuser_fuse(pm_fuser* f, int64_t* count);
int pm_fuser_download(pm_fuser* f, float* points, float* normals,
                      uint8_t* colors);
/* Points before outlier removal, greedy rounds run, device time of the
 * fusion and of the outlier removal in the last fuse. */
int pm_fuser_stats(pm_fuser* f, int64_t* fused_before_sor, int64_t* rounds,
                   float* fuse_ms, float* sor_ms);

/* ---- Neighbour selection (sfm/dense_helpers.h:40-67) ----------------- */

/* BuildCovisibilityGraph (dense_helpers.cc:30-326) over plain arrays:
 * n_points landmarks (FP64 x, y, z) with their observing shots in CSR form
 * (obs_off[n_points + 1], obs_shot: processable shot indices; entries
 * outside [0, n_shots) are not processable).  Points seen by >= 2 shots are
 * merged with their k = max(1, min(fuse_knn, m - 1)) nearest neighbours
 * (float L2, ties to the lower index) within fuse_radius_factor times the
 * median of the per-point (k/2)-th nearest distances (union-find) into
 * super-points: FP64 centroid in member order, sorted shot union, members.
 * Edges are the shot pairs (lo < hi) sharing a super-point, weighted by the
 * count.  Super-points are numbered by their smallest member and edges in
 * (lo, hi) order (the reference's orders come from hash-map iteration).
 * Limits: n_shots <= 16384, k <= 32. */
typedef struct pm_covis pm_covis;
int pm_covis_create(int device, pm_covis** out);
int pm_covis_destroy(pm_covis* c);
int pm_covis_build(pm_covis* c, int64_t n_points, const double* coords,
                   const int64_t* obs_off, const int32_t* obs_shot,
                   int n_shots, int fuse_knn, double fuse_radius_factor,
                   int64_t* n_super_points, int64_t* n_vis,
                   int64_t* n_edges);
/* point_sp[n_points] (-1: not merged), sp_coords[3 n_sp],
 * sp_member_off[n_sp + 1] + sp_members (landmark indices), sp_vis_off
 * [n_sp + 1] + sp_vis[n_vis], edges[2 n_edges], weights[n_edges], the
 * scale (median k-nearest distance).  Any output may be NULL. */
int pm_covis_download(pm_covis* c, int32_t* point_sp, double* sp_coords,
                      int64_t* sp_member_off, int32_t* sp_members,
                      int64_t* sp_vis_off, int32_t* sp_vis, int32_t* edges,
                      int32_t* weights, float* scale);
/* SelectNeighbors (dense_helpers.cc:331-512): camera centres origins
 * [3 n_shots], super-points (centroid, visibility CSR, track count).  Per
 * shot pair: count and score (tracks with the angle at the centroid in
 * (theta_min, theta_max)) and the angle sum (cos rounded to 2^-32, fixed
 * point: independent of order; the reference sums FP64 in a
 * thread-dependent order).  best[i] = i, then up to num_neighbors shots
 * with count > min_point_best, mean angle within [theta_min, theta_max] and
 * score > min_point_best, by score descending (ties: lower index); all[i] =
 * best[i], then the shots with count > min_point_all in index order.
 * Limits: n_shots <= 16384, num_neighbors <= 64. */
int pm_covis_select_neighbors(pm_covis* c, int n_shots, const double* origins,
                              int64_t n_sp, const double* sp_coords,
                              const int64_t* sp_vis_off, const int32_t* sp_vis,
                              const int32_t* sp_ntracks, int num_neighbors,
                              int min_point_best, int min_point_all,
                              double theta_min_deg, double theta_max_deg,
                              int64_t* n_best, int64_t* n_all);
/* best_off[n_shots + 1] + best[n_best], all_off[n_shots + 1] + all[n_all] */
int pm_covis_download_neighbors(pm_covis* c, int64_t* best_off, int32_t* best,
                                int64_t* all_off, int32_t* all);
int pm_covis_last_ms(pm_covis* c, float* ms);

/* ---- LZ4 block codec of the depthmap files (dataset.py:1080-1149) ----- */

/* Host-side helpers with no counterpart in the reference, which calls the
 * lz4 Python module: the block format of the LZ4 specification.  Decoding
 * appends to out[out_pos ..] (matches may reach back into earlier output,
 * as linked blocks do) and returns the new end in *new_pos; corrupt input
 * or too small an output gives PM_ERR_INVALID.  Compression is a greedy
 * single-probe hash matcher (independent block). */
int pm_lz4_block_decompress(const uint8_t* src, int64_t src_size, uint8_t* out,
                            int64_t out_pos, int64_t out_cap, int64_t* new_pos);
int pm_lz4_block_compress(const uint8_t* src, int64_t n, uint8_t* dst,
                          int64_t cap, int64_t* written);

/* ---- Batched estimation ---------------------------------------------- */

/* Many reference views ("jobs") of one pool of views, run together: each
 * launch walks the tiles of every job, and a whole run (state clears,
 * random init, 2 x iterations red/black launches, output packing) is
 * captured in one hipGraph that is replayed while the configuration stays
 * the same.  Job k computes exactly what one estimator computes with
 * add_view(ref), add_view(neighbours[0..n-1]), the same settings and
 * set_seed(seed + k).  No counterpart in the reference, whose callers run
 * one DepthmapEstimator per reference view in a thread pool
 * (opensfm/dense.py:187-241). */
typedef struct pm_batch pm_batch;

/* lane layouts of the red/black launches (lanes across x down; each lane
 * takes one pixel of the colour in a pair of columns) */
#define PM_TILE_8x8 0
#define PM_TILE_16x16 1
#define PM_TILE_32x8 2

int pm_batch_create(int device, pm_batch** out);
int pm_batch_destroy(pm_batch* b);
/* Adds a camera (world->camera K, R, t as in pm_estimator_add_view) and its
 * image and mask to the pool; *view_id receives its index. */
int pm_batch_add_view(pm_batch* b, const double* K, const double* R,
                      const double* t, const uint8_t* image,
                      const uint8_t* mask, int width, int height, int stride,
                      int* view_id);
/* Adds a job: reference pool view `ref_view` against `n_neighbours` pool
 * views, depth range [min_depth, max_depth]; *job_id receives its index. */
int pm_batch_add_job(pm_batch* b, int ref_view, const int* neighbours,
                     int n_neighbours, double min_depth, double max_depth,
                     int* job_id);
/* Settings shared by every job (as the pm_estimator_set_* of the same
 * name; defaults equal the estimator's). */
int pm_batch_set_patchmatch_iterations(pm_batch* b, int n);
int pm_batch_set_patch_size(pm_batch* b, int size);
int pm_batch_set_min_patch_sd(pm_batch* b, float sd);
int pm_batch_set_seed(pm_batch* b, uint64_t seed);
int pm_batch_set_exploration(pm_batch* b, int n);
/* pm_estimator_set_pyramid for every job (default 0 = off): the half level
 * is a batch of the pool views box-averaged by 2 with the same jobs, run
 * first in the same launch sequence (and graph); job k still equals the
 * estimator with seed + k bit for bit. */
int pm_batch_set_pyramid(pm_batch* b, int half_iterations);
/* pm_estimator_set_pyramid_levels for every job (default 2). */
int pm_batch_set_pyramid_levels(pm_batch* b, int levels);
/* pm_estimator_set_pyramid_schedule for every job. */
int pm_batch_set_pyramid_schedule(pm_batch* b, const int32_t* iterations, int n);
/* pm_estimator_set_pyramid_consistency for every job (default 1). */
int pm_batch_set_pyramid_consistency(pm_batch* b, int enable);
/* pm_estimator_set_consistency for every job (default 1). */
int pm_batch_set_consistency(pm_batch* b, int enable);
int pm_batch_set_post_process(pm_batch* b, int enable);
int pm_batch_set_precision(pm_batch* b, int precision);
/* Learned initialisation per job (not in the reference; the batched form of
 * pm_estimator_set_init_map / _set_init_search / _set_init_spread, with the
 * same meaning and defaults): job `job` of a batch with a map equals a
 * pm_estimator given the same map, seed + job and settings.  The map must
 * stay valid through every run made while it is set; NULL clears it. */
/* pm_estimator_initializer_input for job `job` (its reference view and
 * neighbours from the pool, the pool's views box-averaged by `scale` once and
 * shared by every job); the batch's patch size (5 or 7). */
int pm_batch_initializer_input(pm_batch* b, int job, int scale, int nplanes, void* out,
                               int out_w, int out_h);
int pm_batch_set_init_map(pm_batch* b, int job, const float* map, int w, int h, int stride,
                          int scale);
int pm_batch_set_init_search(pm_batch* b, int n);
int pm_batch_set_init_spread(pm_batch* b, float sigma_log_depth, float sigma_slope);
/* Lane layout of the red/black launches (default PM_TILE_16x16).  The maps
 * do not depend on it. */
int pm_batch_set_tile(pm_batch* b, int tile);
/* 1 (default): run through the captured hipGraph; 0: launch eagerly. */
int pm_batch_set_graph(pm_batch* b, int enable);
/* Times init + one red/black iteration on the first jobs (about 8 Mpx) for
 * every tile and every patch size in patch_sizes[0..n-1] (n = 0: the
 * current size only) and keeps the fastest pair; *patch_out and *tile_out
 * receive it.  Patch size changes the maps: list only sizes acceptable to
 * the caller. */
int pm_batch_autotune(pm_batch* b, int mode, const int* patch_sizes, int n,
                      int* patch_out, int* tile_out);
/* Runs every job (PM_MODE_PATCHMATCH or PM_MODE_PATCHMATCH_SAMPLE); the
 * maps stay in HBM. */
int pm_batch_run(pm_batch* b, int mode);
/* Maps of job `job` of the last run, layouts as pm_estimator_compute. */
int pm_batch_download(pm_batch* b, int job, float* depth, float* plane,
                      float* score, int32_t* nghbr);
/* DepthmapCleaner::Clean (depthmap.cc:524-555) of every job's depth against
 * the depths of its neighbour views that are themselves the reference view
 * of a job, on the device. */
int pm_batch_clean(pm_batch* b, float same_depth_threshold,
                   int min_consistent_views);
int pm_batch_download_clean(pm_batch* b, int job, float* depth);
/* Job `job`'s result in the layout of the cluster estimator's DepthmapResult
 * (estimator.h:33-38, depthmap_bind.h:53-73) with its readback rules
 * (estimator.cc:960-993): depth H*W (0 where the plane is not finite or the
 * depth lies outside the job's [min_depth, max_depth]), unit camera-frame
 * normal H*W*3 (plane / |plane|, 0 without depth), cost H*W (1 - NCC score
 * wherever finite, whatever the depth; 0 otherwise), confidence H*W
 * (1 / (1 + exp((cost - 0.5) * 5)) where depth > 0, else 0;
 * estimator.cc:1022-1040). */
int pm_batch_download_cluster(pm_batch* b, int job, float* depth, float* normal,
                              float* cost, float* confidence);
/* Last run: scorings, launches = red/black launches, total_ms = dominant_ms
 * = the whole run on the device. */
int pm_batch_stats(pm_batch* b, pm_stats* out);
/* Device time of the last pm_batch_clean, milliseconds. */
int pm_batch_clean_ms(pm_batch* b, float* ms);

/* ---- Initialiser network kernels ------------------------------------- */

/* BF16 MFMA kernels of the learned PatchMatch initialiser (a U-Net over the
 * quarter-resolution reference image and a coarse plane-sweep cost
 * volume), forward and backward.  No counterpart in the reference.  All
 * tensor pointers are device pointers; `stream` is a hipStream_t (NULL: the
 * default stream).  Activations are NHWC BF16 with a channel stride (xs,
 * ys, as) and offset (yo, ao), so skip connections share concatenation
 * buffers; gradients are FP32; weights are [cout][kp] BF16 with
 * k = tap * cin + ci, tap = 3 (dy + 1) + (dx + 1), zero padded to
 * kp = pm_nn_conv_kpad(cin).  Channel counts, strides and offsets of
 * convolution inputs are multiples of 8. */
int pm_nn_conv_kpad(int cin);
/* 3x3 convolution, zero padding 1: y = conv(x, w) + bias (BF16 or FP32,
 * FP32 optionally accumulated), act = leaky_relu_0.1(y) in BF16 (either
 * may be NULL).  x is BF16, or FP32 (x_f32) rounded to BF16 on load -- the
 * input gradient is this call on the output gradient with
 * pm_nn_weight_flip weights. */
int pm_nn_conv3x3(const void* x, int x_f32, int npix, int H, int W, int cin,
                  int xs, const void* w, int cout, int kp, const float* bias,
                  void* y, int y_f32, int ys, int yo, int accumulate, void* act,
                  int as, int ao, void* stream);
/* dw[cout][kp] += d(conv)/dw for output gradient dz (FP32 [npix][dzs]);
 * db[cout] += sum over pixels (db may be NULL). */
int pm_nn_conv3x3_wgrad(const void* x, int npix, int H, int W, int cin, int xs,
                        const float* dz, int cout, int dzs, float* dw, int kp,
                        float* db, void* stream);
/* wt[ci][tap * cout + co] = w[co][(8 - tap) * cin + ci], kpt =
 * pm_nn_conv_kpad(cout): the weights of the input-gradient convolution. */
int pm_nn_weight_flip(const void* w, int cout, int cin, int kp, void* wt,
                      int kpt, void* stream);
/* dz = da * leaky_relu_0.1'(.), the slope taken from the activation a. */
int pm_nn_lrelu_backward(const float* da, int das, int dao, const void* a,
                         int as, int ao, float* dz, int npix, int c,
                         void* stream);
/* 2x2 max pool (even H, W) into contiguous y; idx records the winner. */
int pm_nn_maxpool2(const void* x, int xs, int xo, int N, int H, int W, int c,
                   void* y, uint8_t* idx, void* stream);
/* dx[winner] += dy. */
int pm_nn_maxpool2_backward(const float* dy, const uint8_t* idx, int N, int H,
                            int W, int c, float* dx, int dxs, int dxo,
                            void* stream);
/* Nearest 2x upsampling of contiguous x ([N][h][w][c]) into y. */
int pm_nn_upsample2(const void* x, int N, int h, int w, int c, void* y, int ys,
                    int yo, void* stream);
/* dx (contiguous, low resolution) = sum of the 4 children of dy. */
int pm_nn_upsample2_backward(const float* dy, int dys, int dyo, int N, int h,
                             int w, int c, float* dx, void* stream);
/* Initialiser loss of out (FP32 [npix][os], channels t, sx, sy, c) against
 * target (FP32 [npix][4]: t*, sx*, sy*, valid): Huber(0.02) on t and
 * w_normal * Huber(0.5) on the slopes over valid pixels (mean over
 * n_valid), w_conf * BCE(sigmoid(c), valid) (mean over npix).  Writes
 * dout and adds the three terms to sums[0..2]. */
int pm_nn_loss(const float* out, int os, const float* target, int npix,
               int n_valid, float w_normal, float w_conf, float* dout,
               float* sums, void* stream);
/* Adam step `step` (1-based, bias-corrected) on FP32 w, refreshing the BF16
 * copy w16 (may be NULL). */
int pm_nn_adam(float* w, void* w16, float* m, float* v, const float* g, int n,
               float lr, float beta1, float beta2, float eps, int step,
               void* stream);
int pm_nn_to_bf16(const float* x, void* y, int n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OPENSFM_AMD_H_ */
