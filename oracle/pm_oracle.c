/*
 * pm_oracle.c -- CPU restatement of OpenSfM's legacy dense depthmap path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine in
 * opensfm_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product never links or calls it.
 *
 * It restates, function by function, opensfm/src/lib/dense/src/depthmap.cc
 * (and the declarations in opensfm/src/lib/dense/depthmap.h) in plain C with
 * the reference's arithmetic: FP64 where the reference uses cv::Matx33d /
 * cv::Vec3d, FP32 where it uses float / cv::Vec3f / cv::Matx33f, the same
 * operation order, and no FMA contraction (the Makefile builds with
 * -ffp-contract=off, as the reference's x86-64 build has no -march flag and
 * therefore no FMA).
 *
 * Third-party arithmetic restated here (OpenCV 4.13.0, conda-linux-64.lock,
 * not vendored in the reference):
 *   - cv::Matx products: accumulate from 0 in k order        (mat3_mul)
 *   - cv::Matx33d::inv(): 3x3 cofactor / determinant inverse (inv3)
 *   - Vec3f / float: multiply by the float reciprocal        (plane_from_depth_and_normal)
 *   - cv::medianBlur(CV_32F, 5): exact 5x5 median, replicate border (median5)
 * exp/sqrt/fabs on float arguments resolve to the float overloads (expf,
 * sqrtf, fabsf); with the double overloads the results differ by < 1 ulp.
 *
 * Deviations from the reference, all on inputs where the reference is
 * undefined or non-reproducible:
 *   - the random source is a seeded splitmix64 stream instead of rand() and
 *     std::mt19937(std::random_device) (depthmap.cc:127-132, 141);
 *   - LinearInterpolation treats a NaN coordinate as outside the image (the
 *     reference casts NaN to int, which is undefined behaviour).
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_PATCH 15

/* ------------------------------------------------------------------ */
/* 3x3 FP64 helpers in OpenCV's accumulation order                     */
/* ------------------------------------------------------------------ */

/* C = A * B, row-major, s = 0; s += a(i,k) * b(k,j)  (cv::Matx_MatMulOp) */
static void mat3_mul(const double* A, const double* B, double* C) {
  double T[9];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[k * 3 + j];
      T[i * 3 + j] = s;
    }
  }
  memcpy(C, T, sizeof(T));
}

/* y = A * x  (cv::Matx33d * cv::Vec3d) */
static void mat3_vec(const double* A, const double* x, double* y) {
  double t[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0;
    for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * x[k];
    t[i] = s;
  }
  memcpy(y, t, sizeof(t));
}

static void mat3_transpose(const double* A, double* B) {
  double T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T[j * 3 + i] = A[i * 3 + j];
  memcpy(B, T, sizeof(T));
}

/* cv::Matx33d::inv() with DECOMP_LU: for 3x3 OpenCV uses the determinant and
 * cofactors (cv::internal::Matx_FastInvOp<_Tp, 3>).  Returns 0 if singular. */
int orc_inv3(const double* a, double* b) {
#define A(r, c) a[(r) * 3 + (c)]
  double d = A(0, 0) * (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) -
             A(0, 1) * (A(1, 0) * A(2, 2) - A(2, 0) * A(1, 2)) +
             A(0, 2) * (A(1, 0) * A(2, 1) - A(2, 0) * A(1, 1));
  if (d == 0) return 0;
  d = 1 / d;
  double t[9];
  t[0] = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) * d;
  t[1] = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) * d;
  t[2] = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * d;
  t[3] = (A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2)) * d;
  t[4] = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * d;
  t[5] = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) * d;
  t[6] = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) * d;
  t[7] = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) * d;
  t[8] = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) * d;
#undef A
  memcpy(b, t, sizeof(t));
  return 1;
}

/* ------------------------------------------------------------------ */
/* Pixel-level primitives                                               */
/* ------------------------------------------------------------------ */

/* LinearInterpolation<unsigned char>  depthmap.cc:14-30 */
float orc_linear_interp_u8(const uint8_t* img, int width, int height, float y,
                           float x) {
  if (!(x >= 0.0f) || x >= width - 1 || !(y >= 0.0f) || y >= height - 1) {
    return 0.0f;
  }
  int ix = (int)x;
  int iy = (int)y;
  float dx = x - ix;
  float dy = y - iy;
  float im00 = img[iy * width + ix];
  float im01 = img[iy * width + ix + 1];
  float im10 = img[(iy + 1) * width + ix];
  float im11 = img[(iy + 1) * width + ix + 1];
  float im0 = (1 - dx) * im00 + dx * im01;
  float im1 = (1 - dx) * im10 + dx * im11;
  return (1 - dy) * im0 + dy * im1;
}

/* LinearInterpolation<float>  depthmap.cc:14-30 (cleaner) */
float orc_linear_interp_f32(const float* img, int width, int height, float y,
                            float x) {
  if (!(x >= 0.0f) || x >= width - 1 || !(y >= 0.0f) || y >= height - 1) {
    return 0.0f;
  }
  int ix = (int)x;
  int iy = (int)y;
  float dx = x - ix;
  float dy = y - iy;
  float im00 = img[iy * width + ix];
  float im01 = img[iy * width + ix + 1];
  float im10 = img[(iy + 1) * width + ix];
  float im11 = img[(iy + 1) * width + ix + 1];
  float im0 = (1 - dx) * im00 + dx * im01;
  float im1 = (1 - dx) * im10 + dx * im11;
  return (1 - dy) * im0 + dy * im1;
}

/* Variance  depthmap.cc:32-44 (two-pass, population) */
float orc_variance(const float* x, int n) {
  float sum = 0;
  for (int i = 0; i < n; ++i) sum += x[i];
  float mean = sum / n;
  float sum2 = 0;
  for (int i = 0; i < n; ++i) sum2 += (x[i] - mean) * (x[i] - mean);
  return sum2 / n;
}

/* NCCEstimator  depthmap.cc:46-74 */
typedef struct {
  float sumx, sumy, sumxx, sumyy, sumxy, sumw;
} ncc_t;

static void ncc_push(ncc_t* e, float x, float y, float w) {
  e->sumx += w * x;
  e->sumy += w * y;
  e->sumxx += w * x * x;
  e->sumyy += w * y * y;
  e->sumxy += w * x * y;
  e->sumw += w;
}

static float ncc_get(const ncc_t* e) {
  if (e->sumw == 0.0) return -1;
  float meanx = e->sumx / e->sumw;
  float meany = e->sumy / e->sumw;
  float meanxx = e->sumxx / e->sumw;
  float meanyy = e->sumyy / e->sumw;
  float meanxy = e->sumxy / e->sumw;
  float varx = meanxx - meanx * meanx;
  float vary = meanyy - meany * meany;
  if (varx < 0.1 || vary < 0.1) return -1;
  return (meanxy - meanx * meany) / sqrtf(varx * vary);
}

/* NCC of explicit samples, for tests: x, y, w arrays of length n. */
float orc_ncc(const float* x, const float* y, const float* w, int n) {
  ncc_t e = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < n; ++k) ncc_push(&e, x[k], y[k], w[k]);
  return ncc_get(&e);
}

/* DepthmapEstimator::BilateralWeight  depthmap.cc:480-487 */
float orc_bilateral_weight(float dcolor, float dx, float dy) {
  const float dcolor_sigma = 50.0f;
  const float dx_sigma = 5.0f;
  const float dcolor_factor = 1.0f / (2 * dcolor_sigma * dcolor_sigma);
  const float dx_factor = 1.0f / (2 * dx_sigma * dx_sigma);
  return expf(-dcolor * dcolor * dcolor_factor -
              (dx * dx + dy * dy) * dx_factor);
}

/* PlaneInducedHomographyBaked  depthmap.cc:96-102:
 * H = K2 * (Q2 + a2 * v^T) * K1inv in FP64, cast to FP32. */
void orc_homography_baked(const double* K1inv, const double* Q2,
                          const double* a2, const double* K2, const float* vf,
                          float* H) {
  double v[3] = {vf[0], vf[1], vf[2]};
  double S[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      s += a2[i] * v[j];
      S[i * 3 + j] = Q2[i * 3 + j] + s;
    }
  double T[9];
  mat3_mul(K2, S, T);
  mat3_mul(T, K1inv, T);
  for (int k = 0; k < 9; ++k) H[k] = (float)T[k];
}

/* PlaneFromDepthAndNormal  depthmap.cc:120-125 */
void orc_plane_from_depth_and_normal(float x, float y, const double* K,
                                     float depth, const float* normal,
                                     float* plane) {
  double Kinv[9];
  orc_inv3(K, Kinv);
  double S[9];
  for (int k = 0; k < 9; ++k) S[k] = Kinv[k] * depth;
  double q[3] = {x, y, 1};
  double pd[3];
  mat3_vec(S, q, pd);
  float point[3] = {(float)pd[0], (float)pd[1], (float)pd[2]};
  float s = 0;
  for (int k = 0; k < 3; ++k) s += normal[k] * point[k];
  float denom = -s;
  float m = (1e-6f < denom) ? denom : 1e-6f; /* std::max(1e-6f, denom) */
  float inv = 1.f / m;
  for (int k = 0; k < 3; ++k) plane[k] = normal[k] * inv;
}

/* DepthOfPlaneBackprojection  depthmap.cc:114-118 */
float orc_depth_of_plane_backprojection(double x, double y, const double* K,
                                        const float* planef) {
  double Kinv[9];
  orc_inv3(K, Kinv);
  double p[3] = {planef[0], planef[1], planef[2]};
  double r[3];
  for (int j = 0; j < 3; ++j) {
    double s = 0;
    for (int k = 0; k < 3; ++k) s += p[k] * Kinv[k * 3 + j];
    r[j] = s;
  }
  double q[3] = {x, y, 1};
  double s = 0;
  for (int k = 0; k < 3; ++k) s += r[k] * q[k];
  float denom = (float)(-s);
  float m = (1e-6f < denom) ? denom : 1e-6f;
  return 1.0f / m;
}

/* 5x5 median with replicate border (cv::medianBlur(src, dst, 5), CV_32F) */
static int cmp_float(const void* a, const void* b) {
  float x = *(const float*)a, y = *(const float*)b;
  return (x > y) - (x < y);
}

void orc_median5(const float* src, float* dst, int width, int height) {
  float win[25];
  for (int i = 0; i < height; ++i) {
    for (int j = 0; j < width; ++j) {
      int c = 0;
      for (int u = -2; u <= 2; ++u) {
        int ii = i + u;
        ii = ii < 0 ? 0 : (ii >= height ? height - 1 : ii);
        for (int v = -2; v <= 2; ++v) {
          int jj = j + v;
          jj = jj < 0 ? 0 : (jj >= width ? width - 1 : jj);
          win[c++] = src[ii * width + jj];
        }
      }
      qsort(win, 25, sizeof(float), cmp_float);
      dst[i * width + j] = win[12];
    }
  }
}

/* ------------------------------------------------------------------ */
/* Seeded random source (replaces rand() / std::mt19937)               */
/* ------------------------------------------------------------------ */

static uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* UniformRand  depthmap.cc:127-132, with a 31-bit integer like glibc rand() */
static float uniform_rand(uint64_t* s, float a, float b) {
  int r = (int)(splitmix64(s) >> 33);
  return a + (b - a) * (float)r / (float)2147483647;
}

/* std::normal_distribution<float>(0, 1): Box-Muller on two uniforms. */
static float unit_normal(uint64_t* s) {
  double u1 = ((splitmix64(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
  double u2 = (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
  return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

/* std::uniform_int_distribution<int>(a, b) */
static int uniform_int(uint64_t* s, int a, int b) {
  uint64_t span = (uint64_t)(b - a) + 1;
  return a + (int)(splitmix64(s) % span);
}

/* ------------------------------------------------------------------ */
/* DepthmapEstimator  depthmap.h:57-111, depthmap.cc:134-502            */
/* ------------------------------------------------------------------ */

typedef struct {
  int n, cap;
  double *K, *R, *t, *Kinv, *Q, *a; /* 9, 9, 3, 9, 9, 3 per view */
  uint8_t** img;
  uint8_t** mask;
  int *w, *h;
  int patch_size;
  double min_depth, max_depth;
  int num_depth_planes;
  int iterations;
  float min_patch_variance;
  uint64_t seed;     /* each compute restarts the random stream from here */
  uint64_t rng;
  int64_t scorings; /* ComputePlaneImageScore calls of the last compute */
  /* result maps of the running compute */
  float *depth, *plane, *score;
  int32_t* nghbr;
} orc_est;

void* orc_create(void) {
  orc_est* e = (orc_est*)calloc(1, sizeof(orc_est));
  e->patch_size = 7;
  e->num_depth_planes = 50;
  e->iterations = 3;
  e->min_patch_variance = 5 * 5;
  e->seed = 0x5EEDull;
  return e;
}

void orc_destroy(void* p) {
  orc_est* e = (orc_est*)p;
  if (!e) return;
  for (int v = 0; v < e->n; ++v) {
    free(e->img[v]);
    free(e->mask[v]);
  }
  free(e->K); free(e->R); free(e->t); free(e->Kinv); free(e->Q); free(e->a);
  free(e->img); free(e->mask); free(e->w); free(e->h);
  free(e);
}

/* DepthmapEstimator::AddView  depthmap.cc:146-162 */
int orc_add_view(void* p, const double* K, const double* R, const double* t,
                 const uint8_t* image, const uint8_t* mask, int width,
                 int height) {
  orc_est* e = (orc_est*)p;
  if (e->n == e->cap) {
    int c = e->cap ? 2 * e->cap : 8;
    e->K = realloc(e->K, sizeof(double) * 9 * c);
    e->R = realloc(e->R, sizeof(double) * 9 * c);
    e->t = realloc(e->t, sizeof(double) * 3 * c);
    e->Kinv = realloc(e->Kinv, sizeof(double) * 9 * c);
    e->Q = realloc(e->Q, sizeof(double) * 9 * c);
    e->a = realloc(e->a, sizeof(double) * 3 * c);
    e->img = realloc(e->img, sizeof(uint8_t*) * c);
    e->mask = realloc(e->mask, sizeof(uint8_t*) * c);
    e->w = realloc(e->w, sizeof(int) * c);
    e->h = realloc(e->h, sizeof(int) * c);
    e->cap = c;
  }
  int v = e->n;
  memcpy(e->K + 9 * v, K, 9 * sizeof(double));
  memcpy(e->R + 9 * v, R, 9 * sizeof(double));
  memcpy(e->t + 3 * v, t, 3 * sizeof(double));
  orc_inv3(e->K + 9 * v, e->Kinv + 9 * v);
  double R0t[9];
  mat3_transpose(e->R, R0t); /* view 0 is the reference */
  mat3_mul(e->R + 9 * v, R0t, e->Q + 9 * v);
  double qt[3];
  mat3_vec(e->Q + 9 * v, e->t, qt);
  for (int k = 0; k < 3; ++k) e->a[3 * v + k] = qt[k] - e->t[3 * v + k];
  size_t np = (size_t)width * height;
  e->img[v] = (uint8_t*)malloc(np);
  e->mask[v] = (uint8_t*)malloc(np);
  memcpy(e->img[v], image, np);
  memcpy(e->mask[v], mask, np);
  e->w[v] = width;
  e->h[v] = height;
  e->n++;
  return v;
}

void orc_set_depth_range(void* p, double mn, double mx, int planes) {
  orc_est* e = (orc_est*)p;
  e->min_depth = mn;
  e->max_depth = mx;
  e->num_depth_planes = planes;
}
void orc_set_patchmatch_iterations(void* p, int n) { ((orc_est*)p)->iterations = n; }
void orc_set_patch_size(void* p, int s) { ((orc_est*)p)->patch_size = s; }
void orc_set_min_patch_sd(void* p, float sd) { ((orc_est*)p)->min_patch_variance = sd * sd; }
void orc_set_seed(void* p, uint64_t seed) { ((orc_est*)p)->seed = seed; }
int64_t orc_last_scorings(void* p) { return ((orc_est*)p)->scorings; }

/* ComputePlaneImageScore  depthmap.cc:441-478 (local-affine warp) */
static float plane_image_score(orc_est* e, int i, int j, const float* plane,
                               int other) {
  e->scorings++;
  float H[9];
  orc_homography_baked(e->Kinv, e->Q + 9 * other, e->a + 3 * other,
                       e->K + 9 * other, plane, H);
  int hpz = (e->patch_size - 1) / 2;
  float fj = (float)j, fi = (float)i;
  float u = H[0] * fj + H[1] * fi + H[2];
  float v = H[3] * fj + H[4] * fi + H[5];
  float w = H[6] * fj + H[7] * fi + H[8];
  if (w == 0.0) return -1.0f;

  float dfdx_x = (H[0] * w - H[6] * u) / (w * w);
  float dfdx_y = (H[3] * w - H[6] * v) / (w * w);
  float dfdy_x = (H[1] * w - H[7] * u) / (w * w);
  float dfdy_y = (H[4] * w - H[7] * v) / (w * w);

  float Hx0 = u / w;
  float Hy0 = v / w;

  const uint8_t* im0 = e->img[0];
  int W0 = e->w[0];
  float im1_center = im0[i * W0 + j];
  ncc_t ncc = {0, 0, 0, 0, 0, 0};
  for (int dy = -hpz; dy <= hpz; ++dy) {
    for (int dx = -hpz; dx <= hpz; ++dx) {
      float im1 = im0[(i + dy) * W0 + j + dx];
      float x2 = Hx0 + dfdx_x * dx + dfdy_x * dy;
      float y2 = Hy0 + dfdx_y * dx + dfdy_y * dy;
      float im2 = orc_linear_interp_u8(e->img[other], e->w[other],
                                       e->h[other], y2, x2);
      float weight = orc_bilateral_weight(im1 - im1_center, dx, dy);
      ncc_push(&ncc, im1, im2, weight);
    }
  }
  return ncc_get(&ncc);
}

float orc_score(void* p, int i, int j, const float* plane, int other) {
  return plane_image_score((orc_est*)p, i, j, plane, other);
}

/* FP64 twin of ComputePlaneImageScore: the same local-affine algorithm with
 * every quantity (homography, Jacobian, bilinear sample, bilateral weight,
 * moments) in double.  Used to state the error of the FP32/FP16/BF16 paths. */
double orc_score_f64(void* p, int i, int j, const double* v, int other) {
  orc_est* e = (orc_est*)p;
  const double* K1inv = e->Kinv;
  const double* Q2 = e->Q + 9 * other;
  const double* a2 = e->a + 3 * other;
  const double* K2 = e->K + 9 * other;
  double S[9], T[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) S[r * 3 + c] = Q2[r * 3 + c] + a2[r] * v[c];
  mat3_mul(K2, S, T);
  mat3_mul(T, K1inv, T);
  const double* H = T;
  int hpz = (e->patch_size - 1) / 2;
  double u = H[0] * j + H[1] * i + H[2];
  double vv = H[3] * j + H[4] * i + H[5];
  double w = H[6] * j + H[7] * i + H[8];
  if (w == 0.0) return -1.0;
  double dfdx_x = (H[0] * w - H[6] * u) / (w * w);
  double dfdx_y = (H[3] * w - H[6] * vv) / (w * w);
  double dfdy_x = (H[1] * w - H[7] * u) / (w * w);
  double dfdy_y = (H[4] * w - H[7] * vv) / (w * w);
  double Hx0 = u / w, Hy0 = vv / w;
  const uint8_t* im0 = e->img[0];
  const uint8_t* im2p = e->img[other];
  int W0 = e->w[0], W2 = e->w[other], H2 = e->h[other];
  double c = im0[i * W0 + j];
  double sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0, sw = 0;
  for (int dy = -hpz; dy <= hpz; ++dy) {
    for (int dx = -hpz; dx <= hpz; ++dx) {
      double x1 = im0[(i + dy) * W0 + j + dx];
      double x2 = Hx0 + dfdx_x * dx + dfdy_x * dy;
      double y2 = Hy0 + dfdx_y * dx + dfdy_y * dy;
      double y = 0;
      if (x2 >= 0 && x2 < W2 - 1 && y2 >= 0 && y2 < H2 - 1) {
        int ix = (int)x2, iy = (int)y2;
        double fx = x2 - ix, fy = y2 - iy;
        double a00 = im2p[iy * W2 + ix], a01 = im2p[iy * W2 + ix + 1];
        double a10 = im2p[(iy + 1) * W2 + ix], a11 = im2p[(iy + 1) * W2 + ix + 1];
        y = (1 - fy) * ((1 - fx) * a00 + fx * a01) + fy * ((1 - fx) * a10 + fx * a11);
      }
      double dc = x1 - c;
      double wgt = exp(-dc * dc / 5000.0 - (double)(dx * dx + dy * dy) / 50.0);
      sx += wgt * x1; sy += wgt * y; sxx += wgt * x1 * x1; syy += wgt * y * y;
      sxy += wgt * x1 * y; sw += wgt;
    }
  }
  if (sw == 0.0) return -1.0;
  double mx = sx / sw, my = sy / sw;
  double vx = sxx / sw - mx * mx, vy = syy / sw - my * my;
  if (vx < 0.1 || vy < 0.1) return -1.0;
  return (sxy / sw - mx * my) / sqrt(vx * vy);
}

/* ComputePlaneScore  depthmap.cc:408-419 */
static void plane_score(orc_est* e, int i, int j, const float* plane,
                        float* score, int* nghbr) {
  *score = -1.0f;
  *nghbr = 0;
  for (int other = 1; other < e->n; ++other) {
    float s = plane_image_score(e, i, j, plane, other);
    if (s > *score) {
      *score = s;
      *nghbr = other;
    }
  }
}

/* AssignPixel  depthmap.cc:398-406 */
static void assign_pixel(orc_est* e, int i, int j, float depth,
                         const float* plane, float score, int nghbr) {
  int W = e->w[0];
  size_t k = (size_t)i * W + j;
  e->depth[k] = depth;
  e->plane[3 * k + 0] = plane[0];
  e->plane[3 * k + 1] = plane[1];
  e->plane[3 * k + 2] = plane[2];
  e->score[k] = score;
  e->nghbr[k] = nghbr;
}

/* CheckPlaneCandidate  depthmap.cc:376-386 */
static void check_plane_candidate(orc_est* e, int i, int j,
                                  const float* plane) {
  float score;
  int nghbr;
  plane_score(e, i, j, plane, &score, &nghbr);
  if (score > e->score[(size_t)i * e->w[0] + j]) {
    float depth = orc_depth_of_plane_backprojection(j, i, e->K, plane);
    assign_pixel(e, i, j, depth, plane, score, nghbr);
  }
}

/* CheckPlaneImageCandidate  depthmap.cc:388-396 */
static void check_plane_image_candidate(orc_est* e, int i, int j,
                                        const float* plane, int nghbr) {
  float score = plane_image_score(e, i, j, plane, nghbr);
  if (score > e->score[(size_t)i * e->w[0] + j]) {
    float depth = orc_depth_of_plane_backprojection(j, i, e->K, plane);
    assign_pixel(e, i, j, depth, plane, score, nghbr);
  }
}

/* ComputeBruteForce  depthmap.cc:184-200 */
static void brute_force(orc_est* e) {
  int hpz = (e->patch_size - 1) / 2;
  int H = e->h[0], W = e->w[0];
  for (int i = hpz; i < H - hpz; ++i) {
    for (int j = hpz; j < W - hpz; ++j) {
      for (int d = 0; d < e->num_depth_planes; ++d) {
        float depth = (float)(1 / (1 / e->min_depth +
                                   d * (1 / e->max_depth - 1 / e->min_depth) /
                                       (e->num_depth_planes - 1)));
        float normal[3] = {0, 0, -1};
        float plane[3];
        orc_plane_from_depth_and_normal((float)j, (float)i, e->K, depth,
                                        normal, plane);
        check_plane_candidate(e, i, j, plane);
      }
    }
  }
}

/* RandomInitialization  depthmap.cc:237-256 */
static void random_initialization(orc_est* e, int sample) {
  int hpz = (e->patch_size - 1) / 2;
  int H = e->h[0], W = e->w[0];
  float lmin = (float)log(e->min_depth), lmax = (float)log(e->max_depth);
  for (int i = hpz; i < H - hpz; ++i) {
    for (int j = hpz; j < W - hpz; ++j) {
      float depth = expf(uniform_rand(&e->rng, lmin, lmax));
      float n0 = uniform_rand(&e->rng, -1, 1);
      float n1 = uniform_rand(&e->rng, -1, 1);
      float normal[3] = {n0, n1, -1};
      float plane[3];
      orc_plane_from_depth_and_normal((float)j, (float)i, e->K, depth, normal,
                                      plane);
      int nghbr;
      float score;
      if (sample) {
        nghbr = uniform_int(&e->rng, e->n > 1 ? 1 : 0, e->n > 1 ? e->n - 1 : 0);
        score = plane_image_score(e, i, j, plane, nghbr);
      } else {
        plane_score(e, i, j, plane, &score, &nghbr);
      }
      assign_pixel(e, i, j, depth, plane, score, nghbr);
    }
  }
}

/* PatchVariance  depthmap.cc:271-281 */
static float patch_variance(orc_est* e, int i, int j) {
  float patch[ORC_MAX_PATCH * ORC_MAX_PATCH];
  int hpz = (e->patch_size - 1) / 2;
  int c = 0;
  int W = e->w[0];
  for (int u = -hpz; u <= hpz; ++u)
    for (int v = -hpz; v <= hpz; ++v) patch[c++] = e->img[0][(i + u) * W + j + v];
  return orc_variance(patch, e->patch_size * e->patch_size);
}

/* ComputeIgnoreMask  depthmap.cc:258-269 */
static void compute_ignore_mask(orc_est* e) {
  int hpz = (e->patch_size - 1) / 2;
  int H = e->h[0], W = e->w[0];
  const float zero[3] = {0, 0, 0};
  for (int i = hpz; i < H - hpz; ++i) {
    for (int j = hpz; j < W - hpz; ++j) {
      int masked = e->mask[0][i * W + j] == 0;
      int low_variance = patch_variance(e, i, j) < e->min_patch_variance;
      if (masked || low_variance) assign_pixel(e, i, j, 0.0f, zero, 0.0f, 0);
    }
  }
}

/* PatchMatchUpdatePixel  depthmap.cc:305-374 */
static void update_pixel(orc_est* e, int i, int j, const int adj[2][2],
                         int sample) {
  int W = e->w[0];
  size_t k0 = (size_t)i * W + j;
  if (e->depth[k0] == 0.0f) return;

  for (int k = 0; k < 2; ++k) {
    int ia = i + adj[k][0], ja = j + adj[k][1];
    size_t ka = (size_t)ia * W + ja;
    if (e->depth[ka] == 0.0f) continue;
    float plane[3] = {e->plane[3 * ka], e->plane[3 * ka + 1], e->plane[3 * ka + 2]};
    if (sample) {
      check_plane_image_candidate(e, i, j, plane, e->nghbr[ka]);
    } else {
      check_plane_candidate(e, i, j, plane);
    }
  }

  float depth_range = 0.02f;
  float normal_range = 0.5f;
  int current_nghbr = e->nghbr[k0];
  for (int k = 0; k < 6; ++k) {
    float current_depth = e->depth[k0];
    float depth = current_depth * expf(depth_range * unit_normal(&e->rng));
    float cp[3] = {e->plane[3 * k0], e->plane[3 * k0 + 1], e->plane[3 * k0 + 2]};
    if (cp[2] == 0.0) continue;
    float g0 = unit_normal(&e->rng);
    float g1 = unit_normal(&e->rng);
    float normal[3] = {-cp[0] / cp[2] + normal_range * g0,
                       -cp[1] / cp[2] + normal_range * g1, -1.0f};
    float plane[3];
    orc_plane_from_depth_and_normal((float)j, (float)i, e->K, depth, normal,
                                    plane);
    if (sample) {
      check_plane_image_candidate(e, i, j, plane, current_nghbr);
    } else {
      check_plane_candidate(e, i, j, plane);
    }
    depth_range *= 0.3f;
    normal_range *= 0.8f;
  }

  if (!sample || e->n <= 2) return;

  int other = uniform_int(&e->rng, 1, e->n - 1);
  while (other == current_nghbr) other = uniform_int(&e->rng, 1, e->n - 1);
  float plane[3] = {e->plane[3 * k0], e->plane[3 * k0 + 1], e->plane[3 * k0 + 2]};
  check_plane_image_candidate(e, i, j, plane, other);
}

/* PatchMatchForwardPass / PatchMatchBackwardPass  depthmap.cc:283-303 */
static void forward_pass(orc_est* e, int sample) {
  static const int adj[2][2] = {{-1, 0}, {0, -1}};
  int hpz = (e->patch_size - 1) / 2;
  for (int i = hpz; i < e->h[0] - hpz; ++i)
    for (int j = hpz; j < e->w[0] - hpz; ++j) update_pixel(e, i, j, adj, sample);
}

static void backward_pass(orc_est* e, int sample) {
  static const int adj[2][2] = {{0, 1}, {1, 0}};
  int hpz = (e->patch_size - 1) / 2;
  for (int i = e->h[0] - hpz - 1; i >= hpz; --i)
    for (int j = e->w[0] - hpz - 1; j >= hpz; --j) update_pixel(e, i, j, adj, sample);
}

/* PostProcess  depthmap.cc:489-502 */
static void post_process(orc_est* e) {
  int H = e->h[0], W = e->w[0];
  float* filtered = (float*)malloc(sizeof(float) * (size_t)H * W);
  orc_median5(e->depth, filtered, W, H);
  for (size_t k = 0; k < (size_t)H * W; ++k) {
    float d = e->depth[k];
    float m = filtered[k];
    if (d == 0.0 || fabsf(d - m) / d > 0.05) e->depth[k] = 0;
  }
  free(filtered);
}

/* mode 0: ComputeBruteForce (depthmap.cc:184-200)
 * mode 1: ComputePatchMatch (depthmap.cc:202-213)
 * mode 2: ComputePatchMatchSample (depthmap.cc:215-227)
 * Output buffers (AssignMatrices, depthmap.cc:229-235): depth H*W, plane
 * H*W*3, score H*W, nghbr H*W, all zero-initialised here.
 * stop_after_iterations < 0 runs everything; otherwise PatchMatch stops
 * after that many iterations and skips PostProcess (used by tests that
 * check the per-iteration score invariant). */
int orc_compute(void* p, int mode, float* depth, float* plane, float* score,
                int32_t* nghbr, int stop_after_iterations) {
  orc_est* e = (orc_est*)p;
  if (e->n < 1) return -1;
  size_t np = (size_t)e->w[0] * e->h[0];
  memset(depth, 0, np * sizeof(float));
  memset(plane, 0, 3 * np * sizeof(float));
  memset(score, 0, np * sizeof(float));
  memset(nghbr, 0, np * sizeof(int32_t));
  e->depth = depth;
  e->plane = plane;
  e->score = score;
  e->nghbr = nghbr;
  e->scorings = 0;
  e->rng = e->seed;
  if (mode == 0) {
    brute_force(e);
  } else {
    int sample = mode == 2;
    random_initialization(e, sample);
    compute_ignore_mask(e);
    int iters = e->iterations;
    if (stop_after_iterations >= 0 && stop_after_iterations < iters)
      iters = stop_after_iterations;
    for (int it = 0; it < iters; ++it) {
      forward_pass(e, sample);
      backward_pass(e, sample);
    }
    if (stop_after_iterations < 0) post_process(e);
  }
  e->depth = e->plane = e->score = NULL;
  e->nghbr = NULL;
  return 0;
}

/* ------------------------------------------------------------------ */
/* DepthmapCleaner  depthmap.h:113-129, depthmap.cc:504-555              */
/* ------------------------------------------------------------------ */

/* Views are given as arrays: Ks (n*9), Rs (n*9), ts (n*3), depths[n],
 * widths[n], heights[n].  View 0 is the one being cleaned. */
int orc_clean(int n, const double* Ks, const double* Rs, const double* ts,
              const float* const* depths, const int* widths,
              const int* heights, float same_depth_threshold,
              int min_consistent_views, float* clean_depth) {
  if (n < 1) return -1;
  int W = widths[0], H = heights[0];
  double Kinv0[9], R0t[9];
  orc_inv3(Ks, Kinv0);
  mat3_transpose(Rs, R0t);
  for (int i = 0; i < H; ++i) {
    for (int j = 0; j < W; ++j) {
      float depth = depths[0][(size_t)i * W + j];
      /* Backproject  depthmap.cc:109-112: R^T (depth K^-1 q - t) */
      double S[9];
      for (int k = 0; k < 9; ++k) S[k] = Kinv0[k] * (double)depth;
      double q[3] = {(double)j, (double)i, 1};
      double X[3];
      mat3_vec(S, q, X);
      for (int k = 0; k < 3; ++k) X[k] = X[k] - ts[k];
      mat3_vec(R0t, X, X);
      float point[3] = {(float)X[0], (float)X[1], (float)X[2]};
      int consistent_views = 1;
      for (int other = 1; other < n; ++other) {
        /* Project  depthmap.cc:104-107: K (R x + t) */
        double pd[3] = {point[0], point[1], point[2]};
        double y[3];
        mat3_vec(Rs + 9 * other, pd, y);
        for (int k = 0; k < 3; ++k) y[k] = y[k] + ts[3 * other + k];
        mat3_vec(Ks + 9 * other, y, y);
        float rp[3] = {(float)y[0], (float)y[1], (float)y[2]};
        if (rp[2] < 1e-8 || isnan(rp[2])) continue;
        float u = rp[0] / rp[2];
        float v = rp[1] / rp[2];
        float depth_of_point = rp[2];
        float depth_at_reprojection = orc_linear_interp_f32(
            depths[other], widths[other], heights[other], v, u);
        if (fabsf(depth_at_reprojection - depth_of_point) <
            depth_of_point * same_depth_threshold) {
          consistent_views++;
        }
      }
      clean_depth[(size_t)i * W + j] =
          consistent_views >= min_consistent_views ? depths[0][(size_t)i * W + j]
                                                   : 0;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------
 * DepthmapPruner::Prune  depthmap.cc:574-636.  Keeps a pixel of view 0
 * unless some other view sees, at the pixel the point projects to (rounded),
 * a depth beyond (1 - threshold) times the point's depth with a larger
 * "area" -n_z / depth * f (a finer sample of the same surface).  Outputs in
 * row-major pixel order: point (world, float), normal rotated to the world
 * (float Matx33f(R0^T) * n), color (3 u8), label.  Types as the reference:
 * cv::normalize of Vec3f (double norm and scale, float result), FP64
 * Backproject/Project, float area. */
static void normalize3f(const float* v, float* out) {
  double s = 0;
  for (int k = 0; k < 3; ++k) s += (double)v[k] * (double)v[k];
  const double nv = sqrt(s);
  const double a = nv != 0 ? 1.0 / nv : 0.0;
  for (int k = 0; k < 3; ++k) out[k] = (float)((double)v[k] * a);
}

/* The reference rounds x / z + 0.5 to std::int64_t and hands it to
 * IsInsideImage(const cv::Mat&, int, int) and Mat::at(int, int): the low 32
 * bits are used.  A value outside int64 (or NaN) is undefined behaviour in
 * C++; x86's truncating convert returns INT64_MIN there, whose low 32 bits
 * are 0, and that is what is returned here. */
static int ref_pixel_index(double u) {
  if (!(fabs(u) < 9.223372036854775808e18)) return 0;
  return (int)(unsigned int)(unsigned long long)(long long)u;
}

long orc_prune(int n, const double* Ks, const double* Rs, const double* ts,
               const float* const* depths, const float* const* planes,
               const unsigned char* const* colors,
               const unsigned char* const* labels, const int* widths,
               const int* heights, float same_depth_threshold, long capacity,
               float* points, float* normals, unsigned char* out_colors,
               unsigned char* out_labels) {
  if (n < 1) return -1;
  const int W = widths[0], H = heights[0];
  double Kinv0[9], R0t[9];
  orc_inv3(Ks, Kinv0);
  mat3_transpose(Rs, R0t);
  float Rinv[9];
  for (int k = 0; k < 9; ++k) Rinv[k] = (float)R0t[k];
  long count = 0;
  for (int i = 0; i < H; ++i) {
    for (int j = 0; j < W; ++j) {
      const size_t p = (size_t)i * W + j;
      const float depth = depths[0][p];
      if (depth <= 0) continue;
      float normal[3];
      normalize3f(planes[0] + 3 * p, normal);
      const float area = (float)((double)(-normal[2] / depth) * Ks[0]);
      double S[9], X[3];
      for (int k = 0; k < 9; ++k) S[k] = (double)depth * Kinv0[k];
      const double q[3] = {(double)j, (double)i, 1};
      mat3_vec(S, q, X);
      for (int k = 0; k < 3; ++k) X[k] = X[k] - ts[k];
      mat3_vec(R0t, X, X);
      const float point[3] = {(float)X[0], (float)X[1], (float)X[2]};
      int keep = 1;
      for (int o = 1; o < n && keep; ++o) {
        const double pd[3] = {point[0], point[1], point[2]};
        double y[3];
        mat3_vec(Rs + 9 * o, pd, y);
        for (int k = 0; k < 3; ++k) y[k] = y[k] + ts[3 * o + k];
        mat3_vec(Ks + 9 * o, y, y);
        if (y[2] < 1e-8 || isnan(y[2])) continue;
        const int iu = ref_pixel_index(y[0] / y[2] + 0.5);
        const int iv = ref_pixel_index(y[1] / y[2] + 0.5);
        const double depth_of_point = y[2];
        if (!(iv >= 0 && iv < heights[o] && iu >= 0 && iu < widths[o])) continue;
        const size_t po = (size_t)iv * widths[o] + (size_t)iu;
        const float d_o = depths[o][po];
        if ((double)d_o > (double)(1 - same_depth_threshold) * depth_of_point) {
          float no[3];
          normalize3f(planes[o] + 3 * po, no);
          if (d_o == 0.0 || (double)(-no[2] / d_o) * Ks[9 * o] > (double)area) keep = 0;
        }
      }
      if (!keep) continue;
      if (count < capacity) {
        for (int r = 0; r < 3; ++r) {
          float s = 0.0f;
          for (int k = 0; k < 3; ++k) s += Rinv[3 * r + k] * normal[k];
          points[3 * count + r] = point[r];
          normals[3 * count + r] = s;
          out_colors[3 * count + r] = colors[0][3 * p + r];
        }
        out_labels[count] = labels[0][p];
      }
      count++;
    }
  }
  return count;
}

/* ------------------------------------------------------------------------
 * DepthmapFuser::Fuse  fuser.cc:113-480, single-threaded (SetNumThreads(1)):
 * primary views in index order, pixels in row-major order inside the border
 * margin, neighbours in the order given.  A reference pixel gathers the
 * neighbour pixels its point projects to (rounded) that pass the depth
 * (asymmetric), forward-backward reprojection and normal tests and are not
 * yet fused; with at least min_consistent observations (itself included)
 * it emits the component-wise medians and marks itself and the used pixels
 * fused.  Then statistical outlier removal over the fused points (mean
 * distance to the k nearest neighbours against mean + factor * std).
 *
 * Float arithmetic as the reference's Eigen 3.4 fixed-size expressions
 * (third-party, absent from the reference tree; Ubuntu 24.04 ships 3.4.0):
 * a 3-term product/dot reduces as c0 + (c1 + c2) (the unrolled reduction
 * splits the range in halves), a 2-term one as c0 + c1, (s * A) * v is
 * sum_k (s * A_ik) * v_k, A.inverse() is the cofactor form: cofactors of
 * column 0, det = c00 m00 + (c10 m10 + c20 m20), invdet = 1 / det,
 * inverse(i, j) = cofactor(j, i) * invdet; v.norm() = sqrt(v0^2 + (v1^2 +
 * v2^2)); v /= s divides.  The k-nearest distances are vlfeat's float L2
 * ((dx^2 + dy^2) + dz^2, dimension 3 takes its scalar tail), exact search
 * (no comparison limit). */

static float fu_dot3(const float* a, const float* b) {
  return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

static void fu_mv3(const float* M, const float* x, float* y) {
  float t[3];
  for (int i = 0; i < 3; ++i) t[i] = fu_dot3(M + 3 * i, x);
  for (int i = 0; i < 3; ++i) y[i] = t[i];
}

static float fu_cof(const float* m, int i, int j) {
  const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return m[3 * i1 + j1] * m[3 * i2 + j2] - m[3 * i1 + j2] * m[3 * i2 + j1];
}

void orc_fuse_inverse3f(const float* m, float* r) {
  const float c00 = fu_cof(m, 0, 0), c10 = fu_cof(m, 1, 0), c20 = fu_cof(m, 2, 0);
  const float det = c00 * m[0] + (c10 * m[3] + c20 * m[6]);
  const float invdet = 1.0f / det;
  r[3] = fu_cof(m, 0, 1) * invdet;
  r[4] = fu_cof(m, 1, 1) * invdet;
  r[6] = fu_cof(m, 0, 2) * invdet;
  r[5] = fu_cof(m, 2, 1) * invdet;
  r[7] = fu_cof(m, 1, 2) * invdet;
  r[8] = fu_cof(m, 2, 2) * invdet;
  r[0] = c00 * invdet;
  r[1] = c10 * invdet;
  r[2] = c20 * invdet;
}

/* Rinv * (d * Kinv * (x, y, 1) - t) */
static void fu_backproject(const float* Kinv, const float* Rinv, const float* t, float d,
                           float x, float y, float* out) {
  float w[3];
  for (int i = 0; i < 3; ++i) {
    const float a = d * Kinv[3 * i], b = d * Kinv[3 * i + 1], c = d * Kinv[3 * i + 2];
    w[i] = a * x + (b * y + c * 1.0f);
    w[i] = w[i] - t[i];
  }
  fu_mv3(Rinv, w, out);
}

/* K * (R * X + t) */
static void fu_project(const float* K, const float* R, const float* t, const float* X,
                       float* out) {
  float a[3];
  fu_mv3(R, X, a);
  for (int i = 0; i < 3; ++i) a[i] = a[i] + t[i];
  fu_mv3(K, a, out);
}

static float fu_norm3(const float* v) { return sqrtf(v[0] * v[0] + (v[1] * v[1] + v[2] * v[2])); }

/* static_cast<int>(std::round(x)): outside int (or NaN) the conversion is
 * undefined in C++; x86 returns INT_MIN, which every bounds test rejects. */
static int fu_round_int(float x) {
  const float r = roundf(x);
  if (!(fabsf(r) < 2147483648.0f)) return INT_MIN;
  return (int)r;
}

/* BilinearDepth (fuser.cc:95-111) over foundation::BilinearInterpolation
 * (foundation/interpolation.h:15-32). */
static float fu_bilinear_depth(const float* D, int W, int H, float u, float v) {
  const int x0 = (int)floorf(u), y0 = (int)floorf(v);
  const int x1 = x0 + 1, y1 = y0 + 1;
  if (x0 < 0 || y0 < 0 || x1 >= W || y1 >= H) return -1.0f;
  if (D[(size_t)y0 * W + x0] <= 0.0f || D[(size_t)y0 * W + x1] <= 0.0f ||
      D[(size_t)y1 * W + x0] <= 0.0f || D[(size_t)y1 * W + x1] <= 0.0f)
    return -1.0f;
  if (u < 0.0f || u >= (float)(W - 1) || v < 0.0f || v >= (float)(H - 1)) return -1.0f;
  const int ix = (int)u, iy = (int)v;
  const float dx = u - (float)ix, dy = v - (float)iy;
  const float v00 = D[(size_t)iy * W + ix], v01 = D[(size_t)iy * W + ix + 1];
  const float v10 = D[(size_t)(iy + 1) * W + ix], v11 = D[(size_t)(iy + 1) * W + ix + 1];
  const float a = (1 - dx) * v00 + dx * v01;
  const float b = (1 - dx) * v10 + dx * v11;
  return (1 - dy) * a + dy * b;
}

/* std::nth_element(v, v + n/2, v + n) then v[n/2]: the (n/2)-th smallest
 * value; ties are broken by position (which only decides between +0 and -0). */
static float fu_median_f(const float* v, int n) {
  const int k = n / 2;
  for (int i = 0; i < n; ++i) {
    int rank = 0;
    for (int j = 0; j < n; ++j)
      if (v[j] < v[i] || (v[j] == v[i] && j < i)) ++rank;
    if (rank == k) return v[i];
  }
  return v[k];  /* unreachable without NaN */
}

static unsigned char fu_median_u8(const unsigned char* v, int n) {
  const int k = n / 2;
  for (int i = 0; i < n; ++i) {
    int rank = 0;
    for (int j = 0; j < n; ++j)
      if (v[j] < v[i] || (v[j] == v[i] && j < i)) ++rank;
    if (rank == k) return v[i];
  }
  return v[k];
}

typedef struct {
  float K[9], R[9], t[3], Kinv[9], Rinv[9];
} FuCam;

static void fu_camera(const double* K, const double* R, const double* t, FuCam* c) {
  for (int k = 0; k < 9; ++k) {
    c->K[k] = (float)K[k];
    c->R[k] = (float)R[k];
  }
  for (int k = 0; k < 3; ++k) c->t[k] = (float)t[k];
  orc_fuse_inverse3f(c->K, c->Kinv);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c->Rinv[3 * i + j] = c->R[3 * j + i];
}

/* Statistical outlier removal (fuser.cc:369-478) over n points in place;
 * returns the kept count.  mean_knn (n floats) receives each point's mean
 * distance to its knn nearest others. */
long orc_sor(long n, float* points, float* normals, unsigned char* colors, int knn,
             float factor, float* mean_knn) {
  if (knn <= 0 || n <= (long)knn + 1) return n;
  const int K = knn;
  float* bd = (float*)malloc(sizeof(float) * (K + 1));
  long* bi = (long*)malloc(sizeof(long) * (K + 1));
  for (long i = 0; i < n; ++i) {
    const float* q = points + 3 * i;
    int m = 0;
    for (long j = 0; j < n; ++j) {
      const float* p = points + 3 * j;
      float acc = 0.0f;
      for (int k = 0; k < 3; ++k) {
        const float d = q[k] - p[k];
        acc += d * d;
      }
      /* keep the K + 1 smallest by (distance, index), sorted */
      if (m < K + 1) {
        int s = m++;
        while (s > 0 && bd[s - 1] > acc) {
          bd[s] = bd[s - 1];
          bi[s] = bi[s - 1];
          --s;
        }
        bd[s] = acc;
        bi[s] = j;
      } else if (acc < bd[K]) {
        int s = K;
        while (s > 0 && bd[s - 1] > acc) {
          bd[s] = bd[s - 1];
          bi[s] = bi[s - 1];
          --s;
        }
        bd[s] = acc;
        bi[s] = j;
      }
    }
    float sum = 0.0f;
    int count = 0;
    for (int j = 0; j < m && count < K; ++j) {
      if (bi[j] == i) continue;
      sum += sqrtf(bd[j]);
      ++count;
    }
    mean_knn[i] = count > 0 ? sum / (float)count : 1e30f;
  }
  free(bd);
  free(bi);
  double sum_d = 0, sum_d2 = 0;
  for (long i = 0; i < n; ++i) {
    const double d = mean_knn[i];
    if (d < 1e20f) {
      sum_d += d;
      sum_d2 += d * d;
    }
  }
  const double mean_d = sum_d / n;
  const double var_d = sum_d2 / n - mean_d * mean_d;
  const double std_d = sqrt(var_d > 0.0 ? var_d : 0.0);
  const float threshold = (float)(mean_d + factor * std_d);
  long kept = 0;
  for (long i = 0; i < n; ++i) {
    if (mean_knn[i] <= threshold) {
      for (int k = 0; k < 3; ++k) {
        points[3 * kept + k] = points[3 * i + k];
        normals[3 * kept + k] = normals[3 * i + k];
        colors[3 * kept + k] = colors[3 * i + k];
      }
      ++kept;
    }
  }
  return kept;
}

/* views: n; per view K, R (row-major 3x3), t, depth H*W, normal H*W*3,
 * color H*W*3, mask H*W or NULL, neighbours nb_ids[nb_off[v] .. nb_off[v+1]),
 * primary flag.  Parameters as the setters take them (fuser.cc:52-81).
 * Returns the number of points written (after SOR), or -1 when capacity is
 * too small (the fused count before SOR is then *n_fused). */
long orc_fuse(int n, const double* Ks, const double* Rs, const double* ts,
              const float* const* depths, const float* const* normals,
              const unsigned char* const* colors, const unsigned char* const* masks,
              const int* widths, const int* heights, const int* nb_off, const int* nb_ids,
              const int* primary, int min_consistent, float max_reproj_px,
              float max_depth_error, float max_normal_deg, int border_margin, int sor_knn,
              float sor_factor, float behind_factor, long capacity, float* out_points,
              float* out_normals, unsigned char* out_colors, long* n_fused) {
  const float max_reproj_sq = max_reproj_px * max_reproj_px;
  const float min_cos = cosf(max_normal_deg * (float)3.14159265358979323846 / 180.0f);
  const int bm = border_margin > 0 ? border_margin : 0;
  const int knn = sor_knn > 0 ? sor_knn : 0;
  const float behind = behind_factor < 0.0f ? 0.0f : (behind_factor > 1.0f ? 1.0f : behind_factor);
  *n_fused = 0;
  if (n <= 0) return 0;
  FuCam* cams = (FuCam*)malloc(sizeof(FuCam) * n);
  unsigned char** fused = (unsigned char**)malloc(sizeof(unsigned char*) * n);
  for (int v = 0; v < n; ++v) {
    fu_camera(Ks + 9 * v, Rs + 9 * v, ts + 3 * v, &cams[v]);
    fused[v] = (unsigned char*)calloc((size_t)widths[v] * heights[v], 1);
  }
  int max_nb = 0;
  for (int v = 0; v < n; ++v)
    if (nb_off[v + 1] - nb_off[v] > max_nb) max_nb = nb_off[v + 1] - nb_off[v];
  const int cap_obs = max_nb + 1;
  float* pts = (float*)malloc(sizeof(float) * 3 * cap_obs);
  float* nrm = (float*)malloc(sizeof(float) * 3 * cap_obs);
  unsigned char* clr = (unsigned char*)malloc(3 * cap_obs);
  int* used_v = (int*)malloc(sizeof(int) * cap_obs);
  long* used_p = (long*)malloc(sizeof(long) * cap_obs);
  float* comp = (float*)malloc(sizeof(float) * cap_obs);
  unsigned char* comp8 = (unsigned char*)malloc(cap_obs);
  long count = 0;
  int overflow = 0;
  for (int vi = 0; vi < n; ++vi) {
    if (!primary[vi]) continue;
    const FuCam* A = &cams[vi];
    const int W = widths[vi], H = heights[vi];
    for (int r = bm; r < H - bm; ++r) {
      for (int c = bm; c < W - bm; ++c) {
        const size_t p = (size_t)r * W + c;
        if (fused[vi][p]) continue;
        if (masks[vi] && masks[vi][p] == 0) continue;
        const float d = depths[vi][p];
        if (d <= 0.0f) continue;
        float P[3], nc[3], nw[3];
        fu_backproject(A->Kinv, A->Rinv, A->t, d, (float)c, (float)r, P);
        for (int k = 0; k < 3; ++k) nc[k] = normals[vi][3 * p + k];
        const float nlen = fu_norm3(nc);
        if (nlen < 1e-6f) continue;
        for (int k = 0; k < 3; ++k) nc[k] = nc[k] / nlen;
        fu_mv3(A->Rinv, nc, nw);
        int m = 0, nu = 0;
        for (int k = 0; k < 3; ++k) {
          pts[k] = P[k];
          nrm[k] = nw[k];
          clr[k] = colors[vi][3 * p + k];
        }
        m = 1;
        for (int s = nb_off[vi]; s < nb_off[vi + 1]; ++s) {
          const int o = nb_ids[s];
          if (o < 0 || o >= n) continue;
          const FuCam* S = &cams[o];
          const int Ws = widths[o], Hs = heights[o];
          float proj[3];
          fu_project(S->K, S->R, S->t, P, proj);
          if (proj[2] < 1e-6f) continue;
          const float su = proj[0] / proj[2], sv = proj[1] / proj[2];
          const int sc = fu_round_int(su), sr = fu_round_int(sv);
          if (sc < bm || sc >= Ws - bm || sr < bm || sr >= Hs - bm) continue;
          const size_t sp = (size_t)sr * Ws + sc;
          if (masks[o] && masks[o][sp] == 0) continue;
          if (fused[o][sp]) continue;
          float sd = fu_bilinear_depth(depths[o], Ws, Hs, su, sv);
          if (sd <= 0.0f) {
            sd = depths[o][sp];
            if (sd <= 0.0f) continue;
          }
          const float dp = proj[2];
          const float front = max_depth_error * dp;
          const float back = front * behind;
          const float diff = sd - dp;
          if (diff > front || diff < -back) continue;
          float SP[3], rp[3];
          fu_backproject(S->Kinv, S->Rinv, S->t, sd, (float)sc, (float)sr, SP);
          fu_project(A->K, A->R, A->t, SP, rp);
          if (rp[2] < 1e-6f) continue;
          const float e0 = rp[0] / rp[2] - (float)c, e1 = rp[1] / rp[2] - (float)r;
          const float err = e0 * e0 + e1 * e1;
          if (err > max_reproj_sq) continue;
          float sn[3], snw[3];
          for (int k = 0; k < 3; ++k) sn[k] = normals[o][3 * sp + k];
          const float snlen = fu_norm3(sn);
          if (snlen < 1e-6f) continue;
          for (int k = 0; k < 3; ++k) sn[k] = sn[k] / snlen;
          fu_mv3(S->Rinv, sn, snw);
          const float cosang = fu_dot3(nw, snw);
          if (cosang < min_cos) continue;
          for (int k = 0; k < 3; ++k) {
            pts[3 * m + k] = SP[k];
            nrm[3 * m + k] = snw[k];
            clr[3 * m + k] = colors[o][3 * sp + k];
          }
          ++m;
          used_v[nu] = o;
          used_p[nu] = (long)sp;
          ++nu;
        }
        if (m < min_consistent) continue;
        float mp[3], mn[3];
        unsigned char mc[3];
        for (int k = 0; k < 3; ++k) {
          for (int j = 0; j < m; ++j) comp[j] = pts[3 * j + k];
          mp[k] = fu_median_f(comp, m);
        }
        for (int k = 0; k < 3; ++k) {
          for (int j = 0; j < m; ++j) comp[j] = nrm[3 * j + k];
          mn[k] = fu_median_f(comp, m);
        }
        const float mlen = fu_norm3(mn);
        if (mlen > 1e-6f)
          for (int k = 0; k < 3; ++k) mn[k] = mn[k] / mlen;
        for (int k = 0; k < 3; ++k) {
          for (int j = 0; j < m; ++j) comp8[j] = clr[3 * j + k];
          mc[k] = fu_median_u8(comp8, m);
        }
        if (count < capacity) {
          for (int k = 0; k < 3; ++k) {
            out_points[3 * count + k] = mp[k];
            out_normals[3 * count + k] = mn[k];
            out_colors[3 * count + k] = mc[k];
          }
        } else {
          overflow = 1;
        }
        ++count;
        fused[vi][p] = 1;
        for (int u = 0; u < nu; ++u) fused[used_v[u]][used_p[u]] = 1;
      }
    }
  }
  for (int v = 0; v < n; ++v) free(fused[v]);
  free(fused);
  free(cams);
  free(pts);
  free(nrm);
  free(clr);
  free(used_v);
  free(used_p);
  free(comp);
  free(comp8);
  *n_fused = count;
  if (overflow) return -1;
  float* mk = (float*)malloc(sizeof(float) * (count > 0 ? count : 1));
  const long kept = orc_sor(count, out_points, out_normals, out_colors, knn, sor_factor, mk);
  free(mk);
  return kept;
}

/* ------------------------------------------------------------------------
 * Neighbour selection (opensfm/src/lib/sfm/src/dense_helpers.cc).
 *
 * orc_covis: BuildCovisibilityGraph (dense_helpers.cc:30-326).  Points seen
 * by >= 2 processable shots (shot indices in [0, n_shots)) are kept in input
 * order; k = max(1, min(fuse_knn, m - 1)) nearest others by brute force over
 * float coordinates (vlfeat's float L2, ties to the lower index); scale =
 * (m/2)-th smallest of the per-point (k/2)-th nearest distances; points i < j
 * with j among i's k nearest at squared distance <= (float)(r * r), r =
 * factor * scale, are merged (union-find); super-points are the components,
 * numbered by their smallest member, with the FP64 centroid summed in member
 * order, the sorted union of the members' shots, and the members.  Edges are
 * the shot pairs (lo < hi) sharing a super-point, in (lo, hi) order, weighted
 * by the count.  The reference's orders of super-points and edges come from
 * hash-map iteration and are unspecified; these are the canonical ones.
 * Returns the number of super-points, or -1 when a capacity is too small. */

static int uf_find(int* parent, int x) {
  while (parent[x] != x) {
    parent[x] = parent[parent[x]];
    x = parent[x];
  }
  return x;
}

static int cmp_int(const void* a, const void* b) {
  const int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

long orc_covis(long n, const double* coords, const long* obs_off, const int* obs_shot,
               int n_shots, int fuse_knn, double fuse_radius_factor, int* point_sp,
               double* sp_coord, long* sp_member_off, int* sp_members, long* sp_vis_off,
               int* sp_vis, long vis_cap, int* edges, int* weights, long edge_cap,
               long* n_edges, float* scale_out) {
  long m = 0;
  long* keep = (long*)malloc(sizeof(long) * (n > 0 ? n : 1));
  for (long i = 0; i < n; ++i) {
    int c = 0;
    for (long o = obs_off[i]; o < obs_off[i + 1]; ++o)
      c += obs_shot[o] >= 0 && obs_shot[o] < n_shots;
    point_sp[i] = -1;
    if (c >= 2) keep[m++] = i;
  }
  *n_edges = 0;
  *scale_out = 0.0f;
  if (m == 0) {
    free(keep);
    return 0;
  }
  int k = fuse_knn < (int)(m - 1) ? fuse_knn : (int)(m - 1);
  if (k < 1) k = 1;
  float* fc = (float*)malloc(sizeof(float) * 3 * m);
  for (long i = 0; i < m; ++i)
    for (int a = 0; a < 3; ++a) fc[3 * i + a] = (float)coords[3 * keep[i] + a];
  int* kidx = (int*)malloc(sizeof(int) * m * k);
  float* kd2 = (float*)malloc(sizeof(float) * m * k);
  float* kdist = (float*)malloc(sizeof(float) * m);
  for (long i = 0; i < m; ++i) {
    int found = 0;
    int* bi = kidx + i * k;
    float* bd = kd2 + i * k;
    for (long j = 0; j < m; ++j) {
      if (j == i) continue;
      float acc = 0.0f;
      for (int a = 0; a < 3; ++a) {
        const float d = fc[3 * i + a] - fc[3 * j + a];
        acc += d * d;
      }
      /* keep the k smallest by (distance, index) */
      if (found == k && !(acc < bd[k - 1])) continue;  /* j > every kept index on ties */
      int s = found < k ? found++ : k - 1;
      while (s > 0 && bd[s - 1] > acc) {
        bd[s] = bd[s - 1];
        bi[s] = bi[s - 1];
        --s;
      }
      bd[s] = acc;
      bi[s] = (int)j;
    }
    for (int t = found; t < k; ++t) {
      bi[t] = -1;
      bd[t] = 0.0f;
    }
    kdist[i] = found > 0 ? sqrtf(bd[found / 2]) : 0.0f;
  }
  float* sorted = (float*)malloc(sizeof(float) * m);
  memcpy(sorted, kdist, sizeof(float) * m);
  qsort(sorted, m, sizeof(float), cmp_float);
  const double scale = (double)sorted[m / 2];
  const double radius = fuse_radius_factor * scale;
  const float r2 = (float)(radius * radius);
  *scale_out = (float)scale;
  int* parent = (int*)malloc(sizeof(int) * m);
  for (long i = 0; i < m; ++i) parent[i] = (int)i;
  for (long i = 0; i < m; ++i)
    for (int s = 0; s < k; ++s) {
      const int o = kidx[i * k + s];
      if (o < 0) break;
      if (o > i && kd2[i * k + s] <= r2) {
        int a = uf_find(parent, (int)i), b = uf_find(parent, o);
        if (a != b) {
          if (a < b) parent[b] = a;
          else parent[a] = b;
        }
      }
    }
  /* components numbered by smallest member */
  int* sp = (int*)malloc(sizeof(int) * m);
  long nsp = 0;
  for (long i = 0; i < m; ++i) {
    const int r = uf_find(parent, (int)i);
    if (r == i) sp[i] = (int)nsp++;
  }
  for (long i = 0; i < m; ++i) sp[i] = sp[uf_find(parent, (int)i)];
  for (long i = 0; i < m; ++i) point_sp[keep[i]] = sp[i];
  /* members in ascending order, centroids summed in that order */
  for (long s = 0; s <= nsp; ++s) sp_member_off[s] = 0;
  for (long i = 0; i < m; ++i) sp_member_off[sp[i] + 1]++;
  for (long s = 0; s < nsp; ++s) sp_member_off[s + 1] += sp_member_off[s];
  long* fill = (long*)malloc(sizeof(long) * (nsp > 0 ? nsp : 1));
  for (long s = 0; s < nsp; ++s) fill[s] = sp_member_off[s];
  for (long i = 0; i < m; ++i) sp_members[fill[sp[i]]++] = (int)keep[i];
  for (long s = 0; s < nsp; ++s) {
    double c[3] = {0, 0, 0};
    for (long t = sp_member_off[s]; t < sp_member_off[s + 1]; ++t)
      for (int a = 0; a < 3; ++a) c[a] += coords[3 * (long)sp_members[t] + a];
    const double cnt = (double)(sp_member_off[s + 1] - sp_member_off[s]);
    for (int a = 0; a < 3; ++a) sp_coord[3 * s + a] = c[a] / cnt;
  }
  /* visibility: sorted unique shots of the members */
  int* buf = (int*)malloc(sizeof(int) * (n_shots > 0 ? n_shots : 1));
  unsigned char* seen = (unsigned char*)calloc(n_shots > 0 ? n_shots : 1, 1);
  long nv = 0;
  int overflow = 0;
  for (long s = 0; s < nsp; ++s) {
    sp_vis_off[s] = nv;
    int nb = 0;
    for (long t = sp_member_off[s]; t < sp_member_off[s + 1]; ++t) {
      const long p = sp_members[t];
      for (long o = obs_off[p]; o < obs_off[p + 1]; ++o) {
        const int sh = obs_shot[o];
        if (sh < 0 || sh >= n_shots || seen[sh]) continue;
        seen[sh] = 1;
        buf[nb++] = sh;
      }
    }
    qsort(buf, nb, sizeof(int), cmp_int);
    for (int b = 0; b < nb; ++b) {
      seen[buf[b]] = 0;
      if (nv < vis_cap) sp_vis[nv] = buf[b];
      else overflow = 1;
      ++nv;
    }
  }
  sp_vis_off[nsp] = nv;
  /* edges */
  if (!overflow) {
    int* cnt = (int*)calloc((size_t)n_shots * n_shots, sizeof(int));
    for (long s = 0; s < nsp; ++s)
      for (long a = sp_vis_off[s]; a < sp_vis_off[s + 1]; ++a)
        for (long b = a + 1; b < sp_vis_off[s + 1]; ++b)
          cnt[(size_t)sp_vis[a] * n_shots + sp_vis[b]]++;
    long ne = 0;
    for (long a = 0; a < n_shots; ++a)
      for (long b = a + 1; b < n_shots; ++b) {
        const int c = cnt[(size_t)a * n_shots + b];
        if (c < 1) continue;
        if (ne < edge_cap) {
          edges[2 * ne] = (int)a;
          edges[2 * ne + 1] = (int)b;
          weights[ne] = c;
        } else {
          overflow = 1;
        }
        ++ne;
      }
    *n_edges = ne;
    free(cnt);
  }
  free(buf);
  free(seen);
  free(fill);
  free(sp);
  free(parent);
  free(sorted);
  free(kdist);
  free(kd2);
  free(kidx);
  free(fc);
  free(keep);
  return overflow ? -1 : nsp;
}

/* orc_select_neighbors: SelectNeighbors (dense_helpers.cc:331-512).  Per
 * super-point and shot pair (vis entries in [0, n_shots) with camera-to-
 * centroid norm >= 1e-12): cos of the angle at the centroid (FP64, clamped),
 * count += n_tracks, score_best += n_tracks when cos(theta_max) < cos <
 * cos(theta_min), and sum_cos += cos * n_tracks in fixed point (cos rounded
 * to a multiple of 2^-32, so the sum does not depend on the order of
 * super-points; the reference sums FP64 in a thread-dependent order).  Per
 * shot: "best" = itself then up to num_neighbors others with count >
 * min_point_best, mean angle within [theta_min, theta_max] (compared as
 * cosines) and score_best > min_point_best, by score descending (ties: lower
 * index first); "all" = best then, in index order, the other shots with
 * count > min_point_all.  Returns the total of the all lists, or -1 when a
 * capacity is too small. */
long orc_select_neighbors(int n_shots, const double* origins, long n_sp, const double* sp_coord,
                          const long* sp_vis_off, const int* sp_vis, const int* sp_ntracks,
                          int num_neighbors, int min_point_best, int min_point_all,
                          double theta_min_deg, double theta_max_deg, long* best_off,
                          int* best, long best_cap, long* all_off, int* all, long all_cap) {
  const double pi = 3.14159265358979323846;
  const double cos_lo = cos(theta_max_deg * pi / 180.0);
  const double cos_hi = cos(theta_min_deg * pi / 180.0);
  const size_t NN = (size_t)n_shots * n_shots;
  long* count = (long*)calloc(NN > 0 ? NN : 1, sizeof(long));
  long* score = (long*)calloc(NN > 0 ? NN : 1, sizeof(long));
  long long* sumfx = (long long*)calloc(NN > 0 ? NN : 1, sizeof(long long));
  int* cams = (int*)malloc(sizeof(int) * (n_shots > 0 ? n_shots : 1));
  double* cv = (double*)malloc(sizeof(double) * 4 * (n_shots > 0 ? n_shots : 1));
  for (long s = 0; s < n_sp; ++s) {
    int nc = 0;
    for (long t = sp_vis_off[s]; t < sp_vis_off[s + 1]; ++t)
      if (sp_vis[t] >= 0 && sp_vis[t] < n_shots) cams[nc++] = sp_vis[t];
    if (nc < 2) continue;
    const long nt = sp_ntracks[s];
    for (int c = 0; c < nc; ++c) {
      double* v = cv + 4 * c;
      for (int a = 0; a < 3; ++a) v[a] = origins[3 * cams[c] + a] - sp_coord[3 * s + a];
      v[3] = sqrt(v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]));
    }
    for (int ci = 0; ci < nc; ++ci) {
      if (cv[4 * ci + 3] < 1e-12) continue;
      for (int cj = ci + 1; cj < nc; ++cj) {
        if (cv[4 * cj + 3] < 1e-12) continue;
        const double* u = cv + 4 * ci;
        const double* w = cv + 4 * cj;
        double c = (u[0] * w[0] + (u[1] * w[1] + u[2] * w[2])) / (u[3] * w[3]);
        c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
        const int good = c > cos_lo && c < cos_hi;
        const long long fx = llrint(c * 4294967296.0) * nt;
        const int a = cams[ci], b = cams[cj];
        const size_t ab = (size_t)a * n_shots + b, ba = (size_t)b * n_shots + a;
        count[ab] += nt;
        count[ba] += nt;
        score[ab] += good ? nt : 0;
        score[ba] += good ? nt : 0;
        sumfx[ab] += fx;
        sumfx[ba] += fx;
      }
    }
  }
  long nb_total = 0, na_total = 0;
  int overflow = 0;
  int* cand = (int*)malloc(sizeof(int) * (n_shots > 0 ? n_shots : 1));
  unsigned char* in_best = (unsigned char*)calloc(n_shots > 0 ? n_shots : 1, 1);
  for (int i = 0; i < n_shots; ++i) {
    int nb = 0;
    for (int j = 0; j < n_shots; ++j) {
      const size_t ij = (size_t)i * n_shots + j;
      if (count[ij] == 0 || count[ij] <= min_point_best) continue;
      double avg = (double)sumfx[ij] / 4294967296.0 / (double)count[ij];
      avg = avg < -1.0 ? -1.0 : (avg > 1.0 ? 1.0 : avg);
      if (avg > cos_hi || avg < cos_lo) continue;   /* mean angle outside [min, max] */
      if (score[ij] > min_point_best) cand[nb++] = j;
    }
    /* score descending, index ascending (insertion sort: stable) */
    for (int a = 1; a < nb; ++a) {
      const int x = cand[a];
      int b = a;
      while (b > 0 && score[(size_t)i * n_shots + cand[b - 1]] < score[(size_t)i * n_shots + x]) {
        cand[b] = cand[b - 1];
        --b;
      }
      cand[b] = x;
    }
    const int take = nb < num_neighbors ? nb : (num_neighbors > 0 ? num_neighbors : 0);
    best_off[i] = nb_total;
    all_off[i] = na_total;
    const int row[1] = {i};
    for (int t = 0; t < 1 + take; ++t) {
      const int v = t == 0 ? row[0] : cand[t - 1];
      in_best[v] = 1;
      if (nb_total < best_cap) best[nb_total] = v;
      else overflow = 1;
      if (na_total < all_cap) all[na_total] = v;
      else overflow = 1;
      ++nb_total;
      ++na_total;
    }
    for (int j = 0; j < n_shots; ++j) {
      const size_t ij = (size_t)i * n_shots + j;
      if (count[ij] > min_point_all && !in_best[j]) {
        if (na_total < all_cap) all[na_total] = j;
        else overflow = 1;
        ++na_total;
      }
    }
    for (int t = 0; t < 1 + take; ++t) in_best[t == 0 ? i : cand[t - 1]] = 0;
  }
  best_off[n_shots] = nb_total;
  all_off[n_shots] = na_total;
  free(cand);
  free(in_best);
  free(cams);
  free(cv);
  free(count);
  free(score);
  free(sumfx);
  return overflow ? -1 : na_total;
}
