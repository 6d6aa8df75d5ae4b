// HIP-free host helpers: the calling thread's error message, FP64 3x3
// algebra in the reference's order (cv::Matx), the bilateral weight table and
// the kernels' view descriptions.  Built with -ffp-contract=off, and built
// on its own under the address and undefined-behaviour sanitizers by
// tests/test_host_sanitizers.py.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

#include "pm_types.h"

namespace pmhost {

inline thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// ---- FP64 3x3 helpers in cv::Matx order (accumulate from zero) ----------

inline void mat3_mul(const double* A, const double* B, double* C) {
  double T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[k * 3 + j];
      T[i * 3 + j] = s;
    }
  memcpy(C, T, sizeof(T));
}

inline void mat3_vec(const double* A, const double* x, double* y) {
  double t[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0;
    for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * x[k];
    t[i] = s;
  }
  memcpy(y, t, sizeof(t));
}

inline void mat3_transpose(const double* A, double* B) {
  double T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T[j * 3 + i] = A[i * 3 + j];
  memcpy(B, T, sizeof(T));
}

// cv::Matx33d::inv() (DECOMP_LU): determinant and cofactors for 3x3.
inline bool inv3(const double* a, double* b) {
#define A(r, c) a[(r) * 3 + (c)]
  double d = A(0, 0) * (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) -
             A(0, 1) * (A(1, 0) * A(2, 2) - A(2, 0) * A(1, 2)) +
             A(0, 2) * (A(1, 0) * A(2, 1) - A(2, 0) * A(1, 1));
  if (d == 0) return false;
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
  return true;
}

inline bool all_finite(const double* p, int n) {
  for (int k = 0; k < n; ++k)
    if (!std::isfinite(p[k])) return false;
  return true;
}

// Bilateral weights of every (|dcolor|, dx^2 + dy^2) pair, with the
// reference's FP32 argument and expf (BilateralWeight, depthmap.cc:480-487).
inline std::vector<float> bilateral_table(int hpz) {
  const float dcolor_sigma = 50.0f;
  const float dx_sigma = 5.0f;
  const float dcolor_factor = 1.0f / (2 * dcolor_sigma * dcolor_sigma);
  const float dx_factor = 1.0f / (2 * dx_sigma * dx_sigma);
  const int nr2 = 2 * hpz * hpz + 1;
  std::vector<float> t(256 * (size_t)nr2);
  for (int dc = 0; dc < 256; ++dc) {
    for (int r2 = 0; r2 < nr2; ++r2) {
      const float dcolor = (float)dc;
      const float rr = (float)r2;  // dx*dx + dy*dy, exact in FP32
      t[(size_t)dc * nr2 + r2] =
          expf(-dcolor * dcolor * dcolor_factor - rr * dx_factor);
    }
  }
  return t;
}

// Pose of view v relative to reference view r (DepthmapEstimator::AddView,
// depthmap.cc:153-155): Q = R_v R_r^T, a = Q t_r - t_v.
inline void relative_pose(const double* Rr, const double* tr, const double* Rv,
                          const double* tv, double* Q, double* a) {
  double Rrt[9], Qtr[3];
  mat3_transpose(Rr, Rrt);
  mat3_mul(Rv, Rrt, Q);
  mat3_vec(Q, tr, Qtr);
  for (int k = 0; k < 3; ++k) a[k] = Qtr[k] - tv[k];
}

// Kernel-side description of view v for reference K_r^-1: the FP32
// homography terms M = K_v Q K_r^-1 and b = K_v a, composed in FP64, and
// the FP64 inputs of the exact path.
inline void fill_view(pm::PmView& T, const uint32_t* quad, int width,
                      int height, const double* Kv, const double* Q,
                      const double* a, const double* Kinv_r) {
  memset(&T, 0, sizeof(T));
  T.quad = quad;
  T.width = width;
  T.height = height;
  double KQ[9], M[9], b[3];
  mat3_mul(Kv, Q, KQ);
  mat3_mul(KQ, Kinv_r, M);
  mat3_vec(Kv, a, b);
  for (int k = 0; k < 9; ++k) T.M[k] = (float)M[k];
  for (int k = 0; k < 3; ++k) T.b[k] = (float)b[k];
  memcpy(T.K, Kv, sizeof(T.K));
  memcpy(T.Q, Q, sizeof(T.Q));
  memcpy(T.a, a, sizeof(T.a));
}

}  // namespace pmhost
