// Host-code checks run under the address and undefined-behaviour sanitizers
// (built by tests/test_host_sanitizers.py with g++, no GPU): the LZ4 block
// codec on round trips and on corrupt, truncated and random inputs, and the
// FP64 camera helpers (inverse, relative pose, view description, bilateral
// table) on random and degenerate cameras.  Exits non-zero on a failed check;
// the sanitizers abort on any memory error or undefined behaviour.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <random>
#include <vector>

#include "../../include/opensfm_amd.h"
#include "../../opensfm_amd/csrc/pm_math.h"

static int failures = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      ++failures;                                                 \
    }                                                             \
  } while (0)

static void lz4_checks(std::mt19937_64& rng) {
  const int64_t sizes[] = {0, 1, 15, 16, 17, 255, 256, 4096, 65536 + 7, 300000};
  for (int64_t n : sizes) {
    for (int kind = 0; kind < 3; ++kind) {
      std::vector<uint8_t> src(n);
      for (int64_t i = 0; i < n; ++i) {
        if (kind == 0) src[i] = (uint8_t)rng();                 // incompressible
        else if (kind == 1) src[i] = (uint8_t)(i % 7);          // periodic
        else src[i] = (uint8_t)((rng() % 16 == 0) ? rng() : 0);  // sparse
      }
      const int64_t cap = n + n / 255 + 64;
      std::vector<uint8_t> comp(cap);
      int64_t written = -1;
      CHECK(pm_lz4_block_compress(src.data(), n, comp.data(), cap, &written) == PM_OK);
      CHECK(written >= 0 && written <= cap);
      std::vector<uint8_t> back(n + 1);
      int64_t pos = -1;
      CHECK(pm_lz4_block_decompress(comp.data(), written, back.data(), 0, n, &pos) == PM_OK);
      CHECK(pos == n && (n == 0 || memcmp(back.data(), src.data(), n) == 0));
      // too small an output must fail cleanly
      if (n > 0) {
        std::vector<uint8_t> small(n - 1 > 0 ? n - 1 : 1);
        CHECK(pm_lz4_block_decompress(comp.data(), written, small.data(), 0, n - 1, &pos) != PM_OK);
      }
      // every truncation of the compressed stream decodes or fails, never
      // reads or writes out of bounds
      for (int64_t cut = 0; cut < written; cut += 1 + written / 64) {
        std::vector<uint8_t> part(comp.begin(), comp.begin() + cut);
        pm_lz4_block_decompress(part.data(), cut, back.data(), 0, n, &pos);
      }
      // too small a compression buffer must fail cleanly
      if (written > 1) {
        std::vector<uint8_t> tiny(written / 2);
        int64_t w2 = 0;
        pm_lz4_block_compress(src.data(), n, tiny.data(), (int64_t)tiny.size(), &w2);
      }
    }
  }
  // random garbage as a compressed stream
  for (int t = 0; t < 2000; ++t) {
    const int64_t m = 1 + (int64_t)(rng() % 200);
    std::vector<uint8_t> junk(m);
    for (auto& b : junk) b = (uint8_t)rng();
    std::vector<uint8_t> out(512);
    int64_t pos = 0;
    pm_lz4_block_decompress(junk.data(), m, out.data(), 0, (int64_t)out.size(), &pos);
  }
  // invalid arguments
  int64_t pos = 0;
  uint8_t b[4] = {0, 0, 0, 0};
  CHECK(pm_lz4_block_decompress(nullptr, 4, b, 0, 4, &pos) != PM_OK);
  CHECK(pm_lz4_block_decompress(b, -1, b, 0, 4, &pos) != PM_OK);
  CHECK(pm_lz4_block_decompress(b, 4, b, 5, 4, &pos) != PM_OK);
}

static void camera_checks(std::mt19937_64& rng) {
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  for (int t = 0; t < 2000; ++t) {
    double A[9], Ai[9], I[9];
    for (double& x : A) x = u(rng);
    if (pmhost::inv3(A, Ai)) {
      pmhost::mat3_mul(A, Ai, I);
      double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                   A[2] * (A[3] * A[7] - A[4] * A[6]);
      if (fabs(det) > 1e-3)
        for (int k = 0; k < 9; ++k) CHECK(fabs(I[k] - (k % 4 == 0 ? 1.0 : 0.0)) < 1e-6);
    }
    double Rr[9], tr[3], Rv[9], tv[3], Q[9], a[3];
    for (double& x : Rr) x = u(rng);
    for (double& x : Rv) x = u(rng);
    for (double& x : tr) x = 10 * u(rng);
    for (double& x : tv) x = 10 * u(rng);
    pmhost::relative_pose(Rr, tr, Rv, tv, Q, a);
    double K[9] = {500, 0, 320, 0, 500, 240, 0, 0, 1}, Ki[9];
    CHECK(pmhost::inv3(K, Ki));
    pm::PmView V;
    pmhost::fill_view(V, nullptr, 640, 480, K, Q, a, Ki);
    CHECK(V.width == 640 && V.height == 480);
  }
  // singular and degenerate matrices
  double Z[9] = {0}, Zi[9];
  CHECK(!pmhost::inv3(Z, Zi));
  double S[9] = {1, 2, 3, 2, 4, 6, 0, 0, 1};
  CHECK(!pmhost::inv3(S, Zi));
  double N[9] = {NAN, 0, 0, 0, 1, 0, 0, 0, 1};
  CHECK(!pmhost::all_finite(N, 9));
  for (int hpz = 1; hpz <= 5; ++hpz) {
    const std::vector<float> tab = pmhost::bilateral_table(hpz);
    CHECK(tab.size() == 256u * (size_t)(2 * hpz * hpz + 1));
    CHECK(tab[0] == 1.0f);
    for (float w : tab) CHECK(w > 0.0f && w <= 1.0f);
  }
}

int main() {
  std::mt19937_64 rng(12345);
  lz4_checks(rng);
  camera_checks(rng);
  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("host checks passed\n");
  return 0;
}
