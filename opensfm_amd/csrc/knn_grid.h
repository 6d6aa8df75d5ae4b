// Host side of the exact k-nearest-neighbour search over a point cloud:
// Morton order and per-level cell hash tables (kernels in fuse_kernels.hip).
// Used by the fuser's outlier removal and by the covisibility graph's
// point merging.
#pragma once

#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "fuse_kernels.h"
#include "pm_host.h"

namespace fu {

using pmhost::HipError;

struct KnnGrid {
  pmhost::DevBuf lohi, codes[2], idx[2], sorted, sort_tmp, ncells;
  std::vector<std::unique_ptr<pmhost::DevBuf>> tables;
  Grid g;
  long long n = 0;

  // Sorted copy of the n points (float x, y, z), their original indices
  // (order()) and the hash tables from the first level whose cells hold
  // about 6 points on average up to the level holding every point.
  void build(const float* points, long long count, hipStream_t s) {
    using pmhost::DevBuf;
    n = count;
    lohi.ensure(6 * sizeof(float));
    HIP_CHECK(launch_bbox(points, n, lohi.as<float>(), s));
    float box[6];
    HIP_CHECK(hipMemcpyAsync(box, lohi.p, sizeof(box), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    float extent = 0.0f;
    for (int a = 0; a < 3; ++a) extent = std::max(extent, box[3 + a] - box[a]);
    const float quantum = extent > 0.0f && std::isfinite(extent) ? extent / 2097152.0f : 1.0f;
    for (int b = 0; b < 2; ++b) {
      codes[b].ensure(n * sizeof(unsigned long long));
      idx[b].ensure(n * sizeof(unsigned int));
    }
    HIP_CHECK(launch_morton(points, n, lohi.as<float>(), quantum,
                            codes[0].as<unsigned long long>(), idx[0].as<unsigned int>(), s));
    size_t tmp_bytes = 0;
    HIP_CHECK(sort_codes(codes[0].as<unsigned long long>(), codes[1].as<unsigned long long>(),
                         idx[0].as<unsigned int>(), idx[1].as<unsigned int>(), (int)n, nullptr,
                         &tmp_bytes, s));
    sort_tmp.ensure(tmp_bytes);
    HIP_CHECK(sort_codes(codes[0].as<unsigned long long>(), codes[1].as<unsigned long long>(),
                         idx[0].as<unsigned int>(), idx[1].as<unsigned int>(), (int)n,
                         sort_tmp.p, &tmp_bytes, s));
    const unsigned long long* sc = codes[1].as<unsigned long long>();
    sorted.ensure(n * 3 * sizeof(float));
    HIP_CHECK(launch_gather_sorted(points, order(), n, sorted.as<float>(), s));
    ncells.ensure(22 * sizeof(unsigned int));
    HIP_CHECK(hipMemsetAsync(ncells.p, 0, 22 * sizeof(unsigned int), s));
    HIP_CHECK(launch_level_count_all(sc, n, ncells.as<unsigned int>(), s));
    unsigned int nc[22];
    HIP_CHECK(hipMemcpyAsync(nc, ncells.p, sizeof(nc), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    int L0 = 21;
    for (int L = 0; L < 22; ++L)
      if ((double)n >= 6.0 * nc[L]) {
        L0 = L;
        break;
      }
    memset(&g, 0, sizeof(g));
    for (int a = 0; a < 3; ++a) g.lo[a] = box[a];
    g.quantum = quantum;
    g.level0 = L0;
    g.levels = 22 - L0;
    tables.clear();
    for (int L = L0; L < 22; ++L) {
      unsigned int size = 1;
      while (size < 2u * std::max(nc[L], 1u)) size <<= 1;
      const size_t bytes = (size_t)size * (sizeof(unsigned long long) + 2 * sizeof(unsigned int));
      std::unique_ptr<DevBuf> t(new DevBuf());
      t->ensure(bytes);
      HIP_CHECK(hipMemsetAsync(t->p, 0, (size_t)size * sizeof(unsigned long long), s));
      const int li = L - L0;
      g.keys[li] = t->as<unsigned long long>();
      g.start[li] = reinterpret_cast<unsigned int*>(g.keys[li] + size);
      g.cnt[li] = g.start[li] + size;
      g.mask[li] = size - 1;
      HIP_CHECK(launch_level_build(sc, n, L, g.keys[li], g.start[li], g.cnt[li], size - 1, s));
      tables.push_back(std::move(t));
    }
  }

  const unsigned int* order() const { return idx[1].as<unsigned int>(); }
  const float* sorted_points() const { return sorted.as<float>(); }
};

}  // namespace fu
