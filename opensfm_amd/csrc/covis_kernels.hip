// Neighbour selection before the depthmaps (sfm::dense_helpers,
// opensfm/src/lib/sfm/src/dense_helpers.cc) on gfx950: the super-point
// covisibility graph (k nearest over the landmarks, union-find merging,
// visibility unions, shot-pair counts) and the per-shot neighbour scores
// and choices.
//
// Shot pairs live in dense N x N tables (N = processable shots, at most
// 16384): one atomic per super-point and pair, then one workgroup per shot
// reads its row.  Angle sums are accumulated in fixed point (cos rounded to
// a multiple of 2^-32), so they do not depend on the order of the atomics.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "covis_kernels.h"

namespace cv {

namespace {

inline int blocks(long long n) { return (int)((n + 255) / 256); }

__global__ __launch_bounds__(256) void k_valid(const long long* __restrict__ off,
                                               const int* __restrict__ shots, long long n,
                                               int n_shots, uint8_t* __restrict__ flag) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  int c = 0;
  for (long long o = off[i]; o < off[i + 1]; ++o) c += (shots[o] >= 0 && shots[o] < n_shots);
  flag[i] = c >= 2;
}

__global__ __launch_bounds__(256) void k_flag_count(const uint8_t* __restrict__ flag,
                                                    long long n, int* __restrict__ counts) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  const int c = __syncthreads_count(i < n && flag[i]);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// rank[i] = position of i among the flagged entries (row-major order), for
// flagged i; index[rank] = i.
__global__ __launch_bounds__(256) void k_flag_rank(const uint8_t* __restrict__ flag, long long n,
                                                   const long long* __restrict__ offsets,
                                                   int* __restrict__ rank,
                                                   int* __restrict__ index) {
  __shared__ int wave_base[4];
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  const bool f = i < n && flag[i];
  const unsigned long long m = __ballot(f);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wave_base[wave] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < 4; ++w) {
      const int c = wave_base[w];
      wave_base[w] = acc;
      acc += c;
    }
  }
  __syncthreads();
  if (!f) return;
  const long long r = offsets[blockIdx.x] + wave_base[wave] + __popcll(m & ((1ull << lane) - 1ull));
  if (rank) rank[i] = (int)r;
  if (index) index[r] = (int)i;
}

__global__ __launch_bounds__(256) void k_gather_points(const double* __restrict__ coords,
                                                       const int* __restrict__ keep, long long m,
                                                       float* __restrict__ fc) {
  const long long j = blockIdx.x * 256ll + threadIdx.x;
  if (j >= m) return;
  const long long i = keep[j];
#pragma unroll
  for (int a = 0; a < 3; ++a) fc[3 * j + a] = (float)coords[3 * i + a];
}

__global__ __launch_bounds__(256) void k_kdist(const int* __restrict__ kidx,
                                               const float* __restrict__ kd2, long long m, int k,
                                               float* __restrict__ kdist) {
  const long long j = blockIdx.x * 256ll + threadIdx.x;
  if (j >= m) return;
  int found = 0;
  for (int s = 0; s < k; ++s) found += kidx[j * k + s] >= 0;
  kdist[j] = found > 0 ? sqrtf(kd2[j * k + found / 2]) : 0.0f;
}

__device__ __forceinline__ int root_of(const int* __restrict__ parent, int x) {
  int p = parent[x];
  while (p != x) {
    x = p;
    p = parent[x];
  }
  return x;
}

// Links the components of i and its merge partners: a root always points to
// a smaller index, so the final root of a component is its smallest member.
__global__ __launch_bounds__(256) void k_hook(const int* __restrict__ kidx,
                                              const float* __restrict__ kd2, long long m, int k,
                                              float r2, int* parent, int* __restrict__ changed) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= m) return;
  for (int s = 0; s < k; ++s) {
    const int o = kidx[i * k + s];
    if (o < 0) break;
    if (o > i && kd2[i * k + s] <= r2) {
      const int a = root_of(parent, (int)i), b = root_of(parent, o);
      if (a != b) {
        atomicMin(&parent[a > b ? a : b], a < b ? a : b);
        *changed = 1;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_compress(int* parent, long long m) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= m) return;
  parent[i] = root_of(parent, (int)i);
}

__global__ __launch_bounds__(256) void k_is_root(const int* __restrict__ parent, long long m,
                                                 uint8_t* __restrict__ flag) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i < m) flag[i] = parent[i] == i;
}

__global__ __launch_bounds__(256) void k_sp_of(const int* __restrict__ parent,
                                               const int* __restrict__ root_rank, long long m,
                                               const int* __restrict__ keep,
                                               unsigned int* __restrict__ sp,
                                               unsigned int* __restrict__ member,
                                               int* __restrict__ point_sp,
                                               unsigned int* __restrict__ counts) {
  const long long j = blockIdx.x * 256ll + threadIdx.x;
  if (j >= m) return;
  const int s = root_rank[parent[j]];
  sp[j] = (unsigned int)s;
  member[j] = (unsigned int)j;
  point_sp[keep[j]] = s;
  atomicAdd(&counts[s], 1u);
}

// One thread per super-point: the FP64 centroid summed in member order.
__global__ __launch_bounds__(256) void k_centroid(const double* __restrict__ coords,
                                                  const int* __restrict__ keep,
                                                  const unsigned int* __restrict__ members,
                                                  const long long* __restrict__ moff,
                                                  long long nsp, double* __restrict__ out,
                                                  int* __restrict__ out_members) {
  const long long s = blockIdx.x * 256ll + threadIdx.x;
  if (s >= nsp) return;
  double c[3] = {0.0, 0.0, 0.0};
  for (long long t = moff[s]; t < moff[s + 1]; ++t) {
    const long long i = keep[members[t]];
    out_members[t] = (int)i;
#pragma unroll
    for (int a = 0; a < 3; ++a) c[a] += coords[3 * i + a];
  }
  const double cnt = (double)(moff[s + 1] - moff[s]);
#pragma unroll
  for (int a = 0; a < 3; ++a) out[3 * s + a] = __ddiv_rn(c[a], cnt);
}

__global__ __launch_bounds__(256) void k_obs_count(const long long* __restrict__ off,
                                                   const int* __restrict__ shots,
                                                   const int* __restrict__ keep, long long m,
                                                   int n_shots, long long* __restrict__ cnt) {
  const long long j = blockIdx.x * 256ll + threadIdx.x;
  if (j >= m) return;
  const long long i = keep[j];
  long long c = 0;
  for (long long o = off[i]; o < off[i + 1]; ++o) c += (shots[o] >= 0 && shots[o] < n_shots);
  cnt[j] = c;
}

__global__ __launch_bounds__(256) void k_obs_keys(const long long* __restrict__ off,
                                                  const int* __restrict__ shots,
                                                  const int* __restrict__ keep,
                                                  const unsigned int* __restrict__ sp,
                                                  const long long* __restrict__ pos, long long m,
                                                  int n_shots,
                                                  unsigned long long* __restrict__ keys) {
  const long long j = blockIdx.x * 256ll + threadIdx.x;
  if (j >= m) return;
  const long long i = keep[j];
  long long w = pos[j];
  for (long long o = off[i]; o < off[i + 1]; ++o) {
    const int sh = shots[o];
    if (sh >= 0 && sh < n_shots) keys[w++] = (unsigned long long)sp[j] * n_shots + sh;
  }
}

__global__ __launch_bounds__(256) void k_unique(const unsigned long long* __restrict__ keys,
                                                long long t, uint8_t* __restrict__ flag) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i < t) flag[i] = i == 0 || keys[i] != keys[i - 1];
}

__global__ __launch_bounds__(256) void k_vis_write(const unsigned long long* __restrict__ keys,
                                                   const int* __restrict__ index, long long nv,
                                                   int n_shots, int* __restrict__ vis,
                                                   unsigned int* __restrict__ counts) {
  const long long r = blockIdx.x * 256ll + threadIdx.x;
  if (r >= nv) return;
  const unsigned long long key = keys[index[r]];
  vis[r] = (int)(key % (unsigned long long)n_shots);
  atomicAdd(&counts[key / (unsigned long long)n_shots], 1u);
}

// One wave per super-point: every pair (a < b) of its visibility list.
__global__ __launch_bounds__(256) void k_pair_count(const long long* __restrict__ voff,
                                                    const int* __restrict__ vis, long long nsp,
                                                    int n_shots,
                                                    unsigned int* __restrict__ table) {
  const long long s = (blockIdx.x * 256ll + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (s >= nsp) return;
  const long long b = voff[s], e = voff[s + 1];
  for (long long x = b; x < e; ++x)
    for (long long y = x + 1 + lane; y < e; y += 64)
      atomicAdd(&table[(size_t)vis[x] * n_shots + vis[y]], 1u);
}

__global__ __launch_bounds__(256) void k_edge_flag(const unsigned int* __restrict__ table,
                                                   int n_shots, uint8_t* __restrict__ flag) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  const long long nn = (long long)n_shots * n_shots;
  if (i >= nn) return;
  const long long a = i / n_shots, b = i - a * n_shots;
  flag[i] = a < b && table[i] > 0;
}

__global__ __launch_bounds__(256) void k_edge_write(const unsigned int* __restrict__ table,
                                                    const int* __restrict__ index, long long ne,
                                                    int n_shots, int* __restrict__ edges,
                                                    int* __restrict__ weights) {
  const long long r = blockIdx.x * 256ll + threadIdx.x;
  if (r >= ne) return;
  const long long i = index[r];
  edges[2 * r] = (int)(i / n_shots);
  edges[2 * r + 1] = (int)(i % n_shots);
  weights[r] = (int)table[i];
}

// ---- SelectNeighbors (dense_helpers.cc:331-512) -------------------------

__global__ __launch_bounds__(256) void k_pair_score(const double* __restrict__ origins,
                                                    int n_shots, const double* __restrict__ sp,
                                                    const long long* __restrict__ voff,
                                                    const int* __restrict__ vis,
                                                    const int* __restrict__ ntracks,
                                                    long long nsp, double cos_lo, double cos_hi,
                                                    unsigned int* __restrict__ count,
                                                    unsigned int* __restrict__ score,
                                                    long long* __restrict__ sumfx) {
  const long long s = (blockIdx.x * 256ll + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (s >= nsp) return;
  const long long b = voff[s], e = voff[s + 1];
  const double P[3] = {sp[3 * s], sp[3 * s + 1], sp[3 * s + 2]};
  const unsigned int nt = (unsigned int)ntracks[s];
  for (long long x = b; x < e; ++x) {
    const int ca = vis[x];
    if (ca < 0 || ca >= n_shots) continue;
    double u[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = origins[3 * ca + a] - P[a];
    const double un = sqrt(u[0] * u[0] + (u[1] * u[1] + u[2] * u[2]));
    if (un < 1e-12) continue;
    for (long long y = x + 1 + lane; y < e; y += 64) {
      const int cb = vis[y];
      if (cb < 0 || cb >= n_shots) continue;
      double w[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) w[a] = origins[3 * cb + a] - P[a];
      const double wn = sqrt(w[0] * w[0] + (w[1] * w[1] + w[2] * w[2]));
      if (wn < 1e-12) continue;
      double c = __ddiv_rn(u[0] * w[0] + (u[1] * w[1] + u[2] * w[2]), un * wn);
      c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
      const bool good = c > cos_lo && c < cos_hi;
      const long long fx = __double2ll_rn(c * 4294967296.0) * (long long)nt;
      const int lo = ca < cb ? ca : cb, hi = ca < cb ? cb : ca;
      const size_t ij = (size_t)lo * n_shots + hi;
      atomicAdd(&count[ij], nt);
      if (good) atomicAdd(&score[ij], nt);
      atomicAdd((unsigned long long*)&sumfx[ij], (unsigned long long)fx);
    }
  }
}

struct PairView {
  const unsigned int* count;
  const unsigned int* score;
  const long long* sumfx;
  int n;
  __device__ __forceinline__ size_t at(int i, int j) const {
    return i < j ? (size_t)i * n + j : (size_t)j * n + i;
  }
};

// One workgroup per shot: best candidates (score descending, index
// ascending) and the count of the remaining "all" candidates.
__global__ __launch_bounds__(256) void k_select_best(PairView pv, int num_neighbors,
                                                     long long min_best, long long min_all,
                                                     double cos_lo, double cos_hi,
                                                     int* __restrict__ best,
                                                     int* __restrict__ best_n,
                                                     long long* __restrict__ all_n) {
  const int i = blockIdx.x;
  const int n = pv.n;
  __shared__ unsigned long long red[256];
  __shared__ unsigned long long picked_s;
  __shared__ int picked[kMaxNeighbours + 1];
  unsigned long long prev = ~0ull;
  int taken = 0;
  if (threadIdx.x == 0) {
    best[(size_t)i * (num_neighbors + 1)] = i;
    picked[0] = i;
  }
  for (int round = 0; round < num_neighbors; ++round) {
    unsigned long long mx = 0;
    for (int j = threadIdx.x; j < n; j += 256) {
      if (j == i) continue;
      const size_t ij = pv.at(i, j);
      const unsigned int c = pv.count[ij];
      if (c == 0 || (long long)c <= min_best) continue;
      double avg = (double)pv.sumfx[ij] / 4294967296.0 / (double)c;
      avg = avg < -1.0 ? -1.0 : (avg > 1.0 ? 1.0 : avg);
      if (avg > cos_hi || avg < cos_lo) continue;
      const unsigned int sc = pv.score[ij];
      if ((long long)sc <= min_best) continue;
      const unsigned long long key =
          ((unsigned long long)sc << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)j);
      if (key < prev && key > mx) mx = key;
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x])
        red[threadIdx.x] = red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) picked_s = red[0];
    __syncthreads();
    const unsigned long long pk = picked_s;
    __syncthreads();
    if (pk == 0) break;
    prev = pk;
    ++taken;
    if (threadIdx.x == 0) {
      const int j = (int)(0xFFFFFFFFu - (unsigned)(pk & 0xFFFFFFFFu));
      best[(size_t)i * (num_neighbors + 1) + taken] = j;
      picked[taken] = j;
    }
  }
  __syncthreads();
  // "all" extras: count > min_all and not among the best
  long long extra = 0;
  for (int j = threadIdx.x; j < n; j += 256) {
    if (j == i) continue;
    const unsigned int c = pv.count[pv.at(i, j)];
    if (c == 0 || (long long)c <= min_all) continue;
    bool in_best = false;
    for (int t = 1; t <= taken; ++t) in_best |= picked[t] == j;
    extra += !in_best;
  }
  __shared__ long long red2[256];
  red2[threadIdx.x] = extra;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red2[threadIdx.x] += red2[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    best_n[i] = 1 + taken;
    all_n[i] = 1 + taken + red2[0];
  }
}

// One workgroup per shot: the "all" list = best, then the extras in index
// order (a block-wide ordered compaction over the row).
__global__ __launch_bounds__(256) void k_select_all(PairView pv, int num_neighbors,
                                                    long long min_all,
                                                    const int* __restrict__ best,
                                                    const int* __restrict__ best_n,
                                                    const long long* __restrict__ all_off,
                                                    int* __restrict__ all) {
  const int i = blockIdx.x;
  const int n = pv.n;
  const int nb = best_n[i];
  const int* bl = best + (size_t)i * (num_neighbors + 1);
  long long w = all_off[i];
  for (int t = threadIdx.x; t < nb; t += 256) all[w + t] = bl[t];
  w += nb;
  __shared__ int wave_cnt[4];
  __shared__ long long base_s;
  if (threadIdx.x == 0) base_s = w;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j0 = 0; j0 < n; j0 += 256) {
    const int j = j0 + threadIdx.x;
    bool f = false;
    const unsigned int c = j < n ? pv.count[pv.at(i, j)] : 0u;
    if (j < n && j != i && c != 0 && (long long)c > min_all) {
      f = true;
      for (int t = 1; t < nb; ++t) f &= bl[t] != j;
    }
    const unsigned long long m = __ballot(f);
    if (lane == 0) wave_cnt[wave] = __popcll(m);
    __syncthreads();
    int before = 0, total = 0;
    for (int x = 0; x < 4; ++x) {
      before += x < wave ? wave_cnt[x] : 0;
      total += wave_cnt[x];
    }
    const long long base = base_s;
    if (f) all[base + before + __popcll(m & ((1ull << lane) - 1ull))] = j;
    __syncthreads();
    if (threadIdx.x == 0) base_s = base + total;
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_valid(const long long* off, const int* shots, long long n, int n_shots,
                        uint8_t* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_valid, dim3(blocks(n)), dim3(256), 0, s, off, shots, n, n_shots, flag);
  return hipGetLastError();
}

hipError_t launch_flag_count(const uint8_t* flag, long long n, int* counts, hipStream_t s) {
  hipLaunchKernelGGL(k_flag_count, dim3(blocks(n)), dim3(256), 0, s, flag, n, counts);
  return hipGetLastError();
}

hipError_t launch_flag_rank(const uint8_t* flag, long long n, const long long* offsets, int* rank,
                            int* index, hipStream_t s) {
  hipLaunchKernelGGL(k_flag_rank, dim3(blocks(n)), dim3(256), 0, s, flag, n, offsets, rank,
                     index);
  return hipGetLastError();
}

hipError_t launch_gather_points(const double* coords, const int* keep, long long m, float* fc,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_gather_points, dim3(blocks(m)), dim3(256), 0, s, coords, keep, m, fc);
  return hipGetLastError();
}

hipError_t launch_kdist(const int* kidx, const float* kd2, long long m, int k, float* kdist,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_kdist, dim3(blocks(m)), dim3(256), 0, s, kidx, kd2, m, k, kdist);
  return hipGetLastError();
}

hipError_t launch_union_round(const int* kidx, const float* kd2, long long m, int k, float r2,
                              int* parent, int* changed, hipStream_t s) {
  hipLaunchKernelGGL(k_hook, dim3(blocks(m)), dim3(256), 0, s, kidx, kd2, m, k, r2, parent,
                     changed);
  hipLaunchKernelGGL(k_compress, dim3(blocks(m)), dim3(256), 0, s, parent, m);
  return hipGetLastError();
}

hipError_t launch_is_root(const int* parent, long long m, uint8_t* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_is_root, dim3(blocks(m)), dim3(256), 0, s, parent, m, flag);
  return hipGetLastError();
}

hipError_t launch_sp_of(const int* parent, const int* root_rank, long long m, const int* keep,
                        unsigned int* sp, unsigned int* member, int* point_sp,
                        unsigned int* counts, hipStream_t s) {
  hipLaunchKernelGGL(k_sp_of, dim3(blocks(m)), dim3(256), 0, s, parent, root_rank, m, keep, sp,
                     member, point_sp, counts);
  return hipGetLastError();
}

hipError_t launch_centroid(const double* coords, const int* keep, const unsigned int* members,
                           const long long* moff, long long nsp, double* out, int* out_members,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_centroid, dim3(blocks(nsp)), dim3(256), 0, s, coords, keep, members, moff,
                     nsp, out, out_members);
  return hipGetLastError();
}

hipError_t launch_obs_count(const long long* off, const int* shots, const int* keep, long long m,
                            int n_shots, long long* cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_obs_count, dim3(blocks(m)), dim3(256), 0, s, off, shots, keep, m, n_shots,
                     cnt);
  return hipGetLastError();
}

hipError_t launch_obs_keys(const long long* off, const int* shots, const int* keep,
                           const unsigned int* sp, const long long* pos, long long m, int n_shots,
                           unsigned long long* keys, hipStream_t s) {
  hipLaunchKernelGGL(k_obs_keys, dim3(blocks(m)), dim3(256), 0, s, off, shots, keep, sp, pos, m,
                     n_shots, keys);
  return hipGetLastError();
}

hipError_t launch_unique(const unsigned long long* keys, long long t, uint8_t* flag,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_unique, dim3(blocks(t)), dim3(256), 0, s, keys, t, flag);
  return hipGetLastError();
}

hipError_t launch_vis_write(const unsigned long long* keys, const int* index, long long nv,
                            int n_shots, int* vis, unsigned int* counts, hipStream_t s) {
  hipLaunchKernelGGL(k_vis_write, dim3(blocks(nv)), dim3(256), 0, s, keys, index, nv, n_shots,
                     vis, counts);
  return hipGetLastError();
}

hipError_t launch_pair_count(const long long* voff, const int* vis, long long nsp, int n_shots,
                             unsigned int* table, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_count, dim3(blocks(nsp * 64)), dim3(256), 0, s, voff, vis, nsp,
                     n_shots, table);
  return hipGetLastError();
}

hipError_t launch_edge_flag(const unsigned int* table, int n_shots, uint8_t* flag,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_edge_flag, dim3(blocks((long long)n_shots * n_shots)), dim3(256), 0, s,
                     table, n_shots, flag);
  return hipGetLastError();
}

hipError_t launch_edge_write(const unsigned int* table, const int* index, long long ne,
                             int n_shots, int* edges, int* weights, hipStream_t s) {
  hipLaunchKernelGGL(k_edge_write, dim3(blocks(ne)), dim3(256), 0, s, table, index, ne, n_shots,
                     edges, weights);
  return hipGetLastError();
}

hipError_t launch_pair_score(const double* origins, int n_shots, const double* sp,
                             const long long* voff, const int* vis, const int* ntracks,
                             long long nsp, double cos_lo, double cos_hi, unsigned int* count,
                             unsigned int* score, long long* sumfx, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_score, dim3(blocks(nsp * 64)), dim3(256), 0, s, origins, n_shots, sp,
                     voff, vis, ntracks, nsp, cos_lo, cos_hi, count, score, sumfx);
  return hipGetLastError();
}

hipError_t launch_select(const unsigned int* count, const unsigned int* score,
                         const long long* sumfx, int n_shots, int num_neighbors,
                         long long min_best, long long min_all, double cos_lo,
                         double cos_hi, int* best, int* best_n, long long* all_n,
                         hipStream_t s) {
  PairView pv{count, score, sumfx, n_shots};
  hipLaunchKernelGGL(k_select_best, dim3(n_shots), dim3(256), 0, s, pv, num_neighbors, min_best,
                     min_all, cos_lo, cos_hi, best, best_n, all_n);
  return hipGetLastError();
}

hipError_t launch_select_all(const unsigned int* count, const unsigned int* score,
                             const long long* sumfx, int n_shots, int num_neighbors,
                             long long min_all, const int* best, const int* best_n,
                             const long long* all_off, int* all, hipStream_t s) {
  PairView pv{count, score, sumfx, n_shots};
  hipLaunchKernelGGL(k_select_all, dim3(n_shots), dim3(256), 0, s, pv, num_neighbors, min_all,
                     best, best_n, all_off, all);
  return hipGetLastError();
}

hipError_t sort_u32_pairs(unsigned int* keys_in, unsigned int* keys_out, unsigned int* vals_in,
                          unsigned int* vals_out, int n, int end_bit, void* temp,
                          size_t* temp_bytes, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, keys_in, keys_out, vals_in,
                                            vals_out, n, 0, end_bit, s);
}

hipError_t sort_u64_keys(unsigned long long* keys_in, unsigned long long* keys_out, int n,
                         int end_bit, void* temp, size_t* temp_bytes, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortKeys(temp, *temp_bytes, keys_in, keys_out, n, 0, end_bit,
                                           s);
}

}  // namespace cv
