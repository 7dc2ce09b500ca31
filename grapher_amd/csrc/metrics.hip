// metrics.hip -- layout quality on the device: normalised stress over BFS
// distances from a set of pivot vertices, the judge of the warm start.
//
// stress = 1 - A^2 / (B C) with, over every pivot s and every vertex j
// reached from s (j != s), D = |x_j - x_s|, d = L * hops(s, j):
// A = sum D / d, B = sum D^2 / d^2, C = number of pairs. This is the
// weighted stress sum (s D - d)^2 / d^2 at its optimal scale s, divided by
// its value at s = 0; it ignores the layout's scale, rotation and
// translation.
#include "../../include/gviz_hip.h"
#include "common.hpp"

#include <vector>

using namespace gvh;

namespace {

constexpr int kB = 256;

// One BFS level for every pivot at once: a vertex still unreached whose
// neighbour sits at `level` is at level + 1. Reads dist[.] == level only, so
// the writes of this pass (level + 1) never feed back into it.
__global__ __launch_bounds__(kB) void k_bfs_level(const int32_t *__restrict__ off,
                                                  const int32_t *__restrict__ adj,
                                                  int32_t *__restrict__ dist, int64_t n,
                                                  int64_t pivots, int32_t level,
                                                  int32_t *__restrict__ changed) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n * pivots)
    return;
  const int64_t s = i / n, v = i % n;
  int32_t *d = dist + s * n;
  if (d[v] >= 0)
    return;
  for (int32_t e = off[v]; e < off[v + 1]; e++)
    if (d[adj[e]] == level) {
      d[v] = level + 1;
      *changed = 1;
      return;
    }
}

__global__ __launch_bounds__(kB) void k_stress_terms(const int32_t *__restrict__ dist,
                                                     const int32_t *__restrict__ piv,
                                                     const double2 *__restrict__ pos,
                                                     int64_t n, int64_t pivots,
                                                     double *__restrict__ part) {
  __shared__ double sm[kB / kWave][3];
  double A = 0.0, B = 0.0, C = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n * pivots;
       i += (int64_t)gridDim.x * kB) {
    const int64_t s = i / n, v = i % n;
    const int32_t h = dist[i];
    if (h <= 0)
      continue;
    const double2 a = pos[piv[s]], b = pos[v];
    const double dx = b.x - a.x, dy = b.y - a.y;
    const double D = sqrt(dx * dx + dy * dy);
    const double d = (double)h;
    A += D / d;
    B += (D * D) / (d * d);
    C += 1.0;
  }
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  A = wave_sum(A);
  B = wave_sum(B);
  C = wave_sum(C);
  if (lane == 0) {
    sm[wave][0] = A;
    sm[wave][1] = B;
    sm[wave][2] = C;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int w = 0; w < kB / kWave; w++)
      t += sm[w][threadIdx.x];
    part[blockIdx.x * 3 + threadIdx.x] = t;
  }
}

} // namespace

struct gvizhip_stress {
  int64_t n = 0, pivots = 0;
  hipStream_t stream = nullptr;
  DevBuf<int32_t> off, adj, dist, piv, changed;
  DevBuf<double> pos, part;
  int blocks = 1;
};

extern "C" {

int gvizhip_stress_create(gvizhip_stress **out, int64_t n, const int64_t *offsets,
                          const int64_t *adj, const int64_t *pivots, int64_t count) {
  if (!out || n <= 0 || !offsets || !pivots || count <= 0 || n > INT32_MAX ||
      offsets[n] > INT32_MAX || (offsets[n] > 0 && !adj)) {
    set_error("gvizhip_stress_create: invalid arguments");
    return -1;
  }
  *out = nullptr;
  int dev = 0;
  if (hipGetDeviceCount(&dev) != hipSuccess || dev <= 0) {
    set_error("gvizhip_stress_create: no GPU");
    return -1;
  }
  const int64_t m = offsets[n];
  std::vector<int32_t> o(n + 1), a(m > 0 ? m : 1), p(count);
  for (int64_t v = 0; v <= n; v++)
    o[v] = (int32_t)offsets[v];
  for (int64_t e = 0; e < m; e++) {
    if (adj[e] < 0 || adj[e] >= n) {
      set_error("gvizhip_stress_create: neighbour id out of range");
      return -1;
    }
    a[e] = (int32_t)adj[e];
  }
  for (int64_t i = 0; i < count; i++) {
    if (pivots[i] < 0 || pivots[i] >= n) {
      set_error("gvizhip_stress_create: pivot out of range");
      return -1;
    }
    p[i] = (int32_t)pivots[i];
  }
  auto *s = new gvizhip_stress;
  s->n = n;
  s->pivots = count;
  s->blocks = 1024;
  if (hipStreamCreate(&s->stream) != hipSuccess || s->off.alloc(n + 1) ||
      s->adj.alloc(m > 0 ? m : 1) || s->dist.alloc(n * count) || s->piv.alloc(count) ||
      s->changed.alloc(1) || s->pos.alloc(2 * n) || s->part.alloc(3 * s->blocks)) {
    gvizhip_stress_destroy(s);
    return -1;
  }
  GVH_CHECK(hipMemcpyAsync(s->off.p, o.data(), sizeof(int32_t) * (n + 1),
                           hipMemcpyHostToDevice, s->stream));
  if (m > 0)
    GVH_CHECK(hipMemcpyAsync(s->adj.p, a.data(), sizeof(int32_t) * m,
                             hipMemcpyHostToDevice, s->stream));
  GVH_CHECK(hipMemcpyAsync(s->piv.p, p.data(), sizeof(int32_t) * count,
                           hipMemcpyHostToDevice, s->stream));
  // dist = -1 everywhere, 0 at each pivot
  std::vector<int32_t> init(n, -1);
  for (int64_t i = 0; i < count; i++) {
    init[p[i]] = 0;
    GVH_CHECK(hipMemcpyAsync(s->dist.p + i * n, init.data(), sizeof(int32_t) * n,
                             hipMemcpyHostToDevice, s->stream));
    GVH_CHECK(hipStreamSynchronize(s->stream));
    init[p[i]] = -1;
  }
  for (int32_t level = 0; level < INT32_MAX; level++) {
    int32_t zero = 0, changed = 0;
    GVH_CHECK(hipMemcpyAsync(s->changed.p, &zero, sizeof(int32_t),
                             hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_bfs_level, dim3(grid_for(n * count, kB)), dim3(kB), 0,
                       s->stream, s->off.p, s->adj.p, s->dist.p, n, count, level,
                       s->changed.p);
    GVH_CHECK_LAUNCH();
    GVH_CHECK(hipMemcpyAsync(&changed, s->changed.p, sizeof(int32_t),
                             hipMemcpyDeviceToHost, s->stream));
    GVH_CHECK(hipStreamSynchronize(s->stream));
    if (!changed)
      break;
  }
  *out = s;
  return 0;
}

int gvizhip_stress_evaluate(gvizhip_stress *s, const double *positions,
                            double *stress) {
  if (!s || !positions || !stress) {
    set_error("gvizhip_stress_evaluate: invalid arguments");
    return -1;
  }
  GVH_CHECK(hipMemcpyAsync(s->pos.p, positions, sizeof(double) * 2 * s->n,
                           hipMemcpyHostToDevice, s->stream));
  hipLaunchKernelGGL(k_stress_terms, dim3(s->blocks), dim3(kB), 0, s->stream, s->dist.p,
                     s->piv.p, (const double2 *)s->pos.p, s->n, s->pivots, s->part.p);
  GVH_CHECK_LAUNCH();
  std::vector<double> part(3 * s->blocks);
  GVH_CHECK(hipMemcpyAsync(part.data(), s->part.p, sizeof(double) * 3 * s->blocks,
                           hipMemcpyDeviceToHost, s->stream));
  GVH_CHECK(hipStreamSynchronize(s->stream));
  double A = 0.0, B = 0.0, C = 0.0;
  for (int b = 0; b < s->blocks; b++) {
    A += part[3 * b];
    B += part[3 * b + 1];
    C += part[3 * b + 2];
  }
  *stress = (C == 0.0 || B == 0.0) ? 0.0 : 1.0 - (A * A) / (B * C);
  return 0;
}

void gvizhip_stress_destroy(gvizhip_stress *s) {
  if (!s)
    return;
  if (s->stream) {
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamDestroy(s->stream);
  }
  delete s;
}

} // extern "C"
