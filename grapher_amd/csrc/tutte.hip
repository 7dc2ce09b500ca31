// tutte.hip -- Jacobi sweep of the Tutte barycentric embedder
// (src/embedders/gvizTutteEmbedder.c:249-335) on gfx950.
//
// One thread per vertex reads its CSR row (insertion order), sums the
// neighbours' previous positions, blends towards the barycentre with
// relaxation alpha and writes the next positions into the other half of a
// ping-pong pair; pinned boundary vertices and isolated vertices are copied.
// Each sweep's max displacement is reduced per block and finished by a
// one-block kernel that records it and raises the stop flag; in a run of
// Jacobi sweeps the first workgroup of the next sweep finishes it instead
// (k_sweep<..., true>), which saves a launch per sweep. Compiled with
// -ffp-contract=off: in FP64 every sum and product rounds as the reference's.
//
// The spring variant (src/embedders/gvizSpringTutteEmbedder.c:289-358) runs
// the same row gather in k_spring and integrates a damped spring towards the
// barycentre with a per-vertex velocity kept in HBM (semi-implicit Euler, in
// the reference's operation order).
//
// Vertices are stored in breadth-first order (relabelled at create, mapped
// back at upload/download), so the neighbours a thread gathers sit close in
// memory. Each row keeps its insertion order, and the sweep's only
// reduction is a max, so the results are those of the original order.
#include "../../include/gviz_hip.h"
#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

using namespace gvh;

namespace {

constexpr int kB = 256;

struct TutteScalars {
  int64_t sweeps;    // sweeps applied (and finished) since upload
  int64_t recordBase;
  double epsilon;
  int32_t stopped;
  int32_t pending;   // a fused sweep ran whose maximum is not finished yet
};

template <class T> struct V2;
template <> struct V2<float> { using type = float2; };
template <> struct V2<double> { using type = double2; };

// Sum of the neighbours' positions over one row, in row order; the
// positions are loaded kGather at a time ahead of the adds so several
// gathers are in flight per thread.
#ifndef GVH_TUTTE_GATHER
#define GVH_TUTTE_GATHER 8
#endif
constexpr int kGather = GVH_TUTTE_GATHER;
template <class Real>
__device__ __forceinline__ void row_sum(const int32_t *__restrict__ adj,
                                        const typename V2<Real>::type *__restrict__ in,
                                        int32_t b, int32_t e, Real &bx, Real &by) {
  typename V2<Real>::type p[kGather];
  for (int32_t k0 = b; k0 < e; k0 += kGather) {
    const int m = e - k0 < kGather ? e - k0 : kGather;
#pragma unroll
    for (int k = 0; k < kGather; k++)
      if (k < m)
        p[k] = in[adj[k0 + k]];
#pragma unroll
    for (int k = 0; k < kGather; k++)
      if (k < m) {
        bx += p[k].x;
        by += p[k].y;
      }
  }
}

// Max over `blocks` partial maxima by one workgroup (every thread calls it;
// the result is valid in thread 0).
__device__ double max_of_partials(const double *__restrict__ partial, int blocks,
                                  double *sm) {
  double m = 0.0;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x)
    m = partial[b] > m ? partial[b] : m;
  m = wave_max(m);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0)
    sm[wave] = m;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x / kWave); w++)
      r = sm[w] > r ? sm[w] : r;
  __syncthreads();
  return r;
}

// Thread 0 of the finishing workgroup: the "tutte.maxDelta" record of sweep
// sc->sweeps, the count, and the stop flag of gvizTutteEmbedderRun
// (converged once maxDelta < epsilon).
__device__ void finish_sweep(TutteScalars *sc, double r, double *rec, int64_t recCap) {
  const int64_t slot = sc->sweeps - sc->recordBase;
  if (slot >= 0 && slot < recCap)
    rec[slot] = r;
  sc->sweeps++;
  if (sc->epsilon >= 0.0 && r < sc->epsilon)
    sc->stopped = 1;
}

// FUSED: a sweep of a run, whose partial maxima go to the half `slot` of
// `partial` (2 x gridDim.x); its first workgroup first finishes the sweep
// before it (in the half sc->sweeps & 1) if one is pending, then marks this
// one pending unless that finish stopped the run. A workgroup that starts
// after the stop flag was raised returns, so a sweep running when the run
// stops may be applied in part; it only writes the buffer the sweep before
// did not, and sc->sweeps keeps pointing at that one.
// The kB-row tiles [t0, t1) a workgroup sweeps: a contiguous range per
// workgroup, and the workgroups of one XCD (they are dealt out in turn,
// blockIdx mod 8) take a contiguous eighth of the ranges, so rows in
// breadth-first order read their neighbours' positions mostly through their
// own XCD's L2.
__device__ __forceinline__ void tile_range(int64_t n, int64_t &t0, int64_t &t1) {
  const int64_t G = gridDim.x, bi = blockIdx.x;
  const int64_t q = G >> 3, r = G & 7, x = bi & 7;
  const int64_t w = x * q + (x < r ? x : r) + (bi >> 3);
  const int64_t tiles = (n + kB - 1) / kB;
  t0 = w * tiles / G;
  t1 = (w + 1) * tiles / G;
}

template <class Real, bool FUSED>
__global__ __launch_bounds__(kB) void k_sweep(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const uint8_t *__restrict__ isBoundary,
    const typename V2<Real>::type *__restrict__ in,
    typename V2<Real>::type *__restrict__ out, int64_t n, Real alpha,
    double *__restrict__ partial, TutteScalars *__restrict__ sc, int slot,
    double *__restrict__ rec, int64_t recCap) {
  __shared__ double sm[16];
  if constexpr (FUSED) {
    if (blockIdx.x == 0) {
      __shared__ int32_t sStop;
      if (sc->pending && !sc->stopped) {
        const double r =
            max_of_partials(partial + (sc->sweeps & 1) * gridDim.x, (int)gridDim.x, sm);
        if (threadIdx.x == 0) {
          finish_sweep(sc, r, rec, recCap);
          sc->pending = 0;
        }
      }
      if (threadIdx.x == 0) {
        sStop = sc->stopped;
        if (!sStop)
          sc->pending = 1;
      }
      __syncthreads();
      if (sStop)
        return;
    } else if (sc->stopped) {
      return;
    }
    partial += slot * gridDim.x;
  } else if (sc->stopped) {
    return;
  }
  double vmax = 0.0;
  int64_t t0, t1;
  tile_range(n, t0, t1);
  for (int64_t t = t0; t < t1; t++) {
    const int64_t u = t * kB + threadIdx.x;
    if (u >= n)
      continue;
    typename V2<Real>::type old = in[u];
    const int32_t b = off[u], e = off[u + 1];
    if (isBoundary[u] || e == b) {
      out[u] = old;
      continue;
    }
    Real bx = 0, by = 0;
    row_sum<Real>(adj, in, b, e, bx, by);
    const Real cnt = (Real)(e - b);
    bx /= cnt;
    by /= cnt;
    Real nx = old.x + alpha * (bx - old.x);
    Real ny = old.y + alpha * (by - old.y);
    Real d0 = nx - old.x, d1 = ny - old.y;
    Real delta = (Real)0;
    delta += d0 * d0;
    delta += d1 * d1;
    double d = (double)sqrt(delta);
    vmax = d > vmax ? d : vmax;
    out[u] = {nx, ny};
  }
  vmax = wave_max(vmax);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0)
    sm[wave] = vmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < kB / kWave; w++)
      m = sm[w] > m ? sm[w] : m;
    partial[blockIdx.x] = m;
  }
}

// Gauss-Seidel (src/embedders/gvizTutteEmbedder.c:258-262, 305-320 with
// useGaussSeidel): the reference relaxes the interior vertices one at a time
// in ascending id, each reading the current positions, so a vertex sees the
// new positions of lower-id neighbours and the old ones of higher-id
// neighbours. Level l holds the vertices whose lower-id interior neighbours
// all sit in levels < l (level = 1 + the highest such neighbour's level), so
// two vertices of one level are never adjacent: a sweep copies the positions
// into the other buffer and relaxes it in place one level per launch, which
// reads exactly what the sequential sweep reads.
template <class Real>
__global__ __launch_bounds__(kB) void k_gs_copy(const typename V2<Real>::type *__restrict__ in,
                                                typename V2<Real>::type *__restrict__ out,
                                                int64_t n, const TutteScalars *__restrict__ sc) {
  if (sc->stopped)
    return;
  for (int64_t u = (int64_t)blockIdx.x * kB + threadIdx.x; u < n;
       u += (int64_t)gridDim.x * kB)
    out[u] = in[u];
}

template <class Real>
__global__ __launch_bounds__(kB) void k_gs_level(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const int32_t *__restrict__ order, int32_t count, typename V2<Real>::type *pos,
    Real alpha, double *__restrict__ partial, const TutteScalars *__restrict__ sc) {
  __shared__ double sm[16];
  if (sc->stopped)
    return;
  double vmax = 0.0;
  const int32_t i = (int32_t)(blockIdx.x * kB + threadIdx.x);
  if (i < count) {
    const int32_t u = order[i];
    const typename V2<Real>::type old = pos[u];
    const int32_t b = off[u], e = off[u + 1];
    Real bx = 0, by = 0;
    row_sum<Real>(adj, pos, b, e, bx, by);
    const Real cnt = (Real)(e - b);
    bx /= cnt;
    by /= cnt;
    Real nx = old.x + alpha * (bx - old.x);
    Real ny = old.y + alpha * (by - old.y);
    Real d0 = nx - old.x, d1 = ny - old.y;
    Real delta = (Real)0;
    delta += d0 * d0;
    delta += d1 * d1;
    vmax = (double)sqrt(delta);
    pos[u] = {nx, ny};
  }
  vmax = wave_max(vmax);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0)
    sm[wave] = vmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < kB / kWave; w++)
      m = sm[w] > m ? sm[w] : m;
    partial[blockIdx.x] = m;
  }
}

// One spring step (src/embedders/gvizSpringTutteEmbedder.c:313-330): for an
// interior vertex with neighbours, a = k (bary - p) - c v, v += a dt,
// p' = p + v dt, per coordinate in that order; pinned and isolated vertices
// keep position and velocity. Each thread owns its vertex's velocity.
template <class Real>
__global__ __launch_bounds__(kB) void k_spring(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const uint8_t *__restrict__ isBoundary,
    const typename V2<Real>::type *__restrict__ in,
    typename V2<Real>::type *__restrict__ out, typename V2<Real>::type *__restrict__ vel,
    int64_t n, Real dt, Real stiffness, Real damping, double *__restrict__ partial,
    const TutteScalars *__restrict__ sc) {
  __shared__ double sm[16];
  if (sc->stopped)
    return;
  double vmax = 0.0;
  int64_t t0, t1;
  tile_range(n, t0, t1);
  for (int64_t t = t0; t < t1; t++) {
    const int64_t u = t * kB + threadIdx.x;
    if (u >= n)
      continue;
    typename V2<Real>::type old = in[u];
    const int32_t b = off[u], e = off[u + 1];
    if (isBoundary[u] || e == b) {
      out[u] = old;
      continue;
    }
    Real bx = 0, by = 0;
    row_sum<Real>(adj, in, b, e, bx, by);
    const Real cnt = (Real)(e - b);
    bx /= cnt;
    by /= cnt;
    typename V2<Real>::type v = vel[u];
    const Real ax = stiffness * (bx - old.x) - damping * v.x;
    v.x += ax * dt;
    const Real nx = old.x + v.x * dt;
    const Real d0 = nx - old.x;
    const Real ay = stiffness * (by - old.y) - damping * v.y;
    v.y += ay * dt;
    const Real ny = old.y + v.y * dt;
    const Real d1 = ny - old.y;
    Real delta = (Real)0;
    delta += d0 * d0;
    delta += d1 * d1;
    double d = (double)sqrt(delta);
    vmax = d > vmax ? d : vmax;
    vel[u] = v;
    out[u] = {nx, ny};
  }
  vmax = wave_max(vmax);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0)
    sm[wave] = vmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < kB / kWave; w++)
      m = sm[w] > m ? sm[w] : m;
    partial[blockIdx.x] = m;
  }
}

// Finishes a sweep: max over blocks, the "tutte.maxDelta" record, the stop
// flag of gvizTutteEmbedderRun (converged once maxDelta < epsilon).
__global__ void k_finish(const double *__restrict__ partial, int blocks,
                         TutteScalars *__restrict__ sc, double *__restrict__ rec,
                         int64_t recCap) {
  __shared__ double sm[16];
  if (sc->stopped)
    return;
  const double r = max_of_partials(partial, blocks, sm);
  if (threadIdx.x == 0)
    finish_sweep(sc, r, rec, recCap);
}

// Finishes the last fused sweep of a run, if it is still pending.
__global__ void k_finish_pending(const double *__restrict__ partial, int blocks,
                                 TutteScalars *__restrict__ sc, double *__restrict__ rec,
                                 int64_t recCap) {
  __shared__ double sm[16];
  if (!sc->pending || sc->stopped)
    return;
  const double r = max_of_partials(partial + (sc->sweeps & 1) * blocks, blocks, sm);
  if (threadIdx.x == 0) {
    finish_sweep(sc, r, rec, recCap);
    sc->pending = 0;
  }
}

__global__ void k_begin(TutteScalars *sc, double epsilon) {
  sc->recordBase = sc->sweeps;
  sc->epsilon = epsilon;
  sc->stopped = 0;
  sc->pending = 0;
}

// ---------------------------------------------------------------- no-alloc
// Scratch of gvizhip_tutte_sweeps_noalloc: the step scalars in the first
// 64 bytes, then one partial maximum per block of the sweep's grid. Each GPU
// holds its own copy; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device.
constexpr int kNoAllocBlocks = 2048;
constexpr int kNoAllocHeader = 64 / sizeof(double);
static_assert(sizeof(TutteScalars) <= 64, "step scalars exceed the scratch header");
static __device__ __align__(256) double scratch[kNoAllocHeader + kNoAllocBlocks];

__global__ void k_noalloc_begin(TutteScalars *sc, double epsilon) {
  sc->sweeps = 0;
  sc->recordBase = 0;
  sc->epsilon = epsilon;
  sc->stopped = 0;
  sc->pending = 0;
}

// After the sweeps: an odd count left the latest positions in b; copy them
// to a so the result is always in a. Reports the count.
__global__ __launch_bounds__(kB) void k_noalloc_end(const TutteScalars *__restrict__ sc,
                                                    double2 *__restrict__ a,
                                                    const double2 *__restrict__ b,
                                                    int64_t n, int64_t *sweepsRun) {
  const int64_t done = sc->sweeps;
  if (blockIdx.x == 0 && threadIdx.x == 0 && sweepsRun)
    *sweepsRun = done;
  if ((done & 1) == 0)
    return;
  for (int64_t u = (int64_t)blockIdx.x * kB + threadIdx.x; u < n;
       u += (int64_t)gridDim.x * kB)
    a[u] = b[u];
}

} // namespace

struct gvizhip_tutte {
  int64_t n = 0, m2 = 0;
  int32_t precision = GVIZHIP_PRECISION_FP64;
  hipStream_t stream = nullptr;
  DevBuf<int32_t> off, adj;
  DevBuf<uint8_t> boundary;
  DevBuf<unsigned char> pos[2];
  DevBuf<double> partial;
  DevBuf<double> records;
  int64_t recordCap = 0;
  DevBuf<TutteScalars> sc;
  int blocks = 1;
  DevBuf<unsigned char> vel; // spring variant: per-vertex velocity
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  double graphKey[4] = {-1.0, 0.0, 0.0, 0.0}; // step kind and its parameters
  TutteScalars *pinned = nullptr;
  std::vector<int32_t> perm; // stored position -> caller's vertex id
  // host copies of the stored CSR and boundary flags, for the Gauss-Seidel
  // schedule (rebuilt after an upload changes the boundary)
  std::vector<int32_t> hostOff, hostAdj;
  std::vector<uint8_t> hostBoundary;
  bool gsDirty = true;
  DevBuf<int32_t> gsOrder;       // stored ids, level by level
  std::vector<int32_t> gsStart;  // level -> first entry of gsOrder (levels + 1)
  std::vector<int64_t> gsSlot;   // level -> its first partial-max slot
  DevBuf<double> gsPartial;      // one slot per (level, block)
  int64_t gsSlots = 0;
};

namespace {

size_t elem(const gvizhip_tutte *t) {
  return t->precision == GVIZHIP_PRECISION_FP64 ? sizeof(double2) : sizeof(float2);
}

// What one step is: kind 0 a Jacobi sweep with relaxation a; kind 1 a
// spring step with dt a, stiffness b, damping c.
struct StepKind {
  int kind;
  double a, b, c;
};

template <class Real> void enqueue_gs(gvizhip_tutte *t, int from, Real alpha) {
  using P = typename V2<Real>::type;
  hipLaunchKernelGGL(k_gs_copy<Real>, dim3((unsigned)t->blocks), dim3(kB), 0, t->stream,
                     (const P *)t->pos[from].p, (P *)t->pos[1 - from].p, t->n, t->sc.p);
  const int levels = (int)t->gsStart.size() - 1;
  for (int l = 0; l < levels; l++) {
    const int32_t cnt = t->gsStart[l + 1] - t->gsStart[l];
    hipLaunchKernelGGL(k_gs_level<Real>, dim3((unsigned)grid_for(cnt, kB)), dim3(kB), 0,
                       t->stream, t->off.p, t->adj.p, t->gsOrder.p + t->gsStart[l], cnt,
                       (P *)t->pos[1 - from].p, alpha, t->gsPartial.p + t->gsSlot[l],
                       t->sc.p);
  }
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, t->stream, t->gsPartial.p,
                     (int)t->gsSlots, t->sc.p, t->records.p, t->recordCap);
}

int enqueue_step(gvizhip_tutte *t, int from, const StepKind &k) {
  const unsigned g = (unsigned)t->blocks;
  const bool f64 = t->precision == GVIZHIP_PRECISION_FP64;
  if (k.kind == 2) {
    if (f64)
      enqueue_gs<double>(t, from, k.a);
    else
      enqueue_gs<float>(t, from, (float)k.a);
    GVH_CHECK_LAUNCH();
    return 0;
  }
  if (k.kind == 3) {
    // a fused Jacobi sweep: finished by the next one or k_finish_pending
    if (f64)
      hipLaunchKernelGGL((k_sweep<double, true>), dim3(g), dim3(kB), 0, t->stream, t->off.p,
                         t->adj.p, t->boundary.p, (const double2 *)t->pos[from].p,
                         (double2 *)t->pos[1 - from].p, t->n, k.a, t->partial.p, t->sc.p,
                         from, t->records.p, t->recordCap);
    else
      hipLaunchKernelGGL((k_sweep<float, true>), dim3(g), dim3(kB), 0, t->stream, t->off.p,
                         t->adj.p, t->boundary.p, (const float2 *)t->pos[from].p,
                         (float2 *)t->pos[1 - from].p, t->n, (float)k.a, t->partial.p,
                         t->sc.p, from, t->records.p, t->recordCap);
    GVH_CHECK_LAUNCH();
    return 0;
  }
  if (k.kind == 0 && f64)
    hipLaunchKernelGGL((k_sweep<double, false>), dim3(g), dim3(kB), 0, t->stream, t->off.p,
                       t->adj.p, t->boundary.p, (const double2 *)t->pos[from].p,
                       (double2 *)t->pos[1 - from].p, t->n, k.a, t->partial.p, t->sc.p, 0,
                       nullptr, (int64_t)0);
  else if (k.kind == 0)
    hipLaunchKernelGGL((k_sweep<float, false>), dim3(g), dim3(kB), 0, t->stream, t->off.p,
                       t->adj.p, t->boundary.p, (const float2 *)t->pos[from].p,
                       (float2 *)t->pos[1 - from].p, t->n, (float)k.a, t->partial.p,
                       t->sc.p, 0, nullptr, (int64_t)0);
  else if (f64)
    hipLaunchKernelGGL(k_spring<double>, dim3(g), dim3(kB), 0, t->stream, t->off.p,
                       t->adj.p, t->boundary.p, (const double2 *)t->pos[from].p,
                       (double2 *)t->pos[1 - from].p, (double2 *)t->vel.p, t->n, k.a, k.b,
                       k.c, t->partial.p, t->sc.p);
  else
    hipLaunchKernelGGL(k_spring<float>, dim3(g), dim3(kB), 0, t->stream, t->off.p,
                       t->adj.p, t->boundary.p, (const float2 *)t->pos[from].p,
                       (float2 *)t->pos[1 - from].p, (float2 *)t->vel.p, t->n,
                       (float)k.a, (float)k.b, (float)k.c, t->partial.p, t->sc.p);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, t->stream, t->partial.p,
                     (int)g, t->sc.p, t->records.p, t->recordCap);
  GVH_CHECK_LAUNCH();
  return 0;
}

void drop_graph(gvizhip_tutte *t) {
  if (t->exec)
    (void)hipGraphExecDestroy(t->exec);
  if (t->graph)
    (void)hipGraphDestroy(t->graph);
  t->exec = nullptr;
  t->graph = nullptr;
  t->graphKey[0] = -1.0;
}

// Two steps (buffer 0 -> 1 -> 0) captured once per step kind and parameters.
int build_graph(gvizhip_tutte *t, const StepKind &k) {
  const double key[4] = {(double)k.kind, k.a, k.b, k.c};
  if (t->exec && memcmp(key, t->graphKey, sizeof key) == 0)
    return 0;
  drop_graph(t);
  GVH_CHECK(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_step(t, 0, k);
  if (rc == 0)
    rc = enqueue_step(t, 1, k);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(t->stream, &g);
  if (rc || e != hipSuccess) {
    if (g)
      (void)hipGraphDestroy(g);
    if (!rc)
      set_error("stream capture of the Tutte step failed: %s", hipGetErrorString(e));
    return -1;
  }
  t->graph = g;
  GVH_CHECK(hipGraphInstantiate(&t->exec, g, nullptr, nullptr, 0));
  memcpy(t->graphKey, key, sizeof key);
  return 0;
}

int read_sc(gvizhip_tutte *t, TutteScalars *out) {
  GVH_CHECK(hipMemcpyAsync(t->pinned, t->sc.p, sizeof(TutteScalars),
                           hipMemcpyDeviceToHost, t->stream));
  GVH_CHECK(hipStreamSynchronize(t->stream));
  *out = *t->pinned;
  return 0;
}

// The Gauss-Seidel levels over the stored vertices: in ascending caller id
// (the reference's order), an interior vertex with neighbours gets one more
// than the highest level of its interior lower-id neighbours.
int build_gs_schedule(gvizhip_tutte *t) {
  if (!t->gsDirty)
    return 0;
  const int64_t n = t->n;
  std::vector<int32_t> inv(n), level(n, -1);
  for (int64_t i = 0; i < n; i++)
    inv[t->perm[i]] = (int32_t)i;
  int32_t levels = 0;
  auto active = [&](int32_t i) {
    return !t->hostBoundary[i] && t->hostOff[i + 1] > t->hostOff[i];
  };
  for (int64_t ou = 0; ou < n; ou++) {
    const int32_t i = inv[ou];
    if (!active(i))
      continue;
    int32_t l = 0;
    for (int32_t e = t->hostOff[i]; e < t->hostOff[i + 1]; e++) {
      const int32_t j = t->hostAdj[e];
      if (active(j) && t->perm[j] < ou && level[j] + 1 > l)
        l = level[j] + 1;
    }
    level[i] = l;
    levels = l + 1 > levels ? l + 1 : levels;
  }
  t->gsStart.assign(levels + 1, 0);
  for (int64_t i = 0; i < n; i++)
    if (level[i] >= 0)
      t->gsStart[level[i] + 1]++;
  for (int32_t l = 0; l < levels; l++)
    t->gsStart[l + 1] += t->gsStart[l];
  std::vector<int32_t> order(t->gsStart[levels] > 0 ? t->gsStart[levels] : 1);
  std::vector<int32_t> at(t->gsStart.begin(), t->gsStart.end() - 1);
  for (int64_t ou = 0; ou < n; ou++) {
    const int32_t i = inv[ou];
    if (level[i] >= 0)
      order[at[level[i]]++] = i;
  }
  t->gsSlot.assign(levels + 1, 0);
  for (int32_t l = 0; l < levels; l++)
    t->gsSlot[l + 1] = t->gsSlot[l] + grid_for(t->gsStart[l + 1] - t->gsStart[l], kB);
  t->gsSlots = t->gsSlot[levels];
  if (t->gsOrder.alloc(order.size()) || t->gsPartial.alloc(t->gsSlots > 0 ? t->gsSlots : 1))
    return -1;
  GVH_CHECK(hipMemcpy(t->gsOrder.p, order.data(), sizeof(int32_t) * order.size(),
                      hipMemcpyHostToDevice));
  drop_graph(t); // a captured Gauss-Seidel step holds the old schedule
  t->gsDirty = false;
  return 0;
}

// `iters` steps of kind k from the current buffer, two per graph replay,
// stopping once a step's max delta falls below epsilon (epsilon >= 0).
// Returns the steps applied; maxDeltas (optional) receives their records.
int64_t run_steps(gvizhip_tutte *t, int64_t iters, double epsilon, double *maxDeltas,
                  const StepKind &k) {
  if (iters == 0 || t->n == 0)
    return 0;
  if (maxDeltas && iters > t->recordCap) {
    if (t->records.alloc(iters))
      return -1;
    t->recordCap = iters;
    drop_graph(t);
  }
  TutteScalars before;
  if (read_sc(t, &before))
    return -1;
  hipLaunchKernelGGL(k_begin, dim3(1), dim3(1), 0, t->stream, t->sc.p, epsilon);
  GVH_CHECK_LAUNCH();
  int64_t done = 0;
  const int cur = (int)(before.sweeps & 1);
  // Jacobi sweeps run fused (each finished by the next); the other kinds
  // keep their own finish kernel.
  const StepKind kr = k.kind == 0 ? StepKind{3, k.a, 0.0, 0.0} : k;
  // An odd starting parity takes one plain step so the two-step graph
  // always starts from buffer 0.
  if (cur == 1) {
    if (enqueue_step(t, 1, kr))
      return -1;
    done++;
  }
  if (iters - done >= 2 && build_graph(t, kr))
    return -1;
  int64_t sinceCheck = 0;
  while (iters - done >= 2) {
    GVH_CHECK(hipGraphLaunch(t->exec, t->stream));
    done += 2;
    sinceCheck += 2;
    if (epsilon >= 0.0 && sinceCheck >= 32 && done < iters) {
      TutteScalars s;
      if (read_sc(t, &s))
        return -1;
      sinceCheck = 0;
      if (s.stopped)
        break;
    }
  }
  if (kr.kind == 3) {
    hipLaunchKernelGGL(k_finish_pending, dim3(1), dim3(256), 0, t->stream, t->partial.p,
                       t->blocks, t->sc.p, t->records.p, t->recordCap);
    GVH_CHECK_LAUNCH();
  }
  if (done < iters) {
    TutteScalars s;
    if (read_sc(t, &s))
      return -1;
    if (!s.stopped && enqueue_step(t, (int)(s.sweeps & 1), k))
      return -1;
  }
  TutteScalars after;
  if (read_sc(t, &after))
    return -1;
  const int64_t ran = after.sweeps - before.sweeps;
  if (maxDeltas && ran > 0)
    GVH_CHECK(hipMemcpy(maxDeltas, t->records.p, sizeof(double) * ran,
                        hipMemcpyDeviceToHost));
  return ran;
}

} // namespace

extern "C" {

int gvizhip_tutte_create(gvizhip_tutte **out, int64_t n, const int64_t *offsets,
                         const int64_t *adj, int32_t precision) {
  if (!out || n < 0 || (n > 0 && !offsets)) {
    set_error("gvizhip_tutte_create: invalid arguments");
    return -1;
  }
  *out = nullptr;
  int dc = 0;
  if (hipGetDeviceCount(&dc) != hipSuccess || dc <= 0) {
    set_error("gvizhip_tutte_create: no GPU is visible to HIP");
    return -1;
  }
  const int64_t m2 = n > 0 ? offsets[n] : 0;
  if (n >= INT32_MAX || m2 >= INT32_MAX) {
    set_error("gvizhip_tutte_create: graph exceeds 2^31 vertices or entries");
    return -1;
  }
  for (int64_t e = 0; e < m2; e++) {
    if (adj[e] < 0 || adj[e] >= n) {
      set_error("gvizhip_tutte_create: adjacency entry %lld out of range",
                (long long)e);
      return -1;
    }
  }
  for (int64_t v = 0; v < n; v++) {
    if (offsets[v + 1] < offsets[v]) {
      set_error("gvizhip_tutte_create: offsets are not monotone");
      return -1;
    }
  }
  // breadth-first order over the components in ascending id of their first
  // vertex; neighbours are enqueued in row order
  std::vector<int32_t> perm(n), inv(n, -1);
  {
    int64_t head = 0, tail = 0;
    for (int64_t r = 0; r < n; r++) {
      if (inv[r] >= 0)
        continue;
      inv[r] = (int32_t)tail;
      perm[tail++] = (int32_t)r;
      while (head < tail) {
        const int32_t u = perm[head++];
        for (int64_t e = offsets[u]; e < offsets[u + 1]; e++) {
          const int64_t w = adj[e];
          if (inv[w] < 0) {
            inv[w] = (int32_t)tail;
            perm[tail++] = (int32_t)w;
          }
        }
      }
    }
  }
  // relabelled CSR: stored vertex i is the caller's perm[i], its row in the
  // original entry order with every neighbour renamed
  std::vector<int32_t> o32(n + 1), a32(m2 > 0 ? m2 : 1);
  o32[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    const int64_t u = perm[i];
    int64_t at = o32[i];
    for (int64_t e = offsets[u]; e < offsets[u + 1]; e++)
      a32[at++] = inv[adj[e]];
    o32[i + 1] = (int32_t)at;
  }
  gvizhip_tutte *t = new gvizhip_tutte();
  t->perm = std::move(perm);
  t->hostOff = o32;
  t->hostAdj = a32;
  t->hostBoundary.assign(n > 0 ? n : 1, 0);
  t->n = n;
  t->m2 = m2;
  t->precision = precision == GVIZHIP_PRECISION_FP64 ? GVIZHIP_PRECISION_FP64
                                                     : GVIZHIP_PRECISION_FP32;
  const int64_t n1 = n > 0 ? n : 1;
  t->blocks = (int)std::min<int64_t>(grid_for(n1, kB), 2048);
  if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&t->pinned, sizeof(TutteScalars), hipHostMallocDefault) !=
          hipSuccess ||
      t->off.alloc(n1 + 1) || t->adj.alloc(m2 > 0 ? m2 : 1) ||
      t->boundary.alloc(n1) || t->pos[0].alloc(n1 * sizeof(double2)) ||
      t->pos[1].alloc(n1 * sizeof(double2)) || t->vel.alloc(n1 * sizeof(double2)) ||
      t->partial.alloc(2 * (size_t)t->blocks) ||
      t->records.alloc(4096) || t->sc.alloc(1)) {
    if (!gvizhip_last_error()[0])
      set_error("gvizhip_tutte_create: allocation failed");
    gvizhip_tutte_destroy(t);
    return -1;
  }
  t->recordCap = 4096;
  TutteScalars s{};
  s.epsilon = -1.0;
  if (hipMemcpy(t->off.p, o32.data(), sizeof(int32_t) * (n + 1),
                hipMemcpyHostToDevice) != hipSuccess ||
      (m2 > 0 && hipMemcpy(t->adj.p, a32.data(), sizeof(int32_t) * m2,
                           hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(t->sc.p, &s, sizeof(s), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(t->boundary.p, 0, n1) != hipSuccess ||
      hipMemset(t->vel.p, 0, n1 * sizeof(double2)) != hipSuccess) {
    set_error("gvizhip_tutte_create: upload failed");
    gvizhip_tutte_destroy(t);
    return -1;
  }
  *out = t;
  return 0;
}

int gvizhip_tutte_upload(gvizhip_tutte *t, const double *positions,
                         const uint8_t *isBoundary) {
  if (!t) {
    set_error("gvizhip_tutte_upload: null context");
    return -1;
  }
  const int64_t n = t->n;
  if (n == 0)
    return 0;
  const std::vector<int32_t> &perm = t->perm;
  if (positions) {
    if (t->precision == GVIZHIP_PRECISION_FP64) {
      std::vector<double> d(2 * n);
      for (int64_t i = 0; i < n; i++) {
        d[2 * i] = positions[2 * (int64_t)perm[i]];
        d[2 * i + 1] = positions[2 * (int64_t)perm[i] + 1];
      }
      GVH_CHECK(hipMemcpy(t->pos[0].p, d.data(), sizeof(double) * 2 * n,
                          hipMemcpyHostToDevice));
    } else {
      std::vector<float> f(2 * n);
      for (int64_t i = 0; i < n; i++) {
        f[2 * i] = (float)positions[2 * (int64_t)perm[i]];
        f[2 * i + 1] = (float)positions[2 * (int64_t)perm[i] + 1];
      }
      GVH_CHECK(hipMemcpy(t->pos[0].p, f.data(), sizeof(float) * 2 * n,
                          hipMemcpyHostToDevice));
    }
  }
  if (isBoundary) {
    std::vector<uint8_t> b(n);
    for (int64_t i = 0; i < n; i++)
      b[i] = isBoundary[perm[i]];
    GVH_CHECK(hipMemcpy(t->boundary.p, b.data(), n, hipMemcpyHostToDevice));
    t->hostBoundary.assign(b.begin(), b.end());
    t->gsDirty = true;
  }
  TutteScalars s{};
  s.epsilon = -1.0;
  GVH_CHECK(hipMemcpy(t->sc.p, &s, sizeof(s), hipMemcpyHostToDevice));
  return 0;
}

int gvizhip_tutte_download(gvizhip_tutte *t, double *positions) {
  if (!t || (!positions && t->n > 0)) {
    set_error("gvizhip_tutte_download: invalid arguments");
    return -1;
  }
  const int64_t n = t->n;
  if (n == 0)
    return 0;
  TutteScalars s;
  if (read_sc(t, &s))
    return -1;
  const int cur = (int)(s.sweeps & 1);
  const std::vector<int32_t> &perm = t->perm;
  if (t->precision == GVIZHIP_PRECISION_FP64) {
    std::vector<double> d(2 * n);
    GVH_CHECK(hipMemcpy(d.data(), t->pos[cur].p, sizeof(double) * 2 * n,
                        hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
      positions[2 * (int64_t)perm[i]] = d[2 * i];
      positions[2 * (int64_t)perm[i] + 1] = d[2 * i + 1];
    }
  } else {
    std::vector<float> f(2 * n);
    GVH_CHECK(hipMemcpy(f.data(), t->pos[cur].p, sizeof(float) * 2 * n,
                        hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
      positions[2 * (int64_t)perm[i]] = (double)f[2 * i];
      positions[2 * (int64_t)perm[i] + 1] = (double)f[2 * i + 1];
    }
  }
  return 0;
}

int64_t gvizhip_tutte_sweeps(gvizhip_tutte *t, int64_t iters, double alpha,
                             double epsilon, double *maxDeltas) {
  if (!t || iters < 0 || !(alpha > 0.0) || alpha > 1.0) {
    set_error("gvizhip_tutte_sweeps: invalid arguments");
    return -1;
  }
  return run_steps(t, iters, epsilon, maxDeltas, StepKind{0, alpha, 0.0, 0.0});
}

int64_t gvizhip_tutte_gs_sweeps(gvizhip_tutte *t, int64_t iters, double alpha,
                                double epsilon, double *maxDeltas) {
  if (!t || iters < 0 || !(alpha > 0.0) || alpha > 1.0) {
    set_error("gvizhip_tutte_gs_sweeps: invalid arguments");
    return -1;
  }
  if (t->n == 0 || iters == 0)
    return 0;
  if (build_gs_schedule(t))
    return -1;
  return run_steps(t, iters, epsilon, maxDeltas, StepKind{2, alpha, 0.0, 0.0});
}

int64_t gvizhip_tutte_spring_steps(gvizhip_tutte *t, int64_t iters, double dt,
                                   double stiffness, double damping, double epsilon,
                                   double *maxDeltas) {
  if (!t || iters < 0 || !(dt > 0.0)) {
    set_error("gvizhip_tutte_spring_steps: invalid arguments");
    return -1;
  }
  return run_steps(t, iters, epsilon, maxDeltas, StepKind{1, dt, stiffness, damping});
}

int gvizhip_tutte_set_velocity(gvizhip_tutte *t, const double *velocity) {
  if (!t) {
    set_error("gvizhip_tutte_set_velocity: null context");
    return -1;
  }
  const int64_t n = t->n;
  if (n == 0)
    return 0;
  const size_t es = elem(t);
  if (!velocity) {
    GVH_CHECK(hipMemset(t->vel.p, 0, es * n));
    return 0;
  }
  const std::vector<int32_t> &perm = t->perm;
  if (t->precision == GVIZHIP_PRECISION_FP64) {
    std::vector<double> d(2 * n);
    for (int64_t i = 0; i < n; i++) {
      d[2 * i] = velocity[2 * (int64_t)perm[i]];
      d[2 * i + 1] = velocity[2 * (int64_t)perm[i] + 1];
    }
    GVH_CHECK(hipMemcpy(t->vel.p, d.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(2 * n);
    for (int64_t i = 0; i < n; i++) {
      f[2 * i] = (float)velocity[2 * (int64_t)perm[i]];
      f[2 * i + 1] = (float)velocity[2 * (int64_t)perm[i] + 1];
    }
    GVH_CHECK(hipMemcpy(t->vel.p, f.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice));
  }
  return 0;
}

int gvizhip_tutte_get_velocity(gvizhip_tutte *t, double *velocity) {
  if (!t || (!velocity && t->n > 0)) {
    set_error("gvizhip_tutte_get_velocity: invalid arguments");
    return -1;
  }
  const int64_t n = t->n;
  if (n == 0)
    return 0;
  GVH_CHECK(hipStreamSynchronize(t->stream));
  const std::vector<int32_t> &perm = t->perm;
  if (t->precision == GVIZHIP_PRECISION_FP64) {
    std::vector<double> d(2 * n);
    GVH_CHECK(hipMemcpy(d.data(), t->vel.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
      velocity[2 * (int64_t)perm[i]] = d[2 * i];
      velocity[2 * (int64_t)perm[i] + 1] = d[2 * i + 1];
    }
  } else {
    std::vector<float> f(2 * n);
    GVH_CHECK(hipMemcpy(f.data(), t->vel.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
      velocity[2 * (int64_t)perm[i]] = (double)f[2 * i];
      velocity[2 * (int64_t)perm[i] + 1] = (double)f[2 * i + 1];
    }
  }
  return 0;
}

int gvizhip_tutte_profile(gvizhip_tutte *t, int64_t sweeps, double alpha,
                          double *sweepMs, double *totalMs) {
  if (!t || sweeps < 0 || !(alpha > 0.0) || alpha > 1.0) {
    set_error("gvizhip_tutte_profile: invalid arguments");
    return -1;
  }
  if (t->n == 0 || sweeps == 0) {
    if (sweepMs)
      *sweepMs = 0.0;
    if (totalMs)
      *totalMs = 0.0;
    return 0;
  }
  TutteScalars s;
  if (read_sc(t, &s))
    return -1;
  hipLaunchKernelGGL(k_begin, dim3(1), dim3(1), 0, t->stream, t->sc.p, -1.0);
  GVH_CHECK_LAUNCH();
  std::vector<hipEvent_t> ev(2 * sweeps + 2, nullptr);
  for (auto &e : ev)
    GVH_CHECK(hipEventCreate(&e));
  int cur = (int)(s.sweeps & 1);
  const unsigned g = (unsigned)t->blocks;
  GVH_CHECK(hipEventRecord(ev[2 * sweeps], t->stream));
  // fused sweeps, as a run applies them: each finishes the one before
  for (int64_t k = 0; k < sweeps; k++) {
    GVH_CHECK(hipEventRecord(ev[2 * k], t->stream));
    if (t->precision == GVIZHIP_PRECISION_FP64)
      hipLaunchKernelGGL((k_sweep<double, true>), dim3(g), dim3(kB), 0, t->stream, t->off.p,
                         t->adj.p, t->boundary.p, (const double2 *)t->pos[cur].p,
                         (double2 *)t->pos[1 - cur].p, t->n, alpha, t->partial.p,
                         t->sc.p, cur, t->records.p, t->recordCap);
    else
      hipLaunchKernelGGL((k_sweep<float, true>), dim3(g), dim3(kB), 0, t->stream, t->off.p,
                         t->adj.p, t->boundary.p, (const float2 *)t->pos[cur].p,
                         (float2 *)t->pos[1 - cur].p, t->n, (float)alpha,
                         t->partial.p, t->sc.p, cur, t->records.p, t->recordCap);
    GVH_CHECK(hipEventRecord(ev[2 * k + 1], t->stream));
    GVH_CHECK_LAUNCH();
    cur = 1 - cur;
  }
  hipLaunchKernelGGL(k_finish_pending, dim3(1), dim3(256), 0, t->stream, t->partial.p,
                     (int)g, t->sc.p, t->records.p, t->recordCap);
  GVH_CHECK_LAUNCH();
  GVH_CHECK(hipEventRecord(ev[2 * sweeps + 1], t->stream));
  GVH_CHECK(hipStreamSynchronize(t->stream));
  double sum = 0.0;
  for (int64_t k = 0; k < sweeps; k++) {
    float ms = 0.0f;
    GVH_CHECK(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
    sum += ms;
  }
  float total = 0.0f;
  GVH_CHECK(hipEventElapsedTime(&total, ev[2 * sweeps], ev[2 * sweeps + 1]));
  for (auto &e : ev)
    (void)hipEventDestroy(e);
  if (sweepMs)
    *sweepMs = sum;
  if (totalMs)
    *totalMs = total;
  return 0;
}

// Device address of this file's scratch array on the current device, looked
// up once per device (the array's name is a host-side stand-in, not an
// address a kernel may use).
static int noalloc_scratch(double **out) {
  static std::mutex mu;
  static double *addr[64] = {};
  int dev = -1;
  GVH_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("gvizhip_tutte_sweeps_noalloc: device %d out of range", dev);
    return -1;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    GVH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    addr[dev] = (double *)p;
  }
  *out = addr[dev];
  return 0;
}

int gvizhip_tutte_sweeps_noalloc(int64_t n, const int32_t *offsets, const int32_t *adj,
                                 const uint8_t *isBoundary, double *posA, double *posB,
                                 int64_t iters, double alpha, double epsilon,
                                 double *maxDeltas, int64_t *sweepsRun, void *stream) {
  if (n < 0 || iters < 0 || !(alpha > 0.0) || alpha > 1.0 ||
      (n > 0 && (!offsets || !adj || !isBoundary || !posA || !posB)) ||
      (iters > 0 && !maxDeltas)) {
    set_error("gvizhip_tutte_sweeps_noalloc: invalid arguments");
    return -1;
  }
  if (n > (int64_t)INT32_MAX - 1) {
    set_error("gvizhip_tutte_sweeps_noalloc: %lld vertices exceed the 32-bit CSR "
              "indices of the sweep (at most %d)", (long long)n, INT32_MAX - 1);
    return -1;
  }
  double *base = nullptr;
  if (noalloc_scratch(&base))
    return -1;
  TutteScalars *sc = (TutteScalars *)base;
  double *partial = base + kNoAllocHeader;
  hipStream_t st = (hipStream_t)stream;
  const unsigned g =
      (unsigned)std::min<int64_t>(grid_for(n > 0 ? n : 1, kB), kNoAllocBlocks);
  hipLaunchKernelGGL(k_noalloc_begin, dim3(1), dim3(1), 0, st, sc, epsilon);
  GVH_CHECK_LAUNCH();
  for (int64_t k = 0; k < iters; k++) {
    const double2 *in = (const double2 *)((k & 1) ? posB : posA);
    double2 *out = (double2 *)((k & 1) ? posA : posB);
    hipLaunchKernelGGL((k_sweep<double, false>), dim3(g), dim3(kB), 0, st, offsets, adj,
                       isBoundary, in, out, n, alpha, partial, sc, 0, nullptr,
                       (int64_t)0);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, st, partial, (int)g, sc,
                       maxDeltas, iters);
    GVH_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_noalloc_end, dim3(g), dim3(kB), 0, st, sc, (double2 *)posA,
                     (const double2 *)posB, n, sweepsRun);
  GVH_CHECK_LAUNCH();
  return 0;
}

void gvizhip_tutte_destroy(gvizhip_tutte *t) {
  if (!t)
    return;
  drop_graph(t);
  if (t->stream)
    (void)hipStreamSynchronize(t->stream);
  if (t->pinned)
    (void)hipHostFree(t->pinned);
  if (t->stream)
    (void)hipStreamDestroy(t->stream);
  delete t;
}

} // extern "C"
