// force_common.hip -- precision-independent stages of the force step:
// bounding box and root cell, quadrant-path keys, key sort, binary radix
// tree, double-double mass prefix sums, the FA2 speed update and the step
// statistics. Compiled with -ffp-contract=off so that the FP64 key
// recurrence and the FA2 scalar update round exactly like the reference.
#include "dd.hpp"
#include "force_engine.hpp"
#include "step_tail.hpp"

#include <cstdarg>

namespace gvh {

static thread_local char g_err[512];

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char *last_error() { return g_err; }

namespace {

constexpr int kBlock = 256;

// ------------------------------------------------------------- root cell
// Min/max over all positions: exact in any order, so a tree reduction gives
// the reference's serial loop result (src/ds/gvizQuadtree.c:165-173).
__global__ __launch_bounds__(kBlock) void k_bbox(const double2 *__restrict__ pos,
                                                 int64_t n,
                                                 double *__restrict__ partial) {
  __shared__ double sm[4 * 16];
  double v[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 p = pos[i];
    v[0] = fmin(v[0], p.x);
    v[1] = fmax(v[1], p.x);
    v[2] = fmin(v[2], p.y);
    v[3] = fmax(v[3], p.y);
  }
  // max over (-min) keeps one reduction kind
  v[0] = -v[0];
  v[2] = -v[2];
  const bool isMax[4] = {true, true, true, true};
  block_reduce<4>(v, isMax, sm);
  if (threadIdx.x == 0) {
    partial[blockIdx.x * 4 + 0] = -v[0];
    partial[blockIdx.x * 4 + 1] = v[1];
    partial[blockIdx.x * 4 + 2] = -v[2];
    partial[blockIdx.x * 4 + 3] = v[3];
  }
}

// Root centre, half size and depth limit: src/ds/gvizQuadtree.c:175-178 and
// the split guard halfSize <= 1e-9 (:6, :134).
__global__ void k_root(const double *__restrict__ partial, int blocks,
                       int64_t n, Scalars *__restrict__ sc) {
  __shared__ double sm[4 * 16];
  double v[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    v[0] = fmax(v[0], -partial[b * 4 + 0]);
    v[1] = fmax(v[1], partial[b * 4 + 1]);
    v[2] = fmax(v[2], -partial[b * 4 + 2]);
    v[3] = fmax(v[3], partial[b * 4 + 3]);
  }
  const bool isMax[4] = {true, true, true, true};
  block_reduce<4>(v, isMax, sm);
  if (threadIdx.x == 0) {
    double minX = -v[0], maxX = v[1], minY = -v[2], maxY = v[3];
    double sx = maxX - minX, sy = maxY - minY;
    double span = sx > sy ? sx : sy;
    double half = span > 0.0 ? span / 2.0 : 1.0;
    sc->rootCx = (minX + maxX) / 2.0;
    sc->rootCy = (minY + maxY) / 2.0;
    sc->rootHalf = half;
    int d = 0;
    double h = half;
    while (!(h <= 1e-9) && d < 64) {
      h = h / 2.0;
      d++;
    }
    sc->maxDepth = d;
    sc->rootRef = n >= 2 ? 0 : -1;
    // every record lies at most d levels deep (a point one more)
    sc->treeDepth = d + 2;
  }
}

// Quadrant path of every point by replaying the reference's FP64 centre
// recurrence (src/ds/gvizQuadtree.c:65-92): digit = 2*(y < cy) + (x >= cx),
// child centre = centre +/- halfSize/2. Level 1 sits in bits 63..62 of keyHi,
// level 33 in bits 63..62 of keyLo; levels past the depth limit stay 0.
__global__ __launch_bounds__(kBlock) void k_keys(const double2 *__restrict__ pos,
                                                 int64_t n,
                                                 const Scalars *__restrict__ sc,
                                                 uint64_t *__restrict__ keyHi,
                                                 uint64_t *__restrict__ keyLo,
                                                 uint32_t *__restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || sc->stopped)
    return;
  double2 p = pos[i];
  double cx = sc->rootCx, cy = sc->rootCy, h = sc->rootHalf;
  const int levels = sc->maxDepth;
  uint64_t hi = 0, lo = 0;
  for (int level = 1; level <= levels; level++) {
    bool east = p.x >= cx;
    bool north = p.y >= cy;
    uint64_t digit = (north ? 0u : 2u) | (east ? 1u : 0u);
    if (level <= 32)
      hi |= digit << (2 * (32 - level));
    else
      lo |= digit << (2 * (64 - level));
    double quarter = h / 2.0;
    cx = cx + (east ? quarter : -quarter);
    cy = cy + (north ? quarter : -quarter);
    h = quarter;
  }
  keyHi[i] = hi;
  keyLo[i] = lo;
  vals[i] = (uint32_t)i;
}

#ifndef GVH_SORT_FROM_BIT
#define GVH_SORT_FROM_BIT 32
#endif
constexpr int kSortFrom = GVH_SORT_FROM_BIT; // keyHi bits below this are not radix-sorted

// The radix sort orders by the top 32 bits of keyHi only (one graph: the
// first 16 levels; a batch: the graph id and 10 levels), stable, so equal
// top halves stay in id order. Runs of equal top halves are then put in full
// (keyHi, keyLo, id) order in place; every key is distinct once the id is
// included. They are common (in a layout whose root box is set by a few far
// vertices, most points share their first 16 levels with a few others), so a
// run of up to kRankRun entries is ordered by rank: a thread per entry counts
// the run's entries below it and writes itself to start + rank. A workgroup
// stages its kBlock positions with kRunHalo entries either side in LDS and
// orders the short runs whose first entry is among its positions, including
// the tail of one that runs into the next workgroup's (threads t < kRankRun
// take position end + t as a second entry). A longer run -- a tight cluster
// -- is heap-sorted in place by the thread at its head (O(L log L)).
//
// Every run has one owner, and only the owner writes its positions. Another
// workgroup that stages them reads only their top halves, which are equal
// within a run whatever order its owner has written so far, so no result
// depends on the order in which workgroups run.
constexpr int kRankRun = 64;
constexpr int kRunHalo = kRankRun + 1;
constexpr int kRunWindow = kBlock + 2 * kRunHalo;

__global__ __launch_bounds__(kBlock) void k_order_runs32(
    uint64_t *__restrict__ keyHi, const uint64_t *__restrict__ keyLo,
    uint32_t *__restrict__ vals, int64_t n, const Scalars *__restrict__ sc) {
  __shared__ uint64_t wk[kRunWindow];
  __shared__ uint32_t wv[kRunWindow];
  if (sc->stopped)
    return;
  // window entry i is sorted position base + i
  const int64_t start = (int64_t)blockIdx.x * kBlock;
  const int64_t end = start + kBlock < n ? start + kBlock : n;
  const int64_t base = start - kRunHalo;
  for (int i = threadIdx.x; i < kRunWindow; i += kBlock) {
    const int64_t g = base + i;
    if (g >= 0 && g < n) {
      wk[i] = keyHi[g];
      wv[i] = vals[g];
    }
  }
  __syncthreads();
  // Orders the entry at position s (start <= s < end + kRankRun). Reads stay
  // in the window: b - 1 >= s - kRunHalo, and e <= b + kRankRun with
  // b < end.
  auto order = [&](int64_t s, bool second) {
    const uint64_t h = wk[s - base];
    const uint32_t v = wv[s - base];
    const uint64_t top = h >> kSortFrom;
    int64_t b = s;
    while (b > 0 && s - b <= kRankRun && (wk[b - 1 - base] >> kSortFrom) == top)
      b--;
    if (second && b >= end)
      return; // the next workgroup's run
    int64_t e = s + 1;
    while (e < n && e - b <= kRankRun && (wk[e - base] >> kSortFrom) == top)
      e++;
    if (e - b <= kRankRun) {
      if (b < start)
        return; // the previous workgroup's run
      int64_t rank = 0;
      bool haveLo = false;
      uint64_t lo = 0;
      for (int64_t j = b; j < e; j++) {
        if (j == s)
          continue;
        const uint64_t hj = wk[j - base];
        const uint32_t vj = wv[j - base];
        bool below;
        if (hj != h) {
          below = hj < h;
        } else {
          if (!haveLo) {
            lo = keyLo[v];
            haveLo = true;
          }
          const uint64_t lj = keyLo[vj];
          below = lj != lo ? lj < lo : vj < v;
        }
        rank += below;
      }
      keyHi[b + rank] = h;
      vals[b + rank] = v;
      return;
    }
    if (second || (s > 0 && (wk[s - 1 - base] >> kSortFrom) == top))
      return; // a long run is ordered by the thread at its head
    e = s + 1;
    while (e < n && (keyHi[e] >> kSortFrom) == top)
      e++;
    const int64_t L = e - s;
    uint64_t *kh = keyHi + s;
    uint32_t *vv = vals + s;
    auto less = [&](uint64_t ha, uint32_t va, uint64_t hb, uint32_t vb) {
      if (ha != hb)
        return ha < hb;
      const uint64_t la = keyLo[va], lb = keyLo[vb];
      if (la != lb)
        return la < lb;
      return va < vb;
    };
    // heap sort: max-heap on the same order, then repeated extraction
    auto sift = [&](int64_t root, int64_t stop) {
      while (2 * root + 1 < stop) {
        int64_t c = 2 * root + 1;
        if (c + 1 < stop && less(kh[c], vv[c], kh[c + 1], vv[c + 1]))
          c++;
        if (!less(kh[root], vv[root], kh[c], vv[c]))
          return;
        const uint64_t th = kh[root];
        kh[root] = kh[c];
        kh[c] = th;
        const uint32_t tv = vv[root];
        vv[root] = vv[c];
        vv[c] = tv;
        root = c;
      }
    };
    for (int64_t r = L / 2 - 1; r >= 0; r--)
      sift(r, L);
    for (int64_t last = L - 1; last > 0; last--) {
      const uint64_t th = kh[0];
      kh[0] = kh[last];
      kh[last] = th;
      const uint32_t tv = vv[0];
      vv[0] = vv[last];
      vv[last] = tv;
      sift(0, last);
    }
  };
  const int64_t s = start + threadIdx.x;
  if (s < end)
    order(s, false);
  const int64_t s2 = end + threadIdx.x;
  if (threadIdx.x < kRankRun && end == start + kBlock && s2 < n)
    order(s2, true);
}

// Longest common prefix in bits of the 128-bit keys at sorted positions i, j;
// identical keys compare their positions (Karras 2012), so delta is unique.
__device__ __forceinline__ int lcp(const uint64_t *__restrict__ hi,
                                   const uint64_t *__restrict__ lo, int64_t n,
                                   int64_t i, int64_t j) {
  if (j < 0 || j >= n)
    return -1;
  uint64_t a = hi[i] ^ hi[j];
  if (a)
    return __clzll(a);
  uint64_t b = lo[i] ^ lo[j];
  if (b)
    return 64 + __clzll(b);
  return 128 + __clz((uint32_t)(i ^ j));
}

// Binary radix tree over sorted keys (Karras, "Maximizing parallelism in the
// construction of BVHs, octrees and k-d trees", HPG 2012). Internal node i
// covers [first, last]; children < 0 encode leaves: -(s + 1).
__global__ __launch_bounds__(kBlock) void k_karras(
    const uint64_t *__restrict__ hi, const uint64_t *__restrict__ lo, int64_t n,
    int32_t *__restrict__ left, int32_t *__restrict__ right,
    int32_t *__restrict__ parent, int32_t *__restrict__ first,
    int32_t *__restrict__ last, int32_t *__restrict__ delta,
    const Scalars *__restrict__ sc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1 || sc->stopped)
    return;
  int d = lcp(hi, lo, n, i, i + 1) - lcp(hi, lo, n, i, i - 1) > 0 ? 1 : -1;
  int dmin = lcp(hi, lo, n, i, i - d);
  int64_t lmax = 2;
  while (lcp(hi, lo, n, i, i + lmax * d) > dmin)
    lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (lcp(hi, lo, n, i, i + (l + t) * d) > dmin)
      l += t;
  int64_t j = i + l * d;
  int dnode = lcp(hi, lo, n, i, j);
  int64_t s = 0;
  int64_t t = l;
  do {
    t = (t + 1) / 2;
    if (lcp(hi, lo, n, i, i + (s + t) * d) > dnode)
      s += t;
  } while (t > 1);
  int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
  int64_t lo_ = i < j ? i : j;
  int64_t hi_ = i < j ? j : i;
  int32_t lc = (lo_ == gamma) ? -(int32_t)(gamma + 1) : (int32_t)gamma;
  int32_t rc = (hi_ == gamma + 1) ? -(int32_t)(gamma + 2) : (int32_t)(gamma + 1);
  left[i] = lc;
  right[i] = rc;
  first[i] = (int32_t)lo_;
  last[i] = (int32_t)hi_;
  delta[i] = dnode;
  if (lc >= 0)
    parent[lc] = (int32_t)i;
  if (rc >= 0)
    parent[rc] = (int32_t)i;
  if (i == 0)
    parent[0] = -1;
}

// ----------------------------------------------- mass prefix sums (dd)
// One wavefront per chunk of 512 points in key order, taken as 8 rows of 64:
// every load and store is a coalesced row, and the chunk's offset comes from
// the scanned per-chunk totals, so no block-wide synchronisation is needed.
constexpr int kScanRows = 8;
constexpr int kScanTile = kScanRows * kWave; // points per chunk (one wave)
constexpr int kScanWaves = kBlock / kWave;   // chunks per block

struct Acc {
  double m;
  dd x, y;
};

__device__ __forceinline__ Acc acc_add(const Acc &a, const Acc &b) {
  return {a.m + b.m, dd_add(a.x, b.x), dd_add(a.y, b.y)};
}

__device__ __forceinline__ Acc acc_zero() { return Acc{0.0, {0.0, 0.0}, {0.0, 0.0}}; }

__device__ __forceinline__ Acc load_elem(const double *__restrict__ spos64,
                                         const double *__restrict__ smass64,
                                         int64_t s) {
  double m = smass64[s];
  return {m, two_prod(m, spos64[2 * s]), two_prod(m, spos64[2 * s + 1])};
}

__device__ __forceinline__ Acc acc_shfl_up(const Acc &a, int off) {
  return {__shfl_up(a.m, off, kWave),
          {__shfl_up(a.x.hi, off, kWave), __shfl_up(a.x.lo, off, kWave)},
          {__shfl_up(a.y.hi, off, kWave), __shfl_up(a.y.lo, off, kWave)}};
}

__device__ __forceinline__ Acc acc_shfl_xor(const Acc &a, int off) {
  return {__shfl_xor(a.m, off, kWave),
          {__shfl_xor(a.x.hi, off, kWave), __shfl_xor(a.x.lo, off, kWave)},
          {__shfl_xor(a.y.hi, off, kWave), __shfl_xor(a.y.lo, off, kWave)}};
}

__device__ __forceinline__ Acc acc_lane(const Acc &a, int l) {
  return {__shfl(a.m, l, kWave),
          {__shfl(a.x.hi, l, kWave), __shfl(a.x.lo, l, kWave)},
          {__shfl(a.y.hi, l, kWave), __shfl(a.y.lo, l, kWave)}};
}

// Per-chunk totals: each lane sums its column of the chunk's 8 rows, then
// the wave adds the 64 column sums in a fixed butterfly order.
__global__ __launch_bounds__(kBlock) void k_scan_reduce(
    const double *__restrict__ spos64, const double *__restrict__ smass64,
    int64_t n, double *__restrict__ part, const Scalars *__restrict__ sc) {
  if (sc->stopped)
    return;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t chunk = (int64_t)blockIdx.x * kScanWaves + threadIdx.x / kWave;
  const int64_t base = chunk * kScanTile;
  if (base >= n)
    return;
  Acc a = acc_zero();
#pragma unroll
  for (int r = 0; r < kScanRows; r++) {
    const int64_t s = base + r * kWave + lane;
    if (s < n)
      a = acc_add(a, load_elem(spos64, smass64, s));
  }
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1)
    a = acc_add(a, acc_shfl_xor(a, off));
  if (lane == 0) {
    double *p = part + 5 * chunk;
    p[0] = a.m;
    p[1] = a.x.hi;
    p[2] = a.x.lo;
    p[3] = a.y.hi;
    p[4] = a.y.lo;
  }
}

// Exclusive scan of the per-chunk totals, in place (one block of 1024).
__global__ __launch_bounds__(1024) void k_scan_partials(double *__restrict__ part,
                                                        int tiles,
                                                        const Scalars *__restrict__ sc) {
  __shared__ Acc sm[1024];
  if (sc->stopped)
    return;
  const int chunk = (tiles + 1023) / 1024;
  const int b = threadIdx.x * chunk;
  const int e = b + chunk < tiles ? b + chunk : tiles;
  Acc a = acc_zero();
  for (int t = b; t < e; t++) {
    const double *p = part + 5 * t;
    a = acc_add(a, Acc{p[0], {p[1], p[2]}, {p[3], p[4]}});
  }
  sm[threadIdx.x] = a;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    Acc o = acc_zero();
    if (threadIdx.x >= (unsigned)off)
      o = sm[threadIdx.x - off];
    __syncthreads();
    if (threadIdx.x >= (unsigned)off)
      sm[threadIdx.x] = acc_add(o, sm[threadIdx.x]);
    __syncthreads();
  }
  Acc run = threadIdx.x ? sm[threadIdx.x - 1] : acc_zero();
  for (int t = b; t < e; t++) {
    double *p = part + 5 * t;
    Acc v = {p[0], {p[1], p[2]}, {p[3], p[4]}};
    p[0] = run.m;
    p[1] = run.x.hi;
    p[2] = run.x.lo;
    p[3] = run.y.hi;
    p[4] = run.y.lo;
    run = acc_add(run, v);
  }
}

// prefix[s + 1] = sum over sorted points 0..s; prefix[0] = 0. Row by row:
// a wave-wide inclusive scan of the row, plus the running carry.
__global__ __launch_bounds__(kBlock) void k_scan_final(
    const double *__restrict__ spos64, const double *__restrict__ smass64,
    int64_t n, const double *__restrict__ part,
    double *__restrict__ pm, double *__restrict__ pxh, double *__restrict__ pxl,
    double *__restrict__ pyh, double *__restrict__ pyl,
    const Scalars *__restrict__ sc) {
  if (sc->stopped)
    return;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t chunk = (int64_t)blockIdx.x * kScanWaves + threadIdx.x / kWave;
  const int64_t base = chunk * kScanTile;
  if (base >= n)
    return;
  const double *p = part + 5 * chunk;
  Acc carry = {p[0], {p[1], p[2]}, {p[3], p[4]}};
  for (int r = 0; r < kScanRows; r++) {
    const int64_t row = base + r * kWave;
    if (row >= n)
      break;
    const int64_t s = row + lane;
    Acc v = s < n ? load_elem(spos64, smass64, s) : acc_zero();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const Acc o = acc_shfl_up(v, off);
      if (lane >= off)
        v = acc_add(o, v);
    }
    const Acc out = acc_add(carry, v);
    if (s < n) {
      pm[s + 1] = out.m;
      pxh[s + 1] = out.x.hi;
      pxl[s + 1] = out.x.lo;
      pyh[s + 1] = out.y.hi;
      pyl[s + 1] = out.y.lo;
    }
    carry = acc_add(carry, acc_lane(v, kWave - 1));
  }
  if (chunk == 0 && lane == 0) {
    pm[0] = 0.0;
    pxh[0] = pxl[0] = pyh[0] = pyl[0] = 0.0;
  }
}

// -------------------------------------------------------- FA2 speed update
__global__ void k_fa2(const double *__restrict__ partial, int blocks, int64_t n,
                      double jitterTolerance, Scalars *__restrict__ sc,
                      StepRecord *__restrict__ rec, int64_t recCap) {
  __shared__ double smem[2 * 16];
  fa2_tail(partial, blocks, n, jitterTolerance, sc, rec, recCap, smem);
}

__global__ void k_stats(const double *__restrict__ partial, int blocks,
                        int64_t n, double gravityK, Scalars *__restrict__ sc,
                        StepRecord *__restrict__ rec, int64_t recCap) {
  __shared__ double smem[3 * 16];
  stats_tail(partial, blocks, n, gravityK, sc, rec, recCap, nullptr, smem);
}


// ------------------------------------------------------------ batch stages
// Bounding box and root cell of every graph of a batch: one block per graph,
// the same root formulas as k_root.
__global__ __launch_bounds__(kBlock) void k_bbox_root_g(const double2 *__restrict__ pos,
                                                        const int64_t *__restrict__ gStart,
                                                        Scalars *__restrict__ scG) {
  __shared__ double sm[4 * 16];
  const int g = blockIdx.x;
  Scalars *sc = scG + g;
  if (sc->stopped)
    return;
  const int64_t b = gStart[g], e = gStart[g + 1];
  double v[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    double2 p = pos[i];
    v[0] = fmax(v[0], -p.x);
    v[1] = fmax(v[1], p.x);
    v[2] = fmax(v[2], -p.y);
    v[3] = fmax(v[3], p.y);
  }
  const bool isMax[4] = {true, true, true, true};
  block_reduce<4>(v, isMax, sm);
  if (threadIdx.x == 0) {
    double minX = -v[0], maxX = v[1], minY = -v[2], maxY = v[3];
    double sx = maxX - minX, sy = maxY - minY;
    double span = sx > sy ? sx : sy;
    double half = span > 0.0 ? span / 2.0 : 1.0;
    sc->rootCx = (minX + maxX) / 2.0;
    sc->rootCy = (minY + maxY) / 2.0;
    sc->rootHalf = half;
    int d = 0;
    double h = half;
    while (!(h <= 1e-9) && d < 64) {
      h = h / 2.0;
      d++;
    }
    sc->maxDepth = d;
    sc->rootRef = e - b >= 2 ? 0 : -1;
    // every record of the graph lies at most d levels deep (a point one more)
    sc->treeDepth = (d < kBatchLevels ? d : kBatchLevels) + 2;
    if (d > kBatchLevels)
      sc->traversalOverflow = 2;
  }
}

// Batch keys: graph id in bits 63..52 of keyHi, then levels 1..26 of the
// quadrant path in keyHi and levels 27..58 in keyLo (the single-graph layout
// shifted by 12 bits, so the longest common prefix of two keys of one graph
// is 12 + the single-graph one).
__global__ __launch_bounds__(kBlock) void k_keys_g(const double2 *__restrict__ pos,
                                                   int64_t n,
                                                   const int32_t *__restrict__ gid,
                                                   const Scalars *__restrict__ scG,
                                                   uint64_t *__restrict__ keyHi,
                                                   uint64_t *__restrict__ keyLo,
                                                   uint32_t *__restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int32_t g = gid[i];
  const Scalars &sc = scG[g];
  double2 p = pos[i];
  double cx = sc.rootCx, cy = sc.rootCy, h = sc.rootHalf;
  const int levels = sc.maxDepth < kBatchLevels ? sc.maxDepth : kBatchLevels;
  uint64_t hi = (uint64_t)g << 52, lo = 0;
  for (int level = 1; level <= levels; level++) {
    bool east = p.x >= cx;
    bool north = p.y >= cy;
    uint64_t digit = (north ? 0u : 2u) | (east ? 1u : 0u);
    if (level <= 26)
      hi |= digit << (2 * (26 - level));
    else
      lo |= digit << (2 * (58 - level));
    double quarter = h / 2.0;
    cx = cx + (east ? quarter : -quarter);
    cy = cy + (north ? quarter : -quarter);
    h = quarter;
  }
  keyHi[i] = hi;
  keyLo[i] = lo;
  vals[i] = (uint32_t)i;
}

__global__ void k_fa2_g(const double *__restrict__ partial,
                        const int64_t *__restrict__ gStart,
                        const int32_t *__restrict__ gBlocks,
                        const int64_t *__restrict__ gPart, double jitterTolerance,
                        Scalars *__restrict__ scG) {
  __shared__ double smem[2 * 16];
  const int g = blockIdx.x;
  fa2_tail(partial + 2 * gPart[g], gBlocks[g], gStart[g + 1] - gStart[g],
           jitterTolerance, scG + g, nullptr, 0, smem);
}

__global__ void k_stats_g(const double *__restrict__ partial,
                          const int64_t *__restrict__ gStart,
                          const int32_t *__restrict__ gBlocks,
                          const int64_t *__restrict__ gPart, double gravityK,
                          Scalars *__restrict__ scG, StepRecord *__restrict__ last) {
  __shared__ double smem[3 * 16];
  const int g = blockIdx.x;
  stats_tail(partial + 3 * gPart[g], gBlocks[g], gStart[g + 1] - gStart[g], gravityK,
             scG + g, nullptr, 0, last + g, smem);
}

} // namespace

int ForceCommon::alloc(int64_t nv, int64_t ne) {
  n = nv;
  m2 = ne;
  const int64_t n1 = nv > 0 ? nv : 1;
  if (pos64.alloc(2 * n1) || csrOff.alloc(n1 + 1) || csrAdj.alloc(ne > 0 ? ne : 1) ||
      degree.alloc(n1) || mass64.alloc(n1) || radius64.alloc(n1) || mr64.alloc(2 * n1) ||
      smass64.alloc(n1))
    return -1;
  if (keyHi.alloc(n1) || keyLo.alloc(n1) || keyLoS.alloc(n1) || vals.alloc(n1) ||
      sorter.reserve(n1))
    return -1;
  if (left.alloc(n1) || right.alloc(n1) || parent.alloc(n1) || first.alloc(n1) ||
      last.alloc(n1) || delta.alloc(n1))
    return -1;
  if (pm.alloc(n1 + 1) || pmxHi.alloc(n1 + 1) || pmxLo.alloc(n1 + 1) ||
      pmyHi.alloc(n1 + 1) || pmyLo.alloc(n1 + 1))
    return -1;
  const int64_t tiles = (n1 + kScanTile - 1) / kScanTile;
  if (scanPartial.alloc(5 * tiles))
    return -1;
  if (nodes.alloc((size_t)n1 * sizeof(NodeRec<double, double>)))
    return -1;
  const size_t w = sizeof(double);
  if (posT.alloc(2 * n1 * w) || spos.alloc(2 * n1 * w) || smr.alloc(2 * n1 * w) ||
      rep.alloc(2 * n1 * w) || disp.alloc(2 * n1 * w) || sf.alloc(2 * n1 * w) ||
      sw.alloc(n1 * w) || attMag.alloc(n1 * w) || repMag.alloc(n1 * w) ||
      massT.alloc(n1 * w) || radT.alloc(n1 * w) || spos64.alloc(2 * n1))
    return -1;
  reductionBlocks = 1024;
  if (partials.alloc(4 * (size_t)reductionBlocks) || sc.alloc(1) ||
      interactions.alloc(n1))
    return -1;
  return 0;
}

int ForceCommon::open_streams() {
  if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&evFork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&evJoin, hipEventDisableTiming) != hipSuccess) {
    set_error("stream or event creation failed");
    return -1;
  }
  return 0;
}

void ForceCommon::close_streams() {
  if (side) {
    (void)hipStreamSynchronize(side);
    (void)hipStreamDestroy(side);
    side = nullptr;
  }
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
  if (evFork)
    (void)hipEventDestroy(evFork);
  if (evJoin)
    (void)hipEventDestroy(evJoin);
  evFork = evJoin = nullptr;
}

// Full key order within runs of equal top halves (after the radix sort).
static int enqueue_order_runs(ForceCommon &c, hipStream_t st) {
  const int64_t n = c.n;
  hipLaunchKernelGGL(k_order_runs32, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st,
                     c.keyHi.p, c.keyLo.p, c.vals.p, n, c.sc.p);
  GVH_CHECK_LAUNCH();
  return 0;
}

// Bounding box, keys, sort and radix tree (everything before the typed
// gather). Typed gather + prefix sums + quad records follow in the typed TU.
int enqueue_keys_and_sort(ForceCommon &c) {
  const int64_t n = c.n;
  hipStream_t st = c.stream;
  const int bb = (int)(grid_for(n, kBlock) < 1024 ? grid_for(n, kBlock) : 1024);
  hipLaunchKernelGGL(k_bbox, dim3(bb), dim3(kBlock), 0, st,
                     (const double2 *)c.pos64.p, n, c.partials.p);
  hipLaunchKernelGGL(k_root, dim3(1), dim3(256), 0, st, c.partials.p, bb, n,
                     c.sc.p);
  hipLaunchKernelGGL(k_keys, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st,
                     (const double2 *)c.pos64.p, n, c.sc.p, c.keyHi.p, c.keyLo.p,
                     c.vals.p);
  GVH_CHECK_LAUNCH();
  if (c.sorter.sort(c.keyHi.p, c.vals.p, n, kSortFrom, 64, st))
    return -1;
  return enqueue_order_runs(c, st);
}

int enqueue_batch_keys_and_sort(ForceCommon &c, BatchState &b) {
  const int64_t n = c.n;
  hipStream_t st = c.stream;
  hipLaunchKernelGGL(k_bbox_root_g, dim3((unsigned)b.graphs), dim3(kBlock), 0, st,
                     (const double2 *)c.pos64.p, b.gStart.p, b.scG.p);
  hipLaunchKernelGGL(k_keys_g, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st,
                     (const double2 *)c.pos64.p, n, b.gid.p, b.scG.p, c.keyHi.p,
                     c.keyLo.p, c.vals.p);
  GVH_CHECK_LAUNCH();
  // graph id and the first 10 levels order the keys; runs that share them
  // are put in full order as for one graph
  if (c.sorter.sort(c.keyHi.p, c.vals.p, n, kSortFrom, 64, st))
    return -1;
  return enqueue_order_runs(c, st);
}

int enqueue_batch_fa2(ForceCommon &c, BatchState &b) {
  hipLaunchKernelGGL(k_fa2_g, dim3((unsigned)b.graphs), dim3(256), 0, c.stream,
                     b.partials.p, b.gStart.p, b.gBlocks.p, b.gPart.p,
                     c.prm.jitterTolerance, b.scG.p);
  GVH_CHECK_LAUNCH();
  return 0;
}

int enqueue_batch_stats(ForceCommon &c, BatchState &b) {
  hipLaunchKernelGGL(k_stats_g, dim3((unsigned)b.graphs), dim3(256), 0, c.stream,
                     b.partials.p, b.gStart.p, b.gBlocks.p, b.gPart.p, c.prm.gravityK,
                     b.scG.p, b.last.p);
  GVH_CHECK_LAUNCH();
  return 0;
}

int enqueue_radix_tree(ForceCommon &c) {
  const int64_t n = c.n;
  hipStream_t st = c.stream;
  if (n >= 2)
    hipLaunchKernelGGL(k_karras, dim3(grid_for(n - 1, kBlock)), dim3(kBlock), 0,
                       st, c.keyHi.p, c.keyLoS.p, n, c.left.p, c.right.p,
                       c.parent.p, c.first.p, c.last.p, c.delta.p, c.sc.p);
  const int tiles = (int)((n + kScanTile - 1) / kScanTile);
  const unsigned sb = (unsigned)((tiles + kScanWaves - 1) / kScanWaves);
  hipLaunchKernelGGL(k_scan_reduce, dim3(sb > 0 ? sb : 1), dim3(kBlock), 0,
                     st, c.spos64.p, c.smass64.p, n, c.scanPartial.p,
                     c.sc.p);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, st, c.scanPartial.p,
                     tiles, c.sc.p);
  hipLaunchKernelGGL(k_scan_final, dim3(sb > 0 ? sb : 1), dim3(kBlock), 0,
                     st, c.spos64.p, c.smass64.p, n, c.scanPartial.p,
                     c.pm.p, c.pmxHi.p, c.pmxLo.p, c.pmyHi.p, c.pmyLo.p, c.sc.p);
  GVH_CHECK_LAUNCH();
  return 0;
}

int enqueue_fa2(ForceCommon &c, int blocks) {
  hipLaunchKernelGGL(k_fa2, dim3(1), dim3(256), 0, c.stream, c.partials.p, blocks,
                     c.n, c.prm.jitterTolerance, c.sc.p, c.records.p, c.recordCap);
  GVH_CHECK_LAUNCH();
  return 0;
}

int enqueue_stats(ForceCommon &c, int blocks) {
  hipLaunchKernelGGL(k_stats, dim3(1), dim3(256), 0, c.stream, c.partials.p,
                     blocks, c.n, c.prm.gravityK, c.sc.p, c.records.p,
                     c.recordCap);
  GVH_CHECK_LAUNCH();
  return 0;
}

} // namespace gvh
