// force_typed.hpp -- precision-typed stages of the force step.
//
// Real = arithmetic type of the force math and its accumulation;
// PosT = type positions are differenced in. Instantiated as
//   <double, double>  FP64: the reference's formulas in the reference's
//                     evaluation order (force_f64.hip, no FMA contraction);
//   <float,  double>  mixed: FP64 position differences, FP32 force math;
//   <float,  float>   FP32 throughout (force_f32.hip).
#pragma once

#include "dd.hpp"
#include "force_engine.hpp"

#include <algorithm>

namespace gvh {

template <class T> struct V2;
template <> struct V2<float> { using type = float2; };
template <> struct V2<double> { using type = double2; };

constexpr int kTB = 256;            // threads per block
constexpr int kTW = kTB / kWave;    // waves per block
constexpr int kStack = 256;         // traversal stack entries per wave
constexpr int kHeavyDegree = 32;    // rows above this get a wavefront each
constexpr int kAllPairsRows = 16;   // most source partitions of the all-pairs launch
constexpr int64_t kAllPairsWaves = 16384; // wavefronts the all-pairs launch aims for

// Direction for exactly coincident points: splitmix64 finaliser over the two
// vertex ids mapped to an angle. The reference hashes buffer addresses here
// (include/core/gvizVec.h:385-394), which nothing else can reproduce.
// Returned by value: an out-parameter would put the call's results in
// scratch memory in every kernel that inlines a force evaluation.
static __device__ __noinline__ double2 coincident_dir(int64_t selfId, int64_t otherId) {
  uint64_t bits = (uint64_t)selfId ^ ((uint64_t)otherId * 0x9E3779B97F4A7C15ull);
  bits ^= bits >> 33;
  bits *= 0xff51afd7ed558ccdull;
  bits ^= bits >> 33;
  double angle = (double)(bits % 1000000u) / 1000000.0 * 2.0 * M_PI;
  return double2{cos(angle), sin(angle)};
}

// The same direction for the FP32 force math, inline: the hardware cosine
// and sine take the angle in revolutions (|error| ~1e-6, far below FP32
// force tolerances), so no library call sits in the force loops.
__device__ __forceinline__ float2 coincident_dir_f(int64_t selfId, int64_t otherId) {
  uint64_t bits = (uint64_t)selfId ^ ((uint64_t)otherId * 0x9E3779B97F4A7C15ull);
  bits ^= bits >> 33;
  bits *= 0xff51afd7ed558ccdull;
  bits ^= bits >> 33;
  const float rev = (float)(uint32_t)(bits % 1000000u) * 1e-6f;
  return float2{__builtin_amdgcn_cosf(rev), __builtin_amdgcn_sinf(rev)};
}

// ------------------------------------------------------- pairwise physics
// MODEL 0 = Fruchterman-Reingold, 1 = LinLog. `mOther` is the other body's
// mass (a pseudo-body's aggregated mass); FR scales by its integer part
// (src/embedders/gvizForceModel.c:15-23), LinLog multiplies the masses
// (include/core/gvizVec.h:552-575).
template <class Real> struct Phys;

// FP32 force math over FP64 positions: the FP32 gap d - (r_v + r_u) carries
// an absolute error of about 4e-7 d (the FP32 distance and radius sum each
// round a few times), and the magnitude m / gap amplifies it by d / gap.
// Gaps closer to zero than kNearGap * d are recomputed from the FP64
// difference and FP64 radii, which bounds a term's relative error by about
// 4e-7 / kNearGap = 4e-6.
constexpr float kNearGap = 0.1f;

template <> struct Phys<double> {
  // FP64: the reference's operation order; sqrt and division are correctly
  // rounded, and this TU is compiled without FMA contraction.
  template <int MODEL, bool NEAR>
  __device__ __forceinline__ static void rep(double dx, double dy, double dist,
                                             double /*inv*/, double radiusSum, double mSelf,
                                             double mOther, double kSq,
                                             double oc, int64_t selfId,
                                             int64_t otherId, double &ax,
                                             double &ay, double, double,
                                             const double *, int64_t) {
    double gap = dist - radiusSum;
    double ux, uy;
    if (dist > 0.0) {
      ux = dx / dist;
      uy = dy / dist;
    } else {
      const double2 u = coincident_dir(selfId, otherId);
      ux = u.x;
      uy = u.y;
    }
    if (MODEL == 1) {
      double mp = mSelf * mOther;
      double mag = gap > 0.0 ? mp / gap : mp * oc;
      ax -= mag * ux;
      ay -= mag * uy;
    } else {
      double mag = gap > 0.0 ? kSq / gap : kSq * oc;
      double t0 = 0.0 - mag * ux;
      double t1 = 0.0 - mag * uy;
      double m = trunc(mOther);
      ax += m * t0;
      ay += m * t1;
    }
  }
  __device__ __forceinline__ static double dist(double dx, double dy) {
    return sqrt(dx * dx + dy * dy);
  }
  // distance; `inv` is not used by this precision
  __device__ __forceinline__ static double dist_inv(double dx, double dy, double &inv) {
    inv = 0.0;
    return sqrt(dx * dx + dy * dy);
  }
  // opening test of src/embedders/gvizForceEmbedder.c:58-63
  __device__ __forceinline__ static bool accept(double twoh, double dist,
                                                double theta) {
    return (twoh / dist) < theta;
  }
  template <int MODEL>
  __device__ __forceinline__ static void att(double dx, double dy, double k,
                                             double &ax, double &ay) {
    double dist = sqrt(dx * dx + dy * dy);
    if (dist < 1e-9)
      return;
    double s = MODEL == 1 ? log(1.0 + dist) / dist : dist / k;
    ax += s * dx;
    ay += s * dy;
  }
  template <int MODEL>
  __device__ __forceinline__ static void att_term(double dx, double dy, double k,
                                                  double &cx, double &cy) {
    cx = 0.0;
    cy = 0.0;
    double dist = sqrt(dx * dx + dy * dy);
    if (dist < 1e-9)
      return;
    double s = MODEL == 1 ? log(1.0 + dist) / dist : dist / k;
    cx = s * dx;
    cy = s * dy;
  }
  __device__ __forceinline__ static void gravity(double x, double y, double mag,
                                                 double &fx, double &fy) {
    double dist = sqrt(x * x + y * y);
    if (dist < 1e-9)
      return;
    double s = mag / dist;
    fx -= s * x;
    fy -= s * y;
  }
  __device__ __forceinline__ static double norm(double x, double y) {
    return sqrt(x * x + y * y);
  }
};

template <> struct Phys<float> {
  // FP32: hardware reciprocal and reciprocal square root (1 ulp each).
  // NEAR (FP64 positions): when two circles nearly touch, the 1/gap
  // magnitude is ill-conditioned, so the gap is recomputed from the FP64
  // position difference and FP64 radii.
  template <int MODEL, bool NEAR>
  __device__ __forceinline__ static void rep(float dx, float dy, float dist, float inv,
                                             float radiusSum, float mSelf,
                                             float mOther, float kSq, float oc,
                                             int64_t selfId, int64_t otherId,
                                             float &ax, float &ay, double dxd,
                                             double dyd,
                                             const double *__restrict__ r64,
                                             int64_t n) {
    float gap = dist - radiusSum;
    if (NEAR && radiusSum > 0.0f && fabsf(gap) < kNearGap * dist) {
      double rs = r64[selfId] + (otherId < n ? r64[otherId] : 0.0);
      gap = (float)(sqrt(dxd * dxd + dyd * dyd) - rs);
    }
    float num = MODEL == 1 ? mSelf * mOther : kSq * truncf(mOther);
    float mag = gap > 0.0f ? num * __builtin_amdgcn_rcpf(gap) : num * oc;
    if (dist > 0.0f) {
      const float s = mag * inv; // mag times the unit vector (dx, dy) / dist
      ax -= s * dx;
      ay -= s * dy;
    } else {
      const float2 u = coincident_dir_f(selfId, otherId);
      ax -= mag * u.x;
      ay -= mag * u.y;
    }
  }
  __device__ __forceinline__ static float dist(float dx, float dy) {
    float d2 = dx * dx + dy * dy;
    return d2 > 0.0f ? d2 * __builtin_amdgcn_rsqf(d2) : 0.0f;
  }
  // distance and its reciprocal from one reciprocal square root
  __device__ __forceinline__ static float dist_inv(float dx, float dy, float &inv) {
    float d2 = dx * dx + dy * dy;
    inv = d2 > 0.0f ? __builtin_amdgcn_rsqf(d2) : 0.0f;
    return d2 * inv;
  }
  __device__ __forceinline__ static bool accept(float twoh, float dist,
                                                float theta) {
    return twoh < theta * dist;
  }
  template <int MODEL>
  __device__ __forceinline__ static void att(float dx, float dy, float k,
                                             float &ax, float &ay) {
    float d2 = dx * dx + dy * dy;
    float dist = d2 > 0.0f ? d2 * __builtin_amdgcn_rsqf(d2) : 0.0f;
    if (dist < 1e-9f)
      return;
    float s = MODEL == 1 ? log1pf(dist) * __builtin_amdgcn_rcpf(dist)
                         : dist * __builtin_amdgcn_rcpf(k);
    ax += s * dx;
    ay += s * dy;
  }
  template <int MODEL>
  __device__ __forceinline__ static void att_term(float dx, float dy, float k,
                                                  float &cx, float &cy) {
    cx = 0.0f;
    cy = 0.0f;
    float d2 = dx * dx + dy * dy;
    float dist = d2 > 0.0f ? d2 * __builtin_amdgcn_rsqf(d2) : 0.0f;
    if (dist < 1e-9f)
      return;
    float s = MODEL == 1 ? log1pf(dist) * __builtin_amdgcn_rcpf(dist)
                         : dist * __builtin_amdgcn_rcpf(k);
    cx = s * dx;
    cy = s * dy;
  }
  __device__ __forceinline__ static void gravity(float x, float y, float mag,
                                                 float &fx, float &fy) {
    float d2 = x * x + y * y;
    float dist = d2 > 0.0f ? d2 * __builtin_amdgcn_rsqf(d2) : 0.0f;
    if (dist < 1e-9f)
      return;
    float s = mag * __builtin_amdgcn_rcpf(dist);
    fx -= s * x;
    fy -= s * y;
  }
  __device__ __forceinline__ static float norm(float x, float y) {
    float d2 = x * x + y * y;
    return d2 > 0.0f ? d2 * __builtin_amdgcn_rsqf(d2) : 0.0f;
  }
};

// ------------------------------------------------------------ kernels

// Sorted copies: FP64 positions (prefix sums; the traversal's positions too
// when they are FP64), FP32 positions where the traversal takes those,
// (mass, radius) for the traversal, and the low key word for the tree.
template <class Real, class PosT>
__global__ __launch_bounds__(kTB) void k_gather(
    const double2 *__restrict__ pos64, const double2 *__restrict__ mr64,
    double *__restrict__ smass64, const uint32_t *__restrict__ vals,
    const uint64_t *__restrict__ keyLo, int64_t n, double2 *__restrict__ spos64,
    typename V2<PosT>::type *__restrict__ spos,
    typename V2<Real>::type *__restrict__ smr, uint64_t *__restrict__ keyLoS,
    const Scalars *__restrict__ sc) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || sc->stopped)
    return;
  uint32_t i = vals[s];
  double2 p = pos64[i];
  spos64[s] = p;
  if constexpr (sizeof(PosT) != sizeof(double))
    spos[s] = {(PosT)p.x, (PosT)p.y}; // FP64 positions are read from spos64
  const double2 mr = mr64[i];
  smr[s] = {(Real)mr.x, (Real)mr.y};
  smass64[s] = mr.x;
  keyLoS[s] = keyLo[i];
}

// Positions in compact order for attraction and gravity (FP32 mode only).
template <class PosT>
__global__ __launch_bounds__(kTB) void k_post(const double2 *__restrict__ pos64,
                                              int64_t n,
                                              typename V2<PosT>::type *__restrict__ out,
                                              const Scalars *__restrict__ sc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || sc->stopped)
    return;
  double2 p = pos64[i];
  out[i] = {(PosT)p.x, (PosT)p.y};
}

__device__ __forceinline__ int quad_depth(int delta, int maxDepth) {
  int q = delta >> 1;
  return q < maxDepth ? q : maxDepth;
}

// One traversal record per quad head: a binary radix node whose quad depth
// floor(delta/2) exceeds its parent's. Heads holding <= nodesPerCell points,
// or sitting at the depth limit, are leaves (the reference's leaf cells and
// overflow leaves, src/ds/gvizQuadtree.c:129-139); the others are the
// reference's branching cells (chains of single-child cells collapse onto
// the deepest one, which carries the same mass and centre of mass and a
// smaller size, so it takes the same opening decisions).
// Leaf records (count > 0). A head holding <= npc points is a leaf of the
// reference directly under its branching parent. A head at the depth limit
// holding more (an overflow leaf) hangs, when its parent cell lies above
// depth maxDepth - 1, under a chain of single-child cells that hold exactly
// its points; the reference's walk tests each of them
// (src/embedders/gvizForceEmbedder.c:50-63), and the deepest (twice the
// leaf's size) accepts whenever any does, so the record carries that cell's
// size, mass and centre of mass (twoh > 0): a target that accepts it takes
// one pseudo-body, one that rejects it takes the points. twoh = 0: the
// points always. Returns false for a branching head.
template <class Real, class PosT>
__device__ __forceinline__ bool leaf_record(NodeRec<Real, PosT> &r, int32_t count,
                                            int64_t npc, int q, int qp, int maxDepth,
                                            int32_t f, int32_t l,
                                            const double *__restrict__ pm,
                                            const double *__restrict__ pxh,
                                            const double *__restrict__ pxl,
                                            const double *__restrict__ pyh,
                                            const double *__restrict__ pyl,
                                            double rootHalf) {
  if (count > npc && q < maxDepth)
    return false;
  r.count = count;
  r.comx = r.comy = (PosT)0;
  r.mass = r.twoh = (Real)0;
  if (count > npc && qp <= maxDepth - 2) {
    const double M = pm[l + 1] - pm[f];
    const dd sx = dd_add({pxh[l + 1], pxl[l + 1]}, {-pxh[f], -pxl[f]});
    const dd sy = dd_add({pyh[l + 1], pyl[l + 1]}, {-pyh[f], -pyl[f]});
    r.comx = (PosT)((sx.hi + sx.lo) / M);
    r.comy = (PosT)((sy.hi + sy.lo) / M);
    r.mass = (Real)M;
    r.twoh = (Real)(2.0 * ldexp(rootHalf, -(maxDepth - 1)));
  }
  return true;
}

template <class Real, class PosT>
__global__ __launch_bounds__(kTB) void k_quad(
    const int32_t *__restrict__ left, const int32_t *__restrict__ right,
    const int32_t *__restrict__ parent, const int32_t *__restrict__ first,
    const int32_t *__restrict__ last, const int32_t *__restrict__ delta,
    const double *__restrict__ pm, const double *__restrict__ pxh,
    const double *__restrict__ pxl, const double *__restrict__ pyh,
    const double *__restrict__ pyl, int64_t n, int64_t npc,
    NodeRec<Real, PosT> *__restrict__ nodes, const Scalars *__restrict__ sc) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n - 1 || sc->stopped)
    return;
  const int maxDepth = sc->maxDepth;
  const int q = quad_depth(delta[b], maxDepth);
  if (b != 0 && quad_depth(delta[parent[b]], maxDepth) >= q)
    return;
  NodeRec<Real, PosT> r;
  const int32_t f = first[b], l = last[b];
  const int32_t count = l - f + 1;
  r.first = f;
  r.child[0] = r.child[1] = r.child[2] = r.child[3] = kNone;
  const int qp = b == 0 ? -1 : quad_depth(delta[parent[b]], maxDepth);
  if (leaf_record(r, count, npc, q, qp, maxDepth, f, l, pm, pxh, pxl, pyh, pyl,
                  sc->rootHalf)) {
    nodes[b] = r;
    return;
  }
  r.count = 0;
  double M = pm[l + 1] - pm[f];
  dd sx = dd_add({pxh[l + 1], pxl[l + 1]}, {-pxh[f], -pxl[f]});
  dd sy = dd_add({pyh[l + 1], pyl[l + 1]}, {-pyh[f], -pyl[f]});
  r.comx = (PosT)((sx.hi + sx.lo) / M);
  r.comy = (PosT)((sy.hi + sy.lo) / M);
  r.mass = (Real)M;
  r.twoh = (Real)(2.0 * ldexp(sc->rootHalf, -q));
  int k = 0;
  const int32_t kids[2] = {left[b], right[b]};
  for (int c = 0; c < 2; c++) {
    int32_t ch = kids[c];
    if (ch < 0 || quad_depth(delta[ch], maxDepth) > q) {
      r.child[k++] = ch;
    } else {
      r.child[k++] = left[ch];
      r.child[k++] = right[ch];
    }
  }
  nodes[b] = r;
}

// ---- lane broadcasts (v_readlane with a wave-uniform lane index)
__device__ __forceinline__ int32_t rl(int32_t v, int l) {
  return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ float rl(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint64_t rl(uint64_t v, int l) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double rl(double v, int l) {
  return __longlong_as_double((long long)rl((uint64_t)__double_as_longlong(v), l));
}

#ifndef GVH_TRAV_STACK
#define GVH_TRAV_STACK 384
#endif
#ifndef GVH_TRAV_DEFER
#define GVH_TRAV_DEFER 72
#endif
// stack entries encode points and tagged cells as negative int32
constexpr int64_t kTravMaxVertices = int64_t(1) << 29;
#ifndef GVH_TRAV_MINBLOCKS
#define GVH_TRAV_MINBLOCKS 1
#endif
constexpr int kTravStack = GVH_TRAV_STACK;      // pending (node, mask) per wave
constexpr int kTravDefer = GVH_TRAV_DEFER;      // deferred bodies per wave
// Stack entries held back from the wide phase. Once the stack holds more than
// kTravStack - reserve - 3 entries the wavefront takes one entry per step,
// which is a depth-first walk: each popped cell adds at most 3 net entries
// per level it descends, so a tree D levels deep needs at most 3 D + 3 slots
// above that point and the stack can never overflow. The traversal takes
// reserve = 3 D + 4 with D the levels the root cell's size allows before the
// depth limit, plus 2 (Scalars::treeDepth), at most kTravReserve: a cell lies
// at most 64 levels deep.
constexpr int kTravReserve = 3 * 64 + 4;
static_assert(kTravStack - kTravReserve >= kWave, "traversal stack too small");

// A deferred body as every lane reads it back from LDS: position, the lanes
// it acts on, mass, radius (0 for a cell) and vertex id (n for a cell).
template <class Real, class PosT> struct alignas(16) TravBody {
  PosT x, y;
  uint64_t mask;
  Real m, r;
  int32_t id;
};

// FP32 pair term of the traversal with the target's own factor left out
// (LinLog: the target's mass, multiplied in once at the end) and no guard
// for d2 == 0. num is the source's part of the numerator (FR:
// k^2 * trunc(mass), LinLog: its mass), formed once when the body is listed.
// A pair with d2 == 0 makes the term non-finite; rep_pair_guarded gives it
// the reference's treatment (coincident direction, magnitude num * oc).
// Without overlap prevention every radius is 0, the gap is the distance and
// the term is num * d / |d|^2. With it, a gap below kNearGap of the distance is
// recomputed from the FP64 difference and radii when positions are FP64
// (a zero radius sum never meets that test).
template <bool OVERLAP, bool NEAR>
__device__ __forceinline__ void rep_pair(float dx, float dy, float d2, float num,
                                         float radiusSum, float oc, double dxd, double dyd,
                                         const double *__restrict__ r64, int64_t selfId,
                                         int64_t otherId, int64_t n, float &ax, float &ay) {
  if constexpr (!OVERLAP) {
    const float s = num * __builtin_amdgcn_rcpf(d2);
    ax -= s * dx;
    ay -= s * dy;
  } else {
    const float inv = __builtin_amdgcn_rsqf(d2);
    const float dist = d2 * inv;
    float gap = dist - radiusSum;
    if (NEAR && fabsf(gap) < kNearGap * dist) {
      const double rs = r64[selfId] + (otherId < n ? r64[otherId] : 0.0);
      gap = (float)(sqrt(dxd * dxd + dyd * dyd) - rs);
    }
    const float s = num * (gap > 0.0f ? __builtin_amdgcn_rcpf(gap) : oc) * inv;
    ax -= s * dx;
    ay -= s * dy;
  }
}

template <bool OVERLAP, bool NEAR>
__device__ __forceinline__ void rep_pair_guarded(float dx, float dy, float d2, float num,
                                                 float radiusSum, float oc, double dxd,
                                                 double dyd, const double *__restrict__ r64,
                                                 int64_t selfId, int64_t otherId, int64_t n,
                                                 float &ax, float &ay) {
  if (d2 > 0.0f) {
    rep_pair<OVERLAP, NEAR>(dx, dy, d2, num, radiusSum, oc, dxd, dyd, r64, selfId, otherId,
                            n, ax, ay);
  } else {
    const float2 u = coincident_dir_f(selfId, otherId);
    const float mag = num * oc;
    ax -= mag * u.x;
    ay -= mag * u.y;
  }
}

// wave-uniform double held in scalar registers
__device__ __forceinline__ double uniform(double v) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Barnes-Hut repulsion, one wavefront per 64 consecutive targets in key
// order. Every lane ends up with exactly the interaction set of the
// reference's per-vertex recursion (src/embedders/gvizForceEmbedder.c:35-74);
// the wavefront walks the union of those sets with a shared stack of
// (node, lane mask) entries, up to 64 entries per step, one per lane:
//   * each lane fetches one entry's record and classifies it against the
//     bounding box of the wavefront's targets: a cell that every target would
//     accept (or open), with a relative margin far above the test's rounding,
//     is accepted (or opened) for all lanes of its mask at once;
//   * accepted cells and leaf points are deferred: the lane that meets one
//     writes its body to a per-wave LDS list, which is later evaluated
//     densely, every lane reading each body with one broadcast LDS load;
//   * cells the box cannot decide are tested by each lane, as the reference
//     does, and opened for the lanes that reject them.
// Stack entries: a cell c >= 0; a point p as -1 - p; a cell to be opened
// for the entry's lanes (put back by the undecided test) as -1 - n - c,
// which keeps every entry an int32 for n < 2^29 vertices.
// WPB wavefronts per workgroup: the waves never synchronise with each
// other, so WPB only changes how the hardware packs them onto the CUs.
template <class Real, class PosT, int MODEL, bool OVERLAP, int WPB>
__global__ __launch_bounds__(WPB * kWave, GVH_TRAV_MINBLOCKS) void k_traverse(
    const NodeRec<Real, PosT> *__restrict__ nodes,
    const typename V2<PosT>::type *__restrict__ spos,
    const typename V2<Real>::type *__restrict__ smr,
    const uint32_t *__restrict__ sidx, int64_t n, Real theta, double theta64, Real kSq,
    Real oc, typename V2<Real>::type *__restrict__ rep, int32_t *__restrict__ inter,
    const double *__restrict__ r64, Scalars *__restrict__ sc,
    unsigned long long *__restrict__ tstats, SegView sv) {
  constexpr bool NEAR = sizeof(PosT) == sizeof(double);
  // FP32 force math: pair terms from rep_pair, flushes replayed with guards
  // when their sum is not finite; FP64 keeps the reference's formulas
  constexpr bool FAST = sizeof(Real) == sizeof(float);
  __shared__ int32_t sNode[WPB][kTravStack];
  __shared__ uint64_t sMask[WPB][kTravStack];
  __shared__ TravBody<Real, PosT> dBody[WPB][kTravDefer];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  int64_t s;
  bool valid;
  int32_t root;
  if (sv.waveGraph) {
    // batch: this wave serves 64 consecutive positions of one graph
    const int64_t w = (int64_t)blockIdx.x * WPB + wave;
    if (w >= sv.waves)
      return;
    const int32_t g = sv.waveGraph[w];
    sc += g;
    s = sv.waveBase[w] + lane;
    valid = s < sv.gStart[g + 1];
    root = sv.rootOf[g];
    // a one-vertex graph has no cell: its only body is the target itself
    if (root == -1)
      root = -(int32_t)(sv.gStart[g] + 1);
  } else {
    // workgroups go to the 8 XCDs in turn (b mod 8): XCD x takes the x-th
    // contiguous eighth of the target groups, so each L2 serves one region
    // of the tree
    const int64_t G = gridDim.x, b = blockIdx.x;
    const int64_t q = G >> 3, r = G & 7, x = b & 7;
    const int64_t wb = x * q + (x < r ? x : r) + (b >> 3);
    s = (wb * WPB + wave) * kWave + lane;
    valid = s < n;
    root = n >= 2 ? 0 : -1;
  }
  if (sc->stopped)
    return;
  const uint64_t live = __ballot(valid);
  if (live == 0)
    return;
  const uint64_t ltMask = lane ? (~0ull >> (kWave - lane)) : 0ull;
  // A point never acts on itself: its own lane, if it is one of this wave's
  // targets, is cleared from the mask when the point is deferred.
  const int64_t base = (int64_t)rl((uint64_t)(s - lane), 0);
  auto notSelf = [&](int64_t p, uint64_t m) -> uint64_t {
    const int64_t rel = p - base;
    return (rel >= 0 && rel < kWave) ? (m & ~(1ull << rel)) : m;
  };
  typename V2<PosT>::type me = valid ? spos[s] : typename V2<PosT>::type{0, 0};
  typename V2<Real>::type mr = valid ? smr[s] : typename V2<Real>::type{0, 0};
  const int64_t selfId = valid ? (int64_t)sidx[s] : 0;
  const double bx0 = uniform(wave_min(valid ? (double)me.x : INFINITY));
  const double bx1 = uniform(wave_max(valid ? (double)me.x : -INFINITY));
  const double by0 = uniform(wave_min(valid ? (double)me.y : INFINITY));
  const double by1 = uniform(wave_max(valid ? (double)me.y : -INFINITY));
  const double margin = sizeof(Real) == sizeof(double) ? 1e-12 : 1e-5;
  // squared thresholds: the box tests compare squares (no square roots); the
  // margin is far above the rounding of either form
  const double thLo = (double)theta * (1.0 - margin);
  const double thHi = (double)theta * (1.0 + margin);
  // (in scalar registers: theta reaches the kernel in the force math's type)
  const double thLo2 = uniform(thLo * thLo), thHi2 = uniform(thHi * thHi);
  // Largest target radius of the wavefront. A body farther from the targets'
  // box than (rHi + its radius) / 0.9, with a margin far above the FP32
  // rounding of the gap, meets every target with a positive gap outside the
  // near band (kNearGap): its terms take num * inv / gap with no tests.
  const double rHi = uniform(wave_max(valid ? (double)mr.y : -INFINITY));
  // interactions per target are counted only when asked for
  const bool counting = inter != nullptr;
  // stack entries held back from the wide phase for this step's tree
  const int32_t depth = __builtin_amdgcn_readfirstlane(sc->treeDepth);
  const int reserve = 3 * depth + 4 < kTravReserve ? 3 * depth + 4 : kTravReserve;
  Real ax = 0, ay = 0;
  int32_t cnt = 0;
  int nd = 0;
  // optional counters (GVIZHIP_TRAVERSAL_STATS): steps, deferred bodies,
  // their mask lanes, undecided cells, their mask lanes, max stack depth,
  // bodies taken in the plain far form, those of them with every lane in
  // their mask
  unsigned long long st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // the far test above for a body at (x, y) with radius rb
  auto farBody = [&](double x, double y, double rb) -> bool {
    if constexpr (!OVERLAP) {
      return true;
    } else {
      const double gx = fmax(fmax(bx0 - x, x - bx1), 0.0);
      const double gy = fmax(fmax(by0 - y, y - by1), 0.0);
      const double sr = rHi + rb;
      return 0.8099 * (gx * gx + gy * gy) > sr * sr;
    }
  };

  // The source's numerator as listed in a body (rep_pair); FP64 lists the mass.
  auto numOf = [&](Real m) -> Real {
    if constexpr (FAST)
      return MODEL == 1 ? m : kSq * truncf(m);
    else
      return m;
  };
  // Deferred body b against this lane's target, if b's mask holds the lane.
  auto body = [&](const TravBody<Real, PosT> &b, bool guarded) {
    if ((b.mask >> lane) & 1ull) {
      const PosT ex = b.x - me.x, ey = b.y - me.y;
      Real dx = (Real)ex, dy = (Real)ey;
      if constexpr (FAST) {
        const Real d2 = dx * dx + dy * dy;
        if (guarded)
          rep_pair_guarded<OVERLAP, NEAR>(dx, dy, d2, b.m, mr.y + b.r, oc, ex, ey, r64,
                                          selfId, (int64_t)b.id, n, ax, ay);
        else
          rep_pair<OVERLAP, NEAR>(dx, dy, d2, b.m, mr.y + b.r, oc, ex, ey, r64, selfId,
                                  (int64_t)b.id, n, ax, ay);
      } else {
        Real inv;
        Real dist = Phys<Real>::dist_inv(dx, dy, inv);
        Phys<Real>::template rep<MODEL, NEAR>(dx, dy, dist, inv, mr.y + b.r, mr.x, b.m, kSq,
                                              oc, selfId, (int64_t)b.id, ax, ay, ex, ey,
                                              r64, n);
      }
    }
  };
  // A far body (FP32 force math): the gap is positive and outside the near
  // band for every target, so rep_pair's term is num * inv / gap (num / d^2
  // without overlap prevention). ALL: every lane holds a target of its mask.
  auto farTerm = [&](const TravBody<Real, PosT> &b, bool all) {
    if constexpr (FAST) {
      if (all || ((b.mask >> lane) & 1ull)) {
        const PosT ex = b.x - me.x, ey = b.y - me.y;
        const float dx = (float)ex, dy = (float)ey;
        const float d2 = dx * dx + dy * dy;
        float s;
        if constexpr (OVERLAP) {
          const float inv = __builtin_amdgcn_rsqf(d2);
          const float gap = d2 * inv - (mr.y + b.r);
          s = b.m * __builtin_amdgcn_rcpf(gap) * inv;
        } else {
          s = b.m * __builtin_amdgcn_rcpf(d2);
        }
        ax -= s * dx;
        ay -= s * dy;
      }
    }
  };

  // Evaluates the deferred bodies 64 at a time, two per iteration so the
  // second body's LDS loads overlap the first's math. In FP32 each lane first
  // classifies one body of the 64 and the list is reordered into general
  // bodies, far bodies with partial masks and far bodies with full masks, each
  // run evaluated in its own loop in list order; a lane whose sums are no
  // longer finite (a pair with d2 == 0) replays the list from the sums before
  // it with guarded terms.
  auto flush = [&]() {
    if (tstats) {
      st[1] += nd;
      for (int t = 0; t < nd; t++)
        st[2] += __popcll(rl(dBody[wave][t].mask, 0));
    }
    const Real ax0 = ax, ay0 = ay;
    for (int base = 0; base < nd; base += kWave) {
      const int end = nd < base + kWave ? nd : base + kWave;
      int g1 = end, f1 = end;
      if constexpr (FAST) {
        // lane k classifies body base + k and moves it, in place, to its
        // class's run: general bodies first, then far ones with partial
        // masks, then far ones every lane takes
        const bool on = base + lane < end;
        const int src = on ? base + lane : base;
        bool f = false, u = false;
        if (on) {
          const TravBody<Real, PosT> &c = dBody[wave][src];
          f = farBody((double)c.x, (double)c.y, (double)c.r);
          u = f && c.mask == ~0ull;
        }
        const uint64_t gm = __ballot(on && !f), pm = __ballot(f && !u), um = __ballot(u);
        g1 = base + (int)__popcll(gm);
        f1 = g1 + (int)__popcll(pm);
        if (tstats) {
          st[6] += __popcll(pm) + __popcll(um);
          st[7] += __popcll(um);
        }
        // the body is read again rather than held across the test, which
        // keeps it out of registers the classification needs
        __asm__ volatile("" ::: "memory");
        const TravBody<Real, PosT> c = dBody[wave][src];
        if (on)
          dBody[wave][!f ? base + (int)__popcll(gm & ltMask)
                         : !u ? g1 + (int)__popcll(pm & ltMask)
                              : f1 + (int)__popcll(um & ltMask)] = c;
      }
      int t = base;
      for (; t + 2 <= g1; t += 2) {
        const TravBody<Real, PosT> b0 = dBody[wave][t];
        const TravBody<Real, PosT> b1 = dBody[wave][t + 1];
        body(b0, false);
        body(b1, false);
      }
      if (t < g1)
        body(dBody[wave][t], false);
      if constexpr (FAST) {
        for (t = g1; t + 2 <= f1; t += 2) {
          const TravBody<Real, PosT> b0 = dBody[wave][t];
          const TravBody<Real, PosT> b1 = dBody[wave][t + 1];
          farTerm(b0, false);
          farTerm(b1, false);
        }
        if (t < f1)
          farTerm(dBody[wave][t], false);
        for (t = f1; t + 2 <= end; t += 2) {
          const TravBody<Real, PosT> b0 = dBody[wave][t];
          const TravBody<Real, PosT> b1 = dBody[wave][t + 1];
          farTerm(b0, true);
          farTerm(b1, true);
        }
        if (t < end)
          farTerm(dBody[wave][t], true);
      }
    }
    if constexpr (FAST) {
      if (!__builtin_isfinite(ax + ay)) {
        ax = ax0;
        ay = ay0;
        for (int u = 0; u < nd; u++)
          body(dBody[wave][u], true);
      }
    }
    if (counting)
      for (int u = 0; u < nd; u++)
        cnt += (int32_t)((dBody[wave][u].mask >> lane) & 1ull);
    nd = 0;
  };
  // Writes point p's body at list slot `at`.
  auto putPoint = [&](int at, int32_t p, uint64_t mk) {
    const typename V2<PosT>::type o = spos[p];
    const typename V2<Real>::type om = smr[p];
    TravBody<Real, PosT> b;
    b.x = o.x;
    b.y = o.y;
    b.mask = mk;
    b.m = numOf(om.x);
    b.r = om.y;
    b.id = (int32_t)sidx[p];
    dBody[wave][at] = b;
  };

  int sp = 1;
  if (lane == 0) {
    sNode[wave][0] = root;
    sMask[wave][0] = live;
  }
  enum { kDefer = 0, kOpen = 1, kAmbig = 2, kRange = 3 };
  while (sp > 0) {
    if (tstats) {
      st[0]++;
      st[5] = st[5] > (unsigned long long)sp ? st[5] : (unsigned long long)sp;
    }
    // entries the stack can take children for: each open entry adds <= 3
    int take = sp < kWave ? sp : kWave;
    const int room = (kTravStack - reserve - sp) / 3;
    if (take > room)
      take = room > 0 ? room : 1;
    const int j = sp - 1 - lane; // lane 0 takes the top entry
    const bool has = lane < take;
    const int32_t node = has ? sNode[wave][j] : 0;
    const uint64_t msk = has ? sMask[wave][j] : 0ull;
    sp -= take;
    int kind = kDefer;
    // read only by lanes that loaded a record (kinds kRange, kOpen, kAmbig
    // and deferred cells)
    NodeRec<Real, PosT> r;
    if (has && node >= 0) {
      r = nodes[node];
      if (r.count > 0 && r.twoh == (Real)0) {
        kind = kRange;
      } else {
        const double cx = (double)r.comx, cy = (double)r.comy;
        const double gx = fmax(fmax(bx0 - cx, cx - bx1), 0.0);
        const double gy = fmax(fmax(by0 - cy, cy - by1), 0.0);
        const double fx = fmax(fabs(cx - bx0), fabs(cx - bx1));
        const double fy = fmax(fabs(cy - by0), fabs(cy - by1));
        const double twoh = (double)r.twoh;
        if (twoh * twoh < thLo2 * (gx * gx + gy * gy))
          kind = kDefer;
        else if (twoh * twoh >= thHi2 * (fx * fx + fy * fy))
          kind = kOpen;
        else
          kind = kAmbig;
        // an overflow leaf its targets open: its points
        if (kind == kOpen && r.count > 0)
          kind = kRange;
      }
    } else if (has && -1 - (int64_t)node >= n) {
      // a cell put back by the lanes of its mask that rejected it
      r = nodes[-1 - (int64_t)node - n];
      kind = r.count > 0 ? kRange : kOpen;
    }
    // deferred points and cells
    {
      const uint64_t dmsk = (has && node < 0) ? notSelf(-(int64_t)node - 1, msk) : msk;
      const bool dflag = has && kind == kDefer && dmsk != 0;
      const uint64_t dm = __ballot(dflag);
      if (nd + (int)__popcll(dm) > kTravDefer)
        flush();
      if (dflag) {
        const int at = nd + (int)__popcll(dm & ltMask);
        if (node < 0) {
          putPoint(at, -node - 1, dmsk);
        } else {
          TravBody<Real, PosT> b;
          b.x = r.comx;
          b.y = r.comy;
          b.mask = dmsk;
          b.m = numOf(r.mass);
          b.r = (Real)0;
          b.id = (int32_t)n;
          dBody[wave][at] = b;
        }
      }
      nd += (int)__popcll(dm);
    }
    // leaves holding several points (nodesPerCell > 1 or the depth limit)
    uint64_t rm = __ballot(has && kind == kRange);
    while (rm) {
      const int l = __builtin_ctzll(rm);
      rm &= rm - 1;
      const int32_t first = rl(r.first, l), count = rl(r.count, l);
      const uint64_t m = rl(msk, l);
      // the leaf's points, up to 64 at a time, one per lane
      for (int32_t k0 = 0; k0 < count; k0 += kWave) {
        const int32_t p = first + k0 + lane;
        const uint64_t mk = k0 + lane < count ? notSelf(p, m) : 0ull;
        const uint64_t pm = __ballot(mk != 0);
        if (nd + (int)__popcll(pm) > kTravDefer)
          flush();
        if (mk != 0)
          putPoint(nd + (int)__popcll(pm & ltMask), p, mk);
        nd += (int)__popcll(pm);
      }
    }
    // children of cells every target opens
    {
      const bool op = has && kind == kOpen;
      int kids = 0;
      if (op)
        kids = (r.child[0] != kNone) + (r.child[1] != kNone) +
               (r.child[2] != kNone) + (r.child[3] != kNone);
      // each lane's first slot: the children of the lanes below it, counted
      // from ballots of the three bits of the child count (no cross-lane
      // shuffles)
      const uint64_t k1 = __ballot(kids & 1), k2 = __ballot(kids & 2), k4 = __ballot(kids & 4);
      const int below = (int)__popcll(k1 & ltMask) + 2 * (int)__popcll(k2 & ltMask) +
                        4 * (int)__popcll(k4 & ltMask);
      const int total = (int)__popcll(k1) + 2 * (int)__popcll(k2) + 4 * (int)__popcll(k4);
      if (sp + total > kTravStack) {
        sc->traversalOverflow = 1;
        break;
      }
      if (op) {
        // children go on in reverse so NW is popped first
        int at = sp + below + kids - 1;
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (r.child[c] != kNone) {
            sNode[wave][at] = r.child[c];
            sMask[wave][at] = msk;
            at--;
          }
      }
      sp += total;
    }
    // cells the box cannot decide: staged in the free tail of the body list
    // (radius slot = cell size, id = node), then every lane of each cell's
    // mask tests it, as the reference does
    {
      const bool amb = has && kind == kAmbig;
      uint64_t am = __ballot(amb);
      const int na = (int)__popcll(am);
      if (na != 0 && nd + na > kTravDefer)
        flush();
      if (amb) {
        TravBody<Real, PosT> b;
        b.x = r.comx;
        b.y = r.comy;
        b.mask = msk;
        b.m = numOf(r.mass);
        b.r = r.twoh;
        b.id = node;
        dBody[wave][nd + (int)__popcll(am & ltMask)] = b;
      }
      // two staged cells per pass: their LDS reads and per-lane tests
      // overlap; put-backs keep the order of the cells
      auto ambTest = [&](const TravBody<Real, PosT> &b) -> bool {
        if (tstats) {
          st[3]++;
          st[4] += __popcll(rl(b.mask, 0));
        }
        bool open = false;
        if ((b.mask >> lane) & 1ull) {
          const PosT ex = b.x - me.x, ey = b.y - me.y;
          Real dx = (Real)ex, dy = (Real)ey;
          if constexpr (FAST) {
            // d2 == 0 gives a NaN distance, which is not accepted: the cell
            // is opened, as the reference opens it at distance 0
            const Real d2 = dx * dx + dy * dy;
            const Real dist = d2 * __builtin_amdgcn_rsqf(d2);
            bool accept = Phys<Real>::accept(b.r, dist, theta);
            if constexpr (NEAR) {
              // Within the FP32 test's rounding of the threshold (2h and the
              // distance each carry a few ulps) the reference's FP64 test
              // decides, so the interaction set stays the reference's.
              // 2h = RN32(rootHalf) 2^(1-q) exactly, which gives back the
              // FP64 2h = rootHalf 2^(1-q) through an exact power-of-two ratio.
              const float rhs = theta * dist;
              if (fabsf(b.r - rhs) <= 4e-6f * rhs) {
                const double rh = sc->rootHalf;
                const double twoh64 =
                    b.r >= 1.2e-38f ? rh * ((double)b.r / (double)(float)rh) : (double)b.r;
                accept = (twoh64 / sqrt(ex * ex + ey * ey)) < theta64;
              }
            }
            if (accept) {
              rep_pair<OVERLAP, NEAR>(dx, dy, d2, b.m, mr.y, oc, ex, ey, r64, selfId, n, n,
                                      ax, ay);
            } else {
              open = true;
            }
          } else {
            Real inv;
            Real dist = Phys<Real>::dist_inv(dx, dy, inv);
            if (Phys<Real>::accept(b.r, dist, theta)) {
              Phys<Real>::template rep<MODEL, NEAR>(dx, dy, dist, inv, mr.y + (Real)0, mr.x,
                                                    b.m, kSq, oc, selfId, n, ax, ay, ex, ey,
                                                    r64, n);
            } else {
              open = true;
            }
          }
        }
        return open;
      };
      auto putBack = [&](uint64_t om, int32_t id) -> bool {
        if (om) {
          // the cell goes back on the stack once, tagged, and is opened for
          // the lanes of om (in parallel with other entries) when taken
          if (sp + 1 > kTravStack) {
            sc->traversalOverflow = 1;
            sp = 0;
            return false;
          }
          if (lane == 0) {
            sNode[wave][sp] = (int32_t)(-1 - n - (int64_t)id);
            sMask[wave][sp] = om;
          }
          sp++;
        }
        return true;
      };
      for (int k = 0; k < na; k += 2) {
        const bool two = k + 1 < na;
        const TravBody<Real, PosT> b0 = dBody[wave][nd + k];
        const TravBody<Real, PosT> b1 = dBody[wave][nd + k + (two ? 1 : 0)];
        const bool o0 = ambTest(b0);
        const bool o1 = two ? ambTest(b1) : false;
        const uint64_t om0 = __ballot(o0), om1 = __ballot(o1);
        // the lanes of a cell's mask that did not open it took it
        if (counting)
          cnt += (int32_t)(((b0.mask & ~om0) >> lane) & 1ull) +
                 (two ? (int32_t)(((b1.mask & ~om1) >> lane) & 1ull) : 0);
        if (!putBack(om0, b0.id) || !putBack(om1, b1.id))
          break;
      }
    }
  }
  flush();
  if (tstats && lane == 0) {
    for (int k = 0; k < 5; k++)
      atomicAdd(&tstats[k], st[k]);
    atomicMax(&tstats[5], st[5]);
    for (int k = 6; k < 8; k++)
      atomicAdd(&tstats[k], st[k]);
  }
  if (valid) {
    if (FAST && MODEL == 1) {
      ax *= mr.x;
      ay *= mr.x;
    }
    rep[selfId] = {ax, ay};
    if (inter)
      inter[selfId] = cnt;
  }
}

// Exact all-pairs repulsion (src/embedders/gvizForceEmbedder.c:104-115):
// sources stream through LDS in tiles of 256, in ascending id order. Block
// row y takes the sources [y * span, (y + 1) * span): with one row (FP64)
// every target sums all sources in the reference's order straight into
// `out`; with several (FP32 force math) each row writes its partial sum to
// out + y * n and k_allpairs_sum adds the rows in order, which gives the
// launch gridDim.y times more wavefronts than there are target tiles.
// Only the tile holding the block's own targets tests for the self pair.
// OVERLAP false (every radius 0) with FP32 force math: the gap is the
// distance, so mag times the unit vector is num * d / |d|^2, one reciprocal
// per pair; num is formed once per source as it enters LDS (FR:
// k^2 * trunc(mass); LinLog: the source's mass, the target's own mass
// multiplied in once at the end).
template <class Real, class PosT, int MODEL, bool OVERLAP>
__global__ __launch_bounds__(kTB) void k_allpairs(
    const typename V2<PosT>::type *__restrict__ pos,
    const Real *__restrict__ massT, const Real *__restrict__ radT, int64_t n, int64_t span,
    Real kSq, Real oc, typename V2<Real>::type *__restrict__ out,
    int32_t *__restrict__ inter, const double *__restrict__ r64,
    const Scalars *__restrict__ sc) {
  constexpr bool NEAR = sizeof(PosT) == sizeof(double);
  constexpr bool SHORT = !OVERLAP && sizeof(Real) == sizeof(float);
  __shared__ typename V2<PosT>::type tPos[kTB];
  __shared__ Real tMass[kTB];
  __shared__ Real tRad[SHORT ? 1 : kTB];
  if (sc->stopped)
    return;
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  const bool valid = i < n;
  typename V2<PosT>::type me = valid ? pos[i] : typename V2<PosT>::type{0, 0};
  const Real mv = valid ? massT[i] : (Real)0;
  const Real rv = valid ? radT[i] : (Real)0;
  Real ax = 0, ay = 0;
  auto pair = [&](int t, int64_t jj) {
    const PosT ex = tPos[t].x - me.x, ey = tPos[t].y - me.y;
    Real dx = (Real)ex, dy = (Real)ey;
    Real inv;
    Real dist = Phys<Real>::dist_inv(dx, dy, inv);
    Phys<Real>::template rep<MODEL, NEAR>(dx, dy, dist, inv, rv + tRad[t], mv, tMass[t], kSq,
                                          oc, i, jj, ax, ay, ex, ey, r64, n);
  };
  // SHORT: d2 of source t. The main loop has no test: a pair with d2 == 0
  // (the target itself, exactly coincident points) gives inf * 0 = NaN, so
  // a lane whose tile sum is not finite redoes that tile pair by pair from
  // the sums before it, with the coincident rule for d2 <= 0.
  auto d2_of = [&](int t, Real &dx, Real &dy) {
    const PosT ex = tPos[t].x - me.x, ey = tPos[t].y - me.y;
    dx = (Real)ex;
    dy = (Real)ey;
    return dx * dx + dy * dy;
  };
  const int64_t own = (int64_t)blockIdx.x * kTB;
  const int64_t j0 = (int64_t)blockIdx.y * span;
  const int64_t j1 = j0 + span < n ? j0 + span : n;
  for (int64_t jb = j0; jb < j1; jb += kTB) {
    const int64_t j = jb + threadIdx.x;
    if (j < j1) {
      tPos[threadIdx.x] = pos[j];
      if constexpr (SHORT) {
        tMass[threadIdx.x] = MODEL == 1 ? massT[j] : kSq * truncf(massT[j]);
      } else {
        tMass[threadIdx.x] = massT[j];
        tRad[threadIdx.x] = radT[j];
      }
    }
    __syncthreads();
    const int len = (int)((j1 - jb) < kTB ? (j1 - jb) : kTB);
    if (valid) {
      if constexpr (SHORT) {
        const Real bx = ax, by = ay;
#pragma unroll 8
        for (int t = 0; t < len; t++) {
          Real dx, dy;
          const Real d2 = d2_of(t, dx, dy);
          const Real s = tMass[t] * __builtin_amdgcn_rcpf(d2);
          ax -= s * dx;
          ay -= s * dy;
        }
        if (!__builtin_isfinite(ax + ay)) {
          ax = bx;
          ay = by;
          for (int t = 0; t < len; t++) {
            Real dx, dy;
            const Real d2 = d2_of(t, dx, dy);
            if (d2 > (Real)0) {
              const Real s = tMass[t] * __builtin_amdgcn_rcpf(d2);
              ax -= s * dx;
              ay -= s * dy;
            } else if (jb + t != i) {
              const float2 u = coincident_dir_f(i, jb + t);
              const Real mag = tMass[t] * oc;
              ax -= mag * u.x;
              ay -= mag * u.y;
            }
          }
        }
      } else if (jb == own) {
        for (int t = 0; t < len; t++) {
          if (jb + t == i)
            continue;
          pair(t, jb + t);
        }
      } else {
        for (int t = 0; t < len; t++)
          pair(t, jb + t);
      }
    }
    __syncthreads();
  }
  if (valid) {
    if (SHORT && MODEL == 1) {
      ax *= mv;
      ay *= mv;
    }
    out[(int64_t)blockIdx.y * n + i] = {ax, ay};
    if (inter && blockIdx.y == 0)
      inter[i] = (int32_t)(n - 1);
  }
}

// rep[i] = sum of the `parts` partial rows, in row order.
template <class Real>
__global__ __launch_bounds__(kTB) void k_allpairs_sum(
    const typename V2<Real>::type *__restrict__ part, int64_t n, int parts,
    typename V2<Real>::type *__restrict__ rep, const Scalars *__restrict__ sc) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n || sc->stopped)
    return;
  Real ax = 0, ay = 0;
  for (int p = 0; p < parts; p++) {
    const typename V2<Real>::type v = part[(int64_t)p * n + i];
    ax += v.x;
    ay += v.y;
  }
  rep[i] = {ax, ay};
}

template <class Real, class PosT, int MODEL>
__global__ __launch_bounds__(kTB) void k_att_heavy(
    const typename V2<PosT>::type *__restrict__ pos,
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const int32_t *__restrict__ rows, int64_t nrows, Real edgeLength,
    typename V2<Real>::type *__restrict__ out, const Scalars *__restrict__ sc) {
  if (sc->stopped)
    return;
  const int lane = threadIdx.x & (kWave - 1);
  for (int64_t w = (int64_t)blockIdx.x * kTW + threadIdx.x / kWave; w < nrows;
       w += (int64_t)gridDim.x * kTW) {
    const int32_t v = rows[w];
    const typename V2<PosT>::type pv = pos[v];
    const int32_t e0 = off[v], e1 = off[v + 1];
    Real ax = 0, ay = 0;
    for (int32_t base = e0; base < e1; base += kWave) {
      const int32_t e = base + lane;
      Real cx = 0, cy = 0;
      if (e < e1) {
        typename V2<PosT>::type u = pos[adj[e]];
        Phys<Real>::template att_term<MODEL>((Real)(u.x - pv.x), (Real)(u.y - pv.y),
                                             edgeLength, cx, cy);
      }
      if (sizeof(Real) == sizeof(double)) {
        // FP64: the reference's left-to-right sum
        const int cnt = e1 - base < kWave ? e1 - base : kWave;
        for (int l = 0; l < cnt; l++) {
          ax += __shfl(cx, l, kWave);
          ay += __shfl(cy, l, kWave);
        }
      } else {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          cx += __shfl_xor(cx, off, kWave);
          cy += __shfl_xor(cy, off, kWave);
        }
        ax += cx;
        ay += cy;
      }
    }
    if (lane == 0)
      out[v] = {ax, ay};
  }
}

// Attraction along the adjacency in insertion order, then the per-vertex
// epilogue of src/embedders/gvizForceEmbedder.c:117-124 and the swinging /
// traction of :128-138; per-block FP64 partial sums of both go to partial.
// Attraction of the vertices with at most kHeavyDegree neighbours, one
// thread each, summed in adjacency order (k_att_heavy takes the others).
template <class Real, class PosT, int MODEL>
__global__ __launch_bounds__(kTB) void k_att_light(
    const typename V2<PosT>::type *__restrict__ pos, const int32_t *__restrict__ off,
    const int32_t *__restrict__ adj, int64_t n, Real edgeLength,
    typename V2<Real>::type *__restrict__ out, const Scalars *__restrict__ sc) {
  const int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n || sc->stopped)
    return;
  const int32_t e0 = off[i], e1 = off[i + 1];
  if (e1 - e0 > kHeavyDegree)
    return;
  const typename V2<PosT>::type v = pos[i];
  Real ax = 0, ay = 0;
  for (int32_t e = e0; e < e1; e++) {
    const typename V2<PosT>::type u = pos[adj[e]];
    Phys<Real>::template att<MODEL>((Real)(u.x - v.x), (Real)(u.y - v.y), edgeLength, ax,
                                    ay);
  }
  out[i] = {ax, ay};
}

template <class Real, class PosT, int MODEL>
__global__ __launch_bounds__(kTB) void k_forces(
    const typename V2<PosT>::type *__restrict__ pos,
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const int32_t *__restrict__ degree, const Real *__restrict__ massT,
    const typename V2<Real>::type *__restrict__ rep, int64_t n, Real edgeLength,
    double gravityK, typename V2<Real>::type *__restrict__ sf,
    typename V2<Real>::type *__restrict__ disp, Real *__restrict__ swOut,
    Real *__restrict__ attMagOut, Real *__restrict__ repMagOut,
    double *__restrict__ partial, const typename V2<Real>::type *__restrict__ att,
    const Scalars *__restrict__ sc, SegView sv) {
  __shared__ double smem[2 * 16];
  int64_t start = 0, len = n, part = 0;
  int64_t blk = blockIdx.x, nblk = gridDim.x;
  if (sv.vbGraph) {
    const int32_t g = sv.vbGraph[blockIdx.x];
    blk = sv.vbIdx[blockIdx.x];
    start = sv.gStart[g];
    len = sv.gStart[g + 1] - start;
    nblk = sv.gBlocks[g];
    part = sv.gPart[g];
    sc += g;
  }
  if (sc->stopped)
    return;
  double accS = 0.0, accT = 0.0;
  for (int64_t li = blk * kTB + threadIdx.x; li < len; li += nblk * kTB) {
    const int64_t i = start + li;
    typename V2<PosT>::type v = pos[i];
    const typename V2<Real>::type a = att[i];
    const Real ax = a.x, ay = a.y;
    typename V2<Real>::type rp = rep[i];
    Real fx = (Real)0, fy = (Real)0;
    fx += (Real)1 * ax;
    fy += (Real)1 * ay;
    fx += (Real)1 * rp.x;
    fy += (Real)1 * rp.y;
    attMagOut[i] = Phys<Real>::norm(ax, ay);
    repMagOut[i] = Phys<Real>::norm(rp.x, rp.y);
    typename V2<Real>::type old = sf[i];
    sf[i] = {fx, fy};
    const Real m = massT[i];
    Real sw = m * Phys<Real>::norm(fx - old.x, fy - old.y);
    Real tr = m * (Real)0.5 * Phys<Real>::norm(fx + old.x, fy + old.y);
    Real gmag = (Real)(gravityK * (double)(degree[i] + 1));
    Phys<Real>::gravity((Real)v.x, (Real)v.y, gmag, fx, fy);
    disp[i] = {fx, fy};
    swOut[i] = sw;
    accS += (double)sw;
    accT += (double)tr;
  }
  double vals[2] = {accS, accT};
  const bool isMax[2] = {false, false};
  block_reduce<2>(vals, isMax, smem);
  if (threadIdx.x == 0) {
    partial[2 * (part + blk)] = vals[0];
    partial[2 * (part + blk) + 1] = vals[1];
  }
}

// Speed-scaled displacement (src/embedders/gvizForceEmbedder.c:187-197) and
// the position update of :467-477, in FP64 on the master positions.
template <class Real>
__global__ __launch_bounds__(kTB) void k_apply(
    const typename V2<Real>::type *__restrict__ disp,
    const Real *__restrict__ sw, const Real *__restrict__ attMag,
    const Real *__restrict__ repMag, int64_t n, double2 *__restrict__ pos64,
    double *__restrict__ partial, const Scalars *__restrict__ sc, SegView sv) {
  __shared__ double smem[3 * 16];
  int64_t start = 0, len = n, part = 0;
  int64_t blk = blockIdx.x, nblk = gridDim.x;
  if (sv.vbGraph) {
    const int32_t g = sv.vbGraph[blockIdx.x];
    blk = sv.vbIdx[blockIdx.x];
    start = sv.gStart[g];
    len = sv.gStart[g + 1] - start;
    nblk = sv.gBlocks[g];
    part = sv.gPart[g];
    sc += g;
  }
  if (sc->stopped)
    return;
  const double gs = sc->globalSpeed;
  double vmax = 0.0, sa = 0.0, sr = 0.0;
  for (int64_t li = blk * kTB + threadIdx.x; li < len; li += nblk * kTB) {
    const int64_t i = start + li;
    double factor = gs / (1.0 + sqrt(gs * (double)sw[i]));
    typename V2<Real>::type d = disp[i];
    double o0 = (double)d.x * factor;
    double o1 = (double)d.y * factor;
    double applied = sqrt(o0 * o0 + o1 * o1);
    vmax = applied > vmax ? applied : vmax;
    sa += (double)attMag[i];
    sr += (double)repMag[i];
    double2 p = pos64[i];
    p.x += o0;
    p.y += o1;
    pos64[i] = p;
  }
  double vals[3] = {vmax, sa, sr};
  const bool isMax[3] = {true, false, false};
  block_reduce<3>(vals, isMax, smem);
  if (threadIdx.x == 0) {
    partial[3 * (part + blk)] = vals[0];
    partial[3 * (part + blk) + 1] = vals[1];
    partial[3 * (part + blk) + 2] = vals[2];
  }
}

// Typed copies of mass and radius in compact order (all-pairs and forces).
template <class Real>
__global__ __launch_bounds__(kTB) void k_typed_mr(const double *__restrict__ m64,
                                                  const double *__restrict__ r64,
                                                  int64_t n, Real *__restrict__ m,
                                                  Real *__restrict__ r,
                                                  double2 *__restrict__ mr) {
  int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (i >= n)
    return;
  m[i] = (Real)m64[i];
  r[i] = (Real)r64[i];
  mr[i] = {m64[i], r64[i]};
}

// ------------------------------------------------------------ host glue
int enqueue_keys_and_sort(ForceCommon &c);
int enqueue_radix_tree(ForceCommon &c);
int enqueue_fa2(ForceCommon &c, int blocks);
int enqueue_stats(ForceCommon &c, int blocks);

// k_traverse for the configured model; the overlap-free FP32 form when no
// vertex has a radius (FP64 always takes the reference's formulas).
template <class Real, class PosT>
int launch_traverse(ForceCommon &c, int model, int64_t waves, hipStream_t st, int32_t *inter,
                    Real theta, Real kSq, Real oc, Scalars *sc,
                    unsigned long long *tstats, SegView sv) {
  using P2 = typename V2<PosT>::type;
  using R2 = typename V2<Real>::type;
  if (c.n >= kTravMaxVertices) {
    set_error("Barnes-Hut traversal: %lld vertices exceed its limit of %lld",
              (long long)c.n, (long long)kTravMaxVertices);
    return -1;
  }
  constexpr bool F64 = sizeof(Real) == sizeof(double);
  const bool ov = F64 || c.prm.preventOverlap != 0;
  const int wpb = c.travWaves == 2 || c.travWaves == 4 ? c.travWaves : 1;
  const unsigned tb = (unsigned)((waves + wpb - 1) / wpb);
  // FP64 positions in key order: k_gather writes them once, into spos64
  const P2 *sposT = sizeof(PosT) == sizeof(double) ? (const P2 *)c.spos64.p
                                                   : (const P2 *)c.spos.p;
#define GVH_TRAVERSE_W(M, OV, W)                                                           \
  hipLaunchKernelGGL((k_traverse<Real, PosT, M, OV, W>), dim3(tb), dim3(W * kWave), 0, st,  \
                     (const NodeRec<Real, PosT> *)c.nodes.p, sposT,                       \
                     (const R2 *)c.smr.p, c.vals.p, c.n, theta, (double)c.prm.theta, kSq,  \
                     oc, (R2 *)c.rep.p,                                                   \
                     inter, c.radius64.p, sc, tstats, sv)
#define GVH_TRAVERSE(M, OV)                                                              \
  do {                                                                                   \
    if (wpb == 4)                                                                        \
      GVH_TRAVERSE_W(M, OV, 4);                                                          \
    else if (wpb == 2)                                                                   \
      GVH_TRAVERSE_W(M, OV, 2);                                                          \
    else                                                                                 \
      GVH_TRAVERSE_W(M, OV, 1);                                                          \
  } while (0)
  if (model == 1) {
    if (ov)
      GVH_TRAVERSE(1, true);
    else
      GVH_TRAVERSE(1, F64);
  } else {
    if (ov)
      GVH_TRAVERSE(0, true);
    else
      GVH_TRAVERSE(0, F64);
  }
#undef GVH_TRAVERSE
#undef GVH_TRAVERSE_W
  return 0;
}

// Attraction of every vertex into c.attHeavy on the side stream: forked from
// the main stream once the positions are final, joined before the force sums.
template <class Real, class PosT>
int fork_attraction(ForceCommon &c, const typename V2<PosT>::type *posC, int model) {
  using R2 = typename V2<Real>::type;
  hipStream_t s = c.side;
  const int64_t n = c.n;
  GVH_CHECK(hipEventRecord(c.evFork, c.stream));
  GVH_CHECK(hipStreamWaitEvent(s, c.evFork, 0));
  const unsigned hb = (unsigned)std::min<int64_t>((c.nHeavy + kTW - 1) / kTW, 2048);
  const Real el = (Real)c.prm.edgeLength;
  if (model == 1) {
    if (c.nHeavy > 0)
      hipLaunchKernelGGL((k_att_heavy<Real, PosT, 1>), dim3(hb), dim3(kTB), 0, s, posC,
                         c.csrOff.p, c.csrAdj.p, c.heavyRows.p, c.nHeavy, el,
                         (R2 *)c.attHeavy.p, c.sc.p);
    hipLaunchKernelGGL((k_att_light<Real, PosT, 1>), dim3(grid_for(n, kTB)), dim3(kTB), 0,
                       s, posC, c.csrOff.p, c.csrAdj.p, n, el, (R2 *)c.attHeavy.p, c.sc.p);
  } else {
    if (c.nHeavy > 0)
      hipLaunchKernelGGL((k_att_heavy<Real, PosT, 0>), dim3(hb), dim3(kTB), 0, s, posC,
                         c.csrOff.p, c.csrAdj.p, c.heavyRows.p, c.nHeavy, el,
                         (R2 *)c.attHeavy.p, c.sc.p);
    hipLaunchKernelGGL((k_att_light<Real, PosT, 0>), dim3(grid_for(n, kTB)), dim3(kTB), 0,
                       s, posC, c.csrOff.p, c.csrAdj.p, n, el, (R2 *)c.attHeavy.p, c.sc.p);
  }
  GVH_CHECK_LAUNCH();
  GVH_CHECK(hipEventRecord(c.evJoin, s));
  return 0;
}

template <class Real, class PosT>
int typed_step(ForceCommon &c, int mode) {
  using P2 = typename V2<PosT>::type;
  using R2 = typename V2<Real>::type;
  const int64_t n = c.n;
  hipStream_t st = c.stream;
  if (n == 0)
    return 0;
  const Real kSq = (Real)(c.prm.edgeLength * c.prm.edgeLength);
  const Real oc = (Real)c.prm.overlapConstant;
  const Real theta = (Real)c.prm.theta;
  int32_t *inter = c.wantInteractions ? c.interactions.p : nullptr;
  const P2 *posC;
  if (sizeof(PosT) == sizeof(double)) {
    posC = (const P2 *)c.pos64.p;
  } else {
    hipLaunchKernelGGL(k_post<PosT>, dim3(grid_for(n, kTB)), dim3(kTB), 0, st,
                       (const double2 *)c.pos64.p, n, (P2 *)c.posT.p, c.sc.p);
    posC = (const P2 *)c.posT.p;
  }
  const int model = c.prm.model == 1 ? 1 : 0;
  if (c.prm.barnesHut || mode == kStepTreeOnly) {
    if (enqueue_keys_and_sort(c))
      return -1;
    hipLaunchKernelGGL((k_gather<Real, PosT>), dim3(grid_for(n, kTB)), dim3(kTB),
                       0, st, (const double2 *)c.pos64.p, (const double2 *)c.mr64.p,
                       c.smass64.p, c.vals.p, c.keyLo.p, n,
                       (double2 *)c.spos64.p, (P2 *)c.spos.p, (R2 *)c.smr.p,
                       c.keyLoS.p, c.sc.p);
    GVH_CHECK_LAUNCH();
    if (enqueue_radix_tree(c))
      return -1;
    if (n >= 2)
      hipLaunchKernelGGL((k_quad<Real, PosT>), dim3(grid_for(n - 1, kTB)),
                         dim3(kTB), 0, st, c.left.p, c.right.p, c.parent.p,
                         c.first.p, c.last.p, c.delta.p, c.pm.p, c.pmxHi.p,
                         c.pmxLo.p, c.pmyHi.p, c.pmyLo.p, n,
                         c.prm.nodesPerCell > 0 ? c.prm.nodesPerCell : 1,
                         (NodeRec<Real, PosT> *)c.nodes.p, c.sc.p);
    if (mode == kStepTreeOnly) {
      GVH_CHECK_LAUNCH();
      return 0;
    }
    // attraction (memory-bound) overlaps the traversal (ALU-bound); forked at
    // the start of the step it slows the short tree-build kernels more than
    // it gains (959 against 1,027 steps/s at 1M vertices)
    if (fork_attraction<Real, PosT>(c, posC, model))
      return -1;
    // a timed launch runs without counting; the counts (and the optional
    // statistics) come from a second launch over the same tree and
    // positions, which repeats the same interactions and sums
    const bool split = c.timeTraversal && inter;
    const int64_t waves = (n + kWave - 1) / kWave;
    if (c.timeTraversal)
      GVH_CHECK(hipEventRecord(c.travBegin, st));
    if (launch_traverse<Real, PosT>(c, model, waves, st, split ? nullptr : inter, theta, kSq,
                                    oc, c.sc.p, split ? nullptr : c.travStats.p,
                                    SegView{}))
      return -1;
    GVH_CHECK_LAUNCH();
    if (c.timeTraversal)
      GVH_CHECK(hipEventRecord(c.travEnd, st));
    if (split && launch_traverse<Real, PosT>(c, model, waves, st, inter, theta, kSq, oc,
                                             c.sc.p, c.travStats.p, SegView{}))
      return -1;
  } else {
    if (fork_attraction<Real, PosT>(c, posC, model))
      return -1;
    const unsigned tb = (unsigned)((n + kTB - 1) / kTB);
    // FP32 force math: split the sources over up to kAllPairsRows block rows
    // (partials in the quadtree record buffer, which this mode leaves
    // unused), so small target counts still fill the chip; FP64 keeps one
    // row and the reference's summation order.
    int rows = 1;
    if (sizeof(Real) == sizeof(float)) {
      const int64_t fit = (int64_t)(c.nodes.n / (sizeof(R2) * (size_t)n));
      const int64_t want = (kAllPairsWaves + (int64_t)tb * kTW - 1) / ((int64_t)tb * kTW);
      rows = (int)std::max<int64_t>(1, std::min<int64_t>({want, fit, kAllPairsRows}));
    }
    const int64_t span = ((n + rows - 1) / rows + kTB - 1) / kTB * kTB;
    R2 *apOut = rows == 1 ? (R2 *)c.rep.p : (R2 *)c.nodes.p;
    if (c.timeTraversal)
      GVH_CHECK(hipEventRecord(c.travBegin, st));
    // radii are all 0 without overlap prevention; FP64 keeps the
    // reference's formulas either way
    const bool ov = c.prm.preventOverlap != 0 || sizeof(Real) == sizeof(double);
#define GVH_ALLPAIRS(M, OV)                                                      \
  hipLaunchKernelGGL((k_allpairs<Real, PosT, M, OV>), dim3(tb, rows), dim3(kTB), 0, st, \
                     posC, (const Real *)c.massT.p, (const Real *)c.radT.p, n, span,   \
                     kSq, oc, apOut, inter, c.radius64.p, c.sc.p)
    if (model == 1) {
      if (ov)
        GVH_ALLPAIRS(1, true);
      else
        GVH_ALLPAIRS(1, false);
    } else {
      if (ov)
        GVH_ALLPAIRS(0, true);
      else
        GVH_ALLPAIRS(0, false);
    }
#undef GVH_ALLPAIRS
    if (rows > 1)
      hipLaunchKernelGGL(k_allpairs_sum<Real>, dim3(tb), dim3(kTB), 0, st,
                         (const R2 *)c.nodes.p, n, rows, (R2 *)c.rep.p, c.sc.p);
    GVH_CHECK_LAUNCH();
    if (c.timeTraversal)
      GVH_CHECK(hipEventRecord(c.travEnd, st));
  }

  const unsigned rb = grid_for(n, kTB) < (unsigned)c.reductionBlocks
                          ? grid_for(n, kTB)
                          : (unsigned)c.reductionBlocks;
  GVH_CHECK(hipStreamWaitEvent(st, c.evJoin, 0));
  if (model == 1)
  {
    hipLaunchKernelGGL((k_forces<Real, PosT, 1>), dim3(rb), dim3(kTB), 0, st,
                       posC, c.csrOff.p, c.csrAdj.p, c.degree.p,
                       (const Real *)c.massT.p, (const R2 *)c.rep.p, n,
                       (Real)c.prm.edgeLength, c.prm.gravityK, (R2 *)c.sf.p,
                       (R2 *)c.disp.p, (Real *)c.sw.p, (Real *)c.attMag.p,
                       (Real *)c.repMag.p, c.partials.p, (const R2 *)c.attHeavy.p,
                       c.sc.p, SegView{});
  }
  else
  {
    hipLaunchKernelGGL((k_forces<Real, PosT, 0>), dim3(rb), dim3(kTB), 0, st,
                       posC, c.csrOff.p, c.csrAdj.p, c.degree.p,
                       (const Real *)c.massT.p, (const R2 *)c.rep.p, n,
                       (Real)c.prm.edgeLength, c.prm.gravityK, (R2 *)c.sf.p,
                       (R2 *)c.disp.p, (Real *)c.sw.p, (Real *)c.attMag.p,
                       (Real *)c.repMag.p, c.partials.p, (const R2 *)c.attHeavy.p,
                       c.sc.p, SegView{});
  }
  GVH_CHECK_LAUNCH();
  if (mode == kStepForcesOnly)
    return 0;
  if (enqueue_fa2(c, (int)rb))
    return -1;
  hipLaunchKernelGGL(k_apply<Real>, dim3(rb), dim3(kTB), 0, st,
                     (const R2 *)c.disp.p, (const Real *)c.sw.p,
                     (const Real *)c.attMag.p, (const Real *)c.repMag.p, n,
                     (double2 *)c.pos64.p, c.partials.p, c.sc.p, SegView{});
  GVH_CHECK_LAUNCH();
  return enqueue_stats(c, (int)rb);
}


// ------------------------------------------------------------ batch stages
template <int Unused = 0>
__global__ void k_reset_roots(int32_t *__restrict__ rootOf, int64_t graphs) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < graphs)
    rootOf[g] = -1;
}

// k_quad for a batch: radix nodes whose common prefix is shorter than the 12
// graph-id bits span several graphs and are skipped; within a graph the quad
// depth is (delta - 12) / 2 and the graph's root record is the head whose
// parent spans several graphs.
template <class Real, class PosT>
__global__ __launch_bounds__(kTB) void k_quad_g(
    const int32_t *__restrict__ left, const int32_t *__restrict__ right,
    const int32_t *__restrict__ parent, const int32_t *__restrict__ first,
    const int32_t *__restrict__ last, const int32_t *__restrict__ delta,
    const double *__restrict__ pm, const double *__restrict__ pxh,
    const double *__restrict__ pxl, const double *__restrict__ pyh,
    const double *__restrict__ pyl, int64_t n, int64_t npc,
    NodeRec<Real, PosT> *__restrict__ nodes, const int32_t *__restrict__ gid,
    const Scalars *__restrict__ scG, int32_t *__restrict__ rootOf) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n - 1)
    return;
  const int dl = delta[b];
  if (dl < 12)
    return;
  const int32_t f = first[b], l = last[b];
  const int32_t g = gid[f];
  const Scalars &sc = scG[g];
  if (sc.stopped)
    return;
  const int maxDepth = sc.maxDepth < kBatchLevels ? sc.maxDepth : kBatchLevels;
  const int q = quad_depth(dl - 12, maxDepth);
  const int32_t p = parent[b];
  const bool top = p < 0 || delta[p] < 12;
  if (!top && quad_depth(delta[p] - 12, maxDepth) >= q)
    return;
  if (top)
    rootOf[g] = (int32_t)b;
  NodeRec<Real, PosT> r;
  const int32_t count = l - f + 1;
  r.first = f;
  r.child[0] = r.child[1] = r.child[2] = r.child[3] = kNone;
  const int qp = top ? -1 : quad_depth(delta[p] - 12, maxDepth);
  if (leaf_record(r, count, npc, q, qp, maxDepth, f, l, pm, pxh, pxl, pyh, pyl,
                  sc.rootHalf)) {
    nodes[b] = r;
    return;
  }
  r.count = 0;
  double M = pm[l + 1] - pm[f];
  dd sx = dd_add({pxh[l + 1], pxl[l + 1]}, {-pxh[f], -pxl[f]});
  dd sy = dd_add({pyh[l + 1], pyl[l + 1]}, {-pyh[f], -pyl[f]});
  r.comx = (PosT)((sx.hi + sx.lo) / M);
  r.comy = (PosT)((sy.hi + sy.lo) / M);
  r.mass = (Real)M;
  r.twoh = (Real)(2.0 * ldexp(sc.rootHalf, -q));
  int k = 0;
  const int32_t kids[2] = {left[b], right[b]};
  for (int c = 0; c < 2; c++) {
    int32_t ch = kids[c];
    if (ch < 0 || quad_depth(delta[ch] - 12, maxDepth) > q) {
      r.child[k++] = ch;
    } else {
      r.child[k++] = left[ch];
      r.child[k++] = right[ch];
    }
  }
  nodes[b] = r;
}

template <class Real, class PosT>
int typed_batch_step(ForceCommon &c, BatchState &bs) {
  using P2 = typename V2<PosT>::type;
  using R2 = typename V2<Real>::type;
  const int64_t n = c.n;
  hipStream_t st = c.stream;
  if (n == 0)
    return 0;
  const SegView sv = bs.view();
  const Real kSq = (Real)(c.prm.edgeLength * c.prm.edgeLength);
  const Real oc = (Real)c.prm.overlapConstant;
  const Real theta = (Real)c.prm.theta;
  const P2 *posC;
  if (sizeof(PosT) == sizeof(double)) {
    posC = (const P2 *)c.pos64.p;
  } else {
    hipLaunchKernelGGL(k_post<PosT>, dim3(grid_for(n, kTB)), dim3(kTB), 0, st,
                       (const double2 *)c.pos64.p, n, (P2 *)c.posT.p, c.sc.p);
    posC = (const P2 *)c.posT.p;
  }
  const int model = c.prm.model == 1 ? 1 : 0;
  if (enqueue_batch_keys_and_sort(c, bs))
    return -1;
  hipLaunchKernelGGL((k_gather<Real, PosT>), dim3(grid_for(n, kTB)), dim3(kTB), 0, st,
                     (const double2 *)c.pos64.p, (const double2 *)c.mr64.p, c.smass64.p,
                     c.vals.p,
                     c.keyLo.p, n, (double2 *)c.spos64.p, (P2 *)c.spos.p,
                     (typename V2<Real>::type *)c.smr.p, c.keyLoS.p, c.sc.p);
  GVH_CHECK_LAUNCH();
  if (enqueue_radix_tree(c))
    return -1;
  hipLaunchKernelGGL(k_reset_roots<0>, dim3(grid_for(bs.graphs, kTB)), dim3(kTB), 0, st,
                     bs.rootOf.p, bs.graphs);
  if (n >= 2)
    hipLaunchKernelGGL((k_quad_g<Real, PosT>), dim3(grid_for(n - 1, kTB)), dim3(kTB), 0,
                       st, c.left.p, c.right.p, c.parent.p, c.first.p, c.last.p,
                       c.delta.p, c.pm.p, c.pmxHi.p, c.pmxLo.p, c.pmyHi.p, c.pmyLo.p, n,
                       c.prm.nodesPerCell > 0 ? c.prm.nodesPerCell : 1,
                       (NodeRec<Real, PosT> *)c.nodes.p, bs.gid.p, bs.scG.p,
                       bs.rootOf.p);
  GVH_CHECK_LAUNCH();
  // attraction (memory-bound) overlaps the traversal (ALU-bound)
  if (fork_attraction<Real, PosT>(c, posC, model))
    return -1;
  int32_t *inter = c.wantInteractions ? c.interactions.p : nullptr;
  // timed without counting; counted by a second launch over the same tree
  const bool split = c.timeTraversal && inter;
  if (c.timeTraversal)
    GVH_CHECK(hipEventRecord(c.travBegin, st));
  if (launch_traverse<Real, PosT>(c, model, bs.waves, st, split ? nullptr : inter, theta,
                                  kSq, oc, bs.scG.p, (unsigned long long *)nullptr, sv))
    return -1;
  GVH_CHECK_LAUNCH();
  if (c.timeTraversal)
    GVH_CHECK(hipEventRecord(c.travEnd, st));
  if (split && launch_traverse<Real, PosT>(c, model, bs.waves, st, inter, theta, kSq, oc,
                                           bs.scG.p, (unsigned long long *)nullptr, sv))
    return -1;
  const unsigned vb = (unsigned)bs.vblocks;
  GVH_CHECK(hipStreamWaitEvent(st, c.evJoin, 0));
  if (model == 1) {
    hipLaunchKernelGGL((k_forces<Real, PosT, 1>), dim3(vb), dim3(kTB), 0, st, posC,
                       c.csrOff.p, c.csrAdj.p, c.degree.p, (const Real *)c.massT.p,
                       (const R2 *)c.rep.p, n, (Real)c.prm.edgeLength, c.prm.gravityK,
                       (R2 *)c.sf.p, (R2 *)c.disp.p, (Real *)c.sw.p, (Real *)c.attMag.p,
                       (Real *)c.repMag.p, bs.partials.p, (const R2 *)c.attHeavy.p,
                       bs.scG.p, sv);
  } else {
    hipLaunchKernelGGL((k_forces<Real, PosT, 0>), dim3(vb), dim3(kTB), 0, st, posC,
                       c.csrOff.p, c.csrAdj.p, c.degree.p, (const Real *)c.massT.p,
                       (const R2 *)c.rep.p, n, (Real)c.prm.edgeLength, c.prm.gravityK,
                       (R2 *)c.sf.p, (R2 *)c.disp.p, (Real *)c.sw.p, (Real *)c.attMag.p,
                       (Real *)c.repMag.p, bs.partials.p, (const R2 *)c.attHeavy.p,
                       bs.scG.p, sv);
  }
  GVH_CHECK_LAUNCH();
  if (enqueue_batch_fa2(c, bs))
    return -1;
  hipLaunchKernelGGL(k_apply<Real>, dim3(vb), dim3(kTB), 0, st, (const R2 *)c.disp.p,
                     (const Real *)c.sw.p, (const Real *)c.attMag.p,
                     (const Real *)c.repMag.p, n, (double2 *)c.pos64.p, bs.partials.p,
                     bs.scG.p, sv);
  GVH_CHECK_LAUNCH();
  return enqueue_batch_stats(c, bs);
}

template <class Real>
int typed_upload(ForceCommon &c) {
  if (c.n == 0)
    return 0;
  hipLaunchKernelGGL(k_typed_mr<Real>, dim3(grid_for(c.n, kTB)), dim3(kTB), 0,
                     c.stream, c.mass64.p, c.radius64.p, c.n, (Real *)c.massT.p,
                     (Real *)c.radT.p, (double2 *)c.mr64.p);
  GVH_CHECK_LAUNCH();
  return 0;
}

} // namespace gvh
