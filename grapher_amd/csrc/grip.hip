// grip.hip -- GRIP multilevel layout on gfx950 (the reference's
// src/embedders/gvizGRIPEmbedder.c and src/algorithms/search/gvizKNearest.c):
//
//   * wavefront BFS: one wavefront per query walks its FIFO queue, expanding
//     a vertex's adjacency 64 entries at a time; new vertices are claimed in a
//     per-wavefront visited bitmap and appended in adjacency (discovery)
//     order with ballots, so kNN lists come out in the reference's order;
//   * MIS filtration as the lexicographically-first independent set of each
//     layer's conflict graph (two candidates conflict when their graph
//     distance is <= 2^(i-1)), which is exactly what the reference's
//     ascending greedy pass produces: conflicts come from bounded BFS balls,
//     the independent set from parallel decision rounds;
//   * the refinement round as a map over the active vertices followed by
//     FP64 block reductions for translation / rotation removal and the stats.
// Compiled with -ffp-contract=off: the FP64 arithmetic rounds like the
// reference's.
#include "../../include/gviz_hip.h"
#include "common.hpp"
#include "radix_sort.hpp"

#include <algorithm>
#include <cstdlib>
#include <deque>
#include <vector>

using namespace gvh;

namespace {

constexpr int kB = 256;
constexpr int kGather = 4; // neighbour loads in flight per thread in k_refine
constexpr int kWpb = kB / kWave;
constexpr double kEdge = 10.0;      // GVIZ_EDGE_LENGTH (gvizGRIPEmbedder.c:31)
constexpr double kGripR = 0.15;     // gvizGRIPr (:45)
constexpr double kGripS = 3.0;      // gvizGRIPs (:46)
constexpr double kFrScale = 0.05;   // FR_SCALE_FACTOR (gvizForceDirected.c:4)
constexpr double kEps = 1e-7;       // gvizNumericEpsilon

// The BFS queue is written and read back by the same wavefront; both go
// through device-scope atomics so the read-after-write order holds at L2.
__device__ __forceinline__ int32_t ld_l2(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_l2(int32_t *p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A search slot's visited bits are read and written by one wavefront (or
// one workgroup) within a kernel, and kernel boundaries order them between
// kernels, so their atomics need only workgroup scope.
__device__ __forceinline__ uint32_t bits_or(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void bits_clear(uint32_t *p) {
  __hip_atomic_store(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct BfsSlots {
  uint32_t *bits;  // S * words
  int32_t *queue;  // S * cap
  int32_t *qdist;  // S * cap
  int64_t words, cap;
};

// One wavefront's BFS from `source` (gvizKNearest.c:244-306 and, with
// maxDepth > 0, verticesWithinRadius of gvizGRIPEmbedder.c:599-670).
// Discovered vertices passing `filter` (and, when lowerOnly, with an id below
// the source) are written to out[0..k) in discovery order with their depth.
// kNN mode (stopAtK) ends the search at the k-th hit; ball mode explores the
// whole radius and reports the total hit count (may exceed k). When the
// filter admits `filterTotal` vertices (> 0), the search also ends once all
// of them (but the source) are found: the rest of the graph holds no hit, so
// the result is the same.
//
// The FIFO is walked as one stream of (queued vertex, adjacency entry)
// pairs, 64 pairs per step: lane j takes the j-th pair after the cursor
// (queue position `head`, entry `cur` of that vertex's row). The pairs keep
// the sequential order (queue position, then adjacency order), so with
// duplicates inside a step resolved in favour of the earlier lane the set of
// fresh vertices and their order are exactly those of the one-vertex-at-a-time
// search.
// A step's first pair per vertex: lane indices through a per-wave LDS hash
// (2 * 64 slots, open addressing, smallest lane kept).
// The graph may be a renumbered copy (breadth-first order, which keeps a
// ball's vertices, rows and visited bits in few cache lines): `source`,
// `filter` and the visited bits use its ids, `orig` maps them back to the
// caller's for the output and the lower-id test (null: the same ids). The
// rows keep their order, so the discovery order is the original graph's.
constexpr int kWaveHash = 2 * kWave;

__device__ __forceinline__ int32_t orig_id(const int32_t *__restrict__ orig, int32_t v) {
  return orig ? orig[v] : v;
}

__device__ int wave_bfs(const int32_t *__restrict__ off,
                        const int32_t *__restrict__ adj,
                        const uint8_t *__restrict__ filter, int32_t source, int k,
                        int maxDepth, bool lowerOnly, bool stopAtK,
                        int64_t filterTotal, uint32_t *bits, int32_t *queue,
                        int32_t *qdist, int32_t *outV, int32_t *outD,
                        int32_t *hKey, int32_t *hVal, const int32_t *__restrict__ orig) {
  const int lane = threadIdx.x & (kWave - 1);
  const int32_t srcOrig = lowerOnly ? orig_id(orig, source) : 0;
  const uint64_t lt = lane ? (~0ull >> (kWave - lane)) : 0ull;
  if (lane == 0) {
    bits_or(&bits[source >> 5], 1u << (source & 31));
    st_l2(queue, source);
    st_l2(qdist, 0);
  }
  const int64_t stopCount =
      filterTotal > 0 ? filterTotal - ((filter && filter[source]) ? 1 : 0) : -1;
  int head = 0, tail = 1, found = 0, cur = 0;
  bool done = stopCount == 0;
  while (head < tail && !done) {
    // this step's queued vertices: positions head .. head + avail - 1
    const int avail = tail - head < kWave ? tail - head : kWave;
    int32_t u = -1, du = 0, start = 0, deg = 0;
    if (lane < avail) {
      u = ld_l2(queue + head + lane);
      du = ld_l2(qdist + head + lane);
      if (!(maxDepth > 0 && du >= maxDepth)) {
        start = off[u] + (lane == 0 ? cur : 0);
        deg = off[u + 1] - start;
      }
    }
    // inclusive scan of the row lengths
    int32_t incl = deg;
    for (int o = 1; o < kWave; o <<= 1) {
      const int32_t v = __shfl_up(incl, o, kWave);
      if (lane >= o)
        incl += v;
    }
    const int32_t total = __shfl(incl, kWave - 1, kWave);
    const int pairs = total < kWave ? total : kWave;
    // lane j < pairs: owner = first lane whose inclusive sum exceeds j
    // (a fixed six-step search, every lane taking part in each shuffle)
    int owner = 0;
    for (int step = kWave / 2; step >= 1; step >>= 1) {
      const int32_t v = __shfl(incl, owner + step - 1, kWave);
      if (v <= lane)
        owner += step;
    }
    const int32_t ownStart = __shfl(start, owner, kWave);
    const int32_t ownExcl = __shfl(incl - deg, owner, kWave);
    const int32_t ownDepth = __shfl(du, owner, kWave);
    const bool valid = lane < pairs;
    const int32_t w = valid ? adj[ownStart + (lane - ownExcl)] : -1;
    // a vertex listed twice among this step's pairs: only its first counts
    hKey[lane] = -1;
    hKey[lane + kWave] = -1;
    hVal[lane] = kWave;
    hVal[lane + kWave] = kWave;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint32_t h = ((uint32_t)w * 2654435761u) & (kWaveHash - 1);
    if (valid) {
      while (true) {
        const int32_t prev = atomicCAS(&hKey[h], -1, w);
        if (prev == -1 || prev == w) {
          atomicMin(&hVal[h], lane);
          break;
        }
        h = (h + 1) & (kWaveHash - 1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const bool dup = valid && hVal[h] != lane;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    bool fresh = false;
    if (valid && !dup) {
      const uint32_t bit = 1u << (w & 31);
      fresh = !(bits_or(&bits[w >> 5], bit) & bit);
    }
    const uint64_t fm = __ballot(fresh);
    const bool pass = fresh && (!filter || filter[w]) &&
                      (!lowerOnly || orig_id(orig, w) < srcOrig);
    const uint64_t pm = __ballot(pass);
    if (pass) {
      const int slot = found + (int)__popcll(pm & lt);
      if (slot < k) {
        outV[slot] = orig_id(orig, w);
        outD[slot] = ownDepth + 1;
      }
    }
    found += (int)__popcll(pm);
    if (fresh) {
      const int at = tail + (int)__popcll(fm & lt);
      st_l2(queue + at, w);
      st_l2(qdist + at, ownDepth + 1);
    }
    tail += (int)__popcll(fm);
    // advance the cursor past the pairs taken
    if (pairs == total) {
      head += avail;
      cur = 0;
    } else {
      const int last = __shfl(owner, pairs - 1, kWave);
      const int32_t lastIncl = __shfl(incl, last, kWave);
      const int32_t lastDeg = __shfl(deg, last, kWave);
      const int32_t took = lastDeg - (lastIncl - pairs);
      if (took == lastDeg) {
        head += last + 1;
        cur = 0;
      } else {
        cur = (last == 0 ? cur : 0) + took;
        head += last;
      }
    }
    if (stopAtK && found >= k)
      done = true;
    if (stopCount > 0 && found >= stopCount)
      done = true;
  }
  // clear this query's bits: every visited vertex sits in the queue
  for (int i = lane; i < tail; i += kWave) {
    const int32_t w = ld_l2(queue + i);
    bits_clear(&bits[w >> 5]);
  }
  return found;
}

// ---- block-cooperative search for calls with few queries: one block of
// kBB threads takes kBB * kPairs (queued vertex, adjacency entry) pairs per
// step (kPairs consecutive pairs per thread, so each thread has that many
// loads in flight) of the same FIFO stream wave_bfs walks 64 at a time. A vertex is fresh at its
// first pair of the step (the smallest pair index, found through an LDS hash
// table) when no earlier step visited it, so the fresh vertices, their queue
// order and the hits are exactly wave_bfs's; only the step size differs.
#ifndef GVH_GRIP_BB
#define GVH_GRIP_BB 1024
#endif
constexpr int kBB = GVH_GRIP_BB;
constexpr int kBW = kBB / kWave;
constexpr int kPairs = 4;                // adjacency pairs per thread per pass
constexpr int kHash = 2 * kBB * kPairs;  // LDS hash over one pass's pairs

// Inclusive block scan of one int per thread; `total` gets the block sum.
__device__ __forceinline__ int block_scan(int v, int *sWave, int &total) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  int incl = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int x = __shfl_up(incl, o, kWave);
    if (lane >= o)
      incl += x;
  }
  if (lane == kWave - 1)
    sWave[wave] = incl;
  __syncthreads();
  int pre = 0, sum = 0;
#pragma unroll
  for (int w = 0; w < kBW; w++) {
    const int x = sWave[w];
    pre += w < wave ? x : 0;
    sum += x;
  }
  __syncthreads();
  total = sum;
  return incl + pre;
}

__device__ int block_bfs(const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                         const uint8_t *__restrict__ filter, int32_t source, int k,
                         int maxDepth, bool lowerOnly, bool stopAtK, int64_t filterTotal,
                         uint32_t *bits, int32_t *queue, int32_t *qdist, int32_t *outV,
                         int32_t *outD, const int32_t *__restrict__ orig) {
  const int32_t srcOrig = lowerOnly ? orig_id(orig, source) : 0;
  __shared__ int32_t sIncl[kBB], sStart[kBB], sDepth[kBB];
  __shared__ int32_t hKey[kHash], hVal[kHash];
  __shared__ int sWave[kBW];
  __shared__ int sLast;
  const int t = threadIdx.x;
  if (t == 0) {
    bits_or(&bits[source >> 5], 1u << (source & 31));
    st_l2(queue, source);
    st_l2(qdist, 0);
  }
  __syncthreads();
  const int64_t stopCount =
      filterTotal > 0 ? filterTotal - ((filter && filter[source]) ? 1 : 0) : -1;
  int head = 0, tail = 1, found = 0, cur = 0;
  bool done = stopCount == 0;
  while (head < tail && !done) {
    const int avail = tail - head < kBB ? tail - head : kBB;
    int32_t start = 0, deg = 0, du = 0;
    if (t < avail) {
      const int32_t u = ld_l2(queue + head + t);
      du = ld_l2(qdist + head + t);
      if (!(maxDepth > 0 && du >= maxDepth)) {
        start = off[u] + (t == 0 ? cur : 0);
        deg = off[u + 1] - start;
      }
    }
    int total;
    const int incl = block_scan(deg, sWave, total);
    sIncl[t] = incl;
    sStart[t] = start;
    sDepth[t] = du;
    for (int i = t; i < kHash; i += kBB) {
      hKey[i] = -1;
      hVal[i] = INT32_MAX;
    }
    __syncthreads();
    // this thread's pairs are p0 .. p0 + kPairs - 1, in queue order
    const int pairs = total < kBB * kPairs ? total : kBB * kPairs;
    const int p0 = t * kPairs;
    int32_t w[kPairs], ownDepth[kPairs];
    if (p0 < pairs) {
      // owner = first queued vertex whose inclusive sum exceeds p0
      int lo = 0, hi = avail - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sIncl[mid] > p0)
          hi = mid;
        else
          lo = mid + 1;
      }
      int owner = lo;
#pragma unroll
      for (int i = 0; i < kPairs; i++) {
        const int p = p0 + i;
        w[i] = -1;
        ownDepth[i] = 0;
        if (p < pairs) {
          while (sIncl[owner] <= p)
            owner++;
          const int excl = owner ? sIncl[owner - 1] : 0;
          w[i] = adj[sStart[owner] + (p - excl)];
          ownDepth[i] = sDepth[owner];
          if (p == pairs - 1)
            sLast = owner;
        }
      }
#pragma unroll
      for (int i = 0; i < kPairs; i++) {
        if (w[i] < 0)
          continue;
        // smallest pair index per vertex (open addressing)
        uint32_t h = ((uint32_t)w[i] * 2654435761u) & (kHash - 1);
        while (true) {
          const int32_t prev = atomicCAS(&hKey[h], -1, w[i]);
          if (prev == -1 || prev == w[i]) {
            atomicMin(&hVal[h], p0 + i);
            break;
          }
          h = (h + 1) & (kHash - 1);
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < kPairs; i++) {
        w[i] = -1;
        ownDepth[i] = 0;
      }
    }
    __syncthreads();
    bool fresh[kPairs], pass[kPairs];
    int nFresh = 0, nPass = 0;
#pragma unroll
    for (int i = 0; i < kPairs; i++) {
      fresh[i] = false;
      if (w[i] >= 0) {
        uint32_t h = ((uint32_t)w[i] * 2654435761u) & (kHash - 1);
        while (hKey[h] != w[i])
          h = (h + 1) & (kHash - 1);
        if (hVal[h] == p0 + i) {
          const uint32_t bit = 1u << (w[i] & 31);
          fresh[i] = !(bits_or(&bits[w[i] >> 5], bit) & bit);
        }
      }
      pass[i] = fresh[i] && (!filter || filter[w[i]]) &&
                (!lowerOnly || orig_id(orig, w[i]) < srcOrig);
      nFresh += fresh[i] ? 1 : 0;
      nPass += pass[i] ? 1 : 0;
    }
    int counts;
    const int packed = block_scan(nFresh | (nPass << 16), sWave, counts);
    int atFresh = (packed & 0xFFFF) - nFresh;
    int atPass = (packed >> 16) - nPass;
#pragma unroll
    for (int i = 0; i < kPairs; i++) {
      if (pass[i]) {
        const int slot = found + atPass++;
        if (slot < k) {
          outV[slot] = orig_id(orig, w[i]);
          outD[slot] = ownDepth[i] + 1;
        }
      }
      if (fresh[i]) {
        st_l2(queue + tail + atFresh, w[i]);
        st_l2(qdist + tail + atFresh, ownDepth[i] + 1);
        atFresh++;
      }
    }
    found += counts >> 16;
    tail += counts & 0xFFFF;
    // advance the cursor past the pairs taken
    if (pairs == total) {
      head += avail;
      cur = 0;
    } else {
      const int last = sLast;
      const int32_t lastIncl = sIncl[last];
      const int32_t lastDeg = lastIncl - (last ? sIncl[last - 1] : 0);
      const int32_t took = lastDeg - (lastIncl - pairs);
      if (took == lastDeg) {
        head += last + 1;
        cur = 0;
      } else {
        cur = (last == 0 ? cur : 0) + took;
        head += last;
      }
    }
    if (stopAtK && found >= k)
      done = true;
    if (stopCount > 0 && found >= stopCount)
      done = true;
    __syncthreads();
  }
  // clear this query's bits: every visited vertex sits in the queue
  for (int i = t; i < tail; i += kBB) {
    const int32_t w = ld_l2(queue + i);
    bits_clear(&bits[w >> 5]);
  }
  __syncthreads();
  return found;
}

// Persistent kernel, one block per slot: block `slot` serves queries slot,
// slot + S, ... with block_bfs.
__global__ __launch_bounds__(kBB) void k_bfs_queries_block(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const uint8_t *__restrict__ filter, const int32_t *__restrict__ sources,
    const int32_t *__restrict__ outIndex, int64_t count, int k, int64_t stride,
    int maxDepth, int lowerOnly, int stopAtK, int64_t filterTotal, BfsSlots sl,
    int64_t slots, int32_t *__restrict__ outV, int32_t *__restrict__ outD,
    int32_t *__restrict__ outCount, int32_t *__restrict__ overflow,
    const int32_t *__restrict__ newId, const int32_t *__restrict__ orig) {
  const int64_t slot = blockIdx.x;
  if (slot >= slots)
    return;
  uint32_t *bits = sl.bits + slot * sl.words;
  int32_t *q = sl.queue + slot * sl.cap;
  int32_t *qd = sl.qdist + slot * sl.cap;
  for (int64_t qi = slot; qi < count; qi += slots) {
    const int32_t src = newId ? newId[sources[qi]] : sources[qi];
    const int64_t o = (outIndex ? (int64_t)outIndex[qi] : qi) * stride;
    const int found = block_bfs(off, adj, filter, src, k, maxDepth, lowerOnly != 0,
                                stopAtK != 0, filterTotal, bits, q, qd, outV + o, outD + o,
                                orig);
    if (threadIdx.x == 0) {
      outCount[outIndex ? outIndex[qi] : qi] = found < k ? found : k;
      if (!stopAtK && found > k)
        atomicOr(overflow, 1);
    }
  }
}

// Persistent kernel: wavefront `slot` serves queries slot, slot + S, ...
__global__ __launch_bounds__(kB) void k_bfs_queries(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const uint8_t *__restrict__ filter, const int32_t *__restrict__ sources,
    const int32_t *__restrict__ outIndex, int64_t count, int k, int64_t stride,
    int maxDepth, int lowerOnly, int stopAtK, int64_t filterTotal, BfsSlots sl,
    int64_t slots, int32_t *__restrict__ outV, int32_t *__restrict__ outD,
    int32_t *__restrict__ outCount, int32_t *__restrict__ overflow,
    const int32_t *__restrict__ newId, const int32_t *__restrict__ orig) {
  __shared__ int32_t sHashK[kWpb][kWaveHash], sHashV[kWpb][kWaveHash];
  const int64_t slot = (int64_t)blockIdx.x * kWpb + threadIdx.x / kWave;
  if (slot >= slots)
    return;
  uint32_t *bits = sl.bits + slot * sl.words;
  int32_t *q = sl.queue + slot * sl.cap;
  int32_t *qd = sl.qdist + slot * sl.cap;
  const int wv = threadIdx.x / kWave;
  for (int64_t qi = slot; qi < count; qi += slots) {
    const int32_t src = newId ? newId[sources[qi]] : sources[qi];
    const int64_t o = (outIndex ? (int64_t)outIndex[qi] : qi) * stride;
    int found = wave_bfs(off, adj, filter, src, k, maxDepth, lowerOnly != 0,
                         stopAtK != 0, filterTotal, bits, q, qd, outV + o, outD + o,
                         sHashK[wv], sHashV[wv], orig);
    if ((threadIdx.x & (kWave - 1)) == 0) {
      outCount[outIndex ? outIndex[qi] : qi] = found < k ? found : k;
      if (!stopAtK && found > k)
        atomicOr(overflow, 1);
    }
  }
}

// ---- multi-source ball search for MIS layers with few candidates (the
// coarse layers, whose balls of radius 2^(i-1) cover much of the graph):
// up to 64 candidates at once, one bit each per vertex, expanded level by
// level with every CU taking part, where one search per block leaves most
// of the chip idle. `seen` holds the sources whose ball holds the vertex,
// `front` the bits that reached it at the current level.

// Level d: every frontier vertex offers its new bits to its neighbours; the
// bits a neighbour did not hold yet form its frontier at level d + 1. The
// frontier entry is cleared once read. act[d] = 1 when any bit moved.
__global__ __launch_bounds__(kB) void k_ms_push(const int32_t *__restrict__ off,
                                                const int32_t *__restrict__ adj, int64_t n,
                                                unsigned long long *__restrict__ seen,
                                                unsigned long long *__restrict__ front,
                                                unsigned long long *__restrict__ next,
                                                int32_t *__restrict__ act) {
  const int64_t u = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (u >= n)
    return;
  const unsigned long long f = front[u];
  if (!f)
    return;
  front[u] = 0ull;
  bool moved = false;
  for (int32_t e = off[u]; e < off[u + 1]; e++) {
    const int32_t w = adj[e];
    const unsigned long long nb = f & ~seen[w];
    if (!nb)
      continue;
    const unsigned long long fresh = nb & ~atomicOr(&seen[w], nb);
    if (fresh) {
      atomicOr(&next[w], fresh);
      moved = true;
    }
  }
  if (moved)
    *act = 1;
}

// Conflict list of source j of the batch (candidate index s0 + j; the
// candidates are ascending, so its lower-id candidates are cand[0 .. s0+j)):
// those whose ball bit j is set, in ascending id, one block per source.
// count = min(found, stride); found > stride sets *overflow.
__global__ __launch_bounds__(kB) void k_ms_lists(const int32_t *__restrict__ cand,
                                                 int64_t s0, int cnt,
                                                 const unsigned long long *__restrict__ seen,
                                                 int64_t stride, int32_t *__restrict__ outV,
                                                 int32_t *__restrict__ outD,
                                                 int32_t *__restrict__ outCount,
                                                 int32_t *__restrict__ overflow) {
  __shared__ int sWave[kWpb];
  __shared__ int sBase;
  const int j = blockIdx.x;
  if (j >= cnt)
    return;
  const int64_t qi = s0 + j;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (threadIdx.x == 0)
    sBase = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < qi; c0 += kB) {
    const int64_t c = c0 + threadIdx.x;
    bool hit = false;
    int32_t v = 0;
    if (c < qi) {
      v = cand[c];
      hit = (seen[v] >> j) & 1ull;
    }
    const uint64_t b = __ballot(hit);
    if (lane == 0)
      sWave[wave] = __popcll(b);
    __syncthreads();
    int before = sBase;
    for (int w = 0; w < wave; w++)
      before += sWave[w];
    if (hit) {
      const int at = before + (int)__popcll(b & (lane ? (~0ull >> (kWave - lane)) : 0ull));
      if (at < stride) {
        outV[qi * stride + at] = v;
        outD[qi * stride + at] = 0;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 0; w < kWpb; w++)
        sBase += sWave[w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    outCount[qi] = sBase < stride ? sBase : (int32_t)stride;
    if (sBase > stride)
      atomicOr(overflow, 1);
  }
}

// ------------------------------------------------------- MIS decisions
// Lexicographically-first MIS over conflict lists of lower-id neighbours:
// IN when every lower conflict is OUT, OUT as soon as one is IN.
constexpr uint8_t kUndecided = 0, kIn = 1, kOut = 2;

__global__ __launch_bounds__(kB) void k_lfmis_lists(
    const int32_t *__restrict__ lists, const int32_t *__restrict__ counts,
    int64_t m, int64_t stride, uint8_t *__restrict__ state,
    int32_t *__restrict__ undecided) {
  int64_t c = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (c >= m || state[c] != kUndecided)
    return;
  bool allOut = true;
  const int32_t cnt = counts[c];
  for (int32_t j = 0; j < cnt; j++) {
    uint8_t s = ((volatile uint8_t *)state)[lists[c * stride + j]];
    if (s == kIn) {
      state[c] = kOut;
      return;
    }
    if (s != kOut)
      allOut = false;
  }
  if (allOut)
    state[c] = kIn;
  else
    atomicAdd(undecided, 1);
}

// Layer 1: the conflict graph is the subgraph itself.
__global__ __launch_bounds__(kB) void k_lfmis_graph(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const uint8_t *__restrict__ active, int64_t n, uint8_t *__restrict__ state,
    int32_t *__restrict__ undecided) {
  int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (v >= n || !active[v] || state[v] != kUndecided)
    return;
  bool allOut = true;
  for (int32_t e = off[v]; e < off[v + 1]; e++) {
    const int32_t u = adj[e];
    if (u >= v)
      continue;
    uint8_t s = ((volatile uint8_t *)state)[u];
    if (s == kIn) {
      state[v] = kOut;
      return;
    }
    if (s != kOut)
      allOut = false;
  }
  if (allOut)
    state[v] = kIn;
  else
    atomicAdd(undecided, 1);
}

// Layer-1 tail order (gvizGRIPEmbedder.c:462-478): a vertex left out of the
// MIS is written when its smallest-id MIS neighbour p scans its adjacency, at
// p's adjacency position: sort key (p, first index of v in adj(p)).
__global__ __launch_bounds__(kB) void k_tail_keys(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const uint8_t *__restrict__ active, const uint8_t *__restrict__ state,
    int64_t n, uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
    int32_t *__restrict__ count) {
  int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (v >= n || !active[v] || state[v] != kOut)
    return;
  int32_t p = INT32_MAX;
  for (int32_t e = off[v]; e < off[v + 1]; e++) {
    const int32_t u = adj[e];
    if (state[u] == kIn && u < p)
      p = u;
  }
  int32_t idx = 0;
  for (int32_t e = off[p]; e < off[p + 1]; e++)
    if (adj[e] == (int32_t)v) {
      idx = e - off[p];
      break;
    }
  const int32_t at = atomicAdd(count, 1);
  keys[at] = ((uint64_t)(uint32_t)p << 32) | (uint32_t)idx;
  vals[at] = (uint32_t)v;
}

// ------------------------------------------------------ refinement round
struct GripRound {
  double sdx, sdy, spx, spy; // sums of normalised disp and positions
  double meanX, meanY, comX, comY;
  double L, I, omega;
  int32_t rotate;
  int32_t pad;
  double maxDisp, sumDisp, sumHeat, sumForce;
};

// refinementPass1Range + calculateSpringForces + updateLocalTemp
// (gvizGRIPEmbedder.c:1071-1099, 892-934, 870-889).
__global__ __launch_bounds__(kB) void k_order_keys(const int32_t *__restrict__ filt,
                                                   int64_t active,
                                                   uint64_t *__restrict__ keys,
                                                   uint32_t *__restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < active) {
    keys[i] = (uint64_t)(uint32_t)filt[i];
    vals[i] = (uint32_t)i;
  }
}

__global__ __launch_bounds__(kB) void k_order_ids(const uint64_t *__restrict__ keys,
                                                  int64_t active,
                                                  int32_t *__restrict__ order) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < active)
    order[i] = (int32_t)keys[i];
}

__global__ __launch_bounds__(kB) void k_refine(
    const int32_t *__restrict__ filt, int64_t active, int64_t layer,
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const double2 *__restrict__ pos, const int32_t *__restrict__ knnV,
    const int32_t *__restrict__ knnD, const int32_t *__restrict__ knnCount,
    int64_t kcap, double2 *__restrict__ disp, double2 *__restrict__ oldDisp,
    double *__restrict__ heat, double *__restrict__ oldCos,
    double *__restrict__ lastF, uint8_t *__restrict__ dispCalc,
    double *__restrict__ partial) {
  __shared__ double sm[4 * 16];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < active;
       i += (int64_t)gridDim.x * kB) {
    const int32_t v = filt ? filt[i] : (int32_t)i;
    const double2 d0 = disp[v];
    oldDisp[v] = d0;
    const double2 pv = pos[v];
    double f0 = 0.0, f1 = 0.0;
    const int32_t kc = knnCount[v];
    const int64_t kb = (int64_t)v * kcap;
    // neighbour positions are gathered kGather at a time ahead of the
    // (in-order) sums, so each thread keeps several loads in flight
    double2 pu[kGather];
    if (layer > 0) {
      for (int32_t j0 = 0; j0 < kc; j0 += kGather) {
        const int m = kc - j0 < kGather ? kc - j0 : kGather;
#pragma unroll
        for (int k = 0; k < kGather; k++)
          if (k < m)
            pu[k] = pos[knnV[kb + j0 + k]];
#pragma unroll
        for (int k = 0; k < kGather; k++) {
          if (k >= m)
            break;
          const double L = (double)knnD[kb + j0 + k] * kEdge;
          if (L < 1e-9)
            continue;
          double dx = pu[k].x - pv.x, dy = pu[k].y - pv.y;
          double s = sqrt(dx * dx + dy * dy) / L - 1.0;
          f0 += s * dx;
          f1 += s * dy;
        }
      }
    } else {
      const double eSq = kEdge * kEdge;
      const int32_t e1 = off[v + 1];
      for (int32_t e0 = off[v]; e0 < e1; e0 += kGather) {
        const int m = e1 - e0 < kGather ? e1 - e0 : kGather;
#pragma unroll
        for (int k = 0; k < kGather; k++)
          if (k < m)
            pu[k] = pos[adj[e0 + k]];
#pragma unroll
        for (int k = 0; k < kGather; k++) {
          if (k >= m)
            break;
          double dx = pu[k].x - pv.x, dy = pu[k].y - pv.y;
          double dsq = dx * dx + dy * dy;
          if (dsq < 1e-18)
            continue;
          double s = dsq / eSq;
          f0 += s * dx;
          f1 += s * dy;
        }
      }
      for (int32_t j0 = 0; j0 < kc; j0 += kGather) {
        const int m = kc - j0 < kGather ? kc - j0 : kGather;
#pragma unroll
        for (int k = 0; k < kGather; k++)
          if (k < m)
            pu[k] = pos[knnV[kb + j0 + k]];
#pragma unroll
        for (int k = 0; k < kGather; k++) {
          if (k >= m)
            break;
          double dx = pv.x - pu[k].x, dy = pv.y - pu[k].y;
          double dsq = dx * dx + dy * dy;
          if (dsq < 1e-18)
            continue;
          double s = kFrScale * eSq / dsq;
          f0 += s * dx;
          f1 += s * dy;
        }
      }
    }
    double h = heat[v];
    if (!dispCalc[v]) {
      dispCalc[v] = 1;
      h = kEdge / 6.0;
    } else {
      double nrm = sqrt(f0 * f0 + f1 * f1);
      double oldNrm = sqrt(d0.x * d0.x + d0.y * d0.y);
      if (!(fabs(nrm) < kEps || fabs(oldNrm) < kEps)) {
        double cs = (f0 * d0.x + f1 * d0.y) / (nrm * oldNrm);
        if (cs > 0 && oldCos[v] > 0)
          h *= (1 + cs * kGripR * kGripS);
        else
          h *= (1 + cs * kGripR);
        oldCos[v] = cs;
      }
    }
    h = fmin(h, kEdge * 10);
    h = fmax(h, kEdge * 1e-4);
    heat[v] = h;
    double nrm = sqrt(f0 * f0 + f1 * f1);
    lastF[v] = nrm;
    if (nrm > kEps) {
      double s = h / nrm;
      f0 *= s;
      f1 *= s;
    }
    disp[v] = {f0, f1};
    acc[0] += f0;
    acc[1] += f1;
    acc[2] += pv.x;
    acc[3] += pv.y;
  }
  const bool isMax[4] = {false, false, false, false};
  block_reduce<4>(acc, isMax, sm);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++)
      partial[4 * blockIdx.x + k] = acc[k];
}

__global__ void k_round_means(const double *__restrict__ partial, int blocks,
                              int64_t active, GripRound *__restrict__ r) {
  __shared__ double sm[4 * 16];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < blocks; b += blockDim.x)
    for (int k = 0; k < 4; k++)
      v[k] += partial[4 * b + k];
  const bool isMax[4] = {false, false, false, false};
  block_reduce<4>(v, isMax, sm);
  if (threadIdx.x == 0) {
    // removeNetTranslation (gvizGRIPEmbedder.c:1106-1121): scale by 1/count
    double inv = 1.0 / (double)active;
    r->meanX = v[0] * inv;
    r->meanY = v[1] * inv;
    // removeNetRotation's centre (:1205-1214): sum / count
    r->comX = v[2] / (double)active;
    r->comY = v[3] / (double)active;
  }
}

// Translation removal and the rotation sums of the 2-D branch of
// removeNetRotation (gvizGRIPEmbedder.c:1216-1245).
__global__ __launch_bounds__(kB) void k_rotation_sums(
    const int32_t *__restrict__ filt, int64_t active,
    const double2 *__restrict__ pos, double2 *__restrict__ disp,
    const GripRound *__restrict__ r, double *__restrict__ partial) {
  __shared__ double sm[2 * 16];
  double acc[2] = {0.0, 0.0};
  const double mx = r->meanX, my = r->meanY, cx = r->comX, cy = r->comY;
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < active;
       i += (int64_t)gridDim.x * kB) {
    const int32_t v = filt ? filt[i] : (int32_t)i;
    double2 d = disp[v];
    d.x += -1.0 * mx;
    d.y += -1.0 * my;
    disp[v] = d;
    const double2 p = pos[v];
    double r0 = p.x - cx, r1 = p.y - cy, r2v = 0.0;
    acc[0] += r0 * d.y - r1 * d.x;
    double r2 = r0 * r0 + r1 * r1 + r2v * r2v;
    acc[1] += r2 - r2v * r2v;
  }
  const bool isMax[2] = {false, false};
  block_reduce<2>(acc, isMax, sm);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = acc[0];
    partial[2 * blockIdx.x + 1] = acc[1];
  }
}

__global__ void k_round_omega(const double *__restrict__ partial, int blocks,
                              int64_t active, GripRound *__restrict__ r) {
  __shared__ double sm[2 * 16];
  double v[2] = {0.0, 0.0};
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    v[0] += partial[2 * b];
    v[1] += partial[2 * b + 1];
  }
  const bool isMax[2] = {false, false};
  block_reduce<2>(v, isMax, sm);
  if (threadIdx.x == 0) {
    r->L = v[0];
    r->I = v[1];
    r->rotate = active >= 3 && !(fabs(v[1]) < kEps);
    r->omega = r->rotate ? v[0] / v[1] : 0.0;
  }
}

// Rotation removal, position update and round stats
// (gvizGRIPEmbedder.c:1246-1260, 1299-1318).
__global__ __launch_bounds__(kB) void k_round_apply(
    const int32_t *__restrict__ filt, int64_t active, double2 *__restrict__ pos,
    double2 *__restrict__ disp, const double *__restrict__ heat,
    const double *__restrict__ lastF, const GripRound *__restrict__ r,
    double *__restrict__ partial) {
  __shared__ double sm[4 * 16];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const bool rot = r->rotate != 0;
  const double w = r->omega, cx = r->comX, cy = r->comY;
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < active;
       i += (int64_t)gridDim.x * kB) {
    const int32_t v = filt ? filt[i] : (int32_t)i;
    double2 d = disp[v];
    double2 p = pos[v];
    if (rot) {
      double r0 = p.x - cx, r1 = p.y - cy, r2v = 0.0;
      d.x -= 0.0 * r2v - w * r1;
      d.y -= w * r0 - 0.0 * r2v;
      disp[v] = d;
    }
    double nrm = sqrt(d.x * d.x + d.y * d.y);
    acc[0] = nrm > acc[0] ? nrm : acc[0];
    acc[1] += nrm;
    acc[2] += heat[v];
    acc[3] += lastF[v];
    p.x += d.x;
    p.y += d.y;
    pos[v] = p;
  }
  const bool isMax[4] = {true, false, false, false};
  block_reduce<4>(acc, isMax, sm);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++)
      partial[4 * blockIdx.x + k] = acc[k];
}

__global__ void k_round_stats(const double *__restrict__ partial, int blocks,
                              int64_t active, GripRound *__restrict__ r) {
  __shared__ double sm[4 * 16];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    v[0] = fmax(v[0], partial[4 * b]);
    for (int k = 1; k < 4; k++)
      v[k] += partial[4 * b + k];
  }
  const bool isMax[4] = {true, false, false, false};
  block_reduce<4>(v, isMax, sm);
  if (threadIdx.x == 0) {
    r->maxDisp = v[0];
    r->sumDisp = v[1];
    r->sumHeat = v[2];
    r->sumForce = v[3];
  }
}

// ------------------------------------------------ compact 2-D rounds
// A layer's active vertices renumbered 0..active-1 in the Morton order of
// their positions (built at the layer's first round). Vertex ids follow the
// order the graph came in, which for most inputs is unrelated to where the
// vertices lie, so the neighbour gathers of one wavefront touch a cache line
// each; in layout order they share lines. The round kernels run unchanged on
// the compact copy (filt == nullptr: vertex i is compact index i).

// Block partials of the bounding box of the active vertices' positions:
// (-min x, -min y, max x, max y), all reduced as maxima.
__global__ __launch_bounds__(kB) void k_active_bbox(const int32_t *__restrict__ filt,
                                                    int64_t active,
                                                    const double2 *__restrict__ pos,
                                                    double *__restrict__ partial) {
  __shared__ double sm[4 * 16];
  double acc[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < active;
       i += (int64_t)gridDim.x * kB) {
    const double2 p = pos[filt[i]];
    acc[0] = fmax(acc[0], -p.x);
    acc[1] = fmax(acc[1], -p.y);
    acc[2] = fmax(acc[2], p.x);
    acc[3] = fmax(acc[3], p.y);
  }
  const bool isMax[4] = {true, true, true, true};
  block_reduce<4>(acc, isMax, sm);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++)
      partial[4 * blockIdx.x + k] = acc[k];
}

__global__ void k_bbox_finish(const double *__restrict__ partial, int blocks,
                              double *__restrict__ box) {
  __shared__ double sm[4 * 16];
  double v[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < blocks; b += blockDim.x)
    for (int k = 0; k < 4; k++)
      v[k] = fmax(v[k], partial[4 * b + k]);
  const bool isMax[4] = {true, true, true, true};
  block_reduce<4>(v, isMax, sm);
  if (threadIdx.x == 0) {
    box[0] = -v[0];
    box[1] = -v[1];
    box[2] = v[2];
    box[3] = v[3];
  }
}

__device__ __forceinline__ uint32_t spread16(uint32_t x) {
  x &= 0xFFFFu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

// 32-bit Morton key of each active vertex's position in the box (16 bits per
// axis); value = its index in filt.
__global__ __launch_bounds__(kB) void k_layout_keys(const int32_t *__restrict__ filt,
                                                    int64_t active,
                                                    const double2 *__restrict__ pos,
                                                    const double *__restrict__ box,
                                                    uint64_t *__restrict__ keys,
                                                    uint32_t *__restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= active)
    return;
  const double2 p = pos[filt[i]];
  const double wx = box[2] - box[0], wy = box[3] - box[1];
  const double fx = wx > 0.0 ? (p.x - box[0]) / wx : 0.0;
  const double fy = wy > 0.0 ? (p.y - box[1]) / wy : 0.0;
  const uint32_t qx = (uint32_t)fmin(fmax(fx * 65535.0, 0.0), 65535.0);
  const uint32_t qy = (uint32_t)fmin(fmax(fy * 65535.0, 0.0), 65535.0);
  keys[i] = (uint64_t)(spread16(qx) | (spread16(qy) << 1));
  vals[i] = (uint32_t)i;
}

// Compact copies of the per-vertex round state, in sorted key order.
__global__ __launch_bounds__(kB) void k_compact_gather(
    const uint32_t *__restrict__ sortedVals, const int32_t *__restrict__ filt,
    int64_t active, int32_t *__restrict__ cOrder, int32_t *__restrict__ cMap,
    const double2 *__restrict__ pos, const double2 *__restrict__ disp,
    const double2 *__restrict__ oldDisp, const double *__restrict__ heat,
    const double *__restrict__ oldCos, const double *__restrict__ lastF,
    const uint8_t *__restrict__ dispCalc, double2 *__restrict__ cPos,
    double2 *__restrict__ cDisp, double2 *__restrict__ cOldDisp,
    double *__restrict__ cHeat, double *__restrict__ cOldCos, double *__restrict__ cLastF,
    uint8_t *__restrict__ cDispCalc) {
  const int64_t c = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (c >= active)
    return;
  const int32_t v = filt[sortedVals[c]];
  cOrder[c] = v;
  cMap[v] = (int32_t)c;
  cPos[c] = pos[v];
  cDisp[c] = disp[v];
  cOldDisp[c] = oldDisp[v];
  cHeat[c] = heat[v];
  cOldCos[c] = oldCos[v];
  cLastF[c] = lastF[v];
  cDispCalc[c] = dispCalc[v];
}

// The compact state written back to the id-ordered arrays.
__global__ __launch_bounds__(kB) void k_compact_scatter(
    const int32_t *__restrict__ cOrder, int64_t active, const double2 *__restrict__ cPos,
    const double2 *__restrict__ cDisp, const double2 *__restrict__ cOldDisp,
    const double *__restrict__ cHeat, const double *__restrict__ cOldCos,
    const double *__restrict__ cLastF, const uint8_t *__restrict__ cDispCalc,
    double2 *__restrict__ pos, double2 *__restrict__ disp, double2 *__restrict__ oldDisp,
    double *__restrict__ heat, double *__restrict__ oldCos, double *__restrict__ lastF,
    uint8_t *__restrict__ dispCalc) {
  const int64_t c = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (c >= active)
    return;
  const int32_t v = cOrder[c];
  pos[v] = cPos[c];
  disp[v] = cDisp[c];
  oldDisp[v] = cOldDisp[c];
  heat[v] = cHeat[c];
  oldCos[v] = cOldCos[c];
  lastF[v] = cLastF[c];
  dispCalc[v] = cDispCalc[c];
}

// kNN lists in compact indices. A target outside the layer (never the case
// for lists built over the layer's visible vertices) sets *bad and is
// replaced by the vertex itself so every later access stays in bounds.
__global__ __launch_bounds__(kB) void k_compact_knn(
    const int32_t *__restrict__ cOrder, int64_t active, const int32_t *__restrict__ cMap,
    int64_t n, const int32_t *__restrict__ knnV, const int32_t *__restrict__ knnD,
    const int32_t *__restrict__ knnCount, int64_t kcap, int32_t *__restrict__ cKnnV,
    int32_t *__restrict__ cKnnD, int32_t *__restrict__ cKnnCount,
    int32_t *__restrict__ bad) {
  const int64_t c = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (c >= active)
    return;
  const int32_t v = cOrder[c];
  int32_t cnt = knnCount[v];
  cnt = cnt < 0 ? 0 : (cnt > kcap ? (int32_t)kcap : cnt);
  cKnnCount[c] = cnt;
  for (int32_t j = 0; j < cnt; j++) {
    const int32_t u = knnV[(int64_t)v * kcap + j];
    int32_t m = (u >= 0 && u < n) ? cMap[u] : -1;
    if (m < 0 || m >= active) {
      *bad = 1;
      m = (int32_t)c;
    }
    cKnnV[c * kcap + j] = m;
    cKnnD[c * kcap + j] = knnD[(int64_t)v * kcap + j];
  }
}

// Layer-0 adjacency rows in compact order (row starts cOff from the host),
// neighbours in compact indices; a neighbour outside the layer sets *bad.
__global__ __launch_bounds__(kB) void k_compact_adj(
    const int32_t *__restrict__ cOrder, int64_t active, const int32_t *__restrict__ cMap,
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const int32_t *__restrict__ cOff, int32_t *__restrict__ cAdj,
    int32_t *__restrict__ bad) {
  const int64_t c = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (c >= active)
    return;
  const int32_t v = cOrder[c];
  const int32_t b = off[v], e = off[v + 1], o = cOff[c];
  for (int32_t k = b; k < e; k++) {
    int32_t m = cMap[adj[k]];
    if (m < 0 || m >= active) {
      *bad = 1;
      m = (int32_t)c;
    }
    cAdj[o + (k - b)] = m;
  }
}

// ------------------------------------ refinement in three and four dimensions
// The same round as above on positions of D = 3 or 4 coordinates, stored
// D-strided per vertex. Sums of squares, dot products and the pairwise
// forces run over the coordinates left to right as gvizVecDot and the
// gvizVecAcc*Force cases for n = 3, 4 do (include/core/gvizVec.h:42-57,
// 138-290).
struct GripRoundN {
  double mean[4], com[4];
  double omega[6]; // D = 3: the solved rotation; D = 4: one per coordinate plane
  int32_t rotate[6];
  double maxDisp, sumDisp, sumHeat, sumForce;
};

template <int D> struct PtN {
  double c[D];
};

template <int D>
__device__ __forceinline__ PtN<D> ld_pt(const double *__restrict__ a, int64_t v) {
  PtN<D> p;
#pragma unroll
  for (int k = 0; k < D; k++)
    p.c[k] = a[v * D + k];
  return p;
}

template <int D>
__device__ __forceinline__ void st_pt(double *__restrict__ a, int64_t v, const PtN<D> &p) {
#pragma unroll
  for (int k = 0; k < D; k++)
    a[v * D + k] = p.c[k];
}

template <int D> __device__ __forceinline__ double sq_sum(const double (&d)[D]) {
  double s = d[0] * d[0];
#pragma unroll
  for (int k = 1; k < D; k++)
    s += d[k] * d[k];
  return s;
}

// refinementPass1Range, calculateSpringForces and updateLocalTemp
// (gvizGRIPEmbedder.c:1071-1099, 892-934, 870-889) for D coordinates.
template <int D>
__global__ __launch_bounds__(kB) void k_refine_nd(
    const int32_t *__restrict__ order, int64_t active, int64_t layer,
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const double *__restrict__ pos, const int32_t *__restrict__ knnV,
    const int32_t *__restrict__ knnD, const int32_t *__restrict__ knnCount,
    int64_t kcap, double *__restrict__ disp, double *__restrict__ oldDisp,
    double *__restrict__ heat, double *__restrict__ oldCos,
    double *__restrict__ lastF, uint8_t *__restrict__ dispCalc,
    double *__restrict__ partial) {
  __shared__ double sm[2 * D * 16];
  double acc[2 * D];
#pragma unroll
  for (int k = 0; k < 2 * D; k++)
    acc[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < active;
       i += (int64_t)gridDim.x * kB) {
    const int32_t v = order[i];
    const PtN<D> d0 = ld_pt<D>(disp, v);
    st_pt<D>(oldDisp, v, d0);
    const PtN<D> pv = ld_pt<D>(pos, v);
    double f[D], dd[D];
#pragma unroll
    for (int k = 0; k < D; k++)
      f[k] = 0.0;
    const int32_t kc = knnCount[v];
    const int64_t kb = (int64_t)v * kcap;
    // neighbour positions are gathered kGather at a time ahead of the
    // (in-order) sums, as in the 2-D kernel
    PtN<D> pu[kGather];
    if (layer > 0) {
      for (int32_t j0 = 0; j0 < kc; j0 += kGather) {
        const int m = kc - j0 < kGather ? kc - j0 : kGather;
#pragma unroll
        for (int k = 0; k < kGather; k++)
          if (k < m)
            pu[k] = ld_pt<D>(pos, knnV[kb + j0 + k]);
#pragma unroll
        for (int k = 0; k < kGather; k++) {
          if (k >= m)
            break;
          const double L = (double)knnD[kb + j0 + k] * kEdge;
          if (L < 1e-9)
            continue;
#pragma unroll
          for (int c = 0; c < D; c++)
            dd[c] = pu[k].c[c] - pv.c[c];
          const double s = sqrt(sq_sum<D>(dd)) / L - 1.0;
#pragma unroll
          for (int c = 0; c < D; c++)
            f[c] += s * dd[c];
        }
      }
    } else {
      const double eSq = kEdge * kEdge;
      const int32_t e1 = off[v + 1];
      for (int32_t e0 = off[v]; e0 < e1; e0 += kGather) {
        const int m = e1 - e0 < kGather ? e1 - e0 : kGather;
#pragma unroll
        for (int k = 0; k < kGather; k++)
          if (k < m)
            pu[k] = ld_pt<D>(pos, adj[e0 + k]);
#pragma unroll
        for (int k = 0; k < kGather; k++) {
          if (k >= m)
            break;
#pragma unroll
          for (int c = 0; c < D; c++)
            dd[c] = pu[k].c[c] - pv.c[c];
          const double dsq = sq_sum<D>(dd);
          if (dsq < 1e-18)
            continue;
          const double s = dsq / eSq;
#pragma unroll
          for (int c = 0; c < D; c++)
            f[c] += s * dd[c];
        }
      }
      for (int32_t j0 = 0; j0 < kc; j0 += kGather) {
        const int m = kc - j0 < kGather ? kc - j0 : kGather;
#pragma unroll
        for (int k = 0; k < kGather; k++)
          if (k < m)
            pu[k] = ld_pt<D>(pos, knnV[kb + j0 + k]);
#pragma unroll
        for (int k = 0; k < kGather; k++) {
          if (k >= m)
            break;
#pragma unroll
          for (int c = 0; c < D; c++)
            dd[c] = pv.c[c] - pu[k].c[c];
          const double dsq = sq_sum<D>(dd);
          if (dsq < 1e-18)
            continue;
          const double s = kFrScale * eSq / dsq;
#pragma unroll
          for (int c = 0; c < D; c++)
            f[c] += s * dd[c];
        }
      }
    }
    double h = heat[v];
    if (!dispCalc[v]) {
      dispCalc[v] = 1;
      h = kEdge / 6.0;
    } else {
      const double nrm = sqrt(sq_sum<D>(f));
      const double oldNrm = sqrt(sq_sum<D>(d0.c));
      if (!(fabs(nrm) < kEps || fabs(oldNrm) < kEps)) {
        double dot = f[0] * d0.c[0];
#pragma unroll
        for (int k = 1; k < D; k++)
          dot += f[k] * d0.c[k];
        const double cs = dot / (nrm * oldNrm);
        if (cs > 0 && oldCos[v] > 0)
          h *= (1 + cs * kGripR * kGripS);
        else
          h *= (1 + cs * kGripR);
        oldCos[v] = cs;
      }
    }
    h = fmin(h, kEdge * 10);
    h = fmax(h, kEdge * 1e-4);
    heat[v] = h;
    const double nrm = sqrt(sq_sum<D>(f));
    lastF[v] = nrm;
    if (nrm > kEps) {
      const double s = h / nrm;
#pragma unroll
      for (int k = 0; k < D; k++)
        f[k] *= s;
    }
    PtN<D> out;
#pragma unroll
    for (int k = 0; k < D; k++) {
      out.c[k] = f[k];
      acc[k] += f[k];
      acc[D + k] += pv.c[k];
    }
    st_pt<D>(disp, v, out);
  }
  bool isMax[2 * D];
#pragma unroll
  for (int k = 0; k < 2 * D; k++)
    isMax[k] = false;
  block_reduce<2 * D>(acc, isMax, sm);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 2 * D; k++)
      partial[2 * D * blockIdx.x + k] = acc[k];
}

// removeNetTranslation's mean (gvizGRIPEmbedder.c:1106-1121) and the centre
// of removeNetRotation (:1198-1206).
template <int D>
__global__ void k_round_means_nd(const double *__restrict__ partial, int blocks,
                                 int64_t active, GripRoundN *__restrict__ r) {
  __shared__ double sm[2 * D * 16];
  double v[2 * D];
  bool isMax[2 * D];
#pragma unroll
  for (int k = 0; k < 2 * D; k++) {
    v[k] = 0.0;
    isMax[k] = false;
  }
  for (int b = threadIdx.x; b < blocks; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < 2 * D; k++)
      v[k] += partial[2 * D * b + k];
  block_reduce<2 * D>(v, isMax, sm);
  if (threadIdx.x == 0) {
    const double inv = 1.0 / (double)active;
    for (int k = 0; k < D; k++) {
      r->mean[k] = v[k] * inv;
      r->com[k] = v[D + k] / (double)active;
    }
  }
}

// Translation removal and the sums of the 3-D branch of removeNetRotation
// (gvizGRIPEmbedder.c:1216-1235): the angular momentum L and the inertia
// tensor, whose off-diagonal entries are symmetric term by term.
__global__ __launch_bounds__(kB) void k_rotation3_sums(
    const int32_t *__restrict__ order, int64_t active, const double *__restrict__ pos,
    double *__restrict__ disp, const GripRoundN *__restrict__ r,
    double *__restrict__ partial) {
  __shared__ double sm[9 * 16];
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double m0 = r->mean[0], m1 = r->mean[1], m2 = r->mean[2];
  const double c0 = r->com[0], c1 = r->com[1], c2 = r->com[2];
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < active;
       i += (int64_t)gridDim.x * kB) {
    const int32_t v = order[i];
    PtN<3> d = ld_pt<3>(disp, v);
    d.c[0] += -1.0 * m0;
    d.c[1] += -1.0 * m1;
    d.c[2] += -1.0 * m2;
    st_pt<3>(disp, v, d);
    const PtN<3> p = ld_pt<3>(pos, v);
    const double r0 = p.c[0] - c0, r1 = p.c[1] - c1, r2 = p.c[2] - c2;
    acc[0] += r1 * d.c[2] - r2 * d.c[1];
    acc[1] += r2 * d.c[0] - r0 * d.c[2];
    acc[2] += r0 * d.c[1] - r1 * d.c[0];
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    acc[3] += rr - r0 * r0;   // I00
    acc[4] += 0.0 - r0 * r1;  // I01 = I10
    acc[5] += 0.0 - r0 * r2;  // I02 = I20
    acc[6] += rr - r1 * r1;   // I11
    acc[7] += 0.0 - r1 * r2;  // I12 = I21
    acc[8] += rr - r2 * r2;   // I22
  }
  const bool isMax[9] = {false, false, false, false, false, false, false, false, false};
  block_reduce<9>(acc, isMax, sm);
  if (threadIdx.x == 0)
    for (int k = 0; k < 9; k++)
      partial[9 * blockIdx.x + k] = acc[k];
}

// solve3x3 (gvizGRIPEmbedder.c:1122-1152): Gaussian elimination with the
// row of largest magnitude as pivot, rows swapped through an index table;
// -1 when a pivot is below gvizNumericEpsilon.
__device__ int solve3(double (&a)[3][3], const double (&b)[3], double (&x)[3]) {
  int idx[3] = {0, 1, 2};
  double rhs[3] = {b[0], b[1], b[2]};
  for (int col = 0; col < 3; col++) {
    int piv = col;
    for (int row = col + 1; row < 3; row++)
      if (fabs(a[idx[row]][col]) > fabs(a[idx[piv]][col]))
        piv = row;
    if (fabs(a[idx[piv]][col]) < kEps)
      return -1;
    const int t = idx[col];
    idx[col] = idx[piv];
    idx[piv] = t;
    for (int row = col + 1; row < 3; row++) {
      const double f = a[idx[row]][col] / a[idx[col]][col];
      for (int j = col; j < 3; j++)
        a[idx[row]][j] -= f * a[idx[col]][j];
      rhs[idx[row]] -= f * rhs[idx[col]];
    }
  }
  for (int col = 2; col >= 0; col--) {
    double s = rhs[idx[col]];
    for (int j = col + 1; j < 3; j++)
      s -= a[idx[col]][j] * x[j];
    x[col] = s / a[idx[col]][col];
  }
  return 0;
}

__global__ void k_rotation3_solve(const double *__restrict__ partial, int blocks,
                                  int64_t active, GripRoundN *__restrict__ r) {
  __shared__ double sm[9 * 16];
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < blocks; b += blockDim.x)
    for (int k = 0; k < 9; k++)
      v[k] += partial[9 * b + k];
  const bool isMax[9] = {false, false, false, false, false, false, false, false, false};
  block_reduce<9>(v, isMax, sm);
  if (threadIdx.x == 0) {
    double a[3][3] = {{v[3], v[4], v[5]}, {v[4], v[6], v[7]}, {v[5], v[7], v[8]}};
    const double L[3] = {v[0], v[1], v[2]};
    double w[3] = {0.0, 0.0, 0.0};
    const int ok = active >= 3 && solve3(a, L, w) == 0;
    r->rotate[0] = ok;
    for (int k = 0; k < 3; k++)
      r->omega[k] = ok ? w[k] : 0.0;
  }
}

// The six coordinate planes of the 4-D branch, removed one after another
// (removeNetRotationPlane, gvizGRIPEmbedder.c:1156-1185, called from
// :1208-1215).
__constant__ int kPlaneA[6] = {0, 0, 0, 1, 1, 2};
__constant__ int kPlaneB[6] = {1, 2, 3, 2, 3, 3};

// The translation removal and the removals of planes [0, count) applied to
// one vertex's displacement, in the reference's order.
__device__ __forceinline__ void apply_planes(PtN<4> &d, const PtN<4> &p,
                                             const GripRoundN *__restrict__ r, int count) {
#pragma unroll
  for (int k = 0; k < 4; k++)
    d.c[k] += -1.0 * r->mean[k];
  for (int q = 0; q < count; q++) {
    if (!r->rotate[q])
      continue;
    const int qa = kPlaneA[q], qb = kPlaneB[q];
    const double ra = p.c[qa] - r->com[qa], rb = p.c[qb] - r->com[qb];
    d.c[qa] += r->omega[q] * rb;
    d.c[qb] -= r->omega[q] * ra;
  }
}

// Plane `plane`'s sums over the displacements the earlier removals leave.
// Those are recomputed per vertex from the refinement's output (the same
// operations in the same order), so no pass but the last writes them.
__global__ __launch_bounds__(kB) void k_plane_sums(
    const int32_t *__restrict__ order, int64_t active, const double *__restrict__ pos,
    const double *__restrict__ disp, const GripRoundN *__restrict__ r, int plane,
    double *__restrict__ partial) {
  __shared__ double sm[2 * 16];
  double acc[2] = {0.0, 0.0};
  const int a = kPlaneA[plane], b = kPlaneB[plane];
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < active;
       i += (int64_t)gridDim.x * kB) {
    const int32_t v = order[i];
    PtN<4> d = ld_pt<4>(disp, v);
    const PtN<4> p = ld_pt<4>(pos, v);
    apply_planes(d, p, r, plane);
    const double ra = p.c[a] - r->com[a], rb = p.c[b] - r->com[b];
    acc[0] += ra * d.c[b] - rb * d.c[a];
    acc[1] += ra * ra + rb * rb;
  }
  const bool isMax[2] = {false, false};
  block_reduce<2>(acc, isMax, sm);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = acc[0];
    partial[2 * blockIdx.x + 1] = acc[1];
  }
}

__global__ void k_plane_omega(const double *__restrict__ partial, int blocks,
                              int64_t active, GripRoundN *__restrict__ r, int plane) {
  __shared__ double sm[2 * 16];
  double v[2] = {0.0, 0.0};
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    v[0] += partial[2 * b];
    v[1] += partial[2 * b + 1];
  }
  const bool isMax[2] = {false, false};
  block_reduce<2>(v, isMax, sm);
  if (threadIdx.x == 0) {
    const int ok = active >= 3 && !(v[1] < kEps);
    r->rotate[plane] = ok;
    r->omega[plane] = ok ? v[0] / v[1] : 0.0;
  }
}

// The last rotation fix-up (the solved 3-D rotation, :1246-1260, or for
// 4-D the translation and all six planes on the refinement's output), the
// position update and the round statistics (:1299-1318).
template <int D>
__global__ __launch_bounds__(kB) void k_round_apply_nd(
    const int32_t *__restrict__ order, int64_t active, double *__restrict__ pos,
    double *__restrict__ disp, const double *__restrict__ heat,
    const double *__restrict__ lastF, const GripRoundN *__restrict__ r,
    double *__restrict__ partial) {
  __shared__ double sm[4 * 16];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const bool rot = r->rotate[0] != 0;
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < active;
       i += (int64_t)gridDim.x * kB) {
    const int32_t v = order[i];
    PtN<D> d = ld_pt<D>(disp, v);
    PtN<D> p = ld_pt<D>(pos, v);
    if constexpr (D == 3) {
      if (rot) {
        const double w0 = r->omega[0], w1 = r->omega[1], w2 = r->omega[2];
        const double r0 = p.c[0] - r->com[0], r1 = p.c[1] - r->com[1],
                     r2 = p.c[2] - r->com[2];
        d.c[0] -= w1 * r2 - w2 * r1;
        d.c[1] -= w2 * r0 - w0 * r2;
        d.c[2] -= w0 * r1 - w1 * r0;
        st_pt<D>(disp, v, d);
      }
    } else {
      apply_planes(d, p, r, 6); // translation and all six planes
      st_pt<D>(disp, v, d);
    }
    const double nrm = sqrt(sq_sum<D>(d.c));
    acc[0] = nrm > acc[0] ? nrm : acc[0];
    acc[1] += nrm;
    acc[2] += heat[v];
    acc[3] += lastF[v];
#pragma unroll
    for (int k = 0; k < D; k++)
      p.c[k] += d.c[k];
    st_pt<D>(pos, v, p);
  }
  const bool isMax[4] = {true, false, false, false};
  block_reduce<4>(acc, isMax, sm);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++)
      partial[4 * blockIdx.x + k] = acc[k];
}

__global__ void k_round_stats_nd(const double *__restrict__ partial, int blocks,
                                 GripRoundN *__restrict__ r) {
  __shared__ double sm[4 * 16];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    v[0] = fmax(v[0], partial[4 * b]);
    for (int k = 1; k < 4; k++)
      v[k] += partial[4 * b + k];
  }
  const bool isMax[4] = {true, false, false, false};
  block_reduce<4>(v, isMax, sm);
  if (threadIdx.x == 0) {
    r->maxDisp = v[0];
    r->sumDisp = v[1];
    r->sumHeat = v[2];
    r->sumForce = v[3];
  }
}

// clearDecorators (gvizGRIPEmbedder.c:936-952) for D coordinates.
template <int D>
__global__ __launch_bounds__(kB) void k_clear_nd(const int32_t *__restrict__ filt,
                                                 int64_t active, int64_t n, double *disp,
                                                 double *oldDisp, double *heat,
                                                 double *oldCos, int32_t *knnCount,
                                                 uint8_t *dispCalc) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n)
    dispCalc[i] = 0;
  if (i < active) {
    const int32_t v = filt[i];
    heat[v] = 0.0;
    oldCos[v] = 0.0;
    knnCount[v] = 0;
    PtN<D> z;
#pragma unroll
    for (int k = 0; k < D; k++)
      z.c[k] = 0.0;
    st_pt<D>(disp, v, z);
    st_pt<D>(oldDisp, v, z);
  }
}

// gripBarycenter (gvizGRIPEmbedder.c:21-27) for D coordinates.
template <int D>
__global__ __launch_bounds__(kB) void k_place_nd(
    const int32_t *__restrict__ filt, int64_t begin, int64_t end,
    const int32_t *__restrict__ hitV, const int32_t *__restrict__ hitCount,
    int64_t stride, double *__restrict__ pos, int32_t *__restrict__ empty) {
  const int64_t i = begin + (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= end)
    return;
  const int64_t q = i - begin;
  const int32_t c = hitCount[q];
  if (c <= 0) {
    atomicOr(empty, 1);
    return;
  }
  double o[D];
#pragma unroll
  for (int k = 0; k < D; k++)
    o[k] = 0.0;
  for (int32_t j = 0; j < c; j++) {
    const PtN<D> p = ld_pt<D>(pos, hitV[q * stride + j]);
#pragma unroll
    for (int k = 0; k < D; k++)
      o[k] += 1.0 * p.c[k];
  }
  const double s = 1.0 / (double)c;
  PtN<D> out;
#pragma unroll
  for (int k = 0; k < D; k++)
    out.c[k] = o[k] * s;
  st_pt<D>(pos, filt[i], out);
}

// clearDecorators (gvizGRIPEmbedder.c:936-952)
__global__ __launch_bounds__(kB) void k_clear(const int32_t *__restrict__ filt,
                                              int64_t active, int64_t n,
                                              double2 *disp, double2 *oldDisp,
                                              double *heat, double *oldCos,
                                              int32_t *knnCount, uint8_t *dispCalc) {
  int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n)
    dispCalc[i] = 0;
  if (i < active) {
    const int32_t v = filt[i];
    heat[v] = 0.0;
    oldCos[v] = 0.0;
    knnCount[v] = 0;
    disp[v] = {0.0, 0.0};
    oldDisp[v] = {0.0, 0.0};
  }
}

// gripBarycenter (gvizGRIPEmbedder.c:21-27) of each placed vertex's hits.
__global__ __launch_bounds__(kB) void k_place(
    const int32_t *__restrict__ filt, int64_t begin, int64_t end,
    const int32_t *__restrict__ hitV, const int32_t *__restrict__ hitCount,
    int64_t stride, double2 *__restrict__ pos, int32_t *__restrict__ empty) {
  int64_t i = begin + (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= end)
    return;
  const int64_t q = i - begin;
  const int32_t c = hitCount[q];
  if (c <= 0) {
    atomicOr(empty, 1);
    return;
  }
  double o0 = 0.0, o1 = 0.0;
  for (int32_t j = 0; j < c; j++) {
    const double2 p = pos[hitV[q * stride + j]];
    o0 += 1.0 * p.x;
    o1 += 1.0 * p.y;
  }
  const double s = 1.0 / (double)c;
  o0 *= s;
  o1 *= s;
  pos[filt[i]] = {o0, o1};
}

// Batched search, per target: the k best (depth, seed order) among the seeds
// processed before every target was full (gvizKNearest.c:62-89, 168-210).
__global__ __launch_bounds__(kB) void k_batch_kth(const int32_t *__restrict__ table,
                                                  int64_t nseeds, int64_t ntargets,
                                                  int k, int32_t *__restrict__ stopSeed,
                                                  int32_t *__restrict__ notFull) {
  int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (t >= ntargets)
    return;
  int reached = 0;
  for (int64_t s = 0; s < nseeds; s++) {
    if (table[s * ntargets + t] >= 0 && ++reached == k) {
      atomicMax(stopSeed, (int32_t)s);
      return;
    }
  }
  atomicOr(notFull, 1);
}

__global__ __launch_bounds__(kB) void k_batch_select(
    const int32_t *__restrict__ table, const int32_t *__restrict__ seeds,
    int64_t nseeds, int64_t ntargets, int k, const int32_t *__restrict__ stopSeed,
    const int32_t *__restrict__ notFull, const int32_t *__restrict__ outIndex,
    int64_t stride, int32_t *__restrict__ outV, int32_t *__restrict__ outD,
    int32_t *__restrict__ outCount) {
  int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (t >= ntargets)
    return;
  const int64_t last = *notFull ? nseeds - 1 : (int64_t)*stopSeed;
  const int64_t o = (outIndex ? (int64_t)outIndex[t] : t) * stride;
  int cnt = 0;
  for (int64_t s = 0; s <= last; s++) {
    const int32_t d = table[s * ntargets + t];
    if (d < 0)
      continue;
    // knn_insert_sorted: append or replace the last, then bubble up while
    // strictly smaller; equal depths keep the earlier seed first
    int i;
    if (cnt < k) {
      i = cnt++;
    } else {
      if (d >= outD[o + k - 1])
        continue;
      i = k - 1;
    }
    outV[o + i] = seeds[s];
    outD[o + i] = d;
    while (i > 0 && outD[o + i] < outD[o + i - 1]) {
      int32_t tv = outV[o + i], td = outD[o + i];
      outV[o + i] = outV[o + i - 1];
      outD[o + i] = outD[o + i - 1];
      outV[o + i - 1] = tv;
      outD[o + i - 1] = td;
      i--;
    }
  }
  outCount[outIndex ? outIndex[t] : t] = cnt;
}

// Full BFS per seed writing target depths (the batch path's seed sweep).
__global__ __launch_bounds__(kB) void k_seed_depths(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const int32_t *__restrict__ seeds, int64_t nseeds,
    const int32_t *__restrict__ targetMap, int64_t ntargets, BfsSlots sl,
    int64_t slots, int32_t *__restrict__ table) {
  const int64_t slot = (int64_t)blockIdx.x * kWpb + threadIdx.x / kWave;
  if (slot >= slots)
    return;
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t *bits = sl.bits + slot * sl.words;
  int32_t *q = sl.queue + slot * sl.cap;
  int32_t *qd = sl.qdist + slot * sl.cap;
  for (int64_t si = slot; si < nseeds; si += slots) {
    int32_t *row = table + si * ntargets;
    for (int64_t t = lane; t < ntargets; t += kWave)
      row[t] = -1;
    int32_t ov = 0, od = 0;
    const int32_t seed = seeds[si];
    // a full BFS; hits are not collected (k = 0), the queue keeps every
    // visited vertex with its depth
    int head = 0, tail = 1;
    if (lane == 0) {
      atomicOr(&bits[seed >> 5], 1u << (seed & 31));
      st_l2(q, seed);
      st_l2(qd, 0);
    }
    const uint64_t lt = lane ? (~0ull >> (kWave - lane)) : 0ull;
    while (head < tail) {
      const int32_t u = __builtin_amdgcn_readfirstlane(ld_l2(q + head));
      const int32_t du = __builtin_amdgcn_readfirstlane(ld_l2(qd + head));
      head++;
      const int32_t e0 = off[u], e1 = off[u + 1];
      for (int32_t base = e0; base < e1; base += kWave) {
        const int32_t e = base + lane;
        const bool valid = e < e1;
        const int32_t w = valid ? adj[e] : -1;
        bool dup = false;
        for (int o2 = 1; o2 < kWave; o2++) {
          int32_t other = __shfl_up(w, o2, kWave);
          if (lane >= o2 && other == w)
            dup = true;
        }
        bool fresh = false;
        if (valid && !dup) {
          const uint32_t bit = 1u << (w & 31);
          fresh = !(atomicOr(&bits[w >> 5], bit) & bit);
        }
        const uint64_t fm = __ballot(fresh);
        if (fresh) {
          const int at = tail + (int)__popcll(fm & lt);
          st_l2(q + at, w);
          st_l2(qd + at, du + 1);
          const int32_t ti = targetMap[w];
          if (ti >= 0)
            row[ti] = du + 1;
        }
        tail += (int)__popcll(fm);
      }
    }
    for (int i = lane; i < tail; i += kWave)
      atomicAnd(&bits[ld_l2(q + i) >> 5], 0u);
    (void)ov;
    (void)od;
  }
}

} // namespace

struct gvizhip_grip {
  int64_t n = 0, m2 = 0, kcap = 0, activeCount = 0, edgeCount = 0;
  hipStream_t stream = nullptr;
  std::vector<int64_t> hoff, hadj;
  std::vector<uint8_t> hactive;
  DevBuf<int32_t> off, adj, filt;
  // the graph in breadth-first order for the BFS searches: rows in that
  // order with their entries in the original order, ids mapped both ways,
  // and a search's filter flags in that order
  DevBuf<int32_t> bfsOff, bfsAdj, bfsNew, bfsOrig;
  DevBuf<uint8_t> bfsFilter;
  DevBuf<uint8_t> active, visible, state;
  DevBuf<double2> pos, disp, oldDisp;
  DevBuf<double> heat, oldCos, lastF, partial;
  DevBuf<uint8_t> dispCalc;
  DevBuf<int32_t> knnV, knnD, knnCount;
  DevBuf<int32_t> qSrc, qOut, hitV, hitD, hitCount, flags;
  DevBuf<uint32_t> bits;
  DevBuf<int32_t> queue, qdist;
  DevBuf<int32_t> table, targetMap;
  DevBuf<GripRound> round;
  // three- and four-dimensional layouts: dim-strided positions and
  // displacements (the double2 buffers above serve dim == 2)
  int64_t dim = 2;
  DevBuf<double> posN, dispN, oldDispN;
  DevBuf<GripRoundN> roundN;
  DevBuf<uint64_t> keys;
  DevBuf<uint32_t> vals;
  RadixSorter sorter;
  // the active vertices of the current layer in ascending id: the order the
  // round kernels walk them in, so per-vertex state is read in sequence
  DevBuf<int32_t> order;
  int64_t orderActive = -1;
  int64_t slots = 0;
  int64_t filtCount = 0;
  int64_t blockQueries = -1; // see kBlockQueries
  // compact 2-D rounds (k_compact_*): the active vertices of the current
  // layer in layout order, their round state, kNN lists and (layer 0)
  // adjacency in compact indices
  DevBuf<int32_t> cOrder, cMap, cOff, cAdj, cKnnV, cKnnD, cKnnCount;
  DevBuf<double2> cPos, cDisp, cOldDisp;
  DevBuf<double> cHeat, cOldCos, cLastF, cBox;
  DevBuf<uint8_t> cDispCalc;
  int64_t cActive = -1;  // layer size the compact copy holds (-1: none)
  bool cAdjReady = false; // cOff / cAdj built for cActive
  bool cAhead = false;    // the compact state is newer than the id-ordered one
  int64_t compactMin = -1; // see kCompactMin
};

namespace {

// the compact round state (defined with the rounds below)
int compact_sync(gvizhip_grip *g, bool keep);

// calls with fewer searches than this give each search a block (the
// environment variable GVIZHIP_GRIP_BLOCK_QUERIES overrides it per engine;
// 0 keeps every search on one wavefront)
constexpr int64_t kBlockQueries = 2048;

int ensure_slots(gvizhip_grip *g) {
  if (g->blockQueries < 0) {
    const char *e = getenv("GVIZHIP_GRIP_BLOCK_QUERIES");
    g->blockQueries = e ? std::max<int64_t>(0, atoll(e)) : kBlockQueries;
  }
  if (g->slots)
    return 0;
  const int64_t words = (g->n + 31) / 32 + 1;
  const int64_t cap = g->n + 1;
  const double perSlot = (double)words * 4 + (double)cap * 8;
  // search scratch: up to a quarter of the free device memory, at most 48 GiB
  size_t freeB = 0, totalB = 0;
  double budget = 12.0 * (1ull << 30);
  if (hipMemGetInfo(&freeB, &totalB) == hipSuccess && freeB > 0)
    budget = std::min(48.0 * (1ull << 30), 0.25 * (double)freeB);
  int64_t s = (int64_t)std::min(1024.0, std::max(4.0, budget / perSlot));
  s = (s / kWpb) * kWpb;
  if (s < kWpb)
    s = kWpb;
  if (g->bits.alloc(s * words) || g->queue.alloc(s * cap) || g->qdist.alloc(s * cap))
    return -1;
  GVH_CHECK(hipMemsetAsync(g->bits.p, 0, sizeof(uint32_t) * s * words, g->stream));
  g->slots = s;
  return 0;
}

// Number of vertices a host filter admits (0 for no filter: no early stop).
int64_t count_set(const uint8_t *flags, int64_t n) {
  if (!flags)
    return 0;
  int64_t c = 0;
  for (int64_t v = 0; v < n; v++)
    c += flags[v] != 0;
  return c;
}

BfsSlots slot_view(gvizhip_grip *g) {
  BfsSlots v;
  v.bits = g->bits.p;
  v.queue = g->queue.p;
  v.qdist = g->qdist.p;
  v.words = (g->n + 31) / 32 + 1;
  v.cap = g->n + 1;
  return v;
}

int upload_visible(gvizhip_grip *g, const uint8_t *visible) {
  if (!visible)
    return 0;
  GVH_CHECK(hipMemcpyAsync(g->visible.p, visible, g->n, hipMemcpyHostToDevice,
                           g->stream));
  return 0;
}

// Filter flags in the searches' breadth-first order.
__global__ __launch_bounds__(kB) void k_flags_in_order(const uint8_t *__restrict__ in,
                                                      const int32_t *__restrict__ orig,
                                                      int64_t n, uint8_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n)
    out[i] = in[orig[i]];
}

// BFS queries on the device: sources (device int32), results into outV/outD
// rows of `stride` selected by outIndex (device, may be null). The searches
// run on the breadth-first copy of the graph; sources, filter and results
// are mapped at its boundary.

int run_queries(gvizhip_grip *g, const int32_t *srcDev, const int32_t *outIndex,
                int64_t count, int k, int64_t stride, const uint8_t *filterDev,
                int64_t filterTotal, int maxDepth, bool lowerOnly, bool stopAtK,
                int32_t *outV, int32_t *outD, int32_t *outCount) {
  if (count == 0)
    return 0;
  if (ensure_slots(g))
    return -1;
  const uint8_t *filter = nullptr;
  if (filterDev) {
    hipLaunchKernelGGL(k_flags_in_order, dim3(grid_for(g->n, kB)), dim3(kB), 0, g->stream,
                       filterDev, g->bfsOrig.p, g->n, g->bfsFilter.p);
    filter = g->bfsFilter.p;
  }
  if (count < g->blockQueries) {
    // few searches (the coarse layers): a block per search
    const int64_t nb = std::min<int64_t>(g->slots, count);
    hipLaunchKernelGGL(k_bfs_queries_block, dim3((unsigned)nb), dim3(kBB), 0, g->stream,
                       g->bfsOff.p, g->bfsAdj.p, filter, srcDev, outIndex, count, k, stride,
                       maxDepth, lowerOnly ? 1 : 0, stopAtK ? 1 : 0, filterTotal,
                       slot_view(g), nb, outV, outD, outCount, g->flags.p, g->bfsNew.p,
                       g->bfsOrig.p);
    GVH_CHECK_LAUNCH();
    return 0;
  }
  const int64_t waves = std::min<int64_t>(g->slots, count);
  const unsigned blocks = (unsigned)((waves + kWpb - 1) / kWpb);
  hipLaunchKernelGGL(k_bfs_queries, dim3(blocks), dim3(kB), 0, g->stream,
                     g->bfsOff.p, g->bfsAdj.p, filter, srcDev, outIndex, count, k,
                     stride, maxDepth, lowerOnly ? 1 : 0, stopAtK ? 1 : 0,
                     filterTotal, slot_view(g), (int64_t)blocks * kWpb, outV, outD,
                     outCount, g->flags.p, g->bfsNew.p, g->bfsOrig.p);
  GVH_CHECK_LAUNCH();
  return 0;
}

// The batched search of gvizSearchKNearestFromVisibleBatch for targets
// (device ids, count), writing rows selected by outIndex.
int run_batch(gvizhip_grip *g, const uint8_t *visibleHost, const int32_t *targetsDev,
              const std::vector<int32_t> &targetsHost, const int32_t *outIndex,
              int64_t count, int k, int64_t stride, int32_t *outV, int32_t *outD,
              int32_t *outCount) {
  std::vector<int32_t> seeds;
  for (int64_t v = 0; v < g->n; v++)
    if (visibleHost[v])
      seeds.push_back((int32_t)v);
  const int64_t ns = (int64_t)seeds.size();
  if (ns == 0 || count == 0)
    return 0;
  if (ensure_slots(g))
    return -1;
  std::vector<int32_t> tmap(g->n, -1);
  for (int64_t t = 0; t < count; t++)
    tmap[targetsHost[t]] = (int32_t)t;
  if (g->table.alloc(ns * count) || g->targetMap.alloc(g->n) ||
      g->qSrc.alloc(std::max<int64_t>(ns, count)))
    return -1;
  GVH_CHECK(hipMemcpyAsync(g->targetMap.p, tmap.data(), sizeof(int32_t) * g->n,
                           hipMemcpyHostToDevice, g->stream));
  DevBuf<int32_t> seedDev;
  if (seedDev.alloc(ns))
    return -1;
  GVH_CHECK(hipMemcpyAsync(seedDev.p, seeds.data(), sizeof(int32_t) * ns,
                           hipMemcpyHostToDevice, g->stream));
  const int64_t waves = std::min<int64_t>(g->slots, ns);
  const unsigned blocks = (unsigned)((waves + kWpb - 1) / kWpb);
  hipLaunchKernelGGL(k_seed_depths, dim3(blocks), dim3(kB), 0, g->stream, g->off.p,
                     g->adj.p, seedDev.p, ns, g->targetMap.p, count, slot_view(g),
                     (int64_t)blocks * kWpb, g->table.p);
  int32_t init[4] = {0, -1, 0, 0}; // overflow, stopSeed, notFull
  GVH_CHECK(hipMemcpyAsync(g->flags.p, init, sizeof(init), hipMemcpyHostToDevice,
                           g->stream));
  hipLaunchKernelGGL(k_batch_kth, dim3(grid_for(count, kB)), dim3(kB), 0, g->stream,
                     g->table.p, ns, count, k, g->flags.p + 1, g->flags.p + 2);
  hipLaunchKernelGGL(k_batch_select, dim3(grid_for(count, kB)), dim3(kB), 0,
                     g->stream, g->table.p, seedDev.p, ns, count, k, g->flags.p + 1,
                     g->flags.p + 2, outIndex, stride, outV, outD, outCount);
  GVH_CHECK_LAUNCH();
  GVH_CHECK(hipStreamSynchronize(g->stream));
  (void)targetsDev;
  return 0;
}

int read_flag(gvizhip_grip *g, int idx, int32_t *out) {
  GVH_CHECK(hipMemcpyAsync(out, g->flags.p + idx, sizeof(int32_t),
                           hipMemcpyDeviceToHost, g->stream));
  GVH_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int set_flag(gvizhip_grip *g, int idx, int32_t v) {
  GVH_CHECK(hipMemcpyAsync(g->flags.p + idx, &v, sizeof(int32_t),
                           hipMemcpyHostToDevice, g->stream));
  GVH_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

// Layer 1 (makeFirstMISPartition, gvizGRIPEmbedder.c:450-483): the
// lexicographically-first MIS of the subgraph in ascending id (`mis`), and
// the other subgraph vertices in the order the reference marks them (each
// MIS vertex marks its unmarked neighbours in adjacency order), which it
// writes backwards from the end of misFiltration (`tail`).
int first_layer(gvizhip_grip *g, std::vector<int64_t> &mis, std::vector<int64_t> &tail) {
  const int64_t n = g->n;
  hipStream_t st = g->stream;
  GVH_CHECK(hipMemsetAsync(g->state.p, 0, n, st));
  int32_t undecided = 1;
  int rounds = 0;
  while (undecided) {
    if (set_flag(g, 0, 0))
      return -1;
    for (int r = 0; r < 4; r++)
      hipLaunchKernelGGL(k_lfmis_graph, dim3(grid_for(n, kB)), dim3(kB), 0, st,
                         g->off.p, g->adj.p, g->active.p, n, g->state.p, g->flags.p);
    GVH_CHECK_LAUNCH();
    if (read_flag(g, 0, &undecided))
      return -1;
    if (++rounds > 4 * (n + 4)) {
      set_error("gvizhip_grip_mis_filtration: layer-1 rounds did not settle");
      return -1;
    }
  }
  std::vector<uint8_t> state(n);
  GVH_CHECK(hipMemcpy(state.data(), g->state.p, n, hipMemcpyDeviceToHost));
  mis.clear();
  for (int64_t v = 0; v < n; v++)
    if (g->hactive[v] && state[v] == kIn)
      mis.push_back(v);
  const int64_t tailCount = g->activeCount - (int64_t)mis.size();
  tail.assign(tailCount > 0 ? tailCount : 0, 0);
  if (tailCount > 0) {
    if (g->keys.alloc(tailCount) || g->vals.alloc(tailCount))
      return -1;
    if (set_flag(g, 0, 0))
      return -1;
    hipLaunchKernelGGL(k_tail_keys, dim3(grid_for(n, kB)), dim3(kB), 0, st, g->off.p,
                       g->adj.p, g->active.p, g->state.p, n, g->keys.p, g->vals.p,
                       g->flags.p);
    GVH_CHECK_LAUNCH();
    if (g->sorter.sort(g->keys.p, g->vals.p, tailCount, 0, 64, st))
      return -1;
    std::vector<uint32_t> order(tailCount);
    GVH_CHECK(hipMemcpyAsync(order.data(), g->vals.p, sizeof(uint32_t) * tailCount,
                             hipMemcpyDeviceToHost, st));
    GVH_CHECK(hipStreamSynchronize(st));
    for (int64_t r = 0; r < tailCount; r++)
      tail[r] = order[r];
  }
  return 0;
}

// Device buffers of the layer selection, kept across the layers of one
// filtration (the conflict-list stride only grows).
struct LayerScratch {
  DevBuf<int32_t> candDev, lists, listsD, listCount;
  DevBuf<uint8_t> candFlag, cstate;
  int64_t stride = 32;
  // multi-source ball search (layers with at most kMultiSources candidates)
  DevBuf<unsigned long long> msSeen, msFront, msNext;
  DevBuf<int32_t> msAct;
};

// Layers with at most this many candidates take the multi-source search
// (one pass over the chip per level for all of them); larger ones give each
// candidate its own search.
constexpr int64_t kMultiSources = 64;

__global__ void k_ms_seed(const int32_t *__restrict__ cand, int64_t m,
                          unsigned long long *__restrict__ seen,
                          unsigned long long *__restrict__ front) {
  const int j = threadIdx.x;
  if (j < m) {
    const int32_t v = cand[j];
    seen[v] = 1ull << j;
    front[v] = 1ull << j;
  }
}

// The balls of radius `radius` of the m (<= 64) candidates: ls.msSeen holds,
// per vertex, the candidates whose ball contains it. Levels run in groups
// of 16; a group whose last level moved no bit ends the search.
int multi_ball(gvizhip_grip *g, LayerScratch &ls, int64_t m, int64_t radius) {
  const int64_t n = g->n;
  hipStream_t st = g->stream;
  constexpr int kGroup = 16;
  if (ls.msSeen.alloc(n) || ls.msFront.alloc(n) || ls.msNext.alloc(n) ||
      ls.msAct.alloc(kGroup))
    return -1;
  const size_t bytes = sizeof(unsigned long long) * n;
  GVH_CHECK(hipMemsetAsync(ls.msSeen.p, 0, bytes, st));
  GVH_CHECK(hipMemsetAsync(ls.msFront.p, 0, bytes, st));
  GVH_CHECK(hipMemsetAsync(ls.msNext.p, 0, bytes, st));
  hipLaunchKernelGGL(k_ms_seed, dim3(1), dim3(kWave), 0, st, ls.candDev.p, m, ls.msSeen.p,
                     ls.msFront.p);
  GVH_CHECK_LAUNCH();
  unsigned long long *front = ls.msFront.p, *next = ls.msNext.p;
  const int64_t levels = std::min<int64_t>(radius, n);
  for (int64_t d0 = 1; d0 <= levels; d0 += kGroup) {
    GVH_CHECK(hipMemsetAsync(ls.msAct.p, 0, sizeof(int32_t) * kGroup, st));
    const int64_t d1 = std::min<int64_t>(d0 + kGroup - 1, levels);
    for (int64_t d = d0; d <= d1; d++) {
      hipLaunchKernelGGL(k_ms_push, dim3(grid_for(n, kB)), dim3(kB), 0, st, g->off.p,
                         g->adj.p, n, ls.msSeen.p, front, next, ls.msAct.p + (d - d0));
      std::swap(front, next);
    }
    GVH_CHECK_LAUNCH();
    int32_t moved = 0;
    GVH_CHECK(hipMemcpyAsync(&moved, ls.msAct.p + (d1 - d0), sizeof(int32_t),
                             hipMemcpyDeviceToHost, st));
    GVH_CHECK(hipStreamSynchronize(st));
    if (!moved)
      break;
  }
  return 0;
}

// Layer i >= 2 (iterMISFiltration, gvizGRIPEmbedder.c:672-739): of the
// candidates `cur` (ascending), the lexicographically-first set in which no
// two lie within BFS distance 2^(i-1) through the whole subgraph; picked[j]
// = 1 for the chosen.
int next_layer(gvizhip_grip *g, int64_t i, const std::vector<int64_t> &cur,
               std::vector<uint8_t> &picked, LayerScratch &ls) {
  const int64_t n = g->n;
  hipStream_t st = g->stream;
  const int64_t m = (int64_t)cur.size();
  if (ls.candFlag.alloc(n))
    return -1;
  const int64_t radius = (int64_t)1 << (i - 1 < 62 ? i - 1 : 62);
  std::vector<int32_t> c32(m > 0 ? m : 1);
  std::vector<uint8_t> flag(n, 0);
  for (int64_t j = 0; j < m; j++) {
    c32[j] = (int32_t)cur[j];
    flag[cur[j]] = 1;
  }
  if (ls.candDev.alloc(m > 0 ? m : 1) || ls.cstate.alloc(m > 0 ? m : 1) ||
      ls.listCount.alloc(m > 0 ? m : 1))
    return -1;
  GVH_CHECK(hipMemcpyAsync(ls.candDev.p, c32.data(), sizeof(int32_t) * (m > 0 ? m : 1),
                           hipMemcpyHostToDevice, st));
  GVH_CHECK(hipMemcpyAsync(ls.candFlag.p, flag.data(), n, hipMemcpyHostToDevice, st));
  // conflict lists: lower-id candidates within the radius (candidate index
  // space: the candidates are ascending, so a lower id is a lower index)
  int64_t &stride = ls.stride;
  const bool multi = m > 0 && m <= kMultiSources;
  if (multi && multi_ball(g, ls, m, radius))
    return -1;
  for (;;) {
    if (ls.lists.alloc(m * stride) || ls.listsD.alloc(m * stride) || set_flag(g, 0, 0))
      return -1;
    const int maxDepth = radius > INT32_MAX ? INT32_MAX : (int)radius;
    if (multi) {
      hipLaunchKernelGGL(k_ms_lists, dim3((unsigned)m), dim3(kB), 0, st, ls.candDev.p,
                         (int64_t)0, (int)m, ls.msSeen.p, stride, ls.lists.p, ls.listsD.p,
                         ls.listCount.p, g->flags.p);
      GVH_CHECK_LAUNCH();
    } else if (run_queries(g, ls.candDev.p, nullptr, m, (int)stride, stride,
                           ls.candFlag.p, 0, maxDepth, true, false, ls.lists.p,
                           ls.listsD.p, ls.listCount.p)) {
      return -1;
    }
    int32_t ov = 0;
    if (read_flag(g, 0, &ov))
      return -1;
    if (!ov)
      break;
    stride *= 2;
  }
  // map conflict vertex ids to candidate indices on the host
  std::vector<int32_t> lst(m * stride), cnt(m > 0 ? m : 1);
  GVH_CHECK(hipMemcpy(lst.data(), ls.lists.p, sizeof(int32_t) * m * stride,
                      hipMemcpyDeviceToHost));
  GVH_CHECK(hipMemcpy(cnt.data(), ls.listCount.p, sizeof(int32_t) * (m > 0 ? m : 1),
                      hipMemcpyDeviceToHost));
  for (int64_t j = 0; j < m; j++)
    for (int32_t t = 0; t < cnt[j]; t++) {
      int32_t v = lst[j * stride + t];
      lst[j * stride + t] =
          (int32_t)(std::lower_bound(cur.begin(), cur.end(), (int64_t)v) - cur.begin());
    }
  GVH_CHECK(hipMemcpy(ls.lists.p, lst.data(), sizeof(int32_t) * m * stride,
                      hipMemcpyHostToDevice));
  GVH_CHECK(hipMemsetAsync(ls.cstate.p, 0, m > 0 ? m : 1, st));
  int32_t undecided = m > 0;
  int64_t guard = 0;
  while (undecided) {
    if (set_flag(g, 0, 0))
      return -1;
    for (int r = 0; r < 4; r++)
      hipLaunchKernelGGL(k_lfmis_lists, dim3(grid_for(m, kB)), dim3(kB), 0, st,
                         ls.lists.p, ls.listCount.p, m, stride, ls.cstate.p, g->flags.p);
    GVH_CHECK_LAUNCH();
    if (read_flag(g, 0, &undecided))
      return -1;
    if (++guard > 4 * (m + 4)) {
      set_error("gvizhip_grip_mis_filtration: layer rounds did not settle");
      return -1;
    }
  }
  std::vector<uint8_t> cs(m > 0 ? m : 1);
  GVH_CHECK(hipMemcpy(cs.data(), ls.cstate.p, m > 0 ? m : 1, hipMemcpyDeviceToHost));
  picked.assign(m, 0);
  for (int64_t j = 0; j < m; j++)
    picked[j] = cs[j] == kIn;
  return 0;
}

// gripPickFarCandidate / migrateOneToFinalLayer (gvizGRIPEmbedder.c:507-597):
// a multi-source BFS of depth <= 64 from the final set, run on the host (the
// final layer holds at most dim + 1 vertices).
void migrate_one(gvizhip_grip *g, std::vector<int64_t> &filt,
                 std::vector<int64_t> &borders, int64_t count) {
  const int64_t finalEnd = borders[count - 1];
  int64_t srcLayer = count - 1, candEnd;
  do {
    if (srcLayer == 0)
      return;
    srcLayer--;
    candEnd = borders[srcLayer];
  } while (candEnd <= finalEnd);
  std::vector<int64_t> depth(g->n, -1);
  std::deque<int64_t> q;
  for (int64_t i = 0; i < finalEnd; i++) {
    depth[filt[i]] = 0;
    q.push_back(filt[i]);
  }
  while (!q.empty()) {
    int64_t v = q.front();
    q.pop_front();
    if (depth[v] >= 64)
      continue;
    for (int64_t e = g->hoff[v]; e < g->hoff[v + 1]; e++) {
      int64_t u = g->hadj[e];
      if (depth[u] >= 0)
        continue;
      depth[u] = depth[v] + 1;
      q.push_back(u);
    }
  }
  int64_t bestIdx = finalEnd, bestScore = 0;
  for (int64_t j = finalEnd; j < candEnd; j++) {
    int64_t v = filt[j];
    int64_t score = depth[v] >= 0 ? depth[v] : 65;
    if (score > bestScore || (score == bestScore && j > bestIdx)) {
      bestScore = score;
      bestIdx = j;
    }
  }
  std::swap(filt[bestIdx], filt[finalEnd]);
  borders[count - 1]++;
}

} // namespace

extern "C" {

int gvizhip_grip_create(gvizhip_grip **out, int64_t n, const int64_t *offsets,
                        const int64_t *adj, const uint8_t *active, int64_t kcap) {
  if (!out || n <= 0 || !offsets || kcap <= 0) {
    set_error("gvizhip_grip_create: invalid arguments");
    return -1;
  }
  *out = nullptr;
  int dc = 0;
  if (hipGetDeviceCount(&dc) != hipSuccess || dc <= 0) {
    set_error("gvizhip_grip_create: no GPU is visible to HIP");
    return -1;
  }
  const int64_t m2 = offsets[n];
  if (n >= INT32_MAX || m2 >= INT32_MAX || n * kcap >= ((int64_t)1 << 40)) {
    set_error("gvizhip_grip_create: graph too large");
    return -1;
  }
  gvizhip_grip *g = new gvizhip_grip();
  g->n = n;
  g->m2 = m2;
  g->kcap = kcap;
  g->hoff.assign(offsets, offsets + n + 1);
  g->hadj.assign(adj, adj + m2);
  g->hactive.assign(n, 1);
  if (active)
    g->hactive.assign(active, active + n);
  std::vector<int32_t> o32(n + 1), a32(m2 > 0 ? m2 : 1);
  for (int64_t i = 0; i <= n; i++)
    o32[i] = (int32_t)offsets[i];
  for (int64_t e = 0; e < m2; e++) {
    if (adj[e] < 0 || adj[e] >= n) {
      set_error("gvizhip_grip_create: adjacency entry out of range");
      delete g;
      return -1;
    }
    a32[e] = (int32_t)adj[e];
  }
  for (int64_t v = 0; v < n; v++)
    g->activeCount += g->hactive[v] != 0;
  g->edgeCount = m2 / 2;
  // breadth-first order (from vertex 0, then from each vertex not reached
  // yet) for the BFS searches: a ball's vertices get nearby ids there
  std::vector<int32_t> bfsOrig(n), bfsNew(n, -1);
  {
    int64_t tail = 0;
    for (int64_t s0 = 0; s0 < n; s0++) {
      if (bfsNew[s0] >= 0)
        continue;
      bfsNew[s0] = (int32_t)tail;
      bfsOrig[tail++] = (int32_t)s0;
      for (int64_t h = tail - 1; h < tail; h++) {
        const int32_t u = bfsOrig[h];
        for (int32_t e = o32[u]; e < o32[u + 1]; e++) {
          const int32_t w = a32[e];
          if (bfsNew[w] < 0) {
            bfsNew[w] = (int32_t)tail;
            bfsOrig[tail++] = w;
          }
        }
      }
    }
  }
  std::vector<int32_t> bo(n + 1), ba(m2 > 0 ? m2 : 1);
  bo[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    const int32_t u = bfsOrig[i];
    int32_t at = bo[i];
    for (int32_t e = o32[u]; e < o32[u + 1]; e++)
      ba[at++] = bfsNew[a32[e]];
    bo[i + 1] = at;
  }
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
      g->off.alloc(n + 1) || g->adj.alloc(m2 > 0 ? m2 : 1) || g->filt.alloc(n) ||
      g->active.alloc(n) || g->visible.alloc(n) || g->state.alloc(n) ||
      g->pos.alloc(n) || g->disp.alloc(n) || g->oldDisp.alloc(n) ||
      g->heat.alloc(n) || g->oldCos.alloc(n) || g->lastF.alloc(n) ||
      g->dispCalc.alloc(n) || g->knnV.alloc(n * kcap) || g->knnD.alloc(n * kcap) ||
      g->knnCount.alloc(n) || g->partial.alloc(4 * std::max<int64_t>(1024, (int64_t)grid_for(n > 0 ? n : 1, kB))) || g->flags.alloc(8) ||
      g->round.alloc(1) || g->qSrc.alloc(n) || g->qOut.alloc(n) ||
      g->hitCount.alloc(n) || g->bfsOff.alloc(n + 1) || g->bfsAdj.alloc(m2 > 0 ? m2 : 1) ||
      g->bfsNew.alloc(n) || g->bfsOrig.alloc(n) || g->bfsFilter.alloc(n)) {
    gvizhip_grip_destroy(g);
    return -1;
  }
  hipStream_t st = g->stream;
  if (hipMemcpyAsync(g->off.p, o32.data(), sizeof(int32_t) * (n + 1),
                     hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(g->bfsOff.p, bo.data(), sizeof(int32_t) * (n + 1),
                     hipMemcpyHostToDevice, st) != hipSuccess ||
      (m2 > 0 && hipMemcpyAsync(g->bfsAdj.p, ba.data(), sizeof(int32_t) * m2,
                                hipMemcpyHostToDevice, st) != hipSuccess) ||
      hipMemcpyAsync(g->bfsNew.p, bfsNew.data(), sizeof(int32_t) * n,
                     hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(g->bfsOrig.p, bfsOrig.data(), sizeof(int32_t) * n,
                     hipMemcpyHostToDevice, st) != hipSuccess ||
      (m2 > 0 && hipMemcpyAsync(g->adj.p, a32.data(), sizeof(int32_t) * m2,
                                hipMemcpyHostToDevice, st) != hipSuccess) ||
      hipMemcpyAsync(g->active.p, g->hactive.data(), n, hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      hipMemsetAsync(g->pos.p, 0, sizeof(double2) * n, st) != hipSuccess ||
      hipMemsetAsync(g->disp.p, 0, sizeof(double2) * n, st) != hipSuccess ||
      hipMemsetAsync(g->oldDisp.p, 0, sizeof(double2) * n, st) != hipSuccess ||
      hipMemsetAsync(g->heat.p, 0, sizeof(double) * n, st) != hipSuccess ||
      hipMemsetAsync(g->oldCos.p, 0, sizeof(double) * n, st) != hipSuccess ||
      hipMemsetAsync(g->lastF.p, 0, sizeof(double) * n, st) != hipSuccess ||
      hipMemsetAsync(g->dispCalc.p, 0, n, st) != hipSuccess ||
      hipMemsetAsync(g->knnCount.p, 0, sizeof(int32_t) * n, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    set_error("gvizhip_grip_create: upload failed");
    gvizhip_grip_destroy(g);
    return -1;
  }
  *out = g;
  return 0;
}

int gvizhip_grip_set_filtration(gvizhip_grip *g, const int64_t *filtration,
                                int64_t count) {
  if (!g || !filtration || count < 0 || count > g->n) {
    set_error("gvizhip_grip_set_filtration: invalid arguments");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  std::vector<int32_t> f32(count > 0 ? count : 1);
  for (int64_t i = 0; i < count; i++)
    f32[i] = (int32_t)filtration[i];
  GVH_CHECK(hipStreamSynchronize(g->stream));
  if (count > 0)
    GVH_CHECK(hipMemcpy(g->filt.p, f32.data(), sizeof(int32_t) * count,
                        hipMemcpyHostToDevice));
  g->filtCount = count;
  g->orderActive = -1;
  return 0;
}

int64_t gvizhip_grip_mis_filtration(gvizhip_grip *g, int64_t dim,
                                    int64_t *filtration, int64_t *borders) {
  if (!g || !filtration || !borders) {
    set_error("gvizhip_grip_mis_filtration: invalid arguments");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  std::vector<int64_t> bord;
  std::vector<int64_t> filt(g->activeCount);
  bord.push_back(g->activeCount);
  // ---- layer 1: lexicographically-first MIS of the subgraph
  std::vector<int64_t> cur, tail;
  if (first_layer(g, cur, tail))
    return -1;
  for (size_t r = 0; r < tail.size(); r++)
    filt[g->activeCount - 1 - (int64_t)r] = tail[r];
  bord.push_back((int64_t)cur.size());
  // ---- layers i >= 2: LF-MIS of the distance-2^(i-1) conflict graph
  LayerScratch ls;
  int64_t i = 2;
  for (;;) {
    std::vector<uint8_t> picked;
    if (next_layer(g, i, cur, picked, ls))
      return -1;
    std::vector<int64_t> next;
    int64_t ptr = bord[i - 1] - 1;
    for (size_t j = 0; j < cur.size(); j++) {
      if (picked[j])
        next.push_back(cur[j]);
      else
        filt[ptr--] = cur[j];
    }
    bord.push_back((int64_t)next.size());
    cur.swap(next);
    if ((int64_t)cur.size() > dim + 1) {
      i++;
      if (i > 126) {
        set_error("gvizhip_grip_mis_filtration: too many layers");
        return -1;
      }
      continue;
    }
    break;
  }
  for (int64_t k = 0; k < (int64_t)cur.size(); k++)
    filt[k] = cur[k];
  while (bord[i] < dim + 1) {
    int64_t before = bord[i];
    migrate_one(g, filt, bord, i + 1);
    if (bord[i] == before)
      break;
  }
  const int64_t layers = i + 1;
  for (int64_t k = 0; k < g->activeCount; k++)
    filtration[k] = filt[k];
  for (int64_t k = 0; k < layers; k++)
    borders[k] = bord[k];
  if (gvizhip_grip_set_filtration(g, filtration, g->activeCount))
    return -1;
  return layers;
}

int64_t gvizhip_grip_mis_first(gvizhip_grip *g, uint8_t *mis, int64_t *tail) {
  if (!g || !mis || !tail) {
    set_error("gvizhip_grip_mis_first: invalid arguments");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  std::vector<int64_t> cur, t;
  if (first_layer(g, cur, t))
    return -1;
  memset(mis, 0, (size_t)g->n);
  for (int64_t v : cur)
    mis[v] = 1;
  for (size_t r = 0; r < t.size(); r++)
    tail[r] = t[r];
  return (int64_t)cur.size();
}

int64_t gvizhip_grip_mis_layer(gvizhip_grip *g, int64_t layer, const int64_t *cand,
                               int64_t count, uint8_t *picked) {
  if (!g || layer < 2 || count < 0 || (count > 0 && (!cand || !picked))) {
    set_error("gvizhip_grip_mis_layer: invalid arguments");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  std::vector<int64_t> cur(cand, cand + count);
  for (int64_t j = 0; j < count; j++)
    if (cur[j] < 0 || cur[j] >= g->n || (j > 0 && cur[j] <= cur[j - 1])) {
      set_error("gvizhip_grip_mis_layer: candidates must be ascending vertex ids");
      return -1;
    }
  LayerScratch ls;
  std::vector<uint8_t> pk;
  if (next_layer(g, layer, cur, pk, ls))
    return -1;
  int64_t c = 0;
  for (int64_t j = 0; j < count; j++) {
    picked[j] = pk[j];
    c += pk[j];
  }
  return c;
}

int gvizhip_grip_migrate(gvizhip_grip *g, int64_t *filtration, int64_t filtrationCount,
                         int64_t *borders, int64_t layerCount, int64_t count) {
  if (!g || !filtration || !borders || count < 1 || count > layerCount ||
      filtrationCount != g->activeCount) {
    set_error("gvizhip_grip_migrate: invalid arguments");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  std::vector<int64_t> filt(filtration, filtration + filtrationCount);
  std::vector<int64_t> bord(borders, borders + layerCount);
  for (int64_t k = 0; k < layerCount; k++)
    if (bord[k] < 0 || bord[k] > filtrationCount) {
      set_error("gvizhip_grip_migrate: border out of range");
      return -1;
    }
  for (int64_t v : filt)
    if (v < 0 || v >= g->n) {
      set_error("gvizhip_grip_migrate: filtration entry out of range");
      return -1;
    }
  const int64_t before = bord[count - 1];
  migrate_one(g, filt, bord, count);
  if (bord[count - 1] == before)
    return -1; // every pool is exhausted
  for (int64_t k = 0; k < filtrationCount; k++)
    filtration[k] = filt[k];
  for (int64_t k = 0; k < layerCount; k++)
    borders[k] = bord[k];
  return 0;
}

int gvizhip_grip_knn(gvizhip_grip *g, const uint8_t *visible, const int64_t *sources,
                     int64_t count, int64_t k, int64_t *outVertex, int64_t *outDist,
                     int64_t *counts) {
  if (!g || count < 0 || k < 0 || (count > 0 && (!sources || !counts))) {
    set_error("gvizhip_grip_knn: invalid arguments");
    return -1;
  }
  if (compact_sync(g, true))
    return -1;
  if (count == 0)
    return 0;
  for (int64_t q = 0; q < count; q++)
    if (sources[q] < 0 || sources[q] >= g->n || !g->hactive[sources[q]]) {
      set_error("gvizhip_grip_knn: source %lld is not a subgraph vertex",
                (long long)sources[q]);
      return -1;
    }
  std::vector<int32_t> s32(count);
  for (int64_t q = 0; q < count; q++)
    s32[q] = (int32_t)sources[q];
  const int64_t kk = k > 0 ? k : 1;
  if (g->qSrc.alloc(count) || g->hitV.alloc(count * kk) || g->hitD.alloc(count * kk) ||
      g->hitCount.alloc(count))
    return -1;
  GVH_CHECK(hipMemcpyAsync(g->qSrc.p, s32.data(), sizeof(int32_t) * count,
                           hipMemcpyHostToDevice, g->stream));
  if (upload_visible(g, visible))
    return -1;
  if (k == 0) {
    for (int64_t q = 0; q < count; q++)
      counts[q] = 0;
    return 0;
  }
  if (run_queries(g, g->qSrc.p, nullptr, count, (int)k, k,
                  visible ? g->visible.p : nullptr, count_set(visible, g->n), 0,
                  false, true, g->hitV.p, g->hitD.p, g->hitCount.p))
    return -1;
  std::vector<int32_t> hv(count * k), hd(count * k), hc(count);
  GVH_CHECK(hipMemcpyAsync(hv.data(), g->hitV.p, sizeof(int32_t) * count * k,
                           hipMemcpyDeviceToHost, g->stream));
  GVH_CHECK(hipMemcpyAsync(hd.data(), g->hitD.p, sizeof(int32_t) * count * k,
                           hipMemcpyDeviceToHost, g->stream));
  GVH_CHECK(hipMemcpyAsync(hc.data(), g->hitCount.p, sizeof(int32_t) * count,
                           hipMemcpyDeviceToHost, g->stream));
  GVH_CHECK(hipStreamSynchronize(g->stream));
  for (int64_t q = 0; q < count; q++) {
    counts[q] = hc[q];
    for (int64_t j = 0; j < k; j++) {
      bool ok = j < hc[q];
      if (outVertex)
        outVertex[q * k + j] = ok ? hv[q * k + j] : -1;
      if (outDist)
        outDist[q * k + j] = ok ? hd[q * k + j] : -1;
    }
  }
  return 0;
}

int gvizhip_grip_knn_batch(gvizhip_grip *g, const uint8_t *visible,
                           const int64_t *targets, int64_t count, int64_t k,
                           int64_t *outVertex, int64_t *outDist, int64_t *counts) {
  if (!g || !visible || count < 0 || k <= 0 || (count > 0 && (!targets || !counts))) {
    set_error("gvizhip_grip_knn_batch: invalid arguments");
    return -1;
  }
  if (compact_sync(g, true))
    return -1;
  if (count == 0)
    return 0;
  std::vector<int32_t> t32(count);
  for (int64_t t = 0; t < count; t++)
    t32[t] = (int32_t)targets[t];
  if (g->hitV.alloc(count * k) || g->hitD.alloc(count * k) ||
      g->hitCount.alloc(count))
    return -1;
  if (run_batch(g, visible, nullptr, t32, nullptr, count, (int)k, k, g->hitV.p,
                g->hitD.p, g->hitCount.p))
    return -1;
  std::vector<int32_t> hv(count * k), hd(count * k), hc(count);
  GVH_CHECK(hipMemcpy(hv.data(), g->hitV.p, sizeof(int32_t) * count * k,
                      hipMemcpyDeviceToHost));
  GVH_CHECK(hipMemcpy(hd.data(), g->hitD.p, sizeof(int32_t) * count * k,
                      hipMemcpyDeviceToHost));
  GVH_CHECK(hipMemcpy(hc.data(), g->hitCount.p, sizeof(int32_t) * count,
                      hipMemcpyDeviceToHost));
  for (int64_t q = 0; q < count; q++) {
    counts[q] = hc[q];
    for (int64_t j = 0; j < k; j++) {
      bool ok = j < hc[q];
      if (outVertex)
        outVertex[q * k + j] = ok ? hv[q * k + j] : -1;
      if (outDist)
        outDist[q * k + j] = ok ? hd[q * k + j] : -1;
    }
  }
  return 0;
}

int gvizhip_grip_set_dimension(gvizhip_grip *g, int64_t dim) {
  if (!g || dim < 2 || dim > 4) {
    set_error("gvizhip_grip_set_dimension: dimension must be 2, 3 or 4");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  g->dim = dim;
  if (dim == 2)
    return 0;
  const int64_t n = g->n;
  const size_t nd = (size_t)(n * dim);
  const int64_t parts =
      std::max<int64_t>(9 * 1024, 2 * dim * (int64_t)grid_for(n > 0 ? n : 1, kB));
  if (g->posN.alloc(nd) || g->dispN.alloc(nd) || g->oldDispN.alloc(nd) ||
      g->roundN.alloc(1) || g->partial.alloc((size_t)parts))
    return -1;
  GVH_CHECK(hipMemsetAsync(g->posN.p, 0, sizeof(double) * nd, g->stream));
  GVH_CHECK(hipMemsetAsync(g->dispN.p, 0, sizeof(double) * nd, g->stream));
  GVH_CHECK(hipMemsetAsync(g->oldDispN.p, 0, sizeof(double) * nd, g->stream));
  GVH_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int gvizhip_grip_upload_nd(gvizhip_grip *g, const double *positions,
                           const gvizhip_grip_vertex_nd *dec, const int64_t *knnVertex,
                           const int64_t *knnDist) {
  if (!g) {
    set_error("gvizhip_grip_upload_nd: null context");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  // The copies below run on the null stream, which does not order against
  // the engine's non-blocking stream: let queued kernels finish first.
  GVH_CHECK(hipStreamSynchronize(g->stream));
  const int64_t n = g->n, D = g->dim;
  if (D == 2) {
    std::vector<gvizhip_grip_vertex> d2(dec ? n : 0);
    for (int64_t v = 0; dec && v < n; v++)
      d2[v] = {dec[v].disp[0], dec[v].disp[1], dec[v].oldDisp[0], dec[v].oldDisp[1],
               dec[v].heat, dec[v].oldCos, dec[v].lastForceMag, dec[v].knnCount,
               dec[v].dispCalculated};
    return gvizhip_grip_upload(g, positions, dec ? d2.data() : nullptr, knnVertex,
                               knnDist);
  }
  if (positions)
    GVH_CHECK(hipMemcpy(g->posN.p, positions, sizeof(double) * n * D,
                        hipMemcpyHostToDevice));
  if (dec) {
    std::vector<double> d(n * D), od(n * D), h(n), oc(n), lf(n);
    std::vector<uint8_t> dc(n);
    std::vector<int32_t> kc(n);
    for (int64_t v = 0; v < n; v++) {
      for (int64_t k = 0; k < D; k++) {
        d[v * D + k] = dec[v].disp[k];
        od[v * D + k] = dec[v].oldDisp[k];
      }
      h[v] = dec[v].heat;
      oc[v] = dec[v].oldCos;
      lf[v] = dec[v].lastForceMag;
      dc[v] = dec[v].dispCalculated ? 1 : 0;
      kc[v] = (int32_t)dec[v].knnCount;
    }
    GVH_CHECK(hipMemcpy(g->dispN.p, d.data(), sizeof(double) * n * D, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->oldDispN.p, od.data(), sizeof(double) * n * D,
                        hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->heat.p, h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->oldCos.p, oc.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->lastF.p, lf.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->dispCalc.p, dc.data(), n, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->knnCount.p, kc.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  }
  return gvizhip_grip_upload(g, nullptr, nullptr, knnVertex, knnDist);
}

int gvizhip_grip_download_nd(gvizhip_grip *g, double *positions,
                             gvizhip_grip_vertex_nd *dec, int64_t *knnVertex,
                             int64_t *knnDist) {
  if (!g) {
    set_error("gvizhip_grip_download_nd: null context");
    return -1;
  }
  if (compact_sync(g, true))
    return -1;
  const int64_t n = g->n, D = g->dim;
  if (D == 2) {
    std::vector<gvizhip_grip_vertex> d2(dec ? n : 0);
    if (gvizhip_grip_download(g, positions, dec ? d2.data() : nullptr, knnVertex,
                              knnDist))
      return -1;
    for (int64_t v = 0; dec && v < n; v++) {
      dec[v] = gvizhip_grip_vertex_nd{};
      dec[v].disp[0] = d2[v].dispX;
      dec[v].disp[1] = d2[v].dispY;
      dec[v].oldDisp[0] = d2[v].oldDispX;
      dec[v].oldDisp[1] = d2[v].oldDispY;
      dec[v].heat = d2[v].heat;
      dec[v].oldCos = d2[v].oldCos;
      dec[v].lastForceMag = d2[v].lastForceMag;
      dec[v].knnCount = d2[v].knnCount;
      dec[v].dispCalculated = d2[v].dispCalculated;
    }
    return 0;
  }
  GVH_CHECK(hipStreamSynchronize(g->stream));
  if (positions)
    GVH_CHECK(hipMemcpy(positions, g->posN.p, sizeof(double) * n * D,
                        hipMemcpyDeviceToHost));
  if (dec) {
    std::vector<double> d(n * D), od(n * D), h(n), oc(n), lf(n);
    std::vector<uint8_t> dc(n);
    std::vector<int32_t> kc(n);
    GVH_CHECK(hipMemcpy(d.data(), g->dispN.p, sizeof(double) * n * D, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(od.data(), g->oldDispN.p, sizeof(double) * n * D,
                        hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(h.data(), g->heat.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(oc.data(), g->oldCos.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(lf.data(), g->lastF.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(dc.data(), g->dispCalc.p, n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(kc.data(), g->knnCount.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    for (int64_t v = 0; v < n; v++) {
      dec[v] = gvizhip_grip_vertex_nd{};
      for (int64_t k = 0; k < D; k++) {
        dec[v].disp[k] = d[v * D + k];
        dec[v].oldDisp[k] = od[v * D + k];
      }
      dec[v].heat = h[v];
      dec[v].oldCos = oc[v];
      dec[v].lastForceMag = lf[v];
      dec[v].dispCalculated = dc[v];
      dec[v].knnCount = kc[v];
    }
  }
  return gvizhip_grip_download(g, nullptr, nullptr, knnVertex, knnDist);
}

int gvizhip_grip_upload(gvizhip_grip *g, const double *positions,
                        const gvizhip_grip_vertex *dec, const int64_t *knnVertex,
                        const int64_t *knnDist) {
  if (!g) {
    set_error("gvizhip_grip_upload: null context");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  // Null-stream copies: wait for kernels queued on the engine's stream.
  GVH_CHECK(hipStreamSynchronize(g->stream));
  const int64_t n = g->n;
  if (g->dim != 2 && (positions || dec)) {
    if (dec) {
      set_error("gvizhip_grip_upload: the vertex state of a %lld-dimensional layout "
                "goes through gvizhip_grip_upload_nd", (long long)g->dim);
      return -1;
    }
    GVH_CHECK(hipMemcpy(g->posN.p, positions, sizeof(double) * n * g->dim,
                        hipMemcpyHostToDevice));
    positions = nullptr;
  }
  if (positions)
    GVH_CHECK(hipMemcpy(g->pos.p, positions, sizeof(double2) * n,
                        hipMemcpyHostToDevice));
  if (dec) {
    std::vector<double2> d(n), od(n);
    std::vector<double> h(n), oc(n), lf(n);
    std::vector<uint8_t> dc(n);
    std::vector<int32_t> kc(n);
    for (int64_t v = 0; v < n; v++) {
      d[v] = {dec[v].dispX, dec[v].dispY};
      od[v] = {dec[v].oldDispX, dec[v].oldDispY};
      h[v] = dec[v].heat;
      oc[v] = dec[v].oldCos;
      lf[v] = dec[v].lastForceMag;
      dc[v] = dec[v].dispCalculated ? 1 : 0;
      kc[v] = (int32_t)dec[v].knnCount;
    }
    GVH_CHECK(hipMemcpy(g->disp.p, d.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->oldDisp.p, od.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->heat.p, h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->oldCos.p, oc.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->lastF.p, lf.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->dispCalc.p, dc.data(), n, hipMemcpyHostToDevice));
    GVH_CHECK(hipMemcpy(g->knnCount.p, kc.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  }
  const int64_t tot = n * g->kcap;
  if (knnVertex) {
    std::vector<int32_t> v(tot);
    for (int64_t i = 0; i < tot; i++)
      v[i] = (int32_t)knnVertex[i];
    GVH_CHECK(hipMemcpy(g->knnV.p, v.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice));
  }
  if (knnDist) {
    std::vector<int32_t> v(tot);
    for (int64_t i = 0; i < tot; i++)
      v[i] = (int32_t)knnDist[i];
    GVH_CHECK(hipMemcpy(g->knnD.p, v.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice));
  }
  return 0;
}

int gvizhip_grip_download(gvizhip_grip *g, double *positions, gvizhip_grip_vertex *dec,
                          int64_t *knnVertex, int64_t *knnDist) {
  if (!g) {
    set_error("gvizhip_grip_download: null context");
    return -1;
  }
  if (compact_sync(g, true))
    return -1;
  const int64_t n = g->n;
  GVH_CHECK(hipStreamSynchronize(g->stream));
  if (g->dim != 2 && (positions || dec)) {
    if (dec) {
      set_error("gvizhip_grip_download: the vertex state of a %lld-dimensional layout "
                "goes through gvizhip_grip_download_nd", (long long)g->dim);
      return -1;
    }
    GVH_CHECK(hipMemcpy(positions, g->posN.p, sizeof(double) * n * g->dim,
                        hipMemcpyDeviceToHost));
    positions = nullptr;
  }
  if (positions)
    GVH_CHECK(hipMemcpy(positions, g->pos.p, sizeof(double2) * n,
                        hipMemcpyDeviceToHost));
  if (dec) {
    std::vector<double2> d(n), od(n);
    std::vector<double> h(n), oc(n), lf(n);
    std::vector<uint8_t> dc(n);
    std::vector<int32_t> kc(n);
    GVH_CHECK(hipMemcpy(d.data(), g->disp.p, sizeof(double2) * n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(od.data(), g->oldDisp.p, sizeof(double2) * n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(h.data(), g->heat.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(oc.data(), g->oldCos.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(lf.data(), g->lastF.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(dc.data(), g->dispCalc.p, n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(kc.data(), g->knnCount.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    for (int64_t v = 0; v < n; v++) {
      dec[v].dispX = d[v].x;
      dec[v].dispY = d[v].y;
      dec[v].oldDispX = od[v].x;
      dec[v].oldDispY = od[v].y;
      dec[v].heat = h[v];
      dec[v].oldCos = oc[v];
      dec[v].lastForceMag = lf[v];
      dec[v].dispCalculated = dc[v];
      dec[v].knnCount = kc[v];
    }
  }
  const int64_t tot = n * g->kcap;
  if (knnVertex || knnDist) {
    std::vector<int32_t> v(tot), d(tot);
    GVH_CHECK(hipMemcpy(v.data(), g->knnV.p, sizeof(int32_t) * tot, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(d.data(), g->knnD.p, sizeof(int32_t) * tot, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < tot; i++) {
      if (knnVertex)
        knnVertex[i] = v[i];
      if (knnDist)
        knnDist[i] = d[i];
    }
  }
  return 0;
}

int gvizhip_grip_clear_decorators(gvizhip_grip *g, int64_t active) {
  if (!g || active < 0 || active > g->filtCount) {
    set_error("gvizhip_grip_clear_decorators: invalid arguments");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  if (g->dim == 3)
    hipLaunchKernelGGL(k_clear_nd<3>, dim3(grid_for(g->n, kB)), dim3(kB), 0, g->stream,
                       g->filt.p, active, g->n, g->dispN.p, g->oldDispN.p, g->heat.p,
                       g->oldCos.p, g->knnCount.p, g->dispCalc.p);
  else if (g->dim == 4)
    hipLaunchKernelGGL(k_clear_nd<4>, dim3(grid_for(g->n, kB)), dim3(kB), 0, g->stream,
                       g->filt.p, active, g->n, g->dispN.p, g->oldDispN.p, g->heat.p,
                       g->oldCos.p, g->knnCount.p, g->dispCalc.p);
  else
    hipLaunchKernelGGL(k_clear, dim3(grid_for(g->n, kB)), dim3(kB), 0, g->stream,
                       g->filt.p, active, g->n, g->disp.p, g->oldDisp.p, g->heat.p,
                       g->oldCos.p, g->knnCount.p, g->dispCalc.p);
  GVH_CHECK_LAUNCH();
  return 0;
}

int gvizhip_grip_update_knn(gvizhip_grip *g, int64_t active, const uint8_t *visible,
                            int64_t k, int32_t batch) {
  if (!g || !visible || active < 0 || active > g->filtCount || k < 0 ||
      k > g->kcap) {
    set_error("gvizhip_grip_update_knn: invalid arguments");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  if (active == 0 || k == 0)
    return 0;
  if (upload_visible(g, visible))
    return -1;
  if (batch) {
    std::vector<int32_t> f(active);
    GVH_CHECK(hipMemcpy(f.data(), g->filt.p, sizeof(int32_t) * active,
                        hipMemcpyDeviceToHost));
    return run_batch(g, visible, g->filt.p, f, g->filt.p, active, (int)k, g->kcap,
                     g->knnV.p, g->knnD.p, g->knnCount.p);
  }
  return run_queries(g, g->filt.p, g->filt.p, active, (int)k, g->kcap, g->visible.p,
                     count_set(visible, g->n), 0, false, true, g->knnV.p, g->knnD.p,
                     g->knnCount.p);
}

int gvizhip_grip_place(gvizhip_grip *g, int64_t begin, int64_t end,
                       const uint8_t *visible, int64_t k, int32_t batch) {
  if (!g || !visible || begin < 0 || end < begin || end > g->filtCount || k <= 0) {
    set_error("gvizhip_grip_place: invalid arguments");
    return -1;
  }
  if (compact_sync(g, false))
    return -1;
  const int64_t count = end - begin;
  if (count == 0)
    return 0;
  if (g->hitV.alloc(count * k) || g->hitD.alloc(count * k) || g->hitCount.alloc(count))
    return -1;
  if (upload_visible(g, visible))
    return -1;
  if (batch) {
    std::vector<int32_t> f(count);
    GVH_CHECK(hipMemcpy(f.data(), g->filt.p + begin, sizeof(int32_t) * count,
                        hipMemcpyDeviceToHost));
    if (run_batch(g, visible, g->filt.p + begin, f, nullptr, count, (int)k, k,
                  g->hitV.p, g->hitD.p, g->hitCount.p))
      return -1;
  } else if (run_queries(g, g->filt.p + begin, nullptr, count, (int)k, k,
                         g->visible.p, count_set(visible, g->n), 0, false, true,
                         g->hitV.p, g->hitD.p, g->hitCount.p)) {
    return -1;
  }
  if (set_flag(g, 3, 0))
    return -1;
  if (g->dim == 3)
    hipLaunchKernelGGL(k_place_nd<3>, dim3(grid_for(count, kB)), dim3(kB), 0, g->stream,
                       g->filt.p, begin, end, g->hitV.p, g->hitCount.p, k, g->posN.p,
                       g->flags.p + 3);
  else if (g->dim == 4)
    hipLaunchKernelGGL(k_place_nd<4>, dim3(grid_for(count, kB)), dim3(kB), 0, g->stream,
                       g->filt.p, begin, end, g->hitV.p, g->hitCount.p, k, g->posN.p,
                       g->flags.p + 3);
  else
    hipLaunchKernelGGL(k_place, dim3(grid_for(count, kB)), dim3(kB), 0, g->stream,
                       g->filt.p, begin, end, g->hitV.p, g->hitCount.p, k, g->pos.p,
                       g->flags.p + 3);
  GVH_CHECK_LAUNCH();
  int32_t empty = 0;
  if (read_flag(g, 3, &empty))
    return -1;
  if (empty) {
    set_error("gvizhip_grip_place: a vertex has no visible vertex in reach");
    return -1;
  }
  return 0;
}

} // extern "C"

namespace {

// One refinement round (src/embedders/gvizGRIPEmbedder.c:1289-1333) on the
// engine's stream; f0/f1 (may be null) are recorded around the spring-force
// kernel.
// The layer's active vertices (filtration[0..active)) in ascending id, built
// once per layer with the radix sort. Every vertex's round update reads only
// the previous state, so the walk order changes nothing but the order of the
// round's sums (means, rotation terms, statistics).
int build_order(gvizhip_grip *g, int64_t active) {
  if (g->orderActive == active)
    return 0;
  hipStream_t st = g->stream;
  if (g->order.alloc((size_t)active) || g->keys.alloc((size_t)active) ||
      g->vals.alloc((size_t)active))
    return -1;
  const unsigned b = grid_for(active, kB);
  hipLaunchKernelGGL(k_order_keys, dim3(b), dim3(kB), 0, st, g->filt.p, active, g->keys.p,
                     g->vals.p);
  GVH_CHECK_LAUNCH();
  if (g->sorter.sort(g->keys.p, g->vals.p, active, 0, 32, st))
    return -1;
  hipLaunchKernelGGL(k_order_ids, dim3(b), dim3(kB), 0, st, g->keys.p, active, g->order.p);
  GVH_CHECK_LAUNCH();
  g->orderActive = active;
  return 0;
}

// The round for D = 3 or 4 coordinates: the refinement pass, the means,
// the rotation removal (one solved 3-D rotation, or the six 4-D coordinate
// planes in turn), then the position update and statistics.
template <int D>
int enqueue_round_nd(gvizhip_grip *g, int64_t active, int64_t layer, hipEvent_t f0,
                     hipEvent_t f1) {
  hipStream_t st = g->stream;
  const unsigned blocks = std::min<unsigned>(grid_for(active, kB), 1024);
  const unsigned rblocks = grid_for(active, kB);
  if (f0)
    GVH_CHECK(hipEventRecord(f0, st));
  hipLaunchKernelGGL(k_refine_nd<D>, dim3(rblocks), dim3(kB), 0, st, g->order.p, active,
                     layer, g->off.p, g->adj.p, g->posN.p, g->knnV.p, g->knnD.p,
                     g->knnCount.p, g->kcap, g->dispN.p, g->oldDispN.p, g->heat.p,
                     g->oldCos.p, g->lastF.p, g->dispCalc.p, g->partial.p);
  if (f1)
    GVH_CHECK(hipEventRecord(f1, st));
  hipLaunchKernelGGL(k_round_means_nd<D>, dim3(1), dim3(256), 0, st, g->partial.p,
                     (int)rblocks, active, g->roundN.p);
  if (D == 3) {
    hipLaunchKernelGGL(k_rotation3_sums, dim3(blocks), dim3(kB), 0, st, g->order.p,
                       active, g->posN.p, g->dispN.p, g->roundN.p, g->partial.p);
    hipLaunchKernelGGL(k_rotation3_solve, dim3(1), dim3(256), 0, st, g->partial.p,
                       (int)blocks, active, g->roundN.p);
  } else {
    for (int plane = 0; plane < 6; plane++) {
      hipLaunchKernelGGL(k_plane_sums, dim3(blocks), dim3(kB), 0, st, g->order.p,
                         active, g->posN.p, g->dispN.p, g->roundN.p, plane, g->partial.p);
      hipLaunchKernelGGL(k_plane_omega, dim3(1), dim3(256), 0, st, g->partial.p,
                         (int)blocks, active, g->roundN.p, plane);
    }
  }
  hipLaunchKernelGGL(k_round_apply_nd<D>, dim3(blocks), dim3(kB), 0, st, g->order.p,
                     active, g->posN.p, g->dispN.p, g->heat.p, g->lastF.p, g->roundN.p,
                     g->partial.p);
  hipLaunchKernelGGL(k_round_stats_nd, dim3(1), dim3(256), 0, st, g->partial.p,
                     (int)blocks, g->roundN.p);
  GVH_CHECK_LAUNCH();
  return 0;
}

constexpr int kCompactFlag = 4; // flags slot of the compact builds
// layers with at least this many active vertices run their rounds compact
// (smaller ones fit the caches in any order); the environment variable
// GVIZHIP_GRIP_COMPACT_MIN overrides it per engine (0: every layer)
constexpr int64_t kCompactMin = 32768;

int64_t compact_min(gvizhip_grip *g) {
  if (g->compactMin < 0) {
    const char *e = getenv("GVIZHIP_GRIP_COMPACT_MIN");
    g->compactMin = e ? std::max<int64_t>(0, atoll(e)) : kCompactMin;
  }
  return g->compactMin;
}

// Writes the compact state back to the id-ordered arrays when it is newer;
// `keep` leaves the compact copy in use (a read of the state), otherwise the
// next round rebuilds it (the state, the kNN lists or the layers change).
int compact_sync(gvizhip_grip *g, bool keep) {
  if (g->cAhead && g->cActive > 0) {
    hipLaunchKernelGGL(k_compact_scatter, dim3(grid_for(g->cActive, kB)), dim3(kB), 0,
                       g->stream, g->cOrder.p, g->cActive, g->cPos.p, g->cDisp.p,
                       g->cOldDisp.p, g->cHeat.p, g->cOldCos.p, g->cLastF.p,
                       g->cDispCalc.p, g->pos.p, g->disp.p, g->oldDisp.p, g->heat.p,
                       g->oldCos.p, g->lastF.p, g->dispCalc.p);
    GVH_CHECK_LAUNCH();
  }
  g->cAhead = false;
  if (!keep) {
    g->cActive = -1;
    g->cAdjReady = false;
  }
  return 0;
}

// Builds the compact copy of layer `active` (layer 0 also needs the
// adjacency). Returns 1 when the layer cannot run compact (a kNN target or
// neighbour outside it), 0 when built, -1 on error.
int ensure_compact(gvizhip_grip *g, int64_t active, int64_t layer) {
  hipStream_t st = g->stream;
  const int64_t n = g->n;
  if (g->cActive != active) {
    if (compact_sync(g, false))
      return -1;
    if (g->cOrder.alloc(active) || g->cMap.alloc(n) || g->cPos.alloc(active) ||
        g->cDisp.alloc(active) || g->cOldDisp.alloc(active) || g->cHeat.alloc(active) ||
        g->cOldCos.alloc(active) || g->cLastF.alloc(active) ||
        g->cDispCalc.alloc(active) || g->cKnnV.alloc(active * g->kcap) ||
        g->cKnnD.alloc(active * g->kcap) || g->cKnnCount.alloc(active) ||
        g->cBox.alloc(4) || g->keys.alloc(active) || g->vals.alloc(active))
      return -1;
    const unsigned bb = std::min<unsigned>(grid_for(active, kB), 1024);
    hipLaunchKernelGGL(k_active_bbox, dim3(bb), dim3(kB), 0, st, g->filt.p, active,
                       g->pos.p, g->partial.p);
    hipLaunchKernelGGL(k_bbox_finish, dim3(1), dim3(256), 0, st, g->partial.p, (int)bb,
                       g->cBox.p);
    hipLaunchKernelGGL(k_layout_keys, dim3(grid_for(active, kB)), dim3(kB), 0, st,
                       g->filt.p, active, g->pos.p, g->cBox.p, g->keys.p, g->vals.p);
    GVH_CHECK_LAUNCH();
    if (g->sorter.sort(g->keys.p, g->vals.p, active, 0, 32, st))
      return -1;
    GVH_CHECK(hipMemsetAsync(g->cMap.p, 0xFF, sizeof(int32_t) * n, st));
    hipLaunchKernelGGL(k_compact_gather, dim3(grid_for(active, kB)), dim3(kB), 0, st,
                       g->vals.p, g->filt.p, active, g->cOrder.p, g->cMap.p, g->pos.p,
                       g->disp.p, g->oldDisp.p, g->heat.p, g->oldCos.p, g->lastF.p,
                       g->dispCalc.p, g->cPos.p, g->cDisp.p, g->cOldDisp.p, g->cHeat.p,
                       g->cOldCos.p, g->cLastF.p, g->cDispCalc.p);
    if (set_flag(g, kCompactFlag, 0))
      return -1;
    hipLaunchKernelGGL(k_compact_knn, dim3(grid_for(active, kB)), dim3(kB), 0, st,
                       g->cOrder.p, active, g->cMap.p, n, g->knnV.p, g->knnD.p,
                       g->knnCount.p, g->kcap, g->cKnnV.p, g->cKnnD.p, g->cKnnCount.p,
                       g->flags.p + kCompactFlag);
    GVH_CHECK_LAUNCH();
    g->cActive = active;
    g->cAdjReady = false;
    g->cAhead = false;
    int32_t bad = 0;
    if (read_flag(g, kCompactFlag, &bad))
      return -1;
    if (bad) {
      g->cActive = -1;
      return 1;
    }
  }
  if (layer == 0 && !g->cAdjReady) {
    // row starts on the host (it keeps the offsets), rows on the device
    std::vector<int32_t> order(active), cOff(active + 1);
    GVH_CHECK(hipMemcpyAsync(order.data(), g->cOrder.p, sizeof(int32_t) * active,
                             hipMemcpyDeviceToHost, st));
    GVH_CHECK(hipStreamSynchronize(st));
    int64_t acc = 0;
    for (int64_t c = 0; c < active; c++) {
      cOff[c] = (int32_t)acc;
      acc += g->hoff[order[c] + 1] - g->hoff[order[c]];
    }
    cOff[active] = (int32_t)acc;
    if (g->cOff.alloc(active + 1) || g->cAdj.alloc(acc > 0 ? acc : 1))
      return -1;
    GVH_CHECK(hipMemcpyAsync(g->cOff.p, cOff.data(), sizeof(int32_t) * (active + 1),
                             hipMemcpyHostToDevice, st));
    if (set_flag(g, kCompactFlag, 0))
      return -1;
    hipLaunchKernelGGL(k_compact_adj, dim3(grid_for(active, kB)), dim3(kB), 0, st,
                       g->cOrder.p, active, g->cMap.p, g->off.p, g->adj.p, g->cOff.p,
                       g->cAdj.p, g->flags.p + kCompactFlag);
    GVH_CHECK_LAUNCH();
    int32_t bad = 0;
    if (read_flag(g, kCompactFlag, &bad))
      return -1;
    if (bad) {
      g->cActive = -1;
      return 1;
    }
    g->cAdjReady = true;
  }
  return 0;
}

int enqueue_round(gvizhip_grip *g, int64_t active, int64_t layer, unsigned *blocksOut,
                  hipEvent_t f0, hipEvent_t f1) {
  hipStream_t st = g->stream;
  if (build_order(g, active))
    return -1;
  if (g->dim == 3 || g->dim == 4) {
    if (blocksOut)
      *blocksOut = std::min<unsigned>(grid_for(active, kB), 1024);
    return g->dim == 3 ? enqueue_round_nd<3>(g, active, layer, f0, f1)
                       : enqueue_round_nd<4>(g, active, layer, f0, f1);
  }
  const unsigned blocks = std::min<unsigned>(grid_for(active, kB), 1024);
  // the compact copy of the layer when it applies, else the id-ordered state
  int cm = 1;
  if (active >= compact_min(g))
    cm = ensure_compact(g, active, layer);
  else if (compact_sync(g, false))
    return -1;
  if (cm < 0)
    return -1;
  const bool compact = cm == 0;
  const int32_t *ord = compact ? nullptr : g->order.p;
  const int32_t *off = compact ? g->cOff.p : g->off.p;
  const int32_t *adj = compact ? g->cAdj.p : g->adj.p;
  double2 *pos = compact ? g->cPos.p : g->pos.p;
  const int32_t *kv = compact ? g->cKnnV.p : g->knnV.p;
  const int32_t *kd = compact ? g->cKnnD.p : g->knnD.p;
  const int32_t *kc = compact ? g->cKnnCount.p : g->knnCount.p;
  double2 *disp = compact ? g->cDisp.p : g->disp.p;
  double2 *oldDisp = compact ? g->cOldDisp.p : g->oldDisp.p;
  double *heat = compact ? g->cHeat.p : g->heat.p;
  double *oldCos = compact ? g->cOldCos.p : g->oldCos.p;
  double *lastF = compact ? g->cLastF.p : g->lastF.p;
  uint8_t *dispCalc = compact ? g->cDispCalc.p : g->dispCalc.p;
  if (f0)
    GVH_CHECK(hipEventRecord(f0, st));
  // one thread per vertex for the gathers: every CU's wave slots filled
  const unsigned rblocks = grid_for(active, kB);
  hipLaunchKernelGGL(k_refine, dim3(rblocks), dim3(kB), 0, st, ord, active, layer, off,
                     adj, pos, kv, kd, kc, g->kcap, disp, oldDisp, heat, oldCos, lastF,
                     dispCalc, g->partial.p);
  if (f1)
    GVH_CHECK(hipEventRecord(f1, st));
  hipLaunchKernelGGL(k_round_means, dim3(1), dim3(256), 0, st, g->partial.p,
                     (int)rblocks, active, g->round.p);
  hipLaunchKernelGGL(k_rotation_sums, dim3(blocks), dim3(kB), 0, st, ord, active, pos,
                     disp, g->round.p, g->partial.p);
  hipLaunchKernelGGL(k_round_omega, dim3(1), dim3(256), 0, st, g->partial.p,
                     (int)blocks, active, g->round.p);
  hipLaunchKernelGGL(k_round_apply, dim3(blocks), dim3(kB), 0, st, ord, active, pos, disp,
                     heat, lastF, g->round.p, g->partial.p);
  if (compact)
    g->cAhead = true;
  hipLaunchKernelGGL(k_round_stats, dim3(1), dim3(256), 0, st, g->partial.p,
                     (int)blocks, active, g->round.p);
  GVH_CHECK_LAUNCH();
  if (blocksOut)
    *blocksOut = blocks;
  return 0;
}

} // namespace

extern "C" {

int gvizhip_grip_refine_round(gvizhip_grip *g, int64_t active, int64_t layer,
                              gvizhip_grip_round_stats *stats) {
  if (!g || active < 0 || active > g->filtCount) {
    set_error("gvizhip_grip_refine_round: invalid arguments");
    return -1;
  }
  hipStream_t st = g->stream;
  GripRound r{};
  GripRoundN rn{};
  if (active > 0) {
    if (enqueue_round(g, active, layer, nullptr, nullptr, nullptr))
      return -1;
    if (g->dim == 2)
      GVH_CHECK(hipMemcpyAsync(&r, g->round.p, sizeof(GripRound), hipMemcpyDeviceToHost,
                               st));
    else
      GVH_CHECK(hipMemcpyAsync(&rn, g->roundN.p, sizeof(GripRoundN),
                               hipMemcpyDeviceToHost, st));
  }
  GVH_CHECK(hipStreamSynchronize(st));
  if (g->dim != 2) {
    r.maxDisp = rn.maxDisp;
    r.sumDisp = rn.sumDisp;
    r.sumHeat = rn.sumHeat;
    r.sumForce = rn.sumForce;
  }
  if (stats) {
    stats->maxDisp = r.maxDisp;
    stats->meanDisp = active ? r.sumDisp / (double)active : 0.0;
    stats->meanForce = active ? r.sumForce / (double)active : 0.0;
    stats->meanHeat = active ? r.sumHeat / (double)active : 0.0;
  }
  return 0;
}

int gvizhip_grip_profile_rounds(gvizhip_grip *g, int64_t active, int64_t layer,
                                int64_t rounds, double *forceMs, double *roundMs) {
  if (!g || active < 0 || active > g->filtCount || rounds < 0) {
    set_error("gvizhip_grip_profile_rounds: invalid arguments");
    return -1;
  }
  double fsum = 0.0, rsum = 0.0;
  if (active > 0 && rounds > 0) {
    hipEvent_t e[4];
    for (auto &x : e)
      GVH_CHECK(hipEventCreate(&x));
    int rc = 0;
    for (int64_t k = 0; k < rounds && rc == 0; k++) {
      if (hipEventRecord(e[2], g->stream) != hipSuccess ||
          enqueue_round(g, active, layer, nullptr, e[0], e[1]) ||
          hipEventRecord(e[3], g->stream) != hipSuccess ||
          hipStreamSynchronize(g->stream) != hipSuccess) {
        rc = -1;
        break;
      }
      float a = 0.0f, b = 0.0f;
      (void)hipEventElapsedTime(&a, e[0], e[1]);
      (void)hipEventElapsedTime(&b, e[2], e[3]);
      fsum += a;
      rsum += b;
    }
    for (auto &x : e)
      (void)hipEventDestroy(x);
    if (rc) {
      if (!gvizhip_last_error()[0])
        set_error("gvizhip_grip_profile_rounds failed");
      return -1;
    }
  }
  if (forceMs)
    *forceMs = fsum;
  if (roundMs)
    *roundMs = rsum;
  return 0;
}

void gvizhip_grip_destroy(gvizhip_grip *g) {
  if (!g)
    return;
  if (g->stream) {
    (void)hipStreamSynchronize(g->stream);
    (void)hipStreamDestroy(g->stream);
  }
  delete g;
}

} // extern "C"
