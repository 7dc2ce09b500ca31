// force_engine.hpp -- device state and pipeline of the force-directed step
// (the reference's gvizForceEmbedderStep, src/embedders/gvizForceEmbedder.c:
// 445-495) on one gfx950 GPU.
//
// Data layout in HBM (n = active vertices, compact ids in ascending parent id
// order, m2 = directed adjacency entries):
//   pos64      double2[n]   master positions (FP64, compact order)
//   csrOff/Adj int32        adjacency in insertion order (attraction order)
//   keyHi/Lo   uint64[n]    quadrant-path keys, 2 bits per level
//   vals       uint32[n]    original ids, permuted by the key sort
//   sorted SoA  spos/smr    positions and (mass, radius) in key order
//   tree       int32        binary radix tree over the sorted keys (Karras);
//                           the quadtree is its subset of "quad heads"
//   prefix     double-double prefix sums of m, m*x, m*y in key order: the
//                           mass and centre of mass of any key range in O(1)
//   nodes      NodeRec[n]   traversal records of the quad heads
//   forces     typed        rep, disp, structForce, swinging, |att|, |rep|
#pragma once

#include "common.hpp"
#include "radix_sort.hpp"

namespace gvh {

enum Precision : int32_t { kFP32 = 0, kFP64 = 1, kMixed = 2 };

struct ForceParams {
  int32_t model;          // 0 FR, 1 LinLog
  int32_t barnesHut;
  int32_t preventOverlap;
  int32_t precision;
  int64_t nodesPerCell;
  double edgeLength;
  double theta;
  double gravityK;
  double overlapConstant;
  double radiusBase;
  double radiusPerDegree;
  double jitterTolerance;
};

// One record per step, written on the device by the stats kernel.
struct StepRecord {
  double maxDisp, speed, meanAtt, meanRep, meanGrav;
  double speedEfficiency, totalSwinging, totalTraction;
};

// Device-resident scalars of one embedder.
struct Scalars {
  double rootCx, rootCy, rootHalf;
  int32_t maxDepth;  // levels that carry digits (capped at 64)
  int32_t rootRef;
  double globalSpeed, speedEfficiency;
  double epsilon;    // Run's threshold; < 0 disables the stop flag
  int64_t iteration; // steps taken since Begin
  int64_t recordBase;// iteration whose record lands in records[0]
  int32_t stopped;
  int32_t traversalOverflow; // 1 stack overflow, 2 tree too deep
  double meanDegreePlusOne;
  int32_t treeDepth; // digit levels + 2: a bound on the traversal's deepest walk
  int32_t pad32;
  double pad[3];
};

constexpr int32_t kNone = INT32_MIN;
// quadrant-path levels a batch key holds (12 bits go to the graph id)
constexpr int kBatchLevels = 58;

// Segments of a batch of independent graphs laid out together (all null for
// a single graph). Graph g owns compact (and sorted) positions
// [gStart[g], gStart[g+1]); its reductions use the same block decomposition a
// run of that graph alone uses (gBlocks[g] blocks, partial slots from
// gPart[g]); its traversal waves start at multiples of 64 of its own range.
struct SegView {
  const int32_t *vbGraph = nullptr;   // reduction block -> graph
  const int32_t *vbIdx = nullptr;     // reduction block -> index within graph
  const int64_t *gStart = nullptr;    // G + 1
  const int32_t *gBlocks = nullptr;   // G
  const int64_t *gPart = nullptr;     // G
  const int32_t *waveGraph = nullptr; // traversal wave -> graph
  const int64_t *waveBase = nullptr;  // traversal wave -> first sorted position
  const int32_t *rootOf = nullptr;    // graph -> root quad record (-1: none)
  int64_t waves = 0;
};

template <class Real, class PosT> struct NodeRec {
  PosT comx, comy;
  Real mass, twoh;
  int32_t child[4];
  int32_t first, count; // count > 0: leaf holding sorted points [first, first+count)
};

struct ForceCommon {
  int64_t n = 0;
  int64_t m2 = 0;
  ForceParams prm{};
  hipStream_t stream = nullptr;

  DevBuf<double> pos64;       // 2n
  DevBuf<int32_t> csrOff;     // n+1
  DevBuf<int32_t> csrAdj;     // m2
  DevBuf<int32_t> degree;     // n
  DevBuf<double> mass64;      // n  (exact: 1 or 1+deg)
  DevBuf<double> radius64;    // n  (0 when overlap prevention is off)
  DevBuf<double> mr64;        // 2n (mass, radius) pairs, one load per vertex
  DevBuf<double> smass64;     // n  masses in key order (prefix sums)

  DevBuf<uint64_t> keyHi, keyLo, keyLoS;
  DevBuf<uint32_t> vals;
  RadixSorter sorter;

  DevBuf<int32_t> left, right, parent, first, last, delta;
  DevBuf<double> pm, pmxHi, pmxLo, pmyHi, pmyLo; // n+1 prefix sums
  DevBuf<double> scanPartial;                    // 5 per scan block
  DevBuf<unsigned char> nodes;                   // NodeRec[n] (typed)

  // typed (allocated for the widest type)
  DevBuf<unsigned char> posT, spos, smr, rep, disp, sf, sw, attMag, repMag,
      massT, radT;
  DevBuf<double> spos64;   // 2n, FP64 sorted positions (COM prefix sums)
  DevBuf<double> partials; // reduction partials
  DevBuf<Scalars> sc;
  DevBuf<StepRecord> records;
  int64_t recordCap = 0;
  DevBuf<int32_t> interactions; // optional per-vertex repulsion count
  int wantInteractions = 0;
  DevBuf<int32_t> heavyRows;            // vertices with degree > kHeavyDegree
  int64_t nHeavy = 0;
  DevBuf<unsigned char> attHeavy;       // attraction of every vertex (typed, n)
  // attraction runs on a side stream, forked after the positions are ready
  // and joined before the force sums, so it overlaps the tree and traversal
  hipStream_t side = nullptr;
  hipEvent_t evFork = nullptr, evJoin = nullptr;
  // measurement hook: events around the traversal kernel (never captured)
  hipEvent_t travBegin = nullptr, travEnd = nullptr;
  int timeTraversal = 0;
  int travWaves = 1; // traversal wavefronts per workgroup (1, 2 or 4)
  DevBuf<unsigned long long> travStats; // null unless GVIZHIP_TRAVERSAL_STATS
  int64_t stepsSinceRecordBase = 0;
  int reductionBlocks = 0;

  int alloc(int64_t nVertices, int64_t nEntries);
  int open_streams();   // main stream, side stream and fork/join events
  void close_streams(); // synchronizes and releases them
};

// A batch of independent graphs laid out together (graph g owns compact
// positions [gStart[g], gStart[g+1]); up to 4096 graphs, whose ids take the
// top 12 bits of the sort key). Device tables for SegView plus per-graph
// scalars, last step records and reduction partials.
struct BatchState {
  int64_t graphs = 0, vblocks = 0, waves = 0;
  DevBuf<int32_t> gid;                  // n
  DevBuf<int64_t> gStart;               // G + 1
  DevBuf<int32_t> gBlocks;              // G
  DevBuf<int64_t> gPart;                // G
  DevBuf<int32_t> vbGraph, vbIdx;       // vblocks
  DevBuf<int32_t> waveGraph;            // waves
  DevBuf<int64_t> waveBase;             // waves
  DevBuf<int32_t> rootOf;               // G
  DevBuf<Scalars> scG;                  // G
  DevBuf<StepRecord> last;              // G
  DevBuf<double> partials;              // 3 * vblocks
  SegView view() const {
    SegView v;
    v.vbGraph = vbGraph.p;
    v.vbIdx = vbIdx.p;
    v.gStart = gStart.p;
    v.gBlocks = gBlocks.p;
    v.gPart = gPart.p;
    v.waveGraph = waveGraph.p;
    v.waveBase = waveBase.p;
    v.rootOf = rootOf.p;
    v.waves = waves;
    return v;
  }
};

// Step modes: the whole step, the force evaluation without the speed update
// and position update, or only the quadtree build.
enum StepMode : int { kStepFull = 0, kStepForcesOnly = 1, kStepTreeOnly = 2 };

// One step of every graph of a batch (Barnes-Hut only); c.sc must hold
// neutral scalars (not stopped, depth limit 64). Instantiated next to
// typed_step, so each precision is compiled exactly as its single-graph step.
template <class Real, class PosT> int typed_batch_step(ForceCommon &c, BatchState &b);
// Bounding box, root cell, keys and key sort of c.pos64 (c.n points).
int enqueue_keys_and_sort(ForceCommon &c);
int enqueue_batch_keys_and_sort(ForceCommon &c, BatchState &b);
int enqueue_batch_fa2(ForceCommon &c, BatchState &b);
int enqueue_batch_stats(ForceCommon &c, BatchState &b);

// Enqueues one step on c.stream. Instantiated in force_f64.hip
// (<double,double>) and force_f32.hip (<float,float>, <float,double>).
template <class Real, class PosT> int typed_step(ForceCommon &c, int mode);
// Refreshes the typed compact-order mass and radius arrays.
template <class Real> int typed_upload(ForceCommon &c);

const char *last_error();

} // namespace gvh
