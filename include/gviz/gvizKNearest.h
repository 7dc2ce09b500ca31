/*
 * gvizKNearest.h -- the reference's graph-distance kNN search
 * (include/algorithms/search/gvizKNearest.h). Single queries run on the
 * host (one breadth-first search that stops at its k-th hit); the batched
 * search over a small visible seed set runs on the GPU
 * (gvizhip_grip_knn_batch). Same results, return codes and tie rules as the
 * reference.
 */
#ifndef GVIZ_HOST_KNEAREST_H
#define GVIZ_HOST_KNEAREST_H

#include "gvizGraph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/algorithms/search/gvizKNearest.h:8-13 */
#define GVIZ_KNN_BATCH_VISIBLE_MAX 256
#define GVIZ_KNN_BATCH_MIN_AVG_DEGREE 32.0
#define GVIZ_KNN_BATCH_TARGET_RATIO 8

typedef struct gvizFoundVertex {
  size_t v;
  size_t dist;
} gvizFoundVertex;

typedef struct gvizKNearestScratch {
  size_t *stamp;
  size_t nvertices;
  size_t epoch;
  gvizDeque queue;
} gvizKNearestScratch;

typedef struct gvizKNNBatchTarget {
  size_t vertex;
  gvizFoundVertex *out;
  size_t count;
} gvizKNNBatchTarget;

/* Visited stamps and a queue for graphs of up to nvertices vertices. 0 / -1. */
int gvizKNearestScratchInit(gvizKNearestScratch *scratch, size_t nvertices);
void gvizKNearestScratchRelease(gvizKNearestScratch *scratch);
/* Up to k vertices of sg nearest to source by edge count, in discovery
 * order (only vertices in filter count; NULL: all). Returns the number
 * written, or -1 (invalid subgraph, source outside it, scratch too small). */
int gvizSearchKNearest(const gvizSubgraph *sg, gvizFoundVertex *out, size_t k,
                       size_t source, gvizVertexSubset filter);
int gvizSearchKNearestScratch(const gvizSubgraph *sg, gvizFoundVertex *out,
                              size_t k, size_t source, gvizVertexSubset filter,
                              gvizKNearestScratch *scratch);
/* For each target, its k nearest visible vertices (one BFS per visible seed,
 * on the device). 0, -1 on error or without a usable GPU, 1 when the batch
 * path does not apply (visible NULL or empty, more than
 * GVIZ_KNN_BATCH_VISIBLE_MAX seeds, or not fewer seeds than targets). */
int gvizSearchKNearestFromVisibleBatch(const gvizSubgraph *sg,
                                       gvizVertexSubset visible, size_t k,
                                       gvizKNNBatchTarget *targets,
                                       size_t targetCount,
                                       gvizKNearestScratch *scratch);
/* Whether the batched search is likely faster than per-vertex searches. */
int gvizSearchKNearestPreferBatch(const gvizSubgraph *sg, size_t visibleCount,
                                  size_t targetCount);
/* Counters of host queries, kept while GVIZ_KNN_PROFILE is set. */
void gvizKNNProfileReset(void);
void gvizKNNProfileSnapshot(unsigned long long *queries,
                            unsigned long long *visited,
                            unsigned long long *maxVisited);

#ifdef __cplusplus
}
#endif
#endif
