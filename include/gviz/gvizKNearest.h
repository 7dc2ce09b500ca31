/*
 * gvizKNearest.h -- the kNN result and scratch types of the reference
 * (include/algorithms/search/gvizKNearest.h:15-66), present so
 * gvizGRIPState has the reference's layout. The searches themselves run on
 * the device (gvizhip_grip_knn / gvizhip_grip_knn_batch in gviz_hip.h); no
 * gvizSearchKNearest* host functions are exported.
 */
#ifndef GVIZ_HOST_KNEAREST_H
#define GVIZ_HOST_KNEAREST_H

#include "gvizGraph.h"

#ifdef __cplusplus
extern "C" {
#endif

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

#ifdef __cplusplus
}
#endif
#endif
