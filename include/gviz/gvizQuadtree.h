/*
 * gvizQuadtree.h -- the host quadtree types of the reference
 * (include/ds/gvizQuadtree.h:20-80), present so gvizForceEmbedderState has
 * the reference's layout. This library builds its trees on the device and
 * exports no gvizQuadtree* functions: the state's `quadtree` stays empty
 * (root NULL, quadtreeReady 0). gvizhip_force_quadtree_leaves and
 * gvizhip_force_quadtree_cells (gviz_hip.h) read the device tree.
 */
#ifndef GVIZ_HOST_QUADTREE_H
#define GVIZ_HOST_QUADTREE_H

#include "gvizGraph.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GVIZ_QUAD_COUNT 4

typedef struct gvizQuadtreeNode {
  double cx, cy;
  double halfSize;
  double comX, comY;
  double mass;
  size_t *pointIndices;
  size_t pointCount;
  size_t pointCapacity;
  struct gvizQuadtreeNode *children[GVIZ_QUAD_COUNT];
} gvizQuadtreeNode;

typedef struct gvizQuadtree {
  gvizQuadtreeNode *root;
  size_t nodesPerCell;
  const double *points;
  const double *masses;
  size_t pointCount;
  gvizArray nodeBlocks;
  gvizArray pointBlocks;
  size_t nodesUsed;
  gvizArray overflowBuffers;
} gvizQuadtree;

/* The reference's thread pool; this library runs no host pool (the state's
 * `pool` stays NULL). */
typedef struct gvizThreadPool gvizThreadPool;

#ifdef __cplusplus
}
#endif
#endif
