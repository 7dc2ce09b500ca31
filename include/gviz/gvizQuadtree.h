/*
 * gvizQuadtree.h -- the reference's host quadtree (include/ds/gvizQuadtree.h)
 * as a host view of a tree built on the device. gvizQuadtreeInit and
 * gvizQuadtreeRebuild send the points to the GPU, which computes the root
 * cell and every point's quadrant path by the reference's FP64 centre
 * recurrence and sorts the paths (gvizhip_qtree_build); the host lays the
 * sorted paths out as the reference's nodes (every split makes four
 * children, empty ones included; a leaf holds <= nodesPerCell points in
 * ascending index order; a cell whose half size is <= 1e-9 holds any number
 * in overflow buffers) and fills mass and centre of mass with the
 * reference's running mean in index order, so node records equal the
 * reference's bit for bit. Arena blocks of 1024 nodes persist across
 * rebuilds as in the reference; nodes are laid out parent before children
 * (the reference's arena order follows its insertion sequence instead).
 *
 * gvizForceEmbedderState.quadtree is filled on demand: gvizQuadtreeRoot on
 * it reads the tree of the embedder's last step (the positions at that
 * step's start, as the reference's Step rebuilds it) from the device, the
 * first time it is called after Begin or a step.
 */
#ifndef GVIZ_HOST_QUADTREE_H
#define GVIZ_HOST_QUADTREE_H

#include "gvizGraph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Default leaf capacity (include/ds/gvizQuadtree.h:8). */
#define GVIZ_QUADTREE_NODES_PER_CELL_DEFAULT 1

/* A child's position in its parent's square, y-axis up. */
typedef enum gvizQuadrant {
  GVIZ_QUAD_NW = 0,
  GVIZ_QUAD_NE = 1,
  GVIZ_QUAD_SW = 2,
  GVIZ_QUAD_SE = 3,
  GVIZ_QUAD_COUNT = 4,
} gvizQuadrant;

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

/* Builds a tree over count points (x, y interleaved; borrowed, as is
 * masses[count]). 0, or -1 on failure (no usable GPU, allocation, or
 * points closer than the 64 path levels the device keys hold can separate
 * while the cell half size is still above 1e-9); the tree is releasable
 * either way. */
int gvizQuadtreeInit(gvizQuadtree *tree, const double *points,
                     const double *masses, size_t count, size_t nodesPerCell);
/* Rebuilds an initialised tree over new points / masses, reusing its node
 * arena (no allocation unless the tree outgrows every earlier build). */
int gvizQuadtreeRebuild(gvizQuadtree *tree, const double *points,
                        const double *masses, size_t count);
void gvizQuadtreeRelease(gvizQuadtree *tree);
/* Root node, or NULL for a tree over zero points. */
const gvizQuadtreeNode *gvizQuadtreeRoot(const gvizQuadtree *tree);
int gvizQuadtreeNodeIsLeaf(const gvizQuadtreeNode *node);
/* The quadrant child, or NULL on a leaf. */
const gvizQuadtreeNode *gvizQuadtreeNodeChild(const gvizQuadtreeNode *node,
                                              gvizQuadrant quadrant);
/* Points held directly (0 for internal nodes). */
size_t gvizQuadtreeNodePointCount(const gvizQuadtreeNode *node);
size_t gvizQuadtreeNodePointAt(const gvizQuadtreeNode *node, size_t i);
void gvizQuadtreeNodeCenterOfMass(const gvizQuadtreeNode *node, double *outX,
                                  double *outY);

/* The reference's thread pool; this library runs no host pool (the state's
 * `pool` stays NULL). */
typedef struct gvizThreadPool gvizThreadPool;

#ifdef __cplusplus
}
#endif
#endif
