/*
 * gvizGRIPEmbedder.h -- the reference's GRIP multilevel embedder API
 * (include/embedders/gvizGRIPEmbedder.h, src/embedders/gvizGRIPInternal.h)
 * with the MIS filtration, kNN searches, placement and refinement rounds on
 * the GPU. Dimensions 2, 3 and 4; Init returns -2 for another.
 */
#ifndef GVIZ_HOST_GRIP_EMBEDDER_H
#define GVIZ_HOST_GRIP_EMBEDDER_H

#include "gvizEmbeddedGraph.h"
#include "gvizKNearest.h"
#include "gvizQuadtree.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gvizGRIPKPolicy {
  GVIZ_GRIP_K_CONSTANT = 0,
  GVIZ_GRIP_K_LAYER_DECAY = 1,
  GVIZ_GRIP_K_LAYER_GROW = 2,
  GVIZ_GRIP_K_PLACEMENT_DECAY = 3,
  GVIZ_GRIP_K_BUDGET = 4,
} gvizGRIPKPolicy;

typedef struct gvizGRIPRoundStats {
  double maxDisplacement;
  double meanDisplacement;
  double meanForce;
} gvizGRIPRoundStats;

typedef struct gvizGRIPDecorators {
  gvizArray knn;
  double *disp;
  double *oldDisp;
  double oldCos;
  double heat;
  double lastForceMag;
} gvizGRIPDecorators;

/* The reference's layout (include/embedders/gvizGRIPEmbedder.h:50-78),
 * field for field. The per-vertex decorators, the computed-this-round bits,
 * the BFS scratch and the kNN scratch live on the device (dec,
 * dispCalculated, radiusBfsStamp, radiusBfsQueue, knnScratch and pool stay
 * NULL / empty here; gvizGRIPEmbedderSaveState writes them out); the
 * device handle and residency flag live beside the state, keyed by its
 * address. */
typedef struct gvizGRIPState {
  gvizEmbeddedGraph graph; /* MUST be first */
  size_t layerCount;
  size_t currLayer;
  size_t currRound;
  size_t *misFiltration;   /* activeCount entries */
  gvizArray misBorder;     /* size_t, layerCount entries */
  gvizGRIPDecorators *dec;
  GVIZ_BIT_ARRAY dispCalculated;
  size_t *radiusBfsStamp;
  size_t radiusBfsEpoch;
  gvizDeque radiusBfsQueue;
  gvizKNearestScratch *knnScratch;
  size_t knnScratchCount;
  gvizThreadPool *pool;
  size_t placementKMax;
  size_t refinementKMax;
  size_t knnCapacity;
  gvizGRIPKPolicy kPolicy;
  gvizGRIPRoundStats lastRoundStats;
  uint8_t statsRegisterMode; /* survives Init (pre-Init configuration) */
  bool statsEnabled;
} gvizGRIPState;

int gvizGRIPEmbedderInit(gvizGRIPState *state, gvizSubgraph subgraph,
                         size_t diameter, size_t dimension);
void gvizGRIPEmbedderConfigureKnnCapacity(gvizGRIPState *state,
                                          size_t knnCapacity);
void gvizGRIPEmbedderConfigureK(gvizGRIPState *state, size_t placementKMax,
                                size_t refinementKMax, gvizGRIPKPolicy policy);
void gvizGRIPEmbedderConfigureStats(gvizGRIPState *state, bool enable);
gvizGRIPRoundStats gvizGRIPEmbedderLastRoundStats(const gvizGRIPState *state);
int gvizGRIPEmbedderEmbed(gvizGRIPState *state);
void gvizGRIPEmbedderRelease(gvizGRIPState *state);
void gvizGRIPEmbedderBegin(gvizGRIPState *state);
void gvizGRIPEmbedderNextStage(gvizGRIPState *state);
void gvizGRIPEmbedderRefineRound(gvizGRIPState *state);
/* Positions computed on the device are copied to the host positions. */
int gvizGRIPEmbedderSyncPositions(gvizGRIPState *state);
/* Resident mode (no reference counterpart): rounds and stages keep the
 * positions on the device; gvizGRIPEmbedderSyncPositions fetches them. */
void gvizGRIPEmbedderSetResidentPositions(gvizGRIPState *state, int enabled);
/* Exact binary checkpoint (no reference counterpart): layer, round, last
 * round stats, draw mask, positions, every vertex's displacement / heat /
 * cosine / kNN state and the kNN lists. LoadState needs Begin on the same
 * graph and configuration first (the filtration is checked); NextStage and
 * RefineRound then continue bit for bit. 0 or -1. */
int gvizGRIPEmbedderSaveState(gvizGRIPState *state, const char *filename);
int gvizGRIPEmbedderLoadState(gvizGRIPState *state, const char *filename);

#ifdef __cplusplus
}
#endif
#endif
