/*
 * gvizGRIPEmbedder.h -- the reference's GRIP multilevel embedder API
 * (include/embedders/gvizGRIPEmbedder.h, src/embedders/gvizGRIPInternal.h)
 * with the MIS filtration, kNN searches, placement and refinement rounds on
 * the GPU. Dimensions 2, 3 and 4; Init returns -2 for another.
 */
#ifndef GVIZ_HOST_GRIP_EMBEDDER_H
#define GVIZ_HOST_GRIP_EMBEDDER_H

#include "gvizEmbeddedGraph.h"

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

struct gvizhip_grip;

typedef struct gvizGRIPState {
  gvizEmbeddedGraph graph; /* MUST be first */
  size_t layerCount;
  size_t currLayer;
  size_t currRound;
  size_t *misFiltration;   /* activeCount entries */
  size_t *misBorder;       /* layerCount entries */
  size_t misBorderCount;
  size_t placementKMax;
  size_t refinementKMax;
  size_t knnCapacity;
  gvizGRIPKPolicy kPolicy;
  gvizGRIPRoundStats lastRoundStats;
  uint8_t statsRegisterMode; /* survives Init (pre-Init configuration) */
  bool statsEnabled;
  int residentPositions;     /* 1: RefineRound keeps positions on the device */
  int devicePositionsFresh;
  struct gvizhip_grip *device;
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
