/*
 * gvizForceEmbedder.h -- the reference's force-directed embedder API
 * (include/embedders/gvizForceEmbedder.h, include/embedders/gvizForceModel.h)
 * on the gfx950 engine: same function names, argument meanings, defaults,
 * return codes, registered actions and stat series. Step and Run execute on
 * the GPU through include/gviz_hip.h; nothing here falls back to the CPU.
 */
#ifndef GVIZ_HOST_FORCE_EMBEDDER_H
#define GVIZ_HOST_FORCE_EMBEDDER_H

#include "gvizEmbeddedGraph.h"
#include "gvizQuadtree.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GVIZ_FORCE_EMBEDDER_EDGE_LENGTH_DEFAULT 10.0
#define GVIZ_FORCE_EMBEDDER_THETA_DEFAULT 1.0
#define GVIZ_FORCE_EMBEDDER_OVERLAP_CONSTANT_DEFAULT 100.0
#define GVIZ_FORCE_EMBEDDER_JITTER_TOLERANCE_DEFAULT 1.0
#define GVIZ_FORCE_EMBEDDER_SPEED_INITIAL 1.0
#define GVIZ_FORCE_EMBEDDER_SPEED_EFFICIENCY_INITIAL 1.0
#define GVIZ_FORCE_EMBEDDER_SPEED_EFFICIENCY_MIN 0.05
#define GVIZ_FORCE_EMBEDDER_SPEED_EFFICIENCY_MAX 1.0
#define GVIZ_FORCE_EMBEDDER_SPEED_MAX_RISE 0.5

typedef enum gvizForceModelKind {
  GVIZ_FORCE_MODEL_FRUCHTERMAN_REINGOLD = 0,
  GVIZ_FORCE_MODEL_LINLOG = 1,
} gvizForceModelKind;

/* Pairwise force vtable (reference: include/embedders/gvizForceModel.h).
 * The GPU engine runs the two built-in models; the table's host functions
 * give the same formulas to callers that evaluate single pairs. */
typedef struct gvizForceModel {
  double (*vertexMass)(size_t degree);
  void (*attractive)(int n, double *vPos, double *uPos, double edgeLength,
                     double *acc);
  void (*repulsive)(int n, double *vPos, double *otherPos, double vMass,
                    double otherMass, double vRadius, double otherRadius,
                    double overlapConstant, double edgeLength, double *acc);
} gvizForceModel;

const gvizForceModel *gvizForceModelGet(gvizForceModelKind kind);

/* Precision of the GPU force math (gviz_hip.h GVIZHIP_PRECISION_*). */
typedef enum gvizForcePrecision {
  GVIZ_FORCE_PRECISION_FP32 = 0,
  GVIZ_FORCE_PRECISION_FP64 = 1,
  GVIZ_FORCE_PRECISION_MIXED = 2,
} gvizForcePrecision;

/* The reference's layout (include/embedders/gvizForceEmbedder.h:42-138),
 * field for field, so callers compiled against the reference's headers read
 * the same offsets. The gfx950 engine's handle, precision and residency
 * flag live beside the state (keyed by its address), set through
 * gvizForceEmbedderSetPrecision / SetResidentPositions. The per-vertex
 * force buffers of the reference (disp, attForceMag, repForceMag,
 * structForce, oldStructForce, appliedDisp, swinging, traction) stay on the
 * device and are NULL here (gvizhip_force_evaluate reads them); the host
 * quadtree stays empty and `pool` NULL. */
typedef struct gvizForceEmbedderState {
  gvizEmbeddedGraph graph; /* MUST be first */
  size_t *vertices;        /* compact index i -> parent id (ascending) */
  size_t vertexCount;
  const gvizForceModel *model;
  double *mass;            /* vertexCount */
  size_t *degree;          /* vertexCount, subgraph degree */
  double radiusBase;
  double radiusPerDegree;
  double overlapConstant;
  int preventOverlap;
  double meanDegreePlusOne;
  double *disp;
  double *positionsScratch; /* compact positions, 2 * vertexCount */
  double *attForceMag;
  double *repForceMag;
  double *structForce;
  double *oldStructForce;
  double *appliedDisp;
  double *swinging;
  double *traction;
  double jitterTolerance;
  double globalSpeed;      /* mirrored from the device after each Step/Run */
  double speedEfficiency;
  double edgeLength;
  double boxExtent;
  double gravityK;
  size_t iteration;
  double lastMaxDisplacement;
  int begun;
  double theta;
  size_t nodesPerCell;
  int barnesHutEnabled;
  gvizQuadtree quadtree;
  int quadtreeReady;
  gvizThreadPool *pool;
} gvizForceEmbedderState;

/* Returns 0, -1 (allocation) or -2 (dimension != 2). No GPU is touched
 * until Begin. Registers "forceEmbedder.step",
 * "forceEmbedder.toggleOverlapPrevention" and the five
 * "forceEmbedder.*" stat series. */
int gvizForceEmbedderInit(gvizForceEmbedderState *state, gvizSubgraph subgraph,
                          size_t dimension, gvizForceModelKind model);
void gvizForceEmbedderRelease(gvizForceEmbedderState *state);
void gvizForceEmbedderConfigure(gvizForceEmbedderState *state,
                                double edgeLength, double boxExtent);
void gvizForceEmbedderConfigureSpeed(gvizForceEmbedderState *state,
                                     double tolerance);
void gvizForceEmbedderConfigureBarnesHut(gvizForceEmbedderState *state,
                                         double theta, size_t nodesPerCell);
void gvizForceEmbedderSetBarnesHutEnabled(gvizForceEmbedderState *state,
                                          int enabled);
void gvizForceEmbedderConfigureGravity(gvizForceEmbedderState *state,
                                       double k);
double gvizForceEmbedderVertexRadius(const gvizForceEmbedderState *state,
                                     size_t index);
void gvizForceEmbedderConfigureRadius(gvizForceEmbedderState *state,
                                      double base, double perDegree);
void gvizForceEmbedderConfigureOverlapPrevention(gvizForceEmbedderState *state,
                                                 double constant);
void gvizForceEmbedderSetPreventOverlapEnabled(gvizForceEmbedderState *state,
                                               int enabled);
/* Random positions in the box (seed 0: time-seeded), speed 1, efficiency 1,
 * forces zero; creates the device engine on first use. 0 or -1 (no GPU). */
int gvizForceEmbedderBegin(gvizForceEmbedderState *state, unsigned int seed);
/* One step; returns maxDisp, or a negative value if the device failed. */
double gvizForceEmbedderStep(gvizForceEmbedderState *state);
/* Steps until maxDisp < epsilon or maxIters; returns the steps taken or -1
 * if not begun (or on device failure). */
int gvizForceEmbedderRun(gvizForceEmbedderState *state, size_t maxIters,
                         double epsilon);

/* Engine settings with no reference counterpart. */
void gvizForceEmbedderSetPrecision(gvizForceEmbedderState *state,
                                   gvizForcePrecision precision);
/* The precision Step and Run use (GVIZ_FORCE_PRECISION_MIXED after Init). */
gvizForcePrecision gvizForceEmbedderPrecision(const gvizForceEmbedderState *state);
/* Resident mode: Step neither reads nor writes the host positions; call
 * gvizForceEmbedderSyncPositions to fetch them. */
void gvizForceEmbedderSetResidentPositions(gvizForceEmbedderState *state,
                                           int enabled);
int gvizForceEmbedderSyncPositions(gvizForceEmbedderState *state);
/* Exact binary checkpoint of a begun embedder: positions, the previous
 * step's structForce, global speed, speed efficiency, step count and last
 * maxDisp. LoadState needs Begin on the same graph first; the run then
 * continues bit for bit as if never interrupted (the reference's text
 * checkpoint, gvizEmbeddedGraphSaveEmbedding, keeps positions only, as
 * "%f"). 0 or -1. */
int gvizForceEmbedderSaveState(gvizForceEmbedderState *state, const char *filename);
int gvizForceEmbedderLoadState(gvizForceEmbedderState *state, const char *filename);
const char *gvizForceEmbedderLastError(void);

#ifdef __cplusplus
}
#endif
#endif
