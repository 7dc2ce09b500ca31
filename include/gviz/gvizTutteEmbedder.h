/*
 * gvizTutteEmbedder.h -- the reference's Tutte barycentric embedder API
 * (include/embedders/gvizTutteEmbedder.h) with the Jacobi sweep on the GPU.
 * Begin tests planarity on the host (left-right test) and fixes the largest
 * face on a convex polygon: 0, -2 for a graph that is not planar. Gauss-
 * Seidel sweeps and FixOuterFace (a face picked in a viewer) are outside
 * this engine: they return -1 (Step: a negative value).
 */
#ifndef GVIZ_HOST_TUTTE_EMBEDDER_H
#define GVIZ_HOST_TUTTE_EMBEDDER_H

#include "gvizEmbeddedGraph.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GVIZ_TUTTE_DEFAULT_EPSILON 1e-5

struct gvizhip_tutte;

typedef struct gvizTutteState {
  gvizEmbeddedGraph graph; /* MUST be first */
  size_t *boundary;
  size_t boundaryCount;
  uint8_t *isBoundary;     /* one byte per parent vertex */
  size_t iteration;
  double lastMaxDelta;
  int converged;
  int useGaussSeidel;
  double epsilon;
  double relaxationRate;
  int begun;
  int precision;           /* GVIZHIP_PRECISION_FP64 (default) or _FP32 */
  int devicePositionsFresh;/* device copy matches the host positions */
  struct gvizhip_tutte *device;
} gvizTutteState;

int gvizTutteEmbedderInit(gvizTutteState *s, gvizSubgraph subgraph,
                          size_t dimension, double epsilon);
void gvizTutteEmbedderRelease(gvizTutteState *s);
int gvizTutteEmbedderBegin(gvizTutteState *s);
int gvizTutteEmbedderSetBoundary(gvizTutteState *s, const size_t *boundary,
                                 size_t count, const double *positions);
void gvizTutteEmbedderSeedInterior(gvizTutteState *s);
void gvizTutteFixConvexPolygon(gvizTutteState *s, const size_t *boundary,
                               size_t count, double radius);
double gvizTutteEmbedderStep(gvizTutteState *s, double dt);
int gvizTutteEmbedderRun(gvizTutteState *s, size_t maxIters);
int gvizTutteEmbedderFixOuterFace(gvizTutteState *s);

#ifdef __cplusplus
}
#endif
#endif
