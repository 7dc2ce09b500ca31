/*
 * gvizTutteEmbedder.h -- the reference's Tutte barycentric embedder API
 * (include/embedders/gvizTutteEmbedder.h) with the Jacobi and Gauss-Seidel
 * sweeps on the GPU. Begin tests planarity on the host (left-right test),
 * writes the rotation system into the graph's adjacency order (as the
 * reference's planarity step does; planarEmbedded is set) and fixes the
 * largest face on a convex polygon: 0, -2 for a graph that is not planar.
 * FixOuterFace pins the face walked from the highlight's first edge.
 */
#ifndef GVIZ_HOST_TUTTE_EMBEDDER_H
#define GVIZ_HOST_TUTTE_EMBEDDER_H

#include "gvizEmbeddedGraph.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GVIZ_TUTTE_DEFAULT_EPSILON 1e-5


/* The reference's layout (include/embedders/gvizTutteEmbedder.h:22-35). The
 * sweep's double buffer lives on the device (`scratch` stays NULL); the
 * device engine and its precision live beside the state, keyed by its
 * address. */
typedef struct gvizTutteState {
  gvizEmbeddedGraph graph; /* MUST be first */
  size_t *boundary;
  size_t boundaryCount;
  GVIZ_BIT_ARRAY isBoundary; /* one bit per parent vertex */
  double *scratch;
  size_t iteration;
  double lastMaxDelta;
  int converged;
  int useGaussSeidel;
  double epsilon;
  double relaxationRate;
  int begun;
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
/* The highlighted face as the boundary (after Begin; a highlight set with
 * gvizEmbeddedGraphSetHighlight). 0, or -1 without either. */
int gvizTutteEmbedderFixOuterFace(gvizTutteState *s);

#ifdef __cplusplus
}
#endif
#endif
