/*
 * gvizSpringTutteEmbedder.h -- the reference's spring-driven Tutte embedder
 * API (include/embedders/gvizSpringTutteEmbedder.h) with the step on the
 * GPU: every interior vertex accelerates towards its neighbours' barycentre,
 * a = stiffness (bary - P) - damping v, v += a dt, P += v dt, from a Jacobi
 * snapshot of the positions. Same fixed point as the Tutte sweep.
 *
 * Begin tests planarity on the host (left-right test), writes the rotation
 * system into the graph's adjacency order and fixes the largest face on a
 * convex polygon (0, or -2 for a graph that is not planar); FixOuterFace
 * pins the face walked from the highlight's first edge instead. The Jacobi
 * snapshot lives on the
 * device, so the state has no scratchPos; velocity is mirrored on the host
 * after every Step / Run. The device step is two-dimensional: Step and Run
 * fail (negative / -1) for another dimension.
 */
#ifndef GVIZ_HOST_SPRING_TUTTE_EMBEDDER_H
#define GVIZ_HOST_SPRING_TUTTE_EMBEDDER_H

#include "gvizEmbeddedGraph.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GVIZ_SPRING_TUTTE_DEFAULT_EPSILON 1e-5
#define GVIZ_SPRING_TUTTE_DEFAULT_STIFFNESS 30.0
#define GVIZ_SPRING_TUTTE_DEFAULT_DAMPING 6.0


/* The reference's layout (include/embedders/gvizSpringTutteEmbedder.h:15-30).
 * The position double buffer lives on the device (`scratchPos` stays
 * NULL); the device engine and its precision live beside the state, keyed
 * by its address. */
typedef struct gvizSpringTutteState {
  gvizEmbeddedGraph graph; /* MUST be first */
  size_t *boundary;
  size_t boundaryCount;
  GVIZ_BIT_ARRAY isBoundary; /* one bit per parent vertex */
  double *scratchPos;
  double *velocity;        /* N * dim; zero for boundary vertices */
  size_t iteration;
  double lastMaxDelta;
  int converged;
  double epsilon;          /* converged once a step's max delta < epsilon */
  double stiffness;
  double damping;
  int begun;
} gvizSpringTutteState;

/* Registers action "springTutte.step" (and "springTutte.fixOuterFace") and
 * stat series "springTutte.maxDelta". 0 or -1. */
int gvizSpringTutteEmbedderInit(gvizSpringTutteState *s, gvizSubgraph subgraph,
                                size_t dimension, double epsilon);
void gvizSpringTutteEmbedderRelease(gvizSpringTutteState *s);
/* Values <= 0 keep the current stiffness / damping. */
void gvizSpringTutteEmbedderConfigure(gvizSpringTutteState *s, double stiffness,
                                      double damping);
int gvizSpringTutteEmbedderBegin(gvizSpringTutteState *s);
int gvizSpringTutteEmbedderSetBoundary(gvizSpringTutteState *s, const size_t *boundary,
                                       size_t count, const double *polygonPositions);
void gvizSpringTutteEmbedderSeedInterior(gvizSpringTutteState *s);
void gvizSpringTutteFixConvexPolygon(gvizSpringTutteState *s, const size_t *boundary,
                                     size_t count, double radius);
/* One step of size dt; returns the max displacement (0 for dt <= 0,
 * negative if the device failed). */
double gvizSpringTutteEmbedderStep(gvizSpringTutteState *s, double dt);
/* Steps until converged or iteration == maxIters; returns the iteration
 * count, or -1 without a boundary (or on device failure). */
int gvizSpringTutteEmbedderRun(gvizSpringTutteState *s, size_t maxIters, double dt);
/* The highlighted face as the boundary (after Begin; a highlight set with
 * gvizEmbeddedGraphSetHighlight). 0, or -1 without either. */
int gvizSpringTutteEmbedderFixOuterFace(gvizSpringTutteState *s);

#ifdef __cplusplus
}
#endif
#endif
