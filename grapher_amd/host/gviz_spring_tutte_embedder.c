/*
 * gviz_spring_tutte_embedder.c -- gvizSpringTutteEmbedder* with the step on
 * the GPU (reference: src/embedders/gvizSpringTutteEmbedder.c).
 *
 * Boundary placement, seeding and the convex polygon are host bookkeeping
 * done as the reference does them; every step runs on the device
 * (gvizhip_tutte_spring_steps), which keeps positions, the Jacobi snapshot
 * and the velocities in HBM. Positions and velocities are uploaded when the
 * host changed them and read back after each Step / Run.
 */
#include "gviz/gvizSpringTutteEmbedder.h"

#include "gviz_hip.h"
#include "gviz_side.h"
#include "gviz_planar.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>


/* The engine's extras beside the reference's state layout: the device
 * engine, its precision, whether its positions match the host's, and a
 * byte-per-vertex copy of the boundary bits (the form the device takes). */
typedef struct spring_side {
  gvizhip_tutte *device;
  int precision; /* GVIZHIP_PRECISION_FP64 (default) or _FP32 */
  int fresh;
  uint8_t *bytes;
} spring_side;

static spring_side *side_of(const gvizSpringTutteState *s) {
  return (spring_side *)gviz_side_get(s);
}

static void mark_stale(gvizSpringTutteState *s) {
  spring_side *x = side_of(s);
  if (x)
    x->fresh = 0;
}

static const uint8_t *boundary_bytes(gvizSpringTutteState *s) {
  spring_side *x = side_of(s);
  gvizVertexSubsetToBytes(s->isBoundary,
                          gvizGraphSize(((gvizEmbeddedGraph *)s)->subgraph.g), x->bytes);
  return x->bytes;
}

static int attach_side(gvizSpringTutteState *s, size_t n) {
  spring_side *x = calloc(1, sizeof(spring_side)), *prev = NULL;
  if (!x || gviz_side_put(s, x, (void **)&prev) != 0) {
    free(x);
    return -1;
  }
  if (prev) { /* a state initialised again without Release */
    if (prev->device)
      gvizhip_tutte_destroy(prev->device);
    free(prev->bytes);
    free(prev);
  }
  x->precision = GVIZHIP_PRECISION_FP64;
  x->bytes = malloc(n ? n : 1);
  return x->bytes ? 0 : -1;
}

static void detach_side(gvizSpringTutteState *s) {
  spring_side *x = (spring_side *)gviz_side_take(s);
  if (!x)
    return;
  if (x->device)
    gvizhip_tutte_destroy(x->device);
  free(x->bytes);
  free(x);
}

static size_t num_vertices(const gvizSpringTutteState *s) {
  return gvizGraphSize(((const gvizEmbeddedGraph *)s)->subgraph.g);
}

static size_t dim_of(const gvizSpringTutteState *s) {
  return ((const gvizEmbeddedGraph *)s)->embedding.dim;
}

static void action_step(gvizEmbeddedGraph *embedding, void *userData,
                        const gvizActionPayload *payload) {
  (void)userData;
  gvizSpringTutteState *s = (gvizSpringTutteState *)embedding;
  if (!s->begun || !s->boundary)
    return;
  double dt = payload && payload->deltaTime > 0.0 ? payload->deltaTime : 1.0 / 60.0;
  gvizSpringTutteEmbedderStep(s, dt);
}

static void action_fix_outer_face(gvizEmbeddedGraph *embedding, void *userData,
                                  const gvizActionPayload *payload) {
  (void)userData;
  (void)payload;
  gvizSpringTutteEmbedderFixOuterFace((gvizSpringTutteState *)embedding);
}

/* src/embedders/gvizSpringTutteEmbedder.c:56-104 */
int gvizSpringTutteEmbedderInit(gvizSpringTutteState *s, gvizSubgraph subgraph,
                                size_t dimension, double epsilon) {
  memset(s, 0, sizeof(*s));
  int res = gvizEmbeddedGraphInit((gvizEmbeddedGraph *)s, subgraph, dimension);
  if (res < 0)
    return res;
  const size_t n = gvizGraphSize(subgraph.g);
  s->isBoundary = gvizVertexSubsetCreateEmpty(subgraph.g);
  s->velocity = calloc(n * dimension > 0 ? n * dimension : 1, sizeof(double));
  if (!s->isBoundary || !s->velocity || attach_side(s, n) != 0)
    return -1;
  s->epsilon = epsilon > 0.0 ? epsilon : GVIZ_SPRING_TUTTE_DEFAULT_EPSILON;
  s->stiffness = GVIZ_SPRING_TUTTE_DEFAULT_STIFFNESS;
  s->damping = GVIZ_SPRING_TUTTE_DEFAULT_DAMPING;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  if (gvizEmbeddedGraphAddAction(e, "springTutte.step", action_step, NULL) < 0)
    return -1;
  if (gvizEmbeddedGraphAddAction(e, "springTutte.fixOuterFace", action_fix_outer_face,
                                 NULL) < 0)
    return -1;
  if (!gvizEmbeddedGraphAddStatSeries(e, "springTutte.maxDelta", GVIZ_STAT_CHART_LINE_LOG))
    return -1;
  return 0;
}

void gvizSpringTutteEmbedderRelease(gvizSpringTutteState *s) {
  if (!s)
    return;
  detach_side(s);
  free(s->boundary);
  gvizVertexSubsetRelease(s->isBoundary);
  free(s->velocity);
  free(s->scratchPos);
  s->boundary = NULL;
  s->isBoundary = NULL;
  s->velocity = NULL;
  s->scratchPos = NULL;
  gvizEmbeddedGraphRelease((gvizEmbeddedGraph *)s);
}

/* src/embedders/gvizSpringTutteEmbedder.c:118-124 */
void gvizSpringTutteEmbedderConfigure(gvizSpringTutteState *s, double stiffness,
                                      double damping) {
  if (stiffness > 0.0)
    s->stiffness = stiffness;
  if (damping > 0.0)
    s->damping = damping;
}

/* src/embedders/gvizSpringTutteEmbedder.c:164-191: planarity test and rotation system
 * (gviz_planar.c), the rotation written into the graph's adjacency order
 * as the reference's planarity step writes its own (planarEmbedded set),
 * then the largest face fixed on a convex polygon of radius 200 and the
 * interior seeded at its centroid. -2 for a graph that is not planar, -1 for
 * another dimension than 2 or a failure. The device engine is rebuilt from
 * the new adjacency order at the next step. */
int gvizSpringTutteEmbedderBegin(gvizSpringTutteState *s) {
  if (!s || dim_of(s) != 2)
    return -1;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  const size_t N = num_vertices(s);
  size_t *boundary = malloc(sizeof(size_t) * (N ? N : 1));
  size_t count = 0;
  int rc = boundary ? gviz_planar_embed_largest_face(e, boundary, N, &count) : -1;
  if (rc != 0) {
    free(boundary);
    return rc == -2 ? -2 : -1;
  }
  spring_side *x = side_of(s);
  if (x && x->device) {
    gvizhip_tutte_destroy(x->device);
    x->device = NULL;
  }
  gvizSpringTutteFixConvexPolygon(s, boundary, count, 200.0);
  free(boundary);
  gvizSpringTutteEmbedderSeedInterior(s);
  s->begun = 1;
  return 0;
}

/* src/embedders/gvizSpringTutteEmbedder.c:371-397: the face walked from the highlight's first
 * edge through the rotation Begin installed, fixed on a convex polygon of
 * radius 200 (interior positions kept); -1 without a planar embedding or a
 * highlight. */
int gvizSpringTutteEmbedderFixOuterFace(gvizSpringTutteState *s) {
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  if (!s || !gvizEmbeddedGraphIsPlanarEmbedded(e) || !gvizEmbeddedGraphHasHighlight(e))
    return -1;
  const size_t N = num_vertices(s);
  size_t *boundary = malloc(sizeof(size_t) * (N ? N : 1));
  size_t count = 0;
  if (!boundary || gviz_planar_face_from_highlight(e, boundary, N, &count) != 0) {
    free(boundary);
    return -1;
  }
  gvizSpringTutteFixConvexPolygon(s, boundary, count, 200.0);
  free(boundary);
  s->iteration = 0;
  s->lastMaxDelta = 0.0;
  s->converged = 0;
  s->begun = 1;
  return 0;
}

/* src/embedders/gvizSpringTutteEmbedder.c:193-223: pins the vertices at the
 * given positions and zeroes their velocity. */
int gvizSpringTutteEmbedderSetBoundary(gvizSpringTutteState *s, const size_t *boundary,
                                       size_t count, const double *positions) {
  if (count < 3)
    return -1;
  const size_t n = num_vertices(s), d = dim_of(s);
  for (size_t i = 0; i < count; i++)
    if (boundary[i] >= n)
      return -1;
  gvizVertexSubsetClearAll(s->isBoundary, n);
  free(s->boundary);
  s->boundary = malloc(sizeof(size_t) * count);
  if (!s->boundary) {
    s->boundaryCount = 0;
    return -1;
  }
  memcpy(s->boundary, boundary, sizeof(size_t) * count);
  s->boundaryCount = count;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  for (size_t i = 0; i < count; i++) {
    gvizSetBit(s->isBoundary, boundary[i]);
    gvizEmbeddedGraphSetVPosition(e, boundary[i], positions + i * d);
    memset(s->velocity + boundary[i] * d, 0, sizeof(double) * d);
  }
  mark_stale(s);
  return 0;
}

/* src/embedders/gvizSpringTutteEmbedder.c:225-250: interior vertices at the
 * boundary centroid, every velocity zero, counters reset. */
void gvizSpringTutteEmbedderSeedInterior(gvizSpringTutteState *s) {
  const size_t n = num_vertices(s), d = dim_of(s);
  double *centroid = calloc(d, sizeof(double));
  if (!centroid)
    return;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  for (size_t i = 0; i < s->boundaryCount; i++) {
    const double *p = gvizEmbeddedGraphGetVPosition(e, s->boundary[i]);
    for (size_t k = 0; k < d; k++)
      centroid[k] += p[k];
  }
  if (s->boundaryCount > 0)
    for (size_t k = 0; k < d; k++)
      centroid[k] /= (double)s->boundaryCount;
  for (size_t u = 0; u < n; u++)
    if (!gvizTestBit(s->isBoundary, u))
      gvizEmbeddedGraphSetVPosition(e, u, centroid);
  free(centroid);
  memset(s->velocity, 0, sizeof(double) * n * d);
  s->iteration = 0;
  s->lastMaxDelta = 0.0;
  s->converged = 0;
  mark_stale(s);
}

/* src/embedders/gvizSpringTutteEmbedder.c:252-270 */
void gvizSpringTutteFixConvexPolygon(gvizSpringTutteState *s, const size_t *boundary,
                                     size_t count, double radius) {
  const size_t d = dim_of(s);
  double *positions = calloc(count > 0 ? count * d : 1, sizeof(double));
  if (!positions)
    return;
  for (size_t k = 0; k < count; k++) {
    double angle = 2.0 * M_PI * (double)k / (double)count;
    positions[k * d] = radius * cos(angle);
    if (d > 1)
      positions[k * d + 1] = radius * sin(angle);
  }
  gvizSpringTutteEmbedderSetBoundary(s, boundary, count, positions);
  free(positions);
}

/* Device engine over parent ids with the subgraph's edges, created on the
 * first step; positions and velocities re-uploaded when the host changed
 * them. */
static int ensure_device(gvizSpringTutteState *s) {
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  if (dim_of(s) != 2)
    return -1;
  spring_side *x = side_of(s);
  if (!x)
    return -1;
  if (!x->device) {
    int64_t *off = NULL, *adj = NULL;
    if (gvizSubgraphToCSR(&e->subgraph, &off, &adj) != 0)
      return -1;
    int rc = gvizhip_tutte_create(&x->device, (int64_t)num_vertices(s), off, adj,
                                  x->precision);
    free(off);
    free(adj);
    if (rc != 0) {
      x->device = NULL;
      return -1;
    }
    x->fresh = 0;
  }
  if (!x->fresh) {
    if (gvizhip_tutte_upload(x->device, e->embedding.vertexPositions, boundary_bytes(s)) != 0 ||
        gvizhip_tutte_set_velocity(x->device, s->velocity) != 0)
      return -1;
    x->fresh = 1;
  }
  return 0;
}

static int read_back(gvizSpringTutteState *s) {
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  if (gvizhip_tutte_download(side_of(s)->device, e->embedding.vertexPositions) != 0 ||
      gvizhip_tutte_get_velocity(side_of(s)->device, s->velocity) != 0)
    return -1;
  return 0;
}

static void record_step(gvizSpringTutteState *s, double maxDelta) {
  s->iteration++;
  s->lastMaxDelta = maxDelta;
  if (maxDelta < s->epsilon)
    s->converged = 1;
  gvizEmbeddedGraphStatAppend((gvizEmbeddedGraph *)s, "springTutte.maxDelta", maxDelta);
}

/* src/embedders/gvizSpringTutteEmbedder.c:333-358. Host positions are
 * uploaded before the step (a caller may have moved vertices) and read back
 * after it, with the velocities. */
double gvizSpringTutteEmbedderStep(gvizSpringTutteState *s, double dt) {
  if (dt <= 0.0)
    return 0.0;
  mark_stale(s);
  if (ensure_device(s) != 0)
    return -1.0;
  double maxDelta = 0.0;
  if (gvizhip_tutte_spring_steps(side_of(s)->device, 1, dt, s->stiffness, s->damping, -1.0,
                                 &maxDelta) != 1)
    return -1.0;
  if (read_back(s) != 0)
    return -1.0;
  record_step(s, maxDelta);
  return maxDelta;
}

/* src/embedders/gvizSpringTutteEmbedder.c:360-369: steps until converged or
 * the iteration count reaches maxIters, the loop and the stop test on the
 * device. */
int gvizSpringTutteEmbedderRun(gvizSpringTutteState *s, size_t maxIters, double dt) {
  if (!s || !s->boundary)
    return -1;
  if (s->converged || s->iteration >= maxIters)
    return (int)s->iteration;
  if (dt <= 0.0) {
    /* the reference's Step returns at once and its loop never ends; here
     * the call returns without stepping */
    return (int)s->iteration;
  }
  mark_stale(s);
  if (ensure_device(s) != 0)
    return -1;
  enum { kChunk = 65536 };
  double *deltas = malloc(sizeof(double) * kChunk);
  if (!deltas)
    return -1;
  int failed = 0;
  while (!s->converged && s->iteration < maxIters) {
    const size_t left = maxIters - s->iteration;
    const int64_t want = (int64_t)(left < kChunk ? left : kChunk);
    const int64_t ran = gvizhip_tutte_spring_steps(side_of(s)->device, want, dt, s->stiffness,
                                                   s->damping, s->epsilon, deltas);
    if (ran <= 0) {
      failed = 1;
      break;
    }
    for (int64_t i = 0; i < ran; i++)
      record_step(s, deltas[i]);
  }
  free(deltas);
  if (read_back(s) != 0)
    failed = 1;
  return failed ? -1 : (int)s->iteration;
}
