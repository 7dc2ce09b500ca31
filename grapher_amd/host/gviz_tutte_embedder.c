/*
 * gviz_tutte_embedder.c -- gvizTutteEmbedder* with the Jacobi sweep on the
 * GPU (reference: src/embedders/gvizTutteEmbedder.c).
 *
 * Boundary placement, interior seeding and the convex polygon are host
 * bookkeeping done as the reference does them; every relaxation sweep runs
 * on the device (gvizhip_tutte_sweeps for Jacobi, gvizhip_tutte_gs_sweeps
 * when useGaussSeidel is set), which keeps the live positions and the
 * Jacobi snapshot in HBM. Positions cross the bus when the host changed
 * them (boundary, seeding) and after each Step / Run. Begin's planarity
 * test and face search run on the host (gviz_planar.c).
 */
#include "gviz/gvizTutteEmbedder.h"

#include "gviz_hip.h"
#include "gviz_side.h"
#include "gviz_planar.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>


/* The engine's extras beside the reference's state layout: the device
 * engine, its precision, whether its positions match the host's, and a
 * byte-per-vertex copy of the boundary bits (the form the device takes). */
typedef struct tutte_side {
  gvizhip_tutte *device;
  int precision; /* GVIZHIP_PRECISION_FP64 (default) or _FP32 */
  int fresh;
  uint8_t *bytes;
} tutte_side;

static tutte_side *side_of(const gvizTutteState *s) {
  return (tutte_side *)gviz_side_get(s);
}

static void mark_stale(gvizTutteState *s) {
  tutte_side *x = side_of(s);
  if (x)
    x->fresh = 0;
}

static const uint8_t *boundary_bytes(gvizTutteState *s) {
  tutte_side *x = side_of(s);
  gvizVertexSubsetToBytes(s->isBoundary,
                          gvizGraphSize(((gvizEmbeddedGraph *)s)->subgraph.g), x->bytes);
  return x->bytes;
}

static int attach_side(gvizTutteState *s, size_t n) {
  tutte_side *x = calloc(1, sizeof(tutte_side)), *prev = NULL;
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

static void detach_side(gvizTutteState *s) {
  tutte_side *x = (tutte_side *)gviz_side_take(s);
  if (!x)
    return;
  if (x->device)
    gvizhip_tutte_destroy(x->device);
  free(x->bytes);
  free(x);
}

static size_t num_vertices(const gvizTutteState *s) {
  return gvizGraphSize(((const gvizEmbeddedGraph *)s)->subgraph.g);
}

static size_t dim_of(const gvizTutteState *s) {
  return ((const gvizEmbeddedGraph *)s)->embedding.dim;
}

static void action_step(gvizEmbeddedGraph *embedding, void *userData,
                        const gvizActionPayload *payload) {
  (void)userData;
  gvizTutteState *s = (gvizTutteState *)embedding;
  if (!s->begun || !s->boundary)
    return;
  double dt = payload && payload->deltaTime > 0.0 ? payload->deltaTime : 1.0 / 60.0;
  gvizTutteEmbedderStep(s, dt);
}

static void action_fix_outer_face(gvizEmbeddedGraph *embedding, void *userData,
                                  const gvizActionPayload *payload) {
  (void)userData;
  (void)payload;
  gvizTutteEmbedderFixOuterFace((gvizTutteState *)embedding);
}

/* src/embedders/gvizTutteEmbedder.c:55-103 */
int gvizTutteEmbedderInit(gvizTutteState *s, gvizSubgraph subgraph,
                          size_t dimension, double epsilon) {
  memset(s, 0, sizeof(*s));
  int res = gvizEmbeddedGraphInit((gvizEmbeddedGraph *)s, subgraph, dimension);
  if (res < 0)
    return res;
  size_t n = gvizGraphSize(subgraph.g);
  s->isBoundary = gvizVertexSubsetCreateEmpty(subgraph.g);
  if (!s->isBoundary || attach_side(s, n) != 0)
    return -1;
  s->epsilon = epsilon > 0.0 ? epsilon : GVIZ_TUTTE_DEFAULT_EPSILON;
  s->relaxationRate = 5.0;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  if (gvizEmbeddedGraphAddAction(e, "tutte.step", action_step, NULL) < 0)
    return -1;
  if (gvizEmbeddedGraphAddAction(e, "tutte.fixOuterFace", action_fix_outer_face,
                                 NULL) < 0)
    return -1;
  if (!gvizEmbeddedGraphAddStatSeries(e, "tutte.maxDelta", GVIZ_STAT_CHART_LINE_LOG))
    return -1;
  return 0;
}

void gvizTutteEmbedderRelease(gvizTutteState *s) {
  if (!s)
    return;
  detach_side(s);
  free(s->boundary);
  gvizVertexSubsetRelease(s->isBoundary);
  free(s->scratch);
  s->boundary = NULL;
  s->isBoundary = NULL;
  s->scratch = NULL;
  gvizEmbeddedGraphRelease((gvizEmbeddedGraph *)s);
}

/* src/embedders/gvizTutteEmbedder.c:145-173: planarity test and rotation system
 * (gviz_planar.c), the rotation written into the graph's adjacency order
 * as the reference's planarity step writes its own (planarEmbedded set),
 * then the largest face fixed on a convex polygon of radius 200 and the
 * interior seeded at its centroid. -2 for a graph that is not planar, -1 for
 * another dimension than 2 or a failure. The device engine is rebuilt from
 * the new adjacency order at the next step. */
int gvizTutteEmbedderBegin(gvizTutteState *s) {
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
  tutte_side *x = side_of(s);
  if (x && x->device) {
    gvizhip_tutte_destroy(x->device);
    x->device = NULL;
  }
  gvizTutteFixConvexPolygon(s, boundary, count, 200.0);
  free(boundary);
  gvizTutteEmbedderSeedInterior(s);
  s->begun = 1;
  return 0;
}

/* src/embedders/gvizTutteEmbedder.c:347-373: the face walked from the highlight's first
 * edge through the rotation Begin installed, fixed on a convex polygon of
 * radius 200 (interior positions kept); -1 without a planar embedding or a
 * highlight. */
int gvizTutteEmbedderFixOuterFace(gvizTutteState *s) {
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
  gvizTutteFixConvexPolygon(s, boundary, count, 200.0);
  free(boundary);
  s->iteration = 0;
  s->lastMaxDelta = 0.0;
  s->converged = 0;
  s->begun = 1;
  return 0;
}

/* src/embedders/gvizTutteEmbedder.c:178-205 */
int gvizTutteEmbedderSetBoundary(gvizTutteState *s, const size_t *boundary,
                                 size_t count, const double *positions) {
  if (count < 3)
    return -1;
  size_t n = num_vertices(s), d = dim_of(s);
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
  }
  mark_stale(s);
  return 0;
}

/* src/embedders/gvizTutteEmbedder.c:207-230: every non-boundary vertex at
 * the boundary centroid. */
void gvizTutteEmbedderSeedInterior(gvizTutteState *s) {
  size_t n = num_vertices(s), d = dim_of(s);
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
  s->iteration = 0;
  s->lastMaxDelta = 0.0;
  s->converged = 0;
  mark_stale(s);
}

/* src/embedders/gvizTutteEmbedder.c:232-249 */
void gvizTutteFixConvexPolygon(gvizTutteState *s, const size_t *boundary,
                               size_t count, double radius) {
  size_t d = dim_of(s);
  double *positions = calloc(count > 0 ? count * d : 1, sizeof(double));
  if (!positions)
    return;
  for (size_t k = 0; k < count; k++) {
    double angle = 2.0 * M_PI * (double)k / (double)count;
    positions[k * d] = radius * cos(angle);
    if (d > 1)
      positions[k * d + 1] = radius * sin(angle);
  }
  gvizTutteEmbedderSetBoundary(s, boundary, count, positions);
  free(positions);
}

/* Device engine over parent ids with the subgraph's edges, created on the
 * first sweep; positions re-uploaded whenever the host changed them. */
static int ensure_device(gvizTutteState *s) {
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  if (dim_of(s) != 2)
    return -1;
  tutte_side *x = side_of(s);
  if (!x)
    return -1;
  if (!x->device) {
    int64_t *off = NULL, *adj = NULL;
    if (gvizSubgraphToCSR(&e->subgraph, &off, &adj) != 0)
      return -1;
    int rc = gvizhip_tutte_create(&x->device, (int64_t)num_vertices(s), off,
                                  adj, x->precision);
    free(off);
    free(adj);
    if (rc != 0) {
      x->device = NULL;
      return -1;
    }
    x->fresh = 0;
  }
  if (!x->fresh) {
    if (gvizhip_tutte_upload(x->device, e->embedding.vertexPositions,
                             boundary_bytes(s)) != 0)
      return -1;
    x->fresh = 1;
  }
  return 0;
}

static void record_sweep(gvizTutteState *s, double maxDelta) {
  s->iteration++;
  s->lastMaxDelta = maxDelta;
  if (maxDelta < s->epsilon)
    s->converged = 1;
  gvizEmbeddedGraphStatAppend((gvizEmbeddedGraph *)s, "tutte.maxDelta", maxDelta);
}

/* src/embedders/gvizTutteEmbedder.c:305-335 (Jacobi, or Gauss-Seidel with
 * useGaussSeidel set). Host positions are
 * re-uploaded before the sweep (a caller may have moved vertices) and read
 * back after it. */
double gvizTutteEmbedderStep(gvizTutteState *s, double dt) {
  double alpha = s->relaxationRate * dt;
  if (alpha <= 0.0)
    return 0.0;
  if (alpha > 1.0)
    alpha = 1.0;
  mark_stale(s);
  if (ensure_device(s) != 0)
    return -1.0;
  double maxDelta = 0.0;
  gvizhip_tutte *d = side_of(s)->device;
  if ((s->useGaussSeidel ? gvizhip_tutte_gs_sweeps(d, 1, alpha, -1.0, &maxDelta)
                         : gvizhip_tutte_sweeps(d, 1, alpha, -1.0, &maxDelta)) != 1)
    return -1.0;
  if (gvizhip_tutte_download(side_of(s)->device, ((gvizEmbeddedGraph *)s)->embedding.vertexPositions) != 0)
    return -1.0;
  record_sweep(s, maxDelta);
  return maxDelta;
}

/* src/embedders/gvizTutteEmbedder.c:337-345: sweeps with dt = 1 until the
 * max delta drops below epsilon or the iteration count reaches maxIters; the
 * loop and the stop test run on the device. Returns the iteration count. */
int gvizTutteEmbedderRun(gvizTutteState *s, size_t maxIters) {
  if (!s || !s->boundary)
    return -1;
  if (s->converged || s->iteration >= maxIters)
    return (int)s->iteration;
  double alpha = s->relaxationRate;
  if (alpha <= 0.0)
    return (int)s->iteration; /* the reference's loop never advances here */
  if (alpha > 1.0)
    alpha = 1.0;
  mark_stale(s);
  if (ensure_device(s) != 0)
    return -1;
  enum { kChunk = 65536 };
  double *deltas = malloc(sizeof(double) * kChunk);
  if (!deltas)
    return -1;
  int failed = 0;
  while (!s->converged && s->iteration < maxIters) {
    size_t left = maxIters - s->iteration;
    int64_t want = (int64_t)(left < kChunk ? left : kChunk);
    gvizhip_tutte *d = side_of(s)->device;
    int64_t ran = s->useGaussSeidel
                      ? gvizhip_tutte_gs_sweeps(d, want, alpha, s->epsilon, deltas)
                      : gvizhip_tutte_sweeps(d, want, alpha, s->epsilon, deltas);
    if (ran <= 0) {
      failed = 1;
      break;
    }
    for (int64_t i = 0; i < ran; i++)
      record_sweep(s, deltas[i]);
  }
  free(deltas);
  if (gvizhip_tutte_download(side_of(s)->device, ((gvizEmbeddedGraph *)s)->embedding.vertexPositions) != 0)
    failed = 1;
  return failed ? -1 : (int)s->iteration;
}
