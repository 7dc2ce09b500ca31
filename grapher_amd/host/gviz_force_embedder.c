/*
 * gviz_force_embedder.c -- gvizForceEmbedder* on the gfx950 engine.
 *
 * The state keeps the reference's configuration fields and defaults
 * (src/embedders/gvizForceEmbedder.c:219-325, 357-410); Begin chooses the
 * initial positions on the host exactly as the reference does (rand_r over
 * the subgraph in ascending vertex id, :412-443) and hands them to the
 * device; Step and Run execute on the GPU (gvizhip_force_step / _run) and
 * append the same five stat series per step (:445-510). There is no CPU
 * path: without a GPU, Begin fails and Step returns a negative value.
 */
#include "gviz/gvizForceEmbedder.h"

#include "gviz_hip.h"
#include "gviz_quadtree_view.h"
#include "gviz_side.h"

#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------ force model table */

/* Direction for coincident points: a hash of the two coordinate buffers'
 * addresses mapped to an angle (include/core/gvizVec.h:385-394). */
static void coincident_dir(const double *a, const double *b, double *ux,
                           double *uy) {
  uint64_t bits = (uint64_t)(uintptr_t)a ^
                  ((uint64_t)(uintptr_t)b * 0x9E3779B97F4A7C15ULL);
  bits ^= bits >> 33;
  bits *= 0xff51afd7ed558ccdULL;
  bits ^= bits >> 33;
  double angle = (double)(bits % 1000000u) / 1000000.0 * 2.0 * M_PI;
  *ux = cos(angle);
  *uy = sin(angle);
}

static double norm_diff(int n, const double *v, const double *u, double *d) {
  double s = 0.0;
  for (int i = 0; i < n; i++) {
    d[i] = u[i] - v[i];
    s += d[i] * d[i];
  }
  return sqrt(s);
}

/* Shared shape of both repulsions (include/core/gvizVec.h:404-427, 552-575):
 * magnitude over the surface gap, overlap constant once in contact. */
static void push_away(int n, const double *v, const double *u, double num,
                      double radiusSum, double overlapConstant, double weight,
                      double *acc) {
  double d[16];
  if (n > 16)
    return;
  double dist = norm_diff(n, v, u, d);
  double gap = dist - radiusSum;
  double mag = gap > 0.0 ? num / gap : num * overlapConstant;
  if (dist > 0.0) {
    for (int i = 0; i < n; i++)
      acc[i] -= weight * (mag * (d[i] / dist));
  } else if (n == 2) {
    double ux, uy;
    coincident_dir(v, u, &ux, &uy);
    acc[0] -= weight * (mag * ux);
    acc[1] -= weight * (mag * uy);
  }
}

static double fr_mass(size_t degree) {
  (void)degree;
  return 1.0;
}

static void fr_att(int n, double *v, double *u, double k, double *acc) {
  double d[16];
  if (n > 16)
    return;
  double dist = norm_diff(n, v, u, d);
  if (dist < 1e-9)
    return;
  for (int i = 0; i < n; i++)
    acc[i] += dist / k * d[i];
}

static void fr_rep(int n, double *v, double *other, double vMass,
                   double otherMass, double vRadius, double otherRadius,
                   double overlapConstant, double edgeLength, double *acc) {
  (void)vMass;
  push_away(n, v, other, edgeLength * edgeLength, vRadius + otherRadius,
            overlapConstant, (double)(size_t)otherMass, acc);
}

static double linlog_mass(size_t degree) { return 1.0 + (double)degree; }

static void linlog_att(int n, double *v, double *u, double k, double *acc) {
  (void)k;
  double d[16];
  if (n > 16)
    return;
  double dist = norm_diff(n, v, u, d);
  if (dist < 1e-9)
    return;
  for (int i = 0; i < n; i++)
    acc[i] += log(1.0 + dist) / dist * d[i];
}

static void linlog_rep(int n, double *v, double *other, double vMass,
                       double otherMass, double vRadius, double otherRadius,
                       double overlapConstant, double edgeLength, double *acc) {
  (void)edgeLength;
  push_away(n, v, other, vMass * otherMass, vRadius + otherRadius,
            overlapConstant, 1.0, acc);
}

static const gvizForceModel kModelFR = {fr_mass, fr_att, fr_rep};
static const gvizForceModel kModelLinLog = {linlog_mass, linlog_att,
                                            linlog_rep};

/* src/embedders/gvizForceModel.c:57-65 */
const gvizForceModel *gvizForceModelGet(gvizForceModelKind kind) {
  return kind == GVIZ_FORCE_MODEL_LINLOG ? &kModelLinLog : &kModelFR;
}

/* ------------------------------------------------------------- embedder */

static __thread const char *host_error = "";

const char *gvizForceEmbedderLastError(void) {
  return host_error[0] ? host_error : gvizhip_last_error();
}

/* src/embedders/gvizForceEmbedder.c:12-17 */
static double default_box_extent(size_t vertexCount, double edgeLength) {
  if (vertexCount == 0)
    return edgeLength;
  return 0.5 * sqrt((double)vertexCount) * edgeLength;
}

/* The engine's extras beside the reference's state layout: the device
 * handle, the configuration last handed to it (a setter call costs nothing
 * until the next step; a change re-records the step's graph), the
 * precision, the residency flag and the compact CSR the device was built
 * from. */
typedef struct force_side {
  gvizhip_force *f; /* NULL before Begin */
  gvizhip_force_config applied;
  gvizForcePrecision precision;
  int residentPositions;
  int64_t *csrOffsets, *csrAdj;
} force_side;

static force_side *side_of(const gvizForceEmbedderState *s) {
  return (force_side *)gviz_side_get(s);
}

static void device_config(const gvizForceEmbedderState *s, const force_side *x,
                          gvizhip_force_config *c) {
  memset(c, 0, sizeof(*c));
  /* the built-in models are recognised by identity (their vtable pointers) */
  c->model = s->model == gvizForceModelGet(GVIZ_FORCE_MODEL_LINLOG) ? 1 : 0;
  c->barnesHut = s->barnesHutEnabled ? 1 : 0;
  c->preventOverlap = s->preventOverlap ? 1 : 0;
  c->precision = (int32_t)x->precision;
  c->nodesPerCell = (int64_t)s->nodesPerCell;
  c->edgeLength = s->edgeLength;
  c->theta = s->theta;
  c->gravityK = s->gravityK;
  c->overlapConstant = s->overlapConstant;
  c->radiusBase = s->radiusBase;
  c->radiusPerDegree = s->radiusPerDegree;
  c->jitterTolerance = s->jitterTolerance;
}

static int sync_config(gvizForceEmbedderState *s, force_side *x) {
  gvizhip_force_config c;
  device_config(s, x, &c);
  if (memcmp(&c, &x->applied, sizeof(c)) == 0)
    return 0;
  if (gvizhip_force_configure(x->f, &c) != 0)
    return -1;
  x->applied = c;
  return 0;
}

static void gather_positions(gvizForceEmbedderState *s) {
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  for (size_t i = 0; i < s->vertexCount; i++) {
    const double *p = gvizEmbeddedGraphGetVPosition(e, s->vertices[i]);
    s->positionsScratch[2 * i] = p[0];
    s->positionsScratch[2 * i + 1] = p[1];
  }
}

static void scatter_positions(gvizForceEmbedderState *s) {
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  for (size_t i = 0; i < s->vertexCount; i++)
    gvizEmbeddedGraphSetVPosition(e, s->vertices[i], s->positionsScratch + 2 * i);
}

static void record_step(gvizForceEmbedderState *s, const gvizhip_force_stats *r) {
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  s->globalSpeed = r->speed;
  s->speedEfficiency = r->speedEfficiency;
  s->iteration++;
  s->lastMaxDisplacement = r->maxDisp;
  gvizEmbeddedGraphStatAppend(e, "forceEmbedder.maxDisp", r->maxDisp);
  gvizEmbeddedGraphStatAppend(e, "forceEmbedder.speed", r->speed);
  gvizEmbeddedGraphStatAppend(e, "forceEmbedder.attractiveForce", r->meanAtt);
  gvizEmbeddedGraphStatAppend(e, "forceEmbedder.repulsiveForce", r->meanRep);
  gvizEmbeddedGraphStatAppend(e, "forceEmbedder.gravityForce", r->meanGrav);
}

static void action_step(gvizEmbeddedGraph *embedding, void *userData,
                        const gvizActionPayload *payload) {
  (void)userData;
  (void)payload;
  gvizForceEmbedderState *s = (gvizForceEmbedderState *)embedding;
  if (s->begun)
    gvizForceEmbedderStep(s);
}

static void action_toggle_overlap(gvizEmbeddedGraph *embedding, void *userData,
                                  const gvizActionPayload *payload) {
  (void)userData;
  (void)payload;
  gvizForceEmbedderState *s = (gvizForceEmbedderState *)embedding;
  s->preventOverlap = !s->preventOverlap;
}

int gvizForceEmbedderInit(gvizForceEmbedderState *s, gvizSubgraph subgraph,
                          size_t dimension, gvizForceModelKind model) {
  memset(s, 0, sizeof(*s));
  gvizQuadtreeRelease(&s->quadtree); /* drops a host-view hook left at this address */
  if (dimension != 2)
    return -2;
  int res = gvizEmbeddedGraphInit((gvizEmbeddedGraph *)s, subgraph, dimension);
  if (res < 0)
    return res;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  s->model = gvizForceModelGet(model == GVIZ_FORCE_MODEL_LINLOG
                                   ? GVIZ_FORCE_MODEL_LINLOG
                                   : GVIZ_FORCE_MODEL_FRUCHTERMAN_REINGOLD);
  force_side *x = calloc(1, sizeof(force_side)), *prev = NULL;
  if (!x || gviz_side_put(s, x, (void **)&prev) != 0) {
    free(x);
    return -1;
  }
  free(prev); /* a state initialised again without Release */
  x->precision = GVIZ_FORCE_PRECISION_MIXED;
  if (gvizSubgraphToCompactCSR(&e->subgraph, &x->csrOffsets, &x->csrAdj,
                               &s->vertices, &s->vertexCount) != 0)
    return -1;
  size_t n = s->vertexCount;
  s->degree = malloc(sizeof(size_t) * (n ? n : 1));
  s->mass = malloc(sizeof(double) * (n ? n : 1));
  s->positionsScratch = malloc(sizeof(double) * 2 * (n ? n : 1));
  if (!s->degree || !s->mass || !s->positionsScratch)
    return -1;
  double sumDegPlusOne = 0.0;
  for (size_t i = 0; i < n; i++) {
    s->degree[i] = (size_t)(x->csrOffsets[i + 1] - x->csrOffsets[i]);
    s->mass[i] = s->model->vertexMass(s->degree[i]);
    sumDegPlusOne += (double)(s->degree[i] + 1);
  }
  s->meanDegreePlusOne = n ? sumDegPlusOne / (double)n : 0.0;

  s->edgeLength = GVIZ_FORCE_EMBEDDER_EDGE_LENGTH_DEFAULT;
  s->boxExtent = default_box_extent(n, s->edgeLength);
  s->jitterTolerance = GVIZ_FORCE_EMBEDDER_JITTER_TOLERANCE_DEFAULT;
  s->theta = GVIZ_FORCE_EMBEDDER_THETA_DEFAULT;
  s->nodesPerCell = GVIZ_QUADTREE_NODES_PER_CELL_DEFAULT;
  s->gravityK = 0.0;
  s->radiusBase = 0.0;
  s->radiusPerDegree = 0.0;
  s->overlapConstant = GVIZ_FORCE_EMBEDDER_OVERLAP_CONSTANT_DEFAULT;
  s->preventOverlap = 0;
  s->barnesHutEnabled = 1;

  if (gvizEmbeddedGraphAddAction(e, "forceEmbedder.step", action_step, NULL) < 0)
    return -1;
  if (gvizEmbeddedGraphAddAction(e, "forceEmbedder.toggleOverlapPrevention",
                                 action_toggle_overlap, NULL) < 0)
    return -1;
  static const char *const series[] = {
      "forceEmbedder.maxDisp", "forceEmbedder.speed",
      "forceEmbedder.attractiveForce", "forceEmbedder.repulsiveForce",
      "forceEmbedder.gravityForce"};
  for (size_t i = 0; i < sizeof(series) / sizeof(series[0]); i++)
    if (!gvizEmbeddedGraphAddStatSeries(e, series[i], GVIZ_STAT_CHART_LINE_LOG))
      return -1;
  return 0;
}

void gvizForceEmbedderRelease(gvizForceEmbedderState *s) {
  if (!s)
    return;
  force_side *x = (force_side *)gviz_side_take(s);
  if (x) {
    if (x->f)
      gvizhip_force_destroy(x->f);
    free(x->csrOffsets);
    free(x->csrAdj);
    free(x);
  }
  gvizQuadtreeRelease(&s->quadtree);
  s->quadtreeReady = 0;
  gvizEmbeddedGraphRelease((gvizEmbeddedGraph *)s);
  free(s->vertices);
  free(s->degree);
  free(s->mass);
  free(s->positionsScratch);
  s->vertices = NULL;
  s->degree = NULL;
  s->mass = NULL;
  s->positionsScratch = NULL;
}

/* Setters: src/embedders/gvizForceEmbedder.c:357-410 */
void gvizForceEmbedderConfigure(gvizForceEmbedderState *s, double edgeLength,
                                double boxExtent) {
  if (edgeLength > 0.0)
    s->edgeLength = edgeLength;
  if (boxExtent > 0.0)
    s->boxExtent = boxExtent;
}

void gvizForceEmbedderConfigureSpeed(gvizForceEmbedderState *s,
                                     double tolerance) {
  if (tolerance > 0.0)
    s->jitterTolerance = tolerance;
}

void gvizForceEmbedderConfigureBarnesHut(gvizForceEmbedderState *s,
                                         double theta, size_t nodesPerCell) {
  if (theta > 0.0)
    s->theta = theta;
  if (nodesPerCell > 0)
    s->nodesPerCell = nodesPerCell;
}

void gvizForceEmbedderSetBarnesHutEnabled(gvizForceEmbedderState *s,
                                          int enabled) {
  s->barnesHutEnabled = enabled;
}

void gvizForceEmbedderConfigureGravity(gvizForceEmbedderState *s, double k) {
  s->gravityK = k;
}

double gvizForceEmbedderVertexRadius(const gvizForceEmbedderState *s,
                                     size_t index) {
  return s->radiusBase *
         (1.0 + s->radiusPerDegree * sqrt((double)s->degree[index]));
}

void gvizForceEmbedderConfigureRadius(gvizForceEmbedderState *s, double base,
                                      double perDegree) {
  s->radiusBase = base;
  s->radiusPerDegree = perDegree;
}

void gvizForceEmbedderConfigureOverlapPrevention(gvizForceEmbedderState *s,
                                                 double constant) {
  if (constant > 0.0)
    s->overlapConstant = constant;
}

void gvizForceEmbedderSetPreventOverlapEnabled(gvizForceEmbedderState *s,
                                               int enabled) {
  s->preventOverlap = enabled;
}

void gvizForceEmbedderSetPrecision(gvizForceEmbedderState *s,
                                   gvizForcePrecision precision) {
  force_side *x = side_of(s);
  if (x)
    x->precision = precision;
}

gvizForcePrecision gvizForceEmbedderPrecision(const gvizForceEmbedderState *s) {
  force_side *x = side_of(s);
  return x ? x->precision : GVIZ_FORCE_PRECISION_MIXED;
}

void gvizForceEmbedderSetResidentPositions(gvizForceEmbedderState *s,
                                           int enabled) {
  force_side *x = side_of(s);
  if (!x)
    return;
  if (x->residentPositions && !enabled && s->begun)
    gvizForceEmbedderSyncPositions(s);
  x->residentPositions = enabled;
}

int gvizForceEmbedderSyncPositions(gvizForceEmbedderState *s) {
  force_side *x = side_of(s);
  if (!x || !x->f)
    return s->vertexCount == 0 ? 0 : -1;
  if (gvizhip_force_download_positions(x->f, s->positionsScratch) != 0)
    return -1;
  scatter_positions(s);
  return 0;
}

static void quadtree_changed(gvizForceEmbedderState *s);

/* Binary checkpoint file: magic, the host's step count and last maxDisp,
 * then the device state (gvizhip_force_save_state). */
static const char kStateFileMagic[8] = {'G', 'V', 'Z', 'F', 'E', 'S', '0', '1'};

int gvizForceEmbedderSaveState(gvizForceEmbedderState *s, const char *filename) {
  host_error = "";
  force_side *x = side_of(s);
  if (!s->begun || !x || (!x->f && s->vertexCount > 0) || !filename) {
    host_error = "gvizForceEmbedderSaveState: not begun or no file name";
    return -1;
  }
  int64_t bytes = 0;
  void *blob = NULL;
  if (x->f) {
    if (!x->residentPositions) {
      gather_positions(s);
      if (gvizhip_force_upload_positions(x->f, s->positionsScratch) != 0)
        return -1;
    }
    bytes = gvizhip_force_state_bytes(x->f);
    blob = bytes > 0 ? malloc((size_t)bytes) : NULL;
    if (!blob || gvizhip_force_save_state(x->f, blob, bytes) != 0) {
      free(blob);
      host_error = "gvizForceEmbedderSaveState: device state not read";
      return -1;
    }
  }
  FILE *f = fopen(filename, "wb");
  if (!f) {
    free(blob);
    host_error = "gvizForceEmbedderSaveState: cannot open the file";
    return -1;
  }
  const int64_t it = (int64_t)s->iteration;
  int ok = fwrite(kStateFileMagic, 1, 8, f) == 8 && fwrite(&it, 8, 1, f) == 1 &&
           fwrite(&s->lastMaxDisplacement, 8, 1, f) == 1 && fwrite(&bytes, 8, 1, f) == 1 &&
           (bytes == 0 || fwrite(blob, 1, (size_t)bytes, f) == (size_t)bytes);
  ok = (fclose(f) == 0) && ok;
  free(blob);
  if (!ok) {
    host_error = "gvizForceEmbedderSaveState: write failed";
    return -1;
  }
  return 0;
}

int gvizForceEmbedderLoadState(gvizForceEmbedderState *s, const char *filename) {
  host_error = "";
  force_side *x = side_of(s);
  if (!s->begun || !x || (!x->f && s->vertexCount > 0) || !filename) {
    host_error = "gvizForceEmbedderLoadState: Begin the embedder on the same graph first";
    return -1;
  }
  FILE *f = fopen(filename, "rb");
  if (!f) {
    host_error = "gvizForceEmbedderLoadState: cannot open the file";
    return -1;
  }
  char magic[8];
  int64_t it = 0, bytes = -1;
  double last = 0.0;
  int ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kStateFileMagic, 8) == 0 &&
           fread(&it, 8, 1, f) == 1 && fread(&last, 8, 1, f) == 1 &&
           fread(&bytes, 8, 1, f) == 1 && bytes >= 0 && it >= 0;
  void *blob = NULL;
  if (ok && bytes > 0) {
    blob = malloc((size_t)bytes);
    ok = blob && fread(blob, 1, (size_t)bytes, f) == (size_t)bytes;
  }
  fclose(f);
  if (ok && x->f && sync_config(s, x) != 0)
    ok = 0;
  if (ok && x->f) {
    ok = gvizhip_force_load_state(x->f, blob, bytes) == 0;
    double gs = 0.0, eff = 0.0;
    int64_t dit = 0;
    if (ok)
      ok = gvizhip_force_get_speed(x->f, &gs, &eff, &dit) == 0;
    if (ok) {
      s->globalSpeed = gs;
      s->speedEfficiency = eff;
    }
    if (ok && !x->residentPositions)
      ok = gvizForceEmbedderSyncPositions(s) == 0;
  } else if (ok && bytes != 0) {
    ok = 0;
  }
  free(blob);
  if (!ok) {
    host_error = "gvizForceEmbedderLoadState: not a state of this embedder's graph";
    return -1;
  }
  s->iteration = (size_t)it;
  s->lastMaxDisplacement = last;
  quadtree_changed(s);
  return 0;
}

/* The state's quadtree on demand (gvizQuadtreeRoot): the device tree of the
 * last step's starting positions, which the reference's Step rebuilds from
 * positionsScratch and mass (src/embedders/gvizForceEmbedder.c:430-452);
 * positionsScratch receives those positions. The arena keeps the
 * nodesPerCell of its first build, as the reference's Rebuild does. */
static int refresh_quadtree(void *owner) {
  gvizForceEmbedderState *s = (gvizForceEmbedderState *)owner;
  force_side *x = side_of(s);
  size_t n = s->vertexCount;
  if (!x || !x->f || n == 0)
    return -1;
  gvizQuadtree *t = &s->quadtree;
  if (!s->quadtreeReady) {
    if (gviz_quadtree_prepare(t, s->nodesPerCell) != 0)
      return -1;
    s->quadtreeReady = 1;
  }
  uint32_t *order = malloc(sizeof(uint32_t) * n);
  uint64_t *hi = malloc(sizeof(uint64_t) * n), *lo = malloc(sizeof(uint64_t) * n);
  double root[4];
  int rc = -1;
  if (order && hi && lo &&
      gvizhip_force_quadtree_keys(x->f, root, order, hi, lo, s->positionsScratch) == 0) {
    t->points = s->positionsScratch;
    t->masses = s->mass;
    t->pointCount = n;
    rc = gviz_quadtree_from_keys(t, root, order, hi, lo);
  }
  free(order);
  free(hi);
  free(lo);
  return rc;
}

static void quadtree_changed(gvizForceEmbedderState *s) {
  if (s->barnesHutEnabled && s->vertexCount > 0)
    gviz_quadtree_set_refresh(&s->quadtree, refresh_quadtree, s);
}

static int ensure_device(gvizForceEmbedderState *s, force_side *x) {
  if (x->f)
    return sync_config(s, x);
  device_config(s, x, &x->applied);
  if (gvizhip_force_create(&x->f, (int64_t)s->vertexCount, x->csrOffsets,
                           x->csrAdj, &x->applied) != 0) {
    x->f = NULL;
    return -1;
  }
  return 0;
}

/* src/embedders/gvizForceEmbedder.c:412-443 */
int gvizForceEmbedderBegin(gvizForceEmbedderState *s, unsigned int seed) {
  host_error = "";
  gvizEmbeddedGraphRandomizePositions((gvizEmbeddedGraph *)s, s->boxExtent, seed);
  s->globalSpeed = GVIZ_FORCE_EMBEDDER_SPEED_INITIAL;
  s->speedEfficiency = GVIZ_FORCE_EMBEDDER_SPEED_EFFICIENCY_INITIAL;
  s->iteration = 0;
  s->lastMaxDisplacement = 0.0;
  s->begun = 1;
  if (s->vertexCount == 0)
    return 0;
  force_side *x = side_of(s);
  if (!x) {
    host_error = "gvizForceEmbedderBegin: the state was not initialised";
    return -1;
  }
  gather_positions(s);
  if (ensure_device(s, x) != 0)
    return -1;
  if (gvizhip_force_begin(x->f, s->positionsScratch) != 0)
    return -1;
  quadtree_changed(s);
  return 0;
}

/* Appends the stats of a step on an empty subgraph (every mean is 0). */
static double empty_step(gvizForceEmbedderState *s) {
  gvizhip_force_stats r = {0};
  r.speed = s->globalSpeed;
  r.speedEfficiency = s->speedEfficiency;
  r.meanGrav = s->gravityK * s->meanDegreePlusOne;
  record_step(s, &r);
  return 0.0;
}

/* src/embedders/gvizForceEmbedder.c:445-495 */
double gvizForceEmbedderStep(gvizForceEmbedderState *s) {
  host_error = "";
  if (s->vertexCount == 0)
    return empty_step(s);
  force_side *x = side_of(s);
  if (!x || !x->f) {
    host_error = "gvizForceEmbedderStep: Begin has not created the GPU engine";
    return -1.0;
  }
  if (sync_config(s, x) != 0)
    return -1.0;
  if (!x->residentPositions) {
    gather_positions(s);
    if (gvizhip_force_upload_positions(x->f, s->positionsScratch) != 0)
      return -1.0;
  }
  gvizhip_force_stats r;
  if (gvizhip_force_step(x->f, &r) != 0)
    return -1.0;
  quadtree_changed(s);
  if (!x->residentPositions) {
    if (gvizhip_force_download_positions(x->f, s->positionsScratch) != 0)
      return -1.0;
    scatter_positions(s);
  }
  record_step(s, &r);
  return r.maxDisp;
}

/* src/embedders/gvizForceEmbedder.c:497-510, with the loop on the device:
 * positions cross the bus once each way per call, and the stop test on
 * maxDisp runs on the GPU. */
int gvizForceEmbedderRun(gvizForceEmbedderState *s, size_t maxIters,
                         double epsilon) {
  host_error = "";
  if (!s->begun)
    return -1;
  if (s->vertexCount == 0) {
    size_t rounds = 0;
    while (rounds < maxIters) {
      double d = empty_step(s);
      rounds++;
      if (d < epsilon)
        break;
    }
    return (int)rounds;
  }
  force_side *x = side_of(s);
  if (!x || !x->f)
    return -1;
  if (sync_config(s, x) != 0)
    return -1;
  if (!x->residentPositions) {
    gather_positions(s);
    if (gvizhip_force_upload_positions(x->f, s->positionsScratch) != 0)
      return -1;
  }
  enum { kChunk = 4096 };
  gvizhip_force_stats *rec = malloc(sizeof(gvizhip_force_stats) * kChunk);
  if (!rec) {
    host_error = "gvizForceEmbedderRun: out of host memory";
    return -1;
  }
  size_t rounds = 0;
  int failed = 0;
  while (rounds < maxIters) {
    size_t want = maxIters - rounds < kChunk ? maxIters - rounds : kChunk;
    int64_t ran = gvizhip_force_run(x->f, (int64_t)want, epsilon, rec);
    if (ran < 0) {
      failed = 1;
      break;
    }
    for (int64_t i = 0; i < ran; i++)
      record_step(s, &rec[i]);
    rounds += (size_t)ran;
    if ((size_t)ran < want || (ran > 0 && rec[ran - 1].maxDisp < epsilon))
      break;
  }
  free(rec);
  if (rounds > 0)
    quadtree_changed(s);
  if (!x->residentPositions) {
    if (gvizhip_force_download_positions(x->f, s->positionsScratch) != 0)
      return -1;
    scatter_positions(s);
  }
  return failed ? -1 : (int)rounds;
}
