/*
 * gviz_grip_embedder.c -- gvizGRIPEmbedder* with the multilevel schedule's
 * heavy phases on the GPU (reference: src/embedders/gvizGRIPEmbedder.c).
 *
 * The host keeps the schedule (current layer, round, K policy, draw mask,
 * stat series) exactly as the reference does; the MIS filtration, the kNN
 * searches, the barycentric placement of each new layer and every
 * refinement round run on the device (include/gviz_hip.h, GRIP section),
 * which keeps positions, per-vertex state and kNN lists in HBM. Positions
 * are uploaded before and read back after each device phase that uses
 * them, unless resident positions are on (gvizGRIPEmbedderSetResidentPositions;
 * then gvizGRIPEmbedderSyncPositions fetches them). Dimensions 2, 3 and 4 (the ones whose rotation removal
 * the reference defines, gvizGRIPEmbedder.c:1192-1261); Init returns -2
 * for another.
 */
#include "gviz/gvizGRIPEmbedder.h"
#include "gviz/gvizGRIPInternal.h"

#include "gviz_hip.h"
#include "gviz_side.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GRIP_EDGE_LENGTH 10.0              /* gvizGRIPEmbedder.c:31 */
#define GRIP_KNN_CAPACITY_DEFAULT 256      /* :33 */
#define GRIP_K_MAX_DEFAULT 128             /* :34 */
#define GRIP_EMBED_MAX_ROUNDS_PER_LAYER 30 /* :1263 */
#define GRIP_EMBED_CONVERGENCE_FACTOR 1e-3 /* :1264 */
#define MIS_MAX_LAYERS 128

static size_t border_at(const gvizGRIPState *s, size_t i) {
  return *(const size_t *)gvizArrayAtIndex(&s->misBorder, i);
}

/* The engine's extras beside the reference's state layout: the device
 * state, the residency flags, and a byte-per-vertex copy of a vertex subset
 * (the form the device takes). */
typedef struct grip_side {
  gvizhip_grip *device;
  int residentPositions; /* 1: rounds keep positions on the device */
  int devicePositionsFresh;
  uint8_t *bytes;        /* n entries */
} grip_side;

static grip_side *side_of(const gvizGRIPState *s) {
  return (grip_side *)gviz_side_get(s);
}

static const uint8_t *subset_bytes(gvizGRIPState *s, const gvizVertexSubset vs) {
  grip_side *x = side_of(s);
  gvizVertexSubsetToBytes(vs, gvizGraphSize(((gvizEmbeddedGraph *)s)->subgraph.g), x->bytes);
  return x->bytes;
}

static const uint8_t *visible_bytes(gvizGRIPState *s) {
  return subset_bytes(s, ((gvizEmbeddedGraph *)s)->drawMask.visibleVertices);
}

static size_t clamp_k(size_t k, size_t minK, size_t maxK) {
  if (k < minK)
    k = minK;
  if (k > maxK)
    k = maxK;
  return k;
}

/* src/embedders/gvizGRIPEmbedder.c:63-111 */
static size_t compute_k(const gvizGRIPState *s, size_t maxK, int forPlacement) {
  size_t minK = ((const gvizEmbeddedGraph *)s)->embedding.dim + 1;
  size_t cap = s->knnCapacity;
  if (maxK > cap)
    maxK = cap;
  gvizGRIPKPolicy policy = s->kPolicy;
  if (forPlacement && policy == GVIZ_GRIP_K_PLACEMENT_DECAY)
    policy = GVIZ_GRIP_K_LAYER_DECAY;
  else if (!forPlacement && policy == GVIZ_GRIP_K_PLACEMENT_DECAY)
    policy = GVIZ_GRIP_K_CONSTANT;
  else if (forPlacement && policy == GVIZ_GRIP_K_BUDGET)
    policy = GVIZ_GRIP_K_LAYER_DECAY;
  size_t k;
  switch (policy) {
  case GVIZ_GRIP_K_LAYER_DECAY:
    k = s->currLayer >= 8 * sizeof(size_t) ? 0 : maxK >> s->currLayer;
    break;
  case GVIZ_GRIP_K_LAYER_GROW: {
    size_t depth = s->layerCount > 0 ? s->layerCount - 1 - s->currLayer : 0;
    k = depth >= 8 * sizeof(size_t) ? 0 : maxK >> depth;
    break;
  }
  case GVIZ_GRIP_K_BUDGET: {
    size_t active = border_at(s, s->currLayer);
    size_t finest = border_at(s, 0);
    if (active == 0)
      return clamp_k(maxK, minK, cap);
    double scaled = (double)maxK * ((double)finest / (double)active);
    k = scaled >= (double)cap ? cap : (size_t)scaled;
    return clamp_k(k, minK, cap);
  }
  case GVIZ_GRIP_K_CONSTANT:
  case GVIZ_GRIP_K_PLACEMENT_DECAY:
  default:
    k = maxK;
    break;
  }
  return clamp_k(k, minK, maxK);
}

static size_t placement_k(const gvizGRIPState *s) {
  return compute_k(s, s->placementKMax, 1);
}

static size_t refinement_k(const gvizGRIPState *s) {
  return compute_k(s, s->refinementKMax, 0);
}

/* src/algorithms/search/gvizKNearest.c:149-166: whether one BFS per visible
 * seed beats one BFS per target. */
static int prefer_batch(const gvizGRIPState *s, size_t visibleCount,
                        size_t targetCount) {
  return gvizSearchKNearestPreferBatch(&((const gvizEmbeddedGraph *)s)->subgraph,
                                       visibleCount, targetCount);
}

void gvizGRIPEmbedderConfigureStats(gvizGRIPState *s, bool enable) {
  if (s)
    s->statsRegisterMode = enable ? 0 : 1;
}

void gvizGRIPEmbedderConfigureKnnCapacity(gvizGRIPState *s, size_t knnCapacity) {
  if (s)
    s->knnCapacity = knnCapacity;
}

/* src/embedders/gvizGRIPEmbedder.c:134-147 */
void gvizGRIPEmbedderConfigureK(gvizGRIPState *s, size_t placementKMax,
                                size_t refinementKMax, gvizGRIPKPolicy policy) {
  if (!s)
    return;
  s->kPolicy = policy;
  if (placementKMax > 0)
    s->placementKMax = placementKMax > s->knnCapacity ? s->knnCapacity : placementKMax;
  if (refinementKMax > 0)
    s->refinementKMax =
        refinementKMax > s->knnCapacity ? s->knnCapacity : refinementKMax;
}

gvizGRIPRoundStats gvizGRIPEmbedderLastRoundStats(const gvizGRIPState *s) {
  return s ? s->lastRoundStats : (gvizGRIPRoundStats){0, 0, 0};
}

static void action_refine_round(gvizEmbeddedGraph *embedding, void *userData,
                                const gvizActionPayload *payload) {
  (void)userData;
  (void)payload;
  gvizGRIPState *s = (gvizGRIPState *)embedding;
  if (s->layerCount == 0)
    return;
  gvizGRIPEmbedderRefineRound(s);
}

static void action_next_stage(gvizEmbeddedGraph *embedding, void *userData,
                              const gvizActionPayload *payload) {
  (void)userData;
  (void)payload;
  gvizGRIPState *s = (gvizGRIPState *)embedding;
  if (s->layerCount == 0)
    return;
  gvizGRIPEmbedderNextStage(s);
}

/* src/embedders/gvizGRIPEmbedder.c:194-332. The device engine (graph, kNN
 * arrays of knnCapacity entries per vertex, per-vertex state) is created by
 * the first Begin. */
int gvizGRIPEmbedderInit(gvizGRIPState *s, gvizSubgraph subgraph,
                         size_t diameter, size_t dimension) {
  (void)diameter; /* sized the reference's border array; ours is fixed */
  size_t activeCount = gvizSubgraphVertexCount(&subgraph);
  if (activeCount < dimension + 1)
    return -1;
  if (dimension < 2 || dimension > 4)
    return -2;
  uint8_t statsRegisterMode = s->statsRegisterMode;
  size_t knnCapacityRequested = s->knnCapacity;
  memset(s, 0, sizeof(*s));
  s->statsRegisterMode = statsRegisterMode;
  s->placementKMax = GRIP_K_MAX_DEFAULT;
  s->refinementKMax = GRIP_K_MAX_DEFAULT;
  s->knnCapacity =
      knnCapacityRequested > 0 ? knnCapacityRequested : GRIP_KNN_CAPACITY_DEFAULT;
  s->kPolicy = GVIZ_GRIP_K_CONSTANT;
  s->statsEnabled = s->statsRegisterMode != 1;

  int res = gvizEmbeddedGraphInit((gvizEmbeddedGraph *)s, subgraph, dimension);
  if (res < 0)
    return res;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  grip_side *x = calloc(1, sizeof(grip_side)), *prev = NULL;
  if (!x || gviz_side_put(s, x, (void **)&prev) != 0) {
    free(x);
    return -1;
  }
  if (prev) { /* a state initialised again without Release */
    if (prev->device)
      gvizhip_grip_destroy(prev->device);
    free(prev->bytes);
    free(prev);
  }
  size_t nAll = gvizGraphSize(subgraph.g);
  x->bytes = malloc(nAll ? nAll : 1);
  s->misFiltration = calloc(activeCount ? activeCount : 1, sizeof(size_t));
  if (!x->bytes || !s->misFiltration ||
      gvizArrayInitAtCapacity(&s->misBorder, sizeof(size_t), MIS_MAX_LAYERS) != 0 ||
      gvizArrayPush(&s->misBorder, &activeCount) != 0)
    return -1;
  gvizEmbeddedGraphDrawMaskClearVertices(e);

  if (gvizEmbeddedGraphAddAction(e, "grip.refineRound", action_refine_round, NULL) < 0)
    return -1;
  if (gvizEmbeddedGraphAddAction(e, "grip.nextStage", action_next_stage, NULL) < 0)
    return -1;
  if (s->statsEnabled) {
    if (!gvizEmbeddedGraphAddStatSeries(e, "grip.heat", GVIZ_STAT_CHART_LINE_LOG) ||
        !gvizEmbeddedGraphAddStatSeries(e, "grip.meanDisp", GVIZ_STAT_CHART_LINE_LOG) ||
        !gvizEmbeddedGraphAddStatSeries(e, "grip.maxDisp", GVIZ_STAT_CHART_LINE_LOG) ||
        !gvizEmbeddedGraphAddStatSeries(e, "grip.meanForce", GVIZ_STAT_CHART_LINE_LOG))
      return -1;
  }
  return 0;
}

void gvizGRIPEmbedderRelease(gvizGRIPState *s) {
  if (!s)
    return;
  grip_side *x = (grip_side *)gviz_side_take(s);
  if (x) {
    if (x->device)
      gvizhip_grip_destroy(x->device);
    free(x->bytes);
    free(x);
  }
  free(s->misFiltration);
  gvizArrayRelease(&s->misBorder);
  s->misFiltration = NULL;
  gvizEmbeddedGraphRelease((gvizEmbeddedGraph *)s);
}

/* src/embedders/gvizGRIPEmbedder.c:335-420: n+1 vertices of a regular
 * n-simplex centred at the origin with the given side length (component-wise
 * construction on the unit sphere, pairwise dot product -1/n, then scaled). */
static void regular_simplex(size_t n, double side, double *out) {
  memset(out, 0, sizeof(double) * (n + 1) * n);
  double c = -1.0 / (double)n;
  out[0] = 1.0;
  for (size_t k = 1; k <= n; k++) {
    double *vk = out + k * n;
    for (size_t j = 0; j < k; j++) {
      const double *vj = out + j * n;
      double dot = 0.0;
      for (size_t m = 0; m < j; m++)
        dot += vk[m] * vj[m];
      vk[j] = (c - dot) / vj[j];
    }
    if (k < n) {
      double sq = 0.0;
      for (size_t m = 0; m < k; m++)
        sq += vk[m] * vk[m];
      double rem = 1.0 - sq;
      vk[k] = rem > 0.0 ? sqrt(rem) : 0.0;
    }
  }
  double edge = 0.0;
  for (size_t m = 0; m < n; m++) {
    double d = out[m] - out[n + m];
    edge += d * d;
  }
  double scale = side / sqrt(edge);
  for (size_t i = 0; i < (n + 1) * n; i++)
    out[i] *= scale;
}

static double *host_positions(gvizGRIPState *s) {
  return ((gvizEmbeddedGraph *)s)->embedding.vertexPositions;
}

static int push_positions(gvizGRIPState *s) {
  grip_side *x = side_of(s);
  if (x->residentPositions && x->devicePositionsFresh)
    return 0;
  if (gvizhip_grip_upload(x->device, host_positions(s), NULL, NULL, NULL) != 0)
    return -1;
  x->devicePositionsFresh = 1;
  return 0;
}

static int pull_positions(gvizGRIPState *s) {
  grip_side *x = side_of(s);
  if (x->residentPositions)
    return 0;
  return gvizhip_grip_download(x->device, host_positions(s), NULL, NULL, NULL);
}

int gvizGRIPEmbedderSyncPositions(gvizGRIPState *s) {
  grip_side *x = s ? side_of(s) : NULL;
  if (!x || !x->device)
    return -1;
  return gvizhip_grip_download(x->device, host_positions(s), NULL, NULL, NULL);
}

void gvizGRIPEmbedderSetResidentPositions(gvizGRIPState *s, int enabled) {
  grip_side *x = s ? side_of(s) : NULL;
  if (!x)
    return;
  if (x->residentPositions && !enabled && x->device && s->layerCount > 0)
    gvizGRIPEmbedderSyncPositions(s);
  if (enabled && !x->residentPositions)
    x->devicePositionsFresh = 0;
  x->residentPositions = enabled;
}

/* Binary checkpoint: magic, n, kNN capacity, layer count, current layer and
 * round, last round stats, the filtration (checked on load), the draw mask,
 * positions, the per-vertex decorators and the kNN lists. */
static void sync_draw_mask(gvizGRIPState *s);

static const char kGripStateMagic[8] = {'G', 'V', 'Z', 'G', 'R', 'S', '0', '2'};

int gvizGRIPEmbedderSaveState(gvizGRIPState *s, const char *filename) {
  grip_side *x = s ? side_of(s) : NULL;
  if (!x || !x->device || !filename || s->layerCount == 0)
    return -1;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  const int64_t n = (int64_t)gvizGraphSize(e->subgraph.g);
  const int64_t kcap = (int64_t)s->knnCapacity, active = (int64_t)border_at(s, 0);
  const int64_t dim = (int64_t)e->embedding.dim;
  if (x->residentPositions && gvizGRIPEmbedderSyncPositions(s) != 0)
    return -1;
  gvizhip_grip_vertex_nd *dec =
      malloc(sizeof(gvizhip_grip_vertex_nd) * (size_t)(n ? n : 1));
  int64_t *kv = malloc(sizeof(int64_t) * (size_t)(n * kcap > 0 ? n * kcap : 1));
  int64_t *kd = malloc(sizeof(int64_t) * (size_t)(n * kcap > 0 ? n * kcap : 1));
  int ok = dec && kv && kd && gvizhip_grip_download_nd(x->device, NULL, dec, kv, kd) == 0;
  FILE *f = ok ? fopen(filename, "wb") : NULL;
  ok = ok && f;
  if (ok) {
    const int64_t head[7] = {n, kcap, (int64_t)s->layerCount, (int64_t)s->currLayer,
                             (int64_t)s->currRound, active, dim};
    const double st[3] = {s->lastRoundStats.maxDisplacement,
                          s->lastRoundStats.meanDisplacement, s->lastRoundStats.meanForce};
    ok = fwrite(kGripStateMagic, 1, 8, f) == 8 && fwrite(head, 8, 7, f) == 7 &&
         fwrite(st, 8, 3, f) == 3;
    for (int64_t i = 0; ok && i < active; i++) {
      const int64_t v = (int64_t)s->misFiltration[i];
      ok = fwrite(&v, 8, 1, f) == 1;
    }
    ok = ok && fwrite(visible_bytes(s), 1, (size_t)n, f) == (size_t)n &&
         fwrite(host_positions(s), sizeof(double), (size_t)(dim * n), f) ==
             (size_t)(dim * n) &&
         fwrite(dec, sizeof(gvizhip_grip_vertex_nd), (size_t)n, f) == (size_t)n &&
         fwrite(kv, 8, (size_t)(n * kcap), f) == (size_t)(n * kcap) &&
         fwrite(kd, 8, (size_t)(n * kcap), f) == (size_t)(n * kcap);
  }
  if (f && fclose(f) != 0)
    ok = 0;
  free(dec);
  free(kv);
  free(kd);
  return ok ? 0 : -1;
}

int gvizGRIPEmbedderLoadState(gvizGRIPState *s, const char *filename) {
  grip_side *x = s ? side_of(s) : NULL;
  if (!x || !x->device || !filename || s->layerCount == 0)
    return -1;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  const int64_t n = (int64_t)gvizGraphSize(e->subgraph.g);
  FILE *f = fopen(filename, "rb");
  if (!f)
    return -1;
  const int64_t dim = (int64_t)e->embedding.dim;
  char magic[8];
  int64_t head[7];
  double st[3];
  int ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kGripStateMagic, 8) == 0 &&
           fread(head, 8, 7, f) == 7 && fread(st, 8, 3, f) == 3 && head[0] == n &&
           head[1] == (int64_t)s->knnCapacity && head[2] == (int64_t)s->layerCount &&
           head[3] >= 0 && head[3] < head[2] && head[5] == (int64_t)border_at(s, 0) &&
           head[6] == dim;
  /* the same graph and configuration give the same filtration */
  for (int64_t i = 0; ok && i < head[5]; i++) {
    int64_t v = -1;
    ok = fread(&v, 8, 1, f) == 1 && v == (int64_t)s->misFiltration[i];
  }
  const int64_t kcap = ok ? head[1] : 0;
  uint8_t *vis = malloc((size_t)(n ? n : 1));
  double *pos = malloc(sizeof(double) * (size_t)(dim * n > 0 ? dim * n : 1));
  gvizhip_grip_vertex_nd *dec =
      malloc(sizeof(gvizhip_grip_vertex_nd) * (size_t)(n ? n : 1));
  int64_t *kv = malloc(sizeof(int64_t) * (size_t)(n * kcap > 0 ? n * kcap : 1));
  int64_t *kd = malloc(sizeof(int64_t) * (size_t)(n * kcap > 0 ? n * kcap : 1));
  ok = ok && vis && pos && dec && kv && kd &&
       fread(vis, 1, (size_t)n, f) == (size_t)n &&
       fread(pos, sizeof(double), (size_t)(dim * n), f) == (size_t)(dim * n) &&
       fread(dec, sizeof(gvizhip_grip_vertex_nd), (size_t)n, f) == (size_t)n &&
       fread(kv, 8, (size_t)(n * kcap), f) == (size_t)(n * kcap) &&
       fread(kd, 8, (size_t)(n * kcap), f) == (size_t)(n * kcap);
  fclose(f);
  if (ok) {
    memcpy(host_positions(s), pos, sizeof(double) * (size_t)(dim * n));
    ok = gvizhip_grip_upload_nd(x->device, pos, dec, kv, kd) == 0;
  }
  if (ok) {
    for (int64_t u = 0; u < n; u++) {
      if (vis[u])
        gvizVertexSubsetShowVertex(e->drawMask.visibleVertices, (size_t)u);
      else
        gvizVertexSubsetHideVertex(e->drawMask.visibleVertices, (size_t)u);
    }
    gvizEmbeddedGraphDrawMaskNotifyChanged(e);
    s->currLayer = (size_t)head[3];
    s->currRound = (size_t)head[4];
    s->lastRoundStats.maxDisplacement = st[0];
    s->lastRoundStats.meanDisplacement = st[1];
    s->lastRoundStats.meanForce = st[2];
    x->devicePositionsFresh = 1;
    sync_draw_mask(s);
  }
  free(vis);
  free(pos);
  free(dec);
  free(kv);
  free(kd);
  return ok ? 0 : -1;
}

/* clearDecorators then updateKNNs (src/embedders/gvizGRIPEmbedder.c:936-1006)
 * for the current layer against the draw mask. */
static int refresh_layer(gvizGRIPState *s) {
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  size_t end = border_at(s, s->currLayer);
  grip_side *x = side_of(s);
  if (gvizhip_grip_clear_decorators(x->device, (int64_t)end) != 0)
    return -1;
  size_t n = gvizGraphSize(e->subgraph.g);
  size_t visibleCount = gvizVertexSubsetCount(e->drawMask.visibleVertices, n);
  int batch = end > 0 && visibleCount > 0 && visibleCount < end &&
              prefer_batch(s, visibleCount, end);
  return gvizhip_grip_update_knn(x->device, (int64_t)end, visible_bytes(s),
                                 (int64_t)refinement_k(s), batch);
}

/* gripSyncDrawMask (:186-192) */
static void sync_draw_mask(gvizGRIPState *s) {
  gvizEmbeddedGraphSetDrawMaskEdgePolicy(
      (gvizEmbeddedGraph *)s,
      s->currLayer == 0 ? GVIZ_DRAW_EDGES_IF_BOTH_VISIBLE : GVIZ_DRAW_EDGES_NONE);
}

static int ensure_device(gvizGRIPState *s) {
  grip_side *x = side_of(s);
  if (!x)
    return -1;
  if (x->device)
    return 0;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  int64_t *off = NULL, *adj = NULL;
  if (gvizSubgraphToCSR(&e->subgraph, &off, &adj) != 0)
    return -1;
  size_t n = gvizGraphSize(e->subgraph.g);
  for (size_t u = 0; u < n; u++)
    x->bytes[u] = gvizSubgraphHasVertex(&e->subgraph, u) ? 1 : 0;
  int rc = gvizhip_grip_create(&x->device, (int64_t)n, off, adj, x->bytes,
                               (int64_t)s->knnCapacity);
  free(off);
  free(adj);
  if (rc != 0) {
    x->device = NULL;
    return -1;
  }
  if (gvizhip_grip_set_dimension(x->device, (int64_t)e->embedding.dim) != 0) {
    gvizhip_grip_destroy(x->device);
    x->device = NULL;
    return -1;
  }
  return 0;
}

/* ------------------------------------------------- white-box stages */

static size_t active_count(const gvizGRIPState *s) { return border_at(s, 0); }

/* src/embedders/gvizGRIPEmbedder.c:450-483 */
void makeFirstMISPartition(gvizGRIPState *state, gvizVertexSubset out) {
  if (ensure_device(state) != 0)
    return;
  grip_side *x = side_of(state);
  const gvizSubgraph *sg = &((gvizEmbeddedGraph *)state)->subgraph;
  size_t n = gvizGraphSize(sg->g), active = active_count(state);
  uint8_t *mis = malloc(n ? n : 1);
  int64_t *tail = malloc(sizeof(int64_t) * (active ? active : 1));
  int64_t cnt = mis && tail ? gvizhip_grip_mis_first(x->device, mis, tail) : -1;
  if (cnt >= 0) {
    for (size_t v = 0; v < n; v++)
      if (mis[v])
        gvizVertexSubsetShowVertex(out, v);
    for (size_t r = 0; r + (size_t)cnt < active; r++)
      state->misFiltration[active - 1 - r] = (size_t)tail[r];
    size_t border = gvizVertexSubsetCount(out, n);
    gvizArrayPush(&state->misBorder, &border);
  }
  free(mis);
  free(tail);
}

/* :672-739 */
int iterMISFiltration(gvizGRIPState *state, size_t i, gvizVertexSubset vertices) {
  if (i < 2 || i > state->misBorder.count || ensure_device(state) != 0)
    return 0;
  grip_side *x = side_of(state);
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)state;
  size_t n = gvizGraphSize(e->subgraph.g);
  int64_t *cand = malloc(sizeof(int64_t) * (n ? n : 1));
  uint8_t *picked = malloc(n ? n : 1);
  int cont = 0;
  if (!cand || !picked)
    goto done;
  size_t m = 0;
  for (size_t v = 0; v < n; v++)
    if (gvizVertexSubsetTest(vertices, v))
      cand[m++] = (int64_t)v;
  int64_t chosen = gvizhip_grip_mis_layer(x->device, (int64_t)i, cand, (int64_t)m, picked);
  if (chosen < 0)
    goto done;
  /* the rest, in ascending order, backwards from the previous layer's end */
  size_t ptr = border_at(state, i - 1);
  for (size_t j = 0; j < m; j++)
    if (!picked[j])
      state->misFiltration[--ptr] = (size_t)cand[j];
  size_t border = (size_t)chosen;
  if (gvizArrayPush(&state->misBorder, &border) != 0)
    goto done;
  gvizVertexSubsetClearAll(vertices, n);
  for (size_t j = 0; j < m; j++)
    if (picked[j])
      gvizVertexSubsetShowVertex(vertices, (size_t)cand[j]);
  cont = border > e->embedding.dim + 1;
done:
  free(cand);
  free(picked);
  return cont;
}

/* :576-597 */
int migrateOneToFinalLayer(gvizGRIPState *state, GVIZ_BIT_ARRAY finalLayer, size_t count) {
  (void)finalLayer;
  if (count < 1 || count > state->misBorder.count || ensure_device(state) != 0)
    return -1;
  grip_side *x = side_of(state);
  size_t active = active_count(state), layers = state->misBorder.count;
  int64_t *filt = malloc(sizeof(int64_t) * (active ? active : 1));
  int64_t *bord = malloc(sizeof(int64_t) * layers);
  int rc = -1;
  if (filt && bord) {
    for (size_t k = 0; k < active; k++)
      filt[k] = (int64_t)state->misFiltration[k];
    for (size_t k = 0; k < layers; k++)
      bord[k] = (int64_t)border_at(state, k);
    rc = gvizhip_grip_migrate(x->device, filt, (int64_t)active, bord, (int64_t)layers,
                              (int64_t)count);
    if (rc == 0) {
      for (size_t k = 0; k < active; k++)
        state->misFiltration[k] = (size_t)filt[k];
      for (size_t k = 0; k < layers; k++)
        *(size_t *)gvizArrayAtIndex(&state->misBorder, k) = (size_t)bord[k];
    }
  }
  free(filt);
  free(bord);
  return rc;
}

/* :741-777 */
size_t createMISFiltration(gvizGRIPState *state) {
  if (ensure_device(state) != 0)
    return 0;
  grip_side *x = side_of(state);
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)state;
  size_t active = active_count(state);
  int64_t *filt = malloc(sizeof(int64_t) * (active ? active : 1));
  int64_t borders[MIS_MAX_LAYERS];
  if (!filt)
    return 0;
  int64_t layers = gvizhip_grip_mis_filtration(x->device, (int64_t)e->embedding.dim, filt,
                                               borders);
  if (layers <= 0 || layers > MIS_MAX_LAYERS) {
    free(filt);
    return 0;
  }
  for (size_t i = 0; i < active; i++)
    state->misFiltration[i] = (size_t)filt[i];
  free(filt);
  state->misBorder.count = 0;
  for (int64_t i = 0; i < layers; i++) {
    size_t b = (size_t)borders[i];
    gvizArrayPush(&state->misBorder, &b);
  }
  return (size_t)layers;
}

/* src/embedders/gvizGRIPEmbedder.c:1008-1030. Without a usable GPU, or for
 * a filtration the device cannot finish (a subgraph that is not connected),
 * layerCount stays 0 (gvizhip_last_error() says why). */
void gvizGRIPEmbedderBegin(gvizGRIPState *s) {
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  s->layerCount = 0;
  size_t layers = createMISFiltration(s);
  if (layers == 0)
    return;
  grip_side *x = side_of(s);
  s->layerCount = layers;
  s->currLayer = s->layerCount - 1;
  s->currRound = 0;

  size_t dim = e->embedding.dim;
  double simplex[4 * 5];
  regular_simplex(dim, GRIP_EDGE_LENGTH * 1000, simplex);
  for (size_t j = 0; j < dim + 1; j++) {
    gvizEmbeddedGraphDrawMaskShowVertex(e, s->misFiltration[j]);
    gvizEmbeddedGraphSetVPosition(e, s->misFiltration[j], simplex + j * dim);
  }
  x->devicePositionsFresh = 0;
  if (push_positions(s) != 0 || refresh_layer(s) != 0) {
    s->layerCount = 0;
    return;
  }
  sync_draw_mask(s);
}

/* src/embedders/gvizGRIPEmbedder.c:1034-1069 with placeLayerVertices
 * (:816-868): the new layer's vertices go to the barycentre of their
 * placement-K nearest visible vertices, then become visible. */
void gvizGRIPEmbedderNextStage(gvizGRIPState *s) {
  if (s->currLayer == 0 || s->layerCount == 0)
    return;
  s->currLayer--;
  s->currRound = 0;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
  size_t begin = border_at(s, s->currLayer + 1);
  size_t end = border_at(s, s->currLayer);
  size_t placeCount = end - begin;
  size_t n = gvizGraphSize(e->subgraph.g);
  size_t visibleCount = gvizVertexSubsetCount(e->drawMask.visibleVertices, n);
  int batch = placeCount > 0 && visibleCount > 0 && visibleCount < placeCount &&
              prefer_batch(s, visibleCount, placeCount);
  if (push_positions(s) != 0)
    return;
  if (placeCount > 0 &&
      gvizhip_grip_place(side_of(s)->device, (int64_t)begin, (int64_t)end,
                         visible_bytes(s), (int64_t)placement_k(s), batch) != 0)
    return;
  for (size_t i = begin; i < end; i++)
    gvizEmbeddedGraphDrawMaskShowVertex(e, s->misFiltration[i]);
  if (pull_positions(s) != 0)
    return;
  if (refresh_layer(s) != 0)
    return;
  sync_draw_mask(s);
}

/* src/embedders/gvizGRIPEmbedder.c:1289-1333 */
void gvizGRIPEmbedderRefineRound(gvizGRIPState *s) {
  if (s->layerCount == 0)
    return;
  size_t layer = s->currLayer;
  size_t active = border_at(s, layer);
  if (push_positions(s) != 0)
    return;
  gvizhip_grip_round_stats st;
  if (gvizhip_grip_refine_round(side_of(s)->device, (int64_t)active, (int64_t)layer,
                                &st) != 0)
    return;
  if (pull_positions(s) != 0)
    return;
  s->lastRoundStats.maxDisplacement = st.maxDisp;
  s->lastRoundStats.meanDisplacement = st.meanDisp;
  s->lastRoundStats.meanForce = st.meanForce;
  if (s->statsEnabled) {
    gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)s;
    gvizEmbeddedGraphStatAppend(e, "grip.heat", st.meanHeat);
    gvizEmbeddedGraphStatAppend(e, "grip.meanDisp", st.meanDisp);
    gvizEmbeddedGraphStatAppend(e, "grip.maxDisp", st.maxDisp);
    gvizEmbeddedGraphStatAppend(e, "grip.meanForce", st.meanForce);
  }
  s->currRound++;
}

/* src/embedders/gvizGRIPEmbedder.c:1266-1287 */
int gvizGRIPEmbedderEmbed(gvizGRIPState *s) {
  if (!s)
    return -1;
  gvizGRIPEmbedderBegin(s);
  if (s->layerCount == 0)
    return -1;
  double epsilon = GRIP_EDGE_LENGTH * GRIP_EMBED_CONVERGENCE_FACTOR;
  for (;;) {
    for (size_t round = 0; round < GRIP_EMBED_MAX_ROUNDS_PER_LAYER; round++) {
      gvizGRIPEmbedderRefineRound(s);
      if (s->lastRoundStats.maxDisplacement < epsilon)
        break;
    }
    if (s->currLayer == 0)
      break;
    gvizGRIPEmbedderNextStage(s);
  }
  return 0;
}
