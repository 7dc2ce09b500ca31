/*
 * gviz_embedded_graph.c -- positions, action and stat registries, draw mask
 * and text checkpoints shared by every embedder of the host layer
 * (reference: src/embedders/gvizEmbeddedGraph.c).
 */
#include "gviz/gvizEmbeddedGraph.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

int gvizEmbeddedGraphInit(gvizEmbeddedGraph *e, gvizSubgraph subgraph,
                          size_t dimension) {
  if (!e || !subgraph.g || dimension == 0)
    return -1;
  memset(e, 0, sizeof(*e));
  e->subgraph = subgraph;
  e->embedding.dim = dimension;
  size_t n = gvizGraphSize(subgraph.g);
  e->embedding.vertexPositions = calloc((n ? n : 1) * dimension, sizeof(double));
  e->drawMask.visibleVertices = gvizVertexSubsetCreateEmpty(subgraph.g);
  if (!e->embedding.vertexPositions || !e->drawMask.visibleVertices)
    return -1;
  for (size_t u = 0; u < n; u++)
    if (gvizSubgraphHasVertex(&subgraph, u))
      gvizVertexSubsetShowVertex(e->drawMask.visibleVertices, u);
  e->drawMask.edgePolicy = GVIZ_DRAW_EDGES_ALL;
  return 0;
}

void gvizEmbeddedGraphRelease(gvizEmbeddedGraph *e) {
  if (!e)
    return;
  free(e->embedding.vertexPositions);
  free(e->actions.actions);
  for (size_t i = 0; i < e->stats.count; i++)
    free(e->stats.series[i].samples);
  free(e->stats.series);
  gvizVertexSubsetRelease(e->drawMask.visibleVertices);
  gvizSubgraphRelease(&e->subgraph);
  gvizSubgraphRelease(&e->highlight);
  e->embedding.vertexPositions = NULL;
  memset(&e->actions, 0, sizeof(e->actions));
  memset(&e->stats, 0, sizeof(e->stats));
  memset(&e->drawMask, 0, sizeof(e->drawMask));
}

size_t gvizEmbeddedGraphDim(const gvizEmbeddedGraph *e) { return e->embedding.dim; }
size_t gvizEmbeddedGraphPositionCount(const gvizEmbeddedGraph *e) {
  return gvizGraphSize(e->subgraph.g);
}
const double *gvizEmbeddedGraphPositions(const gvizEmbeddedGraph *e) {
  return e->embedding.vertexPositions;
}
double *gvizEmbeddedGraphGetVPosition(gvizEmbeddedGraph *e, size_t u) {
  return e->embedding.vertexPositions + u * e->embedding.dim;
}
void gvizEmbeddedGraphSetVPosition(gvizEmbeddedGraph *e, size_t u,
                                   const double *pos) {
  memcpy(e->embedding.vertexPositions + u * e->embedding.dim, pos,
         sizeof(double) * e->embedding.dim);
}
void gvizEmbeddedGraphAddVPosition(gvizEmbeddedGraph *e, size_t u,
                                   const double *delta) {
  double *p = e->embedding.vertexPositions + u * e->embedding.dim;
  for (size_t k = 0; k < e->embedding.dim; k++)
    p[k] += delta[k];
}

/* src/embedders/gvizEmbeddedGraph.c:406-426 */
void gvizEmbeddedGraphRandomizePositions(gvizEmbeddedGraph *e, double boxExtent,
                                         unsigned int seed) {
  if (!e)
    return;
  if (seed == 0)
    seed = (unsigned int)time(NULL);
  size_t dim = e->embedding.dim, n = gvizGraphSize(e->subgraph.g);
  for (size_t u = 0; u < n; u++) {
    if (!gvizSubgraphHasVertex(&e->subgraph, u))
      continue;
    double *p = gvizEmbeddedGraphGetVPosition(e, u);
    for (size_t d = 0; d < dim; d++) {
      double unit = (double)rand_r(&seed) / ((double)RAND_MAX + 1.0);
      p[d] = boxExtent * (2.0 * unit - 1.0);
    }
  }
}

static gvizAction *find_action(const gvizEmbeddedGraph *e, const char *name) {
  for (size_t i = 0; i < e->actions.count; i++)
    if (strcmp(e->actions.actions[i].name, name) == 0)
      return &e->actions.actions[i];
  return NULL;
}

int gvizEmbeddedGraphAddAction(gvizEmbeddedGraph *e, const char *name,
                               gvizActionHandler handler, void *userData) {
  if (!e || !name || !handler)
    return -1;
  gvizAction *existing = find_action(e, name);
  if (existing) {
    existing->handler = handler;
    existing->userData = userData;
    return 0;
  }
  if (e->actions.count == e->actions.capacity) {
    size_t cap = e->actions.capacity ? e->actions.capacity * 2 : 4;
    gvizAction *grown = realloc(e->actions.actions, cap * sizeof(gvizAction));
    if (!grown)
      return -1;
    e->actions.actions = grown;
    e->actions.capacity = cap;
  }
  e->actions.actions[e->actions.count++] = (gvizAction){name, handler, userData};
  return 0;
}

int gvizEmbeddedGraphRemoveAction(gvizEmbeddedGraph *e, const char *name) {
  if (!e || !name)
    return -1;
  gvizAction *a = find_action(e, name);
  if (!a)
    return -1;
  *a = e->actions.actions[--e->actions.count];
  return 0;
}

const gvizAction *gvizEmbeddedGraphFindAction(const gvizEmbeddedGraph *e,
                                              const char *name) {
  if (!e || !name)
    return NULL;
  return find_action(e, name);
}

size_t gvizEmbeddedGraphActionCount(const gvizEmbeddedGraph *e) {
  return e->actions.count;
}

const gvizAction *gvizEmbeddedGraphActionAt(const gvizEmbeddedGraph *e,
                                            size_t idx) {
  return e && idx < e->actions.count ? &e->actions.actions[idx] : NULL;
}

int gvizEmbeddedGraphInvokeAction(gvizEmbeddedGraph *e, const char *name,
                                  const gvizActionPayload *payload) {
  const gvizAction *a = gvizEmbeddedGraphFindAction(e, name);
  if (!a)
    return -1;
  gvizActionPayload zero = {0};
  a->handler(e, a->userData, payload ? payload : &zero);
  return 0;
}

static gvizStatSeries *find_series(const gvizEmbeddedGraph *e, const char *name) {
  for (size_t i = 0; i < e->stats.count; i++)
    if (strcmp(e->stats.series[i].name, name) == 0)
      return &e->stats.series[i];
  return NULL;
}

gvizStatSeries *gvizEmbeddedGraphAddStatSeries(gvizEmbeddedGraph *e,
                                               const char *name,
                                               gvizStatChartKind kind) {
  if (!e || !name)
    return NULL;
  gvizStatSeries *s = find_series(e, name);
  if (s) {
    s->kind = kind;
    return s;
  }
  if (e->stats.count == e->stats.capacity) {
    size_t cap = e->stats.capacity ? e->stats.capacity * 2 : 4;
    gvizStatSeries *grown = realloc(e->stats.series, cap * sizeof(gvizStatSeries));
    if (!grown)
      return NULL;
    e->stats.series = grown;
    e->stats.capacity = cap;
  }
  s = &e->stats.series[e->stats.count++];
  memset(s, 0, sizeof(*s));
  s->name = name;
  s->kind = kind;
  return s;
}

int gvizEmbeddedGraphStatAppend(gvizEmbeddedGraph *e, const char *name,
                                double value) {
  if (!e || !name)
    return -1;
  gvizStatSeries *s = find_series(e, name);
  if (!s)
    s = gvizEmbeddedGraphAddStatSeries(e, name, GVIZ_STAT_CHART_LINE);
  if (!s)
    return -1;
  if (s->count == s->capacity) {
    size_t cap = s->capacity ? s->capacity * 2 : 64;
    double *grown = realloc(s->samples, cap * sizeof(double));
    if (!grown)
      return -1;
    s->samples = grown;
    s->capacity = cap;
  }
  s->samples[s->count++] = value;
  s->revision++;
  return 0;
}

void gvizEmbeddedGraphStatClear(gvizEmbeddedGraph *e, const char *name) {
  gvizStatSeries *s = e && name ? find_series(e, name) : NULL;
  if (!s)
    return;
  s->count = 0;
  s->revision++;
}

size_t gvizEmbeddedGraphStatSeriesCount(const gvizEmbeddedGraph *e) {
  return e->stats.count;
}

const gvizStatSeries *gvizEmbeddedGraphStatSeriesAt(const gvizEmbeddedGraph *e,
                                                    size_t idx) {
  return e && idx < e->stats.count ? &e->stats.series[idx] : NULL;
}

const gvizStatSeries *gvizEmbeddedGraphFindStatSeries(const gvizEmbeddedGraph *e,
                                                      const char *name) {
  return e && name ? find_series(e, name) : NULL;
}

/* Draw mask (src/embedders/gvizEmbeddedGraph.c:74-122): vertex edits do not
 * bump the revision; policy changes, resets and explicit notifications do. */
void gvizEmbeddedGraphDrawMaskShowVertex(gvizEmbeddedGraph *e, size_t u) {
  gvizVertexSubsetShowVertex(e->drawMask.visibleVertices, u);
}
void gvizEmbeddedGraphDrawMaskHideVertex(gvizEmbeddedGraph *e, size_t u) {
  gvizVertexSubsetHideVertex(e->drawMask.visibleVertices, u);
}
void gvizEmbeddedGraphDrawMaskClearVertices(gvizEmbeddedGraph *e) {
  gvizVertexSubsetClearAll(e->drawMask.visibleVertices, gvizGraphSize(e->subgraph.g));
}
void gvizEmbeddedGraphDrawMaskNotifyChanged(gvizEmbeddedGraph *e) {
  e->drawMask.revision++;
}
void gvizEmbeddedGraphResetDrawMask(gvizEmbeddedGraph *e) {
  e->drawMask.edgePolicy = GVIZ_DRAW_EDGES_ALL;
  size_t n = gvizGraphSize(e->subgraph.g);
  for (size_t u = 0; u < n; u++) {
    if (gvizSubgraphHasVertex(&e->subgraph, u))
      gvizVertexSubsetShowVertex(e->drawMask.visibleVertices, u);
    else
      gvizVertexSubsetHideVertex(e->drawMask.visibleVertices, u);
  }
  e->drawMask.revision++;
}
uint64_t gvizEmbeddedGraphDrawMaskRevision(const gvizEmbeddedGraph *e) {
  return e->drawMask.revision;
}
void gvizEmbeddedGraphSetDrawMaskEdgePolicy(gvizEmbeddedGraph *e,
                                            gvizDrawEdgePolicy policy) {
  e->drawMask.edgePolicy = policy;
  e->drawMask.revision++;
}
bool gvizEmbeddedGraphIsVertexVisible(const gvizEmbeddedGraph *e, size_t u) {
  if (!gvizSubgraphHasVertex(&e->subgraph, u))
    return false;
  return gvizVertexSubsetTest(e->drawMask.visibleVertices, u);
}
bool gvizEmbeddedGraphIsEdgeVisible(const gvizEmbeddedGraph *e, size_t u,
                                    size_t v) {
  switch (e->drawMask.edgePolicy) {
  case GVIZ_DRAW_EDGES_NONE:
    return false;
  case GVIZ_DRAW_EDGES_IF_BOTH_VISIBLE:
    return gvizEmbeddedGraphIsVertexVisible(e, u) &&
           gvizEmbeddedGraphIsVertexVisible(e, v);
  case GVIZ_DRAW_EDGES_ALL:
  default: {
    if (!gvizSubgraphHasVertex(&e->subgraph, u) ||
        !gvizSubgraphHasVertex(&e->subgraph, v))
      return false;
    return gvizSubgraphHasEdge(&e->subgraph, u, v);
  }
  }
}

/* src/embedders/gvizEmbeddedGraph.c:428-448 */
int gvizEmbeddedGraphSaveEmbedding(gvizEmbeddedGraph *e, const char *name,
                                   const char *filename) {
  FILE *f = fopen(filename, "w");
  if (!f)
    return -1;
  fprintf(f, "%s\n", name);
  size_t n = gvizGraphSize(e->subgraph.g);
  fprintf(f, "%zu %zu\n", n, e->embedding.dim);
  for (size_t i = 0; i < n; i++) {
    double *p = gvizEmbeddedGraphGetVPosition(e, i);
    for (size_t j = 0; j < e->embedding.dim; j++)
      fprintf(f, "%f ", p[j]);
    fprintf(f, "\n");
  }
  fclose(f);
  return 0;
}

/* src/embedders/gvizEmbeddedGraph.c:450-485 */
int gvizEmbeddedGraphLoadEmbedding(gvizEmbeddedGraph *e, const char *filename) {
  FILE *f = fopen(filename, "r");
  if (!f)
    return -1;
  char name[256];
  size_t n, dim;
  if (!fgets(name, sizeof(name), f) || fscanf(f, "%zu %zu", &n, &dim) != 2 ||
      n != gvizGraphSize(e->subgraph.g) || dim != e->embedding.dim) {
    fclose(f);
    return -1;
  }
  for (size_t i = 0; i < n; i++) {
    double *p = gvizEmbeddedGraphGetVPosition(e, i);
    for (size_t j = 0; j < dim; j++)
      if (fscanf(f, "%lf", &p[j]) != 1) {
        fclose(f);
        return -1;
      }
  }
  fclose(f);
  return 0;
}

int gvizEmbeddedGraphIsPlanarEmbedded(const gvizEmbeddedGraph *e) {
  return e && e->planarEmbedded;
}

/* ------------------------------------------------ structure and growth */

const gvizSubgraph *gvizEmbeddedGraphStructure(const gvizEmbeddedGraph *e) {
  return &e->subgraph;
}

const gvizDrawMask *gvizEmbeddedGraphGetDrawMask(const gvizEmbeddedGraph *e) {
  return &e->drawMask;
}

/* src/embedders/gvizEmbeddedGraph.c:154-192: a bare vertex, present in the
 * subgraph and visible; positions and mask grow with the subgraph's vertex
 * capacity (new positions zero). */
int gvizEmbeddedGraphAddVertex(gvizEmbeddedGraph *e, void *data) {
  if (!e)
    return -1;
  gvizGraph *g = (gvizGraph *)e->subgraph.g;
  size_t oldCap = e->subgraph.vertexCapacity;
  if (gvizGraphAddVertex(g, data, NULL, NULL) != 0)
    return -1;
  size_t idx = gvizGraphSize(g) - 1;
  if (gvizSubgraphRebuild(&e->subgraph) != 0)
    return -1;
  size_t newCap = e->subgraph.vertexCapacity, dim = e->embedding.dim;
  if (newCap > oldCap) {
    double *pos = realloc(e->embedding.vertexPositions, sizeof(double) * newCap * dim);
    if (!pos)
      return -1;
    memset(pos + oldCap * dim, 0, sizeof(double) * (newCap - oldCap) * dim);
    e->embedding.vertexPositions = pos;
    GVIZ_BIT_ARRAY mask = calloc(GVIZ_ARRAY_UNITS(newCap) ? GVIZ_ARRAY_UNITS(newCap) : 1, 1);
    if (!mask)
      return -1;
    memcpy(mask, e->drawMask.visibleVertices, GVIZ_ARRAY_UNITS(oldCap));
    gvizVertexSubsetRelease(e->drawMask.visibleVertices);
    e->drawMask.visibleVertices = mask;
  }
  gvizSubgraphShowVertex(&e->subgraph, idx);
  gvizEmbeddedGraphDrawMaskShowVertex(e, idx);
  gvizEmbeddedGraphDrawMaskNotifyChanged(e);
  return 0;
}

/* :194-211: a full subgraph rebuilds its edge subset and shows the edge */
int gvizEmbeddedGraphAddEdge(gvizEmbeddedGraph *e, size_t from, size_t to) {
  if (!e)
    return -1;
  if (gvizGraphAddEdge((gvizGraph *)e->subgraph.g, from, to) != 0)
    return -1;
  if (gvizSubgraphIsFull(&e->subgraph)) {
    if (gvizSubgraphRebuild(&e->subgraph) != 0)
      return -1;
    gvizSubgraphShowEdge(&e->subgraph, from, to);
    gvizEmbeddedGraphDrawMaskNotifyChanged(e);
  }
  return 0;
}

/* --------------------------------------------------------------- highlight */
/* :491-516: the highlight owns its subsets */
void gvizEmbeddedGraphClearHighlight(gvizEmbeddedGraph *e) {
  if (!e)
    return;
  if (e->highlight.vs || e->highlight.es.bitset)
    gvizSubgraphRelease(&e->highlight);
  memset(&e->highlight, 0, sizeof(e->highlight));
}

void gvizEmbeddedGraphSetHighlight(gvizEmbeddedGraph *e, gvizSubgraph highlight) {
  if (!e)
    return;
  gvizEmbeddedGraphClearHighlight(e);
  e->highlight = highlight;
}

int gvizEmbeddedGraphHasHighlight(const gvizEmbeddedGraph *e) {
  return e && e->highlight.vs != NULL;
}

const gvizSubgraph *gvizEmbeddedGraphGetHighlight(const gvizEmbeddedGraph *e) {
  return gvizEmbeddedGraphHasHighlight(e) ? &e->highlight : NULL;
}
