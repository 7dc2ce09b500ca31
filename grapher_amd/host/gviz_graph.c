/*
 * gviz_graph.c -- adjacency lists, vertex subsets and subgraph views of the
 * host layer, and their flattening to CSR for the device engine. Neighbour
 * order is insertion order (reference: src/ds/gvizGraph.c:163-172,
 * src/ds/gvizSubgraph.c:466-519).
 */
#include "gviz/gvizGraph.h"

#include <stdlib.h>
#include <string.h>

static int adj_push(gvizAdjList *l, size_t v) {
  if (l->count == l->capacity) {
    size_t cap = l->capacity ? l->capacity * 2 : 4;
    size_t *grown = realloc(l->items, cap * sizeof(size_t));
    if (!grown)
      return -1;
    l->items = grown;
    l->capacity = cap;
  }
  l->items[l->count++] = v;
  return 0;
}

int gvizGraphInit(gvizGraph *g, int directed) {
  if (!g)
    return -1;
  memset(g, 0, sizeof(*g));
  g->directed = directed;
  return 0;
}

void gvizGraphRelease(gvizGraph *g) {
  if (!g)
    return;
  for (size_t i = 0; i < g->nvertices; i++)
    free(g->adj[i].items);
  free(g->adj);
  memset(g, 0, sizeof(*g));
}

size_t gvizGraphSize(const gvizGraph *g) { return g ? g->nvertices : 0; }
int gvizGraphIsDirected(const gvizGraph *g) { return g->directed; }

int gvizGraphAddVertex(gvizGraph *g, void *data, void *in, void *out) {
  (void)data;
  (void)in;
  (void)out;
  if (!g)
    return -1;
  if (g->nvertices == g->capacity) {
    size_t cap = g->capacity ? g->capacity * 2 : 16;
    gvizAdjList *grown = realloc(g->adj, cap * sizeof(gvizAdjList));
    if (!grown)
      return -1;
    g->adj = grown;
    g->capacity = cap;
  }
  memset(&g->adj[g->nvertices], 0, sizeof(gvizAdjList));
  g->nvertices++;
  return 0;
}

int gvizGraphAddEdge(gvizGraph *g, size_t from, size_t to) {
  if (!g || from >= g->nvertices || to >= g->nvertices)
    return -1;
  if (adj_push(&g->adj[from], to) != 0)
    return -1;
  if (!g->directed)
    return adj_push(&g->adj[to], from);
  return 0;
}

size_t gvizGraphEdgeCount(const gvizGraph *g) {
  size_t total = 0;
  for (size_t i = 0; i < g->nvertices; i++)
    total += g->adj[i].count;
  return g->directed ? total : total / 2;
}

const size_t *gvizGraphNeighbors(const gvizGraph *g, size_t u, size_t *count) {
  if (!g || u >= g->nvertices) {
    if (count)
      *count = 0;
    return NULL;
  }
  if (count)
    *count = g->adj[u].count;
  return g->adj[u].items;
}

gvizVertexSubset gvizVertexSubsetCreateEmpty(const gvizGraph *g) {
  size_t n = gvizGraphSize(g);
  return calloc(n ? n : 1, 1);
}

void gvizVertexSubsetShowVertex(gvizVertexSubset vs, size_t u) { vs[u] = 1; }
void gvizVertexSubsetHideVertex(gvizVertexSubset vs, size_t u) { vs[u] = 0; }
bool gvizVertexSubsetTest(const gvizVertexSubset vs, size_t u) {
  return vs[u] != 0;
}
size_t gvizVertexSubsetCount(const gvizVertexSubset vs, size_t nvertices) {
  size_t c = 0;
  for (size_t i = 0; i < nvertices; i++)
    c += vs[i] != 0;
  return c;
}
void gvizVertexSubsetRelease(gvizVertexSubset vs) { free(vs); }

gvizSubgraph gvizSubgraphCreateFull(const gvizGraph *g) {
  gvizSubgraph sg = {g, NULL};
  return sg;
}

gvizSubgraph gvizSubgraphCreateVertexInduced(const gvizGraph *g,
                                             gvizVertexSubset v) {
  gvizSubgraph sg = {g, v};
  return sg;
}

void gvizSubgraphRelease(gvizSubgraph *sg) {
  if (!sg)
    return;
  free(sg->vs);
  sg->vs = NULL;
}

bool gvizSubgraphHasVertex(const gvizSubgraph *sg, size_t u) {
  if (u >= gvizGraphSize(sg->g))
    return false;
  return sg->vs ? sg->vs[u] != 0 : true;
}

size_t gvizSubgraphVertexCount(const gvizSubgraph *sg) {
  size_t n = gvizGraphSize(sg->g);
  return sg->vs ? gvizVertexSubsetCount(sg->vs, n) : n;
}

/* src/ds/gvizSubgraph.c:407-425: neighbours inside the subgraph */
size_t gvizSubgraphDegree(const gvizSubgraph *sg, size_t u) {
  if (!gvizSubgraphHasVertex(sg, u))
    return 0;
  const gvizAdjList *l = &sg->g->adj[u];
  if (!sg->vs)
    return l->count;
  size_t d = 0;
  for (size_t k = 0; k < l->count; k++)
    d += sg->vs[l->items[k]] != 0;
  return d;
}

size_t gvizSubgraphEdgeCount(const gvizSubgraph *sg) {
  size_t n = gvizGraphSize(sg->g), total = 0;
  for (size_t u = 0; u < n; u++)
    total += gvizSubgraphDegree(sg, u);
  return sg->g->directed ? total : total / 2;
}

int gvizSubgraphToCSR(const gvizSubgraph *sg, int64_t **offsets,
                      int64_t **adj) {
  size_t n = gvizGraphSize(sg->g);
  int64_t *off = malloc(sizeof(int64_t) * (n + 1));
  if (!off)
    return -1;
  size_t total = 0;
  for (size_t u = 0; u < n; u++) {
    off[u] = (int64_t)total;
    total += gvizSubgraphDegree(sg, u);
  }
  off[n] = (int64_t)total;
  int64_t *a = malloc(sizeof(int64_t) * (total ? total : 1));
  if (!a) {
    free(off);
    return -1;
  }
  size_t k = 0;
  for (size_t u = 0; u < n; u++) {
    if (!gvizSubgraphHasVertex(sg, u))
      continue;
    const gvizAdjList *l = &sg->g->adj[u];
    for (size_t j = 0; j < l->count; j++)
      if (!sg->vs || sg->vs[l->items[j]])
        a[k++] = (int64_t)l->items[j];
  }
  *offsets = off;
  *adj = a;
  return 0;
}

int gvizSubgraphToCompactCSR(const gvizSubgraph *sg, int64_t **offsets,
                             int64_t **adj, size_t **compact, size_t *count) {
  size_t n = gvizGraphSize(sg->g);
  size_t m = gvizSubgraphVertexCount(sg);
  size_t *vert = malloc(sizeof(size_t) * (m ? m : 1));
  int64_t *index = malloc(sizeof(int64_t) * (n ? n : 1));
  if (!vert || !index) {
    free(vert);
    free(index);
    return -1;
  }
  size_t c = 0;
  for (size_t u = 0; u < n; u++) {
    index[u] = -1;
    if (gvizSubgraphHasVertex(sg, u)) {
      index[u] = (int64_t)c;
      vert[c++] = u;
    }
  }
  int64_t *off = malloc(sizeof(int64_t) * (m + 1));
  if (!off) {
    free(vert);
    free(index);
    return -1;
  }
  size_t total = 0;
  for (size_t i = 0; i < m; i++) {
    off[i] = (int64_t)total;
    total += gvizSubgraphDegree(sg, vert[i]);
  }
  off[m] = (int64_t)total;
  int64_t *a = malloc(sizeof(int64_t) * (total ? total : 1));
  if (!a) {
    free(vert);
    free(index);
    free(off);
    return -1;
  }
  size_t k = 0;
  for (size_t i = 0; i < m; i++) {
    const gvizAdjList *l = &sg->g->adj[vert[i]];
    for (size_t j = 0; j < l->count; j++)
      if (index[l->items[j]] >= 0)
        a[k++] = index[l->items[j]];
  }
  free(index);
  *offsets = off;
  *adj = a;
  *compact = vert;
  *count = m;
  return 0;
}
