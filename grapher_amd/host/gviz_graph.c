/*
 * gviz_graph.c -- arrays, adjacency lists, bit-array vertex/edge subsets and
 * subgraph views of the host layer, and their flattening to CSR for the
 * device engines. Neighbour order is insertion order (reference:
 * src/ds/gvizGraph.c:131-172, src/ds/gvizSubgraph.c:105-540).
 */
#include "gviz/gvizGraph.h"

#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- arrays */

int gvizArrayInitAtCapacity(gvizArray *a, size_t elementSize, size_t capacity) {
  if (!a || elementSize == 0)
    return -1;
  a->elementSize = elementSize;
  a->count = 0;
  a->capacity = capacity;
  a->arr = capacity ? malloc(elementSize * capacity) : NULL;
  if (capacity && !a->arr) {
    a->capacity = 0;
    return -1;
  }
  return 0;
}

int gvizArrayInit(gvizArray *a, size_t elementSize) {
  return gvizArrayInitAtCapacity(a, elementSize, 4);
}

static int array_reserve(gvizArray *a, size_t need) {
  if (need <= a->capacity)
    return 0;
  size_t cap = a->capacity ? a->capacity : 4;
  while (cap < need)
    cap *= 2;
  void *grown = realloc(a->arr, a->elementSize * cap);
  if (!grown)
    return -1;
  a->arr = grown;
  a->capacity = cap;
  return 0;
}

int gvizArrayPush(gvizArray *a, void *item) {
  if (!a || !item || array_reserve(a, a->count + 1) != 0)
    return -1;
  memcpy((char *)a->arr + a->count * a->elementSize, item, a->elementSize);
  a->count++;
  return 0;
}

int gvizArrayPop(gvizArray *a, void *res) {
  if (!a || a->count == 0)
    return -1;
  a->count--;
  if (res)
    memcpy(res, (char *)a->arr + a->count * a->elementSize, a->elementSize);
  return 0;
}

int gvizArrayIsEmpty(const gvizArray *a) { return !a || a->count == 0; }

int gvizArrayCopy(gvizArray *dest, const gvizArray *src) {
  if (!dest || !src || dest->elementSize != src->elementSize ||
      array_reserve(dest, src->count) != 0)
    return -1;
  if (src->count)
    memcpy(dest->arr, src->arr, src->count * src->elementSize);
  dest->count = src->count;
  return 0;
}

void gvizArrayRelease(gvizArray *a) {
  if (!a)
    return;
  free(a->arr);
  a->arr = NULL;
  a->count = a->capacity = 0;
}

void *gvizArrayAtIndex(const gvizArray *a, size_t i) {
  if (!a || i >= a->count)
    return NULL;
  return (char *)a->arr + i * a->elementSize;
}

void *gvizArrayTail(const gvizArray *a) {
  return a && a->count ? gvizArrayAtIndex(a, a->count - 1) : NULL;
}

/* ------------------------------------------------------------ bit arrays */

static GVIZ_BIT_ARRAY bits_alloc(size_t nbits) {
  size_t units = GVIZ_ARRAY_UNITS(nbits);
  return calloc(units ? units : 1, 1);
}

static void bits_fill(GVIZ_BIT_ARRAY a, size_t nbits) {
  if (!a || nbits == 0)
    return;
  memset(a, 0xFF, nbits / 8);
  for (size_t k = nbits / 8 * 8; k < nbits; k++)
    gvizSetBit(a, k);
}

static size_t bits_count(const GVIZ_BIT_UNIT *a, size_t from, size_t to) {
  size_t c = 0;
  for (size_t k = from; k < to; k++)
    c += gvizTestBit(a, k) != 0;
  return c;
}

/* ---------------------------------------------------------------- graphs */

int gvizVertexInitAtCapacity(gvizVertex *v, void *data, size_t initialCapacity) {
  if (!v)
    return -1;
  v->data = data;
  return gvizArrayInitAtCapacity(&v->neighbors, sizeof(size_t), initialCapacity);
}

int gvizVertexInit(gvizVertex *v, void *data) {
  return gvizVertexInitAtCapacity(v, data, 4);
}

void gvizVertexRelease(gvizVertex *v) {
  if (v)
    gvizArrayRelease(&v->neighbors);
}

int gvizGraphInitAtCapacity(gvizGraph *g, int directed, size_t initialCapacity) {
  if (!g)
    return -1;
  memset(g, 0, sizeof(*g));
  g->directed = directed;
  return gvizArrayInitAtCapacity(&g->vertices, sizeof(gvizVertex),
                                 initialCapacity ? initialCapacity : 16);
}

int gvizGraphInit(gvizGraph *g, int directed) {
  return gvizGraphInitAtCapacity(g, directed, 16);
}

static gvizVertex *vertex_at(const gvizGraph *g, size_t u) {
  return (gvizVertex *)gvizArrayAtIndex(&g->vertices, u);
}

void gvizGraphRelease(gvizGraph *g) {
  if (!g)
    return;
  for (size_t i = 0; i < g->vertices.count; i++)
    gvizVertexRelease(vertex_at(g, i));
  gvizArrayRelease(&g->vertices);
  if (g->layout) {
    free(g->layout->vertexOffsets);
    free(g->layout);
  }
  free(g->map);
  memset(g, 0, sizeof(*g));
}

size_t gvizGraphSize(const gvizGraph *g) { return g ? g->vertices.count : 0; }
int gvizGraphIsDirected(const gvizGraph *g) { return g->directed; }

int gvizGraphAddVertex(gvizGraph *g, void *data, gvizArray *in, gvizArray *out) {
  if (!g)
    return -1;
  gvizVertex v;
  if (gvizVertexInitAtCapacity(&v, data, out && out->count ? out->count : 4) != 0)
    return -1;
  if (out && gvizArrayCopy(&v.neighbors, out) != 0) {
    gvizVertexRelease(&v);
    return -1;
  }
  if (gvizArrayPush(&g->vertices, &v) != 0) {
    gvizVertexRelease(&v);
    return -1;
  }
  if (in) {
    size_t idx = g->vertices.count - 1;
    for (size_t i = 0; i < in->count; i++) {
      size_t u = *(size_t *)gvizArrayAtIndex(in, i);
      if (u >= g->vertices.count || gvizArrayPush(&vertex_at(g, u)->neighbors, &idx) != 0)
        return -1;
    }
  }
  return 0;
}

int gvizGraphAddEdge(gvizGraph *g, size_t from, size_t to) {
  if (!g || from >= gvizGraphSize(g) || to >= gvizGraphSize(g))
    return -1;
  if (gvizArrayPush(&vertex_at(g, from)->neighbors, &to) != 0)
    return -1;
  if (!g->directed)
    return gvizArrayPush(&vertex_at(g, to)->neighbors, &from);
  return 0;
}

int gvizGraphEdgeExists(gvizGraph *g, size_t from, size_t to) {
  size_t c = 0;
  const size_t *nb = gvizGraphNeighbors(g, from, &c);
  for (size_t k = 0; k < c; k++)
    if (nb[k] == to)
      return 1;
  return 0;
}

size_t gvizGraphEdgeCount(const gvizGraph *g) {
  size_t total = 0;
  for (size_t i = 0; i < gvizGraphSize(g); i++)
    total += vertex_at(g, i)->neighbors.count;
  return g->directed ? total : total / 2;
}

/* src/ds/gvizGraph.c:96-129: prefix sums of the adjacency lengths */
void gvizGraphBuildLayout(gvizGraph *g) {
  if (!g)
    return;
  size_t n = gvizGraphSize(g);
  if (!g->layout) {
    g->layout = calloc(1, sizeof(gvizGraphLayout));
    if (!g->layout)
      return;
  }
  size_t *off = realloc(g->layout->vertexOffsets, (n + 1) * sizeof(size_t));
  if (!off) {
    free(g->layout->vertexOffsets);
    free(g->layout);
    g->layout = NULL;
    return;
  }
  off[0] = 0;
  for (size_t u = 0; u < n; u++)
    off[u + 1] = off[u] + vertex_at(g, u)->neighbors.count;
  g->layout->vertexOffsets = off;
  g->layout->nvertices = n;
  g->layout->edgeCount = off[n];
}

gvizArray *gvizGraphGetVertexNeighbors(const gvizGraph *g, size_t idx) {
  gvizVertex *v = g ? vertex_at(g, idx) : NULL;
  return v ? &v->neighbors : NULL;
}

void *gvizGraphGetVertexData(gvizGraph *g, size_t idx) {
  gvizVertex *v = g ? vertex_at(g, idx) : NULL;
  return v ? v->data : NULL;
}

void gvizGraphSetVertexData(gvizGraph *g, size_t idx, void *data) {
  gvizVertex *v = g ? vertex_at(g, idx) : NULL;
  if (v)
    v->data = data;
}

const size_t *gvizGraphNeighbors(const gvizGraph *g, size_t u, size_t *count) {
  gvizArray *nb = g ? gvizGraphGetVertexNeighbors(g, u) : NULL;
  if (count)
    *count = nb ? nb->count : 0;
  return nb ? (const size_t *)nb->arr : NULL;
}

/* ------------------------------------------------------- vertex subsets */

gvizVertexSubset gvizVertexSubsetCreateEmpty(const struct gvizGraph *g) {
  return bits_alloc(gvizGraphSize(g));
}
void gvizVertexSubsetShowVertex(const gvizVertexSubset vs, size_t u) { gvizSetBit(vs, u); }
void gvizVertexSubsetHideVertex(const gvizVertexSubset vs, size_t u) { gvizClearBit(vs, u); }
void gvizVertexSubsetClearAll(const gvizVertexSubset vs, size_t nvertices) {
  if (vs)
    memset(vs, 0, GVIZ_ARRAY_UNITS(nvertices));
}
void gvizVertexSubsetRelease(const gvizVertexSubset vs) { free(vs); }
bool gvizVertexSubsetTest(const gvizVertexSubset vs, size_t u) {
  return gvizTestBit(vs, u) != 0;
}
size_t gvizVertexSubsetCount(const gvizVertexSubset vs, size_t nvertices) {
  return vs ? bits_count(vs, 0, nvertices) : 0;
}
void gvizVertexSubsetToBytes(const gvizVertexSubset vs, size_t n, uint8_t *out) {
  for (size_t u = 0; u < n; u++)
    out[u] = vs && gvizTestBit(vs, u) ? 1 : 0;
}

/* --------------------------------------------------------- edge subsets */

gvizEdgeSubset gvizEdgeSubsetCreateEmpty(const struct gvizGraph *g) {
  gvizEdgeSubset es = {NULL, NULL};
  if (!g || !g->layout)
    return es;
  es.bitset = bits_alloc(g->layout->edgeCount);
  es.layout = es.bitset ? g->layout : NULL;
  return es;
}
void gvizEdgeSubsetShowEdge(const gvizEdgeSubset es, size_t u, size_t idx) {
  gvizSetBit(es.bitset, es.layout->vertexOffsets[u] + idx);
}
void gvizEdgeSubsetHideEdge(const gvizEdgeSubset es, size_t u, size_t idx) {
  gvizClearBit(es.bitset, es.layout->vertexOffsets[u] + idx);
}
bool gvizEdgeSubsetTestEdge(const gvizEdgeSubset es, size_t u, size_t idx) {
  return gvizTestBit(es.bitset, es.layout->vertexOffsets[u] + idx) != 0;
}
void gvizEdgeSubsetRelease(const gvizEdgeSubset es) { free(es.bitset); }

/* ------------------------------------------------------------ subgraphs */

gvizSubgraph gvizSubgraphCreateEmpty(const struct gvizGraph *g) {
  gvizSubgraph sg = {NULL, NULL, {NULL, NULL}, 0};
  if (!g)
    return sg;
  sg.vs = gvizVertexSubsetCreateEmpty(g);
  if (!sg.vs)
    return sg;
  sg.es = gvizEdgeSubsetCreateEmpty(g);
  if (!sg.es.bitset) {
    gvizVertexSubsetRelease(sg.vs);
    sg.vs = NULL;
    return sg;
  }
  sg.g = g;
  sg.vertexCapacity = gvizGraphSize(g);
  return sg;
}

gvizSubgraph gvizSubgraphCreateFull(const struct gvizGraph *g) {
  gvizSubgraph sg = {NULL, NULL, {NULL, NULL}, 0};
  if (!g || !g->layout)
    return sg;
  sg = gvizSubgraphCreateEmpty(g);
  if (!sg.vs)
    return sg;
  bits_fill(sg.vs, gvizGraphSize(g));
  bits_fill(sg.es.bitset, g->layout->edgeCount);
  return sg;
}

gvizSubgraph gvizSubgraphCreateVertexInduced(const struct gvizGraph *g,
                                             gvizVertexSubset vs) {
  gvizSubgraph sg = {g, vs, {NULL, NULL}, gvizGraphSize(g)};
  return sg;
}

void gvizSubgraphRelease(gvizSubgraph *sg) {
  if (!sg)
    return;
  gvizVertexSubsetRelease(sg->vs);
  gvizEdgeSubsetRelease(sg->es);
  memset(sg, 0, sizeof(*sg));
}

bool gvizSubgraphIsFull(const gvizSubgraph *sg) {
  return sg && sg->g && sg->vs && sg->es.bitset;
}

bool gvizSubgraphHasVertex(const gvizSubgraph *sg, size_t u) {
  if (!sg || !sg->g || !sg->vs || u >= sg->vertexCapacity)
    return false;
  return gvizVertexSubsetTest(sg->vs, u);
}

size_t gvizSubgraphVertexCount(const gvizSubgraph *sg) {
  if (!sg || !sg->g || !sg->vs)
    return 0;
  return gvizVertexSubsetCount(sg->vs, sg->vertexCapacity);
}

/* Whether adjacency entry idx of u is an edge of the subgraph: its edge bit
 * in a full subgraph, its far end in the subset of an induced one. */
static bool edge_kept(const gvizSubgraph *sg, size_t u, size_t idx, size_t v) {
  if (gvizSubgraphIsFull(sg))
    return gvizEdgeSubsetTestEdge(sg->es, u, idx);
  return gvizSubgraphHasVertex(sg, v);
}

bool gvizSubgraphHasEdge(const gvizSubgraph *sg, size_t u, size_t v) {
  if (!gvizSubgraphHasVertex(sg, u) || !gvizSubgraphHasVertex(sg, v))
    return false;
  size_t c = 0;
  const size_t *nb = gvizGraphNeighbors(sg->g, u, &c);
  for (size_t k = 0; k < c; k++)
    if (nb[k] == v && edge_kept(sg, u, k, v))
      return true;
  return false;
}

/* src/ds/gvizSubgraph.c:407-425: adjacency entries inside the subgraph */
size_t gvizSubgraphDegree(const gvizSubgraph *sg, size_t u) {
  if (!gvizSubgraphHasVertex(sg, u))
    return 0;
  size_t c = 0, d = 0;
  const size_t *nb = gvizGraphNeighbors(sg->g, u, &c);
  for (size_t k = 0; k < c; k++)
    d += edge_kept(sg, u, k, nb[k]) ? 1 : 0;
  return d;
}

size_t gvizSubgraphEdgeCount(const gvizSubgraph *sg) {
  if (!sg || !sg->g)
    return 0;
  size_t n = gvizGraphSize(sg->g), total = 0;
  for (size_t u = 0; u < n; u++)
    total += gvizSubgraphDegree(sg, u);
  return sg->g->directed ? total : total / 2;
}

void gvizSubgraphShowVertex(const gvizSubgraph *sg, size_t u) {
  if (sg && sg->vs && u < sg->vertexCapacity)
    gvizVertexSubsetShowVertex(sg->vs, u);
}

/* Hiding a vertex of a full subgraph also drops its incident edges. */
void gvizSubgraphHideVertex(const gvizSubgraph *sg, size_t u) {
  if (!sg || !sg->vs || u >= sg->vertexCapacity)
    return;
  gvizVertexSubsetHideVertex(sg->vs, u);
  if (!gvizSubgraphIsFull(sg))
    return;
  size_t c = 0;
  const size_t *nb = gvizGraphNeighbors(sg->g, u, &c);
  for (size_t k = 0; k < c; k++) {
    gvizEdgeSubsetHideEdge(sg->es, u, k);
    size_t cv = 0;
    const size_t *back = gvizGraphNeighbors(sg->g, nb[k], &cv);
    for (size_t j = 0; j < cv; j++)
      if (back[j] == u)
        gvizEdgeSubsetHideEdge(sg->es, nb[k], j);
  }
}

/* ----------------------------------------------------------------- CSR */

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
    size_t c = 0;
    const size_t *nb = gvizGraphNeighbors(sg->g, u, &c);
    for (size_t j = 0; j < c; j++)
      if (edge_kept(sg, u, j, nb[j]))
        a[k++] = (int64_t)nb[j];
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
    size_t cn = 0;
    const size_t *nb = gvizGraphNeighbors(sg->g, vert[i], &cn);
    for (size_t j = 0; j < cn; j++)
      if (edge_kept(sg, vert[i], j, nb[j]))
        a[k++] = index[nb[j]];
  }
  free(index);
  *offsets = off;
  *adj = a;
  *compact = vert;
  *count = m;
  return 0;
}

/* ------------------------------------------------------------ iterators */
/* src/ds/gvizBitArray.c:44-110: set bits in ascending index order */

gvizBitArrayIterator gvizBitArrayIteratorCreate(GVIZ_BIT_ARRAY arr, size_t size) {
  gvizBitArrayIterator it = {arr, arr ? size : 0, SIZE_MAX};
  return it;
}

gvizBitArrayIterator gvizBitArrayIteratorCreateRange(GVIZ_BIT_ARRAY arr, size_t start,
                                                     size_t end) {
  gvizBitArrayIterator it = {arr, arr ? end : 0, start > 0 ? start - 1 : SIZE_MAX};
  return it;
}

bool gvizBitArrayIterate(gvizBitArrayIterator *it, size_t *out_idx) {
  if (!it || !it->arr)
    return false;
  for (size_t k = it->idx + 1; k < it->size; k++) {
    /* skip whole zero bytes */
    if ((k & 7) == 0 && k + 8 <= it->size && it->arr[k >> 3] == 0) {
      k += 7;
      continue;
    }
    if (gvizTestBit(it->arr, k)) {
      it->idx = k;
      *out_idx = k;
      return true;
    }
  }
  it->idx = it->size ? it->size - 1 : SIZE_MAX;
  return false;
}

gvizBitArrayIterator gvizVertexSubsetIteratorCreate(const gvizVertexSubset vs,
                                                    size_t nvertices) {
  return gvizBitArrayIteratorCreate(vs, nvertices);
}

/* src/ds/gvizSubgraph.c:451-519 */
gvizSubgraphVertexIterator gvizSubgraphVertexIteratorCreate(const gvizSubgraph *sg) {
  gvizSubgraphVertexIterator it = {sg, {NULL, 0, SIZE_MAX}};
  if (sg && sg->g && sg->vs)
    it.it = gvizVertexSubsetIteratorCreate(sg->vs, sg->vertexCapacity);
  return it;
}

bool gvizSubgraphVertexIterate(gvizSubgraphVertexIterator *it, size_t *out_u) {
  if (!it || !it->sg || !it->sg->g || !it->sg->vs)
    return false;
  return gvizBitArrayIterate(&it->it, out_u);
}

gvizSubgraphNeighborIterator gvizSubgraphNeighborIteratorCreate(const gvizSubgraph *sg,
                                                                size_t u) {
  gvizSubgraphNeighborIterator it = {sg, u, 0, {NULL, 0, SIZE_MAX}, 0, NULL,
                                     GVIZ_SUBGRAPH_NEIGHBOR_ITER_NONE};
  if (!gvizSubgraphHasVertex(sg, u))
    return it;
  it.nb = gvizGraphGetVertexNeighbors(sg->g, u);
  if (!it.nb)
    return it;
  if (gvizSubgraphIsFull(sg)) {
    it.base = sg->es.layout->vertexOffsets[u];
    it.it = gvizBitArrayIteratorCreateRange(sg->es.bitset, it.base,
                                            sg->es.layout->vertexOffsets[u + 1]);
    it.adj_idx = SIZE_MAX;
    it.mode = GVIZ_SUBGRAPH_NEIGHBOR_ITER_FULL;
  } else {
    it.mode = GVIZ_SUBGRAPH_NEIGHBOR_ITER_INDUCED;
  }
  return it;
}

bool gvizSubgraphNeighborIterate(gvizSubgraphNeighborIterator *it, size_t *out_v) {
  if (!it)
    return false;
  if (it->mode == GVIZ_SUBGRAPH_NEIGHBOR_ITER_FULL) {
    size_t pos;
    if (!gvizBitArrayIterate(&it->it, &pos))
      return false;
    size_t idx = pos - it->base;
    if (idx >= it->nb->count)
      return false;
    *out_v = *(size_t *)gvizArrayAtIndex(it->nb, idx);
    return true;
  }
  if (it->mode == GVIZ_SUBGRAPH_NEIGHBOR_ITER_INDUCED) {
    while (it->adj_idx < it->nb->count) {
      size_t v = *(size_t *)gvizArrayAtIndex(it->nb, it->adj_idx++);
      if (gvizVertexSubsetTest(it->sg->vs, v)) {
        *out_v = v;
        return true;
      }
    }
  }
  return false;
}

/* ------------------------------------------- edge subsets and rebuilds */

size_t gvizEdgeSubsetCount(const gvizEdgeSubset es) {
  if (!es.bitset || !es.layout)
    return 0;
  return bits_count(es.bitset, 0, es.layout->vertexOffsets[es.layout->nvertices]);
}

size_t gvizEdgeSubsetVertexEdgeCount(const gvizEdgeSubset es, size_t u) {
  if (!es.bitset || !es.layout || u >= es.layout->nvertices)
    return 0;
  return bits_count(es.bitset, es.layout->vertexOffsets[u], es.layout->vertexOffsets[u + 1]);
}

void gvizEdgeSubsetClearVertex(const gvizEdgeSubset es, size_t u) {
  if (!es.bitset || !es.layout || u >= es.layout->nvertices)
    return;
  for (size_t k = es.layout->vertexOffsets[u]; k < es.layout->vertexOffsets[u + 1]; k++)
    gvizClearBit(es.bitset, k);
}

int gvizArrayFindOne(const gvizArray *a, void *item) {
  if (!a || !item)
    return -1;
  for (size_t i = 0; i < a->count; i++)
    if (!memcmp((char *)a->arr + i * a->elementSize, item, a->elementSize))
      return (int)i;
  return -1;
}

static int find_entry(const gvizGraph *g, size_t u, size_t v, size_t *idx) {
  gvizArray *nb = gvizGraphGetVertexNeighbors(g, u);
  int i = nb ? gvizArrayFindOne(nb, &v) : -1;
  if (i < 0)
    return -1;
  *idx = (size_t)i;
  return 0;
}

/* src/ds/gvizSubgraph.c:273-289: the first (u, v) entry's bit */
void gvizSubgraphShowEdge(const gvizSubgraph *sg, size_t u, size_t v) {
  size_t idx;
  if (gvizSubgraphIsFull(sg) && find_entry(sg->g, u, v, &idx) == 0)
    gvizEdgeSubsetShowEdge(sg->es, u, idx);
}

void gvizSubgraphHideEdge(const gvizSubgraph *sg, size_t u, size_t v) {
  size_t idx;
  if (gvizSubgraphIsFull(sg) && find_entry(sg->g, u, v, &idx) == 0)
    gvizEdgeSubsetHideEdge(sg->es, u, idx);
}

/* :521-528 */
void gvizSubgraphMakeFull(gvizSubgraph *sg) {
  if (!sg || !sg->g || !sg->g->layout || !sg->vs || !sg->es.bitset)
    return;
  bits_fill(sg->vs, sg->g->layout->nvertices);
  bits_fill(sg->es.bitset, sg->g->layout->edgeCount);
}

/* :236-257: an edge subset holding every parent edge between subset vertices */
void gvizSubgraphMakeEdgeSubset(gvizSubgraph *sg) {
  if (!sg || !sg->g || !sg->g->layout || !sg->vs || sg->es.bitset)
    return;
  sg->es = gvizEdgeSubsetCreateEmpty(sg->g);
  if (!sg->es.bitset)
    return;
  size_t n = sg->g->layout->nvertices;
  for (size_t u = 0; u < n; u++) {
    if (!gvizVertexSubsetTest(sg->vs, u))
      continue;
    size_t c = 0;
    const size_t *nb = gvizGraphNeighbors(sg->g, u, &c);
    for (size_t k = 0; k < c; k++)
      if (gvizVertexSubsetTest(sg->vs, nb[k]))
        gvizEdgeSubsetShowEdge(sg->es, u, k);
  }
}

/* :291-378: after the parent graph changed. A vertex-induced view grows its
 * vertex capacity (doubling) to the graph's size; a full view's edge bits
 * keep their (vertex, adjacency index) under the rebuilt layout, new
 * entries start hidden. */
int gvizSubgraphRebuild(gvizSubgraph *sg) {
  if (!sg || !sg->g || !sg->vs)
    return -1;
  size_t need = gvizGraphSize(sg->g);
  if (need > sg->vertexCapacity) {
    size_t cap = sg->vertexCapacity ? sg->vertexCapacity * 2 : 64;
    while (cap < need)
      cap *= 2;
    GVIZ_BIT_ARRAY vs = bits_alloc(cap);
    if (!vs)
      return -1;
    memcpy(vs, sg->vs, GVIZ_ARRAY_UNITS(sg->vertexCapacity));
    free(sg->vs);
    sg->vs = vs;
    sg->vertexCapacity = cap;
  }
  if (!sg->es.bitset)
    return 0;
  const gvizGraphLayout *old = sg->es.layout;
  size_t oldN = old ? old->nvertices : 0;
  size_t *oldOff = NULL;
  if (old) {
    oldOff = malloc(sizeof(size_t) * (oldN + 1));
    if (!oldOff)
      return -1;
    memcpy(oldOff, old->vertexOffsets, sizeof(size_t) * (oldN + 1));
  }
  gvizGraphBuildLayout((gvizGraph *)sg->g);
  const gvizGraphLayout *L = sg->g->layout;
  GVIZ_BIT_ARRAY bits = L ? bits_alloc(L->edgeCount) : NULL;
  if (!bits) {
    free(oldOff);
    return -1;
  }
  for (size_t u = 0; u < oldN && u < L->nvertices; u++) {
    size_t len = oldOff[u + 1] - oldOff[u], now = L->vertexOffsets[u + 1] - L->vertexOffsets[u];
    for (size_t k = 0; k < len && k < now; k++)
      if (gvizTestBit(sg->es.bitset, oldOff[u] + k))
        gvizSetBit(bits, L->vertexOffsets[u] + k);
  }
  free(oldOff);
  free(sg->es.bitset);
  sg->es.bitset = bits;
  sg->es.layout = L;
  return 0;
}

/* src/ds/gvizGraph.c:174-184: removes the first (from, to) entry, and for
 * an undirected graph the first (to, from). 0, or -1 when absent. */
int gvizGraphRemoveEdge(gvizGraph *g, size_t from, size_t to) {
  if (!g || from >= gvizGraphSize(g) || to >= gvizGraphSize(g))
    return -1;
  size_t idx;
  if (find_entry(g, from, to, &idx) != 0)
    return -1;
  gvizArray *a = gvizGraphGetVertexNeighbors(g, from);
  memmove((size_t *)a->arr + idx, (size_t *)a->arr + idx + 1,
          sizeof(size_t) * (a->count - idx - 1));
  a->count--;
  if (g->directed)
    return 0;
  if (find_entry(g, to, from, &idx) != 0)
    return -1;
  a = gvizGraphGetVertexNeighbors(g, to);
  memmove((size_t *)a->arr + idx, (size_t *)a->arr + idx + 1,
          sizeof(size_t) * (a->count - idx - 1));
  a->count--;
  return 0;
}
