/*
 * gvizGraph.h -- arrays, bit arrays, adjacency-list graphs, vertex and edge
 * subsets and subgraph views of the host layer (libgviz.so), with the
 * reference's type layouts, names and semantics (reference:
 * include/ds/gvizArray.h, gvizBitArray.h, gvizDeque.h, gvizGraph.h,
 * gvizSubgraph.h), so that a caller compiled against the reference's
 * headers can read the same fields.
 *
 * Each vertex keeps its neighbours in the order edges were added
 * (src/ds/gvizGraph.c:163-172); an undirected edge (u, v) appends v to u's
 * list and u to v's. gvizGraphBuildLayout fixes the edge numbering
 * (vertexOffsets[u] + adjacency index) that edge subsets address. A
 * subgraph is full (every vertex bit and every edge bit set; needs the
 * layout) or induced by a vertex subset (no edge subset); its neighbour
 * order is the parent's.
 */
#ifndef GVIZ_HOST_GRAPH_H
#define GVIZ_HOST_GRAPH_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- arrays */
typedef struct {
  void *arr;          /* elements */
  size_t elementSize; /* bytes per element */
  size_t count;       /* elements in use */
  size_t capacity;    /* elements allocated */
} gvizArray;

int gvizArrayInit(gvizArray *arr, size_t elementSize);
int gvizArrayInitAtCapacity(gvizArray *arr, size_t elementSize, size_t capacity);
int gvizArrayPush(gvizArray *v, void *item);
int gvizArrayPop(gvizArray *v, void *res);
int gvizArrayIsEmpty(const gvizArray *v);
int gvizArrayCopy(gvizArray *dest, const gvizArray *src);
void gvizArrayRelease(gvizArray *v);
void *gvizArrayAtIndex(const gvizArray *v, size_t i);
void *gvizArrayTail(const gvizArray *v);

/* A ring buffer (reference include/ds/gvizDeque.h); present for the layout
 * of the GRIP state, which holds one. */
typedef struct gvizDeque {
  void *arr;
  void *begin;
  size_t count;
  size_t capacity;
  size_t elementSize;
} gvizDeque;

/* ------------------------------------------------------------ bit arrays */
/* Bit k lives in byte k / 8 at position k % 8 (least significant first). */
typedef unsigned char GVIZ_BIT_UNIT;
#define GVIZ_BIT_ARRAY GVIZ_BIT_UNIT *
#define GVIZ_BITS_PER_UNIT (sizeof(GVIZ_BIT_UNIT) * 8)
#define gvizSetBit(A, k) (A[(k) / GVIZ_BITS_PER_UNIT] |= (1 << ((k) % GVIZ_BITS_PER_UNIT)))
#define gvizClearBit(A, k) \
  (A[(k) / GVIZ_BITS_PER_UNIT] &= ~(1 << ((k) % GVIZ_BITS_PER_UNIT)))
#define gvizTestBit(A, k) (A[(k) / GVIZ_BITS_PER_UNIT] & (1 << ((k) % GVIZ_BITS_PER_UNIT)))
#define GVIZ_ARRAY_UNITS(N) ((N + GVIZ_BITS_PER_UNIT - 1) / GVIZ_BITS_PER_UNIT)

/* Set bits in ascending index order (reference include/ds/gvizBitArray.h:
 * 111-115): idx is the last index returned, SIZE_MAX before the first. */
typedef struct {
  GVIZ_BIT_ARRAY arr;
  size_t size;
  size_t idx;
} gvizBitArrayIterator;

gvizBitArrayIterator gvizBitArrayIteratorCreate(GVIZ_BIT_ARRAY arr, size_t size);
/* Bits [start, end). */
gvizBitArrayIterator gvizBitArrayIteratorCreateRange(GVIZ_BIT_ARRAY arr, size_t start,
                                                     size_t end);
bool gvizBitArrayIterate(gvizBitArrayIterator *it, size_t *out_idx);

/* Index of the first element equal (bytewise) to *item, or -1. */
int gvizArrayFindOne(const gvizArray *a, void *item);

/* ---------------------------------------------------------------- graphs */
typedef struct Vertex {
  void *data;          /* caller's data, not owned */
  gvizArray neighbors; /* size_t vertex ids in insertion order */
} gvizVertex;

/* Edge numbering of a graph: edge (u, idx) is bit vertexOffsets[u] + idx. */
typedef struct {
  size_t *vertexOffsets; /* nvertices + 1 */
  size_t nvertices;
  size_t edgeCount; /* adjacency entries */
} gvizGraphLayout;

typedef struct gvizGraph {
  gvizArray vertices;      /* gvizVertex */
  size_t *map;             /* a subgraph copy's map to its parent ids (unused here) */
  int directed;
  gvizGraphLayout *layout; /* NULL until gvizGraphBuildLayout */
} gvizGraph;

int gvizVertexInit(gvizVertex *v, void *data);
int gvizVertexInitAtCapacity(gvizVertex *v, void *data, size_t initialCapacity);
void gvizVertexRelease(gvizVertex *v);

int gvizGraphInit(gvizGraph *g, int directed);
int gvizGraphInitAtCapacity(gvizGraph *g, int directed, size_t initialCapacity);
void gvizGraphRelease(gvizGraph *g);
size_t gvizGraphSize(const gvizGraph *g);
int gvizGraphIsDirected(const gvizGraph *g);
/* Appends a vertex holding `data`; its adjacency is a copy of `out` (NULL:
 * empty), and every vertex listed in `in` gets an entry for it
 * (src/ds/gvizGraph.c:131-161). 0 or -1. */
int gvizGraphAddVertex(gvizGraph *g, void *data, gvizArray *in, gvizArray *out);
/* Adds (from, to); undirected graphs also add (to, from). 0 / -1. */
int gvizGraphAddEdge(gvizGraph *g, size_t from, size_t to);
int gvizGraphEdgeExists(gvizGraph *g, size_t from, size_t to);
/* Undirected: half the adjacency entries; directed: all of them. */
size_t gvizGraphEdgeCount(const gvizGraph *g);
/* (Re)computes the edge numbering; call after the graph's edges change. */
void gvizGraphBuildLayout(gvizGraph *g);
gvizArray *gvizGraphGetVertexNeighbors(const gvizGraph *g, size_t idx);
void *gvizGraphGetVertexData(gvizGraph *g, size_t idx);
void gvizGraphSetVertexData(gvizGraph *g, size_t idx, void *data);
/* Removes the first (from, to) entry and, undirected, the first (to, from).
 * 0 / -1. */
int gvizGraphRemoveEdge(gvizGraph *g, size_t from, size_t to);
/* Neighbours of u in insertion order (no reference counterpart). */
const size_t *gvizGraphNeighbors(const gvizGraph *g, size_t u, size_t *count);

/* ------------------------------------------------------- subsets, views */
typedef GVIZ_BIT_ARRAY gvizVertexSubset;
typedef struct {
  GVIZ_BIT_ARRAY bitset;         /* one bit per adjacency entry */
  const gvizGraphLayout *layout; /* numbering the bits follow */
} gvizEdgeSubset;

typedef struct {
  const struct gvizGraph *g;
  gvizVertexSubset vs;   /* one bit per parent vertex */
  gvizEdgeSubset es;     /* full subgraphs only */
  size_t vertexCapacity; /* bits allocated in vs */
} gvizSubgraph;

gvizVertexSubset gvizVertexSubsetCreateEmpty(const struct gvizGraph *g);
void gvizVertexSubsetShowVertex(const gvizVertexSubset vs, size_t u);
void gvizVertexSubsetHideVertex(const gvizVertexSubset vs, size_t u);
void gvizVertexSubsetClearAll(const gvizVertexSubset vs, size_t nvertices);
void gvizVertexSubsetRelease(const gvizVertexSubset vs);
bool gvizVertexSubsetTest(const gvizVertexSubset vs, size_t u);
size_t gvizVertexSubsetCount(const gvizVertexSubset vs, size_t nvertices);

/* Needs the graph's layout (an empty subset otherwise). */
gvizEdgeSubset gvizEdgeSubsetCreateEmpty(const struct gvizGraph *g);
void gvizEdgeSubsetShowEdge(const gvizEdgeSubset es, size_t u, size_t idx);
void gvizEdgeSubsetHideEdge(const gvizEdgeSubset es, size_t u, size_t idx);
bool gvizEdgeSubsetTestEdge(const gvizEdgeSubset es, size_t u, size_t idx);
void gvizEdgeSubsetRelease(const gvizEdgeSubset es);

/* Every vertex and edge; needs gvizGraphBuildLayout first (an empty view
 * with g == NULL otherwise, as the reference returns). */
gvizSubgraph gvizSubgraphCreateFull(const struct gvizGraph *g);
/* Takes ownership of vs (released with the subgraph). */
gvizSubgraph gvizSubgraphCreateVertexInduced(const struct gvizGraph *g,
                                             gvizVertexSubset vs);
gvizSubgraph gvizSubgraphCreateEmpty(const struct gvizGraph *g);
void gvizSubgraphRelease(gvizSubgraph *sg);
bool gvizSubgraphIsFull(const gvizSubgraph *sg);
bool gvizSubgraphHasVertex(const gvizSubgraph *sg, size_t u);
size_t gvizSubgraphVertexCount(const gvizSubgraph *sg);
bool gvizSubgraphHasEdge(const gvizSubgraph *sg, size_t u, size_t v);
size_t gvizSubgraphDegree(const gvizSubgraph *sg, size_t u);
size_t gvizSubgraphEdgeCount(const gvizSubgraph *sg);
void gvizSubgraphShowVertex(const gvizSubgraph *sg, size_t u);
void gvizSubgraphHideVertex(const gvizSubgraph *sg, size_t u);

gvizBitArrayIterator gvizVertexSubsetIteratorCreate(const gvizVertexSubset vs,
                                                    size_t nvertices);
size_t gvizEdgeSubsetCount(const gvizEdgeSubset es);
size_t gvizEdgeSubsetVertexEdgeCount(const gvizEdgeSubset es, size_t u);
void gvizEdgeSubsetClearVertex(const gvizEdgeSubset es, size_t u);
/* Show / hide the first (u, v) adjacency entry of a full subgraph. */
void gvizSubgraphShowEdge(const gvizSubgraph *sg, size_t u, size_t v);
void gvizSubgraphHideEdge(const gvizSubgraph *sg, size_t u, size_t v);
/* Every vertex and edge of a full subgraph shown. */
void gvizSubgraphMakeFull(gvizSubgraph *sg);
/* Gives a vertex-induced subgraph an edge subset of its induced edges
 * (needs the layout). */
void gvizSubgraphMakeEdgeSubset(gvizSubgraph *sg);
/* After the parent graph changed: grows the vertex capacity (doubling);
 * a full subgraph's edge bits move to the rebuilt layout by (vertex,
 * adjacency index), new entries hidden. 0 / -1. */
int gvizSubgraphRebuild(gvizSubgraph *sg);

/* Iterators (reference include/ds/gvizSubgraph.h:40-62, 202-211): the
 * subgraph's vertices in ascending id; a vertex's subgraph neighbours in
 * adjacency order (edge bits of a full subgraph, the vertex subset of an
 * induced one). */
typedef struct {
  const gvizSubgraph *sg;
  gvizBitArrayIterator it;
} gvizSubgraphVertexIterator;

typedef enum {
  GVIZ_SUBGRAPH_NEIGHBOR_ITER_NONE = 0,
  GVIZ_SUBGRAPH_NEIGHBOR_ITER_FULL,
  GVIZ_SUBGRAPH_NEIGHBOR_ITER_INDUCED,
} gvizSubgraphNeighborIterMode;

typedef struct {
  const gvizSubgraph *sg;
  size_t u;
  size_t base;
  gvizBitArrayIterator it;
  size_t adj_idx;
  const gvizArray *nb;
  gvizSubgraphNeighborIterMode mode;
} gvizSubgraphNeighborIterator;

gvizSubgraphVertexIterator gvizSubgraphVertexIteratorCreate(const gvizSubgraph *sg);
bool gvizSubgraphVertexIterate(gvizSubgraphVertexIterator *it, size_t *out_u);
gvizSubgraphNeighborIterator gvizSubgraphNeighborIteratorCreate(const gvizSubgraph *sg,
                                                                size_t u);
bool gvizSubgraphNeighborIterate(gvizSubgraphNeighborIterator *it, size_t *out_v);

/* CSR of a subgraph over parent ids (edges of inactive endpoints dropped),
 * insertion order kept: offsets[n+1], adj[offsets[n]] (malloc'ed; free()).
 * With compact != NULL, also the CSR over compact ids 0..count-1 (ascending
 * parent id): compact receives the parent id of each compact id. No
 * reference counterpart: the flattening the device engines take. */
int gvizSubgraphToCSR(const gvizSubgraph *sg, int64_t **offsets,
                      int64_t **adj);
int gvizSubgraphToCompactCSR(const gvizSubgraph *sg, int64_t **offsets,
                             int64_t **adj, size_t **compact,
                             size_t *count);
/* The vertex subset as one byte per parent vertex (1: in the subset), the
 * form the device engines take; n bytes are written. */
void gvizVertexSubsetToBytes(const gvizVertexSubset vs, size_t n, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
