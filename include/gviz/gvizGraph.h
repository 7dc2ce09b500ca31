/*
 * gvizGraph.h -- adjacency-list graph and subgraph views of the host layer
 * (libgviz.so), with the reference's names and insertion-order semantics
 * (reference: include/ds/gvizGraph.h, include/ds/gvizSubgraph.h).
 *
 * Each vertex keeps its neighbours in the order edges were added
 * (src/ds/gvizGraph.c:163-172); an undirected edge (u, v) appends v to u's
 * list and u to v's. A subgraph is either the full graph or the subgraph
 * induced by a vertex subset; its neighbour order is the parent's order.
 */
#ifndef GVIZ_HOST_GRAPH_H
#define GVIZ_HOST_GRAPH_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gvizAdjList {
  size_t *items;
  size_t count;
  size_t capacity;
} gvizAdjList;

typedef struct gvizGraph {
  gvizAdjList *adj;  /* one list per vertex */
  size_t nvertices;
  size_t capacity;
  int directed;
} gvizGraph;

int gvizGraphInit(gvizGraph *g, int directed);
void gvizGraphRelease(gvizGraph *g);
size_t gvizGraphSize(const gvizGraph *g);
int gvizGraphIsDirected(const gvizGraph *g);
/* Appends a vertex (data is not stored by this layer; kept for the
 * reference's signature). Returns 0 or -1. */
int gvizGraphAddVertex(gvizGraph *g, void *data, void *in, void *out);
/* Adds (from, to); undirected graphs also add (to, from). 0 / -1. */
int gvizGraphAddEdge(gvizGraph *g, size_t from, size_t to);
/* Undirected: half the adjacency entries; directed: all of them. */
size_t gvizGraphEdgeCount(const gvizGraph *g);
/* Neighbours of u in insertion order. */
const size_t *gvizGraphNeighbors(const gvizGraph *g, size_t u, size_t *count);

/* Vertex subset: one byte per parent vertex (reference: a bit array). */
typedef uint8_t *gvizVertexSubset;
gvizVertexSubset gvizVertexSubsetCreateEmpty(const gvizGraph *g);
void gvizVertexSubsetShowVertex(gvizVertexSubset vs, size_t u);
void gvizVertexSubsetHideVertex(gvizVertexSubset vs, size_t u);
bool gvizVertexSubsetTest(const gvizVertexSubset vs, size_t u);
size_t gvizVertexSubsetCount(const gvizVertexSubset vs, size_t nvertices);
void gvizVertexSubsetRelease(gvizVertexSubset vs);

typedef struct gvizSubgraph {
  const gvizGraph *g;
  gvizVertexSubset vs; /* NULL: every vertex */
} gvizSubgraph;

gvizSubgraph gvizSubgraphCreateFull(const gvizGraph *g);
/* Takes ownership of v (released with the subgraph's owner). */
gvizSubgraph gvizSubgraphCreateVertexInduced(const gvizGraph *g,
                                             gvizVertexSubset v);
void gvizSubgraphRelease(gvizSubgraph *sg);
bool gvizSubgraphHasVertex(const gvizSubgraph *sg, size_t u);
size_t gvizSubgraphVertexCount(const gvizSubgraph *sg);
size_t gvizSubgraphDegree(const gvizSubgraph *sg, size_t u);
size_t gvizSubgraphEdgeCount(const gvizSubgraph *sg);

/* CSR of a subgraph over parent ids (edges of inactive endpoints dropped),
 * insertion order kept: offsets[n+1], adj[offsets[n]] (malloc'ed; free()).
 * With compact != NULL, also the CSR over compact ids 0..count-1 (ascending
 * parent id): compact receives the parent id of each compact id. */
int gvizSubgraphToCSR(const gvizSubgraph *sg, int64_t **offsets,
                      int64_t **adj);
int gvizSubgraphToCompactCSR(const gvizSubgraph *sg, int64_t **offsets,
                             int64_t **adj, size_t **compact,
                             size_t *count);

#ifdef __cplusplus
}
#endif
#endif
