/*
 * gviz_planar.h -- planarity and the largest face of a planar embedding
 * (host side of the Tutte embedders' Begin).
 */
#ifndef GVIZ_HOST_PLANAR_H
#define GVIZ_HOST_PLANAR_H

#include <stddef.h>
#include <stdint.h>

/* The graph in CSR (both directions of every edge; self-loops and repeated
 * edges are ignored). Writes the vertex walk of the largest face (at least
 * three darts; the first in dart order among equals) of a planar embedding
 * into out[0..*count). 0, -2 if the graph is not planar, -1 on error or
 * when the face has more than max darts. */
int gviz_planar_largest_face(int64_t n, const int64_t *off, const int64_t *adj,
                             size_t *out, size_t max, size_t *count);

/* As gviz_planar_largest_face, and (when rotation and rotationCount are
 * not NULL) the rotation system behind the faces: rotation[off[v] ..
 * off[v] + rotationCount[v]) = v's neighbours in the cyclic order the face
 * walk follows (the dart after (u -> v) is v's entry before u), starting
 * at v's earliest adjacency entry; every edge once, loops and repeated
 * edges left out. */
int gviz_planar_embed(int64_t n, const int64_t *off, const int64_t *adj, size_t *out,
                      size_t max, size_t *count, int64_t *rotation,
                      int64_t *rotationCount);

struct gvizEmbeddedGraph;
/* Tutte Begin's planarity step on an embedded graph: the rotation system
 * written into the graph's adjacency lists (each subgraph vertex's list
 * becomes its rotation followed by its entries that are not subgraph
 * edges), the layout and a full subgraph's edge bits rebuilt, and
 * planarEmbedded set; the largest face goes to out. 0, -2 not planar, -1. */
int gviz_planar_embed_largest_face(struct gvizEmbeddedGraph *e, size_t *out, size_t max,
                                   size_t *count);
/* FixOuterFace's boundary: the face walk from the highlight's first edge
 * through the graph's rotation. 0, or -1 (no highlight, no edge in it, a
 * walk longer than max or shorter than 3). */
int gviz_planar_face_from_highlight(const struct gvizEmbeddedGraph *e, size_t *out,
                                    size_t max, size_t *count);

#endif
