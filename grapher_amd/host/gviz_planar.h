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

#endif
