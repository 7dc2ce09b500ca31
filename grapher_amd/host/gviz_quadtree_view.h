/*
 * gviz_quadtree_view.h -- the host layer's internal hooks for host views of
 * device quadtrees (gviz_quadtree.c), used by the force embedder to fill
 * its state's `quadtree` on demand.
 */
#ifndef GVIZ_HOST_QUADTREE_VIEW_H
#define GVIZ_HOST_QUADTREE_VIEW_H

#include "gviz/gvizQuadtree.h"

#include <stdint.h>

/* Empties t and readies its arena arrays (no points yet). 0 / -1. */
int gviz_quadtree_prepare(gvizQuadtree *t, size_t nodesPerCell);
/* Lays out t's nodes from device keys: root[4] = centre x, centre y, half
 * size, path levels; order/keyHi/keyLo[t->pointCount] = point ids and
 * their paths in path order. t->points / masses / pointCount /
 * nodesPerCell must be set. 0 / -1. */
int gviz_quadtree_from_keys(gvizQuadtree *t, const double *root, const uint32_t *order,
                            const uint64_t *keyHi, const uint64_t *keyLo);
/* gvizQuadtreeRoot(t) calls refresh(owner) first whenever t is marked
 * stale (set here and by gviz_quadtree_mark_stale). */
void gviz_quadtree_set_refresh(gvizQuadtree *t, int (*refresh)(void *), void *owner);
void gviz_quadtree_mark_stale(gvizQuadtree *t);
const char *gviz_quadtree_last_error(void);

#endif
