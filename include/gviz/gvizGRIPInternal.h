/*
 * gvizGRIPInternal.h -- the GRIP pipeline stages the reference exposes for
 * white-box tests and benches (src/embedders/gvizGRIPInternal.h), here over
 * the device filtration (gvizhip_grip_mis_first / _mis_layer / _migrate /
 * _mis_filtration). They read and write the state's misFiltration and
 * misBorder exactly as the reference's do; the first call creates the
 * state's device engine.
 */
#ifndef GVIZ_HOST_GRIP_INTERNAL_H
#define GVIZ_HOST_GRIP_INTERNAL_H

#include "gvizGRIPEmbedder.h"

#ifdef __cplusplus
extern "C" {
#endif

static inline void gvizGRIPSwap(size_t *a, size_t *b) {
  if (a != b) {
    size_t t = *a;
    *a = *b;
    *b = t;
  }
}

/* Layer 1: shows its vertices in `out`, writes the others backwards from
 * the end of misFiltration, appends the layer's size to misBorder. */
void makeFirstMISPartition(gvizGRIPState *state, gvizVertexSubset out);
/* Layer i >= 2 from the candidates in `vertices` (replaced by the chosen);
 * the rest are written backwards from misFiltration[misBorder[i-1] - 1].
 * Returns 1 while the layer holds more than dim + 1 vertices, else 0. */
int iterMISFiltration(gvizGRIPState *state, size_t i, gvizVertexSubset vertices);
/* Moves the candidate farthest from the final layer (count - 1) into it.
 * 0, or -1 when every pool is exhausted. finalLayer is unused, as in the
 * reference. */
int migrateOneToFinalLayer(gvizGRIPState *state, GVIZ_BIT_ARRAY finalLayer, size_t count);
/* The whole filtration: misFiltration and misBorder (reset to the active
 * count first) as Begin builds them; returns the layer count (0 on failure). */
size_t createMISFiltration(gvizGRIPState *state);

#ifdef __cplusplus
}
#endif
#endif
