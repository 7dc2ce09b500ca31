/*
 * gvizGraphLoader.h -- .edges file ingestion for the host layer (libgviz.so).
 *
 * gvizEdgesFileOptions, gvizEdgesFileOptionsInit and gvizGraphLoadFromEdgesFile
 * keep the reference's names, options and results
 * (include/utils/graphLoader.h:7-34, src/utils/graphLoader.c:10-214,790-825):
 * one edge per line ("u v", anything after the second id ignored), blank and
 * '%' lines skipped, optionally the first data line skipped as a header,
 * negative ids or a line without two integers rejected, external ids of any
 * base compacted to dense ids in order of first appearance (u before v), and
 * each edge added in file order (undirected: both directions), so every
 * adjacency list is in insertion order.
 *
 * gvizGraphLoadEdgesFileCSR has no reference counterpart: the same graph
 * straight into the CSR arrays the device engines take (gvizhip_*_create),
 * without building per-vertex lists.
 */
#ifndef GVIZ_HOST_GRAPH_LOADER_H
#define GVIZ_HOST_GRAPH_LOADER_H

#include <stdint.h>

#include "gvizGraph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gvizEdgesFileOptions {
  int directed;    /* non-zero: directed edges */
  int skip_header; /* non-zero: skip the first non-comment line */
} gvizEdgesFileOptions;

/* Network Repository defaults: undirected, no header. */
void gvizEdgesFileOptionsInit(gvizEdgesFileOptions *opts);

/* Loads `path` into `out` (uninitialised on entry; zeroed on failure).
 * 0 on success, -1 on I/O, parse or allocation failure. */
int gvizGraphLoadFromEdgesFile(const char *path, const gvizEdgesFileOptions *opts,
                               gvizGraph *out);

/* The same graph as CSR over the dense ids: *n vertices, offsets[n+1] and
 * adj[offsets[n]] (malloc'ed; free()), each row in insertion order.
 * 0 on success, -1 on failure (outputs untouched). */
int gvizGraphLoadEdgesFileCSR(const char *path, const gvizEdgesFileOptions *opts,
                              int64_t *n, int64_t **offsets, int64_t **adj);

#ifdef __cplusplus
}
#endif
#endif
