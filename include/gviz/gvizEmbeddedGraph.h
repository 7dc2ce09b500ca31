/*
 * gvizEmbeddedGraph.h -- the shared base of every embedder in the host layer:
 * FP64 positions indexed by parent vertex id, named actions, named stat
 * series and a draw mask (reference: include/embedders/gvizEmbeddedGraph.h).
 * Every embedder state starts with this struct, so a state pointer can be
 * used as a gvizEmbeddedGraph pointer.
 */
#ifndef GVIZ_HOST_EMBEDDED_GRAPH_H
#define GVIZ_HOST_EMBEDDED_GRAPH_H

#include "gvizGraph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gvizEmbedding {
  size_t dim;
  double *vertexPositions; /* nParent * dim, vertex-major */
} gvizEmbedding;

struct gvizEmbeddedGraph;

typedef struct gvizActionPayload {
  double worldX, worldY;
  double deltaTime;
  int64_t iarg;
  double darg;
} gvizActionPayload;

typedef void (*gvizActionHandler)(struct gvizEmbeddedGraph *embedding,
                                  void *userData,
                                  const gvizActionPayload *payload);

typedef struct gvizAction {
  const char *name;
  gvizActionHandler handler;
  void *userData;
} gvizAction;

typedef struct gvizActionRegistry {
  gvizAction *actions;
  size_t count;
  size_t capacity;
} gvizActionRegistry;

typedef enum gvizStatChartKind {
  GVIZ_STAT_CHART_LINE = 0,
  GVIZ_STAT_CHART_LINE_LOG = 1,
} gvizStatChartKind;

typedef struct gvizStatSeries {
  const char *name; /* not copied: must outlive the registry */
  gvizStatChartKind kind;
  double *samples;
  size_t count;
  size_t capacity;
  uint64_t revision;
} gvizStatSeries;

typedef struct gvizStatRegistry {
  gvizStatSeries *series;
  size_t count;
  size_t capacity;
} gvizStatRegistry;

typedef enum gvizDrawEdgePolicy {
  GVIZ_DRAW_EDGES_ALL = 0,
  GVIZ_DRAW_EDGES_NONE = 1,
  GVIZ_DRAW_EDGES_IF_BOTH_VISIBLE = 2,
} gvizDrawEdgePolicy;

typedef struct gvizDrawMask {
  gvizVertexSubset visibleVertices; /* owned; all subgraph vertices after Init */
  gvizDrawEdgePolicy edgePolicy;
  uint64_t revision; /* bumped on reset, policy change and notification */
} gvizDrawMask;

typedef struct gvizEmbeddedGraph {
  gvizSubgraph subgraph;      /* owned */
  gvizEmbedding embedding;
  gvizActionRegistry actions;
  gvizDrawMask drawMask;
  gvizStatRegistry stats;
  int planarEmbedded;         /* 1 once a Tutte Begin installed its rotation */
  gvizSubgraph highlight;     /* a view the front end highlights (owned) */
} gvizEmbeddedGraph;

int gvizEmbeddedGraphInit(gvizEmbeddedGraph *e, gvizSubgraph subgraph,
                          size_t dimension);
/* The embedded subgraph / the draw mask (read-only). */
const gvizSubgraph *gvizEmbeddedGraphStructure(const gvizEmbeddedGraph *e);
const gvizDrawMask *gvizEmbeddedGraphGetDrawMask(const gvizEmbeddedGraph *e);
/* Grows the underlying graph: a bare vertex (shown in the subgraph and the
 * mask; positions zero) or an edge (shown when the subgraph is full).
 * Embedders keep the graph they saw at Init / Begin. 0 / -1. */
int gvizEmbeddedGraphAddVertex(gvizEmbeddedGraph *e, void *data);
int gvizEmbeddedGraphAddEdge(gvizEmbeddedGraph *e, size_t from, size_t to);
/* The highlighted view (a face picked in a viewer; Set takes ownership). */
void gvizEmbeddedGraphSetHighlight(gvizEmbeddedGraph *e, gvizSubgraph highlight);
void gvizEmbeddedGraphClearHighlight(gvizEmbeddedGraph *e);
int gvizEmbeddedGraphHasHighlight(const gvizEmbeddedGraph *e);
const gvizSubgraph *gvizEmbeddedGraphGetHighlight(const gvizEmbeddedGraph *e);
void gvizEmbeddedGraphRelease(gvizEmbeddedGraph *e);
size_t gvizEmbeddedGraphDim(const gvizEmbeddedGraph *e);
size_t gvizEmbeddedGraphPositionCount(const gvizEmbeddedGraph *e);
const double *gvizEmbeddedGraphPositions(const gvizEmbeddedGraph *e);
double *gvizEmbeddedGraphGetVPosition(gvizEmbeddedGraph *e, size_t u);
void gvizEmbeddedGraphSetVPosition(gvizEmbeddedGraph *e, size_t u,
                                   const double *pos);
void gvizEmbeddedGraphAddVPosition(gvizEmbeddedGraph *e, size_t u,
                                   const double *delta);
/* Positions uniform in [-boxExtent, boxExtent]^dim from glibc rand_r, x then
 * y per subgraph vertex in ascending id; seed 0 means time(NULL)
 * (src/embedders/gvizEmbeddedGraph.c:406-426). */
void gvizEmbeddedGraphRandomizePositions(gvizEmbeddedGraph *e,
                                         double boxExtent, unsigned int seed);

int gvizEmbeddedGraphAddAction(gvizEmbeddedGraph *e, const char *name,
                               gvizActionHandler handler, void *userData);
int gvizEmbeddedGraphRemoveAction(gvizEmbeddedGraph *e, const char *name);
const gvizAction *gvizEmbeddedGraphFindAction(const gvizEmbeddedGraph *e,
                                              const char *name);
size_t gvizEmbeddedGraphActionCount(const gvizEmbeddedGraph *e);
const gvizAction *gvizEmbeddedGraphActionAt(const gvizEmbeddedGraph *e,
                                            size_t idx);
int gvizEmbeddedGraphInvokeAction(gvizEmbeddedGraph *e, const char *name,
                                  const gvizActionPayload *payload);

gvizStatSeries *gvizEmbeddedGraphAddStatSeries(gvizEmbeddedGraph *e,
                                               const char *name,
                                               gvizStatChartKind kind);
/* Creates the series (GVIZ_STAT_CHART_LINE) when it is not registered. */
int gvizEmbeddedGraphStatAppend(gvizEmbeddedGraph *e, const char *name,
                                double value);
void gvizEmbeddedGraphStatClear(gvizEmbeddedGraph *e, const char *name);
size_t gvizEmbeddedGraphStatSeriesCount(const gvizEmbeddedGraph *e);
const gvizStatSeries *gvizEmbeddedGraphStatSeriesAt(const gvizEmbeddedGraph *e,
                                                    size_t idx);
const gvizStatSeries *gvizEmbeddedGraphFindStatSeries(const gvizEmbeddedGraph *e,
                                                      const char *name);

void gvizEmbeddedGraphDrawMaskShowVertex(gvizEmbeddedGraph *e, size_t u);
void gvizEmbeddedGraphDrawMaskHideVertex(gvizEmbeddedGraph *e, size_t u);
void gvizEmbeddedGraphDrawMaskClearVertices(gvizEmbeddedGraph *e);
void gvizEmbeddedGraphDrawMaskNotifyChanged(gvizEmbeddedGraph *e);
/* Every subgraph vertex visible, edge policy ALL. */
void gvizEmbeddedGraphResetDrawMask(gvizEmbeddedGraph *e);
uint64_t gvizEmbeddedGraphDrawMaskRevision(const gvizEmbeddedGraph *e);
void gvizEmbeddedGraphSetDrawMaskEdgePolicy(gvizEmbeddedGraph *e,
                                            gvizDrawEdgePolicy policy);
/* Subgraph vertex whose mask bit is set. */
bool gvizEmbeddedGraphIsVertexVisible(const gvizEmbeddedGraph *e, size_t u);
/* By edge policy: ALL = edge of the subgraph, NONE = never,
 * IF_BOTH_VISIBLE = both endpoints visible. */
bool gvizEmbeddedGraphIsEdgeVisible(const gvizEmbeddedGraph *e, size_t u,
                                    size_t v);

/* Text checkpoint (src/embedders/gvizEmbeddedGraph.c:428-485): "name\n",
 * "N dim\n", then each coordinate "%f " per vertex line. */
int gvizEmbeddedGraphSaveEmbedding(gvizEmbeddedGraph *e, const char *name,
                                   const char *filename);
int gvizEmbeddedGraphLoadEmbedding(gvizEmbeddedGraph *e,
                                   const char *filename);
int gvizEmbeddedGraphIsPlanarEmbedded(const gvizEmbeddedGraph *e);

#ifdef __cplusplus
}
#endif
#endif
