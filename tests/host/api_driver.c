/*
 * api_driver.c -- drives this repository's C11 host layer (libgviz.so,
 * include/gviz/*.h) for the tests.
 *
 *   api_driver unit                 host-side unit checks (no GPU needed)
 *   api_driver <mode> <in> <out>    mode = api_force | api_grip | tutte
 *
 * The three run modes read and write the same framing as the reference-side
 * harness that wrote the golden fixtures (tests/golden/refio.py), so a test
 * feeds the fixture's inputs here and compares the outputs with the
 * reference's. GVIZ_TEST_PRECISION = fp64 (default) | mixed | fp32 selects
 * the force precision; GVIZ_TEST_RESIDENT=1 keeps force positions on the
 * device between Begin and the end of Run.
 */
#include "gviz/gvizForceEmbedder.h"
#include "gviz/gvizGRIPEmbedder.h"
#include "gviz/gvizGRIPInternal.h"
#include "gviz/gvizTutteEmbedder.h"
#include "gviz/gvizSpringTutteEmbedder.h"
#include "gviz/gvizGraphLoader.h"
#include "gviz/gvizQuadtree.h"
#include "gviz/gvizKNearest.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static FILE *fin, *fout;

static int64_t rd_i(void) {
  int64_t v;
  if (fread(&v, 8, 1, fin) != 1) {
    fprintf(stderr, "short read\n");
    exit(2);
  }
  return v;
}
static double rd_d(void) {
  double v;
  if (fread(&v, 8, 1, fin) != 1) {
    fprintf(stderr, "short read\n");
    exit(2);
  }
  return v;
}
static void wr_i(int64_t v) { fwrite(&v, 8, 1, fout); }
static void wr_d(double v) { fwrite(&v, 8, 1, fout); }

static void read_graph(gvizGraph *g, int64_t *nOut) {
  int64_t n = rd_i(), m = rd_i();
  gvizGraphInit(g, 0);
  for (int64_t i = 0; i < n; i++)
    gvizGraphAddVertex(g, NULL, NULL, NULL);
  for (int64_t e = 0; e < m; e++) {
    int64_t u = rd_i(), v = rd_i();
    gvizGraphAddEdge(g, (size_t)u, (size_t)v);
  }
  gvizGraphBuildLayout(g);
  *nOut = n;
}

static gvizForcePrecision env_precision(void) {
  const char *p = getenv("GVIZ_TEST_PRECISION");
  if (p && !strcmp(p, "mixed"))
    return GVIZ_FORCE_PRECISION_MIXED;
  if (p && !strcmp(p, "fp32"))
    return GVIZ_FORCE_PRECISION_FP32;
  return GVIZ_FORCE_PRECISION_FP64;
}

static void write_series(gvizEmbeddedGraph *e, const char *name) {
  const gvizStatSeries *s = gvizEmbeddedGraphFindStatSeries(e, name);
  wr_i(s ? (int64_t)s->count : -1);
  if (s)
    for (size_t i = 0; i < s->count; i++)
      wr_d(s->samples[i]);
}

static int fail_device(const char *what) {
  fprintf(stderr, "%s failed: %s\n", what, gvizForceEmbedderLastError());
  return 1;
}

/* Same noise as the reference-side harness (coordinate *= 1 + amount * u). */
static void perturb_positions(double *p, int64_t count, double amount, int64_t seed) {
  if (!(amount > 0.0))
    return;
  uint64_t st = (uint64_t)seed * 2654435761u + 1u;
  for (int64_t i = 0; i < count; i++) {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    const double u = (double)(st >> 11) * 0x1p-53 * 2.0 - 1.0;
    p[i] *= 1.0 + amount * u;
  }
}

static int mode_api_force(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t nHidden = rd_i();
  gvizVertexSubset vs = gvizVertexSubsetCreateEmpty(&g);
  for (int64_t i = 0; i < n; i++)
    gvizVertexSubsetShowVertex(vs, (size_t)i);
  for (int64_t i = 0; i < nHidden; i++)
    gvizVertexSubsetHideVertex(vs, (size_t)rd_i());
  int64_t model = rd_i(), bh = rd_i(), overlap = rd_i(), npc = rd_i();
  int64_t seed = rd_i(), runIters = rd_i();
  double edgeLength = rd_d(), theta = rd_d(), gravityK = rd_d();
  double overlapC = rd_d(), rBase = rd_d(), rPerDeg = rd_d(), jitter = rd_d();
  double eps = rd_d();
  double perturb = rd_d();
  int64_t perturbSeed = rd_i();
  gvizForceEmbedderState s;
  if (gvizForceEmbedderInit(&s, gvizSubgraphCreateVertexInduced(&g, vs), 2,
                            (gvizForceModelKind)model) != 0)
    return 1;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)&s;
  gvizForceEmbedderSetPrecision(&s, env_precision());
  int resident = getenv("GVIZ_TEST_RESIDENT") && atoi(getenv("GVIZ_TEST_RESIDENT"));
  gvizForceEmbedderConfigure(&s, edgeLength, 0.0);
  gvizForceEmbedderConfigureSpeed(&s, jitter);
  gvizForceEmbedderConfigureBarnesHut(&s, theta, (size_t)npc);
  gvizForceEmbedderSetBarnesHutEnabled(&s, (int)bh);
  gvizForceEmbedderConfigureGravity(&s, gravityK);
  gvizForceEmbedderConfigureRadius(&s, rBase, rPerDeg);
  gvizForceEmbedderConfigureOverlapPrevention(&s, overlapC);
  gvizForceEmbedderSetPreventOverlapEnabled(&s, (int)overlap);
  wr_d(s.boxExtent);
  int rc = gvizForceEmbedderBegin(&s, (unsigned int)seed);
  if (rc != 0)
    return fail_device("gvizForceEmbedderBegin");
  wr_i(rc);
  const double *p = gvizEmbeddedGraphPositions(e);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  perturb_positions(e->embedding.vertexPositions, 2 * n, perturb, perturbSeed);
  gvizForceEmbedderSetResidentPositions(&s, resident);
  int rounds = gvizForceEmbedderRun(&s, (size_t)runIters, eps);
  if (rounds < 0)
    return fail_device("gvizForceEmbedderRun");
  if (resident && gvizForceEmbedderSyncPositions(&s) != 0)
    return fail_device("gvizForceEmbedderSyncPositions");
  gvizForceEmbedderSetResidentPositions(&s, 0);
  wr_i(rounds);
  static const char *const names[] = {
      "forceEmbedder.maxDisp", "forceEmbedder.speed",
      "forceEmbedder.attractiveForce", "forceEmbedder.repulsiveForce",
      "forceEmbedder.gravityForce"};
  for (int k = 0; k < 5; k++)
    write_series(e, names[k]);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  double d = gvizForceEmbedderStep(&s);
  if (d < 0.0)
    return fail_device("gvizForceEmbedderStep");
  wr_d(d);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  wr_d(s.globalSpeed);
  wr_d(s.speedEfficiency);
  wr_i((int64_t)s.iteration);
  gvizForceEmbedderRelease(&s);
  gvizGraphRelease(&g);
  return 0;
}

static int mode_api_grip(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t kcap = rd_i(), pk = rd_i(), rk = rd_i(), policy = rd_i(), dim = rd_i();
  gvizGRIPState st;
  memset(&st, 0, sizeof(st));
  gvizGRIPEmbedderConfigureKnnCapacity(&st, (size_t)kcap);
  if (gvizGRIPEmbedderInit(&st, gvizSubgraphCreateFull(&g), 0, (size_t)dim) != 0)
    return 1;
  gvizGRIPEmbedderConfigureK(&st, (size_t)pk, (size_t)rk, (gvizGRIPKPolicy)policy);
  int rc = gvizGRIPEmbedderEmbed(&st);
  if (rc != 0)
    return fail_device("gvizGRIPEmbedderEmbed");
  wr_i(rc);
  wr_i((int64_t)st.layerCount);
  for (size_t i = 0; i < st.layerCount; i++)
    wr_i((int64_t) * (size_t *)gvizArrayAtIndex(&st.misBorder, i));
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)&st;
  const double *p = gvizEmbeddedGraphPositions(e);
  for (int64_t i = 0; i < dim * n; i++)
    wr_d(p[i]);
  static const char *const names[] = {"grip.heat", "grip.meanDisp",
                                      "grip.maxDisp", "grip.meanForce"};
  for (int k = 0; k < 4; k++)
    write_series(e, names[k]);
  for (int64_t v = 0; v < n; v++)
    wr_i(gvizEmbeddedGraphIsVertexVisible(e, (size_t)v) ? 1 : 0);
  gvizGRIPEmbedderRelease(&st);
  gvizGraphRelease(&g);
  return 0;
}

static int mode_tutte(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t nb = rd_i(), iters = rd_i(), gs = rd_i();
  size_t *boundary = malloc(sizeof(size_t) * nb);
  for (int64_t i = 0; i < nb; i++)
    boundary[i] = (size_t)rd_i();
  double radius = rd_d(), dt = rd_d();
  gvizTutteState s;
  if (gvizTutteEmbedderInit(&s, gvizSubgraphCreateFull(&g), 2, 1e-300) != 0)
    return 1;
  s.useGaussSeidel = (int)gs;
  gvizTutteFixConvexPolygon(&s, boundary, (size_t)nb, radius);
  gvizTutteEmbedderSeedInterior(&s);
  const double *p = gvizEmbeddedGraphPositions((gvizEmbeddedGraph *)&s);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  for (int64_t it = 0; it < iters; it++) {
    double d = gvizTutteEmbedderStep(&s, dt);
    if (d < 0.0)
      return fail_device("gvizTutteEmbedderStep");
    wr_d(d);
  }
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  gvizTutteEmbedderRelease(&s);
  gvizGraphRelease(&g);
  free(boundary);
  return 0;
}

/* ------------------------------------------------------------ unit checks */

static int failures;
#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      fprintf(stderr, "%s:%d: %s: check failed: %s\n", __FILE__, __LINE__, \
              __func__, #cond);                                            \
      failures++;                                                          \
    }                                                                      \
  } while (0)

static void make_path(gvizGraph *g, size_t n) {
  gvizGraphInit(g, 0);
  for (size_t i = 0; i < n; i++)
    gvizGraphAddVertex(g, NULL, NULL, NULL);
  for (size_t i = 0; i + 1 < n; i++)
    gvizGraphAddEdge(g, i, i + 1);
  gvizGraphBuildLayout(g);
}

static void test_graph_insertionOrderAndCSR(void) {
  gvizGraph g;
  gvizGraphInit(&g, 0);
  for (int i = 0; i < 5; i++)
    gvizGraphAddVertex(&g, NULL, NULL, NULL);
  gvizGraphAddEdge(&g, 0, 3);
  gvizGraphAddEdge(&g, 0, 1);
  gvizGraphAddEdge(&g, 2, 0);
  gvizGraphAddEdge(&g, 3, 4);
  CHECK(gvizGraphSize(&g) == 5);
  CHECK(gvizGraphEdgeCount(&g) == 4);
  size_t c;
  const size_t *nb = gvizGraphNeighbors(&g, 0, &c);
  CHECK(c == 3 && nb[0] == 3 && nb[1] == 1 && nb[2] == 2);
  CHECK(gvizGraphAddEdge(&g, 0, 9) == -1);

  gvizVertexSubset vs = gvizVertexSubsetCreateEmpty(&g);
  gvizVertexSubsetShowVertex(vs, 0);
  gvizVertexSubsetShowVertex(vs, 2);
  gvizVertexSubsetShowVertex(vs, 3);
  gvizSubgraph sg = gvizSubgraphCreateVertexInduced(&g, vs);
  CHECK(gvizSubgraphVertexCount(&sg) == 3);
  CHECK(gvizSubgraphDegree(&sg, 0) == 2);
  CHECK(gvizSubgraphDegree(&sg, 1) == 0);
  CHECK(gvizSubgraphEdgeCount(&sg) == 2);
  int64_t *off, *adj;
  size_t *compact, count;
  CHECK(gvizSubgraphToCompactCSR(&sg, &off, &adj, &compact, &count) == 0);
  CHECK(count == 3 && compact[0] == 0 && compact[1] == 2 && compact[2] == 3);
  /* vertex 0 -> (3, 2) in insertion order: compact ids 2, 1 */
  CHECK(off[0] == 0 && off[1] == 2 && adj[0] == 2 && adj[1] == 1);
  CHECK(off[3] == 4);
  free(off);
  free(adj);
  free(compact);
  CHECK(gvizSubgraphToCSR(&sg, &off, &adj) == 0);
  CHECK(off[1] - off[0] == 2 && off[2] == off[1] && off[5] == 4);
  free(off);
  free(adj);
  gvizSubgraphRelease(&sg);
  /* a full subgraph needs the layout, as in the reference */
  gvizSubgraph none = gvizSubgraphCreateFull(&g);
  CHECK(none.g == NULL && none.vs == NULL);
  gvizGraphBuildLayout(&g);
  CHECK(g.layout && g.layout->nvertices == 5 && g.layout->edgeCount == 8);
  CHECK(g.layout->vertexOffsets[1] == 3 && g.layout->vertexOffsets[5] == 8);
  gvizSubgraph full = gvizSubgraphCreateFull(&g);
  CHECK(gvizSubgraphIsFull(&full) && gvizSubgraphVertexCount(&full) == 5);
  CHECK(gvizSubgraphEdgeCount(&full) == 4 && gvizSubgraphDegree(&full, 0) == 3);
  CHECK(gvizSubgraphHasEdge(&full, 3, 4) && !gvizSubgraphHasEdge(&full, 1, 2));
  gvizSubgraphHideVertex(&full, 3);
  CHECK(gvizSubgraphVertexCount(&full) == 4 && gvizSubgraphEdgeCount(&full) == 2);
  CHECK(gvizSubgraphDegree(&full, 0) == 2 && gvizSubgraphDegree(&full, 4) == 0);
  /* bits are least significant first, one byte per eight vertices */
  CHECK(full.vs[0] == (unsigned char)(0x1F & ~(1u << 3)));
  gvizSubgraphRelease(&full);
  gvizArray in;
  gvizArrayInit(&in, sizeof(size_t));
  size_t u0 = 0, u4 = 4;
  gvizArrayPush(&in, &u0);
  gvizArrayPush(&in, &u4);
  CHECK(gvizGraphAddVertex(&g, NULL, &in, NULL) == 0 && gvizGraphSize(&g) == 6);
  CHECK(gvizGraphEdgeExists(&g, 0, 5) && gvizGraphEdgeExists(&g, 4, 5) &&
        !gvizGraphEdgeExists(&g, 5, 0));
  gvizArrayRelease(&in);
  gvizGraphRelease(&g);
}

static int actionCalls;
static void count_action(gvizEmbeddedGraph *e, void *u, const gvizActionPayload *p) {
  (void)e;
  actionCalls += *(int *)u + (int)p->iarg;
}

static void test_embeddedGraph_actionsStatsMask(void) {
  gvizGraph g;
  make_path(&g, 4);
  gvizEmbeddedGraph e;
  CHECK(gvizEmbeddedGraphInit(&e, gvizSubgraphCreateFull(&g), 2) == 0);
  for (size_t u = 0; u < 4; u++) {
    const double *p = gvizEmbeddedGraphGetVPosition(&e, u);
    CHECK(p[0] == 0.0 && p[1] == 0.0);
  }
  int one = 1, ten = 10;
  CHECK(gvizEmbeddedGraphAddAction(&e, "a", count_action, &one) == 0);
  CHECK(gvizEmbeddedGraphAddAction(&e, "b", count_action, &one) == 0);
  CHECK(gvizEmbeddedGraphAddAction(&e, "a", count_action, &ten) == 0);
  CHECK(gvizEmbeddedGraphActionCount(&e) == 2);
  CHECK(gvizEmbeddedGraphInvokeAction(&e, "a", NULL) == 0 && actionCalls == 10);
  gvizActionPayload pl = {0};
  pl.iarg = 5;
  CHECK(gvizEmbeddedGraphInvokeAction(&e, "b", &pl) == 0 && actionCalls == 16);
  CHECK(gvizEmbeddedGraphInvokeAction(&e, "zz", NULL) == -1);
  CHECK(gvizEmbeddedGraphRemoveAction(&e, "a") == 0);
  CHECK(gvizEmbeddedGraphRemoveAction(&e, "a") == -1);
  CHECK(gvizEmbeddedGraphActionAt(&e, 0) && !gvizEmbeddedGraphActionAt(&e, 1));

  CHECK(gvizEmbeddedGraphAddStatSeries(&e, "s", GVIZ_STAT_CHART_LINE_LOG) != NULL);
  CHECK(gvizEmbeddedGraphStatAppend(&e, "s", 1.5) == 0);
  CHECK(gvizEmbeddedGraphStatAppend(&e, "new", 2.5) == 0);
  const gvizStatSeries *sn = gvizEmbeddedGraphFindStatSeries(&e, "new");
  CHECK(sn && sn->kind == GVIZ_STAT_CHART_LINE && sn->count == 1 && sn->samples[0] == 2.5);
  for (int i = 0; i < 1000; i++)
    gvizEmbeddedGraphStatAppend(&e, "s", (double)i);
  const gvizStatSeries *ss = gvizEmbeddedGraphFindStatSeries(&e, "s");
  CHECK(ss->count == 1001 && ss->samples[1000] == 999.0);
  gvizEmbeddedGraphStatClear(&e, "s");
  CHECK(ss->count == 0 && gvizEmbeddedGraphStatSeriesCount(&e) == 2);

  uint64_t r0 = gvizEmbeddedGraphDrawMaskRevision(&e);
  CHECK(gvizEmbeddedGraphIsVertexVisible(&e, 3));
  gvizEmbeddedGraphDrawMaskHideVertex(&e, 3);
  CHECK(!gvizEmbeddedGraphIsVertexVisible(&e, 3));
  CHECK(gvizEmbeddedGraphDrawMaskRevision(&e) == r0);
  CHECK(gvizEmbeddedGraphIsEdgeVisible(&e, 2, 3));
  gvizEmbeddedGraphSetDrawMaskEdgePolicy(&e, GVIZ_DRAW_EDGES_IF_BOTH_VISIBLE);
  CHECK(gvizEmbeddedGraphDrawMaskRevision(&e) == r0 + 1);
  CHECK(!gvizEmbeddedGraphIsEdgeVisible(&e, 2, 3));
  CHECK(gvizEmbeddedGraphIsEdgeVisible(&e, 1, 2));
  CHECK(!gvizEmbeddedGraphIsEdgeVisible(&e, 0, 3));
  gvizEmbeddedGraphSetDrawMaskEdgePolicy(&e, GVIZ_DRAW_EDGES_NONE);
  CHECK(!gvizEmbeddedGraphIsEdgeVisible(&e, 1, 2));
  gvizEmbeddedGraphResetDrawMask(&e);
  CHECK(gvizEmbeddedGraphIsVertexVisible(&e, 3) && gvizEmbeddedGraphIsEdgeVisible(&e, 0, 1));
  CHECK(!gvizEmbeddedGraphIsEdgeVisible(&e, 0, 2));
  gvizEmbeddedGraphRelease(&e);
  gvizGraphRelease(&g);
}

static void test_embeddedGraph_randomizeSaveLoad(void) {
  gvizGraph g;
  make_path(&g, 50);
  gvizEmbeddedGraph e, f;
  gvizEmbeddedGraphInit(&e, gvizSubgraphCreateFull(&g), 2);
  gvizEmbeddedGraphInit(&f, gvizSubgraphCreateFull(&g), 2);
  gvizEmbeddedGraphRandomizePositions(&e, 7.0, 99);
  gvizEmbeddedGraphRandomizePositions(&f, 7.0, 99);
  int inBox = 1, same = 1, spread = 0;
  for (size_t u = 0; u < 50; u++) {
    const double *p = gvizEmbeddedGraphGetVPosition(&e, u);
    const double *q = gvizEmbeddedGraphGetVPosition(&f, u);
    inBox &= fabs(p[0]) <= 7.0 && fabs(p[1]) <= 7.0;
    same &= p[0] == q[0] && p[1] == q[1];
    spread |= fabs(p[0]) > 3.5;
  }
  CHECK(inBox && same && spread);
  char path[] = "/tmp/gviz_embedding_XXXXXX";
  int fd = mkstemp(path);
  CHECK(fd >= 0);
  if (fd >= 0)
    close(fd);
  CHECK(gvizEmbeddedGraphSaveEmbedding(&e, "layout", path) == 0);
  gvizEmbeddedGraphRandomizePositions(&f, 1.0, 5);
  CHECK(gvizEmbeddedGraphLoadEmbedding(&f, path) == 0);
  double worst = 0.0;
  for (size_t u = 0; u < 50; u++) {
    const double *p = gvizEmbeddedGraphGetVPosition(&e, u);
    const double *q = gvizEmbeddedGraphGetVPosition(&f, u);
    worst = fmax(worst, fmax(fabs(p[0] - q[0]), fabs(p[1] - q[1])));
  }
  CHECK(worst <= 5e-7); /* "%f" keeps six decimals */
  unlink(path);
  gvizEmbeddedGraph h;
  gvizGraph g2;
  make_path(&g2, 3);
  gvizEmbeddedGraphInit(&h, gvizSubgraphCreateFull(&g2), 2);
  CHECK(gvizEmbeddedGraphSaveEmbedding(&e, "x", path) == 0);
  CHECK(gvizEmbeddedGraphLoadEmbedding(&h, path) == -1);
  unlink(path);
  gvizEmbeddedGraphRelease(&h);
  gvizEmbeddedGraphRelease(&e);
  gvizEmbeddedGraphRelease(&f);
  gvizGraphRelease(&g);
  gvizGraphRelease(&g2);
}

/* Mirrors tests/embedders/gvizForceEmbedderTests.c (init, configure). */
static void test_forceEmbedder_initConfigure(void) {
  gvizGraph g;
  make_path(&g, 10);
  gvizForceEmbedderState s;
  CHECK(gvizForceEmbedderInit(&s, gvizSubgraphCreateFull(&g), 3,
                              GVIZ_FORCE_MODEL_FRUCHTERMAN_REINGOLD) == -2);
  CHECK(gvizForceEmbedderInit(&s, gvizSubgraphCreateFull(&g), 2,
                              GVIZ_FORCE_MODEL_LINLOG) == 0);
  CHECK(s.vertexCount == 10);
  CHECK(s.edgeLength == GVIZ_FORCE_EMBEDDER_EDGE_LENGTH_DEFAULT);
  CHECK(s.boxExtent == 0.5 * sqrt(10.0) * 10.0);
  CHECK(s.theta == GVIZ_FORCE_EMBEDDER_THETA_DEFAULT && s.nodesPerCell == 1);
  CHECK(s.barnesHutEnabled == 1 && s.preventOverlap == 0 && s.gravityK == 0.0);
  CHECK(s.mass[0] == 2.0 && s.mass[5] == 3.0 && s.degree[5] == 2);
  CHECK(s.meanDegreePlusOne == (2.0 * 2 + 3.0 * 8) / 10.0);
  CHECK(gvizForceEmbedderPrecision(&s) == GVIZ_FORCE_PRECISION_MIXED);
  CHECK(s.disp == NULL && s.pool == NULL && s.quadtree.root == NULL);
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)&s;
  CHECK(gvizEmbeddedGraphFindAction(e, "forceEmbedder.step") != NULL);
  CHECK(gvizEmbeddedGraphFindAction(e, "forceEmbedder.toggleOverlapPrevention") != NULL);
  const char *series[] = {"forceEmbedder.maxDisp", "forceEmbedder.speed",
                          "forceEmbedder.attractiveForce",
                          "forceEmbedder.repulsiveForce",
                          "forceEmbedder.gravityForce"};
  for (int i = 0; i < 5; i++) {
    const gvizStatSeries *ss = gvizEmbeddedGraphFindStatSeries(e, series[i]);
    CHECK(ss && ss->kind == GVIZ_STAT_CHART_LINE_LOG && ss->count == 0);
  }
  gvizForceEmbedderConfigure(&s, 0.0, -1.0);
  CHECK(s.edgeLength == 10.0 && s.boxExtent == 0.5 * sqrt(10.0) * 10.0);
  gvizForceEmbedderConfigure(&s, 25.0, 100.0);
  CHECK(s.edgeLength == 25.0 && s.boxExtent == 100.0);
  gvizForceEmbedderConfigureSpeed(&s, 0.0);
  CHECK(s.jitterTolerance == 1.0);
  gvizForceEmbedderConfigureSpeed(&s, 2.5);
  CHECK(s.jitterTolerance == 2.5);
  gvizForceEmbedderConfigureBarnesHut(&s, 0.0, 0);
  CHECK(s.theta == 1.0 && s.nodesPerCell == 1);
  gvizForceEmbedderConfigureBarnesHut(&s, 0.3, 4);
  CHECK(s.theta == 0.3 && s.nodesPerCell == 4);
  gvizForceEmbedderConfigureOverlapPrevention(&s, -3.0);
  CHECK(s.overlapConstant == 100.0);
  gvizForceEmbedderConfigureRadius(&s, 0.5, 2.0);
  CHECK(fabs(gvizForceEmbedderVertexRadius(&s, 5) - 0.5 * (1.0 + 2.0 * sqrt(2.0))) < 1e-15);
  CHECK(gvizEmbeddedGraphInvokeAction(e, "forceEmbedder.toggleOverlapPrevention", NULL) == 0);
  CHECK(s.preventOverlap == 1);
  /* not begun: the step action is a no-op and Run refuses */
  CHECK(gvizEmbeddedGraphInvokeAction(e, "forceEmbedder.step", NULL) == 0);
  CHECK(gvizEmbeddedGraphFindStatSeries(e, "forceEmbedder.maxDisp")->count == 0);
  CHECK(gvizForceEmbedderRun(&s, 10, 0.0) == -1);
  gvizForceEmbedderRelease(&s);
  gvizGraphRelease(&g);
}

static void test_forceModel_table(void) {
  const gvizForceModel *fr = gvizForceModelGet(GVIZ_FORCE_MODEL_FRUCHTERMAN_REINGOLD);
  const gvizForceModel *ll = gvizForceModelGet(GVIZ_FORCE_MODEL_LINLOG);
  CHECK(fr->vertexMass(7) == 1.0 && ll->vertexMass(7) == 8.0);
  double v[2] = {0.0, 0.0}, u[2] = {3.0, 4.0}, acc[2] = {0.0, 0.0};
  fr->attractive(2, v, u, 10.0, acc); /* (d / k) * delta = 0.5 * (3, 4) */
  CHECK(fabs(acc[0] - 1.5) < 1e-15 && fabs(acc[1] - 2.0) < 1e-15);
  acc[0] = acc[1] = 0.0;
  fr->repulsive(2, v, u, 1.0, 2.0, 0.0, 0.0, 100.0, 10.0, acc);
  /* weight 2 * k^2 / d along -(3, 4) / 5 */
  CHECK(fabs(acc[0] + 2.0 * 100.0 / 5.0 * 0.6) < 1e-12);
  CHECK(fabs(acc[1] + 2.0 * 100.0 / 5.0 * 0.8) < 1e-12);
  acc[0] = acc[1] = 0.0;
  ll->repulsive(2, v, u, 2.0, 3.0, 1.0, 1.0, 100.0, 10.0, acc);
  /* mass product 6 over the gap 5 - 2 = 3 */
  CHECK(fabs(acc[0] + 6.0 / 3.0 * 0.6) < 1e-12 && fabs(acc[1] + 6.0 / 3.0 * 0.8) < 1e-12);
  acc[0] = acc[1] = 0.0;
  ll->repulsive(2, v, u, 2.0, 3.0, 3.0, 3.0, 100.0, 10.0, acc);
  CHECK(fabs(acc[0] + 600.0 * 0.6) < 1e-9); /* in contact: overlap constant */
  acc[0] = acc[1] = 0.0;
  ll->attractive(2, v, u, 10.0, acc);
  CHECK(fabs(acc[0] - log(6.0) / 5.0 * 3.0) < 1e-15);
}

static void test_tutte_boundaryAndSeed(void) {
  gvizGraph g;
  make_path(&g, 6);
  gvizGraphAddEdge(&g, 5, 0);
  gvizGraphBuildLayout(&g); /* edges changed: renumber them */
  gvizTutteState s;
  CHECK(gvizTutteEmbedderInit(&s, gvizSubgraphCreateFull(&g), 2, 0.0) == 0);
  CHECK(s.epsilon == GVIZ_TUTTE_DEFAULT_EPSILON && s.relaxationRate == 5.0);
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)&s;
  CHECK(gvizEmbeddedGraphFindAction(e, "tutte.step") != NULL);
  CHECK(gvizEmbeddedGraphFindAction(e, "tutte.fixOuterFace") != NULL);
  CHECK(gvizEmbeddedGraphFindStatSeries(e, "tutte.maxDelta") != NULL);
  CHECK(gvizTutteEmbedderRun(&s, 10) == -1);
  size_t two[2] = {0, 1}, bad[3] = {0, 1, 6}, tri[4] = {0, 2, 4, 5};
  double pos[8] = {0};
  CHECK(gvizTutteEmbedderSetBoundary(&s, two, 2, pos) == -1);
  CHECK(gvizTutteEmbedderSetBoundary(&s, bad, 3, pos) == -1);
  gvizTutteFixConvexPolygon(&s, tri, 4, 10.0);
  CHECK(s.boundaryCount == 4 && gvizTestBit(s.isBoundary, 2) && !gvizTestBit(s.isBoundary, 1));
  const double *p2 = gvizEmbeddedGraphGetVPosition(e, 2);
  CHECK(fabs(p2[0]) < 1e-12 && fabs(p2[1] - 10.0) < 1e-12);
  gvizTutteEmbedderSeedInterior(&s);
  const double *p1 = gvizEmbeddedGraphGetVPosition(e, 1);
  CHECK(fabs(p1[0]) < 1e-12 && fabs(p1[1]) < 1e-12);
  CHECK(gvizTutteEmbedderStep(&s, 0.0) == 0.0 && s.iteration == 0);
  /* Begin on the 6-cycle: planar, both faces are the cycle */
  CHECK(gvizTutteEmbedderBegin(&s) == 0 && s.boundaryCount == 6 && s.begun);
  gvizTutteEmbedderRelease(&s);
  gvizGraphRelease(&g);
  /* K5 is not planar */
  gvizGraph k5;
  gvizGraphInit(&k5, 0);
  for (size_t i = 0; i < 5; i++)
    gvizGraphAddVertex(&k5, NULL, NULL, NULL);
  for (size_t i = 0; i < 5; i++)
    for (size_t j = i + 1; j < 5; j++)
      gvizGraphAddEdge(&k5, i, j);
  gvizGraphBuildLayout(&k5);
  CHECK(gvizTutteEmbedderInit(&s, gvizSubgraphCreateFull(&k5), 2, 0.0) == 0);
  CHECK(gvizTutteEmbedderBegin(&s) == -2 && !s.begun);
  gvizTutteEmbedderRelease(&s);
  gvizGraphRelease(&k5);
}

static void test_grip_initAndK(void) {
  gvizGraph g;
  make_path(&g, 2);
  gvizGRIPState s;
  memset(&s, 0, sizeof(s));
  CHECK(gvizGRIPEmbedderInit(&s, gvizSubgraphCreateFull(&g), 0, 2) == -1);
  gvizGraphRelease(&g);
  make_path(&g, 30);
  memset(&s, 0, sizeof(s));
  gvizGRIPEmbedderConfigureStats(&s, false);
  gvizGRIPEmbedderConfigureKnnCapacity(&s, 20);
  CHECK(gvizGRIPEmbedderInit(&s, gvizSubgraphCreateFull(&g), 0, 2) == 0);
  CHECK(s.knnCapacity == 20 && !s.statsEnabled);
  CHECK(s.placementKMax == 128 && s.refinementKMax == 128);
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)&s;
  CHECK(gvizEmbeddedGraphStatSeriesCount(e) == 0);
  CHECK(gvizEmbeddedGraphFindAction(e, "grip.refineRound") != NULL);
  CHECK(gvizEmbeddedGraphFindAction(e, "grip.nextStage") != NULL);
  CHECK(!gvizEmbeddedGraphIsVertexVisible(e, 0));
  gvizGRIPEmbedderConfigureK(&s, 64, 0, GVIZ_GRIP_K_LAYER_DECAY);
  CHECK(s.placementKMax == 20 && s.refinementKMax == 128);
  gvizGRIPEmbedderConfigureK(&s, 5, 7, GVIZ_GRIP_K_BUDGET);
  CHECK(s.placementKMax == 5 && s.refinementKMax == 7 && s.kPolicy == GVIZ_GRIP_K_BUDGET);
  /* before Begin the actions do nothing */
  CHECK(gvizEmbeddedGraphInvokeAction(e, "grip.refineRound", NULL) == 0);
  CHECK(s.currRound == 0);
  gvizGRIPRoundStats st = gvizGRIPEmbedderLastRoundStats(&s);
  CHECK(st.maxDisplacement == 0.0);
  gvizGRIPEmbedderRelease(&s);
  gvizGraphRelease(&g);
}

static int mode_unit(void) {
  test_graph_insertionOrderAndCSR();
  test_embeddedGraph_actionsStatsMask();
  test_embeddedGraph_randomizeSaveLoad();
  test_forceEmbedder_initConfigure();
  test_forceModel_table();
  test_tutte_boundaryAndSeed();
  test_grip_initAndK();
  printf("%s (%d failed checks)\n", failures ? "FAIL" : "OK", failures);
  return failures ? 1 : 0;
}

/* Binary checkpoint through the host API: embedder A runs k1 steps from
 * Begin(seed), saves its state to `path`, runs k2 more; embedder B (same
 * graph, Begin(seed + 1)) loads the state and runs k2 steps. Writes both
 * position arrays, both step counts and both global speeds. Input: graph,
 * model, seed, k1, k2, path length, path. */
static int mode_force_checkpoint(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  const int64_t model = rd_i(), seed = rd_i(), k1 = rd_i(), k2 = rd_i();
  const int64_t len = rd_i();
  char *path = calloc((size_t)len + 1, 1);
  if (fread(path, 1, (size_t)len, fin) != (size_t)len)
    return 2;
  gvizForceEmbedderState a, b;
  if (gvizForceEmbedderInit(&a, gvizSubgraphCreateFull(&g), 2, (gvizForceModelKind)model) ||
      gvizForceEmbedderInit(&b, gvizSubgraphCreateFull(&g), 2, (gvizForceModelKind)model))
    return 1;
  gvizForceEmbedderSetPrecision(&a, env_precision());
  gvizForceEmbedderSetPrecision(&b, env_precision());
  gvizForceEmbedderConfigureGravity(&a, 0.05);
  gvizForceEmbedderConfigureGravity(&b, 0.05);
  if (gvizForceEmbedderBegin(&a, (unsigned int)seed) ||
      gvizForceEmbedderBegin(&b, (unsigned int)seed + 1))
    return fail_device("gvizForceEmbedderBegin");
  if (gvizForceEmbedderRun(&a, (size_t)k1, 0.0) < 0)
    return fail_device("gvizForceEmbedderRun");
  if (gvizForceEmbedderSaveState(&a, path) != 0)
    return fail_device("gvizForceEmbedderSaveState");
  if (gvizForceEmbedderRun(&a, (size_t)k2, 0.0) < 0)
    return fail_device("gvizForceEmbedderRun");
  if (gvizForceEmbedderLoadState(&b, path) != 0)
    return fail_device("gvizForceEmbedderLoadState");
  if (gvizForceEmbedderRun(&b, (size_t)k2, 0.0) < 0)
    return fail_device("gvizForceEmbedderRun");
  free(path);
  gvizForceEmbedderState *st[2] = {&a, &b};
  for (int k = 0; k < 2; k++) {
    const double *p = gvizEmbeddedGraphPositions((gvizEmbeddedGraph *)st[k]);
    for (int64_t i = 0; i < 2 * n; i++)
      wr_d(p[i]);
    wr_i((int64_t)st[k]->iteration);
    wr_d(st[k]->globalSpeed);
  }
  gvizForceEmbedderRelease(&a);
  gvizForceEmbedderRelease(&b);
  return 0;
}

/* GRIP checkpoint through the host API: embedder A (Begin, `stages`
 * NextStage calls, `rounds` RefineRound calls) saves its state to `path`;
 * then A and a second embedder B that Begins on the same graph and loads the
 * state run the same continuation (`rounds` rounds, then up to `stages`
 * times NextStage + `rounds` rounds). Writes, for A then B: current layer,
 * positions, visibility. Input: graph, kcap, placement K, refinement K,
 * policy, stages, rounds, path length, path. */
static void grip_continue(gvizGRIPState *st, int64_t stages, int64_t rounds) {
  for (int64_t r = 0; r < rounds; r++)
    gvizGRIPEmbedderRefineRound(st);
  for (int64_t k = 0; k < stages && st->currLayer > 0; k++) {
    gvizGRIPEmbedderNextStage(st);
    for (int64_t r = 0; r < rounds; r++)
      gvizGRIPEmbedderRefineRound(st);
  }
}

static int mode_grip_checkpoint(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  const int64_t kcap = rd_i(), pk = rd_i(), rk = rd_i(), policy = rd_i();
  const int64_t stages = rd_i(), rounds = rd_i(), len = rd_i();
  char *path = calloc((size_t)len + 1, 1);
  if (fread(path, 1, (size_t)len, fin) != (size_t)len)
    return 2;
  gvizGRIPState a, b;
  memset(&a, 0, sizeof(a));
  memset(&b, 0, sizeof(b));
  gvizGRIPEmbedderConfigureKnnCapacity(&a, (size_t)kcap);
  gvizGRIPEmbedderConfigureKnnCapacity(&b, (size_t)kcap);
  if (gvizGRIPEmbedderInit(&a, gvizSubgraphCreateFull(&g), 0, 2) != 0 ||
      gvizGRIPEmbedderInit(&b, gvizSubgraphCreateFull(&g), 0, 2) != 0)
    return 1;
  gvizGRIPEmbedderConfigureK(&a, (size_t)pk, (size_t)rk, (gvizGRIPKPolicy)policy);
  gvizGRIPEmbedderConfigureK(&b, (size_t)pk, (size_t)rk, (gvizGRIPKPolicy)policy);
  gvizGRIPEmbedderBegin(&a);
  for (int64_t k = 0; k < stages && a.currLayer > 0; k++)
    gvizGRIPEmbedderNextStage(&a);
  for (int64_t r = 0; r < rounds; r++)
    gvizGRIPEmbedderRefineRound(&a);
  if (gvizGRIPEmbedderSaveState(&a, path) != 0)
    return fail_device("gvizGRIPEmbedderSaveState");
  grip_continue(&a, stages, rounds);
  gvizGRIPEmbedderBegin(&b);
  if (gvizGRIPEmbedderLoadState(&b, path) != 0)
    return fail_device("gvizGRIPEmbedderLoadState");
  grip_continue(&b, stages, rounds);
  free(path);
  gvizGRIPState *st[2] = {&a, &b};
  for (int k = 0; k < 2; k++) {
    gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)st[k];
    wr_i((int64_t)st[k]->currLayer);
    const double *p = gvizEmbeddedGraphPositions(e);
    for (int64_t i = 0; i < 2 * n; i++)
      wr_d(p[i]);
    for (int64_t v = 0; v < n; v++)
      wr_i(gvizEmbeddedGraphIsVertexVisible(e, (size_t)v) ? 1 : 0);
  }
  gvizGRIPEmbedderRelease(&a);
  gvizGRIPEmbedderRelease(&b);
  gvizGraphRelease(&g);
  return 0;
}

/* The .edges loader on a file named in the input (directed, skip_header,
 * path length, path), in the reference harness's framing: return code, then
 * the vertex count and every adjacency list; then the CSR loader's return
 * code, vertex count, offsets and adjacency. */
static int mode_edges_file(void) {
  gvizEdgesFileOptions opts;
  gvizEdgesFileOptionsInit(&opts);
  opts.directed = (int)rd_i();
  opts.skip_header = (int)rd_i();
  const int64_t len = rd_i();
  char *path = calloc((size_t)len + 1, 1);
  if (fread(path, 1, (size_t)len, fin) != (size_t)len)
    return 2;
  gvizGraph g;
  const int rc = gvizGraphLoadFromEdgesFile(path, &opts, &g);
  wr_i(rc);
  if (rc == 0) {
    const size_t n = gvizGraphSize(&g);
    wr_i((int64_t)n);
    for (size_t u = 0; u < n; u++) {
      size_t cnt = 0;
      const size_t *nb = gvizGraphNeighbors(&g, u, &cnt);
      wr_i((int64_t)cnt);
      for (size_t k = 0; k < cnt; k++)
        wr_i((int64_t)nb[k]);
    }
    gvizGraphRelease(&g);
  }
  int64_t n = 0, *off = NULL, *adj = NULL;
  const int rc2 = gvizGraphLoadEdgesFileCSR(path, &opts, &n, &off, &adj);
  free(path);
  wr_i(rc2);
  if (rc2 == 0) {
    wr_i(n);
    for (int64_t i = 0; i <= n; i++)
      wr_i(off[i]);
    for (int64_t e = 0; e < off[n]; e++)
      wr_i(adj[e]);
    free(off);
    free(adj);
  }
  return 0;
}

/* Text embedding I/O (gvizEmbeddedGraphSaveEmbedding / LoadEmbedding) on the
 * full subgraph of a graph in `dim` dimensions: positions in, Save's return
 * and the bytes it wrote, then for each file to load: positions reset to the
 * input, Load's return and the positions after it. Input: graph, dim,
 * positions, name, save path, load-path count, load paths (each as
 * length + bytes). */
static char *rd_str(void) {
  const int64_t len = rd_i();
  char *s = calloc((size_t)len + 1, 1);
  if (len > 0 && fread(s, 1, (size_t)len, fin) != (size_t)len) {
    free(s);
    return NULL;
  }
  return s;
}

static int mode_embedding_io(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  const int64_t dim = rd_i();
  double *pos = malloc(sizeof(double) * (size_t)(n * dim + 1));
  for (int64_t i = 0; i < n * dim; i++)
    pos[i] = rd_d();
  char *name = rd_str(), *save = rd_str();
  if (!name || !save)
    return 2;
  gvizEmbeddedGraph e;
  if (gvizEmbeddedGraphInit(&e, gvizSubgraphCreateFull(&g), (size_t)dim) != 0)
    return 1;
  for (int64_t u = 0; u < n; u++)
    gvizEmbeddedGraphSetVPosition(&e, (size_t)u, pos + dim * u);
  wr_i(gvizEmbeddedGraphSaveEmbedding(&e, name, save));
  FILE *f = fopen(save, "rb");
  int64_t len = 0;
  char *buf = NULL;
  if (f) {
    fseek(f, 0, SEEK_END);
    len = ftell(f);
    fseek(f, 0, SEEK_SET);
    buf = malloc((size_t)len + 1);
    if (fread(buf, 1, (size_t)len, f) != (size_t)len)
      len = 0;
    fclose(f);
  }
  wr_i(len);
  if (len > 0)
    fwrite(buf, 1, (size_t)len, fout);
  free(buf);
  const int64_t loads = rd_i();
  for (int64_t k = 0; k < loads; k++) {
    char *path = rd_str();
    if (!path)
      return 2;
    for (int64_t u = 0; u < n; u++)
      gvizEmbeddedGraphSetVPosition(&e, (size_t)u, pos + dim * u);
    wr_i(gvizEmbeddedGraphLoadEmbedding(&e, path));
    for (int64_t u = 0; u < n; u++) {
      const double *p = gvizEmbeddedGraphGetVPosition(&e, (size_t)u);
      for (int64_t j = 0; j < dim; j++)
        wr_d(p[j]);
    }
    free(path);
  }
  gvizEmbeddedGraphRelease(&e);
  gvizGraphRelease(&g);
  free(name);
  free(save);
  free(pos);
  return 0;
}

/* Spring-driven Tutte: FixConvexPolygon + SeedInterior (+ Configure), then
 * `steps` Step(dt) calls and one Run(runTo, dt). Writes the seeded
 * positions, each step's return, positions and velocities after the steps,
 * Run's return, positions and velocities after it, and the
 * "springTutte.maxDelta" series. Input: graph, boundary count, steps,
 * runTo, boundary, radius, dt, stiffness, damping, epsilon. */
/* Tutte and spring-Tutte Begin (planarity, largest face, convex polygon,
 * seeded interior), then Run to convergence. */
static int mode_tutte_begin(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t spring = rd_i(), iters = rd_i();
  double eps = rd_d();
  const double *p;
  int rc, it;
  if (!spring) {
    gvizTutteState s;
    if (gvizTutteEmbedderInit(&s, gvizSubgraphCreateFull(&g), 2, eps) != 0)
      return 1;
    rc = gvizTutteEmbedderBegin(&s);
    wr_i(rc);
    if (rc == 0) {
      wr_i((int64_t)s.boundaryCount);
      for (size_t i = 0; i < s.boundaryCount; i++)
        wr_i((int64_t)s.boundary[i]);
      it = gvizTutteEmbedderRun(&s, (size_t)iters);
      wr_i(it);
      p = gvizEmbeddedGraphPositions((gvizEmbeddedGraph *)&s);
      for (int64_t i = 0; i < 2 * n; i++)
        wr_d(p[i]);
    }
    gvizTutteEmbedderRelease(&s);
  } else {
    gvizSpringTutteState s;
    if (gvizSpringTutteEmbedderInit(&s, gvizSubgraphCreateFull(&g), 2, eps) != 0)
      return 1;
    rc = gvizSpringTutteEmbedderBegin(&s);
    wr_i(rc);
    if (rc == 0) {
      wr_i((int64_t)s.boundaryCount);
      for (size_t i = 0; i < s.boundaryCount; i++)
        wr_i((int64_t)s.boundary[i]);
      it = gvizSpringTutteEmbedderRun(&s, (size_t)iters, 1.0 / 60.0);
      wr_i(it);
      p = gvizEmbeddedGraphPositions((gvizEmbeddedGraph *)&s);
      for (int64_t i = 0; i < 2 * n; i++)
        wr_d(p[i]);
    }
    gvizSpringTutteEmbedderRelease(&s);
  }
  gvizGraphRelease(&g);
  return 0;
}

/* Begin, then the adjacency lists Begin left (the rotation), then a
 * highlight (vertices and edges read from the input) fixed as the outer
 * face, then Run: rc of Begin and FixOuterFace, planarEmbedded, per vertex
 * its list, the boundary, the iterations and positions. */
static void write_lists(const gvizGraph *g, int64_t n) {
  for (int64_t u = 0; u < n; u++) {
    size_t c = 0;
    const size_t *nb = gvizGraphNeighbors(g, (size_t)u, &c);
    wr_i((int64_t)c);
    for (size_t k = 0; k < c; k++)
      wr_i((int64_t)nb[k]);
  }
}

static int mode_tutte_outer_face(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t spring = rd_i(), iters = rd_i(), hv = rd_i(), he = rd_i();
  int64_t *verts = malloc(sizeof(int64_t) * (hv ? hv : 1));
  int64_t *edges = malloc(sizeof(int64_t) * 2 * (he ? he : 1));
  for (int64_t i = 0; i < hv; i++)
    verts[i] = rd_i();
  for (int64_t i = 0; i < 2 * he; i++)
    edges[i] = rd_i();
  union {
    gvizTutteState t;
    gvizSpringTutteState sp;
  } st;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)&st;
  int rc = spring ? gvizSpringTutteEmbedderInit(&st.sp, gvizSubgraphCreateFull(&g), 2, 1e-5)
                  : gvizTutteEmbedderInit(&st.t, gvizSubgraphCreateFull(&g), 2, 1e-5);
  if (rc != 0)
    return 1;
  rc = spring ? gvizSpringTutteEmbedderBegin(&st.sp) : gvizTutteEmbedderBegin(&st.t);
  wr_i(rc);
  wr_i(e->planarEmbedded);
  write_lists(&g, n);
  /* the highlight: a subgraph with an edge subset, as a viewer's face pick */
  gvizSubgraph h = gvizSubgraphCreateEmpty(&g);
  for (int64_t i = 0; i < hv; i++)
    gvizSubgraphShowVertex(&h, (size_t)verts[i]);
  for (int64_t i = 0; i < he; i++)
    gvizSubgraphShowEdge(&h, (size_t)edges[2 * i], (size_t)edges[2 * i + 1]);
  gvizEmbeddedGraphSetHighlight(e, h);
  rc = spring ? gvizSpringTutteEmbedderFixOuterFace(&st.sp)
              : gvizTutteEmbedderFixOuterFace(&st.t);
  wr_i(rc);
  size_t bc = spring ? st.sp.boundaryCount : st.t.boundaryCount;
  const size_t *b = spring ? st.sp.boundary : st.t.boundary;
  wr_i((int64_t)bc);
  for (size_t i = 0; i < bc; i++)
    wr_i((int64_t)b[i]);
  int it = spring ? gvizSpringTutteEmbedderRun(&st.sp, (size_t)iters, 1.0 / 60.0)
                  : gvizTutteEmbedderRun(&st.t, (size_t)iters);
  wr_i(it);
  const double *p = gvizEmbeddedGraphPositions(e);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  if (spring)
    gvizSpringTutteEmbedderRelease(&st.sp);
  else
    gvizTutteEmbedderRelease(&st.t);
  gvizGraphRelease(&g);
  free(verts);
  free(edges);
  return 0;
}

static int mode_spring_tutte(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  const int64_t nb = rd_i(), steps = rd_i(), runTo = rd_i();
  size_t *boundary = malloc(sizeof(size_t) * nb);
  for (int64_t i = 0; i < nb; i++)
    boundary[i] = (size_t)rd_i();
  const double radius = rd_d(), dt = rd_d(), stiffness = rd_d(), damping = rd_d();
  const double epsilon = rd_d();
  gvizSpringTutteState s;
  if (gvizSpringTutteEmbedderInit(&s, gvizSubgraphCreateFull(&g), 2, epsilon) != 0)
    return 1;
  gvizSpringTutteEmbedderConfigure(&s, stiffness, damping);
  gvizSpringTutteFixConvexPolygon(&s, boundary, (size_t)nb, radius);
  gvizSpringTutteEmbedderSeedInterior(&s);
  s.begun = 1;
  const double *p = gvizEmbeddedGraphPositions((gvizEmbeddedGraph *)&s);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  for (int64_t it = 0; it < steps; it++)
    wr_d(gvizSpringTutteEmbedderStep(&s, dt));
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(s.velocity[i]);
  wr_i(gvizSpringTutteEmbedderRun(&s, (size_t)runTo, dt));
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(s.velocity[i]);
  write_series((gvizEmbeddedGraph *)&s, "springTutte.maxDelta");
  gvizSpringTutteEmbedderRelease(&s);
  gvizGraphRelease(&g);
  free(boundary);
  return 0;
}

/* ---------------------------------------------------------------- quadtree */
/* Node records in the reference-side harness's framing: parent before
 * children in quadrant order; depth, path key, leaf flag, point count,
 * mass, centre of mass, half size; then per point its leaf's depth and key,
 * then the node count. */
static void qt_dump_node(const gvizQuadtreeNode *nd, int depth, uint64_t kh, uint64_t kl,
                         int64_t *leafDepth, uint64_t *leafHi, uint64_t *leafLo,
                         int64_t *nodeCount) {
  if (!nd)
    return;
  (*nodeCount)++;
  int leaf = gvizQuadtreeNodeIsLeaf(nd);
  wr_i(depth);
  wr_i((int64_t)kh);
  wr_i((int64_t)kl);
  wr_i(leaf);
  wr_i(leaf ? (int64_t)gvizQuadtreeNodePointCount(nd) : 0);
  wr_d(nd->mass);
  wr_d(nd->comX);
  wr_d(nd->comY);
  wr_d(nd->halfSize);
  if (leaf) {
    for (size_t i = 0; i < gvizQuadtreeNodePointCount(nd); i++) {
      size_t p = gvizQuadtreeNodePointAt(nd, i);
      leafDepth[p] = depth;
      leafHi[p] = kh;
      leafLo[p] = kl;
    }
    return;
  }
  for (int q = 0; q < 4; q++) {
    uint64_t h = kh, l = kl;
    int level = depth + 1;
    if (level <= 32)
      h |= (uint64_t)q << (2 * (32 - level));
    else if (level <= 64)
      l |= (uint64_t)q << (2 * (64 - level));
    qt_dump_node(gvizQuadtreeNodeChild(nd, (gvizQuadrant)q), depth + 1, h, l, leafDepth,
                 leafHi, leafLo, nodeCount);
  }
}

static void qt_dump(const gvizQuadtree *t, int64_t n) {
  int64_t *ld = calloc(n ? n : 1, 8);
  uint64_t *lh = calloc(n ? n : 1, 8), *ll = calloc(n ? n : 1, 8);
  int64_t nodeCount = 0;
  qt_dump_node(gvizQuadtreeRoot(t), 0, 0, 0, ld, lh, ll, &nodeCount);
  for (int64_t i = 0; i < n; i++) {
    wr_i(ld[i]);
    wr_i((int64_t)lh[i]);
    wr_i((int64_t)ll[i]);
  }
  wr_i(nodeCount);
  free(ld);
  free(lh);
  free(ll);
}

static int mode_quadtree(void) {
  int64_t n = rd_i(), npc = rd_i();
  double *pts = malloc(sizeof(double) * 2 * (n ? n : 1));
  double *ms = malloc(sizeof(double) * (n ? n : 1));
  for (int64_t i = 0; i < 2 * n; i++)
    pts[i] = rd_d();
  for (int64_t i = 0; i < n; i++)
    ms[i] = rd_d();
  gvizQuadtree t;
  if (gvizQuadtreeInit(&t, pts, ms, (size_t)n, (size_t)npc) != 0)
    return fail_device("gvizQuadtreeInit");
  qt_dump(&t, n);
  gvizQuadtreeRelease(&t);
  free(pts);
  free(ms);
  return 0;
}

/* The force state's quadtree after Begin and `steps` steps with the
 * positions overwritten after Begin: writes the positions before the last
 * step (or after Begin), the tree's point buffer, then the tree. */
static int mode_force_quadtree(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t model = rd_i(), bh = rd_i(), npc = rd_i(), steps = rd_i();
  double theta = rd_d();
  double *pos = malloc(sizeof(double) * 2 * n);
  for (int64_t i = 0; i < 2 * n; i++)
    pos[i] = rd_d();
  gvizForceEmbedderState s;
  if (gvizForceEmbedderInit(&s, gvizSubgraphCreateFull(&g), 2, (gvizForceModelKind)model) != 0)
    return 1;
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)&s;
  gvizForceEmbedderSetPrecision(&s, env_precision());
  gvizForceEmbedderConfigureBarnesHut(&s, theta, (size_t)npc);
  gvizForceEmbedderSetBarnesHutEnabled(&s, (int)bh);
  if (gvizForceEmbedderBegin(&s, 7) != 0)
    return fail_device("gvizForceEmbedderBegin");
  double *p = gvizEmbeddedGraphPositions(e);
  if (steps == 0) {
    /* the tree of Begin's own positions */
  } else {
    memcpy(p, pos, sizeof(double) * 2 * n);
  }
  for (int64_t k = 0; k + 1 < steps; k++)
    if (gvizForceEmbedderStep(&s) < 0.0)
      return fail_device("gvizForceEmbedderStep");
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  if (steps > 0 && gvizForceEmbedderStep(&s) < 0.0)
    return fail_device("gvizForceEmbedderStep");
  const gvizQuadtreeNode *root = gvizQuadtreeRoot(&s.quadtree);
  wr_i(root != NULL);
  wr_i(s.quadtreeReady);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(root ? s.quadtree.points[i] : 0.0);
  for (int64_t i = 0; i < n; i++)
    wr_d(root ? s.quadtree.masses[i] : 0.0);
  qt_dump(&s.quadtree, n);
  gvizForceEmbedderRelease(&s);
  gvizGraphRelease(&g);
  free(pos);
  return 0;
}

/* The reference's quadtree unit cases (tests/ds/gvizQuadtreeTests.c:13-343)
 * restated against the host view; prints each failed check. */
#define QT_CHECK(c)                                                           \
  do {                                                                        \
    if (!(c)) {                                                               \
      fprintf(stderr, "quadtree check failed at line %d: %s\n", __LINE__, #c); \
      bad++;                                                                  \
    }                                                                         \
  } while (0)

static int near(double a, double b) { return fabs(a - b) <= 1e-12 * fmax(1.0, fabs(b)); }

static int mode_quadtree_cases(void) {
  int bad = 0;
  gvizQuadtree t;
  double cx, cy;
  QT_CHECK(GVIZ_QUADTREE_NODES_PER_CELL_DEFAULT == 1);
  /* empty */
  QT_CHECK(gvizQuadtreeInit(&t, NULL, NULL, 0, 1) == 0);
  QT_CHECK(gvizQuadtreeRoot(&t) == NULL);
  gvizQuadtreeRelease(&t);
  /* one point */
  {
    double pts[] = {3.0, 4.0}, ms[] = {1.0};
    QT_CHECK(gvizQuadtreeInit(&t, pts, ms, 1, 1) == 0);
    const gvizQuadtreeNode *r = gvizQuadtreeRoot(&t);
    QT_CHECK(r && gvizQuadtreeNodeIsLeaf(r) && gvizQuadtreeNodePointCount(r) == 1 &&
             gvizQuadtreeNodePointAt(r, 0) == 0);
    gvizQuadtreeNodeCenterOfMass(r, &cx, &cy);
    QT_CHECK(cx == 3.0 && cy == 4.0);
    gvizQuadtreeRelease(&t);
  }
  /* four corners split the root into four single-point leaves */
  double corners[] = {-5.0, 5.0, 5.0, 5.0, -5.0, -5.0, 5.0, -5.0}, ones[] = {1, 1, 1, 1, 1};
  {
    QT_CHECK(gvizQuadtreeInit(&t, corners, ones, 4, 1) == 0);
    const gvizQuadtreeNode *r = gvizQuadtreeRoot(&t);
    QT_CHECK(r && !gvizQuadtreeNodeIsLeaf(r) && r->mass == 4.0);
    for (int q = 0; q < 4; q++) {
      const gvizQuadtreeNode *c = gvizQuadtreeNodeChild(r, (gvizQuadrant)q);
      QT_CHECK(c && gvizQuadtreeNodeIsLeaf(c) && gvizQuadtreeNodePointCount(c) == 1 &&
               gvizQuadtreeNodePointAt(c, 0) == (size_t)q);
    }
    gvizQuadtreeNodeCenterOfMass(r, &cx, &cy);
    QT_CHECK(fabs(cx) <= 1e-9 && fabs(cy) <= 1e-9);
    QT_CHECK(t.nodesUsed == 5);
    /* rebuilds reuse the arena */
    size_t blocks = t.nodeBlocks.count;
    QT_CHECK(blocks > 0);
    for (int k = 0; k < 5; k++) {
      QT_CHECK(gvizQuadtreeRebuild(&t, corners, ones, 4) == 0);
      QT_CHECK(t.nodeBlocks.count == blocks && t.pointBlocks.count == blocks);
    }
    double moved[] = {100.0, 100.0, 101.0, 100.0, 100.0, 101.0, 101.0, 101.0};
    QT_CHECK(gvizQuadtreeRebuild(&t, moved, ones, 4) == 0);
    r = gvizQuadtreeRoot(&t);
    QT_CHECK(r && r->mass == 4.0);
    gvizQuadtreeNodeCenterOfMass(r, &cx, &cy);
    QT_CHECK(near(cx, 100.5) && near(cy, 100.5));
    gvizQuadtreeRelease(&t);
  }
  /* a leaf holds up to nodesPerCell points */
  {
    double pts[] = {1.0, 1.0, 2.0, 2.0, 3.0, 3.0};
    QT_CHECK(gvizQuadtreeInit(&t, pts, ones, 3, 3) == 0);
    const gvizQuadtreeNode *r = gvizQuadtreeRoot(&t);
    QT_CHECK(r && gvizQuadtreeNodeIsLeaf(r) && gvizQuadtreeNodePointCount(r) == 3 &&
             r->mass == 3.0);
    gvizQuadtreeRelease(&t);
  }
  /* coincident points end in an overflow leaf; rebuilds do not pile up
   * overflow buffers */
  {
    double dup[] = {2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0};
    QT_CHECK(gvizQuadtreeInit(&t, dup, ones, 3, 1) == 0);
    const gvizQuadtreeNode *r = gvizQuadtreeRoot(&t);
    QT_CHECK(r && r->mass == 3.0);
    gvizQuadtreeNodeCenterOfMass(r, &cx, &cy);
    QT_CHECK(near(cx, 2.0) && near(cy, 2.0));
    gvizQuadtreeRelease(&t);
    QT_CHECK(gvizQuadtreeInit(&t, dup, ones, 4, 1) == 0);
    size_t over = t.overflowBuffers.count;
    QT_CHECK(over > 0);
    QT_CHECK(gvizQuadtreeRebuild(&t, dup, ones, 4) == 0);
    QT_CHECK(t.overflowBuffers.count == over);
    gvizQuadtreeRelease(&t);
  }
  /* centre of mass is the mean; weights skew it */
  {
    double pts[] = {0.0, 0.0, 10.0, 0.0, 0.0, 10.0, 10.0, 10.0, 4.0, 4.0};
    QT_CHECK(gvizQuadtreeInit(&t, pts, ones, 5, 1) == 0);
    const gvizQuadtreeNode *r = gvizQuadtreeRoot(&t);
    gvizQuadtreeNodeCenterOfMass(r, &cx, &cy);
    QT_CHECK(near(cx, 24.0 / 5.0) && near(cy, 24.0 / 5.0) && r->mass == 5.0);
    gvizQuadtreeRelease(&t);
    double two[] = {0.0, 0.0, 10.0, 0.0}, w[] = {1.0, 3.0};
    QT_CHECK(gvizQuadtreeInit(&t, two, w, 2, 1) == 0);
    r = gvizQuadtreeRoot(&t);
    gvizQuadtreeNodeCenterOfMass(r, &cx, &cy);
    QT_CHECK(r->mass == 4.0 && near(cx, 7.5) && cy == 0.0);
    gvizQuadtreeRelease(&t);
    double same[] = {2.0, 2.0, 2.0, 2.0}, w2[] = {5.0, 10.0};
    QT_CHECK(gvizQuadtreeInit(&t, same, w2, 2, 1) == 0);
    r = gvizQuadtreeRoot(&t);
    gvizQuadtreeNodeCenterOfMass(r, &cx, &cy);
    QT_CHECK(r->mass == 15.0 && near(cx, 2.0) && near(cy, 2.0));
    gvizQuadtreeRelease(&t);
  }
  /* zero total mass keeps the centre of mass where it was (:118-123) */
  {
    double pts[] = {10.0, 10.0, 0.0, 0.0}, w[] = {1.0, -1.0};
    QT_CHECK(gvizQuadtreeInit(&t, pts, w, 2, 1) == 0);
    const gvizQuadtreeNode *r = gvizQuadtreeRoot(&t);
    gvizQuadtreeNodeCenterOfMass(r, &cx, &cy);
    QT_CHECK(r->mass == 0.0 && cx == 10.0 && cy == 10.0);
    gvizQuadtreeRelease(&t);
  }
  wr_i(bad);
  return bad ? 1 : 0;
}

/* --------------------------------------------------------------------- kNN */
/* Same framing as the reference-side harness: per source the count and k
 * (vertex, dist) pairs of gvizSearchKNearest over the visible filter, then
 * the batched search with every non-visible vertex as a target. */
static int mode_knn(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t k = rd_i(), nsrc = rd_i();
  gvizSubgraph sg = gvizSubgraphCreateFull(&g);
  gvizVertexSubset vis = gvizVertexSubsetCreateEmpty(&g);
  for (int64_t i = 0; i < n; i++)
    if (rd_i())
      gvizVertexSubsetShowVertex(vis, (size_t)i);
  gvizFoundVertex *out = malloc(sizeof(gvizFoundVertex) * (k + 1));
  for (int64_t q = 0; q < nsrc; q++) {
    int64_t src = rd_i();
    int c = gvizSearchKNearest(&sg, out, (size_t)k, (size_t)src, vis);
    wr_i(c);
    for (int64_t j = 0; j < k; j++) {
      wr_i(j < c ? (int64_t)out[j].v : -1);
      wr_i(j < c ? (int64_t)out[j].dist : -1);
    }
  }
  int64_t nt = 0;
  gvizKNNBatchTarget *targets = malloc(sizeof(*targets) * n);
  gvizFoundVertex *found = malloc(sizeof(gvizFoundVertex) * n * k);
  for (int64_t i = 0; i < n; i++)
    if (!gvizVertexSubsetTest(vis, (size_t)i)) {
      targets[nt] = (gvizKNNBatchTarget){(size_t)i, found + nt * k, 0};
      nt++;
    }
  gvizKNearestScratch scratch;
  if (gvizKNearestScratchInit(&scratch, (size_t)n) != 0)
    return 1;
  int rc = gvizSearchKNearestFromVisibleBatch(&sg, vis, (size_t)k, targets, (size_t)nt,
                                              &scratch);
  if (rc < 0)
    return fail_device("gvizSearchKNearestFromVisibleBatch");
  wr_i(rc);
  wr_i(nt);
  for (int64_t t = 0; t < nt; t++) {
    wr_i((int64_t)targets[t].vertex);
    wr_i((int64_t)targets[t].count);
    for (int64_t j = 0; j < k; j++) {
      int ok = j < (int64_t)targets[t].count;
      wr_i(ok ? (int64_t)targets[t].out[j].v : -1);
      wr_i(ok ? (int64_t)targets[t].out[j].dist : -1);
    }
  }
  gvizKNearestScratchRelease(&scratch);
  free(out);
  free(targets);
  free(found);
  gvizSubgraphRelease(&sg);
  gvizVertexSubsetRelease(vis);
  gvizGraphRelease(&g);
  return 0;
}

/* The reference's kNN unit cases (tests/algorithms/search/gvizSearchTests.c:
 * 94-131, 184-207, 287-360) restated. */
static int mode_knn_cases(void) {
  int bad = 0;
  gvizFoundVertex f[8];
  /* triangle: both other vertices at distance 1, in adjacency order */
  {
    gvizGraph g;
    gvizGraphInit(&g, 0);
    for (int i = 0; i < 3; i++)
      gvizGraphAddVertex(&g, NULL, NULL, NULL);
    gvizGraphAddEdge(&g, 0, 1);
    gvizGraphAddEdge(&g, 1, 2);
    gvizGraphAddEdge(&g, 0, 2);
    gvizGraphBuildLayout(&g);
    gvizSubgraph sg = gvizSubgraphCreateFull(&g);
    int c = gvizSearchKNearest(&sg, f, 2, 0, NULL);
    QT_CHECK(c == 2 && f[0].v == 1 && f[0].dist == 1 && f[1].v == 2 && f[1].dist == 1);
    /* a filter admits only its vertices */
    gvizVertexSubset filt = gvizVertexSubsetCreateEmpty(&g);
    gvizVertexSubsetShowVertex(filt, 2);
    c = gvizSearchKNearest(&sg, f, 2, 0, filt);
    QT_CHECK(c == 1 && f[0].v == 2 && f[0].dist == 1);
    /* k beyond the graph, k = 0, a source outside the graph */
    c = gvizSearchKNearest(&sg, f, 5, 0, NULL);
    QT_CHECK(c == 2);
    QT_CHECK(gvizSearchKNearest(&sg, NULL, 0, 0, NULL) == 0);
    QT_CHECK(gvizSearchKNearest(&sg, f, 2, 7, NULL) == -1);
    QT_CHECK(gvizSearchKNearest(NULL, f, 2, 0, NULL) == -1);
    gvizVertexSubsetRelease(filt);
    gvizSubgraphRelease(&sg);
    gvizGraphRelease(&g);
  }
  /* path 0-1-2-3: the scratch form equals the stateless one, reused */
  {
    gvizGraph g;
    make_path(&g, 4);
    gvizSubgraph sg = gvizSubgraphCreateFull(&g);
    gvizKNearestScratch sc;
    QT_CHECK(gvizKNearestScratchInit(&sc, 4) == 0);
    gvizFoundVertex a[4], b[4];
    for (size_t src = 0; src < 4; src++)
      for (int rep = 0; rep < 2; rep++) {
        int ca = gvizSearchKNearest(&sg, a, 3, src, NULL);
        int cb = gvizSearchKNearestScratch(&sg, b, 3, src, NULL, &sc);
        QT_CHECK(ca == cb && ca == 3);
        for (int i = 0; i < ca && i < 4; i++)
          QT_CHECK(a[i].v == b[i].v && a[i].dist == b[i].dist);
      }
    gvizFoundVertex c3[3];
    QT_CHECK(gvizSearchKNearest(&sg, c3, 3, 0, NULL) == 3 && c3[2].v == 3 &&
             c3[2].dist == 3);
    /* a scratch too small for the graph */
    gvizKNearestScratch small;
    QT_CHECK(gvizKNearestScratchInit(&small, 2) == 0);
    QT_CHECK(gvizSearchKNearestScratch(&sg, a, 2, 0, NULL, &small) == -1);
    gvizKNearestScratchRelease(&small);
    QT_CHECK(gvizKNearestScratchInit(&small, 0) == -1);
    gvizKNearestScratchRelease(&sc);
    gvizSubgraphRelease(&sg);
    gvizGraphRelease(&g);
  }
  /* the batched search gives every target the per-vertex distances */
  {
    enum { N = 64, K = 4 };
    gvizGraph g;
    gvizGraphInit(&g, 0);
    for (int i = 0; i < N; i++)
      gvizGraphAddVertex(&g, NULL, NULL, NULL);
    for (int i = 0; i + 1 < N; i++)
      gvizGraphAddEdge(&g, i, i + 1);
    for (int i = 0; i + 8 < N; i += 8)
      gvizGraphAddEdge(&g, i, i + 8);
    gvizGraphBuildLayout(&g);
    gvizSubgraph sg = gvizSubgraphCreateFull(&g);
    gvizVertexSubset vis = gvizVertexSubsetCreateEmpty(&g);
    for (int i = 0; i < N; i += 16)
      gvizVertexSubsetShowVertex(vis, i);
    gvizKNNBatchTarget tg[N];
    gvizFoundVertex store[N * K];
    size_t nt = 0;
    for (int i = 0; i < N; i++)
      if (!gvizVertexSubsetTest(vis, i)) {
        tg[nt] = (gvizKNNBatchTarget){(size_t)i, store + nt * K, 0};
        nt++;
      }
    gvizKNearestScratch sc;
    QT_CHECK(gvizKNearestScratchInit(&sc, N) == 0);
    QT_CHECK(gvizSearchKNearestPreferBatch(&sg, gvizVertexSubsetCount(vis, N), nt) == 1);
    QT_CHECK(gvizSearchKNearestFromVisibleBatch(&sg, vis, K, tg, nt, &sc) == 0);
    for (size_t t = 0; t < nt; t++) {
      gvizFoundVertex direct[K];
      int dc = gvizSearchKNearest(&sg, direct, K, tg[t].vertex, vis);
      QT_CHECK(dc == (int)tg[t].count);
      for (int i = 0; i < dc; i++)
        QT_CHECK(direct[i].dist == tg[t].out[i].dist);
    }
    /* the batch path does not apply: no seeds, or not fewer seeds than targets */
    gvizVertexSubset none = gvizVertexSubsetCreateEmpty(&g);
    QT_CHECK(gvizSearchKNearestFromVisibleBatch(&sg, none, K, tg, nt, &sc) == 1);
    QT_CHECK(gvizSearchKNearestFromVisibleBatch(&sg, vis, K, tg, 2, &sc) == 1);
    QT_CHECK(gvizSearchKNearestPreferBatch(&sg, 0, nt) == 0);
    gvizVertexSubsetRelease(none);
    gvizKNearestScratchRelease(&sc);
    gvizVertexSubsetRelease(vis);
    gvizSubgraphRelease(&sg);
    gvizGraphRelease(&g);
  }
  wr_i(bad);
  return bad ? 1 : 0;
}

/* ------------------------------------------------------- GRIP stages */
static void write_filtration(const gvizGRIPState *st, size_t layers) {
  size_t active = *(const size_t *)gvizArrayAtIndex(&st->misBorder, 0);
  wr_i((int64_t)layers);
  wr_i((int64_t)st->misBorder.count);
  for (size_t k = 0; k < st->misBorder.count; k++)
    wr_i((int64_t)*(const size_t *)gvizArrayAtIndex(&st->misBorder, k));
  for (size_t k = 0; k < active; k++)
    wr_i((int64_t)st->misFiltration[k]);
}

/* The filtration twice: built stage by stage through the white-box
 * functions as the reference's createMISFiltration composes them
 * (src/embedders/gvizGRIPEmbedder.c:741-777), then by createMISFiltration
 * on a fresh state. Each: layer count, borders, filtration. */
static int mode_grip_stages(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t dim = rd_i();
  for (int pass = 0; pass < 2; pass++) {
    gvizGRIPState st;
    memset(&st, 0, sizeof st);
    if (gvizGRIPEmbedderInit(&st, gvizSubgraphCreateFull(&g), 0, (size_t)dim) != 0)
      return 1;
    size_t layers;
    if (pass == 0) {
      gvizVertexSubset cur = gvizVertexSubsetCreateEmpty(&g);
      makeFirstMISPartition(&st, cur);
      if (st.misBorder.count != 2)
        return fail_device("makeFirstMISPartition");
      size_t i = 2;
      while (iterMISFiltration(&st, i, cur))
        i++;
      if (st.misBorder.count != i + 1)
        return fail_device("iterMISFiltration");
      size_t k = 0;
      for (size_t v = 0; v < (size_t)n; v++)
        if (gvizVertexSubsetTest(cur, v))
          st.misFiltration[k++] = v;
      while (*(size_t *)gvizArrayAtIndex(&st.misBorder, i) < (size_t)dim + 1)
        if (migrateOneToFinalLayer(&st, cur, i + 1) < 0)
          break;
      layers = i + 1;
      gvizVertexSubsetRelease(cur);
    } else {
      layers = createMISFiltration(&st);
      if (layers == 0)
        return fail_device("createMISFiltration");
    }
    write_filtration(&st, layers);
    gvizGRIPEmbedderRelease(&st);
  }
  gvizGraphRelease(&g);
  return 0;
}

int main(int argc, char **argv) {
  if (argc == 2 && !strcmp(argv[1], "unit"))
    return mode_unit();
  if (argc != 4) {
    fprintf(stderr, "usage: %s unit | %s <mode> <in> <out>\n", argv[0], argv[0]);
    return 2;
  }
  fin = fopen(argv[2], "rb");
  fout = fopen(argv[3], "wb");
  if (!fin || !fout)
    return 2;
  int rc = 2;
  if (!strcmp(argv[1], "api_force"))
    rc = mode_api_force();
  else if (!strcmp(argv[1], "api_grip"))
    rc = mode_api_grip();
  else if (!strcmp(argv[1], "tutte"))
    rc = mode_tutte();
  else if (!strcmp(argv[1], "edges_file"))
    rc = mode_edges_file();
  else if (!strcmp(argv[1], "embedding_io"))
    rc = mode_embedding_io();
  else if (!strcmp(argv[1], "spring_tutte"))
    rc = mode_spring_tutte();
  else if (!strcmp(argv[1], "tutte_begin"))
    rc = mode_tutte_begin();
  else if (!strcmp(argv[1], "grip_checkpoint"))
    rc = mode_grip_checkpoint();
  else if (!strcmp(argv[1], "force_checkpoint"))
    rc = mode_force_checkpoint();
  else if (!strcmp(argv[1], "quadtree"))
    rc = mode_quadtree();
  else if (!strcmp(argv[1], "force_quadtree"))
    rc = mode_force_quadtree();
  else if (!strcmp(argv[1], "quadtree_cases"))
    rc = mode_quadtree_cases();
  else if (!strcmp(argv[1], "knn"))
    rc = mode_knn();
  else if (!strcmp(argv[1], "knn_cases"))
    rc = mode_knn_cases();
  else if (!strcmp(argv[1], "grip_stages"))
    rc = mode_grip_stages();
  else if (!strcmp(argv[1], "tutte_outer_face"))
    rc = mode_tutte_outer_face();
  fclose(fin);
  fclose(fout);
  return rc;
}
