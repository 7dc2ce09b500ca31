/*
 * ref_harness.c -- drives the reference library (oracle/_ref/libgviz_ref.so,
 * compiled from the reference's own sources) on inputs written by
 * tests/golden/make_golden.py and writes its outputs as raw little-endian
 * arrays. Used only in a container that has the reference tree: to write the
 * committed golden fixtures and to cross-check the FP64 restatement.
 *
 * Usage: ref_harness <mode> <in.bin> <out.bin>
 * Input files are a sequence of int64 / double fields in the order each mode
 * reads them; outputs likewise (see tests/golden/refio.py).
 */
#include "algorithms/search/gvizKNearest.h"
#include "ds/gvizGraph.h"
#include "ds/gvizQuadtree.h"
#include "ds/gvizSubgraph.h"
#include "embedders/gvizEmbeddedGraph.h"
#include "embedders/gvizForceEmbedder.h"
#include "embedders/gvizGRIPEmbedder.h"
#include "embedders/gvizGRIPInternal.h"
#include "embedders/gvizTutteEmbedder.h"
#include "embedders/gvizSpringTutteEmbedder.h"
#include "core/gvizThreadPool.h"
#include "utils/graphLoader.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

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

/* graph: n, m, then m (u, v) pairs in insertion order */
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

/* ---------------------------------------------------------------- quadtree */
static void dump_node(const gvizQuadtreeNode *nd, int depth, uint64_t kh,
                      uint64_t kl, int64_t *leafDepth, uint64_t *leafHi,
                      uint64_t *leafLo, int64_t *nodeCount) {
  if (!nd)
    return;
  (*nodeCount)++;
  /* node record: depth, keyHi, keyLo, isLeaf, count, mass, comX, comY, h */
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
    dump_node(gvizQuadtreeNodeChild(nd, (gvizQuadrant)q), depth + 1, h, l,
              leafDepth, leafHi, leafLo, nodeCount);
  }
}

static int mode_quadtree(void) {
  int64_t n = rd_i(), npc = rd_i();
  double *pts = malloc(sizeof(double) * 2 * n);
  double *ms = malloc(sizeof(double) * n);
  for (int64_t i = 0; i < 2 * n; i++)
    pts[i] = rd_d();
  for (int64_t i = 0; i < n; i++)
    ms[i] = rd_d();
  gvizQuadtree t;
  if (gvizQuadtreeInit(&t, pts, ms, (size_t)n, (size_t)npc) != 0)
    return 1;
  int64_t *ld = calloc(n, 8);
  uint64_t *lh = calloc(n, 8), *ll = calloc(n, 8);
  int64_t nodeCount = 0;
  dump_node(gvizQuadtreeRoot(&t), 0, 0, 0, ld, lh, ll, &nodeCount);
  for (int64_t i = 0; i < n; i++) {
    wr_i(ld[i]);
    wr_i((int64_t)lh[i]);
    wr_i((int64_t)ll[i]);
  }
  wr_i(nodeCount);
  gvizQuadtreeRelease(&t);
  return 0;
}

/* ------------------------------------------------------------------- force */
static int mode_force(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t model = rd_i(), bh = rd_i(), overlap = rd_i(), npc = rd_i();
  int64_t iters = rd_i(), dumpEvery = rd_i();
  double edgeLength = rd_d(), theta = rd_d(), gravityK = rd_d();
  double overlapC = rd_d(), rBase = rd_d(), rPerDeg = rd_d(), jitter = rd_d();
  double *pos = malloc(sizeof(double) * 2 * n);
  for (int64_t i = 0; i < 2 * n; i++)
    pos[i] = rd_d();

  gvizForceEmbedderState s;
  int rc = gvizForceEmbedderInit(&s, gvizSubgraphCreateFull(&g), 2,
                                 (gvizForceModelKind)model);
  if (rc != 0)
    return 1;
  gvizForceEmbedderConfigure(&s, edgeLength, 0.0);
  gvizForceEmbedderConfigureSpeed(&s, jitter);
  gvizForceEmbedderConfigureBarnesHut(&s, theta, (size_t)npc);
  gvizForceEmbedderSetBarnesHutEnabled(&s, (int)bh);
  gvizForceEmbedderConfigureGravity(&s, gravityK);
  gvizForceEmbedderConfigureRadius(&s, rBase, rPerDeg);
  gvizForceEmbedderConfigureOverlapPrevention(&s, overlapC);
  gvizForceEmbedderSetPreventOverlapEnabled(&s, (int)overlap);
  gvizForceEmbedderBegin(&s, 12345u);
  for (int64_t i = 0; i < n; i++)
    gvizEmbeddedGraphSetVPosition((gvizEmbeddedGraph *)&s, (size_t)i,
                                  pos + 2 * i);
  for (int64_t it = 0; it < iters; it++) {
    double maxDisp = gvizForceEmbedderStep(&s);
    const gvizStatSeries *sp =
        gvizEmbeddedGraphFindStatSeries((gvizEmbeddedGraph *)&s,
                                        "forceEmbedder.speed");
    const gvizStatSeries *sa =
        gvizEmbeddedGraphFindStatSeries((gvizEmbeddedGraph *)&s,
                                        "forceEmbedder.attractiveForce");
    const gvizStatSeries *sr =
        gvizEmbeddedGraphFindStatSeries((gvizEmbeddedGraph *)&s,
                                        "forceEmbedder.repulsiveForce");
    const gvizStatSeries *sgv =
        gvizEmbeddedGraphFindStatSeries((gvizEmbeddedGraph *)&s,
                                        "forceEmbedder.gravityForce");
    wr_d(maxDisp);
    wr_d(sp->samples[sp->count - 1]);
    wr_d(sa->samples[sa->count - 1]);
    wr_d(sr->samples[sr->count - 1]);
    wr_d(sgv->samples[sgv->count - 1]);
    wr_d(s.speedEfficiency);
    int dump = dumpEvery > 0 && (it % dumpEvery == 0 || it == iters - 1);
    if (dump) {
      for (int64_t i = 0; i < 2 * n; i++)
        wr_d(s.disp[i]);
      for (int64_t i = 0; i < 2 * n; i++)
        wr_d(s.structForce[i]);
      for (int64_t i = 0; i < n; i++)
        wr_d(s.attForceMag[i]);
      for (int64_t i = 0; i < n; i++)
        wr_d(s.repForceMag[i]);
      const double *p = gvizEmbeddedGraphPositions((gvizEmbeddedGraph *)&s);
      for (int64_t i = 0; i < 2 * n; i++)
        wr_d(p[i]);
    }
  }
  gvizForceEmbedderRelease(&s);
  gvizGraphRelease(&g);
  return 0;
}

/* CPU baseline: wall time of each gvizForceEmbedderStep with the reference's
 * own thread pool of `workers` threads (plus the calling thread). */
static int mode_time_force(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t model = rd_i(), bh = rd_i(), overlap = rd_i(), npc = rd_i();
  int64_t iters = rd_i(), workers = rd_i();
  double edgeLength = rd_d(), theta = rd_d(), gravityK = rd_d();
  double overlapC = rd_d(), rBase = rd_d(), rPerDeg = rd_d(), jitter = rd_d();
  double *pos = malloc(sizeof(double) * 2 * n);
  for (int64_t i = 0; i < 2 * n; i++)
    pos[i] = rd_d();
  gvizForceEmbedderState s;
  if (gvizForceEmbedderInit(&s, gvizSubgraphCreateFull(&g), 2,
                            (gvizForceModelKind)model) != 0)
    return 1;
  gvizThreadPoolDestroy(s.pool);
  s.pool = workers > 0 ? gvizThreadPoolCreate((size_t)workers) : NULL;
  gvizForceEmbedderConfigure(&s, edgeLength, 0.0);
  gvizForceEmbedderConfigureSpeed(&s, jitter);
  gvizForceEmbedderConfigureBarnesHut(&s, theta, (size_t)npc);
  gvizForceEmbedderSetBarnesHutEnabled(&s, (int)bh);
  gvizForceEmbedderConfigureGravity(&s, gravityK);
  gvizForceEmbedderConfigureRadius(&s, rBase, rPerDeg);
  gvizForceEmbedderConfigureOverlapPrevention(&s, overlapC);
  gvizForceEmbedderSetPreventOverlapEnabled(&s, (int)overlap);
  gvizForceEmbedderBegin(&s, 12345u);
  for (int64_t i = 0; i < n; i++)
    gvizEmbeddedGraphSetVPosition((gvizEmbeddedGraph *)&s, (size_t)i,
                                  pos + 2 * i);
  for (int64_t it = 0; it < iters; it++) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    gvizForceEmbedderStep(&s);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    wr_d((double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec));
    fflush(fout);
  }
  gvizForceEmbedderRelease(&s);
  gvizGraphRelease(&g);
  return 0;
}

/* ------------------------------------------------------------------- tutte */
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
  for (int64_t it = 0; it < iters; it++)
    wr_d(gvizTutteEmbedderStep(&s, dt));
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  gvizTutteEmbedderRelease(&s);
  gvizGraphRelease(&g);
  return 0;
}

/* --------------------------------------------------------------------- kNN */
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
  for (int64_t s = 0; s < nsrc; s++) {
    int64_t src = rd_i();
    int c = gvizSearchKNearest(&sg, out, (size_t)k, (size_t)src, vis);
    wr_i(c);
    for (int64_t j = 0; j < k; j++) {
      wr_i(j < c ? (int64_t)out[j].v : -1);
      wr_i(j < c ? (int64_t)out[j].dist : -1);
    }
  }
  /* batch path over all non-visible vertices as targets */
  int64_t nt = 0;
  gvizKNNBatchTarget *targets = malloc(sizeof(*targets) * n);
  gvizFoundVertex *found = malloc(sizeof(gvizFoundVertex) * n * k);
  for (int64_t i = 0; i < n; i++)
    if (!gvizVertexSubsetTest(vis, (size_t)i)) {
      targets[nt].vertex = (size_t)i;
      targets[nt].out = found + nt * k;
      targets[nt].count = 0;
      nt++;
    }
  gvizKNearestScratch scratch;
  gvizKNearestScratchInit(&scratch, (size_t)n);
  int rc = gvizSearchKNearestFromVisibleBatch(&sg, vis, (size_t)k, targets,
                                              (size_t)nt, &scratch);
  wr_i(rc);
  wr_i(nt);
  for (int64_t t = 0; t < nt; t++) {
    wr_i((int64_t)targets[t].vertex);
    wr_i((int64_t)targets[t].count);
    for (int64_t j = 0; j < k; j++) {
      wr_i(j < (int64_t)targets[t].count ? (int64_t)targets[t].out[j].v : -1);
      wr_i(j < (int64_t)targets[t].count ? (int64_t)targets[t].out[j].dist
                                          : -1);
    }
  }
  gvizKNearestScratchRelease(&scratch);
  return 0;
}

/* -------------------------------------------------------------------- GRIP */
/* The GRIP refinement pass sets dispCalculated bits of a shared bit array
 * from every pool thread with a plain read-modify-write
 * (src/embedders/gvizGRIPEmbedder.c:1084-1086, include/ds/gvizBitArray.h:15),
 * so under the pool a bit is now and then lost and that vertex's heat
 * restarts in the next round. The fixtures record the reference's defined
 * serial behaviour: the pool is removed (NULL pool: every phase runs on the
 * calling thread). */
static void serial_grip(gvizGRIPState *st) {
  gvizThreadPoolDestroy(st->pool);
  st->pool = NULL;
}

static void dump_grip_state(gvizGRIPState *st, int64_t n, int64_t dim) {
  const double *p = gvizEmbeddedGraphPositions((gvizEmbeddedGraph *)st);
  for (int64_t i = 0; i < dim * n; i++)
    wr_d(p[i]);
  for (int64_t v = 0; v < n; v++) {
    gvizGRIPDecorators *d = &st->dec[v];
    for (int64_t k = 0; k < dim; k++)
      wr_d(d->disp[k]);
    for (int64_t k = 0; k < dim; k++)
      wr_d(d->oldDisp[k]);
    wr_d(d->heat);
    wr_d(d->oldCos);
    wr_d(d->lastForceMag);
    wr_i((int64_t)d->knn.count);
    wr_i(gvizTestBit(st->dispCalculated, (size_t)v) ? 1 : 0);
  }
}

static int mode_grip(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t kcap = rd_i(), kmax = rd_i(), roundsPerLayer = rd_i(), dim = rd_i();
  gvizGRIPState st;
  memset(&st, 0, sizeof(st));
  gvizGRIPEmbedderConfigureKnnCapacity(&st, (size_t)kcap);
  if (gvizGRIPEmbedderInit(&st, gvizSubgraphCreateFull(&g), 0, (size_t)dim) != 0)
    return 1;
  serial_grip(&st);
  gvizGRIPEmbedderConfigureK(&st, (size_t)kmax, (size_t)kmax,
                             GVIZ_GRIP_K_CONSTANT);
  gvizGRIPEmbedderBegin(&st);
  wr_i((int64_t)st.layerCount);
  for (size_t i = 0; i < st.layerCount; i++)
    wr_i((int64_t)*(size_t *)gvizArrayAtIndex(&st.misBorder, i));
  size_t active0 = *(size_t *)gvizArrayAtIndex(&st.misBorder, 0);
  for (size_t i = 0; i < active0; i++)
    wr_i((int64_t)st.misFiltration[i]);
  for (;;) {
    for (int64_t r = 0; r < roundsPerLayer; r++) {
      wr_i((int64_t)st.currLayer);
      dump_grip_state(&st, n, dim);
      /* kNN lists of every vertex, kcap slots */
      for (int64_t v = 0; v < n; v++) {
        gvizGRIPDecorators *d = &st.dec[v];
        for (int64_t j = 0; j < kcap; j++) {
          gvizFoundVertex *f = (gvizFoundVertex *)d->knn.arr + j;
          int ok = j < (int64_t)d->knn.count;
          wr_i(ok ? (int64_t)f->v : -1);
          wr_i(ok ? (int64_t)f->dist : -1);
        }
      }
      gvizGRIPEmbedderRefineRound(&st);
      dump_grip_state(&st, n, dim);
      gvizGRIPRoundStats rs = gvizGRIPEmbedderLastRoundStats(&st);
      wr_d(rs.maxDisplacement);
      wr_d(rs.meanDisplacement);
      wr_d(rs.meanForce);
    }
    if (st.currLayer == 0)
      break;
    gvizGRIPEmbedderNextStage(&st);
  }
  wr_i(-1);
  gvizGRIPEmbedderRelease(&st);
  return 0;
}

/* ------------------------------------------------------- embedder API runs */
static void write_series(gvizEmbeddedGraph *e, const char *name) {
  const gvizStatSeries *s = gvizEmbeddedGraphFindStatSeries(e, name);
  wr_i(s ? (int64_t)s->count : -1);
  if (s)
    for (size_t i = 0; i < s->count; i++)
      wr_d(s->samples[i]);
}

/* Multiplies every coordinate by 1 + amount * u, u uniform in [-1, 1) from a
 * 64-bit LCG: the same start up to rounding-size noise, to measure how far a
 * run's chaotic FA2 dynamics carry such differences. amount 0: no change. */
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

/* Force embedder through its public API on a vertex-induced subgraph:
 * Init, Configure*, Begin(seed) (the reference's own random placement),
 * Run(maxIters, epsilon), then one more Step. */
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
  gvizForceEmbedderConfigure(&s, edgeLength, 0.0);
  gvizForceEmbedderConfigureSpeed(&s, jitter);
  gvizForceEmbedderConfigureBarnesHut(&s, theta, (size_t)npc);
  gvizForceEmbedderSetBarnesHutEnabled(&s, (int)bh);
  gvizForceEmbedderConfigureGravity(&s, gravityK);
  gvizForceEmbedderConfigureRadius(&s, rBase, rPerDeg);
  gvizForceEmbedderConfigureOverlapPrevention(&s, overlapC);
  gvizForceEmbedderSetPreventOverlapEnabled(&s, (int)overlap);
  wr_d(s.boxExtent);
  wr_i(gvizForceEmbedderBegin(&s, (unsigned int)seed));
  const double *p = gvizEmbeddedGraphPositions(e);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  perturb_positions(e->embedding.vertexPositions, 2 * n, perturb, perturbSeed);
  wr_i(gvizForceEmbedderRun(&s, (size_t)runIters, eps));
  static const char *const names[] = {
      "forceEmbedder.maxDisp", "forceEmbedder.speed",
      "forceEmbedder.attractiveForce", "forceEmbedder.repulsiveForce",
      "forceEmbedder.gravityForce"};
  for (int k = 0; k < 5; k++)
    write_series(e, names[k]);
  p = gvizEmbeddedGraphPositions(e);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  wr_d(gvizForceEmbedderStep(&s));
  p = gvizEmbeddedGraphPositions(e);
  for (int64_t i = 0; i < 2 * n; i++)
    wr_d(p[i]);
  wr_d(s.globalSpeed);
  wr_d(s.speedEfficiency);
  wr_i((int64_t)s.iteration);
  gvizForceEmbedderRelease(&s);
  gvizGraphRelease(&g);
  return 0;
}

/* gvizGRIPEmbedderEmbed end to end with a K policy. */
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
  serial_grip(&st);
  gvizGRIPEmbedderConfigureK(&st, (size_t)pk, (size_t)rk, (gvizGRIPKPolicy)policy);
  wr_i(gvizGRIPEmbedderEmbed(&st));
  wr_i((int64_t)st.layerCount);
  for (size_t i = 0; i < st.layerCount; i++)
    wr_i((int64_t)*(size_t *)gvizArrayAtIndex(&st.misBorder, i));
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
  return 0;
}

/* The reference's Embed schedule (gvizGRIPEmbedder.c:1262-1285: rounds of a
 * layer until maxDisplacement < 10 * 1e-3 or 30 rounds, then NextStage) from
 * Begin's placement with every coordinate scaled by 1 + perturb * u: the
 * spread of the final layouts under rounding-size start differences.
 * Writes the final positions. */
static int mode_api_grip_spread(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t kcap = rd_i(), pk = rd_i(), rk = rd_i(), policy = rd_i(), dim = rd_i();
  double perturb = rd_d();
  int64_t perturbSeed = rd_i();
  gvizGRIPState st;
  memset(&st, 0, sizeof(st));
  gvizGRIPEmbedderConfigureKnnCapacity(&st, (size_t)kcap);
  if (gvizGRIPEmbedderInit(&st, gvizSubgraphCreateFull(&g), 0, (size_t)dim) != 0)
    return 1;
  serial_grip(&st);
  gvizGRIPEmbedderConfigureK(&st, (size_t)pk, (size_t)rk, (gvizGRIPKPolicy)policy);
  gvizGRIPEmbedderBegin(&st);
  gvizEmbeddedGraph *e = (gvizEmbeddedGraph *)&st;
  perturb_positions(e->embedding.vertexPositions, dim * n, perturb, perturbSeed);
  const double epsilon = 10.0 * 1e-3;
  for (;;) {
    for (int round = 0; round < 30; round++) {
      gvizGRIPEmbedderRefineRound(&st);
      if (st.lastRoundStats.maxDisplacement < epsilon)
        break;
    }
    if (st.currLayer == 0)
      break;
    gvizGRIPEmbedderNextStage(&st);
  }
  const double *p = gvizEmbeddedGraphPositions(e);
  for (int64_t i = 0; i < dim * n; i++)
    wr_d(p[i]);
  gvizGRIPEmbedderRelease(&st);
  return 0;
}

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* CPU baseline for GRIP: the reference's own schedule on the subgraph
 * induced by `active`, with its thread pool of `workers` threads. Writes
 * the Begin time, then per stage (coarsest first) the NextStage time (0 for
 * the first) and the times of `rounds` RefineRound calls; stops after
 * `maxStages` stages. */
static int mode_time_grip(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  gvizVertexSubset vs = gvizVertexSubsetCreateEmpty(&g);
  for (int64_t i = 0; i < n; i++)
    if (rd_i())
      gvizVertexSubsetShowVertex(vs, (size_t)i);
  int64_t kcap = rd_i(), rounds = rd_i(), maxStages = rd_i(), workers = rd_i();
  gvizGRIPState st;
  memset(&st, 0, sizeof(st));
  gvizGRIPEmbedderConfigureKnnCapacity(&st, (size_t)kcap);
  if (gvizGRIPEmbedderInit(&st, gvizSubgraphCreateVertexInduced(&g, vs), 0, 2) != 0)
    return 1;
  gvizThreadPoolDestroy(st.pool);
  st.pool = workers > 0 ? gvizThreadPoolCreate((size_t)workers) : NULL;
  gvizGRIPEmbedderConfigureK(&st, (size_t)kcap, (size_t)kcap, GVIZ_GRIP_K_CONSTANT);
  double t0 = now_s();
  gvizGRIPEmbedderBegin(&st);
  wr_d(now_s() - t0);
  wr_i((int64_t)st.layerCount);
  fflush(fout);
  for (int64_t stage = 0; stage < maxStages; stage++) {
    double tn = 0.0;
    if (stage > 0) {
      if (st.currLayer == 0)
        break;
      t0 = now_s();
      gvizGRIPEmbedderNextStage(&st);
      tn = now_s() - t0;
    }
    wr_i((int64_t)st.currLayer);
    wr_i((int64_t)*(size_t *)gvizArrayAtIndex(&st.misBorder, st.currLayer));
    wr_d(tn);
    for (int64_t r = 0; r < rounds; r++) {
      t0 = now_s();
      gvizGRIPEmbedderRefineRound(&st);
      wr_d(now_s() - t0);
    }
    fflush(fout);
  }
  wr_i(-1);
  return 0;
}

/* CPU baseline for the Tutte sweep: `iters` Jacobi Step(1.0) calls after
 * FixConvexPolygon + SeedInterior; writes each call's wall time. */
static int mode_time_tutte(void) {
  gvizGraph g;
  int64_t n;
  read_graph(&g, &n);
  int64_t nb = rd_i(), iters = rd_i();
  size_t *boundary = malloc(sizeof(size_t) * nb);
  for (int64_t i = 0; i < nb; i++)
    boundary[i] = (size_t)rd_i();
  double radius = rd_d();
  gvizTutteState s;
  if (gvizTutteEmbedderInit(&s, gvizSubgraphCreateFull(&g), 2, 1e-300) != 0)
    return 1;
  gvizTutteFixConvexPolygon(&s, boundary, (size_t)nb, radius);
  gvizTutteEmbedderSeedInterior(&s);
  for (int64_t it = 0; it < iters; it++) {
    double t0 = now_s();
    gvizTutteEmbedderStep(&s, 1.0);
    wr_d(now_s() - t0);
    fflush(fout);
  }
  return 0;
}

/* The reference's .edges loader on a file named in the input: directed,
 * skip_header, path length and path; writes the return code and, on
 * success, the vertex count and every adjacency list in order. */
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
  free(path);
  wr_i(rc);
  if (rc != 0)
    return 0;
  const size_t n = gvizGraphSize(&g);
  wr_i((int64_t)n);
  for (size_t u = 0; u < n; u++) {
    gvizArray *nb = gvizGraphGetVertexNeighbors(&g, u);
    wr_i((int64_t)nb->count);
    for (size_t k = 0; k < nb->count; k++)
      wr_i((int64_t) * (size_t *)gvizArrayAtIndex(nb, k));
  }
  gvizGraphRelease(&g);
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

int main(int argc, char **argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s mode in out\n", argv[0]);
    return 2;
  }
  fin = fopen(argv[2], "rb");
  fout = fopen(argv[3], "wb");
  if (!fin || !fout)
    return 2;
  int rc = 2;
  if (!strcmp(argv[1], "quadtree"))
    rc = mode_quadtree();
  else if (!strcmp(argv[1], "force"))
    rc = mode_force();
  else if (!strcmp(argv[1], "time_force"))
    rc = mode_time_force();
  else if (!strcmp(argv[1], "tutte"))
    rc = mode_tutte();
  else if (!strcmp(argv[1], "knn"))
    rc = mode_knn();
  else if (!strcmp(argv[1], "grip"))
    rc = mode_grip();
  else if (!strcmp(argv[1], "api_force"))
    rc = mode_api_force();
  else if (!strcmp(argv[1], "api_grip"))
    rc = mode_api_grip();
  else if (!strcmp(argv[1], "api_grip_spread"))
    rc = mode_api_grip_spread();
  else if (!strcmp(argv[1], "time_grip"))
    rc = mode_time_grip();
  else if (!strcmp(argv[1], "time_tutte"))
    rc = mode_time_tutte();
  else if (!strcmp(argv[1], "edges_file"))
    rc = mode_edges_file();
  else if (!strcmp(argv[1], "embedding_io"))
    rc = mode_embedding_io();
  else if (!strcmp(argv[1], "spring_tutte"))
    rc = mode_spring_tutte();
  else if (!strcmp(argv[1], "tutte_begin"))
    rc = mode_tutte_begin();
  fclose(fin);
  fclose(fout);
  return rc;
}
