/*
 * gviz_oracle.c -- FP64 CPU restatement of the reference's force embedder,
 * quadtree, Tutte sweep and layout metrics.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine (tests/,
 * __graft_entry__.smoke(), bench.py cpu_baseline). The product never links it.
 *
 * Arithmetic is written in the reference's evaluation order so that, built
 * with gcc on x86-64 (no FMA contraction), results match the reference
 * bit for bit. Citations are file:line in the reference tree.
 */
#include "gviz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OC_NUMERIC_EPS 1e-7      /* include/embedders/gvizForceDirected.h:6 */
#define OC_QT_MIN_HALF 1e-9      /* src/ds/gvizQuadtree.c:6 */
#define OC_QT_MIN_MASS 1e-9      /* src/ds/gvizQuadtree.c:7 */
#define OC_EFF_MIN 0.05          /* gvizForceEmbedder.h SPEED_EFFICIENCY_MIN */
#define OC_EFF_MAX 1.0
#define OC_MAX_RISE 0.5

/* ------------------------------------------------------ pairwise forces -- */

/* Replacement for include/core/gvizVec.h:385-394, which hashes buffer
 * addresses. Here the hash runs over the two vertex ids (splitmix64 finaliser)
 * and maps to an angle the same way; the GPU engine uses the identical rule. */
void oc_coincident_dir(int64_t selfId, int64_t otherId, double *ux,
                       double *uy) {
  uint64_t bits = (uint64_t)selfId ^ ((uint64_t)otherId * 0x9E3779B97F4A7C15ULL);
  bits ^= bits >> 33;
  bits *= 0xff51afd7ed558ccdULL;
  bits ^= bits >> 33;
  double angle = (double)(bits % 1000000u) / 1000000.0 * 2.0 * M_PI;
  *ux = cos(angle);
  *uy = sin(angle);
}

/* include/core/gvizVec.h:320-335 (n = 2) */
void oc_fr_att(const double *v, const double *u, double k, double *acc) {
  double dx = u[0] - v[0];
  double dy = u[1] - v[1];
  double dist = sqrt(dx * dx + dy * dy);
  if (dist < 1e-9)
    return;
  double s = dist / k;
  acc[0] += s * dx;
  acc[1] += s * dy;
}

/* include/core/gvizVec.h:486-500 (n = 2) */
void oc_linlog_att(const double *v, const double *u, double *acc) {
  double dx = u[0] - v[0];
  double dy = u[1] - v[1];
  double dist = sqrt(dx * dx + dy * dy);
  if (dist < 1e-9)
    return;
  double s = log(1.0 + dist) / dist;
  acc[0] += s * dx;
  acc[1] += s * dy;
}

/* include/core/gvizVec.h:404-427 (n = 2) */
void oc_fr_rep(const double *v, const double *u, double k, double radiusSum,
               double overlapConstant, int64_t selfId, int64_t otherId,
               double *acc) {
  double kSq = k * k;
  double dx = u[0] - v[0];
  double dy = u[1] - v[1];
  double dist_sq = dx * dx + dy * dy;
  double dist = sqrt(dist_sq);
  double gap = dist - radiusSum;
  double mag = gap > 0.0 ? kSq / gap : kSq * overlapConstant;
  double ux, uy;
  if (dist > 0.0) {
    ux = dx / dist;
    uy = dy / dist;
  } else {
    oc_coincident_dir(selfId, otherId, &ux, &uy);
  }
  acc[0] -= mag * ux;
  acc[1] -= mag * uy;
}

/* include/core/gvizVec.h:552-575 (n = 2) */
void oc_linlog_rep(const double *v, const double *u, double vMass,
                   double otherMass, double radiusSum, double overlapConstant,
                   int64_t selfId, int64_t otherId, double *acc) {
  double massProduct = vMass * otherMass;
  double dx = u[0] - v[0];
  double dy = u[1] - v[1];
  double dist_sq = dx * dx + dy * dy;
  double dist = sqrt(dist_sq);
  double gap = dist - radiusSum;
  double mag = gap > 0.0 ? massProduct / gap : massProduct * overlapConstant;
  double ux, uy;
  if (dist > 0.0) {
    ux = dx / dist;
    uy = dy / dist;
  } else {
    oc_coincident_dir(selfId, otherId, &ux, &uy);
  }
  acc[0] -= mag * ux;
  acc[1] -= mag * uy;
}

/* include/core/gvizVec.h:636-647 (n = 2) */
void oc_gravity(const double *v, double magnitude, double *acc) {
  double dist = sqrt(v[0] * v[0] + v[1] * v[1]);
  if (dist < 1e-9)
    return;
  double s = magnitude / dist;
  acc[0] -= s * v[0];
  acc[1] -= s * v[1];
}

/* include/core/gvizVec.h:138-150 (n = 2) */
void oc_kk(const double *v, const double *u, double L, double *acc) {
  if (L < 1e-9)
    return;
  double dx = u[0] - v[0];
  double dy = u[1] - v[1];
  double s = sqrt(dx * dx + dy * dy) / L - 1.0;
  acc[0] += s * dx;
  acc[1] += s * dy;
}

/* include/core/gvizVec.h:192-207 (n = 2) */
void oc_grip_fr_rep(const double *v, const double *u, double edgeLength,
                    double *acc) {
  double edgeLenSq = edgeLength * edgeLength;
  double dx = u[0] - v[0];
  double dy = u[1] - v[1];
  double dist_sq = dx * dx + dy * dy;
  if (dist_sq < 1e-18)
    return;
  double s = dist_sq / edgeLenSq;
  acc[0] += s * dx;
  acc[1] += s * dy;
}

/* include/core/gvizVec.h:257-272 (n = 2) */
void oc_grip_fr_att(const double *v, const double *u, double edgeLength,
                    double frScale, double *acc) {
  double edgeLenSq = edgeLength * edgeLength;
  double dx = v[0] - u[0];
  double dy = v[1] - u[1];
  double dist_sq = dx * dx + dy * dy;
  if (dist_sq < 1e-18)
    return;
  double s = frScale * edgeLenSq / dist_sq;
  acc[0] += s * dx;
  acc[1] += s * dy;
}

/* src/embedders/gvizForceModel.c:5-8,31 */
double oc_vertex_mass(int32_t model, int64_t degree) {
  return model == OC_MODEL_LINLOG ? 1.0 + (double)degree : 1.0;
}

/* src/embedders/gvizForceEmbedder.c:389-393 */
double oc_vertex_radius(const oc_force_params *p, int64_t degree) {
  return p->radiusBase * (1.0 + p->radiusPerDegree * sqrt((double)degree));
}

/* The model's repulsive() callback: src/embedders/gvizForceModel.c:15-23
 * (FR: scratch vector, then acc += (size_t)otherMass * scratch, see
 * src/embedders/gvizForceDirected.c:96-104) and :39-46 (LinLog). */
static inline void model_repulsive(int32_t model, const double *v,
                                   const double *other, double vMass,
                                   double otherMass, double vRadius,
                                   double otherRadius, double overlapConstant,
                                   double edgeLength, int64_t selfId,
                                   int64_t otherId, double *acc) {
  if (model == OC_MODEL_LINLOG) {
    oc_linlog_rep(v, other, vMass, otherMass, vRadius + otherRadius,
                  overlapConstant, selfId, otherId, acc);
  } else {
    double scratch[2] = {0.0, 0.0};
    oc_fr_rep(v, other, edgeLength, vRadius + otherRadius, overlapConstant,
              selfId, otherId, scratch);
    double m = (double)(size_t)otherMass;
    acc[0] += m * scratch[0];
    acc[1] += m * scratch[1];
  }
}

static inline void model_attractive(int32_t model, const double *v,
                                    const double *u, double edgeLength,
                                    double *acc) {
  if (model == OC_MODEL_LINLOG)
    oc_linlog_att(v, u, acc);
  else
    oc_fr_att(v, u, edgeLength, acc);
}

/* ------------------------------------------------------------ quadtree -- */

struct oc_quadtree {
  const double *points;
  const double *masses;
  int64_t n;
  int64_t npc;
  oc_qnode *nodes;
  int64_t nodeCount, nodeCap;
  /* per-node point lists while building */
  int64_t **plist;
  int64_t *pcap;
  int64_t *leafPts; /* compacted, filled at the end */
  int32_t root;
  double rootCx, rootCy, rootHalf;
  int32_t maxDepth;
};

static int32_t qt_alloc(struct oc_quadtree *t, double cx, double cy, double h,
                        int32_t depth, uint64_t kh, uint64_t kl) {
  if (t->nodeCount == t->nodeCap) {
    int64_t cap = t->nodeCap ? t->nodeCap * 2 : 1024;
    t->nodes = realloc(t->nodes, sizeof(oc_qnode) * (size_t)cap);
    t->plist = realloc(t->plist, sizeof(int64_t *) * (size_t)cap);
    t->pcap = realloc(t->pcap, sizeof(int64_t) * (size_t)cap);
    t->nodeCap = cap;
  }
  int32_t id = (int32_t)t->nodeCount++;
  oc_qnode *nd = &t->nodes[id];
  nd->cx = cx;
  nd->cy = cy;
  nd->halfSize = h;
  nd->comX = 0.0;
  nd->comY = 0.0;
  nd->mass = 0.0;
  nd->child[0] = nd->child[1] = nd->child[2] = nd->child[3] = -1;
  nd->depth = depth;
  nd->ptStart = 0;
  nd->ptCount = 0;
  nd->keyHi = kh;
  nd->keyLo = kl;
  t->plist[id] = malloc(sizeof(int64_t) * (size_t)t->npc);
  t->pcap[id] = t->npc;
  return id;
}

static void key_with_digit(uint64_t kh, uint64_t kl, int32_t level, int digit,
                           uint64_t *oh, uint64_t *ol) {
  *oh = kh;
  *ol = kl;
  if (level <= 32)
    *oh |= (uint64_t)digit << (2 * (32 - level));
  else if (level <= 64)
    *ol |= (uint64_t)digit << (2 * (64 - level));
}

/* src/ds/gvizQuadtree.c:65-73: east = px >= cx, north = py >= cy;
 * NW 0, NE 1, SW 2, SE 3. */
static inline int qt_quadrant(double cx, double cy, double px, double py) {
  int east = px >= cx;
  int north = py >= cy;
  if (north)
    return east ? 1 : 0;
  return east ? 3 : 2;
}

/* src/ds/gvizQuadtree.c:75-92 */
static void qt_subdivide(struct oc_quadtree *t, int32_t id) {
  double quarter = t->nodes[id].halfSize / 2.0;
  double offs[4][2] = {{-quarter, quarter},
                       {quarter, quarter},
                       {-quarter, -quarter},
                       {quarter, -quarter}};
  for (int i = 0; i < 4; i++) {
    oc_qnode p = t->nodes[id]; /* realloc-safe copy */
    uint64_t kh, kl;
    key_with_digit(p.keyHi, p.keyLo, p.depth + 1, i, &kh, &kl);
    int32_t c = qt_alloc(t, p.cx + offs[i][0], p.cy + offs[i][1], quarter,
                         p.depth + 1, kh, kl);
    t->nodes[id].child[i] = c;
  }
}

/* src/ds/gvizQuadtree.c:112-157 */
static void qt_insert(struct oc_quadtree *t, int32_t id, int64_t idx) {
  for (;;) {
    double px = t->points[2 * idx];
    double py = t->points[2 * idx + 1];
    double ptMass = t->masses[idx];
    oc_qnode *nd = &t->nodes[id];
    double newMass = nd->mass + ptMass;
    if (newMass > OC_QT_MIN_MASS || newMass < -OC_QT_MIN_MASS) {
      nd->comX = (nd->comX * nd->mass + px * ptMass) / newMass;
      nd->comY = (nd->comY * nd->mass + py * ptMass) / newMass;
    }
    nd->mass = newMass;

    if (nd->child[0] >= 0) {
      id = nd->child[qt_quadrant(nd->cx, nd->cy, px, py)];
      continue;
    }
    if (nd->ptCount < t->pcap[id]) {
      t->plist[id][nd->ptCount++] = idx;
      return;
    }
    if (nd->halfSize <= OC_QT_MIN_HALF) {
      int64_t cap = t->pcap[id] > 0 ? t->pcap[id] * 2 : 1;
      t->plist[id] = realloc(t->plist[id], sizeof(int64_t) * (size_t)cap);
      t->pcap[id] = cap;
      t->plist[id][nd->ptCount++] = idx;
      return;
    }
    qt_subdivide(t, id);
    nd = &t->nodes[id];
    int64_t held = nd->ptCount;
    for (int64_t i = 0; i < held; i++) {
      int64_t h = t->plist[id][i];
      oc_qnode *cur = &t->nodes[id];
      int32_t c = cur->child[qt_quadrant(cur->cx, cur->cy, t->points[2 * h],
                                         t->points[2 * h + 1])];
      qt_insert(t, c, h);
    }
    t->nodes[id].ptCount = 0;
    nd = &t->nodes[id];
    id = nd->child[qt_quadrant(nd->cx, nd->cy, px, py)];
  }
}

static int32_t depth_limit(double halfSize) {
  /* first depth whose half size (exact halvings) is <= 1e-9 */
  int32_t d = 0;
  double h = halfSize;
  while (!(h <= OC_QT_MIN_HALF) && d < 4096) {
    h = h / 2.0;
    d++;
  }
  return d;
}

/* src/ds/gvizQuadtree.c:159-192 */
oc_quadtree *oc_quadtree_build(const double *points, const double *masses,
                               int64_t n, int64_t nodesPerCell) {
  struct oc_quadtree *t = calloc(1, sizeof(*t));
  t->points = points;
  t->masses = masses;
  t->n = n;
  t->npc = nodesPerCell > 0 ? nodesPerCell : 1;
  t->root = -1;
  if (n == 0)
    return t;
  double minX = points[0], maxX = points[0];
  double minY = points[1], maxY = points[1];
  for (int64_t i = 1; i < n; i++) {
    double x = points[2 * i], y = points[2 * i + 1];
    minX = x < minX ? x : minX;
    maxX = x > maxX ? x : maxX;
    minY = y < minY ? y : minY;
    maxY = y > maxY ? y : maxY;
  }
  double sx = maxX - minX, sy = maxY - minY;
  double span = sx > sy ? sx : sy;
  double halfSize = span > 0.0 ? span / 2.0 : 1.0;
  double cx = (minX + maxX) / 2.0;
  double cy = (minY + maxY) / 2.0;
  t->rootCx = cx;
  t->rootCy = cy;
  t->rootHalf = halfSize;
  t->maxDepth = depth_limit(halfSize);
  t->root = qt_alloc(t, cx, cy, halfSize, 0, 0, 0);
  for (int64_t i = 0; i < n; i++)
    qt_insert(t, t->root, i);
  /* compact leaf point lists in node order */
  t->leafPts = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  int64_t off = 0;
  for (int64_t i = 0; i < t->nodeCount; i++) {
    oc_qnode *nd = &t->nodes[i];
    nd->ptStart = off;
    if (nd->child[0] < 0) {
      memcpy(t->leafPts + off, t->plist[i], sizeof(int64_t) * (size_t)nd->ptCount);
      off += nd->ptCount;
    } else {
      nd->ptCount = 0;
    }
  }
  for (int64_t i = 0; i < t->nodeCount; i++)
    free(t->plist[i]);
  free(t->plist);
  free(t->pcap);
  t->plist = NULL;
  t->pcap = NULL;
  return t;
}

void oc_quadtree_free(oc_quadtree *t) {
  if (!t)
    return;
  free(t->nodes);
  free(t->leafPts);
  free(t);
}

int64_t oc_quadtree_node_count(const oc_quadtree *t) { return t->nodeCount; }
const oc_qnode *oc_quadtree_nodes(const oc_quadtree *t) { return t->nodes; }
const int64_t *oc_quadtree_leaf_points(const oc_quadtree *t) {
  return t->leafPts;
}
int32_t oc_quadtree_root(const oc_quadtree *t) { return t->root; }

void oc_quadtree_root_params(const oc_quadtree *t, double *cx, double *cy,
                             double *halfSize, int32_t *maxDepth) {
  *cx = t->rootCx;
  *cy = t->rootCy;
  *halfSize = t->rootHalf;
  *maxDepth = t->maxDepth;
}

void oc_quadtree_point_leaves(const oc_quadtree *t, int32_t *depth,
                              uint64_t *keyHi, uint64_t *keyLo) {
  for (int64_t i = 0; i < t->nodeCount; i++) {
    const oc_qnode *nd = &t->nodes[i];
    if (nd->child[0] >= 0)
      continue;
    for (int64_t j = 0; j < nd->ptCount; j++) {
      int64_t p = t->leafPts[nd->ptStart + j];
      depth[p] = nd->depth;
      keyHi[p] = nd->keyHi;
      keyLo[p] = nd->keyLo;
    }
  }
}

void oc_path_key(double px, double py, double cx, double cy, double halfSize,
                 int32_t maxDepth, uint64_t *keyHi, uint64_t *keyLo) {
  uint64_t kh = 0, kl = 0;
  double h = halfSize;
  for (int32_t level = 1; level <= 64 && level <= maxDepth; level++) {
    int east = px >= cx;
    int north = py >= cy;
    int digit = (north ? 0 : 2) + (east ? 1 : 0);
    key_with_digit(kh, kl, level, digit, &kh, &kl);
    double quarter = h / 2.0;
    cx = cx + (east ? quarter : -quarter);
    cy = cy + (north ? quarter : -quarter);
    h = quarter;
  }
  *keyHi = kh;
  *keyLo = kl;
}

/* ---------------------------------------------------------- force step -- */

typedef struct force_ctx {
  const oc_csr *g;
  const double *pos;
  const double *mass;
  const int64_t *degree;
  const oc_force_params *p;
  const oc_quadtree *tree;
  double *disp, *structForce, *attMag, *repMag;
  int64_t *interactions;
  double *termSum; /* optional: per vertex, sum of |term| over every term */
} force_ctx;

static inline double radius_if(const force_ctx *c, int64_t idx) {
  if (!c->p->preventOverlap)
    return 0.0;
  return oc_vertex_radius(c->p, c->degree[idx]);
}

/* Adds one pair term computed into a zeroed temporary t. The sum rounds
 * exactly as accumulating in place (0 + x and a + (-b) are exact rewrites of
 * the in-place updates), and |t| goes to *termAcc. */
static inline void add_term(double *acc, const double *t, double *termAcc) {
  acc[0] += t[0];
  acc[1] += t[1];
  *termAcc += sqrt(t[0] * t[0] + t[1] * t[1]);
}

/* src/embedders/gvizForceEmbedder.c:35-74 (recursion made explicit; the
 * stack pops children in NW, NE, SW, SE order, as the recursion does). */
static int64_t bh_accumulate(const force_ctx *c, int64_t self, double vRadius,
                             const double *vPos, double *acc, double *termAcc) {
  const oc_quadtree *t = c->tree;
  if (t->root < 0)
    return 0;
  int32_t stack[4 * 4096 + 8];
  int sp = 0;
  int64_t count = 0;
  stack[sp++] = t->root;
  while (sp > 0) {
    int32_t id = stack[--sp];
    const oc_qnode *nd = &t->nodes[id];
    if (nd->mass == 0)
      continue;
    if (nd->child[0] < 0) {
      for (int64_t k = 0; k < nd->ptCount; k++) {
        int64_t idx = t->leafPts[nd->ptStart + k];
        if (idx == self)
          continue;
        double t[2] = {0.0, 0.0};
        double *dst = termAcc ? t : acc;
        model_repulsive(c->p->model, vPos, c->pos + 2 * idx, c->mass[self],
                        c->mass[idx], vRadius, radius_if(c, idx),
                        c->p->overlapConstant, c->p->edgeLength, self, idx,
                        dst);
        if (termAcc)
          add_term(acc, t, termAcc);
        count++;
      }
      continue;
    }
    double dx = nd->comX - vPos[0];
    double dy = nd->comY - vPos[1];
    double dist = sqrt(dx * dx + dy * dy);
    double ratio = (2.0 * nd->halfSize) / dist;
    if (ratio < c->p->theta) {
      double com[2] = {nd->comX, nd->comY};
      double t[2] = {0.0, 0.0};
      double *dst = termAcc ? t : acc;
      model_repulsive(c->p->model, vPos, com, c->mass[self], nd->mass, vRadius,
                      0.0, c->p->overlapConstant, c->p->edgeLength, self,
                      c->g->n, dst);
      if (termAcc)
        add_term(acc, t, termAcc);
      count++;
      continue;
    }
    for (int q = 3; q >= 0; q--)
      stack[sp++] = nd->child[q];
  }
  return count;
}

/* src/embedders/gvizForceEmbedder.c:76-126 */
static void force_range(void *vctx, int64_t begin, int64_t end) {
  const force_ctx *c = vctx;
  const oc_force_params *p = c->p;
  for (int64_t i = begin; i < end; i++) {
    double f[2] = {0.0, 0.0};
    double attF[2] = {0.0, 0.0};
    double repF[2] = {0.0, 0.0};
    const double *vPos = c->pos + 2 * i;
    double termSum = 0.0;
    double *termAcc = c->termSum ? &termSum : NULL;
    for (int64_t e = c->g->offsets[i]; e < c->g->offsets[i + 1]; e++) {
      int64_t u = c->g->adj[e];
      if (termAcc) {
        double t[2] = {0.0, 0.0};
        model_attractive(p->model, vPos, c->pos + 2 * u, p->edgeLength, t);
        add_term(attF, t, termAcc);
      } else {
        model_attractive(p->model, vPos, c->pos + 2 * u, p->edgeLength, attF);
      }
    }
    double vRadius = radius_if(c, i);
    int64_t count = 0;
    if (p->barnesHut) {
      count = bh_accumulate(c, i, vRadius, vPos, repF, termAcc);
    } else {
      for (int64_t j = 0; j < c->g->n; j++) {
        if (j == i)
          continue;
        double t[2] = {0.0, 0.0};
        double *dst = termAcc ? t : repF;
        model_repulsive(p->model, vPos, c->pos + 2 * j, c->mass[i], c->mass[j],
                        vRadius, radius_if(c, j), p->overlapConstant,
                        p->edgeLength, i, j, dst);
        if (termAcc)
          add_term(repF, t, termAcc);
      }
      count = c->g->n - 1;
    }
    f[0] += 1.0 * attF[0];
    f[1] += 1.0 * attF[1];
    f[0] += 1.0 * repF[0];
    f[1] += 1.0 * repF[1];
    if (c->attMag)
      c->attMag[i] = sqrt(attF[0] * attF[0] + attF[1] * attF[1]);
    if (c->repMag)
      c->repMag[i] = sqrt(repF[0] * repF[0] + repF[1] * repF[1]);
    if (c->structForce) {
      c->structForce[2 * i] = f[0];
      c->structForce[2 * i + 1] = f[1];
    }
    double gravityMag = p->gravityK * (double)(c->degree[i] + 1);
    if (termAcc) {
      double t[2] = {0.0, 0.0};
      oc_gravity(vPos, gravityMag, t);
      add_term(f, t, termAcc);
      c->termSum[i] = termSum;
    } else {
      oc_gravity(vPos, gravityMag, f);
    }
    if (c->disp) {
      c->disp[2 * i] = f[0];
      c->disp[2 * i + 1] = f[1];
    }
    if (c->interactions)
      c->interactions[i] = count;
  }
}

void oc_force_compute(const oc_csr *g, const double *pos, const double *mass,
                      const int64_t *degree, const oc_force_params *p,
                      const oc_quadtree *tree, oc_pool *pool, double *disp,
                      double *structForce, double *attMag, double *repMag,
                      int64_t *interactions) {
  oc_quadtree *own = NULL;
  if (p->barnesHut && !tree) {
    own = oc_quadtree_build(pos, mass, g->n, p->nodesPerCell);
    tree = own;
  }
  force_ctx c = {g,    pos,         mass,   degree, p,      tree,
                 disp, structForce, attMag, repMag, interactions, NULL};
  oc_pool_for_range(pool, 0, g->n, 1, force_range, &c);
  oc_quadtree_free(own);
}

void oc_force_term_sums(const oc_csr *g, const double *pos, const double *mass,
                        const int64_t *degree, const oc_force_params *p,
                        oc_pool *pool, double *disp, int64_t *interactions,
                        double *termSum) {
  oc_quadtree *tree = NULL;
  if (p->barnesHut)
    tree = oc_quadtree_build(pos, mass, g->n, p->nodesPerCell);
  force_ctx c = {g,    pos,  mass, degree, p,            tree,
                 disp, NULL, NULL, NULL,   interactions, termSum};
  oc_pool_for_range(pool, 0, g->n, 1, force_range, &c);
  oc_quadtree_free(tree);
}

/* src/embedders/gvizForceEmbedder.c:140-185 */
void oc_fa2_update(int64_t n, double jitterTolerance, double totalSwinging,
                   double totalEffectiveTraction, double *globalSpeed,
                   double *speedEfficiency) {
  if (totalEffectiveTraction < OC_NUMERIC_EPS)
    return;
  double nn = (double)n;
  double estimatedOptimalJT = 0.05 * sqrt(nn);
  double minJT = sqrt(estimatedOptimalJT);
  double maxJT = 10.0;
  double jt = jitterTolerance *
              fmax(minJT, fmin(maxJT, estimatedOptimalJT *
                                          totalEffectiveTraction / (nn * nn)));
  if (totalSwinging / totalEffectiveTraction > 2.0) {
    if (*speedEfficiency > OC_EFF_MIN)
      *speedEfficiency *= 0.5;
    jt = fmax(jt, jitterTolerance);
  }
  double targetSpeed = totalSwinging > OC_NUMERIC_EPS
                           ? jt * *speedEfficiency * totalEffectiveTraction /
                                 totalSwinging
                           : *globalSpeed;
  if (totalSwinging > jt * totalEffectiveTraction) {
    if (*speedEfficiency > OC_EFF_MIN)
      *speedEfficiency *= 0.7;
  } else if (*globalSpeed < 1000.0) {
    *speedEfficiency *= 1.3;
  }
  if (*speedEfficiency < OC_EFF_MIN)
    *speedEfficiency = OC_EFF_MIN;
  else if (*speedEfficiency > OC_EFF_MAX)
    *speedEfficiency = OC_EFF_MAX;
  *globalSpeed += fmin(targetSpeed - *globalSpeed, OC_MAX_RISE * *globalSpeed);
}

oc_force_state *oc_force_state_create(const oc_csr *g,
                                      const oc_force_params *p) {
  oc_force_state *s = calloc(1, sizeof(*s));
  int64_t n = g->n;
  s->n = n;
  size_t n1 = (size_t)(n > 0 ? n : 1);
  s->mass = malloc(sizeof(double) * n1);
  s->degree = malloc(sizeof(int64_t) * n1);
  s->disp = calloc(2 * n1, sizeof(double));
  s->structForce = calloc(2 * n1, sizeof(double));
  s->oldStructForce = calloc(2 * n1, sizeof(double));
  s->appliedDisp = calloc(2 * n1, sizeof(double));
  s->attMag = calloc(n1, sizeof(double));
  s->repMag = calloc(n1, sizeof(double));
  s->swinging = calloc(n1, sizeof(double));
  s->traction = calloc(n1, sizeof(double));
  /* src/embedders/gvizForceEmbedder.c:251-259 */
  double sumDegPlusOne = 0.0;
  for (int64_t i = 0; i < n; i++) {
    s->degree[i] = g->offsets[i + 1] - g->offsets[i];
    s->mass[i] = oc_vertex_mass(p->model, s->degree[i]);
    sumDegPlusOne += (double)(s->degree[i] + 1);
  }
  s->meanDegreePlusOne = n ? sumDegPlusOne / (double)n : 0.0;
  oc_force_state_reset(s);
  return s;
}

void oc_force_state_free(oc_force_state *s) {
  if (!s)
    return;
  free(s->mass);
  free(s->degree);
  free(s->disp);
  free(s->structForce);
  free(s->oldStructForce);
  free(s->appliedDisp);
  free(s->attMag);
  free(s->repMag);
  free(s->swinging);
  free(s->traction);
  free(s);
}

/* src/embedders/gvizForceEmbedder.c:416-424 */
void oc_force_state_reset(oc_force_state *s) {
  s->globalSpeed = 1.0;
  s->speedEfficiency = 1.0;
  memset(s->disp, 0, sizeof(double) * 2 * (size_t)s->n);
  memset(s->structForce, 0, sizeof(double) * 2 * (size_t)s->n);
  memset(s->oldStructForce, 0, sizeof(double) * 2 * (size_t)s->n);
  s->iteration = 0;
}

typedef struct swing_ctx {
  oc_force_state *s;
} swing_ctx;

/* src/embedders/gvizForceEmbedder.c:128-138 */
static void swing_range(void *vctx, int64_t begin, int64_t end) {
  oc_force_state *s = ((swing_ctx *)vctx)->s;
  for (int64_t i = begin; i < end; i++) {
    double *f = s->structForce + 2 * i;
    double *old = s->oldStructForce + 2 * i;
    double diff[2] = {f[0] - old[0], f[1] - old[1]};
    double sum[2] = {f[0] + old[0], f[1] + old[1]};
    s->swinging[i] = s->mass[i] * sqrt(diff[0] * diff[0] + diff[1] * diff[1]);
    s->traction[i] =
        s->mass[i] * 0.5 * sqrt(sum[0] * sum[0] + sum[1] * sum[1]);
  }
}

/* src/embedders/gvizForceEmbedder.c:187-197 */
static void apply_range(void *vctx, int64_t begin, int64_t end) {
  oc_force_state *s = ((swing_ctx *)vctx)->s;
  for (int64_t i = begin; i < end; i++) {
    double factor =
        s->globalSpeed / (1.0 + sqrt(s->globalSpeed * s->swinging[i]));
    s->appliedDisp[2 * i] = s->disp[2 * i] * factor;
    s->appliedDisp[2 * i + 1] = s->disp[2 * i + 1] * factor;
  }
}

/* src/embedders/gvizForceEmbedder.c:445-495 */
double oc_force_step(oc_force_state *s, const oc_csr *g,
                     const oc_force_params *p, double *pos, oc_pool *pool,
                     oc_step_stats *stats) {
  int64_t n = s->n;
  memcpy(s->oldStructForce, s->structForce, sizeof(double) * 2 * (size_t)n);
  oc_force_compute(g, pos, s->mass, s->degree, p, NULL, pool, s->disp,
                   s->structForce, s->attMag, s->repMag, NULL);
  swing_ctx sc = {s};
  oc_pool_for_range(pool, 0, n, 1, swing_range, &sc);
  double totalSwinging = 0.0, totalTraction = 0.0;
  for (int64_t i = 0; i < n; i++) {
    totalSwinging += s->swinging[i];
    totalTraction += s->traction[i];
  }
  oc_fa2_update(n, p->jitterTolerance, totalSwinging, totalTraction,
                &s->globalSpeed, &s->speedEfficiency);
  oc_pool_for_range(pool, 0, n, 1, apply_range, &sc);
  double maxDisp = 0.0, sumAtt = 0.0, sumRep = 0.0;
  for (int64_t i = 0; i < n; i++) {
    double *f = s->appliedDisp + 2 * i;
    double applied = sqrt(f[0] * f[0] + f[1] * f[1]);
    if (applied > maxDisp)
      maxDisp = applied;
    sumAtt += s->attMag[i];
    sumRep += s->repMag[i];
    pos[2 * i] += f[0];
    pos[2 * i + 1] += f[1];
  }
  s->iteration++;
  if (stats) {
    stats->maxDisp = maxDisp;
    stats->speed = s->globalSpeed;
    stats->meanAtt = n ? sumAtt / (double)n : 0.0;
    stats->meanRep = n ? sumRep / (double)n : 0.0;
    stats->meanGrav = p->gravityK * s->meanDegreePlusOne;
    stats->speedEfficiency = s->speedEfficiency;
    stats->totalSwinging = totalSwinging;
    stats->totalTraction = totalTraction;
  }
  return maxDisp;
}

/* ---------------------------------------------------------------- Tutte -- */

/* src/embedders/gvizTutteEmbedder.c:249-324 (dimension 2). */
double oc_tutte_step(const oc_csr *g, const uint8_t *isBoundary, double *pos,
                     double *scratch, double relaxationRate, double dt,
                     int useGaussSeidel) {
  int64_t n = g->n;
  double alpha = relaxationRate * dt;
  if (alpha <= 0.0)
    return 0.0;
  if (alpha > 1.0)
    alpha = 1.0;
  if (!useGaussSeidel)
    for (int64_t u = 0; u < n; u++)
      if (!isBoundary[u]) {
        scratch[2 * u] = pos[2 * u];
        scratch[2 * u + 1] = pos[2 * u + 1];
      }
  double maxDelta = 0.0;
  for (int64_t u = 0; u < n; u++) {
    if (isBoundary[u])
      continue;
    int64_t b = g->offsets[u], e = g->offsets[u + 1];
    if (e == b)
      continue;
    double bary[2] = {0.0, 0.0};
    size_t count = 0;
    for (int64_t k = b; k < e; k++) {
      int64_t v = g->adj[k];
      const double *vp =
          (!useGaussSeidel && !isBoundary[v]) ? scratch + 2 * v : pos + 2 * v;
      bary[0] += vp[0];
      bary[1] += vp[1];
      count++;
    }
    bary[0] /= (double)count;
    bary[1] /= (double)count;
    double *old = pos + 2 * u;
    double np0 = old[0] + alpha * (bary[0] - old[0]);
    double np1 = old[1] + alpha * (bary[1] - old[1]);
    double d0 = np0 - old[0];
    double d1 = np1 - old[1];
    double delta = 0.0;
    delta += d0 * d0;
    delta += d1 * d1;
    old[0] = np0;
    old[1] = np1;
    double d = sqrt(delta);
    if (d > maxDelta)
      maxDelta = d;
  }
  return maxDelta;
}

/* src/embedders/gvizTutteEmbedder.c:230-247 */
void oc_tutte_polygon(int64_t count, double radius, double *out) {
  for (int64_t k = 0; k < count; k++) {
    double angle = 2.0 * M_PI * (double)k / (double)count;
    out[2 * k] = radius * cos(angle);
    out[2 * k + 1] = radius * sin(angle);
  }
}

/* -------------------------------------------------------------- metrics -- */
/* The reference defines no layout-quality metric; these are this project's
 * definitions, pinned by closed-form tests.
 *   stress(x) = min_s sum_{pairs} d^-2 (s*|x_i-x_j| - L*d)^2 / #pairs
 *             = 1 - A^2/(B*C) with A = sum D/d', B = sum D^2/d'^2, C = #pairs,
 *   where D = |x_i - x_j| and d' = L * hops. */
typedef struct stress_ctx {
  const oc_csr *g;
  const double *pos;
  double L;
  const int64_t *sources;
  int64_t nsources;
  double *A, *B, *C; /* per-source partials */
} stress_ctx;

static void stress_range(void *vctx, int64_t begin, int64_t end) {
  stress_ctx *c = vctx;
  int64_t n = c->g->n;
  int64_t *dist = malloc(sizeof(int64_t) * (size_t)n);
  int64_t *queue = malloc(sizeof(int64_t) * (size_t)n);
  for (int64_t si = begin; si < end; si++) {
    int64_t s = c->sources ? c->sources[si] : si;
    for (int64_t i = 0; i < n; i++)
      dist[i] = -1;
    int64_t qh = 0, qt = 0;
    dist[s] = 0;
    queue[qt++] = s;
    while (qh < qt) {
      int64_t u = queue[qh++];
      for (int64_t k = c->g->offsets[u]; k < c->g->offsets[u + 1]; k++) {
        int64_t v = c->g->adj[k];
        if (dist[v] < 0) {
          dist[v] = dist[u] + 1;
          queue[qt++] = v;
        }
      }
    }
    double A = 0.0, B = 0.0, C = 0.0;
    for (int64_t j = 0; j < n; j++) {
      if (j == s || dist[j] <= 0)
        continue;
      double dx = c->pos[2 * j] - c->pos[2 * s];
      double dy = c->pos[2 * j + 1] - c->pos[2 * s + 1];
      double D = sqrt(dx * dx + dy * dy);
      double d = c->L * (double)dist[j];
      A += D / d;
      B += (D * D) / (d * d);
      C += 1.0;
    }
    c->A[si] = A;
    c->B[si] = B;
    c->C[si] = C;
  }
  free(dist);
  free(queue);
}

double oc_stress(const oc_csr *g, const double *pos, double edgeLength,
                 const int64_t *sources, int64_t nsources, oc_pool *pool) {
  int64_t ns = sources ? nsources : g->n;
  if (ns <= 0)
    return 0.0;
  stress_ctx c = {g, pos, edgeLength, sources, ns, NULL, NULL, NULL};
  c.A = calloc((size_t)ns, sizeof(double));
  c.B = calloc((size_t)ns, sizeof(double));
  c.C = calloc((size_t)ns, sizeof(double));
  oc_pool_for_range(pool, 0, ns, 1, stress_range, &c);
  double A = 0.0, B = 0.0, C = 0.0;
  for (int64_t i = 0; i < ns; i++) {
    A += c.A[i];
    B += c.B[i];
    C += c.C[i];
  }
  free(c.A);
  free(c.B);
  free(c.C);
  if (C == 0.0 || B == 0.0)
    return 0.0;
  return 1.0 - (A * A) / (B * C);
}

double oc_edge_energy(const oc_csr *g, const double *pos) {
  double sum = 0.0;
  int64_t m = 0;
  for (int64_t u = 0; u < g->n; u++)
    for (int64_t k = g->offsets[u]; k < g->offsets[u + 1]; k++) {
      int64_t v = g->adj[k];
      if (v <= u)
        continue;
      double dx = pos[2 * v] - pos[2 * u], dy = pos[2 * v + 1] - pos[2 * u + 1];
      sum += sqrt(dx * dx + dy * dy);
      m++;
    }
  if (m == 0 || sum == 0.0)
    return 0.0;
  double mean = sum / (double)m;
  double e = 0.0;
  for (int64_t u = 0; u < g->n; u++)
    for (int64_t k = g->offsets[u]; k < g->offsets[u + 1]; k++) {
      int64_t v = g->adj[k];
      if (v <= u)
        continue;
      double dx = pos[2 * v] - pos[2 * u], dy = pos[2 * v + 1] - pos[2 * u + 1];
      double r = sqrt(dx * dx + dy * dy) / mean - 1.0;
      e += r * r;
    }
  return e / (double)m;
}
