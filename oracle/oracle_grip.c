/*
 * oracle_grip.c -- FP64 CPU restatement of the reference's GRIP pieces:
 * BFS k-nearest search (per-vertex and batched), the MIS filtration and the
 * refinement round.
 *
 * TEST INFRASTRUCTURE ONLY (see gviz_oracle.h). Graph = CSR over parent ids
 * restricted to the subgraph (inactive vertices have no adjacency). Citations
 * are file:line in the reference tree.
 */
#include "gviz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OC_NUMERIC_EPS 1e-7
#define GRIP_EDGE_LENGTH 10.0   /* src/embedders/gvizGRIPEmbedder.c:31 */
#define GRIP_R 0.15             /* :45 */
#define GRIP_S 3.0              /* :46 */
#define GRIP_FR_SCALE 0.05      /* src/embedders/gvizForceDirected.c:4 */
#define KNN_BATCH_VISIBLE_MAX 256
#define KNN_BATCH_MIN_AVG_DEGREE 32.0
#define KNN_BATCH_TARGET_RATIO 8

/* ------------------------------------------------------------------ kNN -- */

/* src/algorithms/search/gvizKNearest.c:244-306 */
int64_t oc_knn_bfs(const oc_csr *g, int64_t source, int64_t k,
                   const uint8_t *filter, oc_found *out) {
  if (k == 0)
    return 0;
  int64_t n = g->n;
  uint8_t *seen = calloc((size_t)n, 1);
  oc_found *queue = malloc(sizeof(oc_found) * (size_t)n);
  int64_t qh = 0, qt = 0, count = 0;
  seen[source] = 1;
  queue[qt++] = (oc_found){source, 0};
  while (qh < qt) {
    oc_found cur = queue[qh++];
    for (int64_t e = g->offsets[cur.v]; e < g->offsets[cur.v + 1]; e++) {
      int64_t nb = g->adj[e];
      if (seen[nb])
        continue;
      seen[nb] = 1;
      oc_found next = {nb, cur.dist + 1};
      if (!filter || filter[nb]) {
        out[count++] = next;
        if (count >= k)
          goto done;
      }
      queue[qt++] = next;
    }
  }
done:
  free(seen);
  free(queue);
  return count;
}

/* src/algorithms/search/gvizKNearest.c:62-89 */
static void knn_insert_sorted(oc_found *out, int64_t *count, int64_t k,
                              int64_t v, int64_t dist) {
  oc_found cand = {v, dist};
  int64_t n = *count;
  if (n < k) {
    out[n] = cand;
    (*count)++;
    for (int64_t i = n; i > 0 && out[i].dist < out[i - 1].dist; i--) {
      oc_found t = out[i];
      out[i] = out[i - 1];
      out[i - 1] = t;
    }
    return;
  }
  if (dist >= out[k - 1].dist)
    return;
  out[k - 1] = cand;
  for (int64_t i = k - 1; i > 0 && out[i].dist < out[i - 1].dist; i--) {
    oc_found t = out[i];
    out[i] = out[i - 1];
    out[i - 1] = t;
  }
}

/* src/algorithms/search/gvizKNearest.c:149-166 */
int oc_knn_prefer_batch(const oc_csr *g, int64_t activeCount, int64_t edgeCount,
                        int64_t visibleCount, int64_t targetCount) {
  (void)g;
  if (visibleCount == 0 || visibleCount >= targetCount ||
      visibleCount > KNN_BATCH_VISIBLE_MAX)
    return 0;
  if (activeCount == 0)
    return 0;
  double avgDeg = (2.0 * (double)edgeCount) / (double)activeCount;
  if (avgDeg > KNN_BATCH_MIN_AVG_DEGREE)
    return 1;
  return targetCount >= KNN_BATCH_TARGET_RATIO * visibleCount;
}

/* src/algorithms/search/gvizKNearest.c:91-138,168-210 */
int oc_knn_batch(const oc_csr *g, const uint8_t *visible, int64_t k,
                 const int64_t *targets, int64_t targetCount, oc_found *out,
                 int64_t *counts) {
  int64_t n = g->n;
  int64_t *targetMap = malloc(sizeof(int64_t) * (size_t)n);
  for (int64_t i = 0; i < n; i++)
    targetMap[i] = -1;
  for (int64_t i = 0; i < targetCount; i++) {
    targetMap[targets[i]] = i;
    counts[i] = 0;
  }
  uint8_t *seen = malloc((size_t)n);
  oc_found *queue = malloc(sizeof(oc_found) * (size_t)n);
  for (int64_t seed = 0; seed < n; seed++) {
    if (!visible[seed])
      continue;
    memset(seen, 0, (size_t)n);
    int64_t qh = 0, qt = 0;
    seen[seed] = 1;
    queue[qt++] = (oc_found){seed, 0};
    while (qh < qt) {
      oc_found cur = queue[qh++];
      for (int64_t e = g->offsets[cur.v]; e < g->offsets[cur.v + 1]; e++) {
        int64_t nb = g->adj[e];
        if (seen[nb])
          continue;
        seen[nb] = 1;
        int64_t dist = cur.dist + 1;
        int64_t ti = targetMap[nb];
        if (ti >= 0)
          knn_insert_sorted(out + ti * k, &counts[ti], k, seed, dist);
        queue[qt++] = (oc_found){nb, dist};
      }
    }
    int full = 1;
    for (int64_t i = 0; i < targetCount && full; i++)
      if (counts[i] < k)
        full = 0;
    if (full)
      break;
  }
  free(targetMap);
  free(seen);
  free(queue);
  return 0;
}

/* ------------------------------------------------------ MIS filtration -- */

typedef struct mis_ctx {
  const oc_csr *g;
  const uint8_t *active;
  int64_t n;
  int64_t activeCount;
  int64_t *filtration;
  int64_t *borders;
  int64_t nborders;
  /* BFS scratch */
  int64_t *stamp;
  int64_t epoch;
  oc_found *queue;
} mis_ctx;

static int is_active(const mis_ctx *c, int64_t v) {
  return c->active ? c->active[v] != 0 : 1;
}

/* src/embedders/gvizGRIPEmbedder.c:450-483 */
static void mis_first(mis_ctx *c, uint8_t *out) {
  uint8_t *states = calloc((size_t)c->n, 1);
  int64_t cur = c->activeCount - 1;
  for (int64_t i = 0; i < c->n; i++) {
    if (!is_active(c, i))
      continue;
    if (!states[i]) {
      out[i] = 1;
      for (int64_t e = c->g->offsets[i]; e < c->g->offsets[i + 1]; e++) {
        int64_t nb = c->g->adj[e];
        if (!states[nb]) {
          states[nb] = 1;
          c->filtration[cur--] = nb;
        }
      }
    }
  }
  int64_t border = 0;
  for (int64_t i = 0; i < c->n; i++)
    border += out[i] != 0;
  c->borders[c->nborders++] = border;
  free(states);
}

/* src/embedders/gvizGRIPEmbedder.c:599-670 */
static void within_radius(mis_ctx *c, int64_t source, int64_t maxDepth,
                          uint8_t *out) {
  int64_t epoch = ++c->epoch;
  int64_t qh = 0, qt = 0;
  c->stamp[source] = epoch;
  c->queue[qt++] = (oc_found){source, 0};
  while (qh < qt) {
    oc_found nd = c->queue[qh++];
    if (maxDepth && nd.dist >= maxDepth)
      continue;
    for (int64_t e = c->g->offsets[nd.v]; e < c->g->offsets[nd.v + 1]; e++) {
      int64_t nb = c->g->adj[e];
      if (c->stamp[nb] == epoch)
        continue;
      c->stamp[nb] = epoch;
      int64_t next = nd.dist + 1;
      if (next <= maxDepth)
        out[nb] = 1;
      c->queue[qt++] = (oc_found){nb, next};
    }
  }
}

/* src/embedders/gvizGRIPEmbedder.c:672-739 */
static int mis_iter(mis_ctx *c, int64_t i, uint8_t *vertices, int64_t dim) {
  uint8_t *newVertices = calloc((size_t)c->n, 1);
  uint8_t *newStates = calloc((size_t)c->n, 1);
  int64_t radius = (int64_t)1 << (i - 1);
  for (int64_t v = 0; v < c->n; v++) {
    if (!vertices[v])
      continue;
    if (newStates[v])
      continue;
    newVertices[v] = 1;
    newStates[v] = 1;
    within_radius(c, v, radius, newStates);
  }
  int64_t ptr = c->borders[i - 1] - 1;
  for (int64_t v = 0; v < c->n; v++)
    if (vertices[v] && !newVertices[v])
      c->filtration[ptr--] = v;
  int64_t border = 0;
  for (int64_t v = 0; v < c->n; v++)
    border += newVertices[v] != 0;
  c->borders[c->nborders++] = border;
  memcpy(vertices, newVertices, (size_t)c->n);
  free(newVertices);
  free(newStates);
  return border > dim + 1;
}

/* src/embedders/gvizGRIPEmbedder.c:507-565 */
static int64_t pick_far(mis_ctx *c, int64_t finalEnd, int64_t candBegin,
                        int64_t candEnd, int64_t maxDepth) {
  int64_t epoch = ++c->epoch;
  int64_t *depth = malloc(sizeof(int64_t) * (size_t)c->n);
  int64_t qh = 0, qt = 0;
  for (int64_t i = 0; i < finalEnd; i++) {
    int64_t v = c->filtration[i];
    c->stamp[v] = epoch;
    depth[v] = 0;
    c->queue[qt++] = (oc_found){v, 0};
  }
  while (qh < qt) {
    oc_found nd = c->queue[qh++];
    if (nd.dist >= maxDepth)
      continue;
    for (int64_t e = c->g->offsets[nd.v]; e < c->g->offsets[nd.v + 1]; e++) {
      int64_t nb = c->g->adj[e];
      if (c->stamp[nb] == epoch)
        continue;
      c->stamp[nb] = epoch;
      depth[nb] = nd.dist + 1;
      c->queue[qt++] = (oc_found){nb, nd.dist + 1};
    }
  }
  int64_t bestIdx = candBegin, bestScore = 0, beyond = maxDepth + 1;
  for (int64_t j = candBegin; j < candEnd; j++) {
    int64_t v = c->filtration[j];
    int64_t score = c->stamp[v] == epoch ? depth[v] : beyond;
    if (score > bestScore || (score == bestScore && j > bestIdx)) {
      bestScore = score;
      bestIdx = j;
    }
  }
  free(depth);
  return bestIdx;
}

/* src/embedders/gvizGRIPEmbedder.c:576-597 */
static int migrate_one(mis_ctx *c, int64_t count) {
  int64_t finalEnd = c->borders[count - 1];
  int64_t srcLayer = count - 1, candEnd;
  do {
    if (srcLayer == 0)
      return -1;
    srcLayer--;
    candEnd = c->borders[srcLayer];
  } while (candEnd <= finalEnd);
  int64_t pick = pick_far(c, finalEnd, finalEnd, candEnd, 64);
  int64_t t = c->filtration[pick];
  c->filtration[pick] = c->filtration[finalEnd];
  c->filtration[finalEnd] = t;
  c->borders[count - 1]++;
  return 0;
}

static void mis_init(mis_ctx *c, const oc_csr *g, const uint8_t *active,
                     int64_t *filtration, int64_t *borders) {
  memset(c, 0, sizeof(*c));
  c->g = g;
  c->active = active;
  c->n = g->n;
  for (int64_t v = 0; v < g->n; v++)
    c->activeCount += active ? active[v] != 0 : 1;
  c->filtration = filtration;
  c->borders = borders;
  c->borders[c->nborders++] = c->activeCount;
  c->stamp = calloc((size_t)g->n, sizeof(int64_t));
  c->queue = malloc(sizeof(oc_found) * (size_t)(g->n + 1));
}

static void mis_free(mis_ctx *c) {
  free(c->stamp);
  free(c->queue);
}

int64_t oc_mis_first(const oc_csr *g, const uint8_t *active,
                     int64_t *filtration, uint8_t *inMis) {
  int64_t borders[4];
  mis_ctx c;
  mis_init(&c, g, active, filtration, borders);
  memset(inMis, 0, (size_t)g->n);
  mis_first(&c, inMis);
  mis_free(&c);
  return borders[1];
}

/* src/embedders/gvizGRIPEmbedder.c:741-777 */
int64_t oc_mis_filtration(const oc_csr *g, const uint8_t *active,
                          int64_t dim, int64_t *filtration, int64_t *borders) {
  mis_ctx c;
  mis_init(&c, g, active, filtration, borders);
  uint8_t *cur = calloc((size_t)g->n, 1);
  mis_first(&c, cur);
  int64_t i = 2;
  /* A subgraph with more than dim + 1 components never coarsens to dim + 1
   * vertices (the reference keeps adding layers forever); stop at 126. */
  while (mis_iter(&c, i, cur, dim)) {
    if (++i > 126) {
      free(cur);
      mis_free(&c);
      return -1;
    }
  }
  int64_t k = 0;
  for (int64_t v = 0; v < g->n; v++)
    if (cur[v])
      filtration[k++] = v;
  while (borders[i] < dim + 1)
    if (migrate_one(&c, i + 1) < 0)
      break;
  free(cur);
  mis_free(&c);
  return i + 1;
}

/* ------------------------------------------------------ refinement round -- */

/* src/embedders/gvizGRIPEmbedder.c:870-889 */
static void update_local_temp(oc_grip_vertex *d) {
  double nrm = sqrt(d->dispX * d->dispX + d->dispY * d->dispY);
  double oldNrm = sqrt(d->oldDispX * d->oldDispX + d->oldDispY * d->oldDispY);
  if (fabs(nrm) < OC_NUMERIC_EPS || fabs(oldNrm) < OC_NUMERIC_EPS)
    return;
  double cs = (d->dispX * d->oldDispX + d->dispY * d->oldDispY) / (nrm * oldNrm);
  if (cs > 0 && d->oldCos > 0)
    d->heat *= (1 + cs * GRIP_R * GRIP_S);
  else
    d->heat *= (1 + cs * GRIP_R);
  d->oldCos = cs;
}

/* src/embedders/gvizGRIPEmbedder.c:892-934,1071-1099 */
static void refine_vertex(const oc_csr *g, int64_t v, int64_t layer,
                          const oc_found *knn, int64_t kcap, oc_grip_vertex *d,
                          const double *pos) {
  d->oldDispX = d->dispX;
  d->oldDispY = d->dispY;
  double f[2] = {0.0, 0.0};
  const double *vp = pos + 2 * v;
  const oc_found *list = knn + v * kcap;
  if (layer > 0) {
    for (int64_t i = 0; i < d->knnCount; i++)
      oc_kk(vp, pos + 2 * list[i].v, (double)list[i].dist * GRIP_EDGE_LENGTH, f);
  } else {
    for (int64_t e = g->offsets[v]; e < g->offsets[v + 1]; e++)
      oc_grip_fr_rep(vp, pos + 2 * g->adj[e], GRIP_EDGE_LENGTH, f);
    for (int64_t i = 0; i < d->knnCount; i++)
      oc_grip_fr_att(vp, pos + 2 * list[i].v, GRIP_EDGE_LENGTH, GRIP_FR_SCALE,
                     f);
  }
  d->dispX = f[0];
  d->dispY = f[1];
  if (!d->dispCalculated) {
    d->dispCalculated = 1;
    d->heat = GRIP_EDGE_LENGTH / 6.0;
  } else {
    update_local_temp(d);
  }
  d->heat = fmin(d->heat, GRIP_EDGE_LENGTH * 10);
  d->heat = fmax(d->heat, GRIP_EDGE_LENGTH * 1e-4);
  double nrm = sqrt(d->dispX * d->dispX + d->dispY * d->dispY);
  d->lastForceMag = nrm;
  if (nrm > OC_NUMERIC_EPS) {
    double s = d->heat / nrm;
    d->dispX *= s;
    d->dispY *= s;
  }
}

/* src/embedders/gvizGRIPEmbedder.c:1289-1333 with removeNetTranslation
 * (:1106-1121) and the 2-D branch of removeNetRotation (:1192-1261). */
void oc_grip_refine_round(const oc_csr *g, const int64_t *filtration,
                          int64_t active, int64_t layer, const oc_found *knn,
                          int64_t kcap, oc_grip_vertex *dec, double *pos,
                          oc_grip_round_stats *stats) {
  for (int64_t i = 0; i < active; i++) {
    int64_t v = filtration[i];
    refine_vertex(g, v, layer, knn, kcap, &dec[v], pos);
  }
  if (active > 0) {
    double mean[2] = {0.0, 0.0};
    for (int64_t i = 0; i < active; i++) {
      oc_grip_vertex *d = &dec[filtration[i]];
      mean[0] += 1.0 * d->dispX;
      mean[1] += 1.0 * d->dispY;
    }
    double inv = 1.0 / (double)active;
    mean[0] *= inv;
    mean[1] *= inv;
    for (int64_t i = 0; i < active; i++) {
      oc_grip_vertex *d = &dec[filtration[i]];
      d->dispX += -1.0 * mean[0];
      d->dispY += -1.0 * mean[1];
    }
  }
  if (active >= 3) {
    double com[2] = {0.0, 0.0};
    for (int64_t i = 0; i < active; i++) {
      const double *p = pos + 2 * filtration[i];
      com[0] += p[0];
      com[1] += p[1];
    }
    com[0] /= (double)active;
    com[1] /= (double)active;
    double L2 = 0.0, I22 = 0.0;
    for (int64_t i = 0; i < active; i++) {
      int64_t v = filtration[i];
      const double *p = pos + 2 * v;
      oc_grip_vertex *d = &dec[v];
      double r0 = p[0] - com[0], r1 = p[1] - com[1], r2v = 0.0;
      L2 += r0 * d->dispY - r1 * d->dispX;
      double r2 = r0 * r0 + r1 * r1 + r2v * r2v;
      I22 += r2 - r2v * r2v;
    }
    if (!(fabs(I22) < OC_NUMERIC_EPS)) {
      double omega = L2 / I22;
      for (int64_t i = 0; i < active; i++) {
        int64_t v = filtration[i];
        const double *p = pos + 2 * v;
        oc_grip_vertex *d = &dec[v];
        double r0 = p[0] - com[0], r1 = p[1] - com[1], r2v = 0.0;
        d->dispX -= 0.0 * r2v - omega * r1;
        d->dispY -= omega * r0 - 0.0 * r2v;
      }
    }
  }
  double maxDisp = 0.0, sumDisp = 0.0, sumHeat = 0.0, sumForce = 0.0;
  for (int64_t i = 0; i < active; i++) {
    int64_t v = filtration[i];
    oc_grip_vertex *d = &dec[v];
    double nrm = sqrt(d->dispX * d->dispX + d->dispY * d->dispY);
    if (nrm > maxDisp)
      maxDisp = nrm;
    sumDisp += nrm;
    sumHeat += d->heat;
    sumForce += d->lastForceMag;
    pos[2 * v] += d->dispX;
    pos[2 * v + 1] += d->dispY;
  }
  if (stats) {
    stats->maxDisp = maxDisp;
    stats->meanDisp = active ? sumDisp / (double)active : 0.0;
    stats->meanForce = active ? sumForce / (double)active : 0.0;
    stats->meanHeat = active ? sumHeat / (double)active : 0.0;
  }
}

/* ------------------------------------ refinement in 2, 3 or 4 dimensions -- */

/* gvizVecDot (include/core/gvizVec.h:42-57): products summed left to right */
static double nd_dot(int64_t dim, const double *a, const double *b) {
  double s = a[0] * b[0];
  for (int64_t k = 1; k < dim; k++)
    s += a[k] * b[k];
  return s;
}

/* updateLocalTemp (gvizGRIPEmbedder.c:870-889) */
static void update_local_temp_nd(int64_t dim, oc_grip_vertex_nd *d) {
  double nrm = sqrt(nd_dot(dim, d->disp, d->disp));
  double oldNrm = sqrt(nd_dot(dim, d->oldDisp, d->oldDisp));
  if (fabs(nrm) < OC_NUMERIC_EPS || fabs(oldNrm) < OC_NUMERIC_EPS)
    return;
  double cs = nd_dot(dim, d->disp, d->oldDisp) / (nrm * oldNrm);
  if (cs > 0 && d->oldCos > 0)
    d->heat *= (1 + cs * GRIP_R * GRIP_S);
  else
    d->heat *= (1 + cs * GRIP_R);
  d->oldCos = cs;
}

/* refinementPass1Range and calculateSpringForces (gvizGRIPEmbedder.c:
 * 1071-1099, 892-934) with the force cases of gvizVecAccKKForce,
 * gvizVecAccGRIPFRRepForce and gvizVecAccGRIPFRAttForce for n = 2..4
 * (include/core/gvizVec.h:138-290). */
static void refine_vertex_nd(const oc_csr *g, int64_t v, int64_t layer,
                             const oc_found *knn, int64_t kcap, int64_t dim,
                             oc_grip_vertex_nd *d, const double *pos) {
  memcpy(d->oldDisp, d->disp, sizeof(double) * (size_t)dim);
  double f[4] = {0.0, 0.0, 0.0, 0.0}, diff[4];
  const double *vp = pos + dim * v;
  const oc_found *list = knn + v * kcap;
  const double eSq = GRIP_EDGE_LENGTH * GRIP_EDGE_LENGTH;
  if (layer > 0) {
    for (int64_t i = 0; i < d->knnCount; i++) {
      double L = (double)list[i].dist * GRIP_EDGE_LENGTH;
      if (L < 1e-9)
        continue;
      const double *up = pos + dim * list[i].v;
      for (int64_t k = 0; k < dim; k++)
        diff[k] = up[k] - vp[k];
      double s = sqrt(nd_dot(dim, diff, diff)) / L - 1.0;
      for (int64_t k = 0; k < dim; k++)
        f[k] += s * diff[k];
    }
  } else {
    for (int64_t e = g->offsets[v]; e < g->offsets[v + 1]; e++) {
      const double *up = pos + dim * g->adj[e];
      for (int64_t k = 0; k < dim; k++)
        diff[k] = up[k] - vp[k];
      double dsq = nd_dot(dim, diff, diff);
      if (dsq < 1e-18)
        continue;
      double s = dsq / eSq;
      for (int64_t k = 0; k < dim; k++)
        f[k] += s * diff[k];
    }
    for (int64_t i = 0; i < d->knnCount; i++) {
      const double *up = pos + dim * list[i].v;
      for (int64_t k = 0; k < dim; k++)
        diff[k] = vp[k] - up[k];
      double dsq = nd_dot(dim, diff, diff);
      if (dsq < 1e-18)
        continue;
      double s = GRIP_FR_SCALE * eSq / dsq;
      for (int64_t k = 0; k < dim; k++)
        f[k] += s * diff[k];
    }
  }
  memcpy(d->disp, f, sizeof(double) * (size_t)dim);
  if (!d->dispCalculated) {
    d->dispCalculated = 1;
    d->heat = GRIP_EDGE_LENGTH / 6.0;
  } else {
    update_local_temp_nd(dim, d);
  }
  d->heat = fmin(d->heat, GRIP_EDGE_LENGTH * 10);
  d->heat = fmax(d->heat, GRIP_EDGE_LENGTH * 1e-4);
  double nrm = sqrt(nd_dot(dim, d->disp, d->disp));
  d->lastForceMag = nrm;
  if (nrm > OC_NUMERIC_EPS) {
    double s = d->heat / nrm;
    for (int64_t k = 0; k < dim; k++)
      d->disp[k] *= s;
  }
}

/* solve3x3 (gvizGRIPEmbedder.c:1122-1152): elimination with the row of
 * largest magnitude as pivot, kept through an index table. */
static int solve3(double a[3][3], const double *b, double *x) {
  int idx[3] = {0, 1, 2};
  double rhs[3] = {b[0], b[1], b[2]};
  for (int col = 0; col < 3; col++) {
    int piv = col;
    for (int row = col + 1; row < 3; row++)
      if (fabs(a[idx[row]][col]) > fabs(a[idx[piv]][col]))
        piv = row;
    if (fabs(a[idx[piv]][col]) < OC_NUMERIC_EPS)
      return -1;
    int t = idx[col];
    idx[col] = idx[piv];
    idx[piv] = t;
    for (int row = col + 1; row < 3; row++) {
      double fct = a[idx[row]][col] / a[idx[col]][col];
      for (int j = col; j < 3; j++)
        a[idx[row]][j] -= fct * a[idx[col]][j];
      rhs[idx[row]] -= fct * rhs[idx[col]];
    }
  }
  for (int col = 2; col >= 0; col--) {
    double s = rhs[idx[col]];
    for (int j = col + 1; j < 3; j++)
      s -= a[idx[col]][j] * x[j];
    x[col] = s / a[idx[col]][col];
  }
  return 0;
}

/* removeNetRotationPlane (gvizGRIPEmbedder.c:1156-1185) */
static void remove_plane_rotation(const int64_t *filtration, int64_t active, int64_t dim,
                                  int a, int b, const double *com,
                                  oc_grip_vertex_nd *dec, const double *pos) {
  double L = 0.0, inertia = 0.0;
  for (int64_t i = 0; i < active; i++) {
    const double *p = pos + dim * filtration[i];
    const double *d = dec[filtration[i]].disp;
    double ra = p[a] - com[a], rb = p[b] - com[b];
    L += ra * d[b] - rb * d[a];
    inertia += ra * ra + rb * rb;
  }
  if (inertia < OC_NUMERIC_EPS)
    return;
  double omega = L / inertia;
  for (int64_t i = 0; i < active; i++) {
    const double *p = pos + dim * filtration[i];
    double *d = dec[filtration[i]].disp;
    double ra = p[a] - com[a], rb = p[b] - com[b];
    d[a] += omega * rb;
    d[b] -= omega * ra;
  }
}

/* gvizGRIPEmbedderRefineRound (gvizGRIPEmbedder.c:1289-1333) with
 * removeNetTranslation (:1106-1121) and removeNetRotation (:1192-1261). */
void oc_grip_refine_round_nd(const oc_csr *g, const int64_t *filtration,
                             int64_t active, int64_t layer, const oc_found *knn,
                             int64_t kcap, int64_t dim, oc_grip_vertex_nd *dec,
                             double *pos, oc_grip_round_stats *stats) {
  for (int64_t i = 0; i < active; i++)
    refine_vertex_nd(g, filtration[i], layer, knn, kcap, dim, &dec[filtration[i]], pos);
  if (active > 0) {
    double mean[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = 0; i < active; i++)
      for (int64_t k = 0; k < dim; k++)
        mean[k] += 1.0 * dec[filtration[i]].disp[k];
    double inv = 1.0 / (double)active;
    for (int64_t k = 0; k < dim; k++)
      mean[k] *= inv;
    for (int64_t i = 0; i < active; i++)
      for (int64_t k = 0; k < dim; k++)
        dec[filtration[i]].disp[k] += -1.0 * mean[k];
  }
  if (active >= 3) {
    double com[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = 0; i < active; i++)
      for (int64_t k = 0; k < dim; k++)
        com[k] += pos[dim * filtration[i] + k];
    for (int64_t k = 0; k < dim; k++)
      com[k] /= (double)active;
    if (dim == 4) {
      static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
      for (int q = 0; q < 6; q++)
        remove_plane_rotation(filtration, active, dim, pa[q], pb[q], com, dec, pos);
    } else {
      /* 2-D and 3-D: omega = I^-1 L about the centre, 2-D through I22 */
      const int three = dim == 3;
      double L[3] = {0, 0, 0}, I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int64_t i = 0; i < active; i++) {
        const double *p = pos + dim * filtration[i];
        const double *d = dec[filtration[i]].disp;
        double r[3] = {p[0] - com[0], p[1] - com[1], three ? p[2] - com[2] : 0.0};
        double dz = three ? d[2] : 0.0;
        L[0] += r[1] * dz - r[2] * d[1];
        L[1] += r[2] * d[0] - r[0] * dz;
        L[2] += r[0] * d[1] - r[1] * d[0];
        double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++)
            I[a][b] += (a == b ? r2 : 0.0) - r[a] * r[b];
      }
      double w[3] = {0.0, 0.0, 0.0};
      int ok;
      if (!three) {
        ok = !(fabs(I[2][2]) < OC_NUMERIC_EPS);
        if (ok)
          w[2] = L[2] / I[2][2];
      } else {
        ok = solve3(I, L, w) == 0;
      }
      if (ok)
        for (int64_t i = 0; i < active; i++) {
          const double *p = pos + dim * filtration[i];
          double *d = dec[filtration[i]].disp;
          double r[3] = {p[0] - com[0], p[1] - com[1], three ? p[2] - com[2] : 0.0};
          d[0] -= w[1] * r[2] - w[2] * r[1];
          d[1] -= w[2] * r[0] - w[0] * r[2];
          if (three)
            d[2] -= w[0] * r[1] - w[1] * r[0];
        }
    }
  }
  double maxDisp = 0.0, sumDisp = 0.0, sumHeat = 0.0, sumForce = 0.0;
  for (int64_t i = 0; i < active; i++) {
    int64_t v = filtration[i];
    oc_grip_vertex_nd *d = &dec[v];
    double nrm = sqrt(nd_dot(dim, d->disp, d->disp));
    if (nrm > maxDisp)
      maxDisp = nrm;
    sumDisp += nrm;
    sumHeat += d->heat;
    sumForce += d->lastForceMag;
    for (int64_t k = 0; k < dim; k++)
      pos[dim * v + k] += d->disp[k];
  }
  if (stats) {
    stats->maxDisp = maxDisp;
    stats->meanDisp = active ? sumDisp / (double)active : 0.0;
    stats->meanForce = active ? sumForce / (double)active : 0.0;
    stats->meanHeat = active ? sumHeat / (double)active : 0.0;
  }
}

/* src/embedders/gvizGRIPEmbedder.c:335-420 */
void oc_regular_simplex(int64_t n, double sideLength, double *out) {
  memset(out, 0, sizeof(double) * (size_t)((n + 1) * n));
  for (int64_t k = 0; k <= n; k++) {
    double *vk = out + k * n;
    if (k == 0) {
      vk[0] = 1.0;
      continue;
    }
    double c = -1.0 / (double)n;
    for (int64_t j = 0; j < k; j++) {
      double *vj = out + j * n;
      double dot = 0.0;
      for (int64_t m = 0; m < j; m++)
        dot += vk[m] * vj[m];
      vk[j] = (c - dot) / vj[j];
    }
    double normSq = 0.0;
    for (int64_t m = 0; m < k; m++)
      normSq += vk[m] * vk[m];
    if (k < n) {
      double rem = 1.0 - normSq;
      vk[k] = rem > 0.0 ? sqrt(rem) : 0.0;
    }
  }
  double unitEdge = 0.0;
  for (int64_t m = 0; m < n; m++) {
    double d = out[m] - out[n + m];
    unitEdge += d * d;
  }
  unitEdge = sqrt(unitEdge);
  double scale = sideLength / unitEdge;
  for (int64_t i = 0; i < (n + 1) * n; i++)
    out[i] *= scale;
}
