/*
 * gviz_knearest.c -- gvizSearchKNearest* (reference:
 * src/algorithms/search/gvizKNearest.c).
 *
 * A single query (gvizSearchKNearest / gvizSearchKNearestScratch) is one
 * breadth-first search that stops at its k-th hit: a few microseconds on
 * the host, less than one kernel launch, and the subgraph would have to be
 * uploaded for it. It runs here, over the subgraph's adjacency in insertion
 * order, with the reference's epoch-stamped visited array and FIFO queue.
 * The many-target form (gvizSearchKNearestFromVisibleBatch: one full BFS
 * per visible seed) runs on the GPU through the GRIP engine's batched
 * search (gvizhip_grip_knn_batch), the device path GRIP's own placement
 * uses. Profile counters (GVIZ_KNN_PROFILE) count the host queries.
 */
#include "gviz/gvizKNearest.h"

#include "gviz_hip.h"

#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------- profile */
static _Atomic unsigned long long prof_queries, prof_visited, prof_max;

void gvizKNNProfileReset(void) {
  atomic_store(&prof_queries, 0);
  atomic_store(&prof_visited, 0);
  atomic_store(&prof_max, 0);
}

void gvizKNNProfileSnapshot(unsigned long long *queries, unsigned long long *visited,
                            unsigned long long *maxVisited) {
  if (queries)
    *queries = atomic_load(&prof_queries);
  if (visited)
    *visited = atomic_load(&prof_visited);
  if (maxVisited)
    *maxVisited = atomic_load(&prof_max);
}

static int profile_on(void) {
  static _Atomic int state = -1;
  int s = atomic_load_explicit(&state, memory_order_relaxed);
  if (s < 0) {
    s = getenv("GVIZ_KNN_PROFILE") != NULL;
    atomic_store_explicit(&state, s, memory_order_relaxed);
  }
  return s;
}

static void profile_add(size_t visited) {
  if (!profile_on())
    return;
  atomic_fetch_add_explicit(&prof_queries, 1, memory_order_relaxed);
  atomic_fetch_add_explicit(&prof_visited, visited, memory_order_relaxed);
  unsigned long long prev = atomic_load(&prof_max);
  while (visited > prev && !atomic_compare_exchange_weak(&prof_max, &prev, visited))
    ;
}

/* ------------------------------------------------- FIFO of found vertices */
/* The scratch's gvizDeque as a ring buffer of gvizFoundVertex: `begin`
 * points at the head element (NULL when the queue was reset). */
static int fifo_init(gvizDeque *q, size_t capacity) {
  q->arr = malloc(sizeof(gvizFoundVertex) * capacity);
  q->begin = NULL;
  q->count = 0;
  q->capacity = q->arr ? capacity : 0;
  q->elementSize = sizeof(gvizFoundVertex);
  return q->arr ? 0 : -1;
}

static void fifo_release(gvizDeque *q) {
  free(q->arr);
  memset(q, 0, sizeof(*q));
}

static size_t fifo_head(const gvizDeque *q) {
  return q->begin ? (size_t)((gvizFoundVertex *)q->begin - (gvizFoundVertex *)q->arr) : 0;
}

static int fifo_push(gvizDeque *q, gvizFoundVertex v) {
  gvizFoundVertex *a = (gvizFoundVertex *)q->arr;
  size_t head = fifo_head(q);
  if (q->count == q->capacity) {
    size_t cap = q->capacity ? 2 * q->capacity : 16;
    gvizFoundVertex *na = malloc(sizeof(gvizFoundVertex) * cap);
    if (!na)
      return -1;
    for (size_t i = 0; i < q->count; i++)
      na[i] = a[(head + i) % q->capacity];
    free(a);
    q->arr = a = na;
    q->capacity = cap;
    head = 0;
  }
  a[(head + q->count) % q->capacity] = v;
  q->count++;
  q->begin = &a[head];
  return 0;
}

static gvizFoundVertex fifo_pop(gvizDeque *q) {
  gvizFoundVertex *a = (gvizFoundVertex *)q->arr;
  size_t head = fifo_head(q);
  gvizFoundVertex v = a[head];
  q->count--;
  q->begin = &a[(head + 1) % q->capacity];
  return v;
}

/* -------------------------------------------------------------- scratch */
int gvizKNearestScratchInit(gvizKNearestScratch *scratch, size_t nvertices) {
  if (!scratch || nvertices == 0)
    return -1;
  scratch->stamp = calloc(nvertices, sizeof(size_t));
  if (!scratch->stamp)
    return -1;
  scratch->nvertices = nvertices;
  scratch->epoch = 1;
  if (fifo_init(&scratch->queue, 256) != 0) {
    free(scratch->stamp);
    scratch->stamp = NULL;
    return -1;
  }
  return 0;
}

void gvizKNearestScratchRelease(gvizKNearestScratch *scratch) {
  if (!scratch)
    return;
  free(scratch->stamp);
  fifo_release(&scratch->queue);
  scratch->stamp = NULL;
  scratch->nvertices = 0;
}

/* -------------------------------------------------------------- queries */
static int input_valid(const gvizSubgraph *sg) {
  return sg && sg->g && sg->g->layout && sg->vs;
}

/* src/algorithms/search/gvizKNearest.c:244-306: hits in discovery order,
 * each at its parent's distance + 1; the first k that pass the filter. */
int gvizSearchKNearestScratch(const gvizSubgraph *sg, gvizFoundVertex *out, size_t k,
                              size_t source, gvizVertexSubset filter,
                              gvizKNearestScratch *scratch) {
  if (k == 0)
    return 0;
  if (!out || !scratch || !scratch->stamp || !input_valid(sg))
    return -1;
  if (!gvizSubgraphHasVertex(sg, source))
    return -1;
  size_t n = sg->g->layout->nvertices;
  if (scratch->nvertices < n)
    return -1;
  size_t epoch = ++scratch->epoch;
  if (epoch == 0) {
    memset(scratch->stamp, 0, sizeof(size_t) * n);
    epoch = scratch->epoch = 1;
  }
  gvizDeque *q = &scratch->queue;
  q->count = 0;
  q->begin = NULL;
  scratch->stamp[source] = epoch;
  if (fifo_push(q, (gvizFoundVertex){source, 0}) != 0)
    return -1;
  const int full = gvizSubgraphIsFull(sg);
  size_t found = 0, visited = 1;
  while (q->count > 0) {
    gvizFoundVertex cur = fifo_pop(q);
    size_t deg = 0;
    const size_t *nb = gvizGraphNeighbors(sg->g, cur.v, &deg);
    for (size_t j = 0; j < deg; j++) {
      size_t w = nb[j];
      if (full ? !gvizEdgeSubsetTestEdge(sg->es, cur.v, j) : !gvizVertexSubsetTest(sg->vs, w))
        continue;
      if (scratch->stamp[w] == epoch)
        continue;
      scratch->stamp[w] = epoch;
      visited++;
      gvizFoundVertex hit = {w, cur.dist + 1};
      if (!filter || gvizVertexSubsetTest(filter, w)) {
        out[found++] = hit;
        if (found >= k) {
          profile_add(visited);
          return (int)k;
        }
      }
      if (fifo_push(q, hit) != 0)
        return -1;
    }
  }
  profile_add(visited);
  return (int)found;
}

int gvizSearchKNearest(const gvizSubgraph *sg, gvizFoundVertex *out, size_t k,
                       size_t source, gvizVertexSubset filter) {
  if (!input_valid(sg))
    return -1;
  gvizKNearestScratch scratch;
  if (gvizKNearestScratchInit(&scratch, sg->g->layout->nvertices) != 0)
    return -1;
  int r = gvizSearchKNearestScratch(sg, out, k, source, filter, &scratch);
  gvizKNearestScratchRelease(&scratch);
  return r;
}

/* :149-166 */
int gvizSearchKNearestPreferBatch(const gvizSubgraph *sg, size_t visibleCount,
                                  size_t targetCount) {
  if (!input_valid(sg) || visibleCount == 0 || visibleCount >= targetCount ||
      visibleCount > GVIZ_KNN_BATCH_VISIBLE_MAX)
    return 0;
  size_t vc = gvizSubgraphVertexCount(sg);
  if (vc == 0)
    return 0;
  double avgDeg = (2.0 * (double)gvizSubgraphEdgeCount(sg)) / (double)vc;
  if (avgDeg > GVIZ_KNN_BATCH_MIN_AVG_DEGREE)
    return 1;
  return targetCount >= GVIZ_KNN_BATCH_TARGET_RATIO * visibleCount;
}

/* :168-210 on the device: seeds are the visible vertices in ascending id,
 * each target keeps its k nearest seeds sorted by distance (ties: the
 * earlier seed), and seeds stop once every target holds k. Returns 0, -1 on
 * error (including no usable GPU), 1 when the batch path does not apply. */
int gvizSearchKNearestFromVisibleBatch(const gvizSubgraph *sg, gvizVertexSubset visible,
                                       size_t k, gvizKNNBatchTarget *targets,
                                       size_t targetCount, gvizKNearestScratch *scratch) {
  if (k == 0 || !visible || targetCount == 0 || !targets || !scratch || !scratch->stamp ||
      !input_valid(sg))
    return 1;
  size_t n = sg->g->layout->nvertices;
  if (scratch->nvertices < n)
    return -1;
  size_t visibleCount = gvizVertexSubsetCount(visible, n);
  if (visibleCount == 0 || visibleCount > GVIZ_KNN_BATCH_VISIBLE_MAX ||
      visibleCount >= targetCount)
    return 1;
  for (size_t i = 0; i < targetCount; i++) {
    targets[i].count = 0;
    if (targets[i].vertex >= n)
      return -1;
  }
  int64_t *off = NULL, *adj = NULL;
  uint8_t *active = malloc(n), *vis = malloc(n);
  int64_t *tg = malloc(sizeof(int64_t) * targetCount);
  int64_t *ov = malloc(sizeof(int64_t) * targetCount * k);
  int64_t *od = malloc(sizeof(int64_t) * targetCount * k);
  int64_t *cnt = malloc(sizeof(int64_t) * targetCount);
  gvizhip_grip *dev = NULL;
  int rc = -1;
  if (!active || !vis || !tg || !ov || !od || !cnt || gvizSubgraphToCSR(sg, &off, &adj) != 0)
    goto done;
  gvizVertexSubsetToBytes(sg->vs, n, active);
  gvizVertexSubsetToBytes(visible, n, vis);
  /* a target listed twice: the reference's map keeps the last entry, the
   * earlier ones stay empty */
  for (size_t i = 0; i < targetCount; i++)
    tg[i] = (int64_t)targets[i].vertex;
  if (gvizhip_grip_create(&dev, (int64_t)n, off, adj, active, (int64_t)k) != 0 ||
      gvizhip_grip_knn_batch(dev, vis, tg, (int64_t)targetCount, (int64_t)k, ov, od,
                             cnt) != 0)
    goto done;
  {
    size_t *last = malloc(sizeof(size_t) * n);
    if (!last)
      goto done;
    for (size_t i = 0; i < targetCount; i++)
      last[targets[i].vertex] = i;
    for (size_t i = 0; i < targetCount; i++) {
      if (last[targets[i].vertex] != i)
        continue;
      size_t c = (size_t)cnt[i];
      for (size_t j = 0; j < c; j++)
        targets[i].out[j] = (gvizFoundVertex){(size_t)ov[i * k + j], (size_t)od[i * k + j]};
      targets[i].count = c;
    }
    free(last);
  }
  rc = 0;
done:
  if (dev)
    gvizhip_grip_destroy(dev);
  free(off);
  free(adj);
  free(active);
  free(vis);
  free(tg);
  free(ov);
  free(od);
  free(cnt);
  return rc;
}
