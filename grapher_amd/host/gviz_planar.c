/*
 * gviz_planar.c -- planarity test, rotation system and largest face for the
 * Tutte embedders' Begin (reference: src/embedders/gvizTutteEmbedder.c:
 * 145-173 with gvizPlanarApplyRotationToEmbedding and
 * gvizPlanarLargestFaceBoundary, src/embedders/gvizPlanarEmbedder.c:437-479,
 * 582-590).
 *
 * The reference tests planarity and embeds with the third-party edge
 * addition planarity suite (third-party/boyerMyrvold, Boyer and Myrvold,
 * "On the cutting edge: simplified O(n) planarity by edge addition", JGAA
 * 8(3), 2004). This file uses the left-right criterion instead (U. Brandes,
 * "The Left-Right Planarity Test", 2009): orientation DFS with lowpoints
 * and nesting depths, the testing DFS over conflict pairs of return-edge
 * intervals, then sign resolution and the embedding DFS that places every
 * back edge beside the left or right reference of its target. Both decide
 * planarity exactly. A planar graph's rotation system is not unique, so the
 * faces can differ from the suite's; for a 3-connected planar graph (the
 * graphs Tutte's theorem is about) the faces are unique up to the mirror
 * image, so the largest face is the same vertex cycle, possibly starting
 * at another vertex or running the other way round.
 *
 * All three DFS passes keep explicit stacks (a path can be as long as the
 * graph). Self-loops and repeated edges are ignored, as a simple graph.
 */
#include "gviz_planar.h"

#include "gviz/gvizEmbeddedGraph.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int64_t lLow, lHigh, rLow, rHigh; /* darts, -1 = none */
} Pair;

typedef struct {
  int64_t n;
  const int64_t *off, *adj;
  int64_t *twin;     /* partner dart, -1 for an ignored dart */
  int64_t *src;      /* source vertex of each dart */
  uint8_t *orient;   /* dart is the oriented direction of its edge */
  int64_t *height, *parent, *ind;
  int64_t *lowpt, *lowpt2, *nest;
  int64_t *ord, *ordStart; /* oriented out-darts sorted by nesting depth */
  int64_t *ref, *lowptEdge, *stackBottom;
  int8_t *side;
  Pair *S;
  int64_t top; /* number of pairs on S */
  int64_t *cw, *ccw, *first, *leftRef, *rightRef;
  int64_t *stack, *tmp;
  uint8_t *skip;
} LR;

static int64_t dst(const LR *g, int64_t d) { return g->adj[d]; }

static int i_empty(int64_t lo, int64_t hi) { return lo < 0 && hi < 0; }

/* Interval [low, high] conflicts with edge b: nonempty and its highest
 * return edge returns above lowpt(b). */
static int conflicting(const LR *g, int64_t lo, int64_t hi, int64_t b) {
  return !i_empty(lo, hi) && g->lowpt[hi] > g->lowpt[b];
}

static int64_t pair_lowest(const LR *g, const Pair *p) {
  if (p->lLow < 0 && p->rLow < 0)
    return INT64_MAX;
  if (i_empty(p->lLow, p->lHigh) || p->lLow < 0)
    return g->lowpt[p->rLow];
  if (i_empty(p->rLow, p->rHigh) || p->rLow < 0)
    return g->lowpt[p->lLow];
  const int64_t a = g->lowpt[p->lLow], b = g->lowpt[p->rLow];
  return a < b ? a : b;
}

static void pair_swap(Pair *p) {
  int64_t a = p->lLow, b = p->lHigh;
  p->lLow = p->rLow;
  p->lHigh = p->rHigh;
  p->rLow = a;
  p->rHigh = b;
}

/* Twin darts of the simple graph underlying the CSR. */
static int pair_darts(LR *g) {
  const int64_t n = g->n, m2 = g->off[n];
  int64_t *key = malloc(sizeof(int64_t) * (size_t)(m2 ? m2 : 1));
  if (!key)
    return -1;
  for (int64_t v = 0; v < n; v++)
    for (int64_t d = g->off[v]; d < g->off[v + 1]; d++) {
      g->src[d] = v;
      g->twin[d] = -1;
    }
  /* per vertex: darts sorted by (target, dart) */
  for (int64_t v = 0; v < n; v++) {
    const int64_t b = g->off[v], e = g->off[v + 1];
    for (int64_t d = b; d < e; d++)
      key[d] = d;
    /* shell sort by (target, dart) */
    for (int64_t gap = (e - b) / 2; gap > 0; gap /= 2)
      for (int64_t i = b + gap; i < e; i++) {
        int64_t t = key[i], j = i;
        while (j - gap >= b && (g->adj[key[j - gap]] > g->adj[t] ||
                                (g->adj[key[j - gap]] == g->adj[t] && key[j - gap] > t))) {
          key[j] = key[j - gap];
          j -= gap;
        }
        key[j] = t;
      }
  }
  /* a dart (v -> w) is kept when w != v and it is the first to w; its twin
   * is the first dart (w -> v), found by binary search in w's sorted list */
  for (int64_t v = 0; v < n; v++) {
    const int64_t b = g->off[v], e = g->off[v + 1];
    for (int64_t i = b; i < e; i++) {
      const int64_t d = key[i], w = g->adj[d];
      if (w == v || (i > b && g->adj[key[i - 1]] == w))
        continue;
      if (w < v)
        continue; /* paired from the lower end */
      int64_t lo = g->off[w], hi = g->off[w + 1];
      while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (g->adj[key[mid]] < v)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < g->off[w + 1] && g->adj[key[lo]] == v) {
        g->twin[d] = key[lo];
        g->twin[key[lo]] = d;
      }
    }
  }
  free(key);
  return 0;
}

/* Phase 1: DFS orientation, heights, lowpoints and nesting depths. */
static void orientation(LR *g, int64_t root) {
  int64_t sp = 0;
  g->stack[sp++] = root;
  while (sp > 0) {
    const int64_t v = g->stack[sp - 1];
    const int64_t e = g->parent[v];
    int descended = 0;
    while (g->ind[v] < g->off[v + 1]) {
      const int64_t d = g->ind[v];
      if (g->twin[d] < 0) {
        g->ind[v]++;
        continue;
      }
      const int64_t w = dst(g, d);
      if (!g->skip[d]) {
        if (g->orient[d] || g->orient[g->twin[d]]) {
          g->ind[v]++;
          continue;
        }
        g->orient[d] = 1;
        g->lowpt[d] = g->height[v];
        g->lowpt2[d] = g->height[v];
        if (g->height[w] < 0) { /* tree edge */
          g->parent[w] = d;
          g->height[w] = g->height[v] + 1;
          g->skip[d] = 1;
          g->stack[sp++] = w;
          descended = 1;
          break;
        }
        g->lowpt[d] = g->height[w]; /* back edge */
      }
      g->nest[d] = 2 * g->lowpt[d] + (g->lowpt2[d] < g->height[v] ? 1 : 0);
      if (e >= 0) {
        if (g->lowpt[d] < g->lowpt[e]) {
          g->lowpt2[e] = g->lowpt[e] < g->lowpt2[d] ? g->lowpt[e] : g->lowpt2[d];
          g->lowpt[e] = g->lowpt[d];
        } else if (g->lowpt[d] > g->lowpt[e]) {
          g->lowpt2[e] = g->lowpt2[e] < g->lowpt[d] ? g->lowpt2[e] : g->lowpt[d];
        } else {
          g->lowpt2[e] = g->lowpt2[e] < g->lowpt2[d] ? g->lowpt2[e] : g->lowpt2[d];
        }
      }
      g->ind[v]++;
    }
    if (!descended)
      sp--;
  }
}

/* a[0..len) stably sorted by key[]: insertion sort on runs of 16, then
 * bottom-up merges through tmp */
static void stable_sort(int64_t *a, int64_t len, const int64_t *key, int64_t *tmp) {
  for (int64_t r = 0; r < len; r += 16) {
    const int64_t e = r + 16 < len ? r + 16 : len;
    for (int64_t i = r + 1; i < e; i++) {
      const int64_t t = a[i];
      int64_t j = i;
      while (j > r && key[a[j - 1]] > key[t]) {
        a[j] = a[j - 1];
        j--;
      }
      a[j] = t;
    }
  }
  for (int64_t w = 16; w < len; w *= 2) {
    for (int64_t lo = 0; lo < len; lo += 2 * w) {
      const int64_t mid = lo + w < len ? lo + w : len;
      const int64_t hi = lo + 2 * w < len ? lo + 2 * w : len;
      int64_t i = lo, j = mid, k = lo;
      while (i < mid && j < hi)
        tmp[k++] = key[a[j]] < key[a[i]] ? a[j++] : a[i++];
      while (i < mid)
        tmp[k++] = a[i++];
      while (j < hi)
        tmp[k++] = a[j++];
    }
    memcpy(a, tmp, sizeof(int64_t) * (size_t)len);
  }
}

/* Oriented out-darts of every vertex, stably sorted by key[]. */
static void order_out_darts(LR *g, const int64_t *key) {
  int64_t k = 0;
  for (int64_t v = 0; v < g->n; v++) {
    g->ordStart[v] = k;
    for (int64_t d = g->off[v]; d < g->off[v + 1]; d++)
      if (g->twin[d] >= 0 && g->orient[d])
        g->ord[k++] = d;
    stable_sort(g->ord + g->ordStart[v], k - g->ordStart[v], key, g->tmp);
  }
  g->ordStart[g->n] = k;
}

static int add_constraints(LR *g, int64_t ei, int64_t e) {
  Pair P = {-1, -1, -1, -1};
  /* merge return edges of ei into P.right */
  do {
    Pair Q = g->S[--g->top];
    if (!i_empty(Q.lLow, Q.lHigh))
      pair_swap(&Q);
    if (!i_empty(Q.lLow, Q.lHigh))
      return 0;
    if (g->lowpt[Q.rLow] > g->lowpt[e]) {
      if (i_empty(P.rLow, P.rHigh)) {
        P.rLow = Q.rLow;
        P.rHigh = Q.rHigh;
      } else if (P.rLow >= 0) {
        g->ref[P.rLow] = Q.rHigh;
      }
      P.rLow = Q.rLow;
    } else if (Q.rLow >= 0) {
      g->ref[Q.rLow] = g->lowptEdge[e];
    }
  } while (g->top != g->stackBottom[ei]);
  /* merge conflicting return edges of the earlier siblings into P.left */
  while (g->top > 0 && (conflicting(g, g->S[g->top - 1].lLow, g->S[g->top - 1].lHigh, ei) ||
                        conflicting(g, g->S[g->top - 1].rLow, g->S[g->top - 1].rHigh, ei))) {
    Pair Q = g->S[--g->top];
    if (conflicting(g, Q.rLow, Q.rHigh, ei))
      pair_swap(&Q);
    if (conflicting(g, Q.rLow, Q.rHigh, ei))
      return 0;
    if (P.rLow >= 0)
      g->ref[P.rLow] = Q.rHigh;
    if (Q.rLow >= 0)
      P.rLow = Q.rLow;
    if (i_empty(P.lLow, P.lHigh)) {
      P.lLow = Q.lLow;
      P.lHigh = Q.lHigh;
    } else if (P.lLow >= 0) {
      g->ref[P.lLow] = Q.lHigh;
    }
    P.lLow = Q.lLow;
  }
  if (!(i_empty(P.lLow, P.lHigh) && i_empty(P.rLow, P.rHigh)))
    g->S[g->top++] = P;
  return 1;
}

static void remove_back_edges(LR *g, int64_t e) {
  const int64_t u = g->src[e];
  while (g->top > 0 && pair_lowest(g, &g->S[g->top - 1]) == g->height[u]) {
    const Pair P = g->S[--g->top];
    if (P.lLow >= 0)
      g->side[P.lLow] = -1;
  }
  if (g->top > 0) {
    Pair P = g->S[--g->top];
    while (P.lHigh >= 0 && dst(g, P.lHigh) == u)
      P.lHigh = g->ref[P.lHigh];
    if (P.lHigh < 0 && P.lLow >= 0) {
      g->ref[P.lLow] = P.rLow;
      g->side[P.lLow] = -1;
      P.lLow = -1;
    }
    while (P.rHigh >= 0 && dst(g, P.rHigh) == u)
      P.rHigh = g->ref[P.rHigh];
    if (P.rHigh < 0 && P.rLow >= 0) {
      g->ref[P.rLow] = P.lLow;
      g->side[P.rLow] = -1;
      P.rLow = -1;
    }
    g->S[g->top++] = P;
  }
  if (g->lowpt[e] < g->height[u] && g->top > 0) {
    const int64_t hl = g->S[g->top - 1].lHigh, hr = g->S[g->top - 1].rHigh;
    if (hl >= 0 && (hr < 0 || g->lowpt[hl] > g->lowpt[hr]))
      g->ref[e] = hl;
    else
      g->ref[e] = hr;
  }
}

/* Phase 2: the testing DFS. 1 planar, 0 not. */
static int testing(LR *g, int64_t root) {
  int64_t sp = 0;
  g->stack[sp++] = root;
  while (sp > 0) {
    const int64_t v = g->stack[--sp];
    const int64_t e = g->parent[v];
    int descended = 0;
    const int64_t b = g->ordStart[v], end = g->ordStart[v + 1];
    for (; g->ind[v] < end; g->ind[v]++) {
      const int64_t ei = g->ord[g->ind[v]], w = dst(g, ei);
      if (!g->skip[ei]) {
        g->stackBottom[ei] = g->top;
        if (ei == g->parent[w]) { /* tree edge: finish w first */
          g->stack[sp++] = v;
          g->stack[sp++] = w;
          g->skip[ei] = 1;
          descended = 1;
          break;
        }
        g->lowptEdge[ei] = ei; /* back edge */
        g->S[g->top++] = (Pair){-1, -1, ei, ei};
      }
      if (g->lowpt[ei] < g->height[v]) {
        if (ei == g->ord[b])
          g->lowptEdge[e] = g->lowptEdge[ei];
        else if (!add_constraints(g, ei, e))
          return 0;
      }
    }
    if (!descended && e >= 0)
      remove_back_edges(g, e);
  }
  return 1;
}

static int64_t sign_of(LR *g, int64_t e) {
  /* follow the reference chain, then resolve it from its end */
  int64_t len = 0;
  for (int64_t x = e; x >= 0; x = g->ref[x])
    g->stack[len++] = x;
  for (int64_t i = len - 2; i >= 0; i--) {
    const int64_t x = g->stack[i];
    g->side[x] = (int8_t)(g->side[x] * g->side[g->stack[i + 1]]);
    g->ref[x] = -1;
  }
  return g->side[e];
}

/* rotation system: circular cw / ccw dart lists per vertex */
static void add_cw(LR *g, int64_t v, int64_t d, int64_t refd) {
  if (refd < 0) {
    g->cw[d] = g->ccw[d] = d;
    g->first[v] = d;
    return;
  }
  const int64_t c = g->cw[refd];
  g->cw[d] = c;
  g->ccw[d] = refd;
  g->cw[refd] = d;
  g->ccw[c] = d;
}

static void add_ccw(LR *g, int64_t v, int64_t d, int64_t refd) {
  if (refd < 0) {
    add_cw(g, v, d, -1);
    return;
  }
  add_cw(g, v, d, g->ccw[refd]);
  if (refd == g->first[v])
    g->first[v] = d;
}

/* Phase 3: place every reverse dart beside its target's references. */
static void embedding(LR *g, int64_t root) {
  int64_t sp = 0;
  g->stack[sp++] = root;
  while (sp > 0) {
    const int64_t v = g->stack[--sp];
    const int64_t end = g->ordStart[v + 1];
    while (g->ind[v] < end) {
      const int64_t ei = g->ord[g->ind[v]++], w = dst(g, ei), back = g->twin[ei];
      if (ei == g->parent[w]) {
        add_ccw(g, w, back, g->first[w]);
        g->leftRef[v] = ei;
        g->rightRef[v] = ei;
        g->stack[sp++] = v;
        g->stack[sp++] = w;
        break;
      }
      if (g->side[ei] == 1) {
        add_cw(g, w, back, g->rightRef[w]);
      } else {
        add_ccw(g, w, back, g->leftRef[w]);
        g->leftRef[w] = back;
      }
    }
  }
}

int gviz_planar_largest_face(int64_t n, const int64_t *off, const int64_t *adj,
                             size_t *out, size_t max, size_t *count) {
  return gviz_planar_embed(n, off, adj, out, max, count, NULL, NULL);
}

int gviz_planar_embed(int64_t n, const int64_t *off, const int64_t *adj, size_t *out,
                      size_t max, size_t *count, int64_t *rotation,
                      int64_t *rotationCount) {
  if (n < 0 || !off || (n > 0 && !adj) || !out || !count)
    return -1;
  *count = 0;
  const int64_t m2 = off[n];
  LR g;
  memset(&g, 0, sizeof(g));
  g.n = n;
  g.off = off;
  g.adj = adj;
  const size_t D = (size_t)(m2 > 0 ? m2 : 1), V = (size_t)(n + 1);
  g.twin = malloc(8 * D);
  g.src = malloc(8 * D);
  g.orient = calloc(D, 1);
  g.skip = calloc(D, 1);
  g.height = malloc(8 * V);
  g.parent = malloc(8 * V);
  g.ind = malloc(8 * V);
  g.lowpt = malloc(8 * D);
  g.lowpt2 = malloc(8 * D);
  g.nest = malloc(8 * D);
  g.ord = malloc(8 * D);
  g.ordStart = malloc(8 * V);
  g.ref = malloc(8 * D);
  g.lowptEdge = malloc(8 * D);
  g.stackBottom = malloc(8 * D);
  g.side = malloc(D);
  g.S = malloc(sizeof(Pair) * D);
  g.cw = malloc(8 * D);
  g.ccw = malloc(8 * D);
  g.first = malloc(8 * V);
  g.leftRef = malloc(8 * V);
  g.rightRef = malloc(8 * V);
  g.stack = malloc(8 * (D > V ? D : V));
  g.tmp = malloc(8 * D);
  uint8_t *seen = calloc(D, 1);
  int rc = -1;
  if (!g.twin || !g.src || !g.orient || !g.skip || !g.height || !g.parent || !g.ind ||
      !g.lowpt || !g.lowpt2 || !g.nest || !g.ord || !g.ordStart || !g.ref ||
      !g.lowptEdge || !g.stackBottom || !g.side || !g.S || !g.cw || !g.ccw ||
      !g.first || !g.leftRef || !g.rightRef || !g.stack || !g.tmp || !seen)
    goto done;
  if (pair_darts(&g) != 0)
    goto done;
  int64_t edges = 0;
  for (int64_t d = 0; d < m2; d++)
    edges += g.twin[d] >= 0;
  edges /= 2;
  if (n > 2 && edges > 3 * n - 6) {
    rc = -2;
    goto done;
  }
  for (int64_t v = 0; v < n; v++) {
    g.height[v] = -1;
    g.parent[v] = -1;
    g.ind[v] = off[v];
    g.first[v] = -1;
    g.leftRef[v] = g.rightRef[v] = -1;
  }
  for (int64_t d = 0; d < m2; d++) {
    g.ref[d] = -1;
    g.side[d] = 1;
    g.lowptEdge[d] = -1;
  }
  int64_t nroots = 0;
  for (int64_t v = 0; v < n; v++)
    if (g.height[v] < 0) {
      g.height[v] = 0;
      orientation(&g, v);
      nroots++;
    }
  order_out_darts(&g, g.nest);
  memset(g.skip, 0, D);
  for (int64_t v = 0; v < n; v++)
    g.ind[v] = g.ordStart[v];
  for (int64_t v = 0; v < n; v++)
    if (g.height[v] == 0 && !testing(&g, v)) {
      rc = -2;
      goto done;
    }
  /* signed nesting depths order each vertex's out-darts in the embedding */
  for (int64_t k = 0; k < g.ordStart[n]; k++) {
    const int64_t d = g.ord[k];
    g.lowpt2[d] = sign_of(&g, d) * g.nest[d]; /* lowpt2 reused as the key */
  }
  order_out_darts(&g, g.lowpt2);
  for (int64_t v = 0; v < n; v++) {
    int64_t prev = -1;
    for (int64_t k = g.ordStart[v]; k < g.ordStart[v + 1]; k++) {
      add_cw(&g, v, g.ord[k], prev);
      prev = g.ord[k];
    }
    g.ind[v] = g.ordStart[v];
  }
  for (int64_t v = 0; v < n; v++)
    if (g.height[v] == 0)
      embedding(&g, v);
  (void)nroots;
  /* faces: the dart after (v -> w) is w's dart ccw of (w -> v); the first
   * largest face in dart order wins */
  size_t best = 0;
  int64_t bestStart = -1;
  for (int64_t d0 = 0; d0 < m2; d0++) {
    if (g.twin[d0] < 0 || seen[d0])
      continue;
    size_t len = 0;
    int64_t d = d0;
    do {
      if (seen[d] || g.twin[d] < 0 || len > (size_t)m2)
        goto done; /* not a rotation system: cannot happen for a planar result */
      seen[d] = 1;
      len++;
      d = g.ccw[g.twin[d]];
    } while (d != d0);
    if (len >= 3 && len > best) {
      best = len;
      bestStart = d0;
    }
  }
  if (bestStart < 0 || best > max)
    goto done;
  {
    int64_t d = bestStart;
    size_t k = 0;
    do {
      out[k++] = (size_t)g.src[d];
      d = g.ccw[g.twin[d]];
    } while (d != bestStart);
    *count = k;
  }
  /* each vertex's darts in cw order (the face walk takes the dart before
   * the reverse one), from its earliest dart in adjacency order */
  if (rotation && rotationCount)
    for (int64_t v = 0; v < n; v++) {
      rotationCount[v] = 0;
      const int64_t f = g.first[v];
      if (f < 0)
        continue;
      int64_t start = f;
      for (int64_t d = g.cw[f]; d != f; d = g.cw[d])
        if (d < start)
          start = d;
      int64_t k = off[v], d = start;
      do {
        rotation[k++] = dst(&g, d);
        d = g.cw[d];
      } while (d != start);
      rotationCount[v] = k - off[v];
    }
  rc = 0;
done:
  free(g.twin);
  free(g.src);
  free(g.orient);
  free(g.skip);
  free(g.height);
  free(g.parent);
  free(g.ind);
  free(g.lowpt);
  free(g.lowpt2);
  free(g.nest);
  free(g.ord);
  free(g.ordStart);
  free(g.ref);
  free(g.lowptEdge);
  free(g.stackBottom);
  free(g.side);
  free(g.S);
  free(g.cw);
  free(g.ccw);
  free(g.first);
  free(g.leftRef);
  free(g.rightRef);
  free(g.stack);
  free(g.tmp);
  free(seen);
  return rc;
}

/* ------------------------------------------------- embedded-graph side */

/* gvizPlanarApplyRotationToEmbedding (src/embedders/gvizPlanarEmbedder.c:
 * 74-104, 160-181, 582-590) with gvizPlanarLargestFaceBoundary. */
int gviz_planar_embed_largest_face(gvizEmbeddedGraph *e, size_t *out, size_t max,
                                   size_t *count) {
  gvizSubgraph *sg = &e->subgraph;
  gvizGraph *G = (gvizGraph *)sg->g;
  const size_t n = gvizGraphSize(G);
  int64_t *off = NULL, *adj = NULL;
  if (gvizSubgraphToCSR(sg, &off, &adj) != 0)
    return -1;
  int64_t *rot = malloc(sizeof(int64_t) * (off[n] > 0 ? off[n] : 1));
  int64_t *rc = malloc(sizeof(int64_t) * (n ? n : 1));
  size_t *list = NULL;
  int res = -1;
  if (!rot || !rc)
    goto done;
  res = gviz_planar_embed((int64_t)n, off, adj, out, max, count, rot, rc);
  if (res != 0)
    goto done;
  res = -1;
  /* each subgraph vertex's list: its rotation, then the entries that are
   * not subgraph edges in their order (a neighbour already listed is not
   * repeated; repeated subgraph edges and loops, which the embedding
   * ignores, are dropped as the reference drops them) */
  {
    size_t cap = 0;
    for (size_t u = 0; u < n; u++) {
      size_t c = 0;
      gvizGraphNeighbors(G, u, &c);
      if (c > cap)
        cap = c;
    }
    list = malloc(sizeof(size_t) * (cap ? cap : 1));
    if (!list)
      goto done;
    for (size_t u = 0; u < n; u++) {
      if (!gvizSubgraphHasVertex(sg, u))
        continue;
      gvizArray *nb = gvizGraphGetVertexNeighbors(G, u);
      size_t k = 0;
      for (int64_t j = 0; j < rc[u]; j++)
        list[k++] = (size_t)rot[off[u] + j];
      for (size_t j = 0; j < nb->count; j++) {
        size_t v = ((size_t *)nb->arr)[j];
        if (gvizSubgraphHasEdge(sg, u, v))
          continue;
        size_t t = 0;
        while (t < k && list[t] != v)
          t++;
        if (t == k)
          list[k++] = v;
      }
      memcpy(nb->arr, list, sizeof(size_t) * k);
      nb->count = k;
    }
  }
  gvizGraphBuildLayout(G);
  if (!G->layout)
    goto done;
  if (gvizSubgraphIsFull(sg)) {
    const gvizGraphLayout *L = G->layout;
    GVIZ_BIT_ARRAY bits = calloc(GVIZ_ARRAY_UNITS(L->edgeCount) ? GVIZ_ARRAY_UNITS(L->edgeCount) : 1, 1);
    if (!bits)
      goto done;
    for (size_t u = 0; u < n; u++)
      if (gvizSubgraphHasVertex(sg, u))
        for (int64_t j = 0; j < rc[u]; j++)
          gvizSetBit(bits, L->vertexOffsets[u] + (size_t)j);
    free(sg->es.bitset);
    sg->es.bitset = bits;
    sg->es.layout = L;
  }
  e->planarEmbedded = 1;
  res = 0;
done:
  free(off);
  free(adj);
  free(rot);
  free(rc);
  free(list);
  return res;
}

/* the adjacency entry of v before u (cyclically): gvizPlanarPrevNeighborCCW
 * (src/embedders/gvizPlanarEmbedder.c:205-211) */
static int prev_neighbor(const gvizGraph *G, size_t v, size_t u, size_t *w) {
  size_t c = 0;
  const size_t *nb = gvizGraphNeighbors(G, v, &c);
  for (size_t i = 0; i < c; i++)
    if (nb[i] == u) {
      *w = nb[i == 0 ? c - 1 : i - 1];
      return 0;
    }
  return -1;
}

/* boundaryCycleFromSubgraph (src/embedders/gvizTutteEmbedder.c:105-143):
 * the first edge (u, v) of the highlight (u ascending, v in adjacency
 * order), then the face walk (u, v) -> (v, entry of v before u) through the
 * graph's rotation until it returns to (u, v). */
int gviz_planar_face_from_highlight(const gvizEmbeddedGraph *e, size_t *out, size_t max,
                                    size_t *count) {
  *count = 0;
  const gvizSubgraph *h = gvizEmbeddedGraphGetHighlight(e);
  if (!h)
    return -1;
  size_t su = 0, sv = 0;
  int found = 0;
  gvizSubgraphVertexIterator vit = gvizSubgraphVertexIteratorCreate(h);
  size_t u;
  while (!found && gvizSubgraphVertexIterate(&vit, &u)) {
    gvizSubgraphNeighborIterator nit = gvizSubgraphNeighborIteratorCreate(h, u);
    size_t v;
    while (gvizSubgraphNeighborIterate(&nit, &v))
      if (gvizSubgraphHasEdge(h, u, v)) {
        su = u;
        sv = v;
        found = 1;
        break;
      }
  }
  if (!found)
    return -1;
  size_t cu = su, cv = sv;
  for (;;) {
    if (*count >= max)
      return -1;
    out[(*count)++] = cv;
    size_t w;
    if (prev_neighbor(h->g, cv, cu, &w) != 0)
      return -1;
    cu = cv;
    cv = w;
    if (cu == su && cv == sv)
      break;
  }
  return *count >= 3 ? 0 : -1;
}
