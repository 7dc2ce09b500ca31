/*
 * gviz_quadtree.c -- gvizQuadtree* (reference: src/ds/gvizQuadtree.c) as a
 * host view of a tree whose keys are built on the device.
 *
 * The device (gvizhip_qtree_build, or gvizhip_force_quadtree_keys for an
 * embedder's tree) returns the root cell and the points sorted by quadrant
 * path. A cell of the reference's tree is a path prefix: it splits iff it
 * holds more than nodesPerCell points and its half size is above 1e-9
 * (src/ds/gvizQuadtree.c:133-147), which depends only on the final point
 * set, not on the insertion sequence. So the host walks the sorted paths
 * top-down, splitting a cell's range of the sorted array by the next digit,
 * and gives each leaf its points in ascending index order (the order the
 * reference's index-order insertion leaves them in). Mass and centre of mass
 * are the reference's running mean over each cell's points in index order
 * (:112-123); each point, in index order, updates the cells on its path,
 * which is the same sequence of operations per cell.
 */
#include "gviz/gvizQuadtree.h"

#include "gviz_hip.h"
#include "gviz_quadtree_view.h"
#include "gviz_side.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define QT_MIN_HALF_SIZE 1e-9 /* src/ds/gvizQuadtree.c:6 */
#define QT_MIN_MASS 1e-9      /* :7 */
#define QT_BLOCK_NODES 1024   /* :8 */

/* The device context of a standalone tree, or the owner that refreshes the
 * tree of an embedder on demand. Keyed by the tree's address. */
typedef struct qt_side {
  gvizhip_qtree *dev;
  int (*refresh)(void *owner);
  void *owner;
  int stale;
} qt_side;

static const char *qt_error = "";

static qt_side *qt_side_of(const gvizQuadtree *t, int create) {
  qt_side *x = (qt_side *)gviz_side_get(t);
  if (x || !create)
    return x;
  x = calloc(1, sizeof(qt_side));
  if (!x)
    return NULL;
  void *prev = NULL;
  if (gviz_side_put(t, x, &prev) != 0) {
    free(x);
    return NULL;
  }
  return x;
}

static gvizQuadtreeNode *qt_node(gvizQuadtree *t, double cx, double cy, double half) {
  size_t b = t->nodesUsed / QT_BLOCK_NODES, o = t->nodesUsed % QT_BLOCK_NODES;
  if (b == t->nodeBlocks.count) {
    gvizQuadtreeNode *nodes = malloc(sizeof(gvizQuadtreeNode) * QT_BLOCK_NODES);
    size_t *pts = t->nodesPerCell > 0
                      ? malloc(sizeof(size_t) * QT_BLOCK_NODES * t->nodesPerCell)
                      : NULL;
    if (!nodes || (t->nodesPerCell > 0 && !pts)) {
      free(nodes);
      free(pts);
      return NULL;
    }
    if (gvizArrayPush(&t->nodeBlocks, &nodes) != 0) {
      free(nodes);
      free(pts);
      return NULL;
    }
    if (gvizArrayPush(&t->pointBlocks, &pts) != 0) {
      free(pts);
      return NULL;
    }
  }
  gvizQuadtreeNode *nodes = *(gvizQuadtreeNode **)gvizArrayAtIndex(&t->nodeBlocks, b);
  size_t *pts = *(size_t **)gvizArrayAtIndex(&t->pointBlocks, b);
  gvizQuadtreeNode *nd = &nodes[o];
  nd->cx = cx;
  nd->cy = cy;
  nd->halfSize = half;
  nd->comX = nd->comY = nd->mass = 0.0;
  nd->pointIndices = t->nodesPerCell > 0 ? &pts[o * t->nodesPerCell] : NULL;
  nd->pointCount = 0;
  nd->pointCapacity = t->nodesPerCell;
  memset(nd->children, 0, sizeof(nd->children));
  t->nodesUsed++;
  return nd;
}

static void qt_free_overflow(gvizQuadtree *t) {
  for (size_t i = 0; i < t->overflowBuffers.count; i++)
    free(*(size_t **)gvizArrayAtIndex(&t->overflowBuffers, i));
  t->overflowBuffers.count = 0;
}

static inline int qt_digit(uint64_t hi, uint64_t lo, int level) {
  return level <= 32 ? (int)((hi >> (2 * (32 - level))) & 3u)
                     : (int)((lo >> (2 * (64 - level))) & 3u);
}

static int cmp_size(const void *a, const void *b) {
  size_t x = *(const size_t *)a, y = *(const size_t *)b;
  return x < y ? -1 : x > y;
}

/* A leaf's points: the block slot while they fit, else overflow buffers
 * grown by doubling as the reference grows them (:94-110; one buffer per
 * growth step). */
static int qt_fill_leaf(gvizQuadtree *t, gvizQuadtreeNode *nd, const uint32_t *order,
                        size_t a, size_t b) {
  size_t cnt = b - a;
  if (cnt > nd->pointCapacity) {
    size_t cap = nd->pointCapacity;
    while (cap < cnt) {
      cap = cap > 0 ? cap * 2 : 1;
      size_t *buf = malloc(sizeof(size_t) * cap);
      if (!buf)
        return -1;
      if (gvizArrayPush(&t->overflowBuffers, &buf) != 0) {
        free(buf);
        return -1;
      }
      nd->pointIndices = buf;
    }
    nd->pointCapacity = cap;
  }
  for (size_t s = a; s < b; s++)
    nd->pointIndices[s - a] = order[s];
  if (cnt > 1) {
    if (cnt <= 16) {
      for (size_t i = 1; i < cnt; i++)
        for (size_t j = i; j > 0 && nd->pointIndices[j - 1] > nd->pointIndices[j]; j--) {
          size_t tmp = nd->pointIndices[j];
          nd->pointIndices[j] = nd->pointIndices[j - 1];
          nd->pointIndices[j - 1] = tmp;
        }
    } else {
      qsort(nd->pointIndices, cnt, sizeof(size_t), cmp_size);
    }
  }
  nd->pointCount = cnt;
  return 0;
}

typedef struct {
  gvizQuadtreeNode *nd;
  size_t a, b;
  int depth;
} qt_item;

int gviz_quadtree_from_keys(gvizQuadtree *t, const double *root, const uint32_t *order,
                            const uint64_t *keyHi, const uint64_t *keyLo) {
  t->root = NULL;
  t->nodesUsed = 0;
  qt_free_overflow(t);
  size_t n = t->pointCount;
  if (n == 0)
    return 0;
  const int levels = (int)root[3];
  gvizQuadtreeNode *r = qt_node(t, root[0], root[1], root[2]);
  if (!r)
    return -1;
  /* depth-first over cells; the stack holds at most 3 siblings per level */
  qt_item stack[4 * 66];
  int top = 0;
  stack[top++] = (qt_item){r, 0, n, 0};
  while (top > 0) {
    qt_item it = stack[--top];
    gvizQuadtreeNode *nd = it.nd;
    size_t cnt = it.b - it.a;
    if (cnt > t->nodesPerCell && nd->halfSize > QT_MIN_HALF_SIZE) {
      if (it.depth >= levels) {
        qt_error = "gvizQuadtree: points closer than 64 path levels separate";
        return -1;
      }
      /* :75-92: children NW, NE, SW, SE at centre -/+ halfSize / 2 */
      double quarter = nd->halfSize / 2.0;
      const double off[4][2] = {{-quarter, quarter}, {quarter, quarter},
                                {-quarter, -quarter}, {quarter, -quarter}};
      for (int q = 0; q < 4; q++) {
        nd->children[q] = qt_node(t, nd->cx + off[q][0], nd->cy + off[q][1], quarter);
        if (!nd->children[q])
          return -1;
      }
      /* the range's next digits are sorted: split at the first of each */
      size_t bound[5];
      bound[0] = it.a;
      bound[4] = it.b;
      const int level = it.depth + 1;
      for (int q = 1; q < 4; q++) {
        size_t lo = bound[q - 1], hi = it.b;
        while (lo < hi) {
          size_t mid = lo + (hi - lo) / 2;
          if (qt_digit(keyHi[mid], keyLo[mid], level) < q)
            lo = mid + 1;
          else
            hi = mid;
        }
        bound[q] = lo;
      }
      for (int q = 3; q >= 0; q--)
        stack[top++] = (qt_item){nd->children[q], bound[q], bound[q + 1], level};
    } else if (cnt > 0) {
      if (qt_fill_leaf(t, nd, order, it.a, it.b) != 0)
        return -1;
    }
  }
  /* :112-123: each cell's running mean over its points in index order */
  const double *pts = t->points;
  const double *ms = t->masses;
  for (size_t i = 0; i < n; i++) {
    double px = pts[2 * i], py = pts[2 * i + 1], m = ms[i];
    gvizQuadtreeNode *nd = r;
    for (;;) {
      double nm = nd->mass + m;
      if (nm > QT_MIN_MASS || nm < -QT_MIN_MASS) {
        nd->comX = (nd->comX * nd->mass + px * m) / nm;
        nd->comY = (nd->comY * nd->mass + py * m) / nm;
      }
      nd->mass = nm;
      if (!nd->children[0])
        break;
      int east = px >= nd->cx, north = py >= nd->cy;
      nd = nd->children[north ? (east ? GVIZ_QUAD_NE : GVIZ_QUAD_NW)
                              : (east ? GVIZ_QUAD_SE : GVIZ_QUAD_SW)];
    }
  }
  t->root = r;
  return 0;
}

int gviz_quadtree_prepare(gvizQuadtree *t, size_t nodesPerCell) {
  memset(t, 0, sizeof(*t));
  t->nodesPerCell = nodesPerCell;
  if (gvizArrayInit(&t->nodeBlocks, sizeof(gvizQuadtreeNode *)) != 0)
    return -1;
  if (gvizArrayInit(&t->pointBlocks, sizeof(size_t *)) != 0 ||
      gvizArrayInit(&t->overflowBuffers, sizeof(size_t *)) != 0)
    return -1;
  return 0;
}

void gviz_quadtree_set_refresh(gvizQuadtree *t, int (*refresh)(void *), void *owner) {
  qt_side *x = qt_side_of(t, 1);
  if (!x)
    return;
  x->refresh = refresh;
  x->owner = owner;
  x->stale = 1;
}

void gviz_quadtree_mark_stale(gvizQuadtree *t) {
  qt_side *x = qt_side_of(t, 0);
  if (x)
    x->stale = 1;
}

const char *gviz_quadtree_last_error(void) { return qt_error; }

static int qt_build(gvizQuadtree *t) {
  size_t n = t->pointCount;
  t->root = NULL;
  t->nodesUsed = 0;
  qt_free_overflow(t);
  if (n == 0)
    return 0;
  qt_side *x = qt_side_of(t, 1);
  if (!x)
    return -1;
  if (!x->dev && gvizhip_qtree_create(&x->dev) != 0) {
    x->dev = NULL;
    qt_error = gvizhip_last_error();
    return -1;
  }
  uint32_t *order = malloc(sizeof(uint32_t) * n);
  uint64_t *hi = malloc(sizeof(uint64_t) * n), *lo = malloc(sizeof(uint64_t) * n);
  double root[4];
  int rc = -1;
  if (order && hi && lo &&
      gvizhip_qtree_build(x->dev, (int64_t)n, t->points, root, order, hi, lo) == 0)
    rc = gviz_quadtree_from_keys(t, root, order, hi, lo);
  else if (order && hi && lo)
    qt_error = gvizhip_last_error();
  free(order);
  free(hi);
  free(lo);
  if (rc != 0)
    t->root = NULL;
  return rc;
}

/* src/ds/gvizQuadtree.c:200-225 */
int gvizQuadtreeInit(gvizQuadtree *tree, const double *points, const double *masses,
                     size_t count, size_t nodesPerCell) {
  if (!tree)
    return -1;
  if (gviz_quadtree_prepare(tree, nodesPerCell) != 0)
    return -1;
  tree->points = points;
  tree->masses = masses;
  tree->pointCount = count;
  return qt_build(tree);
}

/* :227-238 */
int gvizQuadtreeRebuild(gvizQuadtree *tree, const double *points, const double *masses,
                        size_t count) {
  if (!tree)
    return -1;
  tree->points = points;
  tree->masses = masses;
  tree->pointCount = count;
  return qt_build(tree);
}

void gvizQuadtreeRelease(gvizQuadtree *tree) {
  if (!tree)
    return;
  qt_side *x = (qt_side *)gviz_side_take(tree);
  if (x) {
    if (x->dev)
      gvizhip_qtree_destroy(x->dev);
    free(x);
  }
  for (size_t i = 0; i < tree->nodeBlocks.count; i++)
    free(*(gvizQuadtreeNode **)gvizArrayAtIndex(&tree->nodeBlocks, i));
  for (size_t i = 0; i < tree->pointBlocks.count; i++)
    free(*(size_t **)gvizArrayAtIndex(&tree->pointBlocks, i));
  qt_free_overflow(tree);
  gvizArrayRelease(&tree->nodeBlocks);
  gvizArrayRelease(&tree->pointBlocks);
  gvizArrayRelease(&tree->overflowBuffers);
  tree->root = NULL;
  tree->nodesUsed = 0;
}

const gvizQuadtreeNode *gvizQuadtreeRoot(const gvizQuadtree *tree) {
  if (!tree)
    return NULL;
  qt_side *x = qt_side_of(tree, 0);
  if (x && x->refresh && x->stale && x->refresh(x->owner) == 0)
    x->stale = 0;
  return tree->root;
}

int gvizQuadtreeNodeIsLeaf(const gvizQuadtreeNode *node) {
  return node->children[0] == NULL;
}

const gvizQuadtreeNode *gvizQuadtreeNodeChild(const gvizQuadtreeNode *node,
                                              gvizQuadrant quadrant) {
  return node->children[quadrant];
}

size_t gvizQuadtreeNodePointCount(const gvizQuadtreeNode *node) {
  return gvizQuadtreeNodeIsLeaf(node) ? node->pointCount : 0;
}

size_t gvizQuadtreeNodePointAt(const gvizQuadtreeNode *node, size_t i) {
  return node->pointIndices[i];
}

void gvizQuadtreeNodeCenterOfMass(const gvizQuadtreeNode *node, double *outX,
                                  double *outY) {
  if (outX)
    *outX = node->comX;
  if (outY)
    *outY = node->comY;
}
