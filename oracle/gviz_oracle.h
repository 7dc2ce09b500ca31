/*
 * gviz_oracle.h -- FP64 CPU restatement of the reference's layout hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it. Nothing in grapher_amd/ links or calls it.
 *
 * Every routine restates a reference routine (cited per function in
 * gviz_oracle.c) over plain arrays:
 *   - vertices are compact indices 0..n-1 (the reference's state->vertices[]
 *     order, ascending parent id);
 *   - the graph is CSR over compact indices that keeps each vertex's
 *     adjacency insertion order (attraction order, BFS discovery order);
 *   - positions are interleaved x0,y0,x1,y1,... doubles.
 */
#ifndef GVIZ_ORACLE_H
#define GVIZ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- pool -- */
/* Worker pool with the reference's ForRange shape (caller + helpers pulling
 * grain-sized chunks from an atomic counter). threads == 0: online CPUs. */
typedef struct oc_pool oc_pool;
oc_pool *oc_pool_create(int threads);
void oc_pool_destroy(oc_pool *p);
int oc_pool_threads(const oc_pool *p);
typedef void (*oc_range_fn)(void *ctx, int64_t begin, int64_t end);
void oc_pool_for_range(oc_pool *p, int64_t begin, int64_t end, int64_t grain,
                       oc_range_fn fn, void *ctx);

/* ------------------------------------------------------ pairwise forces -- */
/* acc += force on v from u; each mirrors one gvizVecAcc* routine. */
void oc_fr_att(const double *v, const double *u, double k, double *acc);
void oc_linlog_att(const double *v, const double *u, double *acc);
void oc_fr_rep(const double *v, const double *u, double k, double radiusSum,
               double overlapConstant, int64_t selfId, int64_t otherId,
               double *acc);
void oc_linlog_rep(const double *v, const double *u, double vMass,
                   double otherMass, double radiusSum, double overlapConstant,
                   int64_t selfId, int64_t otherId, double *acc);
void oc_gravity(const double *v, double magnitude, double *acc);
void oc_kk(const double *v, const double *u, double L, double *acc);
void oc_grip_fr_rep(const double *v, const double *u, double edgeLength,
                    double *acc);
void oc_grip_fr_att(const double *v, const double *u, double edgeLength,
                    double frScale, double *acc);
/* Direction used when two points coincide exactly (the reference hashes
 * buffer addresses there, which no other program can reproduce). */
void oc_coincident_dir(int64_t selfId, int64_t otherId, double *ux,
                       double *uy);

/* ------------------------------------------------------------ quadtree -- */
typedef struct oc_qnode {
  double cx, cy, halfSize;
  double comX, comY, mass;
  int32_t child[4];   /* -1 on leaves */
  int32_t depth;
  int64_t ptStart;    /* leaf point list: pts[ptStart .. ptStart+ptCount) */
  int64_t ptCount;
  uint64_t keyHi, keyLo; /* quadrant path, 2 bits per level, level 1 = top */
} oc_qnode;

typedef struct oc_quadtree oc_quadtree;
oc_quadtree *oc_quadtree_build(const double *points, const double *masses,
                               int64_t n, int64_t nodesPerCell);
void oc_quadtree_free(oc_quadtree *t);
int64_t oc_quadtree_node_count(const oc_quadtree *t);
const oc_qnode *oc_quadtree_nodes(const oc_quadtree *t);
const int64_t *oc_quadtree_leaf_points(const oc_quadtree *t);
int32_t oc_quadtree_root(const oc_quadtree *t);
/* Root parameters: centre, half size and the depth at which the half size
 * first falls to <= 1e-9 (cells there never split). */
void oc_quadtree_root_params(const oc_quadtree *t, double *cx, double *cy,
                             double *halfSize, int32_t *maxDepth);
/* Per point: depth of the leaf that holds it and that leaf's path key. */
void oc_quadtree_point_leaves(const oc_quadtree *t, int32_t *depth,
                              uint64_t *keyHi, uint64_t *keyLo);

/* Path key of one point by replaying the reference's FP64 centre recurrence
 * from the root for `levels` levels (digits past maxDepth are 0). */
void oc_path_key(double px, double py, double cx, double cy, double halfSize,
                 int32_t maxDepth, uint64_t *keyHi, uint64_t *keyLo);

/* --------------------------------------------------------- force model -- */
enum { OC_MODEL_FR = 0, OC_MODEL_LINLOG = 1 };

typedef struct oc_force_params {
  int32_t model;
  int32_t barnesHut;       /* 1: quadtree walk, 0: exact all-pairs */
  int32_t preventOverlap;
  int32_t pad;
  int64_t nodesPerCell;
  double edgeLength;
  double theta;
  double gravityK;
  double overlapConstant;
  double radiusBase;
  double radiusPerDegree;
  double jitterTolerance;
} oc_force_params;

typedef struct oc_csr {
  int64_t n;
  const int64_t *offsets; /* n+1 */
  const int64_t *adj;     /* offsets[n] entries, compact ids */
} oc_csr;

double oc_vertex_mass(int32_t model, int64_t degree);
double oc_vertex_radius(const oc_force_params *p, int64_t degree);

/* One force evaluation (the reference's computeForceRange over all
 * vertices). Outputs may be NULL. interactions (optional, n entries): number
 * of repulsion evaluations per vertex (leaf pairs + pseudo-bodies). */
void oc_force_compute(const oc_csr *g, const double *pos, const double *mass,
                      const int64_t *degree, const oc_force_params *p,
                      const oc_quadtree *tree, oc_pool *pool, double *disp,
                      double *structForce, double *attMag, double *repMag,
                      int64_t *interactions);

/* The same evaluation, also writing termSum[i] = sum over every term that
 * makes up vertex i's force (each attraction entry, each repulsion leaf point
 * and pseudo-body, gravity) of that term's magnitude: the scale of the
 * per-vertex FP32 error bound |dF_i| <= tol * sum_j |f_ij|. disp and
 * interactions (optional) equal oc_force_compute's bit for bit. */
void oc_force_term_sums(const oc_csr *g, const double *pos, const double *mass,
                        const int64_t *degree, const oc_force_params *p,
                        oc_pool *pool, double *disp, int64_t *interactions,
                        double *termSum);

/* Full-step state (the per-embedder fields Step reads and writes). */
typedef struct oc_force_state {
  int64_t n;
  double globalSpeed;
  double speedEfficiency;
  double meanDegreePlusOne;
  int64_t iteration;
  double *mass;            /* n */
  int64_t *degree;         /* n */
  double *disp;            /* 2n */
  double *structForce;     /* 2n */
  double *oldStructForce;  /* 2n */
  double *appliedDisp;     /* 2n */
  double *attMag, *repMag; /* n */
  double *swinging, *traction; /* n */
} oc_force_state;

typedef struct oc_step_stats {
  double maxDisp, speed, meanAtt, meanRep, meanGrav;
  double speedEfficiency;
  double totalSwinging, totalTraction;
} oc_step_stats;

oc_force_state *oc_force_state_create(const oc_csr *g,
                                      const oc_force_params *p);
void oc_force_state_free(oc_force_state *s);
/* Begin's reset: speed 1, efficiency 1, forces zero, iteration 0. */
void oc_force_state_reset(oc_force_state *s);
/* One gvizForceEmbedderStep: pos (2n) is read and advanced in place. */
double oc_force_step(oc_force_state *s, const oc_csr *g,
                     const oc_force_params *p, double *pos, oc_pool *pool,
                     oc_step_stats *stats);
/* The FA2 scalar update alone (updateGlobalSpeed) on given sums. */
void oc_fa2_update(int64_t n, double jitterTolerance, double totalSwinging,
                   double totalTraction, double *globalSpeed,
                   double *speedEfficiency);

/* ---------------------------------------------------------------- Tutte -- */
/* One gvizTutteEmbedderStep over the parent graph (ids 0..n-1).
 * isBoundary: n bytes. pos: 2n live positions. scratch: 2n. */
double oc_tutte_step(const oc_csr *g, const uint8_t *isBoundary, double *pos,
                     double *scratch, double relaxationRate, double dt,
                     int useGaussSeidel);
/* FixConvexPolygon's positions: radius*(cos,sin)(2*pi*k/count). */
void oc_tutte_polygon(int64_t count, double radius, double *out);

/* -------------------------------------------------------------- metrics -- */
/* Layout stress with the optimal global scale (closed form), over all pairs
 * reachable from the `sources` BFS roots (all vertices when sources == NULL,
 * nsources ignored). Unreachable pairs are skipped. */
double oc_stress(const oc_csr *g, const double *pos, double edgeLength,
                 const int64_t *sources, int64_t nsources, oc_pool *pool);
/* Mean over edges of (|x_u - x_v| / mean_edge_length - 1)^2. */
double oc_edge_energy(const oc_csr *g, const double *pos);

/* ----------------------------------------------------------------- GRIP -- */
typedef struct oc_found { int64_t v; int64_t dist; } oc_found;

/* gvizSearchKNearestScratch: BFS discovery order, filter bytes (NULL = all).
 * Returns the count written to out (<= k) or -1. */
int64_t oc_knn_bfs(const oc_csr *g, int64_t source, int64_t k,
                   const uint8_t *filter, oc_found *out);
/* gvizSearchKNearestFromVisibleBatch: one BFS per visible seed in ascending
 * id; out is targetCount*k; counts per target. */
int oc_knn_batch(const oc_csr *g, const uint8_t *visible, int64_t k,
                 const int64_t *targets, int64_t targetCount, oc_found *out,
                 int64_t *counts);
int oc_knn_prefer_batch(const oc_csr *g, int64_t activeCount, int64_t edgeCount,
                        int64_t visibleCount, int64_t targetCount);

/* MIS filtration over the subgraph given by `active` (n bytes; NULL = all).
 * filtration: n entries (only activeCount used); borders: out array with
 * room for 128 entries. Returns the number of layers. */
int64_t oc_mis_filtration(const oc_csr *g, const uint8_t *active,
                          int64_t dim, int64_t *filtration, int64_t *borders);
/* Only the first greedy partition (layer 1) -- white-box check. */
int64_t oc_mis_first(const oc_csr *g, const uint8_t *active,
                     int64_t *filtration, uint8_t *inMis);

typedef struct oc_grip_vertex {
  double dispX, dispY, oldDispX, oldDispY;
  double heat, oldCos, lastForceMag;
  int64_t knnCount;
  uint8_t dispCalculated;
} oc_grip_vertex;

typedef struct oc_grip_round_stats {
  double maxDisp, meanDisp, meanForce, meanHeat;
} oc_grip_round_stats;

/* One gvizGRIPEmbedderRefineRound over filtration[0..active):
 * knn is n*kcap entries (per vertex, knnCount used). pos: 2n (parent ids). */
void oc_grip_refine_round(const oc_csr *g, const int64_t *filtration,
                          int64_t active, int64_t layer, const oc_found *knn,
                          int64_t kcap, oc_grip_vertex *dec, double *pos,
                          oc_grip_round_stats *stats);

/* The same round for dim = 2, 3 or 4 coordinates (pos: dim*n). dec holds
 * disp / oldDisp of dim coordinates; the 3-D rotation removal solves the
 * inertia system, the 4-D one removes the six coordinate-plane rotations
 * in turn. */
typedef struct oc_grip_vertex_nd {
  double disp[4], oldDisp[4];
  double heat, oldCos, lastForceMag;
  int64_t knnCount;
  int64_t dispCalculated;
} oc_grip_vertex_nd;
void oc_grip_refine_round_nd(const oc_csr *g, const int64_t *filtration,
                             int64_t active, int64_t layer, const oc_found *knn,
                             int64_t kcap, int64_t dim, oc_grip_vertex_nd *dec,
                             double *pos, oc_grip_round_stats *stats);
void oc_regular_simplex(int64_t n, double sideLength, double *out);

#ifdef __cplusplus
}
#endif
#endif
