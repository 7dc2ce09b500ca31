/*
 * gviz_hip.h -- C ABI of the gfx950 layout engine (libgvizhip.so).
 *
 * Plain pointers and sizes only. A caller (the C11 host layer in
 * grapher_amd/host, or the reference's own embedders patched as described in
 * INTEGRATION.md) flattens its graph once into CSR over compact vertex ids,
 * keeps the adjacency in insertion order, and drives the device through these
 * entry points. Every function returns 0 (or a count) on success and a
 * negative value on failure; gvizhip_last_error() describes the failure.
 * Nothing here falls back to the CPU: without a usable GPU the create calls
 * fail.
 */
#ifndef GVIZ_HIP_H
#define GVIZ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Human-readable reason for the last failure on this thread. */
const char *gvizhip_last_error(void);
/* Number of visible GPUs (0 when HIP finds none). */
int gvizhip_device_count(void);

/* Arithmetic of the force math. FP64 reproduces the reference's FP64 formulas
 * in its evaluation order; MIXED differences FP64 positions and does the
 * force math and accumulation in FP32; FP32 is FP32 throughout. The quadtree
 * keys, mass sums, centre-of-mass sums, FA2 speed update and position update
 * are FP64 in every mode. */
enum {
  GVIZHIP_PRECISION_FP32 = 0,
  GVIZHIP_PRECISION_FP64 = 1,
  GVIZHIP_PRECISION_MIXED = 2
};

/* ------------------------------------------------------------------ force */
/* Mirrors the configuration fields of gvizForceEmbedderState
 * (include/embedders/gvizForceEmbedder.h:42-138) that the step reads. */
typedef struct gvizhip_force_config {
  int32_t model;          /* gvizForceModelKind: 0 FR, 1 LinLog */
  int32_t barnesHut;      /* barnesHutEnabled */
  int32_t preventOverlap; /* preventOverlap */
  int32_t precision;      /* GVIZHIP_PRECISION_* */
  int64_t nodesPerCell;
  double edgeLength;
  double theta;
  double gravityK;
  double overlapConstant;
  double radiusBase;
  double radiusPerDegree;
  double jitterTolerance;
} gvizhip_force_config;

/* What gvizForceEmbedderStep reports: its return value and the five stat
 * series it appends (src/embedders/gvizForceEmbedder.c:467-494), plus the
 * FA2 state after the step. */
typedef struct gvizhip_force_stats {
  double maxDisp;
  double speed;
  double meanAtt;
  double meanRep;
  double meanGrav;
  double speedEfficiency;
  double totalSwinging;
  double totalTraction;
} gvizhip_force_stats;

typedef struct gvizhip_force gvizhip_force;

/* Replaces the allocation and setup half of gvizForceEmbedderInit
 * (src/embedders/gvizForceEmbedder.c:219-325): n compact vertices, CSR
 * offsets[n+1] / adj[offsets[n]] over compact ids in insertion order. Masses
 * and degrees follow from the adjacency and cfg->model. */
int gvizhip_force_create(gvizhip_force **out, int64_t n, const int64_t *offsets,
                         const int64_t *adj, const gvizhip_force_config *cfg);
/* Replaces the Configure* setters (:357-410): takes effect at the next step. */
int gvizhip_force_configure(gvizhip_force *f, const gvizhip_force_config *cfg);
/* Replaces gvizForceEmbedderBegin (:412-443) after positions are chosen:
 * uploads positions[2n] (compact order, interleaved x,y), speed 1, efficiency
 * 1, forces zero, iteration 0. */
int gvizhip_force_begin(gvizhip_force *f, const double *positions);
/* Positions in and out, compact order (the reference's gatherPositions and
 * AddVPosition ends of a step, :19-26 and :476). */
int gvizhip_force_upload_positions(gvizhip_force *f, const double *positions);
int gvizhip_force_download_positions(gvizhip_force *f, double *positions);
/* Replaces gvizForceEmbedderStep (:445-495) on device-resident positions. */
int gvizhip_force_step(gvizhip_force *f, gvizhip_force_stats *stats);
/* Replaces gvizForceEmbedderRun (:497-510): steps until maxDisp < epsilon or
 * maxIters steps; returns the number of steps. stats (maxIters entries, may be
 * NULL) receives every step's record. */
int64_t gvizhip_force_run(gvizhip_force *f, int64_t maxIters, double epsilon,
                          gvizhip_force_stats *stats);
/* One force evaluation at the current positions without applying it (the
 * outputs of computeForceRange, :76-126); the FA2 state is left untouched.
 * Any output may be NULL. interactions: repulsion evaluations per vertex. */
int gvizhip_force_evaluate(gvizhip_force *f, double *disp, double *structForce,
                           double *attMag, double *repMag,
                           int32_t *interactions);
/* FA2 state: global speed, speed efficiency, iteration. */
int gvizhip_force_get_speed(gvizhip_force *f, double *globalSpeed,
                            double *speedEfficiency, int64_t *iteration);
int gvizhip_force_set_speed(gvizhip_force *f, double globalSpeed,
                            double speedEfficiency);
/* Exact checkpoint of the embedder's evolving state: FP64 positions, the
 * previous step's structForce (in the engine's precision), global speed,
 * speed efficiency and the step count. No reference counterpart: the
 * reference's checkpoint (src/embedders/gvizEmbeddedGraph.c:428-485) keeps
 * positions as "%f" text and no speed state, so a run resumed from it
 * differs; one resumed from this state continues bit for bit. `bytes` is
 * gvizhip_force_state_bytes(); load takes a state of an engine with the same
 * vertex count (structForce saved in the other precision is converted). */
int64_t gvizhip_force_state_bytes(gvizhip_force *f);
int gvizhip_force_save_state(gvizhip_force *f, void *buf, int64_t bytes);
int gvizhip_force_load_state(gvizhip_force *f, const void *buf, int64_t bytes);
/* The quadtree of the current positions (gvizQuadtreeBuild,
 * src/ds/gvizQuadtree.c:159-192): per vertex the depth and quadrant-path key
 * of its leaf; root[4] = centre x, centre y, half size, depth limit. */
int gvizhip_force_quadtree_leaves(gvizhip_force *f, int32_t *leafDepth,
                                  uint64_t *leafKeyHi, uint64_t *leafKeyLo,
                                  double *root);
/* Branching cells of that quadtree (cells with two or more non-empty
 * children): depth, path key, mass, centre of mass, point count. Returns the
 * number of cells (entries beyond cap are not written). */
int64_t gvizhip_force_quadtree_cells(gvizhip_force *f, int64_t cap,
                                     int32_t *depth, uint64_t *keyHi,
                                     uint64_t *keyLo, double *mass,
                                     double *comX, double *comY,
                                     int64_t *count);
/* The tree behind the host view of gvizForceEmbedderState.quadtree
 * (src/embedders/gvizForceEmbedder.c:430-452 rebuild it from the positions
 * at the start of each Step): the tree of the last step's starting
 * positions, or of the current positions when none was built since they
 * last changed. root[4] = centre x, centre y, half size, depth limit;
 * order[n] = compact vertex ids in key order; keyHi/keyLo[n] = their
 * quadrant-path keys (levels 1-32, 33-64) in that order; points[2n] (may be
 * NULL) = the positions the tree was built from, in compact order. */
int gvizhip_force_quadtree_keys(gvizhip_force *f, double *root, uint32_t *order,
                                uint64_t *keyHi, uint64_t *keyLo, double *points);
/* Measurement hook (no reference counterpart): runs `steps` steps outside
 * the captured graph with HIP events on the engine's stream around the
 * repulsion kernel (Barnes-Hut traversal, or all-pairs when Barnes-Hut is
 * off) and around the whole step. Outputs (may be NULL): summed repulsion
 * kernel time, summed repulsion evaluations over all vertices and steps,
 * summed step time. */
int gvizhip_force_profile(gvizhip_force *f, int64_t steps, double *traversalMs,
                          int64_t *interactions, double *stepMs);
void gvizhip_force_destroy(gvizhip_force *f);

/* ------------------------------------------------------------- Quadtree */
/* The device half of gvizQuadtreeInit / gvizQuadtreeRebuild
 * (src/ds/gvizQuadtree.c:159-238): the root cell (bounding-box centre, half
 * the larger span or 1.0, the depth at which the half size first falls to
 * <= 1e-9) and every point's quadrant path by the reference's FP64 centre
 * recurrence, sorted by (path, id). The host turns the sorted paths into
 * the reference's node records. One context serves any number of builds
 * (its buffers grow to the largest). */
typedef struct gvizhip_qtree gvizhip_qtree;
int gvizhip_qtree_create(gvizhip_qtree **out);
/* count points (x, y interleaved, FP64; count < 2^32). Outputs as
 * gvizhip_force_quadtree_keys. 0 or -1. */
int gvizhip_qtree_build(gvizhip_qtree *q, int64_t count, const double *points,
                        double *root, uint32_t *order, uint64_t *keyHi,
                        uint64_t *keyLo);
void gvizhip_qtree_destroy(gvizhip_qtree *q);

/* ------------------------------------------------------------------ Tutte */
typedef struct gvizhip_tutte gvizhip_tutte;

/* Graph over parent ids 0..n-1 (CSR in insertion order, subgraph edges only)
 * for the Jacobi sweep of gvizTutteEmbedderStep
 * (src/embedders/gvizTutteEmbedder.c:305-335). */
int gvizhip_tutte_create(gvizhip_tutte **out, int64_t n, const int64_t *offsets,
                         const int64_t *adj, int32_t precision);
/* Live positions (2n) and the pinned-boundary flags (n bytes). */
int gvizhip_tutte_upload(gvizhip_tutte *t, const double *positions,
                         const uint8_t *isBoundary);
int gvizhip_tutte_download(gvizhip_tutte *t, double *positions);
/* `iters` Jacobi sweeps with relaxation alpha (already clamped to (0, 1]).
 * Stops after the first sweep whose max displacement is < epsilon when
 * epsilon >= 0. maxDeltas (iters entries, may be NULL) receives each sweep's
 * max displacement. Returns the number of sweeps run. */
int64_t gvizhip_tutte_sweeps(gvizhip_tutte *t, int64_t iters, double alpha,
                             double epsilon, double *maxDeltas);
/* Gauss-Seidel sweeps (gvizTutteEmbedderStep with useGaussSeidel,
 * src/embedders/gvizTutteEmbedder.c:258-262, 305-320): interior vertices
 * relaxed in ascending id, each from the current positions; the same
 * arguments, stop rule and records as gvizhip_tutte_sweeps. The schedule
 * (vertices grouped so that no two of a group are adjacent and every
 * lower-id neighbour sits in an earlier group) is built on the first call
 * after an upload changed the boundary. */
int64_t gvizhip_tutte_gs_sweeps(gvizhip_tutte *t, int64_t iters, double alpha,
                                double epsilon, double *maxDeltas);
/* Spring variant (replaces the loop of gvizSpringTutteEmbedderStep / Run,
 * src/embedders/gvizSpringTutteEmbedder.c:313-358): `iters` steps of
 * a = stiffness (barycentre - p) - damping v, v += a dt, p += v dt for every
 * interior vertex with neighbours, on the same positions and boundary as
 * the sweeps; the velocity (n x 2, caller's ids, zero after create) is set
 * with gvizhip_tutte_set_velocity (NULL: zero) and read with
 * gvizhip_tutte_get_velocity. Stops after the first step whose max delta is
 * below epsilon (epsilon >= 0); returns the steps taken, maxDeltas
 * (optional) receives each one's max delta. */
int64_t gvizhip_tutte_spring_steps(gvizhip_tutte *t, int64_t iters, double dt,
                                   double stiffness, double damping, double epsilon,
                                   double *maxDeltas);
int gvizhip_tutte_set_velocity(gvizhip_tutte *t, const double *velocity);
int gvizhip_tutte_get_velocity(gvizhip_tutte *t, double *velocity);
/* Measurement hook (no reference counterpart): `sweeps` sweeps outside the
 * captured graph, HIP events on the engine's stream around each sweep
 * kernel. Outputs (may be NULL): summed sweep-kernel time, total time. */
int gvizhip_tutte_profile(gvizhip_tutte *t, int64_t sweeps, double alpha,
                          double *sweepMs, double *totalMs);
void gvizhip_tutte_destroy(gvizhip_tutte *t);
/* No counterpart in the reference: an entry point that allocates nothing,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). The FP64 Jacobi
 * sweeps of gvizhip_tutte_sweeps on caller-owned device memory: offsets[n+1]
 * and adj (int32 CSR in insertion order), isBoundary[n] bytes, positions
 * (x, y interleaved) ping-ponged between posA and posB, every call starting
 * from posA; the result is always left in posA. Stops after the first sweep
 * whose max displacement is < epsilon when epsilon >= 0. maxDeltas (device,
 * iters doubles) receives each sweep's max displacement (entries past the
 * last sweep run are untouched); sweepsRun (device int64, may be NULL) the
 * count. Everything is enqueued on `stream` (a hipStream_t, NULL: the null
 * stream) and nothing waits for it. Per-block maxima and the step scalars
 * live in a 16 KiB static device array shared by every call on a GPU: calls
 * that overlap on one GPU race. n is limited to 2^31 - 2 (32-bit indices).
 * Results equal gvizhip_tutte_sweeps's bit for bit. 0 or -1. */
int gvizhip_tutte_sweeps_noalloc(int64_t n, const int32_t *offsets, const int32_t *adj,
                                 const uint8_t *isBoundary, double *posA, double *posB,
                                 int64_t iters, double alpha, double epsilon,
                                 double *maxDeltas, int64_t *sweepsRun, void *stream);

/* ------------------------------------------------------------------- GRIP */
typedef struct gvizhip_grip gvizhip_grip;

/* Graph over parent ids 0..n-1 restricted to the subgraph (CSR in insertion
 * order); active[n] marks subgraph vertices (NULL: all); kcap = kNN capacity
 * per vertex (gvizGRIPEmbedderConfigureKnnCapacity). Replaces the
 * allocations of gvizGRIPEmbedderInit (src/embedders/gvizGRIPEmbedder.c:
 * 194-320). */
int gvizhip_grip_create(gvizhip_grip **out, int64_t n, const int64_t *offsets,
                        const int64_t *adj, const uint8_t *active,
                        int64_t kcap);
/* createMISFiltration (:741-777) for `dim` dimensions: writes
 * filtration[activeCount] and borders[layers] (room for 128) and keeps the
 * filtration on the device; returns the layer count. */
int64_t gvizhip_grip_mis_filtration(gvizhip_grip *g, int64_t dim,
                                    int64_t *filtration, int64_t *borders);
/* The stages of that filtration, for the reference's white-box entry points
 * (src/embedders/gvizGRIPInternal.h:22-40).
 * makeFirstMISPartition (:450-483): mis[n] = 1 on the lexicographically-
 * first MIS of the subgraph; tail[activeCount - misCount] = the other
 * subgraph vertices in the order the reference writes them backwards from
 * the end of misFiltration (tail[r] is misFiltration[activeCount-1-r]).
 * Returns misCount, or -1. */
int64_t gvizhip_grip_mis_first(gvizhip_grip *g, uint8_t *mis, int64_t *tail);
/* iterMISFiltration's selection (:672-739) for layer >= 2: of the
 * candidates cand[count] (ascending ids), picked[j] = 1 on the
 * lexicographically-first set whose members lie more than 2^(layer-1) BFS
 * steps apart through the whole subgraph. Returns the picked count. */
int64_t gvizhip_grip_mis_layer(gvizhip_grip *g, int64_t layer, const int64_t *cand,
                               int64_t count, uint8_t *picked);
/* migrateOneToFinalLayer (:576-597) on a filtration[activeCount] and its
 * borders[layerCount] (borders[0] = activeCount): moves the candidate
 * farthest from the final set (layer count-1) into it. 0, or -1 when every
 * pool is exhausted (arrays unchanged). */
int gvizhip_grip_migrate(gvizhip_grip *g, int64_t *filtration, int64_t filtrationCount,
                         int64_t *borders, int64_t layerCount, int64_t count);
/* Installs a filtration computed elsewhere (count entries). */
int gvizhip_grip_set_filtration(gvizhip_grip *g, const int64_t *filtration,
                                int64_t count);
/* gvizSearchKNearestScratch (src/algorithms/search/gvizKNearest.c:244-306)
 * from each of sources[count] over the `visible` filter (n bytes, NULL: all):
 * out*[count*k], counts[count]. */
int gvizhip_grip_knn(gvizhip_grip *g, const uint8_t *visible,
                     const int64_t *sources, int64_t count, int64_t k,
                     int64_t *outVertex, int64_t *outDist, int64_t *counts);
/* gvizSearchKNearestFromVisibleBatch (:168-210): one BFS per visible seed in
 * ascending id, ties kept in seed order, early stop once every target holds
 * k entries. */
int gvizhip_grip_knn_batch(gvizhip_grip *g, const uint8_t *visible,
                           const int64_t *targets, int64_t count, int64_t k,
                           int64_t *outVertex, int64_t *outDist,
                           int64_t *counts);
/* Per-vertex GRIP state (gvizGRIPDecorators without the kNN array, plus the
 * dispCalculated bit). */
typedef struct gvizhip_grip_vertex {
  double dispX, dispY, oldDispX, oldDispY;
  double heat, oldCos, lastForceMag;
  int64_t knnCount;
  int64_t dispCalculated;
} gvizhip_grip_vertex;
typedef struct gvizhip_grip_round_stats {
  double maxDisp, meanDisp, meanForce, meanHeat;
} gvizhip_grip_round_stats;
/* Host <-> device copies of positions (n*dim, 2n by default), vertex state (n) and kNN lists
 * (n*kcap vertex and dist entries); any pointer may be NULL. */
int gvizhip_grip_upload(gvizhip_grip *g, const double *positions,
                        const gvizhip_grip_vertex *dec,
                        const int64_t *knnVertex, const int64_t *knnDist);
int gvizhip_grip_download(gvizhip_grip *g, double *positions,
                          gvizhip_grip_vertex *dec, int64_t *knnVertex,
                          int64_t *knnDist);
/* Layouts of 3 or 4 coordinates (gvizGRIPEmbedderInit's dimension,
 * src/embedders/gvizGRIPEmbedder.c:1192-1261 for the rotation removal of
 * each): call once after create, before any upload. Positions passed to
 * upload/download are then n*dim doubles; the per-vertex state goes
 * through the _nd pair, whose disp/oldDisp hold dim coordinates (the rest
 * zero). The _nd pair also serves dim 2. 0 or -1. */
int gvizhip_grip_set_dimension(gvizhip_grip *g, int64_t dim);
typedef struct gvizhip_grip_vertex_nd {
  double disp[4], oldDisp[4];
  double heat, oldCos, lastForceMag;
  int64_t knnCount;
  int64_t dispCalculated;
} gvizhip_grip_vertex_nd;
int gvizhip_grip_upload_nd(gvizhip_grip *g, const double *positions,
                           const gvizhip_grip_vertex_nd *dec,
                           const int64_t *knnVertex, const int64_t *knnDist);
int gvizhip_grip_download_nd(gvizhip_grip *g, double *positions,
                             gvizhip_grip_vertex_nd *dec, int64_t *knnVertex,
                             int64_t *knnDist);
/* clearDecorators (:936-952) for filtration[0..active). */
int gvizhip_grip_clear_decorators(gvizhip_grip *g, int64_t active);
/* updateKNNs (:954-1006): kNN lists of filtration[0..active) against the
 * visible set, per-vertex BFS or (batch != 0) the batched search. */
int gvizhip_grip_update_knn(gvizhip_grip *g, int64_t active,
                            const uint8_t *visible, int64_t k, int32_t batch);
/* placeLayerVertices (:816-868): filtration[begin..end) placed at the
 * barycentre of their k visible BFS-nearest vertices. */
int gvizhip_grip_place(gvizhip_grip *g, int64_t begin, int64_t end,
                       const uint8_t *visible, int64_t k, int32_t batch);
/* gvizGRIPEmbedderRefineRound (:1289-1333) over filtration[0..active). */
int gvizhip_grip_refine_round(gvizhip_grip *g, int64_t active, int64_t layer,
                              gvizhip_grip_round_stats *stats);
/* Measurement hook (no reference counterpart): `rounds` refinement rounds
 * as gvizhip_grip_refine_round, with HIP events on the engine's stream
 * around the spring-force kernel and around each whole round. Outputs (may
 * be NULL): summed force-kernel time, summed round time. */
int gvizhip_grip_profile_rounds(gvizhip_grip *g, int64_t active, int64_t layer,
                                int64_t rounds, double *forceMs,
                                double *roundMs);
void gvizhip_grip_destroy(gvizhip_grip *g);

/* ------------------------------------------------------------------ batch */
/* No reference counterpart (the reference lays out one graph per embedder
 * state): up to 4096 independent graphs laid out together. Graph g owns
 * vertices [vertexStart[g], vertexStart[g+1]) of one CSR (offsets[n+1], adj
 * over those global ids; no edge may leave its graph). Every stage of
 * gvizForceEmbedderStep runs once for the whole batch; each graph follows
 * exactly the trajectory gvizhip_force gives it alone from the same start
 * (same keys, tree, interaction sets and summation order; the direction
 * between exactly coincident points hashes batch-wide ids instead).
 * Barnes-Hut only; a graph may need at most 58 quadtree levels. */
typedef struct gvizhip_batch gvizhip_batch;
int gvizhip_batch_create(gvizhip_batch **out, int64_t graphs,
                         const int64_t *vertexStart, const int64_t *offsets,
                         const int64_t *adj, const gvizhip_force_config *cfg);
/* Positions (2n) for every graph; speed 1, efficiency 1, forces zero. */
int gvizhip_batch_begin(gvizhip_batch *b, const double *positions);
/* Schedule knobs that never change results: traversal wave order (0: graph
 * order, 1: largest graphs first), steps per captured launch (1..64) and
 * traversal wavefronts per workgroup (1, 2 or 4; the waves never
 * synchronise, so this only changes how they are packed onto the CUs). */
int gvizhip_batch_schedule(gvizhip_batch *b, int32_t waveOrder,
                           int32_t stepsPerLaunch, int32_t traversalWaves);
/* Times every schedule (2 x 3 x 3) on `trialSteps` steps from the current
 * state, puts the state back and keeps the fastest; outputs may be NULL. */
int gvizhip_batch_autotune(gvizhip_batch *b, int64_t trialSteps,
                           int32_t *waveOrder, int32_t *stepsPerLaunch,
                           int32_t *traversalWaves);
/* Steps every graph until its maxDisp < epsilon (epsilon >= 0) or maxIters
 * steps. Per graph (any may be NULL): steps taken, milliseconds from the
 * call to the check that saw it stop (the call's total when it ran out of
 * steps), last maxDisp. Returns the steps launched. */
int64_t gvizhip_batch_run(gvizhip_batch *b, int64_t maxIters, double epsilon,
                          int64_t *iterations, double *latencyMs,
                          double *maxDisp);
/* Measurement hook: `steps` uncaptured steps with HIP events around the
 * traversal kernel and the whole step; summed times and repulsion
 * evaluations (outputs may be NULL). */
int gvizhip_batch_profile(gvizhip_batch *b, int64_t steps, double *traversalMs,
                          int64_t *interactions, double *stepMs);
int gvizhip_batch_download_positions(gvizhip_batch *b, double *positions);
void gvizhip_batch_destroy(gvizhip_batch *b);

/* ---------------------------------------------------------------- metrics */
/* No reference counterpart: layout quality for judging runs. Normalised
 * stress over BFS hop distances from `count` pivot vertices:
 * 1 - A^2 / (B C), A = sum D/d, B = sum D^2/d^2, C = pairs, over every pivot
 * s and every vertex j it reaches (D = |x_j - x_s|, d = hops). Distances are
 * computed once at create. */
typedef struct gvizhip_stress gvizhip_stress;
int gvizhip_stress_create(gvizhip_stress **out, int64_t n, const int64_t *offsets,
                          const int64_t *adj, const int64_t *pivots,
                          int64_t count);
int gvizhip_stress_evaluate(gvizhip_stress *s, const double *positions,
                            double *stress);
void gvizhip_stress_destroy(gvizhip_stress *s);

/* -------------------------------------------------------- learned start */
/* No reference counterpart: a message-passing network whose output replaces
 * the random positions of gvizForceEmbedderBegin
 * (src/embedders/gvizForceEmbedder.c:412-443) as the force embedder's start.
 * Three layers h' = relu([h | mean of neighbours' h] W + b), hidden width
 * 64, BF16 weights and activations on the matrix cores with FP32
 * accumulation, a linear head to 2 coordinates. Training fits layouts up to
 * rotation, reflection and scale (Procrustes-aligned squared error divided
 * by |T|^2). Parameters are one FP32 vector of gvizhip_gnn_param_count()
 * entries: W1 (32x64), b1, W2 (128x64), b2, W3 (128x64), b3, W4 (64x16),
 * b4 (16); rows of W are inputs, columns outputs, row-major. */
typedef struct gvizhip_gnn gvizhip_gnn;
int gvizhip_gnn_create(gvizhip_gnn **out, int64_t maxVertices,
                       int64_t maxEntries);
int64_t gvizhip_gnn_param_count(void);
int gvizhip_gnn_set_params(gvizhip_gnn *g, const float *params);
int gvizhip_gnn_get_params(gvizhip_gnn *g, float *params);
/* Graph (CSR over 0..n-1, symmetric) and input features (n x 16 FP32, row
 * major; rounded to BF16). */
int gvizhip_gnn_set_graph(gvizhip_gnn *g, int64_t n, const int64_t *offsets,
                          const int64_t *adj, const float *features);
/* The same graph upload with the features built on the device from a start
 * layout (n x 2): columns 0-1 the start, 2-3 / 4-5 / 6-7 / 8-9 its Chebyshev
 * low-pass filters T_d(mu(P)) x of degree 2 / 4 / 8 / 16, where P is the
 * lazy random walk x <- (x + mean of the neighbours' x) / 2 and
 * mu(P) = (2 P - 0.9 I) / 0.9 (recurrence y_1 = mu y_0,
 * y_{k+1} = 2 mu y_k - y_{k-1}), each pair centred and scaled to unit
 * variance per coordinate; column 10 log(1 + degree) / 4, column 11 one, the
 * rest 0. */
int gvizhip_gnn_set_graph_diffused(gvizhip_gnn *g, int64_t n,
                                   const int64_t *offsets, const int64_t *adj,
                                   const double *start);
/* Rebuilds the features of the graph set last from a new start (n x 2), as
 * gvizhip_gnn_set_graph_diffused does, without uploading the graph again;
 * removes the pivot-MDS skip term. */
int gvizhip_gnn_set_start(gvizhip_gnn *g, const double *start);
/* The same from a start drawn on the device: vertex v's coordinates are
 * bits 40-63 and 16-39 of SplitMix64(seed + (v + 1) * 0x9E3779B97F4A7C15),
 * each scaled to [-1, 1) as 2 b / 2^24 - 1. */
int gvizhip_gnn_set_start_seeded(gvizhip_gnn *g, uint64_t seed);
/* A proposed start for the graph set last, with no host work between the
 * steps: seeded start and features as gvizhip_gnn_set_start_seeded, the
 * forward pass, the two outputs centred and scaled to RMS radius `radius`
 * in FP64, one copy to positions (n x 2). */
int gvizhip_gnn_propose(gvizhip_gnn *g, uint64_t seed, double radius,
                        double *positions);
/* Pivot MDS (Brandes & Pich) of the graph set last: BFS hop distances to
 * `count` (<= 64) pivot vertices, D^2 double-centred into C, axes C v / mu^(1/4)
 * for the top two eigenpairs (v, mu) of C^T C, centred and scaled to RMS
 * radius 1. Written to feature columns 12-13 and kept as a skip term added
 * to the head's two outputs by every later forward pass (count 0 removes
 * it; setting a graph removes it). coords (n x 2, optional) receives them.
 * No reference counterpart: part of the warm start in front of
 * gvizForceEmbedderBegin (src/embedders/gvizForceEmbedder.c:412-443). */
int gvizhip_gnn_set_pivots(gvizhip_gnn *g, const int64_t *pivots, int64_t count,
                           double *coords);
/* Network output (n x 2, interleaved) for the current graph. */
int gvizhip_gnn_forward(gvizhip_gnn *g, double *positions);
/* Loss against target (n x 2) and the gradient of every parameter. */
int gvizhip_gnn_gradients(gvizhip_gnn *g, const double *target, double *loss,
                          float *grads);
/* One Adam step (beta 0.9 / 0.999) on the loss against target. */
int gvizhip_gnn_train_step(gvizhip_gnn *g, const double *target, double lr,
                           double *loss);
void gvizhip_gnn_destroy(gvizhip_gnn *g);

#ifdef __cplusplus
}
#endif
#endif
