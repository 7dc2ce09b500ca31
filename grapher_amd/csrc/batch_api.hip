// batch_api.hip -- C ABI of the batched force engine: many independent
// graphs laid out together, every stage of the step launched once for the
// whole batch (keys carry the graph id, reductions and the speed update run
// per graph with the block decomposition a run of that graph alone uses,
// traversal waves never mix graphs), iterations replayed from a captured
// hipGraph. Each graph's trajectory is the one gvizhip_force gives it alone.
#include "../../include/gviz_hip.h"
#include "force_engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <vector>

using namespace gvh;

struct gvizhip_batch {
  ForceCommon c;
  BatchState b;
  std::vector<int64_t> start;   // G + 1
  int32_t waveOrder = 0;        // 0: graph order, 1: largest graphs first
  int32_t stepsPerLaunch = 4;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int32_t graphSteps = 0;
  DevBuf<double> savePos;
  DevBuf<unsigned char> saveSf;
  DevBuf<Scalars> saveSc;
};

namespace {

int batch_step(gvizhip_batch *bt) {
  switch (bt->c.prm.precision) {
  case GVIZHIP_PRECISION_FP64:
    return typed_batch_step<double, double>(bt->c, bt->b);
  case GVIZHIP_PRECISION_MIXED:
    return typed_batch_step<float, double>(bt->c, bt->b);
  default:
    return typed_batch_step<float, float>(bt->c, bt->b);
  }
}

void drop_graph(gvizhip_batch *bt) {
  if (bt->exec)
    (void)hipGraphExecDestroy(bt->exec);
  if (bt->graph)
    (void)hipGraphDestroy(bt->graph);
  bt->exec = nullptr;
  bt->graph = nullptr;
  bt->graphSteps = 0;
}

int build_graph(gvizhip_batch *bt) {
  if (bt->exec && bt->graphSteps == bt->stepsPerLaunch)
    return 0;
  drop_graph(bt);
  GVH_CHECK(hipStreamBeginCapture(bt->c.stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int k = 0; k < bt->stepsPerLaunch && rc == 0; k++)
    rc = batch_step(bt);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(bt->c.stream, &g);
  if (rc || e != hipSuccess) {
    if (g)
      (void)hipGraphDestroy(g);
    if (!rc)
      set_error("stream capture of the batch step failed: %s", hipGetErrorString(e));
    return -1;
  }
  bt->graph = g;
  GVH_CHECK(hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0));
  bt->graphSteps = bt->stepsPerLaunch;
  return 0;
}

// Traversal waves: 64 consecutive positions of one graph each, graphs in id
// order or largest first. The order only schedules work; every wave's
// targets and arithmetic are the same either way.
int upload_waves(gvizhip_batch *bt) {
  const int64_t G = bt->b.graphs;
  std::vector<int64_t> order(G);
  std::iota(order.begin(), order.end(), 0);
  if (bt->waveOrder == 1)
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t c) {
      return bt->start[a + 1] - bt->start[a] > bt->start[c + 1] - bt->start[c];
    });
  std::vector<int32_t> wg;
  std::vector<int64_t> wb;
  for (int64_t g : order)
    for (int64_t s = bt->start[g]; s < bt->start[g + 1]; s += 64) {
      wg.push_back((int32_t)g);
      wb.push_back(s);
    }
  bt->b.waves = (int64_t)wg.size();
  if (bt->b.waveGraph.alloc(wg.size()) || bt->b.waveBase.alloc(wb.size()))
    return -1;
  GVH_CHECK(hipMemcpy(bt->b.waveGraph.p, wg.data(), sizeof(int32_t) * wg.size(),
                      hipMemcpyHostToDevice));
  GVH_CHECK(hipMemcpy(bt->b.waveBase.p, wb.data(), sizeof(int64_t) * wb.size(),
                      hipMemcpyHostToDevice));
  drop_graph(bt);
  return 0;
}

int read_scalars(gvizhip_batch *bt, std::vector<Scalars> &out) {
  out.resize(bt->b.graphs);
  GVH_CHECK(hipMemcpyAsync(out.data(), bt->b.scG.p, sizeof(Scalars) * bt->b.graphs,
                           hipMemcpyDeviceToHost, bt->c.stream));
  GVH_CHECK(hipStreamSynchronize(bt->c.stream));
  return 0;
}

int save_state(gvizhip_batch *bt) {
  ForceCommon &c = bt->c;
  const int64_t n = c.n;
  if (bt->savePos.alloc(2 * n) || bt->saveSf.alloc(2 * n * sizeof(double)) ||
      bt->saveSc.alloc(bt->b.graphs))
    return -1;
  GVH_CHECK(hipMemcpyAsync(bt->savePos.p, c.pos64.p, sizeof(double) * 2 * n,
                           hipMemcpyDeviceToDevice, c.stream));
  GVH_CHECK(hipMemcpyAsync(bt->saveSf.p, c.sf.p, sizeof(double) * 2 * n,
                           hipMemcpyDeviceToDevice, c.stream));
  GVH_CHECK(hipMemcpyAsync(bt->saveSc.p, bt->b.scG.p, sizeof(Scalars) * bt->b.graphs,
                           hipMemcpyDeviceToDevice, c.stream));
  return 0;
}

int restore_state(gvizhip_batch *bt) {
  ForceCommon &c = bt->c;
  const int64_t n = c.n;
  GVH_CHECK(hipMemcpyAsync(c.pos64.p, bt->savePos.p, sizeof(double) * 2 * n,
                           hipMemcpyDeviceToDevice, c.stream));
  GVH_CHECK(hipMemcpyAsync(c.sf.p, bt->saveSf.p, sizeof(double) * 2 * n,
                           hipMemcpyDeviceToDevice, c.stream));
  GVH_CHECK(hipMemcpyAsync(bt->b.scG.p, bt->saveSc.p, sizeof(Scalars) * bt->b.graphs,
                           hipMemcpyDeviceToDevice, c.stream));
  GVH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

} // namespace

extern "C" {

int gvizhip_batch_create(gvizhip_batch **out, int64_t graphs, const int64_t *vertexStart,
                         const int64_t *offsets, const int64_t *adj,
                         const gvizhip_force_config *cfg) {
  if (!out || !cfg || graphs <= 0 || graphs > 4096 || !vertexStart || !offsets) {
    set_error("gvizhip_batch_create: invalid arguments (1..4096 graphs)");
    return -1;
  }
  *out = nullptr;
  if (gvizhip_device_count() <= 0) {
    set_error("gvizhip_batch_create: no GPU is visible to HIP");
    return -1;
  }
  if (!cfg->barnesHut) {
    set_error("gvizhip_batch_create: the batch engine runs Barnes-Hut only");
    return -1;
  }
  const int64_t n = vertexStart[graphs];
  if (vertexStart[0] != 0 || n <= 0 || n >= INT32_MAX) {
    set_error("gvizhip_batch_create: vertex ranges must start at 0 and hold < 2^31");
    return -1;
  }
  const int64_t m2 = offsets[n];
  if (m2 < 0 || m2 >= INT32_MAX || (m2 > 0 && !adj)) {
    set_error("gvizhip_batch_create: bad adjacency size");
    return -1;
  }
  std::vector<int32_t> gid(n), off32(n + 1), adj32(m2 > 0 ? m2 : 1), deg32(n);
  for (int64_t g = 0; g < graphs; g++) {
    const int64_t b = vertexStart[g], e = vertexStart[g + 1];
    if (e <= b) {
      set_error("gvizhip_batch_create: graph %lld has no vertex", (long long)g);
      return -1;
    }
    for (int64_t v = b; v < e; v++) {
      gid[v] = (int32_t)g;
      for (int64_t k = offsets[v]; k < offsets[v + 1]; k++)
        if (adj[k] < b || adj[k] >= e) {
          set_error("gvizhip_batch_create: edge of vertex %lld leaves its graph",
                    (long long)v);
          return -1;
        }
    }
  }
  for (int64_t v = 0; v <= n; v++)
    off32[v] = (int32_t)offsets[v];
  for (int64_t v = 0; v < n; v++) {
    if (offsets[v + 1] < offsets[v]) {
      set_error("gvizhip_batch_create: offsets decrease at %lld", (long long)v);
      return -1;
    }
    deg32[v] = (int32_t)(offsets[v + 1] - offsets[v]);
  }
  for (int64_t e = 0; e < m2; e++)
    adj32[e] = (int32_t)adj[e];

  auto *bt = new gvizhip_batch();
  ForceCommon &c = bt->c;
  BatchState &bs = bt->b;
  c.prm = *(const ForceParams *)cfg;
  bs.graphs = graphs;
  bt->start.assign(vertexStart, vertexStart + graphs + 1);
  if (c.open_streams() || c.alloc(n, m2)) {
    if (!gvizhip_last_error()[0])
      set_error("gvizhip_batch_create: allocation failed");
    gvizhip_batch_destroy(bt);
    return -1;
  }
  // per-graph reduction blocks, as a single-graph run of that size uses
  std::vector<int32_t> gBlocks(graphs), vbGraph, vbIdx;
  std::vector<int64_t> gPart(graphs);
  int64_t vb = 0;
  for (int64_t g = 0; g < graphs; g++) {
    const int64_t ng = vertexStart[g + 1] - vertexStart[g];
    const int64_t rb = std::min<int64_t>((ng + 255) / 256, c.reductionBlocks);
    gBlocks[g] = (int32_t)rb;
    gPart[g] = vb;
    for (int64_t k = 0; k < rb; k++) {
      vbGraph.push_back((int32_t)g);
      vbIdx.push_back((int32_t)k);
    }
    vb += rb;
  }
  bs.vblocks = vb;
  std::vector<int32_t> heavy;
  std::vector<double> mass(n), rad(n);
  std::vector<Scalars> scs(graphs);
  for (int64_t g = 0; g < graphs; g++) {
    double sumDeg1 = 0.0;
    for (int64_t v = vertexStart[g]; v < vertexStart[g + 1]; v++)
      sumDeg1 += (double)(deg32[v] + 1);
    Scalars s{};
    s.globalSpeed = 1.0;
    s.speedEfficiency = 1.0;
    s.epsilon = -1.0;
    s.meanDegreePlusOne = sumDeg1 / (double)(vertexStart[g + 1] - vertexStart[g]);
    scs[g] = s;
  }
  for (int64_t v = 0; v < n; v++) {
    const double deg = (double)deg32[v];
    if (deg32[v] > 32)
      heavy.push_back((int32_t)v);
    mass[v] = c.prm.model == 1 ? 1.0 + deg : 1.0;
    rad[v] = c.prm.preventOverlap
                 ? c.prm.radiusBase * (1.0 + c.prm.radiusPerDegree * std::sqrt(deg))
                 : 0.0;
  }
  c.nHeavy = (int64_t)heavy.size();
  Scalars neutral{};
  neutral.maxDepth = 64;
  neutral.epsilon = -1.0;
  bool bad = bs.gid.alloc(n) || bs.gStart.alloc(graphs + 1) || bs.gBlocks.alloc(graphs) ||
             bs.gPart.alloc(graphs) || bs.vbGraph.alloc(vb) || bs.vbIdx.alloc(vb) ||
             bs.rootOf.alloc(graphs) || bs.scG.alloc(graphs) || bs.last.alloc(graphs) ||
             bs.partials.alloc(3 * vb) ||
             c.heavyRows.alloc(heavy.empty() ? 1 : heavy.size()) ||
             c.attHeavy.alloc(2 * n * sizeof(double));
  if (bad) {
    gvizhip_batch_destroy(bt);
    return -1;
  }
#define UP(dst, src, bytes)                                                       \
  if (hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {          \
    set_error("gvizhip_batch_create: upload failed");                             \
    gvizhip_batch_destroy(bt);                                                    \
    return -1;                                                                    \
  }
  UP(bs.gid.p, gid.data(), sizeof(int32_t) * n);
  UP(bs.gStart.p, vertexStart, sizeof(int64_t) * (graphs + 1));
  UP(bs.gBlocks.p, gBlocks.data(), sizeof(int32_t) * graphs);
  UP(bs.gPart.p, gPart.data(), sizeof(int64_t) * graphs);
  UP(bs.vbGraph.p, vbGraph.data(), sizeof(int32_t) * vb);
  UP(bs.vbIdx.p, vbIdx.data(), sizeof(int32_t) * vb);
  UP(bs.scG.p, scs.data(), sizeof(Scalars) * graphs);
  UP(c.sc.p, &neutral, sizeof(Scalars));
  UP(c.csrOff.p, off32.data(), sizeof(int32_t) * (n + 1));
  if (m2 > 0)
    UP(c.csrAdj.p, adj32.data(), sizeof(int32_t) * m2);
  UP(c.degree.p, deg32.data(), sizeof(int32_t) * n);
  UP(c.mass64.p, mass.data(), sizeof(double) * n);
  UP(c.radius64.p, rad.data(), sizeof(double) * n);
  if (!heavy.empty())
    UP(c.heavyRows.p, heavy.data(), sizeof(int32_t) * heavy.size());
#undef UP
  int rc = c.prm.precision == GVIZHIP_PRECISION_FP64 ? typed_upload<double>(c)
                                                     : typed_upload<float>(c);
  if (rc || upload_waves(bt) ||
      hipMemsetAsync(c.sf.p, 0, 2 * n * sizeof(double), c.stream) != hipSuccess ||
      hipStreamSynchronize(c.stream) != hipSuccess) {
    if (!gvizhip_last_error()[0])
      set_error("gvizhip_batch_create: setup failed");
    gvizhip_batch_destroy(bt);
    return -1;
  }
  *out = bt;
  return 0;
}

int gvizhip_batch_begin(gvizhip_batch *bt, const double *positions) {
  if (!bt || !positions) {
    set_error("gvizhip_batch_begin: invalid arguments");
    return -1;
  }
  ForceCommon &c = bt->c;
  const int64_t G = bt->b.graphs;
  std::vector<Scalars> s;
  if (read_scalars(bt, s))
    return -1;
  for (auto &x : s) {
    x.globalSpeed = 1.0;
    x.speedEfficiency = 1.0;
    x.iteration = 0;
    x.recordBase = 0;
    x.stopped = 0;
    x.epsilon = -1.0;
    x.traversalOverflow = 0;
  }
  GVH_CHECK(hipMemcpyAsync(bt->b.scG.p, s.data(), sizeof(Scalars) * G,
                           hipMemcpyHostToDevice, c.stream));
  GVH_CHECK(hipMemcpyAsync(c.pos64.p, positions, sizeof(double) * 2 * c.n,
                           hipMemcpyHostToDevice, c.stream));
  GVH_CHECK(hipMemsetAsync(c.sf.p, 0, 2 * c.n * sizeof(double), c.stream));
  GVH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

int gvizhip_batch_schedule(gvizhip_batch *bt, int32_t waveOrder, int32_t stepsPerLaunch,
                           int32_t traversalWaves) {
  if (!bt || waveOrder < 0 || waveOrder > 1 || stepsPerLaunch < 1 ||
      stepsPerLaunch > 64 ||
      (traversalWaves != 1 && traversalWaves != 2 && traversalWaves != 4)) {
    set_error("gvizhip_batch_schedule: invalid arguments");
    return -1;
  }
  bt->stepsPerLaunch = stepsPerLaunch;
  bt->c.travWaves = traversalWaves;
  if (waveOrder != bt->waveOrder) {
    bt->waveOrder = waveOrder;
    if (upload_waves(bt))
      return -1;
  }
  drop_graph(bt);
  return 0;
}

int64_t gvizhip_batch_run(gvizhip_batch *bt, int64_t maxIters, double epsilon,
                          int64_t *iterations, double *latencyMs, double *maxDisp) {
  if (!bt || maxIters < 0) {
    set_error("gvizhip_batch_run: invalid arguments");
    return -1;
  }
  ForceCommon &c = bt->c;
  const int64_t G = bt->b.graphs;
  std::vector<Scalars> s;
  if (read_scalars(bt, s))
    return -1;
  std::vector<int64_t> startIter(G);
  for (int64_t g = 0; g < G; g++) {
    startIter[g] = s[g].iteration;
    s[g].epsilon = epsilon;
    s[g].stopped = 0;
  }
  GVH_CHECK(hipMemcpyAsync(bt->b.scG.p, s.data(), sizeof(Scalars) * G,
                           hipMemcpyHostToDevice, c.stream));
  if (build_graph(bt))
    return -1;
  std::vector<double> lat(G, -1.0);
  const auto t0 = std::chrono::steady_clock::now();
  int64_t done = 0;
  int64_t live = G;
  while (done < maxIters && live > 0) {
    if (maxIters - done >= bt->stepsPerLaunch) {
      GVH_CHECK(hipGraphLaunch(bt->exec, c.stream));
      done += bt->stepsPerLaunch;
    } else {
      if (batch_step(bt))
        return -1;
      done++;
    }
    if (epsilon >= 0.0 || latencyMs) {
      if (read_scalars(bt, s))
        return -1;
      const double ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
              .count();
      live = 0;
      for (int64_t g = 0; g < G; g++) {
        if (s[g].stopped) {
          if (lat[g] < 0.0)
            lat[g] = ms;
        } else {
          live++;
        }
      }
    }
  }
  GVH_CHECK(hipStreamSynchronize(c.stream));
  const double total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
          .count();
  if (read_scalars(bt, s))
    return -1;
  for (int64_t g = 0; g < G; g++) {
    if (s[g].traversalOverflow) {
      set_error(s[g].traversalOverflow == 2
                    ? "gvizhip_batch_run: graph %lld needs more than 58 quadtree levels"
                    : "gvizhip_batch_run: traversal stack overflow in graph %lld",
                (long long)g);
      return -1;
    }
    if (iterations)
      iterations[g] = s[g].iteration - startIter[g];
    if (latencyMs)
      latencyMs[g] = lat[g] >= 0.0 ? lat[g] : total;
  }
  if (maxDisp) {
    std::vector<StepRecord> last(G);
    GVH_CHECK(hipMemcpy(last.data(), bt->b.last.p, sizeof(StepRecord) * G,
                        hipMemcpyDeviceToHost));
    for (int64_t g = 0; g < G; g++)
      maxDisp[g] = last[g].maxDisp;
  }
  return done;
}

int gvizhip_batch_autotune(gvizhip_batch *bt, int64_t trialSteps, int32_t *waveOrder,
                           int32_t *stepsPerLaunch, int32_t *traversalWaves) {
  if (!bt || trialSteps < 1) {
    set_error("gvizhip_batch_autotune: invalid arguments");
    return -1;
  }
  if (save_state(bt))
    return -1;
  double best = INFINITY;
  int32_t bo = bt->waveOrder, bsp = bt->stepsPerLaunch, btw = bt->c.travWaves;
  const int32_t spl[3] = {1, 4, 16};
  const int32_t tws[3] = {1, 2, 4};
  for (int32_t order = 0; order <= 1; order++)
    for (int32_t k = 0; k < 3; k++)
      for (int32_t tw : tws) {
      if (gvizhip_batch_schedule(bt, order, spl[k], tw) || restore_state(bt))
        return -1;
      // one launch to build and warm, then the timed trial
      if (gvizhip_batch_run(bt, spl[k], -1.0, nullptr, nullptr, nullptr) < 0)
        return -1;
      const auto t0 = std::chrono::steady_clock::now();
      const int64_t steps = ((trialSteps + spl[k] - 1) / spl[k]) * spl[k];
      if (gvizhip_batch_run(bt, steps, -1.0, nullptr, nullptr, nullptr) < 0)
        return -1;
      const double per =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() /
          (double)steps;
      if (per < best) {
        best = per;
        bo = order;
        bsp = spl[k];
        btw = tw;
      }
    }
  if (gvizhip_batch_schedule(bt, bo, bsp, btw) || restore_state(bt))
    return -1;
  if (waveOrder)
    *waveOrder = bo;
  if (stepsPerLaunch)
    *stepsPerLaunch = bsp;
  if (traversalWaves)
    *traversalWaves = btw;
  return 0;
}

int gvizhip_batch_profile(gvizhip_batch *bt, int64_t steps, double *traversalMs,
                          int64_t *interactions, double *stepMs) {
  if (!bt || steps < 0) {
    set_error("gvizhip_batch_profile: invalid arguments");
    return -1;
  }
  ForceCommon &c = bt->c;
  if (!c.travBegin) {
    GVH_CHECK(hipEventCreate(&c.travBegin));
    GVH_CHECK(hipEventCreate(&c.travEnd));
  }
  hipEvent_t s0, s1;
  GVH_CHECK(hipEventCreate(&s0));
  GVH_CHECK(hipEventCreate(&s1));
  std::vector<int32_t> cnt(c.n);
  double trav = 0.0, step = 0.0;
  int64_t inter = 0;
  int rc = 0;
  for (int64_t k = 0; k < steps && rc == 0; k++) {
    GVH_CHECK(hipMemsetAsync(c.interactions.p, 0, sizeof(int32_t) * c.n, c.stream));
    c.timeTraversal = 1;
    c.wantInteractions = 1;
    if (hipEventRecord(s0, c.stream) != hipSuccess || batch_step(bt) ||
        hipEventRecord(s1, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess)
      rc = -1;
    c.timeTraversal = 0;
    c.wantInteractions = 0;
    float ms = 0.0f;
    if (rc == 0) {
      (void)hipEventElapsedTime(&ms, c.travBegin, c.travEnd);
      trav += ms;
      (void)hipEventElapsedTime(&ms, s0, s1);
      step += ms;
      if (hipMemcpy(cnt.data(), c.interactions.p, sizeof(int32_t) * c.n,
                    hipMemcpyDeviceToHost) != hipSuccess)
        rc = -1;
      for (int64_t i = 0; rc == 0 && i < c.n; i++)
        inter += cnt[i];
    }
  }
  (void)hipEventDestroy(s0);
  (void)hipEventDestroy(s1);
  drop_graph(bt);
  if (rc) {
    if (!gvizhip_last_error()[0])
      set_error("gvizhip_batch_profile failed");
    return -1;
  }
  if (traversalMs)
    *traversalMs = trav;
  if (interactions)
    *interactions = inter;
  if (stepMs)
    *stepMs = step;
  return 0;
}

int gvizhip_batch_download_positions(gvizhip_batch *bt, double *positions) {
  if (!bt || !positions) {
    set_error("gvizhip_batch_download_positions: invalid arguments");
    return -1;
  }
  GVH_CHECK(hipMemcpyAsync(positions, bt->c.pos64.p, sizeof(double) * 2 * bt->c.n,
                           hipMemcpyDeviceToHost, bt->c.stream));
  GVH_CHECK(hipStreamSynchronize(bt->c.stream));
  return 0;
}

void gvizhip_batch_destroy(gvizhip_batch *bt) {
  if (!bt)
    return;
  drop_graph(bt);
  if (bt->c.travBegin)
    (void)hipEventDestroy(bt->c.travBegin);
  if (bt->c.travEnd)
    (void)hipEventDestroy(bt->c.travEnd);
  bt->c.close_streams();
  delete bt;
}

} // extern "C"
