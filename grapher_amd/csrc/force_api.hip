// force_api.hip -- C ABI of the force-directed engine (include/gviz_hip.h).
//
// One gvizhip_force per embedder: its own stream, device buffers and a
// hipGraph holding one captured step, replayed by Step and Run.
#include "../../include/gviz_hip.h"
#include "force_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace gvh;

namespace gvh {
template <class Real, class PosT> int typed_step(ForceCommon &c, int mode);
template <class Real> int typed_upload(ForceCommon &c);

__global__ void k_begin_run(Scalars *sc, double epsilon) {
  sc->recordBase = sc->iteration;
  sc->epsilon = epsilon;
  sc->stopped = 0;
  sc->traversalOverflow = 0; // set by any step of the run, read after it
}
} // namespace gvh

struct gvizhip_force {
  ForceCommon c;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graphValid = false;
  DevBuf<unsigned char> sfBackup;
  std::vector<int64_t> degreeHost;
  Scalars *pinned = nullptr; // host mirror for asynchronous reads
  // the key/sort buffers hold the tree of the last step's starting positions
  // (cleared when positions are replaced or another tree is built over them)
  bool treeFresh = false;
};

namespace {

int dispatch_step(gvizhip_force *f, int mode) {
  switch (f->c.prm.precision) {
  case GVIZHIP_PRECISION_FP64:
    return typed_step<double, double>(f->c, mode);
  case GVIZHIP_PRECISION_MIXED:
    return typed_step<float, double>(f->c, mode);
  default:
    return typed_step<float, float>(f->c, mode);
  }
}

size_t real_size(const ForceCommon &c) {
  return c.prm.precision == GVIZHIP_PRECISION_FP64 ? sizeof(double) : sizeof(float);
}

void drop_graph(gvizhip_force *f) {
  if (f->exec)
    (void)hipGraphExecDestroy(f->exec);
  if (f->graph)
    (void)hipGraphDestroy(f->graph);
  f->exec = nullptr;
  f->graph = nullptr;
  f->graphValid = false;
}

int build_graph(gvizhip_force *f) {
  if (f->graphValid)
    return 0;
  drop_graph(f);
  GVH_CHECK(hipStreamBeginCapture(f->c.stream, hipStreamCaptureModeThreadLocal));
  int rc = dispatch_step(f, kStepFull);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(f->c.stream, &g);
  if (rc != 0 || e != hipSuccess) {
    if (g)
      (void)hipGraphDestroy(g);
    if (rc == 0)
      set_error("stream capture of the force step failed: %s",
                hipGetErrorString(e));
    return -1;
  }
  f->graph = g;
  GVH_CHECK(hipGraphInstantiate(&f->exec, g, nullptr, nullptr, 0));
  f->graphValid = true;
  return 0;
}

int read_scalars(gvizhip_force *f, Scalars *out) {
  GVH_CHECK(hipMemcpyAsync(f->pinned, f->c.sc.p, sizeof(Scalars),
                           hipMemcpyDeviceToHost, f->c.stream));
  GVH_CHECK(hipStreamSynchronize(f->c.stream));
  *out = *f->pinned;
  return 0;
}

int write_scalars(gvizhip_force *f, const Scalars &v) {
  *f->pinned = v;
  GVH_CHECK(hipMemcpyAsync(f->c.sc.p, f->pinned, sizeof(Scalars),
                           hipMemcpyHostToDevice, f->c.stream));
  GVH_CHECK(hipStreamSynchronize(f->c.stream));
  return 0;
}

// Masses (src/embedders/gvizForceModel.c:5-8,31) and radii
// (src/embedders/gvizForceEmbedder.c:28-33,389-393) from the degrees.
int upload_vertex_terms(gvizhip_force *f) {
  ForceCommon &c = f->c;
  const int64_t n = c.n;
  std::vector<double> m(n > 0 ? n : 1), r(n > 0 ? n : 1);
  for (int64_t i = 0; i < n; i++) {
    double deg = (double)f->degreeHost[i];
    m[i] = c.prm.model == 1 ? 1.0 + deg : 1.0;
    r[i] = c.prm.preventOverlap
               ? c.prm.radiusBase * (1.0 + c.prm.radiusPerDegree * std::sqrt(deg))
               : 0.0;
  }
  if (n > 0) {
    GVH_CHECK(hipMemcpyAsync(c.mass64.p, m.data(), sizeof(double) * n,
                             hipMemcpyHostToDevice, c.stream));
    GVH_CHECK(hipMemcpyAsync(c.radius64.p, r.data(), sizeof(double) * n,
                             hipMemcpyHostToDevice, c.stream));
  }
  int rc = c.prm.precision == GVIZHIP_PRECISION_FP64 ? typed_upload<double>(c)
                                                     : typed_upload<float>(c);
  if (rc)
    return -1;
  GVH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

int ensure_records(gvizhip_force *f, int64_t count) {
  if (count <= f->c.recordCap && f->c.records.p)
    return 0;
  int64_t cap = std::max<int64_t>(count, 4096);
  if (f->c.records.alloc(cap))
    return -1;
  f->c.recordCap = cap;
  f->graphValid = false;
  return 0;
}

void to_double(const void *src, size_t realSize, int64_t count, double *dst) {
  if (realSize == sizeof(double)) {
    std::memcpy(dst, src, sizeof(double) * count);
  } else {
    const float *s = (const float *)src;
    for (int64_t i = 0; i < count; i++)
      dst[i] = (double)s[i];
  }
}

int download_real(gvizhip_force *f, const DevBuf<unsigned char> &buf,
                  int64_t count, double *dst) {
  if (!dst || count == 0)
    return 0;
  size_t rs = real_size(f->c);
  std::vector<unsigned char> tmp(rs * count);
  GVH_CHECK(hipMemcpyAsync(tmp.data(), buf.p, rs * count, hipMemcpyDeviceToHost,
                           f->c.stream));
  GVH_CHECK(hipStreamSynchronize(f->c.stream));
  to_double(tmp.data(), rs, count, dst);
  return 0;
}

void fill_stats(const StepRecord &r, gvizhip_force_stats *s) {
  s->maxDisp = r.maxDisp;
  s->speed = r.speed;
  s->meanAtt = r.meanAtt;
  s->meanRep = r.meanRep;
  s->meanGrav = r.meanGrav;
  s->speedEfficiency = r.speedEfficiency;
  s->totalSwinging = r.totalSwinging;
  s->totalTraction = r.totalTraction;
}

} // namespace

extern "C" {

const char *gvizhip_last_error(void) { return gvh::last_error(); }

int gvizhip_device_count(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess)
    return 0;
  return count;
}

int gvizhip_force_create(gvizhip_force **out, int64_t n, const int64_t *offsets,
                         const int64_t *adj, const gvizhip_force_config *cfg) {
  if (!out || !cfg || n < 0 || (n > 0 && !offsets)) {
    set_error("gvizhip_force_create: invalid arguments");
    return -1;
  }
  *out = nullptr;
  if (gvizhip_device_count() <= 0) {
    set_error("gvizhip_force_create: no GPU is visible to HIP");
    return -1;
  }
  const int64_t m2 = n > 0 ? offsets[n] : 0;
  if (n >= INT32_MAX || m2 >= INT32_MAX) {
    set_error("gvizhip_force_create: graph exceeds 2^31 vertices or entries");
    return -1;
  }
  std::vector<int32_t> off32(n + 1), adj32(m2 > 0 ? m2 : 1), deg32(n > 0 ? n : 1);
  for (int64_t i = 0; i <= n; i++)
    off32[i] = n > 0 ? (int32_t)offsets[i] : 0;
  for (int64_t i = 0; i < n; i++) {
    if (offsets[i + 1] < offsets[i]) {
      set_error("gvizhip_force_create: offsets decrease at %lld", (long long)i);
      return -1;
    }
    deg32[i] = (int32_t)(offsets[i + 1] - offsets[i]);
  }
  for (int64_t e = 0; e < m2; e++) {
    if (adj[e] < 0 || adj[e] >= n) {
      set_error("gvizhip_force_create: adjacency entry %lld out of range",
                (long long)e);
      return -1;
    }
    adj32[e] = (int32_t)adj[e];
  }
  gvizhip_force *f = new gvizhip_force();
  ForceCommon &c = f->c;
  c.prm = *(const ForceParams *)cfg;
  if (c.open_streams() ||
      hipHostMalloc((void **)&f->pinned, sizeof(Scalars), hipHostMallocDefault) !=
          hipSuccess) {
    set_error("gvizhip_force_create: stream or pinned allocation failed");
    gvizhip_force_destroy(f);
    return -1;
  }
  if (c.alloc(n, m2) || f->sfBackup.alloc(2 * (n > 0 ? n : 1) * sizeof(double)) ||
      ensure_records(f, 4096)) {
    gvizhip_force_destroy(f);
    return -1;
  }
  f->degreeHost.resize(n);
  double sumDeg1 = 0.0;
  std::vector<int32_t> heavy;
  for (int64_t i = 0; i < n; i++) {
    f->degreeHost[i] = deg32[i];
    sumDeg1 += (double)(deg32[i] + 1);
    if (deg32[i] > 32)
      heavy.push_back((int32_t)i);
  }
  c.nHeavy = (int64_t)heavy.size();
  if (c.heavyRows.alloc(heavy.empty() ? 1 : heavy.size()) ||
      c.attHeavy.alloc(2 * (n > 0 ? n : 1) * sizeof(double)) ||
      (!heavy.empty() &&
       hipMemcpy(c.heavyRows.p, heavy.data(), sizeof(int32_t) * heavy.size(),
                 hipMemcpyHostToDevice) != hipSuccess)) {
    set_error("gvizhip_force_create: heavy-row setup failed");
    gvizhip_force_destroy(f);
    return -1;
  }
  if (hipMemcpy(c.csrOff.p, off32.data(), sizeof(int32_t) * (n + 1),
                hipMemcpyHostToDevice) != hipSuccess ||
      (m2 > 0 && hipMemcpy(c.csrAdj.p, adj32.data(), sizeof(int32_t) * m2,
                           hipMemcpyHostToDevice) != hipSuccess) ||
      (n > 0 && hipMemcpy(c.degree.p, deg32.data(), sizeof(int32_t) * n,
                          hipMemcpyHostToDevice) != hipSuccess)) {
    set_error("gvizhip_force_create: upload failed");
    gvizhip_force_destroy(f);
    return -1;
  }
  Scalars s{};
  s.globalSpeed = 1.0;
  s.speedEfficiency = 1.0;
  s.epsilon = -1.0;
  s.meanDegreePlusOne = n ? sumDeg1 / (double)n : 0.0;
  if (write_scalars(f, s) || upload_vertex_terms(f)) {
    gvizhip_force_destroy(f);
    return -1;
  }
  if (hipMemsetAsync(c.sf.p, 0, 2 * (n > 0 ? n : 1) * sizeof(double), c.stream) !=
          hipSuccess ||
      hipStreamSynchronize(c.stream) != hipSuccess) {
    set_error("gvizhip_force_create: clearing forces failed");
    gvizhip_force_destroy(f);
    return -1;
  }
  *out = f;
  return 0;
}

int gvizhip_force_configure(gvizhip_force *f, const gvizhip_force_config *cfg) {
  if (!f || !cfg) {
    set_error("gvizhip_force_configure: invalid arguments");
    return -1;
  }
  const int oldPrecision = f->c.prm.precision;
  f->c.prm = *(const ForceParams *)cfg;
  f->graphValid = false;
  if (upload_vertex_terms(f))
    return -1;
  if (oldPrecision != f->c.prm.precision) {
    // structForce changes representation: the reference keeps it across
    // configuration calls, so convert it.
    size_t oldRs = oldPrecision == GVIZHIP_PRECISION_FP64 ? 8 : 4;
    size_t newRs = real_size(f->c);
    int64_t cnt = 2 * f->c.n;
    std::vector<unsigned char> tmp(oldRs * (cnt > 0 ? cnt : 1));
    std::vector<double> d(cnt > 0 ? cnt : 1);
    if (cnt > 0) {
      GVH_CHECK(hipMemcpy(tmp.data(), f->c.sf.p, oldRs * cnt, hipMemcpyDeviceToHost));
      to_double(tmp.data(), oldRs, cnt, d.data());
      if (newRs == 8) {
        GVH_CHECK(hipMemcpy(f->c.sf.p, d.data(), 8 * cnt, hipMemcpyHostToDevice));
      } else {
        std::vector<float> fl(cnt);
        for (int64_t i = 0; i < cnt; i++)
          fl[i] = (float)d[i];
        GVH_CHECK(hipMemcpy(f->c.sf.p, fl.data(), 4 * cnt, hipMemcpyHostToDevice));
      }
    }
  }
  return 0;
}

int gvizhip_force_begin(gvizhip_force *f, const double *positions) {
  if (!f) {
    set_error("gvizhip_force_begin: null context");
    return -1;
  }
  ForceCommon &c = f->c;
  f->treeFresh = false;
  if (c.n > 0 && positions)
    GVH_CHECK(hipMemcpyAsync(c.pos64.p, positions, sizeof(double) * 2 * c.n,
                             hipMemcpyHostToDevice, c.stream));
  GVH_CHECK(hipMemsetAsync(c.sf.p, 0, 2 * (c.n > 0 ? c.n : 1) * sizeof(double),
                           c.stream));
  Scalars s;
  if (read_scalars(f, &s))
    return -1;
  s.globalSpeed = 1.0;
  s.speedEfficiency = 1.0;
  s.iteration = 0;
  s.recordBase = 0;
  s.stopped = 0;
  s.epsilon = -1.0;
  return write_scalars(f, s);
}

int gvizhip_force_upload_positions(gvizhip_force *f, const double *positions) {
  if (!f || (!positions && f->c.n > 0)) {
    set_error("gvizhip_force_upload_positions: invalid arguments");
    return -1;
  }
  f->treeFresh = false;
  if (f->c.n > 0)
    GVH_CHECK(hipMemcpyAsync(f->c.pos64.p, positions, sizeof(double) * 2 * f->c.n,
                             hipMemcpyHostToDevice, f->c.stream));
  GVH_CHECK(hipStreamSynchronize(f->c.stream));
  return 0;
}

int gvizhip_force_download_positions(gvizhip_force *f, double *positions) {
  if (!f || (!positions && f->c.n > 0)) {
    set_error("gvizhip_force_download_positions: invalid arguments");
    return -1;
  }
  if (f->c.n > 0)
    GVH_CHECK(hipMemcpyAsync(positions, f->c.pos64.p, sizeof(double) * 2 * f->c.n,
                             hipMemcpyDeviceToHost, f->c.stream));
  GVH_CHECK(hipStreamSynchronize(f->c.stream));
  return 0;
}

int64_t gvizhip_force_run(gvizhip_force *f, int64_t maxIters, double epsilon,
                          gvizhip_force_stats *stats) {
  if (!f || maxIters < 0) {
    set_error("gvizhip_force_run: invalid arguments");
    return -1;
  }
  ForceCommon &c = f->c;
  if (maxIters == 0)
    return 0;
  if (stats && ensure_records(f, maxIters))
    return -1;
  if (build_graph(f))
    return -1;
  Scalars before;
  if (read_scalars(f, &before))
    return -1;
  hipLaunchKernelGGL(k_begin_run, dim3(1), dim3(1), 0, c.stream, c.sc.p, epsilon);
  GVH_CHECK_LAUNCH();
  const int64_t checkEvery = 16;
  int64_t launched = 0;
  while (launched < maxIters) {
    int64_t batch = std::min<int64_t>(checkEvery, maxIters - launched);
    for (int64_t k = 0; k < batch; k++)
      GVH_CHECK(hipGraphLaunch(f->exec, c.stream));
    launched += batch;
    if (epsilon >= 0.0 && launched < maxIters) {
      Scalars s;
      if (read_scalars(f, &s))
        return -1;
      if (s.stopped)
        break;
    }
  }
  Scalars after;
  if (read_scalars(f, &after))
    return -1;
  if (after.traversalOverflow) {
    set_error("gvizhip_force_run: traversal stack overflow");
    return -1;
  }
  const int64_t rounds = after.iteration - before.iteration;
  if (rounds > 0)
    f->treeFresh = c.prm.barnesHut != 0;
  if (stats && rounds > 0) {
    std::vector<StepRecord> rec(rounds);
    GVH_CHECK(hipMemcpy(rec.data(), c.records.p, sizeof(StepRecord) * rounds,
                        hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < rounds; i++)
      fill_stats(rec[i], &stats[i]);
  }
  return rounds;
}

int gvizhip_force_step(gvizhip_force *f, gvizhip_force_stats *stats) {
  gvizhip_force_stats s{};
  int64_t r = gvizhip_force_run(f, 1, -1.0, &s);
  if (r != 1)
    return -1;
  if (stats)
    *stats = s;
  return 0;
}

int gvizhip_force_evaluate(gvizhip_force *f, double *disp, double *structForce,
                           double *attMag, double *repMag,
                           int32_t *interactions) {
  if (!f) {
    set_error("gvizhip_force_evaluate: null context");
    return -1;
  }
  ForceCommon &c = f->c;
  const int64_t n = c.n;
  if (n == 0)
    return 0;
  const size_t rs = real_size(c);
  GVH_CHECK(hipMemcpyAsync(f->sfBackup.p, c.sf.p, 2 * n * rs,
                           hipMemcpyDeviceToDevice, c.stream));
  Scalars s;
  if (read_scalars(f, &s))
    return -1;
  int stopped = s.stopped;
  s.stopped = 0;
  s.traversalOverflow = 0;
  if (write_scalars(f, s))
    return -1;
  c.wantInteractions = interactions != nullptr;
  int rc = dispatch_step(f, kStepForcesOnly);
  c.wantInteractions = 0;
  if (rc)
    return -1;
  f->treeFresh = c.prm.barnesHut != 0;
  GVH_CHECK(hipStreamSynchronize(c.stream));
  if (download_real(f, c.disp, 2 * n, disp) ||
      download_real(f, c.sf, 2 * n, structForce) ||
      download_real(f, c.attMag, n, attMag) ||
      download_real(f, c.repMag, n, repMag))
    return -1;
  if (interactions)
    GVH_CHECK(hipMemcpy(interactions, c.interactions.p, sizeof(int32_t) * n,
                        hipMemcpyDeviceToHost));
  GVH_CHECK(hipMemcpyAsync(c.sf.p, f->sfBackup.p, 2 * n * rs,
                           hipMemcpyDeviceToDevice, c.stream));
  Scalars s2;
  if (read_scalars(f, &s2))
    return -1;
  if (s2.traversalOverflow) {
    set_error("gvizhip_force_evaluate: traversal stack overflow");
    return -1;
  }
  s2.stopped = stopped;
  return write_scalars(f, s2);
}

int gvizhip_force_get_speed(gvizhip_force *f, double *globalSpeed,
                            double *speedEfficiency, int64_t *iteration) {
  if (!f) {
    set_error("gvizhip_force_get_speed: null context");
    return -1;
  }
  Scalars s;
  if (read_scalars(f, &s))
    return -1;
  if (globalSpeed)
    *globalSpeed = s.globalSpeed;
  if (speedEfficiency)
    *speedEfficiency = s.speedEfficiency;
  if (iteration)
    *iteration = s.iteration;
  return 0;
}

int gvizhip_force_set_speed(gvizhip_force *f, double globalSpeed,
                            double speedEfficiency) {
  if (!f) {
    set_error("gvizhip_force_set_speed: null context");
    return -1;
  }
  Scalars s;
  if (read_scalars(f, &s))
    return -1;
  s.globalSpeed = globalSpeed;
  s.speedEfficiency = speedEfficiency;
  return write_scalars(f, s);
}

// Exact checkpoint: a header, the FP64 positions and the previous step's
// structForce in the engine's precision.
namespace {
struct StateHeader {
  char magic[8];
  int64_t n;
  int32_t realSize;
  int32_t reserved;
  double globalSpeed, speedEfficiency;
  int64_t iteration;
};
constexpr char kStateMagic[8] = {'G', 'V', 'H', 'F', 'S', 'T', '0', '1'};
} // namespace

int64_t gvizhip_force_state_bytes(gvizhip_force *f) {
  if (!f) {
    set_error("gvizhip_force_state_bytes: null context");
    return -1;
  }
  return (int64_t)sizeof(StateHeader) + 2 * f->c.n * (8 + (int64_t)real_size(f->c));
}

int gvizhip_force_save_state(gvizhip_force *f, void *buf, int64_t bytes) {
  if (!f || !buf || bytes < gvizhip_force_state_bytes(f)) {
    set_error("gvizhip_force_save_state: null argument or buffer below "
              "gvizhip_force_state_bytes");
    return -1;
  }
  ForceCommon &c = f->c;
  Scalars s;
  if (read_scalars(f, &s))
    return -1;
  StateHeader h;
  memcpy(h.magic, kStateMagic, 8);
  h.n = c.n;
  h.realSize = (int32_t)real_size(c);
  h.reserved = 0;
  h.globalSpeed = s.globalSpeed;
  h.speedEfficiency = s.speedEfficiency;
  h.iteration = s.iteration;
  unsigned char *p = (unsigned char *)buf;
  memcpy(p, &h, sizeof h);
  p += sizeof h;
  if (c.n > 0) {
    GVH_CHECK(hipMemcpy(p, c.pos64.p, 16 * c.n, hipMemcpyDeviceToHost));
    GVH_CHECK(hipMemcpy(p + 16 * c.n, c.sf.p, 2 * c.n * h.realSize,
                        hipMemcpyDeviceToHost));
  }
  return 0;
}

int gvizhip_force_load_state(gvizhip_force *f, const void *buf, int64_t bytes) {
  StateHeader h;
  if (!f || !buf || bytes < (int64_t)sizeof h) {
    set_error("gvizhip_force_load_state: invalid arguments");
    return -1;
  }
  memcpy(&h, buf, sizeof h);
  ForceCommon &c = f->c;
  f->treeFresh = false;
  if (memcmp(h.magic, kStateMagic, 8) != 0 || h.n != c.n ||
      (h.realSize != 4 && h.realSize != 8) ||
      bytes < (int64_t)sizeof h + 2 * h.n * (8 + (int64_t)h.realSize)) {
    set_error("gvizhip_force_load_state: not a state of an engine with %lld vertices",
              (long long)c.n);
    return -1;
  }
  const unsigned char *p = (const unsigned char *)buf + sizeof h;
  if (c.n > 0) {
    GVH_CHECK(hipMemcpy(c.pos64.p, p, 16 * c.n, hipMemcpyHostToDevice));
    const size_t rs = real_size(c);
    const int64_t cnt = 2 * c.n;
    if ((size_t)h.realSize == rs) {
      GVH_CHECK(hipMemcpy(c.sf.p, p + 16 * c.n, cnt * rs, hipMemcpyHostToDevice));
    } else {
      // saved in the other precision: converted through FP64
      std::vector<double> d(cnt);
      to_double(p + 16 * c.n, (size_t)h.realSize, cnt, d.data());
      if (rs == 8) {
        GVH_CHECK(hipMemcpy(c.sf.p, d.data(), 8 * cnt, hipMemcpyHostToDevice));
      } else {
        std::vector<float> fl(cnt);
        for (int64_t i = 0; i < cnt; i++)
          fl[i] = (float)d[i];
        GVH_CHECK(hipMemcpy(c.sf.p, fl.data(), 4 * cnt, hipMemcpyHostToDevice));
      }
    }
  }
  Scalars s;
  if (read_scalars(f, &s))
    return -1;
  s.globalSpeed = h.globalSpeed;
  s.speedEfficiency = h.speedEfficiency;
  s.iteration = h.iteration;
  s.recordBase = h.iteration;
  s.stopped = 0;
  s.epsilon = -1.0;
  return write_scalars(f, s);
}

} // extern "C"

// ------------------------------------------------------ quadtree readback
namespace {

struct TreeHost {
  int64_t n = 0;
  int maxDepth = 0;
  int64_t npc = 1;
  double root[4] = {0, 0, 0, 0};
  std::vector<uint32_t> vals;
  std::vector<uint64_t> hi, lo;
  std::vector<int32_t> left, right, parent, first, last, delta;
  std::vector<double> pm, pxh, pxl, pyh, pyl;
};

int read_tree(gvizhip_force *f, TreeHost &t) {
  ForceCommon &c = f->c;
  const int64_t n = c.n;
  t.n = n;
  t.npc = c.prm.nodesPerCell > 0 ? c.prm.nodesPerCell : 1;
  if (n == 0)
    return 0;
  Scalars s;
  if (read_scalars(f, &s))
    return -1;
  s.stopped = 0;
  if (write_scalars(f, s))
    return -1;
  if (dispatch_step(f, kStepTreeOnly))
    return -1;
  f->treeFresh = false; // now the tree of the current positions
  GVH_CHECK(hipStreamSynchronize(c.stream));
  if (read_scalars(f, &s))
    return -1;
  t.maxDepth = s.maxDepth;
  t.root[0] = s.rootCx;
  t.root[1] = s.rootCy;
  t.root[2] = s.rootHalf;
  t.root[3] = s.maxDepth;
  auto get = [&](auto &vec, const auto &buf, int64_t cnt) -> int {
    vec.resize(cnt);
    GVH_CHECK(hipMemcpy(vec.data(), buf.p, sizeof(vec[0]) * cnt,
                        hipMemcpyDeviceToHost));
    return 0;
  };
  const int64_t ni = n - 1 > 0 ? n - 1 : 0;
  if (get(t.vals, c.vals, n) || get(t.hi, c.keyHi, n) || get(t.lo, c.keyLoS, n) ||
      get(t.left, c.left, ni) || get(t.right, c.right, ni) ||
      get(t.parent, c.parent, ni) || get(t.first, c.first, ni) ||
      get(t.last, c.last, ni) || get(t.delta, c.delta, ni) ||
      get(t.pm, c.pm, n + 1) || get(t.pxh, c.pmxHi, n + 1) ||
      get(t.pxl, c.pmxLo, n + 1) || get(t.pyh, c.pmyHi, n + 1) ||
      get(t.pyl, c.pmyLo, n + 1))
    return -1;
  return 0;
}

inline int qd(const TreeHost &t, int32_t b) {
  int q = t.delta[b] >> 1;
  return q < t.maxDepth ? q : t.maxDepth;
}

inline void truncate_key(uint64_t hi, uint64_t lo, int depth, uint64_t *oh,
                         uint64_t *ol) {
  if (depth <= 0) {
    *oh = 0;
    *ol = 0;
  } else if (depth < 32) {
    *oh = hi & ~(~0ull >> (2 * depth));
    *ol = 0;
  } else if (depth == 32) {
    *oh = hi;
    *ol = 0;
  } else if (depth < 64) {
    *oh = hi;
    *ol = lo & ~(~0ull >> (2 * (depth - 32)));
  } else {
    *oh = hi;
    *ol = lo;
  }
}

// Walks the quad heads from the root exactly as k_quad lays them out.
template <class Leaf, class Cell>
void walk(const TreeHost &t, Leaf onLeaf, Cell onCell) {
  if (t.n == 0)
    return;
  if (t.n == 1) {
    onLeaf(0, 0, 0);
    return;
  }
  struct Item {
    int32_t ref;
    int parentQ;
  };
  std::vector<Item> stack;
  stack.push_back({0, -1});
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    if (it.ref < 0) {
      int32_t s = -it.ref - 1;
      onLeaf(s, s, it.parentQ + 1);
      continue;
    }
    int32_t b = it.ref;
    int q = qd(t, b);
    int64_t count = (int64_t)t.last[b] - t.first[b] + 1;
    if (count <= t.npc || q >= t.maxDepth) {
      int depth = count <= t.npc ? it.parentQ + 1 : t.maxDepth;
      onLeaf(t.first[b], t.last[b], depth);
      continue;
    }
    onCell(b, q);
    int32_t kids[2] = {t.left[b], t.right[b]};
    std::vector<int32_t> ch;
    for (int k = 0; k < 2; k++) {
      int32_t c = kids[k];
      if (c < 0 || qd(t, c) > q) {
        ch.push_back(c);
      } else {
        ch.push_back(t.left[c]);
        ch.push_back(t.right[c]);
      }
    }
    for (auto r = ch.rbegin(); r != ch.rend(); ++r)
      stack.push_back({*r, q});
  }
}

} // namespace

extern "C" int gvizhip_force_quadtree_leaves(gvizhip_force *f, int32_t *leafDepth,
                                             uint64_t *leafKeyHi,
                                             uint64_t *leafKeyLo, double *root) {
  if (!f) {
    set_error("gvizhip_force_quadtree_leaves: null context");
    return -1;
  }
  TreeHost t;
  if (read_tree(f, t))
    return -1;
  walk(
      t,
      [&](int32_t s0, int32_t s1, int depth) {
        for (int32_t s = s0; s <= s1; s++) {
          uint32_t v = t.vals[s];
          uint64_t h, l;
          truncate_key(t.hi[s], t.lo[s], depth, &h, &l);
          if (leafDepth)
            leafDepth[v] = depth;
          if (leafKeyHi)
            leafKeyHi[v] = h;
          if (leafKeyLo)
            leafKeyLo[v] = l;
        }
      },
      [&](int32_t, int) {});
  if (root)
    for (int k = 0; k < 4; k++)
      root[k] = t.root[k];
  return 0;
}

extern "C" int64_t gvizhip_force_quadtree_cells(gvizhip_force *f, int64_t cap,
                                                int32_t *depth, uint64_t *keyHi,
                                                uint64_t *keyLo, double *mass,
                                                double *comX, double *comY,
                                                int64_t *count) {
  if (!f) {
    set_error("gvizhip_force_quadtree_cells: null context");
    return -1;
  }
  TreeHost t;
  if (read_tree(f, t))
    return -1;
  int64_t k = 0;
  walk(
      t, [&](int32_t, int32_t, int) {},
      [&](int32_t b, int q) {
        if (k < cap) {
          int32_t f0 = t.first[b], l0 = t.last[b];
          double M = t.pm[l0 + 1] - t.pm[f0];
          // same double-double difference as the device records
          double sxh = t.pxh[l0 + 1], sxl = t.pxl[l0 + 1];
          double syh = t.pyh[l0 + 1], syl = t.pyl[l0 + 1];
          auto ddsum = [](double ah, double al, double bh, double bl) {
            double s = ah + bh;
            double bb = s - ah;
            double e = (ah - (s - bb)) + (bh - bb);
            e = e + al + bl;
            double r = s + e;
            double rl = e - (r - s);
            return r + rl;
          };
          double sx = ddsum(sxh, sxl, -t.pxh[f0], -t.pxl[f0]);
          double sy = ddsum(syh, syl, -t.pyh[f0], -t.pyl[f0]);
          uint64_t h, l;
          truncate_key(t.hi[f0], t.lo[f0], q, &h, &l);
          if (depth)
            depth[k] = q;
          if (keyHi)
            keyHi[k] = h;
          if (keyLo)
            keyLo[k] = l;
          if (mass)
            mass[k] = M;
          if (comX)
            comX[k] = sx / M;
          if (comY)
            comY[k] = sy / M;
          if (count)
            count[k] = (int64_t)l0 - f0 + 1;
        }
        k++;
      });
  return k;
}

extern "C" int gvizhip_force_profile(gvizhip_force *f, int64_t steps,
                                     double *traversalMs, int64_t *interactions,
                                     double *stepMs) {
  if (!f || steps < 0) {
    set_error("gvizhip_force_profile: invalid arguments");
    return -1;
  }
  ForceCommon &c = f->c;
  if (!c.travBegin) {
    GVH_CHECK(hipEventCreate(&c.travBegin));
    GVH_CHECK(hipEventCreate(&c.travEnd));
  }
  if (getenv("GVIZHIP_TRAVERSAL_STATS") && !c.travStats.p) {
    if (c.travStats.alloc(8))
      return -1;
    f->graphValid = false;
  }
  if (c.travStats.p)
    GVH_CHECK(hipMemset(c.travStats.p, 0, 8 * sizeof(unsigned long long)));
  hipEvent_t s0, s1;
  GVH_CHECK(hipEventCreate(&s0));
  GVH_CHECK(hipEventCreate(&s1));
  hipLaunchKernelGGL(k_begin_run, dim3(1), dim3(1), 0, c.stream, c.sc.p, -1.0);
  double trav = 0.0, step = 0.0;
  int64_t inter = 0;
  std::vector<int32_t> cnt(c.n > 0 ? c.n : 1);
  int rc = 0;
  for (int64_t k = 0; k < steps && rc == 0; k++) {
    c.timeTraversal = 1;
    c.wantInteractions = 1;
    if (hipEventRecord(s0, c.stream) != hipSuccess)
      rc = -1;
    if (rc == 0)
      rc = dispatch_step(f, kStepFull);
    if (rc == 0 && hipEventRecord(s1, c.stream) != hipSuccess)
      rc = -1;
    c.timeTraversal = 0;
    c.wantInteractions = 0;
    f->treeFresh = c.prm.barnesHut != 0;
    if (rc == 0 && hipStreamSynchronize(c.stream) != hipSuccess)
      rc = -1;
    float ms = 0.0f;
    if (rc == 0 && c.n > 0) {
      (void)hipEventElapsedTime(&ms, c.travBegin, c.travEnd);
      trav += ms;
    }
    if (rc == 0) {
      (void)hipEventElapsedTime(&ms, s0, s1);
      step += ms;
    }
    if (rc == 0 && c.n > 0 &&
        hipMemcpy(cnt.data(), c.interactions.p, sizeof(int32_t) * c.n,
                  hipMemcpyDeviceToHost) != hipSuccess)
      rc = -1;
    for (int64_t i = 0; rc == 0 && i < c.n; i++)
      inter += cnt[i];
  }
  (void)hipEventDestroy(s0);
  (void)hipEventDestroy(s1);
  if (rc) {
    if (!gvh::last_error()[0])
      set_error("gvizhip_force_profile failed");
    return -1;
  }
  if (c.travStats.p) {
    unsigned long long t[8];
    if (hipMemcpy(t, c.travStats.p, sizeof(t), hipMemcpyDeviceToHost) == hipSuccess)
      fprintf(stderr,
              "[gvizhip] traversal stats: steps %llu, deferred %llu (mask lanes %llu),"
              " undecided %llu (mask lanes %llu), max stack %llu; far form %llu"
              " (%llu with full masks)\n",
              t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7]);
  }
  if (traversalMs)
    *traversalMs = trav;
  if (interactions)
    *interactions = inter;
  if (stepMs)
    *stepMs = step;
  return 0;
}

extern "C" void gvizhip_force_destroy(gvizhip_force *f) {
  if (!f)
    return;
  drop_graph(f);
  if (f->c.travBegin)
    (void)hipEventDestroy(f->c.travBegin);
  if (f->c.travEnd)
    (void)hipEventDestroy(f->c.travEnd);
  f->c.close_streams();
  if (f->pinned)
    (void)hipHostFree(f->pinned);
  delete f;
}

// ------------------------------------------------ quadtree keys for the host
namespace {

// Copies the tree buffers of c (root cell, key order, sorted keys) to the
// host; the keyLo buffer is read in key order from `keyLoSorted`, or
// permuted into it from the per-point buffer when that is null.
int download_keys(ForceCommon &c, Scalars *pinned, const uint64_t *keyLoSorted,
                  double *root, uint32_t *order, uint64_t *keyHi, uint64_t *keyLo) {
  const int64_t n = c.n;
  GVH_CHECK(hipMemcpyAsync(pinned, c.sc.p, sizeof(Scalars), hipMemcpyDeviceToHost,
                           c.stream));
  GVH_CHECK(hipMemcpyAsync(order, c.vals.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                           c.stream));
  GVH_CHECK(hipMemcpyAsync(keyHi, c.keyHi.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost,
                           c.stream));
  if (keyLoSorted) {
    GVH_CHECK(hipMemcpyAsync(keyLo, keyLoSorted, sizeof(uint64_t) * n,
                             hipMemcpyDeviceToHost, c.stream));
    GVH_CHECK(hipStreamSynchronize(c.stream));
  } else {
    std::vector<uint64_t> lo(n);
    GVH_CHECK(hipMemcpyAsync(lo.data(), c.keyLo.p, sizeof(uint64_t) * n,
                             hipMemcpyDeviceToHost, c.stream));
    GVH_CHECK(hipStreamSynchronize(c.stream));
    for (int64_t s = 0; s < n; s++)
      keyLo[s] = lo[order[s]];
  }
  root[0] = pinned->rootCx;
  root[1] = pinned->rootCy;
  root[2] = pinned->rootHalf;
  root[3] = (double)pinned->maxDepth;
  return 0;
}

} // namespace

extern "C" int gvizhip_force_quadtree_keys(gvizhip_force *f, double *root,
                                           uint32_t *order, uint64_t *keyHi,
                                           uint64_t *keyLo, double *points) {
  if (!f || !root || (f->c.n > 0 && (!order || !keyHi || !keyLo))) {
    set_error("gvizhip_force_quadtree_keys: invalid arguments");
    return -1;
  }
  ForceCommon &c = f->c;
  const int64_t n = c.n;
  if (n == 0)
    return 0;
  if (!f->treeFresh) {
    Scalars s;
    if (read_scalars(f, &s))
      return -1;
    s.stopped = 0;
    if (write_scalars(f, s) || dispatch_step(f, kStepTreeOnly))
      return -1;
    f->treeFresh = true;
  }
  if (download_keys(c, f->pinned, c.keyLoS.p, root, order, keyHi, keyLo))
    return -1;
  if (points) {
    std::vector<double> sp(2 * n);
    GVH_CHECK(hipMemcpy(sp.data(), c.spos64.p, sizeof(double) * 2 * n,
                        hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < n; s++) {
      points[2 * (int64_t)order[s]] = sp[2 * s];
      points[2 * (int64_t)order[s] + 1] = sp[2 * s + 1];
    }
  }
  return 0;
}

struct gvizhip_qtree {
  ForceCommon c;
  Scalars *pinned = nullptr;
};

extern "C" int gvizhip_qtree_create(gvizhip_qtree **out) {
  if (!out) {
    set_error("gvizhip_qtree_create: null output");
    return -1;
  }
  *out = nullptr;
  gvizhip_qtree *q = new gvizhip_qtree();
  ForceCommon &c = q->c;
  if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&q->pinned, sizeof(Scalars), hipHostMallocDefault) !=
          hipSuccess) {
    set_error("gvizhip_qtree_create: stream or pinned allocation failed");
    gvizhip_qtree_destroy(q);
    return -1;
  }
  if (c.sc.alloc(1) || c.partials.alloc(4 * 1024)) {
    gvizhip_qtree_destroy(q);
    return -1;
  }
  // k_keys reads the stop flag; nothing here ever sets it
  if (hipMemset(c.sc.p, 0, sizeof(Scalars)) != hipSuccess) {
    set_error("gvizhip_qtree_create: scalar reset failed");
    gvizhip_qtree_destroy(q);
    return -1;
  }
  *out = q;
  return 0;
}

extern "C" int gvizhip_qtree_build(gvizhip_qtree *q, int64_t count, const double *points,
                                   double *root, uint32_t *order, uint64_t *keyHi,
                                   uint64_t *keyLo) {
  if (!q || count < 0 || !root ||
      (count > 0 && (!points || !order || !keyHi || !keyLo))) {
    set_error("gvizhip_qtree_build: invalid arguments");
    return -1;
  }
  if (count > (int64_t)UINT32_MAX) {
    set_error("gvizhip_qtree_build: %lld points exceed the 32-bit point ids",
              (long long)count);
    return -1;
  }
  if (count == 0)
    return 0;
  ForceCommon &c = q->c;
  c.n = count;
  if (c.pos64.alloc(2 * count) || c.keyHi.alloc(count) || c.keyLo.alloc(count) ||
      c.vals.alloc(count) || c.sorter.reserve(count))
    return -1;
  GVH_CHECK(hipMemcpyAsync(c.pos64.p, points, sizeof(double) * 2 * count,
                           hipMemcpyHostToDevice, c.stream));
  if (enqueue_keys_and_sort(c))
    return -1;
  return download_keys(c, q->pinned, nullptr, root, order, keyHi, keyLo);
}

extern "C" void gvizhip_qtree_destroy(gvizhip_qtree *q) {
  if (!q)
    return;
  if (q->c.stream) {
    (void)hipStreamSynchronize(q->c.stream);
    (void)hipStreamDestroy(q->c.stream);
    q->c.stream = nullptr;
  }
  if (q->pinned)
    (void)hipHostFree(q->pinned);
  delete q;
}
