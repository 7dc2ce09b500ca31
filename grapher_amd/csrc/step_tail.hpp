// step_tail.hpp -- the per-step scalar work shared by the single-graph
// engine and the batch engine: the FA2 global speed update and the step
// statistics, each run by one block over the per-block FP64 partials of one
// graph. Compiled without FMA contraction (force_common.hip, batch.hip) so
// that the scalar update rounds like the reference.
#pragma once

#include "force_engine.hpp"

namespace gvh {

// src/embedders/gvizForceEmbedder.c:140-185, with S and T reduced from the
// per-block FP64 partials of the force kernel.
__device__ inline void fa2_tail(const double *__restrict__ partial, int blocks,
                                int64_t n, double jitterTolerance,
                                Scalars *__restrict__ sc,
                                StepRecord *__restrict__ rec, int64_t recCap,
                                double *smem) {
  if (sc->stopped)
    return;
  double v[2] = {0.0, 0.0};
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    v[0] += partial[2 * b];
    v[1] += partial[2 * b + 1];
  }
  const bool isMax[2] = {false, false};
  block_reduce<2>(v, isMax, smem);
  if (threadIdx.x != 0)
    return;
  double totalSwinging = v[0], totalEffectiveTraction = v[1];
  int64_t slot = sc->iteration - sc->recordBase;
  if (rec && slot >= 0 && slot < recCap) {
    rec[slot].totalSwinging = totalSwinging;
    rec[slot].totalTraction = totalEffectiveTraction;
  }
  if (totalEffectiveTraction < 1e-7)
    return;
  double gs = sc->globalSpeed, eff = sc->speedEfficiency;
  double nn = (double)n;
  double estimatedOptimalJT = 0.05 * sqrt(nn);
  double minJT = sqrt(estimatedOptimalJT);
  double maxJT = 10.0;
  double jt = jitterTolerance *
              fmax(minJT, fmin(maxJT, estimatedOptimalJT *
                                          totalEffectiveTraction / (nn * nn)));
  if (totalSwinging / totalEffectiveTraction > 2.0) {
    if (eff > 0.05)
      eff *= 0.5;
    jt = fmax(jt, jitterTolerance);
  }
  double targetSpeed = totalSwinging > 1e-7
                           ? jt * eff * totalEffectiveTraction / totalSwinging
                           : gs;
  if (totalSwinging > jt * totalEffectiveTraction) {
    if (eff > 0.05)
      eff *= 0.7;
  } else if (gs < 1000.0) {
    eff *= 1.3;
  }
  if (eff < 0.05)
    eff = 0.05;
  else if (eff > 1.0)
    eff = 1.0;
  gs += fmin(targetSpeed - gs, 0.5 * gs);
  sc->globalSpeed = gs;
  sc->speedEfficiency = eff;
}

// Step tail (src/embedders/gvizForceEmbedder.c:467-494): max |applied|, mean
// |att|, mean |rep|, mean gravity; advances the iteration and, for Run,
// raises the stop flag once maxDisp < epsilon. `last` (may be null) receives
// the record whatever the record slot.
__device__ inline void stats_tail(const double *__restrict__ partial, int blocks,
                                  int64_t n, double gravityK,
                                  Scalars *__restrict__ sc,
                                  StepRecord *__restrict__ rec, int64_t recCap,
                                  StepRecord *__restrict__ last, double *smem) {
  if (sc->stopped)
    return;
  double v[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    v[0] = fmax(v[0], partial[3 * b]);
    v[1] += partial[3 * b + 1];
    v[2] += partial[3 * b + 2];
  }
  const bool isMax[3] = {true, false, false};
  block_reduce<3>(v, isMax, smem);
  if (threadIdx.x != 0)
    return;
  double maxDisp = v[0];
  StepRecord r{};
  r.maxDisp = maxDisp;
  r.speed = sc->globalSpeed;
  r.meanAtt = n ? v[1] / (double)n : 0.0;
  r.meanRep = n ? v[2] / (double)n : 0.0;
  r.meanGrav = gravityK * sc->meanDegreePlusOne;
  r.speedEfficiency = sc->speedEfficiency;
  int64_t slot = sc->iteration - sc->recordBase;
  if (rec && slot >= 0 && slot < recCap) {
    StepRecord &o = rec[slot];
    o.maxDisp = r.maxDisp;
    o.speed = r.speed;
    o.meanAtt = r.meanAtt;
    o.meanRep = r.meanRep;
    o.meanGrav = r.meanGrav;
    o.speedEfficiency = r.speedEfficiency;
  }
  if (last) {
    last->maxDisp = r.maxDisp;
    last->speed = r.speed;
    last->meanAtt = r.meanAtt;
    last->meanRep = r.meanRep;
    last->meanGrav = r.meanGrav;
    last->speedEfficiency = r.speedEfficiency;
  }
  sc->iteration++;
  if (sc->epsilon >= 0.0 && maxDisp < sc->epsilon)
    sc->stopped = 1;
}

} // namespace gvh
