// g1_commit.h -- per-env commit of one surrogate step: choose the qacc (the
// network's, or ABA's for judged and fallback envs), judge, count, and
// integrate semi-implicitly (src/g1_core.h:212-222 with the quaternion
// exponential of src/spatial.h:73-79).  Shared by the stand-alone integrate
// kernel and the fused surrogate kernel, so both give the same bits.
#pragma once
#include <stdint.h>

#include "g1_fastmath.h"

namespace g1k {

// The rotating judge works on blocks of JUDGE_BLOCK consecutive envs (one
// wavefront of the step kernel, one row tile of the fused network): block b
// is judged at the steps where b % period == phase, so every env is judged
// once per `period` steps, the judge's state loads stay coalesced and whole
// network tiles can be skipped.
constexpr int JUDGE_BLOCK = 64;
G1K_DEV bool judged_env(int e, int phase, int period) { return ((e / JUDGE_BLOCK) % period) == phase; }

// `live` lanes are envs < nenv; every lane of the wave must call this (the
// statistics are reduced across the wave).  sur[35] is the network's qacc in
// physical units.  The env's inputs come through accessors: aba(j) its ABA
// qacc j (read only for judged and fallback envs), qp(m) / qv(m) its qpos /
// qvel row m; the results go to qpos / qvel ([dim][N]).  stats: [0] judged,
// [1] rejected, [2] max judged rel. error (float bits, non-negative, so
// integer max orders them), [3] fallback env-steps.
template <typename ABA, typename QP, typename QV>
G1K_DEV void surrogate_commit(bool live, int e, size_t N, int phase, int period, float tol, float dt,
                              const float* sur, ABA aba, uint8_t* fallback, unsigned* stats, float* qacc_out,
                              QP qp, QV qv, float* qpos, float* qvel) {
  constexpr int NV = 35, NU = 29;
  const bool judged = live && judged_env(e, phase, period);
  const bool fb = live && fallback && fallback[e];
  float qa[NV];
  float err = 0.f;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    float v = live ? sur[j] : 0.f;
    if (judged || fb) {
      const float ref = aba(j);
      err = fmaxf(err, fabsf(v - ref) * frcp(fmaxf(fabsf(ref), 1.0f)));
      v = ref;
    }
    qa[j] = v;
  }
  const bool reject = judged && !(err <= tol);
  if (judged && fallback) fallback[e] = reject ? 1 : 0;
  // wave-level statistics: one atomic per counter per wave
  const unsigned n_judged = (unsigned)__popcll(__ballot(judged));
  const unsigned n_reject = (unsigned)__popcll(__ballot(reject));
  const unsigned n_fb = (unsigned)__popcll(__ballot(fb && !judged));
  unsigned ebits = judged ? __float_as_uint(err) : 0u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned o = (unsigned)__shfl_xor((int)ebits, off, 64);
    ebits = o > ebits ? o : ebits;
  }
  if ((threadIdx.x & 63) == 0 && stats) {
    if (n_judged) atomicAdd(&stats[0], n_judged);
    if (n_reject) atomicAdd(&stats[1], n_reject);
    if (n_judged) atomicMax(&stats[2], ebits);
    if (n_fb) atomicAdd(&stats[3], n_fb);
  }
  if (!live) return;
  if (qacc_out) {
#pragma unroll
    for (int j = 0; j < NV; ++j) qacc_out[j * N + e] = qa[j];
  }
  float vw[3], om[3], bp[3], bq[4];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    vw[m] = qv(m) + dt * qa[m];
    om[m] = qv(3 + m) + dt * qa[3 + m];
    bp[m] = qp(m) + dt * vw[m];
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) bq[m] = qp(3 + m);
  const float wn = fsqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
  const float angle = wn * dt;
  if (angle >= 1e-12f) {
    float sh, ch;
    fsincos(0.5f * angle, &sh, &ch);
    const float s = sh * frcp(angle / dt);
    const float dw = ch, dx = s * om[0], dy = s * om[1], dz = s * om[2];
    const float nw = bq[0] * dw - bq[1] * dx - bq[2] * dy - bq[3] * dz;
    const float nx = bq[0] * dx + bq[1] * dw + bq[2] * dz - bq[3] * dy;
    const float ny = bq[0] * dy - bq[1] * dz + bq[2] * dw + bq[3] * dx;
    const float nz = bq[0] * dz + bq[1] * dy - bq[2] * dx + bq[3] * dw;
    const float inv = frsq(nw * nw + nx * nx + ny * ny + nz * nz);
    bq[0] = nw * inv; bq[1] = nx * inv; bq[2] = ny * inv; bq[3] = nz * inv;
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    qvel[m * N + e] = vw[m];
    qvel[(3 + m) * N + e] = om[m];
    qpos[m * N + e] = bp[m];
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) qpos[(3 + m) * N + e] = bq[m];
#pragma unroll
  for (int h = 0; h < NU; ++h) {
    const float v = qv(6 + h) + dt * qa[6 + h];
    qvel[(6 + h) * N + e] = v;
    qpos[(7 + h) * N + e] = qp(7 + h) + dt * v;
  }
}

// The same with every input read from the SoA state in memory.
G1K_DEV void surrogate_commit(bool live, int e, size_t N, int phase, int period, float tol, float dt,
                              const float* sur, const float* qacc_aba, uint8_t* fallback, unsigned* stats,
                              float* qacc_out, float* qpos, float* qvel) {
  surrogate_commit(
      live, e, N, phase, period, tol, dt, sur, [&](int j) { return qacc_aba[j * N + e]; }, fallback, stats,
      qacc_out, [&](int m) { return qpos[m * N + e]; }, [&](int m) { return qvel[m * N + e]; }, qpos, qvel);
}

}  // namespace g1k
