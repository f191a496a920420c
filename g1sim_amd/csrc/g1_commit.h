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

// Judge part of the commit for one env per lane: the network's qacc sur(j)
// against ABA's aba(j) over all dofs (read only for judged and fallback
// envs), the env's new fallback flag, and the wave's statistics.  `live`
// lanes are envs < nenv; every lane of the wave must call this (the
// statistics are reduced across the wave).  stats: [0] judged, [1] rejected,
// [2] max judged rel. error (float bits, non-negative, so integer max orders
// them), [3] fallback env-steps.
template <typename SUR, typename ABA>
G1K_DEV void surrogate_judge(bool live, bool judged, bool fb, int e, float tol, SUR sur, ABA aba,
                             uint8_t* fallback, unsigned* stats) {
  constexpr int NV = 35;
  float err = 0.f;
  if (judged || fb) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const float ref = aba(j);
      err = fmaxf(err, fabsf(sur(j) - ref) * frcp(fmaxf(fabsf(ref), 1.0f)));
    }
  }
  const bool reject = judged && !(err <= tol);
  if (judged && fallback) fallback[e] = reject ? 1 : 0;
  // wave-level statistics: one atomic per counter per wave
  const unsigned n_judged = (unsigned)__popcll(__ballot(judged));
  const unsigned n_reject = (unsigned)__popcll(__ballot(reject));
  const unsigned n_fb = (unsigned)__popcll(__ballot(live && fb && !judged));
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
}

// Semi-implicit Euler of the floating base (src/g1_core.h:212-222 with the
// quaternion exponential of src/spatial.h:73-79): velocities += dt qacc, then
// position += dt v and q <- q (x) exp(w dt / 2), normalised.  qa: qacc 0-5
// (world linear, body angular); vw / om: world linear / body angular
// velocity; bp / bq: base position and quaternion (float, or double for the
// float64 base-pose accumulation).  Every multiply-add is an explicit fma in a
// fixed order, so the fused step and the surrogate's commit, separate kernel
// instances, produce the same bits from the same inputs.
template <typename BT>
G1K_DEV void euler_base(float dt, const float* qa, float (&vw)[3], float (&om)[3], BT (&bp)[3], BT (&bq)[4]) {
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    vw[m] = fmaf(dt, qa[m], vw[m]);
    om[m] = fmaf(dt, qa[3 + m], om[m]);
    bp[m] = fma((BT)dt, (BT)vw[m], bp[m]);
  }
  const float wn = fsqrt(fmaf(om[2], om[2], fmaf(om[1], om[1], om[0] * om[0])));
  const float angle = wn * dt;
  if (angle >= 1e-12f) {
    float sh, ch;
    fsincos(0.5f * angle, &sh, &ch);
    const float s = sh * frcp(angle / dt);
    const BT dw = ch, dx = s * om[0], dy = s * om[1], dz = s * om[2];
    const BT nw = fma(-bq[3], dz, fma(-bq[2], dy, fma(-bq[1], dx, bq[0] * dw)));
    const BT nx = fma(-bq[3], dy, fma(bq[2], dz, fma(bq[1], dw, bq[0] * dx)));
    const BT ny = fma(bq[3], dx, fma(bq[2], dw, fma(-bq[1], dz, bq[0] * dy)));
    const BT nz = fma(bq[3], dw, fma(-bq[2], dx, fma(bq[1], dy, bq[0] * dz)));
    const BT inv = frsq(fma(nz, nz, fma(ny, ny, fma(nx, nx, nw * nw))));
    bq[0] = nw * inv; bq[1] = nx * inv; bq[2] = ny * inv; bq[3] = nz * inv;
  }
}

// Semi-implicit Euler of one hinge: qd += dt qdd, then q += dt qd (the same
// fma order as euler_base's convention, shared by the step and the commit).
G1K_DEV void euler_hinge(float dt, float qdd, float& qd, float& q) {
  qd = fmaf(dt, qdd, qd);
  q = fmaf(dt, qd, q);
}

// euler_base on an env's state read through qp(m) / qv(m) (its qpos / qvel
// row m), written to rows 0-6 of qpos and 0-5 of qvel ([dim][N]).
template <typename QA, typename QP, typename QV>
G1K_DEV void integrate_base(int e, size_t N, float dt, QA qa, QP qp, QV qv, float* qpos, float* qvel) {
  float a[6], vw[3], om[3], bp[3], bq[4];
#pragma unroll
  for (int m = 0; m < 6; ++m) a[m] = qa(m);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    vw[m] = qv(m);
    om[m] = qv(3 + m);
    bp[m] = qp(m);
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) bq[m] = qp(3 + m);
  euler_base(dt, a, vw, om, bp, bq);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    qvel[m * N + e] = vw[m];
    qvel[(3 + m) * N + e] = om[m];
    qpos[m * N + e] = bp[m];
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) qpos[(3 + m) * N + e] = bq[m];
}

// Semi-implicit Euler of hinge h (0-28): qvel row 6 + h, qpos row 7 + h.
G1K_DEV void integrate_hinge(int h, int e, size_t N, float dt, float qa, float qp, float qv, float* qpos,
                             float* qvel) {
  float v = qv, q = qp;
  euler_hinge(dt, qa, v, q);
  qvel[(6 + h) * N + e] = v;
  qpos[(7 + h) * N + e] = q;
}

// The whole commit of one env per lane: qacc = ABA's for judged and fallback
// envs, else the network's sur[35] (physical units); judge and count; store
// qacc_out (nullable); integrate.  aba(j) is the env's ABA qacc j (read only
// for judged and fallback envs), qp(m) / qv(m) its qpos / qvel row m; the
// results go to qpos / qvel ([dim][N]).  Every lane of the wave must call
// this.
template <typename ABA, typename QP, typename QV>
G1K_DEV void surrogate_commit(bool live, int e, size_t N, int phase, int period, float tol, float dt,
                              const float* sur, ABA aba, uint8_t* fallback, unsigned* stats, float* qacc_out,
                              QP qp, QV qv, float* qpos, float* qvel) {
  constexpr int NV = 35, NU = 29;
  const bool judged = live && judged_env(e, phase, period);
  const bool fb = live && fallback && fallback[e];
  surrogate_judge(live, judged, fb, e, tol, [&](int j) { return live ? sur[j] : 0.f; }, aba, fallback, stats);
  if (!live) return;
  const bool use_aba = judged || fb;
  float qa[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) qa[j] = use_aba ? aba(j) : sur[j];
  if (qacc_out) {
#pragma unroll
    for (int j = 0; j < NV; ++j) qacc_out[j * N + e] = qa[j];
  }
  integrate_base(e, N, dt, [&](int j) { return qa[j]; }, qp, qv, qpos, qvel);
#pragma unroll
  for (int h = 0; h < NU; ++h) integrate_hinge(h, e, N, dt, qa[6 + h], qp(7 + h), qv(6 + h), qpos, qvel);
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
