// ref_bench.cpp -- CPU-baseline driver around the reference physics core
// (TEST INFRASTRUCTURE: built into oracle/_ref/, loaded only by bench.py's
// cpu_baseline leg and by tests).
//
// Compiled against the reference's own headers (src/g1_core.h with
// G1_PRECISION=4) where they lie; nothing of the reference is copied.  It
// steps nenv envs in float SoA ([dim][nenv]) with one OpenMP iteration per
// env and a per-iteration workspace, which is what the reference's CPU shim
// does (src/g1_kernel.cu:25-37, 106-151 under Makefile:21-24).
#include "g1_core.h"
#include <omp.h>
#include <cstddef>

extern "C" double g1ref_bench_f32(int nenv, int nsteps, double dt, float* qpos, float* qvel,
                                  float* anchor, const float* ctrl, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  G1Config cfg = g1_default_config();
  cfg.dt = G1Real(dt);
  double t0 = omp_get_wtime();
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int e = 0; e < nenv; ++e) {
    const size_t N = (size_t)nenv;
    G1Real qp[G1_NQ], qv[G1_NV], ct[G1_NU], an[2 * G1_NC];
    for (int i = 0; i < G1_NQ; ++i) qp[i] = qpos[i * N + e];
    for (int i = 0; i < G1_NV; ++i) qv[i] = qvel[i * N + e];
    for (int i = 0; i < G1_NU; ++i) ct[i] = ctrl[i * N + e];
    for (int i = 0; i < 2 * G1_NC; ++i) an[i] = anchor[i * N + e];
    G1Ws w;
    for (int s = 0; s < nsteps; ++s) g1_step(cfg, qp, qv, ct, an, w, nullptr);
    for (int i = 0; i < G1_NQ; ++i) qpos[i * N + e] = qp[i];
    for (int i = 0; i < G1_NV; ++i) qvel[i * N + e] = qv[i];
    for (int i = 0; i < 2 * G1_NC; ++i) anchor[i * N + e] = an[i];
  }
  return omp_get_wtime() - t0;
}
