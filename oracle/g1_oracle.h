/* g1_oracle.h -- FP64 CPU oracle for the G1 fused step.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or the
 * timed CPU baseline).  The product (libg1sim_amd.so, g1sim_amd.G1Sim) never
 * links, loads or falls back to it.
 *
 * It restates the reference algorithm (src/g1_core.h, src/spatial.h) in plain
 * C, reentrant (all scratch on the caller's stack), in float64 by default.
 * The first four symbols keep the single-env host C-ABI of the reference's
 * FP64 validation library (src/host_lib.cpp:7-53, built as libg1host.so), so
 * code that binds that library binds this one unchanged.
 */
#ifndef G1_ORACLE_H
#define G1_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* src/host_lib.cpp:7-10: sizes; returns sizeof(real) (8 for the oracle). */
int g1_sizes(int* nb, int* nq, int* nv, int* nu, int* nc);
/* src/host_lib.cpp:13-22: world position (30x3) and rotation (30x9 row-major). */
void g1_c_fk(const double* qpos, double* xpos, double* xmat);
/* src/host_lib.cpp:25-35: smooth forward dynamics, qacc from (qpos, qvel, tau). */
void g1_c_fd(const double* qpos, const double* qvel, const double* tau, double* qacc);
/* src/host_lib.cpp:38-53: nsub fused steps (contacts + PD + limits) at dt. */
void g1_c_step(double* qpos, double* qvel, const double* ctrl, int nsub,
               double dt, double* fn_out, double* anchor);

/* ---- oracle-only extensions (no reference counterpart) ---------------- */

/* One env, nsub fused steps with every output of the GPU step:
 * mu = foot friction (the model's 1.0 reproduces g1_c_step);
 * qacc_out[35], fn_out[8], foot_pos[6], foot_vel[6], tau_out[35] refer to the
 * last substep (foot data from its pre-integration kinematics, like
 * src/g1_kernel.cu:128-145); tau_out is the total generalized force
 * (PD + joint limits on hinges, J^T f of the contacts on all dofs).
 * Every output pointer is nullable. */
void g1o_step(double* qpos, double* qvel, const double* ctrl, double* anchor,
              int nsub, double dt, double mu, double* qacc_out, double* fn_out,
              double* foot_pos, double* foot_vel, double* tau_out);

/* Batched teacher-forcing check: nenv envs in float SoA ([dim][nenv]) are cast
 * to float64 and stepped once each (nsub = 1).  Outputs are float64 SoA and
 * nullable.  mu is per env (nullable -> model value).  nthreads <= 0 -> all. */
void g1o_step_batch_f32(int nenv, double dt, const float* qpos, const float* qvel,
                        const float* anchor, const float* ctrl, const float* mu,
                        double* qacc_out, double* qpos_out, double* qvel_out,
                        double* anchor_out, double* fn_out, double* foot_pos,
                        double* foot_vel, double* tau_out, int nthreads);

/* Recursive Newton-Euler inverse dynamics (independent of ABA): tau for given
 * (qpos, qvel, qacc) in the same generalized-coordinate conventions as
 * g1_c_fd (base force world frame, base torque body frame). */
void g1o_rnea(const double* qpos, const double* qvel, const double* qacc, double* tau);
/* Composite-rigid-body joint-space inertia M (35x35 row-major), armature on
 * the hinge diagonal. */
void g1o_crba(const double* qpos, double* M);
/* Packed 6x6 SPD solve exactly as the base solve of the step (21 floats,
 * lower-triangular row-major). */
void g1o_m6_solve(const double* A21, const double* b6, double* x6);
/* Quaternion integration q <- q (x) exp(w dt / 2), normalised. */
void g1o_quat_integrate(double* q4, const double* w3, double dt);

/* Integer hash + uniform draw of the reset kernel (src/g1_kernel.cu:54-62). */
unsigned g1o_hash(unsigned x);
/* Reset of one env (src/g1_kernel.cu:85-103) evaluated in float32 with the
 * noise term as one fused multiply-add: writes qpos[36], qvel[35], anchor[16]. */
void g1o_reset_env_f32(int env, unsigned seed, float noise, float drop,
                       float* qpos, float* qvel, float* anchor);

/* Timed CPU baseline: nsteps fused steps (nsub = 1) of nenv envs in float32
 * SoA with nthreads OpenMP threads; returns wall seconds.  Built from the
 * same source with real = float (see Makefile). */
double g1o_bench_f32(int nenv, int nsteps, double dt, float* qpos, float* qvel,
                     float* anchor, const float* ctrl, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
