/* g1sim_amd.h -- C-ABI of libg1sim_amd.so, the MI355X (gfx950) step engine
 * for the Unitree G1 (29-DoF, floating base).
 *
 * State is float32 structure-of-arrays with the env index fastest
 * (array[i * nenv + e]), exactly the layout of the reference
 * (src/g1_kernel.cu:9-10):
 *   qpos [36][nenv]  base position (world), base quaternion (w,x,y,z), 29 hinges
 *   qvel [35][nenv]  base linear velocity (world), base angular velocity (body), 29 hinges
 *   anchor [16][nenv] stick-slip anchors (x, y) of the 8 foot contacts, 1e30 = free
 *   ctrl [29][nenv]  PD position targets
 * Every device pointer is owned by the caller; state is updated in place.
 * `stream` is a hipStream_t (an opaque pointer at the ABI, as cudaStream_t is
 * in the reference); 0 means the null stream.  Work is asynchronous on it.
 */
#ifndef G1SIM_AMD_H
#define G1SIM_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Replaces g1_cuda_step (src/g1_kernel.cu:154-163), bound by ctypes at
 * src/g1sim.py:15-20 and called at src/g1sim.py:119-130.  nsub substeps at
 * dt = 1e-3 with the model's friction; foot_pos [6][nenv], foot_vel [6][nenv]
 * (world position / linear velocity of bodies 6 and 12 before the last
 * substep's integration) and contact_forces [8][nenv] (last substep's normal
 * forces) are nullable.  Like the reference it returns nothing; a failure is
 * kept for g1_hip_last_error(). */
void g1_cuda_step(int nenv, int nsub, float* qpos, float* qvel, float* anchor, const float* ctrl,
                  float* foot_pos, float* foot_vel, float* contact_forces, void* stream);

/* Replaces g1_cuda_reset_done (src/g1_kernel.cu:165-173), bound at
 * src/g1sim.py:23-33 and called at src/g1sim.py:93-103.  For every env with
 * done[e] != 0: qpos = stand keyframe, z += drop, hinges += noise * U(-1, 1)
 * from the reference's integer hash of (seed, e); qvel = 0; anchors free.
 * done is uint8 [nenv] on the device. */
void g1_cuda_reset_done(int nenv, float* qpos, float* qvel, float* anchor, const uint8_t* done,
                        unsigned seed, float noise, float drop, void* stream);

/* ---- additive entry points (no reference counterpart; the reference
 * launcher hard-codes its config at src/g1_kernel.cu:157) ---------------- */

/* Flags of g1_hip_step_ex. */
#define G1_FLAG_F64_SOLVE 1u   /* floating-base 6x6 Cholesky solve in float64 */
#define G1_FLAG_F64_BASE  2u   /* base position/quaternion accumulated in float64
                                  (requires base_pose64) */
#define G1_FLAG_F64_CONTACT 4u /* foot-contact kinematics in float64: pelvis and
                                  leg world poses and twists, contact points,
                                  penetration depth and contact-point velocity */
#define G1_FLAG_WAVES4    8u   /* force 4 waves (one limb chain each) per 64 envs */
#define G1_FLAG_WAVES2   16u   /* force 2 waves (legs, arms) per 64 envs; by
                                  default 2 from 32,768 envs, 4 below (setting
                                  both wave flags is an invalid argument) */

/* g1_cuda_step with a runtime dt and extra outputs.  Returns 0 or a negative
 * error code (-1 invalid argument, -2 launch failure; see g1_hip_last_error).
 *   qacc_out     [35][nenv] float32, qacc of the last substep (nullable)
 *   tau_out      [35][nenv] float32, total generalized force of the last
 *                substep: servo + joint-limit torques on the hinges plus J^T f
 *                of the foot contacts on every dof (base force in the world
 *                frame, base torque in the body frame, like g1_c_fd's tau);
 *                with it, the step is qacc = (M + dt B)^-1 (tau - C(q, qd))
 *                (nullable)
 *   mu           [nenv] per-env foot friction (nullable: the model's 1.0)
 *   base_pose64  [7][nenv] float64 base position + quaternion, read and
 *                written instead of qpos[0:7] for accumulation when
 *                G1_FLAG_F64_BASE is set (qpos[0:7] still receives the
 *                rounded pose); must be null otherwise. */
int g1_hip_step_ex(int nenv, int nsub, float dt, float* qpos, float* qvel, float* anchor,
                   const float* ctrl, float* foot_pos, float* foot_vel, float* contact_forces,
                   float* qacc_out, float* tau_out, const float* mu, double* base_pose64,
                   unsigned flags, void* stream);

/* A schedule of PD targets in one call: nsteps physics steps of g1_hip_step_ex
 * (no qacc / tau outputs) where step step0 + s uses the target block
 * ctrl_blocks + ((step0 + s) / period) * block_stride ([29][nenv] each,
 * block_stride >= 29 * nenv floats apart) -- the targets of the reference's
 * benchmark loop (src/g1_kernel.cu:227-238) resampled every `period` steps.
 * The library enqueues launches of up to 1,000 steps that switch target
 * blocks themselves (each block's targets loaded one substep before they
 * apply), bit-identical to calling g1_hip_step_ex block by block; a long
 * open-loop schedule pays one call and one state load/store per 1,000
 * steps.  foot_pos / foot_vel / contact_forces (nullable) hold the outputs
 * of the last step.  Returns 0 or a negative error code like
 * g1_hip_step_ex. */
int g1_hip_step_schedule(int nenv, int nsteps, int period, int step0, float dt, float* qpos, float* qvel,
                         float* anchor, const float* ctrl_blocks, long long block_stride, float* foot_pos,
                         float* foot_vel, float* contact_forces, const float* mu, double* base_pose64,
                         unsigned flags, void* stream);

/* ---- learned forward-dynamics surrogate: BF16 MFMA kernels (no reference
 * counterpart; the reference has no learned component) ------------------ */

/* Epilogues of g1_mlp_gemm. */
#define G1_MLP_EPI_F32        0  /* C = acc (fp32)                                  */
#define G1_MLP_EPI_F32_ATOMIC 1  /* C += acc (fp32 atomics; split > 1 allowed)      */
#define G1_MLP_EPI_BIAS_SILU  2  /* C = acc + bias (fp32), H = silu(C) (bf16)       */
#define G1_MLP_EPI_BIAS_SILU_H 3 /* H = silu(acc + bias) (bf16)                     */
#define G1_MLP_EPI_BIAS_F32   4  /* C = acc + bias (fp32)                           */
#define G1_MLP_EPI_SILU_BWD   5  /* C = acc * silu'(Zin) (fp32), H = same (bf16)    */
#define G1_MLP_EPI_F32_SPLIT  6  /* C + z*I*ldc = acc of r-slice z (split partials) */

/* C[I][J] = sum_r A(i, r) B(j, r) with bf16 A, B and fp32 accumulation on
 * gfx950 MFMA.  A(i, r) = a_trans ? A[r*lda + i] : A[i*lda + r], likewise B.
 * I and J must be multiples of 128 and R a multiple of 64*split; split > 1
 * (gridDim.z partial sums) only with G1_MLP_EPI_F32_ATOMIC or
 * G1_MLP_EPI_F32_SPLIT (C then holds split slices of I x ldc).  C and H use
 * the leading dimension ldc.  Supported (a_trans, b_trans, epi): (0,0,
 * BIAS_SILU | BIAS_SILU_H | BIAS_F32 | F32 | SILU_BWD), (0,1, SILU_BWD | F32), (1,1,
 * F32_ATOMIC | F32_SPLIT | F32).  Returns 0, -1 (bad arguments) or -2 (launch
 * failure). */
int g1_mlp_gemm(int a_trans, int b_trans, int epi, const void* A, int lda, const void* B, int ldb,
                int I, int J, int R, int split, float* C, void* H, int ldc, const float* bias,
                const float* Zin, void* stream);
/* g1_mlp_gemm with colpart (nullable; G1_MLP_EPI_SILU_BWD only): the column
 * sums of the epilogue's fp32 output over each block of 64 rows,
 * colpart[(i / 64) * ldc + j], each summed in a fixed order (a bias
 * gradient's partials); with colpart given, C is nullable. */
int g1_mlp_gemm_ex(int a_trans, int b_trans, int epi, const void* A, int lda, const void* B, int ldb,
                   int I, int J, int R, int split, float* C, void* H, int ldc, const float* bias,
                   const float* Zin, float* colpart, void* stream);
/* One hidden layer's backward pass in one launch, the two GEMMs that read
 * the layer's output gradient d16 [B][rows] (bf16) side by side in one grid:
 * the weight gradient, g1_mlp_gemm_ex(1, 1, G1_MLP_EPI_F32_SPLIT, d16, rows,
 * hin, cols, rows, cols, B, split, gW_split, NULL, cols, ...) with hin
 * [B][cols] the layer's input, and the data gradient through SiLU',
 * g1_mlp_gemm_ex(0, 0, G1_MLP_EPI_SILU_BWD, d16, rows, WT16, rows, B, cols,
 * rows, 1, NULL, dZ16, ld_dz, NULL, Z, colpart) with WT16 [cols][rows] the
 * transposed bf16 weights, Z [B][ld_dz] the input's pre-activation and
 * colpart (required) the bias-gradient partials of the layer below.  Same
 * outputs, bit for bit, as the two calls; the weight gradient's workgroups
 * would fill only part of the chip alone.  Returns 0, -1 or -2. */
int g1_mlp_backward_layer(const void* d16, int rows, const void* hin, int cols, const void* WT16, int B, int split,
                          float* gW_split, void* dZ16, int ld_dz, const float* Z, float* colpart, void* stream);
/* dY = 2 (pred - target) / (B * n_out) on columns < n_out, 0 beyond, as
 * bf16 (dY16) and fp32 (dY32, nullable); *sse += sum of squared errors when
 * sse is non-null.  pred, target: fp32 [B][ld]. */
int g1_mlp_mse_grad(const float* pred, const float* target, int B, int ld, int n_out, void* dY16,
                    float* dY32, float* sse, void* stream);
/* The same dY (bf16 only) without atomics: per block of 64 rows, the column
 * sums of dY, colpart[(i / 64) * ld + j], and per block of 64 rows x 32
 * columns the summed squared error, sse_part[(i / 64) * (ld / 32) + j / 32],
 * each in a fixed order.  B a multiple of 64, ld of 32. */
int g1_mlp_mse_grad_part(const float* pred, const float* target, int B, int ld, int n_out, void* dY16,
                         float* colpart, float* sse_part, void* stream);
/* db[j] += sum_i dZ[i][j] over fp32 dZ [B][ld]. */
int g1_mlp_colsum(const float* dZ, int B, int ld, float* db, void* stream);
/* out[i] = sum_{z < split} parts[z * n + i] (the slices of a G1_MLP_EPI_F32_SPLIT
 * GEMM). */
int g1_mlp_sum_splits(const float* parts, int split, long long n, float* out, void* stream);
/* One Adam step (step number t >= 1, L2 weight decay wd) on fp32 master
 * weights p with moments m, v; writes the bf16 copy p16 (nullable).  mask
 * (nullable) marks trainable entries; masked-out entries stay zero. */
int g1_mlp_adam(float* p, float* m, float* v, const float* grad, void* p16, const uint8_t* mask,
                int n, float lr, float b1, float b2, float eps, int t, float wd, void* stream);
/* db[j] = sum_i dZ[i][j] over fp32 dZ [B][ld], without atomics (the same
 * bits on every run): part is fp32 scratch [ceil(B / 256)][ld]. */
int g1_mlp_colsum_det(const float* dZ, int B, int ld, float* part, float* db, void* stream);
/* Device-resident training loop: every per-iteration scalar is in device
 * memory, so an iteration is a fixed launch sequence (graph-capturable).
 * g1_mlp_train_schedule advances state (int [3]: next iteration of the run,
 * Adam step t, this iteration) and writes sched (float [3]: cosine-decayed
 * learning rate from lr to lr_end over iters, 1 - b1^t, 1 - b2^t).
 * g1_mlp_sample draws B (step, env) pairs into idx [2B] (the layout of
 * g1_mlp_gather) for the iteration in state[2]: step in [lo, hi), env in
 * [0, nenv), a counter-based hash of (seed, iteration, b).
 * g1_mlp_adam_multi is g1_mlp_adam over ntensor (<= 16) tensors in one
 * launch (host arrays of device pointers; p16, p16t, cols, mask, wd and their
 * entries nullable), with lr and the bias corrections read from sched; p16t[k]
 * (with cols[k]) also receives the transposed bf16 copy, [cols][rows], of a
 * [rows][cols] tensor. */
int g1_mlp_train_schedule(int* state, float* sched, int iters, float lr, float lr_end, float b1, float b2,
                          void* stream);
int g1_mlp_sample(int B, int lo, int hi, int nenv, unsigned seed, const int* state, int* idx, void* stream);
int g1_mlp_adam_multi(int ntensor, float* const* p, float* const* m, float* const* v, const float* const* grad,
                      void* const* p16, void* const* p16t, const int* cols, const uint8_t* const* mask,
                      const int* n, const float* wd, const float* sched, float b1, float b2, float eps,
                      void* stream);
/* g1_mlp_adam_multi with each gradient given as nsum[k] slices (grad[k] +
 * z * gstride[k], z < nsum[k], gstride[k] >= n[k]) that the step adds in a
 * fixed order: z ascending for up to 16 slices (the slices of a split
 * G1_MLP_EPI_F32_SPLIT GEMM, the same bits as g1_mlp_sum_splits), one wave
 * per element beyond (lane l adds slices l, l + 64, ..., then a butterfly
 * over the wave: the 64-row partials of g1_mlp_gemm_ex and
 * g1_mlp_mse_grad_part).  nsum and gstride nullable (one slice each). */
int g1_mlp_adam_multi_sum(int ntensor, float* const* p, float* const* m, float* const* v, const float* const* grad,
                          void* const* p16, void* const* p16t, const int* cols, const uint8_t* const* mask,
                          const int* n, const int* nsum, const long long* gstride, const float* wd,
                          const float* sched, float b1, float b2, float eps, void* stream);
/* x[0..n) = v (a fill kernel, usable where memset nodes are not). */
int g1_mlp_fill(float* x, float v, long long n, void* stream);
/* y (bf16) = x (fp32), n elements. */
int g1_mlp_to_bf16(const float* x, void* y, long long n, void* stream);
/* Builds z-scored samples from SoA blocks X[step][nfeat][nenv] and
 * Y[step][nout][nenv]: sample b is (step, env) = (idx[2b], idx[2b+1]), or env b
 * of step 0 when idx is null.  Xn: bf16 [B][ldx] (zero padded); Yn (nullable):
 * fp32 [B][ldy].  xinv / yinv are reciprocal standard deviations. */
int g1_mlp_gather(const float* X, const float* Y, const int* idx, int B, int nfeat, int nout, int nenv,
                  const float* xmean, const float* xinv, const float* ymean, const float* yinv,
                  void* Xn, int ldx, float* Yn, int ldy, void* stream);

/* g1_mlp_gather from a sample-major copy of the dataset, rows
 * S[step][env][nfeat + nout] (features, then targets); Yn required.  The
 * same bits as g1_mlp_gather. */
int g1_mlp_gather_rows(const float* S, const int* idx, int B, int nfeat, int nout, int nenv, const float* xmean,
                       const float* xinv, const float* ymean, const float* yinv, void* Xn, int ldx, float* Yn,
                       int ldy, void* stream);

/* Network input of every env: z-scored (qpos, qvel, tau) as bf16 rows of
 * Xn [nrows][ldx] (nrows >= nenv, padding rows and columns >= 106 zero;
 * ldx a multiple of 8).  mean / inv: [106] feature means and reciprocal
 * standard deviations. */
int g1_mlp_features(int nenv, int nrows, const float* qpos, const float* qvel, const float* tau,
                    const float* mean, const float* inv, void* Xn, int ldx, void* stream);

/* ---- surrogate step with ABA judge and fallback (no reference counterpart) */

/* Total generalized force (servo + limits + contact J^T f, as tau_out of
 * g1_hip_step_ex) and the contact update (anchors, normal forces) of one step,
 * without ABA and without integration: the surrogate's input.  foot_pos /
 * foot_vel ([6][nenv], bodies 6 and 12, world frame, as g1_cuda_step reports
 * them) are the feet before the step.  mu, contact_forces, foot_pos and
 * foot_vel are nullable. */
int g1_hip_forces(int nenv, float dt, const float* qpos, const float* qvel, float* anchor,
                  const float* ctrl, const float* mu, float* tau_out, float* contact_forces,
                  float* foot_pos, float* foot_vel, void* stream);
/* The rotating judge works on blocks of 64 consecutive envs: env e is judged
 * at the steps where (e / 64) % period == phase, so each env once per period
 * steps.
 * Device list of the envs that ABA steps this step: the judged envs and those
 * with fallback[e] != 0 (nullable).  list: int [nenv]; *count is reset and
 * filled on the device. */
int g1_hip_judge_list(int nenv, int phase, int period, const uint8_t* fallback, int* list, int* count,
                      void* stream);
/* ABA qacc (one substep at dt, state read-only) of the listed envs of a batch
 * of nenv, written to qacc_out [35][nenv] at their columns.  The contact
 * anchors are read, not updated.  flags: G1_FLAG_F64_SOLVE,
 * G1_FLAG_F64_CONTACT, G1_FLAG_WAVES4/2. */
int g1_hip_step_judge(int nenv, float dt, const float* qpos, const float* qvel, const float* anchor,
                      const float* ctrl, const float* mu, const int* list, const int* count, float* qacc_out,
                      unsigned flags, void* stream);
/* g1_hip_judge_list followed by g1_hip_step_judge in one kernel and without
 * the list: the workgroup of each block of 64 envs steps the block's judged
 * and fallback envs (fallback nullable), or exits when there are none. */
int g1_hip_step_judge_blocks(int nenv, float dt, const float* qpos, const float* qvel, const float* anchor,
                             const float* ctrl, const float* mu, const uint8_t* fallback, int phase, int period,
                             float* qacc_out, unsigned flags, void* stream);
/* g1_hip_step_judge_blocks and g1_hip_forces in one kernel: a block of 64
 * envs holding a judged or fallback env is stepped whole by ABA (qacc_out,
 * tau_out, the contact update and the feet stored for all its envs; qpos and
 * qvel are not written), every other block gets g1_hip_forces' outputs only.
 * The values are those of the two calls up to float rounding (a separate
 * kernel instance may contract multiply-adds differently).  contact_forces,
 * foot_pos, foot_vel, mu and fallback are nullable. */
int g1_hip_judge_forces(int nenv, float dt, const float* qpos, const float* qvel, float* anchor,
                        const float* ctrl, const float* mu, const uint8_t* fallback, int phase, int period,
                        float* qacc_out, float* tau_out, float* contact_forces, float* foot_pos,
                        float* foot_vel, unsigned flags, void* stream);
/* Applies qacc and integrates one step (semi-implicit Euler as the fused
 * step).  pred: the network's z-scored output, row e of [>= nenv][ldp];
 * qacc = pred * ystd + ymean, except for envs judged this step (the block
 * rotation above) or in fallback, which take qacc_aba.  A judged env
 * whose surrogate qacc differs from ABA's by more than tol (max over dofs of
 * |a - b| / max(|b|, 1)) enters fallback (ABA every step) until a later judge
 * accepts it.  stats (unsigned [4], accumulated): judged, rejected, max
 * judged error (float bits), fallback env-steps.  fallback and qacc_out are
 * nullable. */
int g1_hip_surrogate_integrate(int nenv, float dt, int phase, int period, float tol, float* qpos,
                               float* qvel, const float* pred, int ldp, const float* ymean,
                               const float* ystd, const float* qacc_aba, uint8_t* fallback,
                               unsigned* stats, float* qacc_out, void* stream);

/* ---- fused network inference (no reference counterpart) ------------------
 * One 512-thread workgroup per 64 envs runs a whole MLP with its activations
 * in LDS.  Weights are bf16 device arrays in the k-tiled layout
 * [in / 32][out][32] (element (o, i) at ((i / 32) * out + o) * 32 + i % 32),
 * biases fp32; W and b are HOST arrays of device pointers. */

/* Surrogate forward: features as g1_mlp_features (106 z-scored inputs), then
 * 128 -> 512 x4 (SiLU) -> 64 with W of (in, out) = (128, 512), (512, 512) x3,
 * (512, 64) (input columns >= 106 and output rows >= 35 zero) and
 * b = {[512] x4, [>=64]}.  pred: fp32 [nenv][ldp], ldp >= 64; the same
 * values as the layer-by-layer GEMMs (g1_mlp_gemm). */
int g1_hip_surrogate_forward(int nenv, const float* qpos, const float* qvel, const float* tau,
                             const float* xmean, const float* xinv, const void* const* W,
                             const float* const* b, float* pred, int ldp, void* stream);
/* Surrogate step: g1_hip_surrogate_forward followed by the commit of
 * g1_hip_surrogate_integrate (same judge, fallback, statistics and
 * integration), in one kernel; qpos / qvel are updated in place.  A block of
 * 64 envs that is not judged this step and whose envs are all in fallback
 * skips the network (its qacc is ABA's either way). */
int g1_hip_surrogate_step(int nenv, float dt, int phase, int period, float tol, float* qpos, float* qvel,
                          const float* tau, const float* xmean, const float* xinv, const void* const* W,
                          const float* const* b, const float* ymean, const float* ystd,
                          const float* qacc_aba, uint8_t* fallback, unsigned* stats, float* qacc_out,
                          void* stream);
/* nsteps steps of the surrogate stepper in one host call: for s < nsteps,
 * with phase (t0 + s) % period, g1_hip_judge_forces (ABA qacc of the judged
 * and fallback envs into qacc_aba, everyone's tau_total into tau, the contact
 * update) then g1_hip_surrogate_step; the same launches, in the same order,
 * as 2 * nsteps calls (ctrl and mu fixed over the call).  qacc_out,
 * contact_forces, foot_pos, foot_vel, mu and fallback nullable. */
int g1_hip_surrogate_steps(int nenv, int nsteps, int t0, float dt, int period, float tol, float* qpos,
                           float* qvel, float* anchor, const float* ctrl, const float* mu, const float* xmean,
                           const float* xinv, const void* const* W, const float* const* b,
                           const float* ymean, const float* ystd, float* qacc_aba, float* tau,
                           uint8_t* fallback, unsigned* stats, float* qacc_out, float* contact_forces,
                           float* foot_pos, float* foot_vel, unsigned judge_flags, void* stream);
/* Deferred judging (no reference counterpart; the judge of the north
 * star's surrogate + ABA-judge stepper).
 *
 * g1_hip_judge_forces_commit is g1_hip_judge_forces plus a commit: a block of
 * 64 envs whose live envs are all in fallback is integrated and stored in
 * place, as by the plain step (a fallback env takes ABA's qacc whatever its
 * judgement).  qacc_track (nullable) receives those envs' qacc, and stats[3]
 * counts their env-steps.  If such a block is judged this step, it records
 * its network inputs before the step (rows 0-105: qpos 36, qvel 35,
 * tau_total 35) and ABA's qacc (rows 106-140) into snap [141][nenv], and
 * sets sflag [nenv].
 *
 * g1_hip_surrogate_step_deferred, called after it, is g1_hip_surrogate_step
 * for the other blocks (it leaves the committed ones alone).
 *
 * g1_hip_surrogate_judge_batch runs the network on the records (the fused
 * forward kernel, its output kept in LDS) and judges each recorded env with
 * the judge's measure and tolerance.  An env within the tolerance leaves
 * fallback, then its flag is cleared.  stats needs 6 entries: [4] / [5]
 * count batch-judged / rejected envs, [2] takes their max error.
 *
 * g1_hip_surrogate_steps_deferred runs these per step, with the batch
 * judgement after every step whose phase is period - 1.  With blk_run
 * (uint8 [ceil(nenv / 64)], nullable) it goes by period segments: one
 * g1_hip_fallback_run launch per segment steps every block whose envs are
 * all in fallback to the segment's end (such a block stays in fallback until
 * the period's batch judgement, so all its steps are ABA's), and the
 * per-step launches leave those blocks alone.  The results are the per-step
 * path's: state, records, flags and statistics.  all_fallback (host int,
 * in/out, needs blk_run; nullable): 1 on entry says every env is in fallback
 * (as at the stepper's start, with only resets since); the per-step
 * launches, which would then have nothing to do, are not issued while that
 * holds.  After each batch judgement the envs off fallback are counted into
 * stats[6] (stats needs 7 entries) and read back, with one stream
 * synchronisation per period, once the next segment's fallback run is
 * queued; on return all_fallback says whether every env is in fallback.
 *
 * g1_hip_fallback_run: the blocks whose live envs are all in fallback take
 * nsub steps as by the plain step, each writing its record (as
 * g1_hip_judge_forces_commit does) at its judged step, phase (phase0 + s) %
 * period; nsub <= period - phase0.  blk_run[b] = 1 marks the blocks stepped,
 * 0 the others (nsub = 0 or nenv = 0 launches nothing and leaves blk_run as
 * it was).  qacc_track, contact_forces and the feet (nullable) receive
 * the last step's values of the stepped blocks; stats[3] counts their
 * env-steps.
 *
 * Blocks with an env off fallback are judged at once.  So, as with
 * immediate judging, the network's qacc is integrated only for envs whose
 * last judgement, at most period - 1 steps old, was within the tolerance;
 * an env leaves fallback at the end of the period instead of the step after
 * its judgement.  Callers that reset envs clear their sflag (the record
 * would judge the pre-reset state). */
int g1_hip_judge_forces_commit(int nenv, float dt, float* qpos, float* qvel, float* anchor,
                               const float* ctrl, const float* mu, const uint8_t* fallback, int phase,
                               int period, float* qacc_out, float* tau_out, float* contact_forces,
                               float* foot_pos, float* foot_vel, float* qacc_track, float* snap,
                               uint8_t* sflag, unsigned* stats, unsigned flags, void* stream);
int g1_hip_surrogate_step_deferred(int nenv, float dt, int phase, int period, float tol, float* qpos,
                                   float* qvel, const float* tau, const float* xmean, const float* xinv,
                                   const void* const* W, const float* const* b, const float* ymean,
                                   const float* ystd, const float* qacc_aba, uint8_t* fallback,
                                   unsigned* stats, float* qacc_out, void* stream);
int g1_hip_surrogate_judge_batch(int nenv, float tol, const float* xmean, const float* xinv,
                                 const void* const* W, const float* const* b, const float* ymean,
                                 const float* ystd, const float* snap, uint8_t* sflag, uint8_t* fallback,
                                 unsigned* stats, void* stream);
int g1_hip_surrogate_steps_deferred(int nenv, int nsteps, int t0, float dt, int period, float tol,
                                    float* qpos, float* qvel, float* anchor, const float* ctrl,
                                    const float* mu, const float* xmean, const float* xinv,
                                    const void* const* W, const float* const* b, const float* ymean,
                                    const float* ystd, float* qacc_aba, float* tau, uint8_t* fallback,
                                    unsigned* stats, float* qacc_out, float* contact_forces, float* foot_pos,
                                    float* foot_vel, unsigned judge_flags, float* snap, uint8_t* sflag,
                                    uint8_t* blk_run, int* all_fallback, void* stream);
/* g1_hip_surrogate_steps_deferred with a PD-target schedule: step s of the
 * call uses the block ctrl_blocks + ((ctrl_step0 + s) / ctrl_period) *
 * block_stride ([29][nenv] each, block_stride >= 29 * nenv floats apart), as
 * g1_hip_step_schedule; with blk_run the fallback runs switch target blocks
 * themselves and span whole period segments.  Same results as one
 * g1_hip_surrogate_steps_deferred call per target block. */
int g1_hip_surrogate_steps_schedule(int nenv, int nsteps, int t0, float dt, int period, float tol,
                                    float* qpos, float* qvel, float* anchor, const float* ctrl_blocks,
                                    long long block_stride, int ctrl_period, int ctrl_step0, const float* mu,
                                    const float* xmean, const float* xinv, const void* const* W,
                                    const float* const* b, const float* ymean, const float* ystd,
                                    float* qacc_aba, float* tau, uint8_t* fallback, unsigned* stats,
                                    float* qacc_out, float* contact_forces, float* foot_pos, float* foot_vel,
                                    unsigned judge_flags, float* snap, uint8_t* sflag, uint8_t* blk_run,
                                    int* all_fallback, void* stream);
int g1_hip_fallback_run(int nenv, int nsub, float dt, float* qpos, float* qvel, float* anchor,
                        const float* ctrl, const float* mu, const uint8_t* fallback, int phase0, int period,
                        uint8_t* blk_run, float* qacc_track, float* contact_forces, float* foot_pos,
                        float* foot_vel, float* snap, uint8_t* sflag, unsigned* stats, unsigned flags,
                        void* stream);

/* RL observation of every env (replaces the torch code of
 * src/g1sim.py:132-170).  layout 0: 99 values per env in the reference order
 * (commands 3, projected gravity 3, base linear velocity in the base frame 3,
 * base angular velocity 3, q - stand 29, qd 29, last ctrl 29); layout 1: 96
 * values (base angular velocity, projected gravity, commands, q - stand, qd,
 * last ctrl).  commands [3][nenv], last_ctrl [29][nenv]; obs fp32
 * [nenv][99 or 96]. */
int g1_hip_get_obs(int nenv, int layout, const float* qpos, const float* qvel, const float* commands,
                   const float* last_ctrl, float* obs, void* stream);

/* Policy act (examples/train.py:21-29, 69-70, 86-95, 284-310): observation
 * (layout as g1_hip_get_obs), z-scored with obs_mean / obs_inv
 * (1 / sqrt(var + 1e-8), >= 99 entries) and clipped to +-clip, then
 * 128 -> 512 -> 256 -> 128 (ELU) -> 32 with W of (in, out) = (128, 512),
 * (512, 256), (256, 128), (128, 32); action j = out_j (+ exp(log_std_j) N(0, 1) when
 * log_std is non-null, drawn from (seed, t, env)); ctrl = stand + action *
 * action_scale ([29][nenv]).  actions ([29][nenv]) and obs_out
 * ([nenv][nobs], the raw observation) are nullable.  push_interval
 * (nullable, int [nenv]) counts down; at <= 0 the base velocity qvel[0:6]
 * gets a uniform kick of +-(0.5, 0.5, 0.4, 0.52, 0.52, 0.78) and a new
 * interval in [50, 150) (examples/train.py:301-310).  last_ctrl may alias
 * ctrl. */
int g1_policy_act(int nenv, int layout, const float* qpos, float* qvel, const float* commands,
                  const float* last_ctrl, const float* obs_mean, const float* obs_inv, float clip,
                  const void* const* W, const float* const* b, const float* log_std, unsigned seed, int t,
                  float* ctrl, float* actions, float* obs_out, int* push_interval, void* stream);

/* Running observation normaliser update of examples/train.py:49-66 (what
 * ActorCritic.act() does before normalising): batch mean and biased variance
 * of the observation (layout as g1_hip_get_obs) over the nenv envs, merged
 * into mean / var ([nobs] fp32) and count ([1] fp32) in place by Chan's
 * formula; inv ([nobs]) receives 1 / sqrt(var + 1e-8) for g1_policy_act.
 * scratch: double [2 * 64 * nobs + 1] device workspace. */
int g1_policy_obs_norm_update(int nenv, int layout, const float* qpos, const float* qvel, const float* commands,
                              const float* last_ctrl, float* mean, float* var, float* count, float* inv,
                              double* scratch, void* stream);

/* g1_policy_obs_norm_update without the scratch argument, for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc); it has no counterpart in the reference.
 * Its workspace is a static device array of 2 * 64 * 99 + 1 doubles (99 KiB,
 * the widest layout) that each GPU holds until the process ends; its address
 * on the current device is looked up on the first call there.  It allocates
 * nothing.  Every call on one GPU shares that workspace: calls that overlap
 * on one GPU (concurrent streams, or host threads on one device) race.
 * The current device must be the one the buffers and stream belong to. */
int g1_policy_obs_norm_update_noalloc(int nenv, int layout, const float* qpos, const float* qvel,
                                      const float* commands, const float* last_ctrl, float* mean, float* var,
                                      float* count, float* inv, void* stream);

/* Reward and termination (examples/train.py:143-225) of the post-step state,
 * then the rollout bookkeeping (examples/train.py:319-351): prev_actions <-
 * actions (0 for reset envs), command_timeout - 1, episode_length + 1,
 * reset = done | episode_length >= max_episode_length; reset envs get new
 * commands (x, y ~ U(+-cmd_lin), yaw ~ U(+-cmd_yaw)), a timer in [150, 250),
 * episode_length 0 and a push interval in [50, 150) (push_interval
 * nullable); envs whose timer ran out get new commands and timer.  reward
 * fp32 [nenv], done and reset uint8 [nenv]; the caller resets the envs
 * flagged in `reset` (g1_cuda_reset_done).  reward_sum / done_count (one
 * double each, nullable) are incremented by this step's summed reward and
 * number of terminated envs (device atomics: the summation order, and so the
 * last bits of reward_sum, vary from run to run). */
int g1_rollout_post(int nenv, const float* qpos, const float* qvel, const float* foot_pos,
                    const float* foot_vel, const float* contact_forces, const float* actions, float* commands,
                    float* prev_actions, int* command_timeout, int* episode_length, int* push_interval,
                    int max_episode_length, float cmd_lin, float cmd_yaw, unsigned seed, int t,
                    float* reward, uint8_t* done, uint8_t* reset, double* reward_sum,
                    double* done_count, void* stream);
/* Rollout start: commands for every env, timers 0, push intervals in
 * [50, 150) (push_interval nullable). */
int g1_rollout_init(int nenv, float cmd_lin, float cmd_yaw, unsigned seed, float* commands,
                    int* command_timeout, int* episode_length, int* push_interval, void* stream);

/* ---- machine-peak microbenchmarks (measurement only) -----------------------
 * Non-packed FP32 FMA (blocks x 256 threads x iters x 8 FMAs) and dense BF16
 * MFMA (blocks x 4 waves x iters x 16 v_mfma_f32_16x16x32_bf16) on register
 * operands; out: float [blocks * 256]. */
int g1_hip_peak_fp32(int blocks, int iters, float* out, void* stream);
int g1_hip_peak_bf16(int blocks, int iters, float* out, void* stream);

/* Message of the last failure on the calling thread: every int-returning
 * entry point clears it on entry and names itself and the reason when it
 * returns non-zero ("" after a successful call). */
const char* g1_hip_last_error(void);

/* Library version (major * 10000 + minor * 100 + patch). */
int g1_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif
