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

/* Message of the last failure on the calling thread ("" if none). */
const char* g1_hip_last_error(void);

/* Library version (major * 10000 + minor * 100 + patch). */
int g1_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif
