// g1_step.hip -- fused G1 physics step for MI355X (gfx950).
//
// One step = kinematics and body twists -> 8 foot-sphere spring-damper
// contacts with stick-slip anchors -> PD servos with implicit damping and soft
// joint limits -> Featherstone ABA (packed 21-float articulated inertias,
// floating-base 6x6 Cholesky) -> semi-implicit Euler.  The physics is that of
// src/g1_core.h:29-232; the data path is built for CDNA4:
//
// * Work split.  A workgroup holds 64 envs (lane = env) in 4 wave64s.  Each
//   wave owns one limb chain of the tree: wave 0 the left leg (bodies 1-6)
//   and contacts 0-3, wave 1 the right leg (7-12) and contacts 4-7, wave 2
//   the left arm (16-22), wave 3 the right arm (23-29).  Chain code is
//   specialised per wave (uniform branch on the wave index), so every model
//   constant is an instruction literal and every hinge rotation is a planar
//   rotation about a unit axis.  16,384 envs fill 1,024 waves, one per SIMD
//   of the 256 CUs, where one lane per env would leave 3 of 4 SIMDs idle.
// * One LDS exchange per substep.  Each chain reduces to a 27-float
//   articulated inertia + bias force in its root's parent frame; the four
//   contributions go through LDS (double-buffered, one barrier per substep)
//   and every wave then runs the short pelvis/waist part (bodies 0, 13-15,
//   the 6x6 base solve) with identical code, so no second barrier is needed.
// * Registers only.  No per-env workspace in memory: the forward pass keeps
//   7 floats per joint (U/D and u/D) plus the joint's cos/sin and 4-float
//   velocity-product term.
// * Gravity enters as a uniform acceleration offset of the base (the classic
//   ABA identity: bias forces without gravity, a0 = a0' + [0; R0^T g]), so no
//   body needs its world rotation except the two feet chains.
// * HBM traffic is the state itself, coalesced SoA loads/stores
//   ([dim][nenv], env fastest, as src/g1_kernel.cu:9-10), once per launch;
//   nsub substeps run from registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <stdarg.h>

#include <type_traits>

#include "g1_model_kernel.h"
#include "g1_fastmath.h"
#include "g1_commit.h"
#include "g1_err.h"
#include "g1_launch.h"
#include "../../include/g1sim_amd.h"

namespace g1k {

struct Cfg {
  float dt;
  float kn, dn, kt, dtt, inv_kt;
  float limit_kp, limit_kd;
  float gz;
  float mu;
};

// src/g1_types.h:44-55 (dt is a runtime argument here)
static Cfg default_cfg(float dt) {
  Cfg c;
  c.dt = dt;
  c.kn = 2e4f; c.dn = 280.f; c.kt = 1e4f; c.dtt = 280.f; c.inv_kt = 1.0f / 1e4f;
  c.limit_kp = 200.f; c.limit_kd = 5.f;
  c.gz = (float)G1M_GRAVITY_Z;
  c.mu = (float)G1M_CONTACT_MU;
  return c;
}

constexpr float ANCHOR_FREE = 1e30f;     // src/g1_core.h:108

// ---------------------------------------------------------------- rotations
// Planar rotation about axis K: plane (I, J) = ((K+1)%3, (K+2)%3) and
// R_K = identity on K, [[c, -s], [s, c]] on (I, J)  (src/spatial.h:63-70 for a
// unit axis).
template <int K> struct Plane { static constexpr int I = (K + 1) % 3, J = (K + 2) % 3; };

// c * x for a model constant c (a compile-time constant once the templates
// are expanded).  Without fast-math flags the compiler must keep 0 * x (x
// could be inf or NaN): the zero entries of the tree rotations, offsets,
// skew matrices and rigid-body inertias cost one multiply or FMA each.  A
// zero coefficient gives an exact -0 instead, which every sum absorbs, so no
// instruction is issued and finite results are unchanged.  A difference
// a - c x is written a + cm(-c, x) (bit-identical: negation is exact).
template <typename T_> G1K_DEV T_ cm(T_ c, T_ x) { return c == T_(0) ? T_(-0.0) : c * x; }

template <int K, typename T_ = float> G1K_DEV void rot(T_* v, T_ c, T_ s) {          // v <- R_K v
  constexpr int I = Plane<K>::I, J = Plane<K>::J;
  const T_ a = v[I], b = v[J];
  v[I] = c * a - s * b;
  v[J] = s * a + c * b;
}
template <int K, typename T_ = float> G1K_DEV void rotT(T_* v, T_ c, T_ s) {         // v <- R_K^T v
  constexpr int I = Plane<K>::I, J = Plane<K>::J;
  const T_ a = v[I], b = v[J];
  v[I] = c * a + s * b;
  v[J] = c * b - s * a;
}
// Tree rotations folded into the joint angle: planar about the hinge's own
// axis, so RT R_K(q) = R_K(theta + q) and fk_joint composes theta into the
// joint's cos/sin; such bodies, like those without a tree rotation, skip it.
template <int B> constexpr bool rt_none = kM.b[B].rt_ident || kM.b[B].rt_fold;

// constant rotation of body B's tree frame (a planar rotation where it is one)
template <int B> G1K_DEV void rt_mul(float* v) {                          // v <- RT v
  if constexpr (kM.b[B].rt_axis >= 0) {
    rot<kM.b[B].rt_axis>(v, (float)kM.b[B].rt_c, (float)kM.b[B].rt_s);
    return;
  }
  constexpr const double* R = kM.b[B].RT;
  const float x = v[0], y = v[1], z = v[2];
  v[0] = cm((float)R[0], x) + cm((float)R[1], y) + cm((float)R[2], z);
  v[1] = cm((float)R[3], x) + cm((float)R[4], y) + cm((float)R[5], z);
  v[2] = cm((float)R[6], x) + cm((float)R[7], y) + cm((float)R[8], z);
}
template <int B, typename T_ = float> G1K_DEV void rt_mulT(T_* v) {      // v <- RT^T v
  if constexpr (kM.b[B].rt_axis >= 0) {
    rotT<kM.b[B].rt_axis, T_>(v, (T_)kM.b[B].rt_c, (T_)kM.b[B].rt_s);
    return;
  }
  constexpr const double* R = kM.b[B].RT;
  const T_ x = v[0], y = v[1], z = v[2];
  v[0] = cm((T_)R[0], x) + cm((T_)R[3], y) + cm((T_)R[6], z);
  v[1] = cm((T_)R[1], x) + cm((T_)R[4], y) + cm((T_)R[7], z);
  v[2] = cm((T_)R[2], x) + cm((T_)R[5], y) + cm((T_)R[8], z);
}

// ------------------------------------------------------- spatial transforms
// X_B = (R, t) with R = RT_B R_K(q): pose of body B in its parent.
// Motion vector parent -> child (src/spatial.h:102-105).
template <int B, typename T_ = float> G1K_DEV void motion_to_child(const T_* m, T_ c, T_ s, T_* out) {
  constexpr int K = kM.b[B].axis;
  constexpr T_ t0 = (T_)kM.b[B].t[0], t1 = (T_)kM.b[B].t[1], t2 = (T_)kM.b[B].t[2];
  T_ a[3] = {m[0], m[1], m[2]};
  // l = m_l - t x m_a, written m_l + (m_a x t) so that a zero offset leaves no add
  T_ l[3] = {m[3] + (cm(-t1, m[2]) + cm(t2, m[1])), m[4] + (cm(-t2, m[0]) + cm(t0, m[2])),
             m[5] + (cm(-t0, m[1]) + cm(t1, m[0]))};
  if constexpr (!rt_none<B>) { rt_mulT<B, T_>(a); rt_mulT<B, T_>(l); }
  rotT<K, T_>(a, c, s);
  rotT<K, T_>(l, c, s);
  out[0] = a[0]; out[1] = a[1]; out[2] = a[2];
  out[3] = l[0]; out[4] = l[1]; out[5] = l[2];
}
// Force vector child -> parent, accumulated (src/spatial.h:107-110).
template <int B> G1K_DEV void force_to_parent_add(const float* f, float c, float s, float* acc) {
  constexpr int K = kM.b[B].axis;
  constexpr float t0 = (float)kM.b[B].t[0], t1 = (float)kM.b[B].t[1], t2 = (float)kM.b[B].t[2];
  float a[3] = {f[0], f[1], f[2]}, l[3] = {f[3], f[4], f[5]};
  rot<K>(a, c, s);
  rot<K>(l, c, s);
  if constexpr (!rt_none<B>) { rt_mul<B>(a); rt_mul<B>(l); }
  acc[0] += a[0] + (cm(t1, l[2]) + cm(-t2, l[1]));
  acc[1] += a[1] + (cm(t2, l[0]) + cm(-t0, l[2]));
  acc[2] += a[2] + (cm(t0, l[1]) + cm(-t1, l[0]));
  acc[3] += l[0]; acc[4] += l[1]; acc[5] += l[2];
}

template <int B> G1K_DEV void force_to_parent(float (&f)[6], float c, float s) {
  float acc[6] = {-0.f, -0.f, -0.f, -0.f, -0.f, -0.f};   // sums start at -0: -0 + x is x (no add issued)
  force_to_parent_add<B>(f, c, s, acc);
#pragma unroll
  for (int m = 0; m < 6; ++m) f[m] = acc[m];
}

// Symmetric 3x3 S <- R_K S R_K^T, with c2 = cos 2q, s2 = sin 2q.
template <int K> G1K_DEV void rot_sym(float (&S)[3][3], float c, float s, float c2, float s2) {
  constexpr int I = Plane<K>::I, J = Plane<K>::J;
  const float ski = S[K][I], skj = S[K][J], sii = S[I][I], sjj = S[J][J], sij = S[I][J];
  const float nki = c * ski - s * skj, nkj = s * ski + c * skj;
  const float h = 0.5f * (sii + sjj), d = sii - sjj;
  const float e = 0.5f * d * c2 - sij * s2;
  const float nij = 0.5f * d * s2 + sij * c2;
  S[K][I] = S[I][K] = nki;
  S[K][J] = S[J][K] = nkj;
  S[I][I] = h + e;
  S[J][J] = h - e;
  S[I][J] = S[J][I] = nij;
}
// General 3x3 H <- R_K H R_K^T.
template <int K> G1K_DEV void rot_gen(float (&H)[3][3], float c, float s) {
  constexpr int I = Plane<K>::I, J = Plane<K>::J;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const float a = H[I][m], b = H[J][m];
    H[I][m] = c * a - s * b;
    H[J][m] = s * a + c * b;
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const float a = H[m][I], b = H[m][J];
    H[m][I] = c * a - s * b;
    H[m][J] = s * a + c * b;
  }
}
// A <- RT A RT^T for a constant tree rotation (sym: only the lower triangle is used).
template <int B, bool SYM> G1K_DEV void rt_sim(float (&A)[3][3]) {
  if constexpr (kM.b[B].rt_axis >= 0) {
    constexpr double c = kM.b[B].rt_c, s = kM.b[B].rt_s;
    if constexpr (SYM) rot_sym<kM.b[B].rt_axis>(A, (float)c, (float)s, (float)(c * c - s * s), (float)(2.0 * c * s));
    else rot_gen<kM.b[B].rt_axis>(A, (float)c, (float)s);
    return;
  }
  constexpr const double* R = kM.b[B].RT;
  float T[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      T[r][c] = cm((float)R[3 * r], A[0][c]) + cm((float)R[3 * r + 1], A[1][c]) + cm((float)R[3 * r + 2], A[2][c]);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (SYM && c > r) continue;
      A[r][c] = cm((float)R[3 * c], T[r][0]) + cm((float)R[3 * c + 1], T[r][1]) + cm((float)R[3 * c + 2], T[r][2]);
    }
  if (SYM) {
    A[0][1] = A[1][0]; A[0][2] = A[2][0]; A[1][2] = A[2][1];
  }
}

// Articulated inertia child -> parent, accumulated: P += X* Ia X^-1
// (src/spatial.h:219-251): rotate the J/H/M blocks, then shift by t:
// H_A = H + t^ M, J_A = J + t^ H^T - H_A t^, M_A = M.
template <int B>
G1K_DEV void inertia_to_parent_add(const float (&A)[21], float c, float s, float (&P)[21]) {
  constexpr int K = kM.b[B].axis;
  constexpr float t0 = (float)kM.b[B].t[0], t1 = (float)kM.b[B].t[1], t2 = (float)kM.b[B].t[2];
  float Jb[3][3], H[3][3], M[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      Jb[r][q] = A[pk(r, q)];
      H[r][q] = A[pk(r, 3 + q)];
      M[r][q] = A[pk(3 + r, 3 + q)];
    }
  const float c2 = c * c - s * s, s2 = 2.0f * c * s;
  rot_sym<K>(Jb, c, s, c2, s2);
  rot_sym<K>(M, c, s, c2, s2);
  rot_gen<K>(H, c, s);
  if constexpr (!rt_none<B>) {
    rt_sim<B, true>(Jb);
    rt_sim<B, true>(M);
    rt_sim<B, false>(H);
  }
  const float tx[3][3] = {{0.f, -t2, t1}, {t2, 0.f, -t0}, {-t1, t0, 0.f}};
  float HA[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q)
      HA[r][q] = H[r][q] + (cm(tx[r][0], M[0][q]) + cm(tx[r][1], M[1][q]) + cm(tx[r][2], M[2][q]));
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q <= r; ++q) {
      const float tht = cm(tx[r][0], H[q][0]) + cm(tx[r][1], H[q][1]) + cm(tx[r][2], H[q][2]);
      const float hat = cm(tx[0][q], HA[r][0]) + cm(tx[1][q], HA[r][1]) + cm(tx[2][q], HA[r][2]);
      P[pk(r, q)] += Jb[r][q] + tht - hat;
      P[pk(3 + r, 3 + q)] += M[r][q];
    }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) P[pk(r, 3 + q)] += HA[r][q];
}

// bias = v x* (I v) with body B's constant rigid-body inertia
// (src/g1_core.h:119 without its gravity term).
template <int B> G1K_DEV void bias_force(const float* v, float* f) {
  float h[6];
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    float acc = -0.f;
#pragma unroll
    for (int m = 0; m < 6; ++m) acc += cm((float)kM.b[B].I[pk(r, m)], v[m]);
    h[r] = acc;
  }
  // [w x h_a + v_l x h_l ; w x h_l]   (src/spatial.h:116-118)
  f[0] = (v[1] * h[2] - v[2] * h[1]) + (v[4] * h[5] - v[5] * h[4]);
  f[1] = (v[2] * h[0] - v[0] * h[2]) + (v[5] * h[3] - v[3] * h[5]);
  f[2] = (v[0] * h[1] - v[1] * h[0]) + (v[3] * h[4] - v[4] * h[3]);
  f[3] = v[1] * h[5] - v[2] * h[4];
  f[4] = v[2] * h[3] - v[0] * h[5];
  f[5] = v[0] * h[4] - v[1] * h[3];
}

// Zero entries start as -0: contributions added to them later need no add
// (-0 + x is x; +0 + x is not, for x = -0).
template <int B> G1K_DEV void init_inertia(float (&A)[21]) {
#pragma unroll
  for (int m = 0; m < 21; ++m) A[m] = kM.b[B].I[m] == 0.0 ? -0.f : (float)kM.b[B].I[m];
}

// ------------------------------------------------------------ per-joint data
// Kept per hinge from kinematics to the forward pass.
struct Joint {
  float c, s;     // cos/sin of the joint angle
  float cor[4];   // velocity-product acceleration v x (qd S), nonzero (I,J) entries: a_I, a_J, l_I, l_J
  float bias[6];  // velocity-product bias force (consumed by the backward pass)
  float W[6];     // U / D
  float uD;       // u / D
  float tau;      // servo + joint-limit torque (src/g1_core.h:166-174)
};

// cos/sin of the hinge angle -> of theta + q where the tree rotation of body
// B is R_K(theta) (rt_fold): the composite R_K(theta) R_K(q).
template <int B, typename T_ = float> G1K_DEV void fold_tree(T_& c, T_& s) {
  if constexpr (kM.b[B].rt_fold) {
    constexpr T_ ct = (T_)kM.b[B].rt_c, st = (T_)kM.b[B].rt_s;
    const T_ c0 = c, s0 = s;
    c = c0 * ct - s0 * st;
    s = s0 * ct + c0 * st;
  }
}

// Kinematics of hinge body B: twist from the parent's (src/g1_core.h:39-53).
template <int B> G1K_DEV void fk_joint(float q, float qd, const float* vpar, float* v, Joint& jt) {
  constexpr int K = kM.b[B].axis, I = Plane<K>::I, J = Plane<K>::J;
  fsincos(q, &jt.s, &jt.c);
  fold_tree<B>(jt.c, jt.s);
  motion_to_child<B>(vpar, jt.c, jt.s, v);
  v[K] += qd;
  // cross_motion(v, qd S) with S = [e_K; 0] (src/g1_core.h:180)
  jt.cor[0] = qd * v[J];
  jt.cor[1] = -qd * v[I];
  jt.cor[2] = qd * v[3 + J];
  jt.cor[3] = -qd * v[3 + I];
  bias_force<B>(v, jt.bias);
}

// World pose of hinge body B from its parent's (src/g1_core.h:47-50).
template <int B, typename T_>
G1K_DEV void world_pose(const T_ (&Rp)[9], const T_ (&pp)[3], T_ c, T_ s, T_ (&R)[9], T_ (&p)[3]) {
  constexpr int K = kM.b[B].axis, I = Plane<K>::I, J = Plane<K>::J;
  constexpr T_ t0 = (T_)kM.b[B].t[0], t1 = (T_)kM.b[B].t[1], t2 = (T_)kM.b[B].t[2];
#pragma unroll
  for (int r = 0; r < 3; ++r) p[r] = cm(t0, Rp[3 * r]) + cm(t1, Rp[3 * r + 1]) + cm(t2, Rp[3 * r + 2]) + pp[r];
  T_ T[9];
  if constexpr (rt_none<B>) {
#pragma unroll
    for (int m = 0; m < 9; ++m) T[m] = Rp[m];
  } else {
    constexpr const double* Q = kM.b[B].RT;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        T[3 * r + q] = cm((T_)Q[q], Rp[3 * r]) + cm((T_)Q[3 + q], Rp[3 * r + 1]) + cm((T_)Q[6 + q], Rp[3 * r + 2]);
  }
  // columns I, J of T R_K
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const T_ a = T[3 * r + I], b = T[3 * r + J];
    R[3 * r + K] = T[3 * r + K];
    R[3 * r + I] = c * a + s * b;
    R[3 * r + J] = c * b - s * a;
  }
}

// Backward step of hinge body B into its parent: PD servo + soft limits +
// implicit damping (src/g1_core.h:156-186).  A/pA (the body's articulated
// inertia and bias) are consumed.
// Servo + soft-limit torque of hinge h and its velocity coefficient b for
// the implicit damping term dt * b (src/g1_core.h:166-176).
template <int h> G1K_DEV float servo_torque(const Cfg& cf, float q, float qd, float ctl, float& b) {
  constexpr float kp = (float)g1m_act_kp[h], kv = (float)g1m_act_kv[h];
  constexpr float lim = (float)g1m_act_frclim[h];
  constexpr float lo = (float)g1m_jnt_range[2 * h], hi = (float)g1m_jnt_range[2 * h + 1];
  float t = kp * (ctl - q) - kv * qd;
  t = fminf(fmaxf(t, -lim), lim);
  b = kv;
  const bool below = q < lo, above = q > hi;
  const float lt = cf.limit_kp * ((below ? lo : hi) - q) - cf.limit_kd * qd;
  t += (below || above) ? lt : 0.f;
  b += (below || above) ? cf.limit_kd : 0.f;
  return t;
}

template <int B>
G1K_DEV void backward_joint(const Cfg& cf, float (&A)[21], float (&pA)[6], float q, float qd, float ctl,
                            Joint& jt, float (&Ppar)[21], float (&pApar)[6]) {
  constexpr int h = B - 1, K = kM.b[B].axis, I = Plane<K>::I, J = Plane<K>::J;
  constexpr float arm = (float)g1m_armature[h];
  float b;
  const float t = servo_torque<h>(cf, q, qd, ctl, b);
  jt.tau = t;

  float U[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) U[m] = A[pk(m, K)];
  const float D = U[K] + arm + cf.dt * b;
  const float invD = frcp(D);
  const float u = t - pA[K];
#pragma unroll
  for (int m = 0; m < 6; ++m) jt.W[m] = U[m] * invD;
  jt.uD = u * invD;
  // Ia = IA - U U^T / D
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int q2 = 0; q2 <= r; ++q2) A[pk(r, q2)] -= U[r] * jt.W[q2];
  // pa = pA + Ia c + (u/D) U, c nonzero at angular I,J and linear I,J
  float pa[6];
#pragma unroll
  for (int r = 0; r < 6; ++r)
    pa[r] = pA[r] + A[pk(r, I)] * jt.cor[0] + A[pk(r, J)] * jt.cor[1] + A[pk(r, 3 + I)] * jt.cor[2] +
            A[pk(r, 3 + J)] * jt.cor[3] + jt.uD * U[r];
  inertia_to_parent_add<B>(A, jt.c, jt.s, Ppar);
  force_to_parent_add<B>(pa, jt.c, jt.s, pApar);
}

// Forward step of hinge body B (src/g1_core.h:190-201): acceleration from the
// parent's, qdd = u/D - (U/D) . ap.
template <int B> G1K_DEV float forward_joint(const float* apar, const Joint& jt, float* a) {
  constexpr int K = kM.b[B].axis, I = Plane<K>::I, J = Plane<K>::J;
  motion_to_child<B>(apar, jt.c, jt.s, a);
  a[I] += jt.cor[0];
  a[J] += jt.cor[1];
  a[3 + I] += jt.cor[2];
  a[3 + J] += jt.cor[3];
  float dot = -0.f;
#pragma unroll
  for (int m = 0; m < 6; ++m) dot += jt.W[m] * a[m];
  const float qdd = jt.uD - dot;
  a[K] += qdd;
  return qdd;
}

// ------------------------------------------------------------------ chains
// Chain of N bodies FIRST..FIRST+N-1 (each the parent of the next).
template <int FIRST, int N> struct ChainOps {
  template <int i = 0> static G1K_DEV void fk(const float* q, const float* qd, const float* vroot, Joint* jt,
                                              float (&vleaf)[6]) {
    if constexpr (i < N) {
      float v[6];
      fk_joint<FIRST + i>(q[i], qd[i], vroot, v, jt[i]);
      if constexpr (i + 1 < N) fk<i + 1>(q, qd, v, jt, vleaf);
      else {
#pragma unroll
        for (int m = 0; m < 6; ++m) vleaf[m] = v[m];
      }
    }
  }
  // world poses along the chain (feet only); returns the leaf's
  // World poses along the chain (feet only) and, for T_ = double, the body
  // twists again in float64 with the joint rotations re-evaluated from q.
  template <typename T_, int i = 0>
  static G1K_DEV void world(const T_ (&Rp)[9], const T_ (&pp)[3], const T_ (&vp)[6], const Joint* jt, const float* q,
                            const float* qd, T_ (&Rl)[9], T_ (&pl)[3], T_ (&vl)[6]) {
    if constexpr (i < N) {
      T_ R[9], p[3], v[6], c, s;
      if constexpr (sizeof(T_) == sizeof(float)) {
        c = jt[i].c; s = jt[i].s;
      } else {
        sincos((double)q[i], &s, &c);
        fold_tree<FIRST + i, T_>(c, s);
        motion_to_child<FIRST + i, T_>(vp, c, s, v);
        v[kM.b[FIRST + i].axis] += (T_)qd[i];
      }
      world_pose<FIRST + i>(Rp, pp, c, s, R, p);
      if constexpr (i + 1 < N) world<T_, i + 1>(R, p, v, jt, q, qd, Rl, pl, vl);
      else {
#pragma unroll
        for (int m = 0; m < 9; ++m) Rl[m] = R[m];
#pragma unroll
        for (int m = 0; m < 3; ++m) pl[m] = p[m];
        if constexpr (sizeof(T_) != sizeof(float)) {   // float: the caller's twists stand
#pragma unroll
          for (int m = 0; m < 6; ++m) vl[m] = v[m];
        }
      }
    }
  }
  // backward from the leaf, whose articulated quantities start as (A, pA);
  // the chain root's contribution lands in (Pout, pout) in the root's parent frame.
  template <int i = N - 1> static G1K_DEV void backward(const Cfg& cf, float (&A)[21], float (&pA)[6],
                                                        const float* q, const float* qd, const float* ctl,
                                                        Joint* jt, float (&Pout)[21], float (&pout)[6]) {
    if constexpr (i > 0) {
      float Ap[21], pp[6];
      init_inertia<FIRST + i - 1>(Ap);
#pragma unroll
      for (int m = 0; m < 6; ++m) pp[m] = jt[i - 1].bias[m];
      backward_joint<FIRST + i>(cf, A, pA, q[i], qd[i], ctl[i], jt[i], Ap, pp);
      backward<i - 1>(cf, Ap, pp, q, qd, ctl, jt, Pout, pout);
    } else {
      backward_joint<FIRST>(cf, A, pA, q[0], qd[0], ctl[0], jt[0], Pout, pout);
    }
  }
  // generalized force of a body-frame wrench F on the leaf: tau_i = S_i^T F_i
  // up the chain; F ends in the root's parent frame (J^T f of src/g1_core.h:146-148)
  template <int i = N - 1> static G1K_DEV void wrench_to_joints(float (&F)[6], const Joint* jt, float* tc) {
    if constexpr (i >= 0) {
      tc[i] = F[kM.b[FIRST + i].axis];
      force_to_parent<FIRST + i>(F, jt[i].c, jt[i].s);
      wrench_to_joints<i - 1>(F, jt, tc);
    }
  }
  template <int i = 0> static G1K_DEV void forward(const float* aroot, const Joint* jt, float* qdd, float (&aleaf)[6]) {
    if constexpr (i < N) {
      float a[6];
      qdd[i] = forward_joint<FIRST + i>(aroot, jt[i], a);
      if constexpr (i + 1 < N) forward<i + 1>(a, jt, qdd, aleaf);
      else {
#pragma unroll
        for (int m = 0; m < 6; ++m) aleaf[m] = a[m];
      }
    }
  }
};

// ----------------------------------------------------------------- contacts
// Four foot spheres on body FOOT (contacts K0..K0+3) against the plane z = 0
// (src/g1_core.h:122-153): normal spring-damper clamped to [0, 500] N,
// tangential anchor spring-damper projected on the Coulomb cone, the anchor
// dragged on slip and released on lift-off.  The wrench is summed in the
// world frame about the body origin and rotated once.
// With T_ = double the contact-point kinematics (world position, penetration
// depth, anchor offsets, point velocity) are evaluated in float64: the depth
// is a small difference of ~0.75 m terms, the contact stiffness makes qacc
// sensitive to it at the 1e-2 level per float32 ulp of the base height, and
// at impact the damping term dn * v_z (~250 N) needs the velocity to better
// than float32's accumulated ~1e-6.
template <int K0, typename T_>
G1K_DEV void foot_contacts(const Cfg& cf, float mu, const T_ (&Rd)[9], const T_ (&p)[3], const T_* v,
                           float* anchor, float* fn_out, float (&Fb)[6]) {
  float Fw[3] = {-0.f, -0.f, -0.f}, Tw[3] = {-0.f, -0.f, -0.f};
  float R[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) R[m] = (float)Rd[m];
  constexpr T_ radius = (T_)G1M_CONTACT_RADIUS;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const T_ r0 = (T_)g1m_contact_pos[3 * (K0 + k)], r1 = (T_)g1m_contact_pos[3 * (K0 + k) + 1],
             r2 = (T_)g1m_contact_pos[3 * (K0 + k) + 2];
    const T_ rwd[3] = {cm(r0, Rd[0]) + cm(r1, Rd[1]) + cm(r2, Rd[2]), cm(r0, Rd[3]) + cm(r1, Rd[4]) + cm(r2, Rd[5]),
                       cm(r0, Rd[6]) + cm(r1, Rd[7]) + cm(r2, Rd[8])};
    const T_ pwd[3] = {rwd[0] + p[0], rwd[1] + p[1], rwd[2] + p[2]};
    const float rw[3] = {(float)rwd[0], (float)rwd[1], (float)rwd[2]};
    const float pw[3] = {(float)pwd[0], (float)pwd[1], (float)pwd[2]};
    const float depth = (float)(radius - pwd[2]);
    // point velocity in the body frame, then world: R (v_l + w x r)
    const T_ vb[3] = {v[3] + (cm(r2, v[1]) + cm(-r1, v[2])), v[4] + (cm(r0, v[2]) + cm(-r2, v[0])),
                      v[5] + (cm(r1, v[0]) + cm(-r0, v[1]))};
    const float vp[3] = {(float)(Rd[0] * vb[0] + Rd[1] * vb[1] + Rd[2] * vb[2]),
                         (float)(Rd[3] * vb[0] + Rd[4] * vb[1] + Rd[5] * vb[2]),
                         (float)(Rd[6] * vb[0] + Rd[7] * vb[1] + Rd[8] * vb[2])};
    const bool active = depth > 0.f;
    float fn = cf.kn * depth - cf.dn * vp[2];
    fn = fminf(fmaxf(fn, 0.f), 500.f);
    float ax = anchor[2 * k], ay = anchor[2 * k + 1];
    const bool fresh = ax >= ANCHOR_FREE;
    ax = fresh ? pw[0] : ax;
    ay = fresh ? pw[1] : ay;
    float ftx = -cf.kt * (float)(pwd[0] - (T_)ax) - cf.dtt * vp[0];
    float fty = -cf.kt * (float)(pwd[1] - (T_)ay) - cf.dtt * vp[1];
    const float ftn = fsqrt(ftx * ftx + fty * fty), fmx = mu * fn;
    const bool slip = ftn > fmx;
    const float sc = fmx * frcp(ftn + 1e-12f);
    ftx = slip ? ftx * sc : ftx;
    fty = slip ? fty * sc : fty;
    ax = slip ? pw[0] + ftx * cf.inv_kt : ax;
    ay = slip ? pw[1] + fty * cf.inv_kt : ay;
    fn = active ? fn : 0.f;
    ftx = active ? ftx : 0.f;
    fty = active ? fty : 0.f;
    anchor[2 * k] = active ? ax : ANCHOR_FREE;
    anchor[2 * k + 1] = active ? ay : ANCHOR_FREE;
    fn_out[k] = fn;
    Fw[0] += ftx; Fw[1] += fty; Fw[2] += fn;
    Tw[0] += rw[1] * fn - rw[2] * fty;
    Tw[1] += rw[2] * ftx - rw[0] * fn;
    Tw[2] += rw[0] * fty - rw[1] * ftx;
  }
  // body-frame wrench [R^T Tw; R^T Fw]
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    Fb[m] = R[m] * Tw[0] + R[3 + m] * Tw[1] + R[6 + m] * Tw[2];
    Fb[3 + m] = R[m] * Fw[0] + R[3 + m] * Fw[1] + R[6 + m] * Fw[2];
  }
}

// ------------------------------------------------------------ base solve
// IA0 a = -pA0 by packed Cholesky (src/spatial.h:254-278).
template <typename T> G1K_DEV void base_solve(const float (&A)[21], const float (&pA)[6], float (&a)[6]) {
  T L[21], y[6];
#pragma unroll
  for (int m = 0; m < 21; ++m) L[m] = (T)A[m];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
#pragma unroll
    for (int k = 0; k < j; ++k) L[pk(j, j)] -= L[pk(j, k)] * L[pk(j, k)];
    // 1 / sqrt(pivot) in one refined reciprocal square root (the pivots of
    // the positive definite base inertia are positive); in float64, 1 / sqrt
    const T inv = frsq(L[pk(j, j)]);
    L[pk(j, j)] = inv;                         // the diagonal is kept inverted
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
#pragma unroll
      for (int k = 0; k < j; ++k) L[pk(i, j)] -= L[pk(i, k)] * L[pk(j, k)];
      L[pk(i, j)] *= inv;
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    T acc = -(T)pA[i];
#pragma unroll
    for (int k = 0; k < i; ++k) acc -= L[pk(i, k)] * y[k];
    y[i] = acc * L[pk(i, i)];
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    T acc = y[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) acc -= L[pk(k, i)] * y[k];
    y[i] = acc * L[pk(i, i)];
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) a[m] = (float)y[m];
}

// ------------------------------------------------------------------ kernel
constexpr int ENVS_PER_WG = 64;
constexpr int XCH = 27;  // floats per chain contribution (21 + 6)

struct StepArgs {
  int nenv, nsub;
  Cfg cfg;
  float* qpos;
  float* qvel;
  float* anchor;
  const float* ctrl;
  float* foot_pos;
  float* foot_vel;
  float* contact_forces;
  float* qacc;
  float* tau;
  const float* mu;
  double* base64;
  int f64_solve;
  // judge mode: step the envs env_list[0 .. *list_count) of a batch of
  // `stride` envs once, store only their qacc (state is read, not written)
  const int* env_list;
  const int* list_count;
  int stride;
  // block-select judge (step_kernel<..., MODE_JUDGE_BLOCKS>): workgroup b
  // steps the envs of block b that judged_env() or sel_fallback marks, as the
  // list would hold.  Judge + forces (MODE_JUDGE_FORCES): a block with any
  // such env is stepped whole (state read-only; qacc, tau, anchors, contact
  // forces and feet stored) and any other block gets the forces only.
  const uint8_t* sel_fallback;
  int sel_phase, sel_period;
  // PD-target schedule inside one launch (ctrl_stride > 0): the targets of
  // substep s are block (ctrl_phase + s) / ctrl_period, at ctrl + block *
  // ctrl_stride (ctrl is the block of substep 0, ctrl_phase < ctrl_period
  // its substeps done before the launch)
  long long ctrl_stride;
  int ctrl_period, ctrl_phase;
  // Judge + forces with commit (commit_fallback, MODE_JUDGE_FORCES): a block
  // whose live envs are all in fallback is also integrated and stored as by
  // the plain step (a fallback env takes ABA's qacc whatever its
  // judgement).  If that block is judged this step, its deferred-judgement
  // record goes to snap ([141][nenv]: qpos 36, qvel 35 before the step,
  // tau_total 35, qacc 35) with sflag[e] = 1.  qacc_track (nullable) receives
  // the committed envs' qacc; stats[3] counts their env-steps.
  int commit_fallback;
  float* snap;
  uint8_t* sflag;
  float* qacc_track;
  unsigned* stats;
  // Fallback run (MODE_FALLBACK_RUN): a block whose live envs are all in
  // fallback is stepped nsub times as by the plain step (every substep
  // integrates ABA's qacc, as a fallback env's step does) and its record is
  // written at the substep whose phase, (run_phase0 + substep) % sel_period,
  // is the block's judge phase; blk_run[b] = 1 marks such a block, 0 any
  // other (which the kernel leaves alone).  In MODE_JUDGE_FORCES a non-null
  // blk_run makes the marked blocks return at once: the run has stepped them.
  uint8_t* blk_run;
  int run_phase0;
};
// Kernel variants (separate template instances, so the plain step's code is
// not touched by the judge modes).
constexpr int MODE_STEP = 0, MODE_JUDGE_BLOCKS = 1, MODE_JUDGE_FORCES = 2, MODE_FALLBACK_RUN = 3;
static_assert(ENVS_PER_WG == JUDGE_BLOCK, "judge blocks are the step's workgroups");

// Position of the k-th set bit (k from 0, k < popcount(m)) of m.
G1K_DEV int nth_set_bit(unsigned long long m, int k) {
  int lo = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int c = __popcll((m >> lo) & ((1ull << w) - 1ull));
    if (c <= k) { k -= c; lo += w; }
  }
  return lo;
}

// Per-lane chain slots (legs use 6 of 7).
constexpr int MAXJ = 7;
// The 2-wave split keeps 12-14 joints per wave between the backward and the
// forward pass; with every joint's U/D (6 floats) in registers it spills
// ~400 B per lane to scratch.  It parks U/D in LDS instead (43 KB per
// workgroup; two workgroups per CU fit in 160 KB).
template <int SPLIT> constexpr bool W_IN_LDS = SPLIT == 2;

template <int W> struct WaveChain;
template <> struct WaveChain<0> { static constexpr int first = 1, n = 6, root = 0, foot = 6, k0 = 0; };
template <> struct WaveChain<1> { static constexpr int first = 7, n = 6, root = 0, foot = 12, k0 = 4; };
template <> struct WaveChain<2> { static constexpr int first = 16, n = 7, root = 15, foot = -1, k0 = -1; };
template <> struct WaveChain<3> { static constexpr int first = 23, n = 7, root = 15, foot = -1, k0 = -1; };

using Waist = ChainOps<13, 3>;

// Chain part before the exchange: kinematics, contacts, backward pass.
template <int W, typename G>
G1K_DEV void chain_pre(const Cfg& cf, const float* q, const float* qd, const float* ctl, float mu,
                       const float (&v0)[6], const float (&v15)[6], const G (&R0)[9], const G (&p0)[3],
                       const G (&v0g)[6],
                       Joint* jt, float* anchor, float* fn, float (&footp)[3], float (&footv)[3],
                       float (&Pout)[21], float (&pout)[6], bool want_tau, float* tc, float (&Fpel)[6]) {
  using WC = WaveChain<W>;
  using Ch = ChainOps<WC::first, WC::n>;
  float vleaf[6];
  Ch::fk(q, qd, (WC::root == 0) ? v0 : v15, jt, vleaf);
  float A[21], pA[6];
  init_inertia<WC::first + WC::n - 1>(A);
#pragma unroll
  for (int m = 0; m < 6; ++m) pA[m] = jt[WC::n - 1].bias[m];
  float Fb[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if constexpr (WC::foot > 0) {
    G Rf[9], pf[3], vf[6];
    Ch::world(R0, p0, v0g, jt, q, qd, Rf, pf, vf);
    if constexpr (sizeof(G) == sizeof(float)) {
#pragma unroll
      for (int m = 0; m < 6; ++m) vf[m] = vleaf[m];
    }
    foot_contacts<WC::k0>(cf, mu, Rf, pf, vf, anchor, fn, Fb);
#pragma unroll
    for (int m = 0; m < 6; ++m) pA[m] -= Fb[m];
    // foot outputs (src/g1_kernel.cu:128-145): origin position, world linear velocity
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      footp[m] = (float)pf[m];
      footv[m] = (float)Rf[3 * m] * vleaf[3] + (float)Rf[3 * m + 1] * vleaf[4] + (float)Rf[3 * m + 2] * vleaf[5];
    }
  }
  Ch::backward(cf, A, pA, q, qd, ctl, jt, Pout, pout);
  if constexpr (WC::foot > 0) {
    if (want_tau) Ch::wrench_to_joints(Fb, jt, tc);    // Fb ends in the pelvis frame
#pragma unroll
    for (int m = 0; m < 6; ++m) Fpel[m] = Fb[m];
  }
}

template <int W>
G1K_DEV void chain_post(const float (&a0)[6], const float (&a15)[6], const Joint* jt, float* qdd) {
  using WC = WaveChain<W>;
  float aleaf[6];
  ChainOps<WC::first, WC::n>::forward((WC::root == 0) ? a0 : a15, jt, qdd, aleaf);
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = 0 .. N_-1.
template <int N_, typename F> G1K_DEV void static_for(F&& f) {
  if constexpr (N_ > 0) {
    static_for<N_ - 1>(f);
    f(std::integral_constant<int, N_ - 1>{});
  }
}

// Limb chains: 0 left leg, 1 right leg, 2 left arm, 3 right arm.
G1K_DEV int chain_h0(int c) { return c == 0 ? 0 : c == 1 ? 6 : c == 2 ? 15 : 22; }
G1K_DEV int chain_n(int c) { return c < 2 ? 6 : 7; }

template <typename GeoT>
G1K_DEV void dispatch_pre(int chain, const Cfg& cf, const float* q, const float* qd, const float* ctl, float mu,
                          const float (&v0)[6], const float (&v15)[6], const GeoT (&R0g)[9], const GeoT (&p0g)[3],
                          const GeoT (&v0g)[6], Joint* jt, float* anc, float* fn, float (&footp)[3],
                          float (&footv)[3], float (&P)[21], float (&p)[6], bool want_tau, float* tc,
                          float (&Fpel)[6]) {
  switch (chain) {
    case 0: chain_pre<0, GeoT>(cf, q, qd, ctl, mu, v0, v15, R0g, p0g, v0g, jt, anc, fn, footp, footv, P, p, want_tau, tc, Fpel); break;
    case 1: chain_pre<1, GeoT>(cf, q, qd, ctl, mu, v0, v15, R0g, p0g, v0g, jt, anc, fn, footp, footv, P, p, want_tau, tc, Fpel); break;
    case 2: chain_pre<2, GeoT>(cf, q, qd, ctl, mu, v0, v15, R0g, p0g, v0g, jt, anc, fn, footp, footv, P, p, want_tau, tc, Fpel); break;
    default: chain_pre<3, GeoT>(cf, q, qd, ctl, mu, v0, v15, R0g, p0g, v0g, jt, anc, fn, footp, footv, P, p, want_tau, tc, Fpel); break;
  }
}

G1K_DEV void dispatch_post(int chain, const float (&a0)[6], const float (&a15)[6], const Joint* jt, float* qdd) {
  switch (chain) {
    case 0: chain_post<0>(a0, a15, jt, qdd); break;
    case 1: chain_post<1>(a0, a15, jt, qdd); break;
    case 2: chain_post<2>(a0, a15, jt, qdd); break;
    default: chain_post<3>(a0, a15, jt, qdd); break;
  }
}

// SPLIT waves per workgroup of 64 envs: with 4, wave w owns chain w; with 2,
// wave 0 owns both legs (and all 8 contacts) and wave 1 both arms, which
// halves the pelvis/waist work every wave repeats when there are enough envs
// to fill the chip anyway.
//
// ROLE >= 0 compiles the body for one wave index (the 2-wave split, where
// each role keeps two chains in registers); ROLE < 0 dispatches the chain
// code on the wave index at run time and keeps a single copy of the shared
// pelvis/waist code (the 4-wave split).
template <bool F64_BASE, bool F64_CONTACT, int SPLIT, int ROLE, int MODE>
G1K_DEV void step_body(const StepArgs& args, float (&xch)[2][SPLIT][XCH][ENVS_PER_WG],
                       float (&fcx)[2][2][6][ENVS_PER_WG], float* wsh, bool commit = false, bool record = false) {
  constexpr int CPW = 4 / SPLIT;                 // chains per wave
  constexpr bool JF = MODE == MODE_JUDGE_FORCES;
  constexpr bool FR = MODE == MODE_FALLBACK_RUN;
  const int lane = threadIdx.x & 63;
  const int wave = ROLE >= 0 ? ROLE : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool judge = MODE == MODE_JUDGE_BLOCKS || (MODE == MODE_STEP && args.env_list != nullptr);
  bool valid;
  int e;
  size_t N;
  if constexpr (JF || FR) {
    // the whole block, in env order (the kernel has checked it holds a judged
    // or fallback env, or that all its live envs are in fallback)
    const int e_raw = blockIdx.x * ENVS_PER_WG + lane;
    valid = e_raw < args.nenv;
    e = valid ? e_raw : args.nenv - 1;
    N = (size_t)args.nenv;
  } else if constexpr (MODE == MODE_JUDGE_BLOCKS) {
    // this workgroup's block of 64 envs: all of them on the block's judge
    // step, else its fallback envs, packed into the low lanes in env order.
    // Every wave reads the same 64 flags, so the mask and the early exit are
    // uniform per workgroup.
    const int e_raw = blockIdx.x * ENVS_PER_WG + lane;
    const bool blk = (int)(blockIdx.x % (unsigned)args.sel_period) == args.sel_phase;
    const bool in = e_raw < args.nenv && (blk || (args.sel_fallback && args.sel_fallback[e_raw]));
    const unsigned long long m = __ballot(in);
    if (m == 0ull) return;
    const int cnt = __popcll(m);
    valid = lane < cnt;
    e = blockIdx.x * ENVS_PER_WG + nth_set_bit(m, valid ? lane : cnt - 1);
    N = (size_t)args.nenv;
  } else {
    const int nenv = judge ? *args.list_count : args.nenv;
    if (judge && (int)(blockIdx.x * ENVS_PER_WG) >= nenv) return;   // uniform per workgroup
    const int i_raw = blockIdx.x * ENVS_PER_WG + lane;
    valid = i_raw < nenv;
    const int i_env = valid ? i_raw : nenv - 1;  // tail lanes mirror the last env, never store
    e = judge ? args.env_list[i_env] : i_env;
    N = judge ? (size_t)args.stride : (size_t)nenv;
  }
  const Cfg cf = args.cfg;
  const int e_ = e;
  const bool legs = wave * CPW < 2;            // this wave owns the leg chains

  // ---- load (all waves: base + waist; own chains)
  using BaseT = typename std::conditional<F64_BASE, double, float>::type;
  BaseT bp[3], bq[4];
  float vw[3], om[3];
  if constexpr (F64_BASE) {
#pragma unroll
    for (int m = 0; m < 3; ++m) bp[m] = args.base64[m * N + e];
#pragma unroll
    for (int m = 0; m < 4; ++m) bq[m] = args.base64[(3 + m) * N + e];
  } else {
#pragma unroll
    for (int m = 0; m < 3; ++m) bp[m] = args.qpos[m * N + e];
#pragma unroll
    for (int m = 0; m < 4; ++m) bq[m] = args.qpos[(3 + m) * N + e];
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    vw[m] = args.qvel[m * N + e];
    om[m] = args.qvel[(3 + m) * N + e];
  }
  // this wave's PD targets (waist + own chains) from one target block
  auto load_ctrl = [&](const float* cp, float (&w3)[3], float (&ct)[CPW][MAXJ]) {
#pragma unroll
    for (int m = 0; m < 3; ++m) w3[m] = cp[(12 + m) * N + e];
    static_for<CPW>([&](auto C) {
      constexpr int c = decltype(C)::value;
      const int ch = wave * CPW + c, h0 = chain_h0(ch), nj = chain_n(ch);
#pragma unroll
      for (int m = 0; m < MAXJ; ++m) ct[c][m] = cp[(h0 + (m < nj ? m : nj - 1)) * N + e];
    });
  };
  float wq[3], wqd[3], wct[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    wq[m] = args.qpos[(7 + 12 + m) * N + e];
    wqd[m] = args.qvel[(6 + 12 + m) * N + e];
  }
  float cq[CPW][MAXJ], cqd[CPW][MAXJ], cct[CPW][MAXJ];
  load_ctrl(args.ctrl, wct, cct);
  float anc[CPW][8];
  float mu = cf.mu;
  if (legs && args.mu) mu = args.mu[e];
  static_for<CPW>([&](auto C) {
    constexpr int c = decltype(C)::value;
    const int ch = wave * CPW + c, h0 = chain_h0(ch), nj = chain_n(ch);
#pragma unroll
    for (int m = 0; m < MAXJ; ++m) {
      const int mm = m < nj ? m : nj - 1;
      cq[c][m] = args.qpos[(7 + h0 + mm) * N + e];
      cqd[c][m] = args.qvel[(6 + h0 + mm) * N + e];
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) anc[c][m] = legs ? args.anchor[(ch * 8 + m) * N + e] : ANCHOR_FREE;
  });
  // deferred-judgement record, part 1: the state before the step, in the
  // network's input order (qpos rows 0-35, qvel rows 36-70)
  // The record's addresses are computed where it is written (the env index
  // goes through an empty asm): hoisted out of the fallback run's substep
  // loop, its ~140 row addresses stayed live and the 2-wave instance spilled.
  auto opaque = [](int v) __attribute__((always_inline)) {
    asm volatile("" : "+v"(v));
    return v;
  };
  auto record_state = [&]() __attribute__((always_inline)) {
    float* sp = args.snap;
    const int e = opaque(e_);
    if (wave == 0) {
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        sp[m * N + e] = (float)bp[m];
        sp[(36 + m) * N + e] = vw[m];
        sp[(36 + 3 + m) * N + e] = om[m];
        sp[(7 + 12 + m) * N + e] = wq[m];
        sp[(36 + 6 + 12 + m) * N + e] = wqd[m];
      }
#pragma unroll
      for (int m = 0; m < 4; ++m) sp[(3 + m) * N + e] = (float)bq[m];
    }
    static_for<CPW>([&](auto C) {
      constexpr int c = decltype(C)::value;
      const int ch = wave * CPW + c, h0 = chain_h0(ch), nj = chain_n(ch);
#pragma unroll
      for (int m = 0; m < MAXJ; ++m) {
        if (m < nj) {
          sp[(7 + h0 + m) * N + e] = cq[c][m];
          sp[(36 + 6 + h0 + m) * N + e] = cqd[c][m];
        }
      }
    });
  };
  if constexpr (JF) {
    if (record && valid) record_state();
  }
  // target schedule: the next block's targets are loaded one substep before
  // they apply, so their memory latency hides behind that substep
  const bool sched = (MODE == MODE_STEP || MODE == MODE_FALLBACK_RUN) && args.ctrl_stride > 0;
  const float* cnext = args.ctrl;
  int cpos = args.ctrl_phase;                    // substeps of the current block already stepped
  float nwt[3], nct[CPW][MAXJ];

  float qacc_base[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, qdd_w[3] = {0.f, 0.f, 0.f}, qdd_c[CPW][MAXJ];
  float fn[CPW][4], footp[CPW][3], footv[CPW][3];
  // total generalized force of the last substep (servo + limits + J^T f);
  // the fallback run computes it on its record substep only
  const bool tau_out = args.tau != nullptr;
  bool want_tau = tau_out;
  const bool f64_solve = args.f64_solve;
  float tau_c[CPW][MAXJ], tau_w[3] = {0.f, 0.f, 0.f}, tau_b[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < CPW; ++c) {
#pragma unroll
    for (int m = 0; m < MAXJ; ++m) { qdd_c[c][m] = 0.f; tau_c[c][m] = 0.f; }
#pragma unroll
    for (int m = 0; m < 4; ++m) fn[c][m] = 0.f;
#pragma unroll
    for (int m = 0; m < 3; ++m) { footp[c][m] = 0.f; footv[c][m] = 0.f; }
  }
  const float dt = cf.dt;
  int rec_phase = -1;
  if constexpr (FR) rec_phase = (int)(blockIdx.x % (unsigned)args.sel_period);
  for (int sub = 0; sub < args.nsub; ++sub) {
    const int buf = sub & 1;
    // fallback run: this substep is the block's judged step (uniform)
    bool rec_now = false;
    if constexpr (FR) {
      rec_now = args.snap && (args.run_phase0 + sub) % args.sel_period == rec_phase;
      want_tau = rec_now;
      if (rec_now && valid) record_state();
    }
    if (sched) {                                 // uniform
      if (cpos == args.ctrl_period) {            // block boundary: the prefetched targets apply
#pragma unroll
        for (int m = 0; m < 3; ++m) wct[m] = nwt[m];
#pragma unroll
        for (int c = 0; c < CPW; ++c)
#pragma unroll
          for (int m = 0; m < MAXJ; ++m) cct[c][m] = nct[c][m];
        cpos = 0;
      }
      if (cpos == args.ctrl_period - 1 && sub + 1 < args.nsub) {
        cnext += args.ctrl_stride;
        load_ctrl(cnext, nwt, nct);
      }
      ++cpos;
    }
    // ---- pelvis kinematics (src/g1_core.h:30-37)
    float R0[9], p0[3], v0[6];
    {
      const BaseT n2 = bq[0] * bq[0] + bq[1] * bq[1] + bq[2] * bq[2] + bq[3] * bq[3];
      const BaseT inv = frsq(n2);
      const float w = (float)(bq[0] * inv), x = (float)(bq[1] * inv), y = (float)(bq[2] * inv),
                  z = (float)(bq[3] * inv);
      R0[0] = 1 - 2 * (y * y + z * z); R0[1] = 2 * (x * y - w * z); R0[2] = 2 * (x * z + w * y);
      R0[3] = 2 * (x * y + w * z); R0[4] = 1 - 2 * (x * x + z * z); R0[5] = 2 * (y * z - w * x);
      R0[6] = 2 * (x * z - w * y); R0[7] = 2 * (y * z + w * x); R0[8] = 1 - 2 * (x * x + y * y);
#pragma unroll
      for (int m = 0; m < 3; ++m) p0[m] = (float)bp[m];
      v0[0] = om[0]; v0[1] = om[1]; v0[2] = om[2];
#pragma unroll
      for (int m = 0; m < 3; ++m) v0[3 + m] = R0[m] * vw[0] + R0[3 + m] * vw[1] + R0[6 + m] * vw[2];
    }
    // pelvis pose for the foot-contact kinematics (float64 when F64_CONTACT)
    using GeoT = typename std::conditional<F64_CONTACT, double, float>::type;
    GeoT R0g[9], p0g[3], v0g[6];
    if constexpr (F64_CONTACT) {
      const double q0 = bq[0], q1 = bq[1], q2 = bq[2], q3 = bq[3];
      const double inv = 1.0 / sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
      const double w = q0 * inv, x = q1 * inv, y = q2 * inv, z = q3 * inv;
      R0g[0] = 1 - 2 * (y * y + z * z); R0g[1] = 2 * (x * y - w * z); R0g[2] = 2 * (x * z + w * y);
      R0g[3] = 2 * (x * y + w * z); R0g[4] = 1 - 2 * (x * x + z * z); R0g[5] = 2 * (y * z - w * x);
      R0g[6] = 2 * (x * z - w * y); R0g[7] = 2 * (y * z + w * x); R0g[8] = 1 - 2 * (x * x + y * y);
#pragma unroll
      for (int m = 0; m < 3; ++m) p0g[m] = (double)bp[m];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        v0g[m] = (double)om[m];
        v0g[3 + m] = R0g[m] * (double)vw[0] + R0g[3 + m] * (double)vw[1] + R0g[6 + m] * (double)vw[2];
      }
    } else {
#pragma unroll
      for (int m = 0; m < 9; ++m) R0g[m] = R0[m];
#pragma unroll
      for (int m = 0; m < 3; ++m) p0g[m] = p0[m];
#pragma unroll
      for (int m = 0; m < 6; ++m) v0g[m] = v0[m];
    }
    // ---- waist kinematics (bodies 13-15), needed by the arms and the merge
    Joint wj[3];
    float v15[6];
    Waist::fk(wq, wqd, v0, wj, v15);

    // ---- own chains before the exchange; contributions summed per wave
    Joint cj[CPW][MAXJ];
    float Pw[21], pw[6], tc[CPW][MAXJ], Fw[6];
#pragma unroll
    for (int m = 0; m < 21; ++m) Pw[m] = -0.f;   // sums start at -0, which -0 + x drops
#pragma unroll
    for (int m = 0; m < 6; ++m) { pw[m] = -0.f; Fw[m] = -0.f; }
    static_for<CPW>([&](auto C) {
      constexpr int c = decltype(C)::value;
      float Pc[21], pc[6], Fpel[6];
#pragma unroll
      for (int m = 0; m < MAXJ; ++m) tc[c][m] = 0.f;
#pragma unroll
      for (int m = 0; m < 21; ++m) Pc[m] = -0.f;
#pragma unroll
      for (int m = 0; m < 6; ++m) { pc[m] = -0.f; Fpel[m] = -0.f; }
      dispatch_pre<GeoT>(wave * CPW + c, cf, cq[c], cqd[c], cct[c], mu, v0, v15, R0g, p0g, v0g, cj[c], anc[c],
                         fn[c], footp[c], footv[c], Pc, pc, want_tau, tc[c], Fpel);
#pragma unroll
      for (int m = 0; m < 21; ++m) Pw[m] += Pc[m];
#pragma unroll
      for (int m = 0; m < 6; ++m) { pw[m] += pc[m]; Fw[m] += Fpel[m]; }
    });
    if constexpr (W_IN_LDS<SPLIT>) {
      // U/D of every joint parked in this lane's own LDS slots until the
      // forward pass (no other lane or wave reads them)
      static_for<CPW>([&](auto C) {
        constexpr int c = decltype(C)::value;
        const int nj = chain_n(wave * CPW + c);
#pragma unroll
        for (int m = 0; m < MAXJ; ++m)
          if (m < nj) {
#pragma unroll
            for (int k = 0; k < 6; ++k) wsh[(((wave * CPW + c) * MAXJ + m) * 6 + k) * ENVS_PER_WG + lane] = cj[c][m].W[k];
          }
      });
    }
#pragma unroll
    for (int m = 0; m < 21; ++m) xch[buf][wave][m][lane] = Pw[m];
#pragma unroll
    for (int m = 0; m < 6; ++m) xch[buf][wave][21 + m][lane] = pw[m];
    if constexpr (SPLIT == 4) {
      if (want_tau && legs) {
#pragma unroll
        for (int m = 0; m < 6; ++m) fcx[buf][wave][m][lane] = Fw[m];
      }
    }
    __syncthreads();

    // ---- merge: torso <- arms, waist backward, pelvis <- legs + waist
    float A15[21], p15[6], A0[21], pA0[6];
    init_inertia<15>(A15);
    init_inertia<0>(A0);
    bias_force<0>(v0, pA0);
#pragma unroll
    for (int m = 0; m < 6; ++m) p15[m] = wj[2].bias[m];
#pragma unroll
    for (int w = 0; w < SPLIT; ++w) {
      const bool to_pelvis = w * CPW < 2;
      // The 27 reads of one contribution are issued back to back and waited
      // for once. Left to itself the scheduler (at the register limit) pairs
      // each read with its add, one LDS round trip per pair of floats.
      float t[XCH];
#pragma unroll
      for (int m = 0; m < XCH; ++m) t[m] = xch[buf][w][m][lane];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int m = 0; m < 21; ++m) {
        if (to_pelvis) A0[m] += t[m];
        else A15[m] += t[m];
      }
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        if (to_pelvis) pA0[m] += t[21 + m];
        else p15[m] += t[21 + m];
      }
    }
    Waist::backward(cf, A15, p15, wq, wqd, wct, wj, A0, pA0);

    // ---- base: IA0 a0' = -pA0 (gravity-free), forward through the waist
    float a0[6];
    if (f64_solve) base_solve<double>(A0, pA0, a0);
    else base_solve<float>(A0, pA0, a0);
    float a15[6];
    if constexpr (W_IN_LDS<SPLIT>) {
      static_for<CPW>([&](auto C) {
        constexpr int c = decltype(C)::value;
        const int nj = chain_n(wave * CPW + c);
#pragma unroll
        for (int m = 0; m < MAXJ; ++m)
          if (m < nj) {
#pragma unroll
            for (int k = 0; k < 6; ++k) cj[c][m].W[k] = wsh[(((wave * CPW + c) * MAXJ + m) * 6 + k) * ENVS_PER_WG + lane];
          }
      });
    }
    Waist::forward(a0, wj, qdd_w, a15);
    static_for<CPW>([&](auto C) {
      constexpr int c = decltype(C)::value;
      dispatch_post(wave * CPW + c, a0, a15, cj[c], qdd_c[c]);
    });
    if (want_tau) {
      static_for<CPW>([&](auto C) {
        constexpr int c = decltype(C)::value;
        const int nj = chain_n(wave * CPW + c);
#pragma unroll
        for (int m = 0; m < MAXJ; ++m) tau_c[c][m] = (m < nj ? cj[c][m].tau : 0.f) + tc[c][m];
      });
#pragma unroll
      for (int m = 0; m < 3; ++m) tau_w[m] = wj[m].tau;
      // base: world force R0 F.l, body torque F.a (convention of g1_c_fd's tau)
      float F[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        if constexpr (SPLIT == 4) F[m] = fcx[buf][0][m][lane] + fcx[buf][1][m][lane];
        else F[m] = Fw[m];            // wave 0 holds both legs (only wave 0 stores it)
      }
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        tau_b[m] = R0[3 * m] * F[3] + R0[3 * m + 1] * F[4] + R0[3 * m + 2] * F[5];
        tau_b[3 + m] = F[m];
      }
    }
    // base qacc in MuJoCo convention (src/g1_core.h:202-206), gravity restored:
    // world linear = R0 (a0'_l + w x v_l) + g, angular = body frame a0'_a
    {
      const float cl[3] = {a0[3] + (v0[1] * v0[5] - v0[2] * v0[4]), a0[4] + (v0[2] * v0[3] - v0[0] * v0[5]),
                           a0[5] + (v0[0] * v0[4] - v0[1] * v0[3])};
#pragma unroll
      for (int m = 0; m < 3; ++m) qacc_base[m] = R0[3 * m] * cl[0] + R0[3 * m + 1] * cl[1] + R0[3 * m + 2] * cl[2];
      qacc_base[2] += cf.gz;
      qacc_base[3] = a0[0]; qacc_base[4] = a0[1]; qacc_base[5] = a0[2];
    }

    if constexpr (FR) {
      // fallback run, record part 2: this substep's tau_total and qacc
      if (rec_now && valid) {
        float* rp = args.snap;
        const int e = opaque(e_);
        if (wave == 0) {
#pragma unroll
          for (int m = 0; m < 6; ++m) {
            rp[(71 + m) * N + e] = tau_b[m];
            rp[(106 + m) * N + e] = qacc_base[m];
          }
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            rp[(71 + 6 + 12 + m) * N + e] = tau_w[m];
            rp[(106 + 6 + 12 + m) * N + e] = qdd_w[m];
          }
          args.sflag[e] = 1;
        }
        static_for<CPW>([&](auto C) {
          constexpr int c = decltype(C)::value;
          const int ch = wave * CPW + c, h0 = chain_h0(ch), nj = chain_n(ch);
#pragma unroll
          for (int m = 0; m < MAXJ; ++m) {
            if (m < nj) {
              rp[(71 + 6 + h0 + m) * N + e] = tau_c[c][m];
              rp[(106 + 6 + h0 + m) * N + e] = qdd_c[c][m];
            }
          }
        });
      }
    }

    // ---- semi-implicit Euler (src/g1_core.h:212-222; quaternion exponential
    // src/spatial.h:73-79), the same bits as the surrogate's commit
    euler_base(dt, qacc_base, vw, om, bp, bq);
#pragma unroll
    for (int m = 0; m < 3; ++m) euler_hinge(dt, qdd_w[m], wqd[m], wq[m]);
#pragma unroll
    for (int c = 0; c < CPW; ++c)
#pragma unroll
      for (int m = 0; m < MAXJ; ++m) euler_hinge(dt, qdd_c[c][m], cqd[c][m], cq[c][m]);
  }

  // ---- store
  if (!valid || args.nsub <= 0) return;
  if (judge) {
    if (wave == 0) {
#pragma unroll
      for (int m = 0; m < 6; ++m) args.qacc[m * N + e] = qacc_base[m];
#pragma unroll
      for (int m = 0; m < 3; ++m) args.qacc[(6 + 12 + m) * N + e] = qdd_w[m];
    }
    static_for<CPW>([&](auto C) {
      constexpr int c = decltype(C)::value;
      const int ch = wave * CPW + c, h0 = chain_h0(ch), nj = chain_n(ch);
#pragma unroll
      for (int m = 0; m < MAXJ; ++m)
        if (m < nj) args.qacc[(6 + h0 + m) * N + e] = qdd_c[c][m];
    });
    return;
  }
  // judge + forces with commit: the committed block's state, its tracked
  // qacc, and (judged) the rest of its deferred-judgement record
  float* trk = (JF && commit) || FR ? args.qacc_track : nullptr;
  float* rec = JF && record ? args.snap : nullptr;
  if constexpr (JF || FR) {
    if ((FR || commit) && wave == 0) {
      // the active lanes are the valid ones (the others have returned)
      const unsigned long long vm = __ballot(true);
      if (lane == 0 && args.stats) atomicAdd(&args.stats[3], (unsigned)__popcll(vm) * (unsigned)args.nsub);
    }
  }
  if (wave == 0) {
    if (!JF || commit) {
#pragma unroll
      for (int m = 0; m < 3; ++m) args.qpos[m * N + e] = (float)bp[m];
#pragma unroll
      for (int m = 0; m < 4; ++m) args.qpos[(3 + m) * N + e] = (float)bq[m];
      if constexpr (F64_BASE) {
#pragma unroll
        for (int m = 0; m < 3; ++m) args.base64[m * N + e] = (double)bp[m];
#pragma unroll
        for (int m = 0; m < 4; ++m) args.base64[(3 + m) * N + e] = (double)bq[m];
      }
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        args.qvel[m * N + e] = vw[m];
        args.qvel[(3 + m) * N + e] = om[m];
        args.qpos[(7 + 12 + m) * N + e] = wq[m];
        args.qvel[(6 + 12 + m) * N + e] = wqd[m];
      }
    }
    // (a committed block's qacc and tau are read by nobody: the network
    // kernel leaves the block alone, the record has its own copy)
    if (tau_out && !(JF && commit)) {
#pragma unroll
      for (int m = 0; m < 6; ++m) args.tau[m * N + e] = tau_b[m];
#pragma unroll
      for (int m = 0; m < 3; ++m) args.tau[(6 + 12 + m) * N + e] = tau_w[m];
    }
    if (args.qacc && !(JF && commit)) {
#pragma unroll
      for (int m = 0; m < 6; ++m) args.qacc[m * N + e] = qacc_base[m];
#pragma unroll
      for (int m = 0; m < 3; ++m) args.qacc[(6 + 12 + m) * N + e] = qdd_w[m];
    }
    if (trk) {
#pragma unroll
      for (int m = 0; m < 6; ++m) trk[m * N + e] = qacc_base[m];
#pragma unroll
      for (int m = 0; m < 3; ++m) trk[(6 + 12 + m) * N + e] = qdd_w[m];
    }
    if (rec) {
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        rec[(71 + m) * N + e] = tau_b[m];
        rec[(106 + m) * N + e] = qacc_base[m];
      }
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        rec[(71 + 6 + 12 + m) * N + e] = tau_w[m];
        rec[(106 + 6 + 12 + m) * N + e] = qdd_w[m];
      }
      args.sflag[e] = 1;
    }
  }
  static_for<CPW>([&](auto C) {
    constexpr int c = decltype(C)::value;
    const int ch = wave * CPW + c, h0 = chain_h0(ch), nj = chain_n(ch);
#pragma unroll
    for (int m = 0; m < MAXJ; ++m) {
      if (m < nj) {
        if (!JF || commit) {
          args.qpos[(7 + h0 + m) * N + e] = cq[c][m];
          args.qvel[(6 + h0 + m) * N + e] = cqd[c][m];
        }
        if (args.qacc && !(JF && commit)) args.qacc[(6 + h0 + m) * N + e] = qdd_c[c][m];
        if (tau_out && !(JF && commit)) args.tau[(6 + h0 + m) * N + e] = tau_c[c][m];
        if (trk) trk[(6 + h0 + m) * N + e] = qdd_c[c][m];
        if (rec) {
          rec[(71 + 6 + h0 + m) * N + e] = tau_c[c][m];
          rec[(106 + 6 + h0 + m) * N + e] = qdd_c[c][m];
        }
      }
    }
    if (legs) {
#pragma unroll
      for (int m = 0; m < 8; ++m) args.anchor[(ch * 8 + m) * N + e] = anc[c][m];
      if (args.contact_forces) {
#pragma unroll
        for (int m = 0; m < 4; ++m) args.contact_forces[(ch * 4 + m) * N + e] = fn[c][m];
      }
      if (args.foot_pos) {
#pragma unroll
        for (int m = 0; m < 3; ++m) args.foot_pos[(ch * 3 + m) * N + e] = footp[c][m];
      }
      if (args.foot_vel) {
#pragma unroll
        for (int m = 0; m < 3; ++m) args.foot_vel[(ch * 3 + m) * N + e] = footv[c][m];
      }
    }
  });
}

// ------------------------------------------------------- surrogate support
// Total generalized force without ABA: pelvis and leg kinematics, the 8 foot
// contacts (anchors and normal forces updated exactly as by the step),
// J^T f of the contact wrenches and the servo + limit torques of all 29
// hinges.  A workgroup holds 64 envs (lane = env) in 2 waves, one per leg:
// wave L runs leg L's kinematics and 4 contacts and the servos of its leg and
// of half the upper body; the pelvis wrench of the right leg reaches wave 0
// through LDS.
struct ForcesArgs {
  int nenv;
  Cfg cfg;
  const float* qpos;
  const float* qvel;
  float* anchor;
  const float* ctrl;
  const float* mu;
  float* tau;
  float* contact_forces;
  float* foot_pos;
  float* foot_vel;
};

// Servo + limit torques of hinges h0 .. h0+n-1, stored as tau rows 6 + h.
template <int h0, int n> G1K_DEV void servo_store(const ForcesArgs& a, size_t N, int e, bool live) {
  static_for<n>([&](auto H) {
    constexpr int h = h0 + decltype(H)::value;
    float b;
    const float t = servo_torque<h>(a.cfg, a.qpos[(7 + h) * N + e], a.qvel[(6 + h) * N + e], a.ctrl[h * N + e], b);
    if (live) a.tau[(6 + h) * N + e] = t;
  });
}

// Leg L: kinematics, contacts (anchors, normal forces, feet), servo + J^T f
// of its 6 hinges; returns the legs' contact wrench on the pelvis in Fb.
template <int L>
G1K_DEV void leg_forces(const ForcesArgs& a, size_t N, int e, bool live, const float (&R0)[9], const float (&p0)[3],
                        const float (&v0)[6], float (&Fb)[6]) {
  using WC = WaveChain<L>;
  using Ch = ChainOps<WC::first, 6>;
  constexpr int h0 = WC::first - 1;
  const Cfg& cf = a.cfg;
  float q[6], qd[6], ct[6], anc[8], fn[4], tc[6], vleaf[6], Rf[9], pf[3];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    q[m] = a.qpos[(7 + h0 + m) * N + e];
    qd[m] = a.qvel[(6 + h0 + m) * N + e];
    ct[m] = a.ctrl[(h0 + m) * N + e];
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) anc[m] = a.anchor[(L * 8 + m) * N + e];
  const float mu = a.mu ? a.mu[e] : cf.mu;
  Joint jt[6];
  Ch::fk(q, qd, v0, jt, vleaf);
  Ch::world(R0, p0, v0, jt, q, qd, Rf, pf, vleaf);   // float: vleaf kept
  foot_contacts<WC::k0>(cf, mu, Rf, pf, vleaf, anc, fn, Fb);
  Ch::wrench_to_joints(Fb, jt, tc);
  if (!live) return;
  static_for<6>([&](auto H) {
    constexpr int m = decltype(H)::value;
    float b;
    a.tau[(6 + h0 + m) * N + e] = servo_torque<h0 + m>(cf, q[m], qd[m], ct[m], b) + tc[m];
  });
#pragma unroll
  for (int m = 0; m < 8; ++m) a.anchor[(L * 8 + m) * N + e] = anc[m];
  if (a.contact_forces) {
#pragma unroll
    for (int m = 0; m < 4; ++m) a.contact_forces[(L * 4 + m) * N + e] = fn[m];
  }
  // foot body origin and its world linear velocity, as the step reports them
  if (a.foot_pos) {
#pragma unroll
    for (int m = 0; m < 3; ++m) a.foot_pos[(L * 3 + m) * N + e] = pf[m];
  }
  if (a.foot_vel) {
#pragma unroll
    for (int m = 0; m < 3; ++m)
      a.foot_vel[(L * 3 + m) * N + e] = Rf[3 * m] * vleaf[3] + Rf[3 * m + 1] * vleaf[4] + Rf[3 * m + 2] * vleaf[5];
  }
}

// One 64-env block of the forces: waves 0 and 1 run a leg each (with half of
// the upper-body servos); further waves of the workgroup have nothing to do.
// Every thread of the workgroup must call it (one barrier).
G1K_DEV void forces_block(const ForcesArgs& a, int wave, float (&fx)[6][ENVS_PER_WG]) {
  const int lane = threadIdx.x & 63;
  const int e_raw = blockIdx.x * ENVS_PER_WG + lane;
  const bool live = e_raw < a.nenv;
  const int e = live ? e_raw : a.nenv - 1;        // tail lanes mirror the last env, never store
  const size_t N = (size_t)a.nenv;
  float R0[9], p0[3], v0[6], Fb[6];
  if (wave < 2) {
    float bq[4], vw[3], om[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      p0[m] = a.qpos[m * N + e];
      vw[m] = a.qvel[m * N + e];
      om[m] = a.qvel[(3 + m) * N + e];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) bq[m] = a.qpos[(3 + m) * N + e];
    const float inv = frsq(bq[0] * bq[0] + bq[1] * bq[1] + bq[2] * bq[2] + bq[3] * bq[3]);
    const float w = bq[0] * inv, x = bq[1] * inv, y = bq[2] * inv, z = bq[3] * inv;
    R0[0] = 1 - 2 * (y * y + z * z); R0[1] = 2 * (x * y - w * z); R0[2] = 2 * (x * z + w * y);
    R0[3] = 2 * (x * y + w * z); R0[4] = 1 - 2 * (x * x + z * z); R0[5] = 2 * (y * z - w * x);
    R0[6] = 2 * (x * z - w * y); R0[7] = 2 * (y * z + w * x); R0[8] = 1 - 2 * (x * x + y * y);
    v0[0] = om[0]; v0[1] = om[1]; v0[2] = om[2];
#pragma unroll
    for (int m = 0; m < 3; ++m) v0[3 + m] = R0[m] * vw[0] + R0[3 + m] * vw[1] + R0[6 + m] * vw[2];
    if (wave == 0) {
      leg_forces<0>(a, N, e, live, R0, p0, v0, Fb);
      servo_store<12, 9>(a, N, e, live);            // waist and left arm
    } else {
      leg_forces<1>(a, N, e, live, R0, p0, v0, Fb);
      servo_store<21, 8>(a, N, e, live);            // right arm
#pragma unroll
      for (int m = 0; m < 6; ++m) fx[m][lane] = Fb[m];
    }
  }
  __syncthreads();
  if (wave == 0 && live) {
    float F[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) F[m] = Fb[m] + fx[m][lane];
    // base: world force R0 F.l, body torque F.a (convention of g1_c_fd's tau)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      a.tau[m * N + e] = R0[3 * m] * F[3] + R0[3 * m + 1] * F[4] + R0[3 * m + 2] * F[5];
      a.tau[(3 + m) * N + e] = F[m];
    }
  }
}

__global__ __launch_bounds__(128) void forces_kernel(ForcesArgs a) {
  __shared__ float fx[6][ENVS_PER_WG];
  forces_block(a, __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), fx);
}

G1K_DEV ForcesArgs forces_args(const StepArgs& s) {
  ForcesArgs a;
  a.nenv = s.nenv; a.cfg = s.cfg;
  a.qpos = s.qpos; a.qvel = s.qvel; a.anchor = s.anchor; a.ctrl = s.ctrl; a.mu = s.mu; a.tau = s.tau;
  a.contact_forces = s.contact_forces; a.foot_pos = s.foot_pos; a.foot_vel = s.foot_vel;
  return a;
}

template <bool F64_BASE, bool F64_CONTACT, int SPLIT, int MODE>
__global__ __launch_bounds__(64 * SPLIT) void step_kernel(StepArgs args) {
  __shared__ float xch[2][SPLIT][XCH][ENVS_PER_WG];
  __shared__ float fcx[2][2][6][ENVS_PER_WG];   // legs' contact wrenches, pelvis frame (4-wave split)
  __shared__ float wsh[W_IN_LDS<SPLIT> ? 4 * MAXJ * 6 * ENVS_PER_WG : 1];
  bool commit = false, record = false;
  if constexpr (MODE == MODE_JUDGE_FORCES) {
    if (args.blk_run && args.blk_run[blockIdx.x]) return;     // stepped by this period's fallback run
    const int e_raw = blockIdx.x * ENVS_PER_WG + (threadIdx.x & 63);
    const bool live = e_raw < args.nenv;
    const bool blk = (int)(blockIdx.x % (unsigned)args.sel_period) == args.sel_phase;
    const bool fb = live && args.sel_fallback && args.sel_fallback[e_raw];
    const bool in = live && (blk || fb);
    if (__ballot(in) == 0ull) {        // uniform: every wave reads the same 64 flags
      forces_block(forces_args(args), __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), fcx[0][0]);
      return;
    }
    // every live env of the block in fallback: commit here (uniform)
    commit = args.commit_fallback && args.sel_fallback && __ballot(live && !fb) == 0ull;
    record = commit && blk && args.snap;
  } else if constexpr (MODE == MODE_FALLBACK_RUN) {
    const int e_raw = blockIdx.x * ENVS_PER_WG + (threadIdx.x & 63);
    const bool live = e_raw < args.nenv;
    const bool fb = live && args.sel_fallback[e_raw];
    const bool run = __ballot(live && !fb) == 0ull;     // uniform: every live env in fallback
    if (threadIdx.x == 0) args.blk_run[blockIdx.x] = run ? 1 : 0;
    if (!run) return;
  }
  if constexpr (SPLIT == 4) {
    step_body<F64_BASE, F64_CONTACT, SPLIT, -1, MODE>(args, xch, fcx, wsh, commit, record);
  } else {
    if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0)
      step_body<F64_BASE, F64_CONTACT, SPLIT, 0, MODE>(args, xch, fcx, wsh, commit, record);
    else
      step_body<F64_BASE, F64_CONTACT, SPLIT, 1, MODE>(args, xch, fcx, wsh, commit, record);
  }
}

// The envs whose qacc ABA computes this step: the envs of every block of 64
// with block % period == phase (the rotating judge, judged_env) plus every
// env in fallback.
// One atomic per workgroup: each wave counts its selected envs (ballot), the
// workgroup takes one block of slots, and every env's slot is its rank in
// the workgroup, so the list is sorted within each workgroup's run and the
// judge step's loads stay coalesced.  (One atomic per env serialises on the
// counter: 65,536 fallback envs took 13 us.)
__global__ __launch_bounds__(256) void judge_list_kernel(int nenv, int phase, int period, const uint8_t* fallback,
                                                          int* list, int* count) {
  __shared__ int wave_count[4];
  __shared__ int block_base;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool sel = e < nenv && (judged_env(e, phase, period) || (fallback && fallback[e]));
  const unsigned long long m = __ballot(sel);
  if (lane == 0) wave_count[w] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int total = wave_count[0] + wave_count[1] + wave_count[2] + wave_count[3];
    block_base = total ? atomicAdd(count, total) : 0;
  }
  __syncthreads();
  int base = block_base;
  for (int v = 0; v < w; ++v) base += wave_count[v];
  if (sel) list[base + __popcll(m & ((1ull << lane) - 1ull))] = e;
}

__global__ void zero_int_kernel(int* x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

// Applies qacc and integrates (semi-implicit Euler as the step kernel):
// envs in the judge list take ABA's qacc; judged envs compare the surrogate's
// qacc with ABA's (rel. error |a - b| / max(|b|, 1), max over dofs) and enter
// or leave fallback on `tol`; all other envs take the surrogate's qacc
// (pred is the network's z-scored output, row e of [nenv][ldp]).
// stats[0] judged, stats[1] rejected, stats[2] max judged error (float bits),
// stats[3] envs in fallback that used ABA.
struct IntegrateArgs {
  int nenv, phase, period;
  float dt, tol;
  float* qpos;
  float* qvel;
  const float* pred;
  int ldp;
  const float* ymean;
  const float* ystd;
  const float* qacc_aba;
  uint8_t* fallback;
  unsigned* stats;
  float* qacc_out;
};

__global__ __launch_bounds__(256) void surrogate_integrate_kernel(IntegrateArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = e < a.nenv;
  const size_t N = (size_t)a.nenv;
  float sur[NV];
  if (live) {
#pragma unroll
    for (int j = 0; j < NV; ++j) sur[j] = a.pred[(size_t)e * a.ldp + j] * a.ystd[j] + a.ymean[j];
  }
  surrogate_commit(live, e, N, a.phase, a.period, a.tol, a.dt, sur, a.qacc_aba, a.fallback, a.stats, a.qacc_out,
                   a.qpos, a.qvel);
}

// ------------------------------------------------------------------ reset
// src/g1_kernel.cu:54-62: integer hash and uniform draw.
G1K_DEV unsigned hash32(unsigned x) {
  x ^= x >> 16; x *= 0x7feb352du;
  x ^= x >> 15; x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// src/g1_kernel.cu:85-103.  The noise term is one fused multiply-add,
// v = fma(noise, 2u - 1, v) (2u - 1 is exact), which is what a CUDA build of
// the reference source produces by default (contraction of a*b + c); the
// oracle evaluates the same expression, so the reset is bit-exact.
__global__ __launch_bounds__(256) void reset_done_kernel(int nenv, float* qpos, float* qvel, float* anchor,
                                                         const uint8_t* done, unsigned seed, float noise,
                                                         float drop) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nenv || !done[e]) return;
  const size_t N = (size_t)nenv;
  unsigned s = hash32(seed ^ (unsigned)(e + 1));
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    float v = (float)g1m_qpos_stand[i];
    if (i == 2) v = __fadd_rn(v, drop);
    if (i >= 7) {
      s = hash32(s + 0x9E3779B9u);
      const float u = (float)s * 2.3283064365386963e-10f;   // exact scaling by 2^-32
      const float r = 2.0f * u - 1.0f;                          // exact for u in [0, 1]
      v = __builtin_fmaf(noise, r, v);
    }
    qpos[i * N + e] = v;
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) qvel[i * N + e] = 0.f;
#pragma unroll
  for (int i = 0; i < 2 * NC; ++i) anchor[i * N + e] = ANCHOR_FREE;
}

}  // namespace g1k

// ================================================================== C-ABI
namespace {
thread_local char g_err[512] = "";

void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace

namespace g1err {
void clear() { g_err[0] = 0; }
int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return -1;
}
int launch(const char* what) {
  const hipError_t err = hipGetLastError();
  if (err == hipSuccess) return 0;
  snprintf(g_err, sizeof(g_err), "%s launch failed: %s", what, hipGetErrorString(err));
  return -2;
}
}  // namespace g1err

namespace {

using KernelFn = void (*)(g1k::StepArgs);

// Work split: 4 waves (one limb chain each) fill the chip from 16,384 envs
// (1,024 waves, one per SIMD); from 32,768 envs 2 waves (legs, arms) per 64
// envs still give 1,024+ waves and repeat the pelvis/waist part half as often
// (measured on MI355X at 32,768 envs: 2.03e9 vs 1.65e9 env-steps/s).
constexpr int kSplit2MinEnvs = 32768;

template <int SPLIT>
KernelFn pick(bool f64b, bool f64c, int mode) {
  using namespace g1k;
  if (mode == MODE_JUDGE_FORCES)
    return f64c ? step_kernel<false, true, SPLIT, MODE_JUDGE_FORCES>
                : step_kernel<false, false, SPLIT, MODE_JUDGE_FORCES>;
  if (mode == MODE_FALLBACK_RUN)
    return f64c ? step_kernel<false, true, SPLIT, MODE_FALLBACK_RUN>
                : step_kernel<false, false, SPLIT, MODE_FALLBACK_RUN>;
  if (mode == MODE_JUDGE_BLOCKS)
    return f64c ? step_kernel<false, true, SPLIT, MODE_JUDGE_BLOCKS>
                : step_kernel<false, false, SPLIT, MODE_JUDGE_BLOCKS>;
  return f64b ? (f64c ? step_kernel<true, true, SPLIT, MODE_STEP> : step_kernel<true, false, SPLIT, MODE_STEP>)
              : (f64c ? step_kernel<false, true, SPLIT, MODE_STEP> : step_kernel<false, false, SPLIT, MODE_STEP>);
}

int launch_step(g1k::StepArgs a, unsigned flags, hipStream_t stream, int mode = g1k::MODE_STEP) {
  const bool f64b = flags & G1_FLAG_F64_BASE, f64c = flags & G1_FLAG_F64_CONTACT;
  a.f64_solve = (flags & G1_FLAG_F64_SOLVE) ? 1 : 0;
  int split = a.nenv >= kSplit2MinEnvs ? 2 : 4;
  if (flags & G1_FLAG_WAVES4) split = 4;
  if (flags & G1_FLAG_WAVES2) split = 2;
  KernelFn k = split == 2 ? pick<2>(f64b, f64c, mode) : pick<4>(f64b, f64c, mode);
  const int nblocks = (a.nenv + g1k::ENVS_PER_WG - 1) / g1k::ENVS_PER_WG;
  hipLaunchKernelGGL(k, dim3(nblocks), dim3(64 * split), 0, stream, a);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_err("step launch failed: %s", hipGetErrorString(err));
    return -2;
  }
  return 0;
}

// flags of the surrogate judge and fallback-run launches: the float64
// islands of the solve and the contacts, and at most one wave split
bool bad_judge_flags(unsigned flags) {
  return (flags & ~(G1_FLAG_F64_SOLVE | G1_FLAG_F64_CONTACT | G1_FLAG_WAVES4 | G1_FLAG_WAVES2)) ||
         ((flags & G1_FLAG_WAVES4) && (flags & G1_FLAG_WAVES2));
}

int validate(int nenv, int nsub, float dt, const void* qpos, const void* qvel, const void* anchor,
             const void* ctrl) {
  if (nenv < 0 || nsub < 0) { set_err("nenv (%d) and nsub (%d) must be >= 0", nenv, nsub); return -1; }
  if (!(dt > 0.f) || !isfinite(dt)) { set_err("dt must be finite and > 0 (got %g)", (double)dt); return -1; }
  if (nenv > 0 && (!qpos || !qvel || !anchor || !ctrl)) {
    set_err("qpos, qvel, anchor and ctrl are required");
    return -1;
  }
  return 0;
}
}  // namespace

// g1_hip_judge_forces, and a block whose live envs are all in fallback is
// also committed (integrated and stored as by the plain step) with its
// deferred-judgement record written if it is judged this step
// (StepArgs::commit_fallback).  blk_run (nullable): the blocks a fallback run
// has already stepped, left alone.
namespace g1k_host {
int judge_forces_commit(int nenv, float dt, float* qpos, float* qvel, float* anchor, const float* ctrl,
                        const float* mu, const uint8_t* fallback, int phase, int period, float* qacc_out,
                        float* tau_out, float* contact_forces, float* foot_pos, float* foot_vel, float* qacc_track,
                        float* snap, uint8_t* sflag, unsigned* stats, const uint8_t* blk_run, unsigned flags,
                        void* stream) {
  g_err[0] = 0;
  int rc = validate(nenv, 1, dt, qpos, qvel, anchor, ctrl);
  if (rc) return rc;
  if (period <= 0 || phase < 0 || phase >= period || !qacc_out || !tau_out || !fallback || !snap || !sflag ||
      !stats) {
    set_err("judge_forces_commit: bad phase/period, or no qacc / tau / fallback / records / stats");
    return -1;
  }
  if (bad_judge_flags(flags)) {
    set_err("judge_forces_commit: unsupported flags 0x%x", flags);
    return -1;
  }
  if (nenv == 0) return 0;
  g1k::StepArgs a{};
  a.nenv = nenv; a.nsub = 1;
  a.cfg = g1k::default_cfg(dt);
  a.qpos = qpos; a.qvel = qvel; a.anchor = anchor;
  a.ctrl = ctrl;
  a.foot_pos = foot_pos; a.foot_vel = foot_vel; a.contact_forces = contact_forces;
  a.qacc = qacc_out; a.tau = tau_out; a.mu = mu;
  a.stride = nenv;
  a.sel_fallback = fallback; a.sel_phase = phase; a.sel_period = period;
  a.commit_fallback = 1; a.snap = snap; a.sflag = sflag; a.qacc_track = qacc_track; a.stats = stats;
  a.blk_run = const_cast<uint8_t*>(blk_run);
  return launch_step(a, flags, (hipStream_t)stream, g1k::MODE_JUDGE_FORCES);
}
// g1_hip_fallback_run with a PD-target schedule (ctrl_stride > 0): substep s
// of the run uses the block ctrl + ((ctrl_phase + s) / ctrl_period)
// ctrl_stride, ctrl being the block of substep 0 and ctrl_phase < ctrl_period
// the substeps of that block taken before the run.
int fallback_run(int nenv, int nsub, float dt, float* qpos, float* qvel, float* anchor, const float* ctrl,
                 long long ctrl_stride, int ctrl_period, int ctrl_phase, const float* mu, const uint8_t* fallback,
                 int phase0, int period, uint8_t* blk_run, float* qacc_track, float* contact_forces, float* foot_pos,
                 float* foot_vel, float* snap, uint8_t* sflag, unsigned* stats, unsigned flags, void* stream) {
  g_err[0] = 0;
  int rc = validate(nenv, nsub, dt, qpos, qvel, anchor, ctrl);
  if (rc) return rc;
  if (period <= 0 || phase0 < 0 || nsub > period - phase0 || !fallback || !blk_run || !snap || !sflag || !stats) {
    set_err("fallback_run: bad phase/period (the run may not cross a period end), or no fallback / blk_run / "
            "records / stats");
    return -1;
  }
  if (ctrl_stride < 0 || (ctrl_stride > 0 && (ctrl_stride < 29LL * nenv || ctrl_period <= 0 || ctrl_phase < 0 ||
                                             ctrl_phase >= ctrl_period))) {
    set_err("fallback_run: bad target schedule");
    return -1;
  }
  if (bad_judge_flags(flags)) {
    set_err("fallback_run: unsupported flags 0x%x", flags);
    return -1;
  }
  if (nenv == 0 || nsub == 0) return 0;
  g1k::StepArgs a{};
  a.nenv = nenv; a.nsub = nsub;
  a.cfg = g1k::default_cfg(dt);
  a.qpos = qpos; a.qvel = qvel; a.anchor = anchor;
  a.ctrl = ctrl;
  a.ctrl_stride = ctrl_stride; a.ctrl_period = ctrl_period; a.ctrl_phase = ctrl_phase;
  a.foot_pos = foot_pos; a.foot_vel = foot_vel; a.contact_forces = contact_forces;
  a.mu = mu;
  a.stride = nenv;
  a.sel_fallback = fallback; a.sel_phase = phase0; a.sel_period = period;
  a.snap = snap; a.sflag = sflag; a.qacc_track = qacc_track; a.stats = stats;
  a.blk_run = blk_run; a.run_phase0 = phase0;
  return launch_step(a, flags, (hipStream_t)stream, g1k::MODE_FALLBACK_RUN);
}
}  // namespace g1k_host

extern "C" {

int g1_hip_version(void) { return 100; }

const char* g1_hip_last_error(void) { return g_err; }

int g1_hip_step_ex(int nenv, int nsub, float dt, float* qpos, float* qvel, float* anchor, const float* ctrl,
                   float* foot_pos, float* foot_vel, float* contact_forces, float* qacc_out, float* tau_out,
                   const float* mu, double* base_pose64, unsigned flags, void* stream) {
  g_err[0] = 0;
  int rc = validate(nenv, nsub, dt, qpos, qvel, anchor, ctrl);
  if (rc) return rc;
  if (flags & ~(G1_FLAG_F64_SOLVE | G1_FLAG_F64_BASE | G1_FLAG_F64_CONTACT | G1_FLAG_WAVES4 | G1_FLAG_WAVES2) ||
      ((flags & G1_FLAG_WAVES4) && (flags & G1_FLAG_WAVES2))) { set_err("unknown flag bits 0x%x", flags); return -1; }
  if (((flags & G1_FLAG_F64_BASE) != 0) != (base_pose64 != nullptr)) {
    set_err("base_pose64 must be given exactly when G1_FLAG_F64_BASE is set");
    return -1;
  }
  if (nenv == 0 || nsub == 0) return 0;
  g1k::StepArgs a{};
  a.nenv = nenv; a.nsub = nsub;
  a.cfg = g1k::default_cfg(dt);
  a.qpos = qpos; a.qvel = qvel; a.anchor = anchor; a.ctrl = ctrl;
  a.foot_pos = foot_pos; a.foot_vel = foot_vel; a.contact_forces = contact_forces;
  a.qacc = qacc_out; a.tau = tau_out; a.mu = mu; a.base64 = base_pose64;
  a.f64_solve = 0;
  a.env_list = nullptr; a.list_count = nullptr; a.stride = nenv;
  return launch_step(a, flags, (hipStream_t)stream);
}

// Substeps per launch of g1_hip_step_schedule (a launch covers 4 ms of
// simulated time per 1,000 substeps at dt 2e-3; 16,384 envs take ~9 ms).
constexpr int kScheduleMaxSub = 1000;

int g1_hip_step_schedule(int nenv, int nsteps, int period, int step0, float dt, float* qpos, float* qvel,
                         float* anchor, const float* ctrl_blocks, long long block_stride, float* foot_pos,
                         float* foot_vel, float* contact_forces, const float* mu, double* base_pose64,
                         unsigned flags, void* stream) {
  g_err[0] = 0;
  if (nsteps < 0 || period <= 0 || step0 < 0 || block_stride < 29LL * nenv) {
    set_err("step_schedule: nsteps (%d) >= 0, period (%d) > 0, step0 (%d) >= 0 and block_stride (%lld) >= 29 nenv "
            "required", nsteps, period, step0, block_stride);
    return -1;
  }
  int rc = validate(nenv, 1, dt, qpos, qvel, anchor, ctrl_blocks);
  if (rc) return rc;
  if (flags & ~(G1_FLAG_F64_SOLVE | G1_FLAG_F64_BASE | G1_FLAG_F64_CONTACT | G1_FLAG_WAVES4 | G1_FLAG_WAVES2) ||
      ((flags & G1_FLAG_WAVES4) && (flags & G1_FLAG_WAVES2))) { set_err("unknown flag bits 0x%x", flags); return -1; }
  if (((flags & G1_FLAG_F64_BASE) != 0) != (base_pose64 != nullptr)) {
    set_err("base_pose64 must be given exactly when G1_FLAG_F64_BASE is set");
    return -1;
  }
  if (nenv == 0 || nsteps == 0) return 0;
  g1k::StepArgs a{};
  a.nenv = nenv;
  a.cfg = g1k::default_cfg(dt);
  a.qpos = qpos; a.qvel = qvel; a.anchor = anchor;
  a.foot_pos = foot_pos; a.foot_vel = foot_vel; a.contact_forces = contact_forces;
  a.qacc = nullptr; a.tau = nullptr; a.mu = mu; a.base64 = base_pose64;
  a.env_list = nullptr; a.list_count = nullptr; a.stride = nenv;
  // launches of up to kScheduleMaxSub substeps, each switching target blocks
  // itself: step step0 + s uses block (step0 + s) / period
  a.ctrl_stride = block_stride;
  a.ctrl_period = period;
  for (int s = 0; s < nsteps;) {
    const int abs_step = step0 + s;
    const int n = std::min(kScheduleMaxSub, nsteps - s);
    a.nsub = n;
    a.ctrl = ctrl_blocks + (size_t)(abs_step / period) * (size_t)block_stride;
    a.ctrl_phase = abs_step % period;
    rc = launch_step(a, flags, (hipStream_t)stream);
    if (rc) return rc;
    s += n;
  }
  return 0;
}

// src/g1_kernel.cu:154-163: dt = 1e-3 (g1_default_config), model friction.
void g1_cuda_step(int nenv, int nsub, float* qpos, float* qvel, float* anchor, const float* ctrl,
                  float* foot_pos, float* foot_vel, float* contact_forces, void* stream) {
  g1_hip_step_ex(nenv, nsub, 1e-3f, qpos, qvel, anchor, ctrl, foot_pos, foot_vel, contact_forces, nullptr,
                 nullptr, nullptr, nullptr, 0u, stream);
}

int g1_hip_forces(int nenv, float dt, const float* qpos, const float* qvel, float* anchor, const float* ctrl,
                  const float* mu, float* tau_out, float* contact_forces, float* foot_pos, float* foot_vel,
                  void* stream) {
  g_err[0] = 0;
  if (nenv < 0 || !(dt > 0.f)) { set_err("forces: bad nenv/dt"); return -1; }
  if (nenv == 0) return 0;
  if (!qpos || !qvel || !anchor || !ctrl || !tau_out) { set_err("forces: qpos, qvel, anchor, ctrl, tau required"); return -1; }
  g1k::ForcesArgs a;
  a.nenv = nenv; a.cfg = g1k::default_cfg(dt);
  a.qpos = qpos; a.qvel = qvel; a.anchor = anchor; a.ctrl = ctrl; a.mu = mu; a.tau = tau_out;
  a.contact_forces = contact_forces; a.foot_pos = foot_pos; a.foot_vel = foot_vel;
  hipLaunchKernelGGL(g1k::forces_kernel, dim3((nenv + g1k::ENVS_PER_WG - 1) / g1k::ENVS_PER_WG), dim3(128), 0,
                     (hipStream_t)stream, a);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) { set_err("forces launch failed: %s", hipGetErrorString(err)); return -2; }
  return 0;
}

int g1_hip_judge_list(int nenv, int phase, int period, const uint8_t* fallback, int* list, int* count,
                      void* stream) {
  g_err[0] = 0;
  if (nenv < 0 || period <= 0 || phase < 0 || phase >= period || !list || !count) {
    set_err("judge_list: bad arguments");
    return -1;
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(g1k::zero_int_kernel, dim3(1), dim3(64), 0, s, count, 1);
  if (nenv > 0)
    hipLaunchKernelGGL(g1k::judge_list_kernel, dim3((nenv + 255) / 256), dim3(256), 0, s, nenv, phase, period,
                       fallback, list, count);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) { set_err("judge_list launch failed: %s", hipGetErrorString(err)); return -2; }
  return 0;
}

int g1_hip_step_judge(int nenv, float dt, const float* qpos, const float* qvel, const float* anchor,
                      const float* ctrl, const float* mu, const int* list, const int* count, float* qacc_out,
                      unsigned flags, void* stream) {
  g_err[0] = 0;
  int rc = validate(nenv, 1, dt, qpos, qvel, anchor, ctrl);
  if (rc) return rc;
  if (!list || !count || !qacc_out) { set_err("step_judge: list, count and qacc are required"); return -1; }
  if (bad_judge_flags(flags)) {
    set_err("step_judge: unsupported flags 0x%x", flags);
    return -1;
  }
  if (nenv == 0) return 0;
  g1k::StepArgs a{};
  a.nenv = nenv; a.nsub = 1;
  a.cfg = g1k::default_cfg(dt);
  a.qpos = const_cast<float*>(qpos); a.qvel = const_cast<float*>(qvel); a.anchor = const_cast<float*>(anchor);
  a.ctrl = ctrl;
  a.foot_pos = nullptr; a.foot_vel = nullptr; a.contact_forces = nullptr;
  a.qacc = qacc_out; a.tau = nullptr; a.mu = mu; a.base64 = nullptr;
  a.f64_solve = 0;
  a.env_list = list; a.list_count = count; a.stride = nenv;
  return launch_step(a, flags, (hipStream_t)stream);
}

int g1_hip_step_judge_blocks(int nenv, float dt, const float* qpos, const float* qvel, const float* anchor,
                             const float* ctrl, const float* mu, const uint8_t* fallback, int phase, int period,
                             float* qacc_out, unsigned flags, void* stream) {
  g_err[0] = 0;
  int rc = validate(nenv, 1, dt, qpos, qvel, anchor, ctrl);
  if (rc) return rc;
  if (period <= 0 || phase < 0 || phase >= period || !qacc_out) {
    set_err("step_judge_blocks: bad phase/period or no qacc");
    return -1;
  }
  if (bad_judge_flags(flags)) {
    set_err("step_judge_blocks: unsupported flags 0x%x", flags);
    return -1;
  }
  if (nenv == 0) return 0;
  g1k::StepArgs a{};
  a.nenv = nenv; a.nsub = 1;
  a.cfg = g1k::default_cfg(dt);
  a.qpos = const_cast<float*>(qpos); a.qvel = const_cast<float*>(qvel); a.anchor = const_cast<float*>(anchor);
  a.ctrl = ctrl;
  a.qacc = qacc_out; a.mu = mu;
  a.stride = nenv;
  a.sel_fallback = fallback; a.sel_phase = phase; a.sel_period = period;
  return launch_step(a, flags, (hipStream_t)stream, g1k::MODE_JUDGE_BLOCKS);
}

int g1_hip_judge_forces(int nenv, float dt, const float* qpos, const float* qvel, float* anchor, const float* ctrl,
                        const float* mu, const uint8_t* fallback, int phase, int period, float* qacc_out,
                        float* tau_out, float* contact_forces, float* foot_pos, float* foot_vel, unsigned flags,
                        void* stream) {
  g_err[0] = 0;
  int rc = validate(nenv, 1, dt, qpos, qvel, anchor, ctrl);
  if (rc) return rc;
  if (period <= 0 || phase < 0 || phase >= period || !qacc_out || !tau_out) {
    set_err("judge_forces: bad phase/period, or no qacc / tau");
    return -1;
  }
  if (bad_judge_flags(flags)) {
    set_err("judge_forces: unsupported flags 0x%x", flags);
    return -1;
  }
  if (nenv == 0) return 0;
  g1k::StepArgs a{};
  a.nenv = nenv; a.nsub = 1;
  a.cfg = g1k::default_cfg(dt);
  a.qpos = const_cast<float*>(qpos); a.qvel = const_cast<float*>(qvel); a.anchor = anchor;
  a.ctrl = ctrl;
  a.foot_pos = foot_pos; a.foot_vel = foot_vel; a.contact_forces = contact_forces;
  a.qacc = qacc_out; a.tau = tau_out; a.mu = mu;
  a.stride = nenv;
  a.sel_fallback = fallback; a.sel_phase = phase; a.sel_period = period;
  return launch_step(a, flags, (hipStream_t)stream, g1k::MODE_JUDGE_FORCES);
}

// g1k_host::judge_forces_commit without a fallback run.
int g1_hip_judge_forces_commit(int nenv, float dt, float* qpos, float* qvel, float* anchor, const float* ctrl,
                               const float* mu, const uint8_t* fallback, int phase, int period, float* qacc_out,
                               float* tau_out, float* contact_forces, float* foot_pos, float* foot_vel,
                               float* qacc_track, float* snap, uint8_t* sflag, unsigned* stats, unsigned flags,
                               void* stream) {
  return g1k_host::judge_forces_commit(nenv, dt, qpos, qvel, anchor, ctrl, mu, fallback, phase, period, qacc_out,
                                       tau_out, contact_forces, foot_pos, foot_vel, qacc_track, snap, sflag, stats,
                                       nullptr, flags, stream);
}

// The blocks whose live envs are all in fallback, nsub steps as the plain
// step (MODE_FALLBACK_RUN), each writing its deferred-judgement record at
// its judged step; blk_run[b] = 1 for those blocks, 0 for the others.
int g1_hip_fallback_run(int nenv, int nsub, float dt, float* qpos, float* qvel, float* anchor, const float* ctrl,
                        const float* mu, const uint8_t* fallback, int phase0, int period, uint8_t* blk_run,
                        float* qacc_track, float* contact_forces, float* foot_pos, float* foot_vel, float* snap,
                        uint8_t* sflag, unsigned* stats, unsigned flags, void* stream) {
  return g1k_host::fallback_run(nenv, nsub, dt, qpos, qvel, anchor, ctrl, 0, 0, 0, mu, fallback, phase0, period,
                                blk_run, qacc_track, contact_forces, foot_pos, foot_vel, snap, sflag, stats, flags,
                                stream);
}

int g1_hip_surrogate_integrate(int nenv, float dt, int phase, int period, float tol, float* qpos, float* qvel,
                               const float* pred, int ldp, const float* ymean, const float* ystd,
                               const float* qacc_aba, uint8_t* fallback, unsigned* stats, float* qacc_out,
                               void* stream) {
  g_err[0] = 0;
  if (nenv < 0 || !(dt > 0.f) || period <= 0 || phase < 0 || phase >= period || ldp < 35) {
    set_err("surrogate_integrate: bad arguments");
    return -1;
  }
  if (nenv == 0) return 0;
  if (!qpos || !qvel || !pred || !ymean || !ystd || !qacc_aba || !stats) {
    set_err("surrogate_integrate: null pointer");
    return -1;
  }
  g1k::IntegrateArgs a{nenv, phase, period, dt, tol, qpos, qvel, pred, ldp, ymean, ystd, qacc_aba, fallback,
                       stats, qacc_out};
  hipLaunchKernelGGL(g1k::surrogate_integrate_kernel, dim3((nenv + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     a);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) { set_err("surrogate_integrate launch failed: %s", hipGetErrorString(err)); return -2; }
  return 0;
}

void g1_cuda_reset_done(int nenv, float* qpos, float* qvel, float* anchor, const uint8_t* done, unsigned seed,
                        float noise, float drop, void* stream) {
  g_err[0] = 0;
  if (nenv <= 0) return;
  if (!qpos || !qvel || !anchor || !done) { set_err("reset: qpos, qvel, anchor and done are required"); return; }
  const int tpb = 256, nb = (nenv + tpb - 1) / tpb;
  hipLaunchKernelGGL(g1k::reset_done_kernel, dim3(nb), dim3(tpb), 0, (hipStream_t)stream, nenv, qpos, qvel,
                     anchor, done, seed, noise, drop);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) set_err("reset launch failed: %s", hipGetErrorString(err));
}

}  // extern "C"
