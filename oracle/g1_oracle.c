/* g1_oracle.c -- FP64 CPU oracle for the G1 fused step (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from this file, and only as the checker or as the
 * timed CPU baseline; the product never does.
 *
 * A plain-C restatement of the reference algorithm: spatial algebra of
 * src/spatial.h, the three stages of src/g1_core.h (kinematics, fused
 * contacts/PD/ABA, semi-implicit Euler), the host C-ABI of src/host_lib.cpp
 * and the reset RNG of src/g1_kernel.cu.  Each function names the lines it
 * follows.  Unlike src/host_lib.cpp it keeps no static workspace, so it is
 * reentrant and the batched entry points run it on OpenMP threads.
 *
 * `real` is double unless built with -DG1O_REAL=float (the float build is the
 * CPU-baseline "port": the same algorithm at the GPU's precision).
 */
#include "g1_oracle.h"
#include "g1_model_data.h"

#include <math.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef G1O_REAL
#define G1O_REAL double
#endif
typedef G1O_REAL real;

#define NB G1M_NB
#define NJ (G1M_NB - 1)
#define NQ G1M_NQ
#define NV G1M_NV
#define NU G1M_NU
#define NC G1M_NC

/* ------------------------------------------------ src/g1_types.h:33-55 */
typedef struct {
  real gravity_z, kn, dn, kt, dt_t, limit_kp, limit_kd, dt, mu;
} cfg_t;

static cfg_t default_cfg(void) {
  cfg_t c;
  c.gravity_z = (real)G1M_GRAVITY_Z;
  c.kn = (real)2e4;   c.dn = (real)280;
  c.kt = (real)1e4;   c.dt_t = (real)280;
  c.limit_kp = (real)200; c.limit_kd = (real)5;
  c.dt = (real)1e-3;
  c.mu = (real)G1M_CONTACT_MU;
  return c;
}

/* ------------------------------------------------ src/spatial.h:11-118 */
typedef struct { real x, y, z; } v3;
typedef struct { real m[9]; } m3;           /* row-major */
typedef struct { real w, x, y, z; } quat;
typedef struct { m3 R; v3 p; } se3;        /* pose of child in parent */
typedef struct { v3 a, l; } sv;             /* [angular; linear] */
typedef struct { real m[21]; } m6;          /* packed lower triangle */

static v3 V(real x, real y, real z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 vscl(real s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
static real vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 vcross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static v3 vld(const double* p) { return V((real)p[0], (real)p[1], (real)p[2]); }
static v3 vldr(const real* p) { return V(p[0], p[1], p[2]); }

static v3 mv(const m3* A, v3 v) {                 /* A v */
  return V(A->m[0] * v.x + A->m[1] * v.y + A->m[2] * v.z,
           A->m[3] * v.x + A->m[4] * v.y + A->m[5] * v.z,
           A->m[6] * v.x + A->m[7] * v.y + A->m[8] * v.z);
}
static v3 mtv(const m3* A, v3 v) {                /* A^T v */
  return V(A->m[0] * v.x + A->m[3] * v.y + A->m[6] * v.z,
           A->m[1] * v.x + A->m[4] * v.y + A->m[7] * v.z,
           A->m[2] * v.x + A->m[5] * v.y + A->m[8] * v.z);
}
static m3 mm(const m3* A, const m3* B) {          /* A B */
  m3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C.m[3 * i + j] = A->m[3 * i] * B->m[j] + A->m[3 * i + 1] * B->m[3 + j] +
                       A->m[3 * i + 2] * B->m[6 + j];
  return C;
}
static m3 mmt(const m3* A, const m3* B) {         /* A B^T */
  m3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C.m[3 * i + j] = A->m[3 * i] * B->m[3 * j] + A->m[3 * i + 1] * B->m[3 * j + 1] +
                       A->m[3 * i + 2] * B->m[3 * j + 2];
  return C;
}

/* src/spatial.h:42-47 */
static quat qnormalize(quat q) {
  real n = sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  real s = (real)1 / n;
  quat r = {q.w * s, q.x * s, q.y * s, q.z * s};
  return r;
}
/* src/spatial.h:48-54 (Hamilton product, w first) */
static quat qmult(quat a, quat b) {
  quat r = {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
            a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
  return r;
}
/* src/spatial.h:55-62 */
static m3 qmat(quat q) {
  real w = q.w, x = q.x, y = q.y, z = q.z;
  m3 R;
  R.m[0] = 1 - 2 * (y * y + z * z); R.m[1] = 2 * (x * y - w * z); R.m[2] = 2 * (x * z + w * y);
  R.m[3] = 2 * (x * y + w * z); R.m[4] = 1 - 2 * (x * x + z * z); R.m[5] = 2 * (y * z - w * x);
  R.m[6] = 2 * (x * z - w * y); R.m[7] = 2 * (y * z + w * x); R.m[8] = 1 - 2 * (x * x + y * y);
  return R;
}
/* src/spatial.h:63-70: Rodrigues for a unit axis */
static m3 rodrigues(v3 a, real q) {
  real c = cos(q), s = sin(q), t = 1 - c;
  m3 R;
  R.m[0] = c + t * a.x * a.x;       R.m[1] = t * a.x * a.y - s * a.z; R.m[2] = t * a.x * a.z + s * a.y;
  R.m[3] = t * a.x * a.y + s * a.z; R.m[4] = c + t * a.y * a.y;       R.m[5] = t * a.y * a.z - s * a.x;
  R.m[6] = t * a.x * a.z - s * a.y; R.m[7] = t * a.y * a.z + s * a.x; R.m[8] = c + t * a.z * a.z;
  return R;
}
/* src/spatial.h:73-79: q <- q (x) exp(w dt / 2), with the 1e-12 cut-off */
static quat qintegrate(quat q, v3 w, real dt) {
  real angle = sqrt(vdot(w, w)) * dt;
  if (angle < (real)1e-12) return q;
  real s = sin(angle / 2) / (angle / dt);
  quat dq = {(real)cos(angle / 2), s * w.x, s * w.y, s * w.z};
  return qnormalize(qmult(q, dq));
}

static sv S6(v3 a, v3 l) { sv r; r.a = a; r.l = l; return r; }
static sv sadd(sv u, sv w) { return S6(vadd(u.a, w.a), vadd(u.l, w.l)); }
static sv ssub(sv u, sv w) { return S6(vsub(u.a, w.a), vsub(u.l, w.l)); }
static sv sscl(real s, sv u) { return S6(vscl(s, u.a), vscl(s, u.l)); }
static real sdot(sv u, sv w) { return vdot(u.a, w.a) + vdot(u.l, w.l); }
/* src/spatial.h:102-105: motion, parent -> child */
static sv motion_to_child(const se3* X, sv m) {
  return S6(mtv(&X->R, m.a), mtv(&X->R, vsub(m.l, vcross(X->p, m.a))));
}
/* src/spatial.h:107-110: force, child -> parent */
static sv force_to_parent(const se3* X, sv f) {
  v3 l = mv(&X->R, f.l);
  return S6(vadd(mv(&X->R, f.a), vcross(X->p, l)), l);
}
/* src/spatial.h:112-114 */
static sv crm(sv v, sv m) {
  return S6(vcross(v.a, m.a), vadd(vcross(v.a, m.l), vcross(v.l, m.a)));
}
/* src/spatial.h:116-118 */
static sv crf(sv v, sv f) {
  return S6(vadd(vcross(v.a, f.a), vcross(v.l, f.l)), vcross(v.a, f.l));
}

#define PK(i, j) ((i) * ((i) + 1) / 2 + (j))   /* packed index, i >= j */
static real m6get(const m6* A, int i, int j) { return i >= j ? A->m[PK(i, j)] : A->m[PK(j, i)]; }

/* src/spatial.h:125-138: symmetric 6x6 times spatial vector */
static sv m6mul(const m6* A, sv u) {
  real x[6] = {u.a.x, u.a.y, u.a.z, u.l.x, u.l.y, u.l.z}, y[6];
  for (int i = 0; i < 6; ++i) {
    real s = 0;
    for (int j = 0; j < 6; ++j) s += m6get(A, i, j) * x[j];
    y[i] = s;
  }
  return S6(V(y[0], y[1], y[2]), V(y[3], y[4], y[5]));
}
/* src/spatial.h:141-149: A -= U U^T / d */
static void m6_rank1_sub(m6* A, sv U, real inv_d) {
  real u[6] = {U.a.x, U.a.y, U.a.z, U.l.x, U.l.y, U.l.z};
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) A->m[PK(i, j)] -= u[i] * u[j] * inv_d;
}
/* src/spatial.h:152-172: rigid-body inertia about the body origin */
static m6 rb_inertia(real mass, v3 c, const double* Icom) {
  real cx[9] = {0, -c.z, c.y, c.z, 0, -c.x, -c.y, c.x, 0};
  m6 I;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      real val;
      if (i < 3) {
        real s = 0;
        for (int k = 0; k < 3; ++k) s += cx[3 * i + k] * cx[3 * k + j];
        val = (real)Icom[3 * i + j] - mass * s;
      } else if (j < 3) {
        val = -mass * cx[3 * (i - 3) + j];
      } else {
        val = (i == j) ? mass : 0;
      }
      I.m[PK(i, j)] = val;
    }
  return I;
}
/* src/spatial.h:219-251: parent += X* Ia X^-1 in block form
 * (J angular-angular, H = upper-right block, M linear-linear). */
static void m6_add_to_parent(m6* P, const m6* Ia, const se3* X) {
  m3 J, H, Mb;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      J.m[3 * i + j] = m6get(Ia, i, j);
      H.m[3 * i + j] = m6get(Ia, i, 3 + j);
      Mb.m[3 * i + j] = m6get(Ia, 3 + i, 3 + j);
    }
  m3 t;
  t = mm(&X->R, &J);  m3 tJ = mmt(&t, &X->R);
  t = mm(&X->R, &H);  m3 tH = mmt(&t, &X->R);
  t = mm(&X->R, &Mb); m3 tM = mmt(&t, &X->R);
  v3 p = X->p;
  m3 px = {{0, -p.z, p.y, p.z, 0, -p.x, -p.y, p.x, 0}};
  m3 pxM = mm(&px, &tM);
  m3 HA, tHT, JA;
  for (int k = 0; k < 9; ++k) HA.m[k] = tH.m[k] + pxM.m[k];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) tHT.m[3 * i + j] = tH.m[3 * j + i];
  m3 pxHT = mm(&px, &tHT);
  m3 HApx = mm(&HA, &px);
  for (int k = 0; k < 9; ++k) JA.m[k] = tJ.m[k] + pxHT.m[k] - HApx.m[k];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j <= i; ++j) {
      P->m[PK(i, j)] += JA.m[3 * i + j];
      P->m[PK(3 + i, 3 + j)] += tM.m[3 * i + j];
    }
  for (int i = 0; i < 3; ++i)          /* row 3+i, col j  = H^T(i, j) = HA(j, i) */
    for (int j = 0; j < 3; ++j) P->m[PK(3 + i, j)] += HA.m[3 * j + i];
}
/* src/spatial.h:254-278: in-place packed Cholesky, forward + back solve */
static void m6_cholesky_solve(const real* A, const real* b, real* y) {
  real L[21];
  memcpy(L, A, sizeof(L));
  for (int j = 0; j < 6; ++j) {
    for (int k = 0; k < j; ++k) L[PK(j, j)] -= L[PK(j, k)] * L[PK(j, k)];
    L[PK(j, j)] = sqrt(L[PK(j, j)]);
    real inv = (real)1 / L[PK(j, j)];
    for (int i = j + 1; i < 6; ++i) {
      for (int k = 0; k < j; ++k) L[PK(i, j)] -= L[PK(i, k)] * L[PK(j, k)];
      L[PK(i, j)] *= inv;
    }
  }
  for (int i = 0; i < 6; ++i) y[i] = b[i];
  for (int i = 0; i < 6; ++i) {
    for (int k = 0; k < i; ++k) y[i] -= L[PK(i, k)] * y[k];
    y[i] /= L[PK(i, i)];
  }
  for (int i = 5; i >= 0; --i) {
    for (int k = i + 1; k < 6; ++k) y[i] -= L[PK(k, i)] * y[k];
    y[i] /= L[PK(i, i)];
  }
}
static sv m6solve(const m6* A, sv b) {
  real bb[6] = {b.a.x, b.a.y, b.a.z, b.l.x, b.l.y, b.l.z}, y[6];
  m6_cholesky_solve(A->m, bb, y);
  return S6(V(y[0], y[1], y[2]), V(y[3], y[4], y[5]));
}

/* ---------------------------------------------- src/g1_core.h:13-23 */
typedef struct {
  se3 X[NB];          /* liMi: pose of body i in its parent */
  m3 oR[NB];
  v3 oP[NB];
  sv v[NB];
  m6 IA[NB];
  sv pA[NB];
  sv U[NJ];
  real D[NJ], u[NJ];
  sv a[NB];
} ws_t;

static sv hinge_S(int h) {            /* [axis; anchor x axis] */
  v3 ax = vld(g1m_jnt_axis + 3 * h), an = vld(g1m_jnt_anchor + 3 * h);
  return S6(ax, vcross(an, ax));
}

/* src/g1_core.h:29-54: kinematics and body twists */
static void fk_vel(const real* qpos, const real* qvel, ws_t* w) {
  quat bq = {qpos[3], qpos[4], qpos[5], qpos[6]};
  w->X[0].R = qmat(qnormalize(bq));
  w->X[0].p = vldr(qpos);
  w->oR[0] = w->X[0].R;
  w->oP[0] = w->X[0].p;
  w->v[0] = S6(vldr(qvel + 3), mtv(&w->oR[0], vldr(qvel)));
  for (int i = 1; i < NB; ++i) {
    int h = i - 1, par = g1m_parent[i];
    v3 ax = vld(g1m_jnt_axis + 3 * h), an = vld(g1m_jnt_anchor + 3 * h);
    quat tq = {(real)g1m_tree_quat[4 * i], (real)g1m_tree_quat[4 * i + 1],
               (real)g1m_tree_quat[4 * i + 2], (real)g1m_tree_quat[4 * i + 3]};
    m3 RT = qmat(qnormalize(tq));
    m3 RJ = rodrigues(ax, qpos[7 + h]);
    w->X[i].R = mm(&RT, &RJ);
    w->X[i].p = vadd(mv(&RT, vsub(an, mv(&RJ, an))), vld(g1m_tree_pos + 3 * i));
    w->oR[i] = mm(&w->oR[par], &w->X[i].R);
    w->oP[i] = vadd(mv(&w->oR[par], w->X[i].p), w->oP[par]);
    w->v[i] = sadd(motion_to_child(&w->X[i], w->v[par]), sscl(qvel[6 + h], hinge_S(h)));
  }
}

/* gravity + velocity-product bias of every body (src/g1_core.h:61-67, 113-120) */
static void init_bias(const cfg_t* c, ws_t* w) {
  for (int i = 0; i < NB; ++i) {
    v3 com = vld(g1m_com + 3 * i);
    real m = (real)g1m_mass[i];
    w->IA[i] = rb_inertia(m, com, g1m_inertia_com + 9 * i);
    v3 fg = vscl(m, mtv(&w->oR[i], V(0, 0, c->gravity_z)));
    w->pA[i] = ssub(crf(w->v[i], m6mul(&w->IA[i], w->v[i])), S6(vcross(com, fg), fg));
  }
}

/* one backward step of joint h (body i) into its parent, given tau and the
 * extra diagonal term (src/g1_core.h:72-88 and 156-186) */
static void backward_joint(int i, real qd, real tau, real extra_d, ws_t* w) {
  int h = i - 1, par = g1m_parent[i];
  sv Sh = hinge_S(h);
  w->U[h] = m6mul(&w->IA[i], Sh);
  w->D[h] = sdot(Sh, w->U[h]) + (real)g1m_armature[h] + extra_d;
  w->u[h] = tau - (real)g1m_damping[h] * qd - sdot(Sh, w->pA[i]);
  real invd = (real)1 / w->D[h];
  sv cb = crm(w->v[i], sscl(qd, Sh));
  m6 Ia = w->IA[i];
  m6_rank1_sub(&Ia, w->U[h], invd);
  sv pa = sadd(sadd(w->pA[i], m6mul(&Ia, cb)), sscl(w->u[h] * invd, w->U[h]));
  m6_add_to_parent(&w->IA[par], &Ia, &w->X[i]);
  w->pA[par] = sadd(w->pA[par], force_to_parent(&w->X[i], pa));
}

/* base solve + forward pass + base qacc convention (src/g1_core.h:89-104, 187-206) */
static void solve_forward(const real* qvel, ws_t* w, real* qacc) {
  w->a[0] = m6solve(&w->IA[0], ssub(S6(V(0, 0, 0), V(0, 0, 0)), w->pA[0]));
  for (int i = 1; i < NB; ++i) {
    int h = i - 1, par = g1m_parent[i];
    sv Sh = hinge_S(h);
    sv cb = crm(w->v[i], sscl(qvel[6 + h], Sh));
    sv ap = sadd(motion_to_child(&w->X[i], w->a[par]), cb);
    real qdd = (w->u[h] - sdot(w->U[h], ap)) / w->D[h];
    w->a[i] = sadd(ap, sscl(qdd, Sh));
    qacc[6 + h] = qdd;
  }
  v3 aw = w->a[0].a;
  v3 al = mv(&w->oR[0], vadd(w->a[0].l, vcross(w->v[0].a, w->v[0].l)));
  qacc[0] = al.x; qacc[1] = al.y; qacc[2] = al.z;
  qacc[3] = aw.x; qacc[4] = aw.y; qacc[5] = aw.z;
}

/* src/g1_core.h:59-105: ABA with generalized force tau, no contacts/PD */
static void aba_pure(const cfg_t* c, const real* qvel, const real* tau, ws_t* w, real* qacc) {
  init_bias(c, w);
  v3 fb = mtv(&w->oR[0], vldr(tau));
  w->pA[0] = ssub(w->pA[0], S6(vldr(tau + 3), fb));
  for (int i = NB - 1; i >= 1; --i) backward_joint(i, qvel[5 + i], tau[5 + i], 0, w);
  solve_forward(qvel, w, qacc);
}

/* Generalized force of body-frame wrenches fext[NB] (J^T f), base force in
 * world frame and base torque in body frame, like g1_c_fd's tau. */
static void wrench_to_generalized(const ws_t* w, const sv* fext, real* tau) {
  sv F[NB];
  for (int i = 0; i < NB; ++i) F[i] = fext[i];
  for (int i = NB - 1; i >= 1; --i) {
    tau[5 + i] = sdot(hinge_S(i - 1), F[i]);
    F[g1m_parent[i]] = sadd(F[g1m_parent[i]], force_to_parent(&w->X[i], F[i]));
  }
  v3 fw = mv(&w->oR[0], F[0].l);
  tau[0] = fw.x; tau[1] = fw.y; tau[2] = fw.z;
  tau[3] = F[0].a.x; tau[4] = F[0].a.y; tau[5] = F[0].a.z;
}

#define ANCHOR_FREE ((real)1e30)

/* src/g1_core.h:109-207: contacts, PD servos, joint limits, implicit damping
 * and ABA.  tau_out (nullable) = total generalized force. */
static void aba_fused(const cfg_t* c, const real* qpos, const real* qvel, const real* ctrl,
                      real* anchor, ws_t* w, real* qacc, real* fn_out, real* tau_out) {
  init_bias(c, w);
  sv fext[NB];
  if (tau_out) memset(fext, 0, sizeof(fext));
  /* contacts, src/g1_core.h:123-153 */
  for (int k = 0; k < NC; ++k) {
    int b = g1m_contact_body[k];
    v3 r = vld(g1m_contact_pos + 3 * k);
    v3 p = vadd(mv(&w->oR[b], r), w->oP[b]);
    real depth = (real)G1M_CONTACT_RADIUS - p.z;
    real fn = 0;
    if (depth > 0) {
      v3 vp = mv(&w->oR[b], vadd(w->v[b].l, vcross(w->v[b].a, r)));
      fn = c->kn * depth - c->dn * vp.z;
      if (fn < 0) fn = 0;
      if (fn > (real)500) fn = (real)500;
      if (anchor[2 * k] >= ANCHOR_FREE) { anchor[2 * k] = p.x; anchor[2 * k + 1] = p.y; }
      real ftx = -c->kt * (p.x - anchor[2 * k]) - c->dt_t * vp.x;
      real fty = -c->kt * (p.y - anchor[2 * k + 1]) - c->dt_t * vp.y;
      real ftn = sqrt(ftx * ftx + fty * fty), fmax = c->mu * fn;
      if (ftn > fmax) {
        real sc = fmax / (ftn + (real)1e-12);
        ftx *= sc; fty *= sc;
        anchor[2 * k] = p.x + ftx / c->kt;
        anchor[2 * k + 1] = p.y + fty / c->kt;
      }
      v3 fb = mtv(&w->oR[b], V(ftx, fty, fn));
      sv f = S6(vcross(r, fb), fb);
      w->pA[b] = ssub(w->pA[b], f);
      if (tau_out) fext[b] = sadd(fext[b], f);
    } else {
      anchor[2 * k] = ANCHOR_FREE; anchor[2 * k + 1] = ANCHOR_FREE;
    }
    if (fn_out) fn_out[k] = fn;
  }
  if (tau_out) wrench_to_generalized(w, fext, tau_out);
  /* backward pass with PD + limits, src/g1_core.h:156-186 */
  for (int i = NB - 1; i >= 1; --i) {
    int h = i - 1;
    real q = qpos[7 + h], qd = qvel[6 + h];
    real t = (real)g1m_act_kp[h] * (ctrl[h] - q) - (real)g1m_act_kv[h] * qd;
    real lim = (real)g1m_act_frclim[h];
    if (t > lim) t = lim;
    if (t < -lim) t = -lim;
    real lo = (real)g1m_jnt_range[2 * h], hi = (real)g1m_jnt_range[2 * h + 1];
    real bd = (real)g1m_act_kv[h] + (real)g1m_damping[h];
    if (q < lo) { t += c->limit_kp * (lo - q) - c->limit_kd * qd; bd += c->limit_kd; }
    if (q > hi) { t += c->limit_kp * (hi - q) - c->limit_kd * qd; bd += c->limit_kd; }
    if (tau_out) tau_out[6 + h] += t;
    backward_joint(i, qd, t, c->dt * bd, w);
  }
  solve_forward(qvel, w, qacc);
}

/* src/g1_core.h:212-222: semi-implicit Euler */
static void integrate(real dt, const real* qacc, real* qpos, real* qvel) {
  for (int i = 0; i < NV; ++i) qvel[i] += dt * qacc[i];
  qpos[0] += dt * qvel[0];
  qpos[1] += dt * qvel[1];
  qpos[2] += dt * qvel[2];
  quat q = {qpos[3], qpos[4], qpos[5], qpos[6]};
  q = qintegrate(q, vldr(qvel + 3), dt);
  qpos[3] = q.w; qpos[4] = q.x; qpos[5] = q.y; qpos[6] = q.z;
  for (int h = 0; h < NJ; ++h) qpos[7 + h] += dt * qvel[6 + h];
}

/* foot outputs of src/g1_kernel.cu:128-145 (bodies 6 and 12, world frame) */
static void foot_outputs(const ws_t* w, real* fp, real* fv) {
  const int fb[2] = {6, 12};
  for (int s = 0; s < 2; ++s) {
    int b = fb[s];
    v3 vl = mv(&w->oR[b], w->v[b].l);
    if (fp) { fp[3 * s] = w->oP[b].x; fp[3 * s + 1] = w->oP[b].y; fp[3 * s + 2] = w->oP[b].z; }
    if (fv) { fv[3 * s] = vl.x; fv[3 * s + 1] = vl.y; fv[3 * s + 2] = vl.z; }
  }
}

/* src/g1_core.h:225-232 plus the optional outputs */
static void step_env(const cfg_t* c, real* qp, real* qv, const real* ct, real* an, int nsub,
                     real* qacc_last, real* fn, real* fp, real* fv, real* tau) {
  ws_t w;
  real qa[NV];
  for (int s = 0; s < nsub; ++s) {
    int last = (s == nsub - 1);
    real* t = (last && tau) ? tau : NULL;
    if (t) memset(t, 0, NV * sizeof(real));
    fk_vel(qp, qv, &w);
    aba_fused(c, qp, qv, ct, an, &w, qa, last ? fn : NULL, t);
    if (last) {
      foot_outputs(&w, fp, fv);
      if (qacc_last) memcpy(qacc_last, qa, sizeof(qa));
    }
    integrate(c->dt, qa, qp, qv);
  }
}

/* ======================================================== C-ABI */

int g1_sizes(int* nb, int* nq, int* nv, int* nu, int* nc) {
  *nb = NB; *nq = NQ; *nv = NV; *nu = NU; *nc = NC;
  return (int)sizeof(real);
}

void g1_c_fk(const double* qpos, double* xpos, double* xmat) {
  ws_t w;
  real qp[NQ], qv[NV];
  for (int i = 0; i < NQ; ++i) qp[i] = (real)qpos[i];
  for (int i = 0; i < NV; ++i) qv[i] = 0;
  fk_vel(qp, qv, &w);
  for (int b = 0; b < NB; ++b) {
    xpos[3 * b] = w.oP[b].x; xpos[3 * b + 1] = w.oP[b].y; xpos[3 * b + 2] = w.oP[b].z;
    for (int k = 0; k < 9; ++k) xmat[9 * b + k] = w.oR[b].m[k];
  }
}

void g1_c_fd(const double* qpos, const double* qvel, const double* tau, double* qacc) {
  ws_t w;
  cfg_t c = default_cfg();
  real qp[NQ], qv[NV], tu[NV], qa[NV];
  for (int i = 0; i < NQ; ++i) qp[i] = (real)qpos[i];
  for (int i = 0; i < NV; ++i) { qv[i] = (real)qvel[i]; tu[i] = (real)tau[i]; }
  fk_vel(qp, qv, &w);
  aba_pure(&c, qv, tu, &w, qa);
  for (int i = 0; i < NV; ++i) qacc[i] = qa[i];
}

void g1o_step(double* qpos, double* qvel, const double* ctrl, double* anchor, int nsub,
              double dt, double mu, double* qacc_out, double* fn_out, double* foot_pos,
              double* foot_vel, double* tau_out) {
  cfg_t c = default_cfg();
  c.dt = (real)dt;
  c.mu = (real)mu;
  real qp[NQ], qv[NV], ct[NU], an[2 * NC], qa[NV], fn[NC], fp[6], fv[6], tau[NV];
  for (int i = 0; i < NQ; ++i) qp[i] = (real)qpos[i];
  for (int i = 0; i < NV; ++i) qv[i] = (real)qvel[i];
  for (int i = 0; i < NU; ++i) ct[i] = (real)ctrl[i];
  for (int i = 0; i < 2 * NC; ++i) an[i] = (real)anchor[i];
  step_env(&c, qp, qv, ct, an, nsub, qa, fn, fp, fv, tau_out ? tau : NULL);
  for (int i = 0; i < 2 * NC; ++i) anchor[i] = an[i];
  for (int i = 0; i < NQ; ++i) qpos[i] = qp[i];
  for (int i = 0; i < NV; ++i) qvel[i] = qv[i];
  if (nsub > 0) {
    if (qacc_out) for (int i = 0; i < NV; ++i) qacc_out[i] = qa[i];
    if (fn_out) for (int i = 0; i < NC; ++i) fn_out[i] = fn[i];
    if (foot_pos) for (int i = 0; i < 6; ++i) foot_pos[i] = fp[i];
    if (foot_vel) for (int i = 0; i < 6; ++i) foot_vel[i] = fv[i];
    if (tau_out) for (int i = 0; i < NV; ++i) tau_out[i] = tau[i];
  }
}

/* src/host_lib.cpp:38-53 */
void g1_c_step(double* qpos, double* qvel, const double* ctrl, int nsub, double dt,
               double* fn_out, double* anchor) {
  g1o_step(qpos, qvel, ctrl, anchor, nsub, dt, G1M_CONTACT_MU, NULL, fn_out, NULL, NULL, NULL);
}

void g1o_step_batch_f32(int nenv, double dt, const float* qpos, const float* qvel,
                        const float* anchor, const float* ctrl, const float* mu,
                        double* qacc_out, double* qpos_out, double* qvel_out,
                        double* anchor_out, double* fn_out, double* foot_pos,
                        double* foot_vel, double* tau_out, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#else
  (void)nthreads;
#endif
  for (int e = 0; e < nenv; ++e) {
    const size_t N = (size_t)nenv;
    cfg_t c = default_cfg();
    c.dt = (real)dt;
    if (mu) c.mu = (real)mu[e];
    real qp[NQ], qv[NV], ct[NU], an[2 * NC], qa[NV], fn[NC], fp[6], fv[6], tau[NV];
    for (int i = 0; i < NQ; ++i) qp[i] = (real)qpos[i * N + e];
    for (int i = 0; i < NV; ++i) qv[i] = (real)qvel[i * N + e];
    for (int i = 0; i < NU; ++i) ct[i] = (real)ctrl[i * N + e];
    for (int i = 0; i < 2 * NC; ++i) an[i] = (real)anchor[i * N + e];
    step_env(&c, qp, qv, ct, an, 1, qa, fn, fp, fv, tau_out ? tau : NULL);
    for (int i = 0; i < NV; ++i) {
      if (qacc_out) qacc_out[i * N + e] = qa[i];
      if (qvel_out) qvel_out[i * N + e] = qv[i];
      if (tau_out) tau_out[i * N + e] = tau[i];
    }
    for (int i = 0; i < NQ; ++i) if (qpos_out) qpos_out[i * N + e] = qp[i];
    for (int i = 0; i < 2 * NC; ++i) if (anchor_out) anchor_out[i * N + e] = an[i];
    for (int i = 0; i < NC; ++i) if (fn_out) fn_out[i * N + e] = fn[i];
    for (int i = 0; i < 6; ++i) {
      if (foot_pos) foot_pos[i * N + e] = fp[i];
      if (foot_vel) foot_vel[i * N + e] = fv[i];
    }
  }
}

/* ---------------------------------------------- independent checks */

/* Recursive Newton-Euler with the free-base conventions of g1_c_fd. */
void g1o_rnea(const double* qpos, const double* qvel, const double* qacc, double* tau) {
  ws_t w;
  cfg_t c = default_cfg();
  real qp[NQ], qv[NV];
  for (int i = 0; i < NQ; ++i) qp[i] = (real)qpos[i];
  for (int i = 0; i < NV; ++i) qv[i] = (real)qvel[i];
  fk_vel(qp, qv, &w);
  sv f[NB];
  /* base: a0 = [qacc_ang; R^T qacc_lin - omega x v_lin] (inverse of the
   * convention of src/g1_core.h:202-206) */
  v3 al = mtv(&w.oR[0], vld(qacc));
  w.a[0] = S6(vld(qacc + 3), vsub(al, vcross(w.v[0].a, w.v[0].l)));
  for (int i = 0; i < NB; ++i) {
    if (i > 0) {
      int h = i - 1;
      sv Sh = hinge_S(h);
      w.a[i] = sadd(sadd(motion_to_child(&w.X[i], w.a[g1m_parent[i]]), sscl((real)qacc[6 + h], Sh)),
                    crm(w.v[i], sscl(qv[6 + h], Sh)));
    }
    v3 com = vld(g1m_com + 3 * i);
    m6 I = rb_inertia((real)g1m_mass[i], com, g1m_inertia_com + 9 * i);
    sv ag = S6(V(0, 0, 0), mtv(&w.oR[i], V(0, 0, c.gravity_z)));
    f[i] = sadd(m6mul(&I, ssub(w.a[i], ag)), crf(w.v[i], m6mul(&I, w.v[i])));
  }
  for (int i = NB - 1; i >= 1; --i) {
    int h = i - 1;
    tau[6 + h] = sdot(hinge_S(h), f[i]) + g1m_armature[h] * qacc[6 + h] + g1m_damping[h] * qv[6 + h];
    f[g1m_parent[i]] = sadd(f[g1m_parent[i]], force_to_parent(&w.X[i], f[i]));
  }
  v3 fw = mv(&w.oR[0], f[0].l);
  tau[0] = fw.x; tau[1] = fw.y; tau[2] = fw.z;
  tau[3] = f[0].a.x; tau[4] = f[0].a.y; tau[5] = f[0].a.z;
}

/* Composite rigid body algorithm.  Columns of the base use the generalized
 * velocity map v0 = [qvel_ang; R^T qvel_lin]. */
void g1o_crba(const double* qpos, double* M) {
  ws_t w;
  real qp[NQ], qv[NV];
  for (int i = 0; i < NQ; ++i) qp[i] = (real)qpos[i];
  for (int i = 0; i < NV; ++i) qv[i] = 0;
  fk_vel(qp, qv, &w);
  m6 Ic[NB];
  for (int i = 0; i < NB; ++i)
    Ic[i] = rb_inertia((real)g1m_mass[i], vld(g1m_com + 3 * i), g1m_inertia_com + 9 * i);
  for (int i = NB - 1; i >= 1; --i) m6_add_to_parent(&Ic[g1m_parent[i]], &Ic[i], &w.X[i]);
  double Mb[NV * NV];
  memset(Mb, 0, sizeof(Mb));
  /* body-twist coordinates: base dofs 0..5 = [ang; lin] of the base twist */
  for (int i = 1; i < NB; ++i) {
    int di = 5 + i;
    sv F = m6mul(&Ic[i], hinge_S(i - 1));
    Mb[di * NV + di] = sdot(hinge_S(i - 1), F) + g1m_armature[i - 1];
    int j = i;
    while (g1m_parent[j] > 0) {
      F = force_to_parent(&w.X[j], F);
      j = g1m_parent[j];
      double mij = sdot(hinge_S(j - 1), F);
      Mb[di * NV + 5 + j] = mij;
      Mb[(5 + j) * NV + di] = mij;
    }
    F = force_to_parent(&w.X[j], F);          /* into the base frame */
    double col[6] = {F.a.x, F.a.y, F.a.z, F.l.x, F.l.y, F.l.z};
    for (int k = 0; k < 6; ++k) { Mb[k * NV + di] = col[k]; Mb[di * NV + k] = col[k]; }
  }
  for (int r = 0; r < 6; ++r)
    for (int s = 0; s < 6; ++s) Mb[r * NV + s] = m6get(&Ic[0], r, s);
  /* T maps generalized base velocity [lin_w; ang_b] to the twist [ang; lin_b]:
   * twist_ang = qv[3:6], twist_lin = R^T qv[0:3]. M = T^T Mb T on base blocks. */
  double T[36];
  memset(T, 0, sizeof(T));
  for (int r = 0; r < 3; ++r) {
    T[r * 6 + 3 + r] = 1.0;
    for (int s = 0; s < 3; ++s) T[(3 + r) * 6 + s] = w.oR[0].m[3 * s + r];
  }
  double tmp[NV * NV];
  memcpy(tmp, Mb, sizeof(Mb));
  for (int r = 0; r < NV; ++r)         /* columns: tmp[:, base] = Mb[:, base] T */
    for (int s = 0; s < 6; ++s) {
      double acc = 0;
      for (int k = 0; k < 6; ++k) acc += Mb[r * NV + k] * T[k * 6 + s];
      tmp[r * NV + s] = acc;
    }
  memcpy(M, tmp, sizeof(tmp));
  for (int s = 0; s < NV; ++s)         /* rows: M[base, :] = T^T tmp[base, :] */
    for (int r = 0; r < 6; ++r) {
      double acc = 0;
      for (int k = 0; k < 6; ++k) acc += T[k * 6 + r] * tmp[k * NV + s];
      M[r * NV + s] = acc;
    }
}

void g1o_m6_solve(const double* A21, const double* b6, double* x6) {
  real A[21], b[6], y[6];
  for (int i = 0; i < 21; ++i) A[i] = (real)A21[i];
  for (int i = 0; i < 6; ++i) b[i] = (real)b6[i];
  m6_cholesky_solve(A, b, y);
  for (int i = 0; i < 6; ++i) x6[i] = y[i];
}

void g1o_quat_integrate(double* q4, const double* w3, double dt) {
  quat q = {(real)q4[0], (real)q4[1], (real)q4[2], (real)q4[3]};
  q = qintegrate(q, vld(w3), (real)dt);
  q4[0] = q.w; q4[1] = q.x; q4[2] = q.y; q4[3] = q.z;
}

/* src/g1_kernel.cu:54-58 */
unsigned g1o_hash(unsigned x) {
  x ^= x >> 16; x *= 0x7feb352du;
  x ^= x >> 15; x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

/* src/g1_kernel.cu:59-62 and 85-103 in float32.  The noise term is one
 * fused multiply-add, v = fma(noise, 2u - 1, v), the contraction a CUDA build
 * of the reference source applies by default (2u - 1 itself is exact). */
void g1o_reset_env_f32(int env, unsigned seed, float noise, float drop, float* qpos,
                       float* qvel, float* anchor) {
  unsigned s = g1o_hash(seed ^ (unsigned)(env + 1));
  for (int i = 0; i < NQ; ++i) {
    volatile float v = (float)g1m_qpos_stand[i];
    if (i == 2) v = v + drop;
    if (i >= 7) {
      s = g1o_hash(s + 0x9E3779B9u);
      volatile float u = (float)s * 2.3283064365386963e-10f;
      volatile float two_u = 2.0f * u;
      volatile float r = two_u - 1.0f;
      v = fmaf(noise, r, v);
    }
    qpos[i] = v;
  }
  for (int i = 0; i < NV; ++i) qvel[i] = 0.0f;
  for (int i = 0; i < 2 * NC; ++i) anchor[i] = 1e30f;
}

double g1o_bench_f32(int nenv, int nsteps, double dt, float* qpos, float* qvel, float* anchor,
                     const float* ctrl, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  double t0 = omp_get_wtime();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#else
  (void)nthreads;
  double t0 = 0;
#endif
  for (int e = 0; e < nenv; ++e) {
    const size_t N = (size_t)nenv;
    cfg_t c = default_cfg();
    c.dt = (real)dt;
    real qp[NQ], qv[NV], ct[NU], an[2 * NC];
    for (int i = 0; i < NQ; ++i) qp[i] = (real)qpos[i * N + e];
    for (int i = 0; i < NV; ++i) qv[i] = (real)qvel[i * N + e];
    for (int i = 0; i < NU; ++i) ct[i] = (real)ctrl[i * N + e];
    for (int i = 0; i < 2 * NC; ++i) an[i] = (real)anchor[i * N + e];
    for (int s = 0; s < nsteps; ++s) step_env(&c, qp, qv, ct, an, 1, NULL, NULL, NULL, NULL, NULL);
    for (int i = 0; i < NQ; ++i) qpos[i * N + e] = (float)qp[i];
    for (int i = 0; i < NV; ++i) qvel[i * N + e] = (float)qv[i];
    for (int i = 0; i < 2 * NC; ++i) anchor[i * N + e] = (float)an[i];
  }
#ifdef _OPENMP
  return omp_get_wtime() - t0;
#else
  return 0;
#endif
}
