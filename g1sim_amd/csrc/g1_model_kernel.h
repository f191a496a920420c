// g1_model_kernel.h -- compile-time view of the G1 model for the HIP kernels.
//
// Everything here is constexpr and evaluated by the compiler in float64, then
// rounded once to float where a kernel uses it, so model constants become
// instruction literals (no loads, no registers).  State-independent terms of
// the reference step are precomputed here instead of per step:
//   * tree rotations from the (already unit) tree quaternions
//     (src/g1_core.h:44-45 normalises and converts them every step);
//   * rigid-body spatial inertias about the body origin
//     (src/spatial.h:152-172, called per body per step at src/g1_core.h:115);
//   * hinge axes are unit basis vectors and anchors are zero
//     (src/g1_model.h:211-272), asserted below, so S = [e_k; 0].
#pragma once

#define G1M_CONST static constexpr
#include "g1_model_data.h"

namespace g1k {

constexpr int NB = G1M_NB, NQ = G1M_NQ, NV = G1M_NV, NU = G1M_NU, NC = G1M_NC;

struct BodyK {
  int parent;
  int axis;        // hinge axis index (0 x, 1 y, 2 z); -1 for the free base
  bool rt_ident;   // tree rotation is the identity
  int rt_axis;     // the tree rotation is a planar rotation about unit axis rt_axis (-1: general)
  double rt_c, rt_s;  // its cosine and sine (R[I][I], R[J][I] with I, J = axis + 1, + 2)
  bool rt_fold;    // planar about the hinge's own axis: R_tree R_K(q) = R_K(theta + q)
  double RT[9];    // tree rotation (pose of the joint frame in the parent), row-major
  double t[3];     // tree position
  double I[21];    // spatial inertia about the body origin, packed lower triangle
};

struct ModelK {
  BodyK b[NB];
};

constexpr int pk(int r, int c) { return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r; }

constexpr BodyK make_body(int i) {
  BodyK B{};
  B.parent = g1m_parent[i];
  B.axis = -1;
  if (i > 0) {
    int n1 = 0;
    for (int k = 0; k < 3; ++k) {
      double a = g1m_jnt_axis[3 * (i - 1) + k];
      if (a == 1.0) B.axis = k;
      if (a != 0.0) ++n1;
      if (g1m_jnt_anchor[3 * (i - 1) + k] != 0.0) B.axis = -100;  // rejected below
    }
    if (n1 != 1) B.axis = -100;
  }
  const double w = g1m_tree_quat[4 * i], x = g1m_tree_quat[4 * i + 1],
               y = g1m_tree_quat[4 * i + 2], z = g1m_tree_quat[4 * i + 3];
  B.rt_ident = (w == 1.0 && x == 0.0 && y == 0.0 && z == 0.0);
  B.RT[0] = 1 - 2 * (y * y + z * z); B.RT[1] = 2 * (x * y - w * z); B.RT[2] = 2 * (x * z + w * y);
  B.RT[3] = 2 * (x * y + w * z); B.RT[4] = 1 - 2 * (x * x + z * z); B.RT[5] = 2 * (y * z - w * x);
  B.RT[6] = 2 * (x * z - w * y); B.RT[7] = 2 * (y * z + w * x); B.RT[8] = 1 - 2 * (x * x + y * y);
  B.rt_axis = -1;
  B.rt_c = 1.0;
  B.rt_s = 0.0;
  for (int k = 0; k < 3 && !B.rt_ident; ++k) {
    const int I = (k + 1) % 3, J = (k + 2) % 3;
    const bool planar = B.RT[3 * k + k] == 1.0 && B.RT[3 * k + I] == 0.0 && B.RT[3 * k + J] == 0.0 &&
                        B.RT[3 * I + k] == 0.0 && B.RT[3 * J + k] == 0.0 && B.RT[3 * I + I] == B.RT[3 * J + J] &&
                        B.RT[3 * I + J] == -B.RT[3 * J + I];
    if (planar) {
      B.rt_axis = k;
      B.rt_c = B.RT[3 * I + I];
      B.rt_s = B.RT[3 * J + I];
    }
  }
  B.rt_fold = B.rt_axis >= 0 && B.rt_axis == B.axis;
  for (int k = 0; k < 3; ++k) B.t[k] = g1m_tree_pos[3 * i + k];
  const double m = g1m_mass[i];
  const double cx = g1m_com[3 * i], cy = g1m_com[3 * i + 1], cz = g1m_com[3 * i + 2];
  const double C[9] = {0, -cz, cy, cz, 0, -cx, -cy, cx, 0};
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c <= r; ++c) {
      double v = 0;
      if (r < 3) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += C[3 * r + k] * C[3 * k + c];
        v = g1m_inertia_com[9 * i + 3 * r + c] - m * s;
      } else if (c < 3) {
        v = -m * C[3 * (r - 3) + c];
      } else {
        v = (r == c) ? m : 0.0;
      }
      B.I[pk(r, c)] = v;
    }
  return B;
}

constexpr ModelK make_model() {
  ModelK M{};
  for (int i = 0; i < NB; ++i) M.b[i] = make_body(i);
  return M;
}

static constexpr ModelK kM = make_model();

constexpr bool model_ok() {
  for (int i = 1; i < NB; ++i)
    if (kM.b[i].axis < 0) return false;
  for (int h = 0; h < NU; ++h)
    if (g1m_damping[h] != 0.0) return false;  // damping terms are folded out
  return true;
}
static_assert(model_ok(), "kernels assume unit hinge axes, zero anchors and zero joint damping");

// Chains of the kinematic tree (src/g1_model.h:15): each a run of bodies whose
// parent is the previous one, hanging off `root`.
constexpr bool is_chain(int first, int n, int root) {
  if (kM.b[first].parent != root) return false;
  for (int i = 1; i < n; ++i)
    if (kM.b[first + i].parent != first + i - 1) return false;
  return true;
}
static_assert(is_chain(1, 6, 0) && is_chain(7, 6, 0), "legs");
static_assert(is_chain(13, 3, 0), "waist");
static_assert(is_chain(16, 7, 15) && is_chain(23, 7, 15), "arms");
static_assert(g1m_contact_body[0] == 6 && g1m_contact_body[3] == 6 && g1m_contact_body[4] == 12 &&
                  g1m_contact_body[7] == 12,
              "contacts 0-3 on the left foot (body 6), 4-7 on the right foot (body 12)");

}  // namespace g1k
