// g1_fastmath.h -- float helpers shared by the step and inference kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace g1k {

#ifndef G1K_DEV
#define G1K_DEV __device__ __forceinline__
#endif

// ------------------------------------------------------------ fast math
// Hardware reciprocal / square root / reciprocal square root (1 ulp) refined
// by one Newton step, in place of the IEEE division and square-root
// sequences (about 10 instructions each).  Arguments are positive and
// normal where these are used (joint inertias, Cholesky pivots, norms).
G1K_DEV float frcp(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  const float e = __builtin_fmaf(-x, r, 1.0f);
  return __builtin_fmaf(r, e, r);
}
G1K_DEV float frsq(float x) {
  float r = __builtin_amdgcn_rsqf(x);
  const float h = 0.5f * x * r;                    // Newton on 1/sqrt: r (1.5 - 0.5 x r^2)
  return __builtin_fmaf(r, __builtin_fmaf(-h, r, 0.5f), r);
}
G1K_DEV float fsqrt(float x) {                    // x >= 0; sqrt(0) = 0
  const float r = frsq(x);
  return x > 0.f ? x * r : 0.f;
}
G1K_DEV double frcp(double x) { return 1.0 / x; }
G1K_DEV double fsqrt(double x) { return sqrt(x); }
G1K_DEV double frsq(double x) { return 1.0 / sqrt(x); }

// sin and cos of a joint angle: Cody-Waite reduction by pi/2 (three-part
// constant, exact products for |x| < 2^12) and Taylor polynomials on
// |r| <= pi/4 (truncation below 2e-9).  Larger arguments take the library
// routine.
G1K_DEV void fsincos(float x, float* sp, float* cp) {
  if (__builtin_expect(fabsf(x) > 4096.f, 0)) {
    sincosf(x, sp, cp);
    return;
  }
  const float k = __builtin_rintf(x * 0.63661977236758134f);     // 2/pi
  float r = __builtin_fmaf(-k, 1.5707963705062866f, x);          // pi/2 = P1 + P2 + P3
  r = __builtin_fmaf(-k, -4.3711388286737929e-08f, r);
  r = __builtin_fmaf(-k, -1.7151245100059206e-15f, r);
  const float r2 = r * r;
  // sin r = r (1 - r^2/3! + r^4/5! - r^6/7! + r^8/9!)
  float ps = __builtin_fmaf(r2, 2.7557319223985888e-06f, -1.9841269841269841e-04f);
  ps = __builtin_fmaf(ps, r2, 8.3333333333333333e-03f);
  ps = __builtin_fmaf(ps, r2, -1.6666666666666666e-01f);
  const float sn = __builtin_fmaf(ps * r2, r, r);
  // cos r = 1 - r^2/2! + r^4/4! - r^6/6! + r^8/8! - r^10/10!
  float pc = __builtin_fmaf(r2, -2.7557319223985888e-07f, 2.4801587301587302e-05f);
  pc = __builtin_fmaf(pc, r2, -1.3888888888888889e-03f);
  pc = __builtin_fmaf(pc, r2, 4.1666666666666667e-02f);
  pc = __builtin_fmaf(pc, r2, -0.5f);
  const float cs = __builtin_fmaf(pc, r2, 1.0f);
  const int quad = (int)k & 3;
  const float s0 = (quad & 1) ? cs : sn;
  const float c0 = (quad & 1) ? sn : cs;
  *sp = (quad & 2) ? -s0 : s0;
  *cp = ((quad + 1) & 2) ? -c0 : c0;
}

}  // namespace g1k
