// g1_act.h -- activations shared by the GEMM epilogues (g1_mlp.hip) and the
// fused inference kernels (g1_infer.hip), so both produce the same bits.
// Hardware exp2 (v_exp_f32) and reciprocal (v_rcp_f32): about 1 ulp each,
// far below the bf16 rounding the activations then go through.
#pragma once
#include <hip/hip_runtime.h>

namespace g1act {

__device__ __forceinline__ float exp_fast(float x) {
  return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);    // 2^(x log2 e); 0 for x -> -inf, inf for x -> inf
}
// 1 / (1 + e^-z): rcp(inf) = 0 for z -> -inf, rcp(1) = 1 for z -> inf
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + exp_fast(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
// d silu / dz = s (1 + z (1 - s))
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}
// ELU with alpha 1 (torch.nn.ELU): z for z > 0, e^z - 1 otherwise.  As
// e^z - 1 >= z everywhere, that is the median of (z, e^z - 1, 0): one
// v_med3_f32 in place of a compare and a select.  (Below |z| ~ 5e-4 the
// rounded e^z - 1 can fall on the other side of z; the two forms then differ
// by less than 3e-7.)
__device__ __forceinline__ float elu(float z) { return __builtin_amdgcn_fmed3f(z, exp_fast(z) - 1.0f, 0.0f); }

}  // namespace g1act
