// g1_peaks.hip -- machine-peak microbenchmarks (measurement only, not on the
// step path): non-packed FP32 FMA issue rate and dense BF16 MFMA rate on the
// device, so roofline fractions can be read against measured as well as
// published peaks.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1_err.h"

namespace g1peak {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// 8 independent FMA chains per lane (enough to cover the VALU latency at
// any occupancy); every result feeds the output so nothing is dead.
__global__ __launch_bounds__(256) void fma_kernel(int iters, float a, float b, float* out) {
  float x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = (float)(threadIdx.x + k) * 1e-3f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = __builtin_fmaf(x[k], a, b);
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += x[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 16 independent 16x16x32 BF16 MFMA accumulators per wave (a 4 x 4 tile of
// distinct A and B fragments, as a GEMM inner loop holds them) on register
// operands.
__global__ __launch_bounds__(256) void mfma_kernel(int iters, float* out) {
  const int lane = threadIdx.x & 63;
  bf16x8 a[4], b[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[r][k] = (bf16)(0.001f * (lane + k + r));
      b[r][k] = (bf16)(0.002f * (lane - k - r));
    }
  f32x4 acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b[c], acc[r][c], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) s += acc[r][c][0] + acc[r][c][1] + acc[r][c][2] + acc[r][c][3];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace g1peak

extern "C" {

// FP32: blocks x 256 threads x iters x 8 FMAs (2 FLOP each).  out: float
// [blocks * 256].
int g1_hip_peak_fp32(int blocks, int iters, float* out, void* stream) {
  g1err::clear();
  if (blocks <= 0 || iters <= 0 || !out) return g1err::fail("g1_hip_peak_fp32: invalid arguments");
  hipLaunchKernelGGL(g1peak::fma_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, iters, 0.999f, 1e-4f,
                     out);
  return g1err::launch("g1_hip_peak_fp32");
}

// BF16 MFMA: blocks x 4 waves x iters x 16 MFMAs of 16 x 16 x 32 (16,384
// FLOP each).  out: float [blocks * 256].
int g1_hip_peak_bf16(int blocks, int iters, float* out, void* stream) {
  g1err::clear();
  if (blocks <= 0 || iters <= 0 || !out) return g1err::fail("g1_hip_peak_bf16: invalid arguments");
  hipLaunchKernelGGL(g1peak::mfma_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, iters, out);
  return g1err::launch("g1_hip_peak_bf16");
}

}  // extern "C"
