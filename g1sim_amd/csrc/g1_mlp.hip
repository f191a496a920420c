// g1_mlp.hip -- BF16 MFMA kernels of the learned forward-dynamics surrogate.
//
// One tiled GEMM, C[I x J] = sum_r A(i, r) B(j, r), with bf16 operands, fp32
// accumulation on v_mfma_f32_16x16x32_bf16 (gfx950) and fused epilogues, plus
// the small elementwise kernels of training (loss gradient, bias gradient,
// Adam) and of inference (feature gather/normalisation).  All activations are
// sample-major ([batch][features], features contiguous); the three GEMMs of a
// dense layer are then
//   forward        Z  = H W^T          A = H  (row),  B = W  (row)
//   data grad      dH = dZ W           A = dZ (row),  B = W  (transposed)
//   weight grad    dW = dZ^T H         A = dZ (trans), B = H (transposed)
// selected by the A_T / B_T template flags (transposed operands are turned
// into the [row][r] LDS layout while they are staged).
//
// Tiles: 128 x 128 per 256-thread workgroup (2 x 2 waves of 64 x 64, each 4 x 4
// MFMA tiles of 16 x 16), r-steps of 32, double-buffered LDS with a 16-byte row
// pad.  The weight-gradient GEMM reduces over the batch, so it splits r across
// gridDim.z and adds partial tiles with fp32 atomics.  Dimensions are padded by
// the host: I, J multiples of 128, R a multiple of 32 (times the split).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "g1_act.h"
#include "g1_err.h"

namespace g1mlp {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Problem shapes are multiples of BI x BJ (x BR per r-slice); a launch
// tiles them with T x T output tiles, T = 128, or 64 when 128-tiles would
// leave the chip under-filled (a workgroup's few r-steps cannot hide the
// global latency by themselves: four 64-tile workgroups per CU hide it for
// each other).
constexpr int BI = 128, BJ = 128, BR = 64, PADR = 8, LDR = BR + PADR;
template <int T> struct Tiling {
  static constexpr int CHUNKS = T * BR / 8 / 256;   // 16-byte chunks of one operand tile per thread
  static_assert(CHUNKS % 2 == 0, "transposed staging pairs the chunks");
  static constexpr int F = T / 32;                  // 16 x 16 MFMA tiles per wave and dimension
};

enum Epi : int {
  EPI_F32 = 0,         // C (fp32) = acc
  EPI_F32_ATOMIC = 1,  // C (fp32) += acc   (split-r partial tiles)
  EPI_BIAS_SILU = 2,   // Z (fp32) = acc + b; H (bf16) = silu(Z)     (training forward)
  EPI_BIAS_SILU_H = 3, // H (bf16) = silu(acc + b)                   (inference forward)
  EPI_BIAS_F32 = 4,    // C (fp32) = acc + b                          (output layer)
  // (the three bias epilogues start the accumulators at b, so acc + b is the
  // accumulator itself: the fused inference kernels do the same, bit for bit)
  EPI_SILU_BWD = 5,    // dZ = acc * silu'(Z): fp32 to C, bf16 to H   (data grad through SiLU)
  EPI_F32_SPLIT = 6,   // C[z] (fp32) = acc of r-slice z; C[z] at C + z * I * ldc (summed by g1_mlp_sum_splits)
};

struct GemmArgs {
  const bf16* A;
  const bf16* B;
  int lda, ldb;          // leading dimensions of A, B as stored
  int I, J, R;           // problem size (padded)
  int r_per_split;       // r extent of one gridDim.z slice
  float* C;              // fp32 output (EPI_F32*, Z of BIAS_SILU, dZ of SILU_BWD, out of BIAS_F32)
  bf16* H;               // bf16 output (BIAS_SILU*, SILU_BWD)
  int ldc;               // leading dimension of C and H
  const float* bias;     // [J]
  const float* Zin;      // pre-activation for SILU_BWD, leading dimension ldc
  float* colpart;        // SILU_BWD: column sums of dZ per 64-row block, [I / 64][ldc] (or null)
};

// One operand tile (128 rows x BR) of a step goes global -> registers (load,
// issued before the step's MFMAs) -> LDS (put, after them), so the global
// latency hides behind the MFMAs.  The LDS tile is [row][r] either way.
template <int T, bool TRANS>
__device__ __forceinline__ void load(bf16x8 (&v)[Tiling<T>::CHUNKS], const bf16* g, int ld, int row0, int r0,
                                     int tid) {
#pragma unroll
  for (int t = 0; t < Tiling<T>::CHUNKS; ++t) {
    const int c = tid + 256 * t;
    if constexpr (!TRANS) {
      // stored [row][r], r contiguous: BR / 8 chunks per row
      const int row = c / (BR / 8), rc = (c % (BR / 8)) * 8;
      v[t] = *reinterpret_cast<const bf16x8*>(g + (size_t)(row0 + row) * ld + r0 + rc);
    } else {
      // stored [r][row], row contiguous: T / 8 chunks of 8 rows per r; a
      // thread takes the same chunk at two consecutive r (t even, t + 1) so
      // that put() can write (r, r + 1) pairs as 4-byte words
      const int cp = tid + 256 * (t / 2);
      const int rr = 2 * (cp / (T / 8)) + (t & 1), rowc = (cp % (T / 8)) * 8;
      v[t] = *reinterpret_cast<const bf16x8*>(g + (size_t)(r0 + rr) * ld + row0 + rowc);
    }
  }
}

template <int T, bool TRANS>
__device__ __forceinline__ void put(bf16 (*s)[LDR], const bf16x8 (&v)[Tiling<T>::CHUNKS], int tid) {
#pragma unroll
  for (int t = 0; t < Tiling<T>::CHUNKS; ++t) {
    const int c = tid + 256 * t;
    if constexpr (!TRANS) {
      const int row = c / (BR / 8), rc = (c % (BR / 8)) * 8;
      *reinterpret_cast<bf16x8*>(&s[row][rc]) = v[t];
    } else if ((t & 1) == 0) {
      const int cp = tid + 256 * (t / 2);
      const int rr = 2 * (cp / (T / 8)), rowc = (cp % (T / 8)) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const bf16 pair[2] = {v[t][q], v[t + 1][q]};
        // one 4-byte LDS store (aligned: rr is even); memcpy, not a store
        // through an unsigned*, which the compiler may assume does not alias
        // the bf16 fragment reads of the tile (that form gave wrong tiles)
        __builtin_memcpy(&s[rowc + q][rr], pair, 4);
      }
    }
  }
}

// One workgroup's output tile (bx, by) of r-slice bz, with the operand
// tiles staged in sA / sB.
template <bool A_T, bool B_T, int EPI, int T>
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, int bx, int by, int bz, bf16 (*sA)[T][LDR],
                                          bf16 (*sB)[T][LDR]) {
  using TL = Tiling<T>;
  constexpr int F = TL::F;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  const int i0 = bx * T, j0 = by * T;
  const int rb = bz * g.r_per_split;
  const int nsteps = g.r_per_split / BR;

  // accumulators start at 0, or at the bias of the lane's column (lane % 16 of
  // each tile; the bias epilogues run unsplit)
  constexpr bool BIAS_INIT = EPI == EPI_BIAS_SILU || EPI == EPI_BIAS_SILU_H || EPI == EPI_BIAS_F32;
  f32x4 acc[F][F];
#pragma unroll
  for (int b = 0; b < F; ++b) {
    const float v = BIAS_INIT ? g.bias[j0 + wj * (T / 2) + b * 16 + (lane & 15)] : 0.f;
#pragma unroll
    for (int a = 0; a < F; ++a) acc[a][b] = f32x4{v, v, v, v};
  }

  bf16x8 va[TL::CHUNKS], vb[TL::CHUNKS];
  load<T, A_T>(va, g.A, g.lda, i0, rb, tid);
  load<T, B_T>(vb, g.B, g.ldb, j0, rb, tid);
  put<T, A_T>(sA[0], va, tid);
  put<T, B_T>(sB[0], vb, tid);
  __syncthreads();
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  for (int st = 0; st < nsteps; ++st) {
    const int cur = st & 1;
    const bool next = st + 1 < nsteps;
    if (next) {
      load<T, A_T>(va, g.A, g.lda, i0, rb + (st + 1) * BR, tid);
      load<T, B_T>(vb, g.B, g.ldb, j0, rb + (st + 1) * BR, tid);
    }
#pragma unroll
    for (int kk = 0; kk < BR / 32; ++kk) {
      bf16x8 fa[F], fb[F];
#pragma unroll
      for (int a = 0; a < F; ++a)
        fa[a] = *reinterpret_cast<const bf16x8*>(&sA[cur][wi * (T / 2) + a * 16 + lr][kk * 32 + lk]);
#pragma unroll
      for (int b = 0; b < F; ++b)
        fb[b] = *reinterpret_cast<const bf16x8*>(&sB[cur][wj * (T / 2) + b * 16 + lr][kk * 32 + lk]);
#pragma unroll
      for (int a = 0; a < F; ++a)
#pragma unroll
        for (int b = 0; b < F; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
    // the other buffer was last read in step st - 1, before that step's barrier
    if (next) {
      put<T, A_T>(sA[cur ^ 1], va, tid);
      put<T, B_T>(sB[cur ^ 1], vb, tid);
    }
    __syncthreads();
  }

  // column sums of the SILU_BWD output over this lane's rows, per column tile
  float csum[F];
#pragma unroll
  for (int b = 0; b < F; ++b) csum[b] = 0.f;
  // D fragment: lane holds rows 4*(lane/16)+q, column lane%16 of each 16x16 tile
#pragma unroll
  for (int a = 0; a < F; ++a)
#pragma unroll
    for (int b = 0; b < F; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = i0 + wi * (T / 2) + a * 16 + (lane >> 4) * 4 + q;
        const int j = j0 + wj * (T / 2) + b * 16 + lr;
        const size_t o = (size_t)i * g.ldc + j;
        const float v = acc[a][b][q];
        if constexpr (EPI == EPI_F32) {
          g.C[o] = v;
        } else if constexpr (EPI == EPI_F32_ATOMIC) {
          atomicAdd(&g.C[o], v);
        } else if constexpr (EPI == EPI_F32_SPLIT) {
          g.C[(size_t)bz * g.I * g.ldc + o] = v;
        } else if constexpr (EPI == EPI_BIAS_SILU) {
          g.C[o] = v;
          g.H[o] = (bf16)g1act::silu(v);
        } else if constexpr (EPI == EPI_BIAS_SILU_H) {
          g.H[o] = (bf16)g1act::silu(v);
        } else if constexpr (EPI == EPI_BIAS_F32) {
          g.C[o] = v;
        } else if constexpr (EPI == EPI_SILU_BWD) {
          const float d = v * g1act::silu_grad(g.Zin[o]);
          if (g.C) g.C[o] = d;
          g.H[o] = (bf16)d;
          csum[b] += d;
        }
      }
  if constexpr (EPI == EPI_SILU_BWD) {
    // The bias gradient's partial sums, in a fixed order: per lane over its
    // row tiles and rows (a, q ascending), then over the 4 lane groups of a
    // column (xor 16, then xor 32: every lane ends with the same bits), then
    // (64-row tiles) the two waves of a column half, wave row 0 first.
    if (g.colpart) {                          // uniform
#pragma unroll
      for (int b = 0; b < F; ++b) {
        csum[b] += __shfl_xor(csum[b], 16, 64);
        csum[b] += __shfl_xor(csum[b], 32, 64);
      }
      if constexpr (T == 128) {
        // a wave's 64 rows are one 64-row block
        if (lane < 16) {
#pragma unroll
          for (int b = 0; b < F; ++b)
            g.colpart[(size_t)(bx * 2 + wi) * g.ldc + j0 + wj * (T / 2) + b * 16 + lr] = csum[b];
        }
      } else {
        // the workgroup's 64 rows are one block: the waves of a column half
        // (wi = 0, 1) meet in LDS (the operand tiles are no longer read)
        float* red = reinterpret_cast<float*>(&sA[0][0][0]);          // [2][T]
        if (lane < 16) {
#pragma unroll
          for (int b = 0; b < F; ++b) red[wi * T + wj * (T / 2) + b * 16 + lr] = csum[b];
        }
        __syncthreads();
        if (tid < T) g.colpart[(size_t)bx * g.ldc + j0 + tid] = red[tid] + red[T + tid];
      }
    }
  }
}

template <bool A_T, bool B_T, int EPI, int T>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 sA[2][T][LDR];
  __shared__ __attribute__((aligned(16))) bf16 sB[2][T][LDR];
  gemm_tile<A_T, B_T, EPI, T>(g, blockIdx.x, blockIdx.y, blockIdx.z, sA, sB);
}

// Two independent GEMMs with one tile size in one grid: blocks [0, n0) are
// the first's, the rest the second's, each in the order its own launch
// would give them (x fastest, then y, then r-slice).
template <bool A0, bool B0, int E0, bool A1, bool B1, int E1, int T>
__global__ __launch_bounds__(256) void gemm_pair_kernel(GemmArgs g0, GemmArgs g1, int nx0, int ny0, int n0, int nx1,
                                                        int ny1) {
  __shared__ __attribute__((aligned(16))) bf16 sA[2][T][LDR];
  __shared__ __attribute__((aligned(16))) bf16 sB[2][T][LDR];
  const int b = blockIdx.x;
  if (b < n0) {
    gemm_tile<A0, B0, E0, T>(g0, b % nx0, (b / nx0) % ny0, b / (nx0 * ny0), sA, sB);
  } else {
    const int c = b - n0;
    gemm_tile<A1, B1, E1, T>(g1, c % nx1, (c / nx1) % ny1, c / (nx1 * ny1), sA, sB);
  }
}

// ---------------------------------------------------------------- training
// dY = 2 (pred - target) / (B * n_out) on the first n_out columns, 0 on padding;
// also accumulates the summed squared error into *sse (fp32 atomics).  A
// grid-stride loop over at most MSE_BLOCKS workgroups keeps the atomics on
// the one address few (each costs ~13 ns of serialisation there).
constexpr int MSE_BLOCKS = 256;
__global__ __launch_bounds__(256) void mse_grad_kernel(const float* pred, const float* target, int B, int ld,
                                                       int n_out, bf16* dY16, float* dY32, float* sse) {
  __shared__ float wsum[4];
  const int n = B * ld;
  const float scale = 2.0f / (float)(B * n_out);
  float e2 = 0.f;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
    const int j = idx % ld;
    float d = 0.f;
    if (j < n_out) {
      const float e = pred[idx] - target[idx];
      d = scale * e;
      e2 += e * e;
    }
    dY16[idx] = (bf16)d;
    if (dY32) dY32[idx] = d;
  }
  if (!sse) return;                         // uniform
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = e2;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(sse, wsum[0] + wsum[1] + wsum[2] + wsum[3]);
}

// The same loss gradient without atomics, for the training iteration: one
// workgroup per block of 64 rows x 32 columns writes dY as bf16, the block's
// column sums of dY (the output layer's bias-gradient partials,
// colpart[(i / 64) * ld + j]) and its summed squared error
// (sse_part[(i / 64) * (ld / 32) + j / 32]), each in a fixed order: a
// thread takes rows h, h + 8, ... of its column, the 8 row groups then add
// in order.
__global__ __launch_bounds__(256) void mse_grad_part_kernel(const float* pred, const float* target, int ld, int n_out,
                                                            float scale, bf16* dY16, float* colpart,
                                                            float* sse_part) {
  __shared__ float red[8][32];
  __shared__ float wsum[4];
  const int tid = threadIdx.x, jl = tid & 31, h = tid >> 5;
  const int j = blockIdx.y * 32 + jl;
  const size_t r0 = (size_t)blockIdx.x * 64;
  float cs = 0.f, e2 = 0.f;
#pragma unroll
  for (int r = h; r < 64; r += 8) {
    const size_t idx = (r0 + r) * ld + j;
    float d = 0.f;
    if (j < n_out) {
      const float e = pred[idx] - target[idx];
      d = scale * e;
      e2 += e * e;
    }
    dY16[idx] = (bf16)d;
    cs += d;
  }
  red[h][jl] = cs;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
  if ((tid & 63) == 0) wsum[tid >> 6] = e2;
  __syncthreads();
  if (tid < 32) {
    float c = red[0][tid];
#pragma unroll
    for (int k = 1; k < 8; ++k) c += red[k][tid];
    colpart[(size_t)blockIdx.x * ld + j] = c;
  }
  if (tid == 0) sse_part[(size_t)blockIdx.x * gridDim.y + blockIdx.y] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

// out[i] = sum over z < split of parts[z * n + i] (the r-slices of a split
// weight-gradient GEMM), float4 at a time when n is a multiple of 4.
__global__ void sum_splits_kernel(const float* parts, int split, size_t n, float* out) {
  const size_t i4 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i4 >= n) return;
  if (i4 + 4 <= n && (n & 3) == 0) {
    float4 acc = *reinterpret_cast<const float4*>(parts + i4);
    for (int z = 1; z < split; ++z) {
      const float4 v = *reinterpret_cast<const float4*>(parts + (size_t)z * n + i4);
      acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
    }
    *reinterpret_cast<float4*>(out + i4) = acc;
  } else {
    for (size_t i = i4; i < n && i < i4 + 4; ++i) {
      float a = 0.f;
      for (int z = 0; z < split; ++z) a += parts[(size_t)z * n + i];
      out[i] = a;
    }
  }
}

// db[j] = sum_i dZ[i][j]: a workgroup covers 64 columns x rows_per_block
// rows; its 4 waves take every 4th row (coalesced 256-byte row segments),
// reduce through LDS and add one partial per column.
__global__ __launch_bounds__(256) void colsum_kernel(const float* dZ, int B, int ld, int rows_per_block, float* db) {
  __shared__ float part[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  const int r0 = blockIdx.y * rows_per_block;
  const int r1 = min(B, r0 + rows_per_block);
  float s = 0.f;
  if (j < ld) {
#pragma unroll 4
    for (int i = r0 + w; i < r1; i += 4) s += dZ[(size_t)i * ld + j];
  }
  part[w][lane] = s;
  __syncthreads();
  if (w == 0 && j < ld) atomicAdd(&db[j], part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]);
}

// Adam on fp32 master weights; refreshes the bf16 copy used by the GEMMs.
// Entries with mask == 0 (padding) stay exactly zero.
__global__ void adam_kernel(float* p, float* m, float* v, const float* grad, bf16* p16, const uint8_t* mask,
                            int n, float lr, float b1, float b2, float eps, float bc1, float bc2, float wd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mask && !mask[i]) {
    if (p16) p16[i] = (bf16)0.f;
    return;
  }
  const float gi = grad[i] + wd * p[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float upd = lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  const float pi = p[i] - upd;
  p[i] = pi;
  if (p16) p16[i] = (bf16)pi;
}

// ---- device-resident training loop: every per-iteration scalar (learning
// rate, Adam bias corrections, the sampler's iteration) lives in device
// memory, so one iteration is a fixed kernel sequence that a HIP graph can
// replay.

// state: int [3] = {next iteration of the run, Adam step t, this iteration};
// sched: float [3] = {lr, 1 - b1^t, 1 - b2^t}.  Cosine decay from lr to
// lr_end over `iters` iterations.
__global__ void train_schedule_kernel(int* state, float* sched, int iters, float lr, float lr_end, float b1,
                                      float b2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int it = state[0], t = state[1] + 1;
  state[0] = it + 1;
  state[1] = t;
  state[2] = it;
  const float frac = (float)it / (float)(iters > 1 ? iters - 1 : 1);
  sched[0] = lr_end + 0.5f * (lr - lr_end) * (1.f + cosf(3.14159265358979f * frac));
  sched[1] = 1.f - powf(b1, (float)t);
  sched[2] = 1.f - powf(b2, (float)t);
}

__device__ __forceinline__ unsigned mix32(unsigned x) {
  x ^= x >> 16; x *= 0x7feb352du;
  x ^= x >> 15; x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Minibatch indices of the iteration in state[2]: sample b is (step, env) =
// (idx[2b], idx[2b+1]), step uniform in [lo, hi), env in [0, nenv), from a
// counter-based hash (multiply-shift to the range).
__global__ void sample_kernel(int B, int lo, int hi, int nenv, unsigned seed, const int* state, int* idx) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  unsigned h = mix32(seed ^ 0x2545f491u);
  h = mix32(h + (unsigned)state[2] * 0x9E3779B9u);
  h = mix32(h ^ ((unsigned)b * 0x85EBCA6Bu));
  const unsigned hs = mix32(h + 0xC2B2AE35u), he = mix32(h + 0x27D4EB2Fu);
  idx[2 * b] = lo + (int)(((unsigned long long)hs * (unsigned)(hi - lo)) >> 32);
  idx[2 * b + 1] = (int)(((unsigned long long)he * (unsigned)nenv) >> 32);
}

// Column sums without atomics: pass 1 writes one partial per block of
// COLSUM_ROWS rows, part[blk][j]; pass 2 adds the partials of column j in
// block order, so the result does not depend on scheduling.
constexpr int COLSUM_ROWS = 256;
__global__ __launch_bounds__(256) void colsum_part_kernel(const float* dZ, int B, int ld, float* part) {
  __shared__ float ps[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  const int r0 = blockIdx.y * COLSUM_ROWS, r1 = min(B, r0 + COLSUM_ROWS);
  float s = 0.f;
  if (j < ld) {
#pragma unroll 8
    for (int i = r0 + w; i < r1; i += 4) s += dZ[(size_t)i * ld + j];
  }
  ps[w][lane] = s;
  __syncthreads();
  if (w == 0 && j < ld) part[(size_t)blockIdx.y * ld + j] = (ps[0][lane] + ps[1][lane]) + (ps[2][lane] + ps[3][lane]);
}

__global__ void colsum_final_kernel(const float* part, int nblk, int ld, float* db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ld) return;
  float s = 0.f;
#pragma unroll 8
  for (int k = 0; k < nblk; ++k) s += part[(size_t)k * ld + j];
  db[j] = s;
}

// Adam over several tensors in one launch (the update of adam_kernel), with
// the learning rate and bias corrections read from sched.
constexpr int ADAM_MAX = 16;
constexpr int ADAM_WIDE = 16;    // more gradient slices than this: one wave per element
constexpr int ADAM_INFLIGHT = 8; // gradient slices whose loads are issued together (the rest one at a time)
struct AdamSet {
  int ntensor;
  int offset[ADAM_MAX + 1];      // first thread of each tensor's run (a multiple of 64)
  int n[ADAM_MAX];               // elements
  int nsum[ADAM_MAX];            // gradient = sum of nsum slices g + z * gstride, z ascending
  long long gstride[ADAM_MAX];
  float* p[ADAM_MAX];
  float* m[ADAM_MAX];
  float* v[ADAM_MAX];
  const float* g[ADAM_MAX];
  bf16* p16[ADAM_MAX];
  bf16* p16t[ADAM_MAX];          // transposed bf16 copy [cols][rows] of a [rows][cols] tensor, or null
  int cols[ADAM_MAX];
  const uint8_t* mask[ADAM_MAX];
  float wd[ADAM_MAX];
};

// the mask byte of a tensor without a mask
__device__ const uint8_t adam_unmasked = 1;

__global__ __launch_bounds__(256) void adam_multi_kernel(AdamSet s, const float* sched, float b1, float b2,
                                                         float eps) {
  // Every tensor's run of threads starts at a multiple of 64, so a wave
  // lies in one tensor: its index comes from the wave's first thread, and
  // the tensor's fields are scalar loads.
  const int i0 = __builtin_amdgcn_readfirstlane(blockIdx.x * blockDim.x + (threadIdx.x & ~63));
  if (i0 >= s.offset[s.ntensor]) return;
  int k = 0;
  while (k + 1 < s.ntensor && i0 >= s.offset[k + 1]) ++k;
  const int i = i0 + (threadIdx.x & 63);
  const int nk = s.nsum[k];
  const float* gk = s.g[k];
  const long long gs = s.gstride[k];
  const bool wide = nk > ADAM_WIDE;
  // a wide tensor has one wave per element (its run is whole waves long),
  // the others one thread
  const int t = i - s.offset[k];
  const int j = wide ? t >> 6 : t;
  if (j >= s.n[k]) return;                     // gap before the next wave-aligned run
  // the element's state and mask are loaded with its gradient slices (one
  // round trip)
  float* p = s.p[k];
  const uint8_t* mask = s.mask[k];
  const float pj = p[j], mj = s.m[k][j], vj = s.v[k][j];
  const uint8_t mv = *(mask ? mask + j : &adam_unmasked);     // no branch: the load stays in flight
  float gsum;
  if (wide) {
    // lane l sums slices l, l + 64, ... in order (two loads in flight per
    // trip), then a butterfly over the wave; lane 0 carries on
    const int l = t & 63;
    float acc = 0.f;
    for (int z = l; z < nk; z += 128) {
      const float g0 = gk[(size_t)z * gs + j];
      const float g1 = z + 64 < nk ? gk[(size_t)(z + 64) * gs + j] : 0.f;
      acc += g0;
      if (z + 64 < nk) acc += g1;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (l != 0) return;
    gsum = acc;
  } else {
    // the first ADAM_INFLIGHT slices are all loaded before the first add (z
    // ascending, as the one-at-a-time sum)
    float gz[ADAM_INFLIGHT];
#pragma unroll
    for (int z = 0; z < ADAM_INFLIGHT; ++z) gz[z] = z < nk ? gk[(size_t)z * gs + j] : 0.f;
    gsum = gz[0];
#pragma unroll
    for (int z = 1; z < ADAM_INFLIGHT; ++z)
      if (z < nk) gsum += gz[z];
    for (int z = ADAM_INFLIGHT; z < nk; ++z) gsum += gk[(size_t)z * gs + j];
  }
  bf16* p16 = s.p16[k];
  bf16* p16t = s.p16t[k];
  const int tj = p16t ? (j % s.cols[k]) * (s.n[k] / s.cols[k]) + j / s.cols[k] : 0;
  if (mv == 0) {
    if (p16) p16[j] = (bf16)0.f;
    if (p16t) p16t[tj] = (bf16)0.f;
    return;
  }
  const float lr = sched[0], bc1 = sched[1], bc2 = sched[2];
  // each fused multiply-add written out, so the rounding does not depend
  // on which product the compiler contracts
  const float gi = __builtin_fmaf(s.wd[k], pj, gsum);
  const float mi = __builtin_fmaf(b1, mj, (1.f - b1) * gi);
  const float vi = __builtin_fmaf((1.f - b2) * gi, gi, b2 * vj);
  s.m[k][j] = mi;
  s.v[k][j] = vi;
  const float upd = lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  const float pi = pj - upd;
  p[j] = pi;
  if (p16) p16[j] = (bf16)pi;
  if (p16t) p16t[tj] = (bf16)pi;
}

__global__ void fill_kernel(float* x, float v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void f32_to_bf16_kernel(const float* x, bf16* y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (bf16)x[i];
}

// ------------------------------------------------------------ features
// Sample b takes (step, env) = (idx[2b], idx[2b+1]) of a dataset of SoA blocks
// X[step][nfeat][nenv], Y[step][nout][nenv]; writes the z-scored features as
// bf16 row b of Xn [B][ldx] (zero padding) and the z-scored targets as fp32
// row b of Yn [B][ldy].  With idx == null, sample b is env b of step 0.
__global__ void gather_kernel(const float* X, const float* Y, const int* idx, int B, int nfeat, int nout,
                              int nenv, const float* xmean, const float* xinv, const float* ymean,
                              const float* yinv, bf16* Xn, int ldx, float* Yn, int ldy) {
  const int b = blockIdx.x;
  if (b >= B) return;
  const size_t step = idx ? (size_t)idx[2 * b] : 0;
  const size_t env = idx ? (size_t)idx[2 * b + 1] : (size_t)b;
  for (int f = threadIdx.x; f < ldx; f += blockDim.x) {
    float v = 0.f;
    if (f < nfeat) v = (X[(step * nfeat + f) * nenv + env] - xmean[f]) * xinv[f];
    Xn[(size_t)b * ldx + f] = (bf16)v;
  }
  if (Yn) {
    for (int f = threadIdx.x; f < ldy; f += blockDim.x) {
      float v = 0.f;
      if (f < nout) v = (Y[(step * nout + f) * nenv + env] - ymean[f]) * yinv[f];
      Yn[(size_t)b * ldy + f] = v;
    }
  }
}

// The same samples from a sample-major copy of the dataset, rows
// S[step][env][nfeat + nout] (features, then targets): one wave per sample
// reads its row with consecutive lanes on consecutive floats; the same
// arithmetic as gather_kernel, so the same bits.
__global__ __launch_bounds__(256) void gather_rows_kernel(const float* S, const int* idx, int B, int nfeat, int nout,
                                                          int nenv, const float* xmean, const float* xinv,
                                                          const float* ymean, const float* yinv, bf16* Xn, int ldx,
                                                          float* Yn, int ldy) {
  const int lane = threadIdx.x & 63, b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= B) return;
  const float* row = S + ((size_t)idx[2 * b] * nenv + idx[2 * b + 1]) * (size_t)(nfeat + nout);
  for (int f = lane; f < ldx; f += 64) {
    float v = 0.f;
    if (f < nfeat) v = (row[f] - xmean[f]) * xinv[f];
    Xn[(size_t)b * ldx + f] = (bf16)v;
  }
  for (int f = lane; f < ldy; f += 64) {
    float v = 0.f;
    if (f < nout) v = (row[nfeat + f] - ymean[f]) * yinv[f];
    Yn[(size_t)b * ldy + f] = v;
  }
}

// Network input of every env of a SoA state: z-scored (qpos 36, qvel 35,
// tau 35) as bf16 row e of Xn [nrows][ldx]; rows >= nenv and columns >= 106
// are zero.  One thread per row.
__global__ __launch_bounds__(256) void features_kernel(int nenv, int nrows, const float* qpos, const float* qvel,
                                                       const float* tau, const float* mean, const float* inv,
                                                       bf16* Xn, int ldx) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nrows) return;
  const size_t N = (size_t)nenv;
  bf16* row = Xn + (size_t)e * ldx;
  for (int f0 = 0; f0 < ldx; f0 += 8) {
    bf16x8 v;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int f = f0 + q;
      float x = 0.f;
      if (e < nenv && f < 106) {
        const float raw = f < 36 ? qpos[f * N + e] : f < 71 ? qvel[(f - 36) * N + e] : tau[(f - 71) * N + e];
        x = (raw - mean[f]) * inv[f];
      }
      v[q] = (bf16)x;
    }
    *reinterpret_cast<bf16x8*>(row + f0) = v;
  }
}

}  // namespace g1mlp

// ================================================================== C-ABI
namespace {
using namespace g1mlp;
// 128-tiles when they give at least two workgroups per CU of a 256-CU chip
int tile_of(const GemmArgs& a, int split) { return (a.I / 128) * (a.J / 128) * split >= 512 ? 128 : 64; }

template <bool AT, bool BT, int E>
int launch_gemm(const GemmArgs& a, int split, hipStream_t s) {
  if (tile_of(a, split) == 128) {
    hipLaunchKernelGGL((gemm_kernel<AT, BT, E, 128>), dim3(a.I / 128, a.J / 128, split), dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL((gemm_kernel<AT, BT, E, 64>), dim3(a.I / 64, a.J / 64, split), dim3(256), 0, s, a);
  }
  return g1err::launch("g1_mlp_gemm");
}
}  // namespace

extern "C" {

// C-ABI of the MLP kernels; the contract of each entry point is documented with its declaration.
// The arguments of one GEMM, checked, into *out (0), or an error (-1).
static int make_gemm(int epi, const void* A, int lda, const void* B, int ldb, int I, int J, int R, int split, float* C,
                     void* H, int ldc, const float* bias, const float* Zin, float* colpart, GemmArgs* out) {
  if (I <= 0 || J <= 0 || R <= 0 || split <= 0 || I % BI || J % BJ || R % (BR * split))
    return g1err::fail("g1_mlp_gemm: shape %d x %d x %d (split %d) is not a multiple of the %d x %d x %d tile",
                       I, J, R, split, BI, BJ, BR);
  if (!A || !B) return g1err::fail("g1_mlp_gemm: A and B are required");
  if (colpart && (epi != EPI_SILU_BWD || ldc < J))
    return g1err::fail("g1_mlp_gemm: column partial sums are an output of the SILU_BWD epilogue (ldc >= J)");
  const bool needs_c = epi != EPI_BIAS_SILU_H && !(epi == EPI_SILU_BWD && colpart);
  if ((needs_c && !C) || ((epi == EPI_BIAS_SILU || epi == EPI_BIAS_SILU_H || epi == EPI_SILU_BWD) && !H))
    return g1err::fail("g1_mlp_gemm: epilogue %d needs its C / H outputs", epi);
  if ((epi >= EPI_BIAS_SILU && epi <= EPI_BIAS_F32 && !bias) || (epi == EPI_SILU_BWD && !Zin))
    return g1err::fail("g1_mlp_gemm: epilogue %d needs bias / Zin", epi);
  if (split > 1 && epi != EPI_F32_ATOMIC && epi != EPI_F32_SPLIT)
    return g1err::fail("g1_mlp_gemm: split-R needs the F32_ATOMIC or F32_SPLIT epilogue");
  *out = GemmArgs{(const bf16*)A, (const bf16*)B, lda, ldb, I, J, R, R / split, C, (bf16*)H, ldc, bias, Zin, colpart};
  return 0;
}

int g1_mlp_gemm_ex(int a_trans, int b_trans, int epi, const void* A, int lda, const void* B, int ldb, int I, int J,
                   int R, int split, float* C, void* H, int ldc, const float* bias, const float* Zin, float* colpart,
                   void* stream) {
  g1err::clear();
  GemmArgs g;
  if (make_gemm(epi, A, lda, B, ldb, I, J, R, split, C, H, ldc, bias, Zin, colpart, &g)) return -1;
  hipStream_t s = (hipStream_t)stream;
  const int key = (a_trans ? 1 : 0) | (b_trans ? 2 : 0);
  switch (key * 8 + epi) {
    case 0 * 8 + EPI_BIAS_SILU: return launch_gemm<false, false, EPI_BIAS_SILU>(g, split, s);
    case 0 * 8 + EPI_BIAS_SILU_H: return launch_gemm<false, false, EPI_BIAS_SILU_H>(g, split, s);
    case 0 * 8 + EPI_BIAS_F32: return launch_gemm<false, false, EPI_BIAS_F32>(g, split, s);
    case 0 * 8 + EPI_F32: return launch_gemm<false, false, EPI_F32>(g, split, s);
    case 0 * 8 + EPI_SILU_BWD: return launch_gemm<false, false, EPI_SILU_BWD>(g, split, s);
    case 2 * 8 + EPI_SILU_BWD: return launch_gemm<false, true, EPI_SILU_BWD>(g, split, s);
    case 2 * 8 + EPI_F32: return launch_gemm<false, true, EPI_F32>(g, split, s);
    case 3 * 8 + EPI_F32_ATOMIC: return launch_gemm<true, true, EPI_F32_ATOMIC>(g, split, s);
    case 3 * 8 + EPI_F32_SPLIT: return launch_gemm<true, true, EPI_F32_SPLIT>(g, split, s);
    case 3 * 8 + EPI_F32: return launch_gemm<true, true, EPI_F32>(g, split, s);
    default: return g1err::fail("g1_mlp_gemm: invalid arguments");
  }
}

int g1_mlp_backward_layer(const void* d16, int rows, const void* hin, int cols, const void* WT16, int B, int split,
                          float* gW_split, void* dZ16, int ld_dz, const float* Z, float* colpart, void* stream) {
  g1err::clear();
  GemmArgs gw, gd;
  // weight gradient: dW = d^T h over the batch, in r-slices
  if (make_gemm(EPI_F32_SPLIT, d16, rows, hin, cols, rows, cols, B, split, gW_split, nullptr, cols, nullptr, nullptr,
                nullptr, &gw))
    return -1;
  // data gradient through SiLU': dZ = (d W) silu'(Z), with the bias-gradient column partials
  if (make_gemm(EPI_SILU_BWD, d16, rows, WT16, rows, B, cols, rows, 1, nullptr, dZ16, ld_dz, nullptr, Z, colpart, &gd))
    return -1;
  hipStream_t s = (hipStream_t)stream;
  const int tw = tile_of(gw, split), td = tile_of(gd, 1);
  if (tw != td) {
    // tile sizes differ: two launches
    const int rc = launch_gemm<true, true, EPI_F32_SPLIT>(gw, split, s);
    return rc ? rc : launch_gemm<false, false, EPI_SILU_BWD>(gd, 1, s);
  }
  const int nx0 = gw.I / tw, ny0 = gw.J / tw, n0 = nx0 * ny0 * split, nx1 = gd.I / td, ny1 = gd.J / td;
  const int n = n0 + nx1 * ny1;
  if (tw == 128)
    hipLaunchKernelGGL((gemm_pair_kernel<true, true, EPI_F32_SPLIT, false, false, EPI_SILU_BWD, 128>), dim3(n),
                       dim3(256), 0, s, gw, gd, nx0, ny0, n0, nx1, ny1);
  else
    hipLaunchKernelGGL((gemm_pair_kernel<true, true, EPI_F32_SPLIT, false, false, EPI_SILU_BWD, 64>), dim3(n),
                       dim3(256), 0, s, gw, gd, nx0, ny0, n0, nx1, ny1);
  return g1err::launch("g1_mlp_backward_layer");
}

int g1_mlp_gemm(int a_trans, int b_trans, int epi, const void* A, int lda, const void* B, int ldb, int I, int J,
                int R, int split, float* C, void* H, int ldc, const float* bias, const float* Zin, void* stream) {
  return g1_mlp_gemm_ex(a_trans, b_trans, epi, A, lda, B, ldb, I, J, R, split, C, H, ldc, bias, Zin, nullptr, stream);
}

int g1_mlp_mse_grad(const float* pred, const float* target, int B, int ld, int n_out, void* dY16, float* dY32,
                    float* sse, void* stream) {
  g1err::clear();
  if (B <= 0 || ld <= 0 || n_out > ld) return g1err::fail("g1_mlp_mse_grad: invalid arguments");
  const int n = B * ld;
  const int blocks = (n + 255) / 256 < MSE_BLOCKS ? (n + 255) / 256 : MSE_BLOCKS;
  hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, pred, target, B, ld, n_out,
                     (bf16*)dY16, dY32, sse);
  return g1err::launch("g1_mlp_mse_grad");
}

int g1_mlp_mse_grad_part(const float* pred, const float* target, int B, int ld, int n_out, void* dY16,
                         float* colpart, float* sse_part, void* stream) {
  g1err::clear();
  if (B <= 0 || B % 64 || ld <= 0 || ld % 32 || n_out < 0 || n_out > ld || !pred || !target || !dY16 || !colpart ||
      !sse_part)
    return g1err::fail("g1_mlp_mse_grad_part: invalid arguments (B a multiple of 64, ld of 32)");
  hipLaunchKernelGGL(mse_grad_part_kernel, dim3(B / 64, ld / 32), dim3(256), 0, (hipStream_t)stream, pred, target, ld,
                     n_out, 2.0f / (float)(B * n_out), (bf16*)dY16, colpart, sse_part);
  return g1err::launch("g1_mlp_mse_grad_part");
}

int g1_mlp_sum_splits(const float* parts, int split, long long n, float* out, void* stream) {
  g1err::clear();
  if (split <= 0 || n < 0 || !parts || !out) return g1err::fail("g1_mlp_sum_splits: invalid arguments");
  if (n == 0) return 0;
  const size_t n4 = ((size_t)n + 3) / 4;
  hipLaunchKernelGGL(sum_splits_kernel, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, (hipStream_t)stream, parts,
                     split, (size_t)n, out);
  return g1err::launch("g1_mlp_sum_splits");
}

int g1_mlp_colsum(const float* dZ, int B, int ld, float* db, void* stream) {
  g1err::clear();
  if (B <= 0 || ld <= 0) return g1err::fail("g1_mlp_colsum: invalid arguments");
  const int rpb = 64;
  hipLaunchKernelGGL(colsum_kernel, dim3((ld + 63) / 64, (B + rpb - 1) / rpb), dim3(256), 0, (hipStream_t)stream,
                     dZ, B, ld, rpb, db);
  return g1err::launch("g1_mlp_colsum");
}

int g1_mlp_colsum_det(const float* dZ, int B, int ld, float* part, float* db, void* stream) {
  g1err::clear();
  if (B <= 0 || ld <= 0 || !dZ || !part || !db) return g1err::fail("g1_mlp_colsum_det: invalid arguments");
  const int nblk = (B + COLSUM_ROWS - 1) / COLSUM_ROWS;
  hipLaunchKernelGGL(colsum_part_kernel, dim3((ld + 63) / 64, nblk), dim3(256), 0, (hipStream_t)stream, dZ, B, ld,
                     part);
  hipLaunchKernelGGL(colsum_final_kernel, dim3((ld + 255) / 256), dim3(256), 0, (hipStream_t)stream, part, nblk, ld,
                     db);
  return g1err::launch("g1_mlp_colsum_det");
}

int g1_mlp_train_schedule(int* state, float* sched, int iters, float lr, float lr_end, float b1, float b2,
                          void* stream) {
  g1err::clear();
  if (!state || !sched || iters <= 0) return g1err::fail("g1_mlp_train_schedule: invalid arguments");
  hipLaunchKernelGGL(train_schedule_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, state, sched, iters, lr,
                     lr_end, b1, b2);
  return g1err::launch("g1_mlp_train_schedule");
}

int g1_mlp_sample(int B, int lo, int hi, int nenv, unsigned seed, const int* state, int* idx, void* stream) {
  g1err::clear();
  if (B <= 0 || lo < 0 || hi <= lo || nenv <= 0 || !state || !idx)
    return g1err::fail("g1_mlp_sample: invalid arguments");
  hipLaunchKernelGGL(sample_kernel, dim3((B + 255) / 256), dim3(256), 0, (hipStream_t)stream, B, lo, hi, nenv, seed,
                     state, idx);
  return g1err::launch("g1_mlp_sample");
}

int g1_mlp_adam_multi_sum(int ntensor, float* const* p, float* const* m, float* const* v, const float* const* grad,
                          void* const* p16, void* const* p16t, const int* cols, const uint8_t* const* mask,
                          const int* n, const int* nsum, const long long* gstride, const float* wd,
                          const float* sched, float b1, float b2, float eps, void* stream) {
  g1err::clear();
  if (ntensor <= 0 || ntensor > ADAM_MAX || !p || !m || !v || !grad || !n || !sched)
    return g1err::fail("g1_mlp_adam_multi: invalid arguments (1 to %d tensors)", ADAM_MAX);
  AdamSet s{};
  s.ntensor = ntensor;
  long long total = 0;
  for (int k = 0; k < ntensor; ++k) {
    if (!p[k] || !m[k] || !v[k] || !grad[k] || n[k] <= 0)
      return g1err::fail("g1_mlp_adam_multi: tensor %d is missing or empty", k);
    s.n[k] = n[k];
    s.nsum[k] = nsum ? nsum[k] : 1;
    s.gstride[k] = gstride ? gstride[k] : 0;
    if (s.nsum[k] <= 0 || (s.nsum[k] > 1 && s.gstride[k] < n[k]))
      return g1err::fail("g1_mlp_adam_multi: tensor %d: %d gradient slices at stride %lld (>= %d elements)", k,
                         s.nsum[k], s.gstride[k], n[k]);
    const bool wide = s.nsum[k] > ADAM_WIDE;
    total = (total + 63) / 64 * 64;                 // every run starts a wave (a wide one: a wave per element)
    s.offset[k] = (int)total;
    total += wide ? 64LL * n[k] : (long long)n[k];
    s.p[k] = p[k]; s.m[k] = m[k]; s.v[k] = v[k]; s.g[k] = grad[k];
    s.p16[k] = p16 ? (bf16*)p16[k] : nullptr;
    s.p16t[k] = p16t ? (bf16*)p16t[k] : nullptr;
    s.cols[k] = cols ? cols[k] : 1;
    if (s.p16t[k] && (s.cols[k] <= 0 || n[k] % s.cols[k]))
      return g1err::fail("g1_mlp_adam_multi: tensor %d: %d elements are not rows x %d columns", k, n[k], s.cols[k]);
    s.mask[k] = mask ? mask[k] : nullptr;
    s.wd[k] = wd ? wd[k] : 0.f;
  }
  if (total > 0x7fffffffLL) return g1err::fail("g1_mlp_adam_multi: more than 2^31 elements");
  s.offset[ntensor] = (int)total;
  hipLaunchKernelGGL(adam_multi_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, s,
                     sched, b1, b2, eps);
  return g1err::launch("g1_mlp_adam_multi");
}

int g1_mlp_adam_multi(int ntensor, float* const* p, float* const* m, float* const* v, const float* const* grad,
                      void* const* p16, void* const* p16t, const int* cols, const uint8_t* const* mask,
                      const int* n, const float* wd, const float* sched, float b1, float b2, float eps,
                      void* stream) {
  return g1_mlp_adam_multi_sum(ntensor, p, m, v, grad, p16, p16t, cols, mask, n, nullptr, nullptr, wd, sched, b1, b2,
                               eps, stream);
}

int g1_mlp_adam(float* p, float* m, float* v, const float* grad, void* p16, const uint8_t* mask, int n, float lr,
                float b1, float b2, float eps, int t, float wd, void* stream) {
  g1err::clear();
  if (n <= 0 || t <= 0) return g1err::fail("g1_mlp_adam: invalid arguments");
  const float bc1 = 1.f - powf(b1, (float)t), bc2 = 1.f - powf(b2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, m, v, grad,
                     (bf16*)p16, mask, n, lr, b1, b2, eps, bc1, bc2, wd);
  return g1err::launch("g1_mlp_adam");
}

int g1_mlp_fill(float* x, float v, long long n, void* stream) {
  g1err::clear();
  if (n <= 0) return n == 0 ? 0 : -1;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, v,
                     (size_t)n);
  return g1err::launch("g1_mlp_fill");
}

int g1_mlp_to_bf16(const float* x, void* y, long long n, void* stream) {
  g1err::clear();
  if (n <= 0) return n == 0 ? 0 : -1;
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x,
                     (bf16*)y, (size_t)n);
  return g1err::launch("g1_mlp_to_bf16");
}

int g1_mlp_gather(const float* X, const float* Y, const int* idx, int B, int nfeat, int nout, int nenv,
                  const float* xmean, const float* xinv, const float* ymean, const float* yinv, void* Xn, int ldx,
                  float* Yn, int ldy, void* stream) {
  g1err::clear();
  if (B <= 0 || nfeat > ldx || (Yn && nout > ldy) || (!idx && B > nenv))
    return g1err::fail("g1_mlp_gather: invalid arguments");
  hipLaunchKernelGGL(gather_kernel, dim3(B), dim3(128), 0, (hipStream_t)stream, X, Y, idx, B, nfeat, nout, nenv,
                     xmean, xinv, ymean, yinv, (bf16*)Xn, ldx, Yn, ldy);
  return g1err::launch("g1_mlp_gather");
}

int g1_mlp_gather_rows(const float* S, const int* idx, int B, int nfeat, int nout, int nenv, const float* xmean,
                       const float* xinv, const float* ymean, const float* yinv, void* Xn, int ldx, float* Yn, int ldy,
                       void* stream) {
  g1err::clear();
  if (B <= 0 || nfeat <= 0 || nout <= 0 || nenv <= 0 || nfeat > ldx || nout > ldy || !S || !idx || !xmean || !xinv ||
      !ymean || !yinv || !Xn || !Yn)
    return g1err::fail("g1_mlp_gather_rows: invalid arguments");
  hipLaunchKernelGGL(gather_rows_kernel, dim3((B + 3) / 4), dim3(256), 0, (hipStream_t)stream, S, idx, B, nfeat, nout,
                     nenv, xmean, xinv, ymean, yinv, (bf16*)Xn, ldx, Yn, ldy);
  return g1err::launch("g1_mlp_gather_rows");
}

int g1_mlp_features(int nenv, int nrows, const float* qpos, const float* qvel, const float* tau, const float* mean,
                    const float* inv, void* Xn, int ldx, void* stream) {
  g1err::clear();
  if (nenv < 0 || nrows < nenv || ldx < 106 || ldx % 8 || !qpos || !qvel || !tau || !mean || !inv || !Xn)
    return g1err::fail("g1_mlp_features: invalid arguments");
  if (nrows == 0) return 0;
  hipLaunchKernelGGL(features_kernel, dim3((nrows + 255) / 256), dim3(256), 0, (hipStream_t)stream, nenv, nrows,
                     qpos, qvel, tau, mean, inv, (bf16*)Xn, ldx);
  return g1err::launch("g1_mlp_features");
}

}  // extern "C"
