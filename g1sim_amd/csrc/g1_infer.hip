// g1_infer.hip -- fused network inference for the step loop (gfx950).
//
// One workgroup of 8 wave64s takes 64 envs (128 from 32,768 envs on, in the
// forward-only surrogate and the policy) through a whole MLP: it builds the
// network input from the SoA state, runs every dense layer on
// v_mfma_f32_16x16x32_bf16 (bf16 operands, fp32 accumulation) with the
// activations held in LDS (two 64 x 528 bf16 buffers, or one 128 x 528
// buffer updated in place: 135 KB), and finishes with a fused epilogue, so
// no activation ever goes to HBM.  Weights are
// read from L2 (1.7 MB for the surrogate, 0.43 MB for the policy), each
// lane loading 16 contiguous bytes of a weight row as its B fragment, several
// k-steps ahead of its use; a layer's first k-steps are loaded before the
// previous layer's epilogue, so no layer starts on a cold weight fetch.
//
// * Surrogate step: z-scored (q, qd, tau_total) -> 512 x4 SiLU -> 35, then
//   the per-env commit (judge / fallback / semi-implicit Euler) of
//   g1_commit.h.  The forward-only variant writes the raw output instead.
// * Policy act: the RL observation (96- or 99-dim, src/g1sim.py:132-170),
//   normalised and clipped as examples/train.py:69-70, -> 512 -> 256 -> 128
//   (ELU) -> 29 (examples/train.py:21-29), an optional Gaussian sample, and
//   ctrl = default + action * scale (examples/train.py:299); the push of
//   examples/train.py:301-310 is applied to the state in the same kernel.
//
// Weights are stored k-tiled for these kernels ([NIN/32][NOUT][32] bf16; the
// host converts the row-major training copy).  Tiling: a layer of NOUT
// outputs has NOUT/16 column tiles of 16.  With 8 or
// more column tiles each wave owns NOUT/128 of them for all 4 row tiles
// (weight reuse x4 in registers); narrower layers split the 4 x NC tiles so
// every wave has NC/2 row tiles of one column tile.  The k-order of the
// accumulation is that of the tiled GEMM in g1_mlp.hip (32 at a time,
// ascending), so both give the same fp32 results.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <utility>

#include "g1_model_kernel.h"
#include "g1_fastmath.h"
#include "g1_commit.h"
#include "g1_act.h"
#include "g1_err.h"
#include "g1_launch.h"
#include "../../include/g1sim_amd.h"

namespace g1inf {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Workgroup barrier that orders LDS only: each wave's LDS accesses complete
// and become visible, while its global loads stay in flight.  __syncthreads()
// also waits for every outstanding global load, and the next layer's weight
// fragments are loaded just before the layer barriers, so it put a full L2
// latency into every layer.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// f(std::integral_constant<int, i>) for i = 0 .. N_ - 1, unrolled at compile time
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N_, typename F> __device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N_>{});
}

constexpr int ROWS = 64;                 // envs per workgroup of the 64-row kernels
// Activation row stride (bf16): 32 bytes of pad.  The A fragments are
// 16-byte reads, lane l at row l % 16 and k-offset (l / 16) * 8; with an
// 8-element (16-byte) pad the rows' bank offsets step by 4 dwords and lanes
// l and l + 16 of adjacent rows meet on the same banks, with 16 elements they
// step by 8 and never meet (measured on MI355X: the fragment reads of a k-loop
// 1.57x faster, a 512 x 512 layer's k-loop with L2 weights 5-10 % faster).
constexpr int LDA = 512 + 16;
constexpr int LDO = 65;                  // fp32 output row stride in LDS (odd: no bank conflicts)
constexpr int NV = 35, NU = 29;

enum Act : int { ACT_SILU = 0, ACT_ELU = 1 };

// Rows (envs) per workgroup, both with 8 waves (two per SIMD).  64: two
// activation buffers, layer l reads one and writes the other.  128: one
// buffer (128 x 528 bf16 = 135 KB) updated in place -- each layer's output
// stays in the accumulators until every wave has read the layer's input.
// Every weight fragment a wave loads then serves 8 row tiles instead of 4,
// which halves the L2 weight traffic per FLOP (measured on MI355X: the
// surrogate forward at 65,536 envs 156 -> 131 us, at 131,072 264 -> 248 us;
// the policy unchanged); it fills the 256 CUs from 32,768 envs.  (4 waves
// of 128 x 128 tiles, one per SIMD with 512 registers, measured slower:
// 159 and 294 us.)
template <int R> struct Rows {
  static_assert(R == 64 || R == 128, "64 or 128 rows per workgroup");
  static constexpr int RT = R / 16;              // row tiles
  static constexpr int NBUF = R == 64 ? 2 : 1;
  static constexpr int NW = 8;                   // waves per workgroup
  static constexpr int NT = 64 * NW;             // threads per workgroup
};

template <int NOUT, int RTILES> struct Tile {
  static constexpr int NWAVE = Rows<RTILES * 16>::NW;
  static constexpr int NC = NOUT / 16;
  static constexpr bool WIDE = NC >= NWAVE;
  static constexpr int CT = WIDE ? NC / NWAVE : 1;
  static constexpr int RT = WIDE ? RTILES : (RTILES * NC) / NWAVE;
  static_assert(NOUT % 16 == 0, "output width must be a multiple of 16");
  static_assert(WIDE ? NC % NWAVE == 0 : (NWAVE % NC == 0 && (RTILES * NC) % NWAVE == 0), "unsupported width");
  __device__ static int r0(int w) { return WIDE ? 0 : (w / NC) * RT; }
  __device__ static int c0(int w) { return WIDE ? w * CT : w % NC; }
};

// Weight fragments of one layer in flight: a ring of NB k-steps of the
// wave's CT column tiles, FR fragments in all (16 = 64 VGPRs with 64 rows, 8
// with 128 rows, whose accumulators take twice the registers), so a layer
// with fewer column tiles per wave runs further ahead.  The first PF k-steps are loaded by wprefetch() before the layer
// (the next layer's ring is filled before this layer's epilogue and
// barrier: weights do not depend on activations); the k-loop loads the rest
// PF steps ahead.
template <int NIN, int NOUT, int RTILES> struct Ring {
  using T = Tile<NOUT, RTILES>;
  static constexpr int KS = NIN / 32;
  static constexpr int FR = RTILES == 4 ? 16 : 8;
  static constexpr int NB0 = FR / T::CT < 2 ? 2 : FR / T::CT;
  static constexpr int NB = NB0 < KS ? NB0 : KS;
  static constexpr int PF = NB - 1;
  static_assert(NIN % 32 == 0 && KS >= 2 && PF >= 1, "input width must be a multiple of 32, at least 64");
  bf16x8 f[NB][T::CT];
};

// W is k-tiled, [NIN / 32][NOUT][32] bf16: the 32-wide k-slice of every
// output row is one contiguous block, so the fragments of a k-step (16 rows
// x 64 B per load instruction) are contiguous 1 KB runs (row-major
// [NOUT][NIN] puts them NIN * 2 bytes apart and measured at 0.6x the rate on
// MI355X).  Lane l reads row c * 16 + l % 16, k-elements (l / 16) * 8 + 0..7,
// with a buffer load: the layer in a buffer resource (scalar registers), the
// lane's byte offset within a k-slice as the vector offset and the k-slice as
// the scalar offset, so no load carries 64-bit address arithmetic.
template <int NIN, int NOUT, int RTILES>
__device__ __forceinline__ bf16x8 wload(const bf16* W, int w, int lane, int k, int c) {
  using T = Tile<NOUT, RTILES>;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)W, (short)0, NIN * NOUT * 2, 0x00020000);
  const int voff = (((T::c0(w) + c) * 16 + (lane & 15)) * 32 + (lane >> 4) * 8) * 2;
  return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, k * NOUT * 64, 0));
}

template <int NIN, int NOUT, int RTILES>
__device__ __forceinline__ void wprefetch(const bf16* __restrict__ W, int w, int lane, Ring<NIN, NOUT, RTILES>& rg) {
  using R = Ring<NIN, NOUT, RTILES>;
#pragma unroll
  for (int k = 0; k < R::PF; ++k)
#pragma unroll
    for (int c = 0; c < Tile<NOUT, RTILES>::CT; ++c)
      rg.f[k][c] = wload<NIN, NOUT, RTILES>(W, w, lane, k, c);
}

// acc = init + in[rows x NIN] (LDS) . W^T for this wave's tiles (init[c]:
// column tile c's start, the bias), with the first PF k-steps of weights
// already in `rg`.
//
// TR = false: acc[r][c][q] is out[row (r0 + r) * 16 + (lane / 16) * 4 + q]
// [col (c0 + c) * 16 + lane % 16]; TR = true (weights as the first MFMA
// operand, so the tile comes out transposed): acc[r][c][q] is
// out[row (r0 + r) * 16 + lane % 16][col (c0 + c) * 16 + (lane / 16) * 4 + q],
// four adjacent columns of one row per lane.
template <int NIN, int NOUT, int RTILES, bool TR = false>
__device__ __forceinline__ void mma_layer(const bf16* __restrict__ sin, const bf16* __restrict__ W, int w,
                                          int lane, Ring<NIN, NOUT, RTILES>& rg,
                                          f32x4 (&acc)[Tile<NOUT, RTILES>::RT][Tile<NOUT, RTILES>::CT],
                                          const f32x4 (&init)[Tile<NOUT, RTILES>::CT]) {
  using T = Tile<NOUT, RTILES>;
  using R = Ring<NIN, NOUT, RTILES>;
  constexpr int KS = R::KS, PF = R::PF, NB = R::NB;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  const int r0 = T::r0(w);
  const bf16* ap = sin + (r0 * 16 + lr) * LDA + lk;
  // Software pipeline: weight fragments (L2) PF k-steps ahead in the ring,
  // activation fragments (LDS) one step ahead; a scheduling barrier per
  // k-step keeps the compiler from sinking the prefetches next to their uses.
  bf16x8 fa[2][T::RT];
#pragma unroll
  for (int r = 0; r < T::RT; ++r) fa[0][r] = *reinterpret_cast<const bf16x8*>(ap + r * 16 * LDA);
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    if (k + PF < KS) {
#pragma unroll
      for (int c = 0; c < T::CT; ++c)
        rg.f[(k + PF) % NB][c] = wload<NIN, NOUT, RTILES>(W, w, lane, k + PF, c);
    }
    if (k + 1 < KS) {
#pragma unroll
      for (int r = 0; r < T::RT; ++r)
        fa[(k + 1) & 1][r] = *reinterpret_cast<const bf16x8*>(ap + r * 16 * LDA + (k + 1) * 32);
    }
#pragma unroll
    for (int r = 0; r < T::RT; ++r)
#pragma unroll
      for (int c = 0; c < T::CT; ++c) {
        const f32x4 cin = k == 0 ? init[c] : acc[r][c];
        acc[r][c] = TR ? __builtin_amdgcn_mfma_f32_16x16x32_bf16(rg.f[k % NB][c], fa[k & 1][r], cin, 0, 0, 0)
                       : __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[k & 1][r], rg.f[k % NB][c], cin, 0, 0, 0);
      }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// hidden layer: out = act(in W^T + b) as bf16 into LDS, the accumulators
// starting at the bias; the next layer's first weight k-steps (Wn) go into
// `next` before the epilogue.  With one buffer (sout == sin, 128 rows) a
// barrier separates the last read of the input from the first write of the
// output, and `b` is the kernel's LDS copy of the bias.
template <int NIN, int NOUT, int ACT, int NIN2, int NOUT2, int RTILES>
__device__ __forceinline__ void dense(const bf16* sin, bf16* sout, const bf16* __restrict__ W,
                                      const float* __restrict__ b, int w, int lane, Ring<NIN, NOUT, RTILES>& rg,
                                      const bf16* __restrict__ Wn, Ring<NIN2, NOUT2, RTILES>& next) {
  using T = Tile<NOUT, RTILES>;
  const int r0 = T::r0(w), c0 = T::c0(w), lr = lane & 15;
  typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
  // bias of the lane's four columns per column tile
  f32x4 bc[T::CT];
#pragma unroll
  for (int c = 0; c < T::CT; ++c) {
    const int col = (c0 + c) * 16 + (lane >> 4) * 4;
    if constexpr (RTILES == 8) {
      bc[c] = *reinterpret_cast<const f32x4*>(b + col);
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) bc[c][q] = b[col + q];
    }
  }
  f32x4 acc[T::RT][T::CT];
  mma_layer<NIN, NOUT, RTILES, true>(sin, W, w, lane, rg, acc, bc);
  __builtin_amdgcn_sched_barrier(0);
  wprefetch<NIN2, NOUT2, RTILES>(Wn, w, lane, next);
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (RTILES == 8) lds_barrier();        // in place: every wave has read the input
#pragma unroll
  for (int c = 0; c < T::CT; ++c) {
    // four adjacent columns of one row per lane: one 8-byte LDS store
    const int col = (c0 + c) * 16 + (lane >> 4) * 4;
#pragma unroll
    for (int r = 0; r < T::RT; ++r) {
      const int row = (r0 + r) * 16 + lr;
      bf16x4 h;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float z = acc[r][c][q];
        if constexpr (ACT == ACT_SILU) h[q] = (bf16)g1act::silu(z);
        else h[q] = (bf16)g1act::elu(z);
      }
      *reinterpret_cast<bf16x4*>(sout + row * LDA + col) = h;
    }
  }
}

// output layer: out = in W^T + b as fp32 into LDS [rows][LDO] (over the
// input buffer with 128 rows, after a barrier)
template <int NIN, int NOUT, int RTILES>
__device__ __forceinline__ void dense_out(const bf16* sin, float* sout, const bf16* __restrict__ W,
                                          const float* __restrict__ b, int w, int lane,
                                          Ring<NIN, NOUT, RTILES>& rg) {
  using T = Tile<NOUT, RTILES>;
  static_assert(NOUT <= LDO, "output tile wider than the LDS row");
  const int r0 = T::r0(w), c0 = T::c0(w), lr = lane & 15;
  f32x4 bc[T::CT];
#pragma unroll
  for (int c = 0; c < T::CT; ++c) {
    const float v = b[(c0 + c) * 16 + lr];
    bc[c] = f32x4{v, v, v, v};
  }
  f32x4 acc[T::RT][T::CT];
  mma_layer<NIN, NOUT, RTILES>(sin, W, w, lane, rg, acc, bc);
  if constexpr (RTILES == 8) lds_barrier();        // in place: every wave has read the input
#pragma unroll
  for (int c = 0; c < T::CT; ++c) {
    const int col = (c0 + c) * 16 + lr;
#pragma unroll
    for (int r = 0; r < T::RT; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) sout[((r0 + r) * 16 + (lane >> 4) * 4 + q) * LDO + col] = acc[r][c][q];
  }
}

// ------------------------------------------------------------- surrogate
struct SurrogateArgs {
  int nenv;
  const float* qpos;
  const float* qvel;
  const float* tau;
  const float* xmean;
  const float* xinv;
  const bf16* W[5];      // k-tiled [NIN/32][NOUT][32]: (128, 512), (512, 512) x3, (512, 64)
  const float* b[5];     // [512] x4, [>=64]
  // forward-only output (MODE 0)
  float* pred;
  int ldp;
  // commit (MODE 1)
  int phase, period;
  float tol, dt;
  const float* ymean;
  const float* ystd;
  const float* qacc_aba;
  uint8_t* fallback;
  unsigned* stats;
  float* qacc_out;
  float* qpos_rw;
  float* qvel_rw;
  // deferred judging (MODE 1): the blocks whose live envs are all in
  // fallback were committed by the judge + forces kernel
  // (g1_hip_judge_forces_commit), which also recorded the judged ones
  int committed;
  // batch judgement (MODE 2): the pending-record flags (the records are the
  // network input, qacc_aba their ABA qacc rows)
  uint8_t* sflag;
};

constexpr int SUR_NIN = 106, SUR_PIN = 128, SUR_H = 512, SUR_POUT = 64;
constexpr int SNAP_ROWS = SUR_NIN + 35;       // deferred-judgement record: inputs, ABA qacc
static_assert(ROWS == g1k::JUDGE_BLOCK, "a network tile is one judge block");

// ------------------------------------------ split-half chain (128 rows)
// The forward-only surrogate with 128 rows per workgroup runs its layers as
// one unrolled sequence of k-steps in which every SiLU epilogue overlaps
// MFMA work.  Each wave's four column tiles of a 512-wide layer are split
// into half A = tiles {2w, 2w + 1} (features 32w .. 32w + 31, all below 256)
// and half B = tiles {16 + 2w, 17 + 2w} (256 + 32w ..).  A hidden layer is
//   lo: half A, then half B, over k-steps 0-7 (features 0-255 of the
//       input), with the previous layer's half-B epilogue interleaved with
//       the half-A steps and stored into features 256-511 (nobody reads them
//       until the barrier; done by the end of the half-A steps, so no more
//       than two sets of accumulators are live);
//   barrier;
//   hi: half A, then half B, over k-steps 8-15, with this layer's half-A
//       epilogue interleaved and stored into features 0-255 (after the
//       barrier no wave reads them in this layer);
//   barrier.
// The first layer reads the network input from columns 256-383 (its
// half-A epilogue overlaps its half-B k-steps); the output layer's k-steps
// 0-7 overlap the last hidden layer's half-B epilogue.  Every output element
// still accumulates its k-steps in ascending order, so the results are those
// of the layer-by-layer kernel bit for bit.  (Measured on MI355X: the four
// SiLU epilogues took 36 % of a workgroup's time when no MFMA work ran
// beside them.)
namespace split {
constexpr int IN_COL = 256;                  // the network input: columns 256-383
constexpr int PF = 2, NB = PF + 1;           // weight fragments PF steps ahead
constexpr int NL1 = 8, NH = 32, NO = 16;     // steps of the first, each hidden and the output layer
constexpr int NG = NL1 + 3 * NH + NO;
constexpr int OUT0 = NL1 + 3 * NH;
struct Step { int layer, half, k; };
constexpr Step step(int g) {
  if (g < NL1) return Step{0, g / 4, g % 4};
  if (g < OUT0) {
    const int s = (g - NL1) % NH;
    return Step{1 + (g - NL1) / NH, (s / 8) % 2, s < 16 ? s % 8 : 8 + s % 8};
  }
  return Step{4, 0, g - OUT0};
}
constexpr int hstep(int g) { return (g - NL1) % NH; }   // step within a hidden layer
// The accumulators of a half-layer start at its bias (k-step 0 adds the
// first k-slice to it), as in the layer-by-layer GEMMs of training, so no
// epilogue adds a bias.  init_bias(g): the hidden half-layer whose k-step 0
// is step g, as layer * 2 + half, or -1.
constexpr int init_bias(int g) {
  if (g < 0 || g >= OUT0) return -1;
  const Step s = step(g);
  return s.k == 0 ? s.layer * 2 + s.half : -1;
}
// a barrier follows step g: the end of the first layer, the middle and end of
// each hidden layer, the middle and end of the output layer
constexpr bool barrier_after(int g) {
  return g == NL1 - 1 || (g >= NL1 && g < OUT0 && (hstep(g) == 15 || hstep(g) == 31)) || g == OUT0 + 7 ||
         g == OUT0 + NO - 1;
}

// Weight fragments are read with buffer loads: the layer's base in a buffer
// resource (scalar registers), the lane's byte offset within a k-slice (the
// same for every k-step and every hidden layer, five registers in all) and
// the k-step's slice offset as a scalar.  (Flat addresses, one 64-bit pair
// per layer, half, column tile and k-step in this unrolled chain, spilled.)
struct WSrc {
  __amdgpu_buffer_rsrc_t rs[5];
  int voff[2][2];          // hidden layers: half, column tile
  int voffo;               // output layer
};
__device__ __forceinline__ WSrc wsrc(const SurrogateArgs& a, int w, int lane) {
  WSrc r;
  constexpr int nin[5] = {SUR_PIN, SUR_H, SUR_H, SUR_H, SUR_H};
  constexpr int nout[5] = {SUR_H, SUR_H, SUR_H, SUR_H, SUR_POUT};
#pragma unroll
  for (int l = 0; l < 5; ++l)
    r.rs[l] = __builtin_amdgcn_make_buffer_rsrc((void*)a.W[l], (short)0, nin[l] * nout[l] * 2, 0x00020000);
  const int lb = (lane & 15) * 64 + (lane >> 4) * 16;     // row lane % 16 of the tile, 16 bytes per lane group
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int c = 0; c < 2; ++c) r.voff[h][c] = (h * 16 + 2 * w + c) * 16 * 64 + lb;
  r.voffo = (w & 3) * 16 * 64 + lb;
  return r;
}

template <int G> __device__ __forceinline__ void load_w(const WSrc& ws, bf16x8 (&wr)[NB][2]) {
  constexpr Step S = step(G);
  if constexpr (S.layer < 4) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
      wr[G % NB][c] = __builtin_bit_cast(
          bf16x8, __builtin_amdgcn_raw_buffer_load_b128(ws.rs[S.layer], ws.voff[S.half][c], S.k * SUR_H * 64, 0));
  } else {
    wr[G % NB][0] = __builtin_bit_cast(
        bf16x8, __builtin_amdgcn_raw_buffer_load_b128(ws.rs[4], ws.voffo, S.k * SUR_POUT * 64, 0));
  }
}

// The chain from the network input (in LDS, columns 256-383) to the fp32
// output tile in `out` ([128][LDO], over the activation buffer).  `wr` holds
// the weight fragments of steps 0 .. PF - 1 (prefetch()).
template <int UNUSED = 0>
__device__ __forceinline__ void chain(const SurrogateArgs& a, const WSrc& ws, bf16* buf, const float* sbias,
                                      float* out, int w, int lane, bf16x8 (&wr)[NB][2]) {
  typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
  const int lr = lane & 15, lk = (lane >> 4) * 8, lq = (lane >> 4) * 4;
  f32x4 accA[2][8][2], accB[2][8][2], accO[4];
  bf16x8 fa[2][8];
  const float bo = a.b[4][(w & 3) * 16 + lr];
  // Per-lane LDS byte offsets; every LDS address is one of them plus a
  // compile-time constant below 64 KB, which the ds instructions take as
  // their immediate offset: row tiles 0-3 and 4-7 each have their own base
  // (row tile 4 starts 67,584 bytes in), so a step computes no addresses.
  // The offsets are re-declared opaque at every step, so no address derived
  // from them stays live across steps (the compiler otherwise kept each (row
  // tile, column) address of the epilogue stores live through all five
  // layers, which reuse them, and spilled them).
  constexpr int RT4 = 4 * 16 * LDA * 2;                // bytes from row tile 0 to row tile 4
  constexpr int RTB = 16 * LDA * 2;                    // bytes per row tile
  int oa0 = (lr * LDA + lk) * 2, oa1 = oa0 + RT4;      // A fragments
  int oo = (((w >> 2) * 64 + lr) * LDA + lk) * 2;      // A fragments of the output layer
  int oe0 = (lr * LDA + 2 * w * 16 + lq) * 2, oe1 = oe0 + RT4;   // epilogue stores
  int ob = (2 * w * 16 + lq) * 4;                      // bias
  char* bb = reinterpret_cast<char*>(buf);
  const char* sb = reinterpret_cast<const char*>(sbias);
  // A fragments (activations) of step G into fa[G & 1]
  auto load_a = [&](auto GC) __attribute__((always_inline)) {
    constexpr int G = decltype(GC)::value;
    constexpr Step S = step(G);
    constexpr int col0 = (S.layer == 0 ? IN_COL : 0) + S.k * 32;
    if constexpr (S.layer < 4) {
#pragma unroll
      for (int r = 0; r < 8; ++r)
        fa[G & 1][r] = *reinterpret_cast<const bf16x8*>(bb + (r < 4 ? oa0 : oa1) + (r & 3) * RTB + col0 * 2);
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) fa[G & 1][r] = *reinterpret_cast<const bf16x8*>(bb + oo + r * RTB + col0 * 2);
    }
  };
  // The bias of the lane's four columns in each of its two column tiles, for
  // k-step 0 of a half-layer: read from LDS one step before (bsv[half]).
  f32x4 bsv[2][2];
  auto load_bias = [&](int lh) __attribute__((always_inline)) {
    const int l = lh >> 1, hb = lh & 1;
#pragma unroll
    for (int c = 0; c < 2; ++c)
      bsv[hb][c] = *reinterpret_cast<const f32x4*>(sb + ob + (l * SUR_H + hb * IN_COL + c * 16) * 4);
  };
  const f32x4 bov = f32x4{bo, bo, bo, bo};
  // SiLU epilogue of one 16 x 16 tile (row tile rt, column tile c of half hb:
  // the lane's four adjacent columns), stored as bf16 into the buffer
  auto epi = [&](const f32x4& acc, int rt, int c, int hb) __attribute__((always_inline)) {
    const int cc = hb * IN_COL + c * 16;
    bf16x4 h;
    h[0] = (bf16)g1act::silu(acc[0]);
    h[1] = (bf16)g1act::silu(acc[1]);
    h[2] = (bf16)g1act::silu(acc[2]);
    h[3] = (bf16)g1act::silu(acc[3]);
    *reinterpret_cast<bf16x4*>(bb + (rt < 4 ? oe0 : oe1) + (rt & 3) * RTB + cc * 2) = h;
  };
  load_bias(init_bias(0));
  load_a(std::integral_constant<int, 0>{});
  static_for<NG>([&](auto GC) {
    constexpr int G = decltype(GC)::value;
    constexpr Step S = step(G);
    asm volatile("" : "+v"(oa0), "+v"(oa1), "+v"(oo), "+v"(oe0), "+v"(oe1), "+v"(ob));
    if constexpr (G + PF < NG) load_w<G + PF>(ws, wr);
    if constexpr (G + 1 < NG && !barrier_after(G)) load_a(std::integral_constant<int, G + 1>{});
    // MFMAs of the step (k-step 0 starts from the bias)
    if constexpr (S.layer < 4) {
      auto mma = [&](f32x4 (&acc)[8][2]) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
          for (int c = 0; c < 2; ++c)
            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[G % NB][c], fa[G & 1][r],
                                                                S.k == 0 ? bsv[S.half][c] : acc[r][c], 0, 0, 0);
      };
      if constexpr (S.half == 0) mma(accA[S.layer & 1]);
      else mma(accB[S.layer & 1]);
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        accO[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[G & 1][r], wr[G % NB][0], S.k == 0 ? bov : accO[r], 0,
                                                          0, 0);
    }
    if constexpr (init_bias(G + 1) >= 0) load_bias(init_bias(G + 1));
    // the epilogue work interleaved with this step
    if constexpr (S.layer == 0 && S.half == 1) {
      // first layer, half A: row tiles 2 (G - 4) and 2 (G - 4) + 1
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int c = 0; c < 2; ++c) epi(accA[0][2 * (G - 4) + i][c], 2 * (G - 4) + i, c, 0);
    } else if constexpr (S.layer >= 1 && S.layer < 4 && hstep(G) < 8) {
      // the previous layer's half B: one row tile per half-A step
      constexpr int rt = hstep(G);
#pragma unroll
      for (int c = 0; c < 2; ++c) epi(accB[(S.layer - 1) & 1][rt][c], rt, c, 1);
    } else if constexpr (S.layer >= 1 && S.layer < 4 && hstep(G) >= 24) {
      // this layer's half A: one row tile per step
      constexpr int rt = hstep(G) - 24;
#pragma unroll
      for (int c = 0; c < 2; ++c) epi(accA[S.layer & 1][rt][c], rt, c, 0);
    } else if constexpr (S.layer == 4 && S.k < 8) {
      // the last hidden layer's half B: one row tile per step
#pragma unroll
      for (int c = 0; c < 2; ++c) epi(accB[3 & 1][S.k][c], S.k, c, 1);
    }
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (barrier_after(G)) {
      lds_barrier();
      if constexpr (G + 1 < NG) load_a(std::integral_constant<int, G + 1>{});
    }
  });
  // every wave has read the buffer: the fp32 output tile goes over it
  // (row tiles (w / 4) * 4 .. + 3, column tile w % 4)
  const int col = (w & 3) * 16 + lr;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) out[(((w >> 2) * 4 + r) * 16 + lq + q) * LDO + col] = accO[r][q];
}

template <int UNUSED = 0>
__device__ __forceinline__ void prefetch(const WSrc& ws, bf16x8 (&wr)[NB][2]) {
  static_for<PF>([&](auto GC) { load_w<decltype(GC)::value>(ws, wr); });
}
}  // namespace split

// ----------------------------- split-half chain, two buffers (64 rows)
// The same idea for 64 rows per workgroup (the surrogate below 32,768 envs
// and the stepper's network + commit kernel, where a workgroup is one judge
// block and the step time is one workgroup's latency).  Layer l reads buffer
// l % 2 and writes buffer (l + 1) % 2, so an epilogue can store as soon as it
// is computed.  A hidden layer is
//   A-lo (half A, k-steps 0-7), the previous layer's half-B epilogue
//        interleaved and stored into features 256-511 of this layer's input;
//   barrier;
//   A-hi (half A, k-steps 8-15);
//   B (half B, k-steps 0-15), this layer's half-A epilogue interleaved (one
//        tile every other step) and stored into the next layer's input;
//   barrier.
// The first layer runs A then B (its half-A epilogue in its B steps); the
// output layer's k-steps 0-7 overlap the last half-B epilogue.  k-steps
// still accumulate in ascending order: the results are those of the
// layer-by-layer kernel bit for bit.
namespace split2 {
using split::Step;
using split::NB;
using split::PF;
constexpr int NL1 = 8, NH = 32, NO = 16;
constexpr int NG = NL1 + 3 * NH + NO;
constexpr int OUT0 = NL1 + 3 * NH;
constexpr Step step(int g) {
  if (g < NL1) return Step{0, g / 4, g % 4};
  if (g < OUT0) {
    const int s = (g - NL1) % NH;
    return Step{1 + (g - NL1) / NH, s < 16 ? 0 : 1, s < 16 ? s : s - 16};
  }
  return Step{4, 0, g - OUT0};
}
constexpr int hstep(int g) { return (g - NL1) % NH; }
constexpr bool barrier_after(int g) {
  return g == NL1 - 1 || (g >= NL1 && g < OUT0 && (hstep(g) == 7 || hstep(g) == 31)) || g == OUT0 + 7 ||
         g == OUT0 + NO - 1;
}
// the hidden half-layer (layer * 2 + half) whose k-step 0 is step g, or -1:
// its accumulators start at its bias, as in split::chain
constexpr int init_bias(int g) {
  if (g < 0 || g >= OUT0) return -1;
  const Step s = step(g);
  return s.k == 0 ? s.layer * 2 + s.half : -1;
}
constexpr int BLA = 8;        // biases are read from global memory BLA steps ahead

template <int G> __device__ __forceinline__ void load_w(const split::WSrc& ws, bf16x8 (&wr)[NB][2]) {
  constexpr Step S = step(G);
  if constexpr (S.layer < 4) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
      wr[G % NB][c] = __builtin_bit_cast(
          bf16x8, __builtin_amdgcn_raw_buffer_load_b128(ws.rs[S.layer], ws.voff[S.half][c], S.k * SUR_H * 64, 0));
  } else {
    wr[G % NB][0] = __builtin_bit_cast(
        bf16x8, __builtin_amdgcn_raw_buffer_load_b128(ws.rs[4], ws.voffo, S.k * SUR_POUT * 64, 0));
  }
}

template <int UNUSED = 0>
__device__ __forceinline__ void prefetch(const split::WSrc& ws, bf16x8 (&wr)[NB][2]) {
  static_for<PF>([&](auto GC) { split2::load_w<decltype(GC)::value>(ws, wr); });
}

// From the network input (buffer 0, columns 0-127) to the fp32 output tile in
// `out` ([64][LDO], over buffer 1).  `wr` holds the weight fragments of steps
// 0 .. PF - 1 (prefetch()).
template <int UNUSED = 0>
__device__ __forceinline__ void chain(const SurrogateArgs& a, const split::WSrc& ws, bf16* buf0, bf16* buf1,
                                      float* out, int w, int lane, bf16x8 (&wr)[NB][2]) {
  typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
  const int lr = lane & 15, lk = (lane >> 4) * 8, lq = (lane >> 4) * 4;
  f32x4 accA[2][4][2], accB[2][4][2], accO[2];
  bf16x8 fa[2][4];
  const float bo = a.b[4][(w & 3) * 16 + lr];
  // per-lane element offsets, opaque at every step (as in split::chain)
  int oa = lr * LDA + lk;                              // A fragments
  int oo = ((w >> 2) * 32 + lr) * LDA + lk;            // A fragments of the output layer
  int oe = lr * LDA + 2 * w * 16 + lq;                 // epilogue stores
  int ob = 2 * w * 16 + lq;                            // bias
  auto bufl = [&](int l) -> bf16* { return (l & 1) ? buf1 : buf0; };
  auto load_a = [&](auto GC) __attribute__((always_inline)) {
    constexpr int G = decltype(GC)::value;
    constexpr Step S = step(G);
    const bf16* src = bufl(S.layer) + S.k * 32;
    if constexpr (S.layer < 4) {
#pragma unroll
      for (int r = 0; r < 4; ++r) fa[G & 1][r] = *reinterpret_cast<const bf16x8*>(src + oa + r * 16 * LDA);
    } else {
#pragma unroll
      for (int r = 0; r < 2; ++r) fa[G & 1][r] = *reinterpret_cast<const bf16x8*>(src + oo + r * 16 * LDA);
    }
  };
  // bias of the lane's columns for k-step 0 of a half-layer (bsv[half][tile]),
  // read BLA steps ahead
  f32x4 bsv[2][2];
  auto load_bias = [&](int lh) __attribute__((always_inline)) {
    const int l = lh >> 1, hb = lh & 1;
#pragma unroll
    for (int c = 0; c < 2; ++c) bsv[hb][c] = *reinterpret_cast<const f32x4*>(a.b[l] + ob + hb * 256 + c * 16);
  };
  const f32x4 bov = f32x4{bo, bo, bo, bo};
  // SiLU epilogue of one 16 x 16 tile of layer l (row tile rt, column tile c
  // of half hb), stored as bf16 into layer l's output buffer
  auto epi = [&](const f32x4& acc, int l, int rt, int c, int hb) __attribute__((always_inline)) {
    bf16x4 h;
    h[0] = (bf16)g1act::silu(acc[0]);
    h[1] = (bf16)g1act::silu(acc[1]);
    h[2] = (bf16)g1act::silu(acc[2]);
    h[3] = (bf16)g1act::silu(acc[3]);
    *reinterpret_cast<bf16x4*>(bufl(l + 1) + oe + rt * 16 * LDA + hb * 256 + c * 16) = h;
  };
  static_for<BLA>([&](auto GC) {
    constexpr int G = decltype(GC)::value;
    if constexpr (init_bias(G) >= 0) load_bias(init_bias(G));
  });
  load_a(std::integral_constant<int, 0>{});
  static_for<NG>([&](auto GC) {
    constexpr int G = decltype(GC)::value;
    constexpr Step S = step(G);
    asm volatile("" : "+v"(oa), "+v"(oo), "+v"(oe), "+v"(ob));
    if constexpr (G + PF < NG) split2::load_w<G + PF>(ws, wr);
    if constexpr (G + 1 < NG && !barrier_after(G)) load_a(std::integral_constant<int, G + 1>{});
    // MFMAs of the step (k-step 0 starts from the bias)
    if constexpr (S.layer < 4) {
      auto mma = [&](f32x4 (&acc)[4][2]) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 2; ++c)
            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[G % NB][c], fa[G & 1][r],
                                                                S.k == 0 ? bsv[S.half][c] : acc[r][c], 0, 0, 0);
      };
      if constexpr (S.half == 0) mma(accA[S.layer & 1]);
      else mma(accB[S.layer & 1]);
    } else {
#pragma unroll
      for (int r = 0; r < 2; ++r)
        accO[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[G & 1][r], wr[G % NB][0], S.k == 0 ? bov : accO[r], 0,
                                                          0, 0);
    }
    // (after the MFMAs: a half's next bias may replace the one this step used)
    if constexpr (init_bias(G + BLA) >= 0) load_bias(init_bias(G + BLA));
    // the epilogue work interleaved with this step
    if constexpr (S.layer == 0 && S.half == 1) {
      // first layer, half A: two tiles per step
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        constexpr int t0 = 2 * (G - 4);
        epi(accA[0][(t0 + i) / 2][(t0 + i) % 2], 0, (t0 + i) / 2, (t0 + i) % 2, 0);
      }
    } else if constexpr (S.layer >= 1 && S.layer < 4 && hstep(G) < 8) {
      // the previous layer's half B: one tile per A-lo step
      constexpr int t = hstep(G);
      epi(accB[(S.layer - 1) & 1][t / 2][t % 2], S.layer - 1, t / 2, t % 2, 1);
    } else if constexpr (S.layer >= 1 && S.layer < 4 && hstep(G) >= 16 && (hstep(G) & 1) == 0) {
      // this layer's half A: one tile every other B step
      constexpr int t = (hstep(G) - 16) / 2;
      epi(accA[S.layer & 1][t / 2][t % 2], S.layer, t / 2, t % 2, 0);
    } else if constexpr (S.layer == 4 && S.k < 8) {
      // the last hidden layer's half B: one tile per step
      epi(accB[3 & 1][S.k / 2][S.k % 2], 3, S.k / 2, S.k % 2, 1);
    }
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (barrier_after(G)) {
      lds_barrier();
      if constexpr (G + 1 < NG) load_a(std::integral_constant<int, G + 1>{});
    }
  });
  // every wave has read buffer 1 (last by the fourth hidden layer): the fp32
  // output tile goes over it (row tiles (w / 4) * 2 .. + 1, column tile w % 4)
  const int col = (w & 3) * 16 + lr;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) out[(((w >> 2) * 2 + r) * 16 + lq + q) * LDO + col] = accO[r][q];
}
}  // namespace split2

// LDS stash of the commit's inputs (MODE 1), each row [64 envs]: rows 0-70
// qpos and qvel as the feature stage read them, rows 71-105 ABA's qacc (of
// judged and fallback envs), row 106 the env's fallback flag on entry.
constexpr int STASH_ABA = 71, STASH_FB = STASH_ABA + NV, STASH_ROWS = STASH_FB + 1;

template <int MODE, int R>
__global__ __launch_bounds__(Rows<R>::NT) void surrogate_kernel(SurrogateArgs a) {
  constexpr int NTHR = Rows<R>::NT, NWAVE = Rows<R>::NW;
  static_assert(MODE != 1 || R == ROWS, "the step (commit) kernel works on 64-env judge blocks");
  constexpr int RT = Rows<R>::RT;
  __shared__ __attribute__((aligned(16))) bf16 buf[Rows<R>::NBUF][R * LDA];
  __shared__ float stash[MODE == 1 ? STASH_ROWS * ROWS : 1];
  __shared__ float ysc[MODE == 1 ? 2 * NV : 1];          // ystd, ymean
  // with 128 rows the layers run as the split-half chain (split::chain), the
  // network input in columns 256-383 of the buffer, and the hidden layers'
  // biases are staged in LDS ([4][512]); with 64 rows as the two-buffer chain
  // (split2::chain)
  constexpr bool SPLIT = R == 128;
  constexpr int FCOL = SPLIT ? split::IN_COL : 0;
  constexpr bool SBIAS = R == 128;
  __shared__ __attribute__((aligned(16))) float sbias[SBIAS ? 4 * SUR_H : 4];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int e0 = blockIdx.x * R;
  const size_t N = (size_t)a.nenv;
  // thread -> env r of the block and feature group fg: chunks of 8 features,
  // chunk c of the group holding features 8 (fg + G c) to 8 (fg + G c) + 7
  // (the 106 inputs spread evenly over the groups; each chunk goes to LDS in
  // one 16-byte store)
  constexpr int G = NTHR / R;
  const int r = tid % R, fg = tid / R, e = e0 + r, ec = e < a.nenv ? e : a.nenv - 1;
  // The tile's rows are exactly one judge block, so `judged` is uniform.
  const bool judged_blk = MODE == 1 && g1k::judged_env(e0, a.phase, a.period);
  // One batch of loads: the env's fallback flag, the network input (qpos 36,
  // qvel 35, tau 35) and, in a judged block, ABA's qacc of every env; then
  // the first layer's weights.  All are issued before any but the flag is
  // used, so the stage costs about one memory latency.  With 128 rows (two
  // groups of 64 features per thread) the features go in two passes, so the
  // first layer's weights and one pass's inputs fit the registers together.
  constexpr int FPW = SUR_PIN / G;
  constexpr int NPASS = R == 128 ? 2 : 1;
  constexpr int FPP = FPW / NPASS;                       // features per pass
  constexpr int APW = (NV + G - 1) / G;
  static_assert(SUR_PIN % G == 0 && FPW % NPASS == 0 && FPP % 8 == 0, "features split evenly over the groups");
  // the flag is loaded without a branch (from a valid dummy byte when there
  // are no flags) and tested after the other loads are issued
  const bool has_fb = MODE == 1 && a.fallback != nullptr;
  const uint8_t fbyte = *(has_fb ? a.fallback + ec : reinterpret_cast<const uint8_t*>(a.qpos));
  if constexpr (MODE == 1) {
    // deferred judging: a tile whose live envs are all in fallback has been
    // committed by the judge + forces kernel -- leave before any other load
    // (with every env in fallback, every tile)
    if (a.committed && !__syncthreads_or(e < a.nenv && fbyte == 0)) return;
  }
  auto feature_src = [&](int f) -> const float* {
    return f < 36 ? a.qpos + (size_t)f * N : f < 71 ? a.qvel + (size_t)(f - 36) * N : a.tau + (size_t)(f - 71) * N;
  };
  // The feature group is a compile-time constant in the loads and the
  // normalisation (one uniform branch picks the instance): each feature's
  // source row and its statistics' offsets are literals, the 8 statistics of
  // a chunk one scalar load.  (With fg known only at run time the stage took
  // 10k of a 128-env workgroup's 117k cycles.)
  float raw[FPP];
  auto load_pass = [&](auto P) __attribute__((always_inline)) {
    static_for<G>([&](auto FG) {
      if (fg == decltype(FG)::value) {
        static_for<FPP>([&](auto I) {
          constexpr int ii = decltype(P)::value * FPP + decltype(I)::value;
          constexpr int f = 8 * (decltype(FG)::value + G * (ii / 8)) + ii % 8;
          raw[decltype(I)::value] = f < SUR_NIN ? feature_src(f)[ec] : 0.f;
        });
      }
    });
  };
  load_pass(std::integral_constant<int, 0>{});
  // hidden-layer biases for the LDS copy (thread tid: element tid of each layer)
  float bstage[SBIAS ? 4 : 1];
  if constexpr (SBIAS) {
    static_assert(!SBIAS || NTHR == SUR_H, "one bias element per thread and layer");
#pragma unroll
    for (int l = 0; l < 4; ++l) bstage[l] = a.b[l][tid];
  }
  float abav[APW] = {};
  if constexpr (MODE == 1) {
    if (judged_blk) {
#pragma unroll
      for (int i = 0; i < APW; ++i) {
        const int j = fg + G * i;
        abav[i] = j < NV ? a.qacc_aba[(size_t)j * N + ec] : 0.f;
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  const bool fb = has_fb && e < a.nenv && fbyte != 0;
  if constexpr (MODE == 1) {
    // A tile that is not judged this step and whose live envs are all in
    // fallback takes ABA's qacc everywhere: commit without the network.  With
    // deferred judging (a.committed) every tile whose live envs are all in
    // fallback, judged or not, has been committed (and a judged one recorded)
    // by the judge + forces kernel: nothing to do.
    if (a.fallback && (!judged_blk || a.committed)) {
      const int own = e < a.nenv && !fb;
      if (!__syncthreads_or(own)) {
        if (a.committed) return;
        if (w == 0) {
          float sur[NV];
#pragma unroll
          for (int j = 0; j < NV; ++j) sur[j] = 0.f;
          g1k::surrogate_commit(e < a.nenv, e, N, a.phase, a.period, a.tol, a.dt, sur, a.qacc_aba, a.fallback,
                                a.stats, a.qacc_out, a.qpos_rw, a.qvel_rw);
        }
        return;
      }
      // fallback envs of a block that runs the network: their ABA qacc
#pragma unroll
      for (int i = 0; i < APW; ++i) {
        const int j = fg + G * i;
        abav[i] = fb && j < NV ? a.qacc_aba[(size_t)j * N + ec] : 0.f;
      }
    }
  }
  float ys = 0.f;
  if constexpr (MODE == 1) {
    if (tid < 2 * NV) ys = tid < NV ? a.ystd[tid] : a.ymean[tid - NV];
  }
  __builtin_amdgcn_sched_barrier(0);
  bf16x8 swr[split::NB][2];
  const split::WSrc sws = split::wsrc(a, w, lane);
  if constexpr (SPLIT) split::prefetch(sws, swr);
  else split2::prefetch(sws, swr);
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (MODE == 1) {
    // stash the commit's inputs
#pragma unroll
    for (int i = 0; i < APW; ++i) {
      const int j = fg + G * i;
      if (j < NV) stash[(STASH_ABA + j) * ROWS + r] = (judged_blk || fb) ? abav[i] : 0.f;
    }
    if (fg == 0) stash[STASH_FB * ROWS + r] = fb ? 1.f : 0.f;
    if (tid < 2 * NV) ysc[tid] = ys;
  }
  static_for<NPASS>([&](auto P) {
    constexpr int p = decltype(P)::value;
    if constexpr (p > 0) {
      load_pass(P);
      __builtin_amdgcn_sched_barrier(0);
    }
    static_for<G>([&](auto FG) {
      if (fg == decltype(FG)::value) {
        static_for<FPP / 8>([&](auto C8) {
          constexpr int i0 = 8 * decltype(C8)::value, ii0 = p * FPP + i0;
          constexpr int f0 = 8 * (decltype(FG)::value + G * (ii0 / 8));
          bf16x8 x8;
          static_for<8>([&](auto Q) {
            constexpr int q = decltype(Q)::value, f = f0 + q;
            float x = 0.f;
            if constexpr (f < SUR_NIN) {
              x = (raw[i0 + q] - a.xmean[f]) * a.xinv[f];
              if (e >= a.nenv) x = 0.f;
              if constexpr (MODE == 1 && f < STASH_ABA) stash[f * ROWS + r] = raw[i0 + q];   // qpos, qvel as read
            }
            x8[q] = (bf16)x;
          });
          // one 16-byte LDS store (a 2-byte store per feature at the 1,040-byte
          // row stride put 8 lanes on each bank)
          *reinterpret_cast<bf16x8*>(&buf[0][r * LDA + FCOL + f0]) = x8;
        });
      }
    });
    __builtin_amdgcn_sched_barrier(0);
  });
  if constexpr (SBIAS) {
#pragma unroll
    for (int l = 0; l < 4; ++l) sbias[l * SUR_H + tid] = bstage[l];
  }
  lds_barrier();
  // the fp32 output tile goes over the last buffer
  float* out = reinterpret_cast<float*>(buf[Rows<R>::NBUF - 1]);
  if constexpr (SPLIT) split::chain(a, sws, buf[0], sbias, out, w, lane, swr);
  else split2::chain(a, sws, buf[0], buf[1], out, w, lane, swr);
  lds_barrier();
  if constexpr (MODE == 0) {
    // [R][64] fp32 tile to pred rows; 512 threads
    for (int i = tid; i < R * SUR_POUT; i += NTHR) {
      const int rr = i / SUR_POUT, c = i % SUR_POUT, er = e0 + rr;
      if (er < a.nenv) a.pred[(size_t)er * a.ldp + c] = out[rr * LDO + c];
    }
  } else if constexpr (MODE == 2) {
    // The batch judgement of deferred judging: env e with a record of this
    // period (sflag[e]) is judged on its recorded state -- the network's
    // output against the recorded ABA qacc, the judge's measure and
    // tolerance (g1k::surrogate_judge) -- and leaves fallback if within it.
    // The env has stayed in fallback since it was recorded (every env is
    // judged once per period; a reset clears its record), so the network's
    // qacc is still used only by envs whose last judgement is at most
    // period - 1 steps old.  stats[4] / [5]: batch-judged / rejected envs,
    // [2] the max judged error.  One env per thread of the first R.
    if (tid < R) {
      const int er = e0 + tid;
      const bool rec = er < a.nenv && a.sflag[er] != 0;
      float err = 0.f;
      if (rec) {
        const float* orow = out + tid * LDO;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          const float sur = orow[j] * a.ystd[j] + a.ymean[j];
          const float ref = a.qacc_aba[(size_t)j * N + er];
          err = fmaxf(err, fabsf(sur - ref) * g1k::frcp(fmaxf(fabsf(ref), 1.0f)));
        }
      }
      const bool reject = rec && !(err <= a.tol);
      if (rec) {
        a.fallback[er] = reject ? 1 : 0;
        a.sflag[er] = 0;
      }
      const unsigned n_judged = (unsigned)__popcll(__ballot(rec));
      const unsigned n_reject = (unsigned)__popcll(__ballot(reject));
      unsigned ebits = rec ? __float_as_uint(err) : 0u;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = (unsigned)__shfl_xor((int)ebits, off, 64);
        ebits = o > ebits ? o : ebits;
      }
      if (lane == 0 && n_judged) {
        atomicAdd(&a.stats[4], n_judged);
        if (n_reject) atomicAdd(&a.stats[5], n_reject);
        atomicMax(&a.stats[2], ebits);
      }
    }
  } else {
    // The commit (g1_commit.h), spread over the waves: wave 0 judges and
    // counts (all dofs) and integrates the base, waves 1-7 the hinges.
    // Every piece evaluates the same expressions as surrogate_commit().
    const bool live = e < a.nenv;
    const bool judged = live && judged_blk;
    const float* st = stash + lane;
    const bool fbs = live && st[STASH_FB * ROWS] != 0.f;
    const bool use_aba = judged || fbs;
    const float* orow = out + lane * LDO;
    auto sur = [&](int j) { return orow[j] * ysc[j] + ysc[NV + j]; };
    auto aba = [&](int j) { return st[(STASH_ABA + j) * ROWS]; };
    auto qa = [&](int j) { return use_aba ? aba(j) : sur(j); };
    if (w == 0) {
      g1k::surrogate_judge(live, judged, fbs, e, a.tol, sur, aba, a.fallback, a.stats);
      if (live) {
        float q6[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) q6[j] = qa(j);
        if (a.qacc_out) {
#pragma unroll
          for (int j = 0; j < 6; ++j) a.qacc_out[j * N + e] = q6[j];
        }
        g1k::integrate_base(
            e, N, a.dt, [&](int j) { return q6[j]; }, [&](int m) { return st[m * ROWS]; },
            [&](int m) { return st[(36 + m) * ROWS]; }, a.qpos_rw, a.qvel_rw);
      }
    } else if (live) {
#pragma unroll
      for (int i = 0; i < (NU + NWAVE - 2) / (NWAVE - 1); ++i) {
        const int h = (w - 1) + (NWAVE - 1) * i;
        if (h < NU) {
          const float q = qa(6 + h);
          if (a.qacc_out) a.qacc_out[(6 + h) * N + e] = q;
          g1k::integrate_hinge(h, e, N, a.dt, q, st[(7 + h) * ROWS], st[(36 + 6 + h) * ROWS], a.qpos_rw, a.qvel_rw);
        }
      }
    }
  }
}

// ---------------------------------------------------------------- policy
// src/g1sim.py:76-82
constexpr float kActionScale[NU] = {
    0.443200f, 0.274054f, 0.443200f, 0.274054f, 0.365131f, 0.365131f,
    0.443200f, 0.274054f, 0.443200f, 0.274054f, 0.365131f, 0.365131f,
    0.443200f, 0.365131f, 0.365131f,
    0.365131f, 0.365131f, 0.365131f, 0.365131f, 0.365131f, 0.074501f, 0.074501f,
    0.365131f, 0.365131f, 0.365131f, 0.365131f, 0.365131f, 0.074501f, 0.074501f};

// Counter-based uniform draws for the rollout: stream k of env e at step t.
__device__ __forceinline__ unsigned mix32(unsigned x) {
  x ^= x >> 16; x *= 0x7feb352du;
  x ^= x >> 15; x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ unsigned draw_bits(unsigned seed, int t, int e, int k) {
  unsigned h = mix32(seed ^ 0x68bc21ebu);
  h = mix32(h + (unsigned)t * 0x9E3779B9u);
  h = mix32(h ^ ((unsigned)e * 0x85EBCA6Bu));
  return mix32(h + (unsigned)k * 0xC2B2AE35u);
}
// uniform in [0, 1) with 24 random bits
__device__ __forceinline__ float draw_u01(unsigned seed, int t, int e, int k) {
  return (float)(draw_bits(seed, t, e, k) >> 8) * 5.9604644775390625e-08f;
}

// v rotated by the inverse of the unit quaternion (w, x, y, z), written as
// src/g1sim.py:149-156 (u = -xyz; v + 2 (w (u x v) + u x (u x v)))
__device__ __forceinline__ void quat_rotate_inverse(float qw, float qx, float qy, float qz, const float* v,
                                                    float* o) {
  const float ux = -qx, uy = -qy, uz = -qz;
  const float a0 = uy * v[2] - uz * v[1], a1 = uz * v[0] - ux * v[2], a2 = ux * v[1] - uy * v[0];
  const float b0 = uy * a2 - uz * a1, b1 = uz * a0 - ux * a2, b2 = ux * a1 - uy * a0;
  o[0] = v[0] + 2.0f * (qw * a0 + b0);
  o[1] = v[1] + 2.0f * (qw * a1 + b1);
  o[2] = v[2] + 2.0f * (qw * a2 + b2);
}

// The raw per-env inputs of the observation, as rows: quaternion (4), hinge
// q (29), base linear velocity in the world frame (3), base angular velocity
// (3), hinge qd (29), commands (3), last ctrl (29).
constexpr int RAW_Q = 0, RAW_JP = 4, RAW_LV = 33, RAW_AV = 36, RAW_JV = 39, RAW_CMD = 68, RAW_LAST = 71,
              RAW_N = 100;

// Row `row` of the raw inputs in the SoA state ([dim][N]).
__device__ __forceinline__ const float* raw_row(int row, size_t N, const float* qpos, const float* qvel,
                                                const float* commands, const float* last_ctrl) {
  if (row < RAW_JP) return qpos + (size_t)(3 + row - RAW_Q) * N;
  if (row < RAW_LV) return qpos + (size_t)(7 + row - RAW_JP) * N;
  if (row < RAW_AV) return qvel + (size_t)(row - RAW_LV) * N;
  if (row < RAW_JV) return qvel + (size_t)(3 + row - RAW_AV) * N;
  if (row < RAW_CMD) return qvel + (size_t)(6 + row - RAW_JV) * N;
  if (row < RAW_LAST) return commands + (size_t)(row - RAW_CMD) * N;
  return last_ctrl + (size_t)(row - RAW_LAST) * N;
}

// Feature f of one env's observation from its raw inputs (raw(row) returns
// raw row `row`).  layout 0: the 99-dim G1Sim.get_obs order (commands,
// projected gravity, base linear velocity in the base frame, base angular
// velocity, q - default, qd, last ctrl); layout 1: the 96-dim order (base
// angular velocity, projected gravity, commands, q - default, qd, last ctrl).
template <typename Raw> __device__ __forceinline__ float obs_eval(int layout, int f, Raw raw) {
  int kind, i;   // 0 cmd, 1 gravity, 2 lin vel, 3 ang vel, 4 joint pos, 5 joint vel, 6 last ctrl
  if (layout == 0) {
    if (f < 12) { kind = f / 3 == 0 ? 0 : f / 3 == 1 ? 1 : f / 3 == 2 ? 2 : 3; i = f % 3; }
    else { kind = 4 + (f - 12) / NU; i = (f - 12) % NU; }
  } else {
    if (f < 9) { kind = f / 3 == 0 ? 3 : f / 3 == 1 ? 1 : 0; i = f % 3; }
    else { kind = 4 + (f - 9) / NU; i = (f - 9) % NU; }
  }
  switch (kind) {
    case 0: return raw(RAW_CMD + i);
    case 3: return raw(RAW_AV + i);
    case 4: return raw(RAW_JP + i) - (float)g1m_qpos_stand[7 + i];
    case 5: return raw(RAW_JV + i);
    case 6: return raw(RAW_LAST + i);
    default: break;
  }
  float v[3], o[3];
  if (kind == 1) { v[0] = 0.f; v[1] = 0.f; v[2] = -1.f; }
  else { v[0] = raw(RAW_LV); v[1] = raw(RAW_LV + 1); v[2] = raw(RAW_LV + 2); }
  quat_rotate_inverse(raw(RAW_Q), raw(RAW_Q + 1), raw(RAW_Q + 2), raw(RAW_Q + 3), v, o);
  return o[i];
}

// Feature f of env e read straight from the state.
__device__ float obs_feature(int layout, int f, size_t N, int e, const float* qpos, const float* qvel,
                             const float* commands, const float* last_ctrl) {
  return obs_eval(layout, f, [&](int row) { return raw_row(row, N, qpos, qvel, commands, last_ctrl)[e]; });
}

struct PolicyArgs {
  int nenv, layout, nobs;
  const float* qpos;
  float* qvel;                 // read for the observation; pushes add to it
  const float* commands;
  const float* last_ctrl;
  const float* obs_mean;
  const float* obs_inv;        // 1 / sqrt(var + 1e-8)
  float clip;
  const bf16* W[4];            // k-tiled: (128, 512), (512, 256), (256, 128), (128, 32)
  const float* b[4];
  const float* log_std;        // [29] or null (deterministic mean action)
  unsigned seed;
  int t;
  float* ctrl;                 // [29][N]
  float* actions;              // [29][N] or null
  float* obs_out;              // [N][nobs] or null
  int* push_interval;          // [N] or null (no pushes)
};

constexpr int POL_PIN = 128;

template <int R>
__global__ __launch_bounds__(Rows<R>::NT) void policy_kernel(PolicyArgs a) {
  constexpr int NTHR = Rows<R>::NT;
  constexpr int RT = Rows<R>::RT;
  __shared__ __attribute__((aligned(16))) bf16 buf[Rows<R>::NBUF][R * LDA];
  // with 128 rows the hidden layers' biases (512, 256, 128) are staged in LDS
  constexpr bool SBIAS = R == 128;
  constexpr int PB1 = 512, PB2 = PB1 + 256, PB3 = PB2 + 128;
  __shared__ __attribute__((aligned(16))) float sbias[SBIAS ? PB3 : 4];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int e0 = blockIdx.x * R;
  const size_t N = (size_t)a.nenv;
  // thread -> env r of the block and group g of G (raw rows g, g + G, ...;
  // chunks of 8 features, chunk c of the group holding features 8 (g + G c)
  // to 8 (g + G c) + 7, so the 96 or 99 features spread evenly over the
  // groups)
  constexpr int G = NTHR / R;
  const int r = tid % R, g = tid / R;
  // log std of the group's action pairs (j, j + 8), jb = g + G p: uniform,
  // held in scalar registers through the network
  constexpr int NPAIR = (16 + G - 1) / G;
  float lsd[2 * NPAIR];
#pragma unroll
  for (int p = 0; p < NPAIR; ++p) {
    const int jb = g + G * p, j = jb < 8 ? jb : jb + 8;
    const float l0 = a.log_std && jb < 16 && j < NU ? a.log_std[j] : 0.f;
    const float l1 = a.log_std && jb < 16 && j + 8 < NU ? a.log_std[j + 8] : 0.f;
    lsd[2 * p] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, l0)));
    lsd[2 * p + 1] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, l1)));
  }
  Ring<POL_PIN, 512, RT> ring0;
  {
    // stage the block's raw inputs in LDS, raw row q of env r at float
    // q * QS + r (the lanes of a wave, consecutive envs, hit consecutive
    // banks): with 64 rows in buf[1] (free until the first layer writes
    // it), with 128 rows in floats 64-191 of buffer row q, which the
    // features (floats 0-63 of each row) do not touch.  (An env-major stage,
    // env stride 260 floats, put 8 lanes on each bank: the stage and the
    // feature pass took 16k of a workgroup's 61k cycles.)  Every load of a
    // thread is issued before the first is used, so the stage costs one
    // memory latency.
    float* rawbase = Rows<R>::NBUF == 2 ? reinterpret_cast<float*>(buf[Rows<R>::NBUF - 1])
                                        : reinterpret_cast<float*>(buf[0]) + POL_PIN / 2;
    constexpr int QS = Rows<R>::NBUF == 2 ? R : LDA / 2;     // raw-row stride of the stage (floats)
    static_assert(Rows<R>::NBUF == 1 || RAW_N * R * sizeof(float) <= R * LDA * sizeof(bf16), "raw stage");
    static_assert(Rows<R>::NBUF == 2 || (RAW_N <= R && (POL_PIN / 2 + R) * sizeof(float) <= LDA * sizeof(bf16)),
                  "raw stage");
    auto rawat = [&](int row, int env) -> float& { return rawbase[row * QS + env]; };
    const int e = e0 + r, ec = e < a.nenv ? e : a.nenv - 1;
    constexpr int RPG = (RAW_N + G - 1) / G;           // rows per group
    constexpr int FPG = POL_PIN / G;                   // features per group
    static_assert(FPG % 8 == 0, "features go to LDS 8 at a time");
    float v[RPG], om[FPG], oi[FPG];
#pragma unroll
    for (int i = 0; i < RPG; ++i) {
      const int row = g + G * i;
      v[i] = row < RAW_N ? raw_row(row, N, a.qpos, a.qvel, a.commands, a.last_ctrl)[ec] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < FPG; ++i) {
      const int f = 8 * (g + G * (i / 8)) + i % 8;
      om[i] = f < a.nobs ? a.obs_mean[f] : 0.f;
      oi[i] = f < a.nobs ? a.obs_inv[f] : 0.f;
    }
    // hidden-layer biases for the LDS copy: element tid of layer 0, element
    // tid - 512 of the concatenated layers 1 and 2
    float bs0 = 0.f, bs1 = 0.f;
    if constexpr (SBIAS) {
      static_assert(!SBIAS || (NTHR == PB1 && PB3 - PB1 <= NTHR), "bias staging");
      bs0 = a.b[0][tid];
      if (tid < PB2 - PB1) bs1 = a.b[1][tid];
      else if (tid < PB3 - PB1) bs1 = a.b[2][tid - (PB2 - PB1)];
    }
    // the first layer's weights go last: the wait counter is in order
    __builtin_amdgcn_sched_barrier(0);
    wprefetch<POL_PIN, 512, RT>(a.W[0], w, lane, ring0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < RPG; ++i) {
      const int row = g + G * i;
      if (row < RAW_N) rawat(row, r) = v[i];
    }
    if constexpr (SBIAS) {
      sbias[tid] = bs0;
      if (tid < PB3 - PB1) sbias[PB1 + tid] = bs1;
    }
    lds_barrier();
    // The group's features with the layout and the group index as compile-time
    // constants (one uniform branch picks the instance): every feature's kind,
    // raw row and offset fold into its code.  (With both known only at run
    // time, the per-feature selection took 9k of a workgroup's 61k cycles.)
    auto features = [&](auto L, auto G0, auto OUT) __attribute__((always_inline)) {
      constexpr int LAY = decltype(L)::value, NOBS = LAY == 0 ? 99 : 96;
      static_for<FPG / 8>([&](auto C) {
        constexpr int c = decltype(C)::value, f0 = 8 * (decltype(G0)::value + G * c);
        bf16x8 x8;
        static_for<8>([&](auto Q) {
          constexpr int q = decltype(Q)::value, f = f0 + q;
          float x = 0.f;
          if constexpr (f < NOBS) {
            const float o = obs_eval(LAY, f, [&](int row) { return rawat(row, r); });
            if constexpr (decltype(OUT)::value) {
              if (e < a.nenv) a.obs_out[(size_t)e * NOBS + f] = o;
            }
            // (lanes past the last env hold a copy of its raw inputs: finite
            // rows whose outputs are never stored)
            x = fminf(fmaxf((o - om[8 * c + q]) * oi[8 * c + q], -a.clip), a.clip);
          }
          x8[q] = (bf16)x;
        });
        *reinterpret_cast<bf16x8*>(&buf[0][r * LDA + f0]) = x8;   // one 16-byte LDS store
      });
    };
    static_for<2>([&](auto L) {
      static_for<G>([&](auto G0) {
        if (a.layout == decltype(L)::value && g == decltype(G0)::value) {
          if (a.obs_out) features(L, G0, std::true_type{});
          else features(L, G0, std::false_type{});
        }
      });
    });
  }
  lds_barrier();
  if (a.push_interval && tid < R) {
    // examples/train.py:301-310: countdown; on expiry add a random base
    // velocity kick and draw the next interval from [50, 150)
    const int e = e0 + tid;
    if (e < a.nenv) {
      int pi = a.push_interval[e] - 1;
      if (pi <= 0) {
        const float half[6] = {0.5f, 0.5f, 0.4f, 0.52f, 0.52f, 0.78f};
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          const float u = draw_u01(a.seed, a.t, e, 64 + m);
          a.qvel[m * N + e] += u * (2.0f * half[m]) - half[m];
        }
        pi = 50 + (int)(draw_u01(a.seed, a.t, e, 70) * 100.0f);
      }
      a.push_interval[e] = pi;
    }
  }
  bf16* b0 = buf[0];
  bf16* b1 = buf[Rows<R>::NBUF - 1];
  Ring<512, 256, RT> ring1;
  Ring<256, 128, RT> ring2;
  Ring<128, 32, RT> ring3;
  const float* pb0 = SBIAS ? sbias : a.b[0];
  const float* pb1 = SBIAS ? sbias + PB1 : a.b[1];
  const float* pb2 = SBIAS ? sbias + PB2 : a.b[2];
  dense<POL_PIN, 512, ACT_ELU, 512, 256, RT>(b0, b1, a.W[0], pb0, w, lane, ring0, a.W[1], ring1);
  lds_barrier();
  dense<512, 256, ACT_ELU, 256, 128, RT>(b1, b0, a.W[1], pb1, w, lane, ring1, a.W[2], ring2);
  lds_barrier();
  dense<256, 128, ACT_ELU, 128, 32, RT>(b0, b1, a.W[2], pb2, w, lane, ring2, a.W[3], ring3);
  lds_barrier();
  float* out = reinterpret_cast<float*>(b0);
  dense_out<128, 32, RT>(b1, out, a.W[3], a.b[3], w, lane, ring3);
  lds_barrier();
  // epilogue: thread (r, g) -> env r; actions in pairs (j, j + 8) that share
  // one Box-Muller draw, j from {0..7, 16..23}, pair p of group g being
  // jb = g + G p.  The group index is a compile-time constant (one uniform
  // branch picks the instance), so the action scales and default angles are
  // instruction literals; the standard deviations were read with the stage.
  // (A loop over runtime j loaded them from memory on every iteration: 6.5k
  // of a workgroup's 55k cycles.)
  const int e = e0 + r;
  auto pairs = [&](auto G0) __attribute__((always_inline)) {
    static_for<NPAIR>([&](auto P) {
      constexpr int p = decltype(P)::value, jb = decltype(G0)::value + G * p;
      constexpr int j = jb < 8 ? jb : jb + 8, j2 = j + 8;
      if constexpr (jb < 16 && j < NU) {
        constexpr bool two = j2 < NU;
        float act = out[r * LDO + j];
        float act2 = two ? out[r * LDO + j2] : 0.f;
        if (a.log_std) {
          // Box-Muller: u1 in (0, 1], u2 in [0, 1); the cosine and sine
          // outputs are independent standard normals for actions j and j + 8.
          // v_log_f32 is log2 (u1 >= 2^-24: no denormal input); v_sin_f32 and
          // v_cos_f32 take their argument in revolutions, so sin(2 pi u2) is
          // one instruction with no range reduction (u2 in [0, 1))
          const float u1 = 1.0f - draw_u01(a.seed, a.t, e, 2 * j);
          const float u2 = draw_u01(a.seed, a.t, e, 2 * j + 1);
          const float rad = sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));   // -2 ln 2 log2(u1)
          const float sn = __builtin_amdgcn_sinf(u2), cs = __builtin_amdgcn_cosf(u2);
          act = fmaf(expf(lsd[2 * p]), rad * cs, act);
          if constexpr (two) act2 = fmaf(expf(lsd[2 * p + 1]), rad * sn, act2);
        }
        a.ctrl[j * N + e] = (float)g1m_qpos_stand[7 + j] + act * kActionScale[j];
        if (a.actions) a.actions[j * N + e] = act;
        if constexpr (two) {
          a.ctrl[j2 * N + e] = (float)g1m_qpos_stand[7 + j2] + act2 * kActionScale[j2];
          if (a.actions) a.actions[j2 * N + e] = act2;
        }
      }
    });
  };
  if (e < a.nenv) {
    static_for<G>([&](auto G0) {
      if (g == decltype(G0)::value) pairs(G0);
    });
  }
}

// ------------------------------------------------------------ observation
// G1Sim.get_obs as one kernel: obs [N][99] (src/g1sim.py:132-170).
__global__ __launch_bounds__(256) void obs_kernel(int nenv, int layout, int nobs, const float* qpos,
                                                  const float* qvel, const float* commands,
                                                  const float* last_ctrl, float* obs) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nenv * nobs) return;
  const int e = (int)(i / nobs), f = (int)(i % nobs);
  obs[i] = obs_feature(layout, f, (size_t)nenv, e, qpos, qvel, commands, last_ctrl);
}

// ------------------------------------------------- observation normaliser
// examples/train.py:49-66 (ActorCritic.update_obs_norm, called by act()
// before the observation is normalised): batch mean and biased variance over
// the envs, merged into the running statistics by Chan's formula.  Pass 1:
// workgroup (f, chunk) sums feature f over one chunk of envs about the
// running mean, in float64; pass 2 adds the chunks and merges.
constexpr int NORM_CHUNKS = 64;

__global__ __launch_bounds__(256) void obs_batch_stats_kernel(int nenv, int layout, int chunk, const float* qpos,
                                                              const float* qvel, const float* commands,
                                                              const float* last_ctrl, const float* run_mean,
                                                              const float* count, double* part) {
  __shared__ double s1[4], s2[4];
  const int f = blockIdx.x, ch = blockIdx.y;
  // the running count as it was before this update, for every merge block
  if (f == 0 && ch == 0 && threadIdx.x == 0) part[2 * NORM_CHUNKS * gridDim.x] = (double)count[0];
  const double c = run_mean[f];
  const int e0 = ch * chunk, e1 = min(nenv, e0 + chunk);
  double a = 0.0, b = 0.0;
  for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
    const double x = (double)obs_feature(layout, f, (size_t)nenv, e, qpos, qvel, commands, last_ctrl) - c;
    a += x;
    b += x * x;
  }
  // butterfly within each wave (no barriers), then the 4 wave sums in a
  // fixed order: the result does not depend on scheduling
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s1[w] = a;
    s2[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[(size_t)(f * NORM_CHUNKS + ch) * 2] = (s1[0] + s1[1]) + (s1[2] + s1[3]);
    part[(size_t)(f * NORM_CHUNKS + ch) * 2 + 1] = (s2[0] + s2[1]) + (s2[2] + s2[3]);
  }
}

// mean, var: [nobs] running statistics, count: [1] running count (all fp32,
// updated in place); inv = 1 / sqrt(var + 1e-8) for the policy kernel.  One
// wave per feature: lane c holds chunk c's sums, a butterfly adds them (a
// fixed order), lane 0 merges.
static_assert(NORM_CHUNKS == 64, "one chunk per lane");
__global__ __launch_bounds__(64) void obs_norm_merge_kernel(int nenv, int nobs, int nchunks, const double* part,
                                                            float* mean, float* var, float* count, float* inv) {
  const int f = blockIdx.x, lane = threadIdx.x;
  const double n = (double)nenv, cnt = part[2 * NORM_CHUNKS * nobs], tot = cnt + n;
  double s1 = lane < nchunks ? part[(size_t)(f * NORM_CHUNKS + lane) * 2] : 0.0;
  double s2 = lane < nchunks ? part[(size_t)(f * NORM_CHUNKS + lane) * 2 + 1] : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  if (lane == 0) {
    const double m = mean[f], v = var[f];
    const double d = s1 / n;                                   // batch mean - running mean
    const double bvar = fmax(s2 / n - d * d, 0.0);             // biased batch variance
    const double nm = m + d * n / tot;
    const double M2 = v * cnt + bvar * n + d * d * cnt * n / tot;
    const double nv = M2 / tot;
    mean[f] = (float)nm;
    var[f] = (float)nv;
    inv[f] = (float)(1.0 / sqrt((double)(float)nv + 1e-8));
    if (f == 0) count[0] = (float)tot;
  }
}

// ---------------------------------------------------------------- rollout
// examples/train.py:256-270: per-joint posture tolerances while walking / running
__constant__ float kStdWalking[NU] = {
    0.3f, 0.15f, 0.15f, 0.35f, 0.25f, 0.1f, 0.3f, 0.15f, 0.15f, 0.35f, 0.25f, 0.1f, 0.2f, 0.08f, 0.1f,
    0.15f, 0.15f, 0.1f, 0.15f, 0.3f, 0.3f, 0.3f, 0.15f, 0.15f, 0.1f, 0.15f, 0.3f, 0.3f, 0.3f};
__constant__ float kStdRunning[NU] = {
    0.5f, 0.2f, 0.2f, 0.6f, 0.35f, 0.15f, 0.5f, 0.2f, 0.2f, 0.6f, 0.35f, 0.15f, 0.3f, 0.08f, 0.2f,
    0.5f, 0.2f, 0.15f, 0.35f, 0.3f, 0.3f, 0.3f, 0.5f, 0.2f, 0.15f, 0.35f, 0.3f, 0.3f, 0.3f};

// draw streams of the bookkeeping kernels (the policy kernel uses 0..70)
constexpr int K_RESET = 128, K_TIMEOUT = 136, K_INIT = 144;

// examples/train.py:227-233 with the ranges as arguments (the reference
// uses lin 1.0, yaw 0.5)
__device__ __forceinline__ void resample_command(float* commands, size_t N, int e, unsigned seed, int t, int k,
                                                 float lin, float yaw) {
  commands[e] = draw_u01(seed, t, e, k) * (2.0f * lin) - lin;
  commands[N + e] = draw_u01(seed, t, e, k + 1) * (2.0f * lin) - lin;
  commands[2 * N + e] = draw_u01(seed, t, e, k + 2) * (2.0f * yaw) - yaw;
}
// integer in [lo, lo + n)
__device__ __forceinline__ int draw_int(unsigned seed, int t, int e, int k, int lo, int n) {
  return lo + (int)(draw_u01(seed, t, e, k) * (float)n);
}

struct PostArgs {
  int nenv;
  const float* qpos;
  const float* qvel;
  const float* foot_pos;
  const float* foot_vel;
  const float* contact_forces;
  const float* actions;
  float* commands;
  float* prev_actions;
  int* command_timeout;
  int* episode_length;
  int* push_interval;
  int max_episode_length;
  float cmd_lin, cmd_yaw;
  unsigned seed;
  int t;
  float* reward;
  uint8_t* done;
  uint8_t* reset;
  double* reward_sum;    // += sum of the rewards (nullable)
  double* done_count;    // += number of terminated envs (nullable)
};

// Reward and termination of examples/train.py:143-225 on the post-step state,
// then the per-step bookkeeping of examples/train.py:319-351: previous
// actions, command timer, episode length, timeouts, the reset mask and, for
// reset envs, new commands / timers / push interval.
__device__ void post_step(const PostArgs& a, int e, float& rew, bool& died);

__global__ __launch_bounds__(256) void post_step_kernel(PostArgs a) {
  __shared__ float part_r[4];
  __shared__ int part_d[4];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  float rew = 0.f;
  bool died = false;
  if (e < a.nenv) post_step(a, e, rew, died);
  if (!a.reward_sum && !a.done_count) return;              // uniform
  // the rollout's per-step sums: one double atomic per workgroup and counter
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) rew += __shfl_xor(rew, off, 64);
  const int nd = __popcll(__ballot(died));
  if ((threadIdx.x & 63) == 0) {
    part_r[threadIdx.x >> 6] = rew;
    part_d[threadIdx.x >> 6] = nd;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (a.reward_sum) atomicAdd(a.reward_sum, (double)((part_r[0] + part_r[1]) + (part_r[2] + part_r[3])));
    if (a.done_count) atomicAdd(a.done_count, (double)(part_d[0] + part_d[1] + part_d[2] + part_d[3]));
  }
}

__device__ void post_step(const PostArgs& a, int e, float& rew, bool& died) {
  const size_t N = (size_t)a.nenv;
  // Every input of the env is loaded before the first is used: one memory
  // latency instead of one per dependent group (at one thread per env,
  // 65,536 envs are one wave per SIMD, so nothing else hides it).
  float q[NU], act[NU], prev[NU], fv[6], cf[8];
#pragma unroll
  for (int j = 0; j < NU; ++j) {
    q[j] = a.qpos[(7 + j) * N + e];
    act[j] = a.actions[j * N + e];
    prev[j] = a.prev_actions[j * N + e];
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) fv[m] = a.foot_vel[m * N + e];
#pragma unroll
  for (int m = 0; m < 8; ++m) cf[m] = a.contact_forces[m * N + e];
  const float lz = a.foot_pos[2 * N + e], rz = a.foot_pos[5 * N + e];
  const int ct0 = a.command_timeout[e], el0 = a.episode_length[e];
  const float c0 = a.commands[e], c1 = a.commands[N + e], c2 = a.commands[2 * N + e];
  const float qw = a.qpos[3 * N + e], qx = a.qpos[4 * N + e], qy = a.qpos[5 * N + e], qz = a.qpos[6 * N + e];
  const float down[3] = {0.f, 0.f, -1.f};
  const float vw[3] = {a.qvel[e], a.qvel[N + e], a.qvel[2 * N + e]};
  float g[3], v[3];
  quat_rotate_inverse(qw, qx, qy, qz, down, g);
  quat_rotate_inverse(qw, qx, qy, qz, vw, v);
  const float w0 = a.qvel[3 * N + e], w1 = a.qvel[4 * N + e], w2 = a.qvel[5 * N + e];
  __builtin_amdgcn_sched_barrier(0);
  // 1. velocity tracking
  const float lin_err = (c0 - v[0]) * (c0 - v[0]) + (c1 - v[1]) * (c1 - v[1]) + v[2] * v[2];
  const float rew_lin = expf(-lin_err / 0.25f);
  const float ang_err = (c2 - w2) * (c2 - w2) + (w0 * w0 + w1 * w1);
  const float rew_ang = expf(-ang_err / 0.5f);
  // 2. orientation and speed-dependent posture
  const float rew_orient = expf(-(g[0] * g[0] + g[1] * g[1]) / 0.2f);
  const float speed = sqrtf(c0 * c0 + c1 * c1) + fabsf(c2);
  const bool standing = speed < 0.05f, running = speed >= 1.5f;
  float pose = 0.f, dof = 0.f;
#pragma unroll
  for (int j = 0; j < NU; ++j) {
    const float err = q[j] - (float)g1m_qpos_stand[7 + j];
    const float sd = standing ? 0.05f : running ? kStdRunning[j] : kStdWalking[j];
    pose += err * err / (sd * sd);
    const float lo = (float)g1m_jnt_range[2 * j], hi = (float)g1m_jnt_range[2 * j + 1];
    const float range = hi - lo;
    const float slo = lo + 0.05f * range, shi = hi - 0.05f * range;
    dof += fmaxf(q[j] - shi, 0.f) + fmaxf(slo - q[j], 0.f);
  }
  const float rew_pose = expf(-(pose / (float)NU));
  // 3. penalties
  float rate = 0.f;
#pragma unroll
  for (int j = 0; j < NU; ++j) {
    const float d = act[j] - prev[j];
    rate += d * d;
  }
  const float body_ang = w0 * w0 + w1 * w1;
  // 4. feet
  const float lv2 = fv[0] * fv[0] + fv[1] * fv[1];
  const float rv2 = fv[3] * fv[3] + fv[4] * fv[4];
  const float active = speed > 0.05f ? 1.f : 0.f;
  const float clearance =
      (fabsf(lz - 0.1f) * sqrtf(lv2 + 1e-6f) + fabsf(rz - 0.1f) * sqrtf(rv2 + 1e-6f)) * active;
  const float lf = cf[0] + cf[1] + cf[2] + cf[3];
  const float rf = cf[4] + cf[5] + cf[6] + cf[7];
  const float slip = (lv2 * (lf > 1.f ? 1.f : 0.f) + rv2 * (rf > 1.f ? 1.f : 0.f)) * active;
  rew = 2.0f * rew_lin + 2.0f * rew_ang + rew_orient + rew_pose - dof - 2.0f * clearance - 0.5f * slip -
        0.05f * rate - 0.05f * body_ang;
  a.reward[e] = rew;
  // termination: tilt beyond 70 degrees (the reference's 3.14159)
  const float tilt = fabsf(acosf(fminf(fmaxf(-g[2], -1.f), 1.f)));
  const bool done = tilt > (float)(70.0 * 3.14159 / 180.0);
  a.done[e] = done ? 1 : 0;
  died = done;
  // bookkeeping
  int ct = ct0 - 1;
  int el = el0 + 1;
  const bool rs = done || el >= a.max_episode_length;
#pragma unroll
  for (int j = 0; j < NU; ++j) a.prev_actions[j * N + e] = rs ? 0.f : act[j];
  if (rs) {
    resample_command(a.commands, N, e, a.seed, a.t, K_RESET, a.cmd_lin, a.cmd_yaw);
    ct = draw_int(a.seed, a.t, e, K_RESET + 3, 150, 100);
    el = 0;
    if (a.push_interval) a.push_interval[e] = draw_int(a.seed, a.t, e, K_RESET + 4, 50, 100);
  }
  if (ct <= 0) {
    resample_command(a.commands, N, e, a.seed, a.t, K_TIMEOUT, a.cmd_lin, a.cmd_yaw);
    ct = draw_int(a.seed, a.t, e, K_TIMEOUT + 3, 150, 100);
  }
  a.command_timeout[e] = ct;
  a.episode_length[e] = el;
  a.reset[e] = rs ? 1 : 0;
}

// Rollout start (examples/train.py:237-276): commands for every env, timers
// at zero, push intervals in [50, 150).
__global__ __launch_bounds__(256) void rollout_init_kernel(int nenv, float lin, float yaw, unsigned seed,
                                                           float* commands, int* command_timeout,
                                                           int* episode_length, int* push_interval) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nenv) return;
  resample_command(commands, (size_t)nenv, e, seed, -1, K_INIT, lin, yaw);
  command_timeout[e] = 0;
  episode_length[e] = 0;
  if (push_interval) push_interval[e] = draw_int(seed, -1, e, K_INIT + 3, 50, 100);
}

}  // namespace g1inf

// ================================================================== C-ABI
namespace {
using namespace g1inf;

// From 32,768 envs the forward-only surrogate and the policy kernels take
// 128 envs per workgroup (256+ workgroups: every CU busy); below, 64.
constexpr int kRows128MinEnvs = 32768;

bool fill_surrogate_weights(SurrogateArgs& a, const void* const* W, const float* const* b) {
  if (!W || !b) return false;
  for (int l = 0; l < 5; ++l) {
    if (!W[l] || !b[l]) return false;
    a.W[l] = (const bf16*)W[l];
    a.b[l] = b[l];
  }
  return true;
}
}  // namespace

// Envs off fallback (fallback[e] == 0) into *count: one workgroup, 16 flags
// per thread per pass over the first nvec * 16 flags (nvec = 0 unless the
// array is 16-byte aligned), the rest bytewise.  A kernel, not a memset plus
// atomics, so that it can sit in a captured graph.
__global__ __launch_bounds__(1024) void count_off_fallback_kernel(int nenv, int nvec, const uint8_t* fallback,
                                                                    unsigned* count) {
  __shared__ unsigned part[16];
  const int tid = threadIdx.x;
  unsigned c = 0;
  for (int i = tid; i < nvec; i += 1024) {
    const uint4 v = reinterpret_cast<const uint4*>(fallback)[i];
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j) c += ((w[k] >> (8 * j)) & 0xffu) == 0u;
  }
  for (int e = nvec * 16 + tid; e < nenv; e += 1024) c += fallback[e] == 0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((tid & 63) == 0) part[tid >> 6] = c;
  __syncthreads();
  if (tid == 0) {
    unsigned t = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) t += part[k];
    *count = t;
  }
}

extern "C" {

int g1_hip_surrogate_forward(int nenv, const float* qpos, const float* qvel, const float* tau, const float* xmean,
                             const float* xinv, const void* const* W, const float* const* b, float* pred, int ldp,
                             void* stream) {
  g1err::clear();
  if (nenv < 0 || !qpos || !qvel || !tau || !xmean || !xinv || !pred || ldp < SUR_POUT)
    return g1err::fail("g1_hip_surrogate_forward: invalid arguments");
  SurrogateArgs a{};
  if (!fill_surrogate_weights(a, W, b)) return g1err::fail("g1_hip_surrogate_forward: invalid arguments");
  if (nenv == 0) return 0;
  a.nenv = nenv; a.qpos = qpos; a.qvel = qvel; a.tau = tau; a.xmean = xmean; a.xinv = xinv;
  a.pred = pred; a.ldp = ldp;
  if (nenv >= kRows128MinEnvs)
    hipLaunchKernelGGL((surrogate_kernel<0, 128>), dim3((nenv + 127) / 128), dim3(Rows<128>::NT), 0,
                       (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL((surrogate_kernel<0, 64>), dim3((nenv + 63) / 64), dim3(Rows<64>::NT), 0, (hipStream_t)stream,
                       a);
  return g1err::launch("g1_hip_surrogate_forward");
}

int g1_hip_surrogate_step(int nenv, float dt, int phase, int period, float tol, float* qpos, float* qvel,
                          const float* tau, const float* xmean, const float* xinv, const void* const* W,
                          const float* const* b, const float* ymean, const float* ystd, const float* qacc_aba,
                          uint8_t* fallback, unsigned* stats, float* qacc_out, void* stream) {
  g1err::clear();
  if (nenv < 0 || period <= 0 || phase < 0 || phase >= period || !(dt > 0.f))
    return g1err::fail("g1_hip_surrogate_step: invalid arguments");
  if (!qpos || !qvel || !tau || !xmean || !xinv || !ymean || !ystd || !qacc_aba)
    return g1err::fail("g1_hip_surrogate_step: invalid arguments");
  SurrogateArgs a{};
  if (!fill_surrogate_weights(a, W, b)) return g1err::fail("g1_hip_surrogate_step: invalid arguments");
  if (nenv == 0) return 0;
  a.nenv = nenv; a.qpos = qpos; a.qvel = qvel; a.tau = tau; a.xmean = xmean; a.xinv = xinv;
  a.phase = phase; a.period = period; a.tol = tol; a.dt = dt; a.ymean = ymean; a.ystd = ystd;
  a.qacc_aba = qacc_aba; a.fallback = fallback; a.stats = stats; a.qacc_out = qacc_out;
  a.qpos_rw = qpos; a.qvel_rw = qvel;
  hipLaunchKernelGGL((surrogate_kernel<1, ROWS>), dim3((nenv + ROWS - 1) / ROWS), dim3(Rows<ROWS>::NT), 0,
                     (hipStream_t)stream, a);
  return g1err::launch("g1_hip_surrogate_step");
}

int g1_hip_surrogate_step_deferred(int nenv, float dt, int phase, int period, float tol, float* qpos, float* qvel,
                                   const float* tau, const float* xmean, const float* xinv, const void* const* W,
                                   const float* const* b, const float* ymean, const float* ystd,
                                   const float* qacc_aba, uint8_t* fallback, unsigned* stats, float* qacc_out,
                                   void* stream) {
  g1err::clear();
  if (nenv < 0 || period <= 0 || phase < 0 || phase >= period || !(dt > 0.f))
    return g1err::fail("g1_hip_surrogate_step_deferred: invalid arguments");
  if (!qpos || !qvel || !tau || !xmean || !xinv || !ymean || !ystd || !qacc_aba || !fallback || !stats)
    return g1err::fail("g1_hip_surrogate_step_deferred: invalid arguments");
  SurrogateArgs a{};
  if (!fill_surrogate_weights(a, W, b)) return g1err::fail("g1_hip_surrogate_step_deferred: invalid arguments");
  if (nenv == 0) return 0;
  a.nenv = nenv; a.qpos = qpos; a.qvel = qvel; a.tau = tau; a.xmean = xmean; a.xinv = xinv;
  a.phase = phase; a.period = period; a.tol = tol; a.dt = dt; a.ymean = ymean; a.ystd = ystd;
  a.qacc_aba = qacc_aba; a.fallback = fallback; a.stats = stats; a.qacc_out = qacc_out;
  a.qpos_rw = qpos; a.qvel_rw = qvel; a.committed = 1;
  hipLaunchKernelGGL((surrogate_kernel<1, ROWS>), dim3((nenv + ROWS - 1) / ROWS), dim3(Rows<ROWS>::NT), 0,
                     (hipStream_t)stream, a);
  return g1err::launch("g1_hip_surrogate_step_deferred");
}

int g1_hip_surrogate_judge_batch(int nenv, float tol, const float* xmean, const float* xinv, const void* const* W,
                                 const float* const* b, const float* ymean, const float* ystd, const float* snap,
                                 uint8_t* sflag, uint8_t* fallback, unsigned* stats, void* stream) {
  g1err::clear();
  if (nenv < 0 || !xmean || !xinv || !ymean || !ystd || !snap || !sflag || !fallback || !stats)
    return g1err::fail("g1_hip_surrogate_judge_batch: invalid arguments");
  SurrogateArgs a{};
  if (!fill_surrogate_weights(a, W, b)) return g1err::fail("g1_hip_surrogate_judge_batch: invalid arguments");
  if (nenv == 0) return 0;
  const size_t N = (size_t)nenv;
  a.nenv = nenv; a.qpos = snap; a.qvel = snap + 36 * N; a.tau = snap + 71 * N; a.xmean = xmean; a.xinv = xinv;
  a.tol = tol; a.ymean = ymean; a.ystd = ystd; a.qacc_aba = snap + 106 * N; a.fallback = fallback;
  a.stats = stats; a.sflag = sflag;
  if (nenv >= kRows128MinEnvs)
    hipLaunchKernelGGL((surrogate_kernel<2, 128>), dim3((nenv + 127) / 128), dim3(Rows<128>::NT), 0,
                       (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL((surrogate_kernel<2, 64>), dim3((nenv + 63) / 64), dim3(Rows<64>::NT), 0,
                       (hipStream_t)stream, a);
  return g1err::launch("g1_hip_surrogate_judge_batch");
}

// nsteps stepper steps in one host call: per step the judge + forces kernel
// (g1_hip_judge_forces, defined in g1_step.hip) and the network + commit
// kernel, phase (t0 + s) % period, the same launches in the same order as
// nsteps pairs of calls from the host.
int g1_hip_surrogate_steps(int nenv, int nsteps, int t0, float dt, int period, float tol, float* qpos, float* qvel,
                           float* anchor, const float* ctrl, const float* mu, const float* xmean, const float* xinv,
                           const void* const* W, const float* const* b, const float* ymean, const float* ystd,
                           float* qacc_aba, float* tau, uint8_t* fallback, unsigned* stats, float* qacc_out,
                           float* contact_forces, float* foot_pos, float* foot_vel, unsigned judge_flags,
                           void* stream) {
  g1err::clear();
  if (nenv < 0 || nsteps < 0 || t0 < 0 || period <= 0 || !(dt > 0.f) || !qpos || !qvel || !anchor || !ctrl ||
      !qacc_aba || !tau || !stats)
    return g1err::fail("g1_hip_surrogate_steps: invalid arguments");
  for (int s = 0; s < nsteps; ++s) {
    const int phase = (int)(((long long)t0 + s) % period);
    int rc = g1_hip_judge_forces(nenv, dt, qpos, qvel, anchor, ctrl, mu, fallback, phase, period, qacc_aba, tau,
                                 contact_forces, foot_pos, foot_vel, judge_flags, stream);
    if (rc) return rc;
    rc = g1_hip_surrogate_step(nenv, dt, phase, period, tol, qpos, qvel, tau, xmean, xinv, W, b, ymean, ystd,
                               qacc_aba, fallback, stats, qacc_out, stream);
    if (rc) return rc;
  }
  return 0;
}

// The same with deferred judging: per step g1_hip_judge_forces_commit (which
// also commits, and records if judged, every block whose envs are all in
// fallback) and g1_hip_surrogate_step_deferred (the other blocks), and the
// batch judgement (g1_hip_surrogate_judge_batch) after each step whose phase
// is period - 1.
//
// With blk_run (one byte per block of 64 envs) the steps go by period
// segments: an env leaves fallback only at a batch judgement, so a block
// whose envs are all in fallback at a segment's start stays so until the
// period ends, and every one of its steps until then is ABA's.  One
// g1_hip_fallback_run launch steps all such blocks through the segment
// (recording each at its judged step); the per-step launches then run the
// other blocks only.  Same results as without blk_run: the state, records,
// flags and statistics of the per-step path.
//
// all_fallback (host int, in/out, with blk_run; nullable): 1 on entry says
// that every env is in fallback (after the stepper's start, or a call that
// returned 1 and only resets since).  While that holds the per-step launches
// have nothing to do and are not issued.  After each batch judgement the
// envs off fallback are counted into stats[6] and read back once the next
// segment's fallback run is queued (one stream synchronisation per period);
// a count of 0 keeps skipping.  On return it says whether every env is in
// fallback after the last step.
static int steps_deferred(const char* who, int nenv, int nsteps, int t0, float dt, int period, float tol,
                          float* qpos, float* qvel, float* anchor, const float* ctrl, long long ctrl_stride,
                          int ctrl_period, int ctrl_step0, const float* mu, const float* xmean, const float* xinv,
                          const void* const* W, const float* const* b, const float* ymean, const float* ystd,
                          float* qacc_aba, float* tau, uint8_t* fallback, unsigned* stats, float* qacc_out,
                          float* contact_forces, float* foot_pos, float* foot_vel, unsigned judge_flags, float* snap,
                          uint8_t* sflag, uint8_t* blk_run, int* all_fallback, void* stream) {
  g1err::clear();
  if (nenv < 0 || nsteps < 0 || t0 < 0 || period <= 0 || !(dt > 0.f) || !qpos || !qvel || !anchor || !ctrl ||
      !qacc_aba || !tau || !stats || !fallback || !snap || !sflag || (all_fallback && !blk_run) ||
      ctrl_stride < 0 || (ctrl_stride > 0 && (ctrl_stride < 29LL * nenv || ctrl_period <= 0 || ctrl_step0 < 0)))
    return g1err::fail(who);
  // the targets of step s of this call (fixed without a schedule)
  auto ctrl_at = [&](int s) {
    return ctrl_stride > 0 ? ctrl + (size_t)((ctrl_step0 + s) / ctrl_period) * (size_t)ctrl_stride : ctrl;
  };
  const hipStream_t st = (hipStream_t)stream;
  bool skip = all_fallback && *all_fallback == 1;     // per-step launches have nothing to do
  // A count is queued in stats[6] with its copy into `off`: every return
  // waits for the stream first, so the copy never lands in a dead frame.
  struct Pending {
    hipStream_t st;
    bool counted = false;
    unsigned off = 0;
    ~Pending() {
      if (counted) (void)hipStreamSynchronize(st);
    }
  } pend{st};
  bool& counted = pend.counted;
  unsigned& off = pend.off;
  for (int s = 0; s < nsteps;) {
    const int phase0 = (int)(((long long)t0 + s) % period);
    const int seg = blk_run ? std::min(nsteps - s, period - phase0) : 1;
    if (blk_run) {
      int rc = g1k_host::fallback_run(nenv, seg, dt, qpos, qvel, anchor, ctrl_at(s), ctrl_stride, ctrl_period,
                                      ctrl_stride > 0 ? (ctrl_step0 + s) % ctrl_period : 0, mu, fallback, phase0,
                                      period, blk_run, qacc_out, contact_forces, foot_pos, foot_vel, snap, sflag,
                                      stats, judge_flags, stream);
      if (rc) return rc;
    }
    if (counted) {
      // the count of the last batch judgement (queued before this segment's run)
      if (hipStreamSynchronize(st) != hipSuccess) return g1err::fail("surrogate steps: sync failed");
      skip = off == 0;
      counted = false;
    }
    for (int j = 0; j < seg; ++j) {
      const int phase = phase0 + j;
      int rc;
      if (!skip) {
        rc = g1k_host::judge_forces_commit(nenv, dt, qpos, qvel, anchor, ctrl_at(s + j), mu, fallback, phase, period,
                                           qacc_aba, tau, contact_forces, foot_pos, foot_vel, qacc_out, snap,
                                           sflag, stats, blk_run, judge_flags, stream);
        if (rc) return rc;
        rc = g1_hip_surrogate_step_deferred(nenv, dt, phase, period, tol, qpos, qvel, tau, xmean, xinv, W, b,
                                            ymean, ystd, qacc_aba, fallback, stats, qacc_out, stream);
        if (rc) return rc;
      }
      if (phase == period - 1) {
        rc = g1_hip_surrogate_judge_batch(nenv, tol, xmean, xinv, W, b, ymean, ystd, snap, sflag, fallback, stats,
                                          stream);
        if (rc) return rc;
        if (all_fallback) {
          const int nvec = (reinterpret_cast<uintptr_t>(fallback) % 16 == 0) ? nenv / 16 : 0;
          hipLaunchKernelGGL(count_off_fallback_kernel, dim3(1), dim3(1024), 0, st, nenv, nvec, fallback, stats + 6);
          if (hipGetLastError() != hipSuccess)
            return g1err::fail("surrogate steps: count launch failed");
          if (hipMemcpyAsync(&off, stats + 6, sizeof(unsigned), hipMemcpyDeviceToHost, st) != hipSuccess)
            return g1err::fail("surrogate steps: count copy failed");
          counted = true;
        }
      }
    }
    s += seg;
  }
  if (counted) {
    if (hipStreamSynchronize(st) != hipSuccess) return g1err::fail("surrogate steps: sync failed");
    skip = off == 0;
    counted = false;
  }
  if (all_fallback) *all_fallback = skip ? 1 : 0;
  return 0;
}

int g1_hip_surrogate_steps_deferred(int nenv, int nsteps, int t0, float dt, int period, float tol, float* qpos,
                                    float* qvel, float* anchor, const float* ctrl, const float* mu,
                                    const float* xmean, const float* xinv, const void* const* W,
                                    const float* const* b, const float* ymean, const float* ystd, float* qacc_aba,
                                    float* tau, uint8_t* fallback, unsigned* stats, float* qacc_out,
                                    float* contact_forces, float* foot_pos, float* foot_vel, unsigned judge_flags,
                                    float* snap, uint8_t* sflag, uint8_t* blk_run, int* all_fallback, void* stream) {
  return steps_deferred("g1_hip_surrogate_steps_deferred: invalid arguments", nenv, nsteps, t0, dt, period, tol,
                        qpos, qvel, anchor, ctrl, 0, 0, 0, mu, xmean, xinv, W, b, ymean, ystd, qacc_aba, tau,
                        fallback, stats, qacc_out, contact_forces, foot_pos, foot_vel, judge_flags, snap, sflag,
                        blk_run, all_fallback, stream);
}

// The same with a PD-target schedule: step s of the call uses the block
// ctrl_blocks + ((ctrl_step0 + s) / ctrl_period) block_stride, as
// g1_hip_step_schedule; with blk_run the fallback runs switch blocks
// themselves and span whole period segments.
int g1_hip_surrogate_steps_schedule(int nenv, int nsteps, int t0, float dt, int period, float tol, float* qpos,
                                    float* qvel, float* anchor, const float* ctrl_blocks, long long block_stride,
                                    int ctrl_period, int ctrl_step0, const float* mu, const float* xmean,
                                    const float* xinv, const void* const* W, const float* const* b,
                                    const float* ymean, const float* ystd, float* qacc_aba, float* tau,
                                    uint8_t* fallback, unsigned* stats, float* qacc_out, float* contact_forces,
                                    float* foot_pos, float* foot_vel, unsigned judge_flags, float* snap,
                                    uint8_t* sflag, uint8_t* blk_run, int* all_fallback, void* stream) {
  if (block_stride <= 0) {
    g1err::clear();
    return g1err::fail("g1_hip_surrogate_steps_schedule: invalid arguments");
  }
  return steps_deferred("g1_hip_surrogate_steps_schedule: invalid arguments", nenv, nsteps, t0, dt, period, tol,
                        qpos, qvel, anchor, ctrl_blocks, block_stride, ctrl_period, ctrl_step0, mu, xmean, xinv, W,
                        b, ymean, ystd, qacc_aba, tau, fallback, stats, qacc_out, contact_forces, foot_pos, foot_vel,
                        judge_flags, snap, sflag, blk_run, all_fallback, stream);
}

int g1_policy_act(int nenv, int layout, const float* qpos, float* qvel, const float* commands,
                  const float* last_ctrl, const float* obs_mean, const float* obs_inv, float clip,
                  const void* const* W, const float* const* b, const float* log_std, unsigned seed, int t,
                  float* ctrl, float* actions, float* obs_out, int* push_interval, void* stream) {
  g1err::clear();
  if (nenv < 0 || (layout != 0 && layout != 1) || !qpos || !qvel || !commands || !last_ctrl || !obs_mean ||
      !obs_inv || !ctrl || !W || !b || !(clip > 0.f))
    return g1err::fail("g1_policy_act: invalid arguments");
  PolicyArgs a{};
  for (int l = 0; l < 4; ++l) {
    if (!W[l] || !b[l]) return g1err::fail("g1_policy_act: invalid arguments");
    a.W[l] = (const bf16*)W[l];
    a.b[l] = b[l];
  }
  if (nenv == 0) return 0;
  a.nenv = nenv; a.layout = layout; a.nobs = layout == 0 ? 99 : 96;
  a.qpos = qpos; a.qvel = qvel; a.commands = commands; a.last_ctrl = last_ctrl;
  a.obs_mean = obs_mean; a.obs_inv = obs_inv; a.clip = clip; a.log_std = log_std; a.seed = seed; a.t = t;
  a.ctrl = ctrl; a.actions = actions; a.obs_out = obs_out; a.push_interval = push_interval;
  if (nenv >= kRows128MinEnvs)
    hipLaunchKernelGGL(policy_kernel<128>, dim3((nenv + 127) / 128), dim3(Rows<128>::NT), 0, (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL(policy_kernel<64>, dim3((nenv + 63) / 64), dim3(Rows<64>::NT), 0, (hipStream_t)stream, a);
  return g1err::launch("g1_policy_act");
}

int g1_hip_get_obs(int nenv, int layout, const float* qpos, const float* qvel, const float* commands,
                   const float* last_ctrl, float* obs, void* stream) {
  g1err::clear();
  if (nenv < 0 || (layout != 0 && layout != 1) || !qpos || !qvel || !commands || !last_ctrl || !obs)
    return g1err::fail("g1_hip_get_obs: invalid arguments");
  if (nenv == 0) return 0;
  const int nobs = layout == 0 ? 99 : 96;
  const size_t n = (size_t)nenv * nobs;
  hipLaunchKernelGGL(obs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, nenv, layout,
                     nobs, qpos, qvel, commands, last_ctrl, obs);
  return g1err::launch("g1_hip_get_obs");
}

int g1_rollout_post(int nenv, const float* qpos, const float* qvel, const float* foot_pos, const float* foot_vel,
                    const float* contact_forces, const float* actions, float* commands, float* prev_actions,
                    int* command_timeout, int* episode_length, int* push_interval, int max_episode_length,
                    float cmd_lin, float cmd_yaw, unsigned seed, int t, float* reward, uint8_t* done,
                    uint8_t* reset, double* reward_sum, double* done_count, void* stream) {
  g1err::clear();
  if (nenv < 0 || max_episode_length <= 0 || !qpos || !qvel || !foot_pos || !foot_vel || !contact_forces ||
      !actions || !commands || !prev_actions || !command_timeout || !episode_length || !reward || !done || !reset)
    return g1err::fail("g1_rollout_post: invalid arguments");
  if (nenv == 0) return 0;
  PostArgs a{nenv, qpos, qvel, foot_pos, foot_vel, contact_forces, actions, commands, prev_actions,
             command_timeout, episode_length, push_interval, max_episode_length, cmd_lin, cmd_yaw, seed, t,
             reward, done, reset, reward_sum, done_count};
  hipLaunchKernelGGL(post_step_kernel, dim3((nenv + 255) / 256), dim3(256), 0, (hipStream_t)stream, a);
  return g1err::launch("g1_rollout_post");
}

int g1_rollout_init(int nenv, float cmd_lin, float cmd_yaw, unsigned seed, float* commands, int* command_timeout,
                    int* episode_length, int* push_interval, void* stream) {
  g1err::clear();
  if (nenv < 0 || !commands || !command_timeout || !episode_length)
    return g1err::fail("g1_rollout_init: invalid arguments");
  if (nenv == 0) return 0;
  hipLaunchKernelGGL(rollout_init_kernel, dim3((nenv + 255) / 256), dim3(256), 0, (hipStream_t)stream, nenv,
                     cmd_lin, cmd_yaw, seed, commands, command_timeout, episode_length, push_interval);
  return g1err::launch("g1_rollout_init");
}

// The two normaliser kernels on a workspace of 2 * NORM_CHUNKS * nobs + 1
// doubles: every slot the merge reads is written by the batch pass of the
// same call, so nothing needs clearing beforehand.
static int obs_norm_launch(const char* who, int nenv, int layout, const float* qpos, const float* qvel,
                           const float* commands, const float* last_ctrl, float* mean, float* var, float* count,
                           float* inv, double* scratch, void* stream) {
  const int nobs = layout == 0 ? 99 : 96;
  const int chunk = (nenv + NORM_CHUNKS - 1) / NORM_CHUNKS;
  const int nchunks = (nenv + chunk - 1) / chunk;
  hipLaunchKernelGGL(obs_batch_stats_kernel, dim3(nobs, nchunks), dim3(256), 0, (hipStream_t)stream, nenv, layout,
                     chunk, qpos, qvel, commands, last_ctrl, mean, count, scratch);
  hipLaunchKernelGGL(obs_norm_merge_kernel, dim3(nobs), dim3(64), 0, (hipStream_t)stream, nenv, nobs, nchunks, scratch,
                     mean, var, count, inv);
  return g1err::launch(who);
}

// Workspace of g1_policy_obs_norm_update_noalloc, one copy per GPU for the
// life of the process: the per-(feature, chunk) sums of the widest layout
// (99 features) and the running count, 2 * 64 * 99 + 1 doubles (99 KiB).
// In host code `scratch` only names the symbol; its device address on the
// current GPU comes from hipGetSymbolAddress.
constexpr int NORM_MAX_OBS = 99;
constexpr size_t NORM_SCRATCH_DOUBLES = 2 * (size_t)NORM_CHUNKS * NORM_MAX_OBS + 1;
static __device__ __align__(256) double scratch[NORM_SCRATCH_DOUBLES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices];   // device address of `scratch` on each GPU, set once

// Device address of `scratch` on the current GPU (looked up on the first
// call on that GPU), or null with the error text set.
static double* norm_scratch(const char* who) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    g1err::fail("%s: no current HIP device", who);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !p) {
      g1err::fail("%s: hipGetSymbolAddress(scratch) failed on device %d: %s", who, dev, hipGetErrorString(err));
      return nullptr;
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  return g_scratch_ptr[dev];
}

int g1_policy_obs_norm_update(int nenv, int layout, const float* qpos, const float* qvel, const float* commands,
                              const float* last_ctrl, float* mean, float* var, float* count, float* inv,
                              double* scratch_ws, void* stream) {
  g1err::clear();
  if (nenv <= 0 || (layout != 0 && layout != 1) || !qpos || !qvel || !commands || !last_ctrl || !mean || !var ||
      !count || !inv || !scratch_ws)
    return g1err::fail("g1_policy_obs_norm_update: invalid arguments");
  return obs_norm_launch("g1_policy_obs_norm_update", nenv, layout, qpos, qvel, commands, last_ctrl, mean, var,
                         count, inv, scratch_ws, stream);
}

int g1_policy_obs_norm_update_noalloc(int nenv, int layout, const float* qpos, const float* qvel,
                                      const float* commands, const float* last_ctrl, float* mean, float* var,
                                      float* count, float* inv, void* stream) {
  static const char* who = "g1_policy_obs_norm_update_noalloc";
  g1err::clear();
  if (nenv <= 0 || (layout != 0 && layout != 1) || !qpos || !qvel || !commands || !last_ctrl || !mean || !var ||
      !count || !inv)
    return g1err::fail("%s: invalid arguments", who);
  const int nobs = layout == 0 ? 99 : 96;
  if (2 * (size_t)NORM_CHUNKS * nobs + 1 > NORM_SCRATCH_DOUBLES)
    return g1err::fail("%s: %d features exceed the static workspace of %d", who, nobs, NORM_MAX_OBS);
  double* ws = norm_scratch(who);
  if (!ws) return -1;
  return obs_norm_launch(who, nenv, layout, qpos, qvel, commands, last_ctrl, mean, var, count, inv, ws, stream);
}

}  // extern "C"
