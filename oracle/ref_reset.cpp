// ref_reset.cpp -- CPU driver of the reference's reset kernel (TEST
// INFRASTRUCTURE: built into oracle/_ref/, loaded only by the fixture script
// tests/golden/make_reset_golden.py and by tests).
//
// Includes the reference's own src/g1_kernel.cu where it lies, compiled
// through its own CPU shim header model/fake_cuda.h (FAKE_CUDA: device
// pointers are host pointers, thread/block indices are thread-local
// variables).  Nothing of the reference is copied.  g1ref_reset_done runs
// g1_reset_done_kernel (src/g1_kernel.cu:85-103) once per env of the grid
// g1_cuda_reset_done would launch (128 threads per block,
// src/g1_kernel.cu:165-172), one thread at a time.  The recipe
// (oracle/Makefile) compiles it with FMA contraction on, which is what a
// CUDA build of the kernel does by default (nvcc --fmad=true): the noise
// term becomes one fused multiply-add there.
#define main g1ref_kernel_benchmark_main
#include "g1_kernel.cu"
#undef main

extern "C" void g1ref_reset_done(int nenv, float* qpos, float* qvel, float* anchor, const uint8_t* done,
                                 unsigned seed, float noise, float drop) {
  const unsigned tpb = 128, nblocks = (unsigned)((nenv + tpb - 1) / tpb);
  gridDim = dim3(nblocks);
  blockDim = dim3(tpb);
  for (unsigned b = 0; b < nblocks; ++b)
    for (unsigned t = 0; t < tpb; ++t) {
      blockIdx = dim3(b);
      threadIdx = dim3(t);
      g1_reset_done_kernel(nenv, qpos, qvel, anchor, done, seed, noise, drop);
    }
}

// The reference's hash (src/g1_kernel.cu:54-58), for the fixture script.
extern "C" unsigned g1ref_hash(unsigned x) { return g1_hash(x); }
