#!/usr/bin/env python3
"""Writes tests/golden/reset.npz from the reference's own reset kernel
(src/g1_kernel.cu:54-103, g1_reset_done_kernel with its hash and uniform
draw), compiled on the CPU through the reference's FAKE_CUDA shim by
oracle/Makefile into oracle/_ref/libg1ref_reset.so.  Run in a tree where
/root/reference exists:

    make -C oracle ref && python tests/golden/make_reset_golden.py

The fixture is data: for every (seed, noise, drop) of seeds {0, 42, 1234,
2^32 - 1} x noise {0, 0.05, 0.3} x drop {0, 0.02, 1.0}, the kernel is run
over a batch of NENV envs with a done mask, and qpos of ENVS (rows 2 and
7-35, the ones a reset draws or shifts) is stored, plus the hash of a few
words.  The other rows, qvel = 0 and anchors = 1e30 are checked live.
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
OUT = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(ROOT, "oracle", "_ref", "libg1ref_reset.so")

NENV = 65536
SEEDS = np.array([0, 42, 1234, 2 ** 32 - 1], np.uint32)
NOISES = np.array([0.0, 0.05, 0.3], np.float32)
DROPS = np.array([0.0, 0.02, 1.0], np.float32)
ROWS = np.array([2] + list(range(7, 36)), np.int64)


def load():
    lib = C.CDLL(LIB)
    lib.g1ref_hash.restype = C.c_uint
    lib.g1ref_hash.argtypes = [C.c_uint]
    lib.g1ref_reset_done.restype = None
    lib.g1ref_reset_done.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint,
                                     C.c_float, C.c_float]
    return lib


def envs():
    """0, 1, 127, 128 (block edges), NENV - 1 and 27 random envs."""
    rng = np.random.default_rng(2024)
    rest = rng.choice(np.arange(129, NENV - 1), 27, replace=False)
    return np.sort(np.concatenate([[0, 1, 127, 128, NENV - 1], rest])).astype(np.int64)


def main():
    lib = load()
    env = envs()
    done = np.zeros(NENV, np.uint8)
    done[env] = 1
    qpos = np.zeros((SEEDS.size, NOISES.size, DROPS.size, ROWS.size, env.size), np.float32)
    for a, seed in enumerate(SEEDS):
        for b, noise in enumerate(NOISES):
            for c, drop in enumerate(DROPS):
                qp = np.full((36, NENV), 7.0, np.float32)
                qv = np.full((35, NENV), 7.0, np.float32)
                an = np.full((16, NENV), 7.0, np.float32)
                lib.g1ref_reset_done(NENV, qp.ctypes.data, qv.ctypes.data, an.ctypes.data, done.ctypes.data,
                                     int(seed), float(noise), float(drop))
                untouched = np.setdiff1d(np.arange(NENV), env)
                assert (qp[:, untouched] == 7).all() and (qv[:, env] == 0).all() and (an[:, env] == 1e30).all()
                qpos[a, b, c] = qp[np.ix_(ROWS, env)]
    words = np.array([0, 1, 2, 0x9E3779B9, 0xDEADBEEF, 0xFFFFFFFF, 12345678, 0x80000000], np.uint32)
    hashes = np.array([lib.g1ref_hash(int(w)) for w in words], np.uint32)
    np.savez_compressed(os.path.join(OUT, "reset.npz"), nenv=np.int64(NENV), seeds=SEEDS, noises=NOISES,
                        drops=DROPS, rows=ROWS, envs=env, qpos=qpos, hash_in=words, hash_out=hashes)
    print("wrote reset.npz", qpos.shape)


if __name__ == "__main__":
    sys.exit(main())
