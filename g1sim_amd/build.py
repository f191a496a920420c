"""Builds libg1sim_amd.so (HIP, gfx950) in-tree with hipcc.

    python -m g1sim_amd.build            # or: from g1sim_amd.build import build; build()

The library lands next to this file, where g1sim_amd._lib loads it.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libg1sim_amd.so")
SOURCES = [os.path.join(HERE, "csrc", n) for n in ("g1_step.hip", "g1_mlp.hip", "g1_infer.hip", "g1_peaks.hip")]
ARCH = os.environ.get("G1SIM_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # code object v5 loads under both the ROCm runtime bundled with torch and /opt/rocm's
    "-mcode-object-version=5",
    # no packed FP32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def build(verbose=True, extra_flags=()):
    cmd = [hipcc(), *HIPCC_FLAGS, *extra_flags, "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(HERE, "csrc"), *SOURCES, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    # the host pass warns that it does not know the packed-fp32 feature; keep other warnings visible
    noise = [l for l in (r.stderr or "").splitlines()
             if l.strip() and "packed-fp32-ops" not in l and "warning generated" not in l]
    if noise and verbose:
        sys.stderr.write("\n".join(noise) + "\n")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build()
