"""Compile-time model view of the step kernels (g1sim_amd/csrc/g1_model_kernel.h),
checked on the host: the tree-rotation classification that decides which
bodies fold their constant tree rotation into the joint angle (planar about
the hinge's own axis), which use the planar-rotation code, and which keep a
general 3x3 rotation. Built with g++ from the header alone (no GPU)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "g1sim_amd", "csrc")

PROBE = r"""
#include <cstdio>
#define __device__
#define __forceinline__ inline
#include "g1_model_kernel.h"
int main() {
  for (int i = 0; i < g1k::NB; ++i) {
    const auto& b = g1k::kM.b[i];
    std::printf("%d %d %d %d %d %.17g %.17g\n", i, b.axis, (int)b.rt_ident, b.rt_axis, (int)b.rt_fold, b.rt_c,
                b.rt_s);
  }
}
"""


@pytest.fixture(scope="module")
def bodies(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("model_kernel")
    src, exe = d / "probe.cpp", d / "probe"
    src.write_text(PROBE)
    subprocess.run([gxx, "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    return {int(r[0]): dict(axis=int(r[1]), ident=bool(int(r[2])), rt_axis=int(r[3]), fold=bool(int(r[4])),
                            c=float(r[5]), s=float(r[6])) for r in rows}


def _quat_to_mat(w, x, y, z):
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_tree_rotation_classification(bodies):
    # knees and shoulder rolls: planar about their own hinge axis -> folded;
    # hip rolls: planar about y, hinge about x -> planar code; shoulder
    # pitches: a 1e-4 z component -> general 3x3
    assert sorted(i for i, b in bodies.items() if b["fold"]) == [4, 10, 17, 24]
    assert sorted(i for i, b in bodies.items() if b["rt_axis"] >= 0 and not b["fold"]) == [2, 8]
    assert sorted(i for i, b in bodies.items() if not b["ident"] and b["rt_axis"] < 0) == [16, 23]


def test_planar_tree_rotation_matches_quaternion(bodies):
    """R_tree of every planar body is R_k(theta) with the stored cos/sin, the
    rotation convention of the kernels' rot<K> (identity on k, [[c, -s], [s, c]]
    on ((k+1)%3, (k+2)%3)); and for a folded body R_tree R_k(q) = R_k(theta + q)."""
    import json
    m = json.load(open(os.path.join(ROOT, "g1sim_amd", "model", "g1_model.json")))
    tq = np.array(m["tree_quat"]).reshape(-1, 4)
    for i, b in bodies.items():
        if b["rt_axis"] < 0:
            continue
        k = b["rt_axis"]
        I, J = (k + 1) % 3, (k + 2) % 3
        R = np.eye(3)
        R[I, I], R[I, J], R[J, I], R[J, J] = b["c"], -b["s"], b["s"], b["c"]
        np.testing.assert_allclose(R, _quat_to_mat(*tq[i]), atol=1e-15)
        if b["fold"]:
            assert b["axis"] == k
            q = 0.37
            Rq = np.eye(3)
            Rq[I, I], Rq[I, J], Rq[J, I], Rq[J, J] = np.cos(q), -np.sin(q), np.sin(q), np.cos(q)
            th = np.arctan2(b["s"], b["c"])
            Rs = np.eye(3)
            Rs[I, I], Rs[I, J], Rs[J, I], Rs[J, J] = np.cos(th + q), -np.sin(th + q), np.sin(th + q), np.cos(th + q)
            np.testing.assert_allclose(R @ Rq, Rs, atol=1e-14)
