"""Synthetic workloads of the benchmark configurations (host side, numpy).

Every state is drawn in float64 from ``numpy.random.default_rng(seed)`` and
rounded to float32, so the GPU and the oracle receive the same float32 values.
Arrays are structure-of-arrays ``[dim][nenv]`` (env fastest), the layout of
the GPU C-ABI (src/g1_kernel.cu:9-10).

* drop test (1,024 envs): base at 0.80 m, identity quaternion tilted by
  U(+-0.05) rad roll/pitch, joints = stand +- U(0.05), joint qd ~ N(0, 0.1),
  PD targets = stand hinge angles.
* randomized actuation (16,384 envs): as the drop test plus base linear
  velocity U(+-0.5) m/s in x/y and joint qd ~ N(0, 0.3); PD targets
  stand +- U(0.3) rad, resampled every 10 steps.
* contact stress (131,072 envs): randomized-actuation start and targets, a
  horizontal push |dv| ~ U(0, 1) m/s every 100 steps, per-env friction
  mu ~ U(0.4, 1.2).
"""
import numpy as np

from .model import MODEL, NC, NQ, NU, NV

ANCHOR_FREE = np.float32(1e30)
STAND = MODEL.qpos_stand


def quat_roll_pitch(roll, pitch):
    """(w, x, y, z) of R = Ry(pitch) Rx(roll), yaw 0."""
    cr, sr = np.cos(roll / 2), np.sin(roll / 2)
    cp, sp = np.cos(pitch / 2), np.sin(pitch / 2)
    return np.stack([cp * cr, cp * sr, sp * cr, -sp * sr])


def _base_state(rng, nenv, joint_qd_std, base_vxy):
    qpos = np.zeros((NQ, nenv))
    qpos[2] = 0.80
    roll = rng.uniform(-0.05, 0.05, nenv)
    pitch = rng.uniform(-0.05, 0.05, nenv)
    qpos[3:7] = quat_roll_pitch(roll, pitch)
    qpos[7:] = STAND[7:, None] + rng.uniform(-0.05, 0.05, (NU, nenv))
    qvel = np.zeros((NV, nenv))
    if base_vxy:
        qvel[0:2] = rng.uniform(-base_vxy, base_vxy, (2, nenv))
    qvel[6:] = rng.normal(0.0, joint_qd_std, (NU, nenv))
    anchor = np.full((2 * NC, nenv), ANCHOR_FREE, dtype=np.float32)
    return (qpos.astype(np.float32), qvel.astype(np.float32), anchor)


def drop_test(nenv=1024, seed=0):
    """Returns qpos, qvel, anchor, ctrl (float32 SoA)."""
    rng = np.random.default_rng(seed)
    qpos, qvel, anchor = _base_state(rng, nenv, 0.1, 0.0)
    ctrl = np.repeat(STAND[7:, None], nenv, axis=1).astype(np.float32)
    return qpos, qvel, anchor, ctrl


def randomized_actuation(nenv=16384, seed=1, nblocks=25):
    """Returns qpos, qvel, anchor and ctrl blocks [nblocks][29][nenv]
    (block k holds the targets of steps 10k .. 10k+9)."""
    rng = np.random.default_rng(seed)
    qpos, qvel, anchor = _base_state(rng, nenv, 0.3, 0.5)
    ctrl = STAND[None, 7:, None] + rng.uniform(-0.3, 0.3, (nblocks, NU, nenv))
    return qpos, qvel, anchor, ctrl.astype(np.float32)


def contact_stress(nenv=131072, seed=2, nblocks=100, npush=10):
    """Returns qpos, qvel, anchor, ctrl blocks, mu [nenv] and pushes
    [npush][2][nenv] (world-frame dv added to qvel[0:2] before step 100k)."""
    qpos, qvel, anchor, ctrl = randomized_actuation(nenv, seed, nblocks)
    rng = np.random.default_rng(seed + 1000)
    mu = rng.uniform(0.4, 1.2, nenv).astype(np.float32)
    mag = rng.uniform(0.0, 1.0, (npush, nenv))
    ang = rng.uniform(0.0, 2 * np.pi, (npush, nenv))
    pushes = np.stack([mag * np.cos(ang), mag * np.sin(ang)], axis=1).astype(np.float32)
    return qpos, qvel, anchor, ctrl, mu, pushes


def random_state(rng, n=1):
    """States of the reference's MuJoCo validation (tests/test_vs_mujoco.py:23-30):
    base position U(-1, 1), uniform random unit quaternion, joints in the inner
    80 % of their range, qvel U(-2, 2).  float64, shape [n][dim]."""
    qpos = np.zeros((n, NQ))
    qvel = np.zeros((n, NV))
    lo, hi = MODEL.jnt_lo, MODEL.jnt_hi
    for i in range(n):
        qpos[i, :3] = rng.uniform(-1, 1, 3)
        q = rng.normal(size=4)
        qpos[i, 3:7] = q / np.linalg.norm(q)
        qpos[i, 7:] = rng.uniform(lo + 0.1 * (hi - lo), hi - 0.1 * (hi - lo))
        qvel[i] = rng.uniform(-2, 2, NV)
    return qpos, qvel
