#!/usr/bin/env python3
"""Writes the golden vectors of tests/golden/ from the reference's own FP64
host library (src/host_lib.cpp built by oracle/Makefile into
oracle/_ref/libg1host.so).  Run in a tree where /root/reference exists:

    make -C oracle && python tests/golden/make_golden.py

Every fixture is data: float64 inputs and the reference library's outputs.

* fk.npz      -- g1_c_fk on 32 states drawn like tests/test_vs_mujoco.py:23-30.
* fd.npz      -- g1_c_fd on 64 such states with tau ~ U(-30, 30).
* step1.npz   -- one g1_c_step (nsub 1, dt 2e-3) from 96 contact-rich states:
                 drop-test and randomized-actuation rollouts sampled during
                 impact and stance, plus engineered branch cases (airborne,
                 new touchdown, slip, stick, limits, torque clamp, fn clamps).
* settle.npz  -- 600 steps at dt 2e-3 from the stand pose + 0.02 m
                 (tests/test_vs_mujoco.py:114-133 scenario), qpos every 50
                 steps, final fn and anchors.
* quat.npz    -- the quaternion-integration case of
                 tests/test_vs_mujoco.py:102-112 through g1_c_step.
* nsub.npz    -- g1_c_step with nsub 5 from 16 states (substep loop).
"""
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from g1sim_amd import workloads as W  # noqa: E402
from g1sim_amd.model import MODEL  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
DT = 2e-3


def main():
    ref = O.ref_host()
    rng = np.random.default_rng(20261203)

    qp, qv = W.random_state(rng, 32)
    xpos = np.zeros((32, 30, 3))
    xmat = np.zeros((32, 30, 9))
    for i in range(32):
        xpos[i], xmat[i] = O.fk(qp[i], ref)
    np.savez_compressed(os.path.join(OUT, "fk.npz"), qpos=qp, xpos=xpos, xmat=xmat)

    qp, qv = W.random_state(rng, 64)
    tau = rng.uniform(-30, 30, (64, 35))
    qacc = np.stack([O.fd(qp[i], qv[i], tau[i], ref) for i in range(64)])
    np.savez_compressed(os.path.join(OUT, "fd.npz"), qpos=qp, qvel=qv, tau=tau, qacc=qacc)

    # ---- contact-rich states from rollouts of the reference library
    states = []
    for cfg, seed in (("drop", 0), ("rand", 1)):
        if cfg == "drop":
            p, v, a, c = W.drop_test(8, seed)
            ctrl_of = lambda k: c  # noqa: E731
        else:
            p, v, a, cb = W.randomized_actuation(8, seed, nblocks=30)
            ctrl_of = lambda k: cb[k // 10]  # noqa: E731
        p, v, a = p.astype(np.float64), v.astype(np.float64), a.astype(np.float64)
        for e in range(8):
            qpe, qve, ane = p[:, e].copy(), v[:, e].copy(), a[:, e].copy()
            for k in range(300):
                ctl = ctrl_of(k)[:, e].astype(np.float64)
                if k in (0, 20, 45, 70, 150, 299):
                    states.append((qpe.copy(), qve.copy(), ane.copy(), ctl.copy()))
                qpe, qve, ane, _ = O.step(qpe, qve, ctl, ane, 1, DT, ref)
    stand = MODEL.qpos_stand
    base = states[-1]
    # engineered branch cases around a settled stance state
    cases = []
    q0, v0, a0, c0 = (x.copy() for x in base)
    cases.append((q0.copy(), v0.copy(), np.full(16, 1e30), c0.copy()))       # new touchdown
    q = q0.copy(); q[2] += 0.2
    cases.append((q, v0.copy(), a0.copy(), c0.copy()))                          # airborne
    v = v0.copy(); v[0] += 1.5
    cases.append((q0.copy(), v, a0.copy(), c0.copy()))                          # slip
    v = v0.copy(); v[2] -= 3.0
    cases.append((q0.copy(), v, a0.copy(), c0.copy()))                          # fn clamp 500
    v = v0.copy(); v[2] += 2.0
    cases.append((q0.copy(), v, a0.copy(), c0.copy()))                          # fn clamp 0
    q = q0.copy(); q[7 + 3] = MODEL.jnt_hi[3] + 0.05; q[7 + 16] = MODEL.jnt_lo[16] - 0.05
    cases.append((q, v0.copy(), a0.copy(), c0.copy()))                          # limits
    c = c0.copy(); c[:] = stand[7:] + 2.0
    cases.append((q0.copy(), v0.copy(), a0.copy(), c))                          # torque clamp
    an = a0.copy(); an[0] += 0.02
    cases.append((q0.copy(), v0.copy(), an, c0.copy()))                          # stretched anchor
    states += cases
    n = len(states)
    ins = {k: np.stack([s[i] for s in states]) for i, k in enumerate(("qpos", "qvel", "anchor", "ctrl"))}
    outs = {"qpos_out": [], "qvel_out": [], "anchor_out": [], "fn": []}
    for i in range(n):
        p, v, a, fn = O.step(ins["qpos"][i], ins["qvel"][i], ins["ctrl"][i], ins["anchor"][i], 1, DT, ref)
        outs["qpos_out"].append(p); outs["qvel_out"].append(v)
        outs["anchor_out"].append(a); outs["fn"].append(fn)
    np.savez_compressed(os.path.join(OUT, "step1.npz"), dt=DT, **ins,
                        **{k: np.stack(v) for k, v in outs.items()})
    print(f"step1: {n} states")

    qp = stand.copy(); qp[2] += 0.02
    qv = np.zeros(35); an = np.full(16, 1e30); ctrl = stand[7:].copy()
    traj = []
    for k in range(600):
        qp, qv, an, fn = O.step(qp, qv, ctrl, an, 1, DT, ref)
        if (k + 1) % 50 == 0:
            traj.append(qp.copy())
    np.savez_compressed(os.path.join(OUT, "settle.npz"), qpos_traj=np.stack(traj), qvel=qv,
                        anchor=an, fn=fn)

    quat = np.array([0.3, -0.5, 0.7, 0.4]); quat /= np.linalg.norm(quat)
    qp = np.zeros(36); qv = np.zeros(35); qp[3:7] = quat; qp[2] = 5.0; qv[3:6] = [1.3, -2.1, 0.7]
    qp2, qv2, _, _ = O.step(qp, qv, np.zeros(29), np.full(16, 1e30), 1, 0.01, ref)
    np.savez_compressed(os.path.join(OUT, "quat.npz"), qpos=qp, qvel=qv, qpos_out=qp2, qvel_out=qv2)

    idx = np.arange(0, n, max(1, n // 16))[:16]
    outs = [O.step(ins["qpos"][i], ins["qvel"][i], ins["ctrl"][i], ins["anchor"][i], 5, DT, ref) for i in idx]
    np.savez_compressed(os.path.join(OUT, "nsub.npz"), nsub=5, dt=DT,
                        **{k: ins[k][idx] for k in ins},
                        qpos_out=np.stack([o[0] for o in outs]), qvel_out=np.stack([o[1] for o in outs]),
                        anchor_out=np.stack([o[2] for o in outs]), fn=np.stack([o[3] for o in outs]))
    print("wrote", sorted(f for f in os.listdir(OUT) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
