"""GPU parity of the fused step (libg1sim_amd.so through its C-ABI) against
the FP64 oracle on the same float32 inputs (teacher forcing).

Tolerances (float32 step vs float64 oracle, same input state):
  qacc        rel. err |a - ref| / max(|ref|, 1) (tests/test_vs_mujoco.py:74-75)
              below QACC_TOL on every env whose contact decisions agree;
  next state  qpos/qvel within a few float32 ulps of the state scale;
  contacts    identical on/off and stick/slip decisions except for envs at a
              threshold (counted, bounded by FLIP_FRAC);
  integers    reset RNG bit-exact.
"""
import numpy as np
import pytest
import torch

from g1sim_amd import _lib, workloads as W
from g1sim_amd.model import MODEL
from oracle import oracle as O

from .gpu_helpers import GpuState, branch_flips, rel_err, teacher_forced_step, write_metrics

pytestmark = pytest.mark.gpu

QACC_TOL = 2e-3          # max rel. qacc error of the float32 step on non-flipped envs
QPOS_TOL = 2e-5          # next-state qpos abs. error after one step
FLIP_FRAC = 0.01         # envs with a contact decision at a threshold


def _rollout_metrics(st, steps, dt, ctrl_of=None, mu=None, every=1):
    worst = {"qacc_rel": 0.0, "qpos_abs": 0.0, "qvel_rel": 0.0, "fn_abs": 0.0, "foot_pos_abs": 0.0,
             "foot_vel_abs": 0.0}
    flips = 0
    checked = 0
    for s in range(steps):
        if ctrl_of is not None:
            st.ctrl.copy_(torch.as_tensor(ctrl_of(s), device=st.device))
        if s % every:
            st.step(1, dt)
            continue
        before, after, ref = teacher_forced_step(st, dt, mu)
        flip = branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
        ok = ~flip
        flips += int(flip.sum())
        checked += st.n
        worst["qacc_rel"] = max(worst["qacc_rel"], float(rel_err(after["qacc"][:, ok], ref["qacc"][:, ok]).max(initial=0)))
        worst["qpos_abs"] = max(worst["qpos_abs"], float(np.abs(after["qpos"][:, ok] - ref["qpos"][:, ok]).max(initial=0)))
        worst["qvel_rel"] = max(worst["qvel_rel"], float(rel_err(after["qvel"][:, ok], ref["qvel"][:, ok]).max(initial=0)))
        worst["fn_abs"] = max(worst["fn_abs"], float(np.abs(after["fn"][:, ok] - ref["fn"][:, ok]).max(initial=0)))
        worst["foot_pos_abs"] = max(worst["foot_pos_abs"], float(np.abs(after["foot_pos"] - ref["foot_pos"]).max()))
        worst["foot_vel_abs"] = max(worst["foot_vel_abs"], float(np.abs(after["foot_vel"] - ref["foot_vel"]).max()))
    worst["flip_frac"] = flips / max(checked, 1)
    worst["checked_env_steps"] = checked
    return worst


def test_drop_test_teacher_forced(device):
    """Drop test, 1,024 envs x 250 steps at dt 2e-3, every step checked."""
    qp, qv, an, ct = W.drop_test(1024, seed=0)
    st = GpuState(qp, qv, an, ct, device)
    m = _rollout_metrics(st, 250, 2e-3)
    write_metrics("drop_test_1024x250", m)
    h = st.host()
    assert np.isfinite(h["qpos"]).all()
    z = h["qpos"][2]
    assert ((z > 0.2) & (z < 1.2)).all()                   # src/g1_kernel.cu:243-248
    assert m["flip_frac"] < FLIP_FRAC
    assert m["qacc_rel"] < QACC_TOL
    assert m["qpos_abs"] < QPOS_TOL
    assert m["foot_pos_abs"] < 1e-5                        # tests/test_g1sim.py:154-155 bound


def test_randomized_actuation_teacher_forced(device):
    """Randomized actuation, 16,384 envs x 250 steps at dt 2e-3, targets
    resampled every 10 steps; every 5th step checked against the oracle."""
    qp, qv, an, cb = W.randomized_actuation(16384, seed=1, nblocks=25)
    st = GpuState(qp, qv, an, cb[0], device)
    m = _rollout_metrics(st, 250, 2e-3, ctrl_of=lambda s: cb[s // 10], every=5)
    write_metrics("randomized_actuation_16384x250", m)
    assert m["flip_frac"] < FLIP_FRAC
    assert m["qacc_rel"] < QACC_TOL


def test_golden_states(device, golden):
    """The reference library's golden step states (contact branches, limits,
    clamps) in float32 through the GPU vs the oracle on the same inputs."""
    g = golden("step1")
    f = lambda a: np.ascontiguousarray(a.T.astype(np.float32))  # noqa: E731
    st = GpuState(f(g["qpos"]), f(g["qvel"]), f(g["anchor"]), f(g["ctrl"]), device)
    before, after, ref = teacher_forced_step(st, float(g["dt"]))
    flip = branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
    ok = ~flip
    write_metrics("golden_step1", {"qacc_rel": float(rel_err(after["qacc"][:, ok], ref["qacc"][:, ok]).max()),
                                   "flips": int(flip.sum()), "n": int(st.n)})
    assert flip.sum() <= 2
    assert rel_err(after["qacc"][:, ok], ref["qacc"][:, ok]).max() < QACC_TOL
    # the float32-cast inputs are within float rounding of the reference's own
    # inputs, so the reference library's outputs bound the step too
    assert np.abs(after["qpos"][:, ok] - g["qpos_out"].T[:, ok]).max() < 1e-4
    fin = ref["anchor"] < 1e29
    assert np.array_equal(after["anchor"] < 1e29, fin)


def test_nsub_matches_repeated_single_steps(device):
    qp, qv, an, cb = W.randomized_actuation(1000, seed=4, nblocks=1)
    a = GpuState(qp, qv, an, cb[0], device)
    b = GpuState(qp, qv, an, cb[0], device)
    for _ in range(3):
        a.step(7, 2e-3)
        for _ in range(7):
            b.step(1, 2e-3)
    ha, hb = a.host(), b.host()
    for k in ("qpos", "qvel", "anchor", "fn", "foot_pos", "foot_vel", "qacc"):
        assert np.array_equal(ha[k], hb[k]), k


@pytest.mark.parametrize("n", [1, 63, 65, 130])
def test_partial_workgroups(device, n):
    """A tail workgroup computes exactly what a full one does and writes
    nothing past nenv."""
    qp, qv, an, cb = W.randomized_actuation(256, seed=5, nblocks=1)
    full = GpuState(qp, qv, an, cb[0], device)
    full.step(3, 2e-3)
    pad = 64
    sl = lambda a: np.ascontiguousarray(a[:, :n])  # noqa: E731
    part = GpuState(sl(qp), sl(qv), sl(an), sl(cb[0]), device)
    # guard region after every buffer
    big = torch.full((36 * n + pad,), 12345.0, device=device)
    big[:36 * n] = part.qpos.flatten()
    lib = _lib.load()
    stream = torch.cuda.current_stream(device).cuda_stream
    rc = lib.g1_hip_step_ex(n, 3, 2e-3, big.data_ptr(), part.qvel.data_ptr(), part.anchor.data_ptr(),
                            part.ctrl.data_ptr(), None, None, None, None, None, None, None, 0, stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    hf = full.host()
    got = big[:36 * n].view(36, n).cpu().numpy()
    assert np.array_equal(got, hf["qpos"][:, :n])
    assert (big[36 * n:].cpu().numpy() == 12345.0).all()


def test_reset_done_bit_exact(device):
    n = 1000
    lib = _lib.load()
    rng = np.random.default_rng(0)
    done = (rng.random(n) < 0.5).astype(np.uint8)
    for seed, noise, drop in [(42, 0.05, 0.02), (0, 0.0, 0.02), (1234, 0.3, 1.0), (2 ** 32 - 1, 0.1, 0.0)]:
        qpos = torch.full((36, n), 7.0, device=device)
        qvel = torch.full((35, n), 7.0, device=device)
        anc = torch.full((16, n), 7.0, device=device)
        d = torch.as_tensor(done, device=device)
        lib.g1_cuda_reset_done(n, qpos.data_ptr(), qvel.data_ptr(), anc.data_ptr(), d.data_ptr(), seed,
                               noise, drop, torch.cuda.current_stream(device).cuda_stream)
        torch.cuda.synchronize()
        qp, qv, an = qpos.cpu().numpy(), qvel.cpu().numpy(), anc.cpu().numpy()
        for e in range(n):
            if done[e]:
                rp, rv, ra = O.reset_env_f32(e, seed, noise, drop)
                assert np.array_equal(qp[:, e].view(np.uint32), rp.view(np.uint32)), (seed, e)
                assert (qv[:, e] == 0).all() and (an[:, e] == np.float32(1e30)).all()
            else:
                assert (qp[:, e] == 7).all() and (qv[:, e] == 7).all() and (an[:, e] == 7).all()


def test_settles_like_the_oracle(device):
    """1,500 steps of the drop test: every env stands at the oracle's height
    with the robot's weight on the feet (tests/test_vs_mujoco.py:114-133)."""
    qp, qv, an, ct = W.drop_test(512, seed=3)
    st = GpuState(qp, qv, an, ct, device)
    for _ in range(15):
        st.step(100, 2e-3)
    h = st.host()
    weight = MODEL.total_mass * 9.81
    tot = h["fn"].sum(axis=0)
    assert np.abs(tot - weight).max() / weight < 0.02
    e = 0
    o = O.step_full(qp[:, e].astype(float), qv[:, e].astype(float), ct[:, e].astype(float),
                    an[:, e].astype(float), 1500, 2e-3)
    assert abs(h["qpos"][2, e] - o["qpos"][2]) < 2e-3
    assert np.isfinite(h["qpos"]).all()


@pytest.mark.parametrize("flags", [_lib.G1_FLAG_F64_SOLVE, _lib.G1_FLAG_F64_SOLVE | _lib.G1_FLAG_F64_BASE])
def test_f64_islands(device, flags):
    qp, qv, an, ct = W.drop_test(1024, seed=6)
    st = GpuState(qp, qv, an, ct, device)
    if flags & _lib.G1_FLAG_F64_BASE:
        st.enable_f64_base()
    st.step(60, 2e-3, flags=flags)
    before = st.host()
    st.step(1, 2e-3, flags=flags)
    after = st.host()
    ref = O.step_batch_f32(before["qpos"], before["qvel"], before["anchor"], before["ctrl"], 2e-3,
                           want=("qacc", "fn", "anchor"))
    ok = ~branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
    assert rel_err(after["qacc"][:, ok], ref["qacc"][:, ok]).max() < QACC_TOL
    if st.base64 is not None:
        b64 = st.base64.cpu().numpy()
        assert np.abs(b64 - after["qpos"][:7]).max() < 1e-6
        assert np.abs(np.linalg.norm(b64[3:7], axis=0) - 1).max() < 1e-14


def test_per_env_friction(device):
    qp, qv, an, ct = W.drop_test(2048, seed=7)
    mu = np.random.default_rng(1).uniform(0.4, 1.2, 2048).astype(np.float32)
    qv[0:2] = np.random.default_rng(2).uniform(-1, 1, (2, 2048))     # sliding touchdowns
    st = GpuState(qp, qv, an, ct, device, mu=mu)
    st.step(25, 2e-3)
    before, after, ref = teacher_forced_step(st, 2e-3, mu=mu)
    ok = ~branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
    assert ok.mean() > 0.98
    assert rel_err(after["qacc"][:, ok], ref["qacc"][:, ok]).max() < QACC_TOL


def test_total_generalized_force(device):
    """tau_out: servo + limit torques and J^T f of the contacts, against the
    oracle's total generalized force on the same input (pinned on the CPU by
    (M + dt B) qacc + C = tau, tests/test_oracle.py)."""
    qp, qv, an, cb = W.randomized_actuation(2048, seed=8, nblocks=1)
    st = GpuState(qp, qv, an, cb[0], device)
    st.step(40, 2e-3)
    before = st.host()
    st.step(1, 2e-3, tau=True)
    after = st.host()
    ref = O.step_batch_f32(before["qpos"], before["qvel"], before["anchor"], before["ctrl"], 2e-3,
                           want=("tau", "fn", "anchor"))
    ok = ~branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
    scale = np.maximum(np.abs(ref["tau"][:, ok]), 1.0)
    err = (np.abs(after["tau"][:, ok] - ref["tau"][:, ok]) / scale).max()
    write_metrics("tau_total_2048", {"rel_err": float(err), "ok_frac": float(ok.mean())})
    assert ok.mean() > 0.98
    assert err < 1e-3
    assert np.abs(after["tau"][:6]).max() > 1.0          # contacts load the base
