"""GPU parity of the fused step (libg1sim_amd.so through its C-ABI) against
the FP64 oracle on the same float32 inputs (teacher forcing).

Tolerances (float32 step vs float64 oracle, same input state):
  qacc        rel. err |a - ref| / max(|ref|, 1) (tests/test_vs_mujoco.py:74-75)
              on every env whose contact decisions agree: at most F32_FACTOR
              times the error of the reference algorithm evaluated in float32
              on the same states (max, p99 and median); below AIRBORNE_TOL
              without contacts; below F64_CONTACT_TOL with the float64
              contact-geometry island;
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

from .gpu_helpers import (GpuState, branch_flips, f32_reference_step, rel_err, teacher_forced_step,
                          write_metrics)

pytestmark = pytest.mark.gpu

QPOS_TOL = 2e-5          # next-state qpos abs. error after one step
FLIP_FRAC = 0.01         # envs with a contact decision at a threshold
# The stiff foot contacts make qacc sensitive to the penetration depth at the
# 1e-2 level per float32 ulp of the base height (measured on the FP64 oracle),
# so a float32 step cannot be held to a fixed 1e-5: the bar for the float32
# kernel is the reference algorithm itself evaluated in float32 on the same
# states (liboracle32), and the float64 contact-geometry island is held to
# F64_CONTACT_TOL.
F32_FACTOR = 2.0         # GPU error may not exceed 2x the float32 reference algorithm's
AIRBORNE_TOL = 1e-3      # no contacts: plain ABA in float32
F64_CONTACT_TOL = 2e-3   # measured 3.7e-4 on the drop test (float32 reference: 4.5e-2)


def _rollout_metrics(st, steps, dt, ctrl_of=None, mu=None, every=1, flags=0):
    worst = {"qpos_abs": 0.0, "qvel_rel": 0.0, "fn_abs": 0.0, "foot_pos_abs": 0.0, "foot_vel_abs": 0.0}
    errs, errs32 = [], []
    flips = 0
    checked = 0
    for s in range(steps):
        if ctrl_of is not None:
            st.ctrl.copy_(torch.as_tensor(ctrl_of(s), device=st.device))
        if s % every:
            st.step(1, dt, flags=flags)
            continue
        before = st.host()
        st.step(1, dt, flags=flags)
        after = st.host()
        ref = O.step_batch_f32(before["qpos"], before["qvel"], before["anchor"], before["ctrl"], dt, mu=mu,
                               want=("qacc", "qpos", "qvel", "anchor", "fn", "foot_pos", "foot_vel"))
        f32 = f32_reference_step(before, dt, mu)
        flip = branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
        flip32 = branch_flips(f32["fn"], ref["fn"], f32["anchor"], ref["anchor"])
        ok = ~flip
        flips += int(flip.sum())
        checked += st.n
        errs.append(rel_err(after["qacc"][:, ok], ref["qacc"][:, ok]).max(axis=0))
        errs32.append(rel_err(f32["qacc"][:, ~flip32], ref["qacc"][:, ~flip32]).max(axis=0))
        worst["qpos_abs"] = max(worst["qpos_abs"], float(np.abs(after["qpos"][:, ok] - ref["qpos"][:, ok]).max(initial=0)))
        worst["qvel_rel"] = max(worst["qvel_rel"], float(rel_err(after["qvel"][:, ok], ref["qvel"][:, ok]).max(initial=0)))
        worst["fn_abs"] = max(worst["fn_abs"], float(np.abs(after["fn"][:, ok] - ref["fn"][:, ok]).max(initial=0)))
        worst["foot_pos_abs"] = max(worst["foot_pos_abs"], float(np.abs(after["foot_pos"] - ref["foot_pos"]).max()))
        worst["foot_vel_abs"] = max(worst["foot_vel_abs"], float(np.abs(after["foot_vel"] - ref["foot_vel"]).max()))
    e, e32 = np.concatenate(errs), np.concatenate(errs32)
    worst.update({"qacc_rel_max": float(e.max()), "qacc_rel_p99": float(np.percentile(e, 99)),
                  "qacc_rel_median": float(np.median(e)), "f32ref_qacc_rel_max": float(e32.max()),
                  "f32ref_qacc_rel_p99": float(np.percentile(e32, 99)),
                  "f32ref_qacc_rel_median": float(np.median(e32)),
                  "flip_frac": flips / max(checked, 1), "checked_env_steps": checked, "flags": flags})
    return worst


def _assert_float32_bar(m):
    assert m["flip_frac"] < FLIP_FRAC
    assert m["qacc_rel_max"] <= F32_FACTOR * m["f32ref_qacc_rel_max"]
    assert m["qacc_rel_p99"] <= F32_FACTOR * m["f32ref_qacc_rel_p99"]
    assert m["qacc_rel_median"] <= F32_FACTOR * m["f32ref_qacc_rel_median"]


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
    _assert_float32_bar(m)
    assert m["qpos_abs"] < QPOS_TOL
    assert m["foot_pos_abs"] < 1e-5                        # tests/test_g1sim.py:154-155 bound


def test_drop_test_f64_contact_island(device):
    """Same rollout with the float64 foot-contact geometry (+ float64 base
    solve): the contact-depth sensitivity is removed."""
    flags = _lib.G1_FLAG_F64_CONTACT | _lib.G1_FLAG_F64_SOLVE
    qp, qv, an, ct = W.drop_test(1024, seed=0)
    st = GpuState(qp, qv, an, ct, device)
    m = _rollout_metrics(st, 250, 2e-3, flags=flags)
    write_metrics("drop_test_1024x250_f64contact", m)
    assert m["flip_frac"] < FLIP_FRAC
    assert m["qacc_rel_max"] < F64_CONTACT_TOL
    # measured ratios 0.008 (max) and 0.014 (median); a world pose kept in
    # float inside the island measured 0.27 and 0.29
    assert m["qacc_rel_max"] < 0.1 * m["f32ref_qacc_rel_max"]
    assert m["qacc_rel_median"] < 0.1 * m["f32ref_qacc_rel_median"]


def test_airborne_plain_aba(device):
    """No contacts (base 1 m up), randomized joint velocities: the float32 ABA
    alone against float64."""
    qp, qv, an, cb = W.randomized_actuation(4096, seed=9, nblocks=1)
    qp[2] += 1.0
    st = GpuState(qp, qv, an, cb[0], device)
    before, after, ref = teacher_forced_step(st, 2e-3)
    assert (after["fn"] == 0).all()
    e = rel_err(after["qacc"], ref["qacc"])
    write_metrics("airborne_4096", {"qacc_rel_max": float(e.max()), "qacc_rel_median": float(np.median(e))})
    assert e.max() < AIRBORNE_TOL


def test_randomized_actuation_teacher_forced(device):
    """Randomized actuation, 16,384 envs x 250 steps at dt 2e-3, targets
    resampled every 10 steps; every 5th step checked against the oracle."""
    qp, qv, an, cb = W.randomized_actuation(16384, seed=1, nblocks=25)
    st = GpuState(qp, qv, an, cb[0], device)
    m = _rollout_metrics(st, 250, 2e-3, ctrl_of=lambda s: cb[s // 10], every=5)
    write_metrics("randomized_actuation_16384x250", m)
    _assert_float32_bar(m)


def test_golden_states(device, golden):
    """The reference library's golden step states (contact branches, limits,
    clamps) in float32 through the GPU vs the oracle on the same inputs."""
    g = golden("step1")
    f = lambda a: np.ascontiguousarray(a.T.astype(np.float32))  # noqa: E731
    st = GpuState(f(g["qpos"]), f(g["qvel"]), f(g["anchor"]), f(g["ctrl"]), device)
    before, after, ref = teacher_forced_step(st, float(g["dt"]))
    f32 = f32_reference_step(before, float(g["dt"]))
    flip = branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
    ok = ~flip
    e = rel_err(after["qacc"][:, ok], ref["qacc"][:, ok]).max()
    e32 = rel_err(f32["qacc"][:, ok], ref["qacc"][:, ok]).max()
    write_metrics("golden_step1", {"qacc_rel_max": float(e), "f32ref_qacc_rel_max": float(e32),
                                   "flips": int(flip.sum()), "n": int(st.n)})
    assert flip.sum() <= 2
    assert e <= F32_FACTOR * e32
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


def test_reset_done_matches_reference_fixture(device, golden):
    """The GPU reset against the reference kernel's own output
    (tests/golden/reset.npz: g1_reset_done_kernel, src/g1_kernel.cu:85-103,
    over a 65,536-env batch), bit for bit, for every fixture seed, noise and
    drop."""
    g = golden("reset")
    n, envs, rows = int(g["nenv"]), g["envs"], g["rows"]
    lib = _lib.load()
    done = np.zeros(n, np.uint8)
    done[envs] = 1
    d = torch.as_tensor(done, device=device)
    qpos = torch.empty((36, n), device=device)
    qvel = torch.empty((35, n), device=device)
    anc = torch.empty((16, n), device=device)
    for a, seed in enumerate(g["seeds"]):
        for b, noise in enumerate(g["noises"]):
            for c, drop in enumerate(g["drops"]):
                for t in (qpos, qvel, anc):
                    t.fill_(7.0)
                lib.g1_cuda_reset_done(n, qpos.data_ptr(), qvel.data_ptr(), anc.data_ptr(), d.data_ptr(), int(seed),
                                       float(noise), float(drop), torch.cuda.current_stream(device).cuda_stream)
                torch.cuda.synchronize()
                got = qpos[:, torch.as_tensor(envs, device=device)].cpu().numpy()
                want = g["qpos"][a, b, c]
                assert np.array_equal(got[rows].view(np.uint32), want.view(np.uint32)), (seed, noise, drop)
                assert int((qpos != 7).sum()) == 36 * envs.size      # only the done envs are written
                assert bool((qvel[:, torch.as_tensor(envs, device=device)] == 0).all())


@pytest.mark.parametrize("nenv,step0,nsteps", [(4096, 3, 37), (40000, 20, 25), (100, 7, 1205)])
def test_step_schedule_matches_launches(device, nenv, step0, nsteps):
    """g1_hip_step_schedule (one host call; launches of up to 1,000
    substeps that switch target blocks themselves, the next block's targets
    loaded a substep ahead) gives the bits of g1_hip_step_ex launched block
    by block with the same targets, at both work splits, with per-env
    friction and feet outputs; 1,205 steps take two launches."""
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    nblocks = (step0 + nsteps + 9) // 10
    qp, qv, an, cb = W.randomized_actuation(nenv, seed=14, nblocks=nblocks)
    mu = np.random.default_rng(15).uniform(0.4, 1.2, nenv).astype(np.float32)
    ctrl = torch.as_tensor(cb, device=device)
    a = GpuState(qp, qv, an, cb[0], device, mu=mu)
    b = GpuState(qp, qv, an, cb[0], device, mu=mu)
    rc = lib.g1_hip_step_schedule(nenv, nsteps, 10, step0, 2e-3, a.qpos.data_ptr(), a.qvel.data_ptr(),
                                  a.anchor.data_ptr(), ctrl.data_ptr(), 29 * nenv, a.foot_pos.data_ptr(),
                                  a.foot_vel.data_ptr(), a.fn.data_ptr(), a.mu.data_ptr(), None, 0, s)
    _lib.check(rc, "step_schedule")
    k = step0
    while k < step0 + nsteps:
        n = min(10 - k % 10, step0 + nsteps - k)
        b.ctrl = ctrl[k // 10]
        b.step(n, 2e-3, qacc=False)
        k += n
    ha, hb = a.host(), b.host()
    for key in ("qpos", "qvel", "anchor", "foot_pos", "foot_vel", "fn"):
        assert np.array_equal(ha[key], hb[key]), key
    # argument checks return before any launch
    assert lib.g1_hip_step_schedule(nenv, 5, 0, 0, 2e-3, a.qpos.data_ptr(), a.qvel.data_ptr(), a.anchor.data_ptr(),
                                    ctrl.data_ptr(), 29 * nenv, None, None, None, None, None, 0, s) == -1


def test_settles_like_the_oracle(device):
    """10 s (5,000 steps at dt 2e-3) of the drop test: every env stands with
    the robot's weight on the feet (tests/test_vs_mujoco.py:114-133)."""
    qp, qv, an, ct = W.drop_test(512, seed=3)
    st = GpuState(qp, qv, an, ct, device)
    for _ in range(50):
        st.step(100, 2e-3)
    h = st.host()
    weight = MODEL.total_mass * 9.81
    tot = h["fn"].sum(axis=0)
    assert np.abs(tot - weight).max() / weight < 0.02
    e = 0
    o = O.step_full(qp[:, e].astype(float), qv[:, e].astype(float), ct[:, e].astype(float),
                    an[:, e].astype(float), 5000, 2e-3)
    assert abs(h["qpos"][2, e] - o["qpos"][2]) < 2e-3
    assert np.isfinite(h["qpos"]).all()


@pytest.mark.parametrize("flags", [
    _lib.G1_FLAG_F64_SOLVE,
    _lib.G1_FLAG_F64_SOLVE | _lib.G1_FLAG_F64_BASE,
    # the 2-wave split's float64 instances (chosen by default from 32,768 envs)
    _lib.G1_FLAG_F64_SOLVE | _lib.G1_FLAG_F64_BASE | _lib.G1_FLAG_WAVES2,
    _lib.G1_FLAG_F64_SOLVE | _lib.G1_FLAG_F64_CONTACT | _lib.G1_FLAG_WAVES2,
])
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
    f32 = f32_reference_step(before, 2e-3)
    ok = ~branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
    assert rel_err(after["qacc"][:, ok], ref["qacc"][:, ok]).max() <= \
        F32_FACTOR * rel_err(f32["qacc"][:, ok], ref["qacc"][:, ok]).max()
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
    f32 = f32_reference_step(before, 2e-3, mu=mu)
    ok = ~branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
    assert ok.mean() > 0.98
    assert rel_err(after["qacc"][:, ok], ref["qacc"][:, ok]).max() <= \
        F32_FACTOR * rel_err(f32["qacc"][:, ok], ref["qacc"][:, ok]).max()


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
    t32 = np.zeros((35, st.n))
    f32 = [np.ascontiguousarray(before[k], np.float32) for k in ("qpos", "qvel", "anchor", "ctrl")]
    O.lib32().g1o_step_batch_f32(st.n, 2e-3, *(O.fptr(a) for a in f32), None, None, None, None, None, None,
                                 None, None, O.dptr(t32), 0)
    ok = ~branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
    scale = np.maximum(np.abs(ref["tau"][:, ok]), 1.0)
    err = (np.abs(after["tau"][:, ok] - ref["tau"][:, ok]) / scale).max()
    err32 = (np.abs(t32[:, ok] - ref["tau"][:, ok]) / scale).max()
    write_metrics("tau_total_2048", {"rel_err": float(err), "f32ref_rel_err": float(err32), "ok_frac": float(ok.mean())})
    assert ok.mean() > 0.98
    assert err <= F32_FACTOR * err32
    assert np.abs(after["tau"][:6]).max() > 1.0          # contacts load the base


@pytest.mark.parametrize("force", [_lib.G1_FLAG_WAVES2, _lib.G1_FLAG_WAVES4])
def test_work_splits_agree(device, force):
    """2 waves (legs | arms) and 4 waves (one chain each) per 64 envs compute
    the same step: both within the float32 bar of the oracle."""
    qp, qv, an, cb = W.randomized_actuation(4096, seed=12, nblocks=1)
    st = GpuState(qp, qv, an, cb[0], device)
    st.step(30, 2e-3, flags=force)
    before = st.host()
    st.step(1, 2e-3, flags=force, tau=True)
    after = st.host()
    ref = O.step_batch_f32(before["qpos"], before["qvel"], before["anchor"], before["ctrl"], 2e-3,
                           want=("qacc", "fn", "anchor", "tau"))
    f32 = f32_reference_step(before, 2e-3)
    ok = ~branch_flips(after["fn"], ref["fn"], after["anchor"], ref["anchor"])
    e = rel_err(after["qacc"][:, ok], ref["qacc"][:, ok]).max()
    e32 = rel_err(f32["qacc"][:, ok], ref["qacc"][:, ok]).max()
    write_metrics(f"split_flag{force}", {"qacc_rel_max": float(e), "f32ref_qacc_rel_max": float(e32)})
    assert e <= F32_FACTOR * e32
    assert np.abs(after["tau"][:6]).max() > 1.0


def test_empty_and_zero_substep_calls_leave_state(device):
    """nsub = 0 and nenv = 0 are no-ops on the device (every entry point
    accepts them), and a judge list with no entries writes nothing."""
    qp, qv, an, cb = W.randomized_actuation(300, seed=41, nblocks=1)
    st = GpuState(qp, qv, an, cb[0], device)
    before = st.host()
    st.step(0, 2e-3)
    after = st.host()
    for k in ("qpos", "qvel", "anchor"):
        assert np.array_equal(before[k], after[k]), k
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    assert lib.g1_hip_step_ex(0, 3, 2e-3, None, None, None, None, None, None, None, None, None, None, None, 0, s) == 0
    lib.g1_cuda_reset_done(0, None, None, None, None, 1, 0.1, 0.02, s)
    lst = torch.zeros(300, dtype=torch.int32, device=device)
    cnt = torch.full((1,), 123, dtype=torch.int32, device=device)
    fb = torch.zeros(300, dtype=torch.uint8, device=device)
    # period 400 > nenv and phase 350: nobody is judged
    _lib.check(lib.g1_hip_judge_list(300, 350, 400, fb.data_ptr(), lst.data_ptr(), cnt.data_ptr(), s))
    qacc = torch.full((35, 300), 5.0, device=device)
    _lib.check(lib.g1_hip_step_judge(300, 2e-3, st.qpos.data_ptr(), st.qvel.data_ptr(), st.anchor.data_ptr(),
                                     st.ctrl.data_ptr(), None, lst.data_ptr(), cnt.data_ptr(), qacc.data_ptr(), 0, s))
    torch.cuda.synchronize()
    assert int(cnt.item()) == 0
    assert (qacc == 5.0).all()
    assert np.array_equal(st.host()["qpos"], before["qpos"])
