"""The learned forward-dynamics surrogate on the GPU: its force kernel and
ABA judge against the fused step, the integrate/fallback logic, and what a
trained network gains over the untrained one and over ABA alone."""
import numpy as np
import pytest
import torch

from g1sim_amd import _lib, workloads as W
from g1sim_amd.surrogate import (MLP, NOUT, POUT, SurrogateStepper, evaluate, fit_normalization,
                                 generate_dataset, train)

from .gpu_helpers import GpuState, write_metrics

pytestmark = pytest.mark.gpu


def _state(device, n, seed, steps=40):
    qp, qv, an, cb = W.randomized_actuation(n, seed=seed, nblocks=1)
    st = GpuState(qp, qv, an, cb[0], device)
    st.step(steps, 2e-3)
    torch.cuda.synchronize()
    return st


def judged_envs(n, phase, period=16):
    """The rotating judge's envs: blocks of 64 with block % period == phase."""
    e = np.arange(n)
    return e[(e // 64) % period == phase]


def same_at_f32(a, b, what, unset=1e9):
    """Two float32 evaluations of the same quantity by different kernel
    instances of one code (the compiler contracts multiply-adds differently
    where an instance stores fewer outputs): where one is set, the other is
    set, and they agree to float32 rounding -- |a - b| / max(|b|, 1) with
    median <= 1e-6, p99 <= 1e-4 and max <= 3e-2, the float32 bar of the
    stiff contacts (one ulp of foot height moves qacc by ~1e-2)."""
    assert np.array_equal(a == unset, b == unset), what
    m = b != unset
    if not m.any():
        return
    r = (np.abs(a[m] - b[m]) / np.maximum(np.abs(b[m]), 1.0)).astype(np.float64)
    assert np.median(r) <= 1e-6 and np.quantile(r, 0.99) <= 1e-4 and r.max() <= 3e-2, (what, np.median(r),
                                                                                          np.quantile(r, 0.99),
                                                                                          r.max())


def _judge_blocks(st, fallback, phase, qacc, device, flags=0):
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    _lib.check(lib.g1_hip_step_judge_blocks(st.n, 2e-3, st.qpos.data_ptr(), st.qvel.data_ptr(),
                                            st.anchor.data_ptr(), st.ctrl.data_ptr(), None, fallback.data_ptr(),
                                            phase, 16, qacc.data_ptr(), flags, s))


def test_forces_kernel_matches_step(device):
    """g1_hip_forces gives the step's tau_total, anchors, normal forces and
    foot kinematics."""
    st = _state(device, 4096, 21)
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    anchor2 = st.anchor.clone()
    tau = torch.zeros((35, st.n), device=device)
    fn = torch.zeros((8, st.n), device=device)
    fp = torch.zeros((6, st.n), device=device)
    fv = torch.zeros((6, st.n), device=device)
    rc = lib.g1_hip_forces(st.n, 2e-3, st.qpos.data_ptr(), st.qvel.data_ptr(), anchor2.data_ptr(),
                           st.ctrl.data_ptr(), None, tau.data_ptr(), fn.data_ptr(), fp.data_ptr(), fv.data_ptr(), s)
    _lib.check(rc)
    st.step(1, 2e-3, tau=True)
    h = st.host()
    np.testing.assert_allclose(fp.cpu().numpy(), h["foot_pos"], rtol=0, atol=2e-6)
    np.testing.assert_allclose(fv.cpu().numpy(), h["foot_vel"], rtol=1e-5, atol=2e-5)
    t = tau.cpu().numpy()
    scale = np.maximum(np.abs(h["tau"]), 1.0)
    assert (np.abs(t - h["tau"]) / scale).max() < 1e-3
    assert np.abs(fn.cpu().numpy() - h["fn"]).max() < 1e-2
    a2 = anchor2.cpu().numpy()
    assert np.array_equal(a2 >= 1e29, h["anchor"] >= 1e29)
    fin = h["anchor"] < 1e29
    assert np.abs(a2[fin] - h["anchor"][fin]).max(initial=0) < 1e-5


def test_judge_matches_step_qacc(device):
    """The judge mode steps exactly the listed envs, reads the state and
    writes only their qacc."""
    st = _state(device, 5000, 22)
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    fallback = torch.zeros(st.n, dtype=torch.uint8, device=device)
    fallback[::97] = 1
    lst = torch.zeros(st.n, dtype=torch.int32, device=device)
    cnt = torch.zeros(1, dtype=torch.int32, device=device)
    _lib.check(lib.g1_hip_judge_list(st.n, 3, 16, fallback.data_ptr(), lst.data_ptr(), cnt.data_ptr(), s))
    qacc = torch.full((35, st.n), 1e9, device=device)
    before = st.host()
    _lib.check(lib.g1_hip_step_judge(st.n, 2e-3, st.qpos.data_ptr(), st.qvel.data_ptr(), st.anchor.data_ptr(),
                                     st.ctrl.data_ptr(), None, lst.data_ptr(), cnt.data_ptr(), qacc.data_ptr(), 0, s))
    torch.cuda.synchronize()
    after = st.host()
    for k in ("qpos", "qvel", "anchor"):
        assert np.array_equal(before[k], after[k]), k
    c = int(cnt.item())
    e = np.sort(lst[:c].cpu().numpy())
    expect = np.union1d(judged_envs(st.n, 3), np.nonzero(fallback.cpu().numpy())[0])
    assert np.array_equal(e, expect)
    # the list-free form steps the same envs to the same bits
    qacc_b = torch.full((35, st.n), 1e9, device=device)
    _judge_blocks(st, fallback, 3, qacc_b, device)
    torch.cuda.synchronize()
    assert np.array_equal(before["qpos"], st.host()["qpos"])
    st.step(1, 2e-3)
    ref = st.host()["qacc"]
    q = qacc.cpu().numpy()
    assert np.array_equal(q[:, e], ref[:, e])
    others = np.setdiff1d(np.arange(st.n), e)
    assert (q[:, others] == 1e9).all()
    same_at_f32(qacc_b.cpu().numpy(), q, "blocks vs list")


def _stepper_with_oracle_pred(device, n, seed, corrupt=None, accepted_before=True):
    """A stepper whose 'network' output is ABA's own qacc (z-scored), or a
    corrupted copy, to test the integrate/judge/fallback logic alone.  With
    accepted_before, every env starts out of fallback (as after an accepted
    judgement); otherwise from the stepper's fail-safe start (all in
    fallback)."""
    st = _state(device, n, seed)
    mlp = MLP(device, seed=0)
    stepper = SurrogateStepper(mlp, n, device, period=16, tol=1e-3)
    assert bool((stepper.fallback == 1).all())          # fail-safe start
    if accepted_before:
        stepper.fallback.zero_()
    ref = GpuState(st.qpos.cpu().numpy(), st.qvel.cpu().numpy(), st.anchor.cpu().numpy(), st.ctrl.cpu().numpy(),
                   device)
    ref.step(1, 2e-3)
    qa = ref.qacc.clone()
    if corrupt is not None:
        qa[:, corrupt] += 1.0
    stepper.pred[:n, :NOUT] = ((qa - mlp.ymean[:, None]) * mlp.yinv[:, None]).T
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    _judge_blocks(st, stepper.fallback, 0, stepper.qacc_aba, device)
    _lib.check(lib.g1_hip_surrogate_integrate(n, 2e-3, 0, 16, 1e-3, st.qpos.data_ptr(), st.qvel.data_ptr(),
                                              stepper.pred.data_ptr(), POUT, mlp.ymean.data_ptr(),
                                              stepper.ystd.data_ptr(), stepper.qacc_aba.data_ptr(),
                                              stepper.fallback.data_ptr(), stepper.stats.data_ptr(), None, s))
    torch.cuda.synchronize()
    return st, ref, stepper


def test_integrate_with_exact_prediction(device):
    st, ref, stepper = _stepper_with_oracle_pred(device, 4096, 23)
    a, b = st.host(), ref.host()
    assert np.abs(a["qpos"] - b["qpos"]).max() < 1e-5
    assert np.abs(a["qvel"] - b["qvel"]).max() < 1e-3
    sm = stepper.stats.cpu().numpy()
    assert sm[0] == 4096 // 16 and sm[1] == 0
    assert int(stepper.fallback.sum()) == 0


def test_fail_safe_start(device):
    """A new stepper integrates no unjudged network output: every env starts
    in fallback (ABA), and only the envs of the judged blocks whose
    prediction the judge accepts leave it."""
    bad = np.arange(0, 4096, 3)
    st, ref, stepper = _stepper_with_oracle_pred(device, 4096, 30, corrupt=bad, accepted_before=False)
    a, b = st.host(), ref.host()
    fb = stepper.fallback.cpu().numpy().astype(bool)
    judged = judged_envs(4096, 0)
    accepted = np.setdiff1d(judged, bad)
    assert np.array_equal(np.nonzero(~fb)[0], accepted)
    # every env was stepped with ABA's qacc (judged envs: the judge's; the
    # others: fallback), so all match the fused step
    assert np.abs(a["qpos"] - b["qpos"]).max() < 1e-5
    sm = stepper.stats.cpu().numpy()
    assert sm[0] == judged.size and sm[1] == np.intersect1d(judged, bad).size and sm[3] == 4096 - judged.size


def test_judge_rejects_and_falls_back(device):
    bad = np.arange(0, 4096, 2)                      # every other env predicted wrong
    st, ref, stepper = _stepper_with_oracle_pred(device, 4096, 24, corrupt=bad)
    a, b = st.host(), ref.host()
    fb = stepper.fallback.cpu().numpy().astype(bool)
    judged = judged_envs(4096, 0)
    rejected = np.intersect1d(judged, bad)
    assert np.array_equal(np.nonzero(fb)[0], rejected)   # exactly the judged envs predicted wrong
    # judged envs used ABA's qacc: their state matches the fused step
    assert np.abs(a["qpos"][:, judged] - b["qpos"][:, judged]).max() < 1e-5
    sm = stepper.stats.cpu().numpy()
    assert sm[0] == judged.size and sm[1] == rejected.size


def test_trained_network_gain(device):
    """A network trained by the project's kernels on ABA transitions: its
    held-out accuracy against the untrained network, and the surrogate +
    judge path against plain ABA over a short rollout (error stays at the
    judge's tolerance; throughput and fallback rate are reported)."""
    ds = generate_dataset(device, nenv=8192, steps=160, seed=31)
    mlp = MLP(device, seed=1)
    fit_normalization(mlp, ds)
    before = evaluate(mlp, ds, nsamp=32768)
    hist = train(mlp, ds, iters=1500, batch=8192, lr=1e-3)
    after = evaluate(mlp, ds, nsamp=32768)
    # rollout: surrogate + rotating judge vs plain ABA from the same start
    n = 8192
    qp, qv, an, cb = W.randomized_actuation(n, seed=77, nblocks=5)
    a = GpuState(qp, qv, an, cb[0], device)
    b = GpuState(qp, qv, an, cb[0], device)
    stepper = SurrogateStepper(mlp, n, device)
    torch.cuda.synchronize()
    t_sur = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_sur[0].record()
    for k in range(50):
        a.ctrl.copy_(torch.as_tensor(cb[k // 10], device=device))
        stepper.step(a.qpos, a.qvel, a.anchor, a.ctrl)
    t_sur[1].record()
    t_aba = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_aba[0].record()
    for k in range(50):
        b.ctrl.copy_(torch.as_tensor(cb[k // 10], device=device))
        b.step(1, 2e-3, qacc=False)
    t_aba[1].record()
    torch.cuda.synchronize()
    summ = stepper.summary()
    ha, hb = a.host(), b.host()
    drift = float(np.abs(ha["qpos"] - hb["qpos"]).max())
    m = {"untrained": before, "trained": after, "loss_first": hist[0][1], "loss_last": hist[-1][1],
         "rollout": summ, "rollout_qpos_drift_vs_aba": drift,
         "ms_per_step_surrogate": t_sur[0].elapsed_time(t_sur[1]) / 50,
         "ms_per_step_aba": t_aba[0].elapsed_time(t_aba[1]) / 50, "nenv": n}
    write_metrics("surrogate_gain_8192", m)
    assert hist[-1][1] < 0.2 * hist[0][1]                       # it learns
    assert after["median_rel_err"] < 0.2 * before["median_rel_err"]
    assert np.isfinite(ha["qpos"]).all()
    # every judged block is judged at once or recorded for its period's batch
    assert summ["judged"] + summ["pending_records"] + summ["dropped_records"] == 50 * n // 16


def _trained_small(device, seed=4):
    ds = generate_dataset(device, nenv=4096, steps=40, seed=seed)
    mlp = MLP(device, seed=seed)
    fit_normalization(mlp, ds)
    train(mlp, ds, iters=100, batch=4096)
    return mlp


def test_fused_forward_equals_layer_gemms(device):
    """The fused kernel (features + 5 layers, activations in LDS) gives the
    same fp32 outputs as the layer-by-layer MFMA GEMMs: same bf16 operands,
    same accumulation order."""
    mlp = _trained_small(device)
    st = _state(device, 1000, 25)                  # 1000 envs: a partial last workgroup
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    tau = torch.zeros((35, st.n), device=device)
    _lib.check(lib.g1_hip_forces(st.n, 2e-3, st.qpos.data_ptr(), st.qvel.data_ptr(), st.anchor.clone().data_ptr(),
                                 st.ctrl.data_ptr(), None, tau.data_ptr(), None, None, None, s))
    rows = 1024
    X16 = torch.zeros((rows, 128), dtype=torch.bfloat16, device=device)
    _lib.check(lib.g1_mlp_features(st.n, rows, st.qpos.data_ptr(), st.qvel.data_ptr(), tau.data_ptr(),
                                   mlp.xmean.data_ptr(), mlp.xinv.data_ptr(), X16.data_ptr(), 128, s))
    ref = mlp.forward(X16, rows, train=False).clone()
    fused = torch.full((st.n, 64), 7.0, device=device)
    mlp.forward_state(st.n, st.qpos, st.qvel, tau, fused)
    torch.cuda.synchronize()
    a, b = fused.cpu().numpy(), ref[:st.n, :64].cpu().numpy()
    diff = np.abs(a - b)
    write_metrics("surrogate_fused_forward", {"max_abs_diff": float(diff.max()), "equal_frac": float((a == b).mean())})
    assert (a == b).mean() > 0.999
    assert diff.max() <= 1e-5 * max(1.0, np.abs(b).max())


def test_fused_step_equals_unfused(device):
    """Surrogate steps through the fused kernel and through features + GEMMs +
    integrate agree on state, fallback flags and statistics (immediate
    judging in both: the unfused path has no deferred mode)."""
    mlp = _trained_small(device, seed=6)
    n = 3000
    qp, qv, an, cb = W.randomized_actuation(n, seed=26, nblocks=1)
    runs = []
    for fused in (True, False):
        st = GpuState(qp, qv, an, cb[0], device)
        stepper = SurrogateStepper(mlp, n, device, period=16, tol=0.5, fused=fused, defer=False)
        assert stepper.fused == fused
        for _ in range(5):
            stepper.step(st.qpos, st.qvel, st.anchor, st.ctrl)
        torch.cuda.synchronize()
        runs.append((st.host(), stepper.fallback.cpu().numpy(), stepper.stats.cpu().numpy()))
    (ha, fa, sa), (hb, fb, sb) = runs
    for k in ("qpos", "qvel", "anchor"):
        d = np.abs(ha[k] - hb[k])
        fin = np.isfinite(hb[k])
        assert np.array_equal(np.isfinite(ha[k]), fin), k
        assert d[fin].max(initial=0) <= 1e-4 * max(1.0, np.abs(hb[k][fin]).max(initial=0)), k
    assert (fa != fb).sum() <= 2
    assert abs(int(sa[0]) - int(sb[0])) == 0 and abs(int(sa[1]) - int(sb[1])) <= 2


@pytest.mark.parametrize("n", [40000])
def test_judge_two_wave_split_matches_step(device, n):
    """From 32,768 envs the judge and the step run the 2-wave split; the
    judged envs' qacc still equals the step's bit for bit."""
    st = _state(device, n, 27, steps=5)
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    fallback = torch.zeros(n, dtype=torch.uint8, device=device)
    fallback[::5] = 1
    lst = torch.zeros(n, dtype=torch.int32, device=device)
    cnt = torch.zeros(1, dtype=torch.int32, device=device)
    _lib.check(lib.g1_hip_judge_list(n, 7, 16, fallback.data_ptr(), lst.data_ptr(), cnt.data_ptr(), s))
    qacc = torch.full((35, n), 1e9, device=device)
    _lib.check(lib.g1_hip_step_judge(n, 2e-3, st.qpos.data_ptr(), st.qvel.data_ptr(), st.anchor.data_ptr(),
                                     st.ctrl.data_ptr(), None, lst.data_ptr(), cnt.data_ptr(), qacc.data_ptr(), 0, s))
    torch.cuda.synchronize()
    c = int(cnt.item())
    e = lst[:c].cpu().numpy()
    assert np.array_equal(np.sort(e), np.union1d(judged_envs(n, 7), np.arange(0, n, 5)))
    qacc_b = torch.full((35, n), 1e9, device=device)
    _judge_blocks(st, fallback, 7, qacc_b, device)
    st.step(1, 2e-3)
    ref = st.host()["qacc"]
    assert np.array_equal(qacc.cpu().numpy()[:, e], ref[:, e])
    same_at_f32(qacc_b.cpu().numpy(), qacc.cpu().numpy(), "blocks vs list")


def test_fallback_tiles_skip_the_network(device):
    """Blocks of 64 envs all in fallback and not judged this step skip the
    network in the fused step; the result is the same as through features +
    GEMMs + integrate, which run the network everywhere."""
    mlp = _trained_small(device, seed=8)
    n = 64 * 40 + 17
    qp, qv, an, cb = W.randomized_actuation(n, seed=28, nblocks=1)
    fb0 = np.zeros(n, np.uint8)
    for blk in (1, 2, 5, 17, 33):                 # whole blocks, none judged at phase 0
        fb0[64 * blk:64 * blk + 64] = 1
    fb0[64 * 3:64 * 3 + 63] = 1                   # one env short of a whole block
    fb0[64 * 16:64 * 16 + 64] = 1                 # whole block, judged at phase 0
    fb0[64 * 40:] = 1                             # the partial tail block
    runs = []
    for fused in (True, False):
        st = GpuState(qp, qv, an, cb[0], device)
        # immediate judging: the judged all-fallback block runs the network
        stepper = SurrogateStepper(mlp, n, device, period=16, tol=1e30, fused=fused, defer=False)
        stepper.fallback.copy_(torch.as_tensor(fb0, device=device))
        stepper.step(st.qpos, st.qvel, st.anchor, st.ctrl)
        torch.cuda.synchronize()
        runs.append((st.host(), stepper.fallback.cpu().numpy(), stepper.stats.cpu().numpy()))
    (ha, fa, sa), (hb, fb, sb) = runs
    skipped = np.zeros(n, bool)
    for blk in (1, 2, 5, 17, 33, 40):
        skipped[64 * blk:64 * blk + 64] = True
    for k in ("qpos", "qvel", "anchor"):
        assert np.array_equal(ha[k][:, skipped], hb[k][:, skipped]), k
        d = np.abs(ha[k] - hb[k])
        assert d.max() <= 1e-4 * max(1.0, np.abs(hb[k]).max()), k
    assert np.array_equal(fa, fb)
    assert np.array_equal(sa[[0, 1, 3]], sb[[0, 1, 3]])
    # tol 1e30 accepts every judgement: block 16 (judged) leaves fallback
    assert not fa[64 * 16:64 * 17].any() and fa[64 * 3:64 * 3 + 63].all()
    assert sa[3] == int(fb0.sum()) - 64           # fallback env-steps: all but the judged block


@pytest.mark.parametrize("n", [5000, 40000])
def test_judge_forces_matches_its_parts(device, n):
    """The judge + forces kernel: a block of 64 envs with a judged or fallback
    env carries the step's qacc, tau, anchors, normal forces and feet; every
    other block carries g1_hip_forces' outputs; the state is not written.
    At both work splits, to float32 rounding (separate kernel instances)."""
    st = _state(device, n, 29, steps=5)
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    fallback = torch.zeros(n, dtype=torch.uint8, device=device)
    fallback[::97] = 1
    phase = 5
    z = lambda d: torch.zeros((d, n), device=device)  # noqa: E731
    before = st.host()
    qacc, tau, fn, fp, fv = torch.full((35, n), 1e9, device=device), z(35), z(8), z(6), z(6)
    anc = st.anchor.clone()
    _lib.check(lib.g1_hip_judge_forces(n, 2e-3, st.qpos.data_ptr(), st.qvel.data_ptr(), anc.data_ptr(),
                                       st.ctrl.data_ptr(), None, fallback.data_ptr(), phase, 16, qacc.data_ptr(),
                                       tau.data_ptr(), fn.data_ptr(), fp.data_ptr(), fv.data_ptr(), 0, s),
               "judge_forces")
    tau_f, fn_f, fp_f, fv_f = z(35), z(8), z(6), z(6)
    anc_f = st.anchor.clone()
    _lib.check(lib.g1_hip_forces(n, 2e-3, st.qpos.data_ptr(), st.qvel.data_ptr(), anc_f.data_ptr(),
                                 st.ctrl.data_ptr(), None, tau_f.data_ptr(), fn_f.data_ptr(), fp_f.data_ptr(),
                                 fv_f.data_ptr(), s), "forces")
    torch.cuda.synchronize()
    after = st.host()
    for k in ("qpos", "qvel", "anchor"):
        assert np.array_equal(before[k], after[k]), k
    st.step(1, 2e-3, tau=True)
    ref = st.host()
    blk = np.arange(n) // 64
    sel = ((blk % 16) == phase) | (fallback.cpu().numpy() != 0)
    stepped = np.isin(blk, np.unique(blk[sel]))
    assert 0 < stepped.sum() < n
    h = {k: v.cpu().numpy() for k, v in dict(qacc=qacc, tau=tau, fn=fn, fp=fp, fv=fv, anc=anc).items()}
    f = {k: v.cpu().numpy() for k, v in dict(tau=tau_f, fn=fn_f, fp=fp_f, fv=fv_f, anc=anc_f).items()}
    same_at_f32(h["qacc"][:, sel], ref["qacc"][:, sel], "qacc")
    for k, rk in (("tau", "tau"), ("fn", "fn"), ("fp", "foot_pos"), ("fv", "foot_vel"), ("anc", "anchor")):
        same_at_f32(h[k][:, stepped], ref[rk][:, stepped], k + " stepped", unset=1e30)
        same_at_f32(h[k][:, ~stepped], f[k][:, ~stepped], k + " forces", unset=1e30)


@pytest.mark.parametrize("tol", [1e-3, 1e30])
def test_steps_equal_repeated_step(device, tol):
    """SurrogateStepper.steps (one host call, g1_hip_surrogate_steps) and the
    same number of step() calls give the same bits: state, feet, contact
    forces, qacc, fallback flags and statistics; with the judge accepting
    everything (tol 1e30) the network's qacc is integrated, at 1e-3 every
    env stays with ABA."""
    mlp = MLP(device, seed=4)
    n = 5000
    qp, qv, an, cb = W.randomized_actuation(n, seed=27, nblocks=1)
    runs = []
    for one_call in (True, False):
        st = GpuState(qp, qv, an, cb[0], device)
        stepper = SurrogateStepper(mlp, n, device, period=16, tol=tol, track_qacc=True)
        stepper.steps(st.qpos, st.qvel, st.anchor, st.ctrl, 3)          # t = 3: later calls start mid-period
        if one_call:
            stepper.steps(st.qpos, st.qvel, st.anchor, st.ctrl, 21)
        else:
            for _ in range(21):
                stepper.step(st.qpos, st.qvel, st.anchor, st.ctrl)
        torch.cuda.synchronize()
        assert stepper.t == 24
        extra = {k: getattr(stepper, k).cpu().numpy() for k in ("fallback", "stats", "qacc", "foot_pos", "foot_vel",
                                                                  "contact_forces")}
        runs.append((st.host(), extra))
    (ha, ea), (hb, eb) = runs
    for k in ("qpos", "qvel", "anchor"):
        assert np.array_equal(ha[k], hb[k]), k
    ea["stats"], eb["stats"] = ea["stats"][:6], eb["stats"][:6]     # [6]: steps()' own count of envs off fallback
    for k in ea:
        assert np.array_equal(ea[k], eb[k]), k
    summ = None
    if tol == 1e30:
        summ = int(ea["stats"][3])
        assert int(ea["fallback"].sum()) == 0          # every env accepted once judged
    else:
        assert int(ea["fallback"].sum()) == n
    write_metrics(f"surrogate_steps_one_call_tol{tol:g}", {"stats": ea["stats"].tolist(), "fallback_env_steps": summ})


def _steps_run(device, mlp, n, tol, fallback_runs, skip_idle, schedule, reset=None, reset_before=3, flags=0,
               mu=None):
    """SurrogateStepper.steps over the call sizes in `schedule` (calls start
    and end mid-period), optionally resetting envs (mark_reset) before call
    number `reset_before`."""
    qp, qv, an, cb = W.randomized_actuation(n, seed=35, nblocks=1)
    st = GpuState(qp, qv, an, cb[0], device)
    stepper = SurrogateStepper(mlp, n, device, period=16, tol=tol, track_qacc=True, fallback_runs=fallback_runs,
                               skip_idle=skip_idle, judge_flags=flags)
    for i, k in enumerate(schedule):
        if reset is not None and i == reset_before:
            stepper.mark_reset(reset)
        stepper.steps(st.qpos, st.qvel, st.anchor, st.ctrl, k, mu=mu)
    torch.cuda.synchronize()
    extra = {k: getattr(stepper, k).cpu().numpy() for k in ("fallback", "stats", "qacc", "foot_pos", "foot_vel",
                                                            "contact_forces", "snap", "sflag", "dropped")}
    return st.host(), extra, stepper.summary()


@pytest.mark.parametrize("n", [5000, 40000])
@pytest.mark.parametrize("case", ["never_accepted", "accepted", "accepted_then_reset"])
def test_fallback_runs_match_per_step(device, n, case):
    """steps() with fallback runs (one g1_hip_fallback_run launch per period
    segment for the blocks whose envs are all in fallback), with and without
    the idle per-step launches left out, gives the bits of the per-step path:
    state, feet, contact forces, tracked qacc, records, flags and statistics.
    Never accepted: every block runs; accepted (tol 1e30): the blocks leave
    fallback at the first period's end and the later steps take the per-step
    path; then a reset of whole blocks and of single envs makes blocks that
    run and blocks that mix fallback and network envs.  40,000 envs take the
    2-wave split."""
    mlp = MLP(device, seed=4)
    tol = 1e-3 if case == "never_accepted" else 1e30
    reset = None
    if case == "accepted_then_reset":
        rng = np.random.default_rng(8)
        r = (rng.random(n) < 0.2).astype(np.uint8)
        r[64 * 3:64 * 7] = 1                                 # whole blocks 3-6
        r[64 * 9:64 * 10] = 0                                # block 9: none
        reset = torch.as_tensor(r, device=device)
    schedule = [3, 10, 13, 21, 7, 16]                        # t = 3, 13, 26, 47, 54, 70
    hb, eb, sb = _steps_run(device, mlp, n, tol, False, False, schedule, reset)        # per-step path
    eb["stats"] = eb["stats"][:6]
    for skip_idle in (True, False):
        ha, ea, sa = _steps_run(device, mlp, n, tol, True, skip_idle, schedule, reset)
        for k in ("qpos", "qvel", "anchor", "foot_pos", "foot_vel", "fn"):
            assert np.array_equal(ha[k], hb[k]), (skip_idle, k, float(np.abs(ha[k] - hb[k]).max()))
        ea["stats"] = ea["stats"][:6]          # [6]: the idle skip's own count of envs off fallback
        for k in ea:
            assert np.array_equal(ea[k], eb[k]), (skip_idle, k)
        assert sa == sb, skip_idle
    if case == "never_accepted":
        assert sa["aba_frac"] == 1.0 and sa["fallback_envs_now"] == n
        assert sa["judged"] == sa["rejected"] == (70 // 16) * n       # four batch judgements, all rejected
    else:
        assert sa["fallback_envs_now"] < n
    write_metrics(f"surrogate_fallback_runs_{case}_{n}", sa)


@pytest.mark.parametrize("n", [5000, 40000])
def test_fallback_runs_with_islands_and_friction(device, n):
    """The fallback-run kernel's float64-island instances (the judge's
    F64_SOLVE | F64_CONTACT) with per-env friction: the bits of the per-step
    path, as without islands."""
    mlp = MLP(device, seed=4)
    mu = torch.as_tensor(np.random.default_rng(5).uniform(0.4, 1.2, n).astype(np.float32), device=device)
    flags = _lib.G1_FLAG_F64_SOLVE | _lib.G1_FLAG_F64_CONTACT
    schedule = [5, 20, 9]
    hb, eb, sb = _steps_run(device, mlp, n, 1e-3, False, False, schedule, flags=flags, mu=mu)
    ha, ea, sa = _steps_run(device, mlp, n, 1e-3, True, True, schedule, flags=flags, mu=mu)
    for k in ("qpos", "qvel", "anchor"):
        assert np.array_equal(ha[k], hb[k]), (k, float(np.abs(ha[k] - hb[k]).max()))
    ea["stats"], eb["stats"] = ea["stats"][:6], eb["stats"][:6]
    for k in ea:
        assert np.array_equal(ea[k], eb[k]), k
    assert sa == sb and sa["aba_frac"] == 1.0 and sa["judged"] == 2 * n


@pytest.mark.parametrize("n", [5000, 40000])
@pytest.mark.parametrize("tol", [1e-3, 1e30])
def test_steps_schedule_matches_block_calls(device, n, tol):
    """steps_schedule (PD targets resampled every 10 steps, one host call;
    fallback runs switch target blocks inside the kernel) gives the bits of
    one steps() call per target block: state, feet, contact forces, tracked
    qacc, records, flags and statistics.  The calls start mid-block and
    mid-period."""
    mlp = MLP(device, seed=4)
    qp, qv, an, cb = W.randomized_actuation(n, seed=36, nblocks=7)
    ctrl = torch.as_tensor(cb, device=device)
    runs = []
    for sched in (True, False):
        st = GpuState(qp, qv, an, cb[0], device)
        stepper = SurrogateStepper(mlp, n, device, period=16, tol=tol, track_qacc=True)
        for step0, k in ((0, 7), (7, 38), (45, 20)):
            if sched:
                stepper.steps_schedule(st.qpos, st.qvel, st.anchor, ctrl, step0, k, 10)
            else:
                s0 = step0
                while s0 < step0 + k:
                    m = min(10 - s0 % 10, step0 + k - s0)
                    stepper.steps(st.qpos, st.qvel, st.anchor, ctrl[s0 // 10], m)
                    s0 += m
        torch.cuda.synchronize()
        extra = {k: getattr(stepper, k).cpu().numpy() for k in ("fallback", "qacc", "foot_pos", "foot_vel",
                                                                "contact_forces", "snap", "sflag")}
        extra["stats"] = stepper.stats.cpu().numpy()[:6]
        runs.append((st.host(), extra, stepper.summary()))
    (ha, ea, sa), (hb, eb, sb) = runs
    for k in ("qpos", "qvel", "anchor"):
        assert np.array_equal(ha[k], hb[k]), (k, float(np.abs(ha[k] - hb[k]).max()))
    for k in ea:
        assert np.array_equal(ea[k], eb[k]), k
    assert sa == sb
    if tol == 1e-3:
        assert sa["judged"] == sa["batch_judged"] == (65 // 16) * n     # four batch judgements, all rejected
    else:
        assert sa["fallback_envs_now"] == 0


@pytest.mark.parametrize("n", [5000, 40000])
def test_fallback_runs_are_plain_aba(device, n):
    """A network the judge never accepts: steps_schedule (fallback runs,
    target blocks switched in the kernel, batch judgements between periods)
    leaves exactly the state of plain ABA (g1_hip_step_schedule) over 45
    steps from the same start -- the fallback run computes the plain step's
    bits."""
    mlp = MLP(device, seed=4)
    qp, qv, an, cb = W.randomized_actuation(n, seed=37, nblocks=5)
    ctrl = torch.as_tensor(cb, device=device)
    sur = GpuState(qp, qv, an, cb[0], device)
    aba = GpuState(qp, qv, an, cb[0], device)
    stepper = SurrogateStepper(mlp, n, device, period=16, tol=1e-3)
    stepper.steps_schedule(sur.qpos, sur.qvel, sur.anchor, ctrl, 0, 45, 10)
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    _lib.check(lib.g1_hip_step_schedule(n, 45, 10, 0, 2e-3, aba.qpos.data_ptr(), aba.qvel.data_ptr(),
                                        aba.anchor.data_ptr(), ctrl.data_ptr(), ctrl.shape[1] * ctrl.shape[2],
                                        None, None, None, None, None, 0, s), "step_schedule")
    torch.cuda.synchronize()
    ha, hb = sur.host(), aba.host()
    for k in ("qpos", "qvel", "anchor"):
        assert np.array_equal(ha[k], hb[k]), (k, float(np.abs(ha[k] - hb[k]).max()))
    summ = stepper.summary()
    assert summ["aba_frac"] == 1.0 and summ["judged"] == summ["rejected"] == 2 * n


def _stepper_run(device, mlp, n, steps, tol, defer, reset_at=None):
    qp, qv, an, cb = W.randomized_actuation(n, seed=33, nblocks=1)
    st = GpuState(qp, qv, an, cb[0], device)
    stepper = SurrogateStepper(mlp, n, device, period=16, tol=tol, defer=defer, track_qacc=True)
    fb = []
    for k in range(steps):
        if reset_at is not None and k == reset_at:
            stepper.mark_reset(torch.ones(n, dtype=torch.uint8, device=device))
        stepper.step(st.qpos, st.qvel, st.anchor, st.ctrl)
        fb.append(stepper.fallback.clone())
    torch.cuda.synchronize()
    return st, stepper, [f.cpu().numpy() for f in fb]


@pytest.mark.parametrize("n", [5000, 40000])
def test_deferred_judging_matches_immediate_when_never_accepted(device, n):
    """A network the judge never accepts (random weights, tolerance 1e-3):
    deferred and immediate judging integrate ABA everywhere, bit for bit,
    and judge the same envs with the same errors (the batch pass and the
    fused step's network give the same bits; 40,000 envs batch-judge with
    128-env workgroups)."""
    mlp = MLP(device, seed=4)
    runs = [_stepper_run(device, mlp, n, 32, 1e-3, defer) for defer in (True, False)]
    (sd, pd, _), (si, pi, _) = runs
    hd, hi = sd.host(), si.host()
    for k in ("qpos", "qvel", "anchor"):
        assert np.array_equal(hd[k], hi[k]), k
    a, b = pd.summary(), pi.summary()
    assert a["judged"] == b["judged"] == 32 * n // 16
    assert a["rejected"] == b["rejected"] == a["judged"]
    assert a["max_judged_rel_err"] == b["max_judged_rel_err"]
    assert a["batch_judged"] == a["judged"] and a["pending_records"] == 0
    assert a["aba_frac"] == b["aba_frac"] == 1.0


def test_deferred_judging_accepts_at_the_period_end(device):
    """Tolerance 1e30 (every judgement accepts): with deferred judging every
    env stays with ABA through the first period (the state is plain ABA's bit
    for bit), all leave fallback after its last step, and the next period's
    judgements, of envs off fallback, are made at once."""
    mlp = MLP(device, seed=4)
    n = 4096
    st, stepper, fb = _stepper_run(device, mlp, n, 16, 1e30, True)
    ref = _stepper_run(device, mlp, n, 16, 1e-3, False)[0]          # never accepted: ABA
    h, r = st.host(), ref.host()
    for k in ("qpos", "qvel", "anchor"):
        assert np.array_equal(h[k], r[k]), k
    assert all(int(f.sum()) == n for f in fb[:15]) and int(fb[15].sum()) == 0
    s1 = stepper.summary()
    assert s1["batch_judged"] == n and s1["rejected"] == 0 and s1["judged"] == n
    for _ in range(16):
        stepper.step(st.qpos, st.qvel, st.anchor, st.ctrl)
    torch.cuda.synchronize()
    s2 = stepper.summary()
    assert s2["batch_judged"] == n and s2["judged"] == 2 * n and s2["pending_records"] == 0
    assert int(stepper.fallback.sum().item()) == 0 and np.isfinite(st.host()["qpos"]).all()


def test_reset_clears_pending_records(device):
    """mark_reset mid-period: the records made before the reset are dropped
    (they judged the pre-reset state), so those blocks stay in fallback; the
    blocks judged after the reset are accepted at the period end."""
    mlp = MLP(device, seed=4)
    n = 4096
    st, stepper, fb = _stepper_run(device, mlp, n, 16, 1e30, True, reset_at=8)
    blk = (np.arange(n) // 64) % 16
    assert np.array_equal(fb[15], (blk < 8).astype(np.uint8))
    assert stepper.summary()["batch_judged"] == int((blk >= 8).sum())
