"""The surrogate stepper at the evaluation size of the learned-dynamics
configuration (65,536 envs), stepped in lockstep with plain ABA.

Every step, plain ABA (g1_hip_step_ex, the fused FP32 step) takes one step
from the surrogate path's state (teacher forcing), and the two results are
compared at the float32 bar of two kernel instances of the same physics
(same code, different multiply-add contraction): median relative difference
<= 1e-6, p99 <= 1e-4, max <= 5e-2 over all dofs and envs of a step (one ulp
of foot height moves qacc by ~1.4e-2 in stance).

* A network the judge never accepts (random weights at tolerance 1e-3): the
  stepper starts every env in fallback, so each step is ABA's.
* An exact predictor (ABA's own qacc, z-scored, fed through the commit in
  place of the network): every judgement accepts, every env leaves fallback
  within one judge period, and the integrated network qacc is ABA's.
"""
import numpy as np
import pytest
import torch

from g1sim_amd import _lib, workloads as W
from g1sim_amd.surrogate import MLP, NOUT, POUT, SurrogateStepper

from .gpu_helpers import GpuState, write_metrics

pytestmark = pytest.mark.gpu

N = 65536
STEPS = 112                 # 7 judge periods of 16
DT = 2e-3


def _rel_stats(a, b, unset=None):
    """median / p99 / max of |a - b| / max(|b|, 1) over the entries (both
    unset, e.g. free anchors, are skipped; one unset and not the other is a
    mismatch)."""
    if unset is not None:
        ua, ub = a >= unset, b >= unset
        assert bool((ua == ub).all()), "set/unset mismatch"
        m = ~ub
        a, b = a[m], b[m]
    r = ((a - b).abs() / b.abs().clamp_min(1.0)).flatten().double()
    if r.numel() == 0:
        return 0.0, 0.0, 0.0
    srt = torch.sort(r).values
    return float(srt[srt.numel() // 2]), float(srt[int(0.99 * (srt.numel() - 1))]), float(srt[-1])


def _lockstep(device, exact, N=N):
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    qp, qv, an, cb = W.randomized_actuation(N, seed=41, nblocks=(STEPS + 9) // 10)
    ctrls = torch.as_tensor(cb, device=device)
    sur = GpuState(qp, qv, an, cb[0], device)          # the surrogate path
    ref = GpuState(qp, qv, an, cb[0], device)          # one plain ABA step from sur's state, every step
    free = GpuState(qp, qv, an, cb[0], device)         # plain ABA, free running
    mlp = MLP(device, seed=0)                          # random weights: never within 1e-3 of ABA
    # output z-scoring of the scale of real qacc, so that the exact
    # predictor's round trip (z-score, then back) rounds
    rng = np.random.default_rng(3)
    mlp.set_normalization(np.zeros(106), np.ones(106), rng.normal(0.0, 5.0, NOUT), rng.uniform(5.0, 300.0, NOUT))
    stepper = SurrogateStepper(mlp, N, device, period=16, tol=1e-3, fused=not exact, track_qacc=True)
    pred = torch.zeros((stepper.rows, POUT), dtype=torch.float32, device=device)
    ones = torch.ones(N, dtype=torch.uint8, device=device)
    qall = torch.zeros((NOUT, N), dtype=torch.float32, device=device)
    tau_tmp = torch.zeros((NOUT, N), dtype=torch.float32, device=device)
    worst = {k: [0.0, 0.0, 0.0] for k in ("qacc", "qpos", "qvel", "anchor")}
    fb_after_period = None
    for k in range(STEPS):
        for g in (sur, ref, free):
            g.ctrl.copy_(ctrls[k // 10])
        ref.qpos.copy_(sur.qpos)
        ref.qvel.copy_(sur.qvel)
        ref.anchor.copy_(sur.anchor)
        ref.step(1, DT)
        free.step(1, DT, qacc=False)
        prediction = None
        if exact:
            # ABA's qacc of every env from the judge's own kernel instance
            # (every block selected), z-scored: the exact predictor
            anc = sur.anchor.clone()
            _lib.check(lib.g1_hip_judge_forces(N, DT, sur.qpos.data_ptr(), sur.qvel.data_ptr(), anc.data_ptr(),
                                               sur.ctrl.data_ptr(), None, ones.data_ptr(), 0, 16, qall.data_ptr(),
                                               tau_tmp.data_ptr(), None, None, None, 0, s), "judge_forces")
            pred[:N, :NOUT] = ((qall - mlp.ymean[:, None]) * mlp.yinv[:, None]).T
            prediction = pred
        stepper.step(sur.qpos, sur.qvel, sur.anchor, sur.ctrl, prediction=prediction)
        for key, a, b, unset in (("qacc", stepper.qacc, ref.qacc, None), ("qpos", sur.qpos, ref.qpos, None),
                                 ("qvel", sur.qvel, ref.qvel, None), ("anchor", sur.anchor, ref.anchor, 1e29)):
            st = _rel_stats(a, b, unset)
            worst[key] = [max(x, y) for x, y in zip(worst[key], st)]
        if k == 15:
            fb_after_period = int(stepper.fallback.sum().item())
    torch.cuda.synchronize()
    summ = stepper.summary()
    hs, hf = sur.qpos.cpu().numpy(), free.qpos.cpu().numpy()
    drift = np.linalg.norm(hs[0:3] - hf[0:3], axis=0)
    return worst, summ, fb_after_period, drift, np.isfinite(hs).all()


def _check_worst(worst):
    for key, (med, p99, mx) in worst.items():
        assert med <= 1e-6 and p99 <= 1e-4 and mx <= 5e-2, (key, med, p99, mx)


@pytest.mark.parametrize("n", [N, 16384])
def test_never_accepted_network_is_aba(device, n):
    """At 65,536 envs (the 2-wave split) the fallback path is plain ABA bit
    for bit: the judge's kernel instance gives the step's qacc and the commit
    integrates with the step's fma sequence (g1_commit.h euler_base), so
    the free-running rollouts do not drift apart at all."""
    worst, summ, fb16, drift, finite = _lockstep(device, exact=False, N=n)
    write_metrics(f"surrogate_never_accepted_{n}", {
        "steps": STEPS, "worst_per_step_rel_diff_vs_aba (median, p99, max)": worst, "summary": summ,
        "free_running_base_drift_vs_aba_m": {"median": float(np.median(drift)), "p90": float(np.percentile(drift, 90)),
                                             "max": float(drift.max())}})
    assert finite
    assert summ["judged"] == STEPS * n // 16 and summ["rejected"] == summ["judged"]
    assert summ["aba_frac"] == 1.0 and summ["fallback_envs_now"] == n and fb16 == n
    _check_worst(worst)
    if n == N:
        assert all(v == [0.0, 0.0, 0.0] for v in worst.values()) and drift.max() == 0.0


def test_exact_predictor_accepted_at_65536(device):
    worst, summ, fb16, drift, finite = _lockstep(device, exact=True)
    write_metrics("surrogate_exact_predictor_65536", {
        "steps": STEPS, "worst_per_step_rel_diff_vs_aba (median, p99, max)": worst, "summary": summ,
        "free_running_base_drift_vs_aba_m": {"median": float(np.median(drift)), "p90": float(np.percentile(drift, 90)),
                                             "max": float(drift.max())}})
    assert finite
    assert summ["judged"] == STEPS * N // 16 and summ["rejected"] == 0
    assert summ["max_judged_rel_err"] <= 1e-5
    # each block leaves fallback at its first judgement: all out after one period
    assert fb16 == 0 and summ["fallback_envs_now"] == 0
    # fallback env-steps: block b waits b % 16 steps for its first judgement
    assert summ["fallback_env_steps"] == sum(N // 16 * p for p in range(16))
    _check_worst(worst)
