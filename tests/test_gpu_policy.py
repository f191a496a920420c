"""Policy-in-the-loop kernels on the GPU against the RL oracle
(oracle/rl_oracle.py): the observation kernel (and G1Sim.get_obs), the fused
policy kernel (observation, normalisation, actor MLP on MFMA, sampling, ctrl,
pushes), the reward / termination / bookkeeping kernel, and whole rollouts
with ABA and with the surrogate."""
import numpy as np
import pytest
import torch

from g1sim_amd import _lib, workloads as W
from g1sim_amd.model import MODEL
from g1sim_amd.policy import Policy, PolicyRollout
from g1sim_amd.sim import G1Sim
from oracle import rl_oracle as R

from .gpu_helpers import write_metrics

pytestmark = pytest.mark.gpu

DEFAULT = MODEL.qpos_stand[7:36]
DEFAULT32 = DEFAULT.astype(np.float32)


def bf16(a):
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)


def random_rl_state(n, seed):
    rng = np.random.default_rng(seed)
    qp, qv = W.random_state(rng, n)
    qp = np.ascontiguousarray(qp.T, np.float32); qv = np.ascontiguousarray(qv.T, np.float32)
    qv[0:6] += rng.normal(size=(6, n)).astype(np.float32) * 0.5
    cmd = rng.uniform(-1, 1, size=(3, n)).astype(np.float32)
    cmd[:, : n // 10] = 0.0                                 # some standing commands
    last = (DEFAULT32[:, None] + rng.normal(size=(29, n)) * 0.2).astype(np.float32)
    return qp, qv, cmd, last


def dev(a, device):
    return torch.as_tensor(np.ascontiguousarray(a), device=device)


@pytest.mark.parametrize("layout", [0, 1])
def test_obs_kernel_matches_oracle(device, layout):
    n = 1000
    qp, qv, cmd, last = random_rl_state(n, 10 + layout)
    obs = torch.empty((n, 99 if layout == 0 else 96), device=device)
    lib = _lib.load()
    t = [dev(a, device) for a in (qp, qv, cmd, last)]           # keep the inputs alive across the launch
    _lib.check(lib.g1_hip_get_obs(n, layout, t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), t[3].data_ptr(),
                                  obs.data_ptr(), torch.cuda.current_stream(device).cuda_stream), "get_obs")
    torch.cuda.synchronize()
    ref = R.get_obs(qp.astype(np.float64), qv.astype(np.float64), cmd.astype(np.float64), last.astype(np.float64),
                    DEFAULT32.astype(np.float64), layout=layout)
    np.testing.assert_allclose(obs.cpu().numpy(), ref, rtol=1e-5, atol=2e-6)


def test_g1sim_get_obs_matches_oracle(device):
    env = G1Sim(nenv=300, device=device)
    env.reset_all(seed=3, noise=0.2, drop=0.0)
    env.qvel.normal_(0.0, 0.7)
    env.commands.uniform_(-1.0, 1.0)
    env.ctrl.normal_(0.0, 0.3)
    env.step(nsub=2)
    obs = env.get_obs()
    assert obs.shape == (300, 99) and obs.is_contiguous()
    h = lambda t: t.cpu().numpy().astype(np.float64)  # noqa: E731
    ref = R.get_obs(h(env.qpos), h(env.qvel), h(env.commands), h(env.last_actions), h(env.default_joint_pos)[:, 0])
    np.testing.assert_allclose(obs.cpu().numpy(), ref, rtol=1e-5, atol=2e-6)


def _act(pol, n, qp, qv, cmd, last, device, sample, seed=5, t=0, push=None):
    qpos, qvel, cmd_d, last_d = dev(qp, device), dev(qv, device), dev(cmd, device), dev(last, device)
    ctrl = torch.zeros((29, n), device=device)
    actions = torch.zeros((29, n), device=device)
    obs = torch.zeros((n, pol.nobs), device=device)
    pi = None if push is None else dev(push, device)
    pol.act(n, qpos, qvel, cmd_d, last_d, ctrl, actions=actions, obs_out=obs,
            push_interval=pi, seed=seed, t=t, sample=sample)
    torch.cuda.synchronize()
    return dict(ctrl=ctrl.cpu().numpy(), actions=actions.cpu().numpy(), obs=obs.cpu().numpy(),
                qvel=qvel.cpu().numpy(), push=None if pi is None else pi.cpu().numpy())


@pytest.mark.parametrize("layout", [0, 1])
def test_policy_act_matches_cpu(device, layout):
    n = 1000                                    # not a multiple of 64: partial last workgroup
    qp, qv, cmd, last = random_rl_state(n, 20 + layout)
    pol = Policy(device, layout=layout, seed=7)
    rng = np.random.default_rng(4)
    mean = np.zeros(128, np.float32); var = np.ones(128, np.float32)
    mean[:pol.nobs] = rng.normal(size=pol.nobs) * 0.1
    var[:pol.nobs] = rng.uniform(0.2, 3.0, pol.nobs)
    pol.update_normalizer(mean, var)
    out = _act(pol, n, qp, qv, cmd, last, device, sample=False)
    ref_obs = R.get_obs(qp.astype(np.float64), qv.astype(np.float64), cmd.astype(np.float64),
                        last.astype(np.float64), DEFAULT32.astype(np.float64), layout=layout)
    np.testing.assert_allclose(out["obs"], ref_obs, rtol=1e-5, atol=2e-6)
    inv = pol.obs_inv.cpu().numpy()[:pol.nobs].astype(np.float64)
    x = np.clip((out["obs"].astype(np.float64) - mean[:pol.nobs]) * inv, -5, 5)
    ref = R.policy_mean(x, [w.astype(np.float64) for w in pol.W32], [b.astype(np.float64) for b in pol.b32],
                        round_fn=bf16)
    got = out["actions"].T
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(got - ref).max() < 3e-3 * scale
    ctrl = (DEFAULT32[:, None] + out["actions"] * R.ACTION_SCALE.astype(np.float32)[:, None])
    np.testing.assert_allclose(out["ctrl"], ctrl, rtol=1e-6, atol=1e-6)
    write_metrics(f"policy_act_layout{layout}", {"max_abs_err": float(np.abs(got - ref).max()),
                                                 "max_abs_action": float(np.abs(ref).max())})


def test_policy_sampling_and_pushes(device):
    n = 2048
    qp, qv, cmd, last = random_rl_state(n, 30)
    pol = Policy(device, layout=1, seed=8)
    det = _act(pol, n, qp, qv, cmd, last, device, sample=False)
    push = np.where(np.arange(n) % 2 == 0, 1, 10).astype(np.int32)
    s1 = _act(pol, n, qp, qv, cmd, last, device, sample=True, seed=5, t=3, push=push)
    s2 = _act(pol, n, qp, qv, cmd, last, device, sample=True, seed=5, t=3, push=push)
    s3 = _act(pol, n, qp, qv, cmd, last, device, sample=True, seed=5, t=4)
    assert np.array_equal(s1["actions"], s2["actions"])            # a function of (seed, t, env)
    assert not np.array_equal(s1["actions"], s3["actions"])
    z = (s1["actions"] - det["actions"]).ravel()                    # std = exp(0) = 1
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03
    # the observation is taken before the push
    np.testing.assert_array_equal(s1["obs"], det["obs"])
    # pushes: envs whose counter reaches 0 get a bounded kick and a new interval
    dq = s1["qvel"] - qv
    kicked = push == 1
    half = np.array([0.5, 0.5, 0.4, 0.52, 0.52, 0.78])
    assert (np.abs(dq[:6, kicked]) <= half[:, None] + 1e-6).all()
    assert (np.abs(dq[:6, kicked]).mean(axis=1) > 0.2 * half).all()
    assert (dq[:, ~kicked] == 0).all() and (dq[6:] == 0).all()
    assert ((s1["push"][kicked] >= 50) & (s1["push"][kicked] < 150)).all()
    assert (s1["push"][~kicked] == 9).all()


def test_post_step_matches_oracle(device):
    n = 3000
    rng = np.random.default_rng(40)
    qp, qv, cmd, _ = random_rl_state(n, 40)
    # a few strongly tilted envs terminate
    tilt = rng.random(n) < 0.1
    for e in np.nonzero(tilt)[0]:
        ax = rng.normal(size=3); ax[2] = 0; ax /= np.linalg.norm(ax)
        qp[3:7, e] = np.concatenate([[np.cos(0.8)], np.sin(0.8) * ax])      # 91.7 degrees
    foot_pos = rng.uniform(-0.05, 0.3, (6, n)).astype(np.float32)
    foot_vel = rng.normal(size=(6, n)).astype(np.float32)
    cf = np.where(rng.random((8, n)) < 0.5, rng.uniform(0, 60, (8, n)), 0.0).astype(np.float32)
    actions = rng.normal(size=(29, n)).astype(np.float32)
    prev = rng.normal(size=(29, n)).astype(np.float32)
    ct = rng.choice([1, 2, 40], n).astype(np.int32)
    el = rng.choice([3, 998, 999], n).astype(np.int32)
    pi = rng.integers(5, 60, n).astype(np.int32)
    t = {k: dev(v, device) for k, v in dict(qp=qp, qv=qv, fp=foot_pos, fv=foot_vel, cf=cf, a=actions, pa=prev,
                                             cmd=cmd, ct=ct, el=el, pi=pi).items()}
    reward = torch.zeros(n, device=device)
    done = torch.zeros(n, dtype=torch.uint8, device=device)
    reset = torch.zeros(n, dtype=torch.uint8, device=device)
    sums = torch.tensor([5.0, 2.0], dtype=torch.float64, device=device)    # accumulated into
    lib = _lib.load()
    _lib.check(lib.g1_rollout_post(n, t["qp"].data_ptr(), t["qv"].data_ptr(), t["fp"].data_ptr(), t["fv"].data_ptr(),
                                   t["cf"].data_ptr(), t["a"].data_ptr(), t["cmd"].data_ptr(), t["pa"].data_ptr(),
                                   t["ct"].data_ptr(), t["el"].data_ptr(), t["pi"].data_ptr(), 1000, 1.0, 0.5, 99, 7,
                                   reward.data_ptr(), done.data_ptr(), reset.data_ptr(), sums.data_ptr(),
                                   sums[1:].data_ptr(), torch.cuda.current_stream(device).cuda_stream),
               "rollout_post")
    torch.cuda.synchronize()
    f64 = lambda a: a.astype(np.float64)  # noqa: E731
    obs = R.get_obs(f64(qp), f64(qv), f64(cmd), np.zeros((29, n)), f64(DEFAULT32))
    r_ref, d_ref = R.reward_and_done(obs, f64(qp), f64(foot_pos), f64(foot_vel), f64(cf), f64(actions.T),
                                     f64(prev.T), MODEL.jnt_lo.astype(np.float32).astype(np.float64),
                                     MODEL.jnt_hi.astype(np.float32).astype(np.float64))
    r = reward.cpu().numpy()
    np.testing.assert_allclose(r, r_ref, rtol=2e-5, atol=2e-4)
    d = done.cpu().numpy().astype(bool)
    assert np.array_equal(d, d_ref) and d[tilt].all()
    sm = sums.cpu().numpy()
    np.testing.assert_allclose(sm[0], 5.0 + r.astype(np.float64).sum(), rtol=1e-6, atol=1e-3)
    assert sm[1] == 2.0 + d.sum()
    rs_ref, ct_ref, el_ref, resample = R.bookkeeping(d_ref, ct, el, 1000)
    rs = reset.cpu().numpy().astype(bool)
    assert np.array_equal(rs, rs_ref)
    assert np.array_equal(t["el"].cpu().numpy(), el_ref)
    ct_k = t["ct"].cpu().numpy()
    assert np.array_equal(ct_k[~resample], ct_ref[~resample])
    assert ((ct_k[resample] >= 150) & (ct_k[resample] < 250)).all()
    cmd_k = t["cmd"].cpu().numpy()
    assert np.array_equal(cmd_k[:, ~resample], cmd[:, ~resample])
    assert (np.abs(cmd_k[0:2, resample]) <= 1.0).all() and (np.abs(cmd_k[2, resample]) <= 0.5).all()
    pa = t["pa"].cpu().numpy()
    assert (pa[:, rs] == 0).all() and np.array_equal(pa[:, ~rs], actions[:, ~rs])
    pi_k = t["pi"].cpu().numpy()
    assert np.array_equal(pi_k[~rs], pi[~rs]) and ((pi_k[rs] >= 50) & (pi_k[rs] < 150)).all()
    write_metrics("rollout_post_3000", {"max_abs_reward_err": float(np.abs(r - r_ref).max()),
                                        "done": int(d.sum()), "reset": int(rs.sum()), "resample": int(resample.sum())})


def test_rollout_with_aba(device):
    n = 4096
    ro = PolicyRollout(n, device, physics="aba", seed=1, max_episode_length=30)
    rs = torch.zeros(24, dtype=torch.float64, device=device)
    dc = torch.zeros(24, dtype=torch.float64, device=device)
    for _ in range(2):
        ro.rollout(24, rs, dc)
    torch.cuda.synchronize()
    assert torch.isfinite(ro.qpos).all() and torch.isfinite(ro.qvel).all()
    el = ro.episode_length.cpu().numpy()
    assert (el >= 0).all() and (el < 30).all()                       # timeouts at 30 reset every env
    assert dc.sum().item() > 0                                       # a random policy falls
    assert np.isfinite(rs.cpu().numpy()).all()
    ct = ro.command_timeout.cpu().numpy()
    assert (ct >= 100).all() and (ct < 250).all()
    write_metrics("rollout_aba_4096", {"reward_per_step": (rs.cpu().numpy() / (2 * n)).tolist(),
                                       "done_per_step": dc.cpu().numpy().tolist()})


def test_rollout_with_surrogate(device):
    from g1sim_amd.surrogate import MLP, fit_normalization, generate_dataset, train
    ds = generate_dataset(device, nenv=4096, steps=60, seed=3)
    mlp = MLP(device, seed=2)
    fit_normalization(mlp, ds)
    train(mlp, ds, iters=200, batch=4096)
    del ds
    n = 2048
    ro = PolicyRollout(n, device, physics="surrogate", surrogate=mlp, seed=2, nsub=4)
    ro.rollout(6)
    torch.cuda.synchronize()
    summ = ro.stepper.summary()
    assert summ["env_steps"] == 6 * 4 * n
    # every judged block is judged at once, judged with its period's batch,
    # still pending, or its record was dropped by a reset
    assert summ["judged"] + summ["pending_records"] + summ["dropped_records"] == 6 * 4 * n // 16
    write_metrics("rollout_surrogate_2048", summ)


def test_obs_norm_update_matches_oracle(device):
    """Two consecutive running-statistics updates on the device against the
    oracle's restatement of examples/train.py:49-66."""
    pol = Policy(device, layout=1, seed=3)
    mean = np.zeros(96); var = np.ones(96); count = 1e-4
    for k in range(2):
        n = 5000 + 777 * k
        qp, qv, cmd, last = random_rl_state(n, 60 + k)
        t = [dev(a, device) for a in (qp, qv, cmd, last)]
        pol.update_obs_norm(n, *t)
        torch.cuda.synchronize()
        obs = R.get_obs(qp.astype(np.float64), qv.astype(np.float64), cmd.astype(np.float64),
                        last.astype(np.float64), DEFAULT32.astype(np.float64), layout=1)
        mean, var, count = R.update_obs_norm(obs, mean, var, count)
        np.testing.assert_allclose(pol.obs_mean.cpu().numpy()[:96], mean, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(pol.obs_var.cpu().numpy()[:96], var, rtol=1e-5, atol=1e-6)
        assert abs(float(pol.obs_count.item()) - count) <= 1e-6 * count
        np.testing.assert_allclose(pol.obs_inv.cpu().numpy()[:96], 1.0 / np.sqrt(var + 1e-8), rtol=2e-5)
