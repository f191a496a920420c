"""Size-independent properties at full sizes: environments do not interact,
so an env's result may not depend on the batch it is stepped in.  Large
batches (the 131,072-env contact-stress size, and 1,048,576 envs) are
compared bit for bit with small batches holding a copy of some of their
envs, through the same code path (the work split is forced to match)."""
import numpy as np
import pytest
import torch

from g1sim_amd import _lib, workloads as W
from g1sim_amd.surrogate import MLP

from .gpu_helpers import GpuState

pytestmark = pytest.mark.gpu


def _run(qp, qv, an, ct, device, nsub, flags, mu=None):
    st = GpuState(qp, qv, an, ct, device, mu=mu)
    st.step(nsub, 2e-3, flags=flags, qacc=True)
    return st.host()


@pytest.mark.parametrize("nenv,split_flag", [(131072, _lib.G1_FLAG_WAVES2), (1048576, _lib.G1_FLAG_WAVES2),
                                             (131072, _lib.G1_FLAG_WAVES4)])
def test_large_batch_matches_small_batches(device, nenv, split_flag):
    qp, qv, an, cb = W.randomized_actuation(nenv, seed=8, nblocks=1)
    mu = np.random.default_rng(9).uniform(0.4, 1.2, nenv).astype(np.float32)
    big = _run(qp, qv, an, cb[0], device, 4, split_flag, mu=mu)
    # three slices: the start, a run crossing a workgroup boundary in the
    # middle, and the tail (a partial workgroup when sliced from 100 envs)
    for lo, hi in ((0, 1000), (nenv // 2 - 37, nenv // 2 + 91), (nenv - 100, nenv)):
        sl = slice(lo, hi)
        small = _run(qp[:, sl], qv[:, sl], an[:, sl], cb[0][:, sl], device, 4, split_flag, mu=mu[sl])
        for k in ("qpos", "qvel", "anchor", "qacc", "fn"):
            assert np.array_equal(big[k][:, sl], small[k]), (k, lo, hi)


@pytest.mark.parametrize("n", [65536, 40017])
def test_fused_surrogate_batch_independence(device, n):
    """From 32,768 envs the fused forward runs 128 envs per workgroup (one
    in-place activation buffer), below 64 (two buffers); the k-order of
    every accumulation is the same, so every env's output is the same bits
    whichever kernel and batch computes it.  The big batch is covered
    completely by slices below 32,768 envs (and a tail slice)."""
    mlp = MLP(device, seed=5)
    qp, qv, _, _ = W.randomized_actuation(n, seed=10, nblocks=1)
    tau = np.random.default_rng(11).normal(size=(35, n)).astype(np.float32) * 20
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=device)  # noqa: E731
    big = torch.zeros((n, 64), device=device)
    mlp.forward_state(n, d(qp), d(qv), d(tau), big)
    cuts = list(range(0, n, 32767)) + [n]
    slices = [(lo, hi) for lo, hi in zip(cuts[:-1], cuts[1:])] + [(1000, 1100), (n - 100, n)]
    for lo, hi in slices:
        sl = slice(lo, hi)
        small = torch.zeros((hi - lo, 64), device=device)
        mlp.forward_state(hi - lo, d(qp[:, sl]), d(qv[:, sl]), d(tau[:, sl]), small)
        torch.cuda.synchronize()
        assert torch.equal(big[sl], small), (lo, hi)


@pytest.mark.parametrize("sample", [False, True])
def test_policy_batch_independence(device, sample):
    """The policy kernel at 40,017 envs (128 envs per workgroup, a partial
    last workgroup) against 64-env workgroups on smaller batches: the same
    actions, ctrl, observations and pushes, bit for bit.  With sampling the
    noise of env e is drawn from (seed, t, e), so the small batches start at
    env 0 (same indices); without it any slice matches."""
    from g1sim_amd.policy import Policy
    from g1sim_amd.model import MODEL
    n = 40017
    rng = np.random.default_rng(12)
    qp, qv = W.random_state(rng, n)
    qp = np.ascontiguousarray(qp.T, np.float32)
    qv = np.ascontiguousarray(qv.T, np.float32)
    cmd = rng.uniform(-1, 1, size=(3, n)).astype(np.float32)
    last = (MODEL.qpos_stand[7:, None] + rng.normal(size=(29, n)) * 0.2).astype(np.float32)
    push = rng.integers(-2, 5, n).astype(np.int32)
    pol = Policy(device, layout=1, seed=13)
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=device)  # noqa: E731

    def act(lo, hi):
        m = hi - lo
        qvel = d(qv[:, lo:hi])
        ctrl = torch.zeros((29, m), device=device)
        actions = torch.zeros((29, m), device=device)
        obs = torch.zeros((m, 96), device=device)
        pi = d(push[lo:hi])
        pol.act(m, d(qp[:, lo:hi]), qvel, d(cmd[:, lo:hi]), d(last[:, lo:hi]), ctrl, actions=actions, obs_out=obs,
                push_interval=pi, seed=3, t=5, sample=sample)
        torch.cuda.synchronize()
        return [t.cpu().numpy() for t in (ctrl, actions, obs, qvel, pi)]

    big = act(0, n)
    small_slices = [(0, 30000), (30000, n)] if not sample else [(0, 30000), (0, 64)]
    for lo, hi in small_slices:
        small = act(lo, hi)
        for name, b, sm in zip(("ctrl", "actions", "obs", "qvel", "push"), big, small):
            if lo > 0 and name in ("qvel", "push"):
                continue                        # the push kicks are drawn from the env's index in its batch
            bs = b[lo:hi] if name == "obs" else b[..., lo:hi]
            assert np.array_equal(bs, sm), (name, lo, hi)
