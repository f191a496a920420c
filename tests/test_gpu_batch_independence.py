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


def test_fused_surrogate_batch_independence(device):
    mlp = MLP(device, seed=5)
    n = 65536
    qp, qv, _, _ = W.randomized_actuation(n, seed=10, nblocks=1)
    tau = np.random.default_rng(11).normal(size=(35, n)).astype(np.float32) * 20
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=device)  # noqa: E731
    big = torch.zeros((n, 64), device=device)
    mlp.forward_state(n, d(qp), d(qv), d(tau), big)
    for lo, hi in ((0, 64), (1000, 1100), (n - 100, n)):
        sl = slice(lo, hi)
        small = torch.zeros((hi - lo, 64), device=device)
        mlp.forward_state(hi - lo, d(qp[:, sl]), d(qv[:, sl]), d(tau[:, sl]), small)
        torch.cuda.synchronize()
        assert torch.equal(big[sl], small), (lo, hi)
