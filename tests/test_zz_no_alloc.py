"""g1_policy_obs_norm_update_noalloc (static device workspace, allocates
nothing) against g1_policy_obs_norm_update (caller's workspace), both called
through the C-ABI: the same running statistics, bit for bit, over two calls
in a row, the second on a smaller batch, for both observation layouts."""
import numpy as np
import pytest
import torch

from g1sim_amd import _lib, workloads as W
from g1sim_amd.model import MODEL

pytestmark = pytest.mark.gpu

STAND = MODEL.qpos_stand[7:].astype(np.float32)


def _batch(n, seed, device):
    rng = np.random.default_rng(seed)
    qp, qv = W.random_state(rng, n)
    qp = np.ascontiguousarray(qp.T, np.float32)
    qv = np.ascontiguousarray(qv.T, np.float32)
    qv[0:6] += rng.normal(size=(6, n)).astype(np.float32) * 0.5
    cmd = rng.uniform(-1, 1, size=(3, n)).astype(np.float32)
    last = (STAND[:, None] + rng.normal(size=(29, n)) * 0.2).astype(np.float32)
    return [torch.as_tensor(np.ascontiguousarray(a), device=device) for a in (qp, qv, cmd, last)]


@pytest.mark.parametrize("layout", [1, 0])
def test_noalloc_matches_workspace_entry(device, layout):
    lib = _lib.load()
    s = torch.cuda.current_stream(device).cuda_stream
    nobs = 99 if layout == 0 else 96

    def stats():
        mean = torch.zeros(128, dtype=torch.float32, device=device)
        var = torch.ones(128, dtype=torch.float32, device=device)
        count = torch.full((1,), 1e-4, dtype=torch.float32, device=device)
        inv = torch.ones(128, dtype=torch.float32, device=device)
        return mean, var, count, inv

    ws = torch.zeros(2 * 64 * nobs + 1, dtype=torch.float64, device=device)
    a, b = stats(), stats()
    for k, n in enumerate((65536, 3001)):                  # the second call on a smaller batch
        st = _batch(n, 90 + 10 * layout + k, device)
        ptrs = [t.data_ptr() for t in st]
        rc = lib.g1_policy_obs_norm_update(n, layout, *ptrs, *(t.data_ptr() for t in a), ws.data_ptr(), s)
        _lib.check(rc, "g1_policy_obs_norm_update")
        rc = lib.g1_policy_obs_norm_update_noalloc(n, layout, *ptrs, *(t.data_ptr() for t in b), s)
        _lib.check(rc, "g1_policy_obs_norm_update_noalloc")
        torch.cuda.synchronize()
        for x, y, name in zip(a, b, ("mean", "var", "count", "inv")):
            xa, ya = x.cpu().numpy(), y.cpu().numpy()
            assert np.array_equal(xa.view(np.uint32), ya.view(np.uint32)), (layout, k, name)
        assert float(a[2].item()) == pytest.approx(1e-4 + 65536 + (n if k else 0), rel=1e-6)
        assert not np.array_equal(a[0].cpu().numpy()[:nobs], np.zeros(nobs, np.float32))


def test_noalloc_rejects_bad_arguments():
    lib = _lib.load()
    d = 16
    for nenv, layout in ((0, 1), (-5, 1), (100, 2)):
        rc = lib.g1_policy_obs_norm_update_noalloc(nenv, layout, d, d, d, d, d, d, d, d, None)
        assert rc == -1 and "invalid arguments" in _lib.last_error()
