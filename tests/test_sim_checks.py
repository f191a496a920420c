"""Host-side checks of the G1Sim wrapper that run before any launch (CPU)."""
import pytest
import torch

from g1sim_amd.sim import check_soa

CPU = torch.device("cpu")


def test_check_soa_accepts_the_kernel_layout():
    check_soa("qpos", torch.zeros((36, 10)), 36, 10, CPU)


@pytest.mark.parametrize("t", [
    torch.zeros((35, 10)),                       # wrong row count
    torch.zeros((36, 11)),                       # wrong env count
    torch.zeros((36, 10), dtype=torch.float64),  # wrong dtype
    torch.zeros((10, 36)).t(),                   # right shape, not contiguous
    [[0.0] * 10] * 36,                           # not a tensor
])
def test_check_soa_rejects(t):
    with pytest.raises(ValueError):
        check_soa("qpos", t, 36, 10, CPU)
