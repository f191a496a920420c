"""The torch drop-in G1Sim on the GPU: the reference's API tests
(tests/test_g1sim.py:13-169) with the MuJoCo kinematics check replaced by the
FP64 oracle's forward kinematics (the reference validates that against
MuJoCo to < 1e-12, tests/test_vs_mujoco.py:51-62)."""
import numpy as np
import pytest
import torch

from g1sim_amd import G1Sim
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_g1sim_init(device):
    """Restates tests/test_g1sim.py:13-20 (shapes and dtype)."""
    env = G1Sim(nenv=10, device=device)
    assert env.qpos.shape == (36, 10)
    assert env.qvel.shape == (35, 10)
    assert env.anchor.shape == (16, 10)
    assert env.ctrl.shape == (29, 10)
    assert env.done.shape == (10,)
    assert env.done.dtype == torch.uint8


def test_g1sim_reset_all(device):
    """Restates tests/test_g1sim.py:21-36 (reset overwrites the state; the
    stand keyframe's 0.79 m pelvis height)."""
    env = G1Sim(nenv=5, device=device)
    env.qpos.fill_(999.0)
    env.qvel.fill_(999.0)
    env.reset_all(seed=42, noise=0.0, drop=0.0)
    torch.cuda.synchronize()
    assert torch.all(env.qvel == 0.0)
    assert not torch.any(env.qpos == 999.0)
    assert torch.allclose(env.qpos[2, :], torch.tensor(0.79, device=device), atol=0.05)


def test_g1sim_step(device):
    """Restates tests/test_g1sim.py:38-50 (the state moves); last_actions is
    ctrl after a step (src/g1sim.py:115)."""
    env = G1Sim(nenv=2, device=device)
    env.reset_all(seed=42, noise=0.0, drop=0.0)
    torch.cuda.synchronize()
    q0 = env.qpos.clone()
    env.step(nsub=10)
    torch.cuda.synchronize()
    assert not torch.allclose(env.qpos, q0)
    assert torch.equal(env.last_actions, env.ctrl)


def test_g1sim_selective_reset(device):
    """Restates tests/test_g1sim.py:52-76 (only the done env is reset, to
    0.79 + 1.0 m with drop 1.0)."""
    env = G1Sim(nenv=2, device=device)
    env.reset_all(seed=42, noise=0.0, drop=0.0)
    env.step(nsub=50)
    torch.cuda.synchronize()
    stepped = env.qpos.clone()
    env.done[0] = 0
    env.done[1] = 1
    env.reset_done(env.done, seed=42, noise=0.0, drop=1.0)
    torch.cuda.synchronize()
    assert torch.equal(env.qpos[:, 0], stepped[:, 0])
    assert not torch.allclose(env.qpos[:, 1], stepped[:, 1])
    assert torch.allclose(env.qpos[2, 1], torch.tensor(1.79, device=device), atol=0.05)


def test_g1sim_get_obs(device):
    """Restates tests/test_g1sim.py:78-91 (shape, projected gravity at rest,
    zero joint offsets at the keyframe)."""
    env = G1Sim(nenv=5, device=device)
    env.reset_all(noise=0.0, drop=0.0)
    obs = env.get_obs()
    assert obs.shape == (5, 99)
    assert torch.allclose(obs[:, 3:6], torch.tensor([0.0, 0.0, -1.0], device=device).expand(5, 3), atol=1e-5)
    assert torch.allclose(obs[:, 12:41], torch.zeros(5, 29, device=device), atol=1e-5)


def test_g1sim_foot_data(device):
    """Restates tests/test_g1sim.py:93-119 (foot output shapes, foot heights
    in (-0.05, 0.15) m after a short stance, some contact force)."""
    env = G1Sim(nenv=2, device=device)
    assert env.foot_pos.shape == (6, 2) and env.foot_vel.shape == (6, 2)
    assert env.contact_forces.shape == (8, 2)
    env.reset_all(seed=42, noise=0.0, drop=0.0)
    env.step(nsub=10)
    torch.cuda.synchronize()
    assert not torch.all(env.foot_pos == 0.0)
    assert torch.all(env.foot_pos[2, :] < 0.15) and torch.all(env.foot_pos[5, :] < 0.15)
    assert torch.all(env.foot_pos[2, :] > -0.05) and torch.all(env.foot_pos[5, :] > -0.05)
    assert torch.any(env.contact_forces > 0.0)


def test_g1sim_foot_kinematics_vs_oracle(device):
    """Foot outputs are the pre-step kinematics of bodies 6 and 12
    (src/g1_kernel.cu:128-145): world origin and world linear velocity.  The
    reference's check (tests/test_g1sim.py:121-169) compares them with
    MuJoCo at atol 1e-5; here the FP64 oracle's kinematics stand in for
    MuJoCo, with the same tolerance."""
    env = G1Sim(nenv=64, device=device)
    env.reset_all(seed=42, noise=0.1, drop=0.0)
    env.qvel.normal_(0.0, 0.5)
    torch.cuda.synchronize()
    qp = env.qpos.cpu().numpy().astype(np.float64)
    qv = env.qvel.cpu().numpy().astype(np.float64)
    env.step(nsub=1)
    torch.cuda.synchronize()
    fp = env.foot_pos.cpu().numpy()
    fv = env.foot_vel.cpu().numpy()
    for e in range(64):
        xpos, xmat = O.fk(qp[:, e])
        assert np.allclose(fp[0:3, e], xpos[6], atol=1e-5)
        assert np.allclose(fp[3:6, e], xpos[12], atol=1e-5)
        o = O.step_full(qp[:, e], qv[:, e], env.ctrl[:, e].cpu().numpy().astype(float),
                        np.full(16, 1e30), 1, 1e-3)
        assert np.allclose(fv[:, e], o["foot_vel"], atol=1e-5)


def test_g1sim_additive_options(device):
    env = G1Sim(nenv=256, device=device, dt=2e-3, mu=0.8, track_qacc=True)
    env.reset_all(seed=1, noise=0.05, drop=0.02)
    env.step(nsub=5)
    torch.cuda.synchronize()
    assert torch.isfinite(env.qacc).all() and env.qacc.abs().sum() > 0


def test_g1sim_rejects_reassigned_state_before_launch(device):
    """A state attribute replaced by a tensor of another shape is refused on
    the host (the kernels would otherwise index it out of bounds)."""
    sim = G1Sim(128, device)
    sim.reset_all()
    sim.ctrl = torch.zeros((28, 128), device=device)
    with pytest.raises(ValueError):
        sim.step()
    sim.ctrl = torch.zeros((29, 128), device=device, dtype=torch.float64)
    with pytest.raises(ValueError):
        sim.get_obs()
    sim.ctrl = torch.zeros((29, 128), device=device)
    sim.step()
    torch.cuda.synchronize()


def test_g1sim_on_an_unindexed_device(device):
    """G1Sim(n, 'cuda') compares its tensors against the indexed device."""
    sim = G1Sim(64, "cuda")
    sim.reset_all()
    sim.step()
    assert sim.get_obs().shape == (64, 99)
