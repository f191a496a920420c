"""G1Sim: the torch-ROCm drop-in of the reference's Python API (src/g1sim.py:35-170).

Same tensors (float32 structure-of-arrays ``[dim][nenv]`` on the device), same
methods, same argument checks; the physics runs in libg1sim_amd.so on the
caller's current HIP stream.  Model data (stand keyframe, joint ranges) comes
from the package's own model file instead of MuJoCo (the reference reads them
through ``mujoco.MjModel`` at src/g1sim.py:63-72).

Additive, keyword-only options (defaults reproduce the reference exactly):
``dt`` (the reference launcher fixes 1e-3, src/g1_kernel.cu:157), per-env
friction ``mu`` and ``track_qacc`` (keeps the last substep's qacc in
``self.qacc``).
"""
import ctypes

import torch

from . import _lib
from .model import MODEL, NC, NQ, NU, NV

# src/g1sim.py:76-82: 0.25 * effort limit / stiffness per actuator (mjlab gains)
ACTION_SCALE = [
    0.443200, 0.274054, 0.443200, 0.274054, 0.365131, 0.365131,
    0.443200, 0.274054, 0.443200, 0.274054, 0.365131, 0.365131,
    0.443200, 0.365131, 0.365131,
    0.365131, 0.365131, 0.365131, 0.365131, 0.365131, 0.074501, 0.074501,
    0.365131, 0.365131, 0.365131, 0.365131, 0.365131, 0.074501, 0.074501,
]

REFERENCE_DT = 1e-3


def _stream_ptr(device):
    return torch.cuda.current_stream(device).cuda_stream


def norm_device(device):
    """torch.device with its index ('cuda' -> 'cuda:<current>'), so that it
    compares equal to the device of the tensors allocated on it."""
    d = torch.device(device)
    if d.type == "cuda" and d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


def check_soa(name, t, rows, nenv, device, dtype=torch.float32):
    """A state tensor as the kernels index it: [rows][nenv] of dtype
    (float32 by default), contiguous, on the simulator's device (a
    reassigned attribute of another shape would otherwise be read out of
    bounds on the GPU)."""
    if not isinstance(t, torch.Tensor) or t.dtype != dtype or tuple(t.shape) != (rows, nenv) \
            or t.device != device or not t.is_contiguous():
        got = (tuple(t.shape), t.dtype, str(t.device), t.is_contiguous()) if isinstance(t, torch.Tensor) else type(t)
        raise ValueError(f"{name} must be a contiguous {dtype} ({rows}, {nenv}) tensor on {device}; got {got}")


class G1Sim:
    def __init__(self, nenv: int, device: torch.device, *, dt: float = REFERENCE_DT,
                 mu=None, track_qacc: bool = False):
        self.lib = _lib.load()
        self.nenv = nenv
        if torch.device(device).type != "cuda":
            raise ValueError("G1Sim runs on a HIP device (torch 'cuda' device on ROCm)")
        self.device = norm_device(device)
        self.NQ, self.NV, self.NU, self.NC = NQ, NV, NU, NC
        f32 = dict(dtype=torch.float32, device=self.device)
        # SoA layout: (dim, nenv)  (src/g1sim.py:47-53)
        self.qpos = torch.zeros((NQ, nenv), **f32)
        self.qvel = torch.zeros((NV, nenv), **f32)
        self.anchor = torch.full((2 * NC, nenv), 1e30, **f32)
        self.ctrl = torch.zeros((NU, nenv), **f32)
        self.foot_pos = torch.zeros((6, nenv), **f32)
        self.foot_vel = torch.zeros((6, nenv), **f32)
        self.contact_forces = torch.zeros((8, nenv), **f32)
        # RL state (src/g1sim.py:56-60)
        self.commands = torch.zeros((3, nenv), **f32)
        self.last_actions = torch.zeros((NU, nenv), **f32)
        self.done = torch.ones(nenv, dtype=torch.uint8, device=self.device)
        # stand pose and joint limits from the model (src/g1sim.py:62-72)
        self.default_qpos = torch.tensor(MODEL.qpos_stand, **f32).unsqueeze(1)
        self.default_joint_pos = self.default_qpos[7:36]
        self.global_gravity = torch.zeros((3, nenv), **f32)
        self.global_gravity[2, :] = -1.0
        self.joint_pos_lower = torch.tensor(MODEL.jnt_lo, **f32).unsqueeze(1)
        self.joint_pos_upper = torch.tensor(MODEL.jnt_hi, **f32).unsqueeze(1)
        self.action_scale = torch.tensor(ACTION_SCALE, **f32).unsqueeze(1)
        # additive options
        self.dt = float(dt)
        self.mu = None
        if mu is not None:
            self.mu = torch.as_tensor(mu, **f32).expand(nenv).contiguous()
        self.qacc = torch.zeros((NV, nenv), **f32) if track_qacc else None

    def reset_done(self, done: torch.Tensor, seed: int = 42, noise: float = 0.05, drop: float = 0.02):
        """Resets environments where done is True/1 (src/g1sim.py:84-103)."""
        assert done.shape == (self.nenv,)
        assert done.dtype == torch.uint8
        assert done.device == self.device
        assert done.is_contiguous()
        for name, rows in (("qpos", NQ), ("qvel", NV), ("anchor", 2 * NC)):
            check_soa(name, getattr(self, name), rows, self.nenv, self.device)
        self.lib.g1_cuda_reset_done(self.nenv, self.qpos.data_ptr(), self.qvel.data_ptr(),
                                    self.anchor.data_ptr(), done.data_ptr(), ctypes.c_uint(seed & 0xFFFFFFFF),
                                    noise, drop, _stream_ptr(self.device))

    def reset_all(self, seed: int = 42, noise: float = 0.05, drop: float = 0.02):
        """Resets all environments (src/g1sim.py:105-108)."""
        self.done.fill_(1)
        self.reset_done(self.done, seed, noise, drop)

    def _check_state(self):
        n, d = self.nenv, self.device
        for name, rows in (("qpos", NQ), ("qvel", NV), ("anchor", 2 * NC), ("ctrl", NU), ("foot_pos", 6),
                           ("foot_vel", 6), ("contact_forces", 8), ("commands", 3), ("last_actions", NU)):
            check_soa(name, getattr(self, name), rows, n, d)
        if self.qacc is not None:
            check_soa("qacc", self.qacc, NV, n, d)
        if self.mu is not None:
            mu = self.mu
            if not isinstance(mu, torch.Tensor) or mu.dtype != torch.float32 or tuple(mu.shape) != (n,) \
                    or mu.device != d or not mu.is_contiguous():
                raise ValueError(f"mu must be a contiguous float32 ({n},) tensor on {d}")

    def step(self, nsub: int = 1):
        """Steps all environments forward by nsub substeps (src/g1sim.py:110-130)."""
        assert self.ctrl.is_contiguous()
        assert self.qpos.is_contiguous()
        self._check_state()
        self.last_actions.copy_(self.ctrl)
        stream = _stream_ptr(self.device)
        if self.dt == REFERENCE_DT and self.mu is None and self.qacc is None:
            self.lib.g1_cuda_step(self.nenv, nsub, self.qpos.data_ptr(), self.qvel.data_ptr(),
                                  self.anchor.data_ptr(), self.ctrl.data_ptr(), self.foot_pos.data_ptr(),
                                  self.foot_vel.data_ptr(), self.contact_forces.data_ptr(), stream)
            return
        rc = self.lib.g1_hip_step_ex(
            self.nenv, nsub, self.dt, self.qpos.data_ptr(), self.qvel.data_ptr(), self.anchor.data_ptr(),
            self.ctrl.data_ptr(), self.foot_pos.data_ptr(), self.foot_vel.data_ptr(),
            self.contact_forces.data_ptr(), None if self.qacc is None else self.qacc.data_ptr(), None,
            None if self.mu is None else self.mu.data_ptr(), None, 0, stream)
        _lib.check(rc)

    def get_obs(self) -> torch.Tensor:
        """The 99-dim RL observation (src/g1sim.py:132-170): commands (3),
        projected gravity (3), base linear velocity in the base frame (3), base
        angular velocity (3), joint position error (29), joint velocities (29),
        previous actions (29).  Returns (nenv, 99), built by one HIP kernel
        (g1_hip_get_obs) on the current stream."""
        self._check_state()
        obs = torch.empty((self.nenv, 99), dtype=torch.float32, device=self.device)
        rc = self.lib.g1_hip_get_obs(self.nenv, 0, self.qpos.data_ptr(), self.qvel.data_ptr(),
                                     self.commands.data_ptr(), self.last_actions.data_ptr(), obs.data_ptr(),
                                     _stream_ptr(self.device))
        _lib.check(rc, "g1_hip_get_obs")
        return obs
