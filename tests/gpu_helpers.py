"""Helpers of the GPU parity tests: drive libg1sim_amd.so through its C-ABI on
torch tensors, and compare with the FP64 oracle on the same float32 inputs."""
import json
import os

import numpy as np
import torch

from g1sim_amd import _lib
from oracle import oracle as O


def to_dev(a, device):
    return torch.as_tensor(np.ascontiguousarray(a), device=device)


class GpuState:
    """Float32 SoA state on the device plus output buffers."""

    def __init__(self, qpos, qvel, anchor, ctrl, device, mu=None):
        self.device = device
        self.n = qpos.shape[1]
        self.qpos = to_dev(qpos.astype(np.float32), device)
        self.qvel = to_dev(qvel.astype(np.float32), device)
        self.anchor = to_dev(anchor.astype(np.float32), device)
        self.ctrl = to_dev(ctrl.astype(np.float32), device)
        self.mu = None if mu is None else to_dev(np.asarray(mu, np.float32), device)
        z = lambda d: torch.zeros((d, self.n), dtype=torch.float32, device=device)  # noqa: E731
        self.foot_pos, self.foot_vel, self.fn, self.qacc, self.tau = z(6), z(6), z(8), z(35), z(35)
        self.base64 = None

    def enable_f64_base(self):
        self.base64 = self.qpos[:7].double().contiguous()

    def step(self, nsub=1, dt=2e-3, flags=0, qacc=True, tau=False):
        lib = _lib.load()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        rc = lib.g1_hip_step_ex(
            self.n, nsub, dt, self.qpos.data_ptr(), self.qvel.data_ptr(), self.anchor.data_ptr(),
            self.ctrl.data_ptr(), self.foot_pos.data_ptr(), self.foot_vel.data_ptr(), self.fn.data_ptr(),
            self.qacc.data_ptr() if qacc else None, self.tau.data_ptr() if tau else None,
            None if self.mu is None else self.mu.data_ptr(),
            None if self.base64 is None else self.base64.data_ptr(), flags, stream)
        _lib.check(rc)

    def host(self):
        torch.cuda.synchronize(self.device)
        return {k: getattr(self, k).cpu().numpy() for k in
                ("qpos", "qvel", "anchor", "ctrl", "foot_pos", "foot_vel", "fn", "qacc", "tau")}


def rel_err(a, ref):
    """|a - ref| / max(|ref|, 1) elementwise (tests/test_vs_mujoco.py:74-75)."""
    return np.abs(a - ref) / np.maximum(np.abs(ref), 1.0)


def branch_flips(gpu_fn, ref_fn, gpu_anchor, ref_anchor, anchor_tol=1e-4):
    """Envs where a contact decision of the float step differs from the FP64
    step on the same input: on/off (fn == 0), or stick/slip (anchor moved by
    a different amount)."""
    on = (gpu_fn > 0) != (ref_fn > 0)
    fa, ra = gpu_anchor >= 1e29, ref_anchor >= 1e29
    an = (fa != ra) | (~fa & ~ra & (np.abs(gpu_anchor - np.where(ra, 0, ref_anchor)) > anchor_tol))
    return on.any(axis=0) | an.any(axis=0)


def teacher_forced_step(st, dt, mu=None):
    """One GPU step from the current device state and the oracle's step from
    the same float32 state.  Returns (before, gpu_after, ref)."""
    before = st.host()
    st.step(1, dt)
    after = st.host()
    ref = O.step_batch_f32(before["qpos"], before["qvel"], before["anchor"], before["ctrl"], dt,
                           mu=mu, want=("qacc", "qpos", "qvel", "anchor", "fn", "foot_pos", "foot_vel"))
    return before, after, ref


def f32_reference_step(before, dt, mu=None):
    """The reference algorithm evaluated in float32 on the CPU (the oracle
    source built with real = float) from the same state: the yardstick for
    what float32 arithmetic can reach on these states."""
    n = before["qpos"].shape[1]
    out = {"qacc": np.zeros((35, n)), "fn": np.zeros((8, n)), "anchor": np.zeros((16, n))}
    args = [np.ascontiguousarray(before[k], np.float32) for k in ("qpos", "qvel", "anchor", "ctrl")]
    m = None if mu is None else np.ascontiguousarray(mu, np.float32)
    O.lib32().g1o_step_batch_f32(n, dt, *(O.fptr(a) for a in args), O.fptr(m), O.dptr(out["qacc"]), None, None,
                                 O.dptr(out["anchor"]), O.dptr(out["fn"]), None, None, None, 0)
    return out


def write_metrics(name, data):
    d = os.environ.get("G1SIM_METRICS_DIR")
    print(f"[metrics] {name}: {json.dumps(data)}", flush=True)
    if d:
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, name + ".json"), "w") as f:
            json.dump(data, f, indent=1)
