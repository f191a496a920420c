"""Policy-in-the-loop rollouts on the device.

The loop of examples/train.py:284-355 (without the PPO update), with every
per-step piece a HIP kernel on one stream and no host synchronisation:

    g1_policy_act     observation -> normalise/clip -> actor MLP (BF16 MFMA,
                      fused, activations in LDS) -> sampled action -> ctrl;
                      the push countdown and kick
    physics           g1_hip_step_ex (FP32 ABA, nsub substeps), or the
                      learned surrogate with its rotating ABA judge
    g1_rollout_post   reward, termination, timers, reset mask, commands
    g1_cuda_reset_done  resets the flagged envs (mask on the device)

The reference computes `resets.any()` and index lists on the host every step
(examples/train.py:334-344); here the reset kernel takes the device mask.

Network: examples/train.py:21-29 (actor), nn.Linear's default initialisation
U(+-1/sqrt(fan_in)) for weights and biases, log-std 0 (std 1,
examples/train.py:42), observation normaliser starting at mean 0, variance 1,
count 1e-4 and updated from every batch before it is applied, clip 5
(examples/train.py:44-47, 49-70, 86-88).
"""
import math

import numpy as np
import torch

from . import _lib
from .model import NC, NQ, NU, NV
from .sim import norm_device

OBS_DIM = {0: 99, 1: 96}
LAYERS = [(512, 128), (256, 512), (128, 256), (32, 128)]     # padded (out, in) of the fused kernel
HIDDEN = [512, 256, 128]


class Policy:
    """Actor MLP obs -> 512 -> 256 -> 128 -> 29 (ELU): bf16 weights (row-major
    W16 and the fused kernel's k-tiled Wt) and fp32 copies for checking."""

    def __init__(self, device, layout=1, seed=0, log_std=0.0, clip=5.0):
        self.lib = _lib.load()
        self.dev = norm_device(device)
        self.layout = layout
        self.nobs = OBS_DIM[layout]
        rng = np.random.default_rng(seed)
        true = [(HIDDEN[0], self.nobs), (HIDDEN[1], HIDDEN[0]), (HIDDEN[2], HIDDEN[1]), (NU, HIDDEN[2])]
        self.W32, self.b32, self.W16, self.b = [], [], [], []
        for (rows, cols), (out, fan_in) in zip(LAYERS, true):
            bound = 1.0 / math.sqrt(fan_in)
            w = np.zeros((rows, cols), np.float32)
            w[:out, :fan_in] = rng.uniform(-bound, bound, (out, fan_in))
            b = np.zeros(rows, np.float32)
            b[:out] = rng.uniform(-bound, bound, out)
            self.W32.append(w[:out, :fan_in].copy())
            self.b32.append(b[:out].copy())
            self.W16.append(torch.as_tensor(w, device=self.dev).to(torch.bfloat16).contiguous())
            self.b.append(torch.as_tensor(b, device=self.dev))
        self.obs_mean = torch.zeros(128, dtype=torch.float32, device=self.dev)
        self.obs_var = torch.ones(128, dtype=torch.float32, device=self.dev)
        self.obs_inv = torch.ones(128, dtype=torch.float32, device=self.dev)
        self.obs_count = torch.full((1,), 1e-4, dtype=torch.float32, device=self.dev)   # examples/train.py:46
        self._norm_scratch = torch.zeros(2 * 64 * 128 + 1, dtype=torch.float64, device=self.dev)
        self.clip = float(clip)
        self.log_std = None if log_std is None else torch.full((NU,), float(log_std), dtype=torch.float32,
                                                                device=self.dev)
        from .surrogate import k_tiled
        self.Wt = [k_tiled(w, rows) for w, (rows, _) in zip(self.W16, LAYERS)]
        self._ptrs = (_lib.ptr_array(self.Wt), _lib.ptr_array(self.b))
        self.update_normalizer(self.obs_mean, self.obs_var)

    def update_normalizer(self, mean, var):
        """Sets the running statistics (mean, variance) of the observation."""
        self.obs_mean.copy_(torch.as_tensor(mean, dtype=torch.float32))
        self.obs_var.copy_(torch.as_tensor(var, dtype=torch.float32))
        self.obs_inv.copy_((1.0 / torch.sqrt(self.obs_var.double() + 1e-8)).float())

    def update_obs_norm(self, nenv, qpos, qvel, commands, last_ctrl):
        """examples/train.py:49-66 on the device: merges this batch of
        observations into the running mean / variance / count."""
        s = torch.cuda.current_stream(self.dev).cuda_stream
        rc = self.lib.g1_policy_obs_norm_update(nenv, self.layout, qpos.data_ptr(), qvel.data_ptr(),
                                                commands.data_ptr(), last_ctrl.data_ptr(), self.obs_mean.data_ptr(),
                                                self.obs_var.data_ptr(), self.obs_count.data_ptr(),
                                                self.obs_inv.data_ptr(), self._norm_scratch.data_ptr(), s)
        _lib.check(rc, "g1_policy_obs_norm_update")

    def act(self, nenv, qpos, qvel, commands, last_ctrl, ctrl, actions=None, obs_out=None, push_interval=None,
            seed=0, t=0, sample=True):
        W, b = self._ptrs
        s = torch.cuda.current_stream(self.dev).cuda_stream
        rc = self.lib.g1_policy_act(nenv, self.layout, qpos.data_ptr(), qvel.data_ptr(), commands.data_ptr(),
                                    last_ctrl.data_ptr(), self.obs_mean.data_ptr(), self.obs_inv.data_ptr(),
                                    self.clip, W, b,
                                    self.log_std.data_ptr() if (sample and self.log_std is not None) else None,
                                    seed & 0xFFFFFFFF, t, ctrl.data_ptr(),
                                    None if actions is None else actions.data_ptr(),
                                    None if obs_out is None else obs_out.data_ptr(),
                                    None if push_interval is None else push_interval.data_ptr(), s)
        _lib.check(rc, "g1_policy_act")


class PolicyRollout:
    """nenv envs driven by a policy for `steps`-step rollouts (one policy
    step = nsub physics substeps at dt).  physics: "aba" (the fused FP32 step)
    or "surrogate" (a trained surrogate.MLP with the rotating ABA judge)."""

    def __init__(self, nenv, device, policy=None, physics="aba", surrogate=None, nsub=10, dt=2e-3, seed=0,
                 max_episode_length=1000, cmd_lin=1.0, cmd_yaw=1.0, sample=True, pushes=True, period=16,
                 tol=1e-3, update_norm=True):
        self.lib = _lib.load()
        self.dev = norm_device(device)
        self.nenv, self.nsub, self.dt, self.seed = nenv, nsub, float(dt), seed
        self.policy = policy or Policy(device, layout=1, seed=seed)
        self.physics = physics
        self.max_episode_length, self.cmd_lin, self.cmd_yaw = max_episode_length, cmd_lin, cmd_yaw
        self.sample, self.pushes, self.update_norm = sample, pushes, update_norm
        f32 = dict(dtype=torch.float32, device=self.dev)
        i32 = dict(dtype=torch.int32, device=self.dev)
        u8 = dict(dtype=torch.uint8, device=self.dev)
        self.qpos = torch.zeros((NQ, nenv), **f32)
        self.qvel = torch.zeros((NV, nenv), **f32)
        self.anchor = torch.full((2 * NC, nenv), 1e30, **f32)
        self.ctrl = torch.zeros((NU, nenv), **f32)
        self.actions = torch.zeros((NU, nenv), **f32)
        self.prev_actions = torch.zeros((NU, nenv), **f32)
        self.commands = torch.zeros((3, nenv), **f32)
        self.command_timeout = torch.zeros(nenv, **i32)
        self.episode_length = torch.zeros(nenv, **i32)
        self.push_interval = torch.zeros(nenv, **i32)
        self.reward = torch.zeros(nenv, **f32)
        self.done = torch.zeros(nenv, **u8)
        self.reset = torch.ones(nenv, **u8)
        self.foot_pos = torch.zeros((6, nenv), **f32)
        self.foot_vel = torch.zeros((6, nenv), **f32)
        self.contact_forces = torch.zeros((8, nenv), **f32)
        self.stepper = None
        if physics == "surrogate":
            from .surrogate import SurrogateStepper
            if surrogate is None:
                raise ValueError("physics='surrogate' needs a trained surrogate.MLP")
            # no per-period count read-back: the loop never waits for the device
            self.stepper = SurrogateStepper(surrogate, nenv, device, dt=dt, period=period, tol=tol, skip_idle=False)
            self.foot_pos, self.foot_vel = self.stepper.foot_pos, self.stepper.foot_vel
            self.contact_forces = self.stepper.contact_forces
        elif physics != "aba":
            raise ValueError(f"physics must be 'aba' or 'surrogate', not {physics!r}")
        self.t = 0
        s = self.stream
        # examples/train.py:243, 248, 274-276
        self.lib.g1_cuda_reset_done(nenv, self.qpos.data_ptr(), self.qvel.data_ptr(), self.anchor.data_ptr(),
                                    self.reset.data_ptr(), self._seed(-1), 0.1, 0.02, s)
        _lib.check(self.lib.g1_rollout_init(nenv, cmd_lin, cmd_yaw, seed & 0xFFFFFFFF, self.commands.data_ptr(),
                                            self.command_timeout.data_ptr(), self.episode_length.data_ptr(),
                                            self.push_interval.data_ptr() if pushes else None, s),
                   "g1_rollout_init")

    @property
    def stream(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _seed(self, t):
        return (self.seed * 0x9E3779B1 + (t + 2) * 0x85EBCA77) & 0xFFFFFFFF

    def step(self, reward_sum=None, done_count=None):
        """One policy step: act, nsub physics substeps, reward/termination,
        bookkeeping and resets.  reward_sum / done_count: optional float64
        device tensors (one element) that the post-step kernel increments by
        this step's summed reward and number of terminations."""
        self.act_phase()
        self.physics_phase()
        self.post_phase(reward_sum, done_count)
        self.reset_phase()
        self.t += 1

    # The four launches of a step, in order (step() runs them; they are
    # separate so a caller can observe the state between them).
    def act_phase(self):
        """Normaliser update (act(obs) updates it first, examples/train.py:86-88),
        then the policy: actions, ctrl and pushes."""
        n = self.nenv
        if self.update_norm:
            self.policy.update_obs_norm(n, self.qpos, self.qvel, self.commands, self.ctrl)
        self.policy.act(n, self.qpos, self.qvel, self.commands, self.ctrl, self.ctrl, actions=self.actions,
                        push_interval=self.push_interval if self.pushes else None, seed=self.seed, t=self.t,
                        sample=self.sample)

    def physics_phase(self):
        """nsub physics steps: the fused FP32 ABA step, or the surrogate with
        its judge; the feet and contact forces are left for the reward."""
        lib, s, n = self.lib, self.stream, self.nenv
        if self.stepper is None:
            rc = lib.g1_hip_step_ex(n, self.nsub, self.dt, self.qpos.data_ptr(), self.qvel.data_ptr(),
                                    self.anchor.data_ptr(), self.ctrl.data_ptr(), self.foot_pos.data_ptr(),
                                    self.foot_vel.data_ptr(), self.contact_forces.data_ptr(), None, None, None,
                                    None, 0, s)
            _lib.check(rc)
        else:
            self.stepper.steps(self.qpos, self.qvel, self.anchor, self.ctrl, self.nsub)

    def post_phase(self, reward_sum=None, done_count=None):
        """Reward, termination, timers, commands and the reset mask."""
        lib, s, n, t = self.lib, self.stream, self.nenv, self.t
        rc = lib.g1_rollout_post(n, self.qpos.data_ptr(), self.qvel.data_ptr(), self.foot_pos.data_ptr(),
                                 self.foot_vel.data_ptr(), self.contact_forces.data_ptr(), self.actions.data_ptr(),
                                 self.commands.data_ptr(), self.prev_actions.data_ptr(),
                                 self.command_timeout.data_ptr(), self.episode_length.data_ptr(),
                                 self.push_interval.data_ptr() if self.pushes else None, self.max_episode_length,
                                 self.cmd_lin, self.cmd_yaw, self._seed(t) ^ 0x5bd1e995, t,
                                 self.reward.data_ptr(), self.done.data_ptr(), self.reset.data_ptr(),
                                 None if reward_sum is None else reward_sum.data_ptr(),
                                 None if done_count is None else done_count.data_ptr(), s)
        _lib.check(rc, "g1_rollout_post")

    def reset_phase(self):
        """Resets the envs of the device mask (and returns them to the
        surrogate's fallback)."""
        if self.stepper is not None:              # a reset env is stepped by ABA until the judge accepts it
            self.stepper.mark_reset(self.reset)
        self.lib.g1_cuda_reset_done(self.nenv, self.qpos.data_ptr(), self.qvel.data_ptr(), self.anchor.data_ptr(),
                                    self.reset.data_ptr(), self._seed(self.t), 0.1, 0.02, self.stream)

    def rollout(self, steps=24, reward_sums=None, done_counts=None):
        """examples/train.py:284-355 without the buffer writes: prev_actions
        start at zero; per-step reward sums / done counts (contiguous float64
        device tensors of length >= steps, optional) are accumulated by the
        post-step kernel."""
        _lib.check(self.lib.g1_mlp_fill(self.prev_actions.data_ptr(), 0.0, self.prev_actions.numel(), self.stream),
                   "fill")
        for acc in (reward_sums, done_counts):
            if acc is not None and (acc.dtype != torch.float64 or acc.device != self.dev or not acc.is_contiguous()
                                    or acc.numel() < steps):
                raise ValueError(f"reward_sums / done_counts must be contiguous float64 tensors of >= {steps} "
                                 f"elements on {self.dev}")
        for k in range(steps):
            self.step(None if reward_sums is None else reward_sums[k:k + 1],
                      None if done_counts is None else done_counts[k:k + 1])
