"""CPU oracle of the RL side of the step loop (numpy, float64).

TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker.  The
product (g1sim_amd) never imports it.

Restates, from the reference's Python text:
  * the observation of G1Sim.get_obs            src/g1sim.py:132-170
  * quat_rotate_inverse                          src/g1sim.py:149-156
  * reward and termination                       examples/train.py:143-225
  * observation normalisation and its running
    statistics                                   examples/train.py:49-70
  * the policy MLP (96/99 -> 512 -> 256 -> 128 -> 29, ELU)
                                                 examples/train.py:21-29
  * ctrl from an action                          examples/train.py:299
  * the per-step bookkeeping (timers, resets)    examples/train.py:319-351
All arrays are SoA [dim][nenv] like the device state unless stated.
"""
import numpy as np

NU = 29

# src/g1sim.py:76-82
ACTION_SCALE = np.array([
    0.443200, 0.274054, 0.443200, 0.274054, 0.365131, 0.365131,
    0.443200, 0.274054, 0.443200, 0.274054, 0.365131, 0.365131,
    0.443200, 0.365131, 0.365131,
    0.365131, 0.365131, 0.365131, 0.365131, 0.365131, 0.074501, 0.074501,
    0.365131, 0.365131, 0.365131, 0.365131, 0.365131, 0.074501, 0.074501])

# examples/train.py:256-270
STD_WALKING = np.array([0.3, 0.15, 0.15, 0.35, 0.25, 0.1, 0.3, 0.15, 0.15, 0.35, 0.25, 0.1, 0.2, 0.08, 0.1,
                        0.15, 0.15, 0.1, 0.15, 0.3, 0.3, 0.3, 0.15, 0.15, 0.1, 0.15, 0.3, 0.3, 0.3])
STD_RUNNING = np.array([0.5, 0.2, 0.2, 0.6, 0.35, 0.15, 0.5, 0.2, 0.2, 0.6, 0.35, 0.15, 0.3, 0.08, 0.2,
                        0.5, 0.2, 0.15, 0.35, 0.3, 0.3, 0.3, 0.5, 0.2, 0.15, 0.35, 0.3, 0.3, 0.3])


def quat_rotate_inverse(q, v):
    """src/g1sim.py:149-156: v + 2 (w (u x v) + u x (u x v)), u = -xyz;
    q [4][n] (w, x, y, z), v [3][n]."""
    w = q[0]
    u = -q[1:4]
    uv = np.cross(u, v, axis=0)
    uuv = np.cross(u, uv, axis=0)
    return v + 2.0 * (w * uv + uuv)


def get_obs(qpos, qvel, commands, last_actions, default_joint_pos, layout=0):
    """src/g1sim.py:132-170 (layout 0, (n, 99)) and the 96-dim order of the
    policy loop (layout 1: angular velocity, projected gravity, commands,
    q - default, qd, last ctrl).  Returns (n, nobs)."""
    n = qpos.shape[1]
    grav = np.zeros((3, n)); grav[2] = -1.0
    pg = quat_rotate_inverse(qpos[3:7], grav)
    lv = quat_rotate_inverse(qpos[3:7], qvel[0:3])
    jp = qpos[7:36] - np.asarray(default_joint_pos).reshape(NU, 1)
    if layout == 0:
        parts = [commands, pg, lv, qvel[3:6], jp, qvel[6:35], last_actions]
    else:
        parts = [qvel[3:6], pg, commands, jp, qvel[6:35], last_actions]
    return np.concatenate(parts, axis=0).T


def reward_and_done(obs99, qpos, foot_pos, foot_vel, contact_forces, actions, prev_actions, jnt_lo, jnt_hi):
    """examples/train.py:143-225.  obs99: (n, 99) as get_obs; actions and
    prev_actions: (n, 29); jnt_lo / jnt_hi: (29,).  Returns (reward (n,),
    done (n,) bool)."""
    commands = obs99[:, 0:3]
    proj_gravity = obs99[:, 3:6]
    lin_vel = obs99[:, 6:9]
    ang_vel = obs99[:, 9:12]
    xy_error = np.sum((commands[:, :2] - lin_vel[:, :2]) ** 2, axis=1)
    z_error = lin_vel[:, 2] ** 2
    rew_lin_vel = np.exp(-(xy_error + z_error) / 0.25)
    yaw_error = (commands[:, 2] - ang_vel[:, 2]) ** 2
    roll_pitch_error = np.sum(ang_vel[:, :2] ** 2, axis=1)
    rew_ang_vel = np.exp(-(yaw_error + roll_pitch_error) / 0.5)
    rew_orient = np.exp(-np.sum(proj_gravity[:, :2] ** 2, axis=1) / 0.2)
    total_speed = np.linalg.norm(commands[:, :2], axis=1) + np.abs(commands[:, 2])
    standing = (total_speed < 0.05)[:, None]
    running = (total_speed >= 1.5)[:, None]
    std = np.where(standing, 0.05, np.where(running, STD_RUNNING[None, :], STD_WALKING[None, :]))
    pose_error = np.mean(obs99[:, 12:41] ** 2 / std ** 2, axis=1)
    rew_pose = np.exp(-pose_error)
    rew_action_rate = np.sum((actions - prev_actions) ** 2, axis=1)
    rew_body_ang_vel = np.sum(ang_vel[:, :2] ** 2, axis=1)
    q = qpos[7:36]
    rng = (jnt_hi - jnt_lo)[:, None]
    soft_lo = jnt_lo[:, None] + 0.05 * rng
    soft_hi = jnt_hi[:, None] - 0.05 * rng
    rew_dof_limits = np.sum(np.maximum(q - soft_hi, 0.0) + np.maximum(soft_lo - q, 0.0), axis=0)
    left_z, right_z = foot_pos[2], foot_pos[5]
    lv2 = foot_vel[0] ** 2 + foot_vel[1] ** 2
    rv2 = foot_vel[3] ** 2 + foot_vel[4] ** 2
    active = (total_speed > 0.05).astype(np.float64)
    clearance = (np.abs(left_z - 0.1) * np.sqrt(lv2 + 1e-6) + np.abs(right_z - 0.1) * np.sqrt(rv2 + 1e-6)) * active
    lc = (np.sum(contact_forces[0:4], axis=0) > 1.0).astype(np.float64)
    rc = (np.sum(contact_forces[4:8], axis=0) > 1.0).astype(np.float64)
    slip = (lv2 * lc + rv2 * rc) * active
    reward = (2.0 * rew_lin_vel + 2.0 * rew_ang_vel + 1.0 * rew_orient + 1.0 * rew_pose - 1.0 * rew_dof_limits
              - 2.0 * clearance - 0.5 * slip - 0.05 * rew_action_rate - 0.05 * rew_body_ang_vel)
    limit_angle = 70.0 * 3.14159 / 180.0
    tilt = np.abs(np.arccos(np.clip(-proj_gravity[:, 2], -1.0, 1.0)))
    return reward, tilt > limit_angle


def update_obs_norm(obs, mean, var, count):
    """examples/train.py:49-66: merges a batch (n, nobs) into the running
    (mean, var, count); returns the new triple."""
    batch_mean = obs.mean(axis=0)
    batch_var = obs.var(axis=0)
    n = obs.shape[0]
    delta = batch_mean - mean
    tot = count + n
    new_mean = mean + delta * n / tot
    M2 = var * count + batch_var * n + delta ** 2 * count * n / tot
    return new_mean, M2 / tot, tot


def normalize_obs(obs, mean, var, clip=5.0):
    """examples/train.py:69-70."""
    return np.clip((obs - mean) / np.sqrt(var + 1e-8), -clip, clip)


def elu(z):
    return np.where(z > 0, z, np.expm1(np.minimum(z, 0.0)))


def policy_mean(x, weights, biases, round_fn=None):
    """examples/train.py:21-29 forward (actor mean): x (n, nin); weights
    [(out, in)]; round_fn, when given, rounds the layer inputs (the kernel's
    bf16 operands)."""
    r = round_fn or (lambda a: a)
    h = r(x)
    for l, (W, b) in enumerate(zip(weights, biases)):
        z = h @ r(W).T + b
        h = r(elu(z)) if l < len(weights) - 1 else z
    return h


def ctrl_from_action(action, default_joint_pos):
    """examples/train.py:299: ctrl = default + action * scale; action (n, 29),
    returns [29][n]."""
    return np.asarray(default_joint_pos).reshape(NU, 1) + action.T * ACTION_SCALE[:, None]


def bookkeeping(done, command_timeout, episode_length, max_episode_length):
    """examples/train.py:319-351 without the random draws: returns (reset,
    command_timeout, episode_length, resample) where resample marks envs
    whose commands are redrawn (reset envs and envs whose timer ran out; the
    timers of both are then redrawn)."""
    ct = command_timeout - 1
    el = episode_length + 1
    reset = done | (el >= max_episode_length)
    el = np.where(reset, 0, el)
    timer_out = (~reset) & (ct <= 0)
    return reset, ct, el, reset | timer_out
