"""ORACLE — test infrastructure only.  numpy restatement of the go2_omni policy observation of
the active-exploration rollout: gait clock, observation terms, scales, sorted concatenation,
short history and clipping.  Only tests may import this module.

Follows:
  * gait clock: spigym/envs/locomotion/go2_omni.py:348-379 (gait_indices += dt * frequency,
    foot indices from phase / offset / bound, warped by the duration, clock_inputs =
    sin(2 pi foot_index));
  * observation getters: legged_robot_base.py:795-805 (base_ang_vel, projected_gravity,
    dof_pos - default, dof_vel), :843-844 (actions), go2_locomotion.py:273-277 (command lin /
    ang vel), go2_omni.py:627-668 (body height -- with the evaluation-time sinusoid --, gait
    frequency, gait phase 5:9, footswing, body attitude, stance, clock inputs);
  * scales: config/obs/loco/go2_omni.yaml:109-127; concatenation in sorted key order and the
    short history (newest first, per key in sorted order): legged_robot_base.py:503-527,
    819-829; history buffer shift: envs/env_utils/history_handler.py:36-40;
  * clip_observations 100: config/env/legged_base.yaml:21.
"""
from __future__ import annotations

import numpy as np

TERMS = {  # name: (dim, scale)
    "actions": (12, 1.0), "base_ang_vel": (3, 0.25), "clock_inputs": (4, 1.0), "command_ang_vel": (1, 1.0),
    "command_body_attitude": (2, 0.3), "command_body_height": (1, 2.0), "command_footswing_height": (1, 0.15),
    "command_gait_freq": (1, 1.0), "command_gait_phase": (4, 1.0), "command_lin_vel": (2, 1.0),
    "command_stance": (2, 1.0), "dof_pos": (12, 1.0), "dof_vel": (12, 0.05), "projected_gravity": (3, 1.0),
}
HISTORY = 14
CLIP = 100.0
HEIGHT_LIMITS = (-0.25, 0.15)


def quat_rotate_inverse(q, v):
    """v in the frame of quaternion q (xyzw), torch_utils.py quat_rotate_inverse."""
    x, y, z, w = q
    qv = np.array([x, y, z])
    a = v * (2.0 * w * w - 1.0)
    b = np.cross(qv, v) * w * 2.0
    c = qv * (qv @ v) * 2.0
    return a - b + c


def gait_update(gait_index, cmd, dt):
    """One _step_contact_targets: returns (new gait index, clock_inputs [4]).  The foot indices
    are warped in place (stance part to [0, 0.5), swing part to [0.5, 1)) before the clock is
    taken from them (go2_omni.py:366-379)."""
    freq, phase, offset, bound, dur = cmd[4], cmd[5], cmd[6], cmd[7], cmd[8]
    g = np.remainder(gait_index + dt * freq, 1.0)
    feet = [g + phase + offset + bound, g + offset, g + bound, g + phase]
    warped = []
    for f in feet:
        r = np.remainder(f, 1.0)
        if r < dur:
            f = r * (0.5 / dur)
        elif r > dur:
            f = 0.5 + (r - dur) * (0.5 / (1.0 - dur))
        warped.append(f)
    clock = np.array([np.sin(2 * np.pi * f) for f in warped])
    return g, clock


def terms(state37, cmd, clock, actions, gait_index, q_default, evaluating=True):
    """Scaled observation terms (dict name -> vector)."""
    q = state37[3:7]
    ang = quat_rotate_inverse(q, state37[10:13])
    pg = quat_rotate_inverse(q, np.array([0.0, 0.0, -1.0]))
    if evaluating:   # go2_omni.py:630-642
        h = np.clip(cmd[3] + 0.20 * np.sin(2 * np.pi * gait_index), HEIGHT_LIMITS[0], HEIGHT_LIMITS[1])
    else:
        h = cmd[3]
    raw = {"actions": actions, "base_ang_vel": ang, "clock_inputs": clock, "command_ang_vel": cmd[2:3],
           "command_body_attitude": cmd[10:12], "command_body_height": np.array([h]),
           "command_footswing_height": cmd[9:10], "command_gait_freq": cmd[4:5], "command_gait_phase": cmd[5:9],
           "command_lin_vel": cmd[0:2], "command_stance": cmd[12:14], "dof_pos": state37[13:25] - q_default,
           "dof_vel": state37[25:37], "projected_gravity": pg}
    return {k: np.asarray(raw[k], np.float64) * TERMS[k][1] for k in TERMS}


def assemble(t, history):
    """obs = [sorted current terms | short history per sorted key (newest first)], clipped;
    history: dict name -> [HISTORY, dim] (newest first) BEFORE this step's add.  Returns
    (obs, new history with the current terms added)."""
    keys = sorted(TERMS)
    cur = np.concatenate([t[k] for k in keys])
    hist = np.concatenate([history[k].reshape(-1) for k in keys])
    obs = np.clip(np.concatenate([cur, hist]), -CLIP, CLIP)
    new_hist = {k: np.concatenate([t[k][None], history[k][:-1]], axis=0) for k in keys}
    return obs, new_hist


def empty_history():
    return {k: np.zeros((HISTORY, d)) for k, (d, _) in TERMS.items()}


def obs_dim():
    return sum(d for d, _ in TERMS.values()) * (1 + HISTORY)


def mlp(layers, x):
    """Actor MLP in float64: Linear, ELU between layers (modules.py:47-63)."""
    h = np.asarray(x, np.float64)
    for i, (W, b) in enumerate(layers):
        h = h @ np.asarray(W, np.float64).T + np.asarray(b, np.float64)
        if i < len(layers) - 1:
            h = np.where(h > 0, h, np.expm1(np.minimum(h, 0.0)))
    return h


def _pg_xy(state37):
    return quat_rotate_inverse(state37[3:7], np.array([0.0, 0.0, -1.0]))[:2]


def policy_rollout(o, par_rows, init, cmds, layers, group, delta, ksync=5, ksync_mode="reference",
                   termination_rew=0.0, grav=(0.8, 0.8), action_clip=20.0, q_default=None, dt_ctrl=0.02,
                   decimation=4, evaluating=True, n_steps=None):
    """FP64 restatement of one evaluate_policy rollout (active_sysid.py:527-600 over
    active_sysid_openloop.py:116-426 and base_task.py:90-100): a reset step with zero actions,
    then n_steps - 2 policy steps; per step the actor acts on the observation (actions of
    terminated envs zeroed), the physics runs decimation physics steps (the first torque of an
    aux env after a sync sees its main env's joint state in the reference sync mode), gravity
    termination is OR-ed over each group, the reward is ||X||_F^2 (or termination_rew when
    terminated), then the observation and history are updated.  o: spio.Oracle.
    Returns (per-group average reward [M], final states [n_env][37], done [n_env])."""
    par_rows = np.asarray(par_rows, np.float64)
    n = par_rows.shape[0]
    M, T = cmds.shape[0], cmds.shape[1] if n_steps is None else n_steps
    assert n == M * group
    root = [np.asarray(init[i][:13], np.float64).copy() for i in range(n)]
    q = [np.asarray(init[i][13:25], np.float64).copy() for i in range(n)]
    qd = [np.asarray(init[i][25:37], np.float64).copy() for i in range(n)]
    gait = np.zeros(n)
    hist = [empty_history() for _ in range(n)]
    applied = np.zeros((n, 12))
    done = np.zeros(n, bool)
    acc = np.zeros(M)
    qdef = np.asarray(q_default, np.float64)

    def state37(i):
        return np.concatenate([root[i], q[i], qd[i]])

    def physics(actions, view):
        prev = [(q[i].copy(), qd[i].copy()) for i in range(n)]
        for i in range(n):
            a = np.clip(np.where(done[i], 0.0, actions[i]), -action_clip, action_clip)
            applied[i] = a
            m = i - i % group
            for d in range(decimation):
                if d == 0 and view and i != m:
                    tau = o.control_torque(a, prev[m][0], prev[m][1], par_rows[i])
                else:
                    tau = o.control_torque(a, q[i], qd[i], par_rows[i])
                root[i], q[i], qd[i] = o.physics_step(root[i], q[i], qd[i], tau, par_rows[i])

    def observe(now, nxt):
        obs = np.zeros((n, obs_dim()))
        for i in range(n):
            g = i // group
            gait[i], clock = gait_update(gait[i], cmds[g, now], dt_ctrl)
            t = terms(state37(i), cmds[g, nxt], clock, applied[i], gait[i], qdef, evaluating)
            obs[i], hist[i] = assemble(t, hist[i])
        return obs

    def sync(step_idx):
        if not (ksync <= 1 or step_idx % ksync == 1):
            return False
        if ksync_mode == "state":
            for i in range(n):
                m = i - i % group
                if i != m:
                    root[i], q[i], qd[i] = root[m].copy(), q[m].copy(), qd[m].copy()
            return False
        return True

    physics(np.zeros((n, 12)), False)
    obs = observe(0, 1)
    pending = sync(1)
    for it in range(1, T - 1):
        actions = mlp(layers, obs)
        physics(actions, pending)
        for g in range(M):
            mem = range(g * group, (g + 1) * group)
            term = any(abs(_pg_xy(state37(i))[0]) > grav[0] or abs(_pg_xy(state37(i))[1]) > grav[1] for i in mem)
            for i in mem:
                done[i] = term
            if term:
                acc[g] += termination_rew
            else:
                xm = np.concatenate([root[g * group], q[g * group]])
                for k in range(1, group):
                    X = (xm - np.concatenate([root[g * group + k], q[g * group + k]])) / delta
                    acc[g] += float(X @ X)
        obs = observe(it, it + 1)
        pending = sync(it + 1)
    return acc / (T - 1), np.stack([state37(i) for i in range(n)]), done
