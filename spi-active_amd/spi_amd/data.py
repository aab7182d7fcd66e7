"""Trajectory segments: the synthetic ground-truth generator and the npz dataset loader.

A segment is what one rollout replays: the recorded state at a window start, the
``H`` recorded actions of the window, and the recorded state after each of them.
Packed layout (what the fused kernel stages in LDS), floats per segment:
``37 + H * (12 + 37)`` = init state, then per control step (action, target state).
State rows are 37 floats: root (pos, quat xyzw, linvel, angvel; world) + q(12) + qd(12).

The npz schema is the reference's (scripts/data/common.py:83-101); the windowing and
the ``motion_ends`` bookkeeping follow scripts/eval.py:101-171.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from .params import GO2_DEFAULT_JOINT_ANGLES, SimParams
from .model import DOF_NAMES

STATE_DIM = 37


def segment_stride(horizon: int) -> int:
    return STATE_DIM + horizon * (12 + STATE_DIM)


@dataclass
class SegmentSet:
    init: np.ndarray          # [n_seg, 37]
    actions: np.ndarray       # [n_seg, H, 12]
    targets: np.ndarray       # [n_seg, H, 37]
    motion_ends: np.ndarray   # [n_seg] bool (scripts/eval.py:145-146)
    kp: Optional[np.ndarray] = None   # per-file PD gains (scripts/eval.py:148-152)
    kd: Optional[np.ndarray] = None

    @property
    def n_seg(self) -> int:
        return self.init.shape[0]

    @property
    def horizon(self) -> int:
        return self.actions.shape[1]

    def packed(self) -> np.ndarray:
        n, H = self.n_seg, self.horizon
        out = np.empty((n, segment_stride(H)), dtype=np.float32)
        out[:, :STATE_DIM] = self.init
        rec = np.concatenate([self.actions, self.targets], axis=2).reshape(n, H * (12 + STATE_DIM))
        out[:, STATE_DIM:] = rec
        return out

    @staticmethod
    def from_packed(packed: np.ndarray, horizon: int) -> "SegmentSet":
        """Inverse of ``packed()`` (no motion ends, no recorded gains)."""
        packed = np.asarray(packed, np.float32)
        n = packed.shape[0]
        if packed.ndim != 2 or packed.shape[1] != segment_stride(horizon):
            raise ValueError("packed segments do not match the horizon")
        rec = packed[:, STATE_DIM:].reshape(n, horizon, 12 + STATE_DIM)
        return SegmentSet(packed[:, :STATE_DIM].copy(), rec[:, :, :12].copy(), rec[:, :, 12:].copy(),
                          np.zeros(n, dtype=bool))

    def subset(self, idx) -> "SegmentSet":
        return SegmentSet(self.init[idx], self.actions[idx], self.targets[idx], self.motion_ends[idx],
                          None if self.kp is None else self.kp[idx], None if self.kd is None else self.kd[idx])


def standing_state(height: float = 0.30) -> np.ndarray:
    s = np.zeros(STATE_DIM, dtype=np.float32)
    s[2] = height
    s[6] = 1.0
    s[13:25] = [GO2_DEFAULT_JOINT_ANGLES[n] for n in DOF_NAMES]
    return s


def sinusoid_actions(rng: np.random.Generator, n: int, T: int, dt_ctrl: float, action_scale: float,
                     amp=(0.1, 0.4), freq=(0.5, 3.0)) -> np.ndarray:
    """Per-joint sinusoidal PD-target offsets A sin(2 pi f t + phi), as raw actions.

    The PD target is q0 + action_scale * action (legged_robot_base.py:542-545), so the
    action is the offset divided by action_scale.  Amplitude U[amp], frequency U[freq],
    phase U[0, 2 pi), drawn per trajectory and joint.
    """
    A = rng.uniform(amp[0], amp[1], size=(n, 1, 12))
    f = rng.uniform(freq[0], freq[1], size=(n, 1, 12))
    ph = rng.uniform(0.0, 2.0 * np.pi, size=(n, 1, 12))
    t = (np.arange(T) * dt_ctrl).reshape(1, T, 1)
    return (A * np.sin(2.0 * np.pi * f * t + ph) / action_scale).astype(np.float32)


SimulateFn = Callable[[np.ndarray, np.ndarray, np.ndarray], np.ndarray]
"""(param rows [n, 25], init states [n, 37], actions [n, T, 12]) -> states after each step [n, T, 37]."""


def synthetic_segments(simulate: SimulateFn, theta: np.ndarray, sim: SimParams, n_seg: int, horizon: int,
                       seed: int, warmup: Tuple[int, int] = (10, 40), settle_steps: int = 50,
                       sigma_q: float = 0.01, sigma_v: float = 0.05, amp=(0.1, 0.4), freq=(0.5, 3.0),
                       start_height: float = 0.30) -> SegmentSet:
    """Segments cut from ground-truth rollouts under ``theta`` with sinusoidal PD targets.

    Each segment has its own trajectory: the robot first settles from a standing pose
    (zero action), then follows its sinusoids for ``w ~ U[warmup]`` control steps before
    the recorded window of ``horizon`` steps starts.  Gaussian measurement noise
    (sigma_q on joint positions, sigma_v on joint velocities) is added to every recorded
    state, window starts included.
    """
    rng = np.random.default_rng(seed)
    dt_ctrl = sim.dt * sim.decimation
    theta = np.asarray(theta, dtype=np.float64).reshape(1, -1)
    settled = simulate(theta, standing_state(start_height)[None], np.zeros((1, settle_steps, 12), np.float32))
    s0 = settled[0, -1]
    w_max = warmup[1]
    T = w_max + horizon
    acts = sinusoid_actions(rng, n_seg, T, dt_ctrl, sim.action_scale, amp, freq)
    states = simulate(np.repeat(theta, n_seg, axis=0), np.repeat(s0[None], n_seg, axis=0), acts)
    starts = rng.integers(warmup[0], warmup[1] + 1, size=n_seg)
    init = np.empty((n_seg, STATE_DIM), np.float32)
    actions = np.empty((n_seg, horizon, 12), np.float32)
    targets = np.empty((n_seg, horizon, STATE_DIM), np.float32)
    for i, w in enumerate(starts):
        init[i] = states[i, w - 1]
        actions[i] = acts[i, w:w + horizon]
        targets[i] = states[i, w:w + horizon]
    if sigma_q > 0 or sigma_v > 0:
        init[:, 13:25] += rng.normal(0.0, sigma_q, size=init[:, 13:25].shape)
        init[:, 25:37] += rng.normal(0.0, sigma_v, size=init[:, 25:37].shape)
        targets[:, :, 13:25] += rng.normal(0.0, sigma_q, size=targets[:, :, 13:25].shape)
        targets[:, :, 25:37] += rng.normal(0.0, sigma_v, size=targets[:, :, 25:37].shape)
    return SegmentSet(init, actions, targets, np.zeros(n_seg, dtype=bool))


# ---------------------------------------------------------------- npz datasets (reference schema)
NPZ_KEYS = ("base_positions", "base_orientations", "base_linear_velocities", "base_angular_velocities",
            "joint_positions", "joint_velocities", "actions", "pd_gain_kp", "pd_gain_kd")


def load_npz_dataset(paths: Sequence[str], horizon: int) -> SegmentSet:
    """Sliding windows (stride 1) over every recorded trajectory, as scripts/eval.py:101-171.

    For window start t: init = state at t, actions t..t+H-1, targets = states t+1..t+H
    (eval.py compares only the final one, ``target_idx = init_idx + horizon``).  The last
    window of each file is flagged in ``motion_ends``.
    """
    inits, acts, tgts, ends, kps, kds = [], [], [], [], [], []
    for p in paths:
        with np.load(p, allow_pickle=False) as d:
            states = np.concatenate([d["base_positions"], d["base_orientations"], d["base_linear_velocities"],
                                     d["base_angular_velocities"], d["joint_positions"], d["joint_velocities"]],
                                    axis=1).astype(np.float32)
            a = d["actions"].astype(np.float32)
            kp, kd = float(np.asarray(d["pd_gain_kp"]).flat[0]), float(np.asarray(d["pd_gain_kd"]).flat[0])
        T = states.shape[0]
        n = T - horizon
        if n <= 0:
            continue
        idx = np.arange(n)
        inits.append(states[idx])
        acts.append(np.stack([a[idx + k] for k in range(horizon)], axis=1))
        tgts.append(np.stack([states[idx + k + 1] for k in range(horizon)], axis=1))
        e = np.zeros(n, dtype=bool)
        e[-1] = True
        ends.append(e)
        kps.append(np.full((n, 12), kp, np.float32))
        kds.append(np.full((n, 12), kd, np.float32))
    if not inits:
        raise ValueError("no trajectory is longer than the horizon")
    return SegmentSet(np.concatenate(inits), np.concatenate(acts), np.concatenate(tgts), np.concatenate(ends),
                      np.concatenate(kps), np.concatenate(kds))


def reference_segment_weights(motion_ends: np.ndarray, batch_size: int) -> np.ndarray:
    """Per-segment weights reproducing scripts/eval.py:239-310 exactly.

    Inside each chunk of ``batch_size`` segments, ``eval_mask = ~(cumsum(motion_ends) > 0)``
    drops the first motion end of the chunk and everything after it; the masked sums are
    divided by ``total_valid = (~motion_ends).sum()`` over the whole dataset.
    """
    n = motion_ends.shape[0]
    w = np.zeros(n, dtype=np.float64)
    total_valid = float((~motion_ends).sum())
    for start in range(0, n, batch_size):
        m = motion_ends[start:start + batch_size]
        mask = ~(np.cumsum(m.astype(np.float64)) > 0)
        w[start:start + batch_size] = mask / total_valid
    return w.astype(np.float32)


def segment_gains(kp: np.ndarray, kd: np.ndarray, batch_size: Optional[int] = None) -> np.ndarray:
    """Recorded PD gains per segment, [n, 2] float32 (kp, kd), for ``SpiCostArgs.seg_gains``.

    ``kp``/``kd`` are per-segment (``[n]``) or per-segment-and-joint (``[n, 12]``, all joints
    equal, as load_npz_dataset fills them; column 0 is used).  Without ``batch_size`` every
    segment keeps the gains of its own file.  With ``batch_size`` the reference's chunking is
    reproduced: scripts/eval.py:245-250 evaluates the windows in chunks of ``num_envs`` and
    apply_pd_gains_batch (:192-201) sets the env's gains from the chunk's FIRST window, so a
    chunk that straddles two files runs all of its windows with the first file's gains.
    """
    kp = np.asarray(kp, np.float32)
    kd = np.asarray(kd, np.float32)
    kp = kp[:, 0] if kp.ndim == 2 else kp
    kd = kd[:, 0] if kd.ndim == 2 else kd
    if kp.shape != kd.shape or kp.ndim != 1:
        raise ValueError("kp and kd must have one entry (or one row) per segment")
    g = np.stack([kp, kd], axis=1)
    if batch_size is not None:
        if batch_size < 1:
            raise ValueError("batch_size must be >= 1")
        first = (np.arange(g.shape[0]) // batch_size) * batch_size
        g = g[first]
    return np.ascontiguousarray(g, np.float32)


# ---------------------------------------------------------------- the reference's data recipes
# scripts/data/sine.py:15-21, jump.py:15-21, stand.py:13-14: 5 s at the control rate, thigh
# joints (1, 4, 7, 10) driven with raw action A sin(2 pi f t), calf joints (2, 5, 8, 11) with
# its negative, hips 0; the sine and jump runs start with the base raised to 0.32 m.
RECIPES = {"sine": (0.8, 1.0, 0.32), "jump": (1.2, 1.5, 0.32), "stand": (0.0, 0.0, None)}
THIGH_IDX = (1, 4, 7, 10)
CALF_IDX = (2, 5, 8, 11)


def recipe_actions(name: str, dt_ctrl: float, duration: float = 5.0) -> np.ndarray:
    """[steps, 12] raw actions of a reference recipe: steps = round(duration / dt_ctrl), action at
    step k uses t = k dt_ctrl (sine.py:35-46)."""
    if name not in RECIPES:
        raise ValueError(f"unknown recipe {name!r}; one of {sorted(RECIPES)}")
    amp, freq, _ = RECIPES[name]
    steps = int(round(duration / dt_ctrl))
    phase = amp * np.sin(2.0 * np.pi * freq * np.arange(steps) * dt_ctrl)
    a = np.zeros((steps, 12), np.float32)
    a[:, list(THIGH_IDX)] = phase[:, None]
    a[:, list(CALF_IDX)] = -phase[:, None]
    return a


def record_recipe(simulate: SimulateFn, theta: np.ndarray, sim: SimParams, name: str, init: np.ndarray,
                  duration: float = 5.0) -> Tuple[np.ndarray, np.ndarray]:
    """A recipe run recorded the way the reference's collectors do (sine.py:37-70, common.py):
    ``init`` is the state after the env's reset (scripts/data/common.py create_env: reset_all, then
    for sine and jump the base height set to 0.32 m), every step applies the recipe's action and
    frame k holds the state AFTER step k beside action k.  Returns (states [T, 37], actions
    [T, 12]) for ``save_npz_trajectory``.  (scripts/eval.py:101-171 then replays window t from the
    state after step t with actions t .. t + H - 1: the reference's own pairing, kept as is.)"""
    acts = recipe_actions(name, sim.dt * sim.decimation, duration)
    s0 = np.asarray(init, np.float32).copy()
    height = RECIPES[name][2]
    if height is not None:
        s0[2] = height
    theta = np.asarray(theta, np.float64).reshape(1, -1)
    states = simulate(theta, s0[None], acts[None])[0]
    return states.astype(np.float32), acts


def save_npz_trajectory(path: str, states: np.ndarray, actions: np.ndarray, dt: float, kp: float, kd: float) -> None:
    """Write one trajectory in the reference's npz schema (scripts/data/common.py:83-101)."""
    T = states.shape[0]
    np.savez(path,
             timestamps=np.arange(T) * dt,
             base_positions=states[:, 0:3], base_orientations=states[:, 3:7],
             base_linear_velocities=states[:, 7:10], base_angular_velocities=states[:, 10:13],
             joint_positions=states[:, 13:25], joint_velocities=states[:, 25:37],
             joint_torques=np.zeros((T, 12), np.float32), actions=actions,
             sim_duration=T * dt, data_frequency=int(round(1.0 / dt)), robot_type="go2",
             pd_gain_kp=np.full(12, kp, np.float32), pd_gain_kd=np.full(12, kd, np.float32))
