"""Device-resident dataset store for recorded trajectories (the reference's npz schema).

``scripts/eval.py:101-182`` loads every npz, materialises one window per start index on the
host (init at t, actions t..t+H-1, target at t+H) and copies the arrays to the device.  The
store keeps each trajectory once in HBM (``[T, 37]`` states in the IsaacGym root layout +
``[T, 12]`` actions) and packs the windows for any horizon on the device
(``spi_pack_windows_rows``: every window of every chosen trajectory in one launch, from a
back-to-back copy of the trajectories made on the first call after a trajectory was added),
so the host never holds the H-fold expanded windows.  Trajectories
come from npz files (scripts/data/common.py:83-101 keys), real-robot logs in the same schema,
or arrays.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .data import reference_segment_weights, segment_stride

NPZ_STATE_KEYS = ("base_positions", "base_orientations", "base_linear_velocities", "base_angular_velocities",
                  "joint_positions", "joint_velocities")


@dataclass
class Trajectory:
    name: str
    states: torch.Tensor          # [T, 37] device
    actions: torch.Tensor         # [T, 12] device
    kp: float
    kd: float

    @property
    def n_steps(self) -> int:
        return self.states.shape[0]


@dataclass
class SegmentBatch:
    segments: torch.Tensor        # [n, stride(H)] device, the rollout kernels' layout
    horizon: int
    motion_ends: np.ndarray       # [n] bool, last window of each trajectory
    source: np.ndarray            # [n] index of the trajectory
    start: np.ndarray             # [n] window start step
    kp: np.ndarray                # [n] per-window PD gain of its file (eval.py:148-152)
    kd: np.ndarray

    def reference_weights(self, batch_size: int) -> torch.Tensor:
        w = reference_segment_weights(self.motion_ends, batch_size)
        return torch.from_numpy(w).to(self.segments.device)


class SegmentStore:
    def __init__(self, device="cuda"):
        _lib.require_gpu()
        self.device = torch.device(device)
        self.trajectories: List[Trajectory] = []
        self._joined = None           # (states [R, 37], actions [R, 12], first row of each trajectory)

    def _joined_buffers(self):
        """All trajectories back to back in one device buffer (rebuilt after an add)."""
        if self._joined is None or len(self._joined[2]) != len(self.trajectories):
            starts = np.cumsum([0] + [tr.n_steps for tr in self.trajectories])[:-1]
            self._joined = (torch.cat([tr.states for tr in self.trajectories]),
                            torch.cat([tr.actions for tr in self.trajectories]), starts)
        return self._joined

    def add_arrays(self, states: np.ndarray, actions: np.ndarray, kp: float = 20.0, kd: float = 0.5,
                   name: Optional[str] = None) -> int:
        states = np.ascontiguousarray(states, np.float32)
        actions = np.ascontiguousarray(actions, np.float32)
        if states.ndim != 2 or states.shape[1] != 37 or actions.shape != (states.shape[0], 12):
            raise ValueError("states [T, 37] and actions [T, 12] expected")
        self.trajectories.append(Trajectory(name or f"traj{len(self.trajectories)}",
                                            torch.from_numpy(states).to(self.device),
                                            torch.from_numpy(actions).to(self.device), float(kp), float(kd)))
        return len(self.trajectories) - 1

    def add_npz(self, path: str) -> int:
        with np.load(path, allow_pickle=False) as d:
            states = np.concatenate([np.asarray(d[k], np.float32) for k in NPZ_STATE_KEYS], axis=1)
            actions = np.asarray(d["actions"], np.float32)
            kp = float(np.asarray(d["pd_gain_kp"]).flat[0])
            kd = float(np.asarray(d["pd_gain_kd"]).flat[0])
        return self.add_arrays(states, actions, kp, kd, name=path)

    def windows(self, horizon: int, which: Optional[Sequence[int]] = None) -> SegmentBatch:
        """Every stride-1 window of the chosen trajectories, in order (eval.py:125-146)."""
        which = list(range(len(self.trajectories))) if which is None else list(which)
        counts = [max(self.trajectories[i].n_steps - horizon, 0) for i in which]
        n = int(sum(counts))
        if n == 0:
            raise ValueError("no trajectory is longer than the horizon")
        out = torch.empty((n, segment_stride(horizon)), dtype=torch.float32, device=self.device)
        ends = np.zeros(n, bool)
        src, start, kp, kd = (np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float32),
                              np.zeros(n, np.float32))
        states, actions, first = self._joined_buffers()
        if states.shape[0] >= 2 ** 31:
            raise ValueError("the store holds more than 2^31 - 1 trajectory rows")
        rows = np.zeros(n, np.int32)
        o = 0
        for i, c in zip(which, counts):
            if c == 0:
                continue
            tr = self.trajectories[i]
            ends[o + c - 1] = True
            src[o:o + c], start[o:o + c] = i, np.arange(c)
            rows[o:o + c] = first[i] + np.arange(c)
            kp[o:o + c], kd[o:o + c] = tr.kp, tr.kd
            o += c
        rows_dev = torch.from_numpy(rows).to(self.device)
        rc = _lib.lib().spi_pack_windows_rows(_lib.ptr(states), _lib.ptr(actions), states.shape[0],
                                              _lib.ptr(rows_dev, torch.int32), n, horizon, _lib.ptr(out),
                                              _lib.stream())
        _lib.check(rc, "spi_pack_windows_rows")
        return SegmentBatch(out, horizon, ends, src, start, kp, kd)
