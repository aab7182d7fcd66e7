"""Active-exploration command sequences generated on the device.

The reference turns each Optuna trial into a ``[total_steps, 14]`` command sequence with
Python loops (spigym/agents/sysid/active_sysid.py:259-400): constant windows, tanh-squashed
polynomials or Bezier curves per update window, the other dimensions at ``default_command``.
``CommandGenerator`` does the same for a whole batch of trials in one kernel
(``spi_command_sequences``); trial parameters are laid out in the reference's suggestion
order, so ``param_names()`` are the reference's trial keys and ``from_params`` rebuilds
``best_commands`` from a best-params dict (active_sysid.py:412-470).
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Mapping, Sequence

import numpy as np
import torch

from .. import _lib

MODES = {"constant": 0, "polynomial": 1, "bezier": 2}
MAX_DIMS = 16


class SpiCommandSpec(C.Structure):
    _fields_ = [("mode", C.c_int), ("n_dims", C.c_int), ("n_sampled", C.c_int),
                ("sampled_idx", C.c_int * MAX_DIMS), ("default_cmd", C.c_float * MAX_DIMS),
                ("lo", C.c_float * MAX_DIMS), ("hi", C.c_float * MAX_DIMS), ("n_updates", C.c_int),
                ("steps_per_update", C.c_int), ("poly_degree", C.c_int), ("n_bezier", C.c_int)]


class CommandGenerator:
    def __init__(self, command_ranges: Sequence[Sequence[float]], default_command: Sequence[float],
                 command_sampling_idxs: Sequence[int], command_sampling_mode: str = "constant",
                 rollout_length: float = 25.0, horizon_length: float = 5.0, dt: float = 0.02,
                 poly_degree: int = 3, num_bezier_points: int = 4, device="cuda"):
        """Arguments mirror algo.config.command (algo/active_sysid.yaml); dt is the env's control
        step (active_sysid.py:118-119: updates = rollout // horizon, steps = int(horizon / dt))."""
        if command_sampling_mode not in MODES:
            raise ValueError(f"unknown command_sampling_mode {command_sampling_mode!r}")
        ranges = np.asarray(command_ranges, np.float32)
        n = len(default_command)
        if n > MAX_DIMS or ranges.shape != (n, 2):
            raise ValueError("command_ranges must be [n_dims, 2] with n_dims <= 16")
        self.mode = command_sampling_mode
        self.sampled = [int(i) for i in command_sampling_idxs]
        self.n_updates = int(rollout_length // horizon_length)
        self.steps = int(horizon_length / dt)
        self.poly_degree, self.n_bezier = int(poly_degree), int(num_bezier_points)
        self.n_dims = n
        self.device = torch.device(device)
        s = SpiCommandSpec()
        s.mode, s.n_dims, s.n_sampled = MODES[self.mode], n, len(self.sampled)
        for m, k in enumerate(self.sampled):
            s.sampled_idx[m] = k
        for k in range(n):
            s.default_cmd[k] = float(default_command[k])
            s.lo[k], s.hi[k] = float(ranges[k, 0]), float(ranges[k, 1])
        s.n_updates, s.steps_per_update = self.n_updates, self.steps
        s.poly_degree, s.n_bezier = self.poly_degree, self.n_bezier
        self.spec = s
        self.ranges = ranges

    @property
    def total_steps(self) -> int:
        return self.n_updates * self.steps

    def param_names(self):
        """Trial keys in the reference's suggestion order."""
        U = range(self.n_updates)
        if self.mode == "constant":
            return [f"dim_{a}_update_{u}" for a in self.sampled for u in U]
        if self.mode == "polynomial":
            return [f"dim_{a}_update_{u}_coeff_{i}" for u in U for a in self.sampled for i in range(self.poly_degree + 1)]
        return [f"dim_{a}_update_{u}_bezier_{j}" for u in U for a in self.sampled for j in range(self.n_bezier)]

    def param_bounds(self) -> np.ndarray:
        """[n_params, 2] suggestion ranges: the command range (constant, bezier) or [-2, 2]
        for polynomial coefficients (active_sysid.py:323)."""
        if self.mode == "polynomial":
            return np.tile(np.array([[-2.0, 2.0]], np.float32), (len(self.param_names()), 1))
        out = []
        for name in self.param_names():
            a = int(name.split("_")[1])
            out.append(self.ranges[a])
        return np.asarray(out, np.float32)

    def generate(self, params: torch.Tensor) -> torch.Tensor:
        """params [B, n_params] (device) -> commands [B, total_steps, n_dims] (device)."""
        params = params.to(self.device, torch.float32).contiguous()
        B, P = params.shape
        out = torch.empty((B, self.total_steps, self.n_dims), dtype=torch.float32, device=self.device)
        rc = _lib.lib().spi_command_sequences(C.byref(self.spec), _lib.ptr(params), B, P, _lib.ptr(out),
                                              _lib.stream())
        _lib.check(rc, "spi_command_sequences")
        return out

    def from_params(self, best_params: Mapping[str, float]) -> np.ndarray:
        """[total_steps, n_dims] from a best-params dict keyed like the reference's trials."""
        p = torch.tensor([[float(best_params[k]) for k in self.param_names()]], dtype=torch.float32)
        return self.generate(p)[0].cpu().numpy()
