"""Device-side rollout API over the C ABI: fused rollout + cost, record mode, single steps.

Tensors are torch CUDA (ROCm) tensors; parameters are SoA ``[25, ld]`` float32.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .data import SegmentSet, segment_stride
from .params import COST_MODES, P_COUNT, SimParams

# Tracking-cost weights: base pos, quat (raw xyzw), linvel, angvel, joint q, joint qd.
DEFAULT_TRACKING_WEIGHTS = (10.0, 5.0, 1.0, 1.0, 1.0, 0.05)
# scripts/mass_landscape.py:32-36 (reference mode uses w[0], w[1], w[4])
REFERENCE_WEIGHTS = (10.0, 5.0, 0.0, 0.0, 1.0, 0.0)


# Up to this many (candidate, segment) rollouts the four-lanes-per-rollout kernel is faster:
# one rollout per lane leaves SIMDs idle below ~64 K rollouts and has ~3x the instructions per
# lane on the critical path.
LEG_PARALLEL_MAX_ROLLOUTS = 16384


class RolloutEngine:
    """Fused batched rollouts of the Go2 under per-candidate parameters."""

    def __init__(self, sim: SimParams, cost_mode: str = "tracking", weights: Optional[Sequence[float]] = None,
                 device: str = "cuda", state_limit: float = 1.0e4):
        _lib.require_gpu()
        self.sim = sim
        self.device = torch.device(device)
        self.args = _lib.sim_args(sim, state_limit)
        if weights is None:
            weights = DEFAULT_TRACKING_WEIGHTS if cost_mode == "tracking" else REFERENCE_WEIGHTS
        self.weights = [float(w) for w in weights]
        self.cost_mode = COST_MODES[cost_mode]
        self.cargs = _lib.cost_args(self.cost_mode, weights)
        self.L = _lib.lib()

    # ---------------------------------------------------------------- segments
    def upload_segments(self, segs: SegmentSet) -> torch.Tensor:
        return torch.from_numpy(segs.packed()).to(self.device)

    def cost_args_for(self, n_seg: int, seg_gains: Optional[torch.Tensor] = None):
        """The engine's cost settings, with the recorded PD gains of each segment when given
        (device [n_seg, 2] float32 (kp, kd), e.g. from ``data.segment_gains``)."""
        if seg_gains is None:
            return self.cargs
        if tuple(seg_gains.shape) != (n_seg, 2) or not seg_gains.is_cuda:
            raise ValueError(f"seg_gains must be a [{n_seg}, 2] device tensor")
        return _lib.cost_args(self.cost_mode, self.weights, seg_gains)

    def segment_costs(self, params: torch.Tensor, n_cand: int, segs: torch.Tensor, horizon: int,
                      out: Optional[torch.Tensor] = None, physics: str = "auto",
                      seg_gains: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Cost of every (segment, candidate) pair: [n_seg, n_cand].  physics: "lanes" (one
        rollout per lane), "legs" (four lanes per rollout: lower latency when the batch leaves
        SIMDs idle) or "auto" (legs up to LEG_PARALLEL_MAX_ROLLOUTS rollouts).  seg_gains:
        recorded (kp, kd) per segment in place of the configured PD gains."""
        n_seg = segs.shape[0]
        cargs = self.cost_args_for(n_seg, seg_gains)
        if segs.shape[1] != segment_stride(horizon):
            raise ValueError("segment tensor stride does not match the horizon")
        if params.dim() != 2 or params.shape[0] != P_COUNT or params.shape[1] < n_cand:
            raise ValueError("params must be SoA [25, >= n_cand]")
        if out is None:
            out = torch.empty((n_seg, n_cand), dtype=torch.float32, device=self.device)
        if physics not in ("auto", "lanes", "legs"):
            raise ValueError("physics must be 'auto', 'lanes' or 'legs'")
        legs = physics == "legs" or (physics == "auto" and n_cand * n_seg <= LEG_PARALLEL_MAX_ROLLOUTS)
        fn = self.L.spi_rollout_cost_legs if legs else self.L.spi_rollout_cost
        rc = fn(C_ref(self.args), C_ref(cargs), _lib.ptr(params), params.shape[1], n_cand, _lib.ptr(segs), n_seg,
                horizon, _lib.ptr(out), _lib.stream())
        _lib.check(rc, "spi_rollout_cost")
        return out

    def reduce(self, seg_costs: torch.Tensor, seg_weights: torch.Tensor, out: Optional[torch.Tensor] = None):
        n_seg, n_cand = seg_costs.shape
        if out is None:
            out = torch.empty(n_cand, dtype=torch.float32, device=self.device)
        rc = self.L.spi_reduce_segments(_lib.ptr(seg_costs), _lib.ptr(seg_weights), n_seg, n_cand, _lib.ptr(out),
                                        _lib.stream())
        _lib.check(rc, "spi_reduce_segments")
        return out

    def costs(self, params: torch.Tensor, n_cand: int, segs: torch.Tensor, horizon: int,
              seg_weights: Optional[torch.Tensor] = None, seg_gains: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Per-candidate cost (mean over segments unless weights are given)."""
        sc = self.segment_costs(params, n_cand, segs, horizon, seg_gains=seg_gains)
        if seg_weights is None:
            seg_weights = torch.full((segs.shape[0],), 1.0 / segs.shape[0], dtype=torch.float32, device=self.device)
        return self.reduce(sc, seg_weights)

    # ---------------------------------------------------------------- record mode
    def simulate(self, params: torch.Tensor, init: torch.Tensor, actions: torch.Tensor) -> torch.Tensor:
        n, T = actions.shape[0], actions.shape[1]
        if init.shape != (n, 37) or actions.shape[2] != 12 or params.shape[0] != P_COUNT or params.shape[1] < n:
            raise ValueError("simulate: shape mismatch")
        out = torch.empty((n, T, 37), dtype=torch.float32, device=self.device)
        rc = self.L.spi_simulate(C_ref(self.args), _lib.ptr(params), params.shape[1], n, _lib.ptr(init),
                                 _lib.ptr(actions), T, _lib.ptr(out), _lib.stream())
        _lib.check(rc, "spi_simulate")
        return out

    def simulate_np(self, param_rows: np.ndarray, init: np.ndarray, actions: np.ndarray) -> np.ndarray:
        """numpy convenience wrapper (the synthetic-data generator's SimulateFn)."""
        p = torch.from_numpy(np.ascontiguousarray(np.asarray(param_rows, np.float32).T)).to(self.device)
        i = torch.from_numpy(np.ascontiguousarray(init, np.float32)).to(self.device)
        a = torch.from_numpy(np.ascontiguousarray(actions, np.float32)).to(self.device)
        return self.simulate(p, i, a).cpu().numpy()

    # ---------------------------------------------------------------- single steps (BaseSimulator path)
    def physics_step(self, params: torch.Tensor, root: torch.Tensor, dof_state: torch.Tensor, torques: torch.Tensor,
                     foot_force: Optional[torch.Tensor] = None) -> None:
        n = root.shape[0]
        rc = self.L.spi_physics_step(C_ref(self.args), _lib.ptr(params), params.shape[1], n, _lib.ptr(root),
                                     _lib.ptr(dof_state), _lib.ptr(torques), _lib.ptr(foot_force), _lib.stream())
        _lib.check(rc, "spi_physics_step")

    def forward_dynamics(self, params: torch.Tensor, root: torch.Tensor, dof_state: torch.Tensor,
                         torques: torch.Tensor, foot_force: Optional[torch.Tensor] = None) -> torch.Tensor:
        """[n, 18]: base spatial acceleration (base coords, angular first) and qdd."""
        n = root.shape[0]
        out = torch.empty((n, 18), dtype=torch.float32, device=self.device)
        rc = self.L.spi_forward_dynamics(C_ref(self.args), _lib.ptr(params), params.shape[1], n, _lib.ptr(root),
                                         _lib.ptr(dof_state), _lib.ptr(torques), _lib.ptr(out), _lib.ptr(foot_force),
                                         _lib.stream())
        _lib.check(rc, "spi_forward_dynamics")
        return out

    def compute_torques(self, params: torch.Tensor, actions: torch.Tensor, dof_state: torch.Tensor,
                        out: Optional[torch.Tensor] = None) -> torch.Tensor:
        n = actions.shape[0]
        if out is None:
            out = torch.empty((n, 12), dtype=torch.float32, device=self.device)
        rc = self.L.spi_compute_torques(C_ref(self.args), _lib.ptr(params), params.shape[1], n, _lib.ptr(actions),
                                        _lib.ptr(dof_state), _lib.ptr(out), _lib.stream())
        _lib.check(rc, "spi_compute_torques")
        return out


def C_ref(struct):
    import ctypes
    return ctypes.byref(struct)
