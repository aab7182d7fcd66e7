"""System-identification problems over the fused rollout engine.

``SysIdProblem`` is the batched objective: candidate parameter vectors in, per-candidate
tracking cost out (mean over segments), one fused kernel + one fixed-order reduction.
``identification_prior`` gives the 22-parameter search space: Gaussian centred at the
nominal URDF values with a 20 % width (absolute widths where the nominal is ~0).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np
import torch

from ..params import P_COUNT, P_COMX, P_FRIC0, P_IXX, P_IXY, P_KD0, P_KP0, P_MASS, P_MOTOR0, nominal_params
from ..rollout import RolloutEngine


@dataclass
class Prior:
    mean: np.ndarray
    std: np.ndarray
    lo: np.ndarray
    hi: np.ndarray
    min_std: np.ndarray
    free: np.ndarray        # bool mask of identified parameters


def identification_prior(nominal: Optional[np.ndarray] = None, rel: float = 0.2, n_free: int = 22) -> Prior:
    m = nominal_params() if nominal is None else np.asarray(nominal, np.float64)
    std = rel * np.abs(m)
    std[P_COMX:P_COMX + 3] = 0.02                      # CoM: 2 cm
    std[P_IXY:P_IXY + 3] = 0.002                       # products of inertia: 2e-3 kg m^2
    std[P_FRIC0:P_FRIC0 + 4] = 0.05                    # joint friction: 0.05 N m
    lo, hi = m - 4 * std, m + 4 * std
    lo[P_MASS] = max(lo[P_MASS], 0.2 * m[P_MASS])
    lo[P_IXX:P_IXX + 3] = np.maximum(lo[P_IXX:P_IXX + 3], 0.2 * m[P_IXX:P_IXX + 3])
    lo[P_KP0:P_KD0 + 4] = np.maximum(lo[P_KP0:P_KD0 + 4], 0.2)
    lo[P_FRIC0:P_FRIC0 + 4] = 0.0
    free = np.zeros(P_COUNT, bool)
    free[:n_free] = True
    std[~free] = 0.0
    min_std = np.where(free, 1e-4 * np.maximum(np.abs(m), 1e-2), 0.0)
    return Prior(m, std, lo, hi, min_std, free)


def ground_truth(prior: Prior, seed: int, spread: float = 0.5) -> np.ndarray:
    """A seeded 'true robot': nominal + U(-spread, spread) * prior width on the free entries."""
    rng = np.random.default_rng(seed)
    th = prior.mean.copy()
    u = rng.uniform(-spread, spread, size=th.shape)
    th[prior.free] += (u * prior.std)[prior.free]
    return np.clip(th, prior.lo, prior.hi)


class SysIdProblem:
    """Batched objective: SoA candidates [25, n] -> mean tracking cost over the segments [n]."""

    def __init__(self, engine: RolloutEngine, segments: torch.Tensor, horizon: int,
                 seg_weights: Optional[torch.Tensor] = None, max_cand: int = 8192,
                 seg_gains: Optional[torch.Tensor] = None):
        self.engine, self.segments, self.horizon = engine, segments, horizon
        self.seg_gains = seg_gains     # recorded (kp, kd) per segment, or None: configured gains
        n_seg = segments.shape[0]
        dev = engine.device
        self.seg_weights = seg_weights if seg_weights is not None else \
            torch.full((n_seg,), 1.0 / n_seg, dtype=torch.float32, device=dev)
        self._seg_cost = torch.empty((n_seg, max_cand), dtype=torch.float32, device=dev)
        self._cost = torch.empty(max_cand, dtype=torch.float32, device=dev)

    def __call__(self, theta: torch.Tensor, n: int) -> torch.Tensor:
        sc = self._seg_cost[:, :n] if n == self._seg_cost.shape[1] else \
            self._seg_cost.view(-1)[: self.segments.shape[0] * n].view(self.segments.shape[0], n)
        self.engine.segment_costs(theta, n, self.segments, self.horizon, out=sc, seg_gains=self.seg_gains)
        return self.engine.reduce(sc, self.seg_weights, out=self._cost[:n])


# ---------------------------------------------------------------- plain method: 1-D landscape
# scripts/mass_landscape.py:32-36
COST_COEFF = {"base_pos": 10.0, "base_quat": 5.0, "joint_pos": 1.0}


def mass_landscape(engine: RolloutEngine, segments: torch.Tensor, horizon: int, base_params: np.ndarray,
                   masses: np.ndarray, seg_weights: Optional[torch.Tensor] = None,
                   seg_gains: Optional[torch.Tensor] = None) -> Tuple[np.ndarray, int]:
    """Cost of every base mass on a grid, all candidates in one fused launch.

    Counterpart of ``mass_sweep`` + argmin (scripts/mass_landscape.py:111-171), which
    evaluates the candidates one after another through ``scripts/eval.py:evaluate_batch``.
    With a reference-mode engine (``cost_mode='reference'``, weights ``COST_COEFF``) and
    ``data.reference_segment_weights`` the per-candidate value is the script's
    ``10 pos + 5 quat + 1 joint`` total.  Returns (costs [n], argmin index).
    """
    rows = np.repeat(np.asarray(base_params, np.float64)[None], len(masses), axis=0)
    rows[:, P_MASS] = masses
    theta = torch.from_numpy(np.ascontiguousarray(rows.T.astype(np.float32))).to(engine.device)
    prob = SysIdProblem(engine, segments, horizon, seg_weights, max_cand=len(masses), seg_gains=seg_gains)
    c = prob(theta, len(masses)).cpu().numpy()
    return c, int(np.argmin(c))
