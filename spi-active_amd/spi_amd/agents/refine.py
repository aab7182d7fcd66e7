"""Levenberg-Marquardt refinement of identified parameters from exact residual Jacobians.

CEM (agents/cem.py) finds the basin; the tracking cost is a sum of squared residuals, so a
few Gauss-Newton steps with exact derivatives of the discrete rollout converge to the local
minimiser that a sampling optimizer only approaches.  Every step runs on the device:
``spi_residual_jacobian`` (tangent rollouts of all segments, ``[J | r]`` per segment),
``spi_fim_assemble`` (``[J r]^T [J r]`` on matrix cores) and ``spi_gn_step`` (FP64 normal
equations, damped Cholesky solve); the host only accepts or rejects the step and adapts the
damping.  The reference refines with Optuna trials instead (scripts/mass_opt.py:203-219).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from .. import _lib
from ..data import segment_stride
from ..params import P_COUNT
from ..rollout import C_ref, RolloutEngine
from .fisher import IDENTIFIED, _rows


@dataclass
class RefineResult:
    theta: np.ndarray                  # [25]
    cost: float
    history: List[float] = field(default_factory=list)
    iterations: int = 0
    accepted: int = 0


class GaussNewtonRefiner:
    def __init__(self, engine: RolloutEngine, segments: torch.Tensor, horizon: int,
                 seg_weights: Optional[torch.Tensor] = None, names: Sequence[str] = IDENTIFIED,
                 lo: Optional[np.ndarray] = None, hi: Optional[np.ndarray] = None, legs: bool = True,
                 seg_gains: Optional[torch.Tensor] = None):
        if engine.cost_mode != 0:
            raise ValueError("Gauss-Newton refinement needs the tracking (sum of squares) cost")
        if segments.shape[1] != segment_stride(horizon):
            raise ValueError("segment stride does not match the horizon")
        self.engine, self.segments, self.H = engine, segments, horizon
        # residual Jacobian with four lanes per tangent (one per leg) or one lane per tangent
        self.jacobian_entry = "spi_residual_jacobian_legs" if legs else "spi_residual_jacobian"
        self.n_seg = segments.shape[0]
        self.seg_gains = seg_gains
        self.cargs = engine.cost_args_for(self.n_seg, seg_gains)
        dev = engine.device
        self.seg_w = seg_weights if seg_weights is not None else \
            torch.full((self.n_seg,), 1.0 / self.n_seg, dtype=torch.float32, device=dev)
        self.rows = _rows(names)
        self.P = len(self.rows)
        self.spec = _lib.tangent_spec(self.rows)
        self.lo = None if lo is None else np.asarray(lo, np.float64)
        self.hi = None if hi is None else np.asarray(hi, np.float64)
        L = self.P + 1
        self.Jr = torch.empty((self.n_seg, horizon * 37, L), dtype=torch.float32, device=dev)
        self.gram = torch.empty((self.n_seg, L, L), dtype=torch.float32, device=dev)
        self.status = torch.empty(self.n_seg, dtype=torch.int32, device=dev)
        self.lam = torch.empty(1, dtype=torch.float64, device=dev)
        self.delta = torch.empty((1, self.P), dtype=torch.float64, device=dev)
        self.cost = torch.empty(1, dtype=torch.float64, device=dev)
        self.pred = torch.empty(1, dtype=torch.float64, device=dev)
        self._gram_theta: Optional[np.ndarray] = None     # the theta self.gram belongs to

    def _theta_dev(self, theta: np.ndarray) -> torch.Tensor:
        return torch.from_numpy(np.asarray(theta, np.float32).reshape(P_COUNT, 1).copy()).to(self.engine.device)

    def objectives(self, thetas: np.ndarray) -> np.ndarray:
        """Tracking costs sum_s w_s cost_s of k parameter vectors [k, 25], all in one fused launch
        (k candidates x all segments)."""
        th = np.atleast_2d(np.asarray(thetas, np.float32))
        t = torch.from_numpy(np.ascontiguousarray(th.T)).to(self.engine.device)
        c = self.engine.segment_costs(t, th.shape[0], self.segments, self.H, seg_gains=self.seg_gains)
        return (self.seg_w[:, None] * c).sum(0).double().cpu().numpy()

    def objective(self, theta: np.ndarray) -> float:
        """Tracking cost sum_s w_s cost_s of one parameter vector (fused primal kernel)."""
        return float(self.objectives(theta)[0])

    def step(self, theta: np.ndarray, lam: float):
        """One damped Gauss-Newton step: (delta [P], cost at theta, predicted cost after it).
        The residual Jacobian and [J r]^T [J r] are kept for the last theta: a step retried at
        the same point with another damping only re-solves the normal equations."""
        L = _lib.lib()
        if self._gram_theta is None or not np.array_equal(self._gram_theta, theta):
            t = self._theta_dev(theta)
            rc = getattr(L, self.jacobian_entry)(C_ref(self.engine.args), C_ref(self.cargs), C_ref(self.spec),
                                                 _lib.ptr(t), 1, 1, _lib.ptr(self.segments), _lib.ptr(self.seg_w),
                                                 self.n_seg, self.H, _lib.ptr(self.Jr),
                                                 _lib.ptr(self.status, torch.int32), _lib.stream())
            _lib.check(rc, self.jacobian_entry)
            _lib.check(L.spi_fim_assemble(_lib.ptr(self.Jr), self.n_seg, self.H * 37, self.P + 1,
                                          _lib.ptr(self.gram), _lib.stream()), "spi_fim_assemble")
            self._gram_theta = np.array(theta, np.float64, copy=True)
        self.lam.fill_(lam)
        rc = L.spi_gn_step(_lib.ptr(self.gram), _lib.ptr(self.status, torch.int32), 1, self.n_seg, self.P,
                           _lib.ptr(self.lam, torch.float64), _lib.ptr(self.delta, torch.float64),
                           _lib.ptr(self.cost, torch.float64), _lib.ptr(self.pred, torch.float64), _lib.stream())
        _lib.check(rc, "spi_gn_step")
        out = torch.cat([self.delta[0], self.cost, self.pred]).cpu().numpy()      # one read-back
        return out[:self.P], float(out[self.P]), float(out[self.P + 1])

    def refine(self, theta0: np.ndarray, iterations: int = 10, lam0: float = 1e-3, tol: float = 1e-10,
               line_search: int = 4) -> RefineResult:
        """Damped Gauss-Newton from theta0.  Each iteration solves for the step at the current
        damping and evaluates it together with line_search shortened copies (1/2, 1/4, ... of
        its length) in one launch; the lowest cost below the current one is taken (the damping
        then drops by 3 if the full step won), otherwise the damping grows by 4.  The contacts
        make the cost piecewise smooth, and near a crease of it the full step often overshoots
        where a shorter one still descends."""
        theta = np.asarray(theta0, np.float64).copy()
        cost = self.objective(theta)
        res = RefineResult(theta.copy(), cost, [cost])
        lam = lam0
        fracs = 0.5 ** np.arange(line_search + 1)
        for it in range(iterations):
            delta, c_here, pred = self.step(theta, lam)
            cands = np.repeat(theta[None], len(fracs), axis=0)
            cands[:, self.rows] += fracs[:, None] * delta[None, :]
            if self.lo is not None:
                cands = np.clip(cands, self.lo, self.hi)
            c_new = self.objectives(cands)
            c_new = np.where(np.isfinite(c_new), c_new, np.inf)
            k = int(np.argmin(c_new))
            res.iterations = it + 1
            if c_new[k] < cost:
                gain = cost - c_new[k]
                theta, cost = cands[k].copy(), float(c_new[k])
                if k == 0:
                    lam = max(lam / 3.0, 1e-9)
                res.accepted += 1
                res.history.append(cost)
                if gain <= tol * max(cost, 1e-30):
                    break
            else:
                lam = min(lam * 4.0, 1e6)
        res.theta, res.cost = theta, cost
        return res
