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

N_SPECULATE = 2    # line-search points (full step, half step) whose Jacobians start early


@dataclass
class RefineResult:
    theta: np.ndarray                  # [25]
    cost: float
    history: List[float] = field(default_factory=list)
    iterations: int = 0
    accepted: int = 0
    taken: List[int] = field(default_factory=list)   # line-search point of each accepted step (0: full)


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
        # Two buffer sets, each for N_SPECULATE points: the current point's [J r], Gram and
        # status live in one set while the next points' are computed into the other.
        self._sets = [{"Jr": torch.empty((N_SPECULATE * self.n_seg, horizon * 37, L), dtype=torch.float32, device=dev),
                       "gram": torch.empty((N_SPECULATE * self.n_seg, L, L), dtype=torch.float32, device=dev),
                       "status": torch.empty(N_SPECULATE * self.n_seg, dtype=torch.int32, device=dev)}
                      for _ in range(2)]
        self.lam = torch.empty(1, dtype=torch.float64, device=dev)
        self.delta = torch.empty((1, self.P), dtype=torch.float64, device=dev)
        self.cost = torch.empty(1, dtype=torch.float64, device=dev)
        self.pred = torch.empty(1, dtype=torch.float64, device=dev)
        self._cur: Optional[tuple] = None    # (theta, set, slot) whose Gram step() uses
        self._spec: Optional[tuple] = None   # (set, [theta, ...]) computed ahead by speculate()
        self._side = torch.cuda.Stream(device=dev)   # default priority

    @property
    def gram(self) -> torch.Tensor:
        """[n_seg, P + 1, P + 1] augmented Gram matrices of the current point."""
        _, b, k = self._cur
        return self._sets[b]["gram"][k * self.n_seg:(k + 1) * self.n_seg]

    def _jacobians(self, b: int, thetas: np.ndarray) -> None:
        """[J r] and [J r]^T [J r] of every segment at each row of thetas [k, 25] (k <=
        N_SPECULATE) into slots 0..k-1 of buffer set b, on the current stream."""
        L, S = _lib.lib(), self._sets[b]
        k = thetas.shape[0]
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(thetas, np.float32).T)).to(self.engine.device)
        rc = getattr(L, self.jacobian_entry)(C_ref(self.engine.args), C_ref(self.cargs), C_ref(self.spec),
                                             _lib.ptr(t), k, k, _lib.ptr(self.segments), _lib.ptr(self.seg_w),
                                             self.n_seg, self.H, _lib.ptr(S["Jr"]),
                                             _lib.ptr(S["status"], torch.int32), _lib.stream())
        _lib.check(rc, self.jacobian_entry)
        _lib.check(L.spi_fim_assemble(_lib.ptr(S["Jr"]), k * self.n_seg, self.H * 37, self.P + 1,
                                      _lib.ptr(S["gram"]), _lib.stream()), "spi_fim_assemble")

    def _start_at(self, theta: np.ndarray) -> None:
        """Make theta the current point: its Jacobian and Gram into the set the current point
        does not use (any speculation there is dropped)."""
        b = 0 if self._cur is None else 1 - self._cur[1]
        self._jacobians(b, np.asarray(theta, np.float64)[None])
        self._cur = (np.array(theta, np.float64, copy=True), b, 0)
        self._spec = None

    def speculate(self, thetas: np.ndarray) -> None:
        """Start the Jacobians of up to N_SPECULATE likely next points (not waited for): a later
        step() at one of them finds its Gram ready.  The current point's buffers are untouched."""
        thetas = np.atleast_2d(np.asarray(thetas, np.float64))[:N_SPECULATE]
        b = 1 - self._cur[1]
        self._jacobians(b, thetas)
        self._spec = (b, [th.copy() for th in thetas])

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
        the same point with another damping only re-solves the normal equations; at a point
        that speculate() started, they are taken from there."""
        L = _lib.lib()
        theta = np.asarray(theta, np.float64)
        if self._cur is None or not np.array_equal(self._cur[0], theta):
            hit = None
            if self._spec is not None:
                hit = next((k for k, th in enumerate(self._spec[1]) if np.array_equal(th, theta)), None)
            if hit is not None:
                self._cur = (theta.copy(), self._spec[0], hit)
                self._spec = None
            else:
                self._start_at(theta)
        _, b, k = self._cur
        S, g0 = self._sets[b], k * self.n_seg
        self.lam.fill_(lam)
        rc = L.spi_gn_step(_lib.ptr(S["gram"][g0:]), _lib.ptr(S["status"][g0:], torch.int32), 1, self.n_seg, self.P,
                           _lib.ptr(self.lam, torch.float64), _lib.ptr(self.delta, torch.float64),
                           _lib.ptr(self.cost, torch.float64), _lib.ptr(self.pred, torch.float64), _lib.stream())
        _lib.check(rc, "spi_gn_step")
        out = torch.cat([self.delta[0], self.cost, self.pred]).cpu().numpy()      # one read-back
        return out[:self.P], float(out[self.P]), float(out[self.P + 1])

    def refine(self, theta0: np.ndarray, iterations: int = 10, lam0: float = 1e-3, tol: float = 1e-10,
               line_search: int = 4, speculate: bool = True) -> RefineResult:
        """Damped Gauss-Newton from theta0.  Each iteration solves for the step at the current
        damping and evaluates it together with line_search shortened copies (1/2, 1/4, ... of
        its length) in one launch; the lowest cost below the current one is taken (the damping
        then drops by 3 if the full step won), otherwise the damping grows by 4.  The contacts
        make the cost piecewise smooth, and near a crease of it the full step often overshoots
        where a shorter one still descends.

        With speculate, the Jacobians at the full and the half step are started before the
        line search, which runs beside them on a second stream; when one of the two wins, the
        next iteration's Jacobian is already there.  The first Jacobian likewise runs beside the
        cost at theta0.  The iterates are the same either way."""
        theta = np.asarray(theta0, np.float64).copy()
        self._cur = self._spec = None          # every refinement computes its own Jacobians
        if speculate:
            # the side stream sees what the caller's stream wrote so far (the segments, the
            # weights); nothing that this refinement enqueues later
            self._side.wait_stream(torch.cuda.current_stream())
            # the first Jacobian and the cost at theta0 side by side
            self._start_at(theta)
            with torch.cuda.stream(self._side):
                cost = self.objective(theta)
        else:
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
            if speculate and it + 1 < iterations:     # the last iteration has no next step
                self.speculate(cands)
            if speculate:
                # the line search reads only the segments, the weights and host data, none of
                # which the caller's stream writes during the refinement
                with torch.cuda.stream(self._side):
                    c_new = self.objectives(cands)
            else:
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
                res.taken.append(k)
                res.history.append(cost)
                if gain <= tol * max(cost, 1e-30):
                    break
            else:
                lam = min(lam * 4.0, 1e6)
        res.theta, res.cost = theta, cost
        return res
