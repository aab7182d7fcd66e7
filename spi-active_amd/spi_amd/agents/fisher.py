"""Exact trajectory sensitivities, Fisher information and its log-determinant on the GPU.

The reference scores an excitation by finite differences: every main env has P aux envs
with ``param_k + delta`` (active_sysid_openloop.py:29-80) and the reward is
``||(state_main - state_aux) / delta||_F^2`` summed over the rollout (:402-426), i.e. the
trace of a finite-difference Fisher information.  Here the derivatives are exact: the rollout
is differentiated in forward mode (``spi_traj_sensitivity``), giving
``J_t = d y_t / d theta`` with ``y = [root (13) | q (12)]`` (the state the reference's reward
uses) for every control step.  The Fisher information of a (parameter sample, excitation)
group is ``F = sum_t J_t^T J_t`` (``spi_fim_assemble``, FP32 matrix cores); an excitation is
scored by the log-determinant of its Fisher information averaged over the parameter samples
(D-optimal design, ``spi_fim_logdet``: FP64 mean and Cholesky).

``RolloutCost`` puts the fused rollout cost behind ``torch.autograd.Function``: its backward
is the exact gradient of the discrete rollout (``spi_rollout_cost_grad``).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from .. import _lib
from ..data import segment_stride
from ..params import PARAM_NAMES, P_COUNT
from ..rollout import C_ref, RolloutEngine

IDENTIFIED = PARAM_NAMES[:22]
FIM_ROWS = 25          # y = root (13) + q (12)


def _rows(names: Sequence[str]):
    return [PARAM_NAMES.index(n) for n in names]


def fourier_excitations(n_exc: int, horizon: int, dt_ctrl: float, action_scale: float, n_terms: int = 8,
                        coeff: float = 0.3, seed: int = 0, period: Optional[float] = None,
                        device="cuda") -> torch.Tensor:
    """[n_exc, horizon, 12] raw actions whose PD-target offsets (action x action_scale) are
    Fourier series sum_k c_k sin(2 pi k t / period), k = 1..n_terms, c_k ~ U[-coeff, coeff] rad,
    one series per joint; period defaults to the horizon's duration."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-coeff, coeff, size=(n_exc, 12, n_terms)).astype(np.float32)
    cd = torch.from_numpy(c).to(device)
    out = torch.empty((n_exc, horizon, 12), dtype=torch.float32, device=device)
    period = float(period if period is not None else horizon * dt_ctrl)
    rc = _lib.lib().spi_fourier_excitations(_lib.ptr(cd), n_exc, n_terms, horizon, float(dt_ctrl), period,
                                            1.0 / float(action_scale), _lib.ptr(out), _lib.stream())
    _lib.check(rc, "spi_fourier_excitations")
    return out


@dataclass
class FisherScores:
    logdet: torch.Tensor        # [E] float64
    fim_mean: torch.Tensor      # [E, P, P] float64
    n_valid: torch.Tensor       # [E] int32


class FisherObjective:
    """D-optimal excitation scoring from exact sensitivities."""

    def __init__(self, engine: RolloutEngine, explore: Sequence[str] = IDENTIFIED, reg: float = 1e-6,
                 scale: Optional[Sequence[float]] = None, legs: bool = True, pairs: bool = True):
        """scale: per-parameter magnitudes (e.g. the prior mean): the logdet is then that of the
        information about relative parameter changes, comparable across units; None = raw.
        reg: ridge relative to the mean eigenvalue (trace / P).  Tangent rollouts: four lanes
        (one per leg) per two tangents on one value path (spi_traj_sensitivity_pairs, the
        default: 15 % faster than one tangent per unit), per tangent with pairs=False
        (spi_traj_sensitivity_legs), or one lane per tangent with legs=False
        (spi_traj_sensitivity)."""
        self.engine = engine
        self.legs = bool(legs)
        self.pairs = bool(pairs) and self.legs
        self.names = list(explore)
        self.spec = _lib.tangent_spec(_rows(self.names))
        self.P = len(self.names)
        self.reg = float(reg)
        self.scale = None if scale is None else torch.as_tensor(
            np.abs(np.asarray(scale, np.float64)), dtype=torch.float64, device=engine.device).contiguous()
        if self.scale is not None and self.scale.numel() != self.P:
            raise ValueError("one scale per explored parameter")

    # -- the three device stages
    def sensitivities(self, theta: torch.Tensor, actions: torch.Tensor, init: torch.Tensor, with_y: bool = False):
        """theta [25, S] (SoA), actions [E, H, 12], init [37] -> J [E*S, H*25, P], status [E*S],
        y [E*S, H, 25] or None.  Group g = e * S + c."""
        S = theta.shape[1]
        E, H = actions.shape[0], actions.shape[1]
        dev = self.engine.device
        G = E * S
        J = torch.empty((G, H * FIM_ROWS, self.P), dtype=torch.float32, device=dev)
        st = torch.empty(G, dtype=torch.int32, device=dev)
        y = torch.empty((G, H, FIM_ROWS), dtype=torch.float32, device=dev) if with_y else None
        actions = actions.contiguous()   # keep the (possibly copied) operand alive through the launch
        L = _lib.lib()
        if self.pairs:
            fn = L.spi_traj_sensitivity_pairs
        else:
            fn = L.spi_traj_sensitivity_legs if self.legs else L.spi_traj_sensitivity
        rc = fn(C_ref(self.engine.args), C_ref(self.spec), _lib.ptr(theta), S, S, _lib.ptr(init), _lib.ptr(actions),
                E, H, _lib.ptr(J), _lib.ptr(y), _lib.ptr(st, torch.int32), _lib.stream())
        _lib.check(rc, "spi_traj_sensitivity")
        return J, st, y

    def assemble(self, J: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        G, K, P = J.shape
        F = out if out is not None else torch.empty((G, P, P), dtype=torch.float32, device=J.device)
        _lib.check(_lib.lib().spi_fim_assemble(_lib.ptr(J), G, K, P, _lib.ptr(F), _lib.stream()), "spi_fim_assemble")
        return F

    def logdet(self, F: torch.Tensor, status: torch.Tensor, n_sets: int) -> FisherScores:
        G, P = F.shape[0], F.shape[1]
        if G % n_sets:
            raise ValueError("groups must split evenly into sets")
        dev = F.device
        ld = torch.empty(n_sets, dtype=torch.float64, device=dev)
        Fm = torch.empty((n_sets, P, P), dtype=torch.float64, device=dev)
        nv = torch.empty(n_sets, dtype=torch.int32, device=dev)
        rc = _lib.lib().spi_fim_logdet(_lib.ptr(F), _lib.ptr(status, torch.int32), n_sets, G // n_sets, P,
                                       _lib.ptr(self.scale, torch.float64), self.reg,
                                       _lib.ptr(Fm, torch.float64), _lib.ptr(ld, torch.float64),
                                       _lib.ptr(nv, torch.int32), _lib.stream())
        _lib.check(rc, "spi_fim_logdet")
        return FisherScores(ld, Fm, nv)

    # -- whole objective, chunked over excitations to bound the sensitivity buffer
    def score(self, theta: torch.Tensor, actions: torch.Tensor, init: torch.Tensor,
              max_sens_bytes: float = 16e9) -> FisherScores:
        S = theta.shape[1]
        E, H = actions.shape[0], actions.shape[1]
        per_exc = S * H * FIM_ROWS * self.P * 4
        chunk = max(1, min(E, int(max_sens_bytes // max(per_exc, 1))))
        dev = self.engine.device
        F = torch.empty((E * S, self.P, self.P), dtype=torch.float32, device=dev)
        status = torch.empty(E * S, dtype=torch.int32, device=dev)
        for e0 in range(0, E, chunk):
            e1 = min(E, e0 + chunk)
            J, st, _ = self.sensitivities(theta, actions[e0:e1], init)
            self.assemble(J, out=F[e0 * S:e1 * S])
            status[e0 * S:e1 * S] = st
            del J
        return self.logdet(F, status, E)


def adjoint_cost_grad(engine: RolloutEngine, theta: torch.Tensor, segments: torch.Tensor, horizon: int,
                      checkpoint_every: int = 10, legs: bool = True, seg_gains: Optional[torch.Tensor] = None):
    """Cost [n_seg, n] and gradient [n_seg, n, 25] (all parameter rows) by reverse mode with
    checkpoints every ``checkpoint_every`` control steps: four lanes per rollout
    (``spi_rollout_cost_adjoint_legs``, the default: no register spills, 1.6x faster at
    4096 x 256 x H 20) or, with ``legs=False``, one lane per rollout
    (``spi_rollout_cost_adjoint``)."""
    n = theta.shape[1]
    n_seg = segments.shape[0]
    if segments.shape[1] != segment_stride(horizon):
        raise ValueError("segment stride does not match the horizon")
    L = _lib.lib()
    name = "spi_rollout_cost_adjoint_legs" if legs else "spi_rollout_cost_adjoint"
    if legs:
        wf = int(L.spi_adjoint_legs_work_floats(n, n_seg, horizon, checkpoint_every, engine.args.decimation))
    else:
        wf = int(L.spi_adjoint_work_floats(n, n_seg, horizon, checkpoint_every))
    if wf < 0:
        raise ValueError("bad sizes")
    dev = theta.device
    work = torch.empty(max(wf, 1), dtype=torch.float32, device=dev)
    cost = torch.empty((n_seg, n), dtype=torch.float32, device=dev)
    grad = torch.empty((n_seg, n, P_COUNT), dtype=torch.float32, device=dev)
    cargs = engine.cost_args_for(n_seg, seg_gains)
    rc = getattr(L, name)(C_ref(engine.args), C_ref(cargs), _lib.ptr(theta), n, n, _lib.ptr(segments), n_seg,
                          horizon, checkpoint_every, _lib.ptr(work), wf, _lib.ptr(cost), _lib.ptr(grad),
                          _lib.stream())
    _lib.check(rc, name)
    return cost, grad


class RolloutCost(torch.autograd.Function):
    """cost[c] = sum_s w_s cost_s(theta[:, c]) with the exact gradient of the discrete rollout
    with respect to the parameter rows ``rows`` (other rows get zero gradient).

    method "adjoint" (default): the reverse-mode kernel with checkpoints every
    ``checkpoint_every`` control steps, which the forward pass runs together with the cost;
    method "tangent": forward mode, one lane per parameter row."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, engine: RolloutEngine, segments: torch.Tensor, horizon: int,
                seg_weights: torch.Tensor, rows: Sequence[int], method: str = "adjoint", checkpoint_every: int = 10,
                seg_gains: Optional[torch.Tensor] = None):
        n = theta.shape[1]
        theta = theta.detach().contiguous()
        if method == "adjoint":
            cost, grad = adjoint_cost_grad(engine, theta, segments, horizon, checkpoint_every, seg_gains=seg_gains)
            w = seg_weights.to(torch.float32)
            total = (w[:, None] * cost).sum(0)
            full = torch.zeros((P_COUNT, n), dtype=torch.float32, device=theta.device)
            idx = torch.as_tensor(list(rows), dtype=torch.long, device=theta.device)
            full[idx] = (w[:, None, None] * grad).sum(0).t()[idx]
            ctx.save_for_backward(full)
            return total
        if method != "tangent":
            raise ValueError(f"unknown method {method!r}")
        spec = _lib.tangent_spec(rows)
        n_seg = segments.shape[0]
        if segments.shape[1] != segment_stride(horizon):
            raise ValueError("segment stride does not match the horizon")
        cost = torch.empty((n_seg, n), dtype=torch.float32, device=theta.device)
        grad = torch.empty((n_seg, n, spec.n_tan), dtype=torch.float32, device=theta.device)
        cargs = engine.cost_args_for(n_seg, seg_gains)
        rc = _lib.lib().spi_rollout_cost_grad(C_ref(engine.args), C_ref(cargs), C_ref(spec), _lib.ptr(theta),
                                              n, n, _lib.ptr(segments), n_seg, horizon, _lib.ptr(cost),
                                              _lib.ptr(grad), _lib.stream())
        _lib.check(rc, "spi_rollout_cost_grad")
        w = seg_weights.to(torch.float32)
        total = (w[:, None] * cost).sum(0)
        g_total = (w[:, None, None] * grad).sum(0)                       # [n, P]
        full = torch.zeros((P_COUNT, n), dtype=torch.float32, device=theta.device)
        full[list(rows)] = g_total.t()
        ctx.save_for_backward(full)
        return total

    @staticmethod
    def backward(ctx, g_out):
        (full,) = ctx.saved_tensors
        return full * g_out[None, :], None, None, None, None, None, None, None, None
