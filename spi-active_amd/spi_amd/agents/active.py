"""Active system identification: score excitations by the finite-difference Fisher information.

The reference evaluates a batch of command sequences by stepping (1 main + P aux) env groups
through PhysX and, every control step, summing ``||X||_F^2`` with
``X = (state_main - state_aux) / delta`` over root state and joint positions
(spigym/envs/sysid/active_sysid_openloop.py:402-426), averaged over the rollout
(spigym/agents/sysid/active_sysid.py:527-600); aux env k carries ``param_k + delta`` (absolute,
active_sysid_openloop.py:62-68).  Here a whole group lives in neighbouring lanes of one
wavefront and the objective is one fused kernel (``spi_fim_rollout``).
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .. import _lib
from ..params import PARAM_NAMES, P_COUNT
from ..rollout import C_ref, RolloutEngine

KSYNC_MODES = {"reference": 0, "state": 1}


class FDFIMObjective:
    def __init__(self, engine: RolloutEngine, theta: np.ndarray, explore: Sequence[str], delta: float = 0.1,
                 ksync_steps: int = 5, ksync_mode: str = "reference"):
        self.engine = engine
        self.theta = np.asarray(theta, np.float64)
        self.idx = [PARAM_NAMES.index(p) if p in PARAM_NAMES else _alias(p) for p in explore]
        self.P = len(self.idx)
        self.delta = float(delta)
        # active_sysid_openloop.py:246-252: ksync_steps <= 1 means every step
        self.ksync = max(int(ksync_steps), 1)
        self.ksync_mode = KSYNC_MODES[ksync_mode]

    def params(self, n_group: int) -> torch.Tensor:
        G = self.P + 1
        rows = np.repeat(self.theta[None], n_group * G, axis=0)
        for k, p in enumerate(self.idx):
            rows[k + 1::G, p] += self.delta
        return torch.from_numpy(np.ascontiguousarray(rows.T.astype(np.float32))).to(self.engine.device)

    def __call__(self, actions: torch.Tensor, init: torch.Tensor, mean: bool = True) -> torch.Tensor:
        """actions [M, T, 12]; init [M * (P + 1), 37] (or [37] for all envs) -> FIM per group [M]."""
        M, T = actions.shape[0], actions.shape[1]
        G = self.P + 1
        if init.dim() == 1:
            init = init[None].expand(M * G, 37).contiguous()
        par = self.params(M)
        out = torch.empty(M, dtype=torch.float32, device=self.engine.device)
        actions = actions.contiguous()   # keep the (possibly copied) operand alive through the launch
        rc = _lib.lib().spi_fim_rollout(C_ref(self.engine.args), _lib.ptr(par), par.shape[1], M, self.P,
                                        _lib.ptr(init), _lib.ptr(actions), T, self.delta, self.ksync,
                                        self.ksync_mode, _lib.ptr(out), _lib.stream())
        _lib.check(rc, "spi_fim_rollout")
        return out / T if mean else out


def _alias(name: str) -> int:
    # reference key names (env/active_sysid_openloop.yaml:17-27)
    table = {"inertiay": "inertiay", "motor_model_hip_a": "motor_model_hip_a"}
    return PARAM_NAMES.index(table.get(name, name))
