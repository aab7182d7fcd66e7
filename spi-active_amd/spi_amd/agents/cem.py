"""Cross-entropy method with all state on the device.

Keeps the reference's optimizer interface -- ``ask`` a batch of candidates, evaluate them in
one batched rollout, ``tell`` the costs (spigym/agents/sysid/active_sysid.py:196-217, where
Optuna's ``study.ask()/tell()`` are called once per trial) -- but a whole iteration is three
kernel launches and never synchronises with the host:

  spi_cem_sample   theta = clamp(mean + std * z, lo, hi), z from Philox4x32-10 + Box-Muller
  (evaluate)       any batched objective, e.g. spi_rollout_cost + spi_reduce_segments
  spi_cem_update   bitonic sort of (cost, index) in LDS, elite moments, best-so-far

Parameters with ``std0 == 0`` and ``min_std == 0`` stay fixed at their mean.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Dict, Optional

import numpy as np
import torch

from .. import _lib


@dataclass
class CEMConfig:
    n_samples: int = 4096
    elite_frac: float = 0.1
    iterations: int = 12
    alpha: float = 0.0          # new = alpha * old + (1 - alpha) * elite moments
    seed: int = 0

    @property
    def n_elite(self) -> int:
        # ceil(0.1 * 4096) = 410
        return max(1, int(math.ceil(self.elite_frac * self.n_samples - 1e-9)))


class DeviceCEM:
    def __init__(self, mean0, std0, lo, hi, min_std, cfg: CEMConfig, device: str = "cuda"):
        _lib.require_gpu()
        if cfg.n_samples > _lib.SPI_MAX_CEM_SAMPLES:
            raise ValueError(f"n_samples > {_lib.SPI_MAX_CEM_SAMPLES}")
        self.cfg = cfg
        self.dev = torch.device(device)
        f = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=self.dev).contiguous()
        self.n_param = len(mean0)
        self.mean, self.std = f(mean0).clone(), f(std0).clone()
        self.lo, self.hi, self.min_std = f(lo), f(hi), f(min_std)
        self.theta = torch.empty((self.n_param, cfg.n_samples), dtype=torch.float32, device=self.dev)
        self.best = self.mean.clone()
        self.best_cost = torch.full((1,), float("inf"), dtype=torch.float32, device=self.dev)
        self.elite_idx = torch.empty(cfg.n_elite, dtype=torch.int32, device=self.dev)
        self.iteration = 0
        self.L = _lib.lib()

    def ask(self) -> torch.Tensor:
        """Sample a batch: SoA [n_param, n_samples]."""
        rc = self.L.spi_cem_sample(_lib.ptr(self.mean), _lib.ptr(self.std), _lib.ptr(self.lo), _lib.ptr(self.hi),
                                   self.n_param, self.cfg.n_samples, self.cfg.seed, self.iteration,
                                   _lib.ptr(self.theta), self.cfg.n_samples, _lib.stream())
        _lib.check(rc, "spi_cem_sample")
        return self.theta

    def tell(self, cost: torch.Tensor) -> None:
        if cost.shape[0] != self.cfg.n_samples:
            raise ValueError("one cost per sample")
        rc = self.L.spi_cem_update(_lib.ptr(cost), _lib.ptr(self.theta), self.cfg.n_samples, self.n_param,
                                   self.cfg.n_samples, self.cfg.n_elite, self.cfg.alpha, _lib.ptr(self.min_std),
                                   _lib.ptr(self.mean), _lib.ptr(self.std), _lib.ptr(self.best),
                                   _lib.ptr(self.best_cost), _lib.ptr(self.elite_idx, torch.int32), _lib.stream())
        _lib.check(rc, "spi_cem_update")
        self.iteration += 1

    def solve(self, evaluate: Callable[[torch.Tensor, int], torch.Tensor], iterations: Optional[int] = None):
        for _ in range(iterations or self.cfg.iterations):
            theta = self.ask()
            self.tell(evaluate(theta, self.cfg.n_samples))
        return self.mean

    # ---------------------------------------------------------------- checkpoint / resume
    def state_dict(self) -> Dict[str, np.ndarray]:
        return {"mean": self.mean.cpu().numpy(), "std": self.std.cpu().numpy(), "best": self.best.cpu().numpy(),
                "best_cost": self.best_cost.cpu().numpy(), "iteration": np.array(self.iteration),
                "seed": np.array(self.cfg.seed, dtype=np.uint64)}

    def load_state_dict(self, d: Dict[str, np.ndarray]) -> None:
        for name in ("mean", "std", "best", "best_cost"):
            getattr(self, name).copy_(torch.as_tensor(np.asarray(d[name], np.float32), device=self.dev))
        self.iteration = int(d["iteration"])
        self.cfg.seed = int(d["seed"])

    def save(self, path: str) -> None:
        np.savez(path, **self.state_dict())

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as d:
            self.load_state_dict({k: d[k] for k in d.files})
