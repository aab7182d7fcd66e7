"""An Optuna-shaped study over the on-device CEM (ask / tell / best_params / best_value).

``ActiveSysId.optimize`` (spigym/agents/sysid/active_sysid.py:165-217) builds
``optuna.create_study(direction, sampler)`` from ``optimize.sampler._target_`` and then, per
iteration, asks ``num_main_envs`` trials, turns each into commands with ``trial.suggest_float``
and tells the (negated) reward back.  ``create_study`` here returns a study with that surface
whose sampler is the device CEM: a batch of trials is one ``spi_cem_sample`` launch, the tell of
the batch's last trial is one ``spi_cem_update`` launch.

Define-by-run: the first batch records the parameter names and ranges its trials suggest and
draws them uniformly (CEM iteration 0 with mean = centre, std = width / 2, clipped); from then
on the search space is fixed and every batch is sampled on the device.  ``state_dict`` /
``load_state_dict`` resume a study (the reference pickles its Optuna study, active_sysid.py:178-183).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from .. import _lib


@dataclass
class DeviceCEMSampler:
    """``optimize.sampler`` config target: ``_target_: spi_amd.agents.study.DeviceCEMSampler``."""
    batch_size: int = 1024          # trials per CEM iteration (num_main_envs in the reference)
    elite_frac: float = 0.1
    alpha: float = 0.0              # smoothing of the moments
    min_std_frac: float = 1e-3      # std floor, as a fraction of each range
    seed: int = 0


@dataclass
class FrozenTrial:
    number: int
    params: Dict[str, float]
    value: Optional[float] = None


class DeviceTrial:
    def __init__(self, study: "DeviceStudy", number: int, row: Optional[np.ndarray]):
        self.study, self.number, self._row = study, number, row
        self.params: Dict[str, float] = {}
        self._k = 0

    def suggest_float(self, name: str, low: float, high: float) -> float:
        st = self.study
        if st._names is None:                        # recording batch: uniform draw, space grows
            if name not in st._space:
                st._space[name] = (float(low), float(high))
            v = float(st._rng.uniform(low, high)) if high > low else float(low)
        else:
            k = st._index.get(name)
            if k is None:
                raise ValueError(f"parameter {name!r} was not part of the first batch's search space")
            v = float(self._row[k])
        self.params[name] = v
        return v


class DeviceStudy:
    def __init__(self, direction: str = "minimize", sampler: Optional[DeviceCEMSampler] = None,
                 study_name: Optional[str] = None, device="cuda"):
        if direction not in ("minimize", "maximize"):
            raise ValueError(direction)
        self.direction, self.study_name = direction, study_name
        self.sampler = sampler if sampler is not None else DeviceCEMSampler()
        self.device = torch.device(device)
        self._rng = np.random.default_rng(self.sampler.seed)
        self._space: Dict[str, tuple] = {}
        self._names: Optional[List[str]] = None
        self._index: Dict[str, int] = {}
        self.trials: List[FrozenTrial] = []
        self._batch_rows: Optional[np.ndarray] = None
        self._batch: List[DeviceTrial] = []
        self._told: Dict[int, float] = {}
        self._n_told = 0
        self._iteration = 0
        self._best: Optional[FrozenTrial] = None

    # ------------------------------------------------------------ Optuna surface
    def ask(self) -> DeviceTrial:
        B = self.sampler.batch_size
        if self._names is not None and (self._batch_rows is None or len(self._batch) == B):
            self._sample_batch()
        if self._names is None and len(self._batch) == B:
            raise RuntimeError("tell every trial of the first batch before asking more")
        k = len(self._batch)
        t = DeviceTrial(self, len(self.trials), None if self._batch_rows is None else self._batch_rows[k])
        self._batch.append(t)
        self.trials.append(FrozenTrial(t.number, t.params))
        return t

    def tell(self, trial: DeviceTrial, value: float) -> None:
        v = float(value)
        self.trials[trial.number].value = v
        self.trials[trial.number].params = dict(trial.params)
        key = v if self.direction == "minimize" else -v
        if self._best is None or key < (self._best.value if self.direction == "minimize" else -self._best.value):
            self._best = self.trials[trial.number]
        if trial.number not in self._told and self._batch and \
                self._batch[0].number <= trial.number <= self._batch[-1].number:
            self._n_told += 1                        # trials of the current batch told so far
        self._told[trial.number] = v
        if len(self._batch) == self.sampler.batch_size and self._n_told == len(self._batch):
            self._update()

    @property
    def best_trial(self) -> FrozenTrial:
        if self._best is None:
            raise ValueError("no completed trial")
        return self._best

    @property
    def best_value(self) -> float:
        return self.best_trial.value

    @property
    def best_params(self) -> Dict[str, float]:
        return dict(self.best_trial.params)

    # ------------------------------------------------------------ device CEM
    def _init_cem(self):
        self._names = list(self._space)
        self._index = {n: i for i, n in enumerate(self._names)}
        lo = np.array([self._space[n][0] for n in self._names], np.float32)
        hi = np.array([self._space[n][1] for n in self._names], np.float32)
        f = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(self.device)
        self.lo, self.hi = f(lo), f(hi)
        self.mean, self.std = f(0.5 * (lo + hi)), f(0.5 * (hi - lo))
        self.min_std = f(np.maximum(self.sampler.min_std_frac * (hi - lo), 1e-12))
        B, P = self.sampler.batch_size, len(self._names)
        self.theta = torch.empty((P, B), dtype=torch.float32, device=self.device)
        self.cost = torch.empty(B, dtype=torch.float32, device=self.device)
        self.best_dev = self.mean.clone()
        self.best_cost_dev = torch.full((1,), math.inf, dtype=torch.float32, device=self.device)

    def _sample_batch(self):
        P, B = len(self._names), self.sampler.batch_size
        rc = _lib.lib().spi_cem_sample(_lib.ptr(self.mean), _lib.ptr(self.std), _lib.ptr(self.lo), _lib.ptr(self.hi),
                                       P, B, self.sampler.seed, self._iteration, _lib.ptr(self.theta), B,
                                       _lib.stream())
        _lib.check(rc, "spi_cem_sample")
        self._batch_rows = self.theta.t().cpu().numpy().astype(np.float64)
        self._batch, self._told, self._n_told = [], {}, 0

    def _update(self):
        B = self.sampler.batch_size
        if self._names is None:                      # the recording batch fixes the space
            self._init_cem()
            rows = np.array([[t.params[n] for n in self._names] for t in self._batch], np.float32)
            self.theta.copy_(torch.from_numpy(np.ascontiguousarray(rows.T)).to(self.device))
        vals = np.array([self._told[t.number] for t in self._batch], np.float32)
        if self.direction == "maximize":
            vals = -vals
        self.cost.copy_(torch.from_numpy(vals).to(self.device))
        P = len(self._names)
        n_elite = max(1, int(math.ceil(self.sampler.elite_frac * B - 1e-9)))
        rc = _lib.lib().spi_cem_update(_lib.ptr(self.cost), _lib.ptr(self.theta), B, P, B, n_elite,
                                       float(self.sampler.alpha), _lib.ptr(self.min_std), _lib.ptr(self.mean),
                                       _lib.ptr(self.std), _lib.ptr(self.best_dev), _lib.ptr(self.best_cost_dev),
                                       None, _lib.stream())
        _lib.check(rc, "spi_cem_update")
        self._iteration += 1
        self._batch_rows = None

    # ------------------------------------------------------------ resume
    def state_dict(self) -> Dict:
        d = {"direction": self.direction, "iteration": self._iteration, "space": dict(self._space),
             "names": self._names, "trials": [(t.number, t.params, t.value) for t in self.trials]}
        if self._names is not None:
            d.update({"mean": self.mean.cpu().numpy(), "std": self.std.cpu().numpy()})
        return d

    def load_state_dict(self, d: Dict) -> None:
        self.direction, self._iteration = d["direction"], int(d["iteration"])
        self._space = dict(d["space"])
        self.trials = [FrozenTrial(n, dict(p), v) for n, p, v in d["trials"]]
        done = [t for t in self.trials if t.value is not None]
        sgn = 1.0 if self.direction == "minimize" else -1.0
        self._best = min(done, key=lambda t: sgn * t.value) if done else None
        self._names = None
        if d.get("names") is not None:
            self._space = {n: self._space[n] for n in d["names"]}
            self._init_cem()
            self.mean.copy_(torch.from_numpy(np.asarray(d["mean"], np.float32)).to(self.device))
            self.std.copy_(torch.from_numpy(np.asarray(d["std"], np.float32)).to(self.device))
        self._batch, self._told, self._batch_rows, self._n_told = [], {}, None, 0


def create_study(direction: str = "minimize", sampler: Optional[DeviceCEMSampler] = None,
                 study_name: Optional[str] = None, load_if_exists: bool = True, device="cuda") -> DeviceStudy:
    """Signature of ``optuna.create_study`` as active_sysid.py:185-188 calls it."""
    return DeviceStudy(direction, sampler, study_name, device)
