"""Many independent CEM sysID problems in one batched launch per CEM phase.

Each problem has its own segments, horizon, sample count, prior and Philox stream.  One CEM
iteration of the whole batch is three kernels -- ``spi_batch_cem_sample``,
``spi_batch_rollout_cost`` (a persistent grid pulling (problem, segment, candidate-block)
tasks from a device queue, longest horizons first) and ``spi_batch_cem_update`` (segment
reduction, elite sort, moments, iteration counter, one workgroup per problem) -- and can be
captured in a HIP graph and replayed, since every per-iteration quantity lives on the device.
A problem's results do not depend on which other problems share the batch.

The reference solves one problem at a time: a Python loop over candidate masses
(scripts/mass_landscape.py:123-126) or Optuna trials (scripts/mass_opt.py:203-219).
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .. import _lib
from ..params import P_COUNT
from ..rollout import C_ref, RolloutEngine
from .sysid import Prior


class SpiBatchProblem(C.Structure):
    _fields_ = [("segments", C.c_void_p), ("seg_weight", C.c_void_p), ("n_seg", C.c_int), ("horizon", C.c_int),
                ("cost", _lib.SpiCostArgs), ("n_samples", C.c_int), ("n_elite", C.c_int), ("alpha", C.c_float),
                ("iteration", C.c_uint), ("seed", C.c_ulonglong), ("mean", C.c_void_p), ("std", C.c_void_p),
                ("lo", C.c_void_p), ("hi", C.c_void_p), ("min_std", C.c_void_p), ("theta", C.c_void_p),
                ("cost_seg", C.c_void_p), ("cost_out", C.c_void_p), ("best", C.c_void_p), ("best_cost", C.c_void_p)]


@dataclass
class ProblemSpec:
    segments: torch.Tensor            # [n_seg, stride(H)] device
    horizon: int
    prior: Prior
    n_samples: int = 4096
    elite_frac: float = 0.1
    alpha: float = 0.0
    seed: int = 0
    seg_weights: Optional[torch.Tensor] = None
    seg_gains: Optional[torch.Tensor] = None   # recorded (kp, kd) per segment [n_seg, 2], or configured gains


def _validate(k: int, s: ProblemSpec) -> None:
    """The batched kernels read sizes from the device table and cannot check them: check here."""
    from ..data import segment_stride
    if not (1 <= s.n_samples <= _lib.SPI_MAX_CEM_SAMPLES):
        raise ValueError(f"problem {k}: n_samples {s.n_samples} outside [1, {_lib.SPI_MAX_CEM_SAMPLES}]")
    n_elite = max(1, int(math.ceil(s.elite_frac * s.n_samples - 1e-9)))
    if n_elite > s.n_samples:
        raise ValueError(f"problem {k}: elite_frac {s.elite_frac} gives more elites than samples")
    if not (0.0 <= s.alpha < 1.0):
        raise ValueError(f"problem {k}: alpha {s.alpha} outside [0, 1)")
    if s.horizon < 1 or s.segments.dim() != 2 or s.segments.shape[0] < 1 or \
            s.segments.shape[1] != segment_stride(s.horizon):
        raise ValueError(f"problem {k}: segments {tuple(s.segments.shape)} do not match horizon {s.horizon}")
    if s.segments.dtype != torch.float32 or not s.segments.is_contiguous() or s.segments.device.type != "cuda":
        raise ValueError(f"problem {k}: segments must be a contiguous float32 device tensor")
    if s.seg_weights is not None and (s.seg_weights.numel() != s.segments.shape[0] or
                                      s.seg_weights.dtype != torch.float32 or s.seg_weights.device.type != "cuda"):
        raise ValueError(f"problem {k}: seg_weights must be float32 [n_seg] on the device")
    if s.seg_gains is not None and (tuple(s.seg_gains.shape) != (s.segments.shape[0], 2) or
                                    s.seg_gains.dtype != torch.float32 or not s.seg_gains.is_contiguous() or
                                    s.seg_gains.device.type != "cuda"):
        raise ValueError(f"problem {k}: seg_gains must be a contiguous float32 [n_seg, 2] device tensor")


_PRIORS: Dict[int, tuple] = {}


def _prior_on_device(prior: Prior, dev) -> tuple:
    """Device copies of a prior's bounds and initial moments, uploaded once per prior object."""
    hit = _PRIORS.get(id(prior))
    if hit is not None and hit[0] is prior and hit[1] == str(dev):
        return hit[2]
    f = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev).contiguous()
    t = (f(prior.mean), f(prior.std), f(prior.lo), f(prior.hi), f(prior.min_std))
    _PRIORS[id(prior)] = (prior, str(dev), t)
    return t


def validate_spec(spec: ProblemSpec) -> None:
    _validate(0, spec)


class _Buffers:
    """Device state of one problem (CEM moments, samples, costs, best so far)."""
    def __init__(self, spec: ProblemSpec, dev):
        n, n_seg = spec.n_samples, spec.segments.shape[0]
        mean, std, self.lo, self.hi, self.min_std = _prior_on_device(spec.prior, dev)
        self.mean, self.std = mean.clone(), std.clone()
        self.theta = torch.empty((P_COUNT, n), dtype=torch.float32, device=dev)
        self.cost_seg = torch.empty((n_seg, n), dtype=torch.float32, device=dev)
        self.cost = torch.empty(n, dtype=torch.float32, device=dev)
        self.best = mean.clone()
        self.best_cost = torch.full((1,), float("inf"), dtype=torch.float32, device=dev)
        self.seg_w = (spec.seg_weights if spec.seg_weights is not None else
                      torch.full((n_seg,), 1.0 / n_seg, dtype=torch.float32, device=dev))
        # device addresses in SpiBatchProblem field order (taken once)
        self.ptrs = (spec.segments.data_ptr(), self.seg_w.data_ptr(), self.mean.data_ptr(), self.std.data_ptr(),
                     self.lo.data_ptr(), self.hi.data_ptr(), self.min_std.data_ptr(), self.theta.data_ptr(),
                     self.cost_seg.data_ptr(), self.cost.data_ptr(), self.best.data_ptr(), self.best_cost.data_ptr())

    @staticmethod
    def footprint(spec: ProblemSpec) -> int:
        """Device bytes a problem holds while admitted (its segments included)."""
        n, n_seg = spec.n_samples, spec.segments.shape[0]
        return 4 * (P_COUNT * n + n_seg * n + n + 7 * P_COUNT + n_seg) + spec.segments.numel() * 4


_TABLE_DTYPE = np.dtype(SpiBatchProblem)
_PTR_FIELDS = ("segments", "seg_weight", "mean", "std", "lo", "hi", "min_std", "theta", "cost_seg", "cost_out",
               "best", "best_cost")


def pack_problem_table(specs, bufs, start_iter, cargs, block: int, dev):
    n = len(specs)
    arr = np.zeros(n, dtype=_TABLE_DTYPE)
    ptrs = np.array([b.ptrs for b in bufs], dtype=np.uint64).reshape(n, len(_PTR_FIELDS))
    for k, name in enumerate(_PTR_FIELDS):
        arr[name] = ptrs[:, k]
    n_seg = np.array([s.segments.shape[0] for s in specs], np.int64)
    n_samp = np.array([s.n_samples for s in specs], np.int64)
    arr["n_seg"] = n_seg
    arr["horizon"] = [s.horizon for s in specs]
    arr["cost"]["mode"] = cargs.mode
    arr["cost"]["w"] = np.array(list(cargs.w), np.float32)
    arr["cost"]["seg_gains"] = np.array([0 if s.seg_gains is None else s.seg_gains.data_ptr() for s in specs],
                                        np.uint64)
    arr["n_samples"] = n_samp
    arr["n_elite"] = [max(1, int(math.ceil(s.elite_frac * s.n_samples - 1e-9))) for s in specs]
    arr["alpha"] = [s.alpha for s in specs]
    arr["iteration"] = np.asarray(start_iter, np.uint32)
    arr["seed"] = np.array([s.seed for s in specs], np.uint64)
    tasks = n_seg * ((n_samp + block - 1) // block)
    prefix = np.concatenate([[0], np.cumsum(tasks)[:-1]]).astype(np.int64)
    return _upload(arr.view(np.uint8), dev), _upload(prefix, dev), int(tasks.sum())


def _upload(a: np.ndarray, dev) -> torch.Tensor:
    """Host array -> device tensor without a stream synchronisation (pinned staging from the
    caching host allocator, asynchronous copy on the current stream)."""
    t = torch.from_numpy(np.ascontiguousarray(a).copy())
    if torch.device(dev).type == "cuda":
        return t.pin_memory().to(dev, non_blocking=True)
    return t


class BatchCEM:
    """A fixed set of problems solved together; ``iterate()`` runs one CEM iteration of all."""

    def __init__(self, engine: RolloutEngine, specs: Sequence[ProblemSpec], block: int = 256, unroll: int = 1,
                 blocks_per_cu: int = 4, n_cu: int = 256, buffers: Optional[Sequence["_Buffers"]] = None,
                 iterations_done: Optional[Sequence[int]] = None, validate: bool = True):
        """validate=False skips the per-problem checks for specs a caller has already passed
        through ``validate_spec`` (the server does so once per job at admission)."""
        if not specs:
            raise ValueError("no problems")
        if validate:
            for k, s in enumerate(specs):
                _validate(k, s)
        self.engine = engine
        self.dev = engine.device
        # longest horizon first: the persistent grid takes tickets in table order
        self.order = sorted(range(len(specs)), key=lambda i: -specs[i].horizon)
        self.specs = [specs[i] for i in self.order]
        if buffers is None:
            buffers = [_Buffers(s, self.dev) for s in specs]
        self.bufs = [buffers[i] for i in self.order]
        done = list(iterations_done) if iterations_done is not None else [0] * len(specs)
        self.start_iter = [done[i] for i in self.order]
        self.block, self.unroll = block, unroll
        self.n_blocks = blocks_per_cu * n_cu
        self.queue = torch.zeros(2, dtype=torch.int64, device=self.dev)
        self.max_h = max(s.horizon for s in self.specs)
        self.max_n = max(s.n_samples for s in self.specs)
        self._build_table()

    def _build_table(self):
        """The device table of SpiBatchProblem records, filled column-wise (numpy structured
        array with the C layout), and the task prefix sums."""
        self.table, self.prefix, self.n_tasks = pack_problem_table(self.specs, self.bufs, self.start_iter,
                                                                   self.engine.cargs, self.block, self.dev)

    def device_iterations(self) -> List[int]:
        """Iteration counters as the device table holds them (table order)."""
        raw = self.table.cpu().numpy().tobytes()
        arr = (SpiBatchProblem * len(self.specs)).from_buffer_copy(raw)
        return [int(p.iteration) for p in arr]

    def set_launch(self, block: int, unroll: int) -> None:
        """Change envs per block / unroll depth.  Graphs captured before this call are stale."""
        if block != self.block:
            torch.cuda.current_stream().synchronize()
            self.start_iter = self.device_iterations()     # keep each problem's Philox position
            self.block = block
            self._build_table()
        self.unroll = unroll

    # ---------------------------------------------------------------- phases
    def sample(self):
        _lib.check(_lib.lib().spi_batch_cem_sample(self.table.data_ptr(), len(self.specs), self.max_n, _lib.stream()),
                   "spi_batch_cem_sample")

    def rollout(self):
        rc = _lib.lib().spi_batch_rollout_cost(C_ref(self.engine.args), self.table.data_ptr(), self.prefix.data_ptr(),
                                               len(self.specs), self.n_tasks, self.max_h, self.block, self.unroll,
                                               self.n_blocks, self.queue.data_ptr(), _lib.stream())
        _lib.check(rc, "spi_batch_rollout_cost")

    def update(self):
        _lib.check(_lib.lib().spi_batch_cem_update(self.table.data_ptr(), len(self.specs), _lib.stream()),
                   "spi_batch_cem_update")

    def iterate(self):
        self.sample()
        self.rollout()
        self.update()

    def solve(self, iterations: int, graph: bool = True):
        if not graph:
            for _ in range(iterations):
                self.iterate()
            return
        g = self.capture()
        for _ in range(iterations):
            g.replay()

    def capture(self) -> "torch.cuda.CUDAGraph":
        """HIP graph of one iteration (the problem table's iteration counters advance on device)."""
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            # capture without running: the graph's first replay is the next iteration
            with torch.cuda.graph(g, stream=s):
                self.iterate()
        torch.cuda.current_stream().wait_stream(s)
        return g

    # ---------------------------------------------------------------- results (in the caller's order)
    def results(self) -> List[Dict[str, np.ndarray]]:
        out = [None] * len(self.specs)
        for k, i in enumerate(self.order):
            b = self.bufs[k]
            out[i] = {"mean": b.mean.cpu().numpy(), "std": b.std.cpu().numpy(), "best": b.best.cpu().numpy(),
                      "best_cost": float(b.best_cost.item()), "cost": b.cost.cpu().numpy()}
        return out

    # ---------------------------------------------------------------- autotuner
    def autotune(self, candidates: Sequence[Tuple[int, int]] = ((64, 1), (128, 1), (256, 1), (64, 2), (128, 2),
                                                                (256, 2)), repeats: int = 2) -> Dict:
        """Time one rollout phase per (envs per block, physics-step unroll) and keep the fastest."""
        self.sample()
        times = {}
        for blk, un in candidates:
            self.set_launch(blk, un)
            self.rollout()                      # warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(repeats):
                self.rollout()
            torch.cuda.synchronize()
            times[(blk, un)] = (time.perf_counter() - t0) / repeats
        best = min(times, key=times.get)
        self.set_launch(*best)
        return {"best": best, "seconds": {f"{b}x{u}": round(t, 5) for (b, u), t in times.items()}}
