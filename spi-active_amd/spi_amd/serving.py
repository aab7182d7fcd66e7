"""Continuous-batching service for sysID and excitation-scoring jobs on one GPU.

* sysID jobs (a CEM problem: segments, horizon, sample count, iterations) are admitted as
  soon as they have arrived and their device footprint fits the HBM budget; every tick runs
  one CEM iteration of all admitted jobs in one batched launch per phase (agents/batch.py),
  so a job that arrives while others are running joins at the next iteration boundary and
  leaves when its own iteration count is reached.  CEM state stays on the device for the
  whole life of a job; only membership changes rebuild the (small) problem table.
* excitation-scoring jobs (logdet of the exact Fisher information of candidate excitations,
  agents/fisher.py; or the reference's finite-difference trace, agents/active.py) run on a
  second stream and overlap with the CEM iteration.

A job's result does not depend on the other jobs in flight: each problem has its own Philox
stream, iteration counter and buffers, and the batched kernels never mix problems.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from .agents.active import FDFIMObjective
from .agents.batch import BatchCEM, ProblemSpec, _Buffers, validate_spec
from .agents.fisher import IDENTIFIED, FisherObjective
from .params import P_COUNT
from .rollout import RolloutEngine


@dataclass
class SysIdRequest:
    job_id: int
    arrival: float                 # seconds after the service started
    spec: ProblemSpec
    iterations: int = 12


@dataclass
class ScoreRequest:
    job_id: int
    arrival: float
    actions: torch.Tensor          # [M, T, 12] candidate excitations
    init: torch.Tensor             # [37] or [M * (P + 1), 37]
    theta: np.ndarray              # [25] parameters the excitation is scored at
    explore: Optional[Sequence[str]] = None     # default: the 22 identified parameters
    mode: str = "exact"            # "exact": logdet of the exact Fisher information;
    delta: float = 0.1             # "fd": the reference's finite-difference FIM trace


@dataclass
class _Active:
    req: SysIdRequest
    buf: _Buffers
    done: int = 0


@dataclass
class ServiceStats:
    latency_s: Dict[int, float] = field(default_factory=dict)
    results: Dict[int, Dict] = field(default_factory=dict)
    ticks: int = 0
    max_active: int = 0
    wall_s: float = 0.0
    host_s: float = 0.0            # host time spent admitting, rebuilding tables and launching
    wait_s: float = 0.0            # host time blocked on the previous tick

    def summary(self) -> Dict:
        lat = np.array(list(self.latency_s.values())) if self.latency_s else np.zeros(1)
        return {"completed": len(self.latency_s), "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "jobs_per_s": len(self.latency_s) / max(self.wall_s, 1e-9), "ticks": self.ticks,
                "max_active": self.max_active, "wall_s": self.wall_s, "host_s": self.host_s, "wait_s": self.wait_s}


class SysIdServer:
    """Continuous batching on one GPU.

    The host runs one tick ahead of the device: it launches CEM iteration k + 1 of every
    admitted job while iteration k still runs, and it never synchronises the device as a whole.
    A job that reaches its iteration count leaves the batch at once (its results are copied to
    pinned host memory behind an event), which frees its budget for waiting jobs; scoring jobs
    run on their own stream.  Completed copies are polled every tick."""

    def __init__(self, engine: RolloutEngine, hbm_budget_bytes: float = 200e9, max_active: int = 4096,
                 block: int = 256, unroll: int = 1):
        self.engine = engine
        self.budget = float(hbm_budget_bytes)
        self.max_active = max_active
        self.block, self.unroll = block, unroll
        self.score_stream = torch.cuda.Stream()   # default priority, like every stream here

    def run(self, requests: Sequence, clock: str = "wall") -> ServiceStats:
        """Serve requests until all are done.

        clock="wall": ``arrival`` is in seconds since the call; clock="ticks": ``arrival`` counts
        CEM iterations of the service (deterministic membership, for tests and replays).  Latency
        runs from the moment a request has arrived (queueing for memory included) to the moment
        the host sees its result in host memory."""
        if clock not in ("wall", "ticks"):
            raise ValueError(f"clock must be 'wall' or 'ticks', not {clock!r}")
        realtime = clock == "wall"
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            if isinstance(r, SysIdRequest):
                validate_spec(r.spec)
                if _Buffers.footprint(r.spec) > self.budget:
                    raise ValueError(f"job {r.job_id} needs {_Buffers.footprint(r.spec)} B, over the whole budget")
        # device copies of the scoring parameters, made once before serving starts (a pageable
        # upload inside the loop would wait for the scoring stream)
        self._theta_dev = {}
        for r in pending:
            if isinstance(r, ScoreRequest) and id(r.theta) not in self._theta_dev:
                self._theta_dev[id(r.theta)] = torch.as_tensor(np.asarray(r.theta, np.float32).reshape(-1, 1),
                                                               device=self.engine.device).contiguous()
        stats = ServiceStats()
        main = torch.cuda.current_stream()
        self.score_stream.wait_stream(main)       # inputs of every request are ready
        active: List[_Active] = []
        inflight: List = []            # (entries, event, pinned host tensor); entry = (job_id, kind, offset, size, dtype)
        arrived: Dict[int, float] = {}
        batch: Optional[BatchCEM] = None
        prev_tick: Optional[torch.cuda.Event] = None
        used = 0.0
        t0 = time.perf_counter()
        i = 0

        def ship(parts, entries, stream, dtype=None):
            """One device concatenation and one pinned copy for all results of a tick, ordered on
            the stream that produced them."""
            with torch.cuda.stream(stream):
                dev = torch.cat([x.reshape(-1) if dtype is None else x.reshape(-1).to(dtype) for x in parts])
                host = torch.empty(dev.shape, dtype=dev.dtype, pin_memory=True)
                host.copy_(dev, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(stream)
            inflight.append((entries, ev, host))

        def finish_ready(block: bool):
            keep = []
            for entries, ev, host in inflight:
                if block:
                    ev.synchronize()
                if not ev.query():
                    keep.append((entries, ev, host))
                    continue
                h = host.numpy()
                t_done = time.perf_counter() - t0
                for job_id, kind, o, n, dt in entries:
                    v = h[o:o + n]
                    if kind == "cem":
                        res = {"mean": v[0:P_COUNT].copy(), "std": v[P_COUNT:2 * P_COUNT].copy(),
                               "best": v[2 * P_COUNT:3 * P_COUNT].copy(), "best_cost": float(v[3 * P_COUNT])}
                    else:
                        res = {"score": v.astype(dt)}
                    stats.results[job_id] = res
                    stats.latency_s[job_id] = t_done - arrived[job_id]
            inflight[:] = keep

        while i < len(pending) or active or inflight:
            th0 = time.perf_counter()
            now = th0 - t0
            t_adm = now if realtime else stats.ticks
            for r in pending[i:]:
                if r.arrival > t_adm:
                    break
                arrived.setdefault(r.job_id, r.arrival if realtime else now)
            changed = False
            score_parts, score_entries, so = [], [], 0
            while i < len(pending) and pending[i].arrival <= t_adm:      # admit in arrival order
                r = pending[i]
                if isinstance(r, ScoreRequest):
                    out = self._launch_score(r)
                    score_parts.append(out)
                    score_entries.append((r.job_id, "score", so, out.numel(), np.dtype(str(out.dtype).split(".")[-1])))
                    so += out.numel()
                    i += 1
                    continue
                need = _Buffers.footprint(r.spec)
                if used + need > self.budget or len(active) >= self.max_active:
                    break
                active.append(_Active(r, _Buffers(r.spec, self.engine.device)))
                used += need
                changed = True
                i += 1
            if score_parts:
                ship(score_parts, score_entries, self.score_stream, torch.float64)
            if active:
                if changed or batch is None:
                    batch = BatchCEM(self.engine, [a.req.spec for a in active], block=self.block,
                                     unroll=self.unroll, buffers=[a.buf for a in active],
                                     iterations_done=[a.done for a in active], validate=False)
                batch.iterate()
                tick = torch.cuda.Event()
                tick.record(main)
                for a in active:
                    a.done += 1
                stats.ticks += 1
                stats.max_active = max(stats.max_active, len(active))
                keep, parts, entries = [], [], []
                for a in active:
                    if a.done >= a.req.iterations:
                        b = a.buf
                        parts += [b.mean, b.std, b.best, b.best_cost]
                        entries.append((a.req.job_id, "cem", len(entries) * (3 * P_COUNT + 1), 3 * P_COUNT + 1,
                                        np.float32))
                        used -= _Buffers.footprint(a.req.spec)
                    else:
                        keep.append(a)
                if parts:
                    ship(parts, entries, main)
                    batch = None
                active = keep
                # one tick of run-ahead: wait for the previous tick before preparing the next
                tw = time.perf_counter()
                stats.host_s += tw - th0
                if prev_tick is not None:
                    prev_tick.synchronize()
                stats.wait_s += time.perf_counter() - tw
                prev_tick = tick
            finish_ready(block=not active and i >= len(pending))
            if not active and i < len(pending):
                if realtime:
                    wait = pending[i].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.002))
                    finish_ready(block=False)
                else:
                    if inflight:
                        finish_ready(block=True)
                    stats.ticks = max(stats.ticks, int(math.ceil(pending[i].arrival)))
        torch.cuda.synchronize()
        stats.wall_s = time.perf_counter() - t0
        return stats

    def _launch_score(self, r: ScoreRequest) -> torch.Tensor:
        # request tensors were made before run() started; run() orders the scoring stream
        # after them once, so scoring overlaps with the CEM iterations on the main stream
        with torch.cuda.stream(self.score_stream):
            if r.mode == "exact":
                obj = FisherObjective(self.engine, list(r.explore) if r.explore else IDENTIFIED)
                th = self._theta_dev.get(id(r.theta))
                if th is None:
                    th = torch.as_tensor(np.asarray(r.theta, np.float32).reshape(-1, 1),
                                         device=self.engine.device).contiguous()
                out = obj.score(th, r.actions, r.init).logdet
            elif r.mode == "fd":
                obj = FDFIMObjective(self.engine, r.theta, list(r.explore or ("mass",)), delta=r.delta)
                out = obj(r.actions, r.init)
            else:
                raise ValueError(f"unknown scoring mode {r.mode!r}")
        return out


def poisson_arrivals(n: int, rate_hz: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_hz, size=n))
