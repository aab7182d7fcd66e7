"""Measured workloads besides the 22-parameter CEM solve (``bench.py --workload ...``).

* ``landscape`` -- plain method: base-mass landscape, 256 masses on [5, 9] kg x 64 segments
  x H 5, argmin; a latency-bound solve, replayed as one HIP graph (scripts/mass_landscape.py).
* ``batch64`` -- 64 independent CEM problems of uneven size (horizons {5, 10, 20, 50},
  samples {512, 2048, 8192}, segments {32, 128, 512}, fixed seed) in one persistent launch per
  phase, launch shape from the autotuner, iterations replayed from a HIP graph.
* ``fisher`` -- exact trajectory sensitivities of H 50 rollouts for 22 parameters x 4096
  samples x 128 Fourier excitations (8 terms per joint, U[-0.3, 0.3]), Fisher assembly on
  matrix cores, FP64 logdet per excitation.
* ``serve`` -- Poisson stream of sysID and excitation-scoring jobs through the
  continuous-batching server under an HBM budget.
* ``store`` -- dataset windows: 64 trajectories x 2500 steps packed into H 20 segment records.
* ``commands`` -- command sequences of one 1024-trial ask batch in the three sampling modes.
* ``study`` -- the Optuna-shaped ask / suggest_float / tell loop, 5 iterations of 1024 trials.

Each returns the dict ``bench.py`` prints as its JSON line.
"""
from __future__ import annotations

import time
from typing import Dict

import numpy as np
import torch

from .agents.batch import BatchCEM, ProblemSpec
from .agents.fisher import FIM_ROWS, FisherObjective, fourier_excitations
from .agents.sysid import SysIdProblem, ground_truth, identification_prior
from .data import standing_state, synthetic_segments
from .params import P_MASS, SimParams, nominal_params
from .model import load_model
from .rollout import RolloutEngine

PEAK_FP32_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def _events(stream):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _settled_state(eng: RolloutEngine, theta: np.ndarray, steps: int = 50) -> np.ndarray:
    s = eng.simulate_np(np.asarray(theta)[None], standing_state(0.30)[None], np.zeros((1, steps, 12), np.float32))
    return s[0, -1]


def _host_cores() -> int:
    """Cores this process may use: OMP_NUM_THREADS when set, else the affinity mask."""
    import os
    env = os.environ.get("OMP_NUM_THREADS", "").strip()
    return int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))


def _cpu_rollout_lib():
    """The rollout kernels' algorithm built for the host (spi_dyn.h, x86-64-v3, OpenMP:
    tests/host_dyn/spi_rollout_cpu.cpp), the same-algorithm CPU baseline."""
    import ctypes
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    path = os.path.join(root, "tests", "host_dyn", "lib", "libspi_rollout_cpu.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", os.path.join(root, "tests", "host_dyn"), "-s"], check=True)
    L = ctypes.CDLL(path)
    fp = ctypes.POINTER(ctypes.c_float)
    L.spi_cpu_rollout_cost.argtypes = [ctypes.c_void_p, ctypes.c_void_p, fp, ctypes.c_int, ctypes.c_int, fp,
                                       ctypes.c_int, ctypes.c_int, fp, ctypes.c_int]
    return L


# ------------------------------------------------------------------ plain method (1-D landscape)
def landscape(steps: int, warmup: int) -> Dict:
    sim = SimParams()
    eng = RolloutEngine(sim)
    model = load_model()
    truth = nominal_params(model)                        # base-link mass 6.921 kg
    segs = synthetic_segments(eng.simulate_np, truth, sim, n_seg=64, horizon=5, seed=0)
    seg_dev = eng.upload_segments(segs)
    masses = np.linspace(5.0, 9.0, 256)
    rows = np.repeat(truth[None], 256, axis=0)
    rows[:, P_MASS] = masses
    theta = torch.from_numpy(np.ascontiguousarray(rows.T.astype(np.float32))).cuda()
    prob = SysIdProblem(eng, seg_dev, 5, max_cand=256)
    holder = {}

    def solve():
        holder["cost"] = prob(theta, 256)

    for _ in range(max(warmup, 1)):
        solve()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            solve()
    torch.cuda.current_stream().wait_stream(s)
    for _ in range(max(warmup, 1)):                # the first replays upload the graph
        g.replay()
        torch.argmin(holder["cost"])
    torch.cuda.synchronize()
    lat = []
    st = torch.cuda.current_stream()
    t0 = time.perf_counter()
    for _ in range(steps):
        a, b = _events(st)
        a.record(st)
        g.replay()
        idx_dev = torch.argmin(holder["cost"])     # eager, outside the graph
        b.record(st)
        lat.append((a, b))
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    ms = np.array([a.elapsed_time(b) for a, b in lat])
    idx = int(idx_dev.item())
    ctrl = 256 * 64 * 5
    # the same solve by the same algorithm on the host cores (median of 3), and its argmin
    try:
        import ctypes
        from . import _lib
        L = _cpu_rollout_lib()
        fp = ctypes.POINTER(ctypes.c_float)
        a_h, c_h = _lib.sim_args(sim), _lib.cost_args(eng.cost_mode, eng.weights)
        seg = np.ascontiguousarray(segs.packed(), np.float32)
        par = np.ascontiguousarray(rows.T.astype(np.float32))
        out = np.zeros((64, 256), np.float32)
        cores, times = _host_cores(), []
        for _ in range(3):
            t0 = time.perf_counter()
            used = L.spi_cpu_rollout_cost(ctypes.byref(a_h), ctypes.byref(c_h), par.ctypes.data_as(fp), 256, 256,
                                          seg.ctypes.data_as(fp), 64, 5, out.ctypes.data_as(fp), cores)
            times.append(time.perf_counter() - t0)
        cpu_ms = float(np.median(times)) * 1e3
        cpu = {"kind": "port", "ms_per_solve": round(cpu_ms, 3), "cores": int(used),
               "solves_per_s": round(1e3 / cpu_ms, 2),
               "argmin_mass_kg": float(masses[int(np.argmin(out.astype(np.float64).sum(axis=0)))]),
               "gpu_speedup_p50": round(cpu_ms / float(np.percentile(ms, 50)), 1),
               "sample": "the whole solve (256 masses x 64 segments x H 5), median of 3"}
    except Exception as e:  # reported, never fatal
        cpu = {"value": None, "failed": str(e)}
    return {"metric": "env-steps/s, landscape solves/s; p50/p99 solve latency", "value": round(steps * ctrl / total, 1),
            "unit": "env-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": round(total / steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "landscape_mass_256x64_h5", "candidates": 256, "segments": 64, "horizon": 5,
                       "grid_kg": [5.0, 9.0], "control_steps_per_solve": ctrl, "graph": True},
            "solves_per_s": round(steps / total, 2),
            "solve_latency_ms": {"p50": round(float(np.percentile(ms, 50)), 4),
                                 "p99": round(float(np.percentile(ms, 99)), 4)},
            "argmin_mass_kg": float(masses[idx]), "truth_kg": float(truth[P_MASS]),
            "grid_step_kg": float(masses[1] - masses[0]), "cpu_baseline": cpu}


# ------------------------------------------------------------------ batched uneven problems
def batch64(steps: int, warmup: int, iterations: int = 12, n_prob: int = 64, seed: int = 0,
            autotune: bool = True) -> Dict:
    sim = SimParams()
    eng = RolloutEngine(sim)
    prior = identification_prior()
    rng = np.random.default_rng(seed)
    Hs = rng.choice([5, 10, 20, 50], size=n_prob)
    Ns = rng.choice([512, 2048, 8192], size=n_prob)
    Ss = rng.choice([32, 128, 512], size=n_prob)
    specs, truths = [], []
    for k in range(n_prob):
        truth = ground_truth(prior, seed=1000 + k)
        segs = synthetic_segments(eng.simulate_np, truth, sim, n_seg=int(Ss[k]), horizon=int(Hs[k]), seed=2000 + k)
        specs.append(ProblemSpec(eng.upload_segments(segs), int(Hs[k]), prior, n_samples=int(Ns[k]), seed=k))
        truths.append(truth)
    ctrl_per_iter = int(sum(int(n) * int(s) * int(h) for n, s, h in zip(Ns, Ss, Hs)))
    if autotune:
        tune = BatchCEM(eng, specs)
        info = tune.autotune()
        block, unroll = info["best"]
        del tune
    else:                            # the shapes lie within 1 % of each other; skip the sweep
        info = {"best": [128, 1], "seconds": None}
        block, unroll = 128, 1
    lat = []
    results = None
    st = torch.cuda.current_stream()
    t_total = 0.0
    for rep in range(warmup + steps):
        b = BatchCEM(eng, specs, block=block, unroll=unroll)      # fresh problem state per solve
        torch.cuda.synchronize()
        a, e = _events(st)
        t0 = time.perf_counter()
        a.record(st)
        b.solve(iterations, graph=True)
        e.record(st)
        torch.cuda.synchronize()
        if rep >= warmup:
            t_total += time.perf_counter() - t0
            lat.append(a.elapsed_time(e))
            results = b.results()
    mass_err = [abs(r["mean"][P_MASS] - t[P_MASS]) / t[P_MASS] for r, t in zip(results, truths)]
    ctrl = ctrl_per_iter * iterations
    return {"metric": "env-steps/s, sysID solves/s (64 problems per batch); p50/p99 batch latency",
            "value": round(steps * ctrl / t_total, 1), "unit": "env-steps/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(t_total / steps * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "batch64_uneven", "problems": n_prob, "iterations": iterations,
                       "horizons": [5, 10, 20, 50], "samples": [512, 2048, 8192], "segments": [32, 128, 512],
                       "seed": seed, "control_steps_per_batch": ctrl, "substeps_per_batch": ctrl * sim.decimation,
                       "autotune": info},
            "solves_per_s": round(steps * n_prob / t_total, 3),
            "batch_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                                 "p99": round(float(np.percentile(lat, 99)), 2)},
            "base_mass_rel_err": {"median": round(float(np.median(mass_err)), 5),
                                  "max": round(float(np.max(mass_err)), 5)}}


# ------------------------------------------------------------------ Fisher information
def fisher(steps: int, warmup: int, n_samples: int = 4096, n_exc: int = 128, horizon: int = 50) -> Dict:
    sim = SimParams()
    eng = RolloutEngine(sim)
    prior = identification_prior()
    rng = np.random.default_rng(0)
    rows = np.repeat(prior.mean[None], n_samples, axis=0)
    rows[:, prior.free] += rng.normal(size=(n_samples, int(prior.free.sum()))) * prior.std[prior.free]
    rows = np.clip(rows, prior.lo, prior.hi)
    theta = torch.from_numpy(np.ascontiguousarray(rows.T.astype(np.float32))).cuda()
    init = torch.from_numpy(_settled_state(eng, prior.mean).astype(np.float32)).cuda()
    acts = fourier_excitations(n_exc, horizon, sim.dt * sim.decimation, sim.action_scale, n_terms=8, coeff=0.3,
                               seed=0)
    obj = FisherObjective(eng, reg=1e-6, scale=prior.mean[:22])
    st = torch.cuda.current_stream()
    chunk_exc = 16
    sens_ms, asm_ms, ld_ms, totals = [], [], [], []
    scores = None
    for rep in range(warmup + steps):
        F = torch.empty((n_exc * n_samples, obj.P, obj.P), dtype=torch.float32, device="cuda")
        status = torch.empty(n_exc * n_samples, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev = []
        for e0 in range(0, n_exc, chunk_exc):
            a0, a1 = _events(st)
            b1 = torch.cuda.Event(enable_timing=True)
            a0.record(st)
            J, s_, _ = obj.sensitivities(theta, acts[e0:e0 + chunk_exc], init)
            a1.record(st)
            obj.assemble(J, out=F[e0 * n_samples:(e0 + chunk_exc) * n_samples])
            b1.record(st)
            status[e0 * n_samples:(e0 + chunk_exc) * n_samples] = s_
            ev.append((a0, a1, b1))
            del J
        c0, c1 = _events(st)
        c0.record(st)
        scores = obj.logdet(F, status, n_exc)
        c1.record(st)
        torch.cuda.synchronize()
        if rep >= warmup:
            totals.append(time.perf_counter() - t0)
            sens_ms.append(sum(a.elapsed_time(b) for a, b, _ in ev))
            asm_ms.append(sum(b.elapsed_time(c) for _, b, c in ev))
            ld_ms.append(c0.elapsed_time(c1))
    G = n_samples * n_exc
    ctrl = G * horizon
    t = float(np.mean(totals))
    asm_s = float(np.mean(asm_ms)) / 1e3
    flops_asm = 2.0 * horizon * FIM_ROWS * obj.P * obj.P * G
    bytes_j = 4.0 * G * horizon * FIM_ROWS * obj.P
    ld = scores.logdet.cpu().numpy()
    valid = scores.n_valid.cpu().numpy()
    return {"metric": "sensitivity rollouts/s (22 tangents each), excitations scored/s",
            "value": round(ctrl / t, 1), "unit": "control steps/s (x22 tangents)", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(t * 1e3, 2), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 (logdet fp64)", "data": "synthetic",
            "config": {"workload": "fisher_22x4096x128_h50", "params": obj.P, "samples": n_samples,
                       "excitations": n_exc, "horizon": horizon, "fourier_terms": 8, "coeff": 0.3,
                       "excitation_chunk": chunk_exc},
            "phase_ms": {"sensitivities": round(float(np.mean(sens_ms)), 2),
                         "fim_assemble_mfma": round(float(np.mean(asm_ms)), 3),
                         "logdet_fp64": round(float(np.mean(ld_ms)), 3)},
            "fim_assemble_roofline": {"bound": "hbm", "achieved": round(bytes_j / asm_s / 1e9, 1),
                                      "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": round(bytes_j / asm_s / 1e9 / HBM_PEAK_GBS, 4),
                                      "useful_tflops": round(flops_asm / asm_s / 1e12, 2)},
            "excitations_per_s": round(n_exc / t, 2),
            "best_excitation": int(np.argmax(ld)),
            "logdet_range": [float(ld[np.isfinite(ld)].min()), float(ld.max())],
            # excitations whose mean information matrix is not positive definite in FP64 after the
            # ridge: their sensitivities blow up (|J| ~ 1e10, a chaotic rollout), the FP32 J^T J is
            # below its own rounding there, and the excitation is scored -inf (never chosen)
            "not_positive_definite": int((~np.isfinite(ld)).sum()),
            "valid_fraction": round(float(valid.sum()) / G, 4)}


# ------------------------------------------------------------------ serving
def serve(n_requests: int = 10000, rate_hz: float = 200.0, budget_gb: float = 200.0, seed: int = 0,
          sysid_frac: float = 0.8, n_seg: int = 4, iterations: int = 8) -> Dict:
    from .serving import ScoreRequest, SysIdRequest, SysIdServer, poisson_arrivals
    sim = SimParams()
    eng = RolloutEngine(sim)
    prior = identification_prior()
    rng = np.random.default_rng(seed)
    arrivals = poisson_arrivals(n_requests, rate_hz, seed)
    # segment pools per horizon (each job draws its own truth's data from a pool of 16 truths)
    pools = {}
    init = torch.from_numpy(_settled_state(eng, prior.mean).astype(np.float32)).cuda()
    reqs, kinds, work = [], [], 0
    for k in range(n_requests):
        if rng.random() < sysid_frac:
            H = int(rng.integers(5, 51))
            n = int(np.clip(np.exp(rng.uniform(np.log(512), np.log(8192))) // 256 * 256, 512, 8192))
            key = (H, int(rng.integers(16)))
            if key not in pools:
                truth = ground_truth(prior, seed=5000 + key[1])
                segs = synthetic_segments(eng.simulate_np, truth, sim, n_seg=n_seg, horizon=H, seed=6000 + key[1])
                pools[key] = eng.upload_segments(segs)
            reqs.append(SysIdRequest(k, float(arrivals[k]), ProblemSpec(pools[key], H, prior, n_samples=n, seed=k),
                                     iterations=iterations))
            kinds.append("sysid")
            work += n * n_seg * H * iterations
        else:
            M = int(rng.integers(16, 129))
            acts = fourier_excitations(M, 50, sim.dt * sim.decimation, sim.action_scale, seed=k)
            reqs.append(ScoreRequest(k, float(arrivals[k]), acts, init, prior.mean))
            kinds.append("score")
            work += M * 50
    srv = SysIdServer(eng, hbm_budget_bytes=budget_gb * 1e9)
    stats = srv.run(reqs)
    lat = stats.latency_s
    by = {}
    for kind in ("sysid", "score"):
        v = np.array([lat[r.job_id] for r, kd in zip(reqs, kinds) if kd == kind])
        if v.size:
            by[kind] = {"n": int(v.size), "p50_ms": round(float(np.percentile(v, 50)) * 1e3, 2),
                        "p99_ms": round(float(np.percentile(v, 99)) * 1e3, 2)}
    s = stats.summary()
    n_sysid = sum(1 for kd in kinds if kd == "sysid")
    return {"metric": "serving: jobs/s, sysID solves/s, p50/p99 latency", "value": round(s["jobs_per_s"], 2),
            "unit": "jobs/s", "n_gpus": 1, "steps": n_requests, "warmup": 0,
            "ms_per_step": round(s["wall_s"] / n_requests * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "serve_poisson", "requests": n_requests, "rate_hz": rate_hz, "budget_gb": budget_gb,
                       "sysid_fraction": sysid_frac, "sysid_samples": [512, 8192], "sysid_horizon": [5, 50],
                       "sysid_segments": n_seg, "sysid_iterations": iterations, "score_candidates": [16, 128],
                       "score_horizon": 50, "seed": seed},
            "latency_ms": {"p50": round(s["p50_ms"], 2), "p99": round(s["p99_ms"], 2), "by_kind": by},
            "sysid_solves_per_s": round(n_sysid / s["wall_s"], 2), "wall_s": round(s["wall_s"], 2),
            "offered_work_control_steps": int(work), "ticks": s["ticks"], "max_active": s["max_active"],
            "host_s": round(s["host_s"], 2), "wait_s": round(s["wait_s"], 2)}


# ------------------------------------------------------------------ exact gradients
def gradient(steps: int, warmup: int, n_samples: int = 4096, n_seg: int = 256, horizon: int = 20,
             checkpoint_every: int = 10) -> Dict:
    """Gradient of the tracking cost with respect to all parameters for every (sample, segment)
    rollout of the 22-parameter CEM workload: reverse mode (checkpoints every 10 control steps;
    four lanes per rollout, the default, and one lane per rollout) timed over the full
    4096 x 256 x H 20, forward mode (22 tangent lanes) on a 1/16 slice."""
    from .agents.fisher import IDENTIFIED, _rows, adjoint_cost_grad
    from . import _lib
    from .rollout import C_ref
    sim = SimParams()
    eng = RolloutEngine(sim)
    prior = identification_prior()
    segs = synthetic_segments(eng.simulate_np, ground_truth(prior, seed=1), sim, n_seg=n_seg, horizon=horizon, seed=2)
    seg_dev = eng.upload_segments(segs)
    rng = np.random.default_rng(0)
    rows = np.repeat(prior.mean[None], n_samples, axis=0)
    rows[:, prior.free] += rng.normal(size=(n_samples, int(prior.free.sum()))) * prior.std[prior.free]
    rows = np.clip(rows, prior.lo, prior.hi)
    theta = torch.from_numpy(np.ascontiguousarray(rows.T.astype(np.float32))).cuda()
    st = torch.cuda.current_stream()

    def timed(legs: bool) -> float:
        ms = []
        for rep in range(warmup + steps):
            a, b = _events(st)
            a.record(st)
            cost, grad = adjoint_cost_grad(eng, theta, seg_dev, horizon, checkpoint_every, legs=legs)
            b.record(st)
            torch.cuda.synchronize()
            if rep >= warmup:
                ms.append(a.elapsed_time(b))
            del cost, grad
        return float(np.mean(ms))

    lane_ms = timed(False)
    legs_ms = timed(True)
    adj_ms = legs_ms                       # the default form (four lanes per rollout)
    # primal for the ratio
    a, b = _events(st)
    a.record(st)
    eng.segment_costs(theta, n_samples, seg_dev, horizon)
    b.record(st)
    torch.cuda.synchronize()
    primal_ms = a.elapsed_time(b)
    # forward mode on a slice of the segments (22 lanes per rollout)
    spec = _lib.tangent_spec(_rows(IDENTIFIED))
    n_sl = max(1, n_seg // 16)
    seg_sl = seg_dev[:n_sl].contiguous()
    c2 = torch.empty((n_sl, n_samples), dtype=torch.float32, device="cuda")
    g2 = torch.empty((n_sl, n_samples, spec.n_tan), dtype=torch.float32, device="cuda")
    a, b = _events(st)
    a.record(st)
    rc = _lib.lib().spi_rollout_cost_grad(C_ref(eng.args), C_ref(eng.cargs), C_ref(spec), _lib.ptr(theta),
                                          n_samples, n_samples, _lib.ptr(seg_sl), n_sl, horizon,
                                          _lib.ptr(c2), _lib.ptr(g2), _lib.stream())
    _lib.check(rc, "spi_rollout_cost_grad")
    b.record(st)
    torch.cuda.synchronize()
    tan_ms_full = a.elapsed_time(b) * (n_seg / n_sl)
    rollouts = n_samples * n_seg
    return {"metric": "gradient rollouts/s (d cost / d theta, 25 rows, reverse mode)",
            "value": round(rollouts / (adj_ms / 1e3), 1), "unit": "rollouts/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(adj_ms, 2), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": f"gradient_{n_samples}x{n_seg}_h{horizon}", "samples": n_samples, "segments": n_seg,
                       "horizon": horizon, "checkpoint_every": checkpoint_every},
            "adjoint_ms": round(adj_ms, 2), "adjoint_one_lane_ms": round(lane_ms, 2),
            "adjoint_four_lanes_ms": round(legs_ms, 2), "primal_ms": round(primal_ms, 2),
            "adjoint_over_primal": round(adj_ms / primal_ms, 2),
            "forward_mode_ms_extrapolated": round(tan_ms_full, 1),
            "forward_mode_over_adjoint": round(tan_ms_full / adj_ms, 2)}


def policy(steps: int, warmup: int, batch: int = 1024, wide_batch: int = 8192, total_steps: int = 1250,
           obs_layout: str = "ring") -> Dict:
    """Policy-in-the-loop active exploration (run_active_sysid.py end to end on the device): each
    iteration samples ``batch`` command trials (constant mode, dims 0, 2, 5, 5 windows), expands
    them into 1250-step command sequences, rolls out (1 main + 1 aux) env groups driven by the
    go2_omni actor (900 -> 512 -> 256 -> 128 -> 12, ELU; random weights) and updates the CEM with
    the negated FIM.  Timed at the reference's batch (1024 trials) and at a wide batch that
    fills the GPU; the actor's share is measured separately on the same rows."""
    from .agents.commands import CommandGenerator
    from .agents.policy import PolicyActiveSysId, PolicyFIMRollout, PolicyMLP
    import copy
    sim = copy.deepcopy(SimParams())
    sim.control_mode, sim.motor_model = 1, 3
    eng = RolloutEngine(sim)
    theta = nominal_params(load_model())
    ranges = [[-1.0, 1.0], [-1.0, 1.0], [-1.0, 1.0], [-0.25, 0.15], [2.0, 4.0], [0.0, 1.0], [0.0, 1.0], [0.0, 1.0],
              [0.5, 0.5], [0.03, 0.035], [-0.4, 0.4], [-0.0, 0.0], [0.1, 0.45], [0.35, 0.45]]
    default = [0.5, 0.5, 0.5, 0.0, 3.0, 0.0, 0.5, 0.5, 0.5, 0.03, 0.0, 0.0, 0.3, 0.4]
    dt_ctrl = sim.dt * sim.decimation
    gen = CommandGenerator(ranges, default, [0, 2, 5], "constant", rollout_length=total_steps * dt_ctrl,
                           horizon_length=total_steps * dt_ctrl / 5, dt=dt_ctrl)
    net = PolicyMLP.random(900, 12, seed=0)
    st = torch.cuda.current_stream()
    res = {}
    runs = [("reference_batch", batch)] + ([("wide_batch", wide_batch)] if wide_batch else [])
    for name, B in runs:
        roll = PolicyFIMRollout(eng, net, theta, ["mass"], delta=0.1, ksync_steps=5, obs_layout=obs_layout)
        drv = PolicyActiveSysId(roll, gen, batch_size=B, seed=1)
        for _ in range(warmup):
            drv.step()
        torch.cuda.synchronize()
        a, b = _events(st)
        a.record(st)
        for _ in range(steps):
            drv.step()
        b.record(st)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        n_env = B * roll.G
        T = gen.total_steps
        # actor share: the MLP alone on the same rows, replayed per control step
        obs = torch.randn((n_env, 900), device="cuda")
        out = torch.empty((n_env, 12), device="cuda")
        net(obs, out=out)
        a2, b2 = _events(st)
        a2.record(st)
        for _ in range(50):
            net(obs, out=out)
        b2.record(st)
        torch.cuda.synchronize()
        mlp_ms = a2.elapsed_time(b2) / 50
        flops = 2.0 * sum(i * o for i, o in zip(net.dims[:-1], net.dims[1:])) * n_env
        fim = drv.history[-1]
        res[name] = {"trials": B, "envs": n_env, "control_steps": T - 1, "ms_per_iteration": round(ms, 2),
                     "env_steps_per_s": round(n_env * (T - 1) / (ms / 1e3), 1),
                     "trials_per_s": round(B / (ms / 1e3), 1),
                     "actor_ms_per_control_step": round(mlp_ms, 4),
                     "actor_share": round(mlp_ms * (T - 2) / ms, 3),
                     "actor_tflops": round(flops / (mlp_ms / 1e3) / 1e12, 2),
                     "actor_frac_of_fp32_mfma_peak": round(flops / (mlp_ms / 1e3) / 1e12 / PEAK_FP32_TFLOPS, 3),
                     "finite_fraction": round(float(torch.isfinite(fim).float().mean().item()), 4),
                     "best_value": float(drv.cem.best_cost.item())}
    ref = res["reference_batch"]
    return {"metric": "policy-in-the-loop env-steps/s (active exploration, 1250-step rollouts)",
            "value": ref["env_steps_per_s"], "unit": "env-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ref["ms_per_iteration"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32", "data": "synthetic (random actor weights)",
            "config": {"workload": "policy_active_sysid", "explore": ["mass"], "total_steps": total_steps,
                       "command_mode": "constant", "ksync_steps": 5, "obs_layout": obs_layout}, **res}


# ------------------------------------------------------------------ dataset store and commands
def store(steps: int, warmup: int, n_traj: int = 64, n_steps: int = 2500, horizon: int = 20) -> Dict:
    """Dataset store (scripts/eval.py:101-171 on the device): every stride-1 window of 64
    recorded trajectories of 2500 control steps (50 s at 50 Hz) packed into H 20 segment records
    by ``spi_pack_windows_rows`` in one launch (timed with HIP events), and the store's own
    ``windows`` call around it (host bookkeeping and the row table's upload included).
    Synthetic trajectories (seeded normal values: the copy does not look at them).  Algorithmic
    bytes: the windows written once and the trajectories read once."""
    from . import _lib
    from .data import segment_stride
    from .store import SegmentStore
    rng = np.random.default_rng(0)
    st = SegmentStore()
    for _ in range(n_traj):
        st.add_arrays(rng.normal(size=(n_steps, 37)).astype(np.float32),
                      rng.normal(size=(n_steps, 12)).astype(np.float32))
    stride = segment_stride(horizon)
    per = n_steps - horizon
    n_win = n_traj * per
    out = torch.empty((n_win, stride), dtype=torch.float32, device="cuda")
    states, actions, first = st._joined_buffers()
    rows = torch.from_numpy((first[:, None] + np.arange(per)[None, :]).astype(np.int32).ravel()).cuda()
    L = _lib.lib()
    stream = torch.cuda.current_stream()
    kern_ms, wall_ms = [], []
    for rep in range(warmup + steps):
        a0, a1 = _events(stream)
        a0.record(stream)
        _lib.check(L.spi_pack_windows_rows(_lib.ptr(states), _lib.ptr(actions), states.shape[0],
                                           _lib.ptr(rows, torch.int32), n_win, horizon, _lib.ptr(out),
                                           _lib.stream()), "spi_pack_windows_rows")
        a1.record(stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch = st.windows(horizon)                 # the store's own call, with its host bookkeeping
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if rep >= warmup:
            kern_ms.append(a0.elapsed_time(a1))
            wall_ms.append((t1 - t0) * 1e3)
    assert torch.equal(batch.segments, out)
    written = 4.0 * n_win * stride
    read = 4.0 * n_traj * n_steps * (37 + 12)
    k_s = float(np.mean(kern_ms)) / 1e3
    gbs = (written + read) / k_s / 1e9
    return {"metric": "windows packed/s", "value": round(n_win / k_s, 1), "unit": "windows/s", "n_gpus": 1,
            "steps": steps, "warmup": warmup, "ms_per_step": round(float(np.mean(kern_ms)), 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (copy)",
            "data": "synthetic",
            "config": {"workload": f"store_{n_traj}x{n_steps}_h{horizon}", "trajectories": n_traj,
                       "steps_per_trajectory": n_steps, "horizon": horizon, "windows": n_win,
                       "segment_floats": stride},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": written + read},
            "store_windows_call_ms": round(float(np.mean(wall_ms)), 3)}


def commands(steps: int, warmup: int, n_trials: int = 1024) -> Dict:
    """Command sequences of the active-exploration optimizer (active_sysid.py:295-400 on the
    device): one ask batch of 1024 trials of the reference's command block
    (algo/active_sysid.yaml: 14 command dimensions, 3 sampled, 5 updates of 250 control steps),
    in each of the three sampling modes.  Algorithmic bytes: the sequences written once and the
    trial parameters read once."""
    import os
    import yaml
    from .agents.commands import CommandGenerator
    path = os.path.join(os.path.dirname(__file__), "configs", "algo", "active_sysid.yaml")
    with open(path) as f:
        cmd = yaml.safe_load(f)["algo"]["config"]["command"]
    dt = SimParams().dt * SimParams().decimation
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(0)
    modes = {}
    for mode in ("constant", "polynomial", "bezier"):
        gen = CommandGenerator(cmd["command_ranges"], cmd["default_command"], cmd["command_sampling_idxs"], mode,
                               float(cmd["rollout_length"]), float(cmd["horizon_length"]), dt)
        b = gen.param_bounds()
        P = torch.tensor(rng.uniform(b[:, 0], b[:, 1], size=(n_trials, len(b))), dtype=torch.float32,
                         device="cuda")
        ms = []
        for rep in range(warmup + steps):
            a0, a1 = _events(stream)
            a0.record(stream)
            out = gen.generate(P)
            a1.record(stream)
            torch.cuda.synchronize()
            if rep >= warmup:
                ms.append(a0.elapsed_time(a1))
        t = float(np.mean(ms)) / 1e3
        nbytes = 4.0 * out.numel() + 4.0 * P.numel()
        modes[mode] = {"ms": round(t * 1e3, 4), "trials_per_s": round(n_trials / t, 1),
                       "GB_per_s": round(nbytes / t / 1e9, 1), "frac_hbm": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4),
                       "params_per_trial": int(P.shape[1])}
    worst = max(modes.values(), key=lambda m: m["ms"])
    return {"metric": "command sequences/s", "value": worst["trials_per_s"], "unit": "trials/s (slowest mode)",
            "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": worst["ms"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": f"commands_{n_trials}x{gen.total_steps}x{gen.n_dims}", "trials": n_trials,
                       "steps": gen.total_steps, "dims": gen.n_dims, "sampled": list(gen.sampled)},
            "modes": modes}


def study(steps: int, warmup: int, batch: int = 1024, iterations: int = 5) -> Dict:
    """The Optuna-shaped study (active_sysid.py:185-217's ask / suggest_float / tell loop) over
    the device CEM at the reference's shape: 5 iterations of 1024 trials, each suggesting the 15
    floats of the constant command mode (3 sampled dimensions x 5 updates) and told a value.
    The value is a host-side quadratic in the suggestions (the policy rollout that would
    produce it is measured by ``--workload policy``), so the figure is the shim's own cost: the
    per-trial Python surface plus one sample and one update launch per iteration."""
    from .agents.study import DeviceCEMSampler, create_study
    names = [f"dim_{a}_update_{u}" for a in (0, 2, 5) for u in range(5)]
    target = np.random.default_rng(0).uniform(-0.5, 0.5, len(names))
    per_iter, best = [], []
    for rep in range(warmup + steps):
        st = create_study(direction="maximize", sampler=DeviceCEMSampler(batch_size=batch, seed=rep))
        torch.cuda.synchronize()
        times = []
        for it in range(iterations):
            t0 = time.perf_counter()
            trials = [st.ask() for _ in range(batch)]
            for t in trials:
                x = np.array([t.suggest_float(n, -1.0, 1.0) for n in names])
                st.tell(t, -float(((x - target) ** 2).sum()))
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        if rep >= warmup:
            per_iter.append(times)
            best.append(st.best_value)
    it_ms = float(np.mean(per_iter)) * 1e3
    return {"metric": "trials/s through ask, suggest_float and tell", "value": round(batch / (it_ms / 1e3), 1),
            "unit": "trials/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(it_ms, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (CEM), host floats",
            "data": "synthetic",
            "config": {"workload": f"study_{batch}x{iterations}_15params", "trials_per_iteration": batch,
                       "iterations": iterations, "params": len(names)},
            "ms_per_iteration_by_index": [round(float(x) * 1e3, 2) for x in np.mean(per_iter, axis=0)],
            "best_value_after_5_iterations": float(np.mean(best))}
