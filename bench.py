#!/usr/bin/env python3
"""Benchmark: 22-parameter CEM system identification of the Go2 on one MI355X.

Workload (one "step" = one full sysID solve, inputs resident in HBM):
  4096 parameter samples / iteration, 410 elites (ceil 10 %), 12 iterations, then 10
  Levenberg-Marquardt iterations with exact residual Jacobians; 256 segments of horizon 20
  control steps (dt 0.005 s, decimation 4, 4 integration sub-intervals per physics step),
  PD Kp 20 / Kd 0.5, penalty contact k 5e4 N/m with saturating damping c 100 N s/m, mu 0.8.  The segments are the
  committed fixture tests/golden/c1_fp64_minimiser.npz: a seeded ground truth simulated by the
  FP64 oracle with sinusoidal PD targets (amp U[0.1, 0.4] rad, f U[0.5, 3] Hz) and measurement
  noise (sigma_q 0.01 rad, sigma_v 0.05 rad/s), together with the FP64 least-squares
  minimiser of the same objective, against which the identified parameters are scored.
  Per solve: 4096 * 256 * 20 * 12 = 251,658,240 control steps.

Prints one JSON line (metric, value = control steps per second over the timed solves); extra
keys carry the other configurations in short runs: the 1-D landscape's latency, one batch of
64 uneven problems, a 2,000-request serving stream and the Fisher objective's time.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spi-active_amd"))

import numpy as np  # noqa: E402

METRIC = "env-steps/s, sysID solves/s at ≤1% param error; p50/p99 solve latency"

# Algorithmic FLOPs per control step of the fused rollout (PD + motor model + 4 x 4 ABA
# sub-intervals + contact + integration + tracking cost), from the census program
# spi-active_amd/csrc/spi_flop_census.cpp (counting scalar over the kernel's own code).
FLOPS_PER_CONTROL_STEP = 84690.6
PEAK_FP32_TFLOPS = 157.3          # MI355X published FP32 (vector) peak
# HBM bytes per rollout_cost_kernel launch of this workload: rocprofv3 --pmc FETCH_SIZE
# (2,359.5 KiB, x 2.000 from a 256 MiB copy's reading) + WRITE_SIZE (4,096.0 KiB, x 1.000).
TRAFFIC_BYTES_PER_LAUNCH = 9026560

N_SAMPLES, N_SEG, HORIZON, ITERS = 4096, 256, 20, 12
FIXTURE = os.path.join(ROOT, "tests", "golden", "c1_fp64_minimiser.npz")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-candidates", type=int, default=256, help="oracle CPU engine: candidates timed")
    p.add_argument("--no-plain-baseline", action="store_true")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the landscape, batched-problems, serving and Fisher keys")
    p.add_argument("--refine", type=int, default=10,
                   help="Levenberg-Marquardt iterations after CEM, inside each timed solve (0: CEM only)")
    p.add_argument("--workload", default="cem22",
                   choices=["cem22", "landscape", "batch64", "fisher", "gradient", "serve", "store", "commands",
                            "study", "policy"],
                   help="cem22 (default): the 22-parameter CEM solve; the others are the plain-method landscape, "
                        "64 batched uneven problems, Fisher scoring of excitations, the gradient, the serving "
                        "loop, dataset windows, command sequences, the ask/tell study and the policy-in-the-loop rollout")
    p.add_argument("--requests", type=int, default=10000, help="serve: number of requests")
    p.add_argument("--policy-batches", default="1024,8192",
                   help="policy: trials per iteration at the reference batch[,wide batch]")
    p.add_argument("--obs-layout", default="ring", choices=["ring", "dense"], help="policy: observation layout")
    p.add_argument("--grad-shape", default="256,20",
                   help="gradient: segments,horizon of the rollouts differentiated (4096 samples each)")
    p.add_argument("--profile-launch", action="store_true",
                   help="run one rollout_cost launch of the workload and a 256 MiB device copy (for rocprofv3 "
                        "--pmc FETCH_SIZE / WRITE_SIZE calibration), then exit")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(candidates x segments would shard across GPUs with one cost all-reduce per CEM iteration)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if args.workload != "cem22":
        if not torch.cuda.is_available():
            raise SystemExit("bench.py needs a GPU")
        from spi_amd import workloads
        if args.workload == "serve":
            out = workloads.serve(n_requests=args.requests)
        elif args.workload == "gradient":
            n_seg, horizon = (int(x) for x in args.grad_shape.split(","))
            out = workloads.gradient(args.steps, args.warmup, n_seg=n_seg, horizon=horizon)
        elif args.workload == "policy":
            bs = [int(x) for x in args.policy_batches.split(",")]
            out = workloads.policy(args.steps, args.warmup, batch=bs[0], wide_batch=bs[1] if len(bs) > 1 else 0,
                                   obs_layout=args.obs_layout)
        else:
            out = getattr(workloads, args.workload)(args.steps, args.warmup)
        print(json.dumps(out))
        return
    from spi_amd.agents.cem import CEMConfig, DeviceCEM
    from spi_amd.agents.refine import GaussNewtonRefiner
    from spi_amd.agents.sysid import SysIdProblem, identification_prior
    from spi_amd.data import SegmentSet
    from spi_amd.model import load_model
    from spi_amd.params import PARAM_NAMES, SimParams
    from spi_amd.rollout import RolloutEngine

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    dev = torch.device("cuda:0")
    model = load_model()
    sim = SimParams()                      # Kp 20, Kd 0.5, k 5e4, c 100, mu 0.8, dt 5 ms, decimation 4
    prior = identification_prior()
    with np.load(FIXTURE, allow_pickle=False) as fx:
        seg_np, theta_true = fx["segments"], fx["truth"]
        se_rel, cost_indep = fx["se_rel"], float(fx["cost_fp64"])
        # the oracle's FP64 run of this pipeline (same seed, samples and update rule) when the
        # fixture holds it, else the independent FP64 search's minimiser
        same_seed = "theta_fp64_same_seed" in fx.files
        theta64 = fx["theta_fp64_same_seed"] if same_seed else fx["theta_fp64"]
        cost64 = float(fx["cost_fp64_same_seed"]) if same_seed else cost_indep
        cem64 = (fx["cem_means_same_seed"][-1], fx["cem_stds_same_seed"][-1]) if same_seed else None
        # spread of the FP64 local minimisers reached from independent starts
        starts = [fx["theta_fp64_truth_start"], fx["theta_fp64_cem_start"]] + ([theta64] if same_seed else [])
        pts = np.stack(starts)
        spread64 = (pts.max(0) - pts.min(0)) / np.maximum(np.abs(pts).min(0), 1e-12)
    segs = SegmentSet.from_packed(seg_np, HORIZON)
    eng = RolloutEngine(sim)
    seg_dev = torch.from_numpy(seg_np).to(dev)
    problem = SysIdProblem(eng, seg_dev, HORIZON, max_cand=N_SAMPLES)
    refiner = GaussNewtonRefiner(eng, seg_dev, HORIZON, lo=prior.lo, hi=prior.hi) if args.refine > 0 else None
    stream = torch.cuda.current_stream()

    if args.profile_launch:
        cem = DeviceCEM(prior.mean, prior.std, prior.lo, prior.hi, prior.min_std, CEMConfig(N_SAMPLES, 0.1, ITERS))
        theta = cem.ask()
        eng.segment_costs(theta, N_SAMPLES, seg_dev, HORIZON, out=problem._seg_cost)
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)      # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"profile_launch": "done", "copy_bytes": src.numel() * 4}))
        return

    kernel_events = []

    def evaluate_timed(theta, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        eng.segment_costs(theta, n, seg_dev, HORIZON, out=problem._seg_cost)
        e1.record(stream)
        kernel_events.append((e0, e1))
        return eng.reduce(problem._seg_cost, problem.seg_weights, out=problem._cost[:n])

    def solve(seed, timed):
        """One sysID solve: CEM on the device, then (args.refine > 0) Levenberg-Marquardt from
        the CEM mean.  Returns (CEM mean, refined parameters or None, refinement record)."""
        cem = DeviceCEM(prior.mean, prior.std, prior.lo, prior.hi, prior.min_std,
                        CEMConfig(N_SAMPLES, 0.1, ITERS, seed=seed), device="cuda")
        cem.solve(evaluate_timed if timed else problem)
        if refiner is None:
            return cem.mean, None, None
        res = refiner.refine(cem.mean.cpu().numpy().astype(np.float64), iterations=args.refine)
        return cem.mean, res.theta, res

    for w in range(args.warmup):
        solve(1000 + w, timed=False)
    torch.cuda.synchronize()

    solve_ms = []
    t0 = time.perf_counter()
    outs = []
    for k in range(args.steps):
        t_s = time.perf_counter()
        outs.append(solve(k, timed=True))
        torch.cuda.synchronize()
        solve_ms.append((time.perf_counter() - t_s) * 1e3)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0

    # the segment set handed over from host memory (pageable numpy, as a caller holds it): the
    # PCIe-inclusive rate adds one upload per solve; `value` starts with the segments in HBM
    up_ms = []
    for _ in range(5):
        torch.cuda.synchronize()
        t_u = time.perf_counter()
        eng.upload_segments(segs)
        torch.cuda.synchronize()
        up_ms.append((time.perf_counter() - t_u) * 1e3)
    up_s = float(np.median(up_ms)) / 1e3

    lat = np.array(solve_ms)
    kern = np.array([a.elapsed_time(b) for a, b in kernel_events])
    ctrl_per_solve = N_SAMPLES * N_SEG * HORIZON * ITERS
    value = args.steps * ctrl_per_solve / total

    free = prior.free
    # determined to 1 % by the data: twice the asymptotic standard error of the FP64 minimiser
    # below 1 %, and the FP64 local minimisers from independent starts within 1 % of each other
    # (the objective is piecewise smooth; its creases hold local minima the standard error
    # does not see)
    crb = [i for i in range(len(free)) if free[i] and 2 * se_rel[i] < 0.01]
    ident = [i for i in crb if spread64[i] < 0.01]

    def errors(est, ref):
        rel = np.abs(np.asarray(est, np.float64) - ref) / np.maximum(np.abs(ref), 1e-12)
        return rel, {PARAM_NAMES[i]: round(float(rel[i]), 5) for i in range(len(rel)) if free[i]}

    cem_mean, refined, res = outs[0]                # seed 0, the seed of the FP64 run
    est = refined if refined is not None else cem_mean.cpu().numpy().astype(np.float64)
    rel64, per64 = errors(est, theta64)
    rel_cem64, per_cem64 = errors(cem_mean.cpu().numpy(), theta64)
    rel_true, per_true = errors(est, theta_true)
    within = [PARAM_NAMES[i] for i in ident if rel64[i] <= 0.01]
    accuracy = {
        "reference": ("the FP64 oracle's run of the same pipeline on the same data (CEM with this seed, "
                      "4096 samples and update rule, then FP64 LM; tests/golden/c1_fp64_minimiser.npz)"
                      if same_seed else
                      "FP64 least-squares minimiser of the same objective on the same data (independent FP64 "
                      "searches, tests/golden/c1_fp64_minimiser.npz)"),
        "reference_cost_fp64": cost64,
        "independent_fp64_search_cost": cost_indep,
        "device_cost_at_answer": (round(res.cost, 7) if res is not None else None),
        "identifiable": [PARAM_NAMES[i] for i in ident],
        "identifiable_rule": ("twice the asymptotic standard error of the FP64 minimiser < 1 % of its value, "
                              "and the FP64 local minimisers from independent starts within 1 % of each other"),
        "within_1pct": within, "all_identifiable_within_1pct": len(within) == len(ident),
        "max_rel_identifiable": round(float(max(rel64[i] for i in ident)), 5),
        "standard_error_rule_only": {PARAM_NAMES[i]: {"rel_err": round(float(rel64[i]), 5),
                                                      "fp64_minima_spread": round(float(spread64[i]), 5)}
                                     for i in crb if i not in ident},
        "solve": "CEM 12 iterations + LM %d iterations" % args.refine if refiner is not None else "CEM 12 iterations",
        "per_param": per64,
        "cem_only_per_param": per_cem64,
        # the CEM alone against the FP64 CEM of the same seed: the mean's distance in units of
        # the FP64 search's final width (population std), per free parameter, at most
        "cem_mean_vs_fp64_cem_in_std": (round(float(np.max(np.abs(cem_mean.cpu().numpy().astype(np.float64)
                                                                   - cem64[0])[free] / cem64[1][free])), 4)
                                        if cem64 is not None else None),
        "se_rel_fp64": {PARAM_NAMES[i]: float("%.3g" % se_rel[i]) for i in range(len(se_rel)) if free[i]},
    }
    refined_rec = None
    if res is not None:
        refined_rec = {"iterations": res.iterations, "accepted": res.accepted, "cost_before": res.history[0],
                       "cost_after": res.cost, "line_search_point_taken": res.taken}

    launch_ctrl = N_SAMPLES * N_SEG * HORIZON
    kern_s = float(kern.mean()) / 1e3
    achieved = FLOPS_PER_CONTROL_STEP * launch_ctrl / kern_s / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": TRAFFIC_BYTES_PER_LAUNCH,
                "kernel": "rollout_cost_kernel", "kernel_ms_avg": round(float(kern.mean()), 3),
                "flops_per_control_step": FLOPS_PER_CONTROL_STEP, "control_steps_per_launch": launch_ctrl}

    plain = None
    if not args.no_plain_baseline:
        plain = plain_gpu_baseline(model, sim, segs, prior, eng.weights)
        plain["fused_speedup"] = round(value / plain["value"], 1)

    cpu = cpu_oracle = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sim, seg_np, prior, eng.weights)
        cpu_oracle = cpu_baseline_oracle(model, sim, {"init": segs.init, "actions": segs.actions,
                                                      "targets": segs.targets}, prior, args.cpu_candidates)
        for c in (cpu, cpu_oracle):
            if c.get("value"):
                c["gpu_speedup"] = round(value / c["value"], 1)

    extras = {}
    if not args.no_extras:
        from spi_amd import workloads
        try:
            ls = workloads.landscape(steps=200, warmup=5)
            extras["landscape_mass_256x64_h5"] = {"solve_latency_ms": ls["solve_latency_ms"],
                                                  "solves_per_s": ls["solves_per_s"],
                                                  "argmin_mass_kg": ls["argmin_mass_kg"],
                                                  "cpu_baseline": ls["cpu_baseline"]}
        except Exception as e:  # reported, never fatal
            extras["landscape_mass_256x64_h5"] = {"failed": str(e)}
        try:
            b64 = workloads.batch64(steps=1, warmup=0, autotune=False)
            extras["batch64_uneven"] = {"seconds": round(b64["ms_per_step"] / 1e3, 3), "env_steps_per_s": b64["value"],
                                        "solves_per_s": b64["solves_per_s"],
                                        "control_steps_per_batch": b64["config"]["control_steps_per_batch"]}
        except Exception as e:
            extras["batch64_uneven"] = {"failed": str(e)}
        try:
            sv = workloads.serve(n_requests=2000)
            extras["serve_poisson_200hz_2000"] = {"jobs_per_s": sv["value"], "latency_ms": sv["latency_ms"],
                                                  "sysid_solves_per_s": sv["sysid_solves_per_s"],
                                                  "max_active": sv["max_active"]}
        except Exception as e:
            extras["serve_poisson_200hz_2000"] = {"failed": str(e)}
        try:
            fi = workloads.fisher(steps=1, warmup=1)
            extras["fisher_22x4096x128_h50"] = {"objective_ms": fi["ms_per_step"], "phase_ms": fi["phase_ms"],
                                                "excitations_per_s": fi["excitations_per_s"],
                                                "not_positive_definite": fi["not_positive_definite"]}
        except Exception as e:
            extras["fisher_22x4096x128_h50"] = {"failed": str(e)}

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "env-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(total / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (FP64-oracle ground truth + noise, committed fixture)",
        "config": {"workload": "cem22_go2", "samples": N_SAMPLES, "elites": CEMConfig(N_SAMPLES).n_elite,
                   "iterations": ITERS, "refine_iterations": args.refine, "segments": N_SEG, "horizon": HORIZON,
                   "params": int(free.sum()), "dt": sim.dt, "decimation": sim.decimation,
                   "substeps": sim.substeps, "control_steps_per_solve": ctrl_per_solve},
        "solves_per_s": round(args.steps / total, 4),
        "host_upload": {"bytes": int(seg_np.nbytes), "ms_p50": round(up_s * 1e3, 3),
                        "value_pcie_inclusive": round(args.steps * ctrl_per_solve / (total + args.steps * up_s), 1)},
        "solve_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                             "p99": round(float(np.percentile(lat, 99)), 2)},
        "param_error_vs_fp64": accuracy,
        "param_error_vs_truth": {"max_rel_identifiable": round(float(max(rel_true[i] for i in ident)), 5),
                                 "per_param": per_true},
        "refined": refined_rec,
        "roofline": roofline,
        "plain_gpu_baseline": plain,
        "cpu_baseline": cpu,
        "cpu_baseline_oracle": cpu_oracle,
    }
    out.update(extras)
    print(json.dumps(out))


def plain_gpu_baseline(model, sim, segs, prior, weights):
    """The same rollouts as PyTorch eager ops (Python loop over control steps, decimation and
    sub-intervals; spi_amd.baselines.plain_torch), timed over one control step for a few lane
    counts and both ways of forming the 6x6 products; the best rate is reported."""
    import torch
    from spi_amd.baselines.plain_torch import PlainTorchRollout
    rng = np.random.default_rng(0)
    rows = np.repeat(prior.mean[None], N_SAMPLES, axis=0)
    rows[:, prior.free] += rng.normal(size=(N_SAMPLES, int(prior.free.sum()))) * prior.std[prior.free]
    theta = torch.tensor(np.ascontiguousarray(np.clip(rows, prior.lo, prior.hi).T), dtype=torch.float32,
                         device="cuda")
    f = lambda x: torch.tensor(x, dtype=torch.float32, device="cuda")
    init, acts, tgts = f(segs.init), f(segs.actions), f(segs.targets)
    tried = {}
    for products in ("bcast", "bmm"):
        pl = PlainTorchRollout(sim, model, weights=weights, device="cuda", products=products)
        pl.segment_costs(theta[:, :64], init[:2], acts[:2], tgts[:2], steps=1)          # warm
        for n_seg in (4, 32, 256):
            torch.cuda.synchronize()
            t = time.perf_counter()
            pl.segment_costs(theta, init[:n_seg], acts[:n_seg], tgts[:n_seg], steps=1)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            tried[f"{products}_{N_SAMPLES * n_seg}_lanes"] = round(N_SAMPLES * n_seg / dt, 1)
            if dt > 8.0:
                break
        del pl
        torch.cuda.empty_cache()
    best = max(tried, key=tried.get)
    return {"value": tried[best], "unit": "env-steps/s", "best": best, "tried": tried,
            "sample": "1 control step (4 physics steps x 4 sub-intervals) per configuration"}


def host_cores():
    """Cores this process may use: the host's CPU share as the environment sets it for OpenMP
    (OMP_NUM_THREADS), else every core of the affinity mask; and the size of that mask."""
    visible = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "").strip()
    return (int(env) if env.isdigit() and int(env) > 0 else visible), visible


def _rows_from_prior(prior, n, seed=0):
    rng = np.random.default_rng(seed)
    rows = np.repeat(prior.mean[None], n, axis=0)
    rows[:, prior.free] += rng.normal(size=(n, int(prior.free.sum()))) * prior.std[prior.free]
    return np.clip(rows, prior.lo, prior.hi).astype(np.float32)


def cpu_baseline(sim, seg_packed, prior, weights, target_s=10.0):
    """The GPU kernel's algorithm on the host: spi_dyn.h built for x86-64-v3 with OpenMP
    (tests/host_dyn/spi_rollout_cpu.cpp), over all the cores this process may use, on a bounded
    sample of the workload's candidates (all 256 segments x H 20, about target_s seconds)."""
    import ctypes
    try:
        from spi_amd import _lib
        lib_path = os.path.join(ROOT, "tests", "host_dyn", "lib", "libspi_rollout_cpu.so")
        if not os.path.exists(lib_path):
            import subprocess
            subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "host_dyn"), "-s"], check=True)
        L = ctypes.CDLL(lib_path)
        fp = ctypes.POINTER(ctypes.c_float)
        L.spi_cpu_rollout_cost.argtypes = [ctypes.c_void_p, ctypes.c_void_p, fp, ctypes.c_int, ctypes.c_int, fp,
                                           ctypes.c_int, ctypes.c_int, fp, ctypes.c_int]
        cores, visible = host_cores()
        a, c = _lib.sim_args(sim), _lib.cost_args(0, weights)
        seg = np.ascontiguousarray(seg_packed, np.float32)
        n_seg = seg.shape[0]

        def run(n):
            par = np.ascontiguousarray(_rows_from_prior(prior, n).T)
            out = np.zeros((n_seg, n), np.float32)
            t = time.perf_counter()
            used = L.spi_cpu_rollout_cost(ctypes.byref(a), ctypes.byref(c), par.ctypes.data_as(fp), n, n,
                                          seg.ctypes.data_as(fp), n_seg, HORIZON, out.ctypes.data_as(fp), cores)
            return time.perf_counter() - t, used

        dt, _ = run(16)
        n = int(min(N_SAMPLES, max(16, 16 * target_s / max(dt, 1e-3))))
        dt, used = run(n)
        steps = n * n_seg * HORIZON
        return {"value": round(steps / dt, 1), "unit": "env-steps/s", "cores": used, "kind": "port",
                "algorithm": "same as the GPU kernel: spi_dyn.h built for the host (x86-64-v3, OpenMP), one "
                             "rollout per thread iteration",
                "cores_visible": visible,
                "sample": f"{n} of {N_SAMPLES} CEM candidates x {n_seg} segments x H={HORIZON} "
                          f"({steps} control steps, {dt:.1f} s)"}
    except Exception as e:  # the baseline is reported, never fatal
        return {"value": None, "unit": "env-steps/s", "cores": 0, "kind": "port", "sample": f"failed: {e}"}


def cpu_baseline_oracle(model, sim, segs_np, prior, n_cand):
    """The oracle's FP32 engine (generic dense spatial algebra on the unfolded 35-link URDF
    tree, OpenMP over the same cores) on a bounded sample of the workload."""
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import spio
        spio.build()
        cores, visible = host_cores()
        o = spio.Oracle(model, sim, f32=True)
        rows = _rows_from_prior(prior, n_cand)
        w = np.array([10.0, 5.0, 1.0, 1.0, 1.0, 0.05])
        t = time.perf_counter()
        o.rollout_cost_f32(rows, segs_np["init"], segs_np["actions"], segs_np["targets"], 0, w, n_threads=cores)
        dt = time.perf_counter() - t
        n_seg = segs_np["init"].shape[0]
        steps = n_cand * n_seg * HORIZON
        return {"value": round(steps / dt, 1), "unit": "env-steps/s", "cores": cores, "kind": "port",
                "cores_visible": visible,
                "sample": f"{n_cand} of {N_SAMPLES} CEM candidates x {n_seg} segments x H={HORIZON} "
                          f"({steps} control steps, {dt:.1f} s), FP32 oracle engine on the unfolded URDF tree"}
    except Exception as e:  # the baseline is reported, never fatal
        return {"value": None, "unit": "env-steps/s", "cores": 0, "kind": "port", "sample": f"failed: {e}"}


if __name__ == "__main__":
    main()
