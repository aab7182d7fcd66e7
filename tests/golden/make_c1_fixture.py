#!/usr/bin/env python3
"""Writes tests/golden/c1_fp64_minimiser.npz: the 22-parameter CEM workload's data and its FP64
least-squares minimiser, both from the FP64 oracle (oracle/spi_oracle.hpp), for the
identification-parity test (tests/test_gpu_identify.py) and bench.py's param_error_vs_fp64.

Data: the seeded ground truth of the workload (identification_prior(), ground_truth(seed=1)) and
256 segments of horizon 20 cut from its trajectories with per-joint sinusoidal PD targets
(amplitude U[0.1, 0.4] rad, frequency U[0.5, 3] Hz) plus measurement noise (sigma_q 0.01 rad,
sigma_v 0.05 rad/s) on the joint channels (spi_amd.data.synthetic_segments, seed 2), every
trajectory simulated by the FP64 oracle.

Minimiser of the tracking cost mean_s sum_t sum_i w_i (x_i - target_i)^2 with the workload's
weights (10, 5, 1, 1, 1, 0.05 for pos, quat, linvel, angvel, q, qd) over the 22 identified rows
inside the prior's bounds, found entirely in FP64 on the CPU and independently of the device:
  * FP64 Levenberg-Marquardt (spio.gauss_newton_fp64: exact forward-mode Jacobians of the
    oracle's discrete rollout, the device refiner's line search over halved steps) started from
    the truth;
  * the same method the device runs, in FP64: a cross-entropy search (1024 samples per
    iteration from the oracle's Philox restatement of the device sampler, 10 % elites, 12
    iterations, the prior as the initial distribution), then the FP64 Levenberg-Marquardt from
    its mean.
The lower of the two end points is the reference minimiser (the role of the reference's argmin
in scripts/mass_opt.py:136-169 and scripts/mass_landscape.py:161-171).  The contacts make the
objective piecewise smooth: a local search can end where no step lowers the cost although the
gradient of the piece it is on is not zero; the gradient there is recorded.

The device pipeline's own run in FP64 (keys *_same_seed): the cross-entropy search with the
device's configuration and seed (4096 samples, 410 elites, 12 iterations, seed 0: the same
Philox draws as the device), its mean and std after every iteration, then the FP64
Levenberg-Marquardt from its mean.  Against it the test compares the device's run of the same
algorithm on the same inputs, FP32 against FP64.

Per-parameter asymptotic standard error of the minimiser (Gaussian residuals: the Cramer-Rao
bound of the nonlinear least-squares estimate): cov = s^2 (J^T J)^-1, s^2 = r.r / (N - P),
with r and J the weighted residuals and their Jacobian at the minimiser.

Run from the repository root: python tests/golden/make_c1_fixture.py  (about four hours on 8
cores); --same-seed-only adds the same-seed run to an existing fixture (about three hours).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "spi-active_amd"), os.path.join(ROOT, "oracle")]

import spio  # noqa: E402
from spi_amd.agents.sysid import ground_truth, identification_prior  # noqa: E402
from spi_amd.data import synthetic_segments  # noqa: E402
from spi_amd.model import load_model  # noqa: E402
from spi_amd.params import PARAM_NAMES, SimParams  # noqa: E402

N_SEG, HORIZON = 256, 20
TRUTH_SEED, DATA_SEED = 1, 2
WEIGHTS = np.array([10.0, 5.0, 1.0, 1.0, 1.0, 0.05])
GROUP = np.r_[np.zeros(3), np.ones(4), 2 * np.ones(3), 3 * np.ones(3), 4 * np.ones(12), 5 * np.ones(12)].astype(int)
OUT = os.path.join(HERE, "c1_fp64_minimiser.npz")
CEM_SAMPLES, CEM_ITERS, CEM_SEED = 1024, 12, 7
DEVICE_SAMPLES, DEVICE_SEED = 4096, 0            # the device pipeline's configuration


def fp64_cem(o, segs, prior, n_samples, seed, label, n_threads=0):
    """FP64 cross-entropy search: the device CEM's update rule (elites by (cost, index), elite
    mean, population std with the prior's floor, no smoothing) on samples from the oracle's
    restatement of the device sampler.  Returns the mean and std after every iteration."""
    mean, std = prior.mean.copy(), prior.std.copy()
    n_elite = int(np.ceil(0.1 * n_samples - 1e-9))
    means, stds = [], []
    for it in range(CEM_ITERS):
        theta = spio.cem_sample(mean, std, prior.lo, prior.hi, n_samples, seed, it)      # [25, n]
        c = o.rollout_cost_par(theta.T, segs.init, segs.actions, segs.targets, 0, WEIGHTS,
                               n_threads=n_threads).mean(axis=1)
        c = np.where(np.isfinite(c), c, np.inf)
        el = theta[:, np.lexsort((np.arange(n_samples), c))[:n_elite]]
        mean, std = el.mean(axis=1), np.maximum(el.std(axis=1), prior.min_std)
        means.append(mean.copy())
        stds.append(std.copy())
        print(f"{label} iteration {it}: best {c.min():.6g}, elite mean cost {np.sort(c)[:n_elite].mean():.6g}",
              flush=True)
    return np.array(means), np.array(stds)


def same_seed_run(o, segs, prior, lm):
    t = time.time()
    means, stds = fp64_cem(o, segs, prior, DEVICE_SAMPLES, DEVICE_SEED, "FP64 CEM (device seed)")
    th, cost, hist = lm(means[-1])
    print(f"FP64 CEM (device seed) + LM: {len(hist) - 1} accepted steps, cost {hist[0]:.6g} -> {cost:.9g} "
          f"({time.time() - t:.0f} s)", flush=True)
    return {"cem_means_same_seed": means, "cem_stds_same_seed": stds, "theta_fp64_same_seed": th,
            "cost_fp64_same_seed": cost, "history_same_seed": np.asarray(hist),
            "same_seed_config": np.array([DEVICE_SAMPLES, CEM_ITERS, DEVICE_SEED])}


def main(same_seed_only=False):
    t0 = time.time()
    model, sim = load_model(), SimParams()
    o = spio.Oracle(model, sim)
    prior = identification_prior()
    truth = ground_truth(prior, seed=TRUTH_SEED)

    def simulate(rows, init, acts):
        out = []
        for r, i0, a in zip(rows, init, acts):
            traj, rc = o.simulate(np.asarray(r, np.float64), i0, a)
            if rc != 0:
                raise RuntimeError("ground-truth trajectory diverged")
            out.append(traj)
        return np.stack(out)

    segs = synthetic_segments(simulate, truth, sim, n_seg=N_SEG, horizon=HORIZON, seed=DATA_SEED)
    print(f"data: {time.time() - t0:.1f} s", flush=True)
    rows = np.arange(22)
    seg_w = np.full(N_SEG, 1.0 / N_SEG)

    def lm(start):
        return spio.gauss_newton_fp64(o, start, rows, segs, WEIGHTS, seg_w, prior.lo, prior.hi, iterations=60,
                                      lam0=1e-3, tol=1e-14, line_search=4, grad_tol=1e-6)

    if same_seed_only:
        with np.load(OUT, allow_pickle=False) as f:
            old = {k: f[k] for k in f.files}
        if not np.array_equal(old["segments"], segs.packed()):
            raise RuntimeError(f"{OUT} holds other data than this script makes; regenerate it whole")
        old.update(same_seed_run(o, segs, prior, lm))
        np.savez_compressed(OUT, **old)
        print(f"updated {OUT} ({os.path.getsize(OUT)} bytes, {time.time() - t0:.0f} s)")
        return

    fits = {}
    t = time.time()
    th, cost, hist = lm(truth)
    fits["truth"] = (th, cost, np.asarray(hist))
    print(f"LM from truth: {len(hist) - 1} accepted steps, cost {hist[0]:.6g} -> {cost:.9g} "
          f"({time.time() - t:.0f} s)", flush=True)
    # FP64 cross-entropy search, the device CEM's update rule (elite mean, population std with
    # the prior's floor) on samples from the oracle's restatement of the device sampler
    t = time.time()
    cem_mean = fp64_cem(o, segs, prior, CEM_SAMPLES, CEM_SEED, "FP64 CEM")[0][-1]
    th, cost, hist = lm(cem_mean)
    fits["fp64_cem"] = (th, cost, np.asarray(hist))
    print(f"FP64 CEM + LM: {len(hist) - 1} accepted steps, cost {hist[0]:.6g} -> {cost:.9g} "
          f"({time.time() - t:.0f} s)", flush=True)
    best = min(fits, key=lambda k: fits[k][1])          # the lower of the two end points
    th = fits[best][0]
    print(f"reference minimiser: LM from {best}", flush=True)
    y, J, rc = o.traj_sens_batch(th, rows, segs.init, segs.actions, n_rows=37)
    sw = np.sqrt(WEIGHTS[GROUP][None, :] * seg_w[:, None])
    r = (sw[:, None, :] * (y - segs.targets.astype(np.float64))).ravel()
    Jw = (sw[:, None, :, None] * J).reshape(-1, len(rows))
    N, P = r.size, len(rows)
    s2 = float(r @ r) / (N - P)
    cov = s2 * np.linalg.inv(Jw.T @ Jw)
    se = np.zeros(25)
    se[rows] = np.sqrt(np.diag(cov))
    se_rel = se / np.maximum(np.abs(th), 1e-300)
    g = 2.0 * (Jw.T @ r)                                  # gradient of the cost r.r
    # first-order optimality: relative cost change per relative parameter change, at most
    rel_sens = np.abs(g) * np.maximum(np.abs(th[rows]), 1e-3) / (r @ r)
    other = "fp64_cem" if best == "truth" else "truth"
    print(f"param            truth          fp64 minimiser   from {other:10s} se_rel")
    for i in rows:
        print(f"{PARAM_NAMES[i]:14s} {truth[i]: .8e} {th[i]: .8e} {fits[other][0][i]: .8e} {se_rel[i]:.3e}")
    print(f"gradient norm {np.linalg.norm(g):.3e}; max relative sensitivity {rel_sens.max():.3e}; rc {rc}")
    np.savez_compressed(
        OUT, segments=segs.packed(), truth=truth, theta_fp64=th, start=np.array(best),
        theta_fp64_truth_start=fits["truth"][0], theta_fp64_cem_start=fits["fp64_cem"][0], fp64_cem_mean=cem_mean,
        cost_fp64=fits[best][1], cost_fp64_truth_start=fits["truth"][1],
        cost_fp64_cem_start=fits["fp64_cem"][1], history_truth_start=fits["truth"][2],
        history_cem_start=fits["fp64_cem"][2], cost_truth=fits["truth"][2][0], se=se, se_rel=se_rel,
        grad_at_minimiser=g, rows=rows, weights=WEIGHTS, n_seg=N_SEG, horizon=HORIZON, truth_seed=TRUTH_SEED,
        data_seed=DATA_SEED, **same_seed_run(o, segs, prior, lm))
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main(same_seed_only="--same-seed-only" in sys.argv[1:])
