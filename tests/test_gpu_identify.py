"""Identification parity at the 22-parameter CEM workload's full size (4096 samples x 12
iterations x 256 segments x H 20, then Levenberg-Marquardt with exact residual Jacobians), on
the committed FP64-oracle data (tests/golden/c1_fp64_minimiser.npz, written by
tests/golden/make_c1_fixture.py); the reference's counterpart is the argmin of
scripts/mass_opt.py:136-169 / scripts/mass_landscape.py:161-171 on identical data.

* The objective: the device's mean tracking cost equals the oracle's FP64 value at the FP64
  minimisers and at the truth.
* Same algorithm, same inputs, FP32 against FP64: the fixture holds the oracle's FP64 run of
  the device pipeline (the cross-entropy search with the device's seed, 4096 samples, the same
  Philox draws and update rule, then FP64 Levenberg-Marquardt from its mean).  The device's CEM
  mean follows the FP64 one iteration by iteration; the device's LM from the same start ends
  within 1 % of the FP64 LM's end point for every parameter that the data determine to 1 %
  (twice the asymptotic standard error of the FP64 minimiser below 1 %); so does the whole
  device pipeline.
* Against an independent FP64 search (LM from the truth, and a 1024-sample FP64 CEM with
  another seed followed by LM; the lower end point kept): the oracle's FP64 cost at the device's
  answer is within 0.5 % of that search's minimum.  The objective is piecewise smooth (a foot's
  touchdown step moves with the parameters), and searches from different starts end in
  different creases of it: the two independent FP64 end points differ by 0.5 % in cost and
  0.5-1.8 % in inertias and CoM.
"""
import os

import numpy as np
import pytest
import torch

from conftest import ROOT

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(ROOT, "tests", "golden", "c1_fp64_minimiser.npz")
H, N_SAMPLES, ITERS, SEED, LM_ITERS = 20, 4096, 12, 0, 60


@pytest.fixture(scope="module")
def c1():
    with np.load(FIXTURE, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="module")
def device_run(gpu, sim, c1):
    """The device pipeline on the fixture's segments: CEM (seed 0) with its mean and std after
    every iteration, then LM_ITERS Levenberg-Marquardt iterations from the CEM mean."""
    from spi_amd.agents.cem import CEMConfig, DeviceCEM
    from spi_amd.agents.refine import GaussNewtonRefiner
    from spi_amd.agents.sysid import SysIdProblem, identification_prior
    from spi_amd.rollout import RolloutEngine
    eng = RolloutEngine(sim)
    prior = identification_prior()
    segs = torch.from_numpy(c1["segments"]).cuda()
    prob = SysIdProblem(eng, segs, H, max_cand=N_SAMPLES)
    cem = DeviceCEM(prior.mean, prior.std, prior.lo, prior.hi, prior.min_std, CEMConfig(N_SAMPLES, 0.1, ITERS, seed=SEED))
    means, stds = [], []
    for _ in range(ITERS):
        cem.tell(prob(cem.ask(), N_SAMPLES))
        means.append(cem.mean.cpu().numpy().astype(np.float64))
        stds.append(cem.std.cpu().numpy().astype(np.float64))
    ref = GaussNewtonRefiner(eng, segs, H, lo=prior.lo, hi=prior.hi)
    res = ref.refine(means[-1], iterations=LM_ITERS)
    return {"means": np.array(means), "stds": np.array(stds), "res": res}


def test_objective_at_fp64_minimisers_matches(gpu, sim, c1):
    """The device objective (mean over segments of the tracking cost) at the FP64 minimisers and
    at the truth equals the oracle's FP64 value to 1e-4 relative."""
    from spi_amd.agents.sysid import SysIdProblem
    from spi_amd.rollout import RolloutEngine
    eng = RolloutEngine(sim)
    segs = torch.from_numpy(c1["segments"]).cuda()
    pts = [("theta_fp64", "cost_fp64"), ("theta_fp64_truth_start", "cost_fp64_truth_start"),
           ("theta_fp64_cem_start", "cost_fp64_cem_start"), ("truth", "cost_truth")]
    if "theta_fp64_same_seed" in c1:
        pts.append(("theta_fp64_same_seed", "cost_fp64_same_seed"))
    prob = SysIdProblem(eng, segs, H, max_cand=len(pts))
    th = np.stack([c1[p] for p, _ in pts]).astype(np.float32)
    got = prob(torch.from_numpy(np.ascontiguousarray(th.T)).cuda(), len(pts)).cpu().numpy().astype(np.float64)
    ref = np.array([float(c1[c]) for _, c in pts])
    print("device objective", got, "FP64", ref)
    assert np.all(np.abs(got - ref) <= 1e-4 * ref), (got, ref)


def _same_seed(c1):
    if "theta_fp64_same_seed" not in c1:
        pytest.skip("the fixture predates the FP64 same-seed run (tests/golden/make_c1_fixture.py --same-seed-only)")
    assert tuple(c1["same_seed_config"]) == (N_SAMPLES, ITERS, SEED)
    assert len(c1["history_same_seed"]) > 5, "the FP64 Levenberg-Marquardt took too few steps to be a reference"


def test_device_cem_follows_fp64_run_of_same_algorithm(device_run, c1):
    """The CEM mean after every iteration (same seed, samples, Philox draws and update rule;
    FP32 rollouts on the device, FP64 rollouts in the oracle) stays within a tenth of that
    iteration's FP64 search width (the population std) of the FP64 one."""
    _same_seed(c1)
    m64, s64 = c1["cem_means_same_seed"], c1["cem_stds_same_seed"]
    md = device_run["means"]
    free = s64[0] > 0
    drift = np.abs(md - m64)[:, free] / s64[:, free]
    print("CEM mean |device - FP64| / FP64 std per iteration (max over parameters):",
          np.array2string(drift.max(axis=1), precision=3))
    assert drift.max() <= 0.1, drift.max(axis=1)


def test_device_refinement_follows_fp64_refinement(gpu, sim, c1):
    """Levenberg-Marquardt (the step and four halvings per iteration, the damping rule of the
    oracle's FP64 LM) from the FP64 CEM mean of the same-seed run, on the device in FP32 and in
    the oracle in FP64: the device ends within 1 % of the FP64 end point for every parameter
    that the data determine to 1 % (twice the asymptotic standard error below 1 %) and at a cost
    not above the FP64 end point's by more than 1e-3."""
    from spi_amd.agents.refine import GaussNewtonRefiner
    from spi_amd.agents.sysid import identification_prior
    from spi_amd.params import PARAM_NAMES
    from spi_amd.rollout import RolloutEngine
    _same_seed(c1)
    eng = RolloutEngine(sim)
    prior = identification_prior()
    segs = torch.from_numpy(c1["segments"]).cuda()
    start = c1["cem_means_same_seed"][-1]
    res = GaussNewtonRefiner(eng, segs, H, lo=prior.lo, hi=prior.hi).refine(start, iterations=60)
    th64, se_rel, hist64 = c1["theta_fp64_same_seed"], c1["se_rel"], c1["history_same_seed"]
    ident = [i for i in range(22) if 2 * se_rel[i] < 0.01]
    err = np.abs(res.theta - th64) / np.abs(th64)
    print(f"device LM: {res.iterations} iterations, {res.accepted} accepted, cost {res.history[0]:.7g} -> "
          f"{res.cost:.7g}; FP64 LM: {len(hist64) - 1} accepted, {hist64[0]:.7g} -> {hist64[-1]:.7g}")
    for i in range(22):
        print(f"{PARAM_NAMES[i]:14s} 2se {2 * se_rel[i]:.2e} {'*' if i in ident else ' '} |device - FP64| {err[i]:.2e}")
    assert len(ident) >= 12
    bad = [(PARAM_NAMES[i], round(float(err[i]), 4)) for i in ident if err[i] > 0.01]
    assert not bad, bad
    assert res.cost <= float(c1["cost_fp64_same_seed"]) * (1 + 1e-3)


def test_device_pipeline_identifies_fp64_run_of_same_algorithm(device_run, c1):
    """The whole device pipeline (CEM seed 0, then LM from its own CEM mean) against the FP64 run
    of the same algorithm: every parameter that the data determine to 1 % within 1 %."""
    from spi_amd.params import PARAM_NAMES
    _same_seed(c1)
    res = device_run["res"]
    th64, se_rel = c1["theta_fp64_same_seed"], c1["se_rel"]
    ident = [i for i in range(22) if 2 * se_rel[i] < 0.01]
    err = np.abs(res.theta - th64) / np.abs(th64)
    print(f"device pipeline: cost {res.cost:.7g}; FP64 run of the same algorithm {float(c1['cost_fp64_same_seed']):.7g}")
    for i in range(22):
        print(f"{PARAM_NAMES[i]:14s} 2se {2 * se_rel[i]:.2e} {'*' if i in ident else ' '} |device - FP64| {err[i]:.2e}")
    bad = [(PARAM_NAMES[i], round(float(err[i]), 4)) for i in ident if err[i] > 0.01]
    assert not bad, bad


def test_device_minimum_near_independent_fp64_search(device_run, model, sim, c1):
    """The oracle evaluates the device's answer in FP64: its cost is within 0.5 % of the
    minimum the independent FP64 searches reached (their two end points lie 0.5 % apart), and
    the device's own FP32 value of it agrees with the FP64 one to 1e-4."""
    import spio
    from spi_amd.data import SegmentSet
    res = device_run["res"]
    segs = SegmentSet.from_packed(c1["segments"], H)
    o = spio.Oracle(model, sim)
    c64 = o.rollout_cost_par(np.asarray(res.theta, np.float32).astype(np.float64)[None], segs.init, segs.actions,
                             segs.targets, 0, c1["weights"]).mean()
    ref = float(c1["cost_fp64"])
    print(f"device answer: FP32 cost {res.cost:.7g}, FP64 cost {c64:.7g}; independent FP64 search {ref:.7g} "
          f"(from the truth {float(c1['cost_fp64_truth_start']):.7g}, from the 1024-sample CEM "
          f"{float(c1['cost_fp64_cem_start']):.7g})")
    assert abs(res.cost - c64) <= 1e-4 * c64
    assert c64 <= ref * (1 + 5e-3), (c64, ref)
