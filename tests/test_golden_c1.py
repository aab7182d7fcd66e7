"""The committed FP64 identification fixture (tests/golden/c1_fp64_minimiser.npz) has the
properties tests/golden/make_c1_fixture.py states, checked on the CPU with the oracle:

* its segments are the workload's recipe (the seeded truth, sinusoid-driven trajectories of
  the FP64 oracle plus measurement noise) and regenerate bit for bit;
* every stored cost is the oracle's FP64 objective at the stored parameters;
* every FP64 search's cost history decreases, and its end point lies inside the prior's bounds;
* the stored standard errors are positive, consistent with their relative form, and below
  0.5 % for the 14 parameters the identification tests hold to 1 %.
"""
import os

import numpy as np
import pytest

from conftest import ROOT

FIXTURE = os.path.join(ROOT, "tests", "golden", "c1_fp64_minimiser.npz")


@pytest.fixture(scope="module")
def c1():
    with np.load(FIXTURE, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_segments_regenerate(model, sim, c1):
    import spio
    from spi_amd.agents.sysid import ground_truth, identification_prior
    from spi_amd.data import synthetic_segments
    prior = identification_prior()
    truth = ground_truth(prior, seed=int(c1["truth_seed"]))
    assert np.array_equal(truth, c1["truth"])
    o = spio.Oracle(model, sim)

    def simulate(rows, init, acts):
        return np.stack([o.simulate(np.asarray(r, np.float64), i0, a)[0] for r, i0, a in zip(rows, init, acts)])

    segs = synthetic_segments(simulate, truth, sim, n_seg=int(c1["n_seg"]), horizon=int(c1["horizon"]),
                              seed=int(c1["data_seed"]))
    assert np.array_equal(segs.packed(), c1["segments"])


def test_stored_costs_are_the_fp64_objective(model, sim, c1):
    import spio
    from spi_amd.data import SegmentSet
    segs = SegmentSet.from_packed(c1["segments"], int(c1["horizon"]))
    pts = [("theta_fp64", "cost_fp64"), ("theta_fp64_truth_start", "cost_fp64_truth_start"),
           ("theta_fp64_cem_start", "cost_fp64_cem_start"), ("truth", "cost_truth")]
    if "theta_fp64_same_seed" in c1:
        pts.append(("theta_fp64_same_seed", "cost_fp64_same_seed"))
    o = spio.Oracle(model, sim)
    rows = np.stack([c1[p] for p, _ in pts])
    got = o.rollout_cost_par(rows, segs.init, segs.actions, segs.targets, 0, c1["weights"]).mean(axis=1)
    ref = np.array([float(c1[c]) for _, c in pts])
    assert np.allclose(got, ref, rtol=1e-12, atol=0), (got, ref)
    # the reference minimiser is the lower of the two independent end points
    assert float(c1["cost_fp64"]) == min(float(c1["cost_fp64_truth_start"]), float(c1["cost_fp64_cem_start"]))


def test_searches_descend_inside_the_bounds(c1):
    from spi_amd.agents.sysid import identification_prior
    prior = identification_prior()
    runs = [("history_truth_start", "theta_fp64_truth_start"), ("history_cem_start", "theta_fp64_cem_start")]
    if "history_same_seed" in c1:
        runs.append(("history_same_seed", "theta_fp64_same_seed"))
    for h, th in runs:
        hist = c1[h]
        assert len(hist) > 5 and np.all(np.diff(hist) < 0), h
        assert np.all(c1[th] >= prior.lo - 1e-12) and np.all(c1[th] <= prior.hi + 1e-12), th
    if "cem_means_same_seed" in c1:
        m, sd = c1["cem_means_same_seed"], c1["cem_stds_same_seed"]
        assert m.shape == sd.shape == (int(c1["same_seed_config"][1]), 25)
        assert np.all(sd >= prior.min_std - 1e-12)


def test_standard_errors(c1):
    se, se_rel, th = c1["se"], c1["se_rel"], c1["theta_fp64"]
    rows = c1["rows"]
    assert np.all(se[rows] > 0)
    assert np.allclose(se_rel[rows], se[rows] / np.abs(th[rows]))
    # the parameters the identification tests hold to 1 %: mass, CoM x/y, principal inertias,
    # Kp and Kd scales
    ident = [i for i in rows if 2 * se_rel[i] < 0.01]
    assert {0, 1, 2, 4, 5, 6, 10, 11, 12, 13, 14, 15, 16, 17} <= set(int(i) for i in ident)
