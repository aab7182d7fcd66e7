"""Per-lane failure isolation: a candidate whose parameters cannot be integrated (a NaN or an
infinity anywhere, base mass <= 0, base inertia not positive definite) gets cost +inf and a zero
gradient without running, a candidate with an absurd PD gain runs with its torque saturated at
the effort limit, and no other lane of the launch changes by a single bit.  Covered: the
rollout cost with one lane and four lanes per rollout, the persistent batched kernel and the
reverse-mode cost + gradient with four lanes and one lane per rollout; the dynamics are built
with finite-math, so NaN and infinity are detected on exponent bits."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N_CAND, N_SEG, H = 300, 8, 10
INVALID = {17: ("mass", np.nan), 101: ("mass", -1.0), 150: ("mass", 0.0), 299: ("ixx", -0.01),
           42: ("ixy", 10.0), 7: ("comy", np.inf), 260: ("kd_FR", -np.inf)}
ABSURD_KP = 233                                       # kp_scale_FL = 1e6


def _setup(sim, theta0):
    from spi_amd.agents.sysid import identification_prior
    from spi_amd.data import synthetic_segments
    from spi_amd.rollout import RolloutEngine
    eng = RolloutEngine(sim)
    prior = identification_prior(theta0)
    segs = synthetic_segments(eng.simulate_np, theta0, sim, n_seg=N_SEG, horizon=H, seed=41)
    rng = np.random.default_rng(42)
    rows = np.repeat(prior.mean[None], N_CAND, axis=0)
    rows[:, prior.free] += 0.5 * rng.normal(size=(N_CAND, int(prior.free.sum()))) * prior.std[prior.free]
    clean = np.clip(rows, prior.lo, prior.hi).astype(np.float32)
    bad = clean.copy()
    col = {"mass": 0, "comy": 2, "ixx": 4, "ixy": 7, "kd_FR": 15}
    for lane, (name, v) in INVALID.items():
        bad[lane, col[name]] = v
    bad[ABSURD_KP, 10] = 1e6
    f = lambda r: torch.from_numpy(np.ascontiguousarray(r.T)).cuda()
    return eng, prior, eng.upload_segments(segs), f(clean), f(bad)


def _check(name, clean, bad, grad_clean=None, grad_bad=None):
    """clean/bad: [n_seg, n_cand] costs (and [n_seg, n_cand, 25] gradients)."""
    inv = sorted(INVALID)
    assert np.all(np.isposinf(bad[:, inv])), (name, bad[:, inv])
    others = np.setdiff1d(np.arange(N_CAND), inv + [ABSURD_KP])
    assert np.array_equal(bad[:, others], clean[:, others]), name
    kp = bad[:, ABSURD_KP]
    assert np.all(np.isposinf(kp) | (kp > clean[:, ABSURD_KP])), (name, kp, clean[:, ABSURD_KP])
    if grad_bad is not None:
        assert np.all(grad_bad[:, inv] == 0.0), name
        assert np.array_equal(grad_bad[:, others], grad_clean[:, others]), name
        assert np.isfinite(grad_bad).all()
    assert np.isfinite(clean).all(), name


def test_invalid_and_absurd_lanes_are_isolated(gpu, sim, theta0):
    from spi_amd.agents.batch import BatchCEM, ProblemSpec
    from spi_amd.agents.fisher import adjoint_cost_grad
    eng, prior, segs, clean, bad = _setup(sim, theta0)
    for physics in ("lanes", "legs"):
        c = eng.segment_costs(clean, N_CAND, segs, H, physics=physics).cpu().numpy()
        b = eng.segment_costs(bad, N_CAND, segs, H, physics=physics).cpu().numpy()
        _check(f"rollout {physics}", c, b)

    res = {}
    for which, theta in (("clean", clean), ("bad", bad)):
        run = BatchCEM(eng, [ProblemSpec(segs, H, prior, n_samples=N_CAND, seed=1)], block=64)
        run.sample()
        run.bufs[0].theta.copy_(theta)
        run.rollout()
        res[which] = run.bufs[0].cost_seg.cpu().numpy()
    _check("batch", res["clean"], res["bad"])

    for legs in (True, False):
        cc, gc = adjoint_cost_grad(eng, clean, segs, H, checkpoint_every=5, legs=legs)
        cb, gb = adjoint_cost_grad(eng, bad, segs, H, checkpoint_every=5, legs=legs)
        _check(f"adjoint legs={legs}", cc.cpu().numpy(), cb.cpu().numpy(), gc.cpu().numpy(), gb.cpu().numpy())
