"""The oracle's FP64 forward-mode derivatives (the reference for the GPU sensitivities and
gradients) agree with central finite differences of the FP64 oracle itself."""
import numpy as np


def _setup(model, sim, theta0):
    import spio
    from spi_amd.data import standing_state
    o = spio.Oracle(model, sim)
    st, _ = o.simulate(theta0, standing_state(0.30), np.zeros((40, 12), np.float32))
    return o, st[-1]


def test_trajectory_sensitivities_match_central_differences(model, sim, theta0):
    o, init = _setup(model, sim, theta0)
    rng = np.random.default_rng(0)
    acts = (rng.uniform(-0.3, 0.3, size=(8, 12)) / sim.action_scale).astype(np.float32)
    idx = np.arange(22)
    y, J, rc = o.traj_sens(theta0, idx, init, acts)
    assert rc == 0 and np.abs(J).max() > 0
    for m in idx:
        h = 1e-5 * max(abs(theta0[m]), 1e-2)
        tp, tm = theta0.copy(), theta0.copy()
        tp[m] += h
        tm[m] -= h
        fd = (o.traj_sens(tp, idx[:1], init, acts)[0] - o.traj_sens(tm, idx[:1], init, acts)[0]) / (2 * h)
        err = np.abs(fd - J[:, :, m]).max() / (np.abs(J[:, :, m]).max() + 1e-12)
        assert err < 1e-5, (m, err)
    # the value path equals the plain FP64 simulation
    st, _ = o.simulate(theta0, init, acts)
    assert np.allclose(y[:, :13], st[:, :13], atol=1e-6) and np.allclose(y[:, 13:], st[:, 13:25], atol=1e-6)


def test_cost_gradient_matches_central_differences(model, sim, theta0):
    from spi_amd.agents.sysid import ground_truth, identification_prior
    from spi_amd.data import synthetic_segments
    o, _ = _setup(model, sim, theta0)

    def simulate(theta, init, acts):
        return np.stack([o.simulate(np.asarray(theta)[i], init[i], acts[i])[0] for i in range(init.shape[0])])
    prior = identification_prior(theta0)
    segs = synthetic_segments(simulate, ground_truth(prior, seed=3), sim, n_seg=2, horizon=4, seed=4)
    w = np.array([10.0, 5.0, 1.0, 1.0, 1.0, 0.05])
    idx = np.arange(22)
    c0, g = o.cost_grad(theta0, idx, segs.init, segs.actions, segs.targets, 0, w)
    ref = o.rollout_cost(theta0[None], segs.init, segs.actions, segs.targets, 0, w)[0]
    assert np.allclose(c0, ref, rtol=1e-12)
    for m in idx:
        h = 1e-5 * max(abs(theta0[m]), 1e-2)
        tp, tm = theta0.copy(), theta0.copy()
        tp[m] += h
        tm[m] -= h
        fd = (o.rollout_cost(tp[None], segs.init, segs.actions, segs.targets, 0, w)[0] -
              o.rollout_cost(tm[None], segs.init, segs.actions, segs.targets, 0, w)[0]) / (2 * h)
        assert np.allclose(fd, g[:, m], rtol=1e-5, atol=1e-6 * np.abs(g).max()), (m, fd, g[:, m])


def test_fp64_gauss_newton_recovers_truth_without_noise(model, sim, theta0):
    """The oracle's FP64 Levenberg-Marquardt (the reference minimiser for the GPU refiner)
    converges to the ground truth on noise-free data, where the minimum is exactly the truth."""
    import spio
    from spi_amd.agents.sysid import ground_truth, identification_prior
    from spi_amd.data import synthetic_segments
    from spi_amd.params import P_MASS
    o, _ = _setup(model, sim, theta0)

    def simulate(theta, init, acts):
        return np.stack([o.simulate(np.asarray(theta)[i], init[i], acts[i])[0] for i in range(init.shape[0])])
    prior = identification_prior(theta0)
    truth = ground_truth(prior, seed=8)
    segs = synthetic_segments(simulate, truth, sim, n_seg=4, horizon=6, seed=9, sigma_q=0.0, sigma_v=0.0)
    start = truth.copy()
    start[:22] *= 1 + 0.02 * np.random.default_rng(1).normal(size=22)
    w = np.array([10.0, 5.0, 1.0, 1.0, 1.0, 0.05])
    th, cost, hist = spio.gauss_newton_fp64(o, start, np.arange(22), segs, w, np.full(4, 0.25), prior.lo, prior.hi,
                                            iterations=12)
    assert cost < 1e-4 * hist[0], hist
    assert abs(th[P_MASS] - truth[P_MASS]) / truth[P_MASS] < 1e-3
