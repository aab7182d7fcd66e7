"""The workloads at their configured shapes, each anchored to the FP64 oracle:

* the 1-D base-mass landscape (256 masses on [5, 9] kg x 64 noisy segments x H 5): the device
  argmin is the oracle's argmin on the same data (scripts/mass_landscape.py:161-171);
* batched uneven problems at the extreme shapes of the 64-problem mix (H 50 x 8192 samples x
  512 segments beside H 5 x 512 x 32 and two middle ones) under the autotuned launch shape: the
  persistent batched kernel agrees with the single-problem kernel over each problem's whole grid
  (up to the two builds' rounding), and entries at the grid's corners and at random agree with
  the oracle;
* serving: a 40-request Poisson stream (200 / s) drawn from the serving mix (sysID jobs of
  512-8192 samples at H 5-50, scoring jobs of 16-128 excitations at H 50) under the 200 GB
  budget: every result equals the same job run alone, bit for bit, and one sysID cost and one
  scoring logdet are anchored to the oracle."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _oracle_segments(model, sim, theta, n_seg, H, seed):
    import spio
    from spi_amd.data import synthetic_segments
    o = spio.Oracle(model, sim)

    def simulate(rows, init, acts):
        return np.stack([o.simulate(np.asarray(r, np.float64), i0, a)[0] for r, i0, a in zip(rows, init, acts)])
    return o, synthetic_segments(simulate, theta, sim, n_seg, H, seed)


def test_landscape_noisy_argmin_matches_oracle(gpu, model, sim, theta0):
    """256 masses x 64 noisy segments x H 5 (the plain method's configuration): same argmin grid
    index as the oracle, unless the oracle's two best grid costs are within 1e-5 of each other
    (then either is the argmin within FP32 rounding of the costs)."""
    from spi_amd.agents.sysid import mass_landscape
    from spi_amd.rollout import RolloutEngine
    o, segs = _oracle_segments(model, sim, theta0, 64, 5, seed=0)
    eng = RolloutEngine(sim)
    masses = np.linspace(5.0, 9.0, 256)
    gc, gi = mass_landscape(eng, eng.upload_segments(segs), 5, theta0, masses)
    rows = np.repeat(np.asarray(theta0, np.float64)[None], 256, axis=0)
    rows[:, 0] = masses
    rows = rows.astype(np.float32).astype(np.float64)
    oc = o.rollout_cost_par(rows, segs.init, segs.actions, segs.targets, 0, np.asarray(eng.weights)).mean(axis=1)
    oi = int(np.argmin(oc))
    order = np.argsort(oc)
    tie = oc[order[1]] - oc[order[0]] <= 1e-5 * oc[order[0]]
    print("argmin device", masses[gi], "oracle", masses[oi], "truth", theta0[0], "tie", tie)
    assert gi == oi or (tie and gi == order[1]), (gi, oi)
    assert np.abs(gc - oc).max() <= 2e-3 * oc.max()


def test_batch_extreme_shapes_autotuned(gpu, model, sim, theta0):
    from spi_amd.agents.batch import BatchCEM, ProblemSpec
    from spi_amd.agents.sysid import ground_truth, identification_prior
    from spi_amd.data import synthetic_segments
    from spi_amd.rollout import RolloutEngine
    import spio
    eng = RolloutEngine(sim)
    prior = identification_prior(theta0)
    shapes = [(50, 8192, 512), (5, 512, 32), (20, 2048, 128), (10, 512, 32)]
    specs, segsets = [], []
    for k, (H, n, n_seg) in enumerate(shapes):
        truth = ground_truth(prior, seed=700 + k)
        segs = synthetic_segments(eng.simulate_np, truth, sim, n_seg=n_seg, horizon=H, seed=800 + k)
        segsets.append(segs)
        specs.append(ProblemSpec(eng.upload_segments(segs), H, prior, n_samples=n, seed=900 + k))
    b = BatchCEM(eng, specs)
    info = b.autotune(repeats=1)
    print("autotuned (block, unroll):", info["best"])
    b.sample()
    b.rollout()
    torch.cuda.synchronize()
    o = spio.Oracle(model, sim)
    rng = np.random.default_rng(3)
    for spec, buf, segs in zip(b.specs, b.bufs, [segsets[i] for i in b.order]):
        got = buf.cost_seg.cpu().numpy()
        ref = eng.segment_costs(buf.theta, spec.n_samples, spec.segments, spec.horizon,
                                physics="lanes").cpu().numpy()
        assert np.array_equal(np.isfinite(got), np.isfinite(ref))
        f = np.isfinite(ref)
        rel = np.abs(got[f] - ref[f]) / np.maximum(np.abs(ref[f]), 1e-6)
        print(spec.horizon, spec.n_samples, spec.segments.shape[0], "batch vs single median", np.median(rel),
              "max", rel.max())
        # the two kernels are compiled separately and round differently; over 50 control steps a
        # contact-rich rollout of a prior sample can amplify that (measured at H 50 x 8192 x 512:
        # median 0, 99th percentile 2.5e-5, one rollout in 4.2 M at 0.1)
        assert np.median(rel) <= 1e-5 and np.quantile(rel, 0.99) <= 1e-3 and np.mean(rel > 1e-2) <= 1e-4
        theta = buf.theta.cpu().numpy()
        n_seg = spec.segments.shape[0]
        for s, c in [(0, 0), (n_seg - 1, spec.n_samples - 1)] + [tuple(x) for x in
                                                                 zip(rng.integers(0, n_seg, 2),
                                                                     rng.integers(0, spec.n_samples, 2))]:
            oc = o.rollout_cost(theta[:, c].astype(np.float64), segs.init[s:s + 1], segs.actions[s:s + 1],
                                segs.targets[s:s + 1], 0, np.asarray(eng.weights))[0, 0]
            if np.isfinite(oc):
                assert abs(got[s, c] - oc) <= 2e-3 * oc + 1e-4, (spec.horizon, s, c, got[s, c], oc)


def test_serving_poisson_mix_equals_alone(gpu, model, sim, theta0):
    import spio
    from spi_amd.agents.batch import BatchCEM, ProblemSpec
    from spi_amd.agents.fisher import FisherObjective, fourier_excitations
    from spi_amd.agents.sysid import ground_truth, identification_prior
    from spi_amd.data import standing_state, synthetic_segments
    from spi_amd.rollout import RolloutEngine
    from spi_amd.serving import ScoreRequest, SysIdRequest, SysIdServer, poisson_arrivals
    eng = RolloutEngine(sim)
    prior = identification_prior()
    rng = np.random.default_rng(9)
    n_req, n_seg, iters = 40, 4, 8
    arrivals = poisson_arrivals(n_req, 200.0, seed=9)
    settled = eng.simulate_np(prior.mean[None], standing_state(0.30)[None], np.zeros((1, 50, 12), np.float32))[0, -1]
    init = torch.from_numpy(settled.astype(np.float32)).cuda()
    reqs, sysid, score = [], {}, {}
    for k in range(n_req):
        t = float(arrivals[k] * 200.0)            # tick clock: arrival in units of the mean gap
        if rng.random() < 0.8:
            H = int(rng.integers(5, 51))
            n = int(np.clip(np.exp(rng.uniform(np.log(512), np.log(8192))) // 256 * 256, 512, 8192))
            truth = ground_truth(prior, seed=5000 + k)
            segs = synthetic_segments(eng.simulate_np, truth, sim, n_seg=n_seg, horizon=H, seed=6000 + k)
            spec = ProblemSpec(eng.upload_segments(segs), H, prior, n_samples=n, seed=k)
            sysid[k] = (spec, segs)
            reqs.append(SysIdRequest(k, t, spec, iterations=iters))
        else:
            M = int(rng.integers(16, 129))
            acts = fourier_excitations(M, 50, sim.dt * sim.decimation, sim.action_scale, seed=k)
            score[k] = acts
            reqs.append(ScoreRequest(k, t, acts, init, prior.mean))
    srv = SysIdServer(eng, hbm_budget_bytes=200e9)
    stats = srv.run(reqs, clock="ticks")
    assert len(stats.results) == n_req and stats.max_active >= 2
    for k, (spec, _) in sysid.items():
        alone = BatchCEM(eng, [spec])
        alone.solve(iters, graph=False)
        r = alone.results()[0]
        for key in ("mean", "std", "best"):
            assert np.array_equal(r[key], stats.results[k][key]), (k, key)
        assert r["best_cost"] == stats.results[k]["best_cost"], k
    th = torch.tensor(np.asarray(prior.mean, np.float32).reshape(-1, 1), device="cuda")
    for k, acts in score.items():
        ld = FisherObjective(eng).score(th, acts, init).logdet.cpu().numpy()
        assert np.array_equal(ld, stats.results[k]["score"]), k

    # anchors: one sysID job's best cost against the oracle, one excitation's logdet against
    # the oracle's FP64 sensitivities
    o = spio.Oracle(model, sim)
    k, (spec, segs) = next(iter(sysid.items()))
    best = stats.results[k]["best"].astype(np.float64)
    oc = o.rollout_cost(best, segs.init, segs.actions, segs.targets, 0, np.asarray(eng.weights)).mean()
    assert abs(stats.results[k]["best_cost"] - oc) <= 2e-3 * oc, (stats.results[k]["best_cost"], oc)
    k, acts = next(iter(score.items()))
    lds = stats.results[k]["score"]
    e = int(np.flatnonzero(np.isfinite(lds))[0])
    _, J, rc = o.traj_sens(prior.mean, np.arange(22), settled, acts[e].cpu().numpy(), n_rows=25)
    assert rc == 0
    J = J.reshape(-1, 22)
    F = J.T @ J
    F += 1e-6 * np.trace(F) / 22 * np.eye(22)
    ld64 = np.linalg.slogdet(F)[1]
    print("logdet device", lds[e], "oracle FP64", ld64)
    assert abs(lds[e] - ld64) <= 0.1, (lds[e], ld64)
