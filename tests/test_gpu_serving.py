"""Continuous batching: jobs join and leave the running batch at iteration boundaries, queue
for memory, and each job's result is bit-identical to the same job solved alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spec(eng, prior, k, H, n, n_seg):
    from spi_amd.agents.batch import ProblemSpec
    from spi_amd.agents.sysid import ground_truth
    from spi_amd.data import synthetic_segments
    truth = ground_truth(prior, seed=300 + k)
    segs = synthetic_segments(eng.simulate_np, truth, eng.sim, n_seg=n_seg, horizon=H, seed=400 + k)
    return ProblemSpec(eng.upload_segments(segs), H, prior, n_samples=n, seed=17 + k)


def test_serving_mix_equals_alone(gpu, sim, theta0):
    import torch
    from spi_amd.agents.active import FDFIMObjective
    from spi_amd.agents.batch import BatchCEM, _Buffers
    from spi_amd.agents.fisher import FisherObjective
    from spi_amd.agents.sysid import identification_prior
    from spi_amd.rollout import RolloutEngine
    from spi_amd.serving import ScoreRequest, SysIdRequest, SysIdServer

    eng = RolloutEngine(sim)
    prior = identification_prior(theta0)
    jobs = [(0, 0, 5, 256, 8, 4), (1, 1, 10, 512, 4, 3), (2, 1, 3, 128, 16, 2), (3, 2, 5, 256, 4, 3)]
    reqs, specs = [], {}
    for job_id, arrival, H, n, n_seg, iters in jobs:
        specs[job_id] = (_spec(eng, prior, job_id, H, n, n_seg), iters)
        reqs.append(SysIdRequest(job_id, arrival, specs[job_id][0], iterations=iters))
    rng = np.random.default_rng(5)
    acts = torch.tensor(rng.uniform(-1, 1, size=(4, 6, 12)), dtype=torch.float32, device="cuda")
    init = torch.tensor(np.r_[0, 0, 0.3, 0, 0, 0, 1, np.zeros(6), sim.q0, np.zeros(12)],
                        dtype=torch.float32, device="cuda")
    reqs.append(ScoreRequest(10, 1, acts, init, theta0))
    reqs.append(ScoreRequest(11, 2, acts, init, theta0, explore=("mass", "comx"), mode="fd"))
    # a budget that holds the two largest jobs but not three: job 2 / 3 must queue
    fp = sorted(_Buffers.footprint(s) for s, _ in specs.values())
    budget = fp[-1] + fp[-2] + fp[0] // 2
    srv = SysIdServer(eng, hbm_budget_bytes=budget, block=128)
    stats = srv.run(reqs, clock="ticks")
    assert len(stats.results) == len(reqs)
    assert stats.max_active >= 2

    for job_id, (spec, iters) in specs.items():
        alone = BatchCEM(eng, [spec], block=128)
        alone.solve(iters, graph=False)
        r = alone.results()[0]
        got = stats.results[job_id]
        for key in ("mean", "std", "best"):
            assert np.array_equal(r[key], got[key]), (job_id, key)
        assert r["best_cost"] == got["best_cost"], job_id
    th = torch.tensor(np.asarray(theta0, np.float32).reshape(-1, 1), device="cuda")
    ld = FisherObjective(eng).score(th, acts, init).logdet.cpu().numpy()
    assert np.array_equal(ld, stats.results[10]["score"])
    fim = FDFIMObjective(eng, theta0, ["mass", "comx"], delta=0.1)(acts, init).cpu().numpy()
    assert np.array_equal(fim, stats.results[11]["score"])
    s = stats.summary()
    assert s["completed"] == len(reqs) and s["p99_ms"] >= s["p50_ms"] > 0


def test_serving_wall_clock_poisson(gpu, sim, theta0):
    from spi_amd.agents.sysid import identification_prior
    from spi_amd.rollout import RolloutEngine
    from spi_amd.serving import SysIdRequest, SysIdServer, poisson_arrivals

    eng = RolloutEngine(sim)
    prior = identification_prior(theta0)
    t = poisson_arrivals(6, rate_hz=200.0, seed=3)
    reqs = [SysIdRequest(k, float(t[k]), _spec(eng, prior, 20 + k, 5, 128, 4), iterations=2) for k in range(6)]
    stats = SysIdServer(eng, block=64).run(reqs)
    assert len(stats.results) == 6
    assert all(np.isfinite(r["best_cost"]) for r in stats.results.values())
