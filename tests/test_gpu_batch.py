"""Batched uneven problems: a problem's result does not depend on its batch-mates, the batched
rollout agrees with the single-problem kernel, and HIP-graph replays equal eager iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs(eng, theta0, sizes, seed0=0):
    from spi_amd.agents.batch import ProblemSpec
    from spi_amd.agents.sysid import ground_truth, identification_prior
    from spi_amd.data import synthetic_segments
    prior = identification_prior(theta0)
    specs = []
    for k, (H, n, n_seg) in enumerate(sizes):
        truth = ground_truth(prior, seed=100 + k)
        segs = synthetic_segments(eng.simulate_np, truth, eng.sim, n_seg=n_seg, horizon=H, seed=200 + k)
        specs.append(ProblemSpec(eng.upload_segments(segs), H, prior, n_samples=n, seed=seed0 + k))
    return specs


SIZES = [(5, 256, 8), (10, 512, 4), (3, 128, 16)]


def test_batch_result_independent_of_batch_mates(gpu, sim, theta0):
    from spi_amd.agents.batch import BatchCEM
    from spi_amd.rollout import RolloutEngine
    eng = RolloutEngine(sim)
    specs = _specs(eng, theta0, SIZES)
    both = BatchCEM(eng, specs, block=128)
    both.solve(2, graph=False)
    together = both.results()
    for k, s in enumerate(specs):
        alone = BatchCEM(eng, [s], block=128)
        alone.solve(2, graph=False)
        r = alone.results()[0]
        for key in ("mean", "std", "best", "cost"):
            assert np.array_equal(r[key], together[k][key]), (k, key)
        assert r["best_cost"] == together[k]["best_cost"]


def test_batch_rollout_matches_single_kernel(gpu, sim, theta0):
    """The persistent batched kernel against the single-problem kernel in the same form (one
    rollout per lane; small single problems otherwise pick four lanes per rollout).  The two are
    compiled separately and the compiler contracts them differently; contact-rich rollouts of
    prior samples amplify that rounding: median relative difference <= 1e-5, 99th percentile
    <= 1e-3, max <= 5e-2, the same diverged entries (each form is held to the FP64 oracle
    elsewhere)."""
    from spi_amd.agents.batch import BatchCEM
    from spi_amd.rollout import RolloutEngine
    eng = RolloutEngine(sim)
    specs = _specs(eng, theta0, SIZES)
    b = BatchCEM(eng, specs, block=64, unroll=2)
    b.sample()
    b.rollout()
    for spec, buf in zip(b.specs, b.bufs):
        ref = eng.segment_costs(buf.theta, spec.n_samples, spec.segments, spec.horizon,
                                physics="lanes").cpu().numpy()
        got = buf.cost_seg.cpu().numpy()
        assert np.array_equal(np.isfinite(got), np.isfinite(ref))
        f = np.isfinite(ref)
        rel = np.abs(got[f] - ref[f]) / np.maximum(np.abs(ref[f]), 1e-6)
        print("batch vs single: median", np.median(rel), "p99", np.quantile(rel, 0.99), "max", rel.max())
        assert np.median(rel) <= 1e-5 and np.quantile(rel, 0.99) <= 1e-3 and rel.max() <= 5e-2


def test_batch_graph_replays_equal_eager(gpu, sim, theta0):
    """Each replay compared with an eager twin, with eager work between replays."""
    import torch
    from spi_amd.agents.batch import BatchCEM
    from spi_amd.rollout import RolloutEngine
    eng = RolloutEngine(sim)
    specs = _specs(eng, theta0, SIZES[:2])
    g_run, e_run = BatchCEM(eng, specs), BatchCEM(eng, specs)
    graph = g_run.capture()
    for it in range(3):
        graph.replay()
        e_run.iterate()
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()     # unrelated eager work in between
        for a, b in zip(g_run.results(), e_run.results()):
            for key in ("mean", "std", "best", "cost"):
                assert np.array_equal(a[key], b[key]), (it, key)
    assert int(np.frombuffer(g_run.table.cpu().numpy().tobytes(), dtype=np.uint8).size) > 0


def test_autotune_keeps_philox_position(gpu, sim, theta0):
    """Switching the launch shape mid-solve keeps every problem's iteration counter, so the
    tuned solve samples the same stream as an untuned one."""
    from spi_amd.agents.batch import BatchCEM
    from spi_amd.rollout import RolloutEngine
    eng = RolloutEngine(sim)
    specs = _specs(eng, theta0, SIZES)
    a, b = BatchCEM(eng, specs, block=256), BatchCEM(eng, specs, block=256)
    a.solve(1, graph=False)
    b.solve(1, graph=False)
    info = b.autotune(repeats=1)
    assert b.device_iterations() == [1, 1, 1]
    assert info["best"] in [(64, 1), (128, 1), (256, 1), (64, 2), (128, 2), (256, 2)]
    a.solve(1, graph=False)
    b.solve(1, graph=False)
    assert b.device_iterations() == [2, 2, 2]
    ra, rb = a.results(), b.results()
    for x, y in zip(ra, rb):
        # same samples; costs may differ in the last bits between launch shapes
        assert np.allclose(x["mean"], y["mean"], rtol=1e-4, atol=1e-6)
