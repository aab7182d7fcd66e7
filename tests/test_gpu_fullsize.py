"""The 22-parameter CEM configuration at its full size (4096 samples x 256 segments x H 20:
21 M control steps per rollout launch) through properties that do not depend on the size:
repeated launches are bit-identical, every entry of a full launch equals the same kernel run on
just that candidate and segment (which checks the grid, block remapping and indexing at full
size), a sample of entries matches the FP64 oracle, the segment reduction matches a float64 sum,
and the CEM update of 4096 samples selects numpy's elite set.  One chunk of the Fisher workload
(4096 samples x 16 excitations x H 50, 22 tangents) through the same slice property."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N_SAMPLES, N_SEG, H = 4096, 256, 20
CAND = [0, 1, 255, 256, 2049, 4095]      # first, block edges, middle, last
SEGS = [0, 1, 128, 255]


@pytest.fixture(scope="module")
def c1(gpu, sim):
    from spi_amd.agents.sysid import ground_truth, identification_prior
    from spi_amd.data import synthetic_segments
    from spi_amd.rollout import RolloutEngine
    eng = RolloutEngine(sim)
    prior = identification_prior()
    segs = synthetic_segments(eng.simulate_np, ground_truth(prior, seed=1), sim, n_seg=N_SEG, horizon=H, seed=2)
    seg_dev = eng.upload_segments(segs)
    rng = np.random.default_rng(0)
    rows = np.repeat(prior.mean[None], N_SAMPLES, axis=0)
    rows[:, prior.free] += rng.normal(size=(N_SAMPLES, int(prior.free.sum()))) * prior.std[prior.free]
    rows = np.clip(rows, prior.lo, prior.hi).astype(np.float32)
    theta = torch.from_numpy(np.ascontiguousarray(rows.T)).to(gpu)
    return eng, segs, seg_dev, rows, theta


def _slice(theta, seg_dev):
    sub_theta = theta[:, CAND].contiguous()
    sub_segs = seg_dev[SEGS].contiguous()
    return sub_theta, sub_segs


def test_full_size_rollout_costs(c1, model, sim):
    import spio
    from spi_amd.rollout import DEFAULT_TRACKING_WEIGHTS
    eng, segs, seg_dev, rows, theta = c1
    a = eng.segment_costs(theta, N_SAMPLES, seg_dev, H, physics="lanes")
    b = eng.segment_costs(theta, N_SAMPLES, seg_dev, H, physics="lanes")
    assert torch.equal(a, b)                                        # deterministic
    assert torch.isfinite(a).all()                                  # no sampled candidate diverged
    sub_theta, sub_segs = _slice(theta, seg_dev)
    small = eng.segment_costs(sub_theta, len(CAND), sub_segs, H, physics="lanes")
    assert torch.equal(small, a[SEGS][:, CAND])                     # same lanes, same bits
    o = spio.Oracle(model, sim)
    ref = o.rollout_cost(rows[CAND[:3]].astype(np.float64), segs.init[SEGS[:2]], segs.actions[SEGS[:2]],
                         segs.targets[SEGS[:2]], 0, np.asarray(DEFAULT_TRACKING_WEIGHTS, np.float64)).T
    got = a[SEGS[:2]][:, CAND[:3]].cpu().numpy()
    assert np.all(np.abs(got - ref) <= 2e-3 * ref + 1e-4), np.abs(got - ref).max()
    w = torch.full((N_SEG,), 1.0 / N_SEG, device=a.device)
    red = eng.reduce(a, w).double()
    ref_red = (w.double()[:, None] * a.double()).sum(0)
    assert torch.allclose(red, ref_red, rtol=1e-5, atol=0.0)


def test_full_size_adjoint_entries_equal_slice(c1):
    """The four-lanes-per-rollout reverse-mode kernel at full size: every checked (candidate,
    segment) entry of cost and gradient equals the same kernel on the slice, bit for bit."""
    from spi_amd.agents.fisher import adjoint_cost_grad
    eng, segs, seg_dev, rows, theta = c1
    cost, grad = adjoint_cost_grad(eng, theta, seg_dev, H, checkpoint_every=10)
    assert torch.isfinite(cost).all() and torch.isfinite(grad).all()
    sub_theta, sub_segs = _slice(theta, seg_dev)
    c_s, g_s = adjoint_cost_grad(eng, sub_theta, sub_segs, H, checkpoint_every=10)
    assert torch.equal(c_s, cost[SEGS][:, CAND])
    assert torch.equal(g_s, grad[SEGS][:, CAND])


def test_full_size_cem_update_matches_numpy(gpu):
    """4096 samples, 410 elites (10 %, rounded up), 25 parameter rows: numpy's elite set and
    moments."""
    from spi_amd import _lib
    rng = np.random.default_rng(21)
    n_param, n, n_elite, alpha = 25, N_SAMPLES, 410, 0.0
    cost = rng.gamma(2.0, 1.0, n).astype(np.float32)
    cost[::331] = cost[7]                                           # ties
    theta = rng.normal(0, 1, (n_param, n)).astype(np.float32)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(gpu)
    mean, std = t(np.zeros(n_param, np.float32)), t(np.ones(n_param, np.float32))
    best, best_cost = torch.zeros(n_param, device=gpu), torch.full((1,), float("inf"), device=gpu)
    elite = torch.zeros(n_elite, dtype=torch.int32, device=gpu)
    dc, dth, dms = t(cost), t(theta), t(np.zeros(n_param, np.float32))
    _lib.check(_lib.lib().spi_cem_update(_lib.ptr(dc), _lib.ptr(dth), n, n_param, n, n_elite, alpha, _lib.ptr(dms),
                                         _lib.ptr(mean), _lib.ptr(std), _lib.ptr(best), _lib.ptr(best_cost),
                                         _lib.ptr(elite, torch.int32), _lib.stream()), "spi_cem_update")
    order = np.lexsort((np.arange(n), cost))[:n_elite]
    assert np.array_equal(elite.cpu().numpy(), order)
    el = theta[:, order].astype(np.float64)
    assert np.allclose(mean.cpu().numpy(), el.mean(1), atol=1e-5)
    assert np.allclose(std.cpu().numpy(), el.std(1), atol=1e-5)


def test_full_size_fisher_chunk_equals_slice(gpu, sim):
    """One chunk of the Fisher workload at its full size (4096 samples x 16 excitations x H 50,
    22 tangents on the four-lanes-per-two-tangents kernel: J is 7.2 GB): the J rows and status
    of chosen (sample, excitation) groups equal the same kernel on just those samples and
    excitations, bit for bit, and the MFMA assembly of the full chunk equals float64 J^T J of
    those groups to FP32 accumulation error."""
    from spi_amd.agents.fisher import FisherObjective, fourier_excitations
    from spi_amd.agents.sysid import identification_prior
    from spi_amd.data import standing_state
    from spi_amd.rollout import RolloutEngine
    eng = RolloutEngine(sim)
    prior = identification_prior()
    rng = np.random.default_rng(3)
    rows = np.repeat(prior.mean[None], N_SAMPLES, axis=0)
    rows[:, prior.free] += rng.normal(size=(N_SAMPLES, int(prior.free.sum()))) * prior.std[prior.free] * 0.5
    rows = np.clip(rows, prior.lo, prior.hi).astype(np.float32)
    theta = torch.from_numpy(np.ascontiguousarray(rows.T)).to(gpu)
    n_exc, horizon = 16, 50
    acts = fourier_excitations(n_exc, horizon, sim.dt * sim.decimation, sim.action_scale, seed=1)
    init = torch.tensor(standing_state(0.30), dtype=torch.float32, device=gpu)
    obj = FisherObjective(eng)
    assert obj.pairs
    J, st, _ = obj.sensitivities(theta, acts, init)
    F = obj.assemble(J)
    cand, exc = [0, 63, 64, 2049, 4095], [0, 7, 15]
    groups = [e * N_SAMPLES + c for e in exc for c in cand]
    J_s, st_s, _ = obj.sensitivities(theta[:, cand].contiguous(), acts[exc].contiguous(), init)
    assert torch.equal(st_s, st[groups])
    assert torch.equal(J_s, J[groups])
    Jg = J[groups].double()
    ref = Jg.transpose(1, 2) @ Jg
    err = (F[groups].double() - ref).abs().amax(dim=(1, 2))
    scale = ref.abs().amax(dim=(1, 2))
    assert torch.all(err <= 2e-4 * scale), (err / scale).max()
    del J, F

