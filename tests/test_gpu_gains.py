"""Recorded PD gains: every segment-consuming kernel runs a window with the gains recorded in its
npz file (scripts/eval.py:148-152 loads them, apply_pd_gains_batch :192-201 applies them per
chunk :245-250), checked against the FP64 oracle configured with those gains."""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 5
GAINS = ((20.0, 0.5), (25.0, 0.6))      # (kp, kd) of file A and file B
LENGTHS = (24, 21)                      # control steps per file: 19 + 16 windows


def _oracle(model, sim, kp, kd):
    import spio
    s = copy.deepcopy(sim)
    s.kp = np.full(12, kp)
    s.kd = np.full(12, kd)
    return spio.Oracle(model, s)


def _files(model, sim, theta0, tmp_path):
    """Two noise-free trajectories simulated by the FP64 oracle under their own gains, written in
    the reference's npz schema."""
    from spi_amd.data import sinusoid_actions, save_npz_trajectory, standing_state
    rng = np.random.default_rng(5)
    files = []
    for k, ((kp, kd), T) in enumerate(zip(GAINS, LENGTHS)):
        o = _oracle(model, sim, kp, kd)
        settle, rc = o.simulate(theta0, standing_state(), np.zeros((40, 12), np.float32))
        assert rc == 0
        acts = sinusoid_actions(rng, 1, T, sim.dt * sim.decimation, sim.action_scale)[0]
        traj, rc = o.simulate(theta0, settle[-1], acts)
        assert rc == 0
        states = np.concatenate([settle[-1][None], traj[:-1]])      # state before each action
        f = str(tmp_path / f"gains{k}.npz")
        save_npz_trajectory(f, states.astype(np.float32), acts, sim.dt * sim.decimation, kp, kd)
        files.append(f)
    return files


def _rows(theta0, n, seed):
    rng = np.random.default_rng(seed)
    rows = np.repeat(theta0[None], n, axis=0)
    rows[1:, 0] *= rng.uniform(0.9, 1.1, n - 1)
    rows[1:, 10:18] *= rng.uniform(0.9, 1.1, (n - 1, 8))
    rows[1:, 4:7] *= rng.uniform(0.9, 1.1, (n - 1, 3))
    return rows.astype(np.float32)


def _oracle_costs(model, sim, ds, gains, rows, mode, w):
    """[n_seg, n_cand] oracle costs, each segment under its (kp, kd)."""
    out = np.empty((ds.n_seg, rows.shape[0]))
    for kp, kd in {tuple(g) for g in gains.tolist()}:
        sel = np.where((gains[:, 0] == kp) & (gains[:, 1] == kd))[0]
        o = _oracle(model, sim, kp, kd)
        out[sel] = o.rollout_cost(rows.astype(np.float64), ds.init[sel], ds.actions[sel], ds.targets[sel], mode,
                                  np.asarray(w, np.float64)).T
    return out


def _soa(rows, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(rows.T)).to(dev)


@pytest.mark.parametrize("batch_size", [None, 8])
def test_recorded_gains_match_oracle(gpu, model, sim, theta0, tmp_path, batch_size):
    """Rollout cost (one lane and four lanes per rollout), tracking and reference cost modes,
    with each file's gains (batch_size None) or the reference's first-window-of-chunk gains
    (batch_size 8: windows 16..23 form one chunk that starts in file A, so file B's windows
    19..23 run at file A's gains): |c_gpu - c_oracle| <= 2e-3 c_oracle + 1e-4.  With recorded
    gains the true parameters reproduce the noise-free data (cost <= 1e-3); file B's windows
    under the configured gains (kp 20, kd 0.5) cost far more."""
    import torch
    from spi_amd.data import load_npz_dataset, segment_gains
    from spi_amd.rollout import DEFAULT_TRACKING_WEIGHTS, REFERENCE_WEIGHTS, RolloutEngine
    ds = load_npz_dataset(_files(model, sim, theta0, tmp_path), H)
    assert ds.n_seg == sum(LENGTHS) - 2 * H
    gains = segment_gains(ds.kp, ds.kd, batch_size)
    if batch_size == 8:
        assert np.array_equal(gains[19:24], np.repeat(np.array([GAINS[0]], np.float32), 5, axis=0))
        assert np.array_equal(gains[24:], np.repeat(np.array([GAINS[1]], np.float32), ds.n_seg - 24, axis=0))
    rows = _rows(theta0, 6, seed=1)
    g_dev = torch.from_numpy(gains).to(gpu)
    for mode, w in (("tracking", DEFAULT_TRACKING_WEIGHTS), ("reference", REFERENCE_WEIGHTS)):
        eng = RolloutEngine(sim, cost_mode=mode)
        segs = eng.upload_segments(ds)
        ref = _oracle_costs(model, sim, ds, gains, rows, 0 if mode == "tracking" else 1, w)
        for physics in ("lanes", "legs"):
            g = eng.segment_costs(_soa(rows, gpu), rows.shape[0], segs, H, physics=physics,
                                  seg_gains=g_dev).cpu().numpy()
            err = np.abs(g - ref)
            print(mode, physics, "max abs err", err.max())
            assert np.all(err <= 2e-3 * ref + 1e-4), (mode, physics, err.max())
        if batch_size is None:
            g = eng.segment_costs(_soa(rows, gpu), rows.shape[0], segs, H, seg_gains=g_dev).cpu().numpy()
            wrong = eng.segment_costs(_soa(rows, gpu), rows.shape[0], segs, H).cpu().numpy()
            fileb = slice(LENGTHS[0] - H, ds.n_seg)
            assert g[:, 0].max() <= 1e-3, g[:, 0].max()
            assert np.all(wrong[fileb, 0] > 0.03)      # the oracle: >= 0.064 (reference), 0.35 (tracking)
            assert np.array_equal(wrong[:LENGTHS[0] - H], g[:LENGTHS[0] - H])   # file A: same gains


def test_recorded_gains_in_batch_adjoint_tangent_and_refiner(gpu, model, sim, theta0, tmp_path):
    """The other segment consumers under recorded gains: the persistent batched kernel agrees
    with the single-problem kernel with the same gains (median relative difference <= 1e-5,
    max 5e-2: the two are compiled separately and contract differently, and 20 % prior samples
    make contact-rich rollouts that amplify it) and differs from it without the gains; the reverse-mode (four lanes and
    one lane per rollout) and forward-mode cost + gradient kernels and the Gauss-Newton
    residual Jacobian's cost agree with the oracle (2e-3); the adjoint gradient agrees with the
    oracle's FP64 gradient under the same gains (2e-2 of each row's scale)."""
    import torch
    from spi_amd.agents.batch import BatchCEM, ProblemSpec
    from spi_amd.agents.fisher import IDENTIFIED, RolloutCost, _rows as fisher_rows, adjoint_cost_grad
    from spi_amd.agents.refine import GaussNewtonRefiner
    from spi_amd.agents.sysid import identification_prior
    from spi_amd.data import load_npz_dataset, segment_gains
    from spi_amd.rollout import DEFAULT_TRACKING_WEIGHTS, RolloutEngine
    ds = load_npz_dataset(_files(model, sim, theta0, tmp_path), H)
    gains = segment_gains(ds.kp, ds.kd)
    g_dev = torch.from_numpy(gains).to(gpu)
    eng = RolloutEngine(sim)
    segs = eng.upload_segments(ds)
    n_seg = ds.n_seg

    # batched kernel = single-problem kernel (same form: one rollout per lane)
    prior = identification_prior(theta0)
    spec = ProblemSpec(segs, H, prior, n_samples=256, seed=3, seg_gains=g_dev)
    b = BatchCEM(eng, [spec], block=64)
    b.sample()
    b.rollout()
    buf = b.bufs[0]
    single = eng.segment_costs(buf.theta, 256, segs, H, physics="lanes", seg_gains=g_dev).cpu().numpy()
    got = buf.cost_seg.cpu().numpy()
    rel = np.abs(got - single) / np.maximum(np.abs(single), 1e-6)
    print("batch vs single kernel, relative difference: median", np.median(rel), "max", rel.max())
    assert np.median(rel) <= 1e-5 and np.quantile(rel, 0.99) <= 1e-3 and rel.max() <= 5e-2
    nog = BatchCEM(eng, [ProblemSpec(segs, H, prior, n_samples=256, seed=3)], block=64)
    nog.sample()
    nog.rollout()
    fileb = slice(LENGTHS[0] - H, n_seg)
    rel_nog = np.abs(nog.bufs[0].cost_seg.cpu().numpy()[fileb] - single[fileb]) / np.abs(single[fileb])
    assert np.median(rel_nog) > 1e-2, np.median(rel_nog)      # file B's gains are in effect

    rows = _rows(theta0, 4, seed=2)
    ref = _oracle_costs(model, sim, ds, gains, rows, 0, DEFAULT_TRACKING_WEIGHTS)
    theta = _soa(rows, gpu)
    for legs in (True, False):
        c, grad = adjoint_cost_grad(eng, theta, segs, H, checkpoint_every=2, legs=legs, seg_gains=g_dev)
        err = np.abs(c.cpu().numpy() - ref)
        assert np.all(err <= 2e-3 * ref + 1e-4), ("adjoint", legs, err.max())
    w = torch.full((n_seg,), 1.0 / n_seg, dtype=torch.float32, device=gpu)
    tan = RolloutCost.apply(theta, eng, segs, H, w, fisher_rows(IDENTIFIED), "tangent", 10, g_dev)
    tref = ref.mean(axis=0)
    assert np.all(np.abs(tan.detach().cpu().numpy() - tref) <= 2e-3 * tref + 1e-4)

    # adjoint gradient of candidate 1 against the oracle's FP64 forward-mode gradient, per gains
    grad = grad.cpu().numpy()                              # one lane per rollout: [n_seg, n, 25]
    idx = np.arange(22)
    for kp, kd in GAINS:
        sel = np.where(gains[:, 0] == kp)[0]
        o = _oracle(model, sim, kp, kd)
        oc, og = o.cost_grad(rows[1].astype(np.float64), idx, ds.init[sel], ds.actions[sel], ds.targets[sel], 0,
                             np.asarray(DEFAULT_TRACKING_WEIGHTS, np.float64))
        gg = grad[sel, 1][:, idx]
        scale = np.maximum(np.abs(og).max(axis=0), 1e-6)
        assert np.all(np.abs(gg - og) <= 2e-2 * scale[None]), np.max(np.abs(gg - og) / scale[None])

    ref_obj = ref[:, 1].mean()
    r = GaussNewtonRefiner(eng, segs, H, seg_gains=g_dev)
    assert abs(r.objective(rows[1]) - ref_obj) <= 2e-3 * ref_obj + 1e-4
    _, c_here, _ = r.step(rows[1].astype(np.float64), 1e-3)
    assert abs(c_here - ref_obj) <= 2e-3 * ref_obj + 1e-4, (c_here, ref_obj)
