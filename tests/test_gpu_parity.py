"""GPU parity: every HIP entry point against the FP64 oracle (or a closed form) on the same
seeded inputs.  Tolerances are stated per test; FP32 kernels vs the FP64 oracle."""
import numpy as np
import pytest

from conftest import random_state

pytestmark = pytest.mark.gpu


def _oracle(model, sim, **over):
    import copy
    import spio
    s = copy.deepcopy(sim)
    for k, v in over.items():
        setattr(s, k, v)
    return spio.Oracle(model, s), s


def _engine(sim, **kw):
    from spi_amd.rollout import RolloutEngine
    return RolloutEngine(sim, **kw)


def _soa(rows, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(rows, np.float32).T)).to(dev)


def _rand_params(rng, theta0, n, spread=0.1):
    rows = np.repeat(theta0[None], n, axis=0)
    rows[:, 0] *= rng.uniform(1 - 2 * spread, 1 + 2 * spread, n)           # mass
    rows[:, 1:4] += rng.normal(0, 0.01, (n, 3))                             # CoM
    rows[:, 4:7] *= rng.uniform(1 - spread, 1 + spread, (n, 3))              # principal inertia
    rows[:, 10:18] *= rng.uniform(1 - spread, 1 + spread, (n, 8))            # kp/kd scales
    rows[:, 18:22] = rng.uniform(0.0, 0.3, (n, 4))                           # friction
    rows[:, 22:25] = rng.uniform(5.0, 30.0, (n, 3))                          # motor terms
    return rows


@pytest.mark.parametrize("control_mode,motor_model", [(0, 0), (1, 0), (1, 1), (1, 2), (1, 3), (0, 3)])
def test_torques_match_oracle(gpu, model, sim, theta0, control_mode, motor_model):
    """PD + clip + motor model: |tau_gpu - tau_oracle| <= 1e-5 * max(|tau|, g) where g is the
    motor model's output scale (its gain for act2tau_scalar / act2tau_vec3, 1 N m otherwise):
    a gain multiplies the FP32 rounding of the PD terms along with the torque."""
    import torch
    o, s = _oracle(model, sim, control_mode=control_mode, motor_model=motor_model)
    eng = _engine(s)
    rng = np.random.default_rng(10 + 7 * control_mode + motor_model)
    n = 512
    rows = _rand_params(rng, theta0, n)
    q = np.asarray(s.q0)[None] + rng.normal(0, 0.5, (n, 12))
    qd = rng.normal(0, 5, (n, 12))
    act = rng.normal(0, 15, (n, 12))   # exercises the +-20 action clip and the torque clip
    dof = np.stack([q, qd], axis=2).astype(np.float32)
    tau = eng.compute_torques(_soa(rows, gpu), torch.from_numpy(act.astype(np.float32)).to(gpu),
                              torch.from_numpy(dof).to(gpu)).cpu().numpy()
    ref = np.stack([o.control_torque(act[i].astype(np.float32), dof[i, :, 0], dof[i, :, 1],
                                     rows[i].astype(np.float32)) for i in range(n)])
    grp = np.arange(12) % 3
    if motor_model == 1:
        scale = rows[:, 22:23] * np.ones((1, 12))
    elif motor_model == 2:
        scale = rows[:, 22 + grp]
    else:
        scale = np.ones((n, 12))
    err = np.abs(tau - ref) / np.maximum(np.abs(ref), scale)
    assert err.max() <= 1e-5, err.max()


def _realistic_states(model, sim, theta0, n, seed):
    """States taken from an oracle rollout (feet in and out of contact), velocities perturbed."""
    import spio
    from spi_amd.data import sinusoid_actions, standing_state
    o = spio.Oracle(model, sim)
    rng = np.random.default_rng(seed)
    settled, _ = o.simulate(theta0, standing_state(), np.zeros((40, 12), np.float32))
    acts = sinusoid_actions(rng, 1, n, sim.dt * sim.decimation, sim.action_scale)[0]
    traj, rc = o.simulate(theta0, settled[-1], acts)
    assert rc == 0
    states = traj.astype(np.float64)
    states[:, 25:37] += rng.normal(0, 0.5, (n, 12))
    states[:, 7:13] += rng.normal(0, 0.1, (n, 6))
    return states


def test_physics_step_matches_oracle(gpu, model, sim, theta0):
    """One physics step (4 sub-intervals of stiff penalty contact): positions within 5e-6 m/rad,
    velocities within 1e-3 (FP32 acceleration rounding, amplified by the contact stiffness)."""
    import torch
    o, s = _oracle(model, sim)
    eng = _engine(s)
    rng = np.random.default_rng(3)
    n = 128
    states = _realistic_states(model, s, theta0, n, seed=4)
    rows = _rand_params(rng, theta0, n)
    tau = rng.normal(0, 5, (n, 12))
    root = torch.from_numpy(states[:, :13].astype(np.float32)).to(gpu)
    dof = torch.from_numpy(np.stack([states[:, 13:25], states[:, 25:37]], 2).astype(np.float32)).to(gpu).contiguous()
    ff = torch.zeros((n, 12), dtype=torch.float32, device=gpu)
    eng.physics_step(_soa(rows, gpu), root, dof, torch.from_numpy(tau.astype(np.float32)).to(gpu), ff)
    g_root, g_dof = root.cpu().numpy().astype(np.float64), dof.cpu().numpy().astype(np.float64)
    in_contact = 0
    perr, verr = 0.0, 0.0
    for i in range(n):
        s32 = states[i].astype(np.float32).astype(np.float64)
        r, q, qd = o.physics_step(s32[:13], s32[13:25], s32[25:37], tau[i].astype(np.float32),
                                  rows[i].astype(np.float32).astype(np.float64))
        perr = max(perr, np.abs(g_root[i, :7] - r[:7]).max(), np.abs(g_dof[i, :, 0] - q).max())
        verr = max(verr, np.abs(g_root[i, 7:] - r[7:]).max(), np.abs(g_dof[i, :, 1] - qd).max())
        in_contact += int(np.abs(ff[i].cpu().numpy()).sum() > 0)
    print(f"physics step: max pos err {perr:.3g}, max vel err {verr:.3g}, envs in contact {in_contact}/{n}")
    assert in_contact > n // 4
    assert perr <= 5e-6 and verr <= 1e-3, (perr, verr)


def test_forward_dynamics_matches_oracle(gpu, model, sim, theta0):
    """Accelerations (base spatial acceleration + qdd) of the specialised FP32 ABA vs the FP64
    generic-tree ABA on realistic states: |da| <= 1e-3 * max|a| per env."""
    import torch
    o, s = _oracle(model, sim)
    eng = _engine(s)
    rng = np.random.default_rng(15)
    n = 128
    states = _realistic_states(model, s, theta0, n, seed=16).astype(np.float32)
    rows = _rand_params(rng, theta0, n).astype(np.float32)
    tau = rng.normal(0, 5, (n, 12)).astype(np.float32)
    root = torch.from_numpy(states[:, :13].copy()).to(gpu)
    dof = torch.from_numpy(np.stack([states[:, 13:25], states[:, 25:37]], 2).copy()).to(gpu)
    tt = torch.from_numpy(tau).to(gpu)
    ff = torch.zeros((n, 12), dtype=torch.float32, device=gpu)
    par = _soa(rows, gpu)
    acc = eng.forward_dynamics(par, root, dof, tt, ff).cpu().numpy()
    ffg = ff.cpu().numpy().reshape(n, 4, 3)
    rel = []
    for i in range(n):
        a0, qdd, fo, _ = o.forward(states[i, :13], states[i, 13:25], states[i, 25:37], tau[i], rows[i])
        ref = np.r_[a0, qdd]
        rel.append(np.abs(acc[i] - ref).max() / max(np.abs(ref).max(), 1.0))
        assert np.abs(ffg[i] - fo).max() <= 1e-4 * max(np.abs(fo).max(), 1.0)
    rel = np.array(rel)
    print(f"forward dynamics: max rel acc err {rel.max():.3g}, median {np.median(rel):.3g}")
    assert rel.max() <= 1e-3


def test_simulate_matches_oracle(gpu, model, sim, theta0):
    """Record mode over 20 control steps (80 physics steps): stated max state error per horizon.
    Bounds: base position 1e-3 m, quaternion 1e-3, joint angles 5e-3 rad, velocities 0.05."""
    import torch
    from spi_amd.data import sinusoid_actions
    o, s = _oracle(model, sim)
    eng = _engine(s)
    rng = np.random.default_rng(5)
    n, T = 16, 20
    states = _realistic_states(model, s, theta0, n, seed=6)
    states[:, 25:37] = states[:, 25:37] * 0.2
    rows = _rand_params(rng, theta0, n, spread=0.05)
    acts = sinusoid_actions(rng, n, T, s.dt * s.decimation, s.action_scale)
    g = eng.simulate(_soa(rows, gpu), torch.from_numpy(states.astype(np.float32)).to(gpu),
                     torch.from_numpy(acts).to(gpu)).cpu().numpy()
    errs = []
    for i in range(n):
        ref, rc = o.simulate(rows[i].astype(np.float32).astype(np.float64), states[i].astype(np.float32), acts[i])
        assert rc == 0
        errs.append(np.abs(g[i] - ref))
    e = np.max(np.stack(errs), axis=0)  # [T, 37]
    print("per-step max err pos/quat/q/vel:", e[:, 0:3].max(1)[[0, 4, 9, 19]], e[:, 3:7].max(1)[[0, 4, 9, 19]],
          e[:, 13:25].max(1)[[0, 4, 9, 19]], np.concatenate([e[:, 7:13], e[:, 25:]], 1).max(1)[[0, 4, 9, 19]])
    assert e[:, 0:3].max() <= 1e-3
    assert e[:, 3:7].max() <= 1e-3
    assert e[:, 13:25].max() <= 5e-3
    assert np.concatenate([e[:, 7:13], e[:, 25:]], 1).max() <= 0.05


def _gt_segments(model, sim, theta, n_seg, H, seed, sigma=True):
    import spio
    from spi_amd.data import synthetic_segments
    o = spio.Oracle(model, sim)

    def simulate(rows, init, acts):
        out = []
        for r, i0, a in zip(rows, init, acts):
            traj, rc = o.simulate(np.asarray(r, np.float32).astype(np.float64), i0, a)
            out.append(traj)
        return np.stack(out)

    return synthetic_segments(simulate, theta, sim, n_seg, H, seed,
                              sigma_q=0.01 if sigma else 0.0, sigma_v=0.05 if sigma else 0.0)


@pytest.mark.parametrize("cost_mode,physics", [("tracking", "lanes"), ("reference", "lanes"), ("tracking", "legs"),
                                               ("reference", "legs")])
def test_rollout_cost_matches_oracle(gpu, model, sim, theta0, cost_mode, physics):
    """Fused rollout + cost, one rollout per lane and four lanes per rollout:
    |c_gpu - c_oracle| <= 2e-3 * c_oracle + 1e-4 for every (segment, candidate)."""
    import torch
    from spi_amd.rollout import DEFAULT_TRACKING_WEIGHTS, REFERENCE_WEIGHTS
    o, s = _oracle(model, sim)
    eng = _engine(s, cost_mode=cost_mode)
    w = DEFAULT_TRACKING_WEIGHTS if cost_mode == "tracking" else REFERENCE_WEIGHTS
    rng = np.random.default_rng(7)
    H, n_seg, n_cand = 5, 6, 12
    segs = _gt_segments(model, s, theta0, n_seg, H, seed=8)
    rows = _rand_params(rng, theta0, n_cand, spread=0.05)
    rows[0] = theta0
    g = eng.segment_costs(_soa(rows, gpu), n_cand, eng.upload_segments(segs), H,
                          physics=physics).cpu().numpy()  # [seg, cand]
    ref = o.rollout_cost(rows.astype(np.float32).astype(np.float64), segs.init, segs.actions, segs.targets,
                         0 if cost_mode == "tracking" else 1, np.asarray(w, np.float64)).T
    err = np.abs(g - ref)
    print(cost_mode, "max abs err", err.max(), "max rel", (err / np.maximum(ref, 1e-12)).max())
    assert np.all(err <= 2e-3 * ref + 1e-4), err.max()


def test_leg_parallel_costs_match_one_rollout_per_lane(gpu, model, sim, theta0):
    """spi_rollout_cost_legs against spi_rollout_cost on 300 candidates (20 % spread) x 20 noisy
    segments x H 10, both cost modes: the same diverged (+inf) entries; the order of the legs'
    sums is the only difference, a rounding-level change that contact-rich rollouts of
    far-off candidates amplify: median relative difference <= 1e-5, 99th percentile <= 1e-3,
    max <= 5e-2 (each form is held to the FP64 oracle in test_rollout_cost_matches_oracle)."""
    rng = np.random.default_rng(31)
    H, n_seg, n_cand = 10, 20, 300
    for mode in ("tracking", "reference"):
        _, s = _oracle(model, sim)
        eng = _engine(s, cost_mode=mode)
        segs = eng.upload_segments(_gt_segments(model, s, theta0, n_seg, H, seed=32))
        P = _soa(_rand_params(rng, theta0, n_cand, spread=0.2), gpu)
        a = eng.segment_costs(P, n_cand, segs, H, physics="lanes").cpu().numpy()
        b = eng.segment_costs(P, n_cand, segs, H, physics="legs").cpu().numpy()
        assert np.array_equal(np.isfinite(a), np.isfinite(b))
        f = np.isfinite(a)
        rel = np.abs(a[f] - b[f]) / np.maximum(np.abs(a[f]), 1e-6)
        print(mode, "relative difference legs vs lanes: median", np.median(rel), "p99", np.quantile(rel, 0.99),
              "max", rel.max())
        assert np.median(rel) <= 1e-5 and np.quantile(rel, 0.99) <= 1e-3 and rel.max() <= 5e-2


def test_landscape_argmin_matches_oracle(gpu, model, sim, theta0):
    """1-D base-mass landscape on noise-free data: same argmin grid index as the oracle, at the
    grid point nearest the ground truth."""
    import torch
    o, s = _oracle(model, sim)
    eng = _engine(s)
    H, n_seg = 5, 8
    segs = _gt_segments(model, s, theta0, n_seg, H, seed=9, sigma=False)
    masses = np.linspace(5.0, 9.0, 33)
    rows = np.repeat(theta0[None], masses.size, axis=0)
    rows[:, 0] = masses
    gc = eng.costs(_soa(rows, gpu), masses.size, eng.upload_segments(segs), H).cpu().numpy()
    oc = o.rollout_cost(rows, segs.init, segs.actions, segs.targets, 0,
                        np.asarray(eng.cargs.w[:], np.float64)).mean(axis=1)
    truth = int(np.argmin(np.abs(masses - theta0[0])))
    assert int(np.argmin(gc)) == int(np.argmin(oc)) == truth, (np.argmin(gc), np.argmin(oc), truth)


def test_fd_fim_trace_matches_reference_formula(gpu):
    """spi_fd_fim_trace vs a float64 transcription of active_sysid_openloop.py:402-426 (bmm + trace)."""
    import torch
    from spi_amd import _lib
    rng = np.random.default_rng(11)
    M, P = 37, 3
    N = M * (P + 1)
    root = rng.normal(0, 1, (N, 13)).astype(np.float32)
    dof = rng.normal(0, 1, (N, 12, 2)).astype(np.float32)
    origins = rng.normal(0, 5, (N, 3)).astype(np.float32)
    delta = 0.1
    main_idx = np.arange(0, N, P + 1)[:, None]
    aux_idx = main_idx + np.arange(1, P + 1)[None]
    rm = root[main_idx].astype(np.float64).copy()
    ra = root[aux_idx].astype(np.float64).copy()
    rm[..., :3] -= origins[main_idx]
    ra[..., :3] -= origins[aux_idx]
    X = np.concatenate([(rm - ra) / delta, (dof[main_idx, :, 0] - dof[aux_idx, :, 0]) / delta], axis=2)
    ref = np.repeat(np.trace(np.einsum("mpi,mpj->mij", X, X), axis1=1, axis2=2), P + 1)
    out = torch.zeros(N, dtype=torch.float32, device=gpu)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    root_d, dof_d, org_d = t(root), t(dof), t(origins)   # keep the device copies alive
    rc = _lib.lib().spi_fd_fim_trace(_lib.ptr(root_d), _lib.ptr(dof_d), _lib.ptr(org_d), M, P, delta,
                                     _lib.ptr(out), _lib.stream())
    _lib.check(rc, "fd_fim")
    assert np.allclose(out.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)


def test_set_rows_indexed_bit_exact(gpu):
    import torch
    from spi_amd import _lib
    rng = np.random.default_rng(12)
    n = 100
    dst = torch.from_numpy(rng.normal(size=(n, 13)).astype(np.float32)).to(gpu)
    src = torch.from_numpy(rng.normal(size=(n, 13)).astype(np.float32)).to(gpu)
    idx = np.array([0, 5, 99, 42, 5], dtype=np.int32)
    expect = dst.cpu().numpy().copy()
    expect[idx] = src.cpu().numpy()[idx]
    idx_d = torch.from_numpy(idx).to(gpu)
    rc = _lib.lib().spi_set_rows_indexed(_lib.ptr(dst), _lib.ptr(src), _lib.ptr(idx_d, torch.int32), idx.size, n, 13,
                                         _lib.stream())
    _lib.check(rc, "set_rows")
    assert np.array_equal(dst.cpu().numpy(), expect)


def test_rigid_body_states_match_oracle_fk(gpu, model, sim):
    """Positions of the 19 IsaacGym bodies vs the oracle's forward kinematics (1e-5 m)."""
    import torch
    import spio
    from spi_amd import _lib
    o = spio.Oracle(model, sim)
    rng = np.random.default_rng(13)
    n = 32
    states = np.stack([random_state(rng, sim) for _ in range(n)]).astype(np.float32)
    root = torch.from_numpy(states[:, :13].copy()).to(gpu)
    dof = torch.from_numpy(np.stack([states[:, 13:25], states[:, 25:37]], 2).copy()).to(gpu)
    out = torch.zeros((n, 19, 13), dtype=torch.float32, device=gpu)
    _lib.check(_lib.lib().spi_rigid_body_states(_lib.ptr(root), _lib.ptr(dof), n, _lib.ptr(out), _lib.stream()), "rb")
    g = out.cpu().numpy()
    names = [b["name"] for b in model["isaac_bodies"]]
    for i in range(n):
        lp, _ = o.fk(states[i, :13].astype(np.float64), states[i, 13:25].astype(np.float64))
        for k, name in enumerate(names):
            assert np.abs(g[i, k, :3] - lp[o.order.index(name)]).max() <= 1e-5


def test_cem_sample_matches_oracle_stream(gpu):
    """Device Philox + Box-Muller normals equal the oracle's stream (|dz| <= 2e-5)."""
    import torch
    import spio
    from spi_amd import _lib
    n_param, n = 5, 301
    mean = np.zeros(n_param, np.float32)
    std = np.ones(n_param, np.float32)
    lo, hi = -1e9 * np.ones(n_param, np.float32), 1e9 * np.ones(n_param, np.float32)
    t = lambda a: torch.from_numpy(a).to(gpu)
    theta = torch.zeros((n_param, n), dtype=torch.float32, device=gpu)
    seed, it = 0x1234567890ABCDEF, 7
    dm, ds, dl, dh = t(mean), t(std), t(lo), t(hi)    # keep the device copies alive
    _lib.check(_lib.lib().spi_cem_sample(_lib.ptr(dm), _lib.ptr(ds), _lib.ptr(dl), _lib.ptr(dh),
                                         n_param, n, seed, it, _lib.ptr(theta), n, _lib.stream()), "sample")
    ref = spio.cem_sample(mean, std, lo, hi, n, seed, it)
    assert np.abs(theta.cpu().numpy() - ref).max() <= 2e-5


@pytest.mark.parametrize("n_seg,n_cand", [(64, 256), (13, 300), (7, 5)])
def test_segment_reduction_matches_float64(gpu, sim, n_seg, n_cand):
    """cost[c] = sum_s w_s cost[s][c] in segment order: float64 sum to FP32 rounding, segments of
    weight 0 skipped even where their cost is +inf, +inf propagated from weighted segments; the
    segment counts cover whole groups of eight loads and a remainder."""
    import torch
    from spi_amd.rollout import RolloutEngine
    eng = RolloutEngine(sim)
    rng = np.random.default_rng(n_seg)
    cs = rng.gamma(2.0, 1.0, (n_seg, n_cand)).astype(np.float32)
    w = rng.uniform(0.0, 1.0, n_seg).astype(np.float32)
    w[::3] = 0.0
    cs[0, :] = np.inf                               # weight 0: must not reach the sum
    cs[1, 1] = np.inf                               # weighted: candidate 1 is +inf
    got = eng.reduce(torch.from_numpy(cs).cuda(), torch.from_numpy(w).cuda()).cpu().numpy()
    keep = w != 0.0
    ref = (w[keep, None].astype(np.float64) * cs[keep].astype(np.float64)).sum(0)
    assert np.isinf(got[1]) and np.isinf(ref[1])
    fin = np.isfinite(ref)
    assert np.all(np.isfinite(got[fin]))
    assert np.allclose(got[fin], ref[fin], rtol=2e-6, atol=0.0)


def test_cem_update_matches_numpy(gpu):
    """Elite set (sorted by cost, ties by index, NaN last) identical; moments within 1e-5."""
    import torch
    from spi_amd import _lib
    rng = np.random.default_rng(14)
    n_param, n, n_elite, alpha = 7, 3000, 300, 0.25
    cost = rng.normal(0, 1, n).astype(np.float32)
    cost[::97] = cost[3]          # ties
    cost[5] = np.nan
    theta = rng.normal(0, 1, (n_param, n)).astype(np.float32)
    mean0 = rng.normal(0, 1, n_param).astype(np.float32)
    std0 = rng.uniform(0.5, 1.0, n_param).astype(np.float32)
    min_std = np.full(n_param, 0.01, np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    mean, std = t(mean0.copy()), t(std0.copy())
    best, best_cost = torch.zeros(n_param, device=gpu), torch.full((1,), float("inf"), device=gpu)
    elite = torch.zeros(n_elite, dtype=torch.int32, device=gpu)
    dc, dth, dms = t(cost), t(theta), t(min_std)      # keep the device copies alive
    _lib.check(_lib.lib().spi_cem_update(_lib.ptr(dc), _lib.ptr(dth), n, n_param, n, n_elite, alpha,
                                         _lib.ptr(dms), _lib.ptr(mean), _lib.ptr(std), _lib.ptr(best),
                                         _lib.ptr(best_cost), _lib.ptr(elite, torch.int32), _lib.stream()), "update")
    key = np.where(np.isnan(cost), np.inf, cost)
    order = np.lexsort((np.arange(n), key))[:n_elite]
    assert np.array_equal(elite.cpu().numpy(), order)
    el = theta[:, order].astype(np.float64)
    m_ref = alpha * mean0 + (1 - alpha) * el.mean(1)
    s_ref = np.maximum(alpha * std0 + (1 - alpha) * el.std(1), min_std)
    assert np.allclose(mean.cpu().numpy(), m_ref, atol=1e-5)
    assert np.allclose(std.cpu().numpy(), s_ref, atol=1e-5)
    assert best_cost.item() == key[order[0]]
    assert np.array_equal(best.cpu().numpy(), theta[:, order[0]])


def _fim_oracle_group(o, sim, par, init, acts, delta, ksync, mode):
    """The reference's FD-FIM sum over T command steps for one group, stepped with the FP64
    oracle's torque law and physics step.  Command step t has the reference's step_idx t + 2
    (reset_all's zero-action step is step_idx 1, base_task.py:90-100); after it the aux envs
    sync to the main env when step_idx % ksync == 1 (active_sysid_openloop.py:246-252).
    mode "state" copies the whole state; mode "reference" only rewrites the aux envs' joint
    views (:316-330 via _reset_robot_states_callback), which the next step's first PD torque
    reads and the simulator never sees."""
    G = init.shape[0]
    st = [(e[:13].astype(np.float64), e[13:25].astype(np.float64), e[25:37].astype(np.float64)) for e in init]
    view = None
    tot = 0.0
    for t in range(acts.shape[0]):
        for k in range(G):
            root, q, qd = st[k]
            for d in range(sim.decimation):
                if view is not None and k > 0 and d == 0:
                    tau = o.control_torque(acts[t], view[0], view[1], par[k])
                else:
                    tau = o.control_torque(acts[t], q, qd, par[k])
                root, q, qd = o.physics_step(root, q, qd, tau, par[k])
            st[k] = (root, q, qd)
        view = None
        x0 = np.concatenate([st[0][0], st[0][1]])
        for k in range(1, G):
            X = (x0 - np.concatenate([st[k][0], st[k][1]])) / delta
            tot += float(X @ X)
        if ksync and (ksync == 1 or (t + 2) % ksync == 1):
            if mode == "state":
                st = [st[0]] + [tuple(a.copy() for a in st[0]) for _ in range(1, G)]
            else:
                view = (st[0][1].copy(), st[0][2].copy())
    return tot


@pytest.mark.parametrize("ksync,mode", [(0, "state"), (3, "state"), (5, "reference"), (1, "reference")])
def test_fim_rollout_matches_oracle(gpu, model, sim, theta0, ksync, mode):
    """Fused FD-FIM objective vs the oracle stepping main and aux envs and applying the
    reference formula per step, with the reference's sync schedule (step_idx counted from the
    reset step: with ksync 5 the syncs follow command steps 5 and 10) in both sync modes, and
    with syncs disabled: relative 1e-3.  The aux envs start from perturbed states so that a
    sync changes the result."""
    import torch
    from spi_amd.agents.active import FDFIMObjective
    from spi_amd.data import sinusoid_actions, standing_state
    o, s = _oracle(model, sim)
    eng = _engine(s)
    rng = np.random.default_rng(30)
    M, T, delta = 5, 12, 0.1
    obj = FDFIMObjective(eng, theta0, ["mass", "comx"], delta=delta, ksync_steps=max(ksync, 1), ksync_mode=mode)
    if ksync == 0:
        obj.ksync = 0
    settled, _ = o.simulate(theta0, standing_state(), np.zeros((40, 12), np.float32))
    G = 3
    init = np.repeat(settled[-1][None], M * G, axis=0).astype(np.float32)
    init[:, 13:37] += rng.normal(0, 0.02, (M * G, 24)).astype(np.float32)
    acts = sinusoid_actions(rng, M, T, 0.02, 0.25)
    g = obj(torch.from_numpy(acts).to(gpu), torch.from_numpy(init).to(gpu), mean=False).cpu().numpy()
    par = obj.params(M).cpu().numpy().T.astype(np.float64)
    for m in range(M):
        tot = _fim_oracle_group(o, s, par[m * G:(m + 1) * G], init[m * G:(m + 1) * G], acts[m], delta, ksync, mode)
        assert abs(g[m] - tot) <= 1e-3 * tot + 1e-3, (m, g[m], tot)
