"""Policy-in-the-loop active exploration on the device against the FP64 oracle
(oracle/spio_policy.py): the MFMA actor MLP vs float64 torch, the observation / gait / history
kernel, one control step with zeroed actions and the view-only joint sync, the termination +
FIM reward kernel, whole rollouts (both sync modes, with termination), and HIP-graph replay."""
import copy
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import random_state

pytestmark = pytest.mark.gpu


def _policy_sim(sim):
    s = copy.deepcopy(sim)
    s.control_mode, s.motor_model = 1, 3        # go2_omni hip scaling, A tanh(tau / A) motor model
    return s


def _dev(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _commands(rng, M, T):
    c = np.zeros((M, T, 14), np.float32)
    c[..., 0] = rng.uniform(-1, 1, (M, 1))
    c[..., 1] = rng.uniform(-0.6, 0.6, (M, 1))
    c[..., 2] = rng.uniform(-1, 1, (M, 1))
    c[..., 3] = rng.uniform(-0.25, 0.15, (M, 1))
    c[..., 4] = rng.uniform(2, 4, (M, 1))
    c[..., 5:8] = rng.choice([0.0, 0.5], size=(M, 1, 3))
    c[..., 8] = 0.5
    c[..., 9] = 0.03
    c[..., 10] = rng.uniform(-0.4, 0.4, (M, 1))
    c[..., 12] = rng.uniform(0.1, 0.45, (M, 1))
    c[..., 13] = rng.uniform(0.35, 0.45, (M, 1))
    return c


@pytest.mark.parametrize("M,dims", [(1000, [900, 512, 256, 128, 12]), (37, [70, 33, 12]), (5, [64, 1]),
                                    (9001, [900, 512, 256, 128, 12]), (8200, [70, 33, 12]),
                                    (2048, [900, 512, 256, 128, 12]), (600, [130, 257, 12]),
                                    (16400, [64, 96, 128, 12]), (517, [40, 64, 96, 20]),
                                    (300, [64, 200, 160, 12])])
def test_mlp_matches_float64(gpu, M, dims):
    """MFMA FP32 layers (ELU between) vs float64 torch on the CPU, ragged tile edges included;
    the shapes cover the K split in 1, 2 and 4 groups (2048 rows: layer 1 in 2 groups with a
    ragged last chunk; 1000 and 600 rows: 4 groups, 600 with scalar loads for odd K) and the
    last two layers in one launch (hidden width <= 128, output <= 32: the policy's 128 -> 12 with
    the K split in 4 (1000, 2048 rows), 2 (9001) and 1 group (16400); 96 hidden units leave a
    wavefront without columns (517 rows)); 160 hidden units take two launches (300 rows)."""
    from spi_amd.agents.policy import PolicyMLP
    net = PolicyMLP.random(dims[0], dims[-1], hidden=dims[1:-1], seed=3)
    x = torch.randn(M, dims[0], generator=torch.Generator().manual_seed(1)) * 2
    y = net(x.cuda()).cpu().double()
    h = x.double()
    for i, (W, b) in enumerate(net.layers()):
        h = h @ W.cpu().double().T + b.cpu().double()
        if i < net.n_layers - 1:
            h = torch.nn.functional.elu(h)
    err = (y - h).abs().max().item()
    print("max |MLP - float64|:", err, "scale", h.abs().max().item())
    assert err <= 1e-5 * max(1.0, h.abs().max().item()) * np.sqrt(dims[0])


def test_obs_kernel_matches_oracle(gpu, sim):
    """Three observation updates (gait clock, terms, clip, history) of 6 envs in 3 groups."""
    import spio_policy as op
    from spi_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(5)
    n, G, T = 6, 2, 5
    cmds = _commands(rng, n // G, T)
    cmds[0, :, 8] = 0.3                                 # a warped stance fraction
    q0 = np.asarray(sim.q0, np.float32)
    gait, hist = _dev(np.zeros(n)), _dev(np.zeros((n, 14, 60)))
    obs = torch.zeros((n, 900), device="cuda")
    g_ref = np.zeros(n)
    h_ref = [op.empty_history() for _ in range(n)]
    cmds_d, q0_d = _dev(cmds), _dev(q0)               # device buffers stay referenced while in use
    for step in range(3):
        states = np.stack([random_state(rng, sim) for _ in range(n)]).astype(np.float32)
        states[1, 25:37] = 5000.0                       # dof_vel obs 250 -> clipped
        applied = rng.uniform(-3, 3, (n, 12)).astype(np.float32)
        states_d, applied_d = _dev(states), _dev(applied)
        rc = L.spi_policy_obs(n, G, _lib.ptr(states_d), _lib.ptr(cmds_d), T, step, step + 1,
                              sim.dt * sim.decimation, _lib.ptr(q0_d), _lib.ptr(gait), _lib.ptr(applied_d),
                              _lib.ptr(hist), _lib.ptr(obs), 1, -0.25, 0.15, 100.0, _lib.stream())
        _lib.check(rc, "spi_policy_obs")
        got = obs.cpu().numpy()
        for i in range(n):
            g_ref[i], clock = op.gait_update(g_ref[i], cmds[i // G, step].astype(np.float64), sim.dt * sim.decimation)
            t = op.terms(states[i].astype(np.float64), cmds[i // G, step + 1].astype(np.float64), clock,
                         applied[i], g_ref[i], q0, evaluating=True)
            want, h_ref[i] = op.assemble(t, h_ref[i])
            assert np.abs(got[i] - want).max() < 2e-5, (step, i, np.argmax(np.abs(got[i] - want)))
        assert np.abs(gait.cpu().numpy() - g_ref).max() < 1e-6


@pytest.mark.parametrize("entry", ["spi_policy_step", "spi_policy_step_legs"])
def test_policy_step_matches_oracle(gpu, model, sim, theta0, entry):
    """One control step of 2 groups: group 1 terminated (zero actions), aux envs with the
    view-only joint sync on the first torque; one env per lane and four lanes per env."""
    import spio
    from spi_amd import _lib
    from spi_amd.rollout import RolloutEngine
    s = _policy_sim(sim)
    o = spio.Oracle(model, s)
    eng = RolloutEngine(s)
    L = _lib.lib()
    rng = np.random.default_rng(7)
    n, G = 4, 2
    par = np.repeat(np.asarray(theta0)[None], n, axis=0)
    par[1::2, 0] += 0.1
    states = np.stack([random_state(rng, s, spread=0.1, vel=0.2) for _ in range(n)]).astype(np.float32)
    actions = rng.uniform(-2, 2, (n, 12)).astype(np.float32)
    done = np.array([0, 0, 1, 1], np.int32)
    out, applied, bad = torch.zeros((n, 37), device="cuda"), torch.zeros((n, 12), device="cuda"), _dev(
        np.zeros(n), torch.int32)
    P, states_d, actions_d, done_d = _dev(par.T), _dev(states), _dev(actions), _dev(done, torch.int32)
    rc = getattr(L, entry)(C.byref(eng.args), _lib.ptr(P), n, n, G, _lib.ptr(states_d), _lib.ptr(out),
                           _lib.ptr(actions_d), _lib.ptr(done_d, torch.int32), 1,
                           _lib.ptr(applied), _lib.ptr(bad, torch.int32), _lib.stream())
    _lib.check(rc, entry)
    got, app = out.cpu().numpy(), applied.cpu().numpy()
    assert (bad.cpu().numpy() == 0).all()
    for i in range(n):
        a = np.zeros(12) if done[i] else np.clip(actions[i].astype(np.float64), -s.action_clip, s.action_clip)
        assert np.array_equal(app[i], a.astype(np.float32))
        root, q, qd = (states[i][:13].astype(np.float64), states[i][13:25].astype(np.float64),
                       states[i][25:37].astype(np.float64))
        m = i - i % G
        for d in range(s.decimation):
            if d == 0 and i != m:
                tau = o.control_torque(a, states[m][13:25], states[m][25:37], par[i])
            else:
                tau = o.control_torque(a, q, qd, par[i])
            root, q, qd = o.physics_step(root, q, qd, tau, par[i])
        want = np.concatenate([root, q, qd])
        assert np.abs(got[i] - want).max() < 2e-3, (i, np.abs(got[i] - want).max())


def test_reward_kernel_termination_and_fim(gpu):
    """Group 0 upright (FIM of two aux envs), group 1 with one aux rolled 60 deg (|g_y| = 0.87 >
    0.8: the whole group terminates and earns the termination reward), group 2 with a diverged
    member (flagged bad)."""
    from spi_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(9)
    M, G, delta, trew = 3, 3, 0.1, -1.5
    s = np.zeros((M * G, 37), np.float32)
    s[:, 6] = 1.0
    s[:, :3] = rng.normal(0, 0.05, (M * G, 3))
    s[:, 13:37] = rng.normal(0, 0.3, (M * G, 24))
    ang = np.pi / 3
    s[4, 3], s[4, 6] = np.sin(ang / 2), np.cos(ang / 2)
    done, bad = _dev(np.zeros(M * G), torch.int32), _dev([0] * 8 + [1], torch.int32)
    acc, gbad, s_d = _dev(np.full(M, 0.25)), _dev(np.zeros(M), torch.int32), _dev(s)
    rc = L.spi_policy_reward(_lib.ptr(s_d), M, G, delta, 0.8, 0.8, trew, _lib.ptr(done, torch.int32),
                             _lib.ptr(bad, torch.int32), _lib.ptr(acc), _lib.ptr(gbad, torch.int32), _lib.stream())
    _lib.check(rc, "spi_policy_reward")
    assert done.cpu().numpy().tolist() == [0, 0, 0, 1, 1, 1, 0, 0, 0]
    assert gbad.cpu().numpy().tolist() == [0, 0, 1]
    x = s.astype(np.float64)
    for g, a in enumerate(acc.cpu().numpy()):
        if g == 1:
            assert a == pytest.approx(0.25 + trew)
            continue
        f = sum(float(((x[g * G, :25] - x[g * G + k, :25]) / delta) @ ((x[g * G, :25] - x[g * G + k, :25]) / delta))
                for k in range(1, G))
        assert a == pytest.approx(0.25 + f, rel=1e-5)


def _rollout_case(model, sim, theta0, mode, seed, T, tilt=None, physics="auto"):
    import spio
    import spio_policy as op
    from spi_amd.agents.policy import PolicyFIMRollout, PolicyMLP, reset_states
    from spi_amd.data import standing_state
    from spi_amd.rollout import RolloutEngine
    s = _policy_sim(sim)
    o = spio.Oracle(model, s)
    eng = RolloutEngine(s)
    rng = np.random.default_rng(seed)
    M, G = 3, 2
    net = PolicyMLP.random(900, 12, seed=seed)
    with torch.no_grad():
        net.params.mul_(0.5)                      # a calm random policy
    roll = PolicyFIMRollout(eng, net, theta0, ["mass"], delta=0.1, ksync_steps=5, ksync_mode=mode,
                            termination_rew=-2.0, physics=physics)
    settled, _ = o.simulate(theta0, standing_state(0.32), np.zeros((30, 12), np.float32))
    init = np.repeat(settled[-1][None], M * G, axis=0)
    init[:, 7:13] += reset_states(M * G, s.q0, seed=seed, vel_range=0.1)[:, 7:13]
    if tilt is not None:
        ang = np.pi / 2
        init[tilt * G:(tilt + 1) * G, 3], init[tilt * G:(tilt + 1) * G, 6] = np.sin(ang / 2), np.cos(ang / 2)
    cmds = _commands(rng, M, T)
    fim = roll.evaluate(_dev(cmds), _dev(init)).cpu().numpy()
    par = roll.params(M).cpu().numpy().T.astype(np.float64)
    layers = [(W.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64)) for W, b in net.layers()]
    fr, final, done = op.policy_rollout(o, par, init.astype(np.float64), cmds.astype(np.float64), layers, G, 0.1,
                                        ksync=5, ksync_mode=mode, termination_rew=-2.0, q_default=s.q0,
                                        action_clip=s.action_clip, dt_ctrl=s.dt * s.decimation,
                                        decimation=s.decimation)
    return roll, fim, fr, final, done


@pytest.mark.parametrize("mode,physics", [("reference", "lanes"), ("state", "lanes"), ("reference", "legs"),
                                          ("state", "legs")])
def test_policy_rollout_matches_oracle(gpu, model, sim, theta0, mode, physics):
    """16 control steps (reset step + 14 policy steps) of 3 groups under a random policy, with
    one env per lane and with four lanes per env."""
    roll, fim, fr, final, done = _rollout_case(model, sim, theta0, mode, seed=11, T=16, physics=physics)
    got = roll.final_states().cpu().numpy()
    print(mode, "fim", fim, "oracle", fr, "max state err", np.abs(got - final).max())
    assert np.all(np.isfinite(fim))
    assert np.allclose(fim, fr, rtol=5e-3, atol=1e-3)
    assert np.abs(got - final).max() < 2e-2
    d, _, _ = roll.flags()
    assert np.array_equal(d.cpu().numpy().astype(bool), done)


def test_policy_rollout_termination(gpu, model, sim, theta0):
    """Group 1 starts rolled 90 deg: it terminates (the whole group) and terminated steps earn the
    termination reward; the other groups follow the oracle as before."""
    roll, fim, fr, final, done = _rollout_case(model, sim, theta0, "reference", seed=12, T=10, tilt=1)
    assert done[2] and done[3]
    assert fr[1] <= -2.0 / 9 + 1e-9                 # at least one terminated step
    assert np.allclose(fim, fr, rtol=5e-3, atol=1e-3)
    d, _, _ = roll.flags()
    assert np.array_equal(d.cpu().numpy().astype(bool), done)


def test_graph_replay_matches_eager(gpu, model, sim, theta0):
    """The captured rollout replays bit-identically to the eager launches, also after new
    commands and initial states are written into its buffers."""
    from spi_amd.agents.policy import PolicyFIMRollout, PolicyMLP
    from spi_amd.data import standing_state
    from spi_amd.rollout import RolloutEngine
    s = _policy_sim(sim)
    eng = RolloutEngine(s)
    rng = np.random.default_rng(13)
    M, T = 64, 12
    net = PolicyMLP.random(900, 12, seed=2)
    roll = PolicyFIMRollout(eng, net, theta0, ["mass", "comx"])
    init = _dev(standing_state(0.32))
    outs = []
    for rep in range(2):
        cmds = _dev(_commands(rng, M, T))
        eager = roll.evaluate(cmds, init).clone()
        torch.cuda.synchronize()
        for _ in range(2):
            g = roll.evaluate(cmds, init, graph=True).clone()
            torch.cuda.synchronize()
            assert torch.equal(torch.nan_to_num(g, nan=-7.0), torch.nan_to_num(eager, nan=-7.0)), rep
        outs.append(eager)
    assert not torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("M,T", [(40, 40), (4100, 18)])
def test_ring_observation_equals_dense(gpu, model, sim, theta0, M, T):
    """The ring layout (terms written once, the actor's first layer gathers the 900-wide row) and
    the dense layout (materialised observation + shifted history) give bit-identical rollouts:
    the gathered operands are the same floats in the same K order.  40 steps wrap the 15-slot
    ring twice; 8200 envs take the two-row-block tiles."""
    from spi_amd.agents.policy import PolicyFIMRollout, PolicyMLP, reset_states
    from spi_amd.rollout import RolloutEngine
    s = _policy_sim(sim)
    eng = RolloutEngine(s)
    rng = np.random.default_rng(21)
    net = PolicyMLP.random(900, 12, seed=4)
    init = _dev(reset_states(M * 2, s.q0, seed=3))
    cmds = _dev(_commands(rng, M, T))
    out = {}
    for layout in ("ring", "dense"):
        roll = PolicyFIMRollout(eng, net, theta0, ["mass"], obs_layout=layout)
        out[layout] = (roll.evaluate(cmds, init).clone(), roll.final_states().clone())
    assert torch.equal(torch.nan_to_num(out["ring"][0], nan=-7.0), torch.nan_to_num(out["dense"][0], nan=-7.0))
    assert torch.equal(out["ring"][1], out["dense"][1])


def test_stream_slices_equal_one_stream(gpu, model, sim, theta0):
    """Slices of the groups on separate streams (physics of one slice overlapping the actor
    GEMMs of another) give the same bits as one stream, eagerly and from a captured graph; the
    5-group batch makes uneven, unaligned slices."""
    from spi_amd.agents.policy import PolicyFIMRollout, PolicyMLP, reset_states
    from spi_amd.rollout import RolloutEngine
    s = _policy_sim(sim)
    eng = RolloutEngine(s)
    rng = np.random.default_rng(23)
    net = PolicyMLP.random(900, 12, seed=5)
    for M, T in ((64, 24), (5, 20)):
        init = _dev(reset_states(M * 3, s.q0, seed=M))
        cmds = _dev(_commands(rng, M, T))
        ref = None
        for n_streams in (1, 2, 3):
            roll = PolicyFIMRollout(eng, net, theta0, ["mass", "comx"], n_streams=n_streams)
            for graph in (False, True):
                fim = roll.evaluate(cmds, init, graph=graph).clone()
                fin = roll.final_states().clone()
                torch.cuda.synchronize()
                if ref is None:
                    ref = (fim, fin)
                    continue
                assert torch.equal(torch.nan_to_num(fim, nan=-7.0), torch.nan_to_num(ref[0], nan=-7.0)), (M, n_streams)
                assert torch.equal(fin, ref[1]), (M, n_streams, graph)


def test_leg_parallel_step_matches_one_env_per_lane(gpu, sim, theta0):
    """The four-lanes-per-env kernel against the one-env-per-lane kernel on 1000 random envs in
    groups of 2 (some terminated, view sync on): same states up to the rounding of the legs'
    sums, the same applied actions bit for bit and the same divergence flags."""
    from spi_amd import _lib
    from spi_amd.rollout import RolloutEngine
    s = _policy_sim(sim)
    eng = RolloutEngine(s)
    L = _lib.lib()
    rng = np.random.default_rng(17)
    n, G = 1000, 2
    par = np.repeat(np.asarray(theta0)[None], n, axis=0)
    par[:, 0] *= rng.uniform(0.8, 1.2, n)
    par[:, 10:18] *= rng.uniform(0.8, 1.2, (n, 8))
    states = np.stack([random_state(rng, s, spread=0.2, vel=0.3) for _ in range(n)]).astype(np.float32)
    actions = rng.uniform(-3, 3, (n, 12)).astype(np.float32)
    done = (rng.uniform(size=n) < 0.1).astype(np.int32)
    P, states_d, actions_d, done_d = _dev(par.T), _dev(states), _dev(actions), _dev(done, torch.int32)
    res = {}
    for entry in ("spi_policy_step", "spi_policy_step_legs"):
        out, applied = torch.zeros((n, 37), device="cuda"), torch.zeros((n, 12), device="cuda")
        bad = _dev(np.zeros(n), torch.int32)
        rc = getattr(L, entry)(C.byref(eng.args), _lib.ptr(P), n, n, G, _lib.ptr(states_d), _lib.ptr(out),
                               _lib.ptr(actions_d), _lib.ptr(done_d, torch.int32), 1, _lib.ptr(applied),
                               _lib.ptr(bad, torch.int32), _lib.stream())
        _lib.check(rc, entry)
        res[entry] = (out.cpu().numpy(), applied.cpu().numpy(), bad.cpu().numpy())
    a, b = res["spi_policy_step"], res["spi_policy_step_legs"]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    scale = np.abs(a[0]).max(0) + 1e-3
    err = (np.abs(a[0] - b[0]) / scale).max()
    print("max state difference / component scale:", err)
    assert err < 1e-4
