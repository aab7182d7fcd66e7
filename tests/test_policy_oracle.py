"""The numpy restatement of the policy observation and the policy-in-the-loop rollout (the
oracle for spi_policy_* and spi_mlp_forward) against what pins it independently: the
reference's observation config (golden fixture), scipy's rotations, the gait semantics of the
phase / offset / bound parameters (trot, pace, bound, pronk pair the legs as their names say),
the history semantics (newest first), torch's Linear + ELU, and, for a policy whose output is
zero, the open-loop finite-difference FIM computed from the oracle's own control-step simulate."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "go2_omni_actor_obs.json")


def test_layout_matches_reference_config():
    import spio_policy as op
    with open(GOLDEN) as f:
        ref = json.load(f)
    terms = [k for k in ref["actor_obs"] if k != "short_history"]
    assert sorted(terms) == sorted(op.TERMS)
    for k in terms:
        assert op.TERMS[k] == (ref["dims"][k], ref["scales"][k]), k
    assert set(ref["short_history"]) == set(op.TERMS)
    assert set(ref["short_history"].values()) == {op.HISTORY}
    assert op.CLIP == ref["clip_observations"]
    # actor obs = sorted keys, and "short_history" sorts after every term key
    assert sorted(ref["actor_obs"])[-1] == "short_history"
    assert op.obs_dim() == sum(ref["dims"].values()) * (1 + op.HISTORY) == 900


def test_quat_rotate_inverse_matches_scipy():
    from scipy.spatial.transform import Rotation
    import spio_policy as op
    rng = np.random.default_rng(0)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        v = rng.normal(size=3)
        assert np.allclose(op.quat_rotate_inverse(q, v), Rotation.from_quat(q).inv().apply(v), atol=1e-12)


@pytest.mark.parametrize("gait,pairs", [
    ((0.5, 0.0, 0.0), [(0, 3), (1, 2)]),     # trot: diagonal legs FL-RR, FR-RL
    ((0.0, 0.0, 0.5), [(0, 2), (1, 3)]),     # pace: lateral legs FL-RL, FR-RR
    ((0.0, 0.5, 0.0), [(0, 1), (2, 3)]),     # bound: front pair, rear pair
    ((0.0, 0.0, 0.0), [(0, 1), (0, 2), (0, 3)]),  # pronk: all together
])
def test_gait_pairs_legs_as_named(gait, pairs):
    """Foot order FL, FR, RL, RR; paired legs share the clock, unpaired legs are half a cycle apart."""
    import spio_policy as op
    cmd = np.zeros(14)
    cmd[4], cmd[8] = 3.0, 0.5
    cmd[5:8] = gait
    g = 0.0
    for _ in range(7):
        g, clock = op.gait_update(g, cmd, 0.02)
        for a, b in pairs:
            assert abs(clock[a] - clock[b]) < 1e-12
        if gait != (0.0, 0.0, 0.0):
            a0, b0 = pairs[0][0], pairs[1][0]
            assert abs(clock[a0] + clock[b0]) < 1e-12   # sin(x + pi) = -sin(x)


def test_gait_warp_is_monotone_and_continuous():
    """Stance fraction d maps [0, d) onto [0, 0.5) and (d, 1) onto (0.5, 1); at d = 0.5 the warp is
    the identity, so the clock is sin(2 pi frac(g + phase))."""
    import spio_policy as op
    cmd = np.zeros(14)
    cmd[4] = 1.0
    for dur in (0.5, 0.3, 0.7):
        cmd[8] = dur
        vals = []
        for g0 in np.linspace(0.0, 0.98, 50):
            _, clock = op.gait_update(g0, cmd, 0.0)
            vals.append(np.arcsin(np.clip(clock[1], -1, 1)))
        if dur == 0.5:
            for g0 in np.linspace(0.0, 0.98, 11):
                _, clock = op.gait_update(g0, cmd, 0.0)
                assert np.allclose(clock, np.sin(2 * np.pi * g0), atol=1e-12)
    # boundaries of the warp: r -> d- gives 0.5, r -> d+ gives 0.5
    cmd[8] = 0.3
    eps = 1e-9
    _, c1 = op.gait_update(0.3 - eps, cmd, 0.0)
    _, c2 = op.gait_update(0.3 + eps, cmd, 0.0)
    assert abs(c1[1] - c2[1]) < 1e-6 and abs(c1[1]) < 1e-6     # sin(pi) = 0


def test_history_newest_first():
    import spio_policy as op
    hist = op.empty_history()
    seq = []
    for s in range(4):
        t = {k: np.full(d, 10.0 * s + i) for i, (k, (d, _)) in enumerate(sorted(op.TERMS.items()))}
        obs, hist = op.assemble(t, hist)
        seq.append(t)
    # after 4 adds, history position p of each key holds the terms added p + 1 steps before the last
    keys = sorted(op.TERMS)
    off = 60
    for k in keys:
        d = op.TERMS[k][0]
        block = obs[off:off + d * op.HISTORY].reshape(op.HISTORY, d)
        for p in range(op.HISTORY):
            want = seq[2 - p][k] if p <= 2 else np.zeros(d)
            assert np.array_equal(block[p], want), (k, p)
        off += d * op.HISTORY
    assert off == 900
    cur = np.concatenate([seq[3][k] for k in keys])
    assert np.array_equal(obs[:60], cur)


def test_clip_and_evaluation_height():
    import spio_policy as op
    s = np.zeros(37)
    s[6] = 1.0
    s[25:37] = 1e4                           # dof_vel * 0.05 = 500 -> clipped to 100
    cmd = np.zeros(14)
    cmd[3] = 0.1
    t = op.terms(s, cmd, np.zeros(4), np.zeros(12), 0.25, np.zeros(12), evaluating=True)
    assert np.allclose(t["projected_gravity"], [0, 0, -1])
    assert t["command_body_height"][0] == pytest.approx(2.0 * 0.15)   # 0.1 + 0.2 sin(pi/2) clamped
    t2 = op.terms(s, cmd, np.zeros(4), np.zeros(12), 0.25, np.zeros(12), evaluating=False)
    assert t2["command_body_height"][0] == pytest.approx(0.2)
    obs, _ = op.assemble(t, op.empty_history())
    assert obs.max() == 100.0


def test_mlp_matches_torch():
    import torch
    import spio_policy as op
    g = torch.Generator().manual_seed(0)
    net = torch.nn.Sequential(torch.nn.Linear(30, 16), torch.nn.ELU(), torch.nn.Linear(16, 8), torch.nn.ELU(),
                              torch.nn.Linear(8, 5)).double()
    x = torch.randn(7, 30, generator=g, dtype=torch.float64) * 3
    layers = [(m.weight.detach().numpy(), m.bias.detach().numpy()) for m in net if isinstance(m, torch.nn.Linear)]
    assert np.allclose(op.mlp(layers, x.numpy()), net(x).detach().numpy(), atol=1e-12)


def test_zero_policy_rollout_is_open_loop_fim(model, sim, theta0):
    """A policy with zero output drives every env with zero actions: the rollout reward equals the
    open-loop FD-FIM of zero actions (state-copy sync), stepped with spio.Oracle.simulate."""
    import spio
    import spio_policy as op
    from spi_amd.data import standing_state
    o = spio.Oracle(model, sim)
    settled, _ = o.simulate(theta0, standing_state(), np.zeros((40, 12), np.float32))
    init = settled[-1].astype(np.float64)
    M, G, T, delta, ks = 2, 3, 9, 0.1, 3
    par = np.repeat(np.asarray(theta0, np.float64)[None], M * G, axis=0)
    for m in range(M):
        par[m * G + 1, 0] += delta
        par[m * G + 2, 1] += delta
    layers = [(np.zeros((12, 900)), np.zeros(12))]
    cmds = np.zeros((M, T, 14))
    cmds[:, :, 4], cmds[:, :, 8] = 3.0, 0.5
    inits = np.repeat(init[None], M * G, axis=0)
    fim, final, done = op.policy_rollout(o, par, inits, cmds, layers, G, delta, ksync=ks, ksync_mode="state",
                                         q_default=sim.q0, action_clip=sim.action_clip, decimation=sim.decimation)
    assert not done.any()
    for m in range(M):
        states = [init.copy() for _ in range(G)]
        tot = 0.0
        for t in range(T - 1):
            for k in range(G):
                out, _ = o.simulate(par[m * G + k], states[k].astype(np.float32), np.zeros((1, 12), np.float32))
                states[k] = out[0].astype(np.float64)
            if t >= 1:
                for k in range(1, G):
                    X = (states[0][:25] - states[k][:25]) / delta
                    tot += float(X @ X)
            if (t + 1) % ks == 1:
                for k in range(1, G):
                    states[k] = states[0].copy()
        assert fim[m] == pytest.approx(tot / (T - 1), rel=1e-4), m


def test_default_param_rows_follow_the_reference_setters():
    """env/active_sysid_openloop.yaml default_param: mass 9.39 on the base, CoM 0, inertia 0.005,
    motor A 20; 'inertiay' stays unapplied (the reference's setter is set_inertiaiy) unless asked."""
    import yaml
    from spi_amd.agents.policy import theta_from_default_param
    from spi_amd.params import P_COMX, P_IXX, P_IYY, P_IZZ, P_MASS, P_MOTOR0, nominal_params
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "..", "spi-active_amd", "spi_amd", "configs", "env",
                           "active_sysid_openloop.yaml")) as f:
        env = yaml.safe_load(f)["env"]["config"]
    nom = nominal_params()
    th = theta_from_default_param(env["default_param"], nom)
    assert th[P_MASS] == 9.39 and np.all(th[P_COMX:P_COMX + 3] == 0.0)
    assert th[P_IXX] == 0.005 and th[P_IZZ] == 0.005 and th[P_IYY] == nom[P_IYY]
    assert np.all(th[P_MOTOR0:P_MOTOR0 + 3] == 20.0)
    assert theta_from_default_param(env["default_param"], nom, apply_inertiay=True)[P_IYY] == 0.005
    with open(os.path.join(here, "..", "spi-active_amd", "spi_amd", "configs", "algo", "active_sysid.yaml")) as f:
        algo = yaml.safe_load(f)["algo"]["config"]
    assert algo["command"]["command_sampling_idxs"] == [0, 2, 5]


def test_stream_slices_cover_the_groups():
    """Host logic of the stream slices: contiguous cover of [0, M), Python ints (they become
    pointer offsets), slice starts on 4-env boundaries whenever the group count allows."""
    import types
    from spi_amd.agents.policy import PolicyFIMRollout
    for G in (2, 3, 4, 5):
        for P in (1, 2, 3):
            for M in (1, 2, 5, 64, 1024, 8191):
                parts = PolicyFIMRollout._parts(types.SimpleNamespace(G=G, n_streams=P), M)
                assert parts[0][0] == 0 and parts[-1][1] == M and len(parts) <= P
                assert all(a < b and type(a) is int and type(b) is int for a, b in parts)
                assert all(parts[i][1] == parts[i + 1][0] for i in range(len(parts) - 1))
                assert all((a * G) % 4 == 0 for a, _ in parts)
