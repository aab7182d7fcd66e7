"""The device reset draw (spi_reset_states) and the policy-driven optimizer loop that uses it.

Reset: the reference resets every env at the start of each rollout (reset_all, base_task.py:90-100):
joint positions q_default * U[0.5, 1.5], zero joint velocities (active_sysid_openloop.py:274-292),
base at the init position plus the env origin with identity orientation, linear and angular
velocity U[-0.5, 0.5] (legged_robot_base.py:761-784).  torch's generator cannot be reproduced,
so the draw is pinned by a CPU restatement of its Philox mapping on a few envs and by the
distribution over a million envs.

Optimizer: PolicyActiveSysId.optimize (active_sysid.py:165-257): best commands = the generator
applied to the best parameters, re-evaluating them reproduces the best value, and
best_commands.npz is written with the reference's key (:229-231)."""
import copy

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _expected(n, q_default, seed, base_height=0.34, vel=0.5, lo=0.5, hi=1.5):
    """CPU restatement of the draw: uniform k of env e is word k % 4 of Philox4x32-10 block k // 4
    with counter (e, 0, block, 'RSET') and key (seed low, seed high), mapped to (x + 0.5) 2^-32."""
    import spio
    out = np.zeros((n, 37))
    out[:, 2] = base_height
    out[:, 6] = 1.0
    for e in range(n):
        u = []
        for b in range(5):
            words = spio.philox4x32_10([e, 0, b, 0x52534554], [seed & 0xFFFFFFFF, seed >> 32])
            u += [(np.float32(w) + np.float32(0.5)) * np.float32(2.3283064365386963e-10) for w in words]
        u = np.asarray(u, np.float64)
        out[e, 13:25] = np.asarray(q_default) * (lo + (hi - lo) * u[:12])
        out[e, 7:13] = -vel + 2 * vel * u[12:18]
    return out


def test_reset_draw_matches_restatement_and_distribution(gpu, sim):
    from spi_amd.agents.policy import reset_states_device
    q0 = np.asarray(sim.q0, np.float64)
    seed = (7 << 32) | 12345
    small = reset_states_device(64, q0, seed=seed).cpu().numpy()
    exp = _expected(64, q0, seed)
    assert np.abs(small - exp).max() <= 1e-6 * (1 + np.abs(exp)).max()
    assert np.array_equal(small[:, 25:], np.zeros((64, 12), np.float32))

    n = 1 << 20
    origins = torch.from_numpy(np.random.default_rng(0).normal(size=(n, 3)).astype(np.float32)).cuda()
    big = reset_states_device(n, q0, seed=seed, origins=origins).cpu().numpy()
    assert np.array_equal(big[:64, 3:], small[:, 3:])                    # env e's draw ignores n_env
    assert np.array_equal(big[:, :3], (small[:1, :3] + origins.cpu().numpy()).astype(np.float32))
    assert np.all(big[:, 3:7] == np.array([0, 0, 0, 1], np.float32))
    ratio = big[:, 13:25] / q0.astype(np.float32)[None]
    vel = big[:, 7:13]
    assert ratio.min() >= 0.5 - 1e-6 and ratio.max() <= 1.5 + 1e-6
    assert vel.min() >= -0.5 and vel.max() <= 0.5
    # uniform moments: mean 1 / 0 and variance 1/12 per column, within 5 standard errors
    for x, mu in ((ratio, 1.0), (vel, 0.0)):
        m = x.mean(axis=0)
        v = x.var(axis=0)
        assert np.all(np.abs(m - mu) <= 5 * np.sqrt(1 / 12 / n)), m
        assert np.all(np.abs(v - 1 / 12) <= 5 * np.sqrt(1 / 180 / n)), v
    # Kolmogorov-Smirnov distance of every column against U[0, 1] (critical value at 1e-3: 1.95 / sqrt(n))
    u = np.concatenate([ratio - 0.5, vel + 0.5], axis=1)
    grid = (np.arange(n) + 1.0) / n
    for k in range(u.shape[1]):
        d = np.abs(np.sort(u[:, k]) - grid).max()
        assert d <= 1.95 / np.sqrt(n), (k, d)
    # columns and neighbouring envs uncorrelated
    c = np.corrcoef(u[: n // 2].T)
    assert np.abs(c - np.eye(18)).max() <= 5 / np.sqrt(n // 2)
    assert abs(np.corrcoef(u[:-1, 0], u[1:, 0])[0, 1]) <= 5 / np.sqrt(n)
    other = reset_states_device(64, q0, seed=seed + 1).cpu().numpy()
    assert not np.array_equal(other[:, 13:25], small[:, 13:25])


def _optimizer(sim, theta0, batch, seed, init=None, num_rollouts=1):
    from spi_amd.agents.commands import CommandGenerator
    from spi_amd.agents.policy import PolicyActiveSysId, PolicyFIMRollout, PolicyMLP
    from spi_amd.rollout import RolloutEngine
    s = copy.deepcopy(sim)
    s.control_mode, s.motor_model = 1, 3
    eng = RolloutEngine(s)
    net = PolicyMLP.random(900, 12, seed=2)
    roll = PolicyFIMRollout(eng, net, theta0, ["mass"], ksync_steps=5)
    ranges = [[-1, 1], [-1, 1], [-1, 1], [-0.25, 0.15], [2, 4], [0, 1], [0, 1], [0, 1], [0.5, 0.5], [0.03, 0.035],
              [-0.4, 0.4], [0, 0], [0.1, 0.45], [0.35, 0.45]]
    default = [0.5, 0.5, 0.5, 0.0, 3.0, 0.0, 0.5, 0.5, 0.5, 0.03, 0.0, 0.0, 0.3, 0.4]
    gen = CommandGenerator(ranges, default, [0, 2, 5], "constant", rollout_length=0.6, horizon_length=0.2,
                           dt=s.dt * s.decimation)
    return PolicyActiveSysId(roll, gen, batch_size=batch, seed=seed, init=init, num_rollouts=num_rollouts), gen


def test_policy_optimize_best_commands(gpu, sim, theta0, tmp_path):
    """Two iterations of 16 trials from one fixed reset state: best_commands equals the generator
    applied to best_params, best_commands.npz holds it under the reference's key, and the best
    commands re-evaluated from the same state reproduce best_value bit for bit."""
    from spi_amd.agents.policy import reset_states_device
    B = 16
    one = reset_states_device(1, sim.q0, seed=5)
    init = one.expand(B * 2, 37).contiguous()
    drv, gen = _optimizer(sim, theta0, B, seed=3, init=init)
    res = drv.optimize(2, log_dir=str(tmp_path))
    assert len(drv.history) == 2 and drv.draws == 0
    assert np.isfinite(res["best_value"])
    names = gen.param_names()
    assert sorted(res["best_params"]) == sorted(names)
    again = gen.from_params(res["best_params"])
    assert np.array_equal(res["best_commands"], again)
    with np.load(res["best_commands_path"], allow_pickle=False) as f:
        assert list(f.files) == ["best_commands"]
        assert np.array_equal(f["best_commands"], res["best_commands"])
    assert res["best_commands"].shape == (gen.total_steps, 14)
    # every group replays the best commands from the same state: each one reproduces best_value
    cmds = torch.from_numpy(np.repeat(res["best_commands"][None], B, axis=0)).cuda()
    fim = drv.rollout.evaluate(cmds, init, graph=True)
    assert torch.all(-fim == res["best_value"]), (fim, res["best_value"])
    seen = torch.cat([-h for h in drv.history]).cpu().numpy()
    assert res["best_value"] == np.nanmin(seen)


def test_policy_optimize_redraws_reset_per_rollout(gpu, sim, theta0):
    """Without a fixed initial state every rollout starts from a fresh device reset draw (the
    reference resets before each of its num_rollouts_per_command rollouts): 2 iterations x 2
    rollouts = 4 draws, the last one equal to the draw of its Philox key."""
    from spi_amd.agents.policy import reset_states_device
    drv, _ = _optimizer(sim, theta0, 8, seed=11, num_rollouts=2)
    drv.optimize(2)
    assert drv.draws == 4
    last = reset_states_device(16, sim.q0, seed=drv.reset_key(3))
    assert torch.equal(drv._reset_buf, last)
    first = reset_states_device(16, sim.q0, seed=drv.reset_key(0))
    assert not torch.equal(first, last)
