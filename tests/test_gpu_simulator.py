"""The BaseSimulator face (HipSim / HipSimActiveSysId) on the GPU: construction sequence of
spigym/envs/base_task/base_task.py:20-66, per-env parameters from params_dict, IsaacGym's
tensor/refresh semantics and the gym calls of scripts/eval.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _make_sim(n_main=3, params=("mass",), delta=0.1, apply_inertiay=False):
    import torch
    from spi_amd.config import compose, get_class
    cfg = compose("active_sysid")
    cfg.simulator.config.apply_inertiay = apply_inertiay
    P = len(params)
    N = n_main * (P + 1)
    # params_dict exactly as active_sysid_openloop.py:51-71 builds it
    pd = {}
    for key, value in cfg.env.config.default_param.items():
        if "motor_model" in key:
            pd[key] = {"value": [value[0]] * N}
        else:
            pd[key] = {"body_name": value[0], "value": [value[1]] * N}
    idx = np.arange(N) % (P + 1)
    for i, p in enumerate(params):
        v = np.array(pd[p]["value"])
        v[idx == i + 1] += delta
        pd[p]["value"] = v.tolist()
    cfg["params_dict"] = pd
    cls = get_class(cfg.simulator._target_)
    sim = cls(config=cfg, device="cuda:0")
    sim.set_headless(True)
    sim.setup()
    sim.setup_terrain("plane")
    sim.load_assets()
    origins = torch.zeros((N, 3), device="cuda:0")
    origins[:, 0] = 5.0 * torch.arange(N, device="cuda:0")
    init = [0.0, 0.0, 0.34, 0.0, 0.0, 0.0, 1.0] + [0.0] * 6
    sim.create_envs(N, origins, init)
    sim.get_dof_limits_properties()
    sim.prepare_sim()
    return sim, cfg, N


def test_construction_sequence_and_params_dict(gpu):
    sim, cfg, N = _make_sim()
    assert sim.num_dof == 12 and sim.num_bodies == 19
    assert sim.find_rigid_body_indice("FL_foot") == 4
    p = sim.params.cpu().numpy()
    assert np.allclose(p[0], [9.39, 9.49] * 3, atol=1e-5)        # main 9.39, aux +0.1 (absolute delta)
    assert np.allclose(p[1:4], 0.0) and np.allclose(p[4], 0.005) and np.allclose(p[6], 0.005)
    assert np.allclose(p[22:25], 20.0)
    # reference quirk: 'inertiay' has no setter there -> Iyy keeps the URDF value
    assert np.allclose(p[5], 0.098077, atol=1e-6)
    sim2, _, _ = _make_sim(apply_inertiay=True)
    assert np.allclose(sim2.params.cpu().numpy()[5], 0.005)


def test_step_equals_engine_and_refresh_semantics(gpu):
    import torch
    sim, cfg, N = _make_sim()
    rng = np.random.default_rng(0)
    root = sim.all_root_states.clone()
    root[:, 2] = 0.30
    dof = sim.dof_state.view(N, 12, 2).clone()
    dof[:, :, 0] = torch.as_tensor(np.asarray(sim.sim_params.q0, np.float32), device="cuda:0")
    ids = torch.arange(N, device="cuda:0")
    sim.set_actor_root_state_tensor(ids, root)
    sim.set_dof_state_tensor(ids, dof.view(N * 12, 2))
    sim.refresh_sim_tensors()
    tau = torch.as_tensor(rng.normal(0, 3, (N, 12)).astype(np.float32), device="cuda:0")
    ref_root, ref_dof = sim.all_root_states.clone(), sim.dof_state.view(N, 12, 2).clone()
    sim.engine.physics_step(sim.params, ref_root, ref_dof, tau.clone())
    sim.apply_torques_at_dof(tau)
    sim.simulate_at_each_physics_step()
    # dof view refreshed after every physics step; root view only on refresh_sim_tensors
    assert torch.equal(sim.dof_state.view(N, 12, 2), ref_dof)
    assert torch.equal(sim.all_root_states, root)
    sim.refresh_sim_tensors()
    assert torch.equal(sim.all_root_states, ref_root)
    # a write to the view that is not pushed is overwritten by the next step (IsaacGym semantics)
    sim.dof_pos[:] = 0.0
    sim.simulate_at_each_physics_step()
    assert sim.dof_pos.abs().max().item() > 0.1
    # rigid-body rows: base row equals the root state
    rb = sim._rigid_body_state.view(N, 19, 13)
    assert torch.allclose(rb[:, 0, :7], sim.all_root_states[:, :7], atol=1e-6)


def test_eval_script_gym_calls(gpu):
    """capture_reference_masses / apply_base_mass of scripts/eval.py:185-214 through the shim."""
    sim, cfg, N = _make_sim()
    props = sim.gym.get_actor_rigid_body_properties(sim.envs[0], sim.robot_handles[0])
    masses = np.array([p.mass for p in props], dtype=np.float32)
    assert len(masses) == 19 and abs(masses[sim.body_names.index("FL_foot")] - 0.04) < 1e-7
    masses[0] = 7.5
    for env, actor in zip(sim.envs, sim.robot_handles):
        ps = sim.gym.get_actor_rigid_body_properties(env, actor)
        for i, p in enumerate(ps):
            p.mass = float(masses[i])
        sim.gym.set_actor_rigid_body_properties(env, actor, ps, recomputeInertia=True)
    sim.gym.refresh_mass_matrix_tensors(sim.sim)
    assert np.allclose(sim.params.cpu().numpy()[0], 7.5)


def test_set_body_params_and_gym_shim_agree(gpu):
    """The native set_body_params (one device update for many envs) writes the base link's mass,
    centre of mass and full inertia tensor (products included) into the parameter rows, the gym
    shim reads them back and its write-back is a no-op round trip; the shim also carries products
    of inertia; other bodies and non-symmetric tensors are refused."""
    import torch
    from spi_amd.params import P_COMX, P_COMZ, P_IXX, P_IXY, P_IXZ, P_IYY, P_IYZ, P_IZZ, P_MASS
    sim, cfg, N = _make_sim()
    assert N >= 4
    before = sim.params.clone()
    ids = [1, 3]
    mass = np.array([7.1, 8.2])
    com = np.array([[0.01, -0.02, 0.005], [-0.003, 0.004, 0.02]])
    I0 = np.array([[0.03, 0.001, -0.002], [0.001, 0.1, 0.0005], [-0.002, 0.0005, 0.11]])
    I = np.stack([I0, 1.1 * I0])
    sim.set_body_params(ids, "base", mass=mass, com=com, inertia=I)
    p = sim.params.cpu().numpy()
    for k, e in enumerate(ids):
        assert np.allclose(p[P_MASS, e], mass[k]) and np.allclose(p[P_COMX:P_COMZ + 1, e], com[k])
        got = [p[r, e] for r in (P_IXX, P_IYY, P_IZZ, P_IXY, P_IXZ, P_IYZ)]
        want = [I[k, 0, 0], I[k, 1, 1], I[k, 2, 2], I[k, 0, 1], I[k, 0, 2], I[k, 1, 2]]
        assert np.allclose(got, want, rtol=1e-6)
    untouched = [e for e in range(N) if e not in ids]
    assert torch.equal(sim.params[:, untouched], before[:, untouched])
    props = sim.gym.get_actor_rigid_body_properties(sim.envs[3], sim.robot_handles[3])
    b = props[0]
    assert np.isclose(b.mass, 8.2) and np.isclose(b.inertia.x.y, I[1, 0, 1]) and np.isclose(b.inertia.y.z, I[1, 1, 2])
    snap = sim.params.clone()
    sim.gym.set_actor_rigid_body_properties(sim.envs[3], sim.robot_handles[3], props, recomputeInertia=True)
    assert torch.equal(sim.params, snap)
    props[0].inertia.x.z = props[0].inertia.z.x = 0.004
    sim.gym.set_actor_rigid_body_properties(sim.envs[3], sim.robot_handles[3], props)
    assert np.isclose(sim.params[P_IXZ, 3].item(), 0.004)
    # a scalar for every env at once
    sim.set_body_params(range(N), "base", mass=6.0)
    assert torch.allclose(sim.params[P_MASS], torch.full((N,), 6.0, device="cuda:0"))
    with pytest.raises(NotImplementedError):
        sim.set_body_params([0], "FL_thigh", mass=1.0)
    with pytest.raises(ValueError):
        sim.set_body_params([0], "base", inertia=[[0.1, 0.01, 0.0], [0.0, 0.1, 0.0], [0.0, 0.0, 0.1]])
    with pytest.raises(IndexError):
        sim.set_body_params([N], "base", mass=6.0)
