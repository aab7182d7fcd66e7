"""HipSim: the MI355X backend behind the BaseSimulator surface (replaces IsaacGym/PhysX).

Method-for-method counterpart of spigym/simulator/isaacgym/isaacgym.py:18-627 for the
headless sysID path.  State ownership follows IsaacGym's tensor API:

* the engine advances an internal state (root [N,13], dof [N,12,2]) on the GPU;
* ``all_root_states`` / ``dof_state`` (and their views ``robot_root_states``, ``base_quat``,
  ``dof_pos``, ``dof_vel``) are torch tensors the env reads and writes;
* ``refresh_sim_tensors`` copies internal -> views (isaacgym.py:579-589); a physics step
  refreshes the dof view only (isaacgym.py:626); ``set_*_tensor(env_ids, src)`` copies the
  listed rows of ``src`` into the internal state (isaacgym.py:601-620).

So writes into the views that are not pushed with ``set_*_tensor`` are seen by the env's
next torque computation and then overwritten, exactly as with PhysX.
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np
import torch

from .. import _lib
from ..model import ISAAC_BODY_NAMES, load_model
from ..params import (CONTROL_MODES, MOTOR_MODELS, P_COUNT, P_COMX, P_COMZ, P_IXX, P_IXY, P_IXZ, P_IYY,
                      P_IYZ, P_IZZ, P_MASS, SimParams, nominal_params)
from ..rollout import RolloutEngine
from .base import BaseSimulator


class _Vec:
    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.x, self.y, self.z = x, y, z


class _Inertia:
    def __init__(self, m):
        self.x, self.y, self.z = _Vec(*m[0]), _Vec(*m[1]), _Vec(*m[2])


class RigidBodyProps:
    """Per-body properties with the attribute names of gymapi.RigidBodyProperties."""

    def __init__(self, mass, com, inertia):
        self.mass = float(mass)
        self.com = _Vec(*com)
        self.inertia = _Inertia(inertia)


class _GymShim:
    """The few ``simulator.gym`` calls the plain sysID scripts make (scripts/eval.py:185-214)."""

    def __init__(self, sim: "HipSim"):
        self._s = sim

    def get_actor_rigid_body_properties(self, env, actor):
        return self._s._body_props(int(env))

    def set_actor_rigid_body_properties(self, env, actor, props, recomputeInertia=False):
        self._s._set_body_props(int(env), props)

    def refresh_mass_matrix_tensors(self, sim):
        pass

    def destroy_sim(self, sim):
        pass


class HipSim(BaseSimulator):
    def __init__(self, config, device):
        super().__init__(config, device)
        self.simulator_config = config.simulator.config
        self.robot_config = config.robot
        self.visualize_viewer = False
        self.gym = _GymShim(self)
        self.sim = "hipsim"
        self.viewer = None

    # ---------------------------------------------------------------- configuration
    def setup(self):
        if not str(self.sim_device).startswith("cuda"):
            raise RuntimeError("HipSim runs on the GPU only (device must be cuda:N)")
        _lib.require_gpu()
        sc = self.simulator_config
        self.sim_params = SimParams.from_robot_config(self.robot_config, {"sim": dict(sc.sim)})
        contact = sc.get("contact", {}) or {}
        self.sim_params.k_contact = float(contact.get("stiffness", self.sim_params.k_contact))
        self.sim_params.c_contact = float(contact.get("damping", self.sim_params.c_contact))
        self.sim_params.mu = float(contact.get("friction", self.sim_params.mu))
        self.sim_params.c_tangent = float(contact.get("tangential_damping", self.sim_params.c_tangent))
        self.sim_params.fric_vel = float(sc.get("joint_friction_velocity", self.sim_params.fric_vel))
        self.sim_dt = self.sim_params.dt
        self.device = self.sim_device
        self.engine = RolloutEngine(self.sim_params, device=self.sim_device)

    def setup_terrain(self, mesh_type):
        if mesh_type not in (None, "plane"):
            raise ValueError("HipSim simulates a flat ground plane only (mesh_type 'plane')")
        self.terrain = None

    # ---------------------------------------------------------------- assets
    def load_assets(self):
        self.model = load_model()
        self.num_dof = len(self.model["dof_names"])
        self.dof_names = list(self.model["dof_names"])
        self.body_names = [b["name"] for b in self.model["isaac_bodies"]]
        self.num_bodies = len(self.body_names)
        rc = self.robot_config
        assert self.num_dof == len(rc.dof_names), "Number of DOFs must be equal to number of actions"
        assert self.num_bodies == len(rc.body_names), "Number of bodies must be equal to number of body names"
        assert self.dof_names == list(rc.dof_names), "DOF names must match the config"
        assert self.body_names == list(rc.body_names), "Body names must match the config"
        self._body_list = list(self.body_names)

    def create_envs(self, num_envs, env_origins, base_init_state):
        self.num_envs = int(num_envs)
        self.env_origins = env_origins
        self.base_init_state = base_init_state
        self.envs = list(range(self.num_envs))
        self.robot_handles = [0] * self.num_envs
        nom = nominal_params(self.model)
        self._params_np = np.repeat(nom[None], self.num_envs, axis=0)
        self._process_rigid_body_props_all()
        self.params = torch.from_numpy(np.ascontiguousarray(self._params_np.T.astype(np.float32))).to(self.sim_device)
        lower, upper, vel, eff = (np.array(x, np.float32) for x in self._urdf_limits())
        self.hard_dof_pos_limits = torch.tensor(np.stack([lower, upper], 1), device=self.sim_device)
        self.dof_vel_limits = torch.tensor(vel, device=self.sim_device)
        self.torque_limits = torch.tensor(eff, device=self.sim_device)
        m = (self.hard_dof_pos_limits[:, 0] + self.hard_dof_pos_limits[:, 1]) / 2
        r = self.hard_dof_pos_limits[:, 1] - self.hard_dof_pos_limits[:, 0]
        soft = _get(self.config, ["rewards", "reward_limit", "soft_dof_pos_limit"], 1.0)
        term = _get(self.config, ["termination_scales", "termination_close_to_dof_pos_limit"], 1.0)
        self.dof_pos_limits = torch.stack([m - 0.5 * r * soft, m + 0.5 * r * soft], 1)
        self.dof_pos_limits_termination = torch.stack([m - 0.5 * r * term, m + 0.5 * r * term], 1)
        return self.envs, self.robot_handles

    def _urdf_limits(self):
        lims = [b["limit"] for b in self.model["bodies"][1:]]
        return ([l["lower"] for l in lims], [l["upper"] for l in lims], [l["velocity"] for l in lims],
                [l["effort"] for l in lims])

    def _process_rigid_body_props_all(self):
        """Per-env body-property hook (isaacgym.py:247-272); subclasses set parameters here."""

    def get_dof_limits_properties(self):
        rc = self.robot_config
        for i in range(self.num_dof):
            assert abs(self.hard_dof_pos_limits[i, 0].item() - rc.dof_pos_lower_limit_list[i]) < 1e-5
            assert abs(self.hard_dof_pos_limits[i, 1].item() - rc.dof_pos_upper_limit_list[i]) < 1e-5
            assert abs(self.dof_vel_limits[i].item() - rc.dof_vel_limit_list[i]) < 1e-5
            assert abs(self.torque_limits[i].item() - rc.dof_effort_limit_list[i]) < 1e-5
        return self.dof_pos_limits, self.dof_vel_limits, self.torque_limits

    def find_rigid_body_indice(self, body_name):
        return self.body_names.index(body_name)

    # ---------------------------------------------------------------- tensors
    def prepare_sim(self):
        N, dev = self.num_envs, self.sim_device
        self._root = torch.zeros((N, 13), dtype=torch.float32, device=dev)
        self._dof = torch.zeros((N, self.num_dof, 2), dtype=torch.float32, device=dev)
        init = torch.as_tensor(np.asarray(self.base_init_state, np.float32), device=dev)
        self._root[:] = init[:13]
        self._root[:, :3] += torch.as_tensor(self.env_origins, dtype=torch.float32, device=dev)[:, :3]
        self._torques = torch.zeros((N, self.num_dof), dtype=torch.float32, device=dev)
        self._foot_force = torch.zeros((N, 4, 3), dtype=torch.float32, device=dev)
        self.all_root_states = torch.zeros((N, 13), dtype=torch.float32, device=dev)
        self.robot_root_states = self.all_root_states.view(N, 1, 13)[..., 0, :]
        self.base_quat = self.robot_root_states[..., 3:7]
        self.dof_state = torch.zeros((N * self.num_dof, 2), dtype=torch.float32, device=dev)
        self.dof_pos = self.dof_state.view(N, -1, 2)[..., 0]
        self.dof_vel = self.dof_state.view(N, -1, 2)[..., 1]
        self.contact_forces = torch.zeros((N, self.num_bodies, 3), dtype=torch.float32, device=dev)
        self._rigid_body_state = torch.zeros((N * self.num_bodies, 13), dtype=torch.float32, device=dev)
        rb = self._rigid_body_state.view(N, self.num_bodies, 13)
        self._rigid_body_pos = rb[..., 0:3]
        self._rigid_body_rot = rb[..., 3:7]
        self._rigid_body_vel = rb[..., 7:10]
        self._rigid_body_ang_vel = rb[..., 10:13]
        self._foot_bodies = [self.body_names.index(f"{l}_foot") for l in ("FL", "FR", "RL", "RR")]
        self.refresh_sim_tensors()

    def refresh_sim_tensors(self):
        self.all_root_states.copy_(self._root)
        self.dof_state.view(self.num_envs, -1, 2).copy_(self._dof)
        L = _lib.lib()
        _lib.check(L.spi_rigid_body_states(_lib.ptr(self._root), _lib.ptr(self._dof), self.num_envs,
                                           _lib.ptr(self._rigid_body_state), _lib.stream()), "spi_rigid_body_states")
        self.contact_forces[:, self._foot_bodies, :] = self._foot_force

    def apply_torques_at_dof(self, torques):
        self._torques.copy_(torques.view(self.num_envs, self.num_dof))

    def _set_rows(self, dst: torch.Tensor, src: torch.Tensor, set_env_ids, row_len: int):
        ids = set_env_ids.to(device=self.sim_device, dtype=torch.int32).contiguous()
        src = src.contiguous()
        L = _lib.lib()
        _lib.check(L.spi_set_rows_indexed(_lib.ptr(dst), _lib.ptr(src), _lib.ptr(ids, torch.int32), ids.numel(),
                                          self.num_envs, row_len, _lib.stream()), "spi_set_rows_indexed")

    def set_actor_root_state_tensor(self, set_env_ids, root_states):
        self._set_rows(self._root, root_states.view(self.num_envs, 13), set_env_ids, 13)

    def set_dof_state_tensor(self, set_env_ids, dof_states):
        self._set_rows(self._dof, dof_states.view(self.num_envs, self.num_dof, 2), set_env_ids, 2 * self.num_dof)

    def simulate_at_each_physics_step(self):
        self.engine.physics_step(self.params, self._root, self._dof, self._torques, self._foot_force)
        self.dof_state.view(self.num_envs, -1, 2).copy_(self._dof)   # refresh_dof_state_tensor

    # ---------------------------------------------------------------- properties (gym shim)
    def _body_props(self, env: int) -> List[RigidBodyProps]:
        p = self._params_np[env]
        props = []
        for name in self.body_names:
            if name == "base":
                I = np.array([[p[P_IXX], p[P_IXY], p[P_IXZ]], [p[P_IXY], p[P_IYY], p[P_IYZ]],
                              [p[P_IXZ], p[P_IYZ], p[P_IZZ]]])
                props.append(RigidBodyProps(p[P_MASS], p[P_COMX:P_COMZ + 1], I))
            else:
                link = self.model["raw"]["links"][name]
                from ..model import inertia6_to_matrix
                props.append(RigidBodyProps(link["mass"], link["com"], inertia6_to_matrix(link["inertia"])))
        return props

    def _set_body_props(self, env: int, props) -> None:
        ref = self._body_props(env)
        for i, (a, b) in enumerate(zip(props, ref)):
            # scripts/eval.py round-trips masses through float32 (capture_reference_masses)
            if i > 0 and abs(a.mass - b.mass) > 1e-6 * max(abs(b.mass), 1e-3):
                raise NotImplementedError(f"HipSim identifies the base only; body {self.body_names[i]} changed")
        base = props[0]
        I = base.inertia
        self.set_body_params([env], "base", mass=base.mass, com=(base.com.x, base.com.y, base.com.z),
                             inertia=((I.x.x, I.x.y, I.x.z), (I.y.x, I.y.y, I.y.z), (I.z.x, I.z.y, I.z.z)))

    def set_body_params(self, env_ids, name: str, mass=None, com=None, inertia=None) -> None:
        """Mass, centre of mass and inertia tensor (about the centre of mass, body axes) of body
        ``name`` for the envs ``env_ids``, in one update of the device parameter array: the
        native counterpart of the per-env ``gym.get/set_actor_rigid_body_properties`` loop of
        the plain scripts (scripts/eval.py:185-189, 204-214).  ``mass`` is a scalar or one value
        per env, ``com`` [3] or [n, 3], ``inertia`` [3, 3] or [n, 3, 3] (symmetric); None leaves a
        quantity as it is.  The base link is the only body the engine takes as a parameter
        (the identified one); other bodies keep their URDF values."""
        if name != "base":
            raise NotImplementedError(f"HipSim identifies the base only; body {name!r} is fixed")
        ids = np.asarray(env_ids, dtype=np.int64).reshape(-1)
        if ids.size == 0:
            return
        if ids.min() < 0 or ids.max() >= self.num_envs:
            raise IndexError("env id out of range")
        n = ids.size
        rows = self._params_np[ids]                     # [n, 25] copy
        if mass is not None:
            rows[:, P_MASS] = np.broadcast_to(np.asarray(mass, np.float64).reshape(-1), (n,))
        if com is not None:
            rows[:, P_COMX:P_COMZ + 1] = np.broadcast_to(np.asarray(com, np.float64).reshape(-1, 3), (n, 3))
        if inertia is not None:
            I = np.broadcast_to(np.asarray(inertia, np.float64).reshape(-1, 3, 3), (n, 3, 3))
            if not np.allclose(I, np.swapaxes(I, 1, 2), rtol=1e-6, atol=1e-12):
                raise ValueError("inertia tensor must be symmetric")
            rows[:, P_IXX], rows[:, P_IYY], rows[:, P_IZZ] = I[:, 0, 0], I[:, 1, 1], I[:, 2, 2]
            rows[:, P_IXY], rows[:, P_IXZ], rows[:, P_IYZ] = I[:, 0, 1], I[:, 0, 2], I[:, 1, 2]
        self._params_np[ids] = rows
        dev_ids = torch.as_tensor(ids, device=self.sim_device)
        self.params[:, dev_ids] = torch.as_tensor(np.ascontiguousarray(rows.T.astype(np.float32)),
                                                  device=self.sim_device)

    # ---------------------------------------------------------------- viewer (headless only)
    def setup_viewer(self):
        raise NotImplementedError("HipSim is headless")

    def render(self, sync_frame_time=True):
        pass


def _get(cfg, keys, default):
    cur = cfg
    for k in keys:
        try:
            cur = cur[k]
        except (KeyError, TypeError):
            return default
    return cur
