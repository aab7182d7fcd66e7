"""Simulation settings and the per-candidate physical parameter vector.

Parameter vector layout (one row per candidate / env; the kernels read it as SoA
``[P_COUNT][n]``):

====  ===========  =========================================================
idx   name         meaning
====  ===========  =========================================================
0     mass         base link mass [kg]          (isaacgym_active_sysid.py:61-64)
1-3   comx..comz   base link CoM, base frame [m] (isaacgym_active_sysid.py:66-79)
4-6   ixx iyy izz  base link inertia about CoM  (isaacgym_active_sysid.py:81-94)
7-9   ixy ixz iyz  base link products of inertia
10-13 kp_scale     PD stiffness scale per leg FL FR RL RR (legged_robot_base.py:118, 545)
14-17 kd_scale     PD damping scale per leg                (legged_robot_base.py:119, 545)
18-21 friction     Coulomb joint friction per leg [N m]
22-24 motor a      motor-model terms hip/thigh/calf; meaning per model:
                   act2tau_vec3_tanh A (active_sysid_openloop.py:381-400),
                   act2tau_vec3 gains (:365-379), act2tau_scalar gain = a[0] (:356-363)
====  ===========  =========================================================

The first 22 entries are the identified parameters of the 22-parameter CEM
problem; the motor-model terms are fixed there (zero sampling width) but can be
varied like any other entry (the active-sysID exploration params do).

``recomputeInertia=True`` (isaacgym.py:270, scripts/eval.py:211-213) lives in the
closed IsaacGym binary; here inertia is an independent parameter and setting the
mass never rescales it.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .model import DOF_NAMES, load_model

P_COUNT = 25
P_MASS, P_COMX, P_COMY, P_COMZ = 0, 1, 2, 3
P_IXX, P_IYY, P_IZZ, P_IXY, P_IXZ, P_IYZ = 4, 5, 6, 7, 8, 9
P_KP0, P_KD0, P_FRIC0, P_MOTOR0 = 10, 14, 18, 22
N_IDENT = 22

PARAM_NAMES = (["mass", "comx", "comy", "comz", "inertiax", "inertiay", "inertiaz",
                "inertiaxy", "inertiaxz", "inertiayz"]
               + [f"kp_scale_{l}" for l in ("FL", "FR", "RL", "RR")]
               + [f"kd_scale_{l}" for l in ("FL", "FR", "RL", "RR")]
               + [f"friction_{l}" for l in ("FL", "FR", "RL", "RR")]
               + ["motor_model_hip_a", "motor_model_thigh_a", "motor_model_calf_a"])
assert len(PARAM_NAMES) == P_COUNT

CONTROL_MODES = {"legged_base": 0, "go2_omni": 1}
MOTOR_MODELS = {"none": 0, "act2tau_scalar": 1, "act2tau_vec3": 2, "act2tau_vec3_tanh": 3}
COST_MODES = {"tracking": 0, "reference": 1}

# go2.yaml:56-68 default joint angles in DOF order
GO2_DEFAULT_JOINT_ANGLES = {
    "FL_hip_joint": 0.1, "FL_thigh_joint": 0.8, "FL_calf_joint": -1.5,
    "FR_hip_joint": -0.1, "FR_thigh_joint": 0.8, "FR_calf_joint": -1.5,
    "RL_hip_joint": 0.1, "RL_thigh_joint": 1.0, "RL_calf_joint": -1.5,
    "RR_hip_joint": -0.1, "RR_thigh_joint": 1.0, "RR_calf_joint": -1.5,
}


@dataclass
class SimParams:
    """Everything a rollout needs besides the state and the parameter vector."""
    dt: float = 0.005                 # 1 / sim.fps (isaacgym_active_sysid.yaml:15)
    decimation: int = 4               # control_decimation (isaacgym_active_sysid.yaml:16)
    substeps: int = 4                 # integration sub-intervals per physics step, torque held
                                      # (the role of sim.substeps, isaacgym_active_sysid.yaml:17):
                                      # explicit penalty contact at k=5e4, c=100 needs <= ~2 ms
    gravity: float = 9.81             # isaacgym.py:71
    kp: np.ndarray = field(default_factory=lambda: np.full(12, 20.0))
    kd: np.ndarray = field(default_factory=lambda: np.full(12, 0.5))
    q0: np.ndarray = field(default_factory=lambda: np.array([GO2_DEFAULT_JOINT_ANGLES[n] for n in DOF_NAMES]))
    effort: np.ndarray = field(default_factory=lambda: np.array([23.7, 23.7, 35.55] * 4))
    action_scale: float = 0.25        # go2.yaml:107
    action_clip: float = 20.0         # go2.yaml:108
    control_mode: int = 0             # legged_base (scripts/eval.py path) | go2_omni
    motor_model: int = 0              # none | scalar | vec3 | vec3_tanh
    k_contact: float = 5e4            # penalty contact stiffness [N/m]
    c_contact: float = 100.0          # normal damping [N s/m]
    mu: float = 0.8                   # Coulomb cone
    c_tangent: float = 100.0          # tangential viscous slope below the cone [N s/m]
    fric_vel: float = 0.2             # velocity scale of tanh joint friction [rad/s]

    @staticmethod
    def from_robot_config(robot: Dict, sim: Optional[Dict] = None, **over) -> "SimParams":
        """Build from the robot/simulator config dicts (go2.yaml, isaacgym_active_sysid.yaml)."""
        p = SimParams()
        ctrl = robot.get("control", {})
        stiff, damp = ctrl.get("stiffness", {}), ctrl.get("damping", {})
        if stiff:
            p.kp = np.array([_match_gain(stiff, n) for n in DOF_NAMES], dtype=np.float64)
            p.kd = np.array([_match_gain(damp, n) for n in DOF_NAMES], dtype=np.float64)
        init = robot.get("init_state", {}).get("default_joint_angles")
        if init:
            p.q0 = np.array([init[n] for n in DOF_NAMES], dtype=np.float64)
        if "dof_effort_limit_list" in robot:
            p.effort = np.array(robot["dof_effort_limit_list"], dtype=np.float64)
        p.action_scale = float(ctrl.get("action_scale", p.action_scale))
        p.action_clip = float(ctrl.get("action_clip_value", p.action_clip))
        if sim:
            s = sim.get("config", sim).get("sim", {})
            if "fps" in s:
                p.dt = 1.0 / float(s["fps"])
            p.decimation = int(s.get("control_decimation", p.decimation))
            if "substeps" in s and int(s["substeps"]) > 1:
                p.substeps = int(s["substeps"])
        for k, v in over.items():
            setattr(p, k, v)
        return p

    def kernel_floats(self) -> np.ndarray:
        return np.concatenate([self.kp, self.kd, self.q0, self.effort]).astype(np.float32)


def _match_gain(table: Dict, dof_name: str) -> float:
    """Same lookup rule as legged_robot_base.py:67-83: a key that is a substring of the dof name."""
    for key, val in table.items():
        if key in dof_name:
            return float(val)
    raise ValueError(f"PD gain of joint {dof_name} were not defined")


def nominal_params(model: Optional[Dict] = None, friction: float = 0.1, motor_a=(20.0, 20.0, 20.0)) -> np.ndarray:
    """The URDF's base link inertia, unit PD scales, nominal friction and motor terms."""
    m = model or load_model()
    bl = m["base_link"]
    p = np.zeros(P_COUNT, dtype=np.float64)
    p[P_MASS] = bl["mass"]
    p[P_COMX:P_COMZ + 1] = bl["com"]
    ixx, iyy, izz, ixy, ixz, iyz = bl["inertia"]
    p[P_IXX], p[P_IYY], p[P_IZZ], p[P_IXY], p[P_IXZ], p[P_IYZ] = ixx, iyy, izz, ixy, ixz, iyz
    p[P_KP0:P_KP0 + 4] = 1.0
    p[P_KD0:P_KD0 + 4] = 1.0
    p[P_FRIC0:P_FRIC0 + 4] = friction
    p[P_MOTOR0:P_MOTOR0 + 3] = motor_a
    return p


def params_soa(rows: np.ndarray) -> np.ndarray:
    """[n, P_COUNT] rows -> contiguous SoA [P_COUNT, n] float32."""
    rows = np.asarray(rows, dtype=np.float32).reshape(-1, P_COUNT)
    return np.ascontiguousarray(rows.T)
