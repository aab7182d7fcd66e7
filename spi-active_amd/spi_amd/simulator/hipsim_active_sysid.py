"""HipSimActiveSysId: per-env physical parameters from ``params_dict``.

Counterpart of spigym/simulator/isaacgym/isaacgym_active_sysid.py:29-94.  The env builds
``params_dict`` (spigym/envs/sysid/active_sysid_openloop.py:51-71) as
``{key: {"body_name": "base", "value": [per-env values]}}`` and the simulator calls
``set_<key>(props, env_id, value)`` for every key that has a setter.  Here the same setters
exist with the same names and per-env signature; ``create_envs`` applies each key to a whole
parameter row at once instead of looping over envs.

The reference defines ``set_inertiaiy`` (isaacgym_active_sysid.py:86) while the config key is
``inertiay`` (env/active_sysid_openloop.yaml:23), so Iyy is never applied there.  HipSim
reproduces that by default; ``simulator.config.apply_inertiay: true`` applies it.
"""
from __future__ import annotations

import logging

import numpy as np

from ..params import (P_COMX, P_COMY, P_COMZ, P_IXX, P_IYY, P_IZZ, P_MASS, P_MOTOR0)
from .hipsim import HipSim

log = logging.getLogger(__name__)

# setter name -> parameter row
_ROWS = {"set_mass": P_MASS, "set_comx": P_COMX, "set_comy": P_COMY, "set_comz": P_COMZ,
         "set_inertiax": P_IXX, "set_inertiaiy": P_IYY, "set_inertiaz": P_IZZ,
         "set_motor_model_hip_a": P_MOTOR0, "set_motor_model_thigh_a": P_MOTOR0 + 1,
         "set_motor_model_calf_a": P_MOTOR0 + 2}


class HipSimActiveSysId(HipSim):
    def __init__(self, config, device):
        super().__init__(config, device)
        self.params_dict = config.get("params_dict", None) if hasattr(config, "get") else None
        if self.params_dict is None:
            self.params_dict = getattr(config, "params_dict", None)
        if bool(self.simulator_config.get("apply_inertiay", False)):
            self.set_inertiay = self.set_inertiaiy

    def _process_rigid_body_props_all(self):
        if not self.params_dict:
            return
        for key, value in self.params_dict.items():
            name = f"set_{key}"
            if not hasattr(self, name):
                log.debug("No setter method found for key: %s", key)
                if key == "inertiay":
                    log.warning("params_dict key 'inertiay' is not applied (as in the reference); set "
                                "simulator.config.apply_inertiay to apply it")
                continue
            if "body_name" in value and value["body_name"] != "base":
                raise NotImplementedError(f"HipSim identifies the base only (got body {value['body_name']})")
            vals = np.asarray(list(value["value"]), dtype=np.float64)
            if vals.shape[0] != self.num_envs:
                raise ValueError(f"params_dict['{key}'] has {vals.shape[0]} values for {self.num_envs} envs")
            row = _ROWS[name if name in _ROWS else "set_inertiaiy"]
            self._params_np[:, row] = vals

    # per-env setters with the reference's names and signature (props is this env's
    # parameter row; body lookup by name like isaacgym_active_sysid.py:61-94)
    def _set(self, row, props, env_id, value):
        if value.get("body_name", "base") != "base":
            raise NotImplementedError("HipSim identifies the base only")
        props[row] = value["value"][env_id]

    def set_mass(self, props, env_id, value):
        self._set(P_MASS, props, env_id, value)

    def set_comx(self, props, env_id, value):
        self._set(P_COMX, props, env_id, value)

    def set_comy(self, props, env_id, value):
        self._set(P_COMY, props, env_id, value)

    def set_comz(self, props, env_id, value):
        self._set(P_COMZ, props, env_id, value)

    def set_inertiax(self, props, env_id, value):
        self._set(P_IXX, props, env_id, value)

    def set_inertiaiy(self, props, env_id, value):
        self._set(P_IYY, props, env_id, value)

    def set_inertiaz(self, props, env_id, value):
        self._set(P_IZZ, props, env_id, value)

    def set_motor_model_hip_a(self, props, env_id, value):
        props[P_MOTOR0] = value["value"][env_id]

    def set_motor_model_thigh_a(self, props, env_id, value):
        props[P_MOTOR0 + 1] = value["value"][env_id]

    def set_motor_model_calf_a(self, props, env_id, value):
        props[P_MOTOR0 + 2] = value["value"][env_id]
