"""ctypes binding of libspi_hip.so (the C ABI declared in include/spi_hip.h).

torch provides device memory and the stream only: every pointer passed here is a
``tensor.data_ptr()`` and every call is enqueued on ``torch.cuda.current_stream()``.
There is no fallback: without the built library or a GPU the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np
import torch  # noqa: F401  -- must be loaded first so the HIP runtime (soname libamdhip64.so.7) is shared

from . import build
from .params import SimParams

SPI_P_COUNT = 25
SPI_STATE_DIM = 37
SPI_MAX_CEM_SAMPLES = 8192
SPI_MAX_TANGENTS = 32
SPI_MLP_SCRATCH_FLOATS = 1 << 22


class SpiSimArgs(C.Structure):
    _fields_ = [("dt", C.c_float), ("decimation", C.c_int), ("substeps", C.c_int), ("gravity", C.c_float),
                ("kp", C.c_float * 12), ("kd", C.c_float * 12), ("q0", C.c_float * 12), ("effort", C.c_float * 12),
                ("action_scale", C.c_float), ("action_clip", C.c_float), ("hip_scale", C.c_float),
                ("motor_model", C.c_int), ("k_contact", C.c_float), ("c_contact", C.c_float), ("mu", C.c_float),
                ("c_tangent", C.c_float), ("fric_vel_inv", C.c_float), ("state_limit", C.c_float)]


class SpiCostArgs(C.Structure):
    _fields_ = [("mode", C.c_int), ("w", C.c_float * 6), ("seg_gains", C.c_void_p)]


class SpiResetSpec(C.Structure):
    _fields_ = [("q_default", C.c_float * 12), ("root0", C.c_float * 7), ("q_lo", C.c_float), ("q_hi", C.c_float),
                ("vel_range", C.c_float)]


class SpiTangentSpec(C.Structure):
    _fields_ = [("n_tan", C.c_int), ("param", C.c_int * SPI_P_COUNT)]


def tangent_spec(param_rows) -> SpiTangentSpec:
    rows = [int(r) for r in param_rows]
    if not (1 <= len(rows) <= min(SPI_MAX_TANGENTS, SPI_P_COUNT)):
        raise ValueError(f"1..{SPI_P_COUNT} tangent directions, got {len(rows)}")
    t = SpiTangentSpec()
    t.n_tan = len(rows)
    for k, r in enumerate(rows):
        if not 0 <= r < SPI_P_COUNT:
            raise ValueError(f"parameter row {r} out of range")
        t.param[k] = r
    return t


def sim_args(sim: SimParams, state_limit: float = 1.0e4) -> SpiSimArgs:
    a = SpiSimArgs()
    a.dt = sim.dt
    a.decimation = sim.decimation
    a.substeps = sim.substeps
    a.gravity = sim.gravity
    a.kp[:] = [float(x) for x in sim.kp]
    a.kd[:] = [float(x) for x in sim.kd]
    a.q0[:] = [float(x) for x in sim.q0]
    a.effort[:] = [float(x) for x in sim.effort]
    a.action_scale = sim.action_scale
    a.action_clip = sim.action_clip
    a.hip_scale = 0.5 if sim.control_mode == 1 else 1.0
    a.motor_model = sim.motor_model
    a.k_contact = sim.k_contact
    a.c_contact = sim.c_contact
    a.mu = sim.mu
    a.c_tangent = sim.c_tangent
    a.fric_vel_inv = 1.0 / sim.fric_vel
    a.state_limit = state_limit
    return a


def cost_args(mode: int, weights, seg_gains: Optional[torch.Tensor] = None) -> SpiCostArgs:
    """Cost settings; ``seg_gains`` is an optional device [n_seg, 2] float32 tensor of recorded
    (kp, kd) per segment (the caller keeps it alive while launches that use it are queued)."""
    c = SpiCostArgs()
    c.mode = int(mode)
    c.w[:] = [float(x) for x in weights]
    c.seg_gains = None
    if seg_gains is not None:
        if seg_gains.dtype != torch.float32 or seg_gains.dim() != 2 or seg_gains.shape[1] != 2 \
                or not seg_gains.is_contiguous():
            raise ValueError("seg_gains must be a contiguous float32 [n_seg, 2] tensor")
        c.seg_gains = seg_gains.data_ptr()
    return c


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load the library (building it in-tree if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    path = build.ensure_built()
    L = C.CDLL(path)
    vp, ip, fp = C.c_void_p, C.c_int, C.c_float
    sa, ca = C.POINTER(SpiSimArgs), C.POINTER(SpiCostArgs)
    L.spi_last_error.restype = C.c_char_p
    L.spi_abi_version.restype = ip
    L.spi_segment_stride.argtypes = [ip]
    L.spi_segment_stride.restype = ip
    sigs = {
        "spi_physics_step": [sa, vp, ip, ip, vp, vp, vp, vp, vp],
        "spi_forward_dynamics": [sa, vp, ip, ip, vp, vp, vp, vp, vp, vp],
        "spi_set_rows_indexed": [vp, vp, vp, ip, ip, ip, vp],
        "spi_rigid_body_states": [vp, vp, ip, vp, vp],
        "spi_compute_torques": [sa, vp, ip, ip, vp, vp, vp, vp],
        "spi_fd_fim_trace": [vp, vp, vp, ip, ip, fp, vp, vp],
        "spi_rollout_cost": [sa, ca, vp, ip, ip, vp, ip, ip, vp, vp],
        "spi_rollout_cost_legs": [sa, ca, vp, ip, ip, vp, ip, ip, vp, vp],
        "spi_reduce_segments": [vp, vp, ip, ip, vp, vp],
        "spi_simulate": [sa, vp, ip, ip, vp, vp, ip, vp, vp],
        "spi_fim_rollout": [sa, vp, ip, ip, ip, vp, vp, ip, fp, ip, ip, vp, vp],
        "spi_cem_sample": [vp, vp, vp, vp, ip, ip, C.c_ulonglong, C.c_uint, vp, ip, vp],
        "spi_cem_update": [vp, vp, ip, ip, ip, ip, fp, vp, vp, vp, vp, vp, vp, vp],
        "spi_batch_cem_sample": [vp, ip, ip, vp],
        "spi_batch_rollout_cost": [sa, vp, vp, ip, C.c_longlong, ip, ip, ip, ip, vp, vp],
        "spi_batch_cem_update": [vp, ip, vp],
        "spi_traj_sensitivity": [sa, C.POINTER(SpiTangentSpec), vp, ip, ip, vp, vp, ip, ip, vp, vp, vp, vp],
        "spi_traj_sensitivity_legs": [sa, C.POINTER(SpiTangentSpec), vp, ip, ip, vp, vp, ip, ip, vp, vp, vp, vp],
        "spi_traj_sensitivity_pairs": [sa, C.POINTER(SpiTangentSpec), vp, ip, ip, vp, vp, ip, ip, vp, vp, vp, vp],
        "spi_fim_assemble": [vp, C.c_longlong, ip, ip, vp, vp],
        "spi_fim_logdet": [vp, vp, ip, ip, ip, vp, C.c_double, vp, vp, vp, vp],
        "spi_rollout_cost_grad": [sa, ca, C.POINTER(SpiTangentSpec), vp, ip, ip, vp, ip, ip, vp, vp, vp],
        "spi_fourier_excitations": [vp, ip, ip, ip, fp, fp, fp, vp, vp],
        "spi_residual_jacobian": [sa, ca, C.POINTER(SpiTangentSpec), vp, ip, ip, vp, vp, ip, ip, vp, vp, vp],
        "spi_residual_jacobian_legs": [sa, ca, C.POINTER(SpiTangentSpec), vp, ip, ip, vp, vp, ip, ip, vp, vp, vp],
        "spi_gn_step": [vp, vp, ip, ip, ip, vp, vp, vp, vp, vp],
        "spi_command_sequences": [vp, vp, ip, ip, vp, vp],
        "spi_pack_windows": [vp, vp, ip, ip, ip, ip, vp, vp],
        "spi_pack_windows_rows": [vp, vp, ip, vp, ip, ip, vp, vp],
        "spi_rollout_cost_adjoint": [sa, ca, vp, ip, ip, vp, ip, ip, ip, vp, C.c_longlong, vp, vp, vp],
        "spi_rollout_cost_adjoint_legs": [sa, ca, vp, ip, ip, vp, ip, ip, ip, vp, C.c_longlong, vp, vp, vp],
        "spi_policy_step": [sa, vp, ip, ip, ip, vp, vp, vp, vp, ip, vp, vp, vp],
        "spi_policy_step_legs": [sa, vp, ip, ip, ip, vp, vp, vp, vp, ip, vp, vp, vp],
        "spi_policy_reward": [vp, ip, ip, fp, fp, fp, fp, vp, vp, vp, vp, vp],
        "spi_policy_obs": [ip, ip, vp, vp, ip, ip, ip, fp, vp, vp, vp, vp, vp, ip, fp, fp, fp, vp],
        "spi_policy_sync_state": [vp, ip, ip, vp],
        "spi_mlp_forward": [vp, ip, ip, C.POINTER(ip), vp, vp, C.c_longlong, vp, vp],
        "spi_mlp_forward_ring": [vp, ip, ip, ip, ip, vp, fp, ip, ip, C.POINTER(ip), vp, vp, C.c_longlong, vp, vp],
        "spi_policy_terms": [ip, ip, vp, vp, ip, ip, ip, fp, vp, vp, vp, vp, ip, ip, ip, fp, fp, vp],
        "spi_policy_reward_terms": [ip, ip, vp, vp, ip, ip, ip, fp, vp, vp, vp, vp, ip, ip, ip, fp, fp, fp, fp, fp, fp,
                                    vp, vp, vp, vp, vp],
        "spi_policy_obs_kmap": [C.POINTER(ip)],
        "spi_reset_states": [C.POINTER(SpiResetSpec), ip, vp, C.c_ulonglong, vp, vp],
        "spi_mlp_forward_noalloc": [vp, ip, ip, C.POINTER(ip), vp, vp, vp],
    }
    L.spi_adjoint_work_floats.argtypes = [ip, ip, ip, ip]
    L.spi_adjoint_legs_work_floats.argtypes = [ip, ip, ip, ip, ip]
    for name in ("spi_adjoint_work_floats", "spi_adjoint_legs_work_floats"):
        getattr(L, name).restype = C.c_longlong
    L.spi_policy_obs_dim.restype = ip
    L.spi_policy_ring_slots.restype = ip
    for name in ("spi_mlp_param_floats", "spi_mlp_work_floats"):
        getattr(L, name).restype = C.c_longlong
    L.spi_mlp_param_floats.argtypes = [ip, C.POINTER(ip)]
    L.spi_mlp_work_floats.argtypes = [ip, ip, C.POINTER(ip)]
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ip
    if L.spi_abi_version() != 2:
        raise RuntimeError("libspi_hip.so ABI version mismatch")
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().spi_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {rc}): {msg}")


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("spi_amd needs a ROCm GPU: the engine has no CPU path")


def ptr(t: Optional[torch.Tensor], dtype=torch.float32) -> Optional[int]:
    """Device pointer of a contiguous GPU tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("tensor must live on the GPU")
    if t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream
