"""ORACLE ctypes binding — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker (or the timed CPU baseline).  It loads
oracle/lib/libspi_oracle.so (built by oracle/Makefile) and feeds it the UNFOLDED
URDF tree from the raw part of the committed model file, so the oracle's
dynamics do not depend on the product's fixed-joint folding.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libspi_oracle.so")

P_COUNT = 25


class LinkDesc(C.Structure):
    _fields_ = [("parent", C.c_int), ("jtype", C.c_int), ("dof", C.c_int),
                ("xyz", C.c_double * 3), ("rpy", C.c_double * 3), ("axis", C.c_double * 3),
                ("mass", C.c_double), ("com", C.c_double * 3), ("inertia", C.c_double * 6)]


class ContactDesc(C.Structure):
    _fields_ = [("link", C.c_int), ("pos", C.c_double * 3), ("radius", C.c_double)]


class SimConfig(C.Structure):
    _fields_ = [("dt", C.c_double), ("decimation", C.c_int), ("substeps", C.c_int), ("gravity", C.c_double),
                ("kp", C.c_double * 12), ("kd", C.c_double * 12), ("q0", C.c_double * 12),
                ("effort", C.c_double * 12), ("action_scale", C.c_double), ("action_clip", C.c_double),
                ("control_mode", C.c_int), ("motor_model", C.c_int),
                ("k_contact", C.c_double), ("c_contact", C.c_double), ("mu", C.c_double),
                ("c_tangent", C.c_double), ("fric_vel", C.c_double)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


_libs = {}


def mutant_path(k: int) -> str:
    """Oracle build with the planted integrator mistake k (SPIO_MUTATION in spi_oracle.hpp)."""
    return os.path.join(HERE, "lib", f"libspi_oracle_mut{k}.so")


def lib(path: Optional[str] = None):
    """The oracle library (default build, or the one at `path`, e.g. a mutant)."""
    key = path or LIB_PATH
    if key not in _libs:
        if path is None:
            build()
        L = C.CDLL(key)
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        vp = C.c_void_p
        L.spio_new.restype = vp
        L.spio_new.argtypes = [C.POINTER(LinkDesc), C.c_int, C.POINTER(ContactDesc), C.c_int, C.c_int,
                               C.POINTER(SimConfig)]
        L.spio_new_f32.restype = vp
        L.spio_new_f32.argtypes = L.spio_new.argtypes
        L.spio_free.argtypes = [vp]
        L.spio_free_f32.argtypes = [vp]
        L.spio_forward.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.spio_rnea.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.spio_crba.argtypes = [vp, dp, dp, dp, dp]
        L.spio_momentum.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.spio_fk.argtypes = [vp, dp, dp, dp, dp]
        L.spio_physics_step.argtypes = [vp, dp, dp, dp, dp, dp]
        L.spio_control_torque.argtypes = [C.POINTER(SimConfig), dp, dp, dp, dp, dp]
        L.spio_simulate.argtypes = [vp, dp, fp, fp, C.c_int, fp]
        L.spio_rollout_cost.argtypes = [vp, dp, C.c_int, fp, fp, fp, C.c_int, C.c_int, C.c_int, dp, dp]
        ipp = C.POINTER(C.c_int)
        L.spio_traj_sens.argtypes = [vp, dp, ipp, C.c_int, fp, fp, C.c_int, C.c_int, dp, dp]
        L.spio_cost_grad.argtypes = [vp, dp, ipp, C.c_int, fp, fp, fp, C.c_int, C.c_int, C.c_int, dp, dp, dp]
        if hasattr(L, "spio_traj_sens_batch"):
            L.spio_traj_sens_batch.argtypes = [vp, dp, ipp, C.c_int, fp, fp, C.c_int, C.c_int, C.c_int, dp, dp,
                                               C.c_int]
            L.spio_rollout_cost_par.argtypes = [vp, dp, C.c_int, fp, fp, fp, C.c_int, C.c_int, C.c_int, dp, dp,
                                                C.c_int]
            L.spio_simulate_batch.argtypes = [vp, dp, fp, fp, C.c_int, C.c_int, dp, C.c_int]
        L.spio_rollout_cost_f32.argtypes = [vp, fp, C.c_int, fp, fp, fp, C.c_int, C.c_int, C.c_int, dp, fp,
                                            C.c_int]
        L.spio_max_threads.restype = C.c_int
        u32p = C.POINTER(C.c_uint32)
        L.spio_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.spio_philox_normal.argtypes = [C.c_ulonglong, C.c_uint, C.c_ulonglong]
        L.spio_philox_normal.restype = C.c_double
        L.spio_cem_sample.argtypes = [dp, dp, dp, dp, C.c_int, C.c_int, C.c_ulonglong, C.c_uint, dp]
        assert L.spio_sizeof_config() == C.sizeof(SimConfig)
        assert L.spio_sizeof_link() == C.sizeof(LinkDesc)
        assert L.spio_sizeof_contact() == C.sizeof(ContactDesc)
        _libs[key] = L
    return _libs[key]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def unfolded_tree(model: Dict, dof_names):
    """Links of the raw URDF in DFS order (parents before children), as LinkDesc."""
    raw = model["raw"]
    links, joints = raw["links"], raw["joints"]
    by_parent: Dict[str, list] = {}
    child_set = set()
    for j in joints:
        by_parent.setdefault(j["parent"], []).append(j)
        child_set.add(j["child"])
    root = [n for n in links if n not in child_set][0]
    order, joint_of = [], {}

    def dfs(name):
        order.append(name)
        for j in by_parent.get(name, []):
            joint_of[j["child"]] = j
            dfs(j["child"])

    dfs(root)
    index = {n: i for i, n in enumerate(order)}
    arr = (LinkDesc * len(order))()
    for i, name in enumerate(order):
        L = links[name]
        d = arr[i]
        if name == root:
            d.parent, d.jtype, d.dof = -1, 0, -1
        else:
            j = joint_of[name]
            d.parent = index[j["parent"]]
            if j["type"] == "fixed":
                d.jtype, d.dof = 2, -1
            else:
                d.jtype, d.dof = 1, dof_names.index(j["name"])
            d.xyz[:] = j["xyz"]
            d.rpy[:] = j["rpy"]
            d.axis[:] = j["axis"]
        d.mass = L["mass"]
        # inertia in the link frame (com_rpy is zero for every Go2 link; assert it)
        assert all(abs(v) < 1e-15 for v in L["com_rpy"])
        d.com[:] = L["com"]
        d.inertia[:] = L["inertia"]
    contacts = []
    for leg in ("FL", "FR", "RL", "RR"):
        foot = f"{leg}_foot"
        for s in links[foot]["spheres"]:
            c = ContactDesc()
            c.link = index[foot]
            c.pos[:] = s["xyz"]
            c.radius = s["radius"]
            contacts.append(c)
    carr = (ContactDesc * len(contacts))(*contacts)
    return arr, carr, order


def make_config(sim) -> SimConfig:
    """SimConfig from a spi_amd SimParams-like object (attributes named as below)."""
    c = SimConfig()
    c.dt = sim.dt
    c.decimation = sim.decimation
    c.substeps = sim.substeps
    c.gravity = sim.gravity
    c.kp[:] = list(sim.kp)
    c.kd[:] = list(sim.kd)
    c.q0[:] = list(sim.q0)
    c.effort[:] = list(sim.effort)
    c.action_scale = sim.action_scale
    c.action_clip = sim.action_clip
    c.control_mode = sim.control_mode
    c.motor_model = sim.motor_model
    c.k_contact = sim.k_contact
    c.c_contact = sim.c_contact
    c.mu = sim.mu
    c.c_tangent = sim.c_tangent
    c.fric_vel = sim.fric_vel
    return c


def make_links(specs):
    """LinkDesc array from dicts with keys parent, jtype (0 free, 1 revolute, 2 fixed), dof,
    xyz, rpy, axis, mass, com, inertia (xx yy zz xy xz yz about the CoM)."""
    arr = (LinkDesc * len(specs))()
    for i, s in enumerate(specs):
        d = arr[i]
        d.parent, d.jtype, d.dof = s["parent"], s["jtype"], s.get("dof", -1)
        d.xyz[:] = s.get("xyz", [0, 0, 0])
        d.rpy[:] = s.get("rpy", [0, 0, 0])
        d.axis[:] = s.get("axis", [0, 0, 0])
        d.mass = s["mass"]
        d.com[:] = s.get("com", [0, 0, 0])
        d.inertia[:] = s.get("inertia", [0] * 6)
    return arr


class Oracle:
    """FP64 oracle engine for one configuration."""

    def __init__(self, model: Optional[Dict], sim, f32: bool = False, links=None, contacts=None,
                 lib_path: Optional[str] = None):
        L = self._L = lib(lib_path)
        if model is not None:
            self.links, self.contacts, self.order = unfolded_tree(model, model["dof_names"])
        else:
            self.links = links
            self.contacts = contacts if contacts is not None else (ContactDesc * 0)()
            self.order = list(range(len(links)))
        self.cfg = make_config(sim)
        self.f32 = f32
        ctor = L.spio_new_f32 if f32 else L.spio_new
        self.h = ctor(self.links, len(self.links), self.contacts, len(self.contacts), 0, C.byref(self.cfg))
        assert self.h

    def __del__(self):
        try:
            if self.h:
                (self._L.spio_free_f32 if self.f32 else self._L.spio_free)(self.h)
        except Exception:
            pass

    # -- single-state dynamics (double)
    def forward(self, root, q, qd, tau, par):
        root, q, qd, tau, par = (np.ascontiguousarray(x, dtype=np.float64) for x in (root, q, qd, tau, par))
        a0, qdd, ff, ta = np.zeros(6), np.zeros(12), np.zeros(12), np.zeros(12)
        self._L.spio_forward(self.h, _dp(root), _dp(q), _dp(qd), _dp(tau), _dp(par), _dp(a0), _dp(qdd), _dp(ff), _dp(ta))
        return a0, qdd, ff.reshape(4, 3), ta

    def rnea(self, root, q, qd, a0, qdd, par):
        root, q, qd, a0, qdd, par = (np.ascontiguousarray(x, dtype=np.float64) for x in (root, q, qd, a0, qdd, par))
        tb, t = np.zeros(6), np.zeros(12)
        self._L.spio_rnea(self.h, _dp(root), _dp(q), _dp(qd), _dp(a0), _dp(qdd), _dp(par), _dp(tb), _dp(t))
        return tb, t

    def crba(self, root, q, par):
        root, q, par = (np.ascontiguousarray(x, dtype=np.float64) for x in (root, q, par))
        H = np.zeros((18, 18))
        self._L.spio_crba(self.h, _dp(root), _dp(q), _dp(par), _dp(H))
        return H

    def momentum(self, root, q, qd, par):
        root, q, qd, par = (np.ascontiguousarray(x, dtype=np.float64) for x in (root, q, qd, par))
        h, e = np.zeros(6), C.c_double(0.0)
        self._L.spio_momentum(self.h, _dp(root), _dp(q), _dp(qd), _dp(par), _dp(h), C.byref(e))
        return h, e.value

    def fk(self, root, q):
        root, q = (np.ascontiguousarray(x, dtype=np.float64) for x in (root, q))
        lp, fp = np.zeros((len(self.links), 3)), np.zeros((len(self.contacts), 3))
        self._L.spio_fk(self.h, _dp(root), _dp(q), _dp(lp), _dp(fp))
        return lp, fp

    def physics_step(self, root, q, qd, tau, par):
        root, q, qd = (np.array(x, dtype=np.float64) for x in (root, q, qd))
        tau, par = (np.ascontiguousarray(x, dtype=np.float64) for x in (tau, par))
        self._L.spio_physics_step(self.h, _dp(root), _dp(q), _dp(qd), _dp(tau), _dp(par))
        return root, q, qd

    def control_torque(self, action, q, qd, par):
        action, q, qd, par = (np.ascontiguousarray(x, dtype=np.float64) for x in (action, q, qd, par))
        tau = np.zeros(12)
        self._L.spio_control_torque(C.byref(self.cfg), _dp(action), _dp(q), _dp(qd), _dp(par), _dp(tau))
        return tau

    def simulate(self, par, init37, actions):
        par = np.ascontiguousarray(par, dtype=np.float64)
        init37 = np.ascontiguousarray(init37, dtype=np.float32)
        actions = np.ascontiguousarray(actions, dtype=np.float32)
        T = actions.shape[0]
        out = np.zeros((T, 37), dtype=np.float32)
        rc = self._L.spio_simulate(self.h, _dp(par), _fp(init37), _fp(actions), T, _fp(out))
        return out, rc

    def rollout_cost(self, par, seg_init, seg_actions, seg_targets, cost_mode, weights):
        par = np.ascontiguousarray(par, dtype=np.float64).reshape(-1, P_COUNT)
        seg_init = np.ascontiguousarray(seg_init, dtype=np.float32)
        seg_actions = np.ascontiguousarray(seg_actions, dtype=np.float32)
        seg_targets = np.ascontiguousarray(seg_targets, dtype=np.float32)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        n_seg, H = seg_actions.shape[0], seg_actions.shape[1]
        out = np.zeros((par.shape[0], n_seg))
        self._L.spio_rollout_cost(self.h, _dp(par), par.shape[0], _fp(seg_init), _fp(seg_actions), _fp(seg_targets),
                                n_seg, H, cost_mode, _dp(w), _dp(out))
        return out

    def traj_sens(self, par, idx, init37, actions, n_rows=25):
        """FP64 forward-mode sensitivities: y [T, n_rows] (root 13 + q 12 (+ qd 12) after each
        control step) and J [T, n_rows, P] = d y / d par[idx]."""
        par = np.ascontiguousarray(par, dtype=np.float64)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        init37 = np.ascontiguousarray(init37, dtype=np.float32)
        actions = np.ascontiguousarray(actions, dtype=np.float32)
        T, P = actions.shape[0], idx.shape[0]
        y, J = np.zeros((T, n_rows)), np.zeros((T, n_rows, P))
        rc = self._L.spio_traj_sens(self.h, _dp(par), idx.ctypes.data_as(C.POINTER(C.c_int)), P, _fp(init37),
                                  _fp(actions), T, n_rows, _dp(y), _dp(J))
        return y, J, rc

    def traj_sens_batch(self, par, idx, init37, actions, n_rows=25, n_threads=0):
        """traj_sens of one parameter vector over many trajectories, in parallel: y [n, T, n_rows],
        J [n, T, n_rows, P], number of diverged trajectories."""
        par = np.ascontiguousarray(par, dtype=np.float64)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        init37 = np.ascontiguousarray(init37, dtype=np.float32)
        actions = np.ascontiguousarray(actions, dtype=np.float32)
        n, T, P = actions.shape[0], actions.shape[1], idx.shape[0]
        y, J = np.zeros((n, T, n_rows)), np.zeros((n, T, n_rows, P))
        rc = self._L.spio_traj_sens_batch(self.h, _dp(par), idx.ctypes.data_as(C.POINTER(C.c_int)), P, _fp(init37),
                                          _fp(actions), n, T, n_rows, _dp(y), _dp(J), int(n_threads))
        return y, J, rc

    def simulate_batch(self, par, init37, actions, n_threads=0):
        """FP64 states after every control step of many trajectories under one parameter vector,
        in parallel: [n, T, 37] double, number of trajectories that left the finite range."""
        par = np.ascontiguousarray(par, dtype=np.float64)
        init37 = np.ascontiguousarray(init37, dtype=np.float32)
        actions = np.ascontiguousarray(actions, dtype=np.float32)
        n, T = actions.shape[0], actions.shape[1]
        out = np.zeros((n, T, 37))
        rc = self._L.spio_simulate_batch(self.h, _dp(par), _fp(init37), _fp(actions), n, T, _dp(out), int(n_threads))
        return out, rc

    def rollout_cost_par(self, par, seg_init, seg_actions, seg_targets, cost_mode, weights, n_threads=0):
        """rollout_cost with the (candidate, segment) pairs over threads: [n_cand, n_seg]."""
        par = np.ascontiguousarray(par, dtype=np.float64).reshape(-1, P_COUNT)
        seg_init = np.ascontiguousarray(seg_init, dtype=np.float32)
        seg_actions = np.ascontiguousarray(seg_actions, dtype=np.float32)
        seg_targets = np.ascontiguousarray(seg_targets, dtype=np.float32)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        n_seg, H = seg_actions.shape[0], seg_actions.shape[1]
        out = np.zeros((par.shape[0], n_seg))
        self._L.spio_rollout_cost_par(self.h, _dp(par), par.shape[0], _fp(seg_init), _fp(seg_actions),
                                      _fp(seg_targets), n_seg, H, cost_mode, _dp(w), _dp(out), int(n_threads))
        return out

    def cost_grad(self, par, idx, seg_init, seg_actions, seg_targets, cost_mode, weights):
        """FP64 cost of one parameter vector per segment [n_seg] and its gradient [n_seg, P]."""
        par = np.ascontiguousarray(par, dtype=np.float64)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        seg_init = np.ascontiguousarray(seg_init, dtype=np.float32)
        seg_actions = np.ascontiguousarray(seg_actions, dtype=np.float32)
        seg_targets = np.ascontiguousarray(seg_targets, dtype=np.float32)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        n_seg, H = seg_actions.shape[0], seg_actions.shape[1]
        cost, grad = np.zeros(n_seg), np.zeros((n_seg, idx.shape[0]))
        self._L.spio_cost_grad(self.h, _dp(par), idx.ctypes.data_as(C.POINTER(C.c_int)), idx.shape[0], _fp(seg_init),
                             _fp(seg_actions), _fp(seg_targets), n_seg, H, cost_mode, _dp(w), _dp(cost), _dp(grad))
        return cost, grad

    def rollout_cost_f32(self, par, seg_init, seg_actions, seg_targets, cost_mode, weights, n_threads=0):
        assert self.f32
        par = np.ascontiguousarray(par, dtype=np.float32).reshape(-1, P_COUNT)
        seg_init = np.ascontiguousarray(seg_init, dtype=np.float32)
        seg_actions = np.ascontiguousarray(seg_actions, dtype=np.float32)
        seg_targets = np.ascontiguousarray(seg_targets, dtype=np.float32)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        n_seg, H = seg_actions.shape[0], seg_actions.shape[1]
        out = np.zeros((par.shape[0], n_seg), dtype=np.float32)
        self._L.spio_rollout_cost_f32(self.h, _fp(par), par.shape[0], _fp(seg_init), _fp(seg_actions),
                                    _fp(seg_targets), n_seg, H, cost_mode, _dp(w), _fp(out), n_threads)
        return out


def max_threads() -> int:
    return lib().spio_max_threads()


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().spio_philox4x32_10(c, k, o)
    return list(o)


def philox_normal(seed: int, iteration: int, idx: int) -> float:
    return lib().spio_philox_normal(seed, iteration, idx)


def cem_sample(mean, std, lo, hi, n, seed, iteration):
    mean, std, lo, hi = (np.ascontiguousarray(x, dtype=np.float64) for x in (mean, std, lo, hi))
    out = np.zeros((mean.shape[0], n))
    lib().spio_cem_sample(_dp(mean), _dp(std), _dp(lo), _dp(hi), mean.shape[0], n, seed, iteration, _dp(out))
    return out


def gauss_newton_fp64(o: "Oracle", theta0, rows, segs, weights, seg_w, lo=None, hi=None, iterations=8, lam0=1e-3,
                      tol=1e-10, line_search=0, grad_tol=0.0):
    """FP64 Levenberg-Marquardt on the tracking residuals with the oracle's exact derivatives;
    the same damping / acceptance rule as the GPU refiner (spi_amd.agents.refine): the damped
    Gauss-Newton step is evaluated at its full length and at 1/2, 1/4, ... (line_search
    halvings); the lowest cost below the current one is taken, and the damping is divided by 3
    if the full step won; if none is lower the damping is multiplied by 4.  The Jacobian is
    recomputed only after a step was taken.  Stops after `iterations` iterations, when a step
    gains less than tol * cost, or when the largest relative gradient
    |g_i| max(|theta_i|, 1e-3) / cost falls to grad_tol."""
    rows = np.asarray(rows, dtype=np.int32)
    w = np.asarray(weights, np.float64)
    grp = np.r_[np.zeros(3), np.ones(4), 2 * np.ones(3), 3 * np.ones(3), 4 * np.ones(12), 5 * np.ones(12)].astype(int)

    seg_w = np.asarray(seg_w, np.float64)

    def resid(theta, with_jac):
        if with_jac:
            y, J, _ = o.traj_sens_batch(theta, rows, segs.init, segs.actions, n_rows=37)   # [n, T, 37(, P)]
        else:
            y, _ = o.simulate_batch(theta, segs.init, segs.actions)
        sw = np.sqrt(w[grp][None, :] * seg_w[:, None])                                  # [n, 37]
        R = (sw[:, None, :] * (y - segs.targets.astype(np.float64))).ravel()
        return R, ((sw[:, None, :, None] * J).reshape(-1, len(rows)) if with_jac else None)

    theta = np.asarray(theta0, np.float64).copy()
    r, _ = resid(theta, False)
    cost = float(r @ r)
    lam = lam0
    history = [cost]
    A = b = None
    for _ in range(iterations):
        if A is None:                                   # new point: Jacobian and normal equations
            r, J = resid(theta, True)
            A, b = J.T @ J, J.T @ r
            rel_grad = np.abs(2.0 * b) * np.maximum(np.abs(theta[rows]), 1e-3) / max(cost, 1e-300)
            if rel_grad.max() <= grad_tol:
                break
        eps = 1e-12 * np.trace(A) / len(rows)
        Ad = A + lam * np.diag(np.diag(A)) + eps * np.eye(len(rows))
        delta = -np.linalg.solve(Ad, b)
        best, c_new, k_best = None, np.inf, -1
        for k in range(line_search + 1):
            cand = theta.copy()
            cand[rows] += delta * 0.5 ** k
            if lo is not None:
                cand = np.clip(cand, lo, hi)
            rn, _ = resid(cand, False)
            c_k = float(rn @ rn)
            if np.isfinite(c_k) and c_k < c_new:
                best, c_new, k_best = cand, c_k, k
        if best is not None and c_new < cost:
            gain = cost - c_new
            theta, cost = best, c_new
            if k_best == 0:
                lam = max(lam / 3.0, 1e-9)
            history.append(cost)
            A = b = None
            if gain <= tol * max(cost, 1e-30):
                break
        else:
            lam = min(lam * 4.0, 1e6)
    return theta, cost, history
