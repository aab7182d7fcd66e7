"""The plain GPU rollout path: the same Go2 model and tracking cost written as PyTorch eager ops.

This is the comparison point for the fused HIP engine, built the way the reference's GPU
pipeline runs: a Python loop over control steps, decimation and integration sub-intervals,
with one (or a few) framework kernels per operation over all lanes (the reference's
``LeggedRobotBase.step`` decimation loop, legged_robot_base.py:169-209, around
``gym.simulate``, isaacgym.py:622-626).  Physics: floating-base articulated-body algorithm with
6x6 spatial matrices (batched ``matmul`` over lanes x legs), penalty foot contact, smooth
Coulomb joint friction, semi-implicit Euler -- the model the HIP engine integrates, in any
torch dtype and on any device (float64 on the CPU reproduces the FP64 oracle).

It is a measurement baseline only: nothing in the engine calls it.
"""
from __future__ import annotations

import math
from typing import Dict, Optional

import numpy as np
import torch

from ..model import load_model
from ..params import SimParams

_COST_TRACKING = 0


def _skew(v: torch.Tensor) -> torch.Tensor:
    z = torch.zeros_like(v[..., 0])
    return torch.stack([torch.stack([z, -v[..., 2], v[..., 1]], -1),
                        torch.stack([v[..., 2], z, -v[..., 0]], -1),
                        torch.stack([-v[..., 1], v[..., 0], z], -1)], -2)


def _xform(E: torch.Tensor, r: torch.Tensor, mm) -> torch.Tensor:
    """Plucker motion transform [E 0; -E r^ E] (angular part first)."""
    top = torch.cat([E, torch.zeros_like(E)], -1)
    bot = torch.cat([-mm(E, _skew(r).expand(E.shape)), E], -1)
    return torch.cat([top, bot], -2)


def _crm(v: torch.Tensor) -> torch.Tensor:
    w, u = _skew(v[..., :3]), _skew(v[..., 3:])
    return torch.cat([torch.cat([w, torch.zeros_like(w)], -1), torch.cat([u, w], -1)], -2)


def _crf(v: torch.Tensor) -> torch.Tensor:
    return -_crm(v).transpose(-1, -2)


def _mv(M: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
    return (M @ v.unsqueeze(-1)).squeeze(-1)


def _mm(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    return A @ B


# the same products as broadcast multiply + sum (elementwise kernels, no batched GEMM calls)
def _mv_bcast(M: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
    return (M * v.unsqueeze(-2)).sum(-1)


def _mm_bcast(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    return (A.unsqueeze(-1) * B.unsqueeze(-3)).sum(-2)


def _spatial_inertia(m, h, I):
    """[I_o h^; h^T m 1] from mass, first moment h = m c and the 3x3 inertia about the origin."""
    hx = _skew(h)
    eye = torch.eye(3, dtype=I.dtype, device=I.device).expand(I.shape)
    return torch.cat([torch.cat([I, hx], -1), torch.cat([hx.transpose(-1, -2), m[..., None, None] * eye], -1)], -2)


def _inertia_about_origin(m, c, ixx, iyy, izz, ixy, ixz, iyz):
    cx, cy, cz = c[..., 0], c[..., 1], c[..., 2]
    Ixx = ixx + m * (cy * cy + cz * cz)
    Iyy = iyy + m * (cx * cx + cz * cz)
    Izz = izz + m * (cx * cx + cy * cy)
    Ixy = ixy - m * cx * cy
    Ixz = ixz - m * cx * cz
    Iyz = iyz - m * cy * cz
    return torch.stack([torch.stack([Ixx, Ixy, Ixz], -1), torch.stack([Ixy, Iyy, Iyz], -1),
                        torch.stack([Ixz, Iyz, Izz], -1)], -2)


def _axis_rotation(k: int, c: torch.Tensor, s: torch.Tensor) -> torch.Tensor:
    """Rotation (child -> parent) by angle q about coordinate axis k, given cos/sin."""
    o, z = torch.ones_like(c), torch.zeros_like(c)
    if k == 0:
        rows = [[o, z, z], [z, c, -s], [z, s, c]]
    elif k == 1:
        rows = [[c, z, s], [z, o, z], [-s, z, c]]
    else:
        rows = [[c, -s, z], [s, c, z], [z, z, o]]
    return torch.stack([torch.stack(r, -1) for r in rows], -2)


def _solve_spd6(A: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Batched Gaussian elimination without pivoting (A symmetric positive definite) + back substitution."""
    A, b = A.clone(), b.clone()
    n = A.shape[-1]
    for k in range(n - 1):
        l = A[:, k + 1:, k] / A[:, k, k, None]
        A[:, k + 1:, :] -= l[:, :, None] * A[:, k, None, :]
        b[:, k + 1:] -= l * b[:, k, None]
    x = torch.empty_like(b)
    for k in range(n - 1, -1, -1):
        x[:, k] = (b[:, k] - (A[:, k, k + 1:] * x[:, k + 1:]).sum(-1)) / A[:, k, k]
    return x


def _quat_to_rot(q: torch.Tensor) -> torch.Tensor:
    x, y, z, w = q.unbind(-1)
    return torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        torch.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        torch.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)


class PlainTorchRollout:
    """Batched rollouts and tracking cost over (segment, candidate) lanes in torch eager ops."""

    def __init__(self, sim: SimParams, model: Optional[Dict] = None, cost_mode: int = _COST_TRACKING,
                 weights=(10.0, 5.0, 1.0, 1.0, 1.0, 0.05), device="cuda", dtype=torch.float32,
                 state_limit: float = 1.0e4, products: str = "bmm"):
        """products: "bmm" (batched matrix products) or "bcast" (broadcast multiply + sum)."""
        model = model or load_model()
        if products not in ("bmm", "bcast"):
            raise ValueError(products)
        self._mv, self._mm = (_mv, _mm) if products == "bmm" else (_mv_bcast, _mm_bcast)
        self.sim, self.dev, self.dt_ = sim, torch.device(device), dtype
        self.cost_mode, self.w = int(cost_mode), [float(x) for x in weights]
        self.state_limit = state_limit
        t = lambda a: torch.as_tensor(np.asarray(a, np.float64), dtype=dtype, device=self.dev)
        bodies = model["bodies"]
        # legs: bodies 1 + 3 L + j; every leg has the same axis sequence
        self.axes = [int(np.argmax(np.abs(bodies[1 + j]["axis"]))) for j in range(3)]
        for L in range(4):
            for j in range(3):
                assert int(np.argmax(np.abs(bodies[1 + 3 * L + j]["axis"]))) == self.axes[j]
                assert bodies[1 + 3 * L + j]["axis"][self.axes[j]] > 0
        self.r = [t([bodies[1 + 3 * L + j]["origin"] for L in range(4)]) for j in range(3)]          # [4,3]
        self.I = []
        for j in range(3):
            m = t([bodies[1 + 3 * L + j]["mass"] for L in range(4)])
            c = t([bodies[1 + 3 * L + j]["com"] for L in range(4)])
            ii = t([bodies[1 + 3 * L + j]["inertia"] for L in range(4)])
            Io = _inertia_about_origin(m, c, *ii.unbind(-1))
            self.I.append(_spatial_inertia(m, m[:, None] * c, Io))                                 # [4,6,6]
        fx = model["base_fixed"]
        self.fixed_m = float(fx["mass"])
        self.fixed_c = t(fx["com"])
        self.fixed_i = t(fx["inertia"])
        self.foot = t([ct["pos"] for ct in model["contacts"]])                                      # [4,3]
        self.rad = t([ct["radius"] for ct in model["contacts"]])                                    # [4]
        self.kp, self.kd = t(sim.kp).reshape(4, 3), t(sim.kd).reshape(4, 3)
        self.q0, self.effort = t(sim.q0).reshape(4, 3), t(sim.effort).reshape(4, 3)
        self.hip = t([0.5 if sim.control_mode == 1 else 1.0, 1.0, 1.0])
        self.S = [torch.zeros(6, dtype=dtype, device=self.dev) for _ in range(3)]
        for j in range(3):
            self.S[j][self.axes[j]] = 1.0

    # ------------------------------------------------------------------ parameters
    def lane_params(self, theta: torch.Tensor, n_seg: int) -> Dict[str, torch.Tensor]:
        """theta [25, n_cand] (SoA) -> per-lane tensors for lanes ordered (segment, candidate)."""
        th = theta.to(self.dt_).t().repeat(n_seg, 1)                    # [N, 25]
        m, c = th[:, 0], th[:, 1:4]
        Io = _inertia_about_origin(m, c, *th[:, 4:10].unbind(-1))
        fm = torch.full_like(m, self.fixed_m)
        fIo = _inertia_about_origin(fm, self.fixed_c.expand_as(c), *self.fixed_i.expand(m.shape[0], 6).unbind(-1))
        I0 = _spatial_inertia(m + fm, m[:, None] * c + fm[:, None] * self.fixed_c, Io + fIo)
        return {"I0": I0, "kp_s": th[:, 10:14], "kd_s": th[:, 14:18], "fric": th[:, 18:22], "motor": th[:, 22:25]}

    # ------------------------------------------------------------------ control law
    def torques(self, lp, act, q, qd):
        s = self.sim
        a = act.clamp(-s.action_clip, s.action_clip).reshape(-1, 4, 3) * s.action_scale * self.hip
        t = lp["kp_s"][:, :, None] * self.kp * (a + self.q0 - q) - lp["kd_s"][:, :, None] * self.kd * qd
        t = torch.maximum(torch.minimum(t, self.effort), -self.effort)
        if s.motor_model == 1:
            t = t * lp["motor"][:, :1, None]
        elif s.motor_model == 2:
            t = t * lp["motor"][:, None, :]
        elif s.motor_model == 3:
            A = lp["motor"][:, None, :]
            t = A * torch.tanh(t / A)
        return t

    # ------------------------------------------------------------------ dynamics
    def forward_dynamics(self, lp, st, tau_m):
        """st: dict p [N,3], quat [N,4] xyzw, w, v [N,3] (base coords), q, qd [N,4,3]."""
        s = self.sim
        N = st["p"].shape[0]
        R = _quat_to_rot(st["quat"])
        v0 = torch.cat([st["w"], st["v"]], -1)                              # [N,6]
        tau = tau_m - lp["fric"][:, :, None] * torch.tanh(st["qd"] / s.fric_vel)
        cq, sq = torch.cos(st["q"]), torch.sin(st["q"])
        X, v, c, E = [], [], [], []
        vp = v0[:, None, :].expand(N, 4, 6)
        for j in range(3):
            Rj = _axis_rotation(self.axes[j], cq[..., j], sq[..., j])       # child -> parent
            Xj = _xform(Rj.transpose(-1, -2), self.r[j], self._mm)
            vj = self._mv(Xj, vp) + self.S[j] * st["qd"][..., j, None]
            c.append(self._mv(_crm(vj), self.S[j] * st["qd"][..., j, None]))
            X.append(Xj)
            v.append(vj)
            E.append(Rj)
            vp = vj
        # calf pose and foot contact (world frame z up, plane z = 0)
        Rb = R[:, None]
        pc = self.r[0] + self._mv(E[0], self.r[1] + self._mv(E[1], self.r[2]))      # calf origin, base coords
        pcw = st["p"][:, None] + self._mv(Rb, pc)
        Rc = self._mm(self._mm(self._mm(Rb.expand(E[0].shape), E[0]), E[1]), E[2])
        centre = pcw + self._mv(Rc, self.foot)
        pen = self.rad - centre[..., 2]
        down = -self.rad[:, None] * Rc[..., 2, :]                          # Rc^T (0, 0, -rad)
        pcl = self.foot + down
        vl = v[2][..., 3:] + torch.cross(v[2][..., :3], pcl, dim=-1)
        vw = self._mv(Rc, vl)
        kpen = s.k_contact * torch.clamp(pen, min=1e-12)          # masked below where pen <= 0
        fn = kpen * (1.0 - torch.tanh(s.c_contact * vw[..., 2] / kpen))
        vt = torch.sqrt(vw[..., 0] ** 2 + vw[..., 1] ** 2)
        ft = torch.minimum(s.c_tangent * vt, s.mu * fn)
        scale = torch.where(vt > 0, ft / torch.where(vt > 0, vt, torch.ones_like(vt)), torch.zeros_like(vt))
        fw = torch.stack([-scale * vw[..., 0], -scale * vw[..., 1], fn], -1)
        fw = torch.where((pen > 0)[..., None], fw, torch.zeros_like(fw))
        fl = self._mv(Rc.transpose(-1, -2), fw)
        fext = torch.cat([torch.cross(pcl, fl, dim=-1), fl], -1)
        # backward pass
        IA = self.I[2].expand(N, 4, 6, 6)
        pA = self._mv(_crf(v[2]), self._mv(self.I[2], v[2])) - fext
        U, D, u = [None] * 3, [None] * 3, [None] * 3
        for j in (2, 1, 0):
            k = self.axes[j]
            U[j] = IA[..., :, k]
            D[j] = U[j][..., k]
            u[j] = tau[..., j] - pA[..., k]
            Ia = IA - U[j][..., :, None] * U[j][..., None, :] / D[j][..., None, None]
            pa = pA + self._mv(Ia, c[j]) + U[j] * (u[j] / D[j])[..., None]
            XT = X[j].transpose(-1, -2)
            IA_up, p_up = self._mm(self._mm(XT, Ia), X[j]), self._mv(XT, pa)
            if j > 0:
                IA = self.I[j - 1] + IA_up
                pA = self._mv(_crf(v[j - 1]), self._mv(self.I[j - 1], v[j - 1])) + p_up
            else:
                IA0 = lp["I0"] + IA_up.sum(1)
                p0 = self._mv(_crf(v0), self._mv(lp["I0"], v0)) + p_up.sum(1)
        a0 = _solve_spd6(IA0, -p0)                                   # gravity-free
        qdd = []
        ap = a0[:, None, :].expand(N, 4, 6)
        for j in range(3):
            aj = self._mv(X[j], ap) + c[j]
            x = (u[j] - (U[j] * aj).sum(-1)) / D[j]
            qdd.append(x)
            ap = aj + self.S[j] * x[..., None]
        g = torch.zeros_like(st["p"])
        g[:, 2] = -s.gravity
        acc_v = a0[:, 3:] + self._mv(R.transpose(-1, -2), g)
        return a0[:, :3], acc_v, torch.stack(qdd, -1)

    def integrate(self, st, dw, dv, qdd, h):
        st["w"] = st["w"] + h * dw
        st["v"] = st["v"] + h * dv
        st["qd"] = st["qd"] + h * qdd
        st["q"] = st["q"] + h * st["qd"]
        R = _quat_to_rot(st["quat"])
        st["p"] = st["p"] + h * self._mv(R, st["v"])
        hh = 0.5 * h
        t2 = hh * hh * (st["w"] ** 2).sum(-1)
        sfac = hh * (1.0 - t2 / 6.0)
        cw = 1.0 - 0.5 * t2 + t2 * t2 / 24.0
        d = sfac[:, None] * st["w"]
        qx, qy, qz, qw = st["quat"].unbind(-1)
        dx, dy, dz = d.unbind(-1)
        x = qw * dx + qx * cw + qy * dz - qz * dy
        y = qw * dy - qx * dz + qy * cw + qz * dx
        z = qw * dz + qx * dy - qy * dx + qz * cw
        w = qw * cw - qx * dx - qy * dy - qz * dz
        qn = torch.stack([x, y, z, w], -1)
        st["quat"] = qn / qn.norm(dim=-1, keepdim=True)

    # ------------------------------------------------------------------ rollouts
    def state_from37(self, s37: torch.Tensor):
        R = _quat_to_rot(s37[:, 3:7])
        Rt = R.transpose(-1, -2)
        return {"p": s37[:, 0:3].clone(), "quat": s37[:, 3:7].clone(), "v": self._mv(Rt, s37[:, 7:10]),
                "w": self._mv(Rt, s37[:, 10:13]), "q": s37[:, 13:25].reshape(-1, 4, 3).clone(),
                "qd": s37[:, 25:37].reshape(-1, 4, 3).clone()}

    def state_to37(self, st) -> torch.Tensor:
        R = _quat_to_rot(st["quat"])
        return torch.cat([st["p"], st["quat"], self._mv(R, st["v"]), self._mv(R, st["w"]), st["q"].reshape(-1, 12),
                          st["qd"].reshape(-1, 12)], -1)

    def control_step(self, lp, st, act):
        s = self.sim
        h = s.dt / s.substeps
        for _ in range(s.decimation):
            tau = self.torques(lp, act, st["q"], st["qd"])
            for _ in range(s.substeps):
                dw, dv, qdd = self.forward_dynamics(lp, st, tau)
                self.integrate(st, dw, dv, qdd, h)

    def step_cost(self, st, tg):
        s37 = self.state_to37(st)
        d2 = (s37 - tg) ** 2
        e = [d2[:, 0:3].sum(-1), d2[:, 3:7].sum(-1), d2[:, 7:10].sum(-1), d2[:, 10:13].sum(-1),
             d2[:, 13:25].sum(-1), d2[:, 25:37].sum(-1)]
        if self.cost_mode == _COST_TRACKING:
            return sum(wk * ek for wk, ek in zip(self.w, e))
        return self.w[0] * e[0].sqrt() + self.w[1] * e[1].sqrt() + self.w[4] * e[4].sqrt()

    @torch.no_grad()
    def segment_costs(self, theta: torch.Tensor, init: torch.Tensor, actions: torch.Tensor,
                      targets: torch.Tensor, steps: Optional[int] = None) -> torch.Tensor:
        """theta [25, n_cand]; init [n_seg, 37]; actions [n_seg, H, 12]; targets [n_seg, H, 37]
        -> cost [n_seg, n_cand] (+inf for lanes whose state diverged).  ``steps`` < H stops early
        (for bounded timing samples)."""
        n_seg, H = actions.shape[0], actions.shape[1]
        n_cand = theta.shape[1]
        H = H if steps is None else min(H, steps)
        lp = self.lane_params(theta.to(self.dev), n_seg)
        rep = lambda x: x.to(self.dev, self.dt_).repeat_interleave(n_cand, 0)
        st = self.state_from37(rep(init))
        acc = torch.zeros(n_seg * n_cand, dtype=self.dt_, device=self.dev)
        ok = torch.ones(n_seg * n_cand, dtype=torch.bool, device=self.dev)
        for t in range(H):
            self.control_step(lp, st, rep(actions[:, t]))
            if self.cost_mode == _COST_TRACKING or t == H - 1:
                acc = acc + self.step_cost(st, rep(targets[:, t]))
            nrm = st["w"].abs().sum(-1) + st["v"].abs().sum(-1) + st["qd"].abs().sum((-1, -2))
            ok = ok & torch.isfinite(nrm) & (nrm < self.state_limit)
        acc = torch.where(ok & torch.isfinite(acc), acc, torch.full_like(acc, math.inf))
        return acc.reshape(n_seg, n_cand)
