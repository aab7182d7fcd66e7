"""ORACLE — test infrastructure only.  numpy restatement of the active-exploration command
generators (spigym/agents/sysid/active_sysid.py:259-400): trial parameters in the order the
reference suggests them -> command sequences [total_steps, n_dims] in float32.  Only tests may
import this module."""
from __future__ import annotations

import numpy as np


def _binom(n: int, k: int) -> int:
    # active_sysid.py:402-410: float product, then int
    if k > n - k:
        k = n - k
    r = 1
    for i in range(k):
        r *= (n - i) / (i + 1)
    return int(r)


def commands_from_params(p, mode, ranges, default, sampled_idx, n_updates, steps, poly_degree=3, n_bezier=4):
    """p: the trial's parameters in suggestion order (flat).  Returns float32 [n_updates*steps, n_dims]."""
    p = [float(x) for x in p]
    ranges = np.asarray(ranges, np.float32)
    D = len(sampled_idx)
    if mode == "constant":                                   # :295-314, dim-major suggestions
        upd = np.zeros((n_updates, D), np.float32)
        k = 0
        for d in range(D):
            for u in range(n_updates):
                upd[u, d] = p[k]
                k += 1
        sampled = np.repeat(upd, steps, axis=0)
    elif mode == "polynomial":                               # :316-361
        nc = poly_degree + 1
        t = np.linspace(0.0, 1.0, steps, dtype=np.float32)
        parts, k = [], 0
        for u in range(n_updates):
            coeffs = []
            for d in range(D):
                coeffs.append(p[k:k + nc])
                k += nc
            seg = np.zeros((steps, D), np.float32)
            for d in range(D):
                poly = np.zeros_like(t)
                for i, c in enumerate(coeffs[d]):
                    poly += c * (t ** i)
                poly = np.tanh(poly)
                vmin, vmax = ranges[sampled_idx[d]]
                seg[:, d] = vmin + (poly + 1.0) * 0.5 * (vmax - vmin)
            parts.append(seg)
        sampled = np.vstack(parts)
    elif mode == "bezier":                                   # :363-400
        t = np.linspace(0, 1, steps, dtype=np.float32)
        parts, k = [], 0
        n = n_bezier - 1
        for u in range(n_updates):
            cps = np.zeros((n_bezier, D), np.float32)
            for d in range(D):
                for j in range(n_bezier):
                    cps[j, d] = p[k]
                    k += 1
            seg = np.zeros((steps, D), np.float32)
            for d in range(D):
                for i in range(steps):
                    val = 0.0
                    for j in range(n_bezier):
                        val += _binom(n, j) * (t[i] ** j) * ((1 - t[i]) ** (n - j)) * cps[j, d]
                    vmin, vmax = ranges[sampled_idx[d]]
                    seg[i, d] = np.clip(val, vmin, vmax)
            parts.append(seg)
        sampled = np.vstack(parts)
    else:
        raise ValueError(mode)
    full = np.zeros((sampled.shape[0], len(default)), np.float32)
    full[:, :] = default
    for idx, s in enumerate(sampled_idx):
        full[:, s] = sampled[:, idx]
    return full


def param_names(mode, sampled_idx, n_updates, poly_degree=3, n_bezier=4):
    """Trial parameter names in suggestion order (active_sysid.py:306-307, 334-336, 373-377)."""
    if mode == "constant":
        return [f"dim_{a}_update_{u}" for a in sampled_idx for u in range(n_updates)]
    if mode == "polynomial":
        return [f"dim_{a}_update_{u}_coeff_{i}" for u in range(n_updates) for a in sampled_idx
                for i in range(poly_degree + 1)]
    return [f"dim_{a}_update_{u}_bezier_{j}" for u in range(n_updates) for a in sampled_idx for j in range(n_bezier)]

