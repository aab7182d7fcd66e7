"""The rollout adjoint (spi_adj.h, reverse mode with checkpoints) built for the host, against
central differences of the same build in double precision and the FP64 oracle's forward-mode
gradient, over motor models, cost modes and checkpoint intervals (CPU only; the GPU tests run
the device build of the same code)."""
import ctypes as C
import dataclasses
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "host_dyn", "spi_adj_host.cpp")


def _build(tmp, double):
    out = str(tmp / ("adj_d.so" if double else "adj_f.so"))
    cmd = ["/opt/rocm/bin/hipcc", "--offload-host-only", "-O2", "-std=c++17", "-fPIC", "-shared",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "spi-active_amd", "csrc"), SRC, "-o", out]
    if double:
        cmd.insert(1, "-DSPI_HOST_DOUBLE")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("hipcc host build unavailable: " + r.stderr[-300:])
    L = C.CDLL(out)
    dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
    L.spi_host_rollout_grad.argtypes = [C.c_void_p, C.c_void_p, dp, fp, C.c_int, C.c_int, dp, dp]
    L.spi_host_rollout_cost.argtypes = [C.c_void_p, C.c_void_p, dp, fp, C.c_int]
    L.spi_host_rollout_cost.restype = C.c_double
    return L


@pytest.fixture(scope="module")
def libs(tmp_path_factory):
    t = tmp_path_factory.mktemp("hostadj")
    return _build(t, True), _build(t, False)


def _segments(model, sim, theta0, H, seed, unit_motor=False):
    import spio
    from spi_amd.agents.sysid import ground_truth, identification_prior
    from spi_amd.data import synthetic_segments
    o = spio.Oracle(model, sim)

    def simulate(theta, init, acts):
        return np.stack([o.simulate(np.asarray(theta)[i], init[i], acts[i])[0] for i in range(init.shape[0])])
    prior = identification_prior(theta0)
    truth = ground_truth(prior, seed=seed)
    if unit_motor:                     # gain-type motor models: unit-order gains
        truth[22:25] = 1.0
    return o, prior, synthetic_segments(simulate, truth, sim, n_seg=2, horizon=H, seed=seed + 1)


@pytest.mark.parametrize("motor,mode,H,every", [(3, 0, 10, 4), (0, 1, 12, 5), (1, 0, 8, 8)])
def test_host_adjoint_matches_differences_and_oracle(libs, model, sim, theta0, motor, mode, H, every):
    from spi_amd._lib import cost_args, sim_args
    Ld, Lf = libs
    sim = dataclasses.replace(sim, motor_model=motor)
    o, prior, segs = _segments(model, sim, theta0, H, 2 + motor, unit_motor=motor in (1, 2))
    w = [10.0, 5.0, 1.0, 1.0, 1.0, 0.05] if mode == 0 else [10.0, 5.0, 0.0, 0.0, 1.0, 0.0]
    a, c = sim_args(sim), cost_args(mode, w)
    theta = prior.mean.copy()
    if motor in (1, 2):
        theta[22:25] = 1.0
    dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
    packed = segs.packed()
    for s in range(2):
        rec = np.ascontiguousarray(packed[s])
        cost, g, gf = C.c_double(), np.zeros(25), np.zeros(25)
        assert Ld.spi_host_rollout_grad(C.byref(a), C.byref(c), theta.ctypes.data_as(dp), rec.ctypes.data_as(fp), H,
                                        every, C.byref(cost), g.ctypes.data_as(dp)) == 0
        costf = C.c_double()
        Lf.spi_host_rollout_grad(C.byref(a), C.byref(c), theta.ctypes.data_as(dp), rec.ctypes.data_as(fp), H, every,
                                 C.byref(costf), gf.ctypes.data_as(dp))
        fd = np.zeros(25)
        for k in range(25):
            h = 1e-6 * max(abs(theta[k]), 1e-2)
            tp, tm = theta.copy(), theta.copy()
            tp[k] += h
            tm[k] -= h
            f = lambda t: Ld.spi_host_rollout_cost(C.byref(a), C.byref(c), t.ctypes.data_as(dp),
                                                   rec.ctypes.data_as(fp), H)
            fd[k] = (f(tp) - f(tm)) / (2 * h)
        scale = np.abs(fd).max()
        assert np.abs(g - fd).max() < 1e-4 * scale, (s, np.abs(g - fd).max() / scale)
        cr, gr = o.cost_grad(theta, np.arange(25), segs.init[s:s + 1], segs.actions[s:s + 1],
                             segs.targets[s:s + 1], mode, w)
        # the specialised code keeps the model constants in float (spi_go2_model.h): a ~1e-7
        # relative model difference from the oracle's double tree, which the contacts of a
        # rollout amplify to ~1e-6 of the cost (measured 7e-8 .. 1.6e-6 over these six segments)
        assert abs(cost.value - cr[0]) < 5e-6 * max(cr[0], 1e-3)
        assert np.abs(g - gr[0]).max() < 1e-3 * scale
        # the float build (the device arithmetic) against the FP64 oracle
        assert np.abs(gf - gr[0]).max() < 2e-2 * scale, np.abs(gf - gr[0]).max() / scale
