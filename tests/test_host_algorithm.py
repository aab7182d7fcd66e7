"""The kernel's specialised algorithm (spi_dyn.h) built for the host in FP32, against the FP64
generic-tree oracle (CPU only).  Catches algorithm errors without a GPU; the GPU parity tests
then check the device build of the same code."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "host_dyn", "spi_dyn_host.cpp")


@pytest.fixture(scope="module")
def host_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hostdyn") / "libspi_dyn_host.so")
    cmd = ["/opt/rocm/bin/hipcc", "--offload-host-only", "-O2", "-std=c++17", "-fPIC", "-shared",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "spi-active_amd", "csrc"), SRC, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("hipcc host build unavailable: " + r.stderr[-300:])
    L = ctypes.CDLL(out)
    fp = ctypes.POINTER(ctypes.c_float)
    L.spi_host_forward_dynamics.argtypes = [ctypes.c_void_p, fp, fp, fp, fp, fp]
    L.spi_host_control_step.argtypes = [ctypes.c_void_p, fp, fp, fp]
    return L


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def test_host_forward_dynamics_matches_oracle(host_lib, model, sim, theta0):
    import spio
    from spi_amd._lib import sim_args
    from test_gpu_parity import _rand_params, _realistic_states
    o = spio.Oracle(model, sim)
    args = sim_args(sim)
    rng = np.random.default_rng(21)
    n = 64
    states = _f(_realistic_states(model, sim, theta0, n, seed=22))
    rows = _f(_rand_params(rng, theta0, n))
    tau = _f(rng.normal(0, 5, (n, 12)))
    worst = 0.0
    for i in range(n):
        acc, ff = np.zeros(18, np.float32), np.zeros(12, np.float32)
        host_lib.spi_host_forward_dynamics(ctypes.byref(args), _p(rows[i]), _p(states[i]), _p(tau[i]), _p(acc), _p(ff))
        a0, qdd, fo, _ = o.forward(states[i, :13], states[i, 13:25], states[i, 25:37], tau[i], rows[i])
        ref = np.r_[a0, qdd]
        worst = max(worst, np.abs(acc - ref).max() / max(np.abs(ref).max(), 1.0))
        # the penetration is a difference of ~0.3 m heights: its FP32 rounding (~3e-8 m) times the
        # contact stiffness (k = 5e4 N/m, up to 2k where the damping saturates) is ~3e-3 N
        assert np.allclose(ff.reshape(4, 3), fo, rtol=1e-4, atol=1e-2)
    assert worst <= 1e-3, worst


def test_host_rollout_matches_oracle(host_lib, model, sim, theta0):
    import spio
    from spi_amd._lib import sim_args
    from spi_amd.data import sinusoid_actions
    from test_gpu_parity import _realistic_states
    o = spio.Oracle(model, sim)
    args = sim_args(sim)
    rng = np.random.default_rng(23)
    s0 = _f(_realistic_states(model, sim, theta0, 4, seed=24)[2])
    s0[25:37] *= 0.2
    acts = sinusoid_actions(rng, 1, 10, 0.02, 0.25)[0]
    ref, rc = o.simulate(theta0, s0, acts)
    st = s0.copy()
    par = _f(theta0)
    for t in range(10):
        host_lib.spi_host_control_step(ctypes.byref(args), _p(par), _p(st), _p(_f(acts[t])))
        assert np.abs(st[:7] - ref[t, :7]).max() < 1e-4
        assert np.abs(st[13:25] - ref[t, 13:25]).max() < 1e-3


def test_cpu_baseline_rollout_cost_matches_oracle(model, sim, theta0):
    """The host-CPU baseline (spi_dyn.h built for x86 with OpenMP, tests/host_dyn/Makefile) that
    bench.py times: its tracking costs match the FP64 oracle (2e-3 relative), like the GPU
    kernel's, on 6 candidates x 4 noisy segments x H 5, and do not depend on the thread count."""
    import spio
    from spi_amd import _lib
    from spi_amd.data import synthetic_segments
    from spi_amd.rollout import DEFAULT_TRACKING_WEIGHTS
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "host_dyn"), "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("host baseline build unavailable: " + r.stderr[-300:])
    L = ctypes.CDLL(os.path.join(ROOT, "tests", "host_dyn", "lib", "libspi_rollout_cpu.so"))
    o = spio.Oracle(model, sim)

    def simulate(rows, init, acts):
        return np.stack([o.simulate(np.asarray(rw, np.float64), i0, a)[0] for rw, i0, a in zip(rows, init, acts)])

    H, n_seg, n_cand = 5, 4, 6
    segs = synthetic_segments(simulate, theta0, sim, n_seg, H, seed=4)
    rng = np.random.default_rng(3)
    rows = np.repeat(theta0[None], n_cand, axis=0)
    rows[1:, 0] *= rng.uniform(0.8, 1.2, n_cand - 1)
    rows[1:, 10:18] *= rng.uniform(0.9, 1.1, (n_cand - 1, 8))
    rows = rows.astype(np.float32)
    par = _f(rows.T)
    seg = _f(segs.packed())
    a = _lib.sim_args(sim)
    c = _lib.cost_args(0, DEFAULT_TRACKING_WEIGHTS)
    outs = []
    for threads in (1, 3):
        out = np.zeros((n_seg, n_cand), np.float32)
        used = L.spi_cpu_rollout_cost(ctypes.byref(a), ctypes.byref(c), _p(par), ctypes.c_int(n_cand),
                                      ctypes.c_int(n_cand), _p(seg), ctypes.c_int(n_seg), ctypes.c_int(H), _p(out),
                                      ctypes.c_int(threads))
        assert used == threads
        outs.append(out)
    assert np.array_equal(outs[0], outs[1])
    ref = o.rollout_cost(rows.astype(np.float64), segs.init, segs.actions, segs.targets, 0,
                         np.asarray(DEFAULT_TRACKING_WEIGHTS, np.float64)).T
    err = np.abs(outs[0] - ref)
    assert np.all(err <= 2e-3 * ref + 1e-4), err.max()
