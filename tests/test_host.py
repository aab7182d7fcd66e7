"""Host-side logic and the C ABI surface (CPU only: no device calls)."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "spi_hip.h")


def _header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|long long|const char\*)\s+(spi_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_header_symbol():
    import torch  # noqa: F401  (loads the HIP runtime first, as in production)
    from spi_amd import build
    path = build.ensure_built()
    L = ctypes.CDLL(path)
    names = _header_functions()
    assert len(names) >= 13
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (spi_\w+)", out))
    assert exported == set(names), exported ^ set(names)


def test_ctypes_structs_match_c_layout():
    """sizeof/offsetof of SpiSimArgs and SpiCostArgs as compiled by the host compiler."""
    from spi_amd._lib import SpiCostArgs, SpiSimArgs, SpiTangentSpec
    from spi_amd.agents.commands import SpiCommandSpec
    from spi_amd.agents.batch import SpiBatchProblem
    src = r"""
    #include <cstdio>
    #include <cstddef>
    #include "spi_hip.h"
    int main() {
      std::printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu ", sizeof(SpiSimArgs),
                  offsetof(SpiSimArgs, effort), offsetof(SpiSimArgs, motor_model), offsetof(SpiSimArgs, state_limit),
                  sizeof(SpiCostArgs), offsetof(SpiCostArgs, w), sizeof(SpiBatchProblem),
                  offsetof(SpiBatchProblem, seed), offsetof(SpiBatchProblem, mean),
                  offsetof(SpiBatchProblem, best_cost), sizeof(SpiTangentSpec), offsetof(SpiTangentSpec, param));
      std::printf("%zu %zu %zu ", sizeof(SpiCommandSpec), offsetof(SpiCommandSpec, lo), offsetof(SpiCommandSpec, n_bezier));
      std::printf("%zu %zu\n", offsetof(SpiCostArgs, seg_gains), offsetof(SpiBatchProblem, cost));
    }
    """
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.cpp")
        open(c, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["g++", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I",
                        "/opt/rocm/include", c, "-o", exe], check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    expect = [ctypes.sizeof(SpiSimArgs), SpiSimArgs.effort.offset, SpiSimArgs.motor_model.offset,
              SpiSimArgs.state_limit.offset, ctypes.sizeof(SpiCostArgs), SpiCostArgs.w.offset,
              ctypes.sizeof(SpiBatchProblem), SpiBatchProblem.seed.offset, SpiBatchProblem.mean.offset,
              SpiBatchProblem.best_cost.offset, ctypes.sizeof(SpiTangentSpec), SpiTangentSpec.param.offset,
              ctypes.sizeof(SpiCommandSpec), SpiCommandSpec.lo.offset, SpiCommandSpec.n_bezier.offset,
              SpiCostArgs.seg_gains.offset, SpiBatchProblem.cost.offset]
    assert got == expect


def test_segment_gains_known_answer():
    """Recorded gains per segment: each file's own, or the first window of each chunk of
    batch_size windows (scripts/eval.py:192-201, 245-250).  Files A (3 windows, kp 20) and
    B (4 windows, kp 25); chunks of 2: [A A][A B][B B][B] -> kp 20 20 20 20 25 25 25."""
    from spi_amd.data import segment_gains
    kp = np.array([20, 20, 20, 25, 25, 25, 25], np.float32)
    kd = np.array([0.5, 0.5, 0.5, 0.6, 0.6, 0.6, 0.6], np.float32)
    own = segment_gains(kp, kd)
    assert own.dtype == np.float32 and own.shape == (7, 2)
    assert np.array_equal(own[:, 0], kp) and np.array_equal(own[:, 1], kd)
    ch = segment_gains(np.repeat(kp[:, None], 12, 1), np.repeat(kd[:, None], 12, 1), batch_size=2)
    assert np.array_equal(ch[:, 0], [20, 20, 20, 20, 25, 25, 25])
    assert np.array_equal(ch[:, 1], np.array([0.5, 0.5, 0.5, 0.5, 0.6, 0.6, 0.6], np.float32))
    assert np.array_equal(segment_gains(kp, kd, batch_size=100)[:, 0], np.full(7, 20.0))
    with pytest.raises(ValueError):
        segment_gains(kp, kd[:3])


def test_reference_segment_weights_known_answer():
    """scripts/eval.py:279-310 on a hand-worked case: chunks of 4, motion ends at 2 and 6.
    eval_mask per chunk = ~(cumsum(ends) > 0) -> [1,1,0,0]; total_valid = 6."""
    from spi_amd.data import reference_segment_weights
    ends = np.array([0, 0, 1, 0, 0, 0, 1, 0], bool)
    w = reference_segment_weights(ends, 4)
    assert np.allclose(w, np.array([1, 1, 0, 0, 1, 1, 0, 0]) / 6.0)
    # one chunk holding everything: all windows after the first motion end are dropped
    w2 = reference_segment_weights(ends, 8)
    assert np.allclose(w2, np.array([1, 1, 0, 0, 0, 0, 0, 0]) / 6.0)


def test_segment_packing_layout():
    from spi_amd.data import SegmentSet, segment_stride
    n, H = 3, 4
    init = np.arange(n * 37, dtype=np.float32).reshape(n, 37)
    acts = 1000 + np.arange(n * H * 12, dtype=np.float32).reshape(n, H, 12)
    tg = 5000 + np.arange(n * H * 37, dtype=np.float32).reshape(n, H, 37)
    p = SegmentSet(init, acts, tg, np.zeros(n, bool)).packed()
    assert p.shape == (n, segment_stride(H)) == (n, 37 + H * 49)
    for s in range(n):
        assert np.array_equal(p[s, :37], init[s])
        for t in range(H):
            o = 37 + t * 49
            assert np.array_equal(p[s, o:o + 12], acts[s, t])
            assert np.array_equal(p[s, o + 12:o + 49], tg[s, t])


def test_npz_windows_follow_eval_loader(tmp_path):
    """scripts/eval.py:125-152: num_samples = T - H, init at t, target at t + H,
    action window t..t+H-1, last window flagged, per-file gains."""
    from spi_amd.data import load_npz_dataset, save_npz_trajectory
    T, H = 12, 3
    rng = np.random.default_rng(0)
    files = []
    for k, kp in enumerate((20.0, 25.0)):
        st = rng.normal(size=(T, 37)).astype(np.float32)
        ac = rng.normal(size=(T, 12)).astype(np.float32)
        f = str(tmp_path / f"t{k}.npz")
        save_npz_trajectory(f, st, ac, 0.02, kp, 0.5)
        files.append((f, st, ac))
    ds = load_npz_dataset([f for f, _, _ in files], H)
    assert ds.n_seg == 2 * (T - H)
    st0, ac0 = files[0][1], files[0][2]
    assert np.array_equal(ds.init[4], st0[4])
    assert np.array_equal(ds.targets[4, -1], st0[4 + H])
    assert np.array_equal(ds.actions[4], ac0[4:4 + H])
    assert ds.motion_ends.tolist() == ([False] * (T - H - 1) + [True]) * 2
    assert ds.kp[0, 0] == 20.0 and ds.kp[-1, 0] == 25.0


def test_parameter_layout(model):
    from spi_amd.params import PARAM_NAMES, P_COUNT, nominal_params
    assert P_COUNT == 25 and len(PARAM_NAMES) == 25
    p = nominal_params(model)
    assert p[0] == 6.921 and np.allclose(p[1:4], [0.021112, 0, -0.005366])
    assert np.allclose(p[4:10], [0.02448, 0.098077, 0.107, 0.00012166, 0.0014849, -3.12e-05])
    assert np.all(p[10:18] == 1.0)


def test_sinusoid_actions_amplitude():
    from spi_amd.data import sinusoid_actions
    rng = np.random.default_rng(0)
    a = sinusoid_actions(rng, 50, 500, 0.02, 0.25, amp=(0.1, 0.4), freq=(0.5, 3.0))
    target = 0.25 * a                      # PD target offsets
    peak = np.abs(target).max(axis=1)
    assert peak.min() > 0.09 and peak.max() <= 0.4 + 1e-6


def test_config_composition():
    """Hydra-style defaults, _global_ packages and interpolation of the package configs."""
    from spi_amd.config import compose, get_class
    cfg = compose("active_sysid")
    assert cfg.simulator._target_ == "spi_amd.simulator.hipsim_active_sysid.HipSimActiveSysId"
    assert cfg.simulator.config.sim.fps == 200 and cfg.simulator.config.sim.control_decimation == 4
    assert cfg.robot.dof_names[0] == "FL_hip_joint" and len(cfg.robot.body_names) == 19
    assert cfg.env.config.exploration_params == ["mass"] and cfg.env.config.delta_param == 0.1
    from spi_amd.params import SimParams
    sp = SimParams.from_robot_config(cfg.robot, {"sim": dict(cfg.simulator.config.sim)})
    assert sp.dt == 0.005 and sp.decimation == 4 and sp.substeps == 4
    assert np.allclose(sp.kp, 20.0) and np.allclose(sp.kd, 0.5)
    assert get_class("spi_amd.config.AttrDict").__name__ == "AttrDict"


def test_problem_table_packing_matches_c_layout():
    """The column-wise numpy packing of the batched-problem table equals record-by-record
    ctypes packing of SpiBatchProblem (fake device addresses; no GPU needed)."""
    import ctypes
    import math
    from types import SimpleNamespace

    import torch
    from spi_amd._lib import cost_args
    from spi_amd.agents.batch import SpiBatchProblem, pack_problem_table
    rng = np.random.default_rng(0)
    specs, bufs = [], []
    for k in range(5):
        H, n, n_seg = int(rng.choice([5, 20])), int(rng.choice([512, 8192])), int(rng.integers(1, 40))
        specs.append(SimpleNamespace(segments=torch.empty((n_seg, 37 + 49 * H)), horizon=H, n_samples=n,
                                     elite_frac=0.1, alpha=0.25 * k / 5, seed=1000 + k,
                                     seg_gains=torch.empty((n_seg, 2)) if k % 2 else None))
        bufs.append(SimpleNamespace(ptrs=tuple(int(x) for x in rng.integers(1, 2**47, size=12))))
    ca = cost_args(0, [10, 5, 1, 1, 1, 0.05])
    table, prefix, n_tasks = pack_problem_table(specs, bufs, [0, 3, 1, 7, 2], ca, 128, "cpu")
    arr = (SpiBatchProblem * 5)()
    tot = 0
    for i, (s, b) in enumerate(zip(specs, bufs)):
        p = arr[i]
        (p.segments, p.seg_weight, p.mean, p.std, p.lo, p.hi, p.min_std, p.theta, p.cost_seg, p.cost_out, p.best,
         p.best_cost) = b.ptrs
        p.n_seg, p.horizon, p.cost, p.n_samples = s.segments.shape[0], s.horizon, ca, s.n_samples
        p.cost.seg_gains = None if s.seg_gains is None else s.seg_gains.data_ptr()
        p.n_elite = max(1, int(math.ceil(s.elite_frac * s.n_samples - 1e-9)))
        p.alpha, p.iteration, p.seed = s.alpha, [0, 3, 1, 7, 2][i], s.seed
        assert prefix[i].item() == tot
        tot += p.n_seg * ((s.n_samples + 127) // 128)
    assert bytes(table.numpy()) == bytes(arr) and n_tasks == tot


def test_policy_obs_kmap_follows_the_oracle_layout():
    """spi_policy_obs_kmap (host table of the ring gather) against the oracle's assembly: with
    step s holding terms s + term / 100, the oracle's observation after step 20 has in column k
    the term kmap[k][1] of the step kmap[k][0] steps back (0 before the first step)."""
    import torch  # noqa: F401
    import spio_policy as op
    from spi_amd import build
    L = ctypes.CDLL(build.ensure_built())
    km = (ctypes.c_int * 1800)()
    assert L.spi_policy_obs_kmap(km) == 0
    kmap = np.array(list(km)).reshape(900, 2)
    assert L.spi_policy_ring_slots() == op.HISTORY + 1
    keys = sorted(op.TERMS)
    offs = np.cumsum([0] + [op.TERMS[k][0] for k in keys])
    hist = op.empty_history()
    last = 20
    for step in range(last + 1):
        t = {k: step + (offs[i] + np.arange(op.TERMS[k][0])) / 100.0 for i, k in enumerate(keys)}
        obs, hist = op.assemble(t, hist)
    back, term = kmap[:, 0], kmap[:, 1]
    assert back.min() == 0 and back.max() == op.HISTORY and term.min() == 0 and term.max() == 59
    want = np.where(last - back >= 0, (last - back) + term / 100.0, 0.0)
    assert np.allclose(obs, want, atol=1e-12)


def test_policy_entry_points_reject_bad_arguments():
    """Argument checks run before any device work: null pointers, ragged groups and aliased
    state buffers come back as SPI_ERR_INVALID_ARG with a message (no GPU needed)."""
    import torch  # noqa: F401
    from spi_amd import _lib, build
    from spi_amd.params import SimParams
    L = ctypes.CDLL(build.ensure_built())
    L.spi_last_error.restype = ctypes.c_char_p
    args = _lib.sim_args(SimParams())
    vp = ctypes.c_void_p
    fake = vp(4096)
    # ragged: 7 envs in groups of 2
    assert L.spi_policy_step(ctypes.byref(args), fake, 7, 7, 2, fake, vp(8192), fake, fake, 0, fake, fake, None) == 1
    assert b"spi_policy_step" in L.spi_last_error()
    # aliased state buffers
    assert L.spi_policy_step(ctypes.byref(args), fake, 8, 8, 2, fake, fake, fake, fake, 0, fake, fake, None) == 1
    # reward needs groups of at least 2 (a main and an aux env)
    assert L.spi_policy_reward(fake, 4, 1, ctypes.c_float(0.1), ctypes.c_float(0.8), ctypes.c_float(0.8),
                               ctypes.c_float(0.0), fake, fake, fake, fake, None) == 1
    # commands index past the sequence
    assert L.spi_policy_terms(8, 2, fake, fake, 10, 9, 10, ctypes.c_float(0.02), fake, fake, fake, fake, 8, 0, 1,
                              ctypes.c_float(-0.25), ctypes.c_float(0.15), None) == 1
    # ring head outside the ring
    assert L.spi_policy_terms(8, 2, fake, fake, 10, 1, 2, ctypes.c_float(0.02), fake, fake, fake, fake, 8, 15, 1,
                              ctypes.c_float(-0.25), ctypes.c_float(0.15), None) == 1
    # the fused reward + ring terms: groups of at least 2, head inside the ring, delta != 0
    fl = ctypes.c_float
    def reward_terms(group, head, delta):
        return L.spi_policy_reward_terms(8, group, fake, fake, 10, 1, 2, fl(0.02), fake, fake, fake, fake, 8, head, 1,
                                         fl(-0.25), fl(0.15), fl(delta), fl(0.8), fl(0.8), fl(0.0), fake, fake, fake,
                                         fake, None)
    assert reward_terms(1, 0, 0.1) == 1 and b"spi_policy_reward_terms" in L.spi_last_error()
    assert reward_terms(2, 15, 0.1) == 1
    assert reward_terms(2, 0, 0.0) == 1
    dims = (ctypes.c_int * 3)(70, 33, 12)
    # work buffer too small / ring too narrow
    assert L.spi_mlp_forward(fake, 10, 2, dims, fake, fake, ctypes.c_longlong(10), fake, None) == 1
    assert L.spi_mlp_forward_ring(fake, 8, 15, 60, 0, fake, ctypes.c_float(100.0), 10, 2, dims, fake, fake,
                                  ctypes.c_longlong(1000), fake, None) == 1
    assert b"ring" in L.spi_last_error()


def test_derivative_entry_points_reject_bad_arguments():
    """The reverse-mode and residual-Jacobian entries (both lane layouts) check their arguments
    before any device work: work buffers below the size function, decimation above 8, the
    reference cost for a sum-of-squares Jacobian, and tangent counts above the limit come back as
    SPI_ERR_INVALID_ARG with a message naming the entry (no GPU needed)."""
    import torch  # noqa: F401
    from spi_amd import _lib, build
    from spi_amd.params import SimParams
    L = ctypes.CDLL(build.ensure_built())
    L.spi_last_error.restype = ctypes.c_char_p
    for name in ("spi_adjoint_work_floats", "spi_adjoint_legs_work_floats"):
        getattr(L, name).restype = ctypes.c_longlong
    vp, ll = ctypes.c_void_p, ctypes.c_longlong
    fake = vp(4096)
    args = _lib.sim_args(SimParams())
    cost = _lib.SpiCostArgs(0, (ctypes.c_float * 6)(10, 5, 1, 1, 1, 0.05))
    # one lane of 37 floats per rollout: checkpoints + one block's control-step states; four
    # lanes of 19 floats: checkpoints + one block's physics-step states (decimation 4)
    one = L.spi_adjoint_work_floats(64, 4, 20, 10)
    four = L.spi_adjoint_legs_work_floats(64, 4, 20, 10, args.decimation)
    assert args.decimation == 4
    assert one == 64 * 4 * (2 + 10) * 37 and four == 64 * 4 * 4 * (2 + 10 * 4) * 19
    assert L.spi_adjoint_legs_work_floats(64, 4, 20, 10, 9) == -1
    for name, need in (("spi_rollout_cost_adjoint", one), ("spi_rollout_cost_adjoint_legs", four)):
        fn = getattr(L, name)
        assert fn(ctypes.byref(args), ctypes.byref(cost), fake, 64, 64, fake, 4, 20, 10, fake, ll(need - 1), fake,
                  fake, None) == 1
        assert name.encode() in L.spi_last_error() and b"work buffer" in L.spi_last_error()
        bad = _lib.sim_args(SimParams())
        bad.decimation = 9
        assert fn(ctypes.byref(bad), ctypes.byref(cost), fake, 64, 64, fake, 4, 20, 10, fake, ll(need), fake, fake,
                  None) == 1
        assert fn(ctypes.byref(args), ctypes.byref(cost), fake, 64, 64, fake, 4, 20, 0, fake, ll(need), fake, fake,
                  None) == 1
    ref_cost = _lib.SpiCostArgs(1, (ctypes.c_float * 6)(10, 5, 0, 0, 1, 0))
    spec = _lib.tangent_spec(list(range(22)))
    too_many = _lib.tangent_spec(list(range(22)))
    too_many.n_tan = 32                                        # n_tan + 1 lanes > 32
    for name in ("spi_residual_jacobian", "spi_residual_jacobian_legs"):
        fn = getattr(L, name)
        assert fn(ctypes.byref(args), ctypes.byref(ref_cost), ctypes.byref(spec), fake, 1, 1, fake, fake, 4, 5,
                  fake, fake, None) == 1
        assert name.encode() in L.spi_last_error() and b"tracking" in L.spi_last_error()
        assert fn(ctypes.byref(args), ctypes.byref(cost), ctypes.byref(too_many), fake, 1, 1, fake, fake, 4, 5,
                  fake, fake, None) == 1
        assert fn(ctypes.byref(args), ctypes.byref(cost), ctypes.byref(spec), fake, 1, 2, fake, fake, 4, 5,
                  fake, fake, None) == 1                          # ld < n_cand


def test_empty_inputs_return_ok_and_oversized_are_refused():
    """Zero candidates / segments / samples / excitations return SPI_OK without touching the
    device (the calls run here, with no GPU); records that do not fit the 64 KB LDS staging
    (horizon 400 for segments, 1400 for excitations) and CEM updates of more than
    SPI_MAX_CEM_SAMPLES samples are refused on the host."""
    import torch  # noqa: F401
    from spi_amd import _lib
    from spi_amd.params import SimParams
    L = _lib.lib()
    args = _lib.sim_args(SimParams())
    cost = _lib.SpiCostArgs(0, (ctypes.c_float * 6)(10, 5, 1, 1, 1, 0.05))
    spec = _lib.tangent_spec(list(range(22)))
    f = 4096                                                   # never dereferenced
    a, c, t = ctypes.byref(args), ctypes.byref(cost), ctypes.byref(spec)
    OK, TOO_LARGE = 0, 3
    for fn in (L.spi_rollout_cost, L.spi_rollout_cost_legs):
        assert fn(a, c, f, 0, 0, f, 4, 5, f, None) == OK                       # no candidates
        assert fn(a, c, f, 8, 8, f, 0, 5, f, None) == OK                       # no segments
        assert fn(a, c, f, 8, 8, f, 4, 400, f, None) == TOO_LARGE              # record > 64 KB
    for fn, size in ((L.spi_rollout_cost_adjoint, lambda *x: L.spi_adjoint_work_floats(*x)),
                     (L.spi_rollout_cost_adjoint_legs, lambda *x: L.spi_adjoint_legs_work_floats(*x, 4))):
        assert fn(a, c, f, 0, 0, f, 4, 20, 10, f, 0, f, f, None) == OK
        assert fn(a, c, f, 8, 8, f, 4, 400, 10, f, size(8, 4, 400, 10), f, f, None) == TOO_LARGE
    for fn in (L.spi_traj_sensitivity, L.spi_traj_sensitivity_legs):
        assert fn(a, t, f, 0, 0, f, f, 4, 50, f, None, f, None) == OK         # no samples
        assert fn(a, t, f, 8, 8, f, f, 0, 50, f, None, f, None) == OK         # no excitations
        assert fn(a, t, f, 8, 8, f, f, 4, 1400, f, None, f, None) == TOO_LARGE
    for fn in (L.spi_residual_jacobian, L.spi_residual_jacobian_legs):
        assert fn(a, c, t, f, 0, 0, f, f, 4, 5, f, f, None) == OK
    assert L.spi_rollout_cost_grad(a, c, t, f, 0, 0, f, 4, 5, f, f, None) == OK
    assert L.spi_rollout_cost_grad(a, c, t, f, 8, 8, f, 4, 400, f, f, None) == TOO_LARGE
    # the elite sort runs in LDS: at most SPI_MAX_CEM_SAMPLES samples per update
    assert L.spi_cem_update(f, f, 8193, 22, 8193, 820, 0.0, f, f, f, f, f, None, None) == TOO_LARGE
    assert b"spi_cem_update" in L.spi_last_error()


def test_reference_recipes_known_answers(tmp_path, model, sim, theta0):
    """The reference's data recipes (scripts/data/sine.py, jump.py, stand.py) with the oracle as
    the simulator: 250 steps of 20 ms, thigh actions A sin(2 pi f t) with (A, f) = (0.8, 1),
    (1.2, 1.5), calf actions their negative, hips zero; sine and jump start at base height
    0.32 m; frame k is the state after step k (continuing the oracle's own rollout from frame
    k - 1 with action k reproduces frame k); the npz written in the reference's schema loads into
    250 - H windows."""
    import spio
    from spi_amd.data import (RECIPES, load_npz_dataset, record_recipe, recipe_actions, save_npz_trajectory,
                              standing_state)
    o = spio.Oracle(model, sim)

    def simulate(rows, init, acts):
        return np.stack([o.simulate(np.asarray(r, np.float64), i0, a)[0] for r, i0, a in zip(rows, init, acts)])

    dt = sim.dt * sim.decimation
    for name, (amp, freq, height) in RECIPES.items():
        a = recipe_actions(name, dt)
        assert a.shape == (250, 12)
        t = np.arange(250) * dt
        assert np.allclose(a[:, 1], amp * np.sin(2 * np.pi * freq * t), atol=1e-6)
        assert np.array_equal(a[:, [4, 7, 10]], np.repeat(a[:, 1:2], 3, axis=1))
        assert np.array_equal(a[:, [2, 5, 8, 11]], -np.repeat(a[:, 1:2], 4, axis=1))
        assert np.all(a[:, [0, 3, 6, 9]] == 0.0)
    init = standing_state(0.30)
    states, acts = record_recipe(simulate, theta0, sim, "sine", init, duration=0.4)
    assert states.shape == (20, 37) and acts.shape == (20, 12)
    s0 = init.copy()
    s0[2] = 0.32
    first, _ = o.simulate(theta0, s0, acts[:1])
    assert np.array_equal(first[0], states[0])
    again, _ = o.simulate(theta0, states[9], acts[10:11])
    assert np.abs(again[0] - states[10]).max() < 1e-5
    f = str(tmp_path / "sine.npz")
    save_npz_trajectory(f, states, acts, dt, 20.0, 0.5)
    ds = load_npz_dataset([f], 5)
    assert ds.n_seg == 15 and np.array_equal(ds.init[3], states[3]) and np.array_equal(ds.targets[3, -1], states[8])
    assert np.array_equal(ds.actions[3, 0], acts[3])
