"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host code).

`make -C oracle sanitize` builds the oracle with -fsanitize=address,undefined (and no recovery
from undefined behaviour).  A child Python process, with the GCC address-sanitizer runtime
preloaded, drives every oracle entry point the checks use (forward, inverse and composite
rigid-body dynamics, momentum, kinematics, physics and control steps, single and batched
simulation, rollout costs in FP64 and FP32, cost gradients and trajectory sensitivities) on
small random inputs through both builds: any out-of-bounds access, use after free or
undefined operation stops it with a report, and the two builds must agree bit for bit.  The
same harness run on a small library with a planted out-of-bounds read must report it.  CPU
only.
"""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

DRIVER = r'''
import sys
import numpy as np
sys.path[:0] = [ROOT + "/oracle", ROOT + "/spi-active_amd"]
import spio
from spi_amd.model import load_model
from spi_amd.params import SimParams, nominal_params
from spi_amd.data import standing_state

maps = open("/proc/self/maps").read()
print("asan runtime mapped:", "libasan" in maps)
model, sim = load_model(), SimParams()
theta = np.asarray(nominal_params(), np.float64)
san = ROOT + "/oracle/lib/libspi_oracle_san.so"
out = {}
for tag, path in (("plain", None), ("san", san)):
    o = spio.Oracle(model, sim, lib_path=path)
    f = spio.Oracle(model, sim, f32=True, lib_path=path)
    s0 = standing_state(0.30).astype(np.float64)
    rng = np.random.default_rng(0)            # the same draws for both builds
    root, q, qd = s0[:13].copy(), s0[13:25] + 0.1 * rng.standard_normal(12), rng.standard_normal(12)
    tau = 5.0 * rng.standard_normal(12)
    r = []
    a0, qdd, ff, ta = o.forward(root, q, qd, tau, theta)
    r += [a0, qdd, ff, ta]
    r += list(o.rnea(root, q, qd, a0, qdd, theta))
    r.append(o.crba(root, q, theta))
    h, e = o.momentum(root, q, qd, theta)
    r += [h, np.array([e])]
    r += list(o.fk(root, q))
    r += list(o.physics_step(root, q, qd, tau, theta))
    r.append(o.control_torque(0.3 * rng.standard_normal(12), q, qd, theta))
    init = np.repeat(standing_state(0.30)[None].astype(np.float32), 3, axis=0)
    acts = (0.3 * rng.standard_normal((3, 6, 12))).astype(np.float32)
    traj, rc = o.simulate(theta, init[0], acts[0])
    r.append(traj)
    r.append(o.simulate_batch(theta, init, acts, n_threads=2)[0])
    tg = traj[None].repeat(3, axis=0) + 0.01 * rng.standard_normal((3, 6, 37)).astype(np.float32)
    w = np.ones(37)
    pars = np.repeat(theta[None], 4, axis=0)
    pars[:, 0] *= 1 + 0.05 * rng.standard_normal(4)
    r.append(o.rollout_cost(pars, init, acts, tg, 0, w))
    r.append(o.rollout_cost_par(pars, init, acts, tg, 1, w, n_threads=2))
    r.append(f.rollout_cost_f32(pars.astype(np.float32), init, acts, tg, 0, w, n_threads=2))
    idx = np.array([0, 1, 4, 10, 14, 18], np.int32)
    r += list(o.cost_grad(theta, idx, init, acts, tg, 0, w))
    r += list(o.traj_sens(theta, idx, init[0], acts[0]))
    r += list(o.traj_sens_batch(theta, idx, init, acts, n_threads=2))
    out[tag] = r
bad = [k for k, (a, b) in enumerate(zip(out["plain"], out["san"]))
       if not np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)]
print("sanitized library mapped:", "libspi_oracle_san" in open("/proc/self/maps").read())
print("results", len(out["plain"]), "differ", bad)
sys.exit(1 if bad else 0)
'''


def _asan_runtime():
    r = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def _env(rt):
    return dict(os.environ, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
                UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2")


@pytest.mark.parametrize("flag, body, arg, report", [
    ("address", "int a[4] = {1, 2, 3, 4}; volatile int* p = a; return p[i];", 4, "stack-buffer-overflow"),
    ("undefined", "volatile int big = 2147483647; return big + i;", 1, "signed integer overflow"),
])
def test_harness_reports_a_planted_mistake(tmp_path, flag, body, arg, report):
    """The harness itself: each sanitizer stops a child on the mistake planted for it."""
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("the GCC address-sanitizer runtime is not installed")
    src = tmp_path / "planted.cpp"
    src.write_text('extern "C" int planted(int i) { ' + body + ' }\n')
    so = str(tmp_path / "libplanted.so")
    subprocess.run(["g++", "-O1", "-g", "-fPIC", "-shared", f"-fsanitize={flag}", "-fno-sanitize-recover=all",
                    "-o", so, str(src)], check=True)
    code = f"import ctypes; L = ctypes.CDLL({so!r}); print(L.planted(0), flush=True); print(L.planted({arg}), flush=True)"
    p = subprocess.run([sys.executable, "-c", code], env=_env(rt), capture_output=True, text=True, timeout=120)
    assert p.stdout.split() and p.returncode != 0, (p.stdout, p.stderr[-2000:])
    assert report in p.stderr, p.stderr[-2000:]


def test_oracle_clean_under_address_and_undefined_sanitizers():
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("the GCC address-sanitizer runtime is not installed")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "all", "sanitize"], check=True)
    p = subprocess.run([sys.executable, "-c", DRIVER.replace("ROOT", repr(ROOT))], env=_env(rt),
                       capture_output=True, text=True, timeout=600)
    report = p.stdout[-2000:] + p.stderr[-4000:]
    assert "AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr, report
    assert p.returncode == 0, report
    assert "differ []" in p.stdout, report
    assert "asan runtime mapped: True" in p.stdout and "sanitized library mapped: True" in p.stdout, report
