"""Pins the oracle's time stepping (spi_oracle.hpp, Engine::integrate / physics_step /
control_step) by closed forms of the discrete scheme itself, on the CPU.

Every rollout parity test measures the kernels against this integrator, so a convention
mistake shared by the oracle and a kernel (explicit Euler, a world-frame angular velocity in
the quaternion increment, the wrong sub-interval length, a torque held too long) would pass
all of them.  Each test below uses a system whose forward dynamics are known exactly (uniform
gravity on a free tree, a torque-free body spinning about a principal axis, two coaxial
rotors driven by an internal torque), so the discrete sequences of the scheme

    v_{k+1} = v_k + h a_k,   x_{k+1} = x_k + h v_{k+1},   q_{k+1} = q_k (x) exp(h w_b / 2)

with h = dt / substeps have closed forms, and the oracle must reproduce them to rounding.

Each test is also run against oracle builds with one planted mistake (``make mutants``,
SPIO_MUTATION in spi_oracle.hpp) and must FAIL there for the mistakes its docstring names:
    1 explicit Euler          2 world-frame w, right-multiplied     3 body w, left-multiplied
    4 h = dt                  5 torque held for the whole control step
"""
import copy
import os
import subprocess

import numpy as np
import pytest

import spio

ORACLE_DIR = os.path.dirname(spio.__file__)
MUTANTS = (1, 2, 3, 4, 5)


@pytest.fixture(scope="module")
def mutant_libs():
    r = subprocess.run(["make", "-C", ORACLE_DIR, "-s", "-j5", "mutants"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return {k: spio.mutant_path(k) for k in MUTANTS}


def _quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def _qexp(rotvec):
    """Unit quaternion (xyzw) of the rotation by the vector rotvec (angle |rotvec|)."""
    th = np.linalg.norm(rotvec)
    if th == 0.0:
        return np.array([0.0, 0.0, 0.0, 1.0])
    return np.r_[np.sin(th / 2) * rotvec / th, np.cos(th / 2)]


def _same_rotation(qa, qb, tol):
    return min(np.abs(qa - qb).max(), np.abs(qa + qb).max()) < tol


def _unit_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


# ------------------------------------------------------------------------------------------
def check_free_fall(sim, model, theta0, lib_path=None):
    """The whole robot, contact-free, zero joint torque and velocity, no angular velocity:
    every body falls with g, so the joints stay put and the base follows the semi-implicit
    sequences v_k = v_0 + k h g and p_N = p_0 + h sum_{k=1..N} v_k.
    Catches mutation 1 (explicit Euler: p_N = p_0 + h sum_{k=0..N-1} v_k, off by N h^2 g) and
    mutation 4 (h = dt: off by a factor ~substeps in the velocity)."""
    o = spio.Oracle(model, sim, lib_path=lib_path)
    par = theta0.copy()
    par[18:22] = 0.0
    rng = np.random.default_rng(11)
    root = np.zeros(13)
    root[0:3] = [0.3, -0.2, 20.0]
    root[3:7] = _unit_quat(rng)
    v0 = np.array([0.4, -0.7, 1.1])
    root[7:10] = v0
    q = np.asarray(sim.q0, float) + rng.normal(0, 0.2, 12)
    qd = np.zeros(12)
    n_phys = 10
    h = sim.dt / sim.substeps
    g = np.array([0.0, 0.0, -sim.gravity])
    p0, quat0 = root[0:3].copy(), root[3:7].copy()
    for _ in range(n_phys):
        root, q1, qd = o.physics_step(root, q, qd, np.zeros(12), par)
    N = n_phys * sim.substeps
    k = np.arange(1, N + 1)
    v_expect = v0 + N * h * g
    p_expect = p0 + h * (N * v0 + g * h * k.sum())
    assert np.abs(root[7:10] - v_expect).max() < 1e-12 * 10
    assert np.abs(root[0:3] - p_expect).max() < 1e-12 * 20
    assert np.abs(q1 - q).max() < 1e-12 and np.abs(qd).max() < 1e-12
    assert _same_rotation(root[3:7], quat0, 1e-12) and np.abs(root[10:13]).max() < 1e-12


def _single_body(sim, lib_path=None):
    links = spio.make_links([{"parent": -1, "jtype": 0, "mass": 1.0}])
    return spio.Oracle(None, sim, links=links, lib_path=lib_path)


def check_principal_axis_spin(sim, lib_path=None):
    """A torque-free rigid body spinning about a principal axis of its inertia keeps its
    body-frame angular velocity w_b; from any orientation q_0 the scheme gives
    q_N = q_0 (x) exp(N h w_b / 2) (increments about one axis compose exactly), and the world
    angular velocity R(q_N) w_b stays R(q_0) w_b.  The start orientation is far from the
    identity, so R(q_0) w_b differs from w_b.
    Catches mutation 2 (q (x) exp(h R w_b / 2)), mutation 3 (exp(h w_b / 2) (x) q) and
    mutation 4 (N dt instead of N dt / substeps)."""
    s2 = copy.deepcopy(sim)
    s2.gravity = 0.0
    o = _single_body(s2, lib_path)
    rng = np.random.default_rng(12)
    par = np.zeros(25)
    par[0] = 2.5
    par[4:7] = [0.11, 0.23, 0.37]     # principal axes = body axes, CoM at the frame origin
    for axis in range(3):
        wb = np.zeros(3)
        wb[axis] = [3.0, -2.2, 4.1][axis]
        root = np.zeros(13)
        root[2] = 3.0
        root[3:7] = _unit_quat(rng)
        R0 = _quat_matrix(root[3:7])
        assert np.linalg.norm(R0 @ wb - wb) > 0.5 * np.linalg.norm(wb)
        root[10:13] = R0 @ wb
        q0 = root[3:7].copy()
        n_phys = 25
        for _ in range(n_phys):
            root, _, _ = o.physics_step(root, np.zeros(12), np.zeros(12), np.zeros(12), par)
        N = n_phys * sim.substeps
        h = sim.dt / sim.substeps
        q_expect = _qmul(q0, _qexp(N * h * wb))
        assert _same_rotation(root[3:7], q_expect, 1e-12), (root[3:7], q_expect)
        assert np.abs(root[10:13] - R0 @ wb).max() < 1e-12
        assert np.abs(root[0:3] - [0, 0, 3.0]).max() < 1e-12 and np.abs(root[7:10]).max() < 1e-12


def _coaxial_rotors(sim, lib_path=None):
    """Free base plus one rotor on a revolute joint about the base's z axis; both CoMs at the
    base origin, both inertias diagonal, so every rotation stays about body z."""
    links = spio.make_links([
        {"parent": -1, "jtype": 0, "mass": 1.0},
        {"parent": 0, "jtype": 1, "dof": 0, "axis": [0, 0, 1], "mass": 1.7, "com": [0, 0, 0],
         "inertia": [0.12, 0.19, 0.29, 0, 0, 0]},
    ])
    return spio.Oracle(None, sim, links=links, lib_path=lib_path)


def check_internal_torque_sequences(sim, lib_path=None):
    """A joint torque tau between a free base (I_bz) and a coaxial rotor (I_rz) from rest:
    base w_z' = -tau / I_bz, rotor absolute w_z' = tau / I_rz, so qdd = tau (1/I_rz + 1/I_bz)
    exactly.  The scheme gives qd_k = k h qdd, q_N = q_0 + h^2 qdd N (N + 1) / 2, and the base
    orientation q_0 (x) exp(-h^2 (tau / I_bz) N (N + 1) / 2 z / 2).
    Catches mutation 1 (q_N with N (N - 1) / 2), mutations 2 and 3 (the base turns about
    body z from a start orientation far from the identity) and mutation 4 (h = dt)."""
    s2 = copy.deepcopy(sim)
    s2.gravity = 0.0
    o = _coaxial_rotors(s2, lib_path)
    rng = np.random.default_rng(13)
    par = np.zeros(25)
    par[0] = 4.0
    par[4:7] = [0.21, 0.33, 0.47]
    tau_j = 0.8
    I_bz, I_rz = 0.47, 0.29
    alpha = tau_j * (1.0 / I_rz + 1.0 / I_bz)
    root = np.zeros(13)
    root[2] = 2.0
    root[3:7] = _unit_quat(rng)
    quat0 = root[3:7].copy()
    q, qd = np.zeros(12), np.zeros(12)
    q[0] = 0.3
    tau = np.zeros(12)
    tau[0] = tau_j
    n_phys = 12
    for _ in range(n_phys):
        root, q, qd = o.physics_step(root, q, qd, tau, par)
    N = n_phys * sim.substeps
    h = sim.dt / sim.substeps
    assert abs(qd[0] - N * h * alpha) < 1e-12
    assert abs(q[0] - (0.3 + h * h * alpha * N * (N + 1) / 2)) < 1e-12
    wz_b = -N * h * tau_j / I_bz
    R = _quat_matrix(root[3:7])
    assert np.abs(R.T @ root[10:13] - [0, 0, wz_b]).max() < 1e-12
    ang = -h * h * (tau_j / I_bz) * N * (N + 1) / 2
    assert _same_rotation(root[3:7], _qmul(quat0, _qexp(np.array([0.0, 0.0, ang]))), 1e-12)


def check_substeps_hold_torque(model, sim, theta0, lib_path=None):
    """One physics step with `substeps = s` equals s physics steps of an engine configured with
    substeps = 1 and dt / s, under the same held torque, from a state in ground contact.
    Catches mutation 4 (each sub-interval as long as the whole physics step)."""
    from conftest import random_state
    rng = np.random.default_rng(14)
    o4 = spio.Oracle(model, sim, lib_path=lib_path)
    s1 = copy.deepcopy(sim)
    s1.substeps, s1.dt = 1, sim.dt / sim.substeps
    o1 = spio.Oracle(model, s1, lib_path=lib_path)
    for _ in range(4):
        s = random_state(rng, sim, height=0.27)
        tau = rng.normal(0, 4, 12)
        ra, qa, qda = o4.physics_step(s[:13], s[13:25], s[25:37], tau, theta0)
        rb, qb, qdb = s[:13].copy(), s[13:25].copy(), s[25:37].copy()
        for _ in range(sim.substeps):
            rb, qb, qdb = o1.physics_step(rb, qb, qdb, tau, theta0)
        assert np.abs(np.r_[ra, qa, qda] - np.r_[rb, qb, qdb]).max() < 1e-13
        assert np.abs(np.r_[ra, qa, qda] - s).max() > 1e-4       # the step moved the state


def check_torque_per_physics_step(model, sim, theta0, lib_path=None):
    """A control step recomputes the PD + motor-model torque from the current joint state at
    each of its `decimation` physics steps (legged_robot_base.py:201-209): simulate() over
    two control steps equals the loop of control_torque + physics_step (stored as float32).
    Catches mutation 5 (torque computed once per control step)."""
    from spi_amd.data import standing_state
    o = spio.Oracle(model, sim, lib_path=lib_path)
    init = standing_state()
    rng = np.random.default_rng(15)
    acts = rng.uniform(-1.0, 1.0, (2, 12)).astype(np.float32)
    traj, rc = o.simulate(theta0, init, acts)
    assert rc == 0
    s = init.astype(np.float64)
    root, q, qd = s[:13], s[13:25], s[25:37]
    for t in range(2):
        for _ in range(sim.decimation):
            tau = o.control_torque(acts[t].astype(np.float64), q, qd, theta0)
            root, q, qd = o.physics_step(root, q, qd, tau, theta0)
        ref = np.r_[root, q, qd]
        assert np.abs(traj[t] - ref).max() < 2e-6 * (1 + np.abs(ref).max()), np.abs(traj[t] - ref).max()


CHECKS = {
    "free_fall": (lambda L, m, s, t: check_free_fall(s, m, t, L), {1, 4}),
    "principal_axis_spin": (lambda L, m, s, t: check_principal_axis_spin(s, L), {2, 3, 4}),
    "internal_torque": (lambda L, m, s, t: check_internal_torque_sequences(s, L), {1, 2, 3, 4}),
    "substeps_hold_torque": (lambda L, m, s, t: check_substeps_hold_torque(m, s, t, L), {4}),
    "torque_per_physics_step": (lambda L, m, s, t: check_torque_per_physics_step(m, s, t, L), {5}),
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_integrator_closed_form(name, model, sim, theta0):
    CHECKS[name][0](None, model, sim, theta0)


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_closed_form_catches_planted_mistakes(name, mutant_libs, model, sim, theta0):
    """The same check fails on every oracle build whose planted mistake it names."""
    fn, caught = CHECKS[name]
    for k in sorted(caught):
        with pytest.raises(AssertionError):
            fn(mutant_libs[k], model, sim, theta0)


def test_mutants_agree_with_the_oracle_where_no_mistake_applies(mutant_libs, model, sim, theta0):
    """The mutant builds differ from the oracle only through their planted step: a check that
    does not name a mutation passes on that build (so a failure above is the mistake's, not
    a broken build's)."""
    for name, (fn, caught) in CHECKS.items():
        for k in MUTANTS:
            if k not in caught:
                fn(mutant_libs[k], model, sim, theta0)
