"""Pins the FP64 oracle against what the mathematics fixes (CPU only).

The reference's physics is a closed binary (IsaacGym/PhysX) with no tests or fixtures,
so the oracle's dynamics are pinned by invariants and textbook closed forms, chosen so
that a plausible mistake (a sign in a cross product, a transposed transform, a dropped
velocity-product term) fails at least one of them:
  * ABA vs RNEA vs CRBA (three different algorithms on one tree),
  * kinetic-energy balance dT/dt = tau . qd with the mass matrix from CRBA,
  * spatial-momentum conservation under internal torques (RK4 in time),
  * free fall of the centre of mass, Euler's equations of a single free body,
  * the simple pendulum and the planar double pendulum (Lagrangian closed forms),
  * static stance: total normal force = m g,
  * Philox4x32-10 against the published Random123 known-answer vectors.
"""
import copy

import numpy as np
import pytest

import spio
from conftest import random_state


def _zero_par(theta0):
    p = theta0.copy()
    p[18:22] = 0.0   # no joint friction
    return p


def _high_state(rng, sim):
    s = random_state(rng, sim)
    s[2] = 5.0      # no foot touches the ground
    return s


def test_philox_known_answer_vectors():
    """Random123 kat_vectors for philox4x32-10."""
    assert spio.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert spio.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6,
                                                                       0x6d5451fd]
    assert spio.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_philox_normals_are_standard():
    z = np.array([spio.philox_normal(42, 3, i) for i in range(20000)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.02


def test_aba_rnea_crba_consistent(model, sim, theta0):
    o = spio.Oracle(model, sim)
    rng = np.random.default_rng(0)
    for k in range(20):
        s = random_state(rng, sim, height=0.27)
        tau = rng.normal(0, 5, 12)
        par = theta0.copy()
        par[0] *= rng.uniform(0.7, 1.3)
        a0, qdd, ff, ta = o.forward(s[:13], s[13:25], s[25:37], tau, par)
        tb, t = o.rnea(s[:13], s[13:25], s[25:37], a0, qdd, par)
        assert np.abs(tb).max() < 1e-9 and np.abs(t - ta).max() < 1e-9
        H = o.crba(s[:13], s[13:25], par)
        hb, hj = o.rnea(s[:13], s[13:25], s[25:37], np.zeros(6), np.zeros(12), par)
        assert np.allclose(H @ np.r_[a0, qdd] + np.r_[hb, hj], np.r_[np.zeros(6), ta], atol=1e-9)
        assert np.allclose(H, H.T, atol=1e-12) and np.linalg.eigvalsh(H).min() > 0


def test_mass_matrix_total_mass(model, sim, theta0):
    o = spio.Oracle(model, sim)
    H = o.crba(np.r_[0, 0, 1, 0, 0, 0, 1, np.zeros(6)], sim.q0, theta0)
    assert np.allclose(np.diag(H)[3:6], 15.019, atol=1e-12)


def test_kinetic_energy_balance(model, sim, theta0):
    """dT/dt = nu^T H nu_dot + 1/2 nu^T H_dot nu must equal the joint power tau . qd
    (no gravity, no contact, no friction).  H from CRBA, nu_dot from ABA."""
    s2 = copy.deepcopy(sim)
    s2.gravity = 0.0
    o = spio.Oracle(model, s2)
    par = _zero_par(theta0)
    rng = np.random.default_rng(1)
    for _ in range(10):
        s = _high_state(rng, sim)
        tau = rng.normal(0, 3, 12)
        a0, qdd, _, ta = o.forward(s[:13], s[13:25], s[25:37], tau, par)
        R = _quat_matrix(s[3:7])
        nu = np.r_[R.T @ s[10:13], R.T @ s[7:10], s[25:37]]
        H = o.crba(s[:13], s[13:25], par)
        eps = 1e-6
        Hp = o.crba(s[:13], s[13:25] + eps * s[25:37], par)
        Hm = o.crba(s[:13], s[13:25] - eps * s[25:37], par)
        Hdot = (Hp - Hm) / (2 * eps)
        dT = nu @ H @ np.r_[a0, qdd] + 0.5 * nu @ Hdot @ nu
        assert abs(dT - ta @ s[25:37]) < 1e-6 * (1 + abs(ta @ s[25:37])), (dT, ta @ s[25:37])


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


def _deriv(o, s, tau, par):
    a0, qdd, _, _ = o.forward(s[:13], s[13:25], s[25:37], tau, par)
    R = _quat_matrix(s[3:7])
    wb, vb = R.T @ s[10:13], R.T @ s[7:10]
    d = np.zeros(37)
    d[0:3] = s[7:10]
    d[3:7] = 0.5 * _qmul(s[3:7], np.r_[wb, 0.0])
    d[7:10] = R @ (np.cross(wb, vb) + a0[3:6])
    d[10:13] = R @ a0[0:3]
    d[13:25] = s[25:37]
    d[25:37] = qdd
    return d


def _rk4(o, s, tau, par, h, n):
    for _ in range(n):
        k1 = _deriv(o, s, tau, par)
        k2 = _deriv(o, s + 0.5 * h * k1, tau, par)
        k3 = _deriv(o, s + 0.5 * h * k2, tau, par)
        k4 = _deriv(o, s + h * k3, tau, par)
        s = s + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        s[3:7] /= np.linalg.norm(s[3:7])
    return s


def test_momentum_conserved_under_internal_torques(model, sim, theta0):
    """No gravity, no contact: world spatial momentum constant while joints are driven."""
    s2 = copy.deepcopy(sim)
    s2.gravity = 0.0
    o = spio.Oracle(model, s2)
    par = _zero_par(theta0)
    rng = np.random.default_rng(2)
    s = _high_state(rng, sim)
    tau = rng.normal(0, 3, 12)
    h0, e0 = o.momentum(s[:13], s[13:25], s[25:37], par)
    s1 = _rk4(o, s, tau, par, 2e-4, 250)
    h1, e1 = o.momentum(s1[:13], s1[13:25], s1[25:37], par)
    assert np.abs(h1 - h0).max() < 1e-7, (h0, h1)
    assert abs(e1 - e0) > 1e-3  # the torques did work: the test is not vacuous


def test_free_fall_com_acceleration(model, sim, theta0):
    """Zero joint torque, no contact: linear momentum grows at m g exactly."""
    o = spio.Oracle(model, sim)
    par = _zero_par(theta0)
    rng = np.random.default_rng(3)
    s = _high_state(rng, sim)
    zero = np.zeros(12)
    h0, _ = o.momentum(s[:13], s[13:25], s[25:37], par)
    T = 0.05
    s1 = _rk4(o, s, zero, par, 2.5e-4, 200)
    h1, _ = o.momentum(s1[:13], s1[13:25], s1[25:37], par)
    assert np.allclose(h1[3:6] - h0[3:6], [0, 0, -15.019 * 9.81 * T], atol=1e-7)


def _single_body_oracle(sim):
    links = spio.make_links([{"parent": -1, "jtype": 0, "mass": 1.0}])
    return spio.Oracle(None, sim, links=links)


def test_single_free_body_euler_equations(sim):
    """Base alone, CoM at the frame origin: ABA equals Euler's equations
    I w_dot + w x I w = 0 and v_dot_body = -w x v (zero classical acceleration)."""
    s2 = copy.deepcopy(sim)
    s2.gravity = 0.0
    o = _single_body_oracle(s2)
    rng = np.random.default_rng(4)
    for _ in range(5):
        par = np.zeros(25)
        par[0] = 3.0
        par[4:7] = rng.uniform(0.05, 0.5, 3)
        par[7:10] = rng.uniform(-0.01, 0.01, 3)
        I = np.array([[par[4], par[7], par[8]], [par[7], par[5], par[9]], [par[8], par[9], par[6]]])
        s = np.zeros(37)
        s[2] = 10.0
        q = rng.normal(size=4)
        s[3:7] = q / np.linalg.norm(q)
        R = _quat_matrix(s[3:7])
        wb, vb = rng.normal(size=3), rng.normal(size=3)
        s[7:10], s[10:13] = R @ vb, R @ wb
        a0, _, _, _ = o.forward(s[:13], np.zeros(12), np.zeros(12), np.zeros(12), par)
        assert np.allclose(a0[0:3], np.linalg.solve(I, -np.cross(wb, I @ wb)), atol=1e-12)
        assert np.allclose(a0[3:6], -np.cross(wb, vb), atol=1e-12)


def _pendulum_links(lengths, masses):
    """Fixed base, planar chain of point masses about the y axis, hanging along -z."""
    specs = [{"parent": -1, "jtype": 2, "mass": 1.0}]
    for i, (l, m) in enumerate(zip(lengths, masses)):
        specs.append({"parent": i, "jtype": 1, "dof": i, "axis": [0, 1, 0],
                      "xyz": [0, 0, -lengths[i - 1]] if i > 0 else [0, 0, 0],
                      "mass": m, "com": [0, 0, -l], "inertia": [0] * 6})
    return spio.make_links(specs)


def test_simple_pendulum_closed_form(sim):
    """theta_ddot = -(g / l) sin(theta) + tau / (m l^2), theta about +y from hanging down."""
    l, m, g = 0.7, 1.3, 9.81
    o = spio.Oracle(None, sim, links=_pendulum_links([l], [m]))
    par = np.zeros(25)
    par[0] = 1.0
    rng = np.random.default_rng(5)
    for th in rng.uniform(-2.5, 2.5, 6):
        q = np.zeros(12)
        q[0] = th
        qd = np.zeros(12)
        qd[0] = rng.normal()
        tau = np.zeros(12)
        tau[0] = rng.normal()
        root = np.r_[0, 0, 0, 0, 0, 0, 1, np.zeros(6)]
        _, qdd, _, _ = o.forward(root, q, qd, tau, par)
        # rotating +th about y moves the mass towards -x: gravity torque = -m g l sin(th)
        assert abs(qdd[0] - (-(g / l) * np.sin(th) + tau[0] / (m * l * l))) < 1e-12


def test_double_pendulum_closed_form(sim):
    """Planar double pendulum with point masses (Lagrange): pins the velocity-product terms."""
    l1, l2, m1, m2, g = 0.5, 0.4, 1.1, 0.7, 9.81
    o = spio.Oracle(None, sim, links=_pendulum_links([l1, l2], [m1, m2]))
    par = np.zeros(25)
    par[0] = 1.0
    rng = np.random.default_rng(6)
    for _ in range(8):
        t1, t2 = rng.uniform(-2, 2, 2)      # t2 is relative to link 1
        w1, w2 = rng.normal(0, 3, 2)
        q, qd = np.zeros(12), np.zeros(12)
        q[:2], qd[:2] = [t1, t2], [w1, w2]
        _, qdd, _, _ = o.forward(np.r_[0, 0, 0, 0, 0, 0, 1, np.zeros(6)], q, qd, np.zeros(12), par)
        # absolute angles a1 = t1, a2 = t1 + t2; masses at l sin/cos in the x-z plane:
        # x_i = -sum l_k sin(a_k), z_i = -sum l_k cos(a_k)  (rotation about +y moves -z towards -x)
        a1, a2 = t1, t1 + t2
        d1, d2 = w1, w1 + w2
        M = np.array([[(m1 + m2) * l1 * l1, m2 * l1 * l2 * np.cos(a1 - a2)],
                      [m2 * l1 * l2 * np.cos(a1 - a2), m2 * l2 * l2]])
        rhs = np.array([-m2 * l1 * l2 * np.sin(a1 - a2) * d2 * d2 - (m1 + m2) * g * l1 * np.sin(a1),
                        m2 * l1 * l2 * np.sin(a1 - a2) * d1 * d1 - m2 * g * l2 * np.sin(a2)])
        aa = np.linalg.solve(M, rhs)         # absolute angular accelerations
        assert np.allclose([qdd[0], qdd[1]], [aa[0], aa[1] - aa[0]], atol=1e-10)


def test_static_stance_supports_weight(model, sim, theta0):
    """After settling 5 s with zero actions, the four normal forces carry m g (within 0.1 %).
    (Settling is slow: below the friction cone the tangential law is viscous, so loaded feet
    creep until the splay balances.)"""
    from spi_amd.data import standing_state
    o = spio.Oracle(model, sim)
    traj, rc = o.simulate(theta0, standing_state(), np.zeros((250, 12), np.float32))
    assert rc == 0
    s = traj[-1].astype(np.float64)
    assert np.abs(s[7:13]).max() < 5e-3 and np.abs(s[25:37]).max() < 2e-2
    tau = o.control_torque(np.zeros(12), s[13:25], s[25:37], theta0)
    _, _, ff, _ = o.forward(s[:13], s[13:25], s[25:37], tau, theta0)
    assert abs(ff[:, 2].sum() - 15.019 * 9.81) < 0.001 * 15.019 * 9.81, ff


def test_contact_force_known_answer(model, sim, theta0):
    """A foot sphere penetrating by d at rest gets exactly (0, 0, k d); others none."""
    o = spio.Oracle(model, sim)
    s = np.zeros(37)
    s[6] = 1.0
    s[13:25] = sim.q0
    _, feet = o.fk(s[:13], s[13:25])
    d = 0.004
    s[2] = -feet[:, 2].min() + 0.022 - d          # lowest foot penetrates by d
    a0, qdd, ff, _ = o.forward(s[:13], s[13:25], np.zeros(12), np.zeros(12), theta0)
    low = np.argmin(feet[:, 2])
    pen = 0.022 - (feet[:, 2] + s[2])
    for f in range(4):
        expect = [0, 0, sim.k_contact * pen[f]] if pen[f] > 0 else [0, 0, 0]
        assert np.allclose(ff[f], expect, atol=1e-9)
    assert ff[low, 2] > 0


@pytest.mark.parametrize("vz", [-0.05, -1.0, 0.3, 3.0])
def test_contact_force_bounded_damping_known_answer(model, sim, theta0, vz):
    """Translating robot (base velocity (0, 0, vz), joints at rest): every penetrating foot gets
    Fn = k d (1 - tanh(c vz / (k d))): about k d - c vz for a slow approach or separation, never
    more than 2 k d however fast the approach, never a pull.  As d -> 0 the force goes to 0
    whatever vz, the continuity of the law at touchdown (the unsaturated law
    max(0, k d - c vz) would give c |vz| there)."""
    o = spio.Oracle(model, sim)
    s = np.zeros(37)
    s[6] = 1.0
    s[13:25] = sim.q0
    s[9] = vz
    _, feet = o.fk(s[:13], s[13:25])
    k, c = sim.k_contact, sim.c_contact
    for d in (0.004, 1e-4, 1e-7):
        s[2] = -feet[:, 2].min() + 0.022 - d
        _, _, ff, _ = o.forward(s[:13], s[13:25], np.zeros(12), np.zeros(12), theta0)
        pen = 0.022 - (feet[:, 2] + s[2])
        for f in range(4):
            if pen[f] <= 0:
                assert np.allclose(ff[f], 0.0)
                continue
            fn = k * pen[f] * (1.0 - np.tanh(c * vz / (k * pen[f])))
            assert abs(ff[f, 2] - fn) <= 1e-9 * max(1.0, fn), (d, vz, ff[f, 2], fn)
            assert 0.0 <= ff[f, 2] <= 2 * k * pen[f] + 1e-12
            if d == 0.004 and vz == -0.05:          # |c vz| << k d: the plain spring-damper
                assert abs(ff[f, 2] - (k * pen[f] - c * vz)) <= 0.02 * ff[f, 2]
        if d == 1e-7:
            assert np.abs(ff[:, 2]).max() <= 2 * k * 1e-7 + 1e-12


@pytest.mark.parametrize("mode,motor", [(0, 0), (1, 3), (1, 1), (1, 2)])
def test_control_torque_known_answers(model, sim, theta0, mode, motor):
    """Closed forms of legged_robot_base.py:542-560, go2_omni.py:436-462 and the act2tau models."""
    s2 = copy.deepcopy(sim)
    s2.control_mode, s2.motor_model = mode, motor
    o = spio.Oracle(model, s2)
    par = theta0.copy()
    par[22:25] = [11.0, 13.0, 17.0]
    q, qd = np.asarray(sim.q0, float), np.zeros(12)
    # zero action at the default pose: zero torque for every model (A tanh(0) = 0)
    assert np.allclose(o.control_torque(np.zeros(12), q, qd, par), 0.0, atol=1e-15)
    # small action: PD torque kp * 0.25 * a (hips halved in go2_omni), then the motor model
    a = np.full(12, 0.01)
    tau = o.control_torque(a, q, qd, par)
    base = 20.0 * 0.25 * 0.01 * np.array([0.5 if (mode == 1 and j % 3 == 0) else 1.0 for j in range(12)])
    grp = np.arange(12) % 3
    if motor == 0:
        expect = base
    elif motor == 1:
        expect = base * 11.0
    elif motor == 2:
        expect = base * par[22 + grp]
    else:
        A = par[22 + grp]
        expect = A * np.tanh(base / A)
    assert np.allclose(tau, expect, rtol=1e-14)
    # huge action: action clip at 20 then torque clip at the effort limit, then the model
    big = o.control_torque(np.full(12, 1e6), q, qd, par)
    eff = np.array([23.7, 23.7, 35.55] * 4)
    pd = np.minimum(20.0 * (0.25 * 20.0 * np.array([0.5 if (mode == 1 and j % 3 == 0) else 1.0
                                                      for j in range(12)])), eff)
    if motor == 0:
        assert np.allclose(big, pd)
    elif motor == 3:
        A = par[22 + grp]
        assert np.allclose(big, A * np.tanh(pd / A))


def test_tracking_cost_closed_forms(model, sim, theta0):
    """Zero for a target equal to the oracle's own rollout; a hand-built offset gives its
    weighted square (tracking) and its L2 norm (reference mode, scripts/eval.py:290-292)."""
    from spi_amd.data import standing_state
    o = spio.Oracle(model, sim)
    init = standing_state()
    acts = np.zeros((3, 12), np.float32)
    traj, _ = o.simulate(theta0, init, acts)
    w = np.array([10, 5, 1, 1, 1, 0.05])
    c = o.rollout_cost(theta0[None], init[None], acts[None], traj[None], 0, w)
    assert c[0, 0] < 1e-10
    tg = traj.copy()
    tg[:, 0] += 0.01          # base x off by 1 cm at every step
    tg[-1, 13] += 0.02        # one joint off by 0.02 rad at the end
    ct = o.rollout_cost(theta0[None], init[None], acts[None], tg[None], 0, w)[0, 0]
    assert abs(ct - (3 * 10 * 0.01 ** 2 + 1 * 0.02 ** 2)) < 1e-8
    cr = o.rollout_cost(theta0[None], init[None], acts[None], tg[None], 1, np.array([10, 5, 0, 0, 1, 0]))[0, 0]
    assert abs(cr - (10 * 0.01 + 1 * 0.02)) < 1e-6
