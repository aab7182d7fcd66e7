"""The numpy restatement of the command generators (the oracle for spi_command_sequences)
against closed forms: constant windows repeat, zero polynomial coefficients give the range
midpoint, tanh(atanh(x)) maps back, a Bezier curve with equal control points is constant,
passes through its end points and, with two points, is the straight line."""
import numpy as np

RANGES = [[-1.0, 1.0], [-1.0, 1.0], [-1.0, 1.0], [-0.25, 0.15], [2.0, 4.0], [0.0, 1.0], [0.0, 1.0], [0.0, 1.0],
          [0.5, 0.5], [0.03, 0.035], [-0.4, 0.4], [-0.0, 0.0], [0.1, 0.45], [0.35, 0.45]]
DEFAULT = [0.5, 0.5, 0.5, 0.0, 3.0, 0.0, 0.5, 0.5, 0.5, 0.03, 0.0, 0.0, 0.3, 0.4]
IDX = [0, 2, 5]


def test_constant_windows_and_defaults():
    import spio_commands as oc
    names = oc.param_names("constant", IDX, 5)
    assert names[:2] == ["dim_0_update_0", "dim_0_update_1"] and len(names) == 15
    p = np.arange(15, dtype=np.float64) / 20
    out = oc.commands_from_params(p, "constant", RANGES, DEFAULT, IDX, 5, 250)
    assert out.shape == (1250, 14) and out.dtype == np.float32
    for d, a in enumerate(IDX):
        for u in range(5):
            assert np.all(out[u * 250:(u + 1) * 250, a] == np.float32(p[d * 5 + u]))
    for k in set(range(14)) - set(IDX):
        assert np.all(out[:, k] == np.float32(DEFAULT[k]))


def test_polynomial_closed_forms():
    import spio_commands as oc
    nc = 4
    p = np.zeros(5 * 3 * nc)
    out = oc.commands_from_params(p, "polynomial", RANGES, DEFAULT, IDX, 5, 50)
    mid = [(RANGES[a][0] + RANGES[a][1]) / 2 for a in IDX]
    for d, a in enumerate(IDX):
        assert np.allclose(out[:, a], mid[d], atol=1e-7)
    # constant term atanh(x): the command is vmin + (x + 1)/2 (vmax - vmin)
    x = 0.3
    p[0::nc] = np.arctanh(x)
    out = oc.commands_from_params(p, "polynomial", RANGES, DEFAULT, IDX, 5, 50)
    for a in IDX:
        lo, hi = RANGES[a]
        assert np.allclose(out[:, a], lo + (x + 1) / 2 * (hi - lo), atol=1e-6)
    # linear coefficient: at t = 1 the curve is tanh(c1)
    p[:] = 0
    p[1] = 0.7                                   # update 0, dim 0, coeff 1
    out = oc.commands_from_params(p, "polynomial", RANGES, DEFAULT, IDX, 5, 50)
    assert abs(out[49, 0] - (-1 + (np.tanh(0.7) + 1))) < 1e-6 and abs(out[0, 0]) < 1e-7


def test_bezier_closed_forms():
    import spio_commands as oc
    nb = 4
    rng = np.random.default_rng(0)
    p = rng.uniform(-0.5, 0.5, size=5 * 3 * nb)
    out = oc.commands_from_params(p, "bezier", RANGES, DEFAULT, IDX, 5, 60, n_bezier=nb)
    for u in range(5):
        for d, a in enumerate(IDX):
            cps = p[(u * 3 + d) * nb:(u * 3 + d + 1) * nb]
            lo, hi = RANGES[a]
            assert abs(out[u * 60, a] - np.clip(cps[0], lo, hi)) < 1e-6         # starts at the first point
            assert abs(out[u * 60 + 59, a] - np.clip(cps[-1], lo, hi)) < 1e-6   # ends at the last
    q = np.repeat(0.2, 5 * 3 * nb)
    out = oc.commands_from_params(q, "bezier", RANGES, DEFAULT, IDX, 5, 60, n_bezier=nb)
    assert np.allclose(out[:, 0], 0.2, atol=1e-6)
    # two control points: the straight line between them
    two = np.array([-0.4, 0.6] * 15)
    out = oc.commands_from_params(two, "bezier", RANGES, DEFAULT, IDX, 5, 11, n_bezier=2)
    t = np.linspace(0, 1, 11)
    assert np.allclose(out[:11, 0], -0.4 + t, atol=1e-6)
