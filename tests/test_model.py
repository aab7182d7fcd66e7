"""Model layer: URDF parsing, fixed-joint folding and the generated kernel header (CPU)."""
import os

import numpy as np
import pytest

import spio
from conftest import REFERENCE

URDF = os.path.join(REFERENCE, "spigym/data/robots/go2/urdf/go2.urdf")


@pytest.mark.skipif(not os.path.exists(URDF), reason="reference checkout not present")
def test_urdf_parse_matches_committed_model(model):
    from spi_amd.model import fold_model, parse_urdf
    raw = parse_urdf(URDF)
    assert raw == model["raw"]
    folded = fold_model(raw)
    for k in ("bodies", "dof_names", "base_link", "base_fixed", "contacts", "isaac_bodies"):
        assert folded[k] == model[k], k


def test_folded_masses(model):
    """go2.urdf: base 6.921 (+ two 0.001 kg head links), hip 0.678, thigh 1.152,
    calf 0.154 + foot 0.04; total 15.019 kg."""
    b = model["bodies"]
    assert len(b) == 13
    assert abs(b[0]["mass"] - 6.923) < 1e-12
    for leg in range(4):
        assert [round(x["mass"], 12) for x in b[1 + 3 * leg:4 + 3 * leg]] == [0.678, 1.152, 0.194]
    assert abs(sum(x["mass"] for x in b) - 15.019) < 1e-12
    assert abs(model["total_mass"] - 15.019) < 1e-12
    # calf + foot CoM: (0.154 c_calf + 0.04 (0, 0, -0.213)) / 0.194
    c = (0.154 * np.array([0.00548, -0.000975, -0.115]) + 0.04 * np.array([0, 0, -0.213])) / 0.194
    assert np.allclose(b[3]["com"], c, atol=1e-15)


def test_dof_order_is_isaacgym_dfs(model):
    from spi_amd.model import DOF_NAMES
    assert model["dof_names"] == DOF_NAMES


def _folded_links(model):
    specs = []
    for i, b in enumerate(model["bodies"]):
        if b["parent"] < 0:
            specs.append({"parent": -1, "jtype": 0, "mass": b["mass"], "com": b["com"], "inertia": b["inertia"]})
        else:
            specs.append({"parent": b["parent"], "jtype": 1, "dof": i - 1, "xyz": b["origin"], "axis": b["axis"],
                          "mass": b["mass"], "com": b["com"], "inertia": b["inertia"]})
    return spio.make_links(specs)


def test_folding_preserves_the_mass_matrix(model, sim, theta0):
    """The folded 13-body tree and the unfolded URDF tree have the same joint-space inertia."""
    unf = spio.Oracle(model, sim)
    fol = spio.Oracle(None, sim, links=_folded_links(model))
    base = model["bodies"][0]
    par_f = theta0.copy()
    par_f[0] = base["mass"]
    par_f[1:4] = base["com"]
    par_f[4:10] = base["inertia"]
    rng = np.random.default_rng(0)
    for _ in range(5):
        root = np.r_[rng.normal(size=3), [0, 0, 0, 1], np.zeros(6)]
        q = np.asarray(sim.q0) + rng.normal(0, 0.5, 12)
        Hu = unf.crba(root, q, theta0)
        Hf = fol.crba(root, q, par_f)
        assert np.allclose(Hu, Hf, atol=1e-12), np.abs(Hu - Hf).max()


def test_generated_header_is_current(model):
    from spi_amd.build import CSRC
    from spi_amd.model import generate_header
    with open(os.path.join(CSRC, "spi_go2_model.h")) as f:
        assert f.read() == generate_header(model)


def test_foot_spheres(model):
    """go2.urdf:360-365: sphere r = 0.022 at (-0.002, 0, 0) of the foot frame, which sits at
    (0, 0, -0.213) of the calf."""
    for c in model["contacts"]:
        assert c["radius"] == 0.022
        assert np.allclose(c["pos"], [-0.002, 0.0, -0.213])
