"""Go2 rigid-body model: URDF parsing, fixed-joint folding, model file and C header.

The reference loads ``spigym/data/robots/go2/urdf/go2.urdf`` through IsaacGym
(``spigym/simulator/isaacgym/isaacgym.py:170-224``) with
``collapse_fixed_joints: True`` (``spigym/config/robot/go2/go2.yaml:112``).  The
feet and the two head links carry ``dont_collapse="true"`` (``go2.urdf:72,110,369``)
so IsaacGym keeps them as separate rigid bodies, rigidly attached.  For the
dynamics a rigidly attached body is equivalent to folding its inertia into its
parent, so the engine integrates 13 bodies (trunk + 4 x hip/thigh/calf) and
reports the 19 IsaacGym body names (``go2.yaml:44``) through forward kinematics.

The base link's own inertia (mass, CoM, inertia tensor) is the identified
parameter (``isaacgym_active_sysid.py:61-94`` sets exactly these on body
"base"); everything rigidly attached to the base (the head links) stays a
constant that is added on top, as IsaacGym keeps them as separate bodies whose
masses the setters never touch.
"""
from __future__ import annotations

import json
import math
import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL_PATH = os.path.join(_HERE, "data", "go2_model.json")

# DFS order used by IsaacGym for the Go2 (go2.yaml:29).
DOF_NAMES = [
    "FL_hip_joint", "FL_thigh_joint", "FL_calf_joint",
    "FR_hip_joint", "FR_thigh_joint", "FR_calf_joint",
    "RL_hip_joint", "RL_thigh_joint", "RL_calf_joint",
    "RR_hip_joint", "RR_thigh_joint", "RR_calf_joint",
]
# IsaacGym rigid-body order with dont_collapse links kept (go2.yaml:44).
ISAAC_BODY_NAMES = [
    "base", "FL_hip", "FL_thigh", "FL_calf", "FL_foot",
    "FR_hip", "FR_thigh", "FR_calf", "FR_foot",
    "Head_upper", "Head_lower",
    "RL_hip", "RL_thigh", "RL_calf", "RL_foot",
    "RR_hip", "RR_thigh", "RR_calf", "RR_foot",
]
LEGS = ["FL", "FR", "RL", "RR"]


def rpy_to_matrix(rpy) -> np.ndarray:
    """URDF fixed-axis roll-pitch-yaw: R = Rz(yaw) @ Ry(pitch) @ Rx(roll)."""
    r, p, y = (float(v) for v in rpy)
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return rz @ ry @ rx


def inertia6_to_matrix(i6) -> np.ndarray:
    """(ixx, iyy, izz, ixy, ixz, iyz) -> symmetric 3x3."""
    ixx, iyy, izz, ixy, ixz, iyz = (float(v) for v in i6)
    return np.array([[ixx, ixy, ixz], [ixy, iyy, iyz], [ixz, iyz, izz]])


def matrix_to_inertia6(m) -> List[float]:
    return [float(m[0, 0]), float(m[1, 1]), float(m[2, 2]), float(m[0, 1]), float(m[0, 2]), float(m[1, 2])]


def _vec(text: Optional[str], n=3) -> List[float]:
    if text is None:
        return [0.0] * n
    return [float(v) for v in text.split()]


def parse_urdf(path: str) -> Dict:
    """Parse links (inertial, collision spheres) and joints of a URDF file."""
    root = ET.parse(path).getroot()
    links = {}
    for link in root.findall("link"):
        name = link.get("name")
        entry = {"mass": 0.0, "com": [0.0, 0.0, 0.0], "com_rpy": [0.0, 0.0, 0.0],
                 "inertia": [0.0] * 6, "spheres": []}
        inert = link.find("inertial")
        if inert is not None:
            org = inert.find("origin")
            entry["com"] = _vec(org.get("xyz") if org is not None else None)
            entry["com_rpy"] = _vec(org.get("rpy") if org is not None else None)
            entry["mass"] = float(inert.find("mass").get("value"))
            ia = inert.find("inertia").attrib
            entry["inertia"] = [float(ia[k]) for k in ("ixx", "iyy", "izz", "ixy", "ixz", "iyz")]
        for col in link.findall("collision"):
            geom = col.find("geometry")
            sph = geom.find("sphere") if geom is not None else None
            if sph is not None:
                org = col.find("origin")
                entry["spheres"].append({
                    "xyz": _vec(org.get("xyz") if org is not None else None),
                    "radius": float(sph.get("radius")),
                })
        links[name] = entry
    joints = []
    for j in root.findall("joint"):
        org = j.find("origin")
        axis = j.find("axis")
        lim = j.find("limit")
        joints.append({
            "name": j.get("name"),
            "type": j.get("type"),
            "parent": j.find("parent").get("link"),
            "child": j.find("child").get("link"),
            "xyz": _vec(org.get("xyz") if org is not None else None),
            "rpy": _vec(org.get("rpy") if org is not None else None),
            "axis": _vec(axis.get("xyz")) if axis is not None else [0.0, 0.0, 0.0],
            "limit": ({k: float(v) for k, v in lim.attrib.items()} if lim is not None else None),
        })
    return {"links": links, "joints": joints}


@dataclass
class RigidInertia:
    """Mass, CoM (body frame) and inertia about the CoM (body axes)."""
    mass: float
    com: np.ndarray
    inertia: np.ndarray  # 3x3 about com

    @staticmethod
    def zero():
        return RigidInertia(0.0, np.zeros(3), np.zeros((3, 3)))

    def transformed(self, rot: np.ndarray, trans: np.ndarray) -> "RigidInertia":
        """Express in a parent frame where this frame sits at ``trans`` with orientation ``rot``."""
        return RigidInertia(self.mass, rot @ self.com + trans, rot @ self.inertia @ rot.T)

    def __add__(self, other: "RigidInertia") -> "RigidInertia":
        m = self.mass + other.mass
        if m <= 0.0:
            return RigidInertia.zero()
        c = (self.mass * self.com + other.mass * other.com) / m
        total = np.zeros((3, 3))
        for part in (self, other):
            d = part.com - c
            total += part.inertia + part.mass * (np.dot(d, d) * np.eye(3) - np.outer(d, d))
        return RigidInertia(m, c, total)


def link_inertia(link: Dict) -> RigidInertia:
    rot = rpy_to_matrix(link["com_rpy"])
    return RigidInertia(link["mass"], np.array(link["com"], dtype=float),
                        rot @ inertia6_to_matrix(link["inertia"]) @ rot.T)


def fold_model(raw: Dict) -> Dict:
    """Fold every fixed joint into its nearest movable ancestor.

    Returns the folded 13-body model as plain JSON-able data.  Body 0 is the
    floating base.  ``base_fixed`` holds what is rigidly attached to the base
    link (kept constant under identification); ``base_link`` is the base
    link's own URDF inertia (the nominal identified parameters).
    """
    links, joints = raw["links"], raw["joints"]
    by_child = {j["child"]: j for j in joints}
    children: Dict[str, List[Dict]] = {}
    for j in joints:
        children.setdefault(j["parent"], []).append(j)
    root_name = [n for n in links if n not in by_child]
    assert len(root_name) == 1, root_name
    root_name = root_name[0]

    bodies = []          # movable bodies (and base)
    body_of_link = {}    # link name -> (body index, R, p) placing the link in the body frame
    attach: Dict[int, RigidInertia] = {}

    def visit(link_name: str, body_idx: int, rot: np.ndarray, trans: np.ndarray):
        body_of_link[link_name] = (body_idx, rot, trans)
        if not (link_name == bodies[body_idx]["name"]):
            attach[body_idx] = attach.get(body_idx, RigidInertia.zero()) + \
                link_inertia(links[link_name]).transformed(rot, trans)
        for j in children.get(link_name, []):
            jr = rpy_to_matrix(j["rpy"])
            jp = np.array(j["xyz"], dtype=float)
            if j["type"] == "fixed":
                visit(j["child"], body_idx, rot @ jr, rot @ jp + trans)
            elif j["type"] in ("revolute", "continuous"):
                assert np.allclose(rot, np.eye(3)) and np.allclose(jr, np.eye(3)), \
                    "movable joints behind rotated fixed joints are not supported"
                idx = len(bodies)
                bodies.append({
                    "name": j["child"], "parent": body_idx, "joint": j["name"],
                    "axis": [float(v) for v in j["axis"]],
                    "origin": [float(v) for v in (trans + jp)],
                    "limit": j["limit"],
                })
                visit(j["child"], idx, np.eye(3), np.zeros(3))
            else:
                raise ValueError(f"unsupported joint type {j['type']}")

    bodies.append({"name": root_name, "parent": -1, "joint": None, "axis": None,
                   "origin": [0.0, 0.0, 0.0], "limit": None})
    visit(root_name, 0, np.eye(3), np.zeros(3))

    out_bodies = []
    for i, b in enumerate(bodies):
        own = link_inertia(links[b["name"]])
        total = own + attach[i] if i in attach else own
        entry = dict(b)
        entry.update({"mass": total.mass, "com": [float(v) for v in total.com],
                      "inertia": matrix_to_inertia6(total.inertia)})
        out_bodies.append(entry)

    base_link = link_inertia(links[root_name])
    base_fixed = attach.get(0, RigidInertia.zero())

    dof_names = [b["joint"] for b in out_bodies[1:]]
    contacts = []
    for leg in LEGS:
        foot = f"{leg}_foot"
        bidx, rot, trans = body_of_link[foot]
        for s in links[foot]["spheres"]:
            contacts.append({"link": foot, "body": bidx,
                             "pos": [float(v) for v in rot @ np.array(s["xyz"]) + trans],
                             "radius": s["radius"]})
    isaac_bodies = []
    for name in ISAAC_BODY_NAMES:
        bidx, rot, trans = body_of_link[name]
        assert np.allclose(rot, np.eye(3))
        isaac_bodies.append({"name": name, "body": bidx, "offset": [float(v) for v in trans]})

    return {
        "bodies": out_bodies,
        "dof_names": dof_names,
        "base_link": {"mass": base_link.mass, "com": [float(v) for v in base_link.com],
                      "inertia": matrix_to_inertia6(base_link.inertia)},
        "base_fixed": {"mass": base_fixed.mass, "com": [float(v) for v in base_fixed.com],
                       "inertia": matrix_to_inertia6(base_fixed.inertia)},
        "contacts": contacts,
        "isaac_bodies": isaac_bodies,
        "total_mass": float(sum(l["mass"] for l in links.values())),
    }


def build_model_file(urdf_path: str, out_path: str = DEFAULT_MODEL_PATH) -> Dict:
    raw = parse_urdf(urdf_path)
    folded = fold_model(raw)
    model = {"source": "spigym/data/robots/go2/urdf/go2.urdf", "raw": raw, **folded}
    with open(out_path, "w") as f:
        json.dump(model, f, indent=1, sort_keys=True)
    return model


def load_model(path: str = DEFAULT_MODEL_PATH) -> Dict:
    with open(path) as f:
        return json.load(f)


@dataclass
class Go2Model:
    """Convenience view of the folded model used by the host side."""
    data: Dict = field(default_factory=load_model)

    @property
    def num_dof(self) -> int:
        return len(self.data["dof_names"])

    @property
    def dof_names(self) -> List[str]:
        return list(self.data["dof_names"])

    @property
    def body_names(self) -> List[str]:
        return [b["name"] for b in self.data["isaac_bodies"]]

    def dof_limits(self):
        lims = [b["limit"] for b in self.data["bodies"][1:]]
        lower = np.array([l["lower"] for l in lims], dtype=np.float32)
        upper = np.array([l["upper"] for l in lims], dtype=np.float32)
        vel = np.array([l["velocity"] for l in lims], dtype=np.float32)
        effort = np.array([l["effort"] for l in lims], dtype=np.float32)
        return lower, upper, vel, effort


def _fmt(v: float) -> str:
    return repr(float(v)) + "f" if False else "%.17g" % float(v)


def generate_header(model: Dict) -> str:
    """C header with the folded model as compile-time constants for the kernels."""
    b = model["bodies"]
    lines = [
        "/* Generated from go2_model.json by spi_amd.model.generate_header; do not edit. */",
        "/* Folded Go2 tree: body 0 = trunk, bodies 1+3l..3+3l = hip, thigh, calf(+foot) of leg l. */",
        "#pragma once",
        "#define SPI_NB 13",
        "#define SPI_NDOF 12",
        "#define SPI_NLEG 4",
    ]

    def arr(name, vals, ctype="double"):
        return f"static constexpr {ctype} {name}[{len(vals)}] = {{" + ", ".join(_fmt(v) for v in vals) + "};"

    axes = []
    origins, masses, coms, inert = [], [], [], []
    for body in b[1:]:
        ax = body["axis"]
        k = int(np.argmax(np.abs(ax)))
        assert abs(abs(ax[k]) - 1.0) < 1e-12 and abs(sum(abs(a) for a in ax) - 1.0) < 1e-12, \
            "kernels assume coordinate-axis joints"
        axes.append(k if ax[k] > 0 else -(k + 1))
        origins += body["origin"]
        masses.append(body["mass"])
        coms += body["com"]
        inert += body["inertia"]
    lines.append("static constexpr int SPI_JOINT_AXIS[12] = {" + ", ".join(str(a) for a in axes) + "};")
    lines.append(arr("SPI_JOINT_ORIGIN", origins))
    lines.append(arr("SPI_LINK_MASS", masses))
    lines.append(arr("SPI_LINK_COM", coms))
    lines.append(arr("SPI_LINK_INERTIA", inert))
    bf = model["base_fixed"]
    lines.append(arr("SPI_BASE_FIXED_MASS", [bf["mass"]]))
    lines.append(arr("SPI_BASE_FIXED_COM", bf["com"]))
    lines.append(arr("SPI_BASE_FIXED_INERTIA", bf["inertia"]))
    bl = model["base_link"]
    lines.append(arr("SPI_BASE_LINK_MASS", [bl["mass"]]))
    lines.append(arr("SPI_BASE_LINK_COM", bl["com"]))
    lines.append(arr("SPI_BASE_LINK_INERTIA", bl["inertia"]))
    cpos, crad = [], []
    for c in model["contacts"]:
        cpos += c["pos"]
        crad.append(c["radius"])
    lines.append(arr("SPI_FOOT_POS", cpos))
    lines.append(arr("SPI_FOOT_RADIUS", crad))
    lim = [body["limit"] for body in b[1:]]
    lines.append(arr("SPI_DOF_LOWER", [l["lower"] for l in lim]))
    lines.append(arr("SPI_DOF_UPPER", [l["upper"] for l in lim]))
    lines.append(arr("SPI_DOF_VELOCITY", [l["velocity"] for l in lim]))
    lines.append(arr("SPI_DOF_EFFORT", [l["effort"] for l in lim]))
    iso = model["isaac_bodies"]
    lines.append("#define SPI_NISAAC %d" % len(iso))
    lines.append("static constexpr int SPI_ISAAC_BODY[%d] = {" % len(iso) + ", ".join(str(e["body"]) for e in iso) + "};")
    offs = []
    for e in iso:
        offs += e["offset"]
    lines.append(arr("SPI_ISAAC_OFFSET", offs))
    return "\n".join(lines) + "\n"


if __name__ == "__main__":  # regenerate: python -m spi_amd.model <go2.urdf>
    import sys
    m = build_model_file(sys.argv[1])
    hdr = os.path.join(_HERE, "..", "csrc", "spi_go2_model.h")
    with open(hdr, "w") as f:
        f.write(generate_header(m))
    print("wrote", DEFAULT_MODEL_PATH, hdr)
