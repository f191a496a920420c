#!/usr/bin/env python3
"""Model-data generator for the Unitree G1 (29-DoF, floating base).

Reads the MJCF files of the reference (``model/g1_stripped.xml`` for the
kinematic tree, inertials, joints and actuators; ``model/g1_raw.xml`` for the
foot contact spheres) and writes two files that the build commits:

* ``g1_model.json``          -- every constant in float64, one array per name;
* ``../csrc/g1_model_data.h`` -- the same numbers as ``constexpr`` tables for
  the HIP kernels and the C oracle.

Derived quantities are computed here with our own float64 arithmetic, following
what MuJoCo's compiler does with the same inputs (the reference's own generator
``model/gen_model.py:81-141`` reads them back from a compiled ``MjModel``):

* body and inertial quaternions are normalised (``model/gen_model.py:102``);
* the COM inertia is rotated into the body frame, ``R diag(I) R^T``
  (``model/gen_model.py:105-109``);
* position-servo damping from ``dampratio``: ``kv = 2 * dampratio *
  sqrt(kp * M_jj(qpos0))`` with ``M_jj`` the joint-space inertia diagonal at
  the reference configuration, armature included (``model/gen_model.py:132``
  reads MuJoCo's result).  ``M_jj`` is evaluated by a composite-rigid-body pass.

One value has no source in the MJCF files: the foot friction coefficient.  The
reference's generated model uses mu = 1.0 (``src/g1_model.h:351``) while
``model/g1_raw.xml:17`` says 0.6; the generated model is what the reference
simulates, so mu = 1.0 is used and stated here.

``--check`` compares every array with the reference's generated header
(``src/g1_model.h``) parsed as numbers, and fails on any relative difference
above 1e-12.  Run from the repository root:

    python g1sim_amd/model/gen_model.py --ref /root/reference [--check]
"""
import argparse
import json
import os
import re
import sys
import xml.etree.ElementTree as ET

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
JSON_OUT = os.path.join(HERE, "g1_model.json")
HDR_OUT = os.path.join(HERE, "..", "csrc", "g1_model_data.h")

FOOT_MU = 1.0          # src/g1_model.h:351 (see module docstring)
GRAVITY_Z = -9.81      # src/g1_types.h:46


def _vec(s, n, default):
    if s is None:
        return np.array(default, dtype=np.float64)
    v = np.array([float(x) for x in s.split()], dtype=np.float64)
    assert v.shape == (n,), s
    return v


def quat_to_mat(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def parse_stripped(path):
    root = ET.parse(path).getroot()
    # class "g1" defaults: joint armature/frictionloss, position kp/dampratio
    arm, kp, dampratio = 0.0, None, None
    for d in root.iter("default"):
        if d.get("class") == "g1":
            j = d.find("joint")
            if j is not None and j.get("armature") is not None:
                arm = float(j.get("armature"))
            p = d.find("position")
            if p is not None:
                kp = float(p.get("kp"))
                dampratio = float(p.get("dampratio"))
    assert kp is not None and dampratio is not None

    bodies = []   # dicts in MuJoCo (depth-first pre-) order

    def walk(el, parent):
        idx = len(bodies)
        inert = el.find("inertial")
        b = {
            "name": el.get("name"),
            "parent": parent,
            "pos": _vec(el.get("pos"), 3, [0, 0, 0]),
            "quat": _vec(el.get("quat"), 4, [1, 0, 0, 0]),
            "ipos": _vec(inert.get("pos"), 3, [0, 0, 0]),
            "iquat": _vec(inert.get("quat"), 4, [1, 0, 0, 0]),
            "mass": float(inert.get("mass")),
            "diag": _vec(inert.get("diaginertia"), 3, None),
            "joint": None,
        }
        j = el.find("joint")
        if j is not None:
            b["joint"] = {
                "name": j.get("name"),
                "axis": _vec(j.get("axis"), 3, [0, 0, 1]),
                "pos": _vec(j.get("pos"), 3, [0, 0, 0]),
                "range": _vec(j.get("range"), 2, None),
                "frc": _vec(j.get("actuatorfrcrange"), 2, None),
            }
        else:
            assert el.find("freejoint") is not None and parent == -1
        bodies.append(b)
        for c in el.findall("body"):
            walk(c, idx)

    wb = root.find("worldbody")
    tops = wb.findall("body")
    assert len(tops) == 1
    walk(tops[0], -1)

    acts = [a.get("joint") for a in root.find("actuator").findall("position")]
    key = root.find("keyframe").find("key")
    return bodies, acts, arm, kp, dampratio, _vec(key.get("qpos"), 36, None)


def parse_feet(raw_path, body_names):
    root = ET.parse(raw_path).getroot()
    radius = None
    for d in root.iter("default"):
        if d.get("class") == "foot":
            radius = float(d.find("geom").get("size"))
    feet = []
    for body in root.iter("body"):
        for g in body.findall("geom"):
            if g.get("class") == "foot":
                feet.append((body_names.index(body.get("name")),
                             _vec(g.get("pos"), 3, None)))
    assert len(feet) == 8 and radius is not None
    return feet, radius


def build(ref):
    bodies, acts, arm, kp, dampratio, key = parse_stripped(
        os.path.join(ref, "model", "g1_stripped.xml"))
    names = [b["name"] for b in bodies]
    nb = len(bodies)
    assert nb == 30
    hinges = [b for b in bodies[1:]]
    assert [b["joint"]["name"] for b in hinges] == acts, "actuator order"

    parent = [b["parent"] for b in bodies]
    tree_pos = np.array([b["pos"] for b in bodies])
    tree_quat = np.array([b["quat"] / np.linalg.norm(b["quat"]) for b in bodies])
    mass = np.array([b["mass"] for b in bodies])
    com = np.array([b["ipos"] for b in bodies])
    icom = []
    for b in bodies:
        q = b["iquat"] / np.linalg.norm(b["iquat"])
        R = quat_to_mat(q)
        icom.append(R @ np.diag(b["diag"]) @ R.T)
    icom = np.array(icom)
    axis = np.array([b["joint"]["axis"] for b in hinges])
    anchor = np.array([b["joint"]["pos"] for b in hinges])
    rng = np.array([b["joint"]["range"] for b in hinges])
    frclim = np.array([b["joint"]["frc"][1] for b in hinges])
    armature = np.full(29, arm)
    damping = np.zeros(29)
    act_kp = np.full(29, kp)

    # composite inertia of every subtree, in its own body frame, at qpos0
    # (hinges at 0, so each child pose in its parent is the tree pose).
    def spatial_inertia(m, c, Ic):
        cx = np.array([[0, -c[2], c[1]], [c[2], 0, -c[0]], [-c[1], c[0], 0]])
        I = np.zeros((6, 6))
        I[:3, :3] = Ic - m * cx @ cx
        I[:3, 3:] = m * cx
        I[3:, :3] = -m * cx
        I[3:, 3:] = m * np.eye(3)
        return I

    def force_xform(R, p):      # child force -> parent force
        px = np.array([[0, -p[2], p[1]], [p[2], 0, -p[0]], [-p[1], p[0], 0]])
        X = np.zeros((6, 6))
        X[:3, :3] = R
        X[:3, 3:] = px @ R
        X[3:, 3:] = R
        return X

    Ic = [spatial_inertia(mass[i], com[i], icom[i]) for i in range(nb)]
    for i in range(nb - 1, 0, -1):
        R = quat_to_mat(tree_quat[i])
        Xf = force_xform(R, tree_pos[i])
        Ic[parent[i]] = Ic[parent[i]] + Xf @ Ic[i] @ Xf.T
    kv = np.zeros(29)
    for h in range(29):
        S = np.concatenate([axis[h], np.cross(anchor[h], axis[h])])
        mjj = S @ Ic[h + 1] @ S + armature[h]
        kv[h] = dampratio * 2.0 * np.sqrt(kp * mjj)

    feet, radius = parse_feet(os.path.join(ref, "model", "g1_raw.xml"), names)
    model = {
        "names": names,
        "parent": parent,
        "tree_pos": tree_pos.reshape(-1).tolist(),
        "tree_quat": tree_quat.reshape(-1).tolist(),
        "mass": mass.tolist(),
        "com": com.reshape(-1).tolist(),
        "inertia_com": icom.reshape(-1).tolist(),
        "jnt_axis": axis.reshape(-1).tolist(),
        "jnt_anchor": anchor.reshape(-1).tolist(),
        "armature": armature.tolist(),
        "damping": damping.tolist(),
        "jnt_range": rng.reshape(-1).tolist(),
        "act_kp": act_kp.tolist(),
        "act_kv": kv.tolist(),
        "act_frclim": frclim.tolist(),
        "contact_body": [f[0] for f in feet],
        "contact_pos": np.array([f[1] for f in feet]).reshape(-1).tolist(),
        "contact_radius": radius,
        "contact_mu": FOOT_MU,
        "qpos_stand": key.tolist(),
        "gravity_z": GRAVITY_Z,
    }
    return model


def parse_ref_header(path):
    """Numbers of the reference's generated header, by array name."""
    txt = open(path).read()
    out = {}
    for m in re.finditer(r"G1_MODEL_CONST\s+(\w+)\s+(\w+)(\[\d+\])?\s*=\s*"
                         r"(\{[^}]*\}|[^;]*);", txt):
        body = m.group(4)
        nums = re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", body.replace("R(", "("))
        vals = [float(x) for x in nums]
        out[m.group(2)] = vals if m.group(3) else vals[0]
    return out


def check(model, ref):
    hdr = parse_ref_header(os.path.join(ref, "src", "g1_model.h"))
    worst = 0.0
    for key, val in model.items():
        if key == "names" or key == "gravity_z":
            continue
        r = np.asarray(hdr["g1_" + key], dtype=np.float64)
        v = np.asarray(val, dtype=np.float64)
        assert r.shape == v.shape, (key, r.shape, v.shape)
        err = np.max(np.abs(r - v) / np.maximum(np.abs(r), 1e-300 + np.abs(v)).clip(min=1e-300)) if r.size else 0
        # tiny absolute values (e.g. 1e-21 off-diagonals) compare absolutely
        err = min(err, np.max(np.abs(r - v)) / max(np.max(np.abs(r)), 1e-300) if r.size else 0)
        worst = max(worst, err)
        print(f"  {key:15s} n={r.size:4d} max rel diff {err:.3g}")
        assert err < 1e-12, key
    print(f"model check vs reference header: worst rel diff {worst:.3g}")


def emit_header(model, path):
    def arr(name, vals, ctype="double"):
        flat = list(np.asarray(vals, dtype=np.float64).reshape(-1))
        body = ", ".join(repr(float(v)) for v in flat)
        return f"G1M_CONST {ctype} {name}[{len(flat)}] = {{{body}}};\n"

    def iarr(name, vals):
        body = ", ".join(str(int(v)) for v in vals)
        return f"G1M_CONST int {name}[{len(vals)}] = {{{body}}};\n"

    lines = [
        "// Generated by g1sim_amd/model/gen_model.py from g1_model.json -- do not edit.\n",
        "// Unitree G1 29-DoF model constants in float64 (bodies in depth-first order,\n",
        "// 0 = pelvis with the free joint; hinge h drives body h+1).\n",
        "#pragma once\n",
        "#ifndef G1M_CONST\n#define G1M_CONST static const\n#endif\n",
        "#define G1M_NB 30\n#define G1M_NQ 36\n#define G1M_NV 35\n#define G1M_NU 29\n#define G1M_NC 8\n",
        "// " + ", ".join(f"{i}:{n}" for i, n in enumerate(model["names"])) + "\n",
        iarr("g1m_parent", model["parent"]),
    ]
    for key in ["tree_pos", "tree_quat", "mass", "com", "inertia_com", "jnt_axis",
                "jnt_anchor", "armature", "damping", "jnt_range", "act_kp", "act_kv",
                "act_frclim"]:
        lines.append(arr("g1m_" + key, model[key]))
    lines.append(iarr("g1m_contact_body", model["contact_body"]))
    lines.append(arr("g1m_contact_pos", model["contact_pos"]))
    lines.append(f"#define G1M_CONTACT_RADIUS {model['contact_radius']!r}\n")
    lines.append(f"#define G1M_CONTACT_MU {model['contact_mu']!r}\n")
    lines.append(f"#define G1M_GRAVITY_Z {model['gravity_z']!r}\n")
    lines.append(arr("g1m_qpos_stand", model["qpos_stand"]))
    with open(path, "w") as f:
        f.writelines(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--check", action="store_true")
    args = ap.parse_args()
    model = build(args.ref)
    if args.check:
        check(model, args.ref)
    with open(JSON_OUT, "w") as f:
        json.dump(model, f, indent=1)
    emit_header(model, HDR_OUT)
    print(f"wrote {JSON_OUT} and {HDR_OUT}")


if __name__ == "__main__":
    sys.exit(main())
