"""Surface dictionaries (reflDicts) for the PO and ray-trace entry points.

Turns user-level surface descriptions into the internal reflDict that the
C ABI's reflparams is filled from, following the reference's System.py:
addParabola 244-291 (focus mode: a = b = 2*sqrt(f), c = -1, translate to the
vertex, rotate +z onto the axis), addHyperbola 293-353 (a3 = b3 = sqrt(c^2 -
a^2), c3 = a = c/ecc, centred between the foci), addEllipse 355-418,
addPlane 420-442, findRotation 2714-2746 (Rodrigues) and
MatTransform.MatTranslate 64-84.  Lengths in mm, angles in degrees.
"""

import copy

import numpy as np

GMODES = {"xy": 0, "uv": 1, "AoE": 2, 0: 0, 1: 1, 2: 2}

_DEFAULTS = {
    "flip": False,
    "gcenter": np.zeros(2),
    "ecc_uv": 0.0,
    "rot_uv": 0.0,
    "rms": 0.0,
    "rms_seed": 0,
}


def find_rotation(v, u):
    """4x4 matrix rotating v onto u (System.findRotation, System.py:2714-2746)."""
    v = np.asarray(v, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    out = np.eye(4)
    if np.array_equal(v, u):
        return out
    lv, lu = np.linalg.norm(v), np.linalg.norm(u)
    if lv == 0 or lu == 0:
        raise ValueError("zero-length vector in find_rotation")
    w = np.cross(v / lv, u / lu)
    lw = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    out[:3, :3] = np.eye(3) + K + K @ K * (1 - np.dot(v, u)) / lw ** 2
    return out


def mat_translate(trans, mat=None):
    """Left-multiply a translation (MatTransform.MatTranslate 64-84)."""
    mat = np.eye(4) if mat is None else mat
    T = np.eye(4)
    T[:3, 3] = np.asarray(trans, dtype=np.float64)
    return T @ mat


def _base(d, type_):
    r = copy.deepcopy(d)
    for k, v in _DEFAULTS.items():
        r.setdefault(k, copy.deepcopy(v))
    r["type"] = type_
    r["gmode"] = GMODES[r["gmode"]]
    r["gridsize"] = np.asarray(r["gridsize"], dtype=np.int64)
    r.setdefault("transf", np.eye(4))
    return r


def paraboloid(d):
    """reflDict for a paraboloid z = x^2/a^2 + y^2/b^2 (type 0)."""
    r = _base(d, 0)
    if r.get("pmode", "manual") == "focus":
        ve = np.asarray(r["vertex"], dtype=np.float64)
        f1 = np.asarray(r["focus_1"], dtype=np.float64)
        diff = f1 - ve
        a = 2 * np.sqrt(np.sqrt(np.dot(diff, diff)))
        R = find_rotation(np.array([0.0, 0.0, 1.0]), diff / np.linalg.norm(diff))
        r["transf"] = mat_translate(ve, R)
        r["coeffs"] = np.array([a, a, -1.0])
    else:
        c = np.asarray(r["coeffs"], dtype=np.float64)
        r["coeffs"] = np.array([c[0], c[1], -1.0])
    return r


def hyperboloid(d):
    """reflDict for a hyperboloid z = c*sqrt(x^2/a^2 + y^2/b^2 + 1) (type 1)."""
    r = _base(d, 1)
    if r.get("pmode", "manual") == "focus":
        f1 = np.asarray(r["focus_1"], dtype=np.float64)
        f2 = np.asarray(r["focus_2"], dtype=np.float64)
        diff = f1 - f2
        c = np.sqrt(np.dot(diff, diff)) / 2
        a = c / r["ecc"]
        b = np.sqrt(c ** 2 - a ** 2)
        R = find_rotation(np.array([0.0, 0.0, 1.0]), diff / np.linalg.norm(diff))
        r["transf"] = mat_translate((f1 + f2) / 2, R)
        r["coeffs"] = np.array([b, b, a])
    else:
        r["coeffs"] = np.asarray(r["coeffs"], dtype=np.float64)
    return r


def ellipsoid(d):
    """reflDict for an ellipsoid z = c*sqrt(1 - x^2/a^2 - y^2/b^2) (type 2)."""
    r = _base(d, 2)
    if r.get("pmode", "manual") == "focus":
        f1 = np.asarray(r["focus_1"], dtype=np.float64)
        f2 = np.asarray(r["focus_2"], dtype=np.float64)
        diff = f1 - f2
        fn = diff / np.linalg.norm(diff)
        ecc = r["ecc"]
        axis = np.array([1.0, 0, 0]) if r.get("orient", "z") == "x" else np.array([0, 0, 1.0])
        R = find_rotation(axis, fn)
        a = np.sqrt(np.dot(diff, diff)) / (2 * ecc)
        b = a * np.sqrt(1 - ecc ** 2)
        r["transf"] = mat_translate((f1 + f2) / 2, R)
        r["coeffs"] = np.array([a, b, b]) if r.get("orient", "z") == "x" else np.array([b, b, a])
    else:
        r["coeffs"] = np.asarray(r["coeffs"], dtype=np.float64)
    return r


def plane(d):
    """reflDict for a plane z = 0 (type 3); gmode xy, uv or AoE."""
    r = _base(d, 3)
    r["coeffs"] = np.zeros(3)
    return r


def gridsize(refl):
    g = np.asarray(refl["gridsize"]).ravel()
    return int(g[0]), int(g[1])
