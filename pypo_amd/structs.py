"""ctypes mirrors of the C ABI structs in include/pypo_hip.h.

Field order and types are those of the reference's src/PyPO/Structs.py:9-141
(double variants), which mirror src/include/Structs.h:29-161, so a struct
filled here can be handed to either library.
"""

import ctypes

_dp = ctypes.POINTER(ctypes.c_double)


def _fields(names):
    return [(n, _dp) for n in names]


_C2 = ["r1x", "r1y", "r1z", "i1x", "i1y", "i1z", "r2x", "r2y", "r2z", "i2x", "i2y", "i2z"]
_C4 = _C2 + ["r3x", "r3y", "r3z", "i3x", "i3y", "i3z", "r4x", "r4y", "r4z", "i4x", "i4y", "i4z"]
_C2R = _C2 + ["r3x", "r3y", "r3z"]


class arrC1(ctypes.Structure):
    """Complex scalar field, SoA (Structs.py:12-18)."""
    _fields_ = _fields(["x", "y"])


class c2Bundle(ctypes.Structure):
    """Two complex 3-vectors per point (Structs.py:29-47)."""
    _fields_ = _fields(_C2)


class c4Bundle(ctypes.Structure):
    """Four complex 3-vectors per point (Structs.py:49-79)."""
    _fields_ = _fields(_C4)


class c2rBundle(ctypes.Structure):
    """Two complex and one real 3-vector per point (Structs.py:81-101)."""
    _fields_ = _fields(_C2R)


class reflparams(ctypes.Structure):
    """Surface parameters, passed by value (Structs.py:97-114)."""
    _fields_ = [("coeffs", _dp),
                ("lxu", _dp),
                ("lyv", _dp),
                ("n_cells", ctypes.POINTER(ctypes.c_int)),
                ("flip", ctypes.c_bool),
                ("gmode", ctypes.c_int),
                ("gcenter", _dp),
                ("ecc_uv", ctypes.c_double),
                ("rot_uv", ctypes.c_double),
                ("type", ctypes.c_int),
                ("transf", _dp),
                ("rms", ctypes.c_double),
                ("rms_seed", ctypes.c_uint)]


class reflcontainer(ctypes.Structure):
    """Realised grid (Structs.py:116-128)."""
    _fields_ = [("size", ctypes.c_int)] + _fields(["x", "y", "z", "nx", "ny", "nz", "area"])


class cframe(ctypes.Structure):
    """Ray frame (Structs.py:130-141)."""
    _fields_ = [("size", ctypes.c_int)] + _fields(["x", "y", "z", "dx", "dy", "dz"])


BUNDLE_FIELDS = {"c2Bundle": _C2, "c4Bundle": _C4, "c2rBundle": _C2R, "arrC1": ["x", "y"]}


class RTDict(ctypes.Structure):
    """Tubular ray-frame parameters, angles in radians (Structs.h:164-171)."""
    _fields_ = [("nRays", ctypes.c_int), ("nRing", ctypes.c_int), ("angx0", ctypes.c_double),
                ("angy0", ctypes.c_double), ("x0", ctypes.c_double), ("y0", ctypes.c_double)]


class GRTDict(ctypes.Structure):
    """Gaussian ray-frame parameters, angles in radians (Structs.h:174-181)."""
    _fields_ = [("nRays", ctypes.c_int), ("angx0", ctypes.c_double), ("angy0", ctypes.c_double),
                ("x0", ctypes.c_double), ("y0", ctypes.c_double), ("seed", ctypes.c_int)]


class GPODict(ctypes.Structure):
    """Astigmatic Gaussian beam parameters (Structs.h:184-192)."""
    _fields_ = [("lam", ctypes.c_double), ("w0x", ctypes.c_double), ("w0y", ctypes.c_double),
                ("n", ctypes.c_double), ("E0", ctypes.c_double), ("dxyz", ctypes.c_double),
                ("pol", _dp)]


class vecGPODict(ctypes.Structure):
    """Complex-source-point vector Gaussian beam parameters (Structs.h:195-202)."""
    _fields_ = [("lam", ctypes.c_double), ("w0", ctypes.c_double), ("z", ctypes.c_double),
                ("n", ctypes.c_double), ("power", ctypes.c_double), ("mode", ctypes.c_int)]


class ScalarGPODict(ctypes.Structure):
    """Scalar Gaussian beam parameters (Structs.h:206-213)."""
    _fields_ = [("lam", ctypes.c_double), ("w0x", ctypes.c_double), ("w0y", ctypes.c_double),
                ("n", ctypes.c_double), ("E0", ctypes.c_double), ("dxyz", ctypes.c_double)]


class aperparams(ctypes.Structure):
    """Aperture of Efficiencies._generateMask (Efficiencies.py:11-41); shape
    0 ellipse, 1 rectangle, 2 none (every point)."""
    _fields_ = [("shape", ctypes.c_int), ("center", ctypes.c_double * 2), ("outer", ctypes.c_double * 2),
                ("outer_x", ctypes.c_double * 2), ("outer_y", ctypes.c_double * 2)]


# ---- FP32 twins (Structs.py:206-337), for the single-precision entry points ----
_fp = ctypes.POINTER(ctypes.c_float)


def _ffields(names):
    return [(n, _fp) for n in names]


class arrC1f(ctypes.Structure):
    _fields_ = _ffields(["x", "y"])


class c2Bundlef(ctypes.Structure):
    _fields_ = _ffields(_C2)


class c4Bundlef(ctypes.Structure):
    _fields_ = _ffields(_C4)


class c2rBundlef(ctypes.Structure):
    _fields_ = _ffields(_C2R)


class reflparamsf(ctypes.Structure):
    _fields_ = [("coeffs", _fp), ("lxu", _fp), ("lyv", _fp), ("n_cells", ctypes.POINTER(ctypes.c_int)),
                ("flip", ctypes.c_bool), ("gmode", ctypes.c_int), ("gcenter", _fp), ("ecc_uv", ctypes.c_float),
                ("rot_uv", ctypes.c_float), ("type", ctypes.c_int), ("transf", _fp), ("rms", ctypes.c_float),
                ("rms_seed", ctypes.c_uint)]


class reflcontainerf(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int)] + _ffields(["x", "y", "z", "nx", "ny", "nz", "area"])


class cframef(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int)] + _ffields(["x", "y", "z", "dx", "dy", "dz"])
