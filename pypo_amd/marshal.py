"""numpy <-> ctypes marshalling for the C ABI.

Plays the role of the reference's src/PyPO/BindUtils.py (allfill_reflparams
277-327, allocate_reflcontainer 329-352, allfill_c2Bundle 425-454,
allocate_c2Bundle 173-199, c2BundleToObj 52-83, allfill_cframe 401-423,
frameToObj 600-625) with one difference: buffers are numpy arrays whose data
pointers are handed to the library (no per-element .tolist() copies).  Every
struct returned here keeps its arrays alive through a `_keep` attribute.
"""

import ctypes

import numpy as np

from . import structs as S
from . import types as T

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def fill_reflparams(refl):
    """reflDict -> reflparams (BindUtils.py:277-327).  gmode picks which
    limits are read: lims_x/lims_y (xy), lims_u/lims_v (uv), lims_Az/lims_El
    (AoE)."""
    p = S.reflparams()
    gmode = int(refl["gmode"])
    lims = {0: ("lims_x", "lims_y"), 1: ("lims_u", "lims_v"), 2: ("lims_Az", "lims_El")}[gmode]
    keep = {
        "coeffs": _f64(refl.get("coeffs", np.zeros(3)), 3),
        "lxu": _f64(refl[lims[0]], 2),
        "lyv": _f64(refl[lims[1]], 2),
        "gcenter": _f64(refl.get("gcenter", np.zeros(2)), 2),
        "transf": _f64(np.asarray(refl.get("transf", np.eye(4)), dtype=np.float64).ravel(), 16),
        "n_cells": np.ascontiguousarray(np.asarray(refl["gridsize"], dtype=np.int32).ravel()),
    }
    p.coeffs = _ptr(keep["coeffs"])
    p.lxu = _ptr(keep["lxu"])
    p.lyv = _ptr(keep["lyv"])
    p.gcenter = _ptr(keep["gcenter"])
    p.transf = _ptr(keep["transf"])
    p.n_cells = keep["n_cells"].ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    p.flip = bool(refl.get("flip", False))
    p.gmode = gmode
    p.ecc_uv = float(refl.get("ecc_uv", 0.0))
    p.rot_uv = float(refl.get("rot_uv", 0.0))
    p.type = int(refl["type"])
    p.rms = float(refl.get("rms", 0.0))
    p.rms_seed = int(refl.get("rms_seed", 0)) & 0xFFFFFFFF
    p._keep = keep
    return p


def alloc_container(n):
    """Empty reflcontainer of n points (BindUtils.py:329-352)."""
    c = S.reflcontainer()
    c.size = int(n)
    keep = {}
    for name in ("x", "y", "z", "nx", "ny", "nz", "area"):
        keep[name] = np.zeros(int(n), dtype=np.float64)
        setattr(c, name, _ptr(keep[name]))
    c._keep = keep
    return c


def container_to_grids(c, shape):
    k = c._keep
    return T.reflGrids(*(k[n].reshape(shape) for n in ("x", "y", "z", "nx", "ny", "nz", "area")))


def alloc_bundle(kind, n):
    """Zeroed result bundle ("c2Bundle", "c4Bundle", "c2rBundle", "arrC1")."""
    b = getattr(S, kind)()
    keep = {}
    for name in S.BUNDLE_FIELDS[kind]:
        keep[name] = np.zeros(int(n), dtype=np.float64)
        setattr(b, name, _ptr(keep[name]))
    b._keep = keep
    return b


def fill_c2bundle(obj, n):
    """currents/fields object -> c2Bundle (BindUtils.py:425-454)."""
    b = S.c2Bundle()
    keep = {}
    comps = obj.components() if hasattr(obj, "components") else list(obj)
    for ci, c in enumerate(comps):
        c = np.asarray(c).ravel()
        j = ci // 3
        ax = "xyz"[ci % 3]
        keep[f"r{j + 1}{ax}"] = _f64(np.real(c), n)
        keep[f"i{j + 1}{ax}"] = _f64(np.imag(c), n)
    for name, arr in keep.items():
        setattr(b, name, _ptr(arr))
    b._keep = keep
    return b


def fill_arrc1(S_field, n):
    """scalarfield -> arrC1 (BindUtils.py sfieldConv)."""
    b = S.arrC1()
    s = np.asarray(S_field.S if hasattr(S_field, "S") else S_field).ravel()
    keep = {"x": _f64(np.real(s), n), "y": _f64(np.imag(s), n)}
    b.x = _ptr(keep["x"])
    b.y = _ptr(keep["y"])
    b._keep = keep
    return b


def _c3(keep, pre, j, shape):
    return [(keep[f"r{j}{a}"] + 1j * keep[f"i{j}{a}"]).reshape(shape) for a in "xyz"]


def c2bundle_to_obj(b, shape, obj_t="fields"):
    """c2Bundle -> currents/fields (BindUtils.py:52-83)."""
    k = b._keep
    comps = _c3(k, "", 1, shape) + _c3(k, "", 2, shape)
    return T.currents(*comps) if obj_t == "currents" else T.fields(*comps)


def c4bundle_to_objs(b, shape):
    """c4Bundle -> (currents J,M ; fields E,H) (BindUtils.py c4BundleToObj)."""
    k = b._keep
    jm = T.currents(*(_c3(k, "", 1, shape) + _c3(k, "", 2, shape)))
    eh = T.fields(*(_c3(k, "", 3, shape) + _c3(k, "", 4, shape)))
    return jm, eh


def c2rbundle_to_objs(b, shape):
    """c2rBundle -> (fields E_r,H_r ; rfield S_r) (BindUtils.py c2rBundleToObj)."""
    k = b._keep
    eh = T.fields(*(_c3(k, "", 1, shape) + _c3(k, "", 2, shape)))
    pr = T.rfield(*(k[f"r3{a}"].reshape(shape) for a in "xyz"))
    return eh, pr


def arrc1_to_obj(b, shape):
    k = b._keep
    return T.scalarfield((k["x"] + 1j * k["y"]).reshape(shape))


def fill_cframe(fr):
    """frame -> cframe (BindUtils.py:401-423)."""
    c = S.cframe()
    c.size = int(fr.size)
    keep = {}
    for name in ("x", "y", "z", "dx", "dy", "dz"):
        keep[name] = _f64(getattr(fr, name), fr.size).copy()
        setattr(c, name, _ptr(keep[name]))
    c._keep = keep
    return c


def alloc_cframe(n):
    c = S.cframe()
    c.size = int(n)
    keep = {}
    for name in ("x", "y", "z", "dx", "dy", "dz"):
        keep[name] = np.zeros(int(n), dtype=np.float64)
        setattr(c, name, _ptr(keep[name]))
    c._keep = keep
    return c


def cframe_to_frame(c):
    """cframe -> frame (BindUtils.py:600-625)."""
    k = c._keep
    return T.frame(c.size, k["x"], k["y"], k["z"], k["dx"], k["dy"], k["dz"])
