"""Figures of merit of src/PyPO/Efficiencies.py as device reductions.

Each function reduces fields or frames in device memory with a
libpypohip.so kernel (metrics.hip) and returns the reference's number:

  calcSpillover (Efficiencies.py:90-115)  s1 / s0 written as |s1|^2 / (s0 s1)
  calcTaper     (117-145)                 |sum_in f dA|^2 / sum_in |f|^2 dA / sum_in dA
  calcXpol      (147-159)                 1 - P_x / (P_co + P_x)
  calcRTcenter / calcRTtilt / calcRMS (43-88) over the rays with |d|^2 > 0.8

A NaN spillover or taper (an empty aperture or a zero field) is 0, as in the
reference.  Inputs may be host objects (the reference's types: a complex
field component and a reflDict, a frame) or device tensors (device.py
layouts): host data is copied to the device once, device data is reduced in
place.  The aperture follows the aperDict keys ("shape", "center", "outer",
"outer_x", "outer_y"), checked like Checks.check_aperDict (Checks.py:1656-1709).
"""

import copy
import ctypes

import numpy as np
import torch

from . import bind as B
from . import device as D
from . import structs as PS

APER_SHAPES = {"ellipse": 0, "ell": 0, "rectangle": 1, "rect": 1}


class ApertureError(Exception):
    pass


def _declare(lib):
    if getattr(lib, "_metrics_declared", False):
        return
    P, dp, vp = ctypes.POINTER, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p
    lib.pypo_hip_aperture_sums_dev.argtypes = [P(PS.reflcontainer), dp, dp, PS.aperparams, dp, vp]
    lib.pypo_hip_power_sums_dev.argtypes = [dp, dp, dp, dp, ctypes.c_int, dp, vp]
    lib.pypo_hip_frame_stats_dev.argtypes = [P(PS.cframe), dp, vp]
    for f in ("aperture_sums", "power_sums", "frame_stats"):
        getattr(lib, f"pypo_hip_{f}_dev").restype = ctypes.c_int
    lib._metrics_declared = True


def _lib():
    lib = B.loadHIPlib()
    _declare(lib)
    return lib


def _shape_code(shape):
    """AperShapes.ELL / RECT (enum, its name or value) or "ellipse" / "rectangle"."""
    if shape is None:
        return 0
    name = getattr(shape, "name", shape)
    if isinstance(name, str):
        key = name.lower()
        if key in APER_SHAPES:
            return APER_SHAPES[key]
        raise ApertureError(f"unknown aperture shape {shape!r}")
    val = int(getattr(shape, "value", shape))
    if val not in (0, 1):
        raise ApertureError(f"unknown aperture shape {shape!r}")
    return val


def check_aperDict(aperDict):
    """Defaults and required keys of Checks.check_aperDict (Checks.py:1656-1709)."""
    err = []
    if "plot" in aperDict and not isinstance(aperDict["plot"], bool):
        err.append("aperDict: plot must be a bool")
    aperDict.setdefault("plot", True)
    if "center" in aperDict:
        if np.asarray(aperDict["center"]).shape != (2,):
            err.append("aperDict: center must have shape (2,)")
    else:
        aperDict["center"] = np.zeros(2)
    if aperDict.get("shape") is None:
        aperDict["shape"] = "ellipse"
    code = _shape_code(aperDict["shape"])
    needs = ["outer", "inner"] if code == 0 else ["outer_x", "outer_y", "inner_x", "inner_y"]
    err += [f"aperDict: missing {k}" for k in needs if k not in aperDict]
    if err:
        raise ApertureError("\n".join(err))
    return aperDict


def aper_struct(aperDict):
    """aperparams for an aperDict (None or {}: every point)."""
    a = PS.aperparams()
    if not aperDict:
        a.shape = 2
        return a
    d = check_aperDict(copy.copy(aperDict))
    a.shape = _shape_code(d["shape"])
    a.center[:] = [float(v) for v in np.asarray(d["center"], dtype=float)]
    if a.shape == 0:
        a.outer[:] = [float(v) for v in np.asarray(d["outer"], dtype=float)]
    else:
        a.outer_x[:] = [float(v) for v in np.asarray(d["outer_x"], dtype=float)]
        a.outer_y[:] = [float(v) for v in np.asarray(d["outer_y"], dtype=float)]
    return a


def _dptr(t):
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), ctypes.POINTER(ctypes.c_double))


def _call(lib, fn, args, what):
    if fn(*args):
        B.check(lib, what)


def _complex_rows(comp, device="cuda"):
    """(2, n) float64 device tensor (re, im) of a complex component, or the tensor itself."""
    if isinstance(comp, torch.Tensor):
        if comp.dim() != 2 or comp.shape[0] != 2:
            raise ValueError("field component tensor must be (2, n): rows re, im")
        return comp.contiguous()
    c = np.ravel(np.asarray(comp))
    return D.to_device(np.stack([np.real(c), np.imag(c)]), device)


def rest_grid(surfaceObject, device="cuda"):
    """(7, n) device grid of a surface in its rest frame (generateGrid with
    transform=False: the transformation is the identity)."""
    r = copy.copy(surfaceObject)
    r["transf"] = np.eye(4)
    return D.generate_grid(r, device)


# ---- raw reductions -------------------------------------------------------------------

def aperture_sums(grid, comp, aperDict=None):
    """[sum |f|^2 dA, sum_in |f|^2 dA, sum_in Re f dA, sum_in Im f dA, sum_in dA] over a
    (7, n) device grid and a (2, n) component tensor (rows re, im)."""
    lib = _lib()
    D._check_dev(grid, 7, "grid")
    D._check_dev(comp, 2, "field component")
    if comp.shape[1] != grid.shape[1]:
        raise ValueError("field component and grid must have the same number of points")
    out = (ctypes.c_double * 5)()
    g = D.container(grid)
    _call(lib, lib.pypo_hip_aperture_sums_dev, [ctypes.byref(g), _dptr(comp[0]), _dptr(comp[1]),
                                                aper_struct(aperDict), out, D._stream()],
          "pypo_hip_aperture_sums_dev")
    return np.array(out[:])


def power_sums(a, b):
    """[sum |a|^2, sum |b|^2] of two (2, n) component tensors."""
    lib = _lib()
    D._check_dev(a, 2, "co-polar component")
    D._check_dev(b, 2, "cross-polar component")
    if a.shape[1] != b.shape[1]:
        raise ValueError("components must have the same number of points")
    out = (ctypes.c_double * 2)()
    _call(lib, lib.pypo_hip_power_sums_dev, [_dptr(a[0]), _dptr(a[1]), _dptr(b[0]), _dptr(b[1]),
                                             int(a.shape[1]), out, D._stream()], "pypo_hip_power_sums_dev")
    return np.array(out[:])


def frame_stats(frame):
    """[count, sum x, y, z, dx, dy, dz, sum |p - c|^2] over the rays of a (6, n)
    device frame (or a frame object) with |d|^2 > 0.8."""
    lib = _lib()
    t = frame if isinstance(frame, torch.Tensor) else D.frame_tensor(frame)
    out = (ctypes.c_double * 8)()
    fs = D.frame_struct(t)
    _call(lib, lib.pypo_hip_frame_stats_dev, [ctypes.byref(fs), out, D._stream()], "pypo_hip_frame_stats_dev")
    return np.array(out[:])


# ---- the reference's functions ---------------------------------------------------------

def _nan_to_zero(v):
    return 0.0 if np.isnan(v) else float(v)


def calcSpillover(field, surfaceObject, aperDict, grid=None):
    """Spillover efficiency of a field component on an aperture (Efficiencies.py:90-115).
    `grid` may pass the surface's rest-frame device grid to skip generating it."""
    if not aperDict:
        raise ApertureError("calcSpillover needs an aperDict")
    comp = _complex_rows(field)
    g = rest_grid(surfaceObject, comp.device) if grid is None else grid
    s = aperture_sums(g, comp, aperDict)
    with np.errstate(divide="ignore", invalid="ignore"):
        return _nan_to_zero(s[1] * s[1] / (s[0] * s[1]))


def calcTaper(field, surfaceObject, aperDict=None, grid=None):
    """Taper efficiency of a field component, over an aperture or the whole
    surface (Efficiencies.py:117-145)."""
    comp = _complex_rows(field)
    g = rest_grid(surfaceObject, comp.device) if grid is None else grid
    s = aperture_sums(g, comp, aperDict)
    with np.errstate(divide="ignore", invalid="ignore"):
        return _nan_to_zero((s[2] * s[2] + s[3] * s[3]) / s[1] / s[4])


def calcXpol(Cofield, Xfield):
    """Cross-polar efficiency (Efficiencies.py:147-159)."""
    p = power_sums(_complex_rows(Cofield), _complex_rows(Xfield))
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(1 - p[1] / (p[0] + p[1]))


def _frame_mean(s):
    with np.errstate(divide="ignore", invalid="ignore"):
        return s[1:4] / s[0], s[4:7] / s[0]


def calcRTcenter(frame):
    """Mean position of the rays that hit (Efficiencies.py:43-57)."""
    return _frame_mean(frame_stats(frame))[0]


def calcRTtilt(frame):
    """Normalised mean direction of the rays that hit (Efficiencies.py:59-73)."""
    t = _frame_mean(frame_stats(frame))[1]
    with np.errstate(divide="ignore", invalid="ignore"):
        return t / np.linalg.norm(t)


def calcRMS(frame):
    """RMS distance of the rays that hit from their mean position (Efficiencies.py:75-88)."""
    s = frame_stats(frame)
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.sqrt(s[7] / s[0]))
