"""Beam initialisation through libpypohip.so -- the GPU twin of the reference's
src/PyPO/BindBeam.py (makeRTframe 52-80, makeGRTframe 82-112, makeGauss
114-157, makeGaussBeam, makeScalarGauss 200-228, calcCurrents 230-275).

Dictionaries use the reference's keys and units (angles in degrees; the
fillers convert to radians like BindUtils.allfill_RTDict / allfill_GRTDict).
The host-buffer calls below are synchronous; the device variants take and
return torch tensors (device.py layouts).
"""

import ctypes

import numpy as np

from . import bind as B
from . import marshal as MU
from . import structs as PS
from .geometry import gridsize

CURRENT_MODES = {"full": 0, "PMC": 1, "PEC": 2, "source": 3}
BEAM_MODES = ["full", "pmc", "pec"]   # vecGPODict "mode" names (Checks.Beam_modelist)


def _declare(lib):
    if getattr(lib, "_beam_declared", False):
        return
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.makeRTframe.argtypes = [PS.RTDict, P(PS.cframe)]
    lib.makeGRTframe.argtypes = [PS.GRTDict, P(PS.cframe)]
    lib.makeGauss.argtypes = [PS.GPODict, PS.reflparams, P(PS.c2Bundle), P(PS.c2Bundle)]
    lib.makeGaussBeam.argtypes = [PS.vecGPODict, PS.reflparams, P(PS.c2Bundle), P(PS.c2Bundle)]
    lib.makeScalarGauss.argtypes = [PS.ScalarGPODict, PS.reflparams, P(PS.arrC1)]
    lib.calcCurrents.argtypes = [P(PS.c2Bundle), P(PS.c2Bundle), PS.reflparams, ctypes.c_int]
    for f in ("makeRTframe", "makeGRTframe", "makeGauss", "makeGaussBeam", "makeScalarGauss",
              "calcCurrents"):
        getattr(lib, f).restype = None
    lib.pypo_hip_make_gauss_dev.argtypes = [PS.GPODict, P(PS.reflcontainer), P(PS.c2Bundle),
                                            P(PS.c2Bundle), vp]
    lib.pypo_hip_make_gauss_beam_dev.argtypes = [PS.vecGPODict, P(PS.reflcontainer), P(PS.c2Bundle),
                                                 P(PS.c2Bundle), vp]
    lib.pypo_hip_make_scalar_gauss_dev.argtypes = [PS.ScalarGPODict, P(PS.reflcontainer), P(PS.arrC1), vp]
    lib.pypo_hip_calc_currents_dev.argtypes = [P(PS.c2Bundle), P(PS.c2Bundle), P(PS.reflcontainer),
                                               ctypes.c_int, vp]
    lib.pypo_hip_make_rt_frame_dev.argtypes = [PS.RTDict, P(PS.cframe), vp]
    lib.pypo_hip_make_grt_frame_dev.argtypes = [PS.GRTDict, P(PS.cframe), vp]
    for f in ("make_gauss", "make_gauss_beam", "make_scalar_gauss", "calc_currents", "make_rt_frame",
              "make_grt_frame"):
        getattr(lib, f"pypo_hip_{f}_dev").restype = ctypes.c_int
    lib._beam_declared = True


def _lib():
    lib = B.loadHIPlib()
    _declare(lib)
    return lib


# ---- dictionary -> struct (BindUtils.py:456-554) ------------------------------------

def fill_RTDict(d):
    return PS.RTDict(int(d["nRays"]), int(d["nRing"]), float(np.radians(d["angx0"])),
                     float(np.radians(d["angy0"])), float(d["x0"]), float(d["y0"]))


def fill_GRTDict(d):
    return PS.GRTDict(int(d["nRays"]), float(np.radians(d["angx0"])), float(np.radians(d["angy0"])),
                      float(d["x0"]), float(d["y0"]), int(d.get("seed", -1)))


def fill_GPODict(d):
    pol = (ctypes.c_double * 3)(*np.asarray(d.get("pol", [1.0, 0.0, 0.0]), dtype=np.float64).tolist())
    s = PS.GPODict(float(d["lam"]), float(d["w0x"]), float(d["w0y"]), float(d.get("n", 1.0)),
                   float(d.get("E0", 1.0)), float(d.get("dxyz", 0.0)), pol)
    s._keep = pol
    return s


def beam_mode(mode):
    if isinstance(mode, str):
        return BEAM_MODES.index(mode.lower())
    return int(mode)


def fill_vecGPODict(d):
    return PS.vecGPODict(float(d["lam"]), float(d["w0"]), float(d["z"]), float(d.get("n", 1.0)),
                         float(d.get("power", 4 * np.pi)), beam_mode(d.get("mode", 0)))


def fill_SGPODict(d):
    return PS.ScalarGPODict(float(d["lam"]), float(d["w0x"]), float(d["w0y"]), float(d.get("n", 1.0)),
                            float(d.get("E0", 1.0)), float(d.get("dxyz", 0.0)))


def _run(lib, fn, args, what):
    fn(*args)
    B.check(lib, what)


# ---- host-buffer calls (BindBeam.py) ---------------------------------------------------

def makeRTframe(RTDict_py):
    """Tubular ray frame (BindBeam.makeRTframe) -> frame."""
    lib = _lib()
    n = 1 + int(RTDict_py["nRays"]) * 4 * int(RTDict_py["nRing"])
    res = MU.alloc_cframe(n)
    _run(lib, lib.makeRTframe, [fill_RTDict(RTDict_py), ctypes.byref(res)], "makeRTframe")
    return MU.cframe_to_frame(res)


def makeGRTframe(grdict_py):
    """Gaussian ray frame by rejection sampling (BindBeam.makeGRTframe) -> frame."""
    lib = _lib()
    n = int(grdict_py["nRays"])
    res = MU.alloc_cframe(n)
    _run(lib, lib.makeGRTframe, [fill_GRTDict(grdict_py), ctypes.byref(res)], "makeGRTframe")
    return MU.cframe_to_frame(res)


def makeGauss(gdict_py, source):
    """Astigmatic Gaussian beam on `source` -> (fields, currents)."""
    lib = _lib()
    shape = gridsize(source)
    n = int(np.prod(shape))
    f, c = MU.alloc_bundle("c2Bundle", n), MU.alloc_bundle("c2Bundle", n)
    _run(lib, lib.makeGauss, [fill_GPODict(gdict_py), MU.fill_reflparams(source), ctypes.byref(f),
                              ctypes.byref(c)], "makeGauss")
    return MU.c2bundle_to_obj(f, shape, "fields"), MU.c2bundle_to_obj(c, shape, "currents")


def makeGaussBeam(gdict_py, source):
    """Complex-source-point vector Gaussian on `source` -> (fields, currents)."""
    lib = _lib()
    shape = gridsize(source)
    n = int(np.prod(shape))
    f, c = MU.alloc_bundle("c2Bundle", n), MU.alloc_bundle("c2Bundle", n)
    _run(lib, lib.makeGaussBeam, [fill_vecGPODict(gdict_py), MU.fill_reflparams(source), ctypes.byref(f),
                                  ctypes.byref(c)], "makeGaussBeam")
    return MU.c2bundle_to_obj(f, shape, "fields"), MU.c2bundle_to_obj(c, shape, "currents")


def makeScalarGauss(gdict_py, source):
    """Scalar Gaussian on `source` -> scalarfield."""
    lib = _lib()
    shape = gridsize(source)
    n = int(np.prod(shape))
    f = MU.alloc_bundle("arrC1", n)
    _run(lib, lib.makeScalarGauss, [fill_SGPODict(gdict_py), MU.fill_reflparams(source), ctypes.byref(f)],
         "makeScalarGauss")
    return MU.arrc1_to_obj(f, shape)


def calcCurrents(fields, source, mode):
    """Currents from fields on `source`; mode "full", "PMC", "PEC" or "source"."""
    lib = _lib()
    if isinstance(mode, str):
        if mode not in CURRENT_MODES:
            raise ValueError(f"calcCurrents: unknown mode {mode!r}")
        mode = CURRENT_MODES[mode]
    shape = gridsize(source)
    n = int(np.prod(shape))
    fin = MU.fill_c2bundle(fields, n)
    c = MU.alloc_bundle("c2Bundle", n)
    _run(lib, lib.calcCurrents, [ctypes.byref(fin), ctypes.byref(c), MU.fill_reflparams(source),
                                 int(mode)], "calcCurrents")
    return MU.c2bundle_to_obj(c, shape, "currents")


# ---- device-resident variants (torch tensors, device.py layouts) ----------------------

def _dev_call(lib, fn, args, what):
    rc = fn(*args)
    if rc:
        B.check(lib, what)


def make_gauss(gdict_py, grid):
    """(12, n) field and (12, n) current tensors of an astigmatic Gaussian on a device grid."""
    import torch
    from . import device as D
    lib = _lib()
    n = grid.shape[1]
    f = torch.empty((12, n), dtype=torch.float64, device=grid.device)
    c = torch.empty_like(f)
    g, bf, bc = D.container(grid), D.bundle(f), D.bundle(c)
    _dev_call(lib, lib.pypo_hip_make_gauss_dev, [fill_GPODict(gdict_py), ctypes.byref(g), ctypes.byref(bf),
                                                 ctypes.byref(bc), D._stream()], "pypo_hip_make_gauss_dev")
    return f, c


def make_gauss_beam(gdict_py, grid):
    import torch
    from . import device as D
    lib = _lib()
    n = grid.shape[1]
    f = torch.empty((12, n), dtype=torch.float64, device=grid.device)
    c = torch.empty_like(f)
    g, bf, bc = D.container(grid), D.bundle(f), D.bundle(c)
    _dev_call(lib, lib.pypo_hip_make_gauss_beam_dev, [fill_vecGPODict(gdict_py), ctypes.byref(g),
                                                      ctypes.byref(bf), ctypes.byref(bc), D._stream()],
              "pypo_hip_make_gauss_beam_dev")
    return f, c


def make_scalar_gauss(gdict_py, grid):
    import torch
    from . import device as D
    lib = _lib()
    f = torch.empty((2, grid.shape[1]), dtype=torch.float64, device=grid.device)
    g, bf = D.container(grid), D.bundle(f)
    _dev_call(lib, lib.pypo_hip_make_scalar_gauss_dev, [fill_SGPODict(gdict_py), ctypes.byref(g),
                                                        ctypes.byref(bf), D._stream()],
              "pypo_hip_make_scalar_gauss_dev")
    return f


def calc_currents(field, grid, mode):
    import torch
    from . import device as D
    lib = _lib()
    if isinstance(mode, str):
        mode = CURRENT_MODES[mode]
    c = torch.empty_like(field)
    g, bf, bc = D.container(grid), D.bundle(field), D.bundle(c)
    _dev_call(lib, lib.pypo_hip_calc_currents_dev, [ctypes.byref(bf), ctypes.byref(bc), ctypes.byref(g),
                                                    int(mode), D._stream()], "pypo_hip_calc_currents_dev")
    return c


def make_rt_frame(RTDict_py, device="cuda"):
    import torch
    from . import device as D
    lib = _lib()
    n = 1 + int(RTDict_py["nRays"]) * 4 * int(RTDict_py["nRing"])
    fr = torch.empty((6, n), dtype=torch.float64, device=device)
    s = D.frame_struct(fr)
    _dev_call(lib, lib.pypo_hip_make_rt_frame_dev, [fill_RTDict(RTDict_py), ctypes.byref(s), D._stream()],
              "pypo_hip_make_rt_frame_dev")
    return fr


def make_grt_frame(grdict_py, device="cuda"):
    import torch
    from . import device as D
    lib = _lib()
    fr = torch.empty((6, int(grdict_py["nRays"])), dtype=torch.float64, device=device)
    s = D.frame_struct(fr)
    _dev_call(lib, lib.pypo_hip_make_grt_frame_dev, [fill_GRTDict(grdict_py), ctypes.byref(s), D._stream()],
              "pypo_hip_make_grt_frame_dev")
    return fr
