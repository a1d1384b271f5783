"""Single-precision binding: the twin of the reference's src/PyPO/BindGPU.py.

The reference's GPU library takes FP32 structs (src/CUDA/InterfaceCUDA.h:29-69).
libpypohip.so exports the same callKernelf_* / callRTKernel symbols
(widening to FP64 inside, mixed-precision sums, FP32 results), so the
reference's own BindGPU runs unchanged against it; this module is the same
binding written against this package's types: loadGPUlib (BindGPU.py:17-81),
PyPO_GPUf (83-224) and RT_GPUf (226-263), with FP32 buffers and results
returned as float64 arrays like the reference's.
"""

import ctypes

import numpy as np

from . import bind as B
from . import structs as S
from . import types as T
from .geometry import gridsize

_fp = ctypes.POINTER(ctypes.c_float)
_MODES = {"JM": ("callKernelf_JM", S.c2Bundlef), "EH": ("callKernelf_EH", S.c2Bundlef),
          "JMEH": ("callKernelf_JMEH", S.c4Bundlef), "EHP": ("callKernelf_EHP", S.c2rBundlef),
          "FF": ("callKernelf_FF", S.c2Bundlef), "scalar": ("callKernelf_scalar", S.arrC1f)}
# The reference CPU library's FP32 names (InterfaceCPU.h:64-101): same
# arguments with numThreads after k instead of nBlocks, nThreads at the end.
_CPU_NAMES = {"JM": "propagateToGridf_JM", "EH": "propagateToGridf_EH", "JMEH": "propagateToGridf_JMEH",
              "EHP": "propagateToGridf_EHP", "FF": "propagateToFarFieldf", "scalar": "propagateToGridf_scalar"}


def loadGPUlib():
    """libpypohip.so with the FP32 entry points declared as BindGPU.loadGPUlib declares them."""
    lib = B.loadHIPlib()
    if getattr(lib, "_f32_declared", False):
        return lib
    P, f, i = ctypes.POINTER, ctypes.c_float, ctypes.c_int
    for mode, (sym, bundle) in _MODES.items():
        src = S.arrC1f if mode == "scalar" else S.c2Bundlef
        fn = getattr(lib, sym)
        fn.argtypes = [P(bundle), S.reflparamsf, S.reflparamsf, P(S.reflcontainerf), P(S.reflcontainerf),
                       P(src), f, f, f, i, i]
        fn.restype = None
        fn = getattr(lib, _CPU_NAMES[mode])
        fn.argtypes = [P(bundle), S.reflparamsf, S.reflparamsf, P(S.reflcontainerf), P(S.reflcontainerf),
                       P(src), f, i, f, f]
        fn.restype = None
    lib.callRTKernel.argtypes = [S.reflparamsf, P(S.cframef), P(S.cframef), f, f, i, i]
    lib.callRTKernel.restype = None
    lib._f32_declared = True
    return lib


def _f32(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def fill_reflparamsf(refl):
    """reflDict -> reflparamsf (BindUtils.allfill_reflparams with ctypes.c_float)."""
    p = S.reflparamsf()
    gmode = int(refl["gmode"])
    lims = {0: ("lims_x", "lims_y"), 1: ("lims_u", "lims_v"), 2: ("lims_Az", "lims_El")}[gmode]
    keep = {"coeffs": _f32(refl.get("coeffs", np.zeros(3)), 3), "lxu": _f32(refl[lims[0]], 2),
            "lyv": _f32(refl[lims[1]], 2), "gcenter": _f32(refl.get("gcenter", np.zeros(2)), 2),
            "transf": _f32(np.asarray(refl.get("transf", np.eye(4))).ravel(), 16),
            "n_cells": np.ascontiguousarray(np.asarray(refl["gridsize"], dtype=np.int32).ravel())}
    for k in ("coeffs", "lxu", "lyv", "gcenter", "transf"):
        setattr(p, k, keep[k].ctypes.data_as(_fp))
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


def _alloc(struct, names, n, init=None):
    s = struct()
    keep = {}
    for j, name in enumerate(names):
        keep[name] = _f32(init[j], n) if init is not None else np.zeros(int(n), dtype=np.float32)
        setattr(s, name, keep[name].ctypes.data_as(_fp))
    if hasattr(s, "size"):
        s.size = int(n)
    s._keep = keep
    return s


def _names(struct):
    return [n for n, _ in struct._fields_ if n != "size"]


def _complex_rows(obj):
    comps = obj.components() if hasattr(obj, "components") else list(obj)
    rows = []
    for j in range(0, len(comps), 3):
        rows += [np.real(np.ravel(c)) for c in comps[j:j + 3]] + [np.imag(np.ravel(c)) for c in comps[j:j + 3]]
    return rows


def PyPO_GPUf(source, target, runPODict, interface="gpu"):
    """PO propagation through the FP32 entry points; result types as BindGPU.PyPO_GPUf.
    interface "gpu" calls callKernelf_* (InterfaceCUDA.h), "cpu" the CPU
    library's propagateToGridf_* / propagateToFarFieldf (InterfaceCPU.h)."""
    if interface not in ("gpu", "cpu"):
        raise ValueError(f"interface must be 'gpu' or 'cpu', not {interface!r}")
    lib = loadGPUlib()
    mode = runPODict["mode"]
    if mode not in _MODES:
        raise B.PyPOHIPError(f"{mode} is not a valid propagation mode")
    sym, bundle = _MODES[mode]
    gs = int(np.prod(gridsize(source)))
    gt = int(np.prod(gridsize(target)))
    shape = gridsize(target)
    cs = _alloc(S.reflcontainerf, _names(S.reflcontainerf), gs)
    ct = _alloc(S.reflcontainerf, _names(S.reflcontainerf), gt)
    if mode == "scalar":
        sf = np.ravel(np.asarray(getattr(runPODict["s_scalarfield"], "S", runPODict["s_scalarfield"])))
        src = _alloc(S.arrC1f, ["x", "y"], gs, [np.real(sf), np.imag(sf)])
    else:
        src = _alloc(S.c2Bundlef, _names(S.c2Bundlef), gs, _complex_rows(runPODict["s_current"]))
    res = _alloc(bundle, _names(bundle), gt)
    tdir = -1.0 if runPODict.get("exp", "fwd") == "fwd" else 1.0
    nthreads = int(runPODict.get("nThreads", 256))
    if interface == "gpu":
        getattr(lib, sym)(ctypes.byref(res), fill_reflparamsf(source), fill_reflparamsf(target), ctypes.byref(cs),
                          ctypes.byref(ct), ctypes.byref(src), float(runPODict["k"]),
                          float(runPODict.get("epsilon", 1.0)), tdir, (gt + nthreads - 1) // nthreads, nthreads)
    else:
        sym = _CPU_NAMES[mode]
        getattr(lib, sym)(ctypes.byref(res), fill_reflparamsf(source), fill_reflparamsf(target), ctypes.byref(cs),
                          ctypes.byref(ct), ctypes.byref(src), float(runPODict["k"]),
                          int(runPODict.get("nThreads", 1)), float(runPODict.get("epsilon", 1.0)), tdir)
    B.check(lib, sym)
    a = [res._keep[n].astype(np.float64).reshape(shape) for n in _names(bundle)]

    def c6(v):
        return [v[i] + 1j * v[i + 3] for i in range(3)] + [v[6 + i] + 1j * v[9 + i] for i in range(3)]

    if mode == "JM":
        return T.currents(*c6(a))
    if mode in ("EH", "FF"):
        return T.fields(*c6(a))
    if mode == "JMEH":
        return [T.currents(*c6(a[:12])), T.fields(*c6(a[12:]))]
    if mode == "EHP":
        return [T.fields(*c6(a[:12])), T.rfield(*a[12:15])]
    return T.scalarfield(a[0] + 1j * a[1])


def RT_GPUf(runRTDict):
    """Ray trace through callRTKernel (BindGPU.RT_GPUf); returns a frame of float64 arrays."""
    lib = loadGPUlib()
    fr = runRTDict["fr_in"]
    n = int(fr.size)
    names = _names(S.cframef)
    inp = _alloc(S.cframef, names, n, [getattr(fr, a) for a in names])
    out = _alloc(S.cframef, names, n)
    nthreads = int(runRTDict.get("nThreads", 256))
    lib.callRTKernel(fill_reflparamsf(runRTDict["t_name"]), ctypes.byref(inp), ctypes.byref(out),
                     float(runRTDict.get("tol", 1e-3)), float(runRTDict.get("t0", 1.0)),
                     (n + nthreads - 1) // nthreads, nthreads)
    B.check(lib, "callRTKernel")
    return T.frame(n, *[out._keep[a].astype(np.float64) for a in names])
