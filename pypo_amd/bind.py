"""ctypes binding of libpypohip.so -- the GPU twin of the reference's
src/PyPO/BindCPU.py (loadCPUlib 17-80, PyPO_CPUd 82-220, RT_CPUd 222-254).

The FP64 drop-in symbols take exactly the argument lists BindCPU declares,
so PyPO_HIPd / RT_HIPd below are BindCPU's call sequence pointed at the HIP
library.  Extra runPODict keys understood here (ignored by the reference):
  "precision": "mixed" (FP32 integrand, FP64 distance/phase; default) or "fp64".
Errors never exit the process: the library records them and this module
raises PyPOHIPError after the call.
"""

import ctypes
import os
import threading

import numpy as np

from . import marshal as MU
from . import structs as PS
from .geometry import gridsize

LIB_NAME = "libpypohip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

MODES = {"JM": 0, "EH": 1, "JMEH": 2, "EHP": 3, "FF": 4, "scalar": 5}
PRECISIONS = {"mixed": 0, "fp64": 1}

_lib = None
_lock = threading.Lock()


class PyPOHIPError(RuntimeError):
    """A libpypohip.so call failed (bad arguments, no device, HIP error)."""


def loadHIPlib():
    """Load libpypohip.so and declare its argument types (BindCPU.loadCPUlib).

    Raises OSError if the library has not been built: there is no CPU
    fallback inside this package."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        # One HIP runtime per process: PyTorch-ROCm carries its own
        # libamdhip64.so.7 / libhsa-runtime64.so.1.  Loading torch first makes
        # libpypohip.so bind to that same runtime (matching sonames); loading
        # this library first would leave torch's HIP on a foreign HSA runtime
        # with no visible GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `python -m pypo_amd.build_lib`")
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        d, i = ctypes.c_double, ctypes.c_int
        for name, bundle in (("propagateToGrid_JM", PS.c2Bundle), ("propagateToGrid_EH", PS.c2Bundle),
                             ("propagateToGrid_JMEH", PS.c4Bundle), ("propagateToGrid_EHP", PS.c2rBundle),
                             ("propagateToFarField", PS.c2Bundle)):
            f = getattr(lib, name)
            f.argtypes = [P(bundle), PS.reflparams, PS.reflparams, P(PS.reflcontainer),
                          P(PS.reflcontainer), P(PS.c2Bundle), d, i, d, d]
            f.restype = None
        lib.propagateToGrid_scalar.argtypes = [P(PS.arrC1), PS.reflparams, PS.reflparams,
                                               P(PS.reflcontainer), P(PS.reflcontainer), P(PS.arrC1),
                                               d, i, d, d]
        lib.propagateToGrid_scalar.restype = None
        lib.propagateRays.argtypes = [PS.reflparams, P(PS.cframe), P(PS.cframe), i, d, d]
        lib.propagateRays.restype = None
        lib.generateGrid.argtypes = [PS.reflparams, P(PS.reflcontainer), ctypes.c_bool, ctypes.c_bool]
        lib.generateGrid.restype = None
        lib.pypo_hip_last_error.argtypes = [ctypes.c_char_p, i]
        lib.pypo_hip_last_error.restype = i
        lib.pypo_hip_version.argtypes = []
        lib.pypo_hip_version.restype = ctypes.c_char_p
        lib.pypo_hip_device_count.argtypes = []
        lib.pypo_hip_device_count.restype = i
        lib.pypo_hip_stream_ordered_scratch.argtypes = []
        lib.pypo_hip_stream_ordered_scratch.restype = i
        lib.pypo_hip_scratch_bytes.argtypes = [ctypes.c_void_p]
        lib.pypo_hip_scratch_bytes.restype = ctypes.c_longlong
        lib.pypo_hip_release_scratch.argtypes = [ctypes.c_void_p]
        lib.pypo_hip_release_scratch.restype = i
        lib.pypo_hip_set_precision.argtypes = [i]
        lib.pypo_hip_set_precision.restype = i
        lib.pypo_hip_set_rt_max_iter.argtypes = [i]
        lib.pypo_hip_set_rt_max_iter.restype = i
        lib.pypo_hip_set_timing.argtypes = [i]
        lib.pypo_hip_set_timing.restype = i
        lib.pypo_hip_last_kernel_ms.argtypes = []
        lib.pypo_hip_last_kernel_ms.restype = d
        lib.pypo_hip_propagate_dev.argtypes = [i, ctypes.c_void_p, P(PS.reflcontainer),
                                               P(PS.reflcontainer), ctypes.c_void_p, i, i, i, d, d, d,
                                               i, ctypes.c_void_p]
        lib.pypo_hip_propagate_dev.restype = i
        lib.pypo_hip_propagate_noalloc.argtypes = lib.pypo_hip_propagate_dev.argtypes
        lib.pypo_hip_propagate_noalloc.restype = i
        lib.pypo_hip_generate_grid_dev.argtypes = [PS.reflparams, P(PS.reflcontainer), ctypes.c_void_p]
        lib.pypo_hip_generate_grid_dev.restype = i
        lib.pypo_hip_trace_dev.argtypes = [P(PS.reflparams), i, P(PS.cframe), d, d, ctypes.c_void_p]
        lib.pypo_hip_trace_dev.restype = i
        lib.pypo_hip_set_sum_shape.argtypes = [i, i]
        lib.pypo_hip_set_sum_shape.restype = i
        lib.pypo_hip_set_ff_kernel.argtypes = [i]
        lib.pypo_hip_set_ff_kernel.restype = i
        lib.pypo_hip_inject_alloc_failures.argtypes = [i]
        lib.pypo_hip_inject_alloc_failures.restype = i
        RC, CB = P(PS.reflcontainer), P(PS.c2Bundle)
        lib.pypo_hip_displace_dev.argtypes = [RC, P(d), RC, ctypes.c_void_p]
        lib.pypo_hip_displace_dev.restype = i
        lib.pypo_hip_ff_power_grad_dev.argtypes = [RC, i, i, i, RC, i, i, i, CB, RC, CB, P(d), d, d,
                                                   i, P(d), ctypes.c_void_p]
        lib.pypo_hip_ff_power_grad_dev.restype = i
        _lib = lib
        return lib


def check(lib, what=""):
    """Raise PyPOHIPError if the last library call on this thread failed."""
    buf = ctypes.create_string_buffer(1024)
    code = lib.pypo_hip_last_error(buf, 1024)
    if code:
        raise PyPOHIPError(f"{what}: {buf.value.decode(errors='replace')} (code {code})")


def _run(lib, fn, args, what):
    """Run one synchronous FFI call.  The reference wraps each call in a daemon
    thread (Threadmgr.new_sthread) so that Ctrl-C stays responsive; the error
    state is per thread, so it is read on the thread that made the call."""
    box = {}

    def body():
        fn(*args)
        buf = ctypes.create_string_buffer(1024)
        box["code"] = lib.pypo_hip_last_error(buf, 1024)
        box["msg"] = buf.value.decode(errors="replace")

    t = threading.Thread(target=body, daemon=True)
    t.start()
    while t.is_alive():
        t.join(0.1)
    if box.get("code"):
        raise PyPOHIPError(f"{what}: {box['msg']} (code {box['code']})")


_RESULT = {"JM": ("propagateToGrid_JM", "c2Bundle"), "EH": ("propagateToGrid_EH", "c2Bundle"),
           "JMEH": ("propagateToGrid_JMEH", "c4Bundle"), "EHP": ("propagateToGrid_EHP", "c2rBundle"),
           "FF": ("propagateToFarField", "c2Bundle"), "scalar": ("propagateToGrid_scalar", "arrC1")}


def PyPO_HIPd(source, target, runPODict, return_grids=False):
    """PO propagation on the GPU with FP64 inputs and outputs.

    Same call sequence and return types as BindCPU.PyPO_CPUd (BindCPU.py:82-220):
    JM -> currents, EH/FF -> fields, JMEH -> [currents, fields],
    EHP -> [fields, rfield], scalar -> scalarfield."""
    lib = loadHIPlib()
    mode = runPODict["mode"]
    if mode not in _RESULT:
        raise PyPOHIPError(f"{mode} is not a valid propagation mode")
    sym, bundle = _RESULT[mode]
    gs = int(np.prod(gridsize(source)))
    gt = int(np.prod(gridsize(target)))
    shape = gridsize(target)
    csp, ctp = MU.fill_reflparams(source), MU.fill_reflparams(target)
    cs, ct = MU.alloc_container(gs), MU.alloc_container(gt)
    exp = runPODict.get("exp", "fwd")
    t_dir = -1.0 if exp == "fwd" else 1.0
    if mode == "scalar":
        src = MU.fill_arrc1(runPODict["s_scalarfield"], gs)
    else:
        src = MU.fill_c2bundle(runPODict["s_current"], gs)
    res = MU.alloc_bundle(bundle, gt)
    prec = PRECISIONS[runPODict.get("precision", "mixed")]
    with _lock:
        old = lib.pypo_hip_set_precision(prec)
        try:
            _run(lib, getattr(lib, sym),
                 [ctypes.byref(res), csp, ctp, ctypes.byref(cs), ctypes.byref(ct), ctypes.byref(src),
                  float(runPODict["k"]), int(runPODict.get("nThreads", 256)),
                  float(runPODict.get("epsilon", 1.0)), t_dir], sym)
        finally:
            lib.pypo_hip_set_precision(old)
    if mode == "JM":
        out = MU.c2bundle_to_obj(res, shape, "currents")
    elif mode in ("EH", "FF"):
        out = MU.c2bundle_to_obj(res, shape, "fields")
    elif mode == "JMEH":
        out = list(MU.c4bundle_to_objs(res, shape))
    elif mode == "EHP":
        out = list(MU.c2rbundle_to_objs(res, shape))
    else:
        out = MU.arrc1_to_obj(res, shape)
    if return_grids:
        return out, MU.container_to_grids(cs, gridsize(source)), MU.container_to_grids(ct, shape)
    return out


def RT_HIPd(runRTDict, max_iter=None):
    """Ray trace on the GPU in FP64 (BindCPU.RT_CPUd, BindCPU.py:222-254)."""
    lib = loadHIPlib()
    fr = runRTDict["fr_in"]
    inp = MU.fill_cframe(fr)
    out = MU.alloc_cframe(fr.size)
    ctp = MU.fill_reflparams(runRTDict["t_name"])
    with _lock:
        old = None
        if max_iter is not None:
            old = lib.pypo_hip_set_rt_max_iter(int(max_iter))
        try:
            _run(lib, lib.propagateRays,
                 [ctp, ctypes.byref(inp), ctypes.byref(out), int(runRTDict.get("nThreads", 256)),
                  float(runRTDict.get("tol", 1e-3)), float(runRTDict.get("t0", 1.0))], "propagateRays")
        finally:
            if old is not None:
                lib.pypo_hip_set_rt_max_iter(old)
    return MU.cframe_to_frame(out)


def generate_grid(refl, transform=True, spheric=True):
    """Surface grid on the GPU (InterfaceReflector generateGrid) -> reflGrids."""
    lib = loadHIPlib()
    nx, ny = gridsize(refl)
    c = MU.alloc_container(nx * ny)
    with _lock:
        _run(lib, lib.generateGrid, [MU.fill_reflparams(refl), ctypes.byref(c), bool(transform),
                                     bool(spheric)], "generateGrid")
    return MU.container_to_grids(c, (nx, ny))


FF_KERNELS = {"matrix": 0, "vector": 1}


def set_ff_kernel(kind):
    """Kernel of the mixed-precision far-field sum: "matrix" (matrix cores,
    the default) or "vector" (vector ALUs).  Returns nothing; raises on an
    unknown name."""
    lib = loadHIPlib()
    if kind not in FF_KERNELS:
        raise ValueError(f"far-field kernel must be one of {sorted(FF_KERNELS)}")
    if lib.pypo_hip_set_ff_kernel(FF_KERNELS[kind]):
        check(lib, "pypo_hip_set_ff_kernel")
