"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU checkers.

  kind="port" -> oracle/liboracle.so, this project's FP64 restatement of the
                 reference CPU path (pypo_oracle.cpp), symbols "oracle_*".
  kind="ref"  -> oracle/_ref/libpypocpu_ref.so, the reference's own C++
                 sources (src/CPU/InterfaceCPU.cpp & co.) compiled by
                 oracle/Makefile; present only where it was built.

Used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
The call sequence mirrors src/PyPO/BindCPU.py:82-254 (PyPO_CPUd, RT_CPUd).
"""

import ctypes
import os

import numpy as np

from pypo_amd import marshal as MU
from pypo_amd import structs as PS
from pypo_amd.geometry import gridsize

_HERE = os.path.dirname(os.path.abspath(__file__))
PATHS = {
    "port": os.path.join(_HERE, "liboracle.so"),
    "ref": os.path.join(_HERE, "_ref", "libpypocpu_ref.so"),
}
_PREFIX = {"port": "oracle_", "ref": ""}
_cache = {}

_MODE_SYMBOL = {
    "JM": ("propagateToGrid_JM", "c2Bundle"),
    "EH": ("propagateToGrid_EH", "c2Bundle"),
    "JMEH": ("propagateToGrid_JMEH", "c4Bundle"),
    "EHP": ("propagateToGrid_EHP", "c2rBundle"),
    "FF": ("propagateToFarField", "c2Bundle"),
    "scalar": ("propagateToGrid_scalar", "arrC1"),
}


def available(kind):
    return os.path.exists(PATHS[kind])


class Lib:
    def __init__(self, kind):
        if not available(kind):
            raise OSError(f"{PATHS[kind]} is not built (make -C oracle{' ref' if kind == 'ref' else ''})")
        self.kind = kind
        self.pre = _PREFIX[kind]
        self.cdll = ctypes.CDLL(PATHS[kind])
        P = ctypes.POINTER
        for sym, bundle in _MODE_SYMBOL.values():
            f = getattr(self.cdll, self.pre + sym)
            src_t = PS.arrC1 if bundle == "arrC1" else PS.c2Bundle
            f.argtypes = [P(getattr(PS, bundle)), PS.reflparams, PS.reflparams, P(PS.reflcontainer),
                          P(PS.reflcontainer), P(src_t), ctypes.c_double, ctypes.c_int,
                          ctypes.c_double, ctypes.c_double]
            f.restype = None
        rays = getattr(self.cdll, self.pre + "propagateRays")
        rays.argtypes = [PS.reflparams, P(PS.cframe), P(PS.cframe), ctypes.c_int, ctypes.c_double,
                         ctypes.c_double] + ([ctypes.c_int] if kind == "port" else [])
        rays.restype = None
        gg = getattr(self.cdll, self.pre + "generateGrid")
        gg.argtypes = [PS.reflparams, P(PS.reflcontainer), ctypes.c_bool, ctypes.c_bool]
        gg.restype = None
        if kind == "port":
            fs = self.cdll.oracle_farfield_subset
            fs.argtypes = [P(PS.c2Bundle), PS.reflparams, P(PS.reflcontainer), P(PS.reflcontainer),
                           P(PS.c2Bundle), P(ctypes.c_int), ctypes.c_int, ctypes.c_double,
                           ctypes.c_int, ctypes.c_double, ctypes.c_double]
            fs.restype = None

    def fn(self, name):
        return getattr(self.cdll, self.pre + name)


def load(kind="port"):
    if kind not in _cache:
        _cache[kind] = Lib(kind)
    return _cache[kind]


def default_threads():
    """Host threads for the checker: the CPU share of one GPU (16), or fewer
    when the machine has fewer cores."""
    return max(1, min(16, os.cpu_count() or 1))


def generate_grid(lib, refl, transform=True, spheric=True):
    nx, ny = gridsize(refl)
    c = MU.alloc_container(nx * ny)
    lib.fn("generateGrid")(MU.fill_reflparams(refl), ctypes.byref(c), transform, spheric)
    return MU.container_to_grids(c, (nx, ny))


def propagate(lib, mode, source, target, src_obj, k, nthreads=None, epsilon=1.0, exp="fwd"):
    """One PO propagation through the checker library (BindCPU.PyPO_CPUd).
    Returns (result, source_grids, target_grids); result types as PyPO_CPUd."""
    nthreads = nthreads or default_threads()
    sym, bundle = _MODE_SYMBOL[mode]
    gs = int(np.prod(gridsize(source)))
    gt = int(np.prod(gridsize(target)))
    shape = gridsize(target)
    csp, ctp = MU.fill_reflparams(source), MU.fill_reflparams(target)
    cs, ct = MU.alloc_container(gs), MU.alloc_container(gt)
    tdir = -1.0 if exp == "fwd" else 1.0
    src = MU.fill_arrc1(src_obj, gs) if mode == "scalar" else MU.fill_c2bundle(src_obj, gs)
    res = MU.alloc_bundle(bundle, gt)
    lib.fn(sym)(ctypes.byref(res), csp, ctp, ctypes.byref(cs), ctypes.byref(ct), ctypes.byref(src),
                float(k), int(nthreads), float(epsilon), tdir)
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
    return out, MU.container_to_grids(cs, gridsize(source)), MU.container_to_grids(ct, shape)


def trace(lib, target, fr, nthreads=None, tol=1e-3, t0=1.0, max_iter=0):
    """Ray trace through the checker library (BindCPU.RT_CPUd).  max_iter is
    only honoured by the port (0 = reference behaviour, no cap)."""
    nthreads = nthreads or default_threads()
    inp = MU.fill_cframe(fr)
    out = MU.alloc_cframe(fr.size)
    args = [MU.fill_reflparams(target), ctypes.byref(inp), ctypes.byref(out), int(nthreads),
            float(tol), float(t0)]
    if lib.kind == "port":
        args.append(int(max_iter))
    lib.fn("propagateRays")(*args)
    return MU.cframe_to_frame(out)


def farfield_subset(lib, source, target, currents, k, idx, nthreads=None, epsilon=1.0, exp="fwd"):
    """FF on targets idx only (port only).  Grids are generated here, outside
    the timed call.  Returns (6, len(idx)) complex array and the seconds the
    native call took."""
    import time
    nthreads = nthreads or default_threads()
    gs = int(np.prod(gridsize(source)))
    gt = int(np.prod(gridsize(target)))
    csp, ctp = MU.fill_reflparams(source), MU.fill_reflparams(target)
    cs, ct = MU.alloc_container(gs), MU.alloc_container(gt)
    lib.fn("generateGrid")(csp, ctypes.byref(cs), True, True)
    lib.fn("generateGrid")(ctp, ctypes.byref(ct), True, True)
    idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int32))
    res = MU.alloc_bundle("c2Bundle", idx.size)
    cur = MU.fill_c2bundle(currents, gs)
    tdir = -1.0 if exp == "fwd" else 1.0
    t = time.perf_counter()
    lib.cdll.oracle_farfield_subset(ctypes.byref(res), csp, ctypes.byref(cs), ctypes.byref(ct),
                                    ctypes.byref(cur), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                    int(idx.size), float(k), int(nthreads), float(epsilon), tdir)
    dt = time.perf_counter() - t
    o = MU.c2bundle_to_obj(res, (idx.size,), "fields")
    return np.stack(o.components()), dt


def _container_from_rows(rows):
    """reflcontainer over a (7, n) float64 array (rows x y z nx ny nz area)."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    c = PS.reflcontainer()
    c.size = rows.shape[1]
    for i, name in enumerate(("x", "y", "z", "nx", "ny", "nz", "area")):
        setattr(c, name, rows[i].ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    c._keep = rows
    return c


def _bundle_from_rows(rows, struct=PS.c2Bundle):
    """Bundle (c2Bundle by default) over a float64 array whose rows follow the
    struct's field order."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    b = struct()
    for i, (name, _) in enumerate(b._fields_):
        setattr(b, name, rows[i].ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    b._keep = rows
    return b


def propagate_on_grids(lib, mode, src_rows, cells, tgt_rows, cur_rows, k, nthreads=None, epsilon=1.0,
                       exp="fwd"):
    """JM / EH / JMEH / EHP / FF from (7, n_s) source rows onto (7, n_t)
    target rows that the caller generated (port only; no grid generation
    inside).  cells = (gmode, ncx, ncy) of the source grid; cur_rows
    (12, n_s).  Returns the result rows in the field order of the mode's
    bundle: (12, n_t) c2Bundle, (24, n_t) c4Bundle for JMEH, (15, n_t)
    c2rBundle for EHP."""
    modes = {"JM": 0, "EH": 1, "JMEH": 2, "EHP": 3, "FF": 4}
    structs = {"JMEH": PS.c4Bundle, "EHP": PS.c2rBundle}
    nthreads = nthreads or default_threads()
    cs, ct = _container_from_rows(src_rows), _container_from_rows(tgt_rows)
    cur = _bundle_from_rows(cur_rows)
    struct = structs.get(mode, PS.c2Bundle)
    out = np.zeros((len(struct._fields_), ct.size))
    res = _bundle_from_rows(out, struct)
    f = lib.cdll.oracle_propagate_on_grids
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                  ctypes.POINTER(PS.reflcontainer), ctypes.POINTER(PS.reflcontainer),
                  ctypes.POINTER(PS.c2Bundle), ctypes.c_double, ctypes.c_int, ctypes.c_double,
                  ctypes.c_double]
    f.restype = None
    gmode, ncx, ncy = cells
    f(modes[mode], ctypes.cast(ctypes.byref(res), ctypes.c_void_p), int(gmode), int(ncx), int(ncy),
      ctypes.byref(cs), ctypes.byref(ct), ctypes.byref(cur), float(k), int(nthreads), float(epsilon),
      -1.0 if exp == "fwd" else 1.0)
    return res._keep


# ---- beam initialisation (InterfaceBeam.h:21-34) -------------------------------------

def _beam_fn(lib, name, argtypes):
    f = getattr(lib.cdll, lib.pre + name)
    f.argtypes = argtypes
    f.restype = None
    return f


def make_gauss(lib, gdict, source, vector=False):
    """makeGauss (vector=False) or makeGaussBeam (vector=True) through the
    checker library -> (fields, currents)."""
    from pypo_amd import beam as PB
    P = ctypes.POINTER
    shape = gridsize(source)
    n = int(np.prod(shape))
    f, c = MU.alloc_bundle("c2Bundle", n), MU.alloc_bundle("c2Bundle", n)
    if vector:
        fn = _beam_fn(lib, "makeGaussBeam", [PS.vecGPODict, PS.reflparams, P(PS.c2Bundle), P(PS.c2Bundle)])
        fn(PB.fill_vecGPODict(gdict), MU.fill_reflparams(source), ctypes.byref(f), ctypes.byref(c))
    else:
        fn = _beam_fn(lib, "makeGauss", [PS.GPODict, PS.reflparams, P(PS.c2Bundle), P(PS.c2Bundle)])
        g = PB.fill_GPODict(gdict)
        fn(g, MU.fill_reflparams(source), ctypes.byref(f), ctypes.byref(c))
    return MU.c2bundle_to_obj(f, shape, "fields"), MU.c2bundle_to_obj(c, shape, "currents")


def make_scalar_gauss(lib, gdict, source):
    from pypo_amd import beam as PB
    shape = gridsize(source)
    f = MU.alloc_bundle("arrC1", int(np.prod(shape)))
    fn = _beam_fn(lib, "makeScalarGauss", [PS.ScalarGPODict, PS.reflparams, ctypes.POINTER(PS.arrC1)])
    fn(PB.fill_SGPODict(gdict), MU.fill_reflparams(source), ctypes.byref(f))
    return MU.arrc1_to_obj(f, shape)


def calc_currents(lib, fields, source, mode):
    shape = gridsize(source)
    n = int(np.prod(shape))
    fin = MU.fill_c2bundle(fields, n)
    c = MU.alloc_bundle("c2Bundle", n)
    P = ctypes.POINTER
    fn = _beam_fn(lib, "calcCurrents", [P(PS.c2Bundle), P(PS.c2Bundle), PS.reflparams, ctypes.c_int])
    fn(ctypes.byref(fin), ctypes.byref(c), MU.fill_reflparams(source), int(mode))
    return MU.c2bundle_to_obj(c, shape, "currents")


def make_rt_frame(lib, rdict):
    from pypo_amd import beam as PB
    n = 1 + int(rdict["nRays"]) * 4 * int(rdict["nRing"])
    res = MU.alloc_cframe(n)
    fn = _beam_fn(lib, "makeRTframe", [PS.RTDict, ctypes.POINTER(PS.cframe)])
    fn(PB.fill_RTDict(rdict), ctypes.byref(res))
    return MU.cframe_to_frame(res)


def make_grt_frame_ref(rdict):
    """makeGRTframe of the compiled reference (random: it seeds from
    std::random_device whatever the seed field says)."""
    from pypo_amd import beam as PB
    lib = load("ref")
    res = MU.alloc_cframe(int(rdict["nRays"]))
    fn = _beam_fn(lib, "makeGRTframe", [PS.GRTDict, ctypes.POINTER(PS.cframe)])
    fn(PB.fill_GRTDict(rdict), ctypes.byref(res))
    return MU.cframe_to_frame(res)
