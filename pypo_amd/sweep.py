"""Batched JM -> FF sweeps on the GPU (pypo_hip_sweep_* in include/pypo_hip.h).

A sweep is a set of instances -- feed surface + feed currents, dish surface,
wavenumber -- of uneven grid sizes sharing one far-field grid.  Each sum
stage (JM on all dishes, FF of all dishes) is one launch over a static,
cost-sorted work list; a run is five kernels (JM sum, JM finish, pack, FF
sum, FF finish) that can be captured into one hipGraph.
The reference runs such sweeps as Python loops over System.runPO calls.
"""

import ctypes

import numpy as np

from . import bind as B
from . import marshal as MU
from . import structs as PS
from . import types as T
from .geometry import gridsize

C_MM_GHZ = 299.792458      # speed of light in mm * GHz


class SweepDesc(ctypes.Structure):
    _fields_ = [("n_inst", ctypes.c_int),
                ("feeds", ctypes.POINTER(PS.reflparams)),
                ("dishes", ctypes.POINTER(PS.reflparams)),
                ("k", ctypes.POINTER(ctypes.c_double)),
                ("feed_currents", ctypes.POINTER(PS.c2Bundle)),
                ("ff", PS.reflparams),
                ("t_direction", ctypes.c_double),
                ("tpl", ctypes.c_int),
                ("chunk_tiles", ctypes.c_int),
                ("wgs_per_cu", ctypes.c_int),
                ("use_graph", ctypes.c_int)]


def _declare(lib):
    if getattr(lib, "_sweep_declared", False):
        return
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.pypo_hip_sweep_create.argtypes = [P(SweepDesc), P(vp)]
    lib.pypo_hip_sweep_plan.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.pypo_hip_sweep_run.argtypes = [vp, vp]
    lib.pypo_hip_sweep_read.argtypes = [vp, ctypes.c_int, P(PS.c2Bundle)]
    lib.pypo_hip_sweep_read_currents.argtypes = [vp, ctypes.c_int, P(PS.c2Bundle)]
    lib.pypo_hip_sweep_info.argtypes = [vp, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_int),
                                        P(ctypes.c_int), P(P(ctypes.c_double))]
    lib.pypo_hip_sweep_destroy.argtypes = [vp]
    for f in ("create", "plan", "run", "read", "read_currents", "info", "destroy"):
        getattr(lib, f"pypo_hip_sweep_{f}").restype = ctypes.c_int
    lib._sweep_declared = True


class Sweep:
    """One batched sweep.  feeds / dishes: lists of surface dictionaries
    (geometry.py), ks: wavenumbers (1/mm), feed_currents: currents objects on
    the feed grids, ff: the shared far-field surface dictionary."""

    def __init__(self, feeds, dishes, ks, feed_currents, ff, exp="fwd", tpl=0, chunk_tiles=0,
                 wgs_per_cu=0, use_graph=True):
        self.lib = B.loadHIPlib()
        _declare(self.lib)
        n = len(feeds)
        if not (len(dishes) == len(ks) == len(feed_currents) == n and n > 0):
            raise ValueError("feeds, dishes, ks and feed_currents must have the same non-zero length")
        self.n_inst = n
        self.ff_shape = gridsize(ff)
        self.dish_shapes = [gridsize(d) for d in dishes]
        fp = (PS.reflparams * n)()
        dp = (PS.reflparams * n)()
        fc = (PS.c2Bundle * n)()
        keep = []
        for i in range(n):
            a, b = MU.fill_reflparams(feeds[i]), MU.fill_reflparams(dishes[i])
            c = MU.fill_c2bundle(feed_currents[i], int(np.prod(gridsize(feeds[i]))))
            keep += [a, b, c]
            fp[i], dp[i], fc[i] = a, b, c
        kk = (ctypes.c_double * n)(*[float(k) for k in ks])
        ffp = MU.fill_reflparams(ff)
        desc = SweepDesc(n, fp, dp, kk, fc, ffp, -1.0 if exp == "fwd" else 1.0, int(tpl),
                         int(chunk_tiles), int(wgs_per_cu), int(bool(use_graph)))
        h = ctypes.c_void_p()
        rc = self.lib.pypo_hip_sweep_create(ctypes.byref(desc), ctypes.byref(h))
        if rc:
            B.check(self.lib, "pypo_hip_sweep_create")
        self.handle = h
        self._keep = (keep, ffp, desc)

    def plan(self, tpl=0, chunk_tiles=0, wgs_per_cu=0, use_graph=True):
        rc = self.lib.pypo_hip_sweep_plan(self.handle, int(tpl), int(chunk_tiles), int(wgs_per_cu),
                                          int(bool(use_graph)))
        if rc:
            B.check(self.lib, "pypo_hip_sweep_plan")

    def run(self, stream=None):
        """Enqueue one full JM -> FF run (asynchronous on `stream`; default stream if None)."""
        s = ctypes.c_void_p(stream) if stream else None
        rc = self.lib.pypo_hip_sweep_run(self.handle, s)
        if rc:
            B.check(self.lib, "pypo_hip_sweep_run")

    def info(self):
        pj, pf = ctypes.c_double(), ctypes.c_double()
        ij, i_f = ctypes.c_int(), ctypes.c_int()
        ptr = ctypes.POINTER(ctypes.c_double)()
        self.lib.pypo_hip_sweep_info(self.handle, ctypes.byref(pj), ctypes.byref(pf), ctypes.byref(ij),
                                     ctypes.byref(i_f), ctypes.byref(ptr))
        return {"pairs_jm": pj.value, "pairs_ff": pf.value, "items_jm": ij.value,
                "items_ff": i_f.value}

    def read(self, inst):
        """Far field of instance `inst` -> fields (synchronous)."""
        n = int(np.prod(self.ff_shape))
        b = MU.alloc_bundle("c2Bundle", n)
        rc = self.lib.pypo_hip_sweep_read(self.handle, int(inst), ctypes.byref(b))
        if rc:
            B.check(self.lib, "pypo_hip_sweep_read")
        return MU.c2bundle_to_obj(b, self.ff_shape, "fields")

    def read_currents(self, inst):
        """PEC currents of instance `inst` on its dish -> currents (synchronous)."""
        shape = self.dish_shapes[inst]
        b = MU.alloc_bundle("c2Bundle", int(np.prod(shape)))
        rc = self.lib.pypo_hip_sweep_read_currents(self.handle, int(inst), ctypes.byref(b))
        if rc:
            B.check(self.lib, "pypo_hip_sweep_read_currents")
        return MU.c2bundle_to_obj(b, shape, "currents")

    def close(self):
        if getattr(self, "handle", None):
            self.lib.pypo_hip_sweep_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def feed_waist(f_over_d, edge_db, lam):
    """Waist of a Gaussian aperture giving `edge_db` at the rim of a dish of
    focal ratio f/D: far-field amplitude exp(-(k w0 sin t / 2)^2)(1 + cos t)/2
    times the focal-distance space loss (1 + cos t)/2."""
    te = 2 * np.arctan(1 / (4 * f_over_d))
    k = 2 * np.pi / lam
    target = 10 ** (edge_db / 20)
    lo, hi = 1e-4 * lam, 10 * lam
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        e = np.exp(-(k * mid * np.sin(te) / 2) ** 2) * (1 + np.cos(te)) / 2 * (1 + np.cos(te)) / 2
        lo, hi = (mid, hi) if e > target else (lo, mid)
    return 0.5 * (lo + hi)


def huygens_feed(n, half, w0, offset, z_focus):
    """Feed plane (n x n over +-half mm) at the focus, shifted laterally by
    `offset` mm in x, facing -z, with a Gaussian Huygens source
    (E || x travelling -z: J = -x, M = +y)."""
    from . import geometry as G
    feed = G.plane({"gmode": "xy", "lims_x": [-half, half], "lims_y": [-half, half],
                    "gridsize": [n, n], "flip": True})
    feed["transf"] = G.mat_translate([offset, 0.0, z_focus])
    x = np.linspace(-half, half, n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    amp = np.exp(-(X ** 2 + Y ** 2) / w0 ** 2) + 0j
    z = np.zeros_like(amp)
    return feed, T.currents(-amp, z, z, z, amp, z)


def frequency_offset_sweep(n_freq=64, f_lo=80.0, f_hi=120.0, offsets=(0.0, 2.0, -2.0, 4.0),
                           diameter=1000.0, focal=400.0, n_lo=201, n_hi=401, n_feed=25,
                           n_ff=181, ff_deg=1.0, edge_db=-12.0):
    """Instances of the batched-sweep workload: n_freq frequencies in
    [f_lo, f_hi] GHz x lateral feed offsets (mm); the dish uv grid grows with
    frequency, n = rint(n_lo + (n_hi - n_lo)(f - f_lo)/(f_hi - f_lo)); every
    instance goes JM -> FF onto an n_ff x n_ff AoE grid over +-ff_deg.
    Returns (feeds, dishes, ks, currents, ff, meta)."""
    from . import geometry as G
    freqs = np.linspace(f_lo, f_hi, n_freq)
    feeds, dishes, ks, curs, meta = [], [], [], [], []
    for f in freqs:
        lam = C_MM_GHZ / f
        n = int(np.rint(n_lo + (n_hi - n_lo) * (f - f_lo) / (f_hi - f_lo)))
        w0 = feed_waist(focal / diameter, edge_db, lam)
        dish = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, focal],
                             "gmode": "uv", "lims_u": [0, diameter / 2], "lims_v": [0, 360],
                             "gridsize": [n, n]})
        for off in offsets:
            feed, cur = huygens_feed(n_feed, 1.6 * lam, w0, off, focal)
            feeds.append(feed)
            dishes.append(dish)
            ks.append(2 * np.pi / lam)
            curs.append(cur)
            meta.append({"f_ghz": float(f), "offset_mm": float(off), "n_dish": n})
    ff = G.plane({"gmode": "AoE", "lims_Az": [-ff_deg, ff_deg], "lims_El": [-ff_deg, ff_deg],
                  "gridsize": [n_ff, n_ff]})
    return feeds, dishes, ks, curs, ff, meta
