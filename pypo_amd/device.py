"""Device-resident use of libpypohip.so with torch tensors as HBM buffers.

PyTorch is plumbing here: it owns device memory and the stream; every
computation is a libpypohip.so kernel launched on torch's current stream
through the pypo_hip_*_dev entry points.  Layouts:
  grid      (7, n) float64  rows x, y, z, nx, ny, nz, area
  currents  (12, n) float64 rows r1x r1y r1z i1x i1y i1z r2x .. i2z (c2Bundle order)
  scalar    (2, n) float64  rows re, im (arrC1)
  frame     (6, n) float64  rows x, y, z, dx, dy, dz
JM -> (12, n_t), EH/FF -> (12, n_t), JMEH -> (24, n_t), EHP -> (15, n_t),
scalar -> (2, n_t), all in the reference's struct order.
"""

import ctypes

import numpy as np
import torch

from . import bind as B
from . import marshal as MU
from . import structs as PS
from .geometry import gridsize

N_OUT = {"JM": 12, "EH": 12, "FF": 12, "JMEH": 24, "EHP": 15, "scalar": 2}
_BUNDLE = {12: PS.c2Bundle, 24: PS.c4Bundle, 15: PS.c2rBundle, 2: PS.arrC1}
_dp = ctypes.POINTER(ctypes.c_double)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _row_ptr(t, i):
    return ctypes.cast(ctypes.c_void_p(t[i].data_ptr()), _dp)


def _check_dev(t, rows, name):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.dim() == 2
            and t.shape[0] == rows and t.is_contiguous()):
        raise ValueError(f"{name}: expected a contiguous ({rows}, n) float64 CUDA tensor")


def container(grid):
    """reflcontainer struct whose arrays are the rows of a (7, n) tensor."""
    _check_dev(grid, 7, "grid")
    c = PS.reflcontainer()
    c.size = grid.shape[1]
    for i, name in enumerate(("x", "y", "z", "nx", "ny", "nz", "area")):
        setattr(c, name, _row_ptr(grid, i))
    c._keep = grid
    return c


def bundle(t):
    """c2Bundle / c4Bundle / c2rBundle / arrC1 struct over the rows of t."""
    rows = t.shape[0]
    b = _BUNDLE[rows]()
    _check_dev(t, rows, "bundle")
    for i, (name, _) in enumerate(b._fields_):
        setattr(b, name, _row_ptr(t, i))
    b._keep = t
    return b


def frame_struct(t):
    _check_dev(t, 6, "frame")
    c = PS.cframe()
    c.size = t.shape[1]
    for i, name in enumerate(("x", "y", "z", "dx", "dy", "dz")):
        setattr(c, name, _row_ptr(t, i))
    c._keep = t
    return c


def generate_grid(refl, device="cuda"):
    """(7, n) tensor with the surface grid generated on the GPU."""
    lib = B.loadHIPlib()
    nx, ny = gridsize(refl)
    g = torch.empty((7, nx * ny), dtype=torch.float64, device=device)
    c = container(g)
    rc = lib.pypo_hip_generate_grid_dev(MU.fill_reflparams(refl), ctypes.byref(c), _stream())
    if rc:
        B.check(lib, "pypo_hip_generate_grid_dev")
    return g


def propagate(mode, src_grid, tgt_grid, src_values, src_refl, k, epsilon=1.0, exp="fwd",
              precision="mixed", out=None, no_alloc=False):
    """One PO propagation with every buffer in HBM; returns the result tensor.
    no_alloc=True runs pypo_hip_propagate_noalloc, which allocates no device
    memory (its workspace is the library's fixed device scratch; `out` should
    then be given, or this function allocates it with torch)."""
    lib = B.loadHIPlib()
    if mode not in N_OUT:
        raise ValueError(f"{mode} is not a valid propagation mode")
    n_t = tgt_grid.shape[1]
    if out is None:
        out = torch.empty((N_OUT[mode], n_t), dtype=torch.float64, device=tgt_grid.device)
    _check_dev(out, N_OUT[mode], "out")
    rows = 2 if mode == "scalar" else 12
    _check_dev(src_values, rows, "source values")
    ncx, ncy = gridsize(src_refl)
    cs, ct = container(src_grid), container(tgt_grid)
    src_b, res_b = bundle(src_values), bundle(out)
    t_dir = -1.0 if exp == "fwd" else 1.0
    fn = lib.pypo_hip_propagate_noalloc if no_alloc else lib.pypo_hip_propagate_dev
    rc = fn(B.MODES[mode], ctypes.cast(ctypes.byref(res_b), ctypes.c_void_p), ctypes.byref(cs), ctypes.byref(ct),
            ctypes.cast(ctypes.byref(src_b), ctypes.c_void_p), int(src_refl["gmode"]), ncx, ncy, float(k),
            float(epsilon), t_dir, B.PRECISIONS[precision], _stream())
    if rc:
        B.check(lib, fn.__name__)
    return out


def trace(surfaces, frame, tol=1e-3, t0=1.0):
    """Trace a (6, n) frame through up to 4 surfaces in one launch; returns
    the list of (6, n) frames after each surface."""
    lib = B.loadHIPlib()
    _check_dev(frame, 6, "frame")
    outs = [torch.empty_like(frame) for _ in surfaces]
    frames = (PS.cframe * (len(surfaces) + 1))()
    keep = [frame_struct(frame)] + [frame_struct(o) for o in outs]
    for i, f in enumerate(keep):
        frames[i] = f
    params = (PS.reflparams * len(surfaces))()
    pkeep = [MU.fill_reflparams(s) for s in surfaces]
    for i, p in enumerate(pkeep):
        params[i] = p
    rc = lib.pypo_hip_trace_dev(params, len(surfaces), frames, float(tol), float(t0), _stream())
    if rc:
        B.check(lib, "pypo_hip_trace_dev")
    return outs


def to_device(arr, device="cuda"):
    return torch.as_tensor(np.ascontiguousarray(arr, dtype=np.float64), device=device)


def currents_tensor(cur, device="cuda"):
    """currents/fields object -> (12, n) tensor in c2Bundle order.  A result whose
    values still live on the device hands over its own rows (no copy; callers
    treat them as read-only)."""
    rows = cur.device_rows() if hasattr(cur, "device_rows") else None
    if rows is not None:
        return rows
    comps = [np.asarray(c).ravel() for c in cur.components()]
    rows = [np.real(comps[0]), np.real(comps[1]), np.real(comps[2]),
            np.imag(comps[0]), np.imag(comps[1]), np.imag(comps[2]),
            np.real(comps[3]), np.real(comps[4]), np.real(comps[5]),
            np.imag(comps[3]), np.imag(comps[4]), np.imag(comps[5])]
    return to_device(np.stack(rows), device)


def tensor_to_complex6(t):
    """(12, n) c2Bundle-order tensor -> (6, n) complex128 numpy array."""
    a = t.detach().cpu().numpy()
    return np.stack([a[0] + 1j * a[3], a[1] + 1j * a[4], a[2] + 1j * a[5],
                     a[6] + 1j * a[9], a[7] + 1j * a[10], a[8] + 1j * a[11]])


def _declare_hybrid(lib):
    if getattr(lib, "_hybrid_declared", False):
        return
    P = ctypes.POINTER
    lib.pypo_hip_hybrid_dev.argtypes = [PS.reflparams, P(PS.cframe), P(PS.cframe), P(PS.c2Bundle),
                                        P(PS.c2Bundle), P(ctypes.c_double), ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_void_p]
    lib.pypo_hip_hybrid_dev.restype = ctypes.c_int
    lib.pypo_hip_interp_nearest_dev.argtypes = [P(PS.cframe), P(PS.c2Bundle), P(PS.reflcontainer),
                                                P(PS.c2Bundle), ctypes.c_void_p]
    lib.pypo_hip_interp_nearest_dev.restype = ctypes.c_int
    lib._hybrid_declared = True


def hybrid(target, frame, field, k, start=None, tol=1e-3, t0=1.0):
    """Hybrid ray-trace / PO step: trace the (6, n) frame onto `target` and
    carry the (12, n) field samples along the rays (exp(j k L), with the
    start-point amplitude factor when `start` is given).  Returns the traced
    frame and the carried field, both on the device."""
    lib = B.loadHIPlib()
    _declare_hybrid(lib)
    _check_dev(frame, 6, "frame")
    _check_dev(field, 12, "field")
    if field.shape[1] != frame.shape[1]:
        raise ValueError("frame and field must have the same number of points")
    fr_out = torch.empty_like(frame)
    f_out = torch.empty_like(field)
    fi, fo = frame_struct(frame), frame_struct(fr_out)
    bi, bo = bundle(field), bundle(f_out)
    st = None
    if start is not None:
        st = (ctypes.c_double * 3)(*[float(v) for v in np.asarray(start, dtype=np.float64).ravel()[:3]])
    rc = lib.pypo_hip_hybrid_dev(MU.fill_reflparams(target), ctypes.byref(fi), ctypes.byref(fo),
                                 ctypes.byref(bi), ctypes.byref(bo), st, float(k), float(tol), float(t0),
                                 _stream())
    if rc:
        B.check(lib, "pypo_hip_hybrid_dev")
    return fr_out, f_out


def interp_nearest(frame, values, grid):
    """(12, m) values at the (7, m) grid's points: each takes the values of the
    nearest point of the (6, n) frame (ties: smallest point index)."""
    lib = B.loadHIPlib()
    _declare_hybrid(lib)
    _check_dev(frame, 6, "frame")
    _check_dev(values, 12, "values")
    if values.shape[1] != frame.shape[1]:
        raise ValueError("frame and values must have the same number of points")
    out = torch.empty((12, grid.shape[1]), dtype=torch.float64, device=grid.device)
    fs, bv, g, bo = frame_struct(frame), bundle(values), container(grid), bundle(out)
    rc = lib.pypo_hip_interp_nearest_dev(ctypes.byref(fs), ctypes.byref(bv), ctypes.byref(g),
                                         ctypes.byref(bo), _stream())
    if rc:
        B.check(lib, "pypo_hip_interp_nearest_dev")
    return out


def frame_tensor(fr, device="cuda"):
    """frame object -> (6, n) tensor (the frame's own device rows when it has them)."""
    rows = fr.device_rows() if hasattr(fr, "device_rows") else None
    if rows is not None:
        return rows
    return to_device(np.stack([np.ravel(getattr(fr, a)) for a in ("x", "y", "z", "dx", "dy", "dz")]), device)


def scalar_tensor(sf, device="cuda"):
    """scalarfield object -> (2, n) tensor rows Re, Im (its own device rows when it has them)."""
    rows = sf.device_rows() if hasattr(sf, "device_rows") else None
    if rows is not None:
        return rows
    s = np.ravel(np.asarray(sf.S))
    return to_device(np.stack([np.real(s), np.imag(s)]), device)


def grid_key(refl):
    """Bytes that identify the grid a reflDict generates: every reflparams value
    the generator reads (coefficients, limits, gridsize, flip, gmode, centre,
    uv ellipticity and rotation, type, transformation, roughness and seed)."""
    p = MU.fill_reflparams(refl)
    k = p._keep
    parts = [k[n].tobytes() for n in ("coeffs", "lxu", "lyv", "gcenter", "transf", "n_cells")]
    parts.append(np.array([p.flip, p.gmode, p.type, p.rms_seed], dtype=np.int64).tobytes())
    parts.append(np.array([p.ecc_uv, p.rot_uv, p.rms], dtype=np.float64).tobytes())
    return b"".join(parts)


class GridCache:
    """Device grids of surfaces, kept between propagations and regenerated when
    a surface's parameters change (keyed by grid_key); at most `max_entries`
    grids are held, least recently used first out."""

    def __init__(self, max_entries=8):
        import collections
        self.max_entries = max_entries
        self._grids = collections.OrderedDict()

    def get(self, refl):
        key = grid_key(refl)
        g = self._grids.get(key)
        if g is None:
            g = generate_grid(refl)
            self._grids[key] = g
            while len(self._grids) > self.max_entries:
                self._grids.popitem(last=False)
        else:
            self._grids.move_to_end(key)
        return g

    def clear(self):
        self._grids.clear()


def scratch_bytes(stream=None):
    """Device bytes of scratch the library holds for calls on `stream` (a
    torch.cuda.Stream; None: the current stream)."""
    lib = B.loadHIPlib()
    st = ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)
    n = lib.pypo_hip_scratch_bytes(st)
    if n < 0:
        B.check(lib, "pypo_hip_scratch_bytes")
    return int(n)


def release_scratch(stream=None):
    """Free the scratch the library holds for calls on `stream` (None: the
    current stream), e.g. before dropping a short-lived stream.  Waits for the
    stream's last library call; graphs captured from calls on the stream must
    not be launched afterwards."""
    lib = B.loadHIPlib()
    st = ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)
    if lib.pypo_hip_release_scratch(st):
        B.check(lib, "pypo_hip_release_scratch")


def tensor_to_frame(t):
    from . import types as T
    a = t.detach().cpu().numpy()
    return T.frame(a.shape[1], *[a[i].copy() for i in range(6)])
