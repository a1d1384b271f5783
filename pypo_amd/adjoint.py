"""Surface-shape gradients of the far-field power, as a torch.autograd.Function.

The reflector (dish) is perturbed the way the reference perturbs surfaces for
surface errors (src/common/InterfaceReflector.cpp:852-866): every grid point
moves along its unit normal by d[i]; normals and area elements stay.  The
forward pass is the reference's JM -> FF chain (System.runPO with mode "JM"
from the feed onto the dish, then mode "FF"), run on the GPU through
pypo_hip_displace_dev and pypo_hip_propagate_dev; the backward pass is one
call of pypo_hip_ff_power_grad_dev, which recomputes the induced currents
and their derivatives from the feed instead of keeping them from the forward
pass, and sums the adjoint far field back over the dish.

    prob = ShapeProblem(feed, feed_currents, dish, ff, k)
    d = torch.zeros(prob.n_dish, dtype=torch.float64, device="cuda", requires_grad=True)
    P = prob.power(d)                       # |E_ff|^2 on the far-field grid
    loss = ((P - target) ** 2).sum()
    loss.backward()                         # d.grad = dloss/dd (mm^-1 units of loss)

PyTorch is plumbing here (device memory, streams, the autograd tape); every
field computation is a libpypohip.so kernel.
"""

import ctypes

import numpy as np
import torch

from . import bind as B
from . import device as D
from .geometry import gridsize


def _dptr(t):
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), ctypes.POINTER(ctypes.c_double))


class ShapeProblem:
    """Feed currents -> JM on a displaceable dish -> FF.

    feed, dish, ff: surface dictionaries (geometry.py); feed_currents: a
    currents object on the feed grid or a (12, n_feed) CUDA tensor in
    c2Bundle row order; k in 1/mm.  precision: "mixed" (FP32 integrand, FP64
    distance and phase) or "fp64", used by forward and backward alike unless
    overridden per call."""

    def __init__(self, feed, feed_currents, dish, ff, k, exp="fwd", precision="mixed"):
        self.lib = B.loadHIPlib()
        if precision not in B.PRECISIONS:
            raise ValueError(f"unknown precision {precision!r}")
        self.feed, self.dish, self.ff = feed, dish, ff
        self.k = float(k)
        self.exp = exp
        self.tdir = -1.0 if exp == "fwd" else 1.0
        self.precision = precision
        self.g_feed = D.generate_grid(feed)
        self.g_dish0 = D.generate_grid(dish)
        self.g_ff = D.generate_grid(ff)
        if isinstance(feed_currents, torch.Tensor):
            self.c_feed = feed_currents
        else:
            self.c_feed = D.currents_tensor(feed_currents)
        self.n_feed = self.g_feed.shape[1]
        self.n_dish = self.g_dish0.shape[1]
        self.n_ff = self.g_ff.shape[1]
        self.dish_cells = (int(dish["gmode"]),) + tuple(int(v) for v in gridsize(dish))
        self.feed_cells = (int(feed["gmode"]),) + tuple(int(v) for v in gridsize(feed))
        # The backward sum over far-field targets does not depend on their
        # order.  Handed over row by row with every other row reversed, the
        # matrix kernel's groups of four consecutive targets never straddle a
        # jump back across the grid (a row wrap), so all of its tiles may take
        # the follower phase.
        nx, ny = (int(v) for v in gridsize(ff))
        self._tperm = None
        self._g_ff_rev = self.g_ff
        if nx * ny == self.n_ff and nx > 1 and ny > 1:
            idx = torch.arange(self.n_ff, device=self.g_ff.device).view(nx, ny)
            idx[1::2] = idx[1::2].flip(1)
            self._tperm = idx.reshape(-1)
            self._g_ff_rev = self.g_ff[:, self._tperm].contiguous()

    def _prec(self, precision):
        prec = precision or self.precision
        if prec not in B.PRECISIONS:
            raise ValueError(f"unknown precision {prec!r}")
        return prec

    def _check_d(self, d):
        if not (isinstance(d, torch.Tensor) and d.is_cuda and d.dtype == torch.float64
                and d.shape == (self.n_dish,)):
            raise ValueError(f"displacement: expected a ({self.n_dish},) float64 CUDA tensor")
        return d.detach().contiguous()

    def displaced(self, d):
        """(7, n_dish) dish grid moved by d along its normals."""
        d = self._check_d(d)
        out = torch.empty_like(self.g_dish0)
        cb, co = D.container(self.g_dish0), D.container(out)
        rc = self.lib.pypo_hip_displace_dev(ctypes.byref(cb), _dptr(d), ctypes.byref(co), D._stream())
        if rc:
            B.check(self.lib, "pypo_hip_displace_dev")
        return out

    def far_field(self, d, precision=None):
        """(12, n_ff) far field (c2Bundle rows) of the dish displaced by d.
        The dish currents are a temporary of this call."""
        prec = self._prec(precision)
        g = self.displaced(d)
        J = D.propagate("JM", self.g_feed, g, self.c_feed, self.feed, self.k, exp=self.exp, precision=prec)
        return D.propagate("FF", g, self.g_ff, J, self.dish, self.k, exp=self.exp, precision=prec)

    def power_grad(self, d, E, dl_dp, precision=None):
        """dL/dd (n_dish,) for dL/dP_t = dl_dp (n_ff,), E the far field of d."""
        prec = self._prec(precision)
        g = self.displaced(d)
        E = E.detach().contiguous()
        if E.shape != (12, self.n_ff) or not E.is_cuda or E.dtype != torch.float64:
            raise ValueError(f"E: expected a (12, {self.n_ff}) float64 CUDA tensor")
        gp = dl_dp.detach().to(torch.float64).contiguous()
        if gp.shape != (self.n_ff,) or not gp.is_cuda:
            raise ValueError(f"dL/dP: expected a ({self.n_ff},) CUDA tensor")
        grad = torch.empty(self.n_dish, dtype=torch.float64, device=g.device)
        if self._tperm is not None:
            E = E[:, self._tperm].contiguous()
            gp = gp[self._tperm].contiguous()
        cd, cf, cff = D.container(g), D.container(self.g_feed), D.container(self._g_ff_rev)
        bf, be = D.bundle(self.c_feed), D.bundle(E)
        gm, ncx, ncy = self.dish_cells
        fgm, fncx, fncy = self.feed_cells
        rc = self.lib.pypo_hip_ff_power_grad_dev(ctypes.byref(cd), gm, ncx, ncy, ctypes.byref(cf), fgm,
                                                 fncx, fncy, ctypes.byref(bf), ctypes.byref(cff),
                                                 ctypes.byref(be), _dptr(gp), self.k, self.tdir,
                                                 B.PRECISIONS[prec], _dptr(grad), D._stream())
        if rc:
            B.check(self.lib, "pypo_hip_ff_power_grad_dev")
        return grad

    def power(self, d):
        """|E_ff|^2 (n_ff,) as a differentiable function of d."""
        return FarFieldPower.apply(d, self)


class FarFieldPower(torch.autograd.Function):
    """d (n_dish,) -> P = |E_ff|^2 (n_ff,).  Saves d and the far field only;
    the backward pass recomputes the dish currents on the GPU."""

    @staticmethod
    def forward(ctx, d, prob):
        E = prob.far_field(d)
        ctx.prob = prob
        ctx.save_for_backward(d.detach(), E)
        return (E[0:6] ** 2).sum(0)

    @staticmethod
    def backward(ctx, dl_dp):
        d, E = ctx.saved_tensors
        return ctx.prob.power_grad(d, E, dl_dp), None


# ---- the surface-error workload ------------------------------------------------

def zernike_indices(n_modes):
    """(n, m) pairs, radial order n ascending and m = -n, -n+2, .., n within
    an order; orders 0..7 give 36 modes."""
    out = []
    n = 0
    while len(out) < n_modes:
        for m in range(-n, n + 1, 2):
            out.append((n, m))
        n += 1
    return out[:n_modes]


def zernike(n, m, rho, theta):
    """Orthonormal Zernike polynomial Z_n^m on the unit disk (cos for m >= 0,
    sin for m < 0; unit mean square over the disk)."""
    from math import factorial
    am = abs(m)
    R = np.zeros_like(rho)
    for s in range((n - am) // 2 + 1):
        c = (-1) ** s * factorial(n - s) / (factorial(s) * factorial((n + am) // 2 - s)
                                            * factorial((n - am) // 2 - s))
        R = R + c * rho ** (n - 2 * s)
    if m == 0:
        return np.sqrt(n + 1) * R
    ang = np.cos(am * theta) if m > 0 else np.sin(am * theta)
    return np.sqrt(2 * (n + 1)) * R * ang


def zernike_displacement(x, y, radius, n_modes=36, sigma=0.02, seed=0):
    """Seeded random surface: sum of n_modes Zernike modes over the aperture
    (x, y in mm, aperture radius in mm) with standard-normal coefficients,
    scaled to an RMS of sigma (mm) over the given points."""
    rng = np.random.default_rng(seed)
    rho = np.sqrt(x * x + y * y) / radius
    theta = np.arctan2(y, x)
    d = np.zeros_like(rho)
    for (n, m), c in zip(zernike_indices(n_modes), rng.standard_normal(n_modes)):
        d += c * zernike(n, m, rho, theta)
    return d * (sigma / np.sqrt(np.mean(d * d)))


def surface_error_workload(n_dish=601, n_ff=201, n_feed=33, diameter=3000.0, focal=1200.0, lam=1.0,
                           ff_deg=0.2, edge_db=-12.0, n_modes=36, sigma=0.02, seed=2024):
    """Differentiated far field: an on-axis paraboloid (uv grid n_dish^2)
    fed from its focus by a Gaussian Huygens source with edge_db taper, far
    field on an n_ff^2 AoE grid over +-ff_deg.  Returns (feed, feed currents,
    dish, ff, k, radius) -- the displacement is made on the generated grid
    with zernike_displacement."""
    from . import geometry as G
    from .sweep import feed_waist, huygens_feed
    w0 = feed_waist(focal / diameter, edge_db, lam)
    feed, cur = huygens_feed(n_feed, 1.6 * lam, w0, 0.0, focal)
    dish = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, focal],
                         "gmode": "uv", "lims_u": [0, diameter / 2], "lims_v": [0, 360],
                         "gridsize": [n_dish, n_dish]})
    ff = G.plane({"gmode": "AoE", "lims_Az": [-ff_deg, ff_deg], "lims_El": [-ff_deg, ff_deg],
                  "gridsize": [n_ff, n_ff]})
    return feed, cur, dish, ff, 2 * np.pi / lam, diameter / 2
