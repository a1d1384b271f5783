"""Pins the CPU checker of src/PyPO/Efficiencies.py (oracle/efficiencies.py).

Closed forms: a Gaussian exp(-r^2/w^2) on a disc of radius R with a circular
aperture of radius a has spillover (1 - e^{-2a^2/w^2}) / (1 - e^{-2R^2/w^2})
and taper 2 w^2 (1 - e^{-a^2/w^2})^2 / (a^2 (1 - e^{-2a^2/w^2})); the polar
grid converges at first order in the radial step, so two grids are
Richardson-extrapolated.  Cross-polar efficiency of b = c a is
1 / (1 + |c|^2).  Ring frames have a known centre, tilt and RMS.  Masks are
checked on hand-placed points, including the reference's angle taken about
the origin rather than the aperture centre."""

import numpy as np

from oracle import efficiencies as OE
from pypo_amd import geometry as G
from pypo_amd import types as T


def _disc(port, nu, R=4.0):
    pl = G.plane({"name": "p", "gmode": "uv", "lims_u": [0, R], "lims_v": [0, 360], "gridsize": [nu, 360]})
    return OE.rest_grid(port, pl)


def test_gaussian_on_disc_closed_forms(port):
    w, a, R = 1.0, 1.5, 4.0
    ap = {"shape": "ellipse", "outer": [a, a], "inner": [0, 0], "center": [0, 0]}
    vals = {}
    for nu in (401, 801):
        g = _disc(port, nu, R)
        f = np.exp(-(g.x ** 2 + g.y ** 2) / w ** 2) * np.exp(0.3j)
        vals[nu] = (OE.spillover(f, g, ap), OE.taper(f, g, ap), OE.taper(np.ones_like(f), g, ap))
        # spillover is the fraction of power inside the aperture
        m = OE.generate_mask(g.x, g.y, ap)
        p = np.abs(f) ** 2 * g.area
        assert abs(vals[nu][0] - p[m].sum() / p.sum()) < 1e-14
        # a uniform field has taper 1 on any aperture
        assert abs(vals[nu][2] - 1) < 1e-14
    sp = 2 * vals[801][0] - vals[401][0]
    tp = 2 * vals[801][1] - vals[401][1]
    sp_cf = (1 - np.exp(-2 * a ** 2 / w ** 2)) / (1 - np.exp(-2 * R ** 2 / w ** 2))
    tp_cf = 2 * w ** 2 * (1 - np.exp(-a ** 2 / w ** 2)) ** 2 / (a ** 2 * (1 - np.exp(-2 * a ** 2 / w ** 2)))
    assert abs(sp / sp_cf - 1) < 2e-5
    assert abs(tp / tp_cf - 1) < 2e-5


def test_taper_whole_surface_and_empty_aperture(port):
    g = _disc(port, 101)
    f = np.exp(-(g.x ** 2 + g.y ** 2))
    # without an aperture every point counts
    A = g.area.sum()
    ref = abs((f * g.area).sum()) ** 2 / (np.abs(f) ** 2 * g.area).sum() / A
    assert abs(OE.taper(f, g) - ref) < 1e-15
    # an aperture that holds no point gives NaN in the formula, reported as 0
    ap = {"shape": "ellipse", "outer": [1e-9, 1e-9], "inner": [0, 0], "center": [50, 50]}
    assert OE.taper(f, g, ap) == 0.0 and OE.spillover(f, g, ap) == 0.0


def test_xpol_proportional_components():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(13, 9)) + 1j * rng.normal(size=(13, 9))
    for c in (0.1, 1.0, 2 - 1j):
        assert abs(OE.xpol(a, c * a) - 1 / (1 + abs(c) ** 2)) < 1e-14


def test_masks_on_hand_placed_points():
    ell = {"shape": "ellipse", "outer": [2, 1], "inner": [0, 0], "center": [0, 0]}
    x = np.array([1.9, 0.0, 0.0, 2.1, -1.9])
    y = np.array([0.0, 0.9, 1.1, 0.0, 0.0])
    np.testing.assert_array_equal(OE.generate_mask(x, y, ell), [True, True, False, False, True])
    # off-centre ellipse: the angle is taken about the origin, so (2, 5) is
    # outside although (2/3)^2 < 1 about the centre (0, 5)
    off = {"shape": "ellipse", "outer": [3, 1], "inner": [0, 0], "center": [0, 5]}
    np.testing.assert_array_equal(OE.generate_mask(np.array([0.0, 2.0]), np.array([5.5, 5.0]), off),
                                  [True, False])
    rect = {"shape": "rectangle", "center": [1, 1], "outer_x": [-1, 2], "outer_y": [-0.5, 0.5],
            "inner_x": [0, 0], "inner_y": [0, 0]}
    x = np.array([2.9, 3.0, 0.0, 1.0, 1.0])
    y = np.array([1.4, 1.0, 1.0, 1.5, 0.6])
    np.testing.assert_array_equal(OE.generate_mask(x, y, rect), [True, False, False, False, True])


def test_ring_frame_statistics():
    n = 64
    phi = 2 * np.pi * np.arange(n) / n
    r, z0, th = 2.5, 7.0, 0.2
    x, y = 1.0 + r * np.cos(phi), -3.0 + r * np.sin(phi)
    dx, dz = np.full(n, np.sin(th)), np.full(n, np.cos(th))
    # four misses (d = 0) far away must not count
    x = np.r_[x, [1e3] * 4]
    y = np.r_[y, [1e3] * 4]
    z = np.full(n + 4, z0)
    dx, dy, dz = np.r_[dx, [0.0] * 4], np.zeros(n + 4), np.r_[dz, [0.0] * 4]
    fr = T.frame(n + 4, x, y, z, dx, dy, dz)
    np.testing.assert_allclose(OE.rt_center(fr), [1.0, -3.0, z0], atol=1e-14)
    np.testing.assert_allclose(OE.rt_tilt(fr), [np.sin(th), 0, np.cos(th)], atol=1e-15)
    assert abs(OE.rms(fr) - r) < 1e-14
