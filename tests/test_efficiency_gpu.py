"""Figures of merit on the GPU (metrics.hip through pypo_hip_aperture_sums_dev,
pypo_hip_power_sums_dev, pypo_hip_frame_stats_dev; pypo_amd/efficiencies.py
and the System methods) against the CPU checker oracle/efficiencies.py.

Bar: 1e-12 relative for every efficiency and raw sum (FP64 reductions in a
different order); the aperture area sum agreeing to 1e-12 also shows the
masks select the same points.  At 4e6 points: a uniform field has taper 1,
an aperture holding the whole surface gives spillover 1, sums scale with
|c|^2, and two calls agree bit for bit.  findRTfocus on a paraboloid lands on
its focus."""

import numpy as np
import pytest

from oracle import bind as OB
from oracle import efficiencies as OE
from pypo_amd import geometry as G
from pypo_amd import types as T

pytestmark = pytest.mark.gpu

ELL = {"shape": "ellipse", "center": [0.3, -0.2], "outer": [1.7, 2.3], "inner": [0.2, 0.2]}
RECT = {"shape": "rectangle", "center": [0.5, 0.1], "outer_x": [-2.1, 1.3], "outer_y": [-0.7, 2.6],
        "inner_x": [0, 0], "inner_y": [0, 0]}


def _surfaces():
    disc = G.plane({"name": "disc", "gmode": "uv", "lims_u": [0, 4], "lims_v": [0, 360], "gridsize": [121, 90]})
    dish = G.paraboloid({"name": "dish", "pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 6.0],
                         "gmode": "xy", "lims_x": [-3.5, 3.5], "lims_y": [-3, 3.2], "gridsize": [97, 83]})
    # a transformation moves the surface but not its rest-frame grid
    c, s = np.cos(0.4), np.sin(0.4)
    dish["transf"] = np.array([[c, 0, s, 5.0], [0, 1, 0, -2.0], [-s, 0, c, 30.0], [0, 0, 0, 1]])
    return {"disc": disc, "dish": dish}


def _field(g, seed):
    rng = np.random.default_rng(seed)
    amp = np.exp(-(g.x ** 2 + g.y ** 2) / 4.0)
    return amp * (rng.normal(1, 0.2, g.x.shape) + 1j * rng.normal(0, 0.2, g.x.shape))


@pytest.mark.parametrize("surf", ["disc", "dish"])
@pytest.mark.parametrize("aper", [ELL, RECT], ids=["ellipse", "rectangle"])
def test_spillover_taper_match_checker(hip, port, surf, aper):
    from pypo_amd import efficiencies as E
    refl = _surfaces()[surf]
    g = OE.rest_grid(port, refl)
    f = _field(g, 7)
    sp, tp = E.calcSpillover(f, refl, aper), E.calcTaper(f, refl, aper)
    assert isinstance(sp, float) and isinstance(tp, float)
    assert abs(sp / OE.spillover(f, g, aper) - 1) < 1e-12
    assert abs(tp / OE.taper(f, g, aper) - 1) < 1e-12
    assert abs(E.calcTaper(f, refl) / OE.taper(f, g) - 1) < 1e-12
    # the raw sums: the area inside the aperture agrees, so the masks do
    s = E.aperture_sums(E.rest_grid(refl), E._complex_rows(f), aper)
    m = OE.generate_mask(g.x, g.y, aper)
    ref = [np.sum(np.abs(f) ** 2 * g.area), np.sum((np.abs(f) ** 2 * g.area)[m]),
           np.sum((f.real * g.area)[m]), np.sum((f.imag * g.area)[m]), np.sum(g.area[m])]
    np.testing.assert_allclose(s, ref, rtol=1e-12)
    assert 0 < m.sum() < m.size


def test_xpol_matches_checker(hip):
    from pypo_amd import efficiencies as E
    rng = np.random.default_rng(5)
    co = rng.normal(size=(301, 207)) + 1j * rng.normal(size=(301, 207))
    cr = 0.05 * (rng.normal(size=co.shape) + 1j * rng.normal(size=co.shape))
    assert abs(E.calcXpol(co, cr) / OE.xpol(co, cr) - 1) < 1e-12
    assert abs(E.calcXpol(co, (0.5 - 0.5j) * co) - 1 / 1.5) < 1e-13


def _frame(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = rng.normal(3, 2, n), rng.normal(-1, 1, n), rng.normal(40, 0.5, n)
    d = rng.normal(0, 0.1, (3, n))
    d[2] += 1
    d /= np.linalg.norm(d, axis=0)
    miss = rng.random(n) < 0.1
    d[:, miss] = 0.0
    x[miss] = 1e6
    return T.frame(n, x, y, z, d[0], d[1], d[2])


def test_frame_statistics_match_checker(hip):
    from pypo_amd import efficiencies as E
    fr = _frame(200_003, 11)
    np.testing.assert_allclose(E.calcRTcenter(fr), OE.rt_center(fr), rtol=1e-12)
    np.testing.assert_allclose(E.calcRTtilt(fr), OE.rt_tilt(fr), rtol=1e-12, atol=1e-15)
    assert abs(E.calcRMS(fr) / OE.rms(fr) - 1) < 1e-12


def test_traced_frame_statistics(hip, port):
    """A frame traced on the device is scored in place (no copy to the host)."""
    from pypo_amd import device as D
    from pypo_amd import efficiencies as E
    from pypo_amd import beam as PB
    dish = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 10.0], "gmode": "xy",
                         "lims_x": [-5, 5], "lims_y": [-5, 5], "gridsize": [3, 3]})
    rt = {"nRays": 40, "nRing": 25, "angx0": 0.5, "angy0": 0.3, "x0": 3.0, "y0": 2.0}
    fr_t = PB.make_rt_frame(rt)
    hit = D.trace([dish], fr_t, tol=1e-12)[0]
    ref = OB.trace(port, dish, D.tensor_to_frame(fr_t), None, 1e-12, 1.0, max_iter=50)
    np.testing.assert_allclose(E.calcRTcenter(hit), OE.rt_center(ref), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(E.calcRTtilt(hit), OE.rt_tilt(ref), rtol=1e-11, atol=1e-12)
    assert abs(E.calcRMS(hit) / OE.rms(ref) - 1) < 1e-11


def test_full_size_properties(hip):
    import torch
    from pypo_amd import efficiencies as E
    refl = G.plane({"name": "big", "gmode": "xy", "lims_x": [-10, 10], "lims_y": [-10, 10],
                    "gridsize": [2000, 2000]})
    g = E.rest_grid(refl)
    n = g.shape[1]
    x, y = g[0], g[1]
    amp = torch.exp(-(x * x + y * y) / 16.0)
    comp = torch.stack([amp * torch.cos(0.2 * x), amp * torch.sin(0.2 * x)]).contiguous()
    ones = torch.stack([torch.ones(n, dtype=torch.float64, device=g.device),
                        torch.zeros(n, dtype=torch.float64, device=g.device)]).contiguous()
    everything = {"shape": "rectangle", "center": [0, 0], "outer_x": [-11, 11], "outer_y": [-11, 11],
                  "inner_x": [0, 0], "inner_y": [0, 0]}
    assert abs(E.calcTaper(ones, refl, ELL, grid=g) - 1) < 1e-12
    assert abs(E.calcTaper(ones, refl, grid=g) - 1) < 1e-12
    assert abs(E.calcSpillover(comp, refl, everything, grid=g) - 1) < 1e-15
    a = E.aperture_sums(g, comp, ELL)
    b = E.aperture_sums(g, comp, ELL)
    assert np.array_equal(a, b)                      # deterministic order
    c = 1.5 - 2.0j
    scaled = torch.stack([1.5 * comp[0] + 2.0 * comp[1], 1.5 * comp[1] - 2.0 * comp[0]]).contiguous()
    s = E.aperture_sums(g, scaled, ELL)
    np.testing.assert_allclose(s[[0, 1]], abs(c) ** 2 * a[[0, 1]], rtol=1e-13)
    np.testing.assert_allclose(s[2] + 1j * s[3], c * (a[2] + 1j * a[3]), rtol=1e-13)
    assert s[4] == a[4]
    # spillover of a field that is zero outside the aperture is 1
    inside = torch.as_tensor(OE.generate_mask(x.cpu().numpy(), y.cpu().numpy(), ELL), device=g.device)
    masked = (comp * inside).contiguous()
    assert abs(E.calcSpillover(masked, refl, ELL, grid=g) - 1) < 1e-13


def test_system_efficiencies(hip, port):
    from pypo_amd.system import System
    s = System()
    s.addPlane({"name": "pl", "gmode": "xy", "lims_x": [-2, 2], "lims_y": [-2, 2], "gridsize": [101, 101]})
    s.createGaussian({"name": "g", "lam": 0.01, "w0x": 1, "w0y": 1, "n": 1, "E0": 1, "dxyz": 1,
                      "pol": np.array([1, 0, 0])}, "pl")
    g = OE.rest_grid(port, s.system["pl"])
    ex, ey = s.fields["g"].Ex, s.fields["g"].Ey
    aper = {"shape": "ellipse", "center": [0, 0], "outer": [0.9, 0.9], "inner": [0, 0]}
    sp = s.calcSpillover("g", "Ex", aper)
    assert isinstance(sp, float) and abs(sp / OE.spillover(ex, g, aper) - 1) < 1e-12
    assert abs(s.calcTaper("g", 0) / OE.taper(ex, g) - 1) < 1e-12
    assert abs(s.calcTaper("g", "Ex", aperDict=aper) / OE.taper(ex, g, aper) - 1) < 1e-12
    assert abs(s.calcXpol("g", "Ex", "Ey") - OE.xpol(ex, ey)) < 1e-14
    assert s.calcXpol("g", "Ex", "Ex") == 0.5


def test_findRTfocus_on_paraboloid(hip, port):
    """Rays parallel to the axis of a paraboloid meet at its focus."""
    from scipy.optimize import fmin
    from pypo_amd.system import System
    s = System()
    s.addParabola({"name": "par", "pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 10.0],
                   "gmode": "xy", "lims_x": [-4, 4], "lims_y": [-4, 4], "gridsize": [3, 3]})
    s.createTubeFrame({"name": "tube", "nRays": 8, "nRing": 6, "angx0": 0, "angy0": 0, "x0": 3, "y0": 3})
    s.runRayTracer({"fr_in": "tube", "t_name": "par", "fr_out": "refl", "tol": 1e-12})
    focus = s.findRTfocus("refl")
    assert focus.shape == (3,)
    assert np.linalg.norm(focus - [0, 0, 10.0]) < 1e-3
    assert "focal_plane_refl" in s.system and "focus_refl" in s.frames
    assert s.calcSpotRMS("focus_refl") < 1e-3
    # the same search on the CPU checker's tracer ends at the same point
    fr = s.frames["refl"]
    tilt, center = OE.rt_tilt(fr), OE.rt_center(fr)
    plane = dict(s.system["focal_plane_refl"])

    def spot(f):
        T4 = np.eye(4)
        T4[:3, 3] = center + f[0] * tilt
        plane["transf"] = T4
        return OE.rms(OB.trace(port, plane, fr, None, 1e-12, 1.0, max_iter=50))

    ref = fmin(spot, 0.0, full_output=True, disp=False, ftol=1e-12)
    np.testing.assert_allclose(focus, ref[0][0] * tilt + center, atol=1e-9)
