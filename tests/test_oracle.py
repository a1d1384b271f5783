"""Pins the CPU checker (oracle/liboracle.so) before it is trusted.

* golden vectors written by the reference's own compiled CPU library
  (tests/golden/*.npz, made by tests/golden/make_golden.py);
* the compiled reference itself (oracle/_ref), where it is built;
* closed forms that a plausible mistake would break: the far field of a
  uniformly excited plate (trapezoid weights, phase sign, k^2/4pi with the
  reference's float pi, -j factor, transverse projection), translation
  phase, transversality, the near-field -> far-field limit, a single-point
  spherical wave, ray-quadric intersections solved as quadratics, grid
  geometry against the analytic surfaces.
"""

import numpy as np
import pytest

import cases as C
from helpers import flat, group_err
from oracle import bind as OB
from pypo_amd import geometry as G
from pypo_amd import types as T
from pypo_amd.geometry import gridsize

GOLD = "tests/golden/"
PI_F = float(np.float32(3.14159265359))   # Propagation.h:197 stores a float literal


def _src_for(c):
    shape = gridsize(c["source"])
    return (C.random_scalar(shape, c["seed"]) if c["mode"] == "scalar"
            else C.random_currents(shape, c["seed"]))


@pytest.fixture(scope="module")
def gold_po():
    return np.load(GOLD + "po.npz")


@pytest.mark.parametrize("name", sorted(C.po_cases()))
def test_port_matches_golden_po(port, gold_po, name):
    c = C.po_cases()[name]
    np.testing.assert_array_equal(flat(_src_for(c)), gold_po[f"{name}/src"])
    out, gs, gt = OB.propagate(port, c["mode"], c["source"], c["target"], _src_for(c),
                               2 * np.pi / c["lam"], 3, c.get("epsilon", 1.0), c["exp"])
    assert group_err(flat(out), gold_po[f"{name}/out"]) <= 1e-13
    tg = np.stack([np.ravel(getattr(gt, a)) for a in ("x", "y", "z", "nx", "ny", "nz", "area")])
    np.testing.assert_allclose(tg, gold_po[f"{name}/tgt_grid"], rtol=0, atol=1e-12)


def test_port_grids_match_golden(port):
    gold = np.load(GOLD + "grids.npz")
    for name, refl in C.grid_cases().items():
        for tr, sp in ((True, True), (False, False)):
            g = OB.generate_grid(port, refl, tr, sp)
            got = np.stack([np.ravel(getattr(g, a)) for a in ("x", "y", "z", "nx", "ny", "nz", "area")])
            np.testing.assert_allclose(got, gold[f"{name}/{int(tr)}{int(sp)}"], rtol=1e-14, atol=1e-14,
                                       err_msg=name)


def test_port_rt_matches_golden(port):
    gold = np.load(GOLD + "rt.npz")
    for name, c in C.rt_cases().items():
        out = OB.trace(port, c["target"], C.rt_frame(c), 3, c["tol"], c["t0"], max_iter=0)
        got = np.stack([getattr(out, a) for a in ("x", "y", "z", "dx", "dy", "dz")])
        np.testing.assert_allclose(got, gold[f"{name}/out"], rtol=0, atol=1e-12, err_msg=name)


@pytest.mark.skipif(not OB.available("ref"), reason="oracle/_ref not built (needs the reference tree)")
def test_port_bit_exact_with_reference_library():
    ref, port = OB.load("ref"), OB.load("port")
    rng = np.random.default_rng(5)
    dish = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 40.0], "gmode": "uv",
                         "lims_u": [0, 30], "lims_v": [0, 360], "gridsize": [41, 40]})
    ff = G.plane({"gmode": "AoE", "lims_Az": [-4, 4], "lims_El": [-4, 4], "gridsize": [23, 21]})
    cur = T.currents(*[rng.standard_normal((41, 40)) + 1j * rng.standard_normal((41, 40))
                       for _ in range(6)])
    a, _, _ = OB.propagate(port, "FF", dish, ff, cur, 2 * np.pi / 1.3, 5)
    b, _, _ = OB.propagate(ref, "FF", dish, ff, cur, 2 * np.pi / 1.3, 5)
    np.testing.assert_array_equal(flat(a), flat(b))


# ---- closed forms --------------------------------------------------------------

def _trap_sum(alpha, x0, dx, n):
    """sum_i w_i exp(j alpha x_i) dx over x_i = x0 + i dx with trapezoid
    half weights at both ends, summed as a geometric series."""
    z = np.exp(1j * alpha * dx)
    first = np.exp(1j * alpha * x0)
    last = first * z ** (n - 1)
    with np.errstate(invalid="ignore", divide="ignore"):
        geo = np.where(np.abs(1 - z) < 1e-300, n * first, first * (1 - z ** n) / (1 - z))
    return dx * (geo - 0.5 * (first + last))


def _plate_case(nx=17, ny=12):
    src = G.plane({"gmode": "xy", "lims_x": [-2.0, 3.0], "lims_y": [-1.5, 2.5], "gridsize": [nx, ny]})
    ff = G.plane({"gmode": "AoE", "lims_Az": [-25, 25], "lims_El": [-20, 30], "gridsize": [9, 11]})
    return src, ff


@pytest.mark.parametrize("exp", ["fwd", "bwd"])
def test_far_field_of_uniform_plate_closed_form(port, exp):
    lam = 0.9
    k = 2 * np.pi / lam
    t = -1.0 if exp == "fwd" else 1.0
    src, ff = _plate_case()
    nx, ny = gridsize(src)
    J = np.array([0.7 - 0.2j, -0.3 + 0.9j, 0.0])
    comps = [np.full((nx, ny), J[0]), np.full((nx, ny), J[1]), np.zeros((nx, ny), complex)] + \
            [np.zeros((nx, ny), complex)] * 3
    out, _, gt = OB.propagate(port, "FF", src, ff, T.currents(*comps), k, 2, 1.0, exp)
    th, ph = gt.x.ravel(), gt.y.ravel()
    r = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)])
    dx, dy = 5.0 / (nx - 1), 4.0 / (ny - 1)
    S = _trap_sum(-t * k * r[0], -2.0, dx, nx) * _trap_sum(-t * k * r[1], -1.5, dy, ny)
    SJ = J[:, None] * S[None, :]
    proj = SJ - r * np.sum(SJ * r, axis=0)
    E = -1j * k * k / (4 * PI_F) * proj
    H = -1j * k * k / (4 * PI_F) * (-t) * np.cross(r.T, SJ.T).T
    got = flat(out)
    assert group_err(got, np.concatenate([E, H])) < 1e-11
    # teeth: the same sum without trapezoid end weights is far off
    Sbad = (dx * np.exp(1j * (-t * k * r[0])[:, None] * (-2.0 + dx * np.arange(nx))[None, :]).sum(1)) * \
        (dy * np.exp(1j * (-t * k * r[1])[:, None] * (-1.5 + dy * np.arange(ny))[None, :]).sum(1))
    Ebad = -1j * k * k / (4 * PI_F) * (J[:, None] * Sbad - r * np.sum(J[:, None] * Sbad * r, axis=0))
    assert group_err(got[:3], Ebad) > 1e-3
    # and exact pi instead of the reference's float pi is visible at this tolerance
    assert group_err(got[:3], E * PI_F / np.pi) > 1e-9


def test_far_field_translation_phase_and_transversality(port):
    lam, k = 1.1, 2 * np.pi / 1.1
    rng = np.random.default_rng(9)
    src = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 30.0], "gmode": "uv",
                        "lims_u": [0, 10], "lims_v": [0, 360], "gridsize": [13, 14]})
    ff = G.plane({"gmode": "AoE", "lims_Az": [-6, 6], "lims_El": [-6, 6], "gridsize": [7, 7]})
    cur = T.currents(*[rng.standard_normal((13, 14)) + 1j * rng.standard_normal((13, 14)) for _ in range(6)])
    a, _, gt = OB.propagate(port, "FF", src, ff, cur, k, 2)
    d = np.array([3.0, -2.0, 5.0])
    moved = dict(src)
    moved["transf"] = G.mat_translate(d)
    b, _, _ = OB.propagate(port, "FF", moved, ff, cur, k, 2)
    th, ph = gt.x.ravel(), gt.y.ravel()
    r = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)])
    phase = np.exp(1j * k * (d @ r))          # t = -1: exp(-j t k d.r)
    A, Bm = flat(a), flat(b)
    assert group_err(Bm, A * phase[None, :]) < 1e-11
    # E and H are transverse to r
    for g in (0, 3):
        radial = np.abs(np.sum(A[g:g + 3] * r, axis=0))
        assert radial.max() < 1e-12 * np.abs(A[g:g + 3]).max()


def test_near_field_tends_to_far_field(port):
    lam, k = 1.0, 2 * np.pi
    rng = np.random.default_rng(4)
    src = G.plane({"gmode": "xy", "lims_x": [-0.4, 0.4], "lims_y": [-0.4, 0.4], "gridsize": [5, 5]})
    cur = T.currents(*[rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5)) for _ in range(6)])
    ff = G.plane({"gmode": "AoE", "lims_Az": [-30, 30], "lims_El": [-30, 30], "gridsize": [3, 3]})
    Eff, _, gt = OB.propagate(port, "FF", src, ff, cur, k, 1)
    th, ph = gt.x.ravel(), gt.y.ravel()
    r = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)])
    R = 2.0e6
    # target: a 3x3 "plane" grid whose points are R*r_hat (built through transf per point)
    near = []
    for i in range(r.shape[1]):
        tgt = G.plane({"gmode": "xy", "lims_x": [-1e-9, 1e-9], "lims_y": [-1e-9, 1e-9], "gridsize": [2, 2]})
        tgt["transf"] = G.mat_translate(R * r[:, i])
        e, _, _ = OB.propagate(port, "EH", src, tgt, cur, k, 1)
        near.append(flat(e)[:, 0])
    near = np.stack(near, axis=1)
    scale = np.exp(-1j * k * R) / (k * R)     # t = -1: e^{j t k R} / (kR)
    assert group_err(near, flat(Eff) * scale) < 1e-4


def test_scalar_single_point_spherical_wave(port):
    lam, k = 0.8, 2 * np.pi / 0.8
    src = G.plane({"gmode": "xy", "lims_x": [-0.2, 0.2], "lims_y": [-0.2, 0.2], "gridsize": [3, 3]})
    u = np.zeros((3, 3), complex)
    u[1, 1] = 0.6 - 1.1j
    tgt = G.plane({"gmode": "xy", "lims_x": [-5, 5], "lims_y": [-4, 6], "gridsize": [6, 5]})
    tgt["transf"] = G.mat_translate([1.0, 2.0, 12.0])
    for exp, t in (("fwd", -1.0), ("bwd", 1.0)):
        out, _, gt = OB.propagate(port, "scalar", src, tgt, T.scalarfield(u), k, 2, 1.0, exp)
        rr = np.sqrt(gt.x.ravel() ** 2 + gt.y.ravel() ** 2 + gt.z.ravel() ** 2)
        area = 0.2 * 0.2
        want = -k * k * u[1, 1] * area * np.exp(-1j * t * k * rr) / (4 * PI_F * rr)
        assert group_err(flat(out), want[None, :], 1) < 1e-13


def _first_root(A, B, Cc):
    A, B, Cc = (np.asarray(v, dtype=np.longdouble) for v in (A, B, Cc))
    disc = np.sqrt(B * B - 4 * A * Cc)
    r1 = (-B - disc) / (2 * A)
    r2 = (-B + disc) / (2 * A)
    lo, hi = np.minimum(r1, r2), np.maximum(r1, r2)
    return np.where(lo > 0, lo, hi)


def test_ray_paraboloid_intersection_quadratic_and_focus(port):
    f = 50.0
    par = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, f], "gmode": "xy",
                        "lims_x": [-1, 1], "lims_y": [-1, 1], "gridsize": [3, 3]})
    n = 101
    rng = np.random.default_rng(0)
    x, y = rng.uniform(-20, 20, n), rng.uniform(-20, 20, n)
    z = np.full(n, 80.0)
    fr = T.frame(n, x, y, z, np.zeros(n), np.zeros(n), -np.ones(n))
    out = OB.trace(port, par, fr, 2, 1e-13, 10.0, max_iter=0)
    a2 = par["coeffs"][0] ** 2
    t = z - (x * x + y * y) / a2          # vertical rays: the quadratic is linear
    np.testing.assert_allclose(out.z, (80.0 - t).astype(float), atol=1e-10)
    # reflected rays pass through the focus
    s = (f - out.z) / out.dz
    np.testing.assert_allclose(out.x + s * out.dx, 0, atol=1e-9)
    np.testing.assert_allclose(out.y + s * out.dy, 0, atol=1e-9)
    # oblique rays against the quadratic root
    d = np.stack([rng.normal(0, 0.1, n), rng.normal(0, 0.1, n), -np.ones(n)])
    d /= np.linalg.norm(d, axis=0)
    fr = T.frame(n, x, y, z, d[0], d[1], d[2])
    out = OB.trace(port, par, fr, 2, 1e-13, 10.0, max_iter=0)
    A = (d[0] ** 2 + d[1] ** 2) / a2
    B = 2 * (x * d[0] + y * d[1]) / a2 - d[2]
    Cc = (x * x + y * y) / a2 - z
    t = _first_root(A, B, Cc)
    for comp, p0, dd in ((out.x, x, d[0]), (out.y, y, d[1]), (out.z, z, d[2])):
        np.testing.assert_allclose(comp, (p0 + t * dd).astype(float), atol=1e-9)


def test_cassegrain_rays_converge_to_second_focus(port):
    prim = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 120.0], "gmode": "xy",
                         "lims_x": [-1, 1], "lims_y": [-1, 1], "gridsize": [3, 3]})
    sec = G.hyperboloid({"pmode": "focus", "focus_1": [0, 0, 120.0], "focus_2": [0, 0, -30.0],
                         "ecc": 1.1, "gmode": "xy", "lims_x": [-1, 1], "lims_y": [-1, 1],
                         "gridsize": [3, 3]})
    n = 64
    rng = np.random.default_rng(2)
    rr, ph = rng.uniform(20, 140, n), rng.uniform(0, 2 * np.pi, n)
    fr = T.frame(n, rr * np.cos(ph), rr * np.sin(ph), np.full(n, 200.0), np.zeros(n), np.zeros(n),
                 -np.ones(n))
    f1 = OB.trace(port, prim, fr, 2, 1e-13, 100.0, max_iter=0)
    f2 = OB.trace(port, sec, f1, 2, 1e-13, 100.0, max_iter=0)
    s = (-30.0 - f2.z) / f2.dz
    np.testing.assert_allclose(f2.x + s * f2.dx, 0, atol=1e-8)
    np.testing.assert_allclose(f2.y + s * f2.dy, 0, atol=1e-8)
    # reflection law at the secondary: |d_out| = 1 and angle in = angle out
    np.testing.assert_allclose(np.sqrt(f2.dx ** 2 + f2.dy ** 2 + f2.dz ** 2), 1, atol=1e-12)


def test_grid_geometry_matches_analytic_surfaces(port):
    f = 25.0
    par = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, f], "gmode": "uv",
                        "lims_u": [0, 20], "lims_v": [0, 360], "gridsize": [201, 200]})
    g = OB.generate_grid(port, par)
    x, y, z = g.x.ravel(), g.y.ravel(), g.z.ravel()
    np.testing.assert_allclose(z, (x * x + y * y) / (4 * f), atol=1e-12)
    nrm = np.sqrt(g.nx ** 2 + g.ny ** 2 + g.nz ** 2)
    np.testing.assert_allclose(nrm, 1, atol=1e-14)
    # normal is parallel to grad(z - r^2/4f) = (-x/2f, -y/2f, 1)
    gv = np.stack([-x / (2 * f), -y / (2 * f), np.ones_like(x)])
    gv /= np.linalg.norm(gv, axis=0)
    np.testing.assert_allclose(np.stack([g.nx.ravel(), g.ny.ravel(), g.nz.ravel()]), gv, atol=1e-13)
    # trapezoid-weighted area (as the PO sums weight it) vs the cap area
    # 2 pi/3 * 2f * ((1 + R^2/4f^2)^(3/2) - 1) * 2f   (surface of revolution)
    R = 20.0
    exact = (8 * np.pi * f * f / 3) * ((1 + R * R / (4 * f * f)) ** 1.5 - 1)
    w = np.ones((201, 200))
    w[0, :] *= 0.5
    w[-1, :] *= 0.5
    assert abs(np.sum(w * g.area) - exact) / exact < 1e-4
    # AoE spheric mapping: polar = sqrt(Az^2 + El^2), azimuth = atan2(El, Az)
    ff = G.plane({"gmode": "AoE", "lims_Az": [-3, 3], "lims_El": [-2, 4], "gridsize": [7, 7]})
    gf = OB.generate_grid(port, ff)
    Az, El = np.meshgrid(np.linspace(-3, 3, 7), np.linspace(-2, 4, 7), indexing="ij")
    np.testing.assert_allclose(gf.x, np.sqrt(Az ** 2 + El ** 2) * np.pi / 180, atol=1e-15)
    np.testing.assert_allclose(np.where(Az ** 2 + El ** 2 > 0, gf.y, 0),
                               np.where(Az ** 2 + El ** 2 > 0, np.arctan2(El, Az), 0), atol=1e-15)


def test_rough_grid_deviates_follow_normals(port):
    base = C.grid_cases()["par_uv_rough"]
    smooth = dict(base, rms=0.0)
    a = OB.generate_grid(port, base)
    b = OB.generate_grid(port, smooth)
    d = np.stack([a.x - b.x, a.y - b.y, a.z - b.z])
    n = np.stack([b.nx, b.ny, b.nz])
    e = np.sum(d * n, axis=0)
    np.testing.assert_allclose(d, e[None] * n, atol=1e-13)
    # std::mt19937(1234) + a fresh std::normal_distribution(0, rms) per point
    assert 0.01 < np.std(e) < 0.1
