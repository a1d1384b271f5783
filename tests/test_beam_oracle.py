"""Pins the checker's beam initialisation (oracle/pypo_oracle.cpp, restating
src/common/BeamInit.h) against golden vectors written by the compiled
reference (tests/golden/beams.npz) and against closed forms: the Gaussian at
its waist, the scalar Gaussian's normalisation, M = -2 n x E, the ring
geometry of tubular frames and the moments of the reference's (random)
Gaussian ray frame."""

import numpy as np
import pytest

import cases as C
from helpers import flat, group_err
from oracle import bind as OB
from pypo_amd.geometry import gridsize

GOLD = np.load("tests/golden/beams.npz")


def run_case(lib, c):
    if c["kind"] in ("gauss", "vec"):
        f, cur = OB.make_gauss(lib, c["params"], c["source"], vector=c["kind"] == "vec")
        return {"field": flat(f), "current": flat(cur)}
    if c["kind"] == "scalar":
        return {"field": np.ravel(OB.make_scalar_gauss(lib, c["params"], c["source"]).S)}
    if c["kind"] == "currents":
        fields = C.random_fields(gridsize(c["source"]), c["seed"])
        return {"current": flat(OB.calc_currents(lib, fields, c["source"], c["mode"]))}
    fr = OB.make_rt_frame(lib, c["params"])
    return {"frame": np.stack([getattr(fr, a) for a in ("x", "y", "z", "dx", "dy", "dz")])}


@pytest.mark.parametrize("name", sorted(C.beam_cases()))
def test_port_beams_match_golden(port, name):
    c = C.beam_cases()[name]
    got = run_case(port, c)
    for key, val in got.items():
        ref = GOLD[f"{name}/{key}"]
        if key == "frame":
            np.testing.assert_allclose(val, ref, rtol=0, atol=1e-14)
        elif val.ndim == 1:
            assert np.max(np.abs(val - ref)) <= 1e-14 * np.max(np.abs(ref))
        else:
            assert group_err(val, ref) <= 1e-14, key


def test_gauss_at_waist_closed_form(port):
    """z = 0 plane, w0x = w0y = w0: E = E0 exp(-(x^2 + y^2) / w0^2) pol, real;
    plane normal +z, so M = -2 z x E = (0, -2 Ex, 0)."""
    c = C.beam_cases()["gauss_waist"]
    g = OB.generate_grid(port, c["source"])
    x, y = np.ravel(g.x), np.ravel(g.y)
    f, cur = OB.make_gauss(port, c["params"], c["source"])
    ex = np.exp(-(x ** 2 + y ** 2) / 4.0)
    np.testing.assert_allclose(np.ravel(f.Ex), ex, rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(np.ravel(cur.My), -2 * ex, rtol=1e-14, atol=1e-16)
    assert np.all(np.ravel(f.Ey) == 0) and np.all(np.ravel(cur.Jx) == 0)


def test_scalar_gauss_normalisation(port):
    """At the waist the scalar Gaussian is E0 sqrt(2 / (pi w0x w0y)) exp(-x^2/w0x^2 - y^2/w0y^2)."""
    c = C.beam_cases()["scalar_waist"]
    g = OB.generate_grid(port, c["source"])
    x, y = np.ravel(g.x), np.ravel(g.y)
    s = np.ravel(OB.make_scalar_gauss(port, c["params"], c["source"]).S)
    ref = np.sqrt(2 / (np.pi * 2.0 * 1.0)) * np.exp(-x ** 2 / 4.0 - y ** 2 / 1.0)
    np.testing.assert_allclose(s, ref, rtol=1e-14, atol=1e-16)


def test_vector_gauss_mode_factors(port):
    """PEC doubles J, PMC doubles M, full keeps both (BeamInit.h:524-594)."""
    cs = C.beam_cases()
    full = flat(OB.make_gauss(port, cs["vec_gauss_m0"]["params"], cs["vec_gauss_m0"]["source"], True)[1])
    pmc = flat(OB.make_gauss(port, cs["vec_gauss_m1"]["params"], cs["vec_gauss_m1"]["source"], True)[1])
    pec = flat(OB.make_gauss(port, cs["vec_gauss_m2"]["params"], cs["vec_gauss_m2"]["source"], True)[1])
    np.testing.assert_array_equal(pec[:3], 2 * full[:3])
    np.testing.assert_array_equal(pmc[3:], 2 * full[3:])
    assert not pec[3:].any() and not pmc[:3].any()


def test_tube_frame_rings(port):
    """Rays of ring n lie on the ellipse (x/x0)^2 + (y/y0)^2 = (n/nRing)^2 and
    directions are unit vectors; ray 0 is the chief ray."""
    p = C.beam_cases()["rt_tube_3ring"]["params"]
    fr = OB.make_rt_frame(port, p)
    assert fr.size == 1 + 4 * 4 * 3
    assert (fr.x[0], fr.y[0], fr.dz[0]) == (0.0, 0.0, 1.0)
    rings = np.repeat(np.arange(1, 4), 16)
    r2 = (fr.x[1:] / p["x0"]) ** 2 + (fr.y[1:] / p["y0"]) ** 2
    np.testing.assert_allclose(r2, (rings / 3.0) ** 2, rtol=1e-13)
    np.testing.assert_allclose(fr.dx ** 2 + fr.dy ** 2 + fr.dz ** 2, 1.0, rtol=1e-14)


def test_reference_gaussian_frame_moments():
    """The reference's own Gaussian frame (random: std::random_device) has the
    moments of the truncated Gaussian exp(-x^2/s^2) on [-3s, 3s]: mean 0,
    standard deviation s / sqrt(2) (4-sigma bands for 999 samples)."""
    fr = GOLD["grt/frame"]
    c = C.grt_case()
    n = fr.shape[1] - 1
    s = np.array([c["x0"], c["y0"]])
    for k in range(2):
        v = fr[k, 1:]
        sd = s[k] / np.sqrt(2)
        assert abs(v.mean()) < 4 * sd / np.sqrt(n)
        assert abs(v.std() / sd - 1) < 4 / np.sqrt(2 * n)
    np.testing.assert_allclose(fr[3] ** 2 + fr[4] ** 2 + fr[5] ** 2, 1.0, rtol=1e-14)
