"""The hybrid checker's carried factor (oracle/hybrid.carry, restating
System.py:1596-1605 with calcRayLen 1304-1372) against geometry worked out
by hand.  Rays leave the point `start` = origin, are sampled on the plane
z = z1 and traced onto the parallel plane z = z2.  Along the ray through the
sample p = (x, y, z1), direction p / |p|, similar triangles give

    L0 = |p|,    L1 = (z2 - z1) |p| / z1,

so the factor exp(j k L1) sqrt(L0 / (2 L1 + L0)) is
exp(j k (z2 - z1) |p| / z1) sqrt(z1 / (2 z2 - z1)): the same amplitude on
every ray.  Negative controls: swapping L0 and L1, pairing the wrong frames
(the input frame in place of the traced one) and dropping the start point
each give a factor this test rejects."""

import numpy as np
import pytest

from oracle import bind as OB
from oracle import hybrid as OH
from pypo_amd import geometry as G
from pypo_amd import types as T

Z1, Z2, LAM = 100.0, 300.0, 1.7
K = 2 * np.pi / LAM


def _case(n=25, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-60, 60, n * n)
    y = rng.uniform(-60, 60, n * n)
    p = np.stack([x, y, np.full(x.size, Z1)])
    d = p / np.linalg.norm(p, axis=0)
    fr = T.frame(x.size, p[0], p[1], p[2], d[0], d[1], d[2])
    target = G.plane({"gmode": "xy", "lims_x": [-500, 500], "lims_y": [-500, 500], "gridsize": [3, 3]})
    target["transf"] = G.mat_translate([0, 0, Z2])
    comps = [rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size) for _ in range(6)]
    return fr, target, comps, np.linalg.norm(p, axis=0)


def closed_form(rho):
    """exp(j k L1) sqrt(L0 / (2 L1 + L0)) with L0 = rho, L1 = (z2 - z1) rho / z1."""
    return np.exp(1j * K * (Z2 - Z1) * rho / Z1) * np.sqrt(Z1 / (2 * Z2 - Z1))


def swapped_form(rho):
    L0, L1 = rho, (Z2 - Z1) * rho / Z1
    return np.exp(1j * K * L0) * np.sqrt(L1 / (2 * L0 + L1))


def _factor(carried, comps):
    return np.stack(carried) / np.stack(comps)


def test_carry_matches_closed_form(port):
    fr, target, comps, rho = _case()
    fr_out, carried = OH.carry(port, target, fr, comps, K, start=np.zeros(3), tol=1e-10, t0=150.0)
    np.testing.assert_allclose(fr_out.z, Z2, rtol=0, atol=1e-10)
    f = _factor(carried, comps)
    want = closed_form(rho)
    assert np.max(np.abs(f - want[None, :])) <= 1e-9
    # amplitude: one constant for every ray
    assert np.ptp(np.abs(f)) <= 1e-12


def test_carry_without_start_is_the_pure_phase(port):
    fr, target, comps, rho = _case(seed=1)
    _, carried = OH.carry(port, target, fr, comps, K, start=None, tol=1e-10, t0=150.0)
    want = np.exp(1j * K * (Z2 - Z1) * rho / Z1)
    assert np.max(np.abs(_factor(carried, comps) - want[None, :])) <= 1e-9


def test_negative_controls_are_rejected(port, monkeypatch):
    fr, target, comps, rho = _case(seed=2)
    want = closed_form(rho)[None, :]
    # L0 and L1 swapped: the amplitude sqrt(200/400) against sqrt(100/500)
    assert np.max(np.abs(swapped_form(rho)[None, :] - want)) > 0.1
    # no start point: the amplitude factor is missing
    _, no_start = OH.carry(port, target, fr, comps, K, start=None, tol=1e-10, t0=150.0)
    assert np.max(np.abs(_factor(no_start, comps) - want)) > 0.1
    # wrong frame pairing, planted in the checker: the "traced" frame is the input frame
    monkeypatch.setattr(OB, "trace", lambda lib, tgt, f, **kw: f)
    _, wrong = OH.carry(port, target, fr, comps, K, start=np.zeros(3), tol=1e-10, t0=150.0)
    assert np.max(np.abs(_factor(wrong, comps) - want)) > 0.1


@pytest.mark.parametrize("z2", [150.0, 1000.0])
def test_amplitude_factor_over_distances(port, z2):
    """sqrt(L0 / (2 L1 + L0)) = sqrt(z1 / (2 z2 - z1)) for other target distances."""
    fr, target, comps, rho = _case(n=9, seed=3)
    target["transf"] = G.mat_translate([0, 0, z2])
    _, carried = OH.carry(port, target, fr, comps, K, start=np.zeros(3), tol=1e-10, t0=z2 / 2)
    f = _factor(carried, comps)
    np.testing.assert_allclose(np.abs(f), np.sqrt(Z1 / (2 * z2 - Z1)), rtol=1e-12)
    want = np.exp(1j * K * (z2 - Z1) * rho / Z1) * np.sqrt(Z1 / (2 * z2 - Z1))
    assert np.max(np.abs(f - want[None, :])) <= 1e-8
