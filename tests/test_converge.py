"""System.autoConverge / System.convergeOnTarget (System.py:2779-3080).

The loops are host orchestration over scalar runPO calls.  The CPU tests
drive them with a synthetic propagation whose field is a known function of
the grid size, so the step sequence (grow, stop, bisect) can be worked out
by hand; the GPU tests run them on the device and check that they stop at
the same grid sizes as the same loops driven by the CPU checker's
propagation, and the reference test's assertions (an ndarray equal to the
surface's new gridsize)."""

import numpy as np
import pytest

from oracle import bind as OB
from pypo_amd import bind as B
from pypo_amd import types as T
from pypo_amd.geometry import gridsize
from pypo_amd.system import System


def _xy_system():
    s = System()
    s.addPlane({"name": "src", "gmode": "xy", "lims_x": [-1, 1], "lims_y": [-1, 1], "gridsize": [5, 5]})
    s.addPlane({"name": "scat", "gmode": "xy", "lims_x": [-9, 9], "lims_y": [-9, 9], "gridsize": [13, 13]})
    s.addPlane({"name": "tgt", "gmode": "xy", "lims_x": [-3, 3], "lims_y": [-3, 3], "gridsize": [7, 7]})
    z = np.zeros((5, 5), dtype=complex)
    f = T.fields(z + 1, z, z, z, z + 0.5, z)
    f.setMeta("src", 1.0)
    s.fields["beam"] = f
    return s


def test_autoConverge_steps(monkeypatch):
    """S = 1 + 1/g0 on the patch and areas summing to 1, so P(g) =
    ((1 + 1/g) / (1 + 1/11))^2 for the grid g = 11, 21, 31, ...  The steps
    differ by 0.078, 0.027, 0.0136, 0.0082: the first below 1e-2 is at 51,
    which scales back to 51 * 9 = 459 on the full surface."""
    s = _xy_system()
    seen = []

    def fake_po(source, target, d):
        g0, g1 = gridsize(target)
        seen.append((source["name"], target["name"], g0, g1, list(target["lims_x"])))
        assert d["mode"] == "scalar" and d["epsilon"] == 10
        return T.scalarfield(np.full((g0, g1), 1 + 1 / g0, dtype=complex))

    def fake_grid(refl, transform=True, spheric=True):
        g0, g1 = gridsize(refl)
        a = np.full((g0, g1), 1.0 / (g0 * g1))
        return T.reflGrids(a, a, a, a, a, a, a)

    monkeypatch.setattr(System, "_propagate", lambda self, src, tgt, d: fake_po(src, tgt, d))
    monkeypatch.setattr(B, "generate_grid", fake_grid)
    out = s.autoConverge("beam", "scat")
    assert type(out) is np.ndarray
    assert list(out) == [459, 459]
    assert list(s.system["scat"]["gridsize"]) == [459, 459]
    assert [g for (_, _, g, _, _) in seen] == [11, 21, 31, 41, 51]
    # the patch is a ninth of the surface, and the limits come back afterwards
    np.testing.assert_allclose(seen[0][4], [-1, 1])
    np.testing.assert_allclose(s.system["scat"]["lims_x"], [-9, 9])
    assert "_S_conv" not in s.scalarfields
    # the strongest component (Ex, |1| > |0.5|) is the propagated scalar field
    np.testing.assert_array_equal(s.scalarfields["_beam"].S, np.ones((5, 5)))


def test_autoConverge_max_iter(monkeypatch):
    s = _xy_system()
    calls = []

    def fake_po(source, target, d):
        calls.append(1)
        g0, g1 = gridsize(target)
        return T.scalarfield(np.full((g0, g1), (-1.0) ** len(calls), dtype=complex))

    def fake_grid(refl, transform=True, spheric=True):
        g0, g1 = gridsize(refl)
        a = np.full((g0, g1), 1.0)
        return T.reflGrids(a, a, a, a, a, a, a)

    monkeypatch.setattr(System, "_propagate", lambda self, src, tgt, d: fake_po(src, tgt, d))
    monkeypatch.setattr(B, "generate_grid", fake_grid)
    # P grows with the grid and never settles: stops after max_iter steps
    out = s.autoConverge("beam", "scat", add=1, max_iter=4, patch_size=0.5)
    assert len(calls) == 4
    assert list(out) == [10, 10]
    np.testing.assert_allclose(s.system["scat"]["lims_x"], [-9, 9])


def test_convergeOnTarget_steps(monkeypatch):
    """Target field 1 + 4/g0^2 + 1/g1^2 for a scatterer grid (g0, g1).
    Axis 0 (g1 = 10): 20 differs from 10 by 2.9e-2, 40 from 20 by 7.4e-3, so
    40 converges; the bisection compares against the field at 20 and accepts
    30, 25, 22.  Axis 1 (g0 = 10): 20 differs from 10 by 7.2e-3; bisection
    accepts 15, 12, 11."""
    s = _xy_system()

    def fake_po(source, target, d):
        g0, g1 = gridsize(target)
        if target["name"] == "_conv_scatterer":
            return T.scalarfield(np.ones((g0, g1), dtype=complex))
        assert target["name"] == "_conv_target" and source["name"] == "_conv_scatterer"
        assert (g0, g1) == (5, 5)
        s0, s1 = gridsize(source)
        return T.scalarfield(np.full((g0, g1), 1 + 4 / s0 ** 2 + 1 / s1 ** 2, dtype=complex))

    monkeypatch.setattr(System, "_propagate", lambda self, src, tgt, d: fake_po(src, tgt, d))
    out = s.convergeOnTarget("beam", "scat", "tgt")
    assert list(out) == [22, 11]
    assert list(s.system["scat"]["gridsize"]) == [22, 11]
    assert "_conv_target" not in s.system and "_conv_scatterer" not in s.system
    assert "_S_conv_scat" not in s.scalarfields and "_S_conv_target" not in s.scalarfields
    assert list(s.system["tgt"]["gridsize"]) == [7, 7]


def _dish_system():
    """Feed plane with a Gaussian field under a small paraboloid; the
    scattered field is looked at on the feed plane itself."""
    s = System()
    s.addPlane({"name": "feed", "gmode": "xy", "lims_x": [-1.5, 1.5], "lims_y": [-1.5, 1.5],
                "gridsize": [15, 15], "flip": True})
    s.system["feed"]["transf"] = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 12.0], [0, 0, 0, 1]])
    s.addParabola({"name": "dish", "pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 12.0],
                   "gmode": "uv", "lims_u": [0, 6], "lims_v": [0, 360], "gridsize": [25, 24]})
    x = np.linspace(-1.5, 1.5, 15)
    X, Y = np.meshgrid(x, x, indexing="ij")
    amp = np.exp(-(X ** 2 + Y ** 2) / 0.8 ** 2) + 0j
    z = np.zeros_like(amp)
    f = T.fields(amp, z, z, z, amp / 2, z)
    f.setMeta("feed", 2 * np.pi)
    s.fields["beam"] = f
    return s


def _checker_po(port):
    def po(source, target, d):
        out, _, _ = OB.propagate(port, d["mode"], source, target, d["s_scalarfield"], d["k"],
                                 epsilon=d.get("epsilon", 1.0), exp=d.get("exp", "fwd"))
        return out
    return po


def test_converge_loops_with_checker_propagation(port, monkeypatch):
    """Dry run of both loops on the CPU checker: they terminate on a real
    propagation and give grids the GPU test below must reproduce."""
    s = _dish_system()
    monkeypatch.setattr(System, "_propagate", lambda self, src, tgt, d: _checker_po(port)(src, tgt, d))
    monkeypatch.setattr(B, "generate_grid", lambda r, t=True, sp=True: OB.generate_grid(port, r, t, sp))
    a = s.autoConverge("beam", "dish")
    assert a.dtype.kind == "i" and list(a) == [189, 189]
    assert list(s.convergeOnTarget("beam", "dish", "feed")) == [45, 11]
    assert list(s.convergeOnTarget("beam", "dish", "feed", tol=1e-3)) == [90, 12]


@pytest.mark.gpu
def test_converge_loops_gpu(hip, port, monkeypatch):
    s = _dish_system()
    gpu_auto = s.autoConverge("beam", "dish")
    assert type(gpu_auto) is np.ndarray
    assert list(gpu_auto) == list(s.system["dish"]["gridsize"])
    s.system["dish"]["gridsize"] = np.array([25, 24])
    gpu_conv = s.convergeOnTarget("beam", "dish", "feed")
    assert list(gpu_conv) == list(s.system["dish"]["gridsize"])
    # a tighter tolerance runs the grid up to 160 along u and bisects back
    gpu_tight = s.convergeOnTarget("beam", "dish", "feed", tol=1e-3)

    # the same loops on the CPU checker's propagation
    r = _dish_system()
    monkeypatch.setattr(System, "_propagate", lambda self, src, tgt, d: _checker_po(port)(src, tgt, d))
    monkeypatch.setattr(B, "generate_grid", lambda g, t=True, sp=True: OB.generate_grid(port, g, t, sp))
    assert list(r.autoConverge("beam", "dish")) == list(gpu_auto)
    r.system["dish"]["gridsize"] = np.array([25, 24])
    assert list(r.convergeOnTarget("beam", "dish", "feed")) == list(gpu_conv)
    assert list(r.convergeOnTarget("beam", "dish", "feed", tol=1e-3)) == list(gpu_tight)
