"""The reference's own System tests (tests/test_SystemPO_RT.py:47-212 with
the templates of tests/TestTemplates.py) run through pypo_amd.system.System
with the reference's default device="CPU": JM, EH and JMEH fwd and bwd onto
the manual xy paraboloid, FF onto the AoE plane, EHP followed by the hybrid
step (interpolated with a start point, and without interpolation for one
component), scalar mode, a ray trace of the tubular frame, autoConverge onto
the manual hyperboloids, mergeBeams and interpBeam.  As there,
the result types are checked; here the values are also compared with the
CPU checker wherever the checker's value is finite (the templates put a
source and a target point at the origin, where both give NaN)."""

import numpy as np
import pytest

from helpers import flat
from oracle import bind as OB
from pypo_amd import types as T
from pypo_amd.system import System

pytestmark = pytest.mark.gpu

PLANE_XY = {"name": "testPlane_xy", "gmode": "xy", "lims_x": np.array([-1, 1]), "lims_y": np.array([-1, 1]),
            "gridsize": np.array([13, 13])}
PLANE_AOE = {"name": "testPlane_AoE", "gmode": "AoE", "lims_Az": np.array([-1, 1]),
             "lims_El": np.array([-1, 1]), "gridsize": np.array([13, 13])}
PARA_MAN_XY = {"name": "testParaboloid_man_xy", "pmode": "manual", "coeffs": np.array([1, 1]), "gmode": "xy",
               "lims_x": np.array([-1, 1]), "lims_y": np.array([-1, 1]), "gridsize": np.array([13, 13])}
GPO_FIELD = {"name": "testGaussField", "lam": 0.01, "w0x": 1, "w0y": 1, "n": 1, "E0": 1, "dxyz": 1,
             "pol": np.array([1, 0, 0])}
PS_UFIELD = {"name": "testPS_UField", "lam": 1, "E0": 1, "pol": np.array([1, 0, 0])}
TUBE = {"name": "testTubeRTframe", "nRays": 7, "nRing": 7, "angx0": 1, "angy0": 1, "x0": 1, "y0": 1}


@pytest.fixture
def s(hip):
    s = System()
    s.addPlane(dict(PLANE_XY))
    s.createPointSource(dict(PS_UFIELD), PLANE_XY["name"])
    s.createGaussian(dict(GPO_FIELD), PLANE_XY["name"])
    s.createScalarGaussian(dict(GPO_FIELD), PLANE_XY["name"])
    s.addPlane(dict(PLANE_AOE))
    s.addParabola(dict(PARA_MAN_XY))
    s.createTubeFrame(dict(TUBE))
    return s


def _close(got, ref, tol):
    got, ref = np.asarray(got), np.asarray(ref)
    ok = np.isfinite(ref)
    assert np.array_equal(ok, np.isfinite(got))
    pk = np.max(np.abs(ref[ok]))
    assert np.max(np.abs(got[ok] - ref[ok])) <= tol * pk


def _ref(port, s, mode, src, tgt, direction="fwd"):
    obj = s.currents[src] if mode != "scalar" else s.scalarfields[src]
    surf = s.system[obj.surf]
    out, _, _ = OB.propagate(port, mode, surf, s.system[tgt], obj, obj.k, epsilon=10, exp=direction)
    return out


@pytest.mark.parametrize("direction", ["fwd", "bwd"])
def test_runPO_JM_EH_JMEH(s, port, direction):
    tgt = PARA_MAN_XY["name"]
    base = {"t_name": tgt, "s_current": GPO_FIELD["name"], "epsilon": 10, "exp": direction, "device": "CPU"}
    s.runPO(dict(base, name_JM="test_JM", mode="JM"))
    assert type(s.currents["test_JM"]) is T.currents
    _close(flat(s.currents["test_JM"]), flat(_ref(port, s, "JM", GPO_FIELD["name"], tgt, direction)), 1e-5)
    s.runPO(dict(base, name_EH="test_EH", mode="EH"))
    assert type(s.fields["test_EH"]) is T.fields
    _close(flat(s.fields["test_EH"]), flat(_ref(port, s, "EH", GPO_FIELD["name"], tgt, direction)), 1e-5)
    s.runPO(dict(base, name_JM="test_JM", name_EH="test_EH", mode="JMEH"))
    assert type(s.currents["test_JM"]) is T.currents and type(s.fields["test_EH"]) is T.fields


def test_runPO_FF_and_scalar(s, port):
    s.runPO({"t_name": PLANE_AOE["name"], "s_current": GPO_FIELD["name"], "epsilon": 10, "exp": "fwd",
             "device": "CPU", "name_EH": "test_EH", "mode": "FF"})
    assert type(s.fields["test_EH"]) is T.fields
    _close(flat(s.fields["test_EH"]), flat(_ref(port, s, "FF", GPO_FIELD["name"], PLANE_AOE["name"])), 1e-5)
    s.runPO({"t_name": PARA_MAN_XY["name"], "s_scalarfield": GPO_FIELD["name"], "epsilon": 10, "exp": "fwd",
             "device": "CPU", "name_field": "test_S", "mode": "scalar"})
    assert type(s.scalarfields["test_S"]) is T.scalarfield
    _close(flat(s.scalarfields["test_S"]), flat(_ref(port, s, "scalar", GPO_FIELD["name"], PARA_MAN_XY["name"])),
           1e-5)


@pytest.mark.parametrize("interp,comp,start", [(True, None, np.ones(3)), (False, "Ex", None)])
def test_runPO_EHP_hybrid(s, interp, comp, start):
    s.runPO({"t_name": PARA_MAN_XY["name"], "s_current": GPO_FIELD["name"], "epsilon": 10, "exp": "fwd",
             "device": "CPU", "name_EH": "test_EH", "name_P": "test_P", "mode": "EHP"})
    assert type(s.fields["test_EH"]) is T.fields and type(s.frames["test_P"]) is T.frame
    s.runHybridPropagation({"fr_in": "test_P", "t_name": PARA_MAN_XY["name"], "field_in": "test_EH",
                            "fr_out": "test_fr_out", "field_out": "test_field_out", "interp": interp,
                            "comp": comp, "start": start})
    assert type(s.fields["test_field_out"]) is T.fields
    assert type(s.frames["test_fr_out"]) is T.frame


def test_runRT(s, port):
    fr = s.frames[TUBE["name"]]
    s.addFrame("shifted", T.frame(fr.size, fr.x, fr.y, np.asarray(fr.z) - 1, fr.dx, fr.dy, fr.dz))
    s.runRayTracer({"fr_in": "shifted", "t_name": PARA_MAN_XY["name"], "device": "CPU", "fr_out": "test_fr",
                    "t0": 1})
    assert type(s.frames["test_fr"]) is T.frame
    ref = OB.trace(port, s.system[PARA_MAN_XY["name"]], s.frames["shifted"], tol=1e-3, t0=1, max_iter=50)
    got = s.frames["test_fr"]
    # tol 1e-3 (the reference default): a stopping decision at the threshold could
    # differ by one Newton step, which moves a hit point by far less than 1e-6
    np.testing.assert_allclose(np.stack([got.x, got.y, got.z]), np.stack([ref.x, ref.y, ref.z]), atol=1e-6)


def test_interp_linear_and_cubic(s):
    """interpFrame's other griddata methods run on the host: "linear"
    interpolates the carried field, "cubic" raises for 3-D points in scipy
    as it does for the reference."""
    s.runPO({"t_name": PARA_MAN_XY["name"], "s_current": GPO_FIELD["name"], "mode": "EHP", "exp": "fwd",
             "name_EH": "eh", "name_P": "P", "precision": "fp64"})
    fr = s.frames["P"]
    # a field sampled on a tilted plane of points, linear in x and y: linear interpolation is exact
    n = 200
    rng = np.random.default_rng(0)
    x, y = rng.uniform(-1.5, 1.5, n), rng.uniform(-1.5, 1.5, n)
    z = 0.3 * x - 0.2 * y + rng.uniform(-0.4, 0.4, n)
    s.addFrame("cloud", T.frame(n, x, y, z, np.zeros(n), np.zeros(n), np.ones(n)))
    lin = (2.0 * x - 3.0 * y + 0.5 * z) + 1j * (x + y)
    s.fields["cloud_f"] = T.fields(lin, lin, lin, lin, lin, lin)
    s.fields["cloud_f"].setMeta(PLANE_XY["name"], 1.0)
    out = s.interpFrame("cloud", "cloud_f", PLANE_XY["name"], "lin_out", comp="Ex", method="linear")
    g = np.meshgrid(np.linspace(-1, 1, 13), np.linspace(-1, 1, 13), indexing="ij")
    want = (2.0 * g[0] - 3.0 * g[1]) + 1j * (g[0] + g[1])
    ok = np.isfinite(out)
    assert ok.sum() > 100
    np.testing.assert_allclose(out[ok], want[ok], atol=1e-9)
    with pytest.raises(ValueError):
        s.interpFrame("cloud", "cloud_f", PLANE_XY["name"], "cub_out", method="cubic")
    assert fr.size == 169


def test_mergeBeams(s):
    """test_SystemPO_RT.py:198-202: merging the Gaussian and the point source on
    the plane stores a field and a current under the merged name (here also:
    the merge is their sum)."""
    s.mergeBeams(GPO_FIELD["name"], PS_UFIELD["name"], merged_name="test")
    s.mergeBeams(GPO_FIELD["name"], PS_UFIELD["name"], obj="currents", merged_name="test")
    assert "test" in s.fields and "test" in s.currents
    assert type(s.fields["test"]) is T.fields and type(s.currents["test"]) is T.currents
    want = flat(s.fields[GPO_FIELD["name"]]) + flat(s.fields[PS_UFIELD["name"]])
    np.testing.assert_allclose(flat(s.fields["test"]), want, rtol=0, atol=1e-15 * np.abs(want).max())


@pytest.mark.parametrize("name", [GPO_FIELD["name"], PS_UFIELD["name"]])
def test_interpBeam(s, name):
    """test_SystemPO_RT.py:204-212: interpolating a source's field and currents
    onto a 101 x 101 copy of its plane stores both under name + "_interp"."""
    s.interpBeam(name, np.array([101, 101]), obj="fields")
    s.interpBeam(name, np.array([101, 101]), obj="currents")
    assert f"{name}_interp" in s.fields and f"{name}_interp" in s.currents
    assert s.fields[f"{name}_interp"].shape == (101, 101)
    assert s.system[PLANE_XY["name"] + "_interp"]["gridsize"].tolist() == [101, 101]


HYP_MAN_XY = {"name": "testHyperboloid_man_xy", "pmode": "manual", "coeffs": np.array([1, 1, 1]), "ecc": 1.1,
              "gmode": "xy", "lims_x": np.array([-1, 1]), "lims_y": np.array([-1, 1]),
              "gridsize": np.array([13, 13])}
HYP_MAN_UV = {"name": "testHyperboloid_man_uv", "pmode": "manual", "coeffs": np.array([1, 1, 1]), "ecc": 1.1,
              "gmode": "uv", "lims_u": np.array([0, 1]), "lims_v": np.array([0, 360]),
              "gridsize": np.array([13, 12])}


@pytest.mark.parametrize("name", [GPO_FIELD["name"], PS_UFIELD["name"]])
def test_autoConverge(s, name):
    """test_SystemPO_RT.py:187-196: autoConverge from each source onto the
    manual xy and uv hyperboloids returns a gridsize array and sets it on
    the surface."""
    s.addHyperbola(dict(HYP_MAN_XY))
    s.addHyperbola(dict(HYP_MAN_UV))
    for hyp in (HYP_MAN_XY, HYP_MAN_UV):
        gridsize = s.autoConverge(name, hyp["name"])
        assert type(gridsize) is np.ndarray
        assert gridsize[0] == s.system[hyp["name"]]["gridsize"][0]
        assert gridsize[1] == s.system[hyp["name"]]["gridsize"][1]
