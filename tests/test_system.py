"""System.runPO / System.runRayTracer: the reference's dictionary API on the GPU.

Mirrors the reference's tests/test_SystemPO_RT.py (result types for every
mode, fwd and bwd, invalid dictionaries raise RunPOError / RunRTError) and
adds value checks against the CPU checker."""

import numpy as np
import pytest

from helpers import flat, group_err
from oracle import bind as OB
from pypo_amd import types as T
from pypo_amd.system import RunPOError, RunRTError, System


def make_system():
    s = System()
    s.addPlane({"name": "feed", "gmode": "xy", "lims_x": [-2, 2], "lims_y": [-2, 2],
                "gridsize": [17, 17], "flip": True})
    s.system["feed"]["transf"] = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 40.0], [0, 0, 0, 1]])
    s.addParabola({"name": "dish", "pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 40.0],
                   "gmode": "uv", "lims_u": [0, 30], "lims_v": [0, 360], "gridsize": [25, 24]})
    s.addPlane({"name": "ff", "gmode": "AoE", "lims_Az": [-3, 3], "lims_El": [-3, 3],
                "gridsize": [11, 11]})
    x = np.linspace(-2, 2, 17)
    X, Y = np.meshgrid(x, x, indexing="ij")
    amp = np.exp(-(X ** 2 + Y ** 2) / 1.5 ** 2) + 0j
    z = np.zeros_like(amp)
    s.addCurrents("feedJM", T.currents(-amp, z, z, z, amp, z), "feed", 2 * np.pi / 0.8)
    s.addScalarField("feedS", T.scalarfield(amp), "feed", 2 * np.pi / 0.8)
    return s


def test_invalid_runPODict_raises():
    s = make_system()
    with pytest.raises(RunPOError):
        s.runPO({})
    with pytest.raises(RunPOError):
        s.runPO({"t_name": "dish", "mode": "XX", "s_current": "feedJM", "name_JM": "a"})
    with pytest.raises(RunPOError):
        s.runPO({"t_name": "nope", "mode": "JM", "s_current": "feedJM", "name_JM": "a"})
    with pytest.raises(RunPOError):
        s.runPO({"t_name": "dish", "mode": "JM", "s_current": "nope", "name_JM": "a"})
    with pytest.raises(RunPOError):
        s.runPO({"t_name": "dish", "mode": "EHP", "s_current": "feedJM", "name_EH": "a"})
    with pytest.raises(RunPOError):
        s.runPO({"t_name": "dish", "mode": "JM", "s_current": "feedJM", "name_JM": "a",
                 "precision": "half"})


def test_device_key_is_an_alias():
    """device "CPU" (the reference's default) and unknown names are accepted
    and run on the GPU, as the reference accepts them (Checks.py:1566-1591)."""
    s = make_system()
    for dev in ("CPU", "GPU", "HIP", "TPU", None):
        d = {"t_name": "dish", "mode": "JM", "s_current": "feedJM", "name_JM": "a"}
        if dev is not None:
            d["device"] = dev
        s._check_runPODict(d)
        assert d["device"] == "GPU" and d["nThreads"] == 256
    s.addFrame("f", T.frame(1, [0.0], [0.0], [50.0], [0.0], [0.0], [-1.0]))
    d = {"fr_in": "f", "t_name": "dish", "fr_out": "o", "device": "CPU"}
    s._check_runRTDict(d)
    assert d["device"] == "GPU"


def test_invalid_runRTDict_raises():
    s = make_system()
    with pytest.raises(RunRTError):
        s.runRayTracer({})
    with pytest.raises(RunRTError):
        s.runRayTracer({"fr_in": "nope", "t_name": "dish", "fr_out": "o"})
    s.addFrame("f", T.frame(1, [0.0], [0.0], [50.0], [0.0], [0.0], [-1.0]))
    with pytest.raises(RunRTError):
        s.runRayTracer({"fr_in": "f", "t_name": "nope", "fr_out": "o"})


def test_name_suffixing_without_override():
    s = make_system()
    s.override = False
    s.addCurrents("x", s.currents["feedJM"], "feed", 1.0)
    s.addCurrents("x_1", s.currents["feedJM"], "feed", 1.0)
    d = {"t_name": "dish", "mode": "JM", "s_current": "feedJM", "name_JM": "x"}
    s._check_runPODict(d)
    assert d["name_JM"] == "x_2"


@pytest.mark.gpu
@pytest.mark.parametrize("exp", ["fwd", "bwd"])
def test_runPO_all_modes(hip, port, exp):
    s = make_system()
    k = s.currents["feedJM"].k
    feed, dish, ff = s.system["feed"], s.system["dish"], s.system["ff"]
    s.runPO({"t_name": "dish", "mode": "JM", "s_current": "feedJM", "name_JM": "dJM", "exp": exp})
    assert type(s.currents["dJM"]) is T.currents and s.currents["dJM"].surf == "dish"
    s.runPO({"t_name": "dish", "mode": "EH", "s_current": "feedJM", "name_EH": "dEH", "exp": exp})
    assert type(s.fields["dEH"]) is T.fields
    s.runPO({"t_name": "dish", "mode": "JMEH", "s_current": "feedJM", "name_JM": "dJM2",
             "name_EH": "dEH2", "exp": exp})
    assert type(s.currents["dJM2"]) is T.currents and type(s.fields["dEH2"]) is T.fields
    s.runPO({"t_name": "dish", "mode": "EHP", "s_current": "feedJM", "name_EH": "dEHr",
             "name_P": "dP", "exp": exp, "precision": "fp64"})
    assert type(s.fields["dEHr"]) is T.fields and type(s.frames["dP"]) is T.frame
    s.runPO({"t_name": "ff", "mode": "FF", "s_current": "dJM", "name_EH": "ffEH", "exp": exp})
    assert type(s.fields["ffEH"]) is T.fields
    s.runPO({"t_name": "dish", "mode": "scalar", "s_scalarfield": "feedS", "name_field": "dS",
             "exp": exp})
    assert type(s.scalarfields["dS"]) is T.scalarfield
    # values against the CPU checker
    jm_ref, _, _ = OB.propagate(port, "JM", feed, dish, s.currents["feedJM"], k, exp=exp)
    assert group_err(flat(s.currents["dJM"]), flat(jm_ref)) < 1e-5
    ff_ref, _, _ = OB.propagate(port, "FF", dish, ff, s.currents["dJM"], k, exp=exp)
    assert group_err(flat(s.fields["ffEH"]), flat(ff_ref)) < 1e-5
    ehp_ref, _, _ = OB.propagate(port, "EHP", feed, dish, s.currents["feedJM"], k, exp=exp)
    Pr = ehp_ref[1]
    fr = s.frames["dP"]
    np.testing.assert_allclose(np.stack([fr.dx, fr.dy, fr.dz]),
                               np.stack([Pr.x.ravel(), Pr.y.ravel(), Pr.z.ravel()]), atol=1e-9)
    assert group_err(flat(s.fields["dEHr"]), flat(ehp_ref[0])) < 1e-9


@pytest.mark.gpu
def test_runRayTracer_and_frame_metadata(hip, port):
    s = make_system()
    n = 400
    rng = np.random.default_rng(1)
    x, y = rng.uniform(-25, 25, n), rng.uniform(-25, 25, n)
    s.addFrame("rays", T.frame(n, x, y, np.full(n, 60.0), np.zeros(n), np.zeros(n), -np.ones(n)))
    s.runRayTracer({"fr_in": "rays", "t_name": "dish", "fr_out": "hit", "tol": 1e-12, "t0": 10.0})
    fr = s.frames["hit"]
    assert type(fr) is T.frame
    ref = OB.trace(port, s.system["dish"], s.frames["rays"], None, 1e-12, 10.0, max_iter=50)
    np.testing.assert_allclose(np.stack([fr.x, fr.y, fr.z]), np.stack([ref.x, ref.y, ref.z]), atol=1e-9)
    np.testing.assert_allclose(fr.pos, [x.mean(), y.mean(), fr.z.mean()], atol=1e-9)
    # reflected axis-parallel rays point at the focus on average: tilt ~ +z
    assert fr.ori[2] > 0.9


def test_invalid_hybridDict_raises():
    from pypo_amd.system import HybridPropError
    s = make_system()
    with pytest.raises(HybridPropError):
        s.runHybridPropagation({"fr_in": "none", "t_name": "dish", "fr_out": "o", "field_in": "x",
                                "field_out": "y"})


def test_calc_beam_pairs():
    """w0 and z of a vector Gaussian from the other pairs (Checks.calc_beam):
    they reproduce the beam width w(z) = w0 sqrt(1 + (z/zR)^2) and the phase
    radius R(z) = z (1 + (zR/z)^2), zR = pi w0^2 / lam."""
    from pypo_amd.system import calc_beam
    lam, w0, z = 1.0, 2.0, 30.0
    zR = np.pi * w0 ** 2 / lam
    w = w0 * np.sqrt(1 + (z / zR) ** 2)
    R = z * (1 + (zR / z) ** 2)
    for given in ({"w0": w0, "w": w}, {"z": z, "R": R}, {"w": w, "R": R}):
        d = dict(given, lam=lam)
        calc_beam(d)
        assert d["w0"] == pytest.approx(w0, rel=1e-9)
        assert d["z"] == pytest.approx(z, rel=1e-9)
    with pytest.raises(ValueError):
        calc_beam({"lam": 1.0, "w0": 1.0})


@pytest.mark.gpu
def test_system_beams_and_frames(hip, port):
    """createTubeFrame / createGaussian / createPointSource / createUniformSource
    / createScalarGaussian through System, against the CPU checker."""
    import cases as C
    s = make_system()
    rt = C.beam_cases()["rt_tube_3ring"]["params"]
    fr = s.createTubeFrame(dict(rt, name="tube"))
    ref = OB.make_rt_frame(port, rt)
    np.testing.assert_allclose(np.stack([fr.x, fr.y, fr.dz]), np.stack([ref.x, ref.y, ref.dz]), atol=1e-13)
    g = {"name": "gauss", "lam": 0.8, "w0x": 1.0, "w0y": 1.2, "n": 1.0, "E0": 1.0, "dxyz": 0.0,
         "pol": np.array([0.0, 1.0, 0.0])}
    f, c = s.createGaussian(g, "feed")
    fref, cref = OB.make_gauss(port, g, s.system["feed"])
    assert group_err(flat(f), flat(fref)) <= 1e-12 and group_err(flat(c), flat(cref)) <= 1e-12
    assert s.currents["gauss"].surf == "feed"
    ps, pc = s.createPointSource({"name": "ps", "lam": 0.8, "E0": 2.0, "phase": 0.3,
                                  "pol": np.array([1.0, 0.0, 0.0])}, "feed")
    assert np.count_nonzero(flat(ps)) == 1
    assert group_err(flat(pc), flat(OB.calc_currents(port, ps, s.system["feed"], 1))) <= 1e-14
    us, uc = s.createUniformSource({"name": "us", "lam": 0.8, "E0": 1.0, "phase": 0.0,
                                    "pol": np.array([0.0, 1.0, 0.0])}, "feed")
    assert group_err(flat(uc), flat(OB.calc_currents(port, us, s.system["feed"], 1))) <= 1e-14
    sg = {"name": "sg", "lam": 0.8, "w0x": 1.0, "w0y": 1.0, "n": 1.0, "E0": 1.0, "dxyz": 0.0}
    sf = s.createScalarGaussian(sg, "feed")
    sref = OB.make_scalar_gauss(port, sg, s.system["feed"])
    assert np.max(np.abs(np.ravel(sf.S) - np.ravel(sref.S))) <= 1e-12 * np.max(np.abs(sref.S))
    # the Gaussian's currents drive the PO chain like any other currents
    out = s.runPO({"t_name": "dish", "name_JM": "jm", "s_current": "gauss", "mode": "JM", "precision": "fp64"})
    ref_jm, _, _ = OB.propagate(port, "JM", s.system["feed"], s.system["dish"], cref, 2 * np.pi / 0.8)
    assert group_err(flat(out), flat(ref_jm)) <= 1e-10
    lens = s.calcRayLen("tube", "tube")
    assert np.all(lens[0] == 0) and len(lens) == 2


@pytest.mark.gpu
def test_runPO_chains_on_device(hip, monkeypatch):
    """JM then FF through System: the JM currents stay in device memory and
    feed the FF call without any device-to-host copy; both results are
    bit-identical to the drop-in path's (same kernels on the same inputs);
    reading a component copies the values to the host once."""
    from pypo_amd import bind as B
    s = make_system()
    k = s.currents["feedJM"].k
    copies = []
    real = T._DeviceBacked._to_host

    def spy(self):
        if self.device_rows() is not None:
            copies.append(type(self).__name__)
        return real(self)

    monkeypatch.setattr(T._DeviceBacked, "_to_host", spy)
    s.runPO({"t_name": "dish", "mode": "JM", "s_current": "feedJM", "name_JM": "dJM"})
    s.runPO({"t_name": "ff", "mode": "FF", "s_current": "dJM", "name_EH": "ffEH"})
    assert copies == []
    assert s.currents["dJM"].device_rows() is not None and s.fields["ffEH"].device_rows() is not None
    jm_host = B.PyPO_HIPd(s.system["feed"], s.system["dish"], {"mode": "JM", "k": k, "s_current": s.currents["feedJM"]})
    ff_host = B.PyPO_HIPd(s.system["dish"], s.system["ff"], {"mode": "FF", "k": k, "s_current": jm_host})
    np.testing.assert_array_equal(flat(s.fields["ffEH"]), flat(ff_host))
    assert copies == ["fields"] and s.fields["ffEH"].device_rows() is None
    np.testing.assert_array_equal(flat(s.currents["dJM"]), flat(jm_host))
    assert copies == ["fields", "currents"]
    # a surface whose parameters change gets a new grid
    s.system["ff"]["lims_Az"] = np.array([-2.0, 2.0])
    s.runPO({"t_name": "ff", "mode": "FF", "s_current": "dJM", "name_EH": "ffEH2"})
    ff2 = B.PyPO_HIPd(s.system["dish"], s.system["ff"], {"mode": "FF", "k": k, "s_current": jm_host})
    np.testing.assert_array_equal(flat(s.fields["ffEH2"]), flat(ff2))


@pytest.mark.gpu
def test_merge_interp_and_remove_beams(hip):
    """mergeBeams, interpBeam and the remove* / setOverride / copyObj helpers
    (the reference's test_mergeBeams and test_interpBeam,
    tests/test_SystemPO_RT.py:198-214, plus values): a merge is the sum of
    its beams and stays on the device; beams on different surfaces are
    refused; linear interpolation of a field linear in (x, y) returns that
    field on the new grid."""
    from pypo_amd.system import CurrentNameError, FieldNameError, FrameNameError, MergeBeamError
    s = make_system()
    g = {"name": "gauss", "lam": 0.8, "w0x": 1.0, "w0y": 1.2, "n": 1.0, "E0": 1.0, "dxyz": 0.0,
         "pol": np.array([0.0, 1.0, 0.0])}
    s.createGaussian(g, "feed")
    s.createUniformSource({"name": "us", "lam": 0.8, "E0": 1.0, "phase": 0.3, "pol": np.array([1.0, 0.0, 0.0])},
                          "feed")
    for obj in ("fields", "currents"):
        s.mergeBeams("gauss", "us", obj=obj, merged_name="test")
        pool = getattr(s, obj)
        assert pool["test"].device_rows() is not None and pool["test"].surf == "feed"
        want = flat(pool["gauss"]) + flat(pool["us"])
        np.testing.assert_allclose(flat(pool["test"]), want, rtol=0, atol=1e-15 * np.max(np.abs(want)))
    s.createUniformSource({"name": "far", "lam": 0.8, "E0": 1.0, "phase": 0.0, "pol": np.array([1.0, 0.0, 0.0])},
                          "dish")
    with pytest.raises(MergeBeamError):
        s.mergeBeams("gauss", "far")
    with pytest.raises(FieldNameError):
        s.mergeBeams("gauss", "nope")
    # a field linear in (x, y) on the 17 x 17 feed plane, interpolated onto 101 x 101
    grid = s.generateGrids("feed")
    x, y = np.asarray(grid.x), np.asarray(grid.y)
    lin = [(a + 0.5 * x - 2.0 * y) + 1j * (b - x + 0.25 * y) for a, b in
           ((1, 0), (0, 1), (2, 2), (-1, 0), (0, -3), (0.5, 0.5))]
    s.fields["lin"] = T.fields(*lin)
    s.fields["lin"].setMeta("feed", 2 * np.pi / 0.8)
    s.interpBeam("lin", np.array([101, 101]))
    assert s.system["feed_interp"]["gridsize"].tolist() == [101, 101]
    assert s.fields["lin_interp"].surf == "feed_interp"
    g2 = s.generateGrids("feed_interp")
    x2, y2 = np.asarray(g2.x), np.asarray(g2.y)
    inner = (slice(1, -1), slice(1, -1))
    for i, (a, b) in enumerate(((1, 0), (0, 1), (2, 2), (-1, 0), (0, -3), (0.5, 0.5))):
        want = (a + 0.5 * x2 - 2.0 * y2) + 1j * (b - x2 + 0.25 * y2)
        got = s.fields["lin_interp"].components()[i]
        assert np.all(np.isfinite(got[inner]))
        ok = np.isfinite(got)
        np.testing.assert_allclose(got[ok], want[ok], rtol=0, atol=1e-12)
    s.interpBeam("gauss", np.array([31, 29]), obj="currents")
    assert s.currents["gauss_interp"].shape == (31, 29)
    # remove*, with the reference's name errors
    s.removeField("lin")
    s.removeCurrent("gauss_interp")
    with pytest.raises(FieldNameError):
        s.removeField("lin")
    with pytest.raises(CurrentNameError):
        s.removeCurrent("gauss_interp")
    with pytest.raises(FrameNameError):
        s.removeFrame("nothing")
    s.createTubeFrame({"name": "tube", "nRays": 4, "nRing": 2, "angx0": 1.0, "angy0": 1.0, "x0": 2.0, "y0": 2.0})
    s.removeFrame("tube")
    assert "tube" not in s.frames
    # copyObj copies a device-resident result; setOverride(False) suffixes names
    c = s.copyObj(s.fields["test"])
    np.testing.assert_array_equal(flat(c), flat(s.fields["test"]))
    s.setOverride(False)
    s.mergeBeams("gauss", "us", merged_name="test")     # merge stores under the name given, as there
    s.createUniformSource({"name": "us", "lam": 0.8, "E0": 1.0, "phase": 0.0, "pol": np.array([1.0, 0.0, 0.0])},
                          "feed")
    assert "us_1" in s.fields
    s.setLoggingVerbosity(False)
    s.setLoggingVerbosity(True)
