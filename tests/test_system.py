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
        s.runPO({"t_name": "dish", "mode": "JM", "s_current": "feedJM", "name_JM": "a", "device": "CPU"})
    with pytest.raises(RunPOError):
        s.runPO({"t_name": "dish", "mode": "JM", "s_current": "feedJM", "name_JM": "a",
                 "precision": "half"})


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
