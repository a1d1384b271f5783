"""Beam initialisation on the GPU (beam.hip through the drop-in symbols
makeGauss / makeGaussBeam / makeScalarGauss / calcCurrents / makeRTframe /
makeGRTframe and their pypo_hip_*_dev twins) against the golden vectors of
the compiled reference (tests/golden/beams.npz), FP64: 1e-12 of each
field's peak; ray frames to 1e-13 mm.  makeGRTframe is random in the
reference (std::random_device), so it is checked by distribution: moments of
the truncated Gaussian and a two-sample Kolmogorov-Smirnov test against the
reference's own sample."""

import numpy as np
import pytest

import cases as C
from helpers import flat, group_err
from pypo_amd import beam as PB
from pypo_amd.geometry import gridsize

pytestmark = pytest.mark.gpu
GOLD = np.load("tests/golden/beams.npz")


def run_hip(c):
    if c["kind"] == "gauss":
        f, cur = PB.makeGauss(c["params"], c["source"])
        return {"field": flat(f), "current": flat(cur)}
    if c["kind"] == "vec":
        f, cur = PB.makeGaussBeam(c["params"], c["source"])
        return {"field": flat(f), "current": flat(cur)}
    if c["kind"] == "scalar":
        return {"field": np.ravel(PB.makeScalarGauss(c["params"], c["source"]).S)}
    if c["kind"] == "currents":
        fields = C.random_fields(gridsize(c["source"]), c["seed"])
        return {"current": flat(PB.calcCurrents(fields, c["source"], c["mode"]))}
    fr = PB.makeRTframe(c["params"])
    return {"frame": np.stack([getattr(fr, a) for a in ("x", "y", "z", "dx", "dy", "dz")])}


@pytest.mark.parametrize("name", sorted(C.beam_cases()))
def test_dropin_beams_match_reference(hip, name):
    c = C.beam_cases()[name]
    for key, val in run_hip(c).items():
        ref = GOLD[f"{name}/{key}"]
        if key == "frame":
            np.testing.assert_allclose(val, ref, rtol=0, atol=1e-13)
        elif val.ndim == 1:
            assert np.max(np.abs(val - ref)) <= 1e-12 * np.max(np.abs(ref))
        else:
            assert group_err(val, ref) <= 1e-12, key


def test_device_variants_match_dropin(hip):
    import torch
    from pypo_amd import device as D
    cs = C.beam_cases()
    c = cs["gauss_astig"]
    g = D.generate_grid(c["source"])
    f, cur = PB.make_gauss(c["params"], g)
    f0, cur0 = PB.makeGauss(c["params"], c["source"])
    np.testing.assert_array_equal(D.tensor_to_complex6(f), flat(f0))
    np.testing.assert_array_equal(D.tensor_to_complex6(cur), flat(cur0))
    c = cs["vec_gauss_behind"]
    g = D.generate_grid(c["source"])
    f, cur = PB.make_gauss_beam(c["params"], g)
    assert group_err(D.tensor_to_complex6(f), GOLD["vec_gauss_behind/field"]) <= 1e-12
    # currents from the device fields, every mode, against the drop-in call
    fobj = _fields_obj(D.tensor_to_complex6(f), gridsize(c["source"]))
    for mode in range(4):
        jm = PB.calc_currents(f, g, mode)
        np.testing.assert_array_equal(D.tensor_to_complex6(jm), flat(PB.calcCurrents(fobj, c["source"], mode)))
    s = PB.make_scalar_gauss(cs["scalar_astig"]["params"], D.generate_grid(cs["scalar_astig"]["source"]))
    ref = GOLD["scalar_astig/field"]
    got = s[0].cpu().numpy() + 1j * s[1].cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    fr = PB.make_rt_frame(cs["rt_tube_3ring"]["params"]).cpu().numpy()
    np.testing.assert_allclose(fr, GOLD["rt_tube_3ring/frame"], rtol=0, atol=1e-13)
    torch.cuda.synchronize()


def _fields_obj(c6, shape):
    from pypo_amd import types as T
    return T.fields(*[c6[q].reshape(shape) for q in range(6)])


def _ks_2samp(a, b):
    """Two-sample Kolmogorov-Smirnov statistic and its asymptotic p-value."""
    a, b = np.sort(a), np.sort(b)
    grid = np.concatenate([a, b])
    d = np.max(np.abs(np.searchsorted(a, grid, side="right") / a.size -
                      np.searchsorted(b, grid, side="right") / b.size))
    en = np.sqrt(a.size * b.size / (a.size + b.size))
    lam = (en + 0.12 + 0.11 / en) * d
    p = 2 * sum((-1) ** (k - 1) * np.exp(-2 * k * k * lam * lam) for k in range(1, 101))
    return d, min(max(p, 0.0), 1.0)


def test_gaussian_ray_frame_distribution(hip):
    c = dict(C.grt_case(), nRays=20001)
    fr = PB.makeGRTframe(c)
    a = np.stack([fr.x, fr.y, fr.z, fr.dx, fr.dy, fr.dz])
    assert (a[0, 0], a[1, 0], a[5, 0]) == (0.0, 0.0, 1.0)
    n = a.shape[1] - 1
    np.testing.assert_allclose(a[3] ** 2 + a[4] ** 2 + a[5] ** 2, 1.0, rtol=1e-14)
    assert np.all(a[2] == 0)
    for k, s in ((0, c["x0"]), (1, c["y0"])):
        v = a[k, 1:]
        sd = s / np.sqrt(2)
        assert abs(v.mean()) < 4 * sd / np.sqrt(n)
        assert abs(v.std() / sd - 1) < 4 / np.sqrt(2 * n)
        assert np.max(np.abs(v)) <= 3 * s
    ref = GOLD["grt/frame"]
    for k in (0, 1, 3, 4):
        d, p = _ks_2samp(a[k, 1:], ref[k, 1:])
        print(f"component {k}: KS D = {d:.3f}, p = {p:.3f}")
        assert p > 1e-3


def test_gaussian_ray_frame_seeding(hip):
    c = dict(C.grt_case(), nRays=5000)
    f1 = PB.makeGRTframe(c)
    f2 = PB.makeGRTframe(c)
    np.testing.assert_array_equal(f1.x, f2.x)
    np.testing.assert_array_equal(f1.dy, f2.dy)
    f3 = PB.makeGRTframe(dict(c, seed=8))
    assert np.mean(f3.x != f1.x) > 0.99
    f4 = PB.makeGRTframe(dict(c, seed=-1))
    assert np.mean(f4.x != f1.x) > 0.99


def test_beam_errors_raise(hip):
    from pypo_amd import bind as B
    c = C.beam_cases()["currents_m0"]
    fields = C.random_fields(gridsize(c["source"]), 1)
    with pytest.raises(B.PyPOHIPError):
        PB.calcCurrents(fields, c["source"], 7)
    # a pdf peak far below 1 accepts (almost) nothing: each ray stops after its
    # bounded number of trials and the call reports it
    with pytest.raises(B.PyPOHIPError):
        PB.makeGRTframe({"nRays": 64, "x0": 1000.0, "y0": 1000.0, "angx0": 60.0, "angy0": 60.0, "seed": 1})
