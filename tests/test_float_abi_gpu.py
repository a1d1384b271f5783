"""The reference GPU library's single-precision entry points (callKernelf_*,
callRTKernel; src/CUDA/InterfaceCUDA.h:29-69) through pypo_amd.bind_gpu, the
twin of the reference's BindGPU, and the CPU library's FP32 names
(propagateToGridf_*, propagateToFarFieldf; src/CPU/InterfaceCPU.h:64-101).

The inputs cross the boundary in FP32, so the checker runs the FP64 CPU
restatement on the same FP32-rounded inputs (surface parameters, currents,
frames).  Bars: every PO mode within the mixed bar (1e-5 of each field's
peak) of the checker, and hit points within FP32 output rounding."""

import numpy as np
import pytest

import cases as C
from helpers import flat, group_err
from oracle import bind as OB
from pypo_amd import bind_gpu as BG
from pypo_amd import types as T
from pypo_amd.geometry import gridsize

pytestmark = pytest.mark.gpu


def _f32_refl(refl):
    """The surface as the FP32 structs carry it: every float parameter rounded to FP32."""
    r = dict(refl)
    for key in ("coeffs", "lims_x", "lims_y", "lims_u", "lims_v", "lims_Az", "lims_El", "gcenter", "transf"):
        if key in r:
            r[key] = np.asarray(r[key], dtype=np.float32).astype(np.float64)
    for key in ("ecc_uv", "rot_uv", "rms"):
        if key in r:
            r[key] = float(np.float32(r[key]))
    return r


def _f32_complex(a):
    a = np.asarray(a)
    return np.real(a).astype(np.float32).astype(np.float64) + 1j * np.imag(a).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("name", sorted(C.po_cases()))
def test_callKernelf_matches_checker_on_fp32_inputs(hip, port, name):
    c = C.po_cases()[name]
    src, tgt = _f32_refl(c["source"]), _f32_refl(c["target"])
    shape = gridsize(c["source"])
    k = float(np.float32(2 * np.pi / c["lam"]))
    eps = float(np.float32(c.get("epsilon", 1.0)))
    if c["mode"] == "scalar":
        s = _f32_complex(C.random_scalar(shape, c["seed"]).S)
        obj = T.scalarfield(s)
        d = {"mode": "scalar", "k": k, "exp": c["exp"], "epsilon": eps, "s_scalarfield": obj}
    else:
        cur = C.random_currents(shape, c["seed"])
        obj = T.currents(*[_f32_complex(x) for x in cur.components()])
        d = {"mode": c["mode"], "k": k, "exp": c["exp"], "epsilon": eps, "s_current": obj}
    got = BG.PyPO_GPUf(src, tgt, d)
    ref, _, _ = OB.propagate(port, c["mode"], src, tgt, obj, k, epsilon=eps, exp=c["exp"])
    err = group_err(flat(got), flat(ref))
    print(f"{name} fp32 entry point: {err:.2e}")
    assert err <= 1e-5
    assert type(got) is type(ref) or (isinstance(got, list) and [type(x) for x in got] == [type(x) for x in ref])


@pytest.mark.parametrize("name", sorted(C.po_cases()))
def test_cpu_library_fp32_names_match_callKernelf(hip, name):
    """propagateToGridf_* / propagateToFarFieldf (InterfaceCPU.h:64-101) run the
    same FP32 path as callKernelf_*: bit-identical results, whatever numThreads says."""
    c = C.po_cases()[name]
    src, tgt = _f32_refl(c["source"]), _f32_refl(c["target"])
    shape = gridsize(c["source"])
    k = float(np.float32(2 * np.pi / c["lam"]))
    eps = float(np.float32(c.get("epsilon", 1.0)))
    if c["mode"] == "scalar":
        d = {"mode": "scalar", "k": k, "exp": c["exp"], "epsilon": eps,
             "s_scalarfield": T.scalarfield(_f32_complex(C.random_scalar(shape, c["seed"]).S))}
    else:
        cur = C.random_currents(shape, c["seed"])
        d = {"mode": c["mode"], "k": k, "exp": c["exp"], "epsilon": eps,
             "s_current": T.currents(*[_f32_complex(x) for x in cur.components()])}
    a = flat(BG.PyPO_GPUf(src, tgt, d, interface="gpu"))
    b = flat(BG.PyPO_GPUf(src, tgt, dict(d, nThreads=7), interface="cpu"))
    assert len(a) == len(b)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_callRTKernel_matches_checker_on_fp32_inputs(hip, port):
    rng = np.random.default_rng(2)
    n = 5000
    x, y = rng.uniform(-40, 40, n), rng.uniform(-40, 40, n)
    d = np.stack([rng.normal(0, 0.01, n), rng.normal(0, 0.01, n), np.ones(n)])
    d /= np.linalg.norm(d, axis=0)
    f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    fr = T.frame(n, f32(x), f32(y), f32(np.full(n, -30.0)), f32(d[0]), f32(d[1]), f32(d[2]))
    from pypo_amd import geometry as G
    dish = _f32_refl(G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 100.0],
                                   "gmode": "xy", "lims_x": [-50, 50], "lims_y": [-50, 50], "gridsize": [3, 3]}))
    got = BG.RT_GPUf({"fr_in": fr, "t_name": dish, "tol": 1e-7, "t0": 30.0})
    ref = OB.trace(port, dish, fr, None, float(np.float32(1e-7)), 30.0, max_iter=50)
    for a in ("x", "y", "z"):
        g, r = getattr(got, a), getattr(ref, a)
        assert np.max(np.abs(g - r)) <= 2e-6 * max(1.0, np.max(np.abs(r))), a
    for a in ("dx", "dy", "dz"):
        assert np.max(np.abs(getattr(got, a) - getattr(ref, a))) <= 1e-6, a
