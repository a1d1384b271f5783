"""Error path of the library on a live device: injected allocation failures
(pypo_hip_inject_alloc_failures) surface as PyPOHIPError with the library's
message, nothing exits or resets the device, and the next call -- which must
also wait for any work the failed call had already enqueued -- gives the
same result as a clean run."""

import numpy as np
import pytest

from helpers import flat
from pypo_amd import bind as B
from pypo_amd import geometry as G
from pypo_amd import types as T

pytestmark = pytest.mark.gpu


def _case():
    src = G.plane({"gmode": "xy", "lims_x": [-5, 5], "lims_y": [-5, 5], "gridsize": [21, 23]})
    tgt = G.plane({"gmode": "xy", "lims_x": [-6, 6], "lims_y": [-6, 6], "gridsize": [17, 19]})
    tgt["transf"] = G.mat_translate([0, 0, 40.0])
    rng = np.random.default_rng(8)
    cur = T.currents(*[rng.normal(size=(21, 23)) + 1j * rng.normal(size=(21, 23)) for _ in range(6)])
    return src, tgt, {"mode": "EH", "k": 2 * np.pi / 2.0, "s_current": cur}


@pytest.mark.parametrize("n_fail", [1, 2, 4])
def test_injected_allocation_failure_then_recovery(hip, n_fail):
    lib = hip
    src, tgt, d = _case()
    clean = flat(B.PyPO_HIPd(src, tgt, d))
    # a larger target forces fresh pool requests in every slot
    big = dict(tgt, gridsize=np.array([41, 43]))
    assert lib.pypo_hip_inject_alloc_failures(n_fail) == 0
    for _ in range(n_fail):
        with pytest.raises(B.PyPOHIPError, match="injected allocation failure"):
            B.PyPO_HIPd(src, big, d)
    # the budget is spent: calls succeed again and agree with a clean run
    out = flat(B.PyPO_HIPd(src, tgt, d))
    np.testing.assert_array_equal(out, clean)
    assert flat(B.PyPO_HIPd(src, big, d)).shape[1] == 41 * 43


def test_injected_failure_in_device_and_ray_paths(hip):
    import torch
    from pypo_amd import device as D
    lib = hip
    src, tgt, d = _case()
    gs, gt = D.generate_grid(src), D.generate_grid(tgt)
    cur = D.currents_tensor(d["s_current"])
    ref = D.propagate("EH", gs, gt, cur, src, d["k"]).clone()
    lib.pypo_hip_inject_alloc_failures(1)
    with pytest.raises(B.PyPOHIPError, match="injected"):
        D.propagate("EH", gs, gt, cur, src, d["k"])
    out = D.propagate("EH", gs, gt, cur, src, d["k"])
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    fr = T.frame(64, np.linspace(-3, 3, 64), np.zeros(64), np.full(64, -10.0), np.zeros(64), np.zeros(64),
                 np.ones(64))
    dish = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 20.0], "gmode": "xy",
                         "lims_x": [-5, 5], "lims_y": [-5, 5], "gridsize": [3, 3]})
    rt = {"fr_in": fr, "t_name": dish, "tol": 1e-12, "t0": 10.0}
    ok = B.RT_HIPd(rt)
    lib.pypo_hip_inject_alloc_failures(1)
    with pytest.raises(B.PyPOHIPError, match="injected"):
        B.RT_HIPd(dict(rt, fr_in=T.frame(100, *[np.resize(getattr(fr, a), 100) for a in
                                                  ("x", "y", "z", "dx", "dy", "dz")])))
    again = B.RT_HIPd(rt)
    np.testing.assert_array_equal(np.stack([again.x, again.y, again.z]), np.stack([ok.x, ok.y, ok.z]))
    import ctypes
    assert lib.pypo_hip_inject_alloc_failures(-1) != 0
    buf = ctypes.create_string_buffer(128)
    assert lib.pypo_hip_last_error(buf, 128) != 0 and b"inject" in buf.value   # read and cleared


def test_device_entry_points_reject_bad_arguments(hip):
    """Argument errors of the device-resident entry points come back as a
    non-zero code and a message, before anything is enqueued, and the
    library stays usable: the next good call matches a clean run."""
    import ctypes
    import torch
    from pypo_amd import device as D
    from pypo_amd import efficiencies as E
    from pypo_amd import marshal as MU
    from pypo_amd import structs as PS
    lib = hip
    src, tgt, d = _case()
    k = d["k"]
    gs, gt = D.generate_grid(src), D.generate_grid(tgt)
    cur = D.currents_tensor(d["s_current"])
    ref = D.propagate("EH", gs, gt, cur, src, k).clone()
    out = torch.zeros_like(ref)
    cs, ct, cb, rb = D.container(gs), D.container(gt), D.bundle(cur), D.bundle(out)
    st = D._stream()
    buf = ctypes.create_string_buffer(512)

    def vp(x):
        return ctypes.cast(ctypes.byref(x), ctypes.c_void_p)

    def expect(rc, text):
        assert rc != 0
        assert lib.pypo_hip_last_error(buf, 512) != 0
        assert text in buf.value.decode(), buf.value

    lib.pypo_hip_last_error(buf, 512)
    args = [vp(rb), ctypes.byref(cs), ctypes.byref(ct), vp(cb), 0, 21, 23, k, 1.0, -1.0]
    for fn in (lib.pypo_hip_propagate_dev, lib.pypo_hip_propagate_noalloc):
        expect(fn(99, *args, 0, st), "unknown PO mode")
        expect(fn(1, *args, 5, st), "unknown precision")
        expect(fn(1, *args[:5], 20, 23, *args[7:], 0, st), "ncx*ncy")
        expect(fn(1, None, *args[1:], 0, st), "NULL argument")
    # more packed sources than the no-alloc scratch holds: refused before any launch
    n_big = 2000 * 1000
    g_big = torch.zeros((7, n_big), dtype=torch.float64, device="cuda")
    c_big = torch.zeros((12, n_big), dtype=torch.float64, device="cuda")
    cs_big, cb_big = D.container(g_big), D.bundle(c_big)
    expect(lib.pypo_hip_propagate_noalloc(1, vp(rb), ctypes.byref(cs_big), ctypes.byref(ct), vp(cb_big), 0, 2000,
                                          1000, k, 1.0, -1.0, 0, st), "more than its")
    del g_big, c_big, cs_big, cb_big
    # ray tracing takes 1 to 4 surfaces
    surf = MU.fill_reflparams(src)
    surfs = (PS.reflparams * 5)(*([surf] * 5))
    fr = torch.zeros((6, 8), dtype=torch.float64, device="cuda")
    frames = (PS.cframe * 6)(*([D.frame_struct(fr)] * 6))
    for n_surf in (0, 5):
        expect(lib.pypo_hip_trace_dev(surfs, n_surf, frames, 1e-9, 1.0, st), "1 to 4 surfaces")
    # nearest interpolation of an empty point set; reductions over nothing
    D._declare_hybrid(lib)
    vals = torch.zeros((12, 8), dtype=torch.float64, device="cuda")
    res = torch.zeros((12, gt.shape[1]), dtype=torch.float64, device="cuda")
    pe, ve, re_ = D.frame_struct(fr), D.bundle(vals), D.bundle(res)
    pe.size = 0     # valid arrays, no points
    expect(lib.pypo_hip_interp_nearest_dev(ctypes.byref(pe), ctypes.byref(ve), ctypes.byref(ct), ctypes.byref(re_),
                                           st), "empty point set")
    E._declare(lib)
    sums = (ctypes.c_double * 8)()
    row = ctypes.cast(ctypes.c_void_p(ref[0].data_ptr()), ctypes.POINTER(ctypes.c_double))
    expect(lib.pypo_hip_power_sums_dev(row, row, row, row, 0, sums, st), "n must be positive")
    expect(lib.pypo_hip_frame_stats_dev(ctypes.byref(pe), sums, st), "empty frame")
    # nothing was enqueued and the library still works
    D.propagate("EH", gs, gt, cur, src, k, out=out)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    assert lib.pypo_hip_last_error(buf, 512) == 0
