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
