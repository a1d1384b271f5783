"""Device-resident calls on several streams.  Each stream that calls the
library has its own grow-only scratch set, so calls on different streams are
not ordered against each other, and a set that must grow frees and allocates
in its stream's order (no device-wide synchronisation).  Results must be
bit-identical to the same calls made one after another on one stream."""

import ctypes

import numpy as np
import pytest

from pypo_amd import geometry as G
from pypo_amd import types as T

pytestmark = pytest.mark.gpu


def _surfaces(n_src, n_tgt):
    src = G.plane({"gmode": "xy", "lims_x": [-5, 5], "lims_y": [-5, 5], "gridsize": [n_src, n_src + 2]})
    tgt = G.plane({"gmode": "xy", "lims_x": [-6, 6], "lims_y": [-6, 6], "gridsize": [n_tgt, n_tgt + 1]})
    tgt["transf"] = G.mat_translate([0, 0, 40.0])
    rng = np.random.default_rng(n_src)
    shape = (n_src, n_src + 2)
    cur = T.currents(*[rng.normal(size=shape) + 1j * rng.normal(size=shape) for _ in range(6)])
    return src, tgt, cur


def _run(D, case, mode, prec):
    src, tgt, cur = case
    gs, gt = D.generate_grid(src), D.generate_grid(tgt)
    return D.propagate(mode, gs, gt, D.currents_tensor(cur), src, 2 * np.pi / 2.0, precision=prec)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_calls_on_two_streams_match_one_stream(hip, prec):
    import torch
    from pypo_amd import device as D
    cases = [_surfaces(31, 41), _surfaces(57, 23), _surfaces(19, 77)]
    modes = ["EH", "JM", "FF"]
    ref = [_run(D, c, m, prec).clone() for c, m in zip(cases, modes)]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for rep in range(3):             # interleaved, no host synchronisation between calls
        for i, (c, m) in enumerate(zip(cases, modes)):
            with torch.cuda.stream(s1 if (i + rep) % 2 == 0 else s2):
                outs.append((i, _run(D, c, m, prec)))
    torch.cuda.synchronize()
    for i, o in outs:
        assert torch.equal(o, ref[i]), (i, modes[i])


def test_calls_from_several_host_threads(hip):
    """Four host threads, each on its own stream, mix device-resident calls,
    synchronous drop-in calls and a call that fails on a bad argument.  The
    library serialises the host side and keeps scratch per stream and the
    error state per thread: every result is bit-identical to the same call
    made alone, and each thread sees only its own error."""
    import threading
    import torch
    from helpers import flat
    from pypo_amd import bind as B
    from pypo_amd import device as D
    cases = [_surfaces(31, 41), _surfaces(57, 23), _surfaces(19, 77), _surfaces(45, 45)]
    modes = ["EH", "JM", "FF", "JMEH"]
    k = 2 * np.pi / 2.0
    ref_dev = [_run(D, c, m, "mixed").clone() for c, m in zip(cases, modes)]
    ref_drop = [flat(B.PyPO_HIPd(c[0], c[1], {"mode": "EH", "k": k, "s_current": c[2]})) for c in cases]
    torch.cuda.synchronize()
    errors, outs = [], {}

    def worker(i):
        try:
            st = torch.cuda.Stream()
            with torch.cuda.stream(st):
                for rep in range(3):
                    j = (i + rep) % len(cases)
                    dev = _run(D, cases[j], modes[j], "mixed")
                    drop = flat(B.PyPO_HIPd(cases[j][0], cases[j][1], {"mode": "EH", "k": k,
                                                                       "s_current": cases[j][2]}))
                    if rep == 1:
                        with pytest.raises(B.PyPOHIPError, match="unknown PO mode"):
                            _bad_mode(D, cases[j], k)
                    st.synchronize()
                    outs[(i, rep)] = (j, dev.clone(), drop)
        except BaseException as e:   # reported on the main thread
            errors.append((i, repr(e)))

    threads = [threading.Thread(target=worker, args=(i,), daemon=True) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not errors, errors
    assert len(outs) == 12
    for (i, rep), (j, dev, drop) in outs.items():
        assert torch.equal(dev, ref_dev[j]), (i, rep, modes[j])
        np.testing.assert_array_equal(drop, ref_drop[j])


def _bad_mode(D, case, k):
    """A device call with an unknown mode number, through the C ABI."""
    import ctypes
    from pypo_amd import bind as B
    lib = B.loadHIPlib()
    src, tgt, cur = case
    gs, gt = D.generate_grid(src), D.generate_grid(tgt)
    c = D.currents_tensor(cur)
    out = torch_empty_like_rows(gt, 12)
    cs, ct, cb, rb = D.container(gs), D.container(gt), D.bundle(c), D.bundle(out)
    ncx, ncy = (int(v) for v in src["gridsize"])
    rc = lib.pypo_hip_propagate_dev(77, ctypes.cast(ctypes.byref(rb), ctypes.c_void_p), ctypes.byref(cs),
                                    ctypes.byref(ct), ctypes.cast(ctypes.byref(cb), ctypes.c_void_p), 0, ncx, ncy,
                                    k, 1.0, -1.0, 0, D._stream())
    assert rc != 0
    B.check(lib, "pypo_hip_propagate_dev")


def torch_empty_like_rows(t, rows):
    import torch
    return torch.empty((rows, t.shape[1]), dtype=torch.float64, device=t.device)


def test_scratch_grows_with_the_stream_ordered_allocator(hip):
    assert hip.pypo_hip_stream_ordered_scratch() == 1


def test_scratch_growth_in_stream_order(hip):
    """small -> large -> small on a fresh stream without host synchronisation:
    the large call grows the stream's buffers while the small call's work may
    still be queued, and the last call reuses the grown buffers."""
    import torch
    from pypo_amd import device as D
    small, large = _surfaces(23, 29), _surfaces(151, 97)
    ref_s = _run(D, small, "EH", "mixed").clone()
    ref_l = _run(D, large, "EH", "mixed").clone()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        a = _run(D, small, "EH", "mixed")
        b = _run(D, large, "EH", "mixed")
        c = _run(D, small, "EH", "mixed")
    s.synchronize()
    assert torch.equal(a, ref_s)
    assert torch.equal(b, ref_l)
    assert torch.equal(c, ref_s)


def test_release_scratch_of_a_stream(hip):
    """A stream's scratch is freed on request, also after the stream itself
    was dropped, and the next call on a stream allocates a new set with the
    same results; other streams keep theirs."""
    import torch
    from pypo_amd import device as D
    case = _surfaces(41, 37)
    ref = _run(D, case, "EH", "mixed").clone()
    torch.cuda.synchronize()
    keep = D.scratch_bytes()
    assert keep > 0
    s = torch.cuda.Stream()
    assert D.scratch_bytes(s) == 0
    with torch.cuda.stream(s):
        a = _run(D, case, "EH", "mixed")
    held = D.scratch_bytes(s)
    assert held > 0
    D.release_scratch(s)          # waits for the call queued on s
    assert D.scratch_bytes(s) == 0
    assert torch.equal(a, ref)
    with torch.cuda.stream(s):
        b = _run(D, case, "EH", "mixed")
    s.synchronize()
    assert torch.equal(b, ref) and D.scratch_bytes(s) == held
    handle = s.cuda_stream
    del s
    assert hip.pypo_hip_release_scratch(ctypes.c_void_p(handle)) == 0   # its stream may be gone
    assert hip.pypo_hip_scratch_bytes(ctypes.c_void_p(handle)) == 0
    assert hip.pypo_hip_release_scratch(ctypes.c_void_p(handle)) == 0   # nothing held: no-op
    assert D.scratch_bytes() == keep


@pytest.mark.parametrize("no_alloc", [False, True])
@pytest.mark.parametrize("mode", ["JM", "FF"])
def test_propagation_captured_in_a_graph(hip, mode, no_alloc):
    """A device-resident propagation captured into a graph after one eager
    call on the capturing stream: every replay, with unrelated eager work and
    synchronisation between replays, equals the eager result."""
    import torch
    from pypo_amd import device as D
    src, tgt, cur = _surfaces(37, 41)
    if mode == "FF":
        tgt = G.plane({"gmode": "AoE", "lims_Az": [-3, 3], "lims_El": [-3, 3], "gridsize": [43, 47]})
    k = 2 * np.pi / 2.0
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        gs, gt = D.generate_grid(src), D.generate_grid(tgt)
        c = D.currents_tensor(cur)
        ref = D.propagate(mode, gs, gt, c, src, k, no_alloc=no_alloc).clone()   # warms the stream's buffers
        out = torch.empty_like(ref)
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        D.propagate(mode, gs, gt, c, src, k, out=out, no_alloc=no_alloc)
    for rep in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
        junk = torch.randn(4096, device="cuda").cumsum(0)     # unrelated eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk).all()


def test_capture_that_would_grow_scratch_is_refused(hip):
    """Growing a buffer set inside a capture is reported, not attempted."""
    import torch
    from pypo_amd import bind as B
    from pypo_amd import device as D
    src, tgt, cur = _surfaces(61, 67)
    k = 2 * np.pi / 2.0
    gs, gt = D.generate_grid(src), D.generate_grid(tgt)
    c = D.currents_tensor(cur)
    out = torch.empty((12, gt.shape[1]), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    s = torch.cuda.Stream()          # fresh stream: its buffer set is empty
    g = torch.cuda.CUDAGraph()
    with pytest.raises(B.PyPOHIPError, match="inside a stream capture"):
        with torch.cuda.graph(g, stream=s):
            D.propagate("EH", gs, gt, c, src, k, out=out)
    torch.cuda.synchronize()
    again = D.propagate("EH", gs, gt, c, src, k)
    torch.cuda.synchronize()
    assert torch.isfinite(again).all()
