"""pypo_hip_propagate_noalloc (workspace in the library's static device
scratch, no allocation) against pypo_hip_propagate_dev on the same inputs,
through the C ABI: every mode in both precisions, two calls in a row with
the second on smaller grids; the large far field (1001^2 -> 501^2), whose
split plan has to shrink to fit the scratch; and a problem too large for
the scratch, rejected before anything runs."""

import numpy as np
import pytest

from pypo_amd import bind as B
from pypo_amd import geometry as G

pytestmark = pytest.mark.gpu

MODES = ["JM", "EH", "JMEH", "EHP", "FF", "scalar"]


def _surfaces(n_src, n_tgt, mode):
    src = G.plane({"gmode": "xy", "lims_x": [-8, 8], "lims_y": [-8, 8], "gridsize": [n_src, n_src + 1],
                   "flip": True})
    src["transf"] = G.mat_translate([0, 0, 60.0])
    if mode == "FF":
        tgt = G.plane({"gmode": "AoE", "lims_Az": [-4, 4], "lims_El": [-4, 4], "gridsize": [n_tgt, n_tgt - 1]})
    else:
        tgt = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 60.0], "gmode": "uv",
                            "lims_u": [0, 40], "lims_v": [0, 360], "gridsize": [n_tgt, n_tgt + 2]})
    return src, tgt


def _inputs(mode, n_src, n_tgt, seed):
    import torch
    from pypo_amd import device as D
    src, tgt = _surfaces(n_src, n_tgt, mode)
    gs, gt = D.generate_grid(src), D.generate_grid(tgt)
    rows = 2 if mode == "scalar" else 12
    g = torch.Generator(device="cpu").manual_seed(seed)
    vals = torch.randn((rows, gs.shape[1]), generator=g, dtype=torch.float64).cuda()
    return src, gs, gt, vals


def _run(mode, prec, inp, no_alloc):
    import torch
    from pypo_amd import device as D
    src, gs, gt, vals = inp
    out = torch.full((D.N_OUT[mode], gt.shape[1]), 7.0, dtype=torch.float64, device="cuda")
    return D.propagate(mode, gs, gt, vals, src, 2 * np.pi / 1.3, precision=prec, out=out, no_alloc=no_alloc)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("mode", MODES)
def test_noalloc_matches_dev(hip, mode, prec):
    """Two no-alloc calls in a row, the second on smaller grids, against the
    pooled entry point: the split plans fit the scratch, so the results are
    bit-identical."""
    import torch
    inputs = [_inputs(mode, 41, 37, 1), _inputs(mode, 13, 11, 2)]
    refs = [_run(mode, prec, inp, False) for inp in inputs]
    outs = [_run(mode, prec, inp, True) for inp in inputs]
    torch.cuda.synchronize()
    for inp, out, ref in zip(inputs, outs, refs):
        print(f"{mode} {prec} {inp[1].shape[1]} -> {inp[2].shape[1]}: max |diff| "
              f"{(out - ref).abs().max().item():.3e}")
        assert torch.equal(out, ref)


def test_noalloc_large_far_field_and_rejection(hip):
    """The large far field: the partials of the usual split plan do not fit
    the scratch beside the 96 MB of packed records, so the call runs with
    fewer splits (the FP64 sum of the same FP32 tile sums, in another order);
    then a call too large for the scratch is rejected with an error and
    leaves its output untouched; then a smaller call still matches."""
    import torch
    import bench
    from pypo_amd import device as D
    feed, dish, ff = bench.surfaces()
    g_dish, g_ff = D.generate_grid(dish), D.generate_grid(ff)
    g = torch.Generator(device="cpu").manual_seed(5)
    cur = torch.randn((12, g_dish.shape[1]), generator=g, dtype=torch.float64).cuda()
    ref = D.propagate("FF", g_dish, g_ff, cur, dish, bench.K, precision="mixed")
    out = torch.empty_like(ref)
    D.propagate("FF", g_dish, g_ff, cur, dish, bench.K, precision="mixed", out=out, no_alloc=True)
    torch.cuda.synchronize()
    for f in (slice(0, 6), slice(6, 12)):
        pk = ref[f].abs().max().item()
        err = (out[f] - ref[f]).abs().max().item() / pk
        print(f"large FF no-alloc vs dev: {err:.2e} of peak")
        assert err <= 1e-12
    # 1201 x 1201 sources: 138 MB of packed records alone exceed the 128 MiB scratch
    big = G.plane({"gmode": "xy", "lims_x": [-8, 8], "lims_y": [-8, 8], "gridsize": [1201, 1201]})
    g_big = D.generate_grid(big)
    cur_big = torch.zeros((12, g_big.shape[1]), dtype=torch.float64, device="cuda")
    sentinel = torch.full((12, g_ff.shape[1]), 3.0, dtype=torch.float64, device="cuda")
    with pytest.raises(B.PyPOHIPError, match="scratch"):
        D.propagate("FF", g_big, g_ff, cur_big, big, bench.K, out=sentinel, no_alloc=True)
    torch.cuda.synchronize()
    assert torch.all(sentinel == 3.0)
    inp = _inputs("EH", 21, 17, 9)
    assert torch.equal(_run("EH", "mixed", inp, True), _run("EH", "mixed", inp, False))


def test_noalloc_calls_on_two_streams_share_the_scratch_in_turn(hip):
    """No-alloc calls on two streams, enqueued with no host synchronisation,
    share the device's one scratch: the library orders them, so each equals
    its pypo_hip_propagate_dev result (overlapping calls would overwrite each
    other's packed sources and partials)."""
    import torch
    from pypo_amd import device as D
    cases = [("FF", 57, 61, 11), ("EH", 43, 37, 12), ("FF", 31, 45, 13), ("JM", 49, 29, 14)]
    refs, ins = [], []
    for mode, ns, nt, seed in cases:
        src, gs, gt, vals = _inputs(mode, ns, nt, seed)
        ins.append((mode, src, gs, gt, vals))
        refs.append(D.propagate(mode, gs, gt, vals, src, 2 * np.pi / 3.0).clone())
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for rep in range(2):
        for i, (mode, src, gs, gt, vals) in enumerate(ins):
            with torch.cuda.stream(s1 if (i + rep) % 2 == 0 else s2):
                outs.append((i, D.propagate(mode, gs, gt, vals, src, 2 * np.pi / 3.0, no_alloc=True)))
    torch.cuda.synchronize()
    for i, o in outs:
        assert torch.equal(o, refs[i]), cases[i]
