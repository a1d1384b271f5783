"""Parity of the HIP PO kernels (through the C ABI) with the CPU checker.

Bars: FP64 path <= 1e-10 of each field's peak; mixed path (FP32 integrand,
FP64 distance/phase) <= 1e-5 of each field's peak -- the tolerance the
headline far-field benchmark is quoted at.  Peaks are taken per field (E, H, J, M,
S_r separately).
"""

import numpy as np
import pytest

import cases as C
from helpers import flat, group_err
from oracle import bind as OB
from pypo_amd import bind as B
from pypo_amd import geometry as G
from pypo_amd import types as T
from pypo_amd.geometry import gridsize

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "mixed": 1e-5}


def _src_for(c):
    shape = gridsize(c["source"])
    return (C.random_scalar(shape, c["seed"]) if c["mode"] == "scalar"
            else C.random_currents(shape, c["seed"]))


def _run(c, prec):
    d = {"mode": c["mode"], "k": 2 * np.pi / c["lam"], "exp": c["exp"], "precision": prec,
         "epsilon": c.get("epsilon", 1.0)}
    d["s_scalarfield" if c["mode"] == "scalar" else "s_current"] = _src_for(c)
    return B.PyPO_HIPd(c["source"], c["target"], d, return_grids=True)


@pytest.fixture(scope="module")
def gold_po():
    return np.load("tests/golden/po.npz")


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("name", sorted(C.po_cases()))
def test_po_matches_golden(hip, gold_po, name, prec):
    c = C.po_cases()[name]
    out, gs, gt = _run(c, prec)
    err = group_err(flat(out), gold_po[f"{name}/out"])
    print(f"{name} {prec}: {err:.2e}")
    assert err <= TOL[prec]
    # the drop-in writes the generated target grid back like the reference
    tg = np.stack([np.ravel(getattr(gt, a)) for a in ("x", "y", "z", "nx", "ny", "nz", "area")])
    np.testing.assert_allclose(tg, gold_po[f"{name}/tgt_grid"], rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("name", [n for n in sorted(C.po_cases()) if n.startswith("FF")])
def test_ff_matrix_and_vector_kernels(hip, gold_po, name):
    """The mixed far field on the matrix cores (FP16 hi/lo operands, FP32
    accumulation) and on the vector ALUs (FP32 FMAs): both within the mixed
    bar of the reference's output, and within 2e-6 of each other.  The
    matrix kernel scales each tile's operands by a power of two, so scaling
    the currents by 2^40 or 2^-40 scales its output exactly."""
    c = C.po_cases()[name]
    outs = {}
    try:
        for kind in ("vector", "matrix"):
            B.set_ff_kernel(kind)
            outs[kind] = flat(_run(c, "mixed")[0])
            err = group_err(outs[kind], gold_po[f"{name}/out"])
            print(f"{name} {kind}: {err:.2e}")
            assert err <= TOL["mixed"]
        assert group_err(outs["matrix"], outs["vector"]) <= 2e-6
        cur = _src_for(c)
        d = {"mode": "FF", "k": 2 * np.pi / c["lam"], "exp": c["exp"], "precision": "mixed"}
        for f in (2.0 ** 40, 2.0 ** -40):
            scaled = T.currents(*[f * comp for comp in cur.components()])
            out = flat(B.PyPO_HIPd(c["source"], c["target"], dict(d, s_current=scaled)))
            np.testing.assert_array_equal(out, f * outs["matrix"])
    finally:
        B.set_ff_kernel("matrix")


def test_ff_matrix_kernel_edge_tiles(hip, port):
    """Tiles that the matrix kernel treats specially: an all-zero tile (no
    scale), a tile mixing magnitudes 1e-6..1 (small values below FP16's
    normal range after scaling), magnetic currents, and a partial last tile
    (775 = 3 * 256 + 7 sources), against the CPU checker."""
    rng = np.random.default_rng(4)
    src = G.plane({"gmode": "xy", "lims_x": [-20, 20], "lims_y": [-20, 20], "gridsize": [25, 31]})
    n = 25 * 31                             # 3 full tiles and a partial one of 7
    comps = [rng.normal(size=n) + 1j * rng.normal(size=n) for _ in range(6)]
    mags = 10.0 ** rng.uniform(-6, 0, n)
    comps = [cc * mags for cc in comps]
    for cc in comps:
        cc[:256] = 0.0                      # the first tile carries nothing
    cur = T.currents(*[cc.reshape(25, 31) for cc in comps])
    ff = G.plane({"gmode": "AoE", "lims_Az": [-30, 30], "lims_El": [-30, 30], "gridsize": [23, 19]})
    k = 2 * np.pi / 3.0
    B.set_ff_kernel("matrix")
    out = B.PyPO_HIPd(src, ff, {"mode": "FF", "k": k, "s_current": cur, "precision": "mixed"})
    ref, _, _ = OB.propagate(port, "FF", src, ff, cur, k)
    err = group_err(flat(out), flat(ref))
    print(f"edge tiles: {err:.2e}")
    assert err <= 1e-5


def _follower_tiles(src_grid, th, ph, k):
    """Which 256-source tiles the matrix kernel runs in its follower-phase form
    (one FP64 phase per group of four consecutive sources, FP32 offsets for
    the other three): all follower sums bounded by 8 revolutions, the bound
    taken over the cone of the target directions (one workgroup here)."""
    u = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)])
    a = u.sum(axis=1)
    a /= np.linalg.norm(a)
    cos_half = np.min(a @ u)
    ub = np.minimum(1.0, np.abs(a) + np.sqrt(max(0.0, 1 - cos_half ** 2)) + 1e-3) if cos_half > 0 else np.ones(3)
    # the packed records are in serpentine order: odd grid rows reversed
    r = np.stack([src_grid.x, src_grid.y, src_grid.z]).reshape(3, *np.shape(src_grid.x))
    if r.ndim == 3:
        r[:, 1::2] = r[:, 1::2, ::-1].copy()
    r = r.reshape(3, -1)
    n = r.shape[1]
    flags = []
    for t0 in range(0, n, 256):
        ok = True
        for g0 in range(t0, min(n, t0 + 256), 4):
            lead = g0 + 1
            for s in range(g0, g0 + 4):
                if s == lead:
                    continue
                if s >= n or lead >= n:
                    ok = False            # padding record far away
                    continue
                d = np.abs((r[:, s] - r[:, lead]) * k / (2 * np.pi))
                ok &= 1 + d @ ub <= 8.0
        flags.append(ok)
    return np.array(flags)


@pytest.mark.parametrize("case", ["follower", "follower_wide", "fp64", "both"])
def test_ff_matrix_phase_forms(hip, port, case):
    """The matrix kernel's two phase forms against the CPU checker and the
    vector kernel: a 64 x 64 plane at 0.5 mm spacing, lambda = 2 mm, where
    every tile takes the follower form; at lambda = 0.1 mm, still all follower
    tiles but with offsets d.u up to 5 revolutions (phases up to ~6 handed to
    v_sin/v_cos unreduced); the same plane at lambda = 0.05 mm
    (neighbours 10 revolutions apart: every tile takes the full FP64 phase);
    and 16 x 301 points at lambda = 0.2 mm, where neighbours along a row
    are 0.53 revolutions apart but the rows 10.7: the groups of four that
    span a row end (records run in serpentine order, so the step there is
    one row) put the tiles holding them on the FP64 form (12 of 19) and the
    others on the follower form."""
    lam, nx, ny = {"follower": (2.0, 64, 64), "follower_wide": (0.1, 64, 64), "fp64": (0.05, 64, 64),
                   "both": (0.2, 16, 301)}[case]
    src = G.plane({"gmode": "xy", "lims_x": [-16, 16], "lims_y": [-16, 16], "gridsize": [nx, ny]})
    ff = G.plane({"gmode": "AoE", "lims_Az": [-30, 30], "lims_El": [-30, 30], "gridsize": [23, 19]})
    k = 2 * np.pi / lam
    rng = np.random.default_rng(11)
    cur = T.currents(*[rng.normal(size=(nx, ny)) + 1j * rng.normal(size=(nx, ny)) for _ in range(6)])
    gs = B.generate_grid(src)
    gt = B.generate_grid(ff, transform=False)
    flags = _follower_tiles(gs, np.ravel(gt.x), np.ravel(gt.y), k)
    expect = {"follower": flags.all(), "follower_wide": flags.all(), "fp64": not flags.any(),
              "both": flags.any() and not flags.all()}
    assert expect[case], flags
    d = {"mode": "FF", "k": k, "s_current": cur, "precision": "mixed"}
    ref, _, _ = OB.propagate(port, "FF", src, ff, cur, k)
    outs = {}
    try:
        for kind in ("vector", "matrix"):
            B.set_ff_kernel(kind)
            outs[kind] = flat(B.PyPO_HIPd(src, ff, d))
            err = group_err(outs[kind], flat(ref))
            print(f"{case} ({flags.sum()}/{flags.size} follower tiles) {kind}: {err:.2e}")
            assert err <= TOL["mixed"]
    finally:
        B.set_ff_kernel("matrix")
    assert group_err(outs["matrix"], outs["vector"]) <= 2e-6


def _feed_dish_farfield():
    """Plain PO configuration: vector Gaussian-like feed, lambda = 3 mm, on a
    101 x 101 plane over +-20 mm at the focus of a paraboloid (f = 0.5 m,
    D = 0.3 m, uv 151 x 151); JM on the dish, then FF onto 101 x 101 AoE +-1 deg."""
    lam = 3.0
    feed = G.plane({"gmode": "xy", "lims_x": [-20, 20], "lims_y": [-20, 20], "gridsize": [101, 101],
                    "flip": True})
    feed["transf"] = G.mat_translate([0, 0, 500.0])
    dish = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 500.0], "gmode": "uv",
                         "lims_u": [0, 150], "lims_v": [0, 360], "gridsize": [151, 151]})
    ff = G.plane({"gmode": "AoE", "lims_Az": [-1, 1], "lims_El": [-1, 1], "gridsize": [101, 101]})
    x = np.linspace(-20, 20, 101)
    X, Y = np.meshgrid(x, x, indexing="ij")
    amp = np.exp(-(X ** 2 + Y ** 2) / 5.0 ** 2)
    z = np.zeros_like(amp, dtype=complex)
    cur = T.currents(-amp + 0j, z, z, z, amp + 0j, z)
    return lam, feed, dish, ff, cur


@pytest.fixture(scope="module")
def feed_dish_farfield_ref(port):
    lam, feed, dish, ff, cur = _feed_dish_farfield()
    k = 2 * np.pi / lam
    jm, _, _ = OB.propagate(port, "JM", feed, dish, cur, k)
    eh, _, _ = OB.propagate(port, "FF", dish, ff, jm, k)
    return jm, eh


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_feed_101_dish_151_farfield_101_chain(hip, feed_dish_farfield_ref, prec):
    lam, feed, dish, ff, cur = _feed_dish_farfield()
    k = 2 * np.pi / lam
    jm_ref, eh_ref = feed_dish_farfield_ref
    jm = B.PyPO_HIPd(feed, dish, {"mode": "JM", "k": k, "s_current": cur, "precision": prec})
    e1 = group_err(flat(jm), flat(jm_ref))
    eh = B.PyPO_HIPd(dish, ff, {"mode": "FF", "k": k, "s_current": jm_ref, "precision": prec})
    e2 = group_err(flat(eh), flat(eh_ref))
    print(f"feed 101^2 -> dish 151^2 -> FF 101^2, {prec}: JM {e1:.2e} FF {e2:.2e}")
    assert e1 <= TOL[prec] and e2 <= TOL[prec]


def test_linearity_and_device_chaining(hip):
    import torch
    from pypo_amd import device as D
    lam, feed, dish, ff, cur = _feed_dish_farfield()
    k = 2 * np.pi / lam
    g_feed, g_dish, g_ff = D.generate_grid(feed), D.generate_grid(dish), D.generate_grid(ff)
    c1 = D.currents_tensor(cur)
    rng = np.random.default_rng(3)
    c2 = torch.as_tensor(rng.standard_normal(c1.shape) * 0.1, device="cuda")
    jm = D.propagate("JM", g_feed, g_dish, c1, feed, k, precision="fp64")
    ffa = D.propagate("FF", g_dish, g_ff, jm, dish, k, precision="mixed")
    ffb = D.propagate("FF", g_dish, g_ff, jm * 2.5, dish, k, precision="mixed")
    ffc = D.propagate("FF", g_dish, g_ff, jm + 0, dish, k, precision="mixed")
    torch.cuda.synchronize()
    # a zero-copy chain gives the same field as the host drop-in path
    jm_host = B.PyPO_HIPd(feed, dish, {"mode": "JM", "k": k, "s_current": cur, "precision": "fp64"})
    eh_host = B.PyPO_HIPd(dish, ff, {"mode": "FF", "k": k, "s_current": jm_host, "precision": "mixed"})
    assert group_err(D.tensor_to_complex6(ffa), flat(eh_host)) < 1e-12
    # deterministic: the same inputs give bit-identical results
    assert torch.equal(ffa, ffc)
    # linear in the currents (scaling by 2.5 is exact in FP32 up to rounding)
    assert group_err(D.tensor_to_complex6(ffb), 2.5 * D.tensor_to_complex6(ffa)) < 1e-6
    # linear in superposition, mixed precision
    jm2 = D.propagate("JM", g_feed, g_dish, c2, feed, k, precision="fp64")
    ffs = D.propagate("FF", g_dish, g_ff, jm + jm2, dish, k, precision="fp64")
    ff2 = D.propagate("FF", g_dish, g_ff, jm2, dish, k, precision="fp64")
    ff1 = D.propagate("FF", g_dish, g_ff, jm, dish, k, precision="fp64")
    torch.cuda.synchronize()
    assert group_err(D.tensor_to_complex6(ffs), D.tensor_to_complex6(ff1 + ff2)) < 1e-12


def test_large_far_field_mixed_vs_fp64_and_transverse(hip, port):
    """The large configuration at full size (1001 x 1001 dish -> 501 x 501):
    the mixed kernel against the FP64 kernel on every target, E and H
    transverse to r_hat, and the FP64 kernel against the CPU checker on one
    far-field row (501 targets x 1e6 sources)."""
    import torch
    import bench
    from pypo_amd import device as D
    w0 = bench.feed_waist()
    feed, dish, ff = bench.surfaces()
    g_feed, g_dish, g_ff = D.generate_grid(feed), D.generate_grid(dish), D.generate_grid(ff)
    c_feed = D.currents_tensor(bench.feed_currents(w0))
    jm = D.propagate("JM", g_feed, g_dish, c_feed, feed, bench.K, precision="fp64")
    m = D.propagate("FF", g_dish, g_ff, jm, dish, bench.K, precision="mixed")
    f = D.propagate("FF", g_dish, g_ff, jm, dish, bench.K, precision="fp64")
    torch.cuda.synchronize()
    M, F = D.tensor_to_complex6(m), D.tensor_to_complex6(f)
    err = group_err(M, F)
    print(f"large FF mixed vs fp64: {err:.2e}")
    assert err <= 1e-5
    gf = g_ff.cpu().numpy()
    th, ph = gf[0], gf[1]
    r = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)])
    for g in (0, 3):
        assert np.max(np.abs(np.sum(M[g:g + 3] * r, axis=0))) < 1e-6 * np.max(np.abs(M[g:g + 3]))
    # one far-field row against the FP64 CPU checker
    row = 250
    idx = row * bench.N_FF + np.arange(bench.N_FF)
    cur = T.currents(*[c.reshape(bench.N_DISH, bench.N_DISH) for c in D.tensor_to_complex6(jm)])
    vals, _ = OB.farfield_subset(port, dish, ff, cur, bench.K, idx)
    e64 = group_err(F[:, idx], vals)
    em = group_err(M[:, idx], vals)
    print(f"large FF row {row}: fp64 {e64:.2e} mixed {em:.2e} vs CPU FP64")
    assert e64 <= 1e-10 and em <= 1e-5


def test_large_jm_mixed_and_chain_against_cpu(hip, port):
    """The large configuration's mixed-precision step end to end against an
    all-FP64 CPU chain: the JM stage (33 x 33 feed -> 1001 x 1001 dish,
    1.09e9 pairs) on every dish point, rim included, and the mixed JM -> mixed
    FF chain on far-field rows 0, 100, ..., 500 (3006 targets x 1e6 sources)
    against the CPU's FP64 JM -> FP64 FF.  Bar: 1e-5 of each field's peak."""
    import torch
    import bench
    from helpers import grid_rows, current_rows
    from pypo_amd import device as D
    w0 = bench.feed_waist()
    feed, dish, ff = bench.surfaces()
    cur_feed = bench.feed_currents(w0)
    g_feed, g_dish, g_ff = D.generate_grid(feed), D.generate_grid(dish), D.generate_grid(ff)
    jm = D.propagate("JM", g_feed, g_dish, D.currents_tensor(cur_feed), feed, bench.K, precision="mixed")
    out = D.propagate("FF", g_dish, g_ff, jm, dish, bench.K, precision="mixed")
    torch.cuda.synchronize()
    J = D.tensor_to_complex6(jm)
    # FP64 CPU chain: JM over the whole dish, then FF on the strided rows
    jm_cpu = OB.propagate_on_grids(port, "JM", grid_rows(port, feed), (feed["gmode"], bench.N_FEED, bench.N_FEED),
                                   grid_rows(port, dish), current_rows(cur_feed), bench.K)
    Jc = np.stack([jm_cpu[0] + 1j * jm_cpu[3], jm_cpu[1] + 1j * jm_cpu[4], jm_cpu[2] + 1j * jm_cpu[5],
                   jm_cpu[6] + 1j * jm_cpu[9], jm_cpu[7] + 1j * jm_cpu[10], jm_cpu[8] + 1j * jm_cpu[11]])
    assert np.all(Jc[3:] == 0) and np.all(J[3:] == 0)   # PEC currents: no M
    e_jm = group_err(J[:3], Jc[:3])
    rim = np.arange((bench.N_DISH - 1) * bench.N_DISH, bench.N_DISH ** 2)
    e_rim = float(np.max(np.abs(J[:3, rim] - Jc[:3, rim])) / np.max(np.abs(Jc[:3])))
    e_str = group_err(J[:3, ::97], Jc[:3, ::97])
    print(f"large JM mixed vs CPU FP64: all {e_jm:.2e}, every 97th {e_str:.2e}, rim {e_rim:.2e}")
    assert e_jm <= 1e-5
    rows = np.arange(0, bench.N_FF, bench.CPU_ROW_STRIDE)
    idx = (rows[:, None] * bench.N_FF + np.arange(bench.N_FF)[None, :]).ravel()
    cur_cpu = T.currents(*[c.reshape(bench.N_DISH, bench.N_DISH) for c in Jc])
    vals, _ = OB.farfield_subset(port, dish, ff, cur_cpu, bench.K, idx)
    e_chain = group_err(D.tensor_to_complex6(out)[:, idx], vals)
    print(f"large JM -> FF mixed chain vs CPU FP64 chain on {idx.size} targets: {e_chain:.2e}")
    assert e_chain <= 1e-5


@pytest.mark.parametrize("mode", ["EH", "JMEH", "EHP"])
def test_large_near_modes_against_cpu(hip, port, mode):
    """The other near-field modes at the large configuration's full size: the
    33 x 33 Huygens feed (J and M both non-zero) onto the whole 1001 x 1001
    dish on the GPU (1.09e9 pairs), both precisions, checked against the FP64
    CPU checker on every 97th dish point and the rim.  Bars: 1e-5 (mixed) and
    1e-10 (FP64) of each field's peak; EHP's reflected Poynting directions
    (unit vectors) to 1e-5 and 1e-10 absolute."""
    import torch
    import bench
    from helpers import grid_rows, current_rows
    from pypo_amd import device as D
    w0 = bench.feed_waist()
    feed, dish, _ = bench.surfaces()
    cur_feed = bench.feed_currents(w0)
    g_feed, g_dish = D.generate_grid(feed), D.generate_grid(dish)
    n = bench.N_DISH ** 2
    idx = np.unique(np.concatenate([np.arange(0, n, 97), np.arange(n - bench.N_DISH, n)]))
    ref = OB.propagate_on_grids(port, mode, grid_rows(port, feed), (feed["gmode"], bench.N_FEED, bench.N_FEED),
                                grid_rows(port, dish)[:, idx], current_rows(cur_feed), bench.K)
    for prec, tol in (("mixed", 1e-5), ("fp64", 1e-10)):
        out = D.propagate(mode, g_feed, g_dish, D.currents_tensor(cur_feed), feed, bench.K, precision=prec)
        torch.cuda.synchronize()
        got = out[:, idx].cpu().numpy()
        n_field = 24 if mode == "JMEH" else 12
        for g0 in range(0, n_field, 6):          # each complex 3-vector: rows re x y z, im x y z
            a = got[g0:g0 + 3] + 1j * got[g0 + 3:g0 + 6]
            r = ref[g0:g0 + 3] + 1j * ref[g0 + 3:g0 + 6]
            peak = np.max(np.abs(r))
            err = float(np.max(np.abs(a - r)) / peak) if peak > 0 else float(np.max(np.abs(a)))
            print(f"large {mode} {prec} rows {g0}..{g0 + 5}: {err:.2e} of peak {peak:.3e}")
            assert err <= tol
        if mode == "EHP":
            err = float(np.max(np.abs(got[12:15] - ref[12:15])))
            print(f"large EHP {prec} Poynting directions: {err:.2e}")
            assert err <= tol


def test_large_scalar_against_cpu(hip, port):
    """The scalar mode at the large configuration's size: a scalar Gaussian on
    the 33 x 33 feed onto the whole 1001 x 1001 dish (1.09e9 pairs) in both
    precisions against the FP64 CPU checker on every dish point."""
    import torch
    import bench
    from pypo_amd import device as D
    w0 = bench.feed_waist()
    feed, dish, _ = bench.surfaces()
    x = np.linspace(-bench.FEED_HALF, bench.FEED_HALF, bench.N_FEED)
    X, Y = np.meshgrid(x, x, indexing="ij")
    field = T.scalarfield(np.exp(-(X ** 2 + Y ** 2) / w0 ** 2) * np.exp(0.3j * X))
    ref, _, _ = OB.propagate(port, "scalar", feed, dish, field, bench.K)
    r = np.ravel(ref.S)
    g_feed, g_dish = D.generate_grid(feed), D.generate_grid(dish)
    src = D.scalar_tensor(field)
    for prec, tol in (("mixed", 1e-5), ("fp64", 1e-10)):
        out = D.propagate("scalar", g_feed, g_dish, src, feed, bench.K, precision=prec)
        torch.cuda.synchronize()
        a = out[0].cpu().numpy() + 1j * out[1].cpu().numpy()
        err = float(np.max(np.abs(a - r)) / np.max(np.abs(r)))
        print(f"large scalar {prec}: {err:.2e} of peak")
        assert err <= tol


def test_bad_arguments_raise(hip):
    src = G.plane({"gmode": "xy", "lims_x": [-1, 1], "lims_y": [-1, 1], "gridsize": [3, 3]})
    z = np.zeros((3, 3), complex)
    cur = T.currents(z, z, z, z, z, z)
    with pytest.raises(B.PyPOHIPError):
        B.PyPO_HIPd(src, dict(src, gridsize=np.array([0, 3])), {"mode": "EH", "k": 1.0, "s_current": cur})
    with pytest.raises(B.PyPOHIPError):
        B.PyPO_HIPd(src, src, {"mode": "XX", "k": 1.0, "s_current": cur})
    # the library stays usable after an error
    out = B.PyPO_HIPd(src, dict(src, transf=G.mat_translate([0, 0, 5.0])),
                      {"mode": "EH", "k": 1.0, "s_current": cur})
    assert np.all(flat(out) == 0)


def test_kernel_timing_api(hip):
    lib = hip
    c = C.po_cases()["FF_uv_fwd"]
    lib.pypo_hip_set_timing(1)
    try:
        _run(c, "mixed")
        assert lib.pypo_hip_last_kernel_ms() > 0
    finally:
        lib.pypo_hip_set_timing(0)


@pytest.mark.parametrize("seed", list(range(24)) + [44])   # 44: normals of length 1.025 (shadow cones)
def test_ff_matrix_random_cases(hip, port, seed):
    """Seeded random far-field cases for the matrix and the vector kernel
    against the CPU checker: xy or uv source grids of random shape (partial last tiles, row
    lengths that do and do not divide by four), random wavelengths (follower
    and FP64 phase tiles), random far-field cones from +-0.1 to +-80 deg,
    random J and M currents, sometimes a tilted surface."""
    rng = np.random.default_rng(1000 + seed)
    nx, ny = (int(v) for v in rng.integers(3, 70, size=2))
    if rng.random() < 0.5:
        src = G.plane({"gmode": "xy", "lims_x": [-20, 20], "lims_y": [-15, 15], "gridsize": [nx, ny]})
    else:
        src = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 60.0], "gmode": "uv",
                            "lims_u": [0, 25], "lims_v": [0, 360], "gridsize": [nx, ny]})
    if rng.random() < 0.3:
        src["transf"] = G.mat_translate([3.0, -2.0, 5.0], G.find_rotation([0, 0, 1], [0.2, 0.1, 1.0]))
    deg = float(10 ** rng.uniform(-1, np.log10(80)))
    ff = G.plane({"gmode": "AoE", "lims_Az": [-deg, deg], "lims_El": [-deg, deg],
                  "gridsize": [int(rng.integers(3, 30)), int(rng.integers(3, 30))]})
    lam = float(10 ** rng.uniform(np.log10(0.05), np.log10(5.0)))
    k = 2 * np.pi / lam
    comps = [rng.normal(size=(nx, ny)) + 1j * rng.normal(size=(nx, ny)) for _ in range(6)]
    if rng.random() < 0.5:
        comps[3:] = [np.zeros((nx, ny), complex)] * 3
    cur = T.currents(*comps)
    ref, _, _ = OB.propagate(port, "FF", src, ff, cur, k)
    try:
        for kind in ("matrix", "vector"):
            B.set_ff_kernel(kind)
            out = B.PyPO_HIPd(src, ff, {"mode": "FF", "k": k, "s_current": cur, "precision": "mixed"})
            err = group_err(flat(out), flat(ref))
            print(f"seed {seed}: {nx}x{ny} {src['gmode']} lam {lam:.3g} cone {deg:.3g} deg, {kind}: {err:.2e}")
            assert err <= TOL["mixed"]
    finally:
        B.set_ff_kernel("matrix")


@pytest.mark.parametrize("seed", range(15))
def test_near_modes_random_cases(hip, port, seed):
    """Seeded random near-field cases (JM, EH, JMEH, EHP, scalar by seed) in
    mixed and FP64 precision against the CPU checker: random source and
    target grid shapes (partial last tiles of every length, whose padding
    records the sums skip), random wavelengths and distances."""
    rng = np.random.default_rng(2000 + seed)
    mode = ("JM", "EH", "JMEH", "EHP", "scalar")[seed % 5]
    nx, ny = (int(v) for v in rng.integers(2, 40, size=2))
    src = G.plane({"gmode": "xy", "lims_x": [-10, 10], "lims_y": [-8, 8], "gridsize": [nx, ny]})
    tgt = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 80.0], "gmode": "uv",
                        "lims_u": [0, 30], "lims_v": [0, 360],
                        "gridsize": [int(rng.integers(3, 40)), int(rng.integers(3, 40))]})
    tgt["transf"] = G.mat_translate([0, 0, float(rng.uniform(20, 200))])   # in front of the source plane
    lam = float(10 ** rng.uniform(np.log10(0.2), np.log10(5.0)))
    k = 2 * np.pi / lam
    if mode == "scalar":
        obj = T.scalarfield(rng.normal(size=(nx, ny)) + 1j * rng.normal(size=(nx, ny)))
        key = "s_scalarfield"
    else:
        obj = T.currents(*[rng.normal(size=(nx, ny)) + 1j * rng.normal(size=(nx, ny)) for _ in range(6)])
        key = "s_current"
    ref, _, _ = OB.propagate(port, mode, src, tgt, obj, k)
    for prec in ("mixed", "fp64"):
        out = B.PyPO_HIPd(src, tgt, {"mode": mode, "k": k, key: obj, "precision": prec})
        err = group_err(flat(out), flat(ref))
        print(f"seed {seed} {mode} {prec}: src {nx}x{ny}, lam {lam:.3g}: {err:.2e}")
        assert err <= TOL[prec]


@pytest.mark.parametrize("shape", [(2, 2), (1, 5), (5, 1), (1, 1)])
def test_degenerate_source_grids_follow_the_checker(hip, port, shape):
    """Source grids with one or two points per axis.  A one-point xy axis
    divides its extent by zero in the reference's generator (NaN coordinates,
    infinite areas; a uv grid with one v point stays finite); the GPU grids,
    and every mode's result, have the same non-finite entries as the CPU
    checker and agree with it elsewhere."""
    rng = np.random.default_rng(sum(shape))
    tgt = G.plane({"gmode": "xy", "lims_x": [-3, 3], "lims_y": [-3, 3], "gridsize": [4, 3]})
    tgt["transf"] = G.mat_translate([0, 0, 10.0])
    ff = G.plane({"gmode": "AoE", "lims_Az": [-5, 5], "lims_El": [-5, 5], "gridsize": [5, 4]})
    sources = [G.plane({"gmode": "xy", "lims_x": [-2, 2], "lims_y": [-2, 2], "gridsize": list(shape)}),
               G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 60.0], "gmode": "uv",
                             "lims_u": [0, 25], "lims_v": [0, 360], "gridsize": list(shape)})]
    for src in sources:
        g_ref = OB.generate_grid(port, src)
        g_got = B.generate_grid(src)
        for a in ("x", "y", "z", "area"):
            r, o = np.ravel(getattr(g_ref, a)), np.ravel(getattr(g_got, a))
            np.testing.assert_array_equal(np.isfinite(o), np.isfinite(r))
            np.testing.assert_allclose(o[np.isfinite(r)], r[np.isfinite(r)], rtol=1e-13, atol=1e-12)
        cur = T.currents(*[rng.normal(size=shape) + 1j * rng.normal(size=shape) for _ in range(6)])
        for mode, target in (("EH", tgt), ("JM", tgt), ("FF", ff)):
            ref, _, _ = OB.propagate(port, mode, src, target, cur, 2 * np.pi / 1.5)
            r = flat(ref)
            for prec in ("mixed", "fp64"):
                o = flat(B.PyPO_HIPd(src, target, {"mode": mode, "k": 2 * np.pi / 1.5, "s_current": cur,
                                                   "precision": prec}))
                fin = np.isfinite(r)
                np.testing.assert_array_equal(np.isfinite(o), fin, err_msg=f"{src['gmode']} {mode} {prec}")
                if fin.all():
                    assert group_err(o, r) <= TOL[prec], (src["gmode"], mode, prec)


def test_ff_shadow_test_with_scaled_normals(hip, port):
    """A transf whose rotation part scales (1.5 R) gives grid normals of
    length 1.5; the reference's shadow test n.r < 0 only reads their sign.
    The kernels skip that test per tile only where the tile's normal cone and
    the targets' direction cone prove it rejects nothing, and those cones are
    built from directions, so every kernel still agrees with the checker.
    Built from the length-1.5 normals, the tile cones came out too narrow
    and every kernel, FP64 included, skipped tests that reject pairs here
    (0.22-0.28 of peak off).  0.4 % of the pairs are shadowed over the
    +-40 deg far field, and no pair has |n.r| below 9e-6, so the FP32
    per-pair test decides as the reference's FP64 one does."""
    src = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 30.0], "gmode": "uv",
                        "lims_u": [0, 40], "lims_v": [0, 360], "gridsize": [40, 36]})
    a = np.radians(21.0)
    rot = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    m = np.eye(4)
    m[:3, :3] = 1.5 * rot
    m[:3, 3] = [2.0, -1.0, 3.0]
    src["transf"] = m
    ff = G.plane({"gmode": "AoE", "lims_Az": [-40, 40], "lims_El": [-40, 40], "gridsize": [61, 61]})
    rng = np.random.default_rng(7)
    cur = T.currents(*[rng.normal(size=(40, 36)) + 1j * rng.normal(size=(40, 36)) for _ in range(6)])
    k = 2 * np.pi / 2.0
    ref, _, _ = OB.propagate(port, "FF", src, ff, cur, k)
    try:
        for kind, prec in (("matrix", "mixed"), ("vector", "mixed"), ("matrix", "fp64")):
            B.set_ff_kernel(kind)
            out = B.PyPO_HIPd(src, ff, {"mode": "FF", "k": k, "s_current": cur, "precision": prec})
            err = group_err(flat(out), flat(ref))
            print(f"scaled normals, {kind} {prec}: {err:.2e}")
            assert err <= TOL[prec], (kind, prec, err)
    finally:
        B.set_ff_kernel("matrix")


@pytest.mark.parametrize("seed", range(60))
def test_po_random_surfaces_transforms_and_directions(hip, port, seed):
    """Every mode (by seed) on random surfaces of every type and grid mode,
    flipped or not, under random rigid and scaling transforms, with random
    wavelength, exp = fwd or bwd and, in EHP, epsilon; near-field targets are
    random surfaces 60-200 mm away in a random direction, far-field targets
    AoE grids up to +-80 deg.  Through the drop-in call, mixed to 1e-5 and
    FP64 to 1e-10 of each field's peak, same non-finite entries.  A pair
    whose |n.R| is within rounding of 0 (below 1e-5 for the mixed check,
    1e-12 for the FP64 one) makes the shadow test a rounding matter, so such
    cases skip that precision."""
    from helpers import random_surface
    rng = np.random.default_rng(5000 + seed)
    src = random_surface(rng)
    n = gridsize(src)
    lam = float(10 ** rng.uniform(np.log10(0.3), np.log10(5.0)))
    k = 2 * np.pi / lam
    mode = ("FF", "JM", "EH", "JMEH", "EHP", "scalar")[seed % 6]
    exp = "fwd" if rng.random() < 0.7 else "bwd"
    eps = float(rng.uniform(1.0, 4.0)) if mode == "EHP" and rng.random() < 0.5 else 1.0
    gs = OB.generate_grid(port, src)
    nrm = np.stack([np.ravel(gs.nx), np.ravel(gs.ny), np.ravel(gs.nz)])
    pts = np.stack([np.ravel(gs.x), np.ravel(gs.y), np.ravel(gs.z)])
    if mode == "FF":
        deg = float(rng.uniform(1, 80))
        tgt = G.plane({"gmode": "AoE", "lims_Az": [-deg, deg], "lims_El": [-deg, deg],
                       "gridsize": [int(rng.integers(3, 30)), int(rng.integers(3, 30))]})
        gt = OB.generate_grid(port, tgt)
        th, ph = np.ravel(gt.x), np.ravel(gt.y)
        dots = nrm.T @ np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)])
    else:
        tgt = random_surface(rng)
        off = rng.normal(size=3)
        tgt["transf"] = G.mat_translate(off * rng.uniform(60, 200) / np.linalg.norm(off), tgt["transf"])
        gt = OB.generate_grid(port, tgt)
        R = np.stack([np.ravel(gt.x), np.ravel(gt.y), np.ravel(gt.z)])[:, None, :] - pts[:, :, None]
        dots = np.einsum("is,ist->st", nrm, R / np.linalg.norm(R, axis=0))
    mind = 1.0 if mode == "scalar" else float(np.min(np.abs(dots)))   # the scalar sum has no shadow test
    if mode == "scalar":
        obj, key = T.scalarfield(rng.normal(size=n) + 1j * rng.normal(size=n)), "s_scalarfield"
    else:
        comps = [rng.normal(size=n) + 1j * rng.normal(size=n) for _ in range(6)]
        if rng.random() < 0.3:
            comps[3:] = [np.zeros(n, complex)] * 3
        obj, key = T.currents(*comps), "s_current"
    ref = flat(OB.propagate(port, mode, src, tgt, obj, k, epsilon=eps, exp=exp)[0])
    checked = 0
    for prec, floor in (("mixed", 1e-5), ("fp64", 1e-12)):
        if mind < floor:
            continue
        out = flat(B.PyPO_HIPd(src, tgt, {"mode": mode, "k": k, key: obj, "precision": prec, "exp": exp,
                                          "epsilon": eps}))
        fin = np.isfinite(ref)
        np.testing.assert_array_equal(np.isfinite(out), fin)
        cols = fin.all(0)
        err = group_err(out[:, cols], ref[:, cols]) if cols.any() else 0.0
        print(f"seed {seed} {mode} {prec} {exp} eps {eps:.2f}: type {src['type']} gmode {src['gmode']} "
              f"{n}, min|n.R| {mind:.1e}: {err:.2e}")
        assert err <= TOL[prec]
        checked += 1
    assert checked or mind < 1e-12


@pytest.mark.parametrize("seed", range(12))
def test_many_source_splits_every_mode(hip, port, seed):
    """3 600-25 600 sources over one to three target blocks: the sums split
    the sources into eight or more parts, so the workgroups take their work
    items in the XCD-aware order (each XCD a contiguous run of the
    split-major list), with a partial last tile.  Every mode (by seed),
    mixed to 1e-5 and FP64 to 1e-10 of peak against the checker."""
    rng = np.random.default_rng(27000 + seed)
    mode = ("FF", "JM", "EH", "JMEH", "EHP", "scalar")[seed % 6]
    n = [int(rng.integers(60, 160)), int(rng.integers(60, 160))]
    src = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 80.0], "gmode": "uv",
                        "lims_u": [0, 60], "lims_v": [0, 360], "gridsize": n})
    k = 2 * np.pi / float(rng.uniform(1.0, 4.0))
    if mode == "FF":
        deg = float(rng.uniform(1, 20))
        tgt = G.plane({"gmode": "AoE", "lims_Az": [-deg, deg], "lims_El": [-deg, deg],
                       "gridsize": [int(rng.integers(3, 30)), int(rng.integers(3, 30))]})
    else:
        tgt = G.plane({"gmode": "xy", "lims_x": [-10, 10], "lims_y": [-10, 10],
                       "gridsize": [int(rng.integers(3, 25)), int(rng.integers(3, 25))], "flip": True})
        tgt["transf"] = G.mat_translate([0, 0, 80.0])
    if mode == "scalar":
        obj, key = T.scalarfield(rng.normal(size=n) + 1j * rng.normal(size=n)), "s_scalarfield"
    else:
        x = np.linspace(-1, 1, n[0])[:, None]
        amp = np.exp(-x ** 2) * np.ones(n) + 0.1 * rng.normal(size=n)
        z = np.zeros(n, complex)
        obj, key = T.currents(amp + 0j, 0.3 * amp + 0j, z, z, z, z), "s_current"
    ref = flat(OB.propagate(port, mode, src, tgt, obj, k)[0])
    for prec in ("mixed", "fp64"):
        out = flat(B.PyPO_HIPd(src, tgt, {"mode": mode, "k": k, key: obj, "precision": prec}))
        err = group_err(out, ref)
        print(f"seed {seed} {mode} {prec}: {n[0]}x{n[1]} sources: {err:.2e}")
        assert err <= TOL[prec]
