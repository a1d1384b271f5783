"""Far-field power adjoint on the GPU (pypo_hip_displace_dev,
pypo_hip_ff_power_grad_dev, pypo_amd.adjoint.FarFieldPower):

* small case: displaced grid, forward JM -> FF and the gradient against the
  CPU checker (oracle/adjoint.py, itself pinned against finite differences
  of the C port): FP64 to 1e-10, mixed to 1e-5;
* the mixed reverse sum on the matrix cores and on the vector ALUs, with
  the matrix kernel's follower phase on every tile (+-6 deg) and on one
  workgroup of three (+-70 deg, with shadowed pairs);
* the autograd wrapper gives the same gradient as the explicit call;
* full size (601x601 dish, D = 3 m, lambda = 1 mm, 36 Zernike modes of
  RMS 20 um, 201x201 far field): the mixed-precision gradient against
  fourth-order central differences of the FP64 GPU forward on 100 random
  points, and against the FP64 adjoint gradient, to 1e-4 relative.
"""

import time

import numpy as np
import pytest
import torch

from helpers import group_err, small_shape_case
from oracle import adjoint as OA
from pypo_amd import adjoint as PA

pytestmark = pytest.mark.gpu


def _rows_to_c6(t):
    a = t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else t
    return np.stack([a[0] + 1j * a[3], a[1] + 1j * a[4], a[2] + 1j * a[5],
                     a[6] + 1j * a[9], a[7] + 1j * a[10], a[8] + 1j * a[11]])


@pytest.fixture(scope="module")
def small(hip, port):
    c = small_shape_case(port)
    prob = PA.ShapeProblem(c["feed"], c["cur"], c["dish"], c["ff"], c["k"])
    d = torch.tensor(c["d"], dtype=torch.float64, device="cuda")
    E0, _ = OA.forward(port, c["fr"], c["fcells"], c["fc"], c["dr"], c["dcells"], np.zeros_like(c["d"]),
                       c["ffr"], c["k"])
    c["T"] = OA.power(E0)
    c["E"], c["J"] = OA.forward(port, c["fr"], c["fcells"], c["fc"], c["dr"], c["dcells"], c["d"], c["ffr"],
                                c["k"])
    c["g"] = np.random.default_rng(5).standard_normal(c["T"].size) / c["T"].max()
    return c, prob, d


def test_displaced_grid_matches_checker(small):
    c, prob, d = small
    g = prob.displaced(d).cpu().numpy()
    np.testing.assert_allclose(g, OA.displace(c["dr"], c["d"]), rtol=1e-14, atol=1e-12)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_forward_matches_checker(small, precision, tol):
    c, prob, d = small
    E = prob.far_field(d, precision=precision)
    err = group_err(_rows_to_c6(E), _rows_to_c6(c["E"]))
    print(f"{precision}: forward error {err:.2e}")
    assert err <= tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_gradient_matches_checker(small, precision, tol):
    c, prob, d = small
    E = torch.tensor(c["E"], device="cuda")
    g = torch.tensor(c["g"], device="cuda")
    got = prob.power_grad(d, E, g, precision=precision).cpu().numpy()
    ref = OA.power_grad(c["fr"], c["fcells"], c["fc"], OA.displace(c["dr"], c["d"]), c["dcells"], c["ffr"],
                        c["E"], c["g"], c["k"])
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"{precision}: gradient error {rel:.2e} (norm), {np.abs(got - ref).max() / np.abs(ref).max():.2e} (max)")
    assert rel <= tol and np.abs(got - ref).max() <= tol * np.abs(ref).max()


def _serpentine(nx, ny):
    """The order ShapeProblem hands the far-field targets to the backward
    sum in: row by row, every other row reversed."""
    idx = np.arange(nx * ny).reshape(nx, ny)
    idx[1::2] = idx[1::2, ::-1]
    return idx.reshape(-1)


def _reverse_follower_share(dish_xyz, th, ph, k):
    """Share of (256-point workgroup, 256-target tile) pairs that the matrix
    reverse sum runs in its follower-phase form: per group of four targets one
    FP64 phase (the second target), FP32 offsets d = u_t - u_leader for the
    other three (u = r k / 2 pi), taken when 1 + sum_i |d_i| max|p_i| <= 8
    over the workgroup's dish points p for every follower of the tile.  Slots
    past the last target carry its direction and have d = 0."""
    u = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)]) * k / (2 * np.pi)
    n = u.shape[1]
    npad = -(-n // 256) * 256
    up = np.repeat(u[:, -1:], npad, axis=1)
    up[:, :n] = u
    idx = np.arange(npad)
    lead = (idx & ~3) | 1
    d = np.abs(up - up[:, lead])
    d[:, idx >= n] = 0
    d[:, idx == lead] = 0
    nd = dish_xyz.shape[1]
    nw = -(-nd // 256)
    pp = np.zeros((3, nw * 256))
    pp[:, :nd] = np.abs(dish_xyz)
    pw = pp.reshape(3, nw, 256).max(axis=2)
    dt = d.reshape(3, npad // 256, 256)
    fast = np.stack([(1 + pw.T @ dt[:, t, :]).max(axis=1) <= 8 for t in range(npad // 256)], axis=1)
    return fast.mean()


@pytest.mark.parametrize("ff_deg", [6.0, 70.0])
def test_gradient_matrix_and_vector_reverse_sums(hip, port, ff_deg):
    """The mixed backward far-field sum on the matrix cores (24 output columns
    B, r_x A, r_y A, r_z A; G recombined with the normals in FP64) and on the
    vector ALUs, against the checker; the +-70 deg far field puts pairs in
    the shadow test, which the cone test cannot skip there."""
    from pypo_amd import bind as B
    c = small_shape_case(port, ff_deg=ff_deg)
    # +-6 deg: every tile takes the follower phase; +-70 deg: one workgroup of three
    perm = _serpentine(*map(int, c["ff"]["gridsize"]))
    share = _reverse_follower_share(c["dr"][:3], c["ffr"][0][perm], c["ffr"][1][perm], c["k"])
    print(f"ff +-{ff_deg} deg: follower-phase share {share:.3f}")
    assert (share == 1) if ff_deg < 10 else 0 < share < 1
    prob = PA.ShapeProblem(c["feed"], c["cur"], c["dish"], c["ff"], c["k"])
    d = torch.tensor(c["d"], dtype=torch.float64, device="cuda")
    E, _ = OA.forward(port, c["fr"], c["fcells"], c["fc"], c["dr"], c["dcells"], c["d"], c["ffr"], c["k"])
    g = np.random.default_rng(7).standard_normal(E.shape[1]) / np.abs(E).max() ** 2
    ref = OA.power_grad(c["fr"], c["fcells"], c["fc"], OA.displace(c["dr"], c["d"]), c["dcells"], c["ffr"],
                        E, g, c["k"])
    grads = {}
    try:
        for kind in ("vector", "matrix"):
            B.set_ff_kernel(kind)
            got = prob.power_grad(d, torch.tensor(E, device="cuda"), torch.tensor(g, device="cuda"),
                                  precision="mixed").cpu().numpy()
            rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            print(f"ff +-{ff_deg} deg, {kind}: gradient error {rel:.2e}")
            assert rel <= 1e-5 and np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
            grads[kind] = got
    finally:
        B.set_ff_kernel("matrix")
    assert np.linalg.norm(grads["matrix"] - grads["vector"]) <= 2e-6 * np.linalg.norm(ref)


def test_autograd_function_matches_explicit_call(small):
    c, prob, d0 = small
    T = torch.tensor(c["T"], device="cuda")
    d = d0.clone().requires_grad_(True)
    P = prob.power(d)
    loss = ((P - T) ** 2).sum() / T.max() ** 2
    loss.backward()
    E = prob.far_field(d0)
    g = 2 * (P.detach() - T) / T.max() ** 2
    ref = prob.power_grad(d0, E, g)
    torch.testing.assert_close(d.grad, ref, rtol=1e-12, atol=0)
    assert torch.allclose(P.detach(), (E[0:6] ** 2).sum(0))


def test_bad_arguments_raise(small):
    from pypo_amd import bind as B
    c, prob, d = small
    with pytest.raises(ValueError):
        prob.displaced(d[:-1])
    with pytest.raises(ValueError):
        prob.displaced(d.float())
    with pytest.raises(ValueError):
        prob.power_grad(d, prob.far_field(d), torch.zeros(3, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):   # far field of the wrong length
        prob.power_grad(d, prob.far_field(d)[:, :-1], torch.zeros(prob.n_ff, dtype=torch.float64, device="cuda"))
    lib = B.loadHIPlib()
    # NULL arrays and an unknown precision are rejected before any launch
    assert lib.pypo_hip_ff_power_grad_dev(None, 1, 1, 1, None, 0, 1, 1, None, None, None, None, 1.0, -1.0,
                                          0, None, None) != 0
    with pytest.raises(B.PyPOHIPError):
        B.check(lib, "pypo_hip_ff_power_grad_dev")
    g = torch.zeros(prob.n_ff, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        prob.power_grad(d, prob.far_field(d), g, precision="fp16")


@pytest.fixture(scope="module")
def full(hip):
    feed, cur, dish, ff, k, radius = PA.surface_error_workload()
    prob = PA.ShapeProblem(feed, cur, dish, ff, k, precision="mixed")
    g0 = prob.g_dish0.cpu().numpy()
    d_np = PA.zernike_displacement(g0[0], g0[1], radius, sigma=0.02, seed=2024)
    d = torch.tensor(d_np, device="cuda")
    T = (prob.far_field(torch.zeros_like(d), precision="fp64")[0:6] ** 2).sum(0)
    return prob, d, T


def _loss(prob, d, T, precision):
    P = (prob.far_field(d, precision=precision)[0:6] ** 2).sum(0)
    return (((P - T) ** 2).sum() / T.max() ** 2).item()


def test_full_size_gradient_against_finite_differences(full):
    prob, d0, T = full
    assert prob.n_dish == 601 * 601 and prob.n_ff == 201 * 201
    gf = prob.g_ff.cpu().numpy()
    perm = _serpentine(201, 201)
    share = _reverse_follower_share(prob.g_dish0.cpu().numpy()[:3], gf[0][perm], gf[1][perm], prob.k)
    share_rows = _reverse_follower_share(prob.g_dish0.cpu().numpy()[:3], gf[0], gf[1], prob.k)
    print(f"reverse sum: follower-phase share {share:.3f} (row by row: {share_rows:.3f})")
    assert share == 1 and share_rows < 1
    d = d0.clone().requires_grad_(True)
    t = time.perf_counter()
    P = prob.power(d)
    loss = ((P - T) ** 2).sum() / T.max() ** 2
    loss.backward()
    torch.cuda.synchronize()
    print(f"forward + backward (mixed): {(time.perf_counter() - t) * 1e3:.1f} ms, loss {loss.item():.4e}")
    g_mixed = d.grad.cpu().numpy()

    # FP64 adjoint gradient
    E64 = prob.far_field(d0, precision="fp64")
    P64 = (E64[0:6] ** 2).sum(0)
    g64 = prob.power_grad(d0, E64, 2 * (P64 - T) / T.max() ** 2, precision="fp64").cpu().numpy()
    rel64 = np.linalg.norm(g_mixed - g64) / np.linalg.norm(g64)
    print(f"mixed vs FP64 adjoint gradient: {rel64:.2e} (norm)")
    assert rel64 <= 1e-4

    # central differences of the FP64 forward, O(h^4), on 100 random points
    rng = np.random.default_rng(7)
    idx = rng.choice(prob.n_dish, 100, replace=False)
    h = 2e-3
    fd = np.empty(idx.size)
    dd = d0.clone()
    for m, i in enumerate(idx):
        f = []
        for s in (2, 1, -1, -2):
            dd[i] = d0[i] + s * h
            f.append(_loss(prob, dd, T, "fp64"))
        dd[i] = d0[i]
        fd[m] = (-f[0] + 8 * f[1] - 8 * f[2] + f[3]) / (12 * h)
    for name, g in (("mixed", g_mixed), ("fp64", g64)):
        rel = np.linalg.norm(g[idx] - fd) / np.linalg.norm(fd)
        worst = np.abs(g[idx] - fd).max() / np.abs(fd).max()
        print(f"{name} adjoint vs central differences on 100 points: {rel:.2e} (norm), {worst:.2e} (max/peak)")
        assert rel <= 1e-4 and worst <= 1e-4


def test_gradient_with_scaled_dish_normals(hip, port):
    """A dish whose transf scales by 1.23 (grid normals of length 1.23) and a
    +-53.5 deg far field that puts 1.3 % of the (dish point, target) pairs in
    the shadow test (the smallest |n.r| is 1.4e-4, far from rounding): the
    matrix-core reverse sum skips a tile's per-pair tests only where the
    dish-normal and target-direction cones exclude shadowing, and those
    cones are built and widened from directions, so the mixed gradient of
    both reverse sums and the FP64 one agree with the checker.  Built from
    the longer normals, the cones came out too narrow: 335 of 625 dish
    points off by up to 4e-2 of the largest gradient."""
    from pypo_amd import bind as B
    from helpers import grid_rows
    c = small_shape_case(port, seed=29, sigma=0.02933749125690907, ff_deg=53.50493480875001)
    m = np.eye(4)
    m[:3, :3] *= 1.228784237485987
    c["dish"]["transf"] = m @ c["dish"]["transf"]
    c["dr"] = grid_rows(port, c["dish"])
    prob = PA.ShapeProblem(c["feed"], c["cur"], c["dish"], c["ff"], c["k"])
    d = torch.tensor(c["d"], dtype=torch.float64, device="cuda")
    E, _ = OA.forward(port, c["fr"], c["fcells"], c["fc"], c["dr"], c["dcells"], c["d"], c["ffr"], c["k"])
    g = np.random.default_rng(29).standard_normal(E.shape[1]) / np.abs(E).max() ** 2
    ref = OA.power_grad(c["fr"], c["fcells"], c["fc"], OA.displace(c["dr"], c["d"]), c["dcells"], c["ffr"],
                        E, g, c["k"])
    try:
        for kind, prec, bar in (("matrix", "mixed", 1e-5), ("vector", "mixed", 1e-5), ("matrix", "fp64", 1e-10)):
            B.set_ff_kernel(kind)
            got = prob.power_grad(d, torch.tensor(E, device="cuda"), torch.tensor(g, device="cuda"),
                                  precision=prec).cpu().numpy()
            worst = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"scaled dish, {kind} {prec}: largest error {worst:.2e} of the largest gradient")
            assert worst <= bar, (kind, prec, worst)
    finally:
        B.set_ff_kernel("matrix")


@pytest.mark.parametrize("seed", range(8))
def test_random_gradients_against_checker(hip, port, seed):
    """Random small surface-error problems: Zernike displacement RMS 0.01-0.2
    mm, far field over +-1..80 deg (wide ones put pairs in the shadow test),
    four in ten dishes under a scaling transform; the matrix and vector
    mixed reverse sums to 1e-5 and the FP64 one to 1e-10 of the largest
    gradient, against the checker."""
    from pypo_amd import bind as B
    from helpers import grid_rows
    rng = np.random.default_rng(3100 + seed)
    c = small_shape_case(port, seed=seed, sigma=float(rng.uniform(0.01, 0.2)), ff_deg=float(rng.uniform(1, 80)))
    if rng.random() < 0.4:
        m = np.eye(4)
        m[:3, :3] *= rng.uniform(0.7, 1.5)
        c["dish"]["transf"] = m @ c["dish"]["transf"]
        c["dr"] = grid_rows(port, c["dish"])
    prob = PA.ShapeProblem(c["feed"], c["cur"], c["dish"], c["ff"], c["k"])
    d = torch.tensor(c["d"], dtype=torch.float64, device="cuda")
    E, _ = OA.forward(port, c["fr"], c["fcells"], c["fc"], c["dr"], c["dcells"], c["d"], c["ffr"], c["k"])
    g = np.random.default_rng(seed).standard_normal(E.shape[1]) / np.abs(E).max() ** 2
    ref = OA.power_grad(c["fr"], c["fcells"], c["fc"], OA.displace(c["dr"], c["d"]), c["dcells"], c["ffr"],
                        E, g, c["k"])
    try:
        for kind, prec, bar in (("matrix", "mixed", 1e-5), ("vector", "mixed", 1e-5), ("matrix", "fp64", 1e-10)):
            B.set_ff_kernel(kind)
            got = prob.power_grad(d, torch.tensor(E, device="cuda"), torch.tensor(g, device="cuda"),
                                  precision=prec).cpu().numpy()
            worst = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"seed {seed} {kind} {prec}: {worst:.2e}")
            assert worst <= bar
    finally:
        B.set_ff_kernel("matrix")
