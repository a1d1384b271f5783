"""Pins the CPU checker of the far-field power adjoint (oracle/adjoint.py)
against central finite differences of the C port's forward JM -> FF chain
(itself pinned against the reference's outputs), and shows that the check
catches the plausible mistakes: a dropped current-derivative term, a dropped
phase term, a wrong phase sign."""

import numpy as np
import pytest

from helpers import small_shape_case
from oracle import adjoint as OA


@pytest.fixture(scope="module")
def case(port):
    c = small_shape_case(port)
    E0, _ = OA.forward(port, c["fr"], c["fcells"], c["fc"], c["dr"], c["dcells"], np.zeros_like(c["d"]),
                       c["ffr"], c["k"])
    c["T"] = OA.power(E0)
    c["scale"] = c["T"].max() ** 2
    return c


def _loss(port, c, d):
    E, _ = OA.forward(port, c["fr"], c["fcells"], c["fc"], c["dr"], c["dcells"], d, c["ffr"], c["k"])
    P = OA.power(E)
    return ((P - c["T"]) ** 2).sum() / c["scale"], E, P


def _fd(port, c, idx, h=1e-3):
    """Fourth-order central differences (truncation O(h^4))."""
    out = []
    for i in idx:
        def f(s):
            d = c["d"].copy()
            d[i] += s * h
            return _loss(port, c, d)[0]
        out.append((-f(2) + 8 * f(1) - 8 * f(-1) + f(-2)) / (12 * h))
    return np.array(out)


def test_adjoint_oracle_matches_finite_differences(port, case):
    c = case
    L, E, P = _loss(port, c, c["d"])
    g = 2 * (P - c["T"]) / c["scale"]
    ga = OA.power_grad(c["fr"], c["fcells"], c["fc"], OA.displace(c["dr"], c["d"]), c["dcells"], c["ffr"],
                       E, g, c["k"])
    idx = np.random.default_rng(0).choice(c["d"].size, 12, replace=False)
    fd = _fd(port, c, idx)
    assert L > 0
    assert np.linalg.norm(ga[idx] - fd) <= 1e-8 * np.linalg.norm(fd)


def test_finite_difference_check_catches_dropped_terms(port, case, monkeypatch):
    """Each term of the gradient carries weight: the check above fails when
    either is dropped or the far-field phase sign is flipped."""
    c = case
    L, E, P = _loss(port, c, c["d"])
    g = 2 * (P - c["T"]) / c["scale"]
    idx = np.random.default_rng(0).choice(c["d"].size, 12, replace=False)
    fd = _fd(port, c, idx)
    real = OA.currents_and_derivative

    def no_dj(*a, **kw):
        J, dJ = real(*a, **kw)
        return J, 0 * dJ

    def no_j(*a, **kw):
        J, dJ = real(*a, **kw)
        return 0 * J, dJ

    for fake in (no_dj, no_j):
        monkeypatch.setattr(OA, "currents_and_derivative", fake)
        ga = OA.power_grad(c["fr"], c["fcells"], c["fc"], OA.displace(c["dr"], c["d"]), c["dcells"],
                           c["ffr"], E, g, c["k"])
        assert np.linalg.norm(ga[idx] - fd) > 1e-2 * np.linalg.norm(fd), fake.__name__
    monkeypatch.setattr(OA, "currents_and_derivative", real)
    ga = OA.power_grad(c["fr"], c["fcells"], c["fc"], OA.displace(c["dr"], c["d"]), c["dcells"], c["ffr"],
                       E, g, c["k"], exp="bwd")
    assert np.linalg.norm(ga[idx] - fd) > 1e-2 * np.linalg.norm(fd)


def test_zernike_modes_orthonormal():
    """36 modes (orders 0..7), orthonormal over the unit disk (midpoint
    quadrature in polar coordinates)."""
    from pypo_amd import adjoint as PA
    modes = PA.zernike_indices(36)
    assert len(modes) == 36 and modes[-1] == (7, 7) and len(set(modes)) == 36
    nr, nt = 400, 256
    r = (np.arange(nr) + 0.5) / nr
    t = (np.arange(nt) + 0.5) * 2 * np.pi / nt
    R, TH = np.meshgrid(r, t, indexing="ij")
    w = (R / nr * 2 * np.pi / nt).ravel() / np.pi
    Z = np.stack([PA.zernike(n, m, R, TH).ravel() for n, m in modes])
    gram = (Z * w) @ Z.T
    np.testing.assert_allclose(gram, np.eye(36), atol=2e-3)
