"""TEST INFRASTRUCTURE ONLY: CPU checker for the far-field power adjoint.

The reference has no adjoint; its forward path is restated in the C port
(pypo_oracle.cpp, pinned against the reference's own outputs), and this file
adds, in numpy FP64:
  * forward(...): feed -> JM on the displaced dish -> FF through the C port,
    the dish moved as the reference moves surfaces for surface errors
    (InterfaceReflector.cpp:852-866: points along their normals, normals and
    area elements unchanged);
  * power_grad(...): the analytic adjoint of P_t = |E_t|^2 with respect to
    the per-point displacements -- the incident-field derivative along the
    normal of the near-field integrand of Propagation.h:1084-1119 plus the
    far-field phase derivative of Propagation.h:1656-1781.
power_grad is pinned in tests/test_oracle.py against central finite
differences of forward() (the C port), term by term: dropping either the
current-derivative term or the phase term fails that check.

Used only by tests/ (small grids: the pair loops are dense numpy).
"""

import numpy as np

from . import bind as OB

PI_F = float(np.float32(3.14159265359))   # the reference's float literal widened


def trapezoid_weights(gmode, ncx, ncy):
    """Propagation.h:1123-1135: half on the xu edges; half on the yv edges
    unless gmode == 1 (uv)."""
    wx = np.ones(ncx)
    wx[[0, -1]] = 0.5
    wy = np.ones(ncy)
    if gmode != 1:
        wy[[0, -1]] = 0.5
    return np.outer(wx, wy).ravel()


def displace(rows0, d):
    rows = np.array(rows0, dtype=np.float64, copy=True)
    rows[0:3] += d[None, :] * rows0[3:6]
    return rows


def forward(lib, feed_rows, feed_cells, feed_cur, dish_rows0, dish_cells, d, ff_rows, k, exp="fwd"):
    """(E (12, n_ff), J (12, n_dish)) through the C port."""
    dish = displace(dish_rows0, d)
    J = OB.propagate_on_grids(lib, "JM", feed_rows, feed_cells, dish, feed_cur, k, exp=exp)
    E = OB.propagate_on_grids(lib, "FF", dish, dish_cells, ff_rows, J, k, exp=exp)
    return E, J


def power(E):
    return (E[0:6] ** 2).sum(0)


def _c6(rows, off):
    """(12, n) c2Bundle rows -> (3, n) complex vector starting at field `off` (0 or 6)."""
    return rows[off:off + 3] + 1j * rows[off + 3:off + 6]


def _cross(a, b):
    return np.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def currents_and_derivative(feed_rows, feed_cells, feed_cur, dish_rows, k, tdir):
    """J = 2 pref n x H and dJ/dd = 2 pref n x (n.grad)H at each dish point."""
    pref = k * k / (4 * PI_F)
    w = trapezoid_weights(*feed_cells) * feed_rows[6]
    Js, Ms = _c6(feed_cur, 0) * w, _c6(feed_cur, 6) * w          # (3, ns)
    src, sn = feed_rows[0:3], feed_rows[3:6]
    p, n = dish_rows[0:3], dish_rows[3:6]
    nd = p.shape[1]
    H = np.zeros((3, nd), complex)
    dH = np.zeros((3, nd), complex)
    for lo in range(0, nd, 512):
        hi = min(nd, lo + 512)
        dv = p[:, None, lo:hi] - src[:, :, None]                  # (3, ns, m)
        R = np.sqrt((dv ** 2).sum(0))
        h = dv / R
        lit = ((sn[:, :, None] * h).sum(0) >= 0)
        nn = n[:, None, lo:hi]
        nu = (nn * h).sum(0)
        dh = (nn - nu * h) / R
        q = 1 / (k * R)
        dq = -k * q * q * nu
        s1 = tdir * q ** 2 + 1j * (q ** 3 - q)
        s2 = -3 * tdir * q ** 2 + 1j * (q - 3 * q ** 3)
        s3 = -q ** 2 + 1j * tdir * q
        ds1 = (2 * tdir * q + 1j * (3 * q ** 2 - 1)) * dq
        ds2 = (-6 * tdir * q + 1j * (1 - 9 * q ** 2)) * dq
        ds3 = (-2 * q + 1j * tdir) * dq
        G = np.exp(1j * tdir * k * R)
        dG = 1j * tdir * k * nu * G
        M = Ms[:, :, None]
        Jc = Js[:, :, None]
        mh = (M * h).sum(0)
        mdh = (M * dh).sum(0)
        jxh = _cross(Jc, h)
        jxdh = _cross(Jc, dh)
        T = M * s1 + mh * h * s2 - jxh * s3
        dT = M * ds1 + (mdh * h + mh * dh) * s2 + mh * h * ds2 - jxdh * s3 - jxh * ds3
        H[:, lo:hi] = (lit * G * T).sum(1)
        dH[:, lo:hi] = (lit * (dG * T + G * dT)).sum(1)
    return 2 * pref * _cross(n, H), 2 * pref * _cross(n, dH)


def power_grad(feed_rows, feed_cells, feed_cur, dish_rows, dish_cells, ff_rows, E, g, k, exp="fwd"):
    """dL/dd for dL/dP_t = g[t]; dish_rows is the displaced dish, E the FF
    result on ff_rows (12, n_ff)."""
    tdir = -1.0 if exp == "fwd" else 1.0
    pref = k * k / (4 * PI_F)
    J, dJ = currents_and_derivative(feed_rows, feed_cells, feed_cur, dish_rows, k, tdir)
    th, ph = ff_rows[0], ff_rows[1]
    r = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)])   # (3, nt)
    A = 2 * g * (-1j * pref) * np.conj(_c6(E, 0))                                    # (3, nt)
    p, n = dish_rows[0:3], dish_rows[3:6]
    nd = p.shape[1]
    B = np.zeros((3, nd), complex)
    Gs = np.zeros((3, nd), complex)
    for lo in range(0, nd, 512):
        hi = min(nd, lo + 512)
        phase = np.exp(-1j * tdir * k * (p[:, lo:hi].T @ r))         # (m, nt)
        w = n[:, lo:hi].T @ r                                          # (m, nt) = n.r
        e = phase * (w >= 0)
        B[:, lo:hi] = (e @ A.T).T
        Gs[:, lo:hi] = ((e * w) @ A.T).T
    c = trapezoid_weights(*dish_cells) * dish_rows[6]
    return c * (np.real((B * dJ).sum(0)) + tdir * k * np.imag((J * Gs).sum(0)))
