"""Coefficients of sincos_rev_f64 (common.hpp): near-minimax polynomials
sin(2 pi r) = r P(r^2) (P of degree 6) and cos(2 pi r) = Q(r^2) (Q of degree 7)
on |r| <= 1/8, by interpolation at Chebyshev nodes of t = r^2 in [0, 1/64] in
60-digit decimal arithmetic; prints the coefficients rounded to double (hex)
and the largest error of those rounded polynomials, evaluated exactly, over
4001 points.  Run: python3 pypo_amd/csrc/sincos_fit.py"""
from decimal import Decimal as D, getcontext
import math
getcontext().prec = 60
PI = D("3.14159265358979323846264338327950288419716939937510582097494459")
def dsin(x):
    s, term, k = D(0), x, 1
    while abs(term) > D(10) ** -58:
        s += term; term = -term * x * x / ((k + 1) * (k + 2)); k += 2
    return s
def dcos(x):
    s, term, k = D(0), D(1), 0
    while abs(term) > D(10) ** -58:
        s += term; term = -term * x * x / ((k + 1) * (k + 2)); k += 2
    return s
def dsqrt(x): return x.sqrt()
def dcosn(k, n):   # cos(k pi / n) for the nodes
    return dcos(PI * k / n)
def solve(A, b):
    n = len(b); M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for c in range(n):
        p = max(range(c, n), key=lambda r: abs(M[r][c])); M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c:
                f = M[r][c] / M[c][c]
                M[r] = [M[r][j] - f * M[c][j] for j in range(n + 1)]
    return [M[i][n] / M[i][i] for i in range(n)]
T = D(1) / 64   # t = r^2 in [0, 1/64]
def fit(func, deg):
    n = deg + 1
    nodes = [T / 2 * (1 - dcosn(2 * k + 1, 2 * n)) for k in range(n)]   # Chebyshev nodes in [0, T]
    A = [[(t ** j if j else D(1)) for j in range(n)] for t in nodes]
    return solve(A, [func(t) for t in nodes])
sin_over_r = lambda t: dsin(2 * PI * dsqrt(t)) / dsqrt(t) if t > 0 else 2 * PI
cos_of = lambda t: dcos(2 * PI * dsqrt(t))
for name, func, deg, rel in (("S", sin_over_r, 6, True), ("C", cos_of, 7, False)):
    c = fit(func, deg)
    cd = [float(x) for x in c]   # rounded to double
    # max error of the double-rounded coefficients, evaluated exactly, over 4001 points of r in [0, 1/8]
    worst = D(0)
    for i in range(4001):
        r = D(i) / 4000 / 8
        t = r * r
        p = D(0)
        for x in reversed(cd):
            p = p * t + D(x)
        val = r * p if name == "S" else p
        ref = dsin(2 * PI * r) if name == "S" else dcos(2 * PI * r)
        err = abs(val - ref)
        worst = max(worst, err)
    print(name, "deg", 2 * deg + (1 if name == "S" else 0), "max abs err (exact eval)", float(worst))
    for j, x in enumerate(cd):
        print(f"  {name}{j} = {x.hex()}  // {x:.17g}")
