"""Shared comparison helpers for the parity tests."""

import numpy as np


def flat(o):
    """Result object(s) -> (rows, n) complex array in the reference's component order."""
    if isinstance(o, list):
        return np.concatenate([flat(x) for x in o])
    if hasattr(o, "components"):
        return np.stack([np.ravel(c) for c in o.components()])
    if hasattr(o, "S"):
        return np.ravel(o.S)[None, :]
    return np.stack([np.ravel(o.x), np.ravel(o.y), np.ravel(o.z)]).astype(complex)


def group_err(a, b, group=3):
    """Max |a - b| relative to the peak |b|, per group of `group` rows (one
    field: E, H, J, M or S_r each has its own peak).  Returns the worst group."""
    a = np.asarray(a)
    b = np.asarray(b)
    worst = 0.0
    for g in range(0, b.shape[0], group):
        pk = np.max(np.abs(b[g:g + group]))
        if pk == 0:
            d = np.max(np.abs(a[g:g + group]))
        else:
            d = np.max(np.abs(a[g:g + group] - b[g:g + group])) / pk
        worst = max(worst, float(d))
    return worst
