"""TEST INFRASTRUCTURE ONLY: CPU checker for the hybrid ray-trace / PO step
(System.runHybridPropagation, System.py:1570-1622; calcRayLen 1304-1372;
interpFrame 1624-1692).  The ray trace is the C port (pypo_oracle.cpp,
pinned against the compiled reference), the ray lengths and phase factors
follow System.py line by line in numpy, and the interpolation is
scipy.interpolate.griddata(method="nearest") -- the routine interpFrame
itself calls.  Used only by tests/.
"""

import numpy as np
from scipy.interpolate import griddata

from . import bind as OB


def carry(lib, target, fr_in, comps, k, start=None, tol=1e-3, t0=1.0, max_iter=0):
    """Trace fr_in onto target; return (fr_out, carried components (6, n)).
    max_iter caps the Newton iterations (0: no cap, as the reference)."""
    fr_out = OB.trace(lib, target, fr_in, tol=tol, t0=t0, max_iter=max_iter)
    L1 = np.sqrt((fr_out.x - fr_in.x) ** 2 + (fr_out.y - fr_in.y) ** 2 + (fr_out.z - fr_in.z) ** 2)
    if start is not None:
        L0 = np.sqrt((fr_in.x - start[0]) ** 2 + (fr_in.y - start[1]) ** 2 + (fr_in.z - start[2]) ** 2)
        expo = np.exp(1j * k * L1) * np.sqrt(L0 / (2 * L1 + L0))
    else:
        expo = np.exp(1j * k * L1)
    return fr_out, np.stack([expo * np.ravel(c) for c in comps])


def interp_nearest(points, comps, grid_xyz):
    """griddata nearest of each complex component, real and imaginary parts
    separately as interpFrame does (System.py:1662-1668)."""
    out = []
    for c in comps:
        r = griddata(points, np.real(c), grid_xyz, method="nearest")
        i = griddata(points, np.imag(c), grid_xyz, method="nearest")
        out.append(r + 1j * i)
    return np.stack(out)
