"""TEST INFRASTRUCTURE ONLY: CPU checker for the figures of merit of
src/PyPO/Efficiencies.py, restated in numpy FP64 from the reference's text:

  generate_mask   Efficiencies.py:11-41 (ellipse about the centre with the
                  angle taken about the origin; open rectangle; only the outer
                  boundary is returned, the inner one is computed and dropped)
  spillover       90-115   |sum conj(f_ap) f dA|^2 / (sum |f|^2 dA  sum |f_ap|^2 dA_ap)
  taper           117-145  |sum f dA|^2 / sum |f|^2 dA / sum dA over the masked points
  xpol            147-159
  rt_center / rt_tilt / rms   43-88 (rays with |d|^2 > 0.8)

The grids come from the C port's generateGrid with transform=False (the
reference calls BindRefl.generateGrid(surface, transform=False, spheric=True)).
Pinned in tests/test_efficiency_oracle.py against closed forms (a Gaussian
on a disc: spillover and taper; x-pol of proportional components; ring
frames), hand-placed mask points for both shapes, and the identity
spillover = sum_in |f|^2 dA / sum |f|^2 dA.  Used only by tests/.
"""

import numpy as np

from . import bind as OB


def _shape(aperDict):
    s = aperDict.get("shape", "ellipse")
    name = getattr(s, "name", s)
    if isinstance(name, str):
        return 0 if name.lower() in ("ellipse", "ell") else 1
    return int(getattr(s, "value", s))


def generate_mask(x, y, aperDict):
    """Efficiencies._generateMask (Efficiencies.py:11-41)."""
    c = np.asarray(aperDict.get("center", np.zeros(2)), dtype=float)
    if _shape(aperDict) == 0:
        t = np.arctan2(y, x) + np.pi
        o = np.asarray(aperDict["outer"], dtype=float)
        outer = (o[0] * np.cos(t)) ** 2 + (o[1] * np.sin(t)) ** 2
        return (x - c[0]) ** 2 + (y - c[1]) ** 2 < outer
    ox = np.asarray(aperDict["outer_x"], dtype=float)
    oy = np.asarray(aperDict["outer_y"], dtype=float)
    cx = ((x - c[0]) > ox[0]) & ((x - c[0]) < ox[1])
    cy = ((y - c[1]) > oy[0]) & ((y - c[1]) < oy[1])
    return cx & cy


def rest_grid(lib, surface):
    return OB.generate_grid(lib, surface, transform=False, spheric=True)


def spillover(field, grids, aperDict):
    """Efficiencies.calcSpillover (Efficiencies.py:90-115)."""
    x, y, area = grids.x, grids.y, grids.area
    mask = generate_mask(x, y, aperDict)
    field_ap = field * mask.astype(complex)
    area_m = area * mask.astype(int)
    with np.errstate(divide="ignore", invalid="ignore"):
        eff = np.absolute(np.sum(np.conj(field_ap) * field * area)) ** 2 / (
            np.sum(np.absolute(field) ** 2 * area) * np.sum(np.absolute(field_ap) ** 2 * area_m))
    return 0.0 if np.isnan(eff) else float(eff)


def taper(field, grids, aperDict=None):
    """Efficiencies.calcTaper (Efficiencies.py:117-145)."""
    area = grids.area
    if aperDict:
        mask = generate_mask(grids.x, grids.y, aperDict)
        field = field[mask]
        area = area[mask]
    with np.errstate(divide="ignore", invalid="ignore"):
        eff = np.absolute(np.sum(field * area)) ** 2 / np.sum(np.absolute(field) ** 2 * area) / np.sum(area)
    return 0.0 if np.isnan(eff) else float(eff)


def xpol(co, cr):
    """Efficiencies.calcXpol (Efficiencies.py:147-159)."""
    return float(1 - np.sum(np.absolute(cr) ** 2) / (np.sum(np.absolute(co) ** 2) + np.sum(np.absolute(cr) ** 2)))


def _good(fr):
    return np.argwhere((fr.dx ** 2 + fr.dy ** 2 + fr.dz ** 2) > 0.8)


def rt_center(fr):
    """Efficiencies.calcRTcenter (Efficiencies.py:43-57)."""
    g = _good(fr)
    n = len(fr.x[g])
    return np.array([np.sum(fr.x[g]) / n, np.sum(fr.y[g]) / n, np.sum(fr.z[g]) / n])


def rt_tilt(fr):
    """Efficiencies.calcRTtilt (Efficiencies.py:59-73)."""
    g = _good(fr)
    n = len(fr.dx[g])
    t = np.array([np.sum(fr.dx[g]) / n, np.sum(fr.dy[g]) / n, np.sum(fr.dz[g]) / n])
    return t / np.linalg.norm(t)


def rms(fr):
    """Efficiencies.calcRMS (Efficiencies.py:75-88)."""
    g = _good(fr)
    c = rt_center(fr)
    return float(np.sqrt(np.sum((fr.x[g] - c[0]) ** 2 + (fr.y[g] - c[1]) ** 2 + (fr.z[g] - c[2]) ** 2)
                         / len(fr.x[g])))
