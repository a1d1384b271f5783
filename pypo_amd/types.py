"""Result containers of the PO and ray-trace path.

Same attribute names, component order and metadata as the reference's
src/PyPO/PyPOTypes.py (resContainer 8-92, currents 94-134, fields 136-176,
rfield 178-214, scalarfield 216-261, reflGrids 263-308, frame 310-374), so
code written against PyPO's result objects reads these unchanged.
"""

import numpy as np


class resContainer:
    """Six complex components in fixed order; indexable like PyPOTypes.py:49-71."""

    def __init__(self, *args, restype=None):
        self.type = "EH" if restype is None else restype
        self.memlist = []
        self.shape = np.asarray(args[0]).shape
        self.size = int(np.prod(self.shape))
        ax = ["x", "y", "z"]
        for i, arg in enumerate(args):
            n = i // 3
            self.memlist.append(f"{self.type[n]}{ax[i - 3 * n]}")
            setattr(self, self.memlist[i], np.asarray(arg))
        self.surf = None
        self.k = None

    def setMeta(self, surf, k):
        self.surf = surf
        self.k = k

    def __getitem__(self, idx):
        return getattr(self, self.memlist[idx])

    def __setitem__(self, idx, item):
        setattr(self, self.memlist[idx], item)

    def components(self):
        return [getattr(self, m) for m in self.memlist]


class currents(resContainer):
    """J and M surface currents (PyPOTypes.py:94-134)."""

    def __init__(self, Jx, Jy, Jz, Mx, My, Mz):
        super().__init__(Jx, Jy, Jz, Mx, My, Mz, restype="JM")


class fields(resContainer):
    """E and H fields (PyPOTypes.py:136-176)."""

    def __init__(self, Ex, Ey, Ez, Hx, Hy, Hz):
        super().__init__(Ex, Ey, Ez, Hx, Hy, Hz, restype="EH")


class rfield:
    """Real 3-vector field, e.g. the reflected Poynting direction (PyPOTypes.py:178-214)."""

    def __init__(self, Prx, Pry, Prz):
        self.x = np.asarray(Prx)
        self.y = np.asarray(Pry)
        self.z = np.asarray(Prz)
        self.shape = self.x.shape


class scalarfield:
    """Complex scalar field (PyPOTypes.py:216-261)."""

    def __init__(self, S):
        self.S = np.asarray(S)
        self.shape = self.S.shape
        self.surf = None
        self.k = None

    def setMeta(self, surf, k):
        self.surf = surf
        self.k = k


class reflGrids:
    """Realised surface grid (PyPOTypes.py:263-308)."""

    def __init__(self, x, y, z, nx, ny, nz, area):
        self.x, self.y, self.z = x, y, z
        self.nx, self.ny, self.nz = nx, ny, nz
        self.area = area


class frame:
    """Ray frame: positions and directions (PyPOTypes.py:310-374)."""

    def __init__(self, size, x, y, z, dx, dy, dz):
        self.size = int(size)
        self.x, self.y, self.z = (np.asarray(a, dtype=np.float64) for a in (x, y, z))
        self.dx, self.dy, self.dz = (np.asarray(a, dtype=np.float64) for a in (dx, dy, dz))
        self.snapshots = {}
        self.pos = None
        self.ori = None
        self.transf = None

    def setMeta(self, pos, ori, transf):
        self.pos = pos
        self.ori = ori
        self.transf = transf
