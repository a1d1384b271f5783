"""Result containers of the PO and ray-trace path.

Same attribute names, component order and metadata as the reference's
src/PyPO/PyPOTypes.py (resContainer 8-92, currents 94-134, fields 136-176,
rfield 178-214, scalarfield 216-261, reflGrids 263-308, frame 310-374), so
code written against PyPO's result objects reads these unchanged.

Results that System computes on the GPU stay in device memory
(`on_device`): the object holds the library's output rows (a float64 torch
tensor) and copies them to host numpy arrays the first time one of its value
attributes is read or assigned.  From then on the host arrays are the values
(the device rows are dropped, so host edits are never shadowed by a stale
device copy).  Until then `device_rows()` hands the rows to the next
propagation, ray trace or reduction without a PCIe round trip.
"""

import numpy as np


class _DeviceBacked:
    """Lazy host copy of values that live on the device.  Subclasses name their
    value attributes (_value_names) and convert the rows (_host_values)."""

    def _init_device(self, rows):
        self.__dict__["_rows"] = rows

    def device_rows(self):
        """The device tensor holding the values, or None once they live on the host."""
        return self.__dict__.get("_rows")

    def _value_names(self):
        return ()

    def _to_host(self):
        rows = self.__dict__.get("_rows")
        if rows is None:
            return
        a = rows.detach().cpu().numpy()     # one device-to-host copy of every row
        for name, v in zip(self._value_names(), self._host_values(a)):
            self.__dict__[name] = v
        self.__dict__["_rows"] = None

    def __getattr__(self, name):
        # only reached for attributes not set yet: the values of a device result
        d = self.__dict__
        if d.get("_rows") is not None and name in self._value_names():
            self._to_host()
            return d[name]
        raise AttributeError(f"{type(self).__name__} object has no attribute {name!r}")

    def __setattr__(self, name, value):
        if self.__dict__.get("_rows") is not None and name in self._value_names():
            self._to_host()
        object.__setattr__(self, name, value)


def _complex_rows(a, v):
    """Complex 3-vector v (0, 1, ...) of c2Bundle-order rows a."""
    return [a[6 * v + i] + 1j * a[6 * v + 3 + i] for i in range(3)]


class resContainer(_DeviceBacked):
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

    @classmethod
    def on_device(cls, rows, shape):
        """Container over (12, n) float64 device rows in c2Bundle order (Re x y z,
        Im x y z of the first vector, then of the second)."""
        obj = cls.__new__(cls)
        obj._init_device(rows)
        restype = cls.RESTYPE
        obj.type = restype
        obj.memlist = [f"{restype[i // 3]}{'xyz'[i % 3]}" for i in range(6)]
        obj.shape = tuple(int(v) for v in shape)
        obj.size = int(np.prod(obj.shape))
        obj.surf = None
        obj.k = None
        return obj

    def _value_names(self):
        return self.__dict__.get("memlist", ())

    def _host_values(self, a):
        return [c.reshape(self.shape) for c in _complex_rows(a, 0) + _complex_rows(a, 1)]

    def component_rows(self, idx):
        """(2, n) device tensor (Re, Im) of component idx while the values live on
        the device, else None."""
        rows = self.device_rows()
        if rows is None:
            return None
        v, a = divmod(int(idx), 3)
        import torch
        return torch.stack([rows[6 * v + a], rows[6 * v + 3 + a]])

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

    RESTYPE = "JM"

    def __init__(self, Jx, Jy, Jz, Mx, My, Mz):
        super().__init__(Jx, Jy, Jz, Mx, My, Mz, restype="JM")


class fields(resContainer):
    """E and H fields (PyPOTypes.py:136-176)."""

    RESTYPE = "EH"

    def __init__(self, Ex, Ey, Ez, Hx, Hy, Hz):
        super().__init__(Ex, Ey, Ez, Hx, Hy, Hz, restype="EH")


class rfield(_DeviceBacked):
    """Real 3-vector field, e.g. the reflected Poynting direction (PyPOTypes.py:178-214)."""

    def __init__(self, Prx, Pry, Prz):
        self.x = np.asarray(Prx)
        self.y = np.asarray(Pry)
        self.z = np.asarray(Prz)
        self.shape = self.x.shape

    @classmethod
    def on_device(cls, rows, shape):
        """rfield over (3, n) float64 device rows x, y, z."""
        obj = cls.__new__(cls)
        obj._init_device(rows)
        obj.shape = tuple(int(v) for v in shape)
        return obj

    def _value_names(self):
        return ("x", "y", "z")

    def _host_values(self, a):
        return [a[i].reshape(self.shape) for i in range(3)]


class scalarfield(_DeviceBacked):
    """Complex scalar field (PyPOTypes.py:216-261)."""

    def __init__(self, S):
        self.S = np.asarray(S)
        self.shape = self.S.shape
        self.surf = None
        self.k = None

    @classmethod
    def on_device(cls, rows, shape):
        """scalarfield over (2, n) float64 device rows Re, Im."""
        obj = cls.__new__(cls)
        obj._init_device(rows)
        obj.shape = tuple(int(v) for v in shape)
        obj.surf = None
        obj.k = None
        return obj

    def _value_names(self):
        return ("S",)

    def _host_values(self, a):
        return [(a[0] + 1j * a[1]).reshape(self.shape)]

    def setMeta(self, surf, k):
        self.surf = surf
        self.k = k


class reflGrids:
    """Realised surface grid (PyPOTypes.py:263-308)."""

    def __init__(self, x, y, z, nx, ny, nz, area):
        self.x, self.y, self.z = x, y, z
        self.nx, self.ny, self.nz = nx, ny, nz
        self.area = area


class frame(_DeviceBacked):
    """Ray frame: positions and directions (PyPOTypes.py:310-374)."""

    def __init__(self, size, x, y, z, dx, dy, dz):
        self.size = int(size)
        self.x, self.y, self.z = (np.asarray(a, dtype=np.float64) for a in (x, y, z))
        self.dx, self.dy, self.dz = (np.asarray(a, dtype=np.float64) for a in (dx, dy, dz))
        self.snapshots = {}
        self.pos = None
        self.ori = None
        self.transf = None

    @classmethod
    def on_device(cls, rows):
        """Frame over (6, n) float64 device rows x, y, z, dx, dy, dz."""
        obj = cls.__new__(cls)
        obj._init_device(rows)
        obj.size = int(rows.shape[1])
        obj.snapshots = {}
        obj.pos = None
        obj.ori = None
        obj.transf = None
        return obj

    def _value_names(self):
        return ("x", "y", "z", "dx", "dy", "dz")

    def _host_values(self, a):
        return [a[i].copy() for i in range(6)]

    def setMeta(self, pos, ori, transf):
        self.pos = pos
        self.ori = ori
        self.transf = transf
