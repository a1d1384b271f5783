"""System: the user-facing PO / ray-trace entry points on the GPU.

Mirrors the part of the reference's src/PyPO/System.py that drives this
path -- surface registration (addParabola 244-291, addHyperbola 293-353,
addEllipse 355-418, addPlane 420-442), runPO (1113-1193) with the runPODict
checks of Checks.check_runPODict (1445-1602), runRayTracer (1531-1568) with
Checks.check_runRTDict (966-1058), and the EHP Poynting frame
(_loadFramePoynt 3195-3215) -- with the same dictionary keys, defaults,
result containers and error classes.  Every propagation runs on the GPU
through libpypohip.so; there is no CPU fallback.  Extra runPODict key:
"precision" = "mixed" (default) or "fp64".

Also mirrored: beam and frame creation (createTubeFrame 1243-1266,
createGRTFrame 1268-1302, createGaussian 1373-1404, createGaussianBeam
1406-1445, createScalarGaussian 1447-1472, createPointSource /
createUniformSource and their scalar twins 2146-2311), calcRayLen
(1304-1372), runHybridPropagation (1570-1622) and interpFrame (1624-1692;
nearest neighbour on the GPU, scipy's other methods on the host); the frame and beam figures of merit
(calcRTcenter / calcRTtilt / calcSpotRMS 1694-1746, calcSpillover / calcTaper
/ calcXpol 1748-1826) as GPU reductions; findRTfocus (2632-2700) over GPU
traces; the grid-convergence loops autoConverge / convergeOnTarget
(2779-3080) over GPU scalar propagations; and the bookkeeping around them:
mergeBeams (1195-1241, summed on the device), interpBeam (2306-2353, scipy
on the host as there), removeFrame / removeField / removeCurrent (994-1040),
setOverride, setLoggingVerbosity and copyObj.
"""

import copy
import logging
import re

import numpy as np

from . import bind as B
from . import geometry as G
from . import types as T

PO_modelist = ["JM", "EH", "JMEH", "EHP", "FF", "scalar"]
log = logging.getLogger("pypo_amd")


class RunPOError(Exception):
    """A runPODict is invalid (Checks.py:262-270)."""


class RunRTError(Exception):
    """A runRTDict is invalid (Checks.py:248-256)."""


class HybridPropError(Exception):
    """A hybridDict is invalid (Checks.py:1604-1654)."""


class FieldNameError(Exception):
    """No PO field of that name in the system (Checks.py:85-110)."""


class CurrentNameError(Exception):
    """No PO current of that name in the system (Checks.py:112-137)."""


class FrameNameError(Exception):
    """No ray-trace frame of that name in the system (Checks.py:166-191)."""


class MergeBeamError(Exception):
    """Beams to merge are defined on different surfaces (Checks.py:1755-1777)."""


FIELD_COMPONENTS = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}


def _device_alias(d, what):
    """The runPODict / runRTDict "device" key (Checks.py:1566-1591, 1023-1047).
    The reference runs "CPU" on its threaded C++ library and falls back to it,
    with a warning, for an unknown name or a missing GPU library.  This
    library has one implementation, on the GPU: "CPU" is accepted as an alias
    so that scripts written for the reference run unchanged, an unknown name
    is warned about as there, and the key is then set to "GPU".  nThreads is
    accepted and ignored (the launch shape is the library's)."""
    dev = d.setdefault("device", "CPU")
    if dev not in ("CPU", "GPU", "HIP"):
        log.warning(f"Device {dev} unknown. Running {what} on the GPU.")
    elif dev == "CPU":
        log.info(f"device CPU: {what} runs on the GPU (this library has no CPU path).")
    d["device"] = "GPU"
    d.setdefault("nThreads", 256)


def _comp_index(comp):
    """None / "NONE" -> None; "Ex".."Hz", 0..5 or an enum with .value -> index
    (Enums.FieldComponents)."""
    if comp is None:
        return None
    if hasattr(comp, "value"):
        comp = comp.value
        if comp is None:
            return None
    if isinstance(comp, str):
        if comp.upper() == "NONE":
            return None
        return FIELD_COMPONENTS[comp]
    c = int(comp)
    if not 0 <= c <= 5:
        raise ValueError(f"field component {comp} out of range")
    return c


def calc_beam(d):
    """Fill w0 / z of a vecGPODict from any two of w0, z, w, R (Checks.calc_beam,
    Checks.py:1299-1350)."""
    lam = d["lam"]
    if "w0" in d:
        w0 = d["w0"]
        if "z" in d:
            return
        if "w" in d:
            d["z"] = np.pi * w0 / lam * np.sqrt(d["w"] ** 2 - w0 ** 2)
            return
        if "R" in d:
            R = d["R"]
            z = R / 2 * (1 + np.sqrt(1 - (2 * np.pi * w0 ** 2 / (lam * R)) ** 2))
            if np.isnan(z):
                z = R / 2 * (1 - np.sqrt(1 - (2 * np.pi * w0 ** 2 / (lam * R)) ** 2))
            d["z"] = z
            return
    elif "z" in d:
        z = d["z"]
        if "w" in d:
            w = d["w"]
            w0 = np.sqrt(w ** 2 / 2 * (1 - np.sqrt(1 - ((2 * lam * z) / (np.pi * w ** 2)) ** 2)))
            if np.isnan(w0):
                w0 = np.sqrt(w ** 2 / 2 * (1 + np.sqrt(1 - ((2 * lam * z) / (np.pi * w ** 2)) ** 2)))
            d["w0"] = w0
            return
        if "R" in d:
            d["w0"] = np.sqrt(lam / np.pi * np.sqrt(z * (d["R"] - z)))
            return
    elif "w" in d and "R" in d:
        w, R = d["w"], d["R"]
        d["w0"] = w / np.sqrt(1 + ((np.pi * w ** 2) / (lam * R)) ** 2)
        d["z"] = R / (1 + ((lam * R) / (np.pi * w ** 2)) ** 2)
        return
    raise ValueError("Need at least two of w0, w, z, R to define the Gaussian beam.")


def _get_index(name, names, override):
    """Next free numeric suffix for `name` (Checks.getIndex, Checks.py:20-39)."""
    if override:
        return 0
    pat = re.compile(rf"^{re.escape(name)}(_\d+)?$")
    nums = [0 if n == name else int(n[len(name) + 1:]) for n in names if pat.match(n)]
    return max(nums) + 1 if nums else 0


# Limit keys of each grid mode (0 xy, 1 uv, 2 AoE), scaled by autoConverge.
_LIMS = {0: ("lims_x", "lims_y"), 1: ("lims_u", "lims_v"), 2: ("lims_Az", "lims_El")}


class System:
    def __init__(self, override=True):
        self.system = {}
        self.currents = {}
        self.fields = {}
        self.scalarfields = {}
        self.frames = {}
        self.assoc = {}
        self.override = override
        self._grid_cache = None    # device grids of the surfaces, by their parameters

    def _device_grid(self, refl):
        """(7, n) device grid of a reflDict, generated once per parameter set."""
        from . import device as D
        if self._grid_cache is None:
            self._grid_cache = D.GridCache()
        return self._grid_cache.get(refl)

    # ---- surfaces ------------------------------------------------------------
    def _add(self, d, builder):
        if "name" not in d:
            raise ValueError("reflDict needs a name")
        self.system[d["name"]] = builder(d)

    def addParabola(self, reflDict):
        self._add(reflDict, G.paraboloid)

    def addHyperbola(self, reflDict):
        self._add(reflDict, G.hyperboloid)

    def addEllipse(self, reflDict):
        self._add(reflDict, G.ellipsoid)

    def addPlane(self, reflDict):
        self._add(reflDict, G.plane)

    def generateGrids(self, name, transform=True, spheric=True):
        """Realised grid of surface `name` (reflGrids), generated on the GPU."""
        return B.generate_grid(self.system[name], transform, spheric)

    # ---- sources -------------------------------------------------------------
    def addCurrents(self, name, cur, surf, k):
        """Register surface currents (J, M) defined on surface `surf`."""
        if surf not in self.system:
            raise RunPOError(f"{surf} is not a surface in the system")
        c = copy.copy(cur)
        c.setMeta(surf, k)
        self.currents[name] = c

    def addScalarField(self, name, field, surf, k):
        if surf not in self.system:
            raise RunPOError(f"{surf} is not a surface in the system")
        f = copy.copy(field)
        f.setMeta(surf, k)
        self.scalarfields[name] = f

    def addFrame(self, name, fr):
        self.frames[name] = fr

    # ---- beams and frames (beam.py, device kernels) -------------------------------
    def _name(self, d, default, pool):
        d.setdefault("name", default)
        num = _get_index(d["name"], pool.keys(), self.override)
        if num > 0:
            d["name"] = f"{d['name']}_{num}"
        return d["name"]

    def createTubeFrame(self, argDict):
        """Tube of rays (System.createTubeFrame, System.py:1243-1266)."""
        from . import beam as PB
        d = copy.copy(argDict)
        for key in ("nRays", "nRing"):
            d[key] = abs(int(d[key]))
        name = self._name(d, "TubeFrame", self.frames)
        fr = PB.makeRTframe(d)
        fr.setMeta(np.zeros(3), np.array([0.0, 0.0, 1.0]), np.eye(4))
        self.frames[name] = fr
        return fr

    def createGRTFrame(self, argDict):
        """Gaussian ray frame (System.createGRTFrame, System.py:1268-1302): the
        angular spread follows from lam / (pi n w0)."""
        from . import beam as PB
        d = copy.copy(argDict)
        d["nRays"] = abs(int(d["nRays"]))
        d.setdefault("n", 1.0)
        d.setdefault("seed", -1)
        d["angx0"] = np.degrees(d["lam"] / (np.pi * d["n"] * d["x0"]))
        d["angy0"] = np.degrees(d["lam"] / (np.pi * d["n"] * d["y0"]))
        name = self._name(d, "GaussFrame", self.frames)
        fr = PB.makeGRTframe(d)
        fr.setMeta(np.zeros(3), np.array([0.0, 0.0, 1.0]), np.eye(4))
        self.frames[name] = fr
        return fr

    def _surface(self, name):
        if name not in self.system:
            raise RunPOError(f"{name} is not a surface in the system")
        return self.system[name]

    def createGaussian(self, gaussDict, name_surface):
        """Astigmatic Gaussian E field and its currents (System.py:1373-1404)."""
        from . import beam as PB
        d = copy.copy(gaussDict)
        d.setdefault("n", 1.0)
        d.setdefault("dxyz", 0.0)
        d.setdefault("E0", 1.0)
        d.setdefault("pol", np.array([1.0, 0.0, 0.0]))
        name = self._name(d, "GaussianBeamPO", self.fields)
        f, c = PB.makeGauss(d, self._surface(name_surface))
        k = 2 * np.pi / d["lam"]
        f.setMeta(name_surface, k)
        c.setMeta(name_surface, k)
        self.fields[name], self.currents[name] = f, c
        return f, c

    def createGaussianBeam(self, gaussDict, name_surface):
        """Vector Gaussian by the complex source point (System.py:1406-1445)."""
        from . import beam as PB
        d = copy.copy(gaussDict)
        calc_beam(d)
        d.setdefault("n", 1.0)
        d.setdefault("power", 4 * np.pi)
        d.setdefault("mode", 0)
        name = self._name(d, "GaussianBeamPO", self.fields)
        f, c = PB.makeGaussBeam(d, self._surface(name_surface))
        k = 2 * np.pi / d["lam"]
        f.setMeta(name_surface, k)
        c.setMeta(name_surface, k)
        self.fields[name], self.currents[name] = f, c
        return f, c

    def createScalarGaussian(self, gaussDict, name_surface):
        """Scalar Gaussian (System.py:1447-1472)."""
        from . import beam as PB
        d = copy.copy(gaussDict)
        d.setdefault("n", 1.0)
        d.setdefault("dxyz", 0.0)
        d.setdefault("E0", 1.0)
        name = self._name(d, "GaussianBeamPO", self.scalarfields)
        s = PB.makeScalarGauss(d, self._surface(name_surface))
        s.setMeta(name_surface, 2 * np.pi / d["lam"])
        self.scalarfields[name] = s
        return s

    def _source_field(self, d, name_surface, point):
        shape = tuple(int(v) for v in self._surface(name_surface)["gridsize"])
        d.setdefault("E0", 1.0)
        d.setdefault("phase", 0.0)
        amp = d["E0"] * np.exp(1j * d["phase"])
        if point:
            a = np.zeros(shape, dtype=complex)
            a[int((shape[0] - 1) / 2), int((shape[1] - 1) / 2)] = amp
        else:
            a = np.ones(shape, dtype=complex) * amp
        return a

    def _vector_source(self, sDict, name_surface, point, default):
        from . import beam as PB
        d = copy.copy(sDict)
        d.setdefault("pol", np.array([1.0, 0.0, 0.0]))
        a = self._source_field(d, name_surface, point)
        z = a * 0
        f = T.fields(a * d["pol"][0], a * d["pol"][1], a * d["pol"][2], z, z, z)
        c = PB.calcCurrents(f, self.system[name_surface], "PMC")
        k = 2 * np.pi / d["lam"]
        name = self._name(d, default, self.fields)
        f.setMeta(name_surface, k)
        c.setMeta(name_surface, k)
        self.fields[name], self.currents[name] = f, c
        return f, c

    def createPointSource(self, PSDict, name_surface):
        """Point source at the surface grid's centre, PMC currents (System.py:2146-2194)."""
        return self._vector_source(PSDict, name_surface, True, "PointSourcePO")

    def createUniformSource(self, UDict, name_surface):
        """Uniform field on the surface, PMC currents (System.py:2196-2237)."""
        return self._vector_source(UDict, name_surface, False, "UniformSourcePO")

    def createPointSourceScalar(self, PSDict, name_surface):
        d = copy.copy(PSDict)
        a = self._source_field(d, name_surface, True)
        name = self._name(d, "PointSourcePO", self.scalarfields)
        s = T.scalarfield(a)
        s.setMeta(name_surface, 2 * np.pi / d["lam"])
        self.scalarfields[name] = s
        return s

    def createUniformSourceScalar(self, UDict, name_surface):
        d = copy.copy(UDict)
        a = self._source_field(d, name_surface, False)
        name = self._name(d, "UniformSourcePO", self.scalarfields)
        s = T.scalarfield(a)
        s.setMeta(name_surface, 2 * np.pi / d["lam"])
        self.scalarfields[name] = s
        return s

    def calcCurrents(self, name_source, fields, mode="PEC"):
        """Currents of a fields object on surface `name_source` (BindBeam.calcCurrents)."""
        from . import beam as PB
        c = PB.calcCurrents(fields, self._surface(name_source), mode)
        c.setMeta(name_source, fields.k)
        return c

    # ---- PO ----------------------------------------------------------------------
    def _check_runPODict(self, d):
        err = []
        d.setdefault("exp", "fwd")
        if d["exp"] not in ("fwd", "bwd"):
            err.append(f"{d['exp']} is not a valid propagation direction.")
        if "t_name" not in d:
            err.append("Please provide t_name in runPODict.")
        elif d["t_name"] not in self.system:
            err.append(f"Target element {d['t_name']} not in system.")
        if "mode" not in d:
            err.append("Please provide mode in runPODict.")
        elif d["mode"] not in PO_modelist:
            err.append(f"{d['mode']} is not a valid propagation mode.")
        else:
            mode = d["mode"]
            if mode == "scalar":
                if d.get("s_scalarfield") not in self.scalarfields:
                    err.append("Please provide a valid s_scalarfield in runPODict.")
            elif d.get("s_current") not in self.currents:
                err.append("Please provide a valid s_current in runPODict.")
            needs = {"JM": ["name_JM"], "EH": ["name_EH"], "FF": ["name_EH"],
                     "JMEH": ["name_JM", "name_EH"], "EHP": ["name_EH", "name_P"],
                     "scalar": ["name_field"]}[mode]
            pools = {"name_JM": self.currents, "name_EH": self.fields, "name_P": self.frames,
                     "name_field": self.scalarfields}
            for key in needs:
                if key not in d:
                    err.append(f"Please provide {key} in runPODict.")
                else:
                    num = _get_index(d[key], pools[key].keys(), self.override)
                    if num > 0:
                        d[key] = f"{d[key]}_{num}"
        d.setdefault("epsilon", 1.0)
        d.setdefault("precision", "mixed")
        if d["precision"] not in B.PRECISIONS:
            err.append(f"{d['precision']} is not a valid precision (mixed, fp64).")
        _device_alias(d, "PO")
        if err:
            for e in err:
                log.error(e)
            raise RunPOError("\n".join(err))

    def runPO(self, runPODict):
        """PO propagation (System.runPO, System.py:1113-1193); stores and returns the result.

        Runs on device-resident grids (cached per surface parameters) and keeps the
        result in device memory: its values are copied to the host only when read,
        and a following runPO, ray trace or figure of merit that takes it as input
        reads it where it lies."""
        d = copy.copy(runPODict)
        self._check_runPODict(d)
        mode = d["mode"]
        if mode == "scalar":
            src = self.scalarfields[d["s_scalarfield"]]
            d["s_scalarfield"] = src
        else:
            src = self.currents[d["s_current"]]
            d["s_current"] = src
        d["k"] = src.k
        source = self.system[src.surf]
        t, k = d["t_name"], d["k"]
        target = self.system[t]
        out = self._propagate(source, target, d)
        if mode == "JM":
            out.setMeta(t, k)
            self.currents[d["name_JM"]] = out
        elif mode in ("EH", "FF"):
            out.setMeta(t, k)
            self.fields[d["name_EH"]] = out
        elif mode == "JMEH":
            out[0].setMeta(t, k)
            out[1].setMeta(t, k)
            self.currents[d["name_JM"]] = out[0]
            self.fields[d["name_EH"]] = out[1]
        elif mode == "EHP":
            out[0].setMeta(t, k)
            self.fields[d["name_EH"]] = out[0]
            self.frames[d["name_P"]] = self._loadFramePoynt(out[1], target)
            self.assoc[t] = [d["name_EH"], d["name_P"]]
        else:
            out.setMeta(t, k)
            self.scalarfields[d["name_field"]] = out
        return out

    def _propagate(self, source, target, d):
        """One propagation on the GPU from a checked runPODict whose source entry
        holds the source object; returns what PyPO_CPUd returns (BindCPU.py:82-220:
        currents, fields, [currents, fields], [fields, rfield] or scalarfield),
        with the values left in device memory."""
        from . import device as D
        mode = d["mode"]
        vals = D.scalar_tensor(d["s_scalarfield"]) if mode == "scalar" else D.currents_tensor(d["s_current"])
        rows = D.propagate(mode, self._device_grid(source), self._device_grid(target), vals, source, d["k"],
                           epsilon=d["epsilon"], exp=d["exp"], precision=d["precision"])
        shape = tuple(int(v) for v in target["gridsize"])
        if mode == "JM":
            return T.currents.on_device(rows, shape)
        if mode in ("EH", "FF"):
            return T.fields.on_device(rows, shape)
        if mode == "JMEH":
            return [T.currents.on_device(rows[:12], shape), T.fields.on_device(rows[12:], shape)]
        if mode == "EHP":
            return [T.fields.on_device(rows[:12], shape), T.rfield.on_device(rows[12:15], shape)]
        return T.scalarfield.on_device(rows, shape)

    def _loadFramePoynt(self, poynting, target):
        """Frame of reflected Poynting vectors on their surface (System.py:3195-3215):
        the surface's grid points with the Poynting directions, kept on the device
        when the directions are there."""
        rows = poynting.device_rows()
        if rows is not None:
            import torch
            g = self._device_grid(target)
            return T.frame.on_device(torch.cat([g[0:3], rows]).contiguous())
        g = B.generate_grid(target)
        n = np.asarray(poynting.x).size
        return T.frame(n, g.x.ravel(), g.y.ravel(), g.z.ravel(), np.ravel(poynting.x), np.ravel(poynting.y),
                       np.ravel(poynting.z))

    # ---- ray tracing -------------------------------------------------------------
    def _check_runRTDict(self, d):
        err = []
        if "fr_in" not in d:
            err.append("Please provide fr_in in runRTDict.")
        elif d["fr_in"] not in self.frames:
            err.append(f"Frame {d['fr_in']} not in system.")
        if "t_name" not in d:
            err.append("Please provide t_name in runRTDict.")
        elif d["t_name"] not in self.system:
            err.append(f"Target element {d['t_name']} not in system.")
        if "fr_out" not in d:
            err.append("Please provide fr_out in runRTDict.")
        else:
            num = _get_index(d["fr_out"], self.frames.keys(), self.override)
            if num > 0:
                d["fr_out"] = f"{d['fr_out']}_{num}"
        tol = d.get("tol", 1e-3)
        if not isinstance(tol, (int, float)):
            tol = 1e-3
        if tol < 0:
            log.warning("Negative tolerances are not allowed. Changing sign.")
            tol = -tol
        d["tol"] = tol
        t0 = d.get("t0", 1)
        d["t0"] = t0 if isinstance(t0, (int, float)) else 1
        _device_alias(d, "ray tracing")
        if err:
            for e in err:
                log.error(e)
            raise RunRTError("\n".join(err))

    def runRayTracer(self, runRTDict):
        """Ray trace a frame onto a surface (System.runRayTracer, System.py:1531-1568).
        The traced frame stays in device memory until its rays are read."""
        from . import device as D
        d = copy.copy(runRTDict)
        self._check_runRTDict(d)
        name_out = d["fr_out"]
        rays = D.frame_tensor(self.frames[d["fr_in"]])
        fr = T.frame.on_device(D.trace([self.system[d["t_name"]]], rays, tol=d["tol"], t0=d["t0"])[0])
        fr.setMeta(self.calcRTcenter(fr), self.calcRTtilt(fr), np.eye(4))
        self.frames[name_out] = fr
        return fr

    def _frame(self, name_frame):
        if isinstance(name_frame, str):
            if name_frame not in self.frames:
                raise RunRTError(f"Frame {name_frame} not in system.")
            return self.frames[name_frame]
        return name_frame

    def calcRTcenter(self, name_frame):
        """Centroid of the rays that hit (System.calcRTcenter, System.py:1694-1710;
        Efficiencies.calcRTcenter 43-57), reduced on the GPU.  Takes a frame's
        name or a frame."""
        from . import efficiencies as E
        return E.calcRTcenter(self._frame(name_frame))

    def calcRTtilt(self, name_frame):
        """Normalised mean direction of the rays that hit (System.py:1712-1728)."""
        from . import efficiencies as E
        return E.calcRTtilt(self._frame(name_frame))

    def calcSpotRMS(self, name_frame):
        """RMS spot size of a frame (System.calcSpotRMS, System.py:1730-1746)."""
        from . import efficiencies as E
        return E.calcRMS(self._frame(name_frame))

    # ---- efficiencies (System.py:1748-1826, Efficiencies.py) ----------------------------
    def _field_comp(self, name_field, comp):
        if name_field not in self.fields:
            raise RunPOError(f"{name_field} is not a field in the system")
        f = self.fields[name_field]
        c = _comp_index(comp)
        if c is None:
            raise ValueError("a field component is required")
        rows = f.component_rows(c)     # (2, n) on the device while the field lives there
        return f, (f[c] if rows is None else rows)

    def calcSpillover(self, name_field, comp, aperDict):
        """Spillover efficiency of a field component on an aperture of the field's surface."""
        from . import efficiencies as E
        f, fc = self._field_comp(name_field, comp)
        return E.calcSpillover(fc, self.system[f.surf], aperDict)

    def calcTaper(self, name_field, comp, aperDict=None):
        """Taper efficiency of a field component (over the whole surface without aperDict)."""
        from . import efficiencies as E
        f, fc = self._field_comp(name_field, comp)
        return E.calcTaper(fc, self.system[f.surf], aperDict or None)

    def calcXpol(self, name_field, comp_co, comp_cr):
        """Cross-polar efficiency of two components of a field."""
        from . import efficiencies as E
        _, co = self._field_comp(name_field, comp_co)
        _, cr = self._field_comp(name_field, comp_cr)
        return E.calcXpol(co, cr)

    # ---- ray-trace focus (System.findRTfocus 2632-2700, _optimiseFocus 3217-3238) ------
    @staticmethod
    def findRotation(v, u):
        """4x4 rotation taking direction v onto u, by Rodrigues' formula about
        v x u (System.findRotation, System.py:2714-2745)."""
        v, u = np.asarray(v, dtype=float), np.asarray(u, dtype=float)
        out = np.eye(4)
        if np.array_equal(v, u):
            return out
        lv, lu = np.linalg.norm(v), np.linalg.norm(u)
        if lv == 0 or lu == 0:
            log.error("Encountered 0-length vector. Cannot proceed.")
            return None
        w = np.cross(v / lv, u / lu)
        if not np.any(w) and np.dot(v, u) > 0:
            return out   # parallel but not bit-equal: the reference's formula divides 0 by 0 here
        K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
        out[:3, :3] = np.eye(3) + K + K @ K * (1 - np.dot(v, u)) / np.linalg.norm(w) ** 2
        return out

    def findRTfocus(self, name_frame, f0=None, tol=1e-12):
        """Point of least RMS spot size along the mean direction of a frame.

        A 3x3-point plane focal_plane_<name> is placed through the frame's
        centre, perpendicular to its mean direction; Nelder-Mead (scipy fmin,
        ftol = tol) moves it along that direction, each trial tracing the
        frame onto it on the GPU and reducing the spot RMS there, so only one
        number per trial leaves the device.  The plane ends at the focus and
        the frame traced onto it is stored as focus_<name>.  Returns the
        focus."""
        from scipy.optimize import fmin

        from . import device as D
        from . import efficiencies as E
        fr = self._frame(name_frame)
        f0 = 0.0 if f0 is None else f0
        tilt, center = self.calcRTtilt(fr), self.calcRTcenter(fr)
        match = np.array([0.0, 0.0, 1.0])
        R = np.eye(4) if np.all(np.isclose(match, -tilt)) else self.findRotation(match, tilt)
        t_name, fr_out = f"focal_plane_{name_frame}", f"focus_{name_frame}"
        self.addPlane({"name": t_name, "gmode": "xy", "lims_x": np.array([-4.2, 4.2]),
                       "lims_y": np.array([-4.2, 4.2]), "gridsize": np.array([3, 3])})
        plane = self.system[t_name]
        rays = D.frame_tensor(fr)

        def place(p):
            T4 = R.copy()
            T4[:3, 3] = p
            plane["transf"] = T4

        def spot(f):
            place(center + float(np.ravel(f)[0]) * tilt)
            return E.calcRMS(D.trace([plane], rays, tol=tol)[0])

        res = fmin(spot, f0, full_output=True, disp=False, ftol=tol)
        out = res[0][0] * tilt + center
        place(out)
        self.frames[fr_out] = D.tensor_to_frame(D.trace([plane], rays, tol=tol)[0])
        log.info(f"Focus of frame {name_frame}: {out}, RMS: {res[1]:.3e}")
        return out

    # ---- hybrid ray-trace / PO ------------------------------------------------------
    def calcRayLen(self, *frames, start=None):
        """Ray lengths between consecutive frames, plus their sum (System.calcRayLen,
        System.py:1304-1372); with `start`, the first entry is the distance from
        that point to the first frame."""
        for f in frames:
            if f not in self.frames:
                raise RunRTError(f"Frame {f} not in system.")
        out = []
        fr0 = self.frames[frames[0]]
        total = np.zeros(len(fr0.x))
        if isinstance(start, np.ndarray):
            lens = np.sqrt((fr0.x - start[0]) ** 2 + (fr0.y - start[1]) ** 2 + (fr0.z - start[2]) ** 2)
            out.append(lens)
            total = total + lens
        for name in frames[1:]:
            fr = self.frames[name]
            lens = np.sqrt((fr.x - fr0.x) ** 2 + (fr.y - fr0.y) ** 2 + (fr.z - fr0.z) ** 2)
            out.append(lens)
            total = total + lens
            fr0 = fr
        out.append(total)
        return out

    def interpFrame(self, name_fr_in, name_field, name_target, name_out, comp=None, method="nearest"):
        """Interpolate a field carried by a frame onto surface `name_target`'s grid
        (System.interpFrame, System.py:1624-1692).  "nearest" runs on the GPU
        (a nearest-neighbour search with scipy griddata's result); any other
        method is handed, as in the reference, to scipy.interpolate.griddata
        on the host with the frame's points and the grid generated on the
        device (griddata supports "linear" on these 3-D points and raises for
        "cubic", as it does for the reference)."""
        return self._interp_frame(name_fr_in, name_field, name_target, name_out, comp, method, True)

    def _interp_frame(self, name_fr_in, name_field, name_target, name_out, comp, method, host_copy):
        """interpFrame; with host_copy False the one-component form skips
        copying that component to the host for the return value (the hybrid
        step discards it) and returns None."""
        from . import device as D
        if name_fr_in not in self.frames:
            raise RunRTError(f"Frame {name_fr_in} not in system.")
        field = self.fields[name_field]
        grid = self._device_grid(self._surface(name_target))
        shape = tuple(int(v) for v in self.system[name_target]["gridsize"])
        fr = self.frames[name_fr_in]
        ci = _comp_index(comp)
        if method == "nearest":
            out = D.interp_nearest(D.frame_tensor(fr), D.currents_tensor(field), grid)
            if ci is None:
                res = T.fields.on_device(out, shape)
                res.setMeta(name_target, field.k)
                self.fields[name_out] = res
                return res
            # the other five components are zero (System._compToFields); the
            # result stays on the device and only the one component is copied
            import torch
            rows = [6 * (ci // 3) + ci % 3, 6 * (ci // 3) + 3 + ci % 3]
            kept = torch.zeros_like(out)
            kept[rows] = out[rows]
            res = T.fields.on_device(kept, shape)
            res.setMeta(name_target, field.k)
            self.fields[name_out] = res
            if not host_copy:
                return None
            a = out[rows].cpu().numpy()
            return (a[0] + 1j * a[1]).reshape(shape)
        else:
            from scipy.interpolate import griddata
            g = grid.cpu().numpy()
            points = (np.ravel(fr.x), np.ravel(fr.y), np.ravel(fr.z))
            comps = [None] * 6
            for i, c in enumerate(field.components()):
                if ci is not None and i != ci:
                    continue
                c = np.ravel(c)
                re = griddata(points, np.real(c), (g[0], g[1], g[2]), method=method)
                im = griddata(points, np.imag(c), (g[0], g[1], g[2]), method=method)
                comps[i] = re.reshape(shape) + 1j * im.reshape(shape)
        if ci is None:
            res = T.fields(*comps)
            ret = res
        else:
            null = np.zeros(shape, dtype=complex)
            res = T.fields(*[comps[i] if i == ci else null for i in range(6)])
            ret = comps[ci]
        res.setMeta(name_target, field.k)
        self.fields[name_out] = res
        return ret

    def _check_hybridDict(self, d):
        try:
            self._check_runRTDict(d)
        except RunRTError as e:
            raise HybridPropError(str(e))
        err = []
        if d.get("field_in") not in self.fields:
            err.append("Please provide a valid field_in in hybridDict.")
        if "field_out" not in d:
            err.append("Please provide field_out in hybridDict.")
        else:
            num = _get_index(d["field_out"], self.fields.keys(), self.override)
            if num > 0:
                d["field_out"] = f"{d['field_out']}_{num}"
        start = d.setdefault("start", None)
        if start is not None:
            start = np.asarray(start, dtype=np.float64)
            if start.shape != (3,):
                err.append("start must be a length-3 array.")
            d["start"] = start
        d.setdefault("interp", True)
        if not isinstance(d["interp"], bool):
            err.append("interp must be a bool.")
        d.setdefault("comp", None)
        if not err and d["field_in"] in self.fields:
            surf = self.fields[d["field_in"]].surf
            assoc = self.assoc.get(surf)
            if assoc is not None and (assoc[0] != d["field_in"] or assoc[1] != d["fr_in"]):
                err.append(f"{d['field_in']} and {d['fr_in']} are not associated on {surf}.")
        if err:
            for e in err:
                log.error(e)
            raise HybridPropError("\n".join(err))

    def runHybridPropagation(self, hybridDict):
        """Hybrid RT/PO step (System.runHybridPropagation, System.py:1570-1622): trace
        fr_in onto t_name, carry field_in along the rays with exp(j k L) (and the
        start-point amplitude), optionally interpolate onto t_name's grid."""
        from . import device as D
        d = copy.copy(hybridDict)
        self._check_hybridDict(d)
        field = self.fields[d["field_in"]]
        fr_t, f_t = D.hybrid(self.system[d["t_name"]], D.frame_tensor(self.frames[d["fr_in"]]),
                             D.currents_tensor(field), field.k, start=d["start"], tol=d["tol"], t0=d["t0"])
        fr_out = T.frame.on_device(fr_t)
        fr_out.setMeta(self.calcRTcenter(fr_out), self.calcRTtilt(fr_out), np.eye(4))
        self.frames[d["fr_out"]] = fr_out
        carried = T.fields.on_device(f_t, field.shape)
        carried.setMeta(d["t_name"], field.k)
        self.fields[d["field_out"]] = carried
        if d["interp"]:
            self._interp_frame(d["fr_out"], d["field_out"], d["t_name"], d["field_out"], d["comp"], "nearest", False)
        self.assoc[d["t_name"]] = [d["field_out"], d["fr_out"]]
        return self.fields[d["field_out"]]

    # ---- grid-convergence loops (System.py:2779-3080) ------------------------------
    def copyElement(self, name, name_copy):
        self.system[name_copy] = copy.deepcopy(self.system[name])
        self.system[name_copy]["name"] = name_copy

    def removeElement(self, name):
        self.system.pop(name, None)

    def removeScalarField(self, name):
        self.scalarfields.pop(name, None)

    def removeFrame(self, frameName):
        """System.removeFrame (System.py:994-1008)."""
        if frameName not in self.frames:
            raise FrameNameError(f"Frame {frameName} not in system.")
        del self.frames[frameName]

    def removeField(self, fieldName):
        """System.removeField (System.py:1010-1024)."""
        if fieldName not in self.fields:
            raise FieldNameError(f"Field {fieldName} not in system.")
        del self.fields[fieldName]

    def removeCurrent(self, currentName):
        """System.removeCurrent (System.py:1026-1040)."""
        if currentName not in self.currents:
            raise CurrentNameError(f"Current {currentName} not in system.")
        del self.currents[currentName]

    def setOverride(self, override=True):
        """Overwrite results of an existing name (True) or append _n (False)
        (System.setOverride, System.py:231-242)."""
        self.override = bool(override)

    def setLoggingVerbosity(self, verbose=True):
        """Log at DEBUG (True) or only CRITICAL (False) (System.py:213-229)."""
        self.verbosity = verbose
        log.setLevel(logging.DEBUG if verbose else logging.CRITICAL)

    def copyObj(self, obj=None):
        """Deep copy of obj, or of the system (System.copyObj, System.py:2702-2712);
        a result held in device memory is copied in device memory."""
        return copy.deepcopy(self if obj is None else obj)

    def _beam_pool(self, obj):
        """"fields" / "currents" pool for an Enums.Objects value or its string."""
        key = getattr(obj, "value", obj)
        if key not in ("fields", "currents"):
            raise ValueError(f"obj must be fields or currents, not {key}")
        return key, getattr(self, key)

    def mergeBeams(self, *beams, obj="fields", merged_name="combined"):
        """Sum of PO fields or currents defined on one surface, stored under
        merged_name (System.mergeBeams, System.py:1195-1241; obj is
        Enums.Objects.FIELD / CURRENT or "fields" / "currents").  The sum runs
        on the device and the result stays there."""
        key, pool = self._beam_pool(obj)
        missing = [b for b in beams if b not in pool]
        if missing:
            err = FieldNameError if key == "fields" else CurrentNameError
            raise err(f"{missing[0]} not in system {key}.")
        surf0 = pool[beams[0]].surf
        other = [b for b in beams if pool[b].surf != surf0]
        if other:
            raise MergeBeamError(f"{other[0]} is defined on {pool[other[0]].surf}, not on {surf0}.")
        from . import device as D
        ex = pool[beams[0]]
        total = D.currents_tensor(ex).clone()
        for b in beams[1:]:
            total += D.currents_tensor(pool[b])
        cls = T.fields if key == "fields" else T.currents
        res = cls.on_device(total, ex.shape)
        res.setMeta(ex.surf, ex.k)
        pool[merged_name] = res

    def interpBeam(self, name, gridsize_new, obj="fields"):
        """Linear interpolation of a PO field or current on a planar surface onto
        a copy of that surface with gridsize_new, stored as name + "_interp" on
        the surface surf + "_interp" (System.interpBeam, System.py:2306-2353).
        As in the reference, scipy's griddata interpolates real and imaginary
        parts over the grids' (x, y); the grids are generated on the GPU."""
        from scipy.interpolate import griddata
        key, pool = self._beam_pool(obj)
        if name not in pool:
            err = FieldNameError if key == "fields" else CurrentNameError
            raise err(f"{name} not in system {key}.")
        src = pool[name]
        surf_new = src.surf + "_interp"
        self.copyElement(src.surf, surf_new)
        self.system[surf_new]["gridsize"] = np.asarray(gridsize_new)
        g0, g1 = self.generateGrids(src.surf), self.generateGrids(surf_new)
        points = (np.ravel(g0.x), np.ravel(g0.y))
        points_new = (np.ravel(g1.x), np.ravel(g1.y))
        shape = tuple(int(v) for v in np.asarray(gridsize_new))
        comps = []
        for c in src.components():
            c = np.ravel(c)
            re = griddata(points, np.real(c), points_new)
            im = griddata(points, np.imag(c), points_new)
            comps.append((re + 1j * im).reshape(shape))
        res = (T.fields if key == "fields" else T.currents)(*comps)
        res.setMeta(surf_new, src.k)
        pool[name + "_interp"] = res

    def _strongest_component(self, source_field):
        """The source's component with the largest magnitude, as a scalarfield
        registered under "_" + source_field (System.py:2799-2803, 2924-2942)."""
        if source_field in self.fields:
            obj = self.fields[source_field]
        elif source_field in self.currents:
            obj = self.currents[source_field]
        elif source_field in self.scalarfields:
            sf = self.scalarfields[source_field]
            self.scalarfields["_" + source_field] = sf
            return
        else:
            raise RunPOError(f"{source_field} is not a field, current or scalar field in the system")
        comps = list(obj.components())
        best = int(np.argmax([np.max(np.abs(c)) for c in comps]))
        sf = T.scalarfield(comps[best])
        sf.setMeta(obj.surf, obj.k)
        self.scalarfields["_" + source_field] = sf

    def autoConverge(self, source_field, name_target, tol=1e-2, add=10, patch_size=1 / 9, max_iter=1000):
        """Grid size of `name_target` at which the scalar power collected on a
        central patch stops changing by more than tol (System.autoConverge,
        System.py:2779-2880).  The patch is the target's limits times patch_size;
        its grid grows by `add` per step; the returned size is scaled back to the
        full surface and set on it.  Every propagation is a GPU runPO."""
        self._strongest_component(source_field)
        tgt = self.system[name_target]
        xu, yv = _LIMS[int(tgt["gmode"])]
        orig = (tgt[xu], tgt[yv])
        tgt[xu] = patch_size * np.asarray(orig[0], dtype=float)
        tgt[yv] = patch_size * np.asarray(orig[1], dtype=float)
        gridsize = np.array([1, 1])
        P0, diff, norm, n = 1e99, 1e99, 1.0, 0
        d = {"t_name": name_target, "s_scalarfield": "_" + source_field, "epsilon": 10, "exp": "fwd",
             "mode": "scalar", "name_field": "_S_conv"}
        override, self.override = self.override, True
        try:
            while np.abs(diff) > tol:
                gridsize = gridsize + add
                tgt["gridsize"] = gridsize
                S = self.runPO(dict(d)).S
                if n == 0:
                    norm = np.max(np.abs(S))
                grid = self.generateGrids(name_target)
                P = np.sum(np.abs(S / norm) ** 2 * grid.area)
                diff = np.abs(P0 - P)
                P0 = P
                self.removeScalarField("_S_conv")
                log.info(f"Difference : {diff} at gridsize {gridsize}")
                n += 1
                if n >= max_iter:
                    log.error("Could not find converged solution.")
                    break
        finally:
            self.override = override
            tgt[xu], tgt[yv] = orig
        gridsize = (gridsize / patch_size).astype(int)
        tgt["gridsize"] = gridsize
        return gridsize

    def convergeOnTarget(self, source_field, name_scatterer, name_target, tol=1e-2, target_gridsize=None,
                         mult=2, div=3, max_iter=16):
        """Grid size of `name_scatterer` at which the scalar field it scatters onto
        a few test points of `name_target` converges (System.convergeOnTarget,
        System.py:2882-3080): per axis, multiply the size by `mult` until two
        steps differ by less than tol (max relative to the peak), then bisect
        `div` times between the last unconverged and first converged sizes."""
        self._strongest_component(source_field)
        self.copyElement(name_target, "_conv_target")
        gmode = int(self.system["_conv_target"]["gmode"])
        if target_gridsize is not None:
            self.system["_conv_target"]["gridsize"] = np.asarray(target_gridsize)
        else:
            self.system["_conv_target"]["gridsize"] = np.array((3, 6)) if gmode == 1 else np.array((5, 5))
        self.copyElement(name_scatterer, "_conv_scatterer")
        scat = self.system["_conv_scatterer"]
        scat["gridsize"] = np.array((10, 30)) if gmode == 1 else np.array((10, 10))
        d1 = {"t_name": "_conv_scatterer", "s_scalarfield": "_" + source_field, "mode": "scalar",
              "name_field": "_S_conv_scat"}
        d2 = {"t_name": "_conv_target", "s_scalarfield": "_S_conv_scat", "mode": "scalar",
              "name_field": "_S_conv_target"}

        def target_field():
            self.runPO(dict(d1))
            return self.runPO(dict(d2)).S.copy()

        override, self.override = self.override, True
        conv_grid = np.array([10, 10])
        try:
            init_field = target_field()
            for i in range(2):
                field, n, diff = init_field.copy(), 0, 1e99
                scat["gridsize"] = np.array([10, 10])
                gridsize = 10
                while np.abs(diff) > tol:
                    last_field = field.copy()
                    gridsize = int(scat["gridsize"][i] * mult)
                    scat["gridsize"][i] = gridsize
                    field = target_field()
                    diff = np.max(np.abs(np.abs(field - last_field))) / np.max(np.abs(field))
                    log.info(f"Axis {i} Step {n + 1}: Difference {diff:.3e} at gridsize[{gridsize}]")
                    n += 1
                    if n >= max_iter:
                        break
                if n >= max_iter:
                    log.error(f"Could not find converged solution for axis {i}; returning {gridsize}")
                    conv_grid[i] = gridsize
                    break
                conv, nonconv = gridsize, gridsize / mult
                conv_field = last_field.copy()
                for j in range(div):
                    nxt = int((conv + nonconv) / 2)
                    scat["gridsize"][i] = nxt
                    field = target_field()
                    diff = np.max(np.abs(np.abs(field - conv_field))) / np.max(np.abs(field))
                    log.info(f"Axis {i} Step {n + j + 1}: Difference {diff:.3e} at gridsize[{nxt}]")
                    if diff < tol:
                        conv, conv_field = nxt, field.copy()
                    else:
                        nonconv = nxt
                conv_grid[i] = conv
        finally:
            self.override = override
            for name in ("_S_conv_scat", "_S_conv_target"):
                self.removeScalarField(name)
            self.removeElement("_conv_target")
            self.removeElement("_conv_scatterer")
        self.system[name_scatterer]["gridsize"] = conv_grid
        return conv_grid
