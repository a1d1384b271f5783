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


def _get_index(name, names, override):
    """Next free numeric suffix for `name` (Checks.getIndex, Checks.py:20-39)."""
    if override:
        return 0
    pat = re.compile(rf"^{re.escape(name)}(_\d+)?$")
    nums = [0 if n == name else int(n[len(name) + 1:]) for n in names if pat.match(n)]
    return max(nums) + 1 if nums else 0


class System:
    def __init__(self, override=True):
        self.system = {}
        self.currents = {}
        self.fields = {}
        self.scalarfields = {}
        self.frames = {}
        self.assoc = {}
        self.override = override

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
        dev = d.setdefault("device", "GPU")
        if dev not in ("GPU", "HIP"):
            err.append(f"device {dev}: this library runs PO on the GPU only (GPU or HIP).")
        d.setdefault("nThreads", 256)
        if err:
            for e in err:
                log.error(e)
            raise RunPOError("\n".join(err))

    def runPO(self, runPODict):
        """PO propagation (System.runPO, System.py:1113-1193); stores and returns the result."""
        d = copy.copy(runPODict)
        self._check_runPODict(d)
        if d["mode"] == "scalar":
            src = self.scalarfields[d["s_scalarfield"]]
            d["s_scalarfield"] = src
        else:
            src = self.currents[d["s_current"]]
            d["s_current"] = src
        d["k"] = src.k
        source = self.system[src.surf]
        target = self.system[d["t_name"]]
        out = B.PyPO_HIPd(source, target, d)
        t, k = d["t_name"], d["k"]
        mode = d["mode"]
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
            self.frames[d["name_P"]] = self._loadFramePoynt(out[1], t)
            self.assoc[t] = [d["name_EH"], d["name_P"]]
        else:
            out.setMeta(t, k)
            self.scalarfields[d["name_field"]] = out
        return out

    def _loadFramePoynt(self, poynting, name_source):
        """Frame of reflected Poynting vectors on their surface (System.py:3195-3215)."""
        g = self.generateGrids(name_source)
        n = poynting.x.size
        return T.frame(n, g.x.ravel(), g.y.ravel(), g.z.ravel(), poynting.x.ravel(),
                       poynting.y.ravel(), poynting.z.ravel())

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
        dev = d.setdefault("device", "GPU")
        if dev not in ("GPU", "HIP"):
            err.append(f"device {dev}: this library traces rays on the GPU only (GPU or HIP).")
        d.setdefault("nThreads", 256)
        if err:
            for e in err:
                log.error(e)
            raise RunRTError("\n".join(err))

    def runRayTracer(self, runRTDict):
        """Ray trace a frame onto a surface (System.runRayTracer, System.py:1531-1568)."""
        d = copy.copy(runRTDict)
        self._check_runRTDict(d)
        name_out = d["fr_out"]
        d["fr_in"] = self.frames[d["fr_in"]]
        d["t_name"] = self.system[d["t_name"]]
        fr = B.RT_HIPd(d)
        fr.setMeta(self.calcRTcenter(fr), self.calcRTtilt(fr), np.eye(4))
        self.frames[name_out] = fr
        return fr

    @staticmethod
    def calcRTcenter(fr):
        """Centroid of the rays that hit (Efficiencies.calcRTcenter 43-57)."""
        good = (fr.dx ** 2 + fr.dy ** 2 + fr.dz ** 2) > 0.8
        if not np.any(good):
            return np.full(3, np.nan)
        return np.array([fr.x[good].mean(), fr.y[good].mean(), fr.z[good].mean()])

    @staticmethod
    def calcRTtilt(fr):
        """Mean direction of the rays that hit (Efficiencies.calcRTtilt 59-73)."""
        good = (fr.dx ** 2 + fr.dy ** 2 + fr.dz ** 2) > 0.8
        if not np.any(good):
            return np.full(3, np.nan)
        t = np.array([fr.dx[good].mean(), fr.dy[good].mean(), fr.dz[good].mean()])
        return t / np.linalg.norm(t)
