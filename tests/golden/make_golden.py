"""Write the golden vectors: run every case of cases.py through the reference's
own CPU library, oracle/_ref/libpypocpu_ref.so, compiled from the reference's
C++ sources (src/CPU/InterfaceCPU.cpp, src/common/InterfaceReflector.cpp,
src/common/InterfaceBeam.cpp) by `make -C oracle ref`.  Outputs: po.npz,
grids.npz, rt.npz, beams.npz in this directory.

usage: python tests/golden/make_golden.py          (po, grids, rt, beams)
       python tests/golden/make_golden.py beams    (beams.npz only)
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import cases as C  # noqa: E402
from oracle import bind as OB  # noqa: E402
from pypo_amd.geometry import gridsize  # noqa: E402


def flat(o):
    if isinstance(o, list):
        return np.concatenate([flat(x) for x in o])
    if hasattr(o, "components"):
        return np.stack([np.ravel(c) for c in o.components()])
    if hasattr(o, "S"):
        return np.ravel(o.S)[None, :]
    return np.stack([np.ravel(o.x), np.ravel(o.y), np.ravel(o.z)]).astype(complex)


def beams(ref):
    """beams.npz: every beam case through the compiled reference's
    src/common/InterfaceBeam.cpp, plus one (random) Gaussian ray frame."""
    out = {}
    for name, c in C.beam_cases().items():
        if c["kind"] in ("gauss", "vec"):
            f, cur = OB.make_gauss(ref, c["params"], c["source"], vector=c["kind"] == "vec")
            out[f"{name}/field"] = flat(f)
            out[f"{name}/current"] = flat(cur)
        elif c["kind"] == "scalar":
            out[f"{name}/field"] = np.ravel(OB.make_scalar_gauss(ref, c["params"], c["source"]).S)
        elif c["kind"] == "currents":
            fields = C.random_fields(gridsize(c["source"]), c["seed"])
            out[f"{name}/field"] = flat(fields)
            out[f"{name}/current"] = flat(OB.calc_currents(ref, fields, c["source"], c["mode"]))
        else:
            fr = OB.make_rt_frame(ref, c["params"])
            out[f"{name}/frame"] = np.stack([getattr(fr, a) for a in ("x", "y", "z", "dx", "dy", "dz")])
    fr = OB.make_grt_frame_ref(C.grt_case())
    out["grt/frame"] = np.stack([getattr(fr, a) for a in ("x", "y", "z", "dx", "dy", "dz")])
    np.savez_compressed(os.path.join(HERE, "beams.npz"), **out)
    print("beams.npz", os.path.getsize(os.path.join(HERE, "beams.npz")), "bytes")


def main():
    ref = OB.load("ref")
    if sys.argv[1:] == ["beams"]:
        beams(ref)
        return
    po = {}
    for name, c in C.po_cases().items():
        shape = gridsize(c["source"])
        src = (C.random_scalar(shape, c["seed"]) if c["mode"] == "scalar"
               else C.random_currents(shape, c["seed"]))
        k = 2 * np.pi / c["lam"]
        out, gs, gt = OB.propagate(ref, c["mode"], c["source"], c["target"], src, k, 3,
                                   c.get("epsilon", 1.0), c["exp"])
        po[f"{name}/out"] = flat(out)
        po[f"{name}/src"] = flat(src)
        po[f"{name}/tgt_grid"] = np.stack([np.ravel(getattr(gt, a)) for a in
                                           ("x", "y", "z", "nx", "ny", "nz", "area")])
    np.savez_compressed(os.path.join(HERE, "po.npz"), **po)

    grids = {}
    for name, refl in C.grid_cases().items():
        for tr, sp in ((True, True), (False, False)):
            g = OB.generate_grid(ref, refl, tr, sp)
            grids[f"{name}/{int(tr)}{int(sp)}"] = np.stack(
                [np.ravel(getattr(g, a)) for a in ("x", "y", "z", "nx", "ny", "nz", "area")])
    np.savez_compressed(os.path.join(HERE, "grids.npz"), **grids)

    rt = {}
    for name, c in C.rt_cases().items():
        fr = C.rt_frame(c)
        out = OB.trace(ref, c["target"], fr, 3, c["tol"], c["t0"])
        rt[f"{name}/out"] = np.stack([getattr(out, a) for a in ("x", "y", "z", "dx", "dy", "dz")])
    np.savez_compressed(os.path.join(HERE, "rt.npz"), **rt)
    for f in ("po.npz", "grids.npz", "rt.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
    beams(ref)


if __name__ == "__main__":
    main()
