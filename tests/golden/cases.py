"""Input definitions of the golden PO / grid / ray-trace cases.

Pure data: surface dictionaries, wavelengths and seeds.  make_golden.py runs
each case through the reference's own CPU library (compiled from its C++
sources by oracle/Makefile) and stores inputs and outputs in po.npz,
grids.npz and rt.npz next to this file; the tests replay the same cases.
"""

import numpy as np

from pypo_amd import geometry as G


def _feed_plane(gmode="xy", n=(13, 12), z=6.0, rms=0.0):
    if gmode == "xy":
        d = {"gmode": "xy", "lims_x": [-1.5, 1.5], "lims_y": [-1.2, 1.4], "gridsize": list(n)}
    else:
        d = {"gmode": "uv", "lims_u": [0, 1.5], "lims_v": [0, 360], "gridsize": list(n)}
    d.update({"flip": True, "rms": rms, "rms_seed": 11})
    p = G.plane(d)
    p["transf"] = G.mat_translate([0.1, -0.2, z])
    return p


def _dish(n=(11, 10), flip=False, gmode="uv"):
    if gmode == "uv":
        d = {"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 6.0], "gmode": "uv",
             "lims_u": [0.2, 5.0], "lims_v": [0, 360], "gridsize": list(n), "flip": flip,
             "ecc_uv": 0.3, "rot_uv": 20.0}
    else:
        d = {"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 6.0], "gmode": "xy",
             "lims_x": [-4, 4], "lims_y": [-3, 5], "gridsize": list(n), "flip": flip,
             "gcenter": [0.5, -0.5]}
    return G.paraboloid(d)


def _ff(n=(9, 8), lim=8.0):
    return G.plane({"gmode": "AoE", "lims_Az": [-lim, lim], "lims_El": [-lim, 0.6 * lim],
                    "gridsize": list(n)})


def po_cases():
    """name -> dict(mode, exp, source, target, lam, seed)."""
    cases = {}
    for exp in ("fwd", "bwd"):
        for mode in ("JM", "EH", "JMEH", "EHP", "scalar"):
            cases[f"{mode}_xy_{exp}"] = dict(mode=mode, exp=exp, source=_feed_plane("xy"),
                                             target=_dish(), lam=0.7, seed=1)
        cases[f"JM_uvsrc_{exp}"] = dict(mode="JM", exp=exp, source=_feed_plane("uv", (9, 14)),
                                        target=_dish((12, 9), gmode="xy"), lam=0.9, seed=2)
        cases[f"FF_uv_{exp}"] = dict(mode="FF", exp=exp, source=_dish((15, 16)), target=_ff(),
                                     lam=0.5, seed=3)
        cases[f"FF_xy_{exp}"] = dict(mode="FF", exp=exp, source=_dish((13, 14), gmode="xy"),
                                     target=_ff((7, 11), 20.0), lam=0.8, seed=4)
    # shadowing: the dish faces away from half of the far-field directions
    cases["FF_shadow_fwd"] = dict(mode="FF", exp="fwd", source=_dish((13, 12), gmode="xy"),
                                  target=_ff((9, 9), 80.0), lam=0.8, seed=5)
    # near field with rough (rms > 0) source and target grids
    cases["EH_rough_fwd"] = dict(mode="EH", exp="fwd", source=_feed_plane("xy", rms=0.01),
                                 target=dict(_dish(), rms=0.02, rms_seed=3), lam=0.7, seed=6)
    # relative permittivity only enters EHP (Propagation.h:200-203, 892)
    cases["EHP_eps_fwd"] = dict(mode="EHP", exp="fwd", source=_feed_plane("xy"), target=_dish(),
                                lam=0.7, seed=7, epsilon=2.5)
    return cases


def grid_cases():
    cases = {
        "par_xy": _dish((7, 9), gmode="xy"),
        "par_uv_flip": _dish((8, 7), flip=True),
        "hyp_xy": G.hyperboloid({"pmode": "focus", "focus_1": [0, 0, 30], "focus_2": [0, 0, -10],
                                 "ecc": 1.4, "gmode": "xy", "lims_x": [-3, 3], "lims_y": [-2, 4],
                                 "gridsize": [7, 6], "gcenter": [0.3, 0.1]}),
        "hyp_uv": G.hyperboloid({"pmode": "focus", "focus_1": [0, 0, 30], "focus_2": [5, 0, -10],
                                 "ecc": 1.4, "gmode": "uv", "lims_u": [0, 3], "lims_v": [0, 360],
                                 "gridsize": [6, 8], "ecc_uv": 0.5, "rot_uv": 33.0}),
        "ell_xy": G.ellipsoid({"pmode": "focus", "focus_1": [0, 0, 30], "focus_2": [0, 0, -10],
                               "ecc": 0.6, "orient": "z", "gmode": "xy", "lims_x": [-3, 3],
                               "lims_y": [-2, 4], "gridsize": [7, 6], "gcenter": [9.0, 9.0]}),
        "ell_uv_x": G.ellipsoid({"pmode": "focus", "focus_1": [10, 0, 0], "focus_2": [-10, 0, 0],
                                 "ecc": 0.5, "orient": "x", "gmode": "uv", "lims_u": [0.5, 4],
                                 "lims_v": [10, 300], "gridsize": [6, 7]}),
        "plane_xy": _feed_plane("xy", (5, 6)),
        "plane_uv": _feed_plane("uv", (6, 5)),
        "plane_aoe": _ff((6, 7), 3.0),
        "par_uv_rough": dict(_dish((9, 8)), rms=0.05, rms_seed=1234),
        "plane_xy_rough": _feed_plane("xy", (6, 6), rms=0.2),
    }
    return cases


def rt_cases():
    """name -> dict(target, n_rays, tol, t0, seed, start_z, spread)."""
    par = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, 50.0],
                        "gmode": "xy", "lims_x": [-1, 1], "lims_y": [-1, 1], "gridsize": [3, 3]})
    par_tilt = dict(par)
    par_tilt["transf"] = G.mat_translate([3.0, -2.0, 1.0], G.find_rotation([0, 0, 1.0], [0.1, 0.05, 1.0]))
    hyp = G.hyperboloid({"pmode": "focus", "focus_1": [0, 0, 50.0], "focus_2": [0, 0, -20.0],
                         "ecc": 1.2, "gmode": "xy", "lims_x": [-1, 1], "lims_y": [-1, 1],
                         "gridsize": [3, 3]})
    ell = G.ellipsoid({"pmode": "focus", "focus_1": [0, 0, 40.0], "focus_2": [0, 0, -40.0],
                       "ecc": 0.5, "orient": "z", "gmode": "xy", "lims_x": [-1, 1],
                       "lims_y": [-1, 1], "gridsize": [3, 3]})
    pla = G.plane({"gmode": "xy", "lims_x": [-1, 1], "lims_y": [-1, 1], "gridsize": [3, 3],
                   "flip": True})
    pla["transf"] = G.mat_translate([0, 0, -5.0], G.find_rotation([0, 0, 1.0], [0.2, 0, 1.0]))
    return {
        "par": dict(target=par, tol=1e-12, t0=10.0, seed=1, start_z=80.0),
        "par_tilt_flip": dict(target=dict(par_tilt, flip=True), tol=1e-10, t0=1.0, seed=2, start_z=80.0),
        "hyp": dict(target=hyp, tol=1e-12, t0=5.0, seed=3, start_z=-5.0, up=True),
        "ell": dict(target=ell, tol=1e-12, t0=5.0, seed=4, start_z=0.0, up=True),
        "plane": dict(target=pla, tol=1e-12, t0=1.0, seed=5, start_z=20.0),
        "par_rough": dict(target=dict(par, rms=0.01, rms_seed=77), tol=1e-12, t0=10.0, seed=6,
                          start_z=80.0),
    }


def rt_frame(case, n=257):
    """Seeded input frame of a ray-trace case (plus two dead rays, d = 0)."""
    from pypo_amd import types as T
    rng = np.random.default_rng(case["seed"])
    x = rng.uniform(-4, 4, n)
    y = rng.uniform(-4, 4, n)
    z = np.full(n, case["start_z"])
    dz = np.ones(n) if case.get("up") else -np.ones(n)
    d = np.stack([rng.normal(0, 0.02, n), rng.normal(0, 0.02, n), dz])
    d /= np.linalg.norm(d, axis=0)
    d[:, :2] = 0.0   # rays that missed an earlier surface
    return T.frame(n, x, y, z, d[0], d[1], d[2])


def random_currents(shape, seed):
    from pypo_amd import types as T
    rng = np.random.default_rng(seed)
    comps = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(6)]
    return T.currents(*comps)


def random_scalar(shape, seed):
    from pypo_amd import types as T
    rng = np.random.default_rng(seed)
    return T.scalarfield(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def _beam_plane(n=(15, 13), z=0.0, tilt=False, half=(4.0, 3.0)):
    p = G.plane({"gmode": "xy", "lims_x": [-half[0], half[0]], "lims_y": [-half[1], half[1]],
                 "gridsize": list(n)})
    m = np.eye(4)
    if tilt:
        m = G.find_rotation([0, 0, 1], [0.2, -0.1, 1.0])
    p["transf"] = G.mat_translate([0.3, -0.2, z], m)
    return p


def beam_cases():
    """name -> dict(kind, params, source[, mode, seed]) for the beam-initialisation calls."""
    cases = {
        "gauss_waist": {"kind": "gauss", "source": _beam_plane(),
                        "params": {"lam": 1.0, "w0x": 2.0, "w0y": 2.0, "n": 1.0, "E0": 1.0, "dxyz": 0.0,
                                   "pol": np.array([1.0, 0.0, 0.0])}},
        "gauss_astig": {"kind": "gauss", "source": _beam_plane(z=7.0, tilt=True),
                        "params": {"lam": 0.8, "w0x": 1.5, "w0y": 2.5, "n": 1.3, "E0": 2.0, "dxyz": 5.0,
                                   "pol": np.array([0.6, 0.8, 0.0])}},
        "scalar_astig": {"kind": "scalar", "source": _beam_plane(z=-4.0, tilt=True),
                         "params": {"lam": 1.2, "w0x": 1.5, "w0y": 2.0, "n": 1.1, "E0": 3.0, "dxyz": 2.0}},
        "scalar_waist": {"kind": "scalar", "source": _beam_plane(),
                         "params": {"lam": 1.0, "w0x": 2.0, "w0y": 1.0, "n": 1.0, "E0": 1.0, "dxyz": 0.0}},
    }
    for mode in (0, 1, 2):
        cases[f"vec_gauss_m{mode}"] = {
            "kind": "vec", "source": _beam_plane(z=0.0),
            "params": {"lam": 1.0, "w0": 1.5, "z": 0.0, "n": 1.0, "power": 4 * np.pi, "mode": mode}}
    # z + gdict.z < 0 on part of the tilted plane: the conjugated branch
    cases["vec_gauss_behind"] = {
        "kind": "vec", "source": _beam_plane(z=-1.0, tilt=True, half=(6.0, 5.0)),
        "params": {"lam": 0.9, "w0": 2.0, "z": 0.5, "n": 1.0, "power": 2.0, "mode": 0}}
    for mode in (0, 1, 2, 3):
        cases[f"currents_m{mode}"] = {"kind": "currents", "source": _beam_plane(z=2.0, tilt=True),
                                      "mode": mode, "seed": 40 + mode}
    cases["rt_tube_3ring"] = {"kind": "rt", "params": {"nRays": 4, "nRing": 3, "angx0": 2.0, "angy0": 3.0,
                                                       "x0": 5.0, "y0": 4.0}}
    cases["rt_tube_1ring"] = {"kind": "rt", "params": {"nRays": 7, "nRing": 1, "angx0": 1.0, "angy0": 0.5,
                                                       "x0": 2.0, "y0": 2.0}}
    cases["rt_tube_chief"] = {"kind": "rt", "params": {"nRays": 0, "nRing": 2, "angx0": 1.0, "angy0": 1.0,
                                                       "x0": 1.0, "y0": 1.0}}
    return cases


def grt_case():
    """Gaussian ray frame: pdf peak 1/(pi^2 x0 y0 angx0 angy0) = 0.405 < 1, so
    the accepted rays follow the truncated Gaussian undistorted."""
    return {"nRays": 1000, "x0": 5.0, "y0": 5.0, "angx0": float(np.degrees(0.1)),
            "angy0": float(np.degrees(0.1)), "seed": 7}


def random_fields(shape, seed):
    from pypo_amd import types as T
    rng = np.random.default_rng(seed)
    comps = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(6)]
    return T.fields(*comps)
