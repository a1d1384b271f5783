"""Benchmark: PO pair-interactions/s of the large JM -> FF propagation on one GPU.

Workload (the large mixed-precision far-field configuration):
  dish   paraboloid, D = 3000 mm, f = 1200 mm, uv grid 1001 x 1001 (1 002 001 points)
  feed   Huygens aperture at the focus, 33 x 33 xy grid over +-1.6 mm, facing the
         dish, Gaussian amplitude with waist chosen for -12 dB edge illumination
  far    AoE grid 501 x 501 over +-0.2 deg x +-0.2 deg
  lambda 1 mm (k = 2 pi / mm)
One step = JM (feed -> dish currents, near-field kernel) + FF (dish -> far
field), all buffers resident in HBM, mixed precision (FP32 integrand, FP64
distance / phase).  value = (JM pairs + FF pairs) / seconds per step.

Also reported: the all-FP64 GPU variant of the same step (median of 3) and
the same mixed step with the far field on the vector ALUs; the errors of
both precisions against the all-FP64 CPU chain (reference CPU library where
it was built, else this project's restatement) for the JM stage on every
dish point, the FF stage and the whole chain on a strided subset of
far-field rows; the FF kernel's own duration (HIP events on the launch
stream) as the roofline figure, with the flop ceiling and the issue ceiling
of its instruction stream; the ray tracer's rays/s on the Cassegrain
configuration; and the CPU baseline (the chain's FF call on the row subset)
timed on the host's cores.

Secondary objects in the same line: "sweep" (the 256-instance batched sweep,
autotuned, hipGraph replay) and "adjoint" (forward + backward of the
far-field power gradient on the 601^2 surface-error workload).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-fp64] [--no-rt]
                       [--no-sweep] [--no-adjoint]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LAM = 1.0
K = 2 * np.pi / LAM
F_DISH = 1200.0
D_DISH = 3000.0
N_DISH = 1001
N_FF = 501
FF_LIM = 0.2
N_FEED = 33
FEED_HALF = 1.6
EDGE_DB = -12.0
CPU_ROW_STRIDE = 100          # far-field rows 0, 100, ..., 500 for the CPU sample

# Algorithmic work per source-target pair of the far-field sum (FP32 + FP64 flops):
# 6 complex multiply-adds (48) + shadow test (6) in FP32; r'.r_hat (6) and the
# phase scale/reduction (4) in FP64; plus one sin and one cos.
FF_FLOPS_F32 = 54
FF_FLOPS_F64 = 10
FF_FLOPS_MAC = 48             # the 6 complex multiply-adds, part of the 54 above
PEAK_F32 = 157.3e12           # MI355X vector FP32 (AMD published)
PEAK_F64 = 78.6e12            # MI355X vector FP64 (AMD published)
PEAK_F16_MFMA = 2.5e15        # MI355X dense FP16 matrix (AMD published)
FF_FLOPS = FF_FLOPS_F32 + FF_FLOPS_F64
# Time per pair at the published peaks of the units the far-field kernel runs
# the work on: the multiply-adds on the matrix cores (FP16 operands, FP32
# accumulation), the shadow test in FP32 and the phase in FP64 on the vector
# ALUs.  The all-vector floor is kept beside it for the vector kernel.
PAIR_TIME_FLOOR = (FF_FLOPS_MAC / PEAK_F16_MFMA + (FF_FLOPS_F32 - FF_FLOPS_MAC) / PEAK_F32
                   + FF_FLOPS_F64 / PEAK_F64)
PAIR_TIME_FLOOR_VECTOR = FF_FLOPS_F32 / PEAK_F32 + FF_FLOPS_F64 / PEAK_F64
FF_KERNEL = "k_ff_mfma8"      # the far-field sum kernel of the default launch shape

# Issue model of k_ff_mfma8's instruction stream: cycles one SIMD spends
# issuing the vector instructions of 64 pairs (one pair per lane), from
# per-instruction issue costs measured on gfx950 with one-instruction loops
# at 4 waves per SIMD (FP32 FMA 3.6, FP64 FMA / v_fract_f64 / v_cvt_f32_f64
# 4.8, v_cvt_pk_f16_f32 5.4, v_fma_mix_f32 5.5, v_sin_f32 / v_cos_f32 9.8
# cycles per wave-instruction).  The 16x16x32 FP16 MFMAs (3 per 256 pairs)
# run beside this on the matrix core and are not charged.  Unlike the flop
# ceiling above, this one charges the sin/cos and the FP16 hi/lo operand
# split, which the algorithmic flop count leaves out.
ISSUE_CYCLES_PER_64_PAIRS = {
    "leader_phase_fp64": (3 * 4.8 + 4.8 + 4.8) / 4,     # one FP64 phase per group of 4 sources
    "follower_phase_fp32": 3 * 3 * 3.6 / 4,              # 3 FP32 FMAs for 3 of the 4 sources
    "sin_cos": 2 * 9.8,
    "fp16_hi_lo_split": 5.4 + 2 * 5.5 + 5.4,
}
SIMDS = 256 * 4               # MI355X: 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9              # MI355X peak engine clock


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def feed_waist():
    """Gaussian aperture waist w0 giving EDGE_DB at the dish rim: far-field
    amplitude exp(-(k w0 sin t / 2)^2) * (1 + cos t) / 2, divided by the focus
    distance 2f / (1 + cos t), relative to boresight."""
    te = 2 * np.arctan(D_DISH / (4 * F_DISH))
    target = 10 ** (EDGE_DB / 20)

    def edge(w0):
        pat = np.exp(-(K * w0 * np.sin(te) / 2) ** 2) * (1 + np.cos(te)) / 2
        space = F_DISH / (2 * F_DISH / (1 + np.cos(te)))
        return pat * space

    lo, hi = 1e-3, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if edge(mid) > target:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def surfaces():
    from pypo_amd import geometry as G
    dish = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, F_DISH],
                         "gmode": "uv", "lims_u": [0, D_DISH / 2], "lims_v": [0, 360],
                         "gridsize": [N_DISH, N_DISH]})
    feed = G.plane({"gmode": "xy", "lims_x": [-FEED_HALF, FEED_HALF], "lims_y": [-FEED_HALF, FEED_HALF],
                    "gridsize": [N_FEED, N_FEED], "flip": True})
    feed["transf"] = G.mat_translate([0, 0, F_DISH])
    ff = G.plane({"gmode": "AoE", "lims_Az": [-FF_LIM, FF_LIM], "lims_El": [-FF_LIM, FF_LIM],
                  "gridsize": [N_FF, N_FF]})
    return feed, dish, ff


def feed_currents(w0):
    from pypo_amd import types as T
    x = np.linspace(-FEED_HALF, FEED_HALF, N_FEED)
    X, Y = np.meshgrid(x, x, indexing="ij")
    amp = np.exp(-(X ** 2 + Y ** 2) / w0 ** 2)
    z = np.zeros_like(amp, dtype=complex)
    # E || x travelling -z through a surface with normal -z: J = -x, M = +y
    return T.currents(-amp + 0j, z, z, z, amp + 0j, z)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (profiles/pmc_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            return float(json.load(f)[kernel]["bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def run_rt(torch, D):
    """Cassegrain ray trace: 3162 x 3162 rays on a 3 m disk, paraboloid then
    hyperboloid, FP64, tol 1e-12.  Returns rays/s (both surfaces) and ms."""
    from pypo_amd import geometry as G
    n1 = 3162
    g = np.linspace(-D_DISH / 2, D_DISH / 2, n1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    rng = np.random.default_rng(3)
    tx = rng.normal(0, 1e-3, X.size)
    ty = rng.normal(0, 1e-3, X.size)
    d = np.stack([np.sin(tx), np.sin(ty), -np.ones(X.size)])
    d /= np.linalg.norm(d, axis=0)
    fr = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 2000.0), d[0], d[1], d[2]])
    prim = G.paraboloid({"pmode": "focus", "vertex": [0, 0, 0], "focus_1": [0, 0, F_DISH],
                         "gmode": "xy", "lims_x": [-1500, 1500], "lims_y": [-1500, 1500],
                         "gridsize": [3, 3]})
    sec = G.hyperboloid({"pmode": "focus", "focus_1": [0, 0, F_DISH], "focus_2": [0, 0, -300.0],
                         "ecc": 1.1, "gmode": "xy", "lims_x": [-150, 150], "lims_y": [-150, 150],
                         "gridsize": [3, 3]})
    from pypo_amd import bind as B
    lib = B.loadHIPlib()
    t_fr = D.to_device(fr)
    outs = D.trace([prim, sec], t_fr, tol=1e-12, t0=1000.0)
    torch.cuda.synchronize()
    # median of individually synchronised calls: one trace is ~0.4 ms, so a
    # single host-side stall would otherwise dominate a short mean
    walls = []
    for _ in range(10):
        t = time.perf_counter()
        outs = D.trace([prim, sec], t_fr, tol=1e-12, t0=1000.0)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t)
    dt = float(np.median(walls))
    lib.pypo_hip_set_timing(1)
    kms = []
    for _ in range(3):
        D.trace([prim, sec], t_fr, tol=1e-12, t0=1000.0)
        torch.cuda.synchronize()
        kms.append(lib.pypo_hip_last_kernel_ms())
    lib.pypo_hip_set_timing(0)
    del outs
    return X.size / dt, dt * 1e3, X.size, float(np.mean(kms))


def run_system_chain(torch, w0, feed, dish, ff):
    """The same mixed JM -> FF step through System.runPO, the reference's API:
    the JM currents stay in device memory between the two calls.  Median of
    3 synchronised steps after three warm steps: the first generates the
    grids, and torch's allocator reserves more memory in each of the first
    three (195, 292, 388 MB), as each step's results replace the last's."""
    from pypo_amd.system import System
    s = System()
    for name, refl in (("feed", feed), ("dish", dish), ("ff", ff)):
        s.system[name] = dict(refl, name=name)
    s.addCurrents("feedJM", feed_currents(w0), "feed", K)

    def step():
        s.runPO({"t_name": "dish", "mode": "JM", "s_current": "feedJM", "name_JM": "jm", "precision": "mixed"})
        s.runPO({"t_name": "ff", "mode": "FF", "s_current": "jm", "name_EH": "ff", "precision": "mixed"})

    for _ in range(3):
        step()
    return median_time(step, 3)


def run_sweep(torch, log):
    """Batched sweep: 64 frequencies (80-120 GHz) x 4 feed offsets = 256
    JM -> FF instances, dish grids 201^2 .. 401^2 (uneven), one 181^2 far
    field, each sum stage one launch, the run a replayed hipGraph; launch
    shape from the autotuner (fresh tuning, no cache)."""
    from pypo_amd import autotune as AT
    from pypo_amd import sweep as SW
    feeds, dishes, ks, curs, ff, meta = SW.frequency_offset_sweep()
    t = time.perf_counter()
    sw = SW.Sweep(feeds, dishes, ks, curs, ff, use_graph=False)
    setup_s = time.perf_counter() - t
    t = time.perf_counter()
    best, secs, table = AT.tune(sw, use_cache=False, use_graph=True)
    tune_s = time.perf_counter() - t
    sw.run()
    torch.cuda.synchronize()
    reps = 5
    t = time.perf_counter()
    for _ in range(reps):
        sw.run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    info = sw.info()
    pairs = info["pairs_jm"] + info["pairs_ff"]
    worst = max(table, key=lambda r: r["seconds"])
    sw.close()
    table_ms = [dict(AT.describe(r["params"]), ms=round(r["seconds"] * 1e3, 2),
                     range_ms=[round(r["min_s"] * 1e3, 2), round(r["max_s"] * 1e3, 2)]) for r in table]
    log(f"sweep {len(feeds)} instances: {dt * 1e3:.1f} ms/run ({pairs / dt:.3e} pairs/s), tuned {best} "
        f"in {tune_s:.1f} s")
    return {"instances": len(feeds), "pairs_per_run": pairs, "ms_per_run": round(dt * 1e3, 2),
            "pairs_per_s": pairs / dt, "launch": "hipGraph replay: JM sum, JM finish, pack, FF sum, FF finish",
            "tuned": AT.describe(best), "tuning_s": round(tune_s, 1), "setup_s": round(setup_s, 1),
            "slowest_candidate_ms": round(worst["seconds"] * 1e3, 1),
            "tuning_table": table_ms, "tuning_timing": "median of 3 synchronised replays per candidate",
            "workload": "64 freqs 80-120 GHz x offsets (0,2,-2,4) mm; paraboloid D=1000 f=400 uv "
                        "201^2..401^2; feed 25x25; AoE 181^2 +-1deg"}


def hybrid_system():
    """The hybrid section of PyPO's Cassegrain tutorial (tutorials/Tutorial5.ipynb
    cell 19) as a System, lambda = 2 mm: a Gaussian (w0 = 12 mm) on a 101 x 100
    polar plane at z = -1000 mm; the hyperboloid secondary (foci (0, 0, 3500)
    and (0, 0, -1000) mm, ecc 1.1, u up to 300 mm) on a 2501 x 2500 uv grid;
    the paraboloid primary (focus (0, 0, 3500) mm); a 1501 x 1500 polar plane
    at z = 3500 mm over the primary's aperture (u 300..5000 mm)."""
    from pypo_amd import geometry as G
    from pypo_amd.system import System
    s = System()
    s.addParabola({"name": "primary", "pmode": "focus", "gmode": "uv", "vertex": np.zeros(3),
                   "focus_1": np.array([0, 0, 3.5e3]), "lims_u": np.array([300, 5e3]),
                   "lims_v": np.array([0, 360]), "gridsize": np.array([320, 40])})
    s.addHyperbola({"name": "secondary", "pmode": "focus", "gmode": "uv", "flip": True,
                    "focus_1": np.array([0, 0, 3.5e3]), "focus_2": np.array([0, 0, -1e3]), "ecc": 1.1,
                    "lims_u": np.array([0, 300]), "lims_v": np.array([0, 360]), "gridsize": np.array([2501, 2500])})
    s.addPlane({"name": "focus", "gmode": "uv", "lims_u": np.array([0, 100]), "lims_v": np.array([0, 360]),
                "gridsize": np.array([101, 100]), "transf": G.mat_translate([0, 0, -1e3])})
    s.addPlane({"name": "foc_pri", "gmode": "uv", "lims_u": np.array([300, 5e3]), "lims_v": np.array([0, 360]),
                "gridsize": np.array([1501, 1500]), "transf": G.mat_translate([0, 0, 3.5e3])})
    s.createGaussian({"name": "source", "lam": 2.0, "w0x": 12.0, "w0y": 12.0, "n": 1, "E0": 1, "dxyz": 0,
                      "pol": np.array([1, 0, 0])}, "focus")
    return s


HYBRID_START = np.array([0, 0, 3.5e3])   # rays leave the secondary as from the primary focus


def hybrid_calls(s):
    """The three System calls of the hybrid section: EHP onto the secondary,
    the reflected rays to the primary (start-point amplitude, no
    interpolation), then to the focal plane with nearest interpolation of Ex."""
    def ehp():
        s.runPO({"t_name": "secondary", "s_current": "source", "mode": "EHP", "name_EH": "EH_sec",
                 "name_P": "P_sec"})

    def to_primary():
        s.runHybridPropagation({"fr_in": "P_sec", "field_in": "EH_sec", "t_name": "primary", "fr_out": "fr_pri",
                                "field_out": "f_pri", "start": HYBRID_START, "interp": False})

    def to_plane():
        s.runHybridPropagation({"fr_in": "fr_pri", "field_in": "f_pri", "t_name": "foc_pri", "fr_out": "fr_foc",
                                "field_out": "f_foc", "comp": "Ex"})
    return ehp, to_primary, to_plane


def run_hybrid(torch, log):
    """The hybrid section timed call by call through System (median of 3
    synchronised calls after a warm pass, results left in device memory)."""
    s = hybrid_system()
    calls = hybrid_calls(s)
    for c in calls:
        c()
    torch.cuda.synchronize()
    (t_e, r_e), (t_1, r_1), (t_2, r_2) = [median_time(c, 3) for c in calls]
    n_rays = int(np.prod(s.system["secondary"]["gridsize"]))
    n_grid = int(np.prod(s.system["foc_pri"]["gridsize"]))
    n_src = int(np.prod(s.system["focus"]["gridsize"]))
    log(f"hybrid: EHP {t_e * 1e3:.1f} ms, rays to the primary {t_1 * 1e3:.1f} ms, "
        f"rays to the plane + nearest {t_2 * 1e3:.1f} ms")
    return {"ms_ehp": round(t_e * 1e3, 2), "ms_ehp_range": r_e, "ehp_pairs": n_src * n_rays,
            "ms_hybrid_trace": round(t_1 * 1e3, 2), "ms_hybrid_trace_range": r_1,
            "ms_hybrid_trace_nearest": round(t_2 * 1e3, 2), "ms_hybrid_trace_nearest_range": r_2,
            "rays": n_rays, "grid_points": n_grid,
            "workload": "Cassegrain hybrid section: Gaussian 101x100 -> EHP -> hyperboloid uv 2501x2500, "
                        "rays -> paraboloid (start-point amplitude), rays -> plane 1501x1500 + nearest Ex; "
                        "lambda=2mm; System calls, results on the device"}


def run_adjoint(torch, log):
    """Differentiated far field: 601^2 dish (D = 3 m, f = 1.2 m, lambda = 1
    mm) with a 36-mode Zernike displacement of RMS 20 um, 201^2 far field over
    +-0.2 deg, loss sum (|E|^2 - P_ideal)^2; forward (JM + FF) and backward
    (currents recomputed + reverse far-field sum), mixed precision."""
    from pypo_amd import adjoint as PA
    feed, cur, dish, ff, k, radius = PA.surface_error_workload()
    prob = PA.ShapeProblem(feed, cur, dish, ff, k, precision="mixed")
    g0 = prob.g_dish0.cpu().numpy()
    d0 = torch.tensor(PA.zernike_displacement(g0[0], g0[1], radius, sigma=0.02, seed=2024), device="cuda")
    T = (prob.far_field(torch.zeros_like(d0))[0:6] ** 2).sum(0)

    def fwd_bwd():
        d = d0.clone().requires_grad_(True)
        P = prob.power(d)
        loss = ((P - T) ** 2).sum() / T.max() ** 2
        loss.backward()
        return d.grad

    fwd_bwd()
    torch.cuda.synchronize()
    reps = 5
    t = time.perf_counter()
    for _ in range(reps):
        prob.far_field(d0)
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t) / reps
    t = time.perf_counter()
    for _ in range(reps):
        g = fwd_bwd()
    torch.cuda.synchronize()
    t_fb = (time.perf_counter() - t) / reps
    E64 = prob.far_field(d0, precision="fp64")
    P64 = (E64[0:6] ** 2).sum(0)
    g64 = prob.power_grad(d0, E64, 2 * (P64 - T) / T.max() ** 2, precision="fp64")
    rel = (torch.linalg.norm(g - g64) / torch.linalg.norm(g64)).item()
    nf, nd, nt = prob.n_feed, prob.n_dish, prob.n_ff
    pairs_b = nf * nd + nd * nt
    t_b = t_fb - t_f
    log(f"adjoint: forward {t_f * 1e3:.1f} ms, forward+backward {t_fb * 1e3:.1f} ms, "
        f"mixed vs fp64 gradient {rel:.2e}")
    return {"ms_forward": round(t_f * 1e3, 2), "ms_forward_backward": round(t_fb * 1e3, 2),
            "backward_pairs": pairs_b, "backward_pairs_per_s": pairs_b / t_b,
            "grad_mixed_vs_fp64_rel": rel,
            "workload": "feed 33x33 -> paraboloid D=3000 f=1200 uv 601x601, 36 Zernike modes RMS 20um "
                        "-> AoE 201x201 +-0.2deg, lambda=1mm; gradient w.r.t. 361201 normal displacements"}


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on
    (its affinity mask), capped by OMP_NUM_THREADS where the host sets it
    (the CPU share of one GPU on a shared GPU host).  Returns the count and
    the figures it came from."""
    total = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = total
    omp = os.environ.get("OMP_NUM_THREADS", "")
    n = aff
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return n, {"os_cpu_count": total, "affinity_cores": aff, "omp_num_threads": omp or None}


def field_err(a, b):
    """Max |a - b| relative to the peak |b|, per field (rows 0-2 and 3-5 of a
    (6, n) complex array each have their own peak); the worse field."""
    worst = 0.0
    for g in (0, 3):
        pk = np.max(np.abs(b[g:g + 3]))
        if pk > 0:
            worst = max(worst, float(np.max(np.abs(a[g:g + 3] - b[g:g + 3])) / pk))
    return worst


def cpu_chain(feed, dish, cur_feed, sub_rows, threads):
    """The all-FP64 CPU chain with the reference CPU library (or this
    project's restatement where the reference was not built): JM over the
    whole dish, then FF on the strided far-field rows.  The FF call is the
    timed CPU baseline; the same call on half the rows checks that the
    sampled rate is linear in the rows."""
    from oracle import bind as OB
    from pypo_amd import geometry as G
    kind = "ref" if OB.available("ref") else "port"
    lib = OB.load(kind)
    t = time.perf_counter()
    jm, _, _ = OB.propagate(lib, "JM", feed, dish, cur_feed, K, threads)
    jm_s = time.perf_counter() - t
    nrow = len(sub_rows)
    sub = G.plane({"gmode": "AoE", "lims_Az": [-FF_LIM, FF_LIM], "lims_El": [-FF_LIM, FF_LIM],
                   "gridsize": [nrow, N_FF]})
    # median of 3 runs: a CPU job under a minute is timed as the median of at least three
    runs = []
    for _ in range(3):
        t = time.perf_counter()
        out, _, _ = OB.propagate(lib, "FF", dish, sub, jm, K, threads)
        runs.append(time.perf_counter() - t)
    dt = float(np.median(runs))
    vals = np.stack([c.ravel() for c in out.components()])
    pairs = N_DISH * N_DISH * nrow * N_FF
    half = G.plane({"gmode": "AoE", "lims_Az": [-FF_LIM, FF_LIM], "lims_El": [-FF_LIM, FF_LIM],
                    "gridsize": [nrow // 2, N_FF]})
    t = time.perf_counter()
    OB.propagate(lib, "FF", dish, half, jm, K, threads)
    dt_half = time.perf_counter() - t
    rate_half = N_DISH * N_DISH * (nrow // 2) * N_FF / dt_half
    n_thr, counts = threads, cpu_threads()[1]
    base = {"value": pairs / dt, "unit": "pairs/s", "cores": n_thr,
            "kind": "reference" if kind == "ref" else "port",
            "sample": f"FF rows {sub_rows[0]}..{sub_rows[-1]} step {CPU_ROW_STRIDE} of the 501x501 "
                      f"far field ({nrow * N_FF} targets x {N_DISH * N_DISH} sources = {pairs:.3e} "
                      f"pairs), one native call incl. its grid generation, median of 3: {dt:.2f} s "
                      f"(runs {', '.join(f'{r:.2f}' for r in runs)} s)",
            "threads_from": "affinity mask, capped by OMP_NUM_THREADS", **counts,
            "compile_flags": "g++ -O3 -pthread -fPIC (CMakeLists.txt:11 release flags)",
            "jm_full_dish_s": round(jm_s, 2),
            "linearity": {"rows": [nrow // 2, nrow], "pairs_per_s": [rate_half, pairs / dt],
                          "ratio": round((pairs / dt) / rate_half, 3)}}
    jm_vals = np.stack([c.ravel() for c in jm.components()])
    return vals, jm_vals, jm, base


def median_time(fn, reps):
    """Median and spread (ms) of `reps` synchronised calls of fn()."""
    import torch
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), [round(min(ts) * 1e3, 2), round(max(ts) * 1e3, 2)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-rt", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-hybrid", action="store_true")
    ap.add_argument("--profile-adjoint", action="store_true",
                    help="only the adjoint forward + backward (for rocprofv3 kernel statistics)")
    ap.add_argument("--profile-hybrid", action="store_true",
                    help="only the hybrid section's three System calls, warm then once more (for rocprofv3)")
    ap.add_argument("--profile-sweep", action="store_true",
                    help="only the batched sweep at its tuned launch shape (for rocprofv3)")
    ap.add_argument("--profile", action="store_true",
                    help="one launch of each kernel plus a 512 MiB device copy (counter calibration), "
                         "for rocprofv3 runs; prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the PO sums shard over targets with no exchange, "
              "but this benchmark runs on one GPU; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from pypo_amd import bind as B
    from pypo_amd import device as D

    lib = B.loadHIPlib()
    if lib.pypo_hip_device_count() < 1 or not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.profile_adjoint:
        run_adjoint(torch, log)
        return
    if args.profile_hybrid:
        calls = hybrid_calls(hybrid_system())
        for _ in range(2):
            for c in calls:
                c()
        torch.cuda.synchronize()
        log("hybrid profile run done")
        return
    if args.profile_sweep:
        from pypo_amd import sweep as SW
        feeds, dishes, ks, curs, ff_s, _ = SW.frequency_offset_sweep()
        sw = SW.Sweep(feeds, dishes, ks, curs, ff_s, tpl=2, chunk_tiles=4096, wgs_per_cu=64, use_graph=True)
        for _ in range(3):
            sw.run()
        torch.cuda.synchronize()
        sw.close()
        log("sweep profile run done")
        return

    w0 = feed_waist()
    feed, dish, ff = surfaces()
    log(f"setup: feed waist {w0:.4f} mm, grids on device")
    g_feed, g_dish, g_ff = D.generate_grid(feed), D.generate_grid(dish), D.generate_grid(ff)
    c_feed = D.currents_tensor(feed_currents(w0))
    n_feed, n_dish, n_ff = g_feed.shape[1], g_dish.shape[1], g_ff.shape[1]
    pairs_jm, pairs_ff = n_feed * n_dish, n_dish * n_ff
    cur_dish = torch.empty((12, n_dish), dtype=torch.float64, device="cuda")
    out_ff = torch.empty((12, n_ff), dtype=torch.float64, device="cuda")

    def step(prec):
        D.propagate("JM", g_feed, g_dish, c_feed, feed, K, precision=prec, out=cur_dish)
        D.propagate("FF", g_dish, g_ff, cur_dish, dish, K, precision=prec, out=out_ff)

    if args.profile:
        step("mixed")
        step("fp64")
        B.set_ff_kernel("vector")    # the same mixed step with the far field on the vector ALUs
        try:
            step("mixed")
        finally:
            B.set_ff_kernel("matrix")
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = src.clone()            # calibration copy: 512 MiB read + 512 MiB written
        torch.cuda.synchronize()
        if not args.no_rt:
            run_rt(torch, D)
        log(f"profile run done ({dst.numel() * 8 / 2**20:.0f} MiB calibration copy)")
        return

    for _ in range(args.warmup):
        step("mixed")
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step("mixed")
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / args.steps
    log(f"mixed step {dt * 1e3:.1f} ms")
    ff_mixed = D.tensor_to_complex6(out_ff)
    cur_host = D.tensor_to_complex6(cur_dish)

    # FF kernel alone, timed with HIP events on its stream (roofline figure)
    lib.pypo_hip_set_timing(1)
    kms = []
    for _ in range(3):
        D.propagate("FF", g_dish, g_ff, cur_dish, dish, K, precision="mixed", out=out_ff)
        torch.cuda.synchronize()
        kms.append(lib.pypo_hip_last_kernel_ms())
    lib.pypo_hip_set_timing(0)
    ff_ms = float(np.mean(kms))
    ff_pairs_s = pairs_ff / (ff_ms * 1e-3)
    achieved = ff_pairs_s * FF_FLOPS / 1e12
    peak = FF_FLOPS / PAIR_TIME_FLOOR / 1e12
    peak_vector = FF_FLOPS / PAIR_TIME_FLOOR_VECTOR / 1e12

    res = {}
    dts, spreads = run_system_chain(torch, w0, feed, dish, ff)
    res["system_runPO_ms_per_step"] = round(dts * 1e3, 2)
    res["system_runPO_ms_per_step_range"] = spreads
    res["system_runPO_vs_device_step"] = round(dts / dt, 3)
    log(f"System.runPO JM -> FF step {dts * 1e3:.1f} ms ({dts / dt:.3f} x the device-resident step)")
    issue_cycles = sum(ISSUE_CYCLES_PER_64_PAIRS.values())
    issue_ceiling = SIMDS * CLOCK_HZ * 64 / issue_cycles
    ff64 = jm64_host = None
    if not args.no_fp64:
        step("fp64")
        dt64, spread64 = median_time(lambda: step("fp64"), 3)
        ff64 = D.tensor_to_complex6(out_ff)
        jm64_host = D.tensor_to_complex6(cur_dish)
        res["fp64_ms_per_step"] = round(dt64 * 1e3, 2)
        res["fp64_ms_per_step_range"] = spread64
        res["fp64_timing"] = "median of 3 synchronised steps"
        res["fp64_pairs_per_s"] = (pairs_jm + pairs_ff) / dt64
        res["mixed_speedup_vs_fp64"] = round(dt64 / dt, 2)
        res["mixed_vs_fp64_gpu_max_err_rel_peak"] = field_err(ff_mixed, ff64)
        log(f"fp64 step {dt64 * 1e3:.1f} ms (median of 3, range {spread64}), "
            f"mixed vs fp64 err {res['mixed_vs_fp64_gpu_max_err_rel_peak']:.2e}")
        # the same mixed step with the far field on the vector ALUs (FP32 FMAs)
        B.set_ff_kernel("vector")
        try:
            step("mixed")
            dtv, spreadv = median_time(lambda: step("mixed"), 3)
        finally:
            B.set_ff_kernel("matrix")
        res["fp32_vector_ms_per_step"] = round(dtv * 1e3, 2)
        res["fp32_vector_ms_per_step_range"] = spreadv
        res["fp32_vector_speedup_vs_fp64"] = round(dt64 / dtv, 2)
        log(f"vector-ALU far field step {dtv * 1e3:.1f} ms ({dt64 / dtv:.2f}x the fp64 step)")

    cpu = None
    if not args.no_cpu:
        from pypo_amd import types as T
        rows = list(range(0, N_FF, CPU_ROW_STRIDE))
        threads, _ = cpu_threads()
        log(f"cpu chain: JM over the whole dish, then FF on {len(rows)} far-field rows, {threads} threads ...")
        vals, jm_cpu, jm_obj, cpu = cpu_chain(feed, dish, feed_currents(w0), rows, threads)
        idx = (np.asarray(rows)[:, None] * N_FF + np.arange(N_FF)[None, :]).ravel()
        # JM stage: the induced J on every dish point (M is zero for the PEC currents)
        res["jm_mixed_max_err_rel_peak_vs_cpu_fp64"] = field_err(cur_host, jm_cpu)
        # whole chain: the timed step's field (mixed JM -> mixed FF) against the FP64 CPU chain
        res["chain_mixed_max_err_rel_peak_vs_cpu_fp64"] = field_err(ff_mixed[:, idx], vals)
        # FF stage alone: the GPU far field of the CPU's FP64 currents
        cur_cpu = D.currents_tensor(jm_obj)
        D.propagate("FF", g_dish, g_ff, cur_cpu, dish, K, precision="mixed", out=out_ff)
        res["ff_mixed_max_err_rel_peak_vs_cpu_fp64"] = field_err(D.tensor_to_complex6(out_ff)[:, idx], vals)
        if ff64 is not None:
            res["jm_fp64_max_err_rel_peak_vs_cpu_fp64"] = field_err(jm64_host, jm_cpu)
            res["chain_fp64_max_err_rel_peak_vs_cpu_fp64"] = field_err(ff64[:, idx], vals)
            D.propagate("FF", g_dish, g_ff, cur_cpu, dish, K, precision="fp64", out=out_ff)
            res["ff_fp64_max_err_rel_peak_vs_cpu_fp64"] = field_err(D.tensor_to_complex6(out_ff)[:, idx], vals)
        torch.cuda.synchronize()
        res["err_sample"] = (f"JM: all {n_dish} dish points; FF and chain: far-field rows {rows[0]}.."
                             f"{rows[-1]} step {CPU_ROW_STRIDE} ({idx.size} targets, full sums); "
                             f"max |error| / peak per field (J; E and H)")
        res["speedup_vs_cpu_ff"] = round(ff_pairs_s / cpu["value"], 1)
        log(f"cpu {cpu['value']:.3e} pairs/s; {res}")

    if not args.no_rt:
        rays_s, rt_ms, nrays, rt_kms = run_rt(torch, D)
        res["rt_rays_per_s"] = rays_s
        res["rt_ms_cassegrain"] = round(rt_ms, 3)
        res["rt_kernel_ms"] = round(rt_kms, 3)
        res["rt_hbm_GBps"] = round(nrays * (48 + 2 * 48) / (rt_kms * 1e-3) / 1e9, 1)
        log(f"rt {rays_s:.3e} rays/s ({rt_ms:.2f} ms wall, kernel {rt_kms:.3f} ms, for {nrays} rays x 2 surfaces)")

    if not args.no_sweep:
        res["sweep"] = run_sweep(torch, log)
    if not args.no_adjoint:
        res["adjoint"] = run_adjoint(torch, log)
    if not args.no_hybrid:
        res["hybrid"] = run_hybrid(torch, log)

    line = {
        "metric": "PO pair-interactions/s JM->FF (mixed FP32/FP64 vs FP64) at <=1e-5 peak err; rays/s",
        "value": (pairs_jm + pairs_ff) / dt,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 phase + fp32 integrand, fp16 hi/lo MFMA products with fp32 accumulation",
        "data": "synthetic (seedless analytic Gaussian feed currents; grids generated on device)",
        "config": {"workload": "JM->FF large far field: feed 33x33 -> paraboloid D=3000 f=1200 uv "
                               "1001x1001 -> AoE 501x501 +-0.2deg, lambda=1mm",
                   "pairs_per_step": pairs_jm + pairs_ff, "jm_pairs": pairs_jm, "ff_pairs": pairs_ff,
                   "feed_waist_mm": round(w0, 5), "precision": "mixed"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": round(peak, 1),
                     "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                     "traffic": measured_traffic(FF_KERNEL),
                     "kernel": FF_KERNEL, "kernel_ms": round(ff_ms, 3),
                     "work_per_pair": f"{FF_FLOPS_F32} FP32 ({FF_FLOPS_MAC} of them multiply-adds, run on "
                                      f"the FP16 matrix cores) + {FF_FLOPS_F64} FP64 flops + sincos",
                     "peak_all_vector": round(peak_vector, 1),
                     "frac_all_vector": round(achieved / peak_vector, 4),
                     "issue_ceiling": {"pairs_per_s": issue_ceiling,
                                       "frac": round(ff_pairs_s / issue_ceiling, 4),
                                       "cycles_per_64_pairs": {k: round(v, 2)
                                                               for k, v in ISSUE_CYCLES_PER_64_PAIRS.items()},
                                       "basis": "vector issue of this kernel's instruction stream at "
                                                "measured gfx950 per-instruction costs, 1024 SIMDs x 2.4 GHz; "
                                                "charges sin/cos and the fp16 hi/lo split"}},
        "cpu_baseline": cpu,
    }
    line.update(res)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
