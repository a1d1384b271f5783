// po_launch.hip -- launch planning and the mode dispatch of one PO
// propagation: packed sources -> sum kernel (grid = target blocks x source
// splits) -> finish kernel (deterministic split reduction + epilogue).

#include "po_common.hpp"

#include <algorithm>

namespace pypo {

namespace {
int g_tpl_override = 0;    // 0: built-in choice
int g_wgs_override = 0;
int g_ff_kernel = PYPO_FF_MATRIX;
}  // namespace

int set_ff_kernel(int kind) {
    PYPO_REQUIRE(kind == PYPO_FF_MATRIX || kind == PYPO_FF_VECTOR, ERR_ARG,
                 "far-field kernel: 0 (matrix cores) or 1 (vector ALUs)");
    g_ff_kernel = kind;
    return 0;
}

int ff_kernel() { return g_ff_kernel; }

int device_cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        return 256;
    return n;
}

int tail_splits(long long blocks, long long tiles, int lo, int hi, long long slots) {
    int best_s = lo;
    long long best = -1;
    for (int s = lo; s <= hi; s++) {
        const long long chunk = (tiles + s - 1) / s;         // tiles per split
        const long long used = (tiles + chunk - 1) / chunk;  // splits launched
        const long long cost = (blocks * used + slots - 1) / slots * chunk;
        if (best < 0 || cost < best) { best = cost; best_s = s; }
    }
    return best_s;
}

int set_sum_shape(int tpl, int want_wgs) {
    PYPO_REQUIRE(tpl >= 0 && tpl <= 2 && want_wgs >= 0 && want_wgs <= (1 << 20), ERR_ARG,
                 "sum shape: tpl must be 0, 1 or 2 and want_wgs in [0, 2^20]");
    g_tpl_override = tpl;
    g_wgs_override = want_wgs;
    return 0;
}

SumPlan plan_sum(int mode, int precision, int n_src, int n_tgt) {
    SumPlan p;
    // Two targets per lane halve the LDS reads per pair in the far field;
    // the near-field and scalar kernels are register-heavy and keep one (two
    // per lane measured ~5 % slower on a 1089 x 1 002 001 JM: 116 registers,
    // half the waves per SIMD).
    // Measured on MI355X (361 201 sources; 10 201 .. 160 801 targets): two
    // targets per lane win at every size, by 7-15 %.
    p.tpl = (mode == PYPO_MODE_FF && n_tgt > TILE) ? 2 : 1;
    if (mode == PYPO_MODE_FF && g_tpl_override) p.tpl = g_tpl_override;
    p.tblocks = (n_tgt + TILE * p.tpl - 1) / (TILE * p.tpl);
    // Enough workgroups to fill 256 CUs many times over (the far field
    // gains 4-10 % from 2048 to 8192 workgroups: a shorter tail, better
    // balance over the 8 XCDs), down to one tile per split.  Splits only
    // multiply when target blocks are few, so the partials stay near
    // want_wgs x 256 targets x components.  On feed 101^2 -> dish 151^2 ->
    // FF 101^2 one tile per split against at least eight: mixed JM 0.78 ->
    // 0.64 ms, EH 1.12 -> 0.87 ms, FP64 FF 0.48 -> 0.41 ms.
    const int want_wgs = g_wgs_override ? g_wgs_override : (mode == PYPO_MODE_FF ? 8192 : 2048);
    int splits = (want_wgs + p.tblocks - 1) / p.tblocks;
    int max_splits = (n_src + TILE - 1) / TILE;
    if (splits > max_splits) splits = max_splits;
    if (splits < 1) splits = 1;
    // Far field: between half and twice that split count, the one whose last
    // round of concurrent workgroups (the kernel's occupancy x CUs) idles
    // least.  Large FF, 251 001 x 1 002 001 (two workgroups per CU): 17 -> 25
    // splits; matrix kernel 95.0 -> 93.7 ms, FP64 358.0 -> 352.9 ms, vector
    // 125.0 -> 122.9 ms.
    if (mode == PYPO_MODE_FF && !g_wgs_override)
        splits = tail_splits(p.tblocks, (n_src + TILE - 1) / TILE, std::max(1, splits / 2),
                             std::min(max_splits, 2 * splits),
                             (long long)ff_wgs_per_cu(precision, p.tpl) * device_cus());
    int chunk = (n_src + splits - 1) / splits;
    chunk = (chunk + TILE - 1) / TILE * TILE;
    p.chunk = chunk;
    p.splits = (n_src + chunk - 1) / chunk;
    if (p.splits < 1) p.splits = 1;
    return p;
}

size_t partial_bytes(int mode, const SumPlan &p, int n_tgt) {
    int comps = (mode == PYPO_MODE_SCALAR) ? 2 : 12;
    return (size_t)p.splits * comps * (size_t)n_tgt * sizeof(double);
}

SumPlan plan_cap_splits(const SumPlan &p, int n_src, int max_splits) {
    if (p.splits <= max_splits || max_splits < 1) return p;
    SumPlan q = p;
    int chunk = (n_src + max_splits - 1) / max_splits;
    q.chunk = (chunk + TILE - 1) / TILE * TILE;
    q.splits = (n_src + q.chunk - 1) / q.chunk;
    return q;
}

int launch_po(int mode, int precision, const void *packed, int n_src, const GridPtrs &ct,
              int n_tgt, double k, double epsilon, double t_dir, void *res, double *partial_ws,
              size_t ws_bytes, hipStream_t st, bool time_it) {
    PYPO_REQUIRE(n_src > 0 && n_tgt > 0, ERR_ARG, "empty source or target grid");
    return launch_po_plan(plan_sum(mode, precision, n_src, n_tgt), mode, precision, packed, n_src, ct, n_tgt,
                          k, epsilon, t_dir, res, partial_ws, ws_bytes, st, time_it);
}

int launch_po_plan(const SumPlan &p, int mode, int precision, const void *packed, int n_src,
                   const GridPtrs &ct, int n_tgt, double k, double epsilon, double t_dir, void *res,
                   double *partial_ws, size_t ws_bytes, hipStream_t st, bool time_it) {
    PYPO_REQUIRE(n_src > 0 && n_tgt > 0, ERR_ARG, "empty source or target grid");
    PYPO_REQUIRE(mode >= PYPO_MODE_JM && mode <= PYPO_MODE_SCALAR, ERR_ARG, "unknown PO mode");
    PYPO_REQUIRE(p.tpl >= 1 && p.tpl <= 2 && p.tblocks == (n_tgt + TILE * p.tpl - 1) / (TILE * p.tpl) &&
                     p.splits >= 1,
                 ERR_ARG, "internal: bad target plan");
    PYPO_REQUIRE(partial_bytes(mode, p, n_tgt) <= ws_bytes, ERR_ARG, "partial workspace too small");
    PYPO_REQUIRE((long long)p.chunk * p.splits >= n_src && p.chunk % TILE == 0, ERR_ARG,
                 "internal: bad source split plan");
    const PhysConsts pc = phys_consts(k, epsilon);
    int err;

    Timer tm(time_it, st);
    if (mode == PYPO_MODE_FF)
        err = launch_ff_sum(precision, p, packed, n_src, ct, n_tgt, -t_dir * k / kTwoPi, partial_ws, st);
    else if (mode == PYPO_MODE_SCALAR)
        err = launch_scalar_sum(precision, p, packed, n_src, ct, n_tgt, -t_dir * k / kTwoPi,
                                partial_ws, st);
    else
        err = launch_near_sum(mode, precision, p, packed, n_src, ct, n_tgt, k, t_dir, partial_ws, st);
    if (err) return err;
    tm.stop();

    if (mode == PYPO_MODE_FF)
        return launch_ff_finish(p, ct, n_tgt, t_dir, pc.prefactor, *(const c2Bundle *)res, partial_ws, st);
    if (mode == PYPO_MODE_SCALAR) {
        const arrC1 *b = (const arrC1 *)res;
        return launch_scalar_finish(p, n_tgt, -k * k / (4 * pc.pi_f), b->x, b->y, partial_ws, st);
    }
    OutC4 o{};
    const int nres = mode == PYPO_MODE_JMEH ? 24 : (mode == PYPO_MODE_EHP ? 15 : 12);
    double *const *ptrs = (double *const *)res;   // every bundle is a sequence of double*
    for (int q = 0; q < nres; q++) o.p[q] = ptrs[q];
    return launch_near_finish(mode, p, ct, n_tgt, pc.prefactor, pc.zeta_inv, o, partial_ws, st);
}

}  // namespace pypo
