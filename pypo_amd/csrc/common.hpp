// common.hpp -- shared declarations of libpypohip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/pypo_hip.h"

namespace pypo {

// ---- error channel --------------------------------------------------------
// Every failure is recorded per calling thread and surfaced by
// pypo_hip_last_error(); nothing calls exit() or resets the device.
void set_error(int code, const std::string &msg);

#define PYPO_HIP_TRY(expr)                                                        \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            ::pypo::set_error((int)_e, std::string(#expr) + ": " + hipGetErrorString(_e)); \
            return (int)_e;                                                       \
        }                                                                         \
    } while (0)

#define PYPO_CHECK_LAUNCH()                                                       \
    do {                                                                          \
        hipError_t _e = hipGetLastError();                                        \
        if (_e != hipSuccess) {                                                   \
            ::pypo::set_error((int)_e, std::string("kernel launch: ") + hipGetErrorString(_e)); \
            return (int)_e;                                                       \
        }                                                                         \
    } while (0)

#define PYPO_REQUIRE(cond, code, msg)                                             \
    do {                                                                          \
        if (!(cond)) {                                                            \
            ::pypo::set_error((code), (msg));                                     \
            return (code);                                                        \
        }                                                                         \
    } while (0)

enum { ERR_ARG = 1001, ERR_ALLOC = 1002, ERR_NODEV = 1003 };

// ---- plain-pointer views used inside kernels -------------------------------
struct GridPtrs {
    double *x, *y, *z, *nx, *ny, *nz, *area;
};

inline GridPtrs grid_ptrs(const reflcontainer &c) {
    return GridPtrs{c.x, c.y, c.z, c.nx, c.ny, c.nz, c.area};
}

// Complex 3-vector pairs in the c2Bundle order r1x..i2z: index q = 0..11
// is (re/im, field, axis) as [r1x r1y r1z i1x i1y i1z r2x r2y r2z i2x i2y i2z].
struct Bundle12 {
    double *p[12];
};

inline Bundle12 bundle12(const c2Bundle &b) {
    return Bundle12{{b.r1x, b.r1y, b.r1z, b.i1x, b.i1y, b.i1z,
                     b.r2x, b.r2y, b.r2z, b.i2x, b.i2y, b.i2z}};
}

// ---- packed source records -------------------------------------------------
// One source point of the PO sums as the kernels read it from HBM/LDS:
// position in FP64 (the phase and the distance need it), unit normal in
// FP32 (only the shadowing sign), and the six complex current components
// pre-multiplied by the trapezoid weight and the area element, in FP32.
// 96 bytes; the J-only, unmasked far-field loop reads bytes 0-23 and 40-63.
struct __attribute__((aligned(16))) SrcRec32 {
    double x, y, z;   // 0..23
    float nx, ny;     // 24..31
    float nz, pad0;   // 32..39
    float cj[6];      // 40..63  w*A*(Jx.re, Jx.im, Jy.re, Jy.im, Jz.re, Jz.im)
    float cm[6];      // 64..87  w*A*(Mx.re, ..., Mz.im)
    float pad1[2];    // 88..95
};
static_assert(sizeof(SrcRec32) == 96, "SrcRec32 must be 96 bytes");

// Per 256-source tile: the cone that holds the tile's normals (unit axis and
// the cosine of its half-angle; -1 when no cone under 90 degrees exists) and
// whether any magnetic current in the tile is non-zero.  Lets the far-field
// sum skip the per-pair shadow test and the M terms where they cannot matter.
// cmax: the largest |w A current| component of the tile (FP32 records), the
// scale of the matrix-core far-field sum's FP16 operands.
struct __attribute__((aligned(16))) TileMeta {
    float ax, ay, az, cos_half;
    int has_m, n_valid;
    float cmax;
    int pad1;
};
static_assert(sizeof(TileMeta) == 32, "TileMeta must be 32 bytes");

// FP64 twin for the all-FP64 kernels: 6 + 12 doubles = 144 bytes.
struct __attribute__((aligned(16))) SrcRec64 {
    double x, y, z;
    double nx, ny, nz;
    double c[12];
};
static_assert(sizeof(SrcRec64) == 144, "SrcRec64 must be 144 bytes");

// Scalar-mode records: position FP64 and w*A*u (complex) in FP32 / FP64.
struct __attribute__((aligned(16))) SrcRecS32 {
    double x, y, z;
    float ur, ui;
};
static_assert(sizeof(SrcRecS32) == 32, "SrcRecS32 must be 32 bytes");

struct __attribute__((aligned(16))) SrcRecS64 {
    double x, y, z;
    double ur, ui;
    double pad;
};
static_assert(sizeof(SrcRecS64) == 48, "SrcRecS64 must be 48 bytes");

// ---- device math -----------------------------------------------------------

// sin and cos of 2*pi*f, f in revolutions, on the hardware transcendental unit
// (v_sin_f32 / v_cos_f32 take their argument in revolutions).  Callers pass a
// reduced f in [-0.5, 1), except the far field's follower phases, |f| <= 9:
// the unit reduces those itself to within one ulp of its values on fract(f).
__device__ __forceinline__ void sincos_rev_f32(float f, float &s, float &c) {
    s = __builtin_amdgcn_sinf(f);
    c = __builtin_amdgcn_cosf(f);
}

// sin and cos of 2*pi*f in FP64 for a phase |f| <= 1 revolution (the FP64
// kernels pass fract(phase)).  Quadrant n = rint(4 f), r = f - n / 4 (exact,
// |r| <= 1/8); sin(2 pi r) = r P(r^2) with P of degree 6 and cos(2 pi r) =
// Q(r^2) with Q of degree 7, both interpolating at Chebyshev nodes of r^2 in
// [0, 1/64] (fitted in 50-digit arithmetic; with the coefficients rounded to
// double the polynomials are within 2.6e-17 and 1.5e-17 of sin and cos); then
// the quadrant's swap and signs.  No special cases: the argument is a finite
// fraction by construction.
__device__ __forceinline__ void sincos_rev_f64(double f, double &s, double &c) {
    const double n = __builtin_rint(4.0 * f);
    const double r = __builtin_fma(n, -0.25, f);
    const double r2 = r * r;
    double ps = 0x1.e3f38399551bfp+1;
    ps = __builtin_fma(ps, r2, -0x1.e30071afc3e59p+3);
    ps = __builtin_fma(ps, r2, 0x1.50782fda12d96p+5);
    ps = __builtin_fma(ps, r2, -0x1.32d2cce2e5b19p+6);
    ps = __builtin_fma(ps, r2, 0x1.466bc677587f8p+6);
    ps = __builtin_fma(ps, r2, -0x1.4abbce625be41p+5);
    ps = __builtin_fma(ps, r2, 0x1.921fb54442d18p+2);   // 2 pi
    ps *= r;
    double pc = -0x1.b264ba152378ap+0;
    pc = __builtin_fma(pc, r2, 0x1.f9cc41140bb60p+2);
    pc = __builtin_fma(pc, r2, -0x1.a6d1ec7906c20p+4);
    pc = __builtin_fma(pc, r2, 0x1.e1f5068355e15p+5);
    pc = __builtin_fma(pc, r2, -0x1.55d3c7e3c90f8p+6);
    pc = __builtin_fma(pc, r2, 0x1.03c1f081b5aacp+6);
    pc = __builtin_fma(pc, r2, -0x1.3bd3cc9be45dep+4);
    pc = __builtin_fma(pc, r2, 1.0);
    const int q = (int)n & 3;   // sin(2 pi (r + q/4)), cos(2 pi (r + q/4))
    const double sv = (q & 1) ? pc : ps, cv = (q & 1) ? ps : pc;
    s = (q & 2) ? -sv : sv;
    c = ((q + 1) & 2) ? -cv : cv;
}

// FP64 phase in revolutions -> reduced FP32 fraction in [-0.5, 0.5].
__device__ __forceinline__ float reduce_rev(double rev) {
    return (float)(rev - __builtin_rint(rev));
}

// Trapezoid weight of source i on an ncx x ncy grid (Propagation.h:1123-1135:
// half on the xu edges always, half on the yv edges unless gmode == 1).
__host__ __device__ __forceinline__ double trapezoid_weight(int i, int gmode, int ncx, int ncy) {
    int xu = i / ncy, yv = i - xu * ncy;
    double w = (xu == 0 || xu == ncx - 1) ? 0.5 : 1.0;
    if (gmode != 1 && (yv == 0 || yv == ncy - 1)) w *= 0.5;
    return w;
}

// Physical constants exactly as the reference stores them: float literals
// widened to double (Propagation.h:196-203).
struct PhysConsts {
    double pi_f, prefactor, zeta_inv;
};

inline PhysConsts phys_consts(double k, double epsilon) {
    PhysConsts c;
    c.pi_f = (double)3.14159265359f;
    double c_l = (double)2.99792458e8f;
    double mu0 = (double)1.2566370614e-6f;
    double eps_vac = 1 / (mu0 * c_l * c_l);
    double eps = epsilon * eps_vac;
    double zeta = std::sqrt(mu0 / eps);
    c.zeta_inv = 1 / zeta;
    c.prefactor = k * k / (4. * c.pi_f);
    return c;
}

constexpr double kTwoPi = 6.283185307179586476925286766559;

// ---- device buffer pool ----------------------------------------------------
// Grow-only named device allocations, reused across calls so chained calls
// do not pay hipMalloc each time.
void *pool_get(int slot, size_t bytes, int *err);
enum PoolSlot {
    SLOT_SRC_GRID = 0, SLOT_TGT_GRID, SLOT_CURRENTS, SLOT_RESULT, SLOT_PACKED,
    SLOT_PARTIAL, SLOT_ERRS_S, SLOT_ERRS_T, SLOT_RAYS, SLOT_RT_ERRS, SLOT_ADJ, SLOT_FLAG, SLOT_NN, SLOT_METRICS,
    SLOT_COUNT
};

hipStream_t lib_stream(int *err);

// ---- launchers (implemented in the .hip files) -----------------------------
// r's arrays are host memory (read at launch); g and d_errs are device memory.
int launch_generate_grid(const reflparams &r, const GridPtrs &g, int n, const double *d_errs,
                         bool transform, bool spheric, hipStream_t st);

int launch_pack_sources(int precision, const GridPtrs &cs, int n_src, const Bundle12 &cur,
                        int gmode, int ncx, int ncy, void *packed, hipStream_t st);
int launch_pack_scalar(int precision, const GridPtrs &cs, int n_src, const double *ur,
                       const double *ui, int gmode, int ncx, int ncy, void *packed, hipStream_t st);
// Packed records plus (for the vector modes) one TileMeta per 256 sources.
size_t packed_bytes(int mode, int precision, int n_src);

// Source-split plan of a PO sum: grid.y = splits, each split sums `chunk`
// sources (a multiple of the 256-source LDS tile) into its own FP64 partial.
struct SumPlan {
    int tpl;          // targets per lane
    int tblocks;      // workgroups along the targets
    int splits;       // source splits (grid.y)
    int chunk;        // sources per split
};
SumPlan plan_sum(int mode, int precision, int n_src, int n_tgt);
int set_sum_shape(int tpl, int want_wgs);   // 0, 0: built-in choice
int set_ff_kernel(int kind);                // PYPO_FF_MATRIX or PYPO_FF_VECTOR
int ff_kernel();
size_t partial_bytes(int mode, const SumPlan &p, int n_tgt);
// p with at most max_splits source splits (whole tiles per split).
SumPlan plan_cap_splits(const SumPlan &p, int n_src, int max_splits);
// Compute units of the current device (a launch-shape hint; 256 if the
// query fails).
int device_cus();
// Split count s in [lo, hi] that minimises the rounds of `slots` concurrent
// workgroups over blocks x s workgroups times the tiles each one runs
// (ceil(tiles / s)): the idle tail of the last round.  Ties: fewer splits.
int tail_splits(long long blocks, long long tiles, int lo, int hi, long long slots);

// One PO propagation from packed sources to the target grid ct.  `res`
// points to a device-side bundle struct of the mode's result type
// (c2Bundle / c4Bundle / c2rBundle / arrC1) whose arrays are device memory.
int launch_po(int mode, int precision, const void *packed, int n_src, const GridPtrs &ct,
              int n_tgt, double k, double epsilon, double t_dir, void *res, double *partial_ws,
              size_t ws_bytes, hipStream_t st, bool time_it);
// The same with a given launch plan (plan_sum's, or one derived from it).
int launch_po_plan(const SumPlan &p, int mode, int precision, const void *packed, int n_src,
                   const GridPtrs &ct, int n_tgt, double k, double epsilon, double t_dir, void *res,
                   double *partial_ws, size_t ws_bytes, hipStream_t st, bool time_it);

// Trace frames[0] through surfs[0..n_surf) (n_surf <= 4); frames[k+1]
// receives the frame after surface k.  d_errs[k]: per-ray surface error
// deviates on the device, or nullptr.
int launch_trace_chain(const reflparams *surfs, const double *const *d_errs, int n_surf,
                       const cframe *frames, double tol, double t0, int max_iter, hipStream_t st,
                       bool time_it);

// Batched sweeps (sweep.hip).
struct Sweep;
int sweep_create(const pypo_sweep_desc *d, Sweep **out, hipStream_t st);
int sweep_plan(Sweep *s, struct SweepPlanParams prm, bool use_graph, hipStream_t st);
int sweep_run(Sweep *s, hipStream_t st);
int sweep_read(const Sweep *s, int inst, c2Bundle *host, hipStream_t st);
int sweep_read_currents(const Sweep *s, int inst, c2Bundle *host, hipStream_t st);
void sweep_info(const Sweep *s, double *pairs_jm, double *pairs_ff, int *n_items_jm, int *n_items_ff,
                double **ff_out);
void sweep_destroy(Sweep *s);
struct SweepPlanParams {
    int tpl, chunk_tiles, wgs_per_cu;
};

// Adjoint of the far-field power (adjoint.hip).  out = g0 + d n (points
// along their normals; normals and areas copied).
int launch_displace(const GridPtrs &g0, const double *d, int n, const GridPtrs &out, hipStream_t st);
size_t adjoint_workspace_bytes(int precision, int n_feed, int n_dish, int n_tgt);
// grad[i] = dL/dd_i for L with dL/d|E_t|^2 = g[t]; dish is the displaced
// grid, E the far field of its JM currents.  All pointers device memory.
int launch_ff_power_grad(int precision, const GridPtrs &dish, int n_dish, int gmode, int ncx, int ncy,
                         const GridPtrs &feed, int n_feed, int fgmode, int fncx,
                         int fncy, const Bundle12 &feed_cur, const GridPtrs &ff, int n_tgt,
                         const Bundle12 &E, const double *g, double k, double tdir, double *grad,
                         void *ws, size_t ws_bytes, hipStream_t st);

// Beam initialisation (beam.hip).  Grids and bundles are device memory.
int launch_gauss(const GPODict &g, const GridPtrs &gr, int n, const Bundle12 &field, const Bundle12 &cur,
                 hipStream_t st);
int launch_scalar_gauss(const ScalarGPODict &g, const GridPtrs &gr, int n, double *ur, double *ui,
                        hipStream_t st);
int launch_gauss_beam(const vecGPODict &g, const GridPtrs &gr, int n, const Bundle12 &field,
                      const Bundle12 &cur, hipStream_t st);
int launch_calc_jm(const Bundle12 &field, const GridPtrs &gr, int n, int mode, const Bundle12 &cur,
                   hipStream_t st);
int rt_frame_size(const RTDict &r);
int launch_rt_frame(const RTDict &r, const cframe &fr, hipStream_t st);
int launch_grt_frame(const GRTDict &g, uint64_t seed, const cframe &fr, int *d_fail, int max_trials,
                     hipStream_t st);
constexpr int kGrtMaxTrials = 1 << 20;

// Hybrid ray-trace / PO (hybrid.hip).
int launch_hybrid_phase(const cframe &in, const cframe &out, const Bundle12 &fin, const Bundle12 &fout,
                        int n, double k, const double *start, hipStream_t st);
size_t nn_workspace_bytes(int n_points);
int launch_interp_nearest(const double *px, const double *py, const double *pz, int n, const Bundle12 &vin,
                          int nvals, const double *qx, const double *qy, const double *qz, int m,
                          const Bundle12 &vout, void *ws, size_t ws_bytes, hipStream_t st);

// ---- figures of merit (metrics.hip) ----
struct FrameView {
    const double *x, *y, *z, *dx, *dy, *dz;
};
using AperMask = aperparams;
size_t metrics_workspace_bytes();
size_t metrics_result_offset();   // in doubles
int launch_aperture_sums(const double *x, const double *y, const double *area, const double *fr,
                         const double *fi, int n, const AperMask &m, double *ws, double *out5, hipStream_t st);
int launch_power_sums(const double *ar, const double *ai, const double *br, const double *bi, int n, double *ws,
                      double *out2, hipStream_t st);
int launch_frame_stats(const FrameView &f, int n, double *ws, double *out8, hipStream_t st);

// Kernel timing of the last timed launch (HIP events on the launch stream).
void record_kernel_ms(double ms);

}  // namespace pypo
