/*
 * pypo_hip.h -- C ABI of libpypohip.so, the MI355X (gfx950) physical-optics
 * and ray-trace library.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in symbols.  Same names, argument types and argument meaning as
 *     the reference's FP64 CPU library (src/CPU/InterfaceCPU.h:24-64).  All
 *     buffers are host memory owned by the caller (ctypes arrays); the
 *     library generates the source and target grids itself, exactly like
 *     src/CPU/InterfaceCPU.cpp:29-243, and writes the result in place.
 *     `numThreads` is accepted and ignored (the GPU launch shape is chosen
 *     by the library).  Calls are synchronous and never exit the process:
 *     failures are reported through pypo_hip_last_error().
 *
 *  2. pypo_hip_* extensions.  Device-resident variants that take device
 *     pointers and a hipStream_t (passed as void*), for chaining JM -> FF
 *     without host round trips, batched sweeps and the autograd wrapper.
 *
 * Struct layouts are byte-identical to src/include/Structs.h:29-161 so that
 * the reference's ctypes mirrors (src/PyPO/Structs.py:9-141) bind unchanged.
 * Lengths are in mm, k in 1/mm, t_direction = -1 forward / +1 backward
 * (src/PyPO/BindCPU.py:117-121).
 */
#ifndef PYPO_HIP_H
#define PYPO_HIP_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- structs: layout of src/include/Structs.h:29-161 ------------------ */

/* One complex scalar field, SoA real/imag.  Structs.h:29-32 */
typedef struct arrC1 { double *x, *y; } arrC1;

/* Two complex 3-vectors per point (J,M or E,H), SoA.  Structs.h:42-57 */
typedef struct c2Bundle {
    double *r1x, *r1y, *r1z, *i1x, *i1y, *i1z;
    double *r2x, *r2y, *r2z, *i2x, *i2y, *i2z;
} c2Bundle;

/* Four complex 3-vectors (J,M,E,H).  Structs.h:59-88 */
typedef struct c4Bundle {
    double *r1x, *r1y, *r1z, *i1x, *i1y, *i1z;
    double *r2x, *r2y, *r2z, *i2x, *i2y, *i2z;
    double *r3x, *r3y, *r3z, *i3x, *i3y, *i3z;
    double *r4x, *r4y, *r4z, *i4x, *i4y, *i4z;
} c4Bundle;

/* Two complex 3-vectors and one real 3-vector (E_r, H_r, S_r).  Structs.h:90-108 */
typedef struct c2rBundle {
    double *r1x, *r1y, *r1z, *i1x, *i1y, *i1z;
    double *r2x, *r2y, *r2z, *i2x, *i2y, *i2z;
    double *r3x, *r3y, *r3z;
} c2rBundle;

/* Surface parameters, passed BY VALUE.  Structs.h:111-132
 * type: 0 paraboloid, 1 hyperboloid, 2 ellipsoid, 3 plane.
 * gmode: 0 xy, 1 uv, 2 AoE.  transf: 4x4 row-major affine matrix. */
typedef struct reflparams {
    double *coeffs;       /* a, b, c */
    double *lxu;          /* lower/upper x or u (or Az) */
    double *lyv;          /* lower/upper y or v (or El) */
    int *n_cells;         /* grid size along x/u and y/v */
    bool flip;
    int gmode;
    double *gcenter;
    double ecc_uv;
    double rot_uv;        /* degrees */
    int type;
    double *transf;
    double rms;
    unsigned int rms_seed;
} reflparams;

/* Realised grid: points, unit normals and area elements.  Structs.h:135-148 */
typedef struct reflcontainer {
    int size;
    double *x, *y, *z, *nx, *ny, *nz, *area;
} reflcontainer;

/* Ray frame: positions and directions.  Structs.h:151-161 */
typedef struct cframe {
    int size;
    double *x, *y, *z, *dx, *dy, *dz;
} cframe;

/* ---- 1. drop-in FP64 entry points (src/CPU/InterfaceCPU.h:24-64) ------ */

/* PEC currents J = 2 n x H_inc, M = 0 on the target.  InterfaceCPU.h:24-28,
 * InterfaceCPU.cpp:29-42 */
void propagateToGrid_JM(c2Bundle *res, reflparams source, reflparams target,
                        reflcontainer *cs, reflcontainer *ct, c2Bundle *currents,
                        double k, int numThreads, double epsilon, double t_direction);

/* Incident E, H on the target.  InterfaceCPU.h:30-34, InterfaceCPU.cpp:64-77 */
void propagateToGrid_EH(c2Bundle *res, reflparams source, reflparams target,
                        reflcontainer *cs, reflcontainer *ct, c2Bundle *currents,
                        double k, int numThreads, double epsilon, double t_direction);

/* PEC J, M plus incident E, H (res->1,2 = J,M; res->3,4 = E,H).
 * InterfaceCPU.h:36-40, InterfaceCPU.cpp:100-113 */
void propagateToGrid_JMEH(c4Bundle *res, reflparams source, reflparams target,
                          reflcontainer *cs, reflcontainer *ct, c2Bundle *currents,
                          double k, int numThreads, double epsilon, double t_direction);

/* Reflected E_r, H_r and reflected Poynting direction S_r.
 * InterfaceCPU.h:42-46, InterfaceCPU.cpp:136-149 */
void propagateToGrid_EHP(c2rBundle *res, reflparams source, reflparams target,
                         reflcontainer *cs, reflcontainer *ct, c2Bundle *currents,
                         double k, int numThreads, double epsilon, double t_direction);

/* Scalar field.  InterfaceCPU.h:48-52, InterfaceCPU.cpp:171-184 */
void propagateToGrid_scalar(arrC1 *res, reflparams source, reflparams target,
                            reflcontainer *cs, reflcontainer *ct, arrC1 *field,
                            double k, int numThreads, double epsilon, double t_direction);

/* Far field E, H on an (polar, azimuth) target grid.  InterfaceCPU.h:54-58,
 * InterfaceCPU.cpp:206-219 */
void propagateToFarField(c2Bundle *res, reflparams source, reflparams target,
                         reflcontainer *cs, reflcontainer *ct, c2Bundle *currents,
                         double k, int numThreads, double epsilon, double t_direction);

/* Ray trace fr_in onto the surface ctp; epsilon is the Newton tolerance and
 * t0 the starting parameter.  InterfaceCPU.h:60-61, InterfaceCPU.cpp:236-243.
 * Like the reference, fr_in is transformed into the surface frame and back
 * (RayTrace.h:284-304), so it is left numerically as it was. */
void propagateRays(reflparams ctp, cframe *fr_in, cframe *fr_out,
                   int numThreads, double epsilon, double t0);

/* Surface grid generation (device kernels, FP64) into host arrays, same
 * contract as src/common/InterfaceReflector.h:21-22 (generateGrid).
 * transform applies refl.transf; spheric maps AoE (Az, El) to
 * (polar, azimuth) angles in radians.  The PO entry points above call it
 * with transform = spheric = true. */
void generateGrid(reflparams refl, reflcontainer *container, bool transform, bool spheric);

/* ---- beam initialisation: drop-in for the reference's beam library ------
 * (src/common/InterfaceBeam.h:21-34, loaded as libpypobeam.so by
 * src/PyPO/BindBeam.py:18-50).  Parameter structs: Structs.h:164-213;
 * angles in radians (BindUtils.allfill_RTDict / allfill_GRTDict convert).
 * Grids of `plane` are generated on the device with transform = spheric =
 * true, as BeamInit.h does; results are written into the caller's host
 * arrays; synchronous. */
typedef struct RTDict { int nRays; int nRing; double angx0, angy0, x0, y0; } RTDict;
typedef struct GRTDict { int nRays; double angx0, angy0, x0, y0; int seed; } GRTDict;
typedef struct GPODict { double lam, w0x, w0y, n, E0, dxyz; double *pol; } GPODict;
typedef struct vecGPODict { double lam, w0, z, n, power; int mode; } vecGPODict;
typedef struct ScalarGPODict { double lam, w0x, w0y, n, E0, dxyz; } ScalarGPODict;

/* Tubular ray frame, 1 + 4 nRays nRing rays (BeamInit.h:159-214); sets fr->size. */
void makeRTframe(RTDict rdict, cframe *fr);
/* Gaussian ray frame by rejection sampling (BeamInit.h:216-286).  Rays are
 * drawn per ray from a counter-based generator keyed by `seed` (seed = -1:
 * a random key).  The reference ignores the seed (its seeded generator goes
 * out of scope, BeamInit.h:224) and draws from std::random_device, so only
 * the distribution is comparable.  Fails when a ray finds no accepted
 * candidate within 2^20 trials. */
void makeGRTframe(GRTDict grdict, cframe *fr);
/* Astigmatic Gaussian E field on `plane`, H = 0, J = 0, M = -2 n x E
 * (BeamInit.h:288-405). */
void makeGauss(GPODict gdict, reflparams plane, c2Bundle *res_field, c2Bundle *res_current);
/* Vector Gaussian by the complex source point: E, H and J, M by mode
 * 0 full / 1 PMC / 2 PEC (BeamInit.h:407-596). */
void makeGaussBeam(vecGPODict gdict, reflparams plane, c2Bundle *res_field, c2Bundle *res_current);
/* Scalar Gaussian (BeamInit.h:598-658). */
void makeScalarGauss(ScalarGPODict sgdict, reflparams plane, arrC1 *res_field);
/* Currents from fields on `rdict`'s grid, mode 0 full / 1 PMC / 2 PEC /
 * 3 source (BeamInit.h:660-844). */
void calcCurrents(c2Bundle *res_field, c2Bundle *res_current, reflparams rdict, int mode);

/* ---- 2. extensions --------------------------------------------------- */

enum pypo_hip_mode {
    PYPO_MODE_JM = 0, PYPO_MODE_EH = 1, PYPO_MODE_JMEH = 2,
    PYPO_MODE_EHP = 3, PYPO_MODE_FF = 4, PYPO_MODE_SCALAR = 5
};

enum pypo_hip_precision {
    PYPO_PREC_MIXED = 0,   /* FP32 integrand, FP64 distance and phase */
    PYPO_PREC_FP64 = 1     /* everything in FP64 */
};

/* Error of the calling thread: the first failure since the last read (later
 * failures do not overwrite it; a successful call does not clear it).
 * Returns 0 if none, otherwise a non-zero code and the message copied into
 * buf (n bytes); either way the slot is cleared. */
int pypo_hip_last_error(char *buf, int n);

/* Library version string. */
const char *pypo_hip_version(void);

/* Number of visible HIP devices (0 when none).  Does not create a context. */
int pypo_hip_device_count(void);

/* Scratch buffers of the device-resident and drop-in calls: each stream that
 * calls the library keeps its own grow-only set (calls on different streams
 * are not ordered against each other).  Returns 1 when a set grows with the
 * stream-ordered allocator (old buffer freed, new one allocated in the
 * stream's order, nothing else waits), 0 when growth waits for the set's
 * stream, negative on error (no device).  Device-resident calls may be
 * captured into a graph once the same call has run on the capturing stream:
 * inside a capture the library skips its ordering events and reports growth
 * as an error; a replay uses that stream's buffers.  No counterpart in the
 * reference (its CPU library allocates per call). */
int pypo_hip_stream_ordered_scratch(void);

/* Device bytes of scratch the library holds for calls on `stream` (a
 * hipStream_t, NULL: the null stream); 0 for a stream it has not seen,
 * negative on error (no device).  No counterpart in the reference. */
long long pypo_hip_scratch_bytes(void *stream);

/* Frees the scratch buffers the library holds for `stream`, for callers
 * that use many short-lived streams (the sets are grow-only and otherwise
 * kept until the process ends).  Waits until the last call enqueued on the
 * stream outside a capture is done with them (the stream is not passed to
 * HIP, so it may already be destroyed).  A graph captured from calls on the
 * stream uses these buffers: do not call this during such a capture or while
 * such a graph may still be launched.  The next call on the stream
 * allocates a new set.  Returns 0 on success. */
int pypo_hip_release_scratch(void *stream);

/* Precision used by the drop-in PO entry points (default PYPO_PREC_MIXED).
 * Returns the previous value. */
int pypo_hip_set_precision(int precision);

/* Maximum Newton iterations per ray (default 50; the reference loops until
 * converged).  Returns the previous value. */
int pypo_hip_set_rt_max_iter(int max_iter);

/* Device-resident PO propagation.  All pointers are device pointers; the
 * structs themselves live in host memory.  Grids must already be generated
 * (pypo_hip_generate_grid_dev).  Source currents: c2Bundle (or arrC1 for
 * the scalar mode).  res is c2Bundle (JM, EH, FF), c4Bundle (JMEH),
 * c2rBundle (EHP) or arrC1 (scalar) according to mode.  gmode/ncx/ncy are
 * the source grid's gmode and n_cells (trapezoid weights,
 * Propagation.h:1123-1135).  stream is a hipStream_t; NULL means the
 * default (null) stream.  Asynchronous on `stream`, in that stream's order;
 * calls on different streams use different scratch buffers and are not
 * ordered against each other.  Returns 0 on success. */
int pypo_hip_propagate_dev(int mode, void *res, const reflcontainer *cs,
                           const reflcontainer *ct, const void *currents,
                           int gmode, int ncx, int ncy, double k, double epsilon,
                           double t_direction, int precision, void *stream);

/* Size of the device scratch of pypo_hip_propagate_noalloc: 128 MiB per
 * device, held from the library module's load until the process ends. */
#define PYPO_HIP_NOALLOC_SCRATCH_BYTES (128u * 1024u * 1024u)

/* pypo_hip_propagate_dev without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate (serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, results and ordering as
 * pypo_hip_propagate_dev; the packed sources and split partial sums live in
 * the library's static device scratch (PYPO_HIP_NOALLOC_SCRATCH_BYTES per
 * device) instead of its grow-only pool, with as many source splits as fit
 * (bit-identical to pypo_hip_propagate_dev whenever its split plan fits).
 * A problem whose packed sources plus one split of partials exceed the
 * scratch is rejected with ERR_ARG (1001) before anything runs.  Every call
 * on a device shares its scratch, which keeps the last call's contents: the
 * library orders these calls with one completion event per device, so calls
 * from several threads or streams run one after another; replays of a
 * captured call are outside that ordering and race with any other no-alloc
 * work on the device that overlaps them.  Calls no hipMalloc/hipFree; the
 * first call on a device creates that device's completion event (and the
 * first library call of a process the library's stream). */
int pypo_hip_propagate_noalloc(int mode, void *res, const reflcontainer *cs,
                               const reflcontainer *ct, const void *currents,
                               int gmode, int ncx, int ncy, double k, double epsilon,
                               double t_direction, int precision, void *stream);

/* Device-resident grid generation into device arrays of container. */
int pypo_hip_generate_grid_dev(reflparams refl, reflcontainer *container, void *stream);

/* Device-resident ray trace through n_surf surfaces in sequence: frames[0]
 * is the input (device arrays, left unchanged), frames[i+1] receives the
 * result on surface i.  tol/t0 as in propagateRays.  Returns 0 on success. */
int pypo_hip_trace_dev(const reflparams *surfs, int n_surf, cframe *frames,
                       double tol, double t0, void *stream);

/* ---- batched sweeps ---------------------------------------------------- */

/* A sweep of n_inst JM -> FF instances (e.g. frequencies x feed offsets) of
 * uneven grid sizes sharing one far-field grid.  Per instance: feed surface,
 * dish surface, wavenumber k (1/mm) and the feed currents (host arrays of the
 * feed grid's size).  The JM step (feed -> dish, PEC currents) and the FF
 * step (dish -> ff) of every instance run as one launch each over a static
 * work list sorted by cost (a run: these two, a finish kernel after each and
 * one packing kernel between them); mixed precision.  tpl (targets per lane:
 * 1 or 2, or 4 = 16 M-blocks of the matrix-core far field at one wave per
 * SIMD), chunk_tiles (256-source tiles per work item) and wgs_per_cu
 * (workgroups per CU in each launch, 1..4096, never more than the stage's
 * work items) shape the launch; 0 picks defaults (2, whole feeds/dishes,
 * 64).
 * use_graph captures the whole run into a hipGraph that pypo_hip_sweep_run
 * replays.  No counterpart in the reference (its batch loops are Python). */
typedef struct pypo_sweep_desc {
    int n_inst;
    const reflparams *feeds;
    const reflparams *dishes;
    const double *k;
    const c2Bundle *feed_currents;
    reflparams ff;
    double t_direction;
    int tpl, chunk_tiles, wgs_per_cu, use_graph;
} pypo_sweep_desc;

int pypo_hip_sweep_create(const pypo_sweep_desc *desc, void **handle);
/* Rebuild the work lists (and graph) with new launch shapes. */
int pypo_hip_sweep_plan(void *handle, int tpl, int chunk_tiles, int wgs_per_cu, int use_graph);
/* Run (replay) the sweep on `stream` (NULL: default stream); asynchronous. */
int pypo_hip_sweep_run(void *handle, void *stream);
/* Copy instance inst's far field (12 arrays of the ff grid size) / dish
 * currents (12 arrays of the dish size) to host; synchronous. */
int pypo_hip_sweep_read(void *handle, int inst, c2Bundle *ff_out);
int pypo_hip_sweep_read_currents(void *handle, int inst, c2Bundle *dish_currents);
/* Pair counts, work-list lengths and the device pointer of the far-field
 * results ([n_inst][12][n_ff] doubles, c2Bundle row order). */
int pypo_hip_sweep_info(void *handle, double *pairs_jm, double *pairs_ff, int *items_jm,
                        int *items_ff, double **ff_out_dev);
int pypo_hip_sweep_destroy(void *handle);

/* ---- adjoint of the far-field power (no counterpart in the reference) ----
 * Surface-shape gradients for reflector optimisation.  The dish is
 * perturbed as the reference perturbs it for surface errors
 * (InterfaceReflector.cpp:852-866): points move along their normals,
 * normals and area elements stay.  Forward: feed currents -> JM on the
 * displaced dish -> FF; P_t = |E_t|^2.
 *
 * out = base with each point moved by d[i] along its normal (device
 * arrays; base->size points). */
int pypo_hip_displace_dev(const reflcontainer *base, const double *d, reflcontainer *out,
                          void *stream);
/* grad[i] = dL/dd[i] for a loss L of the far-field power with
 * dl_dp[t] = dL/dP_t.  dish: the displaced dish grid (gmode/ncx/ncy its
 * n_cells); ff_fields: the FF result, on the ff grid, of the JM currents
 * that feed_currents induce on the dish, both computed with the same k and
 * t_direction.  The dish currents and their derivatives are recomputed
 * here from the feed, not read.  All arrays device memory; precision
 * PYPO_PREC_FP64 runs every sum in FP64, PYPO_PREC_MIXED keeps FP64 for
 * distances and phases and runs the integrands in FP32.  Asynchronous on
 * `stream`, ordered like pypo_hip_propagate_dev. */
int pypo_hip_ff_power_grad_dev(const reflcontainer *dish, int gmode, int ncx, int ncy,
                               const reflcontainer *feed,
                               int feed_gmode, int feed_ncx, int feed_ncy,
                               const c2Bundle *feed_currents, const reflcontainer *ff,
                               const c2Bundle *ff_fields, const double *dl_dp, double k,
                               double t_direction, int precision, double *grad, void *stream);

/* Device-resident beam initialisation on a grid already in device memory
 * (pypo_hip_generate_grid_dev); all bundle and frame arrays are device
 * pointers; asynchronous on `stream` (NULL: the default stream) except
 * pypo_hip_make_grt_frame_dev, which synchronises to check its acceptance
 * flag.  Same formulas as the drop-in beam calls above. */
int pypo_hip_make_gauss_dev(GPODict gdict, const reflcontainer *grid, c2Bundle *field,
                            c2Bundle *current, void *stream);
int pypo_hip_make_gauss_beam_dev(vecGPODict gdict, const reflcontainer *grid, c2Bundle *field,
                                 c2Bundle *current, void *stream);
int pypo_hip_make_scalar_gauss_dev(ScalarGPODict sgdict, const reflcontainer *grid, arrC1 *field,
                                   void *stream);
int pypo_hip_calc_currents_dev(const c2Bundle *field, c2Bundle *current, const reflcontainer *grid,
                               int mode, void *stream);
int pypo_hip_make_rt_frame_dev(RTDict rdict, cframe *fr, void *stream);
int pypo_hip_make_grt_frame_dev(GRTDict grdict, cframe *fr, void *stream);

/* Hybrid ray-trace / PO step (System.runHybridPropagation, System.py:1570-1622)
 * on device arrays: trace fr_in onto `target` into fr_out (as
 * pypo_hip_trace_dev with one surface), then carry the field samples along
 * the rays: field_out = field_in * exp(j k L1), times sqrt(L0 / (2 L1 + L0))
 * when `start` (3 doubles, host memory) is not NULL; L1 = |p_out - p_in|,
 * L0 = |p_in - start| (calcRayLen, System.py:1304-1372).  Asynchronous. */
int pypo_hip_hybrid_dev(reflparams target, const cframe *fr_in, cframe *fr_out,
                        const c2Bundle *field_in, c2Bundle *field_out, const double *start,
                        double k, double tol, double t0, void *stream);

/* Nearest-neighbour interpolation of the 12 arrays of `values`, defined on
 * the positions of `points`, onto the positions of `grid` (interpFrame with
 * method "nearest", System.py:1624-1692, which uses scipy's griddata):
 * out[q] = values[the point nearest to grid point q]; ties go to the smallest
 * point index.  Device arrays; asynchronous. */
int pypo_hip_interp_nearest_dev(const cframe *points, const c2Bundle *values,
                                const reflcontainer *grid, c2Bundle *out, void *stream);

/* Launch shape of the single-propagation PO sums (pypo_hip_propagate_dev and
 * the drop-in symbols): targets per lane of the far-field kernels (1 or 2)
 * and the workgroup count the source splits aim for.  0 restores the
 * built-in choice (far field: 2 targets per lane above 256 targets, 8192
 * workgroups; other modes: 1 target per lane, 2048 workgroups).  A tuning hook; results do not depend on it beyond
 * rounding.  Returns 0, or ERR_ARG for values out of range. */
int pypo_hip_set_sum_shape(int tpl, int want_wgs);

/* Kernel of the mixed-precision far-field sum: PYPO_FF_MATRIX (default) runs
 * the source products on the matrix cores (FP16 hi/lo operand split, FP32
 * accumulation), PYPO_FF_VECTOR on the vector ALUs (FP32 FMAs).  Both fold
 * per-tile sums into FP64 and meet the same accuracy bar; a tuning and
 * cross-check hook.  Returns 0, or ERR_ARG. */
/* Fault injection for tests of the error path: the next n device-buffer
 * requests fail as a failed hipMalloc would (ERR_ALLOC recorded, the call
 * returns it; nothing is allocated or freed).  The environment variable
 * PYPO_HIP_FAIL_ALLOCS sets n at initialisation.  0 disables.  No
 * counterpart in the reference, which has no error channel. */
int pypo_hip_inject_alloc_failures(int n);

#define PYPO_FF_MATRIX 0
#define PYPO_FF_VECTOR 1
int pypo_hip_set_ff_kernel(int kind);

/* ---- Figures of merit on device data ---------------------------------------
 * Reductions behind src/PyPO/Efficiencies.py, for fields and frames that are
 * already in device memory (none has a C counterpart in the reference, whose
 * efficiencies are numpy code).  Each writes a few doubles to HOST memory and
 * returns after the stream has finished them.  FP64, deterministic order.
 *
 * Aperture of Efficiencies._generateMask (Efficiencies.py:11-41), the aperDict
 * keys "shape", "center", "outer", "outer_x", "outer_y".  Ellipse: a point
 * (x, y) of the surface's rest-frame grid is inside when
 * (x-cx)^2 + (y-cy)^2 < (outer0 cos t)^2 + (outer1 sin t)^2, t = atan2(y, x) + pi;
 * rectangle: outer_x0 < x-cx < outer_x1 and outer_y0 < y-cy < outer_y1 (open);
 * NONE selects every point.  The reference's "inner" keys select nothing. */
#define PYPO_APER_ELLIPSE 0
#define PYPO_APER_RECT 1
#define PYPO_APER_NONE 2
typedef struct {
    int shape;
    double center[2];
    double outer[2];
    double outer_x[2];
    double outer_y[2];
} aperparams;

/* sums[5] = { sum |f|^2 dA, sum_in |f|^2 dA, sum_in Re f dA, sum_in Im f dA, sum_in dA }
 * over the grid (rest frame: pypo_hip_generate_grid_dev with an identity
 * transf) and one complex field component f = f_re + j f_im.  Serves
 * Efficiencies.calcSpillover (90-115: s1 / s0) and calcTaper (117-145:
 * (s2^2 + s3^2) / s1 / s4). */
int pypo_hip_aperture_sums_dev(const reflcontainer *grid, const double *f_re, const double *f_im,
                               aperparams aper, double *sums, void *stream);

/* sums[2] = { sum |a|^2, sum |b|^2 } over n points: Efficiencies.calcXpol
 * (147-159) gives 1 - sums[1] / (sums[0] + sums[1]). */
int pypo_hip_power_sums_dev(const double *a_re, const double *a_im, const double *b_re, const double *b_im,
                            int n, double *sums, void *stream);

/* stats[8] over the rays of a frame with |d|^2 > 0.8: { count, sum x, sum y,
 * sum z, sum dx, sum dy, sum dz, sum |p - c|^2 } with c the mean position.
 * Efficiencies.calcRTcenter (43-57) = sums/count, calcRTtilt (59-73) = the
 * normalised mean direction, calcRMS (75-88) = sqrt(stats[7] / count). */
int pypo_hip_frame_stats_dev(const cframe *fr, double *stats, void *stream);

/* ---- Single-precision entry points of the reference's GPU library ----------
 * The reference's GPU library (src/CUDA/InterfaceCUDA.h:29-69, loaded by
 * BindGPU.loadGPUlib as libpypogpu) takes FP32 structs.  These wrappers keep
 * its names and argument types: FP32 inputs are widened to FP64, the grids
 * are generated and the sums run as in the FP64 entry points above (mixed
 * precision by default, see pypo_hip_set_precision), and the results are
 * narrowed to FP32.  The generated grids are written back to cs / ct like
 * the reference's.  nBlocks and nThreads are launch hints and are ignored.
 * Structs are byte-identical to Structs.h:216-362. */
typedef struct arrC1f { float *x, *y; } arrC1f;
typedef struct c2Bundlef {
    float *r1x, *r1y, *r1z, *i1x, *i1y, *i1z;
    float *r2x, *r2y, *r2z, *i2x, *i2y, *i2z;
} c2Bundlef;
typedef struct c4Bundlef {
    float *r1x, *r1y, *r1z, *i1x, *i1y, *i1z;
    float *r2x, *r2y, *r2z, *i2x, *i2y, *i2z;
    float *r3x, *r3y, *r3z, *i3x, *i3y, *i3z;
    float *r4x, *r4y, *r4z, *i4x, *i4y, *i4z;
} c4Bundlef;
typedef struct c2rBundlef {
    float *r1x, *r1y, *r1z, *i1x, *i1y, *i1z;
    float *r2x, *r2y, *r2z, *i2x, *i2y, *i2z;
    float *r3x, *r3y, *r3z;
} c2rBundlef;
typedef struct reflparamsf {
    float *coeffs;
    float *lxu;
    float *lyv;
    int *n_cells;
    bool flip;
    int gmode;
    float *gcenter;
    float ecc_uv;
    float rot_uv;
    int type;
    float *transf;
    float rms;
    unsigned int rms_seed;
} reflparamsf;
typedef struct reflcontainerf { int size; float *x, *y, *z, *nx, *ny, *nz, *area; } reflcontainerf;
typedef struct cframef { int size; float *x, *y, *z, *dx, *dy, *dz; } cframef;

/* InterfaceCUDA.h:31-59: PO into FP32 host bundles (t_direction -1 fwd, +1 bwd). */
void callKernelf_JM(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                    reflcontainerf *ct, c2Bundlef *currents, float k, float epsilon, float t_direction,
                    int nBlocks, int nThreads);
void callKernelf_EH(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                    reflcontainerf *ct, c2Bundlef *currents, float k, float epsilon, float t_direction,
                    int nBlocks, int nThreads);
void callKernelf_JMEH(c4Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                      reflcontainerf *ct, c2Bundlef *currents, float k, float epsilon, float t_direction,
                      int nBlocks, int nThreads);
void callKernelf_EHP(c2rBundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                     reflcontainerf *ct, c2Bundlef *currents, float k, float epsilon, float t_direction,
                     int nBlocks, int nThreads);
void callKernelf_FF(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                    reflcontainerf *ct, c2Bundlef *currents, float k, float epsilon, float t_direction,
                    int nBlocks, int nThreads);
/* InterfaceCUDA.h:61-64 */
void callKernelf_scalar(arrC1f *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                        reflcontainerf *ct, arrC1f *inp, float k, float epsilon, float t_direction,
                        int nBlocks, int nThreads);
/* InterfaceCPU.h:64-101 (InterfaceCPU.cpp:265-455): the reference CPU library's
 * FP32 interface, for C callers of that library.  Same FP32 path as
 * callKernelf_* (grids generated into cs / ct, mixed-precision sums, FP32
 * results); numThreads is the CPU thread count there and is ignored here. */
void propagateToGridf_JM(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                         reflcontainerf *ct, c2Bundlef *currents, float k, int numThreads, float epsilon,
                         float t_direction);
void propagateToGridf_EH(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                         reflcontainerf *ct, c2Bundlef *currents, float k, int numThreads, float epsilon,
                         float t_direction);
void propagateToGridf_JMEH(c4Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                           reflcontainerf *ct, c2Bundlef *currents, float k, int numThreads, float epsilon,
                           float t_direction);
void propagateToGridf_EHP(c2rBundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                          reflcontainerf *ct, c2Bundlef *currents, float k, int numThreads, float epsilon,
                          float t_direction);
void propagateToGridf_scalar(arrC1f *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                             reflcontainerf *ct, arrC1f *field, float k, int numThreads, float epsilon,
                             float t_direction);
void propagateToFarFieldf(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                          reflcontainerf *ct, c2Bundlef *currents, float k, int numThreads, float epsilon,
                          float t_direction);
/* InterfaceCUDA.h:66-68: ray trace (epsilon = Newton tolerance). */
void callRTKernel(reflparamsf ctp, cframef *fr_in, cframef *fr_out, float epsilon, float t0, int nBlocks,
                  int nThreads);

/* Mean kernel time (ms) of the last pypo_hip_propagate_dev / trace call's
 * dominant kernel, measured with HIP events on the launch stream, when
 * timing is enabled by pypo_hip_set_timing(1).  Returns -1 if unavailable. */
int pypo_hip_set_timing(int enable);
double pypo_hip_last_kernel_ms(void);

#ifdef __cplusplus
}
#endif

#endif /* PYPO_HIP_H */
