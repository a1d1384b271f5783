// abi.hip -- the C ABI of libpypohip.so: drop-in FP64 entry points with the
// reference's names and argument lists (src/CPU/InterfaceCPU.h:24-64) and the
// device-resident pypo_hip_* extensions.  Host orchestration only: buffer
// pool, library stream, error channel, host-side random deviates.

#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <vector>

#include <algorithm>
#include <cstdlib>

#include "common.hpp"

namespace pypo {

// ---- error channel -----------------------------------------------------------
namespace {
thread_local int t_err_code = 0;
thread_local std::string t_err_msg;
}  // namespace

void set_error(int code, const std::string &msg) {
    if (t_err_code == 0) {   // keep the first failure of a call
        t_err_code = code;
        t_err_msg = msg;
    }
}

}  // namespace pypo

// Device scratch of pypo_hip_propagate_noalloc: the packed source records,
// their tile metadata and the split partial sums of one propagation.  Each
// device the library runs on holds its own copy from the module's load until
// the process ends.  In host code `scratch` names only a stand-in for the
// device symbol: its device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[PYPO_HIP_NOALLOC_SCRATCH_BYTES];

namespace pypo {

// ---- library state ------------------------------------------------------------
namespace {

std::recursive_mutex g_mu;         // one library call at a time on the host

// Scratch buffers of the calls enqueued on one stream.  Calls on one stream
// run in that stream's order and share its set; calls on different streams
// use different sets and need no ordering against each other, so
// device-resident work on several streams overlaps.
struct PoolSet {
    void *buf[SLOT_COUNT] = {};
    size_t bytes[SLOT_COUNT] = {};
    hipEvent_t done = nullptr;     // completion of the last call that used the set
};
std::map<hipStream_t, PoolSet> g_sets;   // nodes never move: pointers below stay valid
PoolSet *g_cur = nullptr;          // set of the call being enqueued
hipStream_t g_cur_st = nullptr;
PoolSet *g_open = nullptr;         // set of a call that stopped before end_call
hipStream_t g_open_st = nullptr;
bool g_async_alloc = false;        // stream-ordered allocation (hipMallocAsync) available
bool g_capturing = false;          // the call being enqueued is inside a stream capture
hipStream_t g_stream = nullptr;
int g_fail_allocs = 0;              // fault injection: pool requests left to fail
int g_precision = PYPO_PREC_MIXED;
int g_rt_max_iter = 50;
bool g_timing = false;
double g_last_ms = -1.0;

int ensure_init() {
    if (g_stream) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error(ERR_NODEV, "no HIP device visible");
        return ERR_NODEV;
    }
    int dev = 0, pools = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&pools, hipDeviceAttributeMemoryPoolsSupported, dev) == hipSuccess)
        g_async_alloc = pools != 0;
    PYPO_HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    if (const char *v = std::getenv("PYPO_HIP_FAIL_ALLOCS")) g_fail_allocs = std::max(0, std::atoi(v));
    return 0;
}

// Start a call on stream st: select st's buffer set.  A call that stopped on
// an error after enqueuing work never reached end_call; its set's completion
// is marked on its stream now.  Waiting on the set's last completion orders
// this call after every earlier user of the set, which for the same stream is
// already the case (it matters when a new stream reuses a destroyed stream's
// handle).
int begin_call(hipStream_t st) {
    if (g_open) {
        (void)hipEventRecord(g_open->done, g_open_st);
        g_open = nullptr;
    }
    PoolSet &ps = g_sets[st];
    if (!ps.done) PYPO_HIP_TRY(hipEventCreateWithFlags(&ps.done, hipEventDisableTiming));
    // Inside a stream capture the graph's own order stands in for the
    // events: an event recorded outside the capture cannot be waited on in
    // it, nor one recorded in it outside.  A captured call must find its
    // buffers grown already (pool_get refuses to grow during a capture).
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    g_capturing = hipStreamIsCapturing(st, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone;
    if (!g_capturing) PYPO_HIP_TRY(hipStreamWaitEvent(st, ps.done, 0));
    g_cur = &ps;
    g_cur_st = st;
    if (!g_capturing) {
        g_open = &ps;
        g_open_st = st;
    }
    return 0;
}

int end_call(hipStream_t st) {
    if (!g_capturing) PYPO_HIP_TRY(hipEventRecord(g_cur->done, st));
    g_open = nullptr;
    return 0;
}

}  // namespace

void record_kernel_ms(double ms) { g_last_ms = ms; }

hipStream_t lib_stream(int *err) {
    *err = ensure_init();
    return g_stream;
}

void *pool_get(int slot, size_t bytes, int *err) {
    *err = 0;
    if (g_fail_allocs > 0) {
        g_fail_allocs--;
        set_error(ERR_ALLOC, "injected allocation failure (pypo_hip_inject_alloc_failures)");
        *err = ERR_ALLOC;
        return nullptr;
    }
    PoolSet &ps = *g_cur;
    if (bytes == 0) bytes = 256;
    if (ps.bytes[slot] >= bytes) return ps.buf[slot];
    const size_t want = bytes + bytes / 8;   // headroom for slightly larger calls
    hipError_t e;
    if (g_capturing) {
        set_error(ERR_ARG, "a scratch buffer would have to grow inside a stream capture: run the "
                           "same call once on this stream before capturing it");
        *err = ERR_ARG;
        return nullptr;
    }
    if (g_async_alloc) {
        // Stream order frees the old buffer after the earlier calls on this
        // set are done with it and makes the new one ready for what follows;
        // nothing else on the device waits.
        if (ps.buf[slot]) (void)hipFreeAsync(ps.buf[slot], g_cur_st);
        ps.buf[slot] = nullptr;
        ps.bytes[slot] = 0;
        e = hipMallocAsync(&ps.buf[slot], want, g_cur_st);
    } else {
        if (ps.buf[slot]) {
            // earlier calls on this set may still read the old buffer
            (void)hipEventSynchronize(ps.done);
            (void)hipStreamSynchronize(g_cur_st);
            (void)hipFree(ps.buf[slot]);
        }
        ps.buf[slot] = nullptr;
        ps.bytes[slot] = 0;
        e = hipMalloc(&ps.buf[slot], want);
    }
    if (e != hipSuccess) {
        ps.buf[slot] = nullptr;
        set_error(ERR_ALLOC, std::string("allocation of ") + std::to_string(want) +
                                 " device bytes failed: " + hipGetErrorString(e));
        *err = ERR_ALLOC;
        return nullptr;
    }
    ps.bytes[slot] = want;
    return ps.buf[slot];
}

namespace {

// Surface-error deviates of add_surferr_uncorr (InterfaceReflector.cpp:852-866):
// a new normal_distribution per point, so each draw starts the polar method
// afresh (Random.h:111-115).
std::vector<double> grid_deviates(const reflparams &r, int n) {
    std::vector<double> e(n);
    std::mt19937 gen(r.rms_seed);
    for (int i = 0; i < n; i++) {
        std::normal_distribution<double> d(0.0, r.rms);
        e[i] = d(gen);
    }
    return e;
}

// Ray surface errors of RayTracer::parallelRays (RayTrace.h:269-274): one
// distribution for all rays (Random.h:124-137).
std::vector<double> ray_deviates(const reflparams &r, int n) {
    std::vector<double> e(n);
    std::mt19937 gen(r.rms_seed);
    std::normal_distribution<double> d(0.0, r.rms);
    for (int i = 0; i < n; i++) e[i] = d(gen);
    return e;
}

GridPtrs carve_grid(double *base, int n) {
    return GridPtrs{base, base + n, base + 2 * (size_t)n, base + 3 * (size_t)n,
                    base + 4 * (size_t)n, base + 5 * (size_t)n, base + 6 * (size_t)n};
}

int check_refl(const reflparams &r, int n, const char *what) {
    if (!r.n_cells || !r.lxu || !r.lyv || !r.gcenter || !r.coeffs || !r.transf) {
        set_error(ERR_ARG, std::string(what) + ": reflparams has a NULL array");
        return ERR_ARG;
    }
    if ((long long)r.n_cells[0] * r.n_cells[1] != n || n <= 0) {
        set_error(ERR_ARG, std::string(what) + ": n_cells[0]*n_cells[1] != container size");
        return ERR_ARG;
    }
    return 0;
}

int check_container(const reflcontainer *c, const char *what) {
    if (!c || !c->x || !c->y || !c->z || !c->nx || !c->ny || !c->nz || !c->area) {
        set_error(ERR_ARG, std::string(what) + ": NULL reflcontainer or array");
        return ERR_ARG;
    }
    return 0;
}

// Grid into device memory (with surface errors when rms > 0).
int device_grid(const reflparams &r, const GridPtrs &g, int n, int err_slot, bool transform,
                bool spheric, hipStream_t st) {
    const double *d_errs = nullptr;
    if (r.rms > 0) {
        std::vector<double> e = grid_deviates(r, n);
        int err = 0;
        double *d = (double *)pool_get(err_slot, n * sizeof(double), &err);
        if (err) return err;
        PYPO_HIP_TRY(hipMemcpyAsync(d, e.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
        PYPO_HIP_TRY(hipStreamSynchronize(st));   // e is a temporary
        d_errs = d;
    }
    return launch_generate_grid(r, g, n, d_errs, transform, spheric, st);
}

int copy_grid_to_host(const GridPtrs &g, int n, reflcontainer *c, hipStream_t st) {
    double *src[7] = {g.x, g.y, g.z, g.nx, g.ny, g.nz, g.area};
    double *dst[7] = {c->x, c->y, c->z, c->nx, c->ny, c->nz, c->area};
    for (int q = 0; q < 7; q++)
        PYPO_HIP_TRY(hipMemcpyAsync(dst[q], src[q], n * sizeof(double), hipMemcpyDeviceToHost, st));
    return 0;
}

int n_result_arrays(int mode) {
    switch (mode) {
        case PYPO_MODE_JMEH: return 24;
        case PYPO_MODE_EHP: return 15;
        case PYPO_MODE_SCALAR: return 2;
        default: return 12;
    }
}

// Pointer list of a result/current bundle in field order (struct layout order).
void bundle_ptrs(int mode, void *b, double **out) {
    double **p = (double **)b;   // every bundle is a plain sequence of double*
    int n = n_result_arrays(mode);
    for (int q = 0; q < n; q++) out[q] = p[q];
}

// The drop-in PO call: host buffers in, host buffers out, synchronous.
int po_dropin(int mode, void *res, const reflparams &sp, const reflparams &tp, reflcontainer *cs,
              reflcontainer *ct, const void *cur, double k, double epsilon, double tdir) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(res && cur, ERR_ARG, "NULL result or source-current bundle");
    if ((err = check_container(cs, "source")) || (err = check_container(ct, "target"))) return err;
    const int gs = cs->size, gt = ct->size;
    if ((err = check_refl(sp, gs, "source")) || (err = check_refl(tp, gt, "target"))) return err;
    const int nres = n_result_arrays(mode);
    double *hres[24];
    bundle_ptrs(mode, res, hres);
    for (int q = 0; q < nres; q++) PYPO_REQUIRE(hres[q], ERR_ARG, "result bundle has a NULL array");
    const int ncur = (mode == PYPO_MODE_SCALAR) ? 2 : 12;
    double *hcur[12];
    {
        double **p = (double **)cur;
        for (int q = 0; q < ncur; q++) {
            hcur[q] = p[q];
            PYPO_REQUIRE(hcur[q], ERR_ARG, "source bundle has a NULL array");
        }
    }
    hipStream_t st = g_stream;
    if ((err = begin_call(st))) return err;

    double *dsg = (double *)pool_get(SLOT_SRC_GRID, 7 * (size_t)gs * sizeof(double), &err);
    if (err) return err;
    double *dtg = (double *)pool_get(SLOT_TGT_GRID, 7 * (size_t)gt * sizeof(double), &err);
    if (err) return err;
    GridPtrs gsrc = carve_grid(dsg, gs), gtgt = carve_grid(dtg, gt);
    if ((err = device_grid(sp, gsrc, gs, SLOT_ERRS_S, true, true, st))) return err;
    if ((err = device_grid(tp, gtgt, gt, SLOT_ERRS_T, true, true, st))) return err;

    double *dcur = (double *)pool_get(SLOT_CURRENTS, ncur * (size_t)gs * sizeof(double), &err);
    if (err) return err;
    for (int q = 0; q < ncur; q++)
        PYPO_HIP_TRY(hipMemcpyAsync(dcur + q * (size_t)gs, hcur[q], gs * sizeof(double),
                                    hipMemcpyHostToDevice, st));

    const int prec = g_precision;
    void *packed = pool_get(SLOT_PACKED, packed_bytes(mode, prec, gs), &err);
    if (err) return err;
    if (mode == PYPO_MODE_SCALAR) {
        err = launch_pack_scalar(prec, gsrc, gs, dcur, dcur + gs, sp.gmode, sp.n_cells[0],
                                 sp.n_cells[1], packed, st);
    } else {
        Bundle12 b;
        for (int q = 0; q < 12; q++) b.p[q] = dcur + q * (size_t)gs;
        err = launch_pack_sources(prec, gsrc, gs, b, sp.gmode, sp.n_cells[0], sp.n_cells[1], packed,
                                  st);
    }
    if (err) return err;

    SumPlan plan = plan_sum(mode, prec, gs, gt);
    size_t wsb = partial_bytes(mode, plan, gt);
    double *ws = (double *)pool_get(SLOT_PARTIAL, wsb, &err);
    if (err) return err;
    double *dres = (double *)pool_get(SLOT_RESULT, nres * (size_t)gt * sizeof(double), &err);
    if (err) return err;
    double *dptr[24];
    for (int q = 0; q < nres; q++) dptr[q] = dres + q * (size_t)gt;
    // device-side bundle with the same layout as the host one
    c4Bundle dev_b;
    std::memset(&dev_b, 0, sizeof(dev_b));
    std::memcpy(&dev_b, dptr, nres * sizeof(double *));
    err = launch_po(mode, prec, packed, gs, gtgt, gt, k, epsilon, tdir, &dev_b, ws, wsb, st, g_timing);
    if (err) return err;

    for (int q = 0; q < nres; q++)
        PYPO_HIP_TRY(hipMemcpyAsync(hres[q], dptr[q], gt * sizeof(double), hipMemcpyDeviceToHost, st));
    if ((err = copy_grid_to_host(gsrc, gs, cs, st))) return err;
    if ((err = copy_grid_to_host(gtgt, gt, ct, st))) return err;
    if ((err = end_call(st))) return err;
    PYPO_HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

// ---- beam initialisation into host buffers ------------------------------------------
enum BeamKind { BEAM_GAUSS, BEAM_VEC_GAUSS, BEAM_SCALAR, BEAM_CURRENTS };

// Generate `plane`'s grid on the device, run one beam kernel, copy the
// outputs (12 field + 12 current arrays, or 2 scalar arrays, or 12 current
// arrays) into the host arrays.  h_in: the 12 field arrays of calcCurrents.
int beam_dropin(int kind, const void *params, const reflparams &plane, double *const *h_out, int n_out,
                double *const *h_in, int mode) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(plane.n_cells, ERR_ARG, "beam: reflparams has a NULL n_cells");
    const long long nn = (long long)plane.n_cells[0] * plane.n_cells[1];
    PYPO_REQUIRE(nn > 0 && nn < (1LL << 31), ERR_ARG, "beam: grid size out of range");
    const int n = (int)nn;
    if ((err = check_refl(plane, n, "beam"))) return err;
    for (int q = 0; q < n_out; q++) PYPO_REQUIRE(h_out[q], ERR_ARG, "beam: result bundle has a NULL array");
    if (h_in)
        for (int q = 0; q < 12; q++) PYPO_REQUIRE(h_in[q], ERR_ARG, "beam: field bundle has a NULL array");
    hipStream_t st = g_stream;
    if ((err = begin_call(st))) return err;
    double *dg = (double *)pool_get(SLOT_SRC_GRID, 7 * (size_t)n * sizeof(double), &err);
    if (err) return err;
    GridPtrs g = carve_grid(dg, n);
    if ((err = device_grid(plane, g, n, SLOT_ERRS_S, true, true, st))) return err;
    double *dout = (double *)pool_get(SLOT_RESULT, (size_t)n_out * n * sizeof(double), &err);
    if (err) return err;
    Bundle12 o1{}, o2{};
    for (int q = 0; q < 12 && q < n_out; q++) o1.p[q] = dout + (size_t)q * n;
    for (int q = 12; q < n_out; q++) o2.p[q - 12] = dout + (size_t)q * n;
    switch (kind) {
        case BEAM_GAUSS: err = launch_gauss(*(const GPODict *)params, g, n, o1, o2, st); break;
        case BEAM_VEC_GAUSS: err = launch_gauss_beam(*(const vecGPODict *)params, g, n, o1, o2, st); break;
        case BEAM_SCALAR:
            err = launch_scalar_gauss(*(const ScalarGPODict *)params, g, n, o1.p[0], o1.p[1], st);
            break;
        default: {
            double *din = (double *)pool_get(SLOT_CURRENTS, 12 * (size_t)n * sizeof(double), &err);
            if (err) return err;
            Bundle12 fin{};
            for (int q = 0; q < 12; q++) {
                fin.p[q] = din + (size_t)q * n;
                PYPO_HIP_TRY(hipMemcpyAsync(fin.p[q], h_in[q], n * sizeof(double), hipMemcpyHostToDevice, st));
            }
            err = launch_calc_jm(fin, g, n, mode, o1, st);
        }
    }
    if (err) return err;
    for (int q = 0; q < n_out; q++)
        PYPO_HIP_TRY(hipMemcpyAsync(h_out[q], dout + (size_t)q * n, n * sizeof(double),
                                    hipMemcpyDeviceToHost, st));
    if ((err = end_call(st))) return err;
    PYPO_HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

uint64_t grt_key(int seed) {
    if (seed != -1) return 0x9E3779B97F4A7C15ull ^ (uint64_t)(uint32_t)seed;
    std::random_device rd;
    return ((uint64_t)rd() << 32) | rd();
}

// Ray frames into device arrays (the frame struct holds device pointers).
int frame_dev(bool gauss, const void *params, const cframe &fr, hipStream_t st) {
    if (!gauss) return launch_rt_frame(*(const RTDict *)params, fr, st);
    const GRTDict &g = *(const GRTDict *)params;
    int err = 0;
    int *flag = (int *)pool_get(SLOT_FLAG, 256, &err);
    if (err) return err;
    PYPO_HIP_TRY(hipMemsetAsync(flag, 0, sizeof(int), st));
    if ((err = launch_grt_frame(g, grt_key(g.seed), fr, flag, kGrtMaxTrials, st))) return err;
    int h = 0;
    PYPO_HIP_TRY(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    PYPO_HIP_TRY(hipStreamSynchronize(st));
    PYPO_REQUIRE(h == 0, ERR_ARG,
                 "makeGRTframe: some rays found no accepted candidate within 2^20 trials "
                 "(the Gaussian pdf peak 1/(pi^2 x0 y0 angx0 angy0) is too small)");
    return 0;
}

int frame_dropin(bool gauss, const void *params, int n, cframe *fr) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(fr && fr->x && fr->y && fr->z && fr->dx && fr->dy && fr->dz, ERR_ARG,
                 "ray frame: NULL cframe or array");
    PYPO_REQUIRE(n >= 1, ERR_ARG, "ray frame: no rays");
    hipStream_t st = g_stream;
    if ((err = begin_call(st))) return err;
    double *d = (double *)pool_get(SLOT_RAYS, 6 * (size_t)n * sizeof(double), &err);
    if (err) return err;
    cframe df{n, d, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n, d + 4 * (size_t)n, d + 5 * (size_t)n};
    if ((err = frame_dev(gauss, params, df, st))) return err;
    double *hs[6] = {fr->x, fr->y, fr->z, fr->dx, fr->dy, fr->dz};
    for (int q = 0; q < 6; q++)
        PYPO_HIP_TRY(hipMemcpyAsync(hs[q], d + (size_t)q * n, n * sizeof(double), hipMemcpyDeviceToHost, st));
    if ((err = end_call(st))) return err;
    PYPO_HIP_TRY(hipStreamSynchronize(st));
    fr->size = n;
    return 0;
}

int check_bundle12(const c2Bundle *b, const char *what) {
    if (!b) { set_error(ERR_ARG, std::string(what) + ": NULL bundle"); return ERR_ARG; }
    double *const *p = (double *const *)b;
    for (int q = 0; q < 12; q++)
        if (!p[q]) { set_error(ERR_ARG, std::string(what) + ": bundle has a NULL array"); return ERR_ARG; }
    return 0;
}

}  // namespace
}  // namespace pypo

using namespace pypo;

extern "C" {

int pypo_hip_last_error(char *buf, int n) {
    int code = t_err_code;
    if (buf && n > 0) {
        std::snprintf(buf, (size_t)n, "%s", code ? t_err_msg.c_str() : "");
    }
    t_err_code = 0;
    t_err_msg.clear();
    return code;
}

const char *pypo_hip_version(void) { return "pypo-hip 0.1 (gfx950)"; }

int pypo_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pypo_hip_stream_ordered_scratch(void) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return -err;
    return g_async_alloc ? 1 : 0;
}

long long pypo_hip_scratch_bytes(void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return -err;
    auto it = g_sets.find((hipStream_t)stream);
    if (it == g_sets.end()) return 0;
    long long total = 0;
    for (int s = 0; s < SLOT_COUNT; s++) total += (long long)it->second.bytes[s];
    return total;
}

int pypo_hip_release_scratch(void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    auto it = g_sets.find((hipStream_t)stream);
    if (it == g_sets.end()) return 0;
    PoolSet &ps = it->second;
    if (g_open == &ps) {   // a call on it stopped before end_call: mark its end first
        (void)hipEventRecord(ps.done, g_open_st);
        g_open = nullptr;
    }
    // Only the library's own event and stream are used: the caller's stream
    // may be gone.  After the event, no call outside a capture uses the set.
    PYPO_HIP_TRY(hipEventSynchronize(ps.done));
    for (int s = 0; s < SLOT_COUNT; s++) {
        if (!ps.buf[s]) continue;
        PYPO_HIP_TRY(g_async_alloc ? hipFreeAsync(ps.buf[s], g_stream) : hipFree(ps.buf[s]));
        ps.buf[s] = nullptr;
        ps.bytes[s] = 0;
    }
    if (g_async_alloc) PYPO_HIP_TRY(hipStreamSynchronize(g_stream));
    PYPO_HIP_TRY(hipEventDestroy(ps.done));
    if (g_cur == &ps) g_cur = nullptr;
    g_sets.erase(it);
    return 0;
}

int pypo_hip_set_precision(int precision) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int old = g_precision;
    if (precision == PYPO_PREC_MIXED || precision == PYPO_PREC_FP64) g_precision = precision;
    else set_error(ERR_ARG, "precision must be 0 (mixed) or 1 (fp64)");
    return old;
}

int pypo_hip_set_rt_max_iter(int max_iter) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int old = g_rt_max_iter;
    if (max_iter >= 1) g_rt_max_iter = max_iter;
    else set_error(ERR_ARG, "max_iter must be >= 1");
    return old;
}

int pypo_hip_set_timing(int enable) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int old = g_timing ? 1 : 0;
    g_timing = enable != 0;
    g_last_ms = -1.0;
    return old;
}

double pypo_hip_last_kernel_ms(void) { return g_last_ms; }

// ---- drop-in PO entry points (InterfaceCPU.h:24-58) ---------------------------

void propagateToGrid_JM(c2Bundle *res, reflparams source, reflparams target, reflcontainer *cs,
                        reflcontainer *ct, c2Bundle *currents, double k, int numThreads,
                        double epsilon, double t_direction) {
    (void)numThreads;
    po_dropin(PYPO_MODE_JM, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void propagateToGrid_EH(c2Bundle *res, reflparams source, reflparams target, reflcontainer *cs,
                        reflcontainer *ct, c2Bundle *currents, double k, int numThreads,
                        double epsilon, double t_direction) {
    (void)numThreads;
    po_dropin(PYPO_MODE_EH, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void propagateToGrid_JMEH(c4Bundle *res, reflparams source, reflparams target, reflcontainer *cs,
                          reflcontainer *ct, c2Bundle *currents, double k, int numThreads,
                          double epsilon, double t_direction) {
    (void)numThreads;
    po_dropin(PYPO_MODE_JMEH, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void propagateToGrid_EHP(c2rBundle *res, reflparams source, reflparams target, reflcontainer *cs,
                         reflcontainer *ct, c2Bundle *currents, double k, int numThreads,
                         double epsilon, double t_direction) {
    (void)numThreads;
    po_dropin(PYPO_MODE_EHP, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void propagateToGrid_scalar(arrC1 *res, reflparams source, reflparams target, reflcontainer *cs,
                            reflcontainer *ct, arrC1 *field, double k, int numThreads,
                            double epsilon, double t_direction) {
    (void)numThreads;
    po_dropin(PYPO_MODE_SCALAR, res, source, target, cs, ct, field, k, epsilon, t_direction);
}

void propagateToFarField(c2Bundle *res, reflparams source, reflparams target, reflcontainer *cs,
                         reflcontainer *ct, c2Bundle *currents, double k, int numThreads,
                         double epsilon, double t_direction) {
    (void)numThreads;
    po_dropin(PYPO_MODE_FF, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

// ---- drop-in ray trace (InterfaceCPU.h:60-61) -----------------------------------

// Host frames in and out, synchronous; returns 0 or the recorded error code.
static int rays_dropin(const reflparams &ctp, cframe *fr_in, cframe *fr_out, double epsilon, double t0) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(fr_in && fr_out && fr_in->size == fr_out->size && fr_in->size >= 0, ERR_ARG,
                 "propagateRays: NULL frame or frame sizes differ");
    PYPO_REQUIRE(ctp.coeffs && ctp.transf, ERR_ARG, "propagateRays: reflparams has a NULL array");
    const int n = fr_in->size;
    if (n == 0) return 0;
    double *hin[6] = {fr_in->x, fr_in->y, fr_in->z, fr_in->dx, fr_in->dy, fr_in->dz};
    double *hout[6] = {fr_out->x, fr_out->y, fr_out->z, fr_out->dx, fr_out->dy, fr_out->dz};
    for (int q = 0; q < 6; q++) PYPO_REQUIRE(hin[q] && hout[q], ERR_ARG, "propagateRays: frame has a NULL array");
    hipStream_t st = g_stream;
    if ((err = begin_call(st))) return err;
    double *d = (double *)pool_get(SLOT_RAYS, 12 * (size_t)n * sizeof(double), &err);
    if (err) return err;
    for (int q = 0; q < 6; q++)
        PYPO_HIP_TRY(hipMemcpyAsync(d + q * (size_t)n, hin[q], n * sizeof(double), hipMemcpyHostToDevice, st));
    const double *d_errs[1] = {nullptr};
    if (ctp.rms > 0) {
        std::vector<double> e = ray_deviates(ctp, n);
        double *de = (double *)pool_get(SLOT_RT_ERRS, n * sizeof(double), &err);
        if (err) return err;
        PYPO_HIP_TRY(hipMemcpyAsync(de, e.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
        PYPO_HIP_TRY(hipStreamSynchronize(st));
        d_errs[0] = de;
    }
    cframe fr[2];
    fr[0].size = n; fr[1].size = n;
    double **f0 = &fr[0].x, **f1 = &fr[1].x;
    for (int q = 0; q < 6; q++) {
        f0[q] = d + q * (size_t)n;
        f1[q] = d + (6 + q) * (size_t)n;
    }
    if ((err = launch_trace_chain(&ctp, d_errs, 1, fr, epsilon, t0, g_rt_max_iter, st, g_timing))) return err;
    for (int q = 0; q < 6; q++)
        PYPO_HIP_TRY(hipMemcpyAsync(hout[q], f1[q], n * sizeof(double), hipMemcpyDeviceToHost, st));
    if ((err = end_call(st))) return err;
    PYPO_HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

void propagateRays(reflparams ctp, cframe *fr_in, cframe *fr_out, int numThreads, double epsilon,
                   double t0) {
    (void)numThreads;
    (void)rays_dropin(ctp, fr_in, fr_out, epsilon, t0);
}

// ---- single-precision drop-ins (InterfaceCUDA.h:29-69) ------------------------------

// FP32 reflparams widened to FP64, with storage for the arrays it points to.
struct WideRefl {
    double coeffs[3], lxu[2], lyv[2], gcenter[2], transf[16];
    int n_cells[2];
    reflparams p;
};

static bool widen_refl(const reflparamsf &f, WideRefl &w, const char *what) {
    if (!f.coeffs || !f.lxu || !f.lyv || !f.n_cells || !f.gcenter || !f.transf) {
        set_error(ERR_ARG, std::string(what) + ": reflparamsf has a NULL array");
        return false;
    }
    for (int i = 0; i < 3; i++) w.coeffs[i] = f.coeffs[i];
    for (int i = 0; i < 2; i++) {
        w.lxu[i] = f.lxu[i]; w.lyv[i] = f.lyv[i]; w.gcenter[i] = f.gcenter[i]; w.n_cells[i] = f.n_cells[i];
    }
    for (int i = 0; i < 16; i++) w.transf[i] = f.transf[i];
    w.p.coeffs = w.coeffs; w.p.lxu = w.lxu; w.p.lyv = w.lyv; w.p.n_cells = w.n_cells;
    w.p.flip = f.flip; w.p.gmode = f.gmode; w.p.gcenter = w.gcenter;
    w.p.ecc_uv = f.ecc_uv; w.p.rot_uv = f.rot_uv; w.p.type = f.type;
    w.p.transf = w.transf; w.p.rms = f.rms; w.p.rms_seed = f.rms_seed;
    return true;
}

// n FP64 arrays of length len on the host, addressed as a pointer list.
struct HostArrays {
    std::vector<double> buf;
    double *ptr[24];
    HostArrays(int n, size_t len) : buf((size_t)n * len) {
        for (int q = 0; q < n; q++) ptr[q] = buf.data() + q * len;
    }
};

static bool float_arrays_ok(float *const *p, int n, const char *what) {
    for (int q = 0; q < n; q++)
        if (!p[q]) {
            set_error(ERR_ARG, std::string(what) + ": FP32 bundle has a NULL array");
            return false;
        }
    return true;
}

// FP32 PO call: widen, run po_dropin, narrow the result and the grids.
static void po_float(int mode, void *resf, const reflparamsf &sf, const reflparamsf &tf, reflcontainerf *csf,
                     reflcontainerf *ctf, const void *curf, float k, float epsilon, float tdir) {
    if (!resf || !curf || !csf || !ctf) {
        set_error(ERR_ARG, "callKernelf: NULL result, current or grid argument");
        return;
    }
    const int gs = csf->size, gt = ctf->size;
    if (gs <= 0 || gt <= 0) {
        set_error(ERR_ARG, "callKernelf: empty source or target grid");
        return;
    }
    const int nres = n_result_arrays(mode), ncur = mode == PYPO_MODE_SCALAR ? 2 : 12;
    float *const *rf = (float *const *)resf;
    float *const *cf = (float *const *)curf;
    float *const gsf[7] = {csf->x, csf->y, csf->z, csf->nx, csf->ny, csf->nz, csf->area};
    float *const gtf[7] = {ctf->x, ctf->y, ctf->z, ctf->nx, ctf->ny, ctf->nz, ctf->area};
    if (!float_arrays_ok(rf, nres, "result") || !float_arrays_ok(cf, ncur, "currents") ||
        !float_arrays_ok(gsf, 7, "source grid") || !float_arrays_ok(gtf, 7, "target grid"))
        return;
    WideRefl ws, wt;
    if (!widen_refl(sf, ws, "source") || !widen_refl(tf, wt, "target")) return;
    HostArrays cur(ncur, gs), res(nres, gt), gsd(7, gs), gtd(7, gt);
    for (int q = 0; q < ncur; q++)
        for (int i = 0; i < gs; i++) cur.ptr[q][i] = cf[q][i];
    reflcontainer cs{gs, gsd.ptr[0], gsd.ptr[1], gsd.ptr[2], gsd.ptr[3], gsd.ptr[4], gsd.ptr[5], gsd.ptr[6]};
    reflcontainer ct{gt, gtd.ptr[0], gtd.ptr[1], gtd.ptr[2], gtd.ptr[3], gtd.ptr[4], gtd.ptr[5], gtd.ptr[6]};
    if (po_dropin(mode, res.ptr, ws.p, wt.p, &cs, &ct, cur.ptr, k, epsilon, tdir)) return;
    for (int q = 0; q < nres; q++)
        for (int i = 0; i < gt; i++) rf[q][i] = (float)res.ptr[q][i];
    for (int q = 0; q < 7; q++) {
        for (int i = 0; i < gs; i++) gsf[q][i] = (float)gsd.ptr[q][i];
        for (int i = 0; i < gt; i++) gtf[q][i] = (float)gtd.ptr[q][i];
    }
}

void callKernelf_JM(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                    reflcontainerf *ct, c2Bundlef *currents, float k, float epsilon, float t_direction, int, int) {
    po_float(PYPO_MODE_JM, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void callKernelf_EH(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                    reflcontainerf *ct, c2Bundlef *currents, float k, float epsilon, float t_direction, int, int) {
    po_float(PYPO_MODE_EH, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void callKernelf_JMEH(c4Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                      reflcontainerf *ct, c2Bundlef *currents, float k, float epsilon, float t_direction, int,
                      int) {
    po_float(PYPO_MODE_JMEH, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void callKernelf_EHP(c2rBundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                     reflcontainerf *ct, c2Bundlef *currents, float k, float epsilon, float t_direction, int,
                     int) {
    po_float(PYPO_MODE_EHP, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void callKernelf_FF(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                    reflcontainerf *ct, c2Bundlef *currents, float k, float epsilon, float t_direction, int, int) {
    po_float(PYPO_MODE_FF, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void callKernelf_scalar(arrC1f *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                        reflcontainerf *ct, arrC1f *inp, float k, float epsilon, float t_direction, int, int) {
    po_float(PYPO_MODE_SCALAR, res, source, target, cs, ct, inp, k, epsilon, t_direction);
}

// InterfaceCPU.h:64-101: the CPU library's FP32 names, argument order with
// numThreads after k; the thread count has no meaning on the GPU.
void propagateToGridf_JM(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                         reflcontainerf *ct, c2Bundlef *currents, float k, int, float epsilon, float t_direction) {
    po_float(PYPO_MODE_JM, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void propagateToGridf_EH(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                         reflcontainerf *ct, c2Bundlef *currents, float k, int, float epsilon, float t_direction) {
    po_float(PYPO_MODE_EH, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void propagateToGridf_JMEH(c4Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                           reflcontainerf *ct, c2Bundlef *currents, float k, int, float epsilon,
                           float t_direction) {
    po_float(PYPO_MODE_JMEH, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void propagateToGridf_EHP(c2rBundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                          reflcontainerf *ct, c2Bundlef *currents, float k, int, float epsilon,
                          float t_direction) {
    po_float(PYPO_MODE_EHP, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void propagateToGridf_scalar(arrC1f *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                             reflcontainerf *ct, arrC1f *field, float k, int, float epsilon, float t_direction) {
    po_float(PYPO_MODE_SCALAR, res, source, target, cs, ct, field, k, epsilon, t_direction);
}

void propagateToFarFieldf(c2Bundlef *res, reflparamsf source, reflparamsf target, reflcontainerf *cs,
                          reflcontainerf *ct, c2Bundlef *currents, float k, int, float epsilon,
                          float t_direction) {
    po_float(PYPO_MODE_FF, res, source, target, cs, ct, currents, k, epsilon, t_direction);
}

void callRTKernel(reflparamsf ctp, cframef *fr_in, cframef *fr_out, float epsilon, float t0, int, int) {
    if (!fr_in || !fr_out || fr_in->size != fr_out->size || fr_in->size < 0) {
        set_error(ERR_ARG, "callRTKernel: NULL frame or frame sizes differ");
        return;
    }
    const int n = fr_in->size;
    float *const fi[6] = {fr_in->x, fr_in->y, fr_in->z, fr_in->dx, fr_in->dy, fr_in->dz};
    float *const fo[6] = {fr_out->x, fr_out->y, fr_out->z, fr_out->dx, fr_out->dy, fr_out->dz};
    if (!float_arrays_ok(fi, 6, "fr_in") || !float_arrays_ok(fo, 6, "fr_out")) return;
    WideRefl w;
    if (!widen_refl(ctp, w, "callRTKernel")) return;
    if (n == 0) return;
    HostArrays in(6, n), out(6, n);
    for (int q = 0; q < 6; q++)
        for (int i = 0; i < n; i++) in.ptr[q][i] = fi[q][i];
    cframe a{n, in.ptr[0], in.ptr[1], in.ptr[2], in.ptr[3], in.ptr[4], in.ptr[5]};
    cframe b{n, out.ptr[0], out.ptr[1], out.ptr[2], out.ptr[3], out.ptr[4], out.ptr[5]};
    if (rays_dropin(w.p, &a, &b, epsilon, t0)) return;
    for (int q = 0; q < 6; q++)
        for (int i = 0; i < n; i++) fo[q][i] = (float)out.ptr[q][i];
}

// ---- drop-in grid generation (InterfaceReflector.h:21-22) ---------------------------

void generateGrid(reflparams refl, reflcontainer *container, bool transform, bool spheric) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    if (ensure_init()) return;
    if (check_container(container, "generateGrid")) return;
    const int n = container->size;
    if (check_refl(refl, n, "generateGrid")) return;
    hipStream_t st = g_stream;
    if (begin_call(st)) return;
    int err = 0;
    double *d = (double *)pool_get(SLOT_TGT_GRID, 7 * (size_t)n * sizeof(double), &err);
    if (err) return;
    GridPtrs g = carve_grid(d, n);
    if (device_grid(refl, g, n, SLOT_ERRS_T, transform, spheric, st)) return;
    if (copy_grid_to_host(g, n, container, st)) return;
    end_call(st);
    (void)hipStreamSynchronize(st);
}

// ---- device-resident extensions --------------------------------------------------------

int pypo_hip_generate_grid_dev(reflparams refl, reflcontainer *container, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    if ((err = check_container(container, "grid"))) return err;
    const int n = container->size;
    if ((err = check_refl(refl, n, "grid"))) return err;
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream
    if ((err = begin_call(st))) return err;
    if ((err = device_grid(refl, grid_ptrs(*container), n, SLOT_ERRS_S, true, true, st))) return err;
    return end_call(st);
}

int pypo_hip_propagate_dev(int mode, void *res, const reflcontainer *cs, const reflcontainer *ct,
                           const void *currents, int gmode, int ncx, int ncy, double k,
                           double epsilon, double t_direction, int precision, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(mode >= PYPO_MODE_JM && mode <= PYPO_MODE_SCALAR, ERR_ARG, "unknown PO mode");
    PYPO_REQUIRE(precision == PYPO_PREC_MIXED || precision == PYPO_PREC_FP64, ERR_ARG,
                 "unknown precision");
    PYPO_REQUIRE(res && currents && cs && ct, ERR_ARG, "NULL argument");
    if ((err = check_container(cs, "source")) || (err = check_container(ct, "target"))) return err;
    const int gs = cs->size, gt = ct->size;
    PYPO_REQUIRE(gs > 0 && gt > 0 && (long long)ncx * ncy == gs, ERR_ARG,
                 "ncx*ncy must equal the source size and both grids must be non-empty");
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream
    if ((err = begin_call(st))) return err;
    void *packed = pool_get(SLOT_PACKED, packed_bytes(mode, precision, gs), &err);
    if (err) return err;
    GridPtrs gsrc = grid_ptrs(*cs);
    if (mode == PYPO_MODE_SCALAR) {
        const arrC1 *f = (const arrC1 *)currents;
        err = launch_pack_scalar(precision, gsrc, gs, f->x, f->y, gmode, ncx, ncy, packed, st);
    } else {
        err = launch_pack_sources(precision, gsrc, gs, bundle12(*(const c2Bundle *)currents), gmode,
                                  ncx, ncy, packed, st);
    }
    if (err) return err;
    SumPlan plan = plan_sum(mode, precision, gs, gt);
    size_t wsb = partial_bytes(mode, plan, gt);
    double *ws = (double *)pool_get(SLOT_PARTIAL, wsb, &err);
    if (err) return err;
    err = launch_po(mode, precision, packed, gs, grid_ptrs(*ct), gt, k, epsilon, t_direction, res, ws,
                    wsb, st, g_timing);
    if (err) return err;
    return end_call(st);
}

// ---- allocation-free propagation ---------------------------------------------------

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};        // device address of each device's scratch
hipEvent_t g_scratch_done[kMaxDevices] = {};   // completion of the last no-alloc call on each device

// This device's scratch address and ordering event, made on the first call
// on the device with that device current.
int scratch_of_current_device(void **out, hipEvent_t **done) {
    int dev = -1;
    PYPO_HIP_TRY(hipGetDevice(&dev));
    PYPO_REQUIRE(dev >= 0 && dev < kMaxDevices, ERR_ARG, "no-alloc propagation: device index out of range");
    if (!g_scratch[dev]) {
        void *p = nullptr;
        PYPO_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PYPO_REQUIRE(p != nullptr, ERR_ARG, "no-alloc propagation: scratch symbol has no device address");
        g_scratch[dev] = p;
    }
    if (!g_scratch_done[dev]) PYPO_HIP_TRY(hipEventCreateWithFlags(&g_scratch_done[dev], hipEventDisableTiming));
    *out = g_scratch[dev];
    *done = &g_scratch_done[dev];
    return 0;
}

// Records the scratch's completion event on the call's stream when the call
// ends, also when it stops on an error after enqueuing work.
struct ScratchMark {
    hipEvent_t ev;
    hipStream_t st;
    bool on;
    ~ScratchMark() {
        if (on) (void)hipEventRecord(ev, st);
    }
};
}  // namespace

int pypo_hip_propagate_noalloc(int mode, void *res, const reflcontainer *cs, const reflcontainer *ct,
                               const void *currents, int gmode, int ncx, int ncy, double k,
                               double epsilon, double t_direction, int precision, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(mode >= PYPO_MODE_JM && mode <= PYPO_MODE_SCALAR, ERR_ARG, "unknown PO mode");
    PYPO_REQUIRE(precision == PYPO_PREC_MIXED || precision == PYPO_PREC_FP64, ERR_ARG,
                 "unknown precision");
    PYPO_REQUIRE(res && currents && cs && ct, ERR_ARG, "NULL argument");
    if ((err = check_container(cs, "source")) || (err = check_container(ct, "target"))) return err;
    const int gs = cs->size, gt = ct->size;
    PYPO_REQUIRE(gs > 0 && gt > 0 && (long long)ncx * ncy == gs, ERR_ARG,
                 "ncx*ncy must equal the source size and both grids must be non-empty");
    // scratch = [packed records + tile metadata | split partials]; the
    // partials take at most as many source splits as fit behind the records
    const size_t cap = (size_t)PYPO_HIP_NOALLOC_SCRATCH_BYTES;
    const size_t pk = (packed_bytes(mode, precision, gs) + 255) / 256 * 256;
    const size_t per_split = (size_t)(mode == PYPO_MODE_SCALAR ? 2 : 12) * gt * sizeof(double);
    if (pk + per_split > cap) {
        set_error(ERR_ARG, "pypo_hip_propagate_noalloc: " + std::to_string(gs) + " sources -> " +
                               std::to_string(gt) + " targets need " + std::to_string(pk + per_split) +
                               " bytes of scratch, more than its " + std::to_string(cap) +
                               " bytes; use pypo_hip_propagate_dev");
        return ERR_ARG;
    }
    void *base = nullptr;
    hipEvent_t *scratch_done = nullptr;
    if ((err = scratch_of_current_device(&base, &scratch_done))) return err;
    const long long fit = (long long)((cap - pk) / per_split);
    const SumPlan plan = plan_cap_splits(plan_sum(mode, precision, gs, gt), gs, (int)std::min(fit, 1LL << 30));
    double *partial = (double *)((char *)base + pk);
    const size_t wsb = cap - pk;
    PYPO_REQUIRE(partial_bytes(mode, plan, gt) <= wsb, ERR_ARG, "internal: split plan exceeds the scratch");
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream
    if ((err = begin_call(st))) return err;
    // Calls on every stream of the device share its scratch: each waits for
    // the previous one (outside a capture, whose replays are not ordered).
    if (!g_capturing) PYPO_HIP_TRY(hipStreamWaitEvent(st, *scratch_done, 0));
    ScratchMark mark{*scratch_done, st, !g_capturing};
    GridPtrs gsrc = grid_ptrs(*cs);
    if (mode == PYPO_MODE_SCALAR) {
        const arrC1 *f = (const arrC1 *)currents;
        err = launch_pack_scalar(precision, gsrc, gs, f->x, f->y, gmode, ncx, ncy, base, st);
    } else {
        err = launch_pack_sources(precision, gsrc, gs, bundle12(*(const c2Bundle *)currents), gmode, ncx,
                                  ncy, base, st);
    }
    if (err) return err;
    err = launch_po_plan(plan, mode, precision, base, gs, grid_ptrs(*ct), gt, k, epsilon, t_direction, res,
                         partial, wsb, st, false);
    if (err) return err;
    return end_call(st);
}

int pypo_hip_trace_dev(const reflparams *surfs, int n_surf, cframe *frames, double tol, double t0,
                       void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(surfs && frames && n_surf >= 1 && n_surf <= 4, ERR_ARG,
                 "trace: 1 to 4 surfaces and non-NULL frames");
    const int n = frames[0].size;
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream
    if ((err = begin_call(st))) return err;
    const double *d_errs[4] = {nullptr, nullptr, nullptr, nullptr};
    int n_rough = 0;
    for (int k = 0; k < n_surf; k++) n_rough += surfs[k].rms > 0 ? 1 : 0;
    if (n_rough && n > 0) {
        double *de = (double *)pool_get(SLOT_RT_ERRS, (size_t)n_surf * n * sizeof(double), &err);
        if (err) return err;
        for (int k = 0; k < n_surf; k++) {
            if (surfs[k].rms <= 0) continue;
            std::vector<double> e = ray_deviates(surfs[k], n);
            PYPO_HIP_TRY(hipMemcpyAsync(de + k * (size_t)n, e.data(), n * sizeof(double),
                                        hipMemcpyHostToDevice, st));
            PYPO_HIP_TRY(hipStreamSynchronize(st));
            d_errs[k] = de + k * (size_t)n;
        }
    }
    if ((err = launch_trace_chain(surfs, d_errs, n_surf, frames, tol, t0, g_rt_max_iter, st,
                                  g_timing)))
        return err;
    return end_call(st);
}

// ---- batched sweeps ---------------------------------------------------------------

int pypo_hip_sweep_create(const pypo_sweep_desc *desc, void **handle) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(handle, ERR_ARG, "sweep: NULL handle");
    *handle = nullptr;
    Sweep *s = nullptr;
    if ((err = sweep_create(desc, &s, g_stream))) return err;
    *handle = s;
    return 0;
}

int pypo_hip_sweep_plan(void *handle, int tpl, int chunk_tiles, int wgs_per_cu, int use_graph) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    PYPO_REQUIRE(handle, ERR_ARG, "sweep: NULL handle");
    SweepPlanParams prm{tpl, chunk_tiles, wgs_per_cu};
    return sweep_plan((Sweep *)handle, prm, use_graph != 0, g_stream);
}

int pypo_hip_sweep_run(void *handle, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    PYPO_REQUIRE(handle, ERR_ARG, "sweep: NULL handle");
    return sweep_run((Sweep *)handle, (hipStream_t)stream);
}

int pypo_hip_sweep_read(void *handle, int inst, c2Bundle *ff_out) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    PYPO_REQUIRE(handle && ff_out, ERR_ARG, "sweep: NULL argument");
    PYPO_HIP_TRY(hipDeviceSynchronize());
    return sweep_read((Sweep *)handle, inst, ff_out, g_stream);
}

int pypo_hip_sweep_read_currents(void *handle, int inst, c2Bundle *dish_currents) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    PYPO_REQUIRE(handle && dish_currents, ERR_ARG, "sweep: NULL argument");
    PYPO_HIP_TRY(hipDeviceSynchronize());
    return sweep_read_currents((Sweep *)handle, inst, dish_currents, g_stream);
}

int pypo_hip_sweep_info(void *handle, double *pairs_jm, double *pairs_ff, int *items_jm,
                        int *items_ff, double **ff_out_dev) {
    PYPO_REQUIRE(handle, ERR_ARG, "sweep: NULL handle");
    sweep_info((const Sweep *)handle, pairs_jm, pairs_ff, items_jm, items_ff, ff_out_dev);
    return 0;
}

int pypo_hip_sweep_destroy(void *handle) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    if (handle) sweep_destroy((Sweep *)handle);
    return 0;
}

int pypo_hip_set_sum_shape(int tpl, int want_wgs) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    return set_sum_shape(tpl, want_wgs);
}

int pypo_hip_inject_alloc_failures(int n) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    PYPO_REQUIRE(n >= 0, ERR_ARG, "inject_alloc_failures: n must be >= 0");
    g_fail_allocs = n;
    return 0;
}

int pypo_hip_set_ff_kernel(int kind) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    return set_ff_kernel(kind);
}

// ---- beam initialisation (InterfaceBeam.h:21-34) -----------------------------------

void makeGauss(GPODict gdict, reflparams plane, c2Bundle *res_field, c2Bundle *res_current) {
    if (check_bundle12(res_field, "makeGauss field") || check_bundle12(res_current, "makeGauss current"))
        return;
    double *out[24];
    for (int q = 0; q < 12; q++) {
        out[q] = ((double **)res_field)[q];
        out[12 + q] = ((double **)res_current)[q];
    }
    beam_dropin(BEAM_GAUSS, &gdict, plane, out, 24, nullptr, 0);
}

void makeGaussBeam(vecGPODict gdict, reflparams plane, c2Bundle *res_field, c2Bundle *res_current) {
    if (check_bundle12(res_field, "makeGaussBeam field") ||
        check_bundle12(res_current, "makeGaussBeam current"))
        return;
    double *out[24];
    for (int q = 0; q < 12; q++) {
        out[q] = ((double **)res_field)[q];
        out[12 + q] = ((double **)res_current)[q];
    }
    beam_dropin(BEAM_VEC_GAUSS, &gdict, plane, out, 24, nullptr, 0);
}

void makeScalarGauss(ScalarGPODict sgdict, reflparams plane, arrC1 *res_field) {
    if (!res_field) {
        set_error(ERR_ARG, "makeScalarGauss: NULL result");
        return;
    }
    double *out[2] = {res_field->x, res_field->y};
    beam_dropin(BEAM_SCALAR, &sgdict, plane, out, 2, nullptr, 0);
}

void calcCurrents(c2Bundle *res_field, c2Bundle *res_current, reflparams rdict, int mode) {
    if (check_bundle12(res_field, "calcCurrents field") ||
        check_bundle12(res_current, "calcCurrents current"))
        return;
    beam_dropin(BEAM_CURRENTS, nullptr, rdict, (double *const *)res_current, 12,
                (double *const *)res_field, mode);
}

void makeRTframe(RTDict rdict, cframe *fr) {
    if (rdict.nRays < 0 || rdict.nRing < 0) {
        set_error(ERR_ARG, "makeRTframe: nRays and nRing must be >= 0");
        return;
    }
    frame_dropin(false, &rdict, rt_frame_size(rdict), fr);
}

void makeGRTframe(GRTDict grdict, cframe *fr) { frame_dropin(true, &grdict, grdict.nRays, fr); }

int pypo_hip_make_gauss_dev(GPODict gdict, const reflcontainer *grid, c2Bundle *field,
                            c2Bundle *current, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    if ((err = check_container(grid, "grid")) || (err = check_bundle12(field, "field")) ||
        (err = check_bundle12(current, "current")))
        return err;
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream
    if ((err = begin_call(st))) return err;
    if ((err = launch_gauss(gdict, grid_ptrs(*grid), grid->size, bundle12(*field), bundle12(*current), st)))
        return err;
    return end_call(st);
}

int pypo_hip_make_gauss_beam_dev(vecGPODict gdict, const reflcontainer *grid, c2Bundle *field,
                                 c2Bundle *current, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    if ((err = check_container(grid, "grid")) || (err = check_bundle12(field, "field")) ||
        (err = check_bundle12(current, "current")))
        return err;
    hipStream_t st = (hipStream_t)stream;
    if ((err = begin_call(st))) return err;
    if ((err = launch_gauss_beam(gdict, grid_ptrs(*grid), grid->size, bundle12(*field), bundle12(*current),
                                 st)))
        return err;
    return end_call(st);
}

int pypo_hip_make_scalar_gauss_dev(ScalarGPODict sgdict, const reflcontainer *grid, arrC1 *field,
                                   void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    if ((err = check_container(grid, "grid"))) return err;
    PYPO_REQUIRE(field && field->x && field->y, ERR_ARG, "field: NULL arrC1 or array");
    hipStream_t st = (hipStream_t)stream;
    if ((err = begin_call(st))) return err;
    if ((err = launch_scalar_gauss(sgdict, grid_ptrs(*grid), grid->size, field->x, field->y, st))) return err;
    return end_call(st);
}

int pypo_hip_calc_currents_dev(const c2Bundle *field, c2Bundle *current, const reflcontainer *grid,
                               int mode, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    if ((err = check_container(grid, "grid")) || (err = check_bundle12(field, "field")) ||
        (err = check_bundle12(current, "current")))
        return err;
    hipStream_t st = (hipStream_t)stream;
    if ((err = begin_call(st))) return err;
    if ((err = launch_calc_jm(bundle12(*field), grid_ptrs(*grid), grid->size, mode, bundle12(*current), st)))
        return err;
    return end_call(st);
}

int pypo_hip_make_rt_frame_dev(RTDict rdict, cframe *fr, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(fr && fr->x && fr->y && fr->z && fr->dx && fr->dy && fr->dz, ERR_ARG,
                 "ray frame: NULL cframe or array");
    hipStream_t st = (hipStream_t)stream;
    if ((err = begin_call(st))) return err;
    if ((err = frame_dev(false, &rdict, *fr, st))) return err;
    return end_call(st);
}

int pypo_hip_make_grt_frame_dev(GRTDict grdict, cframe *fr, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(fr && fr->x && fr->y && fr->z && fr->dx && fr->dy && fr->dz, ERR_ARG,
                 "ray frame: NULL cframe or array");
    hipStream_t st = (hipStream_t)stream;
    if ((err = begin_call(st))) return err;
    if ((err = frame_dev(true, &grdict, *fr, st))) return err;
    return end_call(st);
}

// ---- hybrid ray-trace / PO (System.runHybridPropagation) -----------------------------

static int check_frame(const cframe *f, const char *what) {
    if (!f || !f->x || !f->y || !f->z || !f->dx || !f->dy || !f->dz) {
        set_error(ERR_ARG, std::string(what) + ": NULL cframe or array");
        return ERR_ARG;
    }
    return 0;
}

int pypo_hip_hybrid_dev(reflparams target, const cframe *fr_in, cframe *fr_out, const c2Bundle *field_in,
                        c2Bundle *field_out, const double *start, double k, double tol, double t0,
                        void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    if ((err = check_frame(fr_in, "fr_in")) || (err = check_frame(fr_out, "fr_out")) ||
        (err = check_bundle12(field_in, "field_in")) || (err = check_bundle12(field_out, "field_out")))
        return err;
    PYPO_REQUIRE(fr_in->size > 0 && fr_out->size == fr_in->size, ERR_ARG,
                 "hybrid: fr_in and fr_out must have the same non-zero size");
    cframe frames[2] = {*fr_in, *fr_out};
    if ((err = pypo_hip_trace_dev(&target, 1, frames, tol, t0, stream))) return err;
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream
    if ((err = begin_call(st))) return err;
    if ((err = launch_hybrid_phase(*fr_in, *fr_out, bundle12(*field_in), bundle12(*field_out), fr_in->size,
                                   k, start, st)))
        return err;
    return end_call(st);
}

int pypo_hip_interp_nearest_dev(const cframe *points, const c2Bundle *values, const reflcontainer *grid,
                                c2Bundle *out, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    if ((err = check_frame(points, "points")) || (err = check_bundle12(values, "values")) ||
        (err = check_container(grid, "grid")) || (err = check_bundle12(out, "out")))
        return err;
    PYPO_REQUIRE(points->size > 0 && grid->size > 0, ERR_ARG, "interpolation: empty point set or grid");
    hipStream_t st = (hipStream_t)stream;
    if ((err = begin_call(st))) return err;
    const size_t wsb = nn_workspace_bytes(points->size);
    void *ws = pool_get(SLOT_NN, wsb, &err);
    if (err) return err;
    if ((err = launch_interp_nearest(points->x, points->y, points->z, points->size, bundle12(*values), 12,
                                     grid->x, grid->y, grid->z, grid->size, bundle12(*out), ws, wsb, st)))
        return err;
    return end_call(st);
}

// ---- adjoint of the far-field power -------------------------------------------------

int pypo_hip_displace_dev(const reflcontainer *base, const double *d, reflcontainer *out,
                          void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(d, ERR_ARG, "displace: NULL displacement");
    if ((err = check_container(base, "base")) || (err = check_container(out, "out"))) return err;
    PYPO_REQUIRE(base->size > 0 && out->size == base->size, ERR_ARG,
                 "displace: base and out must have the same non-zero size");
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream
    if ((err = begin_call(st))) return err;
    if ((err = launch_displace(grid_ptrs(*base), d, base->size, grid_ptrs(*out), st))) return err;
    return end_call(st);
}

int pypo_hip_ff_power_grad_dev(const reflcontainer *dish, int gmode, int ncx, int ncy,
                               const reflcontainer *feed,
                               int feed_gmode, int feed_ncx, int feed_ncy,
                               const c2Bundle *feed_currents, const reflcontainer *ff,
                               const c2Bundle *ff_fields, const double *dl_dp, double k,
                               double t_direction, int precision, double *grad, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(precision == PYPO_PREC_MIXED || precision == PYPO_PREC_FP64, ERR_ARG,
                 "unknown precision");
    PYPO_REQUIRE(feed_currents && ff_fields && dl_dp && grad, ERR_ARG,
                 "adjoint: NULL argument");
    if ((err = check_container(dish, "dish")) || (err = check_container(feed, "feed")) ||
        (err = check_container(ff, "ff")))
        return err;
    PYPO_REQUIRE(k > 0 && (t_direction == 1.0 || t_direction == -1.0), ERR_ARG,
                 "adjoint: k must be positive and t_direction +-1");
    const size_t wsb = adjoint_workspace_bytes(precision, feed->size, dish->size, ff->size);
    hipStream_t st = (hipStream_t)stream;   // NULL: the default stream
    if ((err = begin_call(st))) return err;
    void *ws = pool_get(SLOT_ADJ, wsb, &err);
    if (err) return err;
    if ((err = launch_ff_power_grad(precision, grid_ptrs(*dish), dish->size, gmode, ncx, ncy,
                                    grid_ptrs(*feed), feed->size, feed_gmode,
                                    feed_ncx, feed_ncy, bundle12(*feed_currents), grid_ptrs(*ff),
                                    ff->size, bundle12(*ff_fields), dl_dp, k, t_direction, grad, ws,
                                    wsb, st)))
        return err;
    return end_call(st);
}

// ---- figures of merit (Efficiencies.py) -------------------------------------------------

// Device workspace of the reductions and the address of its result slots.
static int metrics_ws(double **ws, double **res) {
    int err = 0;
    *ws = (double *)pool_get(SLOT_METRICS, metrics_workspace_bytes(), &err);
    if (err) return err;
    *res = *ws + metrics_result_offset();
    return 0;
}

// Copies nres folded results to the host and waits for them.
static int metrics_finish(hipStream_t st, const double *res, double *host, int nres) {
    PYPO_HIP_TRY(hipMemcpyAsync(host, res, nres * sizeof(double), hipMemcpyDeviceToHost, st));
    int err = end_call(st);
    if (err) return err;
    PYPO_HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int pypo_hip_aperture_sums_dev(const reflcontainer *grid, const double *f_re, const double *f_im,
                               aperparams aper, double *sums, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    if ((err = check_container(grid, "grid"))) return err;
    PYPO_REQUIRE(f_re && f_im && sums, ERR_ARG, "aperture sums: NULL field or output");
    PYPO_REQUIRE(grid->size > 0, ERR_ARG, "aperture sums: empty grid");
    PYPO_REQUIRE(aper.shape == PYPO_APER_ELLIPSE || aper.shape == PYPO_APER_RECT || aper.shape == PYPO_APER_NONE,
                 ERR_ARG, "aperture sums: unknown aperture shape");
    hipStream_t st = (hipStream_t)stream;
    if ((err = begin_call(st))) return err;
    double *ws, *res;
    if ((err = metrics_ws(&ws, &res))) return err;
    if ((err = launch_aperture_sums(grid->x, grid->y, grid->area, f_re, f_im, grid->size, aper, ws, res, st)))
        return err;
    return metrics_finish(st, res, sums, 5);
}

int pypo_hip_power_sums_dev(const double *a_re, const double *a_im, const double *b_re, const double *b_im,
                            int n, double *sums, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    PYPO_REQUIRE(a_re && a_im && b_re && b_im && sums, ERR_ARG, "power sums: NULL argument");
    PYPO_REQUIRE(n > 0, ERR_ARG, "power sums: n must be positive");
    hipStream_t st = (hipStream_t)stream;
    if ((err = begin_call(st))) return err;
    double *ws, *res;
    if ((err = metrics_ws(&ws, &res))) return err;
    if ((err = launch_power_sums(a_re, a_im, b_re, b_im, n, ws, res, st))) return err;
    return metrics_finish(st, res, sums, 2);
}

int pypo_hip_frame_stats_dev(const cframe *fr, double *stats, void *stream) {
    std::lock_guard<std::recursive_mutex> lock(g_mu);
    int err = ensure_init();
    if (err) return err;
    if ((err = check_frame(fr, "frame"))) return err;
    PYPO_REQUIRE(stats, ERR_ARG, "frame stats: NULL output");
    PYPO_REQUIRE(fr->size > 0, ERR_ARG, "frame stats: empty frame");
    hipStream_t st = (hipStream_t)stream;
    if ((err = begin_call(st))) return err;
    double *ws, *res;
    if ((err = metrics_ws(&ws, &res))) return err;
    const FrameView f{fr->x, fr->y, fr->z, fr->dx, fr->dy, fr->dz};
    if ((err = launch_frame_stats(f, fr->size, ws, res, st))) return err;
    return metrics_finish(st, res, stats, 8);
}

}  // extern "C"
