// sweep.hip -- batched JM -> FF sweeps (frequency x feed offset, uneven grid
// sizes): one launch per stage over a static work list.
//
// A sweep holds n_inst instances, each a feed (source of the JM step), a
// dish (target of JM, source of FF) and a wavenumber; all instances share
// one far-field grid.  Each stage is one launch: workgroup w processes work
// items w, w + G, w + 2G, ... of a list sorted by cost (longest first), so
// the launch needs no atomics or counters and every workgroup leaves the
// loop when the list ends.  G may reach one workgroup per item; the
// dispatcher then starts the items longest first as slots free.  A
// work item is (instance, 256*TPL-target block, source chunk); its FP64
// partial sums go to a per-instance slot per chunk and batched finish
// kernels reduce the chunks in order and apply the epilogues (J = 2 n x H
// for JM, the far-field projection for FF).  A run is five kernels -- JM
// sum, JM finish, packing the dish currents, FF sum, FF finish -- that can
// be captured into one hipGraph and replayed.

#include <algorithm>
#include <cstring>
#include <vector>

#include "po_ff_mfma.hpp"

namespace pypo {

struct SweepItem {
    int inst, tblock, s0, s1, chunk;
};

// Per-instance view of one stage.
struct StageInst {
    const SrcRec32 *src;      // packed sources
    const TileMeta *meta;
    int n_src;
    int n_tgt;
    double k, tdir, rev_scale, pref;
    const double *tx, *ty, *tz;            // targets (near field) or th, ph (far field)
    const double *tnx, *tny, *tnz;         // target normals (JM epilogue)
    double *partial;                       // [n_chunks][12][n_tgt]
    int n_chunks;
    double *out;                           // [12][n_tgt] result in c2Bundle row order
};

// Packing view: dish grid + dish currents -> records.
struct PackInst {
    GridPtrs grid;
    const double *cur;        // [12][n]
    int n, gmode, ncx, ncy;
    SrcRec32 *rec;
    TileMeta *meta;
};

namespace {

// The far-field stage on the matrix cores (po_ff_mfma.hpp): one work item at
// a time, 64 MB targets per item (MB = 4 TPL, matching the item's target block).
template <int MB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_sweep_ff_mfma(
    const StageInst *__restrict__ insts, const SweepItem *__restrict__ items, int n_items) {
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const SweepItem item = items[it];
        const StageInst &in = insts[item.inst];
        mfma::ff_mfma_body<MB>(in.src, in.meta, item.s0, item.s1, item.tblock * 64 * MB, in.tx, in.ty, in.n_tgt,
                               in.rev_scale, in.partial + (size_t)item.chunk * 12 * in.n_tgt);
        __syncthreads();   // the next item reuses the LDS tiles
    }
}

// Sixteen M-blocks (1024 targets per item) at one wave per SIMD: twice the
// targets per staged tile, half the waves to hide latency with.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1))) void k_sweep_ff_mfma16(
    const StageInst *__restrict__ insts, const SweepItem *__restrict__ items, int n_items) {
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const SweepItem item = items[it];
        const StageInst &in = insts[item.inst];
        mfma::ff_mfma_body<16>(in.src, in.meta, item.s0, item.s1, item.tblock * 64 * 16, in.tx, in.ty,
                               in.n_tgt, in.rev_scale, in.partial + (size_t)item.chunk * 12 * in.n_tgt);
        __syncthreads();   // the next item reuses the LDS tiles
    }
}

template <int TPL>
__global__ __launch_bounds__(256) void k_sweep_ff(const StageInst *__restrict__ insts,
                                                  const SweepItem *__restrict__ items, int n_items) {
    __shared__ SrcRec32 tile[TILE];
    __shared__ float scratch[4];
    const int tid = threadIdx.x;
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const SweepItem item = items[it];
        const StageInst &in = insts[item.inst];
        const int n_tgt = in.n_tgt;
        double rx[TPL], ry[TPL], rz[TPL];
        float fx[TPL], fy[TPL], fz[TPL];
        bool any = false;
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            int t = item.tblock * (TILE * TPL) + j * TILE + tid;
            double x = 0, y = 0, z = 0;
            if (t < n_tgt) {
                double st, ct, sp, cp;
                sincos(in.tx[t], &st, &ct);
                sincos(in.ty[t], &sp, &cp);
                x = cp * st; y = sp * st; z = ct;
            }
            if (j == 0) any = t < n_tgt;
            fx[j] = (float)x; fy[j] = (float)y; fz[j] = (float)z;
            rx[j] = x * in.rev_scale; ry[j] = y * in.rev_scale; rz[j] = z * in.rev_scale;
        }
        Cone cone = block_cone(fx[0], fy[0], fz[0], any, scratch);
        if (TPL > 1 && cone.cos_half > 0.f) {
            float d = 1.f;
#pragma unroll
            for (int j = 1; j < TPL; j++) {
                int t = item.tblock * (TILE * TPL) + j * TILE + tid;
                if (t < n_tgt) d = fminf(d, fx[j] * cone.ax + fy[j] * cone.ay + fz[j] * cone.az);
            }
            cone.cos_half = fminf(cone.cos_half, block_min(d, scratch));
        }
        float acc[TPL][12];
        double tot[TPL][12];
#pragma unroll
        for (int j = 0; j < TPL; j++)
#pragma unroll
            for (int q = 0; q < 12; q++) { acc[j][q] = 0.f; tot[j][q] = 0.0; }
        const uint4 *g = reinterpret_cast<const uint4 *>(in.src);
        TileLoader<6> ld;
        ld.load(g, item.s0 + tid, item.s1);
        for (int base = item.s0; base < item.s1; base += TILE) {
            __syncthreads();
            ld.store(reinterpret_cast<uint4 *>(tile), tid);
            __syncthreads();
            if (base + TILE < item.s1) ld.load(g, base + TILE + tid, item.s1);
            const TileMeta m = in.meta[base / TILE];
            const bool mask = cones_may_shadow(m.ax, m.ay, m.az, m.cos_half, cone);
            if (m.has_m) {
                if (mask) ff_tile_mixed<TPL, true, true>(tile, rx, ry, rz, fx, fy, fz, acc);
                else ff_tile_mixed<TPL, true, false>(tile, rx, ry, rz, fx, fy, fz, acc);
            } else {
                if (mask) ff_tile_mixed<TPL, false, true>(tile, rx, ry, rz, fx, fy, fz, acc);
                else ff_tile_mixed<TPL, false, false>(tile, rx, ry, rz, fx, fy, fz, acc);
            }
#pragma unroll
            for (int j = 0; j < TPL; j++)
#pragma unroll
                for (int q = 0; q < 12; q++) { tot[j][q] += (double)acc[j][q]; acc[j][q] = 0.f; }
        }
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            int t = item.tblock * (TILE * TPL) + j * TILE + tid;
            if (t < n_tgt) {
#pragma unroll
                for (int q = 0; q < 12; q++)
                    in.partial[((size_t)item.chunk * 12 + q) * n_tgt + t] = tot[j][q];
            }
        }
        __syncthreads();   // the next item reuses the tile and the scratch words
    }
}

template <int TPL>
__global__ __launch_bounds__(256) void k_sweep_near(const StageInst *__restrict__ insts,
                                                    const SweepItem *__restrict__ items, int n_items) {
    __shared__ SrcRec32 tile[TILE];
    const int tid = threadIdx.x;
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const SweepItem item = items[it];
        const StageInst &in = insts[item.inst];
        const int n_tgt = in.n_tgt;
        const float kinv = (float)(1.0 / in.k);
        const float tf = (float)in.tdir;
        double px[TPL], py[TPL], pz[TPL];
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            int t = item.tblock * (TILE * TPL) + j * TILE + tid;
            px[j] = (t < n_tgt) ? in.tx[t] : 1.0e6;
            py[j] = (t < n_tgt) ? in.ty[t] : 1.0e6;
            pz[j] = (t < n_tgt) ? in.tz[t] : 1.0e6;
        }
        float acc[TPL][12];
        double tot[TPL][12];
#pragma unroll
        for (int j = 0; j < TPL; j++)
#pragma unroll
            for (int q = 0; q < 12; q++) { acc[j][q] = 0.f; tot[j][q] = 0.0; }
        const uint4 *g = reinterpret_cast<const uint4 *>(in.src);
        TileLoader<6> ld;
        ld.load(g, item.s0 + tid, item.s1);
        for (int base = item.s0; base < item.s1; base += TILE) {
            __syncthreads();
            ld.store(reinterpret_cast<uint4 *>(tile), tid);
            __syncthreads();
            if (base + TILE < item.s1) ld.load(g, base + TILE + tid, item.s1);
            // JM needs H only: rows 6..11 of the partials
            const int n = min(TILE, item.s1 - base);
            if (in.meta[base / TILE].has_m)
                near_tile_mixed<TPL, true, false>(tile, px, py, pz, in.rev_scale, kinv, tf, acc, n);
            else
                near_tile_mixed<TPL, false, false>(tile, px, py, pz, in.rev_scale, kinv, tf, acc, n);
#pragma unroll
            for (int j = 0; j < TPL; j++)
#pragma unroll
                for (int q = 6; q < 12; q++) { tot[j][q] += (double)acc[j][q]; acc[j][q] = 0.f; }
        }
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            int t = item.tblock * (TILE * TPL) + j * TILE + tid;
            if (t < n_tgt) {
#pragma unroll
                for (int q = 6; q < 12; q++)
                    in.partial[((size_t)item.chunk * 12 + q) * n_tgt + t] = tot[j][q];
            }
        }
        __syncthreads();
    }
}

// grid: (ceil(max n_tgt / 256), n_inst)
__global__ __launch_bounds__(256) void k_sweep_jm_finish(const StageInst *__restrict__ insts) {
    const StageInst &in = insts[blockIdx.y];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= in.n_tgt) return;
    const int n = in.n_tgt;
    double S[12];
#pragma unroll
    for (int q = 6; q < 12; q++) {   // the JM sums write the H rows only
        double v = 0.0;
        for (int c = 0; c < in.n_chunks; c++) v += in.partial[((size_t)c * 12 + q) * n + t];
        S[q] = in.pref * v;
    }
    // H_a = (S[6+2a], S[7+2a]); J = 2 n x H (propagateBeam_JM_PEC, Propagation.h:305-321)
    const double nr[3] = {in.tnx[t], in.tny[t], in.tnz[t]};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        in.out[(size_t)a * n + t] = 2 * (nr[b] * S[6 + 2 * c] - nr[c] * S[6 + 2 * b]);
        in.out[(size_t)(3 + a) * n + t] = 2 * (nr[b] * S[7 + 2 * c] - nr[c] * S[7 + 2 * b]);
        in.out[(size_t)(6 + a) * n + t] = 0.0;
        in.out[(size_t)(9 + a) * n + t] = 0.0;
    }
}

__global__ __launch_bounds__(256) void k_sweep_ff_finish(const StageInst *__restrict__ insts) {
    const StageInst &in = insts[blockIdx.y];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= in.n_tgt) return;
    const int n = in.n_tgt;
    double S[12];
#pragma unroll
    for (int q = 0; q < 12; q++) {
        double v = 0.0;
        for (int c = 0; c < in.n_chunks; c++) v += in.partial[((size_t)c * 12 + q) * n + t];
        S[q] = v;
    }
    double st, ct, sp, cp;
    sincos(in.tx[t], &st, &ct);
    sincos(in.ty[t], &sp, &cp);
    const double r[3] = {cp * st, sp * st, ct};
    double jdr_re = 0, jdr_im = 0, mdr_re = 0, mdr_im = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        jdr_re += S[2 * a] * r[a]; jdr_im += S[2 * a + 1] * r[a];
        mdr_re += S[6 + 2 * a] * r[a]; mdr_im += S[7 + 2 * a] * r[a];
    }
    const double tdir = in.tdir, pref = in.pref;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        double rxm_re = r[b] * S[6 + 2 * c] - r[c] * S[6 + 2 * b];
        double rxm_im = r[b] * S[7 + 2 * c] - r[c] * S[7 + 2 * b];
        double rxj_re = r[b] * S[2 * c] - r[c] * S[2 * b];
        double rxj_im = r[b] * S[2 * c + 1] - r[c] * S[2 * b + 1];
        double e_re = S[2 * a] - r[a] * jdr_re + tdir * rxm_re;
        double e_im = S[2 * a + 1] - r[a] * jdr_im + tdir * rxm_im;
        double h_re = S[6 + 2 * a] - r[a] * mdr_re - tdir * rxj_re;
        double h_im = S[7 + 2 * a] - r[a] * mdr_im - tdir * rxj_im;
        in.out[(size_t)a * n + t] = pref * e_im;
        in.out[(size_t)(3 + a) * n + t] = -pref * e_re;
        in.out[(size_t)(6 + a) * n + t] = pref * h_im;
        in.out[(size_t)(9 + a) * n + t] = -pref * h_re;
    }
}

// grid: (ceil(max n / 256), n_inst)
__global__ __launch_bounds__(256) void k_sweep_pack(const PackInst *__restrict__ insts) {
    const PackInst &in = insts[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.n) return;
    const GridPtrs &cs = in.grid;
    const size_t n = in.n;
    double wa = trapezoid_weight(i, in.gmode, in.ncx, in.ncy) * cs.area[i];
    SrcRec32 r;
    r.x = cs.x[i]; r.y = cs.y[i]; r.z = cs.z[i];
    r.nx = (float)cs.nx[i]; r.ny = (float)cs.ny[i]; r.nz = (float)cs.nz[i]; r.pad0 = 0.f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r.cj[2 * a] = (float)(wa * in.cur[a * n + i]);
        r.cj[2 * a + 1] = (float)(wa * in.cur[(3 + a) * n + i]);
        r.cm[2 * a] = (float)(wa * in.cur[(6 + a) * n + i]);
        r.cm[2 * a + 1] = (float)(wa * in.cur[(9 + a) * n + i]);
    }
    r.pad1[0] = 0.f; r.pad1[1] = 0.f;
    in.rec[packed_slot(i, in.n, in.ncx, in.ncy)] = r;
}

// grid: (max tiles, n_inst)
__global__ __launch_bounds__(256) void k_sweep_meta(const PackInst *__restrict__ insts) {
    __shared__ float scratch[4];
    const PackInst &in = insts[blockIdx.y];
    if ((int)blockIdx.x * TILE >= in.n) return;   // uniform per workgroup
    const int i = blockIdx.x * TILE + threadIdx.x;
    const bool valid = i < in.n;
    float nx = 0.f, ny = 0.f, nz = 1.f, cm = 0.f;
    bool m = false;
    if (valid) {
        const SrcRec32 &r = in.rec[i];
        nx = r.nx; ny = r.ny; nz = r.nz;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            m |= (r.cm[q] != 0.f);
            cm = fmaxf(cm, fmaxf(fabsf(r.cj[q]), fabsf(r.cm[q])));
        }
    }
    Cone c = block_cone(nx, ny, nz, valid, scratch);
    float any_m = block_sum(m ? 1.f : 0.f, scratch);
    float cnt = block_sum(valid ? 1.f : 0.f, scratch);
    float cmax = -block_min(-cm, scratch);
    if (threadIdx.x == 0) {
        TileMeta t;
        t.ax = c.ax; t.ay = c.ay; t.az = c.az; t.cos_half = c.cos_half;
        t.has_m = any_m > 0.f ? 1 : 0;
        t.n_valid = (int)cnt;
        t.cmax = cmax; t.pad1 = 0;
        in.meta[blockIdx.x] = t;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct Sweep {
    int n_inst = 0, n_ff = 0;
    int n_cu = 256;
    SweepPlanParams prm{2, 1 << 20, 2};
    double tdir = -1.0;
    // device memory
    std::vector<void *> allocs;
    // per instance (host copies of the descriptors)
    std::vector<StageInst> jm, ff;
    std::vector<PackInst> pk;
    std::vector<int> nf, nd;       // feed and dish sizes
    std::vector<double *> dish_cur;
    double *ff_out = nullptr;      // [n_inst][12][n_ff]
    double *ff_grid = nullptr;     // [7][n_ff]
    // work lists
    StageInst *d_jm = nullptr, *d_ff = nullptr;
    PackInst *d_pk = nullptr;
    SweepItem *d_items_jm = nullptr, *d_items_ff = nullptr;
    int n_items_jm = 0, n_items_ff = 0;
    double *partial_jm = nullptr, *partial_ff = nullptr;
    size_t partial_jm_bytes = 0, partial_ff_bytes = 0;
    int max_nd = 0, max_tiles_d = 0;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
    double pairs_jm = 0, pairs_ff = 0;
};

namespace {

void *sweep_alloc(Sweep *s, size_t bytes, int *err) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 256);
    if (e != hipSuccess) {
        set_error(ERR_ALLOC, std::string("sweep: hipMalloc of ") + std::to_string(bytes) +
                                 " bytes failed: " + hipGetErrorString(e));
        *err = ERR_ALLOC;
        return nullptr;
    }
    s->allocs.push_back(p);
    return p;
}

void sweep_free_plan(Sweep *s) {
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->exec = nullptr;
    s->graph = nullptr;
}

// Work list of one stage: items (inst, tblock, chunk) sorted longest first.
void build_items(const std::vector<StageInst> &st, int tpl, int chunk_tiles,
                 std::vector<SweepItem> &items, std::vector<int> &n_chunks) {
    items.clear();
    n_chunks.assign(st.size(), 0);
    for (size_t i = 0; i < st.size(); i++) {
        const int tiles = n_tiles(st[i].n_src);
        const int nch = (tiles + chunk_tiles - 1) / chunk_tiles;
        n_chunks[i] = nch;
        const int tb = (st[i].n_tgt + TILE * tpl - 1) / (TILE * tpl);
        for (int b = 0; b < tb; b++)
            for (int c = 0; c < nch; c++) {
                int s0 = c * chunk_tiles * TILE;
                int s1 = std::min(st[i].n_src, (c + 1) * chunk_tiles * TILE);
                items.push_back(SweepItem{(int)i, b, s0, s1, c});
            }
    }
    std::stable_sort(items.begin(), items.end(), [](const SweepItem &a, const SweepItem &b) {
        return (a.s1 - a.s0) > (b.s1 - b.s0);
    });
}

int launch_stage_sum(bool far, const Sweep *s, hipStream_t st) {
    const StageInst *ins = far ? s->d_ff : s->d_jm;
    const SweepItem *items = far ? s->d_items_ff : s->d_items_jm;
    const int n = far ? s->n_items_ff : s->n_items_jm;
    if (n == 0) return 0;
    // More workgroups than fit at once is deliberate: the dispatcher starts
    // each as a slot frees, so with the list sorted longest first a grid of
    // one workgroup per item schedules the items longest-first onto free
    // slots.  Never more workgroups than items.
    const int grid = (int)std::max<long>(1, std::min<long>(n, (long)s->n_cu * s->prm.wgs_per_cu));
    if (far) {
        if (ff_kernel() == PYPO_FF_MATRIX) {
            if (s->prm.tpl == 4) hipLaunchKernelGGL(k_sweep_ff_mfma16, dim3(grid), dim3(TILE), 0, st, ins, items, n);
            else if (s->prm.tpl == 2) hipLaunchKernelGGL(k_sweep_ff_mfma<8>, dim3(grid), dim3(TILE), 0, st, ins, items, n);
            else hipLaunchKernelGGL(k_sweep_ff_mfma<4>, dim3(grid), dim3(TILE), 0, st, ins, items, n);
        } else if (s->prm.tpl == 2) {
            hipLaunchKernelGGL(k_sweep_ff<2>, dim3(grid), dim3(TILE), 0, st, ins, items, n);
        } else {
            hipLaunchKernelGGL(k_sweep_ff<1>, dim3(grid), dim3(TILE), 0, st, ins, items, n);
        }
    } else {
        // the near-field work list is built with one target per lane
        hipLaunchKernelGGL(k_sweep_near<1>, dim3(grid), dim3(TILE), 0, st, ins, items, n);
    }
    PYPO_CHECK_LAUNCH();
    return 0;
}

// The run sequence: JM sum, JM finish, pack dish currents + tile metas,
// FF sum, FF finish.  No memsets (all partial slots are fully overwritten).
int enqueue_run(const Sweep *s, hipStream_t st) {
    int err;
    if ((err = launch_stage_sum(false, s, st))) return err;
    hipLaunchKernelGGL(k_sweep_jm_finish, dim3((s->max_nd + 255) / 256, s->n_inst), dim3(256), 0, st,
                       s->d_jm);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sweep_pack, dim3((s->max_nd + 255) / 256, s->n_inst), dim3(256), 0, st, s->d_pk);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sweep_meta, dim3(s->max_tiles_d, s->n_inst), dim3(TILE), 0, st, s->d_pk);
    PYPO_CHECK_LAUNCH();
    if ((err = launch_stage_sum(true, s, st))) return err;
    hipLaunchKernelGGL(k_sweep_ff_finish, dim3((s->n_ff + 255) / 256, s->n_inst), dim3(256), 0, st,
                       s->d_ff);
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace

// (Re)build work lists, partial buffers and, if asked, the captured graph.
int sweep_plan(Sweep *s, SweepPlanParams prm, bool use_graph, hipStream_t st) {
    // targets per lane: 1, 2, or 4 (16 M-blocks; matrix-core far field only)
    if (prm.tpl != 1 && prm.tpl != 2 && prm.tpl != 4) prm.tpl = 2;
    if (prm.tpl == 4 && ff_kernel() != PYPO_FF_MATRIX) prm.tpl = 2;
    if (prm.chunk_tiles < 1) prm.chunk_tiles = 1 << 20;
    if (prm.wgs_per_cu < 1) prm.wgs_per_cu = 64;
    PYPO_REQUIRE(prm.wgs_per_cu <= 4096, ERR_ARG, "sweep: wgs_per_cu must be 1..4096");
    PYPO_HIP_TRY(hipStreamSynchronize(st));
    PYPO_HIP_TRY(hipDeviceSynchronize());   // no replay of the old plan may still run
    sweep_free_plan(s);
    s->prm = prm;
    int err = 0;
    std::vector<SweepItem> items_jm, items_ff;
    std::vector<int> ch_jm, ch_ff;
    build_items(s->jm, 1, prm.chunk_tiles, items_jm, ch_jm);   // near field: one target per lane
    build_items(s->ff, prm.tpl, prm.chunk_tiles, items_ff, ch_ff);
    size_t bj = 0, bf = 0;
    for (int i = 0; i < s->n_inst; i++) {
        bj += (size_t)ch_jm[i] * 12 * s->jm[i].n_tgt * sizeof(double);
        bf += (size_t)ch_ff[i] * 12 * s->ff[i].n_tgt * sizeof(double);
    }
    if (bj > s->partial_jm_bytes) {
        if (s->partial_jm) {
            (void)hipFree(s->partial_jm);
            s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), (void *)s->partial_jm));
        }
        s->partial_jm = (double *)sweep_alloc(s, bj, &err);
        if (err) return err;
        s->partial_jm_bytes = bj;
    }
    if (bf > s->partial_ff_bytes) {
        if (s->partial_ff) {
            (void)hipFree(s->partial_ff);
            s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), (void *)s->partial_ff));
        }
        s->partial_ff = (double *)sweep_alloc(s, bf, &err);
        if (err) return err;
        s->partial_ff_bytes = bf;
    }
    size_t oj = 0, of = 0;
    for (int i = 0; i < s->n_inst; i++) {
        s->jm[i].partial = s->partial_jm + oj;
        s->jm[i].n_chunks = ch_jm[i];
        oj += (size_t)ch_jm[i] * 12 * s->jm[i].n_tgt;
        s->ff[i].partial = s->partial_ff + of;
        s->ff[i].n_chunks = ch_ff[i];
        of += (size_t)ch_ff[i] * 12 * s->ff[i].n_tgt;
    }
    // work lists live in one allocation per stage, sized for the largest plan
    if (s->d_items_jm) {
        (void)hipFree(s->d_items_jm);
        s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), (void *)s->d_items_jm));
    }
    if (s->d_items_ff) {
        (void)hipFree(s->d_items_ff);
        s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), (void *)s->d_items_ff));
    }
    s->d_items_jm = (SweepItem *)sweep_alloc(s, items_jm.size() * sizeof(SweepItem), &err);
    if (err) return err;
    s->d_items_ff = (SweepItem *)sweep_alloc(s, items_ff.size() * sizeof(SweepItem), &err);
    if (err) return err;
    s->n_items_jm = (int)items_jm.size();
    s->n_items_ff = (int)items_ff.size();
    PYPO_HIP_TRY(hipMemcpy(s->d_items_jm, items_jm.data(), items_jm.size() * sizeof(SweepItem),
                           hipMemcpyHostToDevice));
    PYPO_HIP_TRY(hipMemcpy(s->d_items_ff, items_ff.data(), items_ff.size() * sizeof(SweepItem),
                           hipMemcpyHostToDevice));
    PYPO_HIP_TRY(hipMemcpy(s->d_jm, s->jm.data(), s->jm.size() * sizeof(StageInst), hipMemcpyHostToDevice));
    PYPO_HIP_TRY(hipMemcpy(s->d_ff, s->ff.data(), s->ff.size() * sizeof(StageInst), hipMemcpyHostToDevice));
    if (use_graph) {
        if (!s->cap_stream) PYPO_HIP_TRY(hipStreamCreateWithFlags(&s->cap_stream, hipStreamNonBlocking));
        PYPO_HIP_TRY(hipStreamBeginCapture(s->cap_stream, hipStreamCaptureModeThreadLocal));
        int e2 = enqueue_run(s, s->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e3 = hipStreamEndCapture(s->cap_stream, &g);
        if (e2) {
            if (g) (void)hipGraphDestroy(g);
            return e2;
        }
        if (e3 != hipSuccess) {
            set_error((int)e3, std::string("sweep: graph capture failed: ") + hipGetErrorString(e3));
            return (int)e3;
        }
        s->graph = g;
        PYPO_HIP_TRY(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
    }
    return 0;
}

int sweep_run(Sweep *s, hipStream_t st) {
    if (s->exec) {
        PYPO_HIP_TRY(hipGraphLaunch(s->exec, st));
        return 0;
    }
    return enqueue_run(s, st);
}

void sweep_destroy(Sweep *s) {
    (void)hipDeviceSynchronize();
    sweep_free_plan(s);
    if (s->cap_stream) (void)hipStreamDestroy(s->cap_stream);
    for (void *p : s->allocs) (void)hipFree(p);
    delete s;
}

int sweep_create(const pypo_sweep_desc *d, Sweep **out, hipStream_t st) {
    PYPO_REQUIRE(d && d->n_inst >= 1 && d->feeds && d->dishes && d->k && d->feed_currents, ERR_ARG,
                 "sweep: NULL descriptor field or no instances");
    PYPO_REQUIRE(d->ff.n_cells && d->ff.lxu && d->ff.lyv && d->ff.gcenter && d->ff.coeffs &&
                     d->ff.transf && d->ff.n_cells[0] > 0 && d->ff.n_cells[1] > 0,
                 ERR_ARG, "sweep: invalid far-field grid");
    for (int i = 0; i < d->n_inst; i++) {
        const reflparams &fp = d->feeds[i], &dp = d->dishes[i];
        PYPO_REQUIRE(fp.n_cells && dp.n_cells && fp.n_cells[0] > 0 && fp.n_cells[1] > 0 &&
                         dp.n_cells[0] > 0 && dp.n_cells[1] > 0,
                     ERR_ARG, "sweep: instance grid without cells");
        PYPO_REQUIRE(fp.rms <= 0 && dp.rms <= 0, ERR_ARG, "sweep: rough surfaces are not batched");
        PYPO_REQUIRE(d->k[i] > 0, ERR_ARG, "sweep: wavenumbers must be positive");
        double *const *hc = (double *const *)&d->feed_currents[i];
        for (int q = 0; q < 12; q++) PYPO_REQUIRE(hc[q], ERR_ARG, "sweep: feed current array is NULL");
    }
    Sweep *s = new Sweep();
    int err = 0;
    auto fail = [&](int e) { sweep_destroy(s); return e; };
    s->n_inst = d->n_inst;
    s->tdir = d->t_direction;
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&s->n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        s->n_cu = 256;
    // far-field grid, shared by all instances
    const reflparams &fr = d->ff;
    s->n_ff = fr.n_cells[0] * fr.n_cells[1];
    s->ff_grid = (double *)sweep_alloc(s, 7 * (size_t)s->n_ff * sizeof(double), &err);
    if (err) return fail(err);
    GridPtrs gff{s->ff_grid, s->ff_grid + s->n_ff, s->ff_grid + 2 * (size_t)s->n_ff,
                 s->ff_grid + 3 * (size_t)s->n_ff, s->ff_grid + 4 * (size_t)s->n_ff,
                 s->ff_grid + 5 * (size_t)s->n_ff, s->ff_grid + 6 * (size_t)s->n_ff};
    if ((err = launch_generate_grid(fr, gff, s->n_ff, nullptr, true, true, st))) return fail(err);
    s->ff_out = (double *)sweep_alloc(s, (size_t)s->n_inst * 12 * s->n_ff * sizeof(double), &err);
    if (err) return fail(err);
    s->jm.resize(s->n_inst);
    s->ff.resize(s->n_inst);
    s->pk.resize(s->n_inst);
    s->nf.resize(s->n_inst);
    s->nd.resize(s->n_inst);
    s->dish_cur.resize(s->n_inst);
    for (int i = 0; i < s->n_inst; i++) {
        const reflparams &fp = d->feeds[i], &dp = d->dishes[i];
        const int nf = fp.n_cells[0] * fp.n_cells[1], nd = dp.n_cells[0] * dp.n_cells[1];
        s->nf[i] = nf;
        s->nd[i] = nd;
        s->max_nd = std::max(s->max_nd, nd);
        s->max_tiles_d = std::max(s->max_tiles_d, n_tiles(nd));
        // feed: grid, currents, packed records (once)
        double *fg = (double *)sweep_alloc(s, 7 * (size_t)nf * sizeof(double), &err);
        if (err) return fail(err);
        GridPtrs gf{fg, fg + nf, fg + 2 * (size_t)nf, fg + 3 * (size_t)nf, fg + 4 * (size_t)nf,
                    fg + 5 * (size_t)nf, fg + 6 * (size_t)nf};
        if ((err = launch_generate_grid(fp, gf, nf, nullptr, true, true, st))) return fail(err);
        double *fc = (double *)sweep_alloc(s, 12 * (size_t)nf * sizeof(double), &err);
        if (err) return fail(err);
        double *const *hc = (double *const *)&d->feed_currents[i];
        for (int q = 0; q < 12; q++)
            if (hipMemcpyAsync(fc + q * (size_t)nf, hc[q], nf * sizeof(double), hipMemcpyHostToDevice,
                               st) != hipSuccess)
                return fail((set_error(ERR_ALLOC, "sweep: feed current upload failed"), ERR_ALLOC));
        void *fpk = sweep_alloc(s, packed_bytes(PYPO_MODE_FF, PYPO_PREC_MIXED, nf), &err);
        if (err) return fail(err);
        Bundle12 fb;
        for (int q = 0; q < 12; q++) fb.p[q] = fc + q * (size_t)nf;
        if ((err = launch_pack_sources(PYPO_PREC_MIXED, gf, nf, fb, fp.gmode, fp.n_cells[0],
                                       fp.n_cells[1], fpk, st)))
            return fail(err);
        // dish: grid, currents (JM output), packed records (per run)
        double *dg = (double *)sweep_alloc(s, 7 * (size_t)nd * sizeof(double), &err);
        if (err) return fail(err);
        GridPtrs gd{dg, dg + nd, dg + 2 * (size_t)nd, dg + 3 * (size_t)nd, dg + 4 * (size_t)nd,
                    dg + 5 * (size_t)nd, dg + 6 * (size_t)nd};
        if ((err = launch_generate_grid(dp, gd, nd, nullptr, true, true, st))) return fail(err);
        double *dc = (double *)sweep_alloc(s, 12 * (size_t)nd * sizeof(double), &err);
        if (err) return fail(err);
        s->dish_cur[i] = dc;
        void *dpk = sweep_alloc(s, packed_bytes(PYPO_MODE_FF, PYPO_PREC_MIXED, nd), &err);
        if (err) return fail(err);
        const double k = d->k[i];
        const PhysConsts pc = phys_consts(k, 1.0);
        StageInst j{};
        j.src = (const SrcRec32 *)fpk;
        j.meta = (const TileMeta *)((char *)fpk + meta_offset(sizeof(SrcRec32), nf));
        j.n_src = nf; j.n_tgt = nd; j.k = k; j.tdir = d->t_direction;
        j.rev_scale = d->t_direction * k / kTwoPi; j.pref = pc.prefactor;
        j.tx = gd.x; j.ty = gd.y; j.tz = gd.z; j.tnx = gd.nx; j.tny = gd.ny; j.tnz = gd.nz;
        j.out = dc;
        s->jm[i] = j;
        StageInst f{};
        f.src = (const SrcRec32 *)dpk;
        f.meta = (const TileMeta *)((char *)dpk + meta_offset(sizeof(SrcRec32), nd));
        f.n_src = nd; f.n_tgt = s->n_ff; f.k = k; f.tdir = d->t_direction;
        f.rev_scale = -d->t_direction * k / kTwoPi; f.pref = pc.prefactor;
        f.tx = gff.x; f.ty = gff.y; f.tz = gff.z;
        f.out = s->ff_out + (size_t)i * 12 * s->n_ff;
        s->ff[i] = f;
        PackInst p{};
        p.grid = gd; p.cur = dc; p.n = nd; p.gmode = dp.gmode; p.ncx = dp.n_cells[0]; p.ncy = dp.n_cells[1];
        p.rec = (SrcRec32 *)dpk;
        p.meta = (TileMeta *)((char *)dpk + meta_offset(sizeof(SrcRec32), nd));
        s->pk[i] = p;
        s->pairs_jm += (double)nf * nd;
        s->pairs_ff += (double)nd * s->n_ff;
    }
    s->d_jm = (StageInst *)sweep_alloc(s, s->n_inst * sizeof(StageInst), &err);
    if (err) return fail(err);
    s->d_ff = (StageInst *)sweep_alloc(s, s->n_inst * sizeof(StageInst), &err);
    if (err) return fail(err);
    s->d_pk = (PackInst *)sweep_alloc(s, s->n_inst * sizeof(PackInst), &err);
    if (err) return fail(err);
    if (hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(s->d_pk, s->pk.data(), s->n_inst * sizeof(PackInst), hipMemcpyHostToDevice) != hipSuccess)
        return fail((set_error(ERR_ALLOC, "sweep: set-up copy failed"), ERR_ALLOC));
    SweepPlanParams prm{d->tpl, d->chunk_tiles, d->wgs_per_cu};
    if ((err = sweep_plan(s, prm, d->use_graph != 0, st))) return fail(err);
    *out = s;
    return 0;
}

int sweep_read(const Sweep *s, int inst, c2Bundle *host, hipStream_t st) {
    PYPO_REQUIRE(inst >= 0 && inst < s->n_inst, ERR_ARG, "sweep: instance index out of range");
    double *const *h = (double *const *)host;
    for (int q = 0; q < 12; q++) {
        PYPO_REQUIRE(h[q], ERR_ARG, "sweep: result array is NULL");
        PYPO_HIP_TRY(hipMemcpyAsync(h[q], s->ff_out + ((size_t)inst * 12 + q) * s->n_ff,
                                    s->n_ff * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    PYPO_HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int sweep_read_currents(const Sweep *s, int inst, c2Bundle *host, hipStream_t st) {
    PYPO_REQUIRE(inst >= 0 && inst < s->n_inst, ERR_ARG, "sweep: instance index out of range");
    double *const *h = (double *const *)host;
    const int nd = s->nd[inst];
    for (int q = 0; q < 12; q++) {
        PYPO_REQUIRE(h[q], ERR_ARG, "sweep: result array is NULL");
        PYPO_HIP_TRY(hipMemcpyAsync(h[q], s->dish_cur[inst] + (size_t)q * nd, nd * sizeof(double),
                                    hipMemcpyDeviceToHost, st));
    }
    PYPO_HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

void sweep_info(const Sweep *s, double *pairs_jm, double *pairs_ff, int *n_items_jm, int *n_items_ff,
                double **ff_out) {
    if (pairs_jm) *pairs_jm = s->pairs_jm;
    if (pairs_ff) *pairs_ff = s->pairs_ff;
    if (n_items_jm) *n_items_jm = s->n_items_jm;
    if (n_items_ff) *n_items_ff = s->n_items_ff;
    if (ff_out) *ff_out = s->ff_out;
}

}  // namespace pypo
