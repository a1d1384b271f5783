// hybrid.hip -- the hybrid ray-trace / PO step of System.runHybridPropagation
// (src/PyPO/System.py:1570-1622) on the device:
//
//  * k_hybrid_phase: each ray carries its field sample from fr_in to fr_out;
//    the sample is multiplied by exp(j k L1) (L1 = |p_out - p_in|), times
//    sqrt(L0 / (2 L1 + L0)) when the rays start from a point s (L0 =
//    |p_in - s|) -- calcRayLen, System.py:1304-1372, and System.py:1599-1603.
//  * nearest-neighbour interpolation of the carried field from the ray hit
//    points onto the target surface's grid (interpFrame, System.py:1624-1692,
//    which calls scipy.interpolate.griddata(method="nearest")): hit points are
//    binned into a uniform grid over their bounding box, and each grid point
//    searches Chebyshev rings of cells around its own cell until no unsearched
//    cell can hold a closer point.  Ties go to the smallest point index.
//
// All FP64; every kernel is HBM-bound or latency-bound (O(N) work).

#include <hipcub/hipcub.hpp>

#include "common.hpp"

namespace pypo {
namespace {

// ---- phase transport ------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hybrid_phase(const double *__restrict__ xi, const double *__restrict__ yi,
                                                      const double *__restrict__ zi, const double *__restrict__ xo,
                                                      const double *__restrict__ yo, const double *__restrict__ zo,
                                                      Bundle12 fin, Bundle12 fout, int n, double k, int has_start,
                                                      double sx, double sy, double sz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dx = xo[i] - xi[i], dy = yo[i] - yi[i], dz = zo[i] - zi[i];
    const double L1 = sqrt(dx * dx + dy * dy + dz * dz);
    double amp = 1.0;
    if (has_start) {
        const double ax = xi[i] - sx, ay = yi[i] - sy, az = zi[i] - sz;
        const double L0 = sqrt(ax * ax + ay * ay + az * az);
        amp = sqrt(L0 / (2 * L1 + L0));
    }
    double s, c;
    sincos(k * L1, &s, &c);
    const double er = c * amp, ei = s * amp;
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double vr = fin.p[6 * f + a][i], vi = fin.p[6 * f + 3 + a][i];
            fout.p[6 * f + a][i] = er * vr - ei * vi;
            fout.p[6 * f + 3 + a][i] = er * vi + ei * vr;
        }
}

// ---- nearest-neighbour interpolation ----------------------------------------------
struct NNGrid {
    double lo[3];
    double h;
    int dims[3];
    int pad;
};

constexpr int kBBoxThreads = 256;

// Per-block min/max of the three coordinates: partial[block][6].
__global__ __launch_bounds__(kBBoxThreads) void k_bbox_partial(const double *__restrict__ x,
                                                               const double *__restrict__ y,
                                                               const double *__restrict__ z, int n,
                                                               double *__restrict__ partial) {
    __shared__ double sh[6][kBBoxThreads];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double p[3] = {x[i], y[i], z[i]};
        for (int a = 0; a < 3; a++) {
            v[a] = fmin(v[a], p[a]);
            v[3 + a] = fmax(v[3 + a], p[a]);
        }
    }
    for (int q = 0; q < 6; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int s = kBBoxThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 6; q++)
                sh[q][threadIdx.x] = q < 3 ? fmin(sh[q][threadIdx.x], sh[q][threadIdx.x + s])
                                           : fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

// One thread: bounding box -> cell size and grid dimensions.  The cell size
// aims at about two points per cell for points on a surface (the two largest
// extents), falls back to a volume estimate when the third extent is not
// small, and grows until the cell count fits max_cells.
__global__ void k_grid_setup(const double *__restrict__ partial, int nblocks, int n, int max_cells,
                             NNGrid *__restrict__ g) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < nblocks; b++)
        for (int a = 0; a < 3; a++) {
            lo[a] = fmin(lo[a], partial[b * 6 + a]);
            hi[a] = fmax(hi[a], partial[b * 6 + 3 + a]);
        }
    double e[3];
    for (int a = 0; a < 3; a++) e[a] = isfinite(hi[a] - lo[a]) ? hi[a] - lo[a] : 0.0;
    double s[3] = {e[0], e[1], e[2]};   // sorted descending
    for (int a = 0; a < 2; a++)
        for (int b = a + 1; b < 3; b++)
            if (s[b] > s[a]) { const double t = s[a]; s[a] = s[b]; s[b] = t; }
    const double target = fmax(1.0, 0.5 * n);
    double h;
    if (s[0] <= 0.0) h = 1.0;
    else if (s[1] <= 0.0) h = s[0] / target;
    else {
        h = sqrt(s[0] * s[1] / target);
        if (s[2] > h) h = cbrt(s[0] * s[1] * s[2] / target);
    }
    // grow h until the cell count fits (h -> inf gives one cell, so this ends)
    int dims[3] = {1, 1, 1};
    for (int it = 0; it < 4096; it++) {
        double cells = 1.0;
        for (int a = 0; a < 3; a++) {
            const double d = fmin(floor(e[a] / h) + 1.0, 1.0e9);
            dims[a] = (int)(d >= 1.0 ? d : 1.0);
            cells *= dims[a];
        }
        if (cells <= (double)max_cells) break;
        dims[0] = dims[1] = dims[2] = 1;
        h *= 2.0;
    }
    for (int a = 0; a < 3; a++) {
        g->lo[a] = isfinite(lo[a]) ? lo[a] : 0.0;
        g->dims[a] = dims[a];
    }
    g->h = h;
    g->pad = 0;
}

// Cell index along one axis, clamped into [0, dim); NaN goes to cell 0.
__device__ __forceinline__ int cell_coord(double v, double lo, double h, int dim) {
    const double c = floor((v - lo) / h);
    if (!(c >= 0.0)) return 0;
    return c >= (double)dim ? dim - 1 : (int)c;
}

__device__ __forceinline__ int cell_of(const NNGrid &g, double x, double y, double z) {
    const int cx = cell_coord(x, g.lo[0], g.h, g.dims[0]);
    const int cy = cell_coord(y, g.lo[1], g.h, g.dims[1]);
    const int cz = cell_coord(z, g.lo[2], g.h, g.dims[2]);
    return (cx * g.dims[1] + cy) * g.dims[2] + cz;
}

__global__ __launch_bounds__(256) void k_fill_int(int *__restrict__ a, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ __launch_bounds__(256) void k_count(const double *__restrict__ x, const double *__restrict__ y,
                                               const double *__restrict__ z, int n,
                                               const NNGrid *__restrict__ gp, int *__restrict__ counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const NNGrid g = *gp;
    atomicAdd(&counts[cell_of(g, x[i], y[i], z[i])], 1);
}

__global__ __launch_bounds__(256) void k_scatter(const double *__restrict__ x, const double *__restrict__ y,
                                                 const double *__restrict__ z, int n,
                                                 const NNGrid *__restrict__ gp, const int *__restrict__ offsets,
                                                 int *__restrict__ cursor, int *__restrict__ idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const NNGrid g = *gp;
    const int c = cell_of(g, x[i], y[i], z[i]);
    idx[offsets[c] + atomicAdd(&cursor[c], 1)] = i;
}

// One thread per query point.  After the block of cells within Chebyshev
// distance r of the query's cell is searched, every other point lies beyond
// one of the block's open faces; the distance from the query to the slab
// behind a face (which spans the whole grid along the other two axes) bounds
// those points from below.  The search stops when the best squared distance
// is below that bound (shrunk by 1e-12 against rounding in the cell
// assignment) or every cell has been searched.
__global__ __launch_bounds__(256) void k_nn_query(const double *__restrict__ qx, const double *__restrict__ qy,
                                                  const double *__restrict__ qz, int m,
                                                  const double *__restrict__ px, const double *__restrict__ py,
                                                  const double *__restrict__ pz, const NNGrid *__restrict__ gp,
                                                  const int *__restrict__ offsets, const int *__restrict__ idx,
                                                  Bundle12 vin, Bundle12 vout, int nvals) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const NNGrid g = *gp;
    const double q[3] = {qx[t], qy[t], qz[t]};
    const int c[3] = {cell_coord(q[0], g.lo[0], g.h, g.dims[0]), cell_coord(q[1], g.lo[1], g.h, g.dims[1]),
                      cell_coord(q[2], g.lo[2], g.h, g.dims[2])};
    int rmax = 0;
    for (int a = 0; a < 3; a++) rmax = max(rmax, max(c[a], g.dims[a] - 1 - c[a]));
    double best = INFINITY;
    int best_i = -1;
    for (int r = 0; r <= rmax; r++) {
        const int x0 = max(0, c[0] - r), x1 = min(g.dims[0] - 1, c[0] + r);
        const int y0 = max(0, c[1] - r), y1 = min(g.dims[1] - 1, c[1] + r);
        const int z0 = max(0, c[2] - r), z1 = min(g.dims[2] - 1, c[2] + r);
        for (int ix = x0; ix <= x1; ix++)
            for (int iy = y0; iy <= y1; iy++)
                for (int iz = z0; iz <= z1; iz++) {
                    const int cheb = max(abs(ix - c[0]), max(abs(iy - c[1]), abs(iz - c[2])));
                    if (cheb != r) continue;   // inner rings are done
                    const int cell = (ix * g.dims[1] + iy) * g.dims[2] + iz;
                    for (int s = offsets[cell]; s < offsets[cell + 1]; s++) {
                        const int j = idx[s];
                        const double dx = px[j] - q[0], dy = py[j] - q[1], dz = pz[j] - q[2];
                        const double d = dx * dx + dy * dy + dz * dz;
                        if (d < best || (d == best && j < best_i)) {
                            best = d;
                            best_i = j;
                        }
                    }
                }
        if (best_i < 0) continue;
        double glo[3], ghi[3], out2[3];
        for (int a = 0; a < 3; a++) {
            glo[a] = g.lo[a];
            ghi[a] = g.lo[a] + g.dims[a] * g.h;
            const double o = fmax(0.0, fmax(glo[a] - q[a], q[a] - ghi[a]));
            out2[a] = o * o;
        }
        double bound = INFINITY;
        for (int a = 0; a < 3; a++) {
            const double side = out2[(a + 1) % 3] + out2[(a + 2) % 3];
            if (c[a] - r > 0) {
                const double d = fmax(0.0, q[a] - (glo[a] + (c[a] - r) * g.h));
                bound = fmin(bound, side + d * d);
            }
            if (c[a] + r < g.dims[a] - 1) {
                const double d = fmax(0.0, glo[a] + (c[a] + r + 1) * g.h - q[a]);
                bound = fmin(bound, side + d * d);
            }
        }
        if (best < bound * (1.0 - 1e-12)) break;   // bound = inf: every cell searched
    }
    // best_i < 0 only when no point lies at a finite distance (NaN
    // coordinates in the query or in every point): the value is NaN
    if (best_i < 0) {
        for (int v = 0; v < nvals; v++) vout.p[v][t] = NAN;
        return;
    }
    for (int v = 0; v < nvals; v++) vout.p[v][t] = vin.p[v][best_i];
}

}  // namespace

int launch_hybrid_phase(const cframe &in, const cframe &out, const Bundle12 &fin, const Bundle12 &fout,
                        int n, double k, const double *start, hipStream_t st) {
    PYPO_REQUIRE(n > 0, ERR_ARG, "hybrid: empty frame");
    hipLaunchKernelGGL(k_hybrid_phase, dim3((n + 255) / 256), dim3(256), 0, st, in.x, in.y, in.z, out.x,
                       out.y, out.z, fin, fout, n, k, start ? 1 : 0, start ? start[0] : 0.0,
                       start ? start[1] : 0.0, start ? start[2] : 0.0);
    PYPO_CHECK_LAUNCH();
    return 0;
}

namespace {
struct NNPlan {
    int nblocks, max_cells;
    size_t off_partial, off_grid, off_counts, off_offsets, off_cursor, off_idx, off_temp, temp_bytes, total;
};

NNPlan nn_plan(int n, int *err) {
    NNPlan p{};
    p.nblocks = std::min(1024, (n + kBBoxThreads - 1) / kBBoxThreads);
    p.max_cells = 2 * n + 8;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    size_t o = 0;
    p.off_partial = o; o += up((size_t)p.nblocks * 6 * sizeof(double));
    p.off_grid = o; o += up(sizeof(NNGrid));
    p.off_counts = o; o += up((size_t)(p.max_cells + 1) * sizeof(int));
    p.off_offsets = o; o += up((size_t)(p.max_cells + 1) * sizeof(int));
    p.off_cursor = o; o += up((size_t)p.max_cells * sizeof(int));
    p.off_idx = o; o += up((size_t)n * sizeof(int));
    size_t tb = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (int *)nullptr, (int *)nullptr,
                                                    p.max_cells + 1);
    if (e != hipSuccess) *err = (int)e;
    p.off_temp = o; p.temp_bytes = up(tb); o += p.temp_bytes;
    p.total = o;
    return p;
}
}  // namespace

size_t nn_workspace_bytes(int n_points) {
    int err = 0;
    return nn_plan(n_points, &err).total;
}

int launch_interp_nearest(const double *px, const double *py, const double *pz, int n, const Bundle12 &vin,
                          int nvals, const double *qx, const double *qy, const double *qz, int m,
                          const Bundle12 &vout, void *ws, size_t ws_bytes, hipStream_t st) {
    PYPO_REQUIRE(n > 0 && m > 0, ERR_ARG, "interpolation: empty point set or grid");
    PYPO_REQUIRE(nvals >= 1 && nvals <= 12, ERR_ARG, "interpolation: 1 to 12 value arrays");
    PYPO_REQUIRE(n <= (1 << 29), ERR_ARG, "interpolation: too many points");
    int err = 0;
    const NNPlan p = nn_plan(n, &err);
    PYPO_REQUIRE(err == 0, err, "interpolation: scan workspace query failed");
    PYPO_REQUIRE(ws_bytes >= p.total, ERR_ARG, "interpolation: workspace too small");
    char *w = (char *)ws;
    double *partial = (double *)(w + p.off_partial);
    NNGrid *grid = (NNGrid *)(w + p.off_grid);
    int *counts = (int *)(w + p.off_counts), *offsets = (int *)(w + p.off_offsets);
    int *cursor = (int *)(w + p.off_cursor), *idx = (int *)(w + p.off_idx);
    const dim3 b256(256);
    hipLaunchKernelGGL(k_bbox_partial, dim3(p.nblocks), dim3(kBBoxThreads), 0, st, px, py, pz, n, partial);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(1), 0, st, partial, p.nblocks, n, p.max_cells, grid);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fill_int, dim3((p.max_cells + 1 + 255) / 256), b256, 0, st, counts, p.max_cells + 1, 0);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fill_int, dim3((p.max_cells + 255) / 256), b256, 0, st, cursor, p.max_cells, 0);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_count, dim3((n + 255) / 256), b256, 0, st, px, py, pz, n, (const NNGrid *)grid, counts);
    PYPO_CHECK_LAUNCH();
    size_t tb = p.temp_bytes;
    PYPO_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(w + p.off_temp, tb, counts, offsets, p.max_cells + 1, st));
    hipLaunchKernelGGL(k_scatter, dim3((n + 255) / 256), b256, 0, st, px, py, pz, n, (const NNGrid *)grid,
                       (const int *)offsets, cursor, idx);
    PYPO_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_nn_query, dim3((m + 255) / 256), b256, 0, st, qx, qy, qz, m, px, py, pz,
                       (const NNGrid *)grid, (const int *)offsets, (const int *)idx, vin, vout, nvals);
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
