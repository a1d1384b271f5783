// po_pack.hip -- source packing for the PO sums.
//
// One elementwise pass turns the source grid (SoA FP64, generated on the
// device) and the source currents (c2Bundle SoA FP64) into fixed-size
// records: position, normal and w*A*currents, where w is the trapezoid
// weight of Propagation.h:1123-1135 (half on xu edges; half on yv edges
// unless the grid is uv) and A the area element.  A second pass writes one
// TileMeta per 256 records (normal cone, any non-zero M).  HBM-bound:
// 152 B read and 96 B (FP32 records) or 144 B (FP64) written per source.
//
// Records go out in serpentine order (grid rows alternately reversed): the
// sums do not depend on the source order, and this way consecutive records
// are grid neighbours across row ends too, which the matrix far field's
// follower phase needs for groups of four that span a row end.

#include "po_common.hpp"

namespace pypo {
namespace {

__global__ __launch_bounds__(256) void k_pack32(GridPtrs cs, int n, Bundle12 cur, int gmode,
                                                int ncx, int ncy, SrcRec32 *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double wa = trapezoid_weight(i, gmode, ncx, ncy) * cs.area[i];
    SrcRec32 r;
    r.x = cs.x[i]; r.y = cs.y[i]; r.z = cs.z[i];
    r.nx = (float)cs.nx[i]; r.ny = (float)cs.ny[i]; r.nz = (float)cs.nz[i]; r.pad0 = 0.f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r.cj[2 * a] = (float)(wa * cur.p[a][i]);           // J re
        r.cj[2 * a + 1] = (float)(wa * cur.p[3 + a][i]);   // J im
        r.cm[2 * a] = (float)(wa * cur.p[6 + a][i]);       // M re
        r.cm[2 * a + 1] = (float)(wa * cur.p[9 + a][i]);   // M im
    }
    r.pad1[0] = 0.f; r.pad1[1] = 0.f;
    out[packed_slot(i, n, ncx, ncy)] = r;
}

__global__ __launch_bounds__(256) void k_pack64(GridPtrs cs, int n, Bundle12 cur, int gmode,
                                                int ncx, int ncy, SrcRec64 *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double wa = trapezoid_weight(i, gmode, ncx, ncy) * cs.area[i];
    SrcRec64 r;
    r.x = cs.x[i]; r.y = cs.y[i]; r.z = cs.z[i];
    r.nx = cs.nx[i]; r.ny = cs.ny[i]; r.nz = cs.nz[i];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r.c[2 * a] = wa * cur.p[a][i];
        r.c[2 * a + 1] = wa * cur.p[3 + a][i];
        r.c[6 + 2 * a] = wa * cur.p[6 + a][i];
        r.c[7 + 2 * a] = wa * cur.p[9 + a][i];
    }
    out[packed_slot(i, n, ncx, ncy)] = r;
}

__global__ __launch_bounds__(256) void k_pack_s32(GridPtrs cs, int n, const double *__restrict__ ur,
                                                  const double *__restrict__ ui, int gmode, int ncx,
                                                  int ncy, SrcRecS32 *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double wa = trapezoid_weight(i, gmode, ncx, ncy) * cs.area[i];
    SrcRecS32 r;
    r.x = cs.x[i]; r.y = cs.y[i]; r.z = cs.z[i];
    r.ur = (float)(wa * ur[i]); r.ui = (float)(wa * ui[i]);
    out[packed_slot(i, n, ncx, ncy)] = r;
}

__global__ __launch_bounds__(256) void k_pack_s64(GridPtrs cs, int n, const double *__restrict__ ur,
                                                  const double *__restrict__ ui, int gmode, int ncx,
                                                  int ncy, SrcRecS64 *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double wa = trapezoid_weight(i, gmode, ncx, ncy) * cs.area[i];
    SrcRecS64 r;
    r.x = cs.x[i]; r.y = cs.y[i]; r.z = cs.z[i];
    r.ur = wa * ur[i]; r.ui = wa * ui[i]; r.pad = 0.0;
    out[packed_slot(i, n, ncx, ncy)] = r;
}

template <class Rec>
__device__ __forceinline__ void rec_normal_m(const Rec &r, float &nx, float &ny, float &nz, bool &m,
                                             float &cmax);

template <>
__device__ __forceinline__ void rec_normal_m<SrcRec32>(const SrcRec32 &r, float &nx, float &ny,
                                                       float &nz, bool &m, float &cmax) {
    nx = r.nx; ny = r.ny; nz = r.nz;
    m = false;
    cmax = 0.f;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        m |= (r.cm[q] != 0.f);
        cmax = fmaxf(cmax, fmaxf(fabsf(r.cj[q]), fabsf(r.cm[q])));
    }
}

template <>
__device__ __forceinline__ void rec_normal_m<SrcRec64>(const SrcRec64 &r, float &nx, float &ny,
                                                       float &nz, bool &m, float &cmax) {
    nx = (float)r.nx; ny = (float)r.ny; nz = (float)r.nz;
    m = false;
    cmax = 0.f;
#pragma unroll
    for (int q = 0; q < 12; q++) {
        if (q >= 6) m |= (r.c[q] != 0.0);
        cmax = fmaxf(cmax, (float)fabs(r.c[q]));
    }
}

// One workgroup per tile of 256 records.
template <class Rec>
__global__ __launch_bounds__(256) void k_tile_meta(const Rec *__restrict__ recs, int n,
                                                   TileMeta *__restrict__ meta) {
    __shared__ float scratch[4];
    const int i = blockIdx.x * TILE + threadIdx.x;
    const bool valid = i < n;
    float nx = 0.f, ny = 0.f, nz = 1.f, cm = 0.f;
    bool m = false;
    if (valid) rec_normal_m(recs[i], nx, ny, nz, m, cm);
    Cone c = block_cone(nx, ny, nz, valid, scratch);
    float any_m = block_sum(m ? 1.f : 0.f, scratch);
    float cnt = block_sum(valid ? 1.f : 0.f, scratch);
    float cmax = -block_min(-cm, scratch);
    if (threadIdx.x == 0) {
        TileMeta t;
        t.ax = c.ax; t.ay = c.ay; t.az = c.az; t.cos_half = c.cos_half;
        t.has_m = any_m > 0.f ? 1 : 0;
        t.n_valid = (int)cnt;
        t.cmax = cmax;
        t.pad1 = 0;
        meta[blockIdx.x] = t;
    }
}

int record_bytes(int mode, int precision) {
    if (mode == PYPO_MODE_SCALAR) return precision == PYPO_PREC_FP64 ? 48 : 32;
    return precision == PYPO_PREC_FP64 ? 144 : 96;
}

}  // namespace

size_t packed_bytes(int mode, int precision, int n_src) {
    return meta_offset(record_bytes(mode, precision), n_src) + (size_t)n_tiles(n_src) * sizeof(TileMeta);
}

int launch_pack_sources(int precision, const GridPtrs &cs, int n_src, const Bundle12 &cur,
                        int gmode, int ncx, int ncy, void *packed, hipStream_t st) {
    PYPO_REQUIRE(n_src > 0 && (long long)ncx * ncy == n_src, ERR_ARG,
                 "source n_cells do not match the source size");
    int blocks = (n_src + 255) / 256;
    TileMeta *meta = (TileMeta *)((char *)packed +
                                  meta_offset(record_bytes(PYPO_MODE_FF, precision), n_src));
    if (precision == PYPO_PREC_FP64) {
        hipLaunchKernelGGL(k_pack64, dim3(blocks), dim3(256), 0, st, cs, n_src, cur, gmode, ncx, ncy,
                           (SrcRec64 *)packed);
        PYPO_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_tile_meta<SrcRec64>, dim3(n_tiles(n_src)), dim3(TILE), 0, st,
                           (const SrcRec64 *)packed, n_src, meta);
    } else {
        hipLaunchKernelGGL(k_pack32, dim3(blocks), dim3(256), 0, st, cs, n_src, cur, gmode, ncx, ncy,
                           (SrcRec32 *)packed);
        PYPO_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_tile_meta<SrcRec32>, dim3(n_tiles(n_src)), dim3(TILE), 0, st,
                           (const SrcRec32 *)packed, n_src, meta);
    }
    PYPO_CHECK_LAUNCH();
    return 0;
}

int launch_pack_scalar(int precision, const GridPtrs &cs, int n_src, const double *ur,
                       const double *ui, int gmode, int ncx, int ncy, void *packed,
                       hipStream_t st) {
    PYPO_REQUIRE(n_src > 0 && (long long)ncx * ncy == n_src, ERR_ARG,
                 "source n_cells do not match the source size");
    int blocks = (n_src + 255) / 256;
    if (precision == PYPO_PREC_FP64)
        hipLaunchKernelGGL(k_pack_s64, dim3(blocks), dim3(256), 0, st, cs, n_src, ur, ui, gmode, ncx,
                           ncy, (SrcRecS64 *)packed);
    else
        hipLaunchKernelGGL(k_pack_s32, dim3(blocks), dim3(256), 0, st, cs, n_src, ur, ui, gmode, ncx,
                           ncy, (SrcRecS32 *)packed);
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
