// po_scalar.hip -- scalar PO sum (fieldScalarAtPoint, Propagation.h:1279-1331):
//   u(p) = sum_s -k^2/(4 pi r) w A u_s exp(-j t k r),  no shadow test.
// The -k^2/(4 pi) factor is applied by the finish kernel.  Partial layout:
// partial[(split*2 + {re, im}) * n_tgt + t].

#include "po_common.hpp"

namespace pypo {
namespace {

template <int TPL>
__global__ __launch_bounds__(256) void k_scalar_mixed(const SrcRecS32 *__restrict__ src, int n_src,
                                                      int chunk, const double *__restrict__ tx,
                                                      const double *__restrict__ ty,
                                                      const double *__restrict__ tz, int n_tgt,
                                                      double rev_scale, double *__restrict__ partial) {
    __shared__ SrcRecS32 tile[TILE];
    const int tid = threadIdx.x;
    int tb, sp;
    xcd_item(tb, sp);
    double px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = tb * (TILE * TPL) + j * TILE + tid;
        px[j] = (t < n_tgt) ? tx[t] : 1.0e6;
        py[j] = (t < n_tgt) ? ty[t] : 1.0e6;
        pz[j] = (t < n_tgt) ? tz[t] : 1.0e6;
    }
    float ar[TPL], ai[TPL];
    double tr[TPL], ti[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) { ar[j] = ai[j] = 0.f; tr[j] = ti[j] = 0.0; }
    const int s_begin = sp * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<2> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        __syncthreads();
        ld.store(reinterpret_cast<uint4 *>(tile), tid);
        __syncthreads();
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
        const int n = min(TILE, s_end - base);   // padding records add exactly zero: skip them
#pragma unroll 2
        for (int s = 0; s < n; s++) {
            const SrcRecS32 &r = tile[s];
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                double dx = px[j] - r.x, dy = py[j] - r.y, dz = pz[j] - r.z;
                double R2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                float y0 = __builtin_amdgcn_rsqf((float)R2);
                double yd = (double)y0;
                double tR = R2 * yd;
                double e = __builtin_fma(-tR, yd, 1.0);
                double R = __builtin_fma(0.5 * tR, e, tR);
                float f = (float)__builtin_amdgcn_fract(R * rev_scale);
                float sn, cs;
                sincos_rev_f32(f, sn, cs);
                float Rinv = __builtin_fmaf(0.5f * y0, (float)e, y0);
                float ur = r.ur * Rinv, ui = r.ui * Rinv;
                ar[j] = __builtin_fmaf(ur, cs, ar[j]);
                ar[j] = __builtin_fmaf(-ui, sn, ar[j]);
                ai[j] = __builtin_fmaf(ur, sn, ai[j]);
                ai[j] = __builtin_fmaf(ui, cs, ai[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < TPL; j++) { tr[j] += ar[j]; ti[j] += ai[j]; ar[j] = ai[j] = 0.f; }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = tb * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
            partial[((size_t)sp * 2) * n_tgt + t] = tr[j];
            partial[((size_t)sp * 2 + 1) * n_tgt + t] = ti[j];
        }
    }
}

template <int TPL>
__global__ __launch_bounds__(256) void k_scalar_fp64(const SrcRecS64 *__restrict__ src, int n_src,
                                                     int chunk, const double *__restrict__ tx,
                                                     const double *__restrict__ ty,
                                                     const double *__restrict__ tz, int n_tgt,
                                                     double rev_scale, double *__restrict__ partial) {
    __shared__ SrcRecS64 tile[TILE];
    const int tid = threadIdx.x;
    int tb, sp;
    xcd_item(tb, sp);
    double px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = tb * (TILE * TPL) + j * TILE + tid;
        px[j] = (t < n_tgt) ? tx[t] : 1.0e6;
        py[j] = (t < n_tgt) ? ty[t] : 1.0e6;
        pz[j] = (t < n_tgt) ? tz[t] : 1.0e6;
    }
    double ar[TPL], ai[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) { ar[j] = ai[j] = 0.0; }
    const int s_begin = sp * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<3> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        __syncthreads();
        ld.store(reinterpret_cast<uint4 *>(tile), tid);
        __syncthreads();
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
        const int n = min(TILE, s_end - base);
#pragma unroll 1
        for (int s = 0; s < n; s++) {
            const SrcRecS64 &r = tile[s];
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                double dx = px[j] - r.x, dy = py[j] - r.y, dz = pz[j] - r.z;
                double R = sqrt(__builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx)));
                double f = __builtin_amdgcn_fract(R * rev_scale);
                double sn, cs;
                sincos_rev_f64(f, sn, cs);
                double Rinv = 1.0 / R;
                double ur = r.ur * Rinv, ui = r.ui * Rinv;
                ar[j] = __builtin_fma(ur, cs, ar[j]);
                ar[j] = __builtin_fma(-ui, sn, ar[j]);
                ai[j] = __builtin_fma(ur, sn, ai[j]);
                ai[j] = __builtin_fma(ui, cs, ai[j]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = tb * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
            partial[((size_t)sp * 2) * n_tgt + t] = ar[j];
            partial[((size_t)sp * 2 + 1) * n_tgt + t] = ai[j];
        }
    }
}

__global__ __launch_bounds__(256) void k_scalar_finish(const double *__restrict__ partial, int splits,
                                                       int n_tgt, double scale,
                                                       double *__restrict__ out_re,
                                                       double *__restrict__ out_im) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tgt) return;
    double re = 0, im = 0;
    for (int s = 0; s < splits; s++) {
        re += partial[((size_t)s * 2) * n_tgt + t];
        im += partial[((size_t)s * 2 + 1) * n_tgt + t];
    }
    out_re[t] = scale * re;
    out_im[t] = scale * im;
}

}  // namespace

int launch_scalar_sum(int precision, const SumPlan &p, const void *packed, int n_src,
                      const GridPtrs &ct, int n_tgt, double rev_scale, double *partial,
                      hipStream_t st) {
    dim3 grid(p.tblocks, p.splits), block(TILE);
    if (precision == PYPO_PREC_FP64)
        hipLaunchKernelGGL(k_scalar_fp64<1>, grid, block, 0, st, (const SrcRecS64 *)packed, n_src,
                           p.chunk, ct.x, ct.y, ct.z, n_tgt, rev_scale, partial);
    else
        hipLaunchKernelGGL(k_scalar_mixed<1>, grid, block, 0, st, (const SrcRecS32 *)packed, n_src,
                           p.chunk, ct.x, ct.y, ct.z, n_tgt, rev_scale, partial);
    PYPO_CHECK_LAUNCH();
    return 0;
}

int launch_scalar_finish(const SumPlan &p, int n_tgt, double scale, double *re, double *im,
                         const double *partial, hipStream_t st) {
    hipLaunchKernelGGL(k_scalar_finish, dim3((n_tgt + 255) / 256), dim3(256), 0, st, partial,
                       p.splits, n_tgt, scale, re, im);
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
