// po_ff_mfma.hip -- launchers of the matrix-core far-field sum (the kernel
// body and its description are in po_ff_mfma.hpp): grid = target blocks of
// 64 MB targets x source splits.

#include "po_ff_mfma.hpp"

namespace pypo {
namespace {

// Both shapes are held to two waves per SIMD (256 registers): one wave's
// MFMAs then overlap the other's vector work.  8 M-blocks amortise the
// operand loads over twice the targets.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_ff_mfma4(
    const SrcRec32 *__restrict__ src, const TileMeta *__restrict__ meta, int n_src, int chunk,
    const double *__restrict__ th, const double *__restrict__ ph, int n_tgt, double rev_scale,
    double *__restrict__ partial) {
    int tb, sp;
    xcd_item(tb, sp);
    const int s0 = sp * chunk;
    mfma::ff_mfma_body<4>(src, meta, s0, min(n_src, s0 + chunk), tb * 256, th, ph, n_tgt, rev_scale,
                          partial + (size_t)sp * 12 * n_tgt);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_ff_mfma8(
    const SrcRec32 *__restrict__ src, const TileMeta *__restrict__ meta, int n_src, int chunk,
    const double *__restrict__ th, const double *__restrict__ ph, int n_tgt, double rev_scale,
    double *__restrict__ partial) {
    int tb, sp;
    xcd_item(tb, sp);
    const int s0 = sp * chunk;
    mfma::ff_mfma_body<8>(src, meta, s0, min(n_src, s0 + chunk), tb * 512, th, ph, n_tgt, rev_scale,
                          partial + (size_t)sp * 12 * n_tgt);
}

}  // namespace

int ff_mfma_wgs_per_cu(int tpl) {
    static const int n8 = blocks_per_cu(reinterpret_cast<const void *>(k_ff_mfma8));
    static const int n4 = blocks_per_cu(reinterpret_cast<const void *>(k_ff_mfma4));
    return tpl == 2 ? n8 : n4;
}

int launch_ff_mfma(const SumPlan &p, const void *packed, int n_src, const GridPtrs &ct, int n_tgt,
                   double rev_scale, double *partial, hipStream_t st) {
    const TileMeta *meta = (const TileMeta *)((const char *)packed + meta_offset((int)sizeof(SrcRec32), n_src));
    dim3 grid(p.tblocks, p.splits), block(TILE);
    if (p.tpl == 2)
        hipLaunchKernelGGL(k_ff_mfma8, grid, block, 0, st, (const SrcRec32 *)packed, meta, n_src, p.chunk,
                           ct.x, ct.y, n_tgt, rev_scale, partial);
    else
        hipLaunchKernelGGL(k_ff_mfma4, grid, block, 0, st, (const SrcRec32 *)packed, meta, n_src, p.chunk,
                           ct.x, ct.y, n_tgt, rev_scale, partial);
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
