// po_near.hip -- near-field PO sum (fieldAtPoint, Propagation.h:977-1144)
// and the per-target epilogues of the JM, EH, JMEH and EHP modes.
//
// Per pair:  R = |p - r'|, R_hat = (p - r')/R, q = 1/(kR),
//   s1 = t q^2 + j(q^3 - q), s2 = -3t q^2 + j(q - 3q^3), s3 = -q^2 + j t q
//   (Propagation.h:1084-1090 with t^2 = 1), G = exp(j t k R),
//   E += G [J s1 + (J.R)R s2 + (M x R) s3],  H += G [M s1 + (M.R)R s2 - (J x R) s3]
// skipped when n.R_hat < 0 (Propagation.h:1079).  The prefactor k^2/(4 pi)
// is applied by the finish kernel; w*A is folded into the records.
// Mixed precision: FP64 difference vector, squared distance, distance (FP32
// v_rsq refined by one FP64 Newton step: relative error ~1e-14) and phase;
// FP32 for the unit vector, the s-coefficients and the accumulation (per
// 256-source tile, folded into FP64).  Tiles without magnetic current skip
// the M terms.

#include "po_tiles.hpp"

namespace pypo {
namespace {

// WANT_E = false (JM): only the H rows 6..11 of the partials are written.
template <int TPL, bool WANT_E>
__global__ __launch_bounds__(256) void k_near_mixed(const SrcRec32 *__restrict__ src,
                                                    const TileMeta *__restrict__ meta, int n_src,
                                                    int chunk, const double *__restrict__ tx,
                                                    const double *__restrict__ ty,
                                                    const double *__restrict__ tz, int n_tgt,
                                                    double k, double tdir,
                                                    double *__restrict__ partial) {
    __shared__ SrcRec32 tile[TILE];
    const int tid = threadIdx.x;
    int tb, sp;
    xcd_item(tb, sp);
    const double rev_scale = tdir * k / kTwoPi;   // exp(j t k R)
    const float kinv = (float)(1.0 / k);
    const float tf = (float)tdir;
    double px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = tb * (TILE * TPL) + j * TILE + tid;
        // padding lanes sit far away so that every distance stays finite
        px[j] = (t < n_tgt) ? tx[t] : 1.0e6;
        py[j] = (t < n_tgt) ? ty[t] : 1.0e6;
        pz[j] = (t < n_tgt) ? tz[t] : 1.0e6;
    }
    float acc[TPL][12];
    double tot[TPL][12];
#pragma unroll
    for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int q = 0; q < 12; q++) { acc[j][q] = 0.f; tot[j][q] = 0.0; }

    const int s_begin = sp * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<6> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        __syncthreads();
        ld.store(reinterpret_cast<uint4 *>(tile), tid);
        __syncthreads();
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
        const int n = min(TILE, s_end - base);   // padding records add exactly zero: skip them
        if (meta[base / TILE].has_m)
            near_tile_mixed<TPL, true, WANT_E>(tile, px, py, pz, rev_scale, kinv, tf, acc, n);
        else
            near_tile_mixed<TPL, false, WANT_E>(tile, px, py, pz, rev_scale, kinv, tf, acc, n);
#pragma unroll
        for (int j = 0; j < TPL; j++)
#pragma unroll
            for (int q = WANT_E ? 0 : 6; q < 12; q++) { tot[j][q] += (double)acc[j][q]; acc[j][q] = 0.f; }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = tb * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
#pragma unroll
            for (int q = WANT_E ? 0 : 6; q < 12; q++)
                partial[((size_t)sp * 12 + q) * n_tgt + t] = tot[j][q];
        }
    }
}

template <int TPL, bool HAS_M>
__device__ __forceinline__ void near_tile_fp64(const SrcRec64 *tile, const double (&px)[TPL],
                                               const double (&py)[TPL], const double (&pz)[TPL],
                                               double rev_scale, double k, double tdir,
                                               double (&acc)[TPL][12], int n) {
#pragma unroll 1
    for (int s = 0; s < n; s++) {
        const SrcRec64 &r = tile[s];
        const double *cj = r.c, *cm = r.c + 6;
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            double dx = px[j] - r.x, dy = py[j] - r.y, dz = pz[j] - r.z;
            double R = sqrt(__builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx)));
            double Rinv = 1.0 / R;
            double f = __builtin_amdgcn_fract(R * rev_scale);
            double sn, cs;
            sincos_rev_f64(f, sn, cs);
            const double h[3] = {dx * Rinv, dy * Rinv, dz * Rinv};
            double m = __builtin_fma(r.nz, h[2], __builtin_fma(r.ny, h[1], r.nx * h[0]));
            cs = (m < 0.0) ? 0.0 : cs;
            sn = (m < 0.0) ? 0.0 : sn;
            double q1 = Rinv / k;
            double q2 = q1 * q1, q3 = q2 * q1;
            double a1 = tdir * q2, b1 = q3 - q1;
            double a2 = -3.0 * a1, b2 = __builtin_fma(-3.0, q3, q1);
            double a3 = -q2, b3 = tdir * q1;
            double g1r = __builtin_fma(a1, cs, -b1 * sn), g1i = __builtin_fma(a1, sn, b1 * cs);
            double g2r = __builtin_fma(a2, cs, -b2 * sn), g2i = __builtin_fma(a2, sn, b2 * cs);
            double g3r = __builtin_fma(a3, cs, -b3 * sn), g3i = __builtin_fma(a3, sn, b3 * cs);
            double jr = __builtin_fma(cj[4], h[2], __builtin_fma(cj[2], h[1], cj[0] * h[0]));
            double ji = __builtin_fma(cj[5], h[2], __builtin_fma(cj[3], h[1], cj[1] * h[0]));
            double Ar = __builtin_fma(jr, g2r, -ji * g2i), Ai = __builtin_fma(jr, g2i, ji * g2r);
            double Br = 0.0, Bi = 0.0;
            if (HAS_M) {
                double mr = __builtin_fma(cm[4], h[2], __builtin_fma(cm[2], h[1], cm[0] * h[0]));
                double mi = __builtin_fma(cm[5], h[2], __builtin_fma(cm[3], h[1], cm[1] * h[0]));
                Br = __builtin_fma(mr, g2r, -mi * g2i);
                Bi = __builtin_fma(mr, g2i, mi * g2r);
            }
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int b = (a + 1) % 3, d = (a + 2) % 3;
                double jxr_r = __builtin_fma(cj[2 * b], h[d], -cj[2 * d] * h[b]);
                double jxr_i = __builtin_fma(cj[2 * b + 1], h[d], -cj[2 * d + 1] * h[b]);
                double &er = acc[j][2 * a], &ei = acc[j][2 * a + 1];
                double &hr = acc[j][6 + 2 * a], &hi = acc[j][7 + 2 * a];
                er = __builtin_fma(cj[2 * a], g1r, er);
                er = __builtin_fma(-cj[2 * a + 1], g1i, er);
                er = __builtin_fma(h[a], Ar, er);
                ei = __builtin_fma(cj[2 * a], g1i, ei);
                ei = __builtin_fma(cj[2 * a + 1], g1r, ei);
                ei = __builtin_fma(h[a], Ai, ei);
                hr = __builtin_fma(-jxr_r, g3r, hr);
                hr = __builtin_fma(jxr_i, g3i, hr);
                hi = __builtin_fma(-jxr_r, g3i, hi);
                hi = __builtin_fma(-jxr_i, g3r, hi);
                if (HAS_M) {
                    double mxr_r = __builtin_fma(cm[2 * b], h[d], -cm[2 * d] * h[b]);
                    double mxr_i = __builtin_fma(cm[2 * b + 1], h[d], -cm[2 * d + 1] * h[b]);
                    er = __builtin_fma(mxr_r, g3r, er);
                    er = __builtin_fma(-mxr_i, g3i, er);
                    ei = __builtin_fma(mxr_r, g3i, ei);
                    ei = __builtin_fma(mxr_i, g3r, ei);
                    hr = __builtin_fma(cm[2 * a], g1r, hr);
                    hr = __builtin_fma(-cm[2 * a + 1], g1i, hr);
                    hr = __builtin_fma(h[a], Br, hr);
                    hi = __builtin_fma(cm[2 * a], g1i, hi);
                    hi = __builtin_fma(cm[2 * a + 1], g1r, hi);
                    hi = __builtin_fma(h[a], Bi, hi);
                }
            }
        }
    }
}

template <int TPL>
__global__ __launch_bounds__(256) void k_near_fp64(const SrcRec64 *__restrict__ src,
                                                   const TileMeta *__restrict__ meta, int n_src,
                                                   int chunk, const double *__restrict__ tx,
                                                   const double *__restrict__ ty,
                                                   const double *__restrict__ tz, int n_tgt,
                                                   double k, double tdir,
                                                   double *__restrict__ partial) {
    __shared__ SrcRec64 tile[TILE];
    const int tid = threadIdx.x;
    int tb, sp;
    xcd_item(tb, sp);
    const double rev_scale = tdir * k / kTwoPi;
    double px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = tb * (TILE * TPL) + j * TILE + tid;
        px[j] = (t < n_tgt) ? tx[t] : 1.0e6;
        py[j] = (t < n_tgt) ? ty[t] : 1.0e6;
        pz[j] = (t < n_tgt) ? tz[t] : 1.0e6;
    }
    double acc[TPL][12];
#pragma unroll
    for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int q = 0; q < 12; q++) acc[j][q] = 0.0;

    const int s_begin = sp * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<9> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        __syncthreads();
        ld.store(reinterpret_cast<uint4 *>(tile), tid);
        __syncthreads();
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
        const int n = min(TILE, s_end - base);
        if (meta[base / TILE].has_m)
            near_tile_fp64<TPL, true>(tile, px, py, pz, rev_scale, k, tdir, acc, n);
        else
            near_tile_fp64<TPL, false>(tile, px, py, pz, rev_scale, k, tdir, acc, n);
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = tb * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
#pragma unroll
            for (int q = 0; q < 12; q++)
                partial[((size_t)sp * 12 + q) * n_tgt + t] = acc[j][q];
        }
    }
}

__device__ __forceinline__ void unit3(const double *v, double *o) {
    double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (nrm <= 4.9406564584124654e-324) nrm = 4.9406564584124654e-324;   // Utils.h:293-296
    o[0] = v[0] / nrm; o[1] = v[1] / nrm; o[2] = v[2] / nrm;
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Per-target epilogues (FP64): propagateBeam_JM_PEC 269-323 (J = 2 n x H,
// M = 0), _EH 574-617, _JMEH_PEC 482-555, _EHP 805-914.
template <int MODE>
__global__ __launch_bounds__(256) void k_near_finish(const double *__restrict__ partial,
                                                     int splits, int n_tgt, GridPtrs ct,
                                                     double pref, double zeta_inv, OutC4 out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tgt) return;
    double S[12];
    // the JM sums write the H rows only
    constexpr int Q0 = (MODE == PYPO_MODE_JM) ? 6 : 0;
    // split sums; each split's loads issue together
#pragma unroll
    for (int q = 0; q < 12; q++) S[q] = q >= Q0 ? partial[(size_t)q * n_tgt + t] : 0.0;
    for (int s = 1; s < splits; s++)
#pragma unroll
        for (int q = Q0; q < 12; q++) S[q] += partial[((size_t)s * 12 + q) * n_tgt + t];
#pragma unroll
    for (int q = 0; q < 12; q++) S[q] *= pref;
    double Er[3], Ei[3], Hr[3], Hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        Er[a] = S[2 * a]; Ei[a] = S[2 * a + 1];
        Hr[a] = S[6 + 2 * a]; Hi[a] = S[7 + 2 * a];
    }
    if (MODE == PYPO_MODE_EH) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            out.p[a][t] = Er[a]; out.p[3 + a][t] = Ei[a];
            out.p[6 + a][t] = Hr[a]; out.p[9 + a][t] = Hi[a];
        }
        return;
    }
    const double n[3] = {ct.nx[t], ct.ny[t], ct.nz[t]};
    if (MODE == PYPO_MODE_JM || MODE == PYPO_MODE_JMEH) {
        double jr[3], ji[3];
        cross3(n, Hr, jr);
        cross3(n, Hi, ji);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            out.p[a][t] = 2 * jr[a]; out.p[3 + a][t] = 2 * ji[a];
            out.p[6 + a][t] = 0.0; out.p[9 + a][t] = 0.0;
        }
        if (MODE == PYPO_MODE_JMEH) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                out.p[12 + a][t] = Er[a]; out.p[15 + a][t] = Ei[a];
                out.p[18 + a][t] = Hr[a]; out.p[21 + a][t] = Hi[a];
            }
        }
        return;
    }
    // EHP: Re(E x H*), incident / reflected polarisation bases, reflection.
    double sre[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        // (E x conj(H))_a real part = Er_b Hr_c + Ei_b Hi_c - (Er_c Hr_b + Ei_c Hi_b)
        sre[a] = (Er[b] * Hr[c] + Ei[b] * Hi[c]) - (Er[c] * Hr[b] + Ei[c] * Hi[b]);
    }
    double Si[3], tmp[3], pip[3], pia[3], Sr[3], prp[3], pra[3];
    unit3(sre, Si);
    cross3(Si, n, tmp); unit3(tmp, pip);
    cross3(pip, Si, pia);
    double f = 2.0 * (Si[0] * n[0] + Si[1] * n[1] + Si[2] * n[2]);
#pragma unroll
    for (int a = 0; a < 3; a++) Sr[a] = Si[a] - n[a] * f;
    cross3(Sr, n, tmp); unit3(tmp, prp);
    cross3(Sr, prp, pra);
    double ep_r = 0, ep_i = 0, ea_r = 0, ea_i = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        ep_r += Er[a] * prp[a]; ep_i += Ei[a] * prp[a];
        ea_r += Er[a] * pra[a]; ea_i += Ei[a] * pra[a];
    }
    double err[3], eri[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        err[a] = -ep_r * pip[a] - ea_r * pia[a];
        eri[a] = -ep_i * pip[a] - ea_i * pia[a];
    }
    double hr[3], hi[3];
    cross3(Sr, err, hr);
    cross3(Sr, eri, hi);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        out.p[a][t] = err[a]; out.p[3 + a][t] = eri[a];
        out.p[6 + a][t] = hr[a] * zeta_inv; out.p[9 + a][t] = hi[a] * zeta_inv;
        out.p[12 + a][t] = Sr[a];
    }
}

}  // namespace

int launch_near_sum(int mode, int precision, const SumPlan &p, const void *packed, int n_src,
                    const GridPtrs &ct, int n_tgt, double k, double tdir, double *partial,
                    hipStream_t st) {
    dim3 grid(p.tblocks, p.splits), block(TILE);
    const int rb = precision == PYPO_PREC_FP64 ? (int)sizeof(SrcRec64) : (int)sizeof(SrcRec32);
    const TileMeta *meta = (const TileMeta *)((const char *)packed + meta_offset(rb, n_src));
    if (precision == PYPO_PREC_FP64)
        hipLaunchKernelGGL(k_near_fp64<1>, grid, block, 0, st, (const SrcRec64 *)packed, meta, n_src,
                           p.chunk, ct.x, ct.y, ct.z, n_tgt, k, tdir, partial);
    else if (mode == PYPO_MODE_JM)
        hipLaunchKernelGGL((k_near_mixed<1, false>), grid, block, 0, st, (const SrcRec32 *)packed, meta,
                           n_src, p.chunk, ct.x, ct.y, ct.z, n_tgt, k, tdir, partial);
    else
        hipLaunchKernelGGL((k_near_mixed<1, true>), grid, block, 0, st, (const SrcRec32 *)packed, meta,
                           n_src, p.chunk, ct.x, ct.y, ct.z, n_tgt, k, tdir, partial);
    PYPO_CHECK_LAUNCH();
    return 0;
}

int launch_near_finish(int mode, const SumPlan &p, const GridPtrs &ct, int n_tgt, double pref,
                       double zeta_inv, const OutC4 &o, const double *partial, hipStream_t st) {
    dim3 grid((n_tgt + 255) / 256), block(256);
    switch (mode) {
        case PYPO_MODE_JM:
            hipLaunchKernelGGL(k_near_finish<PYPO_MODE_JM>, grid, block, 0, st, partial, p.splits, n_tgt,
                               ct, pref, zeta_inv, o);
            break;
        case PYPO_MODE_EH:
            hipLaunchKernelGGL(k_near_finish<PYPO_MODE_EH>, grid, block, 0, st, partial, p.splits, n_tgt,
                               ct, pref, zeta_inv, o);
            break;
        case PYPO_MODE_JMEH:
            hipLaunchKernelGGL(k_near_finish<PYPO_MODE_JMEH>, grid, block, 0, st, partial, p.splits,
                               n_tgt, ct, pref, zeta_inv, o);
            break;
        default:
            hipLaunchKernelGGL(k_near_finish<PYPO_MODE_EHP>, grid, block, 0, st, partial, p.splits,
                               n_tgt, ct, pref, zeta_inv, o);
            break;
    }
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
