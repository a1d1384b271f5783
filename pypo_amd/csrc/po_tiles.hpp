// po_tiles.hpp -- the per-tile inner loops of the mixed-precision PO sums,
// shared by the single-propagation kernels (po_ff.hip, po_near.hip) and the
// persistent batched-sweep kernels (sweep.hip).  Each loop consumes one
// 256-record LDS tile for the TPL targets a lane owns and accumulates in FP32.
#pragma once

#include "po_common.hpp"

namespace pypo {

// Far field: six (or three, J only) complex multiply-adds of w*A*currents
// with exp(-j t k r'.r) per pair; MASK applies the shadow test n.r >= 0
// (Propagation.h:1719).  rx/ry/rz = r * (-t k / 2 pi) in FP64, fx/fy/fz = r.
template <int TPL, bool HAS_M, bool MASK>
__device__ __forceinline__ void ff_tile_mixed(const SrcRec32 *tile, const double (&rx)[TPL],
                                              const double (&ry)[TPL], const double (&rz)[TPL],
                                              const float (&fx)[TPL], const float (&fy)[TPL],
                                              const float (&fz)[TPL], float (&acc)[TPL][12]) {
#pragma unroll 2
    for (int s = 0; s < TILE; s++) {
        const SrcRec32 &r = tile[s];
        const double x = r.x, y = r.y, z = r.z;
        float cj[6], cm[6];
#pragma unroll
        for (int q = 0; q < 6; q++) cj[q] = r.cj[q];
        if (HAS_M) {
#pragma unroll
            for (int q = 0; q < 6; q++) cm[q] = r.cm[q];
        }
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            double rev = __builtin_fma(z, rz[j], __builtin_fma(y, ry[j], x * rx[j]));
            float f = (float)__builtin_amdgcn_fract(rev);
            float sn, cs;
            sincos_rev_f32(f, sn, cs);
            if (MASK) {
                float m = __builtin_fmaf(r.nz, fz[j], __builtin_fmaf(r.ny, fy[j], r.nx * fx[j]));
                cs = (m < 0.f) ? 0.f : cs;
                sn = (m < 0.f) ? 0.f : sn;
            }
#pragma unroll
            for (int q = 0; q < 3; q++) {
                float &ar = acc[j][2 * q], &ai = acc[j][2 * q + 1];
                ar = __builtin_fmaf(cj[2 * q], cs, ar);
                ar = __builtin_fmaf(-cj[2 * q + 1], sn, ar);
                ai = __builtin_fmaf(cj[2 * q], sn, ai);
                ai = __builtin_fmaf(cj[2 * q + 1], cs, ai);
            }
            if (HAS_M) {
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    float &ar = acc[j][6 + 2 * q], &ai = acc[j][7 + 2 * q];
                    ar = __builtin_fmaf(cm[2 * q], cs, ar);
                    ar = __builtin_fmaf(-cm[2 * q + 1], sn, ar);
                    ai = __builtin_fmaf(cm[2 * q], sn, ai);
                    ai = __builtin_fmaf(cm[2 * q + 1], cs, ai);
                }
            }
        }
    }
}

// Near field (fieldAtPoint, Propagation.h:977-1144): see po_near.hip.
// WANT_E = false accumulates H only (the JM mode needs nothing else); the
// E terms and, for tiles without magnetic current, the s1/s2 factors are
// then dead code.
template <int TPL, bool HAS_M, bool WANT_E>
__device__ __forceinline__ void near_tile_mixed(const SrcRec32 *tile, const double (&px)[TPL],
                                                const double (&py)[TPL], const double (&pz)[TPL],
                                                double rev_scale, float kinv, float tf,
                                                float (&acc)[TPL][12], int n = TILE) {
#pragma unroll 1
    for (int s = 0; s < n; s++) {   // n: the tile's records before the padding
        const SrcRec32 &r = tile[s];
        float cj[6], cm[6];
#pragma unroll
        for (int q = 0; q < 6; q++) cj[q] = r.cj[q];
        if (HAS_M) {
#pragma unroll
            for (int q = 0; q < 6; q++) cm[q] = r.cm[q];
        }
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
            const float h[3] = {(float)dx * Rinv, (float)dy * Rinv, (float)dz * Rinv};
            float m = __builtin_fmaf(r.nz, h[2], __builtin_fmaf(r.ny, h[1], r.nx * h[0]));
            cs = (m < 0.f) ? 0.f : cs;
            sn = (m < 0.f) ? 0.f : sn;
            float q1 = Rinv * kinv;
            float q2 = q1 * q1, q3 = q2 * q1;
            // s_i * G as (re, im)
            float a1 = tf * q2, b1 = q3 - q1;
            float a2 = -3.f * a1, b2 = __builtin_fmaf(-3.f, q3, q1);
            float a3 = -q2, b3 = tf * q1;
            float g1r = __builtin_fmaf(a1, cs, -b1 * sn), g1i = __builtin_fmaf(a1, sn, b1 * cs);
            float g2r = __builtin_fmaf(a2, cs, -b2 * sn), g2i = __builtin_fmaf(a2, sn, b2 * cs);
            float g3r = __builtin_fmaf(a3, cs, -b3 * sn), g3i = __builtin_fmaf(a3, sn, b3 * cs);
            // (J.R) s2 G
            float jr = __builtin_fmaf(cj[4], h[2], __builtin_fmaf(cj[2], h[1], cj[0] * h[0]));
            float ji = __builtin_fmaf(cj[5], h[2], __builtin_fmaf(cj[3], h[1], cj[1] * h[0]));
            float Ar = __builtin_fmaf(jr, g2r, -ji * g2i), Ai = __builtin_fmaf(jr, g2i, ji * g2r);
            float Br = 0.f, Bi = 0.f;
            if (HAS_M) {
                float mr = __builtin_fmaf(cm[4], h[2], __builtin_fmaf(cm[2], h[1], cm[0] * h[0]));
                float mi = __builtin_fmaf(cm[5], h[2], __builtin_fmaf(cm[3], h[1], cm[1] * h[0]));
                Br = __builtin_fmaf(mr, g2r, -mi * g2i);
                Bi = __builtin_fmaf(mr, g2i, mi * g2r);
            }
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int b = (a + 1) % 3, d = (a + 2) % 3;
                // (V x R)_a = V_b R_d - V_d R_b
                float jxr_r = __builtin_fmaf(cj[2 * b], h[d], -cj[2 * d] * h[b]);
                float jxr_i = __builtin_fmaf(cj[2 * b + 1], h[d], -cj[2 * d + 1] * h[b]);
                float &er = acc[j][2 * a], &ei = acc[j][2 * a + 1];
                float &hr = acc[j][6 + 2 * a], &hi = acc[j][7 + 2 * a];
                if (WANT_E) {
                    er = __builtin_fmaf(cj[2 * a], g1r, er);
                    er = __builtin_fmaf(-cj[2 * a + 1], g1i, er);
                    er = __builtin_fmaf(h[a], Ar, er);
                    ei = __builtin_fmaf(cj[2 * a], g1i, ei);
                    ei = __builtin_fmaf(cj[2 * a + 1], g1r, ei);
                    ei = __builtin_fmaf(h[a], Ai, ei);
                }
                hr = __builtin_fmaf(-jxr_r, g3r, hr);
                hr = __builtin_fmaf(jxr_i, g3i, hr);
                hi = __builtin_fmaf(-jxr_r, g3i, hi);
                hi = __builtin_fmaf(-jxr_i, g3r, hi);
                if (HAS_M) {
                    float mxr_r = __builtin_fmaf(cm[2 * b], h[d], -cm[2 * d] * h[b]);
                    float mxr_i = __builtin_fmaf(cm[2 * b + 1], h[d], -cm[2 * d + 1] * h[b]);
                    if (WANT_E) {
                        er = __builtin_fmaf(mxr_r, g3r, er);
                        er = __builtin_fmaf(-mxr_i, g3i, er);
                        ei = __builtin_fmaf(mxr_r, g3i, ei);
                        ei = __builtin_fmaf(mxr_i, g3r, ei);
                    }
                    hr = __builtin_fmaf(cm[2 * a], g1r, hr);
                    hr = __builtin_fmaf(-cm[2 * a + 1], g1i, hr);
                    hr = __builtin_fmaf(h[a], Br, hr);
                    hi = __builtin_fmaf(cm[2 * a], g1i, hi);
                    hi = __builtin_fmaf(cm[2 * a + 1], g1r, hi);
                    hi = __builtin_fmaf(h[a], Bi, hi);
                }
            }
        }
    }
}

}  // namespace pypo
