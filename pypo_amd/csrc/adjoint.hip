// adjoint.hip -- gradient of a far-field power loss with respect to per-point
// reflector displacement along the surface normal.
//
// Model (the reference's surface-error model, InterfaceReflector.cpp:852-866:
// points move along their normals, normals and area elements stay):
//   p_i = p0_i + d_i n_i
//   J_i = 2 n_i x H_inc(p_i)                 (JM, Propagation.h:269-323)
//   E_t = -j pref P_t sum_i c_i J_i exp(j phi_ti),  phi_ti = -t k p_i.r_t,
//   c_i = w_i A_i                              (FF, Propagation.h:1656-1781)
//   P_t = |E_t|^2 ;  g_t = dL/dP_t given by the caller.
// Backward:
//   A_t = 2 g_t (-j pref) conj(E_t)                                  (targets)
//   B_i = sum_t A_t exp(j phi_ti),  G_i = sum_t (n_i.r_t) A_t exp(j phi_ti) (reverse sums)
//   J_i, dJ_i = 2 n_i x (n_i . grad) H_inc(p_i)   (one pass over the feed: the
//                                                   currents are recomputed, not stored)
//   dL/dd_i = c_i [ Re(B_i . dJ_i) + t k Im(J_i . G_i) ]
// The shadow tests of the forward sums are applied here too; their own
// derivatives are zero almost everywhere.

#include "po_ff_mfma.hpp"

#include <algorithm>

namespace pypo {
namespace {

// Target record of the reverse far-field sum: r scaled to revolutions
// (FP64), r (FP32) and A_t (3 complex, FP32).  64 bytes.
struct __attribute__((aligned(16))) TgtRec32 {
    double ux, uy, uz;
    float rx, ry, rz;
    float a[6];
    float pad;
};
static_assert(sizeof(TgtRec32) == 64, "TgtRec32 must be 64 bytes");

struct __attribute__((aligned(16))) TgtRec64 {
    double ux, uy, uz;
    double rx, ry, rz;
    double a[6];
};
static_assert(sizeof(TgtRec64) == 96, "TgtRec64 must be 96 bytes");

// One thread per far-field target: r_t, A_t from E_t and g_t.
template <class Rec>
__global__ __launch_bounds__(256) void k_adj_targets(const double *__restrict__ th,
                                                     const double *__restrict__ ph, Bundle12 e,
                                                     const double *__restrict__ g, int n_tgt,
                                                     double rev_scale, double pref, Rec *__restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tgt) return;
    double st, ct, sp, cp;
    sincos(th[t], &st, &ct);
    sincos(ph[t], &sp, &cp);
    const double r[3] = {cp * st, sp * st, ct};
    Rec o{};
    o.ux = r[0] * rev_scale; o.uy = r[1] * rev_scale; o.uz = r[2] * rev_scale;
    o.rx = r[0]; o.ry = r[1]; o.rz = r[2];
    const double s = 2.0 * g[t] * pref;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        // (-j)(Er - j Ei) = -Ei - j Er
        o.a[2 * a] = s * (-e.p[3 + a][t]);
        o.a[2 * a + 1] = s * (-e.p[a][t]);
    }
    out[t] = o;
}

// Reverse far-field sum, mixed precision: lanes own SPL sources (dish
// points) each, targets stream through LDS; every 64-byte target record read
// from LDS serves SPL pairs.  Phase in FP64 revolutions reduced with
// v_fract_f64, FP32 sin/cos and FMAs, per-tile FP32 sums folded into FP64.
// Partial layout [split][12][n_src]: q 0..5 = B (re/im xyz interleaved),
// 6..11 = G.
template <int SPL>
__global__ __launch_bounds__(256) void k_adj_reverse32(const TgtRec32 *__restrict__ tg, int n_tgt,
                                                       int chunk, GridPtrs src, int n_src,
                                                       double *__restrict__ partial) {
    __shared__ TgtRec32 tile[TILE];
    const int tid = threadIdx.x;
    double px[SPL], py[SPL], pz[SPL];
    float nx[SPL], ny[SPL], nz[SPL];
#pragma unroll
    for (int j = 0; j < SPL; j++) {
        const int i = blockIdx.x * (TILE * SPL) + j * TILE + tid;
        const bool valid = i < n_src;
        px[j] = valid ? src.x[i] : 0.0; py[j] = valid ? src.y[i] : 0.0; pz[j] = valid ? src.z[i] : 0.0;
        nx[j] = valid ? (float)src.nx[i] : 0.f; ny[j] = valid ? (float)src.ny[i] : 0.f;
        nz[j] = valid ? (float)src.nz[i] : 0.f;
    }
    float acc[SPL][12];
    double tot[SPL][12];
#pragma unroll
    for (int j = 0; j < SPL; j++)
#pragma unroll
        for (int q = 0; q < 12; q++) { acc[j][q] = 0.f; tot[j][q] = 0.0; }
    const int t_begin = blockIdx.y * chunk;
    const int t_end = min(n_tgt, t_begin + chunk);
    for (int base = t_begin; base < t_end; base += TILE) {
        __syncthreads();
        const int t = base + tid;
        if (t < t_end) {
            tile[tid] = tg[t];
        } else {
            TgtRec32 z{};
            tile[tid] = z;   // zero A: adds nothing
        }
        __syncthreads();
#pragma unroll 2
        for (int s = 0; s < TILE; s++) {
            const TgtRec32 &r = tile[s];
            const float a[6] = {r.a[0], r.a[1], r.a[2], r.a[3], r.a[4], r.a[5]};
#pragma unroll
            for (int j = 0; j < SPL; j++) {
                double rev = __builtin_fma(pz[j], r.uz, __builtin_fma(py[j], r.uy, px[j] * r.ux));
                float f = (float)__builtin_amdgcn_fract(rev);
                float sn, cs;
                sincos_rev_f32(f, sn, cs);
                float w = __builtin_fmaf(nz[j], r.rz, __builtin_fmaf(ny[j], r.ry, nx[j] * r.rx));
                cs = (w < 0.f) ? 0.f : cs;
                sn = (w < 0.f) ? 0.f : sn;
                const float wc = w * cs, ws = w * sn;
                float *c = acc[j];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const float ar = a[2 * q], ai = a[2 * q + 1];
                    c[2 * q] = __builtin_fmaf(ar, cs, c[2 * q]);
                    c[2 * q] = __builtin_fmaf(-ai, sn, c[2 * q]);
                    c[2 * q + 1] = __builtin_fmaf(ar, sn, c[2 * q + 1]);
                    c[2 * q + 1] = __builtin_fmaf(ai, cs, c[2 * q + 1]);
                    c[6 + 2 * q] = __builtin_fmaf(ar, wc, c[6 + 2 * q]);
                    c[6 + 2 * q] = __builtin_fmaf(-ai, ws, c[6 + 2 * q]);
                    c[7 + 2 * q] = __builtin_fmaf(ar, ws, c[7 + 2 * q]);
                    c[7 + 2 * q] = __builtin_fmaf(ai, wc, c[7 + 2 * q]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < SPL; j++)
#pragma unroll
            for (int q = 0; q < 12; q++) { tot[j][q] += (double)acc[j][q]; acc[j][q] = 0.f; }
    }
#pragma unroll
    for (int j = 0; j < SPL; j++) {
        const int i = blockIdx.x * (TILE * SPL) + j * TILE + tid;
        if (i < n_src) {
#pragma unroll
            for (int q = 0; q < 12; q++) partial[((size_t)blockIdx.y * 12 + q) * n_src + i] = tot[j][q];
        }
    }
}

// ---- reverse sum on the matrix cores ----------------------------------------------
// B_i and G_i = sum_a n_i,a G_i,a with G_i,a = sum_t r_t,a A_t W(i, t) form a
// (dish points x targets) by (targets x 24) complex product: rows are dish
// points (16 per M-block), K runs over targets (cos and sin rows), and the
// 24 real columns (A, r_x A | r_y A, r_z A) fill two 16-column blocks.  The
// phase, FP32 sin/cos, the FP16 hi/lo operand split and the per-tile operand
// scale are those of the far-field kernel (po_ff_mfma.hpp); the shadow test
// n_i.r_t >= 0 zeroes W where a cone test cannot exclude it.  G_i is formed
// from the three G_i,a in FP64 by k_adj_combine24.

// Per 256-target tile: the cone of the directions r_t and the largest |A_t|
// component (the B operand's scale).
__global__ __launch_bounds__(256) void k_adj_tgt_meta(const TgtRec32 *__restrict__ tg, int n_tgt,
                                                      TileMeta *__restrict__ meta) {
    __shared__ float scratch[4];
    const int t = blockIdx.x * TILE + threadIdx.x;
    const bool valid = t < n_tgt;
    float rx = 0.f, ry = 0.f, rz = 1.f, am = 0.f;
    if (valid) {
        const TgtRec32 &r = tg[t];
        rx = r.rx; ry = r.ry; rz = r.rz;
#pragma unroll
        for (int q = 0; q < 6; q++) am = fmaxf(am, fabsf(r.a[q]));
    }
    const Cone c = block_cone(rx, ry, rz, valid, scratch);
    const float amax = -block_min(-am, scratch);
    if (threadIdx.x == 0) {
        TileMeta m{};
        m.ax = c.ax; m.ay = c.ay; m.az = c.az; m.cos_half = c.cos_half;
        m.n_valid = min(TILE, n_tgt - (int)blockIdx.x * TILE);
        m.cmax = amax;
        meta[blockIdx.x] = m;
    }
}

struct RevTile {
    double u[3][TILE];       // r_t * rev_scale
    float r[3][TILE];        // r_t (shadow test)
    float du[3][TILE];       // follower offsets u_t - u_leader (FP32 revolutions per mm)
    int fast[4];             // per wave: 1 when all its followers may take the follower form
};

// This lane's target record (tile slot sl) -> its entries of both B blocks.
__device__ __forceinline__ void build_rev_b(const TgtRec32 &r, int sl, float inv, mfma::half8 (*b0h)[64],
                                            mfma::half8 (*b0l)[64], mfma::half8 (*b1h)[64],
                                            mfma::half8 (*b1l)[64]) {
    const int ks = sl >> 4, g = (sl & 15) >> 2, jj = sl & 3;
    const float w[4] = {1.f, r.rx, r.ry, r.rz};
#pragma unroll
    for (int v = 0; v < 4; v++) {   // A, r_x A, r_y A, r_z A
        mfma::half8 (*bh)[64] = v < 2 ? b0h : b1h;
        mfma::half8 (*bl)[64] = v < 2 ? b0l : b1l;
        const int c0 = (v & 1) * 6;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float re = r.a[2 * q] * w[v] * inv, im = r.a[2 * q + 1] * w[v] * inv;
            unsigned h, l;
            const int s0 = mfma::b_slot(ks, 16 * g + c0 + 2 * q), s1 = mfma::b_slot(ks, 16 * g + c0 + 2 * q + 1);
            mfma::split_pair_fast(re, -im, h, l);
            reinterpret_cast<unsigned *>(&bh[ks][s0])[jj] = h;
            reinterpret_cast<unsigned *>(&bl[ks][s0])[jj] = l;
            mfma::split_pair_fast(im, re, h, l);
            reinterpret_cast<unsigned *>(&bh[ks][s1])[jj] = h;
            reinterpret_cast<unsigned *>(&bl[ks][s1])[jj] = l;
        }
    }
}

// One 256-target tile, 16 K-steps of 16 targets.  Phase as in the forward
// matrix kernel, with the roles of the operands swapped: the targets a lane
// takes per K-step are neighbours on the far-field grid, the second of them
// (the leader) gets rev = p . u in FP64, the others (FAST tiles only)
//   f_t = f_leader + p . (u_t - u_leader)       (three FP32 FMAs)
// with |f_t| <= kLocalRev over the workgroup's dish points (checked when the
// tile is staged).
template <int MB, bool MASK, bool FAST>
__device__ __forceinline__ void rev_tile_mfma(const RevTile &tt, const mfma::half8 (*b0h)[64],
                                              const mfma::half8 (*b0l)[64], const mfma::half8 (*b1h)[64],
                                              const mfma::half8 (*b1l)[64], int lane, const double (&px)[MB],
                                              const double (&py)[MB], const double (&pz)[MB],
                                              const float (&pxf)[MB], const float (&pyf)[MB],
                                              const float (&pzf)[MB], const float (*nv)[64 * MB], int lt0,
                                              mfma::floatx4 (&acc0)[MB], mfma::floatx4 (&acc1)[MB]) {
    using mfma::half8;
    const int kb = lane >> 4;
    // j-th computation of an M-block: the leader (target 1) first
    auto tgt_of = [](int j) { return FAST ? (j == 0 ? 1 : (j == 1 ? 0 : j)) : j; };
#pragma unroll 1
    for (int ks = 0; ks < mfma::KSTEPS; ks++) {
        const int slot = mfma::b_slot(ks, lane);
        const half8 h0 = b0h[ks][slot], l0 = b0l[ks][slot], h1 = b1h[ks][slot], l1 = b1l[ks][slot];
        const int t0 = ks * 16 + kb * 4;
        double ux[4], uy[4], uz[4];
        float dx[4], dy[4], dz[4], rx[4], ry[4], rz[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (!FAST || j == 1) { ux[j] = tt.u[0][t0 + j]; uy[j] = tt.u[1][t0 + j]; uz[j] = tt.u[2][t0 + j]; }
            if (FAST && j != 1) { dx[j] = tt.du[0][t0 + j]; dy[j] = tt.du[1][t0 + j]; dz[j] = tt.du[2][t0 + j]; }
            if (MASK) { rx[j] = tt.r[0][t0 + j]; ry[j] = tt.r[1][t0 + j]; rz[j] = tt.r[2][t0 + j]; }
        }
#pragma unroll
        for (int m = 0; m < MB; m++) {
            unsigned ah[4], al[4];
            float f_lead = 0.f;
#pragma unroll
            for (int jj = 0; jj < 4; jj++) {
                const int j = tgt_of(jj);
                float f;
                if (!FAST || j == 1) {
                    const double rev = __builtin_fma(pz[m], uz[j], __builtin_fma(py[m], uy[j], px[m] * ux[j]));
                    f = (float)__builtin_amdgcn_fract(rev);
                    if (FAST) f_lead = f;
                } else {
                    // |f| <= kLocalRev (the bound counts f_lead): in the hardware's own reduction range
                    f = __builtin_fmaf(pzf[m], dz[j], __builtin_fmaf(pyf[m], dy[j], __builtin_fmaf(pxf[m], dx[j], f_lead)));
                }
                float sn, cs;
                sincos_rev_f32(f, sn, cs);
                if (MASK) {
                    const int lt = lt0 + 16 * m;
                    const float d = __builtin_fmaf(nv[2][lt], rz[j], __builtin_fmaf(nv[1][lt], ry[j], nv[0][lt] * rx[j]));
                    cs = (d < 0.f) ? 0.f : cs;
                    sn = (d < 0.f) ? 0.f : sn;
                }
                mfma::split_pair_fast(cs, sn, ah[j], al[j]);
            }
            const half8 a_hi = __builtin_bit_cast(half8, ah), a_lo = __builtin_bit_cast(half8, al);
            acc0[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi, h0, acc0[m], 0, 0, 0);
            acc1[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi, h1, acc1[m], 0, 0, 0);
            acc0[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi, l0, acc0[m], 0, 0, 0);
            acc1[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi, l1, acc1[m], 0, 0, 0);
            acc0[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_lo, h0, acc0[m], 0, 0, 0);
            acc1[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_lo, h1, acc1[m], 0, 0, 0);
        }
    }
}

// Partial layout [split][24][n_src]: q 0..5 = B, 6..11 = G_x, 12..17 = G_y,
// 18..23 = G_z (re/im xyz interleaved).  Grid: (dish blocks of 64 MB points,
// target splits).
template <int MB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_adj_reverse_mfma(
    const TgtRec32 *__restrict__ tg, const TileMeta *__restrict__ tmeta, int n_tgt, int chunk, GridPtrs src,
    int n_src, double *__restrict__ partial) {
    using mfma::floatx4;
    using mfma::half8;
    __shared__ RevTile tt;
    __shared__ half8 b0h[mfma::KSTEPS][64], b0l[mfma::KSTEPS][64], b1h[mfma::KSTEPS][64], b1l[mfma::KSTEPS][64];
    __shared__ float nv[3][64 * MB];   // normals of the workgroup's dish points (shadow test)
    __shared__ float scratch[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row = lane & 15, kb = lane >> 4;
    const int s0 = blockIdx.x * (64 * MB) + wave * (16 * MB);
    {
        const int ks = tid >> 4, l = mfma::b_slot(ks, 16 * ((tid >> 2) & 3) + 12 + (tid & 3));
        b0h[ks][l] = half8{}; b0l[ks][l] = half8{}; b1h[ks][l] = half8{}; b1l[ks][l] = half8{};
    }
    double px[MB], py[MB], pz[MB];
    float pxf[MB], pyf[MB], pzf[MB];
    float pbx, pby, pbz;   // largest |p_x|, |p_y|, |p_z| over the workgroup's dish points
    Cone cone;
    {
        float fx[MB], fy[MB], fz[MB];
        float mx = 0.f, my = 0.f, mz = 0.f;
        bool any = false;
#pragma unroll
        for (int m = 0; m < MB; m++) {
            const int i = s0 + 16 * m + row;
            const bool valid = i < n_src;
            px[m] = valid ? src.x[i] : 0.0; py[m] = valid ? src.y[i] : 0.0; pz[m] = valid ? src.z[i] : 0.0;
            pxf[m] = (float)px[m]; pyf[m] = (float)py[m]; pzf[m] = (float)pz[m];
            mx = fmaxf(mx, fabsf(pxf[m])); my = fmaxf(my, fabsf(pyf[m])); mz = fmaxf(mz, fabsf(pzf[m]));
            fx[m] = valid ? (float)src.nx[i] : 0.f; fy[m] = valid ? (float)src.ny[i] : 0.f;
            fz[m] = valid ? (float)src.nz[i] : 1.f;
            const int lt = wave * 16 * MB + 16 * m + row;
            nv[0][lt] = fx[m]; nv[1][lt] = fy[m]; nv[2][lt] = fz[m];
            if (m == 0) any = valid;
        }
        cone = block_cone(fx[0], fy[0], fz[0], any, scratch);
        if (MB > 1 && cone.cos_half > 0.f) {
            float d = 1.f;
#pragma unroll
            for (int m = 1; m < MB; m++) {
                float ux = fx[m], uy = fy[m], uz = fz[m];   // directions: the normals need not be unit
                to_unit(ux, uy, uz);
                if (s0 + 16 * m + row < n_src) d = fminf(d, ux * cone.ax + uy * cone.ay + uz * cone.az);
            }
            cone.cos_half = fminf(cone.cos_half, block_min(d, scratch));
        }
        pbx = -block_min(-mx, scratch); pby = -block_min(-my, scratch); pbz = -block_min(-mz, scratch);
    }
    floatx4 acc0[MB], acc1[MB];
    double tot0[MB][4], tot1[MB][4];
#pragma unroll
    for (int m = 0; m < MB; m++) {
        acc0[m] = floatx4{0.f, 0.f, 0.f, 0.f};
        acc1[m] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < 4; i++) { tot0[m][i] = 0.0; tot1[m][i] = 0.0; }
    }
    const int t_begin = blockIdx.y * chunk;
    const int t_end = min(n_tgt, t_begin + chunk);
    for (int base = t_begin; base < t_end; base += TILE) {
        const TileMeta tm = tmeta[base / TILE];
        const float inv = mfma::tile_inv_scale(tm.cmax);
        // A slot past the end carries the last target's direction and no
        // amplitude (zero B entries): its phase only has to stay finite, and
        // a follower whose leader lies past the end keeps a small offset.
        const bool valid = base + tid < t_end;
        TgtRec32 r = tg[valid ? base + tid : t_end - 1];
        if (!valid) {
#pragma unroll
            for (int q = 0; q < 6; q++) r.a[q] = 0.f;
        }
        // offset from the group's leader (same wave); 0 past the end
        const int ls = mfma::leader_slot(lane);
        const double lux = __shfl(r.ux, ls, 64), luy = __shfl(r.uy, ls, 64), luz = __shfl(r.uz, ls, 64);
        const float ddx = valid ? (float)(r.ux - lux) : 0.f;
        const float ddy = valid ? (float)(r.uy - luy) : 0.f;
        const float ddz = valid ? (float)(r.uz - luz) : 0.f;
        const float bound = 1.f + fabsf(ddx) * pbx + fabsf(ddy) * pby + fabsf(ddz) * pbz;
        const bool ok = (lane == ls) || (bound <= mfma::kLocalRev);   // NaN: not ok
        const unsigned long long all_ok = __ballot(ok);
        __syncthreads();
        tt.u[0][tid] = r.ux; tt.u[1][tid] = r.uy; tt.u[2][tid] = r.uz;
        tt.r[0][tid] = r.rx; tt.r[1][tid] = r.ry; tt.r[2][tid] = r.rz;
        tt.du[0][tid] = ddx; tt.du[1][tid] = ddy; tt.du[2][tid] = ddz;
        if (lane == 0) tt.fast[wave] = all_ok == ~0ull;
        build_rev_b(r, tid, inv, b0h, b0l, b1h, b1l);
        __syncthreads();
        const bool fast = __builtin_amdgcn_readfirstlane(tt.fast[0] & tt.fast[1] & tt.fast[2] & tt.fast[3]);
        const int lt0 = wave * 16 * MB + row;
        if (cones_may_shadow(tm.ax, tm.ay, tm.az, tm.cos_half, cone)) {
            if (fast)
                rev_tile_mfma<MB, true, true>(tt, b0h, b0l, b1h, b1l, lane, px, py, pz, pxf, pyf, pzf, nv, lt0, acc0, acc1);
            else
                rev_tile_mfma<MB, true, false>(tt, b0h, b0l, b1h, b1l, lane, px, py, pz, pxf, pyf, pzf, nv, lt0, acc0, acc1);
        } else {
            if (fast)
                rev_tile_mfma<MB, false, true>(tt, b0h, b0l, b1h, b1l, lane, px, py, pz, pxf, pyf, pzf, nv, lt0, acc0, acc1);
            else
                rev_tile_mfma<MB, false, false>(tt, b0h, b0l, b1h, b1l, lane, px, py, pz, pxf, pyf, pzf, nv, lt0, acc0, acc1);
        }
        const double scale = 1.0 / (double)inv;
#pragma unroll
        for (int m = 0; m < MB; m++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                tot0[m][i] = __builtin_fma((double)acc0[m][i], scale, tot0[m][i]);
                tot1[m][i] = __builtin_fma((double)acc1[m][i], scale, tot1[m][i]);
            }
            acc0[m] = floatx4{0.f, 0.f, 0.f, 0.f};
            acc1[m] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
    }
    if (row < 12) {
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int p = s0 + 16 * m + 4 * kb + i;
                if (p < n_src) {
                    partial[((size_t)blockIdx.y * 24 + row) * n_src + p] = tot0[m][i];
                    partial[((size_t)blockIdx.y * 24 + 12 + row) * n_src + p] = tot1[m][i];
                }
            }
    }
}

__global__ __launch_bounds__(256) void k_adj_reverse64(const TgtRec64 *__restrict__ tg, int n_tgt,
                                                       int chunk, GridPtrs src, int n_src,
                                                       double *__restrict__ partial) {
    __shared__ TgtRec64 tile[TILE];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * TILE + tid;
    const bool valid = i < n_src;
    const double px = valid ? src.x[i] : 0.0, py = valid ? src.y[i] : 0.0, pz = valid ? src.z[i] : 0.0;
    const double nx = valid ? src.nx[i] : 0.0, ny = valid ? src.ny[i] : 0.0, nz = valid ? src.nz[i] : 0.0;
    double acc[12];
#pragma unroll
    for (int q = 0; q < 12; q++) acc[q] = 0.0;
    const int t_begin = blockIdx.y * chunk;
    const int t_end = min(n_tgt, t_begin + chunk);
    for (int base = t_begin; base < t_end; base += TILE) {
        __syncthreads();
        const int t = base + tid;
        if (t < t_end) {
            tile[tid] = tg[t];
        } else {
            TgtRec64 z{};
            tile[tid] = z;
        }
        __syncthreads();
#pragma unroll 1
        for (int s = 0; s < TILE; s++) {
            const TgtRec64 &r = tile[s];
            double rev = __builtin_fma(pz, r.uz, __builtin_fma(py, r.uy, px * r.ux));
            double f = __builtin_amdgcn_fract(rev);
            double sn, cs;
            sincos_rev_f64(f, sn, cs);
            double w = __builtin_fma(nz, r.rz, __builtin_fma(ny, r.ry, nx * r.rx));
            cs = (w < 0.0) ? 0.0 : cs;
            sn = (w < 0.0) ? 0.0 : sn;
            const double wc = w * cs, ws = w * sn;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const double ar = r.a[2 * a], ai = r.a[2 * a + 1];
                acc[2 * a] = __builtin_fma(ar, cs, acc[2 * a]);
                acc[2 * a] = __builtin_fma(-ai, sn, acc[2 * a]);
                acc[2 * a + 1] = __builtin_fma(ar, sn, acc[2 * a + 1]);
                acc[2 * a + 1] = __builtin_fma(ai, cs, acc[2 * a + 1]);
                acc[6 + 2 * a] = __builtin_fma(ar, wc, acc[6 + 2 * a]);
                acc[6 + 2 * a] = __builtin_fma(-ai, ws, acc[6 + 2 * a]);
                acc[7 + 2 * a] = __builtin_fma(ar, ws, acc[7 + 2 * a]);
                acc[7 + 2 * a] = __builtin_fma(ai, wc, acc[7 + 2 * a]);
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int q = 0; q < 12; q++) partial[((size_t)blockIdx.y * 12 + q) * n_src + i] = acc[q];
    }
}

// Incident H at each dish point and its directional derivative along the dish
// normal, from the feed's packed FP64 records (w*A*J, w*A*M): fieldAtPoint's
// H integrand (Propagation.h:1084-1119) and its derivative with respect to the
// target position moving along n.  Grid (dish blocks, feed splits of
// fchunk points); output rows [split][12][n]: J = 2 pref n x H (re xyz,
// im xyz), then dJ = 2 pref n x dH, each split's share of the feed sum (the
// combine kernels add the splits).  FP64 throughout.
__global__ __launch_bounds__(256) void k_adj_currents(const SrcRec64 *__restrict__ feed, int n_feed,
                                                      int fchunk, GridPtrs dish, int n_dish, double k,
                                                      double tdir, double pref, double *__restrict__ jd) {
    // feed range of this split; its own 12 rows of jd
    const int f0 = blockIdx.y * fchunk, f1 = min(n_feed, f0 + fchunk);
    jd += (size_t)blockIdx.y * 12 * n_dish;
    __shared__ SrcRec64 tile[TILE];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * TILE + tid;
    const bool valid = i < n_dish;
    const double px = valid ? dish.x[i] : 1.0e6, py = valid ? dish.y[i] : 1.0e6,
                 pz = valid ? dish.z[i] : 1.0e6;
    const double n[3] = {valid ? dish.nx[i] : 0.0, valid ? dish.ny[i] : 0.0, valid ? dish.nz[i] : 1.0};
    double hf[6] = {0, 0, 0, 0, 0, 0};   // H: re xyz, im xyz
    double dh[6] = {0, 0, 0, 0, 0, 0};   // dH
    const double rev_scale = tdir * k / kTwoPi;
    for (int base = f0; base < f1; base += TILE) {
        __syncthreads();
        const int s = base + tid;
        if (s < f1) {
            tile[tid] = feed[s];
        } else {
            SrcRec64 z{};
            z.x = 1.0e7; z.y = 1.0e7; z.z = 1.0e7;
            tile[tid] = z;
        }
        __syncthreads();
        const int cnt = min(TILE, f1 - base);
#pragma unroll 1
        for (int s2 = 0; s2 < cnt; s2++) {
            const SrcRec64 &r = tile[s2];
            const double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
            const double R = sqrt(dx * dx + dy * dy + dz * dz);
            const double Ri = 1.0 / R;
            const double h[3] = {dx * Ri, dy * Ri, dz * Ri};
            if (r.nx * h[0] + r.ny * h[1] + r.nz * h[2] < 0.0) continue;   // shadowed
            const double nu = n[0] * h[0] + n[1] * h[1] + n[2] * h[2];
            const double dh_[3] = {(n[0] - nu * h[0]) * Ri, (n[1] - nu * h[1]) * Ri, (n[2] - nu * h[2]) * Ri};
            const double q = Ri / k, q2 = q * q, q3 = q2 * q;
            const double dq = -k * q2 * nu;
            // s-coefficients and their q-derivatives as (re, im)
            const double s1r = tdir * q2, s1i = q3 - q;
            const double s2r = -3.0 * tdir * q2, s2i = q - 3.0 * q3;
            const double s3r = -q2, s3i = tdir * q;
            const double d1r = 2.0 * tdir * q * dq, d1i = (3.0 * q2 - 1.0) * dq;
            const double d2r = -6.0 * tdir * q * dq, d2i = (1.0 - 9.0 * q2) * dq;
            const double d3r = -2.0 * q * dq, d3i = tdir * dq;
            double sn, cs;
            sincos_rev_f64(__builtin_amdgcn_fract(R * rev_scale), sn, cs);   // G = exp(j t k R)
            // dG = j t k nu G
            const double tkn = tdir * k * nu;
            const double dgr = -tkn * sn, dgi = tkn * cs;
            const double *J = r.c, *M = r.c + 6;   // (re, im) interleaved per axis
            double mdh_r = 0, mdh_i = 0, mh_r = 0, mh_i = 0;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                mh_r += M[2 * a] * h[a]; mh_i += M[2 * a + 1] * h[a];
                mdh_r += M[2 * a] * dh_[a]; mdh_i += M[2 * a + 1] * dh_[a];
            }
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int b = (a + 1) % 3, c = (a + 2) % 3;
                const double jxh_r = J[2 * b] * h[c] - J[2 * c] * h[b];
                const double jxh_i = J[2 * b + 1] * h[c] - J[2 * c + 1] * h[b];
                const double jxd_r = J[2 * b] * dh_[c] - J[2 * c] * dh_[b];
                const double jxd_i = J[2 * b + 1] * dh_[c] - J[2 * c + 1] * dh_[b];
                // T = M s1 + (M.h) h s2 - (J x h) s3
                const double v2r = mh_r * h[a], v2i = mh_i * h[a];
                double Tr = (M[2 * a] * s1r - M[2 * a + 1] * s1i) + (v2r * s2r - v2i * s2i) -
                            (jxh_r * s3r - jxh_i * s3i);
                double Ti = (M[2 * a] * s1i + M[2 * a + 1] * s1r) + (v2r * s2i + v2i * s2r) -
                            (jxh_r * s3i + jxh_i * s3r);
                // dT = M ds1 + [(M.dh) h + (M.h) dh] s2 + (M.h) h ds2 - (J x dh) s3 - (J x h) ds3
                const double w2r = mdh_r * h[a] + mh_r * dh_[a], w2i = mdh_i * h[a] + mh_i * dh_[a];
                double dTr = (M[2 * a] * d1r - M[2 * a + 1] * d1i) + (w2r * s2r - w2i * s2i) +
                             (v2r * d2r - v2i * d2i) - (jxd_r * s3r - jxd_i * s3i) -
                             (jxh_r * d3r - jxh_i * d3i);
                double dTi = (M[2 * a] * d1i + M[2 * a + 1] * d1r) + (w2r * s2i + w2i * s2r) +
                             (v2r * d2i + v2i * d2r) - (jxd_r * s3i + jxd_i * s3r) -
                             (jxh_r * d3i + jxh_i * d3r);
                hf[a] += cs * Tr - sn * Ti;
                hf[3 + a] += cs * Ti + sn * Tr;
                // d(G T) = dG T + G dT
                dh[a] += (dgr * Tr - dgi * Ti) + (cs * dTr - sn * dTi);
                dh[3 + a] += (dgr * Ti + dgi * Tr) + (cs * dTi + sn * dTr);
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int b = (a + 1) % 3, c = (a + 2) % 3;
            const double f = 2.0 * pref;
            jd[(size_t)a * n_dish + i] = f * (n[b] * hf[c] - n[c] * hf[b]);
            jd[(size_t)(3 + a) * n_dish + i] = f * (n[b] * hf[3 + c] - n[c] * hf[3 + b]);
            jd[(size_t)(6 + a) * n_dish + i] = f * (n[b] * dh[c] - n[c] * dh[b]);
            jd[(size_t)(9 + a) * n_dish + i] = f * (n[b] * dh[3 + c] - n[c] * dh[3 + b]);
        }
    }
}

// Mixed-precision twin of k_adj_currents for the PYPO_PREC_MIXED path: FP64
// difference vector, distance (v_rsq_f32 + one FP64 Newton step) and phase;
// FP32 unit vectors, s-coefficients, their derivatives and the per-tile
// sums (folded into FP64).  Feed records are the mixed-precision SrcRec32.
__global__ __launch_bounds__(256) void k_adj_currents32(const SrcRec32 *__restrict__ feed, int n_feed,
                                                        int fchunk, GridPtrs dish, int n_dish, double k,
                                                        double tdir, double pref, double *__restrict__ jd) {
    // feed range of this split; its own 12 rows of jd
    const int f0 = blockIdx.y * fchunk, f1 = min(n_feed, f0 + fchunk);
    jd += (size_t)blockIdx.y * 12 * n_dish;
    __shared__ SrcRec32 tile[TILE];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * TILE + tid;
    const bool valid = i < n_dish;
    const double px = valid ? dish.x[i] : 1.0e6, py = valid ? dish.y[i] : 1.0e6,
                 pz = valid ? dish.z[i] : 1.0e6;
    const float n[3] = {valid ? (float)dish.nx[i] : 0.f, valid ? (float)dish.ny[i] : 0.f,
                        valid ? (float)dish.nz[i] : 1.f};
    const double rev_scale = tdir * k / kTwoPi;   // G = exp(j t k R)
    const float kf = (float)k, kinv = (float)(1.0 / k), tf = (float)tdir;
    float hf[6], dh[6];      // per-tile sums: H and dH (re xyz, im xyz)
    double HF[6], DH[6];     // totals
#pragma unroll
    for (int q = 0; q < 6; q++) { hf[q] = 0.f; dh[q] = 0.f; HF[q] = 0.0; DH[q] = 0.0; }
    for (int base = f0; base < f1; base += TILE) {
        __syncthreads();
        const int s = base + tid;
        if (s < f1) {
            tile[tid] = feed[s];
        } else {
            SrcRec32 z{};
            z.x = 1.0e7; z.y = 1.0e7; z.z = 1.0e7;
            tile[tid] = z;
        }
        __syncthreads();
        const int cnt = min(TILE, f1 - base);
#pragma unroll 1
        for (int s2 = 0; s2 < cnt; s2++) {
            const SrcRec32 &r = tile[s2];
            const double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
            const double R2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
            const float y0 = __builtin_amdgcn_rsqf((float)R2);
            const double yd = (double)y0;
            const double tR = R2 * yd;
            const double e = __builtin_fma(-tR, yd, 1.0);
            const double R = __builtin_fma(0.5 * tR, e, tR);
            float sn, cs;
            sincos_rev_f32((float)__builtin_amdgcn_fract(R * rev_scale), sn, cs);
            const float Ri = __builtin_fmaf(0.5f * y0, (float)e, y0);
            const float h[3] = {(float)dx * Ri, (float)dy * Ri, (float)dz * Ri};
            const float m = __builtin_fmaf(r.nz, h[2], __builtin_fmaf(r.ny, h[1], r.nx * h[0]));
            cs = (m < 0.f) ? 0.f : cs;   // shadowed: G = dG = 0
            sn = (m < 0.f) ? 0.f : sn;
            const float nu = __builtin_fmaf(n[2], h[2], __builtin_fmaf(n[1], h[1], n[0] * h[0]));
            const float dv[3] = {(n[0] - nu * h[0]) * Ri, (n[1] - nu * h[1]) * Ri, (n[2] - nu * h[2]) * Ri};
            const float q = Ri * kinv, q2 = q * q, q3 = q2 * q;
            const float dq = -kf * q2 * nu;
            const float s1r = tf * q2, s1i = q3 - q;
            const float s2r = -3.f * tf * q2, s2i = __builtin_fmaf(-3.f, q3, q);
            const float s3r = -q2, s3i = tf * q;
            const float d1r = 2.f * tf * q * dq, d1i = __builtin_fmaf(3.f, q2, -1.f) * dq;
            const float d2r = -6.f * tf * q * dq, d2i = __builtin_fmaf(-9.f, q2, 1.f) * dq;
            const float d3r = -2.f * q * dq, d3i = tf * dq;
            const float tkn = tf * kf * nu;
            const float dgr = -tkn * sn, dgi = tkn * cs;
            const float *J = r.cj, *M = r.cm;
            float mh_r = 0.f, mh_i = 0.f, md_r = 0.f, md_i = 0.f;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                mh_r = __builtin_fmaf(M[2 * a], h[a], mh_r); mh_i = __builtin_fmaf(M[2 * a + 1], h[a], mh_i);
                md_r = __builtin_fmaf(M[2 * a], dv[a], md_r); md_i = __builtin_fmaf(M[2 * a + 1], dv[a], md_i);
            }
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int b = (a + 1) % 3, c = (a + 2) % 3;
                const float xr = J[2 * b] * h[c] - J[2 * c] * h[b];
                const float xi = J[2 * b + 1] * h[c] - J[2 * c + 1] * h[b];
                const float yr = J[2 * b] * dv[c] - J[2 * c] * dv[b];
                const float yi = J[2 * b + 1] * dv[c] - J[2 * c + 1] * dv[b];
                const float vr = mh_r * h[a], vi = mh_i * h[a];
                const float wr = md_r * h[a] + mh_r * dv[a], wi = md_i * h[a] + mh_i * dv[a];
                const float Mr = M[2 * a], Mi = M[2 * a + 1];
                // T = M s1 + (M.h) h s2 - (J x h) s3
                const float Tr = (Mr * s1r - Mi * s1i) + (vr * s2r - vi * s2i) - (xr * s3r - xi * s3i);
                const float Ti = (Mr * s1i + Mi * s1r) + (vr * s2i + vi * s2r) - (xr * s3i + xi * s3r);
                // dT = M ds1 + [(M.dh) h + (M.h) dh] s2 + (M.h) h ds2 - (J x dh) s3 - (J x h) ds3
                const float dTr = (Mr * d1r - Mi * d1i) + (wr * s2r - wi * s2i) + (vr * d2r - vi * d2i) -
                                  (yr * s3r - yi * s3i) - (xr * d3r - xi * d3i);
                const float dTi = (Mr * d1i + Mi * d1r) + (wr * s2i + wi * s2r) + (vr * d2i + vi * d2r) -
                                  (yr * s3i + yi * s3r) - (xr * d3i + xi * d3r);
                hf[a] += cs * Tr - sn * Ti;
                hf[3 + a] += cs * Ti + sn * Tr;
                dh[a] += (dgr * Tr - dgi * Ti) + (cs * dTr - sn * dTi);
                dh[3 + a] += (dgr * Ti + dgi * Tr) + (cs * dTi + sn * dTr);
            }
        }
#pragma unroll
        for (int q = 0; q < 6; q++) { HF[q] += (double)hf[q]; DH[q] += (double)dh[q]; hf[q] = 0.f; dh[q] = 0.f; }
    }
    if (valid) {
        const double nd[3] = {dish.nx[i], dish.ny[i], dish.nz[i]};
        const double f = 2.0 * pref;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int b = (a + 1) % 3, c = (a + 2) % 3;
            jd[(size_t)a * n_dish + i] = f * (nd[b] * HF[c] - nd[c] * HF[b]);
            jd[(size_t)(3 + a) * n_dish + i] = f * (nd[b] * HF[3 + c] - nd[c] * HF[3 + b]);
            jd[(size_t)(6 + a) * n_dish + i] = f * (nd[b] * DH[c] - nd[c] * DH[b]);
            jd[(size_t)(9 + a) * n_dish + i] = f * (nd[b] * DH[3 + c] - nd[c] * DH[3 + b]);
        }
    }
}

// dL/dd_i = c_i [ Re(B_i . dJ_i) + t k Im(J_i . G_i) ]
// The 12 rows of J and dJ/dd at point i, summed over the feed splits (the
// first split's twelve loads issue together).
__device__ __forceinline__ void load_jd(const double *__restrict__ jd, int jsplits, int n, int i, double (&v)[12]) {
#pragma unroll
    for (int q = 0; q < 12; q++) v[q] = jd[(size_t)q * n + i];
    for (int s = 1; s < jsplits; s++)
#pragma unroll
        for (int q = 0; q < 12; q++) v[q] += jd[((size_t)s * 12 + q) * n + i];
}

__global__ __launch_bounds__(256) void k_adj_combine(const double *__restrict__ partial, int splits,
                                                     GridPtrs dish, int n, int gmode, int ncx,
                                                     int ncy, const double *__restrict__ jd,
                                                     int jsplits, double k, double tdir,
                                                     double *__restrict__ grad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double S[12];   // split sums, the first split's loads issued together
#pragma unroll
    for (int q = 0; q < 12; q++) S[q] = partial[(size_t)q * n + i];
    for (int s = 1; s < splits; s++)
#pragma unroll
        for (int q = 0; q < 12; q++) S[q] += partial[((size_t)s * 12 + q) * n + i];
    double JD[12];
    load_jd(jd, jsplits, n, i, JD);
    double re_bdj = 0.0, im_jg = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double br = S[2 * a], bi = S[2 * a + 1], gr = S[6 + 2 * a], gi = S[7 + 2 * a];
        const double jr = JD[a], ji = JD[3 + a], dr = JD[6 + a], di = JD[9 + a];
        re_bdj += br * dr - bi * di;
        im_jg += jr * gi + ji * gr;
    }
    const double c = trapezoid_weight(i, gmode, ncx, ncy) * dish.area[i];
    grad[i] = c * (re_bdj + tdir * k * im_jg);
}

// The 24-row layout of k_adj_reverse_mfma: G = n_x G_x + n_y G_y + n_z G_z in FP64.
__global__ __launch_bounds__(256) void k_adj_combine24(const double *__restrict__ partial, int splits,
                                                       GridPtrs dish, int n, int gmode, int ncx, int ncy,
                                                       const double *__restrict__ jd, int jsplits, double k,
                                                       double tdir, double *__restrict__ grad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double S[24];   // split sums, the first split's loads issued together
#pragma unroll
    for (int q = 0; q < 24; q++) S[q] = partial[(size_t)q * n + i];
    for (int s = 1; s < splits; s++)
#pragma unroll
        for (int q = 0; q < 24; q++) S[q] += partial[((size_t)s * 24 + q) * n + i];
    const double nrm[3] = {dish.nx[i], dish.ny[i], dish.nz[i]};
    double JD[12];
    load_jd(jd, jsplits, n, i, JD);
    double re_bdj = 0.0, im_jg = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double br = S[2 * a], bi = S[2 * a + 1];
        double gr = 0.0, gi = 0.0;
#pragma unroll
        for (int b = 0; b < 3; b++) {
            gr += nrm[b] * S[6 + 6 * b + 2 * a];
            gi += nrm[b] * S[7 + 6 * b + 2 * a];
        }
        const double jr = JD[a], ji = JD[3 + a], dr = JD[6 + a], di = JD[9 + a];
        re_bdj += br * dr - bi * di;
        im_jg += jr * gi + ji * gr;
    }
    const double c = trapezoid_weight(i, gmode, ncx, ncy) * dish.area[i];
    grad[i] = c * (re_bdj + tdir * k * im_jg);
}

__global__ __launch_bounds__(256) void k_displace(GridPtrs g0, const double *__restrict__ d, int n,
                                                  GridPtrs out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double di = d[i];
    out.x[i] = g0.x[i] + di * g0.nx[i];
    out.y[i] = g0.y[i] + di * g0.ny[i];
    out.z[i] = g0.z[i] + di * g0.nz[i];
    out.nx[i] = g0.nx[i];
    out.ny[i] = g0.ny[i];
    out.nz[i] = g0.nz[i];
    out.area[i] = g0.area[i];
}

}  // namespace

int launch_displace(const GridPtrs &g0, const double *d, int n, const GridPtrs &out, hipStream_t st) {
    PYPO_REQUIRE(n > 0 && d, ERR_ARG, "displace: empty grid or NULL displacement");
    hipLaunchKernelGGL(k_displace, dim3((n + 255) / 256), dim3(256), 0, st, g0, d, n, out);
    PYPO_CHECK_LAUNCH();
    return 0;
}

namespace {

// Target splits of the reverse sum: enough workgroups to fill the chip,
// at least four 256-target tiles per split.  The matrix kernel (two
// workgroups per CU) picks the split count, up to 8, that minimises the
// rounds of concurrent workgroups times the tiles each one runs.
struct AdjPlan {
    bool matrix;   // reverse sum on the matrix cores (mixed precision)
    int rows;      // partial rows per split: 24 (matrix) or 12
    int spl;       // dish points per lane in the reverse sum
    int blocks;    // workgroups of one point per lane
    int rblocks;   // reverse-sum workgroups along the dish points
    int splits, chunk;
    int fsplits, fchunk;   // feed splits of the currents pass (grid.y) and feed points per split
    size_t tg_bytes, feed_bytes, part_bytes, jd_bytes;
};

AdjPlan adj_plan(int precision, int n_feed, int n_dish, int n_tgt) {
    AdjPlan p;
    p.matrix = precision != PYPO_PREC_FP64 && ff_kernel() == PYPO_FF_MATRIX;
    p.rows = p.matrix ? 24 : 12;
    p.spl = precision == PYPO_PREC_FP64 ? 1 : 2;
    p.blocks = (n_dish + TILE - 1) / TILE;
    // matrix reverse sum: 4 M-blocks of 16 dish points per wave, 256 per workgroup
    p.rblocks = p.matrix ? (n_dish + TILE - 1) / TILE : (n_dish + TILE * p.spl - 1) / (TILE * p.spl);
    int splits = (2048 + p.rblocks - 1) / p.rblocks;
    const int max_splits = (n_tgt + 4 * TILE - 1) / (4 * TILE);
    if (p.matrix)
        splits = tail_splits(p.rblocks, (n_tgt + TILE - 1) / TILE, 1, std::min(8, std::max(1, max_splits)),
                             2LL * device_cus());
    splits = std::max(1, std::min(splits, max_splits));
    p.chunk = ((n_tgt + splits - 1) / splits + TILE - 1) / TILE * TILE;
    p.splits = (n_tgt + p.chunk - 1) / p.chunk;
    // The currents pass runs one dish point per lane over all feed points:
    // split the feed (up to 8 ways, whole tiles) when the dish blocks alone
    // leave the last round of workgroups part-empty.  Only at low occupancy:
    // at four or more workgroups per CU the waves of a part-empty last round
    // share their SIMDs with fewer others and finish that much sooner (601^2
    // dish, 33^2 feed: FP64 kernel, 2 per CU, 4.99 -> 4.31 ms with 5 splits;
    // mixed kernel, 4 per CU, 2.53 -> 2.58 ms).
    {
        static const int occ64 = blocks_per_cu(reinterpret_cast<const void *>(k_adj_currents));
        static const int occ32 = blocks_per_cu(reinterpret_cast<const void *>(k_adj_currents32));
        const int ftiles = (n_feed + TILE - 1) / TILE;
        const int occ = precision == PYPO_PREC_FP64 ? occ64 : occ32;
        const int fs = occ >= 4 ? 1
                                : tail_splits(p.blocks, ftiles, 1, std::max(1, std::min(8, ftiles)),
                                              (long long)occ * device_cus());
        p.fchunk = (ftiles + fs - 1) / fs * TILE;
        p.fsplits = (n_feed + p.fchunk - 1) / p.fchunk;
    }
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    p.tg_bytes = up((size_t)n_tgt * (precision == PYPO_PREC_FP64 ? sizeof(TgtRec64) : sizeof(TgtRec32))) +
                 up((size_t)((n_tgt + TILE - 1) / TILE) * sizeof(TileMeta));
    p.feed_bytes = up(packed_bytes(PYPO_MODE_JM, precision, n_feed));
    p.part_bytes = up((size_t)p.splits * p.rows * n_dish * sizeof(double));
    p.jd_bytes = up((size_t)p.fsplits * 12 * n_dish * sizeof(double));
    return p;
}

}  // namespace

size_t adjoint_workspace_bytes(int precision, int n_feed, int n_dish, int n_tgt) {
    AdjPlan p = adj_plan(precision, n_feed, n_dish, n_tgt);
    return p.tg_bytes + p.feed_bytes + p.part_bytes + p.jd_bytes;
}

int launch_ff_power_grad(int precision, const GridPtrs &dish, int n_dish, int gmode, int ncx, int ncy,
                         const GridPtrs &feed, int n_feed, int fgmode, int fncx,
                         int fncy, const Bundle12 &feed_cur, const GridPtrs &ff, int n_tgt,
                         const Bundle12 &E, const double *g, double k, double tdir, double *grad,
                         void *ws, size_t ws_bytes, hipStream_t st) {
    PYPO_REQUIRE(n_dish > 0 && n_feed > 0 && n_tgt > 0, ERR_ARG, "adjoint: empty grid");
    PYPO_REQUIRE((long long)ncx * ncy == n_dish && (long long)fncx * fncy == n_feed, ERR_ARG,
                 "adjoint: n_cells do not match the grid sizes");
    PYPO_REQUIRE(ws_bytes >= adjoint_workspace_bytes(precision, n_feed, n_dish, n_tgt), ERR_ARG,
                 "adjoint: workspace too small");
    const PhysConsts pc = phys_consts(k, 1.0);
    const bool fp64 = precision == PYPO_PREC_FP64;
    const AdjPlan p = adj_plan(precision, n_feed, n_dish, n_tgt);
    char *w = (char *)ws;
    void *tg = w;
    void *fpk = w + p.tg_bytes;
    double *part = (double *)(w + p.tg_bytes + p.feed_bytes);
    double *jd = (double *)(w + p.tg_bytes + p.feed_bytes + p.part_bytes);
    const int blocks = p.blocks, splits = p.splits, chunk = p.chunk;
    const double rev_scale = -tdir * k / kTwoPi;
    const int tb = (n_tgt + 255) / 256;
    if (fp64)
        hipLaunchKernelGGL(k_adj_targets<TgtRec64>, dim3(tb), dim3(256), 0, st, ff.x, ff.y, E, g, n_tgt,
                           rev_scale, pc.prefactor, (TgtRec64 *)tg);
    else
        hipLaunchKernelGGL(k_adj_targets<TgtRec32>, dim3(tb), dim3(256), 0, st, ff.x, ff.y, E, g, n_tgt,
                           rev_scale, pc.prefactor, (TgtRec32 *)tg);
    PYPO_CHECK_LAUNCH();
    int err = launch_pack_sources(precision, feed, n_feed, feed_cur, fgmode, fncx, fncy, fpk, st);
    if (err) return err;
    if (fp64)
        hipLaunchKernelGGL(k_adj_currents, dim3(blocks, p.fsplits), dim3(TILE), 0, st, (const SrcRec64 *)fpk,
                           n_feed, p.fchunk, dish, n_dish, k, tdir, pc.prefactor, jd);
    else
        hipLaunchKernelGGL(k_adj_currents32, dim3(blocks, p.fsplits), dim3(TILE), 0, st, (const SrcRec32 *)fpk,
                           n_feed, p.fchunk, dish, n_dish, k, tdir, pc.prefactor, jd);
    PYPO_CHECK_LAUNCH();
    if (fp64) {
        hipLaunchKernelGGL(k_adj_reverse64, dim3(p.rblocks, splits), dim3(TILE), 0, st,
                           (const TgtRec64 *)tg, n_tgt, chunk, dish, n_dish, part);
    } else if (p.matrix) {
        TileMeta *tmeta = (TileMeta *)((char *)tg + (((size_t)n_tgt * sizeof(TgtRec32) + 255) / 256 * 256));
        hipLaunchKernelGGL(k_adj_tgt_meta, dim3((n_tgt + TILE - 1) / TILE), dim3(TILE), 0, st,
                           (const TgtRec32 *)tg, n_tgt, tmeta);
        PYPO_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_adj_reverse_mfma<4>, dim3(p.rblocks, splits), dim3(TILE), 0, st,
                           (const TgtRec32 *)tg, (const TileMeta *)tmeta, n_tgt, chunk, dish, n_dish, part);
    } else {
        hipLaunchKernelGGL(k_adj_reverse32<2>, dim3(p.rblocks, splits), dim3(TILE), 0, st,
                           (const TgtRec32 *)tg, n_tgt, chunk, dish, n_dish, part);
    }
    PYPO_CHECK_LAUNCH();
    if (p.matrix)
        hipLaunchKernelGGL(k_adj_combine24, dim3(blocks), dim3(256), 0, st, part, splits, dish, n_dish, gmode,
                           ncx, ncy, jd, p.fsplits, k, tdir, grad);
    else
        hipLaunchKernelGGL(k_adj_combine, dim3(blocks), dim3(256), 0, st, part, splits, dish, n_dish, gmode,
                           ncx, ncy, jd, p.fsplits, k, tdir, grad);
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
