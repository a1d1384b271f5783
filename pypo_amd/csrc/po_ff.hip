// po_ff.hip -- far-field PO sum (farfieldAtPoint, Propagation.h:1656-1781).
//
// Factored per target direction r:
//   E = -j pref [SJ - (SJ.r) r + t r x SM],  H = -j pref [SM - (SM.r) r - t r x SJ]
//   SJ = sum_s w_s A_s J_s exp(-j t k r_s.r),  SM likewise with M,
// over the sources with n_s.r >= 0.  The sum kernels accumulate only SJ
// and SM -- six complex multiply-adds per pair -- and the finish kernel
// applies the projection, the cross product and -j pref in FP64.
//
// Per-pair work (mixed precision): the phase in revolutions
// rev = x ux + y uy + z uz (FP64, u = r * (-t k / 2 pi)), its fraction
// (v_fract_f64), FP32 sin/cos of the fraction, then FMAs; in tiles where the
// follower bound holds, one such FP64 phase per group of eight neighbouring
// records and FP32 offsets for the other seven (ff_tile_follow).  Per tile (256
// sources) two uniform decisions pick the inner loop: the shadow test is
// skipped when the tile's normal cone and the workgroup's direction cone
// prove n.r > 0 for every pair, and the M terms are skipped when the tile
// carries no magnetic current (always so for PEC currents from the JM mode).
// The per-tile FP32 partial sums are folded into FP64 accumulators.

#include "po_tiles.hpp"

namespace pypo {
namespace {

// Largest bound on a follower phase sum (revolutions): FP32 rounds such a sum
// to 2^-24 * 8 = 5e-7 revolutions at worst, and v_sin/v_cos reduce it
// themselves.
constexpr float kFollowRev = 8.f;

// The per-tile loop with the follower phase (as in the matrix-core kernel,
// po_ff_mfma.hpp): records come in groups of kGroup grid neighbours (the
// packer writes them in serpentine order); the fourth of a group (the
// leader, at most four records from any other) gets the FP64 phase, the
// other seven
//   f = f_leader + d . u        (three FP32 FMAs)
// with d = (r_s - r_leader) * rev_scale in FP32, written into the record's
// padding (pad0, pad1) when the tile was staged, and u the FP32 direction.
// Taken only when every follower's 1 + sum_i |d_i| ub_i <= kFollowRev.
// Measured on the large far field with PEC currents (kernel only): no
// follower form 121.4 ms, groups of 4 116.3 ms, groups of 8 108.1 ms,
// groups of 16 107.9 ms (twice the offsets for 0.2 % more).
constexpr int kGroup = 8, kLeader = 3;
__device__ __forceinline__ int group_leader(int lane) { return (lane & ~(kGroup - 1)) | kLeader; }
template <int TPL, bool HAS_M, bool MASK>
__device__ __forceinline__ void ff_tile_follow(const SrcRec32 *tile, const double (&rx)[TPL],
                                               const double (&ry)[TPL], const double (&rz)[TPL],
                                               const float (&fx)[TPL], const float (&fy)[TPL],
                                               const float (&fz)[TPL], float (&acc)[TPL][12]) {
#pragma unroll 1
    for (int s0 = 0; s0 < TILE; s0 += kGroup) {
        float f_lead[TPL];
#pragma unroll
        for (int k = 0; k < kGroup; k++) {
            const int js = k == 0 ? kLeader : (k <= kLeader ? k - 1 : k);   // the leader first
            const SrcRec32 &r = tile[s0 + js];
            float cj[6], cm[6];
#pragma unroll
            for (int q = 0; q < 6; q++) cj[q] = r.cj[q];
            if (HAS_M) {
#pragma unroll
                for (int q = 0; q < 6; q++) cm[q] = r.cm[q];
            }
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                float f;
                if (js == kLeader) {
                    const double rev = __builtin_fma(r.z, rz[j], __builtin_fma(r.y, ry[j], r.x * rx[j]));
                    f = (float)__builtin_amdgcn_fract(rev);
                    f_lead[j] = f;
                } else {
                    f = __builtin_fmaf(r.pad1[1], fz[j], __builtin_fmaf(r.pad1[0], fy[j],
                                                                        __builtin_fmaf(r.pad0, fx[j], f_lead[j])));
                }
                float sn, cs;
                sincos_rev_f32(f, sn, cs);
                if (MASK) {
                    const float m = __builtin_fmaf(r.nz, fz[j], __builtin_fmaf(r.ny, fy[j], r.nx * fx[j]));
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
}

template <int TPL, bool HAS_M>
__device__ __forceinline__ void ff_tile_vector(bool mask, bool follow, const SrcRec32 *tile,
                                               const double (&rx)[TPL], const double (&ry)[TPL],
                                               const double (&rz)[TPL], const float (&fx)[TPL],
                                               const float (&fy)[TPL], const float (&fz)[TPL],
                                               float (&acc)[TPL][12]) {
    if (follow) {
        if (mask) ff_tile_follow<TPL, HAS_M, true>(tile, rx, ry, rz, fx, fy, fz, acc);
        else ff_tile_follow<TPL, HAS_M, false>(tile, rx, ry, rz, fx, fy, fz, acc);
    } else {
        if (mask) ff_tile_mixed<TPL, HAS_M, true>(tile, rx, ry, rz, fx, fy, fz, acc);
        else ff_tile_mixed<TPL, HAS_M, false>(tile, rx, ry, rz, fx, fy, fz, acc);
    }
}

template <int TPL>
__global__ __launch_bounds__(256) void k_ff_mixed(const SrcRec32 *__restrict__ src,
                                                  const TileMeta *__restrict__ meta, int n_src,
                                                  int chunk, const double *__restrict__ th,
                                                  const double *__restrict__ ph, int n_tgt,
                                                  double rev_scale, double *__restrict__ partial) {
    __shared__ SrcRec32 tile[TILE];
    __shared__ float scratch[4];
    __shared__ int follow_ok[4];   // per wave: every follower of its 64 records within the bound
    const int tid = threadIdx.x, lane = tid & 63;
    int tb, sp;
    xcd_item(tb, sp);
    double rx[TPL], ry[TPL], rz[TPL];
    float fx[TPL], fy[TPL], fz[TPL];
    Cone cone;
    {
        float cx = 0.f, cy = 0.f, cz = 0.f;
        bool any = false;
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            int t = tb * (TILE * TPL) + j * TILE + tid;
            double x = 0, y = 0, z = 0;
            if (t < n_tgt) {
                double st, ct, sp, cp;
                sincos(th[t], &st, &ct);
                sincos(ph[t], &sp, &cp);
                x = cp * st; y = sp * st; z = ct;
            }
            fx[j] = (float)x; fy[j] = (float)y; fz[j] = (float)z;
            rx[j] = x * rev_scale; ry[j] = y * rev_scale; rz[j] = z * rev_scale;
            if (j == 0) { cx = fx[0]; cy = fy[0]; cz = fz[0]; any = t < n_tgt; }
        }
        // direction cone of the workgroup (first target of each lane, then
        // widened by the other targets' deviation)
        cone = block_cone(cx, cy, cz, any, scratch);
        if (TPL > 1 && cone.cos_half > 0.f) {
            float d = 1.f;
#pragma unroll
            for (int j = 1; j < TPL; j++) {
                int t = tb * (TILE * TPL) + j * TILE + tid;
                if (t < n_tgt) d = fminf(d, fx[j] * cone.ax + fy[j] * cone.ay + fz[j] * cone.az);
            }
            cone.cos_half = fminf(cone.cos_half, block_min(d, scratch));
        }
    }
    float acc[TPL][12];
    double tot[TPL][12];
#pragma unroll
    for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int q = 0; q < 12; q++) { acc[j][q] = 0.f; tot[j][q] = 0.0; }
    // per-component bound on |u| over the workgroup's targets: |a_i| + sin(theta)
    float ubx = 1.f, uby = 1.f, ubz = 1.f;
    if (cone.cos_half > 0.f) {
        const float sn = sqrtf(fmaxf(0.f, 1.f - cone.cos_half * cone.cos_half)) + 1e-3f;
        ubx = fminf(1.f, fabsf(cone.ax) + sn);
        uby = fminf(1.f, fabsf(cone.ay) + sn);
        ubz = fminf(1.f, fabsf(cone.az) + sn);
    }

    const int s_begin = sp * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<6> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        {
            // this lane's follower offset from its group's leader (same wave), in revolutions,
            // written into the padding of its record's LDS copy
            const SrcRec32 r = __builtin_bit_cast(SrcRec32, ld.buf);
            const int ls = group_leader(lane);
            const float ddx = (float)((r.x - __shfl(r.x, ls, 64)) * rev_scale);
            const float ddy = (float)((r.y - __shfl(r.y, ls, 64)) * rev_scale);
            const float ddz = (float)((r.z - __shfl(r.z, ls, 64)) * rev_scale);
            const float bound = 1.f + fabsf(ddx) * ubx + fabsf(ddy) * uby + fabsf(ddz) * ubz;
            const bool ok = (lane == ls) || (bound <= kFollowRev);   // NaN: not ok
            const unsigned long long all_ok = __ballot(ok);
            ld.buf[2].y = __float_as_uint(ddx);   // pad0
            ld.buf[5].z = __float_as_uint(ddy);   // pad1[0]
            ld.buf[5].w = __float_as_uint(ddz);   // pad1[1]
            __syncthreads();
            ld.store(reinterpret_cast<uint4 *>(tile), tid);
            if (lane == 0) follow_ok[tid >> 6] = all_ok == ~0ull;
            __syncthreads();
        }
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
        const TileMeta m = meta[base / TILE];
        const bool mask = cones_may_shadow(m.ax, m.ay, m.az, m.cos_half, cone);
        const bool follow =
            __builtin_amdgcn_readfirstlane(follow_ok[0] & follow_ok[1] & follow_ok[2] & follow_ok[3]);
        if (m.has_m) ff_tile_vector<TPL, true>(mask, follow, tile, rx, ry, rz, fx, fy, fz, acc);
        else ff_tile_vector<TPL, false>(mask, follow, tile, rx, ry, rz, fx, fy, fz, acc);
#pragma unroll
        for (int j = 0; j < TPL; j++)
#pragma unroll
            for (int q = 0; q < 12; q++) { tot[j][q] += (double)acc[j][q]; acc[j][q] = 0.f; }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = tb * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
#pragma unroll
            for (int q = 0; q < 12; q++)
                partial[((size_t)sp * 12 + q) * n_tgt + t] = tot[j][q];
        }
    }
}

template <int TPL, bool HAS_M, bool MASK>
__device__ __forceinline__ void ff_tile_fp64(const SrcRec64 *tile, const double (&rx)[TPL],
                                             const double (&ry)[TPL], const double (&rz)[TPL],
                                             const double (&ux)[TPL], const double (&uy)[TPL],
                                             const double (&uz)[TPL], double (&acc)[TPL][12]) {
#pragma unroll 1
    for (int s = 0; s < TILE; s++) {
        const SrcRec64 &r = tile[s];
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            double rev = __builtin_fma(r.z, rz[j], __builtin_fma(r.y, ry[j], r.x * rx[j]));
            double f = __builtin_amdgcn_fract(rev);
            double sn, cs;
            sincos_rev_f64(f, sn, cs);
            if (MASK) {
                double m = __builtin_fma(r.nz, uz[j], __builtin_fma(r.ny, uy[j], r.nx * ux[j]));
                cs = (m < 0.0) ? 0.0 : cs;
                sn = (m < 0.0) ? 0.0 : sn;
            }
#pragma unroll
            for (int q = 0; q < (HAS_M ? 6 : 3); q++) {
                double &ar = acc[j][2 * q], &ai = acc[j][2 * q + 1];
                ar = __builtin_fma(r.c[2 * q], cs, ar);
                ar = __builtin_fma(-r.c[2 * q + 1], sn, ar);
                ai = __builtin_fma(r.c[2 * q], sn, ai);
                ai = __builtin_fma(r.c[2 * q + 1], cs, ai);
            }
        }
    }
}

template <int TPL>
__global__ __launch_bounds__(256) void k_ff_fp64(const SrcRec64 *__restrict__ src,
                                                 const TileMeta *__restrict__ meta, int n_src,
                                                 int chunk, const double *__restrict__ th,
                                                 const double *__restrict__ ph, int n_tgt,
                                                 double rev_scale, double *__restrict__ partial) {
    __shared__ SrcRec64 tile[TILE];
    __shared__ float scratch[4];
    const int tid = threadIdx.x;
    int tb, sp;
    xcd_item(tb, sp);
    double ux[TPL], uy[TPL], uz[TPL];
    double rx[TPL], ry[TPL], rz[TPL];
    Cone cone;
    {
        bool any = false;
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            int t = tb * (TILE * TPL) + j * TILE + tid;
            double x = 0, y = 0, z = 0;
            if (t < n_tgt) {
                double st, ct, sp, cp;
                sincos(th[t], &st, &ct);
                sincos(ph[t], &sp, &cp);
                x = cp * st; y = sp * st; z = ct;
            }
            if (j == 0) any = t < n_tgt;
            ux[j] = x; uy[j] = y; uz[j] = z;
            rx[j] = x * rev_scale; ry[j] = y * rev_scale; rz[j] = z * rev_scale;
        }
        cone = block_cone((float)ux[0], (float)uy[0], (float)uz[0], any, scratch);
        if (TPL > 1 && cone.cos_half > 0.f) {
            float d = 1.f;
#pragma unroll
            for (int j = 1; j < TPL; j++) {
                int t = tb * (TILE * TPL) + j * TILE + tid;
                if (t < n_tgt) d = fminf(d, (float)(ux[j] * cone.ax + uy[j] * cone.ay + uz[j] * cone.az));
            }
            cone.cos_half = fminf(cone.cos_half, block_min(d, scratch));
        }
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
        const TileMeta m = meta[base / TILE];
        const bool mask = cones_may_shadow(m.ax, m.ay, m.az, m.cos_half, cone);
        if (m.has_m) {
            if (mask) ff_tile_fp64<TPL, true, true>(tile, rx, ry, rz, ux, uy, uz, acc);
            else ff_tile_fp64<TPL, true, false>(tile, rx, ry, rz, ux, uy, uz, acc);
        } else {
            if (mask) ff_tile_fp64<TPL, false, true>(tile, rx, ry, rz, ux, uy, uz, acc);
            else ff_tile_fp64<TPL, false, false>(tile, rx, ry, rz, ux, uy, uz, acc);
        }
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

__global__ __launch_bounds__(256) void k_ff_finish(const double *__restrict__ partial, int splits,
                                                   int n_tgt, const double *__restrict__ th,
                                                   const double *__restrict__ ph, double tdir,
                                                   double pref, OutC2 out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tgt) return;
    double S[12];   // split sums; each split's twelve loads issue together
#pragma unroll
    for (int q = 0; q < 12; q++) S[q] = partial[(size_t)q * n_tgt + t];
    for (int s = 1; s < splits; s++)
#pragma unroll
        for (int q = 0; q < 12; q++) S[q] += partial[((size_t)s * 12 + q) * n_tgt + t];
    double st, ct, sp, cp;
    sincos(th[t], &st, &ct);
    sincos(ph[t], &sp, &cp);
    const double r[3] = {cp * st, sp * st, ct};
    // complex components: J_a = (S[2a], S[2a+1]), M_a = (S[6+2a], S[7+2a])
    double jdr_re = 0, jdr_im = 0, mdr_re = 0, mdr_im = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        jdr_re += S[2 * a] * r[a]; jdr_im += S[2 * a + 1] * r[a];
        mdr_re += S[6 + 2 * a] * r[a]; mdr_im += S[7 + 2 * a] * r[a];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        // (r x V)_a = r_b V_c - r_c V_b
        double rxm_re = r[b] * S[6 + 2 * c] - r[c] * S[6 + 2 * b];
        double rxm_im = r[b] * S[7 + 2 * c] - r[c] * S[7 + 2 * b];
        double rxj_re = r[b] * S[2 * c] - r[c] * S[2 * b];
        double rxj_im = r[b] * S[2 * c + 1] - r[c] * S[2 * b + 1];
        double e_re = S[2 * a] - r[a] * jdr_re + tdir * rxm_re;
        double e_im = S[2 * a + 1] - r[a] * jdr_im + tdir * rxm_im;
        double h_re = S[6 + 2 * a] - r[a] * mdr_re - tdir * rxj_re;
        double h_im = S[7 + 2 * a] - r[a] * mdr_im - tdir * rxj_im;
        // multiply by -j*pref: (x + j y)(-j) = y - j x
        out.p[a][t] = pref * e_im;
        out.p[3 + a][t] = -pref * e_re;
        out.p[6 + a][t] = pref * h_im;
        out.p[9 + a][t] = -pref * h_re;
    }
}

}  // namespace

int ff_wgs_per_cu(int precision, int tpl) {
    if (precision != PYPO_PREC_FP64 && ff_kernel() == PYPO_FF_MATRIX) return ff_mfma_wgs_per_cu(tpl);
    static const int f2 = blocks_per_cu(reinterpret_cast<const void *>(k_ff_fp64<2>));
    static const int f1 = blocks_per_cu(reinterpret_cast<const void *>(k_ff_fp64<1>));
    static const int m2 = blocks_per_cu(reinterpret_cast<const void *>(k_ff_mixed<2>));
    static const int m1 = blocks_per_cu(reinterpret_cast<const void *>(k_ff_mixed<1>));
    if (precision == PYPO_PREC_FP64) return tpl == 2 ? f2 : f1;
    return tpl == 2 ? m2 : m1;
}

int launch_ff_sum(int precision, const SumPlan &p, const void *packed, int n_src, const GridPtrs &ct,
                  int n_tgt, double rev_scale, double *partial, hipStream_t st) {
    if (precision != PYPO_PREC_FP64 && ff_kernel() == PYPO_FF_MATRIX)
        return launch_ff_mfma(p, packed, n_src, ct, n_tgt, rev_scale, partial, st);
    dim3 grid(p.tblocks, p.splits), block(TILE);
    const int rb = precision == PYPO_PREC_FP64 ? (int)sizeof(SrcRec64) : (int)sizeof(SrcRec32);
    const TileMeta *meta = (const TileMeta *)((const char *)packed + meta_offset(rb, n_src));
    if (precision == PYPO_PREC_FP64) {
        if (p.tpl == 2)
            hipLaunchKernelGGL(k_ff_fp64<2>, grid, block, 0, st, (const SrcRec64 *)packed, meta, n_src,
                               p.chunk, ct.x, ct.y, n_tgt, rev_scale, partial);
        else
            hipLaunchKernelGGL(k_ff_fp64<1>, grid, block, 0, st, (const SrcRec64 *)packed, meta, n_src,
                               p.chunk, ct.x, ct.y, n_tgt, rev_scale, partial);
    } else {
        if (p.tpl == 2)
            hipLaunchKernelGGL(k_ff_mixed<2>, grid, block, 0, st, (const SrcRec32 *)packed, meta, n_src,
                               p.chunk, ct.x, ct.y, n_tgt, rev_scale, partial);
        else
            hipLaunchKernelGGL(k_ff_mixed<1>, grid, block, 0, st, (const SrcRec32 *)packed, meta, n_src,
                               p.chunk, ct.x, ct.y, n_tgt, rev_scale, partial);
    }
    PYPO_CHECK_LAUNCH();
    return 0;
}

int launch_ff_finish(const SumPlan &p, const GridPtrs &ct, int n_tgt, double tdir, double pref,
                     const c2Bundle &b, const double *partial, hipStream_t st) {
    OutC2 o{{b.r1x, b.r1y, b.r1z, b.i1x, b.i1y, b.i1z, b.r2x, b.r2y, b.r2z, b.i2x, b.i2y, b.i2z}};
    hipLaunchKernelGGL(k_ff_finish, dim3((n_tgt + 255) / 256), dim3(256), 0, st, partial, p.splits,
                       n_tgt, ct.x, ct.y, tdir, pref, o);
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
