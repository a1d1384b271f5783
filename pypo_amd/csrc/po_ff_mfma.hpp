// po_ff_mfma.hpp -- the mixed-precision far-field sum on the matrix cores,
// shared by the single-propagation kernels (po_ff_mfma.hip) and the batched
// sweep (sweep.hip).
//
// The far field factors per target direction r into the source sums
// (farfieldAtPoint, Propagation.h:1656-1781)
//   S_q(r) = sum_s W(s, r) C_q(s),   W = exp(-j t k r_s.r),  C = w A (J, M),
// a (targets x sources) by (sources x 12) complex matrix product whose left
// factor is generated on the fly.  Each lane computes W for four sources and
// one target -- phase in revolutions, FP32 v_sin/v_cos -- and the products
// run as v_mfma_f32_16x16x32_f16 with FP32 accumulation:
//
//   A (16 targets x 32) = [cos, sin] of 16 sources, interleaved along K;
//   B (32 x 16 columns)  = per source s and output column q = 2a + part
//                          (a = Jx..Mz, part 0 = real, 1 = imaginary):
//                          cos row: Re C_a | Im C_a,  sin row: -Im C_a | Re C_a
//                          (columns 12..15 zero).
//
// FP16 has 11 significant bits, so both operands are split into hi + lo
// halves (x = hi + lo to ~22 bits) and three products are accumulated:
// hi*hi + hi*lo + lo*hi (the lo*lo term is below FP32 rounding).  B is
// scaled per 256-source tile by the power of two at or above the tile's
// largest component (TileMeta::cmax), so it lies in [-1, 1]; the FP32 tile
// sums are multiplied back and folded into FP64 after every tile, like the
// vector kernel's.  The operand split costs three VALU instructions per pair
// against the twelve FP32 FMAs the vector kernel spends on the products.
//
// Phase.  The four sources a lane takes per K-step are consecutive records,
// i.e. neighbours on the source grid.  The second of them (the leader) gets
// the full FP64 phase rev = r_s . q (q = t k r / 2 pi, FP64), v_fract_f64.
// The other three (followers) add their FP32 offset from the leader,
// d_s = (r_s - r_leader) t k / 2 pi, dotted with the FP32 unit direction:
//   f_s = f_leader + d_s . u                  (three FP32 FMAs)
// instead of three FP64 FMAs, v_fract_f64 and a conversion each; v_sin_f32
// and v_cos_f32 reduce such an argument (|f_s| <= 9 revolutions) themselves,
// to within one ulp (6e-8) of their values on fract(f_s) (measured on gfx950
// over 2^24 points in [-9, 9]), so no v_fract_f32 is spent.  The FP32
// rounding of that sum is at most a few ulps of its partial sums, which are
// bounded per workgroup by 1 + sum_i |d_i| (|a_i| + sin(theta)) (a, theta:
// the cone of the workgroup's target directions).  A 256-source tile takes
// the follower form only when that bound is <= kLocalRev for all its
// followers (one flag per tile, set when the tile is staged); otherwise
// -- a grid-row wrap to a distant point, past the end of the sources, wide
// target cones on coarse grids -- every source of the tile gets the FP64
// phase.  The error that stays is that of the FP32 direction u (2^-24
// relative) times |d|, a few 1e-7 revolutions for neighbours a few
// wavelengths apart.
//
// Workgroup: 4 waves x MB M-blocks of 16 targets.  Per tile each lane takes
// one FP32 source record from HBM (prefetched during the previous tile),
// writes its coordinates to LDS (one array per coordinate: the four lane
// groups of a wave then read distinct banks) and its B entries (16 K-steps
// x 64 lanes x 16 B, hi and lo in separate arrays: 32 KB of LDS; lane slots
// swizzled within each K-step, b_slot).  Lane layout of the
// 16x16x32 f16 MFMA (checked on gfx950): lane l holds A[l % 16][8 (l / 16) + j],
// B[8 (l / 16) + j][l % 16] (j = 0..7) and D[4 (l / 16) + i][l % 16] (i = 0..3).

#pragma once

#include "po_tiles.hpp"

namespace pypo {
namespace mfma {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int KSTEPS = TILE / 16;   // 16 sources (32 K rows: cos and sin) per MFMA

// Largest bound on the follower phase sums (revolutions): FP32 rounds such a
// sum to 2^-24 * 8 = 5e-7 revolutions at worst.
constexpr float kLocalRev = 8.f;

// 1 / (power of two at or above cmax); 1 for an all-zero tile.
__device__ __forceinline__ float tile_inv_scale(float cmax) {
    if (!(cmax > 0.f)) return 1.f;
    int e;
    (void)frexpf(cmax, &e);   // cmax = m 2^e, m in [0.5, 1)
    return ldexpf(1.f, -e);
}

// Split of two FP32 values into packed FP16 hi and lo halves (x = hi + lo to
// ~22 bits), for both operands: one packed RNE conversion for hi; c - hi and
// s - hi by v_fma_mix_f32 (hi read as FP16 halves, c and s FP32: the
// differences are exact in FP32); one packed RNE conversion for lo.
// Issue cost per wave-instruction measured on gfx950 (4 waves per SIMD):
// v_cvt_pk_f16_f32 5.4 cycles, v_fma_mix_f32 5.5, but v_fma_mix{lo,hi}_f16
// 9.9 -- so this four-instruction form (21.8) beats cvt_pk + mixlo + mixhi
// (25.2), with bit-identical lo (checked on 2^24 inputs).  In the large far
// field the same holds: the mixlo + mixhi form gave bit-identical sums in
// 99.5 ms against 91.0.
__device__ __forceinline__ void split_pair_fast(float c, float s, unsigned &hi, unsigned &lo) {
    hi = __builtin_bit_cast(unsigned, half2v{(_Float16)c, (_Float16)s});
    float dc, ds;
    asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]" : "=v"(dc) : "v"(hi), "v"(c));
    asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(ds) : "v"(hi), "v"(s));
    lo = __builtin_bit_cast(unsigned, half2v{(_Float16)dc, (_Float16)ds});
}

// Slot of lane L's B operand in a K-step row of 64 half8 (16-byte) slots:
// the low four bits of L are XORed with (ks mod 4, L / 16).  build_b's
// stores of one column come from the 16 lanes of a wave that differ in
// (ks mod 4, g): unswizzled they all hit one bank group (a 16-way
// conflict); swizzled they hit 16 different ones.  A bijection within the
// row, so a wave's 64 reads of one K-step stay one contiguous 1 KB block.
__device__ __forceinline__ int b_slot(int ks, int L) { return L ^ (((ks & 3) << 2) | (L >> 4)); }

// This lane's source record (tile slot sl) -> its B entries of every column.
// Source slot sl sits in K-step sl / 16, lane group g = (sl % 16) / 4, K rows
// 2 (sl % 4) (cos) and 2 (sl % 4) + 1 (sin) of the lanes 16 g + q.
__device__ __forceinline__ void build_b(const SrcRec32 &r, int sl, float inv, half8 (*bhi)[64],
                                        half8 (*blo)[64]) {
    const int ks = sl >> 4, g = (sl & 15) >> 2, jj = (sl & 3);
#pragma unroll
    for (int a = 0; a < 6; a++) {
        const float re = (a < 3 ? r.cj[2 * a] : r.cm[2 * (a - 3)]) * inv;
        const float im = (a < 3 ? r.cj[2 * a + 1] : r.cm[2 * (a - 3) + 1]) * inv;
        unsigned h, l;
        const int c0 = b_slot(ks, 16 * g + 2 * a), c1 = b_slot(ks, 16 * g + 2 * a + 1);
        // real column 2a: cos row Re, sin row -Im
        split_pair_fast(re, -im, h, l);
        reinterpret_cast<unsigned *>(&bhi[ks][c0])[jj] = h;
        reinterpret_cast<unsigned *>(&blo[ks][c0])[jj] = l;
        // imaginary column 2a+1: cos row Im, sin row Re
        split_pair_fast(im, re, h, l);
        reinterpret_cast<unsigned *>(&bhi[ks][c1])[jj] = h;
        reinterpret_cast<unsigned *>(&blo[ks][c1])[jj] = l;
    }
}

// Source coordinates of a tile in LDS, one array per coordinate so that the
// four lane groups of a wave (4 consecutive sources each) read distinct banks;
// d* are the follower offsets from their leader in revolutions (FP32).
struct TileXYZ {
    double x[TILE], y[TILE], z[TILE];
    float nx[TILE], ny[TILE], nz[TILE];
    float dx[TILE], dy[TILE], dz[TILE];
    int fast[4];   // per wave: 1 when all its followers may take the follower form
};

// Slot of the leader of source slot `sl`'s group of four.
__device__ __forceinline__ int leader_slot(int sl) { return (sl & ~3) | 1; }

// One K-step (16 sources) for the MB M-blocks: the pair computations and
// 3 MFMAs per M-block.  The MFMAs of M-block m are issued between the pair
// computations of M-block m + 1 (sched_barrier pins the order): the matrix
// core runs while the wave keeps issuing vector work.  Left to itself the
// compiler clusters the three dependent MFMAs and the SIMD runs them and the
// vector work in turn.  FAST: the follower form of the phase (see above).
template <int MB, bool MASK, bool FAST>
__device__ __forceinline__ void ff_kstep(const TileXYZ &tx, const half8 &b_hi, const half8 &b_lo, int s0,
                                         const double (&qx)[MB], const double (&qy)[MB], const double (&qz)[MB],
                                         const float4 *tq, int lt0, floatx4 (&acc)[MB]) {
    double sx[4], sy[4], sz[4];
    float dx[4], dy[4], dz[4], nx[4], ny[4], nz[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (!FAST || j == 1) { sx[j] = tx.x[s0 + j]; sy[j] = tx.y[s0 + j]; sz[j] = tx.z[s0 + j]; }
        if (FAST && j != 1) { dx[j] = tx.dx[s0 + j]; dy[j] = tx.dy[s0 + j]; dz[j] = tx.dz[s0 + j]; }
        if (MASK) { nx[j] = tx.nx[s0 + j]; ny[j] = tx.ny[s0 + j]; nz[j] = tx.nz[s0 + j]; }
    }
    unsigned ah[2][4], al[2][4];
    float f_lead = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
    auto load_u = [&](int m) {
        const float4 &t = tq[lt0 + 16 * m];
        ux = t.x; uy = t.y; uz = t.z;
    };
    // j-th computation of an M-block: the leader (source 1) first
    auto src_of = [](int j) { return FAST ? (j == 0 ? 1 : (j == 1 ? 0 : j)) : j; };
    auto pair = [&](int m, int j, unsigned &h, unsigned &l) {
        float f;
        if (!FAST || j == 1) {
            const double rev = __builtin_fma(sz[j], qz[m], __builtin_fma(sy[j], qy[m], sx[j] * qx[m]));
            f = (float)__builtin_amdgcn_fract(rev);
            if (FAST) f_lead = f;
        } else {
            // |f| <= kLocalRev (the bound counts f_lead): in the hardware's own reduction range
            f = __builtin_fmaf(dz[j], uz, __builtin_fmaf(dy[j], uy, __builtin_fmaf(dx[j], ux, f_lead)));
        }
        float sn, cs;
        sincos_rev_f32(f, sn, cs);
        if (MASK) {
            const float d = __builtin_fmaf(nz[j], uz, __builtin_fmaf(ny[j], uy, nx[j] * ux));
            cs = (d < 0.f) ? 0.f : cs;
            sn = (d < 0.f) ? 0.f : sn;
        }
        split_pair_fast(cs, sn, h, l);
    };
    if (FAST || MASK) load_u(0);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int js = src_of(j);
        pair(0, js, ah[0][js], al[0][js]);
    }
#pragma unroll
    for (int m = 0; m < MB; m++) {
        const int c = m & 1, n = c ^ 1;
        const half8 a_hi = __builtin_bit_cast(half8, ah[c]), a_lo = __builtin_bit_cast(half8, al[c]);
        if ((FAST || MASK) && m + 1 < MB) load_u(m + 1);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j < 3) {
                __builtin_amdgcn_sched_barrier(0);
                if (j == 0) acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi, b_hi, acc[m], 0, 0, 0);
                if (j == 1) acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_hi, b_lo, acc[m], 0, 0, 0);
                if (j == 2) acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a_lo, b_hi, acc[m], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
            }
            const int js = src_of(j);
            if (m + 1 < MB) pair(m + 1, js, ah[n][js], al[n][js]);
        }
    }
}

// One 256-source tile: 16 K-steps, all in the follower form or none (one
// loop body per form keeps the accumulators in place).
template <int MB, bool MASK, bool FAST>
__device__ __forceinline__ void ff_tile_mfma(const TileXYZ &tx, const half8 (*bhi)[64], const half8 (*blo)[64],
                                             int lane, const double (&qx)[MB], const double (&qy)[MB],
                                             const double (&qz)[MB], const float4 *tq, int lt0,
                                             floatx4 (&acc)[MB]) {
    const int kb = lane >> 4;
#pragma unroll 1
    for (int ks = 0; ks < KSTEPS; ks++) {
        const int slot = b_slot(ks, lane);
        const half8 b_hi = bhi[ks][slot], b_lo = blo[ks][slot];
        ff_kstep<MB, MASK, FAST>(tx, b_hi, b_lo, ks * 16 + kb * 4, qx, qy, qz, tq, lt0, acc);
    }
}

// The sum over sources [s_begin, s_end) (a multiple of TILE apart, except
// at the end of the sources) for the 64 MB targets from t_first, written to
// out[q][t] (q = 0..11, c2Bundle-like row order of S_J and S_M).
template <int MB>
__device__ __forceinline__ void ff_mfma_body(const SrcRec32 *__restrict__ src, const TileMeta *__restrict__ meta,
                                             int s_begin, int s_end, int t_first, const double *__restrict__ th,
                                             const double *__restrict__ ph, int n_tgt, double rev_scale,
                                             double *__restrict__ out) {
    __shared__ TileXYZ tx;
    __shared__ half8 bhi[KSTEPS][64], blo[KSTEPS][64];
    __shared__ float4 tq[64 * MB];     // unit directions of the workgroup's targets (x, y, z, 0)
    __shared__ float scratch[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row = lane & 15, kb = lane >> 4;
    const int t0 = t_first + wave * (16 * MB);   // first target of this wave

    // the columns 12..15 of B stay zero
    {
        const int ks = tid >> 4, l = b_slot(ks, 16 * ((tid >> 2) & 3) + 12 + (tid & 3));
        bhi[ks][l] = half8{};
        blo[ks][l] = half8{};
    }

    // A-operand targets of this lane: t0 + 16 m + row
    double qx[MB], qy[MB], qz[MB];
    float fx[MB], fy[MB], fz[MB];
    Cone cone;
    {
        bool any = false;
#pragma unroll
        for (int m = 0; m < MB; m++) {
            const int t = t0 + 16 * m + row;
            double x = 0, y = 0, z = 0;
            if (t < n_tgt) {
                double st, ct, sp, cp;
                sincos(th[t], &st, &ct);
                sincos(ph[t], &sp, &cp);
                x = cp * st; y = sp * st; z = ct;
            }
            fx[m] = (float)x; fy[m] = (float)y; fz[m] = (float)z;
            const int lt = wave * 16 * MB + 16 * m + row;
            tq[lt] = make_float4(fx[m], fy[m], fz[m], 0.f);
            qx[m] = x * rev_scale; qy[m] = y * rev_scale; qz[m] = z * rev_scale;
            if (m == 0) any = t < n_tgt;
        }
        cone = block_cone(fx[0], fy[0], fz[0], any, scratch);
        if (MB > 1 && cone.cos_half > 0.f) {
            float d = 1.f;
#pragma unroll
            for (int m = 1; m < MB; m++)
                if (t0 + 16 * m + row < n_tgt) d = fminf(d, fx[m] * cone.ax + fy[m] * cone.ay + fz[m] * cone.az);
            cone.cos_half = fminf(cone.cos_half, block_min(d, scratch));
        }
    }

    floatx4 acc[MB];
    double tot[MB][4];
#pragma unroll
    for (int m = 0; m < MB; m++) {
        acc[m] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < 4; i++) tot[m][i] = 0.0;
    }

    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<6> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        const TileMeta tm = meta[base / TILE];
        const float inv = tile_inv_scale(tm.cmax);
        {
            const SrcRec32 r = __builtin_bit_cast(SrcRec32, ld.buf);   // this lane's record
            // offset from the group's leader (same wave), in revolutions
            const int ls = leader_slot(lane);
            const float ddx = (float)((r.x - __shfl(r.x, ls, 64)) * rev_scale);
            const float ddy = (float)((r.y - __shfl(r.y, ls, 64)) * rev_scale);
            const float ddz = (float)((r.z - __shfl(r.z, ls, 64)) * rev_scale);
            // per-component bound on |u| over the workgroup's targets: |a_i| + sin(theta)
            float ubx = 1.f, uby = 1.f, ubz = 1.f;
            if (cone.cos_half > 0.f) {
                const float sn = sqrtf(fmaxf(0.f, 1.f - cone.cos_half * cone.cos_half)) + 1e-3f;
                ubx = fminf(1.f, fabsf(cone.ax) + sn);
                uby = fminf(1.f, fabsf(cone.ay) + sn);
                ubz = fminf(1.f, fabsf(cone.az) + sn);
            }
            const float bound = 1.f + fabsf(ddx) * ubx + fabsf(ddy) * uby + fabsf(ddz) * ubz;
            const bool ok = (lane == ls) || (bound <= kLocalRev);   // NaN: not ok
            const unsigned long long all_ok = __ballot(ok);
            __syncthreads();
            tx.x[tid] = r.x; tx.y[tid] = r.y; tx.z[tid] = r.z;
            tx.nx[tid] = r.nx; tx.ny[tid] = r.ny; tx.nz[tid] = r.nz;
            tx.dx[tid] = ddx; tx.dy[tid] = ddy; tx.dz[tid] = ddz;
            if (lane == 0) tx.fast[wave] = all_ok == ~0ull;
            build_b(r, tid, inv, bhi, blo);
            __syncthreads();
        }
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
        const bool fast = __builtin_amdgcn_readfirstlane(tx.fast[0] & tx.fast[1] & tx.fast[2] & tx.fast[3]);
        const int lt0 = wave * 16 * MB + row;
        if (cones_may_shadow(tm.ax, tm.ay, tm.az, tm.cos_half, cone)) {
            if (fast) ff_tile_mfma<MB, true, true>(tx, bhi, blo, lane, qx, qy, qz, tq, lt0, acc);
            else ff_tile_mfma<MB, true, false>(tx, bhi, blo, lane, qx, qy, qz, tq, lt0, acc);
        } else {
            if (fast) ff_tile_mfma<MB, false, true>(tx, bhi, blo, lane, qx, qy, qz, tq, lt0, acc);
            else ff_tile_mfma<MB, false, false>(tx, bhi, blo, lane, qx, qy, qz, tq, lt0, acc);
        }
        const double scale = 1.0 / (double)inv;   // a power of two: exact
#pragma unroll
        for (int m = 0; m < MB; m++) {
#pragma unroll
            for (int i = 0; i < 4; i++) tot[m][i] = __builtin_fma((double)acc[m][i], scale, tot[m][i]);
            acc[m] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
    }
    // D layout: this lane holds column `row` of targets t0 + 16 m + 4 kb + i
    if (row < 12) {
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int t = t0 + 16 * m + 4 * kb + i;
                if (t < n_tgt) out[(size_t)row * n_tgt + t] = tot[m][i];
            }
    }
}

}  // namespace mfma
}  // namespace pypo
