// po_kernels.hip -- the physical-optics sums on gfx950.
//
// Every PO mode is a direct O(N_src x N_tgt) sum (src/CPU/Propagation.h):
//   near field  fieldAtPoint        977-1144  (JM, EH, JMEH, EHP)
//   far field   farfieldAtPoint     1656-1781 (FF)
//   scalar      fieldScalarAtPoint  1279-1331
// Layout of the work on the GPU:
//   * pack: one pass over the sources writes fixed-size records (position
//     FP64, normal, w*A*currents) so the sum streams 96 B per source;
//   * sum: a workgroup of 256 lanes owns 256*TPL targets (TPL targets per
//     lane, kept in registers) and walks a range of sources in 256-source
//     tiles staged through LDS (broadcast reads: every lane reads the same
//     record).  The next tile's global loads are issued before the current
//     tile is consumed.  Mixed precision keeps FP32 partial sums per tile
//     and folds them into FP64 accumulators after each tile, which bounds
//     the FP32 rounding walk to 256 terms;
//   * source splits (grid.y) give enough workgroups when the target grid is
//     small; each split writes its own FP64 partial and a finish kernel
//     reduces the splits in a fixed order (deterministic, no atomics) and
//     applies the per-target epilogue of the mode in FP64.
// Phases are computed in revolutions in FP64 (k/2pi * distance or
// k/2pi * r'.r_hat), reduced by subtracting the nearest integer, and only
// the reduced fraction goes to FP32 sin/cos (hardware v_sin/v_cos take
// revolutions).  The all-FP64 variants share the structure and evaluate
// sin/cos with FP64 sincospi on the same reduced fraction.

#include "common.hpp"

namespace pypo {
namespace {

constexpr int TILE = 256;

// ---------------------------------------------------------------------------
// packing
// ---------------------------------------------------------------------------

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
        r.c[2 * a] = (float)(wa * cur.p[a][i]);           // J re
        r.c[2 * a + 1] = (float)(wa * cur.p[3 + a][i]);   // J im
        r.c[6 + 2 * a] = (float)(wa * cur.p[6 + a][i]);   // M re
        r.c[7 + 2 * a] = (float)(wa * cur.p[9 + a][i]);   // M im
    }
    r.pad1[0] = 0.f; r.pad1[1] = 0.f;
    out[i] = r;
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
    out[i] = r;
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
    out[i] = r;
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
    out[i] = r;
}

// ---------------------------------------------------------------------------
// LDS tile staging shared by all sum kernels: a tile of TILE records of
// REC_U4 16-byte words, loaded one record per lane.  Past the end of the
// split the lane writes a null record instead (all weights zero, position
// 1e7 mm away from anything) so that the inner loop always runs TILE
// iterations and a null record can never produce a zero distance.
// ---------------------------------------------------------------------------

template <int REC_U4>
struct TileLoader {
    uint4 buf[REC_U4];
    __device__ __forceinline__ void load(const uint4 *__restrict__ g, int s, int s_end) {
        if (s < s_end) {
#pragma unroll
            for (int k = 0; k < REC_U4; k++) buf[k] = g[(size_t)s * REC_U4 + k];
        } else {
#pragma unroll
            for (int k = 0; k < REC_U4; k++) buf[k] = make_uint4(0u, 0u, 0u, 0u);
            // every record starts with double x, y, z
            const unsigned long long far = (unsigned long long)__double_as_longlong(1.0e7);
            const unsigned lo = (unsigned)(far & 0xffffffffull), hi = (unsigned)(far >> 32);
            buf[0] = make_uint4(lo, hi, lo, hi);
            buf[1].x = lo;
            buf[1].y = hi;
        }
    }
    __device__ __forceinline__ void store(uint4 *lds, int tid) {
#pragma unroll
        for (int k = 0; k < REC_U4; k++) lds[tid * REC_U4 + k] = buf[k];
    }
};

// ---------------------------------------------------------------------------
// far field -- farfieldAtPoint (Propagation.h:1656-1781) factored per
// target: E = -j*pref*[SJ - (SJ.r)r + t r x SM], H = -j*pref*[SM - (SM.r)r
// - t r x SJ] with SJ = sum w A J e^{-j t k r'.r}, SM likewise, over the
// sources that pass the shadow test n.r >= 0.  The sum therefore costs six
// complex multiply-adds per pair; the projection is in the finish kernel.
// Partial layout: partial[(split*12 + q)*n_tgt + target], q as SrcRec32.c.
// ---------------------------------------------------------------------------

template <int TPL>
__global__ __launch_bounds__(256) void k_ff_mixed(const SrcRec32 *__restrict__ src, int n_src,
                                                  int chunk, const double *__restrict__ th,
                                                  const double *__restrict__ ph, int n_tgt,
                                                  double rev_scale, double *__restrict__ partial) {
    __shared__ SrcRec32 tile[TILE];
    const int tid = threadIdx.x;
    double rx[TPL], ry[TPL], rz[TPL];
    float fx[TPL], fy[TPL], fz[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        double x = 0, y = 0, z = 0;
        if (t < n_tgt) {
            double st, ct, sp, cp;
            sincos(th[t], &st, &ct);
            sincos(ph[t], &sp, &cp);
            x = cp * st; y = sp * st; z = ct;
        }
        fx[j] = (float)x; fy[j] = (float)y; fz[j] = (float)z;
        rx[j] = x * rev_scale; ry[j] = y * rev_scale; rz[j] = z * rev_scale;
    }
    float acc[TPL][12];
    double tot[TPL][12];
#pragma unroll
    for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int q = 0; q < 12; q++) { acc[j][q] = 0.f; tot[j][q] = 0.0; }

    const int s_begin = blockIdx.y * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<6> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        __syncthreads();
        ld.store(reinterpret_cast<uint4 *>(tile), tid);
        __syncthreads();
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
#pragma unroll 2
        for (int s = 0; s < TILE; s++) {
            const SrcRec32 &r = tile[s];
            const double x = r.x, y = r.y, z = r.z;
            const float nx = r.nx, ny = r.ny, nz = r.nz;
            float c[12];
#pragma unroll
            for (int q = 0; q < 12; q++) c[q] = r.c[q];
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                double rev = x * rx[j] + y * ry[j] + z * rz[j];
                float f = reduce_rev(rev);
                float sn, cs;
                sincos_rev_f32(f, sn, cs);
                float m = nx * fx[j] + ny * fy[j] + nz * fz[j];
                cs = (m < 0.f) ? 0.f : cs;
                sn = (m < 0.f) ? 0.f : sn;
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    acc[j][2 * q] += c[2 * q] * cs - c[2 * q + 1] * sn;
                    acc[j][2 * q + 1] += c[2 * q] * sn + c[2 * q + 1] * cs;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < TPL; j++)
#pragma unroll
            for (int q = 0; q < 12; q++) { tot[j][q] += (double)acc[j][q]; acc[j][q] = 0.f; }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
#pragma unroll
            for (int q = 0; q < 12; q++)
                partial[((size_t)blockIdx.y * 12 + q) * n_tgt + t] = tot[j][q];
        }
    }
}

template <int TPL>
__global__ __launch_bounds__(256) void k_ff_fp64(const SrcRec64 *__restrict__ src, int n_src,
                                                 int chunk, const double *__restrict__ th,
                                                 const double *__restrict__ ph, int n_tgt,
                                                 double rev_scale, double *__restrict__ partial) {
    __shared__ SrcRec64 tile[TILE];
    const int tid = threadIdx.x;
    double ux[TPL], uy[TPL], uz[TPL];
    double rx[TPL], ry[TPL], rz[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        double x = 0, y = 0, z = 0;
        if (t < n_tgt) {
            double st, ct, sp, cp;
            sincos(th[t], &st, &ct);
            sincos(ph[t], &sp, &cp);
            x = cp * st; y = sp * st; z = ct;
        }
        ux[j] = x; uy[j] = y; uz[j] = z;
        rx[j] = x * rev_scale; ry[j] = y * rev_scale; rz[j] = z * rev_scale;
    }
    double acc[TPL][12];
#pragma unroll
    for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int q = 0; q < 12; q++) acc[j][q] = 0.0;

    const int s_begin = blockIdx.y * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<9> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        __syncthreads();
        ld.store(reinterpret_cast<uint4 *>(tile), tid);
        __syncthreads();
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
#pragma unroll 1
        for (int s = 0; s < TILE; s++) {
            const SrcRec64 &r = tile[s];
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                double rev = r.x * rx[j] + r.y * ry[j] + r.z * rz[j];
                double f = rev - __builtin_rint(rev);
                double sn, cs;
                sincospi(2.0 * f, &sn, &cs);
                double m = r.nx * ux[j] + r.ny * uy[j] + r.nz * uz[j];
                cs = (m < 0.0) ? 0.0 : cs;
                sn = (m < 0.0) ? 0.0 : sn;
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    acc[j][2 * q] += r.c[2 * q] * cs - r.c[2 * q + 1] * sn;
                    acc[j][2 * q + 1] += r.c[2 * q] * sn + r.c[2 * q + 1] * cs;
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
#pragma unroll
            for (int q = 0; q < 12; q++)
                partial[((size_t)blockIdx.y * 12 + q) * n_tgt + t] = acc[j][q];
        }
    }
}

struct OutC2 { double *p[12]; };                 // r1x r1y r1z i1x i1y i1z r2x .. i2z
struct OutC4 { double *p[24]; };
struct OutC2R { double *p[15]; };

__global__ __launch_bounds__(256) void k_ff_finish(const double *__restrict__ partial, int splits,
                                                   int n_tgt, const double *__restrict__ th,
                                                   const double *__restrict__ ph, double tdir,
                                                   double pref, OutC2 out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tgt) return;
    double S[12];
#pragma unroll
    for (int q = 0; q < 12; q++) {
        double v = 0.0;
        for (int s = 0; s < splits; s++) v += partial[((size_t)s * 12 + q) * n_tgt + t];
        S[q] = v;
    }
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

// ---------------------------------------------------------------------------
// near field -- fieldAtPoint (Propagation.h:977-1144).  Per pair:
//   R = |p - r'|, R_hat, q = 1/(kR),
//   s1 = t q^2 + j(q^3 - q), s2 = -3t q^2 + j(q - 3q^3), s3 = -q^2 + j t q
//   (the reference's 1084-1090 with t^2 = 1), G = e^{j t k R} (prefactor
//   applied in the finish kernel, w*A folded into the record),
//   E += G[J s1 + (J.R)R s2 + (M x R) s3], H += G[M s1 + (M.R)R s2 - (J x R) s3]
// skipped when n.R_hat < 0.  Mixed precision: FP64 difference vector,
// distance (FP32 rsqrt refined by one FP64 Newton step) and phase; the
// rest in FP32.  Partial layout as the far field: q = (E re/im xyz
// interleaved, H ...), i.e. E_a = (S[2a], S[2a+1]), H_a = (S[6+2a], S[7+2a]).
// ---------------------------------------------------------------------------

__device__ __forceinline__ void cmul(float ar, float ai, float br, float bi, float &cr, float &ci) {
    cr = ar * br - ai * bi;
    ci = ar * bi + ai * br;
}

template <int TPL>
__global__ __launch_bounds__(256) void k_near_mixed(const SrcRec32 *__restrict__ src, int n_src,
                                                    int chunk, const double *__restrict__ tx,
                                                    const double *__restrict__ ty,
                                                    const double *__restrict__ tz, int n_tgt,
                                                    double k, double tdir,
                                                    double *__restrict__ partial) {
    __shared__ SrcRec32 tile[TILE];
    const int tid = threadIdx.x;
    const double rev_scale = tdir * k / kTwoPi;
    const float kinv = (float)(1.0 / k);
    const float tf = (float)tdir;
    double px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
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

    const int s_begin = blockIdx.y * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<6> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        __syncthreads();
        ld.store(reinterpret_cast<uint4 *>(tile), tid);
        __syncthreads();
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
#pragma unroll 1
        for (int s = 0; s < TILE; s++) {
            const SrcRec32 &r = tile[s];
            float c[12];
#pragma unroll
            for (int q = 0; q < 12; q++) c[q] = r.c[q];
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                double dx = px[j] - r.x, dy = py[j] - r.y, dz = pz[j] - r.z;
                double R2 = dx * dx + dy * dy + dz * dz;
                float y0 = __builtin_amdgcn_rsqf((float)R2);
                double yd = (double)y0;
                double tR = R2 * yd;
                double e = __builtin_fma(-tR, yd, 1.0);
                double R = __builtin_fma(0.5 * tR, e, tR);
                float f = reduce_rev(R * rev_scale);
                float sn, cs;
                sincos_rev_f32(f, sn, cs);
                float Rinv = __builtin_fmaf(0.5f * y0, (float)e, y0);
                float hx = (float)dx * Rinv, hy = (float)dy * Rinv, hz = (float)dz * Rinv;
                float m = r.nx * hx + r.ny * hy + r.nz * hz;
                cs = (m < 0.f) ? 0.f : cs;
                sn = (m < 0.f) ? 0.f : sn;
                float q1 = Rinv * kinv;
                float q2 = q1 * q1, q3 = q2 * q1;
                // s_i * G
                float g1r, g1i, g2r, g2i, g3r, g3i;
                cmul(tf * q2, q3 - q1, cs, sn, g1r, g1i);
                cmul(-3.f * tf * q2, q1 - 3.f * q3, cs, sn, g2r, g2i);
                cmul(-q2, tf * q1, cs, sn, g3r, g3i);
                const float h[3] = {hx, hy, hz};
                float jdr_r = c[0] * hx + c[2] * hy + c[4] * hz;
                float jdr_i = c[1] * hx + c[3] * hy + c[5] * hz;
                float mdr_r = c[6] * hx + c[8] * hy + c[10] * hz;
                float mdr_i = c[7] * hx + c[9] * hy + c[11] * hz;
                float ar, ai, br, bi;
                cmul(jdr_r, jdr_i, g2r, g2i, ar, ai);
                cmul(mdr_r, mdr_i, g2r, g2i, br, bi);
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const int b = (a + 1) % 3, d = (a + 2) % 3;
                    // (V x R)_a = V_b R_d - V_d R_b
                    float mxr_r = c[6 + 2 * b] * h[d] - c[6 + 2 * d] * h[b];
                    float mxr_i = c[7 + 2 * b] * h[d] - c[7 + 2 * d] * h[b];
                    float jxr_r = c[2 * b] * h[d] - c[2 * d] * h[b];
                    float jxr_i = c[2 * b + 1] * h[d] - c[2 * d + 1] * h[b];
                    acc[j][2 * a] += c[2 * a] * g1r - c[2 * a + 1] * g1i + h[a] * ar +
                                     mxr_r * g3r - mxr_i * g3i;
                    acc[j][2 * a + 1] += c[2 * a] * g1i + c[2 * a + 1] * g1r + h[a] * ai +
                                         mxr_r * g3i + mxr_i * g3r;
                    acc[j][6 + 2 * a] += c[6 + 2 * a] * g1r - c[7 + 2 * a] * g1i + h[a] * br -
                                         (jxr_r * g3r - jxr_i * g3i);
                    acc[j][7 + 2 * a] += c[6 + 2 * a] * g1i + c[7 + 2 * a] * g1r + h[a] * bi -
                                         (jxr_r * g3i + jxr_i * g3r);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < TPL; j++)
#pragma unroll
            for (int q = 0; q < 12; q++) { tot[j][q] += (double)acc[j][q]; acc[j][q] = 0.f; }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
#pragma unroll
            for (int q = 0; q < 12; q++)
                partial[((size_t)blockIdx.y * 12 + q) * n_tgt + t] = tot[j][q];
        }
    }
}

template <int TPL>
__global__ __launch_bounds__(256) void k_near_fp64(const SrcRec64 *__restrict__ src, int n_src,
                                                   int chunk, const double *__restrict__ tx,
                                                   const double *__restrict__ ty,
                                                   const double *__restrict__ tz, int n_tgt,
                                                   double k, double tdir,
                                                   double *__restrict__ partial) {
    __shared__ SrcRec64 tile[TILE];
    const int tid = threadIdx.x;
    const double rev_scale = tdir * k / kTwoPi;
    double px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        px[j] = (t < n_tgt) ? tx[t] : 1.0e6;
        py[j] = (t < n_tgt) ? ty[t] : 1.0e6;
        pz[j] = (t < n_tgt) ? tz[t] : 1.0e6;
    }
    double acc[TPL][12];
#pragma unroll
    for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int q = 0; q < 12; q++) acc[j][q] = 0.0;

    const int s_begin = blockIdx.y * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<9> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        __syncthreads();
        ld.store(reinterpret_cast<uint4 *>(tile), tid);
        __syncthreads();
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
#pragma unroll 1
        for (int s = 0; s < TILE; s++) {
            const SrcRec64 &r = tile[s];
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                double dx = px[j] - r.x, dy = py[j] - r.y, dz = pz[j] - r.z;
                double R = sqrt(dx * dx + dy * dy + dz * dz);
                double Rinv = 1.0 / R;
                double rev = R * rev_scale;
                double f = rev - __builtin_rint(rev);
                double sn, cs;
                sincospi(2.0 * f, &sn, &cs);
                double hx = dx * Rinv, hy = dy * Rinv, hz = dz * Rinv;
                double m = r.nx * hx + r.ny * hy + r.nz * hz;
                cs = (m < 0.0) ? 0.0 : cs;
                sn = (m < 0.0) ? 0.0 : sn;
                double q1 = Rinv / k;
                double q2 = q1 * q1, q3 = q2 * q1;
                double s1r = tdir * q2, s1i = q3 - q1;
                double s2r = -3.0 * tdir * q2, s2i = q1 - 3.0 * q3;
                double s3r = -q2, s3i = tdir * q1;
                double g1r = s1r * cs - s1i * sn, g1i = s1r * sn + s1i * cs;
                double g2r = s2r * cs - s2i * sn, g2i = s2r * sn + s2i * cs;
                double g3r = s3r * cs - s3i * sn, g3i = s3r * sn + s3i * cs;
                const double h[3] = {hx, hy, hz};
                const double *c = r.c;
                double jdr_r = c[0] * hx + c[2] * hy + c[4] * hz;
                double jdr_i = c[1] * hx + c[3] * hy + c[5] * hz;
                double mdr_r = c[6] * hx + c[8] * hy + c[10] * hz;
                double mdr_i = c[7] * hx + c[9] * hy + c[11] * hz;
                double ar = jdr_r * g2r - jdr_i * g2i, ai = jdr_r * g2i + jdr_i * g2r;
                double br = mdr_r * g2r - mdr_i * g2i, bi = mdr_r * g2i + mdr_i * g2r;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const int b = (a + 1) % 3, d = (a + 2) % 3;
                    double mxr_r = c[6 + 2 * b] * h[d] - c[6 + 2 * d] * h[b];
                    double mxr_i = c[7 + 2 * b] * h[d] - c[7 + 2 * d] * h[b];
                    double jxr_r = c[2 * b] * h[d] - c[2 * d] * h[b];
                    double jxr_i = c[2 * b + 1] * h[d] - c[2 * d + 1] * h[b];
                    acc[j][2 * a] += c[2 * a] * g1r - c[2 * a + 1] * g1i + h[a] * ar +
                                     mxr_r * g3r - mxr_i * g3i;
                    acc[j][2 * a + 1] += c[2 * a] * g1i + c[2 * a + 1] * g1r + h[a] * ai +
                                         mxr_r * g3i + mxr_i * g3r;
                    acc[j][6 + 2 * a] += c[6 + 2 * a] * g1r - c[7 + 2 * a] * g1i + h[a] * br -
                                         (jxr_r * g3r - jxr_i * g3i);
                    acc[j][7 + 2 * a] += c[6 + 2 * a] * g1i + c[7 + 2 * a] * g1r + h[a] * bi -
                                         (jxr_r * g3i + jxr_i * g3r);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
#pragma unroll
            for (int q = 0; q < 12; q++)
                partial[((size_t)blockIdx.y * 12 + q) * n_tgt + t] = acc[j][q];
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
#pragma unroll
    for (int q = 0; q < 12; q++) {
        double v = 0.0;
        for (int s = 0; s < splits; s++) v += partial[((size_t)s * 12 + q) * n_tgt + t];
        S[q] = pref * v;
    }
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

// ---------------------------------------------------------------------------
// scalar -- fieldScalarAtPoint (Propagation.h:1279-1331):
//   u(p) = sum -k^2/(4 pi r) w A u_s e^{-j t k r}, no shadowing.
// The -k^2/(4 pi) factor is applied by the finish kernel.  Partial layout:
// partial[(split*2 + {re,im})*n_tgt + t].
// ---------------------------------------------------------------------------

template <int TPL>
__global__ __launch_bounds__(256) void k_scalar_mixed(const SrcRecS32 *__restrict__ src, int n_src,
                                                      int chunk, const double *__restrict__ tx,
                                                      const double *__restrict__ ty,
                                                      const double *__restrict__ tz, int n_tgt,
                                                      double rev_scale, double *__restrict__ partial) {
    __shared__ SrcRecS32 tile[TILE];
    const int tid = threadIdx.x;
    double px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        px[j] = (t < n_tgt) ? tx[t] : 1.0e6;
        py[j] = (t < n_tgt) ? ty[t] : 1.0e6;
        pz[j] = (t < n_tgt) ? tz[t] : 1.0e6;
    }
    float ar[TPL], ai[TPL];
    double tr[TPL], ti[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) { ar[j] = ai[j] = 0.f; tr[j] = ti[j] = 0.0; }
    const int s_begin = blockIdx.y * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<2> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        __syncthreads();
        ld.store(reinterpret_cast<uint4 *>(tile), tid);
        __syncthreads();
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
#pragma unroll 2
        for (int s = 0; s < TILE; s++) {
            const SrcRecS32 &r = tile[s];
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                double dx = px[j] - r.x, dy = py[j] - r.y, dz = pz[j] - r.z;
                double R2 = dx * dx + dy * dy + dz * dz;
                float y0 = __builtin_amdgcn_rsqf((float)R2);
                double yd = (double)y0;
                double tR = R2 * yd;
                double e = __builtin_fma(-tR, yd, 1.0);
                double R = __builtin_fma(0.5 * tR, e, tR);
                float f = reduce_rev(R * rev_scale);
                float sn, cs;
                sincos_rev_f32(f, sn, cs);
                float Rinv = __builtin_fmaf(0.5f * y0, (float)e, y0);
                float ur = r.ur * Rinv, ui = r.ui * Rinv;
                ar[j] += ur * cs - ui * sn;
                ai[j] += ur * sn + ui * cs;
            }
        }
#pragma unroll
        for (int j = 0; j < TPL; j++) { tr[j] += ar[j]; ti[j] += ai[j]; ar[j] = ai[j] = 0.f; }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
            partial[((size_t)blockIdx.y * 2) * n_tgt + t] = tr[j];
            partial[((size_t)blockIdx.y * 2 + 1) * n_tgt + t] = ti[j];
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
    double px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        px[j] = (t < n_tgt) ? tx[t] : 1.0e6;
        py[j] = (t < n_tgt) ? ty[t] : 1.0e6;
        pz[j] = (t < n_tgt) ? tz[t] : 1.0e6;
    }
    double ar[TPL], ai[TPL];
#pragma unroll
    for (int j = 0; j < TPL; j++) { ar[j] = ai[j] = 0.0; }
    const int s_begin = blockIdx.y * chunk;
    const int s_end = min(n_src, s_begin + chunk);
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
    TileLoader<3> ld;
    ld.load(g, s_begin + tid, s_end);
    for (int base = s_begin; base < s_end; base += TILE) {
        __syncthreads();
        ld.store(reinterpret_cast<uint4 *>(tile), tid);
        __syncthreads();
        if (base + TILE < s_end) ld.load(g, base + TILE + tid, s_end);
#pragma unroll 1
        for (int s = 0; s < TILE; s++) {
            const SrcRecS64 &r = tile[s];
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                double dx = px[j] - r.x, dy = py[j] - r.y, dz = pz[j] - r.z;
                double R = sqrt(dx * dx + dy * dy + dz * dz);
                double rev = R * rev_scale;
                double f = rev - __builtin_rint(rev);
                double sn, cs;
                sincospi(2.0 * f, &sn, &cs);
                double Rinv = 1.0 / R;
                double ur = r.ur * Rinv, ui = r.ui * Rinv;
                ar[j] += ur * cs - ui * sn;
                ai[j] += ur * sn + ui * cs;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
        int t = blockIdx.x * (TILE * TPL) + j * TILE + tid;
        if (t < n_tgt) {
            partial[((size_t)blockIdx.y * 2) * n_tgt + t] = ar[j];
            partial[((size_t)blockIdx.y * 2 + 1) * n_tgt + t] = ai[j];
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

// ---------------------------------------------------------------------------

int record_bytes(int mode, int precision) {
    if (mode == PYPO_MODE_SCALAR) return precision == PYPO_PREC_FP64 ? 48 : 32;
    return precision == PYPO_PREC_FP64 ? 144 : 96;
}

struct Timer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool on;
    hipStream_t st;
    Timer(bool enable, hipStream_t s) : on(enable), st(s) {
        if (on) {
            if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) on = false;
            else (void)hipEventRecord(e0, st);
        }
    }
    void stop() {
        if (!on) return;
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) == hipSuccess) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            record_kernel_ms(ms);
        }
    }
    ~Timer() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

}  // namespace

size_t packed_bytes(int mode, int precision, int n_src) {
    return (size_t)record_bytes(mode, precision) * (size_t)n_src;
}

SumPlan plan_sum(int mode, int precision, int n_src, int n_tgt) {
    (void)precision;
    SumPlan p;
    p.tpl = 1;
    p.tblocks = (n_tgt + TILE * p.tpl - 1) / (TILE * p.tpl);
    // Enough workgroups to fill 256 CUs several times over; never split a
    // source range below 8 tiles.
    const int want_wgs = 2048;
    int splits = (want_wgs + p.tblocks - 1) / p.tblocks;
    int max_splits = (n_src + 8 * TILE - 1) / (8 * TILE);
    if (splits > max_splits) splits = max_splits;
    if (splits < 1) splits = 1;
    int chunk = (n_src + splits - 1) / splits;
    chunk = (chunk + TILE - 1) / TILE * TILE;
    p.chunk = chunk;
    p.splits = (n_src + chunk - 1) / chunk;
    if (p.splits < 1) p.splits = 1;
    (void)mode;
    return p;
}

size_t partial_bytes(int mode, const SumPlan &p, int n_tgt) {
    int comps = (mode == PYPO_MODE_SCALAR) ? 2 : 12;
    return (size_t)p.splits * comps * (size_t)n_tgt * sizeof(double);
}

int launch_pack_sources(int precision, const GridPtrs &cs, int n_src, const Bundle12 &cur,
                        int gmode, int ncx, int ncy, void *packed, hipStream_t st) {
    PYPO_REQUIRE(n_src > 0 && (long long)ncx * ncy == n_src, ERR_ARG,
                 "source n_cells do not match the source size");
    int blocks = (n_src + 255) / 256;
    if (precision == PYPO_PREC_FP64)
        hipLaunchKernelGGL(k_pack64, dim3(blocks), dim3(256), 0, st, cs, n_src, cur, gmode, ncx, ncy,
                           (SrcRec64 *)packed);
    else
        hipLaunchKernelGGL(k_pack32, dim3(blocks), dim3(256), 0, st, cs, n_src, cur, gmode, ncx, ncy,
                           (SrcRec32 *)packed);
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

int launch_po(int mode, int precision, const void *packed, int n_src, const GridPtrs &ct,
              int n_tgt, double k, double epsilon, double t_dir, void *res, double *partial_ws,
              size_t ws_bytes, hipStream_t st, bool time_it) {
    PYPO_REQUIRE(n_src > 0 && n_tgt > 0, ERR_ARG, "empty source or target grid");
    PYPO_REQUIRE(mode >= PYPO_MODE_JM && mode <= PYPO_MODE_SCALAR, ERR_ARG, "unknown PO mode");
    SumPlan p = plan_sum(mode, precision, n_src, n_tgt);
    PYPO_REQUIRE(partial_bytes(mode, p, n_tgt) <= ws_bytes, ERR_ARG, "partial workspace too small");
    PYPO_REQUIRE((long long)p.chunk * p.splits >= n_src && p.chunk % TILE == 0, ERR_ARG,
                 "internal: bad source split plan");
    const PhysConsts pc = phys_consts(k, epsilon);
    const bool fp64 = (precision == PYPO_PREC_FP64);
    dim3 grid(p.tblocks, p.splits), block(256);
    const int fin_blocks = (n_tgt + 255) / 256;

    Timer tm(time_it, st);
    if (mode == PYPO_MODE_FF) {
        const double rev_scale = -t_dir * k / kTwoPi;   // e^{-j t k r'.r}
        if (fp64)
            hipLaunchKernelGGL(k_ff_fp64<1>, grid, block, 0, st, (const SrcRec64 *)packed, n_src,
                               p.chunk, ct.x, ct.y, n_tgt, rev_scale, partial_ws);
        else
            hipLaunchKernelGGL(k_ff_mixed<1>, grid, block, 0, st, (const SrcRec32 *)packed, n_src,
                               p.chunk, ct.x, ct.y, n_tgt, rev_scale, partial_ws);
    } else if (mode == PYPO_MODE_SCALAR) {
        const double rev_scale = -t_dir * k / kTwoPi;   // e^{-j t k r}
        if (fp64)
            hipLaunchKernelGGL(k_scalar_fp64<1>, grid, block, 0, st, (const SrcRecS64 *)packed,
                               n_src, p.chunk, ct.x, ct.y, ct.z, n_tgt, rev_scale, partial_ws);
        else
            hipLaunchKernelGGL(k_scalar_mixed<1>, grid, block, 0, st, (const SrcRecS32 *)packed,
                               n_src, p.chunk, ct.x, ct.y, ct.z, n_tgt, rev_scale, partial_ws);
    } else {
        if (fp64)
            hipLaunchKernelGGL(k_near_fp64<1>, grid, block, 0, st, (const SrcRec64 *)packed, n_src,
                               p.chunk, ct.x, ct.y, ct.z, n_tgt, k, t_dir, partial_ws);
        else
            hipLaunchKernelGGL(k_near_mixed<1>, grid, block, 0, st, (const SrcRec32 *)packed, n_src,
                               p.chunk, ct.x, ct.y, ct.z, n_tgt, k, t_dir, partial_ws);
    }
    PYPO_CHECK_LAUNCH();
    tm.stop();

    if (mode == PYPO_MODE_FF) {
        const c2Bundle *b = (const c2Bundle *)res;
        OutC2 o{{b->r1x, b->r1y, b->r1z, b->i1x, b->i1y, b->i1z,
                 b->r2x, b->r2y, b->r2z, b->i2x, b->i2y, b->i2z}};
        hipLaunchKernelGGL(k_ff_finish, dim3(fin_blocks), block, 0, st, partial_ws, p.splits, n_tgt,
                           ct.x, ct.y, t_dir, pc.prefactor, o);
    } else if (mode == PYPO_MODE_SCALAR) {
        const arrC1 *b = (const arrC1 *)res;
        const double scale = -k * k / (4 * pc.pi_f);
        hipLaunchKernelGGL(k_scalar_finish, dim3(fin_blocks), block, 0, st, partial_ws, p.splits,
                           n_tgt, scale, b->x, b->y);
    } else {
        OutC4 o{};
        if (mode == PYPO_MODE_JMEH) {
            const c4Bundle *b = (const c4Bundle *)res;
            double *v[24] = {b->r1x, b->r1y, b->r1z, b->i1x, b->i1y, b->i1z,
                             b->r2x, b->r2y, b->r2z, b->i2x, b->i2y, b->i2z,
                             b->r3x, b->r3y, b->r3z, b->i3x, b->i3y, b->i3z,
                             b->r4x, b->r4y, b->r4z, b->i4x, b->i4y, b->i4z};
            for (int q = 0; q < 24; q++) o.p[q] = v[q];
        } else if (mode == PYPO_MODE_EHP) {
            const c2rBundle *b = (const c2rBundle *)res;
            double *v[15] = {b->r1x, b->r1y, b->r1z, b->i1x, b->i1y, b->i1z,
                             b->r2x, b->r2y, b->r2z, b->i2x, b->i2y, b->i2z,
                             b->r3x, b->r3y, b->r3z};
            for (int q = 0; q < 15; q++) o.p[q] = v[q];
        } else {
            const c2Bundle *b = (const c2Bundle *)res;
            double *v[12] = {b->r1x, b->r1y, b->r1z, b->i1x, b->i1y, b->i1z,
                             b->r2x, b->r2y, b->r2z, b->i2x, b->i2y, b->i2z};
            for (int q = 0; q < 12; q++) o.p[q] = v[q];
        }
        switch (mode) {
            case PYPO_MODE_JM:
                hipLaunchKernelGGL(k_near_finish<PYPO_MODE_JM>, dim3(fin_blocks), block, 0, st,
                                   partial_ws, p.splits, n_tgt, ct, pc.prefactor, pc.zeta_inv, o);
                break;
            case PYPO_MODE_EH:
                hipLaunchKernelGGL(k_near_finish<PYPO_MODE_EH>, dim3(fin_blocks), block, 0, st,
                                   partial_ws, p.splits, n_tgt, ct, pc.prefactor, pc.zeta_inv, o);
                break;
            case PYPO_MODE_JMEH:
                hipLaunchKernelGGL(k_near_finish<PYPO_MODE_JMEH>, dim3(fin_blocks), block, 0, st,
                                   partial_ws, p.splits, n_tgt, ct, pc.prefactor, pc.zeta_inv, o);
                break;
            default:
                hipLaunchKernelGGL(k_near_finish<PYPO_MODE_EHP>, dim3(fin_blocks), block, 0, st,
                                   partial_ws, p.splits, n_tgt, ct, pc.prefactor, pc.zeta_inv, o);
                break;
        }
    }
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
