// beam.hip -- beam initialisation on the device: Gaussian PO beams, the
// complex-source-point vector Gaussian, scalar Gaussians, currents from fields,
// tubular and Gaussian ray frames (src/common/BeamInit.h).  One thread per
// grid point or ray, FP64 throughout; every kernel is HBM-bound (O(N) work).
//
// Complex arithmetic follows the order of operations of the reference's
// std::complex<double> expressions so that results agree to rounding.

#include "common.hpp"

namespace pypo {
namespace {

constexpr double kPi = 3.14159265358979323846;   // M_PI, the constant BeamInit.h uses

struct cd {
    double r, i;
};

__device__ __forceinline__ cd operator+(cd a, cd b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cd operator*(cd a, cd b) {
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__device__ __forceinline__ cd operator*(double s, cd a) { return {s * a.r, s * a.i}; }
__device__ __forceinline__ cd operator*(cd a, double s) { return {a.r * s, a.i * s}; }
__device__ __forceinline__ cd conj(cd a) { return {a.r, -a.i}; }

// a / b, Smith's algorithm (robust against overflow of |b|^2).
__device__ __forceinline__ cd operator/(cd a, cd b) {
    if (fabs(b.r) >= fabs(b.i)) {
        const double t = b.i / b.r, d = b.r + b.i * t;
        return {(a.r + a.i * t) / d, (a.i - a.r * t) / d};
    }
    const double t = b.r / b.i, d = b.r * t + b.i;
    return {(a.r * t + a.i) / d, (a.i * t - a.r) / d};
}

// Principal square root (branch cut on the negative real axis, like std::sqrt).
__device__ __forceinline__ cd csqrt(cd z) {
    if (z.r == 0.0 && z.i == 0.0) return {0.0, z.i};
    const double t = sqrt(0.5 * (fabs(z.r) + hypot(z.r, z.i)));
    if (z.r >= 0.0) return {t, z.i / (2.0 * t)};
    return {fabs(z.i) / (2.0 * t), copysign(t, z.i)};
}

__device__ __forceinline__ cd cexp(cd z) {
    double s, c;
    sincos(z.i, &s, &c);
    const double e = exp(z.r);
    return {e * c, e * s};
}

// n x v for a real n and a complex v (Utils::ext, Utils.h).
__device__ __forceinline__ void cross_rc(const double n[3], const cd v[3], cd out[3]) {
    out[0] = n[1] * v[2] - n[2] * v[1];
    out[1] = n[2] * v[0] - n[0] * v[2];
    out[2] = n[0] * v[1] - n[1] * v[0];
}

__device__ __forceinline__ void put(const Bundle12 &b, int off, int i, const cd v[3], double f) {
    b.p[off][i] = f * v[0].r; b.p[off + 1][i] = f * v[1].r; b.p[off + 2][i] = f * v[2].r;
    b.p[off + 3][i] = f * v[0].i; b.p[off + 4][i] = f * v[1].i; b.p[off + 5][i] = f * v[2].i;
}

__device__ __forceinline__ void put_zero(const Bundle12 &b, int off, int i) {
    for (int q = 0; q < 6; q++) b.p[off + q][i] = 0.0;
}

// Rotation of (0, 0, 1) by Utils::matRot's Euler angles (Utils.h:613-635).
__device__ __forceinline__ void rot_chief(double rx, double ry, double rz, double out[3]) {
    const double cosx = cos(rx), cosy = cos(ry), cosz = cos(rz);
    const double sinx = sin(rx), siny = sin(ry), sinz = sin(rz);
    const double mu = cosz * siny, rho = sinz * siny;
    out[0] = mu * cosx + sinz * sinx;
    out[1] = rho * cosx - cosz * sinx;
    out[2] = cosy * cosx;
}

// ---- astigmatic Gaussian (initGauss, BeamInit.h:288-405) --------------------------
struct GaussP {
    double lam, w0x, w0y, n, E0, dxyz, pol[3];
};

__global__ __launch_bounds__(256) void k_gauss(GaussP g, GridPtrs gr, int n, Bundle12 field,
                                               Bundle12 cur) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double zRx = kPi * g.w0x * g.w0x * g.n / g.lam;
    const double zRy = kPi * g.w0y * g.w0y * g.n / g.lam;
    const double k = 2 * kPi / g.lam;
    const double x = gr.x[i], y = gr.y[i], z = gr.z[i];
    const double zy = z - g.dxyz;
    const double wzx = g.w0x * sqrt(1 + (z / zRx) * (z / zRx));
    const double wzy = g.w0y * sqrt(1 + (zy / zRy) * (zy / zRy));
    const double Rzx_inv = z / (z * z + zRx * zRx);
    const double Rzy_inv = zy / (zy * zy + zRy * zRy);
    const double phizx = atan(z / zRx), phizy = atan(zy / zRy);
    const cd arg = {-(x / wzx) * (x / wzx) - (y / wzy) * (y / wzy),
                    -kPi / g.lam * (x * x * Rzx_inv + y * y * Rzy_inv) - k * z + (phizx - phizy) * 0.5};
    const cd f = g.E0 * cexp(arg);
    const cd e[3] = {f * g.pol[0], f * g.pol[1], f * g.pol[2]};
    const double nv[3] = {gr.nx[i], gr.ny[i], gr.nz[i]};
    cd m[3];
    cross_rc(nv, e, m);
    put(field, 0, i, e, 1.0);
    put_zero(field, 6, i);
    put_zero(cur, 0, i);
    put(cur, 6, i, m, -2.0);
}

// ---- scalar Gaussian (initScalarGauss, BeamInit.h:598-658) -------------------------
__global__ __launch_bounds__(256) void k_scalar_gauss(GaussP g, GridPtrs gr, int n, double *ur,
                                                      double *ui) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double zRx = kPi * g.w0x * g.w0x * g.n / g.lam;
    const double zRy = kPi * g.w0y * g.w0y * g.n / g.lam;
    const double k = 2 * kPi / g.lam;
    const double x = gr.x[i], y = gr.y[i], z = gr.z[i];
    const double zy = z - g.dxyz;
    const double wzx = g.w0x * sqrt(1 + (z / zRx) * (z / zRx));
    const double wzy = g.w0y * sqrt(1 + (zy / zRy) * (zy / zRy));
    const double Rzx_inv = z / (z * z + zRx * zRx);
    const double Rzy_inv = zy / (zy * zy + zRy * zRy);
    const double phizx = atan(z / zRx), phizy = atan(zy / zRy);
    const cd arg = {-(x / wzx) * (x / wzx) - (y / wzy) * (y / wzy),
                    -kPi / g.lam * (x * x * Rzx_inv + y * y * Rzy_inv) - k * z + (phizx - phizy) * 0.5};
    const cd f = (g.E0 * sqrt(2 / (kPi * wzx * wzy))) * cexp(arg);
    ur[i] = f.r;
    ui[i] = f.i;
}

// ---- vector Gaussian by the complex source point (initGaussBeam, BeamInit.h:407-596)
struct VecGaussP {
    double lam, w0, z, n, power;
    int mode;
};

__global__ __launch_bounds__(256) void k_gauss_beam(VecGaussP g, GridPtrs gr, int n, Bundle12 field,
                                                    Bundle12 cur) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double k = 2 * kPi / g.lam;
    const double zc = k * g.w0 * g.w0 / 2.0;
    const double prefactor = g.w0 * k * sqrt(g.power / (8 * kPi));
    const cd j = {0.0, 1.0};
    const double zs = gr.z[i] + g.z;
    // BeamInit.h:473 applies the integer abs() to z + z0 (the only abs its
    // includes declare), so the axial coordinate of the complex source point is
    // |trunc(z + z0)|; reproduced for parity.
    const cd R[3] = {{gr.x[i], 0.0}, {gr.y[i], 0.0}, {fabs(trunc(zs)), zc}};
    const cd r = csqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    // exp(-j k r - k^2 w0^2 / 2)
    const cd mjk = {-0.0 * k, -k};
    const cd expo = cexp(mjk * r - cd{k * k * g.w0 * g.w0 / 2.0, 0.0});
    const cd kr_inv = cd{1.0, 0.0} / (k * r);
    const cd kr2 = kr_inv * kr_inv;
    const cd mj = {-0.0, -1.0};
    // the 1/kr power series, evaluated left to right as BeamInit.h:483-485 writes it
    const cd s1 = (mj * kr_inv - kr2) + ((j * kr_inv) * kr_inv) * kr_inv;
    const cd s2 = (j * kr_inv + (3.0 * kr_inv) * kr_inv) - (((3.0 * j) * kr_inv) * kr_inv) * kr_inv;
    const cd s3 = j * kr_inv + kr2;
    const cd rr = r * r;
    const cd pe = prefactor * expo;
    cd e[3], h[3];
    e[0] = pe * ((s1 + s2 * (R[0] * R[0] / rr)) - s3 * (R[2] / r));
    e[1] = pe * s2 * R[0] * R[1] / rr;
    e[2] = pe * (s2 * R[0] * R[2] / rr - s3 * R[0] / r);
    h[0] = pe * s2 * R[0] * R[1] / rr;
    h[1] = pe * ((s1 + s2 * (R[1] * R[1] / rr)) - s3 * (R[2] / r));
    h[2] = pe * (s2 * R[1] * R[2] / rr - s3 * R[1] / r);
    if (!(zs >= 0.0)) {
        for (int a = 0; a < 3; a++) { e[a] = conj(e[a]); h[a] = conj(h[a]); }
    }
    const double nv[3] = {gr.nx[i], gr.ny[i], gr.nz[i]};
    cd M[3], J[3];
    cross_rc(nv, e, M);
    cross_rc(nv, h, J);
    put(field, 0, i, e, 1.0);
    put(field, 6, i, h, 1.0);
    // The mode only selects values; both current vectors are always stored
    // (see k_calc_jm for why the stores are kept out of branches).
    const double F = (g.mode == 1 || g.mode == 2) ? 2.0 : 1.0;
    const bool keep_j = g.mode == 0 || g.mode == 2, keep_m = g.mode == 0 || g.mode == 1;
    cd js[3], ms[3];
    for (int a = 0; a < 3; a++) {
        js[a] = keep_j ? F * J[a] : cd{0.0, 0.0};
        ms[a] = keep_m ? -F * M[a] : cd{0.0, 0.0};
    }
    put(cur, 0, i, js, 1.0);
    put(cur, 6, i, ms, 1.0);
}

// ---- currents from fields (calcJM, BeamInit.h:660-844) ------------------------------
// mode 0 full: J = n x H, M = -n x E; 1 PMC: M = -2 n x E; 2 PEC: J = 2 n x H;
// 3 source: J = 2 n x E (the reference's choice).
// The mode selects values only; all twelve arrays are stored unconditionally.
// A switch with a store sequence per case made hipcc (ROCm 7.2, -O3, gfx950)
// merge the cases' common store tail and leave the scalar address registers of
// that tail unset on the mode-3 path, so the mode-3 stores went to stray
// addresses (an illegal memory access).
__global__ __launch_bounds__(256) void k_calc_jm(Bundle12 field, GridPtrs gr, int n, int mode,
                                                 Bundle12 cur) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double nv[3] = {gr.nx[i], gr.ny[i], gr.nz[i]};
    const cd e[3] = {{field.p[0][i], field.p[3][i]}, {field.p[1][i], field.p[4][i]},
                     {field.p[2][i], field.p[5][i]}};
    const cd h[3] = {{field.p[6][i], field.p[9][i]}, {field.p[7][i], field.p[10][i]},
                     {field.p[8][i], field.p[11][i]}};
    cd ne[3], nh[3];
    cross_rc(nv, e, ne);
    cross_rc(nv, h, nh);
    const double fj = mode == 0 ? 1.0 : 2.0;            // J factor (unused for PMC)
    const double fm = mode == 0 ? -1.0 : -2.0;          // M factor (PMC and full only)
    const bool j_from_e = mode == 3, j_zero = mode == 1, m_zero = mode >= 2;
    cd js[3], ms[3];
    for (int a = 0; a < 3; a++) {
        const cd jv = j_from_e ? ne[a] : nh[a];
        js[a] = j_zero ? cd{0.0, 0.0} : fj * jv;
        ms[a] = m_zero ? cd{0.0, 0.0} : fm * ne[a];
    }
    put(cur, 0, i, js, 1.0);
    put(cur, 6, i, ms, 1.0);
}

// ---- tubular ray frame (initFrame, BeamInit.h:159-214) ------------------------------
// Ray 0 is the chief ray; ray i >= 1 sits on ring n = (i-1)/B + 1 (B = nTot/nRing,
// integer division) at the angle reached by (i-1) - (n-1)B sequential additions
// of d_alpha, the running sum of the reference's loop.
struct RTP {
    int nRays, nRing;
    double angx0, angy0, x0, y0;
};

__global__ __launch_bounds__(256) void k_rt_frame(RTP r, int nTot, double *x, double *y, double *z,
                                                  double *dx, double *dy, double *dz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nTot) return;
    // every output is computed first and stored from one block (ray 0: chief ray)
    double o[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 1.0};
    if (i > 0) {
        const int B = nTot / r.nRing;
        const int ring = (i - 1) / B + 1;
        const int steps = (i - 1) - (ring - 1) * B;
        const double d_alpha = r.nRays > 0 ? 2 * kPi / (4 * r.nRays) : 0.0;
        double alpha = 0.0;
        for (int s = 0; s < steps; s++) alpha += d_alpha;
        const double nd = (double)ring;
        o[0] = r.x0 * cos(alpha) / r.nRing * nd;
        o[1] = r.y0 * sin(alpha) / r.nRing * nd;
        rot_chief(r.angy0 * sin(alpha) / r.nRing * nd, r.angx0 * cos(alpha) / r.nRing * nd, 2 * alpha,
                  o + 3);
    }
    x[i] = o[0]; y[i] = o[1]; z[i] = o[2];
    dx[i] = o[3]; dy[i] = o[4]; dz[i] = o[5];
}

// ---- Gaussian ray frame (initRTGauss, BeamInit.h:216-286) ---------------------------
// Same rejection test as the reference: candidates xi_k = U(-1, 1) * 3 * scale_k
// for (x, y, angx, angy), accepted when pdfGauss(xi) > U(0, 1).  Each ray draws
// its own stream from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed
// by the seed and counted by (trial, ray), so rays are independent of the
// launch shape.  A ray gives up after max_trials candidates and raises *fail.
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; round++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// 53-bit uniform double in [0, 1) from two 32-bit words.
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

struct GRTP {
    int nRays;
    double angx0, angy0, x0, y0;
    uint32_t k0, k1;
    int max_trials;
};

__global__ __launch_bounds__(256) void k_grt_frame(GRTP g, double *x, double *y, double *z, double *dx,
                                                   double *dy, double *dz, int *fail) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nRays) return;
    // every output is computed first and stored from one block (ray 0: chief ray)
    double o[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 1.0};
    bool ok = true;
    if (i > 0) {
        const double sc[4] = {g.x0, g.y0, g.angx0, g.angy0};
        const double norm = 1 / (kPi * kPi * sc[0] * sc[1] * sc[2] * sc[3]);
        const double thres = 3.0;
        ok = false;
        o[5] = 0.0;
        for (int t = 0; t < g.max_trials && !ok; t++) {
            uint32_t a[4] = {(uint32_t)t, (uint32_t)i, 0x6A09E667u, 0u};
            uint32_t b[4] = {(uint32_t)t, (uint32_t)i, 0x6A09E667u, 1u};
            uint32_t c[4] = {(uint32_t)t, (uint32_t)i, 0x6A09E667u, 2u};
            philox4x32(a, g.k0, g.k1);
            philox4x32(b, g.k0, g.k1);
            philox4x32(c, g.k0, g.k1);
            const double yi = u01(c[0], c[1]);
            const double u[4] = {u01(a[0], a[1]), u01(a[2], a[3]), u01(b[0], b[1]), u01(b[2], b[3])};
            double xi[4], ex = 0.0;
            for (int q = 0; q < 4; q++) {
                xi[q] = (-1.0 + 2.0 * u[q]) * thres * sc[q];
                ex += xi[q] * xi[q] / (sc[q] * sc[q]);
            }
            if (norm * exp(-ex) > yi) {
                ok = true;
                o[0] = xi[0];
                o[1] = xi[1];
                rot_chief(xi[3], xi[2], 0.0, o + 3);
            }
        }
    }
    x[i] = o[0]; y[i] = o[1]; z[i] = o[2];
    dx[i] = o[3]; dy[i] = o[4]; dz[i] = o[5];
    if (!ok) fail[0] = 1;
}

inline dim3 blocks_for(int n) { return dim3((n + 255) / 256); }

}  // namespace

int launch_gauss(const GPODict &g, const GridPtrs &gr, int n, const Bundle12 &field, const Bundle12 &cur,
                 hipStream_t st) {
    PYPO_REQUIRE(n > 0 && g.pol, ERR_ARG, "makeGauss: empty grid or NULL polarisation");
    GaussP p{g.lam, g.w0x, g.w0y, g.n, g.E0, g.dxyz, {g.pol[0], g.pol[1], g.pol[2]}};
    hipLaunchKernelGGL(k_gauss, blocks_for(n), dim3(256), 0, st, p, gr, n, field, cur);
    PYPO_CHECK_LAUNCH();
    return 0;
}

int launch_scalar_gauss(const ScalarGPODict &g, const GridPtrs &gr, int n, double *ur, double *ui,
                        hipStream_t st) {
    PYPO_REQUIRE(n > 0, ERR_ARG, "makeScalarGauss: empty grid");
    GaussP p{g.lam, g.w0x, g.w0y, g.n, g.E0, g.dxyz, {0, 0, 0}};
    hipLaunchKernelGGL(k_scalar_gauss, blocks_for(n), dim3(256), 0, st, p, gr, n, ur, ui);
    PYPO_CHECK_LAUNCH();
    return 0;
}

int launch_gauss_beam(const vecGPODict &g, const GridPtrs &gr, int n, const Bundle12 &field,
                      const Bundle12 &cur, hipStream_t st) {
    PYPO_REQUIRE(n > 0, ERR_ARG, "makeGaussBeam: empty grid");
    VecGaussP p{g.lam, g.w0, g.z, g.n, g.power, g.mode};
    hipLaunchKernelGGL(k_gauss_beam, blocks_for(n), dim3(256), 0, st, p, gr, n, field, cur);
    PYPO_CHECK_LAUNCH();
    return 0;
}

int launch_calc_jm(const Bundle12 &field, const GridPtrs &gr, int n, int mode, const Bundle12 &cur,
                   hipStream_t st) {
    PYPO_REQUIRE(n > 0, ERR_ARG, "calcCurrents: empty grid");
    PYPO_REQUIRE(mode >= 0 && mode <= 3, ERR_ARG,
                 "calcCurrents: mode must be 0 (full), 1 (PMC), 2 (PEC) or 3 (source)");
    hipLaunchKernelGGL(k_calc_jm, blocks_for(n), dim3(256), 0, st, field, gr, n, mode, cur);
    PYPO_CHECK_LAUNCH();
    return 0;
}

int rt_frame_size(const RTDict &r) { return 1 + r.nRing * 4 * r.nRays; }

int launch_rt_frame(const RTDict &r, const cframe &fr, hipStream_t st) {
    PYPO_REQUIRE(r.nRays >= 0 && r.nRing >= 0, ERR_ARG, "makeRTframe: nRays and nRing must be >= 0");
    const int n = rt_frame_size(r);
    PYPO_REQUIRE(fr.size >= n, ERR_ARG, "makeRTframe: frame smaller than 1 + 4 nRays nRing");
    RTP p{r.nRays, r.nRing, r.angx0, r.angy0, r.x0, r.y0};
    hipLaunchKernelGGL(k_rt_frame, blocks_for(n), dim3(256), 0, st, p, n, fr.x, fr.y, fr.z, fr.dx,
                       fr.dy, fr.dz);
    PYPO_CHECK_LAUNCH();
    return 0;
}

int launch_grt_frame(const GRTDict &g, uint64_t seed, const cframe &fr, int *d_fail, int max_trials,
                     hipStream_t st) {
    PYPO_REQUIRE(g.nRays >= 1 && fr.size >= g.nRays, ERR_ARG,
                 "makeGRTframe: nRays must be >= 1 and fit the frame");
    PYPO_REQUIRE(g.x0 > 0 && g.y0 > 0 && g.angx0 > 0 && g.angy0 > 0, ERR_ARG,
                 "makeGRTframe: x0, y0, angx0, angy0 must be positive");
    GRTP p{g.nRays, g.angx0, g.angy0, g.x0, g.y0, (uint32_t)seed, (uint32_t)(seed >> 32), max_trials};
    hipLaunchKernelGGL(k_grt_frame, blocks_for(g.nRays), dim3(256), 0, st, p, fr.x, fr.y, fr.z, fr.dx,
                       fr.dy, fr.dz, d_fail);
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
