// grid.hip -- surface grid generation on the device (FP64).
//
// One thread per grid point computes position, unit normal and area element
// for the surface parametrisations of src/common/InterfaceReflector.cpp:
// Parabola/Hyperbola/Ellipse _xy and _uv (84-624), Plane_xy/_uv/_AoE
// (645-849), the affine transform of transformGrids (19-48, matVec4 of
// Utils.h:549-567), and the uncorrelated surface error along the normal
// (add_surferr_uncorr 852-866; the deviates themselves come from the host,
// where std::mt19937 + std::normal_distribution fix their order).
// Index i = xu * ncy + yv.  Memory-bound elementwise work: 56 B written per
// point, no inputs besides the parameters.

#include "common.hpp"

namespace pypo {
namespace {

constexpr double kPi = 3.14159265358979323846;

struct GridParams {
    double M[16];
    double xlo, xhi, ylo, yhi, xc, yc, ecc, rot, a, b, c;
    int nx, ny, nfac, type, gmode;
    bool transform, spheric;
};

__global__ __launch_bounds__(256) void k_generate_grid(GridParams P, GridPtrs g, int n,
                                                       const double *__restrict__ errs) {
    // No fused multiply-adds: each product and sum rounds on its own, as in
    // the reference's host build, so xy grids come out bit-identical to it
    // (a contracted j * dy + ylo leaves 1.8e-15 where the reference has an
    // exact 0, and atan2 at the origin then takes another branch).
#pragma clang fp contract(off)
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    int i = idx / P.ny, j = idx - i * P.ny;
    double x = 0, y = 0, z = 0, gx = 0, gy = 0, gz = 1, area = 0;
    double nfac = (double)P.nfac;

    if (P.gmode == 2) {
        // AoE far-field grid (Plane_AoE).  Normal is +z irrespective of flip.
        double dA = (P.xhi - P.xlo) / (P.nx - 1);
        double dE = (P.yhi - P.ylo) / (P.ny - 1);
        double Az = i * dA + P.xlo;
        double El = j * dE + P.ylo;
        area = cos(El * kPi / 180) * dA * dE * kPi * kPi / 180 / 180;
        if (P.spheric) {
            x = sqrt(Az * Az + El * El) * kPi / 180;
            y = atan2(El, Az);
            if (y != y) y = 0;
        } else {
            x = Az;
            y = El;
        }
        nfac = 1.0;
    } else if (P.gmode == 0) {
        double dx = (P.xhi - P.xlo) / (P.nx - 1);
        double dy = (P.yhi - P.ylo) / (P.ny - 1);
        // Ellipse_xy and Plane_xy do not add gcenter.
        bool centred = (P.type == 0 || P.type == 1);
        x = i * dx + P.xlo + (centred ? P.xc : 0.0);
        y = j * dy + P.ylo + (centred ? P.yc : 0.0);
        double nrm = 1;
        if (P.type == 0) {
            z = x * x / (P.a * P.a) + y * y / (P.b * P.b);
            gx = -2 * x / (P.a * P.a); gy = -2 * y / (P.b * P.b); gz = 1;
            nrm = sqrt(gx * gx + gy * gy + 1);
        } else if (P.type == 1) {
            z = P.c * sqrt(x * x / (P.a * P.a) + y * y / (P.b * P.b) + 1);
            gx = -2 * x / (P.a * P.a); gy = -2 * y / (P.b * P.b); gz = 2 * z / (P.c * P.c);
            nrm = sqrt(gx * gx + gy * gy + gz * gz);
        } else if (P.type == 2) {
            z = P.c * sqrt(1 - x * x / (P.a * P.a) - y * y / (P.b * P.b));
            gx = 2 * x / (P.a * P.a); gy = 2 * y / (P.b * P.b); gz = 2 * z / (P.c * P.c);
            nrm = sqrt(gx * gx + gy * gy + gz * gz);
        }
        if (P.type == 3) {
            z = 0; gx = 0; gy = 0; gz = nfac * 1.0;
            area = dx * dy;
        } else {
            gx = nfac * gx / nrm; gy = nfac * gy / nrm; gz = nfac * gz / nrm;
            area = nrm * dx * dy;
        }
    } else {
        double du0 = (P.xhi - P.xlo) / (P.nx - 1);
        double dv = (P.yhi - P.ylo) / (P.ny);   // upper v excluded
        double v = (j * dv + P.ylo) * kPi / 180;
        double sv = sin(v), cv = cos(v);
        double efac = sqrt(1 - P.ecc * P.ecc * sv * sv);
        double duv = du0 / efac;
        double u = i * duv + P.xlo / efac;
        double cr = cos(P.rot), sr = sin(P.rot);
        x = u * cv * cr - u * sv * sr + P.xc;
        y = u * cv * sr + u * sv * cr + P.yc;
        if (P.type == 3) {
            z = 0; gx = 0; gy = 0; gz = nfac * 1.0;
            area = u * duv * dv * kPi / 180;
        } else {
            double dudv = u / (efac * efac) * P.ecc * P.ecc * cv * sv;
            double dxdu = cv * cr - sv * sr;
            double dxdv = cv * cr * dudv - u * sv * cr - sv * sr * dudv - u * cv * sr;
            double dydu = cv * sr + sv * cr;
            double dydv = cv * sr * dudv - u * sv * sr + sv * cr * dudv + u * cv * cr;
            double dzdu, dzdv, nrm;
            if (P.type == 0) {
                z = x * x / P.a / P.a + y * y / P.b / P.b;
                dzdu = 2 * x * dxdu / P.a / P.a + 2 * y * dydu / P.b / P.b;
                dzdv = 2 * x * dxdv / P.a / P.a + 2 * y * dydv / P.b / P.b;
                gx = -2 * x / (P.a * P.a); gy = -2 * y / (P.b * P.b); gz = 1;
                nrm = sqrt(gx * gx + gy * gy + 1);
            } else if (P.type == 1) {
                z = P.c * sqrt(x * x / P.a / P.a + y * y / P.b / P.b + 1);
                dzdu = P.c * P.c / z * (x * dxdu / P.a / P.a + y * dydu / P.b / P.b);
                dzdv = P.c * P.c / z * (x * dxdv / P.a / P.a + y * dydv / P.b / P.b);
                gx = -2 * x / (P.a * P.a); gy = -2 * y / (P.b * P.b); gz = 2 * z / (P.c * P.c);
                nrm = sqrt(gx * gx + gy * gy + gz * gz);
            } else {
                z = P.c * sqrt(1 - x * x / P.a / P.a - y * y / P.b / P.b);
                dzdu = -P.c * P.c / z * (x * dxdu / P.a / P.a + y * dydu / P.b / P.b);
                dzdv = -P.c * P.c / z * (x * dxdv / P.a / P.a + y * dydv / P.b / P.b);
                gx = 2 * x / (P.a * P.a); gy = 2 * y / (P.b * P.b); gz = 2 * z / (P.c * P.c);
                nrm = sqrt(gx * gx + gy * gy + gz * gz);
            }
            gx = nfac * gx / nrm; gy = nfac * gy / nrm; gz = nfac * gz / nrm;
            double qx = dydu * dzdv - dzdu * dydv;
            double qy = dzdu * dxdv - dxdu * dzdv;
            double qz = dxdu * dydv - dydu * dxdv;
            area = sqrt(qx * qx + qy * qy + qz * qz) * duv * dv * kPi / 180;
        }
    }

    if (P.transform) {
        const double *M = P.M;
        double px = M[0] * x + M[1] * y + M[2] * z + M[3];
        double py = M[4] * x + M[5] * y + M[6] * z + M[7];
        double pz = M[8] * x + M[9] * y + M[10] * z + M[11];
        double qx = M[0] * gx + M[1] * gy + M[2] * gz;
        double qy = M[4] * gx + M[5] * gy + M[6] * gz;
        double qz = M[8] * gx + M[9] * gy + M[10] * gz;
        x = px; y = py; z = pz; gx = qx; gy = qy; gz = qz;
    }
    if (errs) {
        double e = errs[idx];
        x += e * gx; y += e * gy; z += e * gz;
    }
    g.x[idx] = x; g.y[idx] = y; g.z[idx] = z;
    g.nx[idx] = gx; g.ny[idx] = gy; g.nz[idx] = gz;
    g.area[idx] = area;
}

}  // namespace

int launch_generate_grid(const reflparams &r, const GridPtrs &g, int n, const double *d_errs,
                         bool transform, bool spheric, hipStream_t st) {
    PYPO_REQUIRE(r.n_cells && r.lxu && r.lyv && r.gcenter && r.coeffs && r.transf, ERR_ARG,
                 "reflparams has a NULL array");
    int nx = r.n_cells[0], ny = r.n_cells[1];
    PYPO_REQUIRE(nx > 0 && ny > 0 && (long long)nx * ny == n, ERR_ARG,
                 "grid size n_cells[0]*n_cells[1] does not match the container size");
    PYPO_REQUIRE(r.gmode >= 0 && r.gmode <= 2 && r.type >= 0 && r.type <= 3, ERR_ARG,
                 "unknown gmode or surface type");
    GridParams P;
    P.xlo = r.lxu[0]; P.xhi = r.lxu[1]; P.ylo = r.lyv[0]; P.yhi = r.lyv[1];
    P.xc = r.gcenter[0]; P.yc = r.gcenter[1];
    P.ecc = r.ecc_uv; P.rot = r.rot_uv * kPi / 180;
    P.a = r.coeffs[0]; P.b = r.coeffs[1]; P.c = r.coeffs[2];
    P.nx = nx; P.ny = ny; P.nfac = r.flip ? -1 : 1;
    P.type = r.type; P.gmode = r.gmode;
    for (int q = 0; q < 16; q++) P.M[q] = r.transf[q];
    P.transform = transform;
    P.spheric = spheric;
    int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_generate_grid, dim3(blocks), dim3(256), 0, st, P, g, n, d_errs);
    PYPO_CHECK_LAUNCH();
    return 0;
}

}  // namespace pypo
