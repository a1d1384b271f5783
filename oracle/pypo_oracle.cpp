// pypo_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of the reference's physical-optics and ray-trace
// path, used as the parity checker by tests/, by __graft_entry__.smoke()
// and as the timed CPU baseline in bench.py.  Nothing in the product
// (pypo_amd/, libpypohip.so) links, loads or calls this file.
//
// It restates, in this project's own code, the algorithms of
//   src/CPU/Propagation.h        (PO integrands, loop order, trapezoid rule,
//                                 shadowing test, std::thread split)
//   src/CPU/RayTrace.h           (Newton ray tracer, miss rule, transforms)
//   src/include/RTRefls.h        (quadric Newton updates and normals)
//   src/include/Utils.h          (vector helpers: normalize, snell, matVec4)
//   src/common/InterfaceReflector.cpp (grid generation)
//   src/include/Random.h         (mt19937 + normal_distribution draws)
// with std::complex<double> arithmetic and the same per-target loop order,
// so that it is also a fair CPU baseline.
//
// Pinning (see tests/test_oracle.py): every entry point is compared with
// golden vectors written by the compiled reference (tests/golden/*.npz),
// with oracle/_ref when it is built, and with closed forms (plate far
// field, single-dipole near field, spherical wave, ray-quadric quadratic).
//
// Exported names carry an "oracle_" prefix so the checker can never be
// mistaken for the product's drop-in symbols.

#include <algorithm>
#include <cmath>
#include <complex>
#include <random>
#include <thread>
#include <vector>

#include "../include/pypo_hip.h"

namespace {

typedef std::complex<double> cplx;

// Constants of the reference, which stores float literals in T=double
// members (Propagation.h:196-203): the float rounding is part of the result.
struct Consts {
    double pi_f, c_l, mu0, eps_vac, eps, zeta, zeta_inv, prefactor;
    Consts(double k, double epsilon) {
        pi_f = (double)3.14159265359f;
        c_l = (double)2.99792458e8f;
        mu0 = (double)1.2566370614e-6f;
        eps_vac = 1 / (mu0 * c_l * c_l);
        eps = epsilon * eps_vac;
        zeta = std::sqrt(mu0 / eps);
        zeta_inv = 1 / zeta;
        prefactor = k * k / (4. * pi_f);
    }
};

const double kPi = 3.14159265358979323846;  // InterfaceReflector.cpp:4

struct V3 { double v[3]; };
struct C3 { cplx v[3]; };

inline double dot3(const double *a, const double *b) {
    double s = 0;
    for (int n = 0; n < 3; n++) s += a[n] * b[n];
    return s;
}

inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Utils.h:288-302: divide by |v|, with |v| floored at denorm_min.
inline void unit3(const double *v, double *o) {
    double nrm = std::sqrt(dot3(v, v));
    if (nrm <= std::numeric_limits<double>::denorm_min()) nrm = std::numeric_limits<double>::denorm_min();
    for (int n = 0; n < 3; n++) o[n] = v[n] / nrm;
}

// Static contiguous split of [0, total): step = total / threads (integer
// division, Propagation.h:224), the last thread takes the remainder
// (Propagation.h:1356-1377, RayTrace.h:283-298).
template <class F>
void split_run(int total, int threads, F fn) {
    if (threads < 1) threads = 1;
    int step = total / threads;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        int lo = t * step;
        int hi = (t == threads - 1) ? total : (t + 1) * step;
        pool.emplace_back(fn, lo, hi);
    }
    for (auto &th : pool) th.join();
}

// ---------------------------------------------------------------------
// Grid generation -- InterfaceReflector.cpp:35-972
// ---------------------------------------------------------------------

struct GridArgs {
    double xlo, xhi, ylo, yhi, xc, yc, ecc, rot, a, b, c;
    int nx, ny, nfac, type, gmode;
    const double *M;
    bool transform, spheric;
};

// matVec4 (Utils.h:549-567) applied to point and normal, in place.
inline void apply_transf(const double *M, reflcontainer *g, int i) {
    double p[3] = {g->x[i], g->y[i], g->z[i]};
    double q[3] = {g->nx[i], g->ny[i], g->nz[i]};
    double po[3], qo[3];
    for (int n = 0; n < 3; n++) {
        po[n] = M[4 * n] * p[0] + M[4 * n + 1] * p[1] + M[4 * n + 2] * p[2] + M[4 * n + 3];
        qo[n] = M[4 * n] * q[0] + M[4 * n + 1] * q[1] + M[4 * n + 2] * q[2];
    }
    g->x[i] = po[0]; g->y[i] = po[1]; g->z[i] = po[2];
    g->nx[i] = qo[0]; g->ny[i] = qo[1]; g->nz[i] = qo[2];
}

inline void store_point(reflcontainer *g, int i, double x, double y, double z,
                        double nx, double ny, double nz, double area) {
    g->x[i] = x; g->y[i] = y; g->z[i] = z;
    g->nx[i] = nx; g->ny[i] = ny; g->nz[i] = nz;
    g->area[i] = area;
}

// Quadric surface value and (unnormalised) gradient normal for the xy
// generators: Parabola_xy 84-130, Hyperbola_xy 212-254, Ellipse_xy 335-376.
void grid_xy_quadric(const GridArgs &A, reflcontainer *g) {
    double dx = (A.xhi - A.xlo) / (A.nx - 1);
    double dy = (A.yhi - A.ylo) / (A.ny - 1);
    bool center = (A.type != 2);   // Ellipse_xy ignores gcenter (:353-357)
    for (int i = 0; i < A.nx; i++) {
        double x = i * dx + A.xlo + (center ? A.xc : 0.0);
        for (int j = 0; j < A.ny; j++) {
            double y = j * dy + A.ylo + (center ? A.yc : 0.0);
            int idx = i * A.ny + j;
            double z, gx, gy, gz;
            if (A.type == 0) {
                z = x * x / (A.a * A.a) + y * y / (A.b * A.b);
                gx = -2 * x / (A.a * A.a); gy = -2 * y / (A.b * A.b); gz = 1;
            } else if (A.type == 1) {
                z = A.c * std::sqrt(x * x / (A.a * A.a) + y * y / (A.b * A.b) + 1);
                gx = -2 * x / (A.a * A.a); gy = -2 * y / (A.b * A.b); gz = 2 * z / (A.c * A.c);
            } else {
                z = A.c * std::sqrt(1 - x * x / (A.a * A.a) - y * y / (A.b * A.b));
                gx = 2 * x / (A.a * A.a); gy = 2 * y / (A.b * A.b); gz = 2 * z / (A.c * A.c);
            }
            double nrm = (A.type == 0) ? std::sqrt(gx * gx + gy * gy + 1)
                                       : std::sqrt(gx * gx + gy * gy + gz * gz);
            store_point(g, idx, x, y, z, A.nfac * gx / nrm, A.nfac * gy / nrm, A.nfac * gz / nrm,
                        nrm * dx * dy);
            if (A.transform) apply_transf(A.M, g, idx);
        }
    }
}

// uv generators: Parabola_uv 164-245, Hyperbola_uv 352-438, Ellipse_uv 541-624.
void grid_uv_quadric(const GridArgs &A, reflcontainer *g) {
    double du0 = (A.xhi - A.xlo) / (A.nx - 1);
    double dv = (A.yhi - A.ylo) / (A.ny);       // upper v excluded
    for (int i = 0; i < A.nx; i++) {
        for (int j = 0; j < A.ny; j++) {
            double v = (j * dv + A.ylo) * kPi / 180;
            double sv = std::sin(v), cv = std::cos(v);
            double efac = std::sqrt(1 - A.ecc * A.ecc * sv * sv);
            double duv = du0 / efac;
            double u = i * duv + A.xlo / efac;
            double dudv = u / (efac * efac) * A.ecc * A.ecc * cv * sv;
            double cr = std::cos(A.rot), sr = std::sin(A.rot);
            int idx = i * A.ny + j;
            double x = u * cv * cr - u * sv * sr + A.xc;
            double y = u * cv * sr + u * sv * cr + A.yc;
            double dxdu = cv * cr - sv * sr;
            double dxdv = cv * cr * dudv - u * sv * cr - sv * sr * dudv - u * cv * sr;
            double dydu = cv * sr + sv * cr;
            double dydv = cv * sr * dudv - u * sv * sr + sv * cr * dudv + u * cv * cr;
            double z, dzdu, dzdv, gx, gy, gz;
            if (A.type == 0) {
                z = x * x / A.a / A.a + y * y / A.b / A.b;
                dzdu = 2 * x * dxdu / A.a / A.a + 2 * y * dydu / A.b / A.b;
                dzdv = 2 * x * dxdv / A.a / A.a + 2 * y * dydv / A.b / A.b;
                gx = -2 * x / (A.a * A.a); gy = -2 * y / (A.b * A.b); gz = 1;
            } else if (A.type == 1) {
                z = A.c * std::sqrt(x * x / A.a / A.a + y * y / A.b / A.b + 1);
                dzdu = A.c * A.c / z * (x * dxdu / A.a / A.a + y * dydu / A.b / A.b);
                dzdv = A.c * A.c / z * (x * dxdv / A.a / A.a + y * dydv / A.b / A.b);
                gx = -2 * x / (A.a * A.a); gy = -2 * y / (A.b * A.b); gz = 2 * z / (A.c * A.c);
            } else {
                z = A.c * std::sqrt(1 - x * x / A.a / A.a - y * y / A.b / A.b);
                dzdu = -A.c * A.c / z * (x * dxdu / A.a / A.a + y * dydu / A.b / A.b);
                dzdv = -A.c * A.c / z * (x * dxdv / A.a / A.a + y * dydv / A.b / A.b);
                gx = 2 * x / (A.a * A.a); gy = 2 * y / (A.b * A.b); gz = 2 * z / (A.c * A.c);
            }
            double nrm = (A.type == 0) ? std::sqrt(gx * gx + gy * gy + 1)
                                       : std::sqrt(gx * gx + gy * gy + gz * gz);
            double Qu[3] = {dxdu, dydu, dzdu}, Qv[3] = {dxdv, dydv, dzdv}, Q[3];
            cross3(Qu, Qv, Q);
            double jac = std::sqrt(dot3(Q, Q));
            store_point(g, idx, x, y, z, A.nfac * gx / nrm, A.nfac * gy / nrm, A.nfac * gz / nrm,
                        jac * duv * dv * kPi / 180);
            if (A.transform) apply_transf(A.M, g, idx);
        }
    }
}

// Plane_xy 645-690 (gcenter ignored), Plane_uv 713-770, Plane_AoE 794-849.
void grid_plane(const GridArgs &A, reflcontainer *g) {
    if (A.gmode == 0) {
        double dx = (A.xhi - A.xlo) / (A.nx - 1);
        double dy = (A.yhi - A.ylo) / (A.ny - 1);
        for (int i = 0; i < A.nx; i++) {
            double x = i * dx + A.xlo;
            for (int j = 0; j < A.ny; j++) {
                int idx = i * A.ny + j;
                store_point(g, idx, x, j * dy + A.ylo, 0, 0, 0, A.nfac * 1, dx * dy);
                if (A.transform) apply_transf(A.M, g, idx);
            }
        }
    } else if (A.gmode == 1) {
        double du0 = (A.xhi - A.xlo) / (A.nx - 1);
        double dv = (A.yhi - A.ylo) / (A.ny);
        for (int i = 0; i < A.nx; i++) {
            for (int j = 0; j < A.ny; j++) {
                double v = (j * dv + A.ylo) * kPi / 180;
                double sv = std::sin(v), cv = std::cos(v);
                double efac = std::sqrt(1 - A.ecc * A.ecc * sv * sv);
                double duv = du0 / efac;
                double u = i * duv + A.xlo / efac;
                double cr = std::cos(A.rot), sr = std::sin(A.rot);
                int idx = i * A.ny + j;
                double x = u * cv * cr - u * sv * sr + A.xc;
                double y = u * cv * sr + u * sv * cr + A.yc;
                store_point(g, idx, x, y, 0, 0, 0, A.nfac * 1, u * duv * dv * kPi / 180);
                if (A.transform) apply_transf(A.M, g, idx);
            }
        }
    } else {
        double dA = (A.xhi - A.xlo) / (A.nx - 1);
        double dE = (A.yhi - A.ylo) / (A.ny - 1);
        for (int i = 0; i < A.nx; i++) {
            double Az = i * dA + A.xlo;
            for (int j = 0; j < A.ny; j++) {
                double El = j * dE + A.ylo;
                int idx = i * A.ny + j;
                double area = std::cos(El * kPi / 180) * dA * dE * kPi * kPi / 180 / 180;
                double gx, gy;
                if (A.spheric) {
                    gx = std::sqrt(Az * Az + El * El) * kPi / 180;
                    gy = std::atan2(El, Az);
                    if (gy != gy) gy = 0;
                } else {
                    gx = Az; gy = El;
                }
                store_point(g, idx, gx, gy, 0, 0, 0, 1, area);   // AoE normal ignores flip
                if (A.transform) apply_transf(A.M, g, idx);
            }
        }
    }
}

// add_surferr_uncorr (InterfaceReflector.cpp:852-866): a fresh
// normal_distribution per point (Random.h generateNormal(T stdev)), so the
// polar method's cached second deviate is discarded every time.
void add_roughness(const reflparams &r, reflcontainer *g) {
    std::mt19937 gen(r.rms_seed);
    int num = r.n_cells[0] * r.n_cells[1];
    for (int n = 0; n < num; ++n) {
        std::normal_distribution<double> dist(0.0, r.rms);
        double e = dist(gen);
        g->x[n] += e * g->nx[n];
        g->y[n] += e * g->ny[n];
        g->z[n] += e * g->nz[n];
    }
}

void make_grid(const reflparams &r, reflcontainer *g, bool transform, bool spheric) {
    GridArgs A;
    A.xlo = r.lxu[0]; A.xhi = r.lxu[1]; A.ylo = r.lyv[0]; A.yhi = r.lyv[1];
    A.xc = r.gcenter[0]; A.yc = r.gcenter[1];
    A.ecc = r.ecc_uv; A.rot = r.rot_uv * kPi / 180;
    A.a = r.coeffs[0]; A.b = r.coeffs[1]; A.c = r.coeffs[2];
    A.nx = r.n_cells[0]; A.ny = r.n_cells[1];
    A.nfac = r.flip ? -1 : 1;
    A.type = r.type; A.gmode = r.gmode; A.M = r.transf;
    A.transform = transform; A.spheric = spheric;
    if (r.gmode == 2) {
        grid_plane(A, g);
    } else if (r.type == 3) {
        grid_plane(A, g);
    } else if (r.gmode == 0) {
        grid_xy_quadric(A, g);
    } else if (r.gmode == 1) {
        grid_uv_quadric(A, g);
    }
    if (r.rms > 0) add_roughness(r, g);
}

// ---------------------------------------------------------------------
// PO sums -- Propagation.h
// ---------------------------------------------------------------------

struct Src {
    const reflcontainer *cs;
    const c2Bundle *cur;
    int gmode, ncx, ncy;
};

inline void load_currents(const c2Bundle *c, int i, cplx *J, cplx *M) {
    J[0] = cplx(c->r1x[i], c->i1x[i]); J[1] = cplx(c->r1y[i], c->i1y[i]); J[2] = cplx(c->r1z[i], c->i1z[i]);
    M[0] = cplx(c->r2x[i], c->i2x[i]); M[1] = cplx(c->r2y[i], c->i2y[i]); M[2] = cplx(c->r2z[i], c->i2z[i]);
}

// Near-field E, H at one target point.  fieldAtPoint, Propagation.h:977-1144.
// Trapezoid rule: half weight on yv edges unless gmode == 1 (uv), and on
// xu edges always (:1123-1135).
void near_field_at(const Src &S, const Consts &K, double k, double t, const double *pt,
                   cplx *E, cplx *H) {
    const cplx j(0., 1.);
    const double half = 0.5, three = 3.0;
    for (int n = 0; n < 3; n++) { E[n] = 0; H[n] = 0; }
    for (int xu = 0; xu < S.ncx; xu++) {
        cplx ye[3] = {0., 0., 0.}, yh[3] = {0., 0., 0.};
        for (int yv = 0; yv < S.ncy; yv++) {
            int i = xu * S.ncy + yv;
            const reflcontainer *cs = S.cs;
            double rv[3] = {pt[0] - cs->x[i], pt[1] - cs->y[i], pt[2] - cs->z[i]};
            double R = std::sqrt(dot3(rv, rv));
            double Rinv = 1 / R;
            double kR = k * R;
            double q = 1 / kR;
            double rh[3] = {rv[0] * Rinv, rv[1] * Rinv, rv[2] * Rinv};
            double nrm[3] = {cs->nx[i], cs->ny[i], cs->nz[i]};
            if (dot3(nrm, rh) < 0) continue;
            cplx Jc[3], Mc[3];
            load_currents(S.cur, i, Jc, Mc);
            cplx s1 = q * (-j + t * q + j * q * q);
            cplx s2 = q * (j - t * three * q - three * j * q * q);
            cplx s3 = t * (j * q - t * q * q);
            cplx JdR = 0., MdR = 0.;
            for (int n = 0; n < 3; n++) { JdR += Jc[n] * rh[n]; MdR += Mc[n] * rh[n]; }
            cplx MxR[3], JxR[3];
            MxR[0] = Mc[1] * rh[2] - Mc[2] * rh[1]; MxR[1] = Mc[2] * rh[0] - Mc[0] * rh[2]; MxR[2] = Mc[0] * rh[1] - Mc[1] * rh[0];
            JxR[0] = Jc[1] * rh[2] - Jc[2] * rh[1]; JxR[1] = Jc[2] * rh[0] - Jc[0] * rh[2]; JxR[2] = Jc[0] * rh[1] - Jc[1] * rh[0];
            cplx G = K.prefactor * std::exp(t * j * k * R) * cs->area[i];
            bool edge = (S.gmode != 1) && (yv == 0 || yv == S.ncy - 1);
            for (int n = 0; n < 3; n++) {
                cplx e = (Jc[n] * s1 + (rh[n] * JdR) * s2 + MxR[n] * s3) * G;
                cplx h = (Mc[n] * s1 + (rh[n] * MdR) * s2 - JxR[n] * s3) * G;
                if (edge) { ye[n] += half * e; yh[n] += half * h; }
                else { ye[n] += e; yh[n] += h; }
            }
        }
        bool xedge = (xu == 0 || xu == S.ncx - 1);
        for (int n = 0; n < 3; n++) {
            E[n] += xedge ? half * ye[n] : ye[n];
            H[n] += xedge ? half * yh[n] : yh[n];
        }
    }
}

// Far field at direction rh.  farfieldAtPoint, Propagation.h:1656-1781.
void far_field_at(const Src &S, const Consts &K, double k, double t, const double *rh,
                  cplx *E, cplx *H) {
    const cplx j(0., 1.);
    const double half = 0.5;
    for (int n = 0; n < 3; n++) { E[n] = 0; H[n] = 0; }
    for (int xu = 0; xu < S.ncx; xu++) {
        cplx ye[3] = {0., 0., 0.}, yh[3] = {0., 0., 0.};
        for (int yv = 0; yv < S.ncy; yv++) {
            int i = xu * S.ncy + yv;
            const reflcontainer *cs = S.cs;
            double nrm[3] = {cs->nx[i], cs->ny[i], cs->nz[i]};
            if (dot3(nrm, rh) < 0) continue;
            cplx Jc[3], Mc[3];
            load_currents(S.cur, i, Jc, Mc);
            cplx JdR = 0., MdR = 0.;
            for (int n = 0; n < 3; n++) { JdR += Jc[n] * rh[n]; MdR += Mc[n] * rh[n]; }
            cplx RxM[3], RxJ[3];
            RxM[0] = rh[1] * Mc[2] - rh[2] * Mc[1]; RxM[1] = rh[2] * Mc[0] - rh[0] * Mc[2]; RxM[2] = rh[0] * Mc[1] - rh[1] * Mc[0];
            RxJ[0] = rh[1] * Jc[2] - rh[2] * Jc[1]; RxJ[1] = rh[2] * Jc[0] - rh[0] * Jc[2]; RxJ[2] = rh[0] * Jc[1] - rh[1] * Jc[0];
            double p[3] = {cs->x[i], cs->y[i], cs->z[i]};
            double rdr = dot3(p, rh);
            cplx G = -K.prefactor * j * std::exp(-t * j * k * rdr) * cs->area[i];
            bool edge = (S.gmode != 1) && (yv == 0 || yv == S.ncy - 1);
            for (int n = 0; n < 3; n++) {
                cplx e = (Jc[n] - rh[n] * JdR + t * RxM[n]) * G;
                cplx h = (Mc[n] - rh[n] * MdR - t * RxJ[n]) * G;
                if (edge) { ye[n] += half * e; yh[n] += half * h; }
                else { ye[n] += e; yh[n] += h; }
            }
        }
        bool xedge = (xu == 0 || xu == S.ncx - 1);
        for (int n = 0; n < 3; n++) {
            E[n] += xedge ? half * ye[n] : ye[n];
            H[n] += xedge ? half * yh[n] : yh[n];
        }
    }
}

// Scalar field.  fieldScalarAtPoint, Propagation.h:1279-1331 (no shadowing).
cplx scalar_at(const reflcontainer *cs, const arrC1 *f, int gmode, int ncx, int ncy,
               const Consts &K, double k, double t, const double *pt) {
    const cplx j(0., 1.);
    const double half = 0.5;
    cplx out = 0.;
    for (int xu = 0; xu < ncx; xu++) {
        cplx yacc = 0.;
        for (int yv = 0; yv < ncy; yv++) {
            int i = xu * ncy + yv;
            cplx u(f->x[i], f->y[i]);
            double rv[3] = {pt[0] - cs->x[i], pt[1] - cs->y[i], pt[2] - cs->z[i]};
            double r = std::sqrt(dot3(rv, rv));
            cplx term = -k * k * u * std::exp(t * -j * k * r) / (4 * K.pi_f * r) * cs->area[i];
            if ((gmode != 1) && (yv == 0 || yv == ncy - 1)) yacc += half * term;
            else yacc += term;
        }
        out += (xu == 0 || xu == ncx - 1) ? half * yacc : yacc;
    }
    return out;
}

inline void put_c3(double *const *rr, double *const *ii, int i, const cplx *v, double scale = 1.0) {
    for (int n = 0; n < 3; n++) { rr[n][i] = scale * v[n].real(); ii[n][i] = scale * v[n].imag(); }
}

enum Mode { JM = 0, EH = 1, JMEH = 2, EHP = 3, FF = 4 };

// Per-target epilogues: propagateBeam_JM_PEC 269-323, _EH 574-617,
// _JMEH_PEC 482-555, _EHP 805-914, propagateToFarField 1581-1639.
void propagate_on(int mode, void *res, int gmode, int ncx, int ncy, reflcontainer *cs,
                  reflcontainer *ct, const c2Bundle *cur, double k, int nthreads, double epsilon,
                  double t) {
    Consts K(k, epsilon);
    Src S{cs, cur, gmode, ncx, ncy};
    split_run(ct->size, nthreads, [&](int lo, int hi) {
        for (int i = lo; i < hi; i++) {
            cplx E[3], H[3];
            if (mode == FF) {
                double th = ct->x[i], ph = ct->y[i];
                double rh[3] = {std::cos(ph) * std::sin(th), std::sin(ph) * std::sin(th), std::cos(th)};
                far_field_at(S, K, k, t, rh, E, H);
                c2Bundle *r = (c2Bundle *)res;
                double *er[3] = {r->r1x, r->r1y, r->r1z}, *ei[3] = {r->i1x, r->i1y, r->i1z};
                double *hr[3] = {r->r2x, r->r2y, r->r2z}, *hi_[3] = {r->i2x, r->i2y, r->i2z};
                put_c3(er, ei, i, E); put_c3(hr, hi_, i, H);
                continue;
            }
            double pt[3] = {ct->x[i], ct->y[i], ct->z[i]};
            double nt[3] = {ct->nx[i], ct->ny[i], ct->nz[i]};
            near_field_at(S, K, k, t, pt, E, H);
            if (mode == JM || mode == JMEH) {
                cplx Jt[3];
                Jt[0] = nt[1] * H[2] - nt[2] * H[1];
                Jt[1] = nt[2] * H[0] - nt[0] * H[2];
                Jt[2] = nt[0] * H[1] - nt[1] * H[0];
                cplx zero[3] = {0., 0., 0.};
                double *jr[3], *ji[3], *mr[3], *mi[3];
                if (mode == JM) {
                    c2Bundle *r = (c2Bundle *)res;
                    jr[0] = r->r1x; jr[1] = r->r1y; jr[2] = r->r1z; ji[0] = r->i1x; ji[1] = r->i1y; ji[2] = r->i1z;
                    mr[0] = r->r2x; mr[1] = r->r2y; mr[2] = r->r2z; mi[0] = r->i2x; mi[1] = r->i2y; mi[2] = r->i2z;
                } else {
                    c4Bundle *r = (c4Bundle *)res;
                    jr[0] = r->r1x; jr[1] = r->r1y; jr[2] = r->r1z; ji[0] = r->i1x; ji[1] = r->i1y; ji[2] = r->i1z;
                    mr[0] = r->r2x; mr[1] = r->r2y; mr[2] = r->r2z; mi[0] = r->i2x; mi[1] = r->i2y; mi[2] = r->i2z;
                    double *er[3] = {r->r3x, r->r3y, r->r3z}, *ei[3] = {r->i3x, r->i3y, r->i3z};
                    double *hr[3] = {r->r4x, r->r4y, r->r4z}, *hi_[3] = {r->i4x, r->i4y, r->i4z};
                    put_c3(er, ei, i, E); put_c3(hr, hi_, i, H);
                }
                put_c3(jr, ji, i, Jt, 2.0);
                put_c3(mr, mi, i, zero);
            } else if (mode == EH) {
                c2Bundle *r = (c2Bundle *)res;
                double *er[3] = {r->r1x, r->r1y, r->r1z}, *ei[3] = {r->i1x, r->i1y, r->i1z};
                double *hr[3] = {r->r2x, r->r2y, r->r2z}, *hi_[3] = {r->i2x, r->i2y, r->i2z};
                put_c3(er, ei, i, E); put_c3(hr, hi_, i, H);
            } else {   // EHP
                // incident Poynting direction Re(E x H*), normalised
                cplx Hc[3] = {std::conj(H[0]), std::conj(H[1]), std::conj(H[2])};
                cplx ExH[3];
                ExH[0] = E[1] * Hc[2] - E[2] * Hc[1];
                ExH[1] = E[2] * Hc[0] - E[0] * Hc[2];
                ExH[2] = E[0] * Hc[1] - E[1] * Hc[0];
                double S_re[3] = {ExH[0].real(), ExH[1].real(), ExH[2].real()};
                double Si[3], tmp[3], pip[3], pia[3], Sr[3], prp[3], pra[3];
                unit3(S_re, Si);
                cross3(Si, nt, tmp); unit3(tmp, pip);
                cross3(pip, Si, pia);
                double f = 2. * dot3(Si, nt);                       // snell, Utils.h:421-432
                for (int n = 0; n < 3; n++) Sr[n] = Si[n] - nt[n] * f;
                cross3(Sr, nt, tmp); unit3(tmp, prp);
                cross3(Sr, prp, pra);
                cplx e_perp = 0., e_par = 0.;
                for (int n = 0; n < 3; n++) { e_perp += E[n] * prp[n]; e_par += E[n] * pra[n]; }
                cplx Er[3], Hr[3];
                for (int n = 0; n < 3; n++) Er[n] = -e_perp * pip[n] - e_par * pia[n];
                cplx SxE[3];
                SxE[0] = Sr[1] * Er[2] - Sr[2] * Er[1];
                SxE[1] = Sr[2] * Er[0] - Sr[0] * Er[2];
                SxE[2] = Sr[0] * Er[1] - Sr[1] * Er[0];
                for (int n = 0; n < 3; n++) Hr[n] = SxE[n] * K.zeta_inv;
                c2rBundle *r = (c2rBundle *)res;
                double *er[3] = {r->r1x, r->r1y, r->r1z}, *ei[3] = {r->i1x, r->i1y, r->i1z};
                double *hr[3] = {r->r2x, r->r2y, r->r2z}, *hi_[3] = {r->i2x, r->i2y, r->i2z};
                put_c3(er, ei, i, Er); put_c3(hr, hi_, i, Hr);
                r->r3x[i] = Sr[0]; r->r3y[i] = Sr[1]; r->r3z[i] = Sr[2];
            }
        }
    });
}

void propagate(int mode, void *res, const reflparams &sp, const reflparams &tp,
               reflcontainer *cs, reflcontainer *ct, const c2Bundle *cur,
               double k, int nthreads, double epsilon, double t) {
    make_grid(sp, cs, true, true);
    make_grid(tp, ct, true, true);
    propagate_on(mode, res, sp.gmode, sp.n_cells[0], sp.n_cells[1], cs, ct, cur, k, nthreads, epsilon, t);
}

// ---------------------------------------------------------------------
// Ray tracer -- RayTrace.h:139-305, RTRefls.h:52-297
// ---------------------------------------------------------------------

inline double q1(double t, double x, double y, double dx, double dy, double a, double b) {
    return (x + t * dx) * (x + t * dx) / (a * a) + (y + t * dy) * (y + t * dy) / (b * b);
}
inline double q2(double t, double x, double y, double dx, double dy, double a, double b) {
    return (x + t * dx) * 2 * dx / (a * a) + (y + t * dy) * 2 * dy / (b * b);
}

// One Newton update t <- t - f(t)/f'(t) for the surface type.
inline double newton_step(int type, double t, double x, double y, double z, double dx, double dy,
                          double dz, double a, double b, double c) {
    if (type == 0)
        return t - (z + t * dz - q1(t, x, y, dx, dy, a, b)) / (dz - q2(t, x, y, dx, dy, a, b));
    if (type == 1)
        return t - (z + t * dz - c * std::sqrt(q1(t, x, y, dx, dy, a, b) + 1)) /
                       (dz - c / (2 * std::sqrt(q1(t, x, y, dx, dy, a, b) + 1)) * q2(t, x, y, dx, dy, a, b));
    if (type == 2)
        return t - (z + t * dz - c * std::sqrt(1 - q1(t, x, y, dx, dy, a, b))) /
                       (dz + c / (2 * std::sqrt(1 - q1(t, x, y, dx, dy, a, b))) * q2(t, x, y, dx, dy, a, b));
    return t - (z + t * dz) / dz;
}

inline void surface_normal(int type, double x, double y, double z, int flip, double a, double b,
                           double c, double *n) {
    if (type == 3) { n[0] = 0; n[1] = 0; n[2] = 1; return; }   // npl ignores flip
    if (type == 0) { n[0] = -2 * x / (a * a) * flip; n[1] = -2 * y / (b * b) * flip; n[2] = flip; }
    else if (type == 1) { n[0] = -2 * x / (a * a) * flip; n[1] = -2 * y / (b * b) * flip; n[2] = 2 * z / (c * c) * flip; }
    else { n[0] = 2 * x / (a * a) * flip; n[1] = 2 * y / (b * b) * flip; n[2] = 2 * z / (c * c) * flip; }
    double nrm = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int m = 0; m < 3; m++) n[m] /= nrm;
}

// transfRays (RayTrace.h:90-119): inverse uses the transpose of the
// rotation block and -R^T t (Utils.h:581-600).
void transform_frame(const double *M, cframe *f, bool inverse) {
    for (int i = 0; i < f->size; i++) {
        double p[3] = {f->x[i], f->y[i], f->z[i]}, d[3] = {f->dx[i], f->dy[i], f->dz[i]};
        double po[3], dout[3];
        for (int n = 0; n < 3; n++) {
            if (inverse) {
                double tr = -M[n] * M[3] - M[n + 4] * M[7] - M[n + 8] * M[11];
                po[n] = M[n] * p[0] + M[n + 4] * p[1] + M[n + 8] * p[2] + tr;
                dout[n] = M[n] * d[0] + M[n + 4] * d[1] + M[n + 8] * d[2];
            } else {
                po[n] = M[4 * n] * p[0] + M[4 * n + 1] * p[1] + M[4 * n + 2] * p[2] + M[4 * n + 3];
                dout[n] = M[4 * n] * d[0] + M[4 * n + 1] * d[1] + M[4 * n + 2] * d[2];
            }
        }
        f->x[i] = po[0]; f->y[i] = po[1]; f->z[i] = po[2];
        f->dx[i] = dout[0]; f->dy[i] = dout[1]; f->dz[i] = dout[2];
    }
}

void trace(const reflparams &tp, cframe *in, cframe *out, int nthreads, double tol, double t0,
           int max_iter) {
    int n_tot = in->size;
    std::vector<double> err(n_tot, 0.);
    if (tp.rms > 0) {   // Random.h:124-137: one distribution for all draws
        std::mt19937 gen(tp.rms_seed);
        std::normal_distribution<double> dist(0.0, tp.rms);
        for (int n = 0; n < n_tot; n++) err[n] = dist(gen);
    }
    transform_frame(tp.transf, in, true);
    int flip = tp.flip ? -1 : 1;
    double a = tp.coeffs[0], b = tp.coeffs[1], c = tp.coeffs[2];
    split_run(n_tot, nthreads, [&](int lo, int hi) {
        for (int i = lo; i < hi; i++) {
            double x = in->x[i], y = in->y[i], z = in->z[i];
            double dx = in->dx[i], dy = in->dy[i], dz = in->dz[i];
            double t = t0, t1 = 1e99;
            double chk = std::fabs(t1 - t);
            int it = 0;
            while (chk > tol && (max_iter <= 0 || it < max_iter)) {
                t1 = newton_step(tp.type, t, x, y, z, dx, dy, dz, a, b, c);
                chk = std::fabs(t1 - t);
                t = t1;
                if (max_iter > 0) it++;   // uncapped: no counter, so none to overflow
            }
            bool dead = std::fabs(std::round(dx)) == 0 && std::fabs(std::round(dy)) == 0 &&
                        std::fabs(std::round(dz)) == 0;
            if (dead || std::isnan(t)) {
                out->x[i] = x; out->y[i] = y; out->z[i] = z;
                out->dx[i] = 0; out->dy[i] = 0; out->dz[i] = 0;
                continue;
            }
            double px = x + t * dx, py = y + t * dy, pz = z + t * dz;
            double n[3];
            surface_normal(tp.type, px, py, pz, flip, a, b, c, n);
            double f = 2. * (dx * n[0] + dy * n[1] + dz * n[2]);
            out->dx[i] = dx - n[0] * f;
            out->dy[i] = dy - n[1] * f;
            out->dz[i] = dz - n[2] * f;
            out->x[i] = px + err[i] * n[0];
            out->y[i] = py + err[i] * n[1];
            out->z[i] = pz + err[i] * n[2];
        }
    });
    transform_frame(tp.transf, in, false);
    transform_frame(tp.transf, out, false);
}


// ---- beam initialisation (src/common/BeamInit.h) ----------------------------------

// Grid of the beam plane, as BeamInit.h generates it (transform on, spheric on).
struct OwnedGrid {
    std::vector<double> buf;
    reflcontainer c;
    explicit OwnedGrid(const reflparams &r) {
        int n = r.n_cells[0] * r.n_cells[1];
        buf.assign(7 * (size_t)n, 0.);
        double *b = buf.data();
        c = reflcontainer{n, b, b + n, b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n, b + 6 * n};
        make_grid(r, &c, true, true);
    }
};

// n x v, real n, complex v.
void ncross(const double *n, const cplx *v, cplx *out) {
    out[0] = n[1] * v[2] - n[2] * v[1];
    out[1] = n[2] * v[0] - n[0] * v[2];
    out[2] = n[0] * v[1] - n[1] * v[0];
}

void store6(c2Bundle *b, int field, int i, const cplx *v, double f) {
    double **p = (double **)b + 6 * field;
    for (int a = 0; a < 3; a++) {
        p[a][i] = f * v[a].real();
        p[3 + a][i] = f * v[a].imag();
    }
}

// Astigmatic Gaussian: initGauss, BeamInit.h:288-405 (with its E0-only
// normalisation; the scalar twin below keeps sqrt(2 / (pi wzx wzy))).
cplx astig_gauss(double lam, double w0x, double w0y, double nidx, double dxyz, double x, double y,
                 double z) {
    const cplx j(0., 1.);
    double zRx = M_PI * w0x * w0x * nidx / lam, zRy = M_PI * w0y * w0y * nidx / lam;
    double k = 2 * M_PI / lam;
    double zy = z - dxyz;
    double wzx = w0x * std::sqrt(1 + (z / zRx) * (z / zRx));
    double wzy = w0y * std::sqrt(1 + (zy / zRy) * (zy / zRy));
    double rx = z / (z * z + zRx * zRx), ry = zy / (zy * zy + zRy * zRy);
    double phx = std::atan(z / zRx), phy = std::atan(zy / zRy);
    return std::exp(-(x / wzx) * (x / wzx) - (y / wzy) * (y / wzy) -
                    j * M_PI / lam * (x * x * rx + y * y * ry) - j * k * z + j * (phx - phy) * 0.5);
}

void make_gauss(const GPODict &g, const reflparams &plane, c2Bundle *field, c2Bundle *cur) {
    OwnedGrid G(plane);
    const reflcontainer &c = G.c;
    const cplx zero3[3] = {0., 0., 0.};
    for (int i = 0; i < c.size; i++) {
        cplx f = g.E0 * astig_gauss(g.lam, g.w0x, g.w0y, g.n, g.dxyz, c.x[i], c.y[i], c.z[i]);
        cplx e[3] = {f * g.pol[0], f * g.pol[1], f * g.pol[2]};
        double nv[3] = {c.nx[i], c.ny[i], c.nz[i]};
        cplx m[3];
        ncross(nv, e, m);
        store6(field, 0, i, e, 1.);
        store6(field, 1, i, zero3, 1.);
        store6(cur, 0, i, zero3, 1.);
        store6(cur, 1, i, m, -2.);
    }
}

void make_scalar_gauss(const ScalarGPODict &g, const reflparams &plane, arrC1 *out) {
    OwnedGrid G(plane);
    const reflcontainer &c = G.c;
    for (int i = 0; i < c.size; i++) {
        double z = c.z[i], zy = z - g.dxyz;
        double zRx = M_PI * g.w0x * g.w0x * g.n / g.lam, zRy = M_PI * g.w0y * g.w0y * g.n / g.lam;
        double wzx = g.w0x * std::sqrt(1 + (z / zRx) * (z / zRx));
        double wzy = g.w0y * std::sqrt(1 + (zy / zRy) * (zy / zRy));
        cplx f = g.E0 * std::sqrt(2 / (M_PI * wzx * wzy)) *
                 astig_gauss(g.lam, g.w0x, g.w0y, g.n, g.dxyz, c.x[i], c.y[i], z);
        out->x[i] = f.real();
        out->y[i] = f.imag();
    }
}

// Complex-source-point vector Gaussian: initGaussBeam, BeamInit.h:407-596.
void make_gauss_beam(const vecGPODict &g, const reflparams &plane, c2Bundle *field, c2Bundle *cur) {
    OwnedGrid G(plane);
    const reflcontainer &c = G.c;
    const cplx j(0., 1.);
    const cplx zero3[3] = {0., 0., 0.};
    double k = 2 * M_PI / g.lam;
    double zc = k * g.w0 * g.w0 / 2.0;
    double pref = g.w0 * k * std::sqrt(g.power / (8 * M_PI));
    double F = (g.mode == 1 || g.mode == 2) ? 2. : 1.;
    for (int i = 0; i < c.size; i++) {
        // BeamInit.h:473 calls abs() on a double where only the integer abs is
        // visible: the axial distance is |trunc(z + z0)|, not |z + z0|.
        cplx R[3] = {cplx(c.x[i], 0.), cplx(c.y[i], 0.),
                     cplx(std::fabs(std::trunc(c.z[i] + g.z)), 0.) + j * zc};
        cplx r = std::sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        cplx ex = std::exp(-j * k * r - k * k * g.w0 * g.w0 / 2.0);
        cplx q = 1.0 / (k * r);
        cplx s1 = -j * q - q * q + j * q * q * q;
        cplx s2 = j * q + 3.0 * q * q - 3.0 * j * q * q * q;
        cplx s3 = j * q + q * q;
        cplx e[3], h[3];
        e[0] = pref * ex * (s1 + s2 * (R[0] * R[0] / (r * r)) - s3 * (R[2] / r));
        e[1] = pref * ex * s2 * R[0] * R[1] / (r * r);
        e[2] = pref * ex * (s2 * R[0] * R[2] / (r * r) - s3 * R[0] / r);
        h[0] = pref * ex * s2 * R[0] * R[1] / (r * r);
        h[1] = pref * ex * (s1 + s2 * (R[1] * R[1] / (r * r)) - s3 * (R[2] / r));
        h[2] = pref * ex * (s2 * R[1] * R[2] / (r * r) - s3 * R[1] / r);
        if (!(c.z[i] + g.z >= 0.0))
            for (int a = 0; a < 3; a++) { e[a] = std::conj(e[a]); h[a] = std::conj(h[a]); }
        double nv[3] = {c.nx[i], c.ny[i], c.nz[i]};
        cplx M[3], J[3];
        ncross(nv, e, M);
        ncross(nv, h, J);
        store6(field, 0, i, e, 1.);
        store6(field, 1, i, h, 1.);
        if (g.mode == 0 || g.mode == 2) store6(cur, 0, i, J, F);
        else store6(cur, 0, i, zero3, 1.);
        if (g.mode == 0 || g.mode == 1) store6(cur, 1, i, M, -F);
        else store6(cur, 1, i, zero3, 1.);
    }
}

// calcJM, BeamInit.h:660-844.
void calc_currents(const c2Bundle *field, c2Bundle *cur, const reflparams &plane, int mode) {
    OwnedGrid G(plane);
    const reflcontainer &c = G.c;
    const cplx zero3[3] = {0., 0., 0.};
    double *const *f = (double *const *)field;
    for (int i = 0; i < c.size; i++) {
        cplx e[3], h[3];
        for (int a = 0; a < 3; a++) {
            e[a] = cplx(f[a][i], f[3 + a][i]);
            h[a] = cplx(f[6 + a][i], f[9 + a][i]);
        }
        double nv[3] = {c.nx[i], c.ny[i], c.nz[i]};
        cplx ne[3], nh[3];
        ncross(nv, e, ne);
        ncross(nv, h, nh);
        if (mode == 0) { store6(cur, 1, i, ne, -1.); store6(cur, 0, i, nh, 1.); }
        else if (mode == 1) { store6(cur, 1, i, ne, -2.); store6(cur, 0, i, zero3, 1.); }
        else if (mode == 2) { store6(cur, 0, i, nh, 2.); store6(cur, 1, i, zero3, 1.); }
        else if (mode == 3) { store6(cur, 0, i, ne, 2.); store6(cur, 1, i, zero3, 1.); }
    }
}

// Tubular frame: initFrame, BeamInit.h:159-214 (same running angle sum).
void make_rt_frame(const RTDict &r, cframe *fr) {
    int nTot = 1 + r.nRing * 4 * r.nRays;
    fr->size = nTot;
    fr->x[0] = fr->y[0] = fr->z[0] = 0.;
    fr->dx[0] = fr->dy[0] = 0.;
    fr->dz[0] = 1.;
    double alpha = 0., d_alpha = r.nRays > 0 ? 2 * M_PI / (4 * r.nRays) : 0.;
    int ring = 1;
    for (int i = 1; i < nTot; i++) {
        fr->x[i] = r.x0 * std::cos(alpha) / r.nRing * ring;
        fr->y[i] = r.y0 * std::sin(alpha) / r.nRing * ring;
        fr->z[i] = 0.;
        double rx = r.angy0 * std::sin(alpha) / r.nRing * ring;
        double ry = r.angx0 * std::cos(alpha) / r.nRing * ring;
        double rz = 2 * alpha;
        // (0, 0, 1) rotated by Utils::matRot's Euler angles (Utils.h:613-635)
        double cx = std::cos(rx), cy = std::cos(ry), cz = std::cos(rz);
        double sx = std::sin(rx), sy = std::sin(ry), sz = std::sin(rz);
        fr->dx[i] = cz * sy * cx + sz * sx;
        fr->dy[i] = sz * sy * cx - cz * sx;
        fr->dz[i] = cy * cx;
        alpha += d_alpha;
        if (i == (nTot / r.nRing) * ring) {
            ring++;
            alpha = 0.;
        }
    }
}

}  // namespace

extern "C" {

void oracle_generateGrid(reflparams refl, reflcontainer *container, bool transform, bool spheric) {
    make_grid(refl, container, transform, spheric);
}

void oracle_propagateToGrid_JM(c2Bundle *res, reflparams s, reflparams t, reflcontainer *cs,
                               reflcontainer *ct, c2Bundle *cur, double k, int nt, double eps,
                               double tdir) {
    propagate(JM, res, s, t, cs, ct, cur, k, nt, eps, tdir);
}

void oracle_propagateToGrid_EH(c2Bundle *res, reflparams s, reflparams t, reflcontainer *cs,
                               reflcontainer *ct, c2Bundle *cur, double k, int nt, double eps,
                               double tdir) {
    propagate(EH, res, s, t, cs, ct, cur, k, nt, eps, tdir);
}

void oracle_propagateToGrid_JMEH(c4Bundle *res, reflparams s, reflparams t, reflcontainer *cs,
                                 reflcontainer *ct, c2Bundle *cur, double k, int nt, double eps,
                                 double tdir) {
    propagate(JMEH, res, s, t, cs, ct, cur, k, nt, eps, tdir);
}

void oracle_propagateToGrid_EHP(c2rBundle *res, reflparams s, reflparams t, reflcontainer *cs,
                                reflcontainer *ct, c2Bundle *cur, double k, int nt, double eps,
                                double tdir) {
    propagate(EHP, res, s, t, cs, ct, cur, k, nt, eps, tdir);
}

void oracle_propagateToFarField(c2Bundle *res, reflparams s, reflparams t, reflcontainer *cs,
                                reflcontainer *ct, c2Bundle *cur, double k, int nt, double eps,
                                double tdir) {
    propagate(FF, res, s, t, cs, ct, cur, k, nt, eps, tdir);
}

void oracle_propagateToGrid_scalar(arrC1 *res, reflparams s, reflparams t, reflcontainer *cs,
                                   reflcontainer *ct, arrC1 *field, double k, int nt, double eps,
                                   double tdir) {
    make_grid(s, cs, true, true);
    make_grid(t, ct, true, true);
    Consts K(k, eps);
    int gm = s.gmode, ncx = s.n_cells[0], ncy = s.n_cells[1];
    split_run(ct->size, nt, [&](int lo, int hi) {
        for (int i = lo; i < hi; i++) {
            double pt[3] = {ct->x[i], ct->y[i], ct->z[i]};
            cplx u = scalar_at(cs, field, gm, ncx, ncy, K, k, tdir, pt);
            res->x[i] = u.real();
            res->y[i] = u.imag();
        }
    });
}

// Same semantics as the reference (no iteration cap) when max_iter <= 0.
void oracle_propagateRays(reflparams ctp, cframe *fr_in, cframe *fr_out, int nt, double tol,
                          double t0, int max_iter) {
    trace(ctp, fr_in, fr_out, nt, tol, t0, max_iter);
}

// A vector PO mode (JM, EH, JMEH, EHP, FF) on grids the caller generated
// (and possibly displaced): no grid generation inside.
void oracle_propagate_on_grids(int mode, void *res, int gmode, int ncx, int ncy, reflcontainer *cs,
                               reflcontainer *ct, c2Bundle *cur, double k, int nt, double eps,
                               double tdir) {
    propagate_on(mode, res, gmode, ncx, ncy, cs, ct, cur, k, nt, eps, tdir);
}

// Far field on a strided subset of targets (bench CPU baseline): targets
// idx[0..n) of an already generated target grid, sources already generated.
void oracle_farfield_subset(c2Bundle *res, reflparams s, reflcontainer *cs, reflcontainer *ct,
                            c2Bundle *cur, const int *idx, int n, double k, int nt, double eps,
                            double tdir) {
    Consts K(k, eps);
    Src S{cs, cur, s.gmode, s.n_cells[0], s.n_cells[1]};
    split_run(n, nt, [&](int lo, int hi) {
        for (int m = lo; m < hi; m++) {
            int i = idx[m];
            double th = ct->x[i], ph = ct->y[i];
            double rh[3] = {std::cos(ph) * std::sin(th), std::sin(ph) * std::sin(th), std::cos(th)};
            cplx E[3], H[3];
            far_field_at(S, K, k, tdir, rh, E, H);
            double *er[3] = {res->r1x, res->r1y, res->r1z}, *ei[3] = {res->i1x, res->i1y, res->i1z};
            double *hr[3] = {res->r2x, res->r2y, res->r2z}, *hi_[3] = {res->i2x, res->i2y, res->i2z};
            put_c3(er, ei, m, E); put_c3(hr, hi_, m, H);
        }
    });
}

void oracle_makeGauss(GPODict g, reflparams plane, c2Bundle *field, c2Bundle *cur) {
    make_gauss(g, plane, field, cur);
}
void oracle_makeGaussBeam(vecGPODict g, reflparams plane, c2Bundle *field, c2Bundle *cur) {
    make_gauss_beam(g, plane, field, cur);
}
void oracle_makeScalarGauss(ScalarGPODict g, reflparams plane, arrC1 *field) {
    make_scalar_gauss(g, plane, field);
}
void oracle_calcCurrents(c2Bundle *field, c2Bundle *cur, reflparams plane, int mode) {
    calc_currents(field, cur, plane, mode);
}
void oracle_makeRTframe(RTDict r, cframe *fr) { make_rt_frame(r, fr); }

}  // extern "C"
