// rt.hip -- Newton ray-quadric tracer, FP64, one ray per lane.
//
// Restates src/CPU/RayTrace.h:139-305 and src/include/RTRefls.h:52-297 for
// the GPU: each lane moves its ray into the surface rest frame (inverse of
// the 4x4 transf, Utils.h:581-600), iterates t <- t - f(t)/f'(t) until
// |dt| <= tol (the reference has no iteration cap; here at most max_iter
// steps), applies the miss rule (all direction components round to zero, or
// t is NaN: position copied, direction zeroed), otherwise the hit point,
// unit normal (flip-aware except for planes), specular reflection
// d' = d - 2(d.n)n and the normal-direction surface error, and moves the
// result back to the world frame.  Several surfaces are traced in sequence
// in one launch, with every intermediate frame written out, so a Cassegrain
// pair reads its 48 input bytes per ray once.  HBM-bound: 48 B read plus
// 48 B written per ray and surface.

#include "common.hpp"

namespace pypo {
namespace {

constexpr int MAX_SURF = 4;

struct Surf {
    double M[16];
    double a, b, c;
    double ia2, ib2, ic2;   // 1/a^2, 1/b^2, 1/c^2 (computed once on the host)
    int type, flip;
    const double *errs;     // device, or nullptr when rms == 0
};

struct SurfSet {
    Surf s[MAX_SURF];
    int n;
};

struct FramePtrs {
    double *x, *y, *z, *dx, *dy, *dz;
};

struct FrameSet {
    FramePtrs f[MAX_SURF + 1];
};

// 1/d from v_rcp_f64 and two Newton-Raphson steps (about 1 ulp); inf/NaN
// inputs give NaN, which the miss rule treats like the reference's inf/NaN.
__device__ __forceinline__ double rcp_nr(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    return r;
}

// 1/sqrt(s) from v_rsq_f64 and two Newton-Raphson steps (about 1 ulp);
// negative s gives NaN like sqrt.
__device__ __forceinline__ double rsq_nr(double s) {
    double y = __builtin_amdgcn_rsq(s);
    const double h = 0.5 * s;
    y = y * __builtin_fma(-h * y, y, 1.5);
    y = y * __builtin_fma(-h * y, y, 1.5);
    return y;
}

// One Newton update t <- t - f(t)/f'(t) for the quadric (RTRefls.h gp 96,
// gh 120, ge 145, gpl 170).  common1/common2 (RTRefls.h:52-94) are
// evaluated once per step with the reciprocal squared semi-axes.  f(t) is
// evaluated to about an ulp (its zero is the converged t); the step
// f/f' uses a refined reciprocal instead of an IEEE division, and for the
// hyperboloid and ellipsoid one refined 1/sqrt gives both sqrt(.) = s/sqrt(s)
// and the 1/(2 sqrt(.)) of the derivative.  These change t only at rounding
// level against the reference's divisions and sqrt.
__device__ __forceinline__ double newton(const Surf &s, double t, double x, double y, double z,
                                         double dx, double dy, double dz) {
    const double X = x + t * dx, Y = y + t * dy;
    if (s.type == 3) return t - (z + t * dz) * rcp_nr(dz);
    const double c1 = X * X * s.ia2 + Y * Y * s.ib2;
    const double c2 = 2.0 * (X * dx * s.ia2 + Y * dy * s.ib2);
    if (s.type == 0) return t - (z + t * dz - c1) * rcp_nr(dz - c2);
    if (s.type == 1) {
        const double q = c1 + 1, iy = rsq_nr(q), r = q * iy;
        return t - (z + t * dz - s.c * r) * rcp_nr(dz - 0.5 * s.c * iy * c2);
    }
    const double q = 1 - c1, iy = rsq_nr(q), r = q * iy;
    return t - (z + t * dz - s.c * r) * rcp_nr(dz + 0.5 * s.c * iy * c2);
}

__global__ __launch_bounds__(256) void k_trace(SurfSet S, FrameSet F, int n, double tol, double t0,
                                               int max_iter) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double wx = F.f[0].x[i], wy = F.f[0].y[i], wz = F.f[0].z[i];
    double wdx = F.f[0].dx[i], wdy = F.f[0].dy[i], wdz = F.f[0].dz[i];
    for (int k = 0; k < S.n; k++) {
        const Surf &s = S.s[k];
        const double *M = s.M;
        // world -> rest frame
        double x, y, z, dx, dy, dz;
        {
            double tr0 = -M[0] * M[3] - M[4] * M[7] - M[8] * M[11];
            double tr1 = -M[1] * M[3] - M[5] * M[7] - M[9] * M[11];
            double tr2 = -M[2] * M[3] - M[6] * M[7] - M[10] * M[11];
            x = M[0] * wx + M[4] * wy + M[8] * wz + tr0;
            y = M[1] * wx + M[5] * wy + M[9] * wz + tr1;
            z = M[2] * wx + M[6] * wy + M[10] * wz + tr2;
            dx = M[0] * wdx + M[4] * wdy + M[8] * wdz;
            dy = M[1] * wdx + M[5] * wdy + M[9] * wdz;
            dz = M[2] * wdx + M[6] * wdy + M[10] * wdz;
        }
        double t = t0, t1 = 1e99;
        double chk = fabs(t1 - t);
        int it = 0;
        while (chk > tol && it < max_iter) {
            t1 = newton(s, t, x, y, z, dx, dy, dz);
            chk = fabs(t1 - t);
            t = t1;
            ++it;
        }
        double ox, oy, oz, odx, ody, odz;
        // RayTrace.h:213 tests round(d) == 0 per component; round() goes half
        // away from zero, so that is |d| < 0.5.
        bool dead = fabs(dx) < 0.5 && fabs(dy) < 0.5 && fabs(dz) < 0.5;
        if (dead || isnan(t)) {
            ox = x; oy = y; oz = z;
            odx = 0; ody = 0; odz = 0;
        } else {
            double px = x + t * dx, py = y + t * dy, pz = z + t * dz;
            double nx, ny, nz;
            if (s.type == 3) {
                nx = 0; ny = 0; nz = 1;
            } else {
                double fl = (double)s.flip;
                if (s.type == 0) {
                    nx = -2 * px * s.ia2 * fl; ny = -2 * py * s.ib2 * fl; nz = fl;
                } else if (s.type == 1) {
                    nx = -2 * px * s.ia2 * fl; ny = -2 * py * s.ib2 * fl;
                    nz = 2 * pz * s.ic2 * fl;
                } else {
                    nx = 2 * px * s.ia2 * fl; ny = 2 * py * s.ib2 * fl;
                    nz = 2 * pz * s.ic2 * fl;
                }
                const double inrm = rsq_nr(nx * nx + ny * ny + nz * nz);
                nx *= inrm; ny *= inrm; nz *= inrm;
            }
            double f = 2. * (dx * nx + dy * ny + dz * nz);
            odx = dx - nx * f; ody = dy - ny * f; odz = dz - nz * f;
            double e = s.errs ? s.errs[i] : 0.0;
            ox = px + e * nx; oy = py + e * ny; oz = pz + e * nz;
        }
        // rest -> world frame
        wx = M[0] * ox + M[1] * oy + M[2] * oz + M[3];
        wy = M[4] * ox + M[5] * oy + M[6] * oz + M[7];
        wz = M[8] * ox + M[9] * oy + M[10] * oz + M[11];
        wdx = M[0] * odx + M[1] * ody + M[2] * odz;
        wdy = M[4] * odx + M[5] * ody + M[6] * odz;
        wdz = M[8] * odx + M[9] * ody + M[10] * odz;
        const FramePtrs &o = F.f[k + 1];
        o.x[i] = wx; o.y[i] = wy; o.z[i] = wz;
        o.dx[i] = wdx; o.dy[i] = wdy; o.dz[i] = wdz;
    }
}

}  // namespace

int launch_trace_chain(const reflparams *surfs, const double *const *d_errs, int n_surf,
                       const cframe *frames, double tol, double t0, int max_iter, hipStream_t st,
                       bool time_it) {
    PYPO_REQUIRE(n_surf >= 1 && n_surf <= MAX_SURF, ERR_ARG, "1 to 4 surfaces per trace call");
    int n = frames[0].size;
    PYPO_REQUIRE(n >= 0, ERR_ARG, "negative frame size");
    if (n == 0) return 0;
    PYPO_REQUIRE(max_iter >= 1, ERR_ARG, "max_iter must be >= 1");
    SurfSet S{};
    FrameSet F{};
    S.n = n_surf;
    for (int k = 0; k < n_surf; k++) {
        const reflparams &r = surfs[k];
        PYPO_REQUIRE(r.coeffs && r.transf, ERR_ARG, "reflparams has a NULL coeffs or transf");
        PYPO_REQUIRE(r.type >= 0 && r.type <= 3, ERR_ARG, "unknown surface type");
        for (int q = 0; q < 16; q++) S.s[k].M[q] = r.transf[q];
        S.s[k].a = r.coeffs[0]; S.s[k].b = r.coeffs[1]; S.s[k].c = r.coeffs[2];
        S.s[k].ia2 = 1.0 / (r.coeffs[0] * r.coeffs[0]);
        S.s[k].ib2 = 1.0 / (r.coeffs[1] * r.coeffs[1]);
        S.s[k].ic2 = 1.0 / (r.coeffs[2] * r.coeffs[2]);
        S.s[k].type = r.type;
        S.s[k].flip = r.flip ? -1 : 1;
        S.s[k].errs = d_errs ? d_errs[k] : nullptr;
    }
    for (int k = 0; k <= n_surf; k++) {
        const cframe &c = frames[k];
        PYPO_REQUIRE(c.size == n, ERR_ARG, "all frames of a trace must have the same size");
        PYPO_REQUIRE(c.x && c.y && c.z && c.dx && c.dy && c.dz, ERR_ARG, "frame has a NULL array");
        F.f[k] = FramePtrs{c.x, c.y, c.z, c.dx, c.dy, c.dz};
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (time_it) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, st);
    }
    hipLaunchKernelGGL(k_trace, dim3((n + 255) / 256), dim3(256), 0, st, S, F, n, tol, t0, max_iter);
    PYPO_CHECK_LAUNCH();
    if (time_it) {
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) == hipSuccess) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            record_kernel_ms(ms);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return 0;
}

}  // namespace pypo
