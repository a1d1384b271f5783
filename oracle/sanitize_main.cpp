// sanitize_main.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Runs every entry point of the CPU restatement (pypo_oracle.cpp) on small
// cases, for the sanitizer builds of `make -C oracle sanitize`:
//   _san/oracle_asan   AddressSanitizer + UndefinedBehaviorSanitizer
//   _san/oracle_tsan   ThreadSanitizer
// Both are built with -fno-sanitize-recover, so any report ends the run with
// a non-zero status.
//
// The threaded entry points split the targets (or rays) over std::threads in
// contiguous blocks, as Propagation.h:1356-1377 and RayTrace.h:283-298 do.
// Every case therefore runs with one thread and with three (an uneven
// split), and the two results must agree bit for bit, because each target is
// computed on its own.  The outputs of the propagations must be finite.
//
// Prints one line per case and "ALL OK" at the end; exit status 0 on success.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../include/pypo_hip.h"

extern "C" {
void oracle_generateGrid(reflparams refl, reflcontainer *container, bool transform, bool spheric);
void oracle_propagateToGrid_JM(c2Bundle *, reflparams, reflparams, reflcontainer *, reflcontainer *,
                               c2Bundle *, double, int, double, double);
void oracle_propagateToGrid_EH(c2Bundle *, reflparams, reflparams, reflcontainer *, reflcontainer *,
                               c2Bundle *, double, int, double, double);
void oracle_propagateToGrid_JMEH(c4Bundle *, reflparams, reflparams, reflcontainer *, reflcontainer *,
                                 c2Bundle *, double, int, double, double);
void oracle_propagateToGrid_EHP(c2rBundle *, reflparams, reflparams, reflcontainer *, reflcontainer *,
                                c2Bundle *, double, int, double, double);
void oracle_propagateToFarField(c2Bundle *, reflparams, reflparams, reflcontainer *, reflcontainer *,
                                c2Bundle *, double, int, double, double);
void oracle_propagateToGrid_scalar(arrC1 *, reflparams, reflparams, reflcontainer *, reflcontainer *,
                                   arrC1 *, double, int, double, double);
void oracle_propagateRays(reflparams, cframe *, cframe *, int, double, double, int);
void oracle_propagate_on_grids(int mode, void *res, int gmode, int ncx, int ncy, reflcontainer *cs,
                               reflcontainer *ct, c2Bundle *cur, double k, int nt, double eps,
                               double tdir);
void oracle_farfield_subset(c2Bundle *res, reflparams s, reflcontainer *cs, reflcontainer *ct,
                            c2Bundle *cur, const int *idx, int n, double k, int nt, double eps,
                            double tdir);
void oracle_makeGauss(GPODict g, reflparams plane, c2Bundle *field, c2Bundle *cur);
void oracle_makeGaussBeam(vecGPODict g, reflparams plane, c2Bundle *field, c2Bundle *cur);
void oracle_makeScalarGauss(ScalarGPODict g, reflparams plane, arrC1 *field);
void oracle_calcCurrents(c2Bundle *field, c2Bundle *cur, reflparams plane, int mode);
void oracle_makeRTframe(RTDict r, cframe *fr);
}

namespace {

static_assert(sizeof(arrC1) == 2 * sizeof(double *), "arrC1 is a list of pointers");
static_assert(sizeof(c2Bundle) == 12 * sizeof(double *), "c2Bundle is a list of pointers");
static_assert(sizeof(c2rBundle) == 15 * sizeof(double *), "c2rBundle is a list of pointers");
static_assert(sizeof(c4Bundle) == 24 * sizeof(double *), "c4Bundle is a list of pointers");

int g_failures = 0;

void fail(const std::string &what) {
    std::printf("FAIL %s\n", what.c_str());
    g_failures++;
}

// Storage for a bundle struct B made of `n_arr` arrays of `n` doubles each,
// sized exactly (so that ASan sees any read or write past an array).
template <typename B>
struct Bundle {
    static constexpr int n_arr = (int)(sizeof(B) / sizeof(double *));
    std::vector<std::vector<double>> arr;
    B b;
    explicit Bundle(int n, double fill = 0.0) : arr(n_arr, std::vector<double>(n, fill)) {
        double *p[n_arr];
        for (int i = 0; i < n_arr; i++) p[i] = arr[i].data();
        std::memcpy(&b, p, sizeof(B));
    }
    bool same(const Bundle &o) const { return arr == o.arr; }
    bool finite() const {
        for (const auto &a : arr)
            for (double v : a)
                if (!std::isfinite(v)) return false;
        return true;
    }
};

// A realised grid with exactly `n` points.
struct Grid {
    std::vector<double> x, y, z, nx, ny, nz, area;
    reflcontainer c;
    explicit Grid(int n) : x(n), y(n), z(n), nx(n), ny(n), nz(n), area(n) {
        c = {n, x.data(), y.data(), z.data(), nx.data(), ny.data(), nz.data(), area.data()};
    }
};

struct Frame {
    std::vector<double> x, y, z, dx, dy, dz;
    cframe c;
    explicit Frame(int n) : x(n), y(n), z(n), dx(n), dy(n), dz(n) {
        c = {n, x.data(), y.data(), z.data(), dx.data(), dy.data(), dz.data()};
    }
    bool same(const Frame &o) const {
        return x == o.x && y == o.y && z == o.z && dx == o.dx && dy == o.dy && dz == o.dz;
    }
};

// Surface parameters plus the arrays reflparams points into.
struct Surface {
    double coeffs[3], lxu[2], lyv[2], gcenter[2], transf[16];
    int n_cells[2];
    reflparams p;
    Surface(int type, int gmode, double a, double b, double c, double x0, double x1, double y0,
            double y1, int nx, int ny) {
        coeffs[0] = a; coeffs[1] = b; coeffs[2] = c;
        lxu[0] = x0; lxu[1] = x1; lyv[0] = y0; lyv[1] = y1;
        gcenter[0] = gcenter[1] = 0.0;
        for (int i = 0; i < 16; i++) transf[i] = (i % 5 == 0) ? 1.0 : 0.0;
        n_cells[0] = nx; n_cells[1] = ny;
        p = {coeffs, lxu, lyv, n_cells, false, gmode, gcenter, 0.0, 0.0, type, transf, 0.0, 0u};
    }
    int size() const { return n_cells[0] * n_cells[1]; }
    // Row-major 4x4: a rotation about x by `deg` degrees, then a translation.
    void place(double deg, double tx, double ty, double tz) {
        double a = deg * M_PI / 180.0, c = std::cos(a), s = std::sin(a);
        double m[16] = {1, 0, 0, tx, 0, c, -s, ty, 0, s, c, tz, 0, 0, 0, 1};
        std::memcpy(transf, m, sizeof m);
    }
};

// Random but reproducible surface currents: sums of a few sinusoids.
Bundle<c2Bundle> currents(int n, int seed) {
    Bundle<c2Bundle> cur(n);
    for (int a = 0; a < Bundle<c2Bundle>::n_arr; a++)
        for (int i = 0; i < n; i++)
            cur.arr[a][i] = std::sin(0.37 * i * (a + 1) + seed) + 0.5 * std::cos(0.11 * i + a);
    return cur;
}

typedef void (*VecFn)(void *, reflparams, reflparams, reflcontainer *, reflcontainer *, c2Bundle *,
                      double, int, double, double);

template <typename B>
void po_case(const char *name, VecFn fn, Surface &src, Surface &tgt, double k) {
    Grid gs(src.size()), gt1(tgt.size()), gt3(tgt.size());
    Grid gs3(src.size());
    Bundle<c2Bundle> cur = currents(src.size(), 3);
    Bundle<B> r1(tgt.size()), r3(tgt.size());
    fn(&r1.b, src.p, tgt.p, &gs.c, &gt1.c, &cur.b, k, 1, 1.0, -1.0);
    fn(&r3.b, src.p, tgt.p, &gs3.c, &gt3.c, &cur.b, k, 3, 1.0, -1.0);
    if (!r1.same(r3)) fail(std::string(name) + ": 1 and 3 threads differ");
    if (!r1.finite()) fail(std::string(name) + ": non-finite output");
    std::printf("%-28s %d sources -> %d targets\n", name, src.size(), tgt.size());
}

void scalar_case(Surface &src, Surface &tgt, double k) {
    Grid gs(src.size()), gt(tgt.size()), gs3(src.size()), gt3(tgt.size());
    Bundle<arrC1> f(src.size());
    for (int i = 0; i < src.size(); i++) {
        f.arr[0][i] = std::cos(0.2 * i);
        f.arr[1][i] = std::sin(0.3 * i);
    }
    Bundle<arrC1> r1(tgt.size()), r3(tgt.size());
    oracle_propagateToGrid_scalar(&r1.b, src.p, tgt.p, &gs.c, &gt.c, &f.b, k, 1, 1.0, -1.0);
    oracle_propagateToGrid_scalar(&r3.b, src.p, tgt.p, &gs3.c, &gt3.c, &f.b, k, 3, 1.0, -1.0);
    if (!r1.same(r3)) fail("scalar: 1 and 3 threads differ");
    if (!r1.finite()) fail("scalar: non-finite output");
    std::printf("%-28s %d sources -> %d targets\n", "scalar", src.size(), tgt.size());
}

void on_grids_case(Surface &src, Surface &tgt, double k) {
    Grid gs(src.size()), gt(tgt.size());
    oracle_generateGrid(src.p, &gs.c, true, true);
    oracle_generateGrid(tgt.p, &gt.c, true, true);
    Bundle<c2Bundle> cur = currents(src.size(), 5);
    Bundle<c4Bundle> r1(tgt.size()), r3(tgt.size());
    const int JMEH = 2;   // the mode order of the propagation enum: JM, EH, JMEH, EHP, FF
    oracle_propagate_on_grids(JMEH, &r1.b, src.p.gmode, src.n_cells[0], src.n_cells[1], &gs.c, &gt.c,
                              &cur.b, k, 1, 1.0, -1.0);
    oracle_propagate_on_grids(JMEH, &r3.b, src.p.gmode, src.n_cells[0], src.n_cells[1], &gs.c, &gt.c,
                              &cur.b, k, 3, 1.0, -1.0);
    if (!r1.same(r3)) fail("on_grids JMEH: 1 and 3 threads differ");
    if (!r1.finite()) fail("on_grids JMEH: non-finite output");
    std::printf("%-28s %d sources -> %d targets\n", "on_grids JMEH", src.size(), tgt.size());
}

void ff_subset_case(Surface &src, Surface &ff, double k) {
    Grid gs(src.size()), gt(ff.size());
    oracle_generateGrid(src.p, &gs.c, true, true);
    oracle_generateGrid(ff.p, &gt.c, false, true);
    Bundle<c2Bundle> cur = currents(src.size(), 9);
    std::vector<int> idx;
    for (int i = 0; i < ff.size(); i += 3) idx.push_back(i);
    const int n = (int)idx.size();
    Bundle<c2Bundle> r1(n), r3(n);
    oracle_farfield_subset(&r1.b, src.p, &gs.c, &gt.c, &cur.b, idx.data(), n, k, 1, 1.0, -1.0);
    oracle_farfield_subset(&r3.b, src.p, &gs.c, &gt.c, &cur.b, idx.data(), n, k, 3, 1.0, -1.0);
    if (!r1.same(r3)) fail("farfield_subset: 1 and 3 threads differ");
    if (!r1.finite()) fail("farfield_subset: non-finite output");
    std::printf("%-28s %d sources -> %d targets\n", "farfield_subset", src.size(), n);
}

void beam_cases(Surface &plane) {
    const int n = plane.size();
    double pol[3] = {1.0, 0.0, 0.0};
    GPODict g = {1.0, 5.0, 4.0, 1.0, 1.0, 0.0, pol};
    Bundle<c2Bundle> field(n), cur(n);
    oracle_makeGauss(g, plane.p, &field.b, &cur.b);
    if (!field.finite() || !cur.finite()) fail("makeGauss: non-finite output");
    for (int mode = 0; mode < 3; mode++) {
        vecGPODict v = {1.0, 5.0, 10.0, 1.0, 1.0, mode};
        Bundle<c2Bundle> f(n), c(n);
        oracle_makeGaussBeam(v, plane.p, &f.b, &c.b);
        if (!f.finite() || !c.finite()) fail("makeGaussBeam: non-finite output");
    }
    ScalarGPODict sg = {1.0, 5.0, 4.0, 1.0, 1.0, 0.0};
    Bundle<arrC1> sf(n);
    oracle_makeScalarGauss(sg, plane.p, &sf.b);
    if (!sf.finite()) fail("makeScalarGauss: non-finite output");
    for (int mode = 0; mode < 4; mode++) {
        Bundle<c2Bundle> c(n);
        oracle_calcCurrents(&field.b, &c.b, plane.p, mode);
        if (!c.finite()) fail("calcCurrents: non-finite output");
    }
    std::printf("%-28s %d points\n", "beams and currents", n);
}

void rt_case(const char *name, Surface &tgt, const Frame &in, double t0, int max_iter) {
    Frame o1(in.c.size), o3(in.c.size);
    Frame a = in, b = in;
    a.c = {in.c.size, a.x.data(), a.y.data(), a.z.data(), a.dx.data(), a.dy.data(), a.dz.data()};
    b.c = {in.c.size, b.x.data(), b.y.data(), b.z.data(), b.dx.data(), b.dy.data(), b.dz.data()};
    oracle_propagateRays(tgt.p, &a.c, &o1.c, 1, 1e-12, t0, max_iter);
    oracle_propagateRays(tgt.p, &b.c, &o3.c, 3, 1e-12, t0, max_iter);
    if (!o1.same(o3)) fail(std::string(name) + ": 1 and 3 threads differ");
    std::printf("%-28s %d rays\n", name, in.c.size);
}

}  // namespace

int main() {
    const double k = 2 * M_PI / 1.0;

    // Sources: a plane in xy and a paraboloid gridded in uv (elliptical rim, rotated).
    Surface plane(3, 0, 0, 0, 0, -5, 5, -5, 5, 15, 17);
    Surface dish(0, 1, 20, 20, -1, 0, 12, 0, 360, 11, 13);
    dish.p.ecc_uv = 0.3;
    dish.p.rot_uv = 20.0;
    // Targets: a tilted xy plane 40 mm away, a rough hyperboloid, an ellipsoid.
    Surface screen(3, 0, 0, 0, 0, -6, 6, -6, 6, 9, 7);
    screen.place(10.0, 0.0, 1.0, 40.0);
    Surface hyp(1, 0, 8, 8, 12, -4, 4, -4, 4, 8, 9);
    hyp.p.rms = 0.01;
    hyp.p.rms_seed = 7;
    hyp.place(0.0, 0.0, 0.0, 60.0);
    Surface ell(2, 0, 30, 30, 40, -6, 6, -6, 6, 7, 9);
    ell.place(5.0, 0.0, 0.0, 90.0);
    // A far-field grid in azimuth / elevation (degrees).
    Surface ff(3, 2, 0, 0, 0, -2, 2, -2, 2, 9, 11);

    po_case<c2Bundle>("JM plane -> screen", (VecFn)oracle_propagateToGrid_JM, plane, screen, k);
    po_case<c2Bundle>("EH dish(uv) -> hyperboloid", (VecFn)oracle_propagateToGrid_EH, dish, hyp, k);
    po_case<c4Bundle>("JMEH plane -> ellipsoid", (VecFn)oracle_propagateToGrid_JMEH, plane, ell, k);
    po_case<c2rBundle>("EHP dish(uv) -> screen", (VecFn)oracle_propagateToGrid_EHP, dish, screen, k);
    po_case<c2Bundle>("FF dish(uv) -> AoE", (VecFn)oracle_propagateToFarField, dish, ff, k);
    scalar_case(plane, screen, k);
    on_grids_case(plane, hyp, k);
    ff_subset_case(dish, ff, k);
    beam_cases(plane);

    {   // generated grids with and without the transform, uv and AoE planes
        Surface uvplane(3, 1, 0, 0, 0, 0, 5, 0, 360, 6, 12);
        Surface aoe(3, 2, 0, 0, 0, -3, 3, -1, 1, 5, 4);
        for (Surface *s : {&uvplane, &aoe, &hyp, &ell, &dish}) {
            Grid g(s->size());
            oracle_generateGrid(s->p, &g.c, true, true);
            oracle_generateGrid(s->p, &g.c, false, false);
        }
        std::printf("%-28s 5 surfaces\n", "generateGrid");
    }

    {   // rays: a tubular frame onto a paraboloid, then onto a rough one
        RTDict rd = {8, 3, 0.02, 0.02, 10.0, 10.0};   // angles in radians
        Frame fr(1 + 4 * rd.nRays * rd.nRing);
        oracle_makeRTframe(rd, &fr.c);
        Surface par(0, 0, 20, 20, -1, -30, 30, -30, 30, 3, 3);
        par.place(0.0, 0.0, 0.0, 100.0);
        rt_case("RT tubular -> paraboloid", par, fr, 100.0, 0);
        Surface rough = par;
        rough.p.rms = 0.05;
        rough.p.rms_seed = 11;
        rough.p.coeffs = rough.coeffs;
        rough.p.lxu = rough.lxu;
        rough.p.lyv = rough.lyv;
        rough.p.n_cells = rough.n_cells;
        rough.p.gcenter = rough.gcenter;
        rough.p.transf = rough.transf;
        rt_case("RT tubular -> rough", rough, fr, 100.0, 50);
        // half the rays point away from the surface: the iteration cap ends them
        Frame mixed = fr;
        for (int i = 0; i < mixed.c.size; i += 2) mixed.dz[i] = -mixed.dz[i];
        rt_case("RT with missing rays", par, mixed, 100.0, 50);
    }

    if (g_failures) {
        std::printf("%d FAILURES\n", g_failures);
        return 1;
    }
    std::printf("ALL OK\n");
    return 0;
}
