// The RK-stage update around poi_uy (the reference's `loop`,
// main_par.f:1762-1785) on the GPU: filter_trid, bzone, outuy, poi_uy,
// poi_ux, poi_uz, wx_wall, wz_wall, bzonew, each also as its own entry
// point.  Host side: the x-line coefficient tables (the pentadiagonal LU of
// lhs5 per mode as ludecomp forms it, the integration recurrence, the
// filter's Thomas factors), the work fields, the launch sequence.
#include <cmath>
#include <cstring>
#include <vector>

#include "tbl_ctx.h"
#include "tbl_kernels.h"
#include "tbl_xline.h"

using namespace tbl;

namespace {

#define TRY(x)                         \
    do {                               \
        int rc_ = (x);                 \
        if (rc_ != TBL_OK) return rc_; \
    } while (0)

int fail(const char *msg)
{
    tbl_set_last_error(msg, __FILE__, __LINE__);
    return TBL_ERR_ARG;
}

template <typename T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    TBL_HIP(hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)));
    TBL_HIP(hipMemset(*p, 0, n * sizeof(T)));
    return TBL_OK;
}

// lhs5 (loop_par.f:258-299): row i (1-based) holds the coefficients of
// columns i-2 .. i+2 in a[5 (i-1) .. 5 (i-1) + 4]
void lhs5(int n, double k2b2, std::vector<double> &a)
{
    a.assign(5 * static_cast<size_t>(n), 0.0);
    auto A = [&](int i, int c) -> double & { return a[5 * (i - 1) + (c - 1)]; };
    A(1, 3) = 1.0;
    A(2, 2) = 1.0 + k2b2 / 12.0;
    A(2, 3) = -2.0 + k2b2 / 1.2;
    A(2, 4) = 1.0 + k2b2 / 12.0;
    for (int i = 3; i <= n - 2; ++i) {
        A(i, 1) = 3.0;
        A(i, 2) = 48.0 + 8.0 * k2b2;
        A(i, 3) = -102.0 + 44.0 * k2b2;
        A(i, 4) = 48.0 + 8.0 * k2b2;
        A(i, 5) = 3.0;
    }
    A(n - 1, 2) = 1.0 + k2b2 / 12.0;
    A(n - 1, 3) = -2.0 + k2b2 / 1.2;
    A(n - 1, 4) = 1.0 + k2b2 / 12.0;
    A(n, 3) = 1.0;
}

// Doolittle factors of the pentadiagonal matrix as ludecomp forms them
// (penta_par.f:6-45): columns 1-2 the multipliers of rows i-2, i-1,
// columns 3-5 the upper band; rows n-1 and n with the two subtractions in
// sequence, as the reference writes them out.
void ludecomp5(int n, const std::vector<double> &a, std::vector<double> &lu)
{
    lu.assign(5 * static_cast<size_t>(n), 0.0);
    auto A = [&](int i, int c) { return a[5 * (i - 1) + (c - 1)]; };
    auto U = [&](int i, int c) -> double & { return lu[5 * (i - 1) + (c - 1)]; };
    for (int i = 1; i <= n - 2; ++i) {
        double s3 = 0.0, s4 = 0.0;
        if (i >= 2) s3 = s3 + U(i, 2) * U(i - 1, 4);
        if (i >= 3) {
            s3 = s3 + U(i, 1) * U(i - 2, 5);
            s4 = s4 + U(i, 2) * U(i - 1, 5);
        }
        U(i, 3) = A(i, 3) - s3;
        U(i, 4) = A(i, 4) - s4;
        U(i, 5) = A(i, 5);
        const double s = i >= 2 ? U(i + 1, 1) * U(i - 1, 4) : 0.0;
        U(i + 1, 2) = (A(i + 1, 2) - s) / U(i, 3);
        U(i + 2, 1) = A(i + 2, 1) / U(i, 3);
    }
    int i = n - 1;
    U(i, 3) = A(i, 3) - U(i, 2) * U(i - 1, 4) - U(i, 1) * U(i - 2, 5);
    U(i, 4) = A(i, 4) - U(i, 2) * U(i - 1, 5);
    U(i + 1, 2) = (A(i + 1, 2) - U(i + 1, 1) * U(i - 1, 4)) / U(i, 3);
    i = n;
    U(i, 3) = A(i, 3) - U(i, 2) * U(i - 1, 4) - U(i, 1) * U(i - 2, 5);
}

// filter constants (par.for:116-120)
constexpr double kAlphaF = 0.48;
constexpr double kAf = (11.0 + 10.0 * kAlphaF) / 16.0;
constexpr double kBf = (15.0 + 34.0 * kAlphaF) / 64.0;
constexpr double kCf = (-3.0 + 6.0 * kAlphaF) / 32.0;
constexpr double kDf = (1.0 - 2.0 * kAlphaF) / 64.0;

int check_loop_grid(const tbl_ctx *c)
{
    if (c->g.nx < 8 || c->g.ny < 8) return fail("the loop routines need nx >= 8 and ny >= 8");
    if (c->g.nx > 64 * kXMaxR) return fail("x-lines longer than 4096 points are not supported");
    if (c->K > 1 && c->g.beta == 0.0) return fail("beta = 0 with more than one mode");
    return TBL_OK;
}

int loop_setup(tbl_ctx *c)
{
    if (c->d_xpenta) return TBL_OK;
    TRY(check_loop_grid(c));
    const int nx = c->g.nx, ny = c->g.ny, K = c->K;
    const int R = (nx + 63) / 64, NP = 64 * R;
    auto seg = [R](int i) { return (i % R) * 64 + i / R; };   // row i -> [r][lane]
    const double dxx = c->g.dx * c->g.dx, beta = c->g.beta;
    // pentadiagonal LU per mode k >= 2 (1-based): l2, l1, 1/u0, u1/u0, u2/u0
    std::vector<double> tp(static_cast<size_t>(K > 1 ? K - 1 : 1) * 5 * NP, 0.0);
    std::vector<double> a, lu;
    for (int k = 1; k < K; ++k) {
        const double k2b2 = dxx * (-static_cast<double>(k) * static_cast<double>(k) * beta * beta);
        lhs5(nx, k2b2, a);
        ludecomp5(nx, a, lu);
        double *T = tp.data() + static_cast<size_t>(k - 1) * 5 * NP;
        for (int i = 0; i < nx; ++i) {
            const int I = i + 1;
            auto U = [&](int r, int col) { return lu[5 * (r - 1) + (col - 1)]; };
            const double d = 1.0 / U(I, 3);
            T[seg(i)] = I >= 2 ? U(I, 2) : 0.0;
            T[NP + seg(i)] = I >= 3 ? U(I, 1) : 0.0;
            T[2 * NP + seg(i)] = d;
            T[3 * NP + seg(i)] = I <= nx - 1 ? U(I, 4) * d : 0.0;
            T[4 * NP + seg(i)] = I <= nx - 2 ? U(I, 5) * d : 0.0;
        }
    }
    // integration: x(i) = g - P x(i-1) - Q x(i-2)
    std::vector<double> ti(2 * static_cast<size_t>(NP), 0.0);
    for (int i = 1; i < nx; ++i) {
        double P, Q;
        if (i == 1 || i == nx - 1) { P = -1.0; Q = 0.0; }
        else if (i == nx - 2) { P = -24.0 / 57.0; Q = -33.0 / 57.0; }
        else { P = -16.0 / 27.0; Q = -11.0 / 27.0; }
        ti[seg(i)] = P;
        ti[NP + seg(i)] = Q;
    }
    // filter (lhs_tridf filter_par.f:181-212, tridseqx derivs_par.f:1009-1037):
    // forward a/bet, 1/bet; backward gam(i+1)
    std::vector<double> tf(3 * static_cast<size_t>(NP), 0.0);
    {
        std::vector<double> fa(nx, 0.0), fb(nx, 1.0), fc(nx, 0.0), gam(nx + 1, 0.0);
        for (int i = 3; i <= nx - 4; ++i) { fa[i] = kAlphaF; fc[i] = kAlphaF; }
        double bet = fb[0];
        tf[seg(0)] = 0.0;
        tf[NP + seg(0)] = 1.0 / bet;
        for (int i = 1; i < nx; ++i) {
            gam[i] = fc[i - 1] / bet;
            bet = fb[i] - fa[i] * gam[i];
            tf[seg(i)] = fa[i] / bet;
            tf[NP + seg(i)] = 1.0 / bet;
        }
        for (int i = 0; i < nx - 1; ++i) tf[2 * NP + seg(i)] = gam[i + 1];
    }
    c->xR = R;
    TRY(dalloc(&c->d_xpenta, tp.size()));
    TRY(dalloc(&c->d_xinteg, ti.size()));
    TRY(dalloc(&c->d_xfilter, tf.size()));
    TBL_HIP(hipMemcpy(c->d_xpenta, tp.data(), tp.size() * sizeof(double), hipMemcpyHostToDevice));
    TBL_HIP(hipMemcpy(c->d_xinteg, ti.data(), ti.size() * sizeof(double), hipMemcpyHostToDevice));
    TBL_HIP(hipMemcpy(c->d_xfilter, tf.data(), tf.size() * sizeof(double), hipMemcpyHostToDevice));
    const size_t n3 = static_cast<size_t>(nx) * ny * K, n2 = static_cast<size_t>(nx) * K;
    TRY(dalloc(&c->d_t1, n3));
    TRY(dalloc(&c->d_t2, n3));
    TRY(dalloc(&c->d_lapv, n2));
    TRY(dalloc(&c->d_dwydy, n2));
    TRY(dalloc(&c->d_d2wy, n2));
    return TBL_OK;
}

cx<double> *C(double *p) { return reinterpret_cast<cx<double> *>(p); }
const cx<double> *C(const double *p) { return reinterpret_cast<const cx<double> *>(p); }

int launched() { TBL_HIP(hipGetLastError()); return TBL_OK; }

int poi_ux(tbl_ctx *c, const double *uy, const double *wy, double *ux, double *duydy)
{
    const int nx = c->g.nx, ny = c->g.ny, K = c->K;
    TRY(deriv(c->dargs, 4, uy, duydy, c->stream));           // deryfv
    XInteg in{C(duydy), nullptr, C(ux), c->d_xinteg, nx, ny, K, c->xR, kXIntUx, -c->g.dx,
              c->g.beta};
    TRY(xinteg(in, c->stream));
    if (K > 1) {
        double *d2 = reinterpret_cast<double *>(c->d_t2);
        TRY(deriv(c->dargs, 0, duydy, d2, c->stream));        // derparx
        XPenta xp{C(d2), C(wy), C(ux), c->d_xpenta, nx, ny, K, c->xR, kXUx,
                  c->g.dx * c->g.dx, c->g.beta};
        TRY(xpenta(xp, c->stream));
    }
    return launched();
}

int poi_uz(tbl_ctx *c, const double *wy, const double *duydy, double *uz)
{
    const int nx = c->g.nx, ny = c->g.ny, K = c->K;
    if (K < 2) return TBL_OK;
    double *dw = reinterpret_cast<double *>(c->d_t2);
    TRY(deriv(c->dargs, 0, wy, dw, c->stream));
    XPenta xp{C(dw), C(duydy), C(uz), c->d_xpenta, nx, ny, K, c->xR, kXUz, c->g.dx * c->g.dx,
              c->g.beta};
    TRY(xpenta(xp, c->stream));
    return launched();
}

int wx_wall(tbl_ctx *c, const double *uy, const double *wy, const double *d2uydx2, double *wx,
            double *lapv)
{
    const int nx = c->g.nx, ny = c->g.ny, K = c->K;
    WallArgs w{};
    w.uy = C(uy);
    w.wy = C(wy);
    w.d2uydx2 = C(d2uydx2);
    w.lapv = C(lapv);
    w.dwydy = c->d_dwydy;
    w.d2wydxdy = c->d_d2wy;
    std::memcpy(w.w, c->c.w_at_w, sizeof(w.w));
    std::memcpy(w.c, c->c.dwydy, sizeof(w.c));
    w.dyyw1 = c->g.dy0 * c->g.dy0 * c->c.w_at_w[0];
    w.rdy = 1.0 / c->g.dy0;
    w.dx = c->g.dx;
    w.beta = c->g.beta;
    w.nx = nx;
    w.ny = ny;
    w.K = K;
    TRY(wall_terms(w, c->stream));
    if (K > 1) {
        XPenta xp{c->d_d2wy, C(lapv), C(wx), c->d_xpenta, nx, ny, K, c->xR, kXWall,
                  c->g.dx * c->g.dx, c->g.beta};
        TRY(xpenta(xp, c->stream));
    }
    return launched();
}

int wz_wall(tbl_ctx *c, const double *wx, const double *lapv, double *wz)
{
    XInteg in{C(wx), C(lapv), C(wz), c->d_xinteg, c->g.nx, c->g.ny, c->K, c->xR, kXIntWz,
              c->g.dx, c->g.beta};
    TRY(xinteg(in, c->stream));
    return launched();
}

int bzone_rows(tbl_ctx *c, const double *bdfc, int wall_row, double *wx, double *wy, double *wz,
               double *th)
{
    BzoneArgs b{bdfc, C(wx), C(wy), C(wz), wall_row ? nullptr : C(th), c->g.nx, c->g.ny, c->K,
                wall_row ? 0 : 1};
    TRY(bzone(b, wall_row ? 1 : c->g.ny - 1, c->stream));
    return launched();
}

int filter_trid(tbl_ctx *c, double *wx, double *wy, double *wz)
{
    XFilter f{C(wx), C(wy), C(wz), c->d_xfilter, c->g.nx, c->g.ny, c->K, c->xR,
              kAf, kBf, kCf, kDf};
    TRY(xfilter(f, c->stream));
    return launched();
}

int outuy(tbl_ctx *c, const double *wz, double *uy, double *dwzdx)
{
    TRY(deriv(c->dargs, 0, wz, dwzdx, c->stream));
    TRY(outflow_row(C(uy), c->g.nx, c->g.ny, c->K, c->stream));
    return launched();
}

bool distinct(std::initializer_list<const void *> ps)
{
    for (auto a = ps.begin(); a != ps.end(); ++a)
        for (auto b = a + 1; b != ps.end(); ++b)
            if (*a && *a == *b) return false;
    return true;
}

}  // namespace

extern "C" int tbl_bdfc_generate(int nx, int i0, int i3, int i4, double *bdfc)
{
    if (!bdfc || nx < 2 || i0 < 2 || i0 > i3 || i3 >= i4 || i4 > nx)
        return fail("tbl_bdfc_generate: need 2 <= i0 <= i3 < i4 <= nx");
    auto B = [&](int r, int i) -> double & { return bdfc[(r - 1) + 2 * static_cast<size_t>(i - 1)]; };
    for (int i = 1; i <= i0; ++i) {
        const double ep = static_cast<double>(i - 1) / static_cast<double>(i0 - 1);
        B(1, i) = ((6.0 * ep - 15.0) * ep + 10.0) * (ep * ep * ep);
        B(2, i) = B(1, i);
    }
    for (int i = i0; i <= i3; ++i) B(1, i) = B(2, i) = 1.0;
    for (int i = i3; i <= i4; ++i) {
        const double ep = static_cast<double>(i - i3) / static_cast<double>(i4 - i3);
        const double t = 1.0 - std::pow(ep, 50);
        B(1, i) = (t * t) * (t * t) * std::exp(-std::pow(ep, 4) / 10.0);
        B(2, i) = B(1, i);
    }
    for (int i = i4; i <= nx; ++i) B(1, i) = B(2, i) = 0.0;
    return TBL_OK;
}

extern "C" int tbl_poi_ux_dev(tbl_ctx *c, const double *d_uy, const double *d_wy, double *d_ux,
                              double *d_duydy)
{
    if (!c || !d_uy || !d_wy || !d_ux || !d_duydy) return TBL_ERR_ARG;
    if (!distinct({d_uy, d_wy, d_ux, d_duydy})) return fail("tbl_poi_ux_dev: aliased fields");
    TRY(loop_setup(c));
    return poi_ux(c, d_uy, d_wy, d_ux, d_duydy);
}

extern "C" int tbl_poi_uz_dev(tbl_ctx *c, const double *d_wy, const double *d_duydy, double *d_uz)
{
    if (!c || !d_wy || !d_duydy || !d_uz) return TBL_ERR_ARG;
    if (!distinct({d_wy, d_duydy, d_uz})) return fail("tbl_poi_uz_dev: aliased fields");
    TRY(loop_setup(c));
    return poi_uz(c, d_wy, d_duydy, d_uz);
}

extern "C" int tbl_wx_wall_dev(tbl_ctx *c, const double *d_uy, const double *d_wy,
                               const double *d_d2uydx2, double *d_wx, double *d_lapv)
{
    if (!c || !d_uy || !d_wy || !d_d2uydx2 || !d_wx || !d_lapv) return TBL_ERR_ARG;
    if (!distinct({d_uy, d_wy, d_d2uydx2, d_wx, d_lapv}))
        return fail("tbl_wx_wall_dev: aliased fields");
    TRY(loop_setup(c));
    return wx_wall(c, d_uy, d_wy, d_d2uydx2, d_wx, d_lapv);
}

extern "C" int tbl_wz_wall_dev(tbl_ctx *c, const double *d_wx, const double *d_lapv, double *d_wz)
{
    if (!c || !d_wx || !d_lapv || !d_wz) return TBL_ERR_ARG;
    if (!distinct({d_wx, d_lapv, d_wz})) return fail("tbl_wz_wall_dev: aliased fields");
    TRY(loop_setup(c));
    return wz_wall(c, d_wx, d_lapv, d_wz);
}

extern "C" int tbl_bzone_dev(tbl_ctx *c, const double *d_bdfc, int wall_row, double *d_wx,
                             double *d_wy, double *d_wz, double *d_th)
{
    if (!c || !d_bdfc || !d_wx || !d_wy || !d_wz) return TBL_ERR_ARG;
    if (!distinct({d_wx, d_wy, d_wz, d_th})) return fail("tbl_bzone_dev: aliased fields");
    return bzone_rows(c, d_bdfc, wall_row, d_wx, d_wy, d_wz, d_th);
}

extern "C" int tbl_filter_trid_dev(tbl_ctx *c, double *d_wx, double *d_wy, double *d_wz)
{
    if (!c || !d_wx || !d_wy || !d_wz) return TBL_ERR_ARG;
    if (!distinct({d_wx, d_wy, d_wz})) return fail("tbl_filter_trid_dev: aliased fields");
    TRY(loop_setup(c));
    return filter_trid(c, d_wx, d_wy, d_wz);
}

extern "C" int tbl_outuy_dev(tbl_ctx *c, const double *d_wz, double *d_uy, double *d_dwzdx)
{
    if (!c || !d_wz || !d_uy || !d_dwzdx) return TBL_ERR_ARG;
    if (!distinct({d_wz, d_uy, d_dwzdx})) return fail("tbl_outuy_dev: aliased fields");
    return outuy(c, d_wz, d_uy, d_dwzdx);
}

extern "C" int tbl_loop_dev(tbl_ctx *c, double erro, const double *d_bdfc,
                            const double *d_d2uydx2, double *d_ux, double *d_uy, double *d_uz,
                            double *d_wx, double *d_wy, double *d_wz, double *d_th,
                            const tbl_solve_opts *opts, int *cycles, double *residual,
                            tbl_solve_stats *stats)
{
    if (!c || !d_bdfc || !d_d2uydx2 || !d_ux || !d_uy || !d_uz || !d_wx || !d_wy || !d_wz)
        return TBL_ERR_ARG;
    if (!distinct({d_ux, d_uy, d_uz, d_wx, d_wy, d_wz, d_th}))
        return fail("tbl_loop_dev: aliased fields");
    TRY(loop_setup(c));
    double *t1 = reinterpret_cast<double *>(c->d_t1);
    double *lapv = reinterpret_cast<double *>(c->d_lapv);
    if (erro < 1e-5) TRY(filter_trid(c, d_wx, d_wy, d_wz));
    TRY(bzone_rows(c, d_bdfc, 0, d_wx, d_wy, d_wz, d_th));
    TRY(outuy(c, d_wz, d_uy, t1));
    TRY(solve_dev(c, d_wx, t1, d_uy, 0, erro, opts, cycles, residual, stats));
    TRY(poi_ux(c, d_uy, d_wy, d_ux, t1));
    TRY(poi_uz(c, d_wy, t1, d_uz));
    TRY(wx_wall(c, d_uy, d_wy, d_d2uydx2, d_wx, lapv));
    TRY(wz_wall(c, d_wx, lapv, d_wz));
    TRY(bzone_rows(c, d_bdfc, 1, d_wx, d_wy, d_wz, nullptr));
    return TBL_OK;
}

extern "C" int tbl_loop(tbl_ctx *c, double erro, const double *bdfc, const double *d2uydx2,
                        double *ux, double *uy, double *uz, double *wx, double *wy, double *wz,
                        double *th, const tbl_solve_opts *opts, int *cycles, double *residual,
                        tbl_solve_stats *stats)
{
    if (!c || !bdfc || !d2uydx2 || !ux || !uy || !uz || !wx || !wy || !wz) return TBL_ERR_ARG;
    TRY(loop_setup(c));
    const int nx = c->g.nx, ny = c->g.ny, K = c->K;
    const size_t n3 = static_cast<size_t>(nx) * ny * K * 2, n2 = static_cast<size_t>(nx) * K * 2;
    const int nf = th ? 7 : 6;
    double *d = nullptr;
    TBL_HIP(hipMalloc(reinterpret_cast<void **>(&d),
                      sizeof(double) * (nf * n3 + n2 + 2 * static_cast<size_t>(nx))));
    double *h[7] = {ux, uy, uz, wx, wy, wz, th};
    double *dv[7] = {};
    int rc = TBL_OK;
    for (int f = 0; f < nf && rc == TBL_OK; ++f) {
        dv[f] = d + f * n3;
        if (hipMemcpy(dv[f], h[f], sizeof(double) * n3, hipMemcpyHostToDevice) != hipSuccess)
            rc = TBL_ERR_HIP;
    }
    double *dd2 = d + nf * n3, *db = dd2 + n2;
    if (rc == TBL_OK &&
        (hipMemcpy(dd2, d2uydx2, sizeof(double) * n2, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(db, bdfc, sizeof(double) * 2 * nx, hipMemcpyHostToDevice) != hipSuccess))
        rc = TBL_ERR_HIP;
    if (rc == TBL_OK)
        rc = tbl_loop_dev(c, erro, db, dd2, dv[0], dv[1], dv[2], dv[3], dv[4], dv[5],
                          th ? dv[6] : nullptr, opts, cycles, residual, stats);
    if (rc == TBL_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = TBL_ERR_HIP;
    for (int f = 0; f < nf && rc == TBL_OK; ++f)
        if (hipMemcpy(h[f], dv[f], sizeof(double) * n3, hipMemcpyDeviceToHost) != hipSuccess)
            rc = TBL_ERR_HIP;
    if (rc == TBL_ERR_HIP) tbl_set_last_error("tbl_loop: copy failed", __FILE__, __LINE__);
    (void)hipFree(d);
    return rc;
}
