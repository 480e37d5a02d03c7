// Right-hand side (drv_gv with lterms_gv or nlterms_gv) and the RK4 time
// step of the reference's Goertler case on the GPU: host side -- work
// fields, the z-DFT twiddle table, the launch sequence of drv_gv (terms,
// the eleven derivatives as seven derivative calls, the combination) and
// of the four RK4 stages around tbl_loop_dev (main_par.f:266-349); and
// drv_fi, the flat-plate form (main_par.f:1264-1336: lterms_fi /
// nlterms_fi, ten derivatives, the combination without curvature terms).
#include <cmath>
#include <vector>

#include "tbl_ctx.h"
#include "tbl_kernels.h"
#include "tbl_rhs.h"

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

size_t field(const tbl_ctx *c) { return static_cast<size_t>(c->g.nx) * c->g.ny * c->K; }

int ensure(tbl_ctx *c, cx<double> **p, size_t nfields)
{
    if (*p) return TBL_OK;
    const size_t bytes = sizeof(cx<double>) * field(c) * nfields;
    TBL_HIP(hipMalloc(reinterpret_cast<void **>(p), bytes));
    TBL_HIP(hipMemset(*p, 0, bytes));
    return TBL_OK;
}

int check_kphys(const tbl_ctx *c, int kphys)
{
    if (kphys == 0) return TBL_OK;
    if (kphys < 4 || kphys < 2 * c->K || (kphys & (kphys - 1)))
        return fail("kphys must be 0 (linear terms) or a power of two >= max(4, 2 nmodes)");
    if (nlterms_lds(c->K, kphys) > 160 * 1024) return fail("kphys too large for the LDS");
    return TBL_OK;
}

int twiddles(tbl_ctx *c, int N)
{
    if (c->tw_n == N) return TBL_OK;
    if (c->d_tw) {
        TBL_HIP(hipFree(c->d_tw));
        c->d_tw = nullptr;
    }
    std::vector<double> t(2 * static_cast<size_t>(N));
    const double two_pi = 6.283185307179586476925286766559;
    for (int q = 0; q < N; ++q) {
        t[q] = std::cos(two_pi * q / N);
        t[N + q] = std::sin(two_pi * q / N);
    }
    TBL_HIP(hipMalloc(reinterpret_cast<void **>(&c->d_tw), sizeof(double) * t.size()));
    TBL_HIP(hipMemcpy(c->d_tw, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    c->tw_n = N;
    return TBL_OK;
}

const cx<double> *C(const double *p) { return reinterpret_cast<const cx<double> *>(p); }
cx<double> *C(double *p) { return reinterpret_cast<cx<double> *>(p); }

BaseArgs base_of(const tbl_base *b)
{
    return b ? BaseArgs{b->uxb, b->uyb, b->wzb, b->varg, b->dvargdx} : BaseArgs{};
}

// fi: the flat-plate form (a, b, c; out[3] unused, b unused when kphys > 0)
int do_terms(tbl_ctx *c, int kphys, const tbl_base *b, const double *const *f,
             cx<double> *const *out, bool fi = false)
{
    if (kphys) TRY(twiddles(c, kphys));
    RhsArgs r{base_of(b), C(f[0]), C(f[1]), C(f[2]), C(f[3]), C(f[4]), C(f[5]),
              out[0], out[1], out[2], fi ? nullptr : out[3], kphys ? c->d_tw : nullptr,
              c->g.nx, c->g.ny, c->K, fi ? 1 : 0};
    TRY(terms(r, kphys, c->stream));
    TBL_HIP(hipGetLastError());
    return TBL_OK;
}

// drv_gv (main_par.f:1339-1403), or with fi drv_fi (main_par.f:1264-1336):
// f = ux, uy, uz, wx, wy, wz; dv = dvx, dvy, dvz
int drv_gv(tbl_ctx *c, double Re, int kphys, const tbl_base *b, const double *const *f,
           cx<double> *const *dv, bool fi = false)
{
    TRY(ensure(c, &c->d_rhs, 15));
    const size_t n = field(c);
    cx<double> *w[15];
    for (int q = 0; q < 15; ++q) w[q] = c->d_rhs + q * n;
    cx<double> *a = w[0], *bb = w[1], *cc = w[2], *d = w[3];
    TRY(do_terms(c, kphys, b, f, w, fi));
    auto D = [](cx<double> *p) { return reinterpret_cast<double *>(p); };
    const DerivArgs &A = c->dargs;
    hipStream_t st = c->stream;
    // a: d/dx, d/dy; b, d (not in drv_fi): d/dx; c: d/dy; wx, wy, wz:
    // d2/dx2, d2/dy2
    TRY(derivs4(A, D(a), D(w[4]), nullptr, D(w[7]), nullptr, st));
    TRY(derivs4(A, D(bb), D(w[5]), nullptr, nullptr, nullptr, st));
    if (!fi) TRY(derivs4(A, D(d), D(w[6]), nullptr, nullptr, nullptr, st));
    TRY(derivs4(A, D(cc), nullptr, nullptr, D(w[8]), nullptr, st));
    TRY(derivs4(A, f[3], nullptr, D(w[9]), nullptr, D(w[12]), st));
    TRY(derivs4(A, f[4], nullptr, D(w[10]), nullptr, D(w[13]), st));
    TRY(derivs4(A, f[5], nullptr, D(w[11]), nullptr, D(w[14]), st));
    DrvArgs r{};
    r.b = base_of(b);
    r.wx = C(f[3]);
    r.wy = C(f[4]);
    r.wz = C(f[5]);
    r.bb = bb;
    r.c = cc;
    r.d = d;
    r.dadx = w[4];
    r.dbdx = w[5];
    r.dddx = w[6];
    r.dady = w[7];
    r.dcdy = w[8];
    r.xx = w[9];
    r.yx = w[10];
    r.zx = w[11];
    r.xy = w[12];
    r.yy = w[13];
    r.zy = w[14];
    r.dvx = dv[0];
    r.dvy = dv[1];
    r.dvz = dv[2];
    r.fi = fi ? 1 : 0;
    r.rre = 1.0 / Re;
    r.fac_y = c->g.fac_y;
    r.beta = c->g.beta;
    r.nx = c->g.nx;
    r.ny = c->g.ny;
    r.K = c->K;
    TRY(drv_combine(r, st));
    TBL_HIP(hipGetLastError());
    return TBL_OK;
}

bool base_ok(const tbl_base *b)
{
    return b && b->uxb && b->uyb && b->wzb && b->varg && b->dvargdx;
}

}  // namespace

extern "C" int tbl_terms_dev(tbl_ctx *c, int kphys, const tbl_base *b, const double *d_ux,
                             const double *d_uy, const double *d_uz, const double *d_wx,
                             const double *d_wy, const double *d_wz, double *d_a, double *d_b,
                             double *d_c, double *d_d)
{
    if (!c || !base_ok(b) || !d_ux || !d_uy || !d_uz || !d_wx || !d_wy || !d_wz || !d_a ||
        !d_b || !d_c || !d_d)
        return TBL_ERR_ARG;
    TRY(check_kphys(c, kphys));
    const double *f[6] = {d_ux, d_uy, d_uz, d_wx, d_wy, d_wz};
    cx<double> *o[4] = {C(d_a), C(d_b), C(d_c), C(d_d)};
    for (auto *x : f)
        for (auto *y : o)
            if (reinterpret_cast<const void *>(x) == y) return fail("tbl_terms_dev: aliased fields");
    return do_terms(c, kphys, b, f, o);
}

extern "C" int tbl_drv_gv_dev(tbl_ctx *c, double Re, int kphys, const tbl_base *b,
                              const double *d_ux, const double *d_uy, const double *d_uz,
                              const double *d_wx, const double *d_wy, const double *d_wz,
                              double *d_dvx, double *d_dvy, double *d_dvz)
{
    if (!c || !base_ok(b) || !d_ux || !d_uy || !d_uz || !d_wx || !d_wy || !d_wz || !d_dvx ||
        !d_dvy || !d_dvz || !(Re > 0.0))
        return TBL_ERR_ARG;
    TRY(check_kphys(c, kphys));
    const double *f[6] = {d_ux, d_uy, d_uz, d_wx, d_wy, d_wz};
    cx<double> *dv[3] = {C(d_dvx), C(d_dvy), C(d_dvz)};
    return drv_gv(c, Re, kphys, b, f, dv);
}

// flat-plate form: base needs uxb, uyb, wzb for the linear terms only
static bool fi_base_ok(const tbl_base *b, int kphys)
{
    return kphys > 0 || (b && b->uxb && b->uyb && b->wzb);
}

extern "C" int tbl_terms_fi_dev(tbl_ctx *c, int kphys, const tbl_base *b, const double *d_ux,
                                const double *d_uy, const double *d_uz, const double *d_wx,
                                const double *d_wy, const double *d_wz, double *d_a, double *d_b,
                                double *d_c)
{
    if (!c || !fi_base_ok(b, kphys) || !d_ux || !d_uy || !d_uz || !d_wx || !d_wy || !d_wz ||
        !d_a || !d_b || !d_c)
        return TBL_ERR_ARG;
    TRY(check_kphys(c, kphys));
    const double *f[6] = {d_ux, d_uy, d_uz, d_wx, d_wy, d_wz};
    cx<double> *o[4] = {C(d_a), C(d_b), C(d_c), nullptr};
    for (auto *x : f)
        for (int q = 0; q < 3; ++q)
            if (reinterpret_cast<const void *>(x) == o[q])
                return fail("tbl_terms_fi_dev: aliased fields");
    return do_terms(c, kphys, kphys ? nullptr : b, f, o, true);
}

extern "C" int tbl_drv_fi_dev(tbl_ctx *c, double Re, int kphys, const tbl_base *b,
                              const double *d_ux, const double *d_uy, const double *d_uz,
                              const double *d_wx, const double *d_wy, const double *d_wz,
                              double *d_dvx, double *d_dvy, double *d_dvz)
{
    if (!c || !fi_base_ok(b, kphys) || !d_ux || !d_uy || !d_uz || !d_wx || !d_wy || !d_wz ||
        !d_dvx || !d_dvy || !d_dvz || !(Re > 0.0))
        return TBL_ERR_ARG;
    TRY(check_kphys(c, kphys));
    const double *f[6] = {d_ux, d_uy, d_uz, d_wx, d_wy, d_wz};
    cx<double> *dv[3] = {C(d_dvx), C(d_dvy), C(d_dvz)};
    return drv_gv(c, Re, kphys, kphys ? nullptr : b, f, dv, true);
}

// host arrays of the RHS entry points: fields in (dv in as well, so row 1
// keeps the caller's values), base-flow arrays in when given, dv out
static int drv_host(tbl_ctx *c, double Re, int kphys, bool fi, const double *const *hb, int nb,
             const double *const *hin, double *const *hout)
{
    const size_t n3 = field(c) * 2, n2 = static_cast<size_t>(c->g.nx) * c->g.ny;
    double *d = nullptr;
    TBL_HIP(hipMalloc(reinterpret_cast<void **>(&d), sizeof(double) * (9 * n3 + 5 * n2)));
    double *df[9], *db[5] = {};
    int rc = TBL_OK;
    for (int q = 0; q < 9; ++q) df[q] = d + q * n3;
    for (int q = 0; q < nb; ++q) db[q] = d + 9 * n3 + q * n2;
    for (int q = 0; q < 9 && rc == TBL_OK; ++q)
        if (hipMemcpy(df[q], q < 6 ? hin[q] : hout[q - 6], sizeof(double) * n3,
                      hipMemcpyHostToDevice) != hipSuccess)
            rc = TBL_ERR_HIP;
    for (int q = 0; q < nb && rc == TBL_OK; ++q)
        if (hipMemcpy(db[q], hb[q], sizeof(double) * n2, hipMemcpyHostToDevice) != hipSuccess)
            rc = TBL_ERR_HIP;
    if (rc == TBL_OK) {
        const tbl_base b{db[0], db[1], db[2], db[3], db[4]};
        rc = fi ? tbl_drv_fi_dev(c, Re, kphys, nb ? &b : nullptr, df[0], df[1], df[2], df[3],
                                 df[4], df[5], df[6], df[7], df[8])
                : tbl_drv_gv_dev(c, Re, kphys, &b, df[0], df[1], df[2], df[3], df[4], df[5],
                                 df[6], df[7], df[8]);
    }
    if (rc == TBL_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = TBL_ERR_HIP;
    for (int q = 0; q < 3 && rc == TBL_OK; ++q)
        if (hipMemcpy(hout[q], df[6 + q], sizeof(double) * n3, hipMemcpyDeviceToHost) !=
            hipSuccess)
            rc = TBL_ERR_HIP;
    if (rc == TBL_ERR_HIP) tbl_set_last_error("drv_gv / drv_fi: copy failed", __FILE__, __LINE__);
    (void)hipFree(d);
    return rc;
}

extern "C" int tbl_drv_gv(tbl_ctx *c, double Re, int kphys, const double *uxb,
                          const double *uyb, const double *wzb, const double *varg,
                          const double *dvargdx, const double *ux, const double *uy,
                          const double *uz, const double *wx, const double *wy, const double *wz,
                          double *dvx, double *dvy, double *dvz)
{
    if (!c || !uxb || !uyb || !wzb || !varg || !dvargdx || !ux || !uy || !uz || !wx || !wy ||
        !wz || !dvx || !dvy || !dvz || !(Re > 0.0))
        return TBL_ERR_ARG;
    TRY(check_kphys(c, kphys));
    const double *hb[5] = {uxb, uyb, wzb, varg, dvargdx};
    const double *hin[6] = {ux, uy, uz, wx, wy, wz};
    double *hout[3] = {dvx, dvy, dvz};
    return drv_host(c, Re, kphys, false, hb, 5, hin, hout);
}

extern "C" int tbl_drv_fi(tbl_ctx *c, double Re, int kphys, const double *uxb,
                          const double *uyb, const double *wzb, const double *ux,
                          const double *uy, const double *uz, const double *wx, const double *wy,
                          const double *wz, double *dvx, double *dvy, double *dvz)
{
    const bool with_base = uxb && uyb && wzb;
    if (!c || (kphys == 0 && !with_base) || !ux || !uy || !uz || !wx || !wy || !wz || !dvx ||
        !dvy || !dvz || !(Re > 0.0))
        return TBL_ERR_ARG;
    TRY(check_kphys(c, kphys));
    const double *hb[3] = {uxb, uyb, wzb};
    const double *hin[6] = {ux, uy, uz, wx, wy, wz};
    double *hout[3] = {dvx, dvy, dvz};
    return drv_host(c, Re, kphys, true, hb, kphys == 0 ? 3 : 0, hin, hout);
}

extern "C" int tbl_rk4_step_dev(tbl_ctx *c, double Re, double dt, int kphys, const tbl_base *b,
                                const double *d_bdfc, double *d_d2uydx2,
                                const double *d_pert_uy, const double *d_pert_d2, int pert_i0,
                                int pert_i1, double *d_ux, double *d_uy, double *d_uz,
                                double *d_wx, double *d_wy, double *d_wz, double *d_th,
                                const tbl_solve_opts *opts, int *cycles, double *residual)
{
    if (!c || !base_ok(b) || !d_bdfc || !d_d2uydx2 || !d_ux || !d_uy || !d_uz || !d_wx ||
        !d_wy || !d_wz || !(Re > 0.0))
        return TBL_ERR_ARG;
    if (d_pert_uy && (!d_pert_d2 || pert_i0 < 1 || pert_i1 > c->g.nx))
        return fail("tbl_rk4_step_dev: disturbance rows outside 1..nx");
    TRY(check_kphys(c, kphys));
    TRY(ensure(c, &c->d_rk, 9));
    const size_t n = field(c);
    cx<double> *dv1[3], *dv2[3], *w1[3];
    for (int q = 0; q < 3; ++q) {
        dv1[q] = c->d_rk + q * n;
        dv2[q] = c->d_rk + (3 + q) * n;
        w1[q] = c->d_rk + (6 + q) * n;
    }
    const double *f[6] = {d_ux, d_uy, d_uz, d_wx, d_wy, d_wz};
    const double hs[4] = {0.5 * dt, 0.5 * dt, dt, dt / 6.0};
    for (int s = 1; s <= 4; ++s) {
        TRY(drv_gv(c, Re, kphys, b, f, s == 1 ? dv1 : dv2));
        RkArgs r{};
        r.w[0] = C(d_wx);
        r.w[1] = C(d_wy);
        r.w[2] = C(d_wz);
        for (int q = 0; q < 3; ++q) {
            r.w1[q] = w1[q];
            r.dv1[q] = dv1[q];
            r.dv2[q] = dv2[q];
        }
        r.h = hs[s - 1];
        r.stage = s;
        r.nx = c->g.nx;
        r.ny = c->g.ny;
        r.K = c->K;
        TRY(rk_update(r, c->stream));
        if ((s == 1 || s == 3) && d_pert_uy)
            TRY(pert_rows(C(d_uy), C(d_d2uydx2), C(d_pert_uy), C(d_pert_d2), c->g.nx, c->g.ny,
                          c->K, pert_i0 - 1, pert_i1 - 1, c->stream));
        TBL_HIP(hipGetLastError());
        TRY(tbl_loop_dev(c, s == 4 ? 1e-6 : 1e-5, d_bdfc, d_d2uydx2, d_ux, d_uy, d_uz, d_wx,
                         d_wy, d_wz, d_th, opts, cycles ? cycles + (s - 1) * c->K : nullptr,
                         residual ? residual + (s - 1) * c->K : nullptr, nullptr));
    }
    return TBL_OK;
}
