/*
 * CPU oracle, part 2: the routines of the reference's `loop` (main_par.f:
 * 1762-1785) around poi_uy -- the x-direction Poisson solves for ux and uz,
 * the wall vorticity, the buffer zone, the filter and the outflow uy.
 *
 * TEST INFRASTRUCTURE ONLY (see tbl_oracle.h).  One MPI rank: the
 * reference's halo exchanges (boundary_exchange_1stmode, _1d, _wwall,
 * _derivs; loop_par.f:728-837) do nothing there and are left out, and
 * `shift` is 0.  Expressions keep the reference's evaluation order.
 */
#include "tbl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define LU5(m, i, c) (m)[((i) - 1) * 5 + ((c) - 1)]
/* (i, j, k) 1-based into a Fortran (nx, ny, nk) field */
#define F3(a, i, j, k) (a)[((i) - 1) + (size_t)nx * (((j) - 1) + (size_t)ny * ((k) - 1))]

static double k2b2_of(const or_params *p, int k)
{
    /* v_k2b2(k), main_par.f:1887 */
    return -(double)(k - 1) * (double)(k - 1) * p->beta * p->beta;
}

static or_cplx kb_of(const or_params *p, int k)
{
    /* v_kb(k) = -im (k-1) beta, main_par.f:1888 */
    return -((double)(k - 1) * p->beta) * I;
}

/* lhs5: loop_par.f:258-299 */
void or_lhs5(int n, double k2b2, double *a)
{
    memset(a, 0, sizeof(double) * 5 * (size_t)n);
    LU5(a, 1, 3) = 1.0;
    LU5(a, 2, 2) = 1.0 + k2b2 / 12.0;
    LU5(a, 2, 3) = -2.0 + k2b2 / 1.2;
    LU5(a, 2, 4) = 1.0 + k2b2 / 12.0;
    for (int i = 3; i <= n - 2; ++i) {
        LU5(a, i, 1) = 3.0;
        LU5(a, i, 2) = 48.0 + 8.0 * k2b2;
        LU5(a, i, 3) = -102.0 + 44.0 * k2b2;
        LU5(a, i, 4) = 48.0 + 8.0 * k2b2;
        LU5(a, i, 5) = 3.0;
    }
    LU5(a, n - 1, 2) = 1.0 + k2b2 / 12.0;
    LU5(a, n - 1, 3) = -2.0 + k2b2 / 1.2;
    LU5(a, n - 1, 4) = 1.0 + k2b2 / 12.0;
    LU5(a, n, 3) = 1.0;
}

/* ludecomp: penta_par.f:6-45.  Row i of U is the row of A less the products
   of row i's multipliers with the U rows above it; the multipliers of rows
   i+1 and i+2 divide by U(i,i).  Rows n-1 and n are written out separately
   in the reference, with the two subtractions in sequence. */
void or_ludecomp(int n, const double *a, double *lu)
{
    memset(lu, 0, sizeof(double) * 5 * (size_t)n);
    for (int i = 1; i <= n - 2; ++i) {
        double s3 = 0.0, s4 = 0.0;
        if (i >= 2) s3 = s3 + LU5(lu, i, 2) * LU5(lu, i - 1, 4);
        if (i >= 3) {
            s3 = s3 + LU5(lu, i, 1) * LU5(lu, i - 2, 5);
            s4 = s4 + LU5(lu, i, 2) * LU5(lu, i - 1, 5);
        }
        LU5(lu, i, 3) = LU5(a, i, 3) - s3;
        LU5(lu, i, 4) = LU5(a, i, 4) - s4;
        LU5(lu, i, 5) = LU5(a, i, 5);
        double s = 0.0;
        if (i >= 2) s = LU5(lu, i + 1, 1) * LU5(lu, i - 1, 4);
        LU5(lu, i + 1, 2) = (LU5(a, i + 1, 2) - s) / LU5(lu, i, 3);
        LU5(lu, i + 2, 1) = LU5(a, i + 2, 1) / LU5(lu, i, 3);
    }
    int i = n - 1;
    LU5(lu, i, 3) = LU5(a, i, 3) - LU5(lu, i, 2) * LU5(lu, i - 1, 4)
                    - LU5(lu, i, 1) * LU5(lu, i - 2, 5);
    LU5(lu, i, 4) = LU5(a, i, 4) - LU5(lu, i, 2) * LU5(lu, i - 1, 5);
    LU5(lu, i + 1, 2) = (LU5(a, i + 1, 2) - LU5(lu, i + 1, 1) * LU5(lu, i - 1, 4))
                        / LU5(lu, i, 3);
    i = n;
    LU5(lu, i, 3) = LU5(a, i, 3) - LU5(lu, i, 2) * LU5(lu, i - 1, 4)
                    - LU5(lu, i, 1) * LU5(lu, i - 2, 5);
}

/* lusolver (penta_par.f:139-208) / lusolver_1 (53-136) at one rank */
void or_lusolve(int n, const double *lu, or_cplx *rhs, size_t stride, int variant)
{
    or_cplx *y = malloc(sizeof(or_cplx) * (size_t)n);
#define R(i) rhs[((i) - 1) * stride]
    y[0] = R(1);
    y[1] = R(2) - LU5(lu, 2, 2) * y[0];
    for (int i = 3; i <= n; ++i) {
        if (variant == 1)
            y[i - 1] = R(i) - LU5(lu, i, 2) * y[i - 2] - LU5(lu, i, 1) * y[i - 3];
        else
            y[i - 1] = R(i) - LU5(lu, i, 1) * y[i - 3] - LU5(lu, i, 2) * y[i - 2];
    }
    R(n) = y[n - 1] / LU5(lu, n, 3);
    R(n - 1) = (y[n - 2] - LU5(lu, n - 1, 4) * R(n)) / LU5(lu, n - 1, 3);
    for (int i = n - 2; i >= 1; --i)
        R(i) = (y[i - 1] - LU5(lu, i, 4) * R(i + 1) - LU5(lu, i, 5) * R(i + 2))
               / LU5(lu, i, 3);
#undef R
    free(y);
}

/* The k = 1 integration of poi_ux (loop_par.f:457-502) and of wz_wall
   (loop_par.f:379-418): x(i) from x(i-1), x(i-2) and d; sdx = -dx or dx. */
static void integrate_x(int n, double sdx, const or_cplx *d, size_t sd, or_cplx *x, size_t sx)
{
#define D(i) d[((i) - 1) * sd]
#define X(i) x[((i) - 1) * sx]
    X(2) = (sdx * (251.0 * D(1) + 646.0 * D(2) - 264.0 * D(3) + 106.0 * D(4) - 19.0 * D(5))
            + 720.0 * X(1)) / 720.0;
    for (int i = 3; i <= n - 2; ++i)
        X(i) = (sdx * (281.0 * D(i - 2) + 2056.0 * D(i - 1) + 1176.0 * D(i)
                       - 104.0 * D(i + 1) + 11.0 * D(i + 2)) / 90.0
                + 11.0 * X(i - 2) + 16.0 * X(i - 1)) / 27.0;
    int i = n - 1;
    X(i) = (sdx * (10.0 * D(i - 2) + 57.0 * D(i - 1) + 24.0 * D(i) - D(i + 1))
            + 33.0 * X(i - 2) + 24.0 * X(i - 1)) / 57.0;
    i = n;
    X(i) = (sdx * (251.0 * D(i) + 646.0 * D(i - 1) - 264.0 * D(i - 2) + 106.0 * D(i - 3)
                   - 19.0 * D(i - 4)) + 720.0 * X(i - 1)) / 720.0;
#undef D
#undef X
}

/* poi_ux: loop_par.f:428-556 */
int or_poi_ux(const or_params *p, const or_cplx *uy, const or_cplx *wy, or_cplx *ux,
              or_cplx *duydy)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    if (nx < 7 || ny < 8) return -1;
    const double dx = p->dx, dxx = p->dx * p->dx;
    int rc = or_deryfv(p, uy, duydy);
    if (rc) return rc;
    /* fundamental mode: continuity, rows 2..jmax-1 */
    for (int j = 2; j <= ny - 1; ++j)
        integrate_x(nx, -dx, &F3(duydy, 1, j, 1), 1, &F3(ux, 1, j, 1), 1);
    if (nk < 2) return 0;
    size_t n3 = (size_t)nx * ny * nk;
    or_cplx *d2 = malloc(sizeof(or_cplx) * n3);
    or_cplx *rhs = malloc(sizeof(or_cplx) * (size_t)nx);
    double *a = malloc(sizeof(double) * 10 * (size_t)nx), *lu = a + 5 * (size_t)nx;
    or_derparx(p, duydy, d2);
    for (int k = 2; k <= nk; ++k) {
        double k2b2 = dxx * k2b2_of(p, k);
        or_cplx kb = kb_of(p, k);
        double kk = (double)(k - 1) * (double)(k - 1) * p->beta * p->beta;
        or_lhs5(nx, k2b2, a);
        or_ludecomp(nx, a, lu);
        for (int j = 2; j <= ny; ++j) {
#define DD(i) F3(d2, i, j, k)
#define W(i) F3(wy, i, j, k)
            rhs[0] = F3(ux, 1, j, k);
            rhs[1] = -dxx * (DD(1) + 10.0 * DD(2) + DD(3)
                             + kb * (W(1) + 10.0 * W(2) + W(3))) / 12.0;
            for (int i = 3; i <= nx - 2; ++i)
                rhs[i - 1] = -8.0 * dxx * (DD(i - 1) + 5.5 * DD(i) + DD(i + 1)
                                           + kb * (W(i - 1) + 5.5 * W(i) + W(i + 1)));
            rhs[nx - 2] = -dxx * (DD(nx) + 10.0 * DD(nx - 1) + DD(nx - 2)
                                  + kb * (W(nx) + 10.0 * W(nx - 1) + W(nx - 2))) / 12.0;
            rhs[nx - 1] = (kb * W(nx) + DD(nx)) / kk;
#undef DD
#undef W
            or_lusolve(nx, lu, rhs, 1, 0);
            for (int i = 2; i <= nx; ++i) F3(ux, i, j, k) = rhs[i - 1];
        }
    }
    free(d2);
    free(rhs);
    free(a);
    return 0;
}

/* poi_uz: loop_par.f:559-688 (the k = 1 branch is commented out there) */
int or_poi_uz(const or_params *p, const or_cplx *wy, const or_cplx *duydy, or_cplx *uz)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    if (nx < 7 || ny < 8) return -1;
    if (nk < 2) return 0;
    const double dxx = p->dx * p->dx;
    size_t n3 = (size_t)nx * ny * nk;
    or_cplx *dw = malloc(sizeof(or_cplx) * n3);
    or_cplx *rhs = malloc(sizeof(or_cplx) * (size_t)nx);
    double *a = malloc(sizeof(double) * 10 * (size_t)nx), *lu = a + 5 * (size_t)nx;
    or_derparx(p, wy, dw);
    for (int k = 2; k <= nk; ++k) {
        double k2b2 = dxx * k2b2_of(p, k);
        or_cplx kb = kb_of(p, k);
        double kk = (double)(k - 1) * (double)(k - 1) * p->beta * p->beta;
        or_lhs5(nx, k2b2, a);
        or_ludecomp(nx, a, lu);
        for (int j = 2; j <= ny; ++j) {
#define DW(i) F3(dw, i, j, k)
#define DU(i) F3(duydy, i, j, k)
            rhs[0] = F3(uz, 1, j, k);
            rhs[1] = dxx * (DW(1) + 10.0 * DW(2) + DW(3)
                            - kb * (DU(1) + 10.0 * DU(2) + DU(3))) / 12.0;
            for (int i = 3; i <= nx - 2; ++i)
                rhs[i - 1] = 8.0 * dxx * (DW(i - 1) + 5.5 * DW(i) + DW(i + 1)
                                          - kb * (DU(i - 1) + 5.5 * DU(i) + DU(i + 1)));
            rhs[nx - 2] = dxx * (DW(nx) + 10.0 * DW(nx - 1) + DW(nx - 2)
                                 - kb * (DU(nx) + 10.0 * DU(nx - 1) + DU(nx - 2))) / 12.0;
            rhs[nx - 1] = (-DW(nx) + kb * DU(nx)) / kk;
#undef DW
#undef DU
            or_lusolve(nx, lu, rhs, 1, 0);
            for (int i = 2; i <= nx; ++i) F3(uz, i, j, k) = rhs[i - 1];
        }
    }
    free(dw);
    free(rhs);
    free(a);
    return 0;
}

/* der2wydxdy: loop_par.f:300-361 */
int or_der2wydxdy(const or_params *p, const or_cplx *wy, int k, or_cplx *d)
{
    const int nx = p->nx, ny = p->ny;
    if (nx < 8 || ny < 7) return -1;
    const double dx = p->dx, dy = p->dy0;
    const double *c = p->dwydy;
    or_cplx *w = malloc(sizeof(or_cplx) * (size_t)nx);
    for (int i = 1; i <= nx; ++i)
        w[i - 1] = (1.0 / dy) * (c[1] * F3(wy, i, 1, k) + c[2] * F3(wy, i, 2, k)
                                 + c[3] * F3(wy, i, 3, k) + c[4] * F3(wy, i, 4, k)
                                 + c[5] * F3(wy, i, 5, k) + c[6] * F3(wy, i, 6, k)
                                 + c[7] * F3(wy, i, 7, k));
#define Wd(i) w[(i) - 1]
    d[0] = (-14.7 * Wd(1) + 36.0 * Wd(2) - 45.0 * Wd(3) + 40.0 * Wd(4) - 22.5 * Wd(5)
            + 7.2 * Wd(6) - Wd(7)) / 6.0 / dx;
    d[1] = (-Wd(1) - 7.7 * Wd(2) + 15.0 * Wd(3) - 10.0 * Wd(4) + 5.0 * Wd(5) - 1.5 * Wd(6)
            + 0.2 * Wd(7)) / 6.0 / dx;
    d[2] = (0.2 * Wd(1) - 2.4 * Wd(2) - 3.5 * Wd(3) + 8.0 * Wd(4) - 3.0 * Wd(5)
            + 0.8 * Wd(6) - 0.1 * Wd(7)) / 6.0 / dx;
    for (int i = 4; i <= nx - 3; ++i)
        d[i - 1] = (45.0 * (Wd(i + 1) - Wd(i - 1)) + 9.0 * (Wd(i - 2) - Wd(i + 2)) + Wd(i + 3)
                    - Wd(i - 3)) / 60.0 / dx;
    d[nx - 3] = (Wd(nx - 4) - 8.0 * Wd(nx - 3) + 8.0 * Wd(nx - 1) - Wd(nx)) / 12.0 / dx;
    d[nx - 2] = (2.0 * Wd(nx) + 3.0 * Wd(nx - 1) - 6.0 * Wd(nx - 2) + Wd(nx - 3)) / 6.0 / dx;
    d[nx - 1] = (3.0 * Wd(nx) - 4.0 * Wd(nx - 1) + Wd(nx - 2)) / 2.0 / dx;
#undef Wd
    free(w);
    return 0;
}

/* wx_wall: loop_par.f:174-255 */
int or_wx_wall(const or_params *p, const or_cplx *uy, const or_cplx *wy,
               const or_cplx *d2uydx2, or_cplx *wx, or_cplx *lapv)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    if (nx < 8 || ny < 7) return -1;
    const double dxx = p->dx * p->dx, dyy = p->dy0 * p->dy0;
    const double *w = p->w_at_w;
    for (int k = 1; k <= nk; ++k)
        for (int i = 1; i <= nx; ++i)
            lapv[(i - 1) + (size_t)nx * (k - 1)] =
                d2uydx2[(i - 1) + (size_t)nx * (k - 1)]
                + (w[2] * F3(uy, i, 1, k) + w[3] * F3(uy, i, 2, k) + w[4] * F3(uy, i, 3, k)
                   + w[5] * F3(uy, i, 4, k) + w[6] * F3(uy, i, 5, k) + w[7] * F3(uy, i, 6, k)
                   + w[8] * F3(uy, i, 7, k)) / (dyy * w[0])
                + k2b2_of(p, k) * F3(uy, i, 1, k);
    or_cplx *d = malloc(sizeof(or_cplx) * (size_t)nx * 2), *rhs = d + nx;
    double *a = malloc(sizeof(double) * 10 * (size_t)nx), *lu = a + 5 * (size_t)nx;
    for (int k = 2; k <= nk; ++k) {
        double k2b2 = dxx * k2b2_of(p, k);
        or_cplx kb = kb_of(p, k);
        double kk = (double)(k - 1) * (double)(k - 1) * p->beta * p->beta;
        or_lhs5(nx, k2b2, a);
        or_ludecomp(nx, a, lu);
        or_der2wydxdy(p, wy, k, d);
        const or_cplx *L = lapv + (size_t)nx * (k - 1);
#define DD(i) d[(i) - 1]
#define LV(i) L[(i) - 1]
        rhs[0] = F3(wx, 1, 1, k);
        rhs[1] = dxx * (-DD(1) - 10.0 * DD(2) - DD(3) + kb * (LV(1) + 10.0 * LV(2) + LV(3)))
                 / 12.0;
        for (int i = 3; i <= nx - 2; ++i)
            rhs[i - 1] = 8.0 * dxx * (-DD(i - 1) - 5.5 * DD(i) - DD(i + 1)
                                      + kb * (LV(i - 1) + 5.5 * LV(i) + LV(i + 1)));
        rhs[nx - 2] = dxx * (-DD(nx) - 10.0 * DD(nx - 1) - DD(nx - 2)
                             + kb * (LV(nx) + 10.0 * LV(nx - 1) + LV(nx - 2))) / 12.0;
        rhs[nx - 1] = (DD(nx) - kb * LV(nx)) / kk;
#undef DD
#undef LV
        or_lusolve(nx, lu, rhs, 1, 1);
        for (int i = 2; i <= nx; ++i) F3(wx, i, 1, k) = rhs[i - 1];
    }
    free(d);
    free(a);
    return 0;
}

/* wz_wall: loop_par.f:364-425 */
int or_wz_wall(const or_params *p, const or_cplx *wx, const or_cplx *lapv, or_cplx *wz)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    if (nx < 7) return -1;
    or_cplx *rhs = malloc(sizeof(or_cplx) * (size_t)nx);
    for (int k = 1; k <= nk; ++k) {
        or_cplx kb = kb_of(p, k);
        for (int i = 1; i <= nx; ++i)
            rhs[i - 1] = kb * F3(wx, i, 1, k) - lapv[(i - 1) + (size_t)nx * (k - 1)];
        integrate_x(nx, p->dx, rhs, 1, &F3(wz, 1, 1, k), 1);
    }
    (void)ny;
    free(rhs);
    return 0;
}

/* bzone (loop_par.f:6-45) and bzonew (691-725) for my_form 1 and 2 */
int or_bzone(const or_params *p, const double *bdfc, int wall_row, or_cplx *wx, or_cplx *wy,
             or_cplx *wz, or_cplx *th)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    int j0 = wall_row ? 1 : 2, j1 = wall_row ? 1 : ny;
    for (int k = 1; k <= nk; ++k) {
        int bdt = (k == 3) ? 2 : 1;
        for (int j = j0; j <= j1; ++j)
            for (int i = 1; i <= nx; ++i) {
                double f = bdfc[(bdt - 1) + 2 * (size_t)(i - 1)];
                F3(wx, i, j, k) = F3(wx, i, j, k) * f;
                F3(wy, i, j, k) = F3(wy, i, j, k) * f;
                F3(wz, i, j, k) = F3(wz, i, j, k) * f;
                if (th && !wall_row) F3(th, i, j, k) = F3(th, i, j, k) * f;
            }
    }
    return 0;
}

/* filter_trid: filter_par.f:144-178; rhsf_trid 214-332; lhs_tridf 181-212;
   tridseqx derivs_par.f:1009-1037 */
int or_filter_trid(const or_params *p, or_cplx *wx, or_cplx *wy, or_cplx *wz)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    if (nx < 7) return -1;
    /* par.for:116-120 */
    const double alphaf = 0.48;
    const double af = (11.0 + 10.0 * alphaf) / 16.0;
    const double bf = (15.0 + 34.0 * alphaf) / 64.0;
    const double cf = (-3.0 + 6.0 * alphaf) / 32.0;
    const double df = (1.0 - 2.0 * alphaf) / 64.0;
    double *a = malloc(sizeof(double) * 4 * (size_t)nx), *b = a + nx, *c = b + nx,
           *gam = c + nx;
    for (int i = 0; i < nx; ++i) { a[i] = 0.0; b[i] = 1.0; c[i] = 0.0; }
    for (int i = 4; i <= nx - 3; ++i) { a[i - 1] = alphaf; c[i - 1] = alphaf; }
    or_cplx *r = malloc(sizeof(or_cplx) * (size_t)nx), *u = malloc(sizeof(or_cplx) * (size_t)nx);
    or_cplx *fields[3] = {wx, wy, wz};
    for (int q = 0; q < 3; ++q) {
        or_cplx *w = fields[q];
        for (int k = 1; k <= nk; ++k)
            for (int j = 1; j <= ny; ++j) {
#define Wf(i) F3(w, i, j, k)
                int n = nx;
                r[0] = (15.0 * Wf(1) + 4.0 * Wf(2) - 6.0 * Wf(3) + 4.0 * Wf(4) - Wf(5)) / 16.0;
                r[1] = (Wf(1) + 12.0 * Wf(2) + 6.0 * Wf(3) - 4.0 * Wf(4) + Wf(5)) / 16.0;
                r[2] = (-Wf(1) + 4.0 * Wf(2) + 10.0 * Wf(3) + 4.0 * Wf(4) - Wf(5)) / 16.0;
                for (int i = 4; i <= n - 3; ++i)
                    r[i - 1] = af * Wf(i) + bf * (Wf(i + 1) + Wf(i - 1))
                               + cf * (Wf(i + 2) + Wf(i - 2)) + df * (Wf(i + 3) + Wf(i - 3));
                r[n - 3] = (-Wf(n) + 4.0 * Wf(n - 1) + 10.0 * Wf(n - 2) + 4.0 * Wf(n - 3)
                            - Wf(n - 4)) / 16.0;
                r[n - 2] = (Wf(n) + 12.0 * Wf(n - 1) + 6.0 * Wf(n - 2) - 4.0 * Wf(n - 3)
                            + Wf(n - 4)) / 16.0;
                r[n - 1] = (15.0 * Wf(n) + 4.0 * Wf(n - 1) - 6.0 * Wf(n - 2) + 4.0 * Wf(n - 3)
                            - Wf(n - 4)) / 16.0;
                /* tridseqx */
                double bet = b[0];
                u[0] = r[0] / bet;
                for (int i = 1; i < n; ++i) {
                    gam[i] = c[i - 1] / bet;
                    bet = b[i] - a[i] * gam[i];
                    u[i] = (r[i] - a[i] * u[i - 1]) / bet;
                }
                for (int i = n - 2; i >= 0; --i) u[i] = u[i] - gam[i + 1] * u[i + 1];
                if (j >= 2)
                    for (int i = 2; i <= n; ++i) Wf(i) = u[i - 1];
#undef Wf
            }
    }
    free(a);
    free(r);
    free(u);
    return 0;
}

/* outuy: loop_par.f:48-119 (my_form 1 and 2: the default branch) */
int or_outuy(const or_params *p, const or_cplx *wz, or_cplx *uy, or_cplx *dwzdx)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    int rc = or_derparx(p, wz, dwzdx);
    if (rc) return rc;
    for (int k = 1; k <= nk; ++k)
        for (int j = 2; j <= ny; ++j) F3(uy, nx, j, k) = 0.0;
    return 0;
}

/* main_par.f:738-755 */
void or_bdfc(int nx, int i0, int i3, int i4, double *bdfc)
{
#define B(r, i) bdfc[((r) - 1) + 2 * (size_t)((i) - 1)]
    for (int i = 1; i <= i0; ++i) {
        double ep = (double)(i - 1) / (double)(i0 - 1);
        B(1, i) = ((6.0 * ep - 15.0) * ep + 10.0) * (ep * ep * ep);
        B(2, i) = B(1, i);
    }
    for (int i = i0; i <= i3; ++i) {
        B(1, i) = 1.0;
        B(2, i) = B(1, i);
    }
    for (int i = i3; i <= i4; ++i) {
        double ep = (double)(i - i3) / (double)(i4 - i3);
        double e2 = ep * ep, e4 = e2 * e2, e8 = e4 * e4, e16 = e8 * e8, e32 = e16 * e16;
        double e50 = e32 * e16 * e2;
        double t = 1.0 - e50;
        B(1, i) = (t * t) * (t * t) * exp(-e4 / 10.0);
        B(2, i) = B(1, i);
    }
    for (int i = i4; i <= nx; ++i) {
        B(1, i) = 0.0;
        B(2, i) = B(1, i);
    }
#undef B
}

/* loop: main_par.f:1762-1785 */
int or_loop(const or_params *p, double erro, const double *bdfc, const or_cplx *d2uydx2,
            or_cplx *ux, or_cplx *uy, or_cplx *uz, or_cplx *wx, or_cplx *wy, or_cplx *wz,
            or_cplx *th, int order, int *cycles, double *res)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    size_t n3 = (size_t)nx * ny * nk;
    int rc = 0;
    if (erro < 1e-5) rc = or_filter_trid(p, wx, wy, wz);
    if (rc) return rc;
    or_bzone(p, bdfc, 0, wx, wy, wz, th);
    or_cplx *t = malloc(sizeof(or_cplx) * n3);
    or_cplx *lapv = malloc(sizeof(or_cplx) * (size_t)nx * nk);
    rc = or_outuy(p, wz, uy, t);
    if (!rc) rc = or_poi_uy(p, wx, t, uy, erro, 11, order, cycles, res);
    if (!rc) rc = or_poi_ux(p, uy, wy, ux, t);
    if (!rc) rc = or_poi_uz(p, wy, t, uz);
    if (!rc) rc = or_wx_wall(p, uy, wy, d2uydx2, wx, lapv);
    if (!rc) rc = or_wz_wall(p, wx, lapv, wz);
    if (!rc) rc = or_bzone(p, bdfc, 1, wx, wy, wz, th);
    free(t);
    free(lapv);
    return rc;
}
