/*
 * CPU oracle for the trans_bl v-Poisson / compact-derivative hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see tbl_oracle.h).  A plain-C, single-threaded
 * FP64 restatement of the reference's algorithm at one MPI rank; every
 * function cites the reference lines it follows.  Expressions keep the
 * reference's evaluation order so that rounding matches it closely.
 *
 * Built with -ffp-contract=off (no fused multiply-add), see oracle/Makefile.
 */
#include "tbl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 1-based (I,J) access into an (nx_l, ny_l) Fortran-order plane */
#define AT(a, nxl, I, J) ((a)[((I) - 1) + (size_t)(nxl) * ((J) - 1)])

/* single-precision literal TINY = 1.e-20 assigned to real*8
   (penta_par.f:317) */
static const double OR_TINY = (double)1.e-20f;

static double ipow(double x, int n)
{
    /* x**n for integer n >= 0 by repeated squaring */
    double r = 1.0, b = x;
    while (n > 0) {
        if (n & 1) r *= b;
        b *= b;
        n >>= 1;
    }
    return r;
}

/* ------------------------------------------------------------------------ */
/* create_ctes: main_par.f:1828-1892                                         */
/* ------------------------------------------------------------------------ */
int or_create_ctes(const or_params *p, or_metrics *m)
{
    if (p->msh < 1 || p->msh > OR_MAXLVL || p->ny > 1024) return -1;
    memset(m, 0, sizeof(*m));
    for (int l = 1; l <= p->msh; ++l) {
        int e = 1 << (l - 1);
        if (p->stf != 1.0)
            m->dy0[l - 1] = p->dy0 * ((ipow(p->stf, e) - 1.0) / (p->stf - 1.0));
        else
            m->dy0[l - 1] = p->dy0 * (double)e;
        m->stf[l - 1] = ipow(p->stf, e);
        double t = p->dx * (double)e;
        m->dx2[l - 1] = t * t;
    }
    for (int j = 1; j <= p->ny - 1; ++j) {
        if (p->stf != 1.0)
            m->qdy[j - 1] = 1.0 / (m->dy0[0] * ipow(m->stf[0], j - 1));
        else
            m->qdy[j - 1] = 1.0 / m->dy0[0];
    }
    for (int l = 1; l <= p->msh; ++l) {
        int nj = (p->ny - 1) / (1 << (l - 1));
        for (int j = 1; j <= nj; ++j) {
            double aux = (p->stf != 1.0) ? m->dy0[l - 1] * ipow(m->stf[l - 1], j - 1)
                                         : m->dy0[l - 1];
            m->qdy2[l - 1][j - 1] = 1.0 / (aux * aux);
        }
    }
    m->ptsx[0] = p->nx;
    m->ptsy[0] = p->ny;
    for (int l = 2; l <= p->msh; ++l) {
        m->ptsx[l - 1] = (m->ptsx[l - 2] + 1) / 2;
        m->ptsy[l - 1] = (m->ptsy[l - 2] + 1) / 2;
    }
    return 0;
}

static double k2b2_of(const or_params *p, int k /*1-based*/)
{
    return -(double)(k - 1) * (double)(k - 1) * p->beta * p->beta;
}

static or_cplx kb_of(const or_params *p, int k)
{
    return -((double)(k - 1) * p->beta) * I;
}

/* ------------------------------------------------------------------------ */
/* derivatives: derivs_par.f:80-177, 259-611, 947-975, 1009-1210             */
/* ------------------------------------------------------------------------ */

/* Thomas solve along one line with stride (tridy 947-975 / tridseqx
   1009-1037; tridparx 775-858 reduces to it at one rank) */
static void thomas(const double *a, const double *b, const double *c, int n,
                   or_cplx *r, size_t stride, double *gam, or_cplx *u)
{
    double bet = b[0];
    u[0] = r[0] / bet;
    for (int j = 1; j < n; ++j) {
        gam[j] = c[j - 1] / bet;
        bet = b[j] - a[j] * gam[j];
        u[j] = (r[(size_t)j * stride] - a[j] * u[j - 1]) / bet;
    }
    for (int j = n - 2; j >= 0; --j) u[j] = u[j] - gam[j + 1] * u[j + 1];
    for (int j = 0; j < n; ++j) r[(size_t)j * stride] = u[j];
}

/* coefx / coeffx: derivs_par.f:1145-1210 */
static void coef_x(int n, int second, double *a, double *b, double *c)
{
    for (int i = 0; i < n; ++i) {
        if (second) { a[i] = 2; b[i] = 11; c[i] = 2; }
        else        { a[i] = 1; b[i] = 3;  c[i] = 1; }
    }
    if (second) {
        a[0] = 0;   b[0] = 13;  c[0] = 137;
        a[1] = 1;   b[1] = 12;  c[1] = 3;
        a[n-2] = 3; b[n-2] = 12; c[n-2] = 1;
        a[n-1] = 137; b[n-1] = 13; c[n-1] = 0;
    } else {
        a[0] = 0;   b[0] = 1;  c[0] = 0;
        a[1] = 1;   b[1] = 6;  c[1] = 2;
        a[n-2] = 2; b[n-2] = 6; c[n-2] = 1;
        a[n-1] = 4; b[n-1] = 1; c[n-1] = 0;
    }
}

/* coefy / coefy_fv / coeffy: derivs_par.f:1040-1142 */
static void coef_y(const or_params *p, int kind, double *a, double *b, double *c)
{
    int n = p->ny;
    const double *sp, *cp, *pp, *lp;
    if (kind == 2) { sp = p->sp_sd; cp = p->cp_sd; pp = p->pp_sd; lp = p->lp_sd; }
    else           { sp = p->sp_fd; cp = p->cp_fd; pp = p->pp_fd; lp = p->lp_fd; }
    for (int j = 0; j < n; ++j) { a[j] = cp[0]; b[j] = cp[1]; c[j] = cp[2]; }
    if (kind == 1)      { a[0] = 0; b[0] = p->fp_fd_e[0]; c[0] = 0; }
    else if (kind == 3) { a[0] = 0; b[0] = 1.0;           c[0] = 0; }
    else                { a[0] = 0; b[0] = p->fp_sd[0];   c[0] = p->fp_sd[1]; }
    a[1] = sp[0]; b[1] = sp[1]; c[1] = sp[2];
    a[n-2] = pp[2]; b[n-2] = pp[1]; c[n-2] = pp[0];
    a[n-1] = lp[1]; b[n-1] = lp[0]; c[n-1] = 0;
}

static int deriv_x(const or_params *p, const or_cplx *f, or_cplx *out, int second)
{
    int nx = p->nx, ny = p->ny, nk = p->nk;
    if (nx < 7) return -1;
    double dx = p->dx, dxx = p->dx * p->dx;
    double *a = malloc(sizeof(double) * nx * 4);
    or_cplx *u = malloc(sizeof(or_cplx) * nx);
    double *b = a + nx, *c = b + nx, *gam = c + nx;
    coef_x(nx, second, a, b, c);
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < ny; ++j) {
            const or_cplx *fl = f + (size_t)nx * (j + (size_t)ny * k);
            or_cplx *r = out + (size_t)nx * (j + (size_t)ny * k);
#define FX(I) fl[(I) - 1]
            int n = nx;
            if (!second) {          /* rhsx: derivs_par.f:259-302 */
                r[0] = (-25.0 * FX(1) + 48.0 * FX(2) - 36.0 * FX(3) + 16.0 * FX(4)
                        - 3.0 * FX(5)) / (12.0 * dx);
                r[1] = (-406.0 * FX(1) - 300.0 * FX(2) + 760.0 * FX(3) - 80.0 * FX(4)
                        + 30.0 * FX(5) - 4.0 * FX(6)) / (120.0 * dx);
                for (int i = 3; i <= n - 2; ++i)
                    r[i - 1] = (FX(i + 2) - FX(i - 2) + 28.0 * (FX(i + 1) - FX(i - 1)))
                               / (12.0 * dx);
                r[n - 2] = (-406.0 * FX(n) - 300.0 * FX(n - 1) + 760.0 * FX(n - 2)
                            - 80.0 * FX(n - 3) + 30.0 * FX(n - 4) - 4.0 * FX(n - 5))
                           / (-120.0 * dx);
                r[n - 1] = (-74.0 * FX(n) + 16.0 * FX(n - 1) + 72.0 * FX(n - 2)
                            - 16.0 * FX(n - 3) + 2.0 * FX(n - 4)) / (-24.0 * dx);
            } else {                /* rhsxx: derivs_par.f:351-397 */
                r[0] = (9775.0 * FX(1) - 20285.0 * FX(2) + 11170.0 * FX(3)
                        - 550.0 * FX(4) - 145.0 * FX(5) + 35.0 * FX(6)) / (60.0 * dxx);
                r[1] = (4834.0 * FX(1) - 8424.0 * FX(2) + 1890.0 * FX(3)
                        + 2320.0 * FX(4) - 810.0 * FX(5) + 216.0 * FX(6)
                        - 26.0 * FX(7)) / (360.0 * dxx);
                for (int i = 3; i <= n - 2; ++i)
                    r[i - 1] = (3.0 * (FX(i + 2) + FX(i - 2)) + 48.0 * (FX(i + 1) + FX(i - 1))
                                - 102.0 * FX(i)) / (4.0 * dxx);
                r[n - 2] = (4834.0 * FX(n) - 8424.0 * FX(n - 1) + 1890.0 * FX(n - 2)
                            + 2320.0 * FX(n - 3) - 810.0 * FX(n - 4) + 216.0 * FX(n - 5)
                            - 26.0 * FX(n - 6)) / (360.0 * dxx);
                r[n - 1] = (9775.0 * FX(n) - 20285.0 * FX(n - 1) + 11170.0 * FX(n - 2)
                            - 550.0 * FX(n - 3) - 145.0 * FX(n - 4) + 35.0 * FX(n - 5))
                           / (60.0 * dxx);
            }
#undef FX
            thomas(a, b, c, nx, r, 1, gam, u);
        }
    free(a);
    free(u);
    return 0;
}

int or_derparx(const or_params *p, const or_cplx *f, or_cplx *out)
{
    return deriv_x(p, f, out, 0);
}

int or_derparxx(const or_params *p, const or_cplx *f, or_cplx *out)
{
    return deriv_x(p, f, out, 1);
}

/* kind: 1 = dery (rhsy 400-449), 2 = deryy (rhsyy 552-611),
         3 = deryfv (rhsyfv 504-549) */
static int deriv_y(const or_params *p, const or_cplx *f, or_cplx *out, int kind)
{
    int nx = p->nx, ny = p->ny, nk = p->nk;
    if (ny < 8) return -1;
    or_metrics m;
    if (or_create_ctes(p, &m)) return -1;
    const double *qdy = m.qdy, *qdy2 = m.qdy2[0];
    double *a = malloc(sizeof(double) * ny * 4);
    or_cplx *u = malloc(sizeof(or_cplx) * ny);
    double *b = a + ny, *c = b + ny, *gam = c + ny;
    coef_y(p, kind, a, b, c);
    size_t sj = (size_t)nx;
    for (int k = 0; k < nk; ++k)
        for (int i = 0; i < nx; ++i) {
            const or_cplx *fl = f + i + (size_t)nx * ny * k;
            or_cplx *r = out + i + (size_t)nx * ny * k;
#define FY(J) fl[((J) - 1) * sj]
#define RY(J) r[((J) - 1) * sj]
            int n = ny;
            if (kind == 1 || kind == 3) {
                const double *sp = p->sp_fd, *cp = p->cp_fd, *pp = p->pp_fd, *lp = p->lp_fd;
                if (kind == 1) {
                    const double *fe = p->fp_fd_e;
                    RY(1) = qdy[0] * (fe[1] * FY(1) + fe[2] * FY(2) + fe[3] * FY(3)
                                      + fe[4] * FY(4) + fe[5] * FY(5));
                } else {
                    RY(1) = 0.0;
                }
                RY(2) = qdy[0] * (sp[3] * FY(1) + sp[4] * FY(2) + sp[5] * FY(3)
                                  + sp[6] * FY(4) + sp[7] * FY(5) + sp[8] * FY(6));
                for (int j = 3; j <= n - 2; ++j)
                    RY(j) = qdy[j - 3] * (cp[3] * FY(j - 2) + cp[4] * FY(j - 1)
                                          + cp[5] * FY(j) + cp[6] * FY(j + 1)
                                          + cp[7] * FY(j + 2));
                RY(n - 1) = qdy[n - 6] * (pp[3] * FY(n) + pp[4] * FY(n - 1)
                                          + pp[5] * FY(n - 2) + pp[6] * FY(n - 3)
                                          + pp[7] * FY(n - 4) + pp[8] * FY(n - 5));
                RY(n) = qdy[n - 5] * (lp[2] * FY(n) + lp[3] * FY(n - 1)
                                      + lp[4] * FY(n - 2) + lp[5] * FY(n - 3)
                                      + lp[6] * FY(n - 4));
            } else {
                const double *fp = p->fp_sd, *sp = p->sp_sd, *cp = p->cp_sd,
                             *pp = p->pp_sd, *lp = p->lp_sd;
                RY(1) = (fp[2] * FY(1) + fp[3] * FY(2) + fp[4] * FY(3) + fp[5] * FY(4)
                         + fp[6] * FY(5) + fp[7] * FY(6)) * qdy2[0];
                RY(2) = (sp[3] * FY(1) + sp[4] * FY(2) + sp[5] * FY(3) + sp[6] * FY(4)
                         + sp[7] * FY(5) + sp[8] * FY(6) + sp[9] * FY(7)) * qdy2[0];
                for (int j = 3; j <= n - 2; ++j)
                    RY(j) = (cp[3] * FY(j - 2) + cp[4] * FY(j - 1) + cp[5] * FY(j)
                             + cp[6] * FY(j + 1) + cp[7] * FY(j + 2)) * qdy2[j - 3];
                RY(n - 1) = (pp[3] * FY(n) + pp[4] * FY(n - 1) + pp[5] * FY(n - 2)
                             + pp[6] * FY(n - 3) + pp[7] * FY(n - 4) + pp[8] * FY(n - 5)
                             + pp[9] * FY(n - 6)) * qdy2[n - 7];
                RY(n) = (lp[2] * FY(n) + lp[3] * FY(n - 1) + lp[4] * FY(n - 2)
                         + lp[5] * FY(n - 3) + lp[6] * FY(n - 4) + lp[7] * FY(n - 5))
                        * qdy2[n - 6];
            }
#undef FY
#undef RY
            thomas(a, b, c, ny, r, sj, gam, u);
        }
    free(a);
    free(u);
    return 0;
}

int or_dery(const or_params *p, const or_cplx *f, or_cplx *out)
{
    return deriv_y(p, f, out, 1);
}

int or_deryy(const or_params *p, const or_cplx *f, or_cplx *out)
{
    return deriv_y(p, f, out, 2);
}

int or_deryfv(const or_params *p, const or_cplx *f, or_cplx *out)
{
    return deriv_y(p, f, out, 3);
}

/* ------------------------------------------------------------------------ */
/* banded LU without pivoting: band5_poi / banbk5_poi, penta_par.f:308-404   */
/* a is n x 5, row j holds the coefficients of v(j-2..j+2)                   */
/* ------------------------------------------------------------------------ */
#define A5(r, c) a[((r) - 1) * 5 + ((c) - 1)]
#define AL(r, c) al[((r) - 1) * 2 + ((c) - 1)]

void or_band5_factor(double *a, int n, double *al)
{
    const int m1 = 2, mm = 5;
    int l = m1;
    for (int i = 1; i <= m1; ++i) {         /* left-justify the first rows */
        for (int j = m1 + 2 - i; j <= mm; ++j) A5(i, j - l) = A5(i, j);
        l -= 1;
        for (int j = mm - l; j <= mm; ++j) A5(i, j) = 0.0;
    }
    l = m1;
    for (int k = 1; k <= n; ++k) {
        if (l < n) l += 1;
        if (A5(k, 1) == 0.0) A5(k, 1) = OR_TINY;
        for (int i = k + 1; i <= l; ++i) {
            double dum = A5(i, 1) / A5(k, 1);
            AL(k, i - k) = dum;
            for (int j = 2; j <= mm; ++j) A5(i, j - 1) = A5(i, j) - dum * A5(k, j);
            A5(i, mm) = 0.0;
        }
    }
}

void or_band5_solve(const double *a, int n, const double *al, or_cplx *rhs)
{
    int l = 2;
    for (int k = 1; k <= n; ++k) {
        if (l < n) l += 1;
        for (int i = k + 1; i <= l; ++i) rhs[i - 1] = rhs[i - 1] - AL(k, i - k) * rhs[k - 1];
    }
    l = 1;
    for (int i = n; i >= 1; --i) {
        or_cplx dum = rhs[i - 1];
        for (int k = 2; k <= l; ++k) dum = dum - A5(i, k) * rhs[k + i - 2];
        rhs[i - 1] = dum / A5(i, 1);
        if (l < 5) l += 1;
    }
}
#undef A5
#undef AL

/* ------------------------------------------------------------------------ */
/* Poisson operator pieces                                                   */
/* ------------------------------------------------------------------------ */

/* coef_center / coef_border: poisson_par.f:229-340.  border selects
   qx = -15/(12 dx^2) instead of -30/(12 dx^2). */
static void line_matrix(const or_params *p, const or_metrics *m, int lvl, double k2b2,
                        int border, double *a /* ny_l x 5 */)
{
    int L = lvl - 1, n = m->ptsy[L];
    const double *sp = p->sp_poi[L], *cp = p->cp_poi[L], *pp = p->pp_poi[L],
                 *lp = p->lp_poi[L], *q2 = m->qdy2[L];
    double qx = (border ? -15.0 : -30.0) / (12.0 * m->dx2[L]);
    double add = (qx + k2b2) / p->fac_y;
#define R5(r, c) a[((r) - 1) * 5 + ((c) - 1)]
    memset(a, 0, sizeof(double) * 5 * n);
    R5(1, 3) = 1.0;
    R5(2, 2) = sp[3] * q2[0];
    R5(2, 3) = sp[4] * q2[0] + sp[1] * add;
    R5(2, 4) = sp[5] * q2[0] + sp[2] * add;
    R5(2, 5) = sp[6] * q2[0];
    for (int j = 3; j <= n - 2; ++j) {
        R5(j, 1) = cp[3] * q2[j - 3];
        R5(j, 2) = cp[4] * q2[j - 3] + cp[0] * add;
        R5(j, 3) = cp[5] * q2[j - 3] + cp[1] * add;
        R5(j, 4) = cp[6] * q2[j - 3] + cp[2] * add;
        R5(j, 5) = cp[7] * q2[j - 3];
    }
    int j = n;
    R5(j - 1, 1) = 0.0;
    R5(j - 1, 2) = pp[1] * q2[j - 3];
    R5(j - 1, 3) = pp[2] * q2[j - 3] + add;
    R5(j - 1, 4) = pp[3] * q2[j - 3];
    R5(j - 1, 5) = 0.0;
    R5(j, 1) = lp[4] * q2[j - 3];
    R5(j, 2) = lp[3] * q2[j - 3];
    R5(j, 3) = lp[2] * q2[j - 3] + add
               - lp[1] * sqrt(q2[j - 3]) * sqrt((p->alpha * p->alpha - k2b2) / p->fac_y);
    R5(j, 4) = 0.0;
    R5(j, 5) = 0.0;
#undef R5
}

/* x-direction stencils.  Full operator (defect): centre
   -v(i-2)+16v(i-1)-30v(i)+16v(i+1)-v(i+2); left border (i=2)
   10v(1)-15v(2)-4v(3)+14v(4)-6v(5)+v(6); right border mirrored
   (poisson_par.f:563-781). */
static or_cplx xs_full(const or_cplx *v, int nxl, int i, int j, int side, int grouped)
{
    if (side < 0)
        return 10.0 * AT(v, nxl, i - 1, j) - 15.0 * AT(v, nxl, i, j)
               - 4.0 * AT(v, nxl, i + 1, j) + 14.0 * AT(v, nxl, i + 2, j)
               - 6.0 * AT(v, nxl, i + 3, j) + AT(v, nxl, i + 4, j);
    if (side > 0)
        return 10.0 * AT(v, nxl, i + 1, j) - 15.0 * AT(v, nxl, i, j)
               - 4.0 * AT(v, nxl, i - 1, j) + 14.0 * AT(v, nxl, i - 2, j)
               - 6.0 * AT(v, nxl, i - 3, j) + AT(v, nxl, i - 4, j);
    if (grouped)   /* rows ny-1, ny of the interior columns (741-767) */
        return -1.0 * (AT(v, nxl, i - 2, j) + AT(v, nxl, i + 2, j))
               + 16.0 * (AT(v, nxl, i - 1, j) + AT(v, nxl, i + 1, j))
               - 30.0 * AT(v, nxl, i, j);
    return -1.0 * AT(v, nxl, i - 2, j) + 16.0 * AT(v, nxl, i - 1, j)
           - 30.0 * AT(v, nxl, i, j) + 16.0 * AT(v, nxl, i + 1, j)
           - 1.0 * AT(v, nxl, i + 2, j);
}

/* Off-diagonal x-stencil used on the smoother's right-hand side
   (rhs_border1/2, rhs_center: poisson_par.f:343-540). */
static or_cplx xs_off(const or_cplx *v, int nxl, int i, int j, int side, int grouped)
{
    if (side < 0)
        return 10.0 * AT(v, nxl, i - 1, j) - 4.0 * AT(v, nxl, i + 1, j)
               + 14.0 * AT(v, nxl, i + 2, j) - 6.0 * AT(v, nxl, i + 3, j)
               + AT(v, nxl, i + 4, j);
    if (side > 0)
        return 10.0 * AT(v, nxl, i + 1, j) - 4.0 * AT(v, nxl, i - 1, j)
               + 14.0 * AT(v, nxl, i - 2, j) - 6.0 * AT(v, nxl, i - 3, j)
               + AT(v, nxl, i - 4, j);
    if (grouped)
        return -(AT(v, nxl, i + 2, j) + AT(v, nxl, i - 2, j))
               + 16.0 * (AT(v, nxl, i + 1, j) + AT(v, nxl, i - 1, j));
    return -1.0 * AT(v, nxl, i - 2, j) + 16.0 * AT(v, nxl, i - 1, j)
           + 16.0 * AT(v, nxl, i + 1, j) - 1.0 * AT(v, nxl, i + 2, j);
}

/* defect: poisson_par.f:543-793 */
int or_defect(const or_params *p, const or_metrics *m, int lvl, int k,
              const or_cplx *v, const or_cplx *S, or_cplx *out, double si)
{
    int L = lvl - 1, nxl = m->ptsx[L], nyl = m->ptsy[L];
    const double *sp = p->sp_poi[L], *cp = p->cp_poi[L], *pp = p->pp_poi[L],
                 *lp = p->lp_poi[L], *q2 = m->qdy2[L];
    double k2b2 = k2b2_of(p, k), dx2 = m->dx2[L], fy = p->fac_y;
    double alpha2 = sqrt((p->alpha * p->alpha - k2b2) / fy);
    for (int i = 2; i <= nxl - 1; ++i) {
        int side = (i == 2) ? -1 : (i == nxl - 1) ? 1 : 0;
        or_cplx b, c, d;
        int j = 2;
        b = sp[2] * xs_full(v, nxl, i, j + 1, side, 0) + sp[1] * xs_full(v, nxl, i, j, side, 0);
        c = sp[3] * AT(v, nxl, i, j - 1) + sp[4] * AT(v, nxl, i, j)
            + sp[5] * AT(v, nxl, i, j + 1) + sp[6] * AT(v, nxl, i, j + 2);
        d = k2b2 * (sp[1] * AT(v, nxl, i, j) + sp[2] * AT(v, nxl, i, j + 1));
        AT(out, nxl, i, j) = AT(S, nxl, i, j)
                             + si * (b / (12.0 * dx2 * fy) + c * q2[0] + d / fy);
        for (j = 3; j <= nyl - 2; ++j) {
            b = cp[0] * xs_full(v, nxl, i, j - 1, side, 0)
                + cp[1] * xs_full(v, nxl, i, j, side, 0)
                + cp[2] * xs_full(v, nxl, i, j + 1, side, 0);
            c = cp[3] * AT(v, nxl, i, j - 2) + cp[4] * AT(v, nxl, i, j - 1)
                + cp[5] * AT(v, nxl, i, j) + cp[6] * AT(v, nxl, i, j + 1)
                + cp[7] * AT(v, nxl, i, j + 2);
            d = k2b2 * (cp[0] * AT(v, nxl, i, j - 1) + cp[1] * AT(v, nxl, i, j)
                        + cp[2] * AT(v, nxl, i, j + 1));
            AT(out, nxl, i, j) = AT(S, nxl, i, j)
                                 + si * (b / (12.0 * dx2 * fy) + c * q2[j - 3] + d / fy);
        }
        j = nyl - 1;
        b = xs_full(v, nxl, i, j, side, 1);
        c = pp[3] * AT(v, nxl, i, j + 1) + pp[2] * AT(v, nxl, i, j)
            + pp[1] * AT(v, nxl, i, j - 1);
        d = k2b2 * AT(v, nxl, i, j);
        AT(out, nxl, i, j) = AT(S, nxl, i, j)
                             + si * (b / (12.0 * dx2 * fy) + c * q2[j - 2] + d / fy);
        j = nyl;
        b = xs_full(v, nxl, i, j, side, 1);
        c = lp[2] * AT(v, nxl, i, j) + lp[3] * AT(v, nxl, i, j - 1)
            + lp[4] * AT(v, nxl, i, j - 2);
        d = k2b2 * AT(v, nxl, i, j);
        AT(out, nxl, i, j) = AT(S, nxl, i, j)
                             + si * (b / (12.0 * dx2 * fy) + c * q2[j - 3] + d / fy
                                     - lp[1] * alpha2 * AT(v, nxl, i, j) * sqrt(q2[j - 3]));
    }
    return 0;
}

/* function_source: poisson_par.f:111-164 (my_form = 1) */
int or_source(const or_params *p, int k, const or_cplx *wx, const or_cplx *dwzdx,
              or_cplx *S)
{
    int nx = p->nx, ny = p->ny;
    const double *sp = p->sp_poi[0], *cp = p->cp_poi[0];
    or_cplx kb = kb_of(p, k);
    size_t off = (size_t)nx * ny * (k - 1);
    const or_cplx *w = wx + off, *dz = dwzdx + off;
#define G(I, J) (kb * AT(w, nx, I, J) - AT(dz, nx, I, J))
    for (int i = 1; i <= nx; ++i) {
        AT(S, nx, i, 1) = 0.0;
        AT(S, nx, i, 2) = sp[1] * G(i, 2) + sp[2] * G(i, 3);
        for (int j = 3; j <= ny - 2; ++j)
            AT(S, nx, i, j) = cp[0] * G(i, j - 1) + cp[1] * G(i, j) + cp[2] * G(i, j + 1);
        AT(S, nx, i, ny - 1) = G(i, ny - 1);
        AT(S, nx, i, ny) = 0.0;
    }
#undef G
    return 0;
}

/* one line relaxation (poisson_par.f:197-219) */
static void relax_line(const or_params *p, const or_metrics *m, int lvl, or_cplx *v,
                       const or_cplx *S, int i, int side, const double *a,
                       const double *al, double rf, or_cplx *rhs)
{
    int L = lvl - 1, nxl = m->ptsx[L], n = m->ptsy[L];
    const double *sp = p->sp_poi[L], *cp = p->cp_poi[L];
    double qx = 1.0 / (12.0 * m->dx2[L] * p->fac_y);
    rhs[0] = AT(v, nxl, i, 1);
    int j = 2;
    rhs[j - 1] = AT(S, nxl, i, j) - qx * (sp[2] * xs_off(v, nxl, i, j + 1, side, 0)
                                          + sp[1] * xs_off(v, nxl, i, j, side, 0));
    for (j = 3; j <= n - 2; ++j)
        rhs[j - 1] = AT(S, nxl, i, j) - qx * (cp[0] * xs_off(v, nxl, i, j - 1, side, 0)
                                              + cp[1] * xs_off(v, nxl, i, j, side, 0)
                                              + cp[2] * xs_off(v, nxl, i, j + 1, side, 0));
    for (j = n - 1; j <= n; ++j)
        rhs[j - 1] = AT(S, nxl, i, j) - qx * xs_off(v, nxl, i, j, side, 1);
    or_band5_solve(a, n, al, rhs);
    for (j = 2; j <= n; ++j)
        AT(v, nxl, i, j) = AT(v, nxl, i, j) + rf * (rhs[j - 1] - AT(v, nxl, i, j));
}

/* sor: poisson_par.f:167-226 */
int or_sor(const or_params *p, const or_metrics *m, int lvl, int k,
           or_cplx *v, const or_cplx *S, int itera, int order)
{
    int L = lvl - 1, nxl = m->ptsx[L], n = m->ptsy[L];
    double rf = (itera == 1) ? 1.0 : 1.2;
    double k2b2 = k2b2_of(p, k);
    double *ac = malloc(sizeof(double) * n * 14);
    double *acb = ac + 5 * n, *alc = acb + 5 * n, *alcb = alc + 2 * n;
    or_cplx *rhs = malloc(sizeof(or_cplx) * n);
    memset(alc, 0, sizeof(double) * 4 * n);
    line_matrix(p, m, lvl, k2b2, 0, ac);
    or_band5_factor(ac, n, alc);
    line_matrix(p, m, lvl, k2b2, 1, acb);
    or_band5_factor(acb, n, alcb);
    for (int it = 0; it < itera; ++it) {
        relax_line(p, m, lvl, v, S, 2, -1, acb, alcb, rf, rhs);
        relax_line(p, m, lvl, v, S, nxl - 1, 1, acb, alcb, rf, rhs);
        if (order == 0) {
            for (int i = 3; i <= nxl - 2; ++i)
                relax_line(p, m, lvl, v, S, i, 0, ac, alc, rf, rhs);
        } else {
            for (int c = 0; c < 3; ++c)
                for (int i = 3 + c; i <= nxl - 2; i += 3)
                    relax_line(p, m, lvl, v, S, i, 0, ac, alc, rf, rhs);
        }
    }
    free(ac);
    free(rhs);
    return 0;
}

/* restrict: poisson_par.f:796-874.  vf, rf: fine (level lvl); vc, wc, Sc:
   coarse (level lvl+1).  rf holds the fine residual on i=2..nx-1,
   j=2..ny. */
int or_restrict(const or_params *p, const or_metrics *m, int lvl,
                const or_cplx *vf, const or_cplx *rf, or_cplx *vc, or_cplx *wc,
                or_cplx *Sc)
{
    (void)p;
    int L = lvl - 1, nxf = m->ptsx[L], nyf = m->ptsy[L];
    int nxc = m->ptsx[L + 1], nyc = m->ptsy[L + 1];
    (void)nyf;
    for (int j = 2; j <= nyc - 1; ++j) {
        int fj = 2 * j - 1;
        for (int i = 2; i <= nxc - 1; ++i) {
            int fi = 2 * i - 1;
            AT(vc, nxc, i, j) = AT(vf, nxf, fi, fj);
            AT(wc, nxc, i, j) = AT(vc, nxc, i, j);
            AT(Sc, nxc, i, j) =
                6.25e-2 * (2.0 * (AT(rf, nxf, fi - 1, fj) + AT(rf, nxf, fi + 1, fj)
                                  + AT(rf, nxf, fi, fj - 1) + AT(rf, nxf, fi, fj + 1))
                           + 4.0 * AT(rf, nxf, fi, fj) + AT(rf, nxf, fi + 1, fj + 1)
                           + AT(rf, nxf, fi + 1, fj - 1) + AT(rf, nxf, fi - 1, fj + 1)
                           + AT(rf, nxf, fi - 1, fj - 1));
        }
    }
    for (int i = 2; i <= nxc - 1; ++i) {
        int fi = 2 * i - 1;
        AT(vc, nxc, i, 1) = AT(vf, nxf, fi, 1);
        AT(wc, nxc, i, 1) = AT(vc, nxc, i, 1);
        int j = nyc, fj = 2 * j - 1;
        AT(vc, nxc, i, j) = AT(vf, nxf, fi, fj);
        AT(wc, nxc, i, j) = AT(vc, nxc, i, j);
        AT(Sc, nxc, i, j) = 0.25 * (AT(rf, nxf, fi - 1, fj) + 2.0 * AT(rf, nxf, fi, fj)
                                    + AT(rf, nxf, fi + 1, fj));
    }
    for (int side = 0; side < 2; ++side) {
        int i = side ? nxc : 1, fi = 2 * i - 1;
        for (int j = 1; j <= nyc; ++j) {
            AT(vc, nxc, i, j) = AT(vf, nxf, fi, 2 * j - 1);
            AT(wc, nxc, i, j) = AT(vc, nxc, i, j);
        }
    }
    return 0;
}

/* correction + intpol: poisson_par.f:877-942.  lvl is the coarse level. */
int or_correct_prolong(const or_params *p, const or_metrics *m, int lvl,
                       or_cplx *vc, const or_cplx *wc, or_cplx *vf)
{
    (void)p;
    int L = lvl - 1, nxc = m->ptsx[L], nyc = m->ptsy[L];
    int nxf = m->ptsx[L - 1], nyf = m->ptsy[L - 1];
    for (int j = 1; j <= nyc; ++j)
        for (int i = 2; i <= nxc - 1; ++i)
            AT(vc, nxc, i, j) = AT(vc, nxc, i, j) - AT(wc, nxc, i, j);
    or_cplx *t = malloc(sizeof(or_cplx) * (size_t)nxf * nyf);
    for (int j = 1; j <= nyc; ++j)
        for (int i = 1; i <= nxc; ++i) AT(t, nxf, 2 * i - 1, 2 * j - 1) = AT(vc, nxc, i, j);
    for (int j = 1; j <= nyf; j += 2)
        for (int i = 2; i <= nxf - 1; i += 2)
            AT(t, nxf, i, j) = 0.5 * (AT(t, nxf, i - 1, j) + AT(t, nxf, i + 1, j));
    for (int j = 2; j <= nyf - 1; j += 2)
        for (int i = 1; i <= nxf; ++i)
            AT(t, nxf, i, j) = 0.5 * (AT(t, nxf, i, j + 1) + AT(t, nxf, i, j - 1));
    for (int j = 1; j <= nyf; ++j)
        for (int i = 2; i <= nxf - 1; ++i) AT(vf, nxf, i, j) = AT(vf, nxf, i, j) + AT(t, nxf, i, j);
    free(t);
    return 0;
}

static int check_grid(const or_params *p)
{
    if (p->msh < 1 || p->msh > OR_MAXLVL) return -1;
    int d = 1 << (p->msh - 1);
    if ((p->nx - 1) % d || (p->ny - 1) % d) return -2;
    if ((p->nx - 1) / d + 1 < 7 || (p->ny - 1) / d + 1 < 5) return -3;
    if (p->ny > 1024) return -4;
    return 0;
}

/* poi_uy: poisson_par.f:6-64, the body of the mode loop (poisson_par.f:30)
   for mode k (1-based) with work planes allocated per call */
static int poi_uy_mode(const or_params *p, const or_metrics *m, int k, const or_cplx *wx,
                       const or_cplx *dwzdx, or_cplx *uy, double erromax, int max_cycles,
                       int order, int *cycles, double *res, or_cplx **V, or_cplx **S,
                       or_cplx **W, or_cplx **R)
{
    int msh = p->msh, nx = p->nx, ny = p->ny;
    size_t off = (size_t)nx * ny * (k - 1);
    or_source(p, k, wx, dwzdx, S[0]);
    memcpy(V[0], uy + off, sizeof(or_cplx) * nx * ny);
    double err = 0.0;
    int done = 0, vc;
    for (vc = 1; vc <= max_cycles && !done; ++vc) {
        for (int l = 1; l <= msh - 1; ++l) {
            or_sor(p, m, l, k, V[l - 1], S[l - 1], 3, order);
            or_defect(p, m, l, k, V[l - 1], S[l - 1], R[l - 1], -1.0);
            if (l == 1) {   /* verifica_poi + get_max_value, 67-108, 945-963 */
                err = 0.0;
                for (int j = 2; j <= ny; ++j)
                    for (int i = 2; i <= nx - 1; ++i) {
                        double a = cabs(AT(R[0], nx, i, j));
                        if (a > err) err = a;
                    }
                if (err < erromax) { done = 1; break; }
            }
            or_restrict(p, m, l, V[l - 1], R[l - 1], V[l], W[l], S[l]);
            or_defect(p, m, l + 1, k, V[l], S[l], S[l], 1.0);
        }
        if (done) break;
        or_sor(p, m, msh, k, V[msh - 1], S[msh - 1], 60, order);
        for (int l = msh; l >= 2; --l) {
            or_correct_prolong(p, m, l, V[l - 1], W[l - 1], V[l - 2]);
            or_sor(p, m, l - 1, k, V[l - 2], S[l - 2], 1, order);
        }
    }
    if (cycles) cycles[k - 1] = done ? vc - 1 : max_cycles;
    if (res) res[k - 1] = err;
    for (int j = 2; j <= ny; ++j)
        for (int i = 2; i <= nx - 1; ++i) AT(uy + off, nx, i, j) = AT(V[0], nx, i, j);
    return 0;
}

static int poi_uy_modes(const or_params *p, const or_cplx *wx, const or_cplx *dwzdx,
                        or_cplx *uy, double erromax, int max_cycles, int order,
                        int *cycles, double *res, const int *modes, int nsel)
{
    int rc = check_grid(p);
    if (rc) return rc;
    or_metrics m;
    if (or_create_ctes(p, &m)) return -1;
    int msh = p->msh;
    or_cplx *V[OR_MAXLVL], *S[OR_MAXLVL], *W[OR_MAXLVL], *R[OR_MAXLVL];
    for (int l = 0; l < msh; ++l) {
        size_t n = (size_t)m.ptsx[l] * m.ptsy[l];
        V[l] = calloc(n, sizeof(or_cplx));
        S[l] = calloc(n, sizeof(or_cplx));
        W[l] = calloc(n, sizeof(or_cplx));
        R[l] = calloc(n, sizeof(or_cplx));
    }
    for (int s = 0; s < nsel; ++s) {
        int k = modes ? modes[s] : s + 1;
        if (k < 1 || k > p->nk) { rc = -5; break; }
        poi_uy_mode(p, &m, k, wx, dwzdx, uy, erromax, max_cycles, order, cycles, res, V, S, W,
                    R);
    }
    for (int l = 0; l < msh; ++l) {
        free(V[l]); free(S[l]); free(W[l]); free(R[l]);
    }
    return rc;
}

int or_poi_uy(const or_params *p, const or_cplx *wx, const or_cplx *dwzdx,
              or_cplx *uy, double erromax, int max_cycles, int order,
              int *cycles, double *res)
{
    return poi_uy_modes(p, wx, dwzdx, uy, erromax, max_cycles, order, cycles, res, NULL, p->nk);
}

int or_poi_uy_modes(const or_params *p, const or_cplx *wx, const or_cplx *dwzdx,
                    or_cplx *uy, double erromax, int max_cycles, int order,
                    int *cycles, double *res, const int *modes, int nsel)
{
    return poi_uy_modes(p, wx, dwzdx, uy, erromax, max_cycles, order, cycles, res, modes, nsel);
}

int or_apply_L_level(const or_params *p, int lvl, int k, const or_cplx *v, or_cplx *out)
{
    or_metrics m;
    if (or_create_ctes(p, &m)) return -1;
    size_t n = (size_t)m.ptsx[lvl - 1] * m.ptsy[lvl - 1];
    or_cplx *z = calloc(n, sizeof(or_cplx));
    memset(out, 0, sizeof(or_cplx) * n);
    or_defect(p, &m, lvl, k, v, z, out, 1.0);
    free(z);
    return 0;
}

int or_apply_L(const or_params *p, int k, const or_cplx *v, or_cplx *out)
{
    return or_apply_L_level(p, 1, k, v, out);
}
