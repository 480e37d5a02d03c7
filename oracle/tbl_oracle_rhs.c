/*
 * CPU oracle, part 3: the right-hand side of the vorticity equations and
 * the RK4 time step of the reference's Goertler case (my_form 1):
 * lterms_gv / nlterms_gv with their z transforms f_to_p, p_to_f and four1
 * (nlterms_par.f:89-182, 303-490), drv_gv (main_par.f:1339-1403) and the
 * four-stage update around `loop` (main_par.f:266-349); and the flat-plate
 * form (my_form 0) of the right-hand side: lterms_fi / nlterms_fi
 * (nlterms_par.f:6-86) and drv_fi (main_par.f:1264-1336).
 *
 * TEST INFRASTRUCTURE ONLY (see tbl_oracle.h).  One MPI rank.  Expressions
 * keep the reference's evaluation order.
 */
#include "tbl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define F3(a, i, j, k) (a)[((i) - 1) + (size_t)nx * (((j) - 1) + (size_t)ny * ((k) - 1))]
#define F2(a, i, j) (a)[((i) - 1) + (size_t)nx * ((j) - 1)]

static double k2b2_of(const or_params *p, int k)
{
    return -(double)(k - 1) * (double)(k - 1) * p->beta * p->beta;
}

static or_cplx kb_of(const or_params *p, int k)
{
    return -((double)(k - 1) * p->beta) * I;
}

/* lterms_gv: nlterms_par.f:89-112 */
static void lterms(const or_params *p, const or_base *b, const or_cplx *ux, const or_cplx *uy,
                   const or_cplx *wx, const or_cplx *wy, const or_cplx *wz, or_cplx *a,
                   or_cplx *bb, or_cplx *c, or_cplx *d)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    for (int k = 1; k <= nk; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double uxb = F2(b->uxb, i, j), uyb = F2(b->uyb, i, j), wzb = F2(b->wzb, i, j);
                F3(a, i, j, k) = uyb * F3(wx, i, j, k) - F3(wy, i, j, k) * uxb;
                F3(bb, i, j, k) = uxb * F3(wz, i, j, k) + F3(ux, i, j, k) * wzb;
                F3(c, i, j, k) = -uyb * F3(wz, i, j, k) - F3(uy, i, j, k) * wzb;
                F3(d, i, j, k) = 2.0 * uxb * F3(ux, i, j, k);
            }
}

/* four1 (nlterms_par.f:431-484) on one line of n = kphys reals */
static void four1(double *d, int n, int isig)
{
    int z = 1;
#define D(q) d[(q) - 1]
    for (int i = 1; i <= n; i += 2) {
        if (z > i) {
            double tr = D(z), ti = D(z + 1);
            D(z) = D(i);
            D(z + 1) = D(i + 1);
            D(i) = tr;
            D(i + 1) = ti;
        }
        int mm = n / 2;
        while (mm >= 2 && z > mm) {
            z -= mm;
            mm /= 2;
        }
        z += mm;
    }
    int mmax = 2;
    while (n > mmax) {
        int istep = 2 * mmax;
        double theta = 6.28318530717959 / (double)(isig * mmax);
        double s = sin(0.5 * theta);
        double wpr = -2.0 * (s * s), wpi = sin(theta), wr = 1.0, wi = 0.0;
        for (int mm = 1; mm <= mmax; mm += 2) {
            for (int i = mm; i <= n; i += istep) {
                int q = i + mmax;
                double tr = wr * D(q) - wi * D(q + 1);
                double ti = wr * D(q + 1) + wi * D(q);
                D(q) = D(i) - tr;
                D(q + 1) = D(i + 1) - ti;
                D(i) = D(i) + tr;
                D(i + 1) = D(i + 1) + ti;
            }
            double wt = wr;
            wr = wr * wpr - wi * wpi + wr;
            wi = wi * wpr + wt * wpi + wi;
        }
        mmax = istep;
    }
#undef D
}

/* f_to_p (nlterms_par.f:303-361) of one point: nk modes -> kphys reals */
static void f_to_p_point(const or_cplx *f, size_t stride, int nk, int kphys, double *dp)
{
    const double theta = -6.28318530717959 / (double)kphys, c1 = 0.5;
    double s = sin(0.5 * theta);
    const double wpr = -2.0 * (s * s), wpi = sin(theta);
#define D(q) dp[(q) - 1]
    for (int k = 0; k < nk; ++k) {
        D(2 * k + 1) = creal(f[k * stride]);
        D(2 * k + 2) = cimag(f[k * stride]);
    }
    for (int k = 2 * nk + 1; k <= kphys; ++k) D(k) = 0.0;
    D(1) = D(1) * 2.0;
    double wr = 1.0 + wpr, wi = wpi;
    const int n2p3 = kphys + 3;
    for (int q = 2; q <= kphys / 4; ++q) {
        int p1 = 2 * q - 1, p2 = p1 + 1, p3 = n2p3 - p2, p4 = p3 + 1;
        double wrs = wr, wis = wi;
        double h1r = c1 * (D(p1) + D(p3));
        double h1i = c1 * (D(p2) - D(p4));
        double h2r = -c1 * (D(p2) + D(p4));
        double h2i = c1 * (D(p1) - D(p3));
        D(p1) = h1r + wrs * h2r - wis * h2i;
        D(p2) = h1i + wrs * h2i + wis * h2r;
        D(p3) = h1r - wrs * h2r + wis * h2i;
        D(p4) = -h1i + wrs * h2i + wis * h2r;
        double wt = wr;
        wr = wr * wpr - wi * wpi + wr;
        wi = wi * wpr + wt * wpi + wi;
    }
    double h1r = D(1);
    D(1) = c1 * (h1r + D(2));
    D(2) = c1 * (h1r - D(2));
    four1(dp, kphys, -1);
#undef D
}

/* p_to_f (nlterms_par.f:364-428) of one point: kphys reals -> nk modes */
static void p_to_f_point(double *dp, int nk, int kphys, or_cplx *f, size_t stride)
{
    const double theta = 6.28318530717959 / (double)kphys, c1 = 0.5;
    double s = sin(0.5 * theta);
    const double wpr = -2.0 * (s * s), wpi = sin(theta);
#define D(q) dp[(q) - 1]
    four1(dp, kphys, +1);
    double wr = 1.0 + wpr, wi = wpi;
    const int n2p3 = kphys + 3;
    for (int q = 2; q <= kphys / 4; ++q) {
        int p1 = 2 * q - 1, p2 = p1 + 1, p3 = n2p3 - p2, p4 = p3 + 1;
        double wrs = wr, wis = wi;
        double h1r = c1 * (D(p1) + D(p3));
        double h1i = c1 * (D(p2) - D(p4));
        double h2r = c1 * (D(p2) + D(p4));
        double h2i = -c1 * (D(p1) - D(p3));
        D(p1) = h1r + wrs * h2r - wis * h2i;
        D(p2) = h1i + wrs * h2i + wis * h2r;
        D(p3) = h1r - wrs * h2r + wis * h2i;
        D(p4) = -h1i + wrs * h2i + wis * h2r;
        double wt = wr;
        wr = wr * wpr - wi * wpi + wr;
        wi = wi * wpr + wt * wpi + wi;
    }
    double h1r = D(1);
    D(1) = (h1r + D(2)) / (double)kphys;
    D(2) = h1r - D(2);
    for (int k = 2; k <= kphys; ++k) D(k) = 2.0 * D(k) / (double)kphys;
    for (int k = 0; k < nk; ++k) {
        if (fabs(D(2 * k + 1)) < 1e-14) D(2 * k + 1) = 0.0;
        if (fabs(D(2 * k + 2)) < 1e-14) D(2 * k + 2) = 0.0;
        f[k * stride] = D(2 * k + 1) + D(2 * k + 2) * I;
    }
    f[0] = creal(f[0]);
#undef D
}

int or_f_to_p(const or_params *p, int kphys, const or_cplx *f, double *phys)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    if (kphys < 2 * nk || kphys < 4 || (kphys & (kphys - 1))) return -1;
    double *d = malloc(sizeof(double) * kphys);
    const size_t plane = (size_t)nx * ny;
    for (size_t e = 0; e < plane; ++e) {
        f_to_p_point(f + e, plane, nk, kphys, d);
        for (int q = 0; q < kphys; ++q) phys[e + plane * q] = d[q];
    }
    free(d);
    return 0;
}

int or_p_to_f(const or_params *p, int kphys, const double *phys, or_cplx *f)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    if (kphys < 2 * nk || kphys < 4 || (kphys & (kphys - 1))) return -1;
    double *d = malloc(sizeof(double) * kphys);
    const size_t plane = (size_t)nx * ny;
    for (size_t e = 0; e < plane; ++e) {
        for (int q = 0; q < kphys; ++q) d[q] = phys[e + plane * q];
        p_to_f_point(d, nk, kphys, f + e, plane);
    }
    free(d);
    return 0;
}

/* nlterms_gv: nlterms_par.f:115-180 (products of six fields per point) */
static int nlterms(const or_params *p, int kphys, const or_base *b, const or_cplx *ux,
                   const or_cplx *uy, const or_cplx *uz, const or_cplx *wx, const or_cplx *wy,
                   const or_cplx *wz, or_cplx *a, or_cplx *bb, or_cplx *c, or_cplx *d)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    if (kphys < 2 * nk || kphys < 4 || (kphys & (kphys - 1))) return -1;
    const size_t plane = (size_t)nx * ny;
    double *P = malloc(sizeof(double) * kphys * 10);
    double *up = P, *vp = P + kphys, *zp = P + 2 * kphys, *xp = P + 3 * kphys,
           *yp = P + 4 * kphys, *wp = P + 5 * kphys, *ap = P + 6 * kphys, *bp = P + 7 * kphys,
           *cp = P + 8 * kphys, *dp = P + 9 * kphys;
    for (size_t e = 0; e < plane; ++e) {
        f_to_p_point(ux + e, plane, nk, kphys, up);
        f_to_p_point(uy + e, plane, nk, kphys, vp);
        f_to_p_point(uz + e, plane, nk, kphys, zp);
        f_to_p_point(wx + e, plane, nk, kphys, xp);
        f_to_p_point(wy + e, plane, nk, kphys, yp);
        f_to_p_point(wz + e, plane, nk, kphys, wp);
        for (int q = 0; q < kphys; ++q) {
            ap[q] = vp[q] * xp[q] - up[q] * yp[q];
            bp[q] = up[q] * wp[q] - zp[q] * xp[q];
            cp[q] = zp[q] * yp[q] - vp[q] * wp[q];
            dp[q] = up[q] * up[q];
        }
        p_to_f_point(ap, nk, kphys, a + e, plane);
        p_to_f_point(bp, nk, kphys, bb + e, plane);
        p_to_f_point(cp, nk, kphys, c + e, plane);
        p_to_f_point(dp, nk, kphys, d + e, plane);
    }
    free(P);
    for (int k = 1; k <= nk; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double uxb = F2(b->uxb, i, j), uyb = F2(b->uyb, i, j), wzb = F2(b->wzb, i, j);
                F3(a, i, j, k) = uyb * F3(wx, i, j, k) - uxb * F3(wy, i, j, k) + F3(a, i, j, k);
                F3(bb, i, j, k) = uxb * F3(wz, i, j, k) + F3(ux, i, j, k) * wzb + F3(bb, i, j, k);
                F3(c, i, j, k) = -uyb * F3(wz, i, j, k) - F3(uy, i, j, k) * wzb + F3(c, i, j, k);
                F3(d, i, j, k) = 2.0 * uxb * F3(ux, i, j, k) + F3(d, i, j, k);
            }
    return 0;
}

int or_terms(const or_params *p, int kphys, const or_base *b, const or_cplx *ux,
             const or_cplx *uy, const or_cplx *uz, const or_cplx *wx, const or_cplx *wy,
             const or_cplx *wz, or_cplx *a, or_cplx *bb, or_cplx *c, or_cplx *d)
{
    if (kphys == 0) {
        lterms(p, b, ux, uy, wx, wy, wz, a, bb, c, d);
        return 0;
    }
    return nlterms(p, kphys, b, ux, uy, uz, wx, wy, wz, a, bb, c, d);
}

/* drv_gv: main_par.f:1339-1403 (kphys 0: lterms_gv, the reference's active
   call; otherwise nlterms_gv with that many physical points) */
int or_drv_gv(const or_params *p, double Re, int kphys, const or_base *b, const or_cplx *ux,
              const or_cplx *uy, const or_cplx *uz, const or_cplx *wx, const or_cplx *wy,
              const or_cplx *wz, or_cplx *dvx, or_cplx *dvy, or_cplx *dvz)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    const size_t n3 = (size_t)nx * ny * nk;
    or_cplx *buf = malloc(sizeof(or_cplx) * n3 * 15);
    if (!buf) return -2;
    or_cplx *a = buf, *bb = a + n3, *c = bb + n3, *d = c + n3, *dadx = d + n3, *dbdx = dadx + n3,
            *dddx = dbdx + n3, *dady = dddx + n3, *dcdy = dady + n3, *xx = dcdy + n3,
            *yx = xx + n3, *zx = yx + n3, *xy = zx + n3, *yy = xy + n3, *zy = yy + n3;
    int rc = or_terms(p, kphys, b, ux, uy, uz, wx, wy, wz, a, bb, c, d);
    if (!rc) rc = or_derparx(p, a, dadx);
    if (!rc) rc = or_derparx(p, bb, dbdx);
    if (!rc) rc = or_derparx(p, d, dddx);
    if (!rc) rc = or_dery(p, a, dady);
    if (!rc) rc = or_dery(p, c, dcdy);
    if (!rc) rc = or_derparxx(p, wx, xx);
    if (!rc) rc = or_derparxx(p, wy, yx);
    if (!rc) rc = or_derparxx(p, wz, zx);
    if (!rc) rc = or_deryy(p, wx, xy);
    if (!rc) rc = or_deryy(p, wy, yy);
    if (!rc) rc = or_deryy(p, wz, zy);
    if (rc) {
        free(buf);
        return rc;
    }
    const double fy = p->fac_y;
    for (int k = 1; k <= nk; ++k) {
        const or_cplx kb = kb_of(p, k);
        const double k2 = k2b2_of(p, k);
        for (int j = 2; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double vg = F2(b->varg, i, j), dvg = F2(b->dvargdx, i, j);
                F3(dvx, i, j, k) = -F3(dady, i, j, k) + kb * F3(bb, i, j, k)
                                   - kb * F3(d, i, j, k) * vg
                                   + (F3(xx, i, j, k) + F3(xy, i, j, k) * fy
                                      + k2 * F3(wx, i, j, k)) / Re;
                F3(dvy, i, j, k) = -kb * F3(c, i, j, k) + F3(dadx, i, j, k)
                                   + (F3(yx, i, j, k) + F3(yy, i, j, k) * fy
                                      + k2 * F3(wy, i, j, k)) / Re;
                F3(dvz, i, j, k) = -F3(dbdx, i, j, k) + F3(dcdy, i, j, k)
                                   + vg * F3(dddx, i, j, k) + F3(d, i, j, k) * dvg
                                   + (F3(zx, i, j, k) + F3(zy, i, j, k) * fy
                                      + k2 * F3(wz, i, j, k)) / Re;
            }
    }
    free(buf);
    return 0;
}

/* lterms_fi: nlterms_par.f:6-38 (mode 1 against the base flow's own
   products, modes >= 2 as lterms_gv without d) */
static void lterms_fi(const or_params *p, const or_base *b, const or_cplx *ux,
                      const or_cplx *uy, const or_cplx *wx, const or_cplx *wy, const or_cplx *wz,
                      or_cplx *a, or_cplx *bb, or_cplx *c)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            double uxb = F2(b->uxb, i, j), uyb = F2(b->uyb, i, j), wzb = F2(b->wzb, i, j);
            F3(a, i, j, 1) = uyb * F3(wx, i, j, 1) - F3(wy, i, j, 1) * uxb;
            F3(bb, i, j, 1) = uxb * (F3(wz, i, j, 1) - wzb) + wzb * (F3(ux, i, j, 1) - uxb);
            F3(c, i, j, 1) = -uyb * (F3(wz, i, j, 1) - wzb) - wzb * (F3(uy, i, j, 1) - uyb);
        }
    for (int k = 2; k <= nk; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double uxb = F2(b->uxb, i, j), uyb = F2(b->uyb, i, j), wzb = F2(b->wzb, i, j);
                F3(a, i, j, k) = uyb * F3(wx, i, j, k) - F3(wy, i, j, k) * uxb;
                F3(bb, i, j, k) = uxb * F3(wz, i, j, k) + F3(ux, i, j, k) * wzb;
                F3(c, i, j, k) = -uyb * F3(wz, i, j, k) - F3(uy, i, j, k) * wzb;
            }
}

/* nlterms_fi: nlterms_par.f:41-86 (the three products in physical space,
   no base-flow terms) */
static int nlterms_fi(const or_params *p, int kphys, const or_cplx *ux, const or_cplx *uy,
                      const or_cplx *uz, const or_cplx *wx, const or_cplx *wy,
                      const or_cplx *wz, or_cplx *a, or_cplx *bb, or_cplx *c)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    if (kphys < 2 * nk || kphys < 4 || (kphys & (kphys - 1))) return -1;
    const size_t plane = (size_t)nx * ny;
    double *P = malloc(sizeof(double) * kphys * 9);
    if (!P) return -2;
    double *up = P, *vp = P + kphys, *zp = P + 2 * kphys, *xp = P + 3 * kphys,
           *yp = P + 4 * kphys, *wp = P + 5 * kphys, *ap = P + 6 * kphys, *bp = P + 7 * kphys,
           *cp = P + 8 * kphys;
    for (size_t e = 0; e < plane; ++e) {
        f_to_p_point(ux + e, plane, nk, kphys, up);
        f_to_p_point(uy + e, plane, nk, kphys, vp);
        f_to_p_point(uz + e, plane, nk, kphys, zp);
        f_to_p_point(wx + e, plane, nk, kphys, xp);
        f_to_p_point(wy + e, plane, nk, kphys, yp);
        f_to_p_point(wz + e, plane, nk, kphys, wp);
        for (int q = 0; q < kphys; ++q) {
            ap[q] = vp[q] * xp[q] - up[q] * yp[q];
            bp[q] = up[q] * wp[q] - zp[q] * xp[q];
            cp[q] = zp[q] * yp[q] - vp[q] * wp[q];
        }
        p_to_f_point(ap, nk, kphys, a + e, plane);
        p_to_f_point(bp, nk, kphys, bb + e, plane);
        p_to_f_point(cp, nk, kphys, c + e, plane);
    }
    free(P);
    return 0;
}

int or_terms_fi(const or_params *p, int kphys, const or_base *b, const or_cplx *ux,
                const or_cplx *uy, const or_cplx *uz, const or_cplx *wx, const or_cplx *wy,
                const or_cplx *wz, or_cplx *a, or_cplx *bb, or_cplx *c)
{
    if (kphys == 0) {
        if (!b) return -1;
        lterms_fi(p, b, ux, uy, wx, wy, wz, a, bb, c);
        return 0;
    }
    return nlterms_fi(p, kphys, ux, uy, uz, wx, wy, wz, a, bb, c);
}

/* drv_fi: main_par.f:1264-1336 (kphys > 0: nlterms_fi, the reference's
   active call, with that many physical points; 0: lterms_fi) */
int or_drv_fi(const or_params *p, double Re, int kphys, const or_base *b, const or_cplx *ux,
              const or_cplx *uy, const or_cplx *uz, const or_cplx *wx, const or_cplx *wy,
              const or_cplx *wz, or_cplx *dvx, or_cplx *dvy, or_cplx *dvz)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    const size_t n3 = (size_t)nx * ny * nk;
    or_cplx *buf = malloc(sizeof(or_cplx) * n3 * 13);
    if (!buf) return -2;
    or_cplx *a = buf, *bb = a + n3, *c = bb + n3, *dadx = c + n3, *dbdx = dadx + n3,
            *dady = dbdx + n3, *dcdy = dady + n3, *xx = dcdy + n3, *yx = xx + n3, *zx = yx + n3,
            *xy = zx + n3, *yy = xy + n3, *zy = yy + n3;
    int rc = or_terms_fi(p, kphys, b, ux, uy, uz, wx, wy, wz, a, bb, c);
    if (!rc) rc = or_derparx(p, a, dadx);
    if (!rc) rc = or_derparx(p, bb, dbdx);
    if (!rc) rc = or_dery(p, a, dady);
    if (!rc) rc = or_dery(p, c, dcdy);
    if (!rc) rc = or_derparxx(p, wx, xx);
    if (!rc) rc = or_derparxx(p, wy, yx);
    if (!rc) rc = or_derparxx(p, wz, zx);
    if (!rc) rc = or_deryy(p, wx, xy);
    if (!rc) rc = or_deryy(p, wy, yy);
    if (!rc) rc = or_deryy(p, wz, zy);
    if (rc) {
        free(buf);
        return rc;
    }
    const double fy = p->fac_y;
    for (int k = 1; k <= nk; ++k) {
        const or_cplx kb = kb_of(p, k);
        const double k2 = k2b2_of(p, k);
        for (int j = 2; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                F3(dvx, i, j, k) = -F3(dady, i, j, k) + kb * F3(bb, i, j, k)
                                   + (F3(xx, i, j, k) + F3(xy, i, j, k) * fy
                                      + k2 * F3(wx, i, j, k)) / Re;
                F3(dvy, i, j, k) = -kb * F3(c, i, j, k) + F3(dadx, i, j, k)
                                   + (F3(yx, i, j, k) + F3(yy, i, j, k) * fy
                                      + k2 * F3(wy, i, j, k)) / Re;
                F3(dvz, i, j, k) = -F3(dbdx, i, j, k) + F3(dcdy, i, j, k)
                                   + (F3(zx, i, j, k) + F3(zy, i, j, k) * fy
                                      + k2 * F3(wz, i, j, k)) / Re;
            }
    }
    free(buf);
    return 0;
}

/* One RK4 time step of the Goertler case: main_par.f:266-349 (four
   drv_gv + update stages, each followed by loop; loop(1e-5) after stages
   1-3, loop(1e-6) after stage 4).  The disturbance of stages 1 and 3
   (gv_pert, main_par.f:1577-1610) is given as pert_uy / pert_d2 (nx, nk):
   for i in [pert_i0, pert_i1] (1-based) and k >= 2 they replace uy(i,1,k)
   and d2uydx2(i,k); pert_uy NULL: none.  f[0..6] = ux, uy, uz, wx, wy, wz, th
   (th NULL). */
int or_rk4_step(const or_params *p, double Re, double dt, int kphys, const or_base *b,
                const double *bdfc, or_cplx *d2uydx2, const or_cplx *pert_uy,
                const or_cplx *pert_d2, int pert_i0, int pert_i1, or_cplx *const *f, int order,
                int *cycles)
{
    const int nx = p->nx, ny = p->ny, nk = p->nk;
    const size_t n3 = (size_t)nx * ny * nk;
    or_cplx *buf = calloc(n3 * 9, sizeof(or_cplx));
    if (!buf) return -2;
    or_cplx *dv1[3] = {buf, buf + n3, buf + 2 * n3};
    or_cplx *dv2[3] = {buf + 3 * n3, buf + 4 * n3, buf + 5 * n3};
    or_cplx *w1[3] = {buf + 6 * n3, buf + 7 * n3, buf + 8 * n3};
    or_cplx *ux = f[0], *uy = f[1], *uz = f[2], *w[3] = {f[3], f[4], f[5]};
    const double dt2 = 0.5 * dt, dt6 = dt / 6.0;
    const int i_ini = 3;
    int rc = 0;
    for (int s = 1; s <= 4 && !rc; ++s) {
        or_cplx **dv = s == 1 ? dv1 : dv2;
        rc = or_drv_gv(p, Re, kphys, b, ux, uy, uz, w[0], w[1], w[2], dv[0], dv[1], dv[2]);
        if (rc) break;
        for (int q = 0; q < 3; ++q)
            for (int k = 1; k <= nk; ++k)
                for (int j = 2; j <= ny - 1; ++j)
                    for (int i = i_ini; i <= nx; ++i) {
                        if (s == 1) {
                            F3(w1[q], i, j, k) = F3(w[q], i, j, k);
                            F3(w[q], i, j, k) = F3(w1[q], i, j, k) + F3(dv1[q], i, j, k) * dt2;
                        } else if (s == 2 || s == 3) {
                            double h = s == 2 ? dt2 : dt;
                            F3(w[q], i, j, k) = F3(w1[q], i, j, k) + F3(dv2[q], i, j, k) * h;
                            F3(dv1[q], i, j, k) = F3(dv1[q], i, j, k) + 2.0 * F3(dv2[q], i, j, k);
                        } else {
                            F3(w[q], i, j, k) = F3(w1[q], i, j, k)
                                                + dt6 * (F3(dv1[q], i, j, k) + F3(dv2[q], i, j, k));
                        }
                    }
        if ((s == 1 || s == 3) && pert_uy)
            for (int i = pert_i0; i <= pert_i1; ++i)
                for (int k = 2; k <= nk; ++k) {
                    F3(uy, i, 1, k) = pert_uy[(i - 1) + (size_t)nx * (k - 1)];
                    d2uydx2[(i - 1) + (size_t)nx * (k - 1)] =
                        pert_d2[(i - 1) + (size_t)nx * (k - 1)];
                }
        rc = or_loop(p, s == 4 ? 1e-6 : 1e-5, bdfc, d2uydx2, ux, uy, uz, w[0], w[1], w[2], f[6],
                     order, cycles ? cycles + (size_t)(s - 1) * nk : NULL, NULL);
    }
    free(buf);
    return rc;
}
