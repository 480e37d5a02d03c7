// Argument blocks and launchers of the right-hand-side kernels (tbl_rhs.hip).
#pragma once

#include "tbl_common.h"

namespace tbl {

// points per workgroup of the nonlinear-terms kernel
constexpr int kNlPts = 4;

// base flow and curvature (COMMON /blas/, /vc/), real (nx, ny) device arrays
struct BaseArgs {
    const double *uxb, *uyb, *wzb, *varg, *dvargdx;
};

struct RhsArgs {
    BaseArgs b;
    const cx<double> *ux, *uy, *uz, *wx, *wy, *wz;
    cx<double> *a, *bb, *c, *d;    // d: Goertler form only
    const double *tw;          // nonlinear: cos, sin of 2 pi t / kphys, t < kphys
    int nx, ny, K;
    int fi;                    // flat-plate form (lterms_fi / nlterms_fi): a, b, c
};

struct DrvArgs {
    BaseArgs b;
    const cx<double> *wx, *wy, *wz, *bb, *c, *d;
    const cx<double> *dadx, *dbdx, *dddx, *dady, *dcdy;
    const cx<double> *xx, *yx, *zx, *xy, *yy, *zy;    // d2w/dx2, d2w/dy2 of wx, wy, wz
    cx<double> *dvx, *dvy, *dvz;
    double rre, fac_y, beta;
    int nx, ny, K;
    int fi;                    // drv_fi: no curvature terms (d, dddx, varg unused)
};

struct RkArgs {
    cx<double> *w[3], *w1[3], *dv1[3];
    const cx<double> *dv2[3];
    double h;                  // dt/2, dt/2, dt, dt/6 for stages 1..4
    int stage, nx, ny, K;
};

size_t nlterms_lds(int K, int N);
int terms(const RhsArgs &r, int kphys, hipStream_t st);
int drv_combine(const DrvArgs &r, hipStream_t st);
int rk_update(const RkArgs &r, hipStream_t st);
int pert_rows(cx<double> *uy, cx<double> *d2, const cx<double> *pu, const cx<double> *pd,
              int nx, int ny, int K, int i0, int i1, hipStream_t st);
int rhs_kernel_attributes();

}  // namespace tbl
