// Argument blocks and launchers of the x-line kernels (tbl_xline.hip): the
// routines of the reference's RK-stage `loop` around poi_uy.
#pragma once

#include "tbl_common.h"

namespace tbl {

// pentadiagonal x-Poisson line kinds (right-hand sides of loop_par.f)
constexpr int kXUx = 0;     // poi_ux, k >= 2, rows 2..jmax (428-556)
constexpr int kXUz = 1;     // poi_uz, k >= 2, rows 2..jmax (559-688)
constexpr int kXWall = 2;   // wx_wall, k >= 2, row 1 (174-255)
// continuity integration kinds
constexpr int kXIntUx = 0;  // poi_ux, k = 1, rows 2..jmax-1 (457-502)
constexpr int kXIntWz = 1;  // wz_wall, every k, row 1 (364-425)
// longest x-line: 64 lanes x kXMaxR rows
constexpr int kXMaxR = 64;

// tables, lane-segmented ([row][lane], 64 R rows each):
//   penta, per mode k >= 2: l2, l1 (forward), 1/U(i,i), U(i,i+1)/U(i,i),
//                           U(i,i+2)/U(i,i) (backward)      -> 5 per mode
//   integration:            P, Q of x(i) = g - P x(i-1) - Q x(i-2)
//   filter:                 a/bet, 1/bet (forward), gam(i+1) (backward)
struct XPenta {
    const cx<double> *A, *B;   // right-hand side inputs
    cx<double> *X;             // solution (column 1 read as the boundary value)
    const double *tab;
    int nx, ny, K, R, kind;
    double dxx, beta;
};

struct XInteg {
    const cx<double> *D;       // kXIntUx: duydy; kXIntWz: wx
    const cx<double> *Lp;      // kXIntWz: lapv (nx, K)
    cx<double> *X;             // ux / wz
    const double *tab;
    int nx, ny, K, R, kind;
    double sdx, beta;          // -dx (poi_ux) or dx (wz_wall)
};

struct XFilter {
    cx<double> *w0, *w1, *w2;
    const double *tab;
    int nx, ny, K, R;
    double af, bf, cf, df;
};

struct WallArgs {
    const cx<double> *uy, *wy, *d2uydx2;
    cx<double> *lapv, *dwydy, *d2wydxdy;   // (nx, K) each
    double w[9], c[8];                     // w_at_w_coef, dwydy_coef
    double dyyw1, rdy, dx, beta;
    int nx, ny, K;
};

struct BzoneArgs {
    const double *bdfc;        // Fortran (2, nx)
    cx<double> *wx, *wy, *wz, *th;
    int nx, ny, K, j0;
};

size_t xline_lds(int R);
int xpenta(const XPenta &a, hipStream_t st);
int xinteg(const XInteg &a, hipStream_t st);
int xfilter(const XFilter &a, hipStream_t st);
int wall_terms(const WallArgs &a, hipStream_t st);
int bzone(const BzoneArgs &a, int rows, hipStream_t st);
int outflow_row(cx<double> *uy, int nx, int ny, int K, hipStream_t st);
int xline_kernel_attributes();

}  // namespace tbl
