// Compact finite-difference derivatives (derivs_par.f:80-177) on gfx950.
//
// Arrays are in the caller's layout: complex (nx, ny, K), x fastest.
//  * d/dx, d2/dx2: one wavefront per x-line (j, k).  Lane t owns points
//    [t*R, t*R+R); the tridiagonal Thomas solve (tridparx at one rank,
//    derivs_par.f:775-858) is split into a per-lane affine map, a 6-step
//    shuffle scan and a replay, forward and backward.
//  * d/dy, d2/dy2: one thread per y-line (i, k); neighbouring threads hold
//    neighbouring x so every row access is coalesced (tridy, 947-975).
// The Thomas factors bet_j and gam_j depend only on the scheme and are
// precomputed on the host with the reference's recurrence.
#include "tbl_kernels.h"

namespace tbl {

using C = cx<double>;

// d/dx right-hand side (rhsx, derivs_par.f:259-302) and d2/dx2 (rhsxx, 351-397)
__device__ __forceinline__ C rhs_x(const C *__restrict__ f, int i, int n, double dx, double dxx,
                                   bool second)
{
#define F(I) f[(I) - 1]
    const int I = i + 1;   // 1-based
    if (!second) {
        if (I == 1)
            return (-25.0 * F(1) + 48.0 * F(2) - 36.0 * F(3) + 16.0 * F(4) - 3.0 * F(5)) / (12.0 * dx);
        if (I == 2)
            return (-406.0 * F(1) - 300.0 * F(2) + 760.0 * F(3) - 80.0 * F(4) + 30.0 * F(5)
                    - 4.0 * F(6)) / (120.0 * dx);
        if (I == n - 1)
            return (-406.0 * F(n) - 300.0 * F(n - 1) + 760.0 * F(n - 2) - 80.0 * F(n - 3)
                    + 30.0 * F(n - 4) - 4.0 * F(n - 5)) / (-120.0 * dx);
        if (I == n)
            return (-74.0 * F(n) + 16.0 * F(n - 1) + 72.0 * F(n - 2) - 16.0 * F(n - 3)
                    + 2.0 * F(n - 4)) / (-24.0 * dx);
        return (F(I + 2) - F(I - 2) + 28.0 * (F(I + 1) - F(I - 1))) / (12.0 * dx);
    }
    if (I == 1)
        return (9775.0 * F(1) - 20285.0 * F(2) + 11170.0 * F(3) - 550.0 * F(4) - 145.0 * F(5)
                + 35.0 * F(6)) / (60.0 * dxx);
    if (I == 2)
        return (4834.0 * F(1) - 8424.0 * F(2) + 1890.0 * F(3) + 2320.0 * F(4) - 810.0 * F(5)
                + 216.0 * F(6) - 26.0 * F(7)) / (360.0 * dxx);
    if (I == n - 1)
        return (4834.0 * F(n) - 8424.0 * F(n - 1) + 1890.0 * F(n - 2) + 2320.0 * F(n - 3)
                - 810.0 * F(n - 4) + 216.0 * F(n - 5) - 26.0 * F(n - 6)) / (360.0 * dxx);
    if (I == n)
        return (9775.0 * F(n) - 20285.0 * F(n - 1) + 11170.0 * F(n - 2) - 550.0 * F(n - 3)
                - 145.0 * F(n - 4) + 35.0 * F(n - 5)) / (60.0 * dxx);
    return (3.0 * (F(I + 2) + F(I - 2)) + 48.0 * (F(I + 1) + F(I - 1)) - 102.0 * F(I)) / (4.0 * dxx);
#undef F
}

template <int R>
__global__ void __launch_bounds__(256) deriv_x_kernel(DerivArgs A, int second, const C *f, C *out)
{
    const int lane = threadIdx.x & 63;
    const size_t line = static_cast<size_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (line >= static_cast<size_t>(A.ny) * A.K) return;
    const int n = A.nx;
    const C *fl = f + line * n;
    C *ol = out + line * n;
    const double *ta = A.thomas + static_cast<size_t>(second) * 3 * A.nmax;
    const double *tb = ta + A.nmax, *tg = tb + A.nmax;
    const int i0 = lane * R;
    C u[R];
    // forward sweep u_i = (r_i - a_i u_{i-1}) / bet_i as an affine map of u_{i0-1}
    double M = 1.0;
    C cc = {0, 0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = i0 + r;
        C rr = {0, 0};
        double a = 0.0, bet = 1.0;
        if (i < n) {
            rr = rhs_x(fl, i, n, A.dx, A.dxx, second);
            a = ta[i];
            bet = tb[i];
        }
        u[r] = rr;
        const double m = -a / bet;
        M = m * M;
        cc = (rr - a * cc) / bet;
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double Mo = __shfl_up(M, d, 64);
        const C co = shfl_up(cc, d, 64);
        if (lane >= d) {
            cc = M * co + cc;
            M = M * Mo;
        }
    }
    C prev = shfl_up(cc, 1, 64);
    if (lane == 0) prev = {0, 0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = i0 + r;
        const double a = (i < n) ? ta[i] : 0.0, bet = (i < n) ? tb[i] : 1.0;
        const C v = (u[r] - a * prev) / bet;
        u[r] = v;
        prev = v;
    }
    // backward sweep x_i = u_i - gam_{i+1} x_{i+1}
    M = 1.0;
    cc = {0, 0};
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {
        const int i = i0 + r;
        const double g = (i + 1 < n) ? tg[i + 1] : 0.0;
        M = -g * M;
        cc = u[r] - g * cc;
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double Mo = __shfl_down(M, d, 64);
        const C co = shfl_down(cc, d, 64);
        if (lane + d < 64) {
            cc = M * co + cc;
            M = M * Mo;
        }
    }
    C next = shfl_down(cc, 1, 64);
    if (lane == 63) next = {0, 0};
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {
        const int i = i0 + r;
        const double g = (i + 1 < n) ? tg[i + 1] : 0.0;
        const C v = u[r] - g * next;
        next = v;
        if (i < n) ol[i] = v;
    }
}

// d/dy (rhsy 400-449), d2/dy2 (rhsyy 552-611), d/dy with wall row 0
// (rhsyfv 504-549); kind 2, 3, 4
__global__ void __launch_bounds__(64) deriv_y_kernel(DerivArgs A, int kind, const C *f, C *out)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= A.nx) return;
    const int nx = A.nx, n = A.ny;
    const C *fl = f + static_cast<size_t>(k) * nx * n + i;
    C *ol = out + static_cast<size_t>(k) * nx * n + i;
    const int tk = kind - 2 + 2;   // thomas table index (kinds 2..4)
    const double *ta = A.thomas + static_cast<size_t>(tk) * 3 * A.nmax;
    const double *tb = ta + A.nmax, *tg = tb + A.nmax;
    const double *qdy = A.qdy, *qdy2 = A.qdy2;
#define FY(J) fl[static_cast<size_t>((J) - 1) * nx]
    C prev = {0, 0};
    for (int J = 1; J <= n; ++J) {
        C r;
        if (kind == 2 || kind == 4) {
            const double *sp = A.sp_fd, *cp = A.cp_fd, *pp = A.pp_fd, *lp = A.lp_fd;
            if (J == 1) {
                if (kind == 2) {
                    const double *fe = A.fp_e;
                    r = qdy[0] * (fe[1] * FY(1) + fe[2] * FY(2) + fe[3] * FY(3) + fe[4] * FY(4)
                                  + fe[5] * FY(5));
                } else {
                    r = {0, 0};
                }
            } else if (J == 2) {
                r = qdy[0] * (sp[3] * FY(1) + sp[4] * FY(2) + sp[5] * FY(3) + sp[6] * FY(4)
                              + sp[7] * FY(5) + sp[8] * FY(6));
            } else if (J <= n - 2) {
                r = qdy[J - 3] * (cp[3] * FY(J - 2) + cp[4] * FY(J - 1) + cp[5] * FY(J)
                                  + cp[6] * FY(J + 1) + cp[7] * FY(J + 2));
            } else if (J == n - 1) {
                r = qdy[n - 6] * (pp[3] * FY(n) + pp[4] * FY(n - 1) + pp[5] * FY(n - 2)
                                  + pp[6] * FY(n - 3) + pp[7] * FY(n - 4) + pp[8] * FY(n - 5));
            } else {
                r = qdy[n - 5] * (lp[2] * FY(n) + lp[3] * FY(n - 1) + lp[4] * FY(n - 2)
                                  + lp[5] * FY(n - 3) + lp[6] * FY(n - 4));
            }
        } else {
            const double *fp = A.fp_sd, *sp = A.sp_sd, *cp = A.cp_sd, *pp = A.pp_sd, *lp = A.lp_sd;
            if (J == 1) {
                r = (fp[2] * FY(1) + fp[3] * FY(2) + fp[4] * FY(3) + fp[5] * FY(4) + fp[6] * FY(5)
                     + fp[7] * FY(6)) * qdy2[0];
            } else if (J == 2) {
                r = (sp[3] * FY(1) + sp[4] * FY(2) + sp[5] * FY(3) + sp[6] * FY(4) + sp[7] * FY(5)
                     + sp[8] * FY(6) + sp[9] * FY(7)) * qdy2[0];
            } else if (J <= n - 2) {
                r = (cp[3] * FY(J - 2) + cp[4] * FY(J - 1) + cp[5] * FY(J) + cp[6] * FY(J + 1)
                     + cp[7] * FY(J + 2)) * qdy2[J - 3];
            } else if (J == n - 1) {
                r = (pp[3] * FY(n) + pp[4] * FY(n - 1) + pp[5] * FY(n - 2) + pp[6] * FY(n - 3)
                     + pp[7] * FY(n - 4) + pp[8] * FY(n - 5) + pp[9] * FY(n - 6)) * qdy2[n - 7];
            } else {
                r = (lp[2] * FY(n) + lp[3] * FY(n - 1) + lp[4] * FY(n - 2) + lp[5] * FY(n - 3)
                     + lp[6] * FY(n - 4) + lp[7] * FY(n - 5)) * qdy2[n - 6];
            }
        }
        const C u = (J == 1) ? r / tb[0] : (r - ta[J - 1] * prev) / tb[J - 1];
        ol[static_cast<size_t>(J - 1) * nx] = u;
        prev = u;
    }
#undef FY
    C nxt = prev;
    for (int j = n - 2; j >= 0; --j) {
        C *p = ol + static_cast<size_t>(j) * nx;
        const C v = *p - tg[j + 1] * nxt;
        *p = v;
        nxt = v;
    }
}

#define TBL_XR(X) X(1) X(2) X(3) X(5) X(7) X(9) X(17) X(33)

int deriv(const DerivArgs &A, int kind, const double *f, double *out, hipStream_t st)
{
    const C *fi = reinterpret_cast<const C *>(f);
    C *fo = reinterpret_cast<C *>(out);
    if (kind == 0 || kind == 1) {
        const int need = (A.nx + 63) / 64;
        const size_t lines = static_cast<size_t>(A.ny) * A.K;
        const dim3 grid(static_cast<unsigned>((lines + 3) / 4));
        bool done = false;
#define X(RR)                                                                              \
    if (!done && need <= RR) {                                                             \
        hipLaunchKernelGGL(deriv_x_kernel<RR>, grid, dim3(256), 0, st, A, kind, fi, fo);   \
        done = true;                                                                       \
    }
        TBL_XR(X)
#undef X
        return done ? TBL_OK : TBL_ERR_GRID;
    }
    if (kind < 2 || kind > 4) return TBL_ERR_ARG;
    const dim3 grid((A.nx + 63) / 64, A.K);
    hipLaunchKernelGGL(deriv_y_kernel, grid, dim3(64), 0, st, A, kind, fi, fo);
    return TBL_OK;
}

}  // namespace tbl
