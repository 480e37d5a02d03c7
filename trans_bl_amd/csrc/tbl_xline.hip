// x-direction line kernels of the RK-stage update around poi_uy (the
// reference's `loop`, main_par.f:1762-1785): the pentadiagonal compact
// Poisson solves of poi_ux / poi_uz / wx_wall (loop_par.f:174-255, 428-688;
// lhs5 258-299, ludecomp/lusolver penta_par.f:6-208), the k = 1 continuity
// integration of poi_ux and wz_wall (loop_par.f:364-425, 457-502), the
// compact filter filter_trid (filter_par.f:144-332), the wall terms lapv and
// der2wydxdy, the buffer-zone damping bzone/bzonew and outuy's outflow row.
//
// Every x-line (one (j, k) row of a user-layout field, x contiguous) is one
// wavefront.  Its right-hand side is formed from coalesced global loads (a
// point's stencil neighbours are the next lanes' points, served by the
// cache) into one LDS line of 64 R points, and each linear recurrence along
// x -- the L and U substitutions of the banded LU, the two Thomas passes,
// the integration -- runs as a partitioned scan: lane l owns rows
// [l R, l R + R), solves them
// from a zero incoming state while tracking the 2x2 response of its last
// two values to the incoming ones, the 64 lanes combine those affine maps
// in a 6-step shuffle scan, and a second pass adds each lane's homogeneous
// correction.  Coefficient tables are stored lane-segmented ([row][lane]),
// so the per-row loads of a wavefront are one coalesced 512-byte access.
#include "tbl_xline.h"

namespace tbl {

namespace {

using cxd = cx<double>;

__device__ __forceinline__ cxd cz() { return {0.0, 0.0}; }

// i * s (s real): multiply by the purely imaginary v_kb = (0, s)
__device__ __forceinline__ cxd mul_i(double s, cxd a) { return {-s * a.im, s * a.re}; }

// z(i) = g(i) - P(i) z(i-1) - Q(i) z(i-2) (FWD) or with i+1, i+2 (backward),
// g = Z * S (HS; else g = Z, and HQ false: Q = 0); Z in LDS holds g on
// entry, z on exit.  Rows past the line must hold g = 0 with P = Q = 0 (they
// lead a backward pass).  RT > 0: R = RT rows per lane known at compile time
// (the row loops unroll, so a lane's table loads issue together); RT = 0: R
// at run time, the row loops unrolled by 4 (longer segments would hold too
// many registers; loading blocks of 4-16 rows ahead measured no faster).
template <bool FWD, int RT, bool HQ, bool HS>
__device__ void recur2(cxd *Z, int Rrt, const double *__restrict__ P, const double *__restrict__ Q,
                       const double *__restrict__ S, int lane)
{
    const int R = RT > 0 ? RT : Rrt;
    const int base = lane * R;
    cxd z1 = cz(), z2 = cz();
    double m11 = 1.0, m12 = 0.0, m21 = 0.0, m22 = 1.0;
    auto step = [&](int t) {
        const int r = FWD ? t : R - 1 - t;
        const int e = r * 64 + lane;
        const double p = P[e], q = HQ ? Q[e] : 0.0;
        cxd g = Z[base + r];
        if constexpr (HS) g = g * S[e];
        const cxd z = g - p * z1 - q * z2;
        const double n1 = -p * m11 - q * m21, n2 = -p * m12 - q * m22;
        z2 = z1;
        z1 = z;
        m21 = m11;
        m22 = m12;
        m11 = n1;
        m12 = n2;
        Z[base + r] = z;
    };
    if constexpr (RT > 0) {
#pragma unroll
        for (int t = 0; t < RT; ++t) step(t);
    } else {
#pragma unroll 4
        for (int t = 0; t < R; ++t) step(t);
    }
    // inclusive scan of the lanes' affine maps (state out = M state in + c)
    cxd c1 = z1, c2 = z2;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int src = FWD ? lane - d : lane + d;
        const bool ok = FWD ? lane >= d : lane + d < 64;
        const int sl = ok ? src : lane;
        const double o11 = __shfl(m11, sl), o12 = __shfl(m12, sl);
        const double o21 = __shfl(m21, sl), o22 = __shfl(m22, sl);
        const cxd oc1 = {__shfl(c1.re, sl), __shfl(c1.im, sl)};
        const cxd oc2 = {__shfl(c2.re, sl), __shfl(c2.im, sl)};
        if (ok) {
            // self after other: M = Ms Mo, c = Ms co + cs
            const cxd n1 = m11 * oc1 + m12 * oc2 + c1;
            const cxd n2 = m21 * oc1 + m22 * oc2 + c2;
            const double a11 = m11 * o11 + m12 * o21, a12 = m11 * o12 + m12 * o22;
            const double a21 = m21 * o11 + m22 * o21, a22 = m21 * o12 + m22 * o22;
            m11 = a11; m12 = a12; m21 = a21; m22 = a22;
            c1 = n1;
            c2 = n2;
        }
    }
    // incoming state of this lane = the composite of the lanes before it
    const int prev = FWD ? lane - 1 : lane + 1;
    const bool has = FWD ? lane > 0 : lane < 63;
    const int pl = has ? prev : lane;
    cxd h1 = {__shfl(c1.re, pl), __shfl(c1.im, pl)};
    cxd h2 = {__shfl(c2.re, pl), __shfl(c2.im, pl)};
    if (!has) {
        h1 = cz();
        h2 = cz();
    }
    auto fix = [&](int t) {
        const int r = FWD ? t : R - 1 - t;
        const int e = r * 64 + lane;
        const double p = P[e], q = HQ ? Q[e] : 0.0;
        const cxd h = -p * h1 - q * h2;
        h2 = h1;
        h1 = h;
        Z[base + r] = Z[base + r] + h;
    };
    if constexpr (RT > 0) {
#pragma unroll
        for (int t = 0; t < RT; ++t) fix(t);
    } else {
#pragma unroll 4
        for (int t = 0; t < R; ++t) fix(t);
    }
}

// ---- pentadiagonal x-Poisson: poi_ux (k >= 2), poi_uz, wx_wall -------------
// F = cA A + i cb B; the right-hand side rows are row 1: the line's own
// boundary value, rows 2 and nx-1: sB (F + 10 F + F), rows 3..nx-2:
// sI (F + 5.5 F + F), row nx: sL F.
template <int RT>
__global__ void __launch_bounds__(64) xpenta_kernel(XPenta a)
{
    extern __shared__ cxd lds[];
    const int lane = threadIdx.x;
    const int line = blockIdx.x;
    const int nx = a.nx, R = a.R, NP = 64 * R;
    int k;
    size_t ofA, ofX;
    if (a.kind == kXWall) {
        k = line + 1;
        ofA = static_cast<size_t>(k) * nx;                        // (nx, K) arrays
        ofX = static_cast<size_t>(k) * a.ny * nx;                 // row 1 of mode k
    } else {
        k = line / (a.ny - 1) + 1;
        const int j = line % (a.ny - 1) + 1;
        ofA = (static_cast<size_t>(k) * a.ny + j) * nx;
        ofX = ofA;
    }
    const double kb = -static_cast<double>(k) * a.beta;             // v_kb = i kb
    const double cA = a.kind == kXUx ? 1.0 : (a.kind == kXUz ? 1.0 : -1.0);
    const double cb = a.kind == kXUz ? -kb : kb;
    const double s8 = a.kind == kXUx ? -8.0 * a.dxx : 8.0 * a.dxx;
    const double sB = a.kind == kXUx ? -a.dxx / 12.0 : a.dxx / 12.0;
    const double kk = static_cast<double>(k) * static_cast<double>(k) * a.beta * a.beta;
    const double sL = a.kind == kXUx ? 1.0 / kk : -1.0 / kk;
    cxd *Z = lds;
    const cxd *A = a.A + ofA, *B = a.B + ofA;
    cxd *X = a.X + ofX;
    auto F = [&](int i) { return cA * A[i] + mul_i(cb, B[i]); };
    // rows past the line hold 0; interior rows take the centred form (the
    // boundary rows' branches diverge only in the first and last chunks)
    for (int i = lane; i < NP; i += 64) {
        cxd g = cz();
        if (i == 0) g = X[0];
        else if (i == 1) g = sB * (F(0) + 10.0 * F(1) + F(2));
        else if (i < nx - 2) g = s8 * (F(i - 1) + 5.5 * F(i) + F(i + 1));
        else if (i == nx - 2) g = sB * (F(nx - 1) + 10.0 * F(nx - 2) + F(nx - 3));
        else if (i == nx - 1) g = sL * F(nx - 1);
        Z[i] = g;
    }
    __syncthreads();
    const double *T = a.tab + static_cast<size_t>(k - 1) * 5 * NP;
    recur2<true, RT, true, false>(Z, R, T, T + NP, nullptr, lane);
    __syncthreads();
    recur2<false, RT, true, true>(Z, R, T + 3 * NP, T + 4 * NP, T + 2 * NP, lane);
    __syncthreads();
    for (int i = lane + 1; i < nx; i += 64) X[i] = Z[i];
}

// ---- k = 1 continuity integration (poi_ux) and wz_wall ---------------------
template <int RT>
__global__ void __launch_bounds__(64) xinteg_kernel(XInteg a)
{
    extern __shared__ cxd lds[];
    const int lane = threadIdx.x;
    const int line = blockIdx.x;
    const int nx = a.nx, R = a.R, NP = 64 * R;
    cxd *Z = lds;
    const bool ux = a.kind == kXIntUx;
    // ux: mode 1, row j = line + 2; wz: row 1 of mode k = line
    const size_t of = ux ? static_cast<size_t>(line + 1) * nx
                         : static_cast<size_t>(line) * a.ny * nx;
    cxd *X = a.X + of;
    const double kb = -static_cast<double>(line) * a.beta;
    const cxd *W = a.D + of, *Lp = ux ? nullptr : a.Lp + static_cast<size_t>(line) * nx;
    auto D = [&](int i) { return ux ? W[i] : mul_i(kb, W[i]) - Lp[i]; };
    const double sdx = a.sdx;
    for (int i = lane; i < NP; i += 64) {
        cxd g = cz();
        if (i == 0) {
            g = X[0];
        } else if (i == 1) {
            g = (sdx * (251.0 * D(0) + 646.0 * D(1) - 264.0 * D(2) + 106.0 * D(3) - 19.0 * D(4)))
                * (1.0 / 720.0);
        } else if (i < nx - 2) {
            g = (sdx * (281.0 * D(i - 2) + 2056.0 * D(i - 1) + 1176.0 * D(i) - 104.0 * D(i + 1)
                        + 11.0 * D(i + 2))) * (1.0 / 2430.0);
        } else if (i == nx - 2) {
            g = (sdx * (10.0 * D(i - 2) + 57.0 * D(i - 1) + 24.0 * D(i) - D(i + 1)))
                * (1.0 / 57.0);
        } else if (i == nx - 1) {
            g = (sdx * (251.0 * D(i) + 646.0 * D(i - 1) - 264.0 * D(i - 2) + 106.0 * D(i - 3)
                        - 19.0 * D(i - 4))) * (1.0 / 720.0);
        }
        Z[i] = g;
    }
    __syncthreads();
    recur2<true, RT, true, false>(Z, R, a.tab, a.tab + NP, nullptr, lane);
    __syncthreads();
    for (int i = lane + 1; i < nx; i += 64) X[i] = Z[i];
}

// ---- filter_trid: explicit 7-point right-hand side + tridiagonal solve -----
template <int RT>
__global__ void __launch_bounds__(64) xfilter_kernel(XFilter a)
{
    extern __shared__ cxd lds[];
    const int lane = threadIdx.x;
    const int line = blockIdx.x;
    const int nx = a.nx, R = a.R, NP = 64 * R, rows = a.ny - 1;
    const int q = line / (a.K * rows);
    const int rem = line % (a.K * rows);
    const int k = rem / rows, j = rem % rows + 1;
    cxd *Wl = (q == 0 ? a.w0 : q == 1 ? a.w1 : a.w2) + (static_cast<size_t>(k) * a.ny + j) * nx;
    cxd *Z = lds;
    // the line is read here and written back after the solve, by this wavefront
    const cxd *W = Wl;
    const double af = a.af, bf = a.bf, cf = a.cf, df = a.df;
    const int n = nx;
    // rows past the line hold 0; interior rows take the centred form (the
    // boundary rows' branches diverge only in the first and last chunks;
    // forming every row's candidate forms and selecting measured slower)
    for (int i = lane; i < NP; i += 64) {
        cxd g = cz();
        if (i == 0) g = (15.0 * W[0] + 4.0 * W[1] - 6.0 * W[2] + 4.0 * W[3] - W[4]) * (1.0 / 16.0);
        else if (i == 1) g = (W[0] + 12.0 * W[1] + 6.0 * W[2] - 4.0 * W[3] + W[4]) * (1.0 / 16.0);
        else if (i == 2) g = (-W[0] + 4.0 * W[1] + 10.0 * W[2] + 4.0 * W[3] - W[4]) * (1.0 / 16.0);
        else if (i < n - 3)
            g = af * W[i] + bf * (W[i + 1] + W[i - 1]) + cf * (W[i + 2] + W[i - 2])
                + df * (W[i + 3] + W[i - 3]);
        else if (i == n - 3)
            g = (-W[n - 1] + 4.0 * W[n - 2] + 10.0 * W[n - 3] + 4.0 * W[n - 4] - W[n - 5])
                * (1.0 / 16.0);
        else if (i == n - 2)
            g = (W[n - 1] + 12.0 * W[n - 2] + 6.0 * W[n - 3] - 4.0 * W[n - 4] + W[n - 5])
                * (1.0 / 16.0);
        else if (i == n - 1)
            g = (15.0 * W[n - 1] + 4.0 * W[n - 2] - 6.0 * W[n - 3] + 4.0 * W[n - 4] - W[n - 5])
                * (1.0 / 16.0);
        Z[i] = g;
    }
    __syncthreads();
    recur2<true, RT, false, true>(Z, R, a.tab, nullptr, a.tab + NP, lane);
    __syncthreads();
    recur2<false, RT, false, false>(Z, R, a.tab + 2 * NP, nullptr, nullptr, lane);
    __syncthreads();
    for (int i = lane + 1; i < nx; i += 64) Wl[i] = Z[i];
}

// lapv(i, k) = d2uydx2 + (w3 uy(i,1) + ... + w9 uy(i,7)) / (dy^2 w1) + v_k2b2 uy(i,1)
// (loop_par.f:193-203); dwydy(i, k) of der2wydxdy (loop_par.f:309-317)
__global__ void wall_terms_kernel(WallArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= a.nx) return;
    const size_t pl = static_cast<size_t>(a.nx) * a.ny;
    const cxd *u = a.uy + k * pl + i, *w = a.wy + k * pl + i;
    const size_t s = a.nx;
    cxd sw = a.w[2] * u[0];
    for (int t = 1; t < 7; ++t) sw = sw + a.w[2 + t] * u[t * s];
    const double k2b2 = -static_cast<double>(k) * static_cast<double>(k) * a.beta * a.beta;
    a.lapv[static_cast<size_t>(k) * a.nx + i] =
        a.d2uydx2[static_cast<size_t>(k) * a.nx + i] + sw / a.dyyw1 + k2b2 * u[0];
    cxd sd = a.c[1] * w[0];
    for (int t = 1; t < 7; ++t) sd = sd + a.c[1 + t] * w[t * s];
    a.dwydy[static_cast<size_t>(k) * a.nx + i] = a.rdy * sd;
}

// der2wydxdy's x stencils (loop_par.f:319-357) on dwydy
__global__ void der2wydxdy_kernel(WallArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const int n = a.nx;
    if (i >= n) return;
    const cxd *d = a.dwydy + static_cast<size_t>(k) * n;
    const double rdx = 1.0 / a.dx;
    cxd r;
    if (i == 0)
        r = (-14.7 * d[0] + 36.0 * d[1] - 45.0 * d[2] + 40.0 * d[3] - 22.5 * d[4] + 7.2 * d[5]
             - d[6]) * (rdx / 6.0);
    else if (i == 1)
        r = (-d[0] - 7.7 * d[1] + 15.0 * d[2] - 10.0 * d[3] + 5.0 * d[4] - 1.5 * d[5]
             + 0.2 * d[6]) * (rdx / 6.0);
    else if (i == 2)
        r = (0.2 * d[0] - 2.4 * d[1] - 3.5 * d[2] + 8.0 * d[3] - 3.0 * d[4] + 0.8 * d[5]
             - 0.1 * d[6]) * (rdx / 6.0);
    else if (i < n - 3)
        r = (45.0 * (d[i + 1] - d[i - 1]) + 9.0 * (d[i - 2] - d[i + 2]) + d[i + 3] - d[i - 3])
            * (rdx / 60.0);
    else if (i == n - 3)
        r = (d[n - 5] - 8.0 * d[n - 4] + 8.0 * d[n - 2] - d[n - 1]) * (rdx / 12.0);
    else if (i == n - 2)
        r = (2.0 * d[n - 1] + 3.0 * d[n - 2] - 6.0 * d[n - 3] + d[n - 4]) * (rdx / 6.0);
    else
        r = (3.0 * d[n - 1] - 4.0 * d[n - 2] + d[n - 3]) * (rdx / 2.0);
    a.d2wydxdy[static_cast<size_t>(k) * n + i] = r;
}

// bzone (rows j >= 2) / bzonew (row 1): w *= bdfc(bdt, i), bdt = 2 for k = 3
__global__ void bzone_kernel(BzoneArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y + a.j0;
    const int k = blockIdx.z;
    if (i >= a.nx) return;
    const double f = a.bdfc[2 * i + (k == 2 ? 1 : 0)];
    const size_t e = (static_cast<size_t>(k) * a.ny + j) * a.nx + i;
    a.wx[e] = f * a.wx[e];
    a.wy[e] = f * a.wy[e];
    a.wz[e] = f * a.wz[e];
    if (a.th) a.th[e] = f * a.th[e];
}

// outuy: uy(ptsx, 2..jmax, k) = 0 (loop_par.f:106-111)
__global__ void outflow_row_kernel(cxd *uy, int nx, int ny, int K)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y;
    if (j >= ny || k >= K) return;
    uy[(static_cast<size_t>(k) * ny + j) * nx + nx - 1] = cz();
}

}  // namespace

size_t xline_lds(int R) { return sizeof(cxd) * 64 * static_cast<size_t>(R); }

// rows per lane compiled for (up to 576-point lines, and the 1025-point
// lines of the pentadiagonal solve and the integration: 2-4 % faster there
// on 1025x193x64, while the filter's run-time form stayed faster); longer
// lines run the run-time R form
#define TBL_XLR_F(X) X(1) X(2) X(3) X(5) X(9)
#define TBL_XLR(X) TBL_XLR_F(X) X(17)

int xpenta(const XPenta &a, hipStream_t st)
{
    const int lines = a.kind == kXWall ? a.K - 1 : (a.ny - 1) * (a.K - 1);
    if (lines <= 0) return TBL_OK;
#define X(RR) \
    if (a.R == RR) { hipLaunchKernelGGL(xpenta_kernel<RR>, dim3(lines), dim3(64), xline_lds(a.R), st, a); return TBL_OK; }
    TBL_XLR(X)
#undef X
    hipLaunchKernelGGL(xpenta_kernel<0>, dim3(lines), dim3(64), xline_lds(a.R), st, a);
    return TBL_OK;
}

int xinteg(const XInteg &a, hipStream_t st)
{
    const int lines = a.kind == kXIntUx ? a.ny - 2 : a.K;
    if (lines <= 0) return TBL_OK;
#define X(RR) \
    if (a.R == RR) { hipLaunchKernelGGL(xinteg_kernel<RR>, dim3(lines), dim3(64), xline_lds(a.R), st, a); return TBL_OK; }
    TBL_XLR(X)
#undef X
    hipLaunchKernelGGL(xinteg_kernel<0>, dim3(lines), dim3(64), xline_lds(a.R), st, a);
    return TBL_OK;
}

int xfilter(const XFilter &a, hipStream_t st)
{
    const int lines = 3 * a.K * (a.ny - 1);
#define X(RR) \
    if (a.R == RR) { hipLaunchKernelGGL(xfilter_kernel<RR>, dim3(lines), dim3(64), xline_lds(a.R), st, a); return TBL_OK; }
    TBL_XLR_F(X)
#undef X
    hipLaunchKernelGGL(xfilter_kernel<0>, dim3(lines), dim3(64), xline_lds(a.R), st, a);
    return TBL_OK;
}

int wall_terms(const WallArgs &a, hipStream_t st)
{
    dim3 grid((a.nx + 255) / 256, a.K);
    hipLaunchKernelGGL(wall_terms_kernel, grid, dim3(256), 0, st, a);
    hipLaunchKernelGGL(der2wydxdy_kernel, grid, dim3(256), 0, st, a);
    return TBL_OK;
}

int bzone(const BzoneArgs &a, int rows, hipStream_t st)
{
    dim3 grid((a.nx + 255) / 256, rows, a.K);
    hipLaunchKernelGGL(bzone_kernel, grid, dim3(256), 0, st, a);
    return TBL_OK;
}

int outflow_row(cx<double> *uy, int nx, int ny, int K, hipStream_t st)
{
    dim3 grid((ny - 1 + 63) / 64, K);
    hipLaunchKernelGGL(outflow_row_kernel, grid, dim3(64), 0, st, uy, nx, ny, K);
    return TBL_OK;
}

int xline_kernel_attributes()
{
    const int lds = static_cast<int>(xline_lds(kXMaxR));
#define X(RR)                                                                             \
    for (const void *f : {reinterpret_cast<const void *>(&xpenta_kernel<RR>),             \
                          reinterpret_cast<const void *>(&xinteg_kernel<RR>),             \
                          reinterpret_cast<const void *>(&xfilter_kernel<RR>)})           \
        TBL_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    TBL_XLR(X)
    X(0)
#undef X
    return TBL_OK;
}

}  // namespace tbl
