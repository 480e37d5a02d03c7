// Right-hand side of the vorticity equations and the RK4 update of the
// reference's Goertler case (my_form 1): lterms_gv / nlterms_gv
// (nlterms_par.f:89-182) with the z transforms of f_to_p / p_to_f
// (303-428), the combination of drv_gv (main_par.f:1339-1403), the four
// stage updates (main_par.f:270-341) and the gv_pert disturbance rows; and
// the flat-plate form (my_form 0): lterms_fi / nlterms_fi (nlterms_par.f:
// 6-86) and drv_fi's combination (main_par.f:1264-1336).
//
// User layout (complex (nx, ny, K), x fastest); base-flow arrays real
// (nx, ny).  All kernels but the nonlinear one are point-wise and
// HBM-bound.  The nonlinear terms transform every (x, y) point's K modes to
// kphys points in z, multiply there and transform back: a workgroup takes
// kNlPts points, stages their 6 K modes in LDS, synthesises each physical
// point's six values as a direct real DFT from a twiddle table in LDS,
// forms the four products, and analyses them back mode by mode -- the same
// transform pair as the reference's packed real FFT (mode 0 real, the
// odd-even mode in mode 0's imaginary part), to rounding.
#include "tbl_rhs.h"

namespace tbl {

namespace {

using cxd = cx<double>;

__device__ __forceinline__ cxd mul_i(double s, cxd a) { return {-s * a.im, s * a.re}; }

// lterms_gv (nlterms_par.f:100-108) at one point
__device__ __forceinline__ void lin_terms(double uxb, double uyb, double wzb, cxd ux, cxd uy,
                                          cxd wx, cxd wy, cxd wz, cxd &a, cxd &b, cxd &c,
                                          cxd &d)
{
    a = uyb * wx - wy * uxb;
    b = uxb * wz + ux * wzb;
    c = -uyb * wz - uy * wzb;
    d = (2.0 * uxb) * ux;
}

// lterms_fi's mode 1 (nlterms_par.f:19-24): the disturbance of b and c
// taken against the base flow's own products
__device__ __forceinline__ void lin_terms_fi1(double uxb, double uyb, double wzb, cxd ux, cxd uy,
                                              cxd wx, cxd wy, cxd wz, cxd &a, cxd &b, cxd &c)
{
    const cxd dwz = {wz.re - wzb, wz.im}, dux = {ux.re - uxb, ux.im}, duy = {uy.re - uyb, uy.im};
    a = uyb * wx - wy * uxb;
    b = uxb * dwz + wzb * dux;
    c = -uyb * dwz - wzb * duy;
}

__global__ void lterms_kernel(RhsArgs r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, k = blockIdx.z;
    if (i >= r.nx) return;
    const size_t e2 = static_cast<size_t>(j) * r.nx + i;
    const size_t e = static_cast<size_t>(k) * r.nx * r.ny + e2;
    if (r.fi) {
        // lterms_fi: modes >= 2 as lterms_gv without d (nlterms_par.f:28-34)
        cxd d;
        if (k == 0)
            lin_terms_fi1(r.b.uxb[e2], r.b.uyb[e2], r.b.wzb[e2], r.ux[e], r.uy[e], r.wx[e],
                          r.wy[e], r.wz[e], r.a[e], r.bb[e], r.c[e]);
        else
            lin_terms(r.b.uxb[e2], r.b.uyb[e2], r.b.wzb[e2], r.ux[e], r.uy[e], r.wx[e],
                      r.wy[e], r.wz[e], r.a[e], r.bb[e], r.c[e], d);
        return;
    }
    lin_terms(r.b.uxb[e2], r.b.uyb[e2], r.b.wzb[e2], r.ux[e], r.uy[e], r.wx[e], r.wy[e],
              r.wz[e], r.a[e], r.bb[e], r.c[e], r.d[e]);
}

// nlterms_gv (FI: nlterms_fi, the first three products, no base-flow
// terms): kNlPts points per workgroup of 256 threads; LDS holds the six
// inputs' modes [f][m][p], the four products [q][n][p] and the twiddles
template <bool FI>
__global__ void __launch_bounds__(256) nlterms_kernel(RhsArgs r, int N)
{
    constexpr int NQ = FI ? 3 : 4;
    extern __shared__ double sm[];
    const int K = r.K, P = kNlPts;
    const size_t plane = static_cast<size_t>(r.nx) * r.ny;
    const size_t p0 = static_cast<size_t>(blockIdx.x) * P;
    const int np = static_cast<int>(plane - p0 < static_cast<size_t>(P) ? plane - p0 : P);
    cxd *M = reinterpret_cast<cxd *>(sm);                  // [6][K][P]
    double *Q = sm + 2 * 6 * K * P;                        // [4][N][P]
    double *cs = Q + 4 * N * P, *sn = cs + N;              // cos, sin (2 pi t / N)
    const cxd *in[6] = {r.ux, r.uy, r.uz, r.wx, r.wy, r.wz};
    for (int t = threadIdx.x; t < N; t += blockDim.x) {
        cs[t] = r.tw[t];
        sn[t] = r.tw[N + t];
    }
    for (int t = threadIdx.x; t < 6 * K * P; t += blockDim.x) {
        const int p = t % P, m = (t / P) % K, f = t / (P * K);
        M[t] = p < np ? in[f][static_cast<size_t>(m) * plane + p0 + p] : cxd{0.0, 0.0};
    }
    __syncthreads();
    // synthesis: v_n = Re c_0 + (Im c_0 / 2)(-1)^n + sum_m Re(c_m e^{-2 pi i m n / N})
    for (int t = threadIdx.x; t < N * P; t += blockDim.x) {
        const int p = t % P, n = t / P;
        double v[6];
        const double alt = (n & 1) ? -0.5 : 0.5;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const cxd c0 = M[(f * K) * P + p];
            v[f] = c0.re + alt * c0.im;
        }
        int idx = 0;
        for (int m = 1; m < K; ++m) {
            idx += n;
            if (idx >= N) idx -= N;
            const double cm = cs[idx], sm_ = sn[idx];
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                const cxd c = M[(f * K + m) * P + p];
                v[f] += c.re * cm + c.im * sm_;
            }
        }
        // ux uy uz wx wy wz -> a = uy wx - ux wy, b = ux wz - uz wx,
        // c = uz wy - uy wz, d = ux ux (nlterms_par.f:152-159)
        Q[(0 * N + n) * P + p] = v[1] * v[3] - v[0] * v[4];
        Q[(1 * N + n) * P + p] = v[0] * v[5] - v[2] * v[3];
        Q[(2 * N + n) * P + p] = v[2] * v[4] - v[1] * v[5];
        if (!FI) Q[(3 * N + n) * P + p] = v[0] * v[0];
    }
    __syncthreads();
    // analysis: c_0 = (1/N) sum v_n (real), c_m = (2/N) sum v_n e^{+2 pi i m n / N};
    // entries below 1e-14 flushed (nlterms_par.f:418-423); then the
    // base-flow terms added (nlterms_par.f:167-177)
    cxd *out[4] = {r.a, r.bb, r.c, r.d};
    for (int t = threadIdx.x; t < K * P; t += blockDim.x) {
        const int p = t % P, m = t / P;
        if (p >= np) continue;
        double re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0};
        int idx = 0;
        for (int n = 0; n < N; ++n) {
            const double cm = cs[idx], sm_ = sn[idx];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const double v = Q[(q * N + n) * P + p];
                re[q] += v * cm;
                im[q] += v * sm_;
            }
            idx += m;
            if (idx >= N) idx -= N;
        }
        const double sc = (m == 0 ? 1.0 : 2.0) / static_cast<double>(N);
        const size_t e2 = p0 + p;
        const size_t e = static_cast<size_t>(m) * plane + e2;
        cxd nl[4];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double x = re[q] * sc, y = m == 0 ? 0.0 : im[q] * sc;
            if (fabs(x) < 1e-14) x = 0.0;
            if (fabs(y) < 1e-14) y = 0.0;
            nl[q] = {x, y};
        }
        if (FI) {
            out[0][e] = nl[0];
            out[1][e] = nl[1];
            out[2][e] = nl[2];
            continue;
        }
        cxd la, lb, lc, ld;
        lin_terms(r.b.uxb[e2], r.b.uyb[e2], r.b.wzb[e2], M[(0 * K + m) * P + p],
                  M[(1 * K + m) * P + p], M[(3 * K + m) * P + p], M[(4 * K + m) * P + p],
                  M[(5 * K + m) * P + p], la, lb, lc, ld);
        out[0][e] = la + nl[0];
        out[1][e] = lb + nl[1];
        out[2][e] = lc + nl[2];
        out[3][e] = ld + nl[3];
    }
}

// drv_gv's combination (main_par.f:1380-1397), rows 2..ny
__global__ void drv_combine_kernel(DrvArgs r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y + 1, k = blockIdx.z;
    if (i >= r.nx) return;
    const size_t e2 = static_cast<size_t>(j) * r.nx + i;
    const size_t e = static_cast<size_t>(k) * r.nx * r.ny + e2;
    const double kb = -static_cast<double>(k) * r.beta;      // v_kb = i kb
    const double k2 = -static_cast<double>(k) * static_cast<double>(k) * r.beta * r.beta;
    const double rre = r.rre, fy = r.fac_y;
    if (r.fi) {
        // drv_fi's combination (main_par.f:1317-1331): no curvature terms
        r.dvx[e] = -r.dady[e] + mul_i(kb, r.bb[e]) + rre * (r.xx[e] + fy * r.xy[e] + k2 * r.wx[e]);
        r.dvy[e] = -mul_i(kb, r.c[e]) + r.dadx[e] + rre * (r.yx[e] + fy * r.yy[e] + k2 * r.wy[e]);
        r.dvz[e] = -r.dbdx[e] + r.dcdy[e] + rre * (r.zx[e] + fy * r.zy[e] + k2 * r.wz[e]);
        return;
    }
    const double vg = r.b.varg[e2], dvg = r.b.dvargdx[e2];
    const cxd d = r.d[e];
    r.dvx[e] = -r.dady[e] + mul_i(kb, r.bb[e]) - vg * mul_i(kb, d)
               + rre * (r.xx[e] + fy * r.xy[e] + k2 * r.wx[e]);
    r.dvy[e] = -mul_i(kb, r.c[e]) + r.dadx[e] + rre * (r.yx[e] + fy * r.yy[e] + k2 * r.wy[e]);
    r.dvz[e] = -r.dbdx[e] + r.dcdy[e] + vg * r.dddx[e] + dvg * d
               + rre * (r.zx[e] + fy * r.zy[e] + k2 * r.wz[e]);
}

// the RK4 stage updates (main_par.f:270-341) on rows 2..ny-1, columns i >= 3
__global__ void rk_update_kernel(RkArgs r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y + 1;
    const int k = blockIdx.z % r.K, q = blockIdx.z / r.K;
    if (i >= r.nx) return;
    const size_t e = (static_cast<size_t>(k) * r.ny + j) * r.nx + i;
    cxd *w = r.w[q], *w1 = r.w1[q], *dv1 = r.dv1[q];
    const cxd *dv2 = r.dv2[q];
    if (r.stage == 1) {
        const cxd v = w[e];
        w1[e] = v;
        w[e] = v + dv1[e] * r.h;
    } else if (r.stage == 4) {
        w[e] = w1[e] + r.h * (dv1[e] + dv2[e]);
    } else {
        const cxd d2 = dv2[e];
        w[e] = w1[e] + d2 * r.h;
        dv1[e] = dv1[e] + 2.0 * d2;
    }
}

// gv_pert's rows (main_par.f:1602-1607): uy(i,1,k), d2uydx2(i,k), k >= 2
__global__ void pert_kernel(cxd *uy, cxd *d2, const cxd *pu, const cxd *pd, int nx, int ny,
                            int i0, int n)
{
    const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y + 1;
    if (i >= i0 + n) return;
    const size_t e = static_cast<size_t>(k) * nx + i;
    uy[static_cast<size_t>(k) * ny * nx + i] = pu[e];
    d2[e] = pd[e];
}

}  // namespace

size_t nlterms_lds(int K, int N)
{
    return sizeof(double) * (2 * 6 * static_cast<size_t>(K) * kNlPts + 4 * static_cast<size_t>(N) * kNlPts + 2 * N);
}

int terms(const RhsArgs &r, int kphys, hipStream_t st)
{
    if (kphys == 0) {
        dim3 grid((r.nx + 255) / 256, r.ny, r.K);
        hipLaunchKernelGGL(lterms_kernel, grid, dim3(256), 0, st, r);
    } else {
        const size_t plane = static_cast<size_t>(r.nx) * r.ny;
        dim3 grid(static_cast<unsigned>((plane + kNlPts - 1) / kNlPts));
        if (r.fi)
            hipLaunchKernelGGL(nlterms_kernel<true>, grid, dim3(256), nlterms_lds(r.K, kphys), st,
                               r, kphys);
        else
            hipLaunchKernelGGL(nlterms_kernel<false>, grid, dim3(256), nlterms_lds(r.K, kphys), st,
                               r, kphys);
    }
    return TBL_OK;
}

int drv_combine(const DrvArgs &r, hipStream_t st)
{
    dim3 grid((r.nx + 255) / 256, r.ny - 1, r.K);
    hipLaunchKernelGGL(drv_combine_kernel, grid, dim3(256), 0, st, r);
    return TBL_OK;
}

int rk_update(const RkArgs &r, hipStream_t st)
{
    if (r.nx <= 2 || r.ny <= 2) return TBL_OK;
    dim3 grid((r.nx - 2 + 255) / 256, r.ny - 2, 3 * r.K);
    hipLaunchKernelGGL(rk_update_kernel, grid, dim3(256), 0, st, r);
    return TBL_OK;
}

int pert_rows(cx<double> *uy, cx<double> *d2, const cx<double> *pu, const cx<double> *pd,
              int nx, int ny, int K, int i0, int i1, hipStream_t st)
{
    const int n = i1 - i0 + 1;
    if (n <= 0 || K < 2) return TBL_OK;
    dim3 grid((n + 63) / 64, K - 1);
    hipLaunchKernelGGL(pert_kernel, grid, dim3(64), 0, st, uy, d2, pu, pd, nx, ny, i0, n);
    return TBL_OK;
}

int rhs_kernel_attributes()
{
    TBL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&nlterms_kernel<false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    TBL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&nlterms_kernel<true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    return TBL_OK;
}

}  // namespace tbl
