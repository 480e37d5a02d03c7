// Compact finite-difference derivatives (derivs_par.f:80-177) on gfx950.
//
// Arrays are in the caller's layout: complex (nx, ny, K), x fastest.  Each
// kernel reads its input once, coalesced, into LDS and produces one or both
// derivatives of a direction from it (the first and the second derivative
// of one field share the load: 3 passes for the pair instead of 4).
//  * x (derparx/derparxx, rhsx/rhsxx + tridparx at one rank,
//    derivs_par.f:80-113, 259-397, 775-858): one wavefront per x-line (j, k);
//    the line goes to LDS with lane-strided loads, lane t then owns points
//    [t*R, t*R+R) and the tridiagonal Thomas solve is a per-lane affine map,
//    a 6-step shuffle scan and a replay, forward and backward.
//  * y (dery/deryy/deryfv, rhsy/rhsyy/rhsyfv + tridy, derivs_par.f:116-177,
//    400-611, 947-975): a workgroup holds a tile of 256/G columns x all ny
//    rows of one mode in LDS (rows loaded coalesced); G lanes solve each
//    column's line, R rows per lane, with the same scan form.  Results go
//    back through the tile and out as coalesced rows.
// The Thomas factors bet_j, gam_j (the reference's recurrence) and 1/bet_j
// are precomputed on the host; the solve multiplies by 1/bet_j where tridy
// divides by bet_j (last-bit rounding differs).
#include "tbl_kernels.h"

namespace tbl {

using C = cx<double>;

// a where c, else b (component-wise selects, no indexed lookup)
__device__ __forceinline__ C sel(bool c, C a, C b)
{
    return {c ? a.re : b.re, c ? a.im : b.im};
}
__device__ __forceinline__ void lds_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Thomas factors of a scheme in lane-segment order (built on the host for
// the kernels' G lanes x R rows): entry r*G + t holds row t*R + r of a,
// 1/bet and gam of the next row (rows past n: 0, 1, 0), so a lane's loads
// of one r are coalesced across the lanes
struct Thomas {
    const double *a, *rbet, *gam1;
};
__device__ __forceinline__ Thomas thomas_seg(const DerivArgs &A, int kind)
{
    const bool x = kind <= 1;
    const size_t per = static_cast<size_t>(x ? 64 * A.xR : A.yG * A.yR);
    const double *t = A.tseg + (x ? kind * 3 * per
                                   : 2 * 3 * static_cast<size_t>(64 * A.xR) + (kind - 2) * 3 * per);
    return {t, t + per, t + 2 * per};
}

// Tridiagonal solve of the n rows held R per lane by G lanes (lane t owns
// rows t*R .. t*R+R-1; u holds their right-hand sides on entry and the
// solution on exit): forward u_j = (r_j - a_j u_{j-1}) / bet_j and backward
// x_j = u_j - gam_{j+1} x_{j+1} (tridy), each as a per-lane affine map of the
// incoming state, an inclusive shuffle scan over the lanes and a replay.
template <int G, int R>
__device__ __forceinline__ void thomas_scan(const Thomas &T, int lane, C (&u)[R])
{
    double M = 1.0;
    C cc = {0, 0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const double a = T.a[r * G + lane], rb = T.rbet[r * G + lane];
        M = (-a * rb) * M;
        cc = (u[r] - a * cc) * rb;
    }
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const double Mo = __shfl_up(M, d, G);
        const C co = shfl_up(cc, d, G);
        if (lane >= d) {
            cc = M * co + cc;
            M = M * Mo;
        }
    }
    C prev = shfl_up(cc, 1, G);
    if (lane == 0) prev = {0, 0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const double a = T.a[r * G + lane], rb = T.rbet[r * G + lane];
        const C v = (u[r] - a * prev) * rb;
        u[r] = v;
        prev = v;
    }
    M = 1.0;
    cc = {0, 0};
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {
        const double g = T.gam1[r * G + lane];
        M = -g * M;
        cc = u[r] - g * cc;
    }
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const double Mo = __shfl_down(M, d, G);
        const C co = shfl_down(cc, d, G);
        if (lane + d < G) {
            cc = M * co + cc;
            M = M * Mo;
        }
    }
    C next = shfl_down(cc, 1, G);
    if (lane == G - 1) next = {0, 0};
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {
        const double g = T.gam1[r * G + lane];
        const C v = u[r] - g * next;
        u[r] = v;
        next = v;
    }
}

// Right-hand sides of the x schemes (rhsx, derivs_par.f:259-302; rhsxx,
// 351-397) on a line f of n points, in the reference's sums; the divisions
// by 12 dx, 120 dx, ... are multiplications by their reciprocals.  The four
// closure points (1-based 1, 2, n-1, n) are the same for every lane of the
// line and are formed once per lane as wavefront-uniform values; interior
// points take the centred stencil.
struct XRhs {
    C p1, p2, pn1, pn;   // closure points 1, 2, n-1, n
};
__device__ __forceinline__ XRhs rhs_x_closures(const C *f, int n, double dx, double dxx,
                                               bool second)
{
#define F(I) f[(I) - 1]
    XRhs b;
    if (!second) {
        b.p1 = (-25.0 * F(1) + 48.0 * F(2) - 36.0 * F(3) + 16.0 * F(4) - 3.0 * F(5)) *
               (1.0 / (12.0 * dx));
        b.p2 = (-406.0 * F(1) - 300.0 * F(2) + 760.0 * F(3) - 80.0 * F(4) + 30.0 * F(5)
                - 4.0 * F(6)) * (1.0 / (120.0 * dx));
        b.pn1 = (-406.0 * F(n) - 300.0 * F(n - 1) + 760.0 * F(n - 2) - 80.0 * F(n - 3)
                 + 30.0 * F(n - 4) - 4.0 * F(n - 5)) * (1.0 / (-120.0 * dx));
        b.pn = (-74.0 * F(n) + 16.0 * F(n - 1) + 72.0 * F(n - 2) - 16.0 * F(n - 3)
                + 2.0 * F(n - 4)) * (1.0 / (-24.0 * dx));
    } else {
        b.p1 = (9775.0 * F(1) - 20285.0 * F(2) + 11170.0 * F(3) - 550.0 * F(4) - 145.0 * F(5)
                + 35.0 * F(6)) * (1.0 / (60.0 * dxx));
        b.p2 = (4834.0 * F(1) - 8424.0 * F(2) + 1890.0 * F(3) + 2320.0 * F(4) - 810.0 * F(5)
                + 216.0 * F(6) - 26.0 * F(7)) * (1.0 / (360.0 * dxx));
        b.pn1 = (4834.0 * F(n) - 8424.0 * F(n - 1) + 1890.0 * F(n - 2) + 2320.0 * F(n - 3)
                 - 810.0 * F(n - 4) + 216.0 * F(n - 5) - 26.0 * F(n - 6)) * (1.0 / (360.0 * dxx));
        b.pn = (9775.0 * F(n) - 20285.0 * F(n - 1) + 11170.0 * F(n - 2) - 550.0 * F(n - 3)
                - 145.0 * F(n - 4) + 35.0 * F(n - 5)) * (1.0 / (60.0 * dxx));
    }
#undef F
    return b;
}
// interior point i (0-based, 2 <= i <= n-3) from f[i-2 .. i+2]
__device__ __forceinline__ C rhs_x_centre(C m2, C m1, C c0, C p1, C p2, double s, bool second)
{
    if (!second) return (p2 - m2 + 28.0 * (p1 - m1)) * s;
    return (3.0 * (p2 + m2) + 48.0 * (p1 + m1) - 102.0 * c0) * s;
}

// x-lines: o1 = d/dx, o2 = d2/dx2 of f (either may be null).  LDS: the
// input lines of the workgroup's kDerivXLines wavefronts, ldx complex each,
// then one output chunk of ch points (a multiple of 64) per wavefront.  A
// solved line goes from the lanes' segments back to coalesced rows through
// LDS: the last derivative of the call through its own input line (no
// longer read once its right-hand side is formed), in one pass; the first
// of two through the chunk buffer, ch points a pass.  The launch sizes ch
// so that a workgroup stays within kDerivXWgLds, two workgroups a CU (two
// wavefronts per SIMD, the occupancy the kernel's registers allow).
constexpr int kDerivXLines = 4;
constexpr int kDerivXWgLds = 80 * 1024;

template <int R>
__global__ void __launch_bounds__(64 * kDerivXLines) deriv_x_kernel(DerivArgs A, int ldx, int ch,
                                                                    const C *__restrict__ f,
                                                                    C *o1, C *o2)
{
    extern __shared__ C xline[];
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    const int nx = A.nx;
    const size_t line = static_cast<size_t>(blockIdx.x) * kDerivXLines + w;
    if (line >= static_cast<size_t>(A.ny) * A.K) return;   // whole wavefront
    C *Ls = xline + static_cast<size_t>(w) * ldx;
    const C *fl = f + line * nx;
    {
        // every load issued before the first is waited for (indices clamped
        // into the line; a load under a lane mask would wait per point)
        C v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = fl[min(lane + 64 * r, nx - 1)];
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (lane + 64 * r < nx) Ls[lane + 64 * r] = v[r];
    }
    lds_wave_sync();
    const int i0 = lane * R;
    C *Os = xline + static_cast<size_t>(kDerivXLines) * ldx + w * ch;   // output chunk
#pragma unroll 1
    for (int second = 0; second < 2; ++second) {
        C *o = second ? o2 : o1;
        if (!o) continue;
        const XRhs b = rhs_x_closures(Ls, nx, A.dx, A.dxx, second);
        const double s = second ? 1.0 / (4.0 * A.dxx) : 1.0 / (12.0 * A.dx);
        // the lane's points and two more each side (clamped into the line)
        C wv[R + 4];
#pragma unroll
        for (int q = 0; q < R + 4; ++q) wv[q] = Ls[min(max(i0 + q - 2, 0), nx - 1)];
        C u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = i0 + r;
            C v = rhs_x_centre(wv[r], wv[r + 1], wv[r + 2], wv[r + 3], wv[r + 4], s, second);
            v = sel(i == 0, b.p1, v);
            v = sel(i == 1, b.p2, v);
            v = sel(i == nx - 2, b.pn1, v);
            v = sel(i == nx - 1, b.pn, v);
            u[r] = sel(i < nx, v, C{0, 0});
        }
        thomas_scan<64, R>(thomas_seg(A, second), lane, u);
        C *ol = o + line * nx;
        if (second || !o2) {
            // the call's last derivative: out through the input line
            lds_wave_sync();
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (i0 + r < nx) Ls[i0 + r] = u[r];
            lds_wave_sync();
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (lane + 64 * r < nx) ol[lane + 64 * r] = Ls[lane + 64 * r];
            break;
        }
        // the lane's points through the output chunk, out as coalesced stores
#pragma unroll 1
        for (int c0 = 0; c0 < nx; c0 += ch) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int e = i0 + r - c0;
                if (e >= 0 && e < ch && i0 + r < nx) Os[e] = u[r];
            }
            lds_wave_sync();
#pragma unroll 1
            for (int e = lane; e < ch; e += 64)
                if (c0 + e < nx) ol[c0 + e] = Os[e];
            lds_wave_sync();
        }
    }
}

// Right-hand sides of the y schemes (rhsy 400-449, rhsyfv 504-549 with the
// wall row 0, rhsyy 552-611) on a column f of n rows: the four closure rows
// per column (wavefront-uniform for the column's lanes) and the centred rows
// j (0-based, 2 <= j <= n-3), with the metric of the spacing the generator
// normalised to (v_qdy(j-2), v_qdy2(j-2, 1); 1-based j in the reference).
__device__ __forceinline__ XRhs rhs_y1_closures(const DerivArgs &A, const C *f, int n, bool fv)
{
    const double *qdy = A.qdy;
    const double *sp = A.sp_fd, *pp = A.pp_fd, *lp = A.lp_fd, *fe = A.fp_e;
#define FY(J) f[(J) - 1]
    XRhs b;
    b.p1 = fv ? C{0, 0}
              : qdy[0] * (fe[1] * FY(1) + fe[2] * FY(2) + fe[3] * FY(3) + fe[4] * FY(4)
                          + fe[5] * FY(5));
    b.p2 = qdy[0] * (sp[3] * FY(1) + sp[4] * FY(2) + sp[5] * FY(3) + sp[6] * FY(4)
                     + sp[7] * FY(5) + sp[8] * FY(6));
    b.pn1 = qdy[n - 6] * (pp[3] * FY(n) + pp[4] * FY(n - 1) + pp[5] * FY(n - 2)
                          + pp[6] * FY(n - 3) + pp[7] * FY(n - 4) + pp[8] * FY(n - 5));
    b.pn = qdy[n - 5] * (lp[2] * FY(n) + lp[3] * FY(n - 1) + lp[4] * FY(n - 2)
                         + lp[5] * FY(n - 3) + lp[6] * FY(n - 4));
    return b;
}
__device__ __forceinline__ XRhs rhs_y2_closures(const DerivArgs &A, const C *f, int n)
{
    const double *qdy2 = A.qdy2;
    const double *fp = A.fp_sd, *sp = A.sp_sd, *pp = A.pp_sd, *lp = A.lp_sd;
    XRhs b;
    b.p1 = (fp[2] * FY(1) + fp[3] * FY(2) + fp[4] * FY(3) + fp[5] * FY(4) + fp[6] * FY(5)
            + fp[7] * FY(6)) * qdy2[0];
    b.p2 = (sp[3] * FY(1) + sp[4] * FY(2) + sp[5] * FY(3) + sp[6] * FY(4) + sp[7] * FY(5)
            + sp[8] * FY(6) + sp[9] * FY(7)) * qdy2[0];
    b.pn1 = (pp[3] * FY(n) + pp[4] * FY(n - 1) + pp[5] * FY(n - 2) + pp[6] * FY(n - 3)
             + pp[7] * FY(n - 4) + pp[8] * FY(n - 5) + pp[9] * FY(n - 6)) * qdy2[n - 7];
    b.pn = (lp[2] * FY(n) + lp[3] * FY(n - 1) + lp[4] * FY(n - 2) + lp[5] * FY(n - 3)
            + lp[6] * FY(n - 4) + lp[7] * FY(n - 5)) * qdy2[n - 6];
#undef FY
    return b;
}

// y-lines: o1 = d/dy (fv: with the wall row forced, deryfv), o2 = d2/dy2 of
// f (either may be null).  Workgroup: TX = 256/G columns x all rows of one
// mode in LDS, column-major with stride ldy (odd, against bank conflicts).
template <int G, int R>
__global__ void __launch_bounds__(256) deriv_y_kernel(DerivArgs A, int ldy, const C *__restrict__ f,
                                                      C *o1, int fv, C *o2)
{
    constexpr int TX = 256 / G;
    extern __shared__ C ytile[];
    const int nx = A.nx, ny = A.ny;
    // workgroup n takes tile (n mod 8) * ceil(B/8) + n / 8 of the B tiles
    // (x fastest, then mode): the hardware deals workgroups to the 8 XCDs
    // round-robin, so neighbouring tiles -- whose rows share the cache line
    // at their boundary -- run on one XCD and share its L2
    const int tiles_x = (nx + TX - 1) / TX;
    const int B = tiles_x * A.K, per = (B + 7) / 8;
    const int m = static_cast<int>(blockIdx.x % 8) * per + static_cast<int>(blockIdx.x / 8);
    if (m >= B) return;
    const int x0 = (m % tiles_x) * TX, k = m / tiles_x;
    const size_t mb = static_cast<size_t>(k) * nx * ny;
    // tile load: TX consecutive x of each row (one contiguous run per row);
    // TX * ny <= 256 R, so R steps of the workgroup cover the tile, and all
    // of a thread's loads are issued before the first is waited for
    {
        C v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int e = min(static_cast<int>(threadIdx.x) + 256 * q, TX * ny - 1);
            const int x = e % TX, j = e / TX;
            const int xi = x0 + x < nx ? x0 + x : nx - 1;
            v[q] = f[mb + static_cast<size_t>(j) * nx + xi];
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int e = static_cast<int>(threadIdx.x) + 256 * q;
            if (e < TX * ny) {
                const int x = e % TX, j = e / TX;
                ytile[x * ldy + j] = x0 + x < nx ? v[q] : C{0, 0};
            }
        }
    }
    __syncthreads();
    const int col = threadIdx.x / G, lane = threadIdx.x % G;
    const C *Tc = ytile + col * ldy;
    const int j0 = lane * R;
    // the lane's rows and two more each side (clamped into the column)
    C wv[R + 4];
#pragma unroll
    for (int q = 0; q < R + 4; ++q) wv[q] = Tc[min(max(j0 + q - 2, 0), ny - 1)];
    const double *cp1 = A.cp_fd, *cp2 = A.cp_sd;
    C u1[R], u2[R];
    if (o1) {
        const XRhs b = rhs_y1_closures(A, Tc, ny, fv != 0);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = j0 + r;
            const double q = A.qdy[min(max(j - 2, 0), ny - 1)];
            C v = q * (cp1[3] * wv[r] + cp1[4] * wv[r + 1] + cp1[5] * wv[r + 2]
                       + cp1[6] * wv[r + 3] + cp1[7] * wv[r + 4]);
            v = sel(j == 0, b.p1, v);
            v = sel(j == 1, b.p2, v);
            v = sel(j == ny - 2, b.pn1, v);
            v = sel(j == ny - 1, b.pn, v);
            u1[r] = sel(j < ny, v, C{0, 0});
        }
    }
    if (o2) {
        const XRhs b = rhs_y2_closures(A, Tc, ny);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = j0 + r;
            const double q = A.qdy2[min(max(j - 2, 0), ny - 1)];
            C v = (cp2[3] * wv[r] + cp2[4] * wv[r + 1] + cp2[5] * wv[r + 2] + cp2[6] * wv[r + 3]
                   + cp2[7] * wv[r + 4]) * q;
            v = sel(j == 0, b.p1, v);
            v = sel(j == 1, b.p2, v);
            v = sel(j == ny - 2, b.pn1, v);
            v = sel(j == ny - 1, b.pn, v);
            u2[r] = sel(j < ny, v, C{0, 0});
        }
    }
    __syncthreads();   // the tile is rewritten with the results below
    // solve, results through the tile, coalesced rows out
    auto finish = [&](C(&u)[R], const Thomas &T, C *o) {
        thomas_scan<G, R>(T, lane, u);
        C *Tw = ytile + col * ldy;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = lane * R + r;
            if (j < ny) Tw[j] = u[r];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < TX * ny; e += 256) {
            const int x = e % TX, j = e / TX;
            if (x0 + x < nx) o[mb + static_cast<size_t>(j) * nx + x0 + x] = ytile[x * ldy + j];
        }
        __syncthreads();
    };
    if (o1) finish(u1, thomas_seg(A, fv ? 4 : 2), o1);
    if (o2) finish(u2, thomas_seg(A, 3), o2);
}

// ---------------------------------------------------------------------------
// Long lines: x-lines of more than 64 * 17 points and y-lines of more than
// 64 * 9 rows take these kernels (no LDS stage; the LDS-staged forms would
// need more than 256 registers per lane there).  x: one wavefront per line,
// lane t owns points [t*R, t*R+R) read from global memory, divisions as in
// tridparx; y: one thread per column, the forward sweep stored in the output
// and re-read by the backward sweep (tridy's order).
// ---------------------------------------------------------------------------
// d/dx right-hand side (rhsx, derivs_par.f:259-302) and d2/dx2 (rhsxx, 351-397)
__device__ __forceinline__ C rhs_x_point(const C *__restrict__ f, int i, int n, double dx, double dxx,
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
__global__ void __launch_bounds__(256) deriv_x_line_kernel(DerivArgs A, int second, const C *f, C *out)
{
    const int lane = threadIdx.x & 63;
    const size_t line = static_cast<size_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (line >= static_cast<size_t>(A.ny) * A.K) return;
    const int n = A.nx;
    const C *fl = f + line * n;
    C *ol = out + line * n;
    const double *ta = A.thomas + static_cast<size_t>(second) * kThomasFields * A.nmax;
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
            rr = rhs_x_point(fl, i, n, A.dx, A.dxx, second);
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
__global__ void __launch_bounds__(64) deriv_y_line_kernel(DerivArgs A, int kind, const C *f, C *out)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= A.nx) return;
    const int nx = A.nx, n = A.ny;
    const C *fl = f + static_cast<size_t>(k) * nx * n + i;
    C *ol = out + static_cast<size_t>(k) * nx * n + i;
    const int tk = kind - 2 + 2;   // thomas table index (kinds 2..4)
    const double *ta = A.thomas + static_cast<size_t>(tk) * kThomasFields * A.nmax;
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


#define TBL_XR(X) X(1) X(2) X(3) X(5) X(7) X(9) X(17)
#define TBL_YR(X) X(1) X(2) X(3) X(4) X(5) X(7) X(9)


// Lane shapes of the LDS-staged kernels: x: 64 lanes, R = xR rows (0: the
// long-line kernel); y: yG lanes, yR rows (0: the long-line kernel)
void deriv_shapes(int nx, int ny, int *xR, int *yG, int *yR)
{
    const int needx = (nx + 63) / 64;
    *xR = 0;
    for (int r : {1, 2, 3, 5, 7, 9, 17})
        if (needx <= r) { *xR = r; break; }
    *yG = (ny + 31) / 32 <= 9 ? 32 : 64;
    const int needy = (ny + *yG - 1) / *yG;
    *yR = 0;
    if (ny <= 64 * 9)
        for (int r : {1, 2, 3, 4, 5, 7, 9})
            if (needy <= r) { *yR = r; break; }
}

// output chunk of deriv_x_kernel: the largest multiple of 64 points (at
// least 64, at most the line rounded up) that keeps a workgroup's LDS within
// kDerivXWgLds; lines up to 1088 points leave at least 128
int deriv_x_chunk(int nx)
{
    const int ldx = nx | 1;
    const int room = kDerivXWgLds / static_cast<int>(sizeof(C)) / kDerivXLines - ldx;
    return std::max(64, std::min((nx + 63) / 64 * 64, room / 64 * 64));
}

static int deriv_x(const DerivArgs &A, const C *f, C *o1, C *o2, hipStream_t st)
{
    const int need = (A.nx + 63) / 64;
    const size_t lines = static_cast<size_t>(A.ny) * A.K;
    if (A.xR == 0) {
        if (need > 33) return TBL_ERR_GRID;
        const dim3 g4(static_cast<unsigned>((lines + 3) / 4));
        if (o1) hipLaunchKernelGGL(deriv_x_line_kernel<33>, g4, dim3(256), 0, st, A, 0, f, o1);
        if (o2) hipLaunchKernelGGL(deriv_x_line_kernel<33>, g4, dim3(256), 0, st, A, 1, f, o2);
        return TBL_OK;
    }
    const dim3 grid(static_cast<unsigned>((lines + kDerivXLines - 1) / kDerivXLines));
    const int ldx = A.nx | 1;
    const int ch = deriv_x_chunk(A.nx);
    const size_t lds = static_cast<size_t>(kDerivXLines) * (ldx + ch) * sizeof(C);
#define X(RR)                                                                             \
    if (A.xR == RR) {                                                                     \
        hipLaunchKernelGGL(deriv_x_kernel<RR>, grid, dim3(64 * kDerivXLines), lds, st, A, \
                           ldx, ch, f, o1, o2);                                           \
        return TBL_OK;                                                                    \
    }
    TBL_XR(X)
#undef X
    return TBL_ERR_GRID;
}

static int deriv_y(const DerivArgs &A, const C *f, C *o1, int fv, C *o2, hipStream_t st)
{
    const int ldy = A.ny | 1;
    if (A.yR == 0) {
        const dim3 g2((A.nx + 63) / 64, A.K);
        if (o1) hipLaunchKernelGGL(deriv_y_line_kernel, g2, dim3(64), 0, st, A, fv ? 4 : 2, f, o1);
        if (o2) hipLaunchKernelGGL(deriv_y_line_kernel, g2, dim3(64), 0, st, A, 3, f, o2);
        return TBL_OK;
    }
    const int G = A.yG;
    const int tiles = (A.nx + 256 / G - 1) / (256 / G) * A.K;
    const dim3 grid(static_cast<unsigned>((tiles + 7) / 8 * 8));
    const size_t lds = static_cast<size_t>(256 / G) * ldy * sizeof(C);
#define X(RR)                                                                                \
    if (A.yR == RR) {                                                                        \
        if (G == 32)                                                                         \
            hipLaunchKernelGGL((deriv_y_kernel<32, RR>), grid, dim3(256), lds, st, A, ldy, f, \
                               o1, fv, o2);                                                   \
        else                                                                                 \
            hipLaunchKernelGGL((deriv_y_kernel<64, RR>), grid, dim3(256), lds, st, A, ldy, f, \
                               o1, fv, o2);                                                   \
        return TBL_OK;                                                                       \
    }
    TBL_YR(X)
#undef X
    return TBL_ERR_GRID;
}

int deriv(const DerivArgs &A, int kind, const double *f, double *out, hipStream_t st)
{
    const C *fi = reinterpret_cast<const C *>(f);
    C *fo = reinterpret_cast<C *>(out);
    switch (kind) {
    case 0: return deriv_x(A, fi, fo, nullptr, st);
    case 1: return deriv_x(A, fi, nullptr, fo, st);
    case 2: return deriv_y(A, fi, fo, 0, nullptr, st);
    case 3: return deriv_y(A, fi, nullptr, 0, fo, st);
    case 4: return deriv_y(A, fi, fo, 1, nullptr, st);
    default: return TBL_ERR_ARG;
    }
}

int derivs4(const DerivArgs &A, const double *f, double *dx, double *dxx, double *dy, double *dyy,
            hipStream_t st)
{
    const C *fi = reinterpret_cast<const C *>(f);
    int rc = TBL_OK;
    if (dx || dxx)
        rc = deriv_x(A, fi, reinterpret_cast<C *>(dx), reinterpret_cast<C *>(dxx), st);
    if (rc == TBL_OK && (dy || dyy))
        rc = deriv_y(A, fi, reinterpret_cast<C *>(dy), 0, reinterpret_cast<C *>(dyy), st);
    return rc;
}

int deriv_kernel_attributes()
{
    hipError_t e = hipSuccess;
#define X(RR)                                                                                  \
    {                                                                                          \
        hipError_t r = hipFuncSetAttribute(reinterpret_cast<const void *>(&deriv_x_kernel<RR>), \
                                           hipFuncAttributeMaxDynamicSharedMemorySize,         \
                                           160 * 1024);                                        \
        if (r != hipSuccess) e = r;                                                            \
    }
    TBL_XR(X)
#undef X
#define X(RR)                                                                                  \
    for (const void *fn : {reinterpret_cast<const void *>(&deriv_y_kernel<32, RR>),            \
                           reinterpret_cast<const void *>(&deriv_y_kernel<64, RR>)}) {         \
        hipError_t r = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,     \
                                           160 * 1024);                                        \
        if (r != hipSuccess) e = r;                                                            \
    }
    TBL_YR(X)
#undef X
    return e == hipSuccess ? TBL_OK : TBL_ERR_HIP;
}

}  // namespace tbl
