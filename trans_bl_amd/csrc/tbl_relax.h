// Line relaxation shared by the smoother kernels (gfx950).
//
// Internal layout of every level: complex values, mode-major, then x column,
// then y contiguous with padded column stride ld (tbl::Level).  A y-line
// (one x column of one mode) is therefore one contiguous run of ny values.
//
// Line relaxation (sor, poisson_par.f:167-226): every line of a colour phase
// is solved by a group of G lanes of one wavefront; lane t owns rows
// [t*R, t*R+R).  The pentadiagonal system (LU factors from band5_poi,
// penta_par.f:308-364) is solved with a partitioned scan: each lane composes
// the affine map of its rows, the maps are combined with a log2(G)-step
// shuffle scan, and each lane replays its rows from the incoming state
// (forward substitution, then back substitution; banbk5_poi,
// penta_par.f:367-404).
#pragma once

#include "tbl_buf.h"
#include "tbl_kernels.h"

namespace tbl {

// ---------------------------------------------------------------------------
// x-direction stencils (poisson_par.f:343-540 and 543-793)
// ---------------------------------------------------------------------------

// Off-diagonal part used on the smoother's right-hand side.  side -1/+1 is
// the left/right border column (rhs_border1/2), 0 an interior column
// (rhs_center, grouped form on rows ny-1 and ny).
template <typename T>
__device__ __forceinline__ cx<T> xoff(const cx<T> *__restrict__ col, int ld, int j,
                                      int side, bool grouped)
{
    // col points at column i of the mode's plane; neighbours at +/- ld
    if (side < 0) {
        return T(10) * col[j - ld] - T(4) * col[j + ld] + T(14) * col[j + 2 * ld]
               - T(6) * col[j + 3 * ld] + col[j + 4 * ld];
    }
    if (side > 0) {
        return T(10) * col[j + ld] - T(4) * col[j - ld] + T(14) * col[j - 2 * ld]
               - T(6) * col[j - 3 * ld] + col[j - 4 * ld];
    }
    if (grouped)
        return -(col[j + 2 * ld] + col[j - 2 * ld]) + T(16) * (col[j + ld] + col[j - ld]);
    return T(-1) * col[j - 2 * ld] + T(16) * col[j - ld] + T(16) * col[j + ld]
           - T(1) * col[j + 2 * ld];
}

// ---------------------------------------------------------------------------
// one y-line relaxation by a group of G lanes
// ---------------------------------------------------------------------------
// The line-LU factors of the rows a lane owns, held in registers: l1, l2
// (forward multipliers), u1, u2, pivot (back substitution, as banbk5_poi
// uses them) and -u1/p, -u2/p, 1/p (the scan's map composition).  Rows past
// ny get the identity (zeros, unit pivot).
template <typename T>
constexpr bool kIsF32 = sizeof(T) == 4;

template <typename T, int R>
struct LineLU {
    T l1[R], l2[R], u1[R], u2[R], pv[R], m1[R], m2[R], rp[R];
};

template <typename T, int R>
__device__ __forceinline__ LineLU<T, R> load_line_lu(const T *__restrict__ lu, int ld, int j0,
                                                    int ny)
{
    LineLU<T, R> f;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        const bool ok = j < ny;
        f.l1[r] = ok ? lu[j] : T(0);
        f.l2[r] = ok ? lu[ld + j] : T(0);
        f.u1[r] = ok ? lu[2 * ld + j] : T(0);
        f.u2[r] = ok ? lu[3 * ld + j] : T(0);
        f.pv[r] = ok ? lu[4 * ld + j] : T(1);
        f.m1[r] = ok ? lu[5 * ld + j] : T(0);
        f.m2[r] = ok ? lu[6 * ld + j] : T(0);
        f.rp[r] = ok ? lu[7 * ld + j] : T(1);
    }
    return f;
}

// Same factors through a buffer resource over the LU table: off0 = byte
// offset of the line's table (field 0, row 0).  Rows past ny get the
// identity; their loads may run past the table and then read zero.
template <typename T, int R>
__device__ __forceinline__ LineLU<T, R> load_line_lu_buf(__amdgpu_buffer_rsrc_t rs, unsigned off0,
                                                        int ld, int j0, int ny)
{
    LineLU<T, R> f;
    const unsigned fb = static_cast<unsigned>(ld) * sizeof(T);
    const unsigned o = off0 + static_cast<unsigned>(j0) * sizeof(T);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool ok = j0 + r < ny;
        const unsigned e = o + r * sizeof(T);
        const T l1 = buf_scalar(rs, e, T(0)), l2 = buf_scalar(rs, e + fb, T(0));
        const T u1 = buf_scalar(rs, e + 2 * fb, T(0)), u2 = buf_scalar(rs, e + 3 * fb, T(0));
        const T pv = buf_scalar(rs, e + 4 * fb, T(0)), m1 = buf_scalar(rs, e + 5 * fb, T(0));
        const T m2 = buf_scalar(rs, e + 6 * fb, T(0)), rp = buf_scalar(rs, e + 7 * fb, T(0));
        f.l1[r] = ok ? l1 : T(0);
        f.l2[r] = ok ? l2 : T(0);
        f.u1[r] = ok ? u1 : T(0);
        f.u2[r] = ok ? u2 : T(0);
        f.pv[r] = ok ? pv : T(1);
        f.m1[r] = ok ? m1 : T(0);
        f.m2[r] = ok ? m2 : T(0);
        f.rp[r] = ok ? rp : T(1);
    }
    return f;
}

// The line's data is reached through accessors: getX(j) = off-diagonal x
// term of row j, getV(j) = current v, getS(j, r) = source of row j = lane*R + r,
// putV(j, v) = store.
template <typename T, int G, int R, class FX, class FV, class FS, class FW>
__device__ __forceinline__ void relax_core(const LevelArgs<T> &L, const LineLU<T, R> &f, T rf,
                                           int lane, FX getX, FV getV, FS getS, FW putV)
{
    const int ny = L.ny;
    const int j0 = lane * R;
    const T qx = L.qx;

    cx<T> vo[R], y[R];
    // x off-diagonal terms for rows j0-1 .. j0+R
    cx<T> X[R + 2];
#pragma unroll
    for (int r = -1; r <= R; ++r) {
        const int j = j0 + r;
        cx<T> x = {T(0), T(0)};
        if (j >= 0 && j < ny) x = getX(j);
        X[r + 1] = x;
    }
    // right-hand sides (rhs_border1/2, rhs_center)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        cx<T> v = {T(0), T(0)}, rhs = {T(0), T(0)};
        if (j < ny) {
            v = getV(j);
            if (j == 0) {
                rhs = v;
            } else if (j == 1) {
                rhs = getS(j, r) - qx * (L.sp[2] * X[r + 2] + L.sp[1] * X[r + 1]);
            } else if (j <= ny - 3) {
                rhs = getS(j, r) - qx * (L.cp[0] * X[r] + L.cp[1] * X[r + 1] + L.cp[2] * X[r + 2]);
            } else {
                rhs = getS(j, r) - qx * X[r + 1];
            }
        }
        vo[r] = v;
        y[r] = rhs;
    }

    // ---- forward substitution: y_j = (r_j - l2_j y_{j-2}) - l1_j y_{j-1}
    // state (a, b) = (y_{j-2}, y_{j-1}); map M_j = [[0,1],[-l2,-l1]]
    T p00 = 1, p01 = 0, p10 = 0, p11 = 1;
    cx<T> c0 = {0, 0}, c1 = {0, 0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const T a1 = f.l1[r], a2 = f.l2[r];
        const T n10 = -a2 * p00 - a1 * p10, n11 = -a2 * p01 - a1 * p11;
        p00 = p10; p01 = p11; p10 = n10; p11 = n11;
        const cx<T> nc1 = (y[r] - a2 * c0) - a1 * c1;
        c0 = c1; c1 = nc1;
    }
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const T q00 = __shfl_up(p00, d, G), q01 = __shfl_up(p01, d, G);
        const T q10 = __shfl_up(p10, d, G), q11 = __shfl_up(p11, d, G);
        const cx<T> e0 = shfl_up(c0, d, G), e1 = shfl_up(c1, d, G);
        if (lane >= d) {
            const cx<T> nc0 = p00 * e0 + p01 * e1 + c0;
            const cx<T> nc1 = p10 * e0 + p11 * e1 + c1;
            c0 = nc0; c1 = nc1;
            const T m00 = p00 * q00 + p01 * q10, m01 = p00 * q01 + p01 * q11;
            const T m10 = p10 * q00 + p11 * q10, m11 = p10 * q01 + p11 * q11;
            p00 = m00; p01 = m01; p10 = m10; p11 = m11;
        }
    }
    {
        cx<T> a = shfl_up(c0, 1, G), b = shfl_up(c1, 1, G);
        if (lane == 0) { a = {0, 0}; b = {0, 0}; }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const T a1 = f.l1[r], a2 = f.l2[r];
            const cx<T> yy = (y[r] - a2 * a) - a1 * b;
            a = b; b = yy;
            y[r] = yy;
        }
    }

    // ---- back substitution: x_j = ((y_j - u1_j x_{j+1}) - u2_j x_{j+2}) / piv_j
    // state (a, b) = (x_{j+1}, x_{j+2}); map [[-u1/p, -u2/p],[1, 0]]
    p00 = 1; p01 = 0; p10 = 0; p11 = 1;
    c0 = {0, 0}; c1 = {0, 0};
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {
        // composition only: precomputed -u1/p, -u2/p, 1/p (the replay below
        // divides by the pivot like the reference)
        const T m0 = f.m1[r], m1 = f.m2[r], rp = f.rp[r];
        const T n00 = m0 * p00 + m1 * p10, n01 = m0 * p01 + m1 * p11;
        p10 = p00; p11 = p01; p00 = n00; p01 = n01;
        const cx<T> nc0 = rp * y[r] + m0 * c0 + m1 * c1;
        c1 = c0; c0 = nc0;
    }
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const T q00 = __shfl_down(p00, d, G), q01 = __shfl_down(p01, d, G);
        const T q10 = __shfl_down(p10, d, G), q11 = __shfl_down(p11, d, G);
        const cx<T> e0 = shfl_down(c0, d, G), e1 = shfl_down(c1, d, G);
        if (lane + d < G) {
            const cx<T> nc0 = p00 * e0 + p01 * e1 + c0;
            const cx<T> nc1 = p10 * e0 + p11 * e1 + c1;
            c0 = nc0; c1 = nc1;
            const T m00 = p00 * q00 + p01 * q10, m01 = p00 * q01 + p01 * q11;
            const T m10 = p10 * q00 + p11 * q10, m11 = p10 * q01 + p11 * q11;
            p00 = m00; p01 = m01; p10 = m10; p11 = m11;
        }
    }
    {
        cx<T> a = shfl_down(c0, 1, G), b = shfl_down(c1, 1, G);
        if (lane == G - 1) { a = {0, 0}; b = {0, 0}; }
#pragma unroll
        for (int r = R - 1; r >= 0; --r) {
            const T pj = f.pv[r], w1 = f.u1[r], w2 = f.u2[r];
            // FP64 divides by the pivot as banbk5_poi does; the FP32
            // smoother (a correction solve under FP64 defects) multiplies
            // by the reciprocal instead of a ~10-instruction division
            const cx<T> t = (y[r] - w1 * a) - w2 * b;
            const cx<T> xx = kIsF32<T> ? t * f.rp[r] : t / pj;
            b = a; a = xx;
            y[r] = xx;
        }
    }
    // relaxation v <- v + rf (v* - v) on rows 2..ny (poisson_par.f:215-218)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j >= 1 && j < ny) putV(j, vo[r] + rf * (y[r] - vo[r]));
    }
}

// one y-line relaxation on a column reached directly (global or LDS)
template <typename T, int G, int R>
__device__ __forceinline__ void relax_line(const LevelArgs<T> &L, cx<T> *col,
                                           const cx<T> *__restrict__ Scol, const T *__restrict__ lu,
                                           int side, T rf, int lane)
{
    const int ny = L.ny, ld = L.ld;
    const LineLU<T, R> f = load_line_lu<T, R>(lu, ld, lane * R, ny);
    relax_core<T, G, R>(
        L, f, rf, lane,
        [&](int j) { return xoff<T>(col, ld, j, side, side == 0 && j >= ny - 2); },
        [&](int j) { return col[j]; }, [&](int j, int) { return Scol[j]; },
        [&](int j, cx<T> v) { col[j] = v; });
}

// column of line c in a phase: phase 0 = the two border columns (Fortran 2
// and nx-1), phases 1..3 = interior colours i = 3+p-1, 3+p-1+3, ... (Fortran)
__device__ __forceinline__ int phase_column(int phase, int c, int nx)
{
    if (phase == 0) return c == 0 ? 1 : nx - 2;
    return 2 + (phase - 1) + 3 * c;
}

__host__ __device__ inline int phase_lines(int phase, int nx)
{
    if (phase == 0) return 2;
    const int first = 2 + (phase - 1), last = nx - 3;
    return first > last ? 0 : (last - first) / 3 + 1;
}

// (G, R) table shared by both smoother forms
#define TBL_LINE_SHAPES(X) \
    X(16, 1) X(16, 2) X(16, 3) X(16, 4) X(16, 5) X(16, 7) X(16, 9) X(16, 13) X(16, 17) \
    X(32, 1) X(32, 2) X(32, 3) X(32, 4) X(32, 5) X(32, 7) X(32, 9) \
    X(64, 1) X(64, 2) X(64, 3) X(64, 4) X(64, 5) X(64, 7) X(64, 9) X(64, 13) X(64, 17)

inline unsigned nblk(size_t n, unsigned b = 256) { return static_cast<unsigned>((n + b - 1) / b); }

// Raise a kernel's dynamic-LDS limit to the CU's 160 KB.
inline hipError_t allow_full_lds(const void *f)
{
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

}  // namespace tbl
