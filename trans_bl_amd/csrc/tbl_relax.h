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

template <typename T>
constexpr bool kIsF32 = sizeof(T) == 4;

// Multiply-add forms with a precision-dependent rounding.  FP64 rounds every
// product and sum in the order the reference writes them (the results below
// are bit-identical to the plain expressions in the comments); FP32 -- the
// correction solve under FP64 defects -- fuses them into multiply-adds (one
// rounding each, and about half the vector instructions).
//   mac(a, x, c)           = a x + c
//   dot2(a, x, b, y)       = a x + b y
//   dot2c(a, x, b, y, c)   = (a x + b y) + c
// mac_f / dot2_f take the choice as a parameter (F: fused), for FP64 work
// that is not bound to the reference's rounding.
template <typename T>
__device__ __forceinline__ T fused_ma(T a, T x, T c)
{
    if constexpr (kIsF32<T>) return __builtin_fmaf(a, x, c);
    else return __builtin_fma(a, x, c);
}
template <bool F, typename T>
__device__ __forceinline__ T mac_f(T a, T x, T c)
{
    if constexpr (F) return fused_ma(a, x, c);
    else return a * x + c;
}
template <bool F, typename T>
__device__ __forceinline__ cx<T> mac_f(T a, cx<T> x, cx<T> c)
{
    return {mac_f<F>(a, x.re, c.re), mac_f<F>(a, x.im, c.im)};
}
template <bool F, typename T>
__device__ __forceinline__ T dot2_f(T a, T x, T b, T y)
{
    if constexpr (F) return fused_ma(b, y, a * x);
    else return a * x + b * y;
}
template <bool F, typename T>
__device__ __forceinline__ cx<T> dot2_f(T a, cx<T> x, T b, cx<T> y)
{
    return {dot2_f<F>(a, x.re, b, y.re), dot2_f<F>(a, x.im, b, y.im)};
}
template <typename T>
__device__ __forceinline__ T mac(T a, T x, T c) { return mac_f<kIsF32<T>>(a, x, c); }
template <typename T>
__device__ __forceinline__ cx<T> mac(T a, cx<T> x, cx<T> c) { return mac_f<kIsF32<T>>(a, x, c); }
template <typename T>
__device__ __forceinline__ T dot2(T a, T x, T b, T y) { return dot2_f<kIsF32<T>>(a, x, b, y); }
template <typename T>
__device__ __forceinline__ cx<T> dot2(T a, cx<T> x, T b, cx<T> y)
{
    return dot2_f<kIsF32<T>>(a, x, b, y);
}
template <typename T>
__device__ __forceinline__ T dot2c(T a, T x, T b, T y, T c)
{
    if constexpr (kIsF32<T>) return __builtin_fmaf(a, x, __builtin_fmaf(b, y, c));
    else return (a * x + b * y) + c;
}
template <typename T>
__device__ __forceinline__ cx<T> dot2c(T a, cx<T> x, T b, cx<T> y, cx<T> c)
{
    return {dot2c(a, x.re, b, y.re, c.re), dot2c(a, x.im, b, y.im, c.im)};
}

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
    // sums in the reference's order: 10 a - 4 b + 14 c - 6 d + e (border),
    // -(p2 + m2) + 16 (p1 + m1) (grouped), -m2 + 16 m1 + 16 p1 - p2
    if (side != 0) {
        const int s = side < 0 ? ld : -ld;
        return mac(T(-6), col[j + 3 * s],
                   mac(T(14), col[j + 2 * s], mac(T(-4), col[j + s], T(10) * col[j - s])))
               + col[j + 4 * s];
    }
    if (grouped)
        return mac(T(16), col[j + ld] + col[j - ld], -(col[j + 2 * ld] + col[j - 2 * ld]));
    return mac(T(-1), col[j + 2 * ld],
               mac(T(16), col[j + ld], mac(T(16), col[j - ld], T(-1) * col[j - 2 * ld])));
}

// ---------------------------------------------------------------------------
// one y-line relaxation by a group of G lanes
// ---------------------------------------------------------------------------
// The line-LU factors of the rows a lane owns, held in registers: l1, l2
// (forward multipliers), u1, u2, pivot (back substitution, as banbk5_poi
// uses them) and -u1/p, -u2/p, 1/p (the scan's map composition).  Rows past
// ny get the identity (zeros, unit pivot).

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

// Same factors through a buffer resource over the LU tables (row stride
// lus = lu_stride(ny)): off0 = byte offset of the line's table (field 0,
// row 0).  The tables hold the identity past ny, so every load is
// unconditional; the field offsets are wavefront-uniform.
template <typename T, int R>
__device__ __forceinline__ LineLU<T, R> load_line_lu_buf(__amdgpu_buffer_rsrc_t rs, unsigned off0,
                                                        int lus, int j0)
{
    LineLU<T, R> f;
    const int fb = lus * static_cast<int>(sizeof(T));
    const unsigned o = off0 + static_cast<unsigned>(j0) * sizeof(T);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const unsigned e = o + r * sizeof(T);
        f.l1[r] = buf_scalar(rs, e, T(0), 0);
        f.l2[r] = buf_scalar(rs, e, T(0), fb);
        f.u1[r] = buf_scalar(rs, e, T(0), 2 * fb);
        f.u2[r] = buf_scalar(rs, e, T(0), 3 * fb);
        f.pv[r] = buf_scalar(rs, e, T(0), 4 * fb);
        f.m1[r] = buf_scalar(rs, e, T(0), 5 * fb);
        f.m2[r] = buf_scalar(rs, e, T(0), 6 * fb);
        f.rp[r] = buf_scalar(rs, e, T(0), 7 * fb);
    }
    return f;
}

// Same from a plain pointer to the line's padded table (stride lus).
template <typename T, int R>
__device__ __forceinline__ LineLU<T, R> load_line_lu_tbl(const T *__restrict__ lu, int lus, int j0)
{
    LineLU<T, R> f;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        f.l1[r] = lu[j];
        f.l2[r] = lu[lus + j];
        f.u1[r] = lu[2 * lus + j];
        f.u2[r] = lu[3 * lus + j];
        f.pv[r] = lu[4 * lus + j];
        f.m1[r] = lu[5 * lus + j];
        f.m2[r] = lu[6 * lus + j];
        f.rp[r] = lu[7 * lus + j];
    }
    return f;
}

// ---------------------------------------------------------------------------
// the partitioned scan across the G lanes of a line
// ---------------------------------------------------------------------------
// Affine map of a run of rows: state' = P state + c (P 2x2 real, c two
// complex values).
template <typename T>
struct Aff {
    T p00, p01, p10, p11;
    cx<T> c0, c1;
};
template <typename T>
__device__ __forceinline__ Aff<T> aff_identity()
{
    return {T(1), T(0), T(0), T(1), {T(0), T(0)}, {T(0), T(0)}};
}
// a <- a after q (q's run feeds a's): c = P_a c_q + c_a, P = P_a P_q.  With
// q the identity the result is a again, bit for bit.
template <typename T>
__device__ __forceinline__ void aff_fold(Aff<T> &a, const Aff<T> &q)
{
    // c0 = P00 q.c0 + P01 q.c1 + c0, ..., P = P_a P_q
    const cx<T> nc0 = dot2c(a.p00, q.c0, a.p01, q.c1, a.c0);
    const cx<T> nc1 = dot2c(a.p10, q.c0, a.p11, q.c1, a.c1);
    a.c0 = nc0; a.c1 = nc1;
    const T m00 = dot2(a.p00, q.p00, a.p01, q.p10), m01 = dot2(a.p00, q.p01, a.p01, q.p11);
    const T m10 = dot2(a.p10, q.p00, a.p11, q.p10), m11 = dot2(a.p10, q.p01, a.p11, q.p11);
    a.p00 = m00; a.p01 = m01; a.p10 = m10; a.p11 = m11;
}

// DPP lane moves (a VALU operand modifier: no LDS round trip).  CTRL: row
// shift right by d (from lane - d of the same 16-lane row) = 0x110 + d, row
// shift left (from lane + d) = 0x100 + d, wavefront shift right / left by
// one = 0x138 / 0x130.  Lanes without a source keep `id`.
template <int CTRL>
__device__ __forceinline__ float dpp(float id, float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, id),
                                                                 __builtin_bit_cast(int, x), CTRL,
                                                                 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double id, double x)
{
    const unsigned long long i = __builtin_bit_cast(unsigned long long, id);
    const unsigned long long v = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_update_dpp(
        static_cast<int>(static_cast<unsigned>(i)), static_cast<int>(static_cast<unsigned>(v)), CTRL,
        0xF, 0xF, false));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_update_dpp(
        static_cast<int>(static_cast<unsigned>(i >> 32)),
        static_cast<int>(static_cast<unsigned>(v >> 32)), CTRL, 0xF, 0xF, false));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
template <int CTRL, typename T>
__device__ __forceinline__ cx<T> dpp(cx<T> id, cx<T> x)
{
    return {dpp<CTRL>(id.re, x.re), dpp<CTRL>(id.im, x.im)};
}
// Same with zero for lanes without a source (the move's bound-control
// form: no register to preset, and the move can merge into its user)
template <int CTRL>
__device__ __forceinline__ float dpp0(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x),
                                                                 CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp0(double x)
{
    const unsigned long long v = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_update_dpp(
        0, static_cast<int>(static_cast<unsigned>(v)), CTRL, 0xF, 0xF, true));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_update_dpp(
        0, static_cast<int>(static_cast<unsigned>(v >> 32)), CTRL, 0xF, 0xF, true));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
template <int CTRL, typename T>
__device__ __forceinline__ cx<T> dpp0(cx<T> x)
{
    return {dpp0<CTRL>(x.re), dpp0<CTRL>(x.im)};
}
// The map of the source lane, the identity where there is none: the zero
// entries of the identity (p01, p10, c) through the zero-filling form
template <int CTRL, typename T>
__device__ __forceinline__ Aff<T> dpp_aff(const Aff<T> &x)
{
    return {dpp<CTRL>(T(1), x.p00), dpp0<CTRL>(x.p01), dpp0<CTRL>(x.p10),
            dpp<CTRL>(T(1), x.p11), dpp0<CTRL>(x.c0),  dpp0<CTRL>(x.c1)};
}
template <typename T>
__device__ __forceinline__ Aff<T> aff_shfl(const Aff<T> &x, int src, int width)
{
    return {__shfl(x.p00, src, width), __shfl(x.p01, src, width), __shfl(x.p10, src, width),
            __shfl(x.p11, src, width), {__shfl(x.c0.re, src, width), __shfl(x.c0.im, src, width)},
            {__shfl(x.c1.re, src, width), __shfl(x.c1.im, src, width)}};
}
// value of lane - 1 / lane + 1 (the caller replaces it on a group's first /
// last lane)
template <typename T>
__device__ __forceinline__ cx<T> lane_prev(cx<T> x) { return dpp0<0x138>(x); }
template <typename T>
__device__ __forceinline__ cx<T> lane_next(cx<T> x) { return dpp0<0x130>(x); }

// Inclusive scan of the lanes' maps in lane order (lane t ends with the map
// of lanes 0..t): four DPP steps inside each 16-lane row, then the rows are
// joined through lane shuffles (one step for 32 lanes, two for 64).
template <typename T, int G>
__device__ __forceinline__ void scan_up(Aff<T> &m, int lane)
{
    aff_fold(m, dpp_aff<0x111>(m));
    aff_fold(m, dpp_aff<0x112>(m));
    aff_fold(m, dpp_aff<0x114>(m));
    aff_fold(m, dpp_aff<0x118>(m));
    if (G >= 32) {
        const Aff<T> q = aff_shfl(m, (lane & 32) + 15, G);   // end of the row below
        if (lane & 16) aff_fold(m, q);
    }
    if (G == 64) {
        const Aff<T> q = aff_shfl(m, 31, G);
        if (lane & 32) aff_fold(m, q);
    }
}
// Same in reverse lane order (lane t ends with the map of lanes t..G-1).
template <typename T, int G>
__device__ __forceinline__ void scan_down(Aff<T> &m, int lane)
{
    aff_fold(m, dpp_aff<0x101>(m));
    aff_fold(m, dpp_aff<0x102>(m));
    aff_fold(m, dpp_aff<0x104>(m));
    aff_fold(m, dpp_aff<0x108>(m));
    if (G >= 32) {
        const Aff<T> q = aff_shfl(m, (lane & 32) + 16, G);   // start of the row above
        if (!(lane & 16)) aff_fold(m, q);
    }
    if (G == 64) {
        const Aff<T> q = aff_shfl(m, 32, G);
        if (!(lane & 32)) aff_fold(m, q);
    }
}

// Rows of a y-line held by its lanes: row 0 (the wall value, an identity
// row of the line system) is not relaxed, so lane t owns rows
// line_j0(t, R) .. + R - 1 of rows 1..ny-1 and a shape needs G R >= ny - 1
// (ny = 1 + 2^k m leaves no idle lane slot where G divides 2^k m).  Loops
// that stage a line lane-strided (row lane + s G) take R + 1 steps, which
// reach row ny - 1.
__host__ __device__ inline int line_j0(int lane, int R) { return 1 + lane * R; }

// The line's data is reached through accessors: getX(j) = off-diagonal x
// term of row j, getV(j) = current v, getS(j, r) = source of row
// j = line_j0(lane, R) + r, putV(j, v) = store.
template <typename T, int G, int R, class FX, class FV, class FS, class FW>
__device__ __forceinline__ void relax_core(const LevelArgs<T> &L, const LineLU<T, R> &f, T rf,
                                           int lane, FX getX, FV getV, FS getS, FW putV)
{
    const int ny = L.ny;
    const int j0 = line_j0(lane, R);
    const T qx = L.qx;
    // row 0's solution is its own value (identity row: y_0 = x_0 = v_0); it
    // enters the rows above as the forward substitution's initial state
    const cx<T> v0 = getV(0);

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
            if (j == 1) {
                // S - qx (sp2 X(j+1) + sp1 X(j))
                rhs = mac(-qx, dot2(L.sp[2], X[r + 2], L.sp[1], X[r + 1]), getS(j, r));
            } else if (j <= ny - 3) {
                // S - qx ((cp0 X(j-1) + cp1 X(j)) + cp2 X(j+1))
                rhs = mac(-qx, mac(L.cp[2], X[r + 2], dot2(L.cp[0], X[r], L.cp[1], X[r + 1])),
                          getS(j, r));
            } else {
                rhs = mac(-qx, X[r + 1], getS(j, r));
            }
        }
        vo[r] = v;
        y[r] = rhs;
    }

    // ---- forward substitution: y_j = (r_j - l2_j y_{j-2}) - l1_j y_{j-1}
    // state (a, b) = (y_{j-2}, y_{j-1}); map M_j = [[0,1],[-l2,-l1]]
    Aff<T> m = aff_identity<T>();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const T a1 = f.l1[r], a2 = f.l2[r];
        // -a2 p0x - a1 p1x; (y - a2 c0) - a1 c1
        const T n10 = dot2(-a2, m.p00, -a1, m.p10), n11 = dot2(-a2, m.p01, -a1, m.p11);
        m.p00 = m.p10; m.p01 = m.p11; m.p10 = n10; m.p11 = n11;
        const cx<T> nc1 = mac(-a1, m.c1, mac(-a2, m.c0, y[r]));
        m.c0 = m.c1; m.c1 = nc1;
    }
    // the line's initial state (y_{-1}, y_0) = (0, v_0) enters through the
    // first lane's map: c += P (0, v_0)
    if (lane == 0) {
        m.c0 = mac(m.p01, v0, m.c0);
        m.c1 = mac(m.p11, v0, m.c1);
    }
    scan_up<T, G>(m, lane);
    {
        // state entering this lane's rows: the previous lane's inclusive map
        cx<T> a = lane_prev(m.c0), b = lane_prev(m.c1);
        if (lane == 0) { a = {0, 0}; b = v0; }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const T a1 = f.l1[r], a2 = f.l2[r];
            const cx<T> yy = mac(-a1, b, mac(-a2, a, y[r]));   // (y - a2 a) - a1 b
            a = b; b = yy;
            y[r] = yy;
        }
    }

    // ---- back substitution: x_j = ((y_j - u1_j x_{j+1}) - u2_j x_{j+2}) / piv_j
    // state (a, b) = (x_{j+1}, x_{j+2}); map [[-u1/p, -u2/p],[1, 0]]
    m = aff_identity<T>();
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {
        // composition only: precomputed -u1/p, -u2/p, 1/p (the replay below
        // divides by the pivot like the reference)
        const T m0 = f.m1[r], m1 = f.m2[r], rp = f.rp[r];
        const T n00 = dot2(m0, m.p00, m1, m.p10), n01 = dot2(m0, m.p01, m1, m.p11);
        m.p10 = m.p00; m.p11 = m.p01; m.p00 = n00; m.p01 = n01;
        // (rp y + m0 c0) + m1 c1
        const cx<T> nc0 = mac(m1, m.c1, dot2(rp, y[r], m0, m.c0));
        m.c1 = m.c0; m.c0 = nc0;
    }
    scan_down<T, G>(m, lane);
    {
        cx<T> a = lane_next(m.c0), b = lane_next(m.c1);
        if (lane == G - 1) { a = {0, 0}; b = {0, 0}; }
#pragma unroll
        for (int r = R - 1; r >= 0; --r) {
            const T pj = f.pv[r], w1 = f.u1[r], w2 = f.u2[r];
            // FP64 divides by the pivot as banbk5_poi does; the FP32
            // smoother (a correction solve under FP64 defects) multiplies
            // by the reciprocal instead of a ~10-instruction division
            const cx<T> t = mac(-w2, b, mac(-w1, a, y[r]));   // (y - w1 a) - w2 b
            const cx<T> xx = kIsF32<T> ? t * f.rp[r] : t / pj;
            b = a; a = xx;
            y[r] = xx;
        }
    }
    // relaxation v <- v + rf (v* - v) on rows 2..ny (poisson_par.f:215-218)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j < ny) putV(j, mac(rf, y[r] - vo[r], vo[r]));
    }
}

// one y-line relaxation on a column reached directly (global or LDS)
template <typename T, int G, int R>
__device__ __forceinline__ void relax_line(const LevelArgs<T> &L, cx<T> *col,
                                           const cx<T> *__restrict__ Scol, const T *__restrict__ lu,
                                           int side, T rf, int lane)
{
    const int ny = L.ny, ld = L.ld;
    const LineLU<T, R> f = load_line_lu_tbl<T, R>(lu, L.lus, line_j0(lane, R));
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
    X(16, 1) X(16, 2) X(16, 3) X(16, 4) X(16, 5) X(16, 6) X(16, 7) X(16, 9) X(16, 13) \
    X(16, 17) \
    X(32, 1) X(32, 2) X(32, 3) X(32, 4) X(32, 5) X(32, 6) X(32, 7) X(32, 9) \
    X(64, 1) X(64, 2) X(64, 3) X(64, 4) X(64, 5) X(64, 6) X(64, 7) X(64, 9) X(64, 13) X(64, 17)

inline unsigned nblk(size_t n, unsigned b = 256) { return static_cast<unsigned>((n + b - 1) / b); }

// Raise a kernel's dynamic-LDS limit to the CU's 160 KB.
inline hipError_t allow_full_lds(const void *f)
{
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

}  // namespace tbl
