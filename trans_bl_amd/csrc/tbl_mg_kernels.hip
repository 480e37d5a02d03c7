// Multigrid kernels of the v-Poisson solve (gfx950).
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
            const cx<T> xx = ((y[r] - w1 * a) - w2 * b) / pj;
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

template <typename T, int G, int R>
__global__ void __launch_bounds__(256) smooth_phase_kernel(LevelArgs<T> L, cx<T> *V,
                                                           const cx<T> *S, const T *lu,
                                                           int phase, T rf, const int *active)
{
    const int nlines = phase_lines(phase, L.nx);
    const int gid = (blockIdx.x * 256 + threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    const int k = gid / nlines;
    if (k >= L.K) return;
    if (active && !active[k]) return;
    const int c = gid - k * nlines;
    const int i = phase_column(phase, c, L.nx);
    const int side = (phase == 0) ? (c == 0 ? -1 : 1) : 0;
    const size_t base = static_cast<size_t>(k) * L.plane + static_cast<size_t>(i) * L.ld;
    const T *lut = lu + (static_cast<size_t>(k) * 2 + (side != 0)) * kLuFields * L.ld;
    relax_line<T, G, R>(L, V + base, S + base, lut, side, rf, lane);
}

constexpr int kStagedLines = 4;   // lines per workgroup of the staged smoother

// Same relaxation with coalesced global traffic: the group first reads its
// line lane-strided (row t, t+G, ...: contiguous across the wavefront),
// forming the off-diagonal x terms, v and S in LDS; the scan then works on
// lane-contiguous segments from LDS; the relaxed v goes back through LDS and
// is stored lane-strided.  Arithmetic identical to smooth_phase_kernel.
template <typename T, int G, int R>
__global__ void __launch_bounds__(256) smooth_staged_kernel(LevelArgs<T> L, cx<T> *V,
                                                            const cx<T> *S, const T *lu,
                                                            int phase, T rf, const int *active)
{
    extern __shared__ unsigned char smem_raw[];
    const int LDW = L.ld;                                // rows held per line
    cx<T> *base_s = reinterpret_cast<cx<T> *>(smem_raw);
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    cx<T> *Xs = base_s + static_cast<size_t>(grp) * 3 * LDW;
    cx<T> *Vs = Xs + LDW, *Ss = Vs + LDW;
    const int nlines = phase_lines(phase, L.nx);
    const int gid = blockIdx.x * kStagedLines + grp;
    const int k = gid / nlines;
    const bool live = k < L.K && !(active && !active[k]);
    const int ny = L.ny, ld = L.ld;
    int side = 0;
    size_t base = 0;
    if (live) {
        const int c = gid - k * nlines;
        const int i = phase_column(phase, c, L.nx);
        side = (phase == 0) ? (c == 0 ? -1 : 1) : 0;
        base = static_cast<size_t>(k) * L.plane + static_cast<size_t>(i) * ld;
        const cx<T> *col = V + base;
        for (int j = lane; j < ny; j += G) {
            Xs[j] = xoff<T>(col, ld, j, side, side == 0 && j >= ny - 2);
            Vs[j] = col[j];
            Ss[j] = S[base + j];
        }
    }
    __syncthreads();
    if (live) {
        const T *lut = lu + (static_cast<size_t>(k) * 2 + (side != 0)) * kLuFields * ld;
        const LineLU<T, R> f = load_line_lu<T, R>(lut, ld, lane * R, ny);
        relax_core<T, G, R>(L, f, rf, lane, [&](int j) { return Xs[j]; },
                            [&](int j) { return Vs[j]; }, [&](int j, int) { return Ss[j]; },
                            [&](int j, cx<T> v) { Vs[j] = v; });
    }
    __syncthreads();
    if (live) {
        cx<T> *col = V + base;
        for (int j = lane; j < ny; j += G)
            if (j >= 1) col[j] = Vs[j];
    }
}

// Fused sweep over column strips: one workgroup owns a strip [a, b) of the
// interior columns of one mode and runs a whole sweep (border phase and the
// three colour phases) in LDS.  It loads the columns [a-6, b+6) (widened by
// one to the boundary column when that range stops one short): each colour
// phase needs its neighbours two columns either side as of the previous
// phase, so phase 1 relaxes its lines on [a-4, b+4), phase 2 on [a-2, b+2)
// and phase 3 on [a, b) -- the halo lines are recomputed, not exchanged --
// and the result on [a, b) equals the phase-by-phase sweep.  Halo columns
// come from a snapshot taken before the sweep (strip_halo_kernel): the
// neighbouring workgroups overwrite those columns of v during the sweep.
__host__ __device__ inline int strip_stride(int ny) { return ny | 1; }

template <typename T>
__global__ void __launch_bounds__(256) strip_halo_kernel(LevelArgs<T> L, const cx<T> *V,
                                                         cx<T> *H, int width, int nb,
                                                         const int *active)
{
    const int k = blockIdx.z, s = blockIdx.y;
    if (active && !active[k]) return;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= kStripHalo * 2 * L.ny) return;
    const int c = e / L.ny, j = e % L.ny;
    const int i = 1 + (s + 1) * width - kStripHalo + c;
    if (i < 1 || i > L.nx - 2) return;
    H[((static_cast<size_t>(k) * nb + s) * 2 * kStripHalo + c) * L.ld + j] =
        V[static_cast<size_t>(k) * L.plane + static_cast<size_t>(i) * L.ld + j];
}

template <typename T, int G, int R>
__global__ void __launch_bounds__(512) smooth_strip_kernel(LevelArgs<T> L, cx<T> *V,
                                                            const cx<T> *S, const T *lu,
                                                            const cx<T> *H, T rf,
                                                            const int *active, int width,
                                                            int nstrips)
{
    extern __shared__ unsigned char smem_raw[];
    cx<T> *sv = reinterpret_cast<cx<T> *>(smem_raw);
    const int k = blockIdx.x / nstrips, sc = blockIdx.x % nstrips;
    if (active && !active[k]) return;
    const int nx = L.nx, ny = L.ny, ld = L.ld, ls = strip_stride(ny);
    const int a = 1 + sc * width, b = min(a + width, nx - 1);
    int lo = max(a - kStripHalo, 0), hi = min(b + kStripHalo, nx);
    // a loaded border column (Fortran 2 / nx-1) is relaxed in the border
    // phase, so its whole one-sided stencil must be loaded too
    if (lo == 1) lo = 0;
    if (hi == nx - 1) hi = nx;
    const int nb = nstrips - 1;
    cx<T> *gv = V + static_cast<size_t>(k) * L.plane;
    const cx<T> *gs = S + static_cast<size_t>(k) * L.plane;
    for (int e = threadIdx.x; e < (hi - lo) * ny; e += blockDim.x) {
        const int i = lo + e / ny, j = e % ny;
        const cx<T> *src;
        if ((i >= a && i < b) || i == 0 || i == nx - 1)
            src = gv + static_cast<size_t>(i) * ld;
        else if (i < a)
            src = H + ((static_cast<size_t>(k) * nb + (sc - 1)) * 2 * kStripHalo
                       + (i - (a - kStripHalo))) * ld;
        else
            src = H + ((static_cast<size_t>(k) * nb + sc) * 2 * kStripHalo
                       + (i - (b - kStripHalo))) * ld;
        sv[(i - lo) * ls + j] = src[j];
    }
    __syncthreads();
    const int grp = threadIdx.x / G, lane = threadIdx.x % G, ngrp = blockDim.x / G;
    const int j0 = lane * R;
    // lines of each phase: first column, count, column step
    int i0s[4], cnts[4], steps[4];
#pragma unroll
    for (int phase = 0; phase < 4; ++phase) {
        if (phase == 0) {
            // border lines: column 2 / nx-1 (Fortran) when their stencils are loaded
            i0s[0] = (lo == 0) ? 1 : nx - 2;
            cnts[0] = (lo == 0) + (hi == nx);
            steps[0] = nx - 3;
        } else {
            const int m = 2 * (3 - phase);
            const int from = max(a - m, 2), to = min(b + m, nx - 2);
            const int cls = (phase + 1) % 3;
            i0s[phase] = from + ((cls - from) % 3 + 3) % 3;
            cnts[phase] = i0s[phase] < to ? (to - 1 - i0s[phase]) / 3 + 1 : 0;
            steps[phase] = 3;
        }
    }
    // this group's lines in sweep order: (phase, c) with c = grp, grp+ngrp, ...
    auto advance = [&](int &ph, int &c) {
        c += ngrp;
        while (ph < 4 && c >= cnts[ph]) {
            ++ph;
            c = grp;
        }
    };
    auto load_s = [&](int ph, int c, cx<T> *out) {
        const int i = i0s[ph] + c * steps[ph];
        const cx<T> *scol = gs + static_cast<size_t>(i) * ld;
#pragma unroll
        for (int r = 0; r < R; ++r)
            out[r] = (j0 + r < ny) ? scol[j0 + r] : cx<T>{T(0), T(0)};
    };
    const LineLU<T, R> fc = load_line_lu<T, R>(
        lu + static_cast<size_t>(k) * 2 * kLuFields * ld, ld, j0, ny);
    int ph = 0, c = grp;
    if (c >= cnts[0]) {
        c -= ngrp;
        advance(ph, c);
    }
    cx<T> Scur[R] = {}, Snext[R] = {};
    if (ph < 4) load_s(ph, c, Scur);
    for (int phase = 0; phase < 4; ++phase) {
        while (ph == phase) {
            // prefetch the source of this group's next line (read-only)
            int phn = ph, cn = c;
            advance(phn, cn);
            if (phn < 4) load_s(phn, cn, Snext);
            const int i = i0s[ph] + c * steps[ph];
            const int side = (ph == 0) ? (i == 1 ? -1 : 1) : 0;
            cx<T> *col = sv + static_cast<size_t>(i - lo) * ls;
            auto run = [&](const LineLU<T, R> &f) {
                relax_core<T, G, R>(
                    L, f, rf, lane,
                    [&](int j) { return xoff<T>(col, ls, j, side, side == 0 && j >= ny - 2); },
                    [&](int j) { return col[j]; }, [&](int, int r) { return Scur[r]; },
                    [&](int j, cx<T> v) { col[j] = v; });
            };
            if (side == 0) {
                run(fc);
            } else {
                run(load_line_lu<T, R>(
                    lu + (static_cast<size_t>(k) * 2 + 1) * kLuFields * ld, ld, j0, ny));
            }
#pragma unroll
            for (int r = 0; r < R; ++r) Scur[r] = Snext[r];
            ph = phn;
            c = cn;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < (b - a) * ny; e += blockDim.x) {
        const int i = a + e / ny, j = e % ny;
        if (j >= 1) gv[static_cast<size_t>(i) * ld + j] = sv[(i - lo) * ls + j];
    }
}

// Coarsest level: one workgroup per mode keeps the whole level in LDS and
// runs every sweep (20*itera, poisson_par.f:45) without leaving the CU.  LDS
// holds v with a compact odd column stride and the mode's line-LU factors;
// the source is read from global memory.  (Scan form, for levels whose lines
// are too long for the one-thread-per-line kernel below.)
__host__ __device__ inline int coarse_stride(int ny) { return ny | 1; }

template <typename T>
__host__ __device__ inline size_t coarse_lds_bytes(int nx, int ny, bool with_s)
{
    const size_t col = static_cast<size_t>(coarse_stride(ny)) * nx * sizeof(cx<T>);
    return col * (with_s ? 2 : 1) + static_cast<size_t>(2 * kLuFields) * ny * sizeof(T);
}

template <typename T, int G, int R>
__global__ void __launch_bounds__(512) smooth_coarsest_kernel(LevelArgs<T> L, cx<T> *V,
                                                               const cx<T> *S, const T *lu,
                                                               int sweeps, T rf,
                                                               const int *active)
{
    extern __shared__ unsigned char smem_raw[];
    const int k = blockIdx.x;
    if (active && !active[k]) return;
    const int nx = L.nx, ny = L.ny, ld = L.ld, ls = coarse_stride(ny);
    cx<T> *sv = reinterpret_cast<cx<T> *>(smem_raw);
    T *slu = reinterpret_cast<T *>(sv + static_cast<size_t>(ls) * nx);
    cx<T> *gv = V + static_cast<size_t>(k) * L.plane;
    const cx<T> *gs = S + static_cast<size_t>(k) * L.plane;
    const T *glu = lu + (static_cast<size_t>(k) * 2) * kLuFields * ld;
    for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
        const int i = e / ny, j = e % ny;
        sv[i * ls + j] = gv[static_cast<size_t>(i) * ld + j];
    }
    for (int e = threadIdx.x; e < 2 * kLuFields * ny; e += blockDim.x)
        slu[e] = glu[static_cast<size_t>(e / ny) * ld + e % ny];
    __syncthreads();
    // the line solve indexes the factors with stride L.ld: give it ny here
    LevelArgs<T> Ls = L;
    Ls.ld = ny;
    const int groups = blockDim.x / G;
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    for (int it = 0; it < sweeps; ++it) {
        for (int phase = 0; phase < 4; ++phase) {
            const int nl = phase_lines(phase, nx);
            for (int c = grp; c < nl; c += groups) {
                const int i = phase_column(phase, c, nx);
                const int side = (phase == 0) ? (c == 0 ? -1 : 1) : 0;
                cx<T> *col = sv + static_cast<size_t>(i) * ls;
                const cx<T> *scol = gs + static_cast<size_t>(i) * ld;
                const LineLU<T, R> f =
                    load_line_lu<T, R>(slu + (side ? kLuFields * ny : 0), ny, lane * R, ny);
                relax_core<T, G, R>(
                    Ls, f, rf, lane,
                    [&](int j) { return xoff<T>(col, ls, j, side, side == 0 && j >= ny - 2); },
                    [&](int j) { return col[j]; }, [&](int j, int) { return scol[j]; },
                    [&](int j, cx<T> v) { col[j] = v; });
            }
            __syncthreads();
        }
    }
    for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
        const int i = e / ny, j = e % ny;
        gv[static_cast<size_t>(i) * ld + j] = sv[i * ls + j];
    }
}

// Coarsest level, one thread per line: each line of a phase is relaxed by one
// thread with the sequential banded substitution of banbk5_poi
// (penta_par.f:367-404).  v (and the source when it fits) stay in LDS; the
// forward-substitution values stay in registers (NY >= ny rows, unrolled);
// the factors are read with wave-uniform global loads, which cannot alias
// the LDS stores, so the row loads are scheduled off the serial chain.  The
// back substitution multiplies by the precomputed -u1/p, -u2/p, 1/p instead
// of dividing by the pivot (last-bit rounding differs from the division).
template <typename T>
__host__ __device__ inline size_t coarse_reg_lds_bytes(int nx, int ny, bool with_s)
{
    return static_cast<size_t>(coarse_stride(ny)) * nx * sizeof(cx<T>) * (with_s ? 2 : 1);
}

// SL: the source is staged in LDS (else read from global memory); a
// compile-time choice, so every access has a known address space.
template <typename T, int NY, bool SL>
__global__ void __launch_bounds__(256) smooth_coarsest_reg_kernel(LevelArgs<T> L, cx<T> *V,
                                                                   const cx<T> *S, const T *lu,
                                                                   int sweeps, T rf,
                                                                   const int *active)
{
    extern __shared__ unsigned char smem_raw[];
    const int k = blockIdx.x;
    if (active && !active[k]) return;
    const int nx = L.nx, ny = L.ny, ld = L.ld, ls = coarse_stride(ny);
    cx<T> *sv = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *ss = sv + static_cast<size_t>(ls) * nx;
    cx<T> *gv = V + static_cast<size_t>(k) * L.plane;
    const cx<T> *gs = S + static_cast<size_t>(k) * L.plane;
    for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
        const int i = e / ny, j = e % ny;
        sv[i * ls + j] = gv[static_cast<size_t>(i) * ld + j];
        if (SL) ss[i * ls + j] = gs[static_cast<size_t>(i) * ld + j];
    }
    __syncthreads();
    const T *lu_centre = lu + (static_cast<size_t>(k) * 2) * kLuFields * ld;
    const T *lu_border = lu_centre + kLuFields * ld;
    const T qx = L.qx;
    const int c = threadIdx.x;
    for (int it = 0; it < sweeps; ++it) {
        for (int phase = 0; phase < 4; ++phase) {
            if (c < phase_lines(phase, nx)) {
                const int i = phase_column(phase, c, nx);
                const int side = (phase == 0) ? (c == 0 ? -1 : 1) : 0;
                const T *f = phase == 0 ? lu_border : lu_centre;
                const T *l1 = f, *l2 = f + ld, *m1 = f + 5 * ld, *m2 = f + 6 * ld, *rp = f + 7 * ld;
                cx<T> *col = sv + static_cast<size_t>(i) * ls;
                const cx<T> *scol = SL ? ss + static_cast<size_t>(i) * ls
                                       : gs + static_cast<size_t>(i) * ld;
                auto X = [&](int j) { return xoff<T>(col, ls, j, side, side == 0 && j >= ny - 2); };
                cx<T> Y[NY];
                cx<T> xm = {T(0), T(0)}, x0 = X(0), xp = X(1);
                cx<T> y1 = {T(0), T(0)}, y2 = {T(0), T(0)};
#pragma unroll
                for (int j = 0; j < NY; ++j) {
                    if (j < ny) {
                        cx<T> rhs;
                        if (j == 0) rhs = col[0];
                        else if (j == 1) rhs = scol[j] - qx * (L.sp[2] * xp + L.sp[1] * x0);
                        else if (j <= ny - 3)
                            rhs = scol[j] - qx * (L.cp[0] * xm + L.cp[1] * x0 + L.cp[2] * xp);
                        else rhs = scol[j] - qx * x0;
                        const cx<T> yy = (rhs - l2[j] * y2) - l1[j] * y1;
                        Y[j] = yy;
                        y2 = y1; y1 = yy;
                        xm = x0; x0 = xp;
                        xp = (j + 2 < ny) ? X(j + 2) : cx<T>{T(0), T(0)};
                    }
                }
                cx<T> a = {T(0), T(0)}, b = {T(0), T(0)};
#pragma unroll
                for (int j = NY - 1; j >= 0; --j) {
                    if (j < ny) {
                        const cx<T> xx = (rp[j] * Y[j] + m2[j] * b) + m1[j] * a;
                        if (j >= 1) col[j] = col[j] + rf * (xx - col[j]);
                        b = a; a = xx;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
        const int i = e / ny, j = e % ny;
        gv[static_cast<size_t>(i) * ld + j] = sv[i * ls + j];
    }
}

#define TBL_COARSE_NY(X) X(16) X(32) X(48)

// ---------------------------------------------------------------------------
// defect (poisson_par.f:543-793): out = S + si * L v on i=2..nx-1, j=2..ny.
// With resmax != NULL the max |out| of each mode is accumulated (the
// verifica_poi / get_max_value test, poisson_par.f:67-108, 945-963).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v)
{
    // non-negative doubles order like their bit patterns
    atomicMax(reinterpret_cast<unsigned long long *>(addr), __double_as_longlong(v));
}

// Defect by column marching: one wavefront owns kDefRows consecutive output
// rows of one strip of columns of one mode.  Its 64 lanes hold rows
// j0-2 .. j0+61 (two halo rows each side); it walks the strip in x keeping
// the five columns i-2..i+2 of its rows in registers (one new column load per
// step) and takes the y-neighbours from lane shuffles, so every value of v is
// loaded about once.  Row forms: Fortran row 2 (sp_poi_coef), interior rows
// (cp_poi_coef), row ny-1 (pp_poi_coef) and row ny (lp_poi_coef, with the
// top boundary term); x part: the border-column one-sided stencils on
// columns 2 and nx-1, the centred fourth-order stencil elsewhere (grouped
// on the two top rows).
constexpr int kDefRows = 60;

__host__ __device__ inline int defect_strip(int nx) { return nx >= 512 ? 32 : 16; }

template <typename T, typename TO>
__global__ void __launch_bounds__(256) defect_march_kernel(LevelArgs<T> L,
                                                           const cx<T> *__restrict__ V,
                                                           const cx<T> *S, cx<TO> *out, T si,
                                                           const int *active, double *resmax,
                                                           int strip, int nstrips, int jchunks)
{
    const int lane = threadIdx.x & 63;
    const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.y;
    if (item >= jchunks * nstrips) return;          // whole wavefront
    if (active && !active[k]) return;
    const int nx = L.nx, ny = L.ny, ld = L.ld;
    const int jc = item % jchunks, sc = item / jchunks;
    const int j = 1 + jc * kDefRows - 2 + lane;
    const bool jin = j >= 0 && j < ny;
    const bool own = lane >= 2 && lane < 2 + kDefRows && j >= 1 && j < ny;
    const int ia = 1 + sc * strip, ib = min(ia + strip, nx - 1);
    const cx<T> *vk = V + static_cast<size_t>(k) * L.plane;
    auto vload = [&](int i) -> cx<T> {
        return (jin && i >= 0 && i < nx) ? vk[static_cast<size_t>(i) * ld + j] : cx<T>{T(0), T(0)};
    };
    const T kk = static_cast<T>(k);
    const T k2b2 = -kk * kk * L.beta * L.beta;
    const T alpha2 = sqrt((L.alpha * L.alpha - k2b2) / L.fac_y);
    const T scx = T(12) * L.dx2 * L.fac_y;
    T q = T(0);
    if (own) q = (j == 1) ? L.qdy2[0] : (j == ny - 2) ? L.qdy2[j - 1] : L.qdy2[j - 2];
    cx<T> w0 = vload(ia - 2), w1 = vload(ia - 1), w2 = vload(ia), w3 = vload(ia + 1),
          w4 = vload(ia + 2);
    double amax = 0.0;
    for (int i = ia; i < ib; ++i) {
        const cx<T> nxt = vload(i + 3);
        cx<T> xs, xo;   // x part at row j: shared (ungrouped) and own
        if (i == 1) {
            xs = T(10) * w1 - T(15) * w2 - T(4) * w3 + T(14) * w4 - T(6) * vload(4) + vload(5);
            xo = xs;
        } else if (i == nx - 2) {
            xs = T(10) * w3 - T(15) * w2 - T(4) * w1 + T(14) * w0 - T(6) * vload(nx - 5)
                 + vload(nx - 6);
            xo = xs;
        } else {
            xs = T(-1) * w0 + T(16) * w1 - T(30) * w2 + T(16) * w3 - T(1) * w4;
            xo = (j >= ny - 2) ? T(-1) * (w0 + w4) + T(16) * (w1 + w3) - T(30) * w2 : xs;
        }
        const cx<T> xm = shfl_up(xs, 1, 64), xp = shfl_down(xs, 1, 64);
        const cx<T> vm1 = shfl_up(w2, 1, 64), vm2 = shfl_up(w2, 2, 64);
        const cx<T> vp1 = shfl_down(w2, 1, 64), vp2 = shfl_down(w2, 2, 64);
        if (own) {
            cx<T> lv;
            if (j == 1) {
                const cx<T> b = L.sp[2] * xp + L.sp[1] * xs;
                const cx<T> c = L.sp[3] * vm1 + L.sp[4] * w2 + L.sp[5] * vp1 + L.sp[6] * vp2;
                const cx<T> d = k2b2 * (L.sp[1] * w2 + L.sp[2] * vp1);
                lv = b / scx + c * q + d / L.fac_y;
            } else if (j <= ny - 3) {
                const cx<T> b = L.cp[0] * xm + L.cp[1] * xs + L.cp[2] * xp;
                const cx<T> c = L.cp[3] * vm2 + L.cp[4] * vm1 + L.cp[5] * w2 + L.cp[6] * vp1
                                + L.cp[7] * vp2;
                const cx<T> d = k2b2 * (L.cp[0] * vm1 + L.cp[1] * w2 + L.cp[2] * vp1);
                lv = b / scx + c * q + d / L.fac_y;
            } else if (j == ny - 2) {
                const cx<T> c = L.pp[3] * vp1 + L.pp[2] * w2 + L.pp[1] * vm1;
                const cx<T> d = k2b2 * w2;
                lv = xo / scx + c * q + d / L.fac_y;
            } else {
                const cx<T> c = L.lp[2] * w2 + L.lp[3] * vm1 + L.lp[4] * vm2;
                const cx<T> d = k2b2 * w2;
                lv = xo / scx + c * q + d / L.fac_y - L.lp[1] * alpha2 * w2 * L.sqq_top;
            }
            const size_t o = static_cast<size_t>(k) * L.plane + static_cast<size_t>(i) * ld + j;
            const cx<T> r = S[o] + si * lv;
            out[o] = cx<TO>{static_cast<TO>(r.re), static_cast<TO>(r.im)};
            if (resmax) amax = fmax(amax, hypot(static_cast<double>(r.re), static_cast<double>(r.im)));
        }
        w0 = w1; w1 = w2; w2 = w3; w3 = w4; w4 = nxt;
    }
    if (resmax) {
        for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_down(amax, o, 64));
        if (lane == 0) atomic_max_nonneg<T>(resmax + k, amax);
    }
}

// ---------------------------------------------------------------------------
// restrict (poisson_par.f:796-874): inject v into the coarse level (and the
// FAS copy w), full-weight the fine residual into the coarse source.
// fas == false: correction scheme (coarse v = 0, no copy).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) restrict_kernel(Level F, Level Cl, int K,
                                                       const cx<T> *Vf, const cx<T> *Rf,
                                                       cx<T> *Vc, cx<T> *Wc, cx<T> *Sc,
                                                       int fas, const int *active)
{
    const size_t idx = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    const size_t per = Cl.plane;
    const int k = static_cast<int>(idx / per);
    if (k >= K) return;
    if (active && !active[k]) return;
    const size_t rem = idx - static_cast<size_t>(k) * per;
    const int i = static_cast<int>(rem / Cl.ld), j = static_cast<int>(rem % Cl.ld);
    if (j >= Cl.ny) return;
    const int fi = 2 * i, fj = 2 * j;
    const cx<T> *vf = Vf + static_cast<size_t>(k) * F.plane;
    const cx<T> *rf = Rf + static_cast<size_t>(k) * F.plane;
    const size_t co = static_cast<size_t>(k) * Cl.plane + static_cast<size_t>(i) * Cl.ld + j;
    if (fas) {
        const cx<T> v = vf[static_cast<size_t>(fi) * F.ld + fj];
        Vc[co] = v;
        Wc[co] = v;
    } else {
        Vc[co] = cx<T>{T(0), T(0)};
    }
    if (i >= 1 && i <= Cl.nx - 2) {
        auto R = [&](int ii, int jj) { return rf[static_cast<size_t>(ii) * F.ld + jj]; };
        if (j >= 1 && j <= Cl.ny - 2) {
            Sc[co] = T(6.25e-2) * (T(2) * (R(fi - 1, fj) + R(fi + 1, fj) + R(fi, fj - 1) + R(fi, fj + 1))
                                   + T(4) * R(fi, fj) + R(fi + 1, fj + 1) + R(fi + 1, fj - 1)
                                   + R(fi - 1, fj + 1) + R(fi - 1, fj - 1));
        } else if (j == Cl.ny - 1) {
            Sc[co] = T(0.25) * (R(fi - 1, fj) + T(2) * R(fi, fj) + R(fi + 1, fj));
        }
    }
}

// correction + intpol (poisson_par.f:877-942).  The coarse correction is
// v_c - w_c on interior columns and v_c itself on the two boundary columns
// (the reference interpolates those unchanged values as well).  For the
// correction scheme (fas == 0) it is v_c.
template <typename T>
__global__ void __launch_bounds__(256) prolong_kernel(Level F, Level Cl, int K, cx<T> *Vf,
                                                      const cx<T> *Vc, const cx<T> *Wc,
                                                      int fas, const int *active)
{
    const size_t idx = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    const int k = static_cast<int>(idx / F.plane);
    if (k >= K) return;
    if (active && !active[k]) return;
    const size_t rem = idx - static_cast<size_t>(k) * F.plane;
    const int i = static_cast<int>(rem / F.ld), j = static_cast<int>(rem % F.ld);
    if (j >= F.ny || i < 1 || i > F.nx - 2) return;
    const cx<T> *vc = Vc + static_cast<size_t>(k) * Cl.plane;
    const cx<T> *wc = Wc + static_cast<size_t>(k) * Cl.plane;
    auto E = [&](int ci, int cj) {
        const size_t o = static_cast<size_t>(ci) * Cl.ld + cj;
        if (fas && ci >= 1 && ci <= Cl.nx - 2) return vc[o] - wc[o];
        return vc[o];
    };
    auto row = [&](int cj) {   // value on an even (0-based) fine row
        if ((i & 1) == 0) return E(i >> 1, cj);
        return T(0.5) * (E(i >> 1, cj) + E((i >> 1) + 1, cj));
    };
    cx<T> t;
    if ((j & 1) == 0) t = row(j >> 1);
    else t = T(0.5) * (row((j >> 1) + 1) + row(j >> 1));
    const size_t o = static_cast<size_t>(k) * F.plane + static_cast<size_t>(i) * F.ld + j;
    Vf[o] = Vf[o] + t;
}

// ---------------------------------------------------------------------------
// user layout (x fastest) <-> internal layout (y fastest) through LDS tiles
// ---------------------------------------------------------------------------
// g = kb*wx - dwzdx (wx may be NULL) into G (internal), uy into V (internal)
__global__ void __launch_bounds__(256) gather_kernel(int nx, int ny, size_t ld, size_t plane,
                                                     const cx<double> *wx, const cx<double> *dz,
                                                     const cx<double> *uy, double beta,
                                                     cx<double> *Gout, cx<double> *Vout,
                                                     int direct_rhs)
{
    __shared__ cx<double> tg[32][33], tv[32][33];
    const int k = blockIdx.z;
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;   // 32 x 8
    const double kb = -static_cast<double>(k) * beta;           // kb = (0, kb)
    const size_t ubase = static_cast<size_t>(k) * nx * ny;
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + tx, j = j0 + r;
        cx<double> g = {0, 0}, v = {0, 0};
        if (i < nx && j < ny) {
            const size_t u = ubase + static_cast<size_t>(j) * nx + i;
            cx<double> d = dz ? dz[u] : cx<double>{0, 0};
            if (direct_rhs) {
                g = d;                      // d holds f directly (solve with S = B f)
            } else {
                cx<double> kw = {0, 0};
                if (wx) { const cx<double> w = wx[u]; kw = cmul(cx<double>{0.0, kb}, w); }
                g = kw - d;
            }
            v = uy ? uy[u] : cx<double>{0, 0};
        }
        tg[r][tx] = g;
        tv[r][tx] = v;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        if (i < nx && j < ny) {
            const size_t o = static_cast<size_t>(k) * plane + static_cast<size_t>(i) * ld + j;
            Gout[o] = tg[tx][r];
            Vout[o] = tv[tx][r];
        }
    }
}

// function_source weights (poisson_par.f:126-144), internal layout
__global__ void __launch_bounds__(256) source_kernel(LevelArgs<double> L, const cx<double> *Gin,
                                                     cx<double> *S)
{
    const size_t idx = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    const int k = static_cast<int>(idx / L.plane);
    if (k >= L.K) return;
    const size_t rem = idx - static_cast<size_t>(k) * L.plane;
    const int j = static_cast<int>(rem % L.ld);
    if (j >= L.ny) return;
    const cx<double> *g = Gin + (idx - j);
    cx<double> s = {0, 0};
    if (j == 1) s = L.sp[1] * g[1] + L.sp[2] * g[2];
    else if (j >= 2 && j <= L.ny - 3) s = L.cp[0] * g[j - 1] + L.cp[1] * g[j] + L.cp[2] * g[j + 1];
    else if (j == L.ny - 2) s = g[j];
    S[idx] = s;
}

// internal V -> user uy on i=2..nx-1, j=2..ny (poisson_par.f:52-57)
__global__ void __launch_bounds__(256) scatter_kernel(int nx, int ny, size_t ld, size_t plane,
                                                      const cx<double> *V, cx<double> *uy)
{
    __shared__ cx<double> t[32][33];
    const int k = blockIdx.z;
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        if (i < nx && j < ny)
            t[r][tx] = V[static_cast<size_t>(k) * plane + static_cast<size_t>(i) * ld + j];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + tx, j = j0 + r;
        if (i >= 1 && i <= nx - 2 && j >= 1 && j < ny)
            uy[static_cast<size_t>(k) * nx * ny + static_cast<size_t>(j) * nx + i] = t[tx][r];
    }
}

// convergence bookkeeping after the level-1 check of cycle `cycle`
__global__ void decide_kernel(int K, const double *resmax, double tol, int cycle, int *active,
                              int *cycles, double *lastres, int *nactive)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    if (!active[k]) return;
    lastres[k] = resmax[k];
    if (resmax[k] < tol) {
        active[k] = 0;
        cycles[k] = cycle - 1;
    } else {
        cycles[k] = cycle;      // will complete this cycle
        atomicAdd(nactive, 1);
    }
}

__global__ void fill_kernel(double *p, size_t n, double v)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_int_kernel(int *p, int n, int v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// mixed precision helpers ----------------------------------------------------
// v64 += e32 on the interior (i=2..nx-1, j=2..ny)
__global__ void __launch_bounds__(256) update_kernel(Level F, int K, cx<double> *V,
                                                     const cx<float> *E, const int *active)
{
    const size_t idx = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    const int k = static_cast<int>(idx / F.plane);
    if (k >= K) return;
    if (active && !active[k]) return;
    const size_t rem = idx - static_cast<size_t>(k) * F.plane;
    const int i = static_cast<int>(rem / F.ld), j = static_cast<int>(rem % F.ld);
    if (i < 1 || i > F.nx - 2 || j < 1 || j >= F.ny) return;
    const cx<float> e = E[idx];
    V[idx] = cx<double>{V[idx].re + static_cast<double>(e.re), V[idx].im + static_cast<double>(e.im)};
}

template <typename TA, typename TB>
__global__ void __launch_bounds__(256) convert_kernel(size_t n, const cx<TA> *a, cx<TB> *b,
                                                      int zero_a_into_b)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    if (zero_a_into_b) { b[i] = cx<TB>{TB(0), TB(0)}; return; }
    b[i] = cx<TB>{static_cast<TB>(a[i].re), static_cast<TB>(a[i].im)};
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(size_t n, unsigned b = 256) { return static_cast<unsigned>((n + b - 1) / b); }

template <typename T, int G, int R>
static void launch_phase(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int phase,
                         T rf, const int *active, hipStream_t st)
{
    const size_t lines = static_cast<size_t>(phase_lines(phase, L.nx)) * L.K;
    if (!lines) return;
    const size_t threads = lines * G;
    hipLaunchKernelGGL((smooth_phase_kernel<T, G, R>), dim3(nblk(threads)), dim3(256), 0, st, L, V, S,
                       lu, phase, rf, active);
}

template <typename T, int G, int R>
static void launch_staged(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int phase,
                          T rf, const int *active, hipStream_t st)
{
    const size_t lines = static_cast<size_t>(phase_lines(phase, L.nx)) * L.K;
    if (!lines) return;
    const size_t shm = static_cast<size_t>(kStagedLines) * 3 * L.ld * sizeof(cx<T>);
    hipLaunchKernelGGL((smooth_staged_kernel<T, G, R>),
                       dim3(static_cast<unsigned>((lines + kStagedLines - 1) / kStagedLines)),
                       dim3(kStagedLines * G), shm, st, L, V, S, lu, phase, rf, active);
}

template <typename T, int G, int R>
static void launch_strip(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, cx<T> *H,
                         const StripGeom &sg, T rf, const int *active, hipStream_t st)
{
    const int nb = sg.nstrips - 1;
    if (nb > 0)
        hipLaunchKernelGGL((strip_halo_kernel<T>),
                           dim3((2 * kStripHalo * L.ny + 255) / 256, nb, L.K), dim3(256), 0, st,
                           L, V, H, sg.width, nb, active);
    const int lines = (sg.width + 4 * 2 + 2) / 3 + 1;
    const int groups = std::min(512 / G, std::max(lines, 1));
    const int threads = (groups * G + 63) / 64 * 64;
    const size_t shm = static_cast<size_t>(std::min(sg.width + 2 * kStripHalo + 2, L.nx))
                       * strip_stride(L.ny) * sizeof(cx<T>);
    hipLaunchKernelGGL((smooth_strip_kernel<T, G, R>), dim3(L.K * sg.nstrips), dim3(threads), shm,
                       st, L, V, S, lu, H, rf, active, sg.width, sg.nstrips);
}

template <typename T, int G, int R>
static void launch_coarsest(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu,
                            int sweeps, T rf, const int *active, hipStream_t st, int threads)
{
    const size_t shm = coarse_lds_bytes<T>(L.nx, L.ny, false);
    hipLaunchKernelGGL((smooth_coarsest_kernel<T, G, R>), dim3(L.K), dim3(threads), shm, st, L, V,
                       S, lu, sweeps, rf, active);
}

// (G, R) table shared by both smoother forms
#define TBL_LINE_SHAPES(X) \
    X(16, 1) X(16, 2) X(16, 3) X(16, 4) X(16, 5) X(16, 7) X(16, 9) X(16, 13) X(16, 17) \
    X(32, 1) X(32, 2) X(32, 3) X(32, 4) X(32, 5) X(32, 7) X(32, 9) \
    X(64, 1) X(64, 2) X(64, 3) X(64, 4) X(64, 5) X(64, 7) X(64, 9) X(64, 13) X(64, 17)

static const int kRs[] = {1, 2, 3, 4, 5, 7, 9, 13, 17};

// Raise the dynamic-LDS limit of every LDS-resident smoother instantiation
// once, before any stream capture can happen.
int set_kernel_attributes()
{
    hipError_t e = hipSuccess;
#define X(GG, RR)                                                                             \
    for (const void *f : {reinterpret_cast<const void *>(&smooth_staged_kernel<double, GG, RR>), \
                          reinterpret_cast<const void *>(&smooth_staged_kernel<float, GG, RR>),  \
                          reinterpret_cast<const void *>(&smooth_coarsest_kernel<double, GG, RR>), \
                          reinterpret_cast<const void *>(&smooth_coarsest_kernel<float, GG, RR>),  \
                          reinterpret_cast<const void *>(&smooth_strip_kernel<double, GG, RR>),    \
                          reinterpret_cast<const void *>(&smooth_strip_kernel<float, GG, RR>)}) { \
        hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,     \
                                           160 * 1024);                                      \
        if (r != hipSuccess) e = r;                                                           \
    }
    TBL_LINE_SHAPES(X)
#undef X
#define X(NY)                                                                                 \
    for (const void *f :                                                                      \
         {reinterpret_cast<const void *>(&smooth_coarsest_reg_kernel<double, NY, true>),       \
          reinterpret_cast<const void *>(&smooth_coarsest_reg_kernel<double, NY, false>),      \
          reinterpret_cast<const void *>(&smooth_coarsest_reg_kernel<float, NY, true>),        \
          reinterpret_cast<const void *>(&smooth_coarsest_reg_kernel<float, NY, false>)}) {    \
        hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,     \
                                           160 * 1024);                                      \
        if (r != hipSuccess) e = r;                                                           \
    }
    TBL_COARSE_NY(X)
#undef X
    return e == hipSuccess ? TBL_OK : TBL_ERR_HIP;
}

StripGeom strip_geom(int nx, int ny, int K, size_t esize)
{
    StripGeom g{0, 0};
    const size_t col = static_cast<size_t>(strip_stride(ny)) * 2 * esize;
    const int maxcols = static_cast<int>(160 * 1024 / col);
    const int wmax = maxcols - 2 * kStripHalo - 2;
    const int n_int = nx - 2;
    if (wmax < 2 * kStripHalo) return g;
    int n = (n_int + wmax - 1) / wmax;
    // enough workgroups for the chip, strips at least 2*halo wide
    const int want = (512 + K - 1) / K;
    n = std::max(n, std::min(want, n_int / (2 * kStripHalo)));
    n = std::max(n, 1);
    g.width = (n_int + n - 1) / n;
    g.nstrips = (n_int + g.width - 1) / g.width;
    return g;
}

int line_rows_for(int G, int ny)
{
    const int need = (ny + G - 1) / G;
    for (int r : kRs) {
        if (r < need) continue;
        if (G == 32 && r > 9) return -1;
        return r;
    }
    return -1;
}

// G > 0: lanes per line with direct (lane-contiguous) global access;
// G < 0: |G| lanes per line with the LDS-staged, coalesced kernel
template <typename T>
int smooth_sweeps(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps, T rf,
                  const int *active, int G, cx<T> *halo, hipStream_t st)
{
    if (G > kStripShape) {
        // fused-sweep strips
        G -= kStripShape;
        const int R = line_rows_for(G, L.ny);
        const StripGeom sg = strip_geom(L.nx, L.ny, L.K, sizeof(T));
        if (R < 0 || sg.nstrips < 1 || (sg.nstrips > 1 && !halo)) return TBL_ERR_GRID;
        for (int it = 0; it < sweeps; ++it) {
            bool done = false;
#define X(GG, RR)                                                                       \
    if (!done && G == GG && R == RR) {                                                  \
        launch_strip<T, GG, RR>(L, V, S, lu, halo, sg, rf, active, st);                 \
        done = true;                                                                    \
    }
            TBL_LINE_SHAPES(X)
#undef X
            if (!done) return TBL_ERR_GRID;
        }
        return TBL_OK;
    }
    const bool staged = G < 0;
    if (staged) G = -G;
    const int R = line_rows_for(G, L.ny);
    if (R < 0) return TBL_ERR_GRID;
    if (staged && static_cast<size_t>(kStagedLines) * 3 * L.ld * sizeof(cx<T>) > 160 * 1024)
        return TBL_ERR_GRID;
    for (int it = 0; it < sweeps; ++it)
        for (int ph = 0; ph < 4; ++ph) {
            bool done = false;
#define X(GG, RR)                                                             \
    if (!done && G == GG && R == RR) {                                        \
        if (staged) launch_staged<T, GG, RR>(L, V, S, lu, ph, rf, active, st); \
        else launch_phase<T, GG, RR>(L, V, S, lu, ph, rf, active, st);        \
        done = true;                                                          \
    }
            TBL_LINE_SHAPES(X)
#undef X
            if (!done) return TBL_ERR_GRID;
        }
    return TBL_OK;
}

template <typename T>
int smooth_coarsest(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps,
                    T rf, const int *active, int G, hipStream_t st)
{
    if (L.ny <= 48 && coarse_reg_lds_bytes<T>(L.nx, L.ny, false) <= 160 * 1024 &&
        phase_lines(1, L.nx) <= 256) {
        const int s_in = coarse_reg_lds_bytes<T>(L.nx, L.ny, true) <= 160 * 1024;
        const int threads = (phase_lines(1, L.nx) + 63) / 64 * 64;
        const size_t shm = coarse_reg_lds_bytes<T>(L.nx, L.ny, s_in);
        bool done = false;
#define X(NY)                                                                                 \
    if (!done && L.ny <= NY) {                                                                \
        if (s_in)                                                                             \
            hipLaunchKernelGGL((smooth_coarsest_reg_kernel<T, NY, true>), dim3(L.K),          \
                               dim3(threads), shm, st, L, V, S, lu, sweeps, rf, active);      \
        else                                                                                  \
            hipLaunchKernelGGL((smooth_coarsest_reg_kernel<T, NY, false>), dim3(L.K),         \
                               dim3(threads), shm, st, L, V, S, lu, sweeps, rf, active);      \
        done = true;                                                                          \
    }
        TBL_COARSE_NY(X)
#undef X
        return TBL_OK;
    }
    if (coarse_lds_bytes<T>(L.nx, L.ny, false) > 160 * 1024)
        return smooth_sweeps<T>(L, V, S, lu, sweeps, rf, active, G > kStripShape ? G - kStripShape : G,
                                nullptr, st);
    if (G > kStripShape) G -= kStripShape;
    if (G < 0) G = -G;
    const int R = line_rows_for(G, L.ny);
    if (R < 0) return TBL_ERR_GRID;
    bool done = false;
#define X(GG, RR)                                                                 \
    if (!done && G == GG && R == RR) {                                            \
        launch_coarsest<T, GG, RR>(L, V, S, lu, sweeps, rf, active, st, 512);     \
        done = true;                                                              \
    }
    TBL_LINE_SHAPES(X)
#undef X
    return done ? TBL_OK : TBL_ERR_GRID;
}

template <typename T, typename TO>
void defect(const LevelArgs<T> &L, const cx<T> *V, const cx<T> *S, cx<TO> *out, T si,
            const int *active, double *resmax, hipStream_t st)
{
    const int strip = defect_strip(L.nx);
    const int nstrips = (L.nx - 2 + strip - 1) / strip;
    const int jchunks = (L.ny - 1 + kDefRows - 1) / kDefRows;
    const int blocks = (nstrips * jchunks + 3) / 4;
    hipLaunchKernelGGL((defect_march_kernel<T, TO>), dim3(blocks, L.K), dim3(256), 0, st, L, V, S,
                       out, si, active, resmax, strip, nstrips, jchunks);
}

template <typename T>
void restrict_level(const Level &F, const Level &Cl, int K, const cx<T> *Vf, const cx<T> *Rf,
                    cx<T> *Vc, cx<T> *Wc, cx<T> *Sc, int fas, const int *active, hipStream_t st)
{
    const size_t n = Cl.plane * K;
    hipLaunchKernelGGL((restrict_kernel<T>), dim3(nblk(n)), dim3(256), 0, st, F, Cl, K, Vf, Rf, Vc,
                       Wc, Sc, fas, active);
}

template <typename T>
void prolong_level(const Level &F, const Level &Cl, int K, cx<T> *Vf, const cx<T> *Vc,
                   const cx<T> *Wc, int fas, const int *active, hipStream_t st)
{
    const size_t n = F.plane * K;
    hipLaunchKernelGGL((prolong_kernel<T>), dim3(nblk(n)), dim3(256), 0, st, F, Cl, K, Vf, Vc, Wc,
                       fas, active);
}

void gather(int nx, int ny, int K, const Level &F, const double *wx, const double *dz,
            const double *uy, double beta, cx<double> *G, cx<double> *V, int direct_rhs,
            hipStream_t st)
{
    dim3 grid((nx + 31) / 32, (ny + 31) / 32, K);
    hipLaunchKernelGGL(gather_kernel, grid, dim3(256), 0, st, nx, ny, static_cast<size_t>(F.ld),
                       F.plane, reinterpret_cast<const cx<double> *>(wx),
                       reinterpret_cast<const cx<double> *>(dz),
                       reinterpret_cast<const cx<double> *>(uy), beta, G, V, direct_rhs);
}

void source(const LevelArgs<double> &L, const cx<double> *G, cx<double> *S, hipStream_t st)
{
    hipLaunchKernelGGL(source_kernel, dim3(nblk(L.plane * L.K)), dim3(256), 0, st, L, G, S);
}

void scatter(int nx, int ny, int K, const Level &F, const cx<double> *V, double *uy,
             hipStream_t st)
{
    dim3 grid((nx + 31) / 32, (ny + 31) / 32, K);
    hipLaunchKernelGGL(scatter_kernel, grid, dim3(256), 0, st, nx, ny, static_cast<size_t>(F.ld),
                       F.plane, V, reinterpret_cast<cx<double> *>(uy));
}

void decide(int K, const double *resmax, double tol, int cycle, int *active, int *cycles,
            double *lastres, int *nactive, hipStream_t st)
{
    hipLaunchKernelGGL(decide_kernel, dim3((K + 255) / 256), dim3(256), 0, st, K, resmax, tol,
                       cycle, active, cycles, lastres, nactive);
}

void fill(double *p, size_t n, double v, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(fill_kernel, dim3(nblk(n)), dim3(256), 0, st, p, n, v);
}

void fill_int(int *p, int n, int v, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

void update(const Level &F, int K, cx<double> *V, const cx<float> *E, const int *active,
            hipStream_t st)
{
    hipLaunchKernelGGL(update_kernel, dim3(nblk(F.plane * K)), dim3(256), 0, st, F, K, V, E, active);
}

template <typename TA, typename TB>
void convert(size_t n, const cx<TA> *a, cx<TB> *b, int zero, hipStream_t st)
{
    hipLaunchKernelGGL((convert_kernel<TA, TB>), dim3(nblk(n)), dim3(256), 0, st, n, a, b, zero);
}

// explicit instantiations
template int smooth_sweeps<double>(const LevelArgs<double> &, cx<double> *, const cx<double> *,
                                   const double *, int, double, const int *, int, cx<double> *,
                                   hipStream_t);
template int smooth_sweeps<float>(const LevelArgs<float> &, cx<float> *, const cx<float> *,
                                  const float *, int, float, const int *, int, cx<float> *,
                                  hipStream_t);
template int smooth_coarsest<double>(const LevelArgs<double> &, cx<double> *, const cx<double> *,
                                     const double *, int, double, const int *, int, hipStream_t);
template void defect<double, double>(const LevelArgs<double> &, const cx<double> *,
                                     const cx<double> *, cx<double> *, double, const int *,
                                     double *, hipStream_t);
template void defect<double, float>(const LevelArgs<double> &, const cx<double> *,
                                    const cx<double> *, cx<float> *, double, const int *, double *,
                                    hipStream_t);
template void defect<float, float>(const LevelArgs<float> &, const cx<float> *, const cx<float> *,
                                   cx<float> *, float, const int *, double *, hipStream_t);
template void restrict_level<double>(const Level &, const Level &, int, const cx<double> *,
                                     const cx<double> *, cx<double> *, cx<double> *, cx<double> *,
                                     int, const int *, hipStream_t);
template void restrict_level<float>(const Level &, const Level &, int, const cx<float> *,
                                    const cx<float> *, cx<float> *, cx<float> *, cx<float> *, int,
                                    const int *, hipStream_t);
template void prolong_level<double>(const Level &, const Level &, int, cx<double> *,
                                    const cx<double> *, const cx<double> *, int, const int *,
                                    hipStream_t);
template void prolong_level<float>(const Level &, const Level &, int, cx<float> *,
                                   const cx<float> *, const cx<float> *, int, const int *,
                                   hipStream_t);
template void convert<double, float>(size_t, const cx<double> *, cx<float> *, int, hipStream_t);
template void convert<float, float>(size_t, const cx<float> *, cx<float> *, int, hipStream_t);
template void convert<double, double>(size_t, const cx<double> *, cx<double> *, int, hipStream_t);
template void convert<float, double>(size_t, const cx<float> *, cx<double> *, int, hipStream_t);

}  // namespace tbl
