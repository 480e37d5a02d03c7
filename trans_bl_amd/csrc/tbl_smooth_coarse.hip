// Coarsest-level smoother: every sweep of the level in one launch, one
// workgroup per mode with the level resident in LDS.
#include <type_traits>

#include "tbl_relax.h"

namespace tbl {

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
    const T *glu = lu + (static_cast<size_t>(k) * 2) * kLuFields * L.lus;
    for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
        const int i = e / ny, j = e % ny;
        sv[i * ls + j] = gv[static_cast<size_t>(i) * ld + j];
    }
    for (int e = threadIdx.x; e < 2 * kLuFields * ny; e += blockDim.x)
        slu[e] = glu[static_cast<size_t>(e / ny) * L.lus + e % ny];
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
                    load_line_lu<T, R>(slu + (side ? kLuFields * ny : 0), ny, line_j0(lane, R), ny);
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
// (penta_par.f:367-404).  v, the factors and (when it fits) the source stay
// in LDS; the forward-substitution values stay in registers.  The row count
// NY is a compile-time constant (one instantiation per supported coarsest
// ny) and so is the line's border side, so a line solve is straight-line
// code and its LDS loads are scheduled ahead of the serial chain.  The back
// substitution multiplies by the precomputed -u1/p, -u2/p, 1/p instead of
// dividing by the pivot (last-bit rounding differs from the division).
template <typename T>
__host__ __device__ inline size_t coarse_reg_lds_bytes(int nx, int ny, bool with_s)
{
    return static_cast<size_t>(coarse_stride(ny)) * nx * sizeof(cx<T>) * (with_s ? 2 : 1) +
           static_cast<size_t>(2 * kLuFields) * ny * sizeof(T);
}

template <typename T, int NY, int SIDE>
__device__ __forceinline__ void coarse_line(const LevelArgs<T> &L, cx<T> *col, const cx<T> *scol,
                                            const T *f, T rf)
{
    constexpr int LS = NY | 1;
    const T qx = L.qx;
    auto X = [&](int j) { return xoff<T>(col, LS, j, SIDE, SIDE == 0 && j >= NY - 2); };
    const T *l1 = f, *l2 = f + NY, *m1 = f + 5 * NY, *m2 = f + 6 * NY, *rp = f + 7 * NY;
    cx<T> Y[NY];
    cx<T> xm = {T(0), T(0)}, x0 = X(0), xp = X(1);
    cx<T> y1 = {T(0), T(0)}, y2 = {T(0), T(0)};
#pragma unroll
    for (int j = 0; j < NY; ++j) {
        cx<T> rhs;
        if (j == 0) rhs = col[0];
        else if (j == 1) rhs = scol[1] - qx * (L.sp[2] * xp + L.sp[1] * x0);
        else if (j <= NY - 3) rhs = scol[j] - qx * (L.cp[0] * xm + L.cp[1] * x0 + L.cp[2] * xp);
        else rhs = scol[j] - qx * x0;
        const cx<T> yy = (rhs - l2[j] * y2) - l1[j] * y1;
        Y[j] = yy;
        y2 = y1; y1 = yy;
        xm = x0; x0 = xp;
        if (j + 2 < NY) xp = X(j + 2);
    }
    // back substitution into Y, then the relaxation stores: no LDS store on
    // the serial chain, so the factor loads need not wait for one
    cx<T> a = {T(0), T(0)}, b = {T(0), T(0)};
#pragma unroll
    for (int j = NY - 1; j >= 0; --j) {
        const cx<T> xx = (rp[j] * Y[j] + m2[j] * b) + m1[j] * a;
        Y[j] = xx;
        b = a; a = xx;
    }
#pragma unroll
    for (int j = 1; j < NY; ++j) col[j] = col[j] + rf * (Y[j] - col[j]);
}

// SL: the source is staged in LDS (else read from global memory); a
// compile-time choice, so every access has a known address space.
template <typename T, int NY, bool SL>
__global__ void __launch_bounds__(256) smooth_coarsest_reg_kernel(LevelArgs<T> L, cx<T> *V,
                                                                   const cx<T> *S, const T *lu,
                                                                   int sweeps, T rf,
                                                                   const int *active)
{
    constexpr int LS = NY | 1;
    extern __shared__ unsigned char smem_raw[];
    const int k = blockIdx.x;
    if (active && !active[k]) return;
    const int nx = L.nx, ld = L.ld;
    cx<T> *sv = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *ss = sv + static_cast<size_t>(LS) * nx;
    cx<T> *gv = V + static_cast<size_t>(k) * L.plane;
    const cx<T> *gs = S + static_cast<size_t>(k) * L.plane;
    T *slu = reinterpret_cast<T *>(SL ? ss + static_cast<size_t>(LS) * nx : ss);
    const T *glu = lu + (static_cast<size_t>(k) * 2) * kLuFields * L.lus;
    for (int e = threadIdx.x; e < nx * NY; e += blockDim.x) {
        const int i = e / NY, j = e % NY;
        sv[i * LS + j] = gv[static_cast<size_t>(i) * ld + j];
        if (SL) ss[i * LS + j] = gs[static_cast<size_t>(i) * ld + j];
    }
    for (int e = threadIdx.x; e < 2 * kLuFields * NY; e += blockDim.x)
        slu[e] = glu[static_cast<size_t>(e / NY) * L.lus + e % NY];
    __syncthreads();
    const T *lu_centre = slu;
    const T *lu_border = slu + kLuFields * NY;
    const int c = threadIdx.x;
    for (int it = 0; it < sweeps; ++it) {
        for (int phase = 0; phase < 4; ++phase) {
            if (c < phase_lines(phase, nx)) {
                const int i = phase_column(phase, c, nx);
                cx<T> *col = sv + static_cast<size_t>(i) * LS;
                const cx<T> *scol = SL ? ss + static_cast<size_t>(i) * LS
                                       : gs + static_cast<size_t>(i) * ld;
                if (phase != 0) coarse_line<T, NY, 0>(L, col, scol, lu_centre, rf);
                else if (c == 0) coarse_line<T, NY, -1>(L, col, scol, lu_border, rf);
                else coarse_line<T, NY, 1>(L, col, scol, lu_border, rf);
            }
            __syncthreads();
        }
    }
    for (int e = threadIdx.x; e < nx * NY; e += blockDim.x) {
        const int i = e / NY, j = e % NY;
        gv[static_cast<size_t>(i) * ld + j] = sv[i * LS + j];
    }
}

// Two-stage form of the same solve, used when its LDS fits: per phase, all
// four wavefronts first form the right-hand sides of every (line, row) of
// the phase into LDS (the bulk of the arithmetic, spread over the CU's
// SIMDs); then two threads per line -- the real and the imaginary part, the
// factors being real -- run the two substitutions and the relaxation.  The
// substitutions are contracted to fused multiply-adds (one dependent
// operation per row and direction); the right-hand sides keep the
// reference's operation order.  v, the source and the right-hand sides are
// in LDS; the factors, the same for every line of a phase, are read from
// global memory at wavefront-uniform addresses (scalar loads, no LDS
// traffic in the chains).
template <typename T, bool PART = false>
__host__ __device__ inline size_t coarse_2s_lds_bytes(int nx, int ny)
{
    const size_t ls = static_cast<size_t>(coarse_stride(ny));
    return (2 * ls * nx + ls * phase_lines(1, nx)) * (PART ? sizeof(T) : sizeof(cx<T>));
}

// one part (real or imaginary; element stride ST: 2 in a complex line, 1
// in a line of one part) of one line; f: the factors' fields, fs apart
template <typename T, int NY, int ST = 2>
__device__ __forceinline__ void coarse_chain(const T *rb, T *col, const T *__restrict__ f, int fs,
                                             T rf)
{
    const T *l1 = f, *l2 = f + fs, *m1 = f + 5 * fs, *m2 = f + 6 * fs, *rp = f + 7 * fs;
    T Y[NY];
    T y1 = T(0), y2 = T(0);
    // y_j = (r_j - l2_j y_{j-2}) - l1_j y_{j-1}
#pragma unroll
    for (int j = 0; j < NY; ++j) {
        const T yy = fma(-l1[j], y1, fma(-l2[j], y2, rb[ST * j]));
        Y[j] = yy;
        y2 = y1; y1 = yy;
    }
    // x_j = (y_j / p_j - (u2_j / p_j) x_{j+2}) - (u1_j / p_j) x_{j+1}
    T a = T(0), b = T(0);
#pragma unroll
    for (int j = NY - 1; j >= 0; --j) {
        const T xx = fma(m1[j], a, fma(m2[j], b, rp[j] * Y[j]));
        Y[j] = xx;
        b = a; a = xx;
    }
#pragma unroll
    for (int j = 1; j < NY; ++j) col[ST * j] = fma(rf, Y[j] - col[ST * j], col[ST * j]);
}

// threads per workgroup: 512 for NY <= 16, 768 up to 32 rows, 1024 when a
// row set takes 64 lanes (NY > 32) -- more line groups per stage-1 pass;
// measured per sweep on 129-column coarsest levels: NY = 25 2-5 % and
// NY = 33 7-10 % faster than 512, while 1024 was slower at NY = 25 and 13.  Stage 2 needs 2 threads per line, so the dispatch requires
// 2 * lines <= kCoarse2sThreads (the smallest count).
constexpr int kCoarse2sThreads = 512;
template <int NY>
constexpr int kCoarse2sThreadsFor = NY > 32 ? 1024 : NY > 16 ? 768 : kCoarse2sThreads;

// PART: the operator's coefficients are real, so the real and imaginary
// parts of a mode are two independent solves; with PART a workgroup takes
// one part (blockIdx.y) of one mode -- twice the workgroups, each with half
// the right-hand-side work, half the LDS and one thread per line in the
// substitutions -- with the same operations on each part as the complex
// form, so the same results bit for bit.
template <typename T, int NY, bool PART = false>
__global__ void __launch_bounds__(kCoarse2sThreadsFor<NY>) smooth_coarsest_2s_kernel(LevelArgs<T> L, cx<T> *__restrict__ V,
                                                                  const cx<T> *__restrict__ S,
                                                                  const T *__restrict__ lu,
                                                                  int sweeps, T rf,
                                                                  const int *active)
{
    using E = std::conditional_t<PART, T, cx<T>>;   // an LDS element
    constexpr int LS = NY | 1;
    extern __shared__ unsigned char smem_raw[];
    const int k = blockIdx.x;
    if (active && !active[k]) return;
    const int part = PART ? static_cast<int>(blockIdx.y) : 0;
    const int nx = L.nx, ld = L.ld;
    E *sv = reinterpret_cast<E *>(smem_raw);
    E *ss = sv + static_cast<size_t>(LS) * nx;
    E *rbuf = ss + static_cast<size_t>(LS) * nx;
    cx<T> *gv = V + static_cast<size_t>(k) * L.plane;
    const cx<T> *gs = S + static_cast<size_t>(k) * L.plane;
    const T *glu = lu + (static_cast<size_t>(k) * 2) * kLuFields * L.lus;
    // the workgroup's part of a complex value, and the store back
    auto get = [&](const cx<T> &z) -> E {
        if constexpr (PART) return part ? z.im : z.re;
        else return z;
    };
    for (int e = threadIdx.x; e < nx * NY; e += blockDim.x) {
        const int i = e / NY, j = e % NY;
        sv[i * LS + j] = get(gv[static_cast<size_t>(i) * ld + j]);
        ss[i * LS + j] = get(gs[static_cast<size_t>(i) * ld + j]);
    }
    __syncthreads();
    const T qx = L.qx;
    // stage-1 lane roles for the colour phases: RW lanes per row set, NG line
    // groups; each lane's row kind is fixed, so its stencil weights are set
    // once: (a, b, c) applied as (a X(j-1) + b X(j)) + c X(j+1), which equals
    // the reference's row forms exactly (zero terms add nothing, and the
    // two-term sum of row 2 only commutes)
    constexpr int RW = NY <= 32 ? 32 : 64;
    constexpr int NG = kCoarse2sThreadsFor<NY> / RW;
    const int jr = threadIdx.x % RW, cs = threadIdx.x / RW;
    T ca = T(0), cb = T(1), cc = T(0);
    if (jr == 1) {
        cb = L.sp[1];
        cc = L.sp[2];
    } else if (jr >= 2 && jr <= NY - 3) {
        ca = L.cp[0];
        cb = L.cp[1];
        cc = L.cp[2];
    }
    const bool grouped = jr >= NY - 2;
    const E zero = E{};
    for (int it = 0; it < sweeps; ++it) {
        for (int phase = 0; phase < 4; ++phase) {
            const int nl = phase_lines(phase, nx);
            if (phase == 0) {
                for (int e = threadIdx.x; e < nl * NY; e += blockDim.x) {
                    const int c = e / NY, j = e % NY;
                    const int i = phase_column(0, c, nx);
                    const int side = c == 0 ? -1 : 1;
                    const E *col = sv + static_cast<size_t>(i) * LS;
                    // xoff's border form: 10 a - 4 b + 14 c - 6 d + e
                    auto X = [&](int jj) -> E {
                        const int s = side < 0 ? LS : -LS;
                        return mac(T(-6), col[jj + 3 * s],
                                   mac(T(14), col[jj + 2 * s],
                                       mac(T(-4), col[jj + s], T(10) * col[jj - s]))) +
                               col[jj + 4 * s];
                    };
                    E rhs;
                    if (j == 0) {
                        rhs = col[0];
                    } else {
                        const E sj = ss[static_cast<size_t>(i) * LS + j];
                        if (j == 1) rhs = sj - qx * (L.sp[2] * X(2) + L.sp[1] * X(1));
                        else if (j <= NY - 3)
                            rhs = sj - qx * (L.cp[0] * X(j - 1) + L.cp[1] * X(j) + L.cp[2] * X(j + 1));
                        else rhs = sj - qx * X(j);
                    }
                    rbuf[c * LS + j] = rhs;
                }
            } else {
                // colour phases: lane (cs, jr) forms row jr of lines cs, cs+NG, ...;
                // the row's own x term from four loads, its neighbours' by shuffles
                const int ncg = (nl + NG - 1) / NG;
                for (int q = 0; q < ncg; ++q) {
                    const int c = cs + q * NG;
                    const int i = 2 + (phase - 1) + 3 * min(c, nl - 1);
                    const E *col = sv + static_cast<size_t>(i) * LS;
                    E x = zero;
                    if (jr < NY) {
                        const E m2 = col[jr - 2 * LS], m1 = col[jr - LS];
                        const E p1 = col[jr + LS], p2 = col[jr + 2 * LS];
                        x = grouped ? -(p2 + m2) + T(16) * (p1 + m1)
                                    : T(-1) * m2 + T(16) * m1 + T(16) * p1 - T(1) * p2;
                    }
                    E xm, xp;
                    if constexpr (PART) {
                        xm = __shfl_up(x, 1, RW);
                        xp = __shfl_down(x, 1, RW);
                    } else {
                        xm = shfl_up(x, 1, RW);
                        xp = shfl_down(x, 1, RW);
                    }
                    if (c < nl && jr < NY) {
                        const E rhs =
                            jr == 0 ? col[0]
                                    : ss[static_cast<size_t>(i) * LS + jr] -
                                          qx * ((ca * xm + cb * x) + cc * xp);
                        rbuf[c * LS + jr] = rhs;
                    }
                }
            }
            __syncthreads();
            // stage 2: the substitutions, one thread per line and part
            const int c = PART ? static_cast<int>(threadIdx.x) : static_cast<int>(threadIdx.x >> 1);
            const int pp = PART ? 0 : static_cast<int>(threadIdx.x & 1);
            if (c < nl) {
                const int i = phase_column(phase, c, nx);
                const T *fb = glu + (phase == 0 ? kLuFields * L.lus : 0);
                if constexpr (PART)
                    coarse_chain<T, NY, 1>(rbuf + c * LS, sv + static_cast<size_t>(i) * LS, fb,
                                           L.lus, rf);
                else
                    coarse_chain<T, NY>(reinterpret_cast<const T *>(rbuf + c * LS) + pp,
                                        reinterpret_cast<T *>(sv + static_cast<size_t>(i) * LS) + pp,
                                        fb, L.lus, rf);
            }
            __syncthreads();
        }
    }
    for (int e = threadIdx.x; e < nx * NY; e += blockDim.x) {
        const int i = e / NY, j = e % NY;
        if constexpr (PART) {
            T *g = reinterpret_cast<T *>(gv + static_cast<size_t>(i) * ld + j) + part;
            *g = sv[i * LS + j];
        } else {
            gv[static_cast<size_t>(i) * ld + j] = sv[i * LS + j];
        }
    }
}

// coarsest row counts with a one-thread-per-line instantiation (ny of the
// coarsest level = (ny-1)/2^(levels-1) + 1 for the usual grids)
#define TBL_COARSE_NY(X) X(9) X(13) X(17) X(25) X(33)

template <typename T, int G, int R>
static void launch_coarsest(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu,
                            int sweeps, T rf, const int *active, hipStream_t st, int threads)
{
    const size_t shm = coarse_lds_bytes<T>(L.nx, L.ny, false);
    hipLaunchKernelGGL((smooth_coarsest_kernel<T, G, R>), dim3(L.K), dim3(threads), shm, st, L, V,
                       S, lu, sweeps, rf, active);
}

template <typename T>
int smooth_coarsest(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps,
                    T rf, const int *active, int G, hipStream_t st)
{
    bool reg_ny = false;
#define X(NY) reg_ny = reg_ny || L.ny == NY;
    TBL_COARSE_NY(X)
#undef X
    // shape kCoarseSeqShape: the sequential kernels where they apply; any other
    // shape, or a level they do not fit, the G-lane scan kernel
    if (G != kCoarseSeqShape) reg_ny = false;
    G = sweep_shape(G);
    if (reg_ny && coarse_2s_lds_bytes<T>(L.nx, L.ny) <= 160 * 1024 &&
        2 * phase_lines(1, L.nx) <= kCoarse2sThreads) {
        // one workgroup per (mode, real or imaginary part)
        const size_t shm = coarse_2s_lds_bytes<T, true>(L.nx, L.ny);
        bool done = false;
#define X(NY)                                                                                 \
    if (!done && L.ny == NY) {                                                                \
        hipLaunchKernelGGL((smooth_coarsest_2s_kernel<T, NY, true>), dim3(L.K, 2),            \
                           dim3(kCoarse2sThreadsFor<NY>),                                     \
                           shm, st,                                                           \
                           L, V, S, lu, sweeps, rf, active);                                  \
        done = true;                                                                          \
    }
        TBL_COARSE_NY(X)
#undef X
        return TBL_OK;
    }
    if (reg_ny && coarse_reg_lds_bytes<T>(L.nx, L.ny, false) <= 160 * 1024 &&
        phase_lines(1, L.nx) <= 256) {
        const int s_in = coarse_reg_lds_bytes<T>(L.nx, L.ny, true) <= 160 * 1024;
        const int threads = (phase_lines(1, L.nx) + 63) / 64 * 64;
        const size_t shm = coarse_reg_lds_bytes<T>(L.nx, L.ny, s_in);
        bool done = false;
#define X(NY)                                                                                 \
    if (!done && L.ny == NY) {                                                                \
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
        return smooth_sweeps<T>(L, V, S, lu, sweeps, rf, active,
                                shape_is_strip(G) ? G - kStripShape : G, nullptr, st);
    G = shape_lanes(G);
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

int coarse_kernel_attributes()
{
    hipError_t e = hipSuccess;
#define X(GG, RR)                                                                             \
    for (const void *f : {reinterpret_cast<const void *>(&smooth_coarsest_kernel<double, GG, RR>), \
                          reinterpret_cast<const void *>(&smooth_coarsest_kernel<float, GG, RR>)}) { \
        hipError_t r = allow_full_lds(f);                                                     \
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
        hipError_t r = allow_full_lds(f);                                                     \
        if (r != hipSuccess) e = r;                                                           \
    }
    TBL_COARSE_NY(X)
#undef X
#define X(NY)                                                                                 \
    for (const void *f : {reinterpret_cast<const void *>(&smooth_coarsest_2s_kernel<double, NY>), \
                          reinterpret_cast<const void *>(&smooth_coarsest_2s_kernel<float, NY>), \
                          reinterpret_cast<const void *>(&smooth_coarsest_2s_kernel<double, NY, true>), \
                          reinterpret_cast<const void *>(&smooth_coarsest_2s_kernel<float, NY, true>)}) { \
        hipError_t r = allow_full_lds(f);                                                     \
        if (r != hipSuccess) e = r;                                                           \
    }
    TBL_COARSE_NY(X)
#undef X
    return e == hipSuccess ? TBL_OK : TBL_ERR_HIP;
}

template int smooth_coarsest<double>(const LevelArgs<double> &, cx<double> *, const cx<double> *,
                                     const double *, int, double, const int *, int, hipStream_t);
template int smooth_coarsest<float>(const LevelArgs<float> &, cx<float> *, const cx<float> *,
                                    const float *, int, float, const int *, int, hipStream_t);

}  // namespace tbl
