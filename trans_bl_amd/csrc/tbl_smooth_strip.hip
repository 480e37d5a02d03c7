// Fused-sweep strip smoother: a whole sweep per workgroup in LDS.
#include "tbl_relax.h"

namespace tbl {

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
    const int j0 = line_j0(lane, R);
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
    const __amdgpu_buffer_rsrc_t lr =
        buf_rsrc(lu, static_cast<size_t>(L.K) * 2 * kLuFields * L.lus * sizeof(T));
    const unsigned tbl = static_cast<unsigned>(kLuFields * L.lus * sizeof(T));
    const LineLU<T, R> fc = load_line_lu_buf<T, R>(lr, (2 * k) * tbl, L.lus, j0);
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
                run(load_line_lu_buf<T, R>(lr, (2 * k + 1) * tbl, L.lus, j0));
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

// Fused-sweep strips: G lanes per line, one launch (plus the halo snapshot)
// per sweep.
template <typename T>
int strip_sweeps(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps, T rf,
                 const int *active, int G, cx<T> *halo, hipStream_t st)
{
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

int strip_kernel_attributes()
{
    hipError_t e = hipSuccess;
#define X(GG, RR)                                                                          \
    for (const void *f : {reinterpret_cast<const void *>(&smooth_strip_kernel<double, GG, RR>), \
                          reinterpret_cast<const void *>(&smooth_strip_kernel<float, GG, RR>)}) { \
        hipError_t r = allow_full_lds(f);                                                  \
        if (r != hipSuccess) e = r;                                                        \
    }
    TBL_LINE_SHAPES(X)
#undef X
    return e == hipSuccess ? TBL_OK : TBL_ERR_HIP;
}

template int strip_sweeps<double>(const LevelArgs<double> &, cx<double> *, const cx<double> *,
                                  const double *, int, double, const int *, int, cx<double> *,
                                  hipStream_t);
template int strip_sweeps<float>(const LevelArgs<float> &, cx<float> *, const cx<float> *,
                                 const float *, int, float, const int *, int, cx<float> *,
                                 hipStream_t);

}  // namespace tbl
