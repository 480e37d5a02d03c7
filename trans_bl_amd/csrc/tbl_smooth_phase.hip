// Colour-phase line smoother: one launch per phase (border columns, then
// the three interior colours), direct or LDS-staged global access.
#include "tbl_relax.h"

namespace tbl {

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
        // G*R >= ny: a fixed trip count lets every load be issued up front
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = lane + r * G;
            if (j < ny) {
                Xs[j] = xoff<T>(col, ld, j, side, side == 0 && j >= ny - 2);
                Vs[j] = col[j];
                Ss[j] = S[base + j];
            }
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
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = lane + r * G;
            if (j >= 1 && j < ny) col[j] = Vs[j];
        }
    }
}

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

// One colour phase per launch.  G > 0: lanes per line with direct
// (lane-contiguous) global access; G < 0: |G| lanes per line, LDS-staged.
template <typename T>
int phase_sweeps(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps, T rf,
                 const int *active, int G, hipStream_t st)
{
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

int phase_kernel_attributes()
{
    hipError_t e = hipSuccess;
#define X(GG, RR)                                                                          \
    for (const void *f : {reinterpret_cast<const void *>(&smooth_staged_kernel<double, GG, RR>), \
                          reinterpret_cast<const void *>(&smooth_staged_kernel<float, GG, RR>)}) { \
        hipError_t r = allow_full_lds(f);                                                  \
        if (r != hipSuccess) e = r;                                                        \
    }
    TBL_LINE_SHAPES(X)
#undef X
    return e == hipSuccess ? TBL_OK : TBL_ERR_HIP;
}

template int phase_sweeps<double>(const LevelArgs<double> &, cx<double> *, const cx<double> *,
                                  const double *, int, double, const int *, int, hipStream_t);
template int phase_sweeps<float>(const LevelArgs<float> &, cx<float> *, const cx<float> *,
                                 const float *, int, float, const int *, int, hipStream_t);

}  // namespace tbl
