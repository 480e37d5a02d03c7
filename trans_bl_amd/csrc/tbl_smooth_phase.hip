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
    const T *lut = lu + (static_cast<size_t>(k) * 2 + (side != 0)) * kLuFields * L.lus;
    relax_line<T, G, R>(L, V + base, S + base, lut, side, rf, lane);
}

constexpr int kStagedLines = 4;   // lines per workgroup of the staged smoother
constexpr int kXcds = 8;          // MI355X: 8 XCDs, one L2 each

// Same relaxation with coalesced global traffic: the group first reads its
// line lane-strided (row t, t+G, ...: contiguous across the wavefront),
// forming the off-diagonal x terms in LDS -- with FULL also v and S, else v
// and S are read and v written lane-contiguous from global memory (a third
// of the LDS, so more lines in flight per CU); the scan then works on
// lane-contiguous segments; with FULL the relaxed v goes back through LDS
// and is stored lane-strided.  Arithmetic identical to smooth_phase_kernel.
// The stage holds rows 1 + lane + s G (s < R: rows 1..G R, the rows the
// lanes relax) and, with FULL, v of row 0 (the line's initial state).
// UPD: vup (the FP64 field of the same layout) += the relaxed v on the
// line's rows 1.., its values loaded with the line's other inputs
template <typename T, int G, int R, bool FULL, bool UPD = false>
__global__ void __launch_bounds__(256) smooth_staged_kernel(LevelArgs<T> L, cx<T> *V,
                                                            const cx<T> *S, const T *lu,
                                                            int phase, T rf, const int *active,
                                                            int zs, cx<double> *vup)
{
    extern __shared__ unsigned char smem_raw[];
    const int LDW = L.ld;                                // rows held per line
    cx<T> *base_s = reinterpret_cast<cx<T> *>(smem_raw);
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    cx<T> *Xs = base_s + static_cast<size_t>(grp) * (FULL ? 3 : 1) * LDW;
    cx<T> *Vs = Xs + LDW, *Ss = Vs + LDW;
    const int nlines = phase_lines(phase, L.nx);
    // XCD-aware order: the hardware deals workgroups round-robin to the 8
    // XCDs (separate L2s), so workgroup n takes block (n % 8) * per + n / 8
    // -- consecutive blocks, whose lines share neighbour columns, run on one
    // XCD one after the other (the grid is padded to 8 * per blocks)
    const int per = static_cast<int>(gridDim.x) / kXcds;
    const int blk = static_cast<int>(blockIdx.x % kXcds) * per + static_cast<int>(blockIdx.x / kXcds);
    const int gid = blk * kStagedLines + grp;
    const int k = gid / nlines;
    const bool live = k < L.K && !(active && !active[k]);
    const int ny = L.ny, ld = L.ld;
    // buffer resources over the (at most two) modes of this workgroup's
    // lines: 32-bit offsets, the row steps of the loads in the instruction
    const int g0 = blk * kStagedLines;
    const int kf = g0 / nlines;
    const int kl = max(kf, min((g0 + kStagedLines - 1) / nlines, L.K - 1));   // padding: kf >= K
    const size_t span = static_cast<size_t>(kl - kf + 1) * L.plane * sizeof(cx<T>);
    const __amdgpu_buffer_rsrc_t vr = buf_rsrc(V + static_cast<size_t>(kf) * L.plane, span);
    const __amdgpu_buffer_rsrc_t sr = buf_rsrc(S + static_cast<size_t>(kf) * L.plane, span);
    int side = 0;
    size_t base = 0;
    cx<double> up0[UPD ? R : 1];
    if (live) {
        const int c = gid - k * nlines;
        const int i = phase_column(phase, c, L.nx);
        side = (phase == 0) ? (c == 0 ? -1 : 1) : 0;
        base = static_cast<size_t>(k) * L.plane + static_cast<size_t>(i) * ld;
        // Every load of the line is issued before any is used: the rows'
        // x-neighbour columns (border: the one-sided stencil's five columns
        // on the inner side; interior: i-2, i-1, i+1, i+2), v and S; rows
        // past ny read padding or zero and are not used.  xoff's arithmetic
        // follows.
        const int sg = -side;
        // neighbour columns i + d0 .. i + d4
        const int d0 = side ? -sg : -2, d1 = side ? sg : -1, d2 = side ? 2 * sg : 1,
                  d3 = side ? 3 * sg : 2, d4 = side ? 4 * sg : 2;
        constexpr unsigned ES = sizeof(cx<T>);
        const unsigned lo = static_cast<unsigned>(((k - kf) * L.plane + static_cast<size_t>(i) * ld + lane) * ES);
        // zero start (zs: the first sweep from v = 0): a column this sweep
        // has not relaxed yet -- the boundary columns, the border columns
        // before phase 1, a colour before its phase, the line's own column
        // -- reads as zero through an out-of-range offset, not from memory
        auto col_off = [&](int d) -> unsigned {
            const int c = i + d;
            bool known = true;
            if (zs) {
                if (d == 0 || c <= 0 || c >= L.nx - 1) known = false;
                else if (c == 1 || c == L.nx - 2) known = phase >= 1;
                else known = (c - 2) % 3 + 1 < phase;
            }
            return known ? lo + static_cast<unsigned>(d * ld) * ES : kBufOff;
        };
        const unsigned b0 = col_off(d0), b1 = col_off(d1), b2 = col_off(d2), b3 = col_off(d3),
                       b4 = col_off(d4), bv = col_off(0);
        cx<T> a0[R], a1[R], a2[R], a3[R], a4[R], vv[R], ss[R];
        if constexpr (UPD) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = 1 + lane + r * G;
                if (j < ny) up0[r] = vup[base + j];
            }
        }
        // row 0's v (lane 0's row-0 offset is lo = bv)
        const cx<T> v0 = FULL && lane == 0 ? buf_load(vr, bv, T(0)) : cx<T>{T(0), T(0)};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // rows past ny (up to G*R) read zero, not the next column
            const bool in = 1 + lane + r * G < ny;
            const unsigned rs = (1 + r * G) * ES;
            auto at = [&](unsigned b) { return in ? b + rs : kBufOff; };
            a0[r] = buf_load(vr, at(b0), T(0));
            a1[r] = buf_load(vr, at(b1), T(0));
            a2[r] = buf_load(vr, at(b2), T(0));
            a3[r] = buf_load(vr, at(b3), T(0));
            a4[r] = buf_load(vr, at(b4), T(0));
            if (FULL) {
                vv[r] = buf_load(vr, at(bv), T(0));
                ss[r] = buf_load(sr, at(lo), T(0));
            }
        }
        if (lane == 0) {
            if (FULL) Vs[0] = v0;
            Xs[0] = cx<T>{T(0), T(0)};   // read by relax_core, used by no row
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = 1 + lane + r * G;
            if (j < ny) {
                cx<T> x;
                // the sums of xoff in its order (multiply-adds in FP32)
                if (side)
                    x = mac(T(-6), a3[r], mac(T(14), a2[r], mac(T(-4), a1[r], T(10) * a0[r])))
                        + a4[r];
                else if (j >= ny - 2)
                    x = mac(T(16), a2[r] + a1[r], -(a3[r] + a0[r]));
                else
                    x = mac(T(-1), a3[r], mac(T(16), a2[r], mac(T(16), a1[r], T(-1) * a0[r])));
                Xs[j] = x;
                if (FULL) {
                    Vs[j] = vv[r];
                    Ss[j] = ss[r];
                }
            }
        }
    }
    __syncthreads();
    if (live) {
        const __amdgpu_buffer_rsrc_t lr =
            buf_rsrc(lu, static_cast<size_t>(L.K) * 2 * kLuFields * L.lus * sizeof(T));
        const unsigned lo = static_cast<unsigned>((static_cast<size_t>(k) * 2 + (side != 0)) *
                                                  kLuFields * L.lus * sizeof(T));
        const LineLU<T, R> f = load_line_lu_buf<T, R>(lr, lo, L.lus, line_j0(lane, R));
        if (FULL) {
            relax_core<T, G, R>(L, f, rf, lane, [&](int j) { return Xs[j]; },
                                [&](int j) { return Vs[j]; }, [&](int j, int) { return Ss[j]; },
                                [&](int j, cx<T> v) { Vs[j] = v; });
        } else {
            cx<T> *col = V + base;
            const cx<T> *scol = S + base;
            relax_core<T, G, R>(L, f, rf, lane, [&](int j) { return Xs[j]; },
                                [&](int j) { return col[j]; }, [&](int j, int) { return scol[j]; },
                                [&](int j, cx<T> v) { col[j] = v; });
        }
    }
    if (FULL) {
        __syncthreads();
        if (live) {
            cx<T> *col = V + base;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = 1 + lane + r * G;
                if (j < ny) {
                    const cx<T> e = Vs[j];
                    // the updating sweep's last colour is read by no later
                    // phase, and the next correction solve starts from zero
                    if (!(UPD && phase == 3)) col[j] = e;
                    // UPD (the last sweep of a correction solve): this
                    // line's e is final, so it is added to the FP64 field
                    // here (same layout), as the update kernel would
                    if constexpr (UPD) {
                        vup[base + j] = cx<double>{up0[r].re + static_cast<double>(e.re),
                                                   up0[r].im + static_cast<double>(e.im)};
                    } else if (vup) {
                        // the same update in runtime form: the dispatch never
                        // passes a field to these instantiations, but with
                        // this path the compiler schedules them as measured
                        // fastest (the FP64 finest-level kernel ran 4.6 %
                        // slower without it)
                        cx<double> &u = vup[base + j];
                        u = cx<double>{u.re + static_cast<double>(e.re),
                                       u.im + static_cast<double>(e.im)};
                    }
                }
            }
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

template <bool FULL>
static size_t staged_lds(int ld, size_t esize)
{
    return static_cast<size_t>(kStagedLines) * (FULL ? 3 : 1) * ld * esize;
}

template <typename T, int G, int R, bool FULL, bool UPD = false>
static void launch_staged(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int phase,
                          T rf, const int *active, hipStream_t st, int zs, cx<double> *vup)
{
    const size_t lines = static_cast<size_t>(phase_lines(phase, L.nx)) * L.K;
    if (!lines) return;
    // blocks padded to a multiple of the XCD count (the kernel's block order)
    const size_t blocks = (lines + kStagedLines - 1) / kStagedLines;
    const size_t per = (blocks + kXcds - 1) / kXcds;
    hipLaunchKernelGGL((smooth_staged_kernel<T, G, R, FULL, UPD>),
                       dim3(static_cast<unsigned>(per * kXcds)),
                       dim3(kStagedLines * G), staged_lds<FULL>(L.ld, sizeof(cx<T>)), st, L, V, S,
                       lu, phase, rf, active, zs, vup);
}

// the FULL staged form, with the update variant (FP32 only) when vup is set
template <typename T, int G, int R>
static void launch_full(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int phase,
                        T rf, const int *active, hipStream_t st, int zs, cx<double> *vup)
{
    if constexpr (sizeof(T) == 4) {
        if (vup) {
            launch_staged<T, G, R, true, true>(L, V, S, lu, phase, rf, active, st, zs, vup);
            return;
        }
    }
    launch_staged<T, G, R, true>(L, V, S, lu, phase, rf, active, st, zs, nullptr);
}

// One colour phase per launch.  G > 0: lanes per line with direct
// (lane-contiguous) global access; G < 0: |G| lanes per line, x terms, v
// and S staged in LDS; G > kXStagedShape: G - kXStagedShape lanes per line,
// only the x terms staged.
template <typename T>
int phase_sweeps(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps, T rf,
                 const int *active, int G, hipStream_t st, bool zero_start, cx<double> *vup)
{
    const int form = G < 0 ? 1 : G > kXStagedShape ? 2 : 0;
    if ((zero_start || vup) && form != 1) return TBL_ERR_ARG;
    if (vup && sizeof(T) != 4) return TBL_ERR_ARG;   // the update variant is FP32 only
    G = form == 1 ? -G : form == 2 ? G - kXStagedShape : G;
    const int R = line_rows_for(G, L.ny);
    if (R < 0) return TBL_ERR_GRID;
    if (form == 1 && staged_lds<true>(L.ld, sizeof(cx<T>)) > 160 * 1024) return TBL_ERR_GRID;
    if (form == 2 && staged_lds<false>(L.ld, sizeof(cx<T>)) > 160 * 1024) return TBL_ERR_GRID;
    for (int it = 0; it < sweeps; ++it)
        for (int ph = 0; ph < 4; ++ph) {
            bool done = false;
            const int zs = zero_start && it == 0;
            cx<double> *up = it == sweeps - 1 ? vup : nullptr;
#define X(GG, RR)                                                                       \
    if (!done && G == GG && R == RR) {                                                  \
        if (form == 1) launch_full<T, GG, RR>(L, V, S, lu, ph, rf, active, st, zs, up);   \
        else if (form == 2) launch_staged<T, GG, RR, false>(L, V, S, lu, ph, rf, active, st, 0, nullptr); \
        else launch_phase<T, GG, RR>(L, V, S, lu, ph, rf, active, st);                  \
        done = true;                                                                    \
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
#define X(GG, RR)                                                                                \
    for (const void *f :                                                                         \
         {reinterpret_cast<const void *>(&smooth_staged_kernel<double, GG, RR, true>),           \
          reinterpret_cast<const void *>(&smooth_staged_kernel<float, GG, RR, true>),            \
          reinterpret_cast<const void *>(&smooth_staged_kernel<float, GG, RR, true, true>),      \
          reinterpret_cast<const void *>(&smooth_staged_kernel<double, GG, RR, false>),          \
          reinterpret_cast<const void *>(&smooth_staged_kernel<float, GG, RR, false>)}) {        \
        hipError_t r = allow_full_lds(f);                                                        \
        if (r != hipSuccess) e = r;                                                              \
    }
    TBL_LINE_SHAPES(X)
#undef X
    return e == hipSuccess ? TBL_OK : TBL_ERR_HIP;
}

template int phase_sweeps<double>(const LevelArgs<double> &, cx<double> *, const cx<double> *,
                                  const double *, int, double, const int *, int, hipStream_t, bool,
                                  cx<double> *);
template int phase_sweeps<float>(const LevelArgs<float> &, cx<float> *, const cx<float> *,
                                 const float *, int, float, const int *, int, hipStream_t, bool,
                                 cx<double> *);

}  // namespace tbl
