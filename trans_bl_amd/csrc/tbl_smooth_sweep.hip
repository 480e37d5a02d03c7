// One-launch line-relaxation sweep, out of place (sor, poisson_par.f:167-226,
// in the colour order of the other smoother forms).
//
// A sweep relaxes the two border columns (Fortran 2 and nx-1), then the
// interior colours i = 3, 6, ..., i = 4, 7, ..., i = 5, 8, ... (Fortran), each
// line's right-hand side taking the current values of its x-neighbours.
// The fourth-order x stencil reaches i +- 2, so a colour's lines read the
// new values of the earlier colours and the old values of the later ones;
// the dependency chain is only four phases deep.  Reading old values from
// Vin and writing every relaxed line to Vout (a second buffer of the same
// layout) removes every write-after-read hazard, so one workgroup can run
// all four phases of a strip of columns of one mode with no other
// workgroup's help: it owns columns [a, b) and recomputes the few lines
// outside them that its own lines read (colour 1 on [a-2, b+4), colour 2 on
// [a-1, b+2), colour 3 on [a, b)).  A recomputed line has the same inputs and
// code as in its owner's workgroup, so both store the same bits.  Between
// phases a workgroup barrier makes the new lines visible; within a phase
// each group of G lanes relaxes its lines one after the other with the
// coalesced, LDS-staged access of smooth_staged_kernel.
//
// Traffic per sweep: Vin and S read and Vout written about once (the
// re-reads of neighbour columns and of the new lines stay in the caches),
// against a full read of v per colour launch in the one-phase-per-launch
// form.  The boundary columns and the wall row are copied from Vin, so Vout
// is a complete field.
//
// ZS: the sweep starts from v = 0 whatever Vin holds (Vin is not read).
// UPD: each owned line's relaxed value is added to vup (the FP64 field of the
// same layout), the correction cycle's update v += e; the last colour's
// lines are then not stored in Vout (nothing reads them).
#include <algorithm>
#include <utility>

#include "tbl_relax.h"

namespace tbl {

constexpr int kSweepThreads = 256;
constexpr int kSweepXcds = 8;   // MI355X: 8 XCDs, one L2 each

// phase of column c: 0 border, 1..3 interior colours, -1 boundary
__device__ __forceinline__ int sweep_phase_of(int c, int nx)
{
    if (c <= 0 || c >= nx - 1) return -1;
    if (c == 1 || c == nx - 2) return 0;
    return (c - 2) % 3 + 1;
}

// wavefront-local ordering of LDS accesses (a group of G lanes never spans
// two wavefronts)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The global inputs of one line: its x-neighbour columns (border lines: the
// one-sided stencil's five columns on the inner side; interior: i-2, i-1,
// i+1, i+2), its own old values and its source, R rows per lane.
template <typename T, int R, bool UPD>
struct SweepIn {
    // rows 1 + lane + r G (the rows the lanes relax); v0: row 0's v (lane 0)
    cx<T> a0[R], a1[R], a2[R], a3[R], a4[R], vv[R], ss[R], v0;
    cx<double> up0[UPD ? R : 1];
};

// vr: one buffer resource over mode k's plane in both buffers of the pair;
// ib / ob: byte offsets of Vin's and Vout's plane in it.  Issues the loads of
// line i (the caller's next use of the registers waits for them).
template <typename T, int G, int R, bool ZS, bool UPD>
__device__ __forceinline__ void sweep_load(const LevelArgs<T> &L, __amdgpu_buffer_rsrc_t vr,
                                           unsigned ib, unsigned ob, __amdgpu_buffer_rsrc_t sr,
                                           const cx<double> *upp, int i, bool own, int lane,
                                           SweepIn<T, R, UPD> &in)
{
    const int ny = L.ny, ld = L.ld, nx = L.nx;
    const int ph = sweep_phase_of(i, nx);
    const int side = ph == 0 ? (i == 1 ? -1 : 1) : 0;
    const int sg = -side;
    const int d0 = side ? -sg : -2, d1 = side ? sg : -1, d2 = side ? 2 * sg : 1,
              d3 = side ? 3 * sg : 2, d4 = side ? 4 * sg : 2;
    constexpr unsigned ES = sizeof(cx<T>);
    const unsigned lo = static_cast<unsigned>((static_cast<size_t>(i) * ld + lane) * ES);
    // a neighbour relaxed in an earlier phase of this sweep is read from
    // Vout, any other (and the line itself) from Vin -- as zero with ZS
    auto off = [&](int d) -> unsigned {
        const int p = sweep_phase_of(i + d, nx);
        const bool fr = p >= 0 && p < ph;
        if (ZS && !fr) return kBufOff;
        return (fr ? ob : ib) + lo + static_cast<unsigned>(d * ld) * ES;
    };
    const unsigned b0 = off(d0), b1 = off(d1), b2 = off(d2), b3 = off(d3), b4 = off(d4);
    const unsigned bv = ZS ? kBufOff : ib + lo;
    if constexpr (UPD) {
        const size_t col = static_cast<size_t>(i) * ld;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = 1 + lane + r * G;
            if (own && j < ny) in.up0[r] = upp[col + j];
        }
    }
    // row 0 (lane 0's offset bv): the line's initial state, copied to Vout
    in.v0 = buf_load(vr, lane == 0 ? bv : kBufOff, T(0));
#pragma unroll
    for (int r = 0; r < R; ++r) {
        // rows past ny (up to G*R) read zero, not the next column
        const bool inr = 1 + lane + r * G < ny;
        const unsigned rs = (1 + r * G) * ES;
        auto at = [&](unsigned b) { return inr ? b + rs : kBufOff; };
        in.a0[r] = buf_load(vr, at(b0), T(0));
        in.a1[r] = buf_load(vr, at(b1), T(0));
        in.a2[r] = buf_load(vr, at(b2), T(0));
        in.a3[r] = buf_load(vr, at(b3), T(0));
        in.a4[r] = buf_load(vr, at(b4), T(0));
        in.vv[r] = buf_load(vr, at(bv), T(0));
        in.ss[r] = buf_load(sr, at(lo), T(0));
    }
}

// Relaxes line i from its loaded inputs: x terms, v and S staged in LDS
// (lane-strided rows to lane-contiguous segments), relax_core, and the
// relaxed line stored lane-strided into Vout (and added to vup).
template <typename T, int G, int R, bool UPD>
__device__ __forceinline__ void sweep_relax(const LevelArgs<T> &L, __amdgpu_buffer_rsrc_t lr,
                                            cx<T> *outp, cx<double> *upp, int k, int i, bool own,
                                            T rf, int lane, const SweepIn<T, R, UPD> &in,
                                            cx<T> *Xs, cx<T> *Vs, cx<T> *Ss)
{
    const int ny = L.ny, ld = L.ld, nx = L.nx;
    const int ph = sweep_phase_of(i, nx);
    const int side = ph == 0 ? (i == 1 ? -1 : 1) : 0;
    if (lane == 0) {
        Vs[0] = in.v0;
        Xs[0] = cx<T>{T(0), T(0)};   // read by relax_core, used by no row
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = 1 + lane + r * G;
        if (j < ny) {
            cx<T> x;
            // the sums of xoff in its order (multiply-adds in FP32)
            if (side)
                x = mac(T(-6), in.a3[r],
                        mac(T(14), in.a2[r], mac(T(-4), in.a1[r], T(10) * in.a0[r]))) +
                    in.a4[r];
            else if (j >= ny - 2)
                x = mac(T(16), in.a2[r] + in.a1[r], -(in.a3[r] + in.a0[r]));
            else
                x = mac(T(-1), in.a3[r],
                        mac(T(16), in.a2[r], mac(T(16), in.a1[r], T(-1) * in.a0[r])));
            Xs[j] = x;
            Vs[j] = in.vv[r];
            Ss[j] = in.ss[r];
        }
    }
    wave_lds_sync();
    const unsigned luo = static_cast<unsigned>((static_cast<size_t>(k) * 2 + (side != 0)) *
                                               kLuFields * L.lus * sizeof(T));
    const LineLU<T, R> f = load_line_lu_buf<T, R>(lr, luo, L.lus, line_j0(lane, R));
    relax_core<T, G, R>(L, f, rf, lane, [&](int j) { return Xs[j]; }, [&](int j) { return Vs[j]; },
                        [&](int j, int) { return Ss[j]; }, [&](int j, cx<T> v) { Vs[j] = v; });
    wave_lds_sync();
    const size_t col = static_cast<size_t>(i) * ld;
    cx<T> *o = outp + col;
    // row 0 (the wall value) is not relaxed: Vout takes Vin's
    if (lane == 0 && !(UPD && ph == 3)) o[0] = in.v0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = 1 + lane + r * G;
        if (j < ny) {
            const cx<T> e = Vs[j];
            if (!(UPD && ph == 3)) o[j] = e;
            if constexpr (UPD) {
                if (own)
                    upp[col + j] = cx<double>{in.up0[r].re + static_cast<double>(e.re),
                                              in.up0[r].im + static_cast<double>(e.im)};
            }
        }
    }
    // the LDS stage is rewritten by this group's next line
    wave_lds_sync();
}

// Vpair: the two buffers, half elements apart; in_half: which holds Vin
template <typename T, int G, int R, bool ZS, bool UPD>
__global__ void __launch_bounds__(kSweepThreads) smooth_sweep_kernel(
    LevelArgs<T> L, cx<T> *Vpair, size_t half, int in_half, const cx<T> *__restrict__ S,
    const T *__restrict__ lu, T rf, const int *active, int width, int nstrips, cx<double> *vup)
{
    extern __shared__ unsigned char smem_raw[];
    constexpr int NG = kSweepThreads / G;
    const int ld = L.ld, ny = L.ny, nx = L.nx;
    // a 64-lane group is one wavefront: its index is wavefront-uniform
    const int grp = G == 64 ? __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / 64)
                            : static_cast<int>(threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    cx<T> *Xs = reinterpret_cast<cx<T> *>(smem_raw) + static_cast<size_t>(grp) * 3 * ld;
    cx<T> *Vs = Xs + ld, *Ss = Vs + ld;
    // XCD-aware order: consecutive strips of a mode (which share halo
    // columns) on one XCD's L2 (the grid is padded to a multiple of 8)
    const int per = static_cast<int>(gridDim.x) / kSweepXcds;
    const int blk = static_cast<int>(blockIdx.x % kSweepXcds) * per +
                    static_cast<int>(blockIdx.x / kSweepXcds);
    const int k = blk / nstrips, s = blk - k * nstrips;
    if (k >= L.K) return;                     // uniform over the workgroup
    if (active && !active[k]) return;
    const int a = 1 + s * width, b = min(a + width, nx - 1);   // owned columns [a, b)
    constexpr unsigned ES = sizeof(cx<T>);
    const size_t poff = static_cast<size_t>(k) * L.plane;
    const size_t pbytes = L.plane * ES;
    const __amdgpu_buffer_rsrc_t vr = buf_rsrc(Vpair + poff, half * ES + pbytes);
    const unsigned ib = static_cast<unsigned>(in_half * half * ES);
    const unsigned ob = static_cast<unsigned>((1 - in_half) * half * ES);
    const __amdgpu_buffer_rsrc_t sr = buf_rsrc(S + poff, pbytes);
    const __amdgpu_buffer_rsrc_t lr =
        buf_rsrc(lu, static_cast<size_t>(L.K) * 2 * kLuFields * L.lus * sizeof(T));
    cx<T> *outp = Vpair + (1 - in_half) * half + poff;
    cx<double> *upp = UPD ? vup + poff : nullptr;
    // boundary columns (Dirichlet data): copied by the strips owning the
    // border columns next to them
    for (int e = 0; e < 2; ++e) {
        const int c = e == 0 ? 0 : nx - 1;
        const int nb = e == 0 ? 1 : nx - 2;
        if (nb < a || nb >= b) continue;
        for (int j = threadIdx.x; j < ny; j += kSweepThreads) {
            const unsigned o = ib + static_cast<unsigned>((static_cast<size_t>(c) * ld + j) * ES);
            outp[static_cast<size_t>(c) * ld + j] = buf_load(vr, ZS ? kBufOff : o, T(0));
        }
    }
    for (int ph = 0; ph < 4; ++ph) {
        int first = 0, step = 3, n = 0;
        if (ph == 0) {
            // border columns read by any line computed here (or owned)
            const bool left = a <= 4, right = b >= nx - 7;
            first = left ? 1 : nx - 2;
            step = nx - 3;
            n = (left ? 1 : 0) + (right ? 1 : 0);
        } else {
            const int lo = max(2, ph == 1 ? a - 2 : ph == 2 ? a - 1 : a);
            const int hi = min(nx - 2, ph == 1 ? b + 4 : ph == 2 ? b + 2 : b);
            const int res = ph == 1 ? 2 : ph == 2 ? 0 : 1;    // column mod 3
            first = lo + ((res - lo % 3) + 3) % 3;
            n = first < hi ? (hi - 1 - first) / 3 + 1 : 0;
        }
        // (loading the next line while relaxing this one was measured slower:
        // the extra registers cost more occupancy than the overlap gains)
        for (int t = grp; t < n; t += NG) {
            const int i = first + t * step;
            const bool own = i >= a && i < b;
            SweepIn<T, R, UPD> in;
            sweep_load<T, G, R, ZS, UPD>(L, vr, ib, ob, sr, upp, i, own, lane, in);
            sweep_relax<T, G, R, UPD>(L, lr, outp, upp, k, i, own, rf, lane, in, Xs, Vs, Ss);
        }
        // this phase's lines are visible to the workgroup's next phase
        __syncthreads();
    }
}

// Strip width (owned columns per workgroup): enough workgroups to give each
// CU several, wide enough that the recomputed lines (about 3 per strip)
// stay a small share.
int sweep_strip_width(int nx, int K, int wsel)
{
    const int cols = nx - 2;
    static const int kWidths[] = {0, 24, 36, 48, 72, 96};
    if (wsel > 0 && wsel < 6) return std::min(kWidths[wsel], cols);
    // default: about 1024 workgroups in total, width within [24, 96]
    const int target = 1024;
    int strips = std::max(1, target / std::max(1, K));
    int w = (cols + strips - 1) / strips;
    w = std::max(24, std::min(96, w));
    return std::min(w, cols);
}

size_t sweep_lds(int ld, size_t esize, int G)
{
    return static_cast<size_t>(kSweepThreads / G) * 3 * ld * esize;
}

template <typename T, int G, int R, bool ZS, bool UPD>
static void launch_sweep(const LevelArgs<T> &L, cx<T> *pair, size_t half, int in_half,
                         const cx<T> *S, const T *lu, T rf, const int *active, int width,
                         hipStream_t st, cx<double> *vup)
{
    const int nstrips = (L.nx - 2 + width - 1) / width;
    const size_t blocks = static_cast<size_t>(nstrips) * L.K;
    const size_t per = (blocks + kSweepXcds - 1) / kSweepXcds;
    hipLaunchKernelGGL((smooth_sweep_kernel<T, G, R, ZS, UPD>),
                       dim3(static_cast<unsigned>(per * kSweepXcds)), dim3(kSweepThreads),
                       sweep_lds(L.ld, sizeof(cx<T>), G), st, L, pair, half, in_half, S, lu, rf,
                       active, width, nstrips, vup);
}

// the four (ZS, UPD) instantiations; UPD exists for FP32 only
template <typename T, int G, int R>
static void launch_sweep_any(const LevelArgs<T> &L, cx<T> *pair, size_t half, int in_half,
                             const cx<T> *S, const T *lu, T rf, const int *active, int width,
                             hipStream_t st, bool zs, cx<double> *up)
{
    if constexpr (sizeof(T) == 4) {
        if (up) {
            if (zs)
                launch_sweep<T, G, R, true, true>(L, pair, half, in_half, S, lu, rf, active, width,
                                                  st, up);
            else
                launch_sweep<T, G, R, false, true>(L, pair, half, in_half, S, lu, rf, active,
                                                   width, st, up);
            return;
        }
    }
    if (zs)
        launch_sweep<T, G, R, true, false>(L, pair, half, in_half, S, lu, rf, active, width, st,
                                           nullptr);
    else
        launch_sweep<T, G, R, false, false>(L, pair, half, in_half, S, lu, rf, active, width, st,
                                            nullptr);
}

// `sweeps` sweeps alternating between Va and Vb, starting from Va (or from
// zero); *cur receives the buffer holding the result.  Va and Vb are the two
// halves of one allocation (either order), so that one buffer resource
// reaches both.  G: lanes per line; wsel: strip width selector (0 =
// default).
template <typename T>
int sweep_sweeps(const LevelArgs<T> &L, cx<T> *Va, cx<T> *Vb, const cx<T> *S, const T *lu,
                 int sweeps, T rf, const int *active, int G, int wsel, hipStream_t st,
                 bool zero_start, cx<double> *vup, cx<T> **cur)
{
    if (!Va || !Vb || Va == Vb) return TBL_ERR_ARG;
    cx<T> *pair = std::min(Va, Vb);
    const size_t half = static_cast<size_t>(std::max(Va, Vb) - pair);
    if (half < L.plane * L.K) return TBL_ERR_ARG;
    // 32-bit buffer offsets below the out-of-range offset 0x80000000
    if ((half + L.plane) * sizeof(cx<T>) >= (size_t{1} << 31)) return TBL_ERR_GRID;
    if (vup && sizeof(T) != 4) return TBL_ERR_ARG;   // the update variant is FP32 only
    const int R = line_rows_for(G, L.ny);
    if (R < 0) return TBL_ERR_GRID;
    if (sweep_lds(L.ld, sizeof(cx<T>), G) > 160 * 1024) return TBL_ERR_GRID;
    const int width = sweep_strip_width(L.nx, L.K, wsel);
    int in_half = Va == pair ? 0 : 1;
    for (int it = 0; it < sweeps; ++it) {
        const bool zs = zero_start && it == 0;
        cx<double> *up = it == sweeps - 1 ? vup : nullptr;
        bool done = false;
#define X(GG, RR)                                                                          \
    if (!done && G == GG && R == RR) {                                                     \
        launch_sweep_any<T, GG, RR>(L, pair, half, in_half, S, lu, rf, active, width, st, zs, up); \
        done = true;                                                                       \
    }
        TBL_LINE_SHAPES(X)
#undef X
        if (!done) return TBL_ERR_GRID;
        in_half = 1 - in_half;
    }
    *cur = pair + in_half * half;   // the last output
    return TBL_OK;
}

int sweep_kernel_attributes()
{
    hipError_t e = hipSuccess;
#define X(GG, RR)                                                                          \
    for (const void *f :                                                                   \
         {reinterpret_cast<const void *>(&smooth_sweep_kernel<double, GG, RR, false, false>), \
          reinterpret_cast<const void *>(&smooth_sweep_kernel<double, GG, RR, true, false>),  \
          reinterpret_cast<const void *>(&smooth_sweep_kernel<float, GG, RR, false, false>),  \
          reinterpret_cast<const void *>(&smooth_sweep_kernel<float, GG, RR, true, false>),   \
          reinterpret_cast<const void *>(&smooth_sweep_kernel<float, GG, RR, false, true>),   \
          reinterpret_cast<const void *>(&smooth_sweep_kernel<float, GG, RR, true, true>)}) { \
        hipError_t r = allow_full_lds(f);                                                  \
        if (r != hipSuccess) e = r;                                                        \
    }
    TBL_LINE_SHAPES(X)
#undef X
    return e == hipSuccess ? TBL_OK : TBL_ERR_HIP;
}

template int sweep_sweeps<double>(const LevelArgs<double> &, cx<double> *, cx<double> *,
                                  const cx<double> *, const double *, int, double, const int *,
                                  int, int, hipStream_t, bool, cx<double> *, cx<double> **);
template int sweep_sweeps<float>(const LevelArgs<float> &, cx<float> *, cx<float> *,
                                 const cx<float> *, const float *, int, float, const int *, int,
                                 int, hipStream_t, bool, cx<double> *, cx<float> **);

}  // namespace tbl
