// Multigrid transfer, defect, layout and bookkeeping kernels of the
// v-Poisson solve (gfx950), and the smoother dispatch.  The smoothers live in
// tbl_smooth_phase.hip, tbl_smooth_strip.hip and tbl_smooth_coarse.hip.
#include "tbl_rhs.h"
#include "tbl_xline.h"
#include "tbl_relax.h"

namespace tbl {

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
// the columns i-2..i+2 of its rows in registers and takes the y-neighbours
// from DPP lane shifts, so every value of v is loaded about once.
// The march goes in blocks of B columns: at a block's start the
// wavefront issues the loads of the next block's v and S columns, then
// computes and stores the block from registers, and only at its end takes
// the loaded columns over.  (The output stores share the load counter, so
// any wait for a load also waits for every store and load before it; one
// such wait per block instead of per column keeps a block's loads in flight
// behind its arithmetic.)  B trades that against registers: 8 for FP32, 2
// for FP64, whose arithmetic needs more wavefronts in flight to keep the
// vector units busy.  Loads go through buffer resources, zero where the
// stencil has no point.  Row forms: Fortran row 2 (sp_poi_coef), interior rows
// (cp_poi_coef), row ny-1 (pp_poi_coef) and row ny (lp_poi_coef, with the
// top boundary term); x part: the border-column one-sided stencils on
// columns 2 and nx-1, the centred fourth-order stencil elsewhere (grouped
// on the two top rows).  The divisions by 12 dx^2 fac_y and fac_y are
// multiplications by their reciprocals, and on interior rows the row's
// coefficients are folded into eight per-lane factors before the march
// (last-bit rounding differs from the reference's expression order).
constexpr int kDefRows = 60;
template <typename T>
constexpr int kDefBlock = sizeof(T) == 4 ? 8 : 2;


// columns per strip: longer strips load fewer halo columns; the mixed
// defect (FP64 in, FP32 out) was measured fastest at 64 on nx = 1025, the
// others at 32
template <typename T, typename TO>
__host__ __device__ inline int defect_strip(int nx)
{
    if (sizeof(T) == 8 && sizeof(TO) == 4 && nx >= 1024) return 64;
    return nx >= 512 ? 32 : 16;
}

// The restricting form (RS): the defect's full-weighting restriction onto
// the next coarser level's source instead of the defect itself.  Its
// wavefronts own kDefRowsRs rows with three halo rows each side, so the
// defect is formed on one halo row each side as well and every even fine
// row has its two neighbours' defects in the wavefront; each strip walks
// one column past its end for the same reason in x.
constexpr int kDefRowsRs = 58;
template <typename T>
struct CoarseOut {
    cx<T> *S;              // the coarser level's source (nullptr: plain defect)
    int nx, ny, ld;
    size_t plane;
};

template <typename T, typename TO, bool RS = false>
__global__ void __launch_bounds__(256) defect_march_kernel(LevelArgs<T> L,
                                                           const cx<T> *__restrict__ V,
                                                           const cx<T> *S, cx<TO> *out, T si,
                                                           const int *active, double *resmax,
                                                           int strip, int nstrips, int jchunks,
                                                           CoarseOut<TO> co)
{
    // fused multiply-adds in FP32 and in the mixed defect (FP64 in, FP32
    // out: no reference counterpart); the FP64 defect rounds as written
    constexpr bool kFuse = sizeof(T) == 4 || sizeof(TO) == 4;
    constexpr int kRows = RS ? kDefRowsRs : kDefRows, kHalo = RS ? 3 : 2;
    const int lane = threadIdx.x & 63;
    const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.y;
    if (item >= jchunks * nstrips) return;          // whole wavefront
    if (active && !active[k]) return;
    const int nx = L.nx, ny = L.ny, ld = L.ld;
    const int jc = item % jchunks, sc = item / jchunks;
    const int j = 1 + jc * kRows - kHalo + lane;
    const bool jin = j >= 0 && j < ny;
    const bool own = lane >= kHalo && lane < kHalo + kRows && j >= 1 && j < ny;
    // rows whose defect is formed: the owned ones, and with RS one more each side
    const bool calc = RS ? lane >= kHalo - 1 && lane <= kHalo + kRows && j >= 1 && j < ny : own;
    const int ia = 1 + sc * strip, ib0 = min(ia + strip, nx - 1);
    // RS: the strip's last even column needs the defect of the next column
    const int ib = RS && ib0 < nx - 1 ? ib0 + 1 : ib0;
    const size_t pbytes = L.plane * sizeof(cx<T>);
    const __amdgpu_buffer_rsrc_t vr = buf_rsrc(V + static_cast<size_t>(k) * L.plane, pbytes);
    const __amdgpu_buffer_rsrc_t sr = buf_rsrc(S + static_cast<size_t>(k) * L.plane, pbytes);
    auto vload = [&](int i) -> cx<T> {
        const bool ok = jin && i >= 0 && i < nx;
        return buf_load(vr, ok ? static_cast<unsigned>((i * ld + j) * sizeof(cx<T>)) : kBufOff, T(0));
    };
    auto sload = [&](int i) -> cx<T> {
        const bool ok = calc && i < nx;
        return buf_load(sr, ok ? static_cast<unsigned>((i * ld + j) * sizeof(cx<T>)) : kBufOff, T(0));
    };
    const T kk = static_cast<T>(k);
    const T k2b2 = -kk * kk * L.beta * L.beta;
    const T alpha2 = sqrt((L.alpha * L.alpha - k2b2) / L.fac_y);
    const T scx = T(12) * L.dx2 * L.fac_y;
    // reciprocals of the two per-level divisors (b/scx, d/fac_y below): the
    // per-point divisions were most of the kernel's VALU work
    const T iscx = T(1) / scx, ifac = T(1) / L.fac_y;
    T q = T(0);
    if (calc) q = (j == 1) ? L.qdy2[0] : (j == ny - 2) ? L.qdy2[j - 1] : L.qdy2[j - 2];
    // interior rows: the row's coefficients folded once --
    // lv = (cp0 xm + cp1 xs + cp2 xp) / scx + (cp3..cp7 . v) q
    //      + k2b2 (cp0 vm1 + cp1 w2 + cp2 vp1) / fac_y
    //    = ax . (xm, xs, xp) + cv . (vm2, vm1, w2, vp1, vp2)
    const T kf = k2b2 * ifac;
    const T ax0 = L.cp[0] * iscx, ax1 = L.cp[1] * iscx, ax2 = L.cp[2] * iscx;
    const T cv0 = L.cp[3] * q, cv1 = L.cp[4] * q + L.cp[0] * kf, cv2 = L.cp[5] * q + L.cp[1] * kf;
    const T cv3 = L.cp[6] * q + L.cp[2] * kf, cv4 = L.cp[7] * q;
    // the restricting form holds two more row sums: 4 columns per block keep
    // it at 96 registers, five wavefronts per SIMD (three at 8; measured
    // 7.26 -> 7.18 ms per 1025x193x64 mixed solve)
    constexpr int B = RS ? 4 : kDefBlock<T>;
    // win[t] = column i0 - 2 + t of the current block (i0 .. i0+B-1 computed),
    // sq[t] = source column i0 + t; nb / ns: the next block's new columns
    cx<T> win[B + 4], sq[B], nb[B], ns[B];
    // RS: the (1,2,1) row sums of the defect at the two previous columns
    cx<T> rc2 = {T(0), T(0)}, rc1 = {T(0), T(0)};
#pragma unroll
    for (int t = 0; t < B + 4; ++t) win[t] = vload(ia - 2 + t);
#pragma unroll
    for (int t = 0; t < B; ++t) sq[t] = sload(ia + t);
    double amax = 0.0;
    for (int i0 = ia; i0 < ib; i0 += B) {
        if (i0 + B < ib) {
#pragma unroll
            for (int t = 0; t < B; ++t) {
                nb[t] = vload(i0 + B + 2 + t);
                ns[t] = sload(i0 + B + t);
            }
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int i = i0 + u;
            if (i >= ib) break;
            const cx<T> w0 = win[u], w1 = win[u + 1], w2 = win[u + 2], w3 = win[u + 3],
                        w4 = win[u + 4];
            const cx<T> s_cur = sq[u];
            // x part at row j, shared with the y-neighbours (ungrouped); the
            // two top rows use their own grouped form (xo below)
            cx<T> xs;
            const bool border = i == 1 || i == nx - 2;
            if (i == 1) {
                xs = T(10) * w1 - T(15) * w2 - T(4) * w3 + T(14) * w4 - T(6) * vload(4) + vload(5);
            } else if (i == nx - 2) {
                xs = T(10) * w3 - T(15) * w2 - T(4) * w1 + T(14) * w0 - T(6) * vload(nx - 5)
                     + vload(nx - 6);
            } else {
                // -w0 + 16 w1 - 30 w2 + 16 w3 - w4 (fused where kFuse)
                xs = mac_f<kFuse>(T(-1), w4,
                                  mac_f<kFuse>(T(16), w3,
                                               mac_f<kFuse>(T(-30), w2,
                                                            mac_f<kFuse>(T(16), w1, T(-1) * w0))));
            }
            auto own_x = [&]() {
                // -(w0 + w4) + 16 (w1 + w3) - 30 w2
                return border ? xs
                              : mac_f<kFuse>(T(-30), w2,
                                             dot2_f<kFuse>(T(-1), w0 + w4, T(16), w1 + w3));
            };
            // y-neighbours from the adjacent lanes by DPP wavefront shifts
            // (no LDS round trip); the edge lanes 0, 1, 62, 63 are halo rows
            // whose results are not used
            const cx<T> z = {T(0), T(0)};
            const cx<T> xm = dpp<0x138>(z, xs), xp = dpp<0x130>(z, xs);
            const cx<T> vm1 = dpp<0x138>(z, w2), vm2 = dpp<0x138>(z, vm1);
            const cx<T> vp1 = dpp<0x130>(z, w2), vp2 = dpp<0x130>(z, vp1);
            cx<T> rv = {T(0), T(0)};   // the defect of this point (RS)
            if (calc) {
                cx<T> lv;
                if (j == 1) {
                    const cx<T> b = L.sp[2] * xp + L.sp[1] * xs;
                    const cx<T> c = L.sp[3] * vm1 + L.sp[4] * w2 + L.sp[5] * vp1 + L.sp[6] * vp2;
                    const cx<T> d = k2b2 * (L.sp[1] * w2 + L.sp[2] * vp1);
                    lv = b * iscx + c * q + d * ifac;
                } else if (j <= ny - 3) {
                    // ax0 xm + ax1 xs + ax2 xp + cv0 vm2 + ... + cv4 vp2, summed left to right
                    cx<T> t = dot2_f<kFuse>(ax0, xm, ax1, xs);
                    t = mac_f<kFuse>(ax2, xp, t);
                    t = mac_f<kFuse>(cv0, vm2, t);
                    t = mac_f<kFuse>(cv1, vm1, t);
                    t = mac_f<kFuse>(cv2, w2, t);
                    t = mac_f<kFuse>(cv3, vp1, t);
                    lv = mac_f<kFuse>(cv4, vp2, t);
                } else if (j == ny - 2) {
                    const cx<T> c = L.pp[3] * vp1 + L.pp[2] * w2 + L.pp[1] * vm1;
                    const cx<T> d = k2b2 * w2;
                    lv = own_x() * iscx + c * q + d * ifac;
                } else {
                    const cx<T> c = L.lp[2] * w2 + L.lp[3] * vm1 + L.lp[4] * vm2;
                    const cx<T> d = k2b2 * w2;
                    lv = own_x() * iscx + c * q + d * ifac - L.lp[1] * alpha2 * w2 * L.sqq_top;
                }
                const size_t o = static_cast<size_t>(k) * L.plane + static_cast<size_t>(i) * ld + j;
                const cx<T> r = mac_f<kFuse>(si, lv, s_cur);   // s + si lv
                rv = r;
                if (!RS) out[o] = cx<TO>{static_cast<TO>(r.re), static_cast<TO>(r.im)};
                if (!RS && resmax) {
                    // max of |r|^2; one square root per wavefront below
                    // (sqrt is monotone and correctly rounded: same maximum)
                    const double rr = static_cast<double>(r.re), ri = static_cast<double>(r.im);
                    amax = fmax(amax, rr * rr + ri * ri);
                }
            }
            if constexpr (RS) {
                // restrict_kernel's full weighting, separated: the row sum
                // (1,2,1) of this column (the top row alone), then on odd i
                // the column sum (1,2,1) of the three previous row sums
                // onto coarse column (i-1)/2 (rows of coarse index
                // 1..cny-2 weighted 1/16, the top row 1/4)
                const cx<T> rdn = dpp<0x138>(z, rv), rup = dpp<0x130>(z, rv);
                const bool top = j == ny - 1;
                const cx<T> rc = top ? rv : rdn + T(2) * rv + rup;
                const int ic = (i - 1) >> 1, jcr = j >> 1;
                if ((i & 1) && i >= ia + 2 && own && !(j & 1) && ic >= 1 && ic <= co.nx - 2 &&
                    jcr >= 1 && jcr <= co.ny - 1) {
                    const cx<T> t = rc2 + T(2) * rc1 + rc;
                    const cx<T> sv = (top ? T(0.25) : T(6.25e-2)) * t;
                    co.S[static_cast<size_t>(k) * co.plane + static_cast<size_t>(ic) * co.ld + jcr] =
                        cx<TO>{static_cast<TO>(sv.re), static_cast<TO>(sv.im)};
                }
                rc2 = rc1;
                rc1 = rc;
            }
        }
        // slide: the block's last four columns lead the next block's window
#pragma unroll
        for (int t = 0; t < 4; ++t) win[t] = win[B + t];
#pragma unroll
        for (int t = 0; t < B; ++t) {
            win[4 + t] = nb[t];
            sq[t] = ns[t];
        }
    }
    if (resmax) {
        for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_down(amax, o, 64));
        if (lane == 0) atomic_max_nonneg<T>(resmax + k, sqrt(amax));
    }
}

// ---------------------------------------------------------------------------
// restrict (poisson_par.f:796-874): inject v into the coarse level (and the
// FAS copy w), full-weight the fine residual into the coarse source.
// fas == 0: correction scheme (coarse v = 0, no copy); fas < 0: correction
// scheme with coarse v left alone (a zero-start smoother follows).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) restrict_kernel(Level F, Level Cl, int K,
                                                       const cx<T> *Vf, const cx<T> *Rf,
                                                       cx<T> *Vc, cx<T> *Wc, cx<T> *Sc,
                                                       int fas, const int *active)
{
    // grid (rows / 32, columns / 8, modes), block 32 x 8: no index division
    const int k = blockIdx.z;
    if (active && !active[k]) return;
    const int j = blockIdx.x * 32 + threadIdx.x, i = blockIdx.y * 8 + threadIdx.y;
    if (j >= Cl.ny || i >= Cl.nx) return;
    const int fi = 2 * i, fj = 2 * j;
    const cx<T> *vf = Vf + static_cast<size_t>(k) * F.plane;
    const cx<T> *rf = Rf + static_cast<size_t>(k) * F.plane;
    const size_t co = static_cast<size_t>(k) * Cl.plane + static_cast<size_t>(i) * Cl.ld + j;
    if (fas > 0) {
        const cx<T> v = vf[static_cast<size_t>(fi) * F.ld + fj];
        Vc[co] = v;
        Wc[co] = v;
    } else if (fas == 0) {
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
// correction scheme (fas == 0) it is v_c.  fas == 2 assigns the bilinear
// interpolant of v_c to the fine interior (rows 2..ny, Fortran) instead of
// adding it.
// two adjacent complex values (an even row and the next), aligned as a pair
template <typename T>
struct alignas(4 * sizeof(T)) cx_pair {
    cx<T> a, b;
};

// A thread handles the even fine row j0 = 2m and the odd row j0 + 1 of one
// column: both interpolate from coarse rows m and m + 1, and the pair is one
// aligned load and store (ld is a multiple of 8, so j0 starts a pair).
template <typename T>
__global__ void __launch_bounds__(256) prolong_kernel(Level F, Level Cl, int K, cx<T> *Vf,
                                                      const cx<T> *Vc, const cx<T> *Wc,
                                                      int fas, const int *active)
{
    // grid (row pairs / 32, columns / 8, modes), block 32 x 8: no index division
    const int k = blockIdx.z;
    if (active && !active[k]) return;
    const int m = blockIdx.x * 32 + threadIdx.x, i = blockIdx.y * 8 + threadIdx.y;
    const int j0 = 2 * m;
    if (j0 >= F.ny || i < 1 || i > F.nx - 2) return;
    const cx<T> *vc = Vc + static_cast<size_t>(k) * Cl.plane;
    const cx<T> *wc = Wc + static_cast<size_t>(k) * Cl.plane;
    auto E = [&](int ci, int cj) {
        const size_t o = static_cast<size_t>(ci) * Cl.ld + cj;
        if (fas == 1 && ci >= 1 && ci <= Cl.nx - 2) return vc[o] - wc[o];
        return vc[o];
    };
    auto row = [&](int cj) {   // value on an even (0-based) fine row
        if ((i & 1) == 0) return E(i >> 1, cj);
        return T(0.5) * (E(i >> 1, cj) + E((i >> 1) + 1, cj));
    };
    const bool has1 = j0 + 1 < F.ny;
    const cx<T> t0 = row(m);
    const size_t o = static_cast<size_t>(k) * F.plane + static_cast<size_t>(i) * F.ld + j0;
    if (!has1) {             // last row of an odd ny: even, alone
        if (fas == 2) {
            if (j0 >= 1) Vf[o] = t0;
        } else {
            Vf[o] = Vf[o] + t0;
        }
        return;
    }
    const cx<T> t1 = T(0.5) * (row(m + 1) + t0);
    cx_pair<T> *pv = reinterpret_cast<cx_pair<T> *>(Vf + o);
    if (fas == 2) {
        // assign the interpolant, wall row (j = 0) untouched
        if (j0 >= 1) *pv = cx_pair<T>{t0, t1};
        else Vf[o + 1] = t1;
    } else {
        const cx_pair<T> v = *pv;
        *pv = cx_pair<T>{v.a + t0, v.b + t1};
    }
}

// ---------------------------------------------------------------------------
// user layout (x fastest) <-> internal layout (y fastest) through LDS tiles
// ---------------------------------------------------------------------------

// function_source (poisson_par.f:111-164) with the layout change: g =
// kb*wx - dwzdx (or f itself) of rows j0-1 .. j0+32 of a 32 x 32 tile goes
// through LDS, and the tile's S = B g (row weights sp2/sp3 at Fortran j = 2,
// cp1..3 on the centre rows, 1 at ny-1, 0 at 1 and ny) and v = uy are
// written in the internal layout; g itself never reaches memory
__global__ void __launch_bounds__(256) gather_source_kernel(LevelArgs<double> L,
                                                            const cx<double> *wx,
                                                            const cx<double> *dz,
                                                            const cx<double> *uy, double beta,
                                                            cx<double> *S, cx<double> *Vout,
                                                            int direct_rhs)
{
    __shared__ cx<double> tg[34][33], tv[32][33];
    const int nx = L.nx, ny = L.ny;
    const int k = blockIdx.z;
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;   // 32 x 8
    const double kb = -static_cast<double>(k) * beta;           // kb = (0, kb)
    const size_t ubase = static_cast<size_t>(k) * nx * ny;
    for (int r = ty; r < 34; r += 8) {
        const int i = i0 + tx, j = j0 - 1 + r;
        cx<double> g = {0, 0};
        if (i < nx && j >= 0 && j < ny) {
            const size_t u = ubase + static_cast<size_t>(j) * nx + i;
            const cx<double> d = dz ? dz[u] : cx<double>{0, 0};
            if (direct_rhs) {
                g = d;
            } else {
                cx<double> kw = {0, 0};
                if (wx) { const cx<double> w = wx[u]; kw = cmul(cx<double>{0.0, kb}, w); }
                g = kw - d;
            }
        }
        tg[r][tx] = g;
    }
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + tx, j = j0 + r;
        cx<double> v = {0, 0};
        if (uy && i < nx && j < ny) v = uy[ubase + static_cast<size_t>(j) * nx + i];
        tv[r][tx] = v;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        if (i < nx && j < ny) {
            // g[j + d] = tg[tx + 1 + d][r]
            cx<double> s = {0, 0};
            if (j == 1) s = L.sp[1] * tg[tx + 1][r] + L.sp[2] * tg[tx + 2][r];
            else if (j >= 2 && j <= ny - 3)
                s = L.cp[0] * tg[tx][r] + L.cp[1] * tg[tx + 1][r] + L.cp[2] * tg[tx + 2][r];
            else if (j == ny - 2) s = tg[tx + 1][r];
            const size_t o = static_cast<size_t>(k) * L.plane + static_cast<size_t>(i) * L.ld + j;
            S[o] = s;
            Vout[o] = tv[tx][r];
        }
    }
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
    // grid (rows / 32, columns / 8, modes), block 32 x 8: no index division
    // (one point per thread: measured faster here than row pairs)
    const int k = blockIdx.z;
    if (active && !active[k]) return;
    const int j = blockIdx.x * 32 + threadIdx.x, i = blockIdx.y * 8 + threadIdx.y;
    if (i < 1 || i > F.nx - 2 || j < 1 || j >= F.ny) return;
    const size_t idx = static_cast<size_t>(k) * F.plane + static_cast<size_t>(i) * F.ld + j;
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

static const int kRs[] = {1, 2, 3, 4, 5, 6, 7, 9, 13, 17};

// grid of the point-wise level kernels: 32 rows x 8 columns per block
static dim3 grid_2d(const Level &l, int K)
{
    return dim3(static_cast<unsigned>((l.ny + 31) / 32), static_cast<unsigned>((l.nx + 7) / 8),
                static_cast<unsigned>(K));
}

// Raise the dynamic-LDS limit of every LDS-resident smoother instantiation
// once, before any stream capture can happen.
int set_kernel_attributes()
{
    int rc = phase_kernel_attributes();
    if (rc == TBL_OK) rc = strip_kernel_attributes();
    if (rc == TBL_OK) rc = coarse_kernel_attributes();
    if (rc == TBL_OK) rc = sweep_kernel_attributes();
    if (rc == TBL_OK) rc = deriv_kernel_attributes();
    if (rc == TBL_OK) rc = xline_kernel_attributes();
    if (rc == TBL_OK) rc = rhs_kernel_attributes();
    return rc;
}

int line_rows_for(int G, int ny)
{
    // rows 1..ny-1 (row 0 is not relaxed)
    const int need = (ny - 1 + G - 1) / G;
    for (int r : kRs) {
        if (r < need) continue;
        if (G == 32 && r > 9) return -1;
        return r;
    }
    return -1;
}

int lu_stride(int ny)
{
    int s = ny;
    for (int G : {16, 32, 64}) {
        const int R = line_rows_for(G, ny);
        if (R > 0 && 1 + G * R > s) s = 1 + G * R;
    }
    return (s + 7) / 8 * 8;
}

// G > 0: lanes per line with direct (lane-contiguous) global access;
// G < 0: |G| lanes per line with the LDS-staged, coalesced kernel;
// kStripShape < G < kXStagedShape: fused-sweep strips with G - kStripShape
// lanes per line; G > kXStagedShape: G - kXStagedShape lanes per line, only
// the x terms staged in LDS
template <typename T>
int smooth_sweeps(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps, T rf,
                  const int *active, int G, cx<T> *halo, hipStream_t st, bool zero_start,
                  cx<double> *vup)
{
    if ((zero_start || vup) && G >= 0) return TBL_ERR_ARG;
    if (G > kStripShape && G < kXStagedShape)
        return strip_sweeps<T>(L, V, S, lu, sweeps, rf, active, G - kStripShape, halo, st);
    return phase_sweeps<T>(L, V, S, lu, sweeps, rf, active, G, st, zero_start, vup);
}

template <typename T, typename TO>
void defect(const LevelArgs<T> &L, const cx<T> *V, const cx<T> *S, cx<TO> *out, T si,
            const int *active, double *resmax, hipStream_t st)
{
    const int strip = defect_strip<T, TO>(L.nx);
    const int nstrips = (L.nx - 2 + strip - 1) / strip;
    const int jchunks = (L.ny - 1 + kDefRows - 1) / kDefRows;
    const int blocks = (nstrips * jchunks + 3) / 4;
    hipLaunchKernelGGL((defect_march_kernel<T, TO>), dim3(blocks, L.K), dim3(256), 0, st, L, V, S,
                       out, si, active, resmax, strip, nstrips, jchunks, CoarseOut<TO>{});
}

template <typename T>
void defect_restrict(const LevelArgs<T> &L, const Level &Cl, const cx<T> *V, const cx<T> *S,
                     cx<T> *Sc, const int *active, hipStream_t st)
{
    const int strip = defect_strip<T, T>(L.nx);   // even: strips start on odd columns
    const int nstrips = (L.nx - 2 + strip - 1) / strip;
    const int jchunks = (L.ny - 1 + kDefRowsRs - 1) / kDefRowsRs;
    const int blocks = (nstrips * jchunks + 3) / 4;
    const CoarseOut<T> co{Sc, Cl.nx, Cl.ny, Cl.ld, Cl.plane};
    hipLaunchKernelGGL((defect_march_kernel<T, T, true>), dim3(blocks, L.K), dim3(256), 0, st, L,
                       V, S, static_cast<cx<T> *>(nullptr), T(-1), active,
                       static_cast<double *>(nullptr), strip, nstrips, jchunks, co);
}

template <typename T>
void restrict_level(const Level &F, const Level &Cl, int K, const cx<T> *Vf, const cx<T> *Rf,
                    cx<T> *Vc, cx<T> *Wc, cx<T> *Sc, int fas, const int *active, hipStream_t st)
{
    hipLaunchKernelGGL((restrict_kernel<T>), grid_2d(Cl, K), dim3(32, 8), 0, st, F, Cl, K, Vf, Rf,
                       Vc, Wc, Sc, fas, active);
}

template <typename T>
void prolong_level(const Level &F, const Level &Cl, int K, cx<T> *Vf, const cx<T> *Vc,
                   const cx<T> *Wc, int fas, const int *active, hipStream_t st)
{
    const dim3 grid(static_cast<unsigned>((F.ny + 63) / 64), static_cast<unsigned>((F.nx + 7) / 8),
                    static_cast<unsigned>(K));
    hipLaunchKernelGGL((prolong_kernel<T>), grid, dim3(32, 8), 0, st, F, Cl, K, Vf, Vc, Wc, fas,
                       active);
}


void gather_source(const LevelArgs<double> &L, const double *wx, const double *dz,
                   const double *uy, double beta, cx<double> *S, cx<double> *V, int direct_rhs,
                   hipStream_t st)
{
    dim3 grid((L.nx + 31) / 32, (L.ny + 31) / 32, L.K);
    hipLaunchKernelGGL(gather_source_kernel, grid, dim3(256), 0, st, L,
                       reinterpret_cast<const cx<double> *>(wx),
                       reinterpret_cast<const cx<double> *>(dz),
                       reinterpret_cast<const cx<double> *>(uy), beta, S, V, direct_rhs);
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
    hipLaunchKernelGGL(update_kernel, grid_2d(F, K), dim3(32, 8), 0, st, F, K, V, E, active);
}

template <typename TA, typename TB>
void convert(size_t n, const cx<TA> *a, cx<TB> *b, int zero, hipStream_t st)
{
    hipLaunchKernelGGL((convert_kernel<TA, TB>), dim3(nblk(n)), dim3(256), 0, st, n, a, b, zero);
}

// explicit instantiations
template int smooth_sweeps<double>(const LevelArgs<double> &, cx<double> *, const cx<double> *,
                                   const double *, int, double, const int *, int, cx<double> *,
                                   hipStream_t, bool, cx<double> *);
template int smooth_sweeps<float>(const LevelArgs<float> &, cx<float> *, const cx<float> *,
                                  const float *, int, float, const int *, int, cx<float> *,
                                  hipStream_t, bool, cx<double> *);
template void defect<double, double>(const LevelArgs<double> &, const cx<double> *,
                                     const cx<double> *, cx<double> *, double, const int *,
                                     double *, hipStream_t);
template void defect<double, float>(const LevelArgs<double> &, const cx<double> *,
                                    const cx<double> *, cx<float> *, double, const int *, double *,
                                    hipStream_t);
template void defect<float, float>(const LevelArgs<float> &, const cx<float> *, const cx<float> *,
                                   cx<float> *, float, const int *, double *, hipStream_t);
template void defect_restrict<float>(const LevelArgs<float> &, const Level &, const cx<float> *,
                                     const cx<float> *, cx<float> *, const int *, hipStream_t);
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
