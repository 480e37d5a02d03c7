// Context creation, the batched FAS / mixed-precision V-cycle drivers and the
// C ABI of trans_bl_amd.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "tbl_ctx.h"

using namespace tbl;

namespace {
thread_local std::string g_last_error;
}

void tbl_set_last_error(const char *msg, const char *file, int line)
{
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s (%s:%d)", msg, file, line);
    g_last_error = buf;
}

extern "C" const char *tbl_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *tbl_version(void) { return "trans_bl_amd 0.1 gfx950"; }

namespace {

template <typename T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return TBL_OK;
    TBL_HIP(hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)));
    TBL_HIP(hipMemset(*p, 0, n * sizeof(T)));
    return TBL_OK;
}

#define TRY(x)                     \
    do {                           \
        int rc_ = (x);             \
        if (rc_ != TBL_OK) return rc_; \
    } while (0)

int default_G(int ny)
{
    if (ny >= 160) return 64;
    if (ny >= 48) return 32;
    return 16;
}

template <typename T>
LevelArgs<T> level_args(const tbl_ctx &c, int l, const T *qdy2)
{
    LevelArgs<T> A{};
    const Level &lv = c.lv[l];
    A.nx = lv.nx;
    A.ny = lv.ny;
    A.ld = lv.ld;
    A.lus = lu_stride(lv.ny);
    A.K = c.K;
    A.plane = lv.plane;
    for (int i = 0; i < 7; ++i) A.sp[i] = static_cast<T>(c.c.sp_poi[l][i]);
    for (int i = 0; i < 8; ++i) A.cp[i] = static_cast<T>(c.c.cp_poi[l][i]);
    for (int i = 0; i < 4; ++i) A.pp[i] = static_cast<T>(c.c.pp_poi[l][i]);
    for (int i = 0; i < 5; ++i) A.lp[i] = static_cast<T>(c.c.lp_poi[l][i]);
    A.qx = static_cast<T>(1.0 / (12.0 * c.m.dx2[l] * c.g.fac_y));
    A.dx2 = static_cast<T>(c.m.dx2[l]);
    A.fac_y = static_cast<T>(c.g.fac_y);
    A.alpha = static_cast<T>(c.g.alpha);
    A.beta = static_cast<T>(c.g.beta);
    A.sqq_top = static_cast<T>(std::sqrt(c.m.qdy2[static_cast<size_t>(l) * c.g.ny + (lv.ny - 3)]));
    A.qdy2 = qdy2;
    return A;
}

// Key of the line shapes of all levels (a captured V-cycle bakes them in).
int shapes_key(const std::vector<int> &G)
{
    unsigned h = 2166136261u;
    for (int g : G) h = (h ^ static_cast<unsigned>(g + 4096)) * 16777619u;
    return static_cast<int>(h & 0x7fffffff);
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static int create_impl(const tbl_grid *grid, const tbl_coefs *coefs, tbl_ctx *c)
{
    static std::once_flag attr_once;
    static int attr_rc = TBL_OK;
    std::call_once(attr_once, [] { attr_rc = set_kernel_attributes(); });
    if (attr_rc != TBL_OK) return attr_rc;
    c->g = *grid;
    TRY(check_grid(c->g));
    if (coefs) c->c = *coefs;
    else if (!generate_coefs(c->g.stf, c->g.levels, &c->c)) return TBL_ERR_ARG;
    build_metrics(c->g, c->m);
    TBL_HIP(hipGetDevice(&c->device));
    // the device address of the allocation-free entry point's array
    TRY(noalloc_setup(c->device));
    c->K = c->g.nmodes;
    c->L = c->g.levels;
    const int K = c->K, L = c->L;
    c->lv.resize(L);
    for (int l = 0; l < L; ++l) c->lv[l] = make_level(c->m.nx[l], c->m.ny[l], c->m.dx2[l]);
    c->V.assign(L, nullptr); c->S.assign(L, nullptr); c->W.assign(L, nullptr); c->R.assign(L, nullptr);
    c->E32.assign(L, nullptr); c->S32.assign(L, nullptr); c->R32.assign(L, nullptr);
    c->E32b.assign(L, nullptr);
    c->qdy2_d.assign(L, nullptr); c->qdy2_f.assign(L, nullptr);
    c->lu_d.assign(L, nullptr); c->lu_f.assign(L, nullptr);
    c->G64.assign(L, 0); c->G32.assign(L, 0);
    c->H64.assign(L, nullptr); c->H32.assign(L, nullptr);
    for (int l = 0; l < L; ++l) {
        const Level &lv = c->lv[l];
        const size_t n = lv.plane * K;
        TRY(dalloc(&c->V[l], n));
        TRY(dalloc(&c->S[l], n));
        TRY(dalloc(&c->W[l], n));
        TRY(dalloc(&c->R[l], n));
        // (E32b: the second half of E32's allocation, for the out-of-place
        // sweep's single buffer resource over both)
        TRY(dalloc(&c->E32[l], l < L - 1 ? 2 * n : n));
        if (l < L - 1) c->E32b[l] = c->E32[l] + n;
        TRY(dalloc(&c->S32[l], n));
        TRY(dalloc(&c->R32[l], n));
        // metric v_qdy2(:, l)
        std::vector<double> q(lv.ld, 0.0);
        std::vector<float> qf(lv.ld, 0.0f);
        for (int j = 0; j < lv.ny - 1; ++j) {
            q[j] = c->m.qdy2[static_cast<size_t>(l) * c->g.ny + j];
            qf[j] = static_cast<float>(q[j]);
        }
        TRY(dalloc(&c->qdy2_d[l], lv.ld));
        TRY(dalloc(&c->qdy2_f[l], lv.ld));
        TBL_HIP(hipMemcpy(c->qdy2_d[l], q.data(), sizeof(double) * lv.ld, hipMemcpyHostToDevice));
        TBL_HIP(hipMemcpy(c->qdy2_f[l], qf.data(), sizeof(float) * lv.ld, hipMemcpyHostToDevice));
        // line LU factors per (mode, centre/border)
        // (row stride lu_stride(ny); rows past ny hold the identity)
        const int d = lu_stride(lv.ny);
        const size_t per = static_cast<size_t>(kLuFields) * d;
        std::vector<double> lu(per * 2 * K, 0.0);
        std::vector<float> luf(lu.size(), 0.0f);
        for (int k = 0; k < K; ++k)
            for (int t = 0; t < 2; ++t) {
                double *b = &lu[(static_cast<size_t>(k) * 2 + t) * per];
                line_lu(c->g, c->c, c->m, l + 1, k, t == 1, b, b + d, b + 2 * d, b + 3 * d,
                        b + 4 * d);
                for (int j = 0; j < lv.ny; ++j) {
                    const double p = b[4 * d + j];
                    b[5 * d + j] = -b[2 * d + j] / p;
                    b[6 * d + j] = -b[3 * d + j] / p;
                    b[7 * d + j] = 1.0 / p;
                }
                for (int j = lv.ny; j < d; ++j) {
                    for (int f = 0; f < kLuFields; ++f) b[f * d + j] = 0.0;
                    b[4 * d + j] = 1.0;
                    b[7 * d + j] = 1.0;
                }
            }
        for (size_t e = 0; e < lu.size(); ++e) luf[e] = static_cast<float>(lu[e]);
        TRY(dalloc(&c->lu_d[l], lu.size()));
        TRY(dalloc(&c->lu_f[l], luf.size()));
        TBL_HIP(hipMemcpy(c->lu_d[l], lu.data(), sizeof(double) * lu.size(), hipMemcpyHostToDevice));
        TBL_HIP(hipMemcpy(c->lu_f[l], luf.data(), sizeof(float) * luf.size(), hipMemcpyHostToDevice));
        c->G64[l] = default_G(lv.ny);
        c->G32[l] = default_G(lv.ny);
        // snapshot buffers for the fused-sweep strip smoother (not on the coarsest level)
        if (l < L - 1) {
            const size_t h64 = strip_halo_elems(strip_geom(lv.nx, lv.ny, K, sizeof(double)), lv.ld, K);
            const size_t h32 = strip_halo_elems(strip_geom(lv.nx, lv.ny, K, sizeof(float)), lv.ld, K);
            if (h64) TRY(dalloc(&c->H64[l], h64));
            if (h32) TRY(dalloc(&c->H32[l], h32));
        }
    }
    // the coarsest level starts on its sequential kernels
    c->G64[L - 1] = c->G32[L - 1] = kCoarseSeqShape;
    c->la_d.resize(L);
    c->la_f.resize(L);
    for (int l = 0; l < L; ++l) {
        c->la_d[l] = level_args<double>(*c, l, c->qdy2_d[l]);
        c->la_f[l] = level_args<float>(*c, l, c->qdy2_f[l]);
    }
    TRY(dalloc(&c->d_active, K));
    TRY(dalloc(&c->d_cycles, K));
    TRY(dalloc(&c->d_nactive, 1));
    TRY(dalloc(&c->d_cycle, 1));
    TRY(dalloc(&c->d_resmax, K));
    TRY(dalloc(&c->d_lastres, K));
    TBL_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->h_nactive), sizeof(int), hipHostMallocDefault));

    // derivatives: metrics and Thomas factors of the five schemes
    const int nmax = std::max(c->g.nx, c->g.ny);
    std::vector<double> th(static_cast<size_t>(5) * kThomasFields * nmax, 0.0);
    for (int kind = 0; kind < 5; ++kind) {
        std::vector<double> a, b, cc, gam, bet;
        deriv_lhs(c->g, c->c, kind, a, b, cc);
        thomas_factors(a, b, cc, gam, bet);
        double *t = &th[static_cast<size_t>(kind) * kThomasFields * nmax];
        for (size_t j = 0; j < a.size(); ++j) {
            t[j] = a[j];
            t[nmax + j] = bet[j];
            t[2 * nmax + j] = gam[j];
            t[3 * nmax + j] = 1.0 / bet[j];
        }
    }
    // lane-segment copies for the LDS-staged kernels
    int xR = 0, yG = 0, yR = 0;
    deriv_shapes(c->g.nx, c->g.ny, &xR, &yG, &yR);
    std::vector<double> ts;
    for (int kind = 0; kind < 5; ++kind) {
        const bool x = kind <= 1;
        const int G = x ? 64 : yG, R = x ? xR : yR, n = x ? c->g.nx : c->g.ny;
        const double *t = &th[static_cast<size_t>(kind) * kThomasFields * nmax];
        const size_t base = ts.size(), per = static_cast<size_t>(G) * R;
        ts.resize(base + 3 * per);
        for (int r = 0; r < R; ++r)
            for (int lane = 0; lane < G; ++lane) {
                const int i = lane * R + r;
                const size_t e = base + static_cast<size_t>(r) * G + lane;
                ts[e] = i < n ? t[i] : 0.0;                            // a
                ts[e + per] = i < n ? t[3 * nmax + i] : 1.0;           // 1/bet
                ts[e + 2 * per] = i + 1 < n ? t[2 * nmax + i + 1] : 0.0;   // gam of row i+1
            }
    }
    if (!ts.empty()) {
        TRY(dalloc(&c->d_tseg, ts.size()));
        TBL_HIP(hipMemcpy(c->d_tseg, ts.data(), sizeof(double) * ts.size(), hipMemcpyHostToDevice));
    }
    TRY(dalloc(&c->d_thomas, th.size()));
    TBL_HIP(hipMemcpy(c->d_thomas, th.data(), sizeof(double) * th.size(), hipMemcpyHostToDevice));
    TRY(dalloc(&c->d_qdy, c->g.ny));
    TBL_HIP(hipMemcpy(c->d_qdy, c->m.qdy.data(), sizeof(double) * c->g.ny, hipMemcpyHostToDevice));
    DerivArgs &A = c->dargs;
    A.nx = c->g.nx;
    A.ny = c->g.ny;
    A.K = K;
    A.dx = c->g.dx;
    A.dxx = c->g.dx * c->g.dx;
    std::memcpy(A.fp_e, c->c.fp_fd_e, sizeof(A.fp_e));
    std::memcpy(A.sp_fd, c->c.sp_fd, sizeof(A.sp_fd));
    std::memcpy(A.cp_fd, c->c.cp_fd, sizeof(A.cp_fd));
    std::memcpy(A.pp_fd, c->c.pp_fd, sizeof(A.pp_fd));
    std::memcpy(A.lp_fd, c->c.lp_fd, sizeof(A.lp_fd));
    std::memcpy(A.fp_sd, c->c.fp_sd, sizeof(A.fp_sd));
    std::memcpy(A.sp_sd, c->c.sp_sd, sizeof(A.sp_sd));
    std::memcpy(A.cp_sd, c->c.cp_sd, sizeof(A.cp_sd));
    std::memcpy(A.pp_sd, c->c.pp_sd, sizeof(A.pp_sd));
    std::memcpy(A.lp_sd, c->c.lp_sd, sizeof(A.lp_sd));
    A.qdy = c->d_qdy;
    A.qdy2 = c->qdy2_d[0];
    A.thomas = c->d_thomas;
    A.nmax = nmax;
    A.tseg = c->d_tseg;
    A.xR = xR;
    A.yG = yG;
    A.yR = yR;
    TBL_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamDefault));
    c->stream = c->own_stream;
    TBL_HIP(hipEventCreate(&c->ev0));
    TBL_HIP(hipEventCreate(&c->ev1));
    TBL_HIP(hipDeviceSynchronize());
    return TBL_OK;
}

extern "C" int tbl_coefs_generate(double stf, int levels, tbl_coefs *out)
{
    if (!out || levels < 1 || levels > TBL_MAX_LEVELS || !(stf > 0)) return TBL_ERR_ARG;
    return generate_coefs(stf, levels, out) ? TBL_OK : TBL_ERR_ARG;
}

extern "C" int tbl_coefs_read(const char *path, int levels, tbl_coefs *out)
{
    if (!path || !out || levels < 1 || levels > TBL_MAX_LEVELS) return TBL_ERR_ARG;
    std::string err;
    const int rc = read_coefs_bin(path, levels, out, &err);
    if (rc != TBL_OK) tbl_set_last_error(err.c_str(), __FILE__, __LINE__);
    return rc;
}

extern "C" int tbl_destroy(tbl_ctx *c)
{
    if (!c) return TBL_ERR_ARG;
    guess_free(c);
    auto fr = [](void *p) { if (p) (void)hipFree(p); };
    for (auto *v : {&c->V, &c->S, &c->W, &c->R})
        for (auto *p : *v) fr(p);
    for (auto *v : {&c->E32, &c->S32, &c->R32})
        for (auto *p : *v) fr(p);
    for (auto *p : c->qdy2_d) fr(p);
    for (auto *p : c->qdy2_f) fr(p);
    for (auto *p : c->lu_d) fr(p);
    for (auto *p : c->lu_f) fr(p);
    for (auto *p : c->H64) fr(p);
    for (auto *p : c->H32) fr(p);
    fr(c->d_active); fr(c->d_cycles); fr(c->d_nactive); fr(c->d_cycle);
    fr(c->d_resmax); fr(c->d_lastres); fr(c->d_qdy); fr(c->d_thomas); fr(c->d_tseg);
    fr(c->d_u0); fr(c->d_u1); fr(c->d_u2);
    fr(c->d_xpenta); fr(c->d_xinteg); fr(c->d_xfilter);
    fr(c->d_t1); fr(c->d_t2); fr(c->d_lapv); fr(c->d_dwydy); fr(c->d_d2wy);
    fr(c->d_rhs); fr(c->d_rk); fr(c->d_tw);
    if (c->h_nactive) (void)hipHostFree(c->h_nactive);
    for (auto &gp : c->graph)
        for (auto &g : gp)
            if (g) (void)hipGraphExecDestroy(g);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    delete c;
    return TBL_OK;
}

extern "C" int tbl_create(const tbl_grid *grid, const tbl_coefs *coefs, tbl_ctx **out)
{
    if (!grid || !out) return TBL_ERR_ARG;
    *out = nullptr;
    tbl_ctx *c = new (std::nothrow) tbl_ctx();
    if (!c) return TBL_ERR_ALLOC;
    int rc = create_impl(grid, coefs, c);
    if (rc != TBL_OK) {
        tbl_destroy(c);
        return rc;
    }
    *out = c;
    return TBL_OK;
}

extern "C" int tbl_synchronize(tbl_ctx *c)
{
    if (!c) return TBL_ERR_ARG;
    TBL_HIP(hipStreamSynchronize(c->stream));
    return TBL_OK;
}

extern "C" int tbl_set_stream(tbl_ctx *c, void *s)
{
    if (!c) return TBL_ERR_ARG;
    c->stream = s ? static_cast<hipStream_t>(s) : c->own_stream;
    return TBL_OK;
}

// ---------------------------------------------------------------------------
// V-cycle drivers
// ---------------------------------------------------------------------------
namespace {

struct Sched {
    int maxc, nu1, nu2, nc;
};

// Schedule of the FP32 correction cycle when the caller leaves nu1 and
// ncoarse at 0.  The FP64 cycle keeps the reference's 3 pre-smoothing and 60
// coarsest sweeps (itera and 20*itera, poisson_par.f:24, 45).  The
// correction cycle only needs an approximate correction, and the converged v
// does not depend on it (it is the fixed point of the fine-level residual):
// on wavepacket workloads from 257x65 to 2049x257, 2 pre-sweeps and 10
// coarsest sweeps took up to one more V-cycle than 3 and 10 but less time
// to the same 1e-9 residual on all but the smallest grids.  5 coarsest
// sweeps took as many cycles as 10 where ny <= 193 (coarsest ny <= 25) but
// up to 15 cycles on grids with ny = 257 (coarsest ny 33), so the count
// follows the coarsest level's ny.
constexpr int kMixedPreSweeps = 2;
constexpr int kMixedCoarseSweeps = 10;
constexpr int kMixedCoarseSweepsShort = 5;   // coarsest ny <= kShortCoarseNy
constexpr int kShortCoarseNy = 25;

Sched sched_of(const tbl_solve_opts *o, int coarsest_ny)
{
    const bool mixed = o && o->precision == 1;
    const int nc_mixed =
        coarsest_ny <= kShortCoarseNy ? kMixedCoarseSweepsShort : kMixedCoarseSweeps;
    Sched s{11, mixed ? kMixedPreSweeps : 3, 1, mixed ? nc_mixed : 60};
    if (o) {
        if (o->max_cycles > 0) s.maxc = o->max_cycles;
        if (o->nu1 > 0) s.nu1 = o->nu1;
        if (o->nu2 > 0) s.nu2 = o->nu2;
        if (o->ncoarse > 0) s.nc = o->ncoarse;
    }
    return s;
}

// rf = 1.2, or 1.0 when a single sweep is done (poisson_par.f:181-182)
inline double rf_of(int sweeps) { return sweeps == 1 ? 1.0 : 1.2; }

__global__ void tick_kernel(int *cycle, double *resmax, int K, int *nactive)
{
    const int t = threadIdx.x;
    if (t == 0) {
        *cycle += 1;
        *nactive = 0;
    }
    for (int k = t; k < K; k += blockDim.x) resmax[k] = 0.0;
}

__global__ void decide2_kernel(int K, const double *resmax, double tol, const int *cycle,
                               int *active, int *cycles, double *lastres, int *nactive)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K || !active[k]) return;
    const int c = *cycle;
    lastres[k] = resmax[k];
    if (resmax[k] < tol) {
        active[k] = 0;
        cycles[k] = c - 1;
    } else {
        cycles[k] = c;
        atomicAdd(nactive, 1);
    }
}

// One FAS V-cycle of every active mode (poisson_par.f:33-50)
int fas_cycle(tbl_ctx *c, const Sched &s, double tol)
{
    hipStream_t st = c->stream;
    const int L = c->L, K = c->K;
    hipLaunchKernelGGL(tick_kernel, dim3(1), dim3(256), 0, st, c->d_cycle, c->d_resmax, K, c->d_nactive);
    for (int l = 0; l < L - 1; ++l) {
        TRY(smooth_sweeps<double>(c->la_d[l], c->V[l], c->S[l], c->lu_d[l], s.nu1, rf_of(s.nu1),
                                  c->d_active, c->G64[l], c->H64[l], st));
        defect<double, double>(c->la_d[l], c->V[l], c->S[l], c->R[l], -1.0, c->d_active,
                               l == 0 ? c->d_resmax : nullptr, st);
        if (l == 0)
            hipLaunchKernelGGL(decide2_kernel, dim3((K + 255) / 256), dim3(256), 0, st, K,
                               c->d_resmax, tol, c->d_cycle, c->d_active, c->d_cycles,
                               c->d_lastres, c->d_nactive);
        restrict_level<double>(c->lv[l], c->lv[l + 1], K, c->V[l], c->R[l], c->V[l + 1],
                               c->W[l + 1], c->S[l + 1], 1, c->d_active, st);
        defect<double, double>(c->la_d[l + 1], c->V[l + 1], c->S[l + 1], c->S[l + 1], 1.0,
                               c->d_active, nullptr, st);
    }
    TRY(smooth_coarsest<double>(c->la_d[L - 1], c->V[L - 1], c->S[L - 1], c->lu_d[L - 1], s.nc,
                                rf_of(s.nc), c->d_active, c->G64[L - 1], st));
    for (int l = L - 1; l >= 1; --l) {
        prolong_level<double>(c->lv[l - 1], c->lv[l], K, c->V[l - 1], c->V[l], c->W[l], 1,
                              c->d_active, st);
        TRY(smooth_sweeps<double>(c->la_d[l - 1], c->V[l - 1], c->S[l - 1], c->lu_d[l - 1], s.nu2,
                                  rf_of(s.nu2), c->d_active, c->G64[l - 1], c->H64[l - 1], st));
    }
    return TBL_OK;
}

// FP32 smoothing sweeps of level l of the correction cycle on e = *cur: in
// place, or out of place between E32 and E32b with the one-launch sweep
// (then *cur moves to the buffer that holds the result)
int mixed_sweeps(tbl_ctx *c, int l, cx<float> **cur, int n, bool zs, cx<double> *vup)
{
    const int G = c->G32[l];
    const float rf = static_cast<float>(rf_of(n));
    if (shape_is_sweep(G)) {
        cx<float> *alt = *cur == c->E32[l] ? c->E32b[l] : c->E32[l];
        return sweep_sweeps<float>(c->la_f[l], *cur, alt, c->S32[l], c->lu_f[l], n, rf,
                                   c->d_active, shape_lanes(G), sweep_wsel(G), c->stream, zs, vup,
                                   cur);
    }
    return smooth_sweeps<float>(c->la_f[l], *cur, c->S32[l], c->lu_f[l], n, rf, c->d_active, G,
                                c->H32[l], c->stream, zs, vup);
}

// FP32 correction-scheme V-cycle on levels t..L-1 for L e = S32[t]: e[t]
// holds its start (zero when start_zero -- the staged and one-launch
// smoothers' first sweep then reads no column it has not relaxed itself,
// so e[t] need not even be cleared on those shapes); the coarser levels
// start from zero; vup: level t's last post-sweep adds its e into that FP64
// field (the update kernel's work)
int correction_vcycle(tbl_ctx *c, const Sched &s, std::vector<cx<float> *> &e, int t,
                      bool start_zero, cx<double> *vup)
{
    hipStream_t st = c->stream;
    const int L = c->L, K = c->K;
    // e's boundary columns are never written and its wall rows only ever
    // receive zeros, all zero since allocation; on the other shapes and the
    // coarsest level e is cleared first (by the caller or the restriction)
    auto zero_start = [&](int l) {
        return s.nu1 > 0 && l < L - 1 && (c->G32[l] < 0 || shape_is_sweep(c->G32[l]));
    };
    for (int l = t; l < L - 1; ++l) {
        TRY(mixed_sweeps(c, l, &e[l], s.nu1, (l > t || start_zero) && zero_start(l), nullptr));
        // the coarser source is the restricted defect (one pass, no stored
        // defect); the coarser e starts from zero -- cleared here unless its
        // first sweep starts from zero by itself
        defect_restrict<float>(c->la_f[l], c->lv[l + 1], e[l], c->S32[l], c->S32[l + 1],
                               c->d_active, st);
        if (!zero_start(l + 1))
            convert<float, float>(c->lv[l + 1].plane * K, e[l + 1], e[l + 1], 1, st);
    }
    // coarsest correction in FP32 too, from the zero the restriction left
    TRY(smooth_coarsest<float>(c->la_f[L - 1], e[L - 1], c->S32[L - 1], c->lu_f[L - 1], s.nc,
                               static_cast<float>(rf_of(s.nc)), c->d_active, c->G32[L - 1], st));
    for (int l = L - 1; l >= t + 1; --l) {
        prolong_level<float>(c->lv[l - 1], c->lv[l], K, e[l - 1], e[l], nullptr, 0, c->d_active,
                             st);
        TRY(mixed_sweeps(c, l - 1, &e[l - 1], s.nu2, false, l - 1 == t ? vup : nullptr));
    }
    return TBL_OK;
}

// the finest level's first sweep from e = 0 without clearing e (the
// staged and one-launch shapes)
bool finest_zero_start(const tbl_ctx *c, const Sched &s)
{
    const int G = c->G32[0];
    return s.nu1 > 0 && c->L > 1 && (G < 0 || shape_is_sweep(G));
}

// v += e: inside the finest level's last post-sweep on those shapes (each
// line adds its final e), else by the update kernel
bool fused_update(const tbl_ctx *c, const Sched &s)
{
    return finest_zero_start(c, s) && s.nu2 > 0;
}

// One FP64 defect-correction step with an FP32 correction-scheme V-cycle
int mixed_cycle(tbl_ctx *c, const Sched &s, double tol)
{
    hipStream_t st = c->stream;
    const int K = c->K;
    hipLaunchKernelGGL(tick_kernel, dim3(1), dim3(256), 0, st, c->d_cycle, c->d_resmax, K, c->d_nactive);
    // FP64 residual of the current v, rounded into the FP32 source
    defect<double, float>(c->la_d[0], c->V[0], c->S[0], c->S32[0], -1.0, c->d_active, c->d_resmax, st);
    hipLaunchKernelGGL(decide2_kernel, dim3((K + 255) / 256), dim3(256), 0, st, K, c->d_resmax,
                       tol, c->d_cycle, c->d_active, c->d_cycles, c->d_lastres, c->d_nactive);
    // e of each level (the out-of-place sweep moves it between two buffers)
    std::vector<cx<float> *> e(c->E32);
    const bool fused = fused_update(c, s);
    if (!finest_zero_start(c, s)) convert<float, float>(c->lv[0].plane * K, e[0], e[0], 1, st);
    TRY(correction_vcycle(c, s, e, 0, true, fused ? c->V[0] : nullptr));
    if (!fused) update(c->lv[0], K, c->V[0], e[0], c->d_active, st);
    return TBL_OK;
}

// The same step with the correction solved by full multigrid: the defect
// restricted to every level, solved on the coarsest, then on each finer
// level the prolonged coarse solution improved by a correction V-cycle of
// that level and below, up to the finest
int mixed_fmg_cycle(tbl_ctx *c, const Sched &s, double tol)
{
    hipStream_t st = c->stream;
    const int L = c->L, K = c->K;
    hipLaunchKernelGGL(tick_kernel, dim3(1), dim3(256), 0, st, c->d_cycle, c->d_resmax, K, c->d_nactive);
    defect<double, float>(c->la_d[0], c->V[0], c->S[0], c->S32[0], -1.0, c->d_active, c->d_resmax, st);
    hipLaunchKernelGGL(decide2_kernel, dim3((K + 255) / 256), dim3(256), 0, st, K, c->d_resmax,
                       tol, c->d_cycle, c->d_active, c->d_cycles, c->d_lastres, c->d_nactive);
    std::vector<cx<float> *> e(c->E32);
    // the source on every level (e of the coarser levels cleared)
    for (int l = 0; l < L - 1; ++l)
        restrict_level<float>(c->lv[l], c->lv[l + 1], K, e[l], c->S32[l], e[l + 1], nullptr,
                              c->S32[l + 1], 0, c->d_active, st);
    TRY(smooth_coarsest<float>(c->la_f[L - 1], e[L - 1], c->S32[L - 1], c->lu_f[L - 1], s.nc,
                               static_cast<float>(rf_of(s.nc)), c->d_active, c->G32[L - 1], st));
    convert<float, float>(c->lv[0].plane * K, e[0], e[0], 1, st);
    const bool fused = fused_update(c, s);
    for (int t = L - 2; t >= 0; --t) {
        prolong_level<float>(c->lv[t], c->lv[t + 1], K, e[t], e[t + 1], nullptr, 0, c->d_active,
                             st);
        TRY(correction_vcycle(c, s, e, t, false, t == 0 && fused ? c->V[0] : nullptr));
    }
    if (!fused) update(c->lv[0], K, c->V[0], e[0], c->d_active, st);
    return TBL_OK;
}

// kind 0: FP64 FAS cycle, 1: mixed cycle, 2: mixed cycle with a full-
// multigrid correction (each kind has its own two cached graphs)
int cycle_of(tbl_ctx *c, int kind, const Sched &s, double tol)
{
    return kind == 0 ? fas_cycle(c, s, tol) : kind == 1 ? mixed_cycle(c, s, tol)
                                                        : mixed_fmg_cycle(c, s, tol);
}

int run_cycle(tbl_ctx *c, int kind, const Sched &s, double tol, bool use_graph)
{
    if (!use_graph) return cycle_of(c, kind, s, tol);
    const int key[5] = {s.nu1, s.nu2, s.nc, shapes_key(c->G64), shapes_key(c->G32)};
    int slot = -1;
    for (int q = 0; q < 2; ++q)
        if (c->graph[kind][q] && std::memcmp(key, c->graph_key[kind][q], sizeof(key)) == 0 &&
            c->graph_tol[kind][q] == tol)
            slot = q;
    if (slot < 0) {
        slot = 1 - c->graph_last[kind];
        if (c->graph[kind][slot]) {
            TBL_HIP(hipGraphExecDestroy(c->graph[kind][slot]));
            c->graph[kind][slot] = nullptr;
        }
        if (!c->stream) return TBL_ERR_STATE;   // capture needs a non-null stream
        hipGraph_t g;
        TBL_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = cycle_of(c, kind, s, tol);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != TBL_OK) return rc;
        TBL_HIP(e);
        TBL_HIP(hipGraphInstantiate(&c->graph[kind][slot], g, nullptr, nullptr, 0));
        TBL_HIP(hipGraphDestroy(g));
        std::memcpy(c->graph_key[kind][slot], key, sizeof(key));
        c->graph_tol[kind][slot] = tol;
    }
    c->graph_last[kind] = slot;
    TBL_HIP(hipGraphLaunch(c->graph[kind][slot], c->stream));
    return TBL_OK;
}

}  // namespace

namespace tbl {

int solve_dev(tbl_ctx *c, const double *wx, const double *dz, double *uy, int direct_rhs,
              double tol, const tbl_solve_opts *opts, int *cycles, double *residual,
              tbl_solve_stats *stats)
{
    if (!c || !dz || !uy) return TBL_ERR_ARG;
    const Sched s = sched_of(opts, c->lv[c->L - 1].ny);
    const int prec = opts ? opts->precision : 0;
    if (prec < 0 || prec > 1) return TBL_ERR_ARG;
    if (opts && (opts->guess < 0 || opts->guess > 3)) return TBL_ERR_ARG;
    if (opts && (opts->fmg < 0 || opts->fmg > 2)) return TBL_ERR_ARG;
    const bool use_graph = opts && opts->use_graph;
    hipStream_t st = c->stream;
    const int K = c->K;
    const Level &F = c->lv[0];
    TBL_HIP(hipEventRecord(c->ev0, st));
    // source assembly: S = B g, v = uy (function_source, poisson_par.f:111-164)
    // guess 1: start from zero (boundary data taken as zero as well)
    const bool zero_start = opts && opts->guess == 1;
    gather_source(c->la_d[0], wx, dz, zero_start ? nullptr : uy, c->g.beta, c->S[0], c->V[0],
                  direct_rhs, st);
    int fallback = 0;
    double guess_res = 0.0;
    if (opts && (opts->guess == 2 || opts->guess == 3))
        TRY(guess_apply(c, st, opts->guess == 3, &fallback, &guess_res));
    fill_int(c->d_active, K, 1, st);
    fill_int(c->d_cycles, K, 0, st);
    fill_int(c->d_cycle, 1, 0, st);
    fill(c->d_lastres, K, 0.0, st);
    int done_cycles = 0;
    // full multigrid for the correction of the mixed cycle (a correction
    // scheme: every cycle solves L e = defect from e = 0), for the first
    // cycle (fmg 1) or every cycle (fmg 2)
    const int fmg = opts && prec == 1 && c->L > 1 ? opts->fmg : 0;
    // The host reads the active-mode count after each cycle and stops at
    // zero; each read idles the device for a round trip.  The first cycles
    // run back to back without it: as many as the previous solve of this
    // context needed before its last two (modes stop on the device, so a
    // solve that converges earlier only launches some cycles that do
    // nothing; after a solve that did not converge, none).
    const int ahead = std::min(std::max(c->last_need - 2, 0), s.maxc);
    bool all = false;
    for (int cyc = 1; cyc <= s.maxc; ++cyc) {
        TRY(run_cycle(c, (cyc == 1 && fmg) || fmg == 2 ? 2 : prec, s, tol, use_graph));
        TBL_HIP(hipGetLastError());
        done_cycles = cyc;
        if (cyc <= ahead && cyc < s.maxc) continue;
        TBL_HIP(hipMemcpyAsync(c->h_nactive, c->d_nactive, sizeof(int), hipMemcpyDeviceToHost, st));
        TBL_HIP(hipStreamSynchronize(st));
        if (*c->h_nactive == 0) {
            all = true;
            break;
        }
    }
    (void)done_cycles;
    scatter(c->g.nx, c->g.ny, K, F, c->V[0], uy, st);
    TBL_HIP(hipGetLastError());
    TBL_HIP(hipEventRecord(c->ev1, st));
    std::vector<int> cyc(K);
    std::vector<double> res(K);
    TBL_HIP(hipMemcpyAsync(cyc.data(), c->d_cycles, sizeof(int) * K, hipMemcpyDeviceToHost, st));
    TBL_HIP(hipMemcpyAsync(res.data(), c->d_lastres, sizeof(double) * K, hipMemcpyDeviceToHost, st));
    TBL_HIP(hipStreamSynchronize(st));
    // the cycles this solve needed: the last mode's count plus the check
    // that found it converged
    int need = 0;
    for (int k = 0; k < K; ++k) need = std::max(need, cyc[k]);
    c->last_need = all ? need + 1 : 0;
    if (cycles) std::memcpy(cycles, cyc.data(), sizeof(int) * K);
    if (residual) std::memcpy(residual, res.data(), sizeof(double) * K);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (int k = 0; k < K; ++k) {
            stats->cycles_max = std::max(stats->cycles_max, cyc[k]);
            stats->cycles_total += cyc[k];
            stats->nconverged += res[k] < tol ? 1 : 0;
            stats->residual_max = std::max(stats->residual_max, res[k]);
        }
        stats->fallback = fallback;
        stats->guess_residual_max = guess_res;
        float ms = 0.f;
        TBL_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        stats->device_ms = ms;
    }
    return TBL_OK;
}

}  // namespace tbl

extern "C" int tbl_poi_uy_dev(tbl_ctx *c, const double *d_wx, const double *d_dwzdx, double *d_uy,
                              double erromax, const tbl_solve_opts *opts, int *cycles,
                              double *residual, tbl_solve_stats *stats)
{
    return solve_dev(c, d_wx, d_dwzdx, d_uy, 0, erromax, opts, cycles, residual, stats);
}

// Solve L v = B f directly (the synthetic form: wx = 0, dwzdx = -f)
extern "C" int tbl_solve_rhs_dev(tbl_ctx *c, const double *d_f, double *d_uy, double erromax,
                                 const tbl_solve_opts *opts, int *cycles, double *residual,
                                 tbl_solve_stats *stats)
{
    return solve_dev(c, nullptr, d_f, d_uy, 1, erromax, opts, cycles, residual, stats);
}

static int staging(tbl_ctx *c)
{
    if (c->d_u0) return TBL_OK;
    const size_t n = static_cast<size_t>(c->g.nx) * c->g.ny * c->K * 2;
    TRY(dalloc(&c->d_u0, n));
    TRY(dalloc(&c->d_u1, n));
    TRY(dalloc(&c->d_u2, n));
    return TBL_OK;
}

extern "C" int tbl_poi_uy(tbl_ctx *c, const double *wx, const double *dwzdx, double *uy,
                          double erromax, const tbl_solve_opts *opts, int *cycles,
                          double *residual, tbl_solve_stats *stats)
{
    if (!c || !dwzdx || !uy) return TBL_ERR_ARG;
    TRY(staging(c));
    const size_t bytes = sizeof(double) * 2 * c->g.nx * c->g.ny * c->K;
    if (wx) TBL_HIP(hipMemcpy(c->d_u0, wx, bytes, hipMemcpyHostToDevice));
    TBL_HIP(hipMemcpy(c->d_u1, dwzdx, bytes, hipMemcpyHostToDevice));
    TBL_HIP(hipMemcpy(c->d_u2, uy, bytes, hipMemcpyHostToDevice));
    TRY(solve_dev(c, wx ? c->d_u0 : nullptr, c->d_u1, c->d_u2, 0, erromax, opts, cycles, residual,
                  stats));
    TBL_HIP(hipMemcpy(uy, c->d_u2, bytes, hipMemcpyDeviceToHost));
    return TBL_OK;
}

extern "C" int tbl_deriv_dev(tbl_ctx *c, int kind, const double *d_fc, double *d_out)
{
    if (!c || !d_fc || !d_out || kind < 0 || kind > 4) return TBL_ERR_ARG;
    if (d_fc == d_out) return TBL_ERR_ARG;
    TRY(deriv(c->dargs, kind, d_fc, d_out, c->stream));
    TBL_HIP(hipGetLastError());
    return TBL_OK;
}

static int deriv_host(tbl_ctx *c, int kind, const double *fc, double *out)
{
    if (!c || !fc || !out) return TBL_ERR_ARG;
    TRY(staging(c));
    const size_t bytes = sizeof(double) * 2 * c->g.nx * c->g.ny * c->K;
    TBL_HIP(hipMemcpy(c->d_u0, fc, bytes, hipMemcpyHostToDevice));
    TRY(deriv(c->dargs, kind, c->d_u0, c->d_u1, c->stream));
    TBL_HIP(hipStreamSynchronize(c->stream));
    TBL_HIP(hipMemcpy(out, c->d_u1, bytes, hipMemcpyDeviceToHost));
    return TBL_OK;
}

extern "C" int tbl_derivs_dev(tbl_ctx *c, const double *d_fc, double *d_ddx, double *d_d2dx2,
                              double *d_ddy, double *d_d2dy2)
{
    if (!c || !d_fc) return TBL_ERR_ARG;
    for (const double *o : {d_ddx, d_d2dx2, d_ddy, d_d2dy2})
        if (o == d_fc) return TBL_ERR_ARG;
    TRY(derivs4(c->dargs, d_fc, d_ddx, d_d2dx2, d_ddy, d_d2dy2, c->stream));
    TBL_HIP(hipGetLastError());
    return TBL_OK;
}

extern "C" int tbl_derparx(tbl_ctx *c, const double *fc, double *o) { return deriv_host(c, 0, fc, o); }
extern "C" int tbl_derparxx(tbl_ctx *c, const double *fc, double *o) { return deriv_host(c, 1, fc, o); }
extern "C" int tbl_dery(tbl_ctx *c, const double *fc, double *o) { return deriv_host(c, 2, fc, o); }
extern "C" int tbl_deryy(tbl_ctx *c, const double *fc, double *o) { return deriv_host(c, 3, fc, o); }
extern "C" int tbl_deryfv(tbl_ctx *c, const double *fc, double *o) { return deriv_host(c, 4, fc, o); }

extern "C" int tbl_residual_dev(tbl_ctx *c, const double *d_wx, const double *d_dwzdx,
                                const double *d_uy, double *d_out, double *residual)
{
    if (!c || !d_dwzdx || !d_uy || !d_out) return TBL_ERR_ARG;
    hipStream_t st = c->stream;
    const int K = c->K;
    const Level &F = c->lv[0];
    gather_source(c->la_d[0], d_wx, d_dwzdx, d_uy, c->g.beta, c->S[0], c->V[0], 0, st);
    fill(c->d_resmax, K, 0.0, st);
    fill(reinterpret_cast<double *>(c->R[0]), F.plane * K * 2, 0.0, st);
    defect<double, double>(c->la_d[0], c->V[0], c->S[0], c->R[0], -1.0, nullptr, c->d_resmax, st);
    fill(d_out, static_cast<size_t>(c->g.nx) * c->g.ny * K * 2, 0.0, st);
    scatter(c->g.nx, c->g.ny, K, F, c->R[0], d_out, st);
    if (residual)
        TBL_HIP(hipMemcpyAsync(residual, c->d_resmax, sizeof(double) * K, hipMemcpyDeviceToHost, st));
    TBL_HIP(hipStreamSynchronize(st));
    return TBL_OK;
}

// Lanes per line of a level's smoother (the autotuner's knob).  prec 0: FP64
// smoother, 1: FP32 smoother.
extern "C" int tbl_set_line_shape(tbl_ctx *c, int prec, int level, int lanes)
{
    if (c && level == c->L && lanes == kCoarseSeqShape) {
        (prec ? c->G32 : c->G64)[level - 1] = lanes;
        return TBL_OK;
    }
    const bool strip = shape_is_strip(lanes);
    const bool sweep = shape_is_sweep(lanes);
    const int a = shape_lanes(lanes);
    if (!c || level < 1 || level > c->L || (a != 16 && a != 32 && a != 64))
        return TBL_ERR_ARG;
    const Level &lv = c->lv[level - 1];
    if (line_rows_for(a, lv.ny) < 0) return TBL_ERR_GRID;
    // one-launch sweep: FP32 correction cycle only, not on the coarsest level
    if (sweep && (prec == 0 || level == c->L || sweep_wsel(lanes) > 5 ||
                  sweep_lds(lv.ld, sizeof(cx<float>), a) > 160 * 1024))
        return TBL_ERR_GRID;
    // strip form: not on the coarsest level, and only for rows per lane that
    // stay in registers with two lines in flight (FP64 <= 3, FP32 <= 5)
    if (strip && (level == c->L || line_rows_for(a, lv.ny) > (prec ? 5 : 3) ||
                  strip_geom(lv.nx, lv.ny, c->K, prec ? 4 : 8).nstrips < 1))
        return TBL_ERR_GRID;
    (prec ? c->G32 : c->G64)[level - 1] = lanes;
    return TBL_OK;
}

extern "C" int tbl_get_line_shape(tbl_ctx *c, int prec, int level)
{
    if (!c || level < 1 || level > c->L) return -1;
    return (prec ? c->G32 : c->G64)[level - 1];
}

extern "C" int tbl_time_smoother(tbl_ctx *c, int prec, int level, int sweeps, double *ms,
                                 int *launches)
{
    if (!c || !ms || level < 1 || level > c->L || sweeps < 1) return TBL_ERR_ARG;
    const int l = level - 1;
    hipStream_t st = c->stream;
    TBL_HIP(hipEventRecord(c->ev0, st));
    if (level == c->L && prec)
        // the coarsest level's solve as each cycle runs it, one launch
        TRY(smooth_coarsest<float>(c->la_f[l], c->E32[l], c->S32[l], c->lu_f[l], sweeps, 1.0f,
                                   nullptr, c->G32[l], st));
    else if (level == c->L)
        TRY(smooth_coarsest<double>(c->la_d[l], c->V[l], c->S[l], c->lu_d[l], sweeps, 1.0,
                                    nullptr, c->G64[l], st));
    else if (prec && shape_is_sweep(c->G32[l])) {
        cx<float> *cur = nullptr;
        TRY(sweep_sweeps<float>(c->la_f[l], c->E32[l], c->E32b[l], c->S32[l], c->lu_f[l], sweeps,
                                1.2f, nullptr, shape_lanes(c->G32[l]), sweep_wsel(c->G32[l]), st,
                                false, nullptr, &cur));
    } else if (prec)
        TRY(smooth_sweeps<float>(c->la_f[l], c->E32[l], c->S32[l], c->lu_f[l], sweeps, 1.2f,
                                 nullptr, c->G32[l], c->H32[l], st));
    else
        TRY(smooth_sweeps<double>(c->la_d[l], c->V[l], c->S[l], c->lu_d[l], sweeps, 1.2,
                                  nullptr, c->G64[l], c->H64[l], st));
    TBL_HIP(hipGetLastError());
    TBL_HIP(hipEventRecord(c->ev1, st));
    TBL_HIP(hipEventSynchronize(c->ev1));
    float t = 0.f;
    TBL_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = t;
    if (launches) {
        if (level == c->L) *launches = 1;
        else if (shape_is_sweep((prec ? c->G32 : c->G64)[l])) *launches = sweeps;
        else if (shape_is_strip((prec ? c->G32 : c->G64)[l])) *launches = 2 * sweeps;
        else *launches = 4 * sweeps;
    }
    return TBL_OK;
}

namespace tbl {

// One RK4 stage of one context, issued asynchronously on its stream (no host
// synchronisation, so it can be captured): the four derivatives of v, then
// the v-Poisson solve with a fixed number of V-cycles, warm-started from v
// and written back into v.
// the cycle kind of cycle cyc (1-based) of a solve with these options
static int kind_of(const tbl_ctx *c, const tbl_solve_opts *opts, int cyc)
{
    const int prec = opts ? opts->precision : 0;
    const int fmg = opts && prec == 1 && c->L > 1 ? opts->fmg : 0;
    return (cyc == 1 && fmg) || fmg == 2 ? 2 : prec;
}

// A stage's first part: the four derivatives of v, the source, the solve's
// flags and its first `cycles` V-cycles from v (no copy-out), launched on
// the context's stream (capturable).
int stage_begin(tbl_ctx *c, const double *f, double *v, double *const *d, double tol,
                const tbl_solve_opts *opts, int cycles)
{
    hipStream_t st = c->stream;
    const Sched s = sched_of(opts, c->lv[c->L - 1].ny);
    const int K = c->K;
    TRY(derivs4(c->dargs, v, d[0], d[1], d[2], d[3], st));
    gather_source(c->la_d[0], nullptr, f, v, c->g.beta, c->S[0], c->V[0], 1, st);
    fill_int(c->d_active, K, 1, st);
    fill_int(c->d_cycles, K, 0, st);
    fill_int(c->d_cycle, 1, 0, st);
    fill(c->d_lastres, K, 0.0, st);
    for (int cyc = 1; cyc <= cycles; ++cyc) TRY(cycle_of(c, kind_of(c, opts, cyc), s, tol));
    return TBL_OK;
}

// One more V-cycle of a stage's solve (cycle number cyc), replayed from the
// context's cached graphs.
int stage_cycle(tbl_ctx *c, double tol, const tbl_solve_opts *opts, int cyc)
{
    const Sched s = sched_of(opts, c->lv[c->L - 1].ny);
    return run_cycle(c, kind_of(c, opts, cyc), s, tol, true);
}

// A stage's copy-out: v = the solution.
int stage_end(tbl_ctx *c, double *v)
{
    scatter(c->g.nx, c->g.ny, c->K, c->lv[0], c->V[0], v, c->stream);
    return TBL_OK;
}

int stage_issue(tbl_ctx *c, const double *f, double *v, double *const *d, double tol,
                const tbl_solve_opts *opts)
{
    TRY(stage_begin(c, f, v, d, tol, opts, sched_of(opts, c->lv[c->L - 1].ny).maxc));
    return stage_end(c, v);
}

}  // namespace tbl
