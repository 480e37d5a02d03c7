// Learned initial guess for the v-Poisson solve.
//
// The U-Net (tbl_net.cpp) works on one coarse multigrid level lg: its input
// is the source S = B g (function_source) fully weighted down to that level
// (scaled by the mode's max |S| there), kz and the normalised height y/ytop;
// its output is v at that level's points scaled by
// sv = max|S| / (kz^2 + kappa^2).  The prediction
// is prolongated to the fine grid with the multigrid's own bilinear
// interpolation (poisson_par.f:877-920) onto the interior of the fine grid;
// the boundary columns and the wall row keep the warm start's values.  The
// candidates are compared on the network's level: after two line sweeps of
// each (for the fully weighted restriction of the source), the max residual
// of the prediction against that of the injected warm start.  (The residual
// of the raw prediction, on any level, is dominated by its pixel-scale
// noise, which the operator amplifies and the first sweeps remove.)  A mode
// whose prediction's residual is more than kGuessAccept times the warm
// start's keeps the warm start; multigrid then runs to the same tolerance
// either way.
#include <algorithm>
#include <cmath>

#include "tbl_ctx.h"

struct tbl_net;

namespace tbl {
int net_forward_async(tbl_net *n, const float *d_x, float *d_y, int N, hipStream_t st);
}

namespace tbl {

struct Guess {
    tbl_net *net = nullptr;
    int lvl = 0, H = 0, W = 0;
    double kappa2 = 1.0, ytop = 1.0;
    float *feat = nullptr, *pred = nullptr;
    double *sf = nullptr, *res0 = nullptr, *resg = nullptr;
    int *fallback = nullptr;
    double *ycoarse = nullptr;    // y of the coarse level's rows
};

namespace {

// max |g| per mode over the fine grid (internal layout): grid (blocks, K),
// per-block maxima combined with an atomic max on the bit pattern (out must
// be cleared first); a mode whose g is zero gets 1 (finish_absmax_kernel)
__global__ void mode_absmax_kernel(Level F, int K, const cx<double> *G, double *out)
{
    const int k = blockIdx.y;
    double m = 0.0;
    const cx<double> *g = G + static_cast<size_t>(k) * F.plane;
    for (size_t e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; e < F.plane;
         e += static_cast<size_t>(gridDim.x) * 256) {
        const int j = static_cast<int>(e % F.ld);
        if (j < F.ny) m = fmax(m, hypot(g[e].re, g[e].im));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        atomicMax(reinterpret_cast<unsigned long long *>(out + k), __double_as_longlong(m));
    }
}

__global__ void finish_absmax_kernel(int K, double *out)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < K && !(out[k] > 0)) out[k] = 1.0;
}

// features (K, H, W, 4) at level lg from the fine g (stride s = 2^(lg-1))
__global__ void features_kernel(Level F, Level Cl, int K, int H, int W, int s, const cx<double> *G,
                                const double *sf, const double *yc, double ytop, double beta,
                                float kzscale, float *feat)
{
    const long total = static_cast<long>(K) * H * W;
    const long e = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= total) return;
    const int w = static_cast<int>(e % W), h = static_cast<int>((e / W) % H);
    const int k = static_cast<int>(e / (static_cast<long>(W) * H));
    float a = 0.f, b = 0.f, kz = 0.f, yy = 0.f;
    if (w < Cl.nx && h < Cl.ny) {
        const cx<double> g = G[static_cast<size_t>(k) * F.plane + static_cast<size_t>(w * s) * F.ld + h * s];
        a = static_cast<float>(g.re / sf[k]);
        b = static_cast<float>(g.im / sf[k]);
        kz = static_cast<float>(k * beta) * kzscale;
        yy = static_cast<float>(yc[h] / ytop);
    }
    float *o = feat + e * 4;
    o[0] = a; o[1] = b; o[2] = kz; o[3] = yy;
}

// targets (K, H, W, 2): v at level lg points / sv
__global__ void targets_kernel(Level F, Level Cl, int K, int H, int W, int s, const cx<double> *V,
                               const double *sf, double beta, double kappa2, float *tgt)
{
    const long total = static_cast<long>(K) * H * W;
    const long e = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= total) return;
    const int w = static_cast<int>(e % W), h = static_cast<int>((e / W) % H);
    const int k = static_cast<int>(e / (static_cast<long>(W) * H));
    float a = 0.f, b = 0.f;
    if (w < Cl.nx && h < Cl.ny) {
        const double kz = k * beta;
        const double sv = sf[k] / (kz * kz + kappa2);
        const cx<double> v = V[static_cast<size_t>(k) * F.plane + static_cast<size_t>(w * s) * F.ld + h * s];
        a = static_cast<float>(v.re / sv);
        b = static_cast<float>(v.im / sv);
    }
    tgt[e * 2] = a;
    tgt[e * 2 + 1] = b;
}

// prediction -> V at level lg (interior only; wall row and side columns 0)
__global__ void place_kernel(Level Cl, int K, int H, int W, const float *pred, const double *sf,
                             double beta, double kappa2, cx<double> *V)
{
    const size_t idx = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= Cl.plane * K) return;
    const int k = static_cast<int>(idx / Cl.plane);
    const size_t rem = idx - static_cast<size_t>(k) * Cl.plane;
    const int i = static_cast<int>(rem / Cl.ld), j = static_cast<int>(rem % Cl.ld);
    cx<double> v = {0, 0};
    if (i >= 1 && i <= Cl.nx - 2 && j >= 1 && j < Cl.ny) {
        const double kz = k * beta;
        const double sv = sf[k] / (kz * kz + kappa2);
        const float *p = pred + ((static_cast<long>(k) * H + j) * W + i) * 2;
        v = {static_cast<double>(p[0]) * sv, static_cast<double>(p[1]) * sv};
    }
    V[idx] = v;
}

// line sweeps applied on the network's level before the comparison
constexpr int kGuessSweeps = 2;
// a prediction is kept unless its residual exceeds the warm start's by this
// factor (or is not finite): even after a few sweeps the residual weights
// the prediction's rough, quickly smoothed error far above the smooth error
// that decides how many V-cycles remain
constexpr double kGuessAccept = 10.0;

// accepted modes: the prolongated prediction on the interior, the warm
// start W on the boundary columns and the wall row; other modes: W
__global__ void fallback_kernel(Level F, int K, const double *res0, const double *resg,
                                const cx<double> *W, cx<double> *V, int accept_all, int *count)
{
    const size_t idx = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= F.plane * K) return;
    const int k = static_cast<int>(idx / F.plane);
    const size_t rem = idx - static_cast<size_t>(k) * F.plane;
    const int i = static_cast<int>(rem / F.ld), j = static_cast<int>(rem % F.ld);
    const bool accept = accept_all || resg[k] < kGuessAccept * res0[k];
    if (!accept || i == 0 || i == F.nx - 1 || j == 0) V[idx] = W[idx];
    if (!accept && rem == 0) atomicAdd(count, 1);
}

unsigned nblk(size_t n) { return static_cast<unsigned>((n + 255) / 256); }


}  // namespace

int guess_free(tbl_ctx *c)
{
    Guess *g = c->guess;
    if (!g) return TBL_OK;
    auto fr = [](void *p) { if (p) (void)hipFree(p); };
    fr(g->feat); fr(g->pred); fr(g->sf); fr(g->res0); fr(g->resg); fr(g->fallback); fr(g->ycoarse);
    delete g;
    c->guess = nullptr;
    return TBL_OK;
}

static int guess_setup(tbl_ctx *c, tbl_net *net, int lvl)
{
    if (lvl < 1 || lvl > c->L) return TBL_ERR_ARG;
    guess_free(c);
    Guess *g = new Guess();
    c->guess = g;
    g->net = net;
    g->lvl = lvl;
    const Level &Cl = c->lv[lvl - 1];
    g->H = round_up(Cl.ny, 8);
    g->W = round_up(Cl.nx, 8);
    std::vector<double> yc(Cl.ny);
    const int s = 1 << (lvl - 1);
    for (int j = 0; j < Cl.ny; ++j) {
        const int jf = j * s;
        yc[j] = (c->g.stf == 1.0) ? c->g.dy0 * jf
                                  : c->g.dy0 * (std::pow(c->g.stf, jf) - 1.0) / (c->g.stf - 1.0);
    }
    g->ytop = (c->g.stf == 1.0) ? c->g.dy0 * (c->g.ny - 1)
                                : c->g.dy0 * (std::pow(c->g.stf, c->g.ny - 1) - 1.0) / (c->g.stf - 1.0);
    g->kappa2 = 4.0 / (g->ytop * g->ytop);
    const size_t px = static_cast<size_t>(c->K) * g->H * g->W;
    TBL_HIP(hipMalloc(&g->feat, px * 4 * sizeof(float)));
    TBL_HIP(hipMalloc(&g->pred, px * 2 * sizeof(float)));
    TBL_HIP(hipMemset(g->feat, 0, px * 4 * sizeof(float)));
    TBL_HIP(hipMalloc(&g->sf, c->K * sizeof(double)));
    TBL_HIP(hipMalloc(&g->res0, c->K * sizeof(double)));
    TBL_HIP(hipMalloc(&g->resg, c->K * sizeof(double)));
    TBL_HIP(hipMalloc(&g->fallback, sizeof(int)));
    TBL_HIP(hipMalloc(&g->ycoarse, Cl.ny * sizeof(double)));
    TBL_HIP(hipMemcpy(g->ycoarse, yc.data(), Cl.ny * sizeof(double), hipMemcpyHostToDevice));
    return TBL_OK;
}

// W[0] = V[0]; for l < L: W[l+1] = injection of W[l], S[l+1] = full weighting
// of S[l] (R[l+1] takes the injection's unused copy).  Needs S[0].
static void restrict_chain(tbl_ctx *c, int L, hipStream_t st)
{
    const Level &F = c->lv[0];
    convert<double, double>(F.plane * c->K, c->V[0], c->W[0], 0, st);
    for (int l = 0; l < L; ++l)
        restrict_level<double>(c->lv[l], c->lv[l + 1], c->K, c->W[l], c->S[l], c->W[l + 1],
                               c->R[l + 1], c->S[l + 1], 1, nullptr, st);
}

// features from the fully weighted source on the network's level (S[L] after
// restrict_chain): point samples of the fine source would alias its
// short-wave content on the coarse levels
static void build_features(tbl_ctx *c, hipStream_t st)
{
    Guess *g = c->guess;
    const int L = g->lvl - 1;
    const Level &Cl = c->lv[L];
    fill(g->sf, c->K, 0.0, st);
    hipLaunchKernelGGL(mode_absmax_kernel, dim3(16, c->K), dim3(256), 0, st, Cl, c->K, c->S[L], g->sf);
    hipLaunchKernelGGL(finish_absmax_kernel, dim3((c->K + 255) / 256), dim3(256), 0, st, c->K, g->sf);
    const long n = static_cast<long>(c->K) * g->H * g->W;
    hipLaunchKernelGGL(features_kernel, dim3(nblk(n)), dim3(256), 0, st, Cl, Cl, c->K, g->H, g->W, 1,
                       c->S[L], g->sf, g->ycoarse, g->ytop, c->g.beta,
                       static_cast<float>(std::sqrt(1.0 / g->kappa2)), g->feat);
}

// Called by the solver after gather+source (R[0] holds g, S[0] the source):
// writes the guess into V[0] and the fallback count into *fallback.
int guess_apply(tbl_ctx *c, hipStream_t st, bool accept_all, int *fallback,
                double *guess_res_max)
{
    Guess *g = c->guess;
    if (!g || !g->net) return TBL_ERR_STATE;
    const Level &F = c->lv[0];
    const int K = c->K;
    const int L = g->lvl - 1;
    const Level &Cl = c->lv[L];
    if (accept_all && L > 0) {
        // every mode starts from the prediction: restrict the source only,
        // and assign the interpolant to the fine interior (the warm start's
        // boundary columns and wall row stay)
        for (int l = 0; l < L; ++l)
            restrict_level<double>(c->lv[l], c->lv[l + 1], K, c->V[l], c->S[l], c->W[l + 1],
                                   nullptr, c->S[l + 1], 0, nullptr, st);
        build_features(c, st);
        int rc = net_forward_async(g->net, g->feat, g->pred, K, st);
        if (rc != TBL_OK) return rc;
        hipLaunchKernelGGL(place_kernel, dim3(nblk(Cl.plane * K)), dim3(256), 0, st, Cl, K, g->H,
                           g->W, g->pred, g->sf, c->g.beta, g->kappa2, c->V[L]);
        for (int l = L; l >= 1; --l) {
            if (l - 1 > 0)
                fill(reinterpret_cast<double *>(c->V[l - 1]), c->lv[l - 1].plane * K * 2, 0.0, st);
            prolong_level<double>(c->lv[l - 1], c->lv[l], K, c->V[l - 1], c->V[l], nullptr,
                                  l - 1 > 0 ? 0 : 2, nullptr, st);
        }
        TBL_HIP(hipGetLastError());
        *fallback = 0;
        *guess_res_max = 0.0;
        return TBL_OK;
    }
    // warm start (V[0] after gather) kept in W[0] and injected down to level
    // L (W[L]); the source fully weighted alongside (S[L])
    restrict_chain(c, L, st);
    build_features(c, st);
    int rc = net_forward_async(g->net, g->feat, g->pred, K, st);
    if (rc != TBL_OK) return rc;
    hipLaunchKernelGGL(place_kernel, dim3(nblk(Cl.plane * K)), dim3(256), 0, st, Cl, K, g->H, g->W,
                       g->pred, g->sf, c->g.beta, g->kappa2, c->V[L]);
    // residuals on level L after kGuessSweeps line sweeps of each candidate
    // (the prediction's pixel-scale noise is what the operator amplifies most;
    // the sweeps remove it, and the smoothed prediction is the one used)
    // (without the test only the prediction is swept)
    if (L > 0) {
        for (cx<double> *cand : {c->V[L], c->W[L]}) {
            if (accept_all && cand == c->W[L]) continue;
            rc = smooth_sweeps<double>(c->la_d[L], cand, c->S[L], c->lu_d[L], kGuessSweeps, 1.2,
                                       nullptr, sweep_shape(c->G64[L]), c->H64[L], st);
            if (rc != TBL_OK) return rc;
        }
    }
    if (!accept_all) {
        fill(g->resg, K, 0.0, st);
        fill(g->res0, K, 0.0, st);
        const int l = L > 0 ? L : 0;
        defect<double, double>(c->la_d[l], c->V[l], c->S[l], c->R[l], -1.0, nullptr, g->resg, st);
        defect<double, double>(c->la_d[l], c->W[l], c->S[l], c->R[l], -1.0, nullptr, g->res0, st);
    }
    for (int l = L; l >= 1; --l) {
        fill(reinterpret_cast<double *>(c->V[l - 1]), c->lv[l - 1].plane * K * 2, 0.0, st);
        prolong_level<double>(c->lv[l - 1], c->lv[l], K, c->V[l - 1], c->V[l], nullptr, 0, nullptr, st);
    }
    fill_int(g->fallback, 1, 0, st);
    hipLaunchKernelGGL(fallback_kernel, dim3(nblk(F.plane * K)), dim3(256), 0, st, F, K, g->res0,
                       g->resg, c->W[0], c->V[0], accept_all ? 1 : 0, g->fallback);
    TBL_HIP(hipGetLastError());
    if (accept_all) {
        *fallback = 0;
        *guess_res_max = 0.0;
        return TBL_OK;
    }
    std::vector<double> rg(K);
    TBL_HIP(hipMemcpyAsync(fallback, g->fallback, sizeof(int), hipMemcpyDeviceToHost, st));
    TBL_HIP(hipMemcpyAsync(rg.data(), g->resg, K * sizeof(double), hipMemcpyDeviceToHost, st));
    TBL_HIP(hipStreamSynchronize(st));
    *guess_res_max = *std::max_element(rg.begin(), rg.end());
    return TBL_OK;
}

}  // namespace tbl

using namespace tbl;

extern "C" int tbl_guess_attach(tbl_ctx *c, tbl_net *net, int level)
{
    if (!c) return TBL_ERR_ARG;
    if (!net) return guess_free(c);
    return guess_setup(c, net, level);
}

extern "C" int tbl_guess_shape(tbl_ctx *c, int *H, int *W)
{
    if (!c || !c->guess || !H || !W) return TBL_ERR_STATE;
    *H = c->guess->H;
    *W = c->guess->W;
    return TBL_OK;
}

// Training pairs from a solved batch: f and v in the user layout (device),
// features (K, H, W, 4) and targets (K, H, W, 2) written as fp32 (device).
extern "C" int tbl_guess_pairs(tbl_ctx *c, const double *d_f, const double *d_v, float *d_x,
                               float *d_t)
{
    if (!c || !c->guess || !d_f || !d_v || !d_x || !d_t) return TBL_ERR_ARG;
    Guess *g = c->guess;
    hipStream_t st = c->stream;
    const Level &F = c->lv[0], &Cl = c->lv[g->lvl - 1];
    gather_source(c->la_d[0], nullptr, d_f, d_v, c->g.beta, c->S[0], c->V[0], 1, st);
    restrict_chain(c, g->lvl - 1, st);
    build_features(c, st);
    const long n = static_cast<long>(c->K) * g->H * g->W;
    hipLaunchKernelGGL(targets_kernel, dim3(nblk(n)), dim3(256), 0, st, F, Cl, c->K, g->H, g->W,
                       1 << (g->lvl - 1), c->V[0], g->sf, c->g.beta, g->kappa2, d_t);
    TBL_HIP(hipMemcpyAsync(d_x, g->feat, n * 4 * sizeof(float), hipMemcpyDeviceToDevice, st));
    TBL_HIP(hipGetLastError());
    TBL_HIP(hipStreamSynchronize(st));
    return TBL_OK;
}
