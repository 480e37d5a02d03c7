// U-Net object of the learned initial guess: topology, parameters, forward,
// backward and Adam, plus its C ABI (tbl_net_*).
//
// Topology (4 levels, base 32 channels, 3x3 convolutions + ReLU):
//   L0: 4 -> 32 -> 32 ----------------------------------------- skip0
//   L1:   pool, 32 -> 64 -> 64 -------------------------- skip1  |
//   L2:     pool, 64 -> 128 -> 128 ------------- skip2       |    |
//   L3:       pool, 128 -> 256 -> 256            |           |    |
//   L2:     up + concat(skip2): 384 -> 128 -> 128            |    |
//   L1:   up + concat(skip1): 192 -> 64 -> 64                     |
//   L0: up + concat(skip0): 96 -> 32 -> 32, 1x1 conv 32 -> 2 (no ReLU)
// Input channels: RHS (re, im), kz, y-metric; output: scaled v (re, im).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/trans_bl_amd.h"
#include "tbl_nn.h"
#include "tbl_common.h"

using namespace tbl::nn;

struct tbl_net;
extern "C" int tbl_net_destroy(tbl_net *n);

namespace {

struct Conv {
    int cin, cout, ks, lvl, relu;
    long woff, boff;   // offsets into the flat parameter vector
    long wpoff;        // offset of the packed 3x3 weights (direct convolution)
};

enum {
    E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, D2A, D2B, D1A, D1B, D0A, D0B, OUT, NCONV
};

template <typename T>
int dmalloc(T **p, size_t n)
{
    *p = nullptr;
    if (!n) return TBL_OK;
    TBL_HIP(hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)));
    TBL_HIP(hipMemset(*p, 0, n * sizeof(T)));
    return TBL_OK;
}

}  // namespace

struct tbl_net {
    int H = 0, W = 0, nmax = 0, base = 32;
    hipStream_t stream = nullptr, own = nullptr;
    Conv L[NCONV];
    long nparam = 0;
    // parameters: fp32 master, bf16 copy, gradient, Adam moments
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16 *wb = nullptr;
    bf16 *wp = nullptr;          // 3x3 weights packed for the direct convolution
    bool wp_stale = true;        // wb changed since the last packing
    long step = 0;
    // activations (bf16, NHWC) and fp32 pre/post-activation scratch
    bf16 *x_in = nullptr;
    bf16 *act[NCONV] = {nullptr};
    bf16 *pool_out[3] = {nullptr};
    bf16 *cat[3] = {nullptr};
    float *z = nullptr;          // fp32 output of the current GEMM (max P*cout)
    float *out = nullptr;        // fp32 network output (P0 x 2)
    bf16 *col = nullptr;         // im2col scratch
    float *dcol = nullptr;       // dcol scratch
    bf16 *dzb = nullptr;         // bf16 gradient of the current conv
    // fp32 gradients of activations
    float *d_act[NCONV] = {nullptr};
    float *d_pool[3] = {nullptr};
    float *d_cat[3] = {nullptr};
    float *d_out = nullptr;
    float *d_loss = nullptr;
    float *d_xin = nullptr;

    int hw(int lvl) const { return (H >> lvl) * (W >> lvl); }
};

namespace {

int net_alloc(tbl_net *n)
{
    const int B = n->base;
    auto set = [&](int id, int cin, int cout, int ks, int lvl, int relu) {
        n->L[id] = Conv{cin, cout, ks, lvl, relu, 0, 0};
    };
    set(E0A, 4, B, 3, 0, 1);          set(E0B, B, B, 3, 0, 1);
    set(E1A, B, 2 * B, 3, 1, 1);      set(E1B, 2 * B, 2 * B, 3, 1, 1);
    set(E2A, 2 * B, 4 * B, 3, 2, 1);  set(E2B, 4 * B, 4 * B, 3, 2, 1);
    set(E3A, 4 * B, 8 * B, 3, 3, 1);  set(E3B, 8 * B, 8 * B, 3, 3, 1);
    set(D2A, 12 * B, 4 * B, 3, 2, 1); set(D2B, 4 * B, 4 * B, 3, 2, 1);
    set(D1A, 6 * B, 2 * B, 3, 1, 1);  set(D1B, 2 * B, 2 * B, 3, 1, 1);
    set(D0A, 3 * B, B, 3, 0, 1);      set(D0B, B, B, 3, 0, 1);
    set(OUT, B, 2, 1, 0, 0);
    long off = 0;
    for (auto &c : n->L) {
        c.woff = off;
        off += static_cast<long>(c.ks) * c.ks * c.cin * c.cout;
        c.boff = off;
        off += c.cout;
    }
    n->nparam = off;
    const size_t np = static_cast<size_t>(off);
    int rc;
#define A_(p, cnt) if ((rc = dmalloc(&(p), (cnt))) != TBL_OK) return rc
    A_(n->w, np); A_(n->g, np); A_(n->m, np); A_(n->v, np); A_(n->wb, np);
    long npk = 0;
    for (auto &c : n->L) {
        c.wpoff = npk;
        if (c.ks == 3) npk += 9L * c.cout * conv3x3_cinp(c.cin);
    }
    A_(n->wp, npk);
    const long N = n->nmax;
    long maxz = 0, maxcol = 0;
    for (auto &c : n->L) {
        const long P = N * n->hw(c.lvl);
        maxz = std::max(maxz, P * c.cout);
        maxcol = std::max(maxcol, P * c.ks * c.ks * c.cin);
    }
    A_(n->x_in, N * n->hw(0) * 4);
    A_(n->d_xin, N * n->hw(0) * 4);
    for (int i = 0; i < NCONV; ++i) {
        const long P = N * n->hw(n->L[i].lvl);
        A_(n->act[i], P * n->L[i].cout);
        A_(n->d_act[i], P * n->L[i].cout);
    }
    const int pc[3] = {B, 2 * B, 4 * B};
    for (int l = 0; l < 3; ++l) {
        A_(n->pool_out[l], N * n->hw(l + 1) * pc[l]);
        A_(n->d_pool[l], N * n->hw(l + 1) * pc[l]);
    }
    const int cc[3] = {3 * B, 6 * B, 12 * B};   // concat channels at L0, L1, L2
    for (int l = 0; l < 3; ++l) {
        A_(n->cat[l], N * n->hw(l) * cc[l]);
        A_(n->d_cat[l], N * n->hw(l) * cc[l]);
    }
    A_(n->z, maxz);
    A_(n->out, N * n->hw(0) * 2);
    A_(n->d_out, N * n->hw(0) * 2);
    A_(n->col, maxcol);
    A_(n->dcol, maxcol);
    A_(n->dzb, maxz);
    A_(n->d_loss, 1);
#undef A_
    return TBL_OK;
}

// He-normal initialisation, deterministic from the seed
void init_params(tbl_net *n, unsigned seed, std::vector<float> &h)
{
    h.assign(static_cast<size_t>(n->nparam), 0.f);
    std::mt19937 rng(seed);
    for (auto &c : n->L) {
        const float sd = std::sqrt(2.0f / static_cast<float>(c.ks * c.ks * c.cin));
        std::normal_distribution<float> nd(0.f, c.relu ? sd : 0.1f * sd);
        const long nw = static_cast<long>(c.ks) * c.ks * c.cin * c.cout;
        for (long e = 0; e < nw; ++e) h[c.woff + e] = nd(rng);
    }
}

// conv forward: input x (bf16 NHWC, cin channels at level lvl) -> act[id]
void conv_fwd(tbl_net *n, int id, const bf16 *x, int N, hipStream_t st)
{
    const Conv &c = n->L[id];
    const int h = n->H >> c.lvl, w = n->W >> c.lvl;
    const long P = static_cast<long>(N) * h * w;
    const int kk = c.ks * c.ks * c.cin;
    if (c.ks == 3) {
        conv3x3(N, h, w, c.cin, c.cout, x, n->wp + c.wpoff, n->w + c.boff, n->act[id], c.relu, st);
        return;
    }
    const bf16 *A = x;
    float *Z = (id == OUT) ? n->out : n->z;
    gemm(static_cast<int>(P), c.cout, kk, A, kk, 1, n->wb + c.woff, c.cout, 1, Z, c.cout, 1.f,
         0.f, st);
    bias_act(P, c.cout, Z, n->w + c.boff, id == OUT ? nullptr : n->act[id], c.relu, st);
}

// conv backward.  dy: fp32 gradient w.r.t. this conv's output; dx: fp32
// gradient w.r.t. its input (written, or added when acc); x: its input
void conv_bwd(tbl_net *n, int id, const bf16 *x, const float *dy, float *dx, int acc, int N,
              hipStream_t st)
{
    const Conv &c = n->L[id];
    const int h = n->H >> c.lvl, w = n->W >> c.lvl;
    const long P = static_cast<long>(N) * h * w;
    const int kk = c.ks * c.ks * c.cin;
    relu_back(P * c.cout, dy, id == OUT ? nullptr : n->act[id], n->dzb, c.relu, st);
    bias_grad(P, c.cout, n->dzb, n->g + c.boff, st);
    const bf16 *col = x;
    if (c.ks == 3) {
        im2col(N, h, w, c.cin, x, n->col, st);
        col = n->col;
    }
    // dW[kk, cout] = col^T dZ   (K = P, split over blocks)
    const long blocks = ((kk + 63) / 64) * ((c.cout + 63) / 64);
    int split = static_cast<int>(std::max(1L, std::min(1024L / std::max(1L, blocks), P / 512)));
    gemm(kk, c.cout, static_cast<int>(P), col, 1, kk, n->dzb, c.cout, 1, n->g + c.woff, c.cout,
         1.f, 0.f, st, split);
    if (!dx) return;
    if (c.ks == 3) {
        gemm(static_cast<int>(P), kk, c.cout, n->dzb, c.cout, 1, n->wb + c.woff, 1, c.cout,
             n->dcol, kk, 1.f, 0.f, st);
        col2im(N, h, w, c.cin, n->dcol, dx, acc, st);
    } else {
        gemm(static_cast<int>(P), kk, c.cout, n->dzb, c.cout, 1, n->wb + c.woff, 1, c.cout, dx, kk,
             1.f, acc ? 1.f : 0.f, st);
    }
}

void forward(tbl_net *n, int N, hipStream_t st)
{
    const int B = n->base;
    // the bf16 weights change with every Adam step and set_params: repack
    if (n->wp_stale) {
        for (auto &c : n->L)
            if (c.ks == 3) pack_conv3x3(c.cin, c.cout, n->wb + c.woff, n->wp + c.wpoff, st);
        n->wp_stale = false;
    }
    conv_fwd(n, E0A, n->x_in, N, st);
    conv_fwd(n, E0B, n->act[E0A], N, st);
    pool(N, n->H, n->W, B, n->act[E0B], n->pool_out[0], st);
    conv_fwd(n, E1A, n->pool_out[0], N, st);
    conv_fwd(n, E1B, n->act[E1A], N, st);
    pool(N, n->H >> 1, n->W >> 1, 2 * B, n->act[E1B], n->pool_out[1], st);
    conv_fwd(n, E2A, n->pool_out[1], N, st);
    conv_fwd(n, E2B, n->act[E2A], N, st);
    pool(N, n->H >> 2, n->W >> 2, 4 * B, n->act[E2B], n->pool_out[2], st);
    conv_fwd(n, E3A, n->pool_out[2], N, st);
    conv_fwd(n, E3B, n->act[E3A], N, st);
    up_concat(N, n->H >> 2, n->W >> 2, 8 * B, 4 * B, n->act[E3B], n->act[E2B], n->cat[2], st);
    conv_fwd(n, D2A, n->cat[2], N, st);
    conv_fwd(n, D2B, n->act[D2A], N, st);
    up_concat(N, n->H >> 1, n->W >> 1, 4 * B, 2 * B, n->act[D2B], n->act[E1B], n->cat[1], st);
    conv_fwd(n, D1A, n->cat[1], N, st);
    conv_fwd(n, D1B, n->act[D1A], N, st);
    up_concat(N, n->H, n->W, 2 * B, B, n->act[D1B], n->act[E0B], n->cat[0], st);
    conv_fwd(n, D0A, n->cat[0], N, st);
    conv_fwd(n, D0B, n->act[D0A], N, st);
    conv_fwd(n, OUT, n->act[D0B], N, st);
}

// gradients of all parameters from d_out (fp32, dL/d(output))
void backward(tbl_net *n, int N, bool input_grad, hipStream_t st)
{
    const int B = n->base;
    conv_bwd(n, OUT, n->act[D0B], n->d_out, n->d_act[D0B], 0, N, st);
    conv_bwd(n, D0B, n->act[D0A], n->d_act[D0B], n->d_act[D0A], 0, N, st);
    conv_bwd(n, D0A, n->cat[0], n->d_act[D0A], n->d_cat[0], 0, N, st);
    up_concat_back(N, n->H, n->W, 2 * B, B, n->d_cat[0], n->d_act[D1B], n->d_act[E0B], 0, st);
    conv_bwd(n, D1B, n->act[D1A], n->d_act[D1B], n->d_act[D1A], 0, N, st);
    conv_bwd(n, D1A, n->cat[1], n->d_act[D1A], n->d_cat[1], 0, N, st);
    up_concat_back(N, n->H >> 1, n->W >> 1, 4 * B, 2 * B, n->d_cat[1], n->d_act[D2B],
                   n->d_act[E1B], 0, st);
    conv_bwd(n, D2B, n->act[D2A], n->d_act[D2B], n->d_act[D2A], 0, N, st);
    conv_bwd(n, D2A, n->cat[2], n->d_act[D2A], n->d_cat[2], 0, N, st);
    up_concat_back(N, n->H >> 2, n->W >> 2, 8 * B, 4 * B, n->d_cat[2], n->d_act[E3B],
                   n->d_act[E2B], 0, st);
    conv_bwd(n, E3B, n->act[E3A], n->d_act[E3B], n->d_act[E3A], 0, N, st);
    conv_bwd(n, E3A, n->pool_out[2], n->d_act[E3A], n->d_pool[2], 0, N, st);
    pool_back(N, n->H >> 2, n->W >> 2, 4 * B, n->d_pool[2], n->d_act[E2B], 1, st);
    conv_bwd(n, E2B, n->act[E2A], n->d_act[E2B], n->d_act[E2A], 0, N, st);
    conv_bwd(n, E2A, n->pool_out[1], n->d_act[E2A], n->d_pool[1], 0, N, st);
    pool_back(N, n->H >> 1, n->W >> 1, 2 * B, n->d_pool[1], n->d_act[E1B], 1, st);
    conv_bwd(n, E1B, n->act[E1A], n->d_act[E1B], n->d_act[E1A], 0, N, st);
    conv_bwd(n, E1A, n->pool_out[0], n->d_act[E1A], n->d_pool[0], 0, N, st);
    pool_back(N, n->H, n->W, B, n->d_pool[0], n->d_act[E0B], 1, st);
    conv_bwd(n, E0B, n->act[E0A], n->d_act[E0B], n->d_act[E0A], 0, N, st);
    conv_bwd(n, E0A, n->x_in, n->d_act[E0A], input_grad ? n->d_xin : nullptr, 0, N, st);
}

int check_n(const tbl_net *n, int N)
{
    return (n && N >= 1 && N <= n->nmax) ? TBL_OK : TBL_ERR_ARG;
}

}  // namespace

extern "C" int tbl_net_create(int H, int W, int max_batch, unsigned seed, tbl_net **out)
{
    if (!out || H < 8 || W < 8 || H % 8 || W % 8 || max_batch < 1) return TBL_ERR_ARG;
    *out = nullptr;
    // the convolutions address an activation tensor (at most 128 bf16
    // channels per pixel) through one buffer resource: 32-bit byte offsets,
    // with 0x80000000 as the out-of-range offset of a masked load
    if (static_cast<long long>(max_batch) * H * W * 128 * 2 >= (1LL << 31)) return TBL_ERR_ARG;
    tbl_net *n = new (std::nothrow) tbl_net();
    if (!n) return TBL_ERR_ALLOC;
    n->H = H;
    n->W = W;
    n->nmax = max_batch;
    int rc = net_alloc(n);
    if (rc == TBL_OK) {
        std::vector<float> h;
        init_params(n, seed, h);
        if (hipMemcpy(n->w, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
            rc = TBL_ERR_HIP;
        if (rc == TBL_OK && hipStreamCreateWithFlags(&n->own, hipStreamDefault) != hipSuccess)
            rc = TBL_ERR_HIP;
        n->stream = n->own;
        if (rc == TBL_OK) {
            to_bf16(n->nparam, n->w, n->wb, n->stream);
            n->wp_stale = true;
            if (hipStreamSynchronize(n->stream) != hipSuccess) rc = TBL_ERR_HIP;
        }
    }
    if (rc != TBL_OK) {
        tbl_net_destroy(n);
        return rc;
    }
    *out = n;
    return TBL_OK;
}

extern "C" int tbl_net_destroy(tbl_net *n)
{
    if (!n) return TBL_ERR_ARG;
    auto fr = [](void *p) { if (p) (void)hipFree(p); };
    fr(n->w); fr(n->g); fr(n->m); fr(n->v); fr(n->wb); fr(n->wp); fr(n->x_in); fr(n->d_xin);
    for (int i = 0; i < NCONV; ++i) { fr(n->act[i]); fr(n->d_act[i]); }
    for (int l = 0; l < 3; ++l) { fr(n->pool_out[l]); fr(n->d_pool[l]); fr(n->cat[l]); fr(n->d_cat[l]); }
    fr(n->z); fr(n->out); fr(n->d_out); fr(n->col); fr(n->dcol); fr(n->dzb); fr(n->d_loss);
    if (n->own) (void)hipStreamDestroy(n->own);
    delete n;
    return TBL_OK;
}

extern "C" long tbl_net_num_params(const tbl_net *n) { return n ? n->nparam : -1; }

extern "C" int tbl_net_get_params(tbl_net *n, float *host)
{
    if (!n || !host) return TBL_ERR_ARG;
    TBL_HIP(hipStreamSynchronize(n->stream));
    TBL_HIP(hipMemcpy(host, n->w, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
    return TBL_OK;
}

extern "C" int tbl_net_set_params(tbl_net *n, const float *host)
{
    if (!n || !host) return TBL_ERR_ARG;
    TBL_HIP(hipMemcpy(n->w, host, sizeof(float) * n->nparam, hipMemcpyHostToDevice));
    to_bf16(n->nparam, n->w, n->wb, n->stream);
    n->wp_stale = true;   // the next forward repacks the 3x3 weights
    TBL_HIP(hipStreamSynchronize(n->stream));
    return TBL_OK;
}

// x: device fp32 NHWC (N, H, W, 4); y: device fp32 (N, H, W, 2)
namespace tbl {
// forward pass queued on `st` (no host synchronisation); the solver's guess
// path runs it on the solve's stream
int net_forward_async(tbl_net *n, const float *d_x, float *d_y, int N, hipStream_t st)
{
    if (check_n(n, N) || !d_x || !d_y) return TBL_ERR_ARG;
    to_bf16(static_cast<long>(N) * n->hw(0) * 4, d_x, n->x_in, st);
    forward(n, N, st);
    TBL_HIP(hipMemcpyAsync(d_y, n->out, sizeof(float) * N * n->hw(0) * 2, hipMemcpyDeviceToDevice, st));
    TBL_HIP(hipGetLastError());
    return TBL_OK;
}
}  // namespace tbl

extern "C" int tbl_net_forward(tbl_net *n, const float *d_x, float *d_y, int N)
{
    const int rc = tbl::net_forward_async(n, d_x, d_y, N, n ? n->stream : nullptr);
    if (rc != TBL_OK) return rc;
    TBL_HIP(hipStreamSynchronize(n->stream));
    return TBL_OK;
}

// forward + MSE loss + backward; gradients to host (may be NULL) and the
// gradient w.r.t. the input (fp32 NHWC, may be NULL)
extern "C" int tbl_net_gradients(tbl_net *n, const float *d_x, const float *d_t, int N, float *loss,
                                 float *host_grad, float *d_xgrad)
{
    if (check_n(n, N) || !d_x || !d_t) return TBL_ERR_ARG;
    hipStream_t st = n->stream;
    const long P0 = static_cast<long>(N) * n->hw(0);
    to_bf16(P0 * 4, d_x, n->x_in, st);
    forward(n, N, st);
    zero(1, n->d_loss, st);
    mse(P0 * 2, n->out, d_t, n->d_out, n->d_loss, st);
    backward(n, N, d_xgrad != nullptr, st);
    TBL_HIP(hipGetLastError());
    if (loss) TBL_HIP(hipMemcpyAsync(loss, n->d_loss, sizeof(float), hipMemcpyDeviceToHost, st));
    if (host_grad)
        TBL_HIP(hipMemcpyAsync(host_grad, n->g, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, st));
    if (d_xgrad)
        TBL_HIP(hipMemcpyAsync(d_xgrad, n->d_xin, sizeof(float) * P0 * 4, hipMemcpyDeviceToDevice, st));
    TBL_HIP(hipStreamSynchronize(st));
    return TBL_OK;
}

// one Adam step on a batch; *loss receives the batch loss before the step
extern "C" int tbl_net_train_step(tbl_net *n, const float *d_x, const float *d_t, int N, float lr,
                                  float *loss)
{
    if (check_n(n, N) || !d_x || !d_t || !(lr > 0)) return TBL_ERR_ARG;
    hipStream_t st = n->stream;
    const long P0 = static_cast<long>(N) * n->hw(0);
    to_bf16(P0 * 4, d_x, n->x_in, st);
    forward(n, N, st);
    zero(1, n->d_loss, st);
    mse(P0 * 2, n->out, d_t, n->d_out, n->d_loss, st);
    backward(n, N, false, st);
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, static_cast<float>(n->step));
    const float c2 = 1.f - std::pow(b2, static_cast<float>(n->step));
    adam(n->nparam, n->w, n->wb, n->g, n->m, n->v, lr, b1, b2, 1e-8f, c1, c2, st);
    n->wp_stale = true;
    TBL_HIP(hipGetLastError());
    if (loss) {
        TBL_HIP(hipMemcpyAsync(loss, n->d_loss, sizeof(float), hipMemcpyDeviceToHost, st));
        TBL_HIP(hipStreamSynchronize(st));
    }
    return TBL_OK;
}

extern "C" int tbl_net_synchronize(tbl_net *n)
{
    if (!n) return TBL_ERR_ARG;
    TBL_HIP(hipStreamSynchronize(n->stream));
    return TBL_OK;
}

// Plain GEMM on the matrix cores, exposed for its unit test:
// C[M,N] (fp32, row-major) = A[M,K] (bf16 bits, row-major) x B[K,N] (bf16, row-major)
extern "C" int tbl_mfma_gemm_bf16(int M, int N, int K, const void *d_a, const void *d_b, float *d_c,
                                  int transpose_a, int transpose_b, int splitk)
{
    if (M < 1 || N < 1 || K < 1 || !d_a || !d_b || !d_c) return TBL_ERR_ARG;
    const bf16 *A = static_cast<const bf16 *>(d_a);
    const bf16 *B = static_cast<const bf16 *>(d_b);
    // transpose_a: A stored as [K, M]; transpose_b: B stored as [N, K]
    gemm(M, N, K, A, transpose_a ? 1 : K, transpose_a ? M : 1, B, transpose_b ? 1 : N,
         transpose_b ? K : 1, d_c, N, 1.f, 0.f, nullptr, splitk);
    TBL_HIP(hipGetLastError());
    TBL_HIP(hipDeviceSynchronize());
    return TBL_OK;
}
