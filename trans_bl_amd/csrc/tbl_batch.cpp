// Uneven batches of independent problems and the captured RK4 stage.
//
// A batch holds one context per problem (its own grid, mode count, level
// count, line shapes and stream).  One stage of the hot path -- four
// derivatives of v (d/dx, d2/dx2, d/dy, d2/dy2: the stage's derparx,
// derparxx, dery, deryy calls) followed by the v-Poisson solve of every mode
// -- is issued for all problems on their own streams, forked from and joined
// to the batch stream with events, and captured once into a single hipGraph.
// The solve inside the graph runs a fixed number of V-cycles with the
// device-side per-mode convergence masks, so no host decision is needed;
// converged modes stop exactly as in the host-driven solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "tbl_ctx.h"

using namespace tbl;

namespace tbl {
int stage_issue(tbl_ctx *c, const double *f, double *v, double *const *d, double tol,
                const tbl_solve_opts *opts);
int stage_begin(tbl_ctx *c, const double *f, double *v, double *const *d, double tol,
                const tbl_solve_opts *opts, int cycles);
int stage_cycle(tbl_ctx *c, double tol, const tbl_solve_opts *opts, int cyc);
int stage_end(tbl_ctx *c, double *v);
}

struct tbl_batch {
    std::vector<tbl_ctx *> ctx;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev;          // fork / join events
    hipGraphExec_t graph = nullptr;
    std::vector<const void *> key;       // pointers the graph was captured with
    double key_tol = 0.0;
    int key_cycles = 0, key_prec = 0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    // the converging stage: its graph (derivatives and the first cycles),
    // keyed like the fixed one plus the solver form and the cycles it holds
    // (ckey_base); cbase: the cycles the next call's graph is to hold
    hipGraphExec_t cgraph = nullptr;
    std::vector<const void *> ckey;
    double ckey_tol = 0.0;
    int ckey_form[6] = {};
    int ckey_base = 0;
    int cbase = 0;
    // each problem's active-mode count and the largest per-mode cycle count
    // (n + 1 values: batch_state_kernel into d_state, read into h_state)
    struct Probe *d_probe = nullptr;
    int *d_state = nullptr;
    int *h_state = nullptr;
};

// where batch_state_kernel reads a problem's counters
struct Probe {
    const int *nactive, *cycles;
    int K;
};

// state[p] = problem p's count of active modes at its last convergence
// check; state[n] = the largest completed-cycle count of any mode
__global__ void __launch_bounds__(256) batch_state_kernel(const Probe *pr, int n, int *state)
{
    __shared__ int need;
    if (threadIdx.x == 0) need = 0;
    __syncthreads();
    int m = 0;
    for (int p = 0; p < n; ++p) {
        if (threadIdx.x == 0) state[p] = *pr[p].nactive;
        for (int k = threadIdx.x; k < pr[p].K; k += blockDim.x) m = max(m, pr[p].cycles[k]);
    }
    atomicMax(&need, m);
    __syncthreads();
    if (threadIdx.x == 0) state[n] = need;
}

#define TRYB(x)                        \
    do {                               \
        int rc_ = (x);                 \
        if (rc_ != TBL_OK) return rc_; \
    } while (0)

extern "C" int tbl_batch_destroy(tbl_batch *b)
{
    if (!b) return TBL_ERR_ARG;
    for (auto *c : b->ctx)
        if (c) tbl_destroy(c);
    for (auto e : b->ev) (void)hipEventDestroy(e);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->cgraph) (void)hipGraphExecDestroy(b->cgraph);
    if (b->d_probe) (void)hipFree(b->d_probe);
    if (b->d_state) (void)hipFree(b->d_state);
    if (b->h_state) (void)hipHostFree(b->h_state);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    delete b;
    return TBL_OK;
}

extern "C" int tbl_batch_create(int n, const tbl_grid *grids, tbl_batch **out)
{
    if (n < 1 || !grids || !out) return TBL_ERR_ARG;
    *out = nullptr;
    tbl_batch *b = new (std::nothrow) tbl_batch();
    if (!b) return TBL_ERR_ALLOC;
    int rc = TBL_OK;
    for (int p = 0; p < n && rc == TBL_OK; ++p) {
        tbl_ctx *c = nullptr;
        rc = tbl_create(&grids[p], nullptr, &c);
        if (rc == TBL_OK) b->ctx.push_back(c);
    }
    if (rc == TBL_OK && hipStreamCreateWithFlags(&b->stream, hipStreamDefault) != hipSuccess)
        rc = TBL_ERR_HIP;
    for (int e = 0; e < 2 * n + 1 && rc == TBL_OK; ++e) {
        hipEvent_t ev;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) rc = TBL_ERR_HIP;
        else b->ev.push_back(ev);
    }
    if (rc == TBL_OK && (hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess))
        rc = TBL_ERR_HIP;
    if (rc == TBL_OK) {
        std::vector<Probe> pr;
        for (auto *c : b->ctx) pr.push_back({c->d_nactive, c->d_cycles, c->K});
        if (hipMalloc(reinterpret_cast<void **>(&b->d_probe), sizeof(Probe) * n) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&b->d_state), sizeof(int) * (n + 1)) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void **>(&b->h_state), sizeof(int) * (n + 1),
                          hipHostMallocDefault) != hipSuccess ||
            hipMemcpy(b->d_probe, pr.data(), sizeof(Probe) * n, hipMemcpyHostToDevice) != hipSuccess)
            rc = TBL_ERR_HIP;
    }
    if (rc != TBL_OK) {
        tbl_batch_destroy(b);
        return rc;
    }
    *out = b;
    return TBL_OK;
}

extern "C" tbl_ctx *tbl_batch_context(tbl_batch *b, int p)
{
    if (!b || p < 0 || p >= static_cast<int>(b->ctx.size())) return nullptr;
    return b->ctx[p];
}

// Issue the whole stage on the batch stream (fork to every problem's stream
// and join back).  f[p], v[p]: device (nx, ny, K) complex of problem p;
// d[4*p + 0..3]: its four derivative outputs.
static int issue(tbl_batch *b, const double *const *f, double *const *v, double *const *d,
                 double tol, const tbl_solve_opts *o)
{
    const int n = static_cast<int>(b->ctx.size());
    TBL_HIP(hipEventRecord(b->ev[0], b->stream));
    for (int p = 0; p < n; ++p) {
        tbl_ctx *c = b->ctx[p];
        TBL_HIP(hipStreamWaitEvent(c->stream, b->ev[0], 0));
        TRYB(stage_issue(c, f[p], v[p], d + 4 * p, tol, o));
        TBL_HIP(hipEventRecord(b->ev[1 + p], c->stream));
        TBL_HIP(hipStreamWaitEvent(b->stream, b->ev[1 + p], 0));
    }
    return TBL_OK;
}

// One RK4 stage of every problem: derivatives of v (input state), then v
// solved in place from itself as the initial guess (the reference's warm
// start) with `cycles` fixed V-cycles (convergence masked per mode).
// use_graph: capture once, replay afterwards (re-captured when pointers,
// tolerance or cycle count change).  *ms: GPU time of the stage.
extern "C" int tbl_batch_stage(tbl_batch *b, const double *const *f, double *const *v,
                               double *const *d, double tol, int cycles, int precision,
                               int use_graph, double *ms)
{
    if (!b || !f || !v || !d || cycles < 1 || precision < 0 || precision > 1) return TBL_ERR_ARG;
    const int n = static_cast<int>(b->ctx.size());
    tbl_solve_opts o{};
    o.precision = precision;
    o.max_cycles = cycles;
    std::vector<const void *> key;
    for (int p = 0; p < n; ++p) {
        key.push_back(f[p]);
        key.push_back(v[p]);
        for (int q = 0; q < 4; ++q) key.push_back(d[4 * p + q]);
    }
    TBL_HIP(hipEventRecord(b->t0, b->stream));
    if (!use_graph) {
        TRYB(issue(b, f, v, d, tol, &o));
    } else {
        if (!b->graph || key != b->key || tol != b->key_tol || cycles != b->key_cycles ||
            precision != b->key_prec) {
            if (b->graph) {
                TBL_HIP(hipGraphExecDestroy(b->graph));
                b->graph = nullptr;
            }
            hipGraph_t g;
            TBL_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = issue(b, f, v, d, tol, &o);
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            TRYB(rc);
            TBL_HIP(e);
            TBL_HIP(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
            TBL_HIP(hipGraphDestroy(g));
            b->key = key;
            b->key_tol = tol;
            b->key_cycles = cycles;
            b->key_prec = precision;
        }
        TBL_HIP(hipGraphLaunch(b->graph, b->stream));
    }
    TBL_HIP(hipEventRecord(b->t1, b->stream));
    TBL_HIP(hipEventSynchronize(b->t1));
    TBL_HIP(hipGetLastError());
    if (ms) {
        float t = 0.f;
        TBL_HIP(hipEventElapsedTime(&t, b->t0, b->t1));
        *ms = t;
    }
    return TBL_OK;
}

// The converging stage's first part on the batch stream (fork to every
// problem's stream and join back): derivatives, source, `cycles` V-cycles.
static int issue_begin(tbl_batch *b, const double *const *f, double *const *v,
                       double *const *d, double tol, const tbl_solve_opts *o, int cycles)
{
    const int n = static_cast<int>(b->ctx.size());
    TBL_HIP(hipEventRecord(b->ev[0], b->stream));
    for (int p = 0; p < n; ++p) {
        tbl_ctx *c = b->ctx[p];
        TBL_HIP(hipStreamWaitEvent(c->stream, b->ev[0], 0));
        TRYB(stage_begin(c, f[p], v[p], d + 4 * p, tol, o, std::min(cycles, o->max_cycles)));
        TBL_HIP(hipEventRecord(b->ev[1 + p], c->stream));
        TBL_HIP(hipStreamWaitEvent(b->stream, b->ev[1 + p], 0));
    }
    return TBL_OK;
}

// One RK4 stage of every problem, solved to the tolerance: the derivatives
// and the solve's first cycles replayed as one captured graph, then further
// cycles only for the problems that still have unconverged modes (one host
// check of the problems' active-mode counts per cycle), up to
// opts->max_cycles; the copy-out last.  The graph holds as many cycles as
// the previous call needed (the largest per-mode count, plus the check
// that finds it converged), the first call opts->max_cycles; it is
// re-captured when that number, the pointers, the tolerance or the solver
// form change.  Per mode the result is that of tbl_batch_stage with the
// same options and max_cycles fixed cycles.  *cycles_run: the cycles the
// slowest problem ran.
extern "C" int tbl_batch_stage_converge(tbl_batch *b, const double *const *f, double *const *v,
                                        double *const *d, double tol,
                                        const tbl_solve_opts *opts, double *ms,
                                        int *cycles_run)
{
    if (!b || !f || !v || !d || !opts) return TBL_ERR_ARG;
    if (opts->precision < 0 || opts->precision > 1 || opts->fmg < 0 || opts->fmg > 2 ||
        opts->max_cycles < 1 || opts->nu1 < 0 || opts->nu2 < 0 || opts->ncoarse < 0)
        return TBL_ERR_ARG;
    const int n = static_cast<int>(b->ctx.size());
    const int maxc = opts->max_cycles;
    std::vector<const void *> key;
    for (int p = 0; p < n; ++p) {
        key.push_back(f[p]);
        key.push_back(v[p]);
        for (int q = 0; q < 4; ++q) key.push_back(d[4 * p + q]);
    }
    const int form[6] = {opts->precision, opts->nu1, opts->nu2, opts->ncoarse, opts->fmg, maxc};
    if (b->cbase < 1 || std::memcmp(form, b->ckey_form, sizeof(form)) != 0) b->cbase = maxc;
    const int base = std::min(b->cbase, maxc);
    if (!b->cgraph || key != b->ckey || tol != b->ckey_tol ||
        std::memcmp(form, b->ckey_form, sizeof(form)) != 0 || base != b->ckey_base) {
        if (b->cgraph) {
            TBL_HIP(hipGraphExecDestroy(b->cgraph));
            b->cgraph = nullptr;
        }
        hipGraph_t g;
        TBL_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = issue_begin(b, f, v, d, tol, opts, base);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        TRYB(rc);
        TBL_HIP(e);
        TBL_HIP(hipGraphInstantiate(&b->cgraph, g, nullptr, nullptr, 0));
        TBL_HIP(hipGraphDestroy(g));
        b->ckey = key;
        b->ckey_tol = tol;
        std::memcpy(b->ckey_form, form, sizeof(form));
        b->ckey_base = base;
    }
    // (a capture is host work: *ms starts after it)
    TBL_HIP(hipEventRecord(b->t0, b->stream));
    TBL_HIP(hipGraphLaunch(b->cgraph, b->stream));
    // after the graph and after each further cycle: one read of every
    // problem's active count (and the cycles needed so far); a further cycle
    // for each problem whose last check found active modes
    int run = base;
    for (;;) {
        hipLaunchKernelGGL(batch_state_kernel, dim3(1), dim3(256), 0, b->stream, b->d_probe, n,
                           b->d_state);
        TBL_HIP(hipMemcpyAsync(b->h_state, b->d_state, sizeof(int) * (n + 1),
                               hipMemcpyDeviceToHost, b->stream));
        TBL_HIP(hipStreamSynchronize(b->stream));
        if (run >= maxc) break;
        bool any = false;
        TBL_HIP(hipEventRecord(b->ev[0], b->stream));
        for (int p = 0; p < n; ++p) {
            if (b->h_state[p] <= 0) continue;
            tbl_ctx *c = b->ctx[p];
            TBL_HIP(hipStreamWaitEvent(c->stream, b->ev[0], 0));
            TRYB(stage_cycle(c, tol, opts, run + 1));
            TBL_HIP(hipEventRecord(b->ev[1 + p], c->stream));
            TBL_HIP(hipStreamWaitEvent(b->stream, b->ev[1 + p], 0));
            any = true;
        }
        if (!any) break;
        ++run;
    }
    // the cycles this call needed (and the check that found them converged):
    // the graph of the next call holds them
    b->cbase = std::min(b->h_state[n] + 1, maxc);
    TBL_HIP(hipEventRecord(b->ev[0], b->stream));
    for (int p = 0; p < n; ++p) {
        tbl_ctx *c = b->ctx[p];
        TBL_HIP(hipStreamWaitEvent(c->stream, b->ev[0], 0));
        TRYB(stage_end(c, v[p]));
        TBL_HIP(hipEventRecord(b->ev[1 + p], c->stream));
        TBL_HIP(hipStreamWaitEvent(b->stream, b->ev[1 + p], 0));
    }
    TBL_HIP(hipEventRecord(b->t1, b->stream));
    TBL_HIP(hipEventSynchronize(b->t1));
    TBL_HIP(hipGetLastError());
    if (ms) {
        float t = 0.f;
        TBL_HIP(hipEventElapsedTime(&t, b->t0, b->t1));
        *ms = t;
    }
    if (cycles_run) *cycles_run = run;
    return TBL_OK;
}

// Per-mode results of problem p's last stage (cycles, last residual).
extern "C" int tbl_batch_results(tbl_batch *b, int p, int *cycles, double *residual)
{
    if (!b || p < 0 || p >= static_cast<int>(b->ctx.size())) return TBL_ERR_ARG;
    tbl_ctx *c = b->ctx[p];
    if (cycles) TBL_HIP(hipMemcpy(cycles, c->d_cycles, sizeof(int) * c->K, hipMemcpyDeviceToHost));
    if (residual)
        TBL_HIP(hipMemcpy(residual, c->d_lastres, sizeof(double) * c->K, hipMemcpyDeviceToHost));
    return TBL_OK;
}
