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
};

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
