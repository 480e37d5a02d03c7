// The residual entry point without device allocation: the same kernels as
// tbl_residual_dev, with its work fields (the source S, v and the defect in
// the multigrid layout, and the per-mode maxima) in a fixed device array of
// the library instead of the context's buffers.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "tbl_ctx.h"

using namespace tbl;

namespace {

// 64 MiB, held by every GPU the library runs on until the process ends
constexpr size_t kScratchBytes = size_t{64} << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void *g_ptr[kMaxDevices] = {};
std::mutex g_mu;

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

namespace tbl {

// The device address of this device's `scratch` (the host-side name is a
// stand-in, not a device address).  Called by tbl_create with the context's
// device current; once per device.
int noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return TBL_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return TBL_OK;
    void *p = nullptr;
    TBL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_ptr[device] = p;
    return TBL_OK;
}

}  // namespace tbl

extern "C" int tbl_residual_dev_noalloc(tbl_ctx *c, const double *d_wx, const double *d_dwzdx,
                                        const double *d_uy, double *d_out, double *residual)
{
    if (!c || !d_dwzdx || !d_uy || !d_out) return TBL_ERR_ARG;
    if (c->device < 0 || c->device >= kMaxDevices) return TBL_ERR_ARG;
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        base = static_cast<unsigned char *>(g_ptr[c->device]);
    }
    if (!base) {
        tbl_set_last_error("scratch array not set up on this device", __FILE__, __LINE__);
        return TBL_ERR_STATE;
    }
    const int K = c->K;
    const Level &F = c->lv[0];
    const size_t field = align256(F.plane * K * sizeof(cx<double>));
    const size_t need = 3 * field + align256(K * sizeof(double));
    if (need > kScratchBytes) {
        char msg[160];
        std::snprintf(msg, sizeof(msg),
                      "grid needs %zu bytes of work space, above the %zu-byte scratch array", need,
                      kScratchBytes);
        tbl_set_last_error(msg, __FILE__, __LINE__);
        return TBL_ERR_ARG;
    }
    cx<double> *S = reinterpret_cast<cx<double> *>(base);
    cx<double> *V = reinterpret_cast<cx<double> *>(base + field);
    cx<double> *R = reinterpret_cast<cx<double> *>(base + 2 * field);
    double *resmax = reinterpret_cast<double *>(base + 3 * field);
    hipStream_t st = c->stream;
    // the array keeps the previous call's contents: S and V are written
    // whole by the gather, R and the maxima are cleared here (fill kernels)
    gather_source(c->la_d[0], d_wx, d_dwzdx, d_uy, c->g.beta, S, V, 0, st);
    fill(resmax, K, 0.0, st);
    fill(reinterpret_cast<double *>(R), F.plane * K * 2, 0.0, st);
    defect<double, double>(c->la_d[0], V, S, R, -1.0, nullptr, resmax, st);
    fill(d_out, static_cast<size_t>(c->g.nx) * c->g.ny * K * 2, 0.0, st);
    scatter(c->g.nx, c->g.ny, K, F, R, d_out, st);
    TBL_HIP(hipGetLastError());
    if (residual)
        TBL_HIP(hipMemcpyAsync(residual, resmax, sizeof(double) * K, hipMemcpyDeviceToHost, st));
    TBL_HIP(hipStreamSynchronize(st));
    return TBL_OK;
}
