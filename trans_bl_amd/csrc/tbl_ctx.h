// Context object behind the tbl_ctx handle of the C ABI.
#pragma once

#include <vector>

#include "tbl_host.h"
#include "tbl_kernels.h"

namespace tbl {
struct Guess;
}

struct tbl_ctx {
    tbl_grid g{};
    tbl_coefs c{};
    tbl::Metrics m;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int K = 0, L = 0;

    std::vector<tbl::Level> lv;
    // FP64 multigrid state per level: v, source, FAS injected copy, residual
    std::vector<tbl::cx<double> *> V, S, W, R;
    // FP32 correction-scheme state per level (mixed precision); E32b is
    // the second e buffer of the out-of-place sweep (all but the coarsest)
    std::vector<tbl::cx<float> *> E32, S32, R32, E32b;
    std::vector<double *> qdy2_d;
    std::vector<float *> qdy2_f;
    std::vector<double *> lu_d;   // [K][2][5][ld] per level
    std::vector<float *> lu_f;
    std::vector<tbl::LevelArgs<double>> la_d;
    std::vector<tbl::LevelArgs<float>> la_f;
    std::vector<int> G64, G32;    // lanes per line for each level
    // halo snapshots of the fused-sweep strip smoother per level
    std::vector<tbl::cx<double> *> H64;
    std::vector<tbl::cx<float> *> H32;

    int *d_active = nullptr, *d_cycles = nullptr, *d_nactive = nullptr, *d_cycle = nullptr;
    double *d_resmax = nullptr, *d_lastres = nullptr;
    int *h_nactive = nullptr;     // pinned
    int last_need = 0;            // cycles the last solve needed (0: it did not converge)

    // derivatives
    double *d_qdy = nullptr, *d_thomas = nullptr, *d_tseg = nullptr;
    tbl::DerivArgs dargs{};

    // staging for the host-pointer entry points (user layout)
    double *d_u0 = nullptr, *d_u1 = nullptr, *d_u2 = nullptr;

    // captured V-cycle graphs (precision 0 / 1) and their parameters
    // two cached graphs per precision (the RK loop alternates two
    // tolerances); the least recently used one is replaced
    // (third kind: the mixed cycle with a full-multigrid correction)
    hipGraphExec_t graph[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    int graph_key[3][2][5] = {};
    double graph_tol[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    int graph_last[3] = {0, 0, 0};

    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // learned initial guess (optional)
    tbl::Guess *guess = nullptr;

    // routines of the RK-stage loop (tbl_loop.cpp), set up on first use:
    // lane-segmented x-line tables (64 * xR rows each) and work fields
    int xR = 0;
    double *d_xpenta = nullptr, *d_xinteg = nullptr, *d_xfilter = nullptr;
    tbl::cx<double> *d_t1 = nullptr, *d_t2 = nullptr;               // (nx, ny, K)
    tbl::cx<double> *d_lapv = nullptr, *d_dwydy = nullptr, *d_d2wy = nullptr;  // (nx, K)

    // right-hand side and RK4 step (tbl_rk4.cpp): 15 work fields of drv_gv,
    // 9 of the RK4 update (dv1, dv2, w1 of wx, wy, wz), the z-DFT twiddles
    tbl::cx<double> *d_rhs = nullptr, *d_rk = nullptr;
    double *d_tw = nullptr;
    int tw_n = 0;
};

namespace tbl {
int guess_apply(tbl_ctx *c, hipStream_t st, bool accept_all, int *fallback,
                double *guess_res_max);
int guess_free(tbl_ctx *c);
int noalloc_setup(int device);
int solve_dev(tbl_ctx *ctx, const double *wx, const double *dz, double *uy, int direct_rhs,
              double tol, const tbl_solve_opts *opts, int *cycles, double *residual,
              tbl_solve_stats *stats);
}
