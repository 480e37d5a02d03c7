// Kernel argument blocks and launcher declarations of trans_bl_amd.
#pragma once

#include "tbl_common.h"

namespace tbl {

// fields of the line-LU table per (mode, centre/border), each ld long:
// l1, l2 (forward multipliers), u1, u2 (super-diagonals of U), pivot,
// -u1/pivot, -u2/pivot, 1/pivot
constexpr int kLuFields = 8;

// Per-level constants of the Poisson operator (poisson_par.f:229-793)
template <typename T>
struct LevelArgs {
    int nx, ny, ld, K;
    int lus;                      // row stride of the line-LU tables (lu_stride)
    size_t plane;                 // nx * ld
    T sp[7], cp[8], pp[4], lp[5]; // sp_poi_coef(:,l) ... lp_poi_coef(:,l)
    T qx;                         // 1 / (12 dx_l^2 fac_y)
    T dx2, fac_y, alpha, beta;
    T sqq_top;                    // sqrt(v_qdy2(ny_l-2, l))
    const T *qdy2;                // v_qdy2(:, l) on the device
};

inline Level make_level(int nx, int ny, double dx2)
{
    Level L;
    L.nx = nx;
    L.ny = ny;
    L.ld = round_up(ny, 8);
    L.dx2 = dx2;
    L.plane = static_cast<size_t>(nx) * L.ld;
    return L;
}

int line_rows_for(int G, int ny);

// Row stride of a level's line-LU tables: at least ny and every line shape's
// 1 + G*R (its lanes hold rows 1..G*R), so a lane's R rows are always inside
// the table; rows past ny hold the identity (zero factors, unit pivot and
// 1/pivot).
int lu_stride(int ny);

// Fused-sweep strip smoother: line shape kStripShape + G selects it (G lanes
// per line); halo of kStripHalo columns each side of a strip.
constexpr int kStripShape = 1000;
// Colour-phase smoother with only the x terms staged in LDS: line shape
// kXStagedShape + G (G lanes per line).
constexpr int kXStagedShape = 2000;

// Coarsest level only: shape kCoarseSeqShape selects the sequential kernels
// (one or two threads per line, smooth_coarsest_2s/_reg); any G form there
// selects the G-lane scan kernel (smooth_coarsest_kernel).
constexpr int kCoarseSeqShape = 1;
// lanes for a colour-phase sweep on a level whose shape code may be the
// sequential coarsest form
inline int sweep_shape(int code) { return code == kCoarseSeqShape ? 16 : code; }

// One-launch, out-of-place sweep (smooth_sweep_kernel): line shape
// kSweepShape + 100 * wsel + G, G lanes per line, wsel the strip-width
// selector (0 = default width).
constexpr int kSweepShape = 3000;
inline bool shape_is_sweep(int code) { return code > kSweepShape && code < kSweepShape + 600; }
inline int sweep_wsel(int code) { return (code - kSweepShape) / 100; }

// Line-shape codes: G, -G, kStripShape + G, kXStagedShape + G or
// kSweepShape + 100 * wsel + G.
inline bool shape_is_strip(int code) { return code > kStripShape && code < kXStagedShape; }
inline int shape_lanes(int code)
{
    return shape_is_sweep(code) ? (code - kSweepShape) % 100
           : code > kXStagedShape ? code - kXStagedShape
           : code > kStripShape ? code - kStripShape
           : code < 0           ? -code
                                : code;
}
constexpr int kStripHalo = 6;
struct StripGeom {
    int width, nstrips;   // strip width in interior columns, number of strips
};
// nstrips == 0: the level's columns do not fit the workgroup's LDS
StripGeom strip_geom(int nx, int ny, int K, size_t esize);
// snapshot buffer of the strip smoother, in complex elements
inline size_t strip_halo_elems(const StripGeom &g, int ld, int K)
{
    return g.nstrips > 1 ? static_cast<size_t>(K) * (g.nstrips - 1) * 2 * kStripHalo * ld : 0;
}
int set_kernel_attributes();
int sweep_kernel_attributes();
int phase_kernel_attributes();
int strip_kernel_attributes();
int coarse_kernel_attributes();

// zero_start: the sweeps start from v = 0 whatever v's memory holds (the
// first sweep reads no column it has not relaxed itself); vup: the last
// sweep also adds each relaxed value to this FP64 field of the same layout
// (vup += v on rows 1.., interior columns -- the correction cycle's update)
// and does not store its last colour's lines into V, which nothing reads
// afterwards (V is left incomplete).  Staged shapes (G < 0) only, vup FP32
// only; any other use returns TBL_ERR_ARG.
template <typename T>
int smooth_sweeps(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps, T rf,
                  const int *active, int G, cx<T> *halo, hipStream_t st, bool zero_start = false,
                  cx<double> *vup = nullptr);
// `sweeps` one-launch sweeps alternating between Va and Vb (out of place),
// starting from Va, or from zero with zero_start; *cur receives the buffer
// that holds the result.  vup: as for smooth_sweeps (FP32 only).
template <typename T>
int sweep_sweeps(const LevelArgs<T> &L, cx<T> *Va, cx<T> *Vb, const cx<T> *S, const T *lu,
                 int sweeps, T rf, const int *active, int G, int wsel, hipStream_t st,
                 bool zero_start, cx<double> *vup, cx<T> **cur);
int sweep_strip_width(int nx, int K, int wsel);
size_t sweep_lds(int ld, size_t esize, int G);   // dynamic LDS of one workgroup
template <typename T>
int phase_sweeps(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps, T rf,
                 const int *active, int G, hipStream_t st, bool zero_start = false,
                 cx<double> *vup = nullptr);
template <typename T>
int strip_sweeps(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps, T rf,
                 const int *active, int G, cx<T> *halo, hipStream_t st);
template <typename T>
int smooth_coarsest(const LevelArgs<T> &L, cx<T> *V, const cx<T> *S, const T *lu, int sweeps,
                    T rf, const int *active, int G, hipStream_t st);
template <typename T, typename TO>
void defect(const LevelArgs<T> &L, const cx<T> *V, const cx<T> *S, cx<TO> *out, T si,
            const int *active, double *resmax, hipStream_t st);
// Sc (the coarser level Cl's source) = the full-weighting restriction of
// the defect S - L v of level L, in one pass (the defect is not stored)
template <typename T>
void defect_restrict(const LevelArgs<T> &L, const Level &Cl, const cx<T> *V, const cx<T> *S,
                     cx<T> *Sc, const int *active, hipStream_t st);
template <typename T>
void restrict_level(const Level &F, const Level &Cl, int K, const cx<T> *Vf, const cx<T> *Rf,
                    cx<T> *Vc, cx<T> *Wc, cx<T> *Sc, int fas, const int *active, hipStream_t st);
template <typename T>
void prolong_level(const Level &F, const Level &Cl, int K, cx<T> *Vf, const cx<T> *Vc,
                   const cx<T> *Wc, int fas, const int *active, hipStream_t st);
// user layout -> internal layout and S = B g (function_source) in one kernel
void gather_source(const LevelArgs<double> &L, const double *wx, const double *dz,
                   const double *uy, double beta, cx<double> *S, cx<double> *V, int direct_rhs,
                   hipStream_t st);
void scatter(int nx, int ny, int K, const Level &F, const cx<double> *V, double *uy,
             hipStream_t st);
void decide(int K, const double *resmax, double tol, int cycle, int *active, int *cycles,
            double *lastres, int *nactive, hipStream_t st);
void fill(double *p, size_t n, double v, hipStream_t st);
void fill_int(int *p, int n, int v, hipStream_t st);
void update(const Level &F, int K, cx<double> *V, const cx<float> *E, const int *active,
            hipStream_t st);
template <typename TA, typename TB>
void convert(size_t n, const cx<TA> *a, cx<TB> *b, int zero, hipStream_t st);

// ---- derivatives (user layout) ------------------------------------------------
struct DerivArgs {
    int nx, ny, K;
    double dx, dxx;
    // y-scheme coefficients (COMMON /coefs/)
    double fp_e[6], sp_fd[9], cp_fd[8], pp_fd[9], lp_fd[7];
    double fp_sd[8], sp_sd[10], cp_sd[8], pp_sd[10], lp_sd[8];
    const double *qdy;     // v_qdy
    const double *qdy2;    // v_qdy2(:,1)
    // Thomas factors per scheme: a, bet, gam, 1/bet (kind-major, stride nmax)
    const double *thomas;  // [kind][kThomasFields][nmax]
    int nmax;
    // the same factors in lane-segment order for the LDS-staged kernels:
    // x kinds 0, 1 ([kind][a, 1/bet, next gam][xR][64]), then y kinds 2, 3, 4
    // ([kind][3][yR][yG]); xR / yR = 0: that direction takes the long-line
    // kernel
    const double *tseg;
    int xR, yG, yR;
};
constexpr int kThomasFields = 4;
void deriv_shapes(int nx, int ny, int *xR, int *yG, int *yR);

int deriv(const DerivArgs &A, int kind, const double *f, double *out, hipStream_t st);
// d/dx, d2/dx2, d/dy, d2/dy2 of one field (any output may be null): two
// launches, each reading f once
int derivs4(const DerivArgs &A, const double *f, double *dx, double *dxx, double *dy, double *dyy,
            hipStream_t st);
int deriv_kernel_attributes();

}  // namespace tbl
