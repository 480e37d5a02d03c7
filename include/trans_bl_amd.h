/*
 * trans_bl_amd -- MI355X (gfx950) implementation of trans_bl's per-spanwise-
 * mode v-Poisson multigrid and compact finite-difference derivatives.
 *
 * C ABI: plain pointers, sizes and small structs; no C++ or torch types.
 * Complex fields are interleaved (re, im) doubles in the reference's Fortran
 * order (ptsx, jmax, kfour): x index fastest, then y, then mode
 * (comm.var:2-5).  Every call returns a tbl_status; the library keeps no
 * global state besides per-context allocations.
 *
 * Reference interfaces replaced (file:line into the reference tree):
 *   tbl_coefs_generate   pre_processing/coef_generation.f90:18-64 (coefs.bin)
 *   tbl_create           create_ctes main_par.f:1828-1892 + derivs_k
 *                        derivs_par.f:6-29 + the /coefs/ read main_par.f:852-873
 *   tbl_poi_uy(_dev)     poi_uy(dwzdx, erromax)       poisson_par.f:6-64
 *   tbl_derparx(_dev)    derparx(ddx, fc)             derivs_par.f:80-95
 *   tbl_derparxx(_dev)   derparxx(d2dx2, fc)          derivs_par.f:98-113
 *   tbl_dery(_dev)       dery(ddy, fc)                derivs_par.f:116-129
 *   tbl_deryy(_dev)      deryy(d2dy2, fc)             derivs_par.f:164-177
 *   tbl_deryfv(_dev)     deryfv(ddy, fc)              derivs_par.f:148-161
 *   tbl_checkpoint_*     escreve (data_RR.bin)        escreve_par.f:6-33
 *   tbl_baseflow_*       escreve (based_RR.bin)       baseflow2D/main_par_2D.f:578-595
 */
#ifndef TRANS_BL_AMD_H
#define TRANS_BL_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define TBL_MAX_LEVELS 8

typedef enum tbl_status {
    TBL_OK = 0,
    TBL_ERR_ARG = 1,      /* null pointer or out-of-range argument */
    TBL_ERR_GRID = 2,     /* grid not usable by the multigrid (verifica, main_par.f:60-134) */
    TBL_ERR_HIP = 3,      /* a HIP runtime call failed */
    TBL_ERR_ALLOC = 4,    /* device or host allocation failed */
    TBL_ERR_STATE = 5,    /* call not valid in the context's current state */
    TBL_ERR_NOCONV = 6    /* reserved: returned only when the caller asks for it */
} tbl_status;

/* Runtime form of the par.for PARAMETERs the path reads (par.for:41-108). */
typedef struct tbl_grid {
    int nx;          /* ptsx  (one rank)           */
    int ny;          /* jmax                       */
    int nmodes;      /* kfour                      */
    int levels;      /* msh                        */
    double dx;       /* dx                         */
    double dy0;      /* dy  (first wall spacing)   */
    double stf;      /* stf (geometric y stretch)  */
    double alpha;    /* alpha (Robin top BC)       */
    double beta;     /* beta  (kz = (k-1)*beta)    */
    double fac_y;    /* fac_y                      */
} tbl_grid;

/* COMMON /coefs/ (comm.coef:1-29), coefficient index order as in coefs.bin.
   Poisson tables are [level][coef] (Fortran sp_poi_coef(7, msh) etc.). */
typedef struct tbl_coefs {
    double fp_fd_e[6], sp_fd[9], cp_fd[8], pp_fd[9], lp_fd[7];
    double fp_sd[8], sp_sd[10], cp_sd[8], pp_sd[10], lp_sd[8];
    double sp_poi[TBL_MAX_LEVELS][7];
    double cp_poi[TBL_MAX_LEVELS][8];
    double pp_poi[TBL_MAX_LEVELS][4];
    double lp_poi[TBL_MAX_LEVELS][5];
    double w_at_w[9];   /* w_at_w_coef: wall vorticity from uy (coefs.bin record 15) */
    double dwydy[8];    /* dwydy_coef: one-sided d/dy at the wall (record 16)      */
} tbl_coefs;

/* Solver options.  Zero-initialised fields take the reference's values. */
typedef struct tbl_solve_opts {
    int precision;    /* 0: FP64 FAS cycle (poisson_par.f:33-50);
                         1: FP32 line smoothing under FP64 defect correction */
    int max_cycles;   /* 0 -> 11, the reference cap (poisson_par.f:33) */
    int nu1;          /* pre-smoothing sweeps, 0 -> 3 (itera, poisson_par.f:24)
                         for precision 0, 2 for the FP32 correction cycle */
    int nu2;          /* post-smoothing sweeps, 0 -> 1 (poisson_par.f:49) */
    int ncoarse;      /* coarsest-level sweeps, 0 -> 60 (20*itera, :45) for
                         precision 0; for the FP32 correction cycle 5 when the
                         coarsest level has at most 25 rows, else 10 */
    int use_graph;    /* 1: replay the V-cycle as a captured hipGraph */
    int guess;        /* 0: start from uy (the reference's warm start,
                            poisson_par.f:157-161); 1: start from zero;
                         2: learned initial guess (tbl_guess_*) on the interior,
                            uy's boundary values kept; falls back to uy when
                            its residual on the network's multigrid level
                            (against the restricted source, after two line
                            sweeps) exceeds 10x uy's or is not finite;
                         3: learned initial guess for every mode (no test) */
    int fmg;          /* precision 1 only (ignored for 0): 1 -- the first
                         cycle solves its correction by full multigrid (the
                         defect restricted to every level, solved on the
                         coarsest, then each finer level starts from the
                         prolonged coarser solution and runs one correction
                         V-cycle); 2 -- every cycle does; any start.  No
                         counterpart in the reference */
    int reserved[4];
} tbl_solve_opts;

/* Per-call statistics (the reference prints k, vcycles-1, sum_err per mode,
   poisson_par.f:58-60). */
typedef struct tbl_solve_stats {
    int cycles_max;       /* largest completed V-cycle count over modes */
    int cycles_total;     /* sum over modes */
    int nconverged;       /* modes whose last checked residual < erromax */
    int fallback;         /* modes where the learned guess was rejected */
    double residual_max;  /* largest last-checked max |S - L v| */
    double guess_residual_max; /* guess=2: max residual of the learned prediction
                                  on the network's level */
    double device_ms;     /* GPU time of the solve (hipEvents) */
} tbl_solve_stats;

typedef struct tbl_ctx tbl_ctx;

/* Version string of the library build. */
const char *tbl_version(void);

/* Text of the last error raised on the calling thread ("" if none). */
const char *tbl_last_error(void);

/* Coefficient tables for grid stretch stf and `levels` multigrid levels:
   the offline generator (pre_processing/derivative.f90:1762-1882) as a host
   routine (Poisson schemes in 80-bit long double).  Replaces running `coef`
   and reading coefs.bin. */
int tbl_coefs_generate(double stf, int levels, tbl_coefs *out);

/* Read the tables of the path from a coefs.bin (20 sequential unformatted
   records, coef_generation.f90:39-60; the same records main_par.f:852-873
   reads).  Rejects a file whose first record is the older 7-entry implicit
   first-point scheme instead of the 6-entry explicit one. */
int tbl_coefs_read(const char *path, int levels, tbl_coefs *out);

/* The reference's field files (Fortran sequential unformatted records, a
   4-byte length before and after each record; records above 2^31 - 1 bytes
   are refused).  Host arrays in the user layout.
   Checkpoint data_RR.bin (escreve, escreve_par.f:6-33; read back by
   main_par.f:763-767): record 1 the time step t, record 2 nfields complex
   fields (nx, ny, nmodes): ux, uy, uz, wx, wy, wz (nfields 6) and th
   (nfields 7, my_form 2). */
int tbl_checkpoint_write(const char *path, int nx, int ny, int nmodes, int t, int nfields,
                         const double *const *fields);
int tbl_checkpoint_read(const char *path, int nx, int ny, int nmodes, int *t, int nfields,
                        double *const *fields);
/* Base flow based_RR.bin (baseflow2D/main_par_2D.f:578-595; read by
   main_par.f:847-850): one record of ux, uy, wz, real (nx, ny) each. */
int tbl_baseflow_write(const char *path, int nx, int ny, const double *ux, const double *uy,
                       const double *wz);
int tbl_baseflow_read(const char *path, int nx, int ny, double *ux, double *uy, double *wz);

/* Create a context on the current HIP device: checks the grid like
   verifica (main_par.f:60-134), builds the multigrid metrics (create_ctes),
   the derivative LHS (derivs_k) and the line-LU factors, and allocates the
   device work arrays.  coefs == NULL uses tbl_coefs_generate(grid->stf). */
int tbl_create(const tbl_grid *grid, const tbl_coefs *coefs, tbl_ctx **out);
int tbl_destroy(tbl_ctx *ctx);

/* Stream for all following work of the context (hipStream_t as void*).
   NULL selects the context's own stream, a blocking stream: it is ordered
   after work the caller queued on the legacy default stream. */
int tbl_set_stream(tbl_ctx *ctx, void *hip_stream);

/* Wait until all work the context queued on its stream has finished. */
int tbl_synchronize(tbl_ctx *ctx);

/* poi_uy on host arrays (poisson_par.f:6-64): wx, dwzdx, uy are complex
   (nx, ny, nmodes); uy is the initial guess on entry and holds the solution
   on i=2..nx-1, j=2..ny on exit (other entries untouched).  cycles and
   residual (length nmodes, may be NULL) receive each mode's completed
   V-cycle count and last checked max residual.  Synchronous. */
int tbl_poi_uy(tbl_ctx *ctx, const double *wx, const double *dwzdx, double *uy,
               double erromax, const tbl_solve_opts *opts, int *cycles,
               double *residual, tbl_solve_stats *stats);

/* Same on device pointers (inputs resident in HBM).  wx may be NULL (treated
   as zero).  Returns after the solve has finished on the context's stream. */
int tbl_poi_uy_dev(tbl_ctx *ctx, const double *d_wx, const double *d_dwzdx,
                   double *d_uy, double erromax, const tbl_solve_opts *opts,
                   int *cycles, double *residual, tbl_solve_stats *stats);

/* Solve L v = B f for every mode (the v-Poisson with wx = 0, dwzdx = -f):
   same as tbl_poi_uy_dev with d_f = -dwzdx.  Device pointers. */
int tbl_solve_rhs_dev(tbl_ctx *ctx, const double *d_f, double *d_uy, double erromax,
                      const tbl_solve_opts *opts, int *cycles, double *residual,
                      tbl_solve_stats *stats);

/* Compact derivatives on host arrays, complex (nx, ny, nmodes). */
int tbl_derparx(tbl_ctx *ctx, const double *fc, double *ddx);
int tbl_derparxx(tbl_ctx *ctx, const double *fc, double *d2dx2);
int tbl_dery(tbl_ctx *ctx, const double *fc, double *ddy);
int tbl_deryy(tbl_ctx *ctx, const double *fc, double *d2dy2);
int tbl_deryfv(tbl_ctx *ctx, const double *fc, double *ddy);

/* Device-pointer forms, asynchronous on the context's stream.
   kind: 0 derparx, 1 derparxx, 2 dery, 3 deryy, 4 deryfv. */
int tbl_deriv_dev(tbl_ctx *ctx, int kind, const double *d_fc, double *d_out);

/* The four derivatives of one field in two passes over it (derparx,
   derparxx, dery, deryy of the same fc, as drv_gv calls them,
   main_par.f:1364-1377; derivs_par.f:80-129, 164-177).  Device pointers,
   asynchronous on the context's stream; any output may be NULL. */
int tbl_derivs_dev(tbl_ctx *ctx, const double *d_fc, double *d_ddx, double *d_d2dx2,
                   double *d_ddy, double *d_d2dy2);

/* Apply the fine-grid operator: out = S - L v on i=2..nx-1, j=2..ny
   (defect with si=-1, poisson_par.f:543-793) for every mode, with
   S = B(kb*wx - dwzdx) (function_source, poisson_par.f:111-164).  Device
   pointers; out entries outside that range are set to zero.  Also returns
   each mode's max |out| in residual (host, may be NULL). */
int tbl_residual_dev(tbl_ctx *ctx, const double *d_wx, const double *d_dwzdx,
                     const double *d_uy, double *d_out, double *residual);

/* Same as tbl_residual_dev, allocating nothing: no reference counterpart;
   for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  Its work
   fields (three FP64 complex fields of the finest level and one double per
   mode, each 256-byte aligned) live in a fixed 64 MiB device array of the
   library, which every GPU the library runs on holds until the process
   ends; grids that need more are rejected with TBL_ERR_ARG.  All calls on
   one GPU share that array: calls that overlap on one GPU race. */
int tbl_residual_dev_noalloc(tbl_ctx *ctx, const double *d_wx, const double *d_dwzdx,
                             const double *d_uy, double *d_out, double *residual);

/* ---- the rest of the RK-stage update: `loop` (main_par.f:1762-1785) --------
   Device pointers, user layout (complex (nx, ny, nmodes), x fastest);
   (nx, K) arrays are complex (ptsx, kfour).  my_form 1, or 2 with a th
   field (NULL th: my_form 1).  Outputs keep column 1 (x inflow), as the
   reference's routines at rank 0 do.  Needs nx >= 8, ny >= 8. */

/* poi_ux (loop_par.f:428-556): duydy = deryfv(uy) (out); ux of mode 1 by
   continuity on rows 2..ny-1, of modes >= 2 by the x-direction compact
   Poisson equation (lhs5 258-299, ludecomp/lusolver penta_par.f:6-208). */
int tbl_poi_ux_dev(tbl_ctx *ctx, const double *d_uy, const double *d_wy, double *d_ux,
                   double *d_duydy);
/* poi_uz (loop_par.f:559-688): modes >= 2, duydy from tbl_poi_ux_dev. */
int tbl_poi_uz_dev(tbl_ctx *ctx, const double *d_wy, const double *d_duydy, double *d_uz);
/* wx_wall + der2wydxdy (loop_par.f:174-255, 300-361): wall row of wx for
   modes >= 2; lapv (nx, K) out, d2uydx2 (nx, K) in (COMMON /derw/). */
int tbl_wx_wall_dev(tbl_ctx *ctx, const double *d_uy, const double *d_wy,
                    const double *d_d2uydx2, double *d_wx, double *d_lapv);
/* wz_wall (loop_par.f:364-425): wall row of wz for every mode. */
int tbl_wz_wall_dev(tbl_ctx *ctx, const double *d_wx, const double *d_lapv, double *d_wz);
/* bzone (loop_par.f:6-45, rows 2..ny; wall_row 0) or bzonew (691-725, row 1;
   wall_row 1): multiply by the buffer-zone table bdfc (real (2, nx), COMMON
   /bd/); th (my_form 2) only on rows 2..ny. */
int tbl_bzone_dev(tbl_ctx *ctx, const double *d_bdfc, int wall_row, double *d_wx, double *d_wy,
                  double *d_wz, double *d_th);
/* filter_trid (filter_par.f:144-332): compact x filter of wx, wy, wz. */
int tbl_filter_trid_dev(tbl_ctx *ctx, double *d_wx, double *d_wy, double *d_wz);
/* outuy (loop_par.f:48-119, my_form 1 and 2): dwzdx = derparx(wz) (out),
   uy(nx, 2..ny, :) = 0. */
int tbl_outuy_dev(tbl_ctx *ctx, const double *d_wz, double *d_uy, double *d_dwzdx);
/* loop(erro) (main_par.f:1762-1785): filter_trid when erro < 1e-5, bzone,
   outuy, poi_uy (to erro, with opts; cycles/residual/stats as
   tbl_poi_uy_dev), poi_ux, poi_uz, wx_wall, wz_wall, bzonew. */
int tbl_loop_dev(tbl_ctx *ctx, double erro, const double *d_bdfc, const double *d_d2uydx2,
                 double *d_ux, double *d_uy, double *d_uz, double *d_wx, double *d_wy,
                 double *d_wz, double *d_th, const tbl_solve_opts *opts, int *cycles,
                 double *residual, tbl_solve_stats *stats);
/* Same with host arrays (the Fortran drop-in's `loop`): every field is
   copied to the GPU and back. */
int tbl_loop(tbl_ctx *ctx, double erro, const double *bdfc, const double *d2uydx2, double *ux,
             double *uy, double *uz, double *wx, double *wy, double *wz, double *th,
             const tbl_solve_opts *opts, int *cycles, double *residual, tbl_solve_stats *stats);
/* The buffer-zone table bdfc (2, nx) as init_gv builds it
   (main_par.f:738-755): 1-based i0, i3, i4 from par.for. */
int tbl_bdfc_generate(int nx, int i0, int i3, int i4, double *bdfc);

/* ---- right-hand side and RK4 step (Goertler case, my_form 1) ---------------
   Base flow and curvature term: COMMON /blas/ uxb, uyb, wzb and /vc/ varg,
   dvargdx, each a real (nx, ny) device array. */
typedef struct tbl_base {
    const double *uxb, *uyb, *wzb, *varg, *dvargdx;
} tbl_base;

/* kphys 0: lterms_gv (nlterms_par.f:89-112, the reference's active call);
   otherwise nlterms_gv (115-180) with kphys points in z (a power of two,
   >= 2 nmodes): a, b, c, d out. */
int tbl_terms_dev(tbl_ctx *ctx, int kphys, const tbl_base *base, const double *d_ux,
                  const double *d_uy, const double *d_uz, const double *d_wx,
                  const double *d_wy, const double *d_wz, double *d_a, double *d_b, double *d_c,
                  double *d_d);
/* drv_gv (main_par.f:1339-1403): dvx, dvy, dvz on rows 2..ny (row 1 not
   written); Re the Reynolds number of par.for. */
int tbl_drv_gv_dev(tbl_ctx *ctx, double Re, int kphys, const tbl_base *base,
                   const double *d_ux, const double *d_uy, const double *d_uz,
                   const double *d_wx, const double *d_wy, const double *d_wz, double *d_dvx,
                   double *d_dvy, double *d_dvz);
/* Same with host arrays (the Fortran drop-in's drv_gv, main_par.f:1339):
   the base flow and curvature term of COMMON /blas/ and /vc/ as real
   (nx, ny) arrays; every array is copied to the GPU and the three results
   back (their row 1, which drv_gv does not write, keeps the caller's
   values). */
int tbl_drv_gv(tbl_ctx *ctx, double Re, int kphys, const double *uxb, const double *uyb,
               const double *wzb, const double *varg, const double *dvargdx, const double *ux,
               const double *uy, const double *uz, const double *wx, const double *wy,
               const double *wz, double *dvx, double *dvy, double *dvz);
/* Flat-plate form (my_form 0, main_par.f:155-230).  kphys 0: lterms_fi
   (nlterms_par.f:6-38; base->uxb, uyb, wzb used); otherwise nlterms_fi
   (41-86, the reference's active call; base may be NULL) with kphys points
   in z (a power of two, >= 2 nmodes): a, b, c out. */
int tbl_terms_fi_dev(tbl_ctx *ctx, int kphys, const tbl_base *base, const double *d_ux,
                     const double *d_uy, const double *d_uz, const double *d_wx,
                     const double *d_wy, const double *d_wz, double *d_a, double *d_b,
                     double *d_c);
/* drv_fi (main_par.f:1264-1336): dvx, dvy, dvz on rows 2..ny (row 1 not
   written); base as tbl_terms_fi_dev. */
int tbl_drv_fi_dev(tbl_ctx *ctx, double Re, int kphys, const tbl_base *base,
                   const double *d_ux, const double *d_uy, const double *d_uz,
                   const double *d_wx, const double *d_wy, const double *d_wz, double *d_dvx,
                   double *d_dvy, double *d_dvz);
/* Same with host arrays (the Fortran drop-in's drv_fi): uxb, uyb, wzb real
   (nx, ny), used (and required) for kphys 0 only; row 1 of the results
   keeps the caller's values. */
int tbl_drv_fi(tbl_ctx *ctx, double Re, int kphys, const double *uxb, const double *uyb,
               const double *wzb, const double *ux, const double *uy, const double *uz,
               const double *wx, const double *wy, const double *wz, double *dvx, double *dvy,
               double *dvz);
/* One RK4 time step (main_par.f:266-349): four drv_gv + update stages
   (rows 2..ny-1, columns 3..nx), each followed by loop (erro 1e-5, 1e-5,
   1e-5, 1e-6).  The disturbance of gv_pert (main_par.f:1577-1610) before
   the loops of stages 1 and 3: for pert_i0 <= i <= pert_i1 (1-based) and
   modes >= 2, uy(i,1,k) = pert_uy(i,k) and d2uydx2(i,k) = pert_d2(i,k)
   ((nx, K) arrays; NULL pert_uy: none).  cycles / residual: 4 x nmodes
   (per stage), or NULL. */
int tbl_rk4_step_dev(tbl_ctx *ctx, double Re, double dt, int kphys, const tbl_base *base,
                     const double *d_bdfc, double *d_d2uydx2, const double *d_pert_uy,
                     const double *d_pert_d2, int pert_i0, int pert_i1, double *d_ux,
                     double *d_uy, double *d_uz, double *d_wx, double *d_wy, double *d_wz,
                     double *d_th, const tbl_solve_opts *opts, int *cycles, double *residual);

/* Line-solve shape of one multigrid level's smoother: `lanes` (16, 32 or
   64) lanes of a wavefront share one y-line; a negative value selects the
   same shape with LDS-staged, coalesced loads and stores; 2000 + lanes
   stages only the x-direction terms in LDS and reads v and the source
   lane-contiguous; 1000 + lanes selects the fused-sweep strip smoother (a
   whole sweep per launch, one workgroup per strip of columns in LDS; not
   on the coarsest level).  On the coarsest level lanes = 1 (the default)
   selects the sequential one-launch kernels and any lane form the G-lane
   scan kernel; its FP32 shape is unused (that solve is FP64).
   All forms give the same colour-ordered sweep.  prec 0 = FP64 smoother,
   1 = FP32 smoother.  Chosen by the autotuner; no reference counterpart
   (the reference solves each line sequentially, penta_par.f:367-404). */
int tbl_set_line_shape(tbl_ctx *ctx, int prec, int level, int lanes);
int tbl_get_line_shape(tbl_ctx *ctx, int prec, int level);

/* Measurement hook (no reference counterpart): run `sweeps` line-relaxation
   sweeps of one level on the context's current multigrid state, bracketed
   by hipEvents on the context's stream; *ms receives the elapsed time and
   *launches the number of smoother kernel launches it made.  prec 0 times
   the FP64 smoother, 1 the FP32 one.  On the coarsest level it times that
   level's solve as the V-cycle of that precision runs it (all sweeps in one
   launch). */
int tbl_time_smoother(tbl_ctx *ctx, int prec, int level, int sweeps, double *ms,
                      int *launches);

/* ---- uneven batches and the captured RK4 stage (no reference counterpart:
   the reference runs one compile-time grid per executable) ---------------- */
typedef struct tbl_batch tbl_batch;

/* One context per problem (grids may differ in nx, ny, modes, levels). */
int tbl_batch_create(int n, const tbl_grid *grids, tbl_batch **out);
int tbl_batch_destroy(tbl_batch *batch);
/* Context of problem p (for line shapes, options); owned by the batch. */
tbl_ctx *tbl_batch_context(tbl_batch *batch, int p);
/* One RK4 stage of every problem, concurrently: derparx, derparxx, dery,
   deryy of v (into d[4p+0..3]), then the v-Poisson solve of all modes warm-
   started from v and written back to v, with `cycles` V-cycles (modes stop
   at the tolerance through device-side masks).  Device pointers: f[p], v[p],
   d[4p+q] are complex (nx_p, ny_p, K_p).  use_graph: capture the whole stage
   once as one hipGraph and replay it.  *ms: GPU time of the stage. */
int tbl_batch_stage(tbl_batch *batch, const double *const *f, double *const *v,
                    double *const *d, double tol, int cycles, int precision, int use_graph,
                    double *ms);
/* The same stage solved to the tolerance instead of for a fixed cycle count
   (no counterpart in the reference): the derivatives and the solve's first
   V-cycles are replayed as one captured hipGraph, then further cycles run
   only for the problems that still have unconverged modes (a host check of
   the problems' active-mode counts per cycle), up to opts->max_cycles.  The
   graph holds as many cycles as the previous call needed (the first call:
   opts->max_cycles), so a sequence of similar stages launches no masked
   cycles.  opts: precision, nu1, nu2, ncoarse, fmg as for tbl_poi_uy_dev,
   max_cycles >= 1 the cap (use_graph and guess are not used: the solve
   starts from v).  Per mode the result equals tbl_batch_stage's with the
   same options and max_cycles cycles.  *cycles_run: cycles of the slowest
   problem; *ms: time of the stage on the device from the graph's launch,
   host checks included (a re-capture is not). */
int tbl_batch_stage_converge(tbl_batch *batch, const double *const *f, double *const *v,
                             double *const *d, double tol, const tbl_solve_opts *opts,
                             double *ms, int *cycles_run);
/* Per-mode completed V-cycles and last residual of problem p's last stage. */
int tbl_batch_results(tbl_batch *batch, int p, int *cycles, double *residual);

/* ---- learned initial guess (no reference counterpart: the reference
   warm-starts from the previous stage, poisson_par.f:157-161) ------------ */
typedef struct tbl_net tbl_net;

/* BF16 U-Net (4 levels, base 32 channels, 3x3) for H x W images (multiples
   of 8), batches up to max_batch; He-initialised from seed. */
int tbl_net_create(int H, int W, int max_batch, unsigned seed, tbl_net **out);
int tbl_net_destroy(tbl_net *net);
long tbl_net_num_params(const tbl_net *net);
int tbl_net_get_params(tbl_net *net, float *host);
int tbl_net_set_params(tbl_net *net, const float *host);
/* x: device fp32 (N, H, W, 4); y: device fp32 (N, H, W, 2) */
int tbl_net_forward(tbl_net *net, const float *d_x, float *d_y, int N);
/* forward + mean-squared-error against t + backward: loss, parameter
   gradients (host, may be NULL), input gradient (device, may be NULL) */
int tbl_net_gradients(tbl_net *net, const float *d_x, const float *d_t, int N, float *loss,
                      float *host_grad, float *d_xgrad);
/* one Adam step (fp32 master weights, bf16 compute copy) */
int tbl_net_train_step(tbl_net *net, const float *d_x, const float *d_t, int N, float lr,
                       float *loss);
int tbl_net_synchronize(tbl_net *net);
/* C = A x B on the matrix cores (bf16 in, fp32 out), row-major; with
   transpose_a A is stored [K, M], with transpose_b B is stored [N, K]. */
int tbl_mfma_gemm_bf16(int M, int N, int K, const void *d_a, const void *d_b, float *d_c,
                       int transpose_a, int transpose_b, int splitk);

/* Attach a network to a context; it predicts v on multigrid level `level`
   (opts.guess = 2 or 3 then starts each mode from its prolongated
   prediction; see tbl_solve_opts.guess).  The network must stay alive while
   attached: destroy the context or detach (net == NULL) first. */
int tbl_guess_attach(tbl_ctx *ctx, tbl_net *net, int level);
int tbl_guess_shape(tbl_ctx *ctx, int *H, int *W);
/* Training pairs from solved modes: f and v device arrays (user layout);
   writes the network's input (nmodes, H, W, 4) and target (nmodes, H, W, 2). */
int tbl_guess_pairs(tbl_ctx *ctx, const double *d_f, const double *d_v, float *d_x, float *d_t);

#ifdef __cplusplus
}
#endif

#endif /* TRANS_BL_AMD_H */
