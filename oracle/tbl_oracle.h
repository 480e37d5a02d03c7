/*
 * CPU oracle for the trans_bl v-Poisson / compact-derivative hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's FP64
 * algorithm, used by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline as the reference that results are compared against.  The
 * product never links or calls it.
 *
 * Arrays are Fortran order with the first index fastest, complex values
 * interleaved (re, im) -- the reference's complex*16 (ptsx, jmax, kfour)
 * layout (comm.var:2-5).  Indices below are 0-based; comments quote the
 * reference's 1-based indices.
 */
#ifndef TBL_ORACLE_H
#define TBL_ORACLE_H

#include <complex.h>
#include <stddef.h>

#define OR_MAXLVL 8

typedef double _Complex or_cplx;

typedef struct or_params {
    int nx, ny, nk, msh;          /* ptsx, jmax, kfour, msh (par.for:71-108) */
    double dx, dy0, stf;          /* dx, dy, stf (par.for:71-83) */
    double alpha, beta, fac_y;    /* par.for:41, 54, 63 */
    /* COMMON /coefs/ (comm.coef:1-29) */
    double fp_fd_e[6], sp_fd[9], cp_fd[8], pp_fd[9], lp_fd[7];
    double fp_sd[8], sp_sd[10], cp_sd[8], pp_sd[10], lp_sd[8];
    double sp_poi[OR_MAXLVL][7], cp_poi[OR_MAXLVL][8];
    double pp_poi[OR_MAXLVL][4], lp_poi[OR_MAXLVL][5];
    double w_at_w[9], dwydy[8];   /* w_at_w_coef, dwydy_coef (comm.coef:13) */
} or_params;

/* create_ctes (main_par.f:1828-1892).  qdy: ny; qdy2: msh*ny ([lvl][j]). */
typedef struct or_metrics {
    int ptsx[OR_MAXLVL], ptsy[OR_MAXLVL];
    double dy0[OR_MAXLVL], stf[OR_MAXLVL], dx2[OR_MAXLVL];
    double qdy[1024];
    double qdy2[OR_MAXLVL][1024];
} or_metrics;

int or_create_ctes(const or_params *p, or_metrics *m);

/* ---- compact derivatives (derivs_par.f:80-177) --------------------------- */
int or_derparx (const or_params *p, const or_cplx *f, or_cplx *out);
int or_derparxx(const or_params *p, const or_cplx *f, or_cplx *out);
int or_dery    (const or_params *p, const or_cplx *f, or_cplx *out);
int or_deryy   (const or_params *p, const or_cplx *f, or_cplx *out);
int or_deryfv  (const or_params *p, const or_cplx *f, or_cplx *out);

/* ---- banded LU (penta_par.f:308-404) -------------------------------------- */
/* a: n x 5 row-major bands (offsets -2..+2); al: n x 2 multipliers */
void or_band5_factor(double *a, int n, double *al);
void or_band5_solve(const double *a, int n, const double *al, or_cplx *rhs);

/* ---- Poisson pieces on one level (compact nx_l x ny_l planes) ------------- */
/* out(i,j) = S(i,j) + si * (L v)(i,j) for i=2..nx_l-1, j=2..ny_l
   (defect, poisson_par.f:543-793).  Other entries of out are untouched. */
int or_defect(const or_params *p, const or_metrics *m, int lvl, int k,
              const or_cplx *v, const or_cplx *S, or_cplx *out, double si);
/* one-mode source assembly (function_source, poisson_par.f:111-164) */
int or_source(const or_params *p, int k, const or_cplx *wx, const or_cplx *dwzdx,
              or_cplx *S);
/* itera line-relaxation sweeps.  order: 0 = reference lexicographic
   (poisson_par.f:191-221); 1 = colour order used by the GPU smoother. */
int or_sor(const or_params *p, const or_metrics *m, int lvl, int k,
           or_cplx *v, const or_cplx *S, int itera, int order);

/* full v-Poisson solve for every mode (poi_uy, poisson_par.f:6-64).
   uy in/out (nx,ny,nk); cycles[k] = completed V-cycles (printed value),
   res[k] = last max residual checked. */
int or_poi_uy(const or_params *p, const or_cplx *wx, const or_cplx *dwzdx,
              or_cplx *uy, double erromax, int max_cycles, int order,
              int *cycles, double *res);

/* same for the listed modes only (1-based k; the others of uy untouched,
   cycles/res written at index k-1) */
int or_poi_uy_modes(const or_params *p, const or_cplx *wx, const or_cplx *dwzdx,
                    or_cplx *uy, double erromax, int max_cycles, int order,
                    int *cycles, double *res, const int *modes, int nsel);

/* finest-level operator for one mode, all points (i=2..nx-1, j=2..ny get
   (L v); others 0) -- used to assemble the dense system in tests. */
int or_apply_L(const or_params *p, int k, const or_cplx *v, or_cplx *out);

/* same for an arbitrary level (compact nx_l x ny_l) */
int or_apply_L_level(const or_params *p, int lvl, int k, const or_cplx *v,
                     or_cplx *out);

/* restriction, prolongation+correction on one mode (poisson_par.f:796-942) */
int or_restrict(const or_params *p, const or_metrics *m, int lvl,
                const or_cplx *vf, const or_cplx *rf, or_cplx *vc, or_cplx *wc,
                or_cplx *Sc);
int or_correct_prolong(const or_params *p, const or_metrics *m, int lvl,
                       or_cplx *vc, const or_cplx *wc, or_cplx *vf);

/* ---- the rest of loop (main_par.f:1762-1785): tbl_oracle_loop.c --------- */
/* lhs5 (loop_par.f:258-299): a is n x 5, row i holds columns i-2..i+2 */
void or_lhs5(int n, double k2b2, double *a);
/* ludecomp (penta_par.f:6-45): lu n x 5 -- columns 1-2 the unit-lower
   multipliers (offsets -2, -1), 3-5 the upper band (0, +1, +2) */
void or_ludecomp(int n, const double *a, double *lu);
/* lusolver at one rank (penta_par.f:139-208) on one line of stride `stride`;
   variant 1 = lusolver_1's order of the two forward terms (53-136) */
void or_lusolve(int n, const double *lu, or_cplx *rhs, size_t stride, int variant);

/* poi_ux (loop_par.f:428-556): duydy = deryfv(uy) out; ux in/out */
int or_poi_ux(const or_params *p, const or_cplx *uy, const or_cplx *wy, or_cplx *ux,
              or_cplx *duydy);
/* poi_uz (loop_par.f:559-688) */
int or_poi_uz(const or_params *p, const or_cplx *wy, const or_cplx *duydy, or_cplx *uz);
/* der2wydxdy for mode k (loop_par.f:300-361): d (nx) out */
int or_der2wydxdy(const or_params *p, const or_cplx *wy, int k, or_cplx *d);
/* wx_wall (loop_par.f:174-255): lapv (nx, nk) out; wx in/out */
int or_wx_wall(const or_params *p, const or_cplx *uy, const or_cplx *wy,
               const or_cplx *d2uydx2, or_cplx *wx, or_cplx *lapv);
/* wz_wall (loop_par.f:364-425): wz in/out */
int or_wz_wall(const or_params *p, const or_cplx *wx, const or_cplx *lapv, or_cplx *wz);
/* bzone (loop_par.f:6-45, rows j >= 2) or bzonew (691-725, row 1), my_form 1
   (th NULL) or 2; bdfc is Fortran (2, nx) */
int or_bzone(const or_params *p, const double *bdfc, int wall_row, or_cplx *wx, or_cplx *wy,
             or_cplx *wz, or_cplx *th);
/* filter_trid (filter_par.f:144-178, 181-332, derivs_par.f:1009-1037) */
int or_filter_trid(const or_params *p, or_cplx *wx, or_cplx *wy, or_cplx *wz);
/* outuy (loop_par.f:48-119), my_form 1 or 2: dwzdx out, uy(nx, 2..ny, :) = 0 */
int or_outuy(const or_params *p, const or_cplx *wz, or_cplx *uy, or_cplx *dwzdx);
/* buffer-zone table (main_par.f:738-755), Fortran (2, nx); i0, i3, i4 1-based */
void or_bdfc(int nx, int i0, int i3, int i4, double *bdfc);
/* loop(erro) (main_par.f:1762-1785), my_form 1 (th NULL) or 2.  order as
   or_poi_uy; cycles/res per mode of the poi_uy call. */
int or_loop(const or_params *p, double erro, const double *bdfc, const or_cplx *d2uydx2,
            or_cplx *ux, or_cplx *uy, or_cplx *uz, or_cplx *wx, or_cplx *wy, or_cplx *wz,
            or_cplx *th, int order, int *cycles, double *res);

/* ---- right-hand side and RK4 step (tbl_oracle_rhs.c) --------------------- */
/* base flow and curvature: COMMON /blas/ uxb, uyb, wzb and /vc/ varg,
   dvargdx, each real (nx, ny) */
typedef struct or_base {
    const double *uxb, *uyb, *wzb, *varg, *dvargdx;
} or_base;
/* f_to_p / p_to_f (nlterms_par.f:303-428): phys is (nx, ny, kphys) reals */
int or_f_to_p(const or_params *p, int kphys, const or_cplx *f, double *phys);
int or_p_to_f(const or_params *p, int kphys, const double *phys, or_cplx *f);
/* kphys 0: lterms_gv (nlterms_par.f:89-112); else nlterms_gv (115-180) */
int or_terms(const or_params *p, int kphys, const or_base *b, const or_cplx *ux,
             const or_cplx *uy, const or_cplx *uz, const or_cplx *wx, const or_cplx *wy,
             const or_cplx *wz, or_cplx *a, or_cplx *bb, or_cplx *c, or_cplx *d);
/* drv_gv (main_par.f:1339-1403): rows j >= 2 of dvx, dvy, dvz written */
int or_drv_gv(const or_params *p, double Re, int kphys, const or_base *b, const or_cplx *ux,
              const or_cplx *uy, const or_cplx *uz, const or_cplx *wx, const or_cplx *wy,
              const or_cplx *wz, or_cplx *dvx, or_cplx *dvy, or_cplx *dvz);
/* flat-plate form (my_form 0): kphys 0: lterms_fi (nlterms_par.f:6-38,
   needs b); else nlterms_fi (41-86, b unused): a, b, c */
int or_terms_fi(const or_params *p, int kphys, const or_base *b, const or_cplx *ux,
                const or_cplx *uy, const or_cplx *uz, const or_cplx *wx, const or_cplx *wy,
                const or_cplx *wz, or_cplx *a, or_cplx *bb, or_cplx *c);
/* drv_fi (main_par.f:1264-1336): rows j >= 2 of dvx, dvy, dvz written */
int or_drv_fi(const or_params *p, double Re, int kphys, const or_base *b, const or_cplx *ux,
              const or_cplx *uy, const or_cplx *uz, const or_cplx *wx, const or_cplx *wy,
              const or_cplx *wz, or_cplx *dvx, or_cplx *dvy, or_cplx *dvz);
/* one RK4 time step (main_par.f:266-349); cycles: 4 x nk (per stage) */
int or_rk4_step(const or_params *p, double Re, double dt, int kphys, const or_base *b,
                const double *bdfc, or_cplx *d2uydx2, const or_cplx *pert_uy,
                const or_cplx *pert_d2, int pert_i0, int pert_i1, or_cplx *const *f, int order,
                int *cycles);

#endif
