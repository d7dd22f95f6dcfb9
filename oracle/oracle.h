/*
 * oracle.h -- FP64 C restatement of the reference multi-block LBM path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker of libmbl_hip: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product path never links, loads or calls it.
 *
 * Every function restates one reference function as CPU loops in double
 * precision with the same per-node arithmetic order (file:line cited at each
 * definition in oracle.c). Barrier-based device reductions become ordered
 * FP64 sums. Structures and flags are those of include/mbl.h compiled with
 * Lbd = double.
 */
#ifndef MBL_ORACLE_H
#define MBL_ORACLE_H

#define MBL_FP64 1
#include "../include/mbl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- node-level functions (lbm/assit_func.h) ---- */
double orc_feq(int k, double rh, double uxx, double uyy);
void orc_macros(double *rh, double *uxx, double *uyy, const double *ftmp);

/* ---- geometry (geo/init_geo.h) ---- */
int orc_geo_init(const mbl_geometry *g, P4 *AreaDefine, AREA *Areas,
                 Inf_Ff_store *Line_Inf_C, Inf_Ff_store *Line_Inf_F, Inf_Ff_store *Line_Inf_CC);

/* ---- one whole simulation, allocated like Multi-block-LBM.cu:main ---- */
typedef struct orc_sim orc_sim;
/* test hook called in place of a block's stream-collide: (ctx, area, level, substep) */
typedef void (*orc_block_hook)(void *ctx, int a, int level, int substep);

/* Creates the blocks, runs init_General, counts/enumerates the drag links of
 * the last block and initialises the interface rings and tridiagonal arrays. */
orc_sim *orc_sim_create(const mbl_geometry *g, const mbl_config *cfg);
void orc_sim_destroy(orc_sim *s);
/* Runs n coarse steps: evolution(level 0), error_L2, getDrag_block. */
int orc_sim_run(orc_sim *s, int n);
/* Runs n coarse steps of evolution only (no diagnostics). */
int orc_sim_evolve(orc_sim *s, int n);
/* error_L2 + getDrag_block on the current state: residual, Fx, Fy */
int orc_sim_diag(orc_sim *s, double out[3]);
/* velocity L2 distance from the analytic channel profile (FLUID nodes) */
double orc_sim_poiseuille_l2(const orc_sim *s);
int orc_sim_steps_done(const orc_sim *s);
int orc_sim_num_areas(const orc_sim *s);
int orc_sim_num_lines(const orc_sim *s, int which); /* 0: CF lines, 1: CC lines */
void orc_sim_area(const orc_sim *s, int a, AREA *out);
void orc_sim_line(const orc_sim *s, int which, int b, Inf_Ff_store *out);
/* field: 0 solid(as double) 1 x 2 y 3 rho 4 ux 5 uy 6 forcex 7 forcey
 *        8..16 f[0..8] (post-stream) 17..25 F[0..8] (post-collision) 26 upx 27 upy */
int orc_sim_get(const orc_sim *s, int a, int field, double *out);
int orc_sim_set(orc_sim *s, int a, int field, const double *in);
int orc_sim_drag_links(const orc_sim *s, Int3 *out, int cap);
/* series: kind 0 residual, 1 Fx, 2 Fy, 3 Cd, 4 Cl; returns count */
int orc_sim_series(const orc_sim *s, int kind, double *out, int cap);
/* ring slot data of CF line b, slot 0..2 (walk order &f_t1), plane k */
int orc_sim_ring(const orc_sim *s, int b, int slot, int k, double *out);
void orc_sim_markers(const orc_sim *s, int b, int out[3]);
/* spline second-derivative columns of CF line b after the last spline (12*Nc) */
int orc_sim_spline_M(const orc_sim *s, int b, double *out);
void orc_sim_set_block_hook(orc_sim *s, orc_block_hook hook, void *ctx);
void orc_sim_spline_line(orc_sim *s, int b);
void orc_sim_transfer_line(orc_sim *s, int b, int which);

/* ---- standalone kernels for unit tests ---- */
/* tridiag(l, m, u) x = b for nrhs column-major right-hand sides, in place */
void orc_gtsv(int n, int nrhs, const double *l, const double *m, const double *u, double *B, int ldb);
void orc_tri_A(int N, double *l, double *m, double *u);
/* temporal 3-point Lagrange value (assit_multi.h:232-245) */
double orc_temporal_value(double xm, double fm1, double f0, double f1);
/* Spline3_S (assit_multi.h:118-134) */
double orc_spline3_s(double xx, double x_1, double h, double fi_1, double fi, double mi_1, double mi);
/* as-coded spline of one column g[0..Nc-1] evaluated at the Nf = r*(Nc-1)+1
 * fine points: RHS 6g_i-3g_{i-1}-3g_{i+1}, natural ends, 0.1666667 */
void orc_spline_column(int Nc, int r, const double *g, double *M, double *out);
/* counter-based normal pair of the random init (SplitMix64 + Box-Muller) */
void orc_random_velocity(unsigned long long seed, double sigma, int i, int j, double *ux, double *uy);

/* ---- single-block periodic run (bandwidth benchmark workload) ---- */
typedef struct orc_periodic orc_periodic;
orc_periodic *orc_periodic_create(int nx, int ny, double tau, double rho0,
                                  unsigned long long seed, double sigma);
void orc_periodic_destroy(orc_periodic *p);
/* n steps of stb1/stb2/stb3 with periodic pulls; returns wall seconds */
double orc_periodic_run(orc_periodic *p, int n);
int orc_periodic_get(const orc_periodic *p, int field, double *out);
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
