/*
 * oracle.c -- FP64 C restatement of the reference multi-block D2Q9 LBM path.
 *
 * TEST INFRASTRUCTURE: the checker of libmbl_hip. Loaded only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
 * product. Each function cites the reference file:line it restates; kernels
 * become loops over nodes with the same per-node arithmetic order, device
 * tree reductions become ordered FP64 sums. Build: oracle/Makefile
 * (-O3 -fopenmp -ffp-contract=off).
 *
 * Deliberate deviations from the reference, each a defined replacement for
 * undefined or missing behaviour:
 *   - init_tri_A / init_Spline3 use race-free semantics (the reference races
 *     across thread blocks for N > 256, multi/FH/assit_multi.h:73-106, and on
 *     l[N-1] in multi/init_multi.h:42-49); the intended final state is used.
 *   - evolution checks `level == LEVEL` before reading mi[level]
 *     (evolution.h:5-8 reads mi[LEVEL] out of bounds first).
 *   - Boundary modes the reference lacks (channel walls / inlet / outlet,
 *     periodic) are extensions, marked EXT below.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* lbm/lbm.h:43-45 */
static const double w[9] = {4. / 9, 1. / 9, 1. / 9, 1. / 9, 1. / 9, 1. / 36, 1. / 36, 1. / 36, 1. / 36};
static const int e[18] = {0, 1, 0, -1, 0, 1, -1, -1, 1, 0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* lbm/assit_func.h:2-9 */
double orc_feq(int k, double rh, double uxx, double uyy)
{
    double eu, uv;
    eu = e[k] * uxx + e[k + 9] * uyy;
    uv = uxx * uxx + uyy * uyy;
    return w[k] * rh * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uv);
}

/* lbm/assit_func.h:23-31 */
void orc_macros(double *rh, double *uxx, double *uyy, const double *f)
{
    *rh = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    *uxx = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    *uyy = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    *uxx = *uxx / *rh;
    *uyy = *uyy / *rh;
}

/* ------------------------------------------------------------------------ */
/* geometry: geo/init_geo.h:17-86 (Line_init) and :89-165 (geo_init)          */
/* ------------------------------------------------------------------------ */

static void line_init(int dir, const P4 *aC, const AREA *sC, Inf_Ff_store *lC,
                      const P4 *aF, const AREA *sF, Inf_Ff_store *lF)
{
    if (dir == 0 || dir == 1) {
        lC->line_C.xstart = (dir == 0 ? sC->m : sC->nx - 1 - sC->m);
        lC->line_C.xend = lC->line_C.xstart;
        lC->line_F.xstart = (dir == 0 ? sF->nx - 1 : 0);
        lC->line_F.xend = lC->line_F.xstart;
        lF->line_C.xstart = (dir == 0 ? 0 : sC->nx - 1);
        lF->line_C.xend = lF->line_C.xstart;
        lF->line_F.xstart = (dir == 0 ? sF->nx - sF->m - 1 : sF->m);
        lF->line_F.xend = lF->line_F.xstart;
        lC->line_C.ystart = (aF->ystart - aC->ystart <= 0) ? 0 : (aF->ystart - aC->ystart) * sC->m;
        lC->line_C.yend = (aF->yend - aC->yend >= 0) ? sC->ny - 1 : (aF->yend - aC->ystart) * sC->m;
        lC->line_F.ystart = (aC->ystart - aF->ystart <= 0) ? 0 : (aC->ystart - aF->ystart) * sF->m;
        lC->line_F.yend = (aC->yend - aF->yend >= 0) ? sF->ny - 1 : (aC->yend - aF->ystart) * sF->m;
        lF->line_C.ystart = lC->line_C.ystart;
        lF->line_C.yend = lC->line_C.yend;
        lF->line_F.ystart = lC->line_F.ystart;
        lF->line_F.yend = lC->line_F.yend;
        lC->Nc = lC->line_C.yend - lC->line_C.ystart + 1;
        lC->Nf = lC->line_F.yend - lC->line_F.ystart + 1;
        lF->Nc = lC->Nc;
        lF->Nf = lC->Nf;
    } else if (dir == 2 || dir == 3) {
        lC->line_C.ystart = (dir == 2 ? sC->m : sC->ny - 1 - sC->m);
        lC->line_C.yend = lC->line_C.ystart;
        lC->line_F.ystart = (dir == 2 ? sF->ny - 1 : 0);
        lC->line_F.yend = lC->line_F.ystart;
        lF->line_C.ystart = (dir == 2 ? 0 : sC->ny - 1);
        lF->line_C.yend = lF->line_C.ystart;
        lF->line_F.ystart = (dir == 2 ? sF->ny - sF->m - 1 : sF->m);
        lF->line_F.yend = lF->line_F.ystart;
        lC->line_C.xstart = (aF->xstart - aC->xstart <= 0) ? 0 : (aF->xstart - aC->xstart) * sC->m;
        lC->line_C.xend = (aF->xend - aC->xend >= 0) ? sC->nx - 1 : (aF->xend - aC->xstart) * sC->m;
        lC->line_F.xstart = (aC->xstart - aF->xstart <= 0) ? 0 : (aC->xstart - aF->xstart) * sF->m;
        lC->line_F.xend = (aC->xend - aF->xend >= 0) ? sF->nx - 1 : (aC->xend - aF->xstart) * sF->m;
        lF->line_C.xstart = lC->line_C.xstart;
        lF->line_C.xend = lC->line_C.xend;
        lF->line_F.xstart = lC->line_F.xstart;
        lF->line_F.xend = lC->line_F.xend;
        lC->Nc = lC->line_C.xend - lC->line_C.xstart + 1;
        lC->Nf = lC->line_F.xend - lC->line_F.xstart + 1;
        lF->Nc = lC->Nc;
        lF->Nf = lC->Nf;
    }
}

int orc_geo_init(const mbl_geometry *g, P4 *AD, AREA *Areas,
                 Inf_Ff_store *LC, Inf_Ff_store *LF, Inf_Ff_store *LCC)
{
    for (int i = 0; i < g->number_area; i++) {
        if (g->AreaDefine) {
            AD[i] = g->AreaDefine[i];
        } else {
            const P4d4i *z = &g->AreaZone[i];
            AD[i].xstart = (int)(z->reduced.xstart * (g->LX - 1)) + z->move.xstart;
            AD[i].xend = (int)(z->reduced.xend * (g->LX - 1)) + z->move.xend;
            AD[i].ystart = (int)(z->reduced.ystart * (g->LY - 1)) + z->move.ystart;
            AD[i].yend = (int)(z->reduced.yend * (g->LY - 1)) + z->move.yend;
        }
    }
    int mm = 1;
    double niu = g->U0_2 * 2 * g->R / g->Re;
    double tau = 0.5 + 3.0 * niu;
    double rv_tau = 1.0 / tau;
    for (int i = 0; i < g->LEVEL; i++) {
        mm *= g->mi[i];
        tau = 0.5 + g->mi[i] * (tau - 0.5);
        rv_tau = 1.0 / tau;
        for (int j = g->ASP[i]; j < g->ASP[i + 1]; j++) {
            Areas[j].m = mm;
            Areas[j].nx = (AD[j].xend - AD[j].xstart) * mm + 1;
            Areas[j].ny = (AD[j].yend - AD[j].ystart) * mm + 1;
            Areas[j].base_x = AD[j].xstart;
            Areas[j].base_y = AD[j].ystart;
            Areas[j].tau = tau;
            Areas[j].rv_tau = rv_tau;
        }
    }
    for (int i = 0; i < g->number_line; i++) {
        int nc = g->NeborCF[i].nc, nf = g->NeborCF[i].nf;
        LC[i].areac_n = nc;
        LC[i].areaf_n = nf;
        LF[i].areac_n = nc;
        LF[i].areaf_n = nf;
        line_init(g->NeborCF[i].direction, &AD[nc], &Areas[nc], &LC[i], &AD[nf], &Areas[nf], &LF[i]);
    }
    for (int i = 0; i < g->number_line2; i += 2) {
        int nc = g->NeborCC[i].nc, nf = g->NeborCC[i].nf;
        LCC[i].areac_n = nc;
        LCC[i].areaf_n = nf;
        line_init(g->NeborCC[i].direction, &AD[nc], &Areas[nc], &LCC[i], &AD[nf], &Areas[nf], &LCC[i + 1]);
        LCC[i + 1].areac_n = g->NeborCC[i + 1].nc;
        LCC[i + 1].areaf_n = g->NeborCC[i + 1].nf;
        P4 tmp = LCC[i + 1].line_C;
        LCC[i + 1].line_C = LCC[i + 1].line_F;
        LCC[i + 1].line_F = tmp;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* random initial velocity (EXT, bandwidth benchmark): SplitMix64 finaliser   */
/* (Steele, Lea & Flood 2014) on a per-node counter, then Box-Muller.         */
/* ------------------------------------------------------------------------ */
static unsigned long long splitmix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_random_velocity(unsigned long long seed, double sigma, int i, int j, double *ux, double *uy)
{
    unsigned long long key = ((unsigned long long)(unsigned)j << 32) | (unsigned long long)(unsigned)i;
    unsigned long long a = splitmix64(seed + (2 * key + 1) * 0x9E3779B97F4A7C15ULL);
    unsigned long long b = splitmix64(seed + (2 * key + 2) * 0x9E3779B97F4A7C15ULL);
    double u1 = (double)((a >> 11) + 1) * (1.0 / 9007199254740992.0); /* (0, 1] */
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);       /* [0, 1) */
    double r = sigma * sqrt(-2.0 * log(u1));
    double th = 6.283185307179586 * u2;
    *ux = r * cos(th);
    *uy = r * sin(th);
}

/* ------------------------------------------------------------------------ */
/* per-block kernels                                                          */
/* ------------------------------------------------------------------------ */

typedef struct orc_block {
    AREA area;
    D2Q9 pp;
    double *upx, *upy;
} orc_block;

static double *dalloc(size_t n)
{
    double *p = (double *)calloc(n ? n : 1, sizeof(double));
    if (!p) {
        fprintf(stderr, "oracle: out of memory (%zu doubles)\n", n);
        abort();
    }
    return p;
}

static void block_alloc(orc_block *b, const AREA *a)
{
    size_t n = (size_t)a->nx * a->ny;
    b->area = *a;
    b->pp.solid = (int *)calloc(n, sizeof(int));
    double **v = &b->pp.x;
    for (int k = 0; k < 25; k++) v[k] = dalloc(n);
    b->upx = dalloc(n);
    b->upy = dalloc(n);
}

static void block_free(orc_block *b)
{
    free(b->pp.solid);
    double **v = &b->pp.x;
    for (int k = 0; k < 25; k++) free(v[k]);
    free(b->upx);
    free(b->upy);
}

/* EXT: parabolic inlet profile, the form left commented at lbm/init_block.h:20 */
static double inlet_u(const mbl_config *c, double y)
{
    double LY = c->LY;
    return 4.0 * c->U0[0] * (LY - 1.5 - y) * (y - 0.5) / (LY - 2.0) / (LY - 2.0);
}

/* lbm/init_block.h:2-51 (+ EXT flag rules for the channel and periodic modes) */
static void init_general(const mbl_config *c, orc_block *b)
{
    AREA a = b->area;
    D2Q9 pp = b->pp;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < a.ny; j++) {
        for (int i = 0; i < a.nx; i++) {
            int ixy = i + j * a.nx;
            double dx = 1.0 / a.m;
            pp.x[ixy] = i * dx + a.base_x;
            pp.y[ixy] = j * dx + a.base_y;
            double ux = c->U0[0], uy = c->U0[1];
            if (c->init_mode == MBL_INIT_REST) {
                ux = 0.0;
                uy = 0.0;
            } else if (c->init_mode == MBL_INIT_RANDOM) {
                orc_random_velocity(c->seed, c->sigma, i, j, &ux, &uy);
            }
            pp.rho[ixy] = c->rho0;
            pp.ux[ixy] = ux;
            pp.uy[ixy] = uy;
            b->upx[ixy] = ux;
            b->upy[ixy] = uy;
            pp.forcex[ixy] = 0.0;
            pp.forcey[ixy] = 0.0;
            double X = pp.x[ixy], Y = pp.y[ixy];
            int flag;
            if ((X - c->CX) * (X - c->CX) + (Y - c->CY) * (Y - c->CY) < c->R * c->R)
                flag = SOLID_SBB;
            else if (c->bc_mode == MBL_BC_PERIODIC)
                flag = FLUID;
            else if (c->bc_mode == MBL_BC_CHANNEL) {
                if (Y == 0.0 || Y == c->LY - 1.0)
                    flag = MBL_SOLID_WALL;
                else if (X == 0.0)
                    flag = BOUNDARY_INLET;
                else if (X == c->LX - 1.0)
                    flag = BOUNDARY_OUTLET;
                else if (i == 0 || j == 0 || i == a.nx - 1 || j == a.ny - 1)
                    flag = C_BOUNDARY;
                else
                    flag = FLUID;
            } else {
                if (Y == 0.0)
                    flag = FLUID;
                else if (Y == c->LY - 1.0)
                    flag = FLUID;
                else if (X == 0.0)
                    flag = FLUID;
                else if (X == c->LX - 1.0)
                    flag = FLUID;
                else if (i == 0 || j == 0 || i == a.nx - 1 || j == a.ny - 1)
                    flag = C_BOUNDARY;
                else
                    flag = FLUID;
            }
            pp.solid[ixy] = flag;
            if (c->init_mode == MBL_INIT_PROFILE) { /* EXT: developed inlet profile, rest in solids */
                int solid = (flag == SOLID_SBB || flag == MBL_SOLID_WALL);
                ux = solid ? 0.0 : inlet_u(c, Y);
                pp.ux[ixy] = ux;
                pp.uy[ixy] = 0.0;
                b->upx[ixy] = ux;
                b->upy[ixy] = 0.0;
            }
            for (int k = 0; k < 9; k++) {
                pp.F[k][ixy] = orc_feq(k, pp.rho[ixy], pp.ux[ixy], pp.uy[ixy]);
                pp.f[k][ixy] = pp.F[k][ixy];
            }
        }
    }
}

/* multi/FH/evolution_block.h:3-32 (pull stream). Out-of-block neighbours are
 * skipped (stale f), as the reference; EXT: wall nodes take their own F for
 * those directions, periodic blocks wrap. */
static void stb1(orc_block *b, int periodic)
{
    AREA a = b->area;
    D2Q9 pp = b->pp;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < a.ny; j++) {
        for (int i = 0; i < a.nx; i++) {
            int ixy = i + j * a.nx;
            int s = pp.solid[ixy];
            if (s == FLUID || s == SOLID_SBB || s == MBL_SOLID_WALL) {
                for (int k = 0; k < 9; k++) {
                    int id = i - e[k];
                    int jd = j - e[k + 9];
                    if (periodic) {
                        id = (id + a.nx) % a.nx;
                        jd = (jd + a.ny) % a.ny;
                    } else if (id < 0 || id > a.nx - 1 || jd < 0 || jd > a.ny - 1) {
                        if (s == MBL_SOLID_WALL) pp.f[k][ixy] = pp.F[k][ixy];
                        continue;
                    }
                    pp.f[k][ixy] = pp.F[k][id + jd * a.nx];
                }
            }
        }
    }
}

/* multi/FH/evolution_block.h:33-73 (+ EXT inlet) */
static void stb2(const mbl_config *c, orc_block *b, int periodic)
{
    AREA a = b->area;
    D2Q9 pp = b->pp;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < a.ny; j++) {
        for (int i = 0; i < a.nx; i++) {
            int ixy = i + j * a.nx;
            int s = pp.solid[ixy];
            if (s == FLUID) {
                if (!periodic && (i == 0 || i == a.nx - 1 || j == 0 || j == a.ny - 1)) {
                    pp.rho[ixy] = c->rho0;
                    pp.ux[ixy] = c->U0[0];
                    pp.uy[ixy] = c->U0[1];
                    for (int k = 0; k < 9; k++) pp.f[k][ixy] = orc_feq(k, c->rho0, c->U0[0], c->U0[1]);
                } else {
                    double ft[9], rh, ux, uy;
                    for (int k = 0; k < 9; k++) ft[k] = pp.f[k][ixy];
                    orc_macros(&rh, &ux, &uy, ft);
                    pp.rho[ixy] = rh;
                    pp.ux[ixy] = ux;
                    pp.uy[ixy] = uy;
                }
            }
        }
    }
}

/* multi/FH/evolution_block.h:74-119, then the EXT channel ends by
 * non-equilibrium extrapolation (Guo, Zheng & Shi, Chin. Phys. 11, 2002) from
 * the neighbour's post-collision state: inlet F = feq(rho_N, u_in(y), 0) +
 * F_N - feq_N with N = x+1 (density extrapolated, velocity imposed); outlet
 * F = feq(rho0, u_N) + F_N - feq_N with N = x-1 (pressure imposed, velocity
 * extrapolated). Both read N after its collision in the same step. */
static void stb3(const mbl_config *c, orc_block *b)
{
    AREA a = b->area;
    D2Q9 pp = b->pp;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < a.ny; j++) {
        for (int i = 0; i < a.nx; i++) {
            int ixy = i + j * a.nx;
            double dt = 1.0 / a.m;
            double ft[9];
            for (int k = 0; k < 9; k++) ft[k] = pp.f[k][ixy];
            int s = pp.solid[ixy];
            if (s == FLUID) {
                double rh = pp.rho[ixy], ux = pp.ux[ixy], uy = pp.uy[ixy];
                double fx = pp.forcex[ixy], fy = pp.forcey[ixy];
                for (int k = 0; k < 9; k++) {
                    double vx = 3.0 * (e[k] - ux) + 9.0 * (e[k] * ux + e[k + 9] * uy) * e[k];
                    double vy = 3.0 * (e[k + 9] - uy) + 9.0 * (e[k] * ux + e[k + 9] * uy) * e[k + 9];
                    double Fk = (1.0 - 0.5 * a.rv_tau) * w[k] * (vx * fx + vy * fy);
                    pp.F[k][ixy] = ft[k] - (ft[k] - orc_feq(k, rh, ux, uy)) * a.rv_tau + dt * Fk;
                }
            } else if (s == SOLID_SBB || s == MBL_SOLID_WALL) {
                for (int k = 0; k < 9; k++) pp.F[k][ixy] = ft[opp[k]];
            }
        }
    }
    for (int j = 0; j < a.ny; j++) {
        for (int i = 0; i < a.nx; i++) {
            int ixy = i + j * a.nx;
            int s = pp.solid[ixy];
            if (s != BOUNDARY_INLET && s != BOUNDARY_OUTLET) continue;
            int n = (s == BOUNDARY_INLET) ? ixy + 1 : ixy - 1;
            double rN = pp.rho[n], uN = pp.ux[n], vN = pp.uy[n];
            double rs, us, vs;
            if (s == BOUNDARY_INLET) {
                rs = rN;
                us = inlet_u(c, pp.y[ixy]);
                vs = 0.0;
            } else {
                rs = c->rho0;
                us = uN;
                vs = vN;
            }
            for (int k = 0; k < 9; k++)
                pp.F[k][ixy] = orc_feq(k, rs, us, vs) + (pp.F[k][n] - orc_feq(k, rN, uN, vN));
            pp.rho[ixy] = rs;
            pp.ux[ixy] = us;
            pp.uy[ixy] = vs;
        }
    }
}

/* multi/FH/evolution_block.h:121-126 */
static void evolution_area_stb(const mbl_config *c, orc_block *b, int periodic)
{
    stb1(b, periodic);
    stb2(c, b, periodic);
    stb3(c, b);
}

/* ------------------------------------------------------------------------ */
/* interface kernels: multi/FH/assit_multi.h, multi/init_multi.h              */
/* ------------------------------------------------------------------------ */

static void line_node(const P4 *l, int ip, int nx, int *ixy)
{
    int i, j;
    if (l->xstart == l->xend) {
        i = l->xstart;
        j = l->ystart + ip;
    } else {
        i = l->xstart + ip;
        j = l->ystart;
    }
    *ixy = i + j * nx;
}

/* assit_multi.h:2-46 */
static void c_c_boundary_transfer(orc_block *b1, orc_block *b2, const Inf_Ff_store *L)
{
    int xs1 = L->line_C.xstart, xe1 = L->line_C.xend, ys1 = L->line_C.ystart, ye1 = L->line_C.yend;
    int xs2 = L->line_F.xstart, ys2 = L->line_F.ystart;
    for (int ip = 0; ip < L->Nc; ip++) {
        int i1 = 0, j1 = 0, i2 = 0, j2 = 0;
        if (xs1 == xe1) {
            i1 = xs1;
            j1 = ys1 + ip;
            i2 = xs2;
            j2 = ys2 + ip;
        }
        if (ys1 == ye1) {
            i1 = xs1 + ip;
            j1 = ys1;
            i2 = xs2 + ip;
            j2 = ys2;
        }
        int ixy1 = i1 + j1 * b1->area.nx;
        int ixy2 = i2 + j2 * b2->area.nx;
        b2->pp.rho[ixy2] = b1->pp.rho[ixy1];
        b2->pp.ux[ixy2] = b1->pp.ux[ixy1];
        b2->pp.uy[ixy2] = b1->pp.uy[ixy1];
        for (int k = 0; k < 9; k++) b2->pp.F[k][ixy2] = b1->pp.F[k][ixy1];
    }
}

/* multi/init_multi.h:30-51, final (race-free) state */
void orc_tri_A(int N, double *l, double *m, double *u)
{
    for (int ip = 0; ip < N; ip++) {
        l[ip] = 0.5;
        m[ip] = 2.0;
        u[ip] = 0.5;
    }
    l[0] = 0.0;
    u[N - 1] = 0.0;
    u[0] = 0.0;
    if (N > 1) l[1] = 0.0;
    if (N >= 2) {
        u[N - 2] = 0.0;
        l[N - 1] = 0.0;
    }
}

/* Thomas algorithm (no pivoting: the matrix is strictly diagonally dominant);
 * stands in for cusparseDgtsv2 (assit_multi.h:197-204), whose result is the
 * unique solution of the same system. */
void orc_gtsv(int n, int nrhs, const double *l, const double *m, const double *u, double *B, int ldb)
{
    double *cp = dalloc((size_t)n);
    for (int r = 0; r < nrhs; r++) {
        double *d = B + (size_t)r * ldb;
        cp[0] = u[0] / m[0];
        d[0] = d[0] / m[0];
        for (int i = 1; i < n; i++) {
            double den = m[i] - l[i] * cp[i - 1];
            cp[i] = u[i] / den;
            d[i] = (d[i] - l[i] * d[i - 1]) / den;
        }
        for (int i = n - 2; i >= 0; i--) d[i] = d[i] - cp[i] * d[i + 1];
    }
    free(cp);
}

/* assit_multi.h:118-134, including the literal 0.1666667 */
double orc_spline3_s(double xx, double x_1, double h, double fi_1, double fi, double mi_1, double mi)
{
    double x, a, b, c, d;
    x = x_1 + h;
    a = mi_1 * (x - xx) * (x - xx) * (x - xx) * 0.1666667;
    b = mi * (xx - x_1) * (xx - x_1) * (xx - x_1) * 0.1666667;
    c = (fi_1 - mi_1 * h * h * 0.1666667) * (x - xx);
    d = (fi - mi * h * h * 0.1666667) * (xx - x_1);
    return (a + b + c + d);
}

static int spline_I(double xx)
{
    /* assit_multi.h:162-166 */
    if (ceil(xx) == 0)
        return (int)(ceil(xx) + 1.0);
    return (int)(ceil(xx));
}

void orc_spline_column(int Nc, int r, const double *g, double *M, double *out)
{
    double *l = dalloc(Nc), *m = dalloc(Nc), *u = dalloc(Nc);
    orc_tri_A(Nc, l, m, u);
    for (int ip = 0; ip < Nc; ip++)
        M[ip] = (ip == 0 || ip == Nc - 1) ? 0.0 : (6.0 * g[ip] - 3.0 * g[ip - 1] - 3.0 * g[ip + 1]);
    orc_gtsv(Nc, 1, l, m, u, M, Nc);
    int Nf = r * (Nc - 1) + 1;
    for (int ip = 0; ip < Nf; ip++) {
        double xx = (double)ip / (double)r;
        int I = spline_I(xx);
        out[ip] = orc_spline3_s(xx, (I - 1.0), 1.0, g[I - 1], g[I], M[I - 1], M[I]);
    }
    free(l);
    free(m);
    free(u);
}

/* assit_multi.h:48-116 (gather + RHS, race-free) */
static void init_spline3(orc_block *bc, const Inf_Ff_store *L, FM_spline3 *S)
{
    int N = S->N;
    for (int ip = 0; ip < N; ip++) {
        int ixy;
        line_node(&L->line_C, ip, bc->area.nx, &ixy);
        S->rho[ip] = bc->pp.rho[ixy];
        S->ux[ip] = bc->pp.ux[ixy];
        S->uy[ip] = bc->pp.uy[ixy];
        for (int k = 0; k < 9; k++) S->F[k][ip] = bc->pp.F[k][ixy];
    }
    double *col[12] = {S->rho, S->ux, S->uy, S->F[0], S->F[1], S->F[2], S->F[3],
                       S->F[4], S->F[5], S->F[6], S->F[7], S->F[8]};
    for (int ip = 0; ip < N; ip++) {
        for (int c = 0; c < 12; c++) {
            if (ip == 0 || ip == N - 1)
                S->Varibs_M[c * N + ip] = 0.0;
            else
                S->Varibs_M[c * N + ip] = (6.0 * col[c][ip] - 3.0 * col[c][ip - 1] - 3.0 * col[c][ip + 1]);
        }
    }
}

/* assit_multi.h:136-185 */
static void calculate_spline3(const AREA *ac, orc_block *bf, const Inf_Ff_store *L, F_k3 *slot, FM_spline3 *S)
{
    const AREA *af = &bf->area;
    int N = S->N;
    int r = af->m / ac->m;
    for (int ip = 0; ip < L->Nf; ip++) {
        int ixy;
        line_node(&L->line_F, ip, af->nx, &ixy);
        double xx = (double)ip / (double)(af->m / ac->m);
        int I = spline_I(xx);
        double rho = orc_spline3_s(xx, (I - 1.0), 1.0, S->rho[I - 1], S->rho[I], S->Varibs_M[I - 1], S->Varibs_M[I]);
        double ux = orc_spline3_s(xx, (I - 1.0), 1.0, S->ux[I - 1], S->ux[I], S->Varibs_M[N + I - 1], S->Varibs_M[N + I]);
        double uy = orc_spline3_s(xx, (I - 1.0), 1.0, S->uy[I - 1], S->uy[I], S->Varibs_M[2 * N + I - 1], S->Varibs_M[2 * N + I]);
        bf->pp.rho[ixy] = rho;
        bf->pp.ux[ixy] = ux;
        bf->pp.uy[ixy] = uy;
        double para = (af->tau - 1.0) / (r * (ac->tau - 1.0));
        for (int k = 0; k < 9; k++) {
            double ft = orc_spline3_s(xx, (I - 1.0), 1.0, S->F[k][I - 1], S->F[k][I],
                                      S->Varibs_M[(3 + k) * N + I - 1], S->Varibs_M[(3 + k) * N + I]);
            double fe = orc_feq(k, rho, ux, uy);
            slot->F[k][ip] = fe + para * (ft - fe);
        }
    }
}

/* assit_multi.h:232-245: 3-point Lagrange in time at x = xm */
static double temporal_value(double xm, double fm1, double f0, double f1)
{
    double a = 0.5 * (xm - 1.0) * xm;
    double b = (xm - 1.0) * (xm + 1.0);
    double c = 0.5 * xm * (xm + 1.0);
    return a * fm1 - b * f0 + c * f1;
}

double orc_temporal_value(double xm, double fm1, double f0, double f1) { return temporal_value(xm, fm1, f0, f1); }

/* assit_multi.h:213-247 */
static void temporal_interpolation(orc_block *bf, const Inf_Ff_store *L, int N,
                                   const F_k3 *Ft_1, const F_k3 *Ft, const F_k3 *Ft1, double xm)
{
    for (int ip = 0; ip < N; ip++) {
        int ixy;
        line_node(&L->line_F, ip, bf->area.nx, &ixy);
        for (int k = 0; k < 9; k++)
            bf->pp.F[k][ixy] = temporal_value(xm, Ft_1->F[k][ip], Ft->F[k][ip], Ft1->F[k][ip]);
    }
}

/* assit_multi.h:249-271 */
static void c_f_boundary_transfer(orc_block *bf, const Inf_Ff_store *L, int N, const F_k3 *Ft)
{
    for (int ip = 0; ip < N; ip++) {
        int ixy;
        line_node(&L->line_F, ip, bf->area.nx, &ixy);
        for (int k = 0; k < 9; k++) bf->pp.F[k][ixy] = Ft->F[k][ip];
    }
}

/* assit_multi.h:273-314 */
static void f_c_boundary_transfer(orc_block *bc, orc_block *bf, const Inf_Ff_store *L)
{
    const AREA *ac = &bc->area, *af = &bf->area;
    int r = af->m / ac->m;
    for (int ip = 0; ip < L->Nc; ip++) {
        int ic, jc, i_f, j_f;
        if (L->line_C.xstart == L->line_C.xend) {
            ic = L->line_C.xstart;
            jc = L->line_C.ystart + ip;
            i_f = L->line_F.xstart;
            j_f = L->line_F.ystart + r * ip;
        } else {
            ic = L->line_C.xstart + ip;
            jc = L->line_C.ystart;
            i_f = L->line_F.xstart + r * ip;
            j_f = L->line_F.ystart;
        }
        int ixyc = ic + jc * ac->nx;
        int ixyf = i_f + j_f * af->nx;
        bc->pp.rho[ixyc] = bf->pp.rho[ixyf];
        bc->pp.ux[ixyc] = bf->pp.ux[ixyf];
        bc->pp.uy[ixyc] = bf->pp.uy[ixyf];
        double para = r * (ac->tau - 1.0) / (af->tau - 1.0);
        for (int k = 0; k < 9; k++) {
            double ft = bf->pp.F[k][ixyf];
            double fe = orc_feq(k, bf->pp.rho[ixyf], bf->pp.ux[ixyf], bf->pp.uy[ixyf]);
            bc->pp.F[k][ixyc] = fe + para * (ft - fe);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* simulation (Multi-block-LBM.cu:main, evolution.h, force_block.h, error_L2.h) */
/* ------------------------------------------------------------------------ */

struct orc_sim {
    mbl_config cfg;
    mbl_geometry g;
    int LEVEL, number_area, number_line, number_line2;
    int mi[MBL_MAX_LEVEL];
    int ASP[MBL_MAX_LEVEL + 1], LTCFFCP[MBL_MAX_LEVEL + 1], LS3CFFCP[MBL_MAX_LEVEL + 1], LCCP[MBL_MAX_LEVEL + 1];
    P4 *AD;
    orc_block *blk;
    Inf_Ff_store *LC, *LF, *LCC;
    Ff_Store *t3;
    FM_spline3 *s3;
    int N_drag;
    Int3 *links;
    double *ser[5];
    int nser, capser;
    int steps;
    orc_block_hook hook; /* test hook: replaces the stream-collide of a block (NULL: stb1-3) */
    void *hook_ctx;
};

/* lbm/force_block.h:2-75 (count) and :78-127 (enumerate, i-major) */
static int drag_links(const orc_block *b, Int3 *out)
{
    const AREA *a = &b->area;
    int n = 0;
    for (int i = 0; i < a->nx; i++) {
        for (int j = 0; j < a->ny; j++) {
            int ixy = i + j * a->nx;
            if (b->pp.solid[ixy] != SOLID_SBB) continue;
            for (int k = 0; k < 9; k++) {
                int id = i - e[k], jd = j - e[k + 9];
                if (id < 0 || id > a->nx - 1 || jd < 0 || jd > a->ny - 1) continue;
                if (b->pp.solid[id + jd * a->nx] == FLUID) {
                    if (out) {
                        out[n].i = i;
                        out[n].j = j;
                        out[n].k = k;
                    }
                    n++;
                }
            }
        }
    }
    return n;
}

/* lbm/force_block.h:129-216 */
static void get_drag(orc_sim *s, double force[2])
{
    const orc_block *b = &s->blk[s->number_area - 1];
    double fx = 0.0, fy = 0.0;
    for (int ip = 0; ip < s->N_drag; ip++) {
        int i = s->links[ip].i, j = s->links[ip].j, k = s->links[ip].k;
        int id = i - e[k], jd = j - e[9 + k], kd = opp[k];
        int ixy = i + j * b->area.nx, ixyd = id + jd * b->area.nx;
        fx += (b->pp.f[k][ixy] * e[k] - b->pp.f[kd][ixyd] * e[kd]);
        fy += (b->pp.f[k][ixy] * e[k + 9] - b->pp.f[kd][ixyd] * e[kd + 9]);
    }
    force[0] = fx;
    force[1] = fy;
}

/* lbm/error_L2.h:3-91 */
static double error_l2(orc_sim *s)
{
    double uvp2_all = 0.0, uv2_all = 0.0;
    for (int a = 0; a < s->number_area; a++) {
        orc_block *b = &s->blk[a];
        size_t n = (size_t)b->area.nx * b->area.ny;
        double uvp2 = 0.0, uv2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : uvp2, uv2)
        for (size_t ixy = 0; ixy < n; ixy++) {
            double ux = b->pp.ux[ixy], upx = b->upx[ixy], uy = b->pp.uy[ixy], upy = b->upy[ixy];
            uvp2 += (ux - upx) * (ux - upx) + (uy - upy) * (uy - upy);
            uv2 += ux * ux + uy * uy;
            b->upx[ixy] = ux;
            b->upy[ixy] = uy;
        }
        uvp2_all += uvp2 / b->area.m;
        uv2_all += uv2 / b->area.m;
    }
    return sqrt(uvp2_all) / sqrt(uv2_all);
}

static void space_interpolation_spline3(orc_sim *s, int b)
{
    Inf_Ff_store *L = &s->LC[b];
    orc_block *bc = &s->blk[L->areac_n], *bf = &s->blk[L->areaf_n];
    FM_spline3 *S = &s->s3[b];
    init_spline3(bc, L, S);
    orc_gtsv(L->Nc, 12, S->l, S->m, S->u, S->Varibs_M, L->Nc);
    F_k3 *slots = &s->t3[b].f_t1;
    calculate_spline3(&bc->area, bf, L, &slots[s->t3[b].marker[0]], S);
}

/* evolution.h:2-65 */
static void evolution_rec(orc_sim *s, int level)
{
    if (level == s->LEVEL) return;
    int periodic = (s->cfg.bc_mode == MBL_BC_PERIODIC);
    for (int i = 0; i < s->mi[level]; i++) {
        if (!(level == 0 || i == 0)) {
            for (int b = s->LTCFFCP[level]; b < s->LTCFFCP[level + 1]; b++) {
                F_k3 *ft = &s->t3[b].f_t1;
                Ff_Store *t = &s->t3[b];
                temporal_interpolation(&s->blk[s->LC[b].areaf_n], &s->LC[b], t->N, &ft[t->marker[0]],
                                       &ft[t->marker[1]], &ft[t->marker[2]], (1.0 * i / s->mi[level]));
            }
        }
        for (int a = s->ASP[level]; a < s->ASP[level + 1]; a++) {
            if (s->hook)
                s->hook(s->hook_ctx, a, level, i);
            else
                evolution_area_stb(&s->cfg, &s->blk[a], periodic);
        }
        if (s->number_line2 != 0) {
            for (int c = s->LCCP[level]; c < s->LCCP[level + 1]; c++)
                c_c_boundary_transfer(&s->blk[s->LCC[c].areac_n], &s->blk[s->LCC[c].areaf_n], &s->LCC[c]);
        }
        if (level != s->LEVEL - 1) {
            for (int b = s->LS3CFFCP[level]; b < s->LS3CFFCP[level + 1]; b++) {
                space_interpolation_spline3(s, b);
                for (int p = 0; p < 3; p++) {
                    int bb = s->t3[b].marker[p] - 1;
                    s->t3[b].marker[p] = (bb == -1) ? 2 : bb;
                }
            }
        }
        evolution_rec(s, level + 1);
        if (level != 0 && i == s->mi[level] - 1) {
            for (int b = s->LTCFFCP[level]; b < s->LTCFFCP[level + 1]; b++) {
                F_k3 *ft = &s->t3[b].f_t1;
                c_f_boundary_transfer(&s->blk[s->LC[b].areaf_n], &s->LC[b], s->t3[b].N, &ft[s->t3[b].marker[2]]);
                f_c_boundary_transfer(&s->blk[s->LF[b].areac_n], &s->blk[s->LF[b].areaf_n], &s->LF[b]);
            }
        }
    }
}

orc_sim *orc_sim_create(const mbl_geometry *g, const mbl_config *cfg)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    s->cfg = *cfg;
    s->g = *g;
    s->LEVEL = g->LEVEL;
    s->number_area = g->number_area;
    s->number_line = g->number_line;
    s->number_line2 = g->number_line2;
    for (int i = 0; i < MBL_MAX_LEVEL; i++) s->mi[i] = g->mi[i];
    for (int i = 0; i <= MBL_MAX_LEVEL; i++) {
        s->ASP[i] = g->ASP[i];
        s->LTCFFCP[i] = g->LTCFFCP[i];
        s->LS3CFFCP[i] = g->LS3CFFCP[i];
        s->LCCP[i] = g->LCCP[i];
    }
    s->AD = (P4 *)calloc(s->number_area, sizeof(P4));
    AREA *areas = (AREA *)calloc(s->number_area, sizeof(AREA));
    s->LC = (Inf_Ff_store *)calloc(s->number_line + 1, sizeof(Inf_Ff_store));
    s->LF = (Inf_Ff_store *)calloc(s->number_line + 1, sizeof(Inf_Ff_store));
    s->LCC = (Inf_Ff_store *)calloc(s->number_line2 + 1, sizeof(Inf_Ff_store));
    orc_geo_init(g, s->AD, areas, s->LC, s->LF, s->LCC);
    s->blk = (orc_block *)calloc(s->number_area, sizeof(orc_block));
    for (int a = 0; a < s->number_area; a++) {
        block_alloc(&s->blk[a], &areas[a]);
        init_general(&s->cfg, &s->blk[a]);
    }
    free(areas);
    /* Multi-block-LBM.cu:42-58, 64-77 */
    s->t3 = (Ff_Store *)calloc(s->number_line + 1, sizeof(Ff_Store));
    s->s3 = (FM_spline3 *)calloc(s->number_line + 1, sizeof(FM_spline3));
    for (int b = 0; b < s->number_line; b++) {
        Ff_Store *t = &s->t3[b];
        t->marker[0] = 2;
        t->marker[1] = 1;
        t->marker[2] = 0;
        t->N = s->LC[b].Nf;
        F_k3 *ft = &t->f_t1;
        for (int sl = 0; sl < 3; sl++)
            for (int k = 0; k < 9; k++) ft[sl].F[k] = dalloc(t->N);
        FM_spline3 *S = &s->s3[b];
        S->N = s->LC[b].Nc;
        S->l = dalloc(S->N);
        S->m = dalloc(S->N);
        S->u = dalloc(S->N);
        S->Varibs_M = dalloc(12 * (size_t)S->N);
        double **fm = &S->rho;
        for (int i = 0; i < 12; i++) fm[i] = dalloc(S->N);
        /* init_Fstore (init_multi.h:3-28) and init_tri_A (:30-51) */
        orc_block *bf = &s->blk[s->LC[b].areaf_n];
        for (int ip = 0; ip < t->N; ip++) {
            int ixy;
            line_node(&s->LC[b].line_F, ip, bf->area.nx, &ixy);
            for (int k = 0; k < 9; k++) {
                t->f_t1.F[k][ip] = bf->pp.F[k][ixy];
                t->f_t.F[k][ip] = bf->pp.F[k][ixy];
                t->f_t_1.F[k][ip] = bf->pp.F[k][ixy];
            }
        }
        orc_tri_A(S->N, S->l, S->m, S->u);
    }
    /* Multi-block-LBM.cu:125-131 */
    orc_block *last = &s->blk[s->number_area - 1];
    s->N_drag = drag_links(last, NULL);
    s->links = (Int3 *)calloc(s->N_drag + 1, sizeof(Int3));
    drag_links(last, s->links);
    s->capser = 1024;
    for (int k = 0; k < 5; k++) s->ser[k] = dalloc(s->capser);
    return s;
}

void orc_sim_destroy(orc_sim *s)
{
    if (!s) return;
    for (int a = 0; a < s->number_area; a++) block_free(&s->blk[a]);
    for (int b = 0; b < s->number_line; b++) {
        F_k3 *ft = &s->t3[b].f_t1;
        for (int sl = 0; sl < 3; sl++)
            for (int k = 0; k < 9; k++) free(ft[sl].F[k]);
        FM_spline3 *S = &s->s3[b];
        free(S->l);
        free(S->m);
        free(S->u);
        free(S->Varibs_M);
        double **fm = &S->rho;
        for (int i = 0; i < 12; i++) free(fm[i]);
    }
    for (int k = 0; k < 5; k++) free(s->ser[k]);
    free(s->t3);
    free(s->s3);
    free(s->links);
    free(s->blk);
    free(s->LC);
    free(s->LF);
    free(s->LCC);
    free(s->AD);
    free(s);
}

static void push_series(orc_sim *s, const double v[5])
{
    if (s->nser == s->capser) {
        s->capser *= 2;
        for (int k = 0; k < 5; k++) s->ser[k] = (double *)realloc(s->ser[k], sizeof(double) * s->capser);
    }
    for (int k = 0; k < 5; k++) s->ser[k][s->nser] = v[k];
    s->nser++;
}

/* Multi-block-LBM.cu:148-158 */
int orc_sim_run(orc_sim *s, int n)
{
    double U0_2 = sqrt(s->cfg.U0[0] * s->cfg.U0[0] + s->cfg.U0[1] * s->cfg.U0[1]);
    const AREA *al = &s->blk[s->number_area - 1].area;
    for (int t = 0; t < n; t++) {
        evolution_rec(s, 0);
        double v[5];
        v[0] = error_l2(s);
        double f[2] = {0.0, 0.0};
        if (s->N_drag > 0) get_drag(s, f);
        v[1] = f[0];
        v[2] = f[1];
        double nrm = s->cfg.rho0 * U0_2 * U0_2 * al->m * s->cfg.R;
        v[3] = f[0] / nrm;
        v[4] = f[1] / nrm;
        push_series(s, v);
        s->steps++;
    }
    return 0;
}

/* error_L2 and getDrag_block on the current state, without a step:
 * out = residual, Fx, Fy (updates u_prev like the reference) */
int orc_sim_diag(orc_sim *s, double out[3])
{
    out[0] = error_l2(s);
    double f[2] = {0.0, 0.0};
    if (s->N_drag > 0) get_drag(s, f);
    out[1] = f[0];
    out[2] = f[1];
    return 0;
}

/* EXT: velocity L2 distance from the analytic plane Poiseuille profile over
 * FLUID nodes, 1/m block weights (the normalisation of error_L2.h:82-87) */
double orc_sim_poiseuille_l2(const orc_sim *s)
{
    double num = 0.0, den = 0.0;
    for (int a = 0; a < s->number_area; a++) {
        const orc_block *b = &s->blk[a];
        double s1 = 0.0, s2 = 0.0;
        for (int j = 0; j < b->area.ny; j++)
            for (int i = 0; i < b->area.nx; i++) {
                int ixy = i + j * b->area.nx;
                if (b->pp.solid[ixy] != FLUID) continue;
                double ua = inlet_u(&s->cfg, b->pp.y[ixy]);
                double ux = b->pp.ux[ixy], uy = b->pp.uy[ixy];
                s1 += (ux - ua) * (ux - ua) + uy * uy;
                s2 += ua * ua;
            }
        num += s1 / b->area.m;
        den += s2 / b->area.m;
    }
    return sqrt(num) / sqrt(den);
}

int orc_sim_evolve(orc_sim *s, int n)
{
    for (int t = 0; t < n; t++) {
        evolution_rec(s, 0);
        s->steps++;
    }
    return 0;
}

int orc_sim_steps_done(const orc_sim *s) { return s->steps; }
int orc_sim_num_areas(const orc_sim *s) { return s->number_area; }
int orc_sim_num_lines(const orc_sim *s, int which) { return which == 0 ? s->number_line : s->number_line2; }
void orc_sim_area(const orc_sim *s, int a, AREA *out) { *out = s->blk[a].area; }
void orc_sim_line(const orc_sim *s, int which, int b, Inf_Ff_store *out)
{
    if (which == 0)
        *out = s->LC[b];
    else if (which == 1)
        *out = s->LCC[b];
    else
        *out = s->LF[b];
}

static double *field_ptr(const orc_block *b, int field)
{
    if (field >= 1 && field <= 25) {
        double *const *v = &b->pp.x;
        return v[field - 1];
    }
    if (field == 26) return b->upx;
    if (field == 27) return b->upy;
    return NULL;
}

int orc_sim_get(const orc_sim *s, int a, int field, double *out)
{
    const orc_block *b = &s->blk[a];
    size_t n = (size_t)b->area.nx * b->area.ny;
    if (field == 0) {
        for (size_t i = 0; i < n; i++) out[i] = b->pp.solid[i];
        return 0;
    }
    double *p = field_ptr(b, field);
    if (!p) return -1;
    memcpy(out, p, n * sizeof(double));
    return 0;
}

int orc_sim_set(orc_sim *s, int a, int field, const double *in)
{
    orc_block *b = &s->blk[a];
    size_t n = (size_t)b->area.nx * b->area.ny;
    if (field == 0) {
        for (size_t i = 0; i < n; i++) b->pp.solid[i] = (int)in[i];
        return 0;
    }
    double *p = field_ptr(b, field);
    if (!p) return -1;
    memcpy(p, in, n * sizeof(double));
    return 0;
}

int orc_sim_drag_links(const orc_sim *s, Int3 *out, int cap)
{
    int n = s->N_drag < cap ? s->N_drag : cap;
    if (out) memcpy(out, s->links, sizeof(Int3) * n);
    return s->N_drag;
}

int orc_sim_series(const orc_sim *s, int kind, double *out, int cap)
{
    int n = s->nser < cap ? s->nser : cap;
    if (out) memcpy(out, s->ser[kind], sizeof(double) * n);
    return s->nser;
}

int orc_sim_ring(const orc_sim *s, int b, int slot, int k, double *out)
{
    const F_k3 *ft = &s->t3[b].f_t1;
    memcpy(out, ft[slot].F[k], sizeof(double) * s->t3[b].N);
    return s->t3[b].N;
}

void orc_sim_markers(const orc_sim *s, int b, int out[3])
{
    for (int p = 0; p < 3; p++) out[p] = s->t3[b].marker[p];
}

/* Test hooks: the cycle of evolution_rec with each block's stream-collide
 * replaced by a caller function (the caller then sets the block's state
 * through orc_sim_set), and single interface operations on one line. */
void orc_sim_set_block_hook(orc_sim *s, orc_block_hook hook, void *ctx)
{
    s->hook = hook;
    s->hook_ctx = ctx;
}

/* space_interpolation_Spline3 of line b into ring slot marker[0] (no rotation) */
void orc_sim_spline_line(orc_sim *s, int b) { space_interpolation_spline3(s, b); }

/* which 0: c_f_boundary_transfer of line b from slot marker[2]; 1: f_c_boundary_transfer */
void orc_sim_transfer_line(orc_sim *s, int b, int which)
{
    if (which == 0) {
        F_k3 *ft = &s->t3[b].f_t1;
        c_f_boundary_transfer(&s->blk[s->LC[b].areaf_n], &s->LC[b], s->t3[b].N, &ft[s->t3[b].marker[2]]);
    } else {
        f_c_boundary_transfer(&s->blk[s->LF[b].areac_n], &s->blk[s->LF[b].areaf_n], &s->LF[b]);
    }
}

int orc_sim_spline_M(const orc_sim *s, int b, double *out)
{
    memcpy(out, s->s3[b].Varibs_M, sizeof(double) * 12 * s->s3[b].N);
    return 12 * s->s3[b].N;
}

/* ------------------------------------------------------------------------ */
/* single periodic block (bandwidth benchmark workload, CPU baseline)         */
/* ------------------------------------------------------------------------ */
struct orc_periodic {
    mbl_config cfg;
    orc_block b;
};

orc_periodic *orc_periodic_create(int nx, int ny, double tau, double rho0,
                                  unsigned long long seed, double sigma)
{
    orc_periodic *p = (orc_periodic *)calloc(1, sizeof(orc_periodic));
    mbl_config *c = &p->cfg;
    c->rho0 = rho0;
    c->U0[0] = 0.0;
    c->U0[1] = 0.0;
    c->LX = nx;
    c->LY = ny;
    c->R = 0.0;
    c->CX = -1e30;
    c->CY = -1e30;
    c->bc_mode = MBL_BC_PERIODIC;
    c->init_mode = MBL_INIT_RANDOM;
    c->seed = seed;
    c->sigma = sigma;
    AREA a;
    a.m = 1;
    a.nx = nx;
    a.ny = ny;
    a.base_x = 0.0;
    a.base_y = 0.0;
    a.tau = tau;
    a.rv_tau = 1.0 / tau;
    block_alloc(&p->b, &a);
    init_general(c, &p->b);
    return p;
}

void orc_periodic_destroy(orc_periodic *p)
{
    if (!p) return;
    block_free(&p->b);
    free(p);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

double orc_periodic_run(orc_periodic *p, int n)
{
    double t0 = now_s();
    for (int t = 0; t < n; t++) evolution_area_stb(&p->cfg, &p->b, 1);
    return now_s() - t0;
}

int orc_periodic_get(const orc_periodic *p, int field, double *out)
{
    size_t n = (size_t)p->b.area.nx * p->b.area.ny;
    if (field == 0) {
        for (size_t i = 0; i < n; i++) out[i] = p->b.pp.solid[i];
        return 0;
    }
    double *q = field_ptr(&p->b, field);
    if (!q) return -1;
    memcpy(out, q, n * sizeof(double));
    return 0;
}
