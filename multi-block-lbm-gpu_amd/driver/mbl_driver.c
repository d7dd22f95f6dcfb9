/*
 * mbl_driver.c -- C host program of the multi-block LBM solver. It follows
 * the reference driver (Multi-block-LBM.cu:19-230): geometry, allocation of
 * every block and interface buffer, init_General, drag-link set-up,
 * init_Fstore / init_tri_A, then _LOOPS x _T coarse steps of
 * evolution -> error_L2 -> getDrag_block with the per-chunk timing line of
 * time_t.dat. It includes only include/mbl.h and links libmbl_hip.
 *
 *   mbl_driver [layout] [steps] [T] [output] [fused]
 *     layout: GEO_1_1 | GEO_2_2 | GEO_2_5 (default) | GEO_4_9 | C1
 *     steps : total coarse steps (default 1000), T: steps per timing chunk
 *     output: 1 writes steady_<n>.dat at the start and after every chunk, as
 *             the reference does (Multi-block-LBM.cu:120, 170); default 0
 *     fused : 1 runs each chunk as one mbl_evolution_step call (the same loop
 *             body, with the residual summed inside the cycle); default 0,
 *             the reference's three calls per step
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/mbl.h"

static double wall_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void *dalloc(size_t bytes)
{
    void *p = mbl_malloc(bytes);
    if (!p) {
        fprintf(stderr, "device allocation of %zu bytes failed\n", bytes);
        exit(1);
    }
    return p;
}

int main(int argc, char **argv)
{
    const char *layout = argc > 1 ? argv[1] : "GEO_2_5";
    int steps = argc > 2 ? atoi(argv[2]) : 1000;
    int T = argc > 3 ? atoi(argv[3]) : 1000;
    if (T <= 0) T = 1000;
    const int write_fields = argc > 4 ? atoi(argv[4]) : 0;
    const int fused = argc > 5 ? atoi(argv[5]) : 0;

    mbl_geometry g;
    mbl_config cfg;
    mbl_config_default(&cfg);
    P4 ad_buf[4 * MBL_MAX_LEVEL];
    P3i cf_buf[8 * MBL_MAX_LEVEL], cc_buf[8 * MBL_MAX_LEVEL];
    if (!strcmp(layout, "C1")) {
        /* two-level 300x80-cell channel around a D=10 cylinder at Re=20 */
        mbl_channel_spec s;
        memset(&s, 0, sizeof(s));
        s.LX = 301;
        s.LY = 81;
        s.levels = 2;
        s.region[1].xstart = 20;
        s.region[1].xend = 100;
        s.region[1].ystart = 20;
        s.region[1].yend = 60;
        s.Re = 20.0;
        s.U = 0.1;
        s.R = 5.0;
        if (mbl_geometry_channel(&s, &g, ad_buf, cf_buf, cc_buf)) return 1;
        cfg.LX = 301;
        cfg.LY = 81;
        cfg.R = 5.0;
        cfg.CX = 60.0;
        cfg.CY = 40.0;
        cfg.bc_mode = MBL_BC_CHANNEL;
        cfg.init_mode = MBL_INIT_REST;
    } else if (mbl_geometry_preset(layout, &g)) {
        fprintf(stderr, "unknown layout %s\n", layout);
        return 1;
    }
    cfg.LEVEL = g.LEVEL;
    for (int l = 0; l < g.LEVEL; l++) cfg.mi[l] = g.mi[l];
    cfg.number_line2 = g.number_line2;
    cfg.number_area = g.number_area;
    cfg.flush_interval = T;

    const int na = g.number_area, nl = g.number_line, nl2 = g.number_line2;
    P4 *AreaDefine = calloc(na, sizeof(P4));
    AREA *Areas = calloc(na, sizeof(AREA));
    Inf_Ff_store *LC = calloc(nl + 1, sizeof(Inf_Ff_store));
    Inf_Ff_store *LF = calloc(nl + 1, sizeof(Inf_Ff_store));
    Inf_Ff_store *LCC = calloc(nl2 + 1, sizeof(Inf_Ff_store));
    if (mbl_geo_init(&g, AreaDefine, Areas, LC, LF, LCC)) {
        fprintf(stderr, "geo_init rejected the layout\n");
        return 1;
    }
    mbl_handle_t h;
    int rc = mbl_handle_create(&h, &cfg);
    if (rc) {
        fprintf(stderr, "mbl_handle_create: %s\n", mbl_error_string(rc));
        return 1;
    }

    /* blocks (Multi-block-LBM.cu:23-38) */
    D2Q9 *Vd = calloc(na, sizeof(D2Q9)), *Vh = calloc(na, sizeof(D2Q9));
    Lbd **upsx = calloc(na, sizeof(Lbd *)), **upsy = calloc(na, sizeof(Lbd *));
    mbl_dim3 block = {16, 16, 1};
    mbl_dim3 *grid = calloc(na, sizeof(mbl_dim3));
    for (int a = 0; a < na; a++) {
        if (mbl_block_alloc(h, &Areas[a], &Vd[a], &upsx[a], &upsy[a])) return 1;
        Vh[a].solid = malloc(sizeof(int) * (size_t)Areas[a].nx * Areas[a].ny);
        grid[a].x = (Areas[a].nx + 15) / 16;
        grid[a].y = (Areas[a].ny + 15) / 16;
        grid[a].z = 1;
    }
    /* interface rings and spline work arrays (Multi-block-LBM.cu:42-77) */
    Ff_Store *t3 = calloc(nl + 1, sizeof(Ff_Store));
    FM_spline3 *s3 = calloc(nl + 1, sizeof(FM_spline3));
    for (int b = 0; b < nl; b++) {
        t3[b].marker[0] = 2;
        t3[b].marker[1] = 1;
        t3[b].marker[2] = 0;
        t3[b].N = LC[b].Nf;
        F_k3 *ft = &t3[b].f_t1;
        for (int sl = 0; sl < 3; sl++)
            for (int k = 0; k < MBL_Q; k++) ft[sl].F[k] = dalloc(sizeof(Lbd) * t3[b].N);
        s3[b].N = LC[b].Nc;
        s3[b].l = dalloc(sizeof(Lbd) * s3[b].N);
        s3[b].m = dalloc(sizeof(Lbd) * s3[b].N);
        s3[b].u = dalloc(sizeof(Lbd) * s3[b].N);
        s3[b].Varibs_M = dalloc(sizeof(Lbd) * 12 * s3[b].N);
        Lbd **fm = &s3[b].rho;
        for (int i = 0; i < 12; i++) fm[i] = dalloc(sizeof(Lbd) * s3[b].N);
    }
    /* initial state (Multi-block-LBM.cu:116-117) */
    for (int a = 0; a < na; a++) init_General(grid[a], block, Areas[a], Vd[a], upsx[a], upsy[a]);
    /* drag links on the last block (Multi-block-LBM.cu:125-131) */
    const int last = na - 1;
    int N_drag = sum_N_drag(block, grid[last], Areas[last], Vd[last]);
    Int3 *d_links = dalloc(sizeof(Int3) * (N_drag > 0 ? N_drag : 1));
    fdrag_store(Areas[last], Vh[last], d_links, Vd[last], N_drag);
    for (int b = 0; b < nl; b++) { /* Multi-block-LBM.cu:136-140 */
        init_Fstore(grid[0], block, Areas[LC[b].areaf_n], Vd[LC[b].areaf_n], LC[b], t3[b]);
        init_tri_A(grid[0], block, Areas[LC[b].areac_n], Vd[LC[b].areac_n], LC[b], s3[b]);
    }
    if ((rc = mbl_synchronize(h))) {
        fprintf(stderr, "set-up failed: %s\n", mbl_error_string(rc));
        return 1;
    }
    printf("%s: %d blocks, %d coarse/fine lines, %d same-level lines, %d drag links\n", layout, na, nl, nl2, N_drag);
    if (write_fields) output(0, Areas, Vh, Vd); /* Multi-block-LBM.cu:120 */

    int asp[MBL_MAX_LEVEL + 1], ls3[MBL_MAX_LEVEL + 1], ltc[MBL_MAX_LEVEL + 1], lcc[MBL_MAX_LEVEL + 1];
    for (int l = 0; l <= g.LEVEL; l++) {
        asp[l] = g.ASP[l];
        ls3[l] = g.LS3CFFCP[l];
        ltc[l] = g.LTCFFCP[l];
        lcc[l] = g.LCCP[l];
    }
    double updates = 0.0;
    for (int a = 0; a < na; a++) updates += (double)Areas[a].nx * Areas[a].ny * Areas[a].m;
    /* time loop (Multi-block-LBM.cu:144-171) */
    int done = 0;
    while (done < steps) {
        int chunk = steps - done < T ? steps - done : T;
        double t0 = wall_s();
        if (fused) {
            mbl_evolution_step(h, done, chunk, asp, ls3, ltc, lcc, block, grid, Areas, Vh, Vd, LC, LF, LCC, t3, s3,
                               upsx, upsy, d_links, N_drag, NULL);
        } else {
            for (int n = 0; n < chunk; n++) {
                int t = done + n;
                evolution(t, asp, ls3, ltc, lcc, 0, block, grid, Areas, Vh, Vd, LC, LF, LCC, t3, s3, h);
                error_L2(t, na, block, grid, Areas, Vd, upsx, upsy, NULL, NULL, NULL, NULL);
                if (N_drag > 0) getDrag_block(t, d_links, Areas[last], Vd[last], NULL, N_drag);
            }
        }
        mbl_flush_diagnostics(h);
        if ((rc = mbl_synchronize(h))) {
            fprintf(stderr, "step %d failed: %s\n", done, mbl_error_string(rc));
            return 1;
        }
        double dt = wall_s() - t0;
        done += chunk;
        printf("%f seconds\n%d\n", dt, done);
        FILE *fp = fopen("time_t.dat", "a");
        if (fp) {
            fprintf(fp, "%d %g\n", done, dt);
            fclose(fp);
        }
        printf("MLUPS %.1f\n", updates * chunk / dt / 1e6);
        if (write_fields) output(done, Areas, Vh, Vd); /* Multi-block-LBM.cu:170 */
    }
    double diag[5];
    mbl_last_diagnostics(h, diag);
    printf("residual %g Cd %g Cl %g\n", diag[0], diag[3], diag[4]);

    for (int b = 0; b < nl; b++) {
        F_k3 *ft = &t3[b].f_t1;
        for (int sl = 0; sl < 3; sl++)
            for (int k = 0; k < MBL_Q; k++) mbl_free(ft[sl].F[k]);
        mbl_free(s3[b].l);
        mbl_free(s3[b].m);
        mbl_free(s3[b].u);
        mbl_free(s3[b].Varibs_M);
        Lbd **fm = &s3[b].rho;
        for (int i = 0; i < 12; i++) mbl_free(fm[i]);
    }
    mbl_free(d_links);
    for (int a = 0; a < na; a++) {
        mbl_block_free(&Vd[a], upsx[a], upsy[a]);
        free(Vh[a].solid);
    }
    mbl_handle_destroy(h);
    free(Vd);
    free(Vh);
    free(upsx);
    free(upsy);
    free(grid);
    free(t3);
    free(s3);
    free(AreaDefine);
    free(Areas);
    free(LC);
    free(LF);
    free(LCC);
    return 0;
}
