/*
 * mbl_geo.c -- block/level layout of the multi-block grid (pure C, host only).
 *
 * mbl_geo_init restates geo_init and Line_init (geo/init_geo.h:17-165): block
 * extents from the zone table, cumulative refinement m, nx/ny, per-level tau
 * and the interface line descriptors. The preset tables are the data of
 * geo/geometry.h (they have to be identical for the layouts to be the same).
 */
#include <math.h>
#include <string.h>

#include "../../include/mbl.h"

/* One interface: coarse-side (C) and fine-side (F) descriptors of a block
 * pair in direction dir (0 left, 1 right, 2 down, 3 up = where the second
 * block lies relative to the first). Same index rules as init_geo.h:17-86. */
static void mbl_line_init(int dir, const P4 *aC, const AREA *sC, Inf_Ff_store *lC, const P4 *aF, const AREA *sF,
                          Inf_Ff_store *lF)
{
    int vertical = (dir == 0 || dir == 1);
    /* position across the line */
    int posC_C, posC_F, posF_C, posF_F;
    if (vertical) {
        posC_C = (dir == 0) ? sC->m : sC->nx - 1 - sC->m;
        posC_F = (dir == 0) ? sF->nx - 1 : 0;
        posF_C = (dir == 0) ? 0 : sC->nx - 1;
        posF_F = (dir == 0) ? sF->nx - sF->m - 1 : sF->m;
    } else {
        posC_C = (dir == 2) ? sC->m : sC->ny - 1 - sC->m;
        posC_F = (dir == 2) ? sF->ny - 1 : 0;
        posF_C = (dir == 2) ? 0 : sC->ny - 1;
        posF_F = (dir == 2) ? sF->ny - sF->m - 1 : sF->m;
    }
    /* extent along the line: the overlap of the two blocks */
    int c0C = vertical ? aC->ystart : aC->xstart, c1C = vertical ? aC->yend : aC->xend;
    int c0F = vertical ? aF->ystart : aF->xstart, c1F = vertical ? aF->yend : aF->xend;
    int nC = vertical ? sC->ny : sC->nx, nF = vertical ? sF->ny : sF->nx;
    int s_C = (c0F - c0C <= 0) ? 0 : (c0F - c0C) * sC->m;
    int e_C = (c1F - c1C >= 0) ? nC - 1 : (c1F - c0C) * sC->m;
    int s_F = (c0C - c0F <= 0) ? 0 : (c0C - c0F) * sF->m;
    int e_F = (c1C - c1F >= 0) ? nF - 1 : (c1C - c0F) * sF->m;
    P4 *pl[4] = {&lC->line_C, &lC->line_F, &lF->line_C, &lF->line_F};
    int pos[4] = {posC_C, posC_F, posF_C, posF_F};
    int st[4] = {s_C, s_F, s_C, s_F}, en[4] = {e_C, e_F, e_C, e_F};
    for (int q = 0; q < 4; q++) {
        if (vertical) {
            pl[q]->xstart = pl[q]->xend = pos[q];
            pl[q]->ystart = st[q];
            pl[q]->yend = en[q];
        } else {
            pl[q]->ystart = pl[q]->yend = pos[q];
            pl[q]->xstart = st[q];
            pl[q]->xend = en[q];
        }
    }
    lC->Nc = lF->Nc = e_C - s_C + 1;
    lC->Nf = lF->Nf = e_F - s_F + 1;
}

int mbl_geo_init(const mbl_geometry *g, P4 *AD, AREA *Areas, Inf_Ff_store *LC, Inf_Ff_store *LF,
                 Inf_Ff_store *LCC)
{
    if (!g || !AD || !Areas || g->LEVEL < 1 || g->LEVEL > MBL_MAX_LEVEL || g->number_area < 1) return MBL_E_INVALID;
    if (!g->AreaDefine && !g->AreaZone) return MBL_E_INVALID;
    if (g->ASP[0] != 0 || g->ASP[g->LEVEL] != g->number_area) return MBL_E_INVALID;
    for (int a = 0; a < g->number_area; a++) {
        if (g->AreaDefine) {
            AD[a] = g->AreaDefine[a];
        } else {
            const P4d4i *z = &g->AreaZone[a];
            AD[a].xstart = (int)(z->reduced.xstart * (g->LX - 1)) + z->move.xstart;
            AD[a].xend = (int)(z->reduced.xend * (g->LX - 1)) + z->move.xend;
            AD[a].ystart = (int)(z->reduced.ystart * (g->LY - 1)) + z->move.ystart;
            AD[a].yend = (int)(z->reduced.yend * (g->LY - 1)) + z->move.yend;
        }
        if (AD[a].xend <= AD[a].xstart || AD[a].yend <= AD[a].ystart) return MBL_E_INVALID;
    }
    /* kinematic viscosity from Re on the cylinder diameter, tau per level
     * with the rescaling tau_l = 0.5 + mi (tau_{l-1} - 0.5) (init_geo.h:99-120) */
    double nu = g->U0_2 * 2 * g->R / g->Re;
    double tau = 0.5 + 3.0 * nu;
    int m = 1;
    for (int l = 0; l < g->LEVEL; l++) {
        if (g->mi[l] < 1) return MBL_E_INVALID;
        m *= g->mi[l];
        tau = 0.5 + g->mi[l] * (tau - 0.5);
        for (int a = g->ASP[l]; a < g->ASP[l + 1]; a++) {
            Areas[a].m = m;
            Areas[a].nx = (AD[a].xend - AD[a].xstart) * m + 1;
            Areas[a].ny = (AD[a].yend - AD[a].ystart) * m + 1;
            Areas[a].base_x = AD[a].xstart;
            Areas[a].base_y = AD[a].ystart;
            Areas[a].tau = tau;
            Areas[a].rv_tau = 1.0 / tau;
        }
    }
    for (int b = 0; b < g->number_line; b++) {
        int nc = g->NeborCF[b].nc, nf = g->NeborCF[b].nf;
        if (nc < 0 || nc >= g->number_area || nf < 0 || nf >= g->number_area) return MBL_E_INVALID;
        LC[b].areac_n = LF[b].areac_n = nc;
        LC[b].areaf_n = LF[b].areaf_n = nf;
        mbl_line_init(g->NeborCF[b].direction, &AD[nc], &Areas[nc], &LC[b], &AD[nf], &Areas[nf], &LF[b]);
        /* the rescale para = (tau_f-1)/(r(tau_c-1)) is singular at tau = 1 (assit_multi.h:176, 304) */
        if (Areas[nc].tau == 1.0 || Areas[nf].tau == 1.0) return MBL_E_INVALID;
        int r = Areas[nf].m / Areas[nc].m;
        if (LC[b].Nf != r * (LC[b].Nc - 1) + 1) return MBL_E_INVALID;
    }
    for (int c = 0; c < g->number_line2; c += 2) {
        int a = g->NeborCC[c].nc, b = g->NeborCC[c].nf;
        if (a < 0 || a >= g->number_area || b < 0 || b >= g->number_area || c + 1 >= g->number_line2)
            return MBL_E_INVALID;
        LCC[c].areac_n = a;
        LCC[c].areaf_n = b;
        mbl_line_init(g->NeborCC[c].direction, &AD[a], &Areas[a], &LCC[c], &AD[b], &Areas[b], &LCC[c + 1]);
        /* the return line copies from b to a: swap its source and target */
        LCC[c + 1].areac_n = g->NeborCC[c + 1].nc;
        LCC[c + 1].areaf_n = g->NeborCC[c + 1].nf;
        P4 t = LCC[c + 1].line_C;
        LCC[c + 1].line_C = LCC[c + 1].line_F;
        LCC[c + 1].line_F = t;
    }
    return 0;
}

/* ---- geo/geometry.h tables ---- */
static const P4d4i zone_1_1[1] = {{{0.0, 1.0, 0.0, 1.0}, {0, 0, 0, 0}}};

static const P4d4i zone_2_2[2] = {{{0.0, 1.0, 0.0, 0.5}, {0, 0, 0, 1}}, {{0.0, 1.0, 0.5, 1.0}, {0, 0, 0, 0}}};
static const P3i cf_2_2[1] = {{0, 1, 3}};

static const P4d4i zone_2_5[5] = {
    {{0.0, 1.0, 0.75, 1.0}, {0, 0, -1, 0}},
    {{0.0, 1.0, 0.0, 0.25}, {0, 0, 0, 1}},
    {{0.0, 0.25 - 0.0625, 0.25, 0.75}, {0, 1, 0, 0}},
    {{0.25 + 0.0625, 1.0, 0.25, 0.75}, {-1, 0, 0, 0}},
    {{0.25 - 0.0625, 0.25 + 0.0625, 0.25, 0.75}, {0, 0, 0, 0}},
};
static const P3i cf_2_5[4] = {{0, 4, 2}, {1, 4, 3}, {2, 4, 1}, {3, 4, 0}};
static const P3i cc_2_5[8] = {{0, 2, 2}, {2, 0, 2}, {0, 3, 2}, {3, 0, 2}, {1, 2, 3}, {2, 1, 3}, {1, 3, 3}, {3, 1, 3}};

static const P4d4i zone_4_9[10] = {
    {{0.0, 0.1, 0, 1.0}, {0, 1, 0, 0}},          {{0.1, 0.4, 0.75, 1.0}, {0, 0, -1, 0}},
    {{0.4, 1.0, 0.0, 1.0}, {-1, 0, 0, 0}},       {{0.1, 0.4, 0.0, 0.25}, {0, 0, 0, 1}},
    {{0.1, 0.15, 0.25, 0.75}, {0, 1, 0, 0}},     {{0.35, 0.4, 0.25, 0.75}, {-1, 0, 0, 0}},
    {{0.15, 0.25, 0.625, 0.75}, {0, 0, -1, 0}},  {{0.25, 0.35, 0.25, 0.75}, {-1, 0, 0, 0}},
    {{0.15, 0.25, 0.25, 0.375}, {0, 0, 0, 1}},   {{0.15, 0.25, 0.375, 0.625}, {0, 0, 0, 0}},
};
static const P3i cf_4_9[17] = {{0, 4, 1}, {1, 4, 2}, {1, 5, 2}, {2, 5, 0}, {3, 5, 3}, {3, 4, 3},
                               {1, 6, 2}, {1, 7, 2}, {3, 7, 3}, {3, 8, 3}, {4, 6, 1}, {5, 7, 0},
                               {4, 8, 1}, {4, 9, 1}, {6, 9, 2}, {7, 9, 0}, {8, 9, 3}};
static const P3i cc_4_9[12] = {{0, 1, 1}, {1, 0, 1}, {2, 1, 0}, {1, 2, 0}, {2, 3, 0}, {3, 2, 0},
                               {0, 3, 1}, {3, 0, 1}, {7, 6, 0}, {6, 7, 0}, {7, 8, 0}, {8, 7, 0}};

static void preset_common(mbl_geometry *g)
{
    memset(g, 0, sizeof(*g));
    g->LX = 768; /* parameters.h:33-37 */
    g->LY = 320;
    g->Re = 10;
    g->U0_2 = 0.1;
    g->R = 10;
}

static void set_prefix(int *dst, const int *src, int n)
{
    for (int i = 0; i < n; i++) dst[i] = src[i];
}

int mbl_geometry_preset(const char *name, mbl_geometry *g)
{
    if (!name || !g) return MBL_E_INVALID;
    preset_common(g);
    if (!strcmp(name, "GEO_1_1")) {
        static const int asp[2] = {0, 1}, z[2] = {0, 0};
        g->LEVEL = 1;
        g->mi[0] = 1;
        g->number_area = 1;
        set_prefix(g->ASP, asp, 2);
        set_prefix(g->LTCFFCP, z, 2);
        set_prefix(g->LS3CFFCP, z, 2);
        set_prefix(g->LCCP, z, 2);
        g->AreaZone = zone_1_1;
        return 0;
    }
    if (!strcmp(name, "GEO_2_2")) {
        static const int asp[3] = {0, 1, 2}, ltc[3] = {0, 0, 1}, ls3[3] = {0, 1, 1}, lcc[3] = {0, 0, 0};
        g->LEVEL = 2;
        g->mi[0] = 1;
        g->mi[1] = 2;
        g->number_area = 2;
        g->number_line = 1;
        set_prefix(g->ASP, asp, 3);
        set_prefix(g->LTCFFCP, ltc, 3);
        set_prefix(g->LS3CFFCP, ls3, 3);
        set_prefix(g->LCCP, lcc, 3);
        g->AreaZone = zone_2_2;
        g->NeborCF = cf_2_2;
        return 0;
    }
    if (!strcmp(name, "GEO_2_5")) {
        static const int asp[3] = {0, 4, 5}, ltc[3] = {0, 0, 4}, ls3[3] = {0, 4, 4}, lcc[3] = {0, 8, 8};
        g->LEVEL = 2;
        g->mi[0] = 1;
        g->mi[1] = 2;
        g->number_area = 5;
        g->number_line = 4;
        g->number_line2 = 8;
        set_prefix(g->ASP, asp, 3);
        set_prefix(g->LTCFFCP, ltc, 3);
        set_prefix(g->LS3CFFCP, ls3, 3);
        set_prefix(g->LCCP, lcc, 3);
        g->AreaZone = zone_2_5;
        g->NeborCF = cf_2_5;
        g->NeborCC = cc_2_5;
        return 0;
    }
    if (!strcmp(name, "GEO_4_9")) {
        static const int asp[5] = {0, 4, 6, 9, 10}, ltc[5] = {0, 0, 6, 13, 17}, ls3[5] = {0, 10, 14, 17, 17},
                         lcc[5] = {0, 8, 8, 12, 12};
        g->LEVEL = 4;
        g->mi[0] = 1;
        g->mi[1] = 2;
        g->mi[2] = 2;
        g->mi[3] = 2;
        g->number_area = 10;
        g->number_line = 17;
        g->number_line2 = 12;
        set_prefix(g->ASP, asp, 5);
        set_prefix(g->LTCFFCP, ltc, 5);
        set_prefix(g->LS3CFFCP, ls3, 5);
        set_prefix(g->LCCP, lcc, 5);
        g->AreaZone = zone_4_9;
        g->NeborCF = cf_4_9;
        g->NeborCC = cc_4_9;
        return 0;
    }
    return MBL_E_INVALID;
}

/* ---- nested channel builder ---- */
int mbl_geometry_channel(const mbl_channel_spec *s, mbl_geometry *g, P4 *AD, P3i *CF, P3i *CC)
{
    if (!s || !g || !AD || !CF || !CC || s->levels < 1 || s->levels > MBL_MAX_LEVEL) return MBL_E_INVALID;
    memset(g, 0, sizeof(*g));
    g->LX = s->LX;
    g->LY = s->LY;
    g->Re = s->Re;
    g->U0_2 = s->U;
    g->R = s->R;
    g->LEVEL = s->levels;
    g->mi[0] = 1;
    for (int l = 1; l < s->levels; l++) g->mi[l] = 2;
    P4 outer = {0, s->LX - 1, 0, s->LY - 1};
    int na = 0, ncf = 0, ncc = 0;
    int ring_first[MBL_MAX_LEVEL]; /* first block of level l */
    int cf_count[MBL_MAX_LEVEL] = {0}, cc_count[MBL_MAX_LEVEL] = {0};
    for (int l = 0; l < s->levels; l++) {
        ring_first[l] = na;
        g->ASP[l] = na;
        if (l == s->levels - 1) {
            AD[na++] = outer;
        } else {
            P4 in = s->region[l + 1];
            if (in.xstart - outer.xstart < 2 || outer.xend - in.xend < 2 || in.ystart - outer.ystart < 2 ||
                outer.yend - in.yend < 2)
                return MBL_E_INVALID;
            P4 top = {outer.xstart, outer.xend, in.yend - 1, outer.yend};
            P4 bot = {outer.xstart, outer.xend, outer.ystart, in.ystart + 1};
            P4 left = {outer.xstart, in.xstart + 1, in.ystart, in.yend};
            P4 right = {in.xend - 1, outer.xend, in.ystart, in.yend};
            AD[na++] = top;
            AD[na++] = bot;
            AD[na++] = left;
            AD[na++] = right;
            outer = in;
        }
    }
    g->ASP[s->levels] = na;
    /* same-level pairs inside each ring: top/bottom with left/right */
    for (int l = 0; l < s->levels; l++) {
        g->LCCP[l] = ncc;
        if (l < s->levels - 1) {
            int t = ring_first[l], b = t + 1, lf = t + 2, rt = t + 3;
            int pairs[4][3] = {{t, lf, 2}, {t, rt, 2}, {b, lf, 3}, {b, rt, 3}};
            for (int p = 0; p < 4; p++) {
                CC[ncc].nc = pairs[p][0];
                CC[ncc].nf = pairs[p][1];
                CC[ncc].direction = pairs[p][2];
                ncc++;
                CC[ncc].nc = pairs[p][1];
                CC[ncc].nf = pairs[p][0];
                CC[ncc].direction = pairs[p][2];
                ncc++;
            }
            cc_count[l] = 8;
        }
    }
    g->LCCP[s->levels] = ncc;
    /* coarse/fine lines, grouped by coarse level */
    for (int l = 0; l + 1 < s->levels; l++) {
        int t = ring_first[l], b = t + 1, lf = t + 2, rt = t + 3;
        int f0 = ring_first[l + 1];
        int start = ncf;
        if (l + 1 == s->levels - 1) {
            int lines[4][3] = {{t, f0, 2}, {b, f0, 3}, {lf, f0, 1}, {rt, f0, 0}};
            for (int q = 0; q < 4; q++) {
                CF[ncf].nc = lines[q][0];
                CF[ncf].nf = lines[q][1];
                CF[ncf].direction = lines[q][2];
                ncf++;
            }
        } else {
            int ft = f0, fb = f0 + 1, fl = f0 + 2, fr = f0 + 3;
            int lines[8][3] = {{t, ft, 2},  {b, fb, 3},  {lf, fl, 1}, {lf, ft, 1},
                               {lf, fb, 1}, {rt, fr, 0}, {rt, ft, 0}, {rt, fb, 0}};
            for (int q = 0; q < 8; q++) {
                CF[ncf].nc = lines[q][0];
                CF[ncf].nf = lines[q][1];
                CF[ncf].direction = lines[q][2];
                ncf++;
            }
        }
        cf_count[l] = ncf - start;
    }
    /* LS3CFFCP by coarse level, LTCFFCP by fine level (the same lines) */
    int acc = 0;
    for (int l = 0; l <= s->levels; l++) {
        g->LS3CFFCP[l] = acc;
        if (l < s->levels) acc += cf_count[l];
    }
    g->LTCFFCP[0] = 0;
    acc = 0;
    for (int l = 1; l <= s->levels; l++) {
        g->LTCFFCP[l] = acc;
        if (l < s->levels) acc += cf_count[l - 1];
    }
    g->LTCFFCP[s->levels] = ncf;
    (void)cc_count;
    g->number_area = na;
    g->number_line = ncf;
    g->number_line2 = ncc;
    g->AreaDefine = AD;
    g->NeborCF = CF;
    g->NeborCC = CC;
    return 0;
}
