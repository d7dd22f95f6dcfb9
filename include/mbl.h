/*
 * mbl.h -- C ABI of libmbl_hip, the MI355X multi-block D2Q9 BGK lattice
 * Boltzmann time-stepping library.
 *
 * Pure C. No HIP or CUDA header is included; every device buffer is a plain
 * pointer obtained from mbl_malloc() (or any HIP allocation of the same
 * process). Two builds of the same source exist:
 *
 *   libmbl_hip.so    Lbd = float   (FP32 state; FP32/FP16/BF16 plane storage)
 *   libmbl_hip64.so  Lbd = double  (FP64 state and planes)
 *
 * A caller includes this header with MBL_FP64 defined when it links the FP64
 * build. Both builds export exactly the same symbol names.
 *
 * The drop-in entry points keep the names and argument order of the
 * reference program (zhangya9741/Multi-block-LBM-GPU):
 *
 *   init_General   lbm/init_block.h:2         (launched at Multi-block-LBM.cu:117)
 *   init_Fstore    multi/init_multi.h:3       (Multi-block-LBM.cu:138)
 *   init_tri_A     multi/init_multi.h:30      (Multi-block-LBM.cu:139)
 *   evolution      evolution.h:2-3            (Multi-block-LBM.cu:151-152)
 *   sum_N_drag     lbm/force_block.h:50       (Multi-block-LBM.cu:127)
 *   fdrag_store    lbm/force_block.h:78       (Multi-block-LBM.cu:131)
 *   getDrag_block  lbm/force_block.h:180      (Multi-block-LBM.cu:157)
 *   error_L2       lbm/error_L2.h:57-58       (Multi-block-LBM.cu:155)
 *
 * Differences at the boundary (each one deliberate):
 *   - `dim3` becomes the POD mbl_dim3; launch shapes passed in are accepted
 *     and ignored (the library picks its own).
 *   - `cusparseHandle_t` becomes mbl_handle_t, an opaque context carrying the
 *     stream, the graph cache, the physical constants that were __constant__
 *     globals (parameters.h:42-52) and the level tables (mi, LEVEL,
 *     number_line2) that evolution.h reads as globals.  Entry points of the
 *     reference that take no handle use the "current" handle (the last one
 *     created or made current), mirroring the reference's global constants.
 *   - AREA's real members are double in both builds (setup-time values).
 *   - Drag and residual outputs are double in both builds (FP64 accumulation).
 *   - State layout: F and f of a D2Q9 block are a ping-pong pair. After every
 *     evolution() call F holds the post-collision state and f the state one
 *     (sub)step older: evolution() swaps the two pointer sets in the caller's
 *     pps_d[] entries whenever a level took an odd number of substeps. The
 *     post-stream distribution of the reference (f after stb1) is implicit:
 *     f_post[k][x] = pps_d.f[k][x - e_k].
 *   - Diagnostics (error_L2.dat, fdrag_t.dat lines) are reduced on the device
 *     and written in batches (mbl_flush_diagnostics, automatic every
 *     flush_interval records and at handle destruction); the bytes written are
 *     the reference formats "%d\t%g\n" and "%d\t%g\t%g\t\n".
 *
 * Error convention: functions returning int return 0 on success and a
 * negative MBL_E* code on failure; void entry points record the code in the
 * handle (sticky, read with mbl_get_last_error).
 */
#ifndef MBL_H
#define MBL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef MBL_FP64
typedef double Lbd;
#else
typedef float Lbd;
#endif
typedef int Lbi;
typedef unsigned int Lbui;

#define MBL_Q 9
#define MBL_MAX_LEVEL 8

/* Node flags. FLUID, SOLID_SBB, BOUNDARY_INLET, BOUNDARY_OUTLET and
 * C_BOUNDARY keep the reference's values (lbm/lbm.h:26-32). The reference
 * declares INLET/OUTLET without using them; here they are the channel ends,
 * both set by non-equilibrium extrapolation from the neighbour's
 * post-collision state: the inlet imposes the parabolic velocity with the
 * density of its x+1 neighbour, the outlet imposes rho0 with the velocity of
 * its x-1 neighbour (zero-gradient velocity). MBL_SOLID_WALL (new) is a
 * full-way bounce-back node that does not take part in the drag sum. */
#define FLUID 0
#define SOLID_SBB 1
#define MBL_SOLID_WALL 2
#define BOUNDARY_INLET 5
#define BOUNDARY_OUTLET 7
#define C_BOUNDARY 9

/* ---- data model (lbm/data_type_lbm.h:7-51, multi/data_type_multi.h:7-31) ---- */

/* Per-block SoA state. Member order and contiguity are ABI: the 25 real
 * pointers starting at &x may be walked as an array (Multi-block-LBM.cu:31-37).
 * Planes are nx*ny elements, row-major, ixy = i + j*nx. With 16-bit plane
 * storage (mbl_config.storage = MBL_STORE_F16S / MBL_STORE_BF16S) the 18
 * f/F pointers address 2-byte elements holding f - w_k; with MBL_STORE_F32S
 * they hold f - w_k in float. */
typedef struct D2Q9 {
    Lbi *solid;
    Lbd *x, *y;
    Lbd *rho, *ux, *uy;
    Lbd *forcex, *forcey;
    Lbd *f[MBL_Q];
    Lbd *F[MBL_Q];
} D2Q9;

typedef struct P4 { Lbi xstart, xend, ystart, yend; } P4;
typedef struct P3i { Lbi nc, nf, direction; } P3i;
typedef struct P4d { double xstart, xend, ystart, yend; } P4d;
typedef struct P4d4i { P4d reduced; P4 move; } P4d4i;
typedef struct Int3 { Lbi i, j, k; } Int3;

typedef struct AREA {
    Lbi m, nx, ny;
    double base_x, base_y, tau, rv_tau;
} AREA;

/* Coarse/fine (or same-level) interface line: a line is vertical when
 * xstart == xend, else horizontal. */
typedef struct Inf_Ff_store {
    Lbi areac_n;
    Lbi Nc;
    P4 line_C;
    Lbi areaf_n;
    Lbi Nf;
    P4 line_F;
} Inf_Ff_store;

typedef struct F_k3 { Lbd *F[MBL_Q]; } F_k3;

/* 3-slot time ring of the fine-edge distributions; the slots are walked as
 * an array from &f_t1 and selected through marker[] (host state). */
typedef struct Ff_Store {
    Lbi N;
    Lbi marker[3];
    F_k3 f_t1, f_t, f_t_1;
} Ff_Store;

/* Spline work arrays of one coarse line (Nc nodes). The library's spline
 * kernel keeps its whole solve in LDS; l/m/u are filled by init_tri_A for
 * callers that inspect them, Varibs_M receives the second-derivative columns
 * (column-major, 12 columns of N) and rho..F[8] the gathered line, so the
 * arrays hold the same values as after the reference's gtsv2 call. */
typedef struct FM_spline3 {
    Lbi N;
    Lbd *l, *m, *u;
    Lbd *Varibs_M;
    Lbd *rho, *ux, *uy;
    Lbd *F[MBL_Q];
} FM_spline3;

typedef struct mbl_dim3 { unsigned x, y, z; } mbl_dim3;

typedef struct mbl_context *mbl_handle_t;

/* ---- configuration ---- */

enum {
    MBL_STORE_NATIVE = 0, /* Lbd planes holding f (FP32 build: float, FP64 build: double) */
    MBL_STORE_F32S = 1,   /* float planes holding f - w_k             (FP32 build only) */
    MBL_STORE_F16S = 2,   /* IEEE binary16 planes holding f - w_k      (FP32 build only) */
    MBL_STORE_BF16S = 3,  /* bfloat16 planes holding f - w_k           (FP32 build only) */
    MBL_STORE_F16C = 4,   /* 16-bit planes holding f - w_k in a 1-4-11 format: sign, 4
                           * exponent bits (bias 15, no infinities), 11 fraction bits;
                           * normal range 2^-14 .. 2 - 2^-10, subnormal step 2^-25,
                           * round to nearest even, saturating at the magnitude code
                           * 0x7FFE (2 - 2^-10); 0xFFFF is NaN.
                           * (FP32 build only) */
    MBL_STORE_F16D = 5    /* as MBL_STORE_F16C in a 1-3-12 format: 3 exponent bits (bias
                           * 7), 12 fraction bits; normal range 2^-6 .. 2 - 2^-11
                           * (saturating at 2 - 2^-11),
                           * subnormal step 2^-18. (FP32 build only) */
};

enum {
    MBL_BC_REFERENCE = 0, /* equilibrium (rho0, U0) on every outer domain edge, as the reference */
    MBL_BC_CHANNEL = 1,   /* walls y=0, y=LY-1; parabolic velocity inlet x=0; pressure outlet x=LX-1 */
    MBL_BC_PERIODIC = 2   /* single block, periodic in x and y */
};

enum {
    MBL_INIT_UNIFORM = 0, /* rho0, U0 everywhere (reference init_General) */
    MBL_INIT_REST = 1,    /* rho0, u = 0 */
    MBL_INIT_RANDOM = 2,  /* rho0, (u, v) ~ N(0, sigma^2) from a counter-based generator (seed) */
    MBL_INIT_PROFILE = 3  /* rho0, the parabolic inlet profile u_in(y) on every non-solid node, rest in solids */
};

typedef struct mbl_config {
    /* level tables read as globals by evolution.h:5-34 */
    int LEVEL;
    int mi[MBL_MAX_LEVEL];
    int number_line2;
    /* physical constants (parameters.h:24-52) */
    double rho0;
    double U0[2];
    int LX, LY;
    double R, CX, CY;
    /* boundary, initial state and storage modes */
    int bc_mode;
    int init_mode;
    int storage;
    unsigned long long seed;
    double sigma;
    /* execution */
    int use_graph;      /* capture each refinement-cycle phase as a hipGraph (default 1) */
    int sync_diag;      /* 1: error_L2/getDrag_block synchronise and write at once (reference timing) */
    int flush_interval; /* records buffered on the device before a file flush (default 1000) */
    int write_files;    /* 1: append error_L2.dat / fdrag_t.dat in out_dir (default 1) */
    char out_dir[256];  /* directory of the .dat files ("" = current directory) */
    /* block count read as a global by output/output.h:3 (0: the count of the last evolution()) */
    int number_area;
} mbl_config;

/* Fills a config with the reference defaults (parameters.h, _GEO_2_5 levels). */
void mbl_config_default(mbl_config *cfg);

int mbl_handle_create(mbl_handle_t *h, const mbl_config *cfg);
void mbl_handle_destroy(mbl_handle_t h);
void mbl_handle_make_current(mbl_handle_t h);
mbl_handle_t mbl_handle_current(void);
int mbl_get_last_error(mbl_handle_t h);
const char *mbl_error_string(int code);
int mbl_synchronize(mbl_handle_t h);
int mbl_flush_diagnostics(mbl_handle_t h);
/* Runs the library's work on an external HIP stream (0 = the handle's own). */
int mbl_set_stream(mbl_handle_t h, void *hip_stream);

enum {
    MBL_OK = 0,
    MBL_E_INVALID = -1,  /* invalid argument or geometry */
    MBL_E_HIP = -2,      /* HIP runtime error */
    MBL_E_NOMEM = -3,    /* allocation failed */
    MBL_E_NOT_FINITE = -4, /* non-finite residual or force */
    MBL_E_DIVERGED = -5, /* density <= 0 detected */
    MBL_E_NO_HANDLE = -6,
    MBL_E_UNSUPPORTED = -7
};

/* ---- memory helpers (the C host never includes HIP headers) ---- */
void *mbl_malloc(size_t bytes);
void mbl_free(void *p);
int mbl_memcpy_h2d(void *dst, const void *src, size_t bytes);
int mbl_memcpy_d2h(void *dst, const void *src, size_t bytes);
int mbl_memcpy_d2d(void *dst, const void *src, size_t bytes);
int mbl_memset_d(void *dst, int value, size_t bytes);
/* Bytes of one f/F plane of nx*ny nodes in the handle's storage mode. */
size_t mbl_plane_bytes(mbl_handle_t h, int nx, int ny);
/* Allocates every plane of a block (solid, 25 state planes in the handle's
 * storage) and the two previous-velocity planes (may be NULL). */
int mbl_block_alloc(mbl_handle_t h, const AREA *area, D2Q9 *pp, Lbd **upx, Lbd **upy);
/* (A block's planes may hold at most 4 GiB each: nx*ny*max(element, Lbd) bytes;
 * larger blocks are rejected when stepped, MBL_E_INVALID / MBL_E_UNSUPPORTED.) */
void mbl_block_free(D2Q9 *pp, Lbd *upx, Lbd *upy);
/* Reads plane k of F (which=0) or f (which=1) of a block as f values in double,
 * decoding the storage mode. */
int mbl_read_distribution(mbl_handle_t h, const D2Q9 *pp, int which, int k, size_t n, double *out);
/* Writes f values (double) into plane k of F (which=0) or f (which=1). */
int mbl_write_distribution(mbl_handle_t h, const D2Q9 *pp, int which, int k, size_t n, const double *in);

/* ---- geometry (geo/geometry.h, geo/init_geo.h) ---- */

/* Runtime form of the compile-time tables of geo/geometry.h. AreaDefine may be
 * given directly (integer coarse coordinates); otherwise it is derived from
 * AreaZone as (int)(frac*(L-1)) + move (init_geo.h:91-97). */
typedef struct mbl_geometry {
    int LEVEL;
    int mi[MBL_MAX_LEVEL];
    int number_area, number_line, number_line2;
    int ASP[MBL_MAX_LEVEL + 1];
    int LTCFFCP[MBL_MAX_LEVEL + 1];
    int LS3CFFCP[MBL_MAX_LEVEL + 1];
    int LCCP[MBL_MAX_LEVEL + 1];
    const P4d4i *AreaZone;
    const P4 *AreaDefine;
    const P3i *NeborCF;
    const P3i *NeborCC;
    int LX, LY;
    double Re, U0_2, R;
} mbl_geometry;

/* geo_init (init_geo.h:89-165) + Line_init (:17-86): fills AreaDefine_out,
 * Areas, Line_Inf_C/F (number_line) and Line_Inf_CC (number_line2). Rejects
 * layouts with tau == 1 on either side of a coarse/fine line (singular
 * rescale) or Nf != r*(Nc-1)+1. */
int mbl_geo_init(const mbl_geometry *g, P4 *AreaDefine_out, AREA *Areas,
                 Inf_Ff_store *Line_Inf_C, Inf_Ff_store *Line_Inf_F, Inf_Ff_store *Line_Inf_CC);

/* The layouts of geo/geometry.h as runtime data: "GEO_1_1", "GEO_2_2",
 * "GEO_2_5", "GEO_4_9" with the parameters.h constants (LX=768, LY=320,
 * R=10, Re=10, U0=0.1). Fills *g (pointing at static tables); 0 on success. */
int mbl_geometry_preset(const char *name, mbl_geometry *g);

/* Nested-refinement channel around a cylinder (no reference counterpart):
 * level 0 spans the LX x LY node domain; region[l] (integer coarse
 * coordinates, l = 1..levels-1) is refined by 2 per level, each region
 * strictly inside the previous one with a margin of at least 2. Every level
 * but the finest is split into top/bottom/left/right blocks around the next
 * region, as _GEO_2_5 splits level 0. The tables are written into the caller's
 * buffers (AreaDefine: 4*levels entries, NeborCF: 8*levels, NeborCC:
 * 8*levels); g points at them. */
typedef struct mbl_channel_spec {
    int LX, LY;
    int levels;
    P4 region[MBL_MAX_LEVEL];
    double Re, U, R;
} mbl_channel_spec;
int mbl_geometry_channel(const mbl_channel_spec *s, mbl_geometry *g, P4 *AreaDefine_buf, P3i *NeborCF_buf,
                         P3i *NeborCC_buf);

/* ---- drop-in entry points ---- */

/* lbm/init_block.h:2-51. Coordinates, rho/u (init_mode), upx/upy, forces = 0,
 * node flags (bc_mode), F = f = feq. */
void init_General(mbl_dim3 grid, mbl_dim3 block, AREA area, D2Q9 pp, Lbd *upx, Lbd *upy);

/* multi/init_multi.h:3-28: the three ring slots <- the fine edge's F. */
void init_Fstore(mbl_dim3 grid, mbl_dim3 block, AREA area, D2Q9 pp, Inf_Ff_store line_Inf, Ff_Store line_t3);

/* multi/init_multi.h:30-51: tridiag(0.5, 2, 0.5) with decoupled end rows. */
void init_tri_A(mbl_dim3 grid, mbl_dim3 block, AREA area, D2Q9 pp, Inf_Ff_store line_Inf, FM_spline3 line_S3);

/* evolution.h:2-65: one step of `level` (level 0: one coarse step) with the
 * whole recursive refinement cycle below it. `pps_h` is not used. */
void evolution(Lbi n, Lbi *asp, Lbi *ls3cffcp, Lbi *ltcffcp, Lbi *lccp, Lbi level,
               mbl_dim3 block, mbl_dim3 *grid, AREA *areas, D2Q9 *pps_h, D2Q9 *pps_d,
               Inf_Ff_store *line_Inf_c, Inf_Ff_store *line_Inf_f, Inf_Ff_store *line_Inf_cc,
               Ff_Store *line_t3, FM_spline3 *line_S3, mbl_handle_t handle);

/* lbm/force_block.h:50-75: number of momentum-exchange links of a block. */
Lbi sum_N_drag(mbl_dim3 block, mbl_dim3 grid, AREA area, D2Q9 pp);

/* lbm/force_block.h:78-127: builds the (i, j, k) link list, i-major order,
 * into d_pdrag. Only h_pp.solid is written (the reference copies all 26
 * arrays; the other 25 are not needed). */
void fdrag_store(AREA area, D2Q9 h_pp, Int3 *d_pdrag, D2Q9 d_pp, Lbi N);

/* lbm/force_block.h:180-216: momentum-exchange force on the body from the
 * post-stream distribution of the block's last substep; appends
 * "t\tCd\tCl\t" to fdrag_t.dat. force[0..1] (may be NULL) receives Fx, Fy
 * immediately when sync_diag is set, else at the next flush. */
void getDrag_block(Lbi t, Int3 *d_pdrag, AREA area, D2Q9 d_pp, double *force, Lbi N);

/* lbm/error_L2.h:57-91: relative temporal-change residual
 * sqrt(sum_a sum|u-u_prev|^2/m_a) / sqrt(sum_a sum|u|^2/m_a), u_prev <- u;
 * appends "t\tresidual" to error_L2.dat. The d_*_b arrays receive per-tile
 * partial sums (at most ceil(nx/16)*ceil(ny/16) entries per block, as the
 * reference allocates); h_*_b are not used. */
void error_L2(Lbi t, const Lbi N_areas, mbl_dim3 block, mbl_dim3 *grid, AREA *areas, D2Q9 *pps_d,
              Lbd **upsx, Lbd **upsy, double **h_uvp2_b, double **d_uvp2_b,
              double **h_uv2_b, double **d_uv2_b);

/* ---- additions (no counterpart in the reference) ---- */

/* `nsteps` iterations of the reference driver's loop body (Multi-block-LBM.cu:144-158):
 *     evolution(t, ..., 0, ...); error_L2(t, asp[LEVEL], ..., upsx, upsy, ...);
 *     getDrag_block(t, d_pdrag, areas[last], pps_d[last], force, N_drag);
 * for t = n, ..., n + nsteps - 1, with last = asp[LEVEL] - 1 (the reference's
 * number_area - 1) and no drag when N_drag is 0. Same state, records, files and
 * force output as those calls; the residual's sums may differ from error_L2's in
 * the last bits (another summation order). Where the layout allows it (not
 * periodic, one node per thread, no interface op rewrites the macros of a fluid,
 * inlet or outlet node off a block edge), the last stream-collide of each level also sums
 * error_L2's terms of the nodes it updates (u_prev <- u), and one launch after
 * the cycle takes the other nodes, the drag and the final sums: two launches
 * fewer per step than the three calls, and no pass over the nodes for the
 * residual. Otherwise it makes the three calls. Returns the handle's error. */
int mbl_evolution_step(mbl_handle_t handle, Lbi n, Lbi nsteps, Lbi *asp, Lbi *ls3cffcp, Lbi *ltcffcp, Lbi *lccp,
                       mbl_dim3 block, mbl_dim3 *grid, AREA *areas, D2Q9 *pps_h, D2Q9 *pps_d,
                       Inf_Ff_store *line_Inf_c, Inf_Ff_store *line_Inf_f, Inf_Ff_store *line_Inf_cc,
                       Ff_Store *line_t3, FM_spline3 *line_S3, Lbd **upsx, Lbd **upsy, Int3 *d_pdrag,
                       Lbi N_drag, double *force);
/* 1 when mbl_evolution_step takes the fused path on the handle's current layout, 0 when it
 * makes the three calls; MBL_E_INVALID before the first multi-block step. No reference
 * counterpart. */
int mbl_evolution_step_fused(mbl_handle_t h);

/* Allocation-free variant of error_L2. No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It computes
 * the residual of error_L2 (same kernels and summation order, so the same
 * bits) and updates u_prev <- u, but records nothing in the diagnostics series
 * and writes no file: d_out (caller device memory, 4 doubles) receives
 * {residual, sum|u - u_prev|^2/m, sum|u|^2/m, smallest density}. It calls no
 * allocation routine: the block descriptors and per-tile partial sums live in
 * a fixed 4 MiB device array per GPU (held until the process ends), which
 * bounds a layout to 64 blocks and about 170,000 residual tiles of 64 x 16
 * nodes (about 178M nodes); larger layouts are rejected with MBL_E_INVALID
 * before any work. Synchronous on the handle's stream. Every call on one GPU
 * shares that array: calls that overlap on one GPU (other threads, other
 * handles on other streams) race. */
int mbl_error_L2_noalloc(mbl_handle_t h, Lbi N_areas, const AREA *areas, const D2Q9 *pps_d,
                         Lbd *const *upsx, Lbd *const *upsy, double *d_out);

/* Latest values recorded by error_L2 / getDrag_block, after a flush:
 * residual, Fx, Fy, Cd, Cl. */
int mbl_last_diagnostics(mbl_handle_t h, double out[5]);

/* Copies the recorded diagnostics series (all records since the handle was
 * created or cleared) into caller arrays of capacity `cap`; returns the count.
 * kind 0: (t, residual); kind 1: (t, Cd, Cl). Arrays not needed may be NULL. */
int mbl_diagnostics_series(mbl_handle_t h, int kind, int cap, int *t, double *a, double *b);
void mbl_diagnostics_clear(mbl_handle_t h);

/* Divergence check (no reference counterpart; the reference checks nothing):
 * every error_L2 record also carries the smallest density of any node, and a
 * flush that finds one <= 0 records MBL_E_DIVERGED. Returns the step number t
 * of the first such error_L2 call, or -1 while none has been seen. */
int mbl_divergence_step(mbl_handle_t h);

/* Launches of the scheduled refinement cycle that evolution() issues from
 * `level` (no reference counterpart): evolution() records the cycle of
 * evolution.h:2-65 in its order, derives every dependence between its
 * operations and merges independent operations of one kind into shared
 * launches (stream-collides of adjacent levels, interface ops, splines; the
 * interface ops and one-workgroup splines of one position share a launch).
 * Returns the number of launches and, per kind, the launches that carry it
 * (pointers may be NULL; a shared launch counts for both), or MBL_E_INVALID
 * before the first evolution() call from that level. */
int mbl_cycle_launches(mbl_handle_t h, int level, int *collides, int *interface_ops, int *splines);
/* Positions of the same schedule: launches at one position have no dependence among them.
 * No reference counterpart. */
int mbl_cycle_positions(mbl_handle_t h, int level);
/* on != 0: the launches of one position run concurrently, the first on the handle's stream and
 * the others on side streams forked from it and joined before the next position (parallel
 * branches of the captured graph). Default off: one after another on the handle's stream,
 * measured faster on MI355X (the cross-stream joins cost more than the overlap saves). Drops
 * the captured graphs. No reference counterpart. */
int mbl_set_cycle_fork(mbl_handle_t h, int on);

/* Velocity L2 distance from the analytic plane Poiseuille profile
 * u_a(y) = 4 U0[0] (LY - 1.5 - y)(y - 0.5)/(LY - 2)^2 (the profile left
 * commented at lbm/init_block.h:20), over the FLUID nodes of all blocks with
 * the 1/m block weights of error_L2:
 * sqrt(sum_a sum ((ux-u_a)^2 + uy^2)/m_a) / sqrt(sum_a sum u_a^2/m_a).
 * Synchronous; does not touch u_prev. */
int mbl_poiseuille_l2(mbl_handle_t h, Lbi N_areas, AREA *areas, D2Q9 *pps_d, double *out);

/* Advances one block by `nsteps` fused stream-collide steps on its own
 * (no interface work), swapping pp->F / pp->f after every launch. This is the
 * path of the single-block periodic bandwidth benchmark. A periodic all-fluid
 * block whose width is a multiple of 64 and height a multiple of 16 (8 for
 * FP64 planes) takes three steps per launch (4- and 8-byte planes) or two
 * (16-bit planes), a remainder of two as one paired launch and of one as a single step:
 * pp->F then holds the same state as after single steps, bit for bit, and pp->f
 * the state one launch back (one step back after single steps). */
int mbl_step_block(mbl_handle_t h, const AREA *area, D2Q9 *pp, int nsteps);
/* on == 0: mbl_step_block advances one step per launch on every block; on == 1: two steps per
 * launch where the shape allows, in tiles with a one-node ring; on == 2: two steps per launch in
 * strips of 254 columns with a 3-row window where the shape allows (width >= 256, height a
 * multiple of 64), else tiles; on == 3 (default): three steps per launch in tiles with a two-node
 * ring for 4- and 8-byte planes where the shape allows (as for tiled pairs), a remainder of two
 * steps as one paired launch; 16-bit planes as on == 1. Same bits in every case. No reference
 * counterpart. */
int mbl_set_step_pairs(mbl_handle_t h, int on);
/* Launches mbl_step_block makes for nsteps steps of the block it stepped last (for per-launch
 * timing). No reference counterpart. */
int mbl_step_block_launches(mbl_handle_t h, const AREA *area, int nsteps);

/* Invariants of one block's state (the F planes): out[0] = sum of f over every node and
 * direction (the mass), out[1], out[2] = sum of f e_x, f e_y (the momentum), out[3] = number
 * of non-finite entries; FP64 sums on the device. Synchronous. No reference counterpart: a
 * check of full-size runs whose planes are too large to copy back. */
int mbl_block_moments(mbl_handle_t h, const AREA *area, const D2Q9 *pp, double out[4]);

/* Per-launch time of the fused stream-collide kernel: wraps the next
 * mbl_step_block() launches in HIP events on the handle's stream. */
int mbl_timer_begin(mbl_handle_t h);
int mbl_timer_end(mbl_handle_t h, double *elapsed_ms);

/* Launch tuning of the fused stream-collide kernel: nodes per thread along x
 * (1, 2, 4 or 8; a level whose block widths it does not divide falls back to a
 * divisor) and threads per workgroup (64, 128, 192 or 256). 0 keeps the
 * current value. */
int mbl_set_tuning(mbl_handle_t h, int nodes_per_thread, int threads_per_group);
/* Persistent grid of the fused stream-collide launches: `groups` workgroups
 * each walk one contiguous run of tiles (0: one workgroup per tile). */
int mbl_set_persistent(mbl_handle_t h, int groups);

/* ---- batched independent cases (no reference counterpart) ----
 * Many two-level channel cases of different size advance together: every
 * case's level-0 blocks form one fused stream-collide launch, every case's
 * refined block the other, and every interface line of every case shares the
 * per-level interface launches of evolution(). The whole refinement cycle of
 * all cases is one captured hipGraph per phase. */
typedef struct mbl_case_spec {
    int LX, LY;       /* channel nodes in coarse units (x = 0..LX-1, y = 0..LY-1) */
    P4 region;        /* the ratio-2 refined block, coarse coordinates */
    double R, CX, CY; /* cylinder radius and centre, coarse units */
    double Re, U;     /* Reynolds number on D = 2R; inlet centre-line velocity */
} mbl_case_spec;

typedef struct mbl_batch *mbl_batch_t;

/* Builds, allocates (one device allocation) and initialises every case at
 * rest with channel boundaries. The handle's storage mode is used. */
int mbl_batch_create(mbl_handle_t h, int ncases, const mbl_case_spec *cases, mbl_batch_t *out);
void mbl_batch_destroy(mbl_batch_t b);
/* Advances every case by nsteps coarse steps; the momentum-exchange force on
 * every cylinder is reduced on the device after each step (FP64). */
int mbl_batch_step(mbl_batch_t b, int nsteps);
/* on != 0: each mbl_batch_step call is one persistent launch of one workgroup
 * per case (largest first) that runs all nsteps coarse steps of its case -- the
 * refinement cycle in the reference order and the drag, with workgroup
 * barriers only -- and returns the same bits as the captured-graph path. Lines
 * of up to 303 coarse points; other layouts return MBL_E_UNSUPPORTED from
 * mbl_batch_step. on = 0 (default): the captured-graph path. No reference
 * counterpart (the reference has one case). */
int mbl_batch_set_persistent(mbl_batch_t b, int on);
/* Node updates (all levels, all cases) per coarse step. */
double mbl_batch_node_updates(mbl_batch_t b);
/* Cd and Cl of every case (normalised as getDrag_block) for the last `nt`
 * recorded steps, case-major: cd[c*nt + s]; returns the number of steps
 * available (at most the history capacity, 4096). */
int mbl_batch_forces(mbl_batch_t b, int nt, double *cd, double *cl);
/* error_L2 (lbm/error_L2.h:57-91) of every case on its own: out[c] = the
 * relative change sqrt(sum|u - u_prev|^2/m) / sqrt(sum|u|^2/m) over the case's
 * blocks since the previous call, or since creation / the last
 * mbl_batch_init_from_raster, whose state is the first u_prev; u_prev <- u.
 * Synchronous. No reference counterpart (the reference has one case). */
int mbl_batch_residual(mbl_batch_t b, double *out);
/* Times `trial_steps` steps for every (nodes per thread, threads per group,
 * persistent groups) candidate and keeps the fastest; the trial steps advance
 * the cases. Returns 0 and the choice and its time per step. With the
 * one-launch path on (mbl_batch_set_persistent) the candidates are the nodes
 * per thread and threads per tile of its stream-collide items. */
int mbl_batch_autotune(mbl_batch_t b, int trial_steps, int *vx, int *tpb, int *groups, double *ms_per_step);

/* Streaming-copy bandwidth (bytes read + written per second, GB/s) of a
 * hand-written copy kernel over two fresh buffers of `bytes`, timed with HIP
 * events over `reps` launches: the measured HBM ceiling. No reference
 * counterpart. */
int mbl_bandwidth_probe(mbl_handle_t h, size_t bytes, int reps, double *gbs);

/* The same for the fused kernel's access pattern: 9 planes of `plane_bytes`
 * read and 9 written per launch (one 16-byte element per thread and plane).
 * No reference counterpart. */
int mbl_bandwidth_probe_planes(mbl_handle_t h, size_t plane_bytes, int reps, double *gbs);

/* ---- field output and checkpoints ---- */

/* output/output.h:2-38: writes steady_<n>.dat in out_dir, a Tecplot point file
 * with one zone per block ("zone_level<a>", I = nx, J = ny) of lines
 * "x\ty\tsolid\trho\tU\tV" in the reference's "%g" format. Like the
 * reference it also copies the device arrays into pps_h when pps_h and its
 * pointers are non-NULL (solid, x, y, rho, ux, uy, forcex, forcey, and the f/F
 * planes decoded to Lbd). The block count is cfg.number_area of the current
 * handle, else the count of its last evolution() layout. */
void output(Lbi n, AREA *areas, D2Q9 *pps_h, D2Q9 *pps_d);
/* The same with an explicit handle and block count; returns 0 or an error code. */
int mbl_output(mbl_handle_t h, Lbi n, Lbi number_area, const AREA *areas, D2Q9 *pps_h, const D2Q9 *pps_d);

/* Checkpoint of the whole time-stepping state of a multi-block layout (no
 * reference counterpart): per block the 9 F and 9 f planes in their storage
 * encoding, rho/ux/uy and (when given) the previous velocities of error_L2;
 * per coarse/fine line the three ring slots and the markers; and the caller's
 * step counter. Load writes into the caller's existing allocations after
 * checking block sizes, storage mode and element sizes, and restores the
 * markers into line_t3; a run resumed from a checkpoint reproduces the
 * uninterrupted run bit for bit. upsx/upsy may be NULL. */
int mbl_checkpoint_save(mbl_handle_t h, const char *path, long long step, Lbi number_area, const AREA *areas,
                        const D2Q9 *pps_d, Lbd *const *upsx, Lbd *const *upsy, Lbi number_line,
                        const Ff_Store *line_t3);
int mbl_checkpoint_load(mbl_handle_t h, const char *path, long long *step, Lbi number_area, const AREA *areas,
                        D2Q9 *pps_d, Lbd **upsx, Lbd **upsy, Lbi number_line, Ff_Store *line_t3);

/* ---- learned warm start (no reference counterpart) ----
 * A convolutional encoder-decoder over an H x W raster (H, W multiples of 8):
 * 4 levels of base, 2*base, 4*base, 4*base channels (base a multiple of 32),
 * 3x3 convolutions with ReLU, 2x2 max pooling, nearest upsampling with skip
 * concatenation, and a linear 1x1 head. Activations are BF16 (NHWC), every
 * convolution runs on BF16 MFMA with FP32 accumulation, master weights and
 * Adam state are FP32. Host tensors are [N][H][W][channels] FP32. */
typedef struct mbl_net *mbl_net_t;
int mbl_net_create(mbl_handle_t h, int H, int W, int in_ch, int out_ch, int base, int max_batch,
                   unsigned long long seed, mbl_net_t *out);
void mbl_net_destroy(mbl_net_t n);
long long mbl_net_num_params(mbl_net_t n);
/* get = 1 copies the FP32 parameters to host, get = 0 loads them from host */
int mbl_net_params(mbl_net_t n, float *host, int get);
int mbl_net_forward(mbl_net_t n, const float *in, int N, float *out);
/* mean squared error over N x H x W x out_ch and its gradient (FP32, host, may be NULL) */
int mbl_net_loss_grad(mbl_net_t n, const float *in, const float *target, int N, double *loss, float *grad);
/* Adam (lr, betas 0.9/0.999) over `epochs` shuffled passes of `nsamples` samples held in HBM */
int mbl_net_train(mbl_net_t n, const float *in, const float *target, int nsamples, int batch, int epochs, float lr,
                  unsigned long long seed, double *epoch_loss);

/* The AREA and the device D2Q9 of block q of case c (q = 0..3 the coarse
 * blocks, 4 the refined block); pp->F holds the current state. For reading
 * planes with mbl_read_distribution and for checkpoints. */
int mbl_batch_block(mbl_batch_t b, int c, int q, AREA *area, D2Q9 *pp);
/* (rho, ux, uy) of every case of a batch resampled onto an H x W raster
 * covering its channel [0, LX-1] x [0, LY-1] (bilinear, finest block wins);
 * out is [ncases][H][W][3] FP32 on the host. */
int mbl_batch_raster(mbl_batch_t b, int H, int W, float *out);
/* Re-initialises every case from rasters [ncases][H][W][3] (rho, ux, uy):
 * F = feq(rho, u) + the first-order Chapman-Enskog non-equilibrium part
 * -(tau - 1) dt w_k rho / cs^2 (e_k e_k - cs^2 I) : grad u, with the velocity
 * gradient of the raster; solid nodes at rest. Interface rings are refilled. */
int mbl_batch_init_from_raster(mbl_batch_t b, int H, int W, const float *fields);

#ifdef __cplusplus
}
#endif

#endif /* MBL_H */
