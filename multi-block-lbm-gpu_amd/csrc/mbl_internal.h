// mbl_internal.h -- descriptors shared by the HIP kernels (mbl_kernels.hip) and
// the host engine (mbl_engine.hip). Not part of the public ABI.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/mbl.h"

#ifdef MBL_FP64
typedef double real_t; // compute type of the state update
#else
typedef float real_t;
#endif

namespace mbl {

enum StoreMode { ST_NATIVE = 0, ST_F32S = 1, ST_F16S = 2, ST_BF16S = 3, ST_F16C = 4, ST_F16D = 5 };

// Element size in bytes of one f/F plane entry for a storage mode.
inline size_t store_elem_bytes(int mode)
{
    switch (mode) {
    case ST_NATIVE: return sizeof(Lbd);
    case ST_F32S: return 4;
    default: return 2;
    }
}

// Constants of the run read by the kernels (parameters.h:42-52 twins).
struct Consts {
    double rho0;
    double U0x, U0y;
    double LY;   // channel height in coarse nodes, for the parabolic inlet
    double umax; // inlet centre-line velocity
};

// One block (an AREA with its D2Q9 planes) in device memory. P[0] is the
// caller's F plane set when the table was built, P[1] the caller's f set; the
// block's state lives in P[parity] where parity is the level's bit of the
// mask passed to every launch.
struct BlockDesc {
    Consts bc;            // boundary constants of the block's case
    void *P[2][MBL_Q];
    const int *flag;
    Lbd *rho, *ux, *uy;
    Lbd *upx, *upy;       // previous velocity (error_L2), may be null
    int nx, ny;
    int level;
    int m;
    int tiles_x;          // stream-collide tiles along x
    uint32_t tx_rcp;      // 0xFFFFFFFF / tiles_x (tile row and column by a multiply-high)
    int tw;               // tile width in lanes (64, or 32: a wave then covers two rows)
    int trows;            // rows per tile (0: the launch's tile_rows)
    int tile_begin;       // first stream-collide tile of the block inside its launch
    int rtile_begin;      // first residual tile of the block inside the residual launch
    int rtiles_x;
    double base_y, dx;
    double omega;         // 1/tau
    double inv_m;         // residual weight 1/m
};

// Interface line operations (multi/FH/assit_multi.h). An op moves data along
// one line between two blocks or between a block and a ring buffer.
enum OpType { OP_TEMPORAL = 0, OP_CC = 1, OP_CF = 2, OP_FC = 3, OP_SPLINE = 4 };

// Ring-slot markers (Ff_Store.marker) of every group of lines, a group being
// the lines whose coarse block is on one level: they rotate together, at that
// level's rate (evolution.h:41-46). Passed by value to the line-op kernels.
struct MarkerSet {
    int m[MBL_MAX_LEVEL][3];
};

struct LineOp {
    int type;
    int mgroup;       // TEMPORAL, CF: marker group (level of the coarse block of the line)
    int n;            // entries (fine points for TEMPORAL/CF/SPLINE, Nc for CC/FC)
    int src_block;    // table index of the block read (CC, FC, SPLINE coarse block)
    int dst_block;    // table index of the block written
    const int *src_idx; // node index in src block per entry (CC, FC; SPLINE: coarse line nodes, Nc of them)
    const int *dst_idx; // node index in dst block per entry
    const uint8_t *mask; // per entry: bit0 write macros (rho, u), bit1 write F
    Lbd *ring[3][MBL_Q]; // ring slots (TEMPORAL, CF, SPLINE), walk order &f_t1
    int nc;             // SPLINE: coarse points
    int ratio;          // SPLINE/FC: areaf.m / areac.m (integer division, as the reference)
    double para;        // FC: r*(tau_c-1)/(tau_f-1); SPLINE: (tau_f-1)/(r*(tau_c-1))
    // SPLINE work arrays: caller's FM_spline3 (l, m, u read; Varibs_M and the
    // gathered columns written) and the library's FP64 second-derivative columns.
    const Lbd *tl, *tm, *tu;
    Lbd *varibs;
    Lbd *gcol[12];
    double *M64;
    // SPLINE: the elimination factors of parallel cyclic reduction on tridiag(tl, tm, tu),
    // which depend on the matrix only (built once per layout from the caller's l, m, u):
    // level q's factors of row i at pcr[(2q) nc + i] (alpha, the row i - 2^q) and
    // pcr[(2q + 1) nc + i] (gamma, the row i + 2^q), then the reduced diagonal at
    // pcr[2 pcr_levels nc + i]. pcr_levels: the levels after which every row is decoupled.
    const double *pcr;
    int pcr_levels;
    // FC: entries whose fine node a c_f op of the same batch writes (line ends on the fine
    // block's ghost edges), read without the node so that the batch is one launch: type 1, the
    // value that c_f op wrote before (its ring slot, entry red_ce); type 2, the value before
    // that c_f op writes it (the ghost as the last collide carried it over: the other plane set)
    int red_n;
    int red_e[2], red_t[2], red_ce[2];
    int red_mg[2];                    // type 1: the c_f op's marker group
    const Lbd *red_ring[2][3][MBL_Q]; // type 1: the c_f op's ring slots
    // SPLINE: per coarse point, the (block, node) its column values are read from in place of
    // src_idx's node in src_block (a same-level copy's source), block -1: none; null: none at all
    const int2 *gred;
    // SPLINE: 1 + the index (in its launch's op table) of a temporal op on the same line that
    // the spline's workgroups do after storing the line's points (0: none). That TEMPORAL op:
    // its markers are the launch's for its group turned trot times, its block's plane set is
    // the launch's parity bit xor tflip, and it interpolates at txm.
    int fuse_t;
    int trot, tflip;
    double txm;
};

// Most PCR levels the one-workgroup spline keeps in registers (rows decouple after 6 levels for
// the spline matrix of any length; 10 levels solve any system of up to 1024 rows exactly).
constexpr int PCR_MAX_LEVELS = 10;

// The PCR factor table of tridiag(l, m, u) of N rows (layout of LineOp::pcr): the levels are run
// until every row's couplings are below 2^-60 of its diagonal or the system is solved.
// Returns the number of levels; `tab` receives (2 levels + 1) N values.
int pcr_factor_table(const double *l, const double *m, const double *u, int N, std::vector<double> &tab);

struct DragArgs {
    void *Fold[MBL_Q]; // planes holding the state one substep older (the caller's f)
    const Int3 *links;
    int n;
    int nx;
    double *out;       // [0] Fx, [1] Fy
};

// The diagnostics launches of a fused coarse step (launch_step_diag: the nodes the collides do
// not sum and the drag, then the final sums).
struct StepDiagArgs {
    const BlockDesc *descs;
    // every node the stream-collides do not sum: (block, node, block, node whose macros it holds
    // when the step's diagnostics run; the node itself, or the source of the copy that writes it
    // in the same launch)
    const int4 *list;
    int nlist;
    int list_wgs;      // workgroups over the list (256 entries each, at least one)
    int ntiles;        // stream-collide tiles of all levels: their sums at rparts[0, 3 ntiles)
    double *rparts;    // the tiles' 3 ntiles partial sums
    // the step's record slot s = *slot_rd (0 <= s < cap; *slot_wr <- s + 1 for the next step):
    // the list workgroups' 3 list_wgs partial sums go to lbase + 3 lstride s (launch_residual_steps
    // sums them), the drag to drag_base + 2 s
    const int *slot_rd;
    int *slot_wr;
    int cap;
    double *lbase;
    size_t lstride;
    double *drag_base;
    DragArgs drag;     // drag.n == 0: no drag
};

// One block of a batched case for the raster resampler and the raster
// initialiser: node (i, j) sits at coarse coordinates (bx + i/m, by + j/m).
struct RasterBlock {
    void *P[2][MBL_Q]; // both plane sets (the initialiser writes both)
    const int *flag;
    Lbd *rho, *ux, *uy;
    double bx, by;
    double tau;
    int m, nx, ny;
    int cas;           // case index
};

// A case's blocks, finest first, and its channel extent in coarse nodes.
struct RasterCase {
    int blk[5];
    int nblk;
    double LX, LY;
};

// ---- batched cases in one persistent launch ----
// One item of a case's coarse step, in the reference order: a stream-collide over the case's
// tiles of one level, one sub-batch of its interface ops, its spline lines, or its drag.
enum ItemKind { ITEM_COLLIDE = 0, ITEM_OPS = 1, ITEM_SPLINE = 2, ITEM_DRAG = 3 };
struct CaseItem {
    int kind;
    int a, b;      // COLLIDE: first tile, tiles; OPS/SPLINE: first op, ops; DRAG: drag table
    int max_n;     // OPS: entries of the longest op
    uint32_t pm;   // level parities in force
    MarkerSet ms;  // ring markers in force
    double xm;     // OPS: temporal interpolation point
};

// Threads per workgroup (one case each; 8 waves, up to 256 VGPRs, so the inlined items do not
// spill); the splines of a case run in one workgroup, so lines of up to CASE_TPB coarse points.
constexpr int CASE_TPB = 512;

struct CaseArgs {
    const CaseItem *items; // [period][ncases][per_case]
    int per_case;          // items of one case per coarse step
    int period;            // coarse steps in the table (it repeats)
    int step0;             // table step of the launch's first step
    int ncases;
    int nsteps;
    const int *case_of_wg; // workgroup -> case
    const BlockDesc *descs;
    const int *tile_blk;
    int ndesc;
    int rows;              // stream-collide tile rows
    int tpb;               // threads per stream-collide tile
    const LineOp *ops;
    const DragArgs *drag[2];
    double *hist;          // Fx, Fy of case c after step s -> hist[2 (ncases ((head0 + s) % cap) + c)]
    int head0, cap;
};

size_t cases_lds(int max_nc);
// 1 when spline lines of max_nc coarse points run inside the case kernel
int cases_spline_ok(int max_nc);
int launch_cases(void *stream, int mode, int vx, const CaseArgs &a, size_t lds);

// ---- launchers (mbl_kernels.hip), all asynchronous on `stream` ----
// tile_rows rows per tile (at least tpb / 64; the waves take them in turn);
// d_tile_blk[t] = descriptor of tile t, or null to search tile_begin.
// tile_base: the launch covers tiles [tile_base, tile_base + total_tiles) of the numbering of
// the descriptors' tile_begin (and of d_tile_blk).
// rmask != 0 (flagged, macros, one node per thread): the tiles of the levels in rmask also add
// error_L2's terms of their fluid, inlet and outlet nodes off the block edges (u_prev <- u) and
// write the tile's weighted sums to rparts[3 t .. 3 t + 2], t the tile's number across all levels.
// Two time steps of a periodic all-fluid nx x ny block in one launch: S (state at step n) ->
// D (state at step n + 2); periodic2_fits: whether the block's shape allows it.
struct Periodic2Args {
    const void *S[MBL_Q];
    void *D[MBL_Q];
    int nx, ny, tiles_x;
    int form; // 1: tiles with a one-node ring; 2: strips with a 3-row window (where the shape allows)
    double omega;
};
int periodic2_fits(int mode, int nx, int ny, int form);
int launch_periodic2(void *stream, int mode, const Periodic2Args &a);
// three steps per launch (same arguments; D receives step n + 3)
int periodic3_fits(int mode, int nx, int ny);
int launch_periodic3(void *stream, int mode, const Periodic2Args &a);
int launch_stream_collide(void *stream, int mode, int vx, int tpb, bool periodic, bool flags, bool macros,
                          const BlockDesc *d_descs, const int *d_tile_blk, int ndesc, int total_tiles, int tile_rows,
                          uint32_t parity_mask, int groups, int tile_base, uint32_t rmask = 0,
                          double *rparts = nullptr);
int launch_init_general(void *stream, int mode, const BlockDesc &d, const AREA &area, Lbd *x, Lbd *y,
                        Lbd *forcex, Lbd *forcey, int bc_mode, int init_mode, unsigned long long seed,
                        double sigma, double LX, double LY, double R, double CX, double CY, const Consts &c);
int launch_line_ops(void *stream, int mode, const LineOp *d_ops, int nops, int max_n,
                    const BlockDesc *d_blocks, uint32_t parity_mask, const MarkerSet &ms, double xm);
// 0: lines of up to max_nc coarse points take the one-workgroup spline kernel; 1: the two-kernel path
int launch_spline_path(int max_nc);
// each op writes ring slot ms.m[op.mgroup][0]
int launch_spline(void *stream, int mode, const LineOp *d_ops, int nops, int max_nc, int max_nf,
                  const BlockDesc *d_blocks, uint32_t parity_mask, const MarkerSet &ms);
// nspl one-workgroup spline ops (max_nc coarse, max_nf fine points) followed by nops interface ops
// (up to max_n entries) in one launch
// A flagged stream-collide (one node per thread) over tiles [tile_base, tile_base + tiles) and
// the entries of nops interface ops in one launch (launch_collide_ops sets opx).
struct CollideOpsArgs {
    const BlockDesc *descs;
    const int *tile_blk;
    int ndesc, rows;
    uint32_t pm;
    int tiles, tile_base;
    uint32_t rmask;
    double *rparts;
    const LineOp *ops;
    int nops, opx;
    MarkerSet ms;
    double xm;
};
int launch_collide_ops(void *stream, int mode, int tpb, const CollideOpsArgs &a, int max_n);
int launch_iface_mixed(void *stream, int mode, const LineOp *d_ops, int nspl, int nops, int max_nc, int max_nf,
                       int max_n, const BlockDesc *d_blocks, uint32_t parity_mask, const MarkerSet &ms, double xm);
int launch_ring_fill(void *stream, int mode, const LineOp *d_ops, int nops, int max_n,
                     const BlockDesc *d_blocks, uint32_t parity_mask);
int launch_tri_A(void *stream, Lbd *l, Lbd *m, Lbd *u, int N);
// d_tblk[t]: block of residual tile t (null: found from the descriptors' rtile_begin).
int launch_residual(void *stream, const BlockDesc *d_descs, const int *d_tblk, int ndesc, int total_rtiles,
                    double *partials, int mode, double umax, double LY, double *d_out);
// error_L2 per case of a batch: the residual tiles (u_prev <- u), then case c's tiles
// [d_case_tiles[c], d_case_tiles[c+1]) summed into d_out[4c..4c+3].
int launch_residual_cases(void *stream, const BlockDesc *d_descs, const int *d_tblk, int ndesc, int total_rtiles,
                          double *partials, int ncases, const int *d_case_tiles, double *d_out);
int launch_drag(void *stream, int mode, const DragArgs &a);
int launch_step_diag(void *stream, int mode, const StepDiagArgs &a);
// The interface ops of the cycle's last position and the step's diagnostics in one launch
// (the diagnostics read nothing the ops write: launch_plans checks that when it plans it).
int launch_ops_diag(void *stream, int mode, const LineOp *d_ops, int nops, int max_n, const BlockDesc *d_blocks,
                    uint32_t parity_mask, const MarkerSet &ms, double xm, const StepDiagArgs &a);
// error_L2 records of nsteps steps: step b sums partials[3 (b stride) + 3 q], q < nparts, in
// order into out[4 b .. 4 b + 3] (residual, s1, s2, smallest density), as k_residual_final does
int launch_residual_steps(void *stream, int nsteps, int nparts, size_t stride, const double *partials, double *out);
int launch_drag_batch(void *stream, int mode, const DragArgs *d_args, int ncases, double *out);
int launch_stream_copy(void *stream, const void *src, void *dst, size_t bytes);
// 9 planes of plane_bytes (consecutive in src and in dst) copied plane by plane
int launch_plane_copy(void *stream, const void *src, void *dst, size_t plane_bytes);
int launch_raster(void *stream, const RasterBlock *d_blocks, const RasterCase *d_cases, int ncases, int H, int W,
                  float *d_out);
int launch_init_raster(void *stream, int mode, const RasterBlock *d_blocks, int nblocks, size_t max_nodes,
                       const RasterCase *d_cases, int H, int W, const float *d_fields);
int launch_decode(void *stream, int mode, const void *plane, int k, size_t n, double *out);
// d_out[0..3]: sum of f, f e_x, f e_y over the 9 planes d_planes[0..8] of n entries, and the
// number of non-finite entries; nparts workgroups write partials[4 nparts] first.
int launch_moments(void *stream, int mode, const void *const *d_planes, size_t n, int nparts, double *partials,
                   double *d_out);
int launch_encode(void *stream, int mode, void *plane, int k, size_t n, const double *in);

} // namespace mbl
