// mbl_engine.hip -- host side of libmbl_hip: the handle, the device tables
// built from the caller's reference structures, the refinement-cycle driver
// with per-phase hipGraph capture, device-side diagnostics and every
// extern "C" entry point of include/mbl.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "mbl_internal.h"

using namespace mbl;

namespace {

// Rows per stream-collide tile of a multi-block layout beyond the waves of a
// workgroup (a loop over rows lowers occupancy more than it saves: measured).
constexpr int MBL_TILE_ROWS = 1;

constexpr int EXh[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int EYh[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

struct Sub {
    int off = 0, count = 0, max_n = 0, max_nc = 0;
    bool spline = false;
};
typedef std::vector<Sub> Batch;

// What one interface op reads and writes (see build_layout's op_access).
struct OpAccess {
    std::vector<uint64_t> rd, wr; // node units (sorted)
    uint32_t epochs = 0;          // levels whose collide epoch the op belongs to
    int ring_line = -1;           // line whose ring the op writes (spline) or reads (temporal, c_f)
    bool ring_write = false;
};

// One operation of the refinement cycle as evolution.h issues it, recorded with the level
// parities and ring markers in force at that point.
struct CycleItem {
    int kind;       // 0 stream-collide of `level`, 1 interface ops, 2 spline
    int level;
    int off, count; // interface / spline ops: range of the layout's op table
    double xm;      // temporal interpolation point of the interface ops
    bool temporal;  // the range holds temporal ops (xm matters)
    uint32_t pm;
    MarkerSet ms;
};

// A merged launch of the scheduled cycle: items of one kind with no dependence among them
// (kind 3: one-workgroup splines and interface ops of one position in one launch).
struct CycleLaunch {
    int kind; // 0 stream-collide, 1 interface ops, 2 splines, 3 splines + interface ops, 4 collide + interface ops
    int pos = 0;             // position in the schedule (launches at one position are independent)
    std::vector<int> items;  // indices into the recorded item sequence
    int lo = 0, hi = 0;      // collide: levels lo..hi (adjacent, same nodes per thread)
    uint32_t rmask = 0;      // collide: levels whose last collide of the cycle this launch holds
    int op_off = 0, op_count = 0, max_n = 0, max_nc = 0; // interface/spline: range of the program's op table
    int nspl = 0, max_nf = 0; // kind 3: its first nspl ops are splines (of up to max_nf fine points)
    std::vector<int> fused;   // kinds 2, 3: temporal items done by its spline workgroups
};

struct CycleProgram {
    bool built = false;
    int nitems = 0;
    std::vector<int> fused_into; // per item: the spline item whose launch does it, or -1
    std::vector<CycleLaunch> launches;
    LineOp *dops = nullptr; // ops of every interface/spline launch, each launch contiguous
    // the fused step's diagnostics in the cycle's last launch (interface ops only): its list,
    // where nodes whose macros that launch copies are read at the copies' sources, and the
    // blocks whose planes that launch writes (the drag's plane set is checked against them)
    bool dmerge = false;
    int4 *dlist_m = nullptr;
    std::vector<int> last_fw_blocks;
};

// Device tables for one multi-block layout (areas, planes, lines) as passed
// to evolution(). Rebuilt when the caller's pointers or geometry change.
struct Layout {
    uint64_t sig = 0;
    int LEVEL = 0, nblocks = 0, number_line = 0, number_line2 = 0;
    int mi[MBL_MAX_LEVEL] = {0};
    int ASP[MBL_MAX_LEVEL + 1] = {0}, LTC[MBL_MAX_LEVEL + 1] = {0}, LS3[MBL_MAX_LEVEL + 1] = {0},
        LCC[MBL_MAX_LEVEL + 1] = {0};
    std::vector<BlockDesc> hb;
    BlockDesc *db = nullptr;
    int tiles[MBL_MAX_LEVEL] = {0};
    int tile_off[MBL_MAX_LEVEL] = {0}; // first entry of each level in dtile
    int *dtile = nullptr;              // per tile: index of its block in db (tiles numbered across levels)
    int rows = 0;                      // rows per stream-collide tile
    int vx[MBL_MAX_LEVEL] = {0};
    bool periodic = false;
    std::vector<LineOp> hops;
    LineOp *dops = nullptr;
    int *didx = nullptr;
    uint8_t *dmask = nullptr;
    double *dM64 = nullptr;
    double *dpcr = nullptr; // PCR factor tables of the spline lines (one per distinct matrix)
    // fused coarse step (mbl_evolution_step): residual sums from the last collide of each level
    // and one diagnostics launch (the nodes no collide writes the macros of, the drag)
    bool step_ok = false;   // non-periodic, one node per thread, macros of collide nodes not rewritten
    int total_tiles = 0;    // stream-collide tiles of all levels
    int4 *dlist = nullptr;  // (block, node, block, node) of the nodes whose macros no collide writes
    std::vector<int4> hlist; // the same on the host
    // per op of hops: its entries' source and destination nodes and write masks
    std::vector<std::vector<int>> hsrc, hdst;
    std::vector<std::vector<uint8_t>> hmsk;
    int nlist = 0;
    double *dres = nullptr; // 3 (total_tiles + list workgroups) partial sums
    std::vector<Lbd *> up_x, up_y; // u_prev arrays installed in the descriptors
    Batch bT[MBL_MAX_LEVEL], bCC[MBL_MAX_LEVEL], bS[MBL_MAX_LEVEL], bCF[MBL_MAX_LEVEL];
    std::vector<std::vector<int>> batch_lines_T, batch_lines_S, batch_lines_CF; // line ids per level
    std::map<uint64_t, hipGraphExec_t> graphs;
    // graphs of several consecutive fused steps: (phase key of the first, number of steps)
    std::map<std::pair<uint64_t, int>, hipGraphExec_t> chunks;
    uint64_t diag_sig = 0; // the diagnostics arguments the fused graphs were captured with
    std::vector<int> level_of;  // level of each block
    std::vector<OpAccess> acc;  // per op of dops: what it reads and writes (cycle scheduling)
    CycleProgram prog[MBL_MAX_LEVEL]; // scheduled launches of evolution() from each level
    void release()
    {
        for (auto &p : prog) {
            if (p.dops) hipFree(p.dops);
            if (p.dlist_m) hipFree(p.dlist_m);
            p = CycleProgram();
        }
        for (auto &g : graphs) hipGraphExecDestroy(g.second);
        graphs.clear();
        for (auto &g : chunks) hipGraphExecDestroy(g.second);
        chunks.clear();
        if (db) hipFree(db);
        if (dops) hipFree(dops);
        if (didx) hipFree(didx);
        if (dmask) hipFree(dmask);
        if (dM64) hipFree(dM64);
        if (dpcr) hipFree(dpcr);
        if (dlist) hipFree(dlist);
        if (dres) hipFree(dres);
        if (dtile) hipFree(dtile);
        dtile = nullptr;
        db = nullptr;
        dops = nullptr;
        didx = nullptr;
        dmask = nullptr;
        dM64 = nullptr;
        dpcr = nullptr;
        dlist = nullptr;
        dres = nullptr;
    }
};

struct Single {
    uint64_t sig = 0;
    BlockDesc hd{};
    BlockDesc *dd = nullptr;
    int tiles = 0, vx = 1, rows = 4, tpb = 256;
    bool flags = true, periodic = false;
};

struct Residual {
    uint64_t sig = 0;
    std::vector<BlockDesc> hd;
    BlockDesc *dd = nullptr;
    int total = 0;
    double *partials = nullptr;
    int *tblk = nullptr; // block of each residual tile
};

// doubles per error_L2 record: residual, sum |u - u_prev|^2 / m, sum |u|^2 / m, min density
constexpr int L2_REC = 4;

struct PendingDrag {
    int t;
    int m;
};

} // namespace

struct mbl_context {
    mbl_config cfg;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int device = 0;
    int last_error = 0;
    int vx = 4, tpb = 256;
    bool vx_set = false;  // vx chosen by mbl_set_tuning or the autotuner, not by default
    int groups = 0; // persistent grid size of the stream-collide launches (0: one workgroup per tile)
    Layout *layout = nullptr;
    uint64_t layout_gen = 0; // advances each time a layout is built
    Single single;
    Residual resid;
    // diagnostics: device record slots and pinned host mirrors
    int cap = 0;
    double *d_l2 = nullptr, *d_drag = nullptr;
    double *h_l2 = nullptr, *h_drag = nullptr;
    std::vector<int> pend_l2;
    std::vector<PendingDrag> pend_drag;
    // records of mbl_evolution_step whose final sum waits for the flush: (slot, partial sums);
    // slot s keeps its partials at d_lparts + 3 lparts_stride s
    std::vector<std::pair<int, int>> pend_final;
    double *d_lparts = nullptr;
    int lparts_stride = 0;
    // the two record-slot counters of the fused step's diagnostics launch (it reads one and
    // writes the next slot to the other, by the coarse level's parity) and their values as
    // enqueued (-1: unknown)
    int *d_slots = nullptr;
    int slot_host[2] = {-1, -1};
    std::vector<int> ser_l2_t, ser_drag_t;
    std::vector<double> ser_l2, ser_cd, ser_cl;
    double last[5] = {0, 0, 0, 0, 0};
    int diverged_step = -1; // first step whose smallest density was <= 0
    void *scratch = nullptr;  // device address of this device's allocation-free scratch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timing = false;
    // side streams of a cycle's concurrent launches (several launches at one position of the
    // schedule) and the fork/join events; created on first use, default priority
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[3] = {nullptr, nullptr, nullptr};
    bool fork = false; // launches at one position run concurrently (mbl_set_cycle_fork)
    int step_pairs = 3; // mbl_step_block: periodic all-fluid blocks two or three steps per launch (its form)
};

static mbl_context *g_current = nullptr;

// ---------------------------------------------------------------------------
// Device scratch of the allocation-free entry point (mbl_error_L2_noalloc): a
// file-scope array of fixed size, one copy per GPU for the life of the process.
// In host code `scratch` names a host-side stand-in, not the device array: its
// device address is looked up once per device (hipGetSymbolAddress, at handle
// creation with that device current) and only that address is used.
// ---------------------------------------------------------------------------
constexpr size_t MBL_NOALLOC_SCRATCH_BYTES = (size_t)4 << 20;
constexpr int MBL_NOALLOC_MAX_BLOCKS = 64;
static __device__ __align__(256) unsigned char scratch[MBL_NOALLOC_SCRATCH_BYTES];
constexpr int MBL_MAX_DEVICES = 64;
static void *g_scratch_ptr[MBL_MAX_DEVICES];
static std::mutex g_scratch_mu;

// The current device's scratch address (looked up on the first call per device).
static int scratch_ptr(int device, void **out)
{
    if (device < 0 || device >= MBL_MAX_DEVICES) return MBL_E_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return MBL_E_HIP;
        }
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

static int set_err(mbl_context *h, int code)
{
    if (h && code && !h->last_error) h->last_error = code;
    return code;
}

#define HIPCHK(h, call)                                                                                                \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) {                                                                                        \
            fprintf(stderr, "mbl: %s failed: %s\n", #call, hipGetErrorString(e_));                                     \
            return set_err(h, MBL_E_HIP);                                                                              \
        }                                                                                                              \
    } while (0)

static inline uint64_t mix64(uint64_t h, uint64_t v)
{
    h ^= v + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
    return h;
}

static inline Consts make_consts(const mbl_config &c)
{
    Consts k;
    k.rho0 = c.rho0;
    k.U0x = c.U0[0];
    k.U0y = c.U0[1];
    k.LY = (double)c.LY;
    k.umax = c.U0[0];
    return k;
}

static int storage_mode(const mbl_config &c)
{
#ifdef MBL_FP64
    return c.storage == MBL_STORE_NATIVE ? ST_NATIVE : -1;
#else
    return (c.storage >= 0 && c.storage <= 5) ? c.storage : -1;
#endif
}

// ---------------------------------------------------------------------------
// configuration and handle
// ---------------------------------------------------------------------------
extern "C" void mbl_config_default(mbl_config *c)
{
    memset(c, 0, sizeof(*c));
    c->LEVEL = 2; // _GEO_2_5 (geo/geometry.h:57-58)
    c->mi[0] = 1;
    c->mi[1] = 2;
    c->number_line2 = 8;
    c->rho0 = 1.0; // parameters.h:27-43
    c->U0[0] = 0.1;
    c->U0[1] = 0.0;
    c->LX = 768;
    c->LY = 320;
    c->R = 10;
    c->CX = 160;
    c->CY = 160;
    c->bc_mode = MBL_BC_REFERENCE;
    c->init_mode = MBL_INIT_UNIFORM;
    c->storage = MBL_STORE_NATIVE;
    c->seed = 0x5EED2024ULL;
    c->sigma = 0.01;
    c->use_graph = 1;
    c->sync_diag = 0;
    c->flush_interval = 1000;
    c->write_files = 1;
    c->out_dir[0] = 0;
}

extern "C" int mbl_handle_create(mbl_handle_t *out, const mbl_config *cfg)
{
    if (!out || !cfg) return MBL_E_INVALID;
    *out = nullptr;
    if (storage_mode(*cfg) < 0) return MBL_E_UNSUPPORTED;
    if (cfg->LEVEL < 1 || cfg->LEVEL > MBL_MAX_LEVEL) return MBL_E_INVALID;
    mbl_context *h = new mbl_context();
    h->cfg = *cfg;
    if (h->cfg.flush_interval <= 0) h->cfg.flush_interval = 1000;
    h->cap = h->cfg.flush_interval;
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return MBL_E_HIP;
    }
    if (hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return MBL_E_HIP;
    }
    h->stream = h->own_stream;
    // the allocation-free entry point's scratch address on this device
    if (scratch_ptr(h->device, &h->scratch) != 0) {
        hipStreamDestroy(h->own_stream);
        delete h;
        return MBL_E_HIP;
    }
    // default launch shape of the fused kernel per plane storage, from launch-shape sweeps on
    // MI355X (16384^2 periodic block): 4 nodes/thread for 4-byte planes, 1 for 2-byte planes,
    // 128 threads per workgroup. Flagged (non-periodic) blocks default to 1 node/thread
    // whatever the storage (measured equal on 2049-wide blocks, faster on narrow ones).
    const int st = storage_mode(h->cfg);
    h->vx = store_elem_bytes(st) == 2 ? 1 : 4;
    h->tpb = 128;
    if (hipMalloc(&h->d_slots, 2 * sizeof(int)) != hipSuccess ||
        hipMalloc(&h->d_l2, sizeof(double) * L2_REC * h->cap) != hipSuccess ||
        hipMalloc(&h->d_drag, sizeof(double) * 2 * h->cap) != hipSuccess ||
        hipHostMalloc(&h->h_l2, sizeof(double) * L2_REC * h->cap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&h->h_drag, sizeof(double) * 2 * h->cap, hipHostMallocDefault) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        mbl_handle_destroy(h);
        return MBL_E_NOMEM;
    }
    g_current = h;
    *out = h;
    return 0;
}

static void release_single(mbl_context *h)
{
    if (h->single.dd) hipFree(h->single.dd);
    h->single.dd = nullptr;
    h->single.sig = 0;
}

static void release_resid(mbl_context *h)
{
    if (h->resid.dd) hipFree(h->resid.dd);
    if (h->resid.partials) hipFree(h->resid.partials);
    if (h->resid.tblk) hipFree(h->resid.tblk);
    h->resid.dd = nullptr;
    h->resid.partials = nullptr;
    h->resid.tblk = nullptr;
    h->resid.sig = 0;
}

extern "C" void mbl_handle_destroy(mbl_handle_t h)
{
    if (!h) return;
    if (h->stream) {
        mbl_flush_diagnostics(h);
        hipStreamSynchronize(h->stream);
    }
    if (h->layout) {
        h->layout->release();
        delete h->layout;
    }
    release_single(h);
    release_resid(h);
    if (h->d_l2) hipFree(h->d_l2);
    if (h->d_drag) hipFree(h->d_drag);
    if (h->d_lparts) hipFree(h->d_lparts);
    if (h->d_slots) hipFree(h->d_slots);
    if (h->h_l2) hipHostFree(h->h_l2);
    if (h->h_drag) hipHostFree(h->h_drag);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    for (int i = 0; i < 3; i++) {
        if (h->side[i]) hipStreamDestroy(h->side[i]);
        if (h->ev_join[i]) hipEventDestroy(h->ev_join[i]);
    }
    if (h->ev_fork) hipEventDestroy(h->ev_fork);
    if (h->own_stream) hipStreamDestroy(h->own_stream);
    if (g_current == h) g_current = nullptr;
    delete h;
}

extern "C" void mbl_handle_make_current(mbl_handle_t h) { g_current = h; }
extern "C" mbl_handle_t mbl_handle_current(void) { return g_current; }
extern "C" int mbl_get_last_error(mbl_handle_t h) { return h ? h->last_error : MBL_E_NO_HANDLE; }

extern "C" const char *mbl_error_string(int code)
{
    switch (code) {
    case MBL_OK: return "ok";
    case MBL_E_INVALID: return "invalid argument or geometry";
    case MBL_E_HIP: return "HIP runtime error";
    case MBL_E_NOMEM: return "allocation failed";
    case MBL_E_NOT_FINITE: return "non-finite residual or force";
    case MBL_E_DIVERGED: return "density <= 0";
    case MBL_E_NO_HANDLE: return "no current handle";
    case MBL_E_UNSUPPORTED: return "unsupported mode or size";
    default: return "unknown error";
    }
}

extern "C" int mbl_synchronize(mbl_handle_t h)
{
    if (!h) return MBL_E_NO_HANDLE;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return h->last_error;
}

extern "C" int mbl_set_stream(mbl_handle_t h, void *s)
{
    if (!h) return MBL_E_NO_HANDLE;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->stream = s ? (hipStream_t)s : h->own_stream;
    if (h->layout)
        for (auto &g : h->layout->graphs) hipGraphExecDestroy(g.second), (void)0;
    if (h->layout) h->layout->graphs.clear();
    return 0;
}

static void drop_layouts(mbl_context *h)
{
    hipStreamSynchronize(h->stream);
    release_single(h);
    if (h->layout) {
        h->layout->release();
        delete h->layout;
        h->layout = nullptr;
    }
}

extern "C" int mbl_set_persistent(mbl_handle_t h, int groups)
{
    if (!h) return MBL_E_NO_HANDLE;
    if (groups < 0) return MBL_E_INVALID;
    h->groups = groups;
    drop_layouts(h); // captured graphs hold the old launch shape
    return 0;
}

extern "C" int mbl_set_tuning(mbl_handle_t h, int vx, int tpb)
{
    if (!h) return MBL_E_NO_HANDLE;
    if (vx && vx != 1 && vx != 2 && vx != 4 && vx != 8) return MBL_E_INVALID;
    if (tpb && (tpb % 64 || tpb < 64 || tpb > 256)) return MBL_E_INVALID;
    if (vx) {
        h->vx = vx;
        h->vx_set = true;
    }
    if (tpb) h->tpb = tpb;
    drop_layouts(h);
    return 0;
}

// ---------------------------------------------------------------------------
// memory helpers
// ---------------------------------------------------------------------------
extern "C" void *mbl_malloc(size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}
extern "C" void mbl_free(void *p)
{
    if (p) hipFree(p);
}
// The copy helpers order themselves after the current handle's queued work
// (its stream is non-blocking, so a plain hipMemcpy would not wait for it).
static int wait_current()
{
    if (g_current && hipStreamSynchronize(g_current->stream) != hipSuccess) return MBL_E_HIP;
    return 0;
}
extern "C" int mbl_memcpy_h2d(void *d, const void *s, size_t n)
{
    if (wait_current()) return MBL_E_HIP;
    return hipMemcpy(d, s, n, hipMemcpyHostToDevice) == hipSuccess ? 0 : MBL_E_HIP;
}
extern "C" int mbl_memcpy_d2h(void *d, const void *s, size_t n)
{
    if (wait_current()) return MBL_E_HIP;
    return hipMemcpy(d, s, n, hipMemcpyDeviceToHost) == hipSuccess ? 0 : MBL_E_HIP;
}
extern "C" int mbl_memcpy_d2d(void *d, const void *s, size_t n)
{
    if (wait_current()) return MBL_E_HIP;
    return hipMemcpy(d, s, n, hipMemcpyDeviceToDevice) == hipSuccess ? 0 : MBL_E_HIP;
}
extern "C" int mbl_memset_d(void *d, int v, size_t n)
{
    if (wait_current()) return MBL_E_HIP;
    return (hipMemset(d, v, n) == hipSuccess && hipDeviceSynchronize() == hipSuccess) ? 0 : MBL_E_HIP;
}

extern "C" size_t mbl_plane_bytes(mbl_handle_t h, int nx, int ny)
{
    int mode = h ? storage_mode(h->cfg) : ST_NATIVE;
    if (mode < 0) return 0;
    return store_elem_bytes(mode) * (size_t)nx * (size_t)ny;
}

extern "C" int mbl_block_alloc(mbl_handle_t h, const AREA *a, D2Q9 *pp, Lbd **upx, Lbd **upy)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!a || !pp || a->nx <= 0 || a->ny <= 0) return set_err(h, MBL_E_INVALID);
    memset(pp, 0, sizeof(*pp));
    const size_t n = (size_t)a->nx * a->ny;
    const size_t pb = mbl_plane_bytes(h, a->nx, a->ny);
    bool ok = (hipMalloc(&pp->solid, n * sizeof(int)) == hipSuccess);
    Lbd **v = &pp->x;
    for (int b = 0; b < 7 && ok; b++) ok = (hipMalloc(&v[b], n * sizeof(Lbd)) == hipSuccess);
    for (int k = 0; k < 9 && ok; k++)
        ok = (hipMalloc(&pp->f[k], pb) == hipSuccess) && (hipMalloc(&pp->F[k], pb) == hipSuccess);
    if (ok && upx) ok = (hipMalloc(upx, n * sizeof(Lbd)) == hipSuccess);
    if (ok && upy) ok = (hipMalloc(upy, n * sizeof(Lbd)) == hipSuccess);
    if (!ok) {
        mbl_block_free(pp, upx ? *upx : nullptr, upy ? *upy : nullptr);
        return set_err(h, MBL_E_NOMEM);
    }
    return 0;
}

extern "C" void mbl_block_free(D2Q9 *pp, Lbd *upx, Lbd *upy)
{
    if (pp) {
        if (pp->solid) hipFree(pp->solid);
        Lbd **v = &pp->x;
        for (int b = 0; b < 25; b++)
            if (v[b]) hipFree(v[b]);
        memset(pp, 0, sizeof(*pp));
    }
    if (upx) hipFree(upx);
    if (upy) hipFree(upy);
}

static int plane_io(mbl_handle_t h, const D2Q9 *pp, int which, int k, size_t n, double *host, bool read)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!pp || k < 0 || k > 8 || !host) return set_err(h, MBL_E_INVALID);
    void *plane = which ? (void *)pp->f[k] : (void *)pp->F[k];
    if (!plane) return set_err(h, MBL_E_INVALID);
    int mode = storage_mode(h->cfg);
    // D2Q9 carries no dimensions: n is checked against the device allocation that holds the plane
    hipDeviceptr_t base = nullptr;
    size_t alloc = 0;
    if (hipMemGetAddressRange(&base, &alloc, (hipDeviceptr_t)plane) != hipSuccess) {
        (void)hipGetLastError();
        return set_err(h, MBL_E_INVALID);
    }
    const size_t before = (size_t)((const char *)plane - (const char *)base);
    if (before > alloc || n > (alloc - before) / store_elem_bytes(mode)) return set_err(h, MBL_E_INVALID);
    double *tmp = nullptr;
    HIPCHK(h, hipMalloc(&tmp, n * sizeof(double) + 8));
    int rc = 0;
    if (read) {
        rc = launch_decode(h->stream, mode, plane, k, n, tmp);
        if (!rc && hipMemcpyAsync(host, tmp, n * sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess)
            rc = MBL_E_HIP;
    } else {
        if (hipMemcpyAsync(tmp, host, n * sizeof(double), hipMemcpyHostToDevice, h->stream) != hipSuccess)
            rc = MBL_E_HIP;
        if (!rc) rc = launch_encode(h->stream, mode, plane, k, n, tmp);
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) rc = MBL_E_HIP;
    hipFree(tmp);
    return set_err(h, rc);
}

extern "C" int mbl_read_distribution(mbl_handle_t h, const D2Q9 *pp, int which, int k, size_t n, double *out)
{
    return plane_io(h, pp, which, k, n, out, true);
}

extern "C" int mbl_write_distribution(mbl_handle_t h, const D2Q9 *pp, int which, int k, size_t n, const double *in)
{
    return plane_io(h, pp, which, k, n, (double *)in, false);
}

// The fused kernel addresses planes with 32-bit byte offsets: every plane of a
// block (distributions and macros) must fit in 4 GiB.
static bool plane_fits(const mbl_context *h, int nx, int ny)
{
    const size_t eb = std::max(store_elem_bytes(storage_mode(h->cfg)), sizeof(Lbd));
    return (size_t)nx * (size_t)ny * eb <= ((size_t)1 << 32);
}

// ---------------------------------------------------------------------------
// block descriptors
// ---------------------------------------------------------------------------
// Row segments of tw lanes x vx nodes start at multiples of vx of the plane
// index, so a row whose length is not a multiple of vx is shifted by up to
// vx - 1 nodes inside its first segment: tiles along x cover nx + vx - 1.
static int seg_tiles(int nx, int vx, int tw)
{
    const int span = tw * vx, ext = (nx % vx) ? vx - 1 : 0;
    return (nx + ext + span - 1) / span;
}

static int seg_waste(int nx, int vx, int tw) { return seg_tiles(nx, vx, tw) * tw * vx - nx; }

static void set_tiles_x(BlockDesc &d, int tiles_x)
{
    d.tiles_x = tiles_x;
    d.tx_rcp = 0xFFFFFFFFu / (uint32_t)tiles_x;
}

static BlockDesc make_desc(const mbl_config &cfg, const AREA &a, const D2Q9 &pp, int level, Lbd *upx, Lbd *upy)
{
    BlockDesc d;
    memset(&d, 0, sizeof(d));
    d.tw = 64;
    d.bc = make_consts(cfg);
    for (int k = 0; k < 9; k++) {
        d.P[0][k] = (void *)pp.F[k];
        d.P[1][k] = (void *)pp.f[k];
    }
    d.flag = pp.solid;
    d.rho = pp.rho;
    d.ux = pp.ux;
    d.uy = pp.uy;
    d.upx = upx;
    d.upy = upy;
    d.nx = a.nx;
    d.ny = a.ny;
    d.level = level;
    d.m = a.m;
    d.base_y = a.base_y;
    d.dx = 1.0 / a.m;
    d.omega = a.rv_tau;
    d.inv_m = 1.0 / a.m;
    return d;
}

// Reads the node flags of a block and checks the rules the fused kernel
// relies on: bounce-back solids off the block edge, outlets with a fluid or
// ghost x-1 neighbour.
static int validate_flags(mbl_context *h, const AREA &a, const int *d_solid, bool periodic, bool *any_solid,
                          std::vector<int> *flags_out = nullptr)
{
    const size_t n = (size_t)a.nx * a.ny;
    std::vector<int> local;
    std::vector<int> &f = flags_out ? *flags_out : local;
    f.assign(n, 0);
    // the handle's stream does not synchronise with the null stream: wait for
    // the work that wrote the flags before the host copy
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(f.data(), d_solid, n * sizeof(int), hipMemcpyDeviceToHost));
    bool solid = false;
    for (int j = 0; j < a.ny; j++) {
        for (int i = 0; i < a.nx; i++) {
            int s = f[(size_t)j * a.nx + i];
            if (s == SOLID_SBB) {
                solid = true;
                if (!periodic && (i < 1 || j < 1 || i > a.nx - 2 || j > a.ny - 2)) {
                    fprintf(stderr, "mbl: SOLID_SBB node (%d,%d) on a block edge\n", i, j);
                    return set_err(h, MBL_E_INVALID);
                }
            } else if (s == MBL_SOLID_WALL) {
                solid = true;
            } else if (s == BOUNDARY_OUTLET || s == BOUNDARY_INLET) {
                // the extrapolation neighbour (x-1 for an outlet, x+1 for an inlet)
                int in = (s == BOUNDARY_OUTLET) ? i - 1 : i + 1;
                int nb = (in >= 0 && in < a.nx) ? f[(size_t)j * a.nx + in] : -1;
                if (nb != FLUID && nb != C_BOUNDARY) {
                    fprintf(stderr, "mbl: inlet/outlet node (%d,%d) without a fluid or ghost neighbour\n", i, j);
                    return set_err(h, MBL_E_INVALID);
                }
            } else if (s != FLUID && s != C_BOUNDARY) {
                fprintf(stderr, "mbl: unknown node flag %d at (%d,%d)\n", s, i, j);
                return set_err(h, MBL_E_INVALID);
            }
        }
    }
    if (any_solid) *any_solid = solid;
    return 0;
}

static void line_node(const P4 &l, int ip, int &i, int &j)
{
    if (l.xstart == l.xend) {
        i = l.xstart;
        j = l.ystart + ip;
    } else {
        i = l.xstart + ip;
        j = l.ystart;
    }
}

// c_c_boundary_transfer index rules (assit_multi.h:21-37)
static void cc_nodes(const Inf_Ff_store &L, int ip, int &i1, int &j1, int &i2, int &j2)
{
    i1 = j1 = i2 = j2 = 0;
    if (L.line_C.xstart == L.line_C.xend) {
        i1 = L.line_C.xstart;
        j1 = L.line_C.ystart + ip;
        i2 = L.line_F.xstart;
        j2 = L.line_F.ystart + ip;
    }
    if (L.line_C.ystart == L.line_C.yend) {
        i1 = L.line_C.xstart + ip;
        j1 = L.line_C.ystart;
        i2 = L.line_F.xstart + ip;
        j2 = L.line_F.ystart;
    }
}

// ---------------------------------------------------------------------------
// PCR elimination factors of a spline matrix. Level q pairs row i with rows i -/+ 2^q:
// alpha = -a_i / b_{i-s}, gamma = -c_i / b_{i+s}, and the row's new couplings and diagonal
// follow from them; the right-hand sides are then updated on the device with the stored
// factors. The multiply-adds are spelled as fused operations, as the device forms them.
// ---------------------------------------------------------------------------
namespace mbl {
int pcr_factor_table(const double *l, const double *m, const double *u, int N, std::vector<double> &tab)
{
    std::vector<double> a(l, l + N), b(m, m + N), c(u, u + N), na(N), nb(N), nc(N);
    std::vector<double> lev;
    int levels = 0;
    for (int s = 1; s < N; s <<= 1) {
        const size_t base = lev.size();
        lev.resize(base + 2 * (size_t)N, 0.0);
        bool decoupled = true;
        for (int i = 0; i < N; i++) {
            const int im = i - s, iq = i + s;
            double xa = 0.0, xc = 0.0, xb = b[i];
            if (im >= 0) {
                const double alpha = -a[i] / b[im];
                lev[base + i] = alpha;
                xa = alpha * a[im];
                xb = std::fma(alpha, c[im], xb);
            }
            if (iq < N) {
                const double gamma = -c[i] / b[iq];
                lev[base + N + i] = gamma;
                xc = gamma * c[iq];
                xb = std::fma(gamma, a[iq], xb);
            }
            na[i] = xa;
            nb[i] = xb;
            nc[i] = xc;
            decoupled = decoupled && std::fabs(xa) <= 0x1p-60 * std::fabs(xb) && std::fabs(xc) <= 0x1p-60 * std::fabs(xb);
        }
        a.swap(na);
        b.swap(nb);
        c.swap(nc);
        levels++;
        if (decoupled) break;
    }
    tab = lev;
    tab.insert(tab.end(), b.begin(), b.end());
    return levels;
}
} // namespace mbl

// ---------------------------------------------------------------------------
// Hazard scheduler for one batch of interface ops issued in reference order.
// Units are (block, node, field) with field 0 = rho/u, 1 = F. A write that is
// overwritten later in the batch before any read is dead and its entry bit is
// cleared; the remaining ops are levelled so that every read-after-write and
// write-after-read pair lands in successive sub-batches, which then run as
// one launch each.
// ---------------------------------------------------------------------------
struct OpBuild {
    LineOp op;
    int uid = -1;  // serial number (c_f ops named by the redirected line ends of f_c ops)
    int red_uid[2] = {-1, -1};
    int line = -1; // coarse/fine line whose ring the op writes (spline) or reads (temporal, c_f)
    std::vector<int> gred; // SPLINE: per coarse point, the (block, node) read in its place, or -1, -1
    std::vector<int> src, dst;
    std::vector<uint8_t> mask;
    std::vector<uint64_t> rd;                          // read units
    std::vector<std::pair<uint64_t, int>> wr;          // write units, (unit, entry*2+field)
};

static inline uint64_t unit(int blk, size_t node, int field)
{
    return ((uint64_t)blk << 40) | ((uint64_t)node << 1) | (uint64_t)field;
}

static std::vector<int> schedule(std::vector<OpBuild> &ops)
{
    struct St {
        int wop = -1, went = 0;
        bool read_since = false;
        std::vector<int> readers;
    };
    std::unordered_map<uint64_t, St> st;
    std::vector<int> lvl(ops.size(), 0);
    for (size_t j = 0; j < ops.size(); j++) {
        int L = 0;
        for (uint64_t u : ops[j].rd) {
            St &s = st[u];
            if (s.wop >= 0) L = std::max(L, lvl[s.wop] + 1);
            s.read_since = true;
            if (s.readers.empty() || s.readers.back() != (int)j) s.readers.push_back((int)j);
        }
        for (auto &w : ops[j].wr) {
            St &s = st[w.first];
            if (s.wop >= 0 && !s.read_since) {
                int ent = s.went >> 1, fld = s.went & 1;
                ops[s.wop].mask[ent] &= (uint8_t)~(1u << fld);
            }
            for (int r : s.readers)
                if (r != (int)j) L = std::max(L, lvl[r] + 1);
            s.wop = (int)j;
            s.went = w.second;
            s.read_since = false;
            s.readers.clear();
        }
        lvl[j] = L;
    }
    return lvl;
}

static bool in_block(const AREA &a, int i, int j) { return i >= 0 && j >= 0 && i < a.nx && j < a.ny; }

// ---------------------------------------------------------------------------
// layout build
// ---------------------------------------------------------------------------
struct EvoArgs {
    Lbi *asp, *ls3, *ltc, *lcc;
    AREA *areas;
    D2Q9 *pps;
    Inf_Ff_store *lc, *lf, *lccs;
    Ff_Store *t3;
    FM_spline3 *s3;
    const Consts *bc = nullptr; // per-block boundary constants (batched cases); null: the handle's
};

static uint64_t layout_sig(mbl_context *h, const EvoArgs &A, int nblocks, int nline, int nline2)
{
    uint64_t s = 0x1234;
    s = mix64(s, (uintptr_t)A.bc);
    s = mix64(s, (uint64_t)nblocks);
    s = mix64(s, (uint64_t)nline);
    s = mix64(s, (uint64_t)nline2);
    for (int l = 0; l <= h->cfg.LEVEL; l++) {
        s = mix64(s, (uint64_t)A.asp[l]);
        s = mix64(s, (uint64_t)A.ltc[l]);
        s = mix64(s, (uint64_t)A.ls3[l]);
        s = mix64(s, (uint64_t)A.lcc[l]);
    }
    for (int a = 0; a < nblocks; a++) {
        const AREA &ar = A.areas[a];
        s = mix64(s, (uint64_t)ar.m | ((uint64_t)ar.nx << 16) | ((uint64_t)ar.ny << 40));
        uint64_t tb;
        memcpy(&tb, &ar.tau, 8);
        s = mix64(s, tb);
        const D2Q9 &p = A.pps[a];
        for (int k = 0; k < 9; k++) {
            uintptr_t x = (uintptr_t)p.F[k], y = (uintptr_t)p.f[k];
            s = mix64(s, std::min(x, y));
            s = mix64(s, std::max(x, y));
        }
        s = mix64(s, (uintptr_t)p.solid);
        s = mix64(s, (uintptr_t)p.rho);
    }
    auto line = [&](const Inf_Ff_store &L) {
        s = mix64(s, (uint64_t)L.areac_n | ((uint64_t)L.areaf_n << 20) | ((uint64_t)L.Nc << 40));
        s = mix64(s, (uint64_t)L.Nf);
        const int *q = &L.line_C.xstart;
        for (int i = 0; i < 8; i++) s = mix64(s, (uint64_t)(uint32_t)q[i]);
    };
    for (int b = 0; b < nline; b++) {
        line(A.lc[b]);
        line(A.lf[b]);
        const F_k3 *ft = &A.t3[b].f_t1;
        for (int sl = 0; sl < 3; sl++)
            for (int k = 0; k < 9; k++) s = mix64(s, (uintptr_t)ft[sl].F[k]);
        s = mix64(s, (uintptr_t)A.s3[b].Varibs_M);
        s = mix64(s, (uintptr_t)A.s3[b].l);
    }
    for (int c = 0; c < nline2; c++) line(A.lccs[c]);
    return s;
}

static int build_layout(mbl_context *h, const EvoArgs &A, int nblocks, int nline, int nline2, uint64_t sig,
                        Layout *&out)
{
    const mbl_config &cfg = h->cfg;
    Layout *L = new Layout();
    L->sig = sig;
    L->LEVEL = cfg.LEVEL;
    L->nblocks = nblocks;
    L->number_line = nline;
    L->number_line2 = nline2;
    L->periodic = (cfg.bc_mode == MBL_BC_PERIODIC);
    for (int l = 0; l < cfg.LEVEL; l++) L->mi[l] = cfg.mi[l];
    for (int l = 0; l <= cfg.LEVEL; l++) {
        L->ASP[l] = A.asp[l];
        L->LTC[l] = A.ltc[l];
        L->LS3[l] = A.ls3[l];
        L->LCC[l] = A.lcc[l];
    }
    auto fail = [&](const char *why) {
        fprintf(stderr, "mbl: invalid layout: %s\n", why);
        L->release();
        delete L;
        return set_err(h, MBL_E_INVALID);
    };
    if (L->ASP[0] != 0 || L->ASP[cfg.LEVEL] != nblocks) return fail("ASP does not cover the blocks");
    if (L->periodic && nblocks != 1) return fail("periodic mode needs exactly one block");
    // blocks
    L->hb.resize(nblocks);
    std::vector<int> &level_of = L->level_of;
    level_of.assign(nblocks, -1);
    for (int l = 0; l < cfg.LEVEL; l++)
        for (int a = L->ASP[l]; a < L->ASP[l + 1]; a++) level_of[a] = l;
    std::vector<std::vector<int>> hflags(nblocks); // node flags on the host (scheduling rules below)
    const int rows = std::max(h->tpb / 64, MBL_TILE_ROWS);
    L->rows = rows;
    std::vector<int> tile_blk; // per tile: the block's index in the layout's descriptor table
    for (int l = 0; l < cfg.LEVEL; l++) {
        L->tile_off[l] = (int)tile_blk.size();
        int vx = (h->vx_set || L->periodic) ? h->vx : 1;
        if (L->periodic) // periodic rows wrap, so they hold whole segments
            for (int a = L->ASP[l]; a < L->ASP[l + 1]; a++)
                while (vx > 1 && A.areas[a].nx % vx) vx >>= 1;
        L->vx[l] = vx;
        int tiles = 0;
        for (int a = L->ASP[l]; a < L->ASP[l + 1]; a++) {
            const AREA &ar = A.areas[a];
            if (ar.nx < 2 || ar.ny < 2) return fail("block smaller than 2x2");
            if (!plane_fits(h, ar.nx, ar.ny)) return fail("block plane larger than 4 GiB");
            for (int k = 0; k < 9; k++)
                if (!A.pps[a].F[k] || !A.pps[a].f[k] || A.pps[a].F[k] == A.pps[a].f[k]) return fail("F/f planes");
            BlockDesc d = make_desc(cfg, ar, A.pps[a], l, nullptr, nullptr);
            if (A.bc) d.bc = A.bc[a];
            // 32-lane row segments when they waste fewer lanes than 64-lane ones (periodic
            // blocks keep 64: the kernel relies on it there)
            d.tw = 64;
            if (!L->periodic && seg_waste(ar.nx, vx, 32) < seg_waste(ar.nx, vx, 64)) d.tw = 32;
            d.trows = rows * (64 / d.tw);
            set_tiles_x(d, seg_tiles(ar.nx, vx, d.tw));
            // tiles are numbered across all levels (a launch may cover several adjacent levels)
            d.tile_begin = L->tile_off[l] + tiles;
            tiles += d.tiles_x * ((ar.ny + d.trows - 1) / d.trows);
            tile_blk.resize(L->tile_off[l] + tiles, a);
            L->hb[a] = d;
            if (validate_flags(h, ar, A.pps[a].solid, L->periodic, nullptr, &hflags[a])) {
                L->release();
                delete L;
                return MBL_E_INVALID;
            }
        }
        L->tiles[l] = tiles;
    }
    // interface ops, batch by batch, in the order evolution.h issues them
    std::vector<OpBuild> all;
    auto check_cf_line = [&](int b) -> bool {
        const Inf_Ff_store &lc = A.lc[b];
        if (lc.areac_n < 0 || lc.areac_n >= nblocks || lc.areaf_n < 0 || lc.areaf_n >= nblocks) return false;
        const AREA &ac = A.areas[lc.areac_n], &af = A.areas[lc.areaf_n];
        int r = af.m / ac.m;
        if (r < 1 || lc.Nf != r * (lc.Nc - 1) + 1) return false;
        if (A.t3[b].N != lc.Nf || A.s3[b].N != lc.Nc) return false;
        if (ac.tau == 1.0 || af.tau == 1.0) return false;
        return true;
    };
    for (int b = 0; b < nline; b++)
        if (!check_cf_line(b)) return fail("coarse/fine line");

    // marker group of each line: the level whose spline step rotates its ring markers
    std::vector<int> mgroup(std::max(1, nline), 0);
    for (int l = 0; l < cfg.LEVEL; l++)
        for (int b = L->LS3[l]; b < L->LS3[l + 1] && b < nline; b++) mgroup[b] = l;
    auto make_T_or_CF = [&](int b, int type) -> OpBuild {
        OpBuild o;
        memset(&o.op, 0, sizeof(LineOp));
        const Inf_Ff_store &lc = A.lc[b];
        const AREA &af = A.areas[lc.areaf_n];
        o.op.type = type;
        o.op.mgroup = mgroup[b];
        o.line = b;
        o.op.n = A.t3[b].N;
        o.op.dst_block = lc.areaf_n;
        o.op.src_block = lc.areaf_n;
        const F_k3 *ft = &A.t3[b].f_t1;
        for (int sl = 0; sl < 3; sl++)
            for (int k = 0; k < 9; k++) o.op.ring[sl][k] = ft[sl].F[k];
        for (int ip = 0; ip < o.op.n; ip++) {
            int i, j;
            line_node(lc.line_F, ip, i, j);
            if (!in_block(af, i, j)) o.op.n = -1;
            size_t node = (size_t)j * af.nx + i;
            o.dst.push_back((int)node);
            o.mask.push_back(2);
            o.wr.push_back({unit(lc.areaf_n, node, 1), ip * 2 + 1});
        }
        return o;
    };
    auto make_CC = [&](int c) -> OpBuild {
        OpBuild o;
        memset(&o.op, 0, sizeof(LineOp));
        const Inf_Ff_store &l = A.lccs[c];
        o.op.type = OP_CC;
        o.op.n = l.Nc;
        o.op.src_block = l.areac_n;
        o.op.dst_block = l.areaf_n;
        if (l.areac_n < 0 || l.areac_n >= nblocks || l.areaf_n < 0 || l.areaf_n >= nblocks) {
            o.op.n = -1;
            return o;
        }
        const AREA &a1 = A.areas[l.areac_n], &a2 = A.areas[l.areaf_n];
        for (int ip = 0; ip < l.Nc; ip++) {
            int i1, j1, i2, j2;
            cc_nodes(l, ip, i1, j1, i2, j2);
            if (!in_block(a1, i1, j1) || !in_block(a2, i2, j2)) {
                o.op.n = -1;
                return o;
            }
            size_t s = (size_t)j1 * a1.nx + i1, d = (size_t)j2 * a2.nx + i2;
            o.src.push_back((int)s);
            o.dst.push_back((int)d);
            o.mask.push_back(3);
            o.rd.push_back(unit(l.areac_n, s, 0));
            o.rd.push_back(unit(l.areac_n, s, 1));
            o.wr.push_back({unit(l.areaf_n, d, 0), ip * 2 + 0});
            o.wr.push_back({unit(l.areaf_n, d, 1), ip * 2 + 1});
        }
        return o;
    };
    auto make_S = [&](int b) -> OpBuild {
        OpBuild o;
        memset(&o.op, 0, sizeof(LineOp));
        const Inf_Ff_store &lc = A.lc[b];
        const AREA &ac = A.areas[lc.areac_n], &af = A.areas[lc.areaf_n];
        const FM_spline3 &S3 = A.s3[b];
        int r = af.m / ac.m;
        o.op.type = OP_SPLINE;
        o.op.mgroup = mgroup[b];
        o.line = b;
        o.op.n = lc.Nf;
        o.op.nc = lc.Nc;
        o.op.ratio = r;
        o.op.para = (af.tau - 1.0) / (r * (ac.tau - 1.0));
        o.op.src_block = lc.areac_n;
        o.op.dst_block = lc.areaf_n;
        o.op.tl = S3.l;
        o.op.tm = S3.m;
        o.op.tu = S3.u;
        o.op.varibs = S3.Varibs_M;
        Lbd *const *gc = &S3.rho;
        for (int c = 0; c < 12; c++) o.op.gcol[c] = gc[c];
        const F_k3 *ft = &A.t3[b].f_t1;
        for (int sl = 0; sl < 3; sl++)
            for (int k = 0; k < 9; k++) o.op.ring[sl][k] = ft[sl].F[k];
        for (int ip = 0; ip < lc.Nc; ip++) {
            int i, j;
            line_node(lc.line_C, ip, i, j);
            if (!in_block(ac, i, j)) {
                o.op.n = -1;
                return o;
            }
            size_t s = (size_t)j * ac.nx + i;
            o.src.push_back((int)s);
            o.rd.push_back(unit(lc.areac_n, s, 0));
            o.rd.push_back(unit(lc.areac_n, s, 1));
        }
        for (int ip = 0; ip < lc.Nf; ip++) {
            int i, j;
            line_node(lc.line_F, ip, i, j);
            if (!in_block(af, i, j)) {
                o.op.n = -1;
                return o;
            }
            size_t d = (size_t)j * af.nx + i;
            o.dst.push_back((int)d);
            o.mask.push_back(1);
            o.wr.push_back({unit(lc.areaf_n, d, 0), ip * 2 + 0});
        }
        return o;
    };
    auto make_FC = [&](int b) -> OpBuild {
        OpBuild o;
        memset(&o.op, 0, sizeof(LineOp));
        const Inf_Ff_store &l = A.lf[b];
        const AREA &ac = A.areas[l.areac_n], &af = A.areas[l.areaf_n];
        int r = af.m / ac.m;
        o.op.type = OP_FC;
        o.op.n = l.Nc;
        o.op.ratio = r;
        o.op.para = r * (ac.tau - 1.0) / (af.tau - 1.0);
        o.op.src_block = l.areaf_n;
        o.op.dst_block = l.areac_n;
        for (int ip = 0; ip < l.Nc; ip++) {
            int ic, jc, i_f, j_f;
            if (l.line_C.xstart == l.line_C.xend) {
                ic = l.line_C.xstart;
                jc = l.line_C.ystart + ip;
                i_f = l.line_F.xstart;
                j_f = l.line_F.ystart + r * ip;
            } else {
                ic = l.line_C.xstart + ip;
                jc = l.line_C.ystart;
                i_f = l.line_F.xstart + r * ip;
                j_f = l.line_F.ystart;
            }
            if (!in_block(ac, ic, jc) || !in_block(af, i_f, j_f)) {
                o.op.n = -1;
                return o;
            }
            size_t s = (size_t)j_f * af.nx + i_f, d = (size_t)jc * ac.nx + ic;
            o.src.push_back((int)s);
            o.dst.push_back((int)d);
            o.mask.push_back(3);
            o.rd.push_back(unit(l.areaf_n, s, 0));
            o.rd.push_back(unit(l.areaf_n, s, 1));
            o.wr.push_back({unit(l.areac_n, d, 0), ip * 2 + 0});
            o.wr.push_back({unit(l.areac_n, d, 1), ip * 2 + 1});
        }
        return o;
    };

    // What each op touches, for the scheduling of a whole refinement cycle (schedule_cycle):
    // node units, the ring of its line, and the levels whose collide epoch it belongs to. An
    // op that reads a block's planes or macros, or writes its F, must stay between the same two
    // collides of that block's level as in the reference order (the collide rewrites every F of
    // the block and the macros of its fluid nodes, and reads every F). A spline writes only the
    // macros of its fine line; those are C_BOUNDARY nodes, which no collide writes, and a
    // collide reads the macros of a C_BOUNDARY node only next to an inlet/outlet node, so the
    // spline is tied to the fine level's epoch only where a written node is not C_BOUNDARY or
    // the fine block holds inlet/outlet nodes.
    auto op_access = [&](const OpBuild &o) -> OpAccess {
        OpAccess ac;
        for (uint64_t u : o.rd) ac.rd.push_back(u);
        for (auto &w : o.wr) ac.wr.push_back(w.first);
        std::sort(ac.rd.begin(), ac.rd.end());
        ac.rd.erase(std::unique(ac.rd.begin(), ac.rd.end()), ac.rd.end());
        std::sort(ac.wr.begin(), ac.wr.end());
        ac.wr.erase(std::unique(ac.wr.begin(), ac.wr.end()), ac.wr.end());
        ac.ring_line = o.line;
        ac.ring_write = (o.op.type == OP_SPLINE);
        const int ls = level_of[o.op.src_block], ld = level_of[o.op.dst_block];
        switch (o.op.type) {
        case OP_TEMPORAL:
        case OP_CF: ac.epochs = 1u << ld; break;
        case OP_CC:
        case OP_FC: ac.epochs = (1u << ls) | (1u << ld); break;
        case OP_SPLINE: {
            ac.epochs = 1u << ls;
            const std::vector<int> &fl = hflags[o.op.dst_block];
            bool tie = false;
            for (int v : fl) tie = tie || v == BOUNDARY_INLET || v == BOUNDARY_OUTLET;
            for (int d : o.dst) tie = tie || fl[d] != C_BOUNDARY;
            if (tie) ac.epochs |= 1u << ld;
            break;
        }
        default: ac.epochs = (1u << ls) | (1u << ld);
        }
        return ac;
    };
    // Batches, each scheduled on its own; ops appended to `all` grouped by sub-batch.
    int next_uid = 0;
    // A c_f / f_c batch (evolution.h:52-62, line by line): an f_c op reads fine nodes that the
    // c_f op of another line of the batch writes where the lines meet (the ends of the f_c
    // line on the fine block's ghost edges). Those reads would split the batch into two
    // launches; instead each such entry reads the value the reference order gives it: what an
    // earlier c_f op wrote (its ring slot), or, before a later c_f op writes the node, the
    // ghost value the last collide carried over (the other plane set; ghost nodes only). At
    // most two such entries per f_c op; otherwise the reads stay and the batch splits.
    auto redirect_fc_ends = [&](std::vector<OpBuild> &ops) {
        for (auto &o : ops) o.uid = next_uid++;
        std::unordered_map<uint64_t, std::vector<std::pair<int, int>>> cfw; // unit -> (op, entry) writers
        for (size_t j = 0; j < ops.size(); j++)
            if (ops[j].op.type == OP_CF)
                for (size_t e = 0; e < ops[j].dst.size(); e++)
                    if (ops[j].mask[e] & 2) cfw[unit(ops[j].op.dst_block, (size_t)ops[j].dst[e], 1)].push_back({(int)j, (int)e});
        for (size_t j = 0; j < ops.size(); j++) {
            OpBuild &o = ops[j];
            if (o.op.type != OP_FC) continue;
            struct Red { int e, t, op, ce; uint64_t u; };
            std::vector<Red> red;
            bool ok = true;
            for (size_t e = 0; e < o.src.size() && ok; e++) {
                const uint64_t u = unit(o.op.src_block, (size_t)o.src[e], 1);
                auto it = cfw.find(u);
                if (it == cfw.end()) continue;
                int before = -1, before_e = -1;
                for (auto &w : it->second)
                    if (w.first < (int)j) {
                        before = w.first;
                        before_e = w.second;
                    }
                if (before >= 0)
                    red.push_back({(int)e, 1, before, before_e, u});
                else if (hflags[o.op.src_block][o.src[e]] == C_BOUNDARY)
                    red.push_back({(int)e, 2, -1, -1, u});
                else
                    ok = false;
            }
            if (!ok || red.size() > 2 || red.empty()) continue;
            o.op.red_n = (int)red.size();
            for (size_t r = 0; r < red.size(); r++) {
                o.op.red_e[r] = red[r].e;
                o.op.red_t[r] = red[r].t;
                o.op.red_ce[r] = red[r].ce;
                o.red_uid[r] = red[r].op >= 0 ? ops[red[r].op].uid : -1;
                o.rd.erase(std::remove(o.rd.begin(), o.rd.end(), red[r].u), o.rd.end());
            }
        }
    };
    // A spline gathers its coarse line right after the same-level copies of its level
    // (evolution.h:26-33, nothing in between): a node whose two fields a copy wrote holds the
    // copy's source node, which no copy of the batch writes, so the spline reads that source
    // node instead (LineOp::gred) and no longer waits for the copies. One-workgroup splines
    // only (the path that reads LineOp::gred).
    auto redirect_spline_reads = [&](std::vector<OpBuild> &sops, size_t cc0, size_t cc1) {
        std::unordered_map<uint64_t, std::pair<int, int>> live; // unit -> (copy op in all, entry)
        for (size_t j = cc0; j < cc1; j++)
            for (auto &w : all[j].wr) {
                const int ent = w.second >> 1, fld = w.second & 1;
                if ((all[j].mask[ent] >> fld) & 1) live[w.first] = {(int)j, ent};
            }
        if (live.empty()) return;
        for (auto &o : sops) {
            if (o.op.type != OP_SPLINE || launch_spline_path(o.op.nc) != 0) continue;
            std::vector<int> gr(2 * o.src.size(), -1);
            bool any = false;
            for (size_t ip = 0; ip < o.src.size(); ip++) {
                const uint64_t u0 = unit(o.op.src_block, (size_t)o.src[ip], 0);
                const uint64_t u1 = unit(o.op.src_block, (size_t)o.src[ip], 1);
                auto a = live.find(u0), b = live.find(u1);
                if (a == live.end() || b == live.end() || a->second != b->second) continue;
                const OpBuild &c = all[a->second.first];
                const int sb = c.op.src_block, sn = c.src[a->second.second];
                if (live.count(unit(sb, (size_t)sn, 0)) || live.count(unit(sb, (size_t)sn, 1))) continue;
                gr[2 * ip] = sb;
                gr[2 * ip + 1] = sn;
                any = true;
                for (auto &u : o.rd) {
                    if (u == u0)
                        u = unit(sb, (size_t)sn, 0);
                    else if (u == u1)
                        u = unit(sb, (size_t)sn, 1);
                }
            }
            if (any) o.gred = std::move(gr);
        }
    };
    auto add_batch = [&](std::vector<OpBuild> ops, Batch &batch) -> bool {
        for (auto &o : ops)
            if (o.op.n < 0) return false;
        std::vector<int> lv = schedule(ops);
        int nsub = 0;
        for (int v : lv) nsub = std::max(nsub, v + 1);
        for (int s = 0; s < nsub; s++) {
            Sub sb;
            sb.off = (int)all.size();
            for (size_t j = 0; j < ops.size(); j++) {
                if (lv[j] != s) continue;
                sb.count++;
                sb.max_n = std::max(sb.max_n, ops[j].op.n);
                sb.max_nc = std::max(sb.max_nc, ops[j].op.nc);
                sb.spline = (ops[j].op.type == OP_SPLINE);
                L->acc.push_back(op_access(ops[j]));
                all.push_back(std::move(ops[j]));
            }
            if (sb.count) batch.push_back(sb);
        }
        return true;
    };
    for (int l = 0; l < cfg.LEVEL; l++) {
        std::vector<OpBuild> ops;
        for (int b = L->LTC[l]; b < L->LTC[l + 1]; b++) ops.push_back(make_T_or_CF(b, OP_TEMPORAL));
        if (!add_batch(std::move(ops), L->bT[l])) return fail("temporal line outside its block");
        ops.clear();
        if (nline2 != 0)
            for (int c = L->LCC[l]; c < L->LCC[l + 1]; c++) ops.push_back(make_CC(c));
        const size_t cc0 = all.size();
        if (!add_batch(std::move(ops), L->bCC[l])) return fail("same-level line outside its blocks");
        const size_t cc1 = all.size();
        ops.clear();
        if (l != cfg.LEVEL - 1)
            for (int b = L->LS3[l]; b < L->LS3[l + 1]; b++) ops.push_back(make_S(b));
        for (auto &o : ops)
            if (o.op.n < 0) return fail("spline line outside its blocks");
        redirect_spline_reads(ops, cc0, cc1);
        if (!add_batch(std::move(ops), L->bS[l])) return fail("spline line outside its blocks");
        ops.clear();
        if (l != 0)
            for (int b = L->LTC[l]; b < L->LTC[l + 1]; b++) {
                ops.push_back(make_T_or_CF(b, OP_CF));
                ops.push_back(make_FC(b));
            }
        for (auto &o : ops)
            if (o.op.n < 0) return fail("coarse/fine transfer line outside its blocks");
        redirect_fc_ends(ops);
        if (!add_batch(std::move(ops), L->bCF[l])) return fail("coarse/fine transfer line outside its blocks");
    }
    // PCR factor tables of the spline matrices, read once from the caller's l, m, u (init_tri_A
    // writes them before the first evolution(); the spline reads them, never writes them)
    std::vector<double> pcr_all;
    std::vector<std::pair<size_t, int>> pcr_of(all.size(), {0, 0}); // per op: offset, levels
    {
        std::map<std::vector<double>, std::pair<size_t, int>> seen;
        for (size_t j = 0; j < all.size(); j++) {
            const LineOp &op = all[j].op;
            if (op.type != OP_SPLINE) continue;
            const int N = op.nc;
            std::vector<Lbd> hl(N), hm(N), hu(N);
            if (hipMemcpy(hl.data(), op.tl, sizeof(Lbd) * N, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(hm.data(), op.tm, sizeof(Lbd) * N, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(hu.data(), op.tu, sizeof(Lbd) * N, hipMemcpyDeviceToHost) != hipSuccess) {
                L->release();
                delete L;
                return set_err(h, MBL_E_HIP);
            }
            std::vector<double> key(3 * (size_t)N);
            for (int i = 0; i < N; i++) {
                key[i] = (double)hl[i];
                key[N + i] = (double)hm[i];
                key[2 * N + i] = (double)hu[i];
            }
            auto it = seen.find(key);
            if (it == seen.end()) {
                std::vector<double> tab;
                const int lv = pcr_factor_table(key.data(), key.data() + N, key.data() + 2 * N, N, tab);
                if (lv > PCR_MAX_LEVELS && launch_spline_path(N) == 0) return fail("spline matrix does not decouple");
                it = seen.emplace(key, std::make_pair(pcr_all.size(), lv)).first;
                pcr_all.insert(pcr_all.end(), tab.begin(), tab.end());
            }
            pcr_of[j] = it->second;
        }
    }
    // upload tables
    size_t nidx = 0, nmask = 0, nM = 0;
    for (auto &o : all) {
        nidx += o.src.size() + o.dst.size() + (o.gred.empty() ? 0 : o.gred.size() + 1);
        nmask += o.mask.size();
        if (o.op.type == OP_SPLINE) nM += 12 * (size_t)o.op.nc;
    }
    if (hipMalloc(&L->db, sizeof(BlockDesc) * std::max(1, nblocks)) != hipSuccess ||
        hipMalloc(&L->dops, sizeof(LineOp) * std::max<size_t>(1, all.size())) != hipSuccess ||
        hipMalloc(&L->didx, sizeof(int) * std::max<size_t>(1, nidx)) != hipSuccess ||
        hipMalloc(&L->dmask, std::max<size_t>(1, nmask)) != hipSuccess ||
        hipMalloc(&L->dM64, sizeof(double) * std::max<size_t>(1, nM)) != hipSuccess ||
        hipMalloc(&L->dtile, sizeof(int) * std::max<size_t>(1, tile_blk.size())) != hipSuccess ||
        hipMalloc(&L->dpcr, sizeof(double) * std::max<size_t>(1, pcr_all.size())) != hipSuccess) {
        L->release();
        delete L;
        return set_err(h, MBL_E_NOMEM);
    }
    std::vector<int> hidx;
    hidx.reserve(nidx);
    std::vector<uint8_t> hmask;
    hmask.reserve(nmask);
    size_t mo = 0;
    std::unordered_map<int, size_t> uid_index;
    for (size_t j = 0; j < all.size(); j++)
        if (all[j].uid >= 0) uid_index[all[j].uid] = j;
    for (size_t j = 0; j < all.size(); j++) {
        const OpBuild &o = all[j];
        LineOp op = o.op;
        for (int r = 0; r < op.red_n; r++)
            if (op.red_t[r] == 1) {
                const LineOp &cf = all[uid_index.at(o.red_uid[r])].op;
                op.red_mg[r] = cf.mgroup;
                for (int sl = 0; sl < 3; sl++)
                    for (int k = 0; k < MBL_Q; k++) op.red_ring[r][sl][k] = cf.ring[sl][k];
            }
        op.src_idx = o.src.empty() ? nullptr : L->didx + hidx.size();
        hidx.insert(hidx.end(), o.src.begin(), o.src.end());
        op.dst_idx = o.dst.empty() ? nullptr : L->didx + hidx.size();
        hidx.insert(hidx.end(), o.dst.begin(), o.dst.end());
        op.gred = nullptr;
        if (!o.gred.empty()) { // (block, node) pairs, 8-byte aligned
            if (hidx.size() & 1) hidx.push_back(0);
            op.gred = reinterpret_cast<const int2 *>(L->didx + hidx.size());
            hidx.insert(hidx.end(), o.gred.begin(), o.gred.end());
        }
        op.mask = L->dmask + hmask.size();
        hmask.insert(hmask.end(), o.mask.begin(), o.mask.end());
        if (op.type == OP_SPLINE) {
            op.M64 = L->dM64 + mo;
            mo += 12 * (size_t)op.nc;
            op.pcr = L->dpcr + pcr_of[j].first;
            op.pcr_levels = pcr_of[j].second;
        }
        L->hops.push_back(op);
        L->hsrc.push_back(o.src);
        L->hdst.push_back(o.dst);
        L->hmsk.push_back(o.mask);
    }
    bool ok = hipMemcpy(L->db, L->hb.data(), sizeof(BlockDesc) * nblocks, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && !tile_blk.empty())
        ok = hipMemcpy(L->dtile, tile_blk.data(), sizeof(int) * tile_blk.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && !L->hops.empty())
        ok = hipMemcpy(L->dops, L->hops.data(), sizeof(LineOp) * L->hops.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && !hidx.empty())
        ok = hipMemcpy(L->didx, hidx.data(), sizeof(int) * hidx.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && !hmask.empty()) ok = hipMemcpy(L->dmask, hmask.data(), hmask.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && !pcr_all.empty())
        ok = hipMemcpy(L->dpcr, pcr_all.data(), sizeof(double) * pcr_all.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        L->release();
        delete L;
        return set_err(h, MBL_E_HIP);
    }
    // The fused coarse step: the last collide of each level sums the residual terms of the
    // nodes off the block edges it writes the macros of (fluid, inlet, outlet), which no
    // interface op may rewrite; the other nodes (block edges, where same-level copies end on
    // the channel's inlet and outlet; solid and ghost nodes) are summed after the cycle, from
    // a list.
    L->total_tiles = (int)tile_blk.size();
    bool step_ok = !L->periodic;
    for (int l = 0; l < cfg.LEVEL; l++) step_ok = step_ok && L->vx[l] == 1;
    auto summed_by_collide = [&](int a, size_t node) {
        const int f = hflags[a][node];
        const int nx = A.areas[a].nx, ny = A.areas[a].ny, i = (int)(node % nx), j = (int)(node / nx);
        const bool edge = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
        return !edge && (f == FLUID || f == BOUNDARY_INLET || f == BOUNDARY_OUTLET);
    };
    for (const OpBuild &o : all)
        for (size_t e = 0; e < o.mask.size() && step_ok; e++)
            if ((o.mask[e] & 1) && summed_by_collide(o.op.dst_block, (size_t)o.dst[e])) step_ok = false;
    if (step_ok) {
        std::vector<int4> &list = L->hlist;
        list.clear();
        for (int a = 0; a < nblocks; a++)
            for (size_t i = 0; i < hflags[a].size(); i++)
                if (!summed_by_collide(a, i)) list.push_back(make_int4(a, (int)i, a, (int)i));
        L->nlist = (int)list.size();
        const int wgs = std::max(1, (L->nlist + 255) / 256);
        ok = hipMalloc(&L->dlist, sizeof(int4) * std::max<size_t>(1, list.size())) == hipSuccess &&
             hipMalloc(&L->dres, sizeof(double) * 3 * ((size_t)L->total_tiles + wgs)) == hipSuccess &&
             (list.empty() ||
              hipMemcpy(L->dlist, list.data(), sizeof(int4) * list.size(), hipMemcpyHostToDevice) == hipSuccess);
        if (!ok) {
            L->release();
            delete L;
            return set_err(h, MBL_E_NOMEM);
        }
    }
    L->step_ok = step_ok;
    out = L;
    return 0;
}

// ---------------------------------------------------------------------------
// refinement cycle (evolution.h:2-65)
// ---------------------------------------------------------------------------
// cycle() walks the recursion in the reference order, rotates the ring markers as
// evolution.h:41-46 does and records every operation (CycleItem) with the parities and
// markers in force. schedule_cycle() turns the recorded order into launches: it derives
// every read-after-write, write-after-read and write-after-write dependence on the data the
// operations touch (collide epochs per level, interface node units, ring slots), gives each
// operation the earliest position its dependences allow, and merges the operations of one
// position into one launch where their kinds allow (stream-collides of adjacent levels,
// interface ops of any line, splines of any level; interface ops with one-workgroup splines or
// with a stream-collide; temporal interpolations inside the splines that wrote their slots).
// Every read still sees the value it sees in the reference order, so results are
// bit-identical; the coarse step of the three-level C2b layout takes 8 launches instead of 22.
struct CycleState {
    uint32_t pm;
    int rc;
    std::vector<CycleItem> items;
};

static int markers_of(const Layout *L, const Ff_Store *t3, int b0, int b1, int m[3])
{
    if (b0 >= b1) {
        m[0] = 2;
        m[1] = 1;
        m[2] = 0;
        return 0;
    }
    for (int p = 0; p < 3; p++) m[p] = t3[b0].marker[p];
    for (int b = b0 + 1; b < b1; b++)
        for (int p = 0; p < 3; p++)
            if (t3[b].marker[p] != m[p]) return MBL_E_UNSUPPORTED;
    for (int p = 0; p < 3; p++)
        if (m[p] < 0 || m[p] > 2) return MBL_E_INVALID;
    (void)L;
    return 0;
}

// The markers of every line group as they stand (a group's lines rotate together).
static int marker_set(const Layout *L, const Ff_Store *t3, MarkerSet &ms)
{
    for (int l = 0; l < MBL_MAX_LEVEL; l++) {
        ms.m[l][0] = 2;
        ms.m[l][1] = 1;
        ms.m[l][2] = 0;
    }
    for (int l = 0; l < L->LEVEL; l++)
        if (L->LS3[l + 1] > L->LS3[l]) {
            int rc = markers_of(L, t3, L->LS3[l], L->LS3[l + 1], ms.m[l]);
            if (rc) return rc;
        }
    return 0;
}

static void record_ops(Layout *L, const Batch &B, CycleState &cs, const Ff_Store *t3, int kind, double xm,
                       bool temporal)
{
    if (cs.rc || B.empty()) return;
    MarkerSet ms;
    int rc = marker_set(L, t3, ms);
    if (rc) {
        cs.rc = rc;
        return;
    }
    for (const Sub &sb : B)
        if (sb.count > 0) cs.items.push_back(CycleItem{kind, -1, sb.off, sb.count, xm, temporal, cs.pm, ms});
}

static void cycle(Layout *L, int level, Ff_Store *t3, CycleState &cs)
{
    if (level == L->LEVEL || cs.rc) return; // level checked before mi[level] is read
    for (int i = 0; i < L->mi[level]; i++) {
        if (!(level == 0 || i == 0) && !L->bT[level].empty())
            record_ops(L, L->bT[level], cs, t3, 1, 1.0 * i / L->mi[level], true);
        if (L->tiles[level] > 0) {
            MarkerSet ms;
            marker_set(L, t3, ms);
            cs.items.push_back(CycleItem{0, level, 0, 0, 0.0, false, cs.pm, ms});
        }
        cs.pm ^= (1u << level);
        if (L->number_line2 != 0) record_ops(L, L->bCC[level], cs, t3, 1, 0.0, false);
        if (level != L->LEVEL - 1 && L->LS3[level + 1] > L->LS3[level]) {
            int m[3];
            int rc = markers_of(L, t3, L->LS3[level], L->LS3[level + 1], m);
            if (rc) {
                cs.rc = rc;
                return;
            }
            record_ops(L, L->bS[level], cs, t3, 2, 0.0, false);
            for (int b = L->LS3[level]; b < L->LS3[level + 1]; b++)
                for (int p = 0; p < 3; p++) {
                    int bb = t3[b].marker[p] - 1;
                    t3[b].marker[p] = (bb == -1) ? 2 : bb;
                }
        }
        cycle(L, level + 1, t3, cs);
        if (level != 0 && i == L->mi[level] - 1 && !L->bCF[level].empty())
            record_ops(L, L->bCF[level], cs, t3, 1, 0.0, false);
    }
}

static int launch_state(const Layout *L, const std::vector<CycleItem> &items, const CycleLaunch &cl, uint32_t &pm,
                        MarkerSet &ms);

// Marker rotations (evolution.h:41-46, m -> m - 1 mod 3) that take a to b, or -1.
static int marker_turns(const int a[3], const int b[3])
{
    for (int r = 0; r < 3; r++) {
        bool same = true;
        for (int p = 0; p < 3; p++) same = same && ((a[p] + 2 * r) % 3 == b[p]);
        if (same) return r;
    }
    return -1;
}

// Whether temporal item t can be done by the workgroups of spline item w: every temporal op
// reads the ring of one of w's splines (same fine points), on the one-workgroup spline path,
// at most one temporal op per spline, and the item's markers are w's turned a whole number
// of times (the op's slots are then w's markers turned as often, at any step).
static bool fuse_temporal(const Layout *L, const std::vector<CycleItem> &items, int w, int t)
{
    const CycleItem &S = items[w], &T = items[t];
    if (S.kind != 2 || T.kind != 1 || !T.temporal) return false;
    int mx = 0;
    for (int o = S.off; o < S.off + S.count; o++) mx = std::max(mx, L->hops[o].nc);
    if (launch_spline_path(mx) != 0) return false;
    std::vector<int> used(S.count, 0);
    for (int o = T.off; o < T.off + T.count; o++) {
        const LineOp &op = L->hops[o];
        if (op.type != OP_TEMPORAL) return false;
        int hit = -1;
        for (int q = S.off; q < S.off + S.count; q++)
            if (L->hops[q].ring[0][0] == op.ring[0][0] && L->hops[q].n == op.n && L->hops[q].mgroup == op.mgroup)
                hit = q - S.off;
        if (hit < 0 || used[hit]++) return false;
        if (marker_turns(S.ms.m[op.mgroup], T.ms.m[op.mgroup]) < 0) return false;
    }
    return true;
}

// Dependence-driven positions of the recorded items, then the merged launches.
static int schedule_cycle(mbl_context *h, Layout *L, const std::vector<CycleItem> &items, CycleProgram &P)
{
    const int n = (int)items.size();
    struct Res {
        int writer = -1;
        std::vector<int> readers;
    };
    std::unordered_map<uint64_t, Res> res;
    // resources: node units as the interface ops name them; a level's collide epoch; a line's ring
    auto epoch_key = [](int l) { return unit(0xFFFFFE, (size_t)l, 0); };
    auto ring_key = [](int b) { return unit(0xFFFFFD, (size_t)b, 0); };
    std::vector<int> pos(n, 0);
    std::vector<int> &fused_into = P.fused_into;
    fused_into.assign(n, -1);
    for (int j = 0; j < n; j++) {
        const CycleItem &it = items[j];
        std::vector<uint64_t> rd, wr;
        if (it.kind == 0) {
            wr.push_back(epoch_key(it.level));
        } else {
            for (int o = it.off; o < it.off + it.count; o++) {
                const OpAccess &ac = L->acc[o];
                rd.insert(rd.end(), ac.rd.begin(), ac.rd.end());
                wr.insert(wr.end(), ac.wr.begin(), ac.wr.end());
                for (int l = 0; l < L->LEVEL; l++)
                    if (ac.epochs & (1u << l)) rd.push_back(epoch_key(l));
                if (ac.ring_line >= 0) (ac.ring_write ? wr : rd).push_back(ring_key(ac.ring_line));
            }
        }
        std::sort(rd.begin(), rd.end());
        rd.erase(std::unique(rd.begin(), rd.end()), rd.end());
        std::sort(wr.begin(), wr.end());
        wr.erase(std::unique(wr.begin(), wr.end()), wr.end());
        // p: the earliest position; pr: the part of it due to reading ring slots, and their writer
        int p = 0, pr = 0, rw = -2;
        int why_item = -1; // the item and resource that set p (printed with the schedule)
        uint64_t why_unit = 0;
        char why_kind = '-';
        auto bump = [&](int q, int item, uint64_t u, char kind) {
            if (q > p) {
                p = q;
                why_item = item;
                why_unit = u;
                why_kind = kind;
            }
        };
        for (uint64_t u : rd) {
            auto it2 = res.find(u);
            if (it2 == res.end() || it2->second.writer < 0) continue;
            const int w = it2->second.writer;
            if ((u >> 40) == 0xFFFFFD) {
                pr = std::max(pr, pos[w] + 1);
                rw = (rw == -2 || rw == w) ? w : -1;
            } else {
                bump(pos[w] + 1, w, u, 'R');
            }
        }
        for (uint64_t u : wr) {
            auto it2 = res.find(u);
            if (it2 == res.end()) continue;
            if (it2->second.writer >= 0) bump(pos[it2->second.writer] + 1, it2->second.writer, u, 'W');
            for (int r : it2->second.readers)
                if (r != j) bump(pos[r] + 1, r, u, 'A');
        }
        // a temporal interpolation right after the splines that wrote its lines' newest slots:
        // done inside the spline launch (each spline workgroup interpolates its line's points
        // after storing them), when nothing else keeps it from that position
        if (it.kind == 1 && it.temporal && rw >= 0 && pr > p && p <= pos[rw] && fuse_temporal(L, items, rw, j))
            fused_into[j] = rw;
        pos[j] = fused_into[j] >= 0 ? pos[rw] : std::max(p, pr);
        if (getenv("MBL_PRINT_SCHEDULE"))
            fprintf(stderr, "mbl schedule: item %2d after item %d (%c block %d node %d field %d)%s\n", j, why_item,
                    why_kind, (int)(why_unit >> 40), (int)((why_unit >> 1) & 0x7FFFFFFFFFull), (int)(why_unit & 1),
                    pr > p ? " ring" : "");
        for (uint64_t u : rd) {
            Res &r = res[u];
            if (r.readers.empty() || r.readers.back() != j) r.readers.push_back(j);
        }
        for (uint64_t u : wr) {
            Res &r = res[u];
            r.writer = j;
            r.readers.clear();
        }
    }
    int npos = 0;
    for (int v : pos) npos = std::max(npos, v + 1);
    if (getenv("MBL_PRINT_SCHEDULE")) { // the recorded order with each item's position and its ops
        static const char *TN[] = {"T", "CC", "CF", "FC", "S"};
        for (int j = 0; j < n; j++) {
            fprintf(stderr, "mbl schedule: item %2d pos %2d ", j, pos[j]);
            if (items[j].kind == 0) {
                fprintf(stderr, "collide level %d\n", items[j].level);
                continue;
            }
            for (int o = items[j].off; o < items[j].off + items[j].count; o++)
                fprintf(stderr, "%s(%d->%d) ", TN[L->hops[o].type], L->hops[o].src_block, L->hops[o].dst_block);
            fprintf(stderr, "\n");
        }
    }
    P.launches.clear();
    for (int q = 0; q < npos; q++) {
        // stream-collides: runs of adjacent levels with the same nodes per thread
        std::vector<int> col;
        for (int j = 0; j < n; j++)
            if (pos[j] == q && items[j].kind == 0) col.push_back(j);
        std::sort(col.begin(), col.end(), [&](int a, int b) { return items[a].level < items[b].level; });
        std::vector<CycleLaunch> cls;
        for (size_t c = 0; c < col.size();) {
            CycleLaunch cl;
            cl.kind = 0;
            cl.pos = q;
            cl.lo = cl.hi = items[col[c]].level;
            cl.items.push_back(col[c]);
            size_t d = c + 1;
            while (d < col.size() && items[col[d]].level == cl.hi + 1 && L->vx[items[col[d]].level] == L->vx[cl.lo]) {
                cl.hi++;
                cl.items.push_back(col[d]);
                d++;
            }
            cls.push_back(cl);
            c = d;
        }
        // interface ops: one launch per temporal interpolation point (ops without one join the first)
        std::vector<CycleLaunch> ifl;
        for (int j = 0; j < n; j++) {
            if (pos[j] != q || items[j].kind != 1 || fused_into[j] >= 0) continue;
            CycleLaunch *tgt = nullptr;
            for (auto &x : ifl) {
                bool xt = false;
                double xx = 0.0;
                for (int k : x.items)
                    if (items[k].temporal) {
                        xt = true;
                        xx = items[k].xm;
                    }
                if (!items[j].temporal || !xt || xx == items[j].xm) {
                    tgt = &x;
                    break;
                }
            }
            if (!tgt) {
                ifl.push_back(CycleLaunch());
                tgt = &ifl.back();
                tgt->kind = 1;
                tgt->pos = q;
            }
            tgt->items.push_back(j);
        }
        // splines: the one-workgroup path and the two-kernel path separately
        CycleLaunch sp[2];
        for (int j = 0; j < n; j++) {
            if (pos[j] != q || items[j].kind != 2) continue;
            int mx = 0;
            for (int o = items[j].off; o < items[j].off + items[j].count; o++) mx = std::max(mx, L->hops[o].nc);
            CycleLaunch &t = sp[launch_spline_path(mx)];
            t.kind = 2;
            t.pos = q;
            t.items.push_back(j);
        }
        // one interface launch and one-workgroup splines at one position: one launch whose first
        // workgroups run the splines and the others the interface entries (no dependence among
        // them, or the schedule would have put them at different positions); with forked
        // positions they stay two launches, on two streams
        if (!h->fork && ifl.size() == 1 && !sp[0].items.empty()) {
            CycleLaunch m;
            m.kind = 3;
            m.pos = q;
            m.items = sp[0].items;
            m.items.insert(m.items.end(), ifl[0].items.begin(), ifl[0].items.end());
            ifl.clear();
            sp[0] = m;
        }
        // temporal items done by the workgroups of their splines' launch
        for (int j = 0; j < n; j++)
            if (pos[j] == q && fused_into[j] >= 0)
                for (auto &t : sp)
                    if (std::find(t.items.begin(), t.items.end(), fused_into[j]) != t.items.end()) t.fused.push_back(j);
        // one stream-collide (one node per thread) and one interface launch at one position: one
        // launch whose first workgroups are the collide's tiles and the others the interface entries
        if (!h->fork && cls.size() == 1 && ifl.size() == 1 && sp[0].items.empty() && sp[1].items.empty() &&
            !L->periodic && L->vx[cls[0].lo] == 1) {
            cls[0].kind = 4;
            cls[0].items.insert(cls[0].items.end(), ifl[0].items.begin(), ifl[0].items.end());
            ifl.clear();
        }
        for (auto &x : cls) P.launches.push_back(x);
        for (auto &x : ifl) P.launches.push_back(x);
        for (auto &t : sp)
            if (!t.items.empty()) P.launches.push_back(t);
    }
    // the launch that holds each level's last collide of the cycle (its tiles' state and macros
    // are final for the step: the fused step's residual is summed there)
    for (int l = 0; l < L->LEVEL; l++) {
        int lastj = -1;
        for (int j = 0; j < n; j++)
            if (items[j].kind == 0 && items[j].level == l) lastj = j;
        for (auto &cl : P.launches)
            if ((cl.kind == 0 || cl.kind == 4) && std::find(cl.items.begin(), cl.items.end(), lastj) != cl.items.end())
                cl.rmask |= 1u << l;
    }
    // the program's op table: each interface/spline launch's ops contiguous
    std::vector<LineOp> tab;
    for (auto &cl : P.launches) {
        if (cl.kind == 0) continue;
        cl.op_off = (int)tab.size();
        for (int j : cl.items)
            for (int o = items[j].off; o < items[j].off + items[j].count; o++) {
                const LineOp &op = L->hops[o];
                tab.push_back(op);
                if (cl.kind == 3 && op.type == OP_SPLINE) { // the splines come first (items in that order)
                    cl.nspl++;
                    cl.max_nf = std::max(cl.max_nf, op.n);
                    cl.max_nc = std::max(cl.max_nc, op.nc);
                    continue;
                }
                cl.max_n = std::max(cl.max_n, op.n);
                cl.max_nc = std::max(cl.max_nc, op.nc);
            }
        cl.op_count = (int)tab.size() - cl.op_off;
        // fused temporal ops after the launch's own, each named by its spline (LineOp::fuse_t)
        uint32_t lpm = 0;
        MarkerSet lms;
        if (!cl.fused.empty() && launch_state(L, items, cl, lpm, lms)) return MBL_E_INVALID;
        for (int f : cl.fused) {
            const CycleItem &T = items[f];
            for (int o = T.off; o < T.off + T.count; o++) {
                LineOp t = L->hops[o];
                int si = -1;
                for (int q = cl.op_off; q < cl.op_off + cl.op_count; q++)
                    if (tab[q].type == OP_SPLINE && tab[q].ring[0][0] == t.ring[0][0]) si = q;
                if (si < 0 || tab[si].fuse_t) return MBL_E_INVALID;
                t.trot = marker_turns(lms.m[t.mgroup], T.ms.m[t.mgroup]);
                t.tflip = (int)(((T.pm ^ lpm) >> L->level_of[t.dst_block]) & 1u);
                if (t.trot < 0) return MBL_E_INVALID;
                t.txm = T.xm;
                tab[si].fuse_t = (int)tab.size() - cl.op_off + 1;
                tab.push_back(t);
            }
        }
    }
    // the diagnostics in the last launch: it holds interface ops only, and every node of the
    // list whose macros it writes is a copy (same-level or f_c) of a node whose macros it does
    // not write
    if (P.dlist_m) hipFree(P.dlist_m);
    P.dlist_m = nullptr;
    P.dmerge = false;
    P.last_fw_blocks.clear();
    if (L->step_ok && !h->fork && !P.launches.empty() && P.launches.back().kind == 1 && !L->hlist.empty()) {
        const CycleLaunch &cl = P.launches.back();
        std::unordered_map<uint64_t, std::pair<int, int>> w; // macro unit -> its copy's source (-1: no copy)
        for (int j : cl.items)
            for (int o = items[j].off; o < items[j].off + items[j].count; o++) {
                const LineOp &op = L->hops[o];
                for (size_t e = 0; e < L->hmsk[o].size(); e++) {
                    if (L->hmsk[o][e] & 2) P.last_fw_blocks.push_back(op.dst_block);
                    if (!(L->hmsk[o][e] & 1)) continue;
                    const bool copy = (op.type == OP_CC || op.type == OP_FC) && e < L->hsrc[o].size();
                    w[unit(op.dst_block, (size_t)L->hdst[o][e], 0)] =
                        copy ? std::make_pair(op.src_block, L->hsrc[o][e]) : std::make_pair(-1, -1);
                }
            }
        std::sort(P.last_fw_blocks.begin(), P.last_fw_blocks.end());
        P.last_fw_blocks.erase(std::unique(P.last_fw_blocks.begin(), P.last_fw_blocks.end()), P.last_fw_blocks.end());
        std::vector<int4> ml = L->hlist;
        bool ok = true;
        for (int4 &q : ml) {
            auto it = w.find(unit(q.x, (size_t)q.y, 0));
            if (it == w.end()) continue;
            const std::pair<int, int> src = it->second;
            if (src.first < 0 || w.count(unit(src.first, (size_t)src.second, 0))) {
                ok = false;
                break;
            }
            q.z = src.first;
            q.w = src.second;
        }
        if (ok) {
            if (hipMalloc(&P.dlist_m, sizeof(int4) * ml.size()) != hipSuccess) return set_err(h, MBL_E_NOMEM);
            if (hipMemcpy(P.dlist_m, ml.data(), sizeof(int4) * ml.size(), hipMemcpyHostToDevice) != hipSuccess)
                return set_err(h, MBL_E_HIP);
            P.dmerge = true;
        }
    }
    if (P.dops) hipFree(P.dops);
    P.dops = nullptr;
    if (!tab.empty()) {
        if (hipMalloc(&P.dops, sizeof(LineOp) * tab.size()) != hipSuccess) return set_err(h, MBL_E_NOMEM);
        if (hipMemcpy(P.dops, tab.data(), sizeof(LineOp) * tab.size(), hipMemcpyHostToDevice) != hipSuccess)
            return set_err(h, MBL_E_HIP);
    }
    P.nitems = n;
    P.built = true;
    return 0;
}

// The level parities and markers of a merged launch: each level's bit and each line group's
// markers from the item that involves it (items merged into one launch agree on them).
static int launch_state(const Layout *L, const std::vector<CycleItem> &items, const CycleLaunch &cl, uint32_t &pm,
                        MarkerSet &ms)
{
    uint32_t known = 0, soft = 0;
    bool msk[MBL_MAX_LEVEL] = {false};
    pm = items[cl.items[0]].pm;
    ms = items[cl.items[0]].ms;
    auto take_bit = [&](int l, uint32_t from) -> int {
        const uint32_t b = 1u << l;
        if ((known & b) && ((pm ^ from) & b)) return MBL_E_INVALID;
        pm = (pm & ~b) | (from & b);
        known |= b;
        return 0;
    };
    // a spline writes no plane of its fine level (only macros and its ring): that bit is taken
    // where nothing else in the launch decides it (a temporal op done by the spline's
    // workgroups counts from it)
    auto soft_bit = [&](int l, uint32_t from) {
        const uint32_t b = 1u << l;
        if ((known | soft) & b) return;
        pm = (pm & ~b) | (from & b);
        soft |= b;
    };
    for (int j : cl.items) {
        const CycleItem &it = items[j];
        if (it.kind == 0) {
            if (take_bit(it.level, it.pm)) return MBL_E_INVALID;
            continue;
        }
        for (int o = it.off; o < it.off + it.count; o++) {
            const LineOp &op = L->hops[o];
            if (op.type == OP_SPLINE) {
                if (take_bit(L->level_of[op.src_block], it.pm)) return MBL_E_INVALID;
                soft_bit(L->level_of[op.dst_block], it.pm);
            } else if (take_bit(L->level_of[op.src_block], it.pm) || take_bit(L->level_of[op.dst_block], it.pm)) {
                return MBL_E_INVALID;
            }
            // the marker groups the op reads: its own (temporal, c_f, spline), the c_f ops'
            // of an f_c op's redirected line ends
            int gs[2], ng = 0;
            if (op.type == OP_TEMPORAL || op.type == OP_CF || op.type == OP_SPLINE) gs[ng++] = op.mgroup;
            if (op.type == OP_FC)
                for (int r = 0; r < op.red_n; r++)
                    if (op.red_t[r] == 1) gs[ng++] = op.red_mg[r];
            for (int q = 0; q < ng; q++) {
                const int g = gs[q];
                if (msk[g]) {
                    for (int p = 0; p < 3; p++)
                        if (ms.m[g][p] != it.ms.m[g][p]) return MBL_E_INVALID;
                } else {
                    for (int p = 0; p < 3; p++) ms.m[g][p] = it.ms.m[g][p];
                    msk[g] = true;
                }
            }
        }
    }
    return 0;
}

// fused: the collide launches add the residual sums of the levels whose last collide they hold.
static int launch_one(mbl_context *h, Layout *L, const std::vector<CycleItem> &items, const CycleProgram &P,
                      const CycleLaunch &cl, hipStream_t stream, bool fused)
{
    const int mode = storage_mode(h->cfg);
    uint32_t pm;
    MarkerSet ms;
    if (launch_state(L, items, cl, pm, ms)) return MBL_E_INVALID;
    if (cl.kind == 0) {
        int tiles = 0;
        for (int l = cl.lo; l <= cl.hi; l++) tiles += L->tiles[l];
        const uint32_t rmask = fused ? cl.rmask : 0u;
        return launch_stream_collide(stream, mode, L->vx[cl.lo], h->tpb, L->periodic, true, true, L->db, L->dtile,
                                     L->nblocks, tiles, L->rows, pm, h->groups, L->tile_off[cl.lo], rmask,
                                     rmask ? L->dres : nullptr);
    }
    if (cl.kind == 4) {
        CollideOpsArgs a;
        memset(&a, 0, sizeof(a));
        for (int l = cl.lo; l <= cl.hi; l++) a.tiles += L->tiles[l];
        for (int j : cl.items)
            if (items[j].temporal) a.xm = items[j].xm;
        a.descs = L->db;
        a.tile_blk = L->dtile;
        a.ndesc = L->nblocks;
        a.rows = L->rows;
        a.pm = pm;
        a.tile_base = L->tile_off[cl.lo];
        a.rmask = fused ? cl.rmask : 0u;
        a.rparts = a.rmask ? L->dres : nullptr;
        a.ops = P.dops + cl.op_off;
        a.nops = cl.op_count;
        a.ms = ms;
        return launch_collide_ops(stream, mode, h->tpb, a, cl.max_n);
    }
    if (cl.kind == 1 || cl.kind == 3) {
        double xm = 0.0;
        for (int j : cl.items)
            if (items[j].temporal) xm = items[j].xm;
        if (cl.kind == 3)
            return launch_iface_mixed(stream, mode, P.dops + cl.op_off, cl.nspl, cl.op_count - cl.nspl, cl.max_nc,
                                      cl.max_nf, cl.max_n, L->db, pm, ms, xm);
        return launch_line_ops(stream, mode, P.dops + cl.op_off, cl.op_count, cl.max_n, L->db, pm, ms, xm);
    }
    return launch_spline(stream, mode, P.dops + cl.op_off, cl.op_count, cl.max_nc, cl.max_n, L->db, pm, ms);
}

static int side_streams(mbl_context *h)
{
    if (h->ev_fork) return 0;
    for (int i = 0; i < 3; i++) {
        if (hipStreamCreateWithFlags(&h->side[i], hipStreamNonBlocking) != hipSuccess) return MBL_E_HIP;
        if (hipEventCreateWithFlags(&h->ev_join[i], hipEventDisableTiming) != hipSuccess) return MBL_E_HIP;
    }
    if (hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess) return MBL_E_HIP;
    return 0;
}

// The scheduled launches in order. Launches at one position of the schedule have no
// dependence among them: the first runs on the handle's stream, the others on side streams
// forked from it and joined back before the next position (in a captured graph: parallel
// branches).
// mdiag: the step's diagnostics, launched with the last launch (CycleProgram::dmerge).
static int run_cycle(mbl_context *h, Layout *L, const std::vector<CycleItem> &items, const CycleProgram &P,
                     bool fused, const StepDiagArgs *mdiag = nullptr)
{
    if ((int)items.size() != P.nitems) return MBL_E_INVALID;
    const size_t nl = P.launches.size();
    for (size_t i = 0; i < nl;) {
        size_t j = i + 1;
        while (j < nl && P.launches[j].pos == P.launches[i].pos) j++;
        const int k = h->fork ? (int)(j - i) : 1;
        if (k > 1) {
            if (int rc = side_streams(h)) return rc;
            if (hipEventRecord(h->ev_fork, h->stream) != hipSuccess) return MBL_E_HIP;
            for (int s = 0; s < std::min(k - 1, 3); s++)
                if (hipStreamWaitEvent(h->side[s], h->ev_fork, 0) != hipSuccess) return MBL_E_HIP;
        }
        for (size_t m = i; m < j; m++) {
            const int r = (int)(m - i);
            hipStream_t st = (k > 1 && r > 0) ? h->side[(r - 1) % 3] : h->stream;
            const CycleLaunch &cl = P.launches[m];
            if (mdiag && m == nl - 1) {
                if (cl.kind != 1 || k > 1) return MBL_E_INVALID;
                uint32_t pm;
                MarkerSet ms;
                if (launch_state(L, items, cl, pm, ms)) return MBL_E_INVALID;
                double xm = 0.0;
                for (int q : cl.items)
                    if (items[q].temporal) xm = items[q].xm;
                if (int rc = launch_ops_diag(st, storage_mode(h->cfg), P.dops + cl.op_off, cl.op_count, cl.max_n, L->db,
                                             pm, ms, xm, *mdiag))
                    return rc;
                continue;
            }
            if (int rc = launch_one(h, L, items, P, cl, st, fused)) return rc;
        }
        if (k > 1)
            for (int s = 0; s < std::min(k - 1, 3); s++) {
                if (hipEventRecord(h->ev_join[s], h->side[s]) != hipSuccess) return MBL_E_HIP;
                if (hipStreamWaitEvent(h->stream, h->ev_join[s], 0) != hipSuccess) return MBL_E_HIP;
            }
        i = j;
    }
    return 0;
}

static void evolution_impl(mbl_context *h, const EvoArgs &A, int level, int nblocks, int nline, int nline2,
                           bool fused = false, StepDiagArgs *diag = nullptr, int drag_block = -1);

extern "C" void evolution(Lbi n, Lbi *asp, Lbi *ls3cffcp, Lbi *ltcffcp, Lbi *lccp, Lbi level, mbl_dim3 block,
                          mbl_dim3 *grid, AREA *areas, D2Q9 *pps_h, D2Q9 *pps_d, Inf_Ff_store *line_Inf_c,
                          Inf_Ff_store *line_Inf_f, Inf_Ff_store *line_Inf_cc, Ff_Store *line_t3,
                          FM_spline3 *line_S3, mbl_handle_t handle)
{
    (void)n;
    (void)block;
    (void)grid;
    (void)pps_h;
    mbl_context *h = handle ? handle : g_current;
    if (!h) return;
    if (h->last_error) return;
    const mbl_config &cfg = h->cfg;
    if (!asp || !ls3cffcp || !ltcffcp || !lccp || !areas || !pps_d || level < 0 || level > cfg.LEVEL) {
        set_err(h, MBL_E_INVALID);
        return;
    }
    const int nblocks = asp[cfg.LEVEL];
    const int nline = ltcffcp[cfg.LEVEL];
    const int nline2 = lccp[cfg.LEVEL];
    if (nline != ls3cffcp[cfg.LEVEL] || (nline > 0 && (!line_Inf_c || !line_Inf_f || !line_t3 || !line_S3)) ||
        (nline2 > 0 && !line_Inf_cc)) {
        set_err(h, MBL_E_INVALID);
        return;
    }
    EvoArgs A{asp, ls3cffcp, ltcffcp, lccp, areas, pps_d, line_Inf_c, line_Inf_f, line_Inf_cc, line_t3, line_S3};
    evolution_impl(h, A, level, nblocks, nline, nline2);
}

// The layout of a multi-block call: the handle's, rebuilt when the caller's tables changed.
static Layout *layout_for(mbl_context *h, const EvoArgs &A, int nblocks, int nline, int nline2)
{
    const uint64_t sig = layout_sig(h, A, nblocks, nline, nline2);
    if (!h->layout || h->layout->sig != sig) {
        if (h->layout) {
            hipStreamSynchronize(h->stream);
            h->layout->release();
            delete h->layout;
            h->layout = nullptr;
        }
        Layout *L = nullptr;
        if (build_layout(h, A, nblocks, nline, nline2, sig, L)) return nullptr;
        h->layout = L;
        h->layout_gen++;
    }
    return h->layout;
}

// The parity of every level (which plane set holds its state, from the caller's F pointers)
// and the phase key: parity bits, ring markers of every line group, starting level.
static int phase_key(const Layout *L, const D2Q9 *pps_d, const Ff_Store *line_t3, int level, uint32_t &pm,
                     uint64_t &key)
{
    pm = 0;
    for (int l = 0; l < L->LEVEL; l++) {
        int par = -1;
        for (int a = L->ASP[l]; a < L->ASP[l + 1]; a++) {
            int p = (pps_d[a].F[0] == (Lbd *)L->hb[a].P[0][0]) ? 0 : ((pps_d[a].F[0] == (Lbd *)L->hb[a].P[1][0]) ? 1 : -1);
            if (p < 0 || (par >= 0 && p != par)) return MBL_E_INVALID;
            par = p;
        }
        if (par == 1) pm |= (1u << l);
    }
    key = pm;
    int sh = 8;
    for (int l = 0; l < L->LEVEL; l++) {
        int m[3];
        if (L->LS3[l + 1] > L->LS3[l]) {
            if (markers_of(L, line_t3, L->LS3[l], L->LS3[l + 1], m)) return MBL_E_UNSUPPORTED;
            key |= (uint64_t)(m[0] | (m[1] << 2) | (m[2] << 4)) << sh;
        }
        sh += 6;
    }
    key |= (uint64_t)level << 60;
    return 0;
}

// fused (level 0 of a layout with step_ok): the collides also write the residual sums
// (mbl_evolution_step); its graphs are kept apart from the plain cycle's.
// Every argument of a diagnostics launch except the ones the phase sets (slot counters, the
// drag's plane set).
static uint64_t diag_signature(const StepDiagArgs &d, int drag_block)
{
    uint64_t s = 0xD1A6;
    s = mix64(s, (uintptr_t)d.descs);
    s = mix64(s, (uintptr_t)d.list);
    s = mix64(s, ((uint64_t)(uint32_t)d.nlist << 32) | (uint32_t)d.list_wgs);
    s = mix64(s, ((uint64_t)(uint32_t)d.ntiles << 32) | (uint32_t)d.cap);
    s = mix64(s, (uintptr_t)d.rparts);
    s = mix64(s, (uintptr_t)d.lbase);
    s = mix64(s, (uint64_t)d.lstride);
    s = mix64(s, (uintptr_t)d.drag_base);
    s = mix64(s, (uint64_t)(int64_t)drag_block);
    s = mix64(s, (uintptr_t)d.drag.links);
    s = mix64(s, ((uint64_t)(uint32_t)d.drag.n << 32) | (uint32_t)d.drag.nx);
    return s;
}

// One call of evolution() planned on the host: its phase key, the recorded items (the ring
// markers rotate in the caller's tables as evolution.h:41-46 does), the diagnostics launch
// that follows it (fused step), and the caller's F/f handed over for the levels that took an
// odd number of substeps. The launches are issued later (launch_plans), one graph for one or
// several consecutive plans.
struct StepPlan {
    uint64_t key = 0;
    std::vector<CycleItem> items;
    bool diag = false;
    bool merge = false; // the diagnostics go with the cycle's last launch (CycleProgram::dmerge)
    StepDiagArgs d;
};

static void drop_fused_graphs(Layout *L)
{
    for (auto it = L->graphs.begin(); it != L->graphs.end();) {
        if (it->first & (1ull << 57)) {
            hipGraphExecDestroy(it->second);
            it = L->graphs.erase(it);
        } else {
            ++it;
        }
    }
    for (auto &g : L->chunks) hipGraphExecDestroy(g.second);
    L->chunks.clear();
}

// diag (fused only): the step's diagnostics launch, whose record slot is want_slot (its slot
// counter is set to it here when it does not already hold it); its drag reads block
// drag_block's older state (the caller's f after the call).
static int plan_step(mbl_context *h, Layout *L, const EvoArgs &A, int level, int nline, bool fused,
                     const StepDiagArgs *diag, int drag_block, int want_slot, StepPlan &sp)
{
    D2Q9 *pps_d = A.pps;
    Ff_Store *line_t3 = A.t3;
    uint32_t pm = 0;
    uint64_t key = 0;
    if (int rc = phase_key(L, pps_d, line_t3, level, pm, key)) return rc;
    // the cycle in the reference order (markers rotate here)
    std::vector<Ff_Store> saved(line_t3, line_t3 + nline);
    CycleState cs{pm, 0, {}};
    cycle(L, level, line_t3, cs);
    auto fail = [&](int code) {
        std::copy(saved.begin(), saved.end(), line_t3);
        return code;
    };
    if (cs.rc) return fail(cs.rc);
    CycleProgram &P = L->prog[level];
    if (!P.built && schedule_cycle(h, L, cs.items, P)) return fail(h->last_error ? h->last_error : MBL_E_HIP);
    if (fused) key |= 1ull << 57;
    sp.key = key;
    sp.diag = fused && diag;
    if (sp.diag) {
        sp.d = *diag;
        // the slot counter this step reads holds the record's index
        const int p = (int)(pm & 1u);
        sp.d.slot_rd = h->d_slots + p;
        sp.d.slot_wr = h->d_slots + (1 - p);
        if (h->slot_host[p] != want_slot) {
            if (hipMemsetD32Async((hipDeviceptr_t)(h->d_slots + p), want_slot, 1, h->stream) != hipSuccess)
                return fail(MBL_E_HIP);
            h->slot_host[p] = want_slot;
        }
        h->slot_host[1 - p] = want_slot + 1;
        if (sp.d.drag.n > 0) {
            if (drag_block < 0 || drag_block >= L->nblocks) return fail(MBL_E_INVALID);
            // after the cycle the block's state is in plane set cs.pm's bit, the older one in the other
            const int cur = (int)((cs.pm >> L->level_of[drag_block]) & 1u);
            for (int k = 0; k < MBL_Q; k++) sp.d.drag.Fold[k] = L->hb[drag_block].P[1 - cur][k];
        }
        // with the last launch when its ops write no plane the drag reads (the drag block's older set)
        sp.merge = P.dmerge && P.dlist_m;
        if (sp.merge && sp.d.drag.n > 0 &&
            std::binary_search(P.last_fw_blocks.begin(), P.last_fw_blocks.end(), drag_block)) {
            uint32_t lpm = 0;
            MarkerSet lms;
            if (launch_state(L, cs.items, P.launches.back(), lpm, lms)) return fail(MBL_E_INVALID);
            const int cur = (int)((cs.pm >> L->level_of[drag_block]) & 1u);
            if ((int)((lpm >> L->level_of[drag_block]) & 1u) == 1 - cur) sp.merge = false;
        }
        if (sp.merge) sp.d.list = P.dlist_m;
        const uint64_t ds = diag_signature(sp.d, drag_block) ^ (sp.merge ? 0x5A5Aull : 0ull);
        if (L->diag_sig != ds) { // captured with other diagnostics arguments: capture again
            if (hipStreamSynchronize(h->stream) != hipSuccess) return fail(MBL_E_HIP);
            drop_fused_graphs(L);
            L->diag_sig = ds;
        }
    }
    sp.items = std::move(cs.items);
    // hand the new state back: F <-> f for every level that took an odd number of substeps
    for (int l = level; l < h->cfg.LEVEL; l++) {
        if (!((cs.pm ^ pm) & (1u << l))) continue;
        for (int a = L->ASP[l]; a < L->ASP[l + 1]; a++)
            for (int k = 0; k < 9; k++) std::swap(pps_d[a].F[k], pps_d[a].f[k]);
    }
    return 0;
}

// The launches of consecutive plans (each its cycle, then its diagnostics launch): one
// captured graph per phase key of the first plan and number of plans, replayed after that.
static int launch_plans(mbl_context *h, Layout *L, int level, const std::vector<StepPlan> &plans)
{
    if (plans.empty()) return 0;
    const CycleProgram &P = L->prog[level];
    const int mode = storage_mode(h->cfg);
    const bool fused = (plans[0].key >> 57) & 1u;
    auto issue = [&]() -> int {
        for (const StepPlan &sp : plans) {
            if (int rc = run_cycle(h, L, sp.items, P, fused, sp.diag && sp.merge ? &sp.d : nullptr)) return rc;
            if (sp.diag && !sp.merge)
                if (int rc = launch_step_diag(h->stream, mode, sp.d)) return rc;
        }
        return 0;
    };
    if (!h->cfg.use_graph) return issue();
    const int n = (int)plans.size();
    hipGraphExec_t ge = nullptr;
    if (n == 1) {
        auto it = L->graphs.find(plans[0].key);
        if (it != L->graphs.end()) ge = it->second;
    } else {
        auto it = L->chunks.find({plans[0].key, n});
        if (it != L->chunks.end()) ge = it->second;
    }
    if (!ge) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return MBL_E_HIP;
        int rc = issue();
        hipError_t e = hipStreamEndCapture(h->stream, &g);
        if (rc || e != hipSuccess || hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return rc ? rc : MBL_E_HIP;
        }
        hipGraphDestroy(g);
        if (n == 1)
            L->graphs[plans[0].key] = ge;
        else
            L->chunks[{plans[0].key, n}] = ge;
    }
    return hipGraphLaunch(ge, h->stream) == hipSuccess ? 0 : MBL_E_HIP;
}

// diag (fused only): the step's diagnostics launch after the cycle, in the same graph; its drag
// reads block drag_block's older state (the caller's f after the call).
static void evolution_impl(mbl_context *h, const EvoArgs &A, int level, int nblocks, int nline, int nline2,
                           bool fused, StepDiagArgs *diag, int drag_block)
{
    Layout *L = layout_for(h, A, nblocks, nline, nline2);
    if (!L) return;
    std::vector<StepPlan> plans(1);
    if (int rc = plan_step(h, L, A, level, nline, fused, diag, drag_block, (int)h->pend_l2.size(), plans[0])) {
        set_err(h, rc);
        return;
    }
    if (int rc = launch_plans(h, L, level, plans)) set_err(h, rc);
}

// ---------------------------------------------------------------------------
// single-block stepping (bandwidth benchmark path)
// ---------------------------------------------------------------------------
extern "C" int mbl_step_block(mbl_handle_t h, const AREA *area, D2Q9 *pp, int nsteps)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!area || !pp || nsteps < 0) return set_err(h, MBL_E_INVALID);
    Single &S = h->single;
    uint64_t sig = 0x77;
    for (int k = 0; k < 9; k++) {
        uintptr_t x = (uintptr_t)pp->F[k], y = (uintptr_t)pp->f[k];
        sig = mix64(sig, std::min(x, y));
        sig = mix64(sig, std::max(x, y));
    }
    sig = mix64(sig, (uint64_t)area->nx | ((uint64_t)area->ny << 32));
    uint64_t tb;
    memcpy(&tb, &area->rv_tau, 8);
    sig = mix64(sig, tb);
    sig = mix64(sig, (uintptr_t)pp->solid);
    sig = mix64(sig, (uintptr_t)pp->rho);
    if (S.sig != sig || !S.dd) {
        release_single(h);
        const bool periodic = (h->cfg.bc_mode == MBL_BC_PERIODIC);
        if (!plane_fits(h, area->nx, area->ny)) return set_err(h, MBL_E_UNSUPPORTED);
        bool any_solid = true;
        if (validate_flags(h, *area, pp->solid, periodic, &any_solid)) return MBL_E_INVALID;
        S.hd = make_desc(h->cfg, *area, *pp, 0, nullptr, nullptr);
        S.vx = (h->vx_set || periodic) ? h->vx : 1;
        if (periodic)
            while (S.vx > 1 && area->nx % S.vx) S.vx >>= 1;
        S.tpb = h->tpb;
        // rows per wave: the 1-4-11 and 1-3-12 planes are VALU-bound, and four rows per wave
        // share one tile's index arithmetic (16384^2 on MI355X: 127k -> 139k/140k MLUPS); the
        // other storages are bandwidth-bound and lose parallelism with more (measured slower)
        const int st = storage_mode(h->cfg);
        S.rows = S.tpb / 64 * ((st == ST_F16C || st == ST_F16D) ? 4 : 1);
        S.hd.tw = 64;
        S.hd.trows = 0;
        set_tiles_x(S.hd, seg_tiles(area->nx, S.vx, 64));
        S.hd.tile_begin = 0;
        S.tiles = S.hd.tiles_x * ((area->ny + S.rows - 1) / S.rows);
        S.periodic = periodic;
        S.flags = !(periodic && !any_solid);
        HIPCHK(h, hipMalloc(&S.dd, sizeof(BlockDesc)));
        HIPCHK(h, hipMemcpy(S.dd, &S.hd, sizeof(BlockDesc), hipMemcpyHostToDevice));
        S.sig = sig;
    }
    int par = (pp->F[0] == (Lbd *)S.hd.P[0][0]) ? 0 : 1;
    const int mode = storage_mode(h->cfg);
    const bool macros = !S.periodic || S.flags;
    // periodic all-fluid blocks: two steps per launch where the shape allows (the state at
    // step n + 2 lands in the other plane set, which then holds it as one step would have
    // left it; the plane set handed back as f holds step n, not n + 1)
    // step_pairs 3: three steps per launch, a remainder of two as one paired launch
    const bool pairs = h->step_pairs && S.periodic && !S.flags &&
                       periodic2_fits(mode, area->nx, area->ny, h->step_pairs == 3 ? 1 : h->step_pairs);
    const bool triples = h->step_pairs == 3 && S.periodic && !S.flags && periodic3_fits(mode, area->nx, area->ny);
    for (int t = 0; t < nsteps;) {
        int rc;
        if (triples && t + 3 <= nsteps) {
            Periodic2Args a;
            memset(&a, 0, sizeof(a));
            for (int k = 0; k < 9; k++) {
                a.S[k] = S.hd.P[par][k];
                a.D[k] = S.hd.P[par ^ 1][k];
            }
            a.nx = area->nx;
            a.ny = area->ny;
            a.form = 1;
            a.omega = S.hd.omega;
            rc = launch_periodic3(h->stream, mode, a);
            t += 3;
        } else if (pairs && t + 2 <= nsteps) {
            Periodic2Args a;
            memset(&a, 0, sizeof(a));
            for (int k = 0; k < 9; k++) {
                a.S[k] = S.hd.P[par][k];
                a.D[k] = S.hd.P[par ^ 1][k];
            }
            a.nx = area->nx;
            a.ny = area->ny;
            a.form = h->step_pairs == 3 ? 1 : h->step_pairs;
            a.omega = S.hd.omega;
            rc = launch_periodic2(h->stream, mode, a);
            t += 2;
        } else {
            rc = launch_stream_collide(h->stream, mode, S.vx, S.tpb, S.periodic, S.flags, macros, S.dd, nullptr, 1,
                                       S.tiles, S.rows, (uint32_t)par, h->groups, 0);
            t += 1;
        }
        if (rc) return set_err(h, rc);
        par ^= 1;
        for (int k = 0; k < 9; k++) std::swap(pp->F[k], pp->f[k]);
    }
    return 0;
}

extern "C" int mbl_step_block_launches(mbl_handle_t h, const AREA *area, int nsteps)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!area || nsteps < 0) return MBL_E_INVALID;
    const Single &S = h->single;
    const int mode = storage_mode(h->cfg);
    const bool pairs = S.dd && h->step_pairs && S.periodic && !S.flags &&
                       periodic2_fits(mode, area->nx, area->ny, h->step_pairs == 3 ? 1 : h->step_pairs);
    const bool triples = S.dd && h->step_pairs == 3 && S.periodic && !S.flags && periodic3_fits(mode, area->nx, area->ny);
    if (triples) {
        const int r = nsteps % 3;
        return nsteps / 3 + (r == 2 && pairs ? 1 : r);
    }
    return pairs ? nsteps / 2 + nsteps % 2 : nsteps;
}

extern "C" int mbl_set_step_pairs(mbl_handle_t h, int on)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (on < 0 || on > 3) return MBL_E_INVALID;
    h->step_pairs = on;
    return 0;
}

extern "C" int mbl_block_moments(mbl_handle_t h, const AREA *area, const D2Q9 *pp, double out[4])
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!area || !pp || !out || area->nx <= 0 || area->ny <= 0) return set_err(h, MBL_E_INVALID);
    for (int k = 0; k < 9; k++)
        if (!pp->F[k]) return set_err(h, MBL_E_INVALID);
    const size_t n = (size_t)area->nx * area->ny;
    const int nparts = (int)std::min<size_t>(2048, (n + 255) / 256);
    void **dplanes = nullptr;
    double *buf = nullptr;
    void *hp[9];
    for (int k = 0; k < 9; k++) hp[k] = (void *)pp->F[k];
    int rc = 0;
    if (hipMalloc(&dplanes, sizeof(hp)) != hipSuccess || hipMalloc(&buf, sizeof(double) * (4 * (size_t)nparts + 4)) != hipSuccess)
        rc = MBL_E_NOMEM;
    if (!rc && hipMemcpyAsync(dplanes, hp, sizeof(hp), hipMemcpyHostToDevice, h->stream) != hipSuccess) rc = MBL_E_HIP;
    if (!rc) rc = launch_moments(h->stream, storage_mode(h->cfg), dplanes, n, nparts, buf, buf + 4 * (size_t)nparts);
    if (!rc && (hipMemcpyAsync(out, buf + 4 * (size_t)nparts, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess))
        rc = MBL_E_HIP;
    if (dplanes) hipFree(dplanes);
    if (buf) hipFree(buf);
    return rc ? set_err(h, rc) : 0;
}

// Times `reps` launches of a copy between two fresh buffers; the copy is
// `planes` planes of `plane_bytes` each (1: one stream), GB/s of bytes read + written.
static int copy_probe(mbl_context *h, size_t plane_bytes, int planes, int reps, double *gbs)
{
    const size_t bytes = plane_bytes * (size_t)planes;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return set_err(h, MBL_E_NOMEM);
    if (hipMalloc(&b, bytes) != hipSuccess) {
        hipFree(a);
        return set_err(h, MBL_E_NOMEM);
    }
    auto launch = [&]() {
        return planes == 1 ? launch_stream_copy(h->stream, a, b, plane_bytes)
                           : launch_plane_copy(h->stream, a, b, plane_bytes);
    };
    int rc = hipMemsetAsync(a, 1, bytes, h->stream) == hipSuccess ? 0 : MBL_E_HIP;
    for (int w = 0; w < 2 && !rc; w++) rc = launch(); // warm-up
    if (!rc && hipEventRecord(h->ev0, h->stream) != hipSuccess) rc = MBL_E_HIP;
    for (int r = 0; r < reps && !rc; r++) rc = launch();
    if (!rc && hipEventRecord(h->ev1, h->stream) != hipSuccess) rc = MBL_E_HIP;
    float ms = 0.f;
    if (!rc && (hipEventSynchronize(h->ev1) != hipSuccess || hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess))
        rc = MBL_E_HIP;
    hipStreamSynchronize(h->stream);
    hipFree(a);
    hipFree(b);
    if (!rc) *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return set_err(h, rc);
}

extern "C" int mbl_bandwidth_probe(mbl_handle_t h, size_t bytes, int reps, double *gbs)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!gbs || reps <= 0 || bytes < 16) return set_err(h, MBL_E_INVALID);
    return copy_probe(h, bytes & ~(size_t)15, 1, reps, gbs);
}

extern "C" int mbl_bandwidth_probe_planes(mbl_handle_t h, size_t plane_bytes, int reps, double *gbs)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!gbs || reps <= 0 || plane_bytes < 16) return set_err(h, MBL_E_INVALID);
    return copy_probe(h, plane_bytes & ~(size_t)15, MBL_Q, reps, gbs);
}

extern "C" int mbl_timer_begin(mbl_handle_t h)
{
    if (!h) return MBL_E_NO_HANDLE;
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    h->timing = true;
    return 0;
}

extern "C" int mbl_timer_end(mbl_handle_t h, double *ms)
{
    if (!h) return MBL_E_NO_HANDLE;
    if (!h->timing) return set_err(h, MBL_E_INVALID);
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    HIPCHK(h, hipEventSynchronize(h->ev1));
    float f = 0.f;
    HIPCHK(h, hipEventElapsedTime(&f, h->ev0, h->ev1));
    h->timing = false;
    if (ms) *ms = f;
    return 0;
}

// ---------------------------------------------------------------------------
// initialisation entry points
// ---------------------------------------------------------------------------
extern "C" void init_General(mbl_dim3 grid, mbl_dim3 block, AREA area, D2Q9 pp, Lbd *upx, Lbd *upy)
{
    (void)grid;
    (void)block;
    mbl_context *h = g_current;
    if (!h) return;
    if (area.nx <= 0 || area.ny <= 0 || area.m <= 0 || !pp.solid || !pp.rho) {
        set_err(h, MBL_E_INVALID);
        return;
    }
    const mbl_config &c = h->cfg;
    BlockDesc d = make_desc(h->cfg, area, pp, 0, upx, upy);
    int rc = launch_init_general(h->stream, storage_mode(c), d, area, pp.x, pp.y, pp.forcex, pp.forcey, c.bc_mode,
                                 c.init_mode, c.seed, c.sigma, (double)c.LX, (double)c.LY, c.R, c.CX, c.CY,
                                 make_consts(c));
    if (rc) set_err(h, rc);
}

static int one_op(mbl_context *h, const LineOp &op, const BlockDesc &bd, bool ring_fill)
{
    LineOp *dop = nullptr;
    BlockDesc *dbd = nullptr;
    HIPCHK(h, hipMalloc(&dop, sizeof(LineOp)));
    HIPCHK(h, hipMalloc(&dbd, sizeof(BlockDesc)));
    HIPCHK(h, hipMemcpy(dop, &op, sizeof(LineOp), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(dbd, &bd, sizeof(BlockDesc), hipMemcpyHostToDevice));
    int rc = 0;
    if (ring_fill) rc = launch_ring_fill(h->stream, storage_mode(h->cfg), dop, 1, op.n, dbd, 0u);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    hipFree(dop);
    hipFree(dbd);
    return set_err(h, rc);
}

extern "C" void init_Fstore(mbl_dim3 grid, mbl_dim3 block, AREA area, D2Q9 pp, Inf_Ff_store line_Inf,
                            Ff_Store line_t3)
{
    (void)grid;
    (void)block;
    mbl_context *h = g_current;
    if (!h) return;
    LineOp op;
    memset(&op, 0, sizeof(op));
    op.n = line_t3.N;
    op.dst_block = 0;
    const F_k3 *ft = &line_t3.f_t1;
    for (int sl = 0; sl < 3; sl++)
        for (int k = 0; k < 9; k++) op.ring[sl][k] = ft[sl].F[k];
    std::vector<int> idx(op.n);
    for (int ip = 0; ip < op.n; ip++) {
        int i, j;
        line_node(line_Inf.line_F, ip, i, j);
        if (!in_block(area, i, j)) {
            set_err(h, MBL_E_INVALID);
            return;
        }
        idx[ip] = j * area.nx + i;
    }
    int *didx = nullptr;
    if (hipMalloc(&didx, sizeof(int) * std::max(1, op.n)) != hipSuccess ||
        hipMemcpy(didx, idx.data(), sizeof(int) * op.n, hipMemcpyHostToDevice) != hipSuccess) {
        set_err(h, MBL_E_HIP);
        return;
    }
    op.dst_idx = didx;
    BlockDesc bd = make_desc(h->cfg, area, pp, 0, nullptr, nullptr);
    one_op(h, op, bd, true);
    hipFree(didx);
}

extern "C" void init_tri_A(mbl_dim3 grid, mbl_dim3 block, AREA area, D2Q9 pp, Inf_Ff_store line_Inf,
                           FM_spline3 line_S3)
{
    (void)grid;
    (void)block;
    (void)area;
    (void)pp;
    (void)line_Inf;
    mbl_context *h = g_current;
    if (!h) return;
    int rc = launch_tri_A(h->stream, line_S3.l, line_S3.m, line_S3.u, line_S3.N);
    if (rc) set_err(h, rc);
}

// ---------------------------------------------------------------------------
// geometry of the momentum-exchange links (force_block.h:2-127)
// ---------------------------------------------------------------------------
static int host_links(mbl_context *h, const AREA &a, const int *d_solid, int *host_solid, std::vector<Int3> *out)
{
    const size_t n = (size_t)a.nx * a.ny;
    std::vector<int> tmp;
    int *f = host_solid;
    if (!f) {
        tmp.resize(n);
        f = tmp.data();
    }
    HIPCHK(h, hipStreamSynchronize(h->stream)); // init_General may still be writing the flags
    HIPCHK(h, hipMemcpy(f, d_solid, n * sizeof(int), hipMemcpyDeviceToHost));
    int count = 0;
    for (int i = 0; i < a.nx; i++)
        for (int j = 0; j < a.ny; j++) {
            if (f[i + (size_t)j * a.nx] != SOLID_SBB) continue;
            for (int k = 0; k < 9; k++) {
                int id = i - EXh[k], jd = j - EYh[k];
                if (id < 0 || id > a.nx - 1 || jd < 0 || jd > a.ny - 1) continue;
                if (f[id + (size_t)jd * a.nx] == FLUID) {
                    if (out) out->push_back(Int3{i, j, k});
                    count++;
                }
            }
        }
    return count;
}

extern "C" Lbi sum_N_drag(mbl_dim3 block, mbl_dim3 grid, AREA area, D2Q9 pp)
{
    (void)block;
    (void)grid;
    mbl_context *h = g_current;
    if (!h) return 0;
    int n = host_links(h, area, pp.solid, nullptr, nullptr);
    return n < 0 ? 0 : n;
}

extern "C" void fdrag_store(AREA area, D2Q9 h_pp, Int3 *d_pdrag, D2Q9 d_pp, Lbi N)
{
    mbl_context *h = g_current;
    if (!h) return;
    std::vector<Int3> links;
    int n = host_links(h, area, d_pp.solid, h_pp.solid, &links);
    if (n < 0) return;
    if (n != N) { // force_block.h:122 assert(ii == N)
        set_err(h, MBL_E_INVALID);
        return;
    }
    if (N > 0 && hipMemcpy(d_pdrag, links.data(), sizeof(Int3) * N, hipMemcpyHostToDevice) != hipSuccess)
        set_err(h, MBL_E_HIP);
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
static std::string out_path(const mbl_context *h, const char *name)
{
    std::string d = h->cfg.out_dir;
    if (d.empty()) return name;
    if (d.back() != '/') d += '/';
    return d + name;
}

extern "C" int mbl_flush_diagnostics(mbl_handle_t h)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    const size_t nl = h->pend_l2.size(), nd = h->pend_drag.size();
    if (!nl && !nd) return 0;
    // the deferred final sums, one launch per run of consecutive slots with the same partial count
    for (size_t i = 0; i < h->pend_final.size();) {
        size_t j = i + 1;
        while (j < h->pend_final.size() && h->pend_final[j].first == h->pend_final[j - 1].first + 1 &&
               h->pend_final[j].second == h->pend_final[i].second)
            j++;
        const int s0 = h->pend_final[i].first;
        const int rc = launch_residual_steps(h->stream, (int)(j - i), h->pend_final[i].second, (size_t)h->lparts_stride,
                                             h->d_lparts + 3 * (size_t)h->lparts_stride * s0, h->d_l2 + L2_REC * s0);
        if (rc) {
            h->pend_final.clear();
            return set_err(h, rc);
        }
        i = j;
    }
    h->pend_final.clear();
    if (nl) HIPCHK(h, hipMemcpyAsync(h->h_l2, h->d_l2, sizeof(double) * L2_REC * nl, hipMemcpyDeviceToHost, h->stream));
    if (nd) HIPCHK(h, hipMemcpyAsync(h->h_drag, h->d_drag, sizeof(double) * 2 * nd, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const double U0_2 = std::sqrt(h->cfg.U0[0] * h->cfg.U0[0] + h->cfg.U0[1] * h->cfg.U0[1]);
    FILE *fl = nullptr, *fd = nullptr;
    if (h->cfg.write_files) {
        if (nl) fl = fopen(out_path(h, "error_L2.dat").c_str(), "a");
        if (nd) fd = fopen(out_path(h, "fdrag_t.dat").c_str(), "a");
    }
    int rc = 0;
    for (size_t i = 0; i < nl; i++) {
        const double *rec = h->h_l2 + L2_REC * i;
        double v = rec[0];
        if (fl) fprintf(fl, "%d\t%g\n", h->pend_l2[i], v); // error_L2.h:87
        h->ser_l2_t.push_back(h->pend_l2[i]);
        h->ser_l2.push_back(v);
        h->last[0] = v;
        // a non-finite sum means a non-finite velocity somewhere; 0/0 of an all-rest field is not an error
        if (!std::isfinite(rec[1]) || !std::isfinite(rec[2])) rc = MBL_E_NOT_FINITE;
        // density <= 0 on some node: the run has diverged (first such step kept)
        if (rec[3] <= 0.0) {
            if (h->diverged_step < 0) {
                h->diverged_step = h->pend_l2[i];
                fprintf(stderr, "mbl: density %g <= 0 at step %d\n", rec[3], h->pend_l2[i]);
            }
            if (!rc) rc = MBL_E_DIVERGED;
        }
    }
    for (size_t i = 0; i < nd; i++) {
        double fx = h->h_drag[2 * i], fy = h->h_drag[2 * i + 1];
        double nrm = h->cfg.rho0 * U0_2 * U0_2 * h->pend_drag[i].m * h->cfg.R;
        double cd = fx / nrm, cl = fy / nrm;
        if (fd) fprintf(fd, "%d\t%g\t%g\t\n", h->pend_drag[i].t, cd, cl); // force_block.h:214
        h->ser_drag_t.push_back(h->pend_drag[i].t);
        h->ser_cd.push_back(cd);
        h->ser_cl.push_back(cl);
        h->last[1] = fx;
        h->last[2] = fy;
        h->last[3] = cd;
        h->last[4] = cl;
        if (!std::isfinite(fx) || !std::isfinite(fy)) rc = MBL_E_NOT_FINITE;
    }
    if (fl) fclose(fl);
    if (fd) fclose(fd);
    h->pend_l2.clear();
    h->pend_drag.clear();
    return set_err(h, rc);
}

static uint64_t resid_sig(int N, const AREA *areas, const D2Q9 *pps, Lbd *const *upsx, Lbd *const *upsy)
{
    uint64_t s = 0x4242;
    for (int a = 0; a < N; a++) {
        s = mix64(s, (uint64_t)areas[a].nx | ((uint64_t)areas[a].ny << 32));
        s = mix64(s, (uint64_t)areas[a].m);
        s = mix64(s, (uintptr_t)pps[a].ux);
        s = mix64(s, (uintptr_t)pps[a].uy);
        s = mix64(s, upsx ? (uintptr_t)upsx[a] : 0);
        s = mix64(s, upsy ? (uintptr_t)upsy[a] : 0);
        s = mix64(s, (uintptr_t)pps[a].solid);
    }
    return s;
}

static int ensure_resid(mbl_context *h, int N_areas, const AREA *areas, const D2Q9 *pps_d, Lbd **upsx, Lbd **upsy);

extern "C" void error_L2(Lbi t, const Lbi N_areas, mbl_dim3 block, mbl_dim3 *grid, AREA *areas, D2Q9 *pps_d,
                         Lbd **upsx, Lbd **upsy, double **h_uvp2_b, double **d_uvp2_b, double **h_uv2_b,
                         double **d_uv2_b)
{
    (void)block;
    (void)grid;
    (void)h_uvp2_b;
    (void)d_uvp2_b;
    (void)h_uv2_b;
    (void)d_uv2_b;
    mbl_context *h = g_current;
    if (!h) return;
    if (N_areas <= 0 || !areas || !pps_d || !upsx || !upsy) {
        set_err(h, MBL_E_INVALID);
        return;
    }
    if (ensure_resid(h, N_areas, areas, pps_d, upsx, upsy)) return;
    Residual &R = h->resid;
    const size_t slot = h->pend_l2.size();
    int rc = launch_residual(h->stream, R.dd, R.tblk, (int)R.hd.size(), R.total, R.partials, 0, 0.0, 0.0, h->d_l2 + L2_REC * slot);
    if (rc) {
        set_err(h, rc);
        return;
    }
    h->pend_l2.push_back(t);
    if (h->cfg.sync_diag || (int)h->pend_l2.size() >= h->cap) mbl_flush_diagnostics(h);
}

static int ensure_resid(mbl_context *h, int N_areas, const AREA *areas, const D2Q9 *pps_d, Lbd **upsx, Lbd **upsy)
{
    Residual &R = h->resid;
    uint64_t sig = resid_sig(N_areas, areas, pps_d, upsx, upsy);
    if (R.sig != sig || !R.dd) {
        release_resid(h);
        R.hd.resize(N_areas);
        int total = 0;
        std::vector<int> tblk;
        for (int a = 0; a < N_areas; a++) {
            if (areas[a].nx < 1 || areas[a].ny < 1 || areas[a].m < 1) return set_err(h, MBL_E_INVALID);
            BlockDesc d = make_desc(h->cfg, areas[a], pps_d[a], 0, upsx ? upsx[a] : nullptr, upsy ? upsy[a] : nullptr);
            d.rtiles_x = (areas[a].nx + 63) / 64;
            d.rtile_begin = total;
            total += d.rtiles_x * ((areas[a].ny + 15) / 16);
            tblk.resize(total, a);
            R.hd[a] = d;
        }
        R.total = total;
        if (hipMalloc(&R.dd, sizeof(BlockDesc) * N_areas) != hipSuccess ||
            hipMalloc(&R.partials, sizeof(double) * 3 * total) != hipSuccess ||
            hipMalloc(&R.tblk, sizeof(int) * total) != hipSuccess ||
            hipMemcpy(R.dd, R.hd.data(), sizeof(BlockDesc) * N_areas, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(R.tblk, tblk.data(), sizeof(int) * total, hipMemcpyHostToDevice) != hipSuccess) {
            release_resid(h);
            return set_err(h, MBL_E_HIP);
        }
        R.sig = sig;
    }
    return 0;
}

extern "C" int mbl_poiseuille_l2(mbl_handle_t h, Lbi N_areas, AREA *areas, D2Q9 *pps_d, double *out)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (N_areas <= 0 || !areas || !pps_d || !out) return set_err(h, MBL_E_INVALID);
    if (ensure_resid(h, N_areas, areas, pps_d, nullptr, nullptr)) return h->last_error;
    Residual &R = h->resid;
    double *d_out = nullptr;
    HIPCHK(h, hipMalloc(&d_out, L2_REC * sizeof(double)));
    int rc = launch_residual(h->stream, R.dd, R.tblk, (int)R.hd.size(), R.total, R.partials, 1, h->cfg.U0[0], (double)h->cfg.LY,
                             d_out);
    double v[L2_REC] = {0, 0, 0, 0};
    if (!rc && hipMemcpyAsync(v, d_out, sizeof(v), hipMemcpyDeviceToHost, h->stream) != hipSuccess) rc = MBL_E_HIP;
    if (hipStreamSynchronize(h->stream) != hipSuccess) rc = MBL_E_HIP;
    hipFree(d_out);
    *out = v[0];
    return set_err(h, rc);
}

// error_L2 without any allocation: block descriptors and per-tile partial sums in the
// device's fixed scratch array (no tile table: each tile finds its block by a search),
// the record {residual, sum|u-u_prev|^2/m, sum|u|^2/m, min rho} written to the caller's
// device buffer. Same kernels and summation order as error_L2, so the same bits.
extern "C" int mbl_error_L2_noalloc(mbl_handle_t h, Lbi N_areas, const AREA *areas, const D2Q9 *pps_d,
                                    Lbd *const *upsx, Lbd *const *upsy, double *d_out)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (N_areas <= 0 || N_areas > MBL_NOALLOC_MAX_BLOCKS || !areas || !pps_d || !upsx || !upsy || !d_out) {
        fprintf(stderr, "mbl_error_L2_noalloc: invalid arguments (1 to %d blocks)\n", MBL_NOALLOC_MAX_BLOCKS);
        return MBL_E_INVALID;
    }
    BlockDesc hd[MBL_NOALLOC_MAX_BLOCKS];
    long long total = 0;
    for (int a = 0; a < N_areas; a++) {
        if (areas[a].nx < 1 || areas[a].ny < 1 || areas[a].m < 1) return MBL_E_INVALID;
        BlockDesc d = make_desc(h->cfg, areas[a], pps_d[a], 0, upsx[a], upsy[a]);
        d.rtiles_x = (areas[a].nx + 63) / 64;
        d.rtile_begin = (int)std::min<long long>(total, 0x7fffffff);
        total += (long long)d.rtiles_x * ((areas[a].ny + 15) / 16);
        hd[a] = d;
    }
    const size_t desc_bytes = (sizeof(BlockDesc) * N_areas + 255) & ~(size_t)255;
    const size_t need = desc_bytes + sizeof(double) * 3 * (size_t)total;
    if (need > MBL_NOALLOC_SCRATCH_BYTES) {
        fprintf(stderr, "mbl_error_L2_noalloc: layout needs %zu bytes of scratch, the maximum is %zu\n", need,
                MBL_NOALLOC_SCRATCH_BYTES);
        return MBL_E_INVALID;
    }
    for (int a = 0; a < N_areas; a++)
        if (!pps_d[a].ux || !pps_d[a].uy || !pps_d[a].rho || !upsx[a] || !upsy[a]) return MBL_E_INVALID;
    if (!h->scratch) return set_err(h, MBL_E_HIP);
    BlockDesc *d_desc = (BlockDesc *)h->scratch;
    double *partials = (double *)((char *)h->scratch + desc_bytes);
    HIPCHK(h, hipMemcpyAsync(d_desc, hd, sizeof(BlockDesc) * N_areas, hipMemcpyHostToDevice, h->stream));
    int rc = launch_residual(h->stream, d_desc, nullptr, N_areas, (int)total, partials, 0, 0.0, 0.0, d_out);
    // the descriptors' host copy lives on this stack frame: the call returns once the work is done
    if (hipStreamSynchronize(h->stream) != hipSuccess && !rc) rc = MBL_E_HIP;
    return set_err(h, rc);
}

extern "C" void getDrag_block(Lbi t, Int3 *d_pdrag, AREA area, D2Q9 d_pp, double *force, Lbi N)
{
    mbl_context *h = g_current;
    if (!h) return;
    if (N < 0 || (N > 0 && !d_pdrag)) {
        set_err(h, MBL_E_INVALID);
        return;
    }
    DragArgs a;
    memset(&a, 0, sizeof(a));
    for (int k = 0; k < 9; k++) a.Fold[k] = (void *)d_pp.f[k];
    a.links = d_pdrag;
    a.n = N;
    a.nx = area.nx;
    const size_t slot = h->pend_drag.size();
    a.out = h->d_drag + 2 * slot;
    int rc = launch_drag(h->stream, storage_mode(h->cfg), a);
    if (rc) {
        set_err(h, rc);
        return;
    }
    h->pend_drag.push_back(PendingDrag{t, area.m});
    if (h->cfg.sync_diag || (int)h->pend_drag.size() >= h->cap) mbl_flush_diagnostics(h);
    if (force) {
        force[0] = h->last[1];
        force[1] = h->last[2];
    }
}

// Fused steps issued as one graph (the planning of the next steps runs ahead on the host).
constexpr int MBL_CHUNK_STEPS = 6;

// One or more iterations of the reference driver's loop (Multi-block-LBM.cu:144-158):
// evolution -> error_L2 -> getDrag_block, fused where the layout allows (the residual sums
// come from the last collide of each level, the rest in one diagnostics launch), otherwise
// the three calls one after another.
extern "C" int mbl_evolution_step(mbl_handle_t handle, Lbi n, Lbi nsteps, Lbi *asp, Lbi *ls3cffcp, Lbi *ltcffcp,
                                  Lbi *lccp, mbl_dim3 block, mbl_dim3 *grid, AREA *areas, D2Q9 *pps_h, D2Q9 *pps_d,
                                  Inf_Ff_store *line_Inf_c, Inf_Ff_store *line_Inf_f, Inf_Ff_store *line_Inf_cc,
                                  Ff_Store *line_t3, FM_spline3 *line_S3, Lbd **upsx, Lbd **upsy, Int3 *d_pdrag,
                                  Lbi N_drag, double *force)
{
    mbl_context *h = handle ? handle : g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (h->last_error) return h->last_error;
    const mbl_config &cfg = h->cfg;
    if (nsteps < 0 || !asp || !ls3cffcp || !ltcffcp || !lccp || !areas || !pps_d || !upsx || !upsy ||
        N_drag < 0 || (N_drag > 0 && !d_pdrag))
        return set_err(h, MBL_E_INVALID);
    const int nblocks = asp[cfg.LEVEL];
    const int nline = ltcffcp[cfg.LEVEL];
    const int nline2 = lccp[cfg.LEVEL];
    if (nblocks <= 0 || nline != ls3cffcp[cfg.LEVEL] || (nline > 0 && (!line_Inf_c || !line_Inf_f || !line_t3 || !line_S3)) ||
        (nline2 > 0 && !line_Inf_cc))
        return set_err(h, MBL_E_INVALID);
    for (int a = 0; a < nblocks; a++)
        if (!upsx[a] || !upsy[a]) return set_err(h, MBL_E_INVALID);
    mbl_context *prev = g_current;
    g_current = h;
    const int last = nblocks - 1;
    EvoArgs A{asp, ls3cffcp, ltcffcp, lccp, areas, pps_d, line_Inf_c, line_Inf_f, line_Inf_cc, line_t3, line_S3};
    for (Lbi s = 0; s < nsteps && !h->last_error;) {
        const Lbi t = n + s;
        Layout *L = layout_for(h, A, nblocks, nline, nline2);
        if (!L) break;
        if (!L->step_ok) { // the three calls
            evolution(t, asp, ls3cffcp, ltcffcp, lccp, 0, block, grid, areas, pps_h, pps_d, line_Inf_c, line_Inf_f,
                      line_Inf_cc, line_t3, line_S3, h);
            error_L2(t, nblocks, block, grid, areas, pps_d, upsx, upsy, nullptr, nullptr, nullptr, nullptr);
            if (N_drag > 0) getDrag_block(t, d_pdrag, areas[last], pps_d[last], force, N_drag);
            s++;
            continue;
        }
        // the caller's u_prev arrays in the descriptors the collides read
        if (L->up_x.size() != (size_t)nblocks || !std::equal(upsx, upsx + nblocks, L->up_x.begin()) ||
            !std::equal(upsy, upsy + nblocks, L->up_y.begin())) {
            L->up_x.assign(upsx, upsx + nblocks);
            L->up_y.assign(upsy, upsy + nblocks);
            for (int a = 0; a < nblocks; a++) {
                L->hb[a].upx = upsx[a];
                L->hb[a].upy = upsy[a];
            }
            if (hipMemcpyAsync(L->db, L->hb.data(), sizeof(BlockDesc) * nblocks, hipMemcpyHostToDevice, h->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess) {
                set_err(h, MBL_E_HIP);
                break;
            }
        }
        StepDiagArgs d;
        memset(&d, 0, sizeof(d));
        d.descs = L->db;
        d.list = L->dlist;
        d.nlist = L->nlist;
        d.list_wgs = std::max(1, (L->nlist + 255) / 256);
        d.ntiles = L->total_tiles;
        d.rparts = L->dres;
        if (d.list_wgs > h->lparts_stride) { // room for the list partials of every record slot
            mbl_flush_diagnostics(h);
            if (h->last_error) break;
            if (h->d_lparts) hipFree(h->d_lparts);
            h->d_lparts = nullptr;
            h->lparts_stride = 0;
            if (hipMalloc(&h->d_lparts, sizeof(double) * 3 * (size_t)d.list_wgs * h->cap) != hipSuccess) {
                set_err(h, MBL_E_NOMEM);
                break;
            }
            h->lparts_stride = d.list_wgs;
        }
        // the drag's record slot is the residual's (both series advance once per step here)
        if (N_drag > 0 && h->pend_drag.size() != h->pend_l2.size()) mbl_flush_diagnostics(h);
        if (h->last_error) break;
        const int slot = (int)h->pend_l2.size();
        d.cap = h->cap;
        d.lbase = h->d_lparts;
        d.lstride = (size_t)h->lparts_stride;
        d.drag_base = h->d_drag;
        if (N_drag > 0) { // the planes (the caller's f after the call) follow from the phase
            d.drag.links = d_pdrag;
            d.drag.n = N_drag;
            d.drag.nx = areas[last].nx;
        }
        // up to MBL_CHUNK_STEPS steps, each its cycle and its diagnostics launch (record slots
        // slot, slot + 1, ...), planned on the host and issued as one graph
        int K = (int)std::min<Lbi>(nsteps - s, MBL_CHUNK_STEPS);
        K = std::max(1, std::min(K, h->cap - slot));
        if (cfg.sync_diag) K = 1;
        std::vector<StepPlan> plans(K);
        int rc = 0;
        for (int i = 0; i < K && !rc; i++) rc = plan_step(h, L, A, 0, nline, true, &d, last, slot + i, plans[i]);
        if (!rc) rc = launch_plans(h, L, 0, plans);
        if (rc) {
            set_err(h, rc);
            break;
        }
        for (int i = 0; i < K; i++) {
            h->pend_l2.push_back((int)(t + i));
            h->pend_final.push_back({slot + i, d.list_wgs});
            if (N_drag > 0) h->pend_drag.push_back(PendingDrag{(int)(t + i), areas[last].m});
        }
        s += K;
        if (cfg.sync_diag || (int)h->pend_l2.size() >= h->cap || (int)h->pend_drag.size() >= h->cap)
            mbl_flush_diagnostics(h);
        if (force && N_drag > 0) {
            force[0] = h->last[1];
            force[1] = h->last[2];
        }
    }
    g_current = prev;
    return h->last_error;
}

extern "C" int mbl_evolution_step_fused(mbl_handle_t h)
{
    if (!h) return MBL_E_NO_HANDLE;
    if (!h->layout) return MBL_E_INVALID;
    return h->layout->step_ok ? 1 : 0;
}

extern "C" int mbl_last_diagnostics(mbl_handle_t h, double out[5])
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    for (int i = 0; i < 5; i++) out[i] = h->last[i];
    return 0;
}

extern "C" int mbl_diagnostics_series(mbl_handle_t h, int kind, int cap, int *t, double *a, double *b)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (kind == 0) {
        int n = (int)h->ser_l2.size();
        for (int i = 0; i < std::min(n, cap); i++) {
            if (t) t[i] = h->ser_l2_t[i];
            if (a) a[i] = h->ser_l2[i];
        }
        return n;
    }
    int n = (int)h->ser_cd.size();
    for (int i = 0; i < std::min(n, cap); i++) {
        if (t) t[i] = h->ser_drag_t[i];
        if (a) a[i] = h->ser_cd[i];
        if (b) b[i] = h->ser_cl[i];
    }
    return n;
}

extern "C" void mbl_diagnostics_clear(mbl_handle_t h)
{
    if (!h) h = g_current;
    if (!h) return;
    h->ser_l2_t.clear();
    h->ser_l2.clear();
    h->ser_drag_t.clear();
    h->ser_cd.clear();
    h->ser_cl.clear();
}

extern "C" int mbl_cycle_launches(mbl_handle_t h, int level, int *collides, int *interface_ops, int *splines)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!h->layout || level < 0 || level >= MBL_MAX_LEVEL || !h->layout->prog[level].built) return MBL_E_INVALID;
    int c[5] = {0, 0, 0, 0, 0};
    for (const CycleLaunch &cl : h->layout->prog[level].launches) c[cl.kind]++;
    if (collides) *collides = c[0] + c[4];
    if (interface_ops) *interface_ops = c[1] + c[3] + c[4];
    if (splines) *splines = c[2] + c[3];
    return c[0] + c[1] + c[2] + c[3] + c[4];
}

extern "C" int mbl_cycle_positions(mbl_handle_t h, int level)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!h->layout || level < 0 || level >= MBL_MAX_LEVEL || !h->layout->prog[level].built) return MBL_E_INVALID;
    const auto &v = h->layout->prog[level].launches;
    return v.empty() ? 0 : v.back().pos + 1;
}

extern "C" int mbl_set_cycle_fork(mbl_handle_t h, int on)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->fork = on != 0;
    if (h->layout) { // captured graphs hold the other launch order; the schedule is rebuilt
        for (auto &g : h->layout->graphs) hipGraphExecDestroy(g.second);
        h->layout->graphs.clear();
        for (auto &p : h->layout->prog) p.built = false;
    }
    return 0;
}

extern "C" int mbl_divergence_step(mbl_handle_t h)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    return h->diverged_step;
}

// ---------------------------------------------------------------------------
// field output (output/output.h:2-38) and checkpoints
// ---------------------------------------------------------------------------
static int d2h(mbl_context *h, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return MBL_E_HIP;
    return 0;
}

static int h2d(mbl_context *h, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return MBL_E_HIP;
    return 0;
}

extern "C" int mbl_output(mbl_handle_t h, Lbi n, Lbi number_area, const AREA *areas, D2Q9 *pps_h, const D2Q9 *pps_d)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (number_area <= 0 || !areas || !pps_d) return set_err(h, MBL_E_INVALID);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    FILE *fo = fopen(out_path(h, ("steady_" + std::to_string(n) + ".dat").c_str()).c_str(), "w");
    if (!fo) return set_err(h, MBL_E_INVALID);
    fprintf(fo, "Title= \"LBM 2D\"\n");
    fprintf(fo, "VARIABLES= \"X\",\"Y\",\"solid\",\"rho\",\"U\",\"V\",\n");
    int rc = 0;
    for (int a = 0; a < number_area && !rc; a++) {
        const size_t N = (size_t)areas[a].nx * areas[a].ny;
        std::vector<int> solid(N);
        std::vector<Lbd> v[5];
        const Lbd *src[5] = {pps_d[a].x, pps_d[a].y, pps_d[a].rho, pps_d[a].ux, pps_d[a].uy};
        rc = d2h(h, solid.data(), pps_d[a].solid, N * sizeof(int));
        for (int q = 0; q < 5 && !rc; q++) {
            v[q].resize(N);
            rc = d2h(h, v[q].data(), src[q], N * sizeof(Lbd));
        }
        if (rc) break;
        // the reference's side effect: the host mirror of every array
        if (pps_h) {
            D2Q9 &ph = pps_h[a];
            if (ph.solid) memcpy(ph.solid, solid.data(), N * sizeof(int));
            Lbd *dst[5] = {ph.x, ph.y, ph.rho, ph.ux, ph.uy};
            for (int q = 0; q < 5; q++)
                if (dst[q]) memcpy(dst[q], v[q].data(), N * sizeof(Lbd));
            if (ph.forcex && !rc) rc = d2h(h, ph.forcex, pps_d[a].forcex, N * sizeof(Lbd));
            if (ph.forcey && !rc) rc = d2h(h, ph.forcey, pps_d[a].forcey, N * sizeof(Lbd));
            std::vector<double> tmp(N);
            for (int which = 0; which < 2 && !rc; which++)
                for (int k = 0; k < 9 && !rc; k++) {
                    Lbd *dp = which ? ph.f[k] : ph.F[k];
                    if (!dp) continue;
                    rc = plane_io(h, &pps_d[a], which, k, N, tmp.data(), true);
                    for (size_t i = 0; i < N && !rc; i++) dp[i] = (Lbd)tmp[i];
                }
            if (rc) break;
        }
        fprintf(fo, "ZONE T= \"zone_level%d\",I= %d,J= %d,F=POINT\n", a, areas[a].nx, areas[a].ny);
        for (size_t i = 0; i < N; i++)
            fprintf(fo, "%g\t%g\t%d\t%g\t%g\t%g\n", (double)v[0][i], (double)v[1][i], solid[i], (double)v[2][i],
                    (double)v[3][i], (double)v[4][i]);
    }
    if (fclose(fo) != 0 && !rc) rc = MBL_E_INVALID;
    return set_err(h, rc);
}

extern "C" void output(Lbi n, AREA *areas, D2Q9 *pps_h, D2Q9 *pps_d)
{
    mbl_context *h = g_current;
    if (!h) return;
    int na = h->cfg.number_area;
    if (na <= 0 && h->layout) na = h->layout->nblocks;
    if (na <= 0) {
        set_err(h, MBL_E_INVALID);
        return;
    }
    mbl_output(h, n, na, areas, pps_h, pps_d);
}

// Checkpoint file: header, then per block F[9], f[9] (storage encoding), rho, ux, uy and
// optionally upx, upy (Lbd), then per line the markers and 27 ring arrays (Lbd).
struct CkptHeader {
    char magic[8];     // "MBLCKPT1"
    int32_t storage, elem_bytes, lbd_bytes, number_area, number_line, has_up;
    int64_t step;
};

static int ckpt_write(FILE *f, const void *p, size_t bytes) { return fwrite(p, 1, bytes, f) == bytes ? 0 : MBL_E_INVALID; }
static int ckpt_read(FILE *f, void *p, size_t bytes) { return fread(p, 1, bytes, f) == bytes ? 0 : MBL_E_INVALID; }

extern "C" int mbl_checkpoint_save(mbl_handle_t h, const char *path, long long step, Lbi number_area,
                                   const AREA *areas, const D2Q9 *pps_d, Lbd *const *upsx, Lbd *const *upsy,
                                   Lbi number_line, const Ff_Store *line_t3)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!path || number_area <= 0 || !areas || !pps_d || number_line < 0 || (number_line > 0 && !line_t3))
        return set_err(h, MBL_E_INVALID);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    FILE *f = fopen(path, "wb");
    if (!f) return set_err(h, MBL_E_INVALID);
    CkptHeader hd;
    memset(&hd, 0, sizeof(hd));
    memcpy(hd.magic, "MBLCKPT1", 8);
    hd.storage = h->cfg.storage;
    hd.elem_bytes = (int32_t)store_elem_bytes(storage_mode(h->cfg));
    hd.lbd_bytes = (int32_t)sizeof(Lbd);
    hd.number_area = number_area;
    hd.number_line = number_line;
    hd.has_up = (upsx && upsy) ? 1 : 0;
    hd.step = step;
    int rc = ckpt_write(f, &hd, sizeof(hd));
    std::vector<char> buf;
    for (int a = 0; a < number_area && !rc; a++) {
        const size_t N = (size_t)areas[a].nx * areas[a].ny;
        const int32_t dims[3] = {areas[a].m, areas[a].nx, areas[a].ny};
        rc = ckpt_write(f, dims, sizeof(dims));
        buf.resize(N * std::max<size_t>(hd.elem_bytes, sizeof(Lbd)));
        for (int which = 0; which < 2 && !rc; which++)
            for (int k = 0; k < 9 && !rc; k++) {
                const void *src = which ? (const void *)pps_d[a].f[k] : (const void *)pps_d[a].F[k];
                rc = d2h(h, buf.data(), src, N * hd.elem_bytes);
                if (!rc) rc = ckpt_write(f, buf.data(), N * hd.elem_bytes);
            }
        const Lbd *mac[5] = {pps_d[a].rho, pps_d[a].ux, pps_d[a].uy, hd.has_up ? upsx[a] : nullptr,
                             hd.has_up ? upsy[a] : nullptr};
        for (int q = 0; q < (hd.has_up ? 5 : 3) && !rc; q++) {
            rc = mac[q] ? d2h(h, buf.data(), mac[q], N * sizeof(Lbd)) : MBL_E_INVALID;
            if (!rc) rc = ckpt_write(f, buf.data(), N * sizeof(Lbd));
        }
    }
    for (int b = 0; b < number_line && !rc; b++) {
        const Ff_Store &t = line_t3[b];
        const int32_t hdr[4] = {t.N, t.marker[0], t.marker[1], t.marker[2]};
        rc = ckpt_write(f, hdr, sizeof(hdr));
        buf.resize((size_t)t.N * sizeof(Lbd));
        const F_k3 *ft = &t.f_t1;
        for (int sl = 0; sl < 3 && !rc; sl++)
            for (int k = 0; k < 9 && !rc; k++) {
                rc = d2h(h, buf.data(), ft[sl].F[k], (size_t)t.N * sizeof(Lbd));
                if (!rc) rc = ckpt_write(f, buf.data(), (size_t)t.N * sizeof(Lbd));
            }
    }
    if (!rc) rc = ckpt_write(f, "MBLEND\0\0", 8);
    if (fclose(f) != 0 && !rc) rc = MBL_E_INVALID;
    return set_err(h, rc);
}

extern "C" int mbl_checkpoint_load(mbl_handle_t h, const char *path, long long *step, Lbi number_area,
                                   const AREA *areas, D2Q9 *pps_d, Lbd **upsx, Lbd **upsy, Lbi number_line,
                                   Ff_Store *line_t3)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!path || number_area <= 0 || !areas || !pps_d || number_line < 0 || (number_line > 0 && !line_t3))
        return set_err(h, MBL_E_INVALID);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    FILE *f = fopen(path, "rb");
    if (!f) return set_err(h, MBL_E_INVALID);
    CkptHeader hd;
    int rc = ckpt_read(f, &hd, sizeof(hd));
    if (!rc && (memcmp(hd.magic, "MBLCKPT1", 8) != 0 || hd.storage != h->cfg.storage ||
                hd.elem_bytes != (int32_t)store_elem_bytes(storage_mode(h->cfg)) ||
                hd.lbd_bytes != (int32_t)sizeof(Lbd) || hd.number_area != number_area ||
                hd.number_line != number_line || (hd.has_up && !(upsx && upsy))))
        rc = MBL_E_INVALID;
    std::vector<char> buf;
    for (int a = 0; a < number_area && !rc; a++) {
        const size_t N = (size_t)areas[a].nx * areas[a].ny;
        int32_t dims[3];
        rc = ckpt_read(f, dims, sizeof(dims));
        if (!rc && (dims[0] != areas[a].m || dims[1] != areas[a].nx || dims[2] != areas[a].ny)) rc = MBL_E_INVALID;
        buf.resize(N * std::max<size_t>(hd.elem_bytes, sizeof(Lbd)));
        for (int which = 0; which < 2 && !rc; which++)
            for (int k = 0; k < 9 && !rc; k++) {
                void *dst = which ? (void *)pps_d[a].f[k] : (void *)pps_d[a].F[k];
                rc = ckpt_read(f, buf.data(), N * hd.elem_bytes);
                if (!rc) rc = h2d(h, dst, buf.data(), N * hd.elem_bytes);
            }
        Lbd *mac[5] = {pps_d[a].rho, pps_d[a].ux, pps_d[a].uy, hd.has_up ? upsx[a] : nullptr,
                       hd.has_up ? upsy[a] : nullptr};
        for (int q = 0; q < (hd.has_up ? 5 : 3) && !rc; q++) {
            rc = ckpt_read(f, buf.data(), N * sizeof(Lbd));
            if (!rc) rc = mac[q] ? h2d(h, mac[q], buf.data(), N * sizeof(Lbd)) : MBL_E_INVALID;
        }
    }
    for (int b = 0; b < number_line && !rc; b++) {
        Ff_Store &t = line_t3[b];
        int32_t hdr[4];
        rc = ckpt_read(f, hdr, sizeof(hdr));
        if (!rc && hdr[0] != t.N) rc = MBL_E_INVALID;
        if (rc) break;
        buf.resize((size_t)t.N * sizeof(Lbd));
        F_k3 *ft = &t.f_t1;
        for (int sl = 0; sl < 3 && !rc; sl++)
            for (int k = 0; k < 9 && !rc; k++) {
                rc = ckpt_read(f, buf.data(), (size_t)t.N * sizeof(Lbd));
                if (!rc) rc = h2d(h, ft[sl].F[k], buf.data(), (size_t)t.N * sizeof(Lbd));
            }
        if (!rc)
            for (int q = 0; q < 3; q++) t.marker[q] = hdr[1 + q];
    }
    char tail[8];
    if (!rc && (ckpt_read(f, tail, 8) || memcmp(tail, "MBLEND\0\0", 8) != 0)) rc = MBL_E_INVALID;
    fclose(f);
    if (!rc && step) *step = hd.step;
    return set_err(h, rc);
}

// ---------------------------------------------------------------------------
// batched independent cases
// ---------------------------------------------------------------------------
struct mbl_batch {
    mbl_context *h = nullptr;
    int n = 0;
    std::vector<mbl_case_spec> cs;
    std::vector<AREA> areas;
    std::vector<D2Q9> pps;
    std::vector<Consts> bc;
    std::vector<Inf_Ff_store> lc, lf, lcc;
    std::vector<Ff_Store> t3;
    std::vector<FM_spline3> s3;
    int asp[3] = {0, 0, 0}, ls3[3] = {0, 0, 0}, ltc[3] = {0, 0, 0}, lccp[3] = {0, 0, 0};
    char *pool = nullptr;         // one device allocation for every plane, ring and line array
    Int3 *d_links = nullptr;
    DragArgs *d_drag[2] = {nullptr, nullptr}; // per case, for the two pointer orders of the refined block
    void *fine_F0 = nullptr;       // refined block of case 0: its F[0] at creation (parity reference)
    double *d_hist = nullptr;      // force history: [cap][n][2]
    int cap = 4096, nhist = 0, head = 0;
    std::vector<double> norm;      // rho0 U^2 m R per case
    // raster exchange: block/case tables (built on first use) and a staging buffer
    RasterBlock *d_rblk = nullptr;
    RasterCase *d_rcase = nullptr;
    size_t max_nodes = 0;
    float *d_raster = nullptr;
    size_t raster_bytes = 0;
    // per-case residual (error_L2 of each case): previous velocities of every block, and the
    // residual tables in case-major block order (built on first use)
    std::vector<Lbd *> upx, upy;
    BlockDesc *d_rdesc = nullptr;
    int *d_rtblk = nullptr, *d_case_tiles = nullptr;
    double *d_rpart = nullptr, *d_rout = nullptr;
    int rtotal = 0;
    // one launch per mbl_batch_step call, one workgroup per case (mbl_batch_set_persistent):
    // the item tables of every step of the phase period, built for one layout generation
    bool persistent = false;
    struct CaseProgram {
        uint64_t gen = 0;                     // layout generation of the tables (0: none)
        CaseItem *d_items = nullptr;          // [period][n][per_case]
        LineOp *d_ops = nullptr;              // interface and spline ops, each case's contiguous
        int *d_case = nullptr;                // workgroup -> case, largest case first
        int per_case = 0, period = 0;
        size_t lds = 0;
        std::vector<uint64_t> key;            // phase key at the start of each step of the period
        std::vector<uint32_t> flips;          // levels whose parity each step flips
        std::vector<std::vector<int>> markers; // ring markers of every line after each step
        void release()
        {
            if (d_items) hipFree(d_items);
            if (d_ops) hipFree(d_ops);
            if (d_case) hipFree(d_case);
            *this = CaseProgram();
        }
    } cp;
};

// u_prev <- u for every block of a batch (after creation and after every restart), so the
// first per-case residual measures the change since that state
static int batch_seed_uprev(mbl_batch *B)
{
    mbl_context *h = B->h;
    for (size_t a = 0; a < B->areas.size(); a++) {
        const size_t bytes = sizeof(Lbd) * (size_t)B->areas[a].nx * B->areas[a].ny;
        HIPCHK(h, hipMemcpyAsync(B->upx[a], B->pps[a].ux, bytes, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(B->upy[a], B->pps[a].uy, bytes, hipMemcpyDeviceToDevice, h->stream));
    }
    return 0;
}

// The three ring slots of every coarse/fine line <- the fine edge's current
// F (init_Fstore, multi/init_multi.h:3-28, for all lines of all cases at once).
static int batch_ring_fill(mbl_batch *B)
{
    mbl_context *h = B->h;
    const int n = B->n;
    const int mode = storage_mode(h->cfg);
    std::vector<LineOp> ops(4 * n);
    std::vector<int> idx;
    int maxn = 0;
    for (int b = 0; b < 4 * n; b++) {
        LineOp &op = ops[b];
        memset(&op, 0, sizeof(op));
        op.n = B->t3[b].N;
        op.dst_block = B->lc[b].areaf_n;
        const F_k3 *ft = &B->t3[b].f_t1;
        for (int sl = 0; sl < 3; sl++)
            for (int k = 0; k < 9; k++) op.ring[sl][k] = ft[sl].F[k];
        const AREA &af = B->areas[op.dst_block];
        op.src_idx = (const int *)(uintptr_t)idx.size(); // offset, fixed up below
        for (int ip = 0; ip < op.n; ip++) {
            int i, j;
            line_node(B->lc[b].line_F, ip, i, j);
            idx.push_back(j * af.nx + i);
        }
        maxn = std::max(maxn, op.n);
    }
    std::vector<BlockDesc> blocks(5 * n);
    for (int a = 0; a < 5 * n; a++) blocks[a] = make_desc(h->cfg, B->areas[a], B->pps[a], 0, nullptr, nullptr);
    const size_t bi = sizeof(int) * std::max<size_t>(1, idx.size());
    const size_t bo = sizeof(LineOp) * ops.size();
    const size_t bb = sizeof(BlockDesc) * blocks.size();
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    char *tmp = nullptr;
    HIPCHK(h, hipMalloc(&tmp, al(bi) + al(bo) + bb));
    int *didx = (int *)tmp;
    LineOp *dops = (LineOp *)(tmp + al(bi));
    BlockDesc *dbl = (BlockDesc *)(tmp + al(bi) + al(bo));
    for (auto &op : ops) {
        op.dst_idx = didx + (uintptr_t)op.src_idx;
        op.src_idx = nullptr;
    }
    int rc = 0;
    if (hipMemcpyAsync(didx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(dops, ops.data(), bo, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(dbl, blocks.data(), bb, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        rc = MBL_E_HIP;
    if (!rc) rc = launch_ring_fill(h->stream, mode, dops, (int)ops.size(), maxn, dbl, 0u);
    if (hipStreamSynchronize(h->stream) != hipSuccess && !rc) rc = MBL_E_HIP;
    hipFree(tmp);
    return rc;
}

static int batch_build(mbl_batch *B)
{
    mbl_context *h = B->h;
    const int n = B->n;
    const int mode = storage_mode(h->cfg);
    const size_t eb = store_elem_bytes(mode);
    // per-case geometry through the channel builder and geo_init
    std::vector<AREA> ca(5 * n);
    std::vector<Inf_Ff_store> clc(4 * n), clf(4 * n), clcc(8 * n);
    for (int c = 0; c < n; c++) {
        const mbl_case_spec &sp = B->cs[c];
        mbl_channel_spec s;
        memset(&s, 0, sizeof(s));
        s.LX = sp.LX;
        s.LY = sp.LY;
        s.levels = 2;
        s.region[1] = sp.region;
        s.Re = sp.Re;
        s.U = sp.U;
        s.R = sp.R;
        mbl_geometry g;
        P4 ad[8], adef[5];
        P3i cf[16], cc[16];
        if (mbl_geometry_channel(&s, &g, ad, cf, cc) || g.number_area != 5 || g.number_line != 4 ||
            g.number_line2 != 8)
            return MBL_E_INVALID;
        if (mbl_geo_init(&g, adef, &ca[5 * c], &clc[4 * c], &clf[4 * c], &clcc[8 * c])) return MBL_E_INVALID;
    }
    // global block order: all level-0 blocks (case-major), then all refined blocks
    auto gid = [n](int c, int q) { return q < 4 ? 4 * c + q : 4 * n + c; };
    B->areas.resize(5 * n);
    B->bc.resize(5 * n);
    for (int c = 0; c < n; c++)
        for (int q = 0; q < 5; q++) {
            B->areas[gid(c, q)] = ca[5 * c + q];
            Consts k;
            k.rho0 = 1.0;
            k.U0x = B->cs[c].U;
            k.U0y = 0.0;
            k.LY = (double)B->cs[c].LY;
            k.umax = B->cs[c].U;
            B->bc[gid(c, q)] = k;
        }
    B->lc.resize(4 * n);
    B->lf.resize(4 * n);
    B->lcc.resize(8 * n);
    for (int c = 0; c < n; c++) {
        for (int b = 0; b < 4; b++) {
            Inf_Ff_store x = clc[4 * c + b], y = clf[4 * c + b];
            x.areac_n = gid(c, x.areac_n);
            x.areaf_n = gid(c, x.areaf_n);
            y.areac_n = gid(c, y.areac_n);
            y.areaf_n = gid(c, y.areaf_n);
            B->lc[4 * c + b] = x;
            B->lf[4 * c + b] = y;
        }
        for (int b = 0; b < 8; b++) {
            Inf_Ff_store x = clcc[8 * c + b];
            x.areac_n = gid(c, x.areac_n);
            x.areaf_n = gid(c, x.areaf_n);
            B->lcc[8 * c + b] = x;
        }
    }
    B->asp[1] = 4 * n;
    B->asp[2] = 5 * n;
    B->ls3[1] = 4 * n;
    B->ls3[2] = 4 * n;
    B->ltc[2] = 4 * n;
    B->lccp[1] = 8 * n;
    B->lccp[2] = 8 * n;
    // one pool: planes of every block, then ring slots and spline arrays of every line
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t total = 0;
    for (const AREA &a : B->areas) {
        size_t nn = (size_t)a.nx * a.ny;
        total += al(nn * sizeof(int)) + 9 * al(nn * sizeof(Lbd)) + 18 * al(nn * eb);
    }
    for (const Inf_Ff_store &L : B->lc) total += 27 * al(L.Nf * sizeof(Lbd)) + 28 * al(L.Nc * sizeof(Lbd));
    HIPCHK(h, hipMalloc(&B->pool, total));
    char *p = B->pool;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += al(bytes);
        return q;
    };
    B->pps.resize(5 * n);
    for (size_t a = 0; a < B->areas.size(); a++) {
        size_t nn = (size_t)B->areas[a].nx * B->areas[a].ny;
        D2Q9 &pp = B->pps[a];
        pp.solid = (int *)take(nn * sizeof(int));
        Lbd **v = &pp.x;
        for (int q = 0; q < 7; q++) v[q] = (Lbd *)take(nn * sizeof(Lbd));
        for (int k = 0; k < 9; k++) {
            pp.f[k] = (Lbd *)take(nn * eb);
            pp.F[k] = (Lbd *)take(nn * eb);
        }
    }
    B->upx.resize(5 * n);
    B->upy.resize(5 * n);
    for (size_t a = 0; a < B->areas.size(); a++) {
        size_t nn = (size_t)B->areas[a].nx * B->areas[a].ny;
        B->upx[a] = (Lbd *)take(nn * sizeof(Lbd));
        B->upy[a] = (Lbd *)take(nn * sizeof(Lbd));
    }
    B->t3.resize(4 * n);
    B->s3.resize(4 * n);
    for (size_t b = 0; b < B->lc.size(); b++) {
        Ff_Store &t = B->t3[b];
        t.N = B->lc[b].Nf;
        t.marker[0] = 2;
        t.marker[1] = 1;
        t.marker[2] = 0;
        F_k3 *ft = &t.f_t1;
        for (int sl = 0; sl < 3; sl++)
            for (int k = 0; k < 9; k++) ft[sl].F[k] = (Lbd *)take(t.N * sizeof(Lbd));
        FM_spline3 &S = B->s3[b];
        S.N = B->lc[b].Nc;
        S.l = (Lbd *)take(S.N * sizeof(Lbd));
        S.m = (Lbd *)take(S.N * sizeof(Lbd));
        S.u = (Lbd *)take(S.N * sizeof(Lbd));
        S.Varibs_M = (Lbd *)take(12 * S.N * sizeof(Lbd));
        Lbd **fm = &S.rho;
        for (int q = 0; q < 12; q++) fm[q] = (Lbd *)take(S.N * sizeof(Lbd));
    }
    // initial state (init_General with the case's cylinder and channel)
    for (int c = 0; c < n; c++)
        for (int q = 0; q < 5; q++) {
            int a = gid(c, q);
            const mbl_case_spec &sp = B->cs[c];
            BlockDesc d = make_desc(h->cfg, B->areas[a], B->pps[a], 0, nullptr, nullptr);
            d.bc = B->bc[a];
            int rc = launch_init_general(h->stream, mode, d, B->areas[a], B->pps[a].x, B->pps[a].y, B->pps[a].forcex,
                                         B->pps[a].forcey, MBL_BC_CHANNEL, MBL_INIT_PROFILE, 0, 0.0, (double)sp.LX,
                                         (double)sp.LY, sp.R, sp.CX, sp.CY, d.bc);
            if (rc) return rc;
        }
    // tridiagonal arrays and rings of every line
    for (int b = 0; b < 4 * n; b++) {
        int rc = launch_tri_A(h->stream, B->s3[b].l, B->s3[b].m, B->s3[b].u, B->s3[b].N);
        if (rc) return rc;
    }
    if (int rc = batch_ring_fill(B)) return rc;
    if (int rc = batch_seed_uprev(B)) return rc;
    // drag links of every refined block
    std::vector<Int3> links;
    std::vector<int> off(n + 1, 0);
    for (int c = 0; c < n; c++) {
        int a = gid(c, 4);
        std::vector<Int3> L;
        if (host_links(h, B->areas[a], B->pps[a].solid, nullptr, &L) < 0) return MBL_E_HIP;
        links.insert(links.end(), L.begin(), L.end());
        off[c + 1] = (int)links.size();
    }
    HIPCHK(h, hipMalloc(&B->d_links, sizeof(Int3) * std::max<size_t>(1, links.size())));
    if (!links.empty())
        HIPCHK(h, hipMemcpy(B->d_links, links.data(), sizeof(Int3) * links.size(), hipMemcpyHostToDevice));
    HIPCHK(h, hipMalloc(&B->d_hist, sizeof(double) * 2 * (size_t)n * B->cap));
    B->norm.resize(n);
    for (int par = 0; par < 2; par++) {
        std::vector<DragArgs> args(n);
        for (int c = 0; c < n; c++) {
            int a = gid(c, 4);
            DragArgs &d = args[c];
            memset(&d, 0, sizeof(d));
            // after an odd number of refined-level substeps F/f are exchanged; the
            // planes one substep older are the set not holding the state
            for (int k = 0; k < 9; k++) d.Fold[k] = par ? (void *)B->pps[a].F[k] : (void *)B->pps[a].f[k];
            d.links = B->d_links + off[c];
            d.n = off[c + 1] - off[c];
            d.nx = B->areas[a].nx;
            B->norm[c] = 1.0 * B->cs[c].U * B->cs[c].U * B->areas[a].m * B->cs[c].R;
        }
        HIPCHK(h, hipMalloc(&B->d_drag[par], sizeof(DragArgs) * n));
        HIPCHK(h, hipMemcpy(B->d_drag[par], args.data(), sizeof(DragArgs) * n, hipMemcpyHostToDevice));
    }
    B->fine_F0 = (void *)B->pps[gid(0, 4)].F[0];
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int mbl_batch_create(mbl_handle_t h, int ncases, const mbl_case_spec *cases, mbl_batch_t *out)
{
    if (!h) h = g_current;
    if (!h) return MBL_E_NO_HANDLE;
    if (!out || !cases || ncases <= 0) return set_err(h, MBL_E_INVALID);
    if (h->cfg.LEVEL != 2 || h->cfg.mi[0] != 1 || h->cfg.mi[1] != 2 || h->cfg.bc_mode != MBL_BC_CHANNEL)
        return set_err(h, MBL_E_INVALID);
    *out = nullptr;
    mbl_batch *B = new mbl_batch();
    B->h = h;
    B->n = ncases;
    B->cs.assign(cases, cases + ncases);
    int rc = batch_build(B);
    if (rc) {
        mbl_batch_destroy(B);
        return set_err(h, rc);
    }
    *out = B;
    return 0;
}

extern "C" void mbl_batch_destroy(mbl_batch_t B)
{
    if (!B) return;
    if (B->h) {
        hipStreamSynchronize(B->h->stream);
        drop_layouts(B->h);
    }
    if (B->pool) hipFree(B->pool);
    if (B->d_links) hipFree(B->d_links);
    if (B->d_hist) hipFree(B->d_hist);
    for (int p = 0; p < 2; p++)
        if (B->d_drag[p]) hipFree(B->d_drag[p]);
    if (B->d_rblk) hipFree(B->d_rblk);
    if (B->d_rcase) hipFree(B->d_rcase);
    if (B->d_raster) hipFree(B->d_raster);
    for (void *q : {(void *)B->d_rdesc, (void *)B->d_rtblk, (void *)B->d_case_tiles, (void *)B->d_rpart,
                    (void *)B->d_rout})
        if (q) hipFree(q);
    B->cp.release();
    delete B;
}

// error_L2 (lbm/error_L2.h:57-91) of every case of a batch on its own: case c's
// sqrt(sum|u - u_prev|^2/m) / sqrt(sum|u|^2/m) over its five blocks, u_prev <- u. Tables in
// case-major block order are built on the first call.
extern "C" int mbl_batch_residual(mbl_batch_t B, double *out)
{
    if (!B || !out) return MBL_E_INVALID;
    mbl_context *h = B->h;
    const int n = B->n;
    if (!B->d_rdesc) {
        std::vector<BlockDesc> hd;
        std::vector<int> tblk, case_tiles(n + 1, 0);
        int total = 0;
        for (int c = 0; c < n; c++) {
            for (int q = 0; q < 5; q++) {
                const int a = q < 4 ? 4 * c + q : 4 * n + c;
                BlockDesc d = make_desc(h->cfg, B->areas[a], B->pps[a], 0, B->upx[a], B->upy[a]);
                d.rtiles_x = (B->areas[a].nx + 63) / 64;
                d.rtile_begin = total;
                total += d.rtiles_x * ((B->areas[a].ny + 15) / 16);
                tblk.resize(total, (int)hd.size());
                hd.push_back(d);
            }
            case_tiles[c + 1] = total;
        }
        B->rtotal = total;
        if (hipMalloc(&B->d_rdesc, sizeof(BlockDesc) * hd.size()) != hipSuccess ||
            hipMalloc(&B->d_rtblk, sizeof(int) * tblk.size()) != hipSuccess ||
            hipMalloc(&B->d_case_tiles, sizeof(int) * case_tiles.size()) != hipSuccess ||
            hipMalloc(&B->d_rpart, sizeof(double) * 3 * (size_t)total) != hipSuccess ||
            hipMalloc(&B->d_rout, sizeof(double) * 4 * (size_t)n) != hipSuccess)
            return set_err(h, MBL_E_NOMEM);
        HIPCHK(h, hipMemcpy(B->d_rdesc, hd.data(), sizeof(BlockDesc) * hd.size(), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(B->d_rtblk, tblk.data(), sizeof(int) * tblk.size(), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(B->d_case_tiles, case_tiles.data(), sizeof(int) * case_tiles.size(),
                            hipMemcpyHostToDevice));
    }
    // the macros the descriptors point at are the caller-visible planes: the state of the last
    // step (the F/f exchange of a batch does not move them)
    int rc = launch_residual_cases(h->stream, B->d_rdesc, B->d_rtblk, 5 * n, B->rtotal, B->d_rpart, n,
                                   B->d_case_tiles, B->d_rout);
    if (rc) return set_err(h, rc);
    std::vector<double> rec(4 * (size_t)n);
    HIPCHK(h, hipMemcpyAsync(rec.data(), B->d_rout, sizeof(double) * rec.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int c = 0; c < n; c++) out[c] = rec[4 * (size_t)c];
    return 0;
}

extern "C" double mbl_batch_node_updates(mbl_batch_t B)
{
    if (!B) return 0.0;
    double u = 0.0;
    for (const AREA &a : B->areas) u += (double)a.nx * a.ny * a.m;
    return u;
}

// The item tables of the one-launch batch step. From the current state, the refinement
// cycle is recorded in the reference order (cycle(), the same recording evolution() schedules)
// step after step until the phase key (level parities and ring markers) comes back: the
// period. Each case's items are the recorded ones restricted to its own blocks and lines --
// the stream-collide of a level over its tiles, each sub-batch of interface ops and the
// spline lines that touch its blocks -- followed by its drag; a workgroup runs them in that
// order with a barrier between items, so every read sees what it sees in the reference order.
static int batch_case_program(mbl_batch *B, Layout *L)
{
    mbl_context *h = B->h;
    const int n = B->n;
    mbl_batch::CaseProgram &cp = B->cp;
    cp.release();
    if (L->LEVEL != 2 || L->vx[0] != L->vx[1]) return MBL_E_UNSUPPORTED;
    std::vector<int> case_of(5 * (size_t)n);
    for (int c = 0; c < n; c++) {
        for (int q = 0; q < 4; q++) case_of[4 * c + q] = c;
        case_of[4 * n + c] = c;
    }
    // tiles of case c at level l: its blocks there are consecutive in the tile numbering
    auto block_tiles = [&](int a) {
        const int l = L->level_of[a];
        const int end = (a + 1 < L->ASP[l + 1]) ? L->hb[a + 1].tile_begin : L->tile_off[l] + L->tiles[l];
        return end - L->hb[a].tile_begin;
    };
    std::vector<int> t0(2 * (size_t)n), tn(2 * (size_t)n);
    for (int c = 0; c < n; c++) {
        t0[2 * c] = L->hb[4 * c].tile_begin;
        tn[2 * c] = 0;
        for (int q = 0; q < 4; q++) tn[2 * c] += block_tiles(4 * c + q);
        if (L->hb[4 * c + 3].tile_begin + block_tiles(4 * c + 3) - t0[2 * c] != tn[2 * c]) return MBL_E_UNSUPPORTED;
        t0[2 * c + 1] = L->hb[4 * n + c].tile_begin;
        tn[2 * c + 1] = block_tiles(4 * n + c);
    }
    // the steps of one period
    uint32_t pm0;
    uint64_t key0;
    if (int rc = phase_key(L, B->pps.data(), B->t3.data(), 0, pm0, key0)) return rc;
    std::vector<Ff_Store> t3 = B->t3;
    std::vector<D2Q9> pps = B->pps;
    std::vector<std::vector<CycleItem>> steps;
    std::vector<int> par;
    for (int st = 0; st < 12; st++) {
        uint32_t pm;
        uint64_t key;
        if (int rc = phase_key(L, pps.data(), t3.data(), 0, pm, key)) return rc;
        if (st > 0 && key == key0) break;
        CycleState cs{pm, 0, {}};
        cycle(L, 0, t3.data(), cs);
        if (cs.rc) return cs.rc;
        const uint32_t fl = cs.pm ^ pm;
        for (int l = 0; l < L->LEVEL; l++)
            if (fl & (1u << l))
                for (int a = L->ASP[l]; a < L->ASP[l + 1]; a++)
                    for (int k = 0; k < 9; k++) std::swap(pps[a].F[k], pps[a].f[k]);
        std::vector<int> mk(3 * t3.size());
        for (size_t b = 0; b < t3.size(); b++)
            for (int q = 0; q < 3; q++) mk[3 * b + q] = t3[b].marker[q];
        cp.key.push_back(key);
        cp.flips.push_back(fl);
        cp.markers.push_back(mk);
        par.push_back(pps[4 * n].F[0] == (Lbd *)B->fine_F0 ? 0 : 1);
        steps.push_back(std::move(cs.items));
    }
    {
        uint32_t pm;
        uint64_t key;
        if (phase_key(L, pps.data(), t3.data(), 0, pm, key) || key != key0) return MBL_E_UNSUPPORTED;
    }
    const int period = (int)steps.size();
    const std::vector<CycleItem> &it0 = steps[0];
    for (const auto &v : steps) {
        if (v.size() != it0.size()) return MBL_E_UNSUPPORTED;
        for (size_t j = 0; j < v.size(); j++)
            if (v[j].kind != it0[j].kind || v[j].level != it0[j].level || v[j].off != it0[j].off ||
                v[j].count != it0[j].count)
                return MBL_E_UNSUPPORTED;
    }
    // each case's ops of every recorded item, contiguous in the program's op table
    const size_t ni = it0.size();
    std::vector<LineOp> tab;
    std::vector<int> ooff(ni * n, 0), ocnt(ni * n, 0), omax(ni * n, 0);
    int max_nc = 0;
    for (int c = 0; c < n; c++)
        for (size_t j = 0; j < ni; j++) {
            if (it0[j].kind == 0) continue;
            ooff[j * n + c] = (int)tab.size();
            for (int o = it0[j].off; o < it0[j].off + it0[j].count; o++) {
                const LineOp &op = L->hops[o];
                if (case_of[op.dst_block] != c) continue;
                if (case_of[op.src_block] != c) return MBL_E_UNSUPPORTED;
                tab.push_back(op);
                ocnt[j * n + c]++;
                omax[j * n + c] = std::max(omax[j * n + c], op.n);
                if (op.type == OP_SPLINE) max_nc = std::max(max_nc, op.nc);
            }
        }
    if (!cases_spline_ok(max_nc)) return MBL_E_UNSUPPORTED;
    // items: [step][case][k]
    std::vector<CaseItem> items;
    int per_case = -1;
    for (int st = 0; st < period; st++)
        for (int c = 0; c < n; c++) {
            const size_t before = items.size();
            for (size_t j = 0; j < ni; j++) {
                const CycleItem &ci = steps[st][j];
                CaseItem x;
                memset(&x, 0, sizeof(x));
                x.pm = ci.pm;
                x.ms = ci.ms;
                x.xm = ci.xm;
                if (ci.kind == 0) {
                    x.kind = ITEM_COLLIDE;
                    x.a = t0[2 * c + ci.level];
                    x.b = tn[2 * c + ci.level];
                    if (x.b == 0) continue;
                } else {
                    if (ocnt[j * n + c] == 0) continue;
                    x.kind = ci.kind == 2 ? ITEM_SPLINE : ITEM_OPS;
                    x.a = ooff[j * n + c];
                    x.b = ocnt[j * n + c];
                    x.max_n = omax[j * n + c];
                }
                items.push_back(x);
            }
            CaseItem d;
            memset(&d, 0, sizeof(d));
            d.kind = ITEM_DRAG;
            d.a = par[st];
            items.push_back(d);
            const int cnt = (int)(items.size() - before);
            if (per_case >= 0 && cnt != per_case) return MBL_E_UNSUPPORTED;
            per_case = cnt;
        }
    // workgroups take the cases largest first
    std::vector<int> order(n);
    std::vector<double> work(n, 0.0);
    for (int a = 0; a < 5 * n; a++) work[case_of[a]] += (double)B->areas[a].nx * B->areas[a].ny * B->areas[a].m;
    for (int c = 0; c < n; c++) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    if (hipMalloc(&cp.d_items, sizeof(CaseItem) * items.size()) != hipSuccess ||
        hipMalloc(&cp.d_ops, sizeof(LineOp) * std::max<size_t>(1, tab.size())) != hipSuccess ||
        hipMalloc(&cp.d_case, sizeof(int) * n) != hipSuccess) {
        cp.release();
        return MBL_E_NOMEM;
    }
    bool ok = hipMemcpy(cp.d_items, items.data(), sizeof(CaseItem) * items.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && !tab.empty())
        ok = hipMemcpy(cp.d_ops, tab.data(), sizeof(LineOp) * tab.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok) ok = hipMemcpy(cp.d_case, order.data(), sizeof(int) * n, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        cp.release();
        return MBL_E_HIP;
    }
    cp.per_case = per_case;
    cp.period = period;
    cp.lds = cases_lds(max_nc);
    cp.gen = h->layout_gen;
    (void)pm0;
    return 0;
}

// nsteps coarse steps of every case in one launch (one workgroup per case), then the host
// state (plane pointer order, ring markers, force history) advanced as the steps leave it.
static int batch_step_persistent(mbl_batch *B, int nsteps)
{
    mbl_context *h = B->h;
    const int n = B->n;
    EvoArgs A{B->asp, B->ls3, B->ltc, B->lccp, B->areas.data(), B->pps.data(), B->lc.data(), B->lf.data(),
              B->lcc.data(), B->t3.data(), B->s3.data()};
    A.bc = B->bc.data();
    Layout *L = layout_for(h, A, 5 * n, 4 * n, 8 * n);
    if (!L) return h->last_error ? h->last_error : set_err(h, MBL_E_INVALID);
    mbl_batch::CaseProgram &cp = B->cp;
    uint32_t pm;
    uint64_t key;
    if (int rc = phase_key(L, B->pps.data(), B->t3.data(), 0, pm, key)) return set_err(h, rc);
    int q = -1;
    if (cp.gen == h->layout_gen)
        for (int i = 0; i < cp.period; i++)
            if (cp.key[i] == key) q = i;
    if (q < 0) {
        hipStreamSynchronize(h->stream);
        if (int rc = batch_case_program(B, L)) return set_err(h, rc);
        q = 0;
    }
    CaseArgs a;
    memset(&a, 0, sizeof(a));
    a.items = cp.d_items;
    a.per_case = cp.per_case;
    a.period = cp.period;
    a.step0 = q;
    a.ncases = n;
    a.nsteps = nsteps;
    a.case_of_wg = cp.d_case;
    a.descs = L->db;
    a.tile_blk = L->dtile;
    a.ndesc = L->nblocks;
    a.rows = L->rows;
    a.tpb = h->tpb;
    a.ops = cp.d_ops;
    a.drag[0] = B->d_drag[0];
    a.drag[1] = B->d_drag[1];
    a.hist = B->d_hist;
    a.head0 = B->head;
    a.cap = B->cap;
    if (int rc = launch_cases(h->stream, storage_mode(h->cfg), L->vx[0], a, cp.lds)) return set_err(h, rc);
    for (int st = 0; st < nsteps; st++) {
        const uint32_t fl = cp.flips[(q + st) % cp.period];
        for (int l = 0; l < L->LEVEL; l++)
            if (fl & (1u << l))
                for (int ar = L->ASP[l]; ar < L->ASP[l + 1]; ar++)
                    for (int k = 0; k < 9; k++) std::swap(B->pps[ar].F[k], B->pps[ar].f[k]);
    }
    if (nsteps > 0) {
        const std::vector<int> &mk = cp.markers[(q + nsteps - 1) % cp.period];
        for (size_t b = 0; b < B->t3.size(); b++)
            for (int p = 0; p < 3; p++) B->t3[b].marker[p] = mk[3 * b + p];
    }
    B->head = (int)((B->head + (long long)nsteps) % B->cap);
    B->nhist = (int)std::min<long long>((long long)B->nhist + nsteps, B->cap);
    return 0;
}

extern "C" int mbl_batch_set_persistent(mbl_batch_t B, int on)
{
    if (!B) return MBL_E_INVALID;
    B->persistent = on != 0;
    return 0;
}

extern "C" int mbl_batch_step(mbl_batch_t B, int nsteps)
{
    if (!B || nsteps < 0) return MBL_E_INVALID;
    mbl_context *h = B->h;
    if (h->last_error) return h->last_error;
    if (B->persistent) return batch_step_persistent(B, nsteps);
    EvoArgs A{B->asp, B->ls3, B->ltc, B->lccp, B->areas.data(), B->pps.data(), B->lc.data(), B->lf.data(),
              B->lcc.data(), B->t3.data(), B->s3.data()};
    A.bc = B->bc.data();
    const int mode = storage_mode(h->cfg);
    for (int t = 0; t < nsteps; t++) {
        evolution_impl(h, A, 0, 5 * B->n, 4 * B->n, 8 * B->n);
        if (h->last_error) return h->last_error;
        const int par = (B->pps[4 * B->n].F[0] == (Lbd *)B->fine_F0) ? 0 : 1;
        int rc = launch_drag_batch(h->stream, mode, B->d_drag[par], B->n, B->d_hist + 2 * (size_t)B->head * B->n);
        if (rc) return set_err(h, rc);
        B->head = (B->head + 1) % B->cap;
        B->nhist = std::min(B->nhist + 1, B->cap);
    }
    return 0;
}

extern "C" int mbl_batch_forces(mbl_batch_t B, int nt, double *cd, double *cl)
{
    if (!B) return MBL_E_INVALID;
    mbl_context *h = B->h;
    const int n = B->n;
    std::vector<double> hist((size_t)2 * n * B->cap);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(hist.data(), B->d_hist, sizeof(double) * hist.size(), hipMemcpyDeviceToHost));
    const int m = std::min(nt, B->nhist);
    for (int s = 0; s < m; s++) {
        int slot = (B->head - m + s + 2 * B->cap) % B->cap;
        for (int c = 0; c < n; c++) {
            double fx = hist[2 * ((size_t)slot * n + c)], fy = hist[2 * ((size_t)slot * n + c) + 1];
            if (cd) cd[(size_t)c * nt + s] = fx / B->norm[c];
            if (cl) cl[(size_t)c * nt + s] = fy / B->norm[c];
        }
    }
    return m;
}

// Block and case tables of the raster kernels and a staging buffer of
// n x H x W x 3 floats. Plane pointers are taken as they are now; the
// initialiser writes both plane sets, so their order does not matter.
static int batch_raster_prepare(mbl_batch *B, int H, int W)
{
    mbl_context *h = B->h;
    const int n = B->n;
    if (H < 2 || W < 2 || (size_t)H * W > (1u << 26)) return set_err(h, MBL_E_INVALID);
    if (!B->d_rblk) {
        std::vector<RasterBlock> rb(5 * n);
        std::vector<RasterCase> rc(n);
        for (int a = 0; a < 5 * n; a++) {
            RasterBlock &r = rb[a];
            memset(&r, 0, sizeof(r));
            const D2Q9 &pp = B->pps[a];
            for (int k = 0; k < 9; k++) {
                r.P[0][k] = pp.F[k];
                r.P[1][k] = pp.f[k];
            }
            r.flag = pp.solid;
            r.rho = pp.rho;
            r.ux = pp.ux;
            r.uy = pp.uy;
            r.bx = B->areas[a].base_x;
            r.by = B->areas[a].base_y;
            r.tau = B->areas[a].tau;
            r.m = B->areas[a].m;
            r.nx = B->areas[a].nx;
            r.ny = B->areas[a].ny;
            r.cas = a < 4 * n ? a / 4 : a - 4 * n;
            B->max_nodes = std::max(B->max_nodes, (size_t)r.nx * r.ny);
        }
        for (int c = 0; c < n; c++) {
            RasterCase &k = rc[c];
            k.nblk = 5;
            k.blk[0] = 4 * n + c; // the refined block first
            for (int q = 0; q < 4; q++) k.blk[q + 1] = 4 * c + q;
            k.LX = B->cs[c].LX;
            k.LY = B->cs[c].LY;
        }
        HIPCHK(h, hipMalloc(&B->d_rblk, sizeof(RasterBlock) * rb.size()));
        HIPCHK(h, hipMalloc(&B->d_rcase, sizeof(RasterCase) * rc.size()));
        HIPCHK(h, hipMemcpy(B->d_rblk, rb.data(), sizeof(RasterBlock) * rb.size(), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(B->d_rcase, rc.data(), sizeof(RasterCase) * rc.size(), hipMemcpyHostToDevice));
    }
    const size_t bytes = sizeof(float) * 3 * (size_t)n * H * W;
    if (bytes > B->raster_bytes) {
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (B->d_raster) hipFree(B->d_raster);
        B->d_raster = nullptr;
        B->raster_bytes = 0;
        HIPCHK(h, hipMalloc(&B->d_raster, bytes));
        B->raster_bytes = bytes;
    }
    return 0;
}

extern "C" int mbl_batch_raster(mbl_batch_t B, int H, int W, float *out)
{
    if (!B || !out) return MBL_E_INVALID;
    mbl_context *h = B->h;
    if (int rc = batch_raster_prepare(B, H, W)) return rc;
    int rc = launch_raster(h->stream, B->d_rblk, B->d_rcase, B->n, H, W, B->d_raster);
    if (rc) return set_err(h, rc);
    HIPCHK(h, hipMemcpyAsync(out, B->d_raster, sizeof(float) * 3 * (size_t)B->n * H * W, hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int mbl_batch_init_from_raster(mbl_batch_t B, int H, int W, const float *fields)
{
    if (!B || !fields) return MBL_E_INVALID;
    mbl_context *h = B->h;
    if (int rc = batch_raster_prepare(B, H, W)) return rc;
    HIPCHK(h, hipMemcpyAsync(B->d_raster, fields, sizeof(float) * 3 * (size_t)B->n * H * W, hipMemcpyHostToDevice,
                             h->stream));
    int rc = launch_init_raster(h->stream, storage_mode(h->cfg), B->d_rblk, 5 * B->n, B->max_nodes, B->d_rcase, H, W,
                                B->d_raster);
    if (rc) return set_err(h, rc);
    if ((rc = batch_ring_fill(B))) return set_err(h, rc);
    if ((rc = batch_seed_uprev(B))) return set_err(h, rc);
    // a restart: the force history starts again
    B->nhist = 0;
    B->head = 0;
    return 0;
}

extern "C" int mbl_batch_block(mbl_batch_t B, int c, int q, AREA *area, D2Q9 *pp)
{
    if (!B || c < 0 || c >= B->n || q < 0 || q > 4) return MBL_E_INVALID;
    const int a = q < 4 ? 4 * c + q : 4 * B->n + c;
    if (area) *area = B->areas[a];
    if (pp) *pp = B->pps[a];
    return 0;
}

extern "C" int mbl_batch_autotune(mbl_batch_t B, int trial_steps, int *vx_out, int *tpb_out, int *groups_out,
                                  double *ms_out)
{
    if (!B || trial_steps <= 0) return MBL_E_INVALID;
    mbl_context *h = B->h;
    int cu = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess && prop.multiProcessorCount > 0)
        cu = prop.multiProcessorCount;
    const int vxs[3] = {1, 2, 4}, tpbs[3] = {64, 128, 256};
    const int gs[4] = {0, 4 * cu, 8 * cu, 16 * cu};
    double best = 1e30;
    int bv = h->vx, bt = h->tpb, bg = h->groups;
    h->vx_set = true;
    for (int v : vxs)
        for (int t : tpbs)
            for (int g : gs) {
                if (B->persistent && g != 0) continue; // one workgroup per case: no grid size to pick
                h->vx = v;
                h->tpb = t;
                h->groups = g;
                drop_layouts(h);
                if (mbl_batch_step(B, 1)) return h->last_error; // builds tables and captures the graphs
                HIPCHK(h, hipEventRecord(h->ev0, h->stream));
                if (mbl_batch_step(B, trial_steps)) return h->last_error;
                HIPCHK(h, hipEventRecord(h->ev1, h->stream));
                HIPCHK(h, hipEventSynchronize(h->ev1));
                float ms = 0.f;
                HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
                double per = ms / trial_steps;
                if (per < best) {
                    best = per;
                    bv = v;
                    bt = t;
                    bg = g;
                }
            }
    h->vx = bv;
    h->tpb = bt;
    h->groups = bg;
    drop_layouts(h);
    if (vx_out) *vx_out = bv;
    if (tpb_out) *tpb_out = bt;
    if (groups_out) *groups_out = bg;
    if (ms_out) *ms_out = best;
    return 0;
}
