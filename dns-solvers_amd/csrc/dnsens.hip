// dnsens.hip -- host side of the ensemble entry points (include/dnsgpu.h).
//
// An ensemble is a set of case handles (members) that advance together on
// one stream.  Per step there are three launches for all members:
//     ens_step_kernel -> ens_reduce_kernel -> ens_finalize_kernel
// driven by device tables (members, tiles, finisher arguments).  A chunk of
// steps is captured once into a HIP graph and replayed while the member
// set is unchanged; adding or removing a member rebuilds the tables and the
// graph.  A member keeps its handle: buffers, masks, log ring and time
// accumulators are the handle's own, so every dnsgpu_* query works on it
// between dnsens_step calls.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "handle.h"
#include "lbm_ens_kernels.h"

namespace {

struct Member {
    dnsgpu* h;
    size_t bytes;  // footprint charged: the handle plus its tile records
    int ntiles;
    double* rec;   // tile records [ntiles][ens_record(Z)]
    dnsens_member_info info;
};

struct TuneEntry {
    int prec, zmax, tiles_log2, blocks_per_cu;
};
// shared by every ensemble of the process (ensembles may live on different
// host threads): guarded
std::vector<TuneEntry> g_ens_tune;
std::mutex g_ens_tune_mutex;

bool ens_tune_lookup(int prec, int zmax, int tiles_log2, int* bpc)
{
    std::lock_guard<std::mutex> lock(g_ens_tune_mutex);
    for (auto& t : g_ens_tune)
        if (t.prec == prec && t.zmax == zmax && t.tiles_log2 == tiles_log2) {
            if (bpc) *bpc = t.blocks_per_cu;
            return true;
        }
    return false;
}

int ilog2(int v)
{
    int r = 0;
    while ((1 << (r + 1)) <= v) r++;
    return r;
}

}  // namespace

struct dnsens {
    int precision;
    int device;  // the device current at create; members live on it
    size_t budget;
    size_t used;
    int flags;
    std::vector<Member> mem;
    hipStream_t stream;
    int ncu;
    // device tables and their capacities
    EnsMember* d_mem;
    FinalizeArgs* d_fin;
    EnsTile* d_tiles;
    double* d_probe;
    int cap_mem, cap_tiles;
    std::vector<EnsMember> h_mem;
    std::vector<FinalizeArgs> h_fin;
    std::vector<EnsTile> h_tiles;
    int ntiles, nchunks, zmax;
    int zslice;    // planes of one tile's plane sums (the largest ens_plane_rows)
    bool pieces;   // some member is in z-pieces (above 512 rows)
    bool shl_lds;  // the step kernel keeps the members' shifts in LDS
    size_t lds;
    int resident;       // step-kernel blocks resident at once for this lds
    int blocks_per_cu;  // grid = blocks_per_cu * ncu (capped)
    int grid;
    bool tables_stale;  // host tables must be rebuilt and uploaded
    bool layout_changed;  // the graph no longer matches the launch sequence
    hipGraph_t graph;
    hipGraphExec_t exec;
    int graph_steps;
    int last_chunk;
    hipEvent_t ev0, ev1;
};

static int default_blocks_per_cu(int precision)
{
    // about 8 (FP32) or 4 (FP64) waves per CU, the single-case optimum
    return precision == DNSGPU_FP64 ? 1 : 2;
}

extern "C" int dnsens_create(int precision, size_t hbm_budget, int flags, dnsens** out)
{
    if (!out) return set_err(DNSGPU_EARG, "null argument");
    *out = nullptr;
    if (precision != DNSGPU_FP64 && precision != DNSGPU_FP32 && precision != DNSGPU_FP16)
        return set_err(DNSGPU_EARG, "precision %d not built", precision);
    dnsens* e = new (std::nothrow) dnsens();
    if (!e) return set_err(DNSGPU_ENOMEM, "host allocation failed");
    e->precision = precision;
    e->flags = flags;
    auto fail = [&](int code) {
        dnsens_destroy(e);
        return code;
    };
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&e->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return fail(set_err(DNSGPU_EHIP, "device query failed"));
    e->device = dev;
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail(set_err(DNSGPU_EHIP, "hipMemGetInfo failed"));
        hbm_budget = fr / 10 * 9;
    }
    e->budget = hbm_budget;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(DNSGPU_EHIP, "stream creation failed"));
    if (hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess)
        return fail(set_err(DNSGPU_EHIP, "event creation failed"));
    e->blocks_per_cu = default_blocks_per_cu(precision);
    const char* env = getenv("DNSENS_BLOCKS_PER_CU");
    if (env && atoi(env) > 0) e->blocks_per_cu = atoi(env);
    e->tables_stale = true;
    e->layout_changed = true;
    *out = e;
    return 0;
}

static void drop_graph(dnsens* e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    e->graph_steps = 0;
}

static void free_member(Member& m)
{
    if (m.rec) (void)hipFree(m.rec);
    m.rec = nullptr;
}

extern "C" void dnsens_destroy(dnsens* e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    for (auto& m : e->mem) {
        m.h->owner = nullptr;
        free_member(m);
    }
    void* ptrs[] = {e->d_mem, e->d_fin, e->d_tiles, e->d_probe};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

static int find_member(const dnsens* e, const dnsgpu* h)
{
    for (size_t i = 0; i < e->mem.size(); i++)
        if (e->mem[i].h == h) return (int)i;
    return -1;
}

// Column tiles of a member, and its tiles (records): one per column tile and
// z-piece.
static int col_tiles_of(const dnsgpu* h)
{
    const int cpt = ens_tile_cols(h->g.ncols);
    return (h->g.ncols + cpt - 1) / cpt;
}

static int tiles_of(const dnsgpu* h) { return col_tiles_of(h) * h->g.nzp; }

// Bytes a member adds on top of its handle: its tile records.
static size_t extra_bytes(const dnsgpu* h) { return sizeof(double) * (size_t)tiles_of(h) * ens_record(h->g.Z); }

extern "C" size_t dnsens_member_bytes(const dnsgpu* h)
{
    return h ? h->dev_bytes + extra_bytes(h) : 0;
}

extern "C" size_t dnsens_footprint(const dns_case* c)
{
    const size_t fp = dnsgpu_footprint(c);
    if (!fp) return 0;
    const int ncols = c->xdv * c->ydv, cpt = ens_tile_cols(ncols);
    const int pieces = z_pieces((c->zdv + kWave - 1) / kWave);
    return fp + sizeof(double) * (size_t)((ncols + cpt - 1) / cpt) * pieces * ens_record(c->zdv);
}

extern "C" int dnsens_add(dnsens* e, dnsgpu* h)
{
    if (!e || !h) return set_err(DNSGPU_EARG, "null argument");
    if (h->c.precision != e->precision)
        return set_err(DNSGPU_EARG, "member precision %d differs from the ensemble's %d", h->c.precision,
                       e->precision);
    if (h->device != e->device)
        return set_err(DNSGPU_EARG, "member on device %d, the ensemble on device %d", h->device, e->device);
    if (find_member(e, h) >= 0) return set_err(DNSGPU_EARG, "handle is already a member");
    if (h->owner) return set_err(DNSGPU_EARG, "handle is a member of another ensemble");
    if ((int)e->mem.size() >= DNSENS_MAX_MEMBERS)
        return set_err(DNSGPU_EARG, "ensemble holds %d members already", DNSENS_MAX_MEMBERS);
    const size_t need = dnsens_member_bytes(h);
    if (e->used + need > e->budget)
        return set_err(DNSGPU_ENOMEM, "member needs %zu bytes, %zu of %zu in use", need, e->used, e->budget);
    // the ensemble's FP16 step bounces at bands on the write side: a state
    // from the two-row single-case step first gets the pushed values; then
    // the handle's own queued work
    if (int rc = band_pushes(h)) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    Member m;
    memset(&m, 0, sizeof m);
    m.h = h;
    m.bytes = need;
    m.ntiles = tiles_of(h);
    if (hipMalloc(&m.rec, extra_bytes(h)) != hipSuccess) {
        free_member(m);
        return set_err(DNSGPU_ENOMEM, "tile records of %d tiles", m.ntiles);
    }
    auto fail = [&](int code) {
        free_member(m);
        return code;
    };
    // a state fresh from a load or an init holds post-stream populations:
    // its first step (local gather) runs through the single-case kernels
    DevState st;
    if (hipMemcpyAsync(&st, h->st, sizeof st, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return fail(set_err(DNSGPU_EHIP, "reading the member's state failed"));
    double stress = 0.0;
    if (h->poststream) {
        int rc = do_step(h, true, h->ring, (long long)st.steps);
        if (rc) return fail(rc);
        LogRow row;
        if (hipMemcpyAsync(&st, h->st, sizeof st, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipMemcpyAsync(&row, h->ring, sizeof row, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess)
            return fail(set_err(DNSGPU_EHIP, "reading the member's first step failed"));
        stress = row.stress;
    }
    // the member's state as of now (dnsens_info before its first dnsens_step)
    m.info.steps = st.steps;
    m.info.n_rows = st.n_rows;
    m.info.mean_pgrad = st.n_rows > 0 ? st.s_pgrad / st.n_rows : 0.0;
    m.info.force = st.G;
    m.info.last_stress = stress;
    m.info.finite = std::isfinite(st.G) && std::isfinite(stress) && std::isfinite(st.s_pgrad);
    h->owner = e;
    e->mem.push_back(m);
    e->used += need;
    e->tables_stale = true;
    e->layout_changed = true;
    return 0;
}

extern "C" int dnsens_remove(dnsens* e, dnsgpu* h)
{
    if (!e || !h) return set_err(DNSGPU_EARG, "null argument");
    const int i = find_member(e, h);
    if (i < 0) return set_err(DNSGPU_EARG, "handle is not a member");
    HIPCHK(hipStreamSynchronize(e->stream));
    e->used -= e->mem[i].bytes;
    e->mem[i].h->owner = nullptr;
    free_member(e->mem[i]);
    e->mem.erase(e->mem.begin() + i);
    e->tables_stale = true;
    e->layout_changed = true;
    return 0;
}

extern "C" int dnsens_count(const dnsens* e) { return e ? (int)e->mem.size() : 0; }
extern "C" size_t dnsens_bytes_used(const dnsens* e) { return e ? e->used : 0; }
extern "C" size_t dnsens_budget(const dnsens* e) { return e ? e->budget : 0; }

extern "C" int dnsens_bind_thread(const dnsens* e)
{
    if (!e) return set_err(DNSGPU_EARG, "null argument");
    HIPCHK(hipSetDevice(e->device));
    return 0;
}

template <typename T>
static int step_occupancy(bool shl, bool pieces, size_t lds, int* per_cu)
{
    if (pieces)
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ens_step_kernel<T, false, true>,
                                                            kWave * kEnsWaves, lds));
    else if (shl)
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ens_step_kernel<T, true>, kWave * kEnsWaves, lds));
    else
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ens_step_kernel<T, false>, kWave * kEnsWaves, lds));
    return 0;
}

static int pick_grid(dnsens* e)
{
    const int wave_blocks = (e->ntiles + kEnsWaves - 1) / kEnsWaves;  // one wave per tile
    e->grid = std::max(1, std::min(e->blocks_per_cu * e->ncu, std::min(e->resident, wave_blocks)));
    return 0;
}

// Host tables from the member list (each member's current buffer is src),
// uploaded on the ensemble stream; device capacity grows as needed.
static int upload_tables(dnsens* e)
{
    const int nm = (int)e->mem.size();
    e->h_mem.resize(nm);
    e->h_fin.resize(nm);
    e->h_tiles.clear();
    int chunk = 0, zmax = 0, zslice = 0;
    bool pieces = false;
    for (int i = 0; i < nm; i++) {
        Member& m = e->mem[i];
        dnsgpu* h = m.h;
        EnsMember& d = e->h_mem[i];
        d.a = step_args(h, h->buf[h->cur], h->buf[1 - h->cur], m.ntiles);
        d.a.bnbr = h->bnbr;  // zero records without bands (the ensemble kernel reads them always)
        d.a.part = m.rec;
        d.a.gpart = nullptr;
        d.planes = h->planes;
        d.chunk0 = chunk;
        d.nchunks = (ens_record(h->g.Z) + kWave - 1) / kWave;
        d.pieces = h->g.nzp;
        d.cpp = h->g.cpp;
        chunk += d.nchunks;
        FinalizeArgs F = fin_args(h, true, h->ring, 0, 0);
        F.cntr_dev = 1;
        e->h_fin[i] = F;
        const int S = col_tiles_of(h), P = h->g.nzp;
        for (int t = 0; t < S; t++)
            for (int pc = 0; pc < P; pc++) {
                EnsTile tl{};
                tl.m = i;
                // columns t, t + S, ... below ncols: at most ens_tile_cols of them
                tl.col0 = t;
                tl.stride = S;
                tl.ncol = (h->g.ncols - t + S - 1) / S;
                tl.lt = t * P + pc;
                tl.piece = pc;
                e->h_tiles.push_back(tl);
            }
        zmax = std::max(zmax, h->g.Z);
        zslice = std::max(zslice, ens_plane_rows(h->g.Z));
        pieces |= P > 1;
    }
    e->ntiles = (int)e->h_tiles.size();
    e->nchunks = chunk;
    if (nm > e->cap_mem) {
        if (e->d_mem) (void)hipFree(e->d_mem);
        if (e->d_fin) (void)hipFree(e->d_fin);
        if (e->d_probe) (void)hipFree(e->d_probe);
        e->d_mem = nullptr;
        e->d_fin = nullptr;
        e->d_probe = nullptr;
        e->cap_mem = 0;
        const int cap = std::max(16, nm * 2);
        if (hipMalloc(&e->d_mem, sizeof(EnsMember) * cap) != hipSuccess ||
            hipMalloc(&e->d_fin, sizeof(FinalizeArgs) * cap) != hipSuccess ||
            hipMalloc(&e->d_probe, sizeof(double) * 5 * cap) != hipSuccess)
            return set_err(DNSGPU_ENOMEM, "ensemble tables");
        e->cap_mem = cap;
        e->layout_changed = true;
    }
    if (e->ntiles > e->cap_tiles) {
        if (e->d_tiles) (void)hipFree(e->d_tiles);
        e->d_tiles = nullptr;
        e->cap_tiles = 0;
        const int cap = std::max(4096, e->ntiles + e->ntiles / 2);
        if (hipMalloc(&e->d_tiles, sizeof(EnsTile) * cap) != hipSuccess)
            return set_err(DNSGPU_ENOMEM, "tile table of %d tiles", cap);
        e->cap_tiles = cap;
        e->layout_changed = true;
    }
    if (nm) {
        HIPCHK(hipMemcpyAsync(e->d_mem, e->h_mem.data(), sizeof(EnsMember) * nm, hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(e->d_fin, e->h_fin.data(), sizeof(FinalizeArgs) * nm, hipMemcpyHostToDevice,
                              e->stream));
        HIPCHK(hipMemcpyAsync(e->d_tiles, e->h_tiles.data(), sizeof(EnsTile) * e->ntiles, hipMemcpyHostToDevice,
                              e->stream));
        // the host tables are rewritten by the next upload
        HIPCHK(hipStreamSynchronize(e->stream));
    }
    if (zmax != e->zmax || pieces != e->pieces || e->resident == 0) {
        e->zmax = zmax;
        e->zslice = std::max(zslice, 1);
        e->pieces = pieces;
        // the shifts in LDS when they fit (never with z-pieces: a member of
        // more than 512 rows), the plane sums of a tile always
        e->shl_lds = !pieces && ens_lds_bytes(e->zslice, std::max(zmax, 1), true) <= kEnsLdsMax;
        e->lds = ens_lds_bytes(e->zslice, std::max(zmax, 1), e->shl_lds);
        int per = 0;
        int rc = with_storage(e->precision, [&](auto t) {
            return step_occupancy<decltype(t)>(e->shl_lds, e->pieces, e->lds, &per);
        });
        if (rc) return rc;
        e->resident = std::max(1, per) * e->ncu;
        e->layout_changed = true;
    }
    if (!getenv("DNSENS_BLOCKS_PER_CU")) {
        e->blocks_per_cu = default_blocks_per_cu(e->precision);
        ens_tune_lookup(e->precision, e->zmax, ilog2(std::max(1, e->ntiles)), &e->blocks_per_cu);
    }
    const int old_grid = e->grid;
    pick_grid(e);
    if (e->grid != old_grid) e->layout_changed = true;
    e->tables_stale = false;
    return 0;
}

template <typename T>
static void enqueue_step_only(dnsens* e, int phase)
{
    const int zs = e->zslice, zm = std::max(e->zmax, 1);
    if (e->pieces)
        hipLaunchKernelGGL((ens_step_kernel<T, false, true>), dim3(e->grid), dim3(kWave * kEnsWaves), e->lds,
                           e->stream, e->d_mem, e->d_tiles, e->ntiles, zs, zm, phase);
    else if (e->shl_lds)
        hipLaunchKernelGGL((ens_step_kernel<T, true>), dim3(e->grid), dim3(kWave * kEnsWaves), e->lds, e->stream,
                           e->d_mem, e->d_tiles, e->ntiles, zs, zm, phase);
    else
        hipLaunchKernelGGL((ens_step_kernel<T, false>), dim3(e->grid), dim3(kWave * kEnsWaves), e->lds, e->stream,
                           e->d_mem, e->d_tiles, e->ntiles, zs, zm, phase);
}

// k steps of every member: step, reduction, finisher per step; log row of
// step s in ring slot s.
template <typename T>
static int enqueue_steps(dnsens* e, int k)
{
    const int nm = (int)e->mem.size();
    for (int s = 0; s < k; s++) {
        enqueue_step_only<T>(e, s & 1);
        hipLaunchKernelGGL(ens_reduce_kernel, dim3(e->nchunks), dim3(kWave * kEnsReduceWaves), 0, e->stream, e->d_mem,
                           nm);
        hipLaunchKernelGGL(ens_finalize_kernel, dim3(nm), dim3(1024), 0, e->stream, e->d_fin, s);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

static int run_chunk(dnsens* e, int k)
{
    const bool use_graph = !(e->flags & DNSENS_NO_GRAPH);
    if (!use_graph) return with_storage(e->precision, [&](auto t) { return enqueue_steps<decltype(t)>(e, k); });
    if (e->layout_changed || e->graph_steps != k || !e->exec) {
        drop_graph(e);
        HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = with_storage(e->precision, [&](auto t) { return enqueue_steps<decltype(t)>(e, k); });
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(e->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ce != hipSuccess) return set_err(DNSGPU_EHIP, "graph capture failed: %s", hipGetErrorString(ce));
        e->graph = g;
        HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
        e->graph_steps = k;
        e->layout_changed = false;
    }
    HIPCHK(hipGraphLaunch(e->exec, e->stream));
    return 0;
}

static int probe(dnsens* e)
{
    const int nm = (int)e->mem.size();
    if (!nm) return 0;
    hipLaunchKernelGGL(ens_probe_kernel, dim3((nm + 63) / 64), dim3(64), 0, e->stream, e->d_fin, nm,
                       e->last_chunk - 1, e->d_probe);
    HIPCHK(hipGetLastError());
    std::vector<double> p(5 * (size_t)nm);
    HIPCHK(hipMemcpyAsync(p.data(), e->d_probe, sizeof(double) * p.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int i = 0; i < nm; i++) {
        dnsens_member_info& in = e->mem[i].info;
        in.steps = p[5 * i + 0];
        in.n_rows = p[5 * i + 1];
        in.mean_pgrad = in.n_rows > 0 ? p[5 * i + 2] / in.n_rows : 0.0;
        in.force = p[5 * i + 3];
        in.last_stress = p[5 * i + 4];
        if (!(isfinite(in.force) && isfinite(in.last_stress) && isfinite(p[5 * i + 2]))) in.finite = 0;
    }
    return 0;
}

static int maybe_autotune(dnsens* e);

// Enqueue nsteps steps of every member in chunks of at most kRingRows.
static int step_chunks(dnsens* e, int nsteps)
{
    int done = 0;
    while (done < nsteps) {
        const int k = std::min(nsteps - done, kRingRows);
        if (e->tables_stale) {
            int rc = upload_tables(e);
            if (rc) return rc;
            rc = maybe_autotune(e);
            if (rc) return rc;
        }
        int rc = run_chunk(e, k);
        if (rc) return rc;
        for (auto& m : e->mem) {
            if (k & 1) m.h->cur = 1 - m.h->cur;
            m.h->poststream = 0;
        }
        // an odd chunk swaps every member's buffers: new tables, same layout
        if (k & 1) e->tables_stale = true;
        e->last_chunk = k;
        done += k;
    }
    return 0;
}

extern "C" int dnsens_step(dnsens* e, int nsteps)
{
    if (!e) return set_err(DNSGPU_EARG, "null ensemble");
    if (nsteps < 0) return set_err(DNSGPU_EARG, "nsteps < 0");
    if (e->mem.empty() || nsteps == 0) return 0;
    int rc = step_chunks(e, nsteps);
    if (rc) return rc;
    return probe(e);
}

extern "C" int dnsens_info(dnsens* e, const dnsgpu* h, dnsens_member_info* out)
{
    if (!e || !h || !out) return set_err(DNSGPU_EARG, "null argument");
    const int i = find_member(e, h);
    if (i < 0) return set_err(DNSGPU_EARG, "handle is not a member");
    *out = e->mem[i].info;
    return 0;
}

extern "C" int dnsens_rows(dnsens* e, const dnsgpu* h, dns_steplog* rows, int n)
{
    if (!e || !h || !rows) return set_err(DNSGPU_EARG, "null argument");
    const int i = find_member(e, h);
    if (i < 0) return set_err(DNSGPU_EARG, "handle is not a member");
    if (n < 0 || n > e->last_chunk)
        return set_err(DNSGPU_EARG, "%d rows asked, the last dnsens_step chunk kept %d", n, e->last_chunk);
    static_assert(sizeof(LogRow) == sizeof(dns_steplog), "log row layout");
    HIPCHK(hipMemcpyAsync(rows, h->ring + (e->last_chunk - n), sizeof(LogRow) * n, hipMemcpyDeviceToHost,
                          e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return 0;
}

// ---------------------------------------------------------------- autotuner

// Time trial launches of the step kernel at each candidate grid.  The
// trials read each member's current buffer and write the other one, which
// the next real step overwrites: the state, the partials' next use and the
// time accumulators are unaffected.
template <typename T>
static int tune_grid(dnsens* e, int trials, int* best_bpc, double* best_ms)
{
    const int cands[] = {1, 2, 3, 4, 6, 8};
    double bt = 1e30;
    int bb = e->blocks_per_cu;
    int last = -1;
    for (int bpc : cands) {
        e->blocks_per_cu = bpc;
        pick_grid(e);
        if (e->grid == last) continue;
        last = e->grid;
        for (int r = 0; r <= trials; r++) {
            if (r == 1) HIPCHK(hipEventRecord(e->ev0, e->stream));
            enqueue_step_only<T>(e, 0);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e->ev1, e->stream));
        HIPCHK(hipEventSynchronize(e->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        if (ms / trials < bt) {
            bt = ms / trials;
            bb = bpc;
        }
    }
    e->blocks_per_cu = bb;
    pick_grid(e);
    *best_bpc = bb;
    *best_ms = bt;
    return 0;
}

static int tune_now(dnsens* e, int trials, int* best_blocks, double* best_ms)
{
    if (e->tables_stale) {
        int rc = upload_tables(e);
        if (rc) return rc;
    }
    int bpc = e->blocks_per_cu;
    double ms = -1.0;
    if (!e->mem.empty()) {
        int rc = with_storage(e->precision, [&](auto t) { return tune_grid<decltype(t)>(e, trials, &bpc, &ms); });
        if (rc) return rc;
        std::lock_guard<std::mutex> lock(g_ens_tune_mutex);
        g_ens_tune.push_back({e->precision, e->zmax, ilog2(std::max(1, e->ntiles)), bpc});
    }
    e->layout_changed = true;
    if (best_blocks) *best_blocks = e->grid;
    if (best_ms) *best_ms = ms;
    return 0;
}

static int maybe_autotune(dnsens* e)
{
    if (!(e->flags & DNSENS_AUTOTUNE) || getenv("DNSENS_BLOCKS_PER_CU")) return 0;
    if (ens_tune_lookup(e->precision, e->zmax, ilog2(std::max(1, e->ntiles)), nullptr)) return 0;
    return tune_now(e, 3, nullptr, nullptr);
}

extern "C" int dnsens_autotune(dnsens* e, int trial_steps, int* best_blocks, double* best_ms)
{
    if (!e || trial_steps < 1) return set_err(DNSGPU_EARG, "bad argument");
    return tune_now(e, trial_steps, best_blocks, best_ms);
}

extern "C" int dnsens_time_steps(dnsens* e, int nsteps, double* ms)
{
    if (!e || nsteps < 1 || !ms) return set_err(DNSGPU_EARG, "bad argument");
    if (e->mem.empty()) return set_err(DNSGPU_EARG, "empty ensemble");
    if (e->tables_stale) {
        int rc = upload_tables(e);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(e->ev0, e->stream));
    int rc = step_chunks(e, nsteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(e->ev1, e->stream));
    HIPCHK(hipEventSynchronize(e->ev1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e->ev0, e->ev1));
    *ms = t;
    return probe(e);
}

extern "C" int dnsens_time_kernels(dnsens* e, int nsteps, double* ms_step_kernels, double* ms_total)
{
    if (!e || nsteps < 1) return set_err(DNSGPU_EARG, "bad argument");
    if (e->mem.empty()) return set_err(DNSGPU_EARG, "empty ensemble");
    if (e->tables_stale) {
        int rc = upload_tables(e);
        if (rc) return rc;
    }
    const int nm = (int)e->mem.size();
    std::vector<hipEvent_t> ev(2 * (size_t)nsteps);
    for (auto& x : ev) HIPCHK(hipEventCreate(&x));
    HIPCHK(hipEventRecord(e->ev0, e->stream));
    for (int s = 0; s < nsteps; s++) {
        HIPCHK(hipEventRecord(ev[2 * s], e->stream));
        with_storage(e->precision, [&](auto t) { enqueue_step_only<decltype(t)>(e, s & 1); });
        HIPCHK(hipEventRecord(ev[2 * s + 1], e->stream));
        hipLaunchKernelGGL(ens_reduce_kernel, dim3(e->nchunks), dim3(kWave * kEnsReduceWaves), 0, e->stream, e->d_mem,
                           nm);
        hipLaunchKernelGGL(ens_finalize_kernel, dim3(nm), dim3(1024), 0, e->stream, e->d_fin, s);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e->ev1, e->stream));
    HIPCHK(hipEventSynchronize(e->ev1));
    double tk = 0.0;
    for (int s = 0; s < nsteps; s++) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]));
        tk += ms;
    }
    float wall = 0.f;
    HIPCHK(hipEventElapsedTime(&wall, e->ev0, e->ev1));
    for (auto& x : ev) (void)hipEventDestroy(x);
    for (auto& m : e->mem) {
        if (nsteps & 1) m.h->cur = 1 - m.h->cur;
        m.h->poststream = 0;
    }
    if (nsteps & 1) e->tables_stale = true;
    e->last_chunk = nsteps;
    if (ms_step_kernels) *ms_step_kernels = tk;
    if (ms_total) *ms_total = wall;
    return probe(e);
}

extern "C" int dnsens_grid(const dnsens* e, int* blocks, int* tiles)
{
    if (!e) return set_err(DNSGPU_EARG, "null ensemble");
    if (blocks) *blocks = e->grid;
    if (tiles) *tiles = e->ntiles;
    return 0;
}
