// dnsgpu.hip -- host side of the C ABI declared in include/dnsgpu.h.
//
// One handle = one channel case resident on one GPU: two population
// buffers (A-B pull streaming), the two wall masks, the partial-sum
// buffers of the fixed-order reductions, a ring of per-step log rows and
// one HIP stream.  The per-step sequence is
//     step_kernel -> reduce_rows_kernel -> finalize_kernel
// all on the handle's stream; the host only reads back log rows.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "handle.h"

using namespace lbm;

static thread_local char g_err[512] = "";

int set_err(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char* dnsgpu_last_error(void) { return g_err; }

// Staging of dnsgpu_fields_noalloc: one static array per GPU the library runs
// on (each device holds its own copy, at its own address), held until the
// process ends.  In host code `scratch` names a host-side stand-in, never
// the device memory: its device address is resolved once per device with
// hipGetSymbolAddress (scratch_address, called from dnsgpu_create with the
// handle's device current) and only that address is passed to kernels and
// copies.  Every handle on a device shares the array; calls are serialised
// per device by g_scratch_mutex.
constexpr size_t kScratchDoubles = DNSGPU_FIELDS_NOALLOC_MAX_DOUBLES;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_table_mutex;
std::mutex g_scratch_mutex[kMaxDevices];
}  // namespace

// Device address of `scratch` on device dev (the current device), resolved
// on the first call for that device.
static int scratch_address(int dev, void** out)
{
    if (dev < 0 || dev >= kMaxDevices) return set_err(DNSGPU_EARG, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_table_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}


// Two rows per lane: FP16 storage and an even number of rows, one z-piece
// (up to 512 rows).
static bool step_rows_ok(int precision, int Z)
{
    return precision == DNSGPU_FP16 && Z >= 2 && Z % 2 == 0 && Z <= kWave * kMaxWaves;
}

// Blocks of the step kernel per CU at W waves per block, `rows` rows per
// lane and nzp z-pieces (the statistics variant with bands: the largest
// register set).
static hipError_t step_occupancy(int precision, int Z, int W, int rows, int nzp, int* per)
{
    const size_t lds = step_lds_bytes(Z, nzp, W, true);
    if (rows == 2)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(per, step_kernel<_Float16, false, true, true, 2>,
                                                            kWave * W, lds);
    const bool pieces = nzp > 1;
    return with_storage(precision, [&](auto t) {
        if (pieces)
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(
                per, step_kernel<decltype(t), false, true, true, 1, true>, kWave * W, lds);
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(per, step_kernel<decltype(t), false, true, true>,
                                                            kWave * W, lds);
    });
}

template <typename K>
static int occupancy_blocks(K kernel, int block, int* out)
{
    int dev = 0, ncu = 0, per = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0));
    if (per < 1) per = 1;
    *out = per * ncu;
    return 0;
}

static int stage_reserve(dnsgpu* h, size_t bytes)
{
    if (h->stage_bytes >= bytes) return 0;
    if (h->stage) (void)hipFree(h->stage);
    h->stage = nullptr;
    h->stage_bytes = 0;
    if (hipMalloc(&h->stage, bytes) != hipSuccess)
        return set_err(DNSGPU_ENOMEM, "staging allocation of %zu bytes failed", bytes);
    h->stage_bytes = bytes;
    return 0;
}

StepArgs step_args(dnsgpu* h, const void* src, void* dst, int nblk)
{
    StepArgs A;
    A.g = h->g;
    A.src = src;
    A.dst = dst;
    A.mbot = h->mbot;
    A.mtop = h->mtop;
    A.wword = h->wword;
    A.bands = h->has_bands ? h->bands : nullptr;
    A.bnbr = h->has_bands ? h->bnbr : nullptr;
    A.cnbr = h->cnbr;
    A.omega = h->omega;
    A.G = &h->st->G;
    A.shift = h->shift;
    A.part = h->part;
    A.gpart = h->gpart;
    A.nblk = nblk;
    return A;
}

FinalizeArgs fin_args(dnsgpu* h, bool stats, LogRow* row, long long cntr, int from_moments)
{
    FinalizeArgs F;
    F.planes = stats ? h->planes : nullptr;
    F.globals = h->planes + (size_t)kNQ * h->g.Z;
    F.shift = h->shift;
    F.st = h->st;
    F.row = row;
    F.X = h->g.X;
    F.Y = h->g.Y;
    F.Z = h->g.Z;
    F.forcing = h->c.forcing;
    F.ub = h->c.ub;
    F.F = h->c.force;
    F.mu = h->mu;
    F.TM = (double)h->c.TSTR;
    F.dt = h->dt;
    F.cntr = cntr;
    F.from_moments = from_moments;
    F.nf = h->has_bands ? h->nf : nullptr;
    F.nfluid = h->has_bands ? h->nfluid : (double)h->g.X * h->g.Y * h->g.Z;
    F.cntr_dev = 0;
    return F;
}

// Partials -> plane sums and globals in h->planes (fixed order): nplane
// partials per plane sum, nglobal per global sum.
static int reduce_partials(dnsgpu* h, bool stats, int nplane, int nglobal)
{
    const int Z = h->g.Z;
    if (stats)
        hipLaunchKernelGGL(reduce_rows_kernel, dim3(kNQ * Z), dim3(256), 0, h->stream, h->part, nplane, h->planes);
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(kNG), dim3(256), 0, h->stream, h->gpart, nglobal,
                       h->planes + (size_t)kNQ * Z);
    HIPCHK(hipGetLastError());
    return 0;
}

int band_pushes(dnsgpu* h)
{
    if (!h->rs_state) return 0;
    h->rs_state = 0;
    if (!h->has_bands || h->poststream) return 0;
    const long long n = (long long)h->g.ncols * h->g.Z;
    const int nb = (int)std::min<long long>((n + 255) / 256, 8LL * h->ncu);
    int rc = with_storage(h->c.precision, [&](auto t) {
        using T = decltype(t);
        hipLaunchKernelGGL(band_push_kernel<T>, dim3(nb), dim3(256), 0, h->stream, (T*)h->buf[h->cur], h->bands,
                           h->g.X, h->g.Y, h->g.Z);
        return 0;
    });
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return 0;
}

// The step kernel alone, on the handle's stream.  A write-side band step
// (FP16 one row per lane) of a state from a read-side one first gets the
// pushed values it pulls (band_pushes).
template <typename T>
static void enqueue_step_kernel(dnsgpu* h, bool stats)
{
    if constexpr (std::is_same<T, _Float16>::value) {
        if (h->step_rows != 2) (void)band_pushes(h);
    }
    const void* src = h->buf[h->cur];
    void* dst = h->buf[1 - h->cur];
    StepArgs A = step_args(h, src, dst, h->step_nblk);
    // the step's own z-pieces (a launch-shape choice)
    A.g.nzp = h->step_nzp;
    A.g.cpp = (h->g.nzc + h->step_nzp - 1) / h->step_nzp;
    dim3 grid(h->step_nblk), block(kWave * h->step_w);
    const size_t lds = step_lds_bytes(h->g.Z, h->step_nzp, h->step_w, stats);
    if constexpr (std::is_same<T, _Float16>::value) {
        if (h->step_rows == 2 && !h->poststream) {
#define STEP_R2(S, B) hipLaunchKernelGGL((step_kernel<T, false, S, B, 2>), grid, block, lds, h->stream, A)
            if (h->has_bands) {
                if (stats) STEP_R2(true, true);
                else STEP_R2(false, true);
            } else {
                if (stats) STEP_R2(true, false);
                else STEP_R2(false, false);
            }
#undef STEP_R2
            return;
        }
    }
#define STEP_VARIANT(F, S, B)                                                                      \
    do {                                                                                           \
        if (h->step_nzp > 1)                                                                       \
            hipLaunchKernelGGL((step_kernel<T, F, S, B, 1, true>), grid, block, lds, h->stream, A); \
        else                                                                                       \
            hipLaunchKernelGGL((step_kernel<T, F, S, B>), grid, block, lds, h->stream, A);         \
    } while (0)
    if (h->has_bands) {
        if (h->poststream) {
            if (stats) STEP_VARIANT(true, true, true);
            else STEP_VARIANT(true, false, true);
        } else {
            if (stats) STEP_VARIANT(false, true, true);
            else STEP_VARIANT(false, false, true);
        }
    } else {
        if (h->poststream) {
            if (stats) STEP_VARIANT(true, true, false);
            else STEP_VARIANT(true, false, false);
        } else {
            if (stats) STEP_VARIANT(false, true, false);
            else STEP_VARIANT(false, false, false);
        }
    }
#undef STEP_VARIANT
}

template <typename T>
static int launch_step(dnsgpu* h, bool stats, LogRow* row, long long cntr)
{
    // the two-row FP16 step bounces at bands on the read side
    const bool rs = std::is_same<T, _Float16>::value && h->step_rows == 2 && !h->poststream && h->has_bands;
    enqueue_step_kernel<T>(h, stats);
    HIPCHK(hipGetLastError());
    int rc = reduce_partials(h, stats, h->step_nblk / h->step_nzp, h->step_nblk);
    if (rc) return rc;
    FinalizeArgs F = fin_args(h, stats, row, cntr, 0);
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(1024), 0, h->stream, F);
    HIPCHK(hipGetLastError());
    h->cur = 1 - h->cur;
    h->poststream = 0;
    h->rs_state = rs ? 1 : 0;
    return 0;
}

int do_step(dnsgpu* h, bool stats, LogRow* row, long long cntr)
{
    return with_storage(h->c.precision, [&](auto t) { return launch_step<decltype(t)>(h, stats, row, cntr); });
}

// Derive the force of the next step and the plane-mean shifts from the
// current state (encalc.c:29-121 on the loaded/initial state).
template <typename T>
static int launch_moments(dnsgpu* h)
{
    StepArgs A = step_args(h, h->buf[h->cur], nullptr, h->nblk);
    if (h->poststream)
        hipLaunchKernelGGL((moments_kernel<T, true>), dim3(h->nblk, h->g.nzp), dim3(h->block), 0, h->stream, A);
    else
        hipLaunchKernelGGL((moments_kernel<T, false>), dim3(h->nblk, h->g.nzp), dim3(h->block), 0, h->stream, A);
    HIPCHK(hipGetLastError());
    int rc = reduce_partials(h, true, h->nblk, h->nblk * h->g.nzp);
    if (rc) return rc;
    FinalizeArgs F = fin_args(h, true, nullptr, 0, 1);
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(1024), 0, h->stream, F);
    HIPCHK(hipGetLastError());
    return 0;
}

static int refresh_force(dnsgpu* h)
{
    HIPCHK(hipMemsetAsync(h->shift, 0, sizeof(double) * 3 * h->g.Z, h->stream));
    return with_storage(h->c.precision, [&](auto t) { return launch_moments<decltype(t)>(h); });
}

static void count_slip_columns(dnsgpu* h, const int8_t* bottom, const int8_t* top);

static int update_wall_words(dnsgpu* h)
{
    const int xy = h->g.X * h->g.Y;
    hipLaunchKernelGGL(wall_word_kernel, dim3((xy + 255) / 256), dim3(256), 0, h->stream, h->mbot, h->mtop, h->g.X,
                       h->g.Y, h->wword);
    HIPCHK(hipGetLastError());
    return 0;
}

// A step grid of about nblk blocks: a positive multiple of the step's
// z-piece count (each group of nzp consecutive blocks updates all pieces of
// its columns).
static int step_grid(const dnsgpu* h, int nblk)
{
    const int P = h->step_nzp;
    return std::max(P, nblk / P * P);
}

// The step kernel's z-piece counts for this case: the column's own pieces
// (z_pieces), and two where that is one and the column has 6 or 8 chunks
// (321-384 or 449-512 rows: two equal pieces).  Two pieces halve each wave's
// plane-sum slice, so more waves fit in the LDS; each piece's blocks also
// read the column's records and the pulled lines at the piece boundary.
// Same-handle runs at 1536x768x384 with riblets: FP64 27.4-27.5 -> 26.2 ms,
// FP16 8.07-8.11 -> 7.72 ms, FP32 13.13 -> 13.35-13.41 ms; three or six
// pieces slower than two in every storage.  The default is two for FP64 and
// FP16 storage, one for FP32; the autotuner times both.
static int step_piece_options(const dnsgpu* h, int* opts)
{
    const int base = h->g.nzp;
    opts[0] = base;
    if (base == 1 && (h->g.nzc == 6 || h->g.nzc == 8)) {
        opts[1] = 2;
        return 2;
    }
    return 1;
}

static bool step_piece_ok(const dnsgpu* h, int nzp)
{
    int opts[2];
    const int n = step_piece_options(h, opts);
    for (int i = 0; i < n; i++)
        if (opts[i] == nzp) return true;
    return false;
}

// Geometry, constants and launch grids of a case (no allocation).
static int plan_handle(const dns_case* c, dnsgpu* h)
{
    if (c->xdv < 2 || c->ydv < 2 || c->zdv < 4 || c->zdv > kMaxZ)
        return set_err(DNSGPU_EARG, "grid %dx%dx%d unsupported (xdv,ydv>=2, 4<=zdv<=%d)", c->xdv, c->ydv, c->zdv,
                       kMaxZ);
    if ((long long)c->xdv * c->ydv * c->zdv >= (1ll << 31))
        return set_err(DNSGPU_EARG, "grid has more than 2^31 nodes");
    if (c->precision != DNSGPU_FP64 && c->precision != DNSGPU_FP32 && c->precision != DNSGPU_FP16)
        return set_err(DNSGPU_EARG, "precision %d not built", c->precision);
    h->c = *c;
    Geom& g = h->g;
    g.X = c->xdv;
    g.Y = c->ydv;
    g.Z = c->zdv;
    g.nzc = (g.Z + kWave - 1) / kWave;
    g.nzp = z_pieces(g.nzc);
    g.cpp = (g.nzc + g.nzp - 1) / g.nzp;
    g.cpb = std::max(1, 4 / g.cpp);
    g.ncols = g.X * g.Y;
    g.ngroups = (g.ncols + g.cpb - 1) / g.cpb;
    g.total = (long long)g.ncols * Q * g.Z;
    h->block = kWave * g.cpp * g.cpb;
    {
        int opts[2];
        const int n = step_piece_options(h, opts);
        h->step_nzp = n > 1 && c->precision != DNSGPU_FP32 ? opts[1] : opts[0];
        const char* spe = getenv("DNSGPU_STEP_PIECES");
        if (spe && step_piece_ok(h, atoi(spe))) h->step_nzp = atoi(spe);
    }
    h->step_w = step_waves(g.Z, h->step_nzp);
    const char* swe = getenv("DNSGPU_STEP_WAVES");
    if (swe) h->step_w = std::max(1, std::min(atoi(swe), h->step_w));
    h->step_rows = 1;
    const char* sre = getenv("DNSGPU_STEP_ROWS");
    if (sre && atoi(sre) == 2 && step_rows_ok(c->precision, g.Z)) {
        h->step_rows = 2;
        h->step_nzp = 1;  // two rows per lane: one piece
        h->step_w = step_waves(g.Z, 1);
    }
    h->esize = storage_bytes(c->precision);
    // omega and the constants exactly as dns-const-q-xycut.c:211-217,
    // wshearstr.c:24, encalc.c:20 evaluate them.
    const int zd = g.Z + 2;
    if (!std::isfinite(c->ub) || !std::isfinite(c->force))
        return set_err(DNSGPU_EARG, "ub %g / force %g not finite", c->ub, c->force);
    if (c->omega < 0) return set_err(DNSGPU_EARG, "omega %g negative (0 derives it from Rebs)", c->omega);
    if (c->omega > 0)
        h->omega = c->omega;
    else {
        double tau = (3. / c->Rebs) * c->ub * (zd - 2.) + 0.5;
        h->omega = 1. / tau;
    }
    // BGK is stable only for 0 < omega < 2 (tau > 1/2); the reference takes
    // any input and runs into NaN instead
    if (!(h->omega > 0.0 && h->omega < 2.0))
        return set_err(DNSGPU_EARG, "relaxation rate omega = %g outside (0, 2) (Rebs %g, ub %g, omega %g)", h->omega,
                       c->Rebs, c->ub, c->omega);
    if (!(c->CFL > 0.0) || !std::isfinite(c->CFL)) return set_err(DNSGPU_EARG, "CFL %g must be positive", c->CFL);
    if (c->forcing != DNSGPU_CONST_Q && c->forcing != DNSGPU_CONST_F)
        return set_err(DNSGPU_EARG, "forcing %d unknown", c->forcing);
    h->mu = (1. / 6.) * (2. / h->omega - 1);
    h->dt = 1. / ((double)(0.5 * (zd - 2.) / c->CFL));

    // grids: the chunk-mapped kernels fill the device once; the step kernel
    // keeps about 4 (FP64) or 8 (FP32) waves per CU in flight, the measured
    // optimum of the read+write stream (overridden by the autotuner, or by
    // DNSGPU_STEP_BLOCKS and DNSGPU_STEP_WAVES)
    int occ = 0;
    int rc = with_storage(c->precision, [&](auto t) {
        return occupancy_blocks(moments_kernel<decltype(t), false>, h->block, &occ);
    });
    if (rc) return rc;
    h->nblk = std::max(1, std::min(occ, g.ngroups));
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    h->ncu = ncu;
    const int waves_per_cu = c->precision == DNSGPU_FP64 ? 4 : 8;
    int sb = (ncu * waves_per_cu + h->step_w - 1) / h->step_w;
    // never more blocks than can be resident at once (the grid loops)
    int per = 0;
    HIPCHK(step_occupancy(c->precision, g.Z, h->step_w, h->step_rows, h->step_nzp, &per));
    h->step_resident = std::max(1, per) * ncu;
    sb = std::min(sb, h->step_resident);
    const char* sbe = getenv("DNSGPU_STEP_BLOCKS");
    if (sbe) sb = atoi(sbe);
    // the partial-sum buffers hold one row per block for any block shape
    // the autotuner may choose (down to one wave per block); the step grid
    // is a multiple of its z-piece count (at most 2 * g.nzp)
    const int pmax = 2 * g.nzp;
    h->step_nblk_max = std::max(pmax, std::min(g.ncols * pmax, 64 * ncu) / pmax * pmax);
    h->step_nblk = step_grid(h, std::min(sb, h->step_nblk_max));
    return 0;
}

struct DevAlloc {
    void** p;
    size_t n;
};

// Every device allocation of a planned handle (the footprint).
static std::vector<DevAlloc> handle_allocs(dnsgpu* h)
{
    const Geom& g = h->g;
    const size_t popbytes = (size_t)g.total * h->esize;
    const size_t xy = (size_t)g.X * g.Y;
    // partials: nblk per plane of the chunk-mapped kernels and nblk * nzp
    // per global sum; the step grid's (up to step_nblk_max)
    const size_t np = (size_t)std::max(h->nblk * g.nzp, h->step_nblk_max);
    return {
        {&h->buf[0], popbytes},
        {&h->buf[1], popbytes},
        {(void**)&h->mbot, xy},
        {(void**)&h->mtop, xy},
        {(void**)&h->wword, xy * 4},
        {(void**)&h->bands, xy * 4},
        {(void**)&h->bnbr, xy * sizeof(BandNbr)},
        {(void**)&h->cnbr, xy * sizeof(ColNbr)},
        {(void**)&h->nf, sizeof(double) * g.Z},
        {(void**)&h->part, sizeof(double) * kNStat * g.Z * np},
        {(void**)&h->gpart, sizeof(double) * kNG * np},
        {(void**)&h->planes, sizeof(double) * (kNStat * g.Z + kNG)},
        {(void**)&h->shift, sizeof(double) * 3 * g.Z},
        {(void**)&h->means, sizeof(double) * 4 * g.Z},
        {(void**)&h->tprof, sizeof(double) * 9 * g.Z},
        {(void**)&h->slip, sizeof(double) * 4 * kSlipBlocks},
        {(void**)&h->st, sizeof(DevState)},
        {(void**)&h->ring, sizeof(LogRow) * kRingRows},
    };
}

extern "C" size_t dnsgpu_footprint(const dns_case* c)
{
    if (!c) return 0;
    dnsgpu p;
    memset(&p, 0, sizeof p);
    if (plan_handle(c, &p)) return 0;
    size_t n = 0;
    for (auto& a : handle_allocs(&p)) n += a.n;
    return n;
}

extern "C" int dnsgpu_create(const dns_case* c, dnsgpu** out)
{
    if (!c || !out) return set_err(DNSGPU_EARG, "null argument");
    *out = nullptr;
    dnsgpu* h = (dnsgpu*)calloc(1, sizeof(dnsgpu));
    if (!h) return set_err(DNSGPU_ENOMEM, "host allocation failed");
    auto fail = [&](int code) {
        dnsgpu_destroy(h);
        return code;
    };
    int rc = plan_handle(c, h);
    if (rc) return fail(rc);
    const Geom& g = h->g;
    if (hipGetDevice(&h->device) != hipSuccess) return fail(set_err(DNSGPU_EHIP, "no current device"));
    rc = scratch_address(h->device, &h->scratch);
    if (rc) return fail(rc);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(DNSGPU_EHIP, "stream creation failed"));
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
        return fail(set_err(DNSGPU_EHIP, "event creation failed"));
    // a small case's population pair from the device's pair arenas when
    // they are reserved (pairarena.hip), else with the other arrays
    const bool arena = (size_t)g.total * h->esize < kPairMinBytes &&
                       pair_arena_alloc(h, (size_t)g.total * h->esize) == 0;
    for (auto& a : handle_allocs(h)) {
        if (arena && (a.p == &h->buf[0] || a.p == &h->buf[1])) {
            h->dev_bytes += a.n;
            continue;
        }
        if (hipMalloc(a.p, a.n) != hipSuccess)
            return fail(set_err(DNSGPU_ENOMEM, "device allocation of %zu bytes failed", a.n));
        h->dev_bytes += a.n;
    }
    const size_t xy = (size_t)g.X * g.Y;
    h->nf_host = (double*)malloc(sizeof(double) * g.Z);
    if (!h->nf_host) return fail(set_err(DNSGPU_ENOMEM, "host allocation failed"));
    for (int z = 0; z < g.Z; z++) h->nf_host[z] = (double)xy;
    if (hipMemcpyAsync(h->nf, h->nf_host, sizeof(double) * g.Z, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        return fail(set_err(DNSGPU_EHIP, "initial copy failed"));
    h->ring_cap = kRingRows;
    if (hipMemsetAsync(h->mbot, -1, xy, h->stream) != hipSuccess ||
        hipMemsetAsync(h->mtop, -1, xy, h->stream) != hipSuccess ||
        hipMemsetAsync(h->bands, 0, xy * 4, h->stream) != hipSuccess ||
        hipMemsetAsync(h->st, 0, sizeof(DevState), h->stream) != hipSuccess ||
        hipMemsetAsync(h->tprof, 0, sizeof(double) * 9 * g.Z, h->stream) != hipSuccess ||
        hipMemsetAsync(h->shift, 0, sizeof(double) * 3 * g.Z, h->stream) != hipSuccess)
        return fail(set_err(DNSGPU_EHIP, "initial memset failed"));
    if (update_wall_words(h)) return fail(DNSGPU_EHIP);
    hipLaunchKernelGGL(column_nbr_kernel, dim3((unsigned)((xy + 255) / 256)), dim3(256), 0, h->stream, g.X, g.Y,
                       h->cnbr);
    // band records of the all-zero bands (read by the ensemble's kernel)
    hipLaunchKernelGGL(band_nbr_kernel, dim3((unsigned)((xy + 255) / 256)), dim3(256), 0, h->stream, h->bands, g.X,
                       g.Y, h->bnbr);
    if (hipGetLastError() != hipSuccess) return fail(set_err(DNSGPU_EHIP, "column records"));
    // the pop buffers start at rest
    h->cur = 0;
    h->poststream = 1;
    h->rs_state = 0;
    if (hipStreamSynchronize(h->stream) != hipSuccess) return fail(set_err(DNSGPU_EHIP, "sync failed"));
    count_slip_columns(h, nullptr, nullptr);
    rc = dnsgpu_init_rest(h);
    if (rc) return fail(rc);
    *out = h;
    return 0;
}

extern "C" void dnsgpu_destroy(dnsgpu* h)
{
    if (!h) return;
    if (h->owner) (void)dnsens_remove(h->owner, h);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    pair_arena_free(h);  // arena ranges go back (buf[0], buf[1] nulled)
    void* ptrs[] = {h->buf[0], h->buf[1], h->mbot,  h->mtop,  h->wword, h->bands, h->bnbr, h->cnbr, h->nf,
                    h->part,   h->gpart,  h->planes, h->shift, h->means, h->tprof, h->slip,
                    h->st,     h->ring,   h->stage};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    free(h->nf_host);
    free(h);
}

extern "C" size_t dnsgpu_device_bytes(const dnsgpu* h) { return h ? h->dev_bytes : 0; }

// ---------------------------------------------------------------- autotuner

namespace {
struct TuneKey {
    int X, Y, Z, prec, bands;
    bool operator==(const TuneKey& o) const
    {
        return X == o.X && Y == o.Y && Z == o.Z && prec == o.prec && bands == o.bands;
    }
};
// shared by every handle of the process (handles may live on different host
// threads): guarded
// a tuned launch shape: blocks of the grid and waves per block
struct TuneShape {
    int blocks, waves, rows, pieces;  // pieces 0: the case's default
};
std::vector<std::pair<TuneKey, TuneShape>> g_tune_cache;
std::mutex g_tune_mutex;

bool tune_lookup(const TuneKey& k, TuneShape* out)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    for (auto& e : g_tune_cache)
        if (e.first == k) {
            *out = e.second;
            return true;
        }
    const char* path = getenv("DNSGPU_TUNE_CACHE");
    if (!path) return false;
    FILE* f = fopen(path, "r");
    if (!f) return false;
    TuneKey r;
    TuneShape sh;
    bool found = false;
    char line[256];
    while (fgets(line, sizeof line, f)) {
        sh.pieces = 0;  // lines without the ninth field: the default pieces
        const int n = sscanf(line, "%d %d %d %d %d %d %d %d %d", &r.X, &r.Y, &r.Z, &r.prec, &r.bands, &sh.blocks,
                             &sh.waves, &sh.rows, &sh.pieces);
        if (n >= 8 && r == k) {
            *out = sh;
            found = true;
        }
    }
    fclose(f);
    if (found) g_tune_cache.push_back({k, *out});
    return found;
}

void tune_store(const TuneKey& k, TuneShape sh)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune_cache.push_back({k, sh});
    const char* path = getenv("DNSGPU_TUNE_CACHE");
    if (!path) return;
    FILE* f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%d %d %d %d %d %d %d %d %d\n", k.X, k.Y, k.Z, k.prec, k.bands, sh.blocks, sh.waves, sh.rows,
            sh.pieces);
    fclose(f);
}
}  // namespace

// Sets the step kernel's launch shape (W waves per block, nblk blocks,
// rows per lane, z-pieces; pieces 0 keeps the current count) within what the
// LDS, the device and the partial buffers allow.  Two rows per lane take
// one piece.
static int set_step_shape(dnsgpu* h, int W, int nblk, int rows, int pieces)
{
    rows = rows == 2 && step_rows_ok(h->c.precision, h->g.Z) ? 2 : 1;
    int nzp = pieces > 0 && step_piece_ok(h, pieces) ? pieces : h->step_nzp;
    if (rows == 2) nzp = 1;
    W = std::max(1, std::min(W, step_waves(h->g.Z, nzp)));
    int per = 0;
    HIPCHK(step_occupancy(h->c.precision, h->g.Z, W, rows, nzp, &per));
    h->step_w = W;
    h->step_rows = rows;
    h->step_nzp = nzp;
    h->step_resident = std::max(1, per) * h->ncu;
    const int cols = std::max(1, (h->g.ncols + W - 1) / W) * nzp;
    h->step_nblk = step_grid(h, std::min(nblk, std::min(std::min(h->step_nblk_max, h->step_resident), cols)));
    return 0;
}

// Time `trial` launches of the step kernel at the current shape (one
// untimed launch first; the buffers are not flipped).
static int time_shape(dnsgpu* h, int trial, double* ms_per_launch)
{
    for (int r = 0; r <= trial; r++) {
        if (r == 1) HIPCHK(hipEventRecord(h->ev0, h->stream));
        with_storage(h->c.precision, [&](auto t) { enqueue_step_kernel<decltype(t)>(h, true); });
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    *ms_per_launch = ms / trial;
    return 0;
}

// ------------------------------------------------ population pair placement
//
// Measured on MI355X with copy kernels over held allocations (16-34 GiB
// buffers): device memory falls into two halves, and a copy from one buffer
// into another runs at 5.45-5.58 TB/s when the two lie in different halves
// against 5.15-5.28 TB/s within one half; the driver fills one half before
// the other, so the two population buffers of dnsgpu_create, allocated one
// after the other, often share a half.  The step reads buf[cur] and writes
// buf[1 - cur] (then the roles swap), and its time follows the placement of
// that pair: C4 riblets in FP32 take 13.2 ms on some pairs and 14.2 ms on
// others, with the same virtual addresses, code and launch shape.  A copy
// probe does not predict the step well (the step reads three x-planes of
// the source around the offset it writes), so the placement is timed with
// the step kernel itself: the destination buffer is replaced by fresh
// allocations, each held while the next is tried so that the allocator
// moves on, and the fastest is kept when it beats the original pair.
namespace {
constexpr size_t kPairFreeReserve = (size_t)2 << 30;   // device memory the search leaves free
constexpr int kPairMaxCandidates = 4;
constexpr double kPairStop = 0.03;  // a candidate this much faster ends the search
constexpr double kPairKeep = 0.015; // the fastest candidate replaces the original if this much faster
}  // namespace

// Pair placement at the current launch shape (state flag already cleared by
// the autotuner).  The state in buf[cur] is not touched; buf[1 - cur] is
// overwritten by the trials, as by the autotuner's own.
static int place_pair(dnsgpu* h, int trial)
{
    const size_t bytes = (size_t)h->g.total * h->esize;
    const char* env = getenv("DNSGPU_PAIR_PLACEMENT");
    // (pairs below kPairMinBytes may be arena ranges, which are never freed here)
    if ((env && !strcmp(env, "0")) || bytes < kPairMinBytes || h->pair_arena || h->pair_candidates) return 0;
    auto time_pair = [&](double* t) {
        double a = 0.0, b = 0.0;
        int rc = time_shape(h, trial, &a);
        if (!rc) rc = time_shape(h, trial, &b);
        *t = std::min(a, b);
        return rc;
    };
    double t0 = 0.0;
    int rc = time_pair(&t0);
    if (rc) return rc;
    void* const orig = h->buf[1 - h->cur];
    std::vector<void*> held;
    void* best = nullptr;
    double tbest = t0;
    int tried = 0;
    while (tried < kPairMaxCandidates) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b < bytes + kPairFreeReserve) break;
        void* c = nullptr;
        if (hipMalloc(&c, bytes) != hipSuccess) break;
        held.push_back(c);
        tried++;
        h->buf[1 - h->cur] = c;
        double t = 0.0;
        rc = time_pair(&t);
        h->buf[1 - h->cur] = orig;
        if (rc) break;
        if (t < tbest) {
            tbest = t;
            best = c;
        }
        if (t < (1.0 - kPairStop) * t0) break;
    }
    (void)hipGetLastError();  // a refused allocation ends the search, nothing else
    if (!rc && best && tbest < (1.0 - kPairKeep) * t0) {
        (void)hipFree(orig);
        h->buf[1 - h->cur] = best;
    } else {
        best = nullptr;
    }
    for (void* c : held)
        if (c != best) (void)hipFree(c);
    h->pair_candidates = tried;
    h->pair_gain = best ? t0 / tbest - 1.0 : 0.0;
    return rc;
}

extern "C" int dnsgpu_pair_placement(const dnsgpu* h, int* candidates, double* step_gain)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    if (candidates) *candidates = h->pair_candidates;
    if (step_gain) *step_gain = h->pair_gain;
    return 0;
}

extern "C" int dnsgpu_autotune(dnsgpu* h, int trial_steps, int* best_blocks, double* best_ms)
{
    if (!h || trial_steps < 1) return set_err(DNSGPU_EARG, "bad argument");
    NOT_MEMBER(h);
    // the handle's state flag and shape are restored on every return
    struct Restore {
        dnsgpu* h;
        int ps, nblk, w, rows, pieces;
        ~Restore()
        {
            h->poststream = ps;
            (void)set_step_shape(h, w, nblk, rows, pieces);
        }
    } restore{h, h->poststream, h->step_nblk, h->step_w, h->step_rows, h->step_nzp};
    h->poststream = 0;
    const TuneKey key{h->g.X, h->g.Y, h->g.Z, h->c.precision, h->has_bands};
    TuneShape cached;
    const bool have = tune_lookup(key, &cached);
    // the pair first (at the cached or the current shape), so that the
    // shapes are timed on the buffers the steps will use
    int rc = have ? set_step_shape(h, cached.waves, cached.blocks, cached.rows, cached.pieces) : 0;
    if (!rc) rc = place_pair(h, trial_steps);
    if (rc) return rc;
    if (have) {
        restore.nblk = h->step_nblk;
        restore.w = h->step_w;
        restore.rows = h->step_rows;
        restore.pieces = h->step_nzp;
        if (best_blocks) *best_blocks = h->step_nblk;
        if (best_ms) *best_ms = -1.0;
        return 0;
    }
    // time the steady-state kernel over the shapes (rows per lane x
    // z-pieces x waves per block x waves per CU), then the three fastest
    // again, alternating, to keep a noisy first round from choosing
    const int wpcs[] = {2, 3, 4, 6, 8, 12, 16, 20, 24, 32};
    struct Cand {
        TuneShape s;
        double t;
    };
    std::vector<Cand> cands;
    const int max_rows = step_rows_ok(h->c.precision, h->g.Z) ? 2 : 1;
    int popts[2];
    const int npopt = step_piece_options(h, popts);
    for (int R = 1; R <= max_rows; R++)
        for (int pi = 0; pi < (R == 2 ? 1 : npopt); pi++) {
            const int P = R == 2 ? 1 : popts[pi];
            for (int W = 2; W <= step_waves(h->g.Z, P); W++) {
                int last = -1;
                for (int wpc : wpcs) {
                    rc = set_step_shape(h, W, (h->ncu * wpc + W - 1) / W, R, P);
                    if (rc) return rc;
                    if (h->step_nblk == last || h->step_w != W || h->step_nzp != P) continue;
                    last = h->step_nblk;
                    double t = 0.0;
                    rc = time_shape(h, trial_steps, &t);
                    if (rc) return rc;
                    cands.push_back({TuneShape{h->step_nblk, W, R, P}, t});
                }
            }
        }
    if (cands.empty()) return set_err(DNSGPU_EHIP, "autotune: no launch shape");
    std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.t < b.t; });
    const int nfin = std::min<int>(3, (int)cands.size());
    for (int i = 0; i < nfin; i++) cands[i].t = 0.0;
    for (int round = 0; round < 2; round++)
        for (int i = 0; i < nfin; i++) {
            int rc = set_step_shape(h, cands[i].s.waves, cands[i].s.blocks, cands[i].s.rows, cands[i].s.pieces);
            if (rc) return rc;
            double t = 0.0;
            rc = time_shape(h, trial_steps, &t);
            if (rc) return rc;
            cands[i].t += t;
        }
    int bi = 0;
    for (int i = 1; i < nfin; i++)
        if (cands[i].t < cands[bi].t) bi = i;
    const TuneShape best = cands[bi].s;
    restore.nblk = best.blocks;
    restore.w = best.waves;
    restore.rows = best.rows;
    restore.pieces = best.pieces;
    tune_store(key, best);
    if (best_blocks) *best_blocks = best.blocks;
    if (best_ms) *best_ms = cands[bi].t / 2;
    return 0;
}

extern "C" int dnsgpu_set_step_shape(dnsgpu* h, int blocks, int waves_per_block, int rows_per_lane)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    NOT_MEMBER(h);
    const int nzp = rows_per_lane == 2 ? 1 : h->step_nzp;  // two rows per lane: one piece
    if (blocks < 1 || waves_per_block < 1 || waves_per_block > step_waves(h->g.Z, nzp))
        return set_err(DNSGPU_EARG, "launch shape %d blocks x %d waves (1..%d waves at zdv %d, %d z-pieces)", blocks,
                       waves_per_block, step_waves(h->g.Z, nzp), h->g.Z, nzp);
    if (rows_per_lane != 1 && !(rows_per_lane == 2 && step_rows_ok(h->c.precision, h->g.Z)))
        return set_err(DNSGPU_EARG, "%d rows per lane (2 needs FP16 storage and an even zdv <= %d)", rows_per_lane,
                       kWave * kMaxWaves);
    return set_step_shape(h, waves_per_block, blocks, rows_per_lane, nzp);
}

extern "C" int dnsgpu_step_shape(const dnsgpu* h, int* blocks, int* waves_per_block, int* rows_per_lane)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    if (blocks) *blocks = h->step_nblk;
    if (waves_per_block) *waves_per_block = h->step_w;
    if (rows_per_lane) *rows_per_lane = h->step_rows;
    return 0;
}

extern "C" int dnsgpu_set_step_pieces(dnsgpu* h, int pieces)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    NOT_MEMBER(h);
    int opts[2];
    const int n = step_piece_options(h, opts);
    if (!step_piece_ok(h, pieces))
        return set_err(DNSGPU_EARG, "%d z-pieces at zdv %d (%d%s)", pieces, h->g.Z, opts[0],
                       n > 1 ? " or 2" : " only");
    if (h->step_rows == 2 && pieces != 1) return set_err(DNSGPU_EARG, "two rows per lane take one z-piece");
    return set_step_shape(h, h->step_w, h->step_nblk, h->step_rows, pieces);
}

extern "C" int dnsgpu_step_pieces(const dnsgpu* h, int* pieces)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    if (pieces) *pieces = h->step_nzp;
    return 0;
}

static void count_slip_columns(dnsgpu* h, const int8_t* bottom, const int8_t* top)
{
    const size_t xy = (size_t)h->g.X * h->g.Y;
    for (int q = 0; q < 4; q++) h->slip_cols[q] = 0;
    for (size_t i = 0; i < xy; i++) {
        const int b = bottom ? bottom[i] : -1, t = top ? top[i] : -1;
        h->slip_cols[b == -2 ? 0 : 1]++;
        h->slip_cols[t == -2 ? 2 : 3]++;
    }
}

extern "C" int dnsgpu_set_walls(dnsgpu* h, const int8_t* bottom, const int8_t* top)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    NOT_MEMBER(h);
    const size_t xy = (size_t)h->g.X * h->g.Y;
    const int8_t* src[2] = {bottom, top};
    int8_t* dst[2] = {h->mbot, h->mtop};
    // both walls are checked before either is copied: a refused call changes nothing
    for (int w = 0; w < 2; w++) {
        if (!src[w]) continue;
        for (size_t i = 0; i < xy; i++) {
            if (src[w][i] != -1 && src[w][i] != -2)
                return set_err(DNSGPU_EARG, "wall code %d at %zu (must be -1 or -2)", src[w][i], i);
            if (h->has_bands && src[w][i] == -2)
                return set_err(DNSGPU_EARG, "slip wall nodes cannot be combined with solid bands");
        }
    }
    for (int w = 0; w < 2; w++) {
        if (src[w]) {
            HIPCHK(hipMemcpyAsync(dst[w], src[w], xy, hipMemcpyHostToDevice, h->stream));
        } else {
            HIPCHK(hipMemsetAsync(dst[w], -1, xy, h->stream));
        }
    }
    int rc = update_wall_words(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    count_slip_columns(h, bottom, top);
    return 0;
}

extern "C" int dnsgpu_set_solid_bands(dnsgpu* h, const int16_t* bottom, const int16_t* top)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    NOT_MEMBER(h);
    // the FP16 step keeps the bounced-back populations in the solid nodes'
    // slots, which a state advanced without these bands does not have (one
    // rule for every precision, though FP32 and FP64 load their own slots)
    if (!h->poststream)
        return set_err(DNSGPU_EARG, "solid bands are set before the first step (after create, an init or a load)");
    const int X = h->g.X, Y = h->g.Y, Z = h->g.Z;
    const size_t xy = (size_t)X * Y;
    std::vector<uint32_t> bw(xy, 0u);
    std::vector<double> nf(Z, 0.0);
    bool any = false;
    for (size_t i = 0; i < xy; i++) {
        const int hb = bottom ? bottom[i] : 0, ht = top ? top[i] : 0;
        if (hb < 0 || ht < 0 || hb + ht > Z - 2)
            return set_err(DNSGPU_EARG, "solid band %d + %d at column %zu leaves fewer than 2 fluid rows of %d", hb,
                           ht, i, Z);
        bw[i] = (uint32_t)hb | ((uint32_t)ht << 16);
        any |= (hb | ht) != 0;
        for (int z = hb; z < Z - ht; z++) nf[z] += 1.0;
    }
    if (any) {
        // bands are no-slip solids: the walls must be no-slip too
        std::vector<int8_t> m(xy);
        for (int w = 0; w < 2; w++) {
            HIPCHK(hipMemcpyAsync(m.data(), w ? h->mtop : h->mbot, xy, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            for (size_t i = 0; i < xy; i++)
                if (m[i] == -2) return set_err(DNSGPU_EARG, "solid bands need no-slip walls (slip node at %zu)", i);
        }
    }
    double total = 0.0;
    for (int z = 0; z < Z; z++) total += nf[z];
    HIPCHK(hipMemcpyAsync(h->bands, bw.data(), xy * 4, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(band_nbr_kernel, dim3((unsigned)((xy + 255) / 256)), dim3(256), 0, h->stream, h->bands, X, Y,
                       h->bnbr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h->nf, nf.data(), sizeof(double) * Z, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->has_bands = any ? 1 : 0;
    h->nfluid = total;
    for (int z = 0; z < Z; z++) h->nf_host[z] = nf[z];
    if (any) {
        // the wall-adjacent velocity is sampled over the columns whose first
        // row next to each wall is fluid (the walls are no-slip here)
        int nb = 0, nt = 0;
        for (size_t i = 0; i < xy; i++) {
            nb += (bottom ? bottom[i] : 0) == 0;
            nt += (top ? top[i] : 0) == 0;
        }
        h->slip_cols[0] = 0;
        h->slip_cols[1] = nb;
        h->slip_cols[2] = 0;
        h->slip_cols[3] = nt;
    }
    // the force of the first step and the plane-mean shifts were derived by
    // the last init or load over every node; derive them again over the
    // fluid nodes (solid nodes may hold anything, zeros in a checkpoint)
    int rc = refresh_force(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// Tables of the seeded turbulent field (DNSGPU_INIT_TURBULENT), laid out in
// one host array t: [nm][kModeRec] modes, [mx+1][X] cos and sin of
// 2 pi i x / X, [nm][4][Y] cos and sin of ky y + p1 and ky y + p2, [3][Z] the
// log-law mean, the envelope and its derivative.  Every transcendental value
// is one libm call on an argument formed in the order written here, and the
// sin and cos of one argument are taken in separate loops (each a plain libm
// call), so the tables are reproducible by any program that calls the same
// libm.  Modes: 0 <= i <= mx, -my <= j <= my, (i, j) > (0, 0), with
// mx = min(nmodes, (X-1)/2) and my = min(nmodes, (Y-1)/2): no two modes alias
// on the grid and no mode is its own mirror, so the modes are orthogonal over
// the periodic planes and the rms of the field follows from the tables alone.
static void turbulent_tables(const dnsgpu* h, const dns_init* spec, std::vector<double>& t, InitArgs& A,
                             size_t off[5])
{
    const int X = h->g.X, Y = h->g.Y, Z = h->g.Z;
    const double pi = 3.141592653589793, two_pi = 6.283185307179586;
    const double kinv = 1.0 / 0.41, Bplus = 5.2;
    const double uc = spec->uc, amp = spec->amp;
    const double nu = (1. / 3.) * (1. / h->omega - 0.5);
    const double hh = 0.5 * (double)Z;
    const int mx = std::min(spec->nmodes, (X - 1) / 2), my = std::min(spec->nmodes, (Y - 1) / 2);
    struct Mode {
        int i, j;
        double sc, a1, a2, p1, p2, kx, ky;
    };
    std::vector<Mode> modes;
    for (int i = 0; i <= mx; i++)
        for (int j = -my; j <= my; j++) {
            if (i == 0 && j <= 0) continue;
            const unsigned long long m = (unsigned long long)((i * 64 + (j + 32)) & 0xFFFF);
            Mode md;
            md.i = i;
            md.j = j;
            md.sc = 1.0 / (double)(i * i + j * j);
            md.a1 = urand(spec->seed, m * 8 + 0);
            md.a2 = urand(spec->seed, m * 8 + 1);
            md.p1 = pi * urand(spec->seed, m * 8 + 2);
            md.p2 = pi * urand(spec->seed, m * 8 + 3);
            md.kx = (two_pi * (double)i) / (double)X;
            md.ky = (two_pi * (double)j) / (double)Y;
            modes.push_back(md);
        }
    const int nm = (int)modes.size();
    off[0] = 0;
    off[1] = off[0] + (size_t)nm * kModeRec;
    off[2] = off[1] + (size_t)(mx + 1) * X;
    off[3] = off[2] + (size_t)(mx + 1) * X;
    off[4] = off[3] + (size_t)nm * 4 * Y;
    t.assign(off[4] + 3 * (size_t)Z, 0.0);
    for (int m = 0; m < nm; m++) {
        double* r = &t[off[0] + (size_t)m * kModeRec];
        r[0] = (double)modes[m].i;
        r[1] = modes[m].sc;
        r[2] = modes[m].a1;
        r[3] = modes[m].a2;
        r[4] = modes[m].kx;
        r[5] = modes[m].ky;
    }
    for (int i = 0; i <= mx; i++)
        for (int x = 0; x < X; x++) t[off[1] + (size_t)i * X + x] = cos((two_pi * (double)((i * x) % X)) / (double)X);
    for (int i = 0; i <= mx; i++)
        for (int x = 0; x < X; x++) t[off[2] + (size_t)i * X + x] = sin((two_pi * (double)((i * x) % X)) / (double)X);
    for (int m = 0; m < nm; m++)
        for (int q = 0; q < 2; q++)
            for (int y = 0; y < Y; y++) {
                const int n = ((modes[m].j * y) % Y + Y) % Y;
                const double arg = (two_pi * (double)n) / (double)Y + (q ? modes[m].p2 : modes[m].p1);
                t[off[3] + ((size_t)m * 4 + 2 * q) * Y + y] = cos(arg);
            }
    for (int m = 0; m < nm; m++)
        for (int q = 0; q < 2; q++)
            for (int y = 0; y < Y; y++) {
                const int n = ((modes[m].j * y) % Y + Y) % Y;
                const double arg = (two_pi * (double)n) / (double)Y + (q ? modes[m].p2 : modes[m].p1);
                t[off[3] + ((size_t)m * 4 + 2 * q + 1) * Y + y] = sin(arg);
            }
    double* U = &t[off[4]];
    double* env = U + Z;
    double* denv = U + 2 * Z;
    std::vector<double> sn(Z), cs(Z);
    for (int z = 0; z < Z; z++) {
        const double yy = (double)z + 0.5;
        const double yw = yy < hh ? yy : (double)Z - yy;
        const double yplus = yw * uc / nu;
        const double uplus = yplus < 11.0 ? yplus : (log(yplus) * kinv + Bplus);
        U[z] = uc * uplus;
    }
    for (int z = 0; z < Z; z++) sn[z] = sin(pi * ((double)z + 0.5) / (double)Z);
    for (int z = 0; z < Z; z++) cs[z] = cos(pi * ((double)z + 0.5) / (double)Z);
    for (int z = 0; z < Z; z++) {
        env[z] = sn[z] * sn[z];
        denv[z] = (((2.0 * sn[z]) * cs[z]) * pi) / (double)Z;
    }
    // rms of the unscaled fluctuation over the grid: mean over the planes of
    // sum_m sc^2 (denv^2 (a1^2 + a2^2) + env^2 ((a1 ky)^2 + (a2 kx)^2
    //   - 2 (a1 ky)(a2 kx) cos(p1 - p2))) / 2
    std::vector<double> cpp(nm);
    for (int m = 0; m < nm; m++) cpp[m] = cos(modes[m].p1 - modes[m].p2);
    double S = 0.0;
    for (int z = 0; z < Z; z++) {
        const double d2 = denv[z] * denv[z], e2 = env[z] * env[z];
        for (int m = 0; m < nm; m++) {
            const Mode& md = modes[m];
            const double q1 = md.a1 * md.ky, q2 = md.a2 * md.kx;
            const double term = (md.sc * md.sc) * (d2 * ((md.a1 * md.a1 + md.a2 * md.a2) * 0.5) +
                                                   e2 * (((q1 * q1 + q2 * q2) - 2.0 * (q1 * q2) * cpp[m]) * 0.5));
            S = S + term;
        }
    }
    const double ucl = uc * (log(hh * uc / nu) * kinv + Bplus);
    A.nm = nm;
    A.nrm = S > 0.0 ? amp * ucl / sqrt(S / (double)Z) : 0.0;
}

static int launch_init(dnsgpu* h, int kind, const dns_init* spec)
{
    InitArgs A;
    memset(&A, 0, sizeof A);
    A.g = h->g;
    A.cur = h->buf[0];
    A.other = h->buf[1];
    A.kind = kind;
    if (spec) {
        A.uc = spec->uc;
        A.amp = spec->amp;
        A.seed = spec->seed;
    }
    if (kind == DNSGPU_INIT_TURBULENT) {
        std::vector<double> t;
        size_t off[5];
        turbulent_tables(h, spec, t, A, off);
        int rc = stage_reserve(h, sizeof(double) * t.size());
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(h->stage, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, h->stream));
        A.mode = h->stage + off[0];
        A.cx = h->stage + off[1];
        A.sx = h->stage + off[2];
        A.by = h->stage + off[3];
        A.zt = h->stage + off[4];
        // the host array t must outlive the copy
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    const long long n = (long long)h->g.X * h->g.Y * h->g.Z;
    int grid = (int)std::min<long long>((n + 255) / 256, 65536);
    with_storage(h->c.precision, [&](auto t) {
        hipLaunchKernelGGL(init_kernel<decltype(t)>, dim3(grid), dim3(256), 0, h->stream, A);
    });
    HIPCHK(hipGetLastError());
    h->cur = 0;
    h->poststream = 1;
    h->rs_state = 0;
    int rc = refresh_force(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int dnsgpu_init_rest(dnsgpu* h)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    NOT_MEMBER(h);
    return launch_init(h, 0, nullptr);
}

extern "C" int dnsgpu_init_field(dnsgpu* h, const dns_init* spec)
{
    if (!h || !spec) return set_err(DNSGPU_EARG, "null argument");
    NOT_MEMBER(h);
    if (spec->kind < 0 || spec->kind > 2) return set_err(DNSGPU_EARG, "unknown init kind %d", spec->kind);
    if (spec->kind == 2 && (spec->nmodes < 1 || spec->nmodes > 31))
        return set_err(DNSGPU_EARG, "nmodes must be in [1, 31]");
    return launch_init(h, spec->kind, spec);
}

static int check_slab(dnsgpu* h, int x0, int nx)
{
    if (x0 < 0 || nx < 1 || x0 + nx > h->g.X)
        return set_err(DNSGPU_EARG, "slab [%d, %d) outside [0, %d)", x0, x0 + nx, h->g.X);
    return 0;
}

extern "C" int dnsgpu_load_aos(dnsgpu* h, int x0, int nx, const double* ftemp, const double* f)
{
    if (!h || !ftemp || !f) return set_err(DNSGPU_EARG, "null argument");
    NOT_MEMBER(h);
    int rc = check_slab(h, x0, nx);
    if (rc) return rc;
    const size_t half = (size_t)nx * h->g.Y * (h->g.Z + 2) * Q * sizeof(double);
    if ((rc = stage_reserve(h, 2 * half))) return rc;
    HIPCHK(hipMemcpyAsync(h->stage, ftemp, half, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync((char*)h->stage + half, f, half, hipMemcpyHostToDevice, h->stream));
    AosArgs A;
    memset(&A, 0, sizeof A);
    A.g = h->g;
    A.cur_w = h->buf[0];
    A.other_w = h->buf[1];
    A.x0 = x0;
    A.nx = nx;
    A.ftemp = h->stage;
    A.f = (double*)((char*)h->stage + half);
    const long long n = (long long)nx * h->g.Y * h->g.Z;
    int grid = (int)std::min<long long>((n + 255) / 256, 65536);
    with_storage(h->c.precision, [&](auto t) {
        hipLaunchKernelGGL(load_aos_kernel<decltype(t)>, dim3(grid), dim3(256), 0, h->stream, A);
    });
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    h->cur = 0;
    h->poststream = 1;
    h->rs_state = 0;
    return 0;
}

extern "C" int dnsgpu_commit_load(dnsgpu* h)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    NOT_MEMBER(h);
    h->cur = 0;
    h->poststream = 1;
    h->rs_state = 0;
    int rc = refresh_force(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int dnsgpu_set_force(dnsgpu* h, double force)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    NOT_MEMBER(h);
    if (h->c.forcing != DNSGPU_CONST_F)
        return set_err(DNSGPU_EARG, "a constant-flow-rate case takes its force from its state (encalc.c:121)");
    if (!std::isfinite(force)) return set_err(DNSGPU_EARG, "force %g not finite", force);
    h->c.force = force;
    // DevState.G: the force of the next step
    HIPCHK(hipMemcpyAsync(&h->st->G, &force, sizeof force, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int dnsgpu_save_aos(dnsgpu* h, int x0, int nx, double* ftemp, double* f)
{
    if (!h || !ftemp || !f) return set_err(DNSGPU_EARG, "null argument");
    int rc = check_slab(h, x0, nx);
    if (rc) return rc;
    const size_t half = (size_t)nx * h->g.Y * (h->g.Z + 2) * Q * sizeof(double);
    if ((rc = stage_reserve(h, 2 * half))) return rc;
    AosArgs A;
    memset(&A, 0, sizeof A);
    A.g = h->g;
    A.cur = h->buf[h->cur];
    A.other = h->buf[1 - h->cur];
    A.mbot = h->mbot;
    A.mtop = h->mtop;
    A.bands = h->has_bands ? h->bands : nullptr;
    A.local = h->poststream;
    A.x0 = x0;
    A.nx = nx;
    A.ftemp = h->stage;
    A.f = (double*)((char*)h->stage + half);
    const long long n = (long long)nx * h->g.Y * (h->g.Z + 2);
    int grid = (int)std::min<long long>((n + 255) / 256, 65536);
    with_storage(h->c.precision, [&](auto t) {
        hipLaunchKernelGGL(save_aos_kernel<decltype(t)>, dim3(grid), dim3(256), 0, h->stream, A);
    });
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ftemp, h->stage, half, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(f, (char*)h->stage + half, half, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int dnsgpu_get_force(dnsgpu* h, double* force)
{
    if (!h || !force) return set_err(DNSGPU_EARG, "null argument");
    HIPCHK(hipMemcpyAsync(force, &h->st->G, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

static int read_rows(dnsgpu* h, dns_steplog* rows, int n)
{
    static_assert(sizeof(LogRow) == sizeof(dns_steplog), "log row layout");
    HIPCHK(hipMemcpyAsync(rows, h->ring, sizeof(LogRow) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int i = 0; i < n; i++)
        if (isnan(rows[i].stress))
            return set_err(DNSGPU_ENAN, "Eroor: NAN, found in wall shear stress (step row %d)", i);
    return 0;
}

extern "C" int dnsgpu_step(dnsgpu* h, long cntr0, int nsteps, dns_steplog* rows)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    NOT_MEMBER(h);
    if (nsteps < 0) return set_err(DNSGPU_EARG, "nsteps < 0");
    int done = 0;
    while (done < nsteps) {
        const int chunk = rows ? std::min(nsteps - done, h->ring_cap) : nsteps - done;
        for (int s = 0; s < chunk; s++) {
            int rc = do_step(h, rows != nullptr, rows ? h->ring + s : nullptr, (long long)cntr0 + done + s);
            if (rc) return rc;
        }
        if (rows) {
            int rc = read_rows(h, rows + done, chunk);
            if (rc) return rc;
        }
        done += chunk;
    }
    if (!rows) {
        double G;
        int rc = dnsgpu_get_force(h, &G);
        if (rc) return rc;
        if (isnan(G)) return set_err(DNSGPU_ENAN, "Eroor: NAN, found in the force budget");
    }
    return 0;
}

template <typename T>
static int launch_stats_pass(dnsgpu* h, int pass)
{
    StatArgs A;
    A.g = h->g;
    A.cur = h->buf[h->cur];
    A.mbot = h->mbot;
    A.mtop = h->mtop;
    A.bands = h->has_bands ? h->bands : nullptr;
    A.local = h->poststream;
    A.pass = pass;
    A.means = h->means;
    A.part = h->part;
    A.nblk = h->nblk;
    hipLaunchKernelGGL(stats_kernel<T>, dim3(h->nblk, h->g.nzp), dim3(h->block), 0, h->stream, A);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(5 * h->g.Z), dim3(256), 0, h->stream, h->part, h->nblk, h->planes);
    HIPCHK(hipGetLastError());
    return 0;
}

// Fluid nodes of plane z as the divisor of its statistics.c means (1 for a
// plane with none, whose sums are zero).
static double plane_count(const dnsgpu* h, int z)
{
    const double n = h->nf_host[z];
    return n > 0.0 ? n : 1.0;
}

// statistics.c's two passes over the current state: leaves the plane means
// (UAVG, VAVG, WAVG, DAVG) in h->means and m, and the fluctuation sums of
// pass 2 (u'u', v'v', w'w', u'w', rho'rho') in h->planes and s.
static int run_statistics(dnsgpu* h, std::vector<double>& m, std::vector<double>& s)
{
    const int Z = h->g.Z;
    s.assign(5 * (size_t)Z, 0.0);
    m.assign(4 * (size_t)Z, 0.0);
    int rc = with_storage(h->c.precision, [&](auto t) { return launch_stats_pass<decltype(t)>(h, 1); });
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s.data(), h->planes, sizeof(double) * 5 * Z, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    // statistics.c:54-60, over the fluid nodes of each plane (the whole
    // plane with flat walls)
    for (int q = 0; q < 4; q++)
        for (int z = 0; z < Z; z++) m[(size_t)q * Z + z] = s[(size_t)q * Z + z] / plane_count(h, z);
    HIPCHK(hipMemcpyAsync(h->means, m.data(), sizeof(double) * 4 * Z, hipMemcpyHostToDevice, h->stream));
    rc = with_storage(h->c.precision, [&](auto t) { return launch_stats_pass<decltype(t)>(h, 2); });
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s.data(), h->planes, sizeof(double) * 5 * Z, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int dnsgpu_statistics(dnsgpu* h, double* out)
{
    if (!h || !out) return set_err(DNSGPU_EARG, "null argument");
    const int Z = h->g.Z, zd = Z + 2;
    std::vector<double> s, m;
    int rc = run_statistics(h, m, s);
    if (rc) return rc;
    memset(out, 0, sizeof(double) * 9 * zd);
    for (int z = 0; z < Z; z++) {
        const int k = z + 1;
        const double nxy = plane_count(h, z);
        out[0 * zd + k] = m[0 * (size_t)Z + z];
        out[1 * zd + k] = m[1 * (size_t)Z + z];
        out[2 * zd + k] = m[2 * (size_t)Z + z];
        out[3 * zd + k] = sqrt(s[0 * (size_t)Z + z] / nxy);
        out[4 * zd + k] = sqrt(s[1 * (size_t)Z + z] / nxy);
        out[5 * zd + k] = sqrt(s[2 * (size_t)Z + z] / nxy);
        out[6 * zd + k] = s[3 * (size_t)Z + z] / nxy;
        out[7 * zd + k] = m[3 * (size_t)Z + z];
        out[8 * zd + k] = sqrt(s[4 * (size_t)Z + z] / nxy);
    }
    return 0;
}

// ---------------------------------------------------------------- time averages

extern "C" int dnsgpu_tavg_reset(dnsgpu* h)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    HIPCHK(hipMemsetAsync(h->tprof, 0, sizeof(double) * 9 * h->g.Z, h->stream));
    DevState st;
    HIPCHK(hipMemcpyAsync(&st, h->st, sizeof st, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const double G = st.G, rho = st.rho_mean, steps = st.steps;
    memset(&st, 0, sizeof st);
    st.G = G;
    st.rho_mean = rho;
    st.steps = steps;
    HIPCHK(hipMemcpyAsync(h->st, &st, sizeof st, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int q = 0; q < 4; q++) h->slip_sum[q] = 0.0;
    return 0;
}

template <typename T>
static int launch_slip(dnsgpu* h)
{
    StatArgs A;
    A.g = h->g;
    A.cur = h->buf[h->cur];
    A.mbot = h->mbot;
    A.mtop = h->mtop;
    A.bands = h->has_bands ? h->bands : nullptr;
    A.local = h->poststream;
    A.pass = 0;
    A.means = nullptr;
    A.part = nullptr;
    A.nblk = 0;
    hipLaunchKernelGGL(slip_velocity_kernel<T>, dim3(kSlipBlocks), dim3(256), 0, h->stream, A, h->slip);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int dnsgpu_tavg_sample(dnsgpu* h)
{
    if (!h) return set_err(DNSGPU_EARG, "null handle");
    std::vector<double> s, m;
    int rc = run_statistics(h, m, s);
    if (rc) return rc;
    const int Z = h->g.Z;
    hipLaunchKernelGGL(tavg_profile_kernel, dim3((Z + 255) / 256), dim3(256), 0, h->stream, h->means, h->planes, Z,
                       h->nf, h->tprof, h->st);
    HIPCHK(hipGetLastError());
    rc = with_storage(h->c.precision, [&](auto t) { return launch_slip<decltype(t)>(h); });
    if (rc) return rc;
    std::vector<double> part(4 * (size_t)kSlipBlocks);
    HIPCHK(hipMemcpyAsync(part.data(), h->slip, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int q = 0; q < 4; q++) {
        double t = 0.0;
        for (int b = 0; b < kSlipBlocks; b++) t += part[(size_t)q * kSlipBlocks + b];
        h->slip_sum[q] += t;
    }
    return 0;
}

extern "C" int dnsgpu_tavg_get(dnsgpu* h, dns_tavg* out, double* profiles)
{
    if (!h || !out) return set_err(DNSGPU_EARG, "null argument");
    const int Z = h->g.Z, zd = Z + 2;
    DevState st;
    std::vector<double> tp(9 * (size_t)Z);
    HIPCHK(hipMemcpyAsync(&st, h->st, sizeof st, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(tp.data(), h->tprof, sizeof(double) * tp.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    memset(out, 0, sizeof *out);
    out->nsteps = st.n_rows;
    out->nsamples = st.n_prof;
    const double nr = st.n_rows > 0 ? st.n_rows : 1.0;
    out->pgrad = st.s_pgrad / nr;
    out->utau = sqrt(fabs(out->pgrad));  // time_stats.c:114
    out->stress = st.s_stress / nr;
    out->stup = st.s_stup / nr;
    out->stdown = st.s_stdown / nr;
    out->Ub = st.s_U / nr;
    out->K = st.s_K / nr;
    out->rho = st.s_rho / nr;
    // cfriction.c:33,64: Cf = stress / (0.5 ub^2) with the time-mean ub
    out->Cf_w = out->Ub != 0.0 ? out->stress / (0.5 * out->Ub * out->Ub) : 0.0;
    out->Cf_p = out->Ub != 0.0 ? out->pgrad / (0.5 * out->Ub * out->Ub) : 0.0;
    const double nu = (1. / 3.) * (1. / h->omega - 0.5);
    out->Re_tau = out->utau * (0.5 * Z) / nu;  // time_stats.c:117
    const double ns = st.n_prof > 0 ? st.n_prof : 1.0;
    const double cnt[4] = {(double)h->slip_cols[0], (double)h->slip_cols[1], (double)h->slip_cols[2],
                           (double)h->slip_cols[3]};
    double* u4[4] = {&out->u_slip_bottom, &out->u_noslip_bottom, &out->u_slip_top, &out->u_noslip_top};
    for (int q = 0; q < 4; q++) *u4[q] = cnt[q] > 0 ? h->slip_sum[q] / (ns * cnt[q]) : 0.0;
    if (profiles) {
        // turb-fld order: UAVG VAVG WAVG UTIN VTIN WTIN REYS DAVG DTIN, time-averaged
        memset(profiles, 0, sizeof(double) * 9 * zd);
        for (int z = 0; z < Z; z++) {
            const int k = z + 1;
            const double u = tp[0 * (size_t)Z + z] / ns, v = tp[1 * (size_t)Z + z] / ns;
            const double w = tp[2 * (size_t)Z + z] / ns, r = tp[3 * (size_t)Z + z] / ns;
            profiles[0 * zd + k] = u;
            profiles[1 * zd + k] = v;
            profiles[2 * zd + k] = w;
            profiles[3 * zd + k] = sqrt(fmax(tp[4 * (size_t)Z + z] / ns - u * u, 0.0));
            profiles[4 * zd + k] = sqrt(fmax(tp[5 * (size_t)Z + z] / ns - v * v, 0.0));
            profiles[5 * zd + k] = sqrt(fmax(tp[6 * (size_t)Z + z] / ns - w * w, 0.0));
            profiles[6 * zd + k] = tp[7 * (size_t)Z + z] / ns - u * w;
            profiles[7 * zd + k] = r;
            profiles[8 * zd + k] = sqrt(fmax(tp[8 * (size_t)Z + z] / ns - r * r, 0.0));
        }
    }
    return 0;
}

extern "C" int dnsgpu_fields(dnsgpu* h, int x0, int nx, double* vel, double* den)
{
    if (!h || !vel || !den) return set_err(DNSGPU_EARG, "null argument");
    int rc = check_slab(h, x0, nx);
    if (rc) return rc;
    const size_t nn = (size_t)nx * h->g.Y * (h->g.Z + 2);
    if ((rc = stage_reserve(h, nn * 4 * sizeof(double)))) return rc;
    FieldArgs A;
    A.g = h->g;
    A.cur = h->buf[h->cur];
    A.mbot = h->mbot;
    A.mtop = h->mtop;
    A.bands = h->has_bands ? h->bands : nullptr;
    A.local = h->poststream;
    A.x0 = x0;
    A.nx = nx;
    A.vel = h->stage;
    A.den = h->stage + nn * 3;
    int grid = (int)std::min<size_t>((nn + 255) / 256, 65536);
    with_storage(h->c.precision, [&](auto t) {
        hipLaunchKernelGGL(fields_kernel<decltype(t)>, dim3(grid), dim3(256), 0, h->stream, A);
    });
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(vel, A.vel, nn * 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(den, A.den, nn * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// dnsgpu_fields without any allocation: the staging is the device's static
// scratch array.  The fields kernel writes every element of the nn * 4
// doubles it stages (zeros at wall rows and solid nodes), so nothing is
// cleared first.
extern "C" int dnsgpu_fields_noalloc(dnsgpu* h, int x0, int nx, double* vel, double* den)
{
    if (!h || !vel || !den) return set_err(DNSGPU_EARG, "null argument");
    int rc = check_slab(h, x0, nx);
    if (rc) return rc;
    const size_t nn = (size_t)nx * h->g.Y * (h->g.Z + 2);
    if (nn * 4 > kScratchDoubles)
        return set_err(DNSGPU_EARG,
                       "dnsgpu_fields_noalloc: %d x-planes need %zu doubles of staging, more than the %zu of the "
                       "static array (at most %zu x-planes of this grid per call)",
                       nx, nn * 4, kScratchDoubles, kScratchDoubles / ((size_t)h->g.Y * (h->g.Z + 2) * 4));
    if (!h->scratch) return set_err(DNSGPU_EHIP, "dnsgpu_fields_noalloc: scratch address not resolved");
    std::lock_guard<std::mutex> lock(g_scratch_mutex[h->device]);
    FieldArgs A;
    A.g = h->g;
    A.cur = h->buf[h->cur];
    A.mbot = h->mbot;
    A.mtop = h->mtop;
    A.bands = h->has_bands ? h->bands : nullptr;
    A.local = h->poststream;
    A.x0 = x0;
    A.nx = nx;
    A.vel = (double*)h->scratch;
    A.den = A.vel + nn * 3;
    int grid = (int)std::min<size_t>((nn + 255) / 256, 65536);
    with_storage(h->c.precision, [&](auto t) {
        hipLaunchKernelGGL(fields_kernel<decltype(t)>, dim3(grid), dim3(256), 0, h->stream, A);
    });
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(vel, A.vel, nn * 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(den, A.den, nn * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int dnsgpu_time_steps(dnsgpu* h, int nsteps, int with_stats, double* ms_kernels, double* ms_wall)
{
    if (!h || nsteps < 1) return set_err(DNSGPU_EARG, "bad argument");
    NOT_MEMBER(h);
    std::vector<hipEvent_t> ev(2 * (size_t)nsteps);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int s = 0; s < nsteps; s++) {
        HIPCHK(hipEventRecord(ev[2 * s], h->stream));
        const bool st = with_stats != 0;
        with_storage(h->c.precision, [&](auto t) { enqueue_step_kernel<decltype(t)>(h, st); });
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ev[2 * s + 1], h->stream));
        int rc = reduce_partials(h, st, h->step_nblk / h->step_nzp, h->step_nblk);
        if (rc) return rc;
        FinalizeArgs F = fin_args(h, st, nullptr, s, 0);
        hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(1024), 0, h->stream, F);
        HIPCHK(hipGetLastError());
        h->cur = 1 - h->cur;
        h->poststream = 0;
    }
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    double tk = 0.0;
    for (int s = 0; s < nsteps; s++) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]));
        tk += ms;
    }
    float wall = 0.f;
    HIPCHK(hipEventElapsedTime(&wall, h->ev0, h->ev1));
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (ms_kernels) *ms_kernels = tk;
    if (ms_wall) *ms_wall = wall;
    return 0;
}
