// handle.h -- the library's internal view of a case handle, shared by the
// single-case entry points (dnsgpu.hip) and the ensemble (dnsens.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/dnsgpu.h"
#include "lbm_io_kernels.h"
#include "lbm_kernels.h"

using namespace lbm;

constexpr int kNStat = 5;  // per-plane sums of the statistics.c kernel
constexpr int kSlipBlocks = 256;
constexpr int kRingRows = 1024;  // log rows kept on the device per handle

// Records the message for dnsgpu_last_error and returns code.
__attribute__((visibility("hidden"))) int set_err(int code, const char* fmt, ...);

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_err(DNSGPU_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                               \
    } while (0)

struct dnsgpu {
    dns_case c;
    Geom g;
    int esize;  // bytes per stored population
    void* buf[2];
    int cur;         // buffer holding the current state
    int poststream;  // 1: buf[cur] holds post-stream populations (after load/init)
    int rs_state;    // 1: buf[cur] is from a read-side band step (solid slots hold no pushes)
    int8_t* mbot;
    int8_t* mtop;
    uint32_t* wword;
    uint32_t* bands;  // [X*Y] solid band words hb | ht << 16
    uint32_t* bnbr;   // [X*Y][16] band neighbourhood records
    uint32_t* cnbr;   // [X*Y][16] column neighbourhood records
    double* nf;       // [Z] fluid nodes per plane (device)
    double* nf_host;  // [Z] the same on the host
    double nfluid;    // fluid nodes in total
    int has_bands;
    double* part;     // [kNQ][Z][nblk]
    double* gpart;    // [kNG][nblk]
    double* planes;   // [kNQ*Z + kNG] reduced sums (planes then globals)
    double* shift;    // [3][Z]
    double* means;    // [4][Z] statistics pass-1 means
    double* tprof;    // [9][Z] time accumulators of the plane profiles
    double* slip;     // [4][kSlipBlocks] slip-velocity partial sums
    double slip_sum[4];  // host FP64 accumulators of the slip sums
    int slip_cols[4];    // columns behind each slip sum
    DevState* st;
    LogRow* ring;
    int ring_cap;
    int nblk;       // blocks of the chunk-mapped kernels (moments, statistics)
    int step_w;     // waves per block of the step kernel
    int step_rows;  // rows per lane of the step kernel (2: FP16 storage, even Z)
    int step_nzp;   // z-pieces of the step kernel (g.nzp, or 2 at 6 or 8 chunks)
    int step_nblk;  // blocks of the step kernel
    int ncu;
    int step_nblk_max;  // capacity of the partial buffers for the step grid
    int step_resident;  // step blocks resident at once on the device
    int block;
    hipStream_t stream;
    double omega, mu, dt;
    double* stage;
    size_t stage_bytes;
    size_t dev_bytes;
    hipEvent_t ev0, ev1;
    struct dnsens* owner;  // ensemble the handle is a member of, or null
    int device;            // the device current at dnsgpu_create
    void* scratch;         // device address of that device's static staging array
    int pair_arena;        // 1: buf[0], buf[1] are ranges of the device's pair arenas (pairarena.hip)
    int pair_candidates;   // destination buffers the autotuner's placement tried (0: none)
    double pair_gain;      // step-time gain of the kept pair over the pair of dnsgpu_create
};

// Entry points that change or advance a member's state are refused while
// the handle belongs to an ensemble (the ensemble advances it).
#define NOT_MEMBER(h)                                                                  \
    do {                                                                               \
        if ((h)->owner)                                                                \
            return set_err(DNSGPU_EARG, "%s: the handle is an ensemble member", __func__); \
    } while (0)

// Calls f with a value of the storage type of a precision (FP64: double,
// FP32: float, FP16: _Float16) as a type tag.
template <typename F>
inline auto with_storage(int precision, F&& f)
{
    if (precision == DNSGPU_FP64) return f(double{});
    if (precision == DNSGPU_FP16) return f(_Float16{});
    return f(float{});
}

// Bytes per stored population of a precision.
inline int storage_bytes(int precision)
{
    return precision == DNSGPU_FP64 ? 8 : (precision == DNSGPU_FP16 ? 2 : 4);
}

// Step arguments of the handle's current state (src) into dst.
__attribute__((visibility("hidden"))) StepArgs step_args(dnsgpu* h, const void* src, void* dst, int nblk);
// Fills the solid band slots of buf[cur] with the pushed values a
// write-side step pulls, when buf[cur] came from a read-side band step.
__attribute__((visibility("hidden"))) int band_pushes(dnsgpu* h);
// Finisher arguments of the handle (row: log row slot or null).
__attribute__((visibility("hidden"))) FinalizeArgs fin_args(dnsgpu* h, bool stats, LogRow* row,
                                                                long long cntr, int from_moments);
// Pair arenas (pairarena.hip): pair_arena_alloc sets buf[0] and buf[1]
// (bytes each) from blocks in different halves of device memory and returns
// 0, or returns nonzero with nothing taken (no arenas reserved on the
// device, or they are full); pair_arena_free returns them.
__attribute__((visibility("hidden"))) int pair_arena_alloc(dnsgpu* h, size_t bytes);
__attribute__((visibility("hidden"))) void pair_arena_free(dnsgpu* h);
// Population buffers of at least this many bytes each are placed by
// dnsgpu_autotune (dnsgpu.hip); smaller pairs can come from the arenas.
constexpr size_t kPairMinBytes = (size_t)2 << 30;
// One single-case step on the handle's stream (kernels enqueued, no sync).
__attribute__((visibility("hidden"))) int do_step(dnsgpu* h, bool stats, LogRow* row, long long cntr);
