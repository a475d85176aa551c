// pairarena.hip -- population pairs of small cases sub-allocated across the
// two halves of device memory.
//
// A copy between two buffers in different halves of MI355X device memory
// runs about 6% faster than within one half, and the step reads one
// population buffer while writing the other.  dnsgpu_autotune places the
// pair of a large case with fresh whole-buffer allocations; small cases (the
// ensemble's members, 40 MB to 1.3 GB per buffer) are many, come and go, and
// their buffers, allocated one after the other, share a half.  A caller that
// is about to create many small cases (the request loop) reserves pair
// arenas on the device (dnsgpu_reserve_pair_arenas): 8 GiB blocks, each
// classed into one half when it is allocated, until each half has the
// bytes asked for free; while they exist, dnsgpu_create takes a small case's
// buf[0] from a block of one half and buf[1] from a block of the other
// (first fit, 2 MiB granules), and falls back to plain allocations when
// they are full.  Measured on the 64-request sweep pool with all members
// resident: FP32 multi-case step 7.00-7.13 -> 6.68-6.70 ms, FP16 5.11 ->
// 4.95 ms.
//
// Classing a new block: it is allocated while the blocks already classed
// are held (so the allocator moves on; it fills one half before the other),
// and timed as the destination of a 4 GiB copy from a block of each half
// (a copy within one half is the slower one); with blocks of one half only,
// against that block's own within-block copy, timed when it was new and
// empty.  Probes only write into the new block.  A block the probe cannot
// class is freed; the surplus blocks of the half reached first are freed at
// the end of the reservation.  All probing happens in the reservation, before
// any member steps.  dnsgpu_release_pair_arenas frees the empty blocks and
// lets the others go as their last range returns.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "handle.h"
#include "rangealloc.h"

namespace {

constexpr size_t kBlock = (size_t)8 << 30;
constexpr size_t kGranule = (size_t)2 << 20;
constexpr size_t kProbe = kBlock / 2;
constexpr size_t kFreeReserve = (size_t)4 << 30;  // device memory a block allocation leaves free
constexpr double kGap = 0.02;  // relative time gap that decides a class
constexpr int kDevices = 64;

__global__ __launch_bounds__(256) void arena_probe_kernel(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

struct Block {
    char* base;
    int half;
    double t_within;  // ms of the kProbe copy within the block, timed when it was empty
    RangeAlloc ra;
    Block(char* b, int h, double t) : base(b), half(h), t_within(t), ra(kBlock, kGranule) {}
};

struct DeviceArenas {
    std::mutex m;
    std::vector<Block*> blocks;
    bool draining = false;
};

DeviceArenas g_dev[kDevices];

// Better of two timed kProbe copies from a to b on the stream (ms), or < 0.
double probe_ms(hipStream_t s, const char* a, char* b)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return -1.0;
    }
    float best = 1e30f;
    bool ok = true;
    for (int r = 0; r < 2 && ok; r++) {
        ok = hipEventRecord(e0, s) == hipSuccess;
        if (ok) {
            hipLaunchKernelGGL(arena_probe_kernel, dim3(1024), dim3(256), 0, s, (const uint4*)a, (uint4*)b,
                               kProbe / 16);
            ok = hipGetLastError() == hipSuccess && hipEventRecord(e1, s) == hipSuccess &&
                 hipEventSynchronize(e1) == hipSuccess;
        }
        float ms = 0.f;
        if (ok) ok = hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (ok) best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ok ? (double)best : -1.0;
}

// Allocates and classes one more block: 1 kept, 0 no device memory for it,
// -1 freed again (the probe could not class it).
int grow(DeviceArenas& d, hipStream_t s)
{
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b < kBlock + kFreeReserve) return 0;
    char* c = nullptr;
    if (hipMalloc((void**)&c, kBlock) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    const double t_within = probe_ms(s, c, c + kProbe);
    Block* ref[2] = {nullptr, nullptr};
    for (Block* b : d.blocks)
        if (!ref[b->half]) ref[b->half] = b;
    int half = -1;
    if (t_within > 0.0) {
        if (!ref[0] && !ref[1]) {
            half = 0;
        } else if (ref[0] && ref[1]) {
            const double t0 = probe_ms(s, ref[0]->base, c), t1 = probe_ms(s, ref[1]->base, c);
            if (t0 > 0.0 && t1 > 0.0) {
                if (t0 > t1 * (1.0 + kGap)) half = 0;
                else if (t1 > t0 * (1.0 + kGap)) half = 1;
            }
        } else {
            Block* r = ref[0] ? ref[0] : ref[1];
            const double t = probe_ms(s, r->base, c);
            if (t > 0.0) {
                const double g = r->t_within / t - 1.0;
                if (g > 2.0 * kGap) half = 1 - r->half;
                else if (g < kGap) half = r->half;
            }
        }
    }
    if (half < 0) {
        (void)hipFree(c);
        return -1;
    }
    d.blocks.push_back(new Block(c, half, t_within));
    return 1;
}

// Bytes of the half's blocks not held by live ranges.
size_t free_bytes(const DeviceArenas& d, int half)
{
    size_t n = 0;
    for (const Block* b : d.blocks)
        if (b->half == half) n += b->ra.capacity() - b->ra.used();
    return n;
}

void free_block(DeviceArenas& d, size_t i)
{
    (void)hipFree(d.blocks[i]->base);
    delete d.blocks[i];
    d.blocks.erase(d.blocks.begin() + (long)i);
}

bool take(DeviceArenas& d, int half, size_t n, void** out)
{
    for (Block* b : d.blocks) {
        size_t off = 0;
        if (b->half == half && b->ra.alloc(n, &off)) {
            *out = b->base + off;
            return true;
        }
    }
    return false;
}

void give_back(DeviceArenas& d, void* p, size_t n)
{
    for (size_t i = 0; i < d.blocks.size(); i++) {
        Block* b = d.blocks[i];
        if ((char*)p < b->base || (char*)p >= b->base + kBlock) continue;
        (void)b->ra.release((size_t)((char*)p - b->base), n);
        if (d.draining && b->ra.empty()) free_block(d, i);
        return;
    }
}

}  // namespace

int pair_arena_alloc(dnsgpu* h, size_t bytes)
{
    if (bytes == 0 || bytes > kBlock || h->device < 0 || h->device >= kDevices) return 1;
    DeviceArenas& d = g_dev[h->device];
    std::lock_guard<std::mutex> lock(d.m);
    if (d.draining || d.blocks.empty()) return 1;
    void* p[2] = {nullptr, nullptr};
    for (int half = 0; half < 2; half++)
        if (!take(d, half, bytes, &p[half])) p[half] = nullptr;
    if (!p[0] || !p[1]) {
        for (int half = 0; half < 2; half++)
            if (p[half]) give_back(d, p[half], bytes);
        return 1;
    }
    h->buf[0] = p[0];
    h->buf[1] = p[1];
    h->pair_arena = 1;
    return 0;
}

void pair_arena_free(dnsgpu* h)
{
    if (!h->pair_arena) return;
    DeviceArenas& d = g_dev[h->device];
    std::lock_guard<std::mutex> lock(d.m);
    const size_t bytes = (size_t)h->g.total * h->esize;
    for (int b = 0; b < 2; b++)
        if (h->buf[b]) give_back(d, h->buf[b], bytes);
    h->buf[0] = h->buf[1] = nullptr;
    h->pair_arena = 0;
}

extern "C" int dnsgpu_reserve_pair_arenas(size_t bytes_per_half, size_t* reserved_per_half)
{
    if (reserved_per_half) *reserved_per_half = 0;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kDevices) return set_err(DNSGPU_EARG, "device %d beyond the arena table", dev);
    DeviceArenas& d = g_dev[dev];
    std::lock_guard<std::mutex> lock(d.m);
    d.draining = false;
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const size_t want = (bytes_per_half + kBlock - 1) / kBlock * kBlock;
    // every block of the device counts at most once per reservation
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) total_b = 0;
    const int max_tries = (int)(total_b / kBlock) + 4;
    for (int tries = 0; tries < max_tries && (free_bytes(d, 0) < want || free_bytes(d, 1) < want); tries++)
        if (grow(d, s) == 0) break;
    (void)hipStreamDestroy(s);
    // the half reached first holds more than asked: its empty surplus goes
    for (int half = 0; half < 2; half++)
        for (size_t i = d.blocks.size(); i-- > 0;) {
            Block* b = d.blocks[i];
            if (b->half == half && b->ra.empty() && free_bytes(d, half) >= want + kBlock) free_block(d, i);
        }
    if (reserved_per_half) *reserved_per_half = std::min(free_bytes(d, 0), free_bytes(d, 1));
    return 0;
}

extern "C" int dnsgpu_release_pair_arenas(void)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kDevices) return set_err(DNSGPU_EARG, "device %d beyond the arena table", dev);
    DeviceArenas& d = g_dev[dev];
    std::lock_guard<std::mutex> lock(d.m);
    d.draining = true;
    for (size_t i = d.blocks.size(); i-- > 0;)
        if (d.blocks[i]->ra.empty()) free_block(d, i);
    return 0;
}

extern "C" int dnsgpu_pair_arena_stats(int device, int* blocks_half0, int* blocks_half1, size_t* bytes_used)
{
    if (device < 0 || device >= kDevices) return set_err(DNSGPU_EARG, "device %d out of range", device);
    DeviceArenas& d = g_dev[device];
    std::lock_guard<std::mutex> lock(d.m);
    int n[2] = {0, 0};
    size_t used = 0;
    for (Block* b : d.blocks) {
        n[b->half]++;
        used += b->ra.used();
    }
    if (blocks_half0) *blocks_half0 = n[0];
    if (blocks_half1) *blocks_half1 = n[1];
    if (bytes_used) *bytes_used = used;
    return 0;
}
