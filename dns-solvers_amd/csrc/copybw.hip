// copybw.hip -- HBM calibration utility: a plain device-to-device copy of a
// known size (16-byte loads and stores, grid-stride), run at several grid
// sizes.  Used (1) to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE
// counters against bytes that are known exactly and (2) to measure the
// read+write bandwidth a simple streaming kernel reaches on this device,
// the practical ceiling of the step kernel's traffic pattern.
//
// usage: copybw [GiB=4] [reps=5] [all]
//        copybw pairs GiB NBUF
//   pairs: NBUF buffers of GiB allocated one after the other and all held;
//     prints the copy rate within each buffer (first half into the second)
//     and of a whole-buffer copy from every buffer into every other one (a
//     copy between buffers that lie in different halves of device memory
//     is the faster one)
//   all: also the copy variants below, a read-only and a write-only stream:
//     x4       four 16-byte loads in flight per thread
//     chunks   each block copies one contiguous chunk
//     wave64k  each wave copies contiguous 64 KiB chunks (dwordx4), chunks
//              grid-strided over the waves
//     nt       non-temporal loads and stores (__builtin_nontemporal_*)
//     ntstore  plain loads, non-temporal stores
//     b1024    the grid-stride copy with 1024-thread blocks
//     memcpy   hipMemcpyAsync device to device
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef float f4 __attribute__((ext_vector_type(4)));

#define CHK(x)                                                                                   \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));                      \
            return 1;                                                                            \
        }                                                                                        \
    } while (0)

__global__ __launch_bounds__(256) void copy_f4(const f4* __restrict__ a, f4* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ __launch_bounds__(1024) void copy_f4_b1024(const f4* __restrict__ a, f4* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// four independent 16-byte loads in flight per thread (all accesses below n)
__global__ __launch_bounds__(256) void copy_f4_x4(const f4* __restrict__ a, f4* __restrict__ b, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += 4 * stride) {
        const bool k1 = i + stride < n, k2 = i + 2 * stride < n, k3 = i + 3 * stride < n;
        f4 v0 = a[i], v1 = {0, 0, 0, 0}, v2 = v1, v3 = v1;
        if (k1) v1 = a[i + stride];
        if (k2) v2 = a[i + 2 * stride];
        if (k3) v3 = a[i + 3 * stride];
        b[i] = v0;
        if (k1) b[i + stride] = v1;
        if (k2) b[i + 2 * stride] = v2;
        if (k3) b[i + 3 * stride] = v3;
    }
}

// each block copies one contiguous chunk
__global__ __launch_bounds__(256) void copy_f4_chunks(const f4* __restrict__ a, f4* __restrict__ b, size_t n)
{
    const size_t per = (n + gridDim.x - 1) / gridDim.x;
    const size_t lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
    for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) b[i] = a[i];
}

// each wave copies whole 64 KiB chunks (4096 float4), chunk c by wave
// c mod (waves in the grid); the last chunk may be partial (i < n checked)
__global__ __launch_bounds__(256) void copy_f4_wave64k(const f4* __restrict__ a, f4* __restrict__ b, size_t n)
{
    const int lane = threadIdx.x & 63;
    const size_t wave = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6;
    const size_t nwaves = ((size_t)gridDim.x * blockDim.x) >> 6;
    const size_t nchunks = (n + 4095) / 4096;
    for (size_t c = wave; c < nchunks; c += nwaves) {
        const size_t base = c * 4096;
#pragma unroll 4
        for (int k = 0; k < 64; k++) {
            const size_t i = base + (size_t)k * 64 + lane;
            if (i < n) b[i] = a[i];
        }
    }
}

__global__ __launch_bounds__(256) void copy_f4_nt(const f4* __restrict__ a, f4* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

__global__ __launch_bounds__(256) void copy_f4_ntstore(const f4* __restrict__ a, f4* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(a[i], b + i);
}

// read-only stream (the sum is stored only if it equals a value the zeroed input never gives)
__global__ __launch_bounds__(256) void read_f4(const f4* __restrict__ a, f4* __restrict__ out, size_t n)
{
    f4 s = {0, 0, 0, 0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) s += a[i];
    if (s.x == 12345.f) out[0] = s;
}

__global__ __launch_bounds__(256) void write_f4(f4* __restrict__ b, size_t n)
{
    const f4 z = {1, 2, 3, 4};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = z;
}

enum Variant { X4, CHUNKS, WAVE64K, NT, NTSTORE, B1024, MEMCPY, READ, WRITE, NVARIANT };
static const char* kNames[NVARIANT] = {"copy, 4 loads in flight", "copy, block chunks", "copy, wave 64 KiB chunks",
                                       "copy, non-temporal",      "copy, nt stores",    "copy, 1024-thr blocks (g/4)",
                                       "hipMemcpyAsync D2D",      "read only",          "write only"};

static void launch(int k, int g, const f4* a, f4* b, size_t n, size_t bytes)
{
    switch (k) {
    case X4: hipLaunchKernelGGL(copy_f4_x4, dim3(g), dim3(256), 0, 0, a, b, n); break;
    case CHUNKS: hipLaunchKernelGGL(copy_f4_chunks, dim3(g), dim3(256), 0, 0, a, b, n); break;
    case WAVE64K: hipLaunchKernelGGL(copy_f4_wave64k, dim3(g), dim3(256), 0, 0, a, b, n); break;
    case NT: hipLaunchKernelGGL(copy_f4_nt, dim3(g), dim3(256), 0, 0, a, b, n); break;
    case NTSTORE: hipLaunchKernelGGL(copy_f4_ntstore, dim3(g), dim3(256), 0, 0, a, b, n); break;
    case B1024: hipLaunchKernelGGL(copy_f4_b1024, dim3(g / 4 > 0 ? g / 4 : 1), dim3(1024), 0, 0, a, b, n); break;
    case MEMCPY: (void)hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, 0); break;
    case READ: hipLaunchKernelGGL(read_f4, dim3(g), dim3(256), 0, 0, a, b, n); break;
    case WRITE: hipLaunchKernelGGL(write_f4, dim3(g), dim3(256), 0, 0, b, n); break;
    }
}

// best of three timed grid-stride copies (1024 x 256 threads), GB/s of bytes read + written
static double pair_rate(const void* a, void* b, size_t bytes, hipEvent_t e0, hipEvent_t e1)
{
    float best = 1e30f;
    for (int r = 0; r < 3; r++) {
        if (hipEventRecord(e0) != hipSuccess) return -1.0;
        hipLaunchKernelGGL(copy_f4, dim3(1024), dim3(256), 0, 0, (const f4*)a, (f4*)b, bytes / 16);
        if (hipGetLastError() != hipSuccess || hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return -1.0;
        float ms;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
        if (ms < best) best = ms;
    }
    return 2.0 * bytes / best / 1e6;
}

static int pairs(double gib, int nbuf)
{
    const size_t bytes = (size_t)(gib * (1ull << 30)) & ~(size_t)65535;
    if (nbuf < 2 || nbuf > 32) {
        fprintf(stderr, "pairs: 2 to 32 buffers\n");
        return 1;
    }
    void* buf[32];
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    printf("%d buffers of %zu bytes, allocated in order and held\n", nbuf, bytes);
    for (int k = 0; k < nbuf; k++) {
        CHK(hipMalloc(&buf[k], bytes));
        CHK(hipMemset(buf[k], 0, bytes));
    }
    for (int k = 0; k < nbuf; k++) {
        const double r = pair_rate(buf[k], (char*)buf[k] + bytes / 2, bytes / 2, e0, e1);
        if (r < 0) return 1;
        printf("buffer %2d: within (first half -> second half) %7.1f GB/s\n", k, r);
    }
    printf("whole-buffer copies, GB/s (row: source, column: destination)\n    ");
    for (int j = 0; j < nbuf; j++) printf(" %6d", j);
    printf("\n");
    for (int i = 0; i < nbuf; i++) {
        printf("%3d ", i);
        for (int j = 0; j < nbuf; j++) {
            if (i == j) {
                printf(" %6s", "-");
                continue;
            }
            const double r = pair_rate(buf[i], buf[j], bytes, e0, e1);
            if (r < 0) return 1;
            printf(" %6.0f", r);
        }
        printf("\n");
        fflush(stdout);
    }
    for (int k = 0; k < nbuf; k++) CHK(hipFree(buf[k]));
    return 0;
}

int main(int argc, char** argv)
{
    if (argc > 3 && !strcmp(argv[1], "pairs")) return pairs(atof(argv[2]), atoi(argv[3]));
    const double gib = argc > 1 ? atof(argv[1]) : 4.0;
    const int reps = argc > 2 ? atoi(argv[2]) : 5;
    const size_t bytes = (size_t)(gib * (1ull << 30)) & ~(size_t)65535;
    const size_t n = bytes / 16;
    void *a, *b;
    CHK(hipMalloc(&a, bytes));
    CHK(hipMalloc(&b, bytes));
    CHK(hipMemset(a, 0, bytes));
    CHK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    printf("copy of %zu bytes (read %zu + write %zu)\n", bytes, bytes, bytes);
    const int grids[] = {256, 512, 1024, 2048, 8192};
    for (int g : grids) {
        float best = 1e30f;
        for (int r = 0; r < reps; r++) {
            CHK(hipEventRecord(e0));
            hipLaunchKernelGGL(copy_f4, dim3(g), dim3(256), 0, 0, (const f4*)a, (f4*)b, n);
            CHK(hipGetLastError());
            CHK(hipEventRecord(e1));
            CHK(hipEventSynchronize(e1));
            float ms;
            CHK(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0 && ms < best) best = ms;
        }
        printf("grid %5d x 256: %8.1f GB/s  (%.3f ms, best of %d)\n", g, 2.0 * bytes / best / 1e6, best, reps - 1);
    }
    if (argc > 3) {
        const int grids2[] = {512, 1024, 2048, 4096, 16384};
        for (int k = 0; k < NVARIANT; k++) {
            for (int g : grids2) {
                if (k == MEMCPY && g != grids2[0]) break;  // no grid to choose
                float best = 1e30f;
                for (int r = 0; r < reps; r++) {
                    CHK(hipEventRecord(e0));
                    launch(k, g, (const f4*)a, (f4*)b, n, bytes);
                    CHK(hipGetLastError());
                    CHK(hipEventRecord(e1));
                    CHK(hipEventSynchronize(e1));
                    float ms;
                    CHK(hipEventElapsedTime(&ms, e0, e1));
                    if (r > 0 && ms < best) best = ms;
                }
                const double moved = (k == READ || k == WRITE ? 1.0 : 2.0) * bytes;
                if (k == MEMCPY)
                    printf("%-26s            : %8.1f GB/s  (best of %d)\n", kNames[k], moved / best / 1e6, reps - 1);
                else
                    printf("%-26s grid %5d x 256: %8.1f GB/s  (best of %d)\n", kNames[k], g, moved / best / 1e6,
                           reps - 1);
            }
        }
    }
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
