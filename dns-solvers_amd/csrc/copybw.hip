// copybw.hip -- HBM calibration utility: a plain device-to-device copy of a
// known size (16-byte loads and stores, grid-stride), run at several grid
// sizes.  Used (1) to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE
// counters against bytes that are known exactly and (2) to measure the
// read+write bandwidth a simple streaming kernel reaches on this device,
// the practical ceiling of the step kernel's traffic pattern.
//
// usage: copybw [GiB=4] [reps=5] [all]
//   all: also a copy with four 16-byte loads in flight per thread, a copy
//        by contiguous per-block chunks, a read-only and a write-only stream
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

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

int main(int argc, char** argv)
{
    const double gib = argc > 1 ? atof(argv[1]) : 4.0;
    const int reps = argc > 2 ? atoi(argv[2]) : 5;
    const size_t bytes = (size_t)(gib * (1ull << 30)) & ~(size_t)15;
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
        const char* names[] = {"copy, 4 loads in flight", "copy, block chunks", "read only", "write only"};
        const int grids2[] = {512, 1024, 2048, 4096, 16384};
        for (int k = 0; k < 4; k++) {
            for (int g : grids2) {
                float best = 1e30f;
                for (int r = 0; r < reps; r++) {
                    CHK(hipEventRecord(e0));
                    if (k == 0) hipLaunchKernelGGL(copy_f4_x4, dim3(g), dim3(256), 0, 0, (const f4*)a, (f4*)b, n);
                    if (k == 1) hipLaunchKernelGGL(copy_f4_chunks, dim3(g), dim3(256), 0, 0, (const f4*)a, (f4*)b, n);
                    if (k == 2) hipLaunchKernelGGL(read_f4, dim3(g), dim3(256), 0, 0, (const f4*)a, (f4*)b, n);
                    if (k == 3) hipLaunchKernelGGL(write_f4, dim3(g), dim3(256), 0, 0, (f4*)b, n);
                    CHK(hipGetLastError());
                    CHK(hipEventRecord(e1));
                    CHK(hipEventSynchronize(e1));
                    float ms;
                    CHK(hipEventElapsedTime(&ms, e0, e1));
                    if (r > 0 && ms < best) best = ms;
                }
                const double moved = (k <= 1 ? 2.0 : 1.0) * bytes;
                printf("%-24s grid %5d x 256: %8.1f GB/s  (best of %d)\n", names[k], g, moved / best / 1e6, reps - 1);
            }
        }
    }
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
