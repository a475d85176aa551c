// copybw.hip -- HBM calibration utility: a plain device-to-device copy of a
// known size (16-byte loads and stores, grid-stride), run at several grid
// sizes.  Used (1) to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE
// counters against bytes that are known exactly and (2) to measure the
// read+write bandwidth a simple streaming kernel reaches on this device,
// the practical ceiling of the step kernel's traffic pattern.
//
// usage: copybw [GiB=4] [reps=5]
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
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
