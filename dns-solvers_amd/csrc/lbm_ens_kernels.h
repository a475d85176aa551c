// lbm_ens_kernels.h -- multi-case kernels of the ensemble: many resident
// channel cases (members) of different grids, textures and Reynolds
// numbers advance together, one launch of each kernel per step for all of
// them.
//
// Work is described by a tile table: a tile is kEnsTileCols consecutive
// columns of one member.  The step kernel's grid is sized to what the
// device holds at once and loops over the table; each tile writes its own
// set of partial sums, so every member's reductions run in an order fixed
// by the member alone (tile count, then the fixed-order row reduction):
// a member's results do not depend on which other members share the
// launches, nor on the grid.
#pragma once

#include "lbm_kernels.h"

namespace lbm {

constexpr int kEnsTileCols = 16;  // columns per tile
constexpr int kEnsWaves = 4;      // waves per block of the ensemble step kernel

// Largest Z a member may have: the block's per-wave plane sums, global
// partials and shifts in 64 KiB of LDS with kEnsWaves waves.
constexpr int ens_max_z()
{
    return (65536 / (int)sizeof(double) - kEnsWaves * kNG * kWave) / (kEnsWaves * kNQ + 3);
}

// One member as the kernels see it.
struct EnsMember {
    StepArgs a;      // buffers as at even steps of a launch sequence (src = state);
                     // part/gpart/nblk: the member's tile partials and tile count
    double* planes;  // [kNQ*Z + kNG] reduced sums (the finisher's input)
    int row0;        // first row of this member in the reduction's row space
    int nrows;       // kNQ*Z + kNG
};

struct EnsTile {
    int m;     // member
    int col0;  // first column
    int ncol;  // columns (kEnsTileCols, fewer in a member's last tile)
    int lt;    // tile index within the member (partial slot)
};

// Wave-uniform load through the constant address space (scalar loads) of
// a table entry that no kernel writes.
template <typename S>
__device__ __forceinline__ S load_uniform(const S* p)
{
    static_assert(sizeof(S) % 4 == 0, "whole dwords");
    constexpr int n = (int)(sizeof(S) / 4);
    const __attribute__((address_space(4))) uint32_t* q = (const __attribute__((address_space(4))) uint32_t*)p;
    uint32_t w[n];
#pragma unroll
    for (int i = 0; i < n; i++) w[i] = q[i];
    S out;
    __builtin_memcpy(&out, w, sizeof(S));
    return out;
}

// The fused step (step_column) over every tile of every member.  phase
// selects the buffer direction: even steps read a.src, odd steps a.dst.
template <typename T>
static __global__ __launch_bounds__(kWave* kEnsWaves) void ens_step_kernel(const EnsMember* __restrict__ mem,
                                                                   const EnsTile* __restrict__ tiles, int ntiles,
                                                                   int phase)
{
    extern __shared__ double lds[];  // step_lds_bytes(max Z, kEnsWaves, true)
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    constexpr int W = kEnsWaves;
    double* red = lds;
    double* planes = lds + (size_t)W * kNG * kWave;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const EnsTile tl = load_uniform(tiles + t);
        StepArgs A = load_uniform(&mem[tl.m].a);
        if (phase) {
            void* s = const_cast<void*>(A.src);
            A.src = A.dst;
            A.dst = s;
        }
        const int Z = A.g.Z;
        double* mine = planes + (size_t)wave * kNQ * Z;
        double* shl = planes + (size_t)W * kNQ * Z;
        for (int i = lane; i < kNQ * Z; i += kWave) mine[i] = 0.0;
        for (int i = threadIdx.x; i < 3 * Z; i += blockDim.x) shl[i] = A.shift[i];
        __syncthreads();
        StepAcc acc;
        const double G = *A.G;
        const int cend = tl.col0 + tl.ncol;
        if (A.bands) {
            for (int c = tl.col0 + wave; c < cend; c += W)
                step_column<T, false, true, true>(A, c, lane, G, shl, mine, acc);
        } else {
            for (int c = tl.col0 + wave; c < cend; c += W)
                step_column<T, false, true, false>(A, c, lane, G, shl, mine, acc);
        }
        step_block_reduce<true>(red, planes, acc, Z, W, wave, lane, A.part, A.gpart, A.nblk, tl.lt);
        __syncthreads();  // the next tile reuses the LDS
    }
}

// Fixed-order reduction of every member's tile partials; block b is row
// b - row0 of the member whose rows contain b (plane rows kNQ*Z, then
// kNG global rows).  Same arithmetic as reduce_rows_kernel over nblk tiles.
static __global__ __launch_bounds__(256) void ens_reduce_kernel(const EnsMember* __restrict__ mem, int nmem)
{
    __shared__ double sh[256];
    const int b = blockIdx.x;
    int m = 0;
    while (m + 1 < nmem && load_uniform(&mem[m + 1].row0) <= b) m++;
    const int r = b - load_uniform(&mem[m].row0);
    const int Z = load_uniform(&mem[m].a.g.Z);
    const int nt = load_uniform(&mem[m].a.nblk);
    const double* in = r < kNQ * Z ? load_uniform(&mem[m].a.part) + (long long)r * nt
                                   : load_uniform(&mem[m].a.gpart) + (long long)(r - kNQ * Z) * nt;
    double s = 0.0;
    for (int i = threadIdx.x; i < nt; i += 256) s += in[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) load_uniform(&mem[m].planes)[r] = sh[0];
}

// Finisher of every member (block m = member m): log row into slot
// `rowslot` of the member's ring, the next step's force, the new shifts.
static __global__ __launch_bounds__(1024) void ens_finalize_kernel(const FinalizeArgs* __restrict__ fa, int rowslot)
{
    __shared__ FinalizeShared S;
    FinalizeArgs A = fa[blockIdx.x];
    if (A.row) A.row += rowslot;
    finalize_body(A, S);
}

// Per-member probe after a launch sequence: out[m] = {steps, rows
// accumulated, sum of p-grad, force, last row's wall stress}.
static __global__ __launch_bounds__(64) void ens_probe_kernel(const FinalizeArgs* __restrict__ fa, int nmem, int lastslot,
                                                       double* __restrict__ out)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmem) return;
    const DevState* st = fa[m].st;
    out[5 * m + 0] = st->steps;
    out[5 * m + 1] = st->n_rows;
    out[5 * m + 2] = st->s_pgrad;
    out[5 * m + 3] = st->G;
    out[5 * m + 4] = lastslot >= 0 && fa[m].row ? fa[m].row[lastslot].stress : 0.0;
}

}  // namespace lbm
