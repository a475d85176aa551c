// lbm_ens_kernels.h -- multi-case kernels of the ensemble: many resident
// channel cases (members) of different grids, textures and Reynolds
// numbers advance together, one launch of each kernel per step for all of
// them.
//
// Work is described by a tile table: a tile is a few columns of one
// member, updated by one wave.  Tile t of a member with S tiles holds the
// columns t, t + S, t + 2S, ...: the waves that run at the same time then
// update runs of adjacent columns, as the single-case kernel's do (about 2%
// faster than consecutive columns per tile).  The step kernel's grid is sized to
// what the device holds at once and its waves loop over the table with no
// block-level synchronisation; each tile writes one contiguous record of
// its per-plane and global sums, so every member's reductions run in an
// order fixed by the member alone (its tiles, then the fixed-order record
// reduction): a member's results do not depend on which other members
// share the launches, nor on the grid.
#pragma once

#include "lbm_kernels.h"

namespace lbm {

constexpr int kEnsWaves = 4;  // waves per block of the ensemble step kernel

// Columns per tile of a member: a function of the member alone.  Wider
// tiles write fewer records (kNQ Z doubles each) but lengthen the last round
// of tiles a launch waits for and leave a small pool fewer waves.  Measured
// on the 64-solve pool with every member resident, whole step: FP32 8/16
// columns 6.79-6.85 ms, 16/32 columns 7.06 ms, 32/64 columns 7.46 ms; FP16
// 8/16 columns 5.01-5.06 ms, 16/32 columns 4.89-5.00 ms, but the FP16 sweep
// with Poisson arrivals (a few members resident at a time) busy 62.6 -> 75.9 s.
__host__ __device__ inline int ens_tile_cols(int ncols) { return ncols >= 32768 ? 16 : 8; }

// Doubles per tile record: kNQ per-plane sums for each of Z planes, then
// the kNG global sums.
__host__ __device__ inline int ens_record(int Z) { return kNQ * Z + kNG; }

// A member taller than 512 rows is split into z-pieces as in the
// single-case step (z_pieces): each of its column tiles is one tile per
// piece, with a record of its own.  Planes of one tile's plane sums:
__host__ __device__ inline int ens_plane_rows(int Z)
{
    const int nzc = (Z + kWave - 1) / kWave, P = z_pieces(nzc);
    return P == 1 ? Z : (nzc + P - 1) / P * kWave;
}

// LDS of the step kernel: each wave's private per-plane sums (zslice planes:
// the largest ens_plane_rows of the members) and, with SHL, a copy of its
// member's [3][Z] plane-mean shifts (read by every column of the tile: from
// LDS about 5% faster in FP16 storage).
__host__ __device__ inline size_t ens_lds_bytes(int zslice, int zmax, bool shl)
{
    return sizeof(double) * kEnsWaves * ((size_t)kNQ * zslice + (shl ? 3 * (size_t)zmax : 0));
}

// Largest dynamic LDS the step kernel asks for (the shifts stay in global
// memory above it): the plane sums of a 512-row piece fit it.
constexpr size_t kEnsLdsMax = 65536;

// One member as the kernels see it.
struct EnsMember {
    StepArgs a;      // buffers as at even steps of a launch sequence (src = state);
                     // part: the member's tile records [ntiles][ens_record(Z)], nblk: ntiles
    double* planes;  // [kNQ*Z + kNG] reduced sums (the finisher's input)
    int chunk0;      // first 64-row chunk of this member in the reduction's chunk space
    int nchunks;     // ceil(ens_record(Z) / 64)
    int pieces;      // z-pieces per column (z_pieces): records lt * pieces + p
    int cpp;         // 64-row chunks per piece
};

struct EnsTile {
    int m;     // member
    int col0;  // first column
    int ncol;  // columns
    int lt;    // record slot within the member (column tile * pieces + piece)
    int stride;  // column stride (the member's column-tile count)
    int piece;   // z-piece of the columns (0 below 513 rows)
    int pad[2];
};

// Orders the wave's LDS accesses across lanes (the zeroing and record
// loops touch other lanes' plane sums): a wavefront-scope fence that the
// compiler does not move memory operations across.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The fused step (step_column) over every tile of every member, one wave
// per tile.  phase selects the buffer direction: even steps read a.src,
// odd steps a.dst.
// PIECES: some member is taller than 512 rows (its tiles cover one z-piece
// each); without it every tile covers whole columns, with compile-time
// piece bounds.
template <typename T, bool SHL, bool PIECES = false>
static __global__ __launch_bounds__(kWave* kEnsWaves) void ens_step_kernel(const EnsMember* __restrict__ mem,
                                                                          const EnsTile* __restrict__ tiles,
                                                                          int ntiles, int zslice, int zmax, int phase)
{
    extern __shared__ double lds[];  // ens_lds_bytes(zslice, zmax, SHL)
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    double* mine = lds + (size_t)wave * ((size_t)kNQ * zslice + (SHL ? 3 * (size_t)zmax : 0));
    double* shl = mine + kNQ * zslice;
    const int nw = gridDim.x * kEnsWaves;
    for (int t = blockIdx.x * kEnsWaves + wave; t < ntiles; t += nw) {
        const EnsTile tl = load_uniform(tiles + t);
        // only the fields the column update reads (scalar loads; fewer live SGPRs)
        const StepArgs* ma = &mem[tl.m].a;
        StepArgs A;
        A.g.Z = load_uniform(&ma->g.Z);
        A.g.nzc = load_uniform(&ma->g.nzc);
        const void* s0 = load_uniform(&ma->src);
        void* d0 = load_uniform(&ma->dst);
        A.src = phase ? d0 : s0;
        A.dst = phase ? const_cast<void*>(s0) : d0;
        A.wword = load_uniform(&ma->wword);
        A.bnbr = load_uniform(&ma->bnbr);
        A.cnbr = load_uniform(&ma->cnbr);
        A.omega = load_uniform(&ma->omega);
        A.G = load_uniform(&ma->G);
        A.shift = load_uniform(&ma->shift);
        A.part = load_uniform(&ma->part);
        const int Z = A.g.Z;
        // the tile's rows: chunks [c0, c1), plane sums of planes zbase .. zbase + pz - 1
        int c0 = 0, c1 = A.g.nzc, pz = Z, zbase = 0;
        if constexpr (PIECES) {
            if (load_uniform(&mem[tl.m].pieces) > 1) {
                const int cpp = load_uniform(&mem[tl.m].cpp);
                c0 = tl.piece * cpp;
                c1 = min(A.g.nzc, c0 + cpp);
                pz = cpp * kWave;
                zbase = c0 * kWave;
            }
        }
        for (int i = lane; i < kNQ * pz; i += kWave) mine[i] = 0.0;
        if constexpr (SHL)
            for (int i = lane; i < 3 * Z; i += kWave) shl[i] = A.shift[i];
        wave_lds_sync();
        StepAcc acc;
        const double G = load_uniform(A.G);
        // one code path for every member: a member without solid bands has
        // all-zero band records, so no chunk takes the band branch
        for (int k = 0; k < tl.ncol; k++)
            step_column<T, false, true, true>(A, tl.col0 + k * tl.stride, lane, G, SHL ? shl : A.shift, mine, acc,
                                              c0, c1, pz, zbase);
        wave_lds_sync();
        double* rec = A.part + (long long)tl.lt * ens_record(Z);
        if constexpr (PIECES) {
            // this piece's planes only (the reduction reads no other plane of this record)
            for (int i = lane; i < kNQ * pz; i += kWave) {
                const int q = i / pz, z = zbase + (i - q * pz);
                if (z < Z) rec[q * Z + z] = mine[i];
            }
        } else {
            for (int i = lane; i < kNQ * Z; i += kWave) rec[i] = mine[i];
        }
        double g[kNG] = {acc.jx, acc.mass, acc.mom, acc.sq};
#pragma unroll
        for (int q = 0; q < kNG; q++) {
            double s = g[q];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
            g[q] = s;
        }
        if (lane < kNG) {
            double v = g[0];
#pragma unroll
            for (int q = 1; q < kNG; q++)
                if (lane == q) v = g[q];
            rec[kNQ * Z + lane] = v;
        }
        wave_lds_sync();  // the next tile zeroes the slice again
    }
}

constexpr int kEnsReduceWaves = 16;  // waves per block of the record reduction

// Fixed-order reduction of every member's tile records: block b covers 64
// consecutive record rows (one per lane) of the member whose chunks
// contain b.  Wave w sums the records t = w + 16 (4k + j) into accumulator
// j in increasing k, the four accumulators are added pairwise, then the 16
// wave sums in a fixed tree.  Reads are 512-byte record segments.  A member
// in z-pieces sums a plane row over the records of the piece holding the
// plane only (record lt * pieces + piece, in increasing lt), its global rows
// over every record.
static __global__ __launch_bounds__(kWave* kEnsReduceWaves) void ens_reduce_kernel(const EnsMember* __restrict__ mem,
                                                                                  int nmem)
{
    __shared__ double sh[kEnsReduceWaves][kWave];
    const int b = blockIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    int m = 0;
    while (m + 1 < nmem && load_uniform(&mem[m + 1].chunk0) <= b) m++;
    const int Z = load_uniform(&mem[m].a.g.Z);
    const int nt = load_uniform(&mem[m].a.nblk);
    const int nrec = ens_record(Z);
    const int r = (b - load_uniform(&mem[m].chunk0)) * kWave + lane;
    const double* part = load_uniform(&mem[m].a.part);
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (r < nrec) {
        // records first, first + stride, ... (cnt of them) hold row r
        int first = 0, stride = 1, cnt = nt;
        const int P = load_uniform(&mem[m].pieces);
        if (P > 1 && r < kNQ * Z) {
            first = (r % Z) / kWave / load_uniform(&mem[m].cpp);
            stride = P;
            cnt = nt / P;
        }
        constexpr int S = kEnsReduceWaves;
        int t = wave;
        for (; t + 3 * S < cnt; t += 4 * S) {
#pragma unroll
            for (int j = 0; j < 4; j++) s[j] += part[(long long)(first + (t + j * S) * stride) * nrec + r];
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (t + j * S < cnt) s[j] += part[(long long)(first + (t + j * S) * stride) * nrec + r];
    }
    sh[wave][lane] = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();
    for (int w = kEnsReduceWaves / 2; w >= 1; w >>= 1) {
        if (wave < w) sh[wave][lane] += sh[wave + w][lane];
        __syncthreads();
    }
    if (wave == 0 && r < nrec) load_uniform(&mem[m].planes)[r] = sh[0][lane];
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
