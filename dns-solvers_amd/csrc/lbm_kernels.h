// lbm_kernels.h -- the fused D3Q19 collide-stream step and its helpers
// for gfx950 (MI355X).
//
// Device layout (structure of arrays, one plane per direction):
//     pop[a * P + (x*Y + y)*Z + z],  z = 0..Z-1 the fluid rows
// z (wall-normal) is the fastest index, as k is in the reference's
// Fb(...) macro (definitions.h:93); the wall rows are not stored.  P is the
// plane stride in elements (Nx*Ny*Nz rounded up and staggered).
//
// Work decomposition: one wavefront owns one (x, y) column segment of 64
// wall-normal rows, so every address offset that depends on x and y (the
// nine neighbour columns of the pull, the wall-mask entries) is
// wave-uniform and lives in SGPRs; the per-lane part of every load is z.
// A block holds cpb columns x nzc segments; blocks loop over column groups
// (persistent-style grid sized to fill the device once).
//
// Streaming is "pull" (A-B): the kernel gathers the post-collision values
// of the previous step from the 19 upstream neighbours (periodic wrap in x
// and y, the one-rank halo of xcomputations.c:222-289,435-555), applies the
// wall rule of xcomputations.c:368-375 in pull form, collides and writes
// the node's 19 post-collision values contiguously.  The pull rule for the
// wall-adjacent row, node X, direction d with upstream S = X - e_d in the
// wall row, W = (X.x, X.y, wall), Y = (X.x - e_dx, X.y - e_dy, X.z):
//     s(W) == -2, or s(S) == -2 and s(W) == -1  ->  g_mirror(d)(Y)  (specular)
//     s(S) == -1 and s(W) == -1                ->  g_opp(d)(X)     (bounce-back)
// which writes every slot exactly once (it is the inverse of the push).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3q19.h"

namespace lbm {

using namespace d3q19;

constexpr int kWave = 64;
constexpr int kMaxWaves = 8;     // waves per block
constexpr int kNQ = 5;           // per-plane accumulators: su sv sw sdrho sq
constexpr int kNG = 3;           // global accumulators: jx, mass-1, budget momentum
constexpr int kNAcc = kNQ + kNG;

struct Geom {
    int X, Y, Z;
    int nzc;      // 64-row segments per column
    int cpb;      // columns per block iteration
    int ncols;    // X*Y
    int ngroups;  // ceil(ncols / cpb)
    long long P;  // plane stride in elements
};

struct StepArgs {
    Geom g;
    const void* src;
    void* dst;
    const int8_t* mbot;  // [X*Y] bottom wall codes (-1 / -2)
    const int8_t* mtop;  // [X*Y] top wall codes
    const uint32_t* wword;  // [X*Y] per-column wall words
    double omega;
    const double* G;      // force used by this step (device scalar)
    const double* shift;  // [3][Z] plane-mean shifts for the one-pass variances
    double* part;         // [kNQ][Z][nblk] per-plane partials
    double* gpart;        // [kNG][nblk] global partials
    int nblk;
};

template <typename T>
struct Store;
template <>
struct Store<double> {
    // FP64: the populations themselves.
    static __device__ __forceinline__ double to_f(double v, int) { return v; }
};
template <>
struct Store<float> {
    // FP32: shifted populations g = f - w.
    static __device__ __forceinline__ double to_f(float v, int a) { return WD[a] + (double)v; }
};

__device__ __forceinline__ int wrap_dec(int v, int n) { return v == 0 ? n - 1 : v - 1; }
__device__ __forceinline__ int wrap_inc(int v, int n) { return v == n - 1 ? 0 : v + 1; }

// Column geometry of one work item.  All fields are wave-uniform; they are
// forced into SGPRs with readfirstlane so every address below is a scalar
// base plus the lane's z offset.
struct Column {
    int x, y, xm, xp, ym, yp;
    int Y, Z;
    uint32_t n0;  // own column base (x*Y + y)*Z
};

// Per-thread column (the thread's own column; no uniformity assumed).
__device__ __forceinline__ Column make_column_lane(int col, int X, int Y, int Z)
{
    Column c;
    c.x = col / Y;
    c.y = col - c.x * Y;
    c.xm = c.x == 0 ? X - 1 : c.x - 1;
    c.xp = c.x == X - 1 ? 0 : c.x + 1;
    c.ym = c.y == 0 ? Y - 1 : c.y - 1;
    c.yp = c.y == Y - 1 ? 0 : c.y + 1;
    c.Y = Y;
    c.Z = Z;
    c.n0 = (uint32_t)((c.x * Y + c.y) * Z);
    return c;
}

// Wave-uniform column (col must be the same in every lane of the wave).
__device__ __forceinline__ Column make_column(int col, int X, int Y, int Z)
{
    Column c;
    c.x = __builtin_amdgcn_readfirstlane(col / Y);
    c.y = __builtin_amdgcn_readfirstlane(col - c.x * Y);
    c.xm = c.x == 0 ? X - 1 : c.x - 1;
    c.xp = c.x == X - 1 ? 0 : c.x + 1;
    c.ym = c.y == 0 ? Y - 1 : c.y - 1;
    c.yp = c.y == Y - 1 ? 0 : c.y + 1;
    c.Y = Y;
    c.Z = Z;
    c.n0 = (uint32_t)((c.x * Y + c.y) * Z);
    return c;
}

// Base of the column (x - ex, y - ey) that direction (ex, ey) pulls from.
template <int ex, int ey>
__device__ __forceinline__ uint32_t colbase(const Column& c)
{
    const int sx = ex == 1 ? c.xm : (ex == -1 ? c.xp : c.x);
    const int sy = ey == 1 ? c.ym : (ey == -1 ? c.yp : c.y);
    return (uint32_t)((sx * c.Y + sy) * c.Z);
}

// Wall word of a column: the wall rule of its two wall-adjacent nodes,
// precomputed from the wall masks (wall_word_kernel).
//   bits 0-4 : bottom node, directions 5,11,14,15,18: 1 = specular, 0 = bounce-back
//   bits 5-9 : top node, directions 6,12,13,16,17
//   bits 10-13: bounce-back of the pushes 12, 13 (bottom) and 11, 14 (top),
//              which reverse e_x (momentum budget)
constexpr int kBotDirs[5] = {5, 11, 14, 15, 18};
constexpr int kTopDirs[5] = {6, 12, 13, 16, 17};

template <int a>
constexpr int wall_bit()
{
    for (int i = 0; i < 5; i++) {
        if (kBotDirs[i] == a) return i;
        if (kTopDirs[i] == a) return 5 + i;
    }
    return -1;
}

__device__ __forceinline__ uint32_t load_uniform_u32(const uint32_t* p)
{
    return *(const __attribute__((address_space(4))) uint32_t*)p;
}

// Gather the post-stream populations of node (column c, row z) from the
// post-collision array src (pull), wall rule included.
template <typename T>
__device__ __forceinline__ void pull19(const T* __restrict__ src, long long P, const Column& c, int z, int Z, int Y,
                                       const int8_t* __restrict__ mbot, const int8_t* __restrict__ mtop, T (&v)[Q])
{
    const uint32_t n = c.n0 + (uint32_t)z;
    static_for<0, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        uint32_t idx = colbase<EX[a], EY[a]>(c) + (uint32_t)(z - EZ[a]);
        if constexpr (EZ[a] == 1) idx = (z == 0) ? n : idx;
        if constexpr (EZ[a] == -1) idx = (z == Z - 1) ? n : idx;
        v[a] = src[(long long)a * P + idx];
    });
    // Wall-adjacent rows: the upstream node is in the wall row.
    if (z == 0) {
        const int mW = mbot[c.x * Y + c.y];
        static_for<0, Q>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            if constexpr (EZ[a] == 1) {
                const int sx = EX[a] == 1 ? c.xm : (EX[a] == -1 ? c.xp : c.x);
                const int sy = EY[a] == 1 ? c.ym : (EY[a] == -1 ? c.yp : c.y);
                const int mS = mbot[sx * Y + sy];
                if (mW == -2 || (mS == -2 && mW == -1))
                    v[a] = src[(long long)MIR[a] * P + colbase<EX[a], EY[a]>(c) + (uint32_t)z];
                else
                    v[a] = src[(long long)OPP[a] * P + n];
            }
        });
    }
    if (z == Z - 1) {
        const int mW = mtop[c.x * Y + c.y];
        static_for<0, Q>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            if constexpr (EZ[a] == -1) {
                const int sx = EX[a] == 1 ? c.xm : (EX[a] == -1 ? c.xp : c.x);
                const int sy = EY[a] == 1 ? c.ym : (EY[a] == -1 ? c.yp : c.y);
                const int mS = mtop[sx * Y + sy];
                if (mW == -2 || (mS == -2 && mW == -1))
                    v[a] = src[(long long)MIR[a] * P + colbase<EX[a], EY[a]>(c) + (uint32_t)z];
                else
                    v[a] = src[(long long)OPP[a] * P + n];
            }
        });
    }
}

template <typename T>
__device__ __forceinline__ void load_local19(const T* __restrict__ src, long long P, uint32_t n, T (&v)[Q])
{
    static_for<0, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        v[a] = src[(long long)a * P + n];
    });
}

// Deterministic block reduction of the per-lane accumulators: per-plane
// sums over the cpb column slots (fixed order), global sums over the whole
// block (fixed order), written as partials for the fixed-order finishers.
__device__ __forceinline__ void block_reduce_write(const StepArgs& A, double (&acc)[kNAcc], int wave, int lane,
                                                   int zc, int z, bool stats)
{
    __shared__ double red[kMaxWaves][kNAcc][kWave];
    const int nw = A.g.nzc * A.g.cpb;
#pragma unroll
    for (int q = 0; q < kNAcc; q++) red[wave][q][lane] = acc[q];
    __syncthreads();
    if (stats && wave < A.g.nzc) {
        // wave == zc of slot 0 handles segment zc
        if (z < A.g.Z) {
#pragma unroll
            for (int q = 0; q < kNQ; q++) {
                double s = 0.0;
                for (int sl = 0; sl < A.g.cpb; sl++) s += red[sl * A.g.nzc + zc][q][lane];
                A.part[((long long)q * A.g.Z + z) * A.nblk + blockIdx.x] = s;
            }
        }
    }
    if (wave == 0) {
#pragma unroll
        for (int q = kNQ; q < kNAcc; q++) {
            if (!stats && q == kNQ) continue;
            double s = 0.0;
            for (int w = 0; w < nw; w++) s += red[w][q][lane];
            // fixed butterfly over the 64 lanes
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
            if (lane == 0) A.gpart[(long long)(q - kNQ) * A.nblk + blockIdx.x] = s;
        }
    }
}

// Hot-path gather for one wave: lane z of the wave's column segment.
// All 19 source addresses are a wave-uniform base (SGPRs: direction plane
// + upstream column - e_z) plus the lane's z * sizeof(T), so each load is a
// single global_load with a scalar base and one shared VGPR offset.  The
// wall-adjacent lanes read a harmless in-allocation neighbour first and are
// then overwritten by the wall rule (uniform decisions from the column's
// wall word, one active lane).
template <typename T, bool FIRST>
__device__ __forceinline__ void gather_wave(const T* __restrict__ src, long long P, const Column& c, uint32_t zoff,
                                            int zc, int lane, int zc_top, int lane_top, uint32_t ww, T (&v)[Q])
{
    const char* s8 = (const char*)src;
    static_for<0, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        long long e;
        if constexpr (FIRST)
            e = (long long)a * P + c.n0;
        else
            e = (long long)a * P + (long long)colbase<EX[a], EY[a]>(c) - EZ[a];
        const char* base = s8 + e * (long long)sizeof(T);
        v[a] = *(const T*)(base + zoff);
    });
    if constexpr (!FIRST) {
        if (zc == 0) {
            static_for<0, Q>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                if constexpr (EZ[a] == 1) {
                    const bool spec = (ww >> wall_bit<a>()) & 1u;
                    const long long e = spec ? (long long)MIR[a] * P + colbase<EX[a], EY[a]>(c)
                                             : (long long)OPP[a] * P + c.n0;
                    if (lane == 0) v[a] = src[e];
                }
            });
        }
        if (zc == zc_top) {
            const long long zt = (long long)(zc_top * kWave + lane_top);
            static_for<0, Q>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                if constexpr (EZ[a] == -1) {
                    const bool spec = (ww >> wall_bit<a>()) & 1u;
                    const long long e = spec ? (long long)MIR[a] * P + colbase<EX[a], EY[a]>(c) + zt
                                             : (long long)OPP[a] * P + c.n0 + zt;
                    if (lane == lane_top) v[a] = src[e];
                }
            });
        }
    }
}

// The fused step: pull-stream + wall rule + BGK collision with force +
// per-plane statistics + momentum budget of the next state.
//   FIRST: the source holds post-stream populations (after a load or an
//          init), so the gather is local instead of a pull.
//   STATS: accumulate the per-plane sums for the per-step logs.
// Budget of the next state (the force of the next step, encalc.c:121):
// mass and x-momentum are conserved by the collision except for the force
// term, whose x-moment is exactly G, and streaming is a permutation except
// that a halfway bounce-back reverses e_x:
//     sum f_next = sum f,  sum e_x f_next = sum (j_x + G) - 2 sum_{BB} e_x g
template <typename T, bool FIRST, bool STATS>
__global__ __launch_bounds__(kWave* kMaxWaves) void step_kernel(StepArgs A)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nzc = A.g.nzc;
    const int zc = wave % nzc;
    const int slot = wave / nzc;
    const int X = A.g.X, Y = A.g.Y, Z = A.g.Z;
    const long long P = A.g.P;
    const int z = zc * kWave + lane;
    const bool zvalid = z < Z;
    const int zc_top = (Z - 1) / kWave, lane_top = (Z - 1) % kWave;
    const uint32_t zoff = (uint32_t)z * (uint32_t)sizeof(T);
    const T* __restrict__ src = (const T*)A.src;
    char* __restrict__ d8 = (char*)A.dst;

    double acc[kNAcc];
#pragma unroll
    for (int q = 0; q < kNAcc; q++) acc[q] = 0.0;

    const double G = *A.G;
    double cu = 0.0, cv = 0.0, cw = 0.0;
    if (STATS && zvalid) {
        cu = A.shift[z];
        cv = A.shift[Z + z];
        cw = A.shift[2 * Z + z];
    }

    for (int grp = blockIdx.x; grp < A.g.ngroups; grp += gridDim.x) {
        const int col = __builtin_amdgcn_readfirstlane(grp * A.g.cpb + slot);
        if (col >= A.g.ncols) continue;
        const Column c = make_column(col, X, Y, Z);
        if (!zvalid) continue;
        const uint32_t ww = FIRST ? 0u : load_uniform_u32(A.wword + col);
        T v[Q];
        gather_wave<T, FIRST>(src, P, c, zoff, zc, lane, zc_top, lane_top, ww, v);

        double mass_node, mom_node, jx_raw;
        double g11 = 0., g12 = 0., g13 = 0., g14 = 0.;
        if constexpr (std::is_same<T, double>::value) {
            double ux, uy, uz, dens;
            moments_fp64(v, ux, uy, uz, dens);
            const double u0 = ux / dens, u1 = uy / dens, u2 = uz / dens;
            collide_fp64(v, u0, u1, u2, dens, A.omega, G);
            jx_raw = ux;
            mass_node = dens - 1.0;
            if constexpr (STATS) {
                const double du = u0 - cu, dv = u1 - cv, dw = u2 - cw;
                acc[0] += du;
                acc[1] += dv;
                acc[2] += dw;
                acc[3] += mass_node;
                acc[4] += (du * du + dv * dv) + dw * dw;
            }
            g11 = v[11];
            g12 = v[12];
            g13 = v[13];
            g14 = v[14];
        } else {
            float drho, u0, u1, u2, jx;
            collide_fp32(v, (float)A.omega, (float)G, drho, u0, u1, u2, jx);
            jx_raw = (double)jx;
            mass_node = (double)drho;
            if constexpr (STATS) {
                const double du = (double)u0 - cu, dv = (double)u1 - cv, dw = (double)u2 - cw;
                acc[0] += du;
                acc[1] += dv;
                acc[2] += dw;
                acc[3] += mass_node;
                acc[4] += (du * du + dv * dv) + dw * dw;
            }
            g11 = WD[11] + (double)v[11];
            g12 = WD[12] + (double)v[12];
            g13 = WD[13] + (double)v[13];
            g14 = WD[14] + (double)v[14];
        }
        mom_node = jx_raw + G;
        // bounce-back reverses e_x of the populations pushed into a no-slip
        // wall node from a no-slip wall column (xcomputations.c:370)
        // the wall word of the column holding the state after this step
        const uint32_t wn = FIRST ? load_uniform_u32(A.wword + col) : ww;
        if (zc == 0 && lane == 0)
            mom_node = mom_node - (((wn >> 10) & 1u) ? 2.0 * g12 : 0.0) + (((wn >> 11) & 1u) ? 2.0 * g13 : 0.0);
        if (zc == zc_top && lane == lane_top)
            mom_node = mom_node - (((wn >> 12) & 1u) ? 2.0 * g11 : 0.0) + (((wn >> 13) & 1u) ? 2.0 * g14 : 0.0);
        if constexpr (STATS) acc[5] += jx_raw;
        acc[6] += mass_node;
        acc[7] += mom_node;

        static_for<0, Q>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            const long long e = (long long)a * P + c.n0;
            *(T*)(d8 + e * (long long)sizeof(T) + zoff) = v[a];
        });
    }
    block_reduce_write(A, acc, wave, lane, zc, z, STATS);
}

// Per-column wall words from the wall masks (Appendix-B pull table of the
// push rule xcomputations.c:368-375, and its bounce-back pushes).
__global__ __launch_bounds__(256) void wall_word_kernel(const int8_t* __restrict__ mbot,
                                                        const int8_t* __restrict__ mtop, int X, int Y,
                                                        uint32_t* __restrict__ ww)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= X * Y) return;
    const int x = col / Y, y = col - x * Y;
    uint32_t w = 0;
    for (int side = 0; side < 2; side++) {
        const int8_t* m = side ? mtop : mbot;
        const int* dirs = side ? kTopDirs : kBotDirs;
        const int mW = m[x * Y + y];
        for (int i = 0; i < 5; i++) {
            const int a = dirs[i];
            const int sx = (x - EX[a] + X) % X, sy = (y - EY[a] + Y) % Y;
            const int mS = m[sx * Y + sy];
            if (mW == -2 || (mS == -2 && mW == -1)) w |= 1u << (5 * side + i);
        }
        // pushes that bounce back from this node: e_x = +1 then -1 toward the wall
        const int ap = side ? 11 : 12, am = side ? 14 : 13;
        const int xp = (x + 1) % X, xm = (x - 1 + X) % X;
        if (mW == -1 && m[xp * Y + y] == -1) w |= 1u << (side ? 12 : 10);
        if (mW == -1 && m[xm * Y + y] == -1) w |= 1u << (side ? 13 : 11);
        (void)ap;
        (void)am;
    }
    ww[col] = w;
}

// Moments-only pass over the current state (no collision, no write):
// the same accumulators as the step, used after a load/init to derive the
// first step's force (encalc.c:121 on the initial state) and the shifts.
template <typename T, bool LOCAL>
__global__ __launch_bounds__(kWave* kMaxWaves) void moments_kernel(StepArgs A)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nzc = A.g.nzc;
    const int zc = wave % nzc;
    const int slot = wave / nzc;
    const int X = A.g.X, Y = A.g.Y, Z = A.g.Z;
    const long long P = A.g.P;
    const int z = zc * kWave + lane;
    const bool zvalid = z < Z;
    const T* __restrict__ src = (const T*)A.src;
    double acc[kNAcc];
#pragma unroll
    for (int q = 0; q < kNAcc; q++) acc[q] = 0.0;
    double cu = 0.0, cv = 0.0, cw = 0.0;
    if (zvalid) {
        cu = A.shift[z];
        cv = A.shift[Z + z];
        cw = A.shift[2 * Z + z];
    }
    for (int grp = blockIdx.x; grp < A.g.ngroups; grp += gridDim.x) {
        const int col = __builtin_amdgcn_readfirstlane(grp * A.g.cpb + slot);
        if (col >= A.g.ncols || !zvalid) continue;
        const Column c = make_column(col, X, Y, Z);
        const uint32_t n = c.n0 + (uint32_t)z;
        T v[Q];
        if constexpr (LOCAL)
            load_local19(src, P, n, v);
        else
            pull19(src, P, c, z, Z, Y, A.mbot, A.mtop, v);
        double f[Q];
#pragma unroll
        for (int a = 0; a < Q; a++) f[a] = Store<T>::to_f(v[a], a);
        double ux, uy, uz, dens;
        moments_fp64(f, ux, uy, uz, dens);
        const double u0 = ux / dens, u1 = uy / dens, u2 = uz / dens;
        const double du = u0 - cu, dv = u1 - cv, dw = u2 - cw;
        acc[0] += du;
        acc[1] += dv;
        acc[2] += dw;
        if constexpr (std::is_same<T, double>::value) {
            acc[3] += dens - 1.0;
        } else {
            double dr = 0.0;
#pragma unroll
            for (int a = 0; a < Q; a++) dr += (double)v[a];
            acc[3] += dr;
        }
        acc[4] += (du * du + dv * dv) + dw * dw;
        acc[5] += ux;
        // the state itself is the "next state" for the budget
        acc[6] += 0.0;
        acc[7] += ux;
    }
    block_reduce_write(A, acc, wave, lane, zc, z, true);
}

// Fixed-order reduction of the partials: block b sums row b of a
// [rows][nblk] matrix.
__global__ __launch_bounds__(256) void reduce_rows_kernel(const double* __restrict__ in, int nblk,
                                                          double* __restrict__ out)
{
    __shared__ double sh[256];
    const double* row = in + (long long)blockIdx.x * nblk;
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 256) s += row[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// Device-resident per-handle scalars.
struct DevState {
    double G;          // force of the next step
    double rho_mean;   // V.rho of the last logged step
    double pad[6];
};

struct LogRow {  // mirrors dns_steplog
    double t, K, flowrate, massrate, stup, stdown, stress, pgrad, force, rho;
};

struct FinalizeArgs {
    const double* planes;   // [kNQ][Z] plane sums (may be null when !stats)
    const double* globals;  // [kNG]
    double* shift;          // [3][Z] in/out
    DevState* st;
    LogRow* row;            // null: no log row
    int X, Y, Z;
    int forcing;            // 0 const Q, 1 const F
    double ub, F, mu, TM, dt;
    long long cntr;
    int from_moments;       // 1: globals came from moments_kernel (initial state)
};

// Per-step finisher (encalc.c:29-138 and wshearstr.c:26-154 restated on
// plane sums): the log row of this step, the next step's force and the
// new plane-mean shifts.  One block; thread z owns plane z, sums over z in
// a fixed tree.
__global__ __launch_bounds__(1024) void finalize_kernel(FinalizeArgs A)
{
    __shared__ double sU[1024], sK[1024], uav[1024], dav[1024];
    const int Z = A.Z;
    const double Nxy = (double)A.X * (double)A.Y;
    const double N = Nxy * (double)Z;
    const int t = threadIdx.x;
    const bool have_planes = A.planes != nullptr;
    double uz = 0.0, kz = 0.0;
    if (have_planes && t < Z) {
        const double du = A.planes[0 * Z + t] / Nxy;
        const double dv = A.planes[1 * Z + t] / Nxy;
        const double dw = A.planes[2 * Z + t] / Nxy;
        uz = A.shift[t] + du;
        kz = A.planes[4 * Z + t] - Nxy * ((du * du + dv * dv) + dw * dw);
        uav[t] = uz;
        dav[t] = 1.0 + A.planes[3 * Z + t] / Nxy;
        A.shift[t] = uz;
        A.shift[Z + t] = A.shift[Z + t] + dv;
        A.shift[2 * Z + t] = A.shift[2 * Z + t] + dw;
    }
    sU[t] = uz;
    sK[t] = kz;
    __syncthreads();
    for (int w = 512; w >= 1; w >>= 1) {
        if (t < w && t + w < (int)blockDim.x) {
            sU[t] += sU[t + w];
            sK[t] += sK[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double Gn = A.st->G;
        const double jx = A.globals[0];
        if (A.row && have_planes) {
            double srho = 0.0;
            for (int z = 0; z < Z; z++) srho += dav[z] - 1.0;
            const double denst = 1.0 + srho / (double)Z;
            const double momentum = jx / N;
            const double K = 0.5 * sK[0] / N;
            const double U = sU[0] / (double)Z;
            const double stdown = (-2. * uav[0] + 3. * uav[1] - uav[2]) * (A.mu * dav[0]);
            const double stup = -(2. * uav[Z - 1] - 3. * uav[Z - 2] + uav[Z - 3]) * (A.mu * dav[Z - 3]);
            LogRow r;
            r.t = A.TM + A.dt * (double)A.cntr;
            r.K = K;
            r.flowrate = U * (double)(A.Y * Z);
            r.massrate = momentum * (double)(A.Y * Z);
            r.stup = stup;
            r.stdown = stdown;
            r.stress = 0.5 * (stup + stdown);
            r.pgrad = Gn * (double)Z / 2.;
            r.force = Gn;
            r.rho = denst;
            *A.row = r;
            A.st->rho_mean = denst;
        }
        // force of the next step (encalc.c:121 on the next state)
        if (A.forcing == 0) {
            double rho_next, mom_next;
            if (A.from_moments) {
                double srho = 0.0;
                for (int z = 0; z < Z; z++) srho += A.planes[3 * Z + z];
                rho_next = 1.0 + srho / N;
                mom_next = jx / N;
            } else {
                rho_next = 1.0 + A.globals[1] / N;
                mom_next = A.globals[2] / N;
            }
            A.st->G = rho_next * A.ub - mom_next;
        } else {
            A.st->G = A.F;
        }
    }
}

}  // namespace lbm
