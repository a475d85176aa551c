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

// Column geometry of one work item: the nine upstream column bases.
struct Column {
    int x, y, xm, xp, ym, yp;
    uint32_t cb[3][3];  // cb[ex+1][ey+1] = base of column (x - ex, y - ey)
    uint32_t n0;        // own column base
};

__device__ __forceinline__ Column make_column(int col, int X, int Y, int Z)
{
    Column c;
    c.x = col / Y;
    c.y = col - c.x * Y;
    c.xm = wrap_dec(c.x, X);
    c.xp = wrap_inc(c.x, X);
    c.ym = wrap_dec(c.y, Y);
    c.yp = wrap_inc(c.y, Y);
    const int sx[3] = {c.xp, c.x, c.xm};  // ex = -1, 0, +1 -> x - ex
    const int sy[3] = {c.yp, c.y, c.ym};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) c.cb[i][j] = (uint32_t)((sx[i] * Y + sy[j]) * Z);
    c.n0 = c.cb[1][1];
    return c;
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
        uint32_t idx = c.cb[EX[a] + 1][EY[a] + 1] + (uint32_t)(z - EZ[a]);
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
                    v[a] = src[(long long)MIR[a] * P + c.cb[EX[a] + 1][EY[a] + 1] + (uint32_t)z];
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
                    v[a] = src[(long long)MIR[a] * P + c.cb[EX[a] + 1][EY[a] + 1] + (uint32_t)z];
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

// Sign of each population's x-momentum after the next streaming: -1 when
// the push of (node, a) is a halfway bounce-back (xcomputations.c:370), so
// the momentum budget of the next state is sum e_x g * sign.  Only
// directions with e_x != 0 and e_z pointing into a wall matter.
__device__ __forceinline__ void bb_flags(const Column& c, int z, int Z, int Y, const int8_t* __restrict__ mbot,
                                         const int8_t* __restrict__ mtop, bool& f12, bool& f13, bool& f11, bool& f14)
{
    f12 = f13 = f11 = f14 = false;
    if (z == 0) {
        const bool w = mbot[c.x * Y + c.y] == -1;
        f12 = w && mbot[c.xp * Y + c.y] == -1;  // a=12 (1,0,-1) pushes to x+1
        f13 = w && mbot[c.xm * Y + c.y] == -1;  // a=13 (-1,0,-1) pushes to x-1
    }
    if (z == Z - 1) {
        const bool w = mtop[c.x * Y + c.y] == -1;
        f11 = w && mtop[c.xp * Y + c.y] == -1;  // a=11 (1,0,1)
        f14 = w && mtop[c.xm * Y + c.y] == -1;  // a=14 (-1,0,1)
    }
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

// The fused step: pull-stream + wall rule + BGK collision with force +
// per-plane statistics + momentum budget of the next state.
//   FIRST: the source holds post-stream populations (after a load or an
//          init), so the gather is local instead of a pull.
//   STATS: accumulate the per-plane sums for the per-step logs.
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
    const T* __restrict__ src = (const T*)A.src;
    T* __restrict__ dst = (T*)A.dst;

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
        if (col >= A.g.ncols || !zvalid) continue;
        const Column c = make_column(col, X, Y, Z);
        const uint32_t n = c.n0 + (uint32_t)z;
        T v[Q];
        if constexpr (FIRST)
            load_local19(src, P, n, v);
        else
            pull19(src, P, c, z, Z, Y, A.mbot, A.mtop, v);

        double mass_node, mom_node, jx_raw;
        if constexpr (std::is_same<T, double>::value) {
            double ux, uy, uz, dens;
            moments_fp64(v, ux, uy, uz, dens);
            const double u0 = ux / dens, u1 = uy / dens, u2 = uz / dens;
            collide_fp64(v, u0, u1, u2, dens, A.omega, G);
            jx_raw = ux;
            if constexpr (STATS) {
                const double du = u0 - cu, dv = u1 - cv, dw = u2 - cw;
                acc[0] += du;
                acc[1] += dv;
                acc[2] += dw;
                acc[3] += dens - 1.0;
                acc[4] += (du * du + dv * dv) + dw * dw;
            }
            double m = 0.0;
#pragma unroll
            for (int a = 0; a < Q; a++) m += v[a];
            mass_node = m - 1.0;
            mom_node = ((v[1] + v[7]) + (v[10] + v[11]) + v[12]) - ((v[2] + v[8]) + (v[9] + v[13]) + v[14]);
            bool f12, f13, f11, f14;
            bb_flags(c, z, Z, Y, A.mbot, A.mtop, f12, f13, f11, f14);
            if (f12) mom_node -= 2.0 * v[12];
            if (f13) mom_node += 2.0 * v[13];
            if (f11) mom_node -= 2.0 * v[11];
            if (f14) mom_node += 2.0 * v[14];
        } else {
            float drho, u0, u1, u2, jx;
            collide_fp32(v, (float)A.omega, (float)G, drho, u0, u1, u2, jx);
            jx_raw = (double)jx;
            if constexpr (STATS) {
                const double du = (double)u0 - cu, dv = (double)u1 - cv, dw = (double)u2 - cw;
                acc[0] += du;
                acc[1] += dv;
                acc[2] += dw;
                acc[3] += (double)drho;
                acc[4] += (du * du + dv * dv) + dw * dw;
            }
            double m = 0.0;
#pragma unroll
            for (int a = 0; a < Q; a++) m += (double)v[a];
            mass_node = m;
            mom_node = (((double)v[1] + (double)v[7]) + ((double)v[10] + (double)v[11]) + (double)v[12]) -
                       (((double)v[2] + (double)v[8]) + ((double)v[9] + (double)v[13]) + (double)v[14]);
            bool f12, f13, f11, f14;
            bb_flags(c, z, Z, Y, A.mbot, A.mtop, f12, f13, f11, f14);
            if (f12) mom_node -= 2.0 * (WD[12] + (double)v[12]);
            if (f13) mom_node += 2.0 * (WD[13] + (double)v[13]);
            if (f11) mom_node -= 2.0 * (WD[11] + (double)v[11]);
            if (f14) mom_node += 2.0 * (WD[14] + (double)v[14]);
        }
        if constexpr (STATS) acc[5] += jx_raw;
        acc[6] += mass_node;
        acc[7] += mom_node;

        static_for<0, Q>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            dst[(long long)a * P + n] = v[a];
        });
    }
    block_reduce_write(A, acc, wave, lane, zc, z, STATS);
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
