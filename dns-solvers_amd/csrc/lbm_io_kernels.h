// lbm_io_kernels.h -- kernels that run at output cadence or at set-up:
// checkpoint conversion (SoA device state <-> AoS lbf layout), the
// statistics.c two-pass plane statistics, vel_print.c moments and the
// initial-field generators.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_kernels.h"

namespace lbm {

// ---------------------------------------------------------------- helpers

template <typename T>
__device__ __forceinline__ T from_f(double f, int a);
template <>
__device__ __forceinline__ double from_f<double>(double f, int) { return f; }
template <>
__device__ __forceinline__ float from_f<float>(double f, int a) { return (float)(f - WD[a]); }
template <>
__device__ __forceinline__ _Float16 from_f<_Float16>(double f, int a)
{
    return Store<_Float16>::from_g((float)(f - WD[a]));
}

// Populations of the current post-stream state of node (col, z) as f (FP64).
template <typename T>
__device__ __forceinline__ void state_f(const T* src, const Column& c, int z, const int8_t* mbot,
                                        const int8_t* mtop, const uint32_t* bands, bool local, double (&f)[Q])
{
    T v[Q];
    if (local)
        load_local19(src, c, z, v);
    else
        pull19(src, c, z, mbot, mtop, bands, v);
#pragma unroll
    for (int a = 0; a < Q; a++) f[a] = Store<T>::to_f(v[a], a);
}

// ------------------------------------------------------- checkpoint I/O

struct AosArgs {
    Geom g;
    const void* cur;    // state buffer (post-stream if local, else post-collision)
    const void* other;  // the "f" half when local (loaded previous post-collision)
    void* cur_w;
    void* other_w;
    const int8_t* mbot;
    const int8_t* mtop;
    const uint32_t* bands;
    int local;
    int x0, nx;
    double* ftemp;  // staging [nx][Y][Z+2][19]
    double* f;
};

// Device state -> lbf halves for x-planes [x0, x0+nx).  One thread per
// (node, k) including the wall rows, which are written as zeros (the
// reference never touches them; an init or checkpoint from the reference
// holds zeros there).
template <typename T>
static __global__ __launch_bounds__(256) void save_aos_kernel(AosArgs A)
{
    const int Z = A.g.Z, Y = A.g.Y, X = A.g.X;
    const int Zd = Z + 2;
    const long long total = (long long)A.nx * Y * Zd;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i % Zd);
        const long long cl = i / Zd;
        const int y = (int)(cl % Y);
        const int x = A.x0 + (int)(cl / Y);
        double* ft = A.ftemp + i * Q;
        double* fo = A.f + i * Q;
        if (k == 0 || k == Zd - 1 || (A.bands && band_solid(A.bands[x * Y + y], k - 1, Z))) {
            // wall rows and solid band nodes: zeros, as in a reference
            // checkpoint (the solver never writes solid nodes)
#pragma unroll
            for (int a = 0; a < Q; a++) {
                ft[a] = 0.0;
                fo[a] = 0.0;
            }
            continue;
        }
        const int z = k - 1;
        const Column c = make_column_lane(x * Y + y, X, Y, Z);
        double fs[Q];
        state_f<T>((const T*)A.cur, c, z, A.mbot, A.mtop, A.bands, A.local != 0, fs);
        const T* ob = (const T*)(A.local ? A.other : A.cur);
#pragma unroll
        for (int a = 0; a < Q; a++) {
            ft[a] = fs[a];
            fo[a] = Store<T>::to_f(ob[c.cb + (long long)a * Z + z], a);
        }
    }
}

// lbf halves -> device: ftemp into cur_w, f into other_w (both local).
template <typename T>
static __global__ __launch_bounds__(256) void load_aos_kernel(AosArgs A)
{
    const int Z = A.g.Z, Y = A.g.Y;
    const long long total = (long long)A.nx * Y * Z;
    T* cw = (T*)A.cur_w;
    T* ow = (T*)A.other_w;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % Z);
        const long long cl = i / Z;
        const int y = (int)(cl % Y);
        const int x = A.x0 + (int)(cl / Y);
        const long long src = (cl * (Z + 2) + (z + 1)) * Q;
        const long long cb = ((long long)x * Y + y) * Q * Z + z;
#pragma unroll
        for (int a = 0; a < Q; a++) {
            cw[cb + (long long)a * Z] = from_f<T>(A.ftemp[src + a], a);
            ow[cb + (long long)a * Z] = from_f<T>(A.f[src + a], a);
        }
    }
}

// ------------------------------------------------------- vel_print

struct FieldArgs {
    Geom g;
    const void* cur;
    const int8_t* mbot;
    const int8_t* mtop;
    const uint32_t* bands;
    int local;
    int x0, nx;
    double* vel;  // [nx][Y][Z+2][3]
    double* den;  // [nx][Y][Z+2]
};

// vel_print.c:65-99: u = (sum e f)/dens, zeros where dens <= 0 (wall rows).
template <typename T>
static __global__ __launch_bounds__(256) void fields_kernel(FieldArgs A)
{
    const int Z = A.g.Z, Y = A.g.Y, X = A.g.X;
    const int Zd = Z + 2;
    const long long total = (long long)A.nx * Y * Zd;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i % Zd);
        const long long cl = i / Zd;
        const int y = (int)(cl % Y);
        const int x = A.x0 + (int)(cl / Y);
        double ux = 0., uy = 0., uz = 0., dens = 0.;
        if (k >= 1 && k <= Z && !(A.bands && band_solid(A.bands[x * Y + y], k - 1, Z))) {
            const Column c = make_column_lane(x * Y + y, X, Y, Z);
            double f[Q];
            state_f<T>((const T*)A.cur, c, k - 1, A.mbot, A.mtop, A.bands, A.local != 0, f);
            moments_fp64(f, ux, uy, uz, dens);
        }
        if (dens > 0.) {
            ux /= dens;
            uy /= dens;
            uz /= dens;
        } else {
            ux = 0.;
            uy = 0.;
            uz = 0.;
            dens = 0.;
        }
        A.vel[i * 3 + 0] = ux;
        A.vel[i * 3 + 1] = uy;
        A.vel[i * 3 + 2] = uz;
        A.den[i] = dens;
    }
}

// ------------------------------------------------------- statistics.c

struct StatArgs {
    Geom g;
    const void* cur;
    const int8_t* mbot;
    const int8_t* mtop;
    const uint32_t* bands;
    int local;
    int pass;             // 1: sums of u,v,w,rho ; 2: sums of squared fluctuations
    const double* means;  // [4][Z] UAVG,VAVG,WAVG,DAVG (pass 2)
    double* part;         // [5][Z][nblk]
    int nblk;
};

// Two-pass plane statistics (statistics.c:31-100); per-lane sums over the
// wave's columns, then fixed-order block sums written as partials.
template <typename T>
static __global__ __launch_bounds__(kWave* kMaxWaves) void stats_kernel(StatArgs A)
{
    __shared__ double red[kMaxWaves][5][kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int cpp = A.g.cpp;
    const int zcl = wave % cpp;  // chunk within the block's z-piece (blockIdx.y)
    const int slot = wave / cpp;
    const int X = A.g.X, Y = A.g.Y, Z = A.g.Z;
    const int z = (blockIdx.y * cpp + zcl) * kWave + lane;
    const bool zvalid = z < Z;
    double acc[5] = {0., 0., 0., 0., 0.};
    double mu = 0., mv = 0., mw = 0., md = 0.;
    if (A.pass == 2 && zvalid) {
        mu = A.means[z];
        mv = A.means[Z + z];
        mw = A.means[2 * Z + z];
        md = A.means[3 * Z + z];
    }
    for (int grp = blockIdx.x; grp < A.g.ngroups; grp += gridDim.x) {
        const int col = __builtin_amdgcn_readfirstlane(grp * A.g.cpb + slot);
        if (col >= A.g.ncols || !zvalid) continue;
        if (A.bands && band_solid(A.bands[col], z, Z)) continue;  // statistics.c sums fluid nodes
        const Column c = make_column(col, X, Y, Z);
        double f[Q];
        state_f<T>((const T*)A.cur, c, z, A.mbot, A.mtop, A.bands, A.local != 0, f);
        double u, v, w, dens;
        moments_fp64(f, u, v, w, dens);
        if (A.pass == 1) {
            acc[0] += (u / dens);
            acc[1] += (v / dens);
            acc[2] += (w / dens);
            acc[3] += dens;
        } else {
            const double UF = (u / dens) - mu, VF = (v / dens) - mv, WF = (w / dens) - mw, DF = dens - md;
            acc[0] += UF * UF;
            acc[1] += VF * VF;
            acc[2] += WF * WF;
            acc[3] += UF * WF;
            acc[4] += DF * DF;
        }
    }
#pragma unroll
    for (int q = 0; q < 5; q++) red[wave][q][lane] = acc[q];
    __syncthreads();
    if (wave < cpp && z < Z) {
#pragma unroll
        for (int q = 0; q < 5; q++) {
            double s = 0.0;
            for (int sl = 0; sl < A.g.cpb; sl++) s += red[sl * cpp + zcl][q][lane];
            A.part[((long long)q * Z + z) * A.nblk + blockIdx.x] = s;
        }
    }
}

// ------------------------------------------------------- time averages

// Adds one sample of the plane statistics (means from pass 1, fluctuation
// sums from pass 2 of stats_kernel) into FP64 time accumulators:
//   tprof[0..3][z] += plane means of u, v, w, rho
//   tprof[4..6][z] += plane means of u^2, v^2, w^2  (variance + mean^2)
//   tprof[7][z]    += plane mean of u w            (covariance + product)
//   tprof[8][z]    += plane mean of rho^2
// so the time-averaged rms is sqrt(<u^2> - <u>^2) over planes and time; the
// plane means are over the nf[z] fluid nodes of each plane.
static __global__ __launch_bounds__(256) void tavg_profile_kernel(const double* __restrict__ means,
                                                           const double* __restrict__ fl, int Z,
                                                           const double* __restrict__ nf,
                                                           double* __restrict__ tprof, DevState* st)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z < Z) {
        const double nxy = nf[z] > 0.0 ? nf[z] : 1.0;
        const double u = means[z], v = means[Z + z], w = means[2 * Z + z], r = means[3 * Z + z];
        tprof[0 * Z + z] += u;
        tprof[1 * Z + z] += v;
        tprof[2 * Z + z] += w;
        tprof[3 * Z + z] += r;
        tprof[4 * Z + z] += fl[0 * Z + z] / nxy + u * u;
        tprof[5 * Z + z] += fl[1 * Z + z] / nxy + v * v;
        tprof[6 * Z + z] += fl[2 * Z + z] / nxy + w * w;
        tprof[7 * Z + z] += fl[3 * Z + z] / nxy + u * w;
        tprof[8 * Z + z] += fl[4 * Z + z] / nxy + r * r;
    }
    if (z == 0) st->n_prof += 1.0;
}

// Mean streamwise velocity of the wall-adjacent fluid rows over slip
// columns (super-hydrophobic stripes/posts) and over no-slip columns, for
// the slip velocity; one thread per column, fixed-order block sums.
//   out[q][blk], q = sum u over bottom-slip, bottom-noslip, top-slip, top-noslip
template <typename T>
static __global__ __launch_bounds__(256) void slip_velocity_kernel(StatArgs A, double* __restrict__ out)
{
    __shared__ double sh[4][256];
    const int X = A.g.X, Y = A.g.Y, Z = A.g.Z;
    double acc[4] = {0., 0., 0., 0.};
    for (int col = blockIdx.x * blockDim.x + threadIdx.x; col < X * Y; col += gridDim.x * blockDim.x) {
        const Column c = make_column_lane(col, X, Y, Z);
        for (int side = 0; side < 2; side++) {
            const int z = side ? Z - 1 : 0;
            if (A.bands && band_solid(A.bands[col], z, Z)) continue;
            const int8_t m = side ? A.mtop[col] : A.mbot[col];
            double f[Q];
            state_f<T>((const T*)A.cur, c, z, A.mbot, A.mtop, A.bands, A.local != 0, f);
            double ux, uy, uz, dens;
            moments_fp64(f, ux, uy, uz, dens);
            acc[2 * side + (m == -2 ? 0 : 1)] += ux / dens;
        }
    }
    for (int q = 0; q < 4; q++) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 4; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

// ------------------------------------------------------- initial fields

struct InitArgs {
    Geom g;
    void* cur;
    void* other;
    int kind;
    double uc, amp;
    unsigned long long seed;
    // TURBULENT: tables of the seeded field (turbulent_tables in dnsgpu.hip)
    int nm;              // Fourier modes
    double nrm;          // scale of the fluctuation (rms |u'| = amp * centreline velocity)
    const double* mode;  // [nm][kModeRec]: i, sc, a1, a2, kx, ky
    const double* cx;    // [mx+1][X] cos(2 pi i x / X)
    const double* sx;    // [mx+1][X] sin(2 pi i x / X)
    const double* by;    // [nm][4][Y] cos, sin of (ky y + p1), then of (ky y + p2)
    const double* zt;    // [3][Z] log-law mean u(z), envelope env(z), d env / dz
};

constexpr int kModeRec = 6;

__host__ __device__ __forceinline__ unsigned long long splitmix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform in [-1, 1): 53 random bits, exact arithmetic.
__host__ __device__ __forceinline__ double urand(unsigned long long seed, unsigned long long ctr)
{
    const unsigned long long h = splitmix64(seed ^ splitmix64(ctr));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

// feq(rho=1, u) with the reference's expression order (xcomputations.c:331).
__device__ __forceinline__ void feq_fp64(double u0, double u1, double u2, double (&f)[Q])
{
    const double v = u0 * u0 + u1 * u1 + u2 * u2;
    static_for<0, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        const double u = edot<a>(u0, u1, u2);
        if constexpr (a == 0)
            f[a] = (WD[a] * 1.0) * (1. - 1.5 * v);
        else
            f[a] = (WD[a] * 1.0) * (((1. + 3. * u) + (4.5 * u) * u) - 1.5 * v);
    });
}

// Seeded initial state, written as post-stream populations into cur (and
// the same values into other, as init.c fills both f and ftemp).
//  POISEUILLE: u_x = 4 uc y(H-y)/H^2, y = z + 1/2 (walls at 0 and H = Z),
//              plus amp*uc*U[-1,1) in each component (counter = node*4+comp).
//  TURBULENT : mean from the log law u+ = ln(y+)/kappa + B (viscous sublayer
//              u+ = y+ below y+ = 11), plus the curl of a random vector
//              potential built from Fourier modes |kx|, |ky| <= nmodes with a
//              sin^2 wall-normal envelope (divergence-free, zero at walls),
//              scaled so that its rms magnitude over the grid is amp times
//              the log law's centreline velocity.
template <typename T>
static __global__ __launch_bounds__(256) void init_kernel(InitArgs A)
{
    const int X = A.g.X, Y = A.g.Y, Z = A.g.Z;
    const long long total = (long long)X * Y * Z;
    T* cw = (T*)A.cur;
    T* ow = (T*)A.other;
    for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < total;
         n += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(n % Z);
        const long long cl = n / Z;
        const int y = (int)(cl % Y);
        const int x = (int)(cl / Y);
        double u0 = 0., u1 = 0., u2 = 0.;
        if (A.kind == 1) {
            const double yy = (double)z + 0.5;
            const double H = (double)Z;
            u0 = 4.0 * A.uc * yy * (H - yy) / (H * H);
            const unsigned long long node = (unsigned long long)n;
            u0 = u0 + A.amp * A.uc * urand(A.seed, node * 4ull + 0ull);
            u1 = A.amp * A.uc * urand(A.seed, node * 4ull + 1ull);
            u2 = A.amp * A.uc * urand(A.seed, node * 4ull + 2ull);
        } else if (A.kind == 2) {
            // u = U(z) e_x + nrm curl(psi), psi = sum over modes m of
            //   sc_m env(z) (a1 cos(ph1), a2 cos(ph2), 0),  ph = kx x + ky y + p,
            // so u'_x = -d psi_y/dz, u'_y = d psi_x/dz, u'_z = d psi_y/dx - d psi_x/dy.
            // The phases come from the x and y tables by the angle-sum rule;
            // only IEEE products and sums here (-ffp-contract=off), in this order.
            const double U = A.zt[z], env = A.zt[Z + z], denv = A.zt[2 * Z + z];
            double ux = 0., uy = 0., uz = 0.;
            for (int m = 0; m < A.nm; m++) {
                const double* md = A.mode + (size_t)m * kModeRec;
                const int i = (int)md[0];
                const double sc = md[1], a1 = md[2], a2 = md[3], kx = md[4], ky = md[5];
                const double cA = A.cx[(size_t)i * X + x], sA = A.sx[(size_t)i * X + x];
                const double* b = A.by + (size_t)m * 4 * Y;
                const double cb1 = b[y], sb1 = b[Y + y], cb2 = b[2 * Y + y], sb2 = b[3 * Y + y];
                const double c1 = cA * cb1 - sA * sb1, s1 = sA * cb1 + cA * sb1;
                const double c2 = cA * cb2 - sA * sb2, s2 = sA * cb2 + cA * sb2;
                ux = ux + sc * (-((denv * a2) * c2));
                uy = uy + sc * ((denv * a1) * c1);
                uz = uz + sc * (env * ((a1 * ky) * s1 - (a2 * kx) * s2));
            }
            u0 = U + A.nrm * ux;
            u1 = A.nrm * uy;
            u2 = A.nrm * uz;
        }
        double f[Q];
        if (A.kind == 0) {
#pragma unroll
            for (int a = 0; a < Q; a++) f[a] = WD[a];
        } else {
            feq_fp64(u0, u1, u2, f);
        }
        const long long cb = ((long long)x * Y + y) * Q * Z + z;
#pragma unroll
        for (int a = 0; a < Q; a++) {
            const T v = from_f<T>(f[a], a);
            cw[cb + (long long)a * Z] = v;
            ow[cb + (long long)a * Z] = v;
        }
    }
}

}  // namespace lbm
