// lbm_kernels.h -- the fused D3Q19 collide-stream step and its helpers
// for gfx950 (MI355X).
//
// Device layout ("column blocks"): for each (x, y) column the 19 direction
// segments of its Z fluid rows are contiguous,
//     pop[((x*Y + y)*19 + a)*Z + z],   z = 0..Z-1 the fluid rows,
// z (wall-normal) fastest as k is in the reference's Fb(...) macro
// (definitions.h:93); the wall rows are not stored.  A wave that updates a
// column therefore writes one contiguous block of 19*Z values and reads
// from the 9 column blocks of its 3x3 (x, y) neighbourhood, which keeps the
// number of DRAM pages open per wave small (measured on MI355X: fetched
// bytes within 3% of the algorithmic 19*sizeof(T) per node, against 6-27%
// excess for one plane per direction).
//
// Work decomposition of the step: one wavefront owns one column at a time
// and walks it in 64-row chunks, lane = row; every address term that
// depends on x and y (the nine upstream column blocks, the wall word) is
// wave-uniform and lives in SGPRs.  Blocks of W waves loop over columns
// (persistent-style grid sized by the autotuner).
//
// Streaming is "pull" (A-B): the kernel gathers the post-collision values
// of the previous step from the 19 upstream neighbours (periodic wrap in x
// and y, the one-rank halo of xcomputations.c:222-289,435-555), applies the
// wall rule of xcomputations.c:368-375 in pull form, collides and writes
// the node's 19 post-collision values.  The pull rule for the
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
constexpr int kMaxWaves = 8;  // waves per block
constexpr int kNQ = 4;        // per-plane accumulators: su sv sw sdrho
constexpr int kNG = 4;        // global accumulators: jx, mass-1, budget momentum, sq (K)
constexpr int kNAcc = kNQ + kNG;
// Tallest column: the finisher owns one plane per thread of its block.
constexpr int kMaxZ = 1024;

// A column taller than kMaxWaves chunks (512 rows) is split into z-pieces
// of at most kMaxWaves chunks: the step kernel gives each piece to its own
// group of blocks (each wave's plane sums then cover one piece), and the
// chunk-mapped kernels to the y index of their grid.  The step kernel may
// split a shorter column too (a launch-shape choice: smaller plane-sum
// slices, more waves resident); the Geom of the StepArgs it is launched with
// carries its piece count.
__host__ __device__ inline int z_pieces(int nzc) { return (nzc + kMaxWaves - 1) / kMaxWaves; }

struct Geom {
    int X, Y, Z;
    int nzc;          // 64-row chunks per column
    int nzp;          // z-pieces per column (z_pieces; the step kernel's launch: its piece count)
    int cpp;          // chunks per z-piece: ceil(nzc / nzp)
    int cpb;          // columns per block iteration (chunk-mapped kernels)
    int ncols;        // X*Y
    int ngroups;      // ceil(ncols / cpb)
    long long total;  // elements of one population array: ncols*19*Z
};

struct StepArgs {
    Geom g;
    const void* src;
    void* dst;
    const int8_t* mbot;     // [X*Y] bottom wall codes (-1 / -2)
    const int8_t* mtop;     // [X*Y] top wall codes
    const uint32_t* wword;  // [X*Y] per-column wall words
    const uint32_t* bands;  // [X*Y] solid bands hb | ht << 16 (null: flat walls only)
    const uint32_t* bnbr;   // [X*Y][16] the bands of each column's 3x3 neighbourhood (band_nbr_kernel)
    const uint32_t* cnbr;   // [X*Y][16] each column's 3x3 neighbour indices (column_nbr_kernel)
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
    static __device__ __forceinline__ float to_g(float v) { return v; }
    static __device__ __forceinline__ float from_g(float g) { return g; }
};
// FP16 storage of the shifted populations, scaled by 2^16 (exact power of
// two) so that deviations down to 1e-9 stay normal numbers while |g| < 1
// stays below the FP16 maximum; arithmetic in FP32.
constexpr float kHalfScale = 65536.0f;
constexpr float kHalfInv = 1.0f / 65536.0f;
template <>
struct Store<_Float16> {
    static __device__ __forceinline__ float to_g(_Float16 v) { return (float)v * kHalfInv; }
    static __device__ __forceinline__ _Float16 from_g(float g) { return (_Float16)(g * kHalfScale); }
    static __device__ __forceinline__ double to_f(_Float16 v, int a) { return WD[a] + (double)to_g(v); }
};

// Column geometry of one work item.
struct Column {
    int x, y, xm, xp, ym, yp;
    int Y, Z;
    long long cb;  // own column block: col*19*Z
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
    c.cb = (long long)col * Q * Z;
    return c;
}

// Wave-uniform column (col must be the same in every lane of the wave);
// readfirstlane keeps every derived address term in SGPRs.
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
    c.cb = (long long)col * Q * Z;
    return c;
}

// Block of the column (x - ex, y - ey) that direction (ex, ey) pulls from.
template <int ex, int ey>
__device__ __forceinline__ long long colbase(const Column& c)
{
    const int sx = ex == 1 ? c.xm : (ex == -1 ? c.xp : c.x);
    const int sy = ey == 1 ? c.ym : (ey == -1 ? c.yp : c.y);
    return (long long)(sx * c.Y + sy) * Q * c.Z;
}

// A column's neighbourhood, one 64-byte record read with one scalar load:
// w[3*i + j] = index of column (x - ex, y - ey) for i = ex + 1, j = ey + 1
// (periodic wrap in x and y), precomputed once per grid (column_nbr_kernel).
struct alignas(64) ColNbr {
    uint32_t w[16];
};

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

// Bytes of global memory (HBM).  The field pointers are held in this
// address space: a pointer loaded from a table (the ensemble's member
// records) is otherwise generic, and every access through a generic pointer
// is a flat instruction with a per-lane 64-bit address.
typedef __attribute__((address_space(1))) char gbyte;
#define GLOBAL_AS __attribute__((address_space(1)))

// Wave-uniform load through the constant address space (scalar loads) of
// a table entry that no kernel running at the same time writes.
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

// A column's band neighbourhood, one 64-byte record read with one scalar
// load: w[3*i + j] = bands of column (x - ex, y - ey) for i = ex + 1,
// j = ey + 1; w[9] = max hb | max ht << 16 over the nine.
struct alignas(64) BandNbr {
    uint32_t w[16];
};

// Solid bands inside the fluid rows (riblets): column (x, y) is solid for
// rows z < hb and z >= Z - ht, packed as hb | ht << 16.  No reference
// counterpart (the reference has solids only in the wall rows); the rule
// is halfway bounce-back whenever the upstream node is solid, which is the
// reference's wall rule when the solid is a no-slip wall row.
__device__ __forceinline__ bool band_solid(uint32_t bw, int zs, int Z)
{
    const int hb = (int)(bw & 0xFFFFu), ht = (int)(bw >> 16);
    return (zs >= 0 && zs < hb) || (zs < Z && zs >= Z - ht);
}

// Row zs (-1 <= zs <= Z) of a column with band word bw is not fluid: a band
// row, or one of the wall rows -1 and Z (zs < hb holds for zs = -1 and
// zs >= Z - ht for zs = Z, whatever the band heights).
__device__ __forceinline__ bool solid_or_wall(uint32_t bw, int zs, int Z)
{
    const int hb = (int)(bw & 0xFFFFu), ht = (int)(bw >> 16);
    return zs < hb || zs >= Z - ht;
}

// Gather the post-stream populations of node (column c, row z) from the
// post-collision array src (pull), wall rule included.  Generic per-thread
// form used by the cadence kernels.
template <typename T>
__device__ __forceinline__ void pull19(const T* __restrict__ src, const Column& c, int z,
                                       const int8_t* __restrict__ mbot, const int8_t* __restrict__ mtop,
                                       const uint32_t* __restrict__ bands, T (&v)[Q])
{
    const int Z = c.Z, Y = c.Y;
    static_for<0, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        int zs = z - EZ[a];
        zs = zs < 0 ? 0 : (zs >= Z ? Z - 1 : zs);
        v[a] = src[colbase<EX[a], EY[a]>(c) + (long long)a * Z + zs];
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
                    v[a] = src[colbase<EX[a], EY[a]>(c) + (long long)MIR[a] * Z + z];
                else
                    v[a] = src[c.cb + (long long)OPP[a] * Z + z];
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
                    v[a] = src[colbase<EX[a], EY[a]>(c) + (long long)MIR[a] * Z + z];
                else
                    v[a] = src[c.cb + (long long)OPP[a] * Z + z];
            }
        });
    }
    if (bands) {
        static_for<1, Q>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            const int sx = EX[a] == 1 ? c.xm : (EX[a] == -1 ? c.xp : c.x);
            const int sy = EY[a] == 1 ? c.ym : (EY[a] == -1 ? c.yp : c.y);
            if (band_solid(bands[sx * Y + sy], z - EZ[a], Z)) v[a] = src[c.cb + (long long)OPP[a] * Z + z];
        });
    }
}

template <typename T>
__device__ __forceinline__ void load_local19(const T* __restrict__ src, const Column& c, int z, T (&v)[Q])
{
    static_for<0, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        v[a] = src[c.cb + (long long)a * c.Z + z];
    });
}

// Hot-path gather for one wave: lane z of the chunk.  Every source address
// is a wave-uniform base (SGPRs: upstream column block + direction segment)
// plus the lane's upstream row offset (row z - e_z).  The wall-adjacent lanes, whose
// upstream row is in the wall, read a clamped in-segment row (so no extra
// cache line is fetched) and are then overwritten by the wall rule: uniform
// decisions from the column's wall word, one active lane.
template <typename T, bool FIRST>
__device__ __forceinline__ void gather_wave(const gbyte* const (&sp)[3][3], int zb, uint32_t zoff, uint32_t zoff_up,
                                            uint32_t zoff_dn, bool bot_lane, bool top_lane, uint32_t ww,
                                            T (&v)[Q])
{
    // sp[ex + 1][ey + 1]: byte address of the block of column (x - ex, y - ey)
    static_for<0, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        const gbyte* p;
        uint32_t off;
        if constexpr (FIRST) {
            p = sp[1][1] + a * zb;
            off = zoff;
        } else {
            p = sp[EX[a] + 1][EY[a] + 1] + a * zb;
            off = EZ[a] == 1 ? zoff_up : (EZ[a] == -1 ? zoff_dn : zoff);
        }
        v[a] = *(const GLOBAL_AS T*)(p + off);
    });
    if constexpr (!FIRST) {
        if (bot_lane) {
            static_for<0, Q>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                if constexpr (EZ[a] == 1) {
                    // row 0: a wave-uniform address
                    const bool spec = (ww >> wall_bit<a>()) & 1u;
                    const gbyte* p = spec ? sp[EX[a] + 1][EY[a] + 1] + MIR[a] * zb : sp[1][1] + OPP[a] * zb;
                    v[a] = *(const GLOBAL_AS T*)p;
                }
            });
        }
        if (top_lane) {
            static_for<0, Q>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                if constexpr (EZ[a] == -1) {
                    // row Z - 1, the segment's last: a wave-uniform address
                    const bool spec = (ww >> wall_bit<a>()) & 1u;
                    const gbyte* p = spec ? sp[EX[a] + 1][EY[a] + 1] + MIR[a] * zb : sp[1][1] + OPP[a] * zb;
                    v[a] = *(const GLOBAL_AS T*)(p + (zb - (int)sizeof(T)));
                }
            });
        }
    }
}

// gather_wave for a chunk next to a solid band, bounce-back on the read side:
// a node whose upstream node x - e_a is solid takes its own post-collision
// opp(a) of the previous step (the population it pushed into the solid) from
// its own slot.  One load per direction, its address chosen per lane (the
// upstream slot or the node's own slot of opp(a)), instead of the pull and a
// masked reload: measured 0.5% faster at 1536x768x384 with riblets (FP32).
// Upstream wall rows are the wall rule's (band_solid is false there).
template <typename T>
__device__ __forceinline__ void gather_wave_read_side(const gbyte* const (&sp)[3][3], int zb, uint32_t zoff,
                                                      uint32_t zoff_up, uint32_t zoff_dn, bool bot_lane,
                                                      bool top_lane, uint32_t ww, const uint32_t (&bw)[3][3], int z,
                                                      int Z, T (&v)[Q])
{
    static_for<0, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        const gbyte* up =
            sp[EX[a] + 1][EY[a] + 1] + a * zb + (EZ[a] == 1 ? zoff_up : (EZ[a] == -1 ? zoff_dn : zoff));
        if constexpr (a == 0) {
            v[a] = *(const GLOBAL_AS T*)up;
        } else {
            const gbyte* own = sp[1][1] + OPP[a] * zb + zoff;
            v[a] = *(const GLOBAL_AS T*)(band_solid(bw[1 + EX[a]][1 + EY[a]], z - EZ[a], Z) ? own : up);
        }
    });
    if (bot_lane) {
        static_for<0, Q>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            if constexpr (EZ[a] == 1) {
                const bool spec = (ww >> wall_bit<a>()) & 1u;
                const gbyte* p = spec ? sp[EX[a] + 1][EY[a] + 1] + MIR[a] * zb : sp[1][1] + OPP[a] * zb;
                v[a] = *(const GLOBAL_AS T*)p;
            }
        });
    }
    if (top_lane) {
        static_for<0, Q>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            if constexpr (EZ[a] == -1) {
                const bool spec = (ww >> wall_bit<a>()) & 1u;
                const gbyte* p = spec ? sp[EX[a] + 1][EY[a] + 1] + MIR[a] * zb : sp[1][1] + OPP[a] * zb;
                v[a] = *(const GLOBAL_AS T*)(p + (zb - (int)sizeof(T)));
            }
        });
    }
}

// The statistics variant keeps the [3][Z] plane-mean shifts in LDS up to
// Z = 256 and reads them from global memory above: the LDS bounds the blocks
// resident per CU, and at Z = 384 one more block per CU is worth more than
// the faster reads (+2% FP32, +3-8% FP16 at Z = 384; at Z = 192 the LDS
// copy is 3-5% faster).
__host__ __device__ inline bool step_shifts_in_lds(int Z) { return Z <= 256; }

// Planes of one wave's plane sums in the step kernel run with nzp z-pieces:
// the column, or one piece of it (whole chunks).
__host__ __device__ inline int step_plane_rows(int Z, int nzp)
{
    const int nzc = (Z + kWave - 1) / kWave;
    return nzp == 1 ? Z : (nzc + nzp - 1) / nzp * kWave;
}

// Dynamic LDS of the step kernel: [W][kNG] per-wave global sums, then with
// STATS [W][kNQ][PZ] per-wave plane sums (PZ = step_plane_rows) and (see
// above) the [3][Z] shifts.  At Z = 192, 4 waves: 29 KB, five blocks per CU.
__host__ __device__ inline size_t step_lds_bytes(int Z, int nzp, int W, bool stats)
{
    return sizeof(double) * ((size_t)W * kNG + (stats ? (size_t)W * kNQ * step_plane_rows(Z, nzp) +
                                                             (step_shifts_in_lds(Z) ? 3 * (size_t)Z : 0)
                                                       : 0));
}

// Waves per block of the step kernel for a given Z and z-piece count: at
// most 4, and the statistics variant's LDS within 64 KiB.
__host__ __device__ inline int step_waves(int Z, int nzp)
{
    int w = 4;
    while (w > 1 && step_lds_bytes(Z, nzp, w, true) > 65536) w--;
    return w;
}

// FP32 collision of stored shifted populations, in place.  FP16 storage
// converts in and out (the stored, rounded values are the next state).
__device__ __forceinline__ void collide_stored(float (&v)[Q], float omega, float G, float& drho, float& u0,
                                               float& u1, float& u2, float& jx)
{
    collide_fp32(v, omega, G, drho, u0, u1, u2, jx);
}

__device__ __forceinline__ void collide_stored(_Float16 (&v)[Q], float omega, float G, float& drho, float& u0,
                                               float& u1, float& u2, float& jx)
{
    // collide the stored values g 2^16 directly (the collision's weights
    // carry the 2^16)
    float g[Q];
#pragma unroll
    for (int a = 0; a < Q; a++) g[a] = (float)v[a];
    collide_fp32<16>(g, omega, G, drho, u0, u1, u2, jx);
#pragma unroll
    for (int a = 0; a < Q; a++) v[a] = (_Float16)g[a];
}

// Register accumulators of one wave over the columns it updates.
struct StepAcc {
    double jx = 0.0, mass = 0.0, mom = 0.0, sq = 0.0;
};

// The fused step of one column (all Z rows in 64-row chunks, one wave):
// pull-stream + wall rule + BGK collision with force + per-plane
// statistics + momentum budget of the next state.
//   FIRST: the source holds post-stream populations (after a load or an
//          init), so the gather is local instead of a pull.
//   STATS: accumulate the per-plane sums for the per-step logs (in each
//          wave's private LDS slice, so the sums are deterministic).
// Budget of the next state (the force of the next step, encalc.c:121):
// mass and x-momentum are conserved by the collision except for the force
// term, whose x-moment is exactly G, and streaming is a permutation except
// that a halfway bounce-back reverses e_x:
//     sum f_next = sum f,  sum e_x f_next = sum (j_x + G) - 2 sum_{BB} e_x g
//   Chunks c0 <= ch < c1 of the column (the whole column, or one z-piece);
//   mine holds planes zbase .. zbase + pz - 1 as [kNQ][pz].
template <typename T, bool FIRST, bool STATS, bool BANDS>
__device__ __forceinline__ void step_column(const StepArgs& A, int col, int lane, double G, const double* shl,
                                            double* mine, StepAcc& acc, int c0, int c1, int pz, int zbase)
{
    const int Z = A.g.Z;
    const int ctop = (Z - 1) / kWave, ltop = (Z - 1) % kWave;
    // Band surfaces: FP32 and FP64 storage bounce on the read side (the
    // node's own slot loaded in place of the upstream slot, plain stores);
    // FP16 on the write side (a conditional store into the solid
    // neighbour's slot).  Both give the same values.  Measured with riblets
    // in one allocation each (the two rules switched between timed
    // launches): the read side is faster in FP32 (10% at 1536x768x384, 5%
    // at 768x384x192, 7% in a mixed ensemble) and in FP64 (2.9%, 3.1% and
    // 8.7% at 1536x768x384, 768x384x192 and 384x192x96), slower in FP16
    // (4%).  A third rule (a fluid node and its solid neighbour storing
    // into each other's slot: one store per direction) was 4-10% slower
    // than both: store instructions whose lanes split between two places
    // cost more than the extra whole-line stores.
    constexpr bool kReadSide = !std::is_same<T, _Float16>::value;
    // Every store of the step is non-temporal: no kernel of the step reads
    // the destination back, so its lines are not kept in the L2, which then
    // holds the pulled lines a chunk shares with the next one (each e_z = +-1
    // segment's line across a 64-row chunk boundary).  Measured: 3-5% in every
    // precision at 768x384x192 and 1536x768x384 (riblets), and larger grids
    // stop losing to refetches.
    const gbyte* __restrict__ s8 = (const gbyte*)A.src;
    gbyte* __restrict__ d8 = (gbyte*)A.dst;
    // byte addresses of the 3x3 neighbourhood's column blocks in the source
    // and the destination, [ex + 1][ey + 1] = column (x - ex, y - ey)
    const gbyte* sp[3][3];
    gbyte* dp[3][3];
    {
        const ColNbr cr = load_uniform(reinterpret_cast<const ColNbr*>(A.cnbr) + col);
        const uint32_t qzb = (uint32_t)(Q * Z * (int)sizeof(T));  // bytes of a column block (< 2^17)
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const unsigned long long o = (unsigned long long)cr.w[3 * i + j] * qzb;
                sp[i][j] = s8 + o;
                dp[i][j] = d8 + o;
            }
    }
    const int zb = Z * (int)sizeof(T);
    const uint32_t ww = load_uniform_u32(A.wword + col);
    // solid bands of the 3x3 neighbourhood, bw[ex+1][ey+1] = column (x - ex, y - ey)
    uint32_t bw[3][3];
    int maxhb = 0, maxht = 0;
    if constexpr (BANDS) {
        const BandNbr nb = load_uniform(reinterpret_cast<const BandNbr*>(A.bnbr) + col);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) bw[i][j] = nb.w[3 * i + j];
        maxhb = (int)(nb.w[9] & 0xFFFFu);
        maxht = (int)(nb.w[9] >> 16);
    }
    for (int ch = c0; ch < c1; ch++) {
        const int z = ch * kWave + lane;
        if (z >= Z) break;
        const int zm = z - zbase;  // plane slot in mine
        const bool bot_lane = ch == 0 && lane == 0;
        const bool top_lane = ch == ctop && lane == ltop;
        const uint32_t zoff = (uint32_t)z * (uint32_t)sizeof(T);
        // upstream rows of the e_z = +1 and -1 populations (z - 1, z + 1),
        // clamped into the column at the walls
        const uint32_t zoff_up = (uint32_t)(z == 0 ? 0 : z - 1) * (uint32_t)sizeof(T);
        const uint32_t zoff_dn = (uint32_t)(z == Z - 1 ? Z - 1 : z + 1) * (uint32_t)sizeof(T);
        // zb (bytes of a direction segment) is made opaque per chunk, so the
        // 19 + 19 per-direction addresses are formed in each chunk from the
        // nine column bases instead of being held (and spilled) across it
        int zbc = zb;
        asm volatile("" : "+s"(zbc));
        T v[Q];
        // chunk touches a band of the neighbourhood (upstream rows z-1..z+1)
        bool slow = false, solid_own = false;
        if constexpr (BANDS) {
            const int z0 = ch * kWave, z1 = min(z0 + kWave, Z);
            slow = (maxhb > 0 && z0 - 1 < maxhb) || (maxht > 0 && z1 >= Z - maxht);
            solid_own = band_solid(bw[1][1], z, Z);
        }
        if constexpr (BANDS && !FIRST && kReadSide) {
            if (slow)
                gather_wave_read_side<T>(sp, zbc, zoff, zoff_up, zoff_dn, bot_lane, top_lane, ww, bw, z, Z, v);
            else
                gather_wave<T, FIRST>(sp, zbc, zoff, zoff_up, zoff_dn, bot_lane, top_lane, ww, v);
        } else {
            gather_wave<T, FIRST>(sp, zbc, zoff, zoff_up, zoff_dn, bot_lane, top_lane, ww, v);
        }

        double mass_node, jx_raw;
        if constexpr (std::is_same<T, double>::value) {
            double ux, uy, uz, dens;
            moments_fp64(v, ux, uy, uz, dens);
            const double u0 = ux / dens, u1 = uy / dens, u2 = uz / dens;
            collide_fp64(v, u0, u1, u2, dens, A.omega, G);
            jx_raw = ux;
            mass_node = dens - 1.0;
            if constexpr (STATS) {
                if (!solid_own) {
                    const double du = u0 - shl[z], dv = u1 - shl[Z + z], dw = u2 - shl[2 * Z + z];
                    mine[0 * pz + zm] += du;
                    mine[1 * pz + zm] += dv;
                    mine[2 * pz + zm] += dw;
                    mine[3 * pz + zm] += mass_node;
                    acc.sq += (du * du + dv * dv) + dw * dw;
                }
            }
        } else {
            float drho, u0, u1, u2, jx;
            collide_stored(v, (float)A.omega, (float)G, drho, u0, u1, u2, jx);
            jx_raw = (double)jx;
            mass_node = (double)drho;
            if constexpr (STATS) {
                if (!solid_own) {
                    const double du = (double)u0 - shl[z], dv = (double)u1 - shl[Z + z],
                                 dw = (double)u2 - shl[2 * Z + z];
                    mine[0 * pz + zm] += du;
                    mine[1 * pz + zm] += dv;
                    mine[2 * pz + zm] += dw;
                    mine[3 * pz + zm] += mass_node;
                    acc.sq += (du * du + dv * dv) + dw * dw;
                }
            }
        }
        // post-collision f of direction a in FP64 (budget terms), formed
        // only where one is needed
        auto fpost = [&](auto ac) -> double {
            constexpr int a = decltype(ac)::value;
            if constexpr (std::is_same<T, double>::value)
                return v[a];
            else
                return WD[a] + (double)Store<T>::to_g(v[a]);
        };
        using I11 = std::integral_constant<int, 11>;
        using I12 = std::integral_constant<int, 12>;
        using I13 = std::integral_constant<int, 13>;
        using I14 = std::integral_constant<int, 14>;
        double mom_node = jx_raw + G;
        // bounce-back reverses e_x of the populations pushed into a
        // no-slip wall node from a no-slip wall column (xcomputations.c:370)
        if (bot_lane)
            mom_node = mom_node - (((ww >> 10) & 1u) ? 2.0 * fpost(I12{}) : 0.0) +
                       (((ww >> 11) & 1u) ? 2.0 * fpost(I13{}) : 0.0);
        if (top_lane)
            mom_node = mom_node - (((ww >> 12) & 1u) ? 2.0 * fpost(I11{}) : 0.0) +
                       (((ww >> 13) & 1u) ? 2.0 * fpost(I14{}) : 0.0);
        if (BANDS && !kReadSide && slow) {
            // Halfway bounce-back at band surfaces, on the write side: the
            // population a fluid node x pushes toward a solid node t = x + e_a
            // returns to x as opp(a) at the next step, so x stores it in t's
            // slot of direction opp(a), the slot x pulls from.  Every other
            // slot of t is stored by t itself (values no fluid node reads):
            // each slot is written once per step and every cache line of
            // the destination is written whole.  Each such push also reverses
            // e_x of that population in the budget (direction order).
            static_for<0, Q>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                const int zt = z + EZ[a];
                // row zt exists (only the e_z != 0 directions can leave the column)
                const bool inside = EZ[a] == 0 ? true : (EZ[a] > 0 ? zt < Z : zt >= 0);
                // x + e_a is not a fluid node (a band node or a wall row)
                const bool nb_solid = a == 0 ? solid_own : solid_or_wall(bw[1 - EX[a]][1 - EY[a]], zt, Z);
                if (!solid_own || nb_solid) __builtin_nontemporal_store(v[a], (GLOBAL_AS T*)(dp[1][1] + a * zbc + zoff));
                if (a != 0 && !solid_own && nb_solid && inside) {
                    __builtin_nontemporal_store(v[a], (GLOBAL_AS T*)(dp[1 - EX[a]][1 - EY[a]] + (OPP[a] * zbc + EZ[a] * (int)sizeof(T)) + zoff));
                    if constexpr (EX[a] != 0) mom_node -= 2.0 * (double)EX[a] * fpost(ac);
                }
            });
        } else {
            if constexpr (BANDS && kReadSide) {
                // read side (above): every node stores all 19 of its own
                // slots (no fluid node pulls a solid node's values); a push
                // of a fluid node toward a solid node inside the rows
                // returns to it as opp(a) and reverses e_x in the budget
                if (slow) {
                    static_for<1, Q>([&](auto ac) {
                        constexpr int a = decltype(ac)::value;
                        if constexpr (EX[a] != 0) {
                            if (!solid_own && band_solid(bw[1 - EX[a]][1 - EY[a]], z + EZ[a], Z))
                                mom_node -= 2.0 * (double)EX[a] * fpost(ac);
                        }
                    });
                }
            }
            // the row offset as a 32-bit value of this block, so the stores
            // take the (SGPR base + 32-bit VGPR offset) form
            uint32_t zo = zoff;
            asm volatile("" : "+v"(zo));
            static_for<0, Q>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                __builtin_nontemporal_store(v[a], (GLOBAL_AS T*)(dp[1][1] + a * zbc + zo));
            });
        }
        if (!solid_own) {
            if constexpr (STATS) acc.jx += jx_raw;
            acc.mass += mass_node;
            acc.mom += mom_node;
        }
    }
}

// Two rows per lane (FP16 storage, even Z): the wave walks its column in
// 128-row chunks, lane l holding rows z = 128 ch + 2 l (low half) and z + 1
// (high half) of every direction in one dword, so every load and store
// moves 4 bytes per lane (256 bytes per wave instruction), the bytes an
// FP32 access moves, where the one-row form moves 2.  The pulls of the
// e_z = +1 and -1 directions are the dwords at rows (z - 1, z) and
// (z + 1, z + 2): 2-byte aligned, a form the amdhsa target's unaligned
// access mode serves as one dword load.  At row 0 the e_z = +1 pull reads
// the last row of the previous direction segment of the same column block
// (directions 5, 11, 14, 15, 18 all follow another segment) and at row
// Z - 1 the e_z = -1 pull the first row of the next one (directions 6, 12,
// 13, 16, 17 all precede another): in bounds, and replaced by the wall
// rule.  Bands bounce on the read side (the values the write side gives to
// every fluid node).  Each node's arithmetic is the one-row form's.
typedef uint32_t __attribute__((aligned(2))) u32_a2;

__device__ __forceinline__ _Float16 half_of(uint32_t w, int hi)
{
    return __builtin_bit_cast(_Float16, (uint16_t)(hi ? w >> 16 : w & 0xFFFFu));
}

__device__ __forceinline__ uint32_t with_half(uint32_t w, int hi, _Float16 h)
{
    const uint32_t b = (uint32_t)__builtin_bit_cast(uint16_t, h);
    return hi ? (w & 0xFFFFu) | (b << 16) : (w & 0xFFFF0000u) | b;
}

__device__ __forceinline__ uint32_t with_half16(uint32_t w, int hi, uint16_t b)
{
    return hi ? (w & 0xFFFFu) | ((uint32_t)b << 16) : (w & 0xFFFF0000u) | (uint32_t)b;
}

template <bool STATS, bool BANDS>
__device__ __forceinline__ void step_column_r2(const StepArgs& A, int col, int lane, double G, const double* shl,
                                               double* mine, StepAcc& acc)
{
    using T = _Float16;
    const int Z = A.g.Z;
    const int nzc2 = (Z + 2 * kWave - 1) / (2 * kWave);
    // the lane whose high half is row Z - 1
    const int ctop = (Z - 2) / (2 * kWave), ltop = ((Z - 2) % (2 * kWave)) / 2;
    const gbyte* __restrict__ s8 = (const gbyte*)A.src;
    gbyte* __restrict__ d8 = (gbyte*)A.dst;
    const gbyte* sp[3][3];
    gbyte* dst;
    {
        const ColNbr cr = load_uniform(reinterpret_cast<const ColNbr*>(A.cnbr) + col);
        const uint32_t qzb = (uint32_t)(Q * Z * (int)sizeof(T));
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) sp[i][j] = s8 + (unsigned long long)cr.w[3 * i + j] * qzb;
        dst = d8 + (unsigned long long)cr.w[4] * qzb;
    }
    const int zb = Z * (int)sizeof(T);
    const uint32_t ww = load_uniform_u32(A.wword + col);
    uint32_t bw[3][3];
    int maxhb = 0, maxht = 0;
    if constexpr (BANDS) {
        const BandNbr nb = load_uniform(reinterpret_cast<const BandNbr*>(A.bnbr) + col);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) bw[i][j] = nb.w[3 * i + j];
        maxhb = (int)(nb.w[9] & 0xFFFFu);
        maxht = (int)(nb.w[9] >> 16);
    }
    for (int ch = 0; ch < nzc2; ch++) {
        const int z = ch * 2 * kWave + 2 * lane;
        if (z >= Z) break;
        const bool bot_lane = ch == 0 && lane == 0;
        const bool top_lane = ch == ctop && lane == ltop;
        const int zoff = z * (int)sizeof(T);
        int zbc = zb;
        asm volatile("" : "+s"(zbc));
        uint32_t v[Q];
        static_for<0, Q>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            const gbyte* p = sp[EX[a] + 1][EY[a] + 1] + a * zbc;
            if constexpr (EZ[a] == 0)
                v[a] = *(const GLOBAL_AS uint32_t*)(p + zoff);
            else
                v[a] = *(const GLOBAL_AS u32_a2*)(p + (zoff - EZ[a] * (int)sizeof(T)));
        });
        if (bot_lane) {
            static_for<0, Q>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                if constexpr (EZ[a] == 1) {
                    const bool spec = (ww >> wall_bit<a>()) & 1u;
                    const gbyte* p = spec ? sp[EX[a] + 1][EY[a] + 1] + MIR[a] * zbc : sp[1][1] + OPP[a] * zbc;
                    v[a] = with_half16(v[a], 0, *(const GLOBAL_AS uint16_t*)p);
                }
            });
        }
        if (top_lane) {
            static_for<0, Q>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                if constexpr (EZ[a] == -1) {
                    const bool spec = (ww >> wall_bit<a>()) & 1u;
                    const gbyte* p = spec ? sp[EX[a] + 1][EY[a] + 1] + MIR[a] * zbc : sp[1][1] + OPP[a] * zbc;
                    v[a] = with_half16(v[a], 1, *(const GLOBAL_AS uint16_t*)(p + (zbc - (int)sizeof(T))));
                }
            });
        }
        bool slow = false, solid0 = false, solid1 = false;
        if constexpr (BANDS) {
            const int z0 = ch * 2 * kWave, z1 = min(z0 + 2 * kWave, Z);
            slow = (maxhb > 0 && z0 - 1 < maxhb) || (maxht > 0 && z1 >= Z - maxht);
            solid0 = band_solid(bw[1][1], z, Z);
            solid1 = band_solid(bw[1][1], z + 1, Z);
            if (slow) {
                // read-side halfway bounce-back per node (see step_column)
                static_for<1, Q>([&](auto ac) {
                    constexpr int a = decltype(ac)::value;
                    const uint32_t bn = bw[1 + EX[a]][1 + EY[a]];
                    const bool c0 = band_solid(bn, z - EZ[a], Z), c1 = band_solid(bn, z + 1 - EZ[a], Z);
                    if (c0 || c1) {
                        const uint32_t own = *(const GLOBAL_AS uint32_t*)(sp[1][1] + OPP[a] * zbc + zoff);
                        const uint32_t m = (c0 ? 0xFFFFu : 0u) | (c1 ? 0xFFFF0000u : 0u);
                        v[a] = (v[a] & ~m) | (own & m);
                    }
                });
            }
        }
        // the two nodes of the lane, row z then row z + 1 (one at a time:
        // unrolled, the two collisions interleave and double the registers)
#pragma unroll 1
        for (int n = 0; n < 2; n++) {
            const int zn = z + n;
            const bool solid_own = n ? solid1 : solid0;
            T hv[Q];
#pragma unroll
            for (int a = 0; a < Q; a++) hv[a] = half_of(v[a], n);
            float drho, u0, u1, u2, jx;
            collide_stored(hv, (float)A.omega, (float)G, drho, u0, u1, u2, jx);
#pragma unroll
            for (int a = 0; a < Q; a++) v[a] = with_half(v[a], n, hv[a]);
            const double jx_raw = (double)jx, mass_node = (double)drho;
            if constexpr (STATS) {
                if (!solid_own) {
                    const double du = (double)u0 - shl[zn], dv = (double)u1 - shl[Z + zn],
                                 dw = (double)u2 - shl[2 * Z + zn];
                    mine[0 * Z + zn] += du;
                    mine[1 * Z + zn] += dv;
                    mine[2 * Z + zn] += dw;
                    mine[3 * Z + zn] += mass_node;
                    acc.sq += (du * du + dv * dv) + dw * dw;
                }
            }
            auto fpost = [&](auto ac) -> double {
                constexpr int a = decltype(ac)::value;
                return WD[a] + (double)Store<T>::to_g(hv[a]);
            };
            using I11 = std::integral_constant<int, 11>;
            using I12 = std::integral_constant<int, 12>;
            using I13 = std::integral_constant<int, 13>;
            using I14 = std::integral_constant<int, 14>;
            double mom_node = jx_raw + G;
            if (n == 0 && bot_lane)
                mom_node = mom_node - (((ww >> 10) & 1u) ? 2.0 * fpost(I12{}) : 0.0) +
                           (((ww >> 11) & 1u) ? 2.0 * fpost(I13{}) : 0.0);
            if (n == 1 && top_lane)
                mom_node = mom_node - (((ww >> 12) & 1u) ? 2.0 * fpost(I11{}) : 0.0) +
                           (((ww >> 13) & 1u) ? 2.0 * fpost(I14{}) : 0.0);
            if constexpr (BANDS) {
                if (slow) {
                    static_for<1, Q>([&](auto ac) {
                        constexpr int a = decltype(ac)::value;
                        if constexpr (EX[a] != 0) {
                            if (!solid_own && band_solid(bw[1 - EX[a]][1 - EY[a]], zn + EZ[a], Z))
                                mom_node -= 2.0 * (double)EX[a] * fpost(ac);
                        }
                    });
                }
            }
            if (!solid_own) {
                if constexpr (STATS) acc.jx += jx_raw;
                acc.mass += mass_node;
                acc.mom += mom_node;
            }
        }
        uint32_t zo = (uint32_t)zoff;
        asm volatile("" : "+v"(zo));
        static_for<0, Q>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            __builtin_nontemporal_store(v[a], (GLOBAL_AS uint32_t*)(dst + a * zbc + zo));
        });
    }
}

// A state written by a read-side band step leaves the slots of solid band
// nodes holding the solid nodes' own values; a write-side step pulls, at a
// fluid node x whose upstream node t = x - e_b is solid, t's slot b, which
// a write-side step fills with x's post-collision opp(b).  This pass gives
// every such slot that value (x's own slot opp(b), which both sides store):
// after it the state serves either side.  Each solid slot has one writer
// (t's slot b is pushed only from t + e_b) and no fluid slot is written,
// so the pass has no read-write overlap.  One thread per fluid node next to
// a band.
template <typename T>
static __global__ __launch_bounds__(256) void band_push_kernel(T* __restrict__ st, const uint32_t* __restrict__ bands,
                                                               int X, int Y, int Z)
{
    const long long total = (long long)X * Y * Z;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % Z);
        const int col = (int)(i / Z);
        const Column c = make_column_lane(col, X, Y, Z);
        if (band_solid(bands[col], z, Z)) continue;
        static_for<1, Q>([&](auto ac) {
            constexpr int a = decltype(ac)::value;
            // t = x + e_a, the node x pushes a toward
            const int tx = EX[a] == 1 ? c.xp : (EX[a] == -1 ? c.xm : c.x);
            const int ty = EY[a] == 1 ? c.yp : (EY[a] == -1 ? c.ym : c.y);
            const int tz = z + EZ[a];
            if (tz >= 0 && tz < Z && band_solid(bands[tx * Y + ty], tz, Z))
                st[((long long)(tx * Y + ty) * Q + OPP[a]) * Z + tz] = st[c.cb + (long long)a * Z + z];
        });
    }
}

// Deterministic reductions of one block: plane sums over the block's waves
// in wave order, global sums by a fixed butterfly over each wave's lanes and
// then over the waves in order.  The waves' plane sums cover planes zbase ..
// zbase + pz - 1 ([W][kNQ][pz]); they are written as partial number pslot
// of [kNQ][Z][npslot], the global sums as partial number gslot of
// [kNG][ngslot].
template <bool STATS>
__device__ __forceinline__ void step_block_reduce(double* red, const double* planes, const StepAcc& acc, int Z,
                                                  int pz, int zbase, int W, int wave, int lane, double* part,
                                                  double* gpart, int npslot, int pslot, int ngslot, int gslot)
{
    // each wave's global sums over its lanes (fixed butterfly), then over the
    // block's waves in wave order
    double g[kNG] = {acc.jx, acc.mass, acc.mom, acc.sq};
#pragma unroll
    for (int q = 0; q < kNG; q++) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) g[q] += __shfl_xor(g[q], off, 64);
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < kNG; q++) red[wave * kNG + q] = g[q];
    }
    __syncthreads();
    if constexpr (STATS) {
        for (int i = threadIdx.x; i < kNQ * pz; i += blockDim.x) {
            const int q = i / pz, z = zbase + (i - q * pz);
            if (z >= Z) continue;
            double s = 0.0;
            for (int w = 0; w < W; w++) s += planes[(size_t)w * kNQ * pz + i];
            part[((long long)q * Z + z) * npslot + pslot] = s;
        }
    }
    if (threadIdx.x < kNG) {
        const int q = threadIdx.x;
        if (STATS || (q != 0 && q != 3)) {
            double s = 0.0;
            for (int w = 0; w < W; w++) s += red[w * kNG + q];
            gpart[(long long)q * ngslot + gslot] = s;
        }
    }
}

// The single-case step: the block's waves take columns round-robin over
// the grid; one set of partials per block.
template <typename T, bool FIRST, bool STATS, bool BANDS, int R>
__device__ __forceinline__ void step_column_of(const StepArgs& A, int col, int lane, double G, const double* shl,
                                               double* mine, StepAcc& acc, int c0, int c1, int pz, int zbase)
{
    if constexpr (R == 2) {
        static_assert(std::is_same<T, _Float16>::value && !FIRST, "two rows per lane: FP16 pull steps only");
        step_column_r2<STATS, BANDS>(A, col, lane, G, shl, mine, acc);  // one piece (Z <= 512)
    } else {
        step_column<T, FIRST, STATS, BANDS>(A, col, lane, G, shl, mine, acc, c0, c1, pz, zbase);
    }
}

// R: rows per lane (1, or 2 for FP16 storage with an even Z).
// PIECES: columns split into A.g.nzp z-pieces (always above 512 rows; at
// 321-384 and 449-512 rows by launch choice), gridDim.x a multiple of nzp:
// block b takes piece b % nzp of the columns of its group b / nzp; without
// it the whole column, with compile-time piece bounds.
template <typename T, bool FIRST, bool STATS, bool BANDS, int R = 1, bool PIECES = false>
static __global__ __launch_bounds__(kWave* kMaxWaves) void step_kernel(StepArgs A)
{
    static_assert(!PIECES || R == 1, "two rows per lane: one z-piece");
    extern __shared__ double lds[];  // see step_lds_bytes
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int W = blockDim.x / kWave;
    const int Z = A.g.Z;
    const int P = PIECES ? A.g.nzp : 1;
    const int piece = PIECES ? (int)blockIdx.x % P : 0;
    const int bgrp = PIECES ? (int)blockIdx.x / P : (int)blockIdx.x;
    const int ngrp = PIECES ? (int)gridDim.x / P : (int)gridDim.x;
    const int c0 = PIECES ? piece * A.g.cpp : 0;
    const int c1 = PIECES ? min(A.g.nzc, c0 + A.g.cpp) : A.g.nzc;
    const int pz = PIECES ? A.g.cpp * kWave : Z;  // planes of each wave's sums
    const int zbase = PIECES ? c0 * kWave : 0;
    double* red = lds;  // [W][kNG]
    double* planes = lds + (size_t)W * kNG;
    double* mine = planes + (size_t)wave * kNQ * pz;
    double* shl = planes + (size_t)W * kNQ * pz;
    const bool shl_lds = step_shifts_in_lds(Z);

    if constexpr (STATS) {
        for (int i = lane; i < kNQ * pz; i += kWave) mine[i] = 0.0;
        if (shl_lds)
            for (int i = threadIdx.x; i < 3 * Z; i += blockDim.x) shl[i] = A.shift[i];
        __syncthreads();
    }
    StepAcc acc;
    const double G = *A.G;
    if (!STATS || shl_lds) {
        for (int it = bgrp * W + wave; it < A.g.ncols; it += ngrp * W)
            step_column_of<T, FIRST, STATS, BANDS, R>(A, __builtin_amdgcn_readfirstlane(it), lane, G, shl, mine,
                                                      acc, c0, c1, pz, zbase);
    } else {
        // the same loop with the shifts read from global memory
        for (int it = bgrp * W + wave; it < A.g.ncols; it += ngrp * W)
            step_column_of<T, FIRST, STATS, BANDS, R>(A, __builtin_amdgcn_readfirstlane(it), lane, G, A.shift,
                                                      mine, acc, c0, c1, pz, zbase);
    }
    step_block_reduce<STATS>(red, planes, acc, Z, pz, zbase, W, wave, lane, A.part, A.gpart, ngrp, bgrp,
                             (int)gridDim.x, (int)blockIdx.x);
}

// Per-column wall words from the wall masks (the pull table of the push
// rule xcomputations.c:368-375, and its bounce-back pushes).
static __global__ __launch_bounds__(256) void wall_word_kernel(const int8_t* __restrict__ mbot,
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
        const int xp = (x + 1) % X, xm = (x - 1 + X) % X;
        if (mW == -1 && m[xp * Y + y] == -1) w |= 1u << (side ? 12 : 10);
        if (mW == -1 && m[xm * Y + y] == -1) w |= 1u << (side ? 13 : 11);
    }
    ww[col] = w;
}

// Neighbour records (ColNbr) of every column.
static __global__ __launch_bounds__(256) void column_nbr_kernel(int X, int Y, uint32_t* __restrict__ out)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= X * Y) return;
    const int x = col / Y, y = col - x * Y;
    const int sx[3] = {x == X - 1 ? 0 : x + 1, x, x == 0 ? X - 1 : x - 1};
    const int sy[3] = {y == Y - 1 ? 0 : y + 1, y, y == 0 ? Y - 1 : y - 1};
    uint32_t* r = out + (size_t)col * 16;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r[3 * i + j] = (uint32_t)(sx[i] * Y + sy[j]);
    for (int k = 9; k < 16; k++) r[k] = 0u;
}

// Band neighbourhood records (BandNbr) of every column.
static __global__ __launch_bounds__(256) void band_nbr_kernel(const uint32_t* __restrict__ bands, int X, int Y,
                                                             uint32_t* __restrict__ out)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= X * Y) return;
    const int x = col / Y, y = col - x * Y;
    const int sx[3] = {x == X - 1 ? 0 : x + 1, x, x == 0 ? X - 1 : x - 1};
    const int sy[3] = {y == Y - 1 ? 0 : y + 1, y, y == 0 ? Y - 1 : y - 1};
    uint32_t* r = out + (size_t)col * 16;
    uint32_t mb = 0, mt = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const uint32_t w = bands[sx[i] * Y + sy[j]];
            r[3 * i + j] = w;
            mb = max(mb, w & 0xFFFFu);
            mt = max(mt, w >> 16);
        }
    r[9] = mb | (mt << 16);
    for (int k = 10; k < 16; k++) r[k] = 0u;
}

// Per-lane accumulators of the chunk-mapped kernels (one wave per
// (column, 64-row chunk) of the block's z-piece, blockIdx.y): per-plane
// sums over the cpb column slots in fixed order, written as partial
// blockIdx.x of [kNQ][Z][gridDim.x]; global sums in a fixed order, written
// as partial blockIdx.y * gridDim.x + blockIdx.x of [kNG][gridDim.x *
// gridDim.y].  zcl: the wave's chunk within the piece.
__device__ __forceinline__ void block_reduce_write(const StepArgs& A, double (&acc)[kNAcc], int wave, int lane,
                                                   int zcl, int z)
{
    __shared__ double red[kMaxWaves][kNAcc][kWave];
    const int cpp = A.g.cpp;
    const int nw = cpp * A.g.cpb;
#pragma unroll
    for (int q = 0; q < kNAcc; q++) red[wave][q][lane] = acc[q];
    __syncthreads();
    if (wave < cpp && z < A.g.Z) {
#pragma unroll
        for (int q = 0; q < kNQ; q++) {
            double s = 0.0;
            for (int sl = 0; sl < A.g.cpb; sl++) s += red[sl * cpp + zcl][q][lane];
            A.part[((long long)q * A.g.Z + z) * gridDim.x + blockIdx.x] = s;
        }
    }
    if (wave == 0) {
        const long long ng = (long long)gridDim.x * gridDim.y, gs = (long long)blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
        for (int q = kNQ; q < kNAcc; q++) {
            double s = 0.0;
            for (int w = 0; w < nw; w++) s += red[w][q][lane];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
            if (lane == 0) A.gpart[(long long)(q - kNQ) * ng + gs] = s;
        }
    }
}

// Moments-only pass over the current state (no collision, no write):
// the same accumulators as the step, used after a load/init to derive the
// first step's force (encalc.c:121 on the initial state) and the shifts.
template <typename T, bool LOCAL>
static __global__ __launch_bounds__(kWave* kMaxWaves) void moments_kernel(StepArgs A)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int zcl = wave % A.g.cpp;  // chunk within the block's z-piece
    const int slot = wave / A.g.cpp;
    const int X = A.g.X, Y = A.g.Y, Z = A.g.Z;
    const int z = (blockIdx.y * A.g.cpp + zcl) * kWave + lane;
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
        if (A.bands && band_solid(A.bands[col], z, Z)) continue;
        T v[Q];
        if constexpr (LOCAL)
            load_local19(src, c, z, v);
        else
            pull19(src, c, z, A.mbot, A.mtop, A.bands, v);
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
            for (int a = 0; a < Q; a++) dr += (double)Store<T>::to_g(v[a]);
            acc[3] += dr;
        }
        acc[kNQ + 0] += ux;  // jx
        acc[kNQ + 1] += 0.0; // mass budget (unused for the initial state)
        acc[kNQ + 2] += ux;  // momentum of the state itself
        acc[kNQ + 3] += (du * du + dv * dv) + dw * dw;
    }
    block_reduce_write(A, acc, wave, lane, zcl, z);
}

// Fixed-order reduction of the partials: block b sums row b of a
// [rows][nblk] matrix.
static __global__ __launch_bounds__(256) void reduce_rows_kernel(const double* __restrict__ in, int nblk,
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
    double G;         // force of the next step
    double rho_mean;  // V.rho of the last logged step
    // FP64 time accumulators of the logged rows since the last reset
    double n_rows, s_pgrad, s_stress, s_stup, s_stdown, s_U, s_K, s_rho;
    double n_prof;  // profile samples accumulated (tavg_profile_kernel)
    double steps;   // step counter: the counter of the next step's log row
    double pad[6];
};

struct LogRow {  // mirrors dns_steplog
    double t, K, flowrate, massrate, stup, stdown, stress, pgrad, force, rho;
};

struct FinalizeArgs {
    const double* planes;   // [kNQ][Z] plane sums (may be null when !stats)
    const double* globals;  // [kNG]: jx, mass-1, budget momentum, sum |u - c|^2
    double* shift;          // [3][Z] in/out
    DevState* st;
    LogRow* row;            // null: no log row
    int X, Y, Z;
    int forcing;            // 0 const Q, 1 const F
    double ub, F, mu, TM, dt;
    long long cntr;
    int from_moments;       // 1: globals came from moments_kernel (initial state)
    const double* nf;       // [Z] fluid nodes per plane (null: X*Y, flat walls)
    double nfluid;          // fluid nodes in total
    int cntr_dev;           // 1: the row counter is DevState::steps instead of cntr
};

// Per-step finisher (encalc.c:29-138 and wshearstr.c:26-154 restated on
// plane sums): the log row of this step, the next step's force and the
// new plane-mean shifts.  One block; thread z owns plane z, sums over z in
// a fixed tree.
// Scratch of the finisher: four arrays of blockDim.x (<= 1024) doubles.
struct FinalizeShared {
    double sU[1024], sK[1024], uav[1024], dav[1024];
};

__device__ __forceinline__ void finalize_body(const FinalizeArgs& A, FinalizeShared& S)
{
    double* sU = S.sU;
    double* sK = S.sK;
    double* uav = S.uav;
    double* dav = S.dav;
    const int Z = A.Z;
    const double Nxy = (double)A.X * (double)A.Y;
    const double N = Nxy * (double)Z;
    const int t = threadIdx.x;
    const bool have_planes = A.planes != nullptr;
    double uz = 0.0, kz = 0.0;
    if (have_planes && t < Z) {
        // The sums run over the nf fluid nodes of plane t, shifted by c.
        // The flow rate and the wall stress take the plane means over the
        // whole plane, solid nodes counting as u = 0 (encalc.c:54-59 and
        // wshearstr.c divide by the full plane): uav = (nf c + S1) / Nxy.
        // K takes the fluctuations about the mean over the fluid nodes,
        // m = (nf c + S1) / nf, which is also the next shift:
        //   sum_f |u - m|^2 = S2 - 2 d.S1 + nf |d|^2 with d = m - c.
        // With flat walls nf = Nxy and both are the reference's.
        const double nf = A.nf ? A.nf[t] : Nxy;
        const double s1u = A.planes[0 * Z + t], s1v = A.planes[1 * Z + t], s1w = A.planes[2 * Z + t];
        const double cu = A.shift[t], cv = A.shift[Z + t], cw = A.shift[2 * Z + t];
        const double nfi = nf > 0.0 ? nf : 1.0;
        const double mu_ = (nf * cu + s1u) / nfi, mv = (nf * cv + s1v) / nfi, mw = (nf * cw + s1w) / nfi;
        const double du = mu_ - cu, dv = mv - cv, dw = mw - cw;
        uz = (nf * cu + s1u) / Nxy;
        kz = 2.0 * ((du * s1u + dv * s1v) + dw * s1w) - nf * ((du * du + dv * dv) + dw * dw);
        uav[t] = uz;
        dav[t] = (nf + A.planes[3 * Z + t]) / Nxy;
        A.shift[t] = mu_;
        A.shift[Z + t] = mv;
        A.shift[2 * Z + t] = mw;
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
            for (int z = 0; z < Z; z++) srho += A.planes[3 * Z + z];
            const double denst = (A.nfluid + srho) / N;
            const double momentum = jx / N;
            const double K = 0.5 * (A.globals[3] - sK[0]) / A.nfluid;  // per fluid node
            const double U = sU[0] / (double)Z;
            const double stdown = (-2. * uav[0] + 3. * uav[1] - uav[2]) * (A.mu * dav[0]);
            const double stup = -(2. * uav[Z - 1] - 3. * uav[Z - 2] + uav[Z - 3]) * (A.mu * dav[Z - 3]);
            const long long cntr = A.cntr_dev ? (long long)A.st->steps : A.cntr;
            LogRow r;
            r.t = A.TM + A.dt * (double)cntr;
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
            A.st->n_rows += 1.0;
            A.st->s_pgrad += r.pgrad;
            A.st->s_stress += r.stress;
            A.st->s_stup += stup;
            A.st->s_stdown += stdown;
            A.st->s_U += U;
            A.st->s_K += K;
            A.st->s_rho += denst;
        }
        if (!A.from_moments) A.st->steps = (A.cntr_dev ? A.st->steps : (double)A.cntr) + 1.0;
        // force of the next step (encalc.c:121 on the next state)
        if (A.forcing == 0) {
            double rho_next, mom_next;
            if (A.from_moments) {
                double srho = 0.0;
                for (int z = 0; z < Z; z++) srho += A.planes[3 * Z + z];
                rho_next = (A.nfluid + srho) / N;
                mom_next = jx / N;
            } else {
                rho_next = (A.nfluid + A.globals[1]) / N;
                mom_next = A.globals[2] / N;
            }
            A.st->G = rho_next * A.ub - mom_next;
        } else {
            A.st->G = A.F;
        }
    }
}

static __global__ __launch_bounds__(1024) void finalize_kernel(FinalizeArgs A)
{
    __shared__ FinalizeShared S;
    finalize_body(A, S);
}

}  // namespace lbm
