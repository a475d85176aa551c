// d3q19.h -- lattice constants and the per-node BGK update, device side.
//
// Direction order, weights, opposite and wall-normal mirror tables follow
// DNS-solver/src/globals.c:3-14 (0 rest; 1-6 +-x,+-y,+-z; 7-10 xy
// diagonals; 11-14 xz diagonals; 15-18 yz diagonals).  The tables must be
// identical to the reference's because the checkpoint files store the 19
// populations in this order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace d3q19 {

constexpr int Q = 19;
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, 1, -1, -1, 0, 0, 0, 0};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12, 17, 18, 15, 16};
constexpr int MIR[Q] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 12, 11, 14, 13, 16, 15, 18, 17};
constexpr double WD[Q] = {1. / 3.,  1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
                          1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                          1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};

// Signed sum s0*c0 + s1*c1 (+ s2*c2) over the nonzero lattice components,
// in component order, with no multiplication by 0 or +-1.  For the
// reference's `E0[a]*u0 + E1[a]*u1 + E2[a]*u2` (xcomputations.c:330) this
// is bitwise the same value: products with +-1 are exact and adding a
// signed zero leaves a nonzero sum unchanged.  Returns +0 when all three
// components are zero (the rest direction).
template <int a>
__device__ __forceinline__ double edot(double u0, double u1, double u2)
{
    double s = 0.0;
    bool first = true;
    if constexpr (EX[a] != 0) {
        s = (EX[a] > 0) ? u0 : -u0;
        first = false;
    }
    if constexpr (EY[a] != 0) {
        if (first) s = (EY[a] > 0) ? u1 : -u1;
        else s = (EY[a] > 0) ? (s + u1) : (s - u1);
        first = false;
    }
    if constexpr (EZ[a] != 0) {
        if (first) s = (EZ[a] > 0) ? u2 : -u2;
        else s = (EZ[a] > 0) ? (s + u2) : (s - u2);
    }
    return s;
}

template <int a>
__device__ __forceinline__ float edotf(float u0, float u1, float u2)
{
    float s = -0.0f;  // the additive identity: the first term costs no add
    if constexpr (EX[a] != 0) s += (EX[a] > 0) ? u0 : -u0;
    if constexpr (EY[a] != 0) s += (EY[a] > 0) ? u1 : -u1;
    if constexpr (EZ[a] != 0) s += (EZ[a] > 0) ? u2 : -u2;
    return s;
}

// Compile-time loop over directions.
template <int a, int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (a < N) {
        f(std::integral_constant<int, a>{});
        static_for<a + 1, N>(f);
    }
}

// Moments of one node in the reference's accumulation order
// (xcomputations.c:316-322): each of ux, uy, uz, dens is a left-to-right sum
// over a = 0..18 of the nonzero-E terms.
__device__ __forceinline__ void moments_fp64(const double (&f)[Q], double& ux, double& uy, double& uz,
                                             double& dens)
{
    ux = 0.0;
    uy = 0.0;
    uz = 0.0;
    dens = 0.0;
    static_for<0, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        if constexpr (EX[a] == 1) ux = ux + f[a];
        if constexpr (EX[a] == -1) ux = ux - f[a];
        if constexpr (EY[a] == 1) uy = uy + f[a];
        if constexpr (EY[a] == -1) uy = uy - f[a];
        if constexpr (EZ[a] == 1) uz = uz + f[a];
        if constexpr (EZ[a] == -1) uz = uz - f[a];
        dens = dens + f[a];
    });
}

// BGK collision with the reference's force term, FP64, bitwise the
// arithmetic of xcomputations.c:314-342 (Appendix-A order):
//   u   = e.ueq ; feq = (w*dens)*(((1+3u)+(4.5u)u)-1.5v)
//   Fi  = (w*((3(ex-ueq0)) + (9u)ex))*Gx ; f += Fi - omega*(f-feq)
// Must be compiled with -ffp-contract=off.
__device__ __forceinline__ void collide_fp64(double (&f)[Q], double ueq0, double ueq1, double ueq2, double dens,
                                             double omega, double Gx)
{
    const double v = ueq0 * ueq0 + ueq1 * ueq1 + ueq2 * ueq2;
    static_for<0, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        const double u = edot<a>(ueq0, ueq1, ueq2);
        double feq;
        if constexpr (a == 0)
            feq = (WD[a] * dens) * (1. - 1.5 * v);
        else
            feq = (WD[a] * dens) * (((1. + 3. * u) + (4.5 * u) * u) - 1.5 * v);
        double br;
        if constexpr (EX[a] == 0)
            br = 3. * (0. - ueq0);
        else if constexpr (EX[a] == 1)
            br = 3. * (1. - ueq0) + 9. * u;
        else
            br = 3. * (-1. - ueq0) - 9. * u;
        const double Fi = (WD[a] * br) * Gx;
        f[a] = f[a] + (Fi - omega * (f[a] - feq));
    });
}

// The same update on shifted populations g = f - w in FP32 (FMA allowed):
//   drho = sum g, rho = 1 + drho, j = sum e g, u = j/rho
//   g' = g + Fi - omega*(g - (feq - w)),  feq - w = w*(drho + rho*(3eu + 4.5eu^2 - 1.5u^2))
// evaluated a pair of opposite directions (a, b = opp(a)) at a time, which
// share w, |e.u| and e_x u_x:
//   moments from g_a + g_b and g_a - g_b;
//   q = 4.5eu^2 - 1.5u^2, poly_a,b = q +- 3eu;
//   omega (feq - w) = K1 poly + K2 with K1 = omega w rho, K2 = omega w drho;
//   Fi_a,b = t +- 3 e_x w G, t = w G (9 e_x eu - 3 u_x)  (Fi = w G (-3 u_x) when e_x = 0);
//   g'_a = (1 - omega) g_a + K1 poly_a + K2 + Fi_a.
// S: the populations are held multiplied by S (a power of two): S = 1 for
// FP32 storage, 2^16 for the FP16 storage's scaled values; the moments are
// unscaled once and the weights carry S, so no per-population rescaling.
template <int LOG2S = 0>
__device__ __forceinline__ void collide_fp32(float (&g)[Q], float omega, float Gx, float& drho, float& u0,
                                             float& u1, float& u2, float& jx)
{
    constexpr float S = (float)(1 << LOG2S), Sinv = 1.0f / (float)(1 << LOG2S);
    // -0 is the additive identity (x + -0 = x for every x), so the first
    // term of each sum costs no add
    float sx = -0.0f, sy = -0.0f, sz = -0.0f, sr = g[0];
    static_for<1, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        if constexpr (a < OPP[a]) {
            constexpr int b = OPP[a];
            const float pab = g[a] + g[b], mab = g[a] - g[b];
            sr += pab;
            if constexpr (EX[a] == 1) sx += mab;
            if constexpr (EX[a] == -1) sx -= mab;
            if constexpr (EY[a] == 1) sy += mab;
            if constexpr (EY[a] == -1) sy -= mab;
            if constexpr (EZ[a] == 1) sz += mab;
            if constexpr (EZ[a] == -1) sz -= mab;
        }
    });
    if constexpr (LOG2S != 0) {
        sx *= Sinv;
        sy *= Sinv;
        sz *= Sinv;
        sr *= Sinv;
    }
    drho = sr;
    jx = sx;
    const float rho = 1.0f + sr;
    const float rinv = 1.0f / rho;
    u0 = sx * rinv;
    u1 = sy * rinv;
    u2 = sz * rinv;
    const float v15 = 1.5f * __builtin_fmaf(u0, u0, __builtin_fmaf(u1, u1, u2 * u2));
    const float om1 = 1.0f - omega;
    const float m3u0 = -3.0f * u0;
    // per weight class k (rest 1/3, axis 1/18, diagonal 1/36, times S)
    constexpr float Wk[3] = {(float)(1. / 3.) * S, (float)(1. / 18.) * S, (float)(1. / 36.) * S};
    float K1[3], K2C[3], K2p[3], K2m[3], Ck[3], Pk[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float ow = omega * Wk[k], wG = Wk[k] * Gx;
        K1[k] = ow * rho;
        const float K2 = ow * sr;
        Ck[k] = wG * m3u0;
        K2C[k] = K2 + Ck[k];
        K2p[k] = __builtin_fmaf(3.0f, wG, K2);
        K2m[k] = __builtin_fmaf(-3.0f, wG, K2);
        Pk[k] = 9.0f * wG;
    }
    g[0] = __builtin_fmaf(om1, g[0], __builtin_fmaf(K1[0], -v15, K2C[0]));
    static_for<1, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        if constexpr (a < OPP[a]) {
            constexpr int b = OPP[a];
            constexpr int k = (EX[a] != 0) + (EY[a] != 0) + (EZ[a] != 0);  // 1 axis, 2 diagonal
            const float eu = edotf<a>(u0, u1, u2);
            const float q = __builtin_fmaf(4.5f * eu, eu, -v15);
            const float pa = __builtin_fmaf(3.0f, eu, q), pb = __builtin_fmaf(-3.0f, eu, q);
            float ia, ib;
            if constexpr (EX[a] == 0) {
                ia = __builtin_fmaf(K1[k], pa, K2C[k]);
                ib = __builtin_fmaf(K1[k], pb, K2C[k]);
            } else {
                const float t = __builtin_fmaf(EX[a] > 0 ? Pk[k] : -Pk[k], eu, Ck[k]);
                ia = __builtin_fmaf(K1[k], pa, (EX[a] > 0 ? K2p[k] : K2m[k]) + t);
                ib = __builtin_fmaf(K1[k], pb, (EX[a] > 0 ? K2m[k] : K2p[k]) + t);
            }
            g[a] = __builtin_fmaf(om1, g[a], ia);
            g[b] = __builtin_fmaf(om1, g[b], ib);
        }
    });
}

}  // namespace d3q19
