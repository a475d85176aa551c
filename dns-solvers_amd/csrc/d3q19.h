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

// Coefficients of the FP32 collision, per weight class (rest, axis,
// diagonal), chosen as single-precision numbers whose sums are exact in real
// arithmetic, so that the FP32 collision conserves mass and momentum up to
// unbiased rounding (the single-precision weights fl(1/18), fl(1/36) give
// 18 fl(1/18) = 1 + 7.5e-9: a systematic gain of 1.5e-8 of the momentum per
// step, about 0.5% of the body force of a Re_tau 180-395 channel):
//   W (the weights):                 w0 + 6 wa + 12 wd = 1,  wa = 1/18 (1 - 6e-8), wd = wa/2
//   C (3 w, the linear term):        2 ca + 8 cd = 1,        cd = fl(1/12)
//   D (4.5 w, the quadratic term):   2 da + 8 dd = 3/2,      da = 1/4, dd = 1/8 (exact)
//   P (9 w, the force's e.u term):   2 pa + 8 pd = 3,        pa = 1/2, pd = 1/4 (exact)
// A coefficient is never multiplied into a node-independent factor (omega,
// or rho = 1 exactly near rest) before the node's own terms: such a product
// rounds the same way at every node and biases the sums the exact
// coefficients keep.
constexpr float kFW[3] = {0x1.555558p-2f, 0x1.c71c70p-5f, 0x1.c71c70p-6f};
constexpr float kFC[3] = {0.0f, 0x1.555554p-3f, 0x1.555556p-4f};
constexpr float kFD[3] = {0.0f, 0.25f, 0.125f};
constexpr float kFP[3] = {0.0f, 0.5f, 0.25f};

// The BGK update with the reference's force term on shifted populations
// g = f - w in FP32 (FMA allowed):
//   drho = sum g, rho = 1 + drho, j = sum e g, u = j/rho
//   g' = (1 - omega) g + omega (feq - w) + Fi
//   feq - w = W drho + rho (C e.u + D (e.u)^2 - W 1.5 u^2)
//   Fi      = G (-3 W u_x + e_x (C + P e.u))
// (with W, C, D, P = w, 3w, 4.5w, 9w this is xcomputations.c:330-341), evaluated
// a pair of opposite directions (a, b = opp(a)) at a time, which share the
// class coefficients, (e.u)^2 and e_x e.u:
//   moments from g_a + g_b and g_a - g_b;
//   per class: R = W (omega drho - 3 G u_x), R+- = R +- C G, V = -W 1.5 u^2;
//   m = D (e.u)^2 + V;
//   e_x = 0:  g'_a,b = (1 - omega) g_a,b + omega rho (m +- C e.u) + R;
//   e_x = s:  g'_a = (1 - omega) g_a + omega rho (m + C e.u) + (R_s + s P G e.u),
//             g'_b = (1 - omega) g_b + omega rho (m - C e.u) + (R_-s + s P G e.u).
// S: the populations are held multiplied by S (a power of two): S = 1 for
// FP32 storage, 2^16 for the FP16 storage's scaled values; the moments are
// unscaled once and the coefficients carry S, so no per-population rescaling.
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
    // u = j / (1 + drho) as j - j q with q = drho / (1 + drho), and
    // omega rho = omega + omega drho: near rest, 1 + drho rounds to one of a
    // few numbers next to 1 and the rounding of 1/rho would bias u by drho^2
    const float rinv = 1.0f / (1.0f + sr);
    const float q = sr * rinv;
    u0 = __builtin_fmaf(-sx, q, sx);
    u1 = __builtin_fmaf(-sy, q, sy);
    u2 = __builtin_fmaf(-sz, q, sz);
    const float v15 = 1.5f * __builtin_fmaf(u0, u0, __builtin_fmaf(u1, u1, u2 * u2));
    const float om1 = 1.0f - omega;
    const float orho = __builtin_fmaf(omega, sr, omega);
    // omega drho - 3 G u_x: a node's own value, multiplied by each W below
    const float t0 = __builtin_fmaf(-3.0f * Gx, u0, omega * sr);
    float R[3], Rp[3], Rm[3], V[3], pG[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        R[k] = (kFW[k] * S) * t0;
        V[k] = (kFW[k] * S) * -v15;
        const float cG = (kFC[k] * S) * Gx;
        Rp[k] = R[k] + cG;
        Rm[k] = R[k] - cG;
        pG[k] = (kFP[k] * S) * Gx;  // exact: P S is a power of two
    }
    g[0] = __builtin_fmaf(om1, g[0], __builtin_fmaf(orho, V[0], R[0]));
    static_for<1, Q>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        if constexpr (a < OPP[a]) {
            constexpr int b = OPP[a];
            constexpr int k = (EX[a] != 0) + (EY[a] != 0) + (EZ[a] != 0);  // 1 axis, 2 diagonal
            constexpr float C = kFC[k] * S, D = kFD[k] * S;
            const float eu = edotf<a>(u0, u1, u2);
            const float m = __builtin_fmaf(D, eu * eu, V[k]);
            const float ea = __builtin_fmaf(C, eu, m), eb = __builtin_fmaf(-C, eu, m);
            float ra, rb;
            if constexpr (EX[a] == 0) {
                ra = R[k];
                rb = R[k];
            } else if constexpr (EX[a] > 0) {
                ra = __builtin_fmaf(pG[k], eu, Rp[k]);
                rb = __builtin_fmaf(pG[k], eu, Rm[k]);
            } else {
                ra = __builtin_fmaf(-pG[k], eu, Rm[k]);
                rb = __builtin_fmaf(-pG[k], eu, Rp[k]);
            }
            g[a] = __builtin_fmaf(om1, g[a], __builtin_fmaf(orho, ea, ra));
            g[b] = __builtin_fmaf(om1, g[b], __builtin_fmaf(orho, eb, rb));
        }
    });
}

}  // namespace d3q19
