/*
 * lbm_oracle.c -- CPU restatement of DNS-solver's per-rank update loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see
 * lbm_oracle.h for who may load it and how each function is pinned.
 *
 * Floating-point expressions are written in the exact association order of
 * the reference and this file must be compiled with -ffp-contract=off so
 * that the results are bitwise those of the reference binary.  Reductions
 * run sequentially in the reference's loop order for the same reason.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Lattice tables: globals.c:3-14 (D3Q19 ordering, weights, opposite and
 * wall-normal mirror).  Restated from the lattice definition. */
static const double W[19] = {1. / 3.,  1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
                             1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                             1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
static const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, 1, -1, -1, 0, 0, 0, 0};
static const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
static const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static const double DEX[19] = {0., 1., -1., 0., 0., 0., 0., 1., -1., -1., 1., 1., 1., -1., -1., 0., 0., 0., 0.};
static const double DEY[19] = {0., 0., 0., 1., -1., 0., 0., 1., 1., -1., -1., 0., 0., 0., 0., 1., 1., -1., -1.};
static const double DEZ[19] = {0., 0., 0., 0., 0., 1., -1., 0., 0., 0., 0., 1., -1., -1., 1., 1., -1., -1., 1.};
static const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12, 17, 18, 15, 16};
static const int MIR[19] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 12, 11, 14, 13, 16, 15, 18, 17};

#define ZD(c) ((c)->Z + 2)
#define NODE(c, i, j, k) ((((size_t)(i) * (size_t)(c)->Y + (size_t)(j)) * (size_t)ZD(c)) + (size_t)(k))
#define POP(c, i, j, k, a) (NODE(c, i, j, k) * 19u + (size_t)(a))

static inline int mask_at(const oracle_case *c, int i, int j, int k)
{
    return c->s[NODE(c, i, j, k)];
}

double oracle_omega(double Rebs, double ub, int zdv)
{
    /* dns-const-q-xycut.c:216-217 with zd = zdv + 2 */
    int zd = zdv + 2;
    double tau = (3. / Rebs) * ub * (zd - 2.) + 0.5;
    return 1. / tau;
}

static double log_dt(const oracle_case *c)
{
    /* encalc.c:20, wshearstr.c:168 */
    int zd = ZD(c);
    return 1. / ((double)(0.5 * (zd - 2.) / c->CFL));
}

void oracle_encalc(const oracle_case *c, const double *ft, long cntr, oracle_log *row)
{
    const int X = c->X, Y = c->Y, zd = ZD(c);
    double *uavg = calloc((size_t)zd, sizeof(double));
    double *vavg = calloc((size_t)zd, sizeof(double));
    double *wavg = calloc((size_t)zd, sizeof(double));
    double *ui = calloc((size_t)zd, sizeof(double));
    double *vi = calloc((size_t)zd, sizeof(double));
    double *wi = calloc((size_t)zd, sizeof(double));
    long *nf = calloc((size_t)zd, sizeof(long));
    long nfluid = 0;
    double momentum = 0., denst = 0., K = 0., U = 0.;
    const double dt = log_dt(c);

    /* pass 1: encalc.c:29-53 */
    for (int i = 0; i < X; i++)
        for (int j = 0; j < Y; j++)
            for (int k = 1; k < zd - 1; k++) {
                if (mask_at(c, i, j, k) < 0) continue; /* 3-D solids only (never on flat walls) */
                const double *p = ft + POP(c, i, j, k, 0);
                double ux = 0., uy = 0., uz = 0., dens = 0.;
                for (int a = 0; a < 19; a++) {
                    dens += p[a];
                    ux += DEX[a] * p[a];
                    uy += DEY[a] * p[a];
                    uz += DEZ[a] * p[a];
                }
                momentum += ux;
                ux /= dens;
                uy /= dens;
                uz /= dens;
                denst += dens;
                uavg[k] += ux;
                vavg[k] += uy;
                wavg[k] += uz;
                nf[k]++;
                nfluid++;
            }
    /* encalc.c:54-59 divides the plane sums by the whole plane: the
     * superficial means (solid nodes count as u = 0), kept for the flow
     * rate; the fluctuations of K are taken about the means over the fluid
     * nodes of the plane and K per fluid node (with flat walls every node of
     * a plane is fluid and both are the reference's). */
    for (int k = 0; k < zd; k++) {
        ui[k] = nf[k] ? uavg[k] / ((double)nf[k]) : 0.;
        vi[k] = nf[k] ? vavg[k] / ((double)nf[k]) : 0.;
        wi[k] = nf[k] ? wavg[k] / ((double)nf[k]) : 0.;
        uavg[k] /= ((double)(X * Y));
        vavg[k] /= ((double)(X * Y));
        wavg[k] /= ((double)(X * Y));
    }
    momentum /= ((double)(X * Y * (zd - 2)));
    denst /= ((double)(X * Y * (zd - 2)));

    /* pass 2: encalc.c:84-111 */
    for (int i = 0; i < X; i++)
        for (int j = 0; j < Y; j++)
            for (int k = 1; k < zd - 1; k++) {
                if (mask_at(c, i, j, k) < 0) continue;
                const double *p = ft + POP(c, i, j, k, 0);
                double ux = 0., uy = 0., uz = 0., dens = 0.;
                for (int a = 0; a < 19; a++) {
                    dens += p[a];
                    ux += DEX[a] * p[a];
                    uy += DEY[a] * p[a];
                    uz += DEZ[a] * p[a];
                }
                ux /= dens;
                uy /= dens;
                uz /= dens;
                double uf = ux - ui[k], vf = uy - vi[k], wf = uz - wi[k];
                K += 0.5 * (uf * uf + vf * vf + wf * wf);
            }
    K /= ((double)nfluid);
    for (int k = 0; k < zd; k++) U += uavg[k]; /* encalc.c:116-119 */
    U /= ((double)(zd - 2));

    double force = c->forcing ? c->F : denst * c->ub - momentum; /* encalc.c:121 */
    row->t = ((double)c->TM + dt * cntr);
    row->K = K;
    row->flowrate = U * ((Y) * (zd - 2));
    row->massrate = (momentum * ((double)((Y) * (zd - 2))));
    row->force = force;
    row->rho = denst;
    free(uavg);
    free(vavg);
    free(wavg);
    free(ui);
    free(vi);
    free(wi);
    free(nf);
}

int oracle_wshearstr(const oracle_case *c, const double *ft, long cntr, oracle_log *row)
{
    const int X = c->X, Y = c->Y, zd = ZD(c);
    const double tau = c->omega;
    const double mu = (1. / 6.) * (2. / tau - 1); /* wshearstr.c:24 */
    double stdown = 0., stup = 0., davg = 0.;
    (void)cntr;

    for (int i = 0; i < X; i++) /* wshearstr.c:26-59 */
        for (int j = 0; j < Y; j++) {
            double ux[3], dens = 0.;
            for (int q = 0; q < 3; q++) {
                const double *p = ft + POP(c, i, j, 1 + q, 0);
                double u = 0.;
                dens = 0.;
                if (mask_at(c, i, j, 1 + q) < 0) { /* solid inside the fluid rows: u = 0, rho = 0 */
                    ux[q] = 0.;
                    continue;
                }
                for (int a = 0; a < 19; a++) {
                    dens += p[a];
                    u += DEX[a] * p[a];
                }
                ux[q] = u / dens;
                if (q == 0) davg += dens;
            }
            stdown += (-2. * ux[0] + 3. * ux[1] - ux[2]);
        }
    davg /= ((double)(X * Y));
    stdown /= ((double)(X * Y));
    stdown *= mu * davg;

    davg = 0.;
    for (int i = 0; i < X; i++) /* wshearstr.c:72-107: k = zd-2, zd-3, zd-4 */
        for (int j = 0; j < Y; j++) {
            double ux[3], dens = 0.;
            for (int q = 0; q < 3; q++) {
                const double *p = ft + POP(c, i, j, zd - 2 - q, 0);
                double u = 0.;
                dens = 0.;
                if (mask_at(c, i, j, zd - 2 - q) < 0) { /* solid inside the fluid rows */
                    ux[q] = 0.;
                    continue;
                }
                for (int a = 0; a < 19; a++) {
                    dens += p[a];
                    u += DEX[a] * p[a];
                }
                ux[q] = u / dens;
            }
            stup += -(2. * ux[0] - 3. * ux[1] + ux[2]);
            davg += dens; /* density of plane zd-4 (wshearstr.c:105) */
        }
    davg /= ((double)(X * Y));
    stup /= ((double)(X * Y));

    double pgrad = row->force;
    pgrad = pgrad * (zd - 2) / 2.;
    stup *= mu * davg;
    double stress = 0.5 * (stup + stdown);
    row->stup = stup;
    row->stdown = stdown;
    row->stress = stress;
    row->pgrad = pgrad;
    return isnan(stress) ? -1 : 0;
}

void oracle_collide_stream(const oracle_case *c, double *ft, double *f, double Gx)
{
    const int X = c->X, Y = c->Y, zd = ZD(c);
    const double tau = c->omega;

    /* collision in place on ftemp: xcomputations.c:311-343 for every node
     * with s > -1 */
    for (int i = 0; i < X; i++)
        for (int j = 0; j < Y; j++)
            for (int k = 0; k < zd; k++) {
                if (mask_at(c, i, j, k) <= -1) continue;
                double *p = ft + POP(c, i, j, k, 0);
                double ux = 0., uy = 0., uz = 0., dens = 0., ueq[3], v;
                for (int a = 0; a < 19; a++) {
                    ux += DEX[a] * p[a];
                    uy += DEY[a] * p[a];
                    uz += DEZ[a] * p[a];
                    dens += p[a];
                }
                ueq[0] = ux / dens;
                ueq[1] = uy / dens;
                ueq[2] = uz / dens;
                v = ueq[0] * ueq[0] + ueq[1] * ueq[1] + ueq[2] * ueq[2];
                for (int a = 0; a < 19; a++) {
                    double u = DEX[a] * ueq[0] + DEY[a] * ueq[1] + DEZ[a] * ueq[2];
                    double feq = W[a] * dens * (1. + 3. * u + 4.5 * u * u - 1.5 * v);
                    double Fi = W[a] * (3. * (DEX[a] - ueq[0]) + 9. * u * DEX[a]) * Gx;
                    p[a] += (Fi - tau * (p[a] - feq));
                }
            }

    /* push streaming: xcomputations.c:111-220,351-433 with the periodic
     * wrap that the one-rank halo (xcomputations.c:222-289,435-555)
     * implements.  Rule order of xcomputations.c:368-375. */
    for (int i = 0; i < X; i++)
        for (int j = 0; j < Y; j++)
            for (int k = 0; k < zd; k++) {
                if (mask_at(c, i, j, k) != ORC_FLUID) continue;
                const double *p = ft + POP(c, i, j, k, 0);
                for (int a = 0; a < 19; a++) {
                    int it = (i + EX[a] + X) % X, jt = (j + EY[a] + Y) % Y, kt = k + EZ[a];
                    int sld = mask_at(c, it, jt, kt);
                    int sld1 = mask_at(c, i, j, kt);
                    if (sld > -1)
                        f[POP(c, it, jt, kt, a)] = p[a];
                    else if (sld == -1 && sld1 == -1)
                        f[POP(c, i, j, k, OPP[a])] = p[a];
                    else if (sld == -1 && sld1 == -2)
                        f[POP(c, it, jt, k, MIR[a])] = p[a];
                    else if (sld == -2)
                        f[POP(c, it, jt, k, MIR[a])] = p[a];
                    else if (sld == -1)
                        /* only reachable with solids inside the fluid rows
                         * (riblets): halfway bounce-back, the reference has
                         * no rule here (it drops the population). */
                        f[POP(c, i, j, k, OPP[a])] = p[a];
                }
            }
}

void oracle_stream_coverage(const oracle_case *c, int32_t *cnt)
{
    const int X = c->X, Y = c->Y, zd = ZD(c);
    memset(cnt, 0, sizeof(int32_t) * (size_t)X * Y * zd * 19);
    for (int i = 0; i < X; i++)
        for (int j = 0; j < Y; j++)
            for (int k = 0; k < zd; k++) {
                if (mask_at(c, i, j, k) != ORC_FLUID) continue;
                for (int a = 0; a < 19; a++) {
                    int it = (i + EX[a] + X) % X, jt = (j + EY[a] + Y) % Y, kt = k + EZ[a];
                    int sld = mask_at(c, it, jt, kt);
                    int sld1 = mask_at(c, i, j, kt);
                    if (sld > -1)
                        cnt[POP(c, it, jt, kt, a)]++;
                    else if ((sld == -1 && sld1 == -2) || sld == -2)
                        cnt[POP(c, it, jt, k, MIR[a])]++;
                    else
                        cnt[POP(c, i, j, k, OPP[a])]++;
                }
            }
}

int oracle_step(const oracle_case *c, double *ft, double *f, long cntr, oracle_log *row)
{
    oracle_log r;
    memset(&r, 0, sizeof r);
    oracle_encalc(c, ft, cntr, &r);
    int bad = oracle_wshearstr(c, ft, cntr, &r);
    if (row) *row = r;
    if (bad) return -1;
    oracle_collide_stream(c, ft, f, r.force);
    /* swap (xcomputations.c:557-559): ftemp <- streamed, f <- collided */
    size_t n = (size_t)c->X * c->Y * ZD(c) * 19u;
    double *tmp = malloc(n * sizeof(double));
    memcpy(tmp, ft, n * sizeof(double));
    memcpy(ft, f, n * sizeof(double));
    memcpy(f, tmp, n * sizeof(double));
    free(tmp);
    return 0;
}

int oracle_run(const oracle_case *c, double *ft, double *f, long cntr0, long n, oracle_log *rows)
{
    for (long s = 0; s < n; s++) {
        if (oracle_step(c, ft, f, cntr0 + s, rows ? rows + s : NULL)) return -1;
    }
    return 0;
}

void oracle_statistics(const oracle_case *c, const double *ft, double *out)
{
    const int X = c->X, Y = c->Y, zd = ZD(c);
    double *UAVG = out + 0 * zd, *VAVG = out + 1 * zd, *WAVG = out + 2 * zd;
    double *UTIN = out + 3 * zd, *VTIN = out + 4 * zd, *WTIN = out + 5 * zd;
    double *REYS = out + 6 * zd, *DAVG = out + 7 * zd, *DTIN = out + 8 * zd;
    /* statistics.c:31-147 over the nodes of each plane, which with solid
     * bands inside the rows are its fluid nodes (the means and the rms over
     * the fluid; with flat walls the whole plane, as the reference) */
    long *nf = calloc((size_t)zd, sizeof(long));
    memset(out, 0, sizeof(double) * 9u * (size_t)zd);
    for (int i = 0; i < X; i++) /* statistics.c:31-52 */
        for (int j = 0; j < Y; j++)
            for (int k = 1; k < zd - 1; k++) {
                if (mask_at(c, i, j, k) < 0) continue;
                const double *p = ft + POP(c, i, j, k, 0);
                double u = 0., v = 0., w = 0., dens = 0.;
                for (int a = 0; a < 19; a++) {
                    dens += p[a];
                    u += DEX[a] * p[a];
                    v += DEY[a] * p[a];
                    w += DEZ[a] * p[a];
                }
                UAVG[k] += (u / dens);
                VAVG[k] += (v / dens);
                WAVG[k] += (w / dens);
                DAVG[k] += dens;
                nf[k]++;
            }
    for (int k = 0; k < zd; k++) {
        const double n = nf[k] ? (double)nf[k] : 1.;
        UAVG[k] /= n;
        VAVG[k] /= n;
        WAVG[k] /= n;
        DAVG[k] /= n;
    }
    for (int i = 0; i < X; i++) /* statistics.c:74-100 */
        for (int j = 0; j < Y; j++)
            for (int k = 1; k < zd - 1; k++) {
                if (mask_at(c, i, j, k) < 0) continue;
                const double *p = ft + POP(c, i, j, k, 0);
                double u = 0., v = 0., w = 0., dens = 0.;
                for (int a = 0; a < 19; a++) {
                    dens += p[a];
                    u += DEX[a] * p[a];
                    v += DEY[a] * p[a];
                    w += DEZ[a] * p[a];
                }
                double UF = (u / dens) - UAVG[k];
                double VF = (v / dens) - VAVG[k];
                double WF = (w / dens) - WAVG[k];
                double DF = dens - DAVG[k];
                UTIN[k] += UF * UF;
                VTIN[k] += VF * VF;
                WTIN[k] += WF * WF;
                REYS[k] += UF * WF;
                DTIN[k] += DF * DF;
            }
    for (int k = 0; k < zd; k++) {
        const double n = nf[k] ? (double)nf[k] : 1.;
        UTIN[k] /= n;
        VTIN[k] /= n;
        WTIN[k] /= n;
        REYS[k] /= n;
        DTIN[k] /= n;
    }
    free(nf);
    for (int k = 0; k < zd; k++) {
        UTIN[k] = sqrt(UTIN[k]);
        VTIN[k] = sqrt(VTIN[k]);
        WTIN[k] = sqrt(WTIN[k]);
        DTIN[k] = sqrt(DTIN[k]);
    }
}

void oracle_moments(const oracle_case *c, const double *ft, double *vel, double *den)
{
    const int X = c->X, Y = c->Y, zd = ZD(c);
    for (int i = 0; i < X; i++) /* vel_print.c:65-99 */
        for (int j = 0; j < Y; j++)
            for (int k = 0; k < zd; k++) {
                const double *p = ft + POP(c, i, j, k, 0);
                double ux = 0., uy = 0., uz = 0., dens = 0.;
                for (int a = 0; a < 19; a++) {
                    ux += DEX[a] * p[a];
                    uy += DEY[a] * p[a];
                    uz += DEZ[a] * p[a];
                    dens += p[a];
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
                size_t n = NODE(c, i, j, k);
                vel[n * 3 + 0] = ux;
                vel[n * 3 + 1] = uy;
                vel[n * 3 + 2] = uz;
                den[n] = dens;
            }
}

void oracle_twoD(const oracle_case *c, const double *ft, double *exact, double *lb)
{
    const int zd = ZD(c);
    const int i = (c->X + 1) / 2, j = (c->Y + 1) / 2; /* output_twoD.c:5 with xd=X+1, yd=Y+1 */
    const int ii = i % c->X, jj = j % c->Y;          /* ghost plane == plane 0 */
    for (int k = 0; k < zd; k++) {
        exact[k] = 0.;
        lb[k] = 0.;
    }
    for (int k = 1; k < zd - 1; k++) exact[k] = -6. * c->ub * ((k - 0.5) / (zd - 2.)) * ((k - 0.5) / (zd - 2.) - 1);
    for (int k = 1; k < zd - 1; k++) {
        const double *p = ft + POP(c, ii, jj, k, 0);
        double dens = 0.;
        for (int a = 0; a < 19; a++) {
            dens += p[a];
            lb[k] += DEX[a] * p[a];
        }
        lb[k] /= dens;
    }
}
