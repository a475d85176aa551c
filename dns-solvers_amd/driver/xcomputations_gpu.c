/*
 * xcomputations_gpu.c -- the reference's collide-stream step on the GPU
 * library, with the reference's own signature (fun_defs.h:6,
 * DNS-solver/src/xcomputations.c:5), so that the reference's unchanged
 * main, encalc, wshearstr, statistics and vel_print link against it: compile
 * the reference's sources with this file in place of xcomputations.c, the
 * reference's definitions.h for the case on the include path before
 * include/ (oracle/build_ref.sh with REF_SHIM=1), and link -ldnsgpu.
 *
 * One call does what the reference's does on one MPI rank: V.ftemp (the
 * post-stream populations encalc and wshearstr have just read) collides
 * with the force V.force (xcomputations.c:29) and relaxation rate tau (the
 * caller's tau = 1/(3 nu + 1/2), dns-const-q-xycut.c:216-217), streams with
 * the wall rule of the wall rows of V.s (codes -1 no-slip, -2 slip,
 * solid_init.c:3-8), and the two arrays swap (xcomputations.c:557-559): on
 * return V.ftemp holds the next step's post-stream populations and V.f this
 * step's post-collision ones, both over the (xd-1) x (yd-1) planes the
 * checkpoint holds (dns-const-q-xycut.c:491-503), the replica planes
 * x = xd-1 and y = yd-1 filled by the periodic wrap.
 *
 * Both arrays cross PCIe every call: this is the parity harness for the
 * reference's own driver, not a fast path (the library's own loop keeps the
 * state on the device).  One rank (XDIM = YDIM = 1), one case per process.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "definitions.h"
#include "dnsgpu.h"

static dnsgpu *g_h;
static double *g_ft, *g_f; /* [xd-1][yd-1][zd][19] staging */

static void fail(const char *what)
{
    fprintf(stderr, "xcomputations (GPU): %s: %s\n", what, dnsgpu_last_error());
    abort();
}

static void open_case(int xd, int yd, int zd, const POINTER *V, DP tau)
{
    const int X = xd - 1, Y = yd - 1, Z = zd - 2;
    dns_case c = dns_case_make(X, Y, Z, Rebs, ub, CFL, TSTR);
    c.forcing = DNSGPU_CONST_F; /* the force is the caller's V.force every call */
    c.force = V->force;
    c.omega = tau;
    if (dnsgpu_create(&c, &g_h)) fail("create");
    int8_t *bot = malloc((size_t)X * Y), *top = malloc((size_t)X * Y);
    if (!bot || !top) abort();
    for (int i = 0; i < X; i++)
        for (int j = 0; j < Y; j++) {
            bot[i * Y + j] = (int8_t)Fs(V->s, i, j, 0, xd, yd, zd);
            top[i * Y + j] = (int8_t)Fs(V->s, i, j, zd - 1, xd, yd, zd);
        }
    if (dnsgpu_set_walls(g_h, bot, top)) fail("walls");
    free(bot);
    free(top);
    const size_t n = (size_t)X * Y * zd * 19;
    g_ft = malloc(n * sizeof(double));
    g_f = malloc(n * sizeof(double));
    if (!g_ft || !g_f) abort();
}

/* the (xd-1) x (yd-1) planes of a [xd][yd][zd][19] array to or from the
 * staging layout, and the replica planes of the periodic wrap */
static void pack(int xd, int yd, int zd, const DP *a, double *s)
{
    const size_t row = (size_t)zd * 19;
    for (int i = 0; i < xd - 1; i++)
        for (int j = 0; j < yd - 1; j++)
            memcpy(s + ((size_t)i * (yd - 1) + j) * row, a + ((size_t)i * yd + j) * row, row * sizeof(double));
}

static void unpack(int xd, int yd, int zd, const double *s, DP *a)
{
    const size_t row = (size_t)zd * 19;
    for (int i = 0; i < xd; i++)
        for (int j = 0; j < yd; j++) {
            const int si = i == xd - 1 ? 0 : i, sj = j == yd - 1 ? 0 : j;
            memcpy(a + ((size_t)i * yd + j) * row, s + ((size_t)si * (yd - 1) + sj) * row, row * sizeof(double));
        }
}

POINTER xcomputations(int xd, int yd, int zd, POINTER V, PDATA pd, DP Gx, DP Gy, DP Gz, DP rhozero, DP tau,
                      DP ubulk)
{
    (void)pd;
    (void)Gx;
    (void)Gy;
    (void)Gz;
    (void)rhozero;
    (void)ubulk;
    if (!g_h) open_case(xd, yd, zd, &V, tau);
    const int X = xd - 1;
    pack(xd, yd, zd, V.ftemp, g_ft);
    pack(xd, yd, zd, V.f, g_f);
    if (dnsgpu_set_force(g_h, V.force)) fail("force");
    if (dnsgpu_load_aos(g_h, 0, X, g_ft, g_f) || dnsgpu_commit_load(g_h)) fail("load");
    if (dnsgpu_step(g_h, 0, 1, NULL)) fail("step");
    /* ftemp: the next post-stream state; f: this step's post-collision values */
    if (dnsgpu_save_aos(g_h, 0, X, g_ft, g_f)) fail("save");
    DP *next = V.f, *post = V.ftemp;
    unpack(xd, yd, zd, g_ft, next);
    unpack(xd, yd, zd, g_f, post);
    V.ftemp = next;
    V.f = post;
    return V;
}
