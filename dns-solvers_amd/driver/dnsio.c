/* dnsio.c -- the reference solver's output formats (see dnsio.h). */
#include "dnsio.h"

#include <stdlib.h>
#include <string.h>

void dnsio_path(char *buf, size_t len, const char *dir, const char *name)
{
    if (!dir || !*dir || !strcmp(dir, "."))
        snprintf(buf, len, "%s", name);
    else
        snprintf(buf, len, "%s/%s", dir, name);
}

static FILE *open_in(const char *dir, const char *name, const char *mode)
{
    char p[1024];
    dnsio_path(p, sizeof p, dir, name);
    return fopen(p, mode);
}

int dnsio_logs_open(dns_logs *L, const char *dir)
{
    memset(L, 0, sizeof *L);
    const char *names[7] = {"energy.txt", "flowrate.txt", "massrate.txt", "str-up.txt",
                            "str-do.txt", "w-str.txt",    "p-grad.txt"};
    FILE **slot[7] = {&L->energy, &L->flowrate, &L->massrate, &L->stup, &L->stdown, &L->wstr, &L->pgrad};
    for (int i = 0; i < 7; i++) {
        *slot[i] = open_in(dir, names[i], "a");
        if (!*slot[i]) {
            dnsio_logs_close(L);
            return -1;
        }
    }
    return 0;
}

void dnsio_logs_close(dns_logs *L)
{
    FILE *f[7] = {L->energy, L->flowrate, L->massrate, L->stup, L->stdown, L->wstr, L->pgrad};
    for (int i = 0; i < 7; i++)
        if (f[i]) fclose(f[i]);
    memset(L, 0, sizeof *L);
}

int dnsio_format_row(int which, const dns_steplog *r, char *buf, int len)
{
    /* encalc.c:128,132,136 and wshearstr.c:137,142,147,152 */
    switch (which) {
    case 0: return snprintf(buf, len, "%f %.14f\n", r->t, r->K);
    case 1: return snprintf(buf, len, "%f %.12f\n", r->t, r->flowrate);
    case 2: return snprintf(buf, len, "%f %.12f\n", r->t, r->massrate);
    case 3: return snprintf(buf, len, "%f %.12f\n", r->t, r->stup);
    case 4: return snprintf(buf, len, "%f %.12f\n", r->t, r->stdown);
    case 5: return snprintf(buf, len, "%f %.12f\n", r->t, r->stress);
    case 6: return snprintf(buf, len, "%f %.12f\n", r->t, r->pgrad);
    }
    return -1;
}

int dnsio_logs_append(dns_logs *L, const dns_steplog *rows, int n)
{
    FILE *f[7] = {L->energy, L->flowrate, L->massrate, L->stup, L->stdown, L->wstr, L->pgrad};
    char buf[128];
    for (int w = 0; w < 7; w++) {
        for (int i = 0; i < n; i++) {
            int m = dnsio_format_row(w, rows + i, buf, sizeof buf);
            if (m < 0 || fwrite(buf, 1, (size_t)m, f[w]) != (size_t)m) return -1;
        }
        if (fflush(f[w])) return -1;
    }
    return 0;
}

int dnsio_write_turbfld(const char *dir, long ts, const double *stats, int zd)
{
    char name[64];
    snprintf(name, sizeof name, "turb-fld.%.3d.%.3ld", 0, ts);
    FILE *sv = open_in(dir, name, "wb");
    if (!sv) return -1;
    size_t w = fwrite(stats, sizeof(double), 9 * (size_t)zd, sv);
    fclose(sv);
    if (w != 9 * (size_t)zd) return -1;
    /* uavg-temp.dat (statistics.c:150-162) */
    double *zps = calloc((size_t)zd, sizeof(double));
    if (!zps) return -1;
    for (int k = 1; k < zd; k++) zps[k] = k - 0.5;
    zps[zd - 1] = zd - 2;
    for (int k = 0; k < zd; k++) zps[k] /= ((double)(zd - 2.));
    sv = open_in(dir, "uavg-temp.dat", "w");
    if (!sv) {
        free(zps);
        return -1;
    }
    for (int k = 0; k < zd; k++) fprintf(sv, "%.12f %.12f\n", zps[k], stats[k]);
    fclose(sv);
    free(zps);
    return 0;
}

int dnsio_write_twoD(const char *dir, const double *jj, int zd, double ub)
{
    /* output_twoD.c:14-46 */
    double *y = calloc((size_t)zd, sizeof(double));
    if (!y) return -1;
    for (int k = 1; k < zd - 1; k++) y[k] = -6. * ub * ((k - 0.5) / (zd - 2.)) * ((k - 0.5) / (zd - 2.) - 1);
    FILE *f = open_in(dir, "2dresults-exact.dat", "w");
    if (!f) {
        free(y);
        return -1;
    }
    fprintf(f, "0. %.12f\n", 0.);
    for (int k = 1; k < zd - 1; k++) fprintf(f, "%.12f %.12lf\n", (k - 0.5) / (zd - 2.), y[k]);
    fprintf(f, "1. 0.\n");
    fclose(f);
    f = open_in(dir, "2dresults-lb.dat", "w");
    if (!f) {
        free(y);
        return -1;
    }
    fprintf(f, "0. %.12f\n", jj[0]);
    for (int k = 1; k < zd - 1; k++) fprintf(f, "%.12f %.12lf\n", (k - 0.5) / (zd - 2.), jj[k]);
    fprintf(f, "1. 0.\n");
    fclose(f);
    free(y);
    return 0;
}

int dnsio_write_force(const char *dir, long tlabel, double force)
{
    char name[64];
    snprintf(name, sizeof name, "force.%.4ld", tlabel);
    FILE *f = open_in(dir, name, "wb");
    if (!f) return -1;
    size_t w = fwrite(&force, sizeof(double), 1, f);
    fclose(f);
    return w == 1 ? 0 : -1;
}

int dnsio_read_force(const char *dir, long tlabel, double *force)
{
    char name[64];
    snprintf(name, sizeof name, "force.%.4ld", tlabel);
    FILE *f = open_in(dir, name, "rb");
    if (!f) return -1;
    size_t r = fread(force, sizeof(double), 1, f);
    fclose(f);
    return r == 1 ? 0 : -1;
}

int dnsio_write_tavg(const char *dir, long ts, const dns_tavg *t, const double *profiles, int zd)
{
    char name[64];
    snprintf(name, sizeof name, "tavg-fld.%.3d.%.3ld", 0, ts);
    FILE *f = open_in(dir, name, "wb");
    if (!f) return -1;
    size_t w = fwrite(profiles, sizeof(double), 9 * (size_t)zd, f);
    fclose(f);
    if (w != 9 * (size_t)zd) return -1;
    f = open_in(dir, "tavg-summary.txt", "w");
    if (!f) return -1;
    fprintf(f, "steps %.0f\nprofile_samples %.0f\n", t->nsteps, t->nsamples);
    fprintf(f, "pgrad %.12e\nu_tau %.12e\nRe_tau %.6f\n", t->pgrad, t->utau, t->Re_tau);
    fprintf(f, "stress %.12e\nstress_top %.12e\nstress_bottom %.12e\n", t->stress, t->stup, t->stdown);
    fprintf(f, "Ub %.12e\nK %.12e\nrho %.12e\n", t->Ub, t->K, t->rho);
    fprintf(f, "Cf_w %.12e\nCf_p %.12e\n", t->Cf_w, t->Cf_p);
    fprintf(f, "u_slip_bottom %.12e\nu_noslip_bottom %.12e\nu_slip_top %.12e\nu_noslip_top %.12e\n",
            t->u_slip_bottom, t->u_noslip_bottom, t->u_slip_top, t->u_noslip_top);
    fclose(f);
    return 0;
}
