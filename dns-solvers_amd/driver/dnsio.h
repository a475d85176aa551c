/*
 * dnsio.h -- writers and readers for the reference solver's files, byte for
 * byte in its formats (the consumers in TKE-budget/postprocessing and
 * quadrant-analysis read these):
 *
 *   energy.txt flowrate.txt massrate.txt     encalc.c:125-138
 *   str-up.txt str-do.txt w-str.txt p-grad.txt   wshearstr.c:133-154
 *   lbf.%.4ld force.%.4ld                   dns-const-q-xycut.c:491-514
 *   vel.%.4d den.%.4d                       vel_print.c:56-62,100-101
 *   turb-fld.%.3d.%.3ld uavg-temp.dat       statistics.c:137-162
 *   2dresults-exact.dat 2dresults-lb.dat    output_twoD.c:34-46
 *   report.code                             dns-const-q-xycut.c:524-540
 */
#ifndef DNSIO_H
#define DNSIO_H

#include <stddef.h>
#include <stdio.h>

#include "../../include/dnsgpu.h"

typedef struct dns_logs {
    FILE *energy, *flowrate, *massrate, *stup, *stdown, *wstr, *pgrad;
} dns_logs;

/* Open the seven log files for appending in dir. */
int dnsio_logs_open(dns_logs *L, const char *dir);
void dnsio_logs_close(dns_logs *L);
/* Append n rows with the reference's format strings. */
int dnsio_logs_append(dns_logs *L, const dns_steplog *rows, int n);
/* One row formatted as the reference prints it, for tests: which = 0..6 in
 * the order energy, flowrate, massrate, str-up, str-do, w-str, p-grad. */
int dnsio_format_row(int which, const dns_steplog *row, char *buf, int len);

/* statistics.c:137-162: turb-fld.000.NNN (9 x zd doubles) and uavg-temp.dat. */
int dnsio_write_turbfld(const char *dir, long ts, const double *stats, int zd);

/* output_twoD.c:34-46 from the wall-normal column ux (zd values, walls 0). */
int dnsio_write_twoD(const char *dir, const double *lb_ux, int zd, double ubulk);

/* Time averages: tavg-fld.000.NNN (9 x zd doubles, the turb-fld layout of
 * the time-averaged profiles) and tavg-summary.txt (u_tau, Re_tau, Cf,
 * bulk velocity, slip velocities, sample counts). */
int dnsio_write_tavg(const char *dir, long ts, const dns_tavg *t, const double *profiles, int zd);

/* force.NNNN: one native double. */
int dnsio_write_force(const char *dir, long tlabel, double force);
int dnsio_read_force(const char *dir, long tlabel, double *force);

/* Path helpers. */
void dnsio_path(char *buf, size_t len, const char *dir, const char *name);

#endif
