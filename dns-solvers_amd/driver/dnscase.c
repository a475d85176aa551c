/* dnscase.c -- case-file reader and texture masks of the host driver. */
#include "dnscase.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void dns_case_defaults(dns_run_case *c)
{
    memset(c, 0, sizeof *c);
    /* definitions.h:30-37 */
    c->Rebs = 100.0;
    c->ub = 0.1;
    c->zdv = 25;
    c->ydv = 8;
    c->xdv = 16;
    c->TSTR = 420;
    c->TEND = 440;
    c->CFL = 0.1;
    c->ITER = -1;
    c->forcing = 0;
    c->precision = 0;
    c->texture = 0;
    c->walls = 0;
    c->period_x = 4;
    c->period_y = 4;
    c->slip_x = 2;
    c->slip_y = 2;
    c->offset_x = 1;
    c->offset_y = 1;
    c->init = 0;
    c->init_seed = 1;
    c->init_modes = 4;
    c->riblet_spacing = 16;
    c->riblet_height = 8;
    c->block_x = 2;
    c->block_y = 2;
    c->block_height = 4;
    c->tavg_start = -1;
    c->tavg_every = 0;
    strcpy(c->outdir, ".");
}

static char *trim(char *s)
{
    while (isspace((unsigned char)*s)) s++;
    char *e = s + strlen(s);
    while (e > s && isspace((unsigned char)e[-1])) *--e = 0;
    return s;
}

static int set_key(dns_run_case *c, const char *k, const char *v, char *err, int errlen)
{
#define DBL(name)                       \
    if (!strcmp(k, #name)) {            \
        c->name = strtod(v, NULL);      \
        return 0;                       \
    }
#define INT(name)                             \
    if (!strcmp(k, #name)) {                  \
        c->name = (int)strtol(v, NULL, 10);   \
        return 0;                             \
    }
    DBL(Rebs) DBL(ub) DBL(CFL) DBL(force) DBL(omega) DBL(init_uc) DBL(init_amp)
    DBL(riblet_spacing) DBL(riblet_height) DBL(riblet_offset)
    INT(xdv) INT(ydv) INT(zdv) INT(TSTR) INT(TEND) INT(period_x) INT(period_y) INT(slip_x) INT(slip_y)
    INT(offset_x) INT(offset_y) INT(init_modes) INT(tavg_every) INT(block_x) INT(block_y) INT(block_height)
#undef DBL
#undef INT
    if (!strcmp(k, "tavg_start")) {
        c->tavg_start = strtol(v, NULL, 10);
        return 0;
    }
    if (!strcmp(k, "ITER")) {
        c->ITER = strtol(v, NULL, 10);
        return 0;
    }
    if (!strcmp(k, "init_seed")) {
        c->init_seed = strtoull(v, NULL, 10);
        return 0;
    }
    if (!strcmp(k, "PSLIP")) {
        if (atoi(v)) c->texture = 1;
        return 0;
    }
    if (!strcmp(k, "RSLIP")) {
        if (atoi(v)) c->texture = 2;
        return 0;
    }
    if (!strcmp(k, "forcing")) {
        if (!strcmp(v, "constq")) c->forcing = 0;
        else if (!strcmp(v, "constf")) c->forcing = 1;
        else goto bad;
        return 0;
    }
    if (!strcmp(k, "precision")) {
        if (!strcmp(v, "fp64")) c->precision = 0;
        else if (!strcmp(v, "fp32")) c->precision = 1;
        else if (!strcmp(v, "fp16")) c->precision = 2;
        else goto bad;
        return 0;
    }
    if (!strcmp(k, "texture")) {
        if (!strcmp(v, "smooth")) c->texture = 0;
        else if (!strcmp(v, "posts")) c->texture = 1;
        else if (!strcmp(v, "stripes")) c->texture = 2;
        else if (!strcmp(v, "riblets")) c->texture = 3;
        else if (!strcmp(v, "blocks")) c->texture = 4;
        else goto bad;
        return 0;
    }
    if (!strcmp(k, "walls")) {
        if (!strcmp(v, "both")) c->walls = 0;
        else if (!strcmp(v, "bottom")) c->walls = 1;
        else if (!strcmp(v, "top")) c->walls = 2;
        else goto bad;
        return 0;
    }
    if (!strcmp(k, "init")) {
        if (!strcmp(v, "rest")) c->init = 0;
        else if (!strcmp(v, "poiseuille")) c->init = 1;
        else if (!strcmp(v, "turbulent")) c->init = 2;
        else goto bad;
        return 0;
    }
    if (!strcmp(k, "outdir")) {
        snprintf(c->outdir, sizeof c->outdir, "%s", v);
        return 0;
    }
    snprintf(err, errlen, "unknown case key '%s'", k);
    return -1;
bad:
    snprintf(err, errlen, "bad value '%s' for key '%s'", v, k);
    return -1;
}

int dns_case_read(const char *path, dns_run_case *c, char *err, int errlen)
{
    FILE *fh = fopen(path, "r");
    if (!fh) {
        snprintf(err, errlen, "cannot open case file '%s'", path);
        return -1;
    }
    char line[1024];
    int ln = 0;
    while (fgets(line, sizeof line, fh)) {
        ln++;
        char *h = strchr(line, '#');
        if (h) *h = 0;
        char *s = trim(line);
        if (!*s) continue;
        char *eq = strchr(s, '=');
        if (!eq) {
            snprintf(err, errlen, "%s:%d: expected key = value", path, ln);
            fclose(fh);
            return -1;
        }
        *eq = 0;
        char *k = trim(s), *v = trim(eq + 1);
        if (set_key(c, k, v, err, errlen)) {
            fclose(fh);
            return -1;
        }
    }
    fclose(fh);
    if (c->xdv < 2 || c->ydv < 2 || c->zdv < 4) {
        snprintf(err, errlen, "grid %dx%dx%d too small", c->xdv, c->ydv, c->zdv);
        return -1;
    }
    if (c->CFL <= 0 || (int)(0.5 * c->zdv / c->CFL) < 1) {
        snprintf(err, errlen, "CFL %g gives no whole step per time unit", c->CFL);
        return -1;
    }
    if (c->texture == 3 && (c->riblet_spacing <= 0 || c->riblet_height < 0)) {
        snprintf(err, errlen, "riblet spacing must be > 0 and height >= 0");
        return -1;
    }
    if (c->texture == 4 && (c->block_x < 0 || c->block_y < 0 || c->block_height < 0)) {
        snprintf(err, errlen, "block sizes and height must be >= 0");
        return -1;
    }
    if (c->texture && (c->period_x < 1 || c->period_y < 1)) {
        snprintf(err, errlen, "texture periods must be >= 1");
        return -1;
    }
    return 0;
}

static int modp(int a, int p) { return ((a % p) + p) % p; }

void dns_case_masks(const dns_run_case *c, int8_t *bottom, int8_t *top)
{
    const int X = c->xdv, Y = c->ydv;
    for (int x = 0; x < X; x++)
        for (int y = 0; y < Y; y++) {
            int slip = 0;
            if (c->texture == 1)
                slip = modp(x - c->offset_x, c->period_x) < c->slip_x && modp(y - c->offset_y, c->period_y) < c->slip_y;
            else if (c->texture == 2)
                slip = modp(y - c->offset_y, c->period_y) < c->slip_y;
            const int8_t code = slip ? -2 : -1; /* riblet walls are no-slip */
            bottom[x * Y + y] = (c->walls == 0 || c->walls == 1) ? code : -1;
            top[x * Y + y] = (c->walls == 0 || c->walls == 2) ? code : -1;
        }
}

void dns_case_bands(const dns_run_case *c, int16_t *bottom, int16_t *top)
{
    const int X = c->xdv, Y = c->ydv;
    if (c->texture == 4) {
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++) {
                const int in = modp(x - c->offset_x, c->period_x) < c->block_x &&
                               modp(y - c->offset_y, c->period_y) < c->block_y;
                const int16_t h = in ? (int16_t)c->block_height : 0;
                bottom[x * Y + y] = (c->walls == 0 || c->walls == 1) ? h : 0;
                top[x * Y + y] = (c->walls == 0 || c->walls == 2) ? h : 0;
            }
        return;
    }
    for (int y = 0; y < Y; y++) {
        int16_t hb = 0;
        if (c->texture == 3) {
            const double yc = y + 0.5 - c->riblet_offset;
            double ph = fmod(yc, c->riblet_spacing) / c->riblet_spacing;
            if (ph < 0) ph += 1.0;
            const double H = c->riblet_height * (1.0 - fabs(2.0 * ph - 1.0));
            hb = H > 0.5 ? (int16_t)ceil(H - 0.5) : 0;
        }
        for (int x = 0; x < X; x++) {
            bottom[x * Y + y] = (c->walls == 0 || c->walls == 1) ? hb : 0;
            top[x * Y + y] = (c->walls == 0 || c->walls == 2) ? hb : 0;
        }
    }
}
