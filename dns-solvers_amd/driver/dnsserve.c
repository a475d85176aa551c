/* dnsserve.c -- continuous-batching service for sweeps of channel solves
 * (see dnsserve.h). */
#define _POSIX_C_SOURCE 200809L
#include "dnsserve.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <time.h>

#include "dnscase.h"

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void sleep_s(double s)
{
    if (s <= 0) return;
    struct timespec ts;
    ts.tv_sec = (time_t)s;
    ts.tv_nsec = (long)((s - (double)ts.tv_sec) * 1e9);
    nanosleep(&ts, NULL);
}

/* splitmix64 stream */
static uint64_t next_u64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double next_unit(uint64_t *s) { return (double)(next_u64(s) >> 11) * (1.0 / 9007199254740992.0); }

void dnsserve_defaults(dnsserve_config *cfg)
{
    memset(cfg, 0, sizeof *cfg);
    cfg->precision = DNSGPU_FP32;
    cfg->ub = 0.1;
    cfg->transient = 2000;
    cfg->window = 2000;
    cfg->check = 500;
    cfg->tol = 0.005;
    cfg->cap = 10000;
    cfg->budget = 0;
    cfg->realtime = 1;
    cfg->autotune = 1;
    cfg->init_modes = 8;
    cfg->init_amp = 0.1;
    cfg->pair_arenas = 1;
}

void dnsserve_generate(int n, uint64_t seed, double mean_gap, dnsserve_request *out)
{
    static const int xs[3] = {128, 256, 512}, ys[3] = {64, 128, 256}, zs[3] = {64, 96, 128};
    uint64_t s = seed;
    double t = 0.0;
    for (int i = 0; i < n; i++) {
        dnsserve_request *r = out + i;
        memset(r, 0, sizeof *r);
        r->xdv = xs[next_u64(&s) % 3];
        r->ydv = ys[next_u64(&s) % 3];
        r->zdv = zs[next_u64(&s) % 3];
        r->retau = 90.0 + 90.0 * next_unit(&s);
        if (next_unit(&s) < 0.5) {
            r->texture = DNSSERVE_GROOVES;
            r->period = 8 + (int)(next_u64(&s) % 57);
            r->param = 0.3 + 0.6 * next_unit(&s);
        } else {
            r->texture = DNSSERVE_RIBLETS;
            r->period = 8 + (int)(next_u64(&s) % 25);
            r->param = (double)(4 + (int)(next_u64(&s) % 9));
        }
        r->seed = next_u64(&s);
        r->arrival = t;
        t += -mean_gap * log(1.0 - next_unit(&s));
    }
}

void dnsserve_case(const dnsserve_request *r, const dnsserve_config *cfg, dns_case *c, dns_init *init)
{
    const double reb = pow(r->retau / 0.09, 1.0 / 0.88);
    memset(c, 0, sizeof *c);
    c->xdv = r->xdv;
    c->ydv = r->ydv;
    c->zdv = r->zdv;
    c->Rebs = reb;  /* tau = 3 ub zdv / Rebs + 1/2 (dns-const-q-xycut.c:216) */
    c->ub = cfg->ub;
    c->CFL = 0.1;
    c->TSTR = 0;
    c->forcing = DNSGPU_CONST_Q;
    c->force = 0.0;
    c->precision = cfg->precision;
    c->omega = 0.0;
    memset(init, 0, sizeof *init);
    init->kind = DNSGPU_INIT_TURBULENT;
    init->uc = cfg->ub / (reb / (2.0 * r->retau));
    init->amp = cfg->init_amp;
    init->seed = r->seed;
    init->nmodes = cfg->init_modes;
}

int dnsserve_open(const dnsserve_request *r, const dnsserve_config *cfg, dnsgpu **out)
{
    dns_case c;
    dns_init init;
    dnsserve_case(r, cfg, &c, &init);
    dnsgpu *h = NULL;
    int rc = dnsgpu_create(&c, &h);
    if (rc) return rc;
    dns_run_case rc_case;
    dns_case_defaults(&rc_case);
    rc_case.xdv = r->xdv;
    rc_case.ydv = r->ydv;
    rc_case.zdv = r->zdv;
    rc_case.walls = 0;
    const size_t xy = (size_t)r->xdv * r->ydv;
    if (r->texture == DNSSERVE_GROOVES) {
        rc_case.texture = 2;
        rc_case.period_y = r->period;
        rc_case.slip_y = (int)lround(r->param * r->period);
        rc_case.offset_y = 0;
        int8_t *b = malloc(xy), *t = malloc(xy);
        if (!b || !t) {
            free(b);
            free(t);
            dnsgpu_destroy(h);
            return DNSGPU_ENOMEM;
        }
        dns_case_masks(&rc_case, b, t);
        rc = dnsgpu_set_walls(h, b, t);
        free(b);
        free(t);
    } else if (r->texture == DNSSERVE_RIBLETS) {
        rc_case.texture = 3;
        rc_case.riblet_spacing = r->period;
        rc_case.riblet_height = r->param;
        rc_case.riblet_offset = 0.0;
        int16_t *b = malloc(xy * sizeof(int16_t)), *t = malloc(xy * sizeof(int16_t));
        if (!b || !t) {
            free(b);
            free(t);
            dnsgpu_destroy(h);
            return DNSGPU_ENOMEM;
        }
        dns_case_bands(&rc_case, b, t);
        rc = dnsgpu_set_solid_bands(h, b, t);
        free(b);
        free(t);
    }
    if (!rc) rc = dnsgpu_init_field(h, &init);
    if (rc) {
        dnsgpu_destroy(h);
        return rc;
    }
    *out = h;
    return 0;
}

typedef struct resident {
    int req;            /* request index */
    dnsgpu *h;
    double steps0;      /* member counter after admission */
    int averaging;      /* time averages running */
    double avg_start;   /* steps after admission when they started */
    double *hist;       /* running u_tau at each check, indexed by check number */
    int nhist;
} resident;

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static double percentile(const double *sorted, int n, double p)
{
    if (n <= 0) return 0.0;
    /* nearest rank */
    int k = (int)ceil(p / 100.0 * n) - 1;
    if (k < 0) k = 0;
    if (k >= n) k = n - 1;
    return sorted[k];
}

static void retire(dnsens *e, resident *m, const dnsserve_request *req, dnsserve_result *res, double t, int status)
{
    dnsserve_result *r = res + m->req;
    dnsens_member_info info;
    dnsens_info(e, m->h, &info);
    dns_tavg tv;
    memset(&tv, 0, sizeof tv);
    if (status >= 0) dnsgpu_tavg_get(m->h, &tv, NULL);
    r->finished = t;
    r->steps = info.steps - m->steps0 + 1.0; /* the admission's first step included */
    r->nodes = (double)req[m->req].xdv * req[m->req].ydv * req[m->req].zdv;
    r->utau = tv.utau;
    r->Re_tau = tv.Re_tau;
    r->Cf_w = tv.Cf_w;
    r->Cf_p = tv.Cf_p;
    r->Ub = tv.Ub;
    r->stress = tv.stress;
    r->u_slip = 0.5 * (tv.u_slip_bottom + tv.u_slip_top);
    r->status = status;
    dnsens_remove(e, m->h);
    dnsgpu_destroy(m->h);
    free(m->hist);
    m->h = NULL;
    m->hist = NULL;
}

/* Requests are prepared (handle created, textured, initialised: dnsserve_open)
 * by a thread of their own in arrival order while the GPU advances the
 * resident solves, and are admitted to the ensemble at the next chunk
 * boundary.  A request is prepared only when its footprint fits the budget
 * together with the members and the requests already prepared. */
typedef struct prep_state {
    const dnsserve_request *req;
    const dnsserve_config *cfg;
    dnsserve_result *res;
    const dnsens *e;
    int n;
    double t0;
    size_t budget;
    pthread_mutex_t mu;
    pthread_cond_t cv;
    size_t reserved;     /* footprints of the members and of the prepared requests */
    int nready;          /* prepared, not yet admitted: ready[head .. head + nready) */
    int ready_head;
    int *ready_req;
    dnsgpu **ready_h;
    int refused;         /* requests refused by the preparer (status -2) */
    int prepared_all;    /* the preparer has handled every request */
    int stop;            /* the sweep ends: the preparer returns */
    int error;           /* an error of the preparer (a DNSGPU_E* code) */
    char errmsg[512];    /* its message (the library's messages are per thread) */
} prep_state;

static void prep_fail(prep_state *p, int rc)
{
    pthread_mutex_lock(&p->mu);
    p->error = rc;
    snprintf(p->errmsg, sizeof p->errmsg, "%s", dnsgpu_last_error());
    p->stop = 1;
    pthread_cond_broadcast(&p->cv);
    pthread_mutex_unlock(&p->mu);
}

static void *prep_main(void *arg)
{
    prep_state *p = arg;
    const int brc = dnsens_bind_thread(p->e);
    if (brc) {
        prep_fail(p, brc);
        return NULL;
    }
    for (int i = 0; i < p->n; i++) {
        const dnsserve_request *r = p->req + i;
        if (p->cfg->realtime) sleep_s(r->arrival - (now_s() - p->t0));
        dns_case c;
        dns_init init;
        dnsserve_case(r, p->cfg, &c, &init);
        const size_t fp = dnsens_footprint(&c);
        pthread_mutex_lock(&p->mu);
        p->res[i].arrival = p->cfg->realtime ? r->arrival : 0.0;
        if (fp == 0 || fp > p->budget) { /* a case the library refuses, or larger than the budget */
            p->res[i].status = -2;
            p->res[i].finished = now_s() - p->t0;
            p->refused++;
            pthread_cond_broadcast(&p->cv);
            pthread_mutex_unlock(&p->mu);
            continue;
        }
        while (!p->stop && p->reserved + fp > p->budget) pthread_cond_wait(&p->cv, &p->mu);
        if (p->stop) {
            pthread_mutex_unlock(&p->mu);
            break;
        }
        p->reserved += fp;
        pthread_mutex_unlock(&p->mu);
        dnsgpu *h = NULL;
        int rc = dnsserve_open(r, p->cfg, &h);
        pthread_mutex_lock(&p->mu);
        if (rc == DNSGPU_ENOMEM && p->reserved > fp) {
            /* the device is fuller than the budget says: retry when a solve retires */
            const size_t before = p->reserved - fp;
            p->reserved -= fp;
            while (!p->stop && p->reserved >= before) pthread_cond_wait(&p->cv, &p->mu);
            pthread_mutex_unlock(&p->mu);
            if (p->stop) break;
            i--;
            continue;
        }
        if (rc) {
            p->reserved -= fp;
            pthread_mutex_unlock(&p->mu);
            prep_fail(p, rc);
            break;
        }
        /* charge what the member will hold (the same bytes as its footprint) */
        const size_t held = dnsens_member_bytes(h);
        p->reserved = p->reserved - fp + held;
        const int slot = p->ready_head + p->nready;
        p->ready_req[slot] = i;
        p->ready_h[slot] = h;
        p->nready++;
        pthread_cond_broadcast(&p->cv);
        pthread_mutex_unlock(&p->mu);
    }
    pthread_mutex_lock(&p->mu);
    p->prepared_all = 1;
    pthread_cond_broadcast(&p->cv);
    pthread_mutex_unlock(&p->mu);
    return NULL;
}

int dnsserve_sweep(const dnsserve_request *req, int n, const dnsserve_config *cfg, dnsserve_result *res,
                   dnsserve_summary *sum)
{
    if (!req || !cfg || !res || !sum || n < 0 || cfg->check < 1 || cfg->window % cfg->check) return DNSGPU_EARG;
    memset(res, 0, sizeof(*res) * (size_t)n);
    memset(sum, 0, sizeof *sum);
    dnsens *e = NULL;
    int rc = dnsens_create(cfg->precision, cfg->budget, cfg->autotune ? DNSENS_AUTOTUNE : 0, &e);
    if (rc) return rc;
    const int nn = n > 0 ? n : 1;
    resident *live = calloc((size_t)nn, sizeof(resident));
    prep_state p;
    memset(&p, 0, sizeof p);
    p.ready_req = malloc(sizeof(int) * (size_t)nn);
    p.ready_h = malloc(sizeof(dnsgpu *) * (size_t)nn);
    if (!live || !p.ready_req || !p.ready_h) {
        free(live);
        free(p.ready_req);
        free(p.ready_h);
        dnsens_destroy(e);
        return DNSGPU_ENOMEM;
    }
    p.req = req;
    p.cfg = cfg;
    p.res = res;
    p.e = e;
    p.n = n;
    p.budget = dnsens_budget(e);
    pthread_mutex_init(&p.mu, NULL);
    pthread_cond_init(&p.cv, NULL);
    const int max_checks = cfg->cap / cfg->check + 2;
    const int wchecks = cfg->window / cfg->check;
    int nlive = 0, done = 0, retired = 0;
    double used_peak = 0.0;
    const double t0 = now_s();
    p.t0 = t0;
    double busy = 0.0;
    if (cfg->pair_arenas) {
        /* one population buffer of every request in each memory half */
        const size_t esize = cfg->precision == DNSGPU_FP64 ? 8 : (cfg->precision == DNSGPU_FP16 ? 2 : 4);
        size_t want = 0;
        for (int i = 0; i < n; i++) want += (size_t)req[i].xdv * req[i].ydv * req[i].zdv * 19 * esize;
        if (want > p.budget / 2) want = p.budget / 2;
        size_t got = 0;
        rc = dnsgpu_reserve_pair_arenas(want, &got);
        if (rc) goto out_nothread;
        sum->arena_bytes = (double)got;
        sum->arena_s = now_s() - t0;
    }
    pthread_t th;
    if (pthread_create(&th, NULL, prep_main, &p) != 0) {
        rc = DNSGPU_EARG;
        goto out_nothread;
    }
    while (1) {
        /* admit what is prepared (at a chunk boundary); wait when nothing runs */
        pthread_mutex_lock(&p.mu);
        while (nlive == 0 && p.nready == 0 && !p.error && !(p.prepared_all && retired + p.refused >= n))
            pthread_cond_wait(&p.cv, &p.mu);
        done = retired + p.refused;
        if (p.error) {
            rc = p.error;
            /* the library's message is the preparer thread's: repeat it here */
            fprintf(stderr, "dnsserve: preparing a request failed: %s\n", p.errmsg);
            pthread_mutex_unlock(&p.mu);
            goto out;
        }
        if (nlive == 0 && p.nready == 0 && p.prepared_all && done >= n) {
            pthread_mutex_unlock(&p.mu);
            break;
        }
        int nadd = p.nready, head = p.ready_head;
        p.ready_head += nadd;
        p.nready = 0;
        pthread_mutex_unlock(&p.mu);
        for (int k = 0; k < nadd; k++) {
            const int i = p.ready_req[head + k];
            dnsgpu *h = p.ready_h[head + k];
            rc = dnsens_add(e, h);
            if (rc) {
                dnsgpu_destroy(h);
                goto out;
            }
            resident *m = live + nlive++;
            memset(m, 0, sizeof *m);
            m->req = i;
            m->h = h;
            dnsens_member_info info;
            dnsens_info(e, h, &info);
            m->steps0 = info.steps;
            m->hist = calloc((size_t)max_checks, sizeof(double));
            res[i].admitted = now_s() - t0;
            if (cfg->verbose)
                fprintf(stderr, "dnsserve: t=%.2fs admit %d (%dx%dx%d) resident=%d bytes=%zu\n", res[i].admitted, i,
                        req[i].xdv, req[i].ydv, req[i].zdv, nlive, dnsens_bytes_used(e));
            if ((double)dnsens_bytes_used(e) > used_peak) used_peak = (double)dnsens_bytes_used(e);
            if (nlive > sum->max_resident) sum->max_resident = nlive;
        }
        if (nlive == 0) continue;
        const double tb = now_s();
        rc = dnsens_step(e, cfg->check);
        if (rc) goto out;
        /* convergence checks, retirements */
        for (int k = 0; k < nlive;) {
            resident *m = live + k;
            dnsens_member_info info;
            dnsens_info(e, m->h, &info);
            const double s = info.steps - m->steps0; /* steps after admission, a multiple of check */
            int status = 2;                          /* 2: keep running */
            if (!info.finite) {
                status = -1;
            } else if (!m->averaging) {
                if (s >= cfg->transient) {
                    rc = dnsgpu_tavg_reset(m->h);
                    if (rc) goto out;
                    m->averaging = 1;
                    m->avg_start = s;
                }
            } else {
                const double ut = sqrt(fmax(info.mean_pgrad, 0.0));
                const int c = (int)((s - m->avg_start) / cfg->check);
                if (c < max_checks) m->hist[c] = ut;
                if (c - wchecks >= 1) {
                    const double prev = m->hist[c - wchecks];
                    if (ut > 0 && fabs(ut - prev) <= cfg->tol * ut) status = 1;
                }
                /* a profile sample per check for the slip velocity */
                rc = dnsgpu_tavg_sample(m->h);
                if (rc) goto out;
            }
            if (status == 2 && s + 1.0 >= cfg->cap) status = 0;
            if (status != 2) {
                const int ri = m->req;
                const size_t fp = dnsens_member_bytes(m->h);
                retire(e, m, req, res, now_s() - t0, status);
                pthread_mutex_lock(&p.mu);
                p.reserved -= fp;
                retired++;
                pthread_cond_broadcast(&p.cv);
                pthread_mutex_unlock(&p.mu);
                if (cfg->verbose)
                    fprintf(stderr, "dnsserve: t=%.2fs done %d status=%d steps=%.0f u_tau=%.4e\n", res[ri].finished,
                            ri, status, res[ri].steps, res[ri].utau);
                live[k] = live[--nlive];
            } else {
                k++;
            }
        }
        busy += now_s() - tb;
    }
    rc = 0;
    {
        double first = 1e300, last = 0.0, node_steps = 0.0;
        double *ttr = malloc(sizeof(double) * (size_t)nn);
        int nt = 0;
        for (int i = 0; i < n; i++) {
            if (res[i].arrival < first) first = res[i].arrival;
            if (res[i].finished > last) last = res[i].finished;
            if (res[i].status >= 0) {
                node_steps += res[i].nodes * res[i].steps;
                if (ttr) ttr[nt++] = res[i].finished - res[i].arrival;
            }
        }
        if (ttr) {
            qsort(ttr, (size_t)nt, sizeof(double), cmp_double);
            sum->p50 = percentile(ttr, nt, 50.0);
            sum->p99 = percentile(ttr, nt, 99.0);
            double acc = 0.0;
            for (int i = 0; i < nt; i++) acc += ttr[i];
            sum->mean_ttr = nt ? acc / nt : 0.0;
            free(ttr);
        }
        sum->completed = nt;
        sum->wall = n ? last - first : 0.0;
        sum->busy = busy;
        sum->node_steps = node_steps;
        sum->mlups = sum->wall > 0 ? node_steps / sum->wall / 1e6 : 0.0;
        sum->mlups_busy = busy > 0 ? node_steps / busy / 1e6 : 0.0;
        sum->solves_per_hour = sum->wall > 0 ? nt / sum->wall * 3600.0 : 0.0;
        sum->peak_bytes = used_peak;
    }
out:
    pthread_mutex_lock(&p.mu);
    p.stop = 1;
    pthread_cond_broadcast(&p.cv);
    pthread_mutex_unlock(&p.mu);
    pthread_join(th, NULL);
    /* prepared requests never admitted */
    for (int k = 0; k < p.nready; k++) dnsgpu_destroy(p.ready_h[p.ready_head + k]);
out_nothread:
    for (int k = 0; k < nlive; k++) {
        if (live[k].h) {
            dnsens_remove(e, live[k].h);
            dnsgpu_destroy(live[k].h);
        }
        free(live[k].hist);
    }
    free(live);
    free(p.ready_req);
    free(p.ready_h);
    pthread_mutex_destroy(&p.mu);
    pthread_cond_destroy(&p.cv);
    dnsens_destroy(e);
    if (cfg->pair_arenas) (void)dnsgpu_release_pair_arenas();
    return rc;
}

int dnsserve_sequential(const dnsserve_request *req, int n, const dnsserve_config *cfg, int precision,
                        const double *steps, double *seconds)
{
    if (!req || !cfg || !steps || !seconds) return DNSGPU_EARG;
    dnsserve_config c1 = *cfg;
    c1.precision = precision;
    dns_steplog *rows = malloc(sizeof(dns_steplog) * 1024);
    if (!rows) return DNSGPU_ENOMEM;
    int rc = 0;
    for (int i = 0; i < n; i++) {
        seconds[i] = 0.0;
        if (steps[i] < 1) continue;
        const double t0 = now_s();
        dnsgpu *h = NULL;
        rc = dnsserve_open(req + i, &c1, &h);
        if (rc) break;
        long done = 0;
        const long total = (long)steps[i];
        while (done < total) {
            const int k = (int)(total - done < 1024 ? total - done : 1024);
            rc = dnsgpu_step(h, done, k, rows);
            if (rc) break;
            done += k;
        }
        dnsgpu_destroy(h);
        if (rc && rc != DNSGPU_ENAN) break;
        rc = 0;
        seconds[i] = now_s() - t0;
        if (cfg->verbose)
            fprintf(stderr, "dnsserve: sequential %d: %.0f steps in %.3fs\n", i, steps[i], seconds[i]);
    }
    free(rows);
    return rc;
}
