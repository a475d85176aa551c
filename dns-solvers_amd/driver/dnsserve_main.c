/* dnsserve -- serve the seeded texture sweep through one ensemble and
 * print one JSON object per solve and a summary line.
 *
 *   dnsserve [-n solves] [-s seed] [-g mean_gap_s] [-b] [-p fp32|fp16|fp64]
 *            [-c cap] [-w window] [-t transient] [-k check] [-q seq_precision] [-a]
 *
 *   -b   burst: every request arrives at t = 0 (no arrival times)
 *   -a   no pair arenas: the members' population buffers allocated plainly
 *   -q   also run the same solves one after another on the single-case
 *        path at this precision (fp64, fp32 or fp16) and report their time
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "dnsserve.h"

static int prec_of(const char *s)
{
    return !strcmp(s, "fp64") ? DNSGPU_FP64 : (!strcmp(s, "fp16") ? DNSGPU_FP16 : DNSGPU_FP32);
}

int main(int argc, char **argv)
{
    dnsserve_config cfg;
    dnsserve_defaults(&cfg);
    cfg.verbose = 1;
    int n = 64, opt, seq = -1;
    unsigned long long seed = 5;
    double gap = 2.0;
    while ((opt = getopt(argc, argv, "n:s:g:bp:c:w:t:k:q:a")) != -1) {
        switch (opt) {
        case 'n': n = atoi(optarg); break;
        case 's': seed = strtoull(optarg, NULL, 10); break;
        case 'g': gap = atof(optarg); break;
        case 'b': cfg.realtime = 0; break;
        case 'p': cfg.precision = prec_of(optarg); break;
        case 'c': cfg.cap = atoi(optarg); break;
        case 'w': cfg.window = atoi(optarg); break;
        case 't': cfg.transient = atoi(optarg); break;
        case 'k': cfg.check = atoi(optarg); break;
        case 'q': seq = prec_of(optarg); break;
        case 'a': cfg.pair_arenas = 0; break;
        default:
            fprintf(stderr, "usage: %s [-n solves] [-s seed] [-g gap] [-b] [-p fp32|fp16|fp64] [-c cap] [-w window] "
                            "[-t transient] [-k check] [-q fp64|fp32] [-a]\n", argv[0]);
            return 2;
        }
    }
    if (n < 1 || cfg.check < 1 || cfg.window % cfg.check) {
        fprintf(stderr, "dnsserve: need n >= 1, check >= 1 and window a multiple of check\n");
        return 2;
    }
    dnsserve_request *req = calloc((size_t)n, sizeof *req);
    dnsserve_result *res = calloc((size_t)n, sizeof *res);
    if (!req || !res) return 1;
    dnsserve_generate(n, seed, gap, req);
    dnsserve_summary sum;
    int rc = dnsserve_sweep(req, n, &cfg, res, &sum);
    if (rc) {
        fprintf(stderr, "dnsserve: %s\n", dnsgpu_last_error());
        return 1;
    }
    for (int i = 0; i < n; i++) {
        const dnsserve_request *q = req + i;
        const dnsserve_result *r = res + i;
        printf("{\"solve\": %d, \"grid\": [%d, %d, %d], \"Re_tau_target\": %.2f, \"texture\": \"%s\", "
               "\"period\": %d, \"param\": %.3f, \"arrival\": %.3f, \"admitted\": %.3f, \"finished\": %.3f, "
               "\"steps\": %.0f, \"status\": %d, \"u_tau\": %.6e, \"Re_tau\": %.3f, \"Cf_w\": %.6e, "
               "\"Cf_p\": %.6e, \"Ub\": %.6e, \"u_slip\": %.6e}\n",
               i, q->xdv, q->ydv, q->zdv, q->retau, q->texture == DNSSERVE_GROOVES ? "grooves" : "riblets",
               q->period, q->param, r->arrival, r->admitted, r->finished, r->steps, r->status, r->utau, r->Re_tau,
               r->Cf_w, r->Cf_p, r->Ub, r->u_slip);
    }
    double seq_s = 0.0;
    if (seq >= 0) {
        double *steps = calloc((size_t)n, sizeof(double)), *secs = calloc((size_t)n, sizeof(double));
        if (!steps || !secs) return 1;
        for (int i = 0; i < n; i++) steps[i] = res[i].status >= 0 ? res[i].steps : 0;
        rc = dnsserve_sequential(req, n, &cfg, seq, steps, secs);
        if (rc) {
            fprintf(stderr, "dnsserve: sequential: %s\n", dnsgpu_last_error());
            return 1;
        }
        for (int i = 0; i < n; i++) seq_s += secs[i];
        free(steps);
        free(secs);
    }
    printf("{\"summary\": true, \"solves\": %d, \"completed\": %d, \"wall_s\": %.3f, \"busy_s\": %.3f, "
           "\"solves_per_hour\": %.1f, \"p50_s\": %.3f, \"p99_s\": %.3f, \"mean_s\": %.3f, \"mlups\": %.1f, "
           "\"mlups_busy\": %.1f, \"max_resident\": %d, \"peak_bytes\": %.0f, \"pair_arena_s\": %.3f, "
           "\"pair_arena_bytes_per_half\": %.0f, \"sequential_s\": %.3f}\n",
           n, sum.completed, sum.wall, sum.busy, sum.solves_per_hour, sum.p50, sum.p99, sum.mean_ttr, sum.mlups,
           sum.mlups_busy, sum.max_resident, sum.peak_bytes, sum.arena_s, sum.arena_bytes, seq_s);
    free(req);
    free(res);
    return 0;
}
