/*
 * dnsserve.h -- continuous-batching service for sweeps of channel solves.
 *
 * Requests (grid, target Re_tau, wall texture) arrive over time; each is
 * admitted into a dnsens ensemble as soon as its device footprint fits the
 * ensemble's budget, advanced together with every other resident solve,
 * and retired when its statistics converge:
 *
 *   - transient: the first `transient` steps after admission are not
 *     averaged;
 *   - then the running friction velocity u_tau = sqrt(<p-grad>) (the
 *     definition of TKE-budget/postprocessing/time_stats.c:114) is taken
 *     every `check` steps; the solve has converged when it changed by less
 *     than `tol` (relative) over the last `window` steps;
 *   - `cap` steps after admission end a solve that has not converged.
 *
 * The reference has no counterpart (one case per MPI job); each solve is
 * the reference's constant-flow-rate channel (encalc.c:121 forcing) with
 * the build's wall textures.
 */
#ifndef DNSSERVE_H
#define DNSSERVE_H

#include <stddef.h>
#include <stdint.h>

#include "../../include/dnsgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

enum dnsserve_texture { DNSSERVE_SMOOTH = 0, DNSSERVE_GROOVES = 1, DNSSERVE_RIBLETS = 2 };

typedef struct dnsserve_request {
    int xdv, ydv, zdv;  /* streamwise, spanwise, wall-normal fluid nodes */
    double retau;       /* target friction Reynolds number */
    int texture;        /* enum dnsserve_texture, on both walls */
    int period;         /* grooves: spanwise period; riblets: spacing (cells) */
    double param;       /* grooves: gas (slip) fraction; riblets: tip height (cells) */
    double arrival;     /* seconds after the start of the sweep */
    uint64_t seed;      /* initial field */
} dnsserve_request;

typedef struct dnsserve_config {
    int precision;     /* DNSGPU_FP64, DNSGPU_FP32 or DNSGPU_FP16 */
    double ub;         /* lattice bulk velocity */
    int transient;     /* steps not averaged */
    int window;        /* convergence window (steps) */
    int check;         /* steps between convergence checks (and per dnsens_step) */
    double tol;        /* relative change of the running u_tau */
    int cap;           /* steps after which a solve ends unconverged */
    size_t budget;     /* admission budget in device bytes (0: 90% of the free memory) */
    int realtime;      /* 1: requests arrive at their arrival times; 0: all at once */
    int autotune;      /* tune the ensemble's step grid per member-set shape */
    int init_modes;    /* Fourier modes of the initial field per periodic direction */
    double init_amp;   /* their amplitude, fraction of the centreline velocity */
    int verbose;       /* 1: a line on stderr per admission and per result */
    int pair_arenas;   /* 1: reserve pair arenas for the members' population buffers (dnsserve_sweep) */
} dnsserve_config;

typedef struct dnsserve_result {
    double arrival, admitted, finished; /* seconds since the start */
    double steps;                       /* steps taken */
    double nodes;                       /* xdv*ydv*zdv */
    double utau, Re_tau, Cf_w, Cf_p, Ub, stress; /* time averages over the statistics window */
    double u_slip;                      /* <u> of the wall-adjacent rows over slip columns (grooves) */
    int status;                         /* 1 converged, 0 ended at the cap, -1 diverged, -2 refused */
    int pad;
} dnsserve_result;

typedef struct dnsserve_summary {
    double wall;        /* first arrival to last result (s) */
    double busy;        /* time with at least one resident solve (s) */
    double node_steps;  /* sum over solves of nodes x steps */
    double mlups;       /* node_steps / wall */
    double mlups_busy;  /* node_steps / busy */
    double p50, p99;    /* time-to-result (finished - arrival) percentiles (s) */
    double mean_ttr;
    double solves_per_hour;
    double peak_bytes;  /* largest admitted footprint sum */
    double arena_s;     /* seconds spent reserving the pair arenas (inside wall) */
    double arena_bytes; /* bytes reserved per memory half (0: none) */
    int max_resident;
    int completed;
} dnsserve_summary;

/* Defaults: FP32, ub 0.1, transient 2000, window 2000, check 500,
 * tol 0.005, cap 10000, budget 0, realtime 1, autotune 1, 8 modes at 10%. */
void dnsserve_defaults(dnsserve_config *cfg);

/* The seeded texture sweep: n requests with exponential inter-arrival
 * times of mean `mean_gap` s (the first at 0), grids from
 * {128,256,512} x {64,128,256} x {64,96,128}, Re_tau ~ U[90,180], grooves
 * (period 8-64, gas fraction 0.3-0.9) or riblets (spacing 8-32, height 4-12). */
void dnsserve_generate(int n, uint64_t seed, double mean_gap, dnsserve_request *out);

/* The case of a request: constant flow rate at ub, Re_b from Dean's
 * correlation Re_tau = 0.09 Re_b^0.88 (Re_b on the full height), the
 * initial friction velocity u_b / (Re_b / (2 Re_tau)). */
void dnsserve_case(const dnsserve_request *r, const dnsserve_config *cfg, dns_case *c, dns_init *init);

/* Create, texture and initialise the handle of a request. */
int dnsserve_open(const dnsserve_request *r, const dnsserve_config *cfg, dnsgpu **out);

/* Serve n requests (sorted by arrival) through one ensemble.  A thread of
 * its own prepares the requests in arrival order (dnsserve_open: the handle
 * created, textured and initialised on the ensemble's device) once each
 * footprint fits the budget beside the resident and the already prepared
 * solves; the calling thread admits the prepared ones at every check
 * boundary, steps the ensemble and retires.  A request whose footprint
 * exceeds the whole budget is refused (status -2).  Each solve's results
 * do not depend on the mix, on the arrival times or on when it was
 * admitted.  With cfg->pair_arenas the sweep first reserves pair arenas on
 * the device (dnsgpu_reserve_pair_arenas) for the population buffers of all
 * n requests (at most half the budget per memory half), so each member's
 * two buffers lie in different halves of device memory, and releases them
 * at the end; the reservation's time counts in the wall time. */
int dnsserve_sweep(const dnsserve_request *req, int n, const dnsserve_config *cfg, dnsserve_result *res,
                   dnsserve_summary *sum);

/* The same solves one after another on the single-case path
 * (dnsgpu_step, logged steps) at `precision`, each for steps[i] steps
 * (1 + steps[i] - 1: the first logged step included); seconds[i] receives
 * each solve's wall time from create to destroy. */
int dnsserve_sequential(const dnsserve_request *req, int n, const dnsserve_config *cfg, int precision,
                        const double *steps, double *seconds);

#ifdef __cplusplus
}
#endif
#endif
