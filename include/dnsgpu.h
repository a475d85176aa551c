/*
 * dnsgpu.h -- C ABI of the MI355X (gfx950) D3Q19 BGK channel-flow step.
 *
 * The library replaces the per-rank update loop of DNS-solver
 * (DNS-solver/src/dns-const-q-xycut.c:485-519): the three physics calls
 *
 *   POINTER encalc(int xd,int yd,int zd,POINTER V,DP Gx,int TM,long cntr,
 *                  DP ubulk,MPI_Comm)                       fun_defs.h:5
 *   POINTER wshearstr(int xd,int yd,int zd,POINTER V,DP Gx,DP tau,int TM,
 *                     long cntr,MPI_Comm)                   fun_defs.h:9
 *   POINTER xcomputations(int xd,int yd,int zd,POINTER V,PDATA pd,DP Gx,
 *                         DP Gy,DP Gz,DP rhozero,DP tau,DP ubulk)
 *                                                           fun_defs.h:6
 *
 * and the cadence outputs statistics() (fun_defs.h:10), vel_print()
 * (fun_defs.h:11), output_twoD() (fun_defs.h:7) and the lbf/force
 * checkpoint (dns-const-q-xycut.c:405-435,491-514).  Where the reference
 * passes a POINTER struct by value and keeps the populations in host
 * memory, this ABI keeps the state on the GPU behind an opaque handle:
 * host arrays cross the boundary only at output cadence, in the
 * reference's own file layouts.
 *
 * Conventions
 *  - Axes and names as in the reference: x streamwise (xdv), y spanwise
 *    (ydv), z wall-normal (zdv fluid rows; the wall rows k=0 and k=zdv+1
 *    are implicit), definitions.h:30-34.
 *  - Host population arrays use the checkpoint layout
 *    [x][y][k=0..zdv+1][19] float64 (definitions.h:93, ghost planes
 *    excluded as in dns-const-q-xycut.c:144-155).
 *  - Every call returns 0 on success or a negative DNSGPU_E* code;
 *    dnsgpu_last_error() gives the message.  A NaN in the wall stress
 *    returns DNSGPU_ENAN (the reference aborts, wshearstr.c:126-130).
 *  - One host thread per handle; all work is ordered on the handle's
 *    HIP stream.  No entry point calls the HIP runtime's device-wide
 *    synchronisation.
 */
#ifndef DNSGPU_H
#define DNSGPU_H

#include <stddef.h>
#include <stdint.h>

/* The case macros of definitions.h:30-37 share their names with the fields
 * of dns_case: the header sets them aside while it declares, and restores
 * them at its end, so it can be included after definitions.h.  Code that
 * sees those macros fills a dns_case with dns_case_make. */
#pragma push_macro("Rebs")
#pragma push_macro("ub")
#pragma push_macro("xdv")
#pragma push_macro("ydv")
#pragma push_macro("zdv")
#pragma push_macro("TSTR")
#pragma push_macro("TEND")
#pragma push_macro("CFL")
#undef Rebs
#undef ub
#undef xdv
#undef ydv
#undef zdv
#undef TSTR
#undef TEND
#undef CFL

#ifdef __cplusplus
extern "C" {
#endif

#define DNSGPU_OK 0
#define DNSGPU_EHIP (-1)   /* HIP runtime error */
#define DNSGPU_ENAN (-2)   /* NaN in the wall stress (wshearstr.c:126-130) */
#define DNSGPU_EARG (-3)   /* bad size / argument */
#define DNSGPU_ENOMEM (-4) /* device allocation failed */

enum dnsgpu_precision {
    DNSGPU_FP64 = 0,  /* populations f in FP64, reference arithmetic bit for bit */
    DNSGPU_FP32 = 1,  /* shifted populations g = f - w in FP32, FP32 collision */
    DNSGPU_FP16 = 2   /* shifted populations in FP16 storage, FP32 collision */
};

/* FP16 storage against FP64, measured over 10^5 steps of the laminar
 * channel at the workloads' viscosity (the slowest-changing case): under
 * constant force the bulk velocity stays within 0.7% and the mean profile
 * within 1.5% at every zdv up to 512; under constant flow rate the force,
 * a difference of two numbers near u_b whose laminar value falls as
 * 1/zdv^2, carries FP16 rounding noise: its mean over 70000 steps is within
 * 2% of FP64 up to zdv 256, 18% off at 384 and 65% at 512.  dnsrun warns
 * above this zdv for FP16 with constant flow rate. */
#define DNSGPU_FP16_CONST_Q_ZMAX 256

enum dnsgpu_forcing {
    DNSGPU_CONST_Q = 0, /* force = rho_mean*ub - jx_mean every step (encalc.c:121) */
    DNSGPU_CONST_F = 1  /* constant body force */
};

/* Case inputs: the macros of definitions.h:30-37 plus the build's own keys. */
typedef struct dns_case {
    int xdv, ydv, zdv;  /* grid (definitions.h:32-34) */
    double Rebs;        /* bulk Reynolds number (definitions.h:30) */
    double ub;          /* bulk velocity (definitions.h:31) */
    double CFL;         /* output cadence utime = (int)(0.5*zdv/CFL) (definitions.h:37) */
    int TSTR;           /* time label of the first step (definitions.h:35) */
    int forcing;        /* enum dnsgpu_forcing */
    double force;       /* body force for DNSGPU_CONST_F */
    int precision;      /* enum dnsgpu_precision */
    double omega;       /* 1/tau; 0 => (3/Rebs)*ub*zdv+0.5 inverted (dns-const-q-xycut.c:216-217) */
} dns_case;

/* A case from definitions.h's macros, positionally (dns_case_make(xdv, ydv,
 * zdv, Rebs, ub, CFL, TSTR) in a file that includes definitions.h):
 * constant flow rate, FP64, omega from Rebs. */
static inline dns_case dns_case_make(int nx, int ny, int nz, double rebs, double ubulk, double cfl, int tstr)
{
    dns_case c;
    c.xdv = nx;
    c.ydv = ny;
    c.zdv = nz;
    c.Rebs = rebs;
    c.ub = ubulk;
    c.CFL = cfl;
    c.TSTR = tstr;
    c.forcing = DNSGPU_CONST_Q;
    c.force = 0.0;
    c.precision = DNSGPU_FP64;
    c.omega = 0.0;
    return c;
}

/* One row of the seven per-step logs, computed from the state before the
 * step's collision: encalc.c:125-138 and wshearstr.c:133-154. */
typedef struct dns_steplog {
    double t;        /* TM + dt*cntr (encalc.c:128) */
    double K;        /* energy.txt */
    double flowrate; /* flowrate.txt */
    double massrate; /* massrate.txt */
    double stup;     /* str-up.txt */
    double stdown;   /* str-do.txt */
    double stress;   /* w-str.txt */
    double pgrad;    /* p-grad.txt = force * zdv/2 */
    double force;    /* the V.force this step's collision used */
    double rho;      /* V.rho, mean density (encalc.c:79) */
} dns_steplog;

/* Initial fields generated on the device. */
enum dnsgpu_init_kind {
    DNSGPU_INIT_REST = 0,       /* f = w everywhere (init.c:14-20) */
    DNSGPU_INIT_POISEUILLE = 1, /* feq(1, u): u_x = 4 uc y(H-y)/H^2 + seeded noise */
    DNSGPU_INIT_TURBULENT = 2   /* feq(1, u): log-law mean + divergence-free Fourier modes */
};

/* TURBULENT: u = U(z) e_x + u', U the log law u_tau (ln(y+)/0.41 + 5.2)
 * (u_tau y+ below y+ = 11) from the nearer wall, u' the curl of a vector
 * potential sum_m sc_m env(z) (a1 cos(ph1), a2 cos(ph2), 0) with
 * env = sin^2(pi y / zdv), ph = 2 pi (i x / xdv + j y / ydv) + p, seeded
 * a1, a2 in [-1, 1), p in [-pi, pi), sc_m = 1 / (i^2 + j^2), over the modes
 * 0 <= i <= min(nmodes, (xdv-1)/2), |j| <= min(nmodes, (ydv-1)/2),
 * (i, j) > (0, 0): divergence-free and zero at the walls. */
typedef struct dns_init {
    int kind;      /* enum dnsgpu_init_kind */
    double uc;     /* centreline velocity (POISEUILLE) / friction velocity u_tau (TURBULENT) */
    double amp;    /* POISEUILLE: noise amplitude as a fraction of uc; TURBULENT: rms of |u'| over the
                      grid as a fraction of the log law's centreline velocity U(zdv/2) */
    uint64_t seed; /* counter-based RNG seed */
    int nmodes;    /* TURBULENT: Fourier modes per periodic direction, 1..31 */
} dns_init;

typedef struct dnsgpu dnsgpu;

/* Allocate the device state for a case.  Replaces the calloc block of
 * dns-const-q-xycut.c:239-356 (no halo buffers: the periodic x/y wrap is
 * done inside the kernel).  DNSGPU_EARG, with *out = NULL, for a grid
 * outside xdv, ydv >= 2, 4 <= zdv <= 1024, fewer than 2^31 nodes, or for
 * parameters the reference would run into NaN with: a relaxation rate
 * (given, or derived from Rebs and ub) outside (0, 2), CFL <= 0, a
 * non-finite ub or force, an unknown forcing or precision. */
int dnsgpu_create(const dns_case *c, dnsgpu **out);
void dnsgpu_destroy(dnsgpu *h);
const char *dnsgpu_last_error(void);

/* Wall masks [xdv][ydv] for the bottom (k=0) and top (k=zdv+1) rows with
 * the node codes of solid_init.c:3-8: -1 no-slip, -2 slip.  NULL means all
 * no-slip.  Replaces solid_init()/slip_ridges_init()/slip_posts_init()
 * (fun_defs.h:1-4) for the wall rows. */
int dnsgpu_set_walls(dnsgpu *h, const int8_t *bottom, const int8_t *top);

/* Solid bands inside the fluid rows (riblets, posts with side walls): per
 * column the number of solid fluid rows above the bottom wall and below the
 * top wall, [xdv][ydv] each, NULL = 0.  No reference counterpart (the
 * reference only has solids in the wall rows); rule in the step kernel.
 * Set them before the first step: right after create, an init or a load
 * (DNSGPU_EARG otherwise). */
int dnsgpu_set_solid_bands(dnsgpu *h, const int16_t *bottom, const int16_t *top);

/* Rest start, init.c:3-20 (the '-s' branch, dns-const-q-xycut.c:385-398). */
int dnsgpu_init_rest(dnsgpu *h);

/* Seeded initial fields (the build's generator; see dns_init). */
int dnsgpu_init_field(dnsgpu *h, const dns_init *spec);

/* Restart ('-p' branch, dns-const-q-xycut.c:399-454): upload x-planes
 * [x0, x0+nx) of both checkpoint halves.  ftemp is the state (post-stream
 * populations), f the previous post-collision populations, each
 * [nx][ydv][zdv+2][19] float64.  After the last slab call
 * dnsgpu_commit_load() derives the force of the first step. */
int dnsgpu_load_aos(dnsgpu *h, int x0, int nx, const double *ftemp, const double *f);
int dnsgpu_commit_load(dnsgpu *h);

/* Checkpoint (dns-const-q-xycut.c:491-503): download x-planes [x0, x0+nx)
 * of both halves in the same layout.  FP32/FP16 states are written as
 * float64 w + g. */
int dnsgpu_save_aos(dnsgpu *h, int x0, int nx, double *ftemp, double *f);

/* Current V.force (the force the next step will use); force.NNNN. */
int dnsgpu_get_force(dnsgpu *h, double *force);
/* Sets the force of a DNSGPU_CONST_F case from the next step on: the
 * reference's V.force as xcomputations takes it (xcomputations.c:29), for a
 * caller that computes the force itself (encalc.c:121 on its own host
 * copy).  DNSGPU_EARG for a constant-flow-rate case. */
int dnsgpu_set_force(dnsgpu *h, double force);

/* Advance nsteps time steps: for each step, encalc + wshearstr + one fused
 * collide-stream (dns-const-q-xycut.c:487-488,518).  cntr0 is the loop
 * counter `ss` of the first step (time label TM + dt*ss).  rows, if not
 * NULL, receives nsteps log rows; with rows == NULL the per-step plane
 * statistics are skipped (the force budget is always kept). */
int dnsgpu_step(dnsgpu *h, long cntr0, int nsteps, dns_steplog *rows);

/* statistics.c:3-184 of the current state: out[9][zdv+2] in the order
 * UAVG,VAVG,WAVG,UTIN,VTIN,WTIN,REYS,DAVG,DTIN (statistics.c:139-147). */
int dnsgpu_statistics(dnsgpu *h, double *out);

/* vel_print.c:65-99 of the current state for x-planes [x0,x0+nx):
 * vel[nx][ydv][zdv+2][3], den[nx][ydv][zdv+2] (zeros on wall rows). */
int dnsgpu_fields(dnsgpu *h, int x0, int nx, double *vel, double *den);

/* dnsgpu_fields without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid it).  No counterpart in the reference.  The staging lives in a
 * static array of DNSGPU_FIELDS_NOALLOC_MAX_DOUBLES doubles (64 MiB) per GPU,
 * held until the process ends and shared by every handle on that GPU: a call
 * takes nx x-planes with nx * ydv * (zdv+2) * 4 <= that number (DNSGPU_EARG
 * beyond it, nothing changed).  Calls on one GPU are serialised inside the
 * library; the array keeps the last call's contents. */
#define DNSGPU_FIELDS_NOALLOC_MAX_DOUBLES ((size_t)8 << 20)
int dnsgpu_fields_noalloc(dnsgpu *h, int x0, int nx, double *vel, double *den);

/* Time averages, FP64 accumulators on the device.  Every logged step
 * (dnsgpu_step with rows) adds its row; dnsgpu_tavg_sample adds one sample
 * of the statistics.c plane profiles of the current state and of the
 * wall-adjacent velocity over slip / no-slip wall columns.  The summary
 * follows the reference's post-processing: u_tau = sqrt(<p-grad>)
 * (TKE-budget/postprocessing/time_stats.c:101-114), Re_tau = u_tau h / nu
 * (time_stats.c:115-117), Cf = <tau_w>/(0.5 <Ub>^2) from w-str.txt and from
 * p-grad.txt (cfriction.c:15-86).  profiles (may be NULL) receives
 * [9][zdv+2] time-averaged profiles in the turb-fld order, the rms and
 * <u'w'> taken over planes and time. */
typedef struct dns_tavg {
    double nsteps;   /* logged steps accumulated */
    double nsamples; /* profile samples accumulated */
    double pgrad;    /* <F h> */
    double utau;     /* sqrt(<F h>) */
    double stress;   /* <tau_w> */
    double stup, stdown;
    double Ub;       /* <bulk velocity> (flowrate.txt / (ydv zdv)) */
    double K;        /* <K> */
    double rho;      /* <V.rho> */
    double Cf_w;     /* <tau_w> / (0.5 <Ub>^2) */
    double Cf_p;     /* <F h> / (0.5 <Ub>^2) */
    double Re_tau;   /* u_tau (zdv/2) / nu */
    double u_slip_bottom, u_noslip_bottom, u_slip_top, u_noslip_top; /* <u> of the wall-adjacent rows */
} dns_tavg;

int dnsgpu_tavg_reset(dnsgpu *h);
int dnsgpu_tavg_sample(dnsgpu *h);
int dnsgpu_tavg_get(dnsgpu *h, dns_tavg *out, double *profiles);

/* Kernel-only timing helpers for benchmarks: time nsteps steps on the
 * handle's stream with HIP events; ms_total receives the elapsed time of
 * the step kernels alone (sum over steps) and ms_wall the wall time of the
 * whole sequence. */
int dnsgpu_time_steps(dnsgpu *h, int nsteps, int with_stats, double *ms_step_kernels, double *ms_wall);

/* Choose the step kernel's launch shape by timing trial_steps launches of
 * each candidate on this handle's own state and masks: the waves per block
 * (each wave one column at a time; 2 to 4, as the per-wave plane sums fit
 * the LDS: the block shape), the blocks of the grid (2 to 32 waves per CU)
 * and, for FP16 storage with an even zdv, the rows per lane (1, or 2 rows
 * in one 4-byte access per direction: the cells per thread), then the
 * three fastest again, alternating.  Every shape gives the same state:
 * only the summation order of the logged sums depends on it.  The state is not
 * advanced: the trials read the current buffer and write the other one
 * without flipping them, so call it before a restart's first checkpoint is
 * written (the second checkpoint half of a freshly loaded state lives in
 * the other buffer).  Results are cached per (grid, precision, texture) in
 * the process and, when the environment variable DNSGPU_TUNE_CACHE names a
 * file, across processes.  No reference counterpart (the reference's
 * blocking is fixed, definitions.h:117).  best_blocks and best_ms may be
 * NULL. */
int dnsgpu_autotune(dnsgpu *h, int trial_steps, int *best_blocks, double *best_ms);

/* The step kernel's current launch shape: blocks of the grid, waves per
 * block and rows per lane (any pointer may be NULL).  No reference
 * counterpart. */
int dnsgpu_step_shape(const dnsgpu *h, int *blocks, int *waves_per_block, int *rows_per_lane);
/* Sets the launch shape by hand (blocks are capped at what is resident at
 * once and rounded down to a multiple of the step's z-pieces, see
 * dnsgpu_step_pieces); rows_per_lane 2 needs FP16 storage and an even zdv
 * up to 512, and takes one z-piece.  The state and the results do not depend on the
 * shape beyond the summation order of the logged sums.  No reference
 * counterpart. */
int dnsgpu_set_step_shape(dnsgpu *h, int blocks, int waves_per_block, int rows_per_lane);
/* The step kernel's z-pieces per column: each piece of a column goes to
 * its own group of blocks, whose waves keep the plane sums of that piece
 * only.  A column above 512 rows always has pieces of at most 512 rows; a
 * column of 321-384 or 449-512 rows may run as one piece or two (the
 * default: two for FP64 and FP16 storage, one for FP32; the autotuner times
 * both; DNSGPU_STEP_PIECES in the environment sets the default).  Setting
 * another count is DNSGPU_EARG, and so is more than one piece with two rows
 * per lane.  The grid is kept a multiple of the count.  The state does not
 * depend on it; the logged sums only by summation order.  No reference
 * counterpart. */
int dnsgpu_step_pieces(const dnsgpu *h, int *pieces);
int dnsgpu_set_step_pieces(dnsgpu *h, int pieces);

/* Device bytes held by the handle. */
size_t dnsgpu_device_bytes(const dnsgpu *h);

/* Placement of the two population buffers, chosen by dnsgpu_autotune.  On
 * MI355X device memory falls into two halves, and a stream that reads one
 * buffer while writing another at the same offsets runs about 6% faster
 * when the two lie in different halves; the step is such a stream, and the
 * two buffers of dnsgpu_create, allocated one after the other, often share
 * a half.  For a case of at least 2 GiB per buffer the autotuner (the
 * first time it runs on the handle, before it times the launch shapes, at
 * the cached or current shape) times the step into fresh destination buffers, holding each while the next is
 * tried (at most 4, as device memory allows, so it briefly holds more than
 * dnsgpu_footprint), and keeps the fastest when it beats the original pair
 * by 1.5%; the others are freed.  Only where the bytes live changes, never
 * a value; the buffer not holding the state is overwritten, as by the
 * autotuner's trials.  DNSGPU_PAIR_PLACEMENT=0 in the environment switches
 * it off.  Returns the destination buffers tried (0: none) and the step-time
 * gain of the kept pair (0 when the original pair was kept); either pointer
 * may be NULL.  No reference counterpart (the reference's populations live
 * in host memory). */
int dnsgpu_pair_placement(const dnsgpu *h, int *candidates, double *step_gain);

/* Pair arenas for many small cases (ensemble members).  Reserves on the
 * current device 8 GiB blocks, each classed by a timed copy probe into one
 * of the two halves of device memory, until each half has bytes_per_half
 * not held by live cases (as device memory allows; *reserved_per_half
 * receives the smaller of the two halves' free bytes, may be NULL).  While
 * they exist, dnsgpu_create takes the
 * two population buffers of a case below 2 GiB per buffer as ranges of a
 * block of each half (first fit, 2 MiB granules) and falls back to plain
 * allocations when they are full; a destroyed handle returns its ranges.
 * Only where the bytes live changes, never a value.  All probing happens
 * here, before any case steps: reserve before creating the members (the
 * request loop, dnsserve_sweep, does).  dnsgpu_release_pair_arenas frees
 * the empty blocks and each other block when its last range returns;
 * later creations allocate plainly.  No reference counterpart. */
int dnsgpu_reserve_pair_arenas(size_t bytes_per_half, size_t *reserved_per_half);
int dnsgpu_release_pair_arenas(void);
/* Blocks of each half held on `device` and the bytes of their ranges in use
 * (any pointer may be NULL). */
int dnsgpu_pair_arena_stats(int device, int *blocks_half0, int *blocks_half1, size_t *bytes_used);

/* Device bytes dnsgpu_create would allocate for this case (0 when the case
 * is refused), without allocating: the admission test of a request. */
size_t dnsgpu_footprint(const dns_case *c);

/* ------------------------------------------------------------ ensembles
 *
 * An ensemble advances many resident cases (members) of different grids,
 * textures and forcing together: per time step one launch of a multi-case
 * step kernel over a table of column tiles of every member, one launch of
 * the partial-sum reduction and one of the per-member finisher (log row,
 * next force), captured into a HIP graph per chunk of steps.  No
 * counterpart in the reference, which runs one case per MPI job; each
 * member computes exactly the reference's step (the same kernels' column
 * update), and its reductions run in an order fixed by the member alone,
 * so a member's results are the same alone and in any mix.
 *
 * A member stays a dnsgpu handle: its logs go to its time accumulators
 * (dnsgpu_tavg_*), and dnsgpu_save_aos, _statistics, _fields, _tavg_* work
 * on it between dnsens_step calls.  Entry points that change or advance
 * its state (dnsgpu_step, _load_aos, _init_*, _set_walls, ...) return
 * DNSGPU_EARG while it is a member; dnsgpu_destroy removes it first.
 */
typedef struct dnsens dnsens;

#define DNSENS_MAX_MEMBERS 256
#define DNSENS_NO_GRAPH 1 /* flags: launch the kernels directly instead of replaying a graph */
#define DNSENS_AUTOTUNE 2 /* flags: tune the step grid whenever the member set takes a new shape */

/* precision: DNSGPU_FP64 or DNSGPU_FP32, that of every member.
 * hbm_budget: admission limit on the sum of the members' device bytes
 * (0: 90% of the device memory free now). */
int dnsens_create(int precision, size_t hbm_budget, int flags, dnsens **out);
void dnsens_destroy(dnsens *e);

/* Device bytes a handle is charged as a member (its own plus the tile
 * records of its partial sums, about 2-4 bytes per node). */
size_t dnsens_member_bytes(const dnsgpu *h);
/* The same for a case not yet created (0 when the case is refused). */
size_t dnsens_footprint(const dns_case *c);

/* Admit handle h (initialised or loaded, same precision; any grid
 * dnsgpu_create takes, a member above 512 rows in z-pieces as in the
 * single-case step).
 * Returns DNSGPU_ENOMEM, with nothing changed, when its bytes do not fit
 * the budget, and DNSGPU_EARG when h is a member already (of this or of
 * another ensemble: a handle has one owner until dnsens_remove).  A state
 * fresh from an init or a load takes its first step here (one logged
 * step). */
int dnsens_add(dnsens *e, dnsgpu *h);
/* The handle leaves the ensemble (state as of the last dnsens_step). */
int dnsens_remove(dnsens *e, dnsgpu *h);
int dnsens_count(const dnsens *e);
size_t dnsens_bytes_used(const dnsens *e);
size_t dnsens_budget(const dnsens *e);
/* Makes the ensemble's device (the one current at dnsens_create) current on
 * the calling thread, so that handles this thread creates can join it:
 * dnsens_add refuses a member created on another device.  For a thread that
 * prepares members while another steps the ensemble. */
int dnsens_bind_thread(const dnsens *e);

/* Advance every member by nsteps logged steps; returns when they are done.
 * A member whose force or log row turned NaN is flagged in its info (and
 * keeps running until removed). */
int dnsens_step(dnsens *e, int nsteps);

typedef struct dnsens_member_info {
    double steps;       /* the member's step counter (counter of its next log row) */
    double n_rows;      /* rows in its time accumulators since dnsgpu_tavg_reset */
    double mean_pgrad;  /* running mean of p-grad over those rows */
    double force;       /* force of its next step */
    double last_stress; /* wall stress of its last step */
    int finite;         /* 0 once its force or a log row was not finite */
    int pad;
} dnsens_member_info;

/* State of member h: after the last dnsens_step, or as admitted when no
 * dnsens_step has run since its dnsens_add. */
int dnsens_info(dnsens *e, const dnsgpu *h, dnsens_member_info *out);

/* The last n log rows of member h from the last chunk of the last
 * dnsens_step (a chunk holds up to 1024 steps). */
int dnsens_rows(dnsens *e, const dnsgpu *h, dns_steplog *rows, int n);

/* Choose the step kernel's grid for the current member set by timing
 * trial_steps launches per candidate (state not advanced); cached per
 * (precision, largest zdv, tile count) in the process. */
int dnsens_autotune(dnsens *e, int trial_steps, int *best_blocks, double *best_ms);

/* Advance nsteps steps like dnsens_step and return the elapsed device time
 * (HIP events on the ensemble's stream) in ms. */
int dnsens_time_steps(dnsens *e, int nsteps, double *ms);

/* Advance nsteps steps with direct launches, timing each step kernel with
 * HIP events on the ensemble's stream: ms_step_kernels receives the sum of
 * the step-kernel times, ms_total the whole sequence (may be NULL). */
int dnsens_time_kernels(dnsens *e, int nsteps, double *ms_step_kernels, double *ms_total);

/* Current step-kernel grid (blocks) and tile count. */
int dnsens_grid(const dnsens *e, int *blocks, int *tiles);

#ifdef __cplusplus
}
#endif

#pragma pop_macro("Rebs")
#pragma pop_macro("ub")
#pragma pop_macro("xdv")
#pragma pop_macro("ydv")
#pragma pop_macro("zdv")
#pragma pop_macro("TSTR")
#pragma pop_macro("TEND")
#pragma pop_macro("CFL")
#endif
