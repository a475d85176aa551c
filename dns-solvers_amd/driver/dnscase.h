/*
 * dnscase.h -- case inputs of the host driver.
 *
 * The reference takes its case as compile-time macros in definitions.h
 * (definitions.h:30-59: Rebs, ub, zdv, ydv, xdv, TSTR, TEND, CFL and the
 * optional ITER, PSLIP, RSLIP).  The driver reads the same keys at run time
 * from a text case file ("key = value", '#' comments), plus its own keys:
 *
 *   forcing    = constq | constf      (default constq, encalc.c:121)
 *   force      = <double>             body force for constf
 *   precision  = fp64 | fp32 | fp16   (default fp64; fp32/fp16: shifted storage)
 *   omega      = <double>             1/tau; default from Rebs, ub, zdv
 *   texture    = smooth | posts | stripes | riblets | blocks  (PSLIP / RSLIP set posts / stripes)
 *   riblet_spacing, riblet_height, riblet_offset   (triangular riblets, cells)
 *   block_x, block_y, block_height   (blocks: solid posts with side walls, footprint
 *                                     block_x x block_y every period_x x period_y)
 *   period_x, period_y, slip_x, slip_y, offset_x, offset_y   (texture geometry)
 *   walls      = both | bottom | top  (which walls carry the texture)
 *   init       = rest | poiseuille | turbulent   (for -s; rest is init.c)
 *   init_uc, init_amp, init_seed, init_modes
 *   tavg_start = <step>               start of the time averages (default: off)
 *   tavg_every = <steps>              profile sampling cadence (default utime)
 *   outdir     = .                    (where every output file goes)
 */
#ifndef DNSCASE_H
#define DNSCASE_H

#include <stdint.h>

typedef struct dns_run_case {
    /* definitions.h */
    double Rebs, ub, CFL;
    int xdv, ydv, zdv;
    int TSTR, TEND;
    long ITER; /* < 0: not defined */
    /* build keys */
    int forcing;   /* 0 constant Q, 1 constant F */
    double force;
    int precision; /* 0 fp64, 1 fp32, 2 fp16 (shifted storage) */
    double omega;  /* 0 => derived */
    int texture;   /* 0 smooth, 1 posts, 2 stripes, 3 riblets, 4 blocks */
    double riblet_spacing, riblet_height, riblet_offset;
    int walls;     /* 0 both, 1 bottom, 2 top */
    int period_x, period_y, slip_x, slip_y, offset_x, offset_y;
    int init;      /* 0 rest, 1 poiseuille, 2 turbulent */
    double init_uc, init_amp;
    uint64_t init_seed;
    int init_modes;
    int block_x, block_y, block_height;
    long tavg_start; /* < 0: no time averages */
    int tavg_every;
    char outdir[512];
} dns_run_case;

/* Defaults: the values of definitions.h:30-37 as shipped. */
void dns_case_defaults(dns_run_case *c);

/* Parse a case file; returns 0 or -1 with a message in err. */
int dns_case_read(const char *path, dns_run_case *c, char *err, int errlen);

/* Wall masks [xdv][ydv] (codes -1 no-slip, -2 slip) of the case's texture.
 * posts   : slip where (x - offset_x) mod period_x < slip_x and
 *           (y - offset_y) mod period_y < slip_y; the reference's PSLIP
 *           pattern (slip_posts_init.c:14-41) is period 4, slip 2, offset 1.
 * stripes : streamwise-invariant slip stripes, (y - offset_y) mod period_y
 *           < slip_y; the reference's RSLIP pattern (slip_ridges_init.c:11-27)
 *           is period 4, slip 2, offset 1. */
void dns_case_masks(const dns_run_case *c, int8_t *bottom, int8_t *top);

/* Solid bands [xdv][ydv] of triangular riblets (texture 3): the cell
 * (y, z) is solid when its centre z + 1/2 lies below the triangle of
 * spanwise period riblet_spacing and tip height riblet_height at y + 1/2.
 * Streamwise-invariant.  Blocks (texture 4): block_height solid rows over
 * the footprint (x - offset_x) mod period_x < block_x and
 * (y - offset_y) mod period_y < block_y.  Zero for every other texture. */
void dns_case_bands(const dns_run_case *c, int16_t *bottom, int16_t *top);

#endif
