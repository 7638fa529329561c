/*
 * gabe_host.h -- host-side pieces of the kept GABE program that sit above the
 * C ABI: the vacuum initial conditions and the initialisation sequence.
 */
#ifndef GABE_HOST_H
#define GABE_HOST_H

#include "gabe.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gabe_host_init {
    double f0[GABE_MAX_FLDS];   /* homogeneous field values      (g2parameters.h:47) */
    double df0[GABE_MAX_FLDS];  /* homogeneous field velocities  (g2parameters.h:48) */
    double rescale_B;           /* IC amplitude scale            (g2parameters.h:69) */
    unsigned randseed;          /* srand seed                    (g2parameters.h:56) */
    int rand_init;              /* 1: add vacuum fluctuations    (g2parameters.h:71) */
    /* Appended options; zero keeps the reference defaults. */
    int symmetric_modes;        /* 1: field_full_rand 0 (g2parameters.h:72), deterministic real
                                 * modes with no rand() draws; needs spec_cut_off */
    double spec_cut_off;        /* tanh window centre in units of N (g2parameters.h:73);
                                 * <= 0 leaves it undefined (no window) */
    double spec_smooth;         /* window sharpness              (g2parameters.h:74) */
} gabe_host_init;

/* Mode options of g2init.cpp:111-165: full_rand 1 draws, full_rand 0 the
 * deterministic real modes; cut_off > 0 applies the tanh window. */
typedef struct gabe_host_modes {
    int full_rand;
    double cut_off;
    double smooth;
} gabe_host_modes;

/* Draw the half-spectrum vacuum modes of one field with libc rand() in the
 * sequential loop order of g2init.cpp:88-172 (field_full_rand 1).  fk and fkd
 * receive [N][N][N/2+1] (re,im) pairs.  The caller seeds rand(). */
int gabe_host_draw_modes(int N, double L, double rescale_B, double d2vdf2, double H0, double *fk, double *fkd);

/* Number of rand() values one field consumes: 4 per mode except k = 0. */
long gabe_host_draws_per_field(int N);

/* The same modes from pre-drawn rand() values (draws[gabe_host_draws_per_field(N)]). */
int gabe_host_modes_from_draws(int N, double L, double rescale_B, double d2vdf2, double H0, const int *draws,
                               double *fk, double *fkd);

/* The modes under the options of g2init.cpp:111-165 (draws may be NULL when
 * opt->full_rand is 0, which consumes none). */
int gabe_host_modes_opt(int N, double L, double rescale_B, double d2vdf2, double H0, const gabe_host_modes *opt,
                        const int *draws, double *fk, double *fkd);

/* The start-up sequence of g2main.cpp:109-119: homogeneous fields and
 * calcEnergy(0); initexpansion(); srand(randseed) and per field (in order)
 * effMass -> modes -> transform-and-add; calcEnergy(0).  adot[0] keeps its
 * homogeneous value, as in the reference. */
int gabe_host_initialize(gabe_ctx *ctx, const gabe_params *p, const gabe_host_init *init);

/* Where a run stands, saved beside the lattice in a checkpoint file: the
 * loop time t (g2main.cpp:106), steps taken, the rand() seed and the number
 * of values drawn since srand (all draws happen at start-up, g2init.cpp:84),
 * and the slice writer's counters (g2output.cpp:129-186). */
typedef struct gabe_host_run {
    double t;
    long long step;
    unsigned randseed;
    long long rand_draws;
    int slice_first, slice_last, spec_last, slice_wait;
} gabe_host_run;

/* Write the run record and gabe_checkpoint_save's blob to path (through a
 * temporary file renamed into place, so a crash leaves the previous file). */
int gabe_host_checkpoint_write(gabe_ctx *ctx, const char *path, const gabe_host_run *run);

/* Read a file written by gabe_host_checkpoint_write into ctx and *run. */
int gabe_host_checkpoint_read(gabe_ctx *ctx, const char *path, gabe_host_run *run);

/* Friedmann-constraint residual from a per-step scalar history (rows of
 * gabe_history: t, a, adot, edkin, edpot, edgrad, edrho, step).  For each
 * interior row n: R = |rho_dot + 3 H (rho + p)| / (3 H rho), with
 * p = edkin - edgrad/3 - edpot, H = adot/a and rho_dot the central difference
 * (rho[n+1] - rho[n-1]) / (t[n+1] - t[n-1]).  out[n] = R (0 for the first and
 * last rows).  Returns the maximum over interior rows. */
double gabe_host_friedmann_residual(const double *hist, int nrows, double *out);

#ifdef __cplusplus
}
#endif
#endif
