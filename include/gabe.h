/*
 * gabe.h -- C ABI of the MI355X-native lattice evolution (libgabe_amd.so).
 *
 * The reference program (GABE 2.x) keeps its lattice state in globals and
 * advances it through internal C++ calls (g2header.h:52-63,127-177).  This
 * header is the boundary that replaces those calls: the host program keeps
 * its driver loop, initialisation and writers, and calls these entry points
 * instead of step(), calcEnergy(), fft_stresstensor()/evolve_perts() and the
 * computational part of specOut().
 *
 * Conventions
 *   - plain C, no exceptions cross the boundary; every function returns an
 *     int status (GABE_OK = 0, negative on error) and gabe_last_error()
 *     returns the message of the most recent failure on the calling thread;
 *   - the caller owns host buffers, the library owns device memory;
 *   - one host thread per context, one HIP stream per context; calls on one
 *     context are not re-entrant;
 *   - gabe_step() is asynchronous (stream-ordered); every call that returns
 *     host data synchronises the stream first;
 *   - host arrays use the reference's slot-0 layouts: fields are
 *     [nflds][N][N][N] doubles with the last index fastest (g2header.h:52-53);
 *     tensor perturbations are [12][N][N][N/2+1] doubles, planes
 *     (11,12,13,22,23,33) x (Re,Im) (g2header.h:172-174).
 */
#ifndef GABE_H
#define GABE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GABE_OK 0
#define GABE_ERR_ARG (-1)       /* invalid argument or parameter */
#define GABE_ERR_HIP (-2)       /* HIP runtime failure (allocation, launch) */
#define GABE_ERR_FFT (-3)       /* rocFFT plan or execution failure */
#define GABE_ERR_NONFINITE (-4) /* energy density became NaN/Inf: unstable run */
#define GABE_ERR_STATE (-5)     /* call not valid in the current state */

#define GABE_FP64 0             /* fields stored and stepped in FP64 (oracle form) */
#define GABE_FP32 1             /* FP64 mean + FP32 fluctuation per field */

#define GABE_MAX_FLDS 2

typedef struct gabe_ctx gabe_ctx;

/* Runtime mirror of the g2parameters.h surface (g2parameters.h:43-100) and of
 * the model constants used by g2model.cpp:59-108. */
typedef struct gabe_params {
    int N;               /* points per side, N >= 4, even           (N)            */
    int nflds;           /* 1 or 2                                  (num_flds)     */
    double L;            /* box side in program units               (L)            */
    double dt;           /* time step                               (dt)           */
    int calc_gws;        /* 1: evolve tensor perturbations          (calc_gws)     */
    int expansion_type;  /* 0 none, 1 self-consistent, 2 user-defined            */
    int precision;       /* GABE_FP64 or GABE_FP32                                 */
    /* model, program units: V = m2_phi phi^2/2 + lambda phi^4/4
     *                           + g2 phi^2 chi^2/2 + m2_chi chi^2/2
     * (the reference default g2model.cpp:66-87 is m2_phi = 1,
     *  g2 = gsq/mphi^2, lambda = m2_chi = 0)                                      */
    double m2_phi, lambda, g2, m2_chi;
    int gw_spectrum_quirk; /* 1: GW spectrum kz=0 term read from the kz=N/2 plane,
                              as g2spectra.cpp:149-164 does; 0: read kz=0         */
    int planes_per_block;  /* sweep tiling: lattice planes each workgroup marches
                              through (0 = autotuned or library default)           */
    int transform_tile;    /* j-transform of the stress spectra: columns per
                              workgroup, 16 or 32 (0 = autotuned or 16)            */
    int mode_tile;         /* tile variant of the fused i-transform + mode kernel,
                              1..gabe_mode_tiles(N, precision) (0 = autotuned or
                              variant 1, the built-in default)                     */
    int stress_chunk;      /* planes per launch when sweep A and the j-transform
                              of the stress spectra are interleaved (pipelined
                              FP32 steps with tensor modes; rounded down to whole
                              planes_per_block): each launch's spectra are
                              transformed along j while still in the memory-side
                              cache.  0 = autotuned or not interleaved, -1 = not
                              interleaved                                          */
    int reserved[3];
} gabe_params;

/* The reference's expansion/energy globals (g2header.h:59-64) plus bookkeeping. */
typedef struct gabe_scalars {
    double a[2], adot[2], edkin[2], edpot[2], edgrad[2], edrho[2];
    double t;             /* simulation time of the current state */
    long long nstep;      /* steps taken since creation */
    int adot0_valid;      /* 1: adot[0] belongs to the current state (set on write) */
    int error;            /* nonzero once a non-finite energy density was seen */
} gabe_scalars;

/* Allocate device state for one lattice (replaces the allocations of
 * g2main.cpp:49-99).  Fields, h and l start at zero. */
int gabe_create(const gabe_params *params, gabe_ctx **out);
int gabe_destroy(gabe_ctx *ctx);
const char *gabe_last_error(void);

/* Number of spectral bins, (int)(sqrt(3) N/2 + 1) (g2spectra.cpp:31). */
int gabe_numbins(int N);

/* Slot-0 field state, host layout [nflds][N][N][N] (g2main.cpp:24-26).
 * With GABE_FP32 the upload splits each field into its FP64 mean and an FP32
 * fluctuation; the download returns mean + fluctuation. */
int gabe_upload_state(gabe_ctx *ctx, const double *field, const double *dfield);
int gabe_download_state(gabe_ctx *ctx, double *field, double *dfield);

/* Plane i of the slot-0 state, layout [nflds][N][N] each (either pointer may
 * be NULL): what the slice writer of g2output.cpp:25-68 reads for
 * field_outdim 1 and 2, and the i+-1, i+-2 planes the 4th-order gradients of
 * calcrho (g2functions.cpp:70-122,176-186) need.  0 <= i < N. */
int gabe_download_plane(gabe_ctx *ctx, int i, double *field, double *dfield);

/* Mean and variance of each slot-0 field, out[2*fld] = <f>, out[2*fld+1] =
 * <f^2> - <f>^2 (g2output.cpp:70-121), reduced on the device in FP64 with the
 * first site's value as the shift of a one-pass sum. */
int gabe_means_vars(gabe_ctx *ctx, double *out);

/* gabe_means_vars on fields in caller-owned device memory, allocating
 * nothing (no device or host heap memory, no library routine that
 * allocates).  No counterpart in the reference; it is for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  field holds nflds arrays of nsites values each,
 * back to back, as doubles (GABE_FP64) or floats (GABE_FP32); base[fld]
 * (NULL: 0) is added to each mean (the FP64 background of a fluctuation
 * array).  out[2*fld] = mean, out[2*fld+1] = variance, the shifted one-pass
 * reduction of gabe_means_vars.  stream is a hipStream_t (NULL: the null
 * stream); the call synchronises it.  1 <= nsites <= GABE_NOALLOC_MAX_SITES.
 * The partial sums use a fixed 16 KiB device array inside the library, one
 * copy per GPU, shared by every call on that GPU: calls that overlap on one
 * GPU race. */
#define GABE_NOALLOC_MAX_SITES (1LL << 33)
int gabe_means_vars_noalloc(const void *field, int precision, long long nsites, int nflds, const double *base,
                            void *stream, double *out);

/* Checkpoint and resume (no counterpart in the reference, which has none).
 * The blob holds the parameters that fix the dynamics, the device scalars
 * (a, adot, energies, FP64 means, t, step count), the history rows of the
 * current trajectory, the state (fields, velocities and, with tensor modes,
 * h and l) in the device's storage precision and, when a learned closure is
 * attached, that attachment (on/off gate, the step it switched off at, its
 * largest residual, threshold, scales and weights), so a context restored
 * from it continues bit for bit.  gabe_checkpoint_size gives the bytes
 * gabe_checkpoint_save writes; gabe_checkpoint_load accepts only a blob from
 * a context with the same N, nflds, precision, calc_gws, expansion_type, L,
 * dt, model coefficients and tensor-mode storage order, re-attaches the
 * blob's closure (or detaches one the blob does not carry). */
int gabe_checkpoint_size(gabe_ctx *ctx, size_t *bytes);
int gabe_checkpoint_save(gabe_ctx *ctx, void *buf, size_t bytes);
int gabe_checkpoint_load(gabe_ctx *ctx, const void *buf, size_t bytes);

/* Slot-0 tensor perturbations h and l = dh/dt, layout [12][N][N][N/2+1]
 * (g2main.cpp:30-35).  Either pointer may be NULL.  The device keeps each mode
 * as its two transverse-traceless polarisation amplitudes: the reference's
 * zero start and TT source (g2evolutionpert.cpp:215-321) never leave that
 * span, so an upload stores the TT part of h and l (k = 0 mode zero) and a
 * download rebuilds the six components from it. */
int gabe_upload_gw(gabe_ctx *ctx, const double *h, const double *l);
int gabe_download_gw(gabe_ctx *ctx, double *h, double *l);

/* Expansion and energy globals.  gabe_set_scalars marks adot[0] as belonging
 * to the current state (the first step then uses it, as step() does after
 * initexpansion(), g2init.cpp:26-29) and starts a new trajectory: the
 * history (gabe_history) holds only rows recorded after it.  gabe_get_scalars returns the reference's
 * post-step values: a[0], and edkin/edpot/edgrad/edrho[0] and adot[0] = adf(0)
 * of the current state (g2functions.cpp:268-270), computing them if stale. */
int gabe_set_scalars(gabe_ctx *ctx, const gabe_scalars *s);
int gabe_get_scalars(gabe_ctx *ctx, gabe_scalars *s);

/* calcEnergy(0) (g2functions.cpp:168-174): energies of the current state into
 * slot 0; adot[0] is left unchanged.  Only s = 0 is valid between steps. */
int gabe_energy(gabe_ctx *ctx, int s);

/* nsteps midpoint-RK2 steps of fields, tensor modes and a(t)
 * (step(), g2functions.cpp:201-304).  Asynchronous. */
int gabe_step(gabe_ctx *ctx, int nsteps);

/* Wait for all queued work; reports GABE_ERR_NONFINITE if the energy density
 * became non-finite. */
int gabe_sync(gabe_ctx *ctx);

/* Computational part of specOut (g2spectra.cpp:23-197).  out receives
 * [ncols][numbins] doubles: P and k^3 P/(2 pi^2) per field, then, with GW,
 * the GW power, the Omega_gw column and the trailing k column, i.e.
 * ncols = 2*nflds (+3 with GW).  The host keeps the writer. */
int gabe_spectra(gabe_ctx *ctx, double *out);

/* randInit's transform-and-add (g2init.cpp:174-187): fk and fkd are the
 * half-spectrum modes [N][N][N/2+1] as (re,im) pairs, as drawn by the host's
 * sequential rand() loop (g2init.cpp:88-172).  The kz=0 and kz=N/2 planes are
 * used by their Hermitian part.  f += c2r(fk), fdot += c2r(fkd), unnormalised. */
int gabe_add_fluctuation(gabe_ctx *ctx, int fld, const double *fk, const double *fkd);

/* effMass(0, fld) (g2model.cpp:89-108): lattice mean of d2V/dfld2. */
int gabe_effmass(gabe_ctx *ctx, int fld, double *out);

/* The TT-projected source T_gw[s] of the current state as fft_stresstensor(s)
 * computes it (g2evolutionpert.cpp:135-339): s = 0 uses field[0], s = 1 the
 * midpoint field[0] + dt/2 dfield[0].  out is [12][N][N][N/2+1]. */
int gabe_tt_source(gabe_ctx *ctx, int s, double *out);

/* Per-step scalar history recorded on the device at each step's start:
 * rows of (t, a, adot, edkin, edpot, edgrad, edrho, step), for the steps
 * since the last gabe_set_scalars or restriction (a checkpoint load restores
 * the blob's rows), at most the last 65536.  Returns the number of rows written to out (at most
 * max_rows, most recent last). */
int gabe_history(gabe_ctx *ctx, double *out, int max_rows);

/* Timing hook: runs nsteps steps with HIP events around every kernel on the
 * context's stream and writes the mean duration in ms of each kernel class to
 * ms_out[0..GABE_NKCLASS).  Classes: sweep A, finalize, stress FFT, GW modes,
 * sweep B. */
#define GABE_NKCLASS 5
int gabe_kernel_times(gabe_ctx *ctx, int nsteps, double *ms_out);

/* Effective parameters of a context, with the sweep tiling actually used
 * in planes_per_block.  No reference counterpart (introspection). */
int gabe_query(gabe_ctx *ctx, gabe_params *out);

/* Tiling autotuner: times a few steps of a synthetic lattice of these
 * parameters for each candidate planes_per_block; with tensor modes on the
 * fused transform path it then times, at the best planes_per_block, every
 * j-transform tile width and mode-kernel tile variant, and then the
 * stress_chunk candidates (not interleaved, N/8, N/4, N/2).  The fastest choice is
 * remembered for (N, nflds, precision, calc_gws); later contexts created with
 * 0 in those fields use it (gabe_query reports what a context uses).  Either
 * output pointer may be NULL; ms_per_step is the best step time.  No
 * reference counterpart. */
int gabe_autotune_sweep(const gabe_params *params, int *planes_per_block, double *ms_per_step);
/* Number of mode-kernel tile variants for this lattice side and precision
 * (1 where only the default exists, 0 where the fused kernel does not run). */
int gabe_mode_tiles(int N, int precision);

/* Ensembles: independent lattices (sizes, couplings, seeds and a(t) of their
 * own) advanced together, each RK2 phase of every member in one launch and
 * two steps per captured HIP graph.  Members are ordinary contexts for every
 * other call (upload, initialise, scalars, spectra, download) and give
 * bit-identical results to the same lattice stepped alone.  All members share
 * precision and nflds; tensor modes are not batched (calc_gws must be 0).
 * Members created with planes_per_block = 0 march min(N, 32) planes per
 * workgroup (the batched launch fills the GPU without smaller tiles).
 * No reference counterpart: the reference runs one lattice per process. */
typedef struct gabe_ensemble gabe_ensemble;
int gabe_ensemble_create(const gabe_params *params, int count, gabe_ensemble **out);
int gabe_ensemble_member(gabe_ensemble *ens, int idx, gabe_ctx **out);
int gabe_ensemble_step(gabe_ensemble *ens, int nsteps);
int gabe_ensemble_sync(gabe_ensemble *ens);
/* Timing hook: nsteps ensemble steps launched one by one (not from the graph)
 * with HIP events around each batched phase on the ensemble's stream; writes
 * the mean ms of the batched sweep A, finalize and sweep B to ms_out[0..3). */
int gabe_ensemble_kernel_times(gabe_ensemble *ens, int nsteps, double *ms_out);
int gabe_ensemble_destroy(gabe_ensemble *ens);

/* ------------------------------------------------------------------------
 * Learned closure of a coarse lattice (no counterpart in the reference).
 *
 * A 3-layer 3D convolutional network (3^3 periodic kernels, 4 -> 16 -> 16 -> 2
 * channels, SiLU), BF16 weights and activations with FP32 accumulation on the
 * matrix cores.  Input per site: the fluctuations (dphi, dchi, dphidot,
 * dchidot), each divided by its input scale; the network's 2-channel output
 * u enters through its lattice bi-Laplacian y = lap lap u (lap u = sum of
 * the 6 neighbours minus 6 u), the form of the Laplacian's own truncation
 * error, so the correction falls like (k dx)^4 at long wavelengths.  Attached to a lattice it adds
 * dt * s_out[c] * y_c / a^2 to the velocity fluctuation of field c after
 * every RK2 step.  Trained (Adam) on pairs made
 * by the plain method: the state of a fine lattice spectrally restricted onto
 * the coarse one, and as target what the coarse lattice's Laplacian misses of
 * the fine one, R(lap_fine f) - lap_coarse(R f) with lap the 7-point
 * Laplacian over dx^2 and R the restriction (stored, for the loss, as the
 * target of lap u: lap_unit^-1 of it over s_out).  Lattices with a closure: FP32,
 * two fields, N in {32, 64, 96, 128}.
 *
 * Parameters, FP32, GABE_CLOSURE_NPARAMS of them in this order:
 *   W1[16][4][27], b1[16], W2[16][16][27], b2[16], W3[2][16][27], b3[2]
 * with tap t = 9 (dz+1) + 3 (dy+1) + (dx+1), dz along the slowest index.
 * Loss: sum over sites s and channels c of (lap u - t)^2 / (2 N^3), t the
 * target of lap u: the force target carried through lap^-1, so each
 * wavelength weighs by the relative error of the frequency correction it
 * gets rather than by the size of the force.  (b3 drops out of lap u: its
 * gradient is zero.)
 * ---------------------------------------------------------------------- */
#define GABE_CLOSURE_NPARAMS 9538
typedef struct gabe_closure gabe_closure;

/* Spectral restriction: the coarse context's state becomes the fine one's
 * Fourier modes with |n_a| < N_coarse/2 (the coarse Nyquist planes zeroed),
 * with the fine context's a, adot, t and FP64 means.  Same nflds, precision
 * and L; the coarse N even and not larger. */
int gabe_restrict(gabe_ctx *fine, gabe_ctx *coarse);

/* A network for lattices of size N stepped with dt, weights drawn from seed. */
int gabe_closure_create(int N, double dt, unsigned seed, gabe_closure **out);
int gabe_closure_destroy(gabe_closure *cl);
int gabe_closure_params(gabe_closure *cl, float *out);
int gabe_closure_set_params(gabe_closure *cl, const float *in);   /* also resets Adam */
/* Input scales s_in[4] and output scales s_out[2]: set by the first sample
 * (rms of the inputs and of the raw targets) or by gabe_closure_set_scales. */
int gabe_closure_scales(gabe_closure *cl, double *s_in, double *s_out);
int gabe_closure_set_scales(gabe_closure *cl, const double *s_in, const double *s_out);

/* One training pair from the plain method.  With restrict_first the coarse
 * context first takes the restriction of the fine one; the input is the
 * coarse state, the target R(lap_fine f) - lap_coarse(R f) per field (over
 * s_out, carried through the unit stencil's inverse: the target of lap u).  Then
 * the fine context takes fine_steps plain steps (0: none) and the coarse
 * context the restriction of the advanced state, ready for the next call with
 * restrict_first = 0.  Returns the sample index. */
int gabe_closure_sample(gabe_closure *cl, gabe_ctx *fine, gabe_ctx *coarse, int fine_steps, int restrict_first);
int gabe_closure_num_samples(gabe_closure *cl);

/* iters Adam steps (learning rate lr), one whole sample per step, samples in
 * a shuffled order drawn from seed; loss_out[iters] (optional) gets the loss
 * of each step before its update. */
int gabe_closure_train(gabe_closure *cl, int iters, double lr, unsigned seed, double *loss_out);
/* loss on a stored sample */
int gabe_closure_loss(gabe_closure *cl, int sample, double *loss);
/* Network on host data in scaled units: x[N^3][4] -> y = lap lap u, [N^3][2]. */
int gabe_closure_forward(gabe_closure *cl, const float *x, float *y);
/* Loss and its gradient with respect to the parameters for host data
 * x[N^3][4] and targets t[2][N^3] of lap u (the kernels training uses). */
int gabe_closure_gradient(gabe_closure *cl, const float *x, const float *t, float *grad, double *loss);

/* Apply the closure after every step of ctx (weights and scales are copied:
 * cl may be destroyed afterwards).  Fallback: after each step the Friedmann
 * residual R = |rho_dot + 3H(rho + p)| / (3 H rho) of the per-step history
 * (central difference, one step back) is checked on the device; once R
 * exceeds residual_threshold (or is not finite) the closure stays off and the
 * lattice continues with the plain step. */
int gabe_closure_attach(gabe_ctx *ctx, gabe_closure *cl, double residual_threshold);
int gabe_closure_detach(gabe_ctx *ctx);
/* active: 1 while the closure is applied; fallback_step: step count when it
 * switched off (-1 if never); max_residual: largest R checked so far. */
int gabe_closure_status(gabe_ctx *ctx, int *active, long long *fallback_step, double *max_residual);
/* Timing hook for the attached closure (no reference counterpart): runs
 * nsteps steps one kernel after another, as gabe_kernel_times does, with HIP
 * events around each closure kernel on the lattice's stream, and writes the
 * mean ms per step to ms_out[0..GABE_NCLCLASS): first convolution (4 -> 16
 * channels), hidden convolution (16 -> 16), output convolution (16 -> 2),
 * the bi-Laplacian kick.  GABE_ERR_STATE without an attached closure. */
#define GABE_NCLCLASS 4
int gabe_closure_kernel_times(gabe_ctx *ctx, int nsteps, double *ms_out);

#ifdef __cplusplus
}
#endif
#endif
