/*
 * g2parameters.h -- compile-time parameters of the kept GABE host program.
 *
 * Same names and meaning as the reference's parameter file
 * (g2parameters.h:43-100), so a GABE user edits this file the same way.  The
 * model is the polynomial V = m2_phi phi^2/2 + lambda phi^4/4 + g2 phi^2 chi^2/2
 * + m2_chi chi^2/2 in program units (gabe_model.h); the coefficients below
 * are derived from the physical constants.  Every value can be overridden on
 * the compiler line (-DGABE_N=16 ...), which the test build uses.  The
 * switches the reference spells as #define (calc_gws, expansion_type, ...)
 * are constants here: the ABI structs use the same names as fields.
 *
 * Default: lambda phi^4/4 + g^2 phi^2 chi^2/2 with lambda = 9e-14,
 * g^2/lambda = 100, phi0 = 0.342 M_pl in units of B = sqrt(lambda) phi0,
 * 64^3, L = 20, dt = dx/20, 200 steps, no tensor modes.
 */
#ifndef G2PARAMETERS_H
#define G2PARAMETERS_H

#include <cmath>

/* model constants */
#define num_flds 2
const double lambda_phys = 9.e-14;                  /* lambda (M_pl units) */
const double g2_over_lambda = 100.;                 /* g^2 / lambda */
const double phi0 = 0.342;                          /* initial homogeneous phi */
const double rescale_B = std::sqrt(lambda_phys) * phi0;
/* polynomial coefficients in program units (V_pr = V / B^2) */
const double model_m2_phi = 0.;
const double model_lambda = 1. / (phi0 * phi0);
const double model_g2 = g2_over_lambda / (phi0 * phi0);
const double model_m2_chi = 0.;
const double f0[2] = {phi0, 0.};
/* slow roll: dphi/dt = -V'/(3H), H = sqrt(8 pi V / 3) */
const double df0[2] = {-(model_lambda * phi0 * phi0 * phi0) /
                           (3. * std::sqrt(8. * M_PI / 3. * 0.25 * model_lambda * phi0 * phi0 * phi0 * phi0)),
                       0.};

/* model independent parameters */
#ifndef GABE_N
#define GABE_N 64
#endif
#ifndef GABE_CALC_GWS
#define GABE_CALC_GWS 0
#endif
#ifndef GABE_STEPS
#define GABE_STEPS 200
#endif
#ifndef GABE_PRECISION
#define GABE_PRECISION 1        /* 0: FP64 arithmetic of the reference, 1: FP32 fluctuation form */
#endif
#define tot_num_thrds 1         /* host threads (the lattice runs on the GPU) */
const int calc_gws = GABE_CALC_GWS;
const int randseed = 44463132;
const int N = GABE_N;
const double L = 20.;
const double starttime = 0.;
const double dt = L / (double)N / 20.;
const double endtime = (GABE_STEPS - 1) * dt;      /* the loop runs floor(endtime/dt)+1 steps */
const int expansion_type = 1;

const int rand_init = 1;
#ifndef GABE_FIELD_FULL_RAND
#define GABE_FIELD_FULL_RAND 1
#endif
#ifndef GABE_SPEC_CUT_OFF
#define GABE_SPEC_CUT_OFF 0.     /* the reference leaves spec_cut_off undefined: no window */
#endif
const int field_full_rand = GABE_FIELD_FULL_RAND;   /* 0: symmetric k-space modes, needs spec_cut_off */
const double spec_cut_off = GABE_SPEC_CUT_OFF;      /* window centre in units of N, e.g. sqrt(3.)/8. */
const double spec_smooth = 0.5;                     /* sharpness of the tanh window */

/* output parameters */
const double screentime = 60;
#ifndef GABE_SLICEWAIT
#define GABE_SLICEWAIT 0
#endif
#ifndef GABE_SLICENUMBER
#define GABE_SLICENUMBER 20
#endif
const int slicewait = GABE_SLICEWAIT;
const int slicenumber = GABE_SLICENUMBER;
const int field_sliceskip = 2;
const int specnumber = 1;
#ifndef GABE_FIELD_OUTDIM
#define GABE_FIELD_OUTDIM 2
#endif
const int field_outdim = GABE_FIELD_OUTDIM;         /* 1: a line, 2: the plane i = 0, 3: the lattice */
const int spec_output = 1;
const int var_output = 1;

/* checkpoint.bin every this many steps (0: never); not in the reference */
#ifndef GABE_CHECKPOINT_EVERY
#define GABE_CHECKPOINT_EVERY 0
#endif
const int checkpoint_every = GABE_CHECKPOINT_EVERY;

const int nflds = num_flds;
const double dx = L / ((double)N);
const double gridsize = (double)N * N * N;

#endif
