/*
 * g2parameters.h -- the benchmark model in the shape of GABE's parameter file:
 * V = lambda phi^4/4 + g^2 phi^2 chi^2/2 with lambda = 9e-14, g^2/lambda = 100,
 * phi0 = 0.342 M_pl, in units of B = sqrt(lambda) phi0; L = 20, dt = dx/20.
 * The numbers the test builds vary can be set on the compiler line
 * (-DGABE_N=16 -DGABE_CALC_GWS=1 -DGABE_STEPS=12 ...).
 */
#include <cmath>

#define fftw_flag 0
#define gNum double

/* model constants */
#define num_flds 2
const gNum lambda_phys = 9.e-14;          /* lambda (M_pl units) */
const gNum g2_over_lambda = 100.;         /* g^2 / lambda */
const gNum phi0 = 0.342;                  /* initial mean of phi */
const gNum f0[2] = {phi0, 0.};
/* slow roll in program units: dphi/dt = -V'/(3H), H = sqrt(8 pi V / 3),
 * V_pr = lambda_pr phi^4 / 4 with lambda_pr = 1/phi0^2 */
const gNum lambda_pr = 1. / (phi0 * phi0);
const gNum df0[2] = {-(lambda_pr * phi0 * phi0 * phi0) /
                         (3. * std::sqrt(8. * M_PI / 3. * 0.25 * lambda_pr * phi0 * phi0 * phi0 * phi0)),
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
#define parallelize 1
#define tot_num_thrds 1
#define calc_gws GABE_CALC_GWS
const int randseed = 44463132;
const int N = GABE_N;
const gNum L = 20.;
const gNum starttime = 0.;
const gNum dt = L / (gNum)N / 20.;
const gNum endtime = (GABE_STEPS - 1) * dt;   /* the loop runs floor(endtime/dt)+1 steps */
#define expansion_type 1

/* model dependent parameters */
const gNum rescale_B = std::sqrt(lambda_phys) * phi0;

#define rand_init 1
#ifndef GABE_FIELD_FULL_RAND
#define GABE_FIELD_FULL_RAND 1
#endif
#define field_full_rand GABE_FIELD_FULL_RAND
#ifdef GABE_SPEC_CUT_OFF
#define spec_cut_off GABE_SPEC_CUT_OFF
#define spec_smooth 0.5
#endif

/* output parameters */
const gNum screentime = 60;
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
#define field_outdim GABE_FIELD_OUTDIM
#define spec_output 1
#define var_output 1

const int nflds = num_flds;
const gNum dx = L / ((gNum)N);
const gNum gridsize = N * N * N;
