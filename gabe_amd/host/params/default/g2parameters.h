/*
 * g2parameters.h -- parameters of the kept GABE host program, in the shape of
 * GABE's own parameter file (g2parameters.h:17-100): the same names, the same
 * spelling (#define for the switches, const for the numbers), so a GABE
 * user's edited file drops in here unchanged.  This one holds GABE's shipped
 * defaults: V = m^2 phi^2/2 + g^2 phi^2 chi^2/2 with m = 1e-6 and
 * g^2 = 2.5e-7, phi0 = 0.193, a 64^3 box of side 10 stepped with dt = dx/80
 * to t = 150, tensor modes on.
 *
 * The model itself (how these constants enter V) lives in g2model.h beside
 * this file, the counterpart of the user's g2model.cpp.
 *
 * Not used by this build: parallelize and tot_num_thrds (the lattice runs on
 * the GPU; tot_num_thrds is only printed to info.txt), fftw_flag 1
 * (long double builds are refused).
 */

#define fftw_flag 0
#define gNum double

/* model constants */
#define num_flds 2
const gNum mphi = 1.e-6;            /* mass of phi */
const gNum phi0 = 0.193;            /* initial mean of phi */
const gNum gsq = 2.5e-7;            /* g^2 of the phi-chi coupling */
const gNum f0[2] = {phi0, 0.};      /* initial means of phi and chi */
const gNum df0[2] = {-0.142231, 0.};/* initial mean velocities */

/* model independent parameters */
#define parallelize 1
#define tot_num_thrds 4
#define calc_gws 1                  /* 1: evolve the tensor perturbations */
const int randseed = 44463132;
const int N = 64;                   /* points per side */
const gNum L = 10.;                 /* box side, program units */
const gNum starttime = 0.;
const gNum endtime = 150.;
const gNum dt = L / (gNum)N / 80.;
#define expansion_type 1            /* 0 none, 1 from adot (self-consistent), 2 user defined */

/* model dependent parameters */
const gNum rescale_B = mphi;

#define rand_init 1                 /* 1: vacuum fluctuations on top of the means */
#define field_full_rand 1           /* 1: random modes, 0: symmetric k-space modes */
/* #define spec_cut_off sqrt(3.)/8.    leave undefined for no window */
/* #define spec_smooth 0.5             sharpness of the tanh window */

/* output parameters */
const gNum screentime = 60;         /* seconds between screen updates */
const int slicewait = 0;            /* steps between slices; 0: use slicenumber */
const int slicenumber = (int)endtime;
const int field_sliceskip = 2;
const int specnumber = 1;
#define field_outdim 2              /* 0 none, 1 a line, 2 the plane i = 0, 3 the lattice */
#define spec_output 1
#define var_output 1

/* derived */
const int nflds = num_flds;
const gNum dx = L / ((gNum)N);
const gNum gridsize = N * N * N;
