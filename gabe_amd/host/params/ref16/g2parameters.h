/*
 * g2parameters.h -- GABE's shipped parameter file (the defaults of
 * params/default) with the edits a user makes for a short test run: a 16^3
 * box, 12 steps (endtime = 11 dt), a slice every 4 steps.  Used with the
 * default g2model.h by the gabe_ref16 test build.
 */

#define fftw_flag 0
#define gNum double

#define num_flds 2
const gNum mphi = 1.e-6;
const gNum phi0 = 0.193;
const gNum gsq = 2.5e-7;
const gNum f0[2] = {phi0, 0.};
const gNum df0[2] = {-0.142231, 0.};

#define parallelize 1
#define tot_num_thrds 4
#define calc_gws 1
const int randseed = 44463132;
const int N = 16;
const gNum L = 10.;
const gNum starttime = 0.;
const gNum dt = L / (gNum)N / 80.;
const gNum endtime = 11. * dt;
#define expansion_type 1

const gNum rescale_B = mphi;

#define rand_init 1
#define field_full_rand 1

const gNum screentime = 60;
const int slicewait = 4;
const int slicenumber = 3;          /* unused: slicewait is set */
const int field_sliceskip = 2;
const int specnumber = 1;
#define field_outdim 2
#define spec_output 1
#define var_output 1

const int nflds = num_flds;
const gNum dx = L / ((gNum)N);
const gNum gridsize = N * N * N;
