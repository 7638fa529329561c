/*
 * gabe_oracle.h -- CPU FP64 restatement of GABE's lattice evolution.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it,
 * and only as the checker.  The product (libgabe_amd.so) never links or calls it.
 *
 * Each function restates one reference routine (file:line cited at its
 * definition in gabe_oracle.c).  The state mirrors the reference's globals
 * (g2header.h:52-63,171-174): field/dfield are [2][nflds][N][N][N] (slot 0 =
 * step start, slot 1 = RK2 midpoint), h/l/T_gw are [2][12][N][N][N/2+1] with
 * planes (11,12,13,22,23,33) x (Re,Im).
 *
 * The reference links FFTW 3 (absent here).  The oracle carries its own
 * mixed-radix complex FFT; r2c is the unnormalised forward e^{-ikx} transform
 * keeping kz in [0,N/2], c2r is the unnormalised inverse whose kz=0 and kz=N/2
 * planes are read as the Hermitian part of the input (the complex-inverse /
 * real-part definition).
 */
#ifndef GABE_ORACLE_H
#define GABE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int N;              /* points per side (g2parameters.h:57) */
    int nflds;          /* number of fields, 1 or 2 (g2parameters.h:43,92) */
    double L;           /* box side, program units (g2parameters.h:58) */
    double dt;          /* time step (g2parameters.h:61) */
    int calc_gws;       /* 1: evolve tensor perturbations (g2parameters.h:55) */
    int expansion_type; /* 0 none, 1 self-consistent, 2 user (g2parameters.h:62) */
    /* Polynomial model in program units:
     * V = m2_phi phi^2/2 + lambda phi^4/4 + g2 phi^2 chi^2/2 + m2_chi chi^2/2.
     * The reference default (g2model.cpp:66-69) is m2_phi=1, g2=gsq/mphi^2. */
    double m2_phi, lambda, g2, m2_chi;
    double rescale_B;   /* amplitude scale of the vacuum ICs (g2parameters.h:69) */
    int nthreads;       /* OpenMP threads (g2parameters.h:54) */
} orc_params;

typedef struct orc_state {
    orc_params p;
    double dx;
    long N3;            /* N^3 sites */
    long M;             /* N*N*(N/2+1) half-spectrum modes */
    double *field;      /* [2][nflds][N^3] */
    double *dfield;     /* [2][nflds][N^3] */
    double a[2], adot[2], edkin[2], edpot[2], edgrad[2], edrho[2];
    double *h, *l, *T_gw; /* [2][12][M] each, NULL when calc_gws == 0 */
    double t;
} orc_state;

orc_state *orc_create(const orc_params *p);
void orc_destroy(orc_state *st);
int orc_numbins(int N);

/* lattice operators (g2functions.cpp) */
double orc_laplacian(const orc_state *st, int s, int fld, int i, int j, int k);
double orc_potential(const orc_state *st, int s, int i, int j, int k);
double orc_dVdf(const orc_state *st, int s, int fld, int i, int j, int k);
double orc_effMass(const orc_state *st, int s, int fld);
double orc_avgGrad(const orc_state *st, int s);
double orc_avgPot(const orc_state *st, int s);
double orc_avgKin(const orc_state *st, int s);
void orc_calcEnergy(orc_state *st, int s);
double orc_adf(const orc_state *st, int s);
double orc_ddfield(const orc_state *st, int s, int fld, int i, int j, int k);
void orc_step(orc_state *st);
double orc_gradF2(const orc_state *st, int s, int fld, int i, int j, int k);
double orc_calcrho(const orc_state *st, int s, int i, int j, int k);

/* tensor perturbations (g2evolutionpert.cpp) */
double orc_stresstensor(const orc_state *st, int aa, int bb, int s, int i, int j, int k);
void orc_fft_stresstensor(orc_state *st, int s);
void orc_gw_update(orc_state *st, int s, int snew, double deet);
void orc_evolve_perts(orc_state *st, int s, int snew, double deet);

/* spectra (g2spectra.cpp).  out is [numspec_out][numbins] followed by the
 * trailing GW column p[numbins] when calc_gws; gw_quirk=1 reproduces the
 * reference's kz=0 GW term read from the kz=N/2 plane. */
void orc_specout(const orc_state *st, double *out, int gw_quirk);

/* initial conditions (g2init.cpp, g2model.cpp) */
void orc_srand(unsigned seed);
int orc_rand(void);
void orc_initfields_homog(orc_state *st, const double *f0, const double *df0);
void orc_initexpansion(orc_state *st);
void orc_randinit(orc_state *st, int fld, double d2vdf2, double *fk_out, double *fkd_out);
void orc_randinit_opt(orc_state *st, int fld, double d2vdf2, int full_rand, double cut_off, double smooth,
                      double *fk_out, double *fkd_out);
void orc_initfields_fluct(orc_state *st);

/* diagnostics (g2output.cpp) */
void orc_meansvars(const orc_state *st, double *out);

/* FFT used everywhere above: in/out are row-major with the last index fastest */
void orc_r2c3d(int N, const double *in, double *out_cplx, int nthreads);
void orc_c2r3d(int N, const double *in_cplx, double *out, int nthreads);
void orc_fft1d(int n, const double *in_cplx, double *out_cplx, int sign);

#ifdef __cplusplus
}
#endif
#endif
