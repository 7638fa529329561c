/*
 * gabe_oracle.c -- CPU FP64 restatement of GABE's lattice evolution.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every routine cites the reference routine it restates (paths relative to
 * the reference tree).  Deliberate differences from the reference:
 *   - reductions are summed per i-plane and then over planes in index order, so
 *     the result does not depend on the thread count (the reference's OpenMP
 *     reduction order does, g2functions.cpp:129-162);
 *   - the FFT is this file's own (FFTW 3 is not available); the first-call
 *     FFTW_MEASURE clobbering of the T11 input (g2evolutionpert.cpp:183-187)
 *     is not reproduced;
 *   - the vacuum ICs draw rand() in the sequential loop order, i.e. the
 *     reference run with one thread (the OpenMP loop of g2init.cpp:88-172).
 */
#include "gabe_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define NN (st->p.N)
#define FLD(st, s, f) ((st)->field + ((long)(s) * (st)->p.nflds + (f)) * (st)->N3)
#define DFLD(st, s, f) ((st)->dfield + ((long)(s) * (st)->p.nflds + (f)) * (st)->N3)
#define GWP(arr, st, s, c) ((arr) + ((long)(s) * 12 + (c)) * (st)->M)

static inline long sidx(int N, int i, int j, int k) { return ((long)i * N + j) * N + k; }
static inline long midx(int N, int i, int j, int k) { return ((long)i * N + j) * (N / 2 + 1) + k; }
/* periodic neighbours, g2functions.cpp:22-30 / g2evolutionpert.cpp:16-25 */
static inline int incr(int N, int i) { return i == N - 1 ? 0 : i + 1; }
static inline int decr(int N, int i) { return i == 0 ? N - 1 : i - 1; }
static inline double pw2(double x) { return x * x; } /* g2functions.cpp:17-20 */

static int nthr(const orc_state *st) { return st->p.nthreads > 0 ? st->p.nthreads : 1; }

int orc_numbins(int N) { return (int)(sqrt(3.) * N / 2 + 1); } /* g2spectra.cpp:31 */

orc_state *orc_create(const orc_params *p)
{
    orc_state *st = (orc_state *)calloc(1, sizeof(orc_state));
    st->p = *p;
    st->dx = p->L / (double)p->N;                      /* g2parameters.h:93 */
    st->N3 = (long)p->N * p->N * p->N;
    st->M = (long)p->N * p->N * (p->N / 2 + 1);
    st->field = (double *)calloc((size_t)2 * p->nflds * st->N3, sizeof(double));
    st->dfield = (double *)calloc((size_t)2 * p->nflds * st->N3, sizeof(double));
    if (p->calc_gws) {                                  /* g2main.cpp:53-95: zeroed h, l */
        st->h = (double *)calloc((size_t)2 * 12 * st->M, sizeof(double));
        st->l = (double *)calloc((size_t)2 * 12 * st->M, sizeof(double));
        st->T_gw = (double *)calloc((size_t)2 * 12 * st->M, sizeof(double));
    }
    return st;
}

void orc_destroy(orc_state *st)
{
    if (!st) return;
    free(st->field); free(st->dfield); free(st->h); free(st->l); free(st->T_gw);
    free(st);
}

/* ------------------------------------------------------------------ model */
/* Pointwise polynomial model; the reference's plug-in is g2model.cpp:59-87.
 * The term order keeps the reference default (m2_phi=1, lambda=0,
 * g2=gsq/mphi^2, m2_chi=0) bit-identical to its expressions. */
static double model_V(const orc_params *p, double phi, double chi)
{
    return 0.5 * p->m2_phi * phi * phi + 0.5 * p->g2 * phi * phi * chi * chi
         + 0.25 * p->lambda * phi * phi * phi * phi + 0.5 * p->m2_chi * chi * chi;
}
static double model_dV(const orc_params *p, int fld, double phi, double chi)
{
    if (fld == 0) return p->m2_phi * phi + p->g2 * chi * chi * phi + p->lambda * phi * phi * phi;
    if (fld == 1) return p->g2 * phi * phi * chi + p->m2_chi * chi;
    return 0.;
}
static double site_val(const orc_state *st, int s, int fld, long n)
{
    return fld < st->p.nflds ? FLD(st, s, fld)[n] : 0.;
}

double orc_potential(const orc_state *st, int s, int i, int j, int k) /* g2model.cpp:66-69 */
{
    long n = sidx(NN, i, j, k);
    return model_V(&st->p, site_val(st, s, 0, n), site_val(st, s, 1, n));
}

double orc_dVdf(const orc_state *st, int s, int fld, int i, int j, int k) /* g2model.cpp:72-87 */
{
    long n = sidx(NN, i, j, k);
    return model_dV(&st->p, fld, site_val(st, s, 0, n), site_val(st, s, 1, n));
}

double orc_effMass(const orc_state *st, int s, int fld) /* g2model.cpp:89-108: <d2V/df2> */
{
    double avemass = 0.;
    for (long n = 0; n < st->N3; n++) {
        double phi = site_val(st, s, 0, n), chi = site_val(st, s, 1, n);
        if (fld == 0) avemass += (st->p.m2_phi + st->p.g2 * chi * chi + 3. * st->p.lambda * phi * phi);
        else if (fld == 1) avemass += (st->p.g2 * phi * phi + st->p.m2_chi);
        else return 1.;
    }
    return avemass / (double)st->N3;
}

/* --------------------------------------------------------------- operators */
double orc_laplacian(const orc_state *st, int s, int fld, int i, int j, int k) /* g2functions.cpp:35-43 */
{
    const double *f = FLD(st, s, fld);
    int N = NN;
    double dx = st->dx;
    return (f[sidx(N, incr(N, i), j, k)] + f[sidx(N, decr(N, i), j, k)]
          + f[sidx(N, i, incr(N, j), k)] + f[sidx(N, i, decr(N, j), k)]
          + f[sidx(N, i, j, incr(N, k))] + f[sidx(N, i, j, decr(N, k))]
          - 6. * f[sidx(N, i, j, k)]) / (dx * dx);
}

/* 4th-order first derivatives, g2functions.cpp:70-98 */
static double dfd(const orc_state *st, const double *f, int dir, int i, int j, int k)
{
    int N = NN;
    int c[3] = {i, j, k};
    int m2[3] = {i, j, k}, m1[3] = {i, j, k}, p1[3] = {i, j, k}, p2[3] = {i, j, k};
    m1[dir] = decr(N, c[dir]); m2[dir] = decr(N, m1[dir]);
    p1[dir] = incr(N, c[dir]); p2[dir] = incr(N, p1[dir]);
    return ((1. / 12.) * f[sidx(N, m2[0], m2[1], m2[2])]
          + ((-2.) / 3.) * f[sidx(N, m1[0], m1[1], m1[2])]
          + (2. / 3.) * f[sidx(N, p1[0], p1[1], p1[2])]
          + ((-1.) / 12.) * f[sidx(N, p2[0], p2[1], p2[2])]) / st->dx;
}

double orc_gradF2(const orc_state *st, int s, int fld, int i, int j, int k) /* g2functions.cpp:118-122 */
{
    const double *f = FLD(st, s, fld);
    double a = dfd(st, f, 0, i, j, k), b = dfd(st, f, 1, i, j, k), c = dfd(st, f, 2, i, j, k);
    return a * a + b * b + c * c;
}

double orc_calcrho(const orc_state *st, int s, int i, int j, int k) /* g2functions.cpp:176-186 */
{
    double rho = orc_potential(st, s, i, j, k);
    long n = sidx(NN, i, j, k);
    for (int fld = 0; fld < st->p.nflds; fld++)
        rho += (.5 * DFLD(st, s, fld)[n] * DFLD(st, s, fld)[n]
              + .5 * orc_gradF2(st, 0, fld, i, j, k) / st->a[s] / st->a[s]); /* always field[0], Q12 */
    return rho;
}

/* -------------------------------------------------------------- energies */
double orc_avgGrad(const orc_state *st, int s) /* g2functions.cpp:124-136 */
{
    int N = NN;
    double *part = (double *)calloc(N, sizeof(double));
    double grad = 0;
    for (int fld = 0; fld < st->p.nflds; fld++) {
        const double *f = FLD(st, s, fld);
#pragma omp parallel for num_threads(nthr(st))
        for (int i = 0; i < N; i++) {
            double acc = 0;
            for (int j = 0; j < N; j++)
                for (int k = 0; k < N; k++)
                    acc -= f[sidx(N, i, j, k)] * orc_laplacian(st, s, fld, i, j, k);
            part[i] = acc;
        }
        for (int i = 0; i < N; i++) grad += part[i];
    }
    free(part);
    return grad / (double)st->N3 / 2. / st->a[s] / st->a[s];
}

double orc_avgPot(const orc_state *st, int s) /* g2functions.cpp:139-150 */
{
    int N = NN;
    double *part = (double *)calloc(N, sizeof(double));
    double pot = 0;
#pragma omp parallel for num_threads(nthr(st))
    for (int i = 0; i < N; i++) {
        double acc = 0;
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) acc += orc_potential(st, s, i, j, k);
        part[i] = acc;
    }
    for (int i = 0; i < N; i++) pot += part[i];
    free(part);
    return pot / (double)st->N3;
}

double orc_avgKin(const orc_state *st, int s) /* g2functions.cpp:153-165 */
{
    int N = NN;
    long N2 = (long)N * N;
    double *part = (double *)calloc(N, sizeof(double));
    double kin = 0;
    for (int fld = 0; fld < st->p.nflds; fld++) {
        const double *df = DFLD(st, s, fld);
#pragma omp parallel for num_threads(nthr(st))
        for (int i = 0; i < N; i++) {
            double acc = 0;
            for (long n = (long)i * N2; n < (long)(i + 1) * N2; n++) acc += df[n] * df[n];
            part[i] = acc;
        }
        for (int i = 0; i < N; i++) kin += part[i];
    }
    free(part);
    return kin / (double)st->N3 / 2.;
}

void orc_calcEnergy(orc_state *st, int s) /* g2functions.cpp:168-174 */
{
    st->edkin[s] = orc_avgKin(st, s);
    st->edpot[s] = orc_avgPot(st, s);
    st->edgrad[s] = orc_avgGrad(st, s);
    st->edrho[s] = st->edkin[s] + st->edpot[s] + st->edgrad[s];
}

double orc_adf(const orc_state *st, int s) /* g2functions.cpp:188-191 */
{
    return sqrt(8. * M_PI / 3. * st->edrho[s]) * st->a[s];
}

double orc_ddfield(const orc_state *st, int s, int fld, int i, int j, int k) /* g2functions.cpp:193-196 */
{
    return orc_laplacian(st, s, fld, i, j, k) / st->a[s] / st->a[s] - orc_dVdf(st, s, fld, i, j, k)
         - 3 * st->adot[s] / st->a[s] * DFLD(st, s, fld)[sidx(NN, i, j, k)];
}

/* one RK2 half: dst = src0 + c*dsrc, d(dst) = d(src0) + c*ddfield(s) */
static void rk_stage(orc_state *st, int s_rhs, int s_dst, double c)
{
    int N = NN;
    for (int fld = 0; fld < st->p.nflds; fld++) {
        double *f0 = FLD(st, 0, fld), *df0 = DFLD(st, 0, fld);
        double *fd = FLD(st, s_dst, fld), *dfd_ = DFLD(st, s_dst, fld);
        const double *dfr = DFLD(st, s_rhs, fld);
#pragma omp parallel for num_threads(nthr(st))
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++)
                for (int k = 0; k < N; k++) {
                    long n = sidx(N, i, j, k);
                    double acc = orc_ddfield(st, s_rhs, fld, i, j, k);
                    fd[n] = f0[n] + c * dfr[n];
                    dfd_[n] = df0[n] + c * acc;
                }
    }
}

void orc_step(orc_state *st) /* g2functions.cpp:201-304 */
{
    double dt = st->p.dt;
    if (st->p.expansion_type == 0) {             /* :207-232 */
        rk_stage(st, 0, 1, .5 * dt);
        rk_stage(st, 1, 0, dt);
        orc_calcEnergy(st, 0);
    } else {                                     /* :234-270 type 1, :272-301 type 2 */
        rk_stage(st, 0, 1, .5 * dt);
        if (st->p.expansion_type == 1 && st->p.calc_gws) orc_evolve_perts(st, 0, 1, 0.5 * dt);
        st->a[1] = st->a[0] + .5 * dt * st->adot[0];
        orc_calcEnergy(st, 1);
        st->adot[1] = orc_adf(st, 1);
        rk_stage(st, 1, 0, dt);
        if (st->p.expansion_type == 1 && st->p.calc_gws) orc_evolve_perts(st, 1, 0, dt);
        st->a[0] = st->a[0] + dt * st->adot[1];
        orc_calcEnergy(st, 0);
        st->adot[0] = orc_adf(st, 0);
    }
    st->t += dt;
}

/* ---------------------------------------------------------------- tensors */
static double delfld(const orc_state *st, int fld, int s, int i, int j, int k, int n) /* g2evolutionpert.cpp:31-44 */
{
    const double *f = FLD(st, s, fld);
    int N = NN;
    if (n == 1) return (f[sidx(N, incr(N, i), j, k)] - f[sidx(N, decr(N, i), j, k)]) / 2;
    if (n == 2) return (f[sidx(N, i, incr(N, j), k)] - f[sidx(N, i, decr(N, j), k)]) / 2;
    if (n == 3) return (f[sidx(N, i, j, incr(N, k))] - f[sidx(N, i, j, decr(N, k))]) / 2;
    return 0.;
}

double orc_stresstensor(const orc_state *st, int aa, int bb, int s, int i, int j, int k) /* g2evolutionpert.cpp:48-132 */
{
    double deln2 = 1 / pw2(st->dx);
    double out = 0.0;
    for (int m = 0; m < st->p.nflds; m++) {
        double d1 = delfld(st, m, s, i, j, k, 1), d2 = delfld(st, m, s, i, j, k, 2),
               d3 = delfld(st, m, s, i, j, k, 3);
        if (aa == 1 && bb == 1)
            out += (2. / 3.) * deln2 * d1 * d1 - (1. / 3.) * deln2 * d2 * d2 - (1. / 3.) * deln2 * d3 * d3;
        else if (aa == 1 && bb == 2) out += deln2 * d1 * d2;
        else if (aa == 1 && bb == 3) out += (deln2 * d1 * d3);
        else if (aa == 2 && bb == 2)
            out += -(1. / 3.) * deln2 * d1 * d1 + (2. / 3.) * deln2 * d2 * d2 - (1. / 3.) * deln2 * d3 * d3;
        else if (aa == 2 && bb == 3) out += (deln2 * d2 * d3);
        else if (aa == 3 && bb == 3)
            out += -(1. / 3.) * deln2 * d1 * d1 - (1. / 3.) * deln2 * d2 * d2 + (2. / 3.) * deln2 * d3 * d3;
        else return 0.0;
    }
    return out;
}

/* TT projection of the six stress transforms at one mode, one part (Re or
 * Im); the coefficients are those of g2evolutionpert.cpp:215-310.
 * S = {S11,S12,S13,S22,S23,S33}; T = {T11,T12,T13,T22,T23,T33}. */
static void tt_reference(double x, double y, double z, double nm, const double S[6], double T[6])
{
    T[0] = .5 * pw2(1 - x * x) * nm * S[0] - (1 - x * x) * x * y * nm * S[1] - (1 - x * x) * x * z * nm * S[2]
         + pw2(x * y) * nm * S[3] - .5 * (1 - x * x) * (1 - y * y) * nm * S[3]
         + (1 + x * x) * y * z * nm * S[4] + pw2(x * z) * nm * S[5] - .5 * (1 - x * x) * (1 - z * z) * nm * S[5];
    T[3] = .5 * pw2(1 - y * y) * nm * S[3] - (1 - y * y) * x * y * nm * S[1] - (1 - y * y) * y * z * nm * S[4]
         + pw2(x * y) * nm * S[0] - .5 * (1 - x * x) * (1 - y * y) * nm * S[0]
         + (1 + y * y) * x * z * nm * S[2] + pw2(y * z) * nm * S[5] - .5 * (1 - y * y) * (1 - z * z) * nm * S[5];
    T[5] = .5 * pw2(1 - z * z) * nm * S[5] - (1 - z * z) * x * z * nm * S[2] - (1 - z * z) * y * z * nm * S[4]
         + pw2(x * z) * nm * S[0] - .5 * (1 - x * x) * (1 - z * z) * nm * S[0]
         + (1 + z * z) * x * y * nm * S[1] + pw2(y * z) * nm * S[3] - .5 * (1 - y * y) * (1 - z * z) * nm * S[3];
    T[1] = -.5 * x * y * (1 - x * x) * nm * S[0] - .5 * x * y * (1 - y * y) * nm * S[3]
         + .5 * x * y * (1 + z * z) * nm * S[5] + (1 - x * x) * (1 - y * y) * nm * S[1]
         - y * z * (1 - x * x) * nm * S[2] - x * z * (1 - y * y) * nm * S[4];
    T[2] = -.5 * x * z * (1 - x * x) * nm * S[0] + .5 * x * z * (1 + y * y) * nm * S[3]
         - .5 * x * z * (1 - z * z) * nm * S[5] + (1 - x * x) * (1 - z * z) * nm * S[2]
         - y * z * (1 - x * x) * nm * S[1] - x * y * (1 - z * z) * nm * S[4];
    T[4] = +.5 * y * z * (1 + x * x) * nm * S[0] - .5 * y * z * (1 - y * y) * nm * S[3]
         - .5 * y * z * (1 - z * z) * nm * S[5] + (1 - y * y) * (1 - z * z) * nm * S[4]
         - x * y * (1 - z * z) * nm * S[2] - x * z * (1 - y * y) * nm * S[1];
}

void orc_fft_stresstensor(orc_state *st, int s) /* g2evolutionpert.cpp:135-339 */
{
    int N = NN;
    long N3 = st->N3, M = st->M;
    double *tij = (double *)malloc(sizeof(double) * 6 * N3);
    double *tout = (double *)malloc(sizeof(double) * 2 * 6 * M);
    static const int AB[6][2] = {{1, 1}, {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}};
#pragma omp parallel for num_threads(nthr(st))
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++)
                for (int c = 0; c < 6; c++)
                    tij[c * N3 + sidx(N, i, j, k)] = orc_stresstensor(st, AB[c][0], AB[c][1], s, i, j, k);
    for (int c = 0; c < 6; c++) orc_r2c3d(N, tij + c * N3, tout + 2 * c * M, nthr(st));

    double normmm = pow(st->dx, 3);
    /* plane index of component c (T11,T12,T13,T22,T23,T33) for Re: */
    static const int plane[6] = {0, 2, 4, 6, 8, 10};
#pragma omp parallel for num_threads(nthr(st))
    for (int i = 0; i < N; i++) {
        int px = (i <= N / 2 ? i : i - N);
        for (int j = 0; j < N; j++) {
            int py = (j <= N / 2 ? j : j - N);
            for (int k = 0; k < N / 2 + 1; k++) {
                int pz = k;
                double momentum2 = pw2((double)px) + pw2((double)py) + pw2((double)pz);
                double dpz = ((double)pz) / sqrt(momentum2);
                double dpx = ((double)px) / sqrt(momentum2);
                double dpy = ((double)py) / sqrt(momentum2);
                long m = midx(N, i, j, k);
                for (int part = 0; part < 2; part++) {
                    double S[6], T[6];
                    for (int c = 0; c < 6; c++) S[c] = tout[2 * (c * M + m) + part];
                    tt_reference(dpx, dpy, dpz, normmm, S, T);
                    for (int c = 0; c < 6; c++) GWP(st->T_gw, st, s, plane[c] + part)[m] = T[c];
                }
            }
        }
    }
    for (int c = 0; c < 12; c++) GWP(st->T_gw, st, s, c)[0] = 0; /* :319-321, Q11 */
    free(tij);
    free(tout);
}

/* per-mode RK2 update given T_gw[s] (g2evolutionpert.cpp:351-438 with derivs :459-516) */
void orc_gw_update(orc_state *st, int s, int snew, double deet)
{
    int N = NN;
    double L = st->p.L;
#pragma omp parallel for num_threads(nthr(st))
    for (int i = 0; i < N; i++) {
        int ii = (i <= N / 2 ? i : i - N);
        for (int j = 0; j < N; j++) {
            int jj = (j <= N / 2 ? j : j - N);
            for (int k = 0; k < N / 2 + 1; k++) {
                long m = midx(N, i, j, k);
                double norm = (4. * M_PI * M_PI / L / L / st->a[s] / st->a[s]);
                double omega = (8. * M_PI) / st->a[s] / st->a[s];
                double kk = (pw2((double)ii) + pw2((double)jj) + pw2((double)k)) * norm;
                double hd[12], ld[12];
                for (int c = 0; c < 12; c++) {
                    double hv = GWP(st->h, st, s, c)[m], lv = GWP(st->l, st, s, c)[m];
                    ld[c] = -kk * hv - 3. * st->adot[s] / st->a[s] * lv + 2. * omega * GWP(st->T_gw, st, s, c)[m];
                    hd[c] = lv;
                }
                for (int c = 0; c < 12; c++) {
                    GWP(st->h, st, snew, c)[m] = GWP(st->h, st, 0, c)[m] + deet * hd[c];
                    GWP(st->l, st, snew, c)[m] = GWP(st->l, st, 0, c)[m] + deet * ld[c];
                }
            }
        }
    }
}

void orc_evolve_perts(orc_state *st, int s, int snew, double deet) /* g2evolutionpert.cpp:343-455 */
{
    orc_fft_stresstensor(st, s);
    orc_gw_update(st, s, snew, deet);
}

/* ---------------------------------------------------------------- spectra */
void orc_specout(const orc_state *st, double *out, int gw_quirk) /* g2spectra.cpp:23-234 */
{
    int N = NN, nflds = st->p.nflds;
    double L = st->p.L;
    long M = st->M;
    const double spec_norm = pow(L / (double)N, 6.) * pow(2. * M_PI / L, 3.);
    const int numbins = orc_numbins(N);
    int numspec_out = st->p.calc_gws ? 2 * nflds + 2 : 2 * nflds;
    double *spec_power = (double *)calloc(numbins, sizeof(double));
    double *spec_numpts = (double *)calloc(numbins, sizeof(double));
    double *outdft = (double *)malloc(sizeof(double) * 2 * M);

    for (int fld = 0; fld < nflds; fld++) {
        orc_r2c3d(N, FLD(st, 0, fld), outdft, nthr(st));
        for (int b = 0; b < numbins; b++) { spec_power[b] = 0.0; spec_numpts[b] = 0; }
        for (int i = 0; i < N; i++) {
            int px = (i <= N / 2 ? i : i - N);
            for (int j = 0; j < N; j++) {
                int py = (j <= N / 2 ? j : j - N);
                for (int k = 1; k < N / 2; k++) {
                    double pm = sqrt(pw2(px) + pw2(py) + pw2(k));
                    long m = midx(N, i, j, k);
                    spec_power[(int)pm] += 2. * (pw2(outdft[2 * m]) + pw2(outdft[2 * m + 1]));
                    spec_numpts[(int)pm] += 2;
                }
                int kk[2] = {0, N / 2};
                for (int e = 0; e < 2; e++) {
                    double pm = sqrt(pw2(px) + pw2(py) + pw2(kk[e]));
                    long m = midx(N, i, j, kk[e]);
                    spec_power[(int)pm] += (pw2(outdft[2 * m]) + pw2(outdft[2 * m + 1]));
                    spec_numpts[(int)pm] += 1;
                }
            }
        }
        for (int b = 0; b < numbins; b++) {
            if (spec_numpts[b] > 0) {
                out[(2 * fld) * numbins + b] = spec_power[b] * spec_norm / ((double)spec_numpts[b]);
                out[(2 * fld + 1) * numbins + b] = pow((double)b * 2. * M_PI / L, 3) * spec_power[b] * spec_norm
                                                 / ((double)spec_numpts[b]) / 2. / M_PI / M_PI;
            } else {
                out[(2 * fld) * numbins + b] = 0.;
                out[(2 * fld + 1) * numbins + b] = 0.;
            }
        }
    }

    if (st->p.calc_gws) {
        double *f2 = (double *)calloc(numbins, sizeof(double));
        int *np = (int *)calloc(numbins, sizeof(int));
        double hubtemp = sqrt(8. * M_PI / 3.) * sqrt(st->edrho[0]);
        double dp_gw = 2. * M_PI / (double)L;
        double norm1_gw = 4.e-5 / pow(100., .333) / 24. / M_PI / M_PI / pow(L, 3.) / pw2(hubtemp);
        const double *l0 = st->l; /* slot 0 */
        for (int i = 0; i < N; i++) {
            int px = (i <= N / 2 ? i : i - N);
            for (int j = 0; j < N; j++) {
                int py = (j <= N / 2 ? j : j - N);
                for (int kz = 0; kz <= N / 2; kz++) {
                    /* :130-181: interior planes weight 2, kz=0 and N/2 weight 1;
                     * the reference's kz=0 term reads plane N/2 (Q1) */
                    int ksrc = (kz == 0 && gw_quirk) ? N / 2 : kz;
                    long m = midx(N, i, j, ksrc);
                    double e = 0;
                    for (int c = 0; c < 12; c++) {
                        double v = l0[(long)c * M + m];
                        int off = (c == 2 || c == 3 || c == 4 || c == 5 || c == 8 || c == 9);
                        e += (off ? 2. : 1.) * pw2(v);
                    }
                    double pm = sqrt(pw2(px) + pw2(py) + pw2(kz));
                    int w = (kz == 0 || kz == N / 2) ? 1 : 2;
                    np[(int)pm] += w;
                    f2[(int)pm] += w * e;
                }
            }
        }
        for (int b = 0; b < numbins; b++) {
            if (np[b] > 0) {
                out[(2 * nflds) * numbins + b] = f2[b] * spec_norm / ((double)np[b]);
                out[(2 * nflds + 1) * numbins + b] = norm1_gw * pow((double)b * 2 * M_PI / L, 3.) * f2[b] / (double)np[b];
            } else {
                out[(2 * nflds) * numbins + b] = 0.;
                out[(2 * nflds + 1) * numbins + b] = 0.;
            }
            out[numspec_out * numbins + b] = dp_gw * (double)b; /* trailing column p[i], Q16 */
        }
        free(f2);
        free(np);
    }
    free(spec_power);
    free(spec_numpts);
    free(outdft);
}

/* -------------------------------------------------------------------- FFT */
static int smallest_factor(int n)
{
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* Recursive mixed-radix Cooley-Tukey: out[0..n) = sum_x in[x*is] W^(x k). */
static void fft_rec(int n, const cplx *in, long is, cplx *out, const cplx *W, long ws)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; r++) fft_rec(m, in + r * is, is * p, out + (long)r * m, W, ws * p);
    cplx t[p];
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) t[r] = out[(long)r * m + k];
        for (int q = 0; q < p; q++) {
            long kk = k + (long)q * m;
            cplx acc = 0;
            for (int r = 0; r < p; r++) acc += t[r] * W[((r * kk) % n) * ws];
            out[kk] = acc;
        }
    }
}

static cplx *twiddles(int n, int sign)
{
    cplx *W = (cplx *)malloc(sizeof(cplx) * n);
    for (int e = 0; e < n; e++) {
        double ang = 2. * M_PI * (double)e / (double)n;
        W[e] = cos(ang) + I * (sign * sin(ang));
    }
    return W;
}

void orc_fft1d(int n, const double *in_cplx, double *out_cplx, int sign)
{
    cplx *W = twiddles(n, sign);
    fft_rec(n, (const cplx *)in_cplx, 1, (cplx *)out_cplx, W, 1);
    free(W);
}

/* in-place 3D complex FFT of a[N][N][N] */
static void fft3d(int N, cplx *a, int sign, int nthreads)
{
    cplx *W = twiddles(N, sign);
    long N2 = (long)N * N;
    for (int axis = 2; axis >= 0; axis--) {
        long stride = axis == 2 ? 1 : (axis == 1 ? N : N2);
#pragma omp parallel num_threads(nthreads)
        {
            cplx *buf = (cplx *)malloc(sizeof(cplx) * N), *res = (cplx *)malloc(sizeof(cplx) * N);
#pragma omp for
            for (long line = 0; line < N2; line++) {
                long base;
                long u = line / N, v = line % N;
                if (axis == 2) base = line * N;
                else if (axis == 1) base = u * N2 + v;
                else base = line;
                for (int x = 0; x < N; x++) buf[x] = a[base + x * stride];
                fft_rec(N, buf, 1, res, W, 1);
                for (int x = 0; x < N; x++) a[base + x * stride] = res[x];
            }
            free(buf);
            free(res);
        }
    }
    free(W);
}

void orc_r2c3d(int N, const double *in, double *out_cplx, int nthreads)
{
    long N3 = (long)N * N * N;
    cplx *a = (cplx *)malloc(sizeof(cplx) * N3);
    for (long n = 0; n < N3; n++) a[n] = in[n];
    fft3d(N, a, -1, nthreads > 0 ? nthreads : 1);
    int H = N / 2 + 1;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < H; k++) {
                cplx v = a[sidx(N, i, j, k)];
                long m = midx(N, i, j, k);
                out_cplx[2 * m] = creal(v);
                out_cplx[2 * m + 1] = cimag(v);
            }
    free(a);
}

void orc_c2r3d(int N, const double *in_cplx, double *out, int nthreads)
{
    long N3 = (long)N * N * N;
    int H = N / 2 + 1;
    cplx *a = (cplx *)malloc(sizeof(cplx) * N3);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) {
                cplx v;
                if (k < H) {
                    long m = midx(N, i, j, k);
                    v = in_cplx[2 * m] + I * in_cplx[2 * m + 1];
                } else {
                    long m = midx(N, (N - i) % N, (N - j) % N, N - k);
                    v = in_cplx[2 * m] - I * in_cplx[2 * m + 1];
                }
                a[sidx(N, i, j, k)] = v;
            }
    fft3d(N, a, +1, nthreads > 0 ? nthreads : 1);
    for (long n = 0; n < N3; n++) out[n] = creal(a[n]);
    free(a);
}

/* ------------------------------------------------------------------- init */
void orc_srand(unsigned seed) { srand(seed); }
int orc_rand(void) { return rand(); }

void orc_initfields_homog(orc_state *st, const double *f0, const double *df0) /* g2model.cpp:122-125 */
{
    for (int fld = 0; fld < st->p.nflds; fld++)
        for (long n = 0; n < st->N3; n++) {
            FLD(st, 0, fld)[n] = f0[fld];
            DFLD(st, 0, fld)[n] = df0[fld];
        }
    orc_calcEnergy(st, 0); /* :145 */
}

void orc_initexpansion(orc_state *st) /* g2init.cpp:16-43 */
{
    if (st->p.expansion_type == 0) {
        st->a[0] = 1; st->adot[0] = 0; st->a[1] = 1; st->adot[1] = 0;
        orc_calcEnergy(st, 0);
    } else if (st->p.expansion_type == 1) {
        st->a[0] = 1;
        orc_calcEnergy(st, 0);
        st->adot[0] = orc_adf(st, 0);
    }
}

/* g2init.cpp:68-189.  full_rand 1 (field_full_rand 1, g2parameters.h:72):
 * two (modulus, phase) rand() draws per mode in sequential loop order
 * (g2init.cpp:111-150); full_rand 0: the deterministic real modes of
 * g2init.cpp:153-165, no draws.  cut_off > 0 defines spec_cut_off
 * (g2parameters.h:73): every amplitude is multiplied by the tanh window
 * (1 - tanh(smooth (|n| - N cut_off)))/2 of g2init.cpp:117-129,158; full_rand 0
 * requires it, as the reference does not compile without it.  fk_out/fkd_out
 * (optional, [N][N][N/2+1][2]) receive the mode arrays before the inverse
 * transform. */
void orc_randinit_opt(orc_state *st, int fld, double d2vdf2, int full_rand, double cut_off, double smooth,
                      double *fk_out, double *fkd_out)
{
    int N = NN, H = N / 2 + 1;
    double L = st->p.L, B = st->p.rescale_B;
    long M = st->M;
    double dk2 = 4. * M_PI * M_PI / L / L;
    double H0 = sqrt(8. * M_PI / 3. * st->edrho[0]);
    double *fkf = (double *)calloc(2 * M, sizeof(double));
    double *fkd = (double *)calloc(2 * M, sizeof(double));
    for (int i = 0; i < N; i++) {
        double i2 = (i < (N / 2 + 1) ? (i * i) : ((N - i) * (N - i)));
        for (int j = 0; j < N; j++) {
            double j2 = (j < (N / 2 + 1) ? (j * j) : ((N - j) * (N - j)));
            for (int k = 0; k < H; k++) {
                double k2 = (k * k);
                long m = midx(N, i, j, k);
                if (i + j + k == 0) {
                    fkf[2 * m] = 0.; fkf[2 * m + 1] = 0.; fkd[2 * m] = 0.; fkd[2 * m + 1] = 0.;
                    continue;
                }
                double wk = sqrt(dk2 * ((double)i2 + (double)j2 + (double)k2) + d2vdf2);
                double win = cut_off > 0. ? (1. - tanh(smooth * (sqrt(i2 + j2 + k2) - N * cut_off))) / 2. : 1.;
                if (!full_rand) {
                    fkf[2 * m] = B * sqrt(1 / (2. * wk * L * L * L)) * win;
                    fkf[2 * m + 1] = 0.;
                    fkd[2 * m] = -H0 * fkf[2 * m];
                    fkd[2 * m + 1] = 0.;
                    continue;
                }
                double RI1[2], RI2[2];
                double randVar = (double)rand() / (double)RAND_MAX;
                double randAngle = 2. * M_PI * (double)rand() / (double)RAND_MAX;
                RI1[0] = B * cos(randAngle) * sqrt(-log(randVar) / (2. * wk * L * L * L)) * win;
                RI1[1] = B * sin(randAngle) * sqrt(-log(randVar) / (2. * wk * L * L * L)) * win;
                randVar = (double)rand() / (double)RAND_MAX;
                randAngle = 2. * M_PI * (double)rand() / (double)RAND_MAX;
                RI2[0] = B * cos(randAngle) * sqrt(-log(randVar) / (2. * wk * L * L * L)) * win;
                RI2[1] = B * sin(randAngle) * sqrt(-log(randVar) / (2. * wk * L * L * L)) * win;
                fkf[2 * m] = (RI1[0] + RI2[0]) / sqrt(2.);
                fkf[2 * m + 1] = (RI1[1] + RI2[1]) / sqrt(2.);
                fkd[2 * m] = (RI2[1] - RI1[1]) * wk / sqrt(2.) - H0 * fkf[2 * m];
                fkd[2 * m + 1] = (RI1[0] - RI2[0]) * wk / sqrt(2.) - H0 * fkf[2 * m + 1];
            }
        }
    }
    if (fk_out) memcpy(fk_out, fkf, sizeof(double) * 2 * M);
    if (fkd_out) memcpy(fkd_out, fkd, sizeof(double) * 2 * M);
    double *inic = (double *)malloc(sizeof(double) * st->N3);
    orc_c2r3d(N, fkf, inic, nthr(st));
    for (long n = 0; n < st->N3; n++) FLD(st, 0, fld)[n] += inic[n];
    orc_c2r3d(N, fkd, inic, nthr(st));
    for (long n = 0; n < st->N3; n++) DFLD(st, 0, fld)[n] += inic[n];
    free(inic);
    free(fkf);
    free(fkd);
}

/* the reference default: field_full_rand 1, spec_cut_off not defined */
void orc_randinit(orc_state *st, int fld, double d2vdf2, double *fk_out, double *fkd_out)
{
    orc_randinit_opt(st, fld, d2vdf2, 1, 0., 0., fk_out, fkd_out);
}

void orc_initfields_fluct(orc_state *st) /* g2model.cpp:132-145 (seeded by the caller) */
{
    for (int fld = 0; fld < st->p.nflds; fld++)
        orc_randinit(st, fld, orc_effMass(st, 0, fld), NULL, NULL);
    orc_calcEnergy(st, 0);
}

/* --------------------------------------------------------------- outputs */
void orc_meansvars(const orc_state *st, double *out) /* g2output.cpp:70-121 */
{
    for (int fld = 0; fld < st->p.nflds; fld++) {
        const double *f = FLD(st, 0, fld);
        double av = 0., var = 0.;
        for (long n = 0; n < st->N3; n++) av += f[n];
        av = av / (double)st->N3;
        double av_sq = av * av;
        for (long n = 0; n < st->N3; n++) var += f[n] * f[n] - av_sq;
        var = var / (double)st->N3;
        out[2 * fld] = av;
        out[2 * fld + 1] = var;
    }
}
