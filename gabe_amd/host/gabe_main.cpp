/*
 * gabe_main.cpp -- the kept GABE host program on top of libgabe_amd.so.
 *
 * Driver loop, start-up, output cadence and writers follow the reference's
 * g2main.cpp:42-138 and g2output.cpp:25-319 (slice fields with rho,
 * means/variances with the instability abort, the time series, info.txt) and
 * the spectra writer of g2spectra.cpp:200-232.  The lattice itself lives on the
 * GPU: step() is gabe_step(), calcEnergy/adf come back through
 * gabe_get_scalars(), the spectra through gabe_spectra().
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <sys/stat.h>
#include <vector>

#include "gabe.h"
#include "gabe_host.h"
#include "gabe_model.h"
/* the run's parameters in GABE's own file format, then its model plug-in;
 * the build's include path picks the directory holding them (params/...) */
#include "g2parameters.h"
#include "g2switches.h"
#include "g2model.h"

namespace {

gabe_ctx *ctx = nullptr;
gabe_params params{};
gabe_scalars sc{};
double t = starttime;
time_t tStart, tCurrent, tLast, tFinish;
FILE *info = nullptr;
/* Host copy of the slot-0 planes the writers read, fetched at output slices:
 * the whole lattice for field_outdim 3, otherwise the planes i = 0, +-1, +-2
 * that the plane i = 0 and its 4th-order gradients need. */
std::vector<double> field0, dfield0;
std::vector<int> plane_slot;   /* lattice plane -> slot in field0/dfield0 (-1: not held) */
long held_sites = 0;           /* sites per field held on the host */

void die(int rc, const char *where)
{
    std::fprintf(stderr, "%s failed (%d): %s\n", where, rc, gabe_last_error());
    std::exit(2);
}
#define CALL(x)                           \
    do {                                  \
        int rc_ = (x);                    \
        if (rc_ != GABE_OK) die(rc_, #x); \
    } while (0)

const gabe_model model{model_m2_phi, model_lambda, model_g2, model_m2_chi};

inline long sidx(int i, int j, int k) { return ((long)plane_slot[i] * N + j) * N + k; }
inline int incr(int i) { return i == N - 1 ? 0 : i + 1; }
inline int decr(int i) { return i == 0 ? N - 1 : i - 1; }

/* reference-shaped model wrappers (g2model.cpp:66-108) on the host copy */
double potential(int i, int j, int k)
{
    const long n = sidx(i, j, k);
    return gabe_V(model, field0[n], nflds > 1 ? field0[held_sites + n] : 0.);
}

/* 4th-order first derivatives and |grad f|^2 (g2functions.cpp:70-122) */
double dfd(const double *f, int dir, int i, int j, int k)
{
    int c[3] = {i, j, k}, m1[3] = {i, j, k}, m2[3] = {i, j, k}, p1[3] = {i, j, k}, p2[3] = {i, j, k};
    m1[dir] = decr(c[dir]);
    m2[dir] = decr(m1[dir]);
    p1[dir] = incr(c[dir]);
    p2[dir] = incr(p1[dir]);
    return ((1. / 12.) * f[sidx(m2[0], m2[1], m2[2])] + ((-2.) / 3.) * f[sidx(m1[0], m1[1], m1[2])] +
            (2. / 3.) * f[sidx(p1[0], p1[1], p1[2])] + ((-1.) / 12.) * f[sidx(p2[0], p2[1], p2[2])]) /
           dx;
}

double gradF2(const double *f, int i, int j, int k)
{
    const double a = dfd(f, 0, i, j, k), b = dfd(f, 1, i, j, k), c = dfd(f, 2, i, j, k);
    return a * a + b * b + c * c;
}

/* rho at a site (g2functions.cpp:176-186) */
double calcrho(int i, int j, int k)
{
    double rho = potential(i, j, k);
    const long n = sidx(i, j, k);
    for (int fld = 0; fld < nflds; fld++)
        rho += (.5 * dfield0[fld * held_sites + n] * dfield0[fld * held_sites + n] +
                .5 * gradF2(field0.data() + fld * held_sites, i, j, k) / sc.a[0] / sc.a[0]);
    return rho;
}

void modelinfo(FILE *f)
{
    std::fprintf(f, "Polynomial two-field model\n");
    std::fprintf(f, "V = m2_phi phi^2/2 + lambda phi^4/4 + g2 phi^2 chi^2/2 + m2_chi chi^2/2 (program units)\n\n");
    std::fprintf(f, "m2_phi= %e\nlambda= %e\ng2= %e\nm2_chi= %e\n", model_m2_phi, model_lambda, model_g2, model_m2_chi);
    std::fprintf(f, "%ssing random initial conditions", rand_init ? "U" : "Not u");
}

/* g2output.cpp:259-298 */
void readable_time(int tt, FILE *f)
{
    const int tminutes = 60, thours = 60 * tminutes, tdays = 24 * thours;
    if (tt == 0) {
        std::fprintf(f, "less than 1 second\n");
        return;
    }
    if (tt > tdays) {
        std::fprintf(f, "%d days", tt / tdays);
        tt = tt % tdays;
        if (tt > 0) std::fprintf(f, ", ");
    }
    if (tt > thours) {
        std::fprintf(f, "%d hours", tt / thours);
        tt = tt % thours;
        if (tt > 0) std::fprintf(f, ", ");
    }
    if (tt > tminutes) {
        std::fprintf(f, "%d minutes", tt / tminutes);
        tt = tt % tminutes;
        if (tt > 0) std::fprintf(f, ", ");
    }
    if (tt > 0) std::fprintf(f, "%d seconds", tt);
    std::fprintf(f, "\n");
}

/* g2output.cpp:190-256 */
void output_parameters()
{
    static int first = 1;
    if (first) {
        info = std::fopen("info.txt", "a");
        std::fprintf(info, "--------------------------\n");
        std::fprintf(info, "Model Specific Information\n");
        std::fprintf(info, "--------------------------\n");
        modelinfo(info);
        std::fprintf(info, "\n--------------------------\n");
        std::fprintf(info, "General Program Information\n");
        std::fprintf(info, "-----------------------------\n");
        std::fprintf(info, "Grid size=%d^3\n", N);
        std::fprintf(info, "L=%e\n", L);
        std::fprintf(info, "dt=%e, dx=%e\n", dt, dx);
        std::fprintf(info, "end time=%e\n", endtime);
        std::fprintf(info, "B=%e\n", rescale_B);
        std::fprintf(info, "\nUsing %s expansion\n", expansion_type == 1 ? "a dot" : (expansion_type == 2 ? "a duble dot" : "no"));
        std::fprintf(info, "\nUsing %d cores\n", tot_num_thrds);
        std::fprintf(info, "%d momentum bins for spectra\n", gabe_numbins(N));
        time(&tStart);
        std::fprintf(info, "\nRun began at %s", ctime(&tStart));
        first = 0;
    } else {
        time(&tFinish);
        std::fprintf(info, "Run ended at %s", ctime(&tFinish));
        std::fprintf(info, "\nRun from t=%e to t=%e took ", starttime, t);
        readable_time((int)(tFinish - tStart), info);
        std::fprintf(info, "\n");
        std::fprintf(info, "\nFinal scale factor is %e\n", sc.a[0]);
    }
    std::fflush(info);
}

void fetch_state()
{
    const long N2 = (long)N * N;
    plane_slot.assign(N, -1);
    if (field_outdim > 2) {
        for (int i = 0; i < N; i++) plane_slot[i] = i;
        held_sites = N2 * N;
        field0.resize((size_t)nflds * held_sites);
        dfield0.resize((size_t)nflds * held_sites);
        CALL(gabe_download_state(ctx, field0.data(), dfield0.data()));
        return;
    }
    std::vector<int> planes;
    for (int d = -2; d <= 2; d++) {
        const int i = ((d % N) + N) % N;
        if (plane_slot[i] < 0) {
            plane_slot[i] = (int)planes.size();
            planes.push_back(i);
        }
    }
    held_sites = (long)planes.size() * N2;
    field0.resize((size_t)nflds * held_sites);
    dfield0.resize((size_t)nflds * held_sites);
    std::vector<double> pf((size_t)nflds * N2), pd((size_t)nflds * N2);
    for (int i : planes) {
        CALL(gabe_download_plane(ctx, i, pf.data(), pd.data()));
        for (int fld = 0; fld < nflds; fld++) {
            std::copy(pf.begin() + fld * N2, pf.begin() + (fld + 1) * N2,
                      field0.begin() + fld * held_sites + plane_slot[i] * N2);
            std::copy(pd.begin() + fld * N2, pd.begin() + (fld + 1) * N2,
                      dfield0.begin() + fld * held_sites + plane_slot[i] * N2);
        }
    }
}

/* g2output.cpp:25-68: field_outdim 3 writes every field_sliceskip-th site of
 * the lattice, 2 the plane i = 0, 1 the line i = j = 0 */
void outputfield(int first)
{
    char name[64];
    std::snprintf(name, sizeof name, "./slices/slices_fields_%d.dat", first);
    FILE *f = std::fopen(name, "w");
    const int ni = field_outdim > 2 ? N : 1, nj = field_outdim > 1 ? N : 1;
    for (int i = 0; i < ni; i += field_sliceskip)
        for (int j = 0; j < nj; j += field_sliceskip)
            for (int k = 0; k < N; k += field_sliceskip) {
                for (int fld = 0; fld < nflds; fld++) std::fprintf(f, "%e ", field0[fld * held_sites + sidx(i, j, k)]);
                std::fprintf(f, "%e ", calcrho(i, j, k));
                std::fprintf(f, "\n");
            }
    std::fclose(f);
}

/* g2spectra.cpp:200-232 */
void specOut(int first)
{
    const int nb = gabe_numbins(N), ncol = 2 * nflds + (calc_gws ? 3 : 0);
    std::vector<double> out((size_t)ncol * nb);
    CALL(gabe_spectra(ctx, out.data()));
    char name[64];
    std::snprintf(name, sizeof name, "./slices/slices_spectra_%d.dat", first);
    FILE *f = std::fopen(name, "w");
    for (int i = 0; i < nb; i++) {
        std::fprintf(f, "%e", 2. * M_PI * i / L);
        for (int c = 0; c < ncol; c++) std::fprintf(f, " %e", out[(size_t)c * nb + i]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

/* g2output.cpp:70-121 */
void meansvars()
{
    FILE *f = std::fopen("./slices/meansvariances.dat", "a");
    std::fprintf(f, "%f", t);
    std::vector<double> mv(2 * nflds);
    CALL(gabe_means_vars(ctx, mv.data()));   /* reduced on the device */
    for (int fld = 0; fld < nflds; fld++) {
        const double av = mv[2 * fld], var = mv[2 * fld + 1];
        std::fprintf(f, " %e %e", av, var);
        if ((av != 0. && av / av != 1.)) {
            std::printf("Unstable solution developed. Field %d not numerical at t=%e\n", fld, t);
            output_parameters();
            std::fflush(f);
            std::exit(1);
        }
    }
    std::fprintf(f, "\n");
    std::fclose(f);
}

int slicewaitf() { return (int)(endtime / dt / slicenumber); }

/* g2output.cpp:129-186: decides whether this call writes a slice */
/* the writer's counters live at file scope so a checkpoint can carry them */
int slice_first = 0, slice_last = 0, spec_last = 0, slice_waitm = 0;

void outputslice()
{
    int &first = slice_first, &lastslice = slice_last, &lastspec = spec_last, &slicewaitm = slice_waitm;
    if (first == 0) {
        slicewaitm = slicewait == 0 ? slicewaitf() : slicewait;
        lastslice = slicewaitm;
        lastspec = specnumber;
    }
    lastslice++;
    if (lastslice >= slicewaitm) {
        int rc = gabe_get_scalars(ctx, &sc);
        if (rc != GABE_OK && rc != GABE_ERR_NONFINITE) die(rc, "gabe_get_scalars");
        fetch_state();
        if (field_outdim != 0) outputfield(first);
        if (spec_output) {
            lastspec++;
            if (lastspec >= specnumber) {
                specOut(first / specnumber);
                lastspec = 0;
            }
        }
        if (var_output) meansvars();
        FILE *st = std::fopen("./slices/slices_time.dat", "a");
        std::fprintf(st, "%e %e %e %e %e %e %e %e\n", t, sc.a[0], sc.adot[0], sc.adot[0] / sc.a[0], sc.edkin[0],
                     sc.edpot[0], sc.edgrad[0], sc.edrho[0]);
        std::fclose(st);
        first++;
        lastslice = 0;
    }
}

/* g2output.cpp:302-319 */
void screenout()
{
    time(&tCurrent);
    if (screentime == 0) return;
    if (tCurrent - tLast > screentime) {
        std::printf("%f\n", t);
        time(&tLast);
    }
}

/* Checkpoint of the run after `step` steps at loop time t (not in the
 * reference, which has none). */
void write_checkpoint(const char *path, long long step)
{
    gabe_host_run run{};
    run.t = t;
    run.step = step;
    run.randseed = (unsigned)randseed;
    run.rand_draws = (rand_init && field_full_rand) ? nflds * gabe_host_draws_per_field(N) : 0;
    run.slice_first = slice_first;
    run.slice_last = slice_last;
    run.spec_last = spec_last;
    run.slice_wait = slice_waitm;
    CALL(gabe_host_checkpoint_write(ctx, path, &run));
}

} // namespace

int main(int argc, char **argv)
{
    const char *resume = nullptr;
    long long stop_at = -1;
    for (int a = 1; a < argc; a++) {
        if (!std::strcmp(argv[a], "--resume") && a + 1 < argc) {
            resume = argv[++a];
        } else if (!std::strcmp(argv[a], "--checkpoint-at") && a + 1 < argc) {
            stop_at = std::atoll(argv[++a]);
        } else {
            std::fprintf(stderr, "usage: %s [--resume FILE] [--checkpoint-at STEP]\n", argv[0]);
            return 2;
        }
    }
    std::printf("\n\nGABE started!\n\n");
    output_parameters();
    std::printf("Info file made\n");
    mkdir("slices", 0755);

    params.N = N;
    params.nflds = nflds;
    params.L = L;
    params.dt = dt;
    params.calc_gws = calc_gws;
    params.expansion_type = expansion_type;
    params.precision = GABE_PRECISION ? GABE_FP32 : GABE_FP64;
    params.m2_phi = model_m2_phi;
    params.lambda = model_lambda;
    params.g2 = model_g2;
    params.m2_chi = model_m2_chi;
    params.gw_spectrum_quirk = 1;
    CALL(gabe_create(&params, &ctx));
    std::printf("Memory allocated for all arrays\n\n");
    std::printf("Starting run...\n\n");

    long long step = 0;
    bool resumed = false;
    if (resume) {
        gabe_host_run run{};
        CALL(gabe_host_checkpoint_read(ctx, resume, &run));
        t = run.t;
        step = run.step;
        slice_first = run.slice_first;
        slice_last = run.slice_last;
        spec_last = run.spec_last;
        slice_waitm = run.slice_wait;
        resumed = true;
        t += dt;   /* the loop's increment after the step the checkpoint followed */
        std::printf("Resumed from %s at step %lld\n", resume, step);
    }
    for (t = resumed ? t : starttime; t <= endtime; t += dt) {   /* g2main.cpp:106-131 */
        if (!resumed && t == starttime) {
            gabe_host_init init{};
            for (int fld = 0; fld < nflds; fld++) {
                init.f0[fld] = f0[fld];
                init.df0[fld] = df0[fld];
            }
            init.rescale_B = rescale_B;
            init.randseed = (unsigned)randseed;
            init.rand_init = rand_init;
            init.symmetric_modes = field_full_rand ? 0 : 1;
            init.spec_cut_off = spec_cut_off;
            init.spec_smooth = spec_smooth;
            CALL(gabe_host_initialize(ctx, &params, &init));
            std::printf("Fields initialized and fluctuated\n");
            screenout();
            outputslice();
        }
        CALL(gabe_step(ctx, 1));
        step++;
        screenout();
        outputslice();
        if (checkpoint_every > 0 && step % checkpoint_every == 0) write_checkpoint("checkpoint.bin", step);
        if (step == stop_at) {
            write_checkpoint("checkpoint.bin", step);
            std::printf("Checkpoint written after step %lld\n", step);
            gabe_destroy(ctx);
            return 0;
        }
    }
    CALL(gabe_sync(ctx));
    CALL(gabe_get_scalars(ctx, &sc));
    std::printf("\nRun complete\n\n");
    output_parameters();
    gabe_destroy(ctx);
    std::printf("gabe_amd host program finished\n\n");
    return 0;
}
