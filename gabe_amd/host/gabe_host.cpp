/*
 * gabe_host.cpp -- initial conditions and start-up sequence of the kept host
 * program, on top of the C ABI (include/gabe.h).
 */
#include "gabe_host.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <cstdio>
#include <vector>

extern "C" int gabe_host_modes_opt(int N, double L, double rescale_B, double d2vdf2, double H0,
                                   const gabe_host_modes *opt, const int *draws, double *fk, double *fkd)
{
    if (N < 4 || !opt || !fk || !fkd || (opt->full_rand && !draws)) return GABE_ERR_ARG;
    if (!opt->full_rand && !(opt->cut_off > 0.)) return GABE_ERR_ARG;   /* field_full_rand 0 needs spec_cut_off */
    const int H = N / 2 + 1;
    const double dk2 = 4. * M_PI * M_PI / L / L;
    const double s2 = std::sqrt(2.);
    long q = 0;
    for (int i = 0; i < N; i++) {
        const double i2 = (i < (N / 2 + 1) ? (i * i) : ((N - i) * (N - i)));
        for (int j = 0; j < N; j++) {
            const double j2 = (j < (N / 2 + 1) ? (j * j) : ((N - j) * (N - j)));
            for (int k = 0; k < H; k++) {
                const double k2 = (k * k);
                const long m = ((long)i * N + j) * H + k;
                if (i + j + k == 0) {          /* zero mode stays zero (g2init.cpp:95-110) */
                    fk[2 * m] = fk[2 * m + 1] = 0.;
                    fkd[2 * m] = fkd[2 * m + 1] = 0.;
                    continue;
                }
                const double wk = std::sqrt(dk2 * (i2 + j2 + k2) + d2vdf2);
                /* tanh window of g2init.cpp:122-129,158 */
                const double win = opt->cut_off > 0.
                                       ? (1. - std::tanh(opt->smooth * (std::sqrt(i2 + j2 + k2) - N * opt->cut_off))) / 2.
                                       : 1.;
                /* products in the reference's evaluation order, so the modes
                 * agree to the last bit */
                if (!opt->full_rand) {         /* g2init.cpp:153-165 */
                    fk[2 * m] = rescale_B * std::sqrt(1 / (2. * wk * L * L * L)) * win;
                    fk[2 * m + 1] = 0.;
                    fkd[2 * m] = -H0 * fk[2 * m];
                    fkd[2 * m + 1] = 0.;
                    continue;
                }
                double r1[2], r2[2];
                for (double *r : {r1, r2}) {   /* two (modulus, phase) draws per mode */
                    const double u = (double)draws[q++] / (double)RAND_MAX;
                    const double th = 2. * M_PI * (double)draws[q++] / (double)RAND_MAX;
                    const double mag = std::sqrt(-std::log(u) / (2. * wk * L * L * L));
                    r[0] = rescale_B * std::cos(th) * mag * win;
                    r[1] = rescale_B * std::sin(th) * mag * win;
                }
                fk[2 * m] = (r1[0] + r2[0]) / s2;
                fk[2 * m + 1] = (r1[1] + r2[1]) / s2;
                fkd[2 * m] = (r2[1] - r1[1]) * wk / s2 - H0 * fk[2 * m];
                fkd[2 * m + 1] = (r1[0] - r2[0]) * wk / s2 - H0 * fk[2 * m + 1];
            }
        }
    }
    return GABE_OK;
}

extern "C" int gabe_host_modes_from_draws(int N, double L, double rescale_B, double d2vdf2, double H0,
                                          const int *draws, double *fk, double *fkd)
{
    const gabe_host_modes opt{1, 0., 0.};
    return gabe_host_modes_opt(N, L, rescale_B, d2vdf2, H0, &opt, draws, fk, fkd);
}

extern "C" long gabe_host_draws_per_field(int N) { return 4L * ((long)N * N * (N / 2 + 1) - 1); }

extern "C" int gabe_host_draw_modes(int N, double L, double rescale_B, double d2vdf2, double H0, double *fk,
                                    double *fkd)
{
    if (N < 4) return GABE_ERR_ARG;
    std::vector<int> draws((size_t)gabe_host_draws_per_field(N));
    for (auto &d : draws) d = rand();
    return gabe_host_modes_from_draws(N, L, rescale_B, d2vdf2, H0, draws.data(), fk, fkd);
}

extern "C" int gabe_host_initialize(gabe_ctx *ctx, const gabe_params *p, const gabe_host_init *init)
{
    if (!ctx || !p || !init) return GABE_ERR_ARG;
    const int N = p->N, nf = p->nflds;
    const long N3 = (long)N * N * N, M = (long)N * N * (N / 2 + 1);
    int rc;
    {   /* initfields #1 (g2model.cpp:122-125,145) */
        std::vector<double> f((size_t)nf * N3), d((size_t)nf * N3);
        for (int fld = 0; fld < nf; fld++)
            for (long n = 0; n < N3; n++) {
                f[(size_t)fld * N3 + n] = init->f0[fld];
                d[(size_t)fld * N3 + n] = init->df0[fld];
            }
        if ((rc = gabe_upload_state(ctx, f.data(), d.data())) != GABE_OK) return rc;
    }
    if ((rc = gabe_energy(ctx, 0)) != GABE_OK) return rc;
    gabe_scalars s;
    if ((rc = gabe_get_scalars(ctx, &s)) != GABE_OK) return rc;
    /* initexpansion (g2init.cpp:16-43) */
    s.a[0] = 1.;
    if (p->expansion_type == 0) {
        s.adot[0] = 0.; s.a[1] = 1.; s.adot[1] = 0.;
    } else {
        s.adot[0] = std::sqrt(8. * M_PI / 3. * s.edrho[0]) * s.a[0];
    }
    if ((rc = gabe_set_scalars(ctx, &s)) != GABE_OK) return rc;
    if (init->rand_init) {   /* initfields #2 (g2model.cpp:128-141) */
        const double H0 = std::sqrt(8. * M_PI / 3. * s.edrho[0]);
        /* every rand() value of every field is drawn first, in the reference's
         * order (field by field, g2model.cpp:131-133): device and FFT library
         * calls made in between must not be able to shift the sequence */
        const gabe_host_modes opt{init->symmetric_modes ? 0 : 1, init->spec_cut_off, init->spec_smooth};
        if (!opt.full_rand && !(opt.cut_off > 0.)) return GABE_ERR_ARG;
        const long per = opt.full_rand ? gabe_host_draws_per_field(N) : 0;
        std::vector<int> draws((size_t)per * nf);
        srand(init->randseed);
        for (auto &d : draws) d = rand();
        std::vector<double> fk(2 * M), fkd(2 * M);
        for (int fld = 0; fld < nf; fld++) {
            double m2;
            if ((rc = gabe_effmass(ctx, fld, &m2)) != GABE_OK) return rc;
            if ((rc = gabe_host_modes_opt(N, p->L, init->rescale_B, m2, H0, &opt, draws.data() + (size_t)fld * per,
                                          fk.data(), fkd.data())) != GABE_OK)
                return rc;
            if ((rc = gabe_add_fluctuation(ctx, fld, fk.data(), fkd.data())) != GABE_OK) return rc;
        }
    }
    return gabe_energy(ctx, 0);
}

namespace {
const char kRunMagic[8] = {'G', 'A', 'B', 'E', 'R', 'U', 'N', '1'};
}

extern "C" int gabe_host_checkpoint_write(gabe_ctx *ctx, const char *path, const gabe_host_run *run)
{
    if (!ctx || !path || !run) return GABE_ERR_ARG;
    size_t n = 0;
    int rc = gabe_checkpoint_size(ctx, &n);
    if (rc != GABE_OK) return rc;
    std::vector<char> blob(n);
    if ((rc = gabe_checkpoint_save(ctx, blob.data(), n)) != GABE_OK) return rc;
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return GABE_ERR_ARG;
    const unsigned long long len = n;
    bool ok = std::fwrite(kRunMagic, 1, 8, f) == 8 && std::fwrite(run, sizeof *run, 1, f) == 1 &&
              std::fwrite(&len, sizeof len, 1, f) == 1 && std::fwrite(blob.data(), 1, n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return GABE_ERR_ARG;
    }
    return GABE_OK;
}

extern "C" int gabe_host_checkpoint_read(gabe_ctx *ctx, const char *path, gabe_host_run *run)
{
    if (!ctx || !path || !run) return GABE_ERR_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return GABE_ERR_ARG;
    char magic[8];
    gabe_host_run r;
    unsigned long long len = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kRunMagic, 8) == 0 &&
              std::fread(&r, sizeof r, 1, f) == 1 && std::fread(&len, sizeof len, 1, f) == 1 && len < (1ULL << 40);
    std::vector<char> blob;
    if (ok) {
        blob.resize((size_t)len);
        ok = std::fread(blob.data(), 1, (size_t)len, f) == (size_t)len;
    }
    std::fclose(f);
    if (!ok) return GABE_ERR_ARG;
    const int rc = gabe_checkpoint_load(ctx, blob.data(), (size_t)len);
    if (rc != GABE_OK) return rc;
    *run = r;
    return GABE_OK;
}

extern "C" double gabe_host_friedmann_residual(const double *hist, int nrows, double *out)
{
    double worst = 0.;
    for (int n = 0; n < nrows; n++) {
        if (out) out[n] = 0.;
        if (n == 0 || n == nrows - 1) continue;
        const double *r = hist + 8L * n, *rm = hist + 8L * (n - 1), *rp = hist + 8L * (n + 1);
        const double H = r[2] / r[1];
        const double rho = r[6], p = r[3] - r[5] / 3. - r[4];
        const double rhodot = (rp[6] - rm[6]) / (rp[0] - rm[0]);
        const double R = std::fabs(rhodot + 3. * H * (rho + p)) / (3. * H * rho);
        if (out) out[n] = R;
        if (R > worst) worst = R;
    }
    return worst;
}
