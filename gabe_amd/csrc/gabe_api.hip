/*
 * gabe_api.hip -- the C ABI declared in include/gabe.h: device state of one
 * lattice, rocFFT plans and the kernel sequence of a step.
 */
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <mutex>
#include <vector>

#include "gabe.h"
#include "gabe_fft.h"
#include "gabe_gwfft.h"
#include "gabe_ctx.h"
#include "gabe_internal.h"

namespace gabe_detail {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}




bool g_rocfft_ready = false;

/* Neumaier-compensated sum in index order */
double csum(const double *x, long n)
{
    double s = 0., c = 0.;
    for (long i = 0; i < n; i++) {
        double t = s + x[i];
        if (std::fabs(s) >= std::fabs(x[i])) c += (s - t) + x[i];
        else c += (x[i] - t) + s;
        s = t;
    }
    return s + c;
}

} // namespace gabe_detail

using namespace gabe_detail;


namespace gabe_detail {

int ensure_rocfft()
{
    if (!g_rocfft_ready) {
        FFTCHK(rocfft_setup());
        g_rocfft_ready = true;
    }
    return GABE_OK;
}

int make_plan(gabe_ctx *c, FftPlan &P, rocfft_transform_type type, bool dbl, size_t batch, int rank,
              size_t out_distance)
{
    CHK(ensure_rocfft());
    size_t len[3] = {(size_t)c->N, (size_t)c->N, (size_t)c->N};
    rocfft_plan_description desc = nullptr;
    if (out_distance) {   /* rank-1 rows with a padded output row stride */
        FFTCHK(rocfft_plan_description_create(&desc));
        size_t in_stride = 1, out_stride = 1;
        FFTCHK(rocfft_plan_description_set_data_layout(desc, rocfft_array_type_real,
                                                       rocfft_array_type_hermitian_interleaved, nullptr, nullptr, 1,
                                                       &in_stride, (size_t)c->N, 1, &out_stride, out_distance));
    }
    const rocfft_status st = rocfft_plan_create(&P.plan, rocfft_placement_notinplace, type,
                                                dbl ? rocfft_precision_double : rocfft_precision_single,
                                                (size_t)rank, len, batch, desc);
    if (desc) rocfft_plan_description_destroy(desc);
    FFTCHK(st);
    FFTCHK(rocfft_plan_get_work_buffer_size(P.plan, &P.work));
    FFTCHK(rocfft_execution_info_create(&P.info));
    FFTCHK(rocfft_execution_info_set_stream(P.info, c->stream));
    if (P.work > c->fft_work_size) {
        if (c->fft_work) HIPCHK(hipFree(c->fft_work));
        c->fft_work = nullptr;
        HIPCHK(hipMalloc(&c->fft_work, P.work));
        c->fft_work_size = P.work;
    }
    return GABE_OK;
}

int run_plan(gabe_ctx *c, FftPlan &P, void *in, void *out)
{
    if (P.work) FFTCHK(rocfft_execution_info_set_work_buffer(P.info, c->fft_work, P.work));
    void *ib[1] = {in}, *ob[1] = {out};
    FFTCHK(rocfft_execute(P.plan, ib, ob, P.info));
    return GABE_OK;
}

void destroy_plan(FftPlan &P)
{
    if (P.plan) rocfft_plan_destroy(P.plan);
    if (P.info) rocfft_execution_info_destroy(P.info);
    P.plan = nullptr;
    P.info = nullptr;
}

template <typename T>
SweepArgs<T> sweep_args(gabe_ctx *c)
{
    SweepArgs<T> a{};
    for (int f = 0; f < c->nf; f++) {
        a.f[f] = (const T *)c->fbuf[c->cur] + (long)f * c->N3;
        a.v[f] = (const T *)c->vbuf[c->cur] + (long)f * c->N3;
        a.fo[f] = (T *)c->fbuf[1 - c->cur] + (long)f * c->N3;
        a.vo[f] = (T *)c->vbuf[1 - c->cur] + (long)f * c->N3;
    }
    a.stress = (T *)c->stress;
    a.shat = (T *)c->shat;
    a.tw = (const T *)c->twiddle;
    a.kpad = c->kpad;
    a.rowfft = c->rowfft ? c->N : 0;
    a.partials = c->partials;
    a.sc = c->sc;
    a.ss = c->ss;
    a.model = c->model;
    a.N = c->N;
    a.ci = c->ci;
    a.nblk = c->nblk;
    a.gx = (int)c->grid.x;
    a.gy = (int)c->grid.y;
    a.N2 = c->N2;
    a.N3 = c->N3;
    a.dt = c->p.dt;
    a.inv_dx2 = 1. / (c->dx * c->dx);
    return a;
}

int launch_sweep_ctx(gabe_ctx *c, int mode, bool gw, hipStream_t st = nullptr)
{
    if (!st) st = c->stream;
    if (c->fp32) launch_sweep<float>(mode, c->nf, gw, sweep_args<float>(c), c->grid, st);
    else launch_sweep<double>(mode, c->nf, gw, sweep_args<double>(c), c->grid, st);
    HIPCHK(hipGetLastError());
    return GABE_OK;
}

int launch_finalize_ctx(gabe_ctx *c, int mode, int refresh_adot)
{
    FinArgs F{};
    F.partials = c->partials;
    F.nblk = c->nblk;
    F.nf = c->nf;
    F.mode = mode;
    F.exp_type = c->p.expansion_type;
    F.split = c->fp32 ? 1 : 0;
    F.refresh_adot = refresh_adot;
    F.N3 = (double)c->N3;
    F.dt = c->p.dt;
    F.sc = c->sc;
    F.ss = c->ss;
    F.hist = c->hist;
    F.hist_cap = c->hist_cap;
    launch_finalize(F, c->stream);
    HIPCHK(hipGetLastError());
    return GABE_OK;
}

int launch_gw_ctx(gabe_ctx *c)
{
    const double dx3 = c->dx * c->dx * c->dx;
    if (c->fp32) {
        GwArgs<float> g{(const float *)c->shat, (float *)c->h, (float *)c->l, c->ss, c->N, c->M, c->kpad, c->p.L, c->p.dt, dx3};
        launch_gw<float>(g, c->stream);
    } else {
        GwArgs<double> g{(const double *)c->shat, (double *)c->h, (double *)c->l, c->ss, c->N, c->M, c->kpad, c->p.L, c->p.dt, dx3};
        launch_gw<double>(g, c->stream);
    }
    HIPCHK(hipGetLastError());
    return GABE_OK;
}

/* kz = N/2 plane [GABE_NSC][N][N] of the current shat buffer when sweep A packs
 * the rows (after the GABE_NSC row blocks), else null */
void *nyq_plane(gabe_ctx *c)
{
    return c->rowfft ? (char *)c->shat + 2 * c->esz * GABE_NSC * c->N2 * c->kpad : nullptr;
}

/* R2C of the GABE_NSC stress arrays into shat [GABE_NSC][N][N][kpad] (kz < N/2+1 valid,
 * or kz < N/2 and the kz = N/2 plane when sweep A packs the rows).
 * with_i = false stops after the j columns (the fused kernel does i). */
int stress_fft(gabe_ctx *c, bool with_i = true, hipStream_t st = nullptr)
{
    if (!st) st = c->stream;
    if (!c->colfft) return run_plan(c, c->fft_stress, c->stress, c->shat);
    if (!c->rowfft) CHK(run_plan(c, c->fft_stress, c->stress, c->shat));   /* rows along k */
    const long N = c->N, kp = c->kpad;
    const int cols = c->rowfft ? (int)(N / 2) : (int)(N / 2 + 1);
    void *nyq = nyq_plane(c);
    ColFftView vj{c->shat, kp, N * kp, 0, cols, 1, GABE_NSC * N, nyq};       /* along j */
    ColFftView vi{c->shat, N * kp, N * N * kp, kp, cols, (int)N, GABE_NSC};  /* along i */
    ColFftView vn{nyq, N, N * N, 0, (int)N, 1, GABE_NSC};                    /* kz = N/2 plane along i */
    if (c->fp32) {
        launch_colfft<float>(vj, c->N, c->fft_cw, c->twiddle, st);
        if (with_i) launch_colfft<float>(vi, c->N, c->fft_cw, c->twiddle, st);
        if (with_i && nyq) launch_colfft<float>(vn, c->N, c->fft_cw, c->twiddle, st);
    } else {
        launch_colfft<double>(vj, c->N, c->fft_cw, c->twiddle, st);
        if (with_i) launch_colfft<double>(vi, c->N, c->fft_cw, c->twiddle, st);
        if (with_i && nyq) launch_colfft<double>(vn, c->N, c->fft_cw, c->twiddle, st);
    }
    HIPCHK(hipGetLastError());
    return GABE_OK;
}

/* sweep A in launches of c->chunk plane blocks on the main stream, each
 * followed on jst by the j-transform of its planes (kz = N/2 split off), so
 * that the spectra a launch wrote are transformed while still in the
 * memory-side cache.  ev_jdone marks the last transform. */
int chunked_sweep_a(gabe_ctx *c)
{
    const int gy = (int)c->grid.y, bpc = c->chunk;
    const long N = c->N, kp = c->kpad;
    const size_t csz = 2 * c->esz;   /* bytes per complex element */
    char *nyq = (char *)nyq_plane(c);
    for (int q = 0, b0 = 0; b0 < gy; q++, b0 += bpc) {
        const int nb = std::min(bpc, gy - b0);
        const dim3 grid(c->grid.x, (unsigned)nb);
        if (c->fp32) {
            SweepArgs<float> a = sweep_args<float>(c);
            a.by0 = b0;
            launch_sweep<float>(SW_A, c->nf, true, a, grid, c->stream);
        } else {
            SweepArgs<double> a = sweep_args<double>(c);
            a.by0 = b0;
            launch_sweep<double>(SW_A, c->nf, true, a, grid, c->stream);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev_chunk[q], c->stream));
        HIPCHK(hipStreamWaitEvent(c->jst, c->ev_chunk[q], 0));
        const long p0 = (long)b0 * c->ci, np = std::min((long)nb * c->ci, N - p0);
        /* [comp][plane p0 + s][j][kz]: outer = component, sub = plane */
        ColFftView vj{(char *)c->shat + csz * p0 * N * kp, kp, N * N * kp, N * kp, (int)(N / 2), (int)np,
                      GABE_NSC, nyq + csz * p0 * N, N * N};
        if (c->fp32) launch_colfft<float>(vj, c->N, c->fft_cw, c->twiddle, c->jst);
        else launch_colfft<double>(vj, c->N, c->fft_cw, c->twiddle, c->jst);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(c->ev_jdone, c->jst));
    return GABE_OK;
}

/* tensor modes of one step: i transform + TT projection + both RK2 stages */
int gw_modes(gabe_ctx *c, hipStream_t st = nullptr)
{
    if (!st) st = c->stream;
    if (!c->gwfft) return launch_gw_ctx(c);
    const double dx3 = c->dx * c->dx * c->dx;
    if (c->fp32) {
        GwFftArgs<float> g{(const float *)c->shat, (float *)c->h, (float *)c->l, c->ss, (const float *)c->twiddle,
                           c->N, c->M, c->kpad, c->p.L, c->p.dt, dx3, c->lay.nt4, (const float *)nyq_plane(c),
                           c->gw_tile};
        launch_gw_ifft<float>(g, st);
    } else {
        GwFftArgs<double> g{(const double *)c->shat, (double *)c->h, (double *)c->l, c->ss,
                            (const double *)c->twiddle, c->N, c->M, c->kpad, c->p.L, c->p.dt, dx3, c->lay.nt4,
                            (const double *)nyq_plane(c), c->gw_tile};
        launch_gw_ifft<double>(g, st);
    }
    HIPCHK(hipGetLastError());
    return GABE_OK;
}

/* the kernels of one step; ev (optional) receives 2 events per kernel class.
 * Without ev, sweep B (which needs only the finalize) runs on the side stream
 * concurrently with the stress transform and the GW modes; the step ends when
 * both are done.  With ev, the kernels run one after another so that each
 * class is timed alone. */
int one_step(gabe_ctx *c, hipEvent_t *ev, hipEvent_t *cl_ev = nullptr)
{
    auto rec = [&](int k) -> int {
        if (ev) HIPCHK(hipEventRecord(ev[k], c->stream));
        return GABE_OK;
    };
    const bool pipe = c->pipe && !ev;
    /* this slot's spectra and step scalars are free once the side work of
     * the step two back, which read them, is done */
    if (pipe) HIPCHK(hipStreamWaitEvent(c->stream, c->ev_gw[c->slot], 0));
    CHK(rec(0));
    if (pipe && c->chunk) CHK(chunked_sweep_a(c));
    else CHK(launch_sweep_ctx(c, c->gw ? SW_A : c->sweep_a, c->gw));
    CHK(rec(1));
    CHK(rec(2));
    CHK(launch_finalize_ctx(c, SW_A, 0));
    CHK(rec(3));
    if (pipe) {
        HIPCHK(hipEventRecord(c->ev_fin, c->stream));
        HIPCHK(hipStreamWaitEvent(c->side, c->ev_fin, 0));
        if (c->chunk) HIPCHK(hipStreamWaitEvent(c->side, c->ev_jdone, 0));
        else CHK(stress_fft(c, false, c->side));
        CHK(gw_modes(c, c->side));
        HIPCHK(hipEventRecord(c->ev_gw[c->slot], c->side));
        CHK(launch_sweep_ctx(c, c->sweep_b, false));
        c->cur = 1 - c->cur;
        c->slot ^= 1;
        c->shat = c->shat_b[c->slot];
        c->ss = c->ss_b[c->slot];
        if (c->closure) CHK(closure_after_step(c, cl_ev));
        return GABE_OK;
    }
    const bool overlap = c->gw && !ev && c->side;
    if (overlap) {
        HIPCHK(hipEventRecord(c->ev_fin, c->stream));
        HIPCHK(hipStreamWaitEvent(c->side, c->ev_fin, 0));
        CHK(launch_sweep_ctx(c, c->sweep_b, false, c->side));
        HIPCHK(hipEventRecord(c->ev_side, c->side));
    }
    CHK(rec(4));
    if (c->gw) CHK(stress_fft(c, !c->gwfft));
    CHK(rec(5));
    CHK(rec(6));
    if (c->gw) CHK(gw_modes(c));
    CHK(rec(7));
    CHK(rec(8));
    if (overlap) HIPCHK(hipStreamWaitEvent(c->stream, c->ev_side, 0));
    else CHK(launch_sweep_ctx(c, c->sweep_b, false));
    CHK(rec(9));
    c->cur = 1 - c->cur;
    if (c->pipe) {
        c->slot ^= 1;
        c->shat = c->shat_b[c->slot];
        c->ss = c->ss_b[c->slot];
    }
    if (c->closure) CHK(closure_after_step(c, cl_ev));
    return GABE_OK;
}

int read_scalars(gabe_ctx *c, DevScalars &d)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&d, c->sc, sizeof d, hipMemcpyDeviceToHost));
    return GABE_OK;
}

int write_scalars(gabe_ctx *c, const DevScalars &d)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->sc, &d, sizeof d, hipMemcpyHostToDevice));
    return GABE_OK;
}

/* energies (and adot[0] if stale) of the current state */
int refresh_e0(gabe_ctx *c)
{
    DevScalars d;
    CHK(read_scalars(c, d));
    if (d.e0_valid && d.adot0_valid) return GABE_OK;
    CHK(launch_sweep_ctx(c, SW_ENERGY, false));
    CHK(launch_finalize_ctx(c, SW_ENERGY, 1));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GABE_OK;
}

/* tiling chosen by gabe_autotune_sweep, keyed by (N, nflds, FP32, tensor modes) */
struct Tiling {
    int planes = 0;      /* sweep planes per workgroup */
    int cw = 0;          /* j-transform columns per workgroup */
    int mode_tile = 0;   /* fused i-transform + mode kernel variant */
    int chunk = 0;       /* planes per interleaved sweep-A / j-transform launch (-1: none) */
};
std::map<long, Tiling> g_tuned;
std::mutex g_tuned_mu;

long tune_key(int N, int nf, bool fp32, bool gw) { return (((long)N * 4 + nf) * 2 + (fp32 ? 1 : 0)) * 2 + (gw ? 1 : 0); }

Tiling tuned_tiling(int N, int nf, bool fp32, bool gw)
{
    std::lock_guard<std::mutex> lock(g_tuned_mu);
    const auto it = g_tuned.find(tune_key(N, nf, fp32, gw));
    return it == g_tuned.end() ? Tiling{} : it->second;
}

int check_params(const gabe_params *p)
{
    if (!p) return fail(GABE_ERR_ARG, "params is NULL");
    if (p->N < 4 || (p->N & 1)) return fail(GABE_ERR_ARG, "N must be even and >= 4 (got %d)", p->N);
    if (p->nflds < 1 || p->nflds > GABE_MAX_FLDS) return fail(GABE_ERR_ARG, "nflds must be 1 or 2 (got %d)", p->nflds);
    if (!(p->L > 0.) || !(p->dt > 0.)) return fail(GABE_ERR_ARG, "L and dt must be positive");
    if (p->expansion_type < 0 || p->expansion_type > 2) return fail(GABE_ERR_ARG, "expansion_type must be 0, 1 or 2");
    if (p->precision != GABE_FP64 && p->precision != GABE_FP32) return fail(GABE_ERR_ARG, "precision must be GABE_FP64 or GABE_FP32");
    if ((long)p->N * p->N * p->N > (1L << 31)) return fail(GABE_ERR_ARG, "N^3 beyond 2^31 is not supported");
    if (p->transform_tile != 0 && p->transform_tile != 16 && p->transform_tile != 32)
        return fail(GABE_ERR_ARG, "transform_tile must be 0, 16 or 32 (got %d)", p->transform_tile);
    if (p->stress_chunk < -1) return fail(GABE_ERR_ARG, "stress_chunk must be -1, 0 or a plane count (got %d)", p->stress_chunk);
    const int nt = gabe_gwfft_tiles(p->N, p->precision == GABE_FP32);
    if (p->mode_tile < 0 || p->mode_tile > std::max(nt, 1))
        return fail(GABE_ERR_ARG, "mode_tile must be 0..%d at N=%d (got %d)", std::max(nt, 1), p->N, p->mode_tile);
    return GABE_OK;
}

int create_impl(const gabe_params *params, gabe_ctx *c, hipStream_t shared = nullptr)
{
    c->p = *params;
    c->N = params->N;
    c->nf = params->nflds;
    c->N2 = (long)c->N * c->N;
    c->N3 = c->N2 * c->N;
    c->M = c->N2 * (c->N / 2 + 1);
    c->dx = params->L / (double)c->N;
    c->fp32 = params->precision == GABE_FP32;
    c->esz = c->fp32 ? 4 : 8;
    c->alloc_gw = params->calc_gws != 0;
    c->gw = params->calc_gws && params->expansion_type == 1; /* Q15: only type 1 evolves GWs */
    c->model = gabe_model{params->m2_phi, params->lambda, params->g2, params->m2_chi};

    if (shared) {
        c->stream = shared;
        c->own_stream = false;
    } else {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    }
    if (!shared && !std::getenv("GABE_SINGLE_STREAM")) {
        HIPCHK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&c->ev_fin, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->ev_side, hipEventDisableTiming));
    }
    const size_t fbytes = c->esz * c->nf * c->N3;
    for (int s = 0; s < 2; s++) {
        HIPCHK(hipMalloc(&c->fbuf[s], fbytes));
        HIPCHK(hipMalloc(&c->vbuf[s], fbytes));
        HIPCHK(hipMemset(c->fbuf[s], 0, fbytes));
        HIPCHK(hipMemset(c->vbuf[s], 0, fbytes));
    }
    const int narr = c->alloc_gw ? GABE_NSC : c->nf;
    c->colfft = c->alloc_gw && gabe_colfft_supported(c->N) && !std::getenv("GABE_LIBRARY_3D_FFT");
    c->gwfft = c->colfft && gabe_gwfft_supported(c->N, c->fp32) && !std::getenv("GABE_UNFUSED_GW");
    /* k rows in sweep A: up to GABE_BLOCK sites per row in both forms; 512-site
     * rows (one 512-thread workgroup each) in FP32 */
    const bool tiled = gabe_sweep_tiled(c->N) && !std::getenv("GABE_UNTILED_SWEEPS");
    c->sweep_a = tiled && c->fp32 ? SW_A_TILE : SW_A;
    c->sweep_b = tiled ? SW_B_TILE : SW_B;
    c->rowfft = c->gwfft && c->N >= 16 && (c->N <= GABE_BLOCK || (c->N == 512 && c->fp32)) &&
                !std::getenv("GABE_LIBRARY_ROWS");
    /* Row length of shat.  Sweep A's rows are N/2 slots, kz = 0 and N/2 packed
     * into slot 0, and the j pass moves kz = N/2 into a plane of its own after
     * the GABE_NSC row blocks (N^2 (N/2) + N^2 = M slots per component): every
     * column chunk is full and every row a whole number of cache lines.  The
     * library row plan writes N/2+1 values, padded to 16 for the column
     * passes. */
    c->kpad = c->rowfft ? c->N / 2 : (c->colfft ? ((c->N / 2 + 1 + 15) / 16) * 16 : c->N / 2 + 1);
    c->lay.N = c->N;
    c->lay.KP = c->N / 2 + 1;
    c->lay.nt4 = (int)((c->lay.KP + 3) / 4);
    c->lay.tiled = c->gwfft ? 1 : 0;
    c->mslots = c->lay.tiled ? (long)c->N * c->lay.nt4 * c->N * 4 : c->M;
    {
        const Tiling t = tuned_tiling(c->N, c->nf, c->fp32, c->alloc_gw);
        c->fft_cw = params->transform_tile ? params->transform_tile : (t.cw ? t.cw : 16);
        c->gw_tile = params->mode_tile ? params->mode_tile : (t.mode_tile ? t.mode_tile : 1);
    }
    HIPCHK(hipMalloc(&c->stress, c->esz * narr * c->N3));
    const size_t shat_bytes = 2 * c->esz * narr * std::max(c->M, c->N2 * c->kpad);
    HIPCHK(hipMalloc(&c->shat, shat_bytes));
    c->shat_b[0] = c->shat;
    c->pipe = c->rowfft && c->side != nullptr;
    if (c->pipe) {
        HIPCHK(hipMalloc(&c->shat_b[1], shat_bytes));
        for (int q = 0; q < 2; q++) {
            HIPCHK(hipEventCreateWithFlags(&c->ev_gw[q], hipEventDisableTiming));
            HIPCHK(hipEventRecord(c->ev_gw[q], c->side));
        }
    }
    if (c->alloc_gw) {
        HIPCHK(hipMalloc(&c->h, 2 * c->esz * 2 * c->mslots));
        HIPCHK(hipMalloc(&c->l, 2 * c->esz * 2 * c->mslots));
        HIPCHK(hipMemset(c->h, 0, 2 * c->esz * 2 * c->mslots));
        HIPCHK(hipMemset(c->l, 0, 2 * c->esz * 2 * c->mslots));
    }

    /* sweep tiling: one thread per plane column, ci planes per workgroup */
    const int bx = (int)((c->N2 + GABE_BLOCK - 1) / GABE_BLOCK);
    int ci = params->planes_per_block;
    if (ci <= 0) ci = tuned_tiling(c->N, c->nf, c->fp32, c->alloc_gw).planes;
    if (ci <= 0) {
        const int by_target = std::max(1, 2048 / bx);
        ci = (c->N + by_target - 1) / by_target;
    }
    ci = std::min(std::max(ci, 1), c->N);
    const int by = (c->N + ci - 1) / ci;
    c->ci = ci;
    c->grid = dim3(bx, by);
    c->nblk = bx * by;
    if (c->pipe && c->gw) {   /* interleaved sweep A / j-transform launches */
        int planes = params->stress_chunk;
        if (planes == 0) planes = tuned_tiling(c->N, c->nf, c->fp32, c->alloc_gw).chunk;
        const int bpc = planes > 0 ? planes / ci : 0;
        if (bpc >= 1 && bpc < by) {
            c->chunk = bpc;
            HIPCHK(hipStreamCreateWithFlags(&c->jst, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&c->ev_jdone, hipEventDisableTiming));
            c->ev_chunk.assign((size_t)((by + bpc - 1) / bpc), nullptr);
            for (hipEvent_t &e : c->ev_chunk) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
    }
    HIPCHK(hipMalloc(&c->partials, sizeof(double) * GABE_MAXSUMS * std::max(c->nblk, 4096)));

    HIPCHK(hipMalloc(&c->sc, sizeof(DevScalars)));
    HIPCHK(hipMalloc(&c->ss, sizeof(StepScalars)));
    c->ss_b[0] = c->ss;
    if (c->pipe) {
        HIPCHK(hipMalloc(&c->ss_b[1], sizeof(StepScalars)));
        HIPCHK(hipMemset(c->ss_b[1], 0, sizeof(StepScalars)));
    }
    DevScalars d{};
    d.a[0] = d.a[1] = 1.;
    HIPCHK(hipMemcpy(c->sc, &d, sizeof d, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->ss, 0, sizeof(StepScalars)));
    c->hist_cap = 1 << 16;
    HIPCHK(hipMalloc(&c->hist, sizeof(double) * GABE_HIST_COLS * c->hist_cap));
    HIPCHK(hipMemset(c->hist, 0, sizeof(double) * GABE_HIST_COLS * c->hist_cap));
    const int nb = gabe_numbins(c->N);
    HIPCHK(hipMalloc(&c->spec_out, sizeof(double) * nb * (GABE_MAXF + 2)));
    HIPCHK(hipMalloc(&c->spec_part, sizeof(double) * nb * gabe_spec_chunks(c->N) * (GABE_MAXF + 2)));

    if (c->alloc_gw) {
        if (c->colfft) {
            if (!c->rowfft)
                CHK(make_plan(c, c->fft_stress, rocfft_transform_type_real_forward, !c->fp32, GABE_NSC * (size_t)c->N2, 1,
                              (size_t)c->kpad));
            std::vector<double> tw(2 * c->N);
            for (int e = 0; e < c->N; e++) {
                tw[2 * e] = gabe_cos2pi_frac(e, c->N);
                tw[2 * e + 1] = -gabe_sin2pi_frac(e, c->N);
            }
            HIPCHK(hipMalloc(&c->twiddle, 2 * c->esz * c->N));
            if (c->fp32) {
                std::vector<float> twf(tw.begin(), tw.end());
                HIPCHK(hipMemcpy(c->twiddle, twf.data(), 2 * c->esz * c->N, hipMemcpyHostToDevice));
            } else {
                HIPCHK(hipMemcpy(c->twiddle, tw.data(), 2 * c->esz * c->N, hipMemcpyHostToDevice));
            }
        } else {
            CHK(make_plan(c, c->fft_stress, rocfft_transform_type_real_forward, !c->fp32, GABE_NSC));
        }
    }
    CHK(make_plan(c, c->fft_spec, rocfft_transform_type_real_forward, !c->fp32, c->nf));
    return GABE_OK;
}

void destroy_impl(gabe_ctx *c)
{
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->closure) closure_release(c);
    destroy_plan(c->fft_stress);
    destroy_plan(c->fft_spec);
    destroy_plan(c->fft_c2r);
    destroy_plan(c->fft_rx);
    void *bufs[] = {c->fbuf[0], c->fbuf[1], c->vbuf[0], c->vbuf[1], c->stress, c->shat_b[0], c->shat_b[1], c->h,
                    c->l, c->partials, c->sc, c->ss_b[0], c->ss_b[1], c->hist, c->spec_out, c->spec_part, c->fft_work,
                    c->twiddle, c->rx_buf};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->jst) (void)hipStreamSynchronize(c->jst);
    for (hipEvent_t e : c->ev_chunk)
        if (e) (void)hipEventDestroy(e);
    if (c->ev_jdone) (void)hipEventDestroy(c->ev_jdone);
    if (c->jst) (void)hipStreamDestroy(c->jst);
    for (hipEvent_t e : c->ev_gw)
        if (e) (void)hipEventDestroy(e);
    if (c->ev_fin) (void)hipEventDestroy(c->ev_fin);
    if (c->ev_side) (void)hipEventDestroy(c->ev_side);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
}

/* host [nf][N3] doubles -> device storage of slot cur (split into mean + fluctuation for FP32) */
int upload_fields(gabe_ctx *c, void *dst, const double *src, double *means)
{
    const long n3 = c->N3;
    if (c->fp32) {
        std::vector<float> tmp(n3);
        for (int f = 0; f < c->nf; f++) {
            const double *x = src + (long)f * n3;
            const double mean = csum(x, n3) / (double)n3;
            for (long n = 0; n < n3; n++) tmp[n] = (float)(x[n] - mean);
            HIPCHK(hipMemcpy((float *)dst + (long)f * n3, tmp.data(), sizeof(float) * n3, hipMemcpyHostToDevice));
            means[f] = mean;
        }
    } else {
        HIPCHK(hipMemcpy(dst, src, sizeof(double) * c->nf * n3, hipMemcpyHostToDevice));
        for (int f = 0; f < c->nf; f++) means[f] = 0.;
    }
    return GABE_OK;
}

int download_fields(gabe_ctx *c, const void *src, double *dst, const double *means)
{
    const long n3 = c->N3;
    if (c->fp32) {
        std::vector<float> tmp(n3);
        for (int f = 0; f < c->nf; f++) {
            HIPCHK(hipMemcpy(tmp.data(), (const float *)src + (long)f * n3, sizeof(float) * n3, hipMemcpyDeviceToHost));
            for (long n = 0; n < n3; n++) dst[(long)f * n3 + n] = means[f] + (double)tmp[n];
        }
    } else {
        HIPCHK(hipMemcpy(dst, src, sizeof(double) * c->nf * n3, hipMemcpyDeviceToHost));
    }
    return GABE_OK;
}

/* nsteps steps, then the side stream's work joined into the main stream, so
 * every later synchronisation of the main stream covers both */
int run_steps(gabe_ctx *c, int nsteps)
{
    for (int n = 0; n < nsteps; n++) CHK(one_step(c, nullptr));
    if (c->pipe)
        for (hipEvent_t e : c->ev_gw) HIPCHK(hipStreamWaitEvent(c->stream, e, 0));
    return GABE_OK;
}

} // namespace gabe_detail

/* ====================================================================== ABI */
extern "C" {

const char *gabe_last_error(void) { return g_err.c_str(); }

int gabe_numbins(int N) { return (int)(std::sqrt(3.) * N / 2 + 1); }

int gabe_create(const gabe_params *params, gabe_ctx **out)
{
    if (!out) return fail(GABE_ERR_ARG, "out is NULL");
    *out = nullptr;
    CHK(check_params(params));
    gabe_ctx *c = new gabe_ctx;
    int r = create_impl(params, c);
    if (r != GABE_OK) {
        std::string keep = g_err;
        destroy_impl(c);
        delete c;
        g_err = keep;
        return r;
    }
    *out = c;
    return GABE_OK;
}

int gabe_destroy(gabe_ctx *c)
{
    if (!c) return GABE_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    destroy_impl(c);
    delete c;
    return GABE_OK;
}

int gabe_upload_state(gabe_ctx *c, const double *field, const double *dfield)
{
    if (!c || !field || !dfield) return fail(GABE_ERR_ARG, "NULL argument");
    DevScalars d;
    CHK(read_scalars(c, d));
    CHK(upload_fields(c, c->fbuf[c->cur], field, d.pbar));
    CHK(upload_fields(c, c->vbuf[c->cur], dfield, d.vbar));
    d.e0_valid = 0;
    CHK(write_scalars(c, d));
    return GABE_OK;
}

int gabe_download_state(gabe_ctx *c, double *field, double *dfield)
{
    if (!c) return fail(GABE_ERR_ARG, "NULL context");
    DevScalars d;
    CHK(read_scalars(c, d));
    if (field) CHK(download_fields(c, c->fbuf[c->cur], field, d.pbar));
    if (dfield) CHK(download_fields(c, c->vbuf[c->cur], dfield, d.vbar));
    return GABE_OK;
}

int gabe_download_plane(gabe_ctx *c, int i, double *field, double *dfield)
{
    if (!c) return fail(GABE_ERR_ARG, "NULL context");
    if (i < 0 || i >= c->N) return fail(GABE_ERR_ARG, "plane index %d outside [0, %d)", i, c->N);
    DevScalars d;
    CHK(read_scalars(c, d));
    const long N2 = c->N2, off = (long)i * N2;
    const void *src[2] = {c->fbuf[c->cur], c->vbuf[c->cur]};
    double *dst[2] = {field, dfield};
    const double *means[2] = {d.pbar, d.vbar};
    std::vector<float> tmp(c->fp32 ? N2 : 0);
    for (int a = 0; a < 2; a++) {
        if (!dst[a]) continue;
        for (int f = 0; f < c->nf; f++) {
            double *out = dst[a] + (long)f * N2;
            if (c->fp32) {
                HIPCHK(hipMemcpy(tmp.data(), (const float *)src[a] + (long)f * c->N3 + off, sizeof(float) * N2,
                                 hipMemcpyDeviceToHost));
                for (long n = 0; n < N2; n++) out[n] = means[a][f] + (double)tmp[n];
            } else {
                HIPCHK(hipMemcpy(out, (const double *)src[a] + (long)f * c->N3 + off, sizeof(double) * N2,
                                 hipMemcpyDeviceToHost));
            }
        }
    }
    return GABE_OK;
}

int gabe_means_vars(gabe_ctx *c, double *out)
{
    if (!c || !out) return fail(GABE_ERR_ARG, "NULL argument");
    DevScalars d;
    CHK(read_scalars(c, d));
    const int nblk = 1024;
    const long N3 = c->N3;
    std::vector<double> part(2 * nblk);
    for (int f = 0; f < c->nf; f++) {
        double shift = 0.;
        float shift32 = 0.f;
        if (c->fp32) {
            const float *x = (const float *)c->fbuf[c->cur] + (long)f * N3;
            launch_meanvar<float>(x, N3, c->partials, nblk, c->stream);
            HIPCHK(hipMemcpyAsync(&shift32, x, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        } else {
            const double *x = (const double *)c->fbuf[c->cur] + (long)f * N3;
            launch_meanvar<double>(x, N3, c->partials, nblk, c->stream);
            HIPCHK(hipMemcpyAsync(&shift, x, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(part.data(), c->partials, sizeof(double) * 2 * nblk, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->fp32) shift = (double)shift32;
        const double m = csum(part.data(), nblk) / (double)N3;
        const double base = c->fp32 ? d.pbar[f] : 0.;
        out[2 * f] = base + shift + m;
        out[2 * f + 1] = csum(part.data() + nblk, nblk) / (double)N3 - m * m;
    }
    return GABE_OK;
}

/* Checkpoint blob: a header, then the device scalars, the history rows of the
 * current trajectory (oldest first), the current slot's field and velocity
 * arrays, with tensor modes h and l (in the storage order the header names),
 * and the attached learned closure's state (gate, fallback step, largest
 * residual, threshold, scales, weights) if one is attached, all in the
 * device's own storage precision, so a restored context continues bit for
 * bit.  Loading a blob without a closure section detaches any closure. */
struct CkptHeader {
    char magic[8];
    unsigned version;
    int N, nflds, precision, calc_gws, expansion_type;
    double L, dt, m2_phi, lambda, g2, m2_chi;
    int gw_spectrum_quirk, hist_rows;
    int mode_tiled, mode_nt4;     /* storage order of the tensor-mode pairs (ModeLayout) */
    long long mode_kp;
    unsigned long long state_bytes, gw_bytes, closure_bytes, total;
};
static const char kCkptMagic[8] = {'G', 'A', 'B', 'E', 'S', 'T', 'A', 'T'};

static CkptHeader ckpt_header(const gabe_ctx *c, int hist_rows, unsigned long long closure_bytes)
{
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, 8);
    h.version = 2;
    h.N = c->N;
    h.nflds = c->nf;
    h.precision = c->p.precision;
    h.calc_gws = c->p.calc_gws;
    h.expansion_type = c->p.expansion_type;
    h.L = c->p.L;
    h.dt = c->p.dt;
    h.m2_phi = c->p.m2_phi;
    h.lambda = c->p.lambda;
    h.g2 = c->p.g2;
    h.m2_chi = c->p.m2_chi;
    h.gw_spectrum_quirk = c->p.gw_spectrum_quirk;
    h.hist_rows = hist_rows;
    h.mode_tiled = c->alloc_gw ? c->lay.tiled : 0;
    h.mode_nt4 = c->alloc_gw ? c->lay.nt4 : 0;
    h.mode_kp = c->alloc_gw ? c->lay.KP : 0;
    h.state_bytes = (unsigned long long)c->nf * c->N3 * c->esz;
    h.gw_bytes = c->alloc_gw ? 2ULL * c->esz * 2 * c->mslots : 0;
    h.closure_bytes = closure_bytes;
    h.total = sizeof(CkptHeader) + sizeof(DevScalars) + sizeof(double) * GABE_HIST_COLS * (unsigned long long)hist_rows +
              2 * h.state_bytes + 2 * h.gw_bytes + closure_bytes;
    return h;
}

/* history rows of the current trajectory held in the ring */
static int hist_rows_of(const gabe_ctx *c, const DevScalars &d)
{
    return (int)std::min<long long>(std::max<long long>(0, d.nstep - d.hist0), c->hist_cap);
}

int gabe_checkpoint_size(gabe_ctx *c, size_t *bytes)
{
    if (!c || !bytes) return fail(GABE_ERR_ARG, "NULL argument");
    DevScalars d;
    CHK(read_scalars(c, d));
    *bytes = ckpt_header(c, hist_rows_of(c, d), closure_ckpt_bytes(c)).total;
    return GABE_OK;
}

int gabe_checkpoint_save(gabe_ctx *c, void *buf, size_t bytes)
{
    if (!c || !buf) return fail(GABE_ERR_ARG, "NULL argument");
    DevScalars d;
    CHK(read_scalars(c, d));
    const int rows = hist_rows_of(c, d);
    const CkptHeader h = ckpt_header(c, rows, closure_ckpt_bytes(c));
    if (bytes < h.total) return fail(GABE_ERR_ARG, "checkpoint buffer holds %zu bytes, %llu needed", bytes, h.total);
    char *p = (char *)buf;
    std::memcpy(p, &h, sizeof h);
    p += sizeof h;
    std::memcpy(p, &d, sizeof d);
    p += sizeof d;
    if (rows) {
        std::vector<double> ring((size_t)c->hist_cap * GABE_HIST_COLS);
        HIPCHK(hipMemcpy(ring.data(), c->hist, sizeof(double) * ring.size(), hipMemcpyDeviceToHost));
        for (int r = 0; r < rows; r++) {
            const long stepno = (long)d.nstep - rows + r;
            std::memcpy(p, ring.data() + (stepno % c->hist_cap) * GABE_HIST_COLS, sizeof(double) * GABE_HIST_COLS);
            p += sizeof(double) * GABE_HIST_COLS;
        }
    }
    for (const void *src : {c->fbuf[c->cur], c->vbuf[c->cur]}) {
        HIPCHK(hipMemcpy(p, src, h.state_bytes, hipMemcpyDeviceToHost));
        p += h.state_bytes;
    }
    if (c->alloc_gw)
        for (const void *src : {(const void *)c->h, (const void *)c->l}) {
            HIPCHK(hipMemcpy(p, src, h.gw_bytes, hipMemcpyDeviceToHost));
            p += h.gw_bytes;
        }
    CHK(closure_ckpt_save(c, p));
    return GABE_OK;
}

int gabe_checkpoint_load(gabe_ctx *c, const void *buf, size_t bytes)
{
    if (!c || !buf) return fail(GABE_ERR_ARG, "NULL argument");
    if (bytes < sizeof(CkptHeader)) return fail(GABE_ERR_ARG, "checkpoint shorter than its header");
    CkptHeader h;
    std::memcpy(&h, buf, sizeof h);
    if (std::memcmp(h.magic, kCkptMagic, 8) != 0) return fail(GABE_ERR_ARG, "not a gabe checkpoint");
    if (h.version != 2) return fail(GABE_ERR_ARG, "checkpoint format version %u; this library reads version 2", h.version);
    const CkptHeader mine = ckpt_header(c, h.hist_rows, h.closure_bytes);
    if (h.N != mine.N || h.nflds != mine.nflds || h.precision != mine.precision || h.calc_gws != mine.calc_gws ||
        h.expansion_type != mine.expansion_type || h.L != mine.L || h.dt != mine.dt || h.m2_phi != mine.m2_phi ||
        h.lambda != mine.lambda || h.g2 != mine.g2 || h.m2_chi != mine.m2_chi ||
        h.gw_spectrum_quirk != mine.gw_spectrum_quirk)
        return fail(GABE_ERR_ARG, "checkpoint was written by a context with other parameters (N=%d nflds=%d "
                                  "precision=%d calc_gws=%d)", h.N, h.nflds, h.precision, h.calc_gws);
    if (h.mode_tiled != mine.mode_tiled || h.mode_nt4 != mine.mode_nt4 || h.mode_kp != mine.mode_kp)
        return fail(GABE_ERR_ARG, "checkpoint stores its tensor modes in another order (%s, %lld slots per row) than "
                                  "this lattice (%s, %lld): it was written with other GW transform settings",
                    h.mode_tiled ? "tiled" : "row-major", h.mode_kp, mine.mode_tiled ? "tiled" : "row-major",
                    mine.mode_kp);
    if (h.hist_rows < 0 || h.hist_rows > c->hist_cap || h.total != mine.total || bytes < h.total)
        return fail(GABE_ERR_ARG, "checkpoint size does not match its header");
    const char *p = (const char *)buf + sizeof h;
    DevScalars d;
    std::memcpy(&d, p, sizeof d);
    p += sizeof d;
    if (h.hist_rows != hist_rows_of(c, d)) return fail(GABE_ERR_ARG, "checkpoint history does not match its step count");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (h.hist_rows) {
        std::vector<double> ring((size_t)c->hist_cap * GABE_HIST_COLS, 0.);
        for (int r = 0; r < h.hist_rows; r++) {
            const long stepno = (long)d.nstep - h.hist_rows + r;
            std::memcpy(ring.data() + (stepno % c->hist_cap) * GABE_HIST_COLS, p, sizeof(double) * GABE_HIST_COLS);
            p += sizeof(double) * GABE_HIST_COLS;
        }
        HIPCHK(hipMemcpy(c->hist, ring.data(), sizeof(double) * ring.size(), hipMemcpyHostToDevice));
    }
    for (void *dst : {c->fbuf[c->cur], c->vbuf[c->cur]}) {
        HIPCHK(hipMemcpy(dst, p, h.state_bytes, hipMemcpyHostToDevice));
        p += h.state_bytes;
    }
    if (c->alloc_gw)
        for (void *dst : {c->h, c->l}) {
            HIPCHK(hipMemcpy(dst, p, h.gw_bytes, hipMemcpyHostToDevice));
            p += h.gw_bytes;
        }
    CHK(closure_ckpt_load(c, p, h.closure_bytes));
    CHK(write_scalars(c, d));
    return GABE_OK;
}

int gabe_upload_gw(gabe_ctx *c, const double *h, const double *l)
{
    if (!c) return fail(GABE_ERR_ARG, "NULL context");
    if (!c->alloc_gw) return fail(GABE_ERR_STATE, "context created with calc_gws = 0");
    HIPCHK(hipStreamSynchronize(c->stream));
    const long M = c->M;
    double *tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, sizeof(double) * 12 * M));
    const double *src[2] = {h, l};
    void *dst[2] = {c->h, c->l};
    int rc = GABE_OK;
    for (int a = 0; a < 2 && rc == GABE_OK; a++) {
        if (!src[a]) continue;
        if (hipMemcpy(tmp, src[a], sizeof(double) * 12 * M, hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(GABE_ERR_HIP, "upload of tensor modes failed");
            break;
        }
        if (c->fp32) launch_gw_pack<float>(tmp, (float *)dst[a], c->lay, M, c->stream);
        else launch_gw_pack<double>(tmp, (double *)dst[a], c->lay, M, c->stream);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(GABE_ERR_HIP, "tensor mode packing failed");
    }
    (void)hipFree(tmp);
    return rc;
}

int gabe_download_gw(gabe_ctx *c, double *h, double *l)
{
    if (!c) return fail(GABE_ERR_ARG, "NULL context");
    if (!c->alloc_gw) return fail(GABE_ERR_STATE, "context created with calc_gws = 0");
    HIPCHK(hipStreamSynchronize(c->stream));
    const long M = c->M;
    double *tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, sizeof(double) * 12 * M));
    double *dst[2] = {h, l};
    const void *src[2] = {c->h, c->l};
    int rc = GABE_OK;
    for (int a = 0; a < 2 && rc == GABE_OK; a++) {
        if (!dst[a]) continue;
        if (c->fp32) launch_gw_unpack<float>((const float *)src[a], tmp, c->lay, M, c->stream);
        else launch_gw_unpack<double>((const double *)src[a], tmp, c->lay, M, c->stream);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess ||
            hipMemcpy(dst[a], tmp, sizeof(double) * 12 * M, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(GABE_ERR_HIP, "tensor mode download failed");
    }
    (void)hipFree(tmp);
    return rc;
}

int gabe_set_scalars(gabe_ctx *c, const gabe_scalars *s)
{
    if (!c || !s) return fail(GABE_ERR_ARG, "NULL argument");
    DevScalars d;
    CHK(read_scalars(c, d));
    for (int k = 0; k < 2; k++) {
        d.a[k] = s->a[k];
        d.adot[k] = s->adot[k];
        d.edkin[k] = s->edkin[k];
        d.edpot[k] = s->edpot[k];
        d.edgrad[k] = s->edgrad[k];
        d.edrho[k] = s->edrho[k];
    }
    d.t = s->t;
    d.nstep = s->nstep;
    d.hist0 = s->nstep;   /* rows recorded before describe another trajectory */
    d.adot0_valid = 1;
    d.e0_valid = 1;
    d.error = 0;
    CHK(write_scalars(c, d));
    return GABE_OK;
}

int gabe_get_scalars(gabe_ctx *c, gabe_scalars *s)
{
    if (!c || !s) return fail(GABE_ERR_ARG, "NULL argument");
    CHK(refresh_e0(c));
    DevScalars d;
    CHK(read_scalars(c, d));
    for (int k = 0; k < 2; k++) {
        s->a[k] = d.a[k];
        s->adot[k] = d.adot[k];
        s->edkin[k] = d.edkin[k];
        s->edpot[k] = d.edpot[k];
        s->edgrad[k] = d.edgrad[k];
        s->edrho[k] = d.edrho[k];
    }
    s->t = d.t;
    s->nstep = d.nstep;
    s->adot0_valid = d.adot0_valid;
    s->error = d.error;
    if (d.error) return fail(GABE_ERR_NONFINITE, "energy density became non-finite (unstable solution)");
    return GABE_OK;
}

int gabe_energy(gabe_ctx *c, int s)
{
    if (!c) return fail(GABE_ERR_ARG, "NULL context");
    if (s != 0) return fail(GABE_ERR_ARG, "only slot 0 holds a state between steps");
    CHK(launch_sweep_ctx(c, SW_ENERGY, false));
    CHK(launch_finalize_ctx(c, SW_ENERGY, 0));
    return GABE_OK;
}

int gabe_step(gabe_ctx *c, int nsteps)
{
    if (!c) return fail(GABE_ERR_ARG, "NULL context");
    if (nsteps < 0) return fail(GABE_ERR_ARG, "nsteps < 0");
    return run_steps(c, nsteps);
}

int gabe_sync(gabe_ctx *c)
{
    if (!c) return fail(GABE_ERR_ARG, "NULL context");
    DevScalars d;
    CHK(read_scalars(c, d));
    if (d.error) return fail(GABE_ERR_NONFINITE, "energy density became non-finite (unstable solution)");
    return GABE_OK;
}

int gabe_spectra(gabe_ctx *c, double *out)
{
    if (!c || !out) return fail(GABE_ERR_ARG, "NULL argument");
    CHK(refresh_e0(c));
    DevScalars d;
    CHK(read_scalars(c, d));
    const int nf = c->nf, N = c->N, nb = gabe_numbins(N);
    HIPCHK(hipMemcpyAsync(c->stress, c->fbuf[c->cur], c->esz * nf * c->N3, hipMemcpyDeviceToDevice, c->stream));
    CHK(run_plan(c, c->fft_spec, c->stress, c->shat));
    if (c->fp32) {
        SpecArgs<float> a{};
        a.fhat = (const float *)c->shat;
        a.l = (const float *)c->l;
        for (int f = 0; f < nf; f++) a.dc[f] = d.pbar[f] * (double)c->N3;
        a.nf = nf; a.N = N; a.numbins = nb; a.quirk = c->p.gw_spectrum_quirk; a.M = c->M; a.out = c->spec_out;
        a.part = c->spec_part;
        a.nch = gabe_spec_chunks(N);
        a.lay = c->lay;
        launch_spectra<float>(a, c->stream);
    } else {
        SpecArgs<double> a{};
        a.fhat = (const double *)c->shat;
        a.l = (const double *)c->l;
        a.nf = nf; a.N = N; a.numbins = nb; a.quirk = c->p.gw_spectrum_quirk; a.M = c->M; a.out = c->spec_out;
        a.part = c->spec_part;
        a.nch = gabe_spec_chunks(N);
        a.lay = c->lay;
        launch_spectra<double>(a, c->stream);
    }
    HIPCHK(hipGetLastError());
    std::vector<double> raw((size_t)nb * (nf + 2));
    HIPCHK(hipMemcpyAsync(raw.data(), c->spec_out, sizeof(double) * raw.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));

    /* normalisation of g2spectra.cpp:27,94-103,114-117,186-195 */
    const double L = c->p.L;
    const double spec_norm = std::pow(L / (double)N, 6.) * std::pow(2. * M_PI / L, 3.);
    for (int f = 0; f < nf; f++) {
        for (int b = 0; b < nb; b++) {
            const double sum = raw[(size_t)b * (nf + 2) + f], cnt = raw[(size_t)b * (nf + 2) + nf];
            if (cnt > 0) {
                out[(2 * f) * nb + b] = sum * spec_norm / cnt;
                out[(2 * f + 1) * nb + b] = std::pow((double)b * 2. * M_PI / L, 3) * sum * spec_norm / cnt / 2. / M_PI / M_PI;
            } else {
                out[(2 * f) * nb + b] = 0.;
                out[(2 * f + 1) * nb + b] = 0.;
            }
        }
    }
    if (c->alloc_gw) {
        const double hubtemp = std::sqrt(8. * M_PI / 3.) * std::sqrt(d.edrho[0]);
        const double norm1_gw = 4.e-5 / std::pow(100., .333) / 24. / M_PI / M_PI / std::pow(L, 3.) / (hubtemp * hubtemp);
        const int ncol = 2 * nf + 2;
        for (int b = 0; b < nb; b++) {
            const double sum = raw[(size_t)b * (nf + 2) + nf + 1], cnt = raw[(size_t)b * (nf + 2) + nf];
            if (cnt > 0) {
                out[(2 * nf) * nb + b] = sum * spec_norm / cnt;
                out[(2 * nf + 1) * nb + b] = norm1_gw * std::pow((double)b * 2 * M_PI / L, 3.) * sum / cnt;
            } else {
                out[(2 * nf) * nb + b] = 0.;
                out[(2 * nf + 1) * nb + b] = 0.;
            }
            out[ncol * nb + b] = 2. * M_PI / L * (double)b;
        }
    }
    return GABE_OK;
}

int gabe_add_fluctuation(gabe_ctx *c, int fld, const double *fk, const double *fkd)
{
    if (!c || !fk || !fkd) return fail(GABE_ERR_ARG, "NULL argument");
    if (fld < 0 || fld >= c->nf) return fail(GABE_ERR_ARG, "field index %d out of range", fld);
    const int N = c->N, H = N / 2 + 1;
    const long M = c->M, N3 = c->N3;
    if (!c->fft_c2r.plan) CHK(make_plan(c, c->fft_c2r, rocfft_transform_type_real_inverse, true, 1));
    HIPCHK(hipStreamSynchronize(c->stream));
    double *din = nullptr, *dout = nullptr;
    HIPCHK(hipMalloc(&din, sizeof(double) * 2 * M));
    HIPCHK(hipMalloc(&dout, sizeof(double) * N3));
    std::vector<double> sym(2 * M);
    const double *srcs[2] = {fk, fkd};
    void *dsts[2] = {c->fbuf[c->cur], c->vbuf[c->cur]};
    int rc = GABE_OK;
    for (int a = 0; a < 2 && rc == GABE_OK; a++) {
        const double *X = srcs[a];
        std::memcpy(sym.data(), X, sizeof(double) * 2 * M);
        /* Hermitian part of the kz = 0 and kz = N/2 planes: X <- (X(k) + conj X(-k)) / 2 */
        for (int kz : {0, N / 2})
            for (int i = 0; i < N; i++)
                for (int j = 0; j < N; j++) {
                    const long m = ((long)i * N + j) * H + kz;
                    const long mc = ((long)((N - i) % N) * N + (N - j) % N) * H + kz;
                    sym[2 * m] = 0.5 * (X[2 * m] + X[2 * mc]);
                    sym[2 * m + 1] = 0.5 * (X[2 * m + 1] - X[2 * mc + 1]);
                }
        if (hipMemcpy(din, sym.data(), sizeof(double) * 2 * M, hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(GABE_ERR_HIP, "upload of modes failed");
            break;
        }
        rc = run_plan(c, c->fft_c2r, din, dout);
        if (rc != GABE_OK) break;
        if (c->fp32) launch_add_real<float>((float *)dsts[a] + (long)fld * N3, dout, N3, c->stream);
        else launch_add_real<double>((double *)dsts[a] + (long)fld * N3, dout, N3, c->stream);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(GABE_ERR_HIP, "fluctuation add failed");
    }
    (void)hipFree(din);
    (void)hipFree(dout);
    if (rc != GABE_OK) return rc;
    DevScalars d;
    CHK(read_scalars(c, d));
    d.e0_valid = 0;
    CHK(write_scalars(c, d));
    return GABE_OK;
}

int gabe_effmass(gabe_ctx *c, int fld, double *out)
{
    if (!c || !out) return fail(GABE_ERR_ARG, "NULL argument");
    if (fld < 0 || fld >= c->nf) return fail(GABE_ERR_ARG, "field index %d out of range", fld);
    DevScalars d;
    CHK(read_scalars(c, d));
    const int nblk = 1024;
    const long N3 = c->N3;
    if (c->fp32) {
        const float *f0 = (const float *)c->fbuf[c->cur];
        launch_effmass<float>(f0, c->nf > 1 ? f0 + N3 : f0, c->nf, fld, d.pbar[0], c->nf > 1 ? d.pbar[1] : 0., c->model,
                              N3, c->partials, nblk, c->stream);
    } else {
        const double *f0 = (const double *)c->fbuf[c->cur];
        launch_effmass<double>(f0, c->nf > 1 ? f0 + N3 : f0, c->nf, fld, 0., 0., c->model, N3, c->partials, nblk,
                               c->stream);
    }
    HIPCHK(hipGetLastError());
    std::vector<double> part(nblk);
    HIPCHK(hipMemcpyAsync(part.data(), c->partials, sizeof(double) * nblk, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *out = csum(part.data(), nblk) / (double)N3;
    return GABE_OK;
}

int gabe_tt_source(gabe_ctx *c, int s, double *out)
{
    if (!c || !out) return fail(GABE_ERR_ARG, "NULL argument");
    if (!c->alloc_gw) return fail(GABE_ERR_STATE, "context created with calc_gws = 0");
    if (s != 0 && s != 1) return fail(GABE_ERR_ARG, "slot must be 0 or 1");
    double *dout = nullptr;
    HIPCHK(hipMalloc(&dout, sizeof(double) * 12 * c->M));
    int rc = launch_sweep_ctx(c, SW_A, true);
    if (rc == GABE_OK) rc = stress_fft(c);
    if (rc == GABE_OK) {
        const double dx3 = c->dx * c->dx * c->dx;
        if (c->fp32)
            launch_tt_out<float>((const float *)c->shat, (const float *)nyq_plane(c), s, c->N, c->M, c->kpad, dx3,
                                 dout, c->stream);
        else
            launch_tt_out<double>((const double *)c->shat, (const double *)nyq_plane(c), s, c->N, c->M, c->kpad,
                                  dx3, dout, c->stream);
        if (hipGetLastError() != hipSuccess) rc = fail(GABE_ERR_HIP, "tt kernel launch failed");
    }
    if (rc == GABE_OK && hipMemcpyAsync(out, dout, sizeof(double) * 12 * c->M, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = fail(GABE_ERR_HIP, "download of T_gw failed");
    if (hipStreamSynchronize(c->stream) != hipSuccess && rc == GABE_OK) rc = fail(GABE_ERR_HIP, "sync failed");
    (void)hipFree(dout);
    return rc;
}

int gabe_history(gabe_ctx *c, double *out, int max_rows)
{
    if (!c || !out || max_rows < 0) return fail(GABE_ERR_ARG, "bad argument");
    DevScalars d;
    CHK(read_scalars(c, d));
    std::vector<double> h((size_t)c->hist_cap * GABE_HIST_COLS);
    HIPCHK(hipMemcpy(h.data(), c->hist, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    const long total = (long)d.nstep;
    const long avail = std::min<long>(std::max<long long>(0, d.nstep - d.hist0), c->hist_cap);
    const long nrow = std::min<long>(avail, max_rows);
    for (long r = 0; r < nrow; r++) {
        const long stepno = total - nrow + r;
        std::memcpy(out + r * GABE_HIST_COLS, h.data() + (stepno % c->hist_cap) * GABE_HIST_COLS,
                    sizeof(double) * GABE_HIST_COLS);
    }
    return (int)nrow;
}

int gabe_kernel_times(gabe_ctx *c, int nsteps, double *ms_out)
{
    if (!c || !ms_out || nsteps <= 0) return fail(GABE_ERR_ARG, "bad argument");
    const int nev = 2 * GABE_NKCLASS;
    std::vector<hipEvent_t> ev((size_t)nev * nsteps);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    int rc = GABE_OK;
    for (int n = 0; n < nsteps && rc == GABE_OK; n++) rc = one_step(c, ev.data() + (size_t)n * nev);
    if (rc == GABE_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(GABE_ERR_HIP, "sync failed");
    for (int k = 0; k < GABE_NKCLASS; k++) ms_out[k] = 0.;
    if (rc == GABE_OK) {
        for (int n = 0; n < nsteps; n++)
            for (int k = 0; k < GABE_NKCLASS; k++) {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, ev[(size_t)n * nev + 2 * k], ev[(size_t)n * nev + 2 * k + 1]);
                ms_out[k] += ms / nsteps;
            }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

int gabe_closure_kernel_times(gabe_ctx *c, int nsteps, double *ms_out)
{
    if (!c || !ms_out || nsteps <= 0) return fail(GABE_ERR_ARG, "bad argument");
    if (!c->closure) return fail(GABE_ERR_STATE, "no closure is attached to this lattice");
    const int nev = 2 * GABE_NKCLASS, ncl = 2 * GABE_NCLCLASS;
    std::vector<hipEvent_t> ev((size_t)nev * nsteps), cev((size_t)ncl * nsteps);
    int rc = GABE_OK;
    for (auto &e : ev)
        if (rc == GABE_OK && hipEventCreate(&e) != hipSuccess) rc = fail(GABE_ERR_HIP, "event creation failed");
    for (auto &e : cev)
        if (rc == GABE_OK && hipEventCreate(&e) != hipSuccess) rc = fail(GABE_ERR_HIP, "event creation failed");
    for (int n = 0; n < nsteps && rc == GABE_OK; n++)
        rc = one_step(c, ev.data() + (size_t)n * nev, cev.data() + (size_t)n * ncl);
    if (rc == GABE_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(GABE_ERR_HIP, "sync failed");
    for (int k = 0; k < GABE_NCLCLASS; k++) ms_out[k] = 0.;
    if (rc == GABE_OK) {
        for (int n = 0; n < nsteps; n++)
            for (int k = 0; k < GABE_NCLCLASS; k++) {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, cev[(size_t)n * ncl + 2 * k], cev[(size_t)n * ncl + 2 * k + 1]);
                ms_out[k] += ms / nsteps;
            }
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : cev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

} // extern "C"

struct gabe_ensemble {
    int count = 0, nf = 0;
    bool fp32 = false;
    hipStream_t stream = nullptr;
    std::vector<gabe_ctx *> members;
    void *d_desc[2] = {nullptr, nullptr};   /* SweepArgs<T>[count] for state parity 0 / 1 */
    FinArgs *d_fin = nullptr;
    /* the members with LDS-tiled sweeps (N a multiple of 16) come first
     * in the descriptor tables and run in launches of their own, so every
     * member takes the sweeps it would take stepped alone: order[q] is the
     * member of descriptor q, d_start holds the workgroup starts of the ntile
     * tiled members, then those of the others */
    std::vector<int> order;
    int ntile = 0;
    int *d_start = nullptr;
    int nblk_tile = 0, nblk_rest = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};   /* two steps starting at parity 0 / 1 */
};

namespace {

/* one batched sweep phase (sweep A, or sweep B when b) of every member:
 * the tiled members in one launch, the others in a second */
template <typename T>
void ens_sweep(gabe_ensemble *e, int parity, bool b)
{
    const auto *d = (const SweepArgs<T> *)e->d_desc[parity];
    const int nt = e->ntile, nr = e->count - e->ntile;
    if (nt) {
        const gabe_ctx *c = e->members[e->order[0]];
        launch_sweep_batch<T>(b ? c->sweep_b : c->sweep_a, e->nf, d, e->d_start, nt, e->nblk_tile, e->stream);
    }
    if (nr) launch_sweep_batch<T>(b ? SW_B : SW_A, e->nf, d + nt, e->d_start + nt + 1, nr, e->nblk_rest, e->stream);
}

void ens_phase(gabe_ensemble *e, int parity, bool b)
{
    if (e->fp32) ens_sweep<float>(e, parity, b);
    else ens_sweep<double>(e, parity, b);
}

int ens_launch_step(gabe_ensemble *e, int parity)
{
    ens_phase(e, parity, false);
    launch_finalize_batch(e->d_fin, e->nf, e->count, e->stream);
    ens_phase(e, parity, true);
    HIPCHK(hipGetLastError());
    return GABE_OK;
}

template <typename T>
int ens_upload_desc(gabe_ensemble *e)
{
    for (int p = 0; p < 2; p++) {
        std::vector<SweepArgs<T>> h(e->count);
        for (int q = 0; q < e->count; q++) {
            gabe_ctx *c = e->members[e->order[q]];
            const int keep = c->cur;
            c->cur = p;
            h[q] = sweep_args<T>(c);
            c->cur = keep;
        }
        HIPCHK(hipMalloc(&e->d_desc[p], sizeof(SweepArgs<T>) * e->count));
        HIPCHK(hipMemcpy(e->d_desc[p], h.data(), sizeof(SweepArgs<T>) * e->count, hipMemcpyHostToDevice));
    }
    return GABE_OK;
}

void ens_free(gabe_ensemble *e)
{
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto &g : e->graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (gabe_ctx *c : e->members) {
        if (!c) continue;
        destroy_impl(c);
        delete c;
    }
    void *bufs[] = {e->d_desc[0], e->d_desc[1], e->d_fin, e->d_start};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (e->stream) (void)hipStreamDestroy(e->stream);
}

int ens_create(const gabe_params *params, int count, gabe_ensemble *e)
{
    e->count = count;
    e->nf = params[0].nflds;
    e->fp32 = params[0].precision == GABE_FP32;
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    std::vector<FinArgs> fin(count);
    for (int m = 0; m < count; m++) {
        gabe_ctx *c = new gabe_ctx;
        e->members.push_back(c);
        CHK(create_impl(&params[m], c, e->stream));
        FinArgs &F = fin[m];
        F = FinArgs{};
        F.partials = c->partials;
        F.nblk = c->nblk;
        F.nf = c->nf;
        F.mode = SW_A;
        F.exp_type = c->p.expansion_type;
        F.split = c->fp32 ? 1 : 0;
        F.N3 = (double)c->N3;
        F.dt = c->p.dt;
        F.sc = c->sc;
        F.ss = c->ss;
        F.hist = c->hist;
        F.hist_cap = c->hist_cap;
    }
    for (int pass = 0; pass < 2; pass++)
        for (int m = 0; m < count; m++)
            if ((e->members[m]->sweep_b == SW_B_TILE) == (pass == 0)) e->order.push_back(m);
    e->ntile = 0;
    for (gabe_ctx *c : e->members) e->ntile += c->sweep_b == SW_B_TILE;
    /* workgroup starts of each launch, padded to multiples of the XCD count */
    std::vector<int> start;
    for (int g = 0; g < 2; g++) {
        const int q0 = g == 0 ? 0 : e->ntile, q1 = g == 0 ? e->ntile : count;
        int s0 = 0;
        start.push_back(0);
        for (int q = q0; q < q1; q++) {
            const gabe_ctx *c = e->members[e->order[q]];
            const int nb = (int)(c->grid.x * c->grid.y);
            s0 += (nb + GABE_NXCD - 1) / GABE_NXCD * GABE_NXCD;
            start.push_back(s0);
        }
        (g == 0 ? e->nblk_tile : e->nblk_rest) = s0;
    }
    HIPCHK(hipMalloc(&e->d_start, sizeof(int) * start.size()));
    HIPCHK(hipMemcpy(e->d_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&e->d_fin, sizeof(FinArgs) * count));
    HIPCHK(hipMemcpy(e->d_fin, fin.data(), sizeof(FinArgs) * count, hipMemcpyHostToDevice));
    if (e->fp32) CHK(ens_upload_desc<float>(e));
    else CHK(ens_upload_desc<double>(e));
    return GABE_OK;
}

/* time `steps` steps of one lattice with HIP events */
int time_steps(gabe_ctx *c, int steps, float &ms)
{
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = GABE_OK;
    if (hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(GABE_ERR_HIP, "event record failed");
    if (rc == GABE_OK) rc = run_steps(c, steps);
    if (rc == GABE_OK && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
        rc = fail(GABE_ERR_HIP, "event sync failed");
    if (rc == GABE_OK) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

} // namespace

extern "C" {

int gabe_query(gabe_ctx *c, gabe_params *out)
{
    if (!c || !out) return fail(GABE_ERR_ARG, "NULL argument");
    *out = c->p;
    out->planes_per_block = c->ci;
    out->transform_tile = c->colfft ? c->fft_cw : 0;
    out->mode_tile = c->gwfft ? c->gw_tile : 0;
    out->stress_chunk = c->chunk * c->ci;
    return GABE_OK;
}

int gabe_mode_tiles(int N, int precision) { return gabe_gwfft_tiles(N, precision == GABE_FP32); }

int gabe_autotune_sweep(const gabe_params *params, int *planes_per_block, double *ms_per_step)
{
    CHK(check_params(params));
    gabe_params p = *params;
    const bool fp32 = p.precision == GABE_FP32, gw = p.calc_gws != 0;
    /* the best of three timings of 10 steps (after 5) of a context with
     * tiling t (ms per 10 steps); fused says whether it uses the fused
     * transform path (the only one with transform tiles) */
    auto trial = [&](const Tiling &t, float &ms, bool &fused) -> int {
        p.planes_per_block = t.planes;
        p.transform_tile = t.cw;
        p.mode_tile = t.mode_tile;
        p.stress_chunk = t.chunk ? t.chunk : -1;
        gabe_ctx *c = new gabe_ctx;
        int rc = create_impl(&p, c);
        if (rc == GABE_OK) {
            fused = c->gwfft;
            /* a smooth synthetic state: timings do not depend on the values */
            const long n = (long)p.nflds * c->N3;
            std::vector<double> f(n), d(n);
            for (long q = 0; q < n; q++) {
                f[q] = 0.3 + 1e-6 * std::sin(0.01 * (double)q);
                d[q] = -0.2 + 1e-6 * std::cos(0.013 * (double)q);
            }
            rc = gabe_upload_state(c, f.data(), d.data());
            if (rc == GABE_OK) rc = time_steps(c, 5, ms);
            float best = 0.f;
            for (int rep = 0; rep < 3 && rc == GABE_OK; rep++) {
                rc = time_steps(c, 10, ms);
                if (rep == 0 || ms < best) best = ms;
            }
            ms = best;
        }
        destroy_impl(c);
        delete c;
        return rc;
    };
    Tiling best;
    float best_ms = 0.f;
    bool fused = false;
    for (int ci : {1, 2, 4, 8, 16, 32, 64}) {
        if (ci > p.N) break;
        float ms = 0.f;
        CHK(trial(Tiling{ci, 16, 1}, ms, fused));
        if (best.planes == 0 || ms < best_ms) {
            best = Tiling{ci, 16, 1};
            best_ms = ms;
        }
    }
    if (gw && fused) {   /* the transform tiles at the chosen sweep tiling */
        const int nt = gabe_gwfft_tiles(p.N, fp32);
        for (int cw : {16, 32})
            for (int mt = 1; mt <= nt; mt++) {
                float ms = 0.f;
                CHK(trial(Tiling{best.planes, cw, mt}, ms, fused));
                if (ms < best_ms) {
                    best = Tiling{best.planes, cw, mt};
                    best_ms = ms;
                }
            }
        /* sweep A and the j-transform interleaved in launches of N/8, N/4 or
         * N/2 planes (where that is a whole number of at least one block) */
        best.chunk = -1;
        for (int planes : {p.N / 8, p.N / 4, p.N / 2}) {
            if (planes < best.planes || planes % best.planes) continue;
            float ms = 0.f;
            const Tiling t{best.planes, best.cw, best.mode_tile, planes};
            CHK(trial(t, ms, fused));
            if (ms < best_ms) {
                best = t;
                best_ms = ms;
            }
        }
    }
    {
        std::lock_guard<std::mutex> lock(g_tuned_mu);
        g_tuned[tune_key(p.N, p.nflds, fp32, gw)] = best;
    }
    if (planes_per_block) *planes_per_block = best.planes;
    if (ms_per_step) *ms_per_step = best_ms / 10.;
    return GABE_OK;
}

int gabe_ensemble_create(const gabe_params *params, int count, gabe_ensemble **out)
{
    if (!out || !params || count <= 0) return fail(GABE_ERR_ARG, "bad argument");
    *out = nullptr;
    for (int m = 0; m < count; m++) {
        CHK(check_params(&params[m]));
        if (params[m].calc_gws) return fail(GABE_ERR_ARG, "ensemble member %d: tensor modes are not batched", m);
        if (params[m].precision != params[0].precision || params[m].nflds != params[0].nflds)
            return fail(GABE_ERR_ARG, "ensemble members must share precision and nflds");
    }
    /* Batched sweeps default to min(N, 32) planes per workgroup: with every
     * member's workgroups in one launch the GPU is full anyway, and marching
     * more planes re-reads fewer halo planes.  The single-lattice tuning (1,
     * 4, 8 and 16 planes at 32^3 ... 128^3) favours parallelism instead and
     * measured 7 % slower on the 64-lattice ensemble. */
    std::vector<gabe_params> ps(params, params + count);
    for (gabe_params &p : ps)
        if (p.planes_per_block <= 0) p.planes_per_block = std::min(p.N, 32);
    gabe_ensemble *e = new gabe_ensemble;
    const int rc = ens_create(ps.data(), count, e);
    if (rc != GABE_OK) {
        const std::string keep = g_err;
        ens_free(e);
        delete e;
        g_err = keep;
        return rc;
    }
    *out = e;
    return GABE_OK;
}

int gabe_ensemble_member(gabe_ensemble *e, int idx, gabe_ctx **out)
{
    if (!e || !out || idx < 0 || idx >= e->count) return fail(GABE_ERR_ARG, "bad member index");
    *out = e->members[idx];
    return GABE_OK;
}

int gabe_ensemble_step(gabe_ensemble *e, int nsteps)
{
    if (!e || nsteps < 0) return fail(GABE_ERR_ARG, "bad argument");
    const int parity = e->members[0]->cur;
    for (gabe_ctx *c : e->members)
        if (c->cur != parity) return fail(GABE_ERR_STATE, "ensemble members were stepped separately by an odd count");
    int p = parity, n = nsteps;
    if (n >= 2 && !e->graph[p]) {   /* capture two steps: the state returns to parity p */
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = ens_launch_step(e, p);
        if (rc == GABE_OK) rc = ens_launch_step(e, 1 - p);
        const hipError_t ec = hipStreamEndCapture(e->stream, &g);
        if (rc != GABE_OK || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);   /* a partial capture is not kept */
            if (rc != GABE_OK) return rc;
            return fail(GABE_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
        }
        const hipError_t ei = hipGraphInstantiate(&e->graph[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) return fail(GABE_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
    }
    for (; n >= 2; n -= 2) HIPCHK(hipGraphLaunch(e->graph[p], e->stream));
    if (n == 1) {
        CHK(ens_launch_step(e, p));
        p = 1 - p;
    }
    for (gabe_ctx *c : e->members) c->cur = p;
    return GABE_OK;
}

int gabe_ensemble_kernel_times(gabe_ensemble *e, int nsteps, double *ms_out)
{
    if (!e || !ms_out || nsteps <= 0) return fail(GABE_ERR_ARG, "bad argument");
    const int parity = e->members[0]->cur;
    for (gabe_ctx *c : e->members)
        if (c->cur != parity) return fail(GABE_ERR_STATE, "ensemble members were stepped separately by an odd count");
    std::vector<hipEvent_t> ev((size_t)4 * nsteps);
    for (auto &x : ev) HIPCHK(hipEventCreate(&x));
    int rc = GABE_OK, p = parity;
    for (int n = 0; n < nsteps && rc == GABE_OK; n++, p = 1 - p) {
        hipEvent_t *E = ev.data() + 4 * n;
        (void)hipEventRecord(E[0], e->stream);
        ens_phase(e, p, false);
        (void)hipEventRecord(E[1], e->stream);
        launch_finalize_batch(e->d_fin, e->nf, e->count, e->stream);
        (void)hipEventRecord(E[2], e->stream);
        ens_phase(e, p, true);
        (void)hipEventRecord(E[3], e->stream);
        if (hipGetLastError() != hipSuccess) rc = fail(GABE_ERR_HIP, "batched launch failed");
    }
    if (rc == GABE_OK && hipStreamSynchronize(e->stream) != hipSuccess) rc = fail(GABE_ERR_HIP, "sync failed");
    for (int k = 0; k < 3; k++) ms_out[k] = 0.;
    if (rc == GABE_OK)
        for (int n = 0; n < nsteps; n++)
            for (int k = 0; k < 3; k++) {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, ev[4 * n + k], ev[4 * n + k + 1]);
                ms_out[k] += ms / nsteps;
            }
    for (auto &x : ev) (void)hipEventDestroy(x);
    if (rc == GABE_OK)
        for (gabe_ctx *c : e->members) c->cur = p;
    return rc;
}

int gabe_ensemble_sync(gabe_ensemble *e)
{
    if (!e) return fail(GABE_ERR_ARG, "NULL ensemble");
    HIPCHK(hipStreamSynchronize(e->stream));
    for (gabe_ctx *c : e->members) CHK(gabe_sync(c));
    return GABE_OK;
}

int gabe_ensemble_destroy(gabe_ensemble *e)
{
    if (!e) return GABE_OK;
    ens_free(e);
    delete e;
    return GABE_OK;
}

} // extern "C"
