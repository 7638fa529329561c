/*
 * gabe_ctx.h -- the device context behind gabe_ctx and the helpers shared by
 * the translation units that implement the C ABI (internal, not installed).
 */
#pragma once
#include <hip/hip_runtime.h>

#include <vector>
#include <rocfft/rocfft.h>

#include <string>

#include "gabe.h"
#include "gabe_internal.h"

namespace gabe_detail {

/* record the message of a failure for gabe_last_error() and return code */
int fail(int code, const char *fmt, ...);
/* Neumaier-compensated sum in index order */
double csum(const double *x, long n);

} // namespace gabe_detail

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return gabe_detail::fail(GABE_ERR_HIP, "%s failed: %s (%s:%d)", #expr,          \
                                     hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)

#define FFTCHK(expr)                                                                        \
    do {                                                                                    \
        rocfft_status s_ = (expr);                                                          \
        if (s_ != rocfft_status_success)                                                    \
            return gabe_detail::fail(GABE_ERR_FFT, "%s failed: rocfft status %d", #expr, (int)s_); \
    } while (0)

#define CHK(expr)                        \
    do {                                 \
        int r_ = (expr);                 \
        if (r_ != GABE_OK) return r_;    \
    } while (0)

struct ClosureRun;   /* a learned closure attached to a context (gabe_closure.hip) */

struct FftPlan {
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    size_t work = 0;
};

struct gabe_ctx {
    gabe_params p{};
    int N = 0, nf = 0;
    long N2 = 0, N3 = 0, M = 0;
    double dx = 0.;
    bool fp32 = false, gw = false, alloc_gw = false;
    size_t esz = 8;              /* bytes per stored real */
    hipStream_t stream = nullptr;
    bool own_stream = true;          /* false for ensemble members (the ensemble's stream) */
    hipStream_t side = nullptr;      /* sweep B runs here, beside the stress transform and GW chain */
    hipEvent_t ev_fin = nullptr, ev_side = nullptr;
    /* pipelined steps (FP32, k rows in sweep A, fused i-transform): the
     * j-transform and the tensor modes of step n run on the side stream while
     * sweeps and finalize of step n+1 run on the main one; the stress spectra
     * and the step scalars alternate between two slots, and ev_gw[s] marks the
     * side work of the last step that used slot s as done */
    bool pipe = false;
    int slot = 0;
    void *shat_b[2] = {nullptr, nullptr};
    StepScalars *ss_b[2] = {nullptr, nullptr};
    hipEvent_t ev_gw[2] = {nullptr, nullptr};
    /* chunked stress transform (pipelined steps, chunk > 0): sweep A runs in
     * launches of chunk plane blocks and the j-transform of each launch's
     * planes follows on jst as soon as it is done, while the spectra it reads
     * are still in the memory-side cache; ev_chunk[q] marks sweep A's launch q
     * done, ev_jdone the last j-transform */
    int chunk = 0;
    hipStream_t jst = nullptr;
    std::vector<hipEvent_t> ev_chunk;
    hipEvent_t ev_jdone = nullptr;
    void *fbuf[2] = {nullptr, nullptr};  /* ping-pong state, [nf][N3] each */
    void *vbuf[2] = {nullptr, nullptr};
    int cur = 0;
    void *stress = nullptr;      /* [ns_arrays][N3] real */
    void *shat = nullptr;        /* [ns_arrays][M] complex */
    void *h = nullptr, *l = nullptr; /* mslots polarisation pairs (complex) each, order lay */
    ModeLayout lay{};
    long mslots = 0;
    double *partials = nullptr;
    int nblk = 0, ci = 0;
    dim3 grid;
    DevScalars *sc = nullptr;
    StepScalars *ss = nullptr;
    double *hist = nullptr;
    int hist_cap = 0;
    double *spec_out = nullptr;
    double *spec_part = nullptr;   /* per-chunk partial sums of the spectra */
    FftPlan fft_stress, fft_spec, fft_c2r;
    bool colfft = false;         /* k-rows by the library, j and i columns by gabe_fft.hip */
    bool gwfft = false;          /* i columns fused with the tensor-mode update */
    bool rowfft = false;         /* k rows transformed inside sweep A */
    int sweep_a = 1, sweep_b = 2; /* sweep variants: SW_A and SW_B, or SW_A_TILE (without tensor
                                  * modes) and SW_B_TILE with LDS-staged tiles */
    long kpad = 0;               /* row stride of shat in complex elements (>= N/2+1) */
    int fft_cw = 16;             /* column tile width of the j/i transforms */
    int gw_tile = 1;             /* tile variant of the fused i-transform + mode kernel */
    void *twiddle = nullptr;     /* exp(-2 pi i e/N), e < N, complex storage type */
    void *fft_work = nullptr;
    size_t fft_work_size = 0;
    gabe_model model{};
    /* spectral restriction onto a coarser context (gabe_restrict) */
    FftPlan fft_rx;                  /* R2C of one array (fine side) or C2R (coarse side) */
    void *rx_buf = nullptr;          /* one half-spectrum, complex storage type */
    /* learned closure applied after every step (gabe_closure_attach) */
    ClosureRun *closure = nullptr;
};

namespace gabe_detail {

int make_plan(gabe_ctx *c, FftPlan &P, rocfft_transform_type type, bool dbl, size_t batch, int rank = 3,
              size_t out_distance = 0);
int run_plan(gabe_ctx *c, FftPlan &P, void *in, void *out);
void destroy_plan(FftPlan &P);
int read_scalars(gabe_ctx *c, DevScalars &d);
int write_scalars(gabe_ctx *c, const DevScalars &d);
/* closure kernels after a step (gabe_closure.hip) and the attachment's teardown */
int closure_after_step(gabe_ctx *c, hipEvent_t *ev = nullptr);   /* ev: 8 events, before/after each kernel */
void closure_release(gabe_ctx *c);
/* the attached closure's part of a checkpoint: its gate (on/off), the step it
 * switched off at, its largest residual, threshold, scales and weights */
struct ClosureCkpt {
    int N, gate;
    long long fallback;
    double rmax, thr;
    float inv_s[4], dt_s[2];
};
size_t closure_ckpt_bytes(const gabe_ctx *c);            /* 0 without a closure */
int closure_ckpt_save(gabe_ctx *c, char *p);
int closure_ckpt_load(gabe_ctx *c, const char *p, size_t bytes);  /* 0 bytes: detach */

} // namespace gabe_detail
