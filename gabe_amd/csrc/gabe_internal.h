/*
 * gabe_internal.h -- device-side data structures and kernel launchers shared
 * by the kernels (gabe_kernels.hip) and the C-ABI implementation (gabe_api.hip).
 */
#ifndef GABE_INTERNAL_H
#define GABE_INTERNAL_H

#include <hip/hip_runtime.h>

#include "gabe_model.h"
#include "gabe_modes.h"

#define GABE_MAXF 2
#define GABE_BLOCK 256
#define GABE_NXCD 8         /* accelerator complex dies of an MI355X, each with its own L2 */
#define GABE_HIST_COLS 8
/* Stress components transformed per RK2 stage.  The TT projection
 * Lambda_ij,lm = P_il P_jm - P_ij P_lm / 2 annihilates any multiple of
 * delta_lm (tr P = 2 for k != 0), so the source S of g2evolutionpert.cpp:48-132
 * and S - S_zz delta give the same T = Lambda S: the five components
 * (xx - zz, xy, xz, yy - zz, yz) carry it, the zz one is zero.  Stage 0
 * (S(f0)) is components 0-4, stage 1 (S(f1)) 5-9. */
#define GABE_SC 5
#define GABE_NSC (2 * GABE_SC)

/* Persistent device scalars: the reference's a[2], adot[2], ed*[2]
 * (g2header.h:59-64) plus, per field, the FP64 background (mean field pbar and
 * mean velocity vbar) around which the FP32 form stores its fluctuations. */
struct DevScalars {
    double a[2], adot[2], edkin[2], edpot[2], edgrad[2], edrho[2];
    double pbar[GABE_MAXF], vbar[GABE_MAXF];
    double t;
    long long nstep;
    int adot0_valid; /* adot[0] belongs to the current state */
    int e0_valid;    /* ed*[0] belong to the current state */
    int error;       /* non-finite energy density seen */
    int pad;
    long long hist0; /* first step whose history row belongs to the current trajectory
                        (gabe_set_scalars, a checkpoint load or a restriction restart it) */
};

/* Values frozen by the step's finalize for sweep B and the GW mode kernel. */
struct StepScalars {
    double a0, adot0, a1, adot1;
    double pbar0[GABE_MAXF], vbar0[GABE_MAXF], pbar1[GABE_MAXF], vbar1[GABE_MAXF];
    double meanVp0[GABE_MAXF], meanVp1[GABE_MAXF];
};

/* Partial sums emitted per sweep workgroup.  Index layout:
 *   0: sum V(f0)    1: sum V(f1)
 *   per field f, base 2 + 7 f:
 *     +0 sum fdot0^2   +1 sum fdot0 X   +2 sum X^2   (X = lap f0/a0^2 - V'(f0))
 *     +3 sum f0 lap f0 +4 sum f1 lap f1
 *     +5 sum V'(f0)    +6 sum V'(f1)
 * f0 is the step-start state, f1 = f0 + dt/2 fdot0 the midpoint field. */
__host__ __device__ constexpr int gabe_nsums(int nf) { return 2 + 7 * nf; }
#define GABE_MAXSUMS (2 + 7 * GABE_MAXF)

/* SW_B_TILE, SW_A_TILE: sweep B, and sweep A without tensor modes, with
 * LDS-staged plane tiles, the callers' choice where they apply: N a multiple
 * of 16; sweep A in FP32 only (in FP64 its tiled form needs 182 VGPRs, two
 * waves per SIMD, against three for the per-thread one) */
enum SweepMode { SW_ENERGY = 0, SW_A = 1, SW_B = 2, SW_B_TILE = 3, SW_A_TILE = 4 };
inline bool gabe_sweep_tiled(int N) { return N % 16 == 0; }

/* Loads and stores through global-address-space pointers.  A pointer read
 * from a descriptor table in memory (the batched launches) is generic to the
 * compiler, which then emits flat instructions: these also count against the
 * LDS counter, so every LDS wait waits for them as well. */
template <typename T>
__device__ __forceinline__ T gld(const T *p)
{
    return *(const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ void gst(T *p, T v)
{
    *(__attribute__((address_space(1))) T *)p = v;
}

template <typename T>
struct SweepArgs {
    const T *f[GABE_MAXF];  /* field (FP64) or fluctuation (FP32), input state */
    const T *v[GABE_MAXF];  /* time derivative, input state */
    T *fo[GABE_MAXF];       /* sweep B output state */
    T *vo[GABE_MAXF];
    T *stress;              /* [GABE_NSC][N^3]: S(f0) comps 0-4, S(f1) comps 5-9 */
    T *shat;                /* [GABE_NSC][N][N][kpad = N/2] complex: row transforms written by the fused
                             * sweep A, slot 0 = S(kz 0) + i S(kz N/2) */
    const T *tw;            /* exp(-2 pi i e / N), e < N, complex */
    long kpad;
    int rowfft;             /* N when sweep A transforms its k-rows itself, else 0 */
    double *partials;       /* [nsums][nblk] */
    const DevScalars *sc;
    const StepScalars *ss;
    gabe_model model;
    int N, ci, nblk;
    int gx, gy;             /* workgroup grid of this lattice's sweeps */
    int by0;                /* first plane block of the launch: a launch of gridDim.y
                             * blocks covers plane blocks by0 .. by0 + gridDim.y - 1 */
    long N2, N3;
    double dt, inv_dx2;
};

struct FinArgs {
    const double *partials;
    int nblk, nf, mode, exp_type, split, refresh_adot;
    double N3, dt;
    DevScalars *sc;
    StepScalars *ss;
    double *hist;
    int hist_cap;
};

template <typename T>
struct GwArgs {
    const T *shat;      /* [GABE_NSC][N][N][kpad] complex: S(f0) transforms 0-4, S(f1) 5-9 */
    T *h, *l;           /* [M][2] complex: (+, x) polarisation amplitudes per mode */
    const StepScalars *ss;
    int N;
    long M;             /* N*N*(N/2+1) modes */
    long kpad;          /* row stride of shat (complex elements) */
    double L, dt, dx3;
};

template <typename T>
struct SpecArgs {
    const T *fhat;      /* [nf][M] complex field transforms */
    const T *l;         /* polarisation pairs in the order lay, or null */
    ModeLayout lay;
    double dc[GABE_MAXF];   /* added to the real part of each field's k=0 mode */
    int nf, N, numbins, quirk;
    long M;
    double *part;       /* [numbins][nch][nf + 2] per-chunk partial sums */
    int nch;            /* px chunks per shell, gabe_spec_chunks(N) */
    double *out;        /* [numbins][nf + 2]: per-field sum w|F|^2, sum w, sum w E_gw */
};

/* px values per spectrum chunk, and chunks per shell: the widest shell spans N values of px */
constexpr int GABE_SPEC_PX = 4;
inline int gabe_spec_chunks(int N) { return (N + GABE_SPEC_PX - 1) / GABE_SPEC_PX; }

template <typename T>
void launch_sweep(int mode, int nf, bool gw, const SweepArgs<T> &a, dim3 grid, hipStream_t s);
void launch_finalize(const FinArgs &a, hipStream_t s);
template <typename T>
void launch_sweep_batch(int mode, int nf, const SweepArgs<T> *descs, const int *start, int count, int nblk,
                        hipStream_t s);
void launch_finalize_batch(const FinArgs *descs, int nf, int count, hipStream_t s);
template <typename T>
void launch_gw(const GwArgs<T> &a, hipStream_t s);
template <typename T>
void launch_gw_pack(const double *ref, T *pol, ModeLayout lay, long M, hipStream_t s);
template <typename T>
void launch_gw_unpack(const T *pol, double *ref, ModeLayout lay, long M, hipStream_t s);
template <typename T>
void launch_tt_out(const T *shat, const T *nyq, int stage, int N, long M, long kpad, double dx3, double *out,
                   hipStream_t s);
template <typename T>
void launch_spectra(const SpecArgs<T> &a, hipStream_t s);
template <typename T>
void launch_effmass(const T *f0, const T *f1, int nf, int fld, double pbar0, double pbar1,
                    gabe_model m, long n3, double *partials, int nblk, hipStream_t s);
template <typename T>
void launch_meanvar(const T *f, long n3, double *partials, int nblk, hipStream_t s);
template <typename T>
void launch_add_real(T *dst, const double *add, long n, hipStream_t s);

#endif
