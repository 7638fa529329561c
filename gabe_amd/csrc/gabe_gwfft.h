/*
 * gabe_gwfft.h -- the i-direction stress transform fused with the tensor-mode
 * update (gabe_gwfft.hip).
 */
#ifndef GABE_GWFFT_H
#define GABE_GWFFT_H

#include <hip/hip_runtime.h>

#include "gabe_internal.h"

template <typename T>
struct GwFftArgs {
    const T *shat;      /* [12][N][N][kpad] complex, transformed along k and j */
    T *h, *l;           /* polarisation pairs, tiled order (gabe_modes.h) */
    const StepScalars *ss;
    const T *tw;        /* exp(-2 pi i e / N), e < N, complex */
    int N;
    long M, kpad;
    double L, dt, dx3;
    int nt4;            /* kz tiles of 4 of the tiled mode order */
    const T *nyq;       /* null: rows hold kz 0..N/2.  Else rows hold kz < N/2
                         * (kpad = N/2) and nyq [12][N][N] the kz = N/2 plane */
    int tile;           /* tile variant, 1..gabe_gwfft_tiles(N, fp32); 1 = default */
};

/* whether the fused kernel exists for this size and storage type: N a power
 * of two in 16..512 (FP32) or 16..256 (FP64) */
bool gabe_gwfft_supported(int N, bool fp32);
/* number of tile variants of the fused kernel for this size (>= 1 where it is
 * supported, 0 elsewhere) */
int gabe_gwfft_tiles(int N, bool fp32);

template <typename T>
void launch_gw_ifft(const GwFftArgs<T> &a, hipStream_t s);

#endif
