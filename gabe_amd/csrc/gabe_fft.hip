/*
 * gabe_fft.hip -- batched complex column transforms for the stress spectra.
 *
 * After the contiguous-row R2C along k, the tensor source needs complex
 * transforms along j (stride kpad) and along i (stride N kpad).  A generic
 * library plan does these as strided column passes; here one workgroup takes
 * CW adjacent columns (CW complex contiguous per row, a full cache line in
 * FP32) and transforms them with the two-step factorisation N = N1 N2, the
 * first step's inputs loaded into registers and its outputs exchanged through
 * an N x CW tile in LDS:
 *   A[n2][k1] = W_N^(n2 k1) sum_n1 x[N2 n1 + n2] W_N1^(n1 k1)      (step 1)
 *   X[k1 + N1 k2] = sum_n2 A[n2][k1] W_N2^(n2 k2)                   (step 2)
 * Each short DFT runs in registers with compile-time twiddles; the inter-step
 * twiddles come from a table computed on the host in FP64.  Forward sign
 * (e^{-i k x}), unnormalised, the convention of FFTW's r2c that the reference
 * uses (g2evolutionpert.cpp:183-195).
 */
#include <hip/hip_runtime.h>

#include "gabe_fft.h"
#include "gabe_fft_dev.h"

namespace {

using namespace gabe_fft;

constexpr int kColThreads = 256;

/* Element (o, r, s, c) of the view [outer][N][sub][cols] sits at
 * o*outer_stride + r*axis_stride + s*sub_stride + c.  Each workgroup takes CW
 * adjacent columns (c < sub_valid) of one (o, s) and transforms them along r,
 * in place.  The first step reads its inputs straight from the columns (each
 * task's N1 loads in flight together, CW adjacent columns per row across the
 * lanes); only its outputs pass through LDS to the second step. */
template <typename T, int N1, int N2, int CW>
__global__ __launch_bounds__(kColThreads) void k_colfft(typename cplx_t<T>::type *data, long axis_stride,
                                                       long outer_stride, long sub_stride, int sub_valid,
                                                       int chunks_per_sub, const typename cplx_t<T>::type *tw,
                                                       typename cplx_t<T>::type *nyq, long nyq_outer)
{
    using C = typename cplx_t<T>::type;
    constexpr int N = N1 * N2;
    static_assert(sizeof(C) * N * (CW + 3) <= 160 * 1024, "tile exceeds the LDS of one CU");
    __shared__ C tile[N][CW + 1];
    __shared__ C packed[N];   /* transformed column 0 when it holds a packed pair */
    __shared__ C twl[N];      /* the twiddle table, read by the first step */
    /* XCD-aware order: neighbouring column chunks on one XCD (one L2) */
    int b = blockIdx.x;
    if (gridDim.x % 8 == 0) b = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
    const int sidx = b / chunks_per_sub, q = b % chunks_per_sub;
    const int width = min(CW, sub_valid - q * CW);
    const bool split = nyq != nullptr && q == 0;   /* uniform over the workgroup */
    C *base = data + (long)blockIdx.y * outer_stride + (long)sidx * sub_stride + (long)q * CW;

    /* first step: task (n2, cc) loads x[n1] = row N2 n1 + n2, transforms it
     * and multiplies by the twiddles; n2 = 0 multiplies by tw[0] = (1, -0),
     * the value unchanged up to the sign of a zero, the lanes kept together */
    auto first_step = [&](int task, C (&x)[N1]) {
        const int cc = task % CW, n2 = task / CW;
        dft_reg<N1, T>(x);
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) {
            C v = x[k1];
            if (k1 != 0) v = cmul(v, twl[(n2 * k1) % N]);
            tile[N2 * k1 + n2][cc] = v;
        }
    };
    auto load_task = [&](int task, C (&x)[N1]) {
        const int cc = task % CW, n2 = task / CW;
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) {
            if (cc < width) x[n1] = base[(long)(N2 * n1 + n2) * axis_stride + cc];
            else x[n1].x = x[n1].y = (T)0;
        }
    };
    if constexpr (N2 * CW == kColThreads) {
        /* one task per thread: its loads go out before the twiddle table is
         * filled, so the barrier that publishes the table leaves them in flight */
        C x[N1];
        load_task(threadIdx.x, x);
        for (int e = threadIdx.x; e < N; e += kColThreads) twl[e] = tw[e];
        __syncthreads();
        first_step(threadIdx.x, x);
    } else {
        for (int e = threadIdx.x; e < N; e += kColThreads) twl[e] = tw[e];
        __syncthreads();
        for (int task = threadIdx.x; task < N2 * CW; task += kColThreads) {
            C x[N1];
            load_task(task, x);
            first_step(task, x);
        }
    }
    __syncthreads();

    for (int task = threadIdx.x; task < N1 * CW; task += kColThreads) {
        const int cc = task % CW, k1 = task / CW;
        C y[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; n2++) y[n2] = tile[N2 * k1 + n2][cc];
        dft_reg<N2, T>(y);
        if (split && cc == 0) {
#pragma unroll
            for (int k2 = 0; k2 < N2; k2++) packed[k1 + N1 * k2] = y[k2];
        } else if (cc < width) {
#pragma unroll
            for (int k2 = 0; k2 < N2; k2++) base[(long)(k1 + N1 * k2) * axis_stride + cc] = y[k2];
        }
    }
    if (split) {
        __syncthreads();
        C *ny = nyq + (long)blockIdx.y * nyq_outer + (long)sidx * N;
        const T h = (T)0.5;
        for (int n = threadIdx.x; n < N; n += kColThreads) {
            const C Z = packed[n], W = packed[(N - n) & (N - 1)];
            C x, y;
            x.x = h * (Z.x + W.x);
            x.y = h * (Z.y - W.y);
            y.x = h * (Z.y + W.y);
            y.y = -h * (Z.x - W.x);
            base[(long)n * axis_stride] = x;
            ny[n] = y;
        }
    }
}

template <typename T, int N1, int N2, int CW>
void launch_cw(const ColFftView &v, const void *tw, hipStream_t s)
{
    using C = typename cplx_t<T>::type;
    const int chunks = (v.sub_valid + CW - 1) / CW;
    const dim3 grid((unsigned)(v.nsub * chunks), (unsigned)v.outer);
    hipLaunchKernelGGL((k_colfft<T, N1, N2, CW>), grid, dim3(kColThreads), 0, s, (C *)v.data, v.axis_stride,
                       v.outer_stride, v.sub_stride, v.sub_valid, chunks, (const C *)tw, (C *)v.nyq,
                       v.nyq_outer ? v.nyq_outer : (long)N1 * N2);
}

template <typename T, int N1, int N2>
void launch_n(const ColFftView &v, int cw, const void *tw, hipStream_t s)
{
    constexpr size_t csz = 2 * sizeof(T);
    if constexpr (csz * N1 * N2 * 35 <= 160 * 1024) {
        if (cw == 32) { launch_cw<T, N1, N2, 32>(v, tw, s); return; }
    }
    if constexpr (csz * N1 * N2 * 19 <= 160 * 1024) launch_cw<T, N1, N2, 16>(v, tw, s);
    else launch_cw<T, N1, N2, 8>(v, tw, s);
}

} // namespace

bool gabe_colfft_supported(int N)
{
    return N == 8 || N == 16 || N == 32 || N == 64 || N == 128 || N == 256 || N == 512;
}

template <typename T>
void launch_colfft(const ColFftView &v, int N, int cw, const void *tw, hipStream_t s)
{
    switch (N) {
    case 8: launch_n<T, 4, 2>(v, cw, tw, s); break;
    case 16: launch_n<T, 4, 4>(v, cw, tw, s); break;
    case 32: launch_n<T, 8, 4>(v, cw, tw, s); break;
    case 64: launch_n<T, 8, 8>(v, cw, tw, s); break;
    case 128: launch_n<T, 16, 8>(v, cw, tw, s); break;
    case 256: launch_n<T, 16, 16>(v, cw, tw, s); break;
    case 512: launch_n<T, 32, 16>(v, cw, tw, s); break;
    default: break;
    }
}

template void launch_colfft<float>(const ColFftView &, int, int, const void *, hipStream_t);
template void launch_colfft<double>(const ColFftView &, int, int, const void *, hipStream_t);
