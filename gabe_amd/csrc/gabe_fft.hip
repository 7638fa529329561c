/*
 * gabe_fft.hip -- batched complex column transforms for the stress spectra.
 *
 * After the contiguous-row R2C along k, the tensor source needs complex
 * transforms along j (stride N/2+1) and along i (stride N(N/2+1)).  A generic
 * library plan does these as strided column passes; here one workgroup takes
 * CW adjacent columns (CW complex contiguous per row, a full cache line in
 * FP32), stages the N x CW tile in LDS and transforms it with the two-step
 * factorisation N = N1 N2:
 *   A[n2][k1] = W_N^(n2 k1) sum_n1 x[N2 n1 + n2] W_N1^(n1 k1)      (step 1)
 *   X[k1 + N1 k2] = sum_n2 A[n2][k1] W_N2^(n2 k2)                   (step 2)
 * Each short DFT runs in registers with compile-time twiddles; the inter-step
 * twiddles come from a table computed on the host in FP64.  Forward sign
 * (e^{-i k x}), unnormalised, the convention of FFTW's r2c that the reference
 * uses (g2evolutionpert.cpp:183-195).
 */
#include <hip/hip_runtime.h>

#include "gabe_fft.h"

namespace {

template <typename T>
struct cplx_t;
template <>
struct cplx_t<float> { using type = float2; };
template <>
struct cplx_t<double> { using type = double2; };

/* table of exp(-2 pi i e / R), e < R, for the in-register DFTs */
template <int R>
struct Twiddles {
    double re[R], im[R];
    constexpr Twiddles() : re(), im()
    {
        for (int e = 0; e < R; e++) {
            re[e] = gabe_cos2pi_frac(e, R);
            im[e] = -gabe_sin2pi_frac(e, R);
        }
    }
};

template <int R>
constexpr int bitrev(int i)
{
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

template <typename C>
__device__ __forceinline__ C cmul(C a, C b)
{
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

/* in-place forward DFT of R (power of two) values held in registers */
template <int R, typename T, typename C>
__device__ __forceinline__ void dft_reg(C (&x)[R])
{
    constexpr Twiddles<R> tw{};
#pragma unroll
    for (int i = 0; i < R; i++) {
        const int j = bitrev<R>(i);
        if (j > i) {
            const C t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
#pragma unroll
    for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
        for (int i = 0; i < R; i += len) {
#pragma unroll
            for (int k = 0; k < len / 2; k++) {
                const C u = x[i + k];
                C v = x[i + k + len / 2];
                if (k != 0) {
                    C w;
                    w.x = (T)tw.re[k * (R / len)];
                    w.y = (T)tw.im[k * (R / len)];
                    v = cmul(v, w);
                }
                x[i + k].x = u.x + v.x;
                x[i + k].y = u.y + v.y;
                x[i + k + len / 2].x = u.x - v.x;
                x[i + k + len / 2].y = u.y - v.y;
            }
        }
    }
}

constexpr int kColThreads = 256;

/* data viewed as [outer][N][inner]; each workgroup transforms CW adjacent
 * inner columns of one outer slab along the middle axis, in place */
template <typename T, int N1, int N2, int CW>
__global__ __launch_bounds__(kColThreads) void k_colfft(typename cplx_t<T>::type *data, long inner, long outer_stride,
                                                       const typename cplx_t<T>::type *tw)
{
    using C = typename cplx_t<T>::type;
    constexpr int N = N1 * N2;
    __shared__ C tile[N][CW + 1];
    const long col0 = (long)blockIdx.x * CW;
    const int width = (int)min((long)CW, inner - col0);
    C *base = data + (long)blockIdx.y * outer_stride + col0;

    for (int e = threadIdx.x; e < N * CW; e += kColThreads) {
        const int r = e / CW, cc = e % CW;
        if (cc < width) tile[r][cc] = base[(long)r * inner + cc];
    }
    __syncthreads();

    for (int task = threadIdx.x; task < N2 * CW; task += kColThreads) {
        const int cc = task % CW, n2 = task / CW;
        C x[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) x[n1] = tile[N2 * n1 + n2][cc];
        dft_reg<N1, T>(x);
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) {
            C v = x[k1];
            if (k1 != 0 && n2 != 0) v = cmul(v, tw[(n2 * k1) % N]);
            tile[N2 * k1 + n2][cc] = v;
        }
    }
    __syncthreads();

    for (int task = threadIdx.x; task < N1 * CW; task += kColThreads) {
        const int cc = task % CW, k1 = task / CW;
        C y[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; n2++) y[n2] = tile[N2 * k1 + n2][cc];
        dft_reg<N2, T>(y);
        if (cc < width) {
#pragma unroll
            for (int k2 = 0; k2 < N2; k2++) base[(long)(k1 + N1 * k2) * inner + cc] = y[k2];
        }
    }
}

template <typename T, int N1, int N2>
void launch_n(void *data, long inner, long outer, const void *tw, hipStream_t s)
{
    using C = typename cplx_t<T>::type;
    constexpr int CW = 16;
    const dim3 grid((unsigned)((inner + CW - 1) / CW), (unsigned)outer);
    hipLaunchKernelGGL((k_colfft<T, N1, N2, CW>), grid, dim3(kColThreads), 0, s, (C *)data, inner,
                       (long)(N1 * N2) * inner, (const C *)tw);
}

} // namespace

bool gabe_colfft_supported(int N)
{
    return N == 8 || N == 16 || N == 32 || N == 64 || N == 128 || N == 256 || N == 512;
}

template <typename T>
void launch_colfft(void *data, int N, long inner, long outer, const void *tw, hipStream_t s)
{
    switch (N) {
    case 8: launch_n<T, 4, 2>(data, inner, outer, tw, s); break;
    case 16: launch_n<T, 4, 4>(data, inner, outer, tw, s); break;
    case 32: launch_n<T, 8, 4>(data, inner, outer, tw, s); break;
    case 64: launch_n<T, 8, 8>(data, inner, outer, tw, s); break;
    case 128: launch_n<T, 16, 8>(data, inner, outer, tw, s); break;
    case 256: launch_n<T, 16, 16>(data, inner, outer, tw, s); break;
    case 512: launch_n<T, 32, 16>(data, inner, outer, tw, s); break;
    default: break;
    }
}

template void launch_colfft<float>(void *, int, long, long, const void *, hipStream_t);
template void launch_colfft<double>(void *, int, long, long, const void *, hipStream_t);
