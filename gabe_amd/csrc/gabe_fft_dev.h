/*
 * gabe_fft_dev.h -- device building blocks of the column transforms: complex
 * types, compile-time twiddle tables and in-register power-of-two DFTs.
 */
#ifndef GABE_FFT_DEV_H
#define GABE_FFT_DEV_H

#include <hip/hip_runtime.h>

#include "gabe_fft.h"

namespace gabe_fft {

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
                if (k * (R / len) == R / 4) {
                    /* w = -i: a swap and a sign, the same values as the product */
                    const T vx = v.x;
                    v.x = v.y;
                    v.y = -vx;
                } else if (k != 0) {
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

} // namespace gabe_fft

#endif
