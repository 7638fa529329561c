/*
 * gabe_fft.h -- column transforms (gabe_fft.hip) and exact twiddle factors.
 */
#ifndef GABE_FFT_H
#define GABE_FFT_H

#include <hip/hip_runtime.h>

/* cos and sin of 2 pi e / R, reduced to |x| <= pi/4 by exact octant symmetry
 * (integer arithmetic) before a Taylor series: correctly rounded to ~1 ulp in
 * FP64, usable in constant expressions. */
constexpr double gabe_taylor_sin(double x)
{
    double term = x, sum = x;
    for (int n = 1; n < 14; n++) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double gabe_taylor_cos(double x)
{
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 14; n++) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
constexpr void gabe_sincos2pi_frac(long e, long R, double &c, double &s)
{
    long u = ((e % R) + R) % R * 8;   /* angle in units of 2 pi / (8R) */
    const long full = 8 * R;
    double sgn_s = 1.0, sgn_c = 1.0;
    if (u > full / 2) { u = full - u; sgn_s = -1.0; }
    if (u > full / 4) { u = full / 2 - u; sgn_c = -1.0; }
    bool swap = false;
    if (u > full / 8) { u = full / 4 - u; swap = true; }
    const double x = 6.283185307179586476925286766559 * (double)u / (double)full;
    double cc = gabe_taylor_cos(x), ss = gabe_taylor_sin(x);
    if (swap) { const double t = cc; cc = ss; ss = t; }
    c = sgn_c * cc;
    s = sgn_s * ss;
}
constexpr double gabe_cos2pi_frac(long e, long R)
{
    double c = 0, s = 0;
    gabe_sincos2pi_frac(e, R, c, s);
    return c;
}
constexpr double gabe_sin2pi_frac(long e, long R)
{
    double c = 0, s = 0;
    gabe_sincos2pi_frac(e, R, c, s);
    return s;
}

/* N for which the column transform exists (powers of two, 8..512) */
bool gabe_colfft_supported(int N);

/* Column set for one transform launch: element (o, r, s, c) of the view
 * [outer][N][nsub][cols] at o*outer_stride + r*axis_stride + s*sub_stride + c
 * (complex elements), transformed along r for every o, s and c < sub_valid.
 *
 * With nyq set, column 0 of every (o, s) holds two real sequences packed as
 * x + i y (the kz = 0 and kz = N/2 values of real-input rows).  Their
 * transforms X, Y are Hermitian in the transform index n, so
 *   X(n) = (Z(n) + conj Z(N-n)) / 2,   Y(n) = (Z(n) - conj Z(N-n)) / 2i;
 * X goes back into column 0 and Y to nyq[o*nyq_outer + s*N + n]
 * (nyq_outer 0 means N, for views with nsub = 1). */
struct ColFftView {
    void *data;
    long axis_stride, outer_stride, sub_stride;
    int sub_valid, nsub;
    long outer;
    void *nyq = nullptr;
    long nyq_outer = 0;
};

/* In-place forward complex transform of length N along r.  cw is the column
 * tile width (16 or 32 complex).  tw is the device table exp(-2 pi i e / N),
 * e < N, in complex T. */
template <typename T>
void launch_colfft(const ColFftView &v, int N, int cw, const void *tw, hipStream_t s);

#endif
