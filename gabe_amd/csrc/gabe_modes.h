/*
 * gabe_modes.h -- half-spectrum mode indexing and the TT polarisation basis,
 * shared by the tensor-mode kernels.
 */
#ifndef GABE_MODES_H
#define GABE_MODES_H

#include <hip/hip_runtime.h>

/* TT polarisation basis of mode n = (px,py,pz) != 0: u, v orthonormal and
 * orthogonal to n (u from the coordinate axis least aligned with n, v = n x u),
 * e+ = u u - v v, ex = u v + v u, stored as components 11,12,13,22,23,33.
 * With these, Lambda_ij,lm = (e+_ij e+_lm + ex_ij ex_lm) / 2: the TT part of
 * a symmetric S is sum_pol e_pol (e_pol : S) / 2. */
__device__ __forceinline__ void pol_uv(int px, int py, int pz, double u[3], double v[3])
{
    const double inv = 1. / sqrt((double)(px * px + py * py + pz * pz));
    const double n[3] = {px * inv, py * inv, pz * inv};
    int a = 0;
    if (fabs(n[1]) < fabs(n[a])) a = 1;
    if (fabs(n[2]) < fabs(n[a])) a = 2;
    u[0] = -n[a] * n[0]; u[1] = -n[a] * n[1]; u[2] = -n[a] * n[2];
    u[a] += 1.;
    const double un = 1. / sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    u[0] *= un; u[1] *= un; u[2] *= un;
    v[0] = n[1] * u[2] - n[2] * u[1];
    v[1] = n[2] * u[0] - n[0] * u[2];
    v[2] = n[0] * u[1] - n[1] * u[0];
}

/* components 11,12,13,22,23,33 of e+ and ex from u and v */
template <typename R>
__device__ __forceinline__ void pol_components(const R u[3], const R v[3], R ep[6], R ex[6])
{
    const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int c = 0; c < 6; c++) {
        ep[c] = u[ia[c]] * u[ib[c]] - v[ia[c]] * v[ib[c]];
        ex[c] = u[ia[c]] * v[ib[c]] + v[ia[c]] * u[ib[c]];
    }
}

__device__ __forceinline__ void pol_basis(int px, int py, int pz, double ep[6], double ex[6])
{
    double u[3], v[3];
    pol_uv(px, py, pz, u, v);
    pol_components<double>(u, v, ep, ex);
}

/* Storage order of the tensor-mode arrays h, l (one polarisation pair per
 * mode).  Lattice order: slot (i N + j) KP + kz, KP = N/2 + 1.  Tiled order,
 * used with the fused i-transform kernel: [j][kz/4][i][kz%4], so each of its
 * workgroups (one j, four kz, all i) reads and writes one contiguous block. */
struct ModeLayout {
    int N, tiled, nt4;  /* nt4 = ceil(KP / 4) */
    long KP;
};

__host__ __device__ __forceinline__ long mode_slot(const ModeLayout &L, int i, int j, int kz)
{
    if (!L.tiled) return ((long)i * L.N + j) * L.KP + kz;
    return (((long)j * L.nt4 + (kz >> 2)) * L.N + i) * 4 + (kz & 3);
}

__device__ __forceinline__ void mode_of(long m, int N, int &px, int &py, int &pz, long &row)
{
    const int H = N / 2 + 1;
    pz = (int)(m % H);
    row = m / H;
    const int j = (int)(row % N), i = (int)(row / N);
    px = i <= N / 2 ? i : i - N;
    py = j <= N / 2 ? j : j - N;
}

#endif
