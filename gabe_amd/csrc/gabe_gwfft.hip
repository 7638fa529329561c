/*
 * gabe_gwfft.hip -- last stress transform (along i) fused with the tensor-mode
 * update.
 *
 * After the k-row R2C and the j-column transform, one workgroup takes the
 * N x CW tile of modes (all i, CW adjacent kz) at one j, or, when the kz = N/2
 * values sit in a plane of their own, all i at CW adjacent j of that plane.
 * It transforms the GABE_NSC stored stress components (five per RK2 stage)
 * along i in LDS, CPR components per round.  After each round, every thread
 * projects the freshly transformed components of its own modes onto the two
 * TT polarisations and accumulates them.  Once all are in, it runs both
 * midpoint-RK2 stages of those modes
 * (g2evolutionpert.cpp:343-516) and writes h and l.  The fully transformed
 * source never goes back to HBM.
 */
#include <hip/hip_runtime.h>

#include "gabe_fft_dev.h"
#include "gabe_gwfft.h"
#include "gabe_modes.h"

namespace {

using namespace gabe_fft;

/* largest divisor of n not above 12: tile loads in flight per batch (all 12
 * of the FP32 256^3 tile at once) */
constexpr int load_batch(int n)
{
    for (int b = 12; b > 1; b--)
        if (n % b == 0) return b;
    return 1;
}

/* LDS row of lattice index i within a tile: one spare row after every N2
 * rows.  Rows are CW complex wide; the first DFT step's lanes read N2
 * consecutive rows, the second step's lanes rows N2 apart, and with this
 * spacing both (and the loads, stores and projections) fall on distinct
 * banks. */
template <int N2>
__device__ __forceinline__ int trow(int i) { return i + i / N2; }
/* trow(N2 a + b) for b < N2 */
template <int N2>
__device__ __forceinline__ int trow2(int a, int b) { return (N2 + 1) * a + b; }

/* tile elements each thread loads per round */
constexpr int gw_lpt(int N, int CW, int CPR, int NT) { return (CPR * N * CW + NT - 1) / NT; }

/* One round: source components C0 .. C0 + NCR - 1 of the GABE_NSC stored
 * ones (stage c / GABE_SC, component c mod GABE_SC of (xx - zz, xy, xz,
 * yy - zz, yz)): fill the NCR N x CW tiles (loaded now, or with PF_IN from
 * pf, loaded by the round before), transform along i (two steps, twiddles
 * from twl), return to natural order, and project the owned modes on e+ and
 * ex, adding into acc0 or acc1 by the component's stage (the projection is
 * linear in the components, so rounds may split or straddle stages).  With
 * NNX > 0 the next round's NNX tiles are loaded into pf after the first DFT
 * step, so their latency passes behind the second step, the projection and
 * the barriers between them (loaded before the first step they would not fit
 * the register budget). */
template <typename T, int N1, int N2, int CW, int CPR, int NT, int C0, int NCR, int NNX, bool PF_IN,
          typename AfterDft>
__device__ __forceinline__ void source_round(AfterDft after_dft,
                                             typename cplx_t<T>::type (&tile)[CPR][N1 * N2 + N1][CW],
                                             const typename cplx_t<T>::type *twl,
                                             const typename cplx_t<T>::type *Sh, long SM, long rstride, int width,
                                             const T (&bu)[(N1 * N2 * CW + NT - 1) / NT][3],
                                             const T (&bv)[(N1 * N2 * CW + NT - 1) / NT][3],
                                             typename cplx_t<T>::type (&acc0)[(N1 * N2 * CW + NT - 1) / NT][2],
                                             typename cplx_t<T>::type (&acc1)[(N1 * N2 * CW + NT - 1) / NT][2],
                                             typename cplx_t<T>::type (&pf)[gw_lpt(N1 * N2, CW, CPR, NT)])
{
    using C = typename cplx_t<T>::type;
    constexpr int N = N1 * N2;
    constexpr int MPT = (N * CW + NT - 1) / NT;
    constexpr int LPT = (CPR * N * CW + NT - 1) / NT;
    constexpr int LB = load_batch(LPT);
    constexpr int S1T = (NCR * N2 * CW + NT - 1) / NT;
    constexpr int S2T = (NCR * N1 * CW + NT - 1) / NT;
    static_assert(NCR >= 1 && NCR <= CPR && C0 + NCR <= GABE_NSC && NNX <= CPR, "rounds cover the components");
    static_assert(LPT % LB == 0, "loads must split into whole batches");
    {
        auto prefetch = [&] {
#pragma unroll
            for (int t = 0; t < LPT; t++) {
                const int e = threadIdx.x + t * NT;
                const int r = e / (N * CW), rem = e % (N * CW), i = rem / CW, col = rem % CW;
                if (e < NNX * N * CW && col < width)
                    pf[t] = Sh[(long)(C0 + NCR + r) * SM + (long)i * rstride + col];
                else pf[t].x = pf[t].y = (T)0;
            }
        };
        if constexpr (PF_IN) {
#pragma unroll
            for (int t = 0; t < LPT; t++) {
                const int e = threadIdx.x + t * NT;
                const int r = e / (N * CW), rem = e % (N * CW);
                if (e < NCR * N * CW) tile[r][trow<N2>(rem / CW)][rem % CW] = pf[t];
            }
        } else
#pragma unroll
        for (int t0 = 0; t0 < LPT; t0 += LB) {
            C v[LB];
#pragma unroll
            for (int t = 0; t < LB; t++) {
                const int e = threadIdx.x + (t0 + t) * NT;
                const int r = e / (N * CW), rem = e % (N * CW), i = rem / CW, col = rem % CW;
                if (e < NCR * N * CW && col < width)
                    v[t] = Sh[(long)(C0 + r) * SM + (long)i * rstride + col];
                else v[t].x = v[t].y = (T)0;
            }
#pragma unroll
            for (int t = 0; t < LB; t++) {
                const int e = threadIdx.x + (t0 + t) * NT;
                const int r = e / (N * CW), rem = e % (N * CW);
                if (e < NCR * N * CW) tile[r][trow<N2>(rem / CW)][rem % CW] = v[t];
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < S1T; t++) {
            const int task = threadIdx.x + t * NT;
            if (task < NCR * N2 * CW) {
                const int r = task / (N2 * CW), rem = task % (N2 * CW), col = rem % CW, n2 = rem / CW;
                C x[N1];
#pragma unroll
                for (int n1 = 0; n1 < N1; n1++) x[n1] = tile[r][trow2<N2>(n1, n2)][col];
                dft_reg<N1, T>(x);
#pragma unroll
                for (int k1 = 0; k1 < N1; k1++) {
                    C w = x[k1];
                    /* n2 = 0: twl[0] = (1, -0) leaves w unchanged up to the sign of a zero */
                    if (k1 != 0) w = cmul(w, twl[(n2 * k1) % N]);
                    tile[r][trow2<N2>(k1, n2)][col] = w;
                }
            }
        }
        __syncthreads();
        if constexpr (NNX > 0) prefetch();

        C y[S2T][N2];
#pragma unroll
        for (int t = 0; t < S2T; t++) {
            const int task = threadIdx.x + t * NT;
            if (task < NCR * N1 * CW) {
                const int r = task / (N1 * CW), rem = task % (N1 * CW), col = rem % CW, k1 = rem / CW;
#pragma unroll
                for (int n2 = 0; n2 < N2; n2++) y[t][n2] = tile[r][trow2<N2>(k1, n2)][col];
                dft_reg<N2, T>(y[t]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < S2T; t++) {
            const int task = threadIdx.x + t * NT;
            if (task < NCR * N1 * CW) {
                const int r = task / (N1 * CW), rem = task % (N1 * CW), col = rem % CW, k1 = rem / CW;
#pragma unroll
                for (int k2 = 0; k2 < N2; k2++) tile[r][trow<N2>(k1 + N1 * k2)][col] = y[t][k2];
            }
        }
        /* the transform's registers are free here: work that only needs to
         * be in flight (the mode loads of the last round) starts now */
        after_dft();
        __syncthreads();

        /* (e : S)/2 of the owned modes, off-diagonal components counted twice */
#pragma unroll
        for (int u = 0; u < MPT; u++) {
            const int e = threadIdx.x + u * NT;
            if (e < N * CW) {
                const int i = e / CW, col = e % CW;
#pragma unroll
                for (int r = 0; r < NCR; r++) {
                    const int cc = (C0 + r) % GABE_SC;   /* xx - zz, xy, xz, yy - zz, yz */
                    const int ia = cc < 3 ? 0 : 1;
                    const int ib = cc == 0 ? 0 : (cc == 1 || cc == 3 ? 1 : 2);
                    const T ua = ia == 0 ? bu[u][0] : bu[u][1];
                    const T va = ia == 0 ? bv[u][0] : bv[u][1];
                    const T ub = ib == 0 ? bu[u][0] : (ib == 1 ? bu[u][1] : bu[u][2]);
                    const T vb = ib == 0 ? bv[u][0] : (ib == 1 ? bv[u][1] : bv[u][2]);
                    const T w = ia == ib ? (T)0.5 : (T)1;
                    const T ep = w * (ua * ub - va * vb);
                    const T ex = w * (ua * vb + va * ub);
                    const C S = tile[r][trow<N2>(i)][col];
                    C(&acc)[MPT][2] = (C0 + r) / GABE_SC == 0 ? acc0 : acc1;
                    acc[u][0].x += ep * S.x;
                    acc[u][0].y += ep * S.y;
                    acc[u][1].x += ex * S.x;
                    acc[u][1].y += ex * S.y;
                }
            }
        }
        __syncthreads();
    }
}

/* rounds R, R + 1, ... of CPR components until all GABE_NSC are in; the last
 * round runs last_dft after its transform (the mode loads) */
template <typename T, int N1, int N2, int CW, int CPR, int NT, int R, typename Nothing, typename LastDft>
__device__ __forceinline__ void source_rounds(Nothing nothing, LastDft last_dft,
                                              typename cplx_t<T>::type (&tile)[CPR][N1 * N2 + N1][CW],
                                              const typename cplx_t<T>::type *twl,
                                              const typename cplx_t<T>::type *Sh, long SM, long rstride, int width,
                                              const T (&bu)[(N1 * N2 * CW + NT - 1) / NT][3],
                                              const T (&bv)[(N1 * N2 * CW + NT - 1) / NT][3],
                                              typename cplx_t<T>::type (&acc0)[(N1 * N2 * CW + NT - 1) / NT][2],
                                              typename cplx_t<T>::type (&acc1)[(N1 * N2 * CW + NT - 1) / NT][2],
                                              typename cplx_t<T>::type (&pf)[gw_lpt(N1 * N2, CW, CPR, NT)])
{
    constexpr int C0 = R * CPR;
    constexpr int NCR = GABE_NSC - C0 < CPR ? GABE_NSC - C0 : CPR;
    constexpr int LEFT = GABE_NSC - C0 - NCR;
    constexpr int NNX = LEFT < CPR ? LEFT : CPR;
    if constexpr (NNX == 0) {
        source_round<T, N1, N2, CW, CPR, NT, C0, NCR, 0, (R > 0)>(last_dft, tile, twl, Sh, SM, rstride, width, bu, bv,
                                                                 acc0, acc1, pf);
    } else {
        source_round<T, N1, N2, CW, CPR, NT, C0, NCR, NNX, (R > 0)>(nothing, tile, twl, Sh, SM, rstride, width, bu, bv,
                                                                   acc0, acc1, pf);
        source_rounds<T, N1, N2, CW, CPR, NT, R + 1>(nothing, last_dft, tile, twl, Sh, SM, rstride, width, bu, bv,
                                                     acc0, acc1, pf);
    }
}

/* waves per SIMD the register allocation must leave room for: the FP32
 * 256^3 tiles of 4 columns x 2 or 3 components per round on 256 threads fit
 * 3 without spilling */
constexpr int gw_min_waves(int tsize, int N1, int N2, int CW, int CPR, int NT)
{
    return tsize == 4 && N1 == 16 && N2 == 16 && CW == 4 && CPR <= 3 && NT == 256 ? 3 : 1;
}

template <typename T, int N1, int N2, int CW, int CPR, int NT>
__global__ __launch_bounds__(NT)
__attribute__((amdgpu_waves_per_eu(gw_min_waves((int)sizeof(T), N1, N2, CW, CPR, NT))))
void k_gw_ifft(GwFftArgs<T> A)
{
    using C = typename cplx_t<T>::type;
    constexpr int N = N1 * N2, KP = N / 2 + 1;
    constexpr int NCH = (KP + CW - 1) / CW;
    constexpr int MPT = (N * CW + NT - 1) / NT;        /* modes owned per thread */
    static_assert(sizeof(C) * (CPR * (N + N1) * CW + N) <= 160 * 1024, "tile exceeds the LDS of one CU");
    __shared__ C tile[CPR][N + N1][CW];
    __shared__ C twl[N];

    /* XCD-aware order: adjacent kz chunks share 128-byte lines, so keep them
     * on one XCD (one L2), working at the same time */
    int b = blockIdx.x;
    if (gridDim.x % 8 == 0) b = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
    /* Workgroup b < N nch: kz chunk q at one j.  With a separate kz = N/2
     * plane, the last N / CW workgroups take CW adjacent j of that plane.
     * Column col of the tile is the mode (j0 + col jstep, pz0 + col pzstep). */
    const int nch = A.nyq ? N / 2 / CW : NCH;
    const bool plane = b >= N * nch;
    int q, j0, jstep, pz0, pzstep, width;
    long SM, rstride;
    const C *Sh;
    if (!plane) {
        q = b % nch;
        j0 = b / nch; jstep = 0;
        pz0 = q * CW; pzstep = 1;
        width = min(CW, KP - q * CW);
        SM = (long)N * N * A.kpad;
        rstride = (long)N * A.kpad;
        Sh = reinterpret_cast<const C *>(A.shat) + ((long)j0 * A.kpad + (long)q * CW);
    } else {
        q = b - N * nch;
        j0 = q * CW; jstep = 1;
        pz0 = N / 2; pzstep = 0;
        width = CW;
        SM = (long)N * N;
        rstride = N;
        Sh = reinterpret_cast<const C *>(A.nyq) + (long)q * CW;
    }

    /* twiddles exp(-2 pi i e/N) */
    for (int e = threadIdx.x; e < N; e += NT) twl[e] = reinterpret_cast<const C *>(A.tw)[e];

    /* owned modes: e = tid + u NT -> (i = e / CW, col = e % CW) */
    T bu[MPT][3], bv[MPT][3];
    C acc0[MPT][2], acc1[MPT][2];
#pragma unroll
    for (int u = 0; u < MPT; u++) {
        const int e = threadIdx.x + u * NT, i = e / CW, col = e % CW;
        const int jj = j0 + col * jstep, py = jj <= N / 2 ? jj : jj - N;
        const int px = i <= N / 2 ? i : i - N, pz = pz0 + col * pzstep;
        const bool has = e < N * CW && col < width && (px != 0 || py != 0 || pz != 0);
        double du[3] = {0., 0., 0.}, dv[3] = {0., 0., 0.};
        if (has) pol_uv(px, py, pz, du, dv);
#pragma unroll
        for (int a = 0; a < 3; a++) { bu[u][a] = (T)du[a]; bv[u][a] = (T)dv[a]; }
#pragma unroll
        for (int p = 0; p < 2; p++) { acc0[u][p].x = acc0[u][p].y = (T)0; acc1[u][p].x = acc1[u][p].y = (T)0; }
    }
    /* h and l of the owned modes, loaded during the last round (their latency
     * hidden behind its projection) */
    C *hh = reinterpret_cast<C *>(A.h);
    C *ll = reinterpret_cast<C *>(A.l);
    C h0v[MPT][2], l0v[MPT][2];
    auto mode_index = [&](int u) -> long {
        const int e = threadIdx.x + u * NT, i = e / CW, col = e % CW;
        const int jj = j0 + col * jstep, pz = pz0 + col * pzstep;
        /* tiled mode order [j][kz/4][i][kz%4] */
        return (((long)jj * A.nt4 + (pz >> 2)) * N + i) * 4 + (pz & 3);
    };
    auto load_modes = [&]() {
#pragma unroll
        for (int u = 0; u < MPT; u++) {
            const int e = threadIdx.x + u * NT, col = e % CW;
            if (e >= N * CW || col >= width) continue;
            const long m = mode_index(u);
#pragma unroll
            for (int pol = 0; pol < 2; pol++) {
                h0v[u][pol] = hh[2 * m + pol];
                l0v[u][pol] = ll[2 * m + pol];
            }
        }
    };
    auto nothing = [] {};
    /* the next round's tiles, carried in registers from one round to the next */
    C pf[gw_lpt(N, CW, CPR, NT)];
    source_rounds<T, N1, N2, CW, CPR, NT, 0>(nothing, load_modes, tile, twl, Sh, SM, rstride, width, bu, bv, acc0, acc1,
                                             pf);

    /* both midpoint-RK2 stages of the owned modes (derivs, g2evolutionpert.cpp:486-501) */
    const StepScalars *ss = A.ss;
    const double k2u = 4. * M_PI * M_PI / A.L / A.L;
    const T fr0 = (T)(3. * ss->adot0 / ss->a0), fr1 = (T)(3. * ss->adot1 / ss->a1);
    const T om0 = (T)(16. * M_PI / (ss->a0 * ss->a0) * A.dx3), om1 = (T)(16. * M_PI / (ss->a1 * ss->a1) * A.dx3);
    const double ia0 = 1. / (ss->a0 * ss->a0), ia1 = 1. / (ss->a1 * ss->a1);
    const T hdt = (T)(0.5 * A.dt), dt = (T)A.dt;
#pragma unroll
    for (int u = 0; u < MPT; u++) {
        const int e = threadIdx.x + u * NT, i = e / CW, col = e % CW;
        if (e >= N * CW || col >= width) continue;
        const int jj = j0 + col * jstep, py = jj <= N / 2 ? jj : jj - N;
        const int px = i <= N / 2 ? i : i - N, pz = pz0 + col * pzstep;
        const double p2 = (double)(px * px + py * py + pz * pz);
        const T kk0 = (T)(k2u * p2 * ia0), kk1 = (T)(k2u * p2 * ia1);
        const long m = mode_index(u);
#pragma unroll
        for (int pol = 0; pol < 2; pol++) {
            const C h0 = h0v[u][pol], l0 = l0v[u][pol];
            const C T0 = acc0[u][pol], T1 = acc1[u][pol];
            C h1, l1, hn, ln;
            h1.x = h0.x + hdt * l0.x;
            h1.y = h0.y + hdt * l0.y;
            l1.x = l0.x + hdt * (-kk0 * h0.x - fr0 * l0.x + om0 * T0.x);
            l1.y = l0.y + hdt * (-kk0 * h0.y - fr0 * l0.y + om0 * T0.y);
            hn.x = h0.x + dt * l1.x;
            hn.y = h0.y + dt * l1.y;
            ln.x = l0.x + dt * (-kk1 * h1.x - fr1 * l1.x + om1 * T1.x);
            ln.y = l0.y + dt * (-kk1 * h1.y - fr1 * l1.y + om1 * T1.y);
            hh[2 * m + pol] = hn;
            ll[2 * m + pol] = ln;
        }
    }
}

template <typename T, int N1, int N2, int CW, int CPR, int NT>
void launch_cfg(const GwFftArgs<T> &a, hipStream_t s)
{
    constexpr int N = N1 * N2, NCH = (N / 2 + 1 + CW - 1) / CW;
    static_assert((N / 2) % CW == 0 && N % CW == 0, "kz chunks and plane chunks must be whole");
    const int grid = a.nyq ? N * (N / 2 / CW) + N / CW : N * NCH;
    hipLaunchKernelGGL((k_gw_ifft<T, N1, N2, CW, CPR, NT>), dim3(grid), dim3(NT), 0, s, a);
}

} // namespace

bool gabe_gwfft_supported(int N, bool fp32) { return N >= 16 && N <= (fp32 ? 512 : 256) && (N & (N - 1)) == 0; }

/* Tile variants (kz columns per workgroup CW, source components per round
 * CPR, threads NT) at the sizes where more than one was worth keeping;
 * variant 1 is the built-in default, the rest are the alternatives the
 * autotuner times. */
int gabe_gwfft_tiles(int N, bool fp32)
{
    if (!gabe_gwfft_supported(N, fp32)) return 0;
    if (fp32) return N == 256 ? 7 : N == 512 ? 4 : 1;
    return N == 256 ? 3 : 1;
}

template <>
void launch_gw_ifft<float>(const GwFftArgs<float> &a, hipStream_t s)
{
    switch (a.N) {
    case 16: launch_cfg<float, 4, 4, 8, 5, 256>(a, s); break;
    case 32: launch_cfg<float, 8, 4, 8, 5, 256>(a, s); break;
    case 64: launch_cfg<float, 8, 8, 8, 5, 512>(a, s); break;
    case 128: launch_cfg<float, 16, 8, 8, 5, 1024>(a, s); break;
    case 256:
        /* default: 4 columns, 2 of the 10 components per round, 256 threads
         * (3 workgroups per CU); 3 components per round measures the same */
        switch (a.tile) {
        case 2: launch_cfg<float, 16, 16, 4, 3, 256>(a, s); break;
        case 3: launch_cfg<float, 16, 16, 4, 5, 512>(a, s); break;
        case 4: launch_cfg<float, 16, 16, 4, 5, 256>(a, s); break;
        case 5: launch_cfg<float, 16, 16, 8, 5, 512>(a, s); break;
        case 6: launch_cfg<float, 16, 16, 8, 2, 512>(a, s); break;
        case 7: launch_cfg<float, 16, 16, 4, 10, 1024>(a, s); break;
        default: launch_cfg<float, 16, 16, 4, 2, 256>(a, s); break;
        }
        break;
    case 512:
        switch (a.tile) {
        case 2: launch_cfg<float, 32, 16, 4, 2, 512>(a, s); break;
        case 3: launch_cfg<float, 32, 16, 2, 5, 256>(a, s); break;
        case 4: launch_cfg<float, 32, 16, 2, 5, 512>(a, s); break;
        default: launch_cfg<float, 32, 16, 4, 3, 512>(a, s); break;
        }
        break;
    default: break;
    }
}

/* FP64: complex values are 16 bytes, so the tiles are half as wide in kz to
 * keep the LDS and register footprint of the FP32 ones */
template <>
void launch_gw_ifft<double>(const GwFftArgs<double> &a, hipStream_t s)
{
    switch (a.N) {
    case 16: launch_cfg<double, 4, 4, 8, 5, 256>(a, s); break;
    case 32: launch_cfg<double, 8, 4, 8, 5, 256>(a, s); break;
    case 64: launch_cfg<double, 8, 8, 4, 5, 256>(a, s); break;
    case 128: launch_cfg<double, 16, 8, 4, 5, 512>(a, s); break;
    case 256:
        switch (a.tile) {
        case 2: launch_cfg<double, 16, 16, 2, 2, 256>(a, s); break;
        case 3: launch_cfg<double, 16, 16, 2, 3, 256>(a, s); break;
        default: launch_cfg<double, 16, 16, 2, 5, 256>(a, s); break;
        }
        break;
    default: break;
    }
}
