/*
 * gabe_kernels.hip -- CDNA4 (gfx950) kernels of the lattice evolution.
 *
 * One step (expansion_type 1, step(), g2functions.cpp:234-270) is
 *   sweep A  -> finalize -> [R2C of the GABE_NSC stress arrays -> GW mode kernel]
 *                        -> sweep B
 * Sweep A reads the step-start state once (7-point halo) and produces every
 * quantity that needs a global reduction: the energies of the step-start state
 * (the reference's end-of-step calcEnergy(0), g2functions.cpp:269-270), the
 * midpoint energies (calcEnergy(1), :251) and the mean potential forces, plus
 * the gradient stress of both RK2 stages for the tensor source
 * (g2evolutionpert.cpp:165-180).  The FP64 finalize turns the partial sums into
 * a, adot and the energies on the device.  Sweep B recomputes the midpoint from
 * the same 1-deep halo and writes the new state.  The GW mode kernel fuses the
 * TT projection and both RK2 stages of every tensor mode.
 *
 * Storage type T is float (FP32 fluctuations around an FP64 background) or
 * double (full fields, the FP64 form).  Potentials, forces and every
 * reduction are FP64 in both; the stencils and stress accumulate in T.
 */
#include <hip/hip_runtime.h>

#include <type_traits>

#include "gabe_fft_dev.h"
#include "gabe_internal.h"
#include "gabe_modes.h"

namespace {

template <typename T>
struct cplx_of;
template <>
struct cplx_of<float> { using type = float2; };
template <>
struct cplx_of<double> { using type = double2; };

/* midpoint field value f + (dt/2) fdot, one rounding; sweeps A and B use this
 * same function so both see identical midpoint fields */
template <typename T>
__device__ __forceinline__ T midpt(T f, T v, T hdt) { return fma(v, hdt, f); }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

/* fixed-order block reduction of NS sums; one partial per workgroup per sum.
 * A workgroup of WIDE x GABE_BLOCK threads stands for WIDE consecutive
 * GABE_BLOCK workgroups: it writes its sum into the first of their slots and
 * zeros into the others, so the finalize reads the same nblk slots. */
template <int NS, int WIDE>
__device__ void reduce_partials(double (&acc)[NS], double (*red)[NS], double *partials, int nblk, int bid)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        double v = wave_sum(acc[s]);
        if (lane == 0) red[w][s] = v;
    }
    __syncthreads();
    if (threadIdx.x < NS) {
        const int s = threadIdx.x;
        double tot = (red[0][s] + red[1][s]) + (red[2][s] + red[3][s]);
        if constexpr (WIDE == 2) tot += (red[4][s] + red[5][s]) + (red[6][s] + red[7][s]);
        partials[(long)s * nblk + WIDE * bid] = tot;
        for (int q = 1; q < WIDE; q++) partials[(long)s * nblk + WIDE * bid + q] = 0.;
    }
}

/* the same with an LDS area of its own; workgroups with a free LDS buffer
 * call reduce_partials on it instead */
template <int NS, int WIDE = 1>
__device__ void block_store_partials(double (&acc)[NS], double *partials, int nblk, int bid)
{
    __shared__ double red[WIDE * GABE_BLOCK / 64][NS];
    reduce_partials<NS, WIDE>(acc, red, partials, nblk, bid);
}

/* ------------------------------------------------------------------ sweeps */
/* Row R2C of the GABE_NSC stress components of the row_block(N) / N k-rows
 * one sweep workgroup holds, for npl (1 or 2) consecutive planes from plane
 * i, written into the half-spectrum rows of shat (the library row transform
 * it replaces).  The components arrive in LDS as GABE_NSC / 2 complex rows,
 * z = S_c + i S_(c+1) for c = 0, 2, ..., 8 (the pairs (xx,xy), (xz,yy) of
 * stage 0, (yz of stage 0, xx of stage 1), (xy,xz), (yy,yz) of stage 1); one
 * complex transform of length N per pair (two-step N = N1 N2, in-register
 * DFTs) and the split
 *   S_c^(m) = (Z(m) + conj Z(N-m)) / 2,  S_(c+1)^(m) = (Z(m) - conj Z(N-m)) / 2i
 * give both real rows' transforms, m = 0..N/2.  Rows are stored N/2 slots
 * long: S(0) and S(N/2), both real, share slot 0 as S(0) + i S(N/2) (the
 * j-column pass separates them).  Two planes per call halve the barriers per
 * plane, which stall the sweep's loads. */
/* LDS layout of the row sequences: element n of sequence sq at
 * sq * rseq_stride(N) + rpad(n), one padding slot every 16 elements, so the
 * stride-16 reads of the second DFT step and the sequences' starts fall on
 * different banks */
__host__ __device__ constexpr int rpad(int n) { return n + (n >> 4); }
__host__ __device__ constexpr int rseq_stride(int N) { return N + N / 16; }
/* threads of a row-transforming sweep workgroup: it holds at least one whole
 * k-row (512-site rows take a 512-thread workgroup) */
__host__ __device__ constexpr int row_block(int N) { return N > GABE_BLOCK ? N : GABE_BLOCK; }
/* planes per transform call: 2 in FP32, 1 in FP64 or for 512-site rows (the
 * same LDS bytes) */
template <typename T>
__host__ __device__ constexpr int row_planes(int N) { return sizeof(T) == 4 && N <= GABE_BLOCK ? 2 : 1; }
/* complex slots of one buffer: row_planes x GABE_NSC / 2 pairs x rows per workgroup */
template <typename T>
__host__ __device__ constexpr int rbuf_size(int N)
{
    return row_planes<T>(N) * (GABE_NSC / 2) * (row_block(N) / N) * rseq_stride(N);
}

template <typename T, int N>
__device__ __forceinline__ void row_r2c(typename gabe_fft::cplx_t<T>::type *zl, typename gabe_fft::cplx_t<T>::type *zo,
                                       const typename gabe_fft::cplx_t<T>::type *tw, const SweepArgs<T> &A, int i,
                                       int j0, int npl)
{
    using gabe_fft::cmul;
    using C = typename gabe_fft::cplx_t<T>::type;
    constexpr int N1 = N == 16 ? 4 : (N <= 64 ? 8 : 16);
    constexpr int N2 = N / N1;
    constexpr int BLK = row_block(N);
    constexpr int NP = GABE_NSC / 2;                     /* complex sequences (pairs) per row */
    constexpr int R = BLK / N, SPP = NP * R, H = N / 2;  /* sequences per plane, slots per row */
    constexpr int SS = rseq_stride(N);
    static_assert((N1 - 1) * (N2 - 1) < N, "twiddle exponents index one table of N");
    const int nseq = SPP * npl;
    for (int task = threadIdx.x; task < nseq * N2; task += BLK) {
        const int sq = task / N2, n2 = task - sq * N2;
        C *b = zl + sq * SS;
        C x[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) x[n1] = b[rpad(N2 * n1 + n2)];
        gabe_fft::dft_reg<N1, T>(x);
        /* lanes with n2 = 0 multiply by tw[0] = (1, -0), which returns the value
         * unchanged (up to the sign of a zero) and keeps the lanes together */
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) {
            C v = x[k1];
            if (k1 != 0) {
                /* W^e for e = n2 k1 < N, from the half table: W^(e + N/2) = -W^e */
                const int e = n2 * k1;
                const bool ng = e >= N / 2;
                C w = tw[ng ? e - N / 2 : e];
                if (ng) { w.x = -w.x; w.y = -w.y; }
                v = cmul(v, w);
            }
            b[rpad(N2 * k1 + n2)] = v;
        }
    }
    __syncthreads();
    for (int task = threadIdx.x; task < nseq * N1; task += BLK) {
        const int sq = task / N1, k1 = task - sq * N1;
        const C *b = zl + sq * SS;
        C *o = zo + sq * SS;
        C y[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; n2++) y[n2] = b[rpad(N2 * k1 + n2)];
        gabe_fft::dft_reg<N2, T>(y);
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) o[rpad(k1 + N1 * k2)] = y[k2];
    }
    __syncthreads();
    const long comp = (long)N * N * A.kpad;
    C *out = (C *)A.shat;
    const T h = (T)0.5;
    if constexpr (R == 1 && BLK == 2 * H) {
        /* one k-row per workgroup (N = 256, 512): thread t keeps slot
         * m = t mod H and takes sequences s0, s0 + 2, ... (s0 = t / H), i.e.
         * pairs s0, s0 + 2, s0 + 4 (below NP) of plane 0, then of plane 1.
         * The slot's LDS offsets and row address are computed once. */
        const int m = threadIdx.x % H, s0 = threadIdx.x / H;
        const bool slot0 = m == 0;   /* slot 0 packs S(0) + i S(N/2) */
        const C *o = zo + s0 * SS;   /* sequence s0 + 2 q of plane pl at o + (NP pl + 2 q) SS */
        const int iz = rpad(m), iw = rpad(slot0 ? N / 2 : N - m);
        C *dst = out + (long)(2 * s0) * comp + ((long)i * N + j0) * A.kpad + m;
        const long plane_stride = (long)N * A.kpad;
        /* planes one at a time: unrolling both keeps too many values live */
#pragma unroll 1
        for (int pl = 0; pl < npl; pl++, o += NP * SS, dst += plane_stride) {
#pragma unroll
            for (int q = 0; q < (NP + 1) / 2; q++) {   /* pair s0 + 2 q of plane pl */
                if (s0 + 2 * q >= NP) break;
                const C Z = o[2 * q * SS + iz], W = o[2 * q * SS + iw];
                C s0v, s1v;
                s0v.x = slot0 ? Z.x : h * (Z.x + W.x);
                s0v.y = slot0 ? W.x : h * (Z.y - W.y);
                s1v.x = slot0 ? Z.y : h * (Z.y + W.y);
                s1v.y = slot0 ? W.y : -h * (Z.x - W.x);
                C *d = dst + (long)(4 * q) * comp;
                d[0] = s0v;
                d[comp] = s1v;
            }
        }
        return;
    }
    for (int task = threadIdx.x; task < nseq * H; task += BLK) {
        const int sq = task / H, m = task - sq * H;
        const int pl = sq / SPP, ps = sq - pl * SPP, pr = ps / R, r = ps - pr * R;
        const C *o = zo + sq * SS;
        /* m = 0 packs both real values: slot 0 = S(0) + i S(N/2) */
        const C Z = o[rpad(m)], W = o[rpad(m == 0 ? N / 2 : N - m)];
        const long row = ((long)(i + pl) * N + (j0 + r)) * A.kpad + m;
        C s0, s1;
        if (m == 0) {
            s0.x = Z.x; s0.y = W.x;
            s1.x = Z.y; s1.y = W.y;
        } else {
            s0.x = h * (Z.x + W.x);
            s0.y = h * (Z.y - W.y);
            s1.x = h * (Z.y + W.y);
            s1.y = -h * (Z.x - W.x);
        }
        out[(2 * pr) * comp + row] = s0;
        out[(2 * pr + 1) * comp + row] = s1;
    }
}

/* FP32 sweep B: the plane march of sweep_body for the new state, with the
 * centres of plane i + 2 loaded while plane i is evaluated.  Four plane
 * buffers rotate: plane i reads planes i - 1, i, i + 1 from three of them and
 * loads plane i + 2 into the fourth, whose plane i - 2 is no longer needed.
 * The march is unrolled by four with the buffer roles fixed at compile time,
 * so no register is copied (a copy out of a register still being loaded
 * would wait for the load).  The loads are issued after the plane's
 * neighbour loads, so the in-order waits for those leave them in flight
 * through the force evaluation and the stores.  (Loading further ahead
 * gains nothing: the wait for the next plane's neighbour loads also waits
 * for every load issued before them.)  The arithmetic is sweep_body's,
 * expression for expression. */
/* XCD-aware block order over the nhw hardware workgroups a lattice has in
 * its launch (nhw >= gx gy).  Consecutive hardware workgroups go round-robin
 * to the 8 XCDs, each with its own L2; every XCD gets a contiguous range of
 * logical blocks instead, so the j+-1 neighbour rows a workgroup reads are
 * being read by workgroups on the same XCD at the same time.  With nhw a
 * multiple of 8 the map is a bijection of [0, nhw) and logical blocks past
 * gx gy are padding (a batched launch pads every lattice's range to a
 * multiple of 8); otherwise the order is the hardware's.  Returns false for
 * a padding workgroup. */
__device__ __forceinline__ bool xcd_block(int hw, int gx, int gy, int nhw, int &bx, int &by)
{
    int logical = hw;
    if (nhw % GABE_NXCD == 0) logical = (hw % GABE_NXCD) * (nhw / GABE_NXCD) + hw / GABE_NXCD;
    bx = logical % gx;
    by = logical / gx;
    return logical < gx * gy;
}

template <int Q>
using plane_role = std::integral_constant<int, Q>;

template <typename T, int NF>
__device__ __forceinline__ void sweep_b_ahead(const SweepArgs<T> &A, int hw, int gx, int gy, int nhw)
{
    const int N = A.N;
    const long N2 = A.N2;
    int bx, by;
    if (!xcd_block(hw, gx, gy, nhw, bx, by)) return;
    by += A.by0;
    const long p = (long)bx * GABE_BLOCK + threadIdx.x;
    const bool valid = p < N2;
    const long pp = valid ? p : 0;
    const int j = (int)(pp / N), k = (int)(pp - (long)j * N);
    /* byte offsets within a plane (N^2 sites of T fit 32 bits): each load is
     * a uniform plane base plus a 32-bit lane offset */
    const unsigned ocen = (unsigned)pp * sizeof(T);
    const unsigned okp = (unsigned)((long)j * N + (k + 1 == N ? 0 : k + 1)) * sizeof(T);
    const unsigned okm = (unsigned)((long)j * N + (k == 0 ? N - 1 : k - 1)) * sizeof(T);
    const unsigned ojp = (unsigned)((long)(j + 1 == N ? 0 : j + 1) * N + k) * sizeof(T);
    const unsigned ojm = (unsigned)((long)(j == 0 ? N - 1 : j - 1) * N + k) * sizeof(T);
    auto at = [&](const T *arr, int plane_i, unsigned off) -> T {
        return gld(reinterpret_cast<const T *>(reinterpret_cast<const char *>(arr + (long)plane_i * N2) + off));
    };
    const int i0 = by * A.ci;
    const int i1 = min(i0 + A.ci, N);

    const T idx2 = (T)A.inv_dx2;
    const double dt = A.dt, hdt = 0.5 * dt;
    const T hdtT = (T)hdt;
    const StepScalars *ss = A.ss;
    const double ia02 = 1. / (ss->a0 * ss->a0), ia12 = 1. / (ss->a1 * ss->a1);
    const double H0 = ss->adot0 / ss->a0, H1 = ss->adot1 / ss->a1;
    double pb0[2] = {0., 0.}, pb1[2] = {0., 0.}, mvp0[2] = {0., 0.}, mvp1[2] = {0., 0.};
#pragma unroll
    for (int f = 0; f < NF; f++) {
        pb0[f] = ss->pbar0[f]; pb1[f] = ss->pbar1[f];
        mvp0[f] = ss->meanVp0[f]; mvp1[f] = ss->meanVp1[f];
    }

    /* centre values of four consecutive planes, field and velocity */
    T bf[4][NF], bv[4][NF];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int ic = (i0 - 1 + q + N) % N;   /* planes i0 - 1 .. i0 + 1 */
#pragma unroll
        for (int f = 0; f < NF; f++) {
            bf[q][f] = at(A.f[f], ic, ocen);
            bv[q][f] = at(A.v[f], ic, ocen);
        }
    }

    /* plane i with planes i - 1, i, i + 1 in buffers Q, Q + 1, Q + 2 (mod 4);
     * plane i + 2 goes into buffer Q + 3 */
    auto plane = [&](auto role, const int i) {
        constexpr int Q = decltype(role)::value;
        constexpr int QM = Q % 4, QC = (Q + 1) % 4, QP = (Q + 2) % 4, QN = (Q + 3) % 4;
        T lap0[NF], lap1[NF];
        double full0[2] = {0., 0.}, full1[2] = {0., 0.};
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const T fm = bf[QM][f], fc = bf[QC][f], fp = bf[QP][f];
            const T vm = bv[QM][f], vc = bv[QC][f], vp = bv[QP][f];
            const T fkp = at(A.f[f], i, okp), fkm = at(A.f[f], i, okm);
            const T fjp = at(A.f[f], i, ojp), fjm = at(A.f[f], i, ojm);
            lap0[f] = (((fm + fp) + (fjm + fjp)) + (fkm + fkp) - (T)6 * fc) * idx2;
            full0[f] = pb0[f] + (double)fc;
            const T vkp = at(A.v[f], i, okp), vkm = at(A.v[f], i, okm);
            const T vjp = at(A.v[f], i, ojp), vjm = at(A.v[f], i, ojm);
            const T m1 = midpt(fm, vm, hdtT), p1 = midpt(fp, vp, hdtT);
            const T jm1 = midpt(fjm, vjm, hdtT), jp1 = midpt(fjp, vjp, hdtT);
            const T km1 = midpt(fkm, vkm, hdtT), kp1 = midpt(fkp, vkp, hdtT);
            const T c1 = midpt(fc, vc, hdtT);
            lap1[f] = (((m1 + p1) + (jm1 + jp1)) + (km1 + kp1) - (T)6 * c1) * idx2;
            full1[f] = pb1[f] + (double)c1;
        }
        __builtin_amdgcn_sched_barrier(0);
        const int i2 = i + 2 < N ? i + 2 : i + 2 - N;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            bf[QN][f] = at(A.f[f], i2, ocen);
            bv[QN][f] = at(A.v[f], i2, ocen);
        }
        __builtin_amdgcn_sched_barrier(0);
        if (valid) {
            double V0, V1, dV0[2], dV1[2];
            gabe_V_dV(A.model, full0[0], full0[1], V0, dV0[0], dV0[1]);
            gabe_V_dV(A.model, full1[0], full1[1], V1, dV1[0], dV1[1]);
#pragma unroll
            for (int f = 0; f < NF; f++) {
                const double Vp0 = dV0[f];
                const double Vp1 = dV1[f];
                const double dv0 = (double)bv[QC][f];
                /* F(0) and F(1) of ddfield (g2functions.cpp:193-196) minus the
                 * background force */
                const double dF0 = (double)lap0[f] * ia02 - (Vp0 - mvp0[f]) - 3. * H0 * dv0;
                const double dv1 = dv0 + hdt * dF0;
                const double dF1 = (double)lap1[f] * ia12 - (Vp1 - mvp1[f]) - 3. * H1 * dv1;
                gst(reinterpret_cast<T *>(reinterpret_cast<char *>(A.fo[f] + (long)i * N2) + ocen),
                    (T)((double)bf[QC][f] + dt * dv1));
                gst(reinterpret_cast<T *>(reinterpret_cast<char *>(A.vo[f] + (long)i * N2) + ocen),
                    (T)(dv0 + dt * dF1));
            }
        }
    };
    int i = i0;
    for (; i + 3 < i1; i += 4) {
        plane(plane_role<0>{}, i);
        plane(plane_role<1>{}, i + 1);
        plane(plane_role<2>{}, i + 2);
        plane(plane_role<3>{}, i + 3);
    }
    if (i < i1) plane(plane_role<0>{}, i);
    if (i + 1 < i1) plane(plane_role<1>{}, i + 1);
    if (i + 2 < i1) plane(plane_role<2>{}, i + 2);
}

/* Sweeps with every operand from LDS (sweep B, and sweep A without tensor
 * modes).  A workgroup owns a TK x TJ tile of (k, j) columns (TK TJ =
 * GABE_BLOCK, one column per thread; TK = 64, 32 or 16, the largest that
 * divides N) and marches down its planes of i.  Each plane's tile with its
 * four one-site edges, (TJ + 2)(TK + 2) slots less the corners, sits in a
 * ring of four plane slots (planes i - 1, i, i + 1 read while i + 2 is
 * written); planes i + 2 and i + 3 are in flight in registers, two loads per
 * thread and array: its own site and one edge site (threads past the
 * 2 TK + 2 TJ edge sites load their own site again).  So the only global
 * loads a plane waits for were issued two planes earlier, and the lateral
 * neighbours cost LDS reads instead of L2 round trips.  The arithmetic is the
 * per-thread sweeps', operand for operand; sweep A's sums are grouped by
 * tile instead of by row segment, so its partial sums (not its fields) differ
 * from the per-thread sweep's in the last bits. */
constexpr int SBT_SLOT = 6 * 66;   /* (TJ + 2)(TK + 2) at TK = 64, the largest of the shapes */

__device__ __forceinline__ int sbt_tk(int N) { return N % 64 == 0 ? 64 : (N % 32 == 0 ? 32 : 16); }

template <typename T, int NF, int MODE>
__device__ __forceinline__ void sweep_tile(const SweepArgs<T> &A, int hw, int gx, int gy, int nhw)
{
    static_assert(MODE == SW_B_TILE || MODE == SW_A_TILE, "tiled sweeps: B, and A without tensor modes");
    constexpr bool SB = MODE == SW_B_TILE;
    constexpr int NA = 2 * NF;   /* arrays: f[0 .. NF), then v[0 .. NF) */
    constexpr int NS = gabe_nsums(NF);
    __shared__ T ring[4][NA][SBT_SLOT];
    const int N = A.N;
    const long N2 = A.N2;
    int bx, by;
    if (!xcd_block(hw, gx, gy, nhw, bx, by)) return;
    by += A.by0;
    const int bid = by * gx + bx;
    const int TK = sbt_tk(N), TJ = GABE_BLOCK / TK, W = TK + 2;
    const int ntk = N / TK;
    const int k0 = (bx % ntk) * TK, j0 = (bx / ntk) * TJ;
    const int t = threadIdx.x, tr = t / TK, tc = t - tr * TK;
    const unsigned ocen = (unsigned)((j0 + tr) * N + k0 + tc) * sizeof(T);
    const int lc = (tr + 1) * W + tc + 1;
    /* the thread's edge site: rows j0 - 1 and j0 + TJ, then columns k0 - 1
     * and k0 + TK (periodic) */
    unsigned oh = ocen;
    int lh = lc;
    if (t < 2 * TK + 2 * TJ) {
        int hj, hk;
        if (t < TK) {
            hj = j0 - 1; hk = k0 + t; lh = t + 1;
        } else if (t < 2 * TK) {
            hj = j0 + TJ; hk = k0 + t - TK; lh = (TJ + 1) * W + t - TK + 1;
        } else if (t < 2 * TK + TJ) {
            hj = j0 + t - 2 * TK; hk = k0 - 1; lh = (t - 2 * TK + 1) * W;
        } else {
            hj = j0 + t - 2 * TK - TJ; hk = k0 + TK; lh = (t - 2 * TK - TJ + 1) * W + TK + 1;
        }
        hj = hj < 0 ? hj + N : (hj >= N ? hj - N : hj);
        hk = hk < 0 ? hk + N : (hk >= N ? hk - N : hk);
        oh = (unsigned)(hj * N + hk) * sizeof(T);
    }
    auto at = [&](const T *arr, int plane_i, unsigned off) -> T {
        return gld(reinterpret_cast<const T *>(reinterpret_cast<const char *>(arr + (long)plane_i * N2) + off));
    };
    const int i0 = by * A.ci;
    const int i1 = min(i0 + A.ci, N);
    /* plane p of the march, p in [i0 - 1, i1]; loads past i1 read plane i1
     * again (an L2 hit) into a slot that is never read */
    auto wrapped = [&](int p) { p = min(p, i1); return p < 0 ? p + N : (p >= N ? p - N : p); };
    auto load = [&](T (&r)[NA][2], int p) {
        const int pw = wrapped(p);
#pragma unroll
        for (int a = 0; a < NA; a++) {
            const T *arr = a < NF ? A.f[a] : A.v[a - NF];
            r[a][0] = at(arr, pw, ocen);
            r[a][1] = at(arr, pw, oh);
        }
    };
    auto put = [&](const T (&r)[NA][2], int s) {
#pragma unroll
        for (int a = 0; a < NA; a++) {
            ring[s][a][lc] = r[a][0];
            ring[s][a][lh] = r[a][1];
        }
    };

    const T idx2 = (T)A.inv_dx2;
    const double dt = A.dt, hdt = 0.5 * dt;
    const T hdtT = (T)hdt;
    /* scalars through global pointers, so that the batched launch (whose
     * pointers come from its descriptor table) reads them as scalars too:
     * sweep B the frozen step values, sweep A the current state */
    double ia02, ia12 = 0., H0 = 0., H1 = 0.;
    double pb0[2] = {0., 0.}, pb1[2] = {0., 0.}, vb0[2] = {0., 0.}, mvp0[2] = {0., 0.}, mvp1[2] = {0., 0.};
    if constexpr (SB) {
        const StepScalars *ss = A.ss;
        const double a0 = gld(&ss->a0), a1 = gld(&ss->a1);
        ia02 = 1. / (a0 * a0);
        ia12 = 1. / (a1 * a1);
        H0 = gld(&ss->adot0) / a0;
        H1 = gld(&ss->adot1) / a1;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            pb0[f] = gld(&ss->pbar0[f]); pb1[f] = gld(&ss->pbar1[f]);
            mvp0[f] = gld(&ss->meanVp0[f]); mvp1[f] = gld(&ss->meanVp1[f]);
        }
    } else {
        const DevScalars *sc = A.sc;
        const double a0 = gld(&sc->a[0]);
        ia02 = 1. / (a0 * a0);
#pragma unroll
        for (int f = 0; f < NF; f++) {
            pb0[f] = gld(&sc->pbar[f]);
            vb0[f] = gld(&sc->vbar[f]);
            pb1[f] = pb0[f] + hdt * vb0[f];
        }
    }
    double acc[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) acc[q] = 0.;

    /* planes i0 - 1 .. i0 + 1 into slots 0 .. 2; plane i0 + 2 in flight */
    T rs[2][NA][2];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        load(rs[0], i0 - 1 + q);
        put(rs[0], q);
    }
    load(rs[0], i0 + 2);
    __syncthreads();

    /* plane i reads slots Q, Q + 1, Q + 2 (planes i - 1, i, i + 1, mod 4),
     * issues plane i + 3 into register set (Q + 1) & 1 and stores plane
     * i + 2, in flight since plane i - 1 in set Q & 1, into slot Q + 3 */
    auto plane = [&](auto role, const int i) {
        constexpr int Q = decltype(role)::value;
        constexpr int SM = Q % 4, SC = (Q + 1) % 4, SP = (Q + 2) % 4, SN = (Q + 3) % 4;
        constexpr int RL = (Q + 1) & 1, RW = Q & 1;
        load(rs[RL], i + 3);
        T lap0[NF], lap1[NF], fcs[NF], vcs[NF], c1s[NF];
        double full0[2] = {0., 0.}, full1[2] = {0., 0.};
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const T *rf = ring[SC][f], *rv = ring[SC][NF + f];
            const T fm = ring[SM][f][lc], fc = rf[lc], fp = ring[SP][f][lc];
            const T vm = ring[SM][NF + f][lc], vc = rv[lc], vp = ring[SP][NF + f][lc];
            const T fkp = rf[lc + 1], fkm = rf[lc - 1];
            const T fjp = rf[lc + W], fjm = rf[lc - W];
            lap0[f] = (((fm + fp) + (fjm + fjp)) + (fkm + fkp) - (T)6 * fc) * idx2;
            full0[f] = pb0[f] + (double)fc;
            const T vkp = rv[lc + 1], vkm = rv[lc - 1];
            const T vjp = rv[lc + W], vjm = rv[lc - W];
            const T m1 = midpt(fm, vm, hdtT), p1 = midpt(fp, vp, hdtT);
            const T jm1 = midpt(fjm, vjm, hdtT), jp1 = midpt(fjp, vjp, hdtT);
            const T km1 = midpt(fkm, vkm, hdtT), kp1 = midpt(fkp, vkp, hdtT);
            const T c1 = midpt(fc, vc, hdtT);
            lap1[f] = (((m1 + p1) + (jm1 + jp1)) + (km1 + kp1) - (T)6 * c1) * idx2;
            full1[f] = pb1[f] + (double)c1;
            fcs[f] = fc;
            vcs[f] = vc;
            c1s[f] = c1;
        }
        double V0, V1, dV0[2], dV1[2];
        gabe_V_dV(A.model, full0[0], full0[1], V0, dV0[0], dV0[1]);
        gabe_V_dV(A.model, full1[0], full1[1], V1, dV1[0], dV1[1]);
        if constexpr (SB) {
#pragma unroll
            for (int f = 0; f < NF; f++) {
                const double Vp0 = dV0[f];
                const double Vp1 = dV1[f];
                const double dv0 = (double)vcs[f];
                /* F(0) and F(1) of ddfield (g2functions.cpp:193-196) minus the
                 * background force */
                const double dF0 = (double)lap0[f] * ia02 - (Vp0 - mvp0[f]) - 3. * H0 * dv0;
                const double dv1 = dv0 + hdt * dF0;
                const double dF1 = (double)lap1[f] * ia12 - (Vp1 - mvp1[f]) - 3. * H1 * dv1;
                gst(reinterpret_cast<T *>(reinterpret_cast<char *>(A.fo[f] + (long)i * N2) + ocen),
                    (T)((double)fcs[f] + dt * dv1));
                gst(reinterpret_cast<T *>(reinterpret_cast<char *>(A.vo[f] + (long)i * N2) + ocen),
                    (T)(dv0 + dt * dF1));
            }
        } else {
            acc[0] += V0;
            acc[1] += V1;
#pragma unroll
            for (int f = 0; f < NF; f++) {
                const double vel0 = vb0[f] + (double)vcs[f];
                const int b = 2 + 7 * f;
                const double Vp0 = dV0[f];
                const double Vp1 = dV1[f];
                const double X = (double)lap0[f] * ia02 - Vp0;
                acc[b + 0] += vel0 * vel0;
                acc[b + 3] += (double)fcs[f] * (double)lap0[f];
                acc[b + 1] += vel0 * X;
                acc[b + 2] += X * X;
                acc[b + 4] += (double)c1s[f] * (double)lap1[f];
                acc[b + 5] += Vp0;
                acc[b + 6] += Vp1;
            }
        }
        put(rs[RW], SN);
        __syncthreads();
    };
    int i = i0;
    for (; i + 3 < i1; i += 4) {
        plane(plane_role<0>{}, i);
        plane(plane_role<1>{}, i + 1);
        plane(plane_role<2>{}, i + 2);
        plane(plane_role<3>{}, i + 3);
    }
    if (i < i1) plane(plane_role<0>{}, i);
    if (i + 1 < i1) plane(plane_role<1>{}, i + 1);
    if (i + 2 < i1) plane(plane_role<2>{}, i + 2);
    if constexpr (!SB) block_store_partials<NS, 1>(acc, A.partials, A.nblk, bid);
}

template <typename T, int NF, int MODE, bool GW, int ROWN = 0>
__device__ __forceinline__ void sweep_body(const SweepArgs<T> &A, int hw, int gx, int gy, int nhw,
                                           typename gabe_fft::cplx_t<T>::type *zl = nullptr,
                                           typename gabe_fft::cplx_t<T>::type *zo = nullptr,
                                           typename gabe_fft::cplx_t<T>::type *zt = nullptr)
{
    if constexpr (MODE == SW_B_TILE) {
        sweep_tile<T, NF, MODE>(A, hw, gx, gy, nhw);
        return;
    }
    if constexpr (MODE == SW_B && sizeof(T) == 4) {
        sweep_b_ahead<T, NF>(A, hw, gx, gy, nhw);
        return;
    }
    constexpr int NS = gabe_nsums(NF);
    const int N = A.N;
    const long N2 = A.N2;
    int bx, by;
    if (!xcd_block(hw, gx, gy, nhw, bx, by)) return;
    by += A.by0;
    constexpr int BLK = ROWN > 0 ? row_block(ROWN) : GABE_BLOCK;
    const long p = (long)bx * BLK + threadIdx.x;
    const bool valid = p < N2;
    const long pp = valid ? p : 0;
    const int j = (int)(pp / N), k = (int)(pp - (long)j * N);
    const long pkp = (long)j * N + (k + 1 == N ? 0 : k + 1);
    const long pkm = (long)j * N + (k == 0 ? N - 1 : k - 1);
    const long pjp = (long)(j + 1 == N ? 0 : j + 1) * N + k;
    const long pjm = (long)(j == 0 ? N - 1 : j - 1) * N + k;
    const int i0 = by * A.ci;
    const int i1 = min(i0 + A.ci, N);
    const int bid = by * gx + bx;

    const T idx2 = (T)A.inv_dx2;
    const double dt = A.dt, hdt = 0.5 * dt;
    const T hdtT = (T)hdt;

    /* scalars: sweep A / energy read the current state, sweep B the frozen step values */
    double pb0[2] = {0., 0.}, pb1[2] = {0., 0.}, vb0[2] = {0., 0.}, mvp0[2] = {0., 0.}, mvp1[2] = {0., 0.};
    double ia02, ia12 = 0., H0 = 0., H1 = 0.;
    if (MODE == SW_B) {
        const StepScalars *ss = A.ss;
        ia02 = 1. / (ss->a0 * ss->a0);
        ia12 = 1. / (ss->a1 * ss->a1);
        H0 = ss->adot0 / ss->a0;
        H1 = ss->adot1 / ss->a1;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            pb0[f] = ss->pbar0[f]; pb1[f] = ss->pbar1[f];
            mvp0[f] = ss->meanVp0[f]; mvp1[f] = ss->meanVp1[f];
        }
    } else {
        const DevScalars *sc = A.sc;
        ia02 = 1. / (sc->a[0] * sc->a[0]);
#pragma unroll
        for (int f = 0; f < NF; f++) {
            pb0[f] = sc->pbar[f];
            vb0[f] = sc->vbar[f];
            pb1[f] = sc->pbar[f] + hdt * sc->vbar[f];
        }
    }

    double acc[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) acc[s] = 0.;
    if constexpr (ROWN > 0) {
        /* the first half of the row transform's twiddle table, read from LDS
         * by every call; the barrier before the first call publishes it */
        using C = typename gabe_fft::cplx_t<T>::type;
        for (int e = threadIdx.x; e < ROWN / 2; e += BLK) zt[e] = reinterpret_cast<const C *>(A.tw)[e];
    }

    T fm[NF], fc[NF], vm[NF], vc[NF];
    const int im0 = i0 == 0 ? N - 1 : i0 - 1;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        fm[f] = A.f[f][(long)im0 * N2 + pp];
        fc[f] = A.f[f][(long)i0 * N2 + pp];
        vm[f] = A.v[f][(long)im0 * N2 + pp];
        vc[f] = A.v[f][(long)i0 * N2 + pp];
    }

    for (int i = i0; i < i1; i++) {
        const int ip = i + 1 == N ? 0 : i + 1;
        const long base = (long)i * N2;
        T fp[NF], vp[NF];
        T lap0[NF], lap1[NF], c1[NF];
        T g0[NF][3], g1[NF][3];
        double full0[2] = {0., 0.}, full1[2] = {0., 0.};
#pragma unroll
        for (int f = 0; f < NF; f++) {
            fp[f] = A.f[f][(long)ip * N2 + pp];
            vp[f] = A.v[f][(long)ip * N2 + pp];
            const T fkp = A.f[f][base + pkp], fkm = A.f[f][base + pkm];
            const T fjp = A.f[f][base + pjp], fjm = A.f[f][base + pjm];
            lap0[f] = (((fm[f] + fp[f]) + (fjm + fjp)) + (fkm + fkp) - (T)6 * fc[f]) * idx2;
            full0[f] = pb0[f] + (double)fc[f];
            if (MODE != SW_ENERGY) {
                const T vkp = A.v[f][base + pkp], vkm = A.v[f][base + pkm];
                const T vjp = A.v[f][base + pjp], vjm = A.v[f][base + pjm];
                const T m1 = midpt(fm[f], vm[f], hdtT), p1 = midpt(fp[f], vp[f], hdtT);
                const T jm1 = midpt(fjm, vjm, hdtT), jp1 = midpt(fjp, vjp, hdtT);
                const T km1 = midpt(fkm, vkm, hdtT), kp1 = midpt(fkp, vkp, hdtT);
                c1[f] = midpt(fc[f], vc[f], hdtT);
                lap1[f] = (((m1 + p1) + (jm1 + jp1)) + (km1 + kp1) - (T)6 * c1[f]) * idx2;
                full1[f] = pb1[f] + (double)c1[f];
                if (GW) {
                    g0[f][0] = (fp[f] - fm[f]) * (T)0.5;
                    g0[f][1] = (fjp - fjm) * (T)0.5;
                    g0[f][2] = (fkp - fkm) * (T)0.5;
                    g1[f][0] = (p1 - m1) * (T)0.5;
                    g1[f][1] = (jp1 - jm1) * (T)0.5;
                    g1[f][2] = (kp1 - km1) * (T)0.5;
                }
            }
        }

        if (valid) {
            /* V and V' at the step-start and midpoint fields (one evaluation
             * each gives V and both derivatives) */
            double V0, V1 = 0., dV0[2], dV1[2] = {0., 0.};
            gabe_V_dV(A.model, full0[0], full0[1], V0, dV0[0], dV0[1]);
            if (MODE != SW_ENERGY) gabe_V_dV(A.model, full1[0], full1[1], V1, dV1[0], dV1[1]);
            if (MODE == SW_B) {
#pragma unroll
                for (int f = 0; f < NF; f++) {
                    const double Vp0 = dV0[f];
                    const double Vp1 = dV1[f];
                    const double dv0 = (double)vc[f];
                    /* F(0) and F(1) of ddfield (g2functions.cpp:193-196) minus the
                     * background force (zero in the FP64 form) */
                    const double dF0 = (double)lap0[f] * ia02 - (Vp0 - mvp0[f]) - 3. * H0 * dv0;
                    const double dv1 = dv0 + hdt * dF0;
                    const double dF1 = (double)lap1[f] * ia12 - (Vp1 - mvp1[f]) - 3. * H1 * dv1;
                    A.fo[f][base + pp] = (T)((double)fc[f] + dt * dv1);
                    A.vo[f][base + pp] = (T)(dv0 + dt * dF1);
                }
            } else {
                acc[0] += V0;
                if (MODE == SW_A) acc[1] += V1;
#pragma unroll
                for (int f = 0; f < NF; f++) {
                    const double vel0 = vb0[f] + (double)vc[f];
                    const int b = 2 + 7 * f;
                    acc[b + 0] += vel0 * vel0;
                    acc[b + 3] += (double)fc[f] * (double)lap0[f];
                    if (MODE == SW_A) {
                        const double Vp0 = dV0[f];
                        const double Vp1 = dV1[f];
                        const double X = (double)lap0[f] * ia02 - Vp0;
                        acc[b + 1] += vel0 * X;
                        acc[b + 2] += X * X;
                        acc[b + 4] += (double)c1[f] * (double)lap1[f];
                        acc[b + 5] += Vp0;
                        acc[b + 6] += Vp1;
                    }
                }
                if (MODE == SW_A && GW) {
                    /* gradient stress of both stages (g2evolutionpert.cpp:48-132),
                     * S_ab = sum_f D_a f D_b f / dx^2 up to a multiple of delta_ab,
                     * as the GABE_SC components (xx - zz, xy, xz, yy - zz, yz): the
                     * TT projection removes the rest (the reference's trace part
                     * and the zz component alike) */
                    T sv[GABE_NSC];
#pragma unroll
                    for (int st = 0; st < 2; st++) {
                        T sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
#pragma unroll
                        for (int f = 0; f < NF; f++) {
                            const T *g = st == 0 ? g0[f] : g1[f];
                            sxx += g[0] * g[0]; syy += g[1] * g[1]; szz += g[2] * g[2];
                            sxy += g[0] * g[1]; sxz += g[0] * g[2]; syz += g[1] * g[2];
                        }
                        sv[GABE_SC * st + 0] = (sxx - szz) * idx2;
                        sv[GABE_SC * st + 1] = sxy * idx2;
                        sv[GABE_SC * st + 2] = sxz * idx2;
                        sv[GABE_SC * st + 3] = (syy - szz) * idx2;
                        sv[GABE_SC * st + 4] = syz * idx2;
                    }
                    if constexpr (ROWN > 0) {
                        /* complex sequence q of a plane carries components 2q (real
                         * part) and 2q + 1 (imaginary part); sequence (plane, q, row t / N)
                         * at ((plane mod row_planes) GABE_NSC/2 + q) RR + t / N */
                        using C = typename gabe_fft::cplx_t<T>::type;
                        constexpr int RR = BLK / ROWN, SS = rseq_stride(ROWN), PPC = row_planes<T>(ROWN);
                        const int t = threadIdx.x, r = t / ROWN, kk = t - r * ROWN;
                        C *z = zl + ((((i - i0) % PPC) * (GABE_NSC / 2)) * RR + r) * SS + rpad(kk);
#pragma unroll
                        for (int q = 0; q < GABE_NSC / 2; q++) {
                            C zq;
                            zq.x = sv[2 * q];
                            zq.y = sv[2 * q + 1];
                            z[q * RR * SS] = zq;
                        }
                    } else {
                        T *S = A.stress + base + pp;
#pragma unroll
                        for (int c = 0; c < GABE_NSC; c++) S[(long)c * A.N3] = sv[c];
                    }
                }
            }
        }
        if constexpr (ROWN > 0) {
            constexpr int PPC = row_planes<T>(ROWN);
            if ((i - i0) % PPC == PPC - 1 || i + 1 == i1) {
                const int npl = (i - i0) % PPC + 1;
                __syncthreads();
                /* no barrier after the call: the next rows written to zl
                 * follow its last reads of zl (before its final barrier), and
                 * the next writes to zo come after the next call's first two
                 * barriers */
                row_r2c<T, ROWN>(zl, zo, zt, A, i + 1 - npl, bx * (BLK / ROWN), npl);
            }
        }
#pragma unroll
        for (int f = 0; f < NF; f++) {
            fm[f] = fc[f]; fc[f] = fp[f];
            vm[f] = vc[f]; vc[f] = vp[f];
        }
    }
    if constexpr (ROWN > 0) {
        /* zl is free: the last transform's reads of it came before its final
         * barrier; using it keeps three workgroups' LDS within one CU */
        static_assert(sizeof(double) * (BLK / 64) * NS <= sizeof(*zl) * rbuf_size<T>(ROWN), "reduction exceeds zl");
        reduce_partials<NS, BLK / GABE_BLOCK>(acc, reinterpret_cast<double (*)[NS]>(zl), A.partials, A.nblk, bid);
    } else if (MODE != SW_B) {
        block_store_partials<NS, BLK / GABE_BLOCK>(acc, A.partials, A.nblk, bid);
    }
}

/* sweep A with the stress rows transformed in the workgroup (N 16..512) */
template <typename T, int NF, int ROWN>
__global__ __launch_bounds__(row_block(ROWN)) void k_sweep_rowfft(SweepArgs<T> A)
{
    using C = typename gabe_fft::cplx_t<T>::type;
    /* zl, zo and the half twiddle table: 53,248 bytes at 256 sites in FP32,
     * so that three workgroups fit in the LDS of one CU (a whole table, 2 KB
     * more, measured at two) */
    __shared__ C zl[rbuf_size<T>(ROWN)], zo[rbuf_size<T>(ROWN)], zt[ROWN / 2];
    sweep_body<T, NF, SW_A, true, ROWN>(A, blockIdx.y * gridDim.x + blockIdx.x, gridDim.x, gridDim.y,
                                        gridDim.x * gridDim.y, zl, zo, zt);
}

template <typename T, int NF, int MODE, bool GW>
__global__ __launch_bounds__(GABE_BLOCK) void k_sweep(SweepArgs<T> A)
{
    sweep_body<T, NF, MODE, GW>(A, blockIdx.y * gridDim.x + blockIdx.x, gridDim.x, gridDim.y, gridDim.x * gridDim.y);
}

/* every lattice of an ensemble in one launch: workgroups [start[L], start[L+1])
 * run lattice L with the block decomposition of its own single launch (each
 * start is a multiple of the XCD count, so the XCD-aware order is unchanged) */
/* the lattice of workgroup b of a batched launch */
__device__ __forceinline__ int batch_member(const int *start, int count, int b)
{
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T, int NF, int MODE, bool GW>
__global__ __launch_bounds__(GABE_BLOCK) void k_sweep_batch(const SweepArgs<T> *__restrict__ descs, const int *start, int count)
{
    const int b = blockIdx.x, lo = batch_member(start, count, b);
    /* the table is read-only for the whole launch (__restrict__): without
     * that the compiler reloads the descriptor's fields after every store of
     * the march (batched sweep B at 64 lattices: 637 -> 490 us) */
    const SweepArgs<T> &A = descs[lo];
    /* the lattice's range is padded to a multiple of 8 workgroups: the block
     * order spreads the padding over the XCDs' ranges */
    sweep_body<T, NF, MODE, GW>(A, b - start[lo], A.gx, A.gy, start[lo + 1] - start[lo]);
}

/* The tiled sweep A held to 128 VGPRs, four waves per SIMD (without the
 * bound it takes 143, three waves; no spills either way): the 64-lattice
 * ensemble's batched sweep A 0.326 against 0.340-0.349 ms, its step 0.671
 * against 0.688 ms. */
template <typename T, int NF>
__global__ __launch_bounds__(GABE_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_sweep_a_tile(SweepArgs<T> A)
{
    sweep_tile<T, NF, SW_A_TILE>(A, blockIdx.y * gridDim.x + blockIdx.x, gridDim.x, gridDim.y, gridDim.x * gridDim.y);
}

template <typename T, int NF>
__global__ __launch_bounds__(GABE_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void
k_sweep_a_tile_batch(const SweepArgs<T> *__restrict__ descs, const int *start, int count)
{
    const int b = blockIdx.x, lo = batch_member(start, count, b);
    const SweepArgs<T> &A = descs[lo];
    sweep_tile<T, NF, SW_A_TILE>(A, b - start[lo], A.gx, A.gy, start[lo + 1] - start[lo]);
}

/* ---------------------------------------------------------------- finalize */
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

/* One workgroup: compensated fixed-order sum of the partials, then the
 * expansion update of g2functions.cpp:250-252,268-270 and the background
 * update of the FP32 form.  Deterministic: the same partials give the same
 * bits on every run. */
template <int NF>
__device__ __forceinline__ void finalize_body(const FinArgs &F)
{
    constexpr int NS = gabe_nsums(NF);
    /* the state scalars the tail (thread 0) needs, read before the reduction
     * so that their latency hides behind it */
    DevScalars *sc = F.sc;
    double a0 = 0., adot_prev = 0., t_old = 0., pbv[NF], vbv[NF];
    long long nstep = 0;
    int adot0_valid = 0;
#pragma unroll
    for (int f = 0; f < NF; f++) { pbv[f] = 0.; vbv[f] = 0.; }
    if (threadIdx.x == 0) {
        a0 = sc->a[0];
        adot_prev = sc->adot[0];
        adot0_valid = sc->adot0_valid;
        t_old = sc->t;
        nstep = sc->nstep;
#pragma unroll
        for (int f = 0; f < NF; f++) { pbv[f] = sc->pbar[f]; vbv[f] = sc->vbar[f]; }
    }

    /* each thread adds partials b, b + GABE_BLOCK, ... in that order; two of
     * them are loaded at a time so that one wait covers 2 NS loads.  Four at a
     * time took 250 VGPRs per wave: beside the mode kernel (168 per wave, one
     * wave per SIMD per workgroup) the finalize then waited for two of its
     * workgroups to leave one CU; at 164 one is enough */
    double s[NS], c[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) { s[q] = 0.; c[q] = 0.; }
    const int nb = F.nblk;
    int b = threadIdx.x;
    for (; b + GABE_BLOCK < nb; b += 2 * GABE_BLOCK) {
        double v[2][NS];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int q = 0; q < NS; q++) v[u][q] = F.partials[(long)q * nb + b + u * GABE_BLOCK];
        __builtin_amdgcn_sched_barrier(0);   /* all loads issued before the first add */
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int q = 0; q < NS; q++) {
                double t, e;
                two_sum(s[q], v[u][q], t, e);
                s[q] = t;
                c[q] += e;
            }
    }
    for (; b < nb; b += GABE_BLOCK) {
#pragma unroll
        for (int q = 0; q < NS; q++) {
            double t, e;
            two_sum(s[q], F.partials[(long)q * nb + b], t, e);
            s[q] = t;
            c[q] += e;
        }
    }
    /* butterfly across the wavefront, then the four wavefronts in order */
    __shared__ double sh_s[GABE_BLOCK / 64][NS], sh_c[GABE_BLOCK / 64][NS];
    __shared__ double tot[GABE_MAXSUMS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NS; q++) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double so = __shfl_xor(s[q], off), co = __shfl_xor(c[q], off);
            double t, e;
            two_sum(s[q], so, t, e);
            s[q] = t;
            c[q] = c[q] + co + e;
        }
        if (lane == 0) { sh_s[wv][q] = s[q]; sh_c[wv][q] = c[q]; }
    }
    __syncthreads();
    if (threadIdx.x < NS) {
        const int q = threadIdx.x;
        double ss = sh_s[0][q], cc = sh_c[0][q];
        for (int w = 1; w < GABE_BLOCK / 64; w++) {
            double t, e;
            two_sum(ss, sh_s[w][q], t, e);
            ss = t;
            cc = cc + sh_c[w][q] + e;
        }
        tot[q] = ss + cc;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;

    const double N3 = F.N3, dt = F.dt;
    double K0 = 0., G0 = 0.;
    for (int f = 0; f < F.nf; f++) {
        K0 += tot[2 + 7 * f + 0];
        G0 += tot[2 + 7 * f + 3];
    }
    const double edkin0 = K0 / N3 / 2.;
    const double edpot0 = tot[0] / N3;
    const double edgrad0 = -G0 / N3 / 2. / a0 / a0;
    const double edrho0 = edkin0 + edpot0 + edgrad0;

    if (F.mode == SW_ENERGY) {
        sc->edkin[0] = edkin0; sc->edpot[0] = edpot0; sc->edgrad[0] = edgrad0; sc->edrho[0] = edrho0;
        sc->e0_valid = 1;
        if (F.refresh_adot && !adot0_valid) {
            sc->adot[0] = F.exp_type == 0 ? 0. : sqrt(8. * M_PI / 3. * edrho0) * a0;
            sc->adot0_valid = 1;
        }
        if (!isfinite(edrho0)) sc->error = 1;
        return;
    }

    /* adot0 = adf(0) of the step-start state (:270 of the previous step), or
     * the value set by the host for the first step (initexpansion, Q4) */
    double adot0;
    if (F.exp_type == 0) adot0 = 0.;
    else if (adot0_valid) adot0 = adot_prev;
    else adot0 = sqrt(8. * M_PI / 3. * edrho0) * a0;
    const double H0 = adot0 / a0;

    /* midpoint kinetic energy from the expansion of
     * fdot1 = fdot0 + dt/2 (X - 3 H0 fdot0) = cf fdot0 + dt/2 X */
    const double cf = 1. - 1.5 * dt * H0;
    double K1 = 0., G1 = 0.;
    for (int f = 0; f < F.nf; f++) {
        const int b = 2 + 7 * f;
        K1 += cf * cf * tot[b] + cf * dt * tot[b + 1] + 0.25 * dt * dt * tot[b + 2];
        G1 += tot[b + 4];
    }
    const double a1 = F.exp_type == 0 ? a0 : a0 + .5 * dt * adot0;           /* :250 */
    const double edkin1 = K1 / N3 / 2.;
    const double edpot1 = tot[1] / N3;
    const double edgrad1 = -G1 / N3 / 2. / a1 / a1;
    const double edrho1 = edkin1 + edpot1 + edgrad1;                          /* :251 */
    const double adot1 = F.exp_type == 0 ? 0. : sqrt(8. * M_PI / 3. * edrho1) * a1; /* :252 */
    const double H1 = adot1 / a1;
    const double a0n = F.exp_type == 0 ? a0 : a0 + dt * adot1;                /* :268 */

    StepScalars *ss = F.ss;
    ss->a0 = a0; ss->adot0 = adot0; ss->a1 = a1; ss->adot1 = adot1;
    for (int f = 0; f < NF; f++) {   /* NF = F.nf (the launch selects it) */
        const int b = 2 + 7 * f;
        const double mv0 = F.split ? tot[b + 5] / N3 : 0.;
        const double mv1 = F.split ? tot[b + 6] / N3 : 0.;
        const double pb0 = pbv[f], vb0 = vbv[f];
        const double pb1 = pb0 + .5 * dt * vb0;
        const double vb1 = vb0 + .5 * dt * (-mv0 - 3. * H0 * vb0);
        ss->pbar0[f] = pb0; ss->vbar0[f] = vb0; ss->pbar1[f] = pb1; ss->vbar1[f] = vb1;
        ss->meanVp0[f] = mv0; ss->meanVp1[f] = mv1;
        sc->pbar[f] = pb0 + dt * vb1;
        sc->vbar[f] = vb0 + dt * (-mv1 - 3. * H1 * vb1);
    }
    if (F.hist && F.hist_cap > 0) {
        double *row = F.hist + (nstep % F.hist_cap) * GABE_HIST_COLS;
        row[0] = t_old; row[1] = a0; row[2] = adot0; row[3] = edkin0;
        row[4] = edpot0; row[5] = edgrad0; row[6] = edrho0; row[7] = (double)nstep;
    }
    sc->edkin[0] = edkin0; sc->edpot[0] = edpot0; sc->edgrad[0] = edgrad0; sc->edrho[0] = edrho0;
    sc->edkin[1] = edkin1; sc->edpot[1] = edpot1; sc->edgrad[1] = edgrad1; sc->edrho[1] = edrho1;
    sc->a[0] = a0n; sc->a[1] = a1;
    sc->adot[0] = adot0; sc->adot[1] = adot1;
    sc->t = t_old + dt;
    sc->nstep = nstep + 1;
    sc->adot0_valid = 0;
    sc->e0_valid = 0;
    if (!isfinite(edrho0) || !isfinite(edrho1)) sc->error = 1;
}

template <int NF>
__global__ __launch_bounds__(GABE_BLOCK) void k_finalize(FinArgs F)
{
    finalize_body<NF>(F);
}

/* one workgroup per lattice of an ensemble */
template <int NF>
__global__ __launch_bounds__(GABE_BLOCK) void k_finalize_batch(const FinArgs *descs)
{
    finalize_body<NF>(descs[blockIdx.x]);
}

/* ------------------------------------------------------------- GW modes */
template <typename C>
__device__ __forceinline__ C cmul(C a, double s) { a.x *= s; a.y *= s; return a; }
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { a.x += b.x; a.y += b.y; return a; }

/* T = Lambda(n) S dx^3 with Lambda_ij,lm = P_il P_jm - P_ij P_lm / 2 and
 * P = 1 - n n^T, evaluated as T = S - n u^T - u n^T + w n n^T - P (tr S - w)/2
 * with u = S n, w = n.S.n.  Same operator as g2evolutionpert.cpp:215-310.
 * Components ordered 11,12,13,22,23,33. */
template <typename C, typename R>
__device__ __forceinline__ void tt_project(const C S[6], R nx, R ny, R nz, R dx3, C T[6])
{
    C u[3];
    u[0].x = S[0].x * nx + S[1].x * ny + S[2].x * nz;
    u[0].y = S[0].y * nx + S[1].y * ny + S[2].y * nz;
    u[1].x = S[1].x * nx + S[3].x * ny + S[4].x * nz;
    u[1].y = S[1].y * nx + S[3].y * ny + S[4].y * nz;
    u[2].x = S[2].x * nx + S[4].x * ny + S[5].x * nz;
    u[2].y = S[2].y * nx + S[4].y * ny + S[5].y * nz;
    C w;
    w.x = u[0].x * nx + u[1].x * ny + u[2].x * nz;
    w.y = u[0].y * nx + u[1].y * ny + u[2].y * nz;
    C hq;   /* (tr S - w) / 2 */
    hq.x = (S[0].x + S[3].x + S[5].x - w.x) * (R)0.5;
    hq.y = (S[0].y + S[3].y + S[5].y - w.y) * (R)0.5;
    const R n[3] = {nx, ny, nz};
    const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const R ni = n[ia[c]], nj = n[ib[c]];
        const R pij = (ia[c] == ib[c] ? (R)1 : (R)0) - ni * nj;
        T[c].x = (S[c].x - ni * u[ib[c]].x - u[ia[c]].x * nj + w.x * ni * nj - pij * hq.x) * dx3;
        T[c].y = (S[c].y - ni * u[ib[c]].y - u[ia[c]].y * nj + w.y * ni * nj - pij * hq.y) * dx3;
    }
}

/* One thread per mode: the TT source of both stages projected on the two
 * polarisations, then the two midpoint-RK2 stages of evolve_perts
 * (g2evolutionpert.cpp:343-455, derivs :459-516) in registers.  h and l are
 * stored per mode as (h+, hx) complex pairs: a TT tensor has two degrees of
 * freedom, and with a TT source and TT (e.g. zero) start every component of
 * the reference's h_ij stays in that span. */
template <typename T>
__global__ __launch_bounds__(GABE_BLOCK) void k_gw(GwArgs<T> G)
{
    using C = typename cplx_of<T>::type;
    const long m = (long)blockIdx.x * GABE_BLOCK + threadIdx.x;
    if (m >= G.M) return;
    const int N = G.N;
    int px, py, pz;
    long r;
    mode_of(m, N, px, py, pz, r);
    const int p2 = px * px + py * py + pz * pz;

    const C *Sh = reinterpret_cast<const C *>(G.shat);
    C *hh = reinterpret_cast<C *>(G.h);
    C *ll = reinterpret_cast<C *>(G.l);
    const long SM = (long)N * N * G.kpad, ms = r * G.kpad + pz;
    C S0[GABE_SC], S1[GABE_SC];
#pragma unroll
    for (int c = 0; c < GABE_SC; c++) {
        S0[c] = Sh[(long)c * SM + ms];
        S1[c] = Sh[(long)(GABE_SC + c) * SM + ms];
    }
    C h0[2] = {hh[2 * m], hh[2 * m + 1]}, l0[2] = {ll[2 * m], ll[2 * m + 1]};
    C T0[2], T1[2];
#pragma unroll
    for (int q = 0; q < 2; q++) { T0[q].x = T0[q].y = 0; T1[q].x = T1[q].y = 0; }
    if (p2 != 0) {   /* the zero mode has no source (g2evolutionpert.cpp:319-321) */
        double e[2][6];
        pol_basis(px, py, pz, e[0], e[1]);
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int c = 0; c < GABE_SC; c++) {
                /* (e : S)/2 dx^3 with off-diagonal components counted twice; the
                 * zz component of the stored source is zero */
                const T w = (T)((c == 0 || c == 3 ? 0.5 : 1.0) * G.dx3 * e[q][c]);
                T0[q].x += w * S0[c].x; T0[q].y += w * S0[c].y;
                T1[q].x += w * S1[c].x; T1[q].y += w * S1[c].y;
            }
    }
    const StepScalars *ss = G.ss;
    const double k2n = 4. * M_PI * M_PI / G.L / G.L * (double)p2;
    const T kk0 = (T)(k2n / (ss->a0 * ss->a0)), kk1 = (T)(k2n / (ss->a1 * ss->a1));
    const T fr0 = (T)(3. * ss->adot0 / ss->a0), fr1 = (T)(3. * ss->adot1 / ss->a1);
    const T om0 = (T)(16. * M_PI / (ss->a0 * ss->a0)), om1 = (T)(16. * M_PI / (ss->a1 * ss->a1));
    const T hdt = (T)(0.5 * G.dt), dt = (T)G.dt;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        C h1, l1, hn, ln;
        h1.x = h0[q].x + hdt * l0[q].x;
        h1.y = h0[q].y + hdt * l0[q].y;
        l1.x = l0[q].x + hdt * (-kk0 * h0[q].x - fr0 * l0[q].x + om0 * T0[q].x);
        l1.y = l0[q].y + hdt * (-kk0 * h0[q].y - fr0 * l0[q].y + om0 * T0[q].y);
        hn.x = h0[q].x + dt * l1.x;
        hn.y = h0[q].y + dt * l1.y;
        ln.x = l0[q].x + dt * (-kk1 * h1.x - fr1 * l1.x + om1 * T1[q].x);
        ln.y = l0[q].y + dt * (-kk1 * h1.y - fr1 * l1.y + om1 * T1[q].y);
        hh[2 * m + q] = hn;
        ll[2 * m + q] = ln;
    }
}

/* reference planes [12][M] (11,12,13,22,23,33 x Re,Im) <-> polarisation pairs
 * [M][2] complex: h_pol = (e_pol : h)/2, h_ij = sum_pol h_pol e_pol_ij */
template <typename T>
__global__ __launch_bounds__(GABE_BLOCK) void k_gw_pack(const double *ref, T *pol, ModeLayout lay, long M)
{
    using C = typename cplx_of<T>::type;
    const long m = (long)blockIdx.x * GABE_BLOCK + threadIdx.x;
    if (m >= M) return;
    const int N = lay.N;
    int px, py, pz;
    long r;
    mode_of(m, N, px, py, pz, r);
    const long sl = mode_slot(lay, (int)(r / N), (int)(r % N), pz);
    C out[2];
    out[0].x = out[0].y = out[1].x = out[1].y = 0;
    if (px != 0 || py != 0 || pz != 0) {
        double e[2][6];
        pol_basis(px, py, pz, e[0], e[1]);
        for (int q = 0; q < 2; q++) {
            double re = 0., im = 0.;
            for (int c = 0; c < 6; c++) {
                const double w = (c == 0 || c == 3 || c == 5 ? 0.5 : 1.0) * e[q][c];
                re += w * ref[(long)(2 * c) * M + m];
                im += w * ref[(long)(2 * c + 1) * M + m];
            }
            out[q].x = (T)re;
            out[q].y = (T)im;
        }
    }
    reinterpret_cast<C *>(pol)[2 * sl] = out[0];
    reinterpret_cast<C *>(pol)[2 * sl + 1] = out[1];
}

template <typename T>
__global__ __launch_bounds__(GABE_BLOCK) void k_gw_unpack(const T *pol, double *ref, ModeLayout lay, long M)
{
    using C = typename cplx_of<T>::type;
    const long m = (long)blockIdx.x * GABE_BLOCK + threadIdx.x;
    if (m >= M) return;
    const int N = lay.N;
    int px, py, pz;
    long r;
    mode_of(m, N, px, py, pz, r);
    const long sl = mode_slot(lay, (int)(r / N), (int)(r % N), pz);
    double e[2][6] = {{0., 0., 0., 0., 0., 0.}, {0., 0., 0., 0., 0., 0.}};
    if (px != 0 || py != 0 || pz != 0) pol_basis(px, py, pz, e[0], e[1]);
    const C a = reinterpret_cast<const C *>(pol)[2 * sl], b = reinterpret_cast<const C *>(pol)[2 * sl + 1];
    for (int c = 0; c < 6; c++) {
        ref[(long)(2 * c) * M + m] = e[0][c] * (double)a.x + e[1][c] * (double)b.x;
        ref[(long)(2 * c + 1) * M + m] = e[0][c] * (double)a.y + e[1][c] * (double)b.y;
    }
}

/* T_gw[s] in the reference layout [12][M] (diagnostic, not on the step path)
 * from the stored source (xx - zz, xy, xz, yy - zz, yz), whose projection is
 * the reference's.  With nyq set, shat rows hold kz < N/2 and the kz = N/2
 * plane is nyq [GABE_NSC][N][N]. */
template <typename T>
__global__ __launch_bounds__(GABE_BLOCK) void k_tt_out(const T *shat, const T *nyq, int stage, int N, long M, long kpad,
                                                         double dx3, double *out)
{
    using C = typename cplx_of<T>::type;
    const long m = (long)blockIdx.x * GABE_BLOCK + threadIdx.x;
    if (m >= M) return;
    const int H = N / 2 + 1;
    const int kz = (int)(m % H);
    const long r = m / H;
    const int j = (int)(r % N), i = (int)(r / N);
    const int px = i <= N / 2 ? i : i - N, py = j <= N / 2 ? j : j - N, pz = kz;
    const int p2 = px * px + py * py + pz * pz;
    const long SM = (long)N * N * kpad;
    const C *Sh = reinterpret_cast<const C *>(shat) + (long)(GABE_SC * stage) * SM;
    const bool plane = nyq != nullptr && kz == N / 2;
    const long NN = (long)N * N;
    const C *Sn = reinterpret_cast<const C *>(nyq) + (long)(GABE_SC * stage) * NN;
    C S[6], Tt[6];
#pragma unroll
    for (int c = 0; c < GABE_SC; c++) S[c] = plane ? Sn[(long)c * NN + r] : Sh[(long)c * SM + r * kpad + kz];
    S[5].x = S[5].y = (T)0;   /* zz */
    if (p2 == 0) {
#pragma unroll
        for (int c = 0; c < 6; c++) Tt[c].x = Tt[c].y = 0;
    } else {
        const double inv = 1. / sqrt((double)p2);
        tt_project<C, T>(S, (T)(px * inv), (T)(py * inv), (T)(pz * inv), (T)dx3, Tt);
    }
#pragma unroll
    for (int c = 0; c < 6; c++) {
        out[(long)(2 * c) * M + m] = (double)Tt[c].x;
        out[(long)(2 * c + 1) * M + m] = (double)Tt[c].y;
    }
}

/* ------------------------------------------------------------- spectra */
__device__ __forceinline__ int isqrt_floor(long x)
{
    if (x <= 0) return 0;
    long r = (long)sqrt((double)x);
    while (r * r > x) r--;
    while ((r + 1) * (r + 1) <= x) r++;
    return (int)r;
}

/* Workgroup (b, c) takes bin b's modes with floor(|n|) = b (g2spectra.cpp:68-92,
 * 126-184) in px chunk c, enumerated column by column ((px,py) fixed, kz
 * contiguous); k_spectra_reduce then adds the chunks of each bin in chunk
 * order, so the result does not depend on scheduling.  Splitting the shells
 * spreads the large ones over many workgroups.  Weights: 2 for
 * 1 <= kz <= N/2-1, 1 for kz in {0, N/2}. */
template <typename T>
__global__ __launch_bounds__(GABE_BLOCK) void k_spectra(SpecArgs<T> P)
{
    using C = typename cplx_of<T>::type;
    const int b = blockIdx.x;
    const int N = P.N, H = N / 2 + 1, nf = P.nf;
    const long M = P.M;
    const C *F = reinterpret_cast<const C *>(P.fhat);
    const C *Lg = reinterpret_cast<const C *>(P.l);
    const int R = b + 1;
    const int lo = max(-R, -N / 2 + 1), hi = min(R, N / 2);
    const int W = hi - lo + 1;
    /* this workgroup's chunk of the shell: GABE_SPEC_PX values of px */
    const int px0 = lo + GABE_SPEC_PX * (int)blockIdx.y, npx = max(0, min(GABE_SPEC_PX, hi + 1 - px0));
    const long b2 = (long)b * b, r2max = (long)(b + 1) * (b + 1) - 1;
    double acc[GABE_MAXF + 2] = {0., 0., 0., 0.};
    for (long q = threadIdx.x; q < (long)npx * W; q += GABE_BLOCK) {
        const int px = px0 + (int)(q / W), py = lo + (int)(q % W);
        const long r2 = (long)px * px + (long)py * py;
        if (r2 > r2max) continue;
        int kzlo = 0;
        if (b2 - r2 > 0) {
            kzlo = isqrt_floor(b2 - r2);
            if ((long)kzlo * kzlo < b2 - r2) kzlo++;
        }
        const int kzhi = min(isqrt_floor(r2max - r2), N / 2);
        const int i = px >= 0 ? px : px + N, j = py >= 0 ? py : py + N;
        const long row = ((long)i * N + j) * H;
        for (int kz = kzlo; kz <= kzhi; kz++) {
            const double w = (kz == 0 || kz == N / 2) ? 1. : 2.;
            const long m = row + kz;
            for (int f = 0; f < nf; f++) {
                const C v = F[(long)f * M + m];
                double re = (double)v.x, im = (double)v.y;
                if (m == 0) re += P.dc[f];
                acc[f] += w * (re * re + im * im);
            }
            acc[nf] += w;
            if (Lg) {
                /* sum_ij |l_ij|^2 (off-diagonal components twice, g2spectra.cpp:133-144)
                 * = 2 (|l+|^2 + |lx|^2) for the polarisation pair */
                const long ms = mode_slot(P.lay, i, j, (P.quirk && kz == 0) ? N / 2 : kz);
                const C a = Lg[2 * ms], bb = Lg[2 * ms + 1];
                const double e = 2. * ((double)a.x * (double)a.x + (double)a.y * (double)a.y +
                                       (double)bb.x * (double)bb.x + (double)bb.y * (double)bb.y);
                acc[nf + 1] += w * e;
            }
        }
    }
    __shared__ double red[GABE_BLOCK / 64][GABE_MAXF + 2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < nf + 2; q++) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) red[wv][q] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < nf + 2) {
        const int q = threadIdx.x;
        P.part[((long)b * P.nch + blockIdx.y) * (nf + 2) + q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
    }
}

/* per-shell sums over the chunks, in chunk order (bit-reproducible) */
__global__ __launch_bounds__(GABE_BLOCK) void k_spectra_reduce(const double *part, int numbins, int nch, int nq,
                                                                double *out)
{
    const int t = blockIdx.x * GABE_BLOCK + threadIdx.x;
    if (t >= numbins * nq) return;
    const int b = t / nq, q = t - b * nq;
    double s = 0.;
    for (int c = 0; c < nch; c++) s += part[((long)b * nch + c) * nq + q];
    out[(long)b * nq + q] = s;
}

/* --------------------------------------------------------------- helpers */
template <typename T>
__global__ __launch_bounds__(GABE_BLOCK) void k_effmass(const T *f0, const T *f1, int nf, int fld, double pbar0,
                                                          double pbar1, gabe_model m, long n3, double *partials)
{
    double acc[1] = {0.};
    for (long n = (long)blockIdx.x * GABE_BLOCK + threadIdx.x; n < n3; n += (long)gridDim.x * GABE_BLOCK) {
        const double phi = pbar0 + (double)f0[n];
        const double chi = nf > 1 ? pbar1 + (double)f1[n] : 0.;
        acc[0] += gabe_d2Vdf2(m, fld, phi, chi);
    }
    block_store_partials<1>(acc, partials, gridDim.x, blockIdx.x);
}

/* shifted one-pass mean/variance partials of one field: sums of d and d^2 with
 * d = f - f[0], so the variance does not cancel against the squared mean */
template <typename T>
__global__ __launch_bounds__(GABE_BLOCK) void k_meanvar(const T *f, long n3, double *partials)
{
    const double shift = (double)f[0];
    double acc[2] = {0., 0.};
    for (long n = (long)blockIdx.x * GABE_BLOCK + threadIdx.x; n < n3; n += (long)gridDim.x * GABE_BLOCK) {
        const double d = (double)f[n] - shift;
        acc[0] += d;
        acc[1] += d * d;
    }
    block_store_partials<2>(acc, partials, gridDim.x, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(GABE_BLOCK) void k_add_real(T *dst, const double *add, long n)
{
    const long i = (long)blockIdx.x * GABE_BLOCK + threadIdx.x;
    if (i < n) dst[i] = (T)((double)dst[i] + add[i]);
}

} // namespace

/* ---------------------------------------------------------------- launchers */
template <typename T, int NF>
static bool launch_rowfft(const SweepArgs<T> &a, dim3 grid, hipStream_t s)
{
    switch (a.rowfft) {
    case 16: hipLaunchKernelGGL((k_sweep_rowfft<T, NF, 16>), grid, dim3(GABE_BLOCK), 0, s, a); return true;
    case 32: hipLaunchKernelGGL((k_sweep_rowfft<T, NF, 32>), grid, dim3(GABE_BLOCK), 0, s, a); return true;
    case 64: hipLaunchKernelGGL((k_sweep_rowfft<T, NF, 64>), grid, dim3(GABE_BLOCK), 0, s, a); return true;
    case 128: hipLaunchKernelGGL((k_sweep_rowfft<T, NF, 128>), grid, dim3(GABE_BLOCK), 0, s, a); return true;
    case 256: hipLaunchKernelGGL((k_sweep_rowfft<T, NF, 256>), grid, dim3(GABE_BLOCK), 0, s, a); return true;
    case 512:
        /* one 512-thread workgroup per pair of the 256-thread row blocks */
        if constexpr (sizeof(T) == 4) {
            if (grid.x % 2) return false;
            hipLaunchKernelGGL((k_sweep_rowfft<T, NF, 512>), dim3(grid.x / 2, grid.y), dim3(512), 0, s, a);
            return true;
        }
        return false;
    default: return false;
    }
}

template <typename T, int NF>
static void launch_sweep_nf(int mode, bool gw, const SweepArgs<T> &a, dim3 grid, hipStream_t s)
{
    if (mode == SW_A && gw && a.rowfft && launch_rowfft<T, NF>(a, grid, s)) return;
    if (mode == SW_ENERGY) hipLaunchKernelGGL((k_sweep<T, NF, SW_ENERGY, false>), grid, dim3(GABE_BLOCK), 0, s, a);
    else if (mode == SW_A && gw) hipLaunchKernelGGL((k_sweep<T, NF, SW_A, true>), grid, dim3(GABE_BLOCK), 0, s, a);
    else if (mode == SW_A) hipLaunchKernelGGL((k_sweep<T, NF, SW_A, false>), grid, dim3(GABE_BLOCK), 0, s, a);
    else if (mode == SW_B_TILE)
        hipLaunchKernelGGL((k_sweep<T, NF, SW_B_TILE, false>), grid, dim3(GABE_BLOCK), 0, s, a);
    else if (mode == SW_A_TILE) {
        /* FP32 only: the callers pass it for FP32 lattices */
        if constexpr (sizeof(T) == 4) hipLaunchKernelGGL((k_sweep_a_tile<T, NF>), grid, dim3(GABE_BLOCK), 0, s, a);
    }
    else hipLaunchKernelGGL((k_sweep<T, NF, SW_B, false>), grid, dim3(GABE_BLOCK), 0, s, a);
}

template <typename T, int NF>
static void launch_sweep_batch_nf(int mode, const SweepArgs<T> *d, const int *start, int count, int nblk,
                                  hipStream_t s)
{
    if (mode == SW_ENERGY)
        hipLaunchKernelGGL((k_sweep_batch<T, NF, SW_ENERGY, false>), dim3(nblk), dim3(GABE_BLOCK), 0, s, d, start, count);
    else if (mode == SW_A)
        hipLaunchKernelGGL((k_sweep_batch<T, NF, SW_A, false>), dim3(nblk), dim3(GABE_BLOCK), 0, s, d, start, count);
    else if (mode == SW_B_TILE)
        hipLaunchKernelGGL((k_sweep_batch<T, NF, SW_B_TILE, false>), dim3(nblk), dim3(GABE_BLOCK), 0, s, d, start, count);
    else if (mode == SW_A_TILE) {
        if constexpr (sizeof(T) == 4)
            hipLaunchKernelGGL((k_sweep_a_tile_batch<T, NF>), dim3(nblk), dim3(GABE_BLOCK), 0, s, d, start, count);
    }
    else hipLaunchKernelGGL((k_sweep_batch<T, NF, SW_B, false>), dim3(nblk), dim3(GABE_BLOCK), 0, s, d, start, count);
}

template <typename T>
void launch_sweep_batch(int mode, int nf, const SweepArgs<T> *descs, const int *start, int count, int nblk,
                        hipStream_t s)
{
    if (nf == 1) launch_sweep_batch_nf<T, 1>(mode, descs, start, count, nblk, s);
    else launch_sweep_batch_nf<T, 2>(mode, descs, start, count, nblk, s);
}

void launch_finalize_batch(const FinArgs *descs, int nf, int count, hipStream_t s)
{
    if (nf == 1) hipLaunchKernelGGL(k_finalize_batch<1>, dim3(count), dim3(GABE_BLOCK), 0, s, descs);
    else hipLaunchKernelGGL(k_finalize_batch<2>, dim3(count), dim3(GABE_BLOCK), 0, s, descs);
}

template <typename T>
void launch_sweep(int mode, int nf, bool gw, const SweepArgs<T> &a, dim3 grid, hipStream_t s)
{
    if (nf == 1) launch_sweep_nf<T, 1>(mode, gw, a, grid, s);
    else launch_sweep_nf<T, 2>(mode, gw, a, grid, s);
}

void launch_finalize(const FinArgs &a, hipStream_t s)
{
    if (a.nf == 1) hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(GABE_BLOCK), 0, s, a);
    else hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(GABE_BLOCK), 0, s, a);
}

template <typename T>
void launch_gw(const GwArgs<T> &a, hipStream_t s)
{
    const unsigned nb = (unsigned)((a.M + GABE_BLOCK - 1) / GABE_BLOCK);
    hipLaunchKernelGGL((k_gw<T>), dim3(nb), dim3(GABE_BLOCK), 0, s, a);
}

template <typename T>
void launch_gw_pack(const double *ref, T *pol, ModeLayout lay, long M, hipStream_t s)
{
    const unsigned nb = (unsigned)((M + GABE_BLOCK - 1) / GABE_BLOCK);
    hipLaunchKernelGGL((k_gw_pack<T>), dim3(nb), dim3(GABE_BLOCK), 0, s, ref, pol, lay, M);
}

template <typename T>
void launch_gw_unpack(const T *pol, double *ref, ModeLayout lay, long M, hipStream_t s)
{
    const unsigned nb = (unsigned)((M + GABE_BLOCK - 1) / GABE_BLOCK);
    hipLaunchKernelGGL((k_gw_unpack<T>), dim3(nb), dim3(GABE_BLOCK), 0, s, pol, ref, lay, M);
}

template <typename T>
void launch_tt_out(const T *shat, const T *nyq, int stage, int N, long M, long kpad, double dx3, double *out, hipStream_t s)
{
    const unsigned nb = (unsigned)((M + GABE_BLOCK - 1) / GABE_BLOCK);
    hipLaunchKernelGGL((k_tt_out<T>), dim3(nb), dim3(GABE_BLOCK), 0, s, shat, nyq, stage, N, M, kpad, dx3, out);
}

template <typename T>
void launch_spectra(const SpecArgs<T> &a, hipStream_t s)
{
    hipLaunchKernelGGL((k_spectra<T>), dim3(a.numbins, a.nch), dim3(GABE_BLOCK), 0, s, a);
    const int nq = a.nf + 2;
    hipLaunchKernelGGL(k_spectra_reduce, dim3((a.numbins * nq + GABE_BLOCK - 1) / GABE_BLOCK), dim3(GABE_BLOCK), 0, s,
                       a.part, a.numbins, a.nch, nq, a.out);
}

template <typename T>
void launch_effmass(const T *f0, const T *f1, int nf, int fld, double pbar0, double pbar1, gabe_model m, long n3,
                    double *partials, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL((k_effmass<T>), dim3(nblk), dim3(GABE_BLOCK), 0, s, f0, f1, nf, fld, pbar0, pbar1, m, n3,
                       partials);
}

template <typename T>
void launch_meanvar(const T *f, long n3, double *partials, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL((k_meanvar<T>), dim3(nblk), dim3(GABE_BLOCK), 0, s, f, n3, partials);
}

template <typename T>
void launch_add_real(T *dst, const double *add, long n, hipStream_t s)
{
    const unsigned nb = (unsigned)((n + GABE_BLOCK - 1) / GABE_BLOCK);
    hipLaunchKernelGGL((k_add_real<T>), dim3(nb), dim3(GABE_BLOCK), 0, s, dst, add, n);
}

#define GABE_INST(T)                                                                                          \
    template void launch_sweep<T>(int, int, bool, const SweepArgs<T> &, dim3, hipStream_t);                   \
    template void launch_sweep_batch<T>(int, int, const SweepArgs<T> *, const int *, int, int, hipStream_t);  \
    template void launch_gw<T>(const GwArgs<T> &, hipStream_t);                                               \
    template void launch_gw_pack<T>(const double *, T *, ModeLayout, long, hipStream_t);                     \
    template void launch_gw_unpack<T>(const T *, double *, ModeLayout, long, hipStream_t);                   \
    template void launch_tt_out<T>(const T *, const T *, int, int, long, long, double, double *, hipStream_t); \
    template void launch_spectra<T>(const SpecArgs<T> &, hipStream_t);                                        \
    template void launch_effmass<T>(const T *, const T *, int, int, double, double, gabe_model, long,        \
                                    double *, int, hipStream_t);                                              \
    template void launch_meanvar<T>(const T *, long, double *, int, hipStream_t);                            \
    template void launch_add_real<T>(T *, const double *, long, hipStream_t);
GABE_INST(float)
GABE_INST(double)
