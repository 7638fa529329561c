/*
 * gabe_closure.hip -- learned closure of a coarse lattice and the spectral
 * restriction that produces its training data.  No counterpart in the
 * reference.
 *
 * Network (per site, periodic 3^3 stencils, channel-last activations):
 *   x  = (dphi, dchi, dphidot, dchidot) / s_in            4 channels, BF16
 *   a1 = SiLU(conv3(x;  W1[16][4][27])  + b1)            16 channels, BF16
 *   a2 = SiLU(conv3(a1; W2[16][16][27]) + b2)            16 channels, BF16
 *   y  =      conv3(a2; W3[2][16][27])  + b3              2 channels, FP32
 * conv3(u; W)[s][co] = sum_t sum_ci W[co][ci][t] u[s + off(t)][ci], tap
 * t = 9 (dz+1) + 3 (dy+1) + (dx+1), dz along i (slowest), dx along k.
 * After every RK2 step the velocity fluctuations get
 *   dfdot_c += dt s_out[c] y_c / a^2,
 * the 1/a^2 carrying the metric factor of the Laplacian the closure corrects.
 *
 * Every convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16: one
 * wave computes 16 consecutive sites (rows) x 16 output channels (columns);
 * K runs over (tap, input channel): two taps of 16 channels, or eight taps of
 * 4 channels, per instruction.  FP32 accumulation, BF16 operands.  The
 * backward pass reuses the same kernel on the gradients with the weights
 * flipped in space and transposed in channels; the weight gradient is a
 * second implicit GEMM with K over sites, fed from LDS copies of the
 * activations transposed to channel-major, reduced over workgroups in a fixed
 * order.  Training is Adam on FP32 master weights, repacked to BF16 after
 * every update.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "gabe_ctx.h"

using namespace gabe_detail;

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int CL_IN = 4, CL_H = 16, CL_OUT = 2, CL_TAPS = 27;
constexpr int P_W1 = 0;
constexpr int P_B1 = P_W1 + CL_H * CL_IN * CL_TAPS;
constexpr int P_W2 = P_B1 + CL_H;
constexpr int P_B2 = P_W2 + CL_H * CL_H * CL_TAPS;
constexpr int P_W3 = P_B2 + CL_H;
constexpr int P_B3 = P_W3 + CL_OUT * CL_H * CL_TAPS;
constexpr int P_TOTAL = P_B3 + CL_OUT;
static_assert(P_TOTAL == GABE_CLOSURE_NPARAMS, "parameter count of the C ABI");

/* The K dimension of one MFMA step covers (dz, dx) tap columns of the 3x3
 * stencil (9 columns) for one halo row; the three dy taps of a column are
 * three separate weight operands, so one activation read serves three
 * output rows. */
constexpr int TAP_COLS = 9;
constexpr int WIDE_STEPS = 5;    /* 9 columns x 16 channels, 2 columns per MFMA (the last half empty) */
constexpr int NARROW_STEPS = 2;  /* 9 columns x 4 channels, 8 columns per MFMA (the second step one column) */
/* The output layer has 2 channels, too few to fill 16 MFMA rows; there the
 * rows are (plane of a wave's pair, output channel, dy) and K runs over 12
 * (plane, dx) columns of 16 channels: the four input planes the wave's two
 * output planes read, so one MFMA step per halo row covers every tap. */
constexpr int OUT_STEPS = 6;
constexpr int WG_ROWS = 2;       /* lattice rows per weight-gradient slab */
constexpr int WG_SLABS = 8;      /* slabs per weight-gradient workgroup */

enum ConvEpi { EPI_HIDDEN_TRAIN, EPI_HIDDEN, EPI_BWD, EPI_OUT };

/* packed MFMA weight operands, [steps][3 dy][64 lanes] x 8 BF16 ([steps][64
 * lanes] for the output layer): the three forward layers, then the
 * backward-data convolutions of layers 3 and 2; biases are kept as [3][16]
 * FP32, zero padded */
constexpr size_t PK_F1 = 0, PK_F2 = PK_F1 + NARROW_STEPS * 3 * 64, PK_F3 = PK_F2 + WIDE_STEPS * 3 * 64,
                 PK_B3 = PK_F3 + OUT_STEPS * 64, PK_B2 = PK_B3 + NARROW_STEPS * 3 * 64,
                 PK_TOTAL = PK_B2 + WIDE_STEPS * 3 * 64;

/* Friedmann-constraint gate (on the device, no host round trip): from the
 * last three history rows, R = |rho_dot + 3H(rho + p)| / (3 H rho) at the
 * middle row with p = K - G/3 - V; the closure stops for good once R exceeds
 * the threshold or a value is not finite.  Records the largest R, and the
 * step of the fallback, for the host. */
struct GateArgs {
    const DevScalars *sc;
    const double *hist;
    int hist_cap;
    double thr;
    int *gate;               /* [0] active */
    long long *fallback;
    double *rmax;
};

__device__ void gate_eval(const GateArgs &G)
{
    if (G.gate[0] == 0) return;
    const long long n = G.sc->nstep - 1;
    if (n - 2 < G.sc->hist0) return;   /* three rows of this trajectory are needed */
    const double *rm = G.hist + ((n - 2) % G.hist_cap) * GABE_HIST_COLS;
    const double *r0 = G.hist + ((n - 1) % G.hist_cap) * GABE_HIST_COLS;
    const double *rp = G.hist + (n % G.hist_cap) * GABE_HIST_COLS;
    const double H = r0[2] / r0[1], rho = r0[6], p = r0[3] - r0[5] / 3. - r0[4];
    const double rhodot = (rp[6] - rm[6]) / (rp[0] - rm[0]);
    const double R = fabs(rhodot + 3. * H * (rho + p)) / (3. * H * rho);
    if (R > G.rmax[0]) G.rmax[0] = R;
    if (!(R <= G.thr)) {
        G.gate[0] = 0;
        G.fallback[0] = G.sc->nstep;
    }
}

struct ConvArgs {
    const bf16 *in;
    const bf16x8 *w;
    const float *bias;
    int N, depth;            /* depth: planes per workgroup column */
    long N3;
    bf16 *z_out, *a_out;     /* HIDDEN: [N3][16] */
    const bf16 *zin;         /* BWD: pre-activation of the layer below, [N3][16] */
    bf16 *g_out;             /* BWD: [N3][16] */
    float *y_out;            /* OUT: [N3][2] */
    const int *gate;         /* skip everything when gate[0] == 0 */
    GateArgs gatev;          /* first layer after a step: evaluate the gate (gatev.gate set) */
    const float *fin[4];     /* 4-channel input as FP32 arrays (used instead of in when set) */
    float fscale[4];
};

/* logistic function with the hardware reciprocal (1 ulp): the results are
 * rounded to BF16 right after */
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigm(z);
    return s * (1.f + z * (1.f - s));
}

/* XCD-aware workgroup order: consecutive hardware workgroups go round-robin
 * to the 8 XCDs; give each XCD a contiguous range of logical workgroups so
 * the halo sites of neighbours share its L2 */
__device__ __forceinline__ int logical_wg()
{
    const int nwg = gridDim.x, hw = blockIdx.x;
    return nwg % GABE_NXCD == 0 ? (hw % GABE_NXCD) * (nwg / GABE_NXCD) + hw / GABE_NXCD : hw;
}

/* A workgroup owns a column of TX (k) x TY (j) sites and `depth` planes
 * along i, and marches down it TZ planes (two per wave) at a time.  The input
 * planes with their 1-site halo live channel-last in a ring of RING plane
 * slots in LDS: the TZ + 2 planes the current tile reads and the TZ the next
 * tile adds, which are fetched into registers while the current tile's MFMAs
 * run.
 * Each wave keeps its weight operands in registers and walks the TY + 2 halo
 * rows of its plane: the activations of one (halo row, tap-column step) are
 * one aligned LDS read and feed the MFMAs of every output row y = yy - dy.
 * The MFMA computes D[channel][site], so a lane ends with 4 consecutive
 * channels of one site. */
constexpr int TX = 16, TY = 4, ZW = 2, TZ = 4 * ZW;   /* ZW planes per wave */
constexpr int HX = TX + 2, HY = TY + 2, PLANE = HX * HY;
constexpr int RING = 2 * TZ + 2;

__device__ __forceinline__ int wrap(int x, int N) { return x < 0 ? x + N : (x >= N ? x - N : x); }

template <int CI, int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_conv(ConvArgs A)
{
    constexpr bool OUTL = EPI == EPI_OUT;
    static_assert(!OUTL || (CI == 16 && ZW == 2), "the output layer reads 16 channels, two planes per wave");
    constexpr int NSTEP = OUTL ? OUT_STEPS : (CI == 16 ? WIDE_STEPS : NARROW_STEPS);
    constexpr int NDY = OUTL ? 1 : 3;             /* weight operands per step */
    constexpr int CH = CI == 16 ? 2 : 1;          /* copy chunks per site: 16-byte halves, or one 8-byte site */
    constexpr int CE = CI / CH;                   /* bf16 per chunk */
    using Chunk = typename std::conditional<CI == 16, bf16x8, bf16x4>::type;
    constexpr int NEW = TZ * PLANE * CH;          /* chunks of the planes one tile adds */
    constexpr int NLD = (NEW + 255) / 256;
    __shared__ __align__(16) bf16 tin[RING * PLANE * CI];
    if (A.gate && A.gate[0] == 0) return;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int N = A.N, D = A.depth;
    const long N2 = (long)N * N;
    const int wg = logical_wg();
    const int ntx = N / TX, nty = N / TY;
    const int k0 = (wg % ntx) * TX, j0 = ((wg / ntx) % nty) * TY, ib = (wg / (ntx * nty)) * D;

    /* chunk c of the planes starting at plane offset o0 (plane ib - 1 + o) */
    auto src = [&](int c, int o0) -> long {
        const int pl = c / (PLANE * CH), rem = c - pl * (PLANE * CH);
        const int hs = rem / CH, part = rem - hs * CH;
        const int xx = hs % HX, yy = hs / HX;
        const int ii = wrap(ib - 1 + o0 + pl, N), jj = wrap(j0 + yy - 1, N), kk = wrap(k0 + xx - 1, N);
        return ((long)ii * N2 + (long)jj * N + kk) * CI + part * CE;
    };
    auto dst = [&](int c, int o0) -> int {
        const int pl = c / (PLANE * CH), rem = c - pl * (PLANE * CH);
        return ((o0 + pl) % RING) * PLANE * CI + rem * CE;
    };
    /* one chunk at element offset off of the input: BF16 activations, or for
     * the first layer of a step correction the four FP32 lattice arrays,
     * scaled and rounded to BF16 here */
    auto fetch = [&](int off) -> Chunk {
        if constexpr (CI == 4) {
            if (A.fin[0]) {
                const int site = off >> 2;
                bf16x4 x;
#pragma unroll
                for (int q = 0; q < 4; q++) x[q] = (bf16)(A.fin[q][site] * A.fscale[q]);
                return x;
            }
        }
        return *(const Chunk *)(A.in + off);
    };
    for (int c = tid; c < (TZ + 2) * PLANE * CH; c += 256) *(Chunk *)(tin + dst(c, 0)) = fetch((int)src(c, 0));

    /* this thread's chunks of the TZ planes each step adds, fixed through the
     * march: plane within the step and LDS offset within a plane slot
     * (packed), and the (j, k, half) part of the global offset */
    int cpl[NLD], cjk[NLD];
#pragma unroll
    for (int m = 0; m < NLD; m++) {
        const int c = tid + 256 * m;
        const int pl = c / (PLANE * CH), rem = c - pl * (PLANE * CH);
        const int hs = rem / CH, part = rem - hs * CH;
        const int jj = wrap(j0 + hs / HX - 1, N), kk = wrap(k0 + hs % HX - 1, N);
        cpl[m] = c < NEW ? (pl << 16) | (rem * CE) : -1;
        cjk[m] = (jj * N + kk) * CI + part * CE;
    }
    const int N2CI = N * N * CI;

    bf16x8 wr[NSTEP][NDY];
#pragma unroll
    for (int s = 0; s < NSTEP; s++)
#pragma unroll
        for (int dy = 0; dy < NDY; dy++) wr[s][dy] = A.w[(s * NDY + dy) * 64 + lane];
    /* bias of the lane's result channels: 4g .. 4g + 3, or in the output
     * layer output channel g & 1 */
    float bias[4];
#pragma unroll
    for (int q = 0; q < 4; q++) bias[q] = A.bias ? A.bias[OUTL ? (g & 1) : 4 * g + q] : 0.f;

    /* this lane's K slots: per step one tap column (8 of its 16 channels) or
     * two (4 channels each): the column's dz (in the output layer the plane
     * 0..3 of the wave's four input planes), and the offset of its dx and
     * channels within a plane slot.  Slots past the 9 columns carry zero
     * weights; they read the last column, so no operand needs a select. */
    constexpr int NC = CI == 16 ? 1 : 2;
    int cdz[NSTEP][NC], coff[NSTEP][NC];
#pragma unroll
    for (int s = 0; s < NSTEP; s++)
#pragma unroll
        for (int c = 0; c < NC; c++) {
            int col = CI == 16 ? 2 * s + (g >> 1) : 8 * s + 2 * g + c;
            if (!OUTL) col = col < TAP_COLS ? col : TAP_COLS - 1;
            cdz[s][c] = col / 3;
            coff[s][c] = (r + col % 3) * CI + (CI == 16 ? 8 * (g & 1) : 0);
        }
    /* the weights and biases are loop invariant: wait for them here, so the
     * only vector loads in flight inside the march are the next planes */
    __builtin_amdgcn_s_waitcnt(0x0F70);   /* vmcnt(0) */
    __syncthreads();
    /* the first layer after a step evaluates the gate in one thread, after
     * every thread of its workgroup has read the flag; the layers after it
     * (and this layer's workgroups that read the flag later) skip once it
     * closes, and what this layer wrote is then never read */
    if constexpr (CI == 4 && EPI == EPI_HIDDEN)
        if (A.gatev.gate && blockIdx.x == 0 && tid == 0) (void)gate_eval(A.gatev);

    const int nit = D / TZ;
    for (int n = 0; n < nit; n++) {
        const bool more = n + 1 < nit;
        Chunk pre[NLD];
        const int o0 = TZ * n + TZ + 2;   /* plane offset of the first plane the next step adds */
        if (more) {
#pragma unroll
            for (int m = 0; m < NLD; m++)
                if (cpl[m] >= 0) {
                    const int ii = wrap(ib - 1 + o0 + (cpl[m] >> 16), N);
                    pre[m] = fetch(__umul24(ii, N2CI) + cjk[m]);
                }
        }
        if constexpr (OUTL) {
            /* the wave's output planes z0, z0 + 1 read input planes z0 - 1 ..
             * z0 + 2 (slots of plane offsets z0 .. z0 + 3); per halo row yy one
             * chain of NSTEP MFMAs gives P[yy]: lane (r, g) holds, for output
             * channel g & 1 of plane z0 + (g >> 1), the partial sums of dy = 0,
             * 1, 2 over that row, and output row y sums halo rows y .. y + 2 */
            const int z0 = ZW * wave;
            int ad[NSTEP];
#pragma unroll
            for (int s = 0; s < NSTEP; s++) ad[s] = ((TZ * n + z0 + cdz[s][0]) % RING) * PLANE * CI + coff[s][0];
            bf16x8 avc[NSTEP], avn[NSTEP];
#pragma unroll
            for (int s = 0; s < NSTEP; s++) avc[s] = *(const bf16x8 *)(tin + ad[s]);
            f32x4 P[TY + 2];
#pragma unroll
            for (int yy = 0; yy < TY + 2; yy++) {
                if (yy + 1 < TY + 2) {
#pragma unroll
                    for (int s = 0; s < NSTEP; s++) avn[s] = *(const bf16x8 *)(tin + ad[s] + (yy + 1) * HX * 16);
                }
                f32x4 p = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < NSTEP; s++) p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[s][0], avc[s], p, 0, 0, 0);
                P[yy] = p;
#pragma unroll
                for (int s = 0; s < NSTEP; s++) avc[s] = avn[s];
            }
            const int i = ib + TZ * n + z0 + (g >> 1);
#pragma unroll
            for (int y = 0; y < TY; y++) {
                const int site = __umul24(__umul24(i, N) + j0 + y, N) + k0 + r;
                A.y_out[site * 2 + (g & 1)] = (P[y][0] + P[y + 1][1]) + P[y + 2][2] + bias[0];
            }
        } else {
#pragma unroll
            for (int zp = 0; zp < ZW; zp++) {
                const int z = ZW * wave + zp;
                /* plane slots of this wave's planes i = ib + TZ n + z + dz - 1 */
                int slot[3];
#pragma unroll
                for (int dz = 0; dz < 3; dz++) slot[dz] = ((TZ * n + z + dz) % RING) * PLANE * CI;
                int ad[NSTEP][NC];
#pragma unroll
                for (int s = 0; s < NSTEP; s++)
#pragma unroll
                    for (int c = 0; c < NC; c++) {
                        const int d = cdz[s][c];
                        ad[s][c] = (d == 0 ? slot[0] : (d == 1 ? slot[1] : slot[2])) + coff[s][c];
                    }
                /* the activation operands of halo row yy + 1 are read from LDS while
                 * the MFMAs of row yy run */
                auto read_row = [&](bf16x8 *av, int yy) {
#pragma unroll
                    for (int s = 0; s < NSTEP; s++) {
                        if constexpr (CI == 16) {
                            av[s] = *(const bf16x8 *)(tin + ad[s][0] + yy * HX * 16);
                        } else {
                            const bf16x4 lo = *(const bf16x4 *)(tin + ad[s][0] + yy * HX * 4);
                            const bf16x4 hi = *(const bf16x4 *)(tin + ad[s][1] + yy * HX * 4);
                            av[s] = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
                        }
                    }
                };
                bf16x8 avc[NSTEP], avn[NSTEP];
                read_row(avc, 0);
                f32x4 acc[TY];
#pragma unroll
                for (int y = 0; y < TY; y++) acc[y] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int yy = 0; yy < TY + 2; yy++) {
                    if (yy + 1 < TY + 2) read_row(avn, yy + 1);
#pragma unroll
                    for (int s = 0; s < NSTEP; s++)
#pragma unroll
                        for (int dy = 0; dy < 3; dy++) {
                            const int y = yy - dy;
                            if (y >= 0 && y < TY)
                                acc[y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[s][dy], avc[s], acc[y], 0, 0, 0);
                        }
#pragma unroll
                    for (int s = 0; s < NSTEP; s++) avc[s] = avn[s];
                }
                /* lane holds D[channels 4g .. 4g + 3][site k0 + r of row (i, j0 + y)] */
                const int i = ib + TZ * n + z;
#pragma unroll
                for (int y = 0; y < TY; y++) {
                    const int site = __umul24(__umul24(i, N) + j0 + y, N) + k0 + r;
                    float v[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) v[q] = acc[y][q] + bias[q];
                    if constexpr (EPI == EPI_HIDDEN_TRAIN || EPI == EPI_HIDDEN) {
                        bf16x4 h;
#pragma unroll
                        for (int q = 0; q < 4; q++) h[q] = (bf16)silu(v[q]);
                        *(bf16x4 *)(A.a_out + site * 16 + 4 * g) = h;
                        if constexpr (EPI == EPI_HIDDEN_TRAIN) {
                            bf16x4 zq;
#pragma unroll
                            for (int q = 0; q < 4; q++) zq[q] = (bf16)v[q];
                            *(bf16x4 *)(A.z_out + site * 16 + 4 * g) = zq;
                        }
                    } else if constexpr (EPI == EPI_BWD) {
                        const bf16x4 zq = *(const bf16x4 *)(A.zin + site * 16 + 4 * g);
                        bf16x4 o;
#pragma unroll
                        for (int q = 0; q < 4; q++) o[q] = (bf16)(acc[y][q] * dsilu((float)zq[q]));
                        *(bf16x4 *)(A.g_out + site * 16 + 4 * g) = o;
                    }
                }
            }
        }
        if (more) {
            const int sb = o0 % RING;
#pragma unroll
            for (int m = 0; m < NLD; m++)
                if (cpl[m] >= 0) {
                    int sl = sb + (cpl[m] >> 16);
                    sl = sl >= RING ? sl - RING : sl;
                    *(Chunk *)(tin + sl * PLANE * CI + (cpl[m] & 0xFFFF)) = pre[m];
                }
            __syncthreads();
        }
    }
}

/* ------------------------------------------------- the Laplacian output */
/* lap of both channels of u [N3][2] at site s */
/* site index -> (i, j, k) in 32-bit arithmetic (lattices here hold < 2^31 sites) */
__device__ __forceinline__ void site_ijk(long s, int N, int &i, int &j, int &k)
{
    const unsigned n = (unsigned)N, q = (unsigned)s / n;
    k = (int)((unsigned)s - q * n);
    i = (int)(q / n);
    j = (int)(q - (unsigned)i * n);
}

__device__ __forceinline__ float2 lap2(const float *u, long s, int N, long N2)
{
    int i, j, k;
    site_ijk(s, N, i, j, k);
    const int n2 = (int)N2, ss = (int)s;
    const int ip = i + 1 == N ? ss - i * n2 : ss + n2, im = i == 0 ? ss + (N - 1) * n2 : ss - n2;
    const int jp = j + 1 == N ? ss - j * N : ss + N, jm = j == 0 ? ss + (N - 1) * N : ss - N;
    const int kp = k + 1 == N ? ss - k : ss + 1, km = k == 0 ? ss + N - 1 : ss - 1;
    const float2 c = *(const float2 *)(u + 2 * s);
    float2 sum = *(const float2 *)(u + 2 * ip);
    for (int q : {im, jp, jm, kp, km}) {
        const float2 v = *(const float2 *)(u + 2 * (long)q);
        sum.x += v.x;
        sum.y += v.y;
    }
    return make_float2(sum.x - 6.f * c.x, sum.y - 6.f * c.y);
}

/* The correction after a step, v_c += dt s_out[c] (lap w)_c / a^2 with
 * w = lap u, in one march: a workgroup owns a TK x TJ tile of (k, j) columns
 * (64 x 4 or 32 x 8, one column per thread) and D planes of i.  Planes of u
 * with a two-site edge, (TJ + 4)(TK + 4) sites, sit in a ring of four LDS
 * slots and planes of w with a one-site edge in a ring of three; plane i + 3
 * of u is in flight in registers while w(i + 1) is computed and plane i
 * corrected, and plane i's velocities are loaded before w(i + 1) is.  Both
 * Laplacians are lap2's arithmetic, operand for operand.  Workgroups go to
 * the XCDs in contiguous ranges, so the edge reads of neighbouring tiles meet
 * in one L2. */
constexpr int KU = 8 * 68, KW = 6 * 66;   /* slot sizes at 64 x 4, the larger of the two tile shapes */

__host__ __device__ inline int kick_tk(int N) { return N % 64 == 0 ? 64 : 32; }

__global__ __launch_bounds__(256) void k_lap2_kick(const float *u, int N, int D, float *v0, float *v1,
                                                   const DevScalars *sc, float dts0, float dts1, const int *gate)
{
    if (gate && gate[0] == 0) return;
    __shared__ float2 us[4][KU];
    __shared__ float2 ws[3][KW];
    const int TK = kick_tk(N), TJ = 256 / TK, UW = TK + 4, WW = TK + 2;
    const int nu = (TJ + 4) * UW, nw = (TJ + 2) * WW;
    const int tid = threadIdx.x, ntk = N / TK, ntj = N / TJ, wg = logical_wg();
    const int k0 = (wg % ntk) * TK, j0 = ((wg / ntk) % ntj) * TJ, i0 = (wg / (ntk * ntj)) * D;
    const int i1 = i0 + D, N2 = N * N;
    const float2 *u2 = reinterpret_cast<const float2 *>(u);
    /* the thread's u sites of a plane: slot index tid + 256 m, m < 3 */
    int uoff[3];
    bool uval[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const int h = tid + 256 * m;
        uval[m] = h < nu;
        const int hh = uval[m] ? h : 0, r = hh / UW, c = hh - r * UW;
        uoff[m] = wrap(j0 + r - 2, N) * N + wrap(k0 + c - 2, N);
    }
    /* plane p of the march, p in [i0 - 2, i1 + 1]; loads past i1 + 1 read
     * plane i1 + 1 again (an L2 hit) into a slot that is never read */
    auto uload = [&](float2 (&r)[3], int p) {
        const int pw = wrap(min(p, i1 + 1), N);
#pragma unroll
        for (int m = 0; m < 3; m++)
            if (uval[m]) r[m] = u2[pw * N2 + uoff[m]];
    };
    auto uput = [&](const float2 (&r)[3], int slot) {
#pragma unroll
        for (int m = 0; m < 3; m++)
            if (uval[m]) us[slot][tid + 256 * m] = r[m];
    };
    auto su = [&](int p) { return (p - i0 + 2) & 3; };
    auto sw = [&](int p) { return (p - i0 + 1) % 3; };
    /* w at plane p from u planes p - 1, p, p + 1 */
    auto wcomp = [&](int p) {
        const float2 *um = us[su(p - 1)], *uc = us[su(p)], *up = us[su(p + 1)];
        float2 *w = ws[sw(p)];
        for (int h = tid; h < nw; h += 256) {
            const int r = h / WW, c = h - r * WW, q = (r + 1) * UW + c + 1;
            const float2 cc = uc[q];
            float2 sum = up[q];
            for (float2 v : {um[q], uc[q + UW], uc[q - UW], uc[q + 1], uc[q - 1]}) {
                sum.x += v.x;
                sum.y += v.y;
            }
            w[h] = make_float2(sum.x - 6.f * cc.x, sum.y - 6.f * cc.y);
        }
    };
    const double a = sc->a[0];
    const float ia2 = (float)(1. / (a * a));
    const int tr = tid / TK, tc = tid - tr * TK, rc = (tr + 1) * WW + tc + 1;
    const int soff = (j0 + tr) * N + k0 + tc;

    float2 r[3];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uload(r, i0 - 2 + q);
        uput(r, q);
    }
    __syncthreads();
    wcomp(i0 - 1);
    wcomp(i0);
    __syncthreads();
    uload(r, i0 + 2);
    uput(r, su(i0 + 2));
    __syncthreads();
    for (int i = i0; i < i1; i++) {
        const int s = i * N2 + soff;
        const float va = v0[s], vb = v1[s];
        uload(r, i + 3);
        wcomp(i + 1);
        __syncthreads();
        const float2 *wm = ws[sw(i - 1)], *wc = ws[sw(i)], *wp = ws[sw(i + 1)];
        const float2 cc = wc[rc];
        float2 sum = wp[rc];
        for (float2 v : {wm[rc], wc[rc + WW], wc[rc - WW], wc[rc + 1], wc[rc - 1]}) {
            sum.x += v.x;
            sum.y += v.y;
        }
        const float2 y = make_float2(sum.x - 6.f * cc.x, sum.y - 6.f * cc.y);
        v0[s] = va + dts0 * ia2 * y.x;
        v1[s] = vb + dts1 * ia2 * y.y;
        uput(r, su(i + 3));
        __syncthreads();
    }
}

/* planes per kick workgroup: the most of 16, 8, 4 that still gives 1024
 * workgroups (128^3: 8 planes, 20.0-20.7 us against 23.1 at 16 and 22.5 at
 * 4, whose ring fills cost more than its four workgroups per CU gain) */
int kick_depth(int N)
{
    const int TK = kick_tk(N), tiles = (N / TK) * (N / (256 / TK));
    for (int d : {16, 8, 4})
        if (N % d == 0 && tiles * (N / d) >= 1024) return d;
    return 4;
}

/* loss terms (w - t)^2 (block partials) and dL/dw = (w - t) / N^3 for
 * w = lap u [N3][2] and targets t [2][N3] */
__global__ __launch_bounds__(256) void k_w_loss(const float *w, const float *t, long N3, float *gy, double *part)
{
    __shared__ double red[4];
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.;
    if (s < N3) {
        const float2 y = *(const float2 *)(w + 2 * s);
        const float d0 = y.x - t[s], d1 = y.y - t[N3 + s];
        acc = (double)d0 * d0 + (double)d1 * d1;
        const float inv = (float)(1. / (double)N3);
        *(float2 *)(gy + 2 * s) = make_float2(d0 * inv, d1 * inv);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

/* y = lap u, both channels, [N3][2] */
__global__ __launch_bounds__(256) void k_lap_out(const float *u, int N, float *y)
{
    const long N2 = (long)N * N, N3 = N2 * N;
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N3) return;
    *(float2 *)(y + 2 * s) = lap2(u, s, N, N2);
}

/* dL/du = lap dL/d(lap u) (lap is self-adjoint), as the BF16 input [N3][4]
 * of the backward convolution of layer 3 (channels 2, 3 zero) */
__global__ __launch_bounds__(256) void k_lap_grad(const float *gy, int N, bf16 *g3)
{
    const long N2 = (long)N * N, N3 = N2 * N;
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N3) return;
    const float2 g = lap2(gy, s, N, N2);
    bf16x4 o;
    o[0] = (bf16)g.x;
    o[1] = (bf16)g.y;
    o[2] = o[3] = (bf16)0.f;
    *(bf16x4 *)(g3 + 4 * s) = o;
}

/* ------------------------------------------------------ weight gradient */
struct WgradArgs {
    const bf16 *gz;      /* [N3][gzc] */
    const bf16 *act;     /* [N3][CI] */
    int gzc, co;         /* channels per site of gz, channels used */
    int N;
    long N3;
    float *part;         /* [nwg][NG][16][16] */
    float *bias_part;    /* [nwg][16] */
};

/* dW[co][ci][t] = sum_s gz[s][co] act[s + off(t)][ci] and db[co] = sum_s gz[s][co]
 * over WG_SLABS slabs (one i-plane, WG_ROWS rows, all k) per workgroup.  LDS
 * holds gz channel-major and act channel-major in three copies shifted by
 * dx = -1, 0, +1, so every MFMA operand is one aligned 16-byte read. */
template <int CI>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs A)
{
    constexpr int NG = CI == 16 ? CL_TAPS : 7;   /* tap groups: one tap, or four taps of 4 channels */
    constexpr int GPW = (NG + 3) / 4;            /* groups per wave */
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = A.N;
    const long N2 = (long)N * N;
    const int SL = WG_ROWS * N;                  /* sites per slab */
    const int RH = WG_ROWS + 2;
    bf16 *gzT = (bf16 *)smem;                    /* [16][SL] */
    bf16 *actT = gzT + 16 * SL;                  /* [3 dx][CI][3 planes][RH][N] */
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int wg = logical_wg();

    f32x4 acc[GPW];
    for (int q = 0; q < GPW; q++) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    const int slabs_per_plane = N / WG_ROWS;
    for (int sl = 0; sl < WG_SLABS; sl++) {
        const int slab = wg * WG_SLABS + sl;
        const int i = slab / slabs_per_plane, j0 = (slab % slabs_per_plane) * WG_ROWS;
        __syncthreads();
        /* stage gz (channel-major) */
        for (int q = threadIdx.x; q < SL; q += 256) {
            const int row = q / N, k = q % N;
            const long site = (long)i * N2 + (long)(j0 + row) * N + k;
            for (int c = 0; c < 16; c++)
                gzT[c * SL + q] = c < A.co ? A.gz[site * A.gzc + c] : (bf16)0.f;
        }
        /* stage act: planes i-1..i+1, rows j0-1..j0+WG_ROWS, three dx-shifted copies */
        for (int q = threadIdx.x; q < 3 * RH * N; q += 256) {
            const int pl = q / (RH * N), rr = (q / N) % RH, k = q % N;
            int ii = i + pl - 1, jj = j0 + rr - 1;
            ii = ii < 0 ? ii + N : (ii >= N ? ii - N : ii);
            jj = jj < 0 ? jj + N : (jj >= N ? jj - N : jj);
            const long site = (long)ii * N2 + (long)jj * N + k;
            bf16 v[CI];
            if constexpr (CI == 16) {
                *(bf16x8 *)v = *(const bf16x8 *)(A.act + site * 16);
                *(bf16x8 *)(v + 8) = *(const bf16x8 *)(A.act + site * 16 + 8);
            } else {
                *(bf16x4 *)v = *(const bf16x4 *)(A.act + site * 4);
            }
            /* copy dxi holds act at k' = k + dx, so site k lands at k - dx */
            for (int dxi = 0; dxi < 3; dxi++) {
                int kd = k - (dxi - 1);
                kd = kd < 0 ? kd + N : (kd >= N ? kd - N : kd);
                for (int c = 0; c < CI; c++) actT[(((dxi * CI + c) * 3 + pl) * RH + rr) * N + kd] = v[c];
            }
        }
        __syncthreads();
        if (threadIdx.x < 16)
            for (int q = 0; q < SL; q++) bsum += (float)gzT[threadIdx.x * SL + q];
        for (int kk = 0; kk < SL / 32; kk++) {
            const int q0 = kk * 32;
            const int row = q0 / N, kb = q0 % N + 8 * g;
            const bf16x8 av = *(const bf16x8 *)(gzT + r * SL + q0 + 8 * g);
#pragma unroll
            for (int gi = 0; gi < GPW; gi++) {
                const int grp = wave + 4 * gi;
                if (grp >= NG) break;
                int t, c;
                if constexpr (CI == 16) {
                    t = grp;
                    c = r;
                } else {
                    t = 4 * grp + (r >> 2);
                    c = r & 3;
                }
                bf16x8 bv;
                if (t < CL_TAPS) {
                    const int dz = t / 9, dy = (t / 3) % 3, dx = t % 3;
                    bv = *(const bf16x8 *)(actT + (((dx * CI + c) * 3 + dz) * RH + row + dy) * N + kb);
                } else {
                    for (int e = 0; e < 8; e++) bv[e] = (bf16)0.f;
                }
                acc[gi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[gi], 0, 0, 0);
            }
        }
    }
    /* D[co = 4g+q][n = r] of group grp */
    for (int gi = 0; gi < GPW; gi++) {
        const int grp = wave + 4 * gi;
        if (grp >= NG) break;
        for (int q = 0; q < 4; q++) A.part[(((long)wg * NG + grp) * 16 + 4 * g + q) * 16 + r] = acc[gi][q];
    }
    if (threadIdx.x < 16) A.bias_part[(long)wg * 16 + threadIdx.x] = bsum;
}

/* grad[param] = sum over workgroups, fixed order.  layer 1: W1[co][ci][t] from
 * group t/4, column 4 (t%4) + ci; layers 2/3: W[co][ci][t] from group t,
 * column ci. */
__global__ void k_wgrad_reduce(const float *part, const float *bias_part, int nwg, int NG, int CO, int CI,
                               float *gw, float *gb)
{
    const int nw = CO * CI * CL_TAPS;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nw) {
        const int co = e / (CI * CL_TAPS), ci = (e / CL_TAPS) % CI, t = e % CL_TAPS;
        const int grp = CI == 4 ? t / 4 : t, col = CI == 4 ? 4 * (t % 4) + ci : ci;
        double s = 0.;
        for (int w = 0; w < nwg; w++) s += part[(((long)w * NG + grp) * 16 + co) * 16 + col];
        gw[e] = (float)s;
    } else if (e < nw + CO) {
        const int co = e - nw;
        double s = 0.;
        for (int w = 0; w < nwg; w++) s += bias_part[(long)w * 16 + co];
        gb[co] = (float)s;
    }
}

/* ----------------------------------------------------------- packing */
/* Weight operand value, step sd = 3 s + dy, of a conv with input channels
 * cin (padded to 16 or 4 slots) and output channels cout, from W[co][ci][t]
 * (or, flipped, from the layer whose weights are Wf[ci][co][26 - t]). */
__device__ float packed_weight(const float *W, int wco, int wci, bool flip, int wide, int sd, int lane, int e)
{
    const int n = lane & 15, g = lane >> 4, s = sd / 3, dy = sd % 3;
    int col, ci;
    if (wide) {
        col = 2 * s + (g >> 1);
        ci = 8 * (g & 1) + e;
    } else {
        col = 8 * s + 2 * g + (e >> 2);
        ci = e & 3;
    }
    if (col >= TAP_COLS) return 0.f;
    const int t = 9 * (col / 3) + 3 * dy + col % 3;
    if (!flip) {   /* conv from wci channels to wco channels */
        if (n >= wco || ci >= wci) return 0.f;
        return W[(n * wci + ci) * CL_TAPS + t];
    }
    /* transposed: input channels are the layer's outputs (wco), outputs its inputs (wci) */
    if (n >= wci || ci >= wco) return 0.f;
    return W[(ci * wci + n) * CL_TAPS + (CL_TAPS - 1 - t)];
}

/* Output-layer operand of step s: row n = lane & 15 is (plane zp = n >> 3 of
 * the wave's pair, output channel co = (n >> 2) & 1, dy = n & 3, dy = 3
 * empty); K chunk g = lane >> 4 is the column 2 s + (g >> 1) = 3 pl + dx of
 * input plane pl (0..3) with channels 8 (g & 1) .. + 7, the tap dz = pl - zp. */
__device__ float out_weight(const float *W3, int s, int lane, int e)
{
    const int n = lane & 15, g = lane >> 4;
    const int zp = n >> 3, co = (n >> 2) & 1, dy = n & 3;
    const int col = 2 * s + (g >> 1), pl = col / 3, dx = col % 3, ci = 8 * (g & 1) + e;
    const int dz = pl - zp;
    if (dy > 2 || dz < 0 || dz > 2) return 0.f;
    return W3[(co * CL_H + ci) * CL_TAPS + 9 * dz + 3 * dy + dx];
}

__global__ void k_pack(const float *P, bf16x8 *pk, float *bias)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < (long)PK_TOTAL) {
        const float *W;
        int wco, wci, wide, s;   /* wide: 0 four-channel steps, 1 sixteen-channel steps, 2 the output layer */
        bool flip;
        long off;
        if (q < (long)PK_F2) { W = P + P_W1; wco = 16; wci = 4; flip = false; wide = 0; off = PK_F1; }
        else if (q < (long)PK_F3) { W = P + P_W2; wco = 16; wci = 16; flip = false; wide = 1; off = PK_F2; }
        else if (q < (long)PK_B3) { W = P + P_W3; wco = 2; wci = 16; flip = false; wide = 2; off = PK_F3; }
        else if (q < (long)PK_B2) { W = P + P_W3; wco = 2; wci = 16; flip = true; wide = 0; off = PK_B3; }
        else { W = P + P_W2; wco = 16; wci = 16; flip = true; wide = 1; off = PK_B2; }
        s = (int)((q - off) / 64);
        const int lane = (int)((q - off) % 64);
        bf16x8 v;
        for (int e = 0; e < 8; e++)
            v[e] = (bf16)(wide == 2 ? out_weight(W, s, lane, e) : packed_weight(W, wco, wci, flip, wide, s, lane, e));
        pk[q] = v;
    }
    if (q < 48) {
        const int l = (int)(q / 16), c = (int)(q % 16);
        const int off = l == 0 ? P_B1 : (l == 1 ? P_B2 : P_B3), n = l == 2 ? CL_OUT : CL_H;
        bias[q] = c < n ? P[off + c] : 0.f;
    }
}

__global__ void k_adam(float *P, const float *G, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float g = G[e];
    const float mm = b1 * m[e] + (1.f - b1) * g;
    const float vv = b2 * v[e] + (1.f - b2) * g * g;
    m[e] = mm;
    v[e] = vv;
    P[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

/* ------------------------------------------------------ samples */
__global__ void k_make_x(const float *f0, const float *f1, const float *v0, const float *v1, float i0, float i1,
                         float i2, float i3, long N3, bf16 *x, const int *gate)
{
    if (gate && gate[0] == 0) return;
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N3) return;
    bf16x4 o;
    o[0] = (bf16)(f0[s] * i0);
    o[1] = (bf16)(f1[s] * i1);
    o[2] = (bf16)(v0[s] * i2);
    o[3] = (bf16)(v1[s] * i3);
    *(bf16x4 *)(x + s * 4) = o;
}

/* 7-point lattice Laplacian of the two fluctuation fields f [2][N3]:
 * out_c = (sum of the 6 neighbours - 6 f_c) / dx^2 */
__device__ __forceinline__ float lap_soa(const float *f, long s, int N, long N2)
{
    int i, j, k;
    site_ijk(s, N, i, j, k);
    const long base = s - k;
    float sum = f[(i + 1 == N ? s - (long)i * N2 : s + N2)] + f[(i == 0 ? s + (long)(N - 1) * N2 : s - N2)];
    sum += f[(j + 1 == N ? s - (long)j * N : s + N)] + f[(j == 0 ? s + (long)(N - 1) * N : s - N)];
    sum += f[base + (k + 1 == N ? 0 : k + 1)] + f[base + (k == 0 ? N - 1 : k - 1)];
    return sum - 6.f * f[s];
}

__global__ __launch_bounds__(256) void k_lap_field(const float *f, int N, float inv_dx2, float *out)
{
    const long N2 = (long)N * N, N3 = N2 * N;
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N3) return;
    out[s] = lap_soa(f, s, N, N2) * inv_dx2;
    out[N3 + s] = lap_soa(f + N3, s, N, N2) * inv_dx2;
}

/* w-target t_c <- (t_c - f_c) * scale_c: lap^-1 of the restricted fine
 * Laplacian minus the restricted field, i.e. lap^-1 of the force target */
__global__ __launch_bounds__(256) void k_sub_scale2(const float *f, long N3, float sc0, float sc1, float *t)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N3) return;
    t[s] = (t[s] - f[s]) * sc0;
    t[N3 + s] = (t[N3 + s] - f[N3 + s]) * sc1;
}

/* target t_c <- (t_c - lap f_c / dx^2) * scale_c: the restricted fine
 * Laplacian minus the coarse Laplacian of the restricted field */
__global__ __launch_bounds__(256) void k_lap_target(const float *f, int N, float inv_dx2, float sc0, float sc1, float *t)
{
    const long N2 = (long)N * N, N3 = N2 * N;
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N3) return;
    t[s] = (t[s] - lap_soa(f, s, N, N2) * inv_dx2) * sc0;
    t[N3 + s] = (t[N3 + s] - lap_soa(f + N3, s, N, N2) * inv_dx2) * sc1;
}

__global__ void k_sumsq(const float *x, long n, double *part)
{
    __shared__ double red[4];
    double acc = 0.;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        acc += (double)x[q] * x[q];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

/* coarse half-spectrum <- fine half-spectrum on |n_a| < Nc/2, times scale;
 * zero on the coarse Nyquist planes.  With inv_lap set, each mode is also
 * divided by the coarse lattice's Laplacian eigenvalue
 * -(4/dx^2) sum_a sin^2(pi n_a / Nc) (inv_lap = dx^2; zero for n = 0): the
 * restriction of lap^-1 of the array. */
template <typename T>
__global__ void k_restrict_modes(const T *fine, int Nf, T *coarse, int Nc, T scale, double inv_lap = 0.)
{
    const int Hc = Nc / 2 + 1, Hf = Nf / 2 + 1;
    const long Mc = (long)Nc * Nc * Hc;
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= Mc) return;
    const int kc = (int)(m % Hc), jc = (int)((m / Hc) % Nc), ic = (int)(m / ((long)Hc * Nc));
    T re = 0, im = 0;
    if (ic != Nc / 2 && jc != Nc / 2 && kc != Nc / 2) {
        const int i_f = ic < Nc / 2 ? ic : Nf - (Nc - ic), j_f = jc < Nc / 2 ? jc : Nf - (Nc - jc);
        const long mf = ((long)i_f * Nf + j_f) * Hf + kc;
        re = fine[2 * mf] * scale;
        im = fine[2 * mf + 1] * scale;
        if (inv_lap != 0.) {
            const double si = sin(M_PI * ic / Nc), sj = sin(M_PI * jc / Nc), sk = sin(M_PI * kc / Nc);
            const double lam = -4. * (si * si + sj * sj + sk * sk);
            const T w = (ic | jc | kc) == 0 ? (T)0 : (T)(inv_lap / lam);
            re *= w;
            im *= w;
        }
    }
    coarse[2 * m] = re;
    coarse[2 * m + 1] = im;
}

inline unsigned nblocks(long n, int b) { return (unsigned)((n + b - 1) / b); }

/* deterministic initial weights: uniform in +-sqrt(3/fan_in) from a 64-bit
 * LCG, biases zero, the output layer scaled down so training starts near the
 * plain step */
void init_params(std::vector<float> &P, unsigned seed)
{
    unsigned long long st = 0x9E3779B97F4A7C15ULL ^ ((unsigned long long)seed << 1);
    auto uni = [&]() {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    P.assign(P_TOTAL, 0.f);
    const struct { int off, n, fan; double gain; } L[3] = {
        {P_W1, CL_H * CL_IN * CL_TAPS, CL_IN * CL_TAPS, 1.0},
        {P_W2, CL_H * CL_H * CL_TAPS, CL_H * CL_TAPS, 1.0},
        {P_W3, CL_OUT * CL_H * CL_TAPS, CL_H * CL_TAPS, 0.1},
    };
    for (const auto &l : L) {
        const double a = l.gain * std::sqrt(3.0 / l.fan);
        for (int e = 0; e < l.n; e++) P[l.off + e] = (float)(a * uni());
    }
}

} // namespace

/* =================================================================== state */
struct gabe_closure {
    int N = 0;
    long N3 = 0;
    double dt = 0.;
    hipStream_t stream = nullptr;
    double s_in[4] = {0, 0, 0, 0}, s_out[2] = {0, 0};
    bool scaled = false;
    float *P = nullptr, *G = nullptr, *m = nullptr, *v = nullptr;
    long long adam_t = 0;
    bf16x8 *pk = nullptr;
    float *bias = nullptr;
    bf16 *z1 = nullptr, *a1 = nullptr, *z2 = nullptr, *a2 = nullptr, *g3 = nullptr, *gz2 = nullptr, *gz1 = nullptr;
    bf16 *xbuf = nullptr;
    float *tbuf = nullptr, *ybuf = nullptr, *gybuf = nullptr, *wbuf = nullptr;
    float *wpart = nullptr, *bpart = nullptr;
    double *lpart = nullptr, *red = nullptr;
    float *flap = nullptr;      /* Laplacian of the fine lattice's fields [2][N3 fine] */
    long flap_n3 = 0;
    std::vector<bf16 *> xs;     /* dataset: inputs [N3][4] */
    std::vector<float *> ts;    /* dataset: targets [2][N3] */
    int wgrad_wg = 0;
};

struct ClosureRun {
    int N = 0;                     /* weights and scales are copies: the closure may go */
    bf16x8 *pk = nullptr;
    float *bias = nullptr;
    bf16 *a1 = nullptr, *a2 = nullptr;
    float *u = nullptr;            /* network output u, [N3][2] */
    int *gate = nullptr;           /* [0] active */
    long long *fallback = nullptr; /* step count when the gate closed */
    double *rmax = nullptr;
    double thr = 0.;
    float inv_s[4], dt_s[2];
};

namespace {

int cl_check_n(int N)
{
    if (N < 32 || N > 128 || N % 32) return fail(GABE_ERR_ARG, "closure lattices need N in {32, 64, 96, 128} (got %d)", N);
    return GABE_OK;
}

/* planes per workgroup column: the longest march (a multiple of TZ dividing
 * N) that still leaves every CU as many workgroups as fit on it at once (2
 * of the 16-channel kernels, 4 of the 4-channel ones) */
int conv_depth(int N, int CI)
{
    const long target = CI == 16 ? 512 : 1024;
    int best = TZ;
    for (int d = TZ; d <= N; d += TZ)
        if (N % d == 0 && (long)N * N * N / ((long)TX * TY * d) >= target) best = d;
    return best;
}

int conv_workgroups(int N, int CI) { return (int)((long)N * N * N / ((long)TX * TY * conv_depth(N, CI))); }

template <int CI, int EPI>
int launch_conv(const ConvArgs &a, hipStream_t s)
{
    ConvArgs b = a;
    b.depth = conv_depth(a.N, CI);
    hipLaunchKernelGGL((k_conv<CI, EPI>), dim3(conv_workgroups(a.N, CI)), dim3(256), 0, s, b);
    HIPCHK(hipGetLastError());
    return GABE_OK;
}

ConvArgs conv_base(int N, const bf16 *in, const bf16x8 *w, const float *bias)
{
    ConvArgs a{};
    a.in = in;
    a.w = w;
    a.bias = bias;
    a.N = N;
    a.N3 = (long)N * N * N;
    return a;
}

size_t wgrad_lds(int CI, int N) { return (size_t)2 * (16 * WG_ROWS * N + 3 * CI * 3 * (WG_ROWS + 2) * N); }

template <int CI>
int launch_wgrad(gabe_closure *cl, const bf16 *gz, int gzc, int co, const bf16 *act, float *gw, float *gb)
{
    WgradArgs w{gz, act, gzc, co, cl->N, cl->N3, cl->wpart, cl->bpart};
    const size_t lds = wgrad_lds(CI, cl->N);
    hipLaunchKernelGGL((k_wgrad<CI>), dim3(cl->wgrad_wg), dim3(256), lds, cl->stream, w);
    HIPCHK(hipGetLastError());
    const int NG = CI == 16 ? CL_TAPS : 7, CIw = CI == 16 ? 16 : 4;
    const int n = co * CIw * CL_TAPS + co;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(nblocks(n, 256)), dim3(256), 0, cl->stream, cl->wpart, cl->bpart,
                       cl->wgrad_wg, NG, co, CIw, gw, gb);
    HIPCHK(hipGetLastError());
    return GABE_OK;
}

int repack(gabe_closure *cl)
{
    hipLaunchKernelGGL(k_pack, dim3(nblocks((long)PK_TOTAL, 256)), dim3(256), 0, cl->stream, cl->P, cl->pk,
                       cl->bias);
    HIPCHK(hipGetLastError());
    return GABE_OK;
}

/* forward (training form: keeps z and a of both hidden layers) and the loss
 * against targets t, leaving dL/dy in g3 */
int forward_loss(gabe_closure *cl, const bf16 *x, const float *t)
{
    ConvArgs a = conv_base(cl->N, x, cl->pk + PK_F1, cl->bias);
    a.z_out = cl->z1;
    a.a_out = cl->a1;
    CHK((launch_conv<4, EPI_HIDDEN_TRAIN>(a, cl->stream)));
    a = conv_base(cl->N, cl->a1, cl->pk + PK_F2, cl->bias + 16);
    a.z_out = cl->z2;
    a.a_out = cl->a2;
    CHK((launch_conv<16, EPI_HIDDEN_TRAIN>(a, cl->stream)));
    a = conv_base(cl->N, cl->a2, cl->pk + PK_F3, cl->bias + 32);
    a.y_out = cl->ybuf;
    CHK((launch_conv<16, EPI_OUT>(a, cl->stream)));
    /* the loss compares w = lap u with targets for lap u (the force targets
     * carried through lap^-1, so that every wavelength weighs by the relative
     * error of the frequency correction it gets); dL/du = lap dL/dw (lap is
     * self-adjoint) */
    const unsigned nb = nblocks(cl->N3, 256);
    hipLaunchKernelGGL(k_lap_out, dim3(nb), dim3(256), 0, cl->stream, cl->ybuf, cl->N, cl->wbuf);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_w_loss, dim3(nb), dim3(256), 0, cl->stream, cl->wbuf, t, cl->N3, cl->gybuf, cl->lpart);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_lap_grad, dim3(nb), dim3(256), 0, cl->stream, cl->gybuf, cl->N, cl->g3);
    HIPCHK(hipGetLastError());
    return GABE_OK;
}

/* gradients of the loss left by forward_loss, into cl->G */
int backward(gabe_closure *cl, const bf16 *x)
{
    CHK(launch_wgrad<16>(cl, cl->g3, 4, CL_OUT, cl->a2, cl->G + P_W3, cl->G + P_B3));
    ConvArgs a = conv_base(cl->N, cl->g3, cl->pk + PK_B3, nullptr);
    a.zin = cl->z2;
    a.g_out = cl->gz2;
    CHK((launch_conv<4, EPI_BWD>(a, cl->stream)));
    CHK(launch_wgrad<16>(cl, cl->gz2, 16, CL_H, cl->a1, cl->G + P_W2, cl->G + P_B2));
    a = conv_base(cl->N, cl->gz2, cl->pk + PK_B2, nullptr);
    a.zin = cl->z1;
    a.g_out = cl->gz1;
    CHK((launch_conv<16, EPI_BWD>(a, cl->stream)));
    CHK(launch_wgrad<4>(cl, cl->gz1, 16, CL_H, x, cl->G + P_W1, cl->G + P_B1));
    return GABE_OK;
}

int read_loss(gabe_closure *cl, double *loss)
{
    std::vector<double> lp(nblocks(cl->N3, 256));
    HIPCHK(hipMemcpyAsync(lp.data(), cl->lpart, sizeof(double) * lp.size(), hipMemcpyDeviceToHost, cl->stream));
    HIPCHK(hipStreamSynchronize(cl->stream));
    *loss = csum(lp.data(), (long)lp.size()) / (2.0 * (double)cl->N3);
    return GABE_OK;
}

int rms_of(gabe_closure *cl, const float *x, long n, double *out)
{
    const int nb = 512;
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, cl->stream, x, n, cl->red);
    HIPCHK(hipGetLastError());
    std::vector<double> p(nb);
    HIPCHK(hipMemcpyAsync(p.data(), cl->red, sizeof(double) * nb, hipMemcpyDeviceToHost, cl->stream));
    HIPCHK(hipStreamSynchronize(cl->stream));
    *out = std::sqrt(csum(p.data(), nb) / (double)n);
    return GABE_OK;
}

void free_closure(gabe_closure *cl)
{
    if (cl->stream) (void)hipStreamSynchronize(cl->stream);
    void *bufs[] = {cl->P, cl->G, cl->m, cl->v, cl->pk, cl->bias, cl->z1, cl->a1, cl->z2, cl->a2, cl->g3, cl->gz2,
                    cl->gz1, cl->xbuf, cl->tbuf, cl->ybuf, cl->gybuf, cl->wbuf, cl->wpart, cl->bpart, cl->lpart, cl->red,
                    cl->flap};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto *p : cl->xs) (void)hipFree(p);
    for (auto *p : cl->ts) (void)hipFree(p);
    if (cl->stream) (void)hipStreamDestroy(cl->stream);
}

int create_closure(gabe_closure *cl, int N, double dt, unsigned seed)
{
    cl->N = N;
    cl->N3 = (long)N * N * N;
    cl->dt = dt;
    cl->wgrad_wg = (int)((long)N * (N / WG_ROWS) / WG_SLABS);
    HIPCHK(hipStreamCreate(&cl->stream));
    const long N3 = cl->N3;
    HIPCHK(hipMalloc(&cl->P, sizeof(float) * P_TOTAL));
    HIPCHK(hipMalloc(&cl->G, sizeof(float) * P_TOTAL));
    HIPCHK(hipMalloc(&cl->m, sizeof(float) * P_TOTAL));
    HIPCHK(hipMalloc(&cl->v, sizeof(float) * P_TOTAL));
    HIPCHK(hipMalloc(&cl->pk, sizeof(bf16x8) * PK_TOTAL));
    HIPCHK(hipMalloc(&cl->bias, sizeof(float) * 48));
    for (bf16 **b : {&cl->z1, &cl->a1, &cl->z2, &cl->a2, &cl->gz2, &cl->gz1}) HIPCHK(hipMalloc(b, sizeof(bf16) * 16 * N3));
    HIPCHK(hipMalloc(&cl->g3, sizeof(bf16) * 4 * N3));
    HIPCHK(hipMalloc(&cl->xbuf, sizeof(bf16) * 4 * N3));
    HIPCHK(hipMalloc(&cl->tbuf, sizeof(float) * 2 * N3));
    HIPCHK(hipMalloc(&cl->ybuf, sizeof(float) * 2 * N3));
    HIPCHK(hipMalloc(&cl->gybuf, sizeof(float) * 2 * N3));
    HIPCHK(hipMalloc(&cl->wbuf, sizeof(float) * 2 * N3));
    HIPCHK(hipMalloc(&cl->wpart, sizeof(float) * (size_t)cl->wgrad_wg * CL_TAPS * 256));
    HIPCHK(hipMalloc(&cl->bpart, sizeof(float) * (size_t)cl->wgrad_wg * 16));
    HIPCHK(hipMalloc(&cl->lpart, sizeof(double) * nblocks(N3, 256)));
    HIPCHK(hipMalloc(&cl->red, sizeof(double) * 512));
    HIPCHK(hipMemset(cl->m, 0, sizeof(float) * P_TOTAL));
    HIPCHK(hipMemset(cl->v, 0, sizeof(float) * P_TOTAL));
    HIPCHK(hipFuncSetAttribute((const void *)k_wgrad<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)wgrad_lds(16, N)));
    HIPCHK(hipFuncSetAttribute((const void *)k_wgrad<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)wgrad_lds(4, N)));
    std::vector<float> P;
    init_params(P, seed);
    HIPCHK(hipMemcpy(cl->P, P.data(), sizeof(float) * P_TOTAL, hipMemcpyHostToDevice));
    CHK(repack(cl));
    HIPCHK(hipStreamSynchronize(cl->stream));
    return GABE_OK;
}

bool fp32_pair(const gabe_ctx *c) { return c->fp32 && c->nf == 2; }

/* restriction of one array of the fine context into one of the coarse */
/* restriction of the fine array src onto the coarse array dst; with
 * inverse_laplacian, of lap_coarse^-1 of it (zero mean) */
int restrict_array(gabe_ctx *f, gabe_ctx *c, void *src, void *dst, bool inverse_laplacian = false)
{
    CHK(run_plan(f, f->fft_rx, src, f->rx_buf));
    HIPCHK(hipStreamSynchronize(f->stream));
    const long Mc = (long)c->N * c->N * (c->N / 2 + 1);
    const double scale = 1.0 / (double)f->N3;
    const double inv_lap = inverse_laplacian ? c->dx * c->dx : 0.;
    if (c->fp32)
        hipLaunchKernelGGL(k_restrict_modes<float>, dim3(nblocks(Mc, 256)), dim3(256), 0, c->stream,
                           (const float *)f->rx_buf, f->N, (float *)c->rx_buf, c->N, (float)scale, inv_lap);
    else
        hipLaunchKernelGGL(k_restrict_modes<double>, dim3(nblocks(Mc, 256)), dim3(256), 0, c->stream,
                           (const double *)f->rx_buf, f->N, (double *)c->rx_buf, c->N, scale, inv_lap);
    HIPCHK(hipGetLastError());
    CHK(run_plan(c, c->fft_rx, c->rx_buf, dst));
    /* the coarse stream reads f->rx_buf, which the fine stream's transform of
     * the next array rewrites: nothing orders the two streams, so wait here */
    HIPCHK(hipStreamSynchronize(c->stream));
    return GABE_OK;
}

} // namespace

namespace {

/* (Re)attach a closure run to c from its state k and the packed weights and
 * biases at pk, bias (device or host memory, as kind says). */
int run_create(gabe_ctx *c, const gabe_detail::ClosureCkpt &k, const void *pk, const void *bias, hipMemcpyKind kind)
{
    if (c->closure) closure_release(c);
    ClosureRun *R = new ClosureRun;
    c->closure = R;
    R->N = k.N;
    R->thr = k.thr;
    for (int q = 0; q < 4; q++) R->inv_s[q] = k.inv_s[q];
    for (int q = 0; q < 2; q++) R->dt_s[q] = k.dt_s[q];
    const long N3 = c->N3;
    int rc = GABE_OK;
    if (hipMalloc(&R->pk, sizeof(bf16x8) * PK_TOTAL) != hipSuccess || hipMalloc(&R->bias, sizeof(float) * 48) ||
        hipMalloc(&R->a1, sizeof(bf16) * 16 * N3) ||
        hipMalloc(&R->a2, sizeof(bf16) * 16 * N3) || hipMalloc(&R->u, sizeof(float) * 2 * N3) ||
        hipMalloc(&R->gate, sizeof(int)) ||
        hipMalloc(&R->fallback, sizeof(long long)) || hipMalloc(&R->rmax, sizeof(double)))
        rc = fail(GABE_ERR_HIP, "closure attachment allocation failed");
    if (rc == GABE_OK &&
        (hipMemcpy(R->pk, pk, sizeof(bf16x8) * PK_TOTAL, kind) != hipSuccess ||
         hipMemcpy(R->bias, bias, sizeof(float) * 48, kind) != hipSuccess ||
         hipMemcpy(R->gate, &k.gate, sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(R->fallback, &k.fallback, sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(R->rmax, &k.rmax, sizeof(double), hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(GABE_ERR_HIP, "closure attachment copy failed");
    if (rc != GABE_OK) closure_release(c);
    return rc;
}

} // namespace

namespace gabe_detail {

int closure_after_step(gabe_ctx *c, hipEvent_t *ev)
{
    ClosureRun *R = c->closure;
    auto rec = [&](int k) -> int {
        if (ev) HIPCHK(hipEventRecord(ev[k], c->stream));
        return GABE_OK;
    };
    const float *f = (const float *)c->fbuf[c->cur], *v = (const float *)c->vbuf[c->cur];
    ConvArgs a = conv_base(R->N, nullptr, R->pk + PK_F1, R->bias);
    a.fin[0] = f;
    a.fin[1] = f + c->N3;
    a.fin[2] = v;
    a.fin[3] = v + c->N3;
    for (int q = 0; q < 4; q++) a.fscale[q] = R->inv_s[q];
    a.gatev = GateArgs{c->sc, c->hist, c->hist_cap, R->thr, R->gate, R->fallback, R->rmax};
    a.gate = R->gate;
    a.a_out = R->a1;
    CHK(rec(0));
    CHK((launch_conv<4, EPI_HIDDEN>(a, c->stream)));
    CHK(rec(1));
    a = conv_base(R->N, R->a1, R->pk + PK_F2, R->bias + 16);
    a.gate = R->gate;
    a.a_out = R->a2;
    CHK(rec(2));
    CHK((launch_conv<16, EPI_HIDDEN>(a, c->stream)));
    CHK(rec(3));
    a = conv_base(R->N, R->a2, R->pk + PK_F3, R->bias + 32);
    a.gate = R->gate;
    a.y_out = R->u;
    CHK(rec(4));
    CHK((launch_conv<16, EPI_OUT>(a, c->stream)));
    CHK(rec(5));
    float *vw = (float *)c->vbuf[c->cur];
    const int N = c->N;
    CHK(rec(6));
    const int TK = kick_tk(N), D = kick_depth(N);
    hipLaunchKernelGGL(k_lap2_kick, dim3((N / TK) * (N / (256 / TK)) * (N / D)), dim3(256), 0, c->stream, R->u, N, D,
                       vw, vw + c->N3, c->sc, R->dt_s[0], R->dt_s[1], R->gate);
    HIPCHK(hipGetLastError());
    CHK(rec(7));
    return GABE_OK;
}

void closure_release(gabe_ctx *c)
{
    ClosureRun *R = c->closure;
    if (!R) return;
    void *bufs[] = {R->pk, R->bias, R->a1, R->a2, R->u, R->gate, R->fallback, R->rmax};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete R;
    c->closure = nullptr;
}

size_t closure_ckpt_bytes(const gabe_ctx *c)
{
    return c->closure ? sizeof(ClosureCkpt) + sizeof(bf16x8) * PK_TOTAL + sizeof(float) * 48 : 0;
}

int closure_ckpt_save(gabe_ctx *c, char *p)
{
    ClosureRun *R = c->closure;
    if (!R) return GABE_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    ClosureCkpt k{};
    k.N = R->N;
    k.thr = R->thr;
    for (int q = 0; q < 4; q++) k.inv_s[q] = R->inv_s[q];
    for (int q = 0; q < 2; q++) k.dt_s[q] = R->dt_s[q];
    HIPCHK(hipMemcpy(&k.gate, R->gate, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&k.fallback, R->fallback, sizeof(long long), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&k.rmax, R->rmax, sizeof(double), hipMemcpyDeviceToHost));
    std::memcpy(p, &k, sizeof k);
    p += sizeof k;
    HIPCHK(hipMemcpy(p, R->pk, sizeof(bf16x8) * PK_TOTAL, hipMemcpyDeviceToHost));
    p += sizeof(bf16x8) * PK_TOTAL;
    HIPCHK(hipMemcpy(p, R->bias, sizeof(float) * 48, hipMemcpyDeviceToHost));
    return GABE_OK;
}

int closure_ckpt_load(gabe_ctx *c, const char *p, size_t bytes)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    if (bytes == 0) {   /* the checkpointed lattice had no closure attached */
        closure_release(c);
        return GABE_OK;
    }
    if (bytes != sizeof(ClosureCkpt) + sizeof(bf16x8) * PK_TOTAL + sizeof(float) * 48)
        return fail(GABE_ERR_ARG, "checkpoint's closure section has %zu bytes, this library writes %zu", bytes,
                    sizeof(ClosureCkpt) + sizeof(bf16x8) * PK_TOTAL + sizeof(float) * 48);
    ClosureCkpt k;
    std::memcpy(&k, p, sizeof k);
    if (k.N != c->N || !fp32_pair(c) || !c->own_stream)
        return fail(GABE_ERR_ARG, "checkpoint carries a closure for N=%d that this lattice cannot take", k.N);
    return run_create(c, k, p + sizeof k, p + sizeof k + sizeof(bf16x8) * PK_TOTAL, hipMemcpyHostToDevice);
}

} // namespace gabe_detail

/* ====================================================================== ABI */
namespace {

/* checks and the transforms of a fine -> coarse restriction */
int restrict_setup(gabe_ctx *fine, gabe_ctx *coarse)
{
    if (fine->nf != coarse->nf || fine->fp32 != coarse->fp32 || fine->p.L != coarse->p.L)
        return fail(GABE_ERR_ARG, "restriction needs the same nflds, precision and box size L");
    if (coarse->N > fine->N || (coarse->N & 1)) return fail(GABE_ERR_ARG, "restriction needs an even coarse N <= fine N");
    const bool dbl = !fine->fp32;
    const size_t cs = 2 * fine->esz;
    if (!fine->fft_rx.plan) {
        CHK(make_plan(fine, fine->fft_rx, rocfft_transform_type_real_forward, dbl, 1));
        HIPCHK(hipMalloc(&fine->rx_buf, cs * fine->M));
    }
    if (!coarse->fft_rx.plan) {
        CHK(make_plan(coarse, coarse->fft_rx, rocfft_transform_type_real_inverse, dbl, 1));
        HIPCHK(hipMalloc(&coarse->rx_buf, cs * coarse->M));
    }
    return GABE_OK;
}

} // namespace

extern "C" {

int gabe_restrict(gabe_ctx *fine, gabe_ctx *coarse)
{
    if (!fine || !coarse) return fail(GABE_ERR_ARG, "NULL context");
    CHK(restrict_setup(fine, coarse));
    HIPCHK(hipStreamSynchronize(fine->stream));
    HIPCHK(hipStreamSynchronize(coarse->stream));
    for (int a = 0; a < 2; a++) {
        char *src = (char *)(a ? fine->vbuf[fine->cur] : fine->fbuf[fine->cur]);
        char *dst = (char *)(a ? coarse->vbuf[coarse->cur] : coarse->fbuf[coarse->cur]);
        for (int f = 0; f < fine->nf; f++)
            CHK(restrict_array(fine, coarse, src + fine->esz * f * fine->N3, dst + coarse->esz * f * coarse->N3));
    }
    HIPCHK(hipStreamSynchronize(coarse->stream));
    DevScalars d;
    CHK(read_scalars(fine, d));
    DevScalars dc;
    CHK(read_scalars(coarse, dc));
    d.nstep = dc.nstep;    /* the coarse context keeps its own step count; its history */
    d.hist0 = dc.nstep;    /* restarts with the restricted state */
    d.e0_valid = 0;
    CHK(write_scalars(coarse, d));
    return GABE_OK;
}

int gabe_closure_create(int N, double dt, unsigned seed, gabe_closure **out)
{
    if (!out) return fail(GABE_ERR_ARG, "out is NULL");
    *out = nullptr;
    CHK(cl_check_n(N));
    if (!(dt > 0.)) return fail(GABE_ERR_ARG, "dt must be positive");
    gabe_closure *cl = new gabe_closure;
    const int rc = create_closure(cl, N, dt, seed);
    if (rc != GABE_OK) {
        free_closure(cl);
        delete cl;
        return rc;
    }
    *out = cl;
    return GABE_OK;
}

int gabe_closure_destroy(gabe_closure *cl)
{
    if (!cl) return GABE_OK;
    free_closure(cl);
    delete cl;
    return GABE_OK;
}

int gabe_closure_params(gabe_closure *cl, float *out)
{
    if (!cl || !out) return fail(GABE_ERR_ARG, "NULL argument");
    HIPCHK(hipStreamSynchronize(cl->stream));
    HIPCHK(hipMemcpy(out, cl->P, sizeof(float) * P_TOTAL, hipMemcpyDeviceToHost));
    return GABE_OK;
}

int gabe_closure_set_params(gabe_closure *cl, const float *in)
{
    if (!cl || !in) return fail(GABE_ERR_ARG, "NULL argument");
    HIPCHK(hipStreamSynchronize(cl->stream));
    HIPCHK(hipMemcpy(cl->P, in, sizeof(float) * P_TOTAL, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(cl->m, 0, sizeof(float) * P_TOTAL));
    HIPCHK(hipMemset(cl->v, 0, sizeof(float) * P_TOTAL));
    cl->adam_t = 0;
    CHK(repack(cl));
    HIPCHK(hipStreamSynchronize(cl->stream));
    return GABE_OK;
}

int gabe_closure_scales(gabe_closure *cl, double *s_in, double *s_out)
{
    if (!cl || !s_in || !s_out) return fail(GABE_ERR_ARG, "NULL argument");
    for (int c = 0; c < 4; c++) s_in[c] = cl->s_in[c];
    for (int c = 0; c < 2; c++) s_out[c] = cl->s_out[c];
    return GABE_OK;
}

int gabe_closure_set_scales(gabe_closure *cl, const double *s_in, const double *s_out)
{
    if (!cl || !s_in || !s_out) return fail(GABE_ERR_ARG, "NULL argument");
    for (int c = 0; c < 4; c++)
        if (!(s_in[c] > 0.)) return fail(GABE_ERR_ARG, "input scales must be positive");
    for (int c = 0; c < 2; c++)
        if (!(s_out[c] > 0.)) return fail(GABE_ERR_ARG, "output scales must be positive");
    for (int c = 0; c < 4; c++) cl->s_in[c] = s_in[c];
    for (int c = 0; c < 2; c++) cl->s_out[c] = s_out[c];
    cl->scaled = true;
    return GABE_OK;
}

int gabe_closure_sample(gabe_closure *cl, gabe_ctx *fine, gabe_ctx *coarse, int fine_steps, int restrict_first)
{
    if (!cl || !fine || !coarse) return fail(GABE_ERR_ARG, "NULL argument");
    if (coarse->N != cl->N || !fp32_pair(coarse) || !fp32_pair(fine))
        return fail(GABE_ERR_ARG, "samples need FP32 two-field contexts, the coarse one of the closure's N");
    if (fine_steps < 0) return fail(GABE_ERR_ARG, "fine_steps must be >= 0");
    CHK(restrict_setup(fine, coarse));
    if (restrict_first) CHK(gabe_restrict(fine, coarse));
    const long N3 = cl->N3, N3f = fine->N3;
    HIPCHK(hipStreamSynchronize(coarse->stream));
    HIPCHK(hipStreamSynchronize(fine->stream));
    /* input scales from the first sample: rms of each fluctuation array */
    const float *cf = (const float *)coarse->fbuf[coarse->cur], *cv = (const float *)coarse->vbuf[coarse->cur];
    if (!cl->scaled) {
        const float *src[4] = {cf, cf + N3, cv, cv + N3};
        for (int c = 0; c < 4; c++) {
            CHK(rms_of(cl, src[c], N3, &cl->s_in[c]));
            if (!(cl->s_in[c] > 0.)) return fail(GABE_ERR_STATE, "zero fluctuation in input channel %d", c);
        }
    }
    if (cl->flap_n3 != N3f) {
        if (cl->flap) HIPCHK(hipFree(cl->flap));
        cl->flap = nullptr;
        HIPCHK(hipMalloc(&cl->flap, sizeof(float) * 2 * N3f));
        cl->flap_n3 = N3f;
    }
    bf16 *x = nullptr;
    float *t = nullptr;
    HIPCHK(hipMalloc(&x, sizeof(bf16) * 4 * N3));
    if (hipMalloc(&t, sizeof(float) * 2 * N3) != hipSuccess) {
        (void)hipFree(x);
        return fail(GABE_ERR_HIP, "sample allocation failed");
    }
    cl->xs.push_back(x);
    cl->ts.push_back(t);
    hipLaunchKernelGGL(k_make_x, dim3(nblocks(N3, 256)), dim3(256), 0, cl->stream, cf, cf + N3, cv, cv + N3,
                       (float)(1. / cl->s_in[0]), (float)(1. / cl->s_in[1]), (float)(1. / cl->s_in[2]),
                       (float)(1. / cl->s_in[3]), N3, x, (const int *)nullptr);
    HIPCHK(hipGetLastError());
    /* k_make_x reads the coarse state on the closure's stream; the coarse
     * stream rewrites it below (gabe_restrict after the fine steps) */
    HIPCHK(hipStreamSynchronize(cl->stream));
    /* target: restriction of the fine lattice's Laplacian minus the coarse
     * lattice's Laplacian of the restricted field */
    hipLaunchKernelGGL(k_lap_field, dim3(nblocks(N3f, 256)), dim3(256), 0, fine->stream,
                       (const float *)fine->fbuf[fine->cur], fine->N, (float)(1. / (fine->dx * fine->dx)), cl->flap);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(fine->stream));
    if (!cl->scaled) {   /* output scales: rms of the first raw force target */
        for (int c = 0; c < 2; c++) CHK(restrict_array(fine, coarse, cl->flap + c * N3f, t + c * N3));
        hipLaunchKernelGGL(k_lap_target, dim3(nblocks(N3, 256)), dim3(256), 0, coarse->stream, cf, cl->N,
                           (float)(1. / (coarse->dx * coarse->dx)), 1.f, 1.f, t);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(coarse->stream));
        for (int c = 0; c < 2; c++) {
            CHK(rms_of(cl, t + c * N3, N3, &cl->s_out[c]));
            if (!(cl->s_out[c] > 0.)) return fail(GABE_ERR_STATE, "zero target in output channel %d", c);
        }
        HIPCHK(hipStreamSynchronize(cl->stream));
        cl->scaled = true;
    }
    /* target of w = lap u (unit 7-point stencil): lap_unit^-1 (force target / s_out)
     * = (lap_c^-1 R(lap_fine f) - R f) / (dx_c^2 s_out), the restriction
     * carrying lap_c^-1 in Fourier space */
    for (int c = 0; c < 2; c++) CHK(restrict_array(fine, coarse, cl->flap + c * N3f, t + c * N3, true));
    const double dx2 = coarse->dx * coarse->dx;
    hipLaunchKernelGGL(k_sub_scale2, dim3(nblocks(N3, 256)), dim3(256), 0, coarse->stream, cf, N3,
                       (float)(1. / (dx2 * cl->s_out[0])), (float)(1. / (dx2 * cl->s_out[1])), t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(coarse->stream));
    /* advance the fine lattice; the coarse one takes its restriction */
    if (fine_steps > 0) {
        CHK(gabe_step(fine, fine_steps));
        CHK(gabe_restrict(fine, coarse));
    }
    return (int)cl->xs.size() - 1;
}

int gabe_closure_num_samples(gabe_closure *cl)
{
    if (!cl) return fail(GABE_ERR_ARG, "NULL closure");
    return (int)cl->xs.size();
}

int gabe_closure_train(gabe_closure *cl, int iters, double lr, unsigned seed, double *loss_out)
{
    if (!cl || iters < 0 || !(lr > 0.)) return fail(GABE_ERR_ARG, "bad argument");
    const int ns = (int)cl->xs.size();
    if (ns == 0) return fail(GABE_ERR_STATE, "no samples");
    std::vector<int> order(ns);
    unsigned long long st = 0x2545F4914F6CDD1DULL ^ seed;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < iters; it++) {
        if (it % ns == 0) {   /* a new shuffle every pass over the samples */
            for (int q = 0; q < ns; q++) order[q] = q;
            for (int q = ns - 1; q > 0; q--) {
                st = st * 6364136223846793005ULL + 1442695040888963407ULL;
                const int p = (int)((st >> 33) % (unsigned long long)(q + 1));
                std::swap(order[q], order[p]);
            }
        }
        const int smp = order[it % ns];
        CHK(forward_loss(cl, cl->xs[smp], cl->ts[smp]));
        CHK(backward(cl, cl->xs[smp]));
        cl->adam_t++;
        const float c1 = 1.f - std::pow(b1, (float)cl->adam_t), c2 = 1.f - std::pow(b2, (float)cl->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(nblocks(P_TOTAL, 256)), dim3(256), 0, cl->stream, cl->P, cl->G, cl->m, cl->v,
                           P_TOTAL, (float)lr, b1, b2, eps, c1, c2);
        HIPCHK(hipGetLastError());
        CHK(repack(cl));
        if (loss_out) CHK(read_loss(cl, &loss_out[it]));
    }
    HIPCHK(hipStreamSynchronize(cl->stream));
    return GABE_OK;
}

int gabe_closure_loss(gabe_closure *cl, int sample, double *loss)
{
    if (!cl || !loss) return fail(GABE_ERR_ARG, "NULL argument");
    if (sample < 0 || sample >= (int)cl->xs.size()) return fail(GABE_ERR_ARG, "no sample %d", sample);
    CHK(forward_loss(cl, cl->xs[sample], cl->ts[sample]));
    return read_loss(cl, loss);
}

int gabe_closure_forward(gabe_closure *cl, const float *x, float *y)
{
    if (!cl || !x || !y) return fail(GABE_ERR_ARG, "NULL argument");
    const long N3 = cl->N3;
    std::vector<bf16> xb((size_t)4 * N3);
    for (long q = 0; q < 4 * N3; q++) xb[q] = (bf16)x[q];
    HIPCHK(hipMemcpy(cl->xbuf, xb.data(), sizeof(bf16) * 4 * N3, hipMemcpyHostToDevice));
    ConvArgs a = conv_base(cl->N, cl->xbuf, cl->pk + PK_F1, cl->bias);
    a.a_out = cl->a1;
    CHK((launch_conv<4, EPI_HIDDEN>(a, cl->stream)));
    a = conv_base(cl->N, cl->a1, cl->pk + PK_F2, cl->bias + 16);
    a.a_out = cl->a2;
    CHK((launch_conv<16, EPI_HIDDEN>(a, cl->stream)));
    a = conv_base(cl->N, cl->a2, cl->pk + PK_F3, cl->bias + 32);
    a.y_out = cl->ybuf;
    CHK((launch_conv<16, EPI_OUT>(a, cl->stream)));
    hipLaunchKernelGGL(k_lap_out, dim3(nblocks(N3, 256)), dim3(256), 0, cl->stream, cl->ybuf, cl->N, cl->wbuf);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_lap_out, dim3(nblocks(N3, 256)), dim3(256), 0, cl->stream, cl->wbuf, cl->N, cl->gybuf);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, cl->gybuf, sizeof(float) * 2 * N3, hipMemcpyDeviceToHost, cl->stream));
    HIPCHK(hipStreamSynchronize(cl->stream));
    return GABE_OK;
}

int gabe_closure_gradient(gabe_closure *cl, const float *x, const float *t, float *grad, double *loss)
{
    if (!cl || !x || !t || !grad || !loss) return fail(GABE_ERR_ARG, "NULL argument");
    const long N3 = cl->N3;
    std::vector<bf16> xb((size_t)4 * N3);
    for (long q = 0; q < 4 * N3; q++) xb[q] = (bf16)x[q];
    HIPCHK(hipMemcpy(cl->xbuf, xb.data(), sizeof(bf16) * 4 * N3, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(cl->tbuf, t, sizeof(float) * 2 * N3, hipMemcpyHostToDevice));
    CHK(forward_loss(cl, cl->xbuf, cl->tbuf));
    CHK(backward(cl, cl->xbuf));
    CHK(read_loss(cl, loss));
    HIPCHK(hipMemcpy(grad, cl->G, sizeof(float) * P_TOTAL, hipMemcpyDeviceToHost));
    return GABE_OK;
}

int gabe_closure_attach(gabe_ctx *c, gabe_closure *cl, double residual_threshold)
{
    if (!c || !cl) return fail(GABE_ERR_ARG, "NULL argument");
    if (c->N != cl->N || !fp32_pair(c)) return fail(GABE_ERR_ARG, "the closure needs an FP32 two-field lattice of its N");
    if (!c->own_stream) return fail(GABE_ERR_ARG, "ensemble members step in batches: attach to a single lattice");
    if (!cl->scaled) return fail(GABE_ERR_STATE, "closure has no input/output scales (no samples, no set_scales)");
    if (!(residual_threshold > 0.)) return fail(GABE_ERR_ARG, "residual threshold must be positive");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(cl->stream));
    ClosureCkpt k{};
    k.N = cl->N;
    k.gate = 1;
    k.fallback = -1;
    k.rmax = 0.;
    k.thr = residual_threshold;
    for (int q = 0; q < 4; q++) k.inv_s[q] = (float)(1. / cl->s_in[q]);
    for (int q = 0; q < 2; q++) k.dt_s[q] = (float)(c->p.dt * cl->s_out[q]);
    return run_create(c, k, cl->pk, cl->bias, hipMemcpyDeviceToDevice);
}

int gabe_closure_detach(gabe_ctx *c)
{
    if (!c) return fail(GABE_ERR_ARG, "NULL context");
    HIPCHK(hipStreamSynchronize(c->stream));
    closure_release(c);
    return GABE_OK;
}

int gabe_closure_status(gabe_ctx *c, int *active, long long *fallback_step, double *max_residual)
{
    if (!c) return fail(GABE_ERR_ARG, "NULL context");
    HIPCHK(hipStreamSynchronize(c->stream));
    int a = 0;
    long long fb = -1;
    double rm = 0.;
    if (c->closure) {
        HIPCHK(hipMemcpy(&a, c->closure->gate, sizeof a, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&fb, c->closure->fallback, sizeof fb, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&rm, c->closure->rmax, sizeof rm, hipMemcpyDeviceToHost));
    }
    if (active) *active = a;
    if (fallback_step) *fallback_step = fb;
    if (max_residual) *max_residual = rm;
    return GABE_OK;
}

} // extern "C"
