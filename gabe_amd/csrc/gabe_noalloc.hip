/*
 * gabe_noalloc.hip -- gabe_means_vars_noalloc: the mean and variance of
 * fields held in caller-owned device memory, with no device or host
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * The per-workgroup partial sums of the reduction (k_meanvar, the kernel of
 * gabe_means_vars) go to a file-scope device array of fixed size, `scratch`,
 * of which every GPU the library runs on holds its own copy until the process
 * ends.  Host code never passes the array's name to a kernel or a copy: the
 * name is a host-side stand-in, not a device address.  The device address is
 * looked up once per device with hipGetSymbolAddress (with that device
 * current) and kept per device.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (two host threads, or two streams in
 * flight) race.  The partials are fully rewritten by each call, so nothing
 * needs clearing between calls.
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "gabe_ctx.h"

using namespace gabe_detail;

namespace {

constexpr int kNoallocBlocks = 1024;              /* workgroups of the reduction, as gabe_means_vars */
constexpr int kScratchDoubles = 2 * kNoallocBlocks; /* [2][kNoallocBlocks]: sums of d and d^2 */
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

/* the current device's copy of `scratch` (its device address), looked up on
 * the first call on that device */
int scratch_ptr(double **out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(GABE_ERR_ARG, "device %d beyond the %d this library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GABE_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return GABE_OK;
}

} // namespace

extern "C" {

int gabe_means_vars_noalloc(const void *field, int precision, long long nsites, int nflds, const double *base,
                            void *stream, double *out)
{
    if (!field || !out) return fail(GABE_ERR_ARG, "NULL argument");
    if (precision != GABE_FP64 && precision != GABE_FP32) return fail(GABE_ERR_ARG, "precision must be GABE_FP64 or GABE_FP32");
    if (nflds < 1 || nflds > GABE_MAX_FLDS) return fail(GABE_ERR_ARG, "nflds must be 1..%d (got %d)", GABE_MAX_FLDS, nflds);
    if (nsites < 1 || nsites > GABE_NOALLOC_MAX_SITES)
        return fail(GABE_ERR_ARG, "nsites must be 1..%lld (got %lld)", (long long)GABE_NOALLOC_MAX_SITES, nsites);
    double *part = nullptr;
    CHK(scratch_ptr(&part));
    hipStream_t s = (hipStream_t)stream;
    double host_part[kScratchDoubles];
    for (int f = 0; f < nflds; f++) {
        double shift = 0.;
        float shift32 = 0.f;
        if (precision == GABE_FP32) {
            const float *x = (const float *)field + (long)f * nsites;
            launch_meanvar<float>(x, (long)nsites, part, kNoallocBlocks, s);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(&shift32, x, sizeof(float), hipMemcpyDeviceToHost, s));
        } else {
            const double *x = (const double *)field + (long)f * nsites;
            launch_meanvar<double>(x, (long)nsites, part, kNoallocBlocks, s);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(&shift, x, sizeof(double), hipMemcpyDeviceToHost, s));
        }
        HIPCHK(hipMemcpyAsync(host_part, part, sizeof host_part, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (precision == GABE_FP32) shift = (double)shift32;
        const double m = csum(host_part, kNoallocBlocks) / (double)nsites;
        out[2 * f] = (base ? base[f] : 0.) + shift + m;
        out[2 * f + 1] = csum(host_part + kNoallocBlocks, kNoallocBlocks) / (double)nsites - m * m;
    }
    return GABE_OK;
}

} // extern "C"
