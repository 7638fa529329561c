/*
 * copy_probe.hip -- plain device-to-device copy of a known number of bytes.
 *
 * Used to calibrate HBM counters and bandwidth: under rocprofv3 --pmc the
 * FETCH_SIZE / WRITE_SIZE it reports per dispatch are compared with the exact
 * bytes this kernel must move (one read and one write of the buffer).
 * Usage: copy_probe [MiB] [iterations] [mode 0|1|2]; prints MiB, GB/s (HIP events).
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

/* mode 0: grid-stride loop; mode 1: one float4 per thread; mode 2: four
 * float4 per thread, loads issued before stores */
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_copy1(const float4 *__restrict__ src, float4 *__restrict__ dst, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_copy4(const float4 *__restrict__ src, float4 *__restrict__ dst, long n)
{
    const long i0 = (long)blockIdx.x * 1024 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (i0 + u * 256 < n) v[u] = src[i0 + u * 256];
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (i0 + u * 256 < n) dst[i0 + u * 256] = v[u];
}

#define CK(x)                                                                          \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

int main(int argc, char **argv)
{
    const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
    const int iters = argc > 2 ? std::atoi(argv[2]) : 20;
    const int mode = argc > 3 ? std::atoi(argv[3]) : 0;
    if (mib <= 0 || mib > 16384 || iters <= 0 || iters > 10000) {
        std::fprintf(stderr, "usage: copy_probe [MiB 1..16384] [iterations 1..10000]\n");
        return 2;
    }
    const long bytes = mib << 20, n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    CK(hipMemset(a, 1, bytes));
    CK(hipMemset(b, 0, bytes));
    auto launch = [&]() {
        if (mode == 1) hipLaunchKernelGGL(k_copy1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n);
        else if (mode == 2) hipLaunchKernelGGL(k_copy4, dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, a, b, n);
        else hipLaunchKernelGGL(k_copy, dim3(256 * 8 * 4), dim3(256), 0, 0, a, b, n);
    };
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    launch();
    CK(hipGetLastError());
    CK(hipEventRecord(e0, 0));
    for (int it = 0; it < iters; it++) launch();
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / iters;
    std::printf("copy_probe: mode %d, %ld MiB, %.3f ms per copy, %.1f GB/s (read + write)\n", mode, mib, per,
                2.0 * bytes / (per * 1e-3) / 1e9);
    CK(hipFree(a));
    CK(hipFree(b));
    return 0;
}
