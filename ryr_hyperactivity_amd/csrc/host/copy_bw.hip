/*
 * copy_bw - a plain device copy of a known size, used to calibrate the
 * rocprofv3 FETCH_SIZE / WRITE_SIZE counters (bytes the counters report per
 * byte actually moved) before they are applied to the CRU kernel, and to
 * measure the copy bandwidth this GPU reaches.
 *   copy_bw [MiB per buffer (default 1024)] [repeats (default 20)]
 */
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

int main(int argc, char **argv)
{
    const long long mib = argc > 1 ? std::atoll(argv[1]) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    const long long bytes = mib << 20;
    const long long n = bytes / (long long)sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, bytes) != hipSuccess || hipMalloc((void **)&b, bytes) != hipSuccess) {
        std::fprintf(stderr, "copy_bw: hipMalloc failed\n");
        return 1;
    }
    (void)hipMemset(a, 1, bytes);
    (void)hipMemset(b, 0, bytes);
    const int grid = 256 * 8 * 4;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    (void)hipDeviceSynchronize();
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) { std::fprintf(stderr, "copy_bw: kernel failed\n"); return 1; }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double per = ms * 1e-3 / reps;
    std::printf("copy_bw: %lld MiB read + %lld MiB written per launch, %.1f us per launch, %.1f GB/s\n",
                mib, mib, per * 1e6, 2.0 * (double)bytes / per / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
