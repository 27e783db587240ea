/*
 * copy_bw - a plain device copy of a known size, used to calibrate the
 * rocprofv3 FETCH_SIZE / WRITE_SIZE counters (bytes the counters report per
 * byte actually moved) before they are applied to the CRU kernels, and to
 * measure the copy bandwidth this GPU reaches.
 *   copy_bw [MiB per buffer (default 1024)] [repeats (default 20)]
 *
 * k_copy moves one float4 per lane, one lane per element (no grid-stride
 * loop: on this GPU a grid-stride form with a fixed grid reaches ~25% less).
 * k_copy_nt is the same with nontemporal loads and stores, the fastest plain
 * stream measured here.
 */
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_copy_nt(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const float *s = (const float *)(src + i);
        float *d = (float *)(dst + i);
        const float x = __builtin_nontemporal_load(s), y = __builtin_nontemporal_load(s + 1);
        const float z = __builtin_nontemporal_load(s + 2), w = __builtin_nontemporal_load(s + 3);
        __builtin_nontemporal_store(x, d);
        __builtin_nontemporal_store(y, d + 1);
        __builtin_nontemporal_store(z, d + 2);
        __builtin_nontemporal_store(w, d + 3);
    }
}

template <typename K>
static int time_copy(const char *name, K kern, const float4 *a, float4 *b, long long n, long long bytes, int reps)
{
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) { std::fprintf(stderr, "copy_bw: kernel failed\n"); return 1; }
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) { std::fprintf(stderr, "copy_bw: kernel failed\n"); return 1; }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double per = ms * 1e-3 / reps;
    std::printf("copy_bw %s: %lld MiB read + %lld MiB written per launch, %.1f us per launch, %.1f GB/s\n", name,
                bytes >> 20, bytes >> 20, per * 1e6, 2.0 * (double)bytes / per / 1e9);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

int main(int argc, char **argv)
{
    const long long mib = argc > 1 ? std::atoll(argv[1]) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    if (mib < 1 || reps < 1) { std::fprintf(stderr, "copy_bw: MiB and repeats must be positive\n"); return 2; }
    const long long bytes = mib << 20;
    const long long n = bytes / (long long)sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, bytes) != hipSuccess || hipMalloc((void **)&b, bytes) != hipSuccess) {
        std::fprintf(stderr, "copy_bw: hipMalloc failed\n");
        return 1;
    }
    (void)hipMemset(a, 1, bytes);
    (void)hipMemset(b, 0, bytes);
    int rc = time_copy("k_copy", k_copy, a, b, n, bytes, reps);
    if (!rc) rc = time_copy("k_copy_nt", k_copy_nt, a, b, n, bytes, reps);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
