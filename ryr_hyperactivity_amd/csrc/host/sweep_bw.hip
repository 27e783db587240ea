/*
 * sweep_bw - the memory access pattern of the sub-voxel sweep with almost no
 * arithmetic: the bandwidth each form of that pattern can reach on this GPU,
 * and the ceiling set by its number of streams.
 *   sweep_bw [chunks of 256 slots (default 12000, about the 256-cell batch)]
 *
 * Ten float arrays of 8 planes x P slots (P = 256 * chunks): ci read from one
 * array and written to another, cnsr likewise, six buffers read and written
 * in place; 64 bytes per (plane, slot).  Stencil partners follow a 100x20x10
 * lattice (y stride 100, z stride 2000 slots), clamped to the array instead
 * of the lattice's own boundary test.  Forms:
 *   copy10            a copy of ten arrays into ten others (the stream count's
 *                     ceiling, no partners);
 *   streams           one lane per (plane, slot), the ten streams only;
 *   sub-voxel         plus the 12 partner loads of the two 7-point stencils;
 *   lane per CRU      one lane per slot holding its 8 planes: partners in the
 *                     same slot from registers, x partners from the
 *                     neighbouring lanes (memory at a wave's two ends), y and z
 *                     partners loaded.
 * Rates are bytes moved (64 per plane-slot, 80 per element for copy10) over
 * the best of three 10-launch timings.
 */
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

struct Bufs {
    const float *ci0;
    float *ci1;
    const float *cn0;
    float *cn1;
    float *b[6];
    long long P;
};

#define SY 100
#define SZ 2000

__device__ __forceinline__ long long clampi(long long q, long long hi) { return q < 0 ? 0 : (q > hi ? hi : q); }

__device__ __forceinline__ void finish(const Bufs &B, long long e, float cc, float cn, float s, float sn)
{
    float v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = B.b[k][e];
    const float d = 1e-3f * (s - 6.f * cc);
    B.ci1[e] = cc + d;
    B.cn1[e] = cn + 1e-3f * (sn - 6.f * cn);
#pragma unroll
    for (int k = 0; k < 6; ++k) B.b[k][e] = v[k] + d * v[(k + 1) % 6];
}

/* one lane per (plane, slot); grid: chunks * 16 workgroups of 128, planes fastest */
template <bool PARTNERS>
__global__ __launch_bounds__(128) void k_sweep(Bufs B)
{
    const int L = blockIdx.x;
    const int b = L >> 4, r = L & 15, ii = r & 7, part = r >> 3;
    const long long P = B.P, hi = P - 1;
    const long long g = (long long)b * 256 + part * 128 + threadIdx.x;
    const long long e = ii * P + g;
    const float cc = B.ci0[e], cn = B.cn0[e];
    float s = 6.f * cc, sn = 6.f * cn;
    if (PARTNERS) {
        const long long oz = (ii & 1) ? SZ : -SZ, oy = (ii & 2) ? SY : -SY, ox = (ii & 4) ? 1 : -1;
        const long long pz = (ii ^ 1) * P, py = (ii ^ 2) * P, px = (ii ^ 4) * P;
        s = B.ci0[pz + g] + B.ci0[pz + clampi(g + oz, hi)] + B.ci0[py + g] + B.ci0[py + clampi(g + oy, hi)] +
            B.ci0[px + g] + B.ci0[px + clampi(g + ox, hi)];
        sn = B.cn0[pz + g] + B.cn0[pz + clampi(g + oz, hi)] + B.cn0[py + g] + B.cn0[py + clampi(g + oy, hi)] +
             B.cn0[px + g] + B.cn0[px + clampi(g + ox, hi)];
    }
    finish(B, e, cc, cn, s, sn);
}

/* one lane per slot, its 8 planes in registers; grid: chunks workgroups of 256 */
__global__ __launch_bounds__(256) void k_sweep_cru(Bufs B)
{
    const long long P = B.P, hi = P - 1;
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    float cc[8], cn[8];
#pragma unroll
    for (int ii = 0; ii < 8; ++ii) { cc[ii] = B.ci0[ii * P + g]; cn[ii] = B.cn0[ii * P + g]; }
#pragma unroll
    for (int ii = 0; ii < 8; ++ii) {
        const int dx = (ii & 4) ? 1 : -1;
        const int src = (lane + dx) & 63;
        float xc = __shfl(cc[ii ^ 4], src), xn = __shfl(cn[ii ^ 4], src);
        if (lane + dx < 0 || lane + dx > 63) {
            const long long gx = clampi(g + dx, hi);
            xc = B.ci0[(ii ^ 4) * P + gx];
            xn = B.cn0[(ii ^ 4) * P + gx];
        }
        const long long gz = clampi(g + ((ii & 1) ? SZ : -SZ), hi), gy = clampi(g + ((ii & 2) ? SY : -SY), hi);
        const float s = cc[ii ^ 1] + B.ci0[(ii ^ 1) * P + gz] + cc[ii ^ 2] + B.ci0[(ii ^ 2) * P + gy] + cc[ii ^ 4] + xc;
        const float sn = cn[ii ^ 1] + B.cn0[(ii ^ 1) * P + gz] + cn[ii ^ 2] + B.cn0[(ii ^ 2) * P + gy] + cn[ii ^ 4] + xn;
        finish(B, ii * P + g, cc[ii], cn[ii], s, sn);
    }
}

/* ten arrays copied into ten others, one element per lane */
__global__ __launch_bounds__(256) void k_copy10(const float *const *src, float *const *dst, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float v[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) v[k] = src[k][i];
#pragma unroll
    for (int k = 0; k < 10; ++k) dst[k][i] = v[k] + 1.f;
}

template <typename F>
static int timeit(const char *name, double bytes, F launch)
{
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    launch();
    if (hipDeviceSynchronize() != hipSuccess) { std::fprintf(stderr, "sweep_bw: %s failed\n", name); return 1; }
    float best = 1e30f;
    for (int t = 0; t < 3; ++t) {
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < 10; ++r) launch();
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) { std::fprintf(stderr, "sweep_bw: %s failed\n", name); return 1; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    const double per = best * 1e-3 / 10;
    std::printf("sweep_bw %-12s %9.1f us per launch, %7.1f GB/s\n", name, per * 1e6, bytes / per / 1e9);
    std::fflush(stdout);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

int main(int argc, char **argv)
{
    const long long nchunks = argc > 1 ? std::atoll(argv[1]) : 12000;
    if (nchunks < 8 || nchunks > 100000) { std::fprintf(stderr, "sweep_bw: chunks must be in [8, 100000]\n"); return 2; }
    const long long P = nchunks * 256, n8 = 8 * P;
    float *p[20] = {};
    for (int k = 0; k < 20; ++k) {
        if (hipMalloc(&p[k], n8 * sizeof(float)) != hipSuccess) { std::fprintf(stderr, "sweep_bw: hipMalloc failed\n"); return 1; }
        (void)hipMemset(p[k], 0, n8 * sizeof(float));
    }
    Bufs B;
    B.ci0 = p[0]; B.ci1 = p[1]; B.cn0 = p[2]; B.cn1 = p[3];
    for (int k = 0; k < 6; ++k) B.b[k] = p[4 + k];
    B.P = P;
    const double bytes = 64.0 * (double)n8;
    std::printf("sweep_bw: %lld chunks, %lld slots per plane, %.3f GB per sweep\n", nchunks, P, bytes / 1e9);
    float **dsrc = nullptr, **ddst = nullptr;
    if (hipMalloc(&dsrc, 10 * sizeof(float *)) != hipSuccess || hipMalloc(&ddst, 10 * sizeof(float *)) != hipSuccess) return 1;
    (void)hipMemcpy(dsrc, p, 10 * sizeof(float *), hipMemcpyHostToDevice);
    (void)hipMemcpy(ddst, p + 10, 10 * sizeof(float *), hipMemcpyHostToDevice);
    const unsigned g128 = (unsigned)(nchunks * 16), g256 = (unsigned)nchunks, gc = (unsigned)((n8 + 255) / 256);
    int rc = timeit("copy10", 80.0 * (double)n8, [&] { hipLaunchKernelGGL(k_copy10, dim3(gc), dim3(256), 0, 0, dsrc, ddst, n8); });
    if (!rc) rc = timeit("streams", bytes, [&] { hipLaunchKernelGGL(k_sweep<false>, dim3(g128), dim3(128), 0, 0, B); });
    if (!rc) rc = timeit("sub-voxel", bytes, [&] { hipLaunchKernelGGL(k_sweep<true>, dim3(g128), dim3(128), 0, 0, B); });
    if (!rc) rc = timeit("lane per CRU", bytes, [&] { hipLaunchKernelGGL(k_sweep_cru, dim3(g256), dim3(256), 0, 0, B); });
    for (int k = 0; k < 20; ++k) (void)hipFree(p[k]);
    (void)hipFree(dsrc);
    (void)hipFree(ddst);
    return rc;
}
