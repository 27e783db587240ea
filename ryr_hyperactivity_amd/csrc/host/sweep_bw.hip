/*
 * sweep_bw - the memory access pattern of the sub-voxel sweep (k_cyto) with
 * almost no arithmetic: the bandwidth that pattern can reach on this GPU,
 * and what the stencil-partner loads cost in it.
 *   sweep_bw [chunks of 256 slots (default 12000, the 256-cell batch)]
 *
 * One lane per (plane ii, slot g) of 8 planes of P = 256 * chunks slots, as in
 * the sweep: ci (f32) read from one buffer and written to the other, cnsr
 * (f64) likewise, five f32 buffers and one f64 buffer read and written in
 * place: 80 algorithmic bytes per lane.  Partners follow a 100x20x10 lattice
 * (y stride 100, z stride 2000 slots), clamped to the array instead of the
 * lattice's own boundary test.  Variants:
 *   streams            the 10 streams only;
 *   streams+partners   plus the 12 partner loads of the 7-point stencils;
 *   partners via LDS   a workgroup holds all 8 planes of 32 slots (a wave
 *                      holds two planes); the 3 partners in the same slot and
 *                      the x partner inside the span come from LDS after one
 *                      barrier, the y and z partners from memory;
 *   two slots a lane   streams+partners with each lane taking two adjacent
 *                      slots (8- and 16-byte loads and stores);
 *   split f64 pools    streams+partners with each f64 pool held as an f32
 *                      value plus a 16-bit residual in units of 2^-15 of the
 *                      value's ulp (two arrays, 6 bytes instead of 8): 68
 *                      algorithmic bytes per lane, the rate reported against
 *                      the same 80 so that the times compare directly.
 */
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

struct SweepBufs {
    const float *ci0;
    float *ci1;
    const double *cn0;
    double *cn1;
    float *b0, *b1, *b2, *b3, *b4;
    double *b5;
    long long P;
};

#define SY 100
#define SZ 2000

__device__ __forceinline__ long long clampi(long long q, long long hi) { return q < 0 ? 0 : (q > hi ? hi : q); }

__device__ __forceinline__ void finish(const SweepBufs &B, long long e, float cc, double cn, float s, double sn)
{
    const float v0 = B.b0[e], v1 = B.b1[e], v2 = B.b2[e], v3 = B.b3[e], v4 = B.b4[e];
    const double v5 = B.b5[e];
    const float d = 1e-3f * (s - 6.f * cc);
    B.ci1[e] = cc + d;
    B.cn1[e] = cn + 1e-3 * (sn - 6.0 * cn);
    B.b0[e] = v0 + d * v1;
    B.b1[e] = v1 + d * v2;
    B.b2[e] = v2 + d * v3;
    B.b3[e] = v3 + d * v4;
    B.b4[e] = v4 + d * v0;
    B.b5[e] = v5 + (double)d;
}

/* grid: chunks * 16 workgroups of 128 (planes fastest), like k_cyto at 128 */
template <bool PARTNERS>
__global__ __launch_bounds__(128) void k_sweep(SweepBufs B)
{
    const int L = blockIdx.x;
    const int b = L >> 4, r = L & 15, ii = r & 7, part = r >> 3;
    const long long P = B.P, hi = P - 1;
    const long long g = (long long)b * 256 + part * 128 + threadIdx.x;
    const long long e = ii * P + g;
    const float cc = B.ci0[e];
    const double cn = B.cn0[e];
    float s = 6.f * cc;
    double sn = 6.0 * cn;
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

/* streams+partners, two adjacent slots per lane; grid: chunks * 8 workgroups
 * of 128 (a workgroup is one plane of one chunk) */
__global__ __launch_bounds__(128) void k_sweep2(SweepBufs B)
{
    const int L = blockIdx.x;
    const int b = L >> 3, ii = L & 7;
    const long long P = B.P, hi = P - 2;
    const long long g = (long long)b * 256 + 2 * threadIdx.x;
    const long long e = ii * P + g;
    const float2 cc = *(const float2 *)(B.ci0 + e);
    const double2 cn = *(const double2 *)(B.cn0 + e);
    const long long oz = (ii & 1) ? SZ : -SZ, oy = (ii & 2) ? SY : -SY, ox = (ii & 4) ? 1 : -1;
    const long long pz = (ii ^ 1) * P, py = (ii ^ 2) * P, px = (ii ^ 4) * P;
    const long long gz = clampi(g + oz, hi), gy = clampi(g + oy, hi);
    const float2 a0 = *(const float2 *)(B.ci0 + pz + g), a1 = *(const float2 *)(B.ci0 + pz + gz);
    const float2 a2 = *(const float2 *)(B.ci0 + py + g), a3 = *(const float2 *)(B.ci0 + py + gy);
    const float2 a4 = *(const float2 *)(B.ci0 + px + g);
    const double2 n0 = *(const double2 *)(B.cn0 + pz + g), n1 = *(const double2 *)(B.cn0 + pz + gz);
    const double2 n2 = *(const double2 *)(B.cn0 + py + g), n3 = *(const double2 *)(B.cn0 + py + gy);
    const double2 n4 = *(const double2 *)(B.cn0 + px + g);
    /* x partners: g - 1 / g + 2 outside the pair, the pair's other slot inside */
    const long long xo = clampi(ox < 0 ? g - 1 : g + 2, P - 1);
    const float ax = B.ci0[px + xo];
    const double nx = B.cn0[px + xo];
    float s[2], sn_f[2];
    double sn[2];
    s[0] = a0.x + a1.x + a2.x + a3.x + a4.x + (ox < 0 ? ax : a4.y);
    s[1] = a0.y + a1.y + a2.y + a3.y + a4.y + (ox < 0 ? a4.x : ax);
    sn[0] = n0.x + n1.x + n2.x + n3.x + n4.x + (ox < 0 ? nx : n4.y);
    sn[1] = n0.y + n1.y + n2.y + n3.y + n4.y + (ox < 0 ? n4.x : nx);
    (void)sn_f;
    const float2 v0 = *(const float2 *)(B.b0 + e), v1 = *(const float2 *)(B.b1 + e), v2 = *(const float2 *)(B.b2 + e);
    const float2 v3 = *(const float2 *)(B.b3 + e), v4 = *(const float2 *)(B.b4 + e);
    const double2 v5 = *(const double2 *)(B.b5 + e);
    const float d0 = 1e-3f * (s[0] - 6.f * cc.x), d1 = 1e-3f * (s[1] - 6.f * cc.y);
    *(float2 *)(B.ci1 + e) = make_float2(cc.x + d0, cc.y + d1);
    *(double2 *)(B.cn1 + e) = make_double2(cn.x + 1e-3 * (sn[0] - 6.0 * cn.x), cn.y + 1e-3 * (sn[1] - 6.0 * cn.y));
    *(float2 *)(B.b0 + e) = make_float2(v0.x + d0 * v1.x, v0.y + d1 * v1.y);
    *(float2 *)(B.b1 + e) = make_float2(v1.x + d0 * v2.x, v1.y + d1 * v2.y);
    *(float2 *)(B.b2 + e) = make_float2(v2.x + d0 * v3.x, v2.y + d1 * v3.y);
    *(float2 *)(B.b3 + e) = make_float2(v3.x + d0 * v4.x, v3.y + d1 * v4.y);
    *(float2 *)(B.b4 + e) = make_float2(v4.x + d0 * v0.x, v4.y + d1 * v0.y);
    *(double2 *)(B.b5 + e) = make_double2(v5.x + (double)d0, v5.y + (double)d1);
}

/* An f64 pool as f32 + 16-bit residual: x = h + q * ulp(h) * 2^-15, with
 * |x - h| <= ulp(h) / 2 after rounding to nearest, so |q| <= 2^14. */
__device__ __forceinline__ int ulp_exp(float h)
{
    const int ex = (int)((__float_as_uint(h) >> 23) & 255u);
    return ex < 1 ? 1 : ex;                     /* ulp(h) = 2^(ex - 150) */
}
__device__ __forceinline__ double split_ld(const float *h, const short *l, long long e)
{
    const float hv = h[e];
    return (double)hv + __builtin_ldexp((double)l[e], ulp_exp(hv) - 165);
}
__device__ __forceinline__ void split_st(float *h, short *l, long long e, double x)
{
    const float hv = (float)x;
    double q = __builtin_rint(__builtin_ldexp(x - (double)hv, 165 - ulp_exp(hv)));
    q = q > 32767.0 ? 32767.0 : (q < -32767.0 ? -32767.0 : q);
    h[e] = hv;
    l[e] = (short)q;
}

struct SplitBufs {
    const float *ci0;
    float *ci1;
    const float *cn0h;
    const short *cn0l;
    float *cn1h;
    short *cn1l;
    float *b0, *b1, *b2, *b3, *b4, *b5h;
    short *b5l;
    long long P;
};

/* streams+partners with the split pools (cnsr and the f64 buffer) */
__global__ __launch_bounds__(128) void k_sweep_split(SplitBufs B)
{
    const int L = blockIdx.x;
    const int b = L >> 4, r = L & 15, ii = r & 7, part = r >> 3;
    const long long P = B.P, hi = P - 1;
    const long long g = (long long)b * 256 + part * 128 + threadIdx.x;
    const long long e = ii * P + g;
    const float cc = B.ci0[e];
    const double cn = split_ld(B.cn0h, B.cn0l, e);
    const long long oz = (ii & 1) ? SZ : -SZ, oy = (ii & 2) ? SY : -SY, ox = (ii & 4) ? 1 : -1;
    const long long pz = (ii ^ 1) * P, py = (ii ^ 2) * P, px = (ii ^ 4) * P;
    const float s = B.ci0[pz + g] + B.ci0[pz + clampi(g + oz, hi)] + B.ci0[py + g] + B.ci0[py + clampi(g + oy, hi)] +
                    B.ci0[px + g] + B.ci0[px + clampi(g + ox, hi)];
    const double sn = split_ld(B.cn0h, B.cn0l, pz + g) + split_ld(B.cn0h, B.cn0l, pz + clampi(g + oz, hi)) +
                      split_ld(B.cn0h, B.cn0l, py + g) + split_ld(B.cn0h, B.cn0l, py + clampi(g + oy, hi)) +
                      split_ld(B.cn0h, B.cn0l, px + g) + split_ld(B.cn0h, B.cn0l, px + clampi(g + ox, hi));
    const float v0 = B.b0[e], v1 = B.b1[e], v2 = B.b2[e], v3 = B.b3[e], v4 = B.b4[e];
    const double v5 = split_ld(B.b5h, B.b5l, e);
    const float d = 1e-3f * (s - 6.f * cc);
    B.ci1[e] = cc + d;
    split_st(B.cn1h, B.cn1l, e, cn + 1e-3 * (sn - 6.0 * cn));
    B.b0[e] = v0 + d * v1;
    B.b1[e] = v1 + d * v2;
    B.b2[e] = v2 + d * v3;
    B.b3[e] = v3 + d * v4;
    B.b4[e] = v4 + d * v0;
    split_st(B.b5h, B.b5l, e, v5 + (double)d);
}

/* grid: chunks * 8 workgroups of 256; workgroup L holds slots 32L..32L+31 of all 8 planes */
__global__ __launch_bounds__(256) void k_sweep_lds(SweepBufs B)
{
    __shared__ float sci[8][32];
    __shared__ double scn[8][32];
    const int ii = (int)threadIdx.x >> 5, c = (int)threadIdx.x & 31;
    const long long P = B.P, hi = P - 1;
    const long long g = (long long)blockIdx.x * 32 + c;
    const long long e = ii * P + g;
    const float cc = B.ci0[e];
    const double cn = B.cn0[e];
    sci[ii][c] = cc;
    scn[ii][c] = cn;
    const long long oz = (ii & 1) ? SZ : -SZ, oy = (ii & 2) ? SY : -SY, ox = (ii & 4) ? 1 : -1;
    const long long pz = (ii ^ 1) * P, py = (ii ^ 2) * P, px = (ii ^ 4) * P;
    const float gz = B.ci0[pz + clampi(g + oz, hi)], gy = B.ci0[py + clampi(g + oy, hi)];
    const double hz = B.cn0[pz + clampi(g + oz, hi)], hy = B.cn0[py + clampi(g + oy, hi)];
    const int cx = c + (int)ox;
    const bool xin = cx >= 0 && cx < 32;
    float gx = 0.f;
    double hx = 0.0;
    if (!xin) {
        gx = B.ci0[px + clampi(g + ox, hi)];
        hx = B.cn0[px + clampi(g + ox, hi)];
    }
    __syncthreads();
    if (xin) {
        gx = sci[ii ^ 4][cx];
        hx = scn[ii ^ 4][cx];
    }
    const float s = sci[ii ^ 1][c] + gz + sci[ii ^ 2][c] + gy + sci[ii ^ 4][c] + gx;
    const double sn = scn[ii ^ 1][c] + hz + scn[ii ^ 2][c] + hy + scn[ii ^ 4][c] + hx;
    finish(B, e, cc, cn, s, sn);
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
    std::printf("sweep_bw %-18s %9.1f us per sweep, %7.1f GB/s algorithmic\n", name, per * 1e6, bytes / per / 1e9);
    std::fflush(stdout);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

int main(int argc, char **argv)
{
    const long long nchunks = argc > 1 ? std::atoll(argv[1]) : 12000;
    if (nchunks < 8 || nchunks > 200000) { std::fprintf(stderr, "sweep_bw: chunks must be in [8, 200000]\n"); return 2; }
    const long long P = nchunks * 256, n8 = 8 * P;
    const size_t elt[10] = {4, 4, 8, 8, 4, 4, 4, 4, 4, 8};
    void *p[10] = {};
    for (int k = 0; k < 10; ++k) {
        if (hipMalloc(&p[k], n8 * elt[k]) != hipSuccess) { std::fprintf(stderr, "sweep_bw: hipMalloc failed\n"); return 1; }
        (void)hipMemset(p[k], 0, n8 * elt[k]);
    }
    SweepBufs B;
    B.ci0 = (const float *)p[0]; B.ci1 = (float *)p[1];
    B.cn0 = (const double *)p[2]; B.cn1 = (double *)p[3];
    B.b0 = (float *)p[4]; B.b1 = (float *)p[5]; B.b2 = (float *)p[6]; B.b3 = (float *)p[7]; B.b4 = (float *)p[8];
    B.b5 = (double *)p[9];
    B.P = P;
    const double bytes = 80.0 * (double)n8;
    std::printf("sweep_bw: %lld chunks, %lld slots per plane, %.3f GB algorithmic per sweep\n", nchunks, P, bytes / 1e9);
    const unsigned g128 = (unsigned)(nchunks * 16), g256 = (unsigned)(nchunks * 8);
    int rc = timeit("streams", bytes, [&] { hipLaunchKernelGGL(k_sweep<false>, dim3(g128), dim3(128), 0, 0, B); });
    if (!rc) rc = timeit("streams+partners", bytes, [&] { hipLaunchKernelGGL(k_sweep<true>, dim3(g128), dim3(128), 0, 0, B); });
    if (!rc) rc = timeit("partners via LDS", bytes, [&] { hipLaunchKernelGGL(k_sweep_lds, dim3(g256), dim3(256), 0, 0, B); });
    if (!rc) rc = timeit("two slots a lane", bytes, [&] { hipLaunchKernelGGL(k_sweep2, dim3(g256), dim3(128), 0, 0, B); });
    /* the split pools live in the first 6 bytes per element of the f64 allocations */
    SplitBufs S;
    S.ci0 = B.ci0; S.ci1 = B.ci1;
    S.cn0h = (const float *)p[2]; S.cn0l = (const short *)((const char *)p[2] + 4 * n8);
    S.cn1h = (float *)p[3]; S.cn1l = (short *)((char *)p[3] + 4 * n8);
    S.b0 = B.b0; S.b1 = B.b1; S.b2 = B.b2; S.b3 = B.b3; S.b4 = B.b4;
    S.b5h = (float *)p[9]; S.b5l = (short *)((char *)p[9] + 4 * n8);
    S.P = P;
    if (!rc) rc = timeit("split f64 pools", bytes, [&] { hipLaunchKernelGGL(k_sweep_split, dim3(g128), dim3(128), 0, 0, S); });
    for (int k = 0; k < 10; ++k) (void)hipFree(p[k]);
    return rc;
}
