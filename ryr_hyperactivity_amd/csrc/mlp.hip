/*
 * mlp.hip - the beat surrogate's training and inference kernels for gfx950
 * (C ABI in include/ryr_mlp.h; no counterpart in the reference).
 *
 * One wavefront per workgroup; a workgroup takes `tiles` consecutive tiles
 * of 16 samples.  Every product is a v_mfma_f32_16x16x32_bf16 (BF16
 * operands, FP32 accumulation) with the operand layout
 *   A (16 x 32):  lane l holds A[l%16][8*(l/16) + j], j = 0..7
 *   B (32 x 16):  lane l holds B[8*(l/16) + j][l%16]
 *   D (16 x 16):  lane l holds D[4*(l/16) + i][l%16], i = 0..3
 * Products whose reduction runs over the 16 samples of a tile (the weight
 * gradients) use K = 16: lanes 32..63 feed zeros.  Weights (BF16, natural
 * and transposed) and the tile's activations live in LDS; the D layout of
 * one layer is turned into the A layout of the next through LDS.
 *
 * Training: each workgroup writes its gradient (and loss) partial; one
 * thread per parameter then adds the partials in workgroup order (so the
 * result does not depend on scheduling) and applies Adam to the FP32 master
 * weights, refreshing their BF16 copy.
 */
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include "ryr_cru.h"
#include "ryr_mlp.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define NP RYR_MLP_NPARAM
#define OFF_W1 0
#define OFF_B1 2048
#define OFF_W2 2112
#define OFF_B2 6208
#define OFF_W3 6272
#define OFF_B3 7296
#define KIN 32      /* padded input width */
#define NH 64
#define NO 16       /* padded output width */

struct MlpArgs {
    const bf16 *wb;         /* packed parameters, BF16 */
    const float *wf;        /* packed parameters, FP32 (biases are read from here) */
    const bf16 *x;          /* samples, [n][32] BF16 */
    const int *idx;         /* sample of the j-th row (null: j itself) */
    const float *y, *t, *mask, *s;   /* labels, targets, target mask, sample weights (null: 1) */
    long long n;            /* rows in this launch */
    int tiles;              /* 16-row tiles per workgroup */
    float w_apd, inv_wsum;
    float *out;             /* forward: [n][2] */
    float *gpart;           /* training: per workgroup NP gradients + 1 loss */
};

struct __attribute__((aligned(16))) MlpLds {
    bf16 w1t[NH][KIN];      /* W1 transposed: [unit][input] */
    bf16 w2[NH][NH];        /* W2: [input unit][output unit] */
    bf16 w2t[NH][NH];
    bf16 w3[NH][NO];
    bf16 w3t[NO][NH];
    float b1[NH], b2[NH], b3[NO];
    bf16 x[16][KIN];
    bf16 a1[16][NH], a2[16][NH];
    bf16 d1[16][NH], d2[16][NH], d3[16][NO];
};

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

/* tanh-form GELU and its derivative */
__device__ __forceinline__ void gelu(float x, float *g, float *dg)
{
    const float c0 = 0.7978845608028654f, c1 = 0.044715f;
    const float u = c0 * (x + c1 * x * x * x);
    const float th = 1.0f - 2.0f / (fexp(2.0f * u) + 1.0f);
    *g = 0.5f * x * (1.0f + th);
    *dg = 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * c0 * (1.0f + 3.0f * c1 * x * x);
}

/* A operand from a row-major [16][ld] LDS array, columns k0 + 8*(l/16) .. +7 */
__device__ __forceinline__ bf16x8 a_rows(const bf16 *base, int ld, int k0, int lane)
{
    return *(const bf16x8 *)(base + (lane & 15) * ld + k0 + 8 * (lane >> 4));
}
/* B operand whose column c is a contiguous row of `base` ([c][k] layout) */
__device__ __forceinline__ bf16x8 b_cols(const bf16 *base, int ld, int c0, int k0, int lane)
{
    return *(const bf16x8 *)(base + (c0 + (lane & 15)) * ld + k0 + 8 * (lane >> 4));
}
/* K = 16 operands over the samples of a tile: element j of lane l is
 * M[8*(l/16) + j][c0 + l%16] for l < 32 and 0 for the upper lanes */
__device__ __forceinline__ bf16x8 k16_cols(const bf16 *base, int ld, int c0, int lane)
{
    bf16x8 v;
    const int kb = 8 * (lane >> 4);
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = lane < 32 ? base[(kb + j) * ld + c0 + (lane & 15)] : (bf16)0.0f;
    return v;
}

template <bool TRAIN>
__global__ __launch_bounds__(64) void k_mlp(MlpArgs a)
{
    __shared__ MlpLds L;
    const int lane = threadIdx.x;
    /* stage the weights */
    for (int e = lane; e < KIN * NH; e += 64) { const int k = e / NH, j = e % NH; L.w1t[j][k] = a.wb[OFF_W1 + e]; }
    for (int e = lane; e < NH * NH; e += 64) {
        const int k = e / NH, j = e % NH;
        const bf16 w = a.wb[OFF_W2 + e];
        L.w2[k][j] = w;
        L.w2t[j][k] = w;
    }
    for (int e = lane; e < NH * NO; e += 64) {
        const int k = e / NO, j = e % NO;
        const bf16 w = a.wb[OFF_W3 + e];
        L.w3[k][j] = w;
        L.w3t[j][k] = w;
    }
    L.b1[lane] = a.wf[OFF_B1 + lane];
    L.b2[lane] = a.wf[OFF_B2 + lane];
    if (lane < NO) L.b3[lane] = a.wf[OFF_B3 + lane];
    __syncthreads();

    const int rq = 4 * (lane >> 4), cl = lane & 15;
    f32x4 gw1[2][4], gw2[4][4], gw3[4];
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3 = 0.0f, loss = 0.0f;
    if (TRAIN) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            gw3[m] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < 4; ++t) gw2[m][t] = f32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int t = 0; t < 4; ++t) gw1[m][t] = f32x4{0, 0, 0, 0};
    }

    for (int tile = 0; tile < a.tiles; ++tile) {
        const long long r0 = ((long long)blockIdx.x * a.tiles + tile) * 16;
        if (r0 >= a.n) break;
        /* the tile's inputs (rows past n are zero) */
        {
            const int row = lane >> 2, q = lane & 3;          /* 16 rows x 4 chunks of 8 */
            const long long r = r0 + row;
            bf16x8 v;
            if (r < a.n) {
                const long long src = a.idx ? a.idx[r] : r;
                v = *(const bf16x8 *)(a.x + src * KIN + 8 * q);
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = (bf16)0.0f;
            }
            *(bf16x8 *)&L.x[row][8 * q] = v;
        }
        __syncthreads();
        /* layer 1: Z1 = X W1 + b1 (K = 32) */
        f32x4 z1[4], z2[4];
        const bf16x8 ax = a_rows(&L.x[0][0], KIN, 0, lane);
#pragma unroll
        for (int t = 0; t < 4; ++t) z1[t] = mfma(ax, b_cols(&L.w1t[0][0], KIN, 16 * t, 0, lane), f32x4{0, 0, 0, 0});
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float g, dg;
                z1[t][i] += L.b1[16 * t + cl];
                gelu(z1[t][i], &g, &dg);
                L.a1[rq + i][16 * t + cl] = (bf16)g;
                z1[t][i] = dg;                                 /* keep GELU'(Z1) */
            }
        __syncthreads();
        /* layer 2: Z2 = A1 W2 + b2 (K = 64) */
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            z2[t] = mfma(a_rows(&L.a1[0][0], NH, 0, lane), b_cols(&L.w2t[0][0], NH, 16 * t, 0, lane), f32x4{0, 0, 0, 0});
            z2[t] = mfma(a_rows(&L.a1[0][0], NH, 32, lane), b_cols(&L.w2t[0][0], NH, 16 * t, 32, lane), z2[t]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float g, dg;
                z2[t][i] += L.b2[16 * t + cl];
                gelu(z2[t][i], &g, &dg);
                L.a2[rq + i][16 * t + cl] = (bf16)g;
                z2[t][i] = dg;
            }
        __syncthreads();
        /* layer 3: Z3 = A2 W3 + b3 (K = 64, 16 padded outputs) */
        f32x4 z3 = mfma(a_rows(&L.a2[0][0], NH, 0, lane), b_cols(&L.w3t[0][0], NH, 0, 0, lane), f32x4{0, 0, 0, 0});
        z3 = mfma(a_rows(&L.a2[0][0], NH, 32, lane), b_cols(&L.w3t[0][0], NH, 0, 32, lane), z3);
#pragma unroll
        for (int i = 0; i < 4; ++i) z3[i] += L.b3[cl];

        if (!TRAIN) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const long long r = r0 + rq + i;
                if (r < a.n && cl < 2) {
                    const float z = z3[i];
                    a.out[2 * r + cl] = cl == 0 ? 1.0f / (1.0f + fexp(-z)) : z;
                }
            }
            __syncthreads();
            continue;
        }

        /* loss and dL/dZ3 (columns 0: EAD logit, 1: APD90) */
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long r = r0 + rq + i;
            float d = 0.0f;
            if (r < a.n && cl < 2) {
                const long long src = a.idx ? a.idx[r] : r;
                const float sw = (a.s ? a.s[src] : 1.0f) * a.inv_wsum;
                const float z = z3[i];
                if (cl == 0) {
                    const float y = a.y[src];
                    const float sg = 1.0f / (1.0f + fexp(-z));
                    const float az = z < 0.0f ? -z : z;
                    loss += sw * ((z > 0.0f ? z : 0.0f) - z * y + __builtin_amdgcn_logf(1.0f + fexp(-az)) * 0.6931471805599453f);
                    d = sw * (sg - y);
                } else {
                    const float mk = a.mask[src];
                    const float e = z - a.t[src];
                    loss += sw * a.w_apd * mk * e * e;
                    d = sw * 2.0f * a.w_apd * mk * e;
                }
            }
            L.d3[rq + i][cl] = (bf16)d;
            gb3 += d;
        }
        __syncthreads();
        /* dW3 = A2^T dZ3 (K = 16 samples); dA2 = dZ3 W3^T (K = 16) */
#pragma unroll
        for (int m = 0; m < 4; ++m)
            gw3[m] = mfma(k16_cols(&L.a2[0][0], NH, 16 * m, lane), k16_cols(&L.d3[0][0], NO, 0, lane), gw3[m]);
        {
            bf16x8 ad;
#pragma unroll
            for (int j = 0; j < 8; ++j) ad[j] = lane < 32 ? L.d3[cl][8 * (lane >> 4) + j] : (bf16)0.0f;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                bf16x8 bw;
#pragma unroll
                for (int j = 0; j < 8; ++j) bw[j] = lane < 32 ? L.w3[16 * t + cl][8 * (lane >> 4) + j] : (bf16)0.0f;
                const f32x4 da2 = mfma(ad, bw, f32x4{0, 0, 0, 0});
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float d = da2[i] * z2[t][i];         /* dZ2 = dA2 * GELU'(Z2) */
                    L.d2[rq + i][16 * t + cl] = (bf16)d;
                    gb2[t] += d;
                }
            }
        }
        __syncthreads();
        /* dW2 = A1^T dZ2 (K = 16); dA1 = dZ2 W2^T (K = 64) */
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const bf16x8 aa = k16_cols(&L.a1[0][0], NH, 16 * m, lane);
#pragma unroll
            for (int t = 0; t < 4; ++t) gw2[m][t] = mfma(aa, k16_cols(&L.d2[0][0], NH, 16 * t, lane), gw2[m][t]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 da1 = mfma(a_rows(&L.d2[0][0], NH, 0, lane), b_cols(&L.w2[0][0], NH, 16 * t, 0, lane), f32x4{0, 0, 0, 0});
            da1 = mfma(a_rows(&L.d2[0][0], NH, 32, lane), b_cols(&L.w2[0][0], NH, 16 * t, 32, lane), da1);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float d = da1[i] * z1[t][i];
                L.d1[rq + i][16 * t + cl] = (bf16)d;
                gb1[t] += d;
            }
        }
        __syncthreads();
        /* dW1 = X^T dZ1 (K = 16) */
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const bf16x8 ax2 = k16_cols(&L.x[0][0], KIN, 16 * m, lane);
#pragma unroll
            for (int t = 0; t < 4; ++t) gw1[m][t] = mfma(ax2, k16_cols(&L.d1[0][0], NH, 16 * t, lane), gw1[m][t]);
        }
        __syncthreads();
    }
    if (!TRAIN) return;

    /* this workgroup's partial: weights from the D layout, biases summed
     * over the four row groups of each column */
    float *g = a.gpart + (long long)blockIdx.x * (NP + 1);
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) g[OFF_W1 + (16 * m + rq + i) * NH + 16 * t + cl] = gw1[m][t][i];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) g[OFF_W2 + (16 * m + rq + i) * NH + 16 * t + cl] = gw2[m][t][i];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) g[OFF_W3 + (16 * m + rq + i) * NO + cl] = gw3[m][i];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        float v1 = gb1[t], v2 = gb2[t];
        v1 += __shfl_xor(v1, 16, 64); v1 += __shfl_xor(v1, 32, 64);
        v2 += __shfl_xor(v2, 16, 64); v2 += __shfl_xor(v2, 32, 64);
        if (lane < 16) { g[OFF_B1 + 16 * t + cl] = v1; g[OFF_B2 + 16 * t + cl] = v2; }
    }
    {
        float v3 = gb3;
        v3 += __shfl_xor(v3, 16, 64); v3 += __shfl_xor(v3, 32, 64);
        if (lane < 16) g[OFF_B3 + cl] = v3;
        float ls = loss;
        for (int o = 1; o < 64; o <<= 1) ls += __shfl_xor(ls, o, 64);
        if (lane == 0) g[NP] = ls;
    }
}

/* Sum of the workgroup partials in workgroup order; then either the
 * gradient out (adam = 0) or one Adam step (adam = 1, step count t). */
__global__ void k_mlp_reduce(const float *gpart, int nwg, float *grad, float *w, float *mo, float *vo, bf16 *wb,
                             int adam, float lr, float b1, float b2, float eps, float c1, float c2, float *loss_out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NP) return;
    float g = 0.0f;
    for (int q = 0; q < nwg; ++q) g += gpart[(long long)q * (NP + 1) + p];
    if (p == NP) { if (loss_out) *loss_out = g; return; }
    if (grad) grad[p] = g;
    if (!adam) return;
    const float m = b1 * mo[p] + (1.0f - b1) * g;
    const float v = b2 * vo[p] + (1.0f - b2) * g * g;
    mo[p] = m;
    vo[p] = v;
    const float nw = w[p] - lr * (m * c1) / (__builtin_sqrtf(v * c2) + eps);
    w[p] = nw;
    wb[p] = (bf16)nw;
}

/* float samples [n][11] -> BF16 [n][32], zero padded */
__global__ void k_mlp_pack(const float *x, long long n, bf16 *xb)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * KIN) return;
    const long long r = e / KIN;
    const int k = (int)(e % KIN);
    xb[e] = (bf16)(k < RYR_MLP_NIN ? x[r * RYR_MLP_NIN + k] : 0.0f);
}

__global__ void k_mlp_tobf16(const float *w, bf16 *wb, int n)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) wb[p] = (bf16)w[p];
}

/* ------------------------------------------------------------------ host */
struct ryr_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *loss = nullptr;
    bf16 *wb = nullptr;
    float *gpart = nullptr;
    size_t gpart_wg = 0;
    long long adam_t = 0;
    std::string err;
};

#define MLP_CHECK(h, call)                                                         \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return RYR_E_HIP;                                                      \
        }                                                                          \
    } while (0)

/* device scratch of one call, freed on every path out of it */
struct MlpScratch {
    std::vector<void *> p;
    ~MlpScratch() { for (void *q : p) hipFree(q); }
    template <typename T> hipError_t alloc(T **ptr, size_t n)
    {
        hipError_t e = hipMalloc((void **)ptr, n > 0 ? n * sizeof(T) : sizeof(T));
        if (e == hipSuccess) p.push_back(*ptr);
        return e;
    }
};

static const int kTiles = 4;   /* 16-row tiles per workgroup */

static int ensure_gpart(ryr_mlp *h, size_t nwg)
{
    if (nwg <= h->gpart_wg) return RYR_OK;
    if (h->gpart) { hipFree(h->gpart); h->gpart = nullptr; h->gpart_wg = 0; }
    MLP_CHECK(h, hipMalloc((void **)&h->gpart, nwg * (NP + 1) * sizeof(float)));
    h->gpart_wg = nwg;
    return RYR_OK;
}

template <typename T>
static hipError_t upload(MlpScratch &sc, T **dst, const T *src, size_t n, hipStream_t s)
{
    hipError_t e = sc.alloc(dst, n);
    if (e == hipSuccess && n) e = hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s);
    return e;
}

extern "C" {

int ryr_mlp_create(int device, ryr_mlp **out)
{
    if (!out) return RYR_E_ARG;
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return RYR_E_NODEV;
    if (device < 0 || device >= nd) return RYR_E_ARG;
    ryr_mlp *h = new ryr_mlp();
    h->device = device;
    auto fail = [&]() { ryr_mlp_destroy(h); return RYR_E_HIP; };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return fail();
    if (hipMalloc((void **)&h->w, NP * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->m, NP * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->v, NP * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->grad, NP * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->loss, sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->wb, NP * sizeof(bf16)) != hipSuccess)
        return fail();
    if (hipMemset(h->w, 0, NP * sizeof(float)) != hipSuccess || hipMemset(h->m, 0, NP * sizeof(float)) != hipSuccess ||
        hipMemset(h->v, 0, NP * sizeof(float)) != hipSuccess || hipMemset(h->wb, 0, NP * sizeof(bf16)) != hipSuccess)
        return fail();
    *out = h;
    return RYR_OK;
}

void ryr_mlp_destroy(ryr_mlp *h)
{
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    for (void *p : {(void *)h->w, (void *)h->m, (void *)h->v, (void *)h->grad, (void *)h->loss, (void *)h->wb,
                    (void *)h->gpart})
        if (p) hipFree(p);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

const char *ryr_mlp_last_error(const ryr_mlp *h) { return h ? h->err.c_str() : "null handle"; }

int ryr_mlp_set_params(ryr_mlp *h, const float *params, int64_t n)
{
    if (!h || !params || n != NP) return RYR_E_ARG;
    MLP_CHECK(h, hipSetDevice(h->device));
    MLP_CHECK(h, hipMemcpyAsync(h->w, params, NP * sizeof(float), hipMemcpyHostToDevice, h->stream));
    MLP_CHECK(h, hipMemsetAsync(h->m, 0, NP * sizeof(float), h->stream));
    MLP_CHECK(h, hipMemsetAsync(h->v, 0, NP * sizeof(float), h->stream));
    hipLaunchKernelGGL(k_mlp_tobf16, dim3((NP + 255) / 256), dim3(256), 0, h->stream, h->w, h->wb, NP);
    MLP_CHECK(h, hipGetLastError());
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    h->adam_t = 0;
    return RYR_OK;
}

int ryr_mlp_get_params(ryr_mlp *h, float *params, int64_t n)
{
    if (!h || !params || n != NP) return RYR_E_ARG;
    MLP_CHECK(h, hipSetDevice(h->device));
    MLP_CHECK(h, hipMemcpyAsync(params, h->w, NP * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    return RYR_OK;
}

int ryr_mlp_forward(ryr_mlp *h, const float *x, int64_t n, float *out)
{
    if (!h || n < 0 || (n > 0 && (!x || !out))) return RYR_E_ARG;
    if (n == 0) return RYR_OK;
    MLP_CHECK(h, hipSetDevice(h->device));
    MlpScratch sc;
    float *dx = nullptr, *dout = nullptr;
    bf16 *xb = nullptr;
    MLP_CHECK(h, upload(sc, &dx, x, (size_t)n * RYR_MLP_NIN, h->stream));
    MLP_CHECK(h, sc.alloc(&xb, (size_t)n * KIN));
    MLP_CHECK(h, sc.alloc(&dout, (size_t)n * 2));
    hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((n * KIN + 255) / 256)), dim3(256), 0, h->stream, dx, (long long)n, xb);
    MlpArgs a{};
    a.wb = h->wb; a.wf = h->w; a.x = xb; a.n = n; a.tiles = kTiles; a.out = dout;
    const long long ntile = (n + 15) / 16;
    hipLaunchKernelGGL((k_mlp<false>), dim3((unsigned)((ntile + kTiles - 1) / kTiles)), dim3(64), 0, h->stream, a);
    MLP_CHECK(h, hipGetLastError());
    MLP_CHECK(h, hipMemcpyAsync(out, dout, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    return RYR_OK;
}

int ryr_mlp_grad(ryr_mlp *h, const float *x, const float *y, const float *t, const float *mask, const float *s,
                 int64_t n, float w_apd, float *loss, float *grad)
{
    if (!h || n <= 0 || !x || !y || !t || !mask || !loss || !grad) return RYR_E_ARG;
    MLP_CHECK(h, hipSetDevice(h->device));
    double wsum = 0.0;
    for (int64_t i = 0; i < n; ++i) wsum += s ? (double)s[i] : 1.0;
    if (!(wsum > 0.0)) { h->err = "sample weights sum to zero"; return RYR_E_ARG; }
    MlpScratch sc;
    float *dx, *dy, *dt, *dm, *ds = nullptr;
    bf16 *xb;
    MLP_CHECK(h, upload(sc, &dx, x, (size_t)n * RYR_MLP_NIN, h->stream));
    MLP_CHECK(h, upload(sc, &dy, y, (size_t)n, h->stream));
    MLP_CHECK(h, upload(sc, &dt, t, (size_t)n, h->stream));
    MLP_CHECK(h, upload(sc, &dm, mask, (size_t)n, h->stream));
    if (s) MLP_CHECK(h, upload(sc, &ds, s, (size_t)n, h->stream));
    MLP_CHECK(h, sc.alloc(&xb, (size_t)n * KIN));
    hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((n * KIN + 255) / 256)), dim3(256), 0, h->stream, dx, (long long)n, xb);
    const long long ntile = (n + 15) / 16;
    const int nwg = (int)((ntile + kTiles - 1) / kTiles);
    if (int rc = ensure_gpart(h, nwg)) return rc;
    MlpArgs a{};
    a.wb = h->wb; a.wf = h->w; a.x = xb; a.y = dy; a.t = dt; a.mask = dm; a.s = ds; a.n = n;
    a.tiles = kTiles; a.w_apd = w_apd; a.inv_wsum = (float)(1.0 / wsum); a.gpart = h->gpart;
    hipLaunchKernelGGL((k_mlp<true>), dim3(nwg), dim3(64), 0, h->stream, a);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NP + 1 + 255) / 256), dim3(256), 0, h->stream, (const float *)h->gpart, nwg,
                       h->grad, h->w, h->m, h->v, h->wb, 0, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, h->loss);
    MLP_CHECK(h, hipGetLastError());
    MLP_CHECK(h, hipMemcpyAsync(grad, h->grad, NP * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    MLP_CHECK(h, hipMemcpyAsync(loss, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    return RYR_OK;
}

int ryr_mlp_train(ryr_mlp *h, const float *x, const float *y, const float *t, const float *mask, int64_t n,
                  float w_apd, int64_t steps, int64_t batch, float lr, float beta1, float beta2, float eps,
                  uint64_t seed, float *loss_trace)
{
    if (!h || n <= 0 || !x || !y || !t || !mask || steps < 0 || batch <= 0 || !(lr > 0.f)) return RYR_E_ARG;
    if (batch > n) batch = n;
    MLP_CHECK(h, hipSetDevice(h->device));
    MlpScratch sc;
    float *dx, *dy, *dt, *dm, *dtrace;
    bf16 *xb;
    int *didx;
    MLP_CHECK(h, upload(sc, &dx, x, (size_t)n * RYR_MLP_NIN, h->stream));
    MLP_CHECK(h, upload(sc, &dy, y, (size_t)n, h->stream));
    MLP_CHECK(h, upload(sc, &dt, t, (size_t)n, h->stream));
    MLP_CHECK(h, upload(sc, &dm, mask, (size_t)n, h->stream));
    MLP_CHECK(h, sc.alloc(&xb, (size_t)n * KIN));
    MLP_CHECK(h, sc.alloc(&dtrace, (size_t)(steps > 0 ? steps : 1)));
    hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((n * KIN + 255) / 256)), dim3(256), 0, h->stream, dx, (long long)n, xb);
    /* minibatch order: consecutive slices of per-pass permutations */
    const int64_t passes = (steps * batch + n - 1) / n + 1;
    std::vector<int> perm((size_t)(passes * n));
    std::mt19937_64 rng(seed);
    for (int64_t p = 0; p < passes; ++p) {
        int *q = perm.data() + p * n;
        for (int64_t i = 0; i < n; ++i) q[i] = (int)i;
        for (int64_t i = n - 1; i > 0; --i) {
            const int64_t j = (int64_t)(rng() % (uint64_t)(i + 1));
            std::swap(q[i], q[j]);
        }
    }
    MLP_CHECK(h, upload(sc, &didx, perm.data(), perm.size(), h->stream));
    const long long ntile = (batch + 15) / 16;
    const int nwg = (int)((ntile + kTiles - 1) / kTiles);
    if (int rc = ensure_gpart(h, nwg)) return rc;
    MlpArgs a{};
    a.wb = h->wb; a.wf = h->w; a.x = xb; a.y = dy; a.t = dt; a.mask = dm; a.s = nullptr; a.n = batch;
    a.tiles = kTiles; a.w_apd = w_apd; a.inv_wsum = 1.0f / (float)batch; a.gpart = h->gpart;
    for (int64_t st = 0; st < steps; ++st) {
        a.idx = didx + st * batch;
        hipLaunchKernelGGL((k_mlp<true>), dim3(nwg), dim3(64), 0, h->stream, a);
        h->adam_t += 1;
        const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, (double)h->adam_t)));
        const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, (double)h->adam_t)));
        hipLaunchKernelGGL(k_mlp_reduce, dim3((NP + 1 + 255) / 256), dim3(256), 0, h->stream, (const float *)h->gpart,
                           nwg, (float *)nullptr, h->w, h->m, h->v, h->wb, 1, lr, beta1, beta2, eps, c1, c2, dtrace + st);
    }
    MLP_CHECK(h, hipGetLastError());
    if (loss_trace && steps > 0)
        MLP_CHECK(h, hipMemcpyAsync(loss_trace, dtrace, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    return RYR_OK;
}

}  /* extern "C" */
