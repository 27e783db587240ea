/* Launchers of the device kernels, one set per precision translation unit
 * (cru_step_f64.hip built with contraction off, cru_step_f32.hip with
 * contraction on outside the strict-FP regions).
 *
 * The sub-voxel kernel and the CRU kernel come in three workgroup sizes
 * (64, 128, 256 lanes).  Every size computes the same values in the same
 * order (the whole-cell sums are formed per wave), so the shape is a pure
 * speed choice.  With a non-null event array the step kernels are launched
 * with hipExtLaunchKernelGGL, which stamps each kernel's own start and end
 * (ev[0..1] sub-voxel, ev[2..3] CRU, ev[4..5] membrane); such launches are
 * never captured into a graph. */
#ifndef RYR_CRU_LAUNCH_H
#define RYR_CRU_LAUNCH_H

#include <hip/hip_ext.h>
#include "cru_device.h"

#define RYR_DECLARE_LAUNCHERS(sfx)                                                               \
    hipError_t ryr_launch_step_##sfx(const RyrKArgs &a, double amp, int nchunks, int ncells,     \
                                     int wg_cyto, int wg_cru, hipStream_t s, hipEvent_t *ev);   \
    hipError_t ryr_launch_diag_##sfx(const RyrKArgs &a, int nchunks, hipStream_t s,             \
                                     hipEvent_t *ev);                                           \
    hipError_t ryr_launch_init_##sfx(const RyrKArgs &a, const RyrInit *init, int nblocks,        \
                                     hipStream_t s);                                             \
    hipError_t ryr_launch_export_##sfx(const RyrKArgs &a, int cell, int field, int planes,       \
                                       long long nel, double *out, hipStream_t s);               \
    hipError_t ryr_launch_import_##sfx(const RyrKArgs &a, int cell, int field, int planes,       \
                                       long long nel, const double *in, hipStream_t s);

RYR_DECLARE_LAUNCHERS(f64)
RYR_DECLARE_LAUNCHERS(f32)

template <typename TF, typename TS, int WG>
static hipError_t ryr_go_cyto(const RyrKArgs &a, int nchunks, hipStream_t s, hipEvent_t *ev)
{
    const dim3 grid((unsigned)nchunks * 8u * (unsigned)(RYR_BLOCK / WG));
    if (ev) hipExtLaunchKernelGGL(k_cyto<TF, TS, WG>, grid, dim3(WG), 0, s, ev[0], ev[1], 0, a);
    else hipLaunchKernelGGL((k_cyto<TF, TS, WG>), grid, dim3(WG), 0, s, a);
    return hipGetLastError();
}

template <typename TF, typename TS, int WG>
static hipError_t ryr_go_cru(const RyrKArgs &a, double amp, int nchunks, hipStream_t s, hipEvent_t *ev)
{
    const dim3 grid((unsigned)nchunks * (unsigned)(RYR_BLOCK / WG));
    if (ev) hipExtLaunchKernelGGL(k_cru_step<TF, TS, WG>, grid, dim3(WG), 0, s, ev[2], ev[3], 0, a, amp);
    else hipLaunchKernelGGL((k_cru_step<TF, TS, WG>), grid, dim3(WG), 0, s, a, amp);
    return hipGetLastError();
}

template <typename TF, typename TS>
static hipError_t ryr_go_step(const RyrKArgs &a, double amp, int nchunks, int ncells, int wg_cyto,
                              int wg_cru, hipStream_t s, hipEvent_t *ev)
{
    hipError_t e;
    switch (wg_cyto) {
    case 64: e = ryr_go_cyto<TF, TS, 64>(a, nchunks, s, ev); break;
    case 128: e = ryr_go_cyto<TF, TS, 128>(a, nchunks, s, ev); break;
    case 256: e = ryr_go_cyto<TF, TS, 256>(a, nchunks, s, ev); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    switch (wg_cru) {
    case 64: e = ryr_go_cru<TF, TS, 64>(a, amp, nchunks, s, ev); break;
    case 128: e = ryr_go_cru<TF, TS, 128>(a, amp, nchunks, s, ev); break;
    case 256: e = ryr_go_cru<TF, TS, 256>(a, amp, nchunks, s, ev); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (ev) hipExtLaunchKernelGGL(k_membrane<TF, TS>, dim3(ncells), dim3(RYR_BLOCK), 0, s, ev[4], ev[5], 0, a, amp);
    else hipLaunchKernelGGL((k_membrane<TF, TS>), dim3(ncells), dim3(RYR_BLOCK), 0, s, a, amp);
    return hipGetLastError();
}

#define RYR_DEFINE_LAUNCHERS(sfx, TF, TS)                                                        \
    hipError_t ryr_launch_step_##sfx(const RyrKArgs &a, double amp, int nchunks, int ncells,     \
                                     int wg_cyto, int wg_cru, hipStream_t s, hipEvent_t *ev)    \
    {                                                                                            \
        return ryr_go_step<TF, TS>(a, amp, nchunks, ncells, wg_cyto, wg_cru, s, ev);             \
    }                                                                                            \
    hipError_t ryr_launch_diag_##sfx(const RyrKArgs &a, int nchunks, hipStream_t s,             \
                                     hipEvent_t *ev)                                            \
    {                                                                                            \
        /* one workgroup per 256-CRU chunk */                                                    \
        if (ev) hipExtLaunchKernelGGL(k_diag<TF, TS>, dim3(nchunks), dim3(RYR_BLOCK), 0, s,      \
                                      ev[0], ev[1], 0, a);                                      \
        else hipLaunchKernelGGL((k_diag<TF, TS>), dim3(nchunks), dim3(RYR_BLOCK), 0, s, a);      \
        return hipGetLastError();                                                                \
    }                                                                                            \
    hipError_t ryr_launch_init_##sfx(const RyrKArgs &a, const RyrInit *init, int nblocks,        \
                                     hipStream_t s)                                              \
    {                                                                                            \
        hipLaunchKernelGGL((k_init<TF, TS>), dim3(nblocks), dim3(RYR_BLOCK), 0, s, a, init);     \
        return hipGetLastError();                                                                \
    }                                                                                            \
    hipError_t ryr_launch_export_##sfx(const RyrKArgs &a, int cell, int field, int planes,       \
                                       long long nel, double *out, hipStream_t s)                \
    {                                                                                            \
        int nb = (int)((nel + 255) / 256);                                                       \
        hipLaunchKernelGGL((k_export<TF, TS>), dim3(nb), dim3(256), 0, s, a, cell, field,        \
                           planes, out);                                                         \
        return hipGetLastError();                                                                \
    }                                                                                            \
    hipError_t ryr_launch_import_##sfx(const RyrKArgs &a, int cell, int field, int planes,       \
                                       long long nel, const double *in, hipStream_t s)           \
    {                                                                                            \
        int nb = (int)((nel + 255) / 256);                                                       \
        hipLaunchKernelGGL((k_import<TF, TS>), dim3(nb), dim3(256), 0, s, a, cell, field,        \
                           planes, in);                                                          \
        return hipGetLastError();                                                                \
    }

#endif
