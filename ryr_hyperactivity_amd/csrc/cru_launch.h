/* Launchers of the device kernels, one set per precision translation unit
 * (cru_step_f64.hip built with contraction off, cru_step_f32.hip with
 * contraction on outside the strict-FP regions). */
#ifndef RYR_CRU_LAUNCH_H
#define RYR_CRU_LAUNCH_H

#include "cru_device.h"

#define RYR_DECLARE_LAUNCHERS(sfx)                                                              \
    hipError_t ryr_launch_step_##sfx(const RyrKArgs &a, double amp, int nblocks, int ncells, hipStream_t s); \
    hipError_t ryr_launch_diag_##sfx(const RyrKArgs &a, int ncells, hipStream_t s);              \
    hipError_t ryr_launch_init_##sfx(const RyrKArgs &a, const RyrInit *init, int nblocks,       \
                                     hipStream_t s);                                            \
    hipError_t ryr_launch_export_##sfx(const RyrKArgs &a, int cell, int field, int planes,      \
                                       long long nel, double *out, hipStream_t s);              \
    hipError_t ryr_launch_import_##sfx(const RyrKArgs &a, int cell, int field, int planes,      \
                                       long long nel, const double *in, hipStream_t s);

RYR_DECLARE_LAUNCHERS(f64)
RYR_DECLARE_LAUNCHERS(f32)

#define RYR_DEFINE_LAUNCHERS(sfx, TF, TS)                                                       \
    hipError_t ryr_launch_step_##sfx(const RyrKArgs &a, double amp, int nblocks, int ncells,    \
                                     hipStream_t s)                                             \
    {                                                                                           \
        hipLaunchKernelGGL((k_cyto<TF, TS>), dim3(nblocks, 8), dim3(RYR_BLOCK), 0, s, a);         \
        hipLaunchKernelGGL((k_cru_step<TF, TS>), dim3(nblocks), dim3(RYR_BLOCK), 0, s, a, amp); \
        hipLaunchKernelGGL((k_membrane<TF, TS>), dim3(ncells), dim3(RYR_BLOCK), 0, s, a, amp);  \
        return hipGetLastError();                                                               \
    }                                                                                           \
    hipError_t ryr_launch_diag_##sfx(const RyrKArgs &a, int ncells, hipStream_t s)              \
    {                                                                                           \
        hipLaunchKernelGGL((k_diag<TF, TS>), dim3(ncells), dim3(RYR_BLOCK), 0, s, a);   /* ncells = workgroups */           \
        return hipGetLastError();                                                               \
    }                                                                                           \
    hipError_t ryr_launch_init_##sfx(const RyrKArgs &a, const RyrInit *init, int nblocks,       \
                                     hipStream_t s)                                             \
    {                                                                                           \
        hipLaunchKernelGGL((k_init<TF, TS>), dim3(nblocks), dim3(RYR_BLOCK), 0, s, a, init);    \
        return hipGetLastError();                                                               \
    }                                                                                           \
    hipError_t ryr_launch_export_##sfx(const RyrKArgs &a, int cell, int field, int planes,      \
                                       long long nel, double *out, hipStream_t s)               \
    {                                                                                           \
        int nb = (int)((nel + 255) / 256);                                                      \
        hipLaunchKernelGGL((k_export<TF, TS>), dim3(nb), dim3(256), 0, s, a, cell, field,       \
                           planes, out);                                                        \
        return hipGetLastError();                                                               \
    }                                                                                           \
    hipError_t ryr_launch_import_##sfx(const RyrKArgs &a, int cell, int field, int planes,      \
                                       long long nel, const double *in, hipStream_t s)          \
    {                                                                                           \
        int nb = (int)((nel + 255) / 256);                                                      \
        hipLaunchKernelGGL((k_import<TF, TS>), dim3(nb), dim3(256), 0, s, a, cell, field,       \
                           planes, in);                                                         \
        return hipGetLastError();                                                               \
    }

#endif
