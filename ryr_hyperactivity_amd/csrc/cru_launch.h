/* Launchers of the device kernels, one set per precision translation unit
 * (cru_step_f64.hip built with contraction off, cru_step_f32.hip with
 * contraction on outside the strict-FP regions).
 *
 * The sub-voxel kernel and the CRU kernel come in three workgroup sizes
 * (64, 128, 256 lanes); the sub-voxel kernel also in a one-lane-per-CRU form
 * (size code 1024), the CRU kernel in a two-lanes-per-CRU form (512).  Every size computes the same values in the same
 * order (the whole-cell sums are formed per wave), so the shape is a pure
 * speed choice, and so is carrying the membrane step inside the sub-voxel
 * launch.  With a non-null event pair a kernel is launched with
 * hipExtLaunchKernelGGL, which stamps its own start and end; such launches
 * are never captured into a graph. */
#ifndef RYR_CRU_LAUNCH_H
#define RYR_CRU_LAUNCH_H

#include <hip/hip_ext.h>
#include "cru_device.h"

#define RYR_DECLARE_LAUNCHERS(sfx)                                                               \
    hipError_t ryr_launch_cyto_##sfx(const RyrKArgs &a, int nchunks, int wg, int kc, int km,     \
                                     bool gate, double amp, hipStream_t s, hipEvent_t *ev);     \
    hipError_t ryr_launch_cru_##sfx(const RyrKArgs &a, int nchunks, int wg, int kc, bool split,  \
                                    hipStream_t s, hipEvent_t *ev);                             \
    hipError_t ryr_launch_mem_##sfx(const RyrKArgs &a, int ncells, int km, double amp,           \
                                    hipStream_t s, hipEvent_t *ev);                             \
    hipError_t ryr_launch_diag_##sfx(const RyrKArgs &a, int nchunks, hipStream_t s,             \
                                     hipEvent_t *ev);                                           \
    hipError_t ryr_launch_one_##sfx(const RyrKArgs &a, int nchunks, int lpc, int kc, bool next,  \
                                    double amp, hipStream_t s, hipEvent_t *ev);                 \
    hipError_t ryr_launch_init_##sfx(const RyrKArgs &a, const RyrInit *init, int nblocks,        \
                                     hipStream_t s);                                             \
    hipError_t ryr_launch_export_##sfx(const RyrKArgs &a, int cell, int field, int planes,       \
                                       long long nel, double *out, hipStream_t s);               \
    hipError_t ryr_launch_import_##sfx(const RyrKArgs &a, int cell, int field, int planes,       \
                                       long long nel, const double *in, hipStream_t s);

RYR_DECLARE_LAUNCHERS(f64)
RYR_DECLARE_LAUNCHERS(f32)

/* One kernel, optionally stamped by its own start/end events (ev[0], ev[1]). */
#define RYR_GO(KERN, GRID, BLOCK, S, EV, ...)                                                    \
    do {                                                                                         \
        if (EV) hipExtLaunchKernelGGL(KERN, GRID, BLOCK, 0, S, (EV)[0], (EV)[1], 0, __VA_ARGS__); \
        else hipLaunchKernelGGL((KERN), GRID, BLOCK, 0, S, __VA_ARGS__);                         \
    } while (0)

/* Sub-voxel kernel of step gstep+kc; with km >= 0 it also runs the membrane
 * step gstep+km of every cell in its first ncells workgroups, and with gate
 * the RyR gating of step gstep+kc in the next nchunks * 256/wg * RYR_GATE_LPC. */
template <typename TF, typename TS, int WG, int MODE>
static hipError_t ryr_go_cyto_m(const RyrKArgs &a, int nchunks, int kc, int km, double amp, hipStream_t s,
                                hipEvent_t *ev)
{
    const unsigned per = (unsigned)(RYR_BLOCK / WG) * (unsigned)nchunks;
    const unsigned grid = ((MODE & 1) ? (unsigned)a.ncells : 0u) + per * ((MODE & 2) ? 8u + RYR_GATE_LPC : 8u);
    RYR_GO((k_cyto<TF, TS, WG, MODE>), dim3(grid), dim3(WG), s, ev, a, kc, km, amp);
    return hipGetLastError();
}

template <typename TF, typename TS, int MODE>
static hipError_t ryr_go_cyto_w(const RyrKArgs &a, int nchunks, int wg, int kc, int km, double amp,
                                hipStream_t s, hipEvent_t *ev)
{
    switch (wg) {
    case 64: return ryr_go_cyto_m<TF, TS, 64, MODE>(a, nchunks, kc, km, amp, s, ev);
    case 128: return ryr_go_cyto_m<TF, TS, 128, MODE>(a, nchunks, kc, km, amp, s, ev);
    case 256: return ryr_go_cyto_m<TF, TS, 256, MODE>(a, nchunks, kc, km, amp, s, ev);
    case 1024:
        /* one lane per CRU (cyto_cru_block), one 256-lane workgroup per chunk; the sweep alone */
        if constexpr (MODE == 0) {
            RYR_GO((k_cyto_cru<TF, TS>), dim3((unsigned)nchunks), dim3(RYR_BLOCK), s, ev, a, kc);
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

template <typename TF, typename TS>
static hipError_t ryr_go_cyto(const RyrKArgs &a, int nchunks, int wg, int kc, int km, bool gate, double amp,
                              hipStream_t s, hipEvent_t *ev)
{
    switch ((km >= 0 ? 1 : 0) | (gate ? 2 : 0)) {
    case 0: return ryr_go_cyto_w<TF, TS, 0>(a, nchunks, wg, kc, km, amp, s, ev);
    case 1: return ryr_go_cyto_w<TF, TS, 1>(a, nchunks, wg, kc, km, amp, s, ev);
    case 2: return ryr_go_cyto_w<TF, TS, 2>(a, nchunks, wg, kc, km, amp, s, ev);
    default: return ryr_go_cyto_w<TF, TS, 3>(a, nchunks, wg, kc, km, amp, s, ev);
    }
}

template <typename TF, typename TS, bool SCHB>
static hipError_t ryr_go_cru_s(const RyrKArgs &a, int nchunks, int wg, int kc, hipStream_t s, hipEvent_t *ev)
{
    switch (wg) {
    case 64: RYR_GO((k_cru_step<TF, TS, 64, SCHB>), dim3((unsigned)nchunks * 4u), dim3(64), s, ev, a, kc); break;
    case 128: RYR_GO((k_cru_step<TF, TS, 128, SCHB>), dim3((unsigned)nchunks * 2u), dim3(128), s, ev, a, kc); break;
    case 256: RYR_GO((k_cru_step<TF, TS, 256, SCHB>), dim3((unsigned)nchunks), dim3(256), s, ev, a, kc); break;
    /* 512 lanes per 256-CRU chunk: two 256-lane workgroups, two lanes per CRU */
    case 512: RYR_GO((k_cru_step<TF, TS, 256, SCHB, 2>), dim3((unsigned)nchunks * 2u), dim3(256), s, ev, a, kc); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

/* the CRU kernel without the RyR gating (run by the sub-voxel launch) */
template <typename TF, typename TS, bool SCHB>
static hipError_t ryr_go_cru_split(const RyrKArgs &a, int nchunks, int wg, int kc, hipStream_t s, hipEvent_t *ev)
{
    switch (wg) {
    case 128: RYR_GO((k_cru_step<TF, TS, 128, SCHB, 1, true>), dim3((unsigned)nchunks * 2u), dim3(128), s, ev, a, kc); break;
    case 256: RYR_GO((k_cru_step<TF, TS, 256, SCHB, 1, true>), dim3((unsigned)nchunks), dim3(256), s, ev, a, kc); break;
    case 512: RYR_GO((k_cru_step<TF, TS, 256, SCHB, 2, true>), dim3((unsigned)nchunks * 2u), dim3(256), s, ev, a, kc); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename TF, typename TS>
static hipError_t ryr_go_cru(const RyrKArgs &a, int nchunks, int wg, int kc, bool split, hipStream_t s,
                             hipEvent_t *ev)
{
    if (split)
        return a.scheme ? ryr_go_cru_split<TF, TS, true>(a, nchunks, wg, kc, s, ev)
                        : ryr_go_cru_split<TF, TS, false>(a, nchunks, wg, kc, s, ev);
    return a.scheme ? ryr_go_cru_s<TF, TS, true>(a, nchunks, wg, kc, s, ev)
                    : ryr_go_cru_s<TF, TS, false>(a, nchunks, wg, kc, s, ev);
}

#define RYR_DEFINE_LAUNCHERS(sfx, TF, TS)                                                        \
    hipError_t ryr_launch_cyto_##sfx(const RyrKArgs &a, int nchunks, int wg, int kc, int km,     \
                                     bool gate, double amp, hipStream_t s, hipEvent_t *ev)      \
    {                                                                                            \
        return ryr_go_cyto<TF, TS>(a, nchunks, wg, kc, km, gate, amp, s, ev);                    \
    }                                                                                            \
    hipError_t ryr_launch_cru_##sfx(const RyrKArgs &a, int nchunks, int wg, int kc, bool split,  \
                                    hipStream_t s, hipEvent_t *ev)                              \
    {                                                                                            \
        return ryr_go_cru<TF, TS>(a, nchunks, wg, kc, split, s, ev);                             \
    }                                                                                            \
    hipError_t ryr_launch_one_##sfx(const RyrKArgs &a, int nchunks, int lpc, int kc, bool next,  \
                                    double amp, hipStream_t s, hipEvent_t *ev)                  \
    {                                                                                            \
        /* membrane workgroups and lpc CRU workgroups per chunk (lpc lanes per CRU; padded to a  \
         * multiple of eight), then eight sweep workgroups per chunk */                         \
        if (lpc != 1 && lpc != 2) return hipErrorInvalidValue;                                   \
        const dim3 grid((unsigned)one_head(a.ncells, lpc * nchunks) + 8u * (unsigned)nchunks);   \
        const long long *gs = a.gstep;                                                           \
        const int nck = (int)(a.P >> 8);                                                         \
        if (lpc == 2) {                                                                          \
            if (next) RYR_GO((k_one<TF, TS, true, 2>), grid, dim3(RYR_BLOCK), s, ev, gs, a.chunks, \
                             a.cells, kc, a.ncells, nck, a, amp);                                \
            else RYR_GO((k_one<TF, TS, false, 2>), grid, dim3(RYR_BLOCK), s, ev, gs, a.chunks,    \
                        a.cells, kc, a.ncells, nck, a, amp);                                     \
        } else {                                                                                 \
            if (next) RYR_GO((k_one<TF, TS, true, 1>), grid, dim3(RYR_BLOCK), s, ev, gs, a.chunks, \
                             a.cells, kc, a.ncells, nck, a, amp);                                \
            else RYR_GO((k_one<TF, TS, false, 1>), grid, dim3(RYR_BLOCK), s, ev, gs, a.chunks,    \
                        a.cells, kc, a.ncells, nck, a, amp);                                     \
        }                                                                                        \
        return hipGetLastError();                                                                \
    }                                                                                            \
    hipError_t ryr_launch_mem_##sfx(const RyrKArgs &a, int ncells, int km, double amp,           \
                                    hipStream_t s, hipEvent_t *ev)                              \
    {                                                                                            \
        RYR_GO((k_membrane<TF, TS>), dim3(ncells), dim3(RYR_BLOCK), s, ev, a, km, amp);          \
        return hipGetLastError();                                                                \
    }                                                                                            \
    hipError_t ryr_launch_diag_##sfx(const RyrKArgs &a, int nchunks, hipStream_t s,             \
                                     hipEvent_t *ev)                                            \
    {                                                                                            \
        /* one workgroup per 256-CRU chunk, then one per cell */                                  \
        if (ev && hipEventRecord(ev[0], s) != hipSuccess) return hipGetLastError();              \
        hipLaunchKernelGGL((k_diag<TF, TS>), dim3(nchunks), dim3(RYR_BLOCK), 0, s, a);           \
        hipLaunchKernelGGL((k_diag_cell<TF, TS>), dim3(a.ncells), dim3(RYR_BLOCK), 0, s, a);     \
        if (ev && hipEventRecord(ev[1], s) != hipSuccess) return hipGetLastError();              \
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
