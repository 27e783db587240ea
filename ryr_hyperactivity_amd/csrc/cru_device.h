/*
 * cru_device.h - gfx950 device code of the CRU update: one CRU per lane, SoA
 * state in HBM, the per-step kernel, the in-kernel whole-cell reduction and
 * membrane ODE, and the output-row reduction kernel.
 *
 * Reference (colinmrees/RyR_hyperactivity): Compute (cuda_Cell_Detailed.cu:
 * 1187-1692) with its leaf physics (subroutine_cuda.c, buffers.c, reesica.h),
 * the host reduction and membrane ODE (cuda_Cell_Detailed.cu:427-739) and the
 * output block (:777-1081).
 *
 * Layout (per handle; all cells concatenated, each cell's CRUs padded to a
 * multiple of 256 so a workgroup never straddles two cells):
 *   slot g = cell offset + n,  n = i + X*(j + Y*k)   (x fastest)
 *   sub-voxel fields: 8 planes, element (ii, g) at ii*P + g
 *   ci, cnsr, cs: two buffers, read buffer = step parity
 * The reference's ghost shell is not stored: every ghost value it writes is a
 * copy of the interior sub-voxel adjacent to it (:1601-1652), which is the
 * very sub-voxel whose stencil reads it, so a neighbour outside the interior
 * is replaced by the centre value.  The differences the reference forms
 * (x_nb - x_c) are then exactly zero, as with the ghost copies.
 *
 * Precision: TF = type of fast concentrations and buffers, TS = type of the
 * large slowly varying pools (cnsr, cjsr, cjsrb, slow troponin).  FP64 mode:
 * TF = TS = double, with the reference's operand order everywhere, so results
 * are bit-identical to the FP64 oracle.  FP32 mode: TF = float, TS = double.
 * Gating, I_CaL/I_NCX per CRU, the reductions and the membrane ODE are FP64
 * in both modes and compiled with contraction off.
 */
#ifndef RYR_CRU_DEVICE_H
#define RYR_CRU_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "common/philox.h"
#include "common/detmath.h"
#include "common/srround.h"

#define RYR_BLOCK 256
#define RYR_NRYR 100
#define RYR_DEV __device__ __forceinline__

/* A uniform base pointer passed through an empty asm: an access through the
 * result forms its 64-bit address where it stands, rather than reusing one
 * computed (and kept live in two VGPRs) for an earlier access to the same
 * element. */
template <typename T>
RYR_DEV T *late_base(T *p)
{
    __asm__ volatile("" : "+s"(p));
    return p;
}

/* ---------------------------------------------------------------- structs */
struct RyrCellInfo {
    int X, Y, Z, N;
    long long off;          /* first slot */
    int first_block, nblocks;
    unsigned long long seed;
    double Ku, Ku2, Kb, svica, svncx, svikr, sviup_tot, Cm, dN;
    long long ncl, noff, ndur;
    int nlcc, line_j, line_k, vclamp_periodic;
    const double *vclamp;   /* V protocol (device array) or null */
    long long vclamp_n;
    int cell, pad2_;        /* index of the cell (per-chunk copies carry it) */
    /* n / X and n / (X*Y) by multiply-high (ryr_fastdiv): multiplier, shifts */
    uint32_t dX_m, dX_s1, dX_s2, dXY_m, dXY_s1, dXY_s2;
    /* variants (RYR_VAR_* of ryr_params) */
    int variants, pad3_;
    long long iso_n;        /* iso: first step with t >= 5.5 CL */
    long long nstim_end;    /* no stimulus (and no beat onset) from this step on; -1: none */
    double sviup_iso;       /* iso: SERCA scale from iso_n on */
    double ci_bath;         /* perm: ci/cs outside the lattice */
};

#define RYR_VAR_PERM      1
#define RYR_VAR_ISO       2
#define RYR_VAR_SLOWNCX   4
#define RYR_VAR_RANDOMLCC 8
/* aux word: nl (bits 0-15), spark flag (16), LCC slots present (17-24) */
#define RYR_AUX_MASK_SHIFT 17

/* iso: fmod(t + 10 CL, 15.5 CL) < dt (cuda_Cell_Detailed.cu:1214) in steps */
RYR_DEV bool iso_active(long long n, long long ncl)
{
    return ncl > 0 && (2 * (n + 10 * ncl)) % (31 * ncl) < 2;
}

/* Division of a 32-bit unsigned n by an invariant d >= 1 (Granlund and
 * Montgomery, PLDI 1994, Fig. 4.1): with l = ceil(log2 d),
 * m = floor(2^32 (2^l - d) / d) + 1 and t = mulhi(m, n),
 * n / d = (t + ((n - t) >> 1)) >> (l - 1); d = 1 is m = 0 with both shifts 0.
 * The constants are computed on the host (ryr_fastdiv_consts). */
RYR_DEV uint32_t ryr_fastdiv(uint32_t n, uint32_t m, uint32_t s1, uint32_t s2)
{
    const uint32_t t = __umulhi(n, m);
    return (t + ((n - t) >> s1)) >> s2;
}

/* membrane state of a cell plus the V-only factors hoisted for its next step */
struct RyrMem {
    double m[13];           /* v xm xh xj xr xs1 xs2 qks xtos ytos xtof ytof xnai */
    double cur[10];         /* csubt xicatto xinacato xina xiks xikr xik1 xinak xitos xitof */
    double stim, out_ica, out_ncx, out_ina;
    long long step;         /* steps completed */
    /* lcc_markov rates (reesica.h:37-59) */
    double alphad, betad, alphaf, betaf;
    /* lcccurrent GHK factors (subroutine_cuda.c:233-243) */
    double ghk_A, ghk_e2, ghk_D;
    int ghk_small, pad_;
    /* ncx factors (subroutine_cuda.c:56-67) */
    double ncx_n1, ncx_n2, ncx_t1, ncx_t3a, ncx_den2;
    /* beat tracker (definitions in the public header, enum ryr_beat_col) */
    long long beat_on;      /* onset step of the beat in progress, -1 before the first */
    int beat_n, beat_flags; /* beats completed; bit0 repolarising, bit1 EAD seen,
                               bit2 V fell after the upstroke, bit3 literal EAD seen */
    double b_vrest, b_vmax, b_vmin, b_apd90, b_ead_t, b_peak_ci, b_jsr, b_sparks, b_eadl_t;
};

struct RyrDD { double hi, lo; };

/* The membrane step of step n split for the one-launch schedule: the state
 * at the start of step n and what the four current groups form from it alone
 * (the gates of step n+1 and the ionic currents of step n, slots RYR_MS_*);
 * [Na]i and V of step n+1 then need only the step's whole-cell sums.  Two
 * per cell, by step parity. */
#define RYR_PEND_MS 24
struct RyrPend {
    double s[13];           /* v xm xh xj xr xs1 xs2 qks xtos ytos xtof ytof xnai at the start of step n */
    double ms[RYR_PEND_MS];
};

/* One per cell: V of a step and its 13 V-only factors (RyrVF order), which
 * the membrane workgroup of the one-launch schedule publishes for the CRU
 * workgroups of the same launch.  The 14 doubles are held as 28 words, each
 * (tag << 32) | one 32-bit half, with tag the step's low 32 bits, written and
 * read with agent-scope 64-bit atomics.  A 64-bit access is single-copy
 * atomic, so a word shows either the old step's half with the old tag or the
 * new one's with the new tag: a reader that finds all 28 tags equal to its
 * step holds that step's values, and the writer needs no ordering between
 * its stores (waiting for them to complete, or a release, took about 2 us). */
#define RYR_VBOX_WORDS 28
struct RyrVBox {
    unsigned long long w[32];
};

/* The V-only factors of the step that the CRU kernel reads (copied from
 * RyrMem at the kernel's start, so that their loads issue with the state
 * loads instead of inside the LCC and NCX branches) */
struct RyrVF {
    double alphad, betad, alphaf, betaf;
    double ghk_A, ghk_e2, ghk_D;
    double ncx_n1, ncx_n2, ncx_t1, ncx_t3a, ncx_den2;
    int ghk_small;
};

RYR_DEV RyrVF load_vf(const RyrMem *M)
{
    RyrVF f;
    f.alphad = M->alphad; f.betad = M->betad; f.alphaf = M->alphaf; f.betaf = M->betaf;
    f.ghk_A = M->ghk_A; f.ghk_e2 = M->ghk_e2; f.ghk_D = M->ghk_D; f.ghk_small = M->ghk_small;
    f.ncx_n1 = M->ncx_n1; f.ncx_n2 = M->ncx_n2; f.ncx_t1 = M->ncx_t1; f.ncx_t3a = M->ncx_t3a;
    f.ncx_den2 = M->ncx_den2;
    return f;
}

/* kernel arguments (identical for every step: the step index lives in RyrMem,
 * so a captured graph replays without new arguments) */
struct RyrKArgs {
    long long P;            /* padded slots */
    double dt;
    int scheme, iso2, spo, ncells;
    void *ci[2], *cnsr[2], *cs[2], *cp[2], *cjsr[2];  /* double-buffered: read buffer = step parity */
    void *catf, *cats, *cacal, *casr, *camyo, *mgmyo;
    void *cjsrb, *Ancx, *casar, *casarh, *casarj, *casarhj;
    void *sb;               /* scheme B: partial cytosol rates of sub-voxels 0,1 (2 planes) */
    uint32_t *lccw, *ryrw, *aux;          /* aux: nl (bits 0-15) | nspark (bit 16) */
    double *d_xire, *d_xinaca;   /* per CRU, written on output steps */
    double *d_up, *d_leak;       /* per sub-voxel (8 planes), written on output steps */
    const RyrCellInfo *cells;
    const RyrCellInfo *chunks;   /* per 256-CRU chunk: a copy of its cell's entry, so a
                                    workgroup reaches its cell parameters in one load */
    long long *gstep;       /* step index a launch's offsets count from (one per handle) */
    RyrMem *mem;
    RyrPend *pend;          /* per cell x 2 (step parity): the one-launch schedule's pending membrane values */
    RyrVBox *vbox;          /* per cell: V published within a launch of the one-launch schedule */
    const int *block_cell;
    RyrDD *partial;         /* whole-cell sums: 3 per wave, 4 waves per chunk; two sets, by step parity */
    unsigned long long *counters;  /* per cell x RYR_C_COUNT */
    double *diag;           /* per cell x 45 */
    long long *diag_step;   /* per cell */
    double *line;           /* per cell x maxX x 22 (offset line_off[c]) */
    const long long *line_off;
    double *beats;          /* per cell x RYR_MAX_BEATS x RYR_BEAT_COLS */
    RyrDD *dpartial;        /* output-row chunk sums: per block x 21 */
    unsigned long long *dcount;   /* output-row integer counts: per chunk x 13 */
};

/* the whole-cell partial sums of step n: the CRU kernel of step n writes set
 * n mod 2, so that a launch can read step n-1's sums while it writes step n's */
RYR_DEV RyrDD *step_partials(const RyrKArgs &a, long long nstep)
{
    return a.partial + (nstep & 1) * 12LL * (a.P >> 8);
}

#define RYR_NCNT 5
#define RYR_MAXB 64
#define RYR_BCOLS 12

/* ------------------------------------------------------------- reductions */
RYR_DEV RyrDD dd_add(RyrDD a, RyrDD b)
{
    RYR_STRICT_FP
    double s = a.hi + b.hi;
    double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b.hi - bb);
    e = e + (a.lo + b.lo);
    RyrDD r;
    r.hi = s + e;
    r.lo = e - (r.hi - s);
    return r;
}

/* A double moved across lanes by a DPP control (a VALU move, no LDS trip). */
template <int CTRL>
RYR_DEV double dpp_f64(double x)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

/* x formed by this point: an empty asm statement that reads it keeps the
 * compiler from sinking its computation past here (it emits nothing) */
RYR_DEV void ryr_pin(double x) { __asm__ volatile("" : : "v"(x)); }

/* Lane `src` (wave-uniform) of x, broadcast through a scalar register. */
RYR_DEV double lanev(double x, int src)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), src);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <int CTRL>
RYR_DEV RyrDD dd_dpp_add(RyrDD v)
{
    RyrDD w;
    w.hi = dpp_f64<CTRL>(v.hi);
    w.lo = dpp_f64<CTRL>(v.lo);
    return dd_add(v, w);
}

/* Sum over the 64 lanes of a wave of a small unsigned count (exact in any
 * order); every lane gets the total. */
RYR_DEV uint32_t wave_sum_u32(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, true);
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 0) + (uint32_t)__builtin_amdgcn_readlane((int)x, 16) +
           (uint32_t)__builtin_amdgcn_readlane((int)x, 32) + (uint32_t)__builtin_amdgcn_readlane((int)x, 48);
}

/* the same for 64-bit values: each DPP step moves both halves and adds
 * with carry */
RYR_DEV unsigned long long wave_sum_u64(unsigned long long x)
{
#define RYR_DPP_U64(CTRL)                                                                              \
    do {                                                                                               \
        const uint32_t lo_ = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, CTRL, 0xF, 0xF, true); \
        const uint32_t hi_ = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(x >> 32), CTRL, 0xF, 0xF, true); \
        x += ((unsigned long long)hi_ << 32) | lo_;                                                    \
    } while (0)
    RYR_DPP_U64(0xB1);
    RYR_DPP_U64(0x4E);
    RYR_DPP_U64(0x141);
    RYR_DPP_U64(0x140);
#undef RYR_DPP_U64
    unsigned long long t = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 16 * r);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), 16 * r);
        t += ((unsigned long long)hi << 32) | lo;
    }
    return t;
}

/* Pairwise tree over the 64 lanes of a wave in index order: level k joins
 * the aligned blocks of 2^(k-1) lanes pairwise.  After a level every lane of
 * a block holds the block's sum, so any lane of the partner block serves as
 * the partner: quad permutations for blocks of 1 and 2, the half-row mirror
 * (lane ^ 7) for 4, the row mirror (lane ^ 15) for 8, and for the last two
 * levels the four 16-lane sums read through scalar registers.  dd_add is
 * commutative bit for bit, so every lane ends with the same value. */
RYR_DEV RyrDD wave_tree(RyrDD v)
{
    v = dd_dpp_add<0xB1>(v);     /* quad_perm [1,0,3,2]: lane ^ 1 */
    v = dd_dpp_add<0x4E>(v);     /* quad_perm [2,3,0,1]: lane ^ 2 */
    v = dd_dpp_add<0x141>(v);    /* row_half_mirror */
    v = dd_dpp_add<0x140>(v);    /* row_mirror */
    RyrDD g0, g1, g2, g3;
    g0.hi = lanev(v.hi, 0);  g0.lo = lanev(v.lo, 0);
    g1.hi = lanev(v.hi, 16); g1.lo = lanev(v.lo, 16);
    g2.hi = lanev(v.hi, 32); g2.lo = lanev(v.lo, 32);
    g3.hi = lanev(v.hi, 48); g3.lo = lanev(v.lo, 48);
    return dd_add(dd_add(g0, g1), dd_add(g2, g3));
}

/* N wave trees level by level, so that their dependent chains interleave;
 * each value gets exactly wave_tree's operations. */
template <int N>
RYR_DEV void wave_trees(RyrDD (&v)[N])
{
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = dd_dpp_add<0xB1>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = dd_dpp_add<0x4E>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = dd_dpp_add<0x141>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = dd_dpp_add<0x140>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        RyrDD g0, g1, g2, g3;
        g0.hi = lanev(v[i].hi, 0);  g0.lo = lanev(v[i].lo, 0);
        g1.hi = lanev(v[i].hi, 16); g1.lo = lanev(v[i].lo, 16);
        g2.hi = lanev(v[i].hi, 32); g2.lo = lanev(v[i].lo, 32);
        g3.hi = lanev(v[i].hi, 48); g3.lo = lanev(v[i].lo, 48);
        v[i] = dd_add(dd_add(g0, g1), dd_add(g2, g3));
    }
}

/* ------------------------------------------------------------- buffers.c */
template <typename T>
RYR_DEV T buf_simple(T b, T c, T dt, T kon, T koff, T B)
{
    if (kon * c * dt > T(1)) return T(0.95) / dt * (B - b) - koff * b;
    return kon * c * (B - b) - koff * b;
}
#define RYR_BSAR (42 * ((0.5 / 8.) * 8.0 / (0.025)))
#define RYR_BSARH (15. * ((0.5 / 8.) * 8.0 / (0.025)))
template <typename T> RYR_DEV T b_tropf(T b, T c, T dt) { return buf_simple<T>(b, c, dt, T(0.0327), T(0.0196), T(70.)); }
template <typename T> RYR_DEV T b_trops(T b, T c, T dt) { return buf_simple<T>(b, c, dt, T(0.00254), T(0.000033), T(134.)); }
template <typename T> RYR_DEV T b_bucal(T b, T c, T dt) { return buf_simple<T>(b, c, dt, T(0.0543), T(0.238), T(24.)); }
template <typename T> RYR_DEV T b_busr(T b, T c, T dt) { return buf_simple<T>(b, c, dt, T(0.1), T(0.06), T(19.)); }
template <typename T> RYR_DEV T b_busar(T b, T c, T dt) { return buf_simple<T>(b, c, dt, T(0.1), T(1.3), T(RYR_BSAR)); }
template <typename T> RYR_DEV T b_busarh(T b, T c, T dt) { return buf_simple<T>(b, c, dt, T(0.1), T(0.03), T(RYR_BSARH)); }
/* busarj/busarhj carry a factor ba = 1 that changes nothing (buffers.c:137-155) */
template <typename T>
RYR_DEV T b_myoca(T CaMyo, T MgMyo, T c, T dt)
{
    if (T(13.8) * c * dt > T(1)) return T(0.95) / dt * (T(140) - CaMyo - MgMyo) - T(0.46) * CaMyo;
    return T(13.8) * c * (T(140) - CaMyo - MgMyo) - T(0.46) * CaMyo;
}
template <typename T>
RYR_DEV T b_myomg(T CaMyo, T MgMyo, T dt)
{
    if (T(0.0157 * 500) * dt > T(1)) return T(0.95) / dt * (T(140) - CaMyo - MgMyo) - T(0.057) * MgMyo;
    return T(0.0157 * 500) * (T(140) - CaMyo - MgMyo) - T(0.057) * MgMyo;
}

/* --------------------------------------------------- SERCA and leak per voxel */
RYR_DEV double uptake_d(double ci, double cnsr)
{
    RYR_STRICT_FP
    double a = ryr_pow(ci / 0.123, 1.787), b = ryr_pow(cnsr / 1700.0, 1.787);
    return 1.0 * 0.3 * (a - b) / (1.0 + a + b);
}
RYR_DEV float powf_fast(float x, float y)
{
    /* x >= 0; x == 0 gives 0 for y > 0 */
    return x > 0.0f ? __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)) : 0.0f;
}
RYR_DEV float uptake_f(float ci, float cnsr)
{
    RYR_EXPR_FP
    float a = powf_fast(ci * (1.0f / 0.123f), 1.787f), b = powf_fast(cnsr * (1.0f / 1700.0f), 1.787f);
    return 0.3f * (a - b) * __builtin_amdgcn_rcpf(1.0f + a + b);
}
template <typename T> RYR_DEV T leak_t(T cnsr, T ci)
{
    return T(0.00001035) * (cnsr - ci) * (cnsr * cnsr) / ((cnsr * cnsr) + T(500.0) * T(500.0));
}

/* -------------------------------------------- hoisted V-only factors (FP64) */
RYR_DEV void vfactors(RyrMem *m, int iso2)
{
    RYR_EXACT_FP
    const double v = m->m[0], xnai = m->m[12];
    /* reesica.h:21-59 */
    double dv5 = 5, dvk = 8, fv5 = -22.8, fvk = 9.1;
    if (iso2) { dv5 = 0; fv5 = -28; fvk = 8.5; }
    double dinf = 1.0 / (1.0 + ryr_exp(-(v - dv5) / dvk));
    double taudin = 1.0 / ((1.0 - ryr_exp(-(v - dv5) / dvk)) / (0.035 * (v - dv5)) * dinf);
    if ((v < 0.0001) && (v > -0.0001)) taudin = 0.035 * dvk / dinf;
    double finf = 1. - 1.0 / (1.0 + ryr_exp(-(v - fv5) / fvk)) / (1. + ryr_exp((v - 60) / 12.));
    double q = 0.0337 * (v + 10.5);
    double taufin = (0.02 - 0.007 * ryr_exp(-(q * q)));
    m->alphad = dinf * taudin;
    m->betad = (1 - dinf) * taudin;
    m->alphaf = (finf) * taufin;
    m->betaf = (1 - finf) * taufin;
    /* subroutine_cuda.c:233-243 */
    double za = v * 96.485 / 8.314 / 308;
    m->ghk_small = __builtin_fabs(za) < 0.001;
    m->ghk_e2 = ryr_exp(2.0 * za);
    m->ghk_A = 4.0 * 11.9 * za * 96.485 * 0.341;
    m->ghk_D = m->ghk_e2 - 1.0;
    /* subroutine_cuda.c:56-67 */
    double e1 = ryr_exp(0.35 * za);
    double e2n = ryr_exp((0.35 - 1.0) * za);
    double r = xnai / 12.3;
    m->ncx_t1 = 0.00359 * ((140.0) * (140.0) * (140.0)) * (1.0 + r * r * r);
    double p3 = xnai * xnai * xnai;
    m->ncx_t3a = (1.3 + (1 * 1.8)) * p3;
    m->ncx_n1 = e1 * p3 * (1 * 1.8);
    m->ncx_n2 = e2n * ((140.0) * (140.0) * (140.0));
    m->ncx_den2 = 1.0 + 0.27 * e2n;
}

/* ---------------------------------------------------------- gating (FP64) */
/* reesica.h:14-116 with the V-only rates of the step precomputed (vfactors).
 * The reference walks the four gates in order: a gate flips if the scaled
 * uniform lies below its rate, else the rate is subtracted and the next gate
 * is tried.  Here the same three subtractions are always formed and the
 * first gate whose comparison holds is picked by selects, so every lane of a
 * wave runs one path whatever its channel's state; the values compared are
 * those of the sequential walk. */
/* the step's parts that need no V: the uniform over dt, and the fCa gate's
 * closing rate */
RYR_DEV double lcc_rig(double ragg, double dt)
{
    RYR_STRICT_FP
    return ragg / dt;
}

RYR_DEV double lcc_betafca(double cpx)
{
    RYR_STRICT_FP
    return 0.175 / (1 + (25 / cpx) * (25 / cpx));
}

template <typename MV>
RYR_DEV int lcc_pick(int i, double rig, double betafca, const MV &m, double betac)
{
    RYR_STRICT_FP
    const double alphac = 0.22, alphafca = 0.012 / 2.;
    const double r1 = (i & 1) ? alphac : betac;
    const double r2 = (i & 2) ? m.alphad : m.betad;
    const double r4 = (i & 4) ? m.alphaf : m.betaf;
    const double r8 = (i & 8) ? alphafca : betafca;
    const bool c1 = rig < r1;
    rig -= r1;
    const bool c2 = rig < r2;
    rig -= r2;
    const bool c4 = rig < r4;
    rig -= r4;
    const bool c8 = rig < r8;
    const int d = c1 ? 1 : c2 ? 2 : c4 ? 4 : c8 ? 8 : 0;
    return (i & d) ? i - d : i + d;
}

template <typename MV>
RYR_DEV int lcc_step(int i, double cpx, double dt, double ragg, const MV &m, double betac)
{
    return lcc_pick(i, lcc_rig(ragg, dt), lcc_betafca(cpx), m, betac);
}

/* Transition t's count once its first uniform u0 (static slot 8 + t) has
 * failed the early-out (subroutine_cuda.c:332-363 and its seven copies):
 * Knuth's product-of-uniforms Poisson count where ryr_knuth_range holds,
 * else the dithered Gaussian with rejection.  The further uniforms come from
 * t's own substream, so the two samplers can run in any order; *nd counts
 * them. */
RYR_DEV bool ryr_knuth_range(int n, double p) { return n <= 1 || p < 0.2 || (n <= 5 && p < 0.3); }

RYR_DEV int ryr_knuth_job(uint64_t seed, uint32_t cru, uint64_t step, int t, int n, double p, double u0, int *nd)
{
    RYR_STRICT_FP
    uint32_t j = 0;
    ryr_u32x4 w = {{0u, 0u, 0u, 0u}};
    const double lam = ryr_exp(-(n) * p);
    int kk = 1;
    double puu = u0;
    while (puu >= lam && kk < 195) {
        kk++;
        if (!(j & 1u)) w = ryr_block(seed, cru, step, ryr_sub_block(t, j));
        const double re = (j & 1u) ? ryr_slot_uniform(w.v[2], w.v[3]) : ryr_slot_uniform(w.v[0], w.v[1]);
        ++j;
        puu = puu * (re);
    }
    *nd += (int)j;
    return kk - 1;
}

RYR_DEV int ryr_gauss_job(uint64_t seed, uint32_t cru, uint64_t step, int t, int n, double p, double u0, int *fails,
                          int *nd)
{
    RYR_STRICT_FP
    int out = -1;
    uint32_t j = 0;
    ryr_u32x4 w = {{0u, 0u, 0u, 0u}};
    int kk = 0;
    while (out < 0) {
        double u[3];
        uint32_t r = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (q == 0 && kk == 0) { u[0] = u0; continue; }
            if (!(j & 1u)) w = ryr_block(seed, cru, step, ryr_sub_block(t, j));
            const uint32_t wa = (j & 1u) ? w.v[2] : w.v[0], wb = (j & 1u) ? w.v[3] : w.v[1];
            ++j;
            u[q] = ryr_slot_uniform(wa, wb);
            r = wb;
        }
        out = (int)(__builtin_floor(n * p + __builtin_sqrt(n * p * (1.0 - p)) *
                                    __builtin_sqrt(-2.0 * ryr_log(1.0 - u[0])) *
                                    ryr_cos(2.0 * 3.1415926 * u[1])) + r % 2);
        kk++;
        if (kk > 200) { out = 0; (*fails)++; }
    }
    *nd += (int)j;
    return out;
}

/* CSQN monomer fraction (subroutine_cuda.c:170-171, :328-329): the same
 * expression feeds the unbound->bound rate aub (divided by taub = 1, exact)
 * and the luminal buffer factor, so it is evaluated once per CRU-step. */
RYR_DEV double csqn_mono(double cjsr, double *roo2_out)
{
    RYR_STRICT_FP
    double roo2 = 5000.0 / (1.0 + ryr_pow(850.0 / (cjsr), 23.0));
    *roo2_out = roo2;
    return (-1.0 + __builtin_sqrt(1.0 + 8.0 * 460.0 * roo2)) / (4.0 * roo2 * 460.0);
}

/* Per-RyR transition probabilities of one step (subroutine_cuda.c:251-330),
 * in the order OU->CU, CU->OU, OU->OB, OB->OU, CB->CU, CU->CB, OB->CB,
 * CB->OB. */
RYR_DEV void ryr_rates(double Ku, double Ku2, double Kb, double cp, double cjsr, double mono, double dt, double pt[8])
{
    RYR_STRICT_FP
    double q1 = (5000 / cjsr) * (5000 / cjsr);
    double x2 = Ku / cp;
    double q2 = x2 * x2;
    double pku = 4.5 * 1.5 * Ku2 * 1 / (1 + q1) * 1 / (1 + q2);
    double pkb = 4.5 * 1.5 / 36. * Kb * 1 / (1 + q1) * 1 / (1 + q2);
    double pkuminus = 1. / 1., pkbminus = 1. / 1.;
    if (pku * dt > 1.0) pku = 1.0 / dt;
    if (pkb * dt > 1.0) pkb = 1.0 / dt;
    if (pkuminus * dt > 1.0) pkuminus = 1.0 / dt;
    if (pkbminus * dt > 1.0) pkbminus = 1.0 / dt;
    double aub = mono / (1.0);
    double abu = 1.0 / (0.2 * 5 * 220.);
    const double pau = (pkb < 1e-16) ? 0 : aub * dt;
    const double pbu = (pkb < 1e-16) ? abu * dt * 36 : abu * dt * (pku / pkb);
    pt[0] = pkuminus * dt; pt[1] = pku * dt; pt[2] = pau; pt[3] = pbu;
    pt[4] = abu * dt; pt[5] = aub * dt; pt[6] = pkbminus * dt; pt[7] = pkb * dt;
}

/* The partner lane of a CRU that two lanes share (lane ^ 1, a quad
 * permutation). */
RYR_DEV uint32_t lane_pair_u32(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, true);
}

/* Streamed state (read once and written once per step, never re-read by a
 * neighbour): nontemporal loads and stores, which leave the L2 to the values
 * the stencils re-read (ci and cnsr of the step's start, cp, cs and cjsr). */
template <typename T> RYR_DEV T ryr_ld_nt(const T *p) { return __builtin_nontemporal_load(p); }
template <typename T> RYR_DEV void ryr_st_nt(T *p, T v) { __builtin_nontemporal_store(v, p); }
/* the same where NT holds, plain otherwise (the CRU kernel's per-CRU state:
 * nontemporal in the instances batches run, which fill the chip; plain in
 * the single cell's, where the hint measured 1% slower) */
template <bool NT, typename T> RYR_DEV T ryr_ld_if(const T *p) { if constexpr (NT) return ryr_ld_nt(p); else return *p; }
template <bool NT, typename T> RYR_DEV void ryr_st_if(T *p, T v) { if constexpr (NT) ryr_st_nt(p, v); else *p = v; }

/* The value x of lane src of the wave (ds_bpermute; src taken modulo 64):
 * every lane of the wave must execute it. */
RYR_DEV float ryr_lane_from(float x, int src)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute((src & 63) << 2, __float_as_int(x)));
}
RYR_DEV double ryr_lane_from(double x, int src)
{
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_ds_bpermute((src & 63) << 2, (int)v);
    const int hi = __builtin_amdgcn_ds_bpermute((src & 63) << 2, (int)(v >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

/* Lane Q of this lane's quad (quad_perm [Q,Q,Q,Q]). */
template <int Q>
RYR_DEV uint32_t quad_lane_u32(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, Q * 0x55, 0xF, 0xF, true);
}

/* subroutine_cuda.c:251-669; counts in/out: nou, ncu, nob, ncb; mono =
 * csqn_mono(cjsr); *nd counts the draws.  With LPC = 2 or 4 adjacent lanes
 * share the CRU: lane h samples the transitions 8h/LPC .. 8(h+1)/LPC - 1 and
 * the counts are gathered before they are combined.
 *
 * The eight transitions draw from their own addresses (philox.h), so they are
 * sampled in two passes rather than in program order.  Pass 1 draws every
 * first uniform (static slots 8..15, four Philox blocks for all lanes at
 * once) and settles the transitions whose count is 0 without the
 * exponential: an empty population or zero rate (exp(-0) = 1), or a first
 * draw below 1 - x, x = n p (1 - 2x below x = 2^-20), which lies under the
 * rounded exp(-x) by more than the rounding of either.  Pass 2 is a loop in
 * which each lane takes its next unsettled transition, so one exponential and
 * one Knuth loop per round serve every lane, whichever transitions they
 * have left.  The counts equal those of the sequential reference loop. */
template <int LPC = 1>
RYR_DEV void ryr_gating(uint64_t seed, uint32_t cru, uint64_t step, double Ku, double Ku2, double Kb, double cp,
                        double cjsr, double mono, double dt, int &nou, int &ncu, int &nob, int &ncb, int *fails,
                        int *nd, int h = 0)
{
    RYR_STRICT_FP
    double pt[8];
    ryr_rates(Ku, Ku2, Kb, cp, cjsr, mono, dt, pt);
    /* source population of transition t: 0 OU, 1 CU, 2 OB, 3 CB */
    const int pop[4] = {nou, ncu, nob, ncb};
    constexpr int src[8] = {0, 1, 0, 2, 3, 1, 2, 3};
    /* this lane's transitions t = TPL h + l, l < TPL: all eight, or with two
     * lanes per CRU one half each (two of the four first-uniform blocks) */
    constexpr int TPL = 8 / LPC;
    double u0[TPL], pl[TPL];
    int nl[TPL];
#pragma unroll
    for (int l = 0; l < TPL; ++l) {
        pl[l] = pt[l];
        nl[l] = pop[src[l]];
#pragma unroll
        for (int hh = 1; hh < LPC; ++hh)
            if (h == hh) { pl[l] = pt[TPL * hh + l]; nl[l] = pop[src[TPL * hh + l]]; }
    }
#pragma unroll
    for (int b = 0; b < TPL / 2; ++b) {
        const ryr_u32x4 w = ryr_block(seed, cru, step, (uint32_t)(RYR_SLOT_RYR0 / 2 + h * (TPL / 2) + b));
        u0[2 * b] = ryr_slot_uniform(w.v[0], w.v[1]);
        u0[2 * b + 1] = ryr_slot_uniform(w.v[2], w.v[3]);
    }
    uint32_t pend = 0u;
#pragma unroll
    for (int l = 0; l < TPL; ++l) {
        const int n = nl[l];
        const double p = pl[l];
        if (n <= 1 || p < 0.2 || (n <= 5 && p < 0.3)) {
            const double x = (double)(n) * p;
            const double thr = (x >= 0x1p-20) ? 1.0 - x : 1.0 - 2.0 * x;
            const bool zero = (n == 0 || p == 0.0) || (x >= 0x1p-40 && u0[l] < thr);
            if (!zero) pend |= 1u << l;
        } else {
            pend |= 1u << l;
        }
    }
    /* counts in byte fields (Knuth counts stay below 195 and Gaussian ones
     * below n p + 5 sqrt(-2 log 2^-53) + 1 < 150) */
    uint32_t kw[LPC == 1 ? 2 : 1] = {};
    int ndraw = TPL;
    /* pending transitions in Knuth's range; the Gaussian ones (n >= 2 and
     * p >= 0.2, rare at the model's time step) are deferred to a pass after
     * this loop, which re-forms their rates and first uniforms, so that the
     * Gaussian sampler's registers do not add to the loop's */
    uint32_t gpend = 0u;
#pragma unroll
    for (int l = 0; l < TPL; ++l)
        if (((pend >> l) & 1u) && !ryr_knuth_range(nl[l], pl[l])) gpend |= 1u << l;
    pend &= ~gpend;
    while (pend) {
        const int l = __builtin_ctz(pend);
        pend &= pend - 1u;
        int n = nl[0];
        double p = pl[0], u = u0[0];
#pragma unroll
        for (int q = 1; q < TPL; ++q)
            if (l == q) { n = nl[q]; p = pl[q]; u = u0[q]; }
        const uint32_t k = (uint32_t)ryr_knuth_job(seed, cru, step, TPL * h + l, n, p, u, &ndraw) & 255u;
        if (LPC == 1 && l >= 4) kw[LPC == 1 ? 1 : 0] |= k << (8 * (l - 4));
        else kw[0] |= k << (8 * (l & 3));
    }
    while (gpend) {
        const int l = __builtin_ctz(gpend);
        gpend &= gpend - 1u;
        const int t = TPL * h + l;
        double pg[8];
        ryr_rates(Ku, Ku2, Kb, cp, cjsr, mono, dt, pg);
        const int popg[4] = {nou, ncu, nob, ncb};
        int n = popg[src[0]];
        double p = pg[0];
#pragma unroll
        for (int q = 1; q < 8; ++q)
            if (t == q) { n = popg[src[q]]; p = pg[q]; }
        const ryr_u32x4 w = ryr_block(seed, cru, step, (uint32_t)(RYR_SLOT_RYR0 / 2 + (t >> 1)));
        const double u = (t & 1) ? ryr_slot_uniform(w.v[2], w.v[3]) : ryr_slot_uniform(w.v[0], w.v[1]);
        const uint32_t k = (uint32_t)ryr_gauss_job(seed, cru, step, t, n, p, u, fails, &ndraw) & 255u;
        if (LPC == 1 && l >= 4) kw[LPC == 1 ? 1 : 0] |= k << (8 * (l - 4));
        else kw[0] |= k << (8 * (l & 3));
    }
    *nd += ndraw;
    uint32_t klo, khi;
    if constexpr (LPC == 1) {
        klo = kw[0];
        khi = kw[LPC == 1 ? 1 : 0];
    } else if constexpr (LPC == 2) {
        /* lane 0 holds transitions 0-3, lane 1 transitions 4-7 */
        const uint32_t other = lane_pair_u32(kw[0]);
        klo = h ? other : kw[0];
        khi = h ? kw[0] : other;
    } else {
        /* lane q of the quad holds transitions 2q, 2q + 1 in its low 16 bits */
        klo = quad_lane_u32<0>(kw[0]) | (quad_lane_u32<1>(kw[0]) << 16);
        khi = quad_lane_u32<2>(kw[0]) | (quad_lane_u32<3>(kw[0]) << 16);
    }

#define RYR_KT(t) ((int)((((t) < 4 ? klo : khi) >> (8 * ((t) & 3))) & 255u))
    int n_ou_cu = RYR_KT(0) > RYR_NRYR ? RYR_NRYR : RYR_KT(0);
    int n_cu_ou = RYR_KT(1) > RYR_NRYR ? RYR_NRYR : RYR_KT(1);
    int n_ou_ob = RYR_KT(2) > RYR_NRYR ? RYR_NRYR : RYR_KT(2);
    int n_ob_ou = RYR_KT(3) > RYR_NRYR ? RYR_NRYR : RYR_KT(3);
    int n_cb_cu = RYR_KT(4) > RYR_NRYR ? RYR_NRYR : RYR_KT(4);
    int n_cu_cb = RYR_KT(5) > RYR_NRYR ? RYR_NRYR : RYR_KT(5);
    int n_ob_cb = RYR_KT(6);
    int n_cb_ob = RYR_KT(7) > RYR_NRYR ? RYR_NRYR : RYR_KT(7);
#undef RYR_KT

    if (n_ou_ob + n_ou_cu > nou) {
        if (n_ou_cu >= n_ou_ob) n_ou_cu = 0; else n_ou_ob = 0;
        if (n_ou_ob > nou) n_ou_ob = 0;
        else if (n_ou_cu > nou) n_ou_cu = 0;
    }
    if (n_ob_ou + n_ob_cb > nob) {
        if (n_ob_ou >= n_ob_cb) n_ob_ou = 0; else n_ob_cb = 0;
        if (n_ob_cb > nob) n_ob_cb = 0;
        else if (n_ob_ou > nob) n_ob_ou = 0;
    }
    if (n_cu_ou + n_cu_cb > ncu) {
        if (n_cu_cb >= n_cu_ou) n_cu_cb = 0; else n_cu_ou = 0;
        if (n_cu_ou > ncu) n_cu_ou = 0;
        else if (n_cu_cb > ncu) n_cu_cb = 0;
    }
    nou += -n_ou_ob - n_ou_cu + n_ob_ou + n_cu_ou;
    nou = nou < 0 ? 0 : (nou > RYR_NRYR ? RYR_NRYR : nou);
    nob += -n_ob_ou - n_ob_cb + n_ou_ob + n_cb_ob;
    nob = nob < 0 ? 0 : (nob > RYR_NRYR ? RYR_NRYR : nob);
    ncu += -n_cu_ou - n_cu_cb + n_ou_cu + n_cb_cu;
    ncu = ncu < 0 ? 0 : (ncu > RYR_NRYR ? RYR_NRYR : ncu);
    ncb = RYR_NRYR - nou - nob - ncu;
    ncb = ncb < 0 ? 0 : (ncb > RYR_NRYR ? RYR_NRYR : ncb);
}

/* ------------------------------------ membrane ODE (FP64, up to four waves) */
/* cuda_Cell_Detailed.cu:447-739: whole-cell currents from the three sums,
 * gates by Rush-Larsen, [Na]i and V by forward Euler.
 *
 * The currents fall into four independent groups: I_Na; I_Kr and I_Ks; I_K1
 * and I_to; I_NaK with the scaling of the Ca-current sums.  Group g runs on
 * wavefront g mod NW of the membrane workgroup, so with four waves the four
 * groups run side by side on the four SIMDs.  Inside a group the work is
 * spread over the lanes by dependency level: the divisions of one level (the
 * exponent arguments, then the rates, then the time constants and steady
 * states, then -dt/tau) are one division per lane, each lane picking its own
 * numerator and denominator, and the exponentials of one level are one ryr_exp
 * per lane.  Values are broadcast through scalar registers.  Each expression
 * is the reference's, in its operand order, so no result depends on which
 * wave or lane computed it.  Each group leaves its new gates and currents in
 * shared slots RYR_MS_*; every wave then forms [Na]i and V from them. */
#define RYR_MS_N 24
static_assert(RYR_MS_N == RYR_PEND_MS, "pending membrane slots");
enum {
    RYR_MS_V = 0, RYR_MS_XM, RYR_MS_XH, RYR_MS_XJ, RYR_MS_XR, RYR_MS_XS1, RYR_MS_XS2, RYR_MS_QKS,
    RYR_MS_XTOS, RYR_MS_YTOS, RYR_MS_XTOF, RYR_MS_YTOF, RYR_MS_NAI,
    RYR_MS_CSUBT, RYR_MS_ICAT, RYR_MS_INACA, RYR_MS_INA, RYR_MS_IKS, RYR_MS_IKR, RYR_MS_IK1, RYR_MS_INAK,
    RYR_MS_ITOS, RYR_MS_ITOF
};

struct MemIn {
    const double *s;      /* v xm xh xj xr xs1 xs2 qks xtos ytos xtof ytof xnai */
    double svikr, Cm, dN, sum_icat, sum_inaca, sum_cs, dt;
    bool slowncx;
};

#define RYR_XKI 140.0
#define RYR_XKO 5.40
#define RYR_FRT (96.485 / 8.314 / 308.0)

/* I_Na (:460-502): m, h, j */
RYR_DEV void mem_ina(const MemIn &I, int lane, double *sh)
{
    RYR_EXACT_FP
    const double v = I.s[0], xm = I.s[1], xh = I.s[2], xj = I.s[3], xnai = I.s[12], dt = I.dt;
    const double frt = RYR_FRT;
    const bool lo = v < -40.0;
    double n = 1.0, d = 1.0;
    switch (lane) {
    case 1: n = -v; d = 11.0; break;
    case 2: n = 80.0 + v; d = -6.8; break;
    case 10: n = v + 10.66; d = -11.10; break;
    case 13: n = 140.0; d = xnai; break;
    default: break;
    }
    const double q1 = n / d;
    double arg = 0.0;
    switch (lane) {
    case 0: arg = -0.1 * (v + 47.13); break;
    case 1: case 2: case 10: arg = q1; break;
    case 3: arg = 0.079 * v; break;
    case 4: arg = 0.35 * v; break;
    case 5: arg = 0.2444 * v; break;
    case 6: arg = -0.04391 * v; break;
    case 7: arg = 0.311 * (v + 79.23); break;
    case 8: arg = -0.01052 * v; break;
    case 9: arg = -0.1378 * (v + 40.14); break;
    case 11: arg = -0.00000025350 * v; break;
    case 12: arg = -0.10 * (v + 32.00); break;
    default: break;
    }
    const double ex = ryr_exp(arg);
    const double lg = ryr_log(lane == 13 ? q1 : 1.0);
    const double e0 = lanev(ex, 0), e1 = lanev(ex, 1);

    /* rates: am; bj; aj's V factor (V < -40) or bh */
    n = 1.0; d = 1.0;
    if (lane == 0) { n = 0.32 * (v + 47.13); d = 1.0 - e0; }
    if (lo) {
        const double e7 = lanev(ex, 7), e8 = lanev(ex, 8), e9 = lanev(ex, 9);
        if (lane == 1) { n = 0.1212 * e8; d = 1.0 + e9; }
        if (lane == 2) { n = v + 37.78; d = 1.0 + e7; }
    } else {
        const double e10 = lanev(ex, 10), e11 = lanev(ex, 11), e12 = lanev(ex, 12);
        if (lane == 1) { n = 0.3 * e11; d = 1.0 + e12; }
        if (lane == 2) { n = 1.00; d = 0.130 * (1.00 + e10); }
    }
    const double q2 = n / d;
    const double am = lanev(q2, 0);
    const double bm = 0.08 * e1;
    const double bj = lanev(q2, 1);
    double ah, bh, aj;
    if (lo) {
        ah = 0.135 * lanev(ex, 2);
        bh = 3.56 * lanev(ex, 3) + 310000.0 * lanev(ex, 4);
        aj = (-127140.0 * lanev(ex, 5) - 0.00003474 * lanev(ex, 6)) * lanev(q2, 2);
    } else {
        ah = 0.00;
        bh = lanev(q2, 2);
        aj = 0.00;
    }
    /* time constants (lanes 0-2) and steady states (lanes 3-5) */
    n = 1.0; d = 1.0;
    switch (lane) {
    case 0: n = 1.00; d = ah + bh; break;
    case 1: n = 1.00; d = aj + bj; break;
    case 2: n = 1.00; d = am + bm; break;
    case 3: n = ah; d = ah + bh; break;
    case 4: n = aj; d = aj + bj; break;
    case 5: n = am; d = am + bm; break;
    default: break;
    }
    const double q3 = n / d;
    const double a4 = -dt / q3;
    const double ex2 = ryr_exp(lane < 3 ? a4 : 0.0);

    const double ena = (1.0 / frt) * lanev(lg, 13);
    const double gna = 15;
    const double gnaleak = 0.3e-3 * 5;
    const double xina = gna * xh * xj * xm * xm * xm * (v - ena) + gnaleak * (v - ena);
    const double hinf = lanev(q3, 3), jinf = lanev(q3, 4), minf = lanev(q3, 5);
    const double xh1 = hinf - (hinf - xh) * lanev(ex2, 0);
    const double xj1 = jinf - (jinf - xj) * lanev(ex2, 1);
    const double xm1 = minf - (minf - xm) * lanev(ex2, 2);
    if (lane == 0) {
        sh[RYR_MS_XM] = xm1; sh[RYR_MS_XH] = xh1; sh[RYR_MS_XJ] = xj1; sh[RYR_MS_INA] = xina;
    }
}

/* I_Kr (:507-529) and I_Ks (:534-559): xr, xs1, xs2 (qks, which needs the
 * sub-membrane mean, in mem_sums) */
RYR_DEV void mem_ikr_iks(const MemIn &I, int lane, double *sh)
{
    RYR_EXACT_FP
    const double v = I.s[0], xr = I.s[4], xs1 = I.s[5], xs2 = I.s[6], qks = I.s[7], xnai = I.s[12];
    const double dt = I.dt;
    const double xki = RYR_XKI, xko = RYR_XKO;
    const double frt = RYR_FRT;
    const double ek = (1.00 / frt) * ryr_log(xko / xki);
    double n = 1.0, d = 1.0;
    switch (lane) {
    case 2: n = -(v + 50.00); d = 7.50; break;
    case 3: n = v + 33.00; d = 22.40; break;
    case 4: n = -(v - 1.500); d = 16.700; break;
    case 7: n = xko + 0.0183300 * 140.0; d = xki + 0.0183300 * xnai; break;
    default: break;
    }
    const double q1 = n / d;
    double arg = 0.0;
    switch (lane) {
    case 0: arg = -0.123 * (v + 7.00); break;
    case 1: arg = 0.1450 * (v + 10.00); break;
    case 2: case 3: case 4: arg = q1; break;
    case 5: arg = -0.1480 * (v + 30.0); break;
    case 6: arg = 0.06870 * (v + 30.00); break;
    default: break;
    }
    const double ex = ryr_exp(arg);
    const double lg = ryr_log(lane == 7 ? q1 : 1.0);

    /* rates and steady states */
    const double e = ex;                         /* lane k uses its own exponential */
    n = 1.0; d = 1.0;
    switch (lane) {
    case 0: n = 0.001380 * (v + 7.00); d = 1.0 - e; break;
    case 1: n = 0.000610 * (v + 10.00); d = e - 1.00; break;
    case 2: n = 1.00; d = 1.00 + e; break;
    case 3: n = 1.00; d = 1.00 + e; break;
    case 4: n = 1.0; d = 1.0 + e; break;
    case 5: n = 0.0000719 * (v + 30.00); d = 1.00 - e; break;
    case 6: n = 0.0001310 * (v + 30.00); d = e - 1.00; break;
    default: break;
    }
    const double q2 = n / d;
    const double xkrv1 = lanev(q2, 0), xkrv2 = lanev(q2, 1);
    const double tx1 = lanev(q2, 5), tx2 = lanev(q2, 6);
    n = 1.0; d = 1.0;
    switch (lane) {
    case 0: n = 1.00; d = xkrv1 + xkrv2; break;
    case 1: n = 1.00; d = tx1 + tx2; break;
    default: break;
    }
    const double q3 = n / d;
    const double taukr = lanev(q3, 0);
    double tauxs = lanev(q3, 1);
    tauxs = 0.95 * tauxs;
    n = 1.0; d = 1.0;
    switch (lane) {
    case 0: n = -dt; d = taukr; break;
    case 1: n = -dt; d = tauxs; break;
    default: break;
    }
    const double q4 = n / d;
    const double ex2 = ryr_exp(lane < 2 ? q4 : 0.0);

    const double gss = __builtin_sqrt(xko / 5.40);
    const double xkrinf = lanev(q2, 2);
    const double rg = lanev(q2, 3);
    const double gkr = 0.0078360;
    const double xikr = I.svikr * gkr * gss * xr * rg * (v - ek);
    const double eks = (1.00 / frt) * lanev(lg, 7);
    const double xs1ss = lanev(q2, 4);
    const double gksx = 0.2000;
    const double sviks_iso = 1.;
    const double xiks = 1.0 * sviks_iso * gksx * qks * xs1 * xs2 * (v - eks);
    const double xr1 = xkrinf - (xkrinf - xr) * lanev(ex2, 0);
    const double xs11 = xs1ss - (xs1ss - xs1) * lanev(ex2, 1);
    const double xs21 = xs1ss - (xs1ss - xs2) * lanev(ex2, 1);
    if (lane == 0) {
        sh[RYR_MS_XR] = xr1; sh[RYR_MS_XS1] = xs11; sh[RYR_MS_XS2] = xs21;
        sh[RYR_MS_IKR] = xikr; sh[RYR_MS_IKS] = xiks;
    }
}

/* I_K1 (:590-600), I_to,s and I_to,f (:604-659) */
RYR_DEV void mem_ik1_ito(const MemIn &I, int lane, double *sh)
{
    RYR_EXACT_FP
    const double v = I.s[0], xtos = I.s[8], ytos = I.s[9], xtof = I.s[10], ytof = I.s[11], dt = I.dt;
    const double xki = RYR_XKI, xko = RYR_XKO;
    const double frt = RYR_FRT;
    const double ek = (1.00 / frt) * ryr_log(xko / xki);
    /* rt1, rt2 (= rt5), rt3 and v/30 */
    double n = 1.0, d = 1.0;
    switch (lane) {
    case 0: n = -(v + 3.0); d = 15.00; break;
    case 1: n = v + 33.5; d = 10.00; break;
    case 2: n = v + 60.00; d = 10.00; break;
    case 3: n = v; d = 30.00; break;
    default: break;
    }
    const double q1 = n / d;
    const double rt1 = lanev(q1, 0), rt2 = lanev(q1, 1), rt3 = lanev(q1, 2), v30 = lanev(q1, 3);
    const double rt4 = -(v30) * (v30);
    double arg = 0.0;
    switch (lane) {
    case 0: arg = 0.2385 * (v - ek - 59.215); break;
    case 1: arg = 0.08032 * (v - ek + 5.476); break;
    case 2: arg = 0.061750 * (v - ek - 594.31); break;
    case 3: arg = -0.5143 * (v - ek + 4.753); break;
    case 4: arg = rt1; break;
    case 5: arg = rt2; break;
    case 6: arg = -rt1; break;
    case 7: arg = rt3; break;
    case 8: arg = rt4; break;
    case 9: arg = rt2; break;            /* rt5 = (v + 33.50) / 10.00 = rt2 */
    default: break;
    }
    const double ex = ryr_exp(arg);
    const double e1 = lanev(ex, 1), e2 = lanev(ex, 2);
    n = 1.0; d = 1.0;
    switch (lane) {
    case 3: n = 0.49124 * e1 + e2; d = 1.0 + ex; break;           /* bki */
    case 0: n = 1.02; d = 1.0 + ex; break;                         /* aki */
    case 4: n = 1.00; d = 1.0 + ex; break;                         /* xtos_inf */
    case 5: n = 1.00; d = 1.00 + ex; break;                        /* ytos_inf = rs_inf */
    case 6: n = 9.00; d = 1.00 + ex; break;                        /* txs */
    case 7: n = 3000.00; d = 1.0 + ex; break;                      /* tys */
    case 9: n = 20.0; d = 1.0 + ex; break;                         /* tyf */
    default: break;
    }
    const double q2 = n / d;
    const double aki = lanev(q2, 0), bki = lanev(q2, 3);
    const double txs = lanev(q2, 6) + 0.50;
    const double tys = lanev(q2, 7) + 30.00;
    const double txf = 3.50 * lanev(ex, 8) + 1.50;
    const double tyf = lanev(q2, 9) + 20.00;
    n = 1.0; d = 1.0;
    switch (lane) {
    case 0: n = aki; d = aki + bki; break;
    case 1: n = -dt; d = txs; break;
    case 2: n = -dt; d = tys; break;
    case 3: n = -dt; d = txf; break;
    case 4: n = -dt; d = tyf; break;
    default: break;
    }
    const double q3 = n / d;
    const double ex2 = ryr_exp((lane >= 1 && lane <= 4) ? q3 : 0.0);

    const double gkix = 0.600;
    const double gki = gkix * (__builtin_sqrt(xko / 5.4));
    const double xkin = lanev(q3, 0);
    const double xik1 = (1.0) * gki * xkin * (v - ek);
    const double xtos_inf = lanev(q2, 4);
    const double ytos_inf = lanev(q2, 5);
    const double rs_inf = ytos_inf;
    const double gtos = 0.040;
    const double xitos = 1.0 * gtos * xtos * (ytos + 0.50 * rs_inf) * (v - ek);
    const double xtof_inf = xtos_inf;
    const double ytof_inf = ytos_inf;
    const double gtof = 0.10;
    const double xitof = 1.0 * gtof * xtof * ytof * (v - ek);
    const double xtos1 = xtos_inf - (xtos_inf - xtos) * lanev(ex2, 1);
    const double ytos1 = ytos_inf - (ytos_inf - ytos) * lanev(ex2, 2);
    const double xtof1 = xtof_inf - (xtof_inf - xtof) * lanev(ex2, 3);
    const double ytof1 = ytof_inf - (ytof_inf - ytof) * lanev(ex2, 4);
    if (lane == 0) {
        sh[RYR_MS_XTOS] = xtos1; sh[RYR_MS_YTOS] = ytos1; sh[RYR_MS_XTOF] = xtof1; sh[RYR_MS_YTOF] = ytof1;
        sh[RYR_MS_IK1] = xik1; sh[RYR_MS_ITOS] = xitos; sh[RYR_MS_ITOF] = xitof;
    }
}

/* I_NaK (:665-676) and the whole-cell Ca currents from the CRU sums */
RYR_DEV void mem_inak(const MemIn &I, int lane, double *sh)
{
    RYR_EXACT_FP
    const double v = I.s[0], xnai = I.s[12];
    const double xko = RYR_XKO;
    const double frt = RYR_FRT;
    const double xkmko = 1.50, xkmnai = 12.00, xibarnak = 1.5000, hh = 1.00;
    const double q1 = xkmnai / xnai;
    double arg = 0.0;
    switch (lane) {
    case 0: arg = 140.0 / 67.30; break;
    case 1: arg = -0.1 * v * frt; break;
    case 2: arg = -v * frt; break;
    default: break;
    }
    const double ex = ryr_exp(arg);
    double n = 1.0, d = 1.0;
    switch (lane) {
    case 0: n = ex - 1.00; d = 7.00; break;                                 /* sigma */
    case 1: n = 1.0; d = 1.0 + ryr_pow(q1, hh); break;
    default: break;
    }
    const double q2 = n / d;
    const double sigma = lanev(q2, 0);
    const double fnak = 1.00 / (1.0 + 0.1245 * lanev(ex, 1) + 0.0365 * sigma * lanev(ex, 2));
    const double xinak = 1.0 * xibarnak * fnak * lanev(q2, 1) * xko / (xko + xkmko);
    if (lane == 0) sh[RYR_MS_INAK] = xinak;
}

/* The parts that need the whole-cell sums of the CRU kernel: the
 * sub-membrane mean csubt and the I_Ks Ca gate qks (:534-559), and the
 * whole-cell Ca currents I_CaL and I_NCX (:447-456) */
RYR_DEV void mem_sums(const MemIn &I, int lane, double *sh)
{
    RYR_EXACT_FP
    const double qks = I.s[7], dt = I.dt;
    double n = 1.0, d = 1.0;
    switch (lane) {
    case 0: n = I.sum_cs; d = I.dN; break;
    case 1: n = I.sum_icat; d = I.Cm; break;
    case 2: n = I.sum_inaca; d = I.Cm; break;
    default: break;
    }
    const double q1 = n / d;
    const double csubt = lanev(q1, 0);
    /* slowncxchange (:540-543) doubles the Ca dependence and the time constant */
    const double qks_inf = I.slowncx ? 0.2 * (1 + 1.6 / (1 + ryr_pow((0.28 / csubt), 3)))
                                     : 0.2 * (1 + 0.8 / (1 + ryr_pow((0.28 / csubt), 3)));
    const double tauqks = I.slowncx ? 2000 : 1000.00;
    const double qks1 = qks + dt * (qks_inf - qks) / tauqks;
    const double xicatto = lanev(q1, 1) * 0.0965 * (0.00126) * 2.0 * ((0.5) + 1.);
    const double xinacato = lanev(q1, 2) * 0.0965 * (0.025);
    /* Every lane holds the same four values and writes them: the lanes of
     * this wave read these slots next with no barrier in between, so each
     * lane must read its own write (a write by lane 0 alone would be a race
     * between divergent branches that the compiler may order either way). */
    sh[RYR_MS_CSUBT] = csubt; sh[RYR_MS_QKS] = qks1; sh[RYR_MS_ICAT] = xicatto; sh[RYR_MS_INACA] = xinacato;
}

/* [Na]i and V (:701-738) from the groups' currents; fills slots V and NAI */
RYR_DEV void mem_finish(const double *s, const double *sh, double dt, double stim, double *v_out,
                        double *nai_out)
{
    RYR_EXACT_FP
    const double xina = sh[RYR_MS_INA], xik1 = sh[RYR_MS_IK1], xikr = sh[RYR_MS_IKR], xiks = sh[RYR_MS_IKS];
    const double xitos = sh[RYR_MS_ITOS], xitof = sh[RYR_MS_ITOF], xinacato = sh[RYR_MS_INACA];
    const double xicatto = sh[RYR_MS_ICAT], xinak = sh[RYR_MS_INAK];
    const double wcainv = 1.0 / 50.0;
    const double conv = 0.18 * 12500;
    const double xinaca = xinacato;
    const double trick = 1.0;
    const double xrr = trick * (1.0 / wcainv / conv) / 1000.0;
    const double dnai = -xrr * (xina + 3.0 * xinak + 3.0 * xinaca);
    double xnai = s[12];
    xnai += dnai * dt;
    const double dvh = -(xina + xik1 + 1 * xikr + 1 * xiks + xitos + 1 * xitof + 1. * xinacato + 1 * xicatto + xinak) + stim;
    double v = s[0];
    v += dvh * dt;
    *v_out = v;
    *nai_out = xnai;
}

/* The V-only factors of the next step (vfactors): the exponentials and the
 * divisions spread over the lanes; every lane gets them. */
RYR_DEV RyrVF vfactors_regs(int iso2, int lane, const double v, const double xnai)
{
    RYR_EXACT_FP
    double dv5 = 5, dvk = 8, fv5 = -22.8, fvk = 9.1;
    if (iso2) { dv5 = 0; fv5 = -28; fvk = 8.5; }
    const double q = 0.0337 * (v + 10.5);
    const double za = v * 96.485 / 8.314 / 308;
    double n = 1.0, dd = 1.0;
    switch (lane) {
    case 0: n = -(v - dv5); dd = dvk; break;
    case 1: n = -(v - fv5); dd = fvk; break;
    case 2: n = v - 60; dd = 12.; break;
    case 3: n = xnai; dd = 12.3; break;
    default: break;
    }
    const double q1 = n / dd;
    double arg = 0.0;
    switch (lane) {
    case 0: case 1: case 2: arg = q1; break;
    case 3: arg = -(q * q); break;
    case 4: arg = 2.0 * za; break;
    case 5: arg = 0.35 * za; break;
    case 6: arg = (0.35 - 1.0) * za; break;
    default: break;
    }
    const double ex = ryr_exp(arg);
    const double e0 = lanev(ex, 0), e1 = lanev(ex, 1), e2 = lanev(ex, 2), e3 = lanev(ex, 3);
    const double e4 = lanev(ex, 4), e5 = lanev(ex, 5), e6 = lanev(ex, 6);
    n = 1.0; dd = 1.0;
    switch (lane) {
    case 0: n = 1.0; dd = 1.0 + e0; break;                          /* dinf */
    case 1: n = 1.0 - e0; dd = 0.035 * (v - dv5); break;
    case 2: n = 1.0; dd = 1.0 + e1; break;
    default: break;
    }
    const double q2 = n / dd;
    const double dinf = lanev(q2, 0);
    n = 1.0; dd = 1.0;
    switch (lane) {
    case 0: n = 1.0; dd = lanev(q2, 1) * dinf; break;               /* taudin */
    case 1: n = 0.035 * dvk; dd = dinf; break;                      /* taudin near V = 0 */
    case 2: n = lanev(q2, 2); dd = 1. + e2; break;
    default: break;
    }
    const double q3 = n / dd;
    double taudin = lanev(q3, 0);
    if ((v < 0.0001) && (v > -0.0001)) taudin = lanev(q3, 1);
    const double finf = 1. - lanev(q3, 2);
    const double taufin = (0.02 - 0.007 * e3);
    const double r = lanev(q1, 3);
    RyrVF f;
    f.alphad = dinf * taudin;
    f.betad = (1 - dinf) * taudin;
    f.alphaf = (finf) * taufin;
    f.betaf = (1 - finf) * taufin;
    f.ghk_small = __builtin_fabs(za) < 0.001;
    f.ghk_e2 = e4;
    f.ghk_A = 4.0 * 11.9 * za * 96.485 * 0.341;
    f.ghk_D = f.ghk_e2 - 1.0;
    f.ncx_t1 = 0.00359 * ((140.0) * (140.0) * (140.0)) * (1.0 + r * r * r);
    double p3 = xnai * xnai * xnai;
    f.ncx_t3a = (1.3 + (1 * 1.8)) * p3;
    f.ncx_n1 = e5 * p3 * (1 * 1.8);
    f.ncx_n2 = e6 * ((140.0) * (140.0) * (140.0));
    f.ncx_den2 = 1.0 + 0.27 * e6;
    return f;
}

/* the same, stored into M by lane 0 */
RYR_DEV void vfactors_wave(RyrMem *M, int iso2, int lane, const double v, const double xnai)
{
    const RyrVF f = vfactors_regs(iso2, lane, v, xnai);
    if (lane != 0) return;
    M->alphad = f.alphad; M->betad = f.betad; M->alphaf = f.alphaf; M->betaf = f.betaf;
    M->ghk_small = f.ghk_small; M->ghk_e2 = f.ghk_e2; M->ghk_A = f.ghk_A; M->ghk_D = f.ghk_D;
    M->ncx_t1 = f.ncx_t1; M->ncx_t3a = f.ncx_t3a; M->ncx_n1 = f.ncx_n1; M->ncx_n2 = f.ncx_n2;
    M->ncx_den2 = f.ncx_den2;
}

RYR_DEV double stim_of(const RyrCellInfo &C, long long n, double amp)
{
    if (C.ncl <= 0 || (C.nstim_end >= 0 && n >= C.nstim_end)) return 0.0;
    long long ph = (n - C.noff) % C.ncl;
    if (ph < 0) ph += C.ncl;
    return ph < C.ndur ? amp : 0.0;
}

RYR_DEV int xcd_remap(int b, int nb)
{
    /* consecutive logical blocks on the same XCD (8 XCDs, round-robin dispatch) */
    int per = nb >> 3;
    if (b < (per << 3)) return (b & 7) * per + (b >> 3);
    return b;
}

/* load a neighbour value on the fine grid; `inside` false -> the mirrored
 * ghost, i.e. the centre value */
template <typename T>
RYR_DEV T nb_load(const T *__restrict__ f, long long idx, bool inside, T centre)
{
    return inside ? f[idx] : centre;
}

/* ------------------------------------------------------------ beat tracker */
RYR_DEV void beat_close(RyrMem *M, double *beats, double complete)
{
    if (M->beat_n < RYR_MAXB) {
        double *r = beats + (long long)M->beat_n * RYR_BCOLS;
        r[0] = (double)M->beat_on; r[1] = M->b_apd90; r[2] = M->b_vmax; r[3] = M->b_vrest;
        r[4] = (M->beat_flags & 2) ? 1.0 : 0.0; r[5] = M->b_ead_t; r[6] = M->b_peak_ci;
        r[7] = M->b_jsr; r[8] = M->b_sparks; r[9] = complete;
        r[10] = (M->beat_flags & 8) ? 1.0 : 0.0; r[11] = M->b_eadl_t;
    }
    M->beat_n++;
}

/* V-based features of the beat, once per step after V is final (vo: V at
 * the start of step n, vn: after it). Onset column holds the onset step;
 * the host converts it to ms. */
RYR_DEV void beat_track(RyrMem *M, const RyrCellInfo &C, long long n, double dt, double vo, double vn,
                        double *beats)
{
    RYR_EXACT_FP
    const bool onset = C.ncl > 0 && n >= C.noff && ((n - C.noff) % C.ncl) == 0 &&
                       (C.nstim_end < 0 || n < C.nstim_end);
    if (onset) {
        if (M->beat_on >= 0) beat_close(M, beats, 1.0);
        M->beat_on = n;
        M->b_vrest = vo; M->b_vmax = vo; M->b_vmin = vo;
        M->b_apd90 = -1.0; M->b_ead_t = -1.0; M->beat_flags = 0;
        M->b_peak_ci = 0.0; M->b_jsr = -1.0; M->b_sparks = 0.0; M->b_eadl_t = -1.0;
    }
    if (M->beat_on < 0) return;
    /* the literal flag: after the upstroke V starts to fall (bit2); then any rise ending above -40 mV */
    if (!(M->beat_flags & 4)) {
        if (M->b_vmax > -40.0 && vn < vo) M->beat_flags |= 4;
    } else if (!(M->beat_flags & 8) && vn > vo && vn > -40.0) {
        M->beat_flags |= 8;
        M->b_eadl_t = (double)(n + 1) * dt - (double)M->beat_on * dt;
    }
    if (vn > M->b_vmax) M->b_vmax = vn;
    const double amp = M->b_vmax - M->b_vrest;
    if (!(M->beat_flags & 1)) {
        if (M->b_vmax > -40.0 && vn < M->b_vmax - 0.4 * amp &&
            (double)(n + 1) * dt - (double)M->beat_on * dt >= 50.0) { M->beat_flags |= 1; M->b_vmin = vn; }
    } else {
        if (vn < M->b_vmin) M->b_vmin = vn;
        if (!(M->beat_flags & 2) && vn > -40.0 && vn - M->b_vmin >= 1.0) {
            M->beat_flags |= 2;
            M->b_ead_t = (double)(n + 1) * dt - (double)M->beat_on * dt;
        }
    }
    if (M->b_apd90 < 0.0 && M->b_vmax > -40.0) {
        const double thr = M->b_vmax - 0.9 * amp;
        if (vo >= thr && vn < thr)
            M->b_apd90 = ((double)n + (vo - thr) / (vo - vn)) * dt - (double)M->beat_on * dt;
    }
}

/* ------------------------------------------------------ cytosol/NSR kernel */
/* buffers.c flux with the saturated on-rate 0.95/dt precomputed (sat) */
RYR_DEV float fbuf(float b, float c, float dt, float sat, float kon, float koff, float B)
{
    RYR_EXPR_FP
    const float on = kon * c;
    return (on * dt > 1.0f) ? sat * (B - b) - koff * b : on * (B - b) - koff * b;
}

/* One lane per sub-voxel of the 2x-refined cytosol/NSR grid: a workgroup of
 * WG lanes is one part of one plane ii of one 256-CRU chunk.  SERCA and leak,
 * the 7-point stencils of ci and cnsr, the five cytosolic buffers and the
 * ci/cnsr/buffer updates (cuda_Cell_Detailed.cu:1305-1308, :1320-1393 /
 * :1557-1586), plus the JSR->NSR and sub-membrane/cleft->cytosol exchange
 * terms of sub-voxels 0 and 1 (:1309-1314, :1356-1361).  Under scheme B those
 * two sub-voxels receive their cleft terms during the sub-steps, so their
 * partial rate is left in `sb` for the CRU kernel.  Runs before the CRU kernel
 * of the same step (it reads the step-start cp, cs, cjsr).
 * Grid: nchunks * 8 * (256/WG) workgroups, planes fastest, so the eight planes
 * of a chunk (which read each other as stencil partners) are consecutive
 * logical workgroups on one XCD and meet in its L2.
 *
 * With FUSE the same launch also carries the membrane step of the previous
 * time step: the first `ncells` logical workgroups run membrane_block for
 * step gstep + km.  The two are independent (this kernel reads no membrane
 * state and the membrane reads only the CRU kernel's sums), so the
 * latency-bound ODE hides behind the sub-voxel sweep; the price is the
 * membrane code's registers, hence a separate instance. */
template <typename TF, typename TS, int WG>
RYR_DEV void membrane_block(const RyrKArgs &a, int c, long long nstep, double stim_amp);
RYR_DEV void cru_v_next_get(const RyrKArgs &a, const RyrCellInfo &C, long long nstep, double stim_amp,
                            double *v_out, RyrVF *vf_out);
RYR_DEV void cru_v_next_wave(const RyrKArgs &a, const RyrCellInfo &C, long long nstep, double stim_amp,
                             double *v_out, RyrVF *vf_out);
#define RYR_GATE_LPC 4   /* lanes per CRU in the gating part: each takes two transitions */
template <typename TF, typename TS, int WG>
RYR_DEV void gating_block(const RyrKArgs &a, int L, long long nstep);

/* MODE bit 0: the first ncells logical workgroups run the membrane step
 * gstep + km; bit 1: the next nchunks * 256/WG * RYR_GATE_LPC run the RyR
 * gating of step gstep + kc (gating_block).  The sub-voxel sweep follows. */

/* The update of sub-voxel ii of CRU n (slot g, element e = ii P + g) from
 * its centre values and its six stencil partners (N/S along z, E/W along y,
 * T/B along x; a partner outside the lattice already replaced by the centre
 * value, or under perm by the bath).  Both sweep forms call it, so they
 * compute the same bits. */
template <typename TF, typename TS>
RYR_DEV void cyto_point(const RyrKArgs &a, const RyrCellInfo &C, long long nstep, bool diag, int ii, int n,
                        long long g, long long e, TF cc, TS cn, TF inN, TF inS, TF inE, TF inW, TF inT, TF inB,
                        TS nnN, TS nnS, TS nnE, TS nnW, TS nnT, TS nnB)
{
    const int par = (int)(nstep & 1);
    TF *__restrict__ ci1 = (TF *)a.ci[par ^ 1];
    TS *__restrict__ cn1 = (TS *)a.cnsr[par ^ 1];
    TF *catf = (TF *)a.catf, *cacal = (TF *)a.cacal, *casr = (TF *)a.casr;
    TF *camyo = (TF *)a.camyo, *mgmyo = (TF *)a.mgmyo;
    TS *cats = (TS *)a.cats;
    const bool schemeA = a.scheme == 0;
    if constexpr (sizeof(TF) == 8) {
        /* exact path: the reference's operand order throughout */
        const double dt = a.dt;
        const double sviup = ((C.variants & RYR_VAR_ISO) && nstep >= C.iso_n) ? C.sviup_iso : C.sviup_tot;
        const double xiup = sviup * uptake_d(cc, cn);
        const double xileak = 1.0 * leak_t<double>(cn, cc);
        const double xicoupn = ((nnN - cn) / (1.8 * 0.7) + (nnS - cn) / (1.8 * 0.7) +
                                (nnE - cn) / (1.8 * 0.7) + (nnW - cn) / (1.8 * 0.7) +
                                (nnT - cn) / (6. * 0.7) + (nnB - cn) / (6. * 0.7));
        const double xicoupi = ((inN - cc) / ((0.33 * 2.) * 0.7) + (inS - cc) / ((0.33 * 2.) * 0.7) +
                                (inE - cc) / ((0.33 * 2.) * 0.7) + (inW - cc) / ((0.33 * 2.) * 0.7) +
                                (inT - cc) / ((0.7 * 2.) * 0.7) + (inB - cc) / ((0.7 * 2.) * 0.7));
        double dotcnsr = (xiup - xileak) * (0.5 / 8.) / (0.025 / 8.) + xicoupn;
        const double v_catf = ryr_ld_nt(catf + e), v_cats = ryr_ld_nt(cats + e), v_cacal = ryr_ld_nt(cacal + e);
        const double v_casr = ryr_ld_nt(casr + e), v_camyo = ryr_ld_nt(camyo + e), v_mgmyo = ryr_ld_nt(mgmyo + e);
        const double f_tf = b_tropf<double>(v_catf, cc, dt);
        const double f_ts = b_trops<double>(v_cats, cc, dt);
        const double f_cal = b_bucal<double>(v_cacal, cc, dt);
        const double f_sr = b_busr<double>(v_casr, cc, dt);
        const double f_myo = b_myoca<double>(v_camyo, v_mgmyo, cc, dt);
        const double buffers = (f_tf + f_ts + f_cal + f_sr + f_myo);
        double dotci = (-xiup + xileak) - buffers + xicoupi;
        if (ii < 2) {
            const double cj = ((const TS *)a.cjsr[par])[g];
            const double diffjn = (cj - cn) / (25.0 * 1.0 / 4.) / 2.;
            dotcnsr += diffjn * (0.02 * 1) / (0.025 / 8.);
            if (schemeA) {
                const double csv = ((const TF *)a.cs[par])[g], cpv = ((const TF *)a.cp[par])[g];
                dotci += (csv - cc) / (0.04) / 2. * ((0.025) / (0.5 / 8.));
                dotci += (cpv - cc) / (0.1) / 2. * ((0.00126) / (0.5 / 8.));
            } else {
                ((TF *)a.sb)[e] = dotci;
            }
        }
        if (ii >= 2 || schemeA) {
            double cin = cc + dotci * dt;
            if (cin < 0) cin = 0;
            ryr_st_nt(ci1 + e, (TF)cin);
        }
        ryr_st_nt(cn1 + e, (TS)(cn + dotcnsr * dt));
        double n_catf = v_catf + f_tf * dt;
        if (n_catf < 0) n_catf = 0;
        double n_cats = v_cats + f_ts * dt;
        if (n_cats < 0) n_cats = 0;
        double n_cacal = v_cacal + f_cal * dt;
        if (n_cacal < 0) n_cacal = 0;
        double n_casr = v_casr + f_sr * dt;
        if (n_casr < 0) n_casr = 0;
        double n_camyo = v_camyo + f_myo * dt;
        if (n_camyo < 0) n_camyo = 0;
        double n_mgmyo = v_mgmyo + b_myomg<double>(n_camyo, v_mgmyo, dt) * dt;
        if (n_mgmyo < 0) n_mgmyo = 0;
        ryr_st_nt(catf + e, (TF)n_catf); ryr_st_nt(cats + e, (TS)n_cats); ryr_st_nt(cacal + e, (TF)n_cacal);
        ryr_st_nt(casr + e, (TF)n_casr); ryr_st_nt(camyo + e, (TF)n_camyo); ryr_st_nt(mgmyo + e, (TF)n_mgmyo);
        if (diag) { a.d_up[e] = xiup; a.d_leak[e] = xileak; }
    } else {
        /* FP32 production path: the same equations with divisions by
         * constants replaced by reciprocal products.  cnsr and slow troponin
         * change by less than a float ulp per step at rest, so their updates
         * are formed in double and rounded stochastically to float.  Fusion
         * only inside an expression: both sweep forms (and the one-launch
         * kernel) inline this, and results must not depend on which one ran */
        RYR_EXPR_FP
        const float dt = (float)a.dt;
        const float sat = 0.95f * __builtin_amdgcn_rcpf(dt);
        const float rIT = (float)(1.0 / ((0.33 * 2.) * 0.7)), rIL = (float)(1.0 / ((0.7 * 2.) * 0.7));
        const float rNT = (float)(1.0 / (1.8 * 0.7)), rNL = (float)(1.0 / (6. * 0.7));
        const float cf = (float)cc, cnf = (float)cn;
        const double sviup = ((C.variants & RYR_VAR_ISO) && nstep >= C.iso_n) ? C.sviup_iso : C.sviup_tot;
        const float xiup = (float)sviup * uptake_f(cf, cnf);
        const float cn2 = cnf * cnf;
        const float xileak = 1.035e-5f * (cnf - cf) * cn2 * __builtin_amdgcn_rcpf(cn2 + 250000.0f);
        const float xicoupi = (((float)inN - cf) + ((float)inS - cf) + ((float)inE - cf) + ((float)inW - cf)) * rIT +
                              (((float)inT - cf) + ((float)inB - cf)) * rIL;
        const float xicoupn = ((float)(nnN - cn) + (float)(nnS - cn) + (float)(nnE - cn) + (float)(nnW - cn)) * rNT +
                              ((float)(nnT - cn) + (float)(nnB - cn)) * rNL;
        float dotcnsr = (xiup - xileak) * (float)((0.5 / 8.) / (0.025 / 8.)) + xicoupn;
        const float v_catf = (float)ryr_ld_nt(catf + e), v_cacal = (float)ryr_ld_nt(cacal + e);
        const float v_casr = (float)ryr_ld_nt(casr + e);
        const float v_camyo = (float)ryr_ld_nt(camyo + e), v_mgmyo = (float)ryr_ld_nt(mgmyo + e);
        const TS v_cats = ryr_ld_nt(cats + e);
        const float f_tf = fbuf(v_catf, cf, dt, sat, 0.0327f, 0.0196f, 70.0f);
        const float f_ts = fbuf((float)v_cats, cf, dt, sat, 0.00254f, 0.000033f, 134.0f);
        const float f_cal = fbuf(v_cacal, cf, dt, sat, 0.0543f, 0.238f, 24.0f);
        const float f_sr = fbuf(v_casr, cf, dt, sat, 0.1f, 0.06f, 19.0f);
        const float on_my = 13.8f * cf;
        const float f_myo = (on_my * dt > 1.0f ? sat : on_my) * (140.0f - v_camyo - v_mgmyo) - 0.46f * v_camyo;
        float dotci = (xileak - xiup) - (f_tf + f_ts + f_cal + f_sr + f_myo) + xicoupi;
        if (ii < 2) {
            const TS cj = ((const TS *)a.cjsr[par])[g];
            dotcnsr += (float)((cj - cn) * TS(1.0 / (25.0 * 1.0 / 4.) / 2.)) * (float)((0.02 * 1) / (0.025 / 8.));
            if (schemeA) {
                const float csv = (float)((const TF *)a.cs[par])[g], cpv = (float)((const TF *)a.cp[par])[g];
                dotci += (csv - cf) * (float)(1.0 / 0.04 / 2. * ((0.025) / (0.5 / 8.))) +
                         (cpv - cf) * (float)(1.0 / 0.1 / 2. * ((0.00126) / (0.5 / 8.)));
            } else {
                ((TF *)a.sb)[e] = (TF)dotci;
            }
        }
        if (ii >= 2 || schemeA) ryr_st_nt(ci1 + e, (TF)fmaxf(cf + dotci * dt, 0.0f));
        const double cn_new = (double)cn + (double)dotcnsr * a.dt;
        if constexpr (sizeof(TS) == 4)
            ryr_st_nt(cn1 + e, ryr_sr_f32(cn_new, ryr_sr_bits(ryr_sr_key(C.seed, (uint64_t)nstep, RYR_SR_CNSR), (uint32_t)n,
                                                              (uint32_t)ii)));
        else
            ryr_st_nt(cn1 + e, (TS)cn_new);
        const float n_camyo = fmaxf(v_camyo + f_myo * dt, 0.0f);
        const float on_mg = 0.0157f * 500.0f;
        const float f_mg = (on_mg * dt > 1.0f ? sat : on_mg) * (140.0f - n_camyo - v_mgmyo) - 0.057f * v_mgmyo;
        ryr_st_nt(catf + e, (TF)fmaxf(v_catf + f_tf * dt, 0.0f));
        double n_cats = (double)v_cats + (double)f_ts * a.dt;
        if (n_cats < 0.0) n_cats = 0.0;
        if constexpr (sizeof(TS) == 4)
            ryr_st_nt(cats + e, ryr_sr_f32(n_cats, ryr_sr_bits(ryr_sr_key(C.seed, (uint64_t)nstep, RYR_SR_CATS), (uint32_t)n,
                                                                (uint32_t)ii)));
        else
            ryr_st_nt(cats + e, (TS)n_cats);
        ryr_st_nt(cacal + e, (TF)fmaxf(v_cacal + f_cal * dt, 0.0f));
        ryr_st_nt(casr + e, (TF)fmaxf(v_casr + f_sr * dt, 0.0f));
        ryr_st_nt(camyo + e, (TF)n_camyo);
        ryr_st_nt(mgmyo + e, (TF)fmaxf(v_mgmyo + f_mg * dt, 0.0f));
        if (diag) { a.d_up[e] = (double)xiup; a.d_leak[e] = (double)xileak; }
    }
}

/* The sub-voxel sweep for logical workgroup L of a launch: one lane per
 * sub-voxel of plane ii of one part of one 256-CRU chunk. */
template <typename TF, typename TS, int WG>
RYR_DEV void cyto_block(const RyrKArgs &a, int L, const RyrCellInfo *recs, const long long *gbase, int kc)
{
    constexpr int SUB = RYR_BLOCK / WG;
    const int b = L / (8 * SUB);
    const int r = L - b * (8 * SUB);
    const int ii = r & 7;
    const int part = r >> 3;
    /* the chunk record (recs: a.chunks) and the step base (gbase: a.gstep),
     * passed by the caller, which may hold them in registers; both loads
     * are issued before either is used */
    const RyrCellInfo C = recs[b];
    const long long nstep = *gbase + kc;
    const int n = (b - C.first_block) * RYR_BLOCK + part * WG + (int)threadIdx.x;
    if (n >= C.N) return;
    const int par = (int)(nstep & 1);
    const bool diag = (nstep % a.spo) == 0;
    const long long P = a.P, g = C.off + n, e = ii * P + g;
    const int X = C.X, Y = C.Y, Z = C.Z;
    /* position in the lattice: i = n mod X, rxy = n mod XY (j = rxy / X) */
    const int sy = X, sz = X * Y;
    const int i = n - X * (int)ryr_fastdiv((uint32_t)n, C.dX_m, C.dX_s1, C.dX_s2);
    const int rxy = n - sz * (int)ryr_fastdiv((uint32_t)n, C.dXY_m, C.dXY_s1, C.dXY_s2);
    const TF *__restrict__ ci0 = (const TF *)a.ci[par];
    const TS *__restrict__ cn0 = (const TS *)a.cnsr[par];

    /* neighbours on the fine grid (bit0 z, bit1 y, bit2 x); outside the
     * interior the mirrored ghost equals the centre value */
    const TF cc = ci0[e];
    const TS cn = cn0[e];
    /* outside the lattice: the centre value, or under perm the bath ci0
     * (the reference never refreshes those ghosts, :1662-1689) */
    const TF cb = (C.variants & RYR_VAR_PERM) ? (TF)C.ci_bath : cc;
    TF inN, inS, inE, inW, inT, inB;
    TS nnN, nnS, nnE, nnW, nnT, nnB;
    if (ii & 1) {
        const bool in = n + sz < C.N;      /* k + 1 < Z */
        inN = in ? ci0[(ii - 1) * P + g + sz] : cb; nnN = in ? cn0[(ii - 1) * P + g + sz] : cn;
        inS = ci0[(ii - 1) * P + g]; nnS = cn0[(ii - 1) * P + g];
    } else {
        const bool in = n >= sz;           /* k > 0 */
        inN = ci0[(ii + 1) * P + g]; nnN = cn0[(ii + 1) * P + g];
        inS = in ? ci0[(ii + 1) * P + g - sz] : cb; nnS = in ? cn0[(ii + 1) * P + g - sz] : cn;
    }
    if (ii & 2) {
        const bool in = rxy + sy < sz;     /* j + 1 < Y */
        inE = in ? ci0[(ii - 2) * P + g + sy] : cb; nnE = in ? cn0[(ii - 2) * P + g + sy] : cn;
        inW = ci0[(ii - 2) * P + g]; nnW = cn0[(ii - 2) * P + g];
    } else {
        const bool in = rxy >= sy;         /* j > 0 */
        inE = ci0[(ii + 2) * P + g]; nnE = cn0[(ii + 2) * P + g];
        inW = in ? ci0[(ii + 2) * P + g - sy] : cb; nnW = in ? cn0[(ii + 2) * P + g - sy] : cn;
    }
    if (ii & 4) {
        const bool in = i + 1 < X;
        inT = in ? ci0[(ii - 4) * P + g + 1] : cb; nnT = in ? cn0[(ii - 4) * P + g + 1] : cn;
        inB = ci0[(ii - 4) * P + g]; nnB = cn0[(ii - 4) * P + g];
    } else {
        const bool in = i > 0;
        inT = ci0[(ii + 4) * P + g]; nnT = cn0[(ii + 4) * P + g];
        inB = in ? ci0[(ii + 4) * P + g - 1] : cb; nnB = in ? cn0[(ii + 4) * P + g - 1] : cn;
    }

    cyto_point<TF, TS>(a, C, nstep, diag, ii, n, g, e, cc, cn, inN, inS, inE, inW, inT, inB,
                       nnN, nnS, nnE, nnW, nnT, nnB);
}

/* The sub-voxel sweep with one lane per CRU, for grids that fill the GPU:
 * the eight sub-voxels of CRU n are held in registers, so the partners inside
 * the CRU come from registers, the x partners (CRUs n - 1 and n + 1) from the
 * neighbouring lanes of the wave (from memory at the wave's two ends), and
 * only the y and z partners from memory: 4 partner loads per sub-voxel and
 * field fewer than the per-sub-voxel form's 6.  Logical workgroup b is
 * chunk b.  Each sub-voxel's update is cyto_point, as in cyto_block. */
template <typename TF, typename TS>
RYR_DEV void cyto_cru_block(const RyrKArgs &a, int b, long long nstep)
{
    const RyrCellInfo C = a.chunks[b];
    const int n = (b - C.first_block) * RYR_BLOCK + (int)threadIdx.x;
    /* a lane past the cell's last CRU holds its padding slot (inside the
     * chunk) until the lane exchange, then leaves; it loads no partner */
    const bool active = n < C.N;
    const int par = (int)(nstep & 1);
    const bool diag = (nstep % a.spo) == 0;
    const long long P = a.P, g = C.off + n;
    const int X = C.X, sy = C.X, sz = C.X * C.Y;
    const int i = n - X * (int)ryr_fastdiv((uint32_t)n, C.dX_m, C.dX_s1, C.dX_s2);
    const int rxy = n - sz * (int)ryr_fastdiv((uint32_t)n, C.dXY_m, C.dXY_s1, C.dXY_s2);
    const TF *__restrict__ ci0 = (const TF *)a.ci[par];
    const TS *__restrict__ cn0 = (const TS *)a.cnsr[par];
    TF cc[8];
    TS cn[8];
#pragma unroll
    for (int ii = 0; ii < 8; ++ii) {
        cc[ii] = ci0[ii * P + g];
        cn[ii] = cn0[ii * P + g];
    }
    /* x partners: plane q of CRU n + 1 (partner of sub-voxel q + 4) and plane
     * q + 4 of CRU n - 1 (partner of sub-voxel q), q < 4 */
    const int lane = (int)threadIdx.x & 63;
    const bool inxp = active && i + 1 < X, inxm = active && i > 0;
    TF xpc[4], xmc[4];
    TS xpn[4], xmn[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        xpc[q] = ryr_lane_from(cc[q], lane + 1);
        xpn[q] = ryr_lane_from(cn[q], lane + 1);
        xmc[q] = ryr_lane_from(cc[q + 4], lane - 1);
        xmn[q] = ryr_lane_from(cn[q + 4], lane - 1);
    }
    if (!active) return;
    if (lane == 63 && inxp) {
#pragma unroll
        for (int q = 0; q < 4; ++q) { xpc[q] = ci0[q * P + g + 1]; xpn[q] = cn0[q * P + g + 1]; }
    }
    if (lane == 0 && inxm) {
#pragma unroll
        for (int q = 0; q < 4; ++q) { xmc[q] = ci0[(q + 4) * P + g - 1]; xmn[q] = cn0[(q + 4) * P + g - 1]; }
    }
    const bool inzp = n + sz < C.N, inzm = n >= sz;     /* k + 1 < Z, k > 0 */
    const bool inyp = rxy + sy < sz, inym = rxy >= sy;  /* j + 1 < Y, j > 0 */
    const bool perm = (C.variants & RYR_VAR_PERM) != 0;
#pragma unroll
    for (int ii = 0; ii < 8; ++ii) {
        const long long e = ii * P + g;
        /* outside the lattice: the centre value, or under perm the bath ci0 */
        const TF cb = perm ? (TF)C.ci_bath : cc[ii];
        TF inN, inS, inE, inW, inT, inB;
        TS nnN, nnS, nnE, nnW, nnT, nnB;
        if (ii & 1) {
            inN = inzp ? ci0[(ii - 1) * P + g + sz] : cb; nnN = inzp ? cn0[(ii - 1) * P + g + sz] : cn[ii];
            inS = cc[ii - 1]; nnS = cn[ii - 1];
        } else {
            inN = cc[ii + 1]; nnN = cn[ii + 1];
            inS = inzm ? ci0[(ii + 1) * P + g - sz] : cb; nnS = inzm ? cn0[(ii + 1) * P + g - sz] : cn[ii];
        }
        if (ii & 2) {
            inE = inyp ? ci0[(ii - 2) * P + g + sy] : cb; nnE = inyp ? cn0[(ii - 2) * P + g + sy] : cn[ii];
            inW = cc[ii - 2]; nnW = cn[ii - 2];
        } else {
            inE = cc[ii + 2]; nnE = cn[ii + 2];
            inW = inym ? ci0[(ii + 2) * P + g - sy] : cb; nnW = inym ? cn0[(ii + 2) * P + g - sy] : cn[ii];
        }
        if (ii & 4) {
            inT = inxp ? xpc[ii - 4] : cb; nnT = inxp ? xpn[ii - 4] : cn[ii];
            inB = cc[ii - 4]; nnB = cn[ii - 4];
        } else {
            inT = cc[ii + 4]; nnT = cn[ii + 4];
            inB = inxm ? xmc[ii] : cb; nnB = inxm ? xmn[ii] : cn[ii];
        }
        cyto_point<TF, TS>(a, C, nstep, diag, ii, n, g, e, cc[ii], cn[ii], inN, inS, inE, inW, inT, inB,
                           nnN, nnS, nnE, nnW, nnT, nnB);
    }
}

/* the lane-per-CRU sweep on its own: one 256-lane workgroup per chunk */
template <typename TF, typename TS>
__global__ __launch_bounds__(RYR_BLOCK) void k_cyto_cru(RyrKArgs a, int kc)
{
    cyto_cru_block<TF, TS>(a, xcd_remap((int)blockIdx.x, (int)gridDim.x), *a.gstep + kc);
}

template <typename TF, typename TS, int WG, int MODE>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu((MODE & 2) ? 4 : 1)))
void k_cyto(RyrKArgs a, int kc, int km, double stim_amp)
{
    constexpr int SUB = RYR_BLOCK / WG;
    const long long base = *a.gstep;
    const int nmem = (MODE & 1) ? a.ncells : 0;
    const int Lg = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    if constexpr ((MODE & 1) != 0) {
        if (Lg < nmem) {
            membrane_block<TF, TS, WG>(a, Lg, base + km, stim_amp);
            return;
        }
    }
    int L = Lg - nmem;
    if constexpr ((MODE & 2) != 0) {
        /* RYR_GATE_LPC gating workgroups per 8 sub-voxel ones */
        const int ngate = ((int)gridDim.x - nmem) / (8 + RYR_GATE_LPC) * RYR_GATE_LPC;
        if (L < ngate) {
            gating_block<TF, TS, WG>(a, L, base + kc);
            return;
        }
        L -= ngate;
    }
    cyto_block<TF, TS, WG>(a, L, a.chunks, a.gstep, kc);
}

/* Cleft/sub-membrane sub-steps of the FP32 production path (scheme A
 * cuda_Cell_Detailed.cu:1396-1464, scheme B :1479-1553), reciprocal
 * constants and float arithmetic.  Under scheme B dotci01 enters with the
 * cytosol kernel's partial rates of sub-voxels 0 and 1. */
template <bool schemeA>
RYR_DEV void fast_cleft(float dt, float cif0, float cif1, float csf0, float csu, float csd,
                        float ica, float xna, float xre, float bcg, bool perm, float bath, float &cp, float &cs,
                        float &sar, float &sarh, float &sarj, float &sarhj, float &dci0, float &dci1)
{
    const float Vs_i = (float)((0.025) / (0.5 / 8.)), Vp_i = (float)((0.00126) / (0.5 / 8.));
    const float Vp_s = (float)((0.00126) / (0.025));
    const float r_tausi2 = (float)(1.0 / 0.04 / 2.), r_taupi2 = (float)(1.0 / 0.1 / 2.);
    const float r_taups = (float)(1.0 / 0.0283), r_taust = (float)(1.0 / 1.42);
    const float dth = dt * 0.2f;
    const float sath = 0.95f / dth;
    const float cpf0 = cp;
    if constexpr (schemeA) {
        const float csdiff = (csu + csd - 2.0f * csf0) * r_taust;
        const float to_ci = ((csf0 - cif0) + (csf0 - cif1)) * r_tausi2;
        const float cp_ci = ((cif0 - cpf0) + (cif1 - cpf0)) * r_taupi2;
        const float ica_s = 0.5f * ica * Vp_s;
#pragma unroll
        for (int it = 0; it < 5; ++it) {
            const float fsar = fbuf(sar, cs, dth, sath, 0.1f, 1.3f, (float)RYR_BSAR);
            const float fsarh = fbuf(sarh, cs, dth, sath, 0.1f, 0.03f, (float)RYR_BSARH);
            const float dotcs = Vp_s * (cp - cs) * r_taups - ica_s + xna - to_ci + csdiff - (fsar + fsarh) - bcg +
                                (perm ? (bath - cs) * 0.2f : 0.0f);
            const float cpbuff = fbuf(sarj, cp, dth, sath, 0.1f, 1.3f, (float)RYR_BSAR);
            const float cpbuffh = fbuf(sarhj, cp, dth, sath, 0.1f, 0.03f, (float)RYR_BSARH);
            const float cpdot = (cs - cp) * r_taups + cp_ci + xre - ica - cpbuff - cpbuffh;
            sarj = fmaxf(sarj + cpbuff * dth, 0.0f);
            sarhj = fmaxf(sarhj + cpbuffh * dth, 0.0f);
            cp += cpdot * dth;
            sar = fmaxf(sar + fsar * dth, 0.0f);
            sarh = fmaxf(sarh + fsarh * dth, 0.0f);
            cs += dotcs * dth;
        }
    } else {
        const float csum = csu + csd;
#pragma unroll
        for (int it = 0; it < 5; ++it) {
            dci0 += ((cs - cif0) * r_tausi2 * Vs_i + (cp - cif0) * r_taupi2 * Vp_i) * 0.2f;
            dci1 += ((cs - cif1) * r_tausi2 * Vs_i + (cp - cif1) * r_taupi2 * Vp_i) * 0.2f;
            const float fsar = fbuf(sar, cs, dth, sath, 0.1f, 1.3f, (float)RYR_BSAR);
            const float fsarh = fbuf(sarh, cs, dth, sath, 0.1f, 0.03f, (float)RYR_BSARH);
            const float dotcs = Vp_s * (cp - 0.5f * 0.0283f * ica - cs) * r_taups + xna -
                                ((cs - cif0) + (cs - cif1)) * r_tausi2 + (csum - 2.0f * cs) * r_taust - (fsar + fsarh) -
                                bcg + (perm ? (bath - cs) * 0.2f : 0.0f);
            const float cpbuff = fbuf(sarj, cp, dth, sath, 0.1f, 1.3f, (float)RYR_BSAR);
            const float cpbuffh = fbuf(sarhj, cp, dth, sath, 0.1f, 0.03f, (float)RYR_BSARH);
            const float cpdot = (csf0 - cp) * r_taups + ((cif0 - cp) + (cif1 - cp)) * r_taupi2 + xre - ica -
                                cpbuff - cpbuffh;
            sar = fmaxf(sar + fsar * dth, 0.0f);
            sarh = fmaxf(sarh + fsarh * dth, 0.0f);
            cs += dotcs * dth;
            sarj = fmaxf(sarj + cpbuff * dth, 0.0f);
            sarhj = fmaxf(sarhj + cpbuffh * dth, 0.0f);
            cp += cpdot * dth;
        }
    }
}

/* ------------------------------------------------------------ RyR gating */
/* The RyR gating of step nstep for WG CRUs (cuda_Cell_Detailed.cu:1274-1290,
 * subroutine_cuda.c:251-669), as a part of the sub-voxel launch of the same
 * step (schedule RYR_SCHED_FUSED_GATING).  It reads only step-start values
 * that the CRU kernel of the previous step wrote last (cp, cjsr, the RyR
 * counts), so it runs beside the sub-voxel sweep and the previous membrane
 * step, and the CRU kernel of the step (GSPLIT) takes the new counts.  It
 * also clears the output interval's LCC-open count and spark flag at the
 * interval's first step, as the CRU kernel does otherwise, then records a
 * spark (open channels rising through 5). */
template <typename TF, typename TS, int WG>
RYR_DEV void gating_block(const RyrKArgs &a, int L, long long nstep)
{
    constexpr int LPC = RYR_GATE_LPC;
    constexpr int CPW = WG / LPC;
    constexpr int SUBG = RYR_BLOCK / CPW;
    const int bid = L / SUBG;
    const int part = L - bid * SUBG;
    const RyrCellInfo C = a.chunks[bid];
    const int n = (bid - C.first_block) * RYR_BLOCK + part * CPW + (int)threadIdx.x / LPC;
    const int h = (int)threadIdx.x & (LPC - 1);
    if (n >= C.N) return;                  /* both lanes of a CRU together */
    const long long g = C.off + n;
    const int par = (int)(nstep & 1);
    const double cp_d = (double)((const TF *)a.cp[par])[g];
    const double cj = (double)((const TS *)a.cjsr[par])[g];
    const uint32_t rw = a.ryrw[g];
    uint32_t ax = a.aux[g];
    if ((nstep == 0) || ((nstep - 1) % a.spo == 0)) ax &= ~0x1ffffu;
    int nou = (int)(rw & 255u), ncu = (int)((rw >> 8) & 255u), nob = (int)((rw >> 16) & 255u), ncb = (int)(rw >> 24);
    const int prevsp = nou + nob;
    double roo2;
    const double mono = csqn_mono(cj, &roo2);
    int nfail = 0, ndraw = 0;
    ryr_gating<LPC>(C.seed, (uint32_t)n, (uint64_t)nstep, C.Ku, C.Ku2, C.Kb, cp_d, cj, mono, a.dt, nou, ncu, nob,
                    ncb, &nfail, &ndraw, h);
    if constexpr (LPC == 2) nfail += (int)lane_pair_u32((uint32_t)nfail);
    if constexpr (LPC == 4)
        nfail = (int)(quad_lane_u32<0>((uint32_t)nfail) + quad_lane_u32<1>((uint32_t)nfail) +
                      quad_lane_u32<2>((uint32_t)nfail) + quad_lane_u32<3>((uint32_t)nfail));
    if (h != 0) return;
    if (prevsp < 5 && (nou + nob) >= 5) ax |= 1u << 16;
    a.ryrw[g] = (uint32_t)nou | ((uint32_t)ncu << 8) | ((uint32_t)nob << 16) | ((uint32_t)ncb << 24);
    a.aux[g] = ax;
    if (nfail) atomicAdd(a.counters + (long long)C.cell * RYR_NCNT + 0, (unsigned long long)nfail);
}

/* ------------------------------------------------------------ step kernel */
/* The CRU update for logical workgroup L of a launch (step *gbase + kc;
 * recs and gbase are a.chunks and a.gstep, passed by the caller, which may
 * hold them in registers). */
template <typename TF, typename TS, int WG, bool SCHB, int LPC, bool GSPLIT, bool VNEXT = false>
RYR_DEV void cru_block(const RyrKArgs &a, int kc, int L, double stim_amp, const RyrCellInfo *recs,
                       const long long *gbase)
{
    static_assert(LPC == 1 || (LPC == 2 && WG >= 128), "paired lanes need two waves per workgroup");

    constexpr int CPW = WG / LPC;          /* CRUs per workgroup */
    constexpr bool NTC = WG < 256;         /* nontemporal per-CRU state (batch instances) */
    constexpr int SUB = RYR_BLOCK / CPW;
    const int bid = L / SUB;               /* 256-CRU chunk */
    const int part = L - bid * SUB;
    const RyrCellInfo C = recs[bid];
    const int c = C.cell;
    const RyrMem *M = a.mem + c;
    const long long nstep = *gbase + kc;
    const int par = (int)(nstep & 1);
    const int spo = a.spo;
    const bool reset = (nstep == 0) || ((nstep - 1) % spo == 0);
    const bool diag = (nstep % spo) == 0;
    const int n = (bid - C.first_block) * RYR_BLOCK + part * CPW + (int)threadIdx.x / LPC;
    const int hl = LPC == 1 ? 0 : (int)threadIdx.x & (LPC - 1);   /* lane half of a shared CRU */
    const bool active = n < C.N;
    /* the lane that stores: under VNEXT every lane runs the update (see
     * below), and only the leads of real CRUs store */
    const bool lead = hl == 0 && active;
    const long long P = a.P;
    const double dt = a.dt;

    double r_icat = 0.0, r_inaca = 0.0, r_cs = 0.0;

    /* The CRU's slot.  A lane past the cell's last CRU (active false) points
     * at a padding slot of the same 256-slot chunk, which exists, so every
     * lane issues the state loads; under VNEXT they are then in flight while
     * the wave forms V, which needs every lane. */
    const long long g = C.off + n;
    const int X = C.X, Y = C.Y, Z = C.Z;
    const int sz = X * Y;
    const TF *__restrict__ ci0 = (const TF *)a.ci[par];
    TF *__restrict__ ci1 = (TF *)a.ci[par ^ 1];
    const TS *__restrict__ cn0 = (const TS *)a.cnsr[par];
    TS *__restrict__ cn1 = (TS *)a.cnsr[par ^ 1];
    const TF *__restrict__ cs0 = (const TF *)a.cs[par];
    TF *__restrict__ cs1 = (TF *)a.cs[par ^ 1];
    TF *catf = (TF *)a.catf, *cacal = (TF *)a.cacal, *casr = (TF *)a.casr;
    TF *camyo = (TF *)a.camyo, *mgmyo = (TF *)a.mgmyo;
    TS *cats = (TS *)a.cats;
    const TF *cp0 = (const TF *)a.cp[par];
    TF *cp1 = (TF *)a.cp[par ^ 1], *ancx = (TF *)a.Ancx;
    TF *casar = (TF *)a.casar, *casarh = (TF *)a.casarh, *casarj = (TF *)a.casarj, *casarhj = (TF *)a.casarhj;
    const TS *cjsr0 = (const TS *)a.cjsr[par];
    TS *cjsr1 = (TS *)a.cjsr[par ^ 1], *cjsrbp = (TS *)a.cjsrb;
    /* The V-only factors and the NSR values of sub-voxels 0-1 are loaded
     * with the rest of the state where registers allow (the instances that
     * run one wave per SIMD anyway), and at their use sites in the
     * instances held to 168 VGPRs, where keeping them live spills */
    constexpr bool vf_early = WG >= 256 || SCHB || VNEXT;
    TF cil[2];
    TS cnl[2] = {TS(0), TS(0)};
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
        cil[ii] = ci0[ii * P + g];
        if constexpr (vf_early) cnl[ii] = cn0[ii * P + g];
    }
    const TF csv = cs0[g];
    TF cpv = cp0[g];
    const TS cjsrv = cjsr0[g];
    uint32_t lw = ryr_ld_if<NTC>(a.lccw + g), rw = ryr_ld_if<NTC>(a.ryrw + g), ax = ryr_ld_if<NTC>(a.aux + g);

    /* VNEXT (one-launch schedule, after the run's first step): V of this
     * step and its V-only factors come from this launch's membrane workgroup
     * or, failing that, are formed by each wave (cru_v_next_get), which needs
     * every lane; so every lane runs the update (a lane past the cell's last
     * CRU on its padding slot, stores masked by `lead`), and the RyR gating,
     * which needs no V, runs first, while V is being formed */
    if (active || VNEXT) {
        double v = 0.0;
        RyrVF vf0{};
        if constexpr (!VNEXT) {
            v = M->m[0];
            if constexpr (vf_early) vf0 = load_vf(M);
        }
        /* the draw blocks of LCC channels 0-3 need only the key, so in the
         * instances that have the registers they are formed while the state
         * loads are in flight */
        ryr_u32x4 lpre[2] = {};
        if constexpr (LPC == 1 && vf_early) {
            lpre[0] = ryr_block(C.seed, (uint32_t)n, (uint64_t)nstep, 0u);
            lpre[1] = ryr_block(C.seed, (uint32_t)n, (uint64_t)nstep, 1u);
        }
        /* the gating-free instances, which have the registers, and the
         * one-launch ones, whose update after V is the step's critical path,
         * also issue the cleft buffer, JSR buffer, NCX, scheme-B and
         * vertical-neighbour loads here rather than at their use */
        constexpr bool EARLY = GSPLIT || VNEXT;
        const bool has_up = n + sz < C.N, has_dn = n >= sz;
        TF e_sar{}, e_sarh{}, e_sarj{}, e_sarhj{}, e_up{}, e_dn{}, e_sb0{}, e_sb1{}, e_ancx{};
        TS e_cjb{};
        if constexpr (EARLY) {
            e_sar = ryr_ld_if<NTC>(casar + g); e_sarh = ryr_ld_if<NTC>(casarh + g);
            e_sarj = ryr_ld_if<NTC>(casarj + g); e_sarhj = ryr_ld_if<NTC>(casarhj + g);
            e_cjb = ryr_ld_if<NTC>(cjsrbp + g);
            e_ancx = ryr_ld_if<NTC>(ancx + g);
            if (has_up) e_up = cs0[g + sz];
            if (has_dn) e_dn = cs0[g - sz];
            if (SCHB) { e_sb0 = ((const TF *)a.sb)[g]; e_sb1 = ((const TF *)a.sb)[P + g]; }
        }
        if (!GSPLIT && reset) ax &= ~0x1ffffu;      /* nl and spark flag; the LCC slots stay (GSPLIT: done by the gating) */
        int nl = (int)(ax & 0xffffu), nspark = (int)((ax >> 16) & 1u);
        uint32_t lmask = (ax >> RYR_AUX_MASK_SHIFT) & 0xffu;
        int nfail = 0, nclampcp = 0, nclampcs = 0;
        const bool perm = (C.variants & RYR_VAR_PERM) != 0;
        if ((C.variants & RYR_VAR_ISO) && iso_active(nstep, C.ncl)) {
            /* iso (:1214-1219): absent slots 4..7 become channels in state 3 */
            for (int ll = 4; ll < 8; ++ll)
                if (!((lmask >> ll) & 1u)) {
                    lmask |= 1u << ll;
                    lw = (lw & ~(15u << (4 * ll))) | (3u << (4 * ll));
                }
        }

        const double cp_d = (double)cpv, cs_d = (double)csv;
        /* RyR gating (:1274-1284; subroutine_cuda.c:251-669): reads only
         * step-start values, so where it runs does not change a result */
        int nou = (int)(rw & 255u), ncu = (int)((rw >> 8) & 255u), nob = (int)((rw >> 16) & 255u),
            ncb = (int)(rw >> 24);
        const int prevsp = nou + nob;
        double roo2;
        const double mono = csqn_mono((double)cjsrv, &roo2);
        auto gate = [&]() {
            int ndraw = 0;
            ryr_gating<LPC>(C.seed, (uint32_t)n, (uint64_t)nstep, C.Ku, C.Ku2, C.Kb, cp_d, (double)cjsrv, mono, dt,
                            nou, ncu, nob, ncb, &nfail, &ndraw, hl);
            if constexpr (LPC == 2) nfail += (int)lane_pair_u32((uint32_t)nfail);
            if (prevsp < 5 && (nou + nob) >= 5) nspark = 1;
        };
        /* spark release (:1286-1293) */
        double xire;
        auto release = [&]() {
            RYR_STRICT_FP
            double po = (double)(nou + nob) / (double)(RYR_NRYR);
            double Jmax = (0.0147 * 11.5);
            xire = 1.0 * Jmax * po * ((double)cjsrv - cp_d) / (0.00126) / 1.0;
            if (xire < 0) xire = 0;
        };
        /* exchange fluxes and JSR (:1309-1318) */
        TS diffjn0, diffjn1;
        TS dotcjsr, cjsrb_new;
        double dotcjsr_d, cjsrb_new_d;
        double jbeta;                               /* the CSQN buffering factor */
        auto jsr_beta = [&]() {
            if constexpr (!vf_early) { cnl[0] = cn0[g]; cnl[1] = cn0[P + g]; }
            if constexpr (sizeof(TF) == 8) {
                diffjn0 = (cjsrv - cnl[0]) / TS(25.0 * 1.0 / 4.) / TS(2.);
                diffjn1 = (cjsrv - cnl[1]) / TS(25.0 * 1.0 / 4.) / TS(2.);
            } else {
                diffjn0 = (cjsrv - cnl[0]) * TS(1.0 / (25.0 * 1.0 / 4.) / 2.);
                diffjn1 = (cjsrv - cnl[1]) * TS(1.0 / (25.0 * 1.0 / 4.) / 2.);
            }
            {
                RYR_STRICT_FP
                const double cj = (double)cjsrv;
                double ene = mono * 20 + (1.0 - mono) * 30;
                double enedot = (20 - 30) * (-mono + 1.0 / (4.0 * 460.0 * mono * roo2 + 1.0)) * (23.0 / cj)
                                * (1.0 - roo2 / 5000.0);
                jbeta = 1.0 / (1.0 + (460.0 * 600.0 * ene + 460.0 * enedot * cj * (cj + 600.0))
                                      / ((600.0 + cj) * (600.0 + cj)));
            }
        };
        auto jsr_flux = [&]() {
            RYR_STRICT_FP
            double flux = ((-(double)diffjn0 - (double)diffjn1) - xire * (0.00126) * 1.0 / (0.02 * 1));
            dotcjsr_d = jbeta * flux;
            dotcjsr = (TS)dotcjsr_d;
            cjsrb_new_d = (double)(EARLY ? e_cjb : ryr_ld_if<NTC>(cjsrbp + g)) + (1 - jbeta) * flux * dt;
            cjsrb_new = (TS)cjsrb_new_d;
        };
        /* VNEXT: what needs no V is formed while V is being published.  What
         * needs no gating result either comes before the gating, so that a
         * wave whose gating runs long is not delayed further after it: the
         * LCC draws of channels 0-3 over dt with the fCa closing rate
         * (lcc_step's V-free parts) and the CSQN buffering factor; then the
         * release and the JSR update */
        double rig_pre[LPC == 1 ? 4 : 2] = {};
        double bfca_pre = 0.0, cpx_pre = 0.0, cpxk_pre = 0.0;
        if constexpr (VNEXT) {
            if (!perm) {
                RYR_STRICT_FP
                cpx_pre = (cp_d + cs_d * (0.5)) / (1 + (0.5));
                cpxk_pre = cpx_pre / 1000.0;
                bfca_pre = lcc_betafca(cpx_pre);
                if constexpr (LPC == 1) {
#pragma unroll
                    for (int ll = 0; ll < 4; ++ll) {
                        const ryr_u32x4 &lb = lpre[ll >> 1];
                        rig_pre[ll] = lcc_rig((ll & 1) ? ryr_slot_uniform(lb.v[2], lb.v[3])
                                                       : ryr_slot_uniform(lb.v[0], lb.v[1]), dt);
                    }
                } else {
                    /* this lane's channels hl and hl + 2 draw from blocks 0 and 1 */
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const ryr_u32x4 lb = ryr_block(C.seed, (uint32_t)n, (uint64_t)nstep, (uint32_t)k);
                        rig_pre[k] = lcc_rig(hl ? ryr_slot_uniform(lb.v[2], lb.v[3])
                                                : ryr_slot_uniform(lb.v[0], lb.v[1]), dt);
                    }
                }
            }
            jsr_beta();
            /* each value is held where it is formed (ryr_pin), so that the
             * compiler does not sink it past the gating or the wait to its
             * use */
            ryr_pin(jbeta); ryr_pin(bfca_pre); ryr_pin(cpx_pre); ryr_pin(cpxk_pre);
#pragma unroll
            for (int k = 0; k < (LPC == 1 ? 4 : 2); ++k) ryr_pin(rig_pre[k]);
            gate();
            release();
            jsr_flux();
            ryr_pin(xire); ryr_pin(dotcjsr_d); ryr_pin(cjsrb_new_d);
            cru_v_next_get(a, C, nstep, stim_amp, &v, &vf0);
        }

        /* LCC gating and I_CaL (cuda_Cell_Detailed.cu:1248-1272): the slots
         * up to the first absent one, as the reference's loop; channel l
         * draws static slot l */
        double icat1 = 0.0, xicat, bcgcur = 0.0;
        {
            RYR_STRICT_FP
            if (!perm) {
                const RyrVF vf = vf_early ? vf0 : load_vf(M);
                const double cpx = VNEXT ? cpx_pre : (cp_d + cs_d * (0.5)) / (1 + (0.5));
                const double betac = a.iso2 ? 2.0 : 4.0;
                int nlcp = 0;
                uint32_t nw = lw;
                ryr_u32x4 lb = {{0u, 0u, 0u, 0u}};
                if constexpr (LPC == 1) {
                    const int npre = __builtin_ctz(~lmask);          /* present slots before the first absent */
                    /* channels 0-3 (every slot of the default four-LCC CRU)
                     * are stepped with no branch between them, so that their
                     * divisions interleave; a result past npre is dropped */
                    int jlo[4];
#pragma unroll
                    for (int ll = 0; ll < 4; ++ll) {
                        const int st = (int)((lw >> (4 * ll)) & 15u);
                        if (!(ll & 1))
                            lb = vf_early ? lpre[ll >> 1]
                                          : ryr_block(C.seed, (uint32_t)n, (uint64_t)nstep, (uint32_t)(ll >> 1));
                        const double ragg = (ll & 1) ? ryr_slot_uniform(lb.v[2], lb.v[3]) : ryr_slot_uniform(lb.v[0], lb.v[1]);
                        if constexpr (VNEXT) jlo[ll] = lcc_pick(st, rig_pre[ll], bfca_pre, vf, betac);
                        else jlo[ll] = lcc_step(st, cpx, dt, ragg, vf, betac);
                    }
#pragma unroll
                    for (int ll = 0; ll < 4; ++ll)
                        if (ll < npre) {
                            nw = (nw & ~(15u << (4 * ll))) | ((uint32_t)jlo[ll] << (4 * ll));
                            if (jlo[ll] == 0) ++nlcp;
                        }
#pragma unroll
                    for (int ll = 4; ll < 8; ++ll) {
                        if (ll < npre) {
                            int st = (int)((lw >> (4 * ll)) & 15u);
                            if (!(ll & 1)) lb = ryr_block(C.seed, (uint32_t)n, (uint64_t)nstep, (uint32_t)(ll >> 1));
                            const double ragg = (ll & 1) ? ryr_slot_uniform(lb.v[2], lb.v[3]) : ryr_slot_uniform(lb.v[0], lb.v[1]);
                            /* VNEXT: the fCa rate formed before the wait, which the
                             * compiler would otherwise hoist here, onto the
                             * critical path, for channels that are rarely present */
                            int jj = VNEXT ? lcc_pick(st, lcc_rig(ragg, dt), bfca_pre, vf, betac)
                                           : lcc_step(st, cpx, dt, ragg, vf, betac);
                            nw = (nw & ~(15u << (4 * ll))) | ((uint32_t)jj << (4 * ll));
                            if (jj == 0) ++nlcp;
                        }
                    }
                } else {
                    /* channels of this lane's parity among the present slots
                     * before the first absent one */
                    const int npre = __builtin_ctz(~lmask);
                    for (int ll = hl; ll < npre; ll += 2) {
                        int st = (int)((lw >> (4 * ll)) & 15u);
                        bool pre = false;
                        if constexpr (VNEXT) pre = ll < 4;
                        int jj;
                        if (pre) {
                            jj = lcc_pick(st, ll < 2 ? rig_pre[0] : rig_pre[1], bfca_pre, vf, betac);
                        } else {
                            lb = ryr_block(C.seed, (uint32_t)n, (uint64_t)nstep, (uint32_t)(ll >> 1));
                            const double ragg = (ll & 1) ? ryr_slot_uniform(lb.v[2], lb.v[3]) : ryr_slot_uniform(lb.v[0], lb.v[1]);
                            jj = VNEXT ? lcc_pick(st, lcc_rig(ragg, dt), bfca_pre, vf, betac)
                                       : lcc_step(st, cpx, dt, ragg, vf, betac);
                        }
                        nw = (nw & ~(15u << (4 * ll))) | ((uint32_t)jj << (4 * ll));
                        if (jj == 0) ++nlcp;
                    }
                    /* the partner lane gated the other parity's channels */
                    uint32_t pm = 0u;
#pragma unroll
                    for (int ll = 0; ll < 8; ++ll)
                        if (ll < npre && (ll & 1) != hl) pm |= 15u << (4 * ll);
                    nw = (nw & ~pm) | (lane_pair_u32(nw) & pm);
                    nlcp += (int)lane_pair_u32((uint32_t)nlcp);
                }
                nl += nlcp;
                lw = nw;
                double X_ = (VNEXT ? cpxk_pre : cpx / 1000.0) * vf.ghk_e2 - (1 * 1.8);
                double ica = vf.ghk_small ? 2.0 * 11.9 * 96.485 * 0.341 * X_ : vf.ghk_A * X_ / vf.ghk_D;
                if (ica > 0.0) ica = 0.0;
                icat1 = C.svica * (double)(nlcp) * ica;
            }
            /* slowncxchange background Ca current (:1266-1268) */
            if (C.variants & RYR_VAR_SLOWNCX)
                bcgcur = 1.0 * 0.003016 / 32. * (v - ryr_log(180000 / cs_d) / 2. / (96.485 / 8.314 / 308.0));
            xicat = icat1 + 0.0 + bcgcur * ((0.025) / (0.00126));
        }

        /* RyR gating (unless done above, or by the gating launch: GSPLIT),
         * spark, release (:1274-1293) */
        if constexpr (!GSPLIT && !VNEXT) gate();
        if constexpr (!VNEXT) release();

        /* NCX (:1294-1304, subroutine_cuda.c:47-75): FP64 in the exact path,
         * float with reciprocals in the production path (the per-CRU current
         * still enters an FP64 compensated sum) */
        double xinaca;
        TF ancx_new;
        if (perm) {                                 /* no NCX (:1302) */
            xinaca = 0.0;
            ancx_new = EARLY ? e_ancx : ryr_ld_if<NTC>(ancx + g);
        } else if constexpr (sizeof(TF) != 8) {
            const RyrVF vf = vf_early ? vf0 : load_vf(M);
            const float csm = (float)cs_d * 1e-3f;
            const float An = (float)(EARLY ? e_ancx : ryr_ld_if<NTC>(ancx + g));
            /* slowncxchange (subroutine_cuda.c:59-62): the exchanger at 5 x the allosteric factor */
            const float Ka = (C.variants & RYR_VAR_SLOWNCX) ? An * 5.0f : An;
            const float t2 = 669921.875f * csm * (1.0f + csm * (float)(1.0 / 0.00359));
            const float t3 = (float)vf.ncx_t3a + csm * 2744000.0f;
            const float Inaca = Ka * 21.0f * ((float)vf.ncx_n1 - (float)vf.ncx_n2 * csm) *
                                __builtin_amdgcn_rcpf(((float)vf.ncx_t1 + t2 + t3) * (float)vf.ncx_den2);
            const float r = 3e-4f * __builtin_amdgcn_rcpf(csm);
            const float Ancxdot = (__builtin_amdgcn_rcpf(1.0f + r * r * r) - An) * (1.0f / 150.0f);
            ancx_new = (TF)(An + Ancxdot * (float)dt);
            xinaca = (double)((float)C.svncx * Inaca);
        } else {
            RYR_STRICT_FP
            const RyrVF vf = vf_early ? vf0 : load_vf(M);
            double cs = cs_d / 1000.0;
            const double An = (double)(EARLY ? e_ancx : ryr_ld_if<NTC>(ancx + g));
            double Ka = An;
            if (C.variants & RYR_VAR_SLOWNCX) Ka = An * 5;  /* subroutine_cuda.c:59-62 */
            double t2 = (87.5) * (87.5) * (87.5) * cs * (1.0 + cs / 0.00359);
            double t3 = vf.ncx_t3a + cs * ((140.0) * (140.0) * (140.0));
            double Inaca = Ka * (21.0 / 1.0) * (vf.ncx_n1 - vf.ncx_n2 * cs)
                           / ((vf.ncx_t1 + t2 + t3) * vf.ncx_den2);
            double r = 0.0003 / cs;
            double Ancxdot = (1.0 / (1.0 + r * r * r) - An) / 150.;
            ancx_new = (TF)(An + Ancxdot * dt);
            xinaca = C.svncx * 1.0 * Inaca;
        }
        if (lead) ryr_st_if<NTC>(ancx + g, ancx_new);

        if constexpr (!VNEXT) { jsr_beta(); jsr_flux(); }
        const TF cpT = cpv, csT = csv;
        const TF dtT = (TF)dt;

        /* cleft/sub-membrane sub-steps (:1356-1473 / :1475-1597); the
         * cytosol/NSR sub-voxels were advanced by k_cyto */
        TF v_casar = EARLY ? e_sar : ryr_ld_if<NTC>(casar + g), v_casarh = EARLY ? e_sarh : ryr_ld_if<NTC>(casarh + g);
        TF v_casarj = EARLY ? e_sarj : ryr_ld_if<NTC>(casarj + g), v_casarhj = EARLY ? e_sarhj : ryr_ld_if<NTC>(casarhj + g);
        const TF csb = perm ? (TF)C.ci_bath : csv;             /* perm: bath outside the lattice */
        const TF cs_up = has_up ? (EARLY ? e_up : cs0[g + sz]) : csb;   /* k + 1 < Z */
        const TF cs_dn = has_dn ? (EARLY ? e_dn : cs0[g - sz]) : csb;   /* k > 0 */
        const TF bcgT = (TF)bcgcur;
        const TF icatT = (TF)icat1, xinacaT = (TF)xinaca, xireT = (TF)xire;
        TF csn = csv;
        const TF *sbp = (const TF *)a.sb;
        if constexpr (sizeof(TF) != 8) {
            float dci[2] = {0.0f, 0.0f};
            if (SCHB) { dci[0] = (float)(EARLY ? e_sb0 : sbp[g]); dci[1] = (float)(EARLY ? e_sb1 : sbp[P + g]); }
            float cpf = (float)cpv, csf = (float)csv;
            float sar = (float)v_casar, sarh = (float)v_casarh, sarj = (float)v_casarj, sarhj = (float)v_casarhj;
            fast_cleft<!SCHB>((float)dt, (float)cil[0], (float)cil[1], (float)csv, (float)cs_up,
                       (float)cs_dn, (float)icat1, (float)xinaca, (float)xire, (float)bcgcur, perm, (float)C.ci_bath,
                       cpf, csf, sar, sarh, sarj, sarhj, dci[0], dci[1]);
            if (SCHB && lead) {
                ci1[g] = (TF)fmaxf((float)cil[0] + dci[0] * (float)dt, 0.0f);
                ci1[P + g] = (TF)fmaxf((float)cil[1] + dci[1] * (float)dt, 0.0f);
            }
            cpv = (TF)cpf; csn = (TF)csf;
            v_casar = (TF)sar; v_casarh = (TF)sarh; v_casarj = (TF)sarj; v_casarhj = (TF)sarhj;
        } else {
        TF dotci01[2] = {TF(0), TF(0)};
        if (SCHB) { dotci01[0] = EARLY ? e_sb0 : sbp[g]; dotci01[1] = EARLY ? e_sb1 : sbp[P + g]; }
        const TF dth = dtT / TF(5);
        if (!SCHB) {
            /* scheme A (:1396-1464) */
            const TF cpstart = cpT;
            for (int it = 0; it < 5; ++it) {
                const TF fsar = b_busar<TF>(v_casar, csn, dth);
                const TF fsarh = b_busarh<TF>(v_casarh, csn, dth);
                const TF csbuffers = (fsar + fsarh);
                TF csdiff = (cs_up + cs_dn - TF(2) * csv) / (TF(1.42));
                if (perm) csdiff += (TF(C.ci_bath) - csn) / TF(5);      /* :1409-1411 */
                const TF dotcs = TF((0.00126) / (0.025)) * (cpv - csn) / (TF(0.0283))
                                 - TF(0.5) * icatT * TF(0.00126) / TF(0.025)
                                 + xinacaT
                                 - (csv - cil[0]) / TF(0.04) / TF(2.)
                                 - (csv - cil[1]) / TF(0.04) / TF(2.)
                                 + csdiff
                                 - csbuffers
                                 - bcgT;
                const TF cpbuff = b_busar<TF>(v_casarj, cpv, dth);
                const TF cpbuffh = b_busarh<TF>(v_casarhj, cpv, dth);
                const TF cpdot = ((csn - cpv) / TF(0.0283) +
                                  (cil[0] - cpstart) / TF(0.1) / TF(2.) +
                                  (cil[1] - cpstart) / TF(0.1) / TF(2.) +
                                  xireT - icatT - cpbuff - cpbuffh);
                v_casarj += cpbuff * dtT / TF(5);
                if (v_casarj < TF(0)) v_casarj = TF(0);
                v_casarhj += cpbuffh * dtT / TF(5);
                if (v_casarhj < TF(0)) v_casarhj = TF(0);
                cpv += cpdot * dtT / TF(5);
                v_casar += fsar * dtT / TF(5);
                if (v_casar < TF(0)) v_casar = TF(0);
                v_casarh += fsarh * dtT / TF(5);
                if (v_casarh < TF(0)) v_casarh = TF(0);
                csn += (dotcs) * dth;
            }
        } else {
            /* scheme B (:1479-1553) */
            for (int it = 0; it < 5; ++it) {
                const TF dpi0 = (cpv - cil[0]) / TF(0.1) / TF(2.);
                const TF dpi1 = (cpv - cil[1]) / TF(0.1) / TF(2.);
                const TF dsi0 = (csn - cil[0]) / TF(0.04) / TF(2.);
                const TF dsi1 = (csn - cil[1]) / TF(0.04) / TF(2.);
                dotci01[0] += dsi0 * TF((0.025) / (0.5 / 8.)) / TF(5);
                dotci01[1] += dsi1 * TF((0.025) / (0.5 / 8.)) / TF(5);
                dotci01[0] += dpi0 * TF((0.00126) / (0.5 / 8.)) / TF(5);
                dotci01[1] += dpi1 * TF((0.00126) / (0.5 / 8.)) / TF(5);
                const TF fsar = b_busar<TF>(v_casar, csn, dth);
                const TF fsarh = b_busarh<TF>(v_casarh, csn, dth);
                const TF csbuffers = (fsar + fsarh);
                TF csdiff = (cs_up + cs_dn - TF(2) * csn) / (TF(1.42));
                if (perm) csdiff += (TF(C.ci_bath) - csn) / TF(5);      /* :1499-1501 */
                const TF dotcs = TF((0.00126) / (0.025)) * (cpv - TF(0.5 * 0.0283) * icatT - csn) / (TF(0.0283))
                                 + xinacaT
                                 - (csn - cil[0]) / TF(0.04) / TF(2.)
                                 - (csn - cil[1]) / TF(0.04) / TF(2.)
                                 + csdiff
                                 - csbuffers
                                 - bcgT;
                const TF cpbuff = b_busar<TF>(v_casarj, cpv, dth);
                const TF cpbuffh = b_busarh<TF>(v_casarhj, cpv, dth);
                const TF cpdot = ((csT - cpv) / TF(0.0283) +
                                  (cil[0] - cpv) / TF(0.1) / TF(2.) +
                                  (cil[1] - cpv) / TF(0.1) / TF(2.) +
                                  xireT - icatT - cpbuff - cpbuffh);
                v_casar += fsar * dth;
                if (v_casar < TF(0)) v_casar = TF(0);
                v_casarh += fsarh * dth;
                if (v_casarh < TF(0)) v_casarh = TF(0);
                csn += (dotcs) * dth;
                v_casarj += cpbuff * dth;
                if (v_casarj < TF(0)) v_casarj = TF(0);
                v_casarhj += cpbuffh * dth;
                if (v_casarhj < TF(0)) v_casarhj = TF(0);
                cpv += cpdot * dth;
            }
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
                TF cin = cil[ii] + dotci01[ii] * dtT;
                if (cin < TF(0)) cin = TF(0);
                if (lead) ci1[ii * P + g] = cin;
            }
        }
        }   /* exact path */
        if (cpv < TF(0)) { cpv = TF(0); ++nclampcp; }
        if (csn < TF(0)) { csn = TF(1e-6); ++nclampcs; }
        if (lead) {
            ryr_st_if<NTC>(casar + g, v_casar); ryr_st_if<NTC>(casarh + g, v_casarh);
            ryr_st_if<NTC>(casarj + g, v_casarj); ryr_st_if<NTC>(casarhj + g, v_casarhj);
            ryr_st_if<NTC>(cp1 + g, cpv);
            ryr_st_if<NTC>(cs1 + g, csn);
            if constexpr (sizeof(TS) == 4) {
                /* float JSR pools (FP32 production): stochastic rounding of the double update */
                ryr_st_if<NTC>(cjsr1 + g, ryr_sr_f32((double)cjsrv + dotcjsr_d * dt,
                                                ryr_sr_bits(ryr_sr_key(C.seed, (uint64_t)nstep, RYR_SR_CJSR), (uint32_t)n, 0u)));
                ryr_st_if<NTC>(cjsrbp + g, ryr_sr_f32(cjsrb_new_d,
                                                 ryr_sr_bits(ryr_sr_key(C.seed, (uint64_t)nstep, RYR_SR_CJSRB), (uint32_t)n, 0u)));
            } else {
                ryr_st_if<NTC>(cjsr1 + g, (TS)(cjsrv + dotcjsr * (TS)dtT));
                ryr_st_if<NTC>(cjsrbp + g, cjsrb_new);
            }
            /* the gating words' bases pass through late_base, so that these
             * stores form their addresses here instead of keeping the loads'
             * 64-bit addresses live across the kernel (which spills in the
             * 168-VGPR instances) */
            ryr_st_if<NTC>(late_base(a.lccw) + g, lw);
            if (!GSPLIT)
                ryr_st_if<NTC>(late_base(a.ryrw) + g,
                          (uint32_t)nou | ((uint32_t)ncu << 8) | ((uint32_t)nob << 16) | ((uint32_t)ncb << 24));
            ryr_st_if<NTC>(late_base(a.aux) + g,
                      (uint32_t)(nl & 0xffff) | ((uint32_t)nspark << 16) | (lmask << RYR_AUX_MASK_SHIFT));
            if (diag) {
                a.d_xire[g] = xire;
                a.d_xinaca[g] = xinaca;
            }
            if (nfail | nclampcp | nclampcs) {
                unsigned long long *cnt = a.counters + (long long)c * RYR_NCNT;
                if (nfail) atomicAdd(cnt + 0, (unsigned long long)nfail);
                if (nclampcp) atomicAdd(cnt + 1, 1ull);
                if (nclampcs) atomicAdd(cnt + 2, 1ull);
            }
            r_icat = xicat;
            r_inaca = xinaca;
            r_cs = (double)csn;
        }
    }

    /* whole-cell sums: pairwise double-double tree over each group of 64
     * CRUs (index order), one partial per group (4 per chunk, whatever the
     * shape); k_membrane pairs them into chunk sums and adds those.  With
     * paired lanes a wave holds 32 CRUs at its even lanes and zeros at the odd
     * ones (adding an exact zero changes no double-double), so its tree is the
     * group tree's lower or upper half, and the two halves meet in LDS. */
    {
        const int lane = threadIdx.x & 63;
        const int wave = (int)threadIdx.x >> 6;
        RyrDD t0 = {r_icat, 0.0}, t1 = {r_inaca, 0.0}, t2 = {r_cs, 0.0};
        t0 = wave_tree(t0);
        t1 = wave_tree(t1);
        t2 = wave_tree(t2);
        if constexpr (LPC == 1) {
            if (lane == 0) {
                RyrDD *pp = step_partials(a, nstep) + 12LL * bid + 3 * (part * (WG / 64) + wave);
                pp[0] = t0; pp[1] = t1; pp[2] = t2;
            }
        } else {
            __shared__ RyrDD s_half[WG / 64][3];
            if (lane == 0) { s_half[wave][0] = t0; s_half[wave][1] = t1; s_half[wave][2] = t2; }
            __syncthreads();
            if (lane == 0 && !(wave & 1)) {
                RyrDD *pp = step_partials(a, nstep) + 12LL * bid + 3 * ((part * (WG / 64) + wave) / 2);
#pragma unroll
                for (int q = 0; q < 3; ++q) pp[q] = dd_add(s_half[wave][q], s_half[wave + 1][q]);
            }
        }
    }
}

/* The 64- and 128-lane scheme-A instances, which the shape tuner picks for
 * batches (throughput), are held to 168 VGPRs: three waves per SIMD, without
 * spilling.  The 256-lane instance, which it picks for single cells (one
 * wave per SIMD anyway), keeps the registers the allocator wants. */
template <typename TF, typename TS, int WG, bool SCHB, int LPC = 1, bool GSPLIT = false>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu((WG >= 256 || SCHB) ? 1 : 3)))
void k_cru_step(RyrKArgs a, int kc)
{
    cru_block<TF, TS, WG, SCHB, LPC, GSPLIT>(a, kc, xcd_remap((int)blockIdx.x, (int)gridDim.x), 0.0, a.chunks, a.gstep);
}

/* ------------------------------------------------------------ membrane kernel */
/* One workgroup per cell, after k_cru_step: the chunk partials of the three
 * whole-cell sums (pairwise trees over groups of 256 chunks, groups in order),
 * then the membrane ODE's four current groups on the workgroup's waves, and on
 * wavefront 0 [Na]i and V, the clamp protocol, the beat tracker, the membrane
 * columns of an output row and the V-only factors of the next step
 * (cuda_Cell_Detailed.cu:430-739). */
template <typename TF, typename TS, int WG>
RYR_DEV void membrane_block(const RyrKArgs &a, int c, long long nstep, double stim_amp)
{
    constexpr int NW = WG / 64;
    __shared__ RyrDD s_red[3][4];
    const RyrCellInfo C = a.cells[c];
    RyrMem *M = a.mem + c;
    const int spo = a.spo;
    const bool diag = (nstep % spo) == 0;
    const double dt = a.dt;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    RyrDD acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    const RyrDD *pp = step_partials(a, nstep) + 12LL * C.first_block;
    __shared__ double s_ms[RYR_MS_N];
    MemIn I;
    I.s = M->m; I.svikr = C.svikr; I.Cm = C.Cm; I.dN = C.dN; I.dt = dt;
    I.slowncx = (C.variants & RYR_VAR_SLOWNCX) != 0;

    /* Slice sl of a group of 256 chunks (chunks 64 sl .. 64 sl + 63) is summed
     * by wave NW - 1 - (sl mod NW): the slices go first to the waves whose
     * current groups are shortest (I_NaK on wave 3), and a slice past the
     * cell's last chunk is zero without a tree.  Which wave sums a slice does
     * not change the sum.  This wave's first slice of wave partials is loaded
     * before the current groups run, so that the loads' latency hides behind
     * them. */
    const int sl0 = NW - 1 - wv;
    RyrDD w0[3][4];
    {
        const int bb = sl0 * 64 + lane;
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (bb < C.nblocks) w0[q][k] = pp[12LL * bb + q + 3 * k];
                else { w0[q][k].hi = 0.0; w0[q][k].lo = 0.0; }
            }
    }

    /* the four current groups (none needs the whole-cell sums), group g on
     * wave g mod NW */
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        if (g % NW != wv) continue;
        if (g == 0) mem_ina(I, lane, s_ms);
        else if (g == 1) mem_ikr_iks(I, lane, s_ms);
        else if (g == 2) mem_ik1_ito(I, lane, s_ms);
        else mem_inak(I, lane, s_ms);
    }

    /* groups of 256 chunk sums, each a pairwise tree of four 64-chunk wave
     * trees (the same tree whatever the number of waves WG gives) */
    for (int g0 = 0; g0 < C.nblocks; g0 += 256) {
        for (int sl = sl0; sl < 4; sl += NW) {
            const int bb = g0 + sl * 64 + lane;
            RyrDD t[3];
            if (g0 + sl * 64 >= C.nblocks) {
                if (lane == 0)
                    for (int q = 0; q < 3; ++q) { s_red[q][sl].hi = 0.0; s_red[q][sl].lo = 0.0; }
                continue;
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                /* chunk sum = pairwise tree over its four wave partials */
                if (g0 == 0 && sl == sl0) {
                    t[q] = dd_add(dd_add(w0[q][0], w0[q][1]), dd_add(w0[q][2], w0[q][3]));
                } else if (bb < C.nblocks) {
                    const RyrDD *w = pp + 12LL * bb + q;
                    t[q] = dd_add(dd_add(w[0], w[3]), dd_add(w[6], w[9]));
                } else { t[q].hi = 0.0; t[q].lo = 0.0; }
                t[q] = wave_tree(t[q]);
            }
            if (lane == 0) { s_red[0][sl] = t[0]; s_red[1][sl] = t[1]; s_red[2][sl] = t[2]; }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 3; ++q)
            acc[q] = dd_add(acc[q], dd_add(dd_add(s_red[q][0], s_red[q][1]), dd_add(s_red[q][2], s_red[q][3])));
        __syncthreads();
    }
    /* wave 0: the parts that need the sums, then [Na]i and V.  With more than
     * one wave, wave 1 then forms the next step's V-only factors while wave 0
     * stores the membrane state and tracks the beat. */
    __shared__ double s_vx[2];
    const double stim = stim_of(C, nstep, stim_amp);
    double vnew = 0.0, xnai_new = 0.0, mo[13], cu[10];
    if (wv == 0) {
        I.sum_icat = acc[0].hi; I.sum_inaca = acc[1].hi; I.sum_cs = acc[2].hi;
        mem_sums(I, lane, s_ms);
        mem_finish(M->m, s_ms, dt, stim, &mo[0], &mo[12]);
#pragma unroll
        for (int q = 1; q < 12; ++q) mo[q] = s_ms[q];
#pragma unroll
        for (int q = 0; q < 10; ++q) cu[q] = s_ms[RYR_MS_CSUBT + q];
        vnew = mo[0];
        if (C.vclamp_n > 0) {
            /* clamp protocols (cuda_Cell_Detailed.cu:733-773) */
            long long q = C.vclamp_periodic ? nstep % C.vclamp_n : (nstep < C.vclamp_n ? nstep : C.vclamp_n - 1);
            vnew = C.vclamp[q];
        }
        xnai_new = mo[12];
        if (lane == 0) { s_vx[0] = vnew; s_vx[1] = xnai_new; }
    }
    if constexpr (NW > 1) {
        __syncthreads();
        if (wv == 1) vfactors_wave(M, a.iso2, lane, s_vx[0], s_vx[1]);
    }
    if (wv != 0) return;
    if (lane == 0) {
        const double vold = M->m[0];
        for (int q = 0; q < 13; ++q) M->m[q] = mo[q];
        M->m[0] = vnew;
        for (int q = 0; q < 10; ++q) M->cur[q] = cu[q];
        M->stim = stim;
        M->out_ica += M->cur[1];
        M->out_ncx += M->cur[2];
        M->out_ina += M->cur[3];
        beat_track(M, C, nstep, dt, vold, vnew, a.beats + (long long)c * RYR_MAXB * RYR_BCOLS);
        if (!(vnew == vnew) || vnew == __builtin_inf() || vnew == -__builtin_inf())
            atomicAdd(a.counters + (long long)c * RYR_NCNT + 4, 1ull);
        if (diag) {
            /* the membrane columns of the output row (:1041-1072) */
            double *r = a.diag + (long long)c * 45;
            double spod = (double)spo;
            r[0] = (double)nstep * dt; r[2] = vnew;
            r[3] = M->out_ncx / spod; r[4] = M->out_ica / spod; r[6] = M->cur[0];
            r[10] = M->m[12]; r[11] = M->cur[4]; r[12] = M->cur[5]; r[13] = M->cur[6];
            r[14] = M->cur[7]; r[15] = M->cur[8]; r[16] = M->cur[9];
            r[17] = M->out_ina / spod; r[18] = M->m[4]; r[38] = M->m[5]; r[39] = M->m[7];
            M->out_ica = 0; M->out_ncx = 0; M->out_ina = 0;
            a.diag_step[c] = nstep;
        }
    }
    if constexpr (NW == 1) vfactors_wave(M, a.iso2, lane, vnew, xnai_new);
    if (lane == 0) M->step = nstep + 1;
}

/* ------------------------------------------- the one-launch schedule's parts
 * A launch of step n runs, side by side, the sub-voxel sweep of step n, the
 * CRU update of step n and one membrane workgroup per cell.  The CRU
 * workgroups need V of step n, which the membrane step of step n-1 forms
 * from step n-1's whole-cell sums; instead of waiting for it they form it
 * themselves from those sums and from the values the membrane workgroup of
 * the previous launch left pending (RyrPend: the state at the start of step
 * n-1 and its current groups), with the same expressions in the same order,
 * so every workgroup gets the same bits.  The membrane workgroup of the launch
 * finishes step n-1 (stores the state, tracks the beat) and forms the current
 * groups of step n for the next launch.  The first launch of a run, whose
 * membrane state is complete, only forms those groups. */

/* The three whole-cell sums of step nstep (I_CaL, I_NCX, cs) over the cell's
 * chunk partials: the tree and order of membrane_block.  Every thread of the
 * workgroup calls it; each gets the sums.  The 64-chunk slices go to waves
 * W0..NW-1 (a wave below W0 only meets the barriers, so it may arrive late). */
template <int NW, int W0 = 0>
RYR_DEV void reduce_cell_sums(const RyrKArgs &a, const RyrCellInfo &C, long long nstep, RyrDD acc[3])
{
    static_assert(W0 >= 0 && W0 < NW, "at least one wave forms the slices");
    __shared__ RyrDD s_red2[3][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const RyrDD *pp = step_partials(a, nstep) + 12LL * C.first_block;
    for (int q = 0; q < 3; ++q) { acc[q].hi = 0.0; acc[q].lo = 0.0; }
    for (int g0 = 0; g0 < C.nblocks; g0 += 256) {
        for (int sl = wv - W0; sl >= 0 && sl < 4; sl += NW - W0) {
            const int bb = g0 + sl * 64 + lane;
            if (g0 + sl * 64 >= C.nblocks) {
                if (lane == 0)
                    for (int q = 0; q < 3; ++q) { s_red2[q][sl].hi = 0.0; s_red2[q][sl].lo = 0.0; }
                continue;
            }
            RyrDD t[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (bb < C.nblocks) {
                    const RyrDD *w = pp + 12LL * bb + q;
                    t[q] = dd_add(dd_add(w[0], w[3]), dd_add(w[6], w[9]));
                } else { t[q].hi = 0.0; t[q].lo = 0.0; }
                t[q] = wave_tree(t[q]);
            }
            if (lane == 0) { s_red2[0][sl] = t[0]; s_red2[1][sl] = t[1]; s_red2[2][sl] = t[2]; }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 3; ++q)
            acc[q] = dd_add(acc[q], dd_add(dd_add(s_red2[q][0], s_red2[q][1]), dd_add(s_red2[q][2], s_red2[q][3])));
        __syncthreads();
    }
}

/* Wave 0: the sums-dependent part of the membrane step of step np from the
 * pending values pd and the sums acc: fills s_ms's sums slots (csubt, qks,
 * I_CaL, I_NCX), and returns V and [Na]i after it (V replaced by a clamp
 * protocol's value), as membrane_block does. */
RYR_DEV void pend_finish(const RyrCellInfo &C, const RyrPend *pd, const RyrDD acc[3], long long np, double dt,
                         double stim_amp, int lane, double *s_ms, double *v_out, double *nai_out)
{
    MemIn I;
    I.s = pd->s; I.svikr = C.svikr; I.Cm = C.Cm; I.dN = C.dN; I.dt = dt;
    I.slowncx = (C.variants & RYR_VAR_SLOWNCX) != 0;
    I.sum_icat = acc[0].hi; I.sum_inaca = acc[1].hi; I.sum_cs = acc[2].hi;
    mem_sums(I, lane, s_ms);
    const double stim = stim_of(C, np, stim_amp);
    double v, nai;
    mem_finish(pd->s, s_ms, dt, stim, &v, &nai);
    if (C.vclamp_n > 0) {
        long long q = C.vclamp_periodic ? np % C.vclamp_n : (np < C.vclamp_n ? np : C.vclamp_n - 1);
        v = C.vclamp[q];
    }
    *v_out = v;
    *nai_out = nai;
}

/* The current groups of the state s (13 values) on the workgroup's waves,
 * into pending slot pd (all threads call; s must stay unchanged meanwhile). */
template <int NW>
RYR_DEV void pend_groups(const RyrCellInfo &C, const double *s, double dt, RyrPend *pd)
{
    __shared__ double s_ms[RYR_MS_N];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    MemIn I;
    I.s = s; I.svikr = C.svikr; I.Cm = C.Cm; I.dN = C.dN; I.dt = dt;
    I.slowncx = (C.variants & RYR_VAR_SLOWNCX) != 0;
    I.sum_icat = 0.0; I.sum_inaca = 0.0; I.sum_cs = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        if (g % NW != wv) continue;
        if (g == 0) mem_ina(I, lane, s_ms);
        else if (g == 1) mem_ikr_iks(I, lane, s_ms);
        else if (g == 2) mem_ik1_ito(I, lane, s_ms);
        else mem_inak(I, lane, s_ms);
    }
    __syncthreads();
    for (int t = (int)threadIdx.x; t < 13 + RYR_MS_N; t += (int)blockDim.x) {
        if (t < 13) pd->s[t] = s[t];
        else pd->ms[t - 13] = s_ms[t - 13];
    }
}

/* Membrane workgroup of the first launch of a run (step nstep): the current
 * groups of the complete state, pending for the next launch. */
template <int NW>
RYR_DEV void membrane_first_block(const RyrKArgs &a, int c, const long long *gbase, int kc, const RyrCellInfo *cells)
{
    const RyrCellInfo C = cells[c];
    const long long nstep = *gbase + kc;
    __shared__ double s_st[13];
    if (threadIdx.x < 13) s_st[threadIdx.x] = a.mem[c].m[threadIdx.x];
    /* no V is published within this launch, and a value left by an earlier
     * run must not be taken for this run's: every tag becomes this step's,
     * which no later launch of the run waits for */
    if (threadIdx.x < RYR_VBOX_WORDS)
        __hip_atomic_store(&a.vbox[c].w[threadIdx.x], (unsigned long long)(unsigned)nstep << 32, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    pend_groups<NW>(C, s_st, a.dt, a.pend + 2LL * c + (nstep & 1));
}

/* Membrane workgroup of a later launch (step nstep): finishes the membrane
 * step of nstep-1 from its pending values and sums, stores the state and the
 * step's currents, tracks the beat (as membrane_block), then forms the
 * current groups of the new state, pending for the next launch. */
template <int NW>
RYR_DEV void membrane_next_block(const RyrKArgs &a, int c, const long long *gbase, int kc, double stim_amp,
                                 const RyrCellInfo *cells)
{
    const RyrCellInfo C = cells[c];
    const long long nstep = *gbase + kc;
    RyrMem *M = a.mem + c;
    const long long np = nstep - 1;
    const RyrPend *pd = a.pend + 2LL * c + (np & 1);
    const double dt = a.dt;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    /* the pending current groups, loaded while the sums form */
    const double pm = threadIdx.x < RYR_MS_N ? pd->ms[threadIdx.x] : 0.0;
    if (wv == 0) {
        /* V of step nstep and its V-only factors first, in wave 0 alone (the
         * tree and expressions of the block's own sums and ODE below, so the
         * same bits), published for this launch's CRU workgroups, which wait
         * for them after their RyR gating.  Waves 1-3 form the block's sums
         * meanwhile. */
        double vp;
        RyrVF f;
        cru_v_next_wave(a, C, nstep, stim_amp, &vp, &f);
        const double val[14] = {vp, f.alphad, f.betad, f.alphaf, f.betaf, f.ghk_A, f.ghk_e2, f.ghk_D, f.ncx_n1,
                                f.ncx_n2, f.ncx_t1, f.ncx_t3a, f.ncx_den2, (double)f.ghk_small};
        if (lane < RYR_VBOX_WORDS) {
            double x = val[0];
#pragma unroll
            for (int k = 1; k < 14; ++k)
                if ((lane >> 1) == k) x = val[k];
            const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
            const unsigned half = (lane & 1) ? (unsigned)(bits >> 32) : (unsigned)bits;
            __hip_atomic_store(&a.vbox[c].w[lane], ((unsigned long long)(unsigned)nstep << 32) | half,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    RyrDD acc[3];
    reduce_cell_sums<NW, 1>(a, C, np, acc);
    __shared__ double s_ms[RYR_MS_N];
    __shared__ double s_st[13];
    if (threadIdx.x < RYR_MS_N) s_ms[threadIdx.x] = pm;
    __syncthreads();
    if (wv == 0) {
        double vnew, nai;
        pend_finish(C, pd, acc, np, dt, stim_amp, lane, s_ms, &vnew, &nai);
        if (lane == 0) {
            double mo[13];
            mo[0] = vnew;
            for (int q = 1; q < 12; ++q) mo[q] = s_ms[q];
            mo[12] = nai;
            const double vold = pd->s[0];
            const double stim = stim_of(C, np, stim_amp);
            for (int q = 0; q < 13; ++q) { M->m[q] = mo[q]; s_st[q] = mo[q]; }
            for (int q = 0; q < 10; ++q) M->cur[q] = s_ms[RYR_MS_CSUBT + q];
            M->stim = stim;
            M->out_ica += M->cur[1];
            M->out_ncx += M->cur[2];
            M->out_ina += M->cur[3];
            beat_track(M, C, np, dt, vold, vnew, a.beats + (long long)c * RYR_MAXB * RYR_BCOLS);
            if (!(vnew == vnew) || vnew == __builtin_inf() || vnew == -__builtin_inf())
                atomicAdd(a.counters + (long long)c * RYR_NCNT + 4, 1ull);
            M->step = nstep;
        }
    }
    __syncthreads();
    pend_groups<NW>(C, s_st, dt, a.pend + 2LL * c + (nstep & 1));
}

/* In each wave of a CRU workgroup of a later launch (step nstep), with
 * every lane active: V of step nstep and its V-only factors, formed from
 * step nstep-1's pending values and sums with no LDS and no barrier, so that
 * the workgroup's state loads stay in flight meanwhile.  The sums are the
 * same tree as reduce_cell_sums (each 64-chunk slice a wave tree, slices
 * pairwise, groups of 256 chunks in order), and [Na]i and V the same
 * expressions as mem_sums and mem_finish. */
RYR_DEV void cru_v_next_wave(const RyrKArgs &a, const RyrCellInfo &C, long long nstep, double stim_amp,
                             double *v_out, RyrVF *vf_out)
{
    RYR_EXACT_FP
    const long long np = nstep - 1;
    const RyrPend *pd = a.pend + 2LL * C.cell + (np & 1);
    const int lane = threadIdx.x & 63;
    /* the pending state and current groups, loaded with the partial sums */
    double ps[13];
#pragma unroll
    for (int k = 0; k < 13; ++k) ps[k] = pd->s[k];
    const double m_ina = pd->ms[RYR_MS_INA], m_ik1 = pd->ms[RYR_MS_IK1], m_ikr = pd->ms[RYR_MS_IKR],
                 m_iks = pd->ms[RYR_MS_IKS], m_itos = pd->ms[RYR_MS_ITOS], m_itof = pd->ms[RYR_MS_ITOF],
                 m_inak = pd->ms[RYR_MS_INAK];
    const RyrDD *pp = step_partials(a, np) + 12LL * C.first_block;
    RyrDD acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int g0 = 0; g0 < C.nblocks; g0 += 256) {
        RyrDD s[2][4];
#pragma unroll
        for (int sl = 0; sl < 4; ++sl) {
            if (g0 + sl * 64 >= C.nblocks) {
                s[0][sl].hi = s[0][sl].lo = s[1][sl].hi = s[1][sl].lo = 0.0;
                continue;
            }
            /* both sums' trees side by side, with no branch between them (a
             * lane past the cell's last chunk reads that chunk and adds
             * zero), so that their dependent chains interleave */
            const int bb = g0 + sl * 64 + lane;
            const bool in = bb < C.nblocks;
            const RyrDD *w = pp + 12LL * (in ? bb : C.nblocks - 1);
            RyrDD t[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const RyrDD x = dd_add(dd_add(w[q], w[q + 3]), dd_add(w[q + 6], w[q + 9]));
                t[q].hi = in ? x.hi : 0.0;
                t[q].lo = in ? x.lo : 0.0;
            }
            wave_trees<2>(t);
            s[0][sl] = t[0];
            s[1][sl] = t[1];
        }
#pragma unroll
        for (int q = 0; q < 2; ++q)
            acc[q] = dd_add(acc[q], dd_add(dd_add(s[q][0], s[q][1]), dd_add(s[q][2], s[q][3])));
    }
    /* I_CaL and I_NCX of the sums (mem_sums), then [Na]i and V (mem_finish) */
    double n = 1.0, d = 1.0;
    if (lane == 1) { n = acc[0].hi; d = C.Cm; }
    if (lane == 2) { n = acc[1].hi; d = C.Cm; }
    const double q1 = n / d;
    double sh[RYR_MS_N];
#pragma unroll
    for (int k = 0; k < RYR_MS_N; ++k) sh[k] = 0.0;
    sh[RYR_MS_INA] = m_ina; sh[RYR_MS_IK1] = m_ik1;
    sh[RYR_MS_IKR] = m_ikr; sh[RYR_MS_IKS] = m_iks;
    sh[RYR_MS_ITOS] = m_itos; sh[RYR_MS_ITOF] = m_itof;
    sh[RYR_MS_INAK] = m_inak;
    sh[RYR_MS_ICAT] = lanev(q1, 1) * 0.0965 * (0.00126) * 2.0 * ((0.5) + 1.);
    sh[RYR_MS_INACA] = lanev(q1, 2) * 0.0965 * (0.025);
    double v, nai;
    mem_finish(ps, sh, a.dt, stim_of(C, np, stim_amp), &v, &nai);
    if (C.vclamp_n > 0) {
        long long q = C.vclamp_periodic ? np % C.vclamp_n : (np < C.vclamp_n ? np : C.vclamp_n - 1);
        v = C.vclamp[q];
    }
    *v_out = v;
    *vf_out = vfactors_regs(a.iso2, lane, v, nai);
}

/* V of step nstep for a CRU wave of a later launch (every lane active): the
 * membrane workgroup's published values once all 28 mailbox words carry this
 * step's tag (agent-scope loads, so the L2 holds no stale copy); if they do
 * not within a bounded number of polls, the wave forms them itself.  Either
 * way the bits are the same. */
#define RYR_VBOX_POLLS 4096
RYR_DEV void cru_v_next_get(const RyrKArgs &a, const RyrCellInfo &C, long long nstep, double stim_amp,
                            double *v_out, RyrVF *vf_out)
{
    const RyrVBox *bx = a.vbox + C.cell;
    const int lane = threadIdx.x & 63;
    const unsigned tag = (unsigned)nstep;
    /* lane k < 28 polls word k; the wave takes the words when every tag is
     * this step's */
    unsigned long long w = (unsigned long long)tag << 32;
    bool ok = false;
    for (int t = 0; t < RYR_VBOX_POLLS; ++t) {
        if (lane < RYR_VBOX_WORDS) w = __hip_atomic_load(&bx->w[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = __all((unsigned)(w >> 32) == tag);
        if (ok) break;
        __builtin_amdgcn_s_sleep(2);
    }
    if (ok) {
        const int half = (int)(unsigned)w;
        double f[14];
#pragma unroll
        for (int k = 0; k < 14; ++k) {
            const unsigned lo = (unsigned)__builtin_amdgcn_readlane(half, 2 * k);
            const unsigned hi = (unsigned)__builtin_amdgcn_readlane(half, 2 * k + 1);
            f[k] = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
        }
        *v_out = f[0];
        vf_out->alphad = f[1]; vf_out->betad = f[2]; vf_out->alphaf = f[3]; vf_out->betaf = f[4];
        vf_out->ghk_A = f[5]; vf_out->ghk_e2 = f[6]; vf_out->ghk_D = f[7]; vf_out->ncx_n1 = f[8];
        vf_out->ncx_n2 = f[9]; vf_out->ncx_t1 = f[10]; vf_out->ncx_t3a = f[11]; vf_out->ncx_den2 = f[12];
        vf_out->ghk_small = f[13] != 0.0;
        return;
    }
    cru_v_next_wave(a, C, nstep, stim_amp, v_out, vf_out);
}

/* the membrane step on its own: one workgroup per cell, step gstep + km */
template <typename TF, typename TS>
__global__ __launch_bounds__(RYR_BLOCK) void k_membrane(RyrKArgs a, int km, double stim_amp)
{
    membrane_block<TF, TS, RYR_BLOCK>(a, (int)blockIdx.x, *a.gstep + km, stim_amp);
}

/* One launch per step (cleft scheme A): membrane workgroups (one per cell),
 * then the CRU update of every 256-CRU chunk (RyR gating included), then the
 * sub-voxel sweep; NEXT: the membrane step of the previous step is pending
 * (see the one-launch schedule's parts above).  The head (membrane and CRU
 * workgroups, which set the step's length) comes first in dispatch order and
 * round-robin over the XCDs; it is padded to a multiple of eight blocks so
 * that the sweep behind it keeps the XCD-by-XCD layout (xcd_remap) in which
 * a chunk's eight planes share one L2. */
__host__ __device__ inline int one_head(int ncells, int nchunks) { return (ncells + nchunks + 7) & ~7; }

/* The step base, the record tables, the step offset and the grid's counts
 * (a.gstep, a.chunks, a.cells, kc, a.ncells, a.P / 256) lead the argument
 * list as plain values, so that they arrive in scalar registers with the
 * wave (kernel-argument preload) rather than by a load from the argument
 * segment: each workgroup's first dependent loads, the step base and its
 * record, then issue one memory round trip earlier. */
template <typename TF, typename TS, bool NEXT, int LPC>
__global__ __launch_bounds__(RYR_BLOCK) void k_one(const long long *gstep, const RyrCellInfo *chunks,
                                                   const RyrCellInfo *cells, int kc, int ncells, int nchunks,
                                                   RyrKArgs a, double stim_amp)
{
    /* each part reads the step base beside its own record, so that the two
     * loads are in flight together */
    const int b = (int)blockIdx.x;
    const int head = one_head(ncells, LPC * nchunks);
    if (b < ncells) {
        if constexpr (NEXT) membrane_next_block<RYR_BLOCK / 64>(a, b, gstep, kc, stim_amp, cells);
        else membrane_first_block<RYR_BLOCK / 64>(a, b, gstep, kc, cells);
        return;
    }
    if (b < ncells + LPC * nchunks) {
        cru_block<TF, TS, RYR_BLOCK, false, LPC, false, NEXT>(a, kc, b - ncells, stim_amp, chunks, gstep);
        return;
    }
    if (b < head) return;                          /* padding */
    cyto_block<TF, TS, RYR_BLOCK>(a, xcd_remap(b - head, (int)gridDim.x - head), chunks, gstep, kc);
}

/* ------------------------------------------------------- output-row kernel */
/* Output row in two launches, so that no workgroup needs a fence (on this
 * GPU an agent-scope release writes back the L2):
 *   k_diag: one workgroup per 256-CRU chunk forms the pairwise double-double
 *     sums of its chunk, its integer counts and its linescan entries;
 *   k_diag_cell: one workgroup per cell adds the chunk sums in chunk order
 *     (the same order as the step reduction and the oracle) and writes the
 *     row (cuda_Cell_Detailed.cu:839-962, :1031-1072). */
#define RYR_DIAG_Q 21
#define RYR_DIAG_I 13
template <typename TF, typename TS>
__global__ __launch_bounds__(RYR_BLOCK) void k_diag(RyrKArgs a)
{
    __shared__ RyrDD s_red[RYR_DIAG_Q][RYR_BLOCK / 64];
    __shared__ unsigned long long s_cnt[RYR_DIAG_I];
    const int b = blockIdx.x;
    const int c = a.block_cell[b];
    const RyrCellInfo C = a.cells[c];
    RyrMem *M = a.mem + c;
    if (a.diag_step[c] != M->step - 1) return;   /* not an output step */
    const int par = (int)(M->step & 1);          /* buffer holding the new state */
    const long long P = a.P;
    const TF *ci = (const TF *)a.ci[par];
    const TS *cn = (const TS *)a.cnsr[par];
    const TF *cs = (const TF *)a.cs[par];
    const TF *catf = (const TF *)a.catf, *cacal = (const TF *)a.cacal, *casr = (const TF *)a.casr;
    const TF *camyo = (const TF *)a.camyo;
    const TS *cats = (const TS *)a.cats;
    const TF *cp = (const TF *)a.cp[par], *ancx = (const TF *)a.Ancx;
    const TF *casar = (const TF *)a.casar, *casarh = (const TF *)a.casarh;
    const TF *casarj = (const TF *)a.casarj, *casarhj = (const TF *)a.casarhj;
    const TS *cjsr = (const TS *)a.cjsr[par], *cjsrb = (const TS *)a.cjsrb;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x < RYR_DIAG_I) s_cnt[threadIdx.x] = 0;
    __syncthreads();

    unsigned long long ic[RYR_DIAG_I] = {0};  /* tnou tnob tncu tncb tn1..tn7 sparks runaway */
    const int n = (b - C.first_block) * RYR_BLOCK + (int)threadIdx.x;
    double val[RYR_DIAG_Q];
#pragma unroll
    for (int q = 0; q < RYR_DIAG_Q; ++q) val[q] = 0.0;
    if (n < C.N) {
        RYR_STRICT_FP
        const long long g = C.off + n;
        const uint32_t rw = a.ryrw[g], lw = a.lccw[g], ax = a.aux[g];
        const int nou = rw & 255, ncu = (rw >> 8) & 255, nob = (rw >> 16) & 255, ncb = rw >> 24;
        double xin = a.d_xinaca[g];
        val[0] = xin < 0 ? xin : 0.0;
        val[1] = (double)ancx[g];
        val[2] = (double)cp[g];
        val[3] = (double)cjsr[g];
        val[4] = (double)cjsrb[g];
        val[5] = (double)(nou + nob) / (double)(RYR_NRYR);
        val[6] = (double)ncb;
        val[7] = (double)casar[g];
        val[8] = (double)casarh[g];
        val[9] = (double)casarj[g];
        val[10] = (double)casarhj[g];
        double s_ci = 0, s_cn = 0, s_tf = 0, s_ts = 0, s_sr = 0, s_my = 0, s_cal = 0, s_lk = 0, s_up = 0;
        for (int ii = 0; ii < 8; ++ii) {
            const long long e = ii * P + g;
            s_ci = s_ci + (double)ci[e] / 8.;
            s_cn = s_cn + (double)cn[e] / 8.;
            s_tf = s_tf + (double)catf[e] / 8.;
            s_ts = s_ts + (double)cats[e] / 8.;
            s_sr = s_sr + (double)casr[e] / 8.;
            s_my = s_my + (double)camyo[e] / 8.;
            s_cal = s_cal + (double)cacal[e] / 8.;
            s_lk = s_lk + a.d_leak[e] / 8.;
            s_up = s_up + a.d_up[e] / 8.;
            if ((double)ci[e] > 1000) ic[12]++;
        }
        val[11] = s_ci; val[12] = s_cn; val[13] = s_tf; val[14] = s_ts; val[15] = s_sr;
        val[16] = s_my; val[17] = s_cal;
        val[18] = s_lk;
        val[19] = s_up;
        val[20] = a.d_xire[g];
        ic[0] += nou; ic[1] += nob; ic[2] += ncu; ic[3] += ncb;
        ic[11] += (ax >> 16) & 1u;
        for (int jj = 0; jj < 8; ++jj) {
            if (!((ax >> (RYR_AUX_MASK_SHIFT + jj)) & 1u)) continue;   /* absent slot */
            const int st = (lw >> (4 * jj)) & 15;
            /* histogram of cuda_Cell_Detailed.cu:904-923 */
            if (st == 0) { ic[10]++; ic[9]++; continue; }
            if (st & 1) ic[4]++;
            if (st & 2) ic[5]++;
            if (st & 4) ic[6]++;
            if (st & 8) ic[7]++;
            if ((st & 4) && !(st & 8)) ic[8]++;
            if (st == 1) ic[9]++;
        }
        /* linescan row at y = line_j, z = line_k (interior indices) */
        const int row0 = C.X * (C.line_j + C.Y * C.line_k);
        if (n >= row0 && n < row0 + C.X) {
            const int x = n - row0;
            double *o = a.line + a.line_off[c] + (long long)x * 22;
            o[0] = (double)(M->step - 1) * a.dt; o[1] = (double)(x + 1); o[2] = (double)ci[g];
            o[3] = (double)nob; o[4] = (double)nou; o[5] = (double)ncb;
            o[6] = (double)ncu; o[7] = (double)(ax & 0xffff); o[8] = (double)cs[g];
            o[9] = (double)cp[g]; o[10] = (double)casar[g]; o[11] = (double)casarh[g];
            o[12] = (double)camyo[g]; o[13] = (double)((const TF *)a.mgmyo)[g]; o[14] = (double)casr[g];
            o[15] = (double)cacal[g]; o[16] = (double)cats[g]; o[17] = (double)catf[g];
            o[18] = a.d_xire[g]; o[19] = (double)cjsr[g]; o[20] = (double)ancx[g]; o[21] = a.d_xinaca[g];
        }
    }
#pragma unroll
    for (int q = 0; q < RYR_DIAG_Q; ++q) {
        RyrDD t = {val[q], 0.0};
        t = wave_tree(t);
        if (lane == 0) s_red[q][wv] = t;
    }
    /* the counts: per-wave sums (each below 2^32), one LDS atomic per wave */
#pragma unroll
    for (int q = 0; q < RYR_DIAG_I; ++q) {
        const uint32_t w = wave_sum_u32((uint32_t)ic[q]);
        if (lane == 0 && w) atomicAdd(&s_cnt[q], (unsigned long long)w);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        RyrDD *pp = a.dpartial + (long long)b * RYR_DIAG_Q;
#pragma unroll
        for (int q = 0; q < RYR_DIAG_Q; ++q)
            pp[q] = dd_add(dd_add(s_red[q][0], s_red[q][1]), dd_add(s_red[q][2], s_red[q][3]));
        unsigned long long *cc = a.dcount + (long long)b * RYR_DIAG_I;
        for (int q = 0; q < RYR_DIAG_I; ++q) cc[q] = s_cnt[q];
    }
}

/* one workgroup per cell, after k_diag: chunk sums in chunk order, the row */
template <typename TF, typename TS>
__global__ __launch_bounds__(RYR_BLOCK) void k_diag_cell(RyrKArgs a)
{
    __shared__ RyrDD s_red[RYR_DIAG_Q][RYR_BLOCK / 64];
    __shared__ unsigned long long s_cnt[RYR_DIAG_I];
    const int c = blockIdx.x;
    const RyrCellInfo C = a.cells[c];
    RyrMem *M = a.mem + c;
    if (a.diag_step[c] != M->step - 1) return;   /* not an output step */
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x < RYR_DIAG_I) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    /* counts: each lane adds its chunks' (integers, any order) */
    {
        unsigned long long ic[RYR_DIAG_I] = {0};
        for (int bb = (int)threadIdx.x; bb < C.nblocks; bb += RYR_BLOCK) {
            const unsigned long long *cc = a.dcount + (long long)(C.first_block + bb) * RYR_DIAG_I;
#pragma unroll
            for (int q = 0; q < RYR_DIAG_I; ++q) ic[q] += cc[q];
        }
        /* a wave sum first: the lanes would otherwise queue on the same 13
         * LDS words */
#pragma unroll
        for (int q = 0; q < RYR_DIAG_I; ++q) {
            const unsigned long long w = wave_sum_u64(ic[q]);
            if (lane == 0 && w) atomicAdd(&s_cnt[q], w);
        }
    }
    /* chunk sums: pairwise trees over groups of 256 chunks, groups in order.
     * A group is four 64-chunk slices; wave w forms the slice trees of the
     * quantities q = w mod 4, and a slice past the cell's last chunk is zero
     * without a tree.  Which wave forms a tree does not change it. */
    RyrDD acc[RYR_DIAG_Q];
#pragma unroll
    for (int q = 0; q < RYR_DIAG_Q; ++q) { acc[q].hi = 0.0; acc[q].lo = 0.0; }
    for (int g0 = 0; g0 < C.nblocks; g0 += RYR_BLOCK) {
        for (int sl = 0; sl < RYR_BLOCK / 64; ++sl) {
            const int bb = g0 + sl * 64 + lane;
            const bool empty = g0 + sl * 64 >= C.nblocks;
            const RyrDD *pp = a.dpartial + (long long)(C.first_block + bb) * RYR_DIAG_Q;
            for (int q = wv; q < RYR_DIAG_Q; q += RYR_BLOCK / 64) {
                RyrDD t;
                t.hi = 0.0; t.lo = 0.0;
                if (!empty) {
                    if (bb < C.nblocks) t = pp[q];
                    t = wave_tree(t);
                }
                if (lane == 0) s_red[q][sl] = t;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int q = 0; q < RYR_DIAG_Q; ++q)
                acc[q] = dd_add(acc[q], dd_add(dd_add(s_red[q][0], s_red[q][1]), dd_add(s_red[q][2], s_red[q][3])));
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    {
        RYR_STRICT_FP
        unsigned long long cnt[RYR_DIAG_I];
        for (int q = 0; q < RYR_DIAG_I; ++q) cnt[q] = s_cnt[q];
        const double dN = C.dN;
        double *r = a.diag + (long long)c * 45;
        double ncxfwd = acc[0].hi / dN, outAncx = acc[1].hi / dN, cproxit = acc[2].hi / dN;
        double cjsrt = acc[3].hi / dN, csqnbt = acc[4].hi / dN, poto = acc[5].hi / dN;
        double pbcto = acc[6].hi / dN / 100.;
        double casart = acc[7].hi / dN, casarht = acc[8].hi / dN, casarjt = acc[9].hi / dN;
        double casarhjt = acc[10].hi / dN;
        double cit = acc[11].hi / dN, cnsrt = acc[12].hi / dN, catft = acc[13].hi / dN;
        double catst = acc[14].hi / dN, casrt = acc[15].hi / dN, camyot = acc[16].hi / dN;
        double cacalt = acc[17].hi / dN, leakt = acc[18].hi / dN, upt = acc[19].hi / dN;
        double ire = acc[20].hi / dN;
        double csubt = M->cur[0];
        double Total_Ca = (cit + catft + catst + casrt + camyot + cacalt) * (0.5 / 8.) * 8. +
                          (csubt + casart + casarht) * (0.025) +
                          (cproxit + casarjt + casarhjt) * (0.00126) +
                          (cnsrt) * (0.025 / 8.) * 8. +
                          (cjsrt + csqnbt) * (0.02 * 1);
        double Total_Ci = (cit + catft + catst + casrt + camyot + cacalt) * (0.5 / 8.) * 8. +
                          (csubt + casart + casarht) * (0.025) +
                          (cproxit + casarjt + casarhjt) * (0.00126);
        r[1] = cit; r[5] = cproxit; r[7] = cjsrt; r[8] = cnsrt; r[9] = poto;
        r[19] = ncxfwd * ((0.025) / (0.00126)); r[20] = pbcto; r[21] = cacalt; r[22] = catft;
        r[23] = leakt; r[24] = upt; r[25] = ire;
        r[26] = (double)cnt[0]; r[27] = (double)cnt[1]; r[28] = (double)cnt[2]; r[29] = (double)cnt[3];
        for (int q = 0; q < 7; ++q) r[30 + q] = (double)cnt[4 + q];
        r[37] = outAncx; r[40] = (double)cnt[11]; r[41] = catst; r[42] = csqnbt;
        r[43] = Total_Ci; r[44] = Total_Ca;
        if (cnt[12]) a.counters[(long long)c * RYR_NCNT + 3] += cnt[12];
        if (M->beat_on >= 0) {
            if (M->step - 1 == M->beat_on) M->b_jsr = cjsrt;
            else M->b_sparks += (double)cnt[11];
            if (cit > M->b_peak_ci) M->b_peak_ci = cit;
        }
    }
}

/* ---------------------------------------------------- set-up / data access */
/* per-cell initial values, computed on the host from Init
 * (cuda_Cell_Detailed.cu:1108-1182) */
struct RyrInit {
    double ci0, cjsr0, cjsrb0, catf0, cats0, cacal0, casr0, camyo0, mgmyo0;
    double casar0, casarh0, ancx0;
    int ncb0, ncu0;
    uint32_t lcc0;
    int pad_;
};

template <typename TF, typename TS>
__global__ __launch_bounds__(RYR_BLOCK) void k_init(RyrKArgs a, const RyrInit *__restrict__ init)
{
    const int b = blockIdx.x;
    const int c = a.block_cell[b];
    const RyrCellInfo C = a.cells[c];
    const long long g = (long long)b * RYR_BLOCK + threadIdx.x;
    const long long n = g - C.off;
    if (n >= C.N) return;
    const RyrInit I = init[c];
    const long long P = a.P;
    for (int buf = 0; buf < 2; ++buf) {
        TF *ci = (TF *)a.ci[buf];
        TS *cn = (TS *)a.cnsr[buf];
        for (int ii = 0; ii < 8; ++ii) { ci[ii * P + g] = (TF)I.ci0; cn[ii * P + g] = (TS)I.cjsr0; }
        ((TF *)a.cs[buf])[g] = (TF)I.ci0;
    }
    for (int ii = 0; ii < 8; ++ii) {
        const long long e = ii * P + g;
        ((TF *)a.catf)[e] = (TF)I.catf0;
        ((TS *)a.cats)[e] = (TS)I.cats0;
        ((TF *)a.cacal)[e] = (TF)I.cacal0;
        ((TF *)a.casr)[e] = (TF)I.casr0;
        ((TF *)a.camyo)[e] = (TF)I.camyo0;
        ((TF *)a.mgmyo)[e] = (TF)I.mgmyo0;
    }
    for (int buf = 0; buf < 2; ++buf) {
        ((TF *)a.cp[buf])[g] = (TF)I.ci0;
        ((TS *)a.cjsr[buf])[g] = (TS)I.cjsr0;
    }
    ((TS *)a.cjsrb)[g] = (TS)I.cjsrb0;
    ((TF *)a.Ancx)[g] = (TF)I.ancx0;
    ((TF *)a.casar)[g] = (TF)I.casar0;
    ((TF *)a.casarh)[g] = (TF)I.casarh0;
    ((TF *)a.casarj)[g] = (TF)I.casar0;
    ((TF *)a.casarhj)[g] = (TF)I.casarh0;
    int nlc = C.nlcc;
    if (C.variants & RYR_VAR_RANDOMLCC) {
        /* randpoint = curand % 9 (:1150-1152): one 32-bit draw of the CRU's
         * stream at counter step 2^64-1 */
        ryr_stream s0;
        ryr_stream_init(&s0, C.seed, (uint32_t)n, ~(uint64_t)0);
        nlc = (int)(ryr_stream_u32(&s0) % 9u);
    }
    uint32_t lw = 0;
    for (int ll = 0; ll < nlc; ++ll) lw |= 3u << (4 * ll);
    a.lccw[g] = lw;
    a.ryrw[g] = (uint32_t)0 | ((uint32_t)I.ncu0 << 8) | ((uint32_t)0 << 16) | ((uint32_t)I.ncb0 << 24);
    a.aux[g] = ((1u << nlc) - 1u) << RYR_AUX_MASK_SHIFT;
    a.d_xire[g] = 0.0; a.d_xinaca[g] = 0.0;
}

/* Non-template kernels: defined in exactly one unit (the API). */
#ifdef RYR_DEFINE_COMMON_KERNELS
__global__ void k_vfactors(RyrMem *mem, int ncells, int iso2)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncells) vfactors(mem + c, iso2);
}
/* step base of the launches that follow: set (add = 0) or advanced (add = 1) */
__global__ void k_stepbase(long long *gstep, long long v, int add)
{
    if (threadIdx.x == 0) *gstep = add ? *gstep + v : v;
}
#else
__global__ void k_vfactors(RyrMem *mem, int ncells, int iso2);
__global__ void k_stepbase(long long *gstep, long long v, int add);
#endif

/* field ids follow enum ryr_field of the public header */
template <typename TF, typename TS>
RYR_DEV double field_get(const RyrKArgs &a, int f, int cur, long long e, long long g, int p)
{
    switch (f) {
    case 0: return (double)((const TF *)a.ci[cur])[e];
    case 1: return (double)((const TS *)a.cnsr[cur])[e];
    case 2: return (double)((const TF *)a.catf)[e];
    case 3: return (double)((const TS *)a.cats)[e];
    case 4: return (double)((const TF *)a.cacal)[e];
    case 5: return (double)((const TF *)a.casr)[e];
    case 6: return (double)((const TF *)a.camyo)[e];
    case 7: return (double)((const TF *)a.mgmyo)[e];
    case 8: return (double)((const TF *)a.cs[cur])[g];
    case 9: return (double)((const TF *)a.cp[cur])[g];
    case 10: return (double)((const TS *)a.cjsr[cur])[g];
    case 11: return (double)((const TS *)a.cjsrb)[g];
    case 12: return (double)((const TF *)a.Ancx)[g];
    case 13: return (double)((const TF *)a.casar)[g];
    case 14: return (double)((const TF *)a.casarh)[g];
    case 15: return (double)((const TF *)a.casarj)[g];
    case 16: return (double)((const TF *)a.casarhj)[g];
    case 17: return (double)(a.ryrw[g] & 255u);
    case 18: return (double)((a.ryrw[g] >> 8) & 255u);
    case 19: return (double)((a.ryrw[g] >> 16) & 255u);
    case 20: return (double)(a.ryrw[g] >> 24);
    case 21: return (double)((a.lccw[g] >> (4 * p)) & 15u);
    case 22: return (double)(a.aux[g] & 0xffffu);
    case 23: return (double)((a.aux[g] >> 16) & 1u);
    }
    return 0.0;
}

template <typename TF, typename TS>
RYR_DEV void field_set(const RyrKArgs &a, int f, int cur, long long e, long long g, int p, double v)
{
    switch (f) {
    case 0: ((TF *)a.ci[cur])[e] = (TF)v; break;
    case 1: ((TS *)a.cnsr[cur])[e] = (TS)v; break;
    case 2: ((TF *)a.catf)[e] = (TF)v; break;
    case 3: ((TS *)a.cats)[e] = (TS)v; break;
    case 4: ((TF *)a.cacal)[e] = (TF)v; break;
    case 5: ((TF *)a.casr)[e] = (TF)v; break;
    case 6: ((TF *)a.camyo)[e] = (TF)v; break;
    case 7: ((TF *)a.mgmyo)[e] = (TF)v; break;
    case 8: ((TF *)a.cs[cur])[g] = (TF)v; break;
    case 9: ((TF *)a.cp[cur])[g] = (TF)v; break;
    case 10: ((TS *)a.cjsr[cur])[g] = (TS)v; break;
    case 11: ((TS *)a.cjsrb)[g] = (TS)v; break;
    case 12: ((TF *)a.Ancx)[g] = (TF)v; break;
    case 13: ((TF *)a.casar)[g] = (TF)v; break;
    case 14: ((TF *)a.casarh)[g] = (TF)v; break;
    case 15: ((TF *)a.casarj)[g] = (TF)v; break;
    case 16: ((TF *)a.casarhj)[g] = (TF)v; break;
    case 17: a.ryrw[g] = (a.ryrw[g] & ~255u) | ((uint32_t)v & 255u); break;
    case 18: a.ryrw[g] = (a.ryrw[g] & ~(255u << 8)) | (((uint32_t)v & 255u) << 8); break;
    case 19: a.ryrw[g] = (a.ryrw[g] & ~(255u << 16)) | (((uint32_t)v & 255u) << 16); break;
    case 20: a.ryrw[g] = (a.ryrw[g] & ~(255u << 24)) | (((uint32_t)v & 255u) << 24); break;
    case 21:   /* 16 = absent slot */
        if (v >= 16.0) a.aux[g] &= ~(1u << (RYR_AUX_MASK_SHIFT + p));
        else {
            a.aux[g] |= 1u << (RYR_AUX_MASK_SHIFT + p);
            a.lccw[g] = (a.lccw[g] & ~(15u << (4 * p))) | (((uint32_t)v & 15u) << (4 * p));
        }
        break;
    case 22: a.aux[g] = (a.aux[g] & ~0xffffu) | ((uint32_t)v & 0xffffu); break;
    case 23: a.aux[g] = (a.aux[g] & ~(1u << 16)) | (((uint32_t)v & 1u) << 16); break;
    }
}

/* planes*N elements, element e = p*N + n (one thread each) */
template <typename TF, typename TS>
__global__ void k_export(RyrKArgs a, int cell, int field, int planes, double *out)
{
    const RyrCellInfo C = a.cells[cell];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)planes * C.N) return;
    const int p = (int)(e / C.N);
    const long long n = e % C.N;
    const long long g = C.off + n;
    const int cur = (int)(a.mem[cell].step & 1);
    /* slots past the cell's channel count are absent: state 16 (:1157-1162) */
    const bool absent = field == 21 && !((a.aux[g] >> (RYR_AUX_MASK_SHIFT + p)) & 1u);
    out[e] = absent ? 16.0 : field_get<TF, TS>(a, field, cur, (long long)p * a.P + g, g, p);
}

/* planes*N elements, element e = p*N + n (one thread each).  The LCC field
 * packs the 8 slots of a CRU into the words lccw[g] and aux[g], so it is
 * written by one thread per CRU that assembles both words from its 8 inputs
 * (threads of different slots would otherwise read-modify-write the same
 * words). */
template <typename TF, typename TS>
__global__ void k_import(RyrKArgs a, int cell, int field, int planes, const double *in)
{
    const RyrCellInfo C = a.cells[cell];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (field == 21) {
        if (e >= C.N) return;
        const long long g = C.off + e;
        uint32_t w = a.lccw[g], mask = 0u;
        for (int p = 0; p < 8; ++p) {
            const double v = in[(long long)p * C.N + e];
            if (v < 16.0) {
                mask |= 1u << p;
                w = (w & ~(15u << (4 * p))) | (((uint32_t)v & 15u) << (4 * p));
            }
        }
        a.lccw[g] = w;
        a.aux[g] = (a.aux[g] & ~(255u << RYR_AUX_MASK_SHIFT)) | (mask << RYR_AUX_MASK_SHIFT);
        return;
    }
    if (e >= (long long)planes * C.N) return;
    const int p = (int)(e / C.N);
    const long long n = e % C.N;
    const long long g = C.off + n;
    const int cur = (int)(a.mem[cell].step & 1);
    field_set<TF, TS>(a, field, cur, (long long)p * a.P + g, g, p, in[e]);
}

/* frozen-input gating probe: LCC loop (:1248-1261) then ryrgating */
#ifdef RYR_DEFINE_COMMON_KERNELS
__global__ void k_gating_probe(long long nq, unsigned long long seed, const uint32_t *cru,
                               unsigned long long step, double Ku, double Ku2, double Kb, double dt,
                               int iso2, int32_t *counts, int32_t *lcc, const double *conc,
                               int32_t *draws, int32_t *nopen, int32_t *fails)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const double cp = conc[4 * q], cjsr = conc[4 * q + 1], cs = conc[4 * q + 2], v = conc[4 * q + 3];
    RyrMem m;
    m.m[0] = v;
    m.m[12] = 10.0;
    vfactors(&m, iso2);
    int no = 0, nlc = 0;
    {
        RYR_STRICT_FP
        const double cpx = (cp + cs * (0.5)) / (1 + (0.5));
        const double betac = iso2 ? 2.0 : 4.0;
        ryr_u32x4 lb = {{0u, 0u, 0u, 0u}};
        for (int ll = 0; ll < 8 && lcc[8 * q + ll] < 16; ll++) {
            if (!(ll & 1)) lb = ryr_block(seed, cru[q], step, (uint32_t)(ll >> 1));
            const double ragg = (ll & 1) ? ryr_slot_uniform(lb.v[2], lb.v[3]) : ryr_slot_uniform(lb.v[0], lb.v[1]);
            int jj = lcc_step(lcc[8 * q + ll], cpx, dt, ragg, m, betac);
            lcc[8 * q + ll] = jj;
            if (jj == 0) ++no;
            ++nlc;
        }
    }
    int nou = counts[4 * q], ncu = counts[4 * q + 1], nob = counts[4 * q + 2], ncb = counts[4 * q + 3];
    int f = 0;
    double roo2;
    const double mono = csqn_mono(cjsr, &roo2);
    int nd = 0;
    ryr_gating(seed, cru[q], step, Ku, Ku2, Kb, cp, cjsr, mono, dt, nou, ncu, nob, ncb, &f, &nd);
    counts[4 * q] = nou; counts[4 * q + 1] = ncu; counts[4 * q + 2] = nob; counts[4 * q + 3] = ncb;
    draws[q] = (int32_t)(nlc + nd);
    nopen[q] = no;
    fails[q] = f;
}
#endif

#endif /* RYR_CRU_DEVICE_H */
