/*
 * ryr_oracle.c - FP64 CPU restatement of the reference's per-step CRU update,
 * whole-cell reduction, membrane ODE and output-row diagnostics.
 *
 * TEST INFRASTRUCTURE: only tests/, the smoke check and the benchmark's CPU
 * baseline use this file, as the reference results are compared with.
 * Nothing in the product links it.
 *
 * Reference files (colinmrees/RyR_hyperactivity):
 *   cuda_Cell_Detailed.cu  - constants :29-145, host ICs :350-366,
 *                            time loop :399-1085, Init :1108-1182,
 *                            Compute :1187-1692
 *   subroutine_cuda.c      - ncx, uptake, leak, luminal(_SS), release,
 *                            lcccurrent, ryrgating
 *   buffers.c              - buffer fluxes
 *   reesica.h              - lcc_markov
 * The layout is the reference's: a lattice of (X+2)(Y+2)(Z+2) CRUs with a
 * ghost shell, pos(i,j,k) = i + nx*(j + ny*k), sub-voxel arrays at
 * pos*8+ii, double-buffered ci/cnsr/cs, boundary copies into the ghosts.
 * Every arithmetic expression keeps the reference's operand order.
 *
 * Deliberate, documented departures (all shared with the GPU build):
 *  - random numbers: Philox4x32-10 streams keyed (seed, CRU, step) instead of
 *    cuRAND XORWOW (:1700); uniforms on (0,1) instead of (0,1];
 *  - exp/log/cos/pow: the deterministic detmath routines (a few ulp from libm);
 *  - time: t = n*dt from an integer step counter; the stimulus window, the
 *    output step and the per-interval counter reset are integer tests on n
 *    (the reference's fmod on an accumulated t can jitter by one step);
 *  - whole-cell and diagnostic sums: a fixed compensated order (pairwise
 *    double-double trees inside chunks of 256 CRUs, then over groups of 256
 *    chunk sums, groups accumulated in order) instead of the serial loop of
 *    :435-445/:839-929;
 *  - the dead dye buffer (cadye, never printed) is not carried;
 *  - Cm follows ryr_params.cm_mode (density-preserving by default).
 *
 * Pinning: the tests check each part against closed forms, invariants,
 * brute force or published known answers (Philox), and the constants and
 * leaf formulas against the reference's own literals (a fixture taken from
 * its sources): every define, the literal sequence of each leaf function,
 * the LCC gate rates, NCX, SERCA, leak and the membrane currents and gates.
 * Scheme B (:1475-1597) conserves no quantity; it is pinned by its exact
 * calcium balance, the sarcolemmal flux plus the lag its mixed operands
 * leave (orc_read_scheme_b_lag).  Parity unpinned: trajectories against the
 * reference's own runs (cuRAND and CUDA libm are absent and the reference
 * ships no outputs), so agreement with it is statistical.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "ryr_oracle.h"
#include "../ryr_hyperactivity_amd/csrc/common/philox.h"
#include "../ryr_hyperactivity_amd/csrc/common/detmath.h"
#include "../ryr_hyperactivity_amd/csrc/common/srround.h"
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants: cuda_Cell_Detailed.cu:63-145 ---------------------------- */
#define O_PI        3.1415926
#define O_VP        (0.00126)
#define O_VJSR      (0.02*1)
#define O_VI        (0.5/8.)
#define O_VNSR      (0.025/8.)
#define O_NRYR      100
#define O_TAUPS     (0.0283)
#define O_TAUPI     (0.1)
#define O_TAUSI     (0.04)
#define O_TAUTR     (25.0*1.0/4.)
#define O_TAUNL     6.
#define O_TAUNT     1.8
#define O_TAUIL     (0.7*2.)
#define O_TAUIT     (0.33*2.)
#define O_TAUST     1.42
#define O_FRT       (96.485/8.314/308.0)
#define O_XI        0.7
#define O_VS        (0.025)
#define O_ICAGAMMA  (0.5)
#define O_TAUCB     1.
#define O_TAUCU     1.
#define O_SV_IUP    (1.0+0.75)
#define O_SVJMAX    11.5
#define O_TAUU      (0.2*5*220.)
#define O_TAUB      (1.0)
#define O_SVILEAK   1.0
#define O_SVTAUXS   0.95
#define O_XNAO      140.0
/* subroutine_cuda.c:7-38 */
#define O_VNACA     (21.0/1.0)
#define O_KMCAI     0.00359
#define O_KMCAO     1.3
#define O_KMNAI     12.3
#define O_KMNAO     87.5
#define O_KSAT      0.27
#define O_ETA       0.35
#define O_FARAD     96.485
#define O_XR        8.314
#define O_TEMPER    308
#define O_CEXT      (1*1.8)
#define O_NM        20
#define O_ND        30
#define O_RATEDIMER 5000.0
#define O_KDIMER    850.0
#define O_HILLDIMER 23.0
#define O_CSQBERS   460.0
#define O_KBERS     600.0
#define O_PCA       11.9
#define O_GAMMAI    0.341
/* buffers.c:16-52 */
#define O_KTSON     0.00254
#define O_KTSOFF    0.000033
#define O_BTS       ((134.))
#define O_KTFON     0.0327
#define O_KTFOFF    0.0196
#define O_BTF       70.
#define O_KCALON    0.0543
#define O_KCALOFF   0.238
#define O_BCAL      24.
#define O_KSRON     0.1
#define O_KSROFF    0.06
#define O_BSR       19.
#define O_KSARON    0.1
#define O_KSAROFF   1.3
#define O_BSAR      (42*(O_VI*8.0/O_VS))
#define O_KSARHON   0.1
#define O_KSARHOFF  0.03
#define O_BSARH     (15.*(O_VI*8.0/O_VS))
#define O_BMYO      140
#define O_KONMYOMG  0.0157
#define O_KOFFMYOMG 0.057
#define O_KONMYOCA  13.8
#define O_KOFFMYOCA 0.46
#define O_MGI       500

#define P2(x) ((x)*(x))
#define P3(x) ((x)*(x)*(x))

/* ---- compensated canonical sum ------------------------------------------ */
typedef struct { double hi, lo; } orc_dd;

static orc_dd dd_add(orc_dd a, orc_dd b)
{
    double s = a.hi + b.hi;
    double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b.hi - bb);
    e = e + (a.lo + b.lo);
    orc_dd r;
    r.hi = s + e;
    r.lo = e - (r.hi - s);
    return r;
}

static orc_dd dd_tree256(orc_dd *v)
{
    for (int w = 1; w < 256; w *= 2)
        for (int l = 0; l < 256; l += 2 * w)
            v[l] = dd_add(v[l], v[l + w]);
    return v[0];
}

/* Two-level pairwise double-double tree: values in chunks of 256
 * consecutive entries (zero-padded) give chunk sums; chunk sums in groups of
 * 256 (zero-padded) give group sums; group sums are added in order. */
double orc_canonical_sum(const double *x, int64_t n)
{
    orc_dd acc = {0.0, 0.0};
    orc_dd v[256], grp[256];
    int gl = 0;
    for (int64_t c0 = 0; c0 < n; c0 += 256) {
        for (int l = 0; l < 256; ++l) {
            v[l].hi = (c0 + l < n) ? x[c0 + l] : 0.0;
            v[l].lo = 0.0;
        }
        grp[gl++] = dd_tree256(v);
        if (gl == 256 || c0 + 256 >= n) {
            for (int l = gl; l < 256; ++l) { grp[l].hi = 0.0; grp[l].lo = 0.0; }
            acc = dd_add(acc, dd_tree256(grp));
            gl = 0;
        }
    }
    return acc.hi;
}

/* ---- leaf physics: subroutine_cuda.c ------------------------------------- */

/* subroutine_cuda.c:47-75 (cs in mM).  Under slowncxchange (:59-62) the
 * exchanger runs at five times the allosteric factor; the factor itself
 * relaxes as without it (the dead first assignment Ka = 0.025*12 is left out) */
static double orc_ncx_v(double cs, double v, double xnai, double *Ancx, double dt, int slowncx)
{
    double za = v * O_FARAD / O_XR / O_TEMPER;
    double Ka = *Ancx;
    if (slowncx) Ka = *Ancx * 5;
    double t1 = O_KMCAI * P3(O_XNAO) * (1.0 + P3(xnai / O_KMNAI));
    double t2 = P3(O_KMNAO) * cs * (1.0 + cs / O_KMCAI);
    double t3 = (O_KMCAO + O_CEXT) * P3(xnai) + cs * P3(O_XNAO);
    double Inaca = Ka * O_VNACA * (ryr_exp(O_ETA * za) * P3(xnai) * O_CEXT
                                   - ryr_exp((O_ETA - 1.0) * za) * P3(O_XNAO) * cs)
                   / ((t1 + t2 + t3) * (1.0 + O_KSAT * ryr_exp((O_ETA - 1.0) * za)));
    double Ancxdot = (1.0 / (1.0 + P3(0.0003 / cs)) - *Ancx) / 150.;
    *Ancx += Ancxdot * dt;
    return Inaca;
}

double orc_ncx(double cs, double v, double xnai, double *Ancx, double dt)
{
    return orc_ncx_v(cs, v, xnai, Ancx, dt, 0);
}

double orc_ncx_slow(double cs, double v, double xnai, double *Ancx, double dt)
{
    return orc_ncx_v(cs, v, xnai, Ancx, dt, 1);
}

/* subroutine_cuda.c:128-137 */
double orc_uptake(double ci, double cnsr)
{
    double vup = 0.3, Ki = 0.123, Knsr = 1700.0, HH = 1.787, upfactor = 1.0;
    return upfactor * vup * (ryr_pow(ci / Ki, HH) - ryr_pow(cnsr / Knsr, HH))
           / (1.0 + ryr_pow(ci / Ki, HH) + ryr_pow(cnsr / Knsr, HH));
}

/* subroutine_cuda.c:139-146 */
double orc_leak(double cnsr, double ci)
{
    double gleak = 0.00001035, Kjsr = 500.0;
    return gleak * (cnsr - ci) * P2(cnsr) / (P2(cnsr) + P2(Kjsr));
}

/* subroutine_cuda.c:161-177 */
double orc_luminal(double cjsr)
{
    double roo2 = O_RATEDIMER / (1.0 + ryr_pow(O_KDIMER / cjsr, O_HILLDIMER));
    double mono = (-1.0 + sqrt(1.0 + 8.0 * O_CSQBERS * roo2)) / (4.0 * roo2 * O_CSQBERS);
    double ene = mono * O_NM + (1.0 - mono) * O_ND;
    double enedot = (O_NM - O_ND) * (-mono + 1.0 / (4.0 * O_CSQBERS * mono * roo2 + 1.0))
                    * (O_HILLDIMER / cjsr) * (1.0 - roo2 / O_RATEDIMER);
    return 1.0 / (1.0 + (O_CSQBERS * O_KBERS * ene + O_CSQBERS * enedot * cjsr * (cjsr + O_KBERS))
                        / P2((O_KBERS + cjsr)));
}

/* subroutine_cuda.c:179-195 */
double orc_luminal_SS(double cjsr)
{
    double roo2 = O_RATEDIMER / (1.0 + ryr_pow(O_KDIMER / cjsr, O_HILLDIMER));
    double mono = (-1.0 + sqrt(1.0 + 8.0 * O_CSQBERS * roo2)) / (4.0 * roo2 * O_CSQBERS);
    double ene = mono * O_NM + (1.0 - mono) * O_ND;
    return O_CSQBERS * ene * cjsr / (O_KBERS + cjsr);
}

/* subroutine_cuda.c:222-228 */
double orc_release(double po, double cjsr, double cp)
{
    double Jmax = (0.0147 * O_SVJMAX);
    return 1.0 * Jmax * po * (cjsr - cp) / O_VP;
}

/* subroutine_cuda.c:231-247 */
double orc_lcccurrent(double v, double cp)
{
    double za = v * O_FARAD / O_XR / O_TEMPER;
    double ica;
    if (fabs(za) < 0.001)
        ica = 2.0 * O_PCA * O_FARAD * O_GAMMAI * (cp / 1000.0 * ryr_exp(2.0 * za) - O_CEXT);
    else
        ica = 4.0 * O_PCA * za * O_FARAD * O_GAMMAI * (cp / 1000.0 * ryr_exp(2.0 * za) - O_CEXT)
              / (ryr_exp(2.0 * za) - 1.0);
    if (ica > 0.0) ica = 0.0;
    return ica;
}

/* ---- buffers.c:55-155 ----------------------------------------------------- */
static double b_myoca(double CaMyo, double MgMyo, double c, double dt)
{
    if (O_KONMYOCA * c * dt > 1)
        return 0.95 / dt * (O_BMYO - CaMyo - MgMyo) - O_KOFFMYOCA * CaMyo;
    return O_KONMYOCA * c * (O_BMYO - CaMyo - MgMyo) - O_KOFFMYOCA * CaMyo;
}
static double b_myomg(double CaMyo, double MgMyo, double c, double dt)
{
    (void)c;
    if (O_KONMYOMG * O_MGI * dt > 1)
        return 0.95 / dt * (O_BMYO - CaMyo - MgMyo) - O_KOFFMYOMG * MgMyo;
    return O_KONMYOMG * O_MGI * (O_BMYO - CaMyo - MgMyo) - O_KOFFMYOMG * MgMyo;
}
#define ORC_SIMPLE_BUFFER(name, kon, koff, B)                                   \
    static double name(double b, double c, double dt)                            \
    {                                                                            \
        if (kon * c * dt > 1) return 0.95 / dt * (B - b) - koff * b;             \
        return kon * c * (B - b) - koff * b;                                     \
    }
ORC_SIMPLE_BUFFER(b_tropf, O_KTFON, O_KTFOFF, O_BTF)      /* :72-79  */
ORC_SIMPLE_BUFFER(b_trops, O_KTSON, O_KTSOFF, O_BTS)      /* :81-88  */
ORC_SIMPLE_BUFFER(b_bucal, O_KCALON, O_KCALOFF, O_BCAL)   /* :90-97  */
ORC_SIMPLE_BUFFER(b_busr, O_KSRON, O_KSROFF, O_BSR)       /* :108-115 */
ORC_SIMPLE_BUFFER(b_busar, O_KSARON, O_KSAROFF, O_BSAR)   /* :118-125 */
ORC_SIMPLE_BUFFER(b_busarh, O_KSARHON, O_KSARHOFF, O_BSARH) /* :128-135 */

/* buffers.c:137-155 (ba = 1) */
static double b_busarj(double CaSar, double c, double dt)
{
    if (O_KSARON * 1 * c * dt > 1)
        return 0.95 / dt * (O_BSAR - CaSar) - O_KSAROFF * 1 * CaSar;
    return O_KSARON * c * 1 * (O_BSAR - CaSar) - O_KSAROFF * 1 * CaSar;
}
static double b_busarhj(double CaSarh, double c, double dt)
{
    if (O_KSARHON * 1 * c * dt > 1)
        return 0.95 / dt * (O_BSARH - CaSarh) - O_KSARHOFF * 1 * CaSarh;
    return O_KSARHON * 1 * c * (O_BSARH - CaSarh) - O_KSARHOFF * 1 * CaSarh;
}

double orc_buffer_flux(int which, double b, double b2, double c, double dt)
{
    switch (which) {
    case 0: return b_tropf(b, c, dt);
    case 1: return b_trops(b, c, dt);
    case 2: return b_bucal(b, c, dt);
    case 3: return b_busr(b, c, dt);
    case 4: return b_myoca(b, b2, c, dt);
    case 5: return b_myomg(b, b2, c, dt);
    case 6: return b_busar(b, c, dt);
    case 7: return b_busarh(b, c, dt);
    case 8: return b_busarj(b, c, dt);
    case 9: return b_busarhj(b, c, dt);
    }
    return 0.0;
}

/* Init's bound fraction (cuda_Cell_Detailed.cu:1165-1169) */
double orc_fracbound(double cjsr_b)
{
    double roo2 = O_RATEDIMER / (1.0 + ryr_pow(O_KDIMER / (cjsr_b), O_HILLDIMER));
    double aub = (-1.0 + sqrt(1.0 + 8.0 * O_CSQBERS * roo2)) / (4.0 * roo2 * O_CSQBERS) / O_TAUB;
    double abu = 1.0 / O_TAUU;
    return 1 / (1 + abu / aub);
}

/* ---- random draws -------------------------------------------------------- */
typedef struct orc_rng {
    ryr_stream ph;
    const uint64_t *replay;   /* raw slots (a<<32 | b) instead of Philox */
    int nreplay;
    int used;
} orc_rng;

static void orc_slot(orc_rng *g, uint32_t *a, uint32_t *b)
{
    if (g->replay) {
        uint64_t s = (g->used < g->nreplay) ? g->replay[g->used] : 0x8000000000000000ull;
        *a = (uint32_t)(s >> 32);
        *b = (uint32_t)s;
    } else {
        ryr_stream_next_slot(&g->ph, a, b);
    }
    g->used++;
}
static double orc_u(orc_rng *g) { uint32_t a, b; orc_slot(g, &a, &b); return ryr_slot_uniform(a, b); }

/* Draw d (0, 1, ...) of RyR transition t under the addressing of philox.h:
 * d = 0 is static slot 8 + t, d >= 1 is slot d - 1 of the transition's own
 * substream.  A replay hands out its slots in consumption order instead. */
static void orc_tslot(orc_rng *g, int t, uint32_t d, uint32_t *a, uint32_t *b)
{
    if (g->replay) {
        orc_slot(g, a, b);
        return;
    }
    const uint64_t seed = (uint64_t)g->ph.k0 | ((uint64_t)g->ph.k1 << 32);
    const uint64_t step = (uint64_t)g->ph.step_lo | ((uint64_t)g->ph.step_hi << 32);
    uint32_t c0, half;
    if (d == 0) { c0 = (uint32_t)(RYR_SLOT_RYR0 + t) >> 1; half = (uint32_t)(RYR_SLOT_RYR0 + t) & 1u; }
    else { c0 = ryr_sub_block(t, d - 1); half = (d - 1) & 1u; }
    ryr_u32x4 o = ryr_block(seed, g->ph.cru, step, c0);
    *a = o.v[2 * half];
    *b = o.v[2 * half + 1];
    g->used++;
}
static double orc_tu(orc_rng *g, int t, uint32_t d) { uint32_t a, b; orc_tslot(g, t, d, &a, &b); return ryr_slot_uniform(a, b); }
static uint32_t orc_ti(orc_rng *g, int t, uint32_t d) { uint32_t a, b; orc_tslot(g, t, d, &a, &b); return b; }

/* ---- reesica.h:14-116, with the uniform supplied ------------------------ */
int orc_lcc_markov(int i, double cpx, double v, double dt, double ragg, int iso2)
{
    double dv5 = 5, dvk = 8, fv5 = -22.8, fvk = 9.1;
    double alphac = 0.22, betac = 4;
    if (iso2) { betac = 2; dv5 = 0; fv5 = -28; fvk = 8.5; }
    double dinf = 1.0 / (1.0 + ryr_exp(-(v - dv5) / dvk));
    double taudin = 1.0 / ((1.0 - ryr_exp(-(v - dv5) / dvk)) / (0.035 * (v - dv5)) * dinf);
    if ((v < 0.0001) && (v > -0.0001))
        taudin = 0.035 * dvk / dinf;
    double finf = 1. - 1.0 / (1.0 + ryr_exp(-(v - fv5) / fvk)) / (1. + ryr_exp((v - 60) / 12.));
    double taufin = (0.02 - 0.007 * ryr_exp(-P2(0.0337 * (v + 10.5))));
    double alphad = dinf * taudin;
    double betad = (1 - dinf) * taudin;
    double alphaf = (finf) * taufin;
    double betaf = (1 - finf) * taufin;
    double alphafca = 0.012 / 2.;
    double betafca = 0.175 / (1 + P2(25 / cpx));
    double rig = ragg / dt;

    if ((i % 2)) { if (rig < alphac) return i - 1; else rig -= alphac; }
    else         { if (rig < betac) return i + 1; else rig -= betac; }
    if (((i / 2) % 2)) { if (rig < alphad) return i - 2; else rig -= alphad; }
    else               { if (rig < betad) return i + 2; else rig -= betad; }
    if (((i / 4) % 2)) { if (rig < alphaf) return i - 4; else rig -= alphaf; }
    else               { if (rig < betaf) return i + 4; else rig -= betaf; }
    if (((i / 8) % 2)) { if (rig < alphafca) return i - 8; else rig -= alphafca; }
    else               { if (rig < betafca) return i + 8; else rig -= betafca; }
    return i;
}

/* ---- subroutine_cuda.c:251-669 ------------------------------------------- */

/* Transition t of the eight population transitions: Knuth's
 * product-of-uniforms Poisson sampler (:336-347) or, outside its validity
 * range, a dithered Gaussian with rejection of negative values (:348-363).
 * Its uniforms come from its own draw sequence (orc_tslot). */
static int orc_transition(orc_rng *g, int t, int n, double p, int *fails)
{
    double lam = ryr_exp(-(n) * p);
    int out = -1;
    uint32_t d = 0;
    if (n <= 1 || p < 0.2 || (n <= 5 && p < 0.3)) {
        int kk = 0;
        double puu = 1.0;
        while (puu >= lam && kk < 195) {
            kk++;
            double re = orc_tu(g, t, d++);
            puu = puu * (re);
        }
        out = kk - 1;
    } else {
        int kk = 0;
        while (out < 0) {
            double u1 = orc_tu(g, t, d++);
            double u2 = orc_tu(g, t, d++);
            uint32_t r = orc_ti(g, t, d++);
            out = (int)(floor(n * p + sqrt(n * p * (1.0 - p)) * sqrt(-2.0 * ryr_log(1.0 - u1))
                              * ryr_cos(2.0 * O_PI * u2)) + r % 2);
            kk++;
            if (kk > 200) { out = 0; (*fails)++; }
        }
    }
    return out;
}

static void orc_ryrgating(orc_rng *g, double Ku, double Ku2, double Kb, double cp, double cjsr,
                          int *ncu, int *nou, int *ncb, int *nob, double dt, int *fails)
{
    double ryrpref = 4.5 * 1.5;
    double cygatek = Ku;
    double ryrhill = 2;
    double pku = ryrpref * Ku2 * 1 / (1 + P2(5000 / cjsr)) * 1 / (1 + ryr_pow(cygatek / cp, ryrhill));
    double pkb = ryrpref / 36. * Kb * 1 / (1 + P2(5000 / cjsr)) * 1 / (1 + ryr_pow(cygatek / cp, ryrhill));
    double pkuminus = 1. / O_TAUCU;
    double pkbminus = 1. / O_TAUCB;
    if (pku * dt > 1.0) pku = 1.0 / dt;
    if (pkb * dt > 1.0) pkb = 1.0 / dt;
    if (pkuminus * dt > 1.0) pkuminus = 1.0 / dt;
    if (pkbminus * dt > 1.0) pkbminus = 1.0 / dt;

    double roo2 = O_RATEDIMER / (1.0 + ryr_pow(O_KDIMER / (cjsr), O_HILLDIMER));
    double aub = (-1.0 + sqrt(1.0 + 8.0 * O_CSQBERS * roo2)) / (4.0 * roo2 * O_CSQBERS) / O_TAUB;
    double abu = 1.0 / O_TAUU;

    /* OU -> CU (:332-364) */
    double pkuh = pkuminus * dt;
    int n_ou_cu = orc_transition(g, 0, *nou, pkuh, fails);
    if (n_ou_cu > O_NRYR) n_ou_cu = O_NRYR;
    /* CU -> OU (:366-397) */
    double pkum = pku * dt;
    int n_cu_ou = orc_transition(g, 1, *ncu, pkum, fails);
    if (n_cu_ou > O_NRYR) n_cu_ou = O_NRYR;
    /* OU -> OB (:401-436) */
    double pau = (pkb < 1e-16) ? 0 : aub * dt;
    int n_ou_ob = orc_transition(g, 2, *nou, pau, fails);
    if (n_ou_ob > O_NRYR) n_ou_ob = O_NRYR;
    /* OB -> OU (:439-476) */
    double pbu = (pkb < 1e-16) ? abu * dt * 36 : abu * dt * (pku / pkb);
    int n_ob_ou = orc_transition(g, 3, *nob, pbu, fails);
    if (n_ob_ou > O_NRYR) n_ob_ou = O_NRYR;
    /* CB -> CU (:480-511) */
    double pcb = abu * dt;
    int n_cb_cu = orc_transition(g, 4, *ncb, pcb, fails);
    if (n_cb_cu > O_NRYR) n_cb_cu = O_NRYR;
    /* CU -> CB (:513-544) */
    double pcub = aub * dt;
    int n_cu_cb = orc_transition(g, 5, *ncu, pcub, fails);
    if (n_cu_cb > O_NRYR) n_cu_cb = O_NRYR;
    /* OB -> CB (:548-579), no cap */
    double pkbh = pkbminus * dt;
    int n_ob_cb = orc_transition(g, 6, *nob, pkbh, fails);
    /* CB -> OB (:581-613) */
    double pkbm = pkb * dt;
    int n_cb_ob = orc_transition(g, 7, *ncb, pkbm, fails);
    if (n_cb_ob > O_NRYR) n_cb_ob = O_NRYR;

    /* outflow conflicts (:617-639) */
    if (n_ou_ob + n_ou_cu > *nou) {
        if (n_ou_cu >= n_ou_ob) n_ou_cu = 0; else n_ou_ob = 0;
        if (n_ou_ob > *nou) n_ou_ob = 0;
        else if (n_ou_cu > *nou) n_ou_cu = 0;
    }
    if (n_ob_ou + n_ob_cb > *nob) {
        if (n_ob_ou >= n_ob_cb) n_ob_ou = 0; else n_ob_cb = 0;
        if (n_ob_cb > *nob) n_ob_cb = 0;
        else if (n_ob_ou > *nob) n_ob_ou = 0;
    }
    if (n_cu_ou + n_cu_cb > *ncu) {
        if (n_cu_cb >= n_cu_ou) n_cu_cb = 0; else n_cu_ou = 0;
        if (n_cu_ou > *ncu) n_cu_ou = 0;
        else if (n_cu_cb > *ncu) n_cu_cb = 0;
    }
    /* population update and clamps (:642-665) */
    *nou += -n_ou_ob - n_ou_cu + n_ob_ou + n_cu_ou;
    if (*nou < 0) *nou = 0;
    if (*nou > O_NRYR) *nou = O_NRYR;
    *nob += -n_ob_ou - n_ob_cb + n_ou_ob + n_cb_ob;
    if (*nob < 0) *nob = 0;
    if (*nob > O_NRYR) *nob = O_NRYR;
    *ncu += -n_cu_ou - n_cu_cb + n_ou_cu + n_cb_cu;
    if (*ncu < 0) *ncu = 0;
    if (*ncu > O_NRYR) *ncu = O_NRYR;
    *ncb = O_NRYR - *nou - *nob - *ncu;
    if (*ncb < 0) *ncb = 0;
    if (*ncb > O_NRYR) *ncb = O_NRYR;
}

int orc_ryrgating_replay(const uint64_t *slots, int nslots, double Ku, double Ku2, double Kb,
                         double cp, double cjsr, double dt, int32_t *counts,
                         int32_t *used, int32_t *fails)
{
    orc_rng g;
    memset(&g, 0, sizeof g);
    g.replay = slots;
    g.nreplay = nslots;
    int nou = counts[0], ncu = counts[1], nob = counts[2], ncb = counts[3], f = 0;
    orc_ryrgating(&g, Ku, Ku2, Kb, cp, cjsr, &ncu, &nou, &ncb, &nob, dt, &f);
    counts[0] = nou; counts[1] = ncu; counts[2] = nob; counts[3] = ncb;
    *used = g.used;
    *fails = f;
    return 0;
}

int orc_gating_probe(int64_t n, uint64_t seed, const uint32_t *cru, uint64_t step,
                     double Ku, double Ku2, double Kb, double dt, int iso2,
                     int32_t *counts, int32_t *lcc, const double *conc,
                     int32_t *draws, int32_t *nopen, int32_t *fails)
{
    for (int64_t q = 0; q < n; ++q) {
        orc_rng g;
        memset(&g, 0, sizeof g);
        ryr_stream_init(&g.ph, seed, cru[q], step);
        double cp = conc[4 * q], cjsr = conc[4 * q + 1], cs = conc[4 * q + 2], v = conc[4 * q + 3];
        double cpx = (cp + cs * O_ICAGAMMA) / (1 + O_ICAGAMMA);
        int no = 0;
        for (int ll = 0; ll < 8 && lcc[8 * q + ll] < 16; ll++) {
            int jj = orc_lcc_markov(lcc[8 * q + ll], cpx, v, dt, orc_u(&g), iso2);
            lcc[8 * q + ll] = jj;
            if (jj == 0) ++no;
        }
        int nou = counts[4 * q], ncu = counts[4 * q + 1], nob = counts[4 * q + 2], ncb = counts[4 * q + 3];
        int f = 0;
        orc_ryrgating(&g, Ku, Ku2, Kb, cp, cjsr, &ncu, &nou, &ncb, &nob, dt, &f);
        counts[4 * q] = nou; counts[4 * q + 1] = ncu; counts[4 * q + 2] = nob; counts[4 * q + 3] = ncb;
        draws[q] = g.used;
        nopen[q] = no;
        fails[q] = f;
    }
    return 0;
}

/* ---- cell state ---------------------------------------------------------- */
struct orc_cell {
    ryr_params p;
    ryr_cell_desc d;
    int X, Y, Z, nx, ny, nz;
    int64_t ncell;          /* nx*ny*nz incl. ghosts */
    int64_t N;              /* interior CRUs */
    double Cm, svica, svncx, svikr, sviup, sviup_lqt;
    /* per CRU (ghost-inclusive) */
    double *cp, *cjsr, *cjsrb, *xire, *po, *Ancx, *xinaca, *xicat;
    double *casar, *casarh, *casarj, *casarhj;
    double *cs, *csnext;
    double *lagB;           /* scheme B: sum over the sub-steps of (cs at step start - cs at sub-step start) */
    int *lcc, *nl, *nou, *ncu, *nob, *ncb, *nspark;
    /* per sub-voxel */
    double *ci, *cinext, *cnsr, *cnsrnext;
    double *catf, *cats, *cacal, *casr, *camyo, *mgmyo;
    double *xiup, *xileak;
    /* membrane (cuda_Cell_Detailed.cu:350-366) */
    double v, xm, xh, xj, xr, xs1, xs2, qks, xtos, ytos, xtof, ytof, xnai;
    double csubt, xicatto, xinacato, xina, xiks, xikr, xik1, xinak, xitos, xitof, stim;
    double out_ica, out_ncx, out_ina;
    int64_t step;
    int64_t counters[RYR_C_COUNT];
    double diag[RYR_DIAG_COLS];
    int64_t diag_step;
    double *line;           /* X * RYR_LINE_COLS */
    double *scratch;        /* N reduction scratch */
    /* voltage clamp waveform */
    double *vclamp;
    int64_t vclamp_n;
    int vclamp_periodic;
    /* beat analyser (definitions in include/ryr_cru.h, enum ryr_beat_col) */
    int64_t beat_on;
    int beat_n, beat_flags;
    double b_vrest, b_vmax, b_vmin, b_apd90, b_ead_t, b_peak_ci, b_jsr, b_sparks, b_eadl_t;
    double beats[RYR_MAX_BEATS][RYR_BEAT_COLS];
};

#define POS(c, i, j, k) ((int64_t)(i) + (int64_t)(c)->nx * ((int64_t)(j) + (int64_t)(c)->ny * (int64_t)(k)))
#define INTERIOR(c, i, j, k) ((int64_t)((i) - 1) + (int64_t)(c)->X * ((int64_t)((j) - 1) + (int64_t)(c)->Y * (int64_t)((k) - 1)))

static void *orc_alloc(int64_t n, size_t sz) { return calloc((size_t)n, sz); }

/* variant timing in whole steps (t = n dt; the reference tests t with fmod) */
static int64_t orc_ncl(const orc_cell *c) { return (int64_t)llround(c->d.cl / c->p.dt); }
/* iso: t >= 5.5 CL (:1220) */
static int64_t orc_iso_n(const orc_cell *c) { return (11 * orc_ncl(c) + 1) / 2; }
/* iso: fmod(t + 10 CL, 15.5 CL) < dt (:1214) */
static int orc_iso_active(const orc_cell *c, int64_t n)
{
    const int64_t ncl = orc_ncl(c);
    return ncl > 0 && (2 * (n + 10 * ncl)) % (31 * ncl) < 2;
}
/* stimulus phase: iso paces at phase 0 (:684), the default 100 ms in (:689) */
static int64_t orc_stim_offset(const orc_cell *c)
{
    return (c->p.variants & RYR_VAR_ISO) ? 0 : (int64_t)llround(c->p.stim_offset / c->p.dt);
}
/* whether step n may be stimulated: never under perm (:697-699), while
 * t < 11.9 CL under iso (:684) */
static int orc_paced(const orc_cell *c, int64_t n)
{
    if (c->p.variants & RYR_VAR_PERM) return 0;
    if (c->p.variants & RYR_VAR_ISO) return n < (119 * orc_ncl(c) + 9) / 10;
    return 1;
}

orc_cell *orc_create(const ryr_params *p, const ryr_cell_desc *d)
{
    if (d->nx < 1 || d->ny < 1 || d->nz < 1 || p->dt <= 0) return NULL;
    orc_cell *c = (orc_cell *)calloc(1, sizeof *c);
    c->p = *p;
    c->d = *d;
    c->X = d->nx; c->Y = d->ny; c->Z = d->nz;
    c->nx = c->X + 2; c->ny = c->Y + 2; c->nz = c->Z + 2;
    c->ncell = (int64_t)c->nx * c->ny * c->nz;
    c->N = (int64_t)c->X * c->Y * c->Z;
    int64_t n1 = c->ncell, n8 = 8 * c->ncell;
    c->cp = orc_alloc(n1, 8); c->cjsr = orc_alloc(n1, 8); c->cjsrb = orc_alloc(n1, 8);
    c->xire = orc_alloc(n1, 8); c->po = orc_alloc(n1, 8); c->Ancx = orc_alloc(n1, 8);
    c->xinaca = orc_alloc(n1, 8); c->xicat = orc_alloc(n1, 8);
    c->casar = orc_alloc(n1, 8); c->casarh = orc_alloc(n1, 8);
    c->casarj = orc_alloc(n1, 8); c->casarhj = orc_alloc(n1, 8);
    c->cs = orc_alloc(n1, 8); c->csnext = orc_alloc(n1, 8); c->lagB = orc_alloc(n1, 8);
    c->lcc = orc_alloc(n8, sizeof(int)); c->nl = orc_alloc(n1, sizeof(int));
    c->nou = orc_alloc(n1, sizeof(int)); c->ncu = orc_alloc(n1, sizeof(int));
    c->nob = orc_alloc(n1, sizeof(int)); c->ncb = orc_alloc(n1, sizeof(int));
    c->nspark = orc_alloc(n1, sizeof(int));
    c->ci = orc_alloc(n8, 8); c->cinext = orc_alloc(n8, 8);
    c->cnsr = orc_alloc(n8, 8); c->cnsrnext = orc_alloc(n8, 8);
    c->catf = orc_alloc(n8, 8); c->cats = orc_alloc(n8, 8); c->cacal = orc_alloc(n8, 8);
    c->casr = orc_alloc(n8, 8); c->camyo = orc_alloc(n8, 8); c->mgmyo = orc_alloc(n8, 8);
    c->xiup = orc_alloc(n8, 8); c->xileak = orc_alloc(n8, 8);
    c->line = orc_alloc((int64_t)c->X * RYR_LINE_COLS, 8);
    c->scratch = orc_alloc(c->N > 0 ? c->N : 1, 8);

    /* scales (cuda_Cell_Detailed.cu:96,100,107,109,119-143,1205,1230-1237) */
    c->Cm = (p->cm_mode == 0) ? p->cm * (double)c->N / 16120.0 : p->cm;
    c->svica = ((0.666) * (1.2) * (d->ica_scale));
    c->svncx = (2.5 * d->ncx_scale);
    c->svikr = p->lqt ? 0.0 : 3.0 * d->ikr_scale;
    c->sviup = O_SV_IUP;
    if (p->iso2) c->sviup = O_SV_IUP * 1.75;
    if (p->tapsi) c->sviup = 0;
    c->sviup_lqt = (p->lqt && p->tapsi) ? 0.05 : 0.;

    /* Init (cuda_Cell_Detailed.cu:1108-1182) over every CRU incl. ghosts */
    double ci_basal = p->ci0;
    double cjsr_b = p->tapsi ? 10 : d->cjsr0;
    double fb = orc_fracbound(cjsr_b);
    for (int64_t ac = 0; ac < n1; ++ac) {
        c->cp[ac] = ci_basal;
        c->cs[ac] = ci_basal;
        c->csnext[ac] = ci_basal;
        c->cjsr[ac] = cjsr_b;
        c->cjsrb[ac] = orc_luminal_SS(cjsr_b);
        c->nspark[ac] = 0;
        c->Ancx[ac] = 0.025;
        for (int ii = 0; ii < 8; ++ii) {
            int64_t a8 = ac * 8 + ii;
            c->ci[a8] = ci_basal; c->cnsr[a8] = cjsr_b;
            c->cinext[a8] = ci_basal; c->cnsrnext[a8] = cjsr_b;
            c->catf[a8] = O_KTFON * ci_basal * O_BTF / (O_KTFON * ci_basal + O_KTFOFF);
            c->cats[a8] = O_KTSON * ci_basal * O_BTS / (O_KTSON * ci_basal + O_KTSOFF);
            c->cacal[a8] = O_KCALON * ci_basal * O_BCAL / (O_KCALON * ci_basal + O_KCALOFF);
            c->casr[a8] = O_KSRON * ci_basal * O_BSR / (O_KSRON * ci_basal + O_KSROFF);
            c->camyo[a8] = 3.;
            c->mgmyo[a8] = 136.209996;
        }
        c->casar[ac] = O_KSARON * ci_basal * O_BSAR / (O_KSARON * ci_basal + O_KSAROFF);
        c->casarh[ac] = O_KSARHON * ci_basal * O_BSARH / (O_KSARHON * ci_basal + O_KSARHOFF);
        c->casarj[ac] = O_KSARON * ci_basal * O_BSAR / (O_KSARON * ci_basal + O_KSAROFF);
        c->casarhj[ac] = O_KSARHON * ci_basal * O_BSARH / (O_KSARHON * ci_basal + O_KSARHOFF);
        int nlc = p->nlcc;
        if (p->variants & RYR_VAR_RANDOMLCC) {
            /* randpoint = curand % 9 (:1150-1152): one 32-bit draw of the
             * CRU's stream at counter step 2^64-1 */
            const int64_t gi = ac % c->nx, gj = (ac / c->nx) % c->ny, gk = ac / ((int64_t)c->nx * c->ny);
            if (gi >= 1 && gi <= c->X && gj >= 1 && gj <= c->Y && gk >= 1 && gk <= c->Z) {
                ryr_stream s0;
                ryr_stream_init(&s0, d->seed, (uint32_t)INTERIOR(c, gi, gj, gk), ~(uint64_t)0);
                nlc = (int)(ryr_stream_u32(&s0) % 9u);
            }
        }
        for (int ll = 0; ll < 8; ll++) c->lcc[ac * 8 + ll] = (ll < nlc) ? 3 : 16;
        c->ncb[ac] = (int)(fb * O_NRYR);
        c->ncu[ac] = O_NRYR - (int)(fb * O_NRYR);
        c->nob[ac] = 0;
        c->nou[ac] = 0;
    }
    /* host membrane ICs (cuda_Cell_Detailed.cu:353-365) */
    c->xnai = 6.1 * ((9000) / (d->cl + 5000));
    c->v = p->v0;
    c->xm = 0.0010; c->xh = 1.00; c->xj = 1.00; c->xr = 0.00;
    c->xs1 = 0.08433669901; c->xs2 = c->xs1; c->qks = 0.20;
    c->xtos = 0.010; c->ytos = 1.00; c->xtof = 0.020; c->ytof = 0.80;
    c->diag_step = -1;
    c->beat_on = -1;
    return c;
}

void orc_destroy(orc_cell *c)
{
    if (!c) return;
    void *ptrs[] = {c->cp, c->cjsr, c->cjsrb, c->xire, c->po, c->Ancx, c->xinaca, c->xicat,
                    c->casar, c->casarh, c->casarj, c->casarhj, c->cs, c->csnext, c->lagB, c->lcc, c->nl,
                    c->nou, c->ncu, c->nob, c->ncb, c->nspark, c->ci, c->cinext, c->cnsr,
                    c->cnsrnext, c->catf, c->cats, c->cacal, c->casr, c->camyo, c->mgmyo,
                    c->xiup, c->xileak, c->line, c->scratch, c->vclamp};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
    free(c);
}

/* 7-point coupling of one sub-voxel on the 2x-refined grid
 * (cuda_Cell_Detailed.cu:1320-1340): ii bit0 = z, bit1 = y, bit2 = x;
 * tt = transverse (y, z) and tl = longitudinal (x) time constant * xi. */
static double orc_coupling(const orc_cell *c, const double *f, int i, int j, int k, int ii,
                           double tt, double tl)
{
    const int64_t ac = POS(c, i, j, k), act = ac * 8 + ii;
    int64_t north = (ii % 2) ? (POS(c, i, j, k + 1) * 8 + ii - 1) : (ac * 8 + ii + 1);
    int64_t south = (ii % 2) ? (ac * 8 + ii - 1) : (POS(c, i, j, k - 1) * 8 + ii + 1);
    int64_t east = ((ii / 2) % 2) ? (POS(c, i, j + 1, k) * 8 + ii - 2) : (ac * 8 + ii + 2);
    int64_t west = ((ii / 2) % 2) ? (ac * 8 + ii - 2) : (POS(c, i, j - 1, k) * 8 + ii + 2);
    int64_t top = ((ii / 4) % 2) ? (POS(c, i + 1, j, k) * 8 + ii - 4) : (ac * 8 + ii + 4);
    int64_t bottom = ((ii / 4) % 2) ? (ac * 8 + ii - 4) : (POS(c, i - 1, j, k) * 8 + ii + 4);
    return ((f[north] - f[act]) / (tt) +
            (f[south] - f[act]) / (tt) +
            (f[east] - f[act]) / (tt) +
            (f[west] - f[act]) / (tt) +
            (f[top] - f[act]) / (tl) +
            (f[bottom] - f[act]) / (tl));
}

/* ---- Compute body for one interior CRU (cuda_Cell_Detailed.cu:1187-1598) */
static void orc_compute_cru(orc_cell *c, int i, int j, int k, int reset, int64_t *cnt)
{
    const double dt = c->p.dt;
    const double v = c->v, xnai = c->xnai;
    const int64_t ac = POS(c, i, j, k);
    double dotcjsr, dotci[8], dotcnsr[8], xicoupn[8], xicoupi[8];
    double *ci = c->ci, *cnsr = c->cnsr, *cs = c->cs;
    orc_rng g;
    memset(&g, 0, sizeof g);
    ryr_stream_init(&g.ph, c->d.seed, (uint32_t)INTERIOR(c, i, j, k), (uint64_t)c->step);

    const int var = c->p.variants;
    const int perm = (var & RYR_VAR_PERM) != 0;
    double sviup = c->sviup;                                        /* :1205-1238 */
    if (var & RYR_VAR_ISO) {
        if (orc_iso_active(c, c->step)) {                           /* :1214-1219 */
            for (int ll = 4; ll < 8; ++ll)
                if (c->lcc[ac * 8 + ll] == 16) c->lcc[ac * 8 + ll] = 3;
        }
        if (!c->p.iso2 && !c->p.tapsi && c->step >= orc_iso_n(c)) sviup = O_SV_IUP * 2.;
    }
    if (reset) { c->nspark[ac] = 0; c->nl[ac] = 0; }                /* :1244-1247 */

    /* LCC gating and I_CaL (:1248-1263) */
    double icat1 = 0;
    if (!perm) {
        int nlcp = 0;
        for (int ll = 0; ll < 8 && c->lcc[ac * 8 + ll] < 16; ll++) {
            double cpx = (c->cp[ac] + cs[ac] * O_ICAGAMMA) / (1 + O_ICAGAMMA);
            int jj = orc_lcc_markov(c->lcc[ac * 8 + ll], cpx, v, dt, orc_u(&g), c->p.iso2);
            c->lcc[ac * 8 + ll] = jj;
            if (jj == 0) { ++c->nl[ac]; ++nlcp; }
        }
        icat1 = c->svica * (double)(nlcp) * orc_lcccurrent(v, (c->cp[ac] + cs[ac] * O_ICAGAMMA) / (1 + O_ICAGAMMA));
    }
    double bcgcur = 0;                                              /* :1265-1272 */
    if (var & RYR_VAR_SLOWNCX) bcgcur = 1.0 * 0.003016 / 32. * (v - ryr_log(180000 / cs[ac]) / 2. / O_FRT);
    c->xicat[ac] = icat1 + 0 + bcgcur * (O_VS / O_VP);

    /* RyR gating, spark onset, release (:1274-1293) */
    int prevsp = c->nou[ac] + c->nob[ac];
    int fails = 0;
    orc_ryrgating(&g, c->d.Ku, c->d.Ku2, c->d.Kb, c->cp[ac], c->cjsr[ac],
                  &c->ncu[ac], &c->nou[ac], &c->ncb[ac], &c->nob[ac], dt, &fails);
    cnt[RYR_C_GAUSS_FALLBACK] += fails;
    if (prevsp < 5 && (c->nou[ac] + c->nob[ac]) >= 5) c->nspark[ac] = 1;
    c->po[ac] = (double)(c->nou[ac] + c->nob[ac]) / (double)(O_NRYR);
    c->xire[ac] = orc_release(c->po[ac], c->cjsr[ac], c->cp[ac]) / 1.0;
    if (c->xire[ac] < 0) c->xire[ac] = 0;

    /* NCX (:1294-1304) */
    c->xinaca[ac] = perm ? 0 : c->svncx * 1.0 * orc_ncx_v(cs[ac] / 1000.0, v, xnai, &c->Ancx[ac], dt,
                                                         (var & RYR_VAR_SLOWNCX) != 0);

    /* SERCA and leak (:1305-1308) */
    for (int ii = 0; ii < 8; ++ii) {
        c->xiup[ac * 8 + ii] = (sviup + c->sviup_lqt) * orc_uptake(ci[ac * 8 + ii], cnsr[ac * 8 + ii]);
        c->xileak[ac * 8 + ii] = O_SVILEAK * orc_leak(cnsr[ac * 8 + ii], ci[ac * 8 + ii]);
    }
    /* exchange fluxes and JSR (:1309-1318) */
    double diffpi0 = (c->cp[ac] - ci[ac * 8]) / O_TAUPI / 2.;
    double diffpi1 = (c->cp[ac] - ci[ac * 8 + 1]) / O_TAUPI / 2.;
    double diffsi0 = (cs[ac] - ci[ac * 8]) / O_TAUSI / 2.;
    double diffsi1 = (cs[ac] - ci[ac * 8 + 1]) / O_TAUSI / 2.;
    double diffjn0 = (c->cjsr[ac] - cnsr[ac * 8]) / O_TAUTR / 2.;
    double diffjn1 = (c->cjsr[ac] - cnsr[ac * 8 + 1]) / O_TAUTR / 2.;
    dotcjsr = orc_luminal(c->cjsr[ac]) * ((-diffjn0 - diffjn1) - c->xire[ac] * O_VP * 1.0 / O_VJSR);
    c->cjsrb[ac] += (1 - orc_luminal(c->cjsr[ac])) * ((-diffjn0 - diffjn1) - c->xire[ac] * O_VP * 1.0 / O_VJSR) * dt;

    /* diffusion + cytosolic buffers (:1320-1355) */
    for (int ii = 0; ii < 8; ++ii) {
        int64_t act = ac * 8 + ii;
        xicoupn[ii] = orc_coupling(c, cnsr, i, j, k, ii, O_TAUNT * O_XI, O_TAUNL * O_XI);
        xicoupi[ii] = orc_coupling(c, ci, i, j, k, ii, O_TAUIT * O_XI, O_TAUIL * O_XI);
        dotcnsr[ii] = (c->xiup[act] - c->xileak[act]) * O_VI / O_VNSR + xicoupn[ii];
        double buffers = (b_tropf(c->catf[act], ci[act], dt) +
                          b_trops(c->cats[act], ci[act], dt) +
                          b_bucal(c->cacal[act], ci[act], dt) +
                          b_busr(c->casr[act], ci[act], dt) +
                          b_myoca(c->camyo[act], c->mgmyo[act], ci[act], dt));
        dotci[ii] = (-c->xiup[act] + c->xileak[act]) - buffers + xicoupi[ii];
    }

    const double dth = dt / 5;
    double *cinext = c->cinext, *cnsrnext = c->cnsrnext, *csnext = c->csnext;
    if (c->p.scheme == RYR_SCHEME_A) {
        /* scheme A (:1356-1473) */
        dotcnsr[0] += diffjn0 * O_VJSR / O_VNSR;
        dotcnsr[1] += diffjn1 * O_VJSR / O_VNSR;
        dotci[0] += diffsi0 * (O_VS / O_VI);
        dotci[1] += diffsi1 * (O_VS / O_VI);
        dotci[0] += diffpi0 * (O_VP / O_VI);
        dotci[1] += diffpi1 * (O_VP / O_VI);
        for (int ii = 0; ii < 8; ++ii) {
            int64_t act = ac * 8 + ii;
            cinext[act] = ci[act] + dotci[ii] * dt;
            if (cinext[act] < 0) cinext[act] = 0;
            cnsrnext[act] = cnsr[act] + dotcnsr[ii] * dt;
            c->catf[act] += b_tropf(c->catf[act], ci[act], dt) * dt;
            if (c->catf[act] < 0) c->catf[act] = 0;
            c->cats[act] += b_trops(c->cats[act], ci[act], dt) * dt;
            if (c->cats[act] < 0) c->cats[act] = 0;
            c->cacal[act] += b_bucal(c->cacal[act], ci[act], dt) * dt;
            if (c->cacal[act] < 0) c->cacal[act] = 0;
            c->casr[act] += b_busr(c->casr[act], ci[act], dt) * dt;
            if (c->casr[act] < 0) c->casr[act] = 0;
            c->camyo[act] += b_myoca(c->camyo[act], c->mgmyo[act], ci[act], dt) * dt;
            if (c->camyo[act] < 0) c->camyo[act] = 0;
            c->mgmyo[act] += b_myomg(c->camyo[act], c->mgmyo[act], ci[act], dt) * dt;
            if (c->mgmyo[act] < 0) c->mgmyo[act] = 0;
        }
        csnext[ac] = cs[ac];
        double cpstart = c->cp[ac];
        for (int iii = 0; iii < 5; ++iii) {
            double csbuffers = (b_busar(c->casar[ac], csnext[ac], dth) +
                                b_busarh(c->casarh[ac], csnext[ac], dth));
            double csdiff = (cs[POS(c, i, j, k + 1)] + cs[POS(c, i, j, k - 1)] - 2 * cs[ac]) / (O_TAUST);
            if (perm) csdiff += (c->p.ci0 - csnext[ac]) / 5;          /* :1409-1411 */
            double dotcs = O_VP / O_VS * (c->cp[ac] - csnext[ac]) / (O_TAUPS)
                           - O_ICAGAMMA * icat1 * O_VP / O_VS
                           + c->xinaca[ac]
                           - (cs[ac] - ci[ac * 8]) / O_TAUSI / 2.
                           - (cs[ac] - ci[ac * 8 + 1]) / O_TAUSI / 2.
                           + csdiff
                           - csbuffers
                           - bcgcur;
            double cpbuff = b_busarj(c->casarj[ac], c->cp[ac], dth);
            double cpbuffh = b_busarhj(c->casarhj[ac], c->cp[ac], dth);
            double cpdot = ((csnext[ac] - c->cp[ac]) / O_TAUPS +
                            (ci[ac * 8] - cpstart) / O_TAUPI / 2. +
                            (ci[ac * 8 + 1] - cpstart) / O_TAUPI / 2. +
                            c->xire[ac] - icat1 - cpbuff - cpbuffh);
            c->casarj[ac] += cpbuff * dt / 5;
            if (c->casarj[ac] < 0) c->casarj[ac] = 0;
            c->casarhj[ac] += cpbuffh * dt / 5;
            if (c->casarhj[ac] < 0) c->casarhj[ac] = 0;
            c->cp[ac] += cpdot * dt / 5;
            c->casar[ac] += b_busar(c->casar[ac], csnext[ac], dt / 5) * dt / 5;
            if (c->casar[ac] < 0) c->casar[ac] = 0;
            c->casarh[ac] += b_busarh(c->casarh[ac], csnext[ac], dt / 5) * dt / 5;
            if (c->casarh[ac] < 0) c->casarh[ac] = 0;
            csnext[ac] += (dotcs) * dth;
        }
    } else {
        /* scheme B (:1475-1597) */
        dotcnsr[0] += diffjn0 * O_VJSR / O_VNSR;
        dotcnsr[1] += diffjn1 * O_VJSR / O_VNSR;
        csnext[ac] = cs[ac];
        double lag = 0;       /* checker's record, not part of the step: see orc_read_scheme_b_lag */
        for (int iii = 0; iii < 5; ++iii) {
            lag += cs[ac] - csnext[ac];
            double dpi0 = (c->cp[ac] - ci[ac * 8]) / O_TAUPI / 2.;
            double dpi1 = (c->cp[ac] - ci[ac * 8 + 1]) / O_TAUPI / 2.;
            double dsi0 = (csnext[ac] - ci[ac * 8]) / O_TAUSI / 2.;
            double dsi1 = (csnext[ac] - ci[ac * 8 + 1]) / O_TAUSI / 2.;
            dotci[0] += dsi0 * (O_VS / O_VI) / 5;
            dotci[1] += dsi1 * (O_VS / O_VI) / 5;
            dotci[0] += dpi0 * (O_VP / O_VI) / 5;
            dotci[1] += dpi1 * (O_VP / O_VI) / 5;
            double csbuffers = (b_busar(c->casar[ac], csnext[ac], dth) +
                                b_busarh(c->casarh[ac], csnext[ac], dth));
            double csdiff = (cs[POS(c, i, j, k + 1)] + cs[POS(c, i, j, k - 1)] - 2 * csnext[ac]) / (O_TAUST);
            if (perm) csdiff += (c->p.ci0 - csnext[ac]) / 5;          /* :1499-1501 */
            double dotcs = O_VP / O_VS * (c->cp[ac] - O_ICAGAMMA * O_TAUPS * icat1 - csnext[ac]) / (O_TAUPS)
                           + c->xinaca[ac]
                           - (csnext[ac] - ci[ac * 8]) / O_TAUSI / 2.
                           - (csnext[ac] - ci[ac * 8 + 1]) / O_TAUSI / 2.
                           + csdiff
                           - csbuffers
                           - bcgcur;
            double cpbuff = b_busarj(c->casarj[ac], c->cp[ac], dth);
            double cpbuffh = b_busarhj(c->casarhj[ac], c->cp[ac], dth);
            double cpdot = ((cs[ac] - c->cp[ac]) / O_TAUPS +
                            (ci[ac * 8] - c->cp[ac]) / O_TAUPI / 2. +
                            (ci[ac * 8 + 1] - c->cp[ac]) / O_TAUPI / 2. +
                            c->xire[ac] - icat1 - cpbuff - cpbuffh);
            c->casar[ac] += b_busar(c->casar[ac], csnext[ac], dt / 5) * dth;
            if (c->casar[ac] < 0) c->casar[ac] = 0;
            c->casarh[ac] += b_busarh(c->casarh[ac], csnext[ac], dt / 5) * dth;
            if (c->casarh[ac] < 0) c->casarh[ac] = 0;
            csnext[ac] += (dotcs) * dth;
            c->casarj[ac] += cpbuff * dth;
            if (c->casarj[ac] < 0) c->casarj[ac] = 0;
            c->casarhj[ac] += cpbuffh * dth;
            if (c->casarhj[ac] < 0) c->casarhj[ac] = 0;
            c->cp[ac] += cpdot * dth;
        }
        c->lagB[ac] = lag;
        for (int ii = 0; ii < 8; ++ii) {
            int64_t act = ac * 8 + ii;
            cinext[act] = ci[act] + dotci[ii] * dt;
            if (cinext[act] < 0) cinext[act] = 0;
            cnsrnext[act] = cnsr[act] + dotcnsr[ii] * dt;
            c->catf[act] += b_tropf(c->catf[act], ci[act], dt) * dt;
            if (c->catf[act] < 0) c->catf[act] = 0;
            c->cats[act] += b_trops(c->cats[act], ci[act], dt) * dt;
            if (c->cats[act] < 0) c->cats[act] = 0;
            c->cacal[act] += b_bucal(c->cacal[act], ci[act], dt) * dt;
            if (c->cacal[act] < 0) c->cacal[act] = 0;
            c->casr[act] += b_busr(c->casr[act], ci[act], dt) * dt;
            if (c->casr[act] < 0) c->casr[act] = 0;
            c->camyo[act] += b_myoca(c->camyo[act], c->mgmyo[act], ci[act], dt) * dt;
            if (c->camyo[act] < 0) c->camyo[act] = 0;
            c->mgmyo[act] += b_myomg(c->camyo[act], c->mgmyo[act], ci[act], dt) * dt;
            if (c->mgmyo[act] < 0) c->mgmyo[act] = 0;
        }
    }
    if (c->cp[ac] < 0) { c->cp[ac] = 0; cnt[RYR_C_CLAMP_CP]++; }
    if (csnext[ac] < 0) { csnext[ac] = 1e-6; cnt[RYR_C_CLAMP_CS]++; }
    c->cjsr[ac] = c->cjsr[ac] + dotcjsr * dt;
}

/* Ghost copies for one interior CRU (cuda_Cell_Detailed.cu:1601-1652). */
static void orc_boundary_perm(orc_cell *c, int i, int j, int k);
static void orc_boundary(orc_cell *c, int i, int j, int k)
{
    if (c->p.variants & RYR_VAR_PERM) { orc_boundary_perm(c, i, j, k); return; }
    double *cinext = c->cinext, *cnsrnext = c->cnsrnext, *csnext = c->csnext;
    if (k == 1) {
        for (int ii = 1; ii < 8; ii += 2) {
            cinext[POS(c, i, j, 0) * 8 + ii] = cinext[POS(c, i, j, 1) * 8 + ii - 1];
            cnsrnext[POS(c, i, j, 0) * 8 + ii] = cnsrnext[POS(c, i, j, 1) * 8 + ii - 1];
        }
        csnext[POS(c, i, j, 0)] = csnext[POS(c, i, j, 1)];
    }
    if (k == c->nz - 2) {
        for (int ii = 0; ii < 8; ii += 2) {
            cinext[POS(c, i, j, c->nz - 1) * 8 + ii] = cinext[POS(c, i, j, c->nz - 2) * 8 + ii + 1];
            cnsrnext[POS(c, i, j, c->nz - 1) * 8 + ii] = cnsrnext[POS(c, i, j, c->nz - 2) * 8 + ii + 1];
        }
        csnext[POS(c, i, j, c->nz - 1)] = csnext[POS(c, i, j, c->nz - 2)];
    }
    if (j == 1) {
        for (int ii = 2; ii < 8; ii += 4) {
            cinext[POS(c, i, 0, k) * 8 + ii] = cinext[POS(c, i, 1, k) * 8 + ii - 2];
            cinext[POS(c, i, 0, k) * 8 + ii + 1] = cinext[POS(c, i, 1, k) * 8 + ii - 1];
            cnsrnext[POS(c, i, 0, k) * 8 + ii] = cnsrnext[POS(c, i, 1, k) * 8 + ii - 2];
            cnsrnext[POS(c, i, 0, k) * 8 + ii + 1] = cnsrnext[POS(c, i, 1, k) * 8 + ii - 1];
        }
    }
    if (j == c->ny - 2) {
        for (int ii = 0; ii < 8; ii += 4) {
            cinext[POS(c, i, c->ny - 1, k) * 8 + ii] = cinext[POS(c, i, c->ny - 2, k) * 8 + ii + 2];
            cinext[POS(c, i, c->ny - 1, k) * 8 + ii + 1] = cinext[POS(c, i, c->ny - 2, k) * 8 + ii + 3];
            cnsrnext[POS(c, i, c->ny - 1, k) * 8 + ii] = cnsrnext[POS(c, i, c->ny - 2, k) * 8 + ii + 2];
            cnsrnext[POS(c, i, c->ny - 1, k) * 8 + ii + 1] = cnsrnext[POS(c, i, c->ny - 2, k) * 8 + ii + 3];
        }
    }
    if (i == 1) {
        for (int ii = 4; ii < 8; ++ii) {
            cinext[POS(c, 0, j, k) * 8 + ii] = cinext[POS(c, 1, j, k) * 8 + ii - 4];
            cnsrnext[POS(c, 0, j, k) * 8 + ii] = cnsrnext[POS(c, 1, j, k) * 8 + ii - 4];
        }
    }
    if (i == c->nx - 2) {
        for (int ii = 0; ii < 4; ++ii) {
            cinext[POS(c, c->nx - 1, j, k) * 8 + ii] = cinext[POS(c, c->nx - 2, j, k) * 8 + ii + 4];
            cnsrnext[POS(c, c->nx - 1, j, k) * 8 + ii] = cnsrnext[POS(c, c->nx - 2, j, k) * 8 + ii + 4];
        }
    }
}

/* perm build (:1662-1689): only the NSR ghosts are refreshed; the ci and cs
 * ghosts keep their set-up value ci0 for the whole run (a bath) */
static void orc_boundary_perm(orc_cell *c, int i, int j, int k)
{
    double *cnsrnext = c->cnsrnext;
    if (k == 1)
        for (int ii = 1; ii < 8; ii += 2) cnsrnext[POS(c, i, j, 0) * 8 + ii] = cnsrnext[POS(c, i, j, 1) * 8 + ii - 1];
    if (k == c->nz - 2)
        for (int ii = 0; ii < 8; ii += 2)
            cnsrnext[POS(c, i, j, c->nz - 1) * 8 + ii] = cnsrnext[POS(c, i, j, c->nz - 2) * 8 + ii + 1];
    if (j == 1)
        for (int ii = 2; ii < 8; ii += 4) {
            cnsrnext[POS(c, i, 0, k) * 8 + ii] = cnsrnext[POS(c, i, 1, k) * 8 + ii - 2];
            cnsrnext[POS(c, i, 0, k) * 8 + ii + 1] = cnsrnext[POS(c, i, 1, k) * 8 + ii - 1];
        }
    if (j == c->ny - 2)
        for (int ii = 0; ii < 8; ii += 4) {
            cnsrnext[POS(c, i, c->ny - 1, k) * 8 + ii] = cnsrnext[POS(c, i, c->ny - 2, k) * 8 + ii + 2];
            cnsrnext[POS(c, i, c->ny - 1, k) * 8 + ii + 1] = cnsrnext[POS(c, i, c->ny - 2, k) * 8 + ii + 3];
        }
    if (i == 1)
        for (int ii = 4; ii < 8; ++ii) cnsrnext[POS(c, 0, j, k) * 8 + ii] = cnsrnext[POS(c, 1, j, k) * 8 + ii - 4];
    if (i == c->nx - 2)
        for (int ii = 0; ii < 4; ++ii)
            cnsrnext[POS(c, c->nx - 1, j, k) * 8 + ii] = cnsrnext[POS(c, c->nx - 2, j, k) * 8 + ii + 4];
}

/* Whole-cell reduction + membrane ODE (cuda_Cell_Detailed.cu:447-739) on a
 * 13-value state [v xm xh xj xr xs1 xs2 qks xtos ytos xtof ytof xnai].
 * currents out: [csubt xicatto xinacato xina xiks xikr xik1 xinak xitos xitof] */
static void orc_membrane(double *s, double sum_icat, double sum_inaca, double sum_cs,
                         double ncru, double Cm, double dt, double svikr, double stim,
                         double *cur, int slowncx)
{
    double v = s[0], xm = s[1], xh = s[2], xj = s[3], xr = s[4], xs1 = s[5], xs2 = s[6];
    double qks = s[7], xtos = s[8], ytos = s[9], xtof = s[10], ytof = s[11], xnai = s[12];
    const double xki = 140.0, xko = 5.40;

    double csubt = sum_cs / ncru;
    double xicatto = sum_icat / Cm * 0.0965 * O_VP * 2.0 * (O_ICAGAMMA + 1.);
    double xinacato = sum_inaca / Cm * 0.0965 * O_VS;

    /* I_Na (:460-502) */
    double ena = (1.0 / O_FRT) * ryr_log(O_XNAO / xnai);
    double am = 0.32 * (v + 47.13) / (1.0 - ryr_exp(-0.1 * (v + 47.13)));
    double bm = 0.08 * ryr_exp(-v / 11.0);
    double ah, bh, aj, bj;
    if (v < -40.0) {
        ah = 0.135 * ryr_exp((80.0 + v) / (-6.8));
        bh = 3.56 * ryr_exp(0.079 * v) + 310000.0 * ryr_exp(0.35 * v);
        aj = (-127140.0 * ryr_exp(0.2444 * v) - 0.00003474 * ryr_exp(-0.04391 * v))
             * ((v + 37.78) / (1.0 + ryr_exp(0.311 * (v + 79.23))));
        bj = (0.1212 * ryr_exp(-0.01052 * v)) / (1.0 + ryr_exp(-0.1378 * (v + 40.14)));
    } else {
        ah = 0.00;
        bh = 1.00 / (0.130 * (1.00 + ryr_exp((v + 10.66) / (-11.10))));
        aj = 0.00;
        bj = (0.3 * ryr_exp(-0.00000025350 * v)) / (1.0 + ryr_exp(-0.10 * (v + 32.00)));
    }
    double tauh = 1.00 / (ah + bh);
    double tauj = 1.00 / (aj + bj);
    double taum = 1.00 / (am + bm);
    double gna = 15;
    double gnaleak = 0.3e-3 * 5;
    double xina = gna * xh * xj * xm * xm * xm * (v - ena) + gnaleak * (v - ena);
    xh = ah / (ah + bh) - ((ah / (ah + bh)) - xh) * ryr_exp(-dt / tauh);
    xj = aj / (aj + bj) - ((aj / (aj + bj)) - xj) * ryr_exp(-dt / tauj);
    xm = am / (am + bm) - ((am / (am + bm)) - xm) * ryr_exp(-dt / taum);

    /* I_Kr (:507-529) */
    double ek = (1.00 / O_FRT) * ryr_log(xko / xki);
    double gss = sqrt(xko / 5.40);
    double xkrv1 = 0.001380 * (v + 7.00) / (1.0 - ryr_exp(-0.123 * (v + 7.00)));
    double xkrv2 = 0.000610 * (v + 10.00) / (ryr_exp(0.1450 * (v + 10.00)) - 1.00);
    double taukr = 1.00 / (xkrv1 + xkrv2);
    double xkrinf = 1.00 / (1.00 + ryr_exp(-(v + 50.00) / 7.50));
    double rg = 1.00 / (1.00 + ryr_exp((v + 33.00) / 22.40));
    double gkr = 0.0078360;
    double xikr = svikr * gkr * gss * xr * rg * (v - ek);
    xr = xkrinf - (xkrinf - xr) * ryr_exp(-dt / taukr);

    /* I_Ks (:534-559) */
    double prnak = 0.0183300;
    double qks_inf = 0.2 * (1 + 0.8 / (1 + ryr_pow((0.28 / csubt), 3)));
    double tauqks = 1000.00;
    if (slowncx) {                                                  /* :540-543 */
        qks_inf = 0.2 * (1 + 1.6 / (1 + ryr_pow((0.28 / csubt), 3)));
        tauqks = 2000;
    }
    double eks = (1.00 / O_FRT) * ryr_log((xko + prnak * O_XNAO) / (xki + prnak * xnai));
    double xs1ss = 1.0 / (1.0 + ryr_exp(-(v - 1.500) / 16.700));
    double tauxs = 1.00 / (0.0000719 * (v + 30.00) / (1.00 - ryr_exp(-0.1480 * (v + 30.0)))
                           + 0.0001310 * (v + 30.00) / (ryr_exp(0.06870 * (v + 30.00)) - 1.00));
    tauxs = O_SVTAUXS * tauxs;
    double gksx = 0.2000;
    double sviks_iso = 1.;
    double xiks = 1.0 * sviks_iso * gksx * qks * xs1 * xs2 * (v - eks);
    xs1 = xs1ss - (xs1ss - xs1) * ryr_exp(-dt / tauxs);
    xs2 = xs1ss - (xs1ss - xs2) * ryr_exp(-dt / tauxs);
    qks = qks + dt * (qks_inf - qks) / tauqks;

    /* I_K1 (:590-600) */
    double gkix = 0.600;
    double gki = gkix * (sqrt(xko / 5.4));
    double aki = 1.02 / (1.0 + ryr_exp(0.2385 * (v - ek - 59.215)));
    double bki = (0.49124 * ryr_exp(0.08032 * (v - ek + 5.476)) + ryr_exp(0.061750 * (v - ek - 594.31)))
                 / (1.0 + ryr_exp(-0.5143 * (v - ek + 4.753)));
    double xkin = aki / (aki + bki);
    double xik1 = (1.0) * gki * xkin * (v - ek);

    /* I_to,s (:604-629) */
    double rt1 = -(v + 3.0) / 15.00;
    double rt2 = (v + 33.5) / 10.00;
    double rt3 = (v + 60.00) / 10.00;
    double xtos_inf = 1.00 / (1.0 + ryr_exp(rt1));
    double ytos_inf = 1.00 / (1.00 + ryr_exp(rt2));
    double rs_inf = 1.00 / (1.00 + ryr_exp(rt2));
    double txs = 9.00 / (1.00 + ryr_exp(-rt1)) + 0.50;
    double tys = 3000.00 / (1.0 + ryr_exp(rt3)) + 30.00;
    double gtos = 0.040;
    double xitos = 1.0 * gtos * xtos * (ytos + 0.50 * rs_inf) * (v - ek);
    xtos = xtos_inf - (xtos_inf - xtos) * ryr_exp(-dt / txs);
    ytos = ytos_inf - (ytos_inf - ytos) * ryr_exp(-dt / tys);

    /* I_to,f (:633-659) */
    double xtof_inf = xtos_inf;
    double ytof_inf = ytos_inf;
    double rt4 = -(v / 30.00) * (v / 30.00);
    double rt5 = (v + 33.50) / 10.00;
    double txf = 3.50 * ryr_exp(rt4) + 1.50;
    double tyf = 20.0 / (1.0 + ryr_exp(rt5)) + 20.00;
    double gtof = 0.10;
    double xitof = 1.0 * gtof * xtof * ytof * (v - ek);
    xtof = xtof_inf - (xtof_inf - xtof) * ryr_exp(-dt / txf);
    ytof = ytof_inf - (ytof_inf - ytof) * ryr_exp(-dt / tyf);

    /* I_NaK (:665-676) */
    double xkmko = 1.50, xkmnai = 12.00, xibarnak = 1.5000, hh = 1.00;
    double sigma = (ryr_exp(O_XNAO / 67.30) - 1.00) / 7.00;
    double fnak = 1.00 / (1.0 + 0.1245 * ryr_exp(-0.1 * v * O_FRT) + 0.0365 * sigma * ryr_exp(-v * O_FRT));
    double xinak = 1.0 * xibarnak * fnak * (1.0 / (1.0 + ryr_pow((xkmnai / xnai), hh))) * xko / (xko + xkmko);

    /* Na and V (:701-738) */
    double wcainv = 1.0 / 50.0;
    double conv = 0.18 * 12500;
    double xinaca = xinacato;
    double trick = 1.0;
    double xrr = trick * (1.0 / wcainv / conv) / 1000.0;
    double dnai = -xrr * (xina + 3.0 * xinak + 3.0 * xinaca);
    xnai += dnai * dt;
    double dvh = -(xina + xik1 + 1 * xikr + 1 * xiks + xitos + 1 * xitof + 1. * xinacato + 1 * xicatto + xinak) + stim;
    v += dvh * dt;

    s[0] = v; s[1] = xm; s[2] = xh; s[3] = xj; s[4] = xr; s[5] = xs1; s[6] = xs2;
    s[7] = qks; s[8] = xtos; s[9] = ytos; s[10] = xtof; s[11] = ytof; s[12] = xnai;
    cur[0] = csubt; cur[1] = xicatto; cur[2] = xinacato; cur[3] = xina; cur[4] = xiks;
    cur[5] = xikr; cur[6] = xik1; cur[7] = xinak; cur[8] = xitos; cur[9] = xitof;
}

int orc_membrane_probe(double *state13, double sum_icat, double sum_inaca, double sum_cs,
                       double ncru, double cm, double dt, double ikr_scale, double stim,
                       double *currents)
{
    orc_membrane(state13, sum_icat, sum_inaca, sum_cs, ncru, cm, dt, 3.0 * ikr_scale, stim, currents, 0);
    return 0;
}

/* ---- beat analyser (no reference counterpart; definitions in the header) */
static void orc_beat_close(orc_cell *c, double complete)
{
    if (c->beat_n < RYR_MAX_BEATS) {
        double *r = c->beats[c->beat_n];
        r[0] = (double)c->beat_on; r[1] = c->b_apd90; r[2] = c->b_vmax; r[3] = c->b_vrest;
        r[4] = (c->beat_flags & 2) ? 1.0 : 0.0; r[5] = c->b_ead_t; r[6] = c->b_peak_ci;
        r[7] = c->b_jsr; r[8] = c->b_sparks; r[9] = complete;
        r[10] = (c->beat_flags & 8) ? 1.0 : 0.0; r[11] = c->b_eadl_t;
    }
    c->beat_n++;
}

static void orc_beat_track(orc_cell *c, int64_t n, double vo, double vn)
{
    const double dt = c->p.dt;
    int64_t ncl = (int64_t)llround(c->d.cl / dt), noff = orc_stim_offset(c);
    if (ncl > 0 && n >= noff && (n - noff) % ncl == 0 && orc_paced(c, n)) {
        if (c->beat_on >= 0) orc_beat_close(c, 1.0);
        c->beat_on = n;
        c->b_vrest = vo; c->b_vmax = vo; c->b_vmin = vo;
        c->b_apd90 = -1.0; c->b_ead_t = -1.0; c->beat_flags = 0;
        c->b_peak_ci = 0.0; c->b_jsr = -1.0; c->b_sparks = 0.0; c->b_eadl_t = -1.0;
    }
    if (c->beat_on < 0) return;
    /* the literal flag: after the upstroke V starts to fall (bit2); then any rise ending above -40 mV */
    if (!(c->beat_flags & 4)) {
        if (c->b_vmax > -40.0 && vn < vo) c->beat_flags |= 4;
    } else if (!(c->beat_flags & 8) && vn > vo && vn > -40.0) {
        c->beat_flags |= 8;
        c->b_eadl_t = (double)(n + 1) * dt - (double)c->beat_on * dt;
    }
    if (vn > c->b_vmax) c->b_vmax = vn;
    double amp = c->b_vmax - c->b_vrest;
    if (!(c->beat_flags & 1)) {
        if (c->b_vmax > -40.0 && vn < c->b_vmax - 0.4 * amp &&
            (double)(n + 1) * dt - (double)c->beat_on * dt >= 50.0) { c->beat_flags |= 1; c->b_vmin = vn; }
    } else {
        if (vn < c->b_vmin) c->b_vmin = vn;
        if (!(c->beat_flags & 2) && vn > -40.0 && vn - c->b_vmin >= 1.0) {
            c->beat_flags |= 2;
            c->b_ead_t = (double)(n + 1) * dt - (double)c->beat_on * dt;
        }
    }
    if (c->b_apd90 < 0.0 && c->b_vmax > -40.0) {
        double thr = c->b_vmax - 0.9 * amp;
        if (vo >= thr && vn < thr)
            c->b_apd90 = ((double)n + (vo - thr) / (vo - vn)) * dt - (double)c->beat_on * dt;
    }
}

int orc_set_vclamp(orc_cell *c, const double *v, int64_t n, int periodic)
{
    free(c->vclamp);
    c->vclamp = NULL;
    c->vclamp_n = 0;
    c->vclamp_periodic = periodic ? 1 : 0;
    if (n > 0) {
        c->vclamp = (double *)malloc(sizeof(double) * (size_t)n);
        memcpy(c->vclamp, v, sizeof(double) * (size_t)n);
        c->vclamp_n = n;
    }
    return 0;
}

int orc_read_beats(const orc_cell *c, double *out, int maxbeats, int *nbeats)
{
    int done = c->beat_n < RYR_MAX_BEATS ? c->beat_n : RYR_MAX_BEATS, k = 0;
    for (; k < done && k < maxbeats; ++k) {
        memcpy(out + k * RYR_BEAT_COLS, c->beats[k], sizeof(double) * RYR_BEAT_COLS);
        out[k * RYR_BEAT_COLS] *= c->p.dt;
    }
    if (c->beat_on >= 0 && k < maxbeats && c->beat_n < RYR_MAX_BEATS) {
        double *r = out + k * RYR_BEAT_COLS;
        r[0] = (double)c->beat_on * c->p.dt; r[1] = c->b_apd90; r[2] = c->b_vmax; r[3] = c->b_vrest;
        r[4] = (c->beat_flags & 2) ? 1.0 : 0.0; r[5] = c->b_ead_t; r[6] = c->b_peak_ci; r[7] = c->b_jsr;
        r[8] = c->b_sparks; r[9] = 0.0;
        r[10] = (c->beat_flags & 8) ? 1.0 : 0.0; r[11] = c->b_eadl_t;
        ++k;
    }
    *nbeats = k;
    return 0;
}

/* stimulus of step n: t = n*dt inside [offset, offset+dur) modulo CL
 * (cuda_Cell_Detailed.cu:689-692 as an integer-step test) */
static double orc_stim(const orc_cell *c, int64_t n)
{
    int64_t ncl = (int64_t)llround(c->d.cl / c->p.dt);
    int64_t noff = orc_stim_offset(c);
    int64_t ndur = (int64_t)llround(c->p.stim_dur / c->p.dt);
    if (ncl <= 0 || !orc_paced(c, n)) return 0.0;
    int64_t ph = (n - noff) % ncl;
    if (ph < 0) ph += ncl;
    return ph < ndur ? c->p.stim_amp : 0.0;
}

/* output row (cuda_Cell_Detailed.cu:777-1081) */
static void orc_diag(orc_cell *c)
{
    const int64_t N = c->N;
    double *sc = c->scratch;
    double dN = (double)N;
#define ORC_SUM(expr_cru) ({                                               \
        for (int k = 1; k <= c->Z; ++k)                                        \
            for (int j = 1; j <= c->Y; ++j)                                    \
                for (int i = 1; i <= c->X; ++i) {                              \
                    int64_t ac = POS(c, i, j, k);                              \
                    (void)ac;                                                  \
                    sc[INTERIOR(c, i, j, k)] = (expr_cru);                     \
                }                                                              \
        orc_canonical_sum(sc, N); })
#define ORC_SUB8(arr) ({ double s_ = 0; for (int ii = 0; ii < 8; ++ii) s_ = s_ + (arr)[ac * 8 + ii] / 8.; s_; })

    double ncxfwd = ORC_SUM(c->xinaca[ac] < 0 ? c->xinaca[ac] : 0.0) / dN;
    double outAncx = ORC_SUM(c->Ancx[ac]) / dN;
    double cproxit = ORC_SUM(c->cp[ac]) / dN;
    double cjsrt = ORC_SUM(c->cjsr[ac]) / dN;
    double csqnbt = ORC_SUM(c->cjsrb[ac]) / dN;
    double poto = ORC_SUM(c->po[ac]) / dN;
    double pbcto = ORC_SUM((double)c->ncb[ac]) / dN / 100.;
    double casart = ORC_SUM(c->casar[ac]) / dN;
    double casarht = ORC_SUM(c->casarh[ac]) / dN;
    double casarjt = ORC_SUM(c->casarj[ac]) / dN;
    double casarhjt = ORC_SUM(c->casarhj[ac]) / dN;
    double cit = ORC_SUM(ORC_SUB8(c->ci)) / dN;
    double cnsrt = ORC_SUM(ORC_SUB8(c->cnsr)) / dN;
    double catft = ORC_SUM(ORC_SUB8(c->catf)) / dN;
    double catst = ORC_SUM(ORC_SUB8(c->cats)) / dN;
    double casrt = ORC_SUM(ORC_SUB8(c->casr)) / dN;
    double camyot = ORC_SUM(ORC_SUB8(c->camyo)) / dN;
    double cacalt = ORC_SUM(ORC_SUB8(c->cacal)) / dN;
    double leakt = ORC_SUM(ORC_SUB8(c->xileak)) / dN;
    double upt = ORC_SUM(ORC_SUB8(c->xiup)) / dN;
    double ire = ORC_SUM(c->xire[ac]) / dN;
#undef ORC_SUM
#undef ORC_SUB8
    int64_t tnou = 0, tnob = 0, tncu = 0, tncb = 0, sparks = 0;
    int64_t tn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 1; k <= c->Z; ++k)
        for (int j = 1; j <= c->Y; ++j)
            for (int i = 1; i <= c->X; ++i) {
                int64_t ac = POS(c, i, j, k);
                tnou += c->nou[ac]; tnob += c->nob[ac]; tncu += c->ncu[ac]; tncb += c->ncb[ac];
                sparks += c->nspark[ac];
                for (int ii = 0; ii < 8; ++ii)
                    if (c->ci[ac * 8 + ii] > 1000) c->counters[RYR_C_CI_RUNAWAY]++;
                for (int jj = 0; jj < 8; ++jj) {
                    switch (c->lcc[ac * 8 + jj]) {       /* :904-923 */
                    case 1: ++tn[1]; ++tn[6]; break;
                    case 2: ++tn[2]; break;
                    case 3: ++tn[1]; ++tn[2]; break;
                    case 4: ++tn[3]; ++tn[5]; break;
                    case 5: ++tn[1]; ++tn[3]; ++tn[5]; break;
                    case 6: ++tn[2]; ++tn[3]; ++tn[5]; break;
                    case 7: ++tn[1]; ++tn[2]; ++tn[3]; ++tn[5]; break;
                    case 8: ++tn[4]; break;
                    case 9: ++tn[1]; ++tn[4]; break;
                    case 10: ++tn[2]; ++tn[4]; break;
                    case 11: ++tn[1]; ++tn[2]; ++tn[4]; break;
                    case 12: ++tn[3]; ++tn[4]; break;
                    case 13: ++tn[1]; ++tn[3]; ++tn[4]; break;
                    case 14: ++tn[2]; ++tn[3]; ++tn[4]; break;
                    case 15: ++tn[1]; ++tn[2]; ++tn[3]; ++tn[4]; break;
                    case 0: ++tn[7]; ++tn[6]; break;
                    }
                }
            }
    double csubt = c->csubt;
    double Total_Ca = (cit + catft + catst + casrt + camyot + cacalt) * O_VI * 8. +
                      (csubt + casart + casarht) * O_VS +
                      (cproxit + casarjt + casarhjt) * O_VP +
                      (cnsrt) * O_VNSR * 8. +
                      (cjsrt + csqnbt) * O_VJSR;
    double Total_Ci = (cit + catft + catst + casrt + camyot + cacalt) * O_VI * 8. +
                      (csubt + casart + casarht) * O_VS +
                      (cproxit + casarjt + casarhjt) * O_VP;
    double spo = (double)c->p.steps_per_output;
    double *r = c->diag;
    r[0] = (double)c->step * c->p.dt; r[1] = cit; r[2] = c->v;
    r[3] = c->out_ncx / spo; r[4] = c->out_ica / spo; r[5] = cproxit; r[6] = csubt;
    r[7] = cjsrt; r[8] = cnsrt; r[9] = poto; r[10] = c->xnai; r[11] = c->xiks;
    r[12] = c->xikr; r[13] = c->xik1; r[14] = c->xinak; r[15] = c->xitos; r[16] = c->xitof;
    r[17] = c->out_ina / spo; r[18] = c->xr; r[19] = ncxfwd * (O_VS / O_VP); r[20] = pbcto;
    r[21] = cacalt; r[22] = catft; r[23] = leakt; r[24] = upt; r[25] = ire;
    r[26] = (double)tnou; r[27] = (double)tnob; r[28] = (double)tncu; r[29] = (double)tncb;
    for (int q = 1; q <= 7; ++q) r[29 + q] = (double)tn[q];
    r[37] = outAncx; r[38] = c->xs1; r[39] = c->qks; r[40] = (double)sparks;
    r[41] = catst; r[42] = csqnbt; r[43] = Total_Ci; r[44] = Total_Ca;
    c->diag_step = c->step;
    if (c->beat_on >= 0) {                   /* Ca features of the beat */
        if (c->step == c->beat_on) c->b_jsr = cjsrt;
        else c->b_sparks += (double)sparks;
        if (cit > c->b_peak_ci) c->b_peak_ci = cit;
    }

    /* linescan (:967-999): jy = ny/2, jz = 5 in ghost-inclusive numbering,
     * clamped to the interior for lattices with fewer than 5 z-layers */
    int jy = c->ny / 2;
    int jz = 5 < c->Z ? 5 : c->Z;
    for (int jx = 1; jx <= c->X; ++jx) {
        int64_t ac = POS(c, jx, jy, jz);
        double *L = c->line + (int64_t)(jx - 1) * RYR_LINE_COLS;
        L[0] = r[0]; L[1] = (double)jx; L[2] = c->ci[ac * 8]; L[3] = c->nob[ac]; L[4] = c->nou[ac];
        L[5] = c->ncb[ac]; L[6] = c->ncu[ac]; L[7] = c->nl[ac]; L[8] = c->cs[ac]; L[9] = c->cp[ac];
        L[10] = c->casar[ac]; L[11] = c->casarh[ac]; L[12] = c->camyo[ac * 8]; L[13] = c->mgmyo[ac * 8];
        L[14] = c->casr[ac * 8]; L[15] = c->cacal[ac * 8]; L[16] = c->cats[ac * 8];
        L[17] = c->catf[ac * 8]; L[18] = c->xire[ac]; L[19] = c->cjsr[ac]; L[20] = c->Ancx[ac];
        L[21] = c->xinaca[ac];
    }
}

static void orc_one_step(orc_cell *c)
{
    const int64_t n = c->step;
    const int spo = c->p.steps_per_output;
    const int reset = (n == 0) || ((n - 1) % spo == 0);
#ifdef _OPENMP
#pragma omp parallel
    {
        int64_t loc[RYR_C_COUNT] = {0};
#pragma omp for collapse(3) schedule(static)
        for (int k = 1; k <= c->Z; ++k)
            for (int j = 1; j <= c->Y; ++j)
                for (int i = 1; i <= c->X; ++i)
                    orc_compute_cru(c, i, j, k, reset, loc);
#pragma omp critical
        for (int q = 0; q < RYR_C_COUNT; ++q) c->counters[q] += loc[q];
    }
#else
    for (int k = 1; k <= c->Z; ++k)
        for (int j = 1; j <= c->Y; ++j)
            for (int i = 1; i <= c->X; ++i)
                orc_compute_cru(c, i, j, k, reset, c->counters);
#endif
    for (int k = 1; k <= c->Z; ++k)
        for (int j = 1; j <= c->Y; ++j)
            for (int i = 1; i <= c->X; ++i)
                orc_boundary(c, i, j, k);

    /* swap (:422-425) */
    double *t;
    t = c->ci; c->ci = c->cinext; c->cinext = t;
    t = c->cnsr; c->cnsr = c->cnsrnext; c->cnsrnext = t;
    t = c->cs; c->cs = c->csnext; c->csnext = t;

    /* whole-cell sums (:430-447) in the canonical order */
    double *sc = c->scratch;
    for (int k = 1; k <= c->Z; ++k)
        for (int j = 1; j <= c->Y; ++j)
            for (int i = 1; i <= c->X; ++i) sc[INTERIOR(c, i, j, k)] = c->cs[POS(c, i, j, k)];
    double sum_cs = orc_canonical_sum(sc, c->N);
    for (int k = 1; k <= c->Z; ++k)
        for (int j = 1; j <= c->Y; ++j)
            for (int i = 1; i <= c->X; ++i) sc[INTERIOR(c, i, j, k)] = c->xicat[POS(c, i, j, k)];
    double sum_icat = orc_canonical_sum(sc, c->N);
    for (int k = 1; k <= c->Z; ++k)
        for (int j = 1; j <= c->Y; ++j)
            for (int i = 1; i <= c->X; ++i) sc[INTERIOR(c, i, j, k)] = c->xinaca[POS(c, i, j, k)];
    double sum_inaca = orc_canonical_sum(sc, c->N);

    double st[13] = {c->v, c->xm, c->xh, c->xj, c->xr, c->xs1, c->xs2, c->qks, c->xtos, c->ytos,
                     c->xtof, c->ytof, c->xnai};
    double cur[10];
    c->stim = orc_stim(c, n);
    const double vold = c->v;
    orc_membrane(st, sum_icat, sum_inaca, sum_cs, (double)c->N, c->Cm, c->p.dt, c->svikr, c->stim, cur,
                 (c->p.variants & RYR_VAR_SLOWNCX) != 0);
    if (c->vclamp_n > 0)   /* clamp protocols (cuda_Cell_Detailed.cu:733-773) */
        st[0] = c->vclamp[c->vclamp_periodic ? n % c->vclamp_n : (n < c->vclamp_n ? n : c->vclamp_n - 1)];
    c->v = st[0]; c->xm = st[1]; c->xh = st[2]; c->xj = st[3]; c->xr = st[4]; c->xs1 = st[5];
    c->xs2 = st[6]; c->qks = st[7]; c->xtos = st[8]; c->ytos = st[9]; c->xtof = st[10];
    c->ytof = st[11]; c->xnai = st[12];
    c->csubt = cur[0]; c->xicatto = cur[1]; c->xinacato = cur[2]; c->xina = cur[3]; c->xiks = cur[4];
    c->xikr = cur[5]; c->xik1 = cur[6]; c->xinak = cur[7]; c->xitos = cur[8]; c->xitof = cur[9];
    if (!(c->v == c->v) || c->v == __builtin_inf() || c->v == -__builtin_inf())
        c->counters[RYR_C_NONFINITE_V]++;
    orc_beat_track(c, n, vold, c->v);

    /* interval accumulators (:774-776) */
    c->out_ica += c->xicatto;
    c->out_ncx += c->xinacato;
    c->out_ina += c->xina;
    if (n % spo == 0) {
        orc_diag(c);
        c->out_ica = 0; c->out_ncx = 0; c->out_ina = 0;   /* :1074-1077 */
    }
    c->step = n + 1;
}

void orc_step(orc_cell *c, int64_t nsteps)
{
    for (int64_t s = 0; s < nsteps; ++s) orc_one_step(c);
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_read_scalars(const orc_cell *c, double *o)
{
    o[RYR_S_T] = (double)(c->step > 0 ? c->step - 1 : 0) * c->p.dt;
    o[RYR_S_V] = c->v; o[RYR_S_XM] = c->xm; o[RYR_S_XH] = c->xh; o[RYR_S_XJ] = c->xj;
    o[RYR_S_XR] = c->xr; o[RYR_S_XS1] = c->xs1; o[RYR_S_XS2] = c->xs2; o[RYR_S_QKS] = c->qks;
    o[RYR_S_XTOS] = c->xtos; o[RYR_S_YTOS] = c->ytos; o[RYR_S_XTOF] = c->xtof;
    o[RYR_S_YTOF] = c->ytof; o[RYR_S_XNAI] = c->xnai; o[RYR_S_CSUBT] = c->csubt;
    o[RYR_S_ICAL] = c->xicatto; o[RYR_S_INCX] = c->xinacato; o[RYR_S_INA] = c->xina;
    o[RYR_S_IKS] = c->xiks; o[RYR_S_IKR] = c->xikr; o[RYR_S_IK1] = c->xik1;
    o[RYR_S_INAK] = c->xinak; o[RYR_S_ITOS] = c->xitos; o[RYR_S_ITOF] = c->xitof;
    o[RYR_S_STIM] = c->stim; o[RYR_S_OUT_ICA] = c->out_ica; o[RYR_S_OUT_NCX] = c->out_ncx;
    o[RYR_S_OUT_INA] = c->out_ina; o[RYR_S_STEP] = (double)c->step;
    return 0;
}

int orc_write_scalar(orc_cell *c, int s, double v)
{
    double *m[] = {&c->v, &c->xm, &c->xh, &c->xj, &c->xr, &c->xs1, &c->xs2, &c->qks,
                   &c->xtos, &c->ytos, &c->xtof, &c->ytof, &c->xnai};
    if (s < RYR_S_V || s > RYR_S_XNAI) return RYR_E_ARG;
    *m[s - RYR_S_V] = v;
    return 0;
}

/* field access in the GPU's SoA interior layout */
static int orc_field_ptr(const orc_cell *c, int f, double **d, int **iv, int *planes)
{
    *d = NULL; *iv = NULL; *planes = 1;
    switch (f) {
    case RYR_F_CI: *d = c->ci; *planes = 8; break;
    case RYR_F_CNSR: *d = c->cnsr; *planes = 8; break;
    case RYR_F_CATF: *d = c->catf; *planes = 8; break;
    case RYR_F_CATS: *d = c->cats; *planes = 8; break;
    case RYR_F_CACAL: *d = c->cacal; *planes = 8; break;
    case RYR_F_CASR: *d = c->casr; *planes = 8; break;
    case RYR_F_CAMYO: *d = c->camyo; *planes = 8; break;
    case RYR_F_MGMYO: *d = c->mgmyo; *planes = 8; break;
    case RYR_F_CS: *d = c->cs; break;
    case RYR_F_CP: *d = c->cp; break;
    case RYR_F_CJSR: *d = c->cjsr; break;
    case RYR_F_CJSRB: *d = c->cjsrb; break;
    case RYR_F_ANCX: *d = c->Ancx; break;
    case RYR_F_CASAR: *d = c->casar; break;
    case RYR_F_CASARH: *d = c->casarh; break;
    case RYR_F_CASARJ: *d = c->casarj; break;
    case RYR_F_CASARHJ: *d = c->casarhj; break;
    case RYR_F_NOU: *iv = c->nou; break;
    case RYR_F_NCU: *iv = c->ncu; break;
    case RYR_F_NOB: *iv = c->nob; break;
    case RYR_F_NCB: *iv = c->ncb; break;
    case RYR_F_LCC: *iv = c->lcc; *planes = 8; break;
    case RYR_F_NL: *iv = c->nl; break;
    case RYR_F_NSPARK: *iv = c->nspark; break;
    default: return RYR_E_ARG;
    }
    return 0;
}

int orc_read_field(const orc_cell *c, int f, double *out, int64_t n)
{
    double *d; int *iv; int planes;
    if (orc_field_ptr(c, f, &d, &iv, &planes)) return RYR_E_ARG;
    if (n < planes * c->N) return RYR_E_ARG;
    for (int k = 1; k <= c->Z; ++k)
        for (int j = 1; j <= c->Y; ++j)
            for (int i = 1; i <= c->X; ++i) {
                int64_t ac = POS(c, i, j, k), q = INTERIOR(c, i, j, k);
                for (int p = 0; p < planes; ++p) {
                    int64_t src = (planes == 8) ? ac * 8 + p : ac;
                    out[p * c->N + q] = d ? d[src] : (double)iv[src];
                }
            }
    return 0;
}

/* Scheme B's cp<->cs exchange uses the step-start cs on the cp side and the
 * sub-step's cs on the cs side (:1503 vs :1515), and its cs z-diffusion the
 * step-start neighbours against the sub-step's own cs (:1496), so it does not
 * conserve calcium: over a step the total changes by the sarcolemmal flux
 * plus (Vp/taups + 2 Vs/taust) dt/5 * sum over CRUs of the lag
 * sum_m (cs_start - cs_m) (m = sub-steps 0..4, cs_0 = cs_start).  This
 * returns each interior CRU's lag of the last scheme-B step, for the test of
 * that identity. */
int orc_read_scheme_b_lag(const orc_cell *c, double *out, int64_t n)
{
    if (n < c->N) return RYR_E_ARG;
    for (int k = 1; k <= c->Z; ++k)
        for (int j = 1; j <= c->Y; ++j)
            for (int i = 1; i <= c->X; ++i)
                out[INTERIOR(c, i, j, k)] = c->lagB[POS(c, i, j, k)];
    return 0;
}

/* Writes the interior and keeps every ghost equal to the adjacent interior
 * value it mirrors, as the reference's boundary copies would. */
int orc_write_field(orc_cell *c, int f, const double *in, int64_t n)
{
    double *d; int *iv; int planes;
    if (orc_field_ptr(c, f, &d, &iv, &planes)) return RYR_E_ARG;
    if (n < planes * c->N) return RYR_E_ARG;
    for (int k = 1; k <= c->Z; ++k)
        for (int j = 1; j <= c->Y; ++j)
            for (int i = 1; i <= c->X; ++i) {
                int64_t ac = POS(c, i, j, k), q = INTERIOR(c, i, j, k);
                for (int p = 0; p < planes; ++p) {
                    int64_t dst = (planes == 8) ? ac * 8 + p : ac;
                    if (d) d[dst] = in[p * c->N + q];
                    else iv[dst] = (int)in[p * c->N + q];
                }
            }
    if (f == RYR_F_CI || f == RYR_F_CNSR || f == RYR_F_CS) {
        /* refresh ghosts: run the boundary copy on the current buffers */
        double *save_ci = c->cinext, *save_cn = c->cnsrnext, *save_cs = c->csnext;
        c->cinext = c->ci; c->cnsrnext = c->cnsr; c->csnext = c->cs;
        for (int k = 1; k <= c->Z; ++k)
            for (int j = 1; j <= c->Y; ++j)
                for (int i = 1; i <= c->X; ++i) orc_boundary(c, i, j, k);
        c->cinext = save_ci; c->cnsrnext = save_cn; c->csnext = save_cs;
    }
    return 0;
}

int orc_read_diag(const orc_cell *c, double *out, int64_t *step)
{
    memcpy(out, c->diag, sizeof c->diag);
    *step = c->diag_step;
    return 0;
}

int orc_read_linescan(const orc_cell *c, double *out, int64_t n)
{
    if (n < (int64_t)c->X * RYR_LINE_COLS) return RYR_E_ARG;
    memcpy(out, c->line, sizeof(double) * (size_t)c->X * RYR_LINE_COLS);
    return 0;
}

int orc_read_counters(const orc_cell *c, int64_t *out)
{
    memcpy(out, c->counters, sizeof c->counters);
    return 0;
}

/* ---- probes of the shared primitives and of the stencil (tests) ---------- */
double orc_dm_exp(double x) { return ryr_exp(x); }
double orc_dm_log(double x) { return ryr_log(x); }
double orc_dm_cos(double x) { return ryr_cos(x); }
double orc_dm_pow(double x, double y) { return ryr_pow(x, y); }
void orc_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    ryr_u32x4 c;
    for (int q = 0; q < 4; ++q) c.v[q] = ctr[q];
    ryr_u32x4 o = ryr_philox4x32_10(c, key[0], key[1]);
    for (int q = 0; q < 4; ++q) out[q] = o.v[q];
}
double orc_slot_uniform(uint32_t a, uint32_t b) { return ryr_slot_uniform(a, b); }
/* the FP32 production mode's stochastic rounding (the oracle itself is FP64
 * and does not round; these expose the shared helpers to the tests) */
uint32_t orc_sr_bits(uint64_t seed, uint64_t step, uint32_t field, uint32_t n, uint32_t ii)
{
    return ryr_sr_bits(ryr_sr_key(seed, step, field), n, ii);
}
uint32_t orc_mix32(uint32_t x) { return ryr_mix32(x); }
float orc_sr_f32(double v, uint32_t r) { return ryr_sr_f32(v, r); }
void orc_stream(uint64_t seed, uint32_t cru, uint64_t step, int n, double *u, uint32_t *w)
{
    ryr_stream s;
    ryr_stream_init(&s, seed, cru, step);
    for (int q = 0; q < n; ++q) {
        uint32_t a, b;
        ryr_stream_next_slot(&s, &a, &b);
        u[q] = ryr_slot_uniform(a, b);
        w[q] = b;
    }
}

/* stencil of a whole field: planes in the SoA interior layout in, the
 * coupling term of every interior sub-voxel out (same layout) */
int orc_stencil_probe(const ryr_params *p, const ryr_cell_desc *d, const double *field,
                      double tt, double tl, double *out)
{
    orc_cell *c = orc_create(p, d);
    if (!c) return RYR_E_ARG;
    orc_write_field(c, RYR_F_CI, field, 8 * c->N);
    for (int k = 1; k <= c->Z; ++k)
        for (int j = 1; j <= c->Y; ++j)
            for (int i = 1; i <= c->X; ++i)
                for (int ii = 0; ii < 8; ++ii)
                    out[ii * c->N + INTERIOR(c, i, j, k)] = orc_coupling(c, c->ci, i, j, k, ii, tt, tl);
    orc_destroy(c);
    return 0;
}
