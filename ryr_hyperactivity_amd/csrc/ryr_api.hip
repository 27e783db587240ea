/*
 * ryr_api.hip - host side of the C ABI declared in include/ryr_cru.h.
 *
 * Replaces the reference's main() plumbing around Compute
 * (cuda_Cell_Detailed.cu:232-1105): set-up and ICs, the per-step launch,
 * the output-interval copies, teardown.  The per-step host work of the
 * reference (D2H of 516 KB, serial sums, membrane ODE) moved on device, so a
 * step here is one kernel launch; every `spo`-step interval (one output row)
 * is captured once as a HIP graph and replayed.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include "ryr_cru.h"
#define RYR_DEFINE_COMMON_KERNELS
#include "cru_launch.h"

struct ryr_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    ryr_params p{};
    std::vector<ryr_cell_desc> desc;
    std::vector<RyrCellInfo> info;
    std::vector<long long> line_off;
    int ncells = 0, nblocks = 0;
    long long P = 0;
    long long step = 0;
    RyrKArgs a{};
    std::vector<std::pair<void *, size_t>> allocs;   /* simulation state, in snapshot order */
    std::vector<void *> tables;     /* layout tables and scratch: rebuilt, never saved */
    long long gbase = 0;            /* value of *a.gstep as the enqueued work will see it; -1 unknown */
    double *stage = nullptr;
    size_t stage_n = 0;
    /* captured step segments: `n` steps that start on an output step (out0)
     * or hold none, replayed by ryr_step; `seen` counts the uses of a
     * segment that has no graph yet */
    struct Seg { bool out0; long long n; hipGraphExec_t ge; int seen; };
    std::vector<Seg> segs;
    bool graph_failed = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t kev[8] = {};         /* per-kernel start/end stamps: cyto, cru, membrane, diag */
    int wg_cyto = 0, wg_cru = 0;    /* workgroup sizes of the two CRU-parallel kernels; 0: tune */
    int sched = -1;                 /* step schedule RYR_SCHED_*; -1: tune */
    hipStream_t stream2 = nullptr;  /* second stream of the concurrent schedule */
    hipEvent_t fork = nullptr, join = nullptr;
    int tune_i = 0;
    std::vector<float> tune_ms[2][3];
    std::vector<std::array<int, 3>> cand;   /* schedule tuning: candidates and their times */
    std::vector<float> cand_ms;
    size_t cand_i = 0;
    bool timing_open = false;
    long long launches = 0;
    std::string err;
    std::vector<double *> vclamp;   /* per cell clamp waveform (device) */
    bool null_launch = false;       /* enqueue empty kernels of the same shapes (ryr_latency_floor) */
};

static std::string g_create_err;

#define RYR_HIPCHECK(h, call)                                                       \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);           \
            return RYR_E_HIP;                                                       \
        }                                                                           \
    } while (0)

extern "C" {

int ryr_abi_version(void) { return RYR_ABI_VERSION; }

void ryr_params_default(ryr_params *p)
{
    std::memset(p, 0, sizeof *p);
    p->dt = 0.025;
    p->scheme = RYR_SCHEME_A;
    p->precision = RYR_PREC_F64;
    p->iso2 = 1;
    p->tapsi = 0;
    p->lqt = 1;
    p->nlcc = 4;
    p->steps_per_output = 100;
    p->stim_amp = 80.0;
    p->stim_dur = 1.0;
    p->stim_offset = 100.0;
    p->ci0 = 0.0944;
    p->v0 = -80.0;
    p->cm_mode = 0;
    p->cm = 45.0;
    p->variants = 0;
}

void ryr_cell_desc_default(ryr_cell_desc *d)
{
    std::memset(d, 0, sizeof *d);
    d->nx = 62; d->ny = 26; d->nz = 10;
    d->seed = 1;
    d->cl = 4000.0;
    d->cjsr0 = 740.0;
    d->ncx_scale = 1.0;
    d->Ku = 10.0;
    d->Kb = 1.0;
    d->ica_scale = 0.7;
    d->Ku2 = 1.0;
    d->ikr_scale = 1.0;
}

const char *ryr_last_error(const ryr_handle *h)
{
    return h ? h->err.c_str() : g_create_err.c_str();
}

}  /* extern "C" */

/* --- host restatement of Init's closed forms (cuda_Cell_Detailed.cu:1118-1174,
 * subroutine_cuda.c:179-195); contraction off in this unit */
static double host_luminal_SS(double cjsr)
{
    double roo2 = 5000.0 / (1.0 + ryr_pow(850.0 / cjsr, 23.0));
    double mono = (-1.0 + std::sqrt(1.0 + 8.0 * 460.0 * roo2)) / (4.0 * roo2 * 460.0);
    double ene = mono * 20 + (1.0 - mono) * 30;
    return 460.0 * ene * cjsr / (600.0 + cjsr);
}
static double host_fracbound(double cjsr_b)
{
    double roo2 = 5000.0 / (1.0 + ryr_pow(850.0 / (cjsr_b), 23.0));
    double aub = (-1.0 + std::sqrt(1.0 + 8.0 * 460.0 * roo2)) / (4.0 * roo2 * 460.0) / (1.0);
    double abu = 1.0 / (0.2 * 5 * 220.);
    return 1 / (1 + abu / aub);
}

static int dev_alloc(ryr_handle *h, void **ptr, size_t bytes, bool state)
{
    RYR_HIPCHECK(h, hipMalloc(ptr, bytes > 0 ? bytes : 16));
    RYR_HIPCHECK(h, hipMemset(*ptr, 0, bytes > 0 ? bytes : 16));
    if (state) h->allocs.push_back({*ptr, bytes});
    else h->tables.push_back(*ptr);
    return RYR_OK;
}

/* constants of ryr_fastdiv for the divisor d >= 1 */
static void ryr_fastdiv_consts(uint32_t d, uint32_t *m, uint32_t *s1, uint32_t *s2)
{
    if (d <= 1) { *m = 0; *s1 = 0; *s2 = 0; return; }
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    *m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
    *s1 = 1;
    *s2 = l - 1;
}

/* the per-cell table and its per-chunk copies */
static int upload_info(ryr_handle *h)
{
    std::vector<RyrCellInfo> ch(h->nblocks);
    for (int c = 0; c < h->ncells; ++c)
        for (int b = 0; b < h->info[c].nblocks; ++b) ch[h->info[c].first_block + b] = h->info[c];
    RYR_HIPCHECK(h, hipMemcpy((void *)h->a.cells, h->info.data(), sizeof(RyrCellInfo) * h->ncells, hipMemcpyHostToDevice));
    RYR_HIPCHECK(h, hipMemcpy((void *)h->a.chunks, ch.data(), sizeof(RyrCellInfo) * h->nblocks, hipMemcpyHostToDevice));
    return RYR_OK;
}

static int validate(const ryr_params *p, const ryr_cell_desc *cells, int ncells, std::string &msg)
{
    if (!p || !cells || ncells <= 0) { msg = "null parameters or ncells <= 0"; return RYR_E_ARG; }
    if (!(p->dt > 0.0) || p->dt > 1.0) { msg = "dt must be in (0, 1] ms"; return RYR_E_ARG; }
    if (p->scheme != RYR_SCHEME_A && p->scheme != RYR_SCHEME_B) { msg = "scheme must be A (0) or B (1)"; return RYR_E_ARG; }
    if (p->precision != RYR_PREC_F64 && p->precision != RYR_PREC_F32) { msg = "precision must be 0 or 1"; return RYR_E_ARG; }
    if (p->nlcc < 0 || p->nlcc > 8) { msg = "nlcc must be in [0, 8]"; return RYR_E_ARG; }
    if (p->steps_per_output < 1) { msg = "steps_per_output must be >= 1"; return RYR_E_ARG; }
    if (!(p->cm > 0.0)) { msg = "cm must be > 0"; return RYR_E_ARG; }
    if (p->variants & ~(RYR_VAR_PERM | RYR_VAR_ISO | RYR_VAR_SLOWNCX | RYR_VAR_RANDOMLCC)) {
        msg = "unknown bits in variants";
        return RYR_E_ARG;
    }
    for (int c = 0; c < ncells; ++c) {
        const ryr_cell_desc &d = cells[c];
        if (d.nx < 1 || d.ny < 1 || d.nz < 1) { msg = "cell " + std::to_string(c) + ": lattice dims must be >= 1"; return RYR_E_ARG; }
        long long N = (long long)d.nx * d.ny * d.nz;
        if (N > (1LL << 30)) { msg = "cell " + std::to_string(c) + ": lattice too large"; return RYR_E_ARG; }
        if (!(d.cl > 0.0)) { msg = "cell " + std::to_string(c) + ": cl must be > 0"; return RYR_E_ARG; }
        if (!(d.cjsr0 > 0.0) || !(d.Ku > 0.0)) { msg = "cell " + std::to_string(c) + ": cjsr0 and Ku must be > 0"; return RYR_E_ARG; }
        double r = d.cl / p->dt;
        if (std::fabs(r - std::llround(r)) > 1e-6 * r) { msg = "cell " + std::to_string(c) + ": cl must be a multiple of dt"; return RYR_E_ARG; }
    }
    return RYR_OK;
}

extern "C" int ryr_create(int device, const ryr_params *p, const ryr_cell_desc *cells, int ncells,
                          ryr_handle **out)
{
    if (!out) { g_create_err = "out is null"; return RYR_E_ARG; }
    *out = nullptr;
    std::string msg;
    int rc = validate(p, cells, ncells, msg);
    if (rc) { g_create_err = msg; return rc; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_create_err = "no HIP device available";
        return RYR_E_NODEV;
    }
    if (device < 0 || device >= ndev) { g_create_err = "device index out of range"; return RYR_E_ARG; }

    ryr_handle *h = new ryr_handle();
    h->device = device;
    h->p = *p;
    h->desc.assign(cells, cells + ncells);
    h->ncells = ncells;
    auto fail = [&](int code) { g_create_err = h->err; ryr_destroy(h); return code; };
    if (hipSetDevice(device) != hipSuccess) { h->err = "hipSetDevice failed"; return fail(RYR_E_HIP); }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        h->err = "hipStreamCreate failed";
        return fail(RYR_E_HIP);
    }
    (void)hipEventCreate(&h->ev0);
    (void)hipEventCreate(&h->ev1);
    for (auto &e : h->kev) (void)hipEventCreate(&e);
    if (hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->join, hipEventDisableTiming) != hipSuccess) {
        h->err = "hipStreamCreate/hipEventCreate failed";
        return fail(RYR_E_HIP);
    }

    /* layout: each cell padded to whole workgroups */
    std::vector<int> block_cell;
    long long lo = 0;
    h->info.resize(ncells);
    h->line_off.resize(ncells);
    for (int c = 0; c < ncells; ++c) {
        const ryr_cell_desc &d = cells[c];
        RyrCellInfo &I = h->info[c];
        std::memset(&I, 0, sizeof I);
        I.X = d.nx; I.Y = d.ny; I.Z = d.nz;
        I.N = d.nx * d.ny * d.nz;
        I.nblocks = (I.N + RYR_BLOCK - 1) / RYR_BLOCK;
        I.first_block = (int)block_cell.size();
        I.off = (long long)I.first_block * RYR_BLOCK;
        for (int b = 0; b < I.nblocks; ++b) block_cell.push_back(c);
        I.seed = d.seed;
        I.Ku = d.Ku; I.Ku2 = d.Ku2; I.Kb = d.Kb;
        I.svica = ((0.666) * (1.2) * (d.ica_scale));
        I.svncx = (2.5 * d.ncx_scale);
        I.svikr = p->lqt ? 0.0 : 3.0 * d.ikr_scale;
        double sviup = (1.0 + 0.75);
        if (p->iso2) sviup = (1.0 + 0.75) * 1.75;
        if (p->tapsi) sviup = 0;
        double sviup_lqt = (p->lqt && p->tapsi) ? 0.05 : 0.;
        I.sviup_tot = sviup + sviup_lqt;
        I.Cm = (p->cm_mode == 0) ? p->cm * (double)I.N / 16120.0 : p->cm;
        I.dN = (double)I.N;
        I.ncl = (long long)std::llround(d.cl / p->dt);
        I.noff = (long long)std::llround(p->stim_offset / p->dt);
        I.ndur = (long long)std::llround(p->stim_dur / p->dt);
        I.nlcc = p->nlcc;
        /* variants (the reference's -D builds) */
        I.variants = p->variants;
        I.nstim_end = -1;
        I.iso_n = (11 * I.ncl + 1) / 2;                 /* t >= 5.5 CL (:1220) */
        I.sviup_iso = I.sviup_tot;
        if (p->variants & RYR_VAR_ISO) {
            I.noff = 0;                                 /* fmod(t, CL) < 1 (:684) */
            I.nstim_end = (119 * I.ncl + 9) / 10;       /* t < 11.9 CL */
            if (!p->iso2 && !p->tapsi) I.sviup_iso = (1.0 + 0.75) * 2. + sviup_lqt;
        }
        if (p->variants & RYR_VAR_PERM) I.nstim_end = 0;    /* stim = 0 (:697-699) */
        I.ci_bath = p->ci0;
        int jy = (d.ny + 2) / 2;            /* ny/2 with the ghost shell (:973) */
        int jz = 5 < d.nz ? 5 : d.nz;        /* :972, clamped to the interior */
        I.line_j = jy - 1;
        I.line_k = jz - 1;
        I.cell = c;
        ryr_fastdiv_consts((uint32_t)d.nx, &I.dX_m, &I.dX_s1, &I.dX_s2);
        ryr_fastdiv_consts((uint32_t)(d.nx * d.ny), &I.dXY_m, &I.dXY_s1, &I.dXY_s2);
        h->line_off[c] = lo;
        lo += (long long)d.nx * RYR_LINE_COLS;
    }
    h->nblocks = (int)block_cell.size();
    h->P = (long long)h->nblocks * RYR_BLOCK;
    const long long P = h->P;
    const bool f32 = p->precision == RYR_PREC_F32;
    const size_t tf = f32 ? 4 : 8, ts = tf;

    RyrKArgs &a = h->a;
    a.P = P;
    a.dt = p->dt;
    a.scheme = p->scheme;
    a.iso2 = p->iso2;
    a.spo = p->steps_per_output;
    a.ncells = ncells;
#define ALLOCX(ptr, bytes, st) do { void *q_ = nullptr; if (dev_alloc(h, &q_, (bytes), st)) return fail(RYR_E_HIP); (ptr) = decltype(ptr)(q_); } while (0)
#define ALLOC(ptr, bytes) ALLOCX(ptr, bytes, true)
#define ALLOCT(ptr, bytes) ALLOCX(ptr, bytes, false)
    for (int b = 0; b < 2; ++b) {
        ALLOC(a.ci[b], 8 * P * tf);
        ALLOC(a.cnsr[b], 8 * P * ts);
        ALLOC(a.cs[b], P * tf);
        ALLOC(a.cp[b], P * tf);
        ALLOC(a.cjsr[b], P * ts);
    }
    ALLOC(a.catf, 8 * P * tf); ALLOC(a.cats, 8 * P * ts); ALLOC(a.cacal, 8 * P * tf);
    ALLOC(a.casr, 8 * P * tf); ALLOC(a.camyo, 8 * P * tf); ALLOC(a.mgmyo, 8 * P * tf);
    ALLOC(a.cjsrb, P * ts); ALLOC(a.Ancx, P * tf);
    ALLOC(a.casar, P * tf); ALLOC(a.casarh, P * tf); ALLOC(a.casarj, P * tf); ALLOC(a.casarhj, P * tf);
    ALLOC(a.lccw, P * 4); ALLOC(a.ryrw, P * 4); ALLOC(a.aux, P * 4);
    ALLOC(a.d_xire, P * 8); ALLOC(a.d_xinaca, P * 8); ALLOCT(a.d_up, 8 * P * 8); ALLOCT(a.d_leak, 8 * P * 8);
    ALLOCT(a.sb, 2 * P * tf);
    RyrCellInfo *d_cells, *d_chunks; RyrMem *d_mem; int *d_bc; RyrDD *d_part;
    unsigned long long *d_cnt; double *d_diag; long long *d_dstep; double *d_line; long long *d_loff;
    ALLOC(d_mem, sizeof(RyrMem) * ncells);
    ALLOC(d_cnt, sizeof(unsigned long long) * ncells * RYR_NCNT);
    ALLOC(d_diag, sizeof(double) * ncells * RYR_DIAG_COLS);
    ALLOC(d_dstep, sizeof(long long) * ncells);
    ALLOC(d_line, sizeof(double) * (lo > 0 ? lo : 1));
    ALLOCT(d_cells, sizeof(RyrCellInfo) * ncells);
    ALLOCT(d_chunks, sizeof(RyrCellInfo) * h->nblocks);
    ALLOCT(d_bc, sizeof(int) * h->nblocks);
    ALLOCT(d_part, sizeof(RyrDD) * 2 * 12 * h->nblocks);
    ALLOCT(a.pend, sizeof(RyrPend) * 2 * ncells);
    ALLOCT(a.vbox, sizeof(RyrVBox) * ncells);
    ALLOCT(d_loff, sizeof(long long) * ncells);
    ALLOCT(a.gstep, sizeof(long long));
    ALLOC(a.beats, sizeof(double) * ncells * RYR_MAXB * RYR_BCOLS);
    ALLOCT(a.dpartial, sizeof(RyrDD) * RYR_DIAG_Q * h->nblocks);
    ALLOCT(a.dcount, sizeof(unsigned long long) * RYR_DIAG_I * h->nblocks);
#undef ALLOC
#undef ALLOCT
#undef ALLOCX
    h->vclamp.assign(ncells, nullptr);
    a.cells = d_cells; a.chunks = d_chunks; a.mem = d_mem; a.block_cell = d_bc; a.partial = d_part;
    a.counters = d_cnt; a.diag = d_diag; a.diag_step = d_dstep; a.line = d_line; a.line_off = d_loff;

    auto H2D = [&](void *dst, const void *src, size_t n) {
        return hipMemcpy(dst, src, n, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (upload_info(h) != RYR_OK ||
        !H2D(d_bc, block_cell.data(), sizeof(int) * h->nblocks) ||
        !H2D(d_loff, h->line_off.data(), sizeof(long long) * ncells)) {
        h->err = "hipMemcpy of cell tables failed";
        return fail(RYR_E_HIP);
    }
    if (hipMemset(d_dstep, 0xff, sizeof(long long) * ncells) != hipSuccess) {
        h->err = "hipMemset failed";
        return fail(RYR_E_HIP);
    }

    /* initial values (cuda_Cell_Detailed.cu:1108-1182, :350-366) */
    std::vector<RyrInit> init(ncells);
    std::vector<RyrMem> mem(ncells);
    for (int c = 0; c < ncells; ++c) {
        const double ci_basal = p->ci0;
        const double cjsr_b = p->tapsi ? 10 : cells[c].cjsr0;
        RyrInit &I = init[c];
        std::memset(&I, 0, sizeof I);
        I.ci0 = ci_basal;
        I.cjsr0 = cjsr_b;
        I.cjsrb0 = host_luminal_SS(cjsr_b);
        I.catf0 = 0.0327 * ci_basal * 70. / (0.0327 * ci_basal + 0.0196);
        I.cats0 = 0.00254 * ci_basal * ((134.)) / (0.00254 * ci_basal + 0.000033);
        I.cacal0 = 0.0543 * ci_basal * 24. / (0.0543 * ci_basal + 0.238);
        I.casr0 = 0.1 * ci_basal * 19. / (0.1 * ci_basal + 0.06);
        I.camyo0 = 3.;
        I.mgmyo0 = 136.209996;
        I.casar0 = 0.1 * ci_basal * RYR_BSAR / (0.1 * ci_basal + 1.3);
        I.casarh0 = 0.1 * ci_basal * RYR_BSARH / (0.1 * ci_basal + 0.03);
        I.ancx0 = 0.025;
        double fb = host_fracbound(cjsr_b);
        I.ncb0 = (int)(fb * RYR_NRYR);
        I.ncu0 = RYR_NRYR - (int)(fb * RYR_NRYR);
        uint32_t lw = 0;
        for (int ll = 0; ll < 8; ++ll) lw |= (uint32_t)(ll < p->nlcc ? 3 : 0) << (4 * ll);
        I.lcc0 = lw;
        RyrMem &M = mem[c];
        std::memset(&M, 0, sizeof M);
        M.m[0] = p->v0; M.m[1] = 0.0010; M.m[2] = 1.00; M.m[3] = 1.00; M.m[4] = 0.00;
        M.m[5] = 0.08433669901; M.m[6] = M.m[5]; M.m[7] = 0.20; M.m[8] = 0.010; M.m[9] = 1.00;
        M.m[10] = 0.020; M.m[11] = 0.80;
        M.m[12] = 6.1 * ((9000) / (cells[c].cl + 5000));
        M.step = 0;
        M.beat_on = -1;
    }
    RyrInit *d_init = nullptr;
    if (hipMalloc((void **)&d_init, sizeof(RyrInit) * ncells) != hipSuccess ||
        !H2D(d_init, init.data(), sizeof(RyrInit) * ncells) ||
        !H2D(d_mem, mem.data(), sizeof(RyrMem) * ncells)) {
        if (d_init) hipFree(d_init);
        h->err = "upload of initial values failed";
        return fail(RYR_E_HIP);
    }
    hipError_t e = f32 ? ryr_launch_init_f32(a, d_init, h->nblocks, h->stream)
                       : ryr_launch_init_f64(a, d_init, h->nblocks, h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_vfactors, dim3((ncells + 63) / 64), dim3(64), 0, h->stream, d_mem, ncells, p->iso2);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(d_init);
    if (e != hipSuccess) {
        h->err = std::string("initialisation kernels failed: ") + hipGetErrorString(e);
        return fail(RYR_E_HIP);
    }
    *out = h;
    return RYR_OK;
}

/* ---- step sequencing
 * Kernels take their step as an offset k from a base held on the device
 * (*a.gstep), so one captured output interval replays with unchanged
 * arguments.  h->gbase mirrors that base for the work already enqueued. */
static hipError_t set_base(ryr_handle *h, long long v)
{
    hipLaunchKernelGGL(k_stepbase, dim3(1), dim3(64), 0, h->stream, h->a.gstep, v, 0);
    h->gbase = v;
    return hipGetLastError();
}

/* offsets must stay small and non-negative */
static hipError_t ensure_base(ryr_handle *h, long long nahead)
{
    if (h->gbase < 0 || h->step < h->gbase || h->step + nahead - h->gbase > (1LL << 30))
        return set_base(h, h->step);
    return hipSuccess;
}

/* Empty kernel with the grid and workgroup of a step kernel: the launch
 * floor of a step's graph (ryr_latency_floor) */
__global__ void k_null(int) {}

static hipError_t go_null(ryr_handle *h, unsigned grid, unsigned wg, hipStream_t st = nullptr)
{
    hipLaunchKernelGGL(k_null, dim3(grid), dim3(wg), 0, st ? st : h->stream, 0);
    return hipGetLastError();
}

#define RYR_F32(h) ((h)->p.precision == RYR_PREC_F32)
static hipError_t go_cyto(ryr_handle *h, int wg, int kc, int km, hipEvent_t *ev, bool gate = false)
{
    if (h->null_launch) {
        if (wg == 1024) return go_null(h, (unsigned)h->nblocks, RYR_BLOCK);
        const unsigned per = (unsigned)(RYR_BLOCK / wg) * (unsigned)h->nblocks;
        return go_null(h, (km >= 0 ? (unsigned)h->ncells : 0u) + per * (gate ? 8u + RYR_GATE_LPC : 8u), wg);
    }
    return RYR_F32(h) ? ryr_launch_cyto_f32(h->a, h->nblocks, wg, kc, km, gate, h->p.stim_amp, h->stream, ev)
                      : ryr_launch_cyto_f64(h->a, h->nblocks, wg, kc, km, gate, h->p.stim_amp, h->stream, ev);
}
static hipError_t go_cru(ryr_handle *h, int wg, int kc, hipEvent_t *ev, hipStream_t st = nullptr, bool split = false)
{
    if (!st) st = h->stream;
    if (h->null_launch) {
        const unsigned nb = (unsigned)h->nblocks;
        return wg == 64 ? go_null(h, 4 * nb, 64, st) : wg == 128 ? go_null(h, 2 * nb, 128, st)
             : wg == 256 ? go_null(h, nb, 256, st) : go_null(h, 2 * nb, 256, st);
    }
    return RYR_F32(h) ? ryr_launch_cru_f32(h->a, h->nblocks, wg, kc, split, st, ev)
                      : ryr_launch_cru_f64(h->a, h->nblocks, wg, kc, split, st, ev);
}
static hipError_t go_mem(ryr_handle *h, int km, hipEvent_t *ev)
{
    if (h->null_launch) return go_null(h, (unsigned)h->ncells, RYR_BLOCK);
    return RYR_F32(h) ? ryr_launch_mem_f32(h->a, h->ncells, km, h->p.stim_amp, h->stream, ev)
                      : ryr_launch_mem_f64(h->a, h->ncells, km, h->p.stim_amp, h->stream, ev);
}
/* one-launch schedule: the launch of step gstep+k (next: the membrane step of
 * step gstep+k-1 is pending) */
static hipError_t go_one(ryr_handle *h, int k, bool next, hipEvent_t *ev)
{
    const int lpc = h->wg_cru == 512 ? 2 : 1;      /* 512: two lanes per CRU */
    if (h->null_launch)
        return go_null(h, (unsigned)one_head(h->ncells, lpc * h->nblocks) + 8u * (unsigned)h->nblocks, RYR_BLOCK);
    return RYR_F32(h) ? ryr_launch_one_f32(h->a, h->nblocks, lpc, k, next, h->p.stim_amp, h->stream, ev)
                      : ryr_launch_one_f64(h->a, h->nblocks, lpc, k, next, h->p.stim_amp, h->stream, ev);
}

static hipError_t go_diag(ryr_handle *h, hipEvent_t *ev)
{
    if (h->null_launch) {
        hipError_t e = go_null(h, (unsigned)h->nblocks, RYR_BLOCK);
        return e == hipSuccess ? go_null(h, (unsigned)h->ncells, RYR_BLOCK) : e;
    }
    return RYR_F32(h) ? ryr_launch_diag_f32(h->a, h->nblocks, h->stream, ev)
                      : ryr_launch_diag_f64(h->a, h->nblocks, h->stream, ev);
}

/* Steps h->step .. h->step+n-1 (h->step itself is not advanced here), in
 * one of four schedules (every schedule computes identical values):
 *  RYR_SCHED_SEQ: sub-voxel kernel, CRU kernel, membrane step;
 *  RYR_SCHED_FUSED: the sub-voxel kernel of step m carries the membrane step
 *    of m-1; the membrane step runs on its own only at an output step (the
 *    output-row kernel follows it) and at the last step of the run, so every
 *    run ends with the whole state advanced;
 *  RYR_SCHED_FUSED_GATING: as FUSED, and the sub-voxel launch of step m also
 *    carries the RyR gating of step m, which reads only step-start values;
 *    the CRU kernel of m then runs without it;
 *  RYR_SCHED_CONCURRENT (cleft scheme A): the sub-voxel and CRU kernels of a
 *    step read only step-start values of each other's fields (cp, cs, cjsr
 *    and ci/cnsr are double-buffered) and write disjoint fields, so the CRU
 *    kernel runs on the second stream beside the memory-bound sub-voxel
 *    sweep; the membrane step waits for both. */
static hipError_t enqueue_run(ryr_handle *h, long long n, int sched)
{
    const long long spo = h->p.steps_per_output;
    bool pending = false;
    hipError_t e = hipSuccess;
    for (long long i = 0; i < n && e == hipSuccess; ++i) {
        const long long m = h->step + i;
        const int k = (int)(m - h->gbase);
        const bool out = m % spo == 0;
        if (sched == RYR_SCHED_ONE) {
            e = go_one(h, k, pending, nullptr);
            if (e == hipSuccess && (out || i == n - 1)) {
                e = go_mem(h, k, nullptr);
                if (e == hipSuccess && out) e = go_diag(h, nullptr);
                pending = false;
            } else {
                pending = true;
            }
            continue;
        }
        if (sched == RYR_SCHED_CONCURRENT) {
            e = hipEventRecord(h->fork, h->stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(h->stream2, h->fork, 0);
            if (e == hipSuccess) e = go_cru(h, h->wg_cru, k, nullptr, h->stream2);
            if (e == hipSuccess) e = hipEventRecord(h->join, h->stream2);
            if (e == hipSuccess) e = go_cyto(h, h->wg_cyto, k, -1, nullptr);
            if (e == hipSuccess) e = hipStreamWaitEvent(h->stream, h->join, 0);
            if (e == hipSuccess) e = go_mem(h, k, nullptr);
            if (e == hipSuccess && out) e = go_diag(h, nullptr);
            continue;
        }
        const bool gate = sched == RYR_SCHED_FUSED_GATING;
        const bool fuse = sched == RYR_SCHED_FUSED || gate;
        e = go_cyto(h, h->wg_cyto, k, pending ? k - 1 : -1, nullptr, gate);
        if (e == hipSuccess) e = go_cru(h, h->wg_cru, k, nullptr, nullptr, gate);
        if (e == hipSuccess && (out || i == n - 1 || !fuse)) {
            e = go_mem(h, k, nullptr);
            if (e == hipSuccess && out) e = go_diag(h, nullptr);
            pending = false;
        } else {
            pending = true;
        }
    }
    return e;
}

/* One step with every kernel on its own and stamped by its own events
 * (kev[0..1] sub-voxel, [2..3] CRU, [4..5] membrane, [6..7] output row);
 * returns whether it was an output step. */
static hipError_t enqueue_timed_step(ryr_handle *h, int wg_cyto, int wg_cru, bool *out)
{
    const int k = (int)(h->step - h->gbase);
    *out = h->step % h->p.steps_per_output == 0;
    /* under the gating schedule the sub-voxel launch carries the gating and
     * the CRU kernel runs without it, as in production */
    const bool gate = h->sched == RYR_SCHED_FUSED_GATING;
    hipError_t e = go_cyto(h, wg_cyto, k, -1, h->kev, gate);
    if (e == hipSuccess) e = go_cru(h, wg_cru, k, h->kev + 2, nullptr, gate);
    if (e == hipSuccess) e = go_mem(h, k, h->kev + 4);
    if (e == hipSuccess && *out) e = go_diag(h, h->kev + 6);
    return e;
}

/* One step of the production schedule with every launch stamped by its own
 * events, for ryr_profile_kernels: the concurrent schedule runs the CRU
 * kernel beside the sub-voxel kernel on the second stream; the fused
 * schedules carry the previous step's membrane step (*pending) in the
 * sub-voxel launch and run it on its own only at an output step and at the
 * last step (`last`), so each kernel's stamp is its duration as production
 * runs it.  *ran[k] tells which of the four stamps were recorded; *tail is
 * the event that ends the step on the main stream. */
static hipError_t enqueue_profiled_step(ryr_handle *h, bool last, bool *pending, bool ran[4], hipEvent_t *tail)
{
    const int k = (int)(h->step - h->gbase);
    const bool out = h->step % h->p.steps_per_output == 0;
    const int sc = h->sched;
    for (int q = 0; q < 4; ++q) ran[q] = false;
    hipError_t e = hipSuccess;
    if (sc == RYR_SCHED_ONE) {
        /* the one launch carries the sweep, the CRU update and the membrane
         * workgroups: stamped as the sub-voxel launch */
        e = go_one(h, k, *pending, h->kev);
        ran[0] = true;
        *tail = h->kev[1];
        if (e == hipSuccess && (out || last)) {
            e = go_mem(h, k, h->kev + 4);
            ran[2] = true;
            *tail = h->kev[5];
            *pending = false;
        } else {
            *pending = true;
        }
    } else if (sc == RYR_SCHED_CONCURRENT) {
        e = hipEventRecord(h->fork, h->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(h->stream2, h->fork, 0);
        if (e == hipSuccess) e = go_cru(h, h->wg_cru, k, h->kev + 2, h->stream2);
        if (e == hipSuccess) e = hipEventRecord(h->join, h->stream2);
        if (e == hipSuccess) e = go_cyto(h, h->wg_cyto, k, -1, h->kev);
        if (e == hipSuccess) e = hipStreamWaitEvent(h->stream, h->join, 0);
        if (e == hipSuccess) e = go_mem(h, k, h->kev + 4);
        ran[0] = ran[1] = ran[2] = true;
        *tail = h->kev[5];
    } else {
        const bool gate = sc == RYR_SCHED_FUSED_GATING;
        const bool fuse = sc == RYR_SCHED_FUSED || gate;
        e = go_cyto(h, h->wg_cyto, k, *pending ? k - 1 : -1, h->kev, gate);
        if (e == hipSuccess) e = go_cru(h, h->wg_cru, k, h->kev + 2, nullptr, gate);
        ran[0] = ran[1] = true;
        *tail = h->kev[3];
        if (e == hipSuccess && (out || last || !fuse)) {
            e = go_mem(h, k, h->kev + 4);
            ran[2] = true;
            *tail = h->kev[5];
            *pending = false;
        } else {
            *pending = true;
        }
    }
    if (e == hipSuccess && out) {
        e = go_diag(h, h->kev + 6);
        ran[3] = true;
        *tail = h->kev[7];
    }
    return e;
}

static void drop_graph(ryr_handle *h)
{
    for (auto &q : h->segs) if (q.ge) hipGraphExecDestroy(q.ge);
    h->segs.clear();
    h->graph_failed = false;
}

/* Shape tuning: the first steps of a handle alternate the three workgroup
 * sizes, each step's kernels stamped by their own events; after
 * RYR_TUNE_ROUNDS rounds each kernel keeps the size with the smallest median
 * duration, 256 unless another is faster by more than 3%.  The tuning steps
 * are ordinary steps of the simulation (every shape computes identical
 * values), so tuning changes no result. */
#define RYR_TUNE_ROUNDS 3
static int tune_step(ryr_handle *h)
{
    const int v = h->tune_i % 3;
    const int wg = 64 << v;
    bool out = false;
    RYR_HIPCHECK(h, ensure_base(h, 1));
    RYR_HIPCHECK(h, enqueue_timed_step(h, wg, wg, &out));
    RYR_HIPCHECK(h, hipEventSynchronize(out ? h->kev[7] : h->kev[5]));
    for (int k = 0; k < 2; ++k) {
        float ms = 0.0f;
        RYR_HIPCHECK(h, hipEventElapsedTime(&ms, h->kev[2 * k], h->kev[2 * k + 1]));
        h->tune_ms[k][v].push_back(ms);
    }
    if (++h->tune_i == 3 * RYR_TUNE_ROUNDS) {
        int best[2] = {256, 256};
        for (int k = 0; k < 2; ++k) {
            float med[3];
            for (int q = 0; q < 3; ++q) {
                std::vector<float> t = h->tune_ms[k][q];
                std::sort(t.begin(), t.end());
                med[q] = t[t.size() / 2];
                h->tune_ms[k][q].clear();
            }
            float bm = med[2];
            for (int q = 1; q >= 0; --q)
                if (med[q] < 0.97f * bm) { bm = med[q]; best[k] = 64 << q; }
        }
        h->wg_cyto = best[0];
        h->wg_cru = best[1];
        drop_graph(h);
    }
    return RYR_OK;
}

/* Schedule tuning, after the shapes.  Candidates: the tuned sizes with each
 * schedule, plus (scheme A) the concurrent schedule with smaller sizes (two
 * kernels sharing the chip favour smaller workgroups).  Each candidate is
 * timed the way production runs, as a captured graph: RYR_TUNE_GSTEPS steps
 * that hold no output step, replayed three times (the first untimed), events
 * around each replay.  The sequential candidate is kept unless another is
 * faster by more than 3%.  Each candidate advances 3 * RYR_TUNE_GSTEPS
 * steps; candidates run while a call has that many steps left. */
/* 8 steps: in a shorter graph the run's first launch and last membrane step
 * (one-launch schedule) weigh enough to hide a 7% difference in the steps
 * between them */
#define RYR_TUNE_GSTEPS 8
/* whether a shape can run: the concurrent schedule needs cleft scheme A,
 * the gating schedule a CRU size of 128, 256 or 512 */
static bool shape_ok(const ryr_handle *h, int wg_cyto, int wg_cru, int sched)
{
    /* the one-lane-per-CRU sweep (1024) runs as a launch of its own */
    if (wg_cyto == 1024 && sched != RYR_SCHED_SEQ && sched != RYR_SCHED_CONCURRENT) return false;
    if (sched == RYR_SCHED_CONCURRENT) return h->p.scheme == RYR_SCHEME_A;
    if (sched == RYR_SCHED_ONE) return h->p.scheme == RYR_SCHEME_A && (wg_cru == 256 || wg_cru == 512);
    if (sched == RYR_SCHED_FUSED_GATING) return wg_cru == 128 || wg_cru == 256 || wg_cru == 512;
    return sched == RYR_SCHED_SEQ || sched == RYR_SCHED_FUSED;
}

static void sched_candidates(const ryr_handle *h, std::vector<std::array<int, 3>> &c)
{
    c.clear();
    auto add = [&](int a, int b, int s) {
        for (auto &x : c) if (x[0] == a && x[1] == b && x[2] == s) return;
        c.push_back({a, b, s});
    };
    add(h->wg_cyto, h->wg_cru, RYR_SCHED_SEQ);
    add(h->wg_cyto, h->wg_cru, RYR_SCHED_FUSED);
    /* fused, the membrane's four current groups on four waves: the sub-voxel
     * size timed alone does not see that */
    add(256, h->wg_cru, RYR_SCHED_FUSED);
    /* two lanes per CRU: the gating and LCC loops split, for latency-bound
     * (small) grids */
    add(256, 512, RYR_SCHED_FUSED);
    add(h->wg_cyto, 512, RYR_SCHED_SEQ);
    /* the RyR gating beside the sub-voxel sweep and the membrane step */
    add(256, 256, RYR_SCHED_FUSED_GATING);
    add(256, 128, RYR_SCHED_FUSED_GATING);   /* the gating-free CRU kernel over twice as many CUs */
    add(256, 512, RYR_SCHED_FUSED_GATING);
    add(h->wg_cyto, 256, RYR_SCHED_FUSED_GATING);
    if (h->p.scheme == RYR_SCHEME_A) {
        /* one launch per step, one or two lanes per CRU */
        add(256, 256, RYR_SCHED_ONE);
        add(256, 512, RYR_SCHED_ONE);
        add(h->wg_cyto, h->wg_cru, RYR_SCHED_CONCURRENT);
        add(128, 64, RYR_SCHED_CONCURRENT);
        add(256, 64, RYR_SCHED_CONCURRENT);
        add(128, 128, RYR_SCHED_CONCURRENT);
        add(64, 64, RYR_SCHED_CONCURRENT);
        /* the one-lane-per-CRU sweep beside the CRU kernel (grids that fill the GPU) */
        add(1024, 64, RYR_SCHED_CONCURRENT);
        add(1024, 128, RYR_SCHED_CONCURRENT);
    }
}

/* one candidate: advances 3 * RYR_TUNE_GSTEPS steps (none of them an output
 * step) and records the best timed replay in ms; false if it cannot be
 * timed as a graph (then it is skipped) */
static int tune_candidate(ryr_handle *h, const std::array<int, 3> &c, float *best_ms, bool *ok)
{
    const int sw_c = h->wg_cyto, sw_r = h->wg_cru;
    h->wg_cyto = c[0];
    h->wg_cru = c[1];
    *ok = false;
    *best_ms = 1e30f;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    RYR_HIPCHECK(h, set_base(h, h->step));
    bool cap = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (cap) {
        bool run = enqueue_run(h, RYR_TUNE_GSTEPS, c[2]) == hipSuccess;
        cap = hipStreamEndCapture(h->stream, &g) == hipSuccess && g && run;
    }
    if (cap) cap = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) hipGraphDestroy(g);
    h->wg_cyto = sw_c;
    h->wg_cru = sw_r;
    if (!cap) {
        (void)hipGetLastError();
        if (ge) hipGraphExecDestroy(ge);
        return RYR_OK;
    }
    for (int r = 0; r < 3; ++r) {
        RYR_HIPCHECK(h, set_base(h, h->step));
        RYR_HIPCHECK(h, hipEventRecord(h->kev[0], h->stream));
        RYR_HIPCHECK(h, hipGraphLaunch(ge, h->stream));
        RYR_HIPCHECK(h, hipEventRecord(h->kev[1], h->stream));
        RYR_HIPCHECK(h, hipEventSynchronize(h->kev[1]));
        float ms = 0.0f;
        RYR_HIPCHECK(h, hipEventElapsedTime(&ms, h->kev[0], h->kev[1]));
        if (r > 0 && ms < *best_ms) *best_ms = ms;
        h->step += RYR_TUNE_GSTEPS;
        h->launches += RYR_TUNE_GSTEPS;
    }
    hipGraphExecDestroy(ge);
    *ok = true;
    return RYR_OK;
}

/* A step segment as a graph: n steps that start on an output step (out0) or
 * hold none (so the segment never crosses an output step after its first),
 * at offsets 0..n-1 from the base, then a node that advances the base by n.
 * The kernels read the absolute step from the base, so one graph serves
 * every segment of that kind wherever it falls. */
#define RYR_MAX_SEGS 8
static ryr_handle::Seg *find_seg(ryr_handle *h, bool out0, long long n)
{
    for (auto &q : h->segs) if (q.out0 == out0 && q.n == n) return &q;
    return nullptr;
}

static bool capture_seg(ryr_handle *h, ryr_handle::Seg &q)
{
    if (h->graph_failed) return false;
    const long long spo = h->p.steps_per_output;
    /* offsets are formed from h->step - h->gbase: capture as if the segment
     * started at a base of its own (nothing runs during a capture) */
    const long long sv_step = h->step, sv_base = h->gbase;
    h->step = q.out0 ? 0 : 1;
    h->gbase = h->step;
    hipGraph_t g = nullptr;
    bool ok = q.n >= 1 && (q.out0 ? q.n <= spo : q.n < spo) &&
              hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        ok = enqueue_run(h, q.n, h->sched < 0 ? RYR_SCHED_SEQ : h->sched) == hipSuccess;
        if (ok) {
            hipLaunchKernelGGL(k_stepbase, dim3(1), dim3(64), 0, h->stream, h->a.gstep, q.n, 1);
            ok = hipGetLastError() == hipSuccess;
        }
        hipError_t e = hipStreamEndCapture(h->stream, &g);
        ok = ok && e == hipSuccess && g;
    }
    h->step = sv_step;
    h->gbase = sv_base;
    if (ok) ok = hipGraphInstantiate(&q.ge, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) hipGraphDestroy(g);
    if (!ok) {
        q.ge = nullptr;
        h->graph_failed = true;
        (void)hipGetLastError();
    }
    return ok;
}

/* The graph for the next n steps (n within one output interval), or null
 * to run them eagerly.  A whole output interval is captured at its first
 * use; any other segment once it has been asked for before (or prepared by
 * ryr_prepare), so that one-off remainders are not captured. */
static hipGraphExec_t seg_graph(ryr_handle *h, bool out0, long long n, bool capture_now)
{
    if (n < 2 || h->sched < 0) return nullptr;
    ryr_handle::Seg *q = find_seg(h, out0, n);
    if (!q) {
        if (h->segs.size() >= 4 * RYR_MAX_SEGS) return nullptr;
        h->segs.push_back({out0, n, nullptr, 0});
        q = &h->segs.back();
    }
    if (q->ge) return q->ge;
    q->seen += 1;
    const bool whole = out0 && n == h->p.steps_per_output;
    if (!(capture_now || whole || q->seen >= 2)) return nullptr;
    int ncap = 0;
    for (auto &x : h->segs) ncap += x.ge ? 1 : 0;
    if (ncap >= RYR_MAX_SEGS) return nullptr;
    return capture_seg(h, *q) ? q->ge : nullptr;
}

extern "C" int ryr_step(ryr_handle *h, int64_t nsteps)
{
    if (!h || nsteps < 0) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    if (nsteps == 0) return RYR_OK;
    if (!h->timing_open) {
        RYR_HIPCHECK(h, hipEventRecord(h->ev0, h->stream));
        h->timing_open = true;
        h->launches = 0;
    }
    const long long spo = h->p.steps_per_output;
    while (nsteps > 0 && h->wg_cyto == 0) {
        int rc = tune_step(h);
        if (rc) return rc;
        h->step += 1;
        h->launches += 1;
        nsteps -= 1;
    }
    while (h->sched < 0 && nsteps > 0) {
        if (h->cand.empty()) {
            sched_candidates(h, h->cand);
            h->cand_ms.assign(h->cand.size(), 1e30f);
            h->cand_i = 0;
        }
        const long long span = 3 * RYR_TUNE_GSTEPS;
        const long long spo_left = spo - h->step % spo;     /* steps before the next output step */
        if (h->step % spo == 0 || spo_left <= span) {
            /* an output step is at hand: step past it eagerly */
            long long n = h->step % spo == 0 ? 1 : spo_left + 1;
            if (n > nsteps) n = nsteps;
            RYR_HIPCHECK(h, ensure_base(h, n));
            RYR_HIPCHECK(h, enqueue_run(h, n, RYR_SCHED_SEQ));
            h->step += n;
            h->launches += n;
            nsteps -= n;
            continue;
        }
        if (nsteps < span) break;
        float ms = 0.0f;
        bool ok = false;
        int rc = tune_candidate(h, h->cand[h->cand_i], &ms, &ok);
        if (rc) return rc;
        /* a candidate that could not be captured advanced no step */
        if (ok) {
            h->cand_ms[h->cand_i] = ms;
            nsteps -= span;
        }
        if (++h->cand_i == h->cand.size()) {
            size_t best = 0;
            for (size_t q = 1; q < h->cand.size(); ++q)
                if (h->cand_ms[q] < 0.97f * h->cand_ms[best]) best = q;
            h->wg_cyto = h->cand[best][0];
            h->wg_cru = h->cand[best][1];
            h->sched = h->cand[best][2];
            h->cand.clear();
            drop_graph(h);
        }
    }
    while (nsteps > 0) {
        /* segments end at the next output step (or the end of the call) */
        long long n = spo - h->step % spo;
        if (n > nsteps) n = nsteps;
        if (hipGraphExec_t ge = seg_graph(h, h->step % spo == 0, n, false)) {
            if (h->gbase != h->step) RYR_HIPCHECK(h, set_base(h, h->step));
            RYR_HIPCHECK(h, hipGraphLaunch(ge, h->stream));
            h->step += n;
            h->gbase += n;
            h->launches += n;
            nsteps -= n;
            continue;
        }
        RYR_HIPCHECK(h, ensure_base(h, n));
        RYR_HIPCHECK(h, enqueue_run(h, n, h->sched < 0 ? RYR_SCHED_SEQ : h->sched));
        h->step += n;
        h->launches += n;
        nsteps -= n;
    }
    RYR_HIPCHECK(h, hipEventRecord(h->ev1, h->stream));
    return RYR_OK;
}

extern "C" int ryr_prepare(ryr_handle *h, int64_t nsteps)
{
    if (!h || nsteps < 0) return RYR_E_ARG;
    if (h->sched < 0) { h->err = "ryr_prepare: the launch shape is still being tuned"; return RYR_E_STATE; }
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    const long long spo = h->p.steps_per_output;
    long long s = h->step;
    while (nsteps > 0) {
        long long n = spo - s % spo;
        if (n > nsteps) n = nsteps;
        (void)seg_graph(h, s % spo == 0, n, true);
        s += n;
        nsteps -= n;
    }
    if (h->graph_failed) { h->err = "ryr_prepare: graph capture failed"; return RYR_E_HIP; }
    return RYR_OK;
}

extern "C" int ryr_latency_floor(ryr_handle *h, int64_t nsteps, double *us_per_step)
{
    if (!h || nsteps < 2 || !us_per_step) return RYR_E_ARG;
    if (h->sched < 0) { h->err = "ryr_latency_floor: the launch shape is still being tuned"; return RYR_E_STATE; }
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    const long long spo = h->p.steps_per_output;
    const long long n = nsteps < spo ? nsteps : spo;
    /* the same segment the production path replays (whole intervals start on
     * an output step), with every kernel an empty one of the same shape */
    const long long sv_step = h->step, sv_base = h->gbase;
    h->step = n == spo ? 0 : 1;
    h->gbase = h->step;
    h->null_launch = true;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    bool ok = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        ok = enqueue_run(h, n, h->sched) == hipSuccess && go_null(h, 1, 64) == hipSuccess;   /* base advance */
        ok = hipStreamEndCapture(h->stream, &g) == hipSuccess && g && ok;
    }
    h->null_launch = false;
    h->step = sv_step;
    h->gbase = sv_base;
    if (ok) ok = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) hipGraphDestroy(g);
    if (!ok) {
        (void)hipGetLastError();
        h->err = "ryr_latency_floor: graph capture failed";
        return RYR_E_HIP;
    }
    float best = 1e30f;
    int rc = RYR_OK;
    for (int r = 0; r < 6 && rc == RYR_OK; ++r) {
        float ms = 0.0f;
        if (hipEventRecord(h->kev[0], h->stream) != hipSuccess || hipGraphLaunch(ge, h->stream) != hipSuccess ||
            hipEventRecord(h->kev[1], h->stream) != hipSuccess || hipEventSynchronize(h->kev[1]) != hipSuccess ||
            hipEventElapsedTime(&ms, h->kev[0], h->kev[1]) != hipSuccess) {
            h->err = "ryr_latency_floor: replay failed";
            rc = RYR_E_HIP;
        }
        if (r > 0 && ms < best) best = ms;
    }
    hipGraphExecDestroy(ge);
    if (rc) return rc;
    *us_per_step = 1e3 * (double)best / (double)n;
    return RYR_OK;
}

extern "C" int ryr_sync(ryr_handle *h)
{
    if (!h) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    return RYR_OK;
}

extern "C" int ryr_kernel_time(ryr_handle *h, double *ms, int64_t *launches)
{
    if (!h || !ms || !launches) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    *ms = 0.0;
    *launches = 0;
    if (!h->timing_open) return RYR_OK;
    RYR_HIPCHECK(h, hipEventSynchronize(h->ev1));
    float f = 0.0f;
    RYR_HIPCHECK(h, hipEventElapsedTime(&f, h->ev0, h->ev1));
    *ms = (double)f;
    *launches = h->launches;
    h->timing_open = false;
    return RYR_OK;
}

extern "C" int ryr_set_launch_shape(ryr_handle *h, int wg_cyto, int wg_cru, int sched)
{
    if (!h) return RYR_E_ARG;
    auto ok = [](int w) { return w == 64 || w == 128 || w == 256; };
    auto ok_cru = [&](int w) { return ok(w) || w == 512; };
    if (wg_cyto == 0 && wg_cru == 0) {
        h->wg_cyto = h->wg_cru = 0;
        h->sched = -1;
        h->tune_i = 0;
        h->cand.clear();
        for (auto &k : h->tune_ms) for (auto &q : k) q.clear();
    } else if ((ok(wg_cyto) || wg_cyto == 1024) && ok_cru(wg_cru) && shape_ok(h, wg_cyto, wg_cru, sched)) {
        h->wg_cyto = wg_cyto;
        h->wg_cru = wg_cru;
        h->sched = sched;
    } else {
        h->err = "workgroup sizes must be 64, 128 or 256 (sub-voxel kernel also 1024 with the sequential or "
                 "concurrent schedule, CRU kernel also 512) and the schedule one of "
                 "RYR_SCHED_* (concurrent needs cleft scheme A, gating a CRU size of 128, 256 or 512), "
                 "or sizes 0, 0 to tune";
        return RYR_E_ARG;
    }
    drop_graph(h);
    return RYR_OK;
}

extern "C" int ryr_get_launch_shape(const ryr_handle *h, int *wg_cyto, int *wg_cru, int *sched)
{
    if (!h || !wg_cyto || !wg_cru || !sched) return RYR_E_ARG;
    *wg_cyto = h->wg_cyto;
    *wg_cru = h->wg_cru;
    *sched = h->sched;
    return RYR_OK;
}

extern "C" int ryr_profile_kernels(ryr_handle *h, int64_t nsteps, double *avg_us, int64_t *count)
{
    if (!h || nsteps < 0 || !avg_us || !count) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    for (int k = 0; k < RYR_NPROF; ++k) { avg_us[k] = 0.0; count[k] = 0; }
    if (h->sched < 0) {
        h->err = "ryr_profile_kernels: the step schedule is still being tuned; step past the tuning first";
        return RYR_E_STATE;
    }
    double tot[RYR_NPROF] = {0, 0, 0, 0};
    bool pending = false;
    for (int64_t i = 0; i < nsteps; ++i) {
        bool ran[4];
        hipEvent_t tail = nullptr;
        RYR_HIPCHECK(h, ensure_base(h, 1));
        RYR_HIPCHECK(h, enqueue_profiled_step(h, i == nsteps - 1, &pending, ran, &tail));
        RYR_HIPCHECK(h, hipEventSynchronize(tail));
        if (ran[1]) RYR_HIPCHECK(h, hipEventSynchronize(h->kev[3]));
        for (int k = 0; k < 4; ++k) {
            if (!ran[k]) continue;
            float ms = 0.0f;
            RYR_HIPCHECK(h, hipEventElapsedTime(&ms, h->kev[2 * k], h->kev[2 * k + 1]));
            tot[k] += 1e3 * (double)ms;
            count[k] += 1;
        }
        h->step += 1;
    }
    for (int k = 0; k < RYR_NPROF; ++k) avg_us[k] = count[k] ? tot[k] / (double)count[k] : 0.0;
    return RYR_OK;
}

extern "C" int ryr_read_scalars(ryr_handle *h, double *out)
{
    if (!h || !out) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    std::vector<RyrMem> mem(h->ncells);
    RYR_HIPCHECK(h, hipMemcpy(mem.data(), h->a.mem, sizeof(RyrMem) * h->ncells, hipMemcpyDeviceToHost));
    for (int c = 0; c < h->ncells; ++c) {
        const RyrMem &M = mem[c];
        double *o = out + (long long)c * RYR_S_COUNT;
        o[RYR_S_T] = (double)(M.step > 0 ? M.step - 1 : 0) * h->p.dt;
        for (int q = 0; q < 13; ++q) o[RYR_S_V + q] = M.m[q];
        o[RYR_S_CSUBT] = M.cur[0]; o[RYR_S_ICAL] = M.cur[1]; o[RYR_S_INCX] = M.cur[2];
        o[RYR_S_INA] = M.cur[3]; o[RYR_S_IKS] = M.cur[4]; o[RYR_S_IKR] = M.cur[5];
        o[RYR_S_IK1] = M.cur[6]; o[RYR_S_INAK] = M.cur[7]; o[RYR_S_ITOS] = M.cur[8];
        o[RYR_S_ITOF] = M.cur[9]; o[RYR_S_STIM] = M.stim; o[RYR_S_OUT_ICA] = M.out_ica;
        o[RYR_S_OUT_NCX] = M.out_ncx; o[RYR_S_OUT_INA] = M.out_ina; o[RYR_S_STEP] = (double)M.step;
    }
    return RYR_OK;
}

static int field_planes(int f)
{
    if (f < 0 || f >= RYR_F_COUNT) return 0;
    return (f <= RYR_F_MGMYO || f == RYR_F_LCC) ? 8 : 1;
}

static int ensure_stage(ryr_handle *h, size_t n)
{
    if (h->stage_n >= n) return RYR_OK;
    if (h->stage) hipFree(h->stage);
    h->stage = nullptr;
    h->stage_n = 0;
    RYR_HIPCHECK(h, hipMalloc((void **)&h->stage, n * sizeof(double)));
    h->stage_n = n;
    return RYR_OK;
}

extern "C" int ryr_read_field(ryr_handle *h, int cell, int field, double *out, int64_t n)
{
    if (!h || !out || cell < 0 || cell >= h->ncells) return RYR_E_ARG;
    const int planes = field_planes(field);
    if (!planes) { h->err = "unknown field"; return RYR_E_ARG; }
    const long long nel = (long long)planes * h->info[cell].N;
    if (n < nel) { h->err = "output buffer too small"; return RYR_E_ARG; }
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    if (int rc = ensure_stage(h, (size_t)nel)) return rc;
    const bool f32 = h->p.precision == RYR_PREC_F32;
    RYR_HIPCHECK(h, f32 ? ryr_launch_export_f32(h->a, cell, field, planes, nel, h->stage, h->stream)
                        : ryr_launch_export_f64(h->a, cell, field, planes, nel, h->stage, h->stream));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    RYR_HIPCHECK(h, hipMemcpy(out, h->stage, nel * sizeof(double), hipMemcpyDeviceToHost));
    return RYR_OK;
}

extern "C" int ryr_write_field(ryr_handle *h, int cell, int field, const double *in, int64_t n)
{
    if (!h || !in || cell < 0 || cell >= h->ncells) return RYR_E_ARG;
    const int planes = field_planes(field);
    if (!planes) { h->err = "unknown field"; return RYR_E_ARG; }
    const long long nel = (long long)planes * h->info[cell].N;
    if (n < nel) { h->err = "input buffer too small"; return RYR_E_ARG; }
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    if (int rc = ensure_stage(h, (size_t)nel)) return rc;
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    RYR_HIPCHECK(h, hipMemcpy(h->stage, in, nel * sizeof(double), hipMemcpyHostToDevice));
    const bool f32 = h->p.precision == RYR_PREC_F32;
    RYR_HIPCHECK(h, f32 ? ryr_launch_import_f32(h->a, cell, field, planes, nel, h->stage, h->stream)
                        : ryr_launch_import_f64(h->a, cell, field, planes, nel, h->stage, h->stream));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    return RYR_OK;
}

extern "C" int ryr_write_scalar(ryr_handle *h, int cell, int s, double value)
{
    if (!h || cell < 0 || cell >= h->ncells || s < RYR_S_V || s > RYR_S_XNAI) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    RyrMem M;
    RYR_HIPCHECK(h, hipMemcpy(&M, h->a.mem + cell, sizeof M, hipMemcpyDeviceToHost));
    M.m[s - RYR_S_V] = value;
    RYR_HIPCHECK(h, hipMemcpy(h->a.mem + cell, &M, sizeof M, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_vfactors, dim3((h->ncells + 63) / 64), dim3(64), 0, h->stream, h->a.mem,
                       h->ncells, h->p.iso2);
    RYR_HIPCHECK(h, hipGetLastError());
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    return RYR_OK;
}

extern "C" int ryr_read_diag(ryr_handle *h, double *out, int64_t *step_out)
{
    if (!h || !out) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    RYR_HIPCHECK(h, hipMemcpy(out, h->a.diag, sizeof(double) * h->ncells * RYR_DIAG_COLS, hipMemcpyDeviceToHost));
    if (step_out) {
        long long s = -1;
        RYR_HIPCHECK(h, hipMemcpy(&s, h->a.diag_step, sizeof s, hipMemcpyDeviceToHost));
        *step_out = s;
    }
    return RYR_OK;
}

extern "C" int ryr_read_linescan(ryr_handle *h, int cell, double *out, int64_t n)
{
    if (!h || !out || cell < 0 || cell >= h->ncells) return RYR_E_ARG;
    const long long nel = (long long)h->info[cell].X * RYR_LINE_COLS;
    if (n < nel) { h->err = "output buffer too small"; return RYR_E_ARG; }
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    RYR_HIPCHECK(h, hipMemcpy(out, h->a.line + h->line_off[cell], sizeof(double) * nel, hipMemcpyDeviceToHost));
    return RYR_OK;
}

extern "C" int ryr_read_counters(ryr_handle *h, int64_t *out)
{
    if (!h || !out) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    std::vector<unsigned long long> v((size_t)h->ncells * RYR_NCNT);
    RYR_HIPCHECK(h, hipMemcpy(v.data(), h->a.counters, sizeof(unsigned long long) * v.size(), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < v.size(); ++q) out[q] = (int64_t)v[q];
    return RYR_OK;
}

extern "C" int ryr_read_beats(ryr_handle *h, int cell, double *out, int maxbeats, int *nbeats)
{
    if (!h || !out || !nbeats || cell < 0 || cell >= h->ncells || maxbeats < 0) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    RyrMem M;
    RYR_HIPCHECK(h, hipMemcpy(&M, h->a.mem + cell, sizeof M, hipMemcpyDeviceToHost));
    std::vector<double> rec((size_t)RYR_MAXB * RYR_BCOLS);
    RYR_HIPCHECK(h, hipMemcpy(rec.data(), h->a.beats + (size_t)cell * RYR_MAXB * RYR_BCOLS,
                              sizeof(double) * rec.size(), hipMemcpyDeviceToHost));
    int done = M.beat_n < RYR_MAXB ? M.beat_n : RYR_MAXB;
    int k = 0;
    for (; k < done && k < maxbeats; ++k) {
        for (int q = 0; q < RYR_BCOLS; ++q) out[k * RYR_BCOLS + q] = rec[(size_t)k * RYR_BCOLS + q];
        out[k * RYR_BCOLS] *= h->p.dt;
    }
    if (M.beat_on >= 0 && k < maxbeats && M.beat_n < RYR_MAXB) {
        double *r = out + (size_t)k * RYR_BCOLS;
        r[0] = (double)M.beat_on * h->p.dt; r[1] = M.b_apd90; r[2] = M.b_vmax; r[3] = M.b_vrest;
        r[4] = (M.beat_flags & 2) ? 1.0 : 0.0; r[5] = M.b_ead_t; r[6] = M.b_peak_ci; r[7] = M.b_jsr;
        r[8] = M.b_sparks; r[9] = 0.0;
        r[10] = (M.beat_flags & 8) ? 1.0 : 0.0; r[11] = M.b_eadl_t;
        ++k;
    }
    *nbeats = k;
    return RYR_OK;
}

extern "C" int ryr_set_vclamp(ryr_handle *h, int cell, const double *v, int64_t n, int periodic)
{
    if (!h || cell < 0 || cell >= h->ncells || n < 0 || (n > 0 && !v)) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (h->vclamp[cell]) { RYR_HIPCHECK(h, hipFree(h->vclamp[cell])); h->vclamp[cell] = nullptr; }
    RyrCellInfo &I = h->info[cell];
    I.vclamp = nullptr;
    I.vclamp_n = 0;
    I.vclamp_periodic = periodic ? 1 : 0;
    if (n > 0) {
        RYR_HIPCHECK(h, hipMalloc((void **)&h->vclamp[cell], sizeof(double) * n));
        RYR_HIPCHECK(h, hipMemcpy(h->vclamp[cell], v, sizeof(double) * n, hipMemcpyHostToDevice));
        I.vclamp = h->vclamp[cell];
        I.vclamp_n = n;
    }
    return upload_info(h);
}

extern "C" int ryr_cru_steps(ryr_handle *h, int64_t *out)
{
    if (!h || !out) return RYR_E_ARG;
    long long n = 0;
    for (const auto &I : h->info) n += I.N;
    *out = (int64_t)(n * h->step);
    return RYR_OK;
}

/* snapshot: header + every device allocation in allocation order.  `cfg`
 * is a hash of every run parameter and every cell description the handle was
 * created with (field by field, so struct padding plays no part): a
 * snapshot loads only into a handle of the same configuration, since the
 * state arrays alone do not record the lattice shape, the scheme, dt, the
 * variants, the seeds or the channel parameters they were advanced with. */
struct RyrSnapHeader {
    char magic[8];
    int version, precision, ncells, pad_;
    long long P, step, nallocs;
    unsigned long long cfg;
};

static void fnv_bytes(unsigned long long &hsh, const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) { hsh ^= b[i]; hsh *= 0x100000001b3ull; }
}

static unsigned long long config_hash(const ryr_handle *h)
{
    unsigned long long s = 0xcbf29ce484222325ull;
    auto d = [&](double x) { fnv_bytes(s, &x, sizeof x); };
    auto i = [&](long long x) { fnv_bytes(s, &x, sizeof x); };
    const ryr_params &p = h->p;
    d(p.dt); i(p.scheme); i(p.precision); i(p.iso2); i(p.tapsi); i(p.lqt); i(p.nlcc);
    i(p.steps_per_output); d(p.stim_amp); d(p.stim_dur); d(p.stim_offset); d(p.ci0); d(p.v0);
    i(p.cm_mode); d(p.cm); i(p.variants);
    i(h->ncells);
    for (const ryr_cell_desc &c : h->desc) {
        i(c.nx); i(c.ny); i(c.nz); i((long long)c.seed); d(c.cl); d(c.cjsr0); d(c.ncx_scale);
        d(c.Ku); d(c.Kb); d(c.ica_scale); d(c.Ku2); d(c.ikr_scale);
    }
    return s;
}

extern "C" int ryr_save(ryr_handle *h, const char *path)
{
    if (!h || !path) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    FILE *f = std::fopen(path, "wb");
    if (!f) { h->err = std::string("cannot open ") + path; return RYR_E_ARG; }
    RyrSnapHeader hd;
    std::memset(&hd, 0, sizeof hd);
    std::memcpy(hd.magic, "RYRSNP2", 8);
    hd.version = RYR_ABI_VERSION; hd.precision = h->p.precision; hd.ncells = h->ncells;
    hd.P = h->P; hd.step = h->step; hd.nallocs = (long long)h->allocs.size();
    hd.cfg = config_hash(h);
    bool ok = std::fwrite(&hd, sizeof hd, 1, f) == 1;
    std::vector<char> buf;
    for (auto &al : h->allocs) {
        if (!ok) break;
        buf.resize(al.second);
        ok = hipMemcpy(buf.data(), al.first, al.second, hipMemcpyDeviceToHost) == hipSuccess;
        long long sz = (long long)al.second;
        ok = ok && std::fwrite(&sz, sizeof sz, 1, f) == 1 && std::fwrite(buf.data(), 1, al.second, f) == al.second;
    }
    std::fclose(f);
    if (!ok) { h->err = "snapshot write failed"; return RYR_E_HIP; }
    return RYR_OK;
}

extern "C" int ryr_load(ryr_handle *h, const char *path)
{
    if (!h || !path) return RYR_E_ARG;
    RYR_HIPCHECK(h, hipSetDevice(h->device));
    RYR_HIPCHECK(h, hipStreamSynchronize(h->stream));
    FILE *f = std::fopen(path, "rb");
    if (!f) { h->err = std::string("cannot open ") + path; return RYR_E_ARG; }
    RyrSnapHeader hd;
    bool ok = std::fread(&hd, sizeof hd, 1, f) == 1 && std::memcmp(hd.magic, "RYRSNP2", 8) == 0 &&
              hd.version == RYR_ABI_VERSION && hd.precision == h->p.precision && hd.ncells == h->ncells &&
              hd.P == h->P && hd.nallocs == (long long)h->allocs.size() && hd.cfg == config_hash(h);
    std::vector<char> buf;
    for (auto &al : h->allocs) {
        if (!ok) break;
        long long sz = 0;
        ok = std::fread(&sz, sizeof sz, 1, f) == 1 && sz == (long long)al.second;
        if (!ok) break;
        buf.resize(al.second);
        ok = std::fread(buf.data(), 1, al.second, f) == al.second &&
             hipMemcpy(al.first, buf.data(), al.second, hipMemcpyHostToDevice) == hipSuccess;
    }
    std::fclose(f);
    if (!ok) { h->err = "snapshot does not match this handle or is corrupt"; return RYR_E_STATE; }
    h->step = hd.step;
    h->gbase = -1;
    return RYR_OK;
}

extern "C" void ryr_destroy(ryr_handle *h)
{
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    drop_graph(h);
    for (auto &al : h->allocs) hipFree(al.first);
    for (void *t : h->tables) hipFree(t);
    if (h->stage) hipFree(h->stage);
    for (double *q : h->vclamp) if (q) hipFree(q);
    for (auto &e : h->kev) if (e) hipEventDestroy(e);
    if (h->fork) hipEventDestroy(h->fork);
    if (h->join) hipEventDestroy(h->join);
    if (h->stream2) { hipStreamSynchronize(h->stream2); hipStreamDestroy(h->stream2); }
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

extern "C" int ryr_gating_probe(int device, int64_t n, uint64_t seed, const uint32_t *cru, uint64_t step,
                                double Ku, double Ku2, double Kb, double dt, int iso2,
                                int32_t *counts, int32_t *lcc, const double *conc,
                                int32_t *draws, int32_t *nopen, int32_t *fails)
{
    if (n <= 0 || !cru || !counts || !lcc || !conc || !draws || !nopen || !fails || !(dt > 0)) {
        g_create_err = "ryr_gating_probe: bad arguments";
        return RYR_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { g_create_err = "no HIP device"; return RYR_E_NODEV; }
    if (hipSetDevice(device) != hipSuccess) { g_create_err = "hipSetDevice failed"; return RYR_E_HIP; }
    uint32_t *d_cru = nullptr; int32_t *d_counts = nullptr, *d_lcc = nullptr, *d_draws = nullptr;
    int32_t *d_nopen = nullptr, *d_fails = nullptr; double *d_conc = nullptr;
    bool ok = hipMalloc((void **)&d_cru, n * 4) == hipSuccess &&
              hipMalloc((void **)&d_counts, n * 16) == hipSuccess &&
              hipMalloc((void **)&d_lcc, n * 32) == hipSuccess &&
              hipMalloc((void **)&d_conc, n * 32) == hipSuccess &&
              hipMalloc((void **)&d_draws, n * 4) == hipSuccess &&
              hipMalloc((void **)&d_nopen, n * 4) == hipSuccess &&
              hipMalloc((void **)&d_fails, n * 4) == hipSuccess;
    ok = ok && hipMemcpy(d_cru, cru, n * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(d_counts, counts, n * 16, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(d_lcc, lcc, n * 32, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(d_conc, conc, n * 32, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_gating_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, (long long)n,
                           (unsigned long long)seed, d_cru, (unsigned long long)step, Ku, Ku2, Kb, dt, iso2,
                           d_counts, d_lcc, d_conc, d_draws, d_nopen, d_fails);
        ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    }
    ok = ok && hipMemcpy(counts, d_counts, n * 16, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(lcc, d_lcc, n * 32, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(draws, d_draws, n * 4, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(nopen, d_nopen, n * 4, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(fails, d_fails, n * 4, hipMemcpyDeviceToHost) == hipSuccess;
    hipFree(d_cru); hipFree(d_counts); hipFree(d_lcc); hipFree(d_conc);
    hipFree(d_draws); hipFree(d_nopen); hipFree(d_fails);
    if (!ok) { g_create_err = "ryr_gating_probe: HIP error"; return RYR_E_HIP; }
    return RYR_OK;
}

/* ---- the gating probe without allocation
 * ryr_gating_probe_noalloc stages the queries in `scratch`, a file-scope
 * device array that every GPU the library runs on holds until the process
 * ends, instead of device buffers allocated per call.  Host code never
 * passes `scratch` itself to a kernel or a copy (in host code the name is a
 * host-side stand-in, not a device address): the address of each device's
 * copy is looked up once, with that device current, and kept per device. */
/* RYR_PROBE_NOALLOC_MAX (queries per call) comes from the public header */
#define RYR_PROBE_Q_BYTES (4 + 16 + 32 + 32 + 4 + 4 + 4)  /* cru, counts, lcc, conc, draws, nopen, fails */
#define RYR_PROBE_SCRATCH_BYTES (RYR_PROBE_NOALLOC_MAX * RYR_PROBE_Q_BYTES + 7 * 256)
static __device__ __align__(256) unsigned char scratch[RYR_PROBE_SCRATCH_BYTES];

#define RYR_MAX_DEVICES 64
static unsigned char *g_scratch_ptr[RYR_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_on(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_create_err = std::string("ryr_gating_probe_noalloc: hipGetSymbolAddress(scratch): ") +
                           hipGetErrorString(e);
            return RYR_E_HIP;
        }
        g_scratch_ptr[device] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[device];
    return RYR_OK;
}

extern "C" int ryr_gating_probe_noalloc(int device, int64_t n, uint64_t seed, const uint32_t *cru, uint64_t step,
                                        double Ku, double Ku2, double Kb, double dt, int iso2,
                                        int32_t *counts, int32_t *lcc, const double *conc,
                                        int32_t *draws, int32_t *nopen, int32_t *fails)
{
    if (n <= 0 || !cru || !counts || !lcc || !conc || !draws || !nopen || !fails || !(dt > 0)) {
        g_create_err = "ryr_gating_probe_noalloc: bad arguments";
        return RYR_E_ARG;
    }
    if (n > RYR_PROBE_NOALLOC_MAX) {
        g_create_err = "ryr_gating_probe_noalloc: at most " + std::to_string(RYR_PROBE_NOALLOC_MAX) +
                       " queries per call (" + std::to_string((long long)n) + " given); use ryr_gating_probe";
        return RYR_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { g_create_err = "no HIP device"; return RYR_E_NODEV; }
    if (device < 0 || device >= ndev || device >= RYR_MAX_DEVICES) { g_create_err = "bad device"; return RYR_E_ARG; }
    if (hipSetDevice(device) != hipSuccess) { g_create_err = "hipSetDevice failed"; return RYR_E_HIP; }
    unsigned char *base = nullptr;
    if (int rc = scratch_on(device, &base)) return rc;
    /* the seven arrays of this call, each on a 256-byte boundary */
    size_t off = 0;
    auto take = [&](size_t bytes) { unsigned char *p = base + off; off += (bytes + 255) & ~(size_t)255; return p; };
    uint32_t *d_cru = (uint32_t *)take(n * 4);
    int32_t *d_counts = (int32_t *)take(n * 16);
    int32_t *d_lcc = (int32_t *)take(n * 32);
    double *d_conc = (double *)take(n * 32);
    int32_t *d_draws = (int32_t *)take(n * 4);
    int32_t *d_nopen = (int32_t *)take(n * 4);
    int32_t *d_fails = (int32_t *)take(n * 4);
    if (off > RYR_PROBE_SCRATCH_BYTES) { g_create_err = "ryr_gating_probe_noalloc: scratch overflow"; return RYR_E_ARG; }
    /* every output is written by the kernel for each query, so nothing needs clearing */
    bool ok = hipMemcpy(d_cru, cru, n * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_counts, counts, n * 16, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_lcc, lcc, n * 32, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_conc, conc, n * 32, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_gating_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, (long long)n,
                           (unsigned long long)seed, d_cru, (unsigned long long)step, Ku, Ku2, Kb, dt, iso2,
                           d_counts, d_lcc, d_conc, d_draws, d_nopen, d_fails);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(0) == hipSuccess;
    }
    ok = ok && hipMemcpy(counts, d_counts, n * 16, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(lcc, d_lcc, n * 32, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(draws, d_draws, n * 4, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(nopen, d_nopen, n * 4, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(fails, d_fails, n * 4, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) { g_create_err = "ryr_gating_probe_noalloc: HIP error"; return RYR_E_HIP; }
    return RYR_OK;
}
