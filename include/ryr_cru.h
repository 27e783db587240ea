/*
 * ryr_cru.h - C ABI of the MI355X (gfx950) CRU-update engine for the detailed
 * rabbit ventricular myocyte model (Terentyev, Rees et al., Circ Res 2014).
 *
 * The reference has no library interface: its boundary is the process CLI
 *   cuda_Cell_Detailed <Device> <name> <a3> <cjsr0> <ncx> <Ku> <r1> <Kb> <ica> <Ku2>
 * (run.sh:9,11; parsed at cuda_Cell_Detailed.cu:236-241,247-256,59,92,96,109,
 * 126/136,401-403,420) and, inside main(), the contract of the Compute kernel
 * (cuda_Cell_Detailed.cu:217,1187) plus the host loop that swaps buffers,
 * reduces I_CaL/I_NCX/cs and integrates the membrane ODE (:399-739).  Every
 * entry point below replaces one piece of that loop; the citation on each
 * names the piece.  The CLI `ryr_cell` (package csrc/host) keeps the
 * positional argv and the output files on top of this ABI.
 *
 * Conventions: plain C types only; every call returns 0 on success or a
 * negative RYR_E_* code; ryr_last_error() gives the message of the last
 * failure on a handle.  A handle owns its device memory and one HIP stream;
 * calls on one handle must not overlap (not thread-safe per handle), distinct
 * handles are independent.  ryr_step is asynchronous; every read synchronises.
 */
#ifndef RYR_CRU_H
#define RYR_CRU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RYR_ABI_VERSION 4

#define RYR_OK            0
#define RYR_E_ARG        -1   /* invalid argument */
#define RYR_E_HIP        -2   /* HIP runtime error */
#define RYR_E_NODEV      -3   /* no usable GPU */
#define RYR_E_STATE      -4   /* handle in wrong state / corrupt snapshot */

#define RYR_SCHEME_A 0        /* cuda_Cell_Detailed.cu:1356-1473 (conservative) */
#define RYR_SCHEME_B 1        /* cuda_Cell_Detailed.cu:1475-1597 */

#define RYR_PREC_F64 0        /* every concentration in FP64 (the reference's arithmetic) */
#define RYR_PREC_F32 1        /* FP32 concentrations/buffers; FP64 gating, V, reductions */

/* Run-wide parameters.  The reference fixes these by macros and Makefile
 * targets (cuda_Cell_Detailed.cu:29-145, Makefile:8-28); here they are runtime
 * values with the reference's values as defaults (ryr_params_default). */
typedef struct ryr_params {
    double dt;              /* ms, DT (:51) = 0.025 */
    int    scheme;          /* RYR_SCHEME_A / _B (the shipped file concatenates both) */
    int    precision;       /* RYR_PREC_F64 / _F32 */
    int    iso2;            /* -Diso2: LCC shapes (reesica.h:30-35), SERCA x1.75 (:1230-1231) */
    int    tapsi;           /* -Dtapsi: cjsr0 := 10 (:1118-1120), SERCA 0 (:1236-1237) */
    int    lqt;             /* -Dlqt: svikr 0 (:128); with tapsi sviup_lqt 0.05 (:121-122) */
    int    nlcc;            /* LCCs per CRU, svncp (:102) = 4, at most 8 */
    int    steps_per_output;/* outt/dt (:63,:777) = 100 */
    double stim_amp;        /* uA/uF (:690) = 80 */
    double stim_dur;        /* ms (:689 "< 1.0") = 1 */
    double stim_offset;     /* ms (:689 "t+tperid-100") = 100 */
    double ci0;             /* ci_basal (:91) = 0.0944 uM, initial ci/cs/cp */
    double v0;              /* initial V (:354) = -80 mV */
    int    cm_mode;         /* 0: Cm = cm * N/16120 (current density preserved), 1: Cm = cm */
    double cm;              /* Cm (:100) = 45 */
    int    variants;        /* OR of RYR_VAR_* (the reference's remaining -D builds); 0 = none */
    int    reserved_;
} ryr_params;

/* Variants of ryr_params.variants (cuda_Cell_Detailed.cu, compile-time there):
 *  PERM       permeabilised cell (:87-93,697-699,1249-1263,1294-1304,1409-1411,
 *             1499-1501,1662-1689): no LCC gating or I_CaL, no NCX, no
 *             stimulus; cs relaxes to ci0 with tau 5 ms; the ci and cs
 *             neighbours outside the lattice are a bath at ci0 (their ghosts
 *             are never refreshed).  The reference's perm build also fixes
 *             ci_basal 0.3 and cjsr_basal 900: set ci0 / cjsr0 for that.
 *  ISO        isoproterenol protocol (:683-693,1213-1229): stimulus at phase
 *             0 of the cycle and only while t < 11.9 CL; from t >= 5.5 CL
 *             SERCA doubles (unless iso2 or tapsi override it, as in the
 *             reference); at t = 5.5 CL (mod 15.5 CL) absent LCC slots 4..7
 *             become channels in state 3.
 *  SLOWNCX    slowncxchange (:540-543,1267-1272,1420,1509): I_Ks qks gate
 *             qks_inf = 0.2(1 + 1.6/(1 + (0.28/csubt)^3)), tau 2000 ms, and a
 *             background Ca current 0.003016/32 (V - E_Ca(cs)) per CRU in the
 *             sub-membrane balance and in the whole-cell I_CaL sum.
 *  RANDOMLCC  LCCs per CRU drawn uniformly from 0..8 at set-up (:1114-1117,
 *             1150-1155) instead of nlcc, from the Philox stream of the CRU
 *             at counter step 2^64-1 (draw % 9). */
#define RYR_VAR_PERM      1
#define RYR_VAR_ISO       2
#define RYR_VAR_SLOWNCX   4
#define RYR_VAR_RANDOMLCC 8

/* One independent cell.  Lattice dims are interior CRUs (the reference's
 * nx,ny,nz = 64,28,12 of :80-82 include a ghost shell: interior 62x26x10). */
typedef struct ryr_cell_desc {
    int      nx, ny, nz;    /* interior CRUs in x (longitudinal), y, z */
    int      pad_;
    uint64_t seed;          /* Philox key; replaces curand_init(18, pos, 0) (:1700) */
    double   cl;            /* pacing cycle length tperid (:50), ms */
    double   cjsr0;         /* argv[4] cjsr_basal (:92), uM; also initial cnsr */
    double   ncx_scale;     /* argv[5]: svncx = 2.5*ncx_scale (:109) */
    double   Ku;            /* argv[6]: svkpr_lqt (:126/136), uM */
    double   Kb;            /* argv[8] (:402) */
    double   ica_scale;     /* argv[9]: svica = 0.666*1.2*ica_scale (:96) */
    double   Ku2;           /* argv[10] (:401): RyR open-rate multiplier */
    double   ikr_scale;     /* svikr = 3*ikr_scale (:138); forced 0 when lqt */
} ryr_cell_desc;

typedef struct ryr_handle ryr_handle;

/* Membrane/whole-cell scalars of one cell, as ryr_read_scalars returns them
 * (index into the per-cell row).  Values are those of the last completed step,
 * in the reference's units.  Cuda_Cell_Detailed.cu:350-366 (state),
 * :430-456 (reductions), :458-739 (currents). */
enum ryr_scalar {
    RYR_S_T = 0,      /* time at the start of the last step, ms */
    RYR_S_V,          /* membrane potential after the last step, mV */
    RYR_S_XM, RYR_S_XH, RYR_S_XJ, RYR_S_XR, RYR_S_XS1, RYR_S_XS2, RYR_S_QKS,
    RYR_S_XTOS, RYR_S_YTOS, RYR_S_XTOF, RYR_S_YTOF, RYR_S_XNAI,
    RYR_S_CSUBT,      /* mean cs over the interior */
    RYR_S_ICAL,       /* xicatto, uA/uF */
    RYR_S_INCX,       /* xinacato, uA/uF */
    RYR_S_INA, RYR_S_IKS, RYR_S_IKR, RYR_S_IK1, RYR_S_INAK, RYR_S_ITOS, RYR_S_ITOF,
    RYR_S_STIM,
    RYR_S_OUT_ICA, RYR_S_OUT_NCX, RYR_S_OUT_INA, /* sums since the last output row */
    RYR_S_STEP,       /* completed steps (exact integer in a double) */
    RYR_S_COUNT
};

/* Per-CRU state fields for ryr_read_field / ryr_write_field.  Sub-voxel
 * fields hold 8 planes: out[ii*N + n] with ii the reference's sub-voxel
 * index (bit0 z, bit1 y, bit2 x; :1322-1327) and n = i + nx*(j + ny*k) the
 * interior CRU index.  Integer fields are returned as exact doubles. */
enum ryr_field {
    RYR_F_CI = 0,  RYR_F_CNSR,                           /* 8 planes each */
    RYR_F_CATF, RYR_F_CATS, RYR_F_CACAL, RYR_F_CASR, RYR_F_CAMYO, RYR_F_MGMYO, /* 8 planes */
    RYR_F_CS, RYR_F_CP, RYR_F_CJSR, RYR_F_CJSRB, RYR_F_ANCX,
    RYR_F_CASAR, RYR_F_CASARH, RYR_F_CASARJ, RYR_F_CASARHJ,
    RYR_F_NOU, RYR_F_NCU, RYR_F_NOB, RYR_F_NCB,
    RYR_F_LCC,                                             /* 8 planes, 16 = absent */
    RYR_F_NL, RYR_F_NSPARK,
    RYR_F_COUNT
};

/* Output-row columns (ryr_read_diag): the 45 columns of wholetest_<name>
 * (cuda_Cell_Detailed.cu:1041-1072) in order, as doubles. */
#define RYR_DIAG_COLS 45
/* Linescan columns (cuda_Cell_Detailed.cu:974-989), one row per x. */
#define RYR_LINE_COLS 22

/* Per-beat summary rows (ryr_read_beats).  The reference records none of
 * these (its only per-interval statistic is the spark flag, :1283-1284,
 * :901); they are defined here:
 *   onset     stimulus onset of the beat (ms)
 *   apd90     time from onset to the first downward crossing of
 *             Vmax - 0.9 (Vmax - Vrest), linearly interpolated (ms; -1 if none)
 *   vmax      peak V of the beat; vrest = V at onset (mV)
 *   ead       1 if, at least 50 ms after onset and once V has repolarised by
 *             40% of the amplitude (past the notch and dome), V rises
 *             again by >= 1 mV above its running minimum while above -40 mV
 *             (dV/dt > 0 during repolarisation, with a 1 mV noise floor)
 *   ead_t     time of that rise from onset (ms; -1 if none)
 *   peak_ci   largest whole-cell mean ci over the beat's output rows (uM)
 *   jsr_load  whole-cell mean cjsr at the output row of the onset (uM)
 *   sparks    sum over the beat's output rows of the CRUs with a spark onset
 *   complete  1 for a finished beat, 0 for the beat in progress
 *   ead_lit   1 if dV/dt > 0 above -40 mV during repolarisation, read
 *             literally: once V has started to fall after the upstroke
 *             (the first step after the beat's V exceeded -40 mV at which V
 *             decreases), any later step of the beat at which V rises and
 *             ends above -40 mV.  A notch-and-dome or any rise on the plateau
 *             counts; ead above is the flag with a time and size floor
 *   ead_lit_t time of that step's end from onset (ms; -1 if none) */
#define RYR_BEAT_COLS 12
#define RYR_MAX_BEATS 64
enum ryr_beat_col {
    RYR_B_ONSET = 0, RYR_B_APD90, RYR_B_VMAX, RYR_B_VREST, RYR_B_EAD, RYR_B_EAD_T,
    RYR_B_PEAK_CI, RYR_B_JSR_LOAD, RYR_B_SPARKS, RYR_B_COMPLETE, RYR_B_EAD_LIT, RYR_B_EAD_LIT_T
};

/* Soft-error counters per cell (ryr_read_counters). */
enum ryr_counter {
    RYR_C_GAUSS_FALLBACK = 0, /* >200 rejections in a Gaussian draw (subroutine_cuda.c:358-361...) */
    RYR_C_CLAMP_CP,           /* cp < 0 clamped (:1466-1467) */
    RYR_C_CLAMP_CS,           /* cs < 0 set to 1e-6 (:1469-1472) */
    RYR_C_CI_RUNAWAY,         /* ci > 1000 uM at an output row (:878-880) */
    RYR_C_NONFINITE_V,        /* V not finite after a step */
    RYR_C_COUNT
};

/* Fill with the reference's values (run.sh's `make il` build: iso2 + lqt). */
void ryr_params_default(ryr_params *p);
/* run.sh:9 "stab" argv values; seed 1; 62x26x10 interior. */
void ryr_cell_desc_default(ryr_cell_desc *d);

/* Library/ABI identity. */
int ryr_abi_version(void);

/* Allocate and initialise ncells independent cells on GPU `device`.
 * Replaces main()'s setup: cudaSetDevice (:242), the allocations (:259-320),
 * the host membrane ICs (:350-366), Init<<<>>> (:373, :1108-1182) and
 * setup_kernel (:377, :1695-1701). */
int ryr_create(int device, const ryr_params *p, const ryr_cell_desc *cells, int ncells,
               ryr_handle **out);

/* Advance every cell by nsteps time steps of dt.  Replaces one iteration of
 * the time loop per step (:399-1085): the Compute launch (:420), the buffer
 * swap (:422-425), the D2H copies and serial sums (:427-456, here an on-device
 * compensated reduction) and the membrane ODE (:458-739, here on device).
 * Output rows (every steps_per_output steps, :777) are reduced on device and
 * kept; fetch them with ryr_read_diag before the next output step. */
int ryr_step(ryr_handle *h, int64_t nsteps);

/* ryr_step splits its steps into segments that end at output steps and
 * replays each segment from a captured HIP graph: a whole output interval is
 * captured at its first use, a shorter segment once the same length has been
 * asked for before.  ryr_prepare captures now, without advancing anything,
 * every graph that ryr_step(h, nsteps) called next would replay, so that
 * call launches graphs only (a caller timing a few steps, or a loop of equal
 * chunks, prepares once).  At most 8 graphs are kept; RYR_E_STATE while
 * the launch shape is still being tuned.  No reference counterpart (the
 * reference launches Compute once per step, cuda_Cell_Detailed.cu:420). */
int ryr_prepare(ryr_handle *h, int64_t nsteps);

/* Block until queued work on the handle's stream is done. */
int ryr_sync(ryr_handle *h);

/* Scalars of every cell: out[c*RYR_S_COUNT + s]. (:350-366, :430-739) */
int ryr_read_scalars(ryr_handle *h, double *out);

/* One state field of one cell (layout in enum ryr_field).  Replaces the
 * output-interval cudaMemcpy of CRU2/CYT/CBU/SBU/ci/cnsr (:780-785). */
int ryr_read_field(ryr_handle *h, int cell, int field, double *out, int64_t n);
/* Overwrite one state field of one cell (test/restart hook; same layout). */
int ryr_write_field(ryr_handle *h, int cell, int field, const double *in, int64_t n);
/* Overwrite one membrane scalar of a cell (s in RYR_S_V..RYR_S_XNAI). */
int ryr_write_scalar(ryr_handle *h, int cell, int s, double value);

/* Latest output row of every cell: out[c*RYR_DIAG_COLS + col]; *step_out gets
 * the step index the row belongs to (-1 if no row yet).  Columns are those of
 * wholetest_<name> (:1041-1072); the reductions are the interior loops of
 * :839-962, :1031-1039. */
int ryr_read_diag(ryr_handle *h, double *out, int64_t *step_out);

/* Linescan rows of the latest output step for one cell: nx rows of
 * RYR_LINE_COLS (:967-999), taken at y = ny/2, z = min(5, nz) in the
 * reference's ghost-inclusive numbering. */
int ryr_read_linescan(ryr_handle *h, int cell, double *out, int64_t n);

/* Beat summaries of one cell: up to maxbeats rows of RYR_BEAT_COLS, the
 * completed beats in order followed by the beat in progress (if any);
 * *nbeats gets the number of rows written.  Beats past RYR_MAX_BEATS are
 * not recorded. */
int ryr_read_beats(ryr_handle *h, int cell, double *out, int maxbeats, int *nbeats);

/* Voltage-clamp protocols (cuda_Cell_Detailed.cu:733-773: vclamp,
 * brownvclamp, apclamp reading one period from v_1000.txt, linearapclamp).
 * After the membrane ODE of step s the cell's V is replaced by
 * v[s % n] (periodic != 0) or v[min(s, n-1)]; the gates, currents and [Na]i
 * still follow the ODE, as in the reference.  n = 0 switches the clamp off. */
int ryr_set_vclamp(ryr_handle *h, int cell, const double *v, int64_t n, int periodic);

/* Soft-error counters: out[c*RYR_C_COUNT + k] (see enum ryr_counter). */
int ryr_read_counters(ryr_handle *h, int64_t *out);

/* Total CRU-steps advanced so far (all cells). */
int ryr_cru_steps(ryr_handle *h, int64_t *out);

/* Elapsed GPU time (ms) of the step kernels recorded with HIP events on the
 * handle's stream since the last call; *launches gets how many step kernels
 * ran.  Used by the benchmark for the per-launch kernel duration. */
int ryr_kernel_time(ryr_handle *h, double *ms, int64_t *launches);

/* Launch shape of a step: the workgroup sizes (64, 128 or 256 lanes) of the
 * two CRU-parallel kernels (the sub-voxel cytosol/NSR kernel and the CRU
 * kernel; for the CRU kernel also 512: 256-lane workgroups with two lanes per
 * CRU, which split the RyR transitions and the LCC channels between them, for
 * grids too small to fill the GPU; for the sub-voxel kernel also 1024: one
 * lane per CRU holding its eight sub-voxels, so that the stencil partners
 * inside a CRU and along x come from registers and neighbouring lanes, for
 * grids that fill the GPU, sequential and concurrent schedules only), and
 * the schedule of a step's kernels:
 *   RYR_SCHED_SEQ         sub-voxel kernel, CRU kernel, membrane step
 *   RYR_SCHED_FUSED       the membrane step of step n rides in the sub-voxel
 *                         launch of step n+1
 *   RYR_SCHED_CONCURRENT  the sub-voxel and CRU kernels of a step run side by
 *                         side on two streams (cleft scheme A only)
 *   RYR_SCHED_FUSED_GATING as FUSED, and the RyR gating of step n (which reads
 *                         only step-start values) rides in the sub-voxel launch
 *                         of step n too, off the CRU kernel (CRU size 128/256/512)
 *   RYR_SCHED_ONE         one launch per step (cleft scheme A, CRU size 256;
 *                         the sub-voxel size is not used): the sweep, the CRU
 *                         update and a membrane workgroup per cell side by side.
 *                         The CRU workgroups form V themselves from the previous
 *                         step's sums and the membrane values its launch left,
 *                         so no launch waits for the membrane step; the
 *                         membrane workgroup finishes the previous step
 * No reference counterpart (the reference hard-codes one 48x448 launch,
 * :1701).  A new handle tunes the shape itself: its first 9 steps alternate
 * the sizes with each kernel timed by its own events (fastest median kept);
 * then each schedule candidate (the tuned sizes with every schedule, the fused
 * schedule with 256-lane sub-voxel workgroups, the two-lanes-per-CRU kernel,
 * and the concurrent schedule with smaller sizes and with the one-lane-per-CRU
 * sweep) runs 24 steps as three replays of a captured 8-step graph, timed by
 * events, and the fastest is kept.  Output steps are stepped over eagerly, so
 * tuning takes about 450 steps.  Every
 * shape computes identical values, so tuning never changes a result.  Passing
 * sizes (0, 0) re-arms the tuning; the getter reports sizes 0 and schedule -1
 * while it is pending. */
#define RYR_SCHED_SEQ        0
#define RYR_SCHED_FUSED      1
#define RYR_SCHED_CONCURRENT 2
#define RYR_SCHED_FUSED_GATING 3
#define RYR_SCHED_ONE        4
int ryr_set_launch_shape(ryr_handle *h, int wg_cyto, int wg_cru, int sched);
int ryr_get_launch_shape(const ryr_handle *h, int *wg_cyto, int *wg_cru, int *sched);

/* Advances nsteps steps in the tuned schedule with every launch eager (no
 * graph) and stamped by HIP events at its start and end; avg_us[k] / count[k]
 * receive the mean duration (microseconds) and launch count of kernel k:
 * 0 sub-voxel, 1 CRU, 2 membrane, 3 output row.  Each launch is the one
 * production makes: under the concurrent schedule the CRU kernel runs beside
 * the sub-voxel kernel, and under the fused schedules the sub-voxel launch
 * carries the previous step's membrane step (and the RyR gating), so k = 2
 * counts only the membrane launches of output steps and of the last step.
 * The steps are ordinary simulation steps.  RYR_E_STATE while the shape is
 * being tuned.  Used by the benchmark for per-kernel roofline figures. */
#define RYR_NPROF 4
int ryr_profile_kernels(ryr_handle *h, int64_t nsteps, double *avg_us, int64_t *count);

/* Launch floor of the production step: the graph ryr_step would replay for
 * min(nsteps, steps_per_output) steps under the current launch shape, with
 * every kernel replaced by an empty kernel of the same grid and workgroup,
 * replayed and timed by events; *us_per_step gets the best replay's time per
 * step (microseconds).  No simulation step is taken.  RYR_E_STATE while the
 * shape is being tuned.  Measurement only; no reference counterpart. */
int ryr_latency_floor(ryr_handle *h, int64_t nsteps, double *us_per_step);

/* Snapshot/restore of the whole device state (all cells, all fields,
 * membrane scalars, step counter).  The reference has none; with a
 * counter-based generator the random stream position is the step index, so a
 * restored handle continues bit-identically.  ryr_load returns RYR_E_STATE,
 * before it writes any state, unless the snapshot was saved by a handle of
 * the same precision, cells, lattice shapes, parameters, variants and seeds. */
int ryr_save(ryr_handle *h, const char *path);
int ryr_load(ryr_handle *h, const char *path);

void ryr_destroy(ryr_handle *h);

/* Message of the last failure on h (or of the last failed ryr_create when h
 * is NULL). */
const char *ryr_last_error(const ryr_handle *h);

/* Test hook: run the RyR (subroutine_cuda.c:251-669) and LCC
 * (reesica.h:14-116, loop :1248-1261) samplers on caller-supplied frozen
 * inputs, on the GPU, with the Philox stream of (seed, cru[q], step).
 * Inputs per query q (n queries): counts[4q..4q+3] = nou, ncu, nob, ncb;
 * lcc[8q..8q+7] channel states (16 = absent); conc[4q..4q+3] = cp, cjsr, cs, v.
 * Scalars: Ku, Ku2, Kb, dt, iso2.  Outputs: counts and lcc overwritten with the
 * new states, draws[q] = slots consumed, nopen[q] = open LCCs,
 * fails[q] = Gaussian fallbacks. */
int ryr_gating_probe(int device, int64_t n, uint64_t seed, const uint32_t *cru, uint64_t step,
                     double Ku, double Ku2, double Kb, double dt, int iso2,
                     int32_t *counts, int32_t *lcc, const double *conc,
                     int32_t *draws, int32_t *nopen, int32_t *fails);

/* ryr_gating_probe without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as ryr_gating_probe; the device copies of the
 * queries are staged in a fixed device array of the library (about 1.5 MiB,
 * held by each GPU it runs on until the process ends) instead of buffers
 * allocated per call, and the call makes no hipMalloc or hipFree.  At most
 * RYR_PROBE_NOALLOC_MAX queries per call (more: RYR_E_ARG).  Every call on
 * one GPU uses that GPU's one array: calls that overlap on the same GPU (from
 * several host threads) race and must be serialised by the caller. */
#define RYR_PROBE_NOALLOC_MAX 16384
int ryr_gating_probe_noalloc(int device, int64_t n, uint64_t seed, const uint32_t *cru, uint64_t step,
                             double Ku, double Ku2, double Kb, double dt, int iso2,
                             int32_t *counts, int32_t *lcc, const double *conc,
                             int32_t *draws, int32_t *nopen, int32_t *fails);

#ifdef __cplusplus
}
#endif

#endif /* RYR_CRU_H */
