/*
 * ryr_oracle.h - FP64 CPU oracle of the CRU update.  TEST INFRASTRUCTURE
 * ONLY: the test suite, the smoke check and the benchmark's CPU-baseline leg
 * compare results with it; the product path never links it.
 *
 * It restates the reference's algorithm function by function (citations in
 * ryr_oracle.c) in the reference's own data layout: ghost-shell lattice,
 * pos(i,j,k)*8+ii sub-voxel indexing, explicit boundary copies.  It shares
 * only two arithmetic primitives with the GPU code: the Philox4x32-10
 * generator and the deterministic exp/log/cos/pow (both pinned by their own
 * tests), so that identical inputs give identical random draws and
 * thresholds.
 */
#ifndef RYR_ORACLE_H
#define RYR_ORACLE_H

#include <stdint.h>
#include "../include/ryr_cru.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cell orc_cell;

orc_cell *orc_create(const ryr_params *p, const ryr_cell_desc *d);
void orc_destroy(orc_cell *c);
void orc_step(orc_cell *c, int64_t nsteps);
int orc_read_scalars(const orc_cell *c, double *out);
int orc_read_field(const orc_cell *c, int field, double *out, int64_t n);
int orc_write_field(orc_cell *c, int field, const double *in, int64_t n);
int orc_write_scalar(orc_cell *c, int s, double v);
int orc_read_diag(const orc_cell *c, double *out, int64_t *step);
int orc_read_linescan(const orc_cell *c, double *out, int64_t n);
int orc_read_counters(const orc_cell *c, int64_t *out);
int orc_num_threads(void);
void orc_set_num_threads(int n);   /* OpenMP threads of the CRU loop (n > 0) */
int orc_set_vclamp(orc_cell *c, const double *v, int64_t n, int periodic);
int orc_read_beats(const orc_cell *c, double *out, int maxbeats, int *nbeats);

/* leaf functions (unit tests) */
double orc_ncx(double cs_mM, double v, double xnai, double *Ancx, double dt);
/* per interior CRU, the last scheme-B step's sum over sub-steps of
 * (step-start cs - sub-step cs): the term by which scheme B departs from
 * calcium conservation (test hook) */
int orc_read_scheme_b_lag(const orc_cell *c, double *out, int64_t n);
/* the same under slowncxchange (exchanger at 5 x the allosteric factor) */
double orc_ncx_slow(double cs_mM, double v, double xnai, double *Ancx, double dt);
double orc_uptake(double ci, double cnsr);
double orc_leak(double cnsr, double ci);
double orc_luminal(double cjsr);
double orc_luminal_SS(double cjsr);
double orc_release(double po, double cjsr, double cp);
double orc_lcccurrent(double v, double cp);
double orc_buffer_flux(int which, double b, double b2, double c, double dt);
double orc_fracbound(double cjsr);
int orc_lcc_markov(int state, double cpx, double v, double dt, double u, int iso2);
double orc_canonical_sum(const double *x, int64_t n);
int orc_ryrgating_replay(const uint64_t *slots, int nslots, double Ku, double Ku2, double Kb,
                         double cp, double cjsr, double dt, int32_t *counts,
                         int32_t *used, int32_t *fails);
int orc_gating_probe(int64_t n, uint64_t seed, const uint32_t *cru, uint64_t step,
                     double Ku, double Ku2, double Kb, double dt, int iso2,
                     int32_t *counts, int32_t *lcc, const double *conc,
                     int32_t *draws, int32_t *nopen, int32_t *fails);
/* membrane ODE of one step on a standalone state (tests of :458-739) */
int orc_membrane_probe(double *state13, double sum_icat, double sum_inaca, double sum_cs,
                       double ncru, double cm, double dt, double ikr_scale, double stim,
                       double *currents);

double orc_dm_exp(double x);
double orc_dm_log(double x);
double orc_dm_cos(double x);
double orc_dm_pow(double x, double y);
void orc_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out);
double orc_slot_uniform(uint32_t a, uint32_t b);
uint32_t orc_sr_bits(uint64_t seed, uint64_t step, uint32_t field, uint32_t n, uint32_t ii);
uint32_t orc_mix32(uint32_t x);
float orc_sr_f32(double v, uint32_t r);
void orc_stream(uint64_t seed, uint32_t cru, uint64_t step, int n, double *u, uint32_t *w);
int orc_stencil_probe(const ryr_params *p, const ryr_cell_desc *d, const double *field,
                      double tt, double tl, double *out);

#ifdef __cplusplus
}
#endif
#endif
