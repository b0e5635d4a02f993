/*
 * croma_host.h -- internal declarations of the C host runtime of
 * libcroma_hip.so (not installed; the public ABI is include/croma_hip.h).
 */
#ifndef CROMA_HOST_H
#define CROMA_HOST_H

#include <stddef.h>

#include "../csrc/croma_kernels.h"
#include "../../include/croma_hip.h"

#define CH_NMAX 512

/* Host copy of one grid and its per-grid caches. */
typedef struct {
    int n;
    int j_pp;
    double pmin, pmax, pemin, pemax;
    double pp[CH_NMAX], pdp[CH_NMAX], pE[CH_NMAX], beta_p[CH_NMAX];
    double p_p2[CH_NMAX], p_expcut[CH_NMAX], sigma_pp[CH_NMAX], sigmoid[CH_NMAX], p_sig[CH_NMAX];
    double p_ecr_w[CH_NMAX];
    double ep[CH_NMAX], edp[CH_NMAX], eE[CH_NMAX], e_p2[CH_NMAX], e_expcut[CH_NMAX], e_ecr_w[CH_NMAX];
    double ppm1, ppp1, epm1, epp1;
} ChGrid;

/* croma_physics.c */
int ch_host_threads(void);
void ch_parallel_for(long n, long min_chunk, void (*fn)(void *, long, long), void *arg);
double ch_pp_pion_cross_section(int kind, double p);
int ch_grid_build(ChGrid *g, int n, double pmin, double pmax, double pemin, double pemax);
double ch_b_coulomb_p(double nicm, double p, double kBT);
int ch_nsub_estimate(const ChGrid *g, const CkParams *prm, long ntr, double dt_snap, double ric_lo,
                     double ric_hi, const double *n_gas, const double *kbt, const double *b_field,
                     const double *divv, const double *l_turb, const double *dv, int *nsub, int *n_on);
int ch_choose_segments(const CkParams *prm, int nsub, int allow);
int ch_ic_cooling_row(const ChGrid *g, double z, double *row, double *m1, double *p1);
/* CMB photon-energy bounds of the IC integral at z (FP_Coef.c:413-436). */
void ch_ic_bounds(double z, double *e0min, double *e0max);
/* IC cooling rate of the electron bins bins[0..count) at z (same arithmetic
 * as ch_ic_cooling_row, which fills every bin). */
int ch_ic_cooling_bins(const ChGrid *g, double z, int count, const int *bins, double *out);

/* croma_host.c */
double ch_now_s(void);
void ch_set_error(const char *fmt, ...);
void ch_params_to_ck(const CromaHipParams *in, CkParams *out);

/* context internals shared with the serving loop */
#define CH_HIST 8192
const ChGrid *ch_ctx_grid(const CromaHipContext *c);
const CkGrid *ch_ctx_dgrid(const CromaHipContext *c);
const CkParams *ch_ctx_prm(const CromaHipContext *c);
int ch_ctx_cu_count(const CromaHipContext *c);
CromaHipTile ch_ctx_tile(const CromaHipContext *c, int protons);
int ch_ctx_coeff_fp32(const CromaHipContext *c); /* FP32 coefficient evaluation selected */
int ch_ctx_fqe(CromaHipContext *c);              /* builds the secondary weights once */
const double *ch_ctx_d_fqe(const CromaHipContext *c);
const int *ch_ctx_d_npmin(const CromaHipContext *c);
int ch_ctx_jp0(const CromaHipContext *c);
int ch_ctx_plain_ws(CromaHipContext *c, int count, CkPlainWs *out); /* plain-path workspace of >= count tracers */
void *ch_ctx_stream(const CromaHipContext *c);
int *ch_ctx_d_err(const CromaHipContext *c);
int ch_ctx_take_err(CromaHipContext *c); /* reads and clears the error word: 0, EPIVOT or ENONFINITE */
void ch_ctx_ic_ends(CromaHipContext *c, double z, double ends[2]); /* IC rate at bins 0, n-1 (cached) */
int ch_cost_bucket(double cost);                 /* log-scale bucket in [0, CH_HIST) */

/* device-resident state of a batch (read by the file writers) */
typedef struct {
    double *d_cre, *d_crp; /* [ntr * n]; d_crp is NULL for electron-only batches */
    int ntr, n, protons;
    void *stream;
} ChBatchState;
int ch_batch_state(CromaHipBatch *b, ChBatchState *out);

#endif
