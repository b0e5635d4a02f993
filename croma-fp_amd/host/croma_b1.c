/*
 * croma_b1.c -- the reference's GPU backend surface (fp_cuda_backend.h:10-159)
 * on this library: run_tracer_pipeline's bucket-segment update, the rank
 * state, the batched solve and the emission entry points, taking the
 * reference's structs (restated in include/croma_hip_b1.h).
 *
 * A bucket segment (the reference's run_cached_simple,
 * fp_cuda_tracer_backend.cu:335-830, as called from
 * include/tracer_fp_solve.inc:1418-1471) runs on the plain path's kernels:
 * backgrounds (lerped on the device when device_interp_background is set),
 * loss rates, D_pp on / off and the Chang-Cooper rows of every tracer of the
 * bucket, the secondary source of its segment-start protons, then nstep
 * implicit substeps of protons and electrons with the on / off window.  The
 * bucket's spectra stay in the context's plain workspace between calls, so
 * a segment without state inputs continues the previous call's state, as the
 * reference's resident scratch does.
 */
#include "croma_host.h"

#include "../../include/croma_hip_b1.h"

#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

static pthread_mutex_t g_b1_lock = PTHREAD_MUTEX_INITIALIZER;
static CromaHipParams g_b1_prm;
static int g_b1_prm_set;

int croma_hip_b1_set_params(const CromaHipParams *prm)
{
    if (!prm) return -1;
    pthread_mutex_lock(&g_b1_lock);
    g_b1_prm = *prm;
    g_b1_prm_set = 1;
    pthread_mutex_unlock(&g_b1_lock);
    return 0;
}

static CromaHipParams b1_params(void)
{
    CromaHipParams p;
    pthread_mutex_lock(&g_b1_lock);
    if (!g_b1_prm_set) {
        croma_hip_default_params(&g_b1_prm);
        g_b1_prm_set = 1;
    }
    p = g_b1_prm;
    pthread_mutex_unlock(&g_b1_lock);
    return p;
}

/* View of a CRspectrum of n bins (params.h:52-61): p[n], dp[n], pp1, pm1,
 * E[n], N[n], N_pre[n], N_ave[n]. */
typedef struct {
    const double *p, *dp, *E;
    double pp1, pm1;
} B1Spec;

static B1Spec b1_spec(const void *grid, int n)
{
    const double *d = (const double *)grid;
    B1Spec s;
    s.p = d;
    s.dp = d + n;
    s.pp1 = d[2 * n];
    s.pm1 = d[2 * n + 1];
    s.E = d + 2 * n + 2;
    return s;
}

/* Grid bounds of momentum_bin (FP_Coef.c:291-330) from its bins: p_j =
 * 10^(lo + j d), d = (hi - lo) / n; the bounds are the configured decimal
 * values, recovered to 1e-9 and then checked by rebuilding the grid. */
static void b1_bounds(const B1Spec *s, int n, double *lo, double *hi)
{
    const double d = log10(s->p[n - 1] / s->p[0]) / (double)(n - 1);
    const double l = log10(s->p[0]);
    *lo = round(l * 1.0e9) / 1.0e9;
    *hi = round((l + d * (double)n) * 1.0e9) / 1.0e9;
}

static int b1_same_grid(const CromaHipContext *c, int species, const B1Spec *s, int n)
{
    double p[CH_NMAX], dp[CH_NMAX], E[CH_NMAX], pm1, pp1;
    if (croma_hip_context_grid(c, species, p, dp, E, &pm1, &pp1)) return 0;
    return !memcmp(p, s->p, (size_t)n * sizeof(double)) && !memcmp(dp, s->dp, (size_t)n * sizeof(double)) &&
           pm1 == s->pm1 && pp1 == s->pp1;
}

/* A context for the reference's grids (cached by n and bounds). */
typedef struct {
    CromaHipContext *ctx;
    int n;
    double bounds[4];
} B1Ctx;

static int b1_ctx_for(B1Ctx *slot, int np_, int npe_, const void *crp_grid, const void *cre_grid)
{
    B1Spec sp, se;
    double b[4];
    if (np_ != npe_ || !(np_ == 64 || np_ == 128 || np_ == 256 || np_ == 512)) {
        ch_set_error("b1: np = npe in {64, 128, 256, 512} required (np %d, npe %d)", np_, npe_);
        return -1;
    }
    if (!crp_grid || !cre_grid) {
        ch_set_error("b1: grids required");
        return -1;
    }
    sp = b1_spec(crp_grid, np_);
    se = b1_spec(cre_grid, npe_);
    b1_bounds(&sp, np_, &b[0], &b[1]);
    b1_bounds(&se, npe_, &b[2], &b[3]);
    if (slot->ctx && (slot->n != np_ || memcmp(slot->bounds, b, sizeof(b)))) {
        croma_hip_context_destroy(slot->ctx);
        slot->ctx = 0;
    }
    if (!slot->ctx) {
        CromaHipGridSpec gs;
        CromaHipParams prm = b1_params();
        gs.n = np_;
        gs.pmin = b[0];
        gs.pmax = b[1];
        gs.pemin = b[2];
        gs.pemax = b[3];
        slot->ctx = croma_hip_context_create(&gs, &prm);
        if (!slot->ctx) return -1;
        slot->n = np_;
        memcpy(slot->bounds, b, sizeof(b));
        if (!b1_same_grid(slot->ctx, 0, &sp, np_) || !b1_same_grid(slot->ctx, 1, &se, npe_)) {
            ch_set_error("b1: the CRspectrum grids are not momentum_bin grids of decimal bounds");
            croma_hip_context_destroy(slot->ctx);
            slot->ctx = 0;
            return -1;
        }
    } else {
        CromaHipParams prm = b1_params();
        if (croma_hip_context_set_params(slot->ctx, &prm)) return -1;
    }
    return 0;
}

struct CromaB1Workspace {
    B1Ctx c;
    int cap;              /* tracers of the bucket buffers */
    double *d_stage;      /* [4][cap * n] cell-major staging: crp, cre, qpi, qepri */
    double *d_bg;         /* [2][CK_NBG][cap]: current, next */
    double *d_ric;        /* [3][n * cap]: current, next, lerped */
    int *d_ident;         /* [cap] 0..cap-1 */
    int *d_idx;           /* [cap] rank indices */
    int *h_ident;
    int resident_ncell;   /* bucket state in the plain workspace (0: none) */
    int rank_cap, rank_ncell, rank_n;
    double *d_rank;       /* [4][rank_cap * rank_n]: qpi, qepri, crp, cre */
};

CromaB1Workspace *croma_hip_b1_workspace_create(void)
{
    return (CromaB1Workspace *)calloc(1, sizeof(CromaB1Workspace));
}

static void b1_free_bucket(CromaB1Workspace *w)
{
    ck_free(w->d_stage);
    ck_free(w->d_bg);
    ck_free(w->d_ric);
    ck_free(w->d_ident);
    ck_free(w->d_idx);
    free(w->h_ident);
    w->d_stage = w->d_bg = w->d_ric = 0;
    w->d_ident = w->d_idx = w->h_ident = 0;
    w->cap = 0;
}

void croma_hip_b1_workspace_destroy(CromaB1Workspace *w)
{
    if (!w) return;
    b1_free_bucket(w);
    ck_free(w->d_rank);
    if (w->c.ctx) croma_hip_context_destroy(w->c.ctx);
    free(w);
}

static int b1_ensure_bucket(CromaB1Workspace *w, int ncell, int n)
{
    int i;
    void *st = ch_ctx_stream(w->c.ctx);
    if (w->cap >= ncell) return 0;
    ck_stream_sync(st);
    b1_free_bucket(w);
    w->d_stage = (double *)ck_malloc((size_t)4 * ncell * n * sizeof(double));
    w->d_bg = (double *)ck_malloc((size_t)2 * CK_NBG * ncell * sizeof(double));
    w->d_ric = (double *)ck_malloc((size_t)3 * n * ncell * sizeof(double));
    w->d_ident = (int *)ck_malloc((size_t)ncell * sizeof(int));
    w->d_idx = (int *)ck_malloc((size_t)ncell * sizeof(int));
    w->h_ident = (int *)malloc((size_t)ncell * sizeof(int));
    if (!w->d_stage || !w->d_bg || !w->d_ric || !w->d_ident || !w->d_idx || !w->h_ident) {
        b1_free_bucket(w);
        ch_set_error("b1: bucket buffers: %s", ck_last_error());
        return -1;
    }
    for (i = 0; i < ncell; i++) w->h_ident[i] = i;
    if (ck_h2d(w->d_ident, w->h_ident, (size_t)ncell * sizeof(int), st) || ck_stream_sync(st)) return -1;
    w->cap = ncell;
    w->resident_ncell = 0; /* the plain workspace may have moved */
    return 0;
}

int croma_hip_b1_rank_state_upload(int np_, int npe_, CromaB1Workspace *w, int ncell, const double *qpi,
                                   const double *qepri, const double *crp, const double *cre, double *elapsed_ms)
{
    const double t0 = ch_now_s();
    size_t m;
    void *st;
    if (elapsed_ms) *elapsed_ms = 0.0;
    if (!w || ncell < 0 || !qpi || !qepri || !crp || !cre || np_ != npe_ || np_ <= 0) return -1;
    m = (size_t)ncell * np_;
    /* the rank state needs no grid: before the first pipeline call it goes
     * through the null stream (the copies complete before the call returns) */
    st = w->c.ctx ? ch_ctx_stream(w->c.ctx) : 0;
    if (w->rank_cap < ncell || w->rank_n != np_) {
        ck_stream_sync(st);
        ck_free(w->d_rank);
        w->d_rank = (double *)ck_malloc(4 * m * sizeof(double));
        w->rank_cap = w->d_rank ? ncell : 0;
        w->rank_n = np_;
        if (!w->d_rank) return -1;
    }
    {
        const size_t cap = (size_t)w->rank_cap * np_;
        if (ck_h2d(w->d_rank, qpi, m * 8, st) || ck_h2d(w->d_rank + cap, qepri, m * 8, st) ||
            ck_h2d(w->d_rank + 2 * cap, crp, m * 8, st) || ck_h2d(w->d_rank + 3 * cap, cre, m * 8, st) ||
            ck_stream_sync(st))
            return -1;
    }
    w->rank_ncell = ncell;
    if (elapsed_ms) *elapsed_ms = 1e3 * (ch_now_s() - t0);
    return 0;
}

int croma_hip_b1_rank_state_download(int np_, int npe_, CromaB1Workspace *w, int ncell, double *crp, double *cre,
                                     double *elapsed_ms)
{
    const double t0 = ch_now_s();
    void *st;
    size_t m, cap;
    if (elapsed_ms) *elapsed_ms = 0.0;
    if (!w || !crp || !cre || np_ != npe_ || !w->d_rank || np_ != w->rank_n || ncell < 0 || ncell > w->rank_ncell)
        return -1;
    st = w->c.ctx ? ch_ctx_stream(w->c.ctx) : 0;
    m = (size_t)ncell * np_;
    cap = (size_t)w->rank_cap * np_;
    if (ck_d2h(crp, w->d_rank + 2 * cap, m * 8, st) || ck_d2h(cre, w->d_rank + 3 * cap, m * 8, st) ||
        ck_stream_sync(st))
        return -1;
    if (elapsed_ms) *elapsed_ms = 1e3 * (ch_now_s() - t0);
    return 0;
}

static int b1_wants_emission(const CromaB1OutputSchedule *s)
{
    return s && (s->emit_final_only || s->emit_all_steps || s->emit_stride > 0 ||
                 (s->selected_steps && s->nselected_steps > 0));
}

int croma_hip_b1_run_tracer_pipeline(int np_, int npe_, const CromaB1PipelineInput *in,
                                     const CromaB1OutputSchedule *schedule, double *crp_out, double *cre_out,
                                     double *eps_syn_out, double *eps_gamma_out, double *eps_nu_out,
                                     CromaB1PipelineTimes *times)
{
    const double t_start = ch_now_s();
    CromaB1Workspace *w;
    const CromaB1CoeffBatchInput *ci;
    CromaHipContext *c;
    CkPlainWs pw;
    CkSlot S;
    CkParams prm;
    void *st;
    int n, ncell, have_in, any_in, use_rank, interp, f, rc;
    double t_mark;
    CromaB1PipelineTimes tl;
    memset(&tl, 0, sizeof(tl));
    if (times) memset(times, 0, sizeof(*times));
    if (!in || !in->coeff_in || !in->workspace || in->ncell <= 0 || in->nstep < 0) return -1;
    /* as the reference (fp_cuda_tracer_backend.cu:1542-1543): no gamma / nu here */
    if (eps_gamma_out || eps_nu_out || in->gamma_in || in->neutrino_in) return -1;
    if (b1_wants_emission(schedule) || eps_syn_out) {
        ch_set_error("b1: emission inside the pipeline is not served (use the emission entry points)");
        return -1;
    }
    if ((in->fp_cadence_steps > 1) || (in->diff_cadence_steps > 1) || (in->adv_cadence_steps > 1)) {
        ch_set_error("b1: cadences other than 1 are not served");
        return -1;
    }
    w = in->workspace;
    ci = in->coeff_in;
    ncell = in->ncell;
    have_in = in->qpi_batch && in->qepri_batch && in->crp_init && in->cre_init;
    any_in = in->qpi_batch || in->qepri_batch || in->crp_init || in->cre_init;
    use_rank = in->gather_rank_state || in->scatter_rank_state;
    if (any_in && !have_in) return -1;
    if ((crp_out != 0) != (cre_out != 0)) return -1;
    if (use_rank && (!in->rank_state_indices || !w->d_rank || w->rank_n != np_)) return -1;
    if (use_rank) {
        /* every index addresses a row of the uploaded rank state */
        int i;
        for (i = 0; i < ncell; i++)
            if (in->rank_state_indices[i] < 0 || in->rank_state_indices[i] >= w->rank_ncell) {
                ch_set_error("b1: rank_state_indices[%d] = %d outside the rank state of %d tracers", i,
                             in->rank_state_indices[i], w->rank_ncell);
                return -1;
            }
    }
    if (in->gather_rank_state && any_in) return -1;
    if (!ci->n_gas || !ci->kbt || !ci->b_field || !ci->divv_gyr || !ci->l_turb_mpc || !ci->dv_imc || !ci->cs ||
        !ci->beta_pl || !ci->rad_ic_batch)
        return -1;
    interp = ci->device_interp_background != 0;
    if (interp && (!ci->n_gas_next || !ci->kbt_next || !ci->b_field_next || !ci->divv_gyr_next ||
                   !ci->l_turb_mpc_next || !ci->dv_imc_next || !ci->cs_next || !ci->beta_pl_next ||
                   !ci->rad_ic_batch_next))
        return -1;
    if (b1_ctx_for(&w->c, np_, npe_, in->crp_grid, in->cre_grid)) return -1;
    c = w->c.ctx;
    n = np_;
    if (!in->gather_rank_state && !have_in && w->resident_ncell != ncell) {
        ch_set_error("b1: no state: neither inputs, rank gather nor a resident bucket of %d tracers", ncell);
        return -1;
    }
    if (b1_ensure_bucket(w, ncell, n) || ch_ctx_plain_ws(c, ncell, &pw))
        return -1;
    st = ch_ctx_stream(c);
    tl.setup_ms = 1e3 * (ch_now_s() - t_start);
    t_mark = ch_now_s();
    /* backgrounds [CK_NBG][ncell] (current, next) and IC rows [n][ncell] */
    {
        const double *cur[CK_NBG] = {ci->n_gas, ci->kbt, ci->b_field, ci->divv_gyr,
                                     ci->l_turb_mpc, ci->dv_imc, ci->cs, ci->beta_pl};
        const double *nxt[CK_NBG] = {ci->n_gas_next, ci->kbt_next, ci->b_field_next, ci->divv_gyr_next,
                                     ci->l_turb_mpc_next, ci->dv_imc_next, ci->cs_next, ci->beta_pl_next};
        const size_t cap = (size_t)w->cap, rn = (size_t)n * ncell;
        memset(&S, 0, sizeof(S));
        rc = 0;
        for (f = 0; f < CK_NBG; f++) {
            rc |= ck_h2d(w->d_bg + f * cap, cur[f], (size_t)ncell * 8, st);
            S.bgc[f] = w->d_bg + f * cap;
            S.bgn[f] = S.bgc[f];
            if (interp) {
                rc |= ck_h2d(w->d_bg + (CK_NBG + f) * cap, nxt[f], (size_t)ncell * 8, st);
                S.bgn[f] = w->d_bg + (CK_NBG + f) * cap;
            }
        }
        rc |= ck_h2d(w->d_ric, ci->rad_ic_batch, rn * 8, st);
        pw.ric = w->d_ric;
        if (interp) {
            rc |= ck_h2d(w->d_ric + (size_t)n * cap, ci->rad_ic_batch_next, rn * 8, st);
            rc |= ck_lerp_f64(rn, w->d_ric, w->d_ric + (size_t)n * cap, ci->interp_alpha,
                              w->d_ric + 2 * (size_t)n * cap, st);
            pw.ric = w->d_ric + 2 * (size_t)n * cap;
        }
        pw.ric_js = ncell;
        pw.ric_is = 1;
        /* bucket state: inputs, the rank state, or the resident rows */
        if (have_in) {
            const size_t m = (size_t)ncell * n, sc = cap * n;
            CkSlot G;
            memset(&G, 0, sizeof(G));
            rc |= ck_h2d(w->d_stage, in->crp_init, m * 8, st);
            rc |= ck_h2d(w->d_stage + sc, in->cre_init, m * 8, st);
            rc |= ck_h2d(w->d_stage + 2 * sc, in->qpi_batch, m * 8, st);
            rc |= ck_h2d(w->d_stage + 3 * sc, in->qepri_batch, m * 8, st);
            G.crp = w->d_stage;
            G.cre = w->d_stage + sc;
            G.qpi = w->d_stage + 2 * sc;
            G.qepri = w->d_stage + 3 * sc;
            rc |= ck_plain_gather(n, ncell, w->d_ident, &G, pw, st);
        }
        if (use_rank) rc |= ck_h2d(w->d_idx, in->rank_state_indices, (size_t)ncell * sizeof(int), st);
        if (in->gather_rank_state) {
            const size_t rc4 = (size_t)w->rank_cap * n;
            CkSlot G;
            memset(&G, 0, sizeof(G));
            G.qpi = w->d_rank;
            G.qepri = w->d_rank + rc4;
            G.crp = w->d_rank + 2 * rc4;
            G.cre = w->d_rank + 3 * rc4;
            rc |= ck_plain_gather(n, ncell, w->d_idx, &G, pw, st);
        }
        /* per-tracer mass and adiabatic flags (the gather cleared them) */
        if (ci->tracer_mass_msun) rc |= ck_h2d(pw.mass, ci->tracer_mass_msun, (size_t)ncell * 8, st);
        else rc |= ck_fill_f64(pw.mass, 0.0, (size_t)ncell, st);
        if (ci->disable_adiabatic) rc |= ck_h2d(pw.dis, ci->disable_adiabatic, (size_t)ncell, st);
        else rc |= ck_fill_u8(pw.dis, 0, (size_t)ncell, st);
        if (rc || ck_stream_sync(st)) {
            ch_set_error("b1: uploads: %s", ck_last_error());
            return -1;
        }
    }
    w->resident_ncell = ncell;
    tl.h2d_ms = 1e3 * (ch_now_s() - t_mark);
    t_mark = ch_now_s();
    prm = *ch_ctx_prm(c);
    prm.psi = in->psi_value;
    prm.mach_limit = in->mach_limit_value;
    prm.epmax = in->epmax;
    if (ck_plain_lerp(n, ncell, w->d_ident, &S, interp ? ci->interp_alpha : 0.0, pw, st) ||
        ck_plain_coeff(n, ncell, *ch_ctx_dgrid(c), prm, in->dt, prm.t_acc_direct_gyr, 1, pw, ch_ctx_d_err(c), st) ||
        ck_stream_sync(st)) {
        ch_set_error("b1: coefficients: %s", ck_last_error());
        return -1;
    }
    tl.coeff_ms = 1e3 * (ch_now_s() - t_mark);
    t_mark = ch_now_s();
    if (ch_ctx_fqe(c) || ck_plain_secondary(n, ncell, ch_ctx_d_fqe(c), ch_ctx_jp0(c), pw, st) ||
        ck_stream_sync(st)) {
        ch_set_error("b1: secondary source: %s", ck_last_error());
        return -1;
    }
    tl.secondary_ms = 1e3 * (ch_now_s() - t_mark);
    t_mark = ch_now_s();
    if (in->nstep > 0 &&
        ck_plain_thomas(n, ncell, in->nstep, in->dt, in->use_windowed_reacc != 0, in->on_start_step,
                        in->on_end_step, 1, pw, ch_ctx_d_err(c), st)) {
        ch_set_error("b1: solve: %s", ck_last_error());
        return -1;
    }
    if ((rc = ch_ctx_take_err(c)) != 0) return -1;
    tl.solve_ms = 1e3 * (ch_now_s() - t_mark);
    t_mark = ch_now_s();
    if (in->scatter_rank_state) {
        const size_t rc4 = (size_t)w->rank_cap * n;
        CkSlot O;
        memset(&O, 0, sizeof(O));
        O.crp = w->d_rank + 2 * rc4;
        O.cre = w->d_rank + 3 * rc4;
        if (ck_plain_scatter(n, ncell, w->d_idx, &O, pw, st)) return -1;
    }
    if (crp_out) {
        const size_t m = (size_t)ncell * n, sc = (size_t)w->cap * n;
        CkSlot O;
        memset(&O, 0, sizeof(O));
        O.crp = w->d_stage;
        O.cre = w->d_stage + sc;
        if (ck_plain_scatter(n, ncell, w->d_ident, &O, pw, st) || ck_d2h(crp_out, w->d_stage, m * 8, st) ||
            ck_d2h(cre_out, w->d_stage + sc, m * 8, st))
            return -1;
    }
    if (ck_stream_sync(st)) return -1;
    tl.d2h_ms = 1e3 * (ch_now_s() - t_mark);
    tl.total_ms = 1e3 * (ch_now_s() - t_start);
    if (times) *times = tl;
    return 0;
}

int croma_hip_b1_run_tracer_pipeline_group(int np_, int npe_, const CromaB1PipelineInput *inputs, int ninput,
                                           const CromaB1OutputSchedule *schedule, CromaB1PipelineTimes *times)
{
    int i;
    if (times) memset(times, 0, sizeof(*times));
    if (!inputs || ninput < 0) return -1;
    for (i = 0; i < ninput; i++) {
        CromaB1PipelineTimes t;
        const CromaB1PipelineInput *in = &inputs[i];
        if (croma_hip_b1_run_tracer_pipeline(np_, npe_, in, schedule, in->crp_state_out, in->cre_state_out,
                                             in->eps_syn_out, in->eps_gamma_out, in->eps_nu_out, &t))
            return -1;
        if (times) {
            double *a = &times->setup_ms;
            const double *b = &t.setup_ms;
            int k;
            for (k = 0; k < 14; k++) a[k] += b[k];
        }
    }
    return 0;
}

int croma_hip_b1_solve_cc_batch(const CromaB1SolveBatchInput *in, const double *state_init, double *state_out,
                                double *elapsed_ms, int *used_pcr_out)
{
    if (used_pcr_out) *used_pcr_out = 0;
    if (!in || !state_init || !state_out) return -1;
    return croma_hip_solve_cc_batch(in->nsys, in->nrow, in->dt, in->a_batch, in->b_batch, in->c_batch,
                                    in->source_batch, state_init, state_out, elapsed_ms) == 0
               ? 0
               : -1;
}

int croma_hip_b1_tridiag_uses_pcr(int nrow)
{
    (void)nrow;
    return 0;
}

const char *croma_hip_b1_tridiag_solver_name(int nrow)
{
    (void)nrow;
    return "Thomas";
}

/* Emission entry points share one context per grid (they carry no
 * workspace), guarded by its own lock (b1_ctx_for takes g_b1_lock to read
 * the parameters). */
static B1Ctx g_b1_emis;
static pthread_mutex_t g_b1_emis_lock = PTHREAD_MUTEX_INITIALIZER;

int croma_hip_b1_prepare_gamma_emission(int np_, int npe_, const CromaB1HadronicEmissionBatchInput *in,
                                        const void *crp_grid, double *eps_gamma_out, double *elapsed_ms)
{
    int rc;
    if (elapsed_ms) *elapsed_ms = 0.0;
    if (!in || !crp_grid || !eps_gamma_out || in->ncell <= 0 || in->nbins <= 0 || !in->n_gas || !in->crp_batch ||
        !in->f_flat)
        return -1;
    pthread_mutex_lock(&g_b1_emis_lock);
    rc = b1_ctx_for(&g_b1_emis, np_, npe_, crp_grid, crp_grid);
    if (!rc)
        rc = croma_hip_emission_hadronic(g_b1_emis.ctx, in->ncell, in->nbins, in->e_min, in->e_max, in->n_gas,
                                         in->crp_batch, in->f_flat, eps_gamma_out, CROMA_HIP_EMIS_FP64, elapsed_ms);
    pthread_mutex_unlock(&g_b1_emis_lock);
    return rc ? -1 : 0;
}

int croma_hip_b1_prepare_neutrino_emission(int np_, int npe_, const CromaB1HadronicEmissionBatchInput *in,
                                           const void *crp_grid, double *eps_nu_out, double *elapsed_ms)
{
    return croma_hip_b1_prepare_gamma_emission(np_, npe_, in, crp_grid, eps_nu_out, elapsed_ms);
}

int croma_hip_b1_prepare_synch_emission(int np_, int npe_, const CromaB1SynchEmissionBatchInput *in,
                                        const void *cre_grid, double *eps_syn_out, double *elapsed_ms)
{
    int rc;
    if (elapsed_ms) *elapsed_ms = 0.0;
    if (!in || !cre_grid || !eps_syn_out || in->ncell <= 0 || in->nfreq <= 0 || !in->b_dyn || !in->cre_batch)
        return -1;
    if (!in->pitch_kernel_table || in->nlogb <= 1 || !(in->inv_dlogb > 0.0)) {
        ch_set_error("b1: synchrotron needs the pitch-kernel table (the per-cell direct form is not served)");
        return -1;
    }
    pthread_mutex_lock(&g_b1_emis_lock);
    rc = b1_ctx_for(&g_b1_emis, np_, npe_, cre_grid, cre_grid);
    if (!rc)
        rc = croma_hip_emission_synch(g_b1_emis.ctx, in->ncell, in->nfreq, in->nlogb, in->logb_min, in->inv_dlogb,
                                      in->pitch_kernel_table, in->cre_batch, in->b_dyn, eps_syn_out, elapsed_ms);
    pthread_mutex_unlock(&g_b1_emis_lock);
    return rc ? -1 : 0;
}

/* ---- layouts ---------------------------------------------------------- */
typedef struct {
    const char *type, *field;
    size_t off;
} B1Field;

#define B1F(T, RT, F) {RT, #F, offsetof(T, F)}
#define B1A(T, RT, RF, F) {RT, RF, offsetof(T, F)}
#define B1S(T, RT) {RT, "", sizeof(T)}

static const B1Field g_b1_layout[] = {
    B1S(CromaB1OutputSchedule, "FpCudaOutputSchedule"),
    B1F(CromaB1OutputSchedule, "FpCudaOutputSchedule", emit_final_only),
    B1F(CromaB1OutputSchedule, "FpCudaOutputSchedule", emit_all_steps),
    B1F(CromaB1OutputSchedule, "FpCudaOutputSchedule", emit_stride),
    B1F(CromaB1OutputSchedule, "FpCudaOutputSchedule", physical_min),
    B1F(CromaB1OutputSchedule, "FpCudaOutputSchedule", physical_max),
    B1F(CromaB1OutputSchedule, "FpCudaOutputSchedule", selected_steps),
    B1F(CromaB1OutputSchedule, "FpCudaOutputSchedule", nselected_steps),
    B1S(CromaB1PipelineInput, "FpCudaPipelineInput"),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", ncell),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", nstep),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", fp_cadence_steps),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", diff_cadence_steps),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", adv_cadence_steps),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", use_windowed_reacc),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", on_start_step),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", on_end_step),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", nfreq),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", nx_tab),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", dt),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", epmax),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", xmin),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", psi_value),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", mach_limit_value),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", transport_mode),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", crp_grid),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", cre_grid),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", coeff_in),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", synch_in),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", gamma_in),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", neutrino_in),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", qpi_batch),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", qepri_batch),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", beta_p),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", crp_init),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", cre_init),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", crp_state_out),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", cre_state_out),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", eps_syn_out),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", eps_gamma_out),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", eps_nu_out),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", rank_state_indices),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", gather_rank_state),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", scatter_rank_state),
    B1F(CromaB1PipelineInput, "FpCudaPipelineInput", workspace),
    B1S(CromaB1PipelineTimes, "FpCudaPipelineTimes"),
    B1F(CromaB1PipelineTimes, "FpCudaPipelineTimes", setup_ms),
    B1F(CromaB1PipelineTimes, "FpCudaPipelineTimes", total_ms),
    B1F(CromaB1PipelineTimes, "FpCudaPipelineTimes", emitted_synch),
    B1F(CromaB1PipelineTimes, "FpCudaPipelineTimes", last_synch_step),
    B1F(CromaB1PipelineTimes, "FpCudaPipelineTimes", used_pcr_proton),
    B1F(CromaB1PipelineTimes, "FpCudaPipelineTimes", used_pcr_electron),
    B1S(CromaB1CoeffBatchInput, "FpCoeffBatchInput"),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", ncell),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", z),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", dt_gyr),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", epmax),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", n_gas),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", kbt),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", b_field),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", divv_gyr),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", disable_adiabatic),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", l_turb_mpc),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", dv_imc),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", cs),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", beta_pl),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", tracer_mass_msun),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", crp_state_cell_major),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", cre_state_cell_major),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", rad_ic_batch),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", rad_ic_m1),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", rad_ic_p1),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", device_interp_background),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", interp_alpha),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", n_gas_next),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", kbt_next),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", b_field_next),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", divv_gyr_next),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", l_turb_mpc_next),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", dv_imc_next),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", cs_next),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", beta_pl_next),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", rad_ic_batch_next),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", rad_ic_m1_next),
    B1F(CromaB1CoeffBatchInput, "FpCoeffBatchInput", rad_ic_p1_next),
    B1S(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput"),
    B1F(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput", ncell),
    B1F(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput", nlogb),
    B1F(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput", z),
    B1F(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput", inv_dlogb),
    B1F(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput", fx_tab),
    B1F(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput", logy),
    B1F(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput", pitch_kernel_table),
    B1F(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput", b_dyn),
    B1F(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput", cre_batch),
    B1F(CromaB1SynchEmissionBatchInput, "FpSynchEmissionBatchInput", nus),
    B1S(CromaB1HadronicEmissionBatchInput, "FpGammaEmissionBatchInput"),
    B1F(CromaB1HadronicEmissionBatchInput, "FpGammaEmissionBatchInput", ncell),
    B1F(CromaB1HadronicEmissionBatchInput, "FpGammaEmissionBatchInput", nbins),
    B1F(CromaB1HadronicEmissionBatchInput, "FpGammaEmissionBatchInput", z),
    B1A(CromaB1HadronicEmissionBatchInput, "FpGammaEmissionBatchInput", "egamma_min", e_min),
    B1A(CromaB1HadronicEmissionBatchInput, "FpGammaEmissionBatchInput", "egamma_max", e_max),
    B1F(CromaB1HadronicEmissionBatchInput, "FpGammaEmissionBatchInput", n_gas),
    B1F(CromaB1HadronicEmissionBatchInput, "FpGammaEmissionBatchInput", crp_batch),
    B1F(CromaB1HadronicEmissionBatchInput, "FpGammaEmissionBatchInput", beta_p),
    B1A(CromaB1HadronicEmissionBatchInput, "FpGammaEmissionBatchInput", "fga_flat", f_flat),
    B1S(CromaB1HadronicEmissionBatchInput, "FpNeutrinoEmissionBatchInput"),
    B1A(CromaB1HadronicEmissionBatchInput, "FpNeutrinoEmissionBatchInput", "enu_min", e_min),
    B1A(CromaB1HadronicEmissionBatchInput, "FpNeutrinoEmissionBatchInput", "enu_max", e_max),
    B1A(CromaB1HadronicEmissionBatchInput, "FpNeutrinoEmissionBatchInput", "fnu_flat", f_flat),
    B1S(CromaB1SolveBatchInput, "FpChangCooperSolveBatchInput"),
    B1F(CromaB1SolveBatchInput, "FpChangCooperSolveBatchInput", nsys),
    B1F(CromaB1SolveBatchInput, "FpChangCooperSolveBatchInput", nrow),
    B1F(CromaB1SolveBatchInput, "FpChangCooperSolveBatchInput", dt),
    B1F(CromaB1SolveBatchInput, "FpChangCooperSolveBatchInput", a_batch),
    B1F(CromaB1SolveBatchInput, "FpChangCooperSolveBatchInput", source_batch),
};

size_t croma_hip_b1_layout(const char *type_name, const char *field)
{
    size_t i;
    if (!type_name || !field) return (size_t)-1;
    for (i = 0; i < sizeof(g_b1_layout) / sizeof(g_b1_layout[0]); i++)
        if (!strcmp(g_b1_layout[i].type, type_name) && !strcmp(g_b1_layout[i].field, field))
            return g_b1_layout[i].off;
    return (size_t)-1;
}
