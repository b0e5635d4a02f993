/*
 * croma_emission_host.c -- host side of the emission contractions
 * (prepare_{synch,gamma,neutrino}_emission_cuda_batch, fp_cuda_backend.h:97-110,
 * and IC, which the reference runs on the CPU).
 *
 * The kernel matrix is folded on the host (quadrature weights, beta, E_k),
 * laid out bin-major [nbins_pad][K], scaled per bin by a power of two and
 * split into BF16 halves for the MFMA modes; the spectra are contracted on
 * the device (croma_emission.hip).  Host arrays in, host arrays out; the
 * elapsed time reported is the device contraction alone.
 */
#include "croma_host.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint16_t em_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

static float em_bf16f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* kmat given as km(k, j) through a callback-free dense array [nbins][K]. */
static int em_contract(int ntr, int K, int nbins, const double *spectra, const double *kmat_bk,
                       const double *row_scale, double *eps, int mode, double *elapsed_ms)
{
    const int pad = (nbins + 63) / 64 * 64;
    const size_t nk = (size_t)pad * K, ns = (size_t)ntr * K, no = (size_t)ntr * nbins;
    double *kt64 = (double *)calloc(nk, sizeof(double));
    float *kt32 = (float *)calloc(nk, sizeof(float));
    uint16_t *khi = (uint16_t *)calloc(nk, sizeof(uint16_t)), *klo = (uint16_t *)calloc(nk, sizeof(uint16_t));
    int *csh = (int *)calloc((size_t)pad, sizeof(int));
    char *d = 0, *p;
    void *st = 0, *e0 = 0, *e1 = 0;
    int rc = 0, c, j;
    size_t bytes;
    if (!kt64 || !kt32 || !khi || !klo || !csh) {
        rc = CROMA_HIP_ENOMEM;
        goto out;
    }
    for (c = 0; c < nbins; c++) {
        int e = -100000;
        for (j = 0; j < K; j++) {
            const double v = kmat_bk[(size_t)c * K + j];
            kt64[(size_t)c * K + j] = v;
            if (v != 0.0 && isfinite(v)) {
                int ev;
                frexp(v, &ev);
                if (ev > e) e = ev;
            }
        }
        csh[c] = (e == -100000) ? 0 : e;
        for (j = 0; j < K; j++) {
            const float f = (float)ldexp(kt64[(size_t)c * K + j], -csh[c]);
            const uint16_t h = em_bf16(f);
            kt32[(size_t)c * K + j] = f;
            khi[(size_t)c * K + j] = h;
            klo[(size_t)c * K + j] = em_bf16(f - em_bf16f(h));
        }
    }
    bytes = ns * 8 + nk * 8 + nk * 4 + nk * 4 + (size_t)pad * 4 + (size_t)ntr * 8 + no * 8 + 8 * 256;
    d = (char *)ck_malloc(bytes);
    st = ck_stream_create();
    e0 = ck_event_create();
    e1 = ck_event_create();
    if (!d || !st || !e0 || !e1) {
        rc = CROMA_HIP_EDEVICE;
        ch_set_error("emission: device allocation: %s", ck_last_error());
        goto out;
    }
    {
        double *d_s, *d_k64, *d_rs = 0, *d_out;
        float *d_k32;
        uint16_t *d_hi, *d_lo;
        int *d_cs;
        p = d;
#define EM_TAKE(ptr, type, b)                                                                                          \
    do {                                                                                                               \
        ptr = (type *)p;                                                                                               \
        p += ((b) + 255) & ~(size_t)255;                                                                               \
    } while (0)
        EM_TAKE(d_s, double, ns * 8);
        EM_TAKE(d_k64, double, nk * 8);
        EM_TAKE(d_k32, float, nk * 4);
        EM_TAKE(d_hi, uint16_t, nk * 2);
        EM_TAKE(d_lo, uint16_t, nk * 2);
        EM_TAKE(d_cs, int, (size_t)pad * 4);
        if (row_scale) EM_TAKE(d_rs, double, (size_t)ntr * 8);
        EM_TAKE(d_out, double, no * 8);
#undef EM_TAKE
        if (ck_h2d(d_s, spectra, ns * 8, st) || ck_h2d(d_k64, kt64, nk * 8, st) || ck_h2d(d_k32, kt32, nk * 4, st) ||
            ck_h2d(d_hi, khi, nk * 2, st) || ck_h2d(d_lo, klo, nk * 2, st) ||
            ck_h2d(d_cs, csh, (size_t)pad * 4, st) || (row_scale && ck_h2d(d_rs, row_scale, (size_t)ntr * 8, st)) ||
            ck_event_record(e0, st) ||
            ck_emission_contract(ntr, K, nbins, pad, d_s, d_hi, d_lo, d_k32, d_k64, d_cs, d_rs, d_out, mode, st) ||
            ck_event_record(e1, st) || ck_d2h(eps, d_out, no * 8, st) || ck_stream_sync(st)) {
            rc = CROMA_HIP_EDEVICE;
            ch_set_error("emission: %s", ck_last_error());
            goto out;
        }
        if (elapsed_ms) *elapsed_ms = ck_event_elapsed_ms(e0, e1);
    }
out:
    if (st) ck_stream_sync(st);
    ck_free(d);
    ck_event_destroy(e0);
    ck_event_destroy(e1);
    if (st) ck_stream_destroy(st);
    free(kt64);
    free(kt32);
    free(khi);
    free(klo);
    free(csh);
    return rc;
}

int croma_hip_emission_ic(const CromaHipContext *ctx, int ntracer, int nbins, const double *cre,
                          const double *fic_flat, double *eps, int mode, double *elapsed_ms)
{
    const ChGrid *g;
    double *km;
    int k, j, rc, n;
    if (!ctx || ntracer <= 0 || nbins <= 0 || !cre || !fic_flat || !eps || mode < 0 || mode > 3)
        return CROMA_HIP_EINVAL;
    g = ch_ctx_grid(ctx);
    n = g->n;
    km = (double *)malloc((size_t)nbins * n * sizeof(double));
    if (!km) return CROMA_HIP_ENOMEM;
    /* eps[i][k] = sum_j fic[k][j] cre[i][j] dp[j] (fp_shared_core.c:2131-2157) */
    for (k = 0; k < nbins; k++)
        for (j = 0; j < n; j++) km[(size_t)k * n + j] = fic_flat[(size_t)k * n + j] * g->edp[j];
    rc = em_contract(ntracer, n, nbins, cre, km, 0, eps, mode, elapsed_ms);
    free(km);
    return rc;
}

int croma_hip_emission_hadronic(const CromaHipContext *ctx, int ntracer, int nbins, double emin, double emax,
                                const double *n_gas, const double *crp, const double *f_flat, double *eps, int mode,
                                double *elapsed_ms)
{
    const ChGrid *g;
    double *km, *tw, d, a;
    int k, j, rc, n;
    if (!ctx || ntracer <= 0 || nbins <= 0 || !n_gas || !crp || !f_flat || !eps || mode < 0 || mode > 3)
        return CROMA_HIP_EINVAL;
    g = ch_ctx_grid(ctx);
    n = g->n;
    km = (double *)malloc((size_t)nbins * n * sizeof(double));
    tw = (double *)malloc((size_t)n * sizeof(double));
    if (!km || !tw) {
        free(km);
        free(tw);
        return CROMA_HIP_ENOMEM;
    }
    /* trapezoid sum_{i>=1} (f_{i-1} + f_i) dp_i = sum_j f_j tw_j,
     * tw_j = dp_j [j >= 1] + dp_{j+1} [j + 1 <= n - 1] (EMISSION.c:297-308) */
    for (j = 0; j < n; j++) tw[j] = (j >= 1 ? g->pdp[j] : 0.0) + (j + 1 <= n - 1 ? g->pdp[j + 1] : 0.0);
    d = (emax - emin) / (double)nbins;
    a = emin - d;
    for (k = 0; k < nbins; k++) {
        double e;
        a += d;
        e = pow(10.0, a);
        for (j = 0; j < n; j++) km[(size_t)k * n + j] = e * 0.5 * g->beta_p[j] * tw[j] * f_flat[(size_t)k * n + j];
    }
    rc = em_contract(ntracer, n, nbins, crp, km, n_gas, eps, mode, elapsed_ms);
    free(km);
    free(tw);
    return rc;
}

int croma_hip_synch_kernel_table(const CromaHipContext *ctx, int nfreq, int ntheta, int nlogb, double logb_min,
                                 double inv_dlogb, int nx_tab, double xmin, const double *logx_tab,
                                 const double *logfx_tab, const double *fx_tab, const double *logy,
                                 const double *pitch_weight, double *table, double *elapsed_ms)
{
    const ChGrid *g;
    size_t ny, nt, bytes;
    double *d = 0;
    void *st = 0, *e0 = 0, *e1 = 0;
    int n, rc = 0;
    if (!ctx || nfreq <= 0 || ntheta <= 0 || nlogb <= 0 || nx_tab < 1 || !logx_tab || !logfx_tab || !fx_tab ||
        !logy || !pitch_weight || !table || !(inv_dlogb > 0.0))
        return CROMA_HIP_EINVAL;
    g = ch_ctx_grid(ctx);
    n = g->n;
    ny = (size_t)nfreq * n * ntheta;
    nt = (size_t)nlogb * nfreq * n;
    bytes = (3 * (size_t)nx_tab + ny + (size_t)ntheta + nt) * sizeof(double);
    d = (double *)ck_malloc(bytes);
    st = ck_stream_create();
    e0 = ck_event_create();
    e1 = ck_event_create();
    if (!d || !st || !e0 || !e1) {
        rc = CROMA_HIP_EDEVICE;
        ch_set_error("synch table: device allocation: %s", ck_last_error());
        goto out;
    }
    {
        double *d_lx = d, *d_lfx = d + nx_tab, *d_fx = d + 2 * (size_t)nx_tab, *d_ly = d + 3 * (size_t)nx_tab;
        double *d_pw = d_ly + ny, *d_tab = d_pw + ntheta;
        if (ck_h2d(d_lx, logx_tab, (size_t)nx_tab * 8, st) || ck_h2d(d_lfx, logfx_tab, (size_t)nx_tab * 8, st) ||
            ck_h2d(d_fx, fx_tab, (size_t)nx_tab * 8, st) || ck_h2d(d_ly, logy, ny * 8, st) ||
            ck_h2d(d_pw, pitch_weight, (size_t)ntheta * 8, st) || ck_event_record(e0, st) ||
            ck_synch_table(n, nfreq, ntheta, nlogb, logb_min, inv_dlogb, nx_tab, xmin, d_lx, d_lfx, d_fx, d_ly,
                           d_pw, d_tab, st) ||
            ck_event_record(e1, st) || ck_d2h(table, d_tab, nt * 8, st) || ck_stream_sync(st)) {
            rc = CROMA_HIP_EDEVICE;
            ch_set_error("synch table: %s", ck_last_error());
            goto out;
        }
        if (elapsed_ms) *elapsed_ms = ck_event_elapsed_ms(e0, e1);
    }
out:
    if (st) ck_stream_sync(st);
    ck_free(d);
    ck_event_destroy(e0);
    ck_event_destroy(e1);
    if (st) ck_stream_destroy(st);
    return rc;
}

int croma_hip_emission_synch(const CromaHipContext *ctx, int ntracer, int nfreq, int nlogb, double logb_min,
                             double inv_dlogb, const double *table, const double *cre, const double *b_field,
                             double *eps, double *elapsed_ms)
{
    return croma_hip_emission_synch_mode(ctx, ntracer, nfreq, nlogb, logb_min, inv_dlogb, table, cre, b_field, eps,
                                         CROMA_HIP_EMIS_FP64, elapsed_ms);
}

/* Synchrotron on MFMA (ck_synch_mfma_prepare / ck_synch_mfma_run): the
 * Chebyshev expansion of the log-B interpolation needs table planes no
 * coarser than 0.02 in log10 B and n <= 256 (the kernel's 64 columns of K = n
 * bins in LDS).  A plan holds the table's device form; its per-call buffers
 * grow to the largest call. */
struct CromaHipSynchPlan {
    const CromaHipContext *ctx;
    int n, nfreq, nlogb;
    double logb_min, inv_dlogb;
    char *d_const;
    double *d_tab, *d_dp;
    uint16_t *d_hi, *d_lo;
    int *d_cs;
    float *d_lt;
    char *d_run;
    long cap;
    double *d_cre, *d_b, *d_eps;
    int *d_work, *d_fix;
    void *stream, *e0, *e1;
};

#define SY_R(b) ((((size_t)(b)) + 255) & ~(size_t)255)

void croma_hip_synch_plan_destroy(CromaHipSynchPlan *p)
{
    if (!p) return;
    if (p->stream) ck_stream_sync(p->stream);
    ck_free(p->d_const);
    ck_free(p->d_run);
    if (p->e0) ck_event_destroy(p->e0);
    if (p->e1) ck_event_destroy(p->e1);
    if (p->stream) ck_stream_destroy(p->stream);
    free(p);
}

CromaHipSynchPlan *croma_hip_synch_plan_create(const CromaHipContext *ctx, int nfreq, int nlogb, double logb_min,
                                               double inv_dlogb, const double *table, double *elapsed_ms)
{
    CromaHipSynchPlan *p;
    const ChGrid *g;
    size_t ntab, nkt, ncs;
    char *q;
    if (!ctx || nfreq <= 0 || nlogb < 2 || !table || !isfinite(logb_min) || !isfinite(inv_dlogb)) {
        ch_set_error("synch plan: invalid arguments");
        return 0;
    }
    g = ch_ctx_grid(ctx);
    if (g->n > 256 || !(inv_dlogb >= 50.0)) {
        ch_set_error("synch MFMA: needs n <= 256 and planes no coarser than 0.02 in log10 B (n %d, 1/dlogb %g)", g->n,
                     inv_dlogb);
        return 0;
    }
    p = (CromaHipSynchPlan *)calloc(1, sizeof(*p));
    if (!p) return 0;
    p->ctx = ctx;
    p->n = g->n;
    p->nfreq = nfreq;
    p->nlogb = nlogb;
    p->logb_min = logb_min;
    p->inv_dlogb = inv_dlogb;
    ntab = (size_t)nlogb * nfreq * p->n;
    nkt = (size_t)(nlogb - 1) * 2 * ((nfreq + 15) / 16) * 64 * p->n;
    ncs = (size_t)(nlogb - 1) * nfreq;
    p->d_const = (char *)ck_malloc(SY_R(ntab * 8) + SY_R((size_t)p->n * 8) + 2 * SY_R(nkt * 2) + SY_R(ncs * 4) +
                                   SY_R(ntab * 4));
    p->stream = ck_stream_create();
    p->e0 = ck_event_create();
    p->e1 = ck_event_create();
    if (!p->d_const || !p->stream || !p->e0 || !p->e1) {
        ch_set_error("synch plan: device allocation: %s", ck_last_error());
        croma_hip_synch_plan_destroy(p);
        return 0;
    }
    q = p->d_const;
    p->d_tab = (double *)q;
    q += SY_R(ntab * 8);
    p->d_dp = (double *)q;
    q += SY_R((size_t)p->n * 8);
    p->d_hi = (uint16_t *)q;
    q += SY_R(nkt * 2);
    p->d_lo = (uint16_t *)q;
    q += SY_R(nkt * 2);
    p->d_cs = (int *)q;
    q += SY_R(ncs * 4);
    p->d_lt = (float *)q;
    if (ck_h2d(p->d_tab, table, ntab * 8, p->stream) || ck_h2d(p->d_dp, g->edp, (size_t)p->n * 8, p->stream) ||
        ck_event_record(p->e0, p->stream) ||
        ck_synch_mfma_prepare(p->n, nfreq, nlogb, p->d_tab, p->d_dp, p->d_hi, p->d_lo, p->d_cs, p->d_lt, p->stream) ||
        ck_event_record(p->e1, p->stream) || ck_stream_sync(p->stream)) {
        ch_set_error("synch plan: %s", ck_last_error());
        croma_hip_synch_plan_destroy(p);
        return 0;
    }
    if (elapsed_ms) *elapsed_ms = ck_event_elapsed_ms(p->e0, p->e1);
    return p;
}

int croma_hip_synch_plan_run(CromaHipSynchPlan *p, int ntr, const double *cre, const double *b_field, double *eps,
                             double *elapsed_ms)
{
    const int n = p ? p->n : 0;
    if (!p || ntr <= 0 || !cre || !b_field || !eps) return CROMA_HIP_EINVAL;
    if (ntr > p->cap) {
        const size_t nt = (size_t)ntr;
        char *q;
        ck_stream_sync(p->stream);
        ck_free(p->d_run);
        p->cap = 0;
        p->d_run = (char *)ck_malloc(SY_R(nt * n * 8) + SY_R(nt * 8) + SY_R(nt * p->nfreq * 8) +
                                     SY_R((2 * nt + 2 * (size_t)p->nlogb + 8) * 4) + SY_R((nt + 1) * 4));
        if (!p->d_run) {
            ch_set_error("synch plan: device allocation: %s", ck_last_error());
            return CROMA_HIP_EDEVICE;
        }
        q = p->d_run;
        p->d_cre = (double *)q;
        q += SY_R(nt * n * 8);
        p->d_b = (double *)q;
        q += SY_R(nt * 8);
        p->d_eps = (double *)q;
        q += SY_R(nt * p->nfreq * 8);
        p->d_work = (int *)q;
        q += SY_R((2 * nt + 2 * (size_t)p->nlogb + 8) * 4);
        p->d_fix = (int *)q;
        p->cap = ntr;
    }
    if (ck_h2d(p->d_cre, cre, (size_t)ntr * n * 8, p->stream) ||
        ck_h2d(p->d_b, b_field, (size_t)ntr * 8, p->stream) || ck_event_record(p->e0, p->stream) ||
        ck_synch_mfma_run(ntr, n, p->nfreq, p->nlogb, p->logb_min, p->inv_dlogb, p->d_tab, p->d_hi, p->d_lo, p->d_cs,
                          p->d_lt, p->d_work, p->d_fix, p->d_cre, p->d_dp, p->d_b, p->d_eps, p->stream) ||
        ck_event_record(p->e1, p->stream) || ck_d2h(eps, p->d_eps, (size_t)ntr * p->nfreq * 8, p->stream) ||
        ck_stream_sync(p->stream)) {
        ch_set_error("synch MFMA: %s", ck_last_error());
        return CROMA_HIP_EDEVICE;
    }
    if (elapsed_ms) *elapsed_ms = ck_event_elapsed_ms(p->e0, p->e1);
    return 0;
}
#undef SY_R

static int synch_mfma(const CromaHipContext *ctx, int ntr, int nfreq, int nlogb, double logb_min, double inv_dlogb,
                      const double *table, const double *cre, const double *b_field, double *eps, double *elapsed_ms)
{
    double t0 = 0.0, t1 = 0.0;
    int rc;
    CromaHipSynchPlan *p = croma_hip_synch_plan_create(ctx, nfreq, nlogb, logb_min, inv_dlogb, table, &t0);
    if (!p) return (ch_ctx_grid(ctx)->n > 256 || !(inv_dlogb >= 50.0)) ? CROMA_HIP_EUNSUPPORTED : CROMA_HIP_EDEVICE;
    rc = croma_hip_synch_plan_run(p, ntr, cre, b_field, eps, &t1);
    croma_hip_synch_plan_destroy(p);
    if (!rc && elapsed_ms) *elapsed_ms = t0 + t1;
    return rc;
}

int croma_hip_emission_synch_mode(const CromaHipContext *ctx, int ntracer, int nfreq, int nlogb, double logb_min,
                                  double inv_dlogb, const double *table, const double *cre, const double *b_field,
                                  double *eps, int mode, double *elapsed_ms)
{
    const ChGrid *g;
    const size_t nt_tab = (size_t)nlogb * nfreq;
    char *d = 0, *p;
    void *st = 0, *e0 = 0, *e1 = 0;
    int n, rc = 0;
    if (!ctx || ntracer <= 0 || nfreq <= 0 || nlogb < 2 || !table || !cre || !b_field || !eps ||
        (mode != CROMA_HIP_EMIS_FP64 && mode != CROMA_HIP_EMIS_FP32 && mode != CROMA_HIP_EMIS_MFMA_XF32))
        return CROMA_HIP_EINVAL;
    g = ch_ctx_grid(ctx);
    n = g->n;
    if (mode == CROMA_HIP_EMIS_MFMA_XF32) return synch_mfma(ctx, ntracer, nfreq, nlogb, logb_min, inv_dlogb, table,
                                                            cre, b_field, eps, elapsed_ms);
    d = (char *)ck_malloc(2 * nt_tab * n * 8 + (size_t)ntracer * n * 8 + (size_t)n * 8 + (size_t)ntracer * 8 +
                          (size_t)ntracer * nfreq * 8 + (3 * (size_t)ntracer + (size_t)nlogb) * 4 + 7 * 256);
    st = ck_stream_create();
    e0 = ck_event_create();
    e1 = ck_event_create();
    if (!d || !st || !e0 || !e1) {
        rc = CROMA_HIP_EDEVICE;
        ch_set_error("synch emission: device allocation: %s", ck_last_error());
        goto out;
    }
    {
        double *d_tab, *d_logt, *d_cre, *d_dp, *d_b, *d_eps;
        int *d_work;
        p = d;
        d_tab = (double *)p;
        p += (nt_tab * n * 8 + 255) & ~(size_t)255;
        d_logt = (double *)p;
        p += (nt_tab * n * 8 + 255) & ~(size_t)255;
        d_cre = (double *)p;
        p += ((size_t)ntracer * n * 8 + 255) & ~(size_t)255;
        d_dp = (double *)p;
        p += ((size_t)n * 8 + 255) & ~(size_t)255;
        d_b = (double *)p;
        p += ((size_t)ntracer * 8 + 255) & ~(size_t)255;
        d_eps = (double *)p;
        p += ((size_t)ntracer * nfreq * 8 + 255) & ~(size_t)255;
        d_work = (int *)p;
        if (ck_h2d(d_tab, table, nt_tab * n * 8, st) || ck_h2d(d_cre, cre, (size_t)ntracer * n * 8, st) ||
            ck_h2d(d_dp, g->edp, (size_t)n * 8, st) || ck_h2d(d_b, b_field, (size_t)ntracer * 8, st) ||
            ck_event_record(e0, st) ||
            ck_emission_synch(ntracer, n, nfreq, nlogb, logb_min, inv_dlogb, d_tab, d_logt, d_work, d_cre, d_dp, d_b,
                              d_eps, mode == CROMA_HIP_EMIS_FP64, st) ||
            ck_event_record(e1, st) || ck_d2h(eps, d_eps, (size_t)ntracer * nfreq * 8, st) || ck_stream_sync(st)) {
            rc = CROMA_HIP_EDEVICE;
            ch_set_error("synch emission: %s", ck_last_error());
            goto out;
        }
        if (elapsed_ms) *elapsed_ms = ck_event_elapsed_ms(e0, e1);
    }
out:
    if (st) ck_stream_sync(st);
    ck_free(d);
    ck_event_destroy(e0);
    ck_event_destroy(e1);
    if (st) ck_stream_destroy(st);
    return rc;
}
