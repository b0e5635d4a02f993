/*
 * croma_host.c -- C host runtime of libcroma_hip.so: device binding,
 * contexts (grid tables on the device), device-resident tracer batches, the
 * per-snapshot scheduler (nsub -> cost-ordered task list -> persistent fused
 * kernel, or the plain per-bucket path), and the snapshot-level drop-in for
 * tracer_fp_batch_execute_parallel.
 *
 * Per snapshot the host work (IC cooling rows, nsub, task ordering) of
 * snapshot s+1 is prepared while the GPU runs snapshot s.
 */
#include "croma_host.h"

#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static __thread char g_err[512];

void ch_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *croma_hip_last_error(void) { return g_err; }

double ch_now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1.0e-9 * (double)ts.tv_nsec;
}

static int dev_fail(const char *what)
{
    ch_set_error("%s: %s", what, ck_last_error());
    return CROMA_HIP_EDEVICE;
}

int croma_hip_backend_parse(const char *spec)
{
    if (!spec) return 0;
    return strcmp(spec, "hip") == 0 || strcmp(spec, "HIP") == 0 || strcmp(spec, "rocm") == 0;
}

/* ------------------------------------------------------------------ */
int croma_hip_is_available(void) { return ck_device_count() > 0 ? 1 : 0; }

int croma_hip_bind_local_rank(int local_rank, int *device_out)
{
    const int n = ck_device_count();
    int dev;
    if (n <= 0) {
        ch_set_error("no HIP device visible");
        return CROMA_HIP_EDEVICE;
    }
    if (local_rank < 0) local_rank = 0;
    dev = local_rank % n;
    if (ck_set_device(dev) != 0) return dev_fail("bind");
    if (device_out) *device_out = dev;
    return 0;
}

int croma_hip_get_binding_state(int *count, int *current)
{
    if (count) *count = ck_device_count();
    if (current) *current = ck_get_device();
    return 0;
}

int croma_hip_get_current_device_id(char *buf, int len)
{
    const int d = ck_get_device();
    if (!buf || len <= 0) return CROMA_HIP_EINVAL;
    if (d < 0) return dev_fail("device id");
    return ck_device_name(d, buf, len) == 0 ? 0 : CROMA_HIP_EDEVICE;
}

void croma_hip_default_params(CromaHipParams *p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->dpp_model = 3;
    p->ttd_tacc_model = 1;
    p->reacc_window_mode = 0;
    p->tracer_nsub_max = 0;
    p->coeff_interp_min_steps = 4;
    p->coeff_interp_max_segments = 16;
    p->psi = 0.5;
    p->mach_limit = 0.5;
    p->eta_dpp_cap = 0.0;
    p->t_acc_direct_gyr = 0.0;
    p->t_off_gyr = 0.0;
    p->t_acc_off_gyr = 3.0;
    p->f_eddy = 2.0;
    p->tracer_nsub_safety = 0.25;
    p->epmax = 1.0e7;
}

void ch_params_to_ck(const CromaHipParams *in, CkParams *o)
{
    o->dpp_model = in->dpp_model;
    o->ttd_tacc_model = in->ttd_tacc_model;
    o->reacc_window_mode = in->reacc_window_mode;
    o->tracer_nsub_max = in->tracer_nsub_max;
    o->coeff_interp_min_steps = in->coeff_interp_min_steps;
    o->coeff_interp_max_segments = in->coeff_interp_max_segments;
    o->psi = in->psi;
    o->mach_limit = in->mach_limit;
    o->eta_dpp_cap = in->eta_dpp_cap;
    o->t_acc_direct_gyr = in->t_acc_direct_gyr;
    o->t_off_gyr = in->t_off_gyr;
    o->t_acc_off_gyr = in->t_acc_off_gyr;
    o->f_eddy = in->f_eddy;
    o->tracer_nsub_safety = in->tracer_nsub_safety;
    o->epmax = in->epmax;
}

static int valid_params(const CromaHipParams *p)
{
    if (!p) return 0;
    if (p->dpp_model < 1 || p->dpp_model > 4) return 0;
    if (!(p->tracer_nsub_safety > 0.0)) return 0;
    if (!(p->epmax > 0.0) || !isfinite(p->epmax)) return 0;
    return 1;
}

/* ------------------------------------------------------------------ */
#define CH_IC_CACHE 8
#define CH_IC_ENDS 256

/* A launch graph slot: the executable of the last launch key seen. */
#define CH_GKEY 8

typedef struct {
    void *exec;
    long key[CH_GKEY];
    int state; /* 0 empty, 1 key seen once, 2 captured, -1 capture refused */
} ChGraph;

struct CromaHipContext {
    int device;
    int cu_count;
    ChGrid g;
    CromaHipParams uprm;
    CkParams prm;
    CkGrid dgrid;
    double *d_grid_mem;
    void *stream;
    void *stream_e; /* split windows: the electron pass, beside the next window's proton pass */
    int *d_counter;
    int *d_err;
    CromaHipTile tile[2]; /* [protons] */
    pthread_mutex_t ic_lock;
    struct {
        int valid;
        double z, m1, p1;
        double row[CH_NMAX];
    } ic[CH_IC_CACHE];
    int ic_next;
    /* end bins (0, n-1) of IC rows, for the nsub estimate */
    struct {
        int valid;
        double z, v[2];
    } ic_end[CH_IC_ENDS];
    int ic_end_next;
    /* secondary weights, built on first use */
    int fqe_ready;
    int jp0;
    int *h_npmin;
    double *d_fqe;
    /* the last IC row a window built (device, n doubles) and its (z, lo, hi):
     * the next window's first snapshot is usually that z */
    double *d_ic_last;
    double ic_last_zb[3];
    int ic_last_valid;
    /* plain-path workspace */
    CkPlainWs pws;
    void *pws_mem;
    int pws_cap;
    /* window launches replayed as hipGraphs */
    int use_graphs;
    int pipeline; /* split windows: electron pass beside the next proton pass */
    /* batch runs estimate nsub and order the queue on the device */
    int nsub_device;
    /* coefficient evaluation in FP32 (losses, D_pp, B, C, face inputs) feeding
     * the FP64 weights, factorisation and substeps */
    int coeff_fp32;
    void **retired; /* graph executables replaced while their last launch may still run */
    int nretired, retired_cap;
    /* croma_hip_execute_snapshot state kept between calls */
    CromaHipBatch *xb;
    void *x_ws;
    size_t x_ws_bytes;
    ChGraph x_graph;
};

/* Derived tables of the Chang-Cooper build for one species: face width
 * h_j = x_{j+1} - x_j (0 at j = n-1), 1/h_j, 1/dx_j, 2/x_j. */
static void derived_tables(int n, const double *x, const double *dx, double *h, double *inv_h, double *inv_dx,
                           double *two_op)
{
    int j;
    for (j = 0; j < n; j++) {
        h[j] = (j + 1 < n) ? x[j + 1] - x[j] : 0.0;
        inv_h[j] = (j + 1 < n) ? 1.0 / h[j] : 0.0;
        inv_dx[j] = 1.0 / dx[j];
        two_op[j] = 2.0 / x[j];
    }
}

#define CH_NTAB 28 /* 21 uploaded, 6 segment-build factor tables and a zero row built on the device */
#define CH_NPOW10 81 /* 10^k, k = -40..40, for the nsub ladder on the device */

static int ctx_upload_grid(CromaHipContext *c)
{
    const int n = c->g.n;
    const size_t nb = (size_t)n * sizeof(double);
    double *tab = (double *)malloc(CH_NTAB * nb + CH_NPOW10 * sizeof(double));
    const double *src[13] = {c->g.pp, c->g.pdp, c->g.pE, c->g.p_p2, c->g.p_expcut, c->g.p_sig, c->g.ep,
                             c->g.edp, c->g.eE, c->g.e_p2, c->g.e_expcut, c->g.p_ecr_w, c->g.e_ecr_w};
    const double *d;
    int k;
    if (!tab) return CROMA_HIP_ENOMEM;
    for (k = 0; k < 13; k++) memcpy(tab + (size_t)k * n, src[k], nb);
    derived_tables(n, c->g.pp, c->g.pdp, tab + 13 * (size_t)n, tab + 14 * (size_t)n, tab + 15 * (size_t)n,
                   tab + 16 * (size_t)n);
    derived_tables(n, c->g.ep, c->g.edp, tab + 17 * (size_t)n, tab + 18 * (size_t)n, tab + 19 * (size_t)n,
                   tab + 20 * (size_t)n);
    for (k = 0; k < CH_NPOW10; k++) tab[21 * (size_t)n + k] = pow(10.0, (double)(k - 40));
    c->d_grid_mem = (double *)ck_malloc(CH_NTAB * nb + CH_NPOW10 * sizeof(double));
    if (!c->d_grid_mem) {
        free(tab);
        return dev_fail("grid alloc");
    }
    if (ck_h2d(c->d_grid_mem, tab, 21 * nb, c->stream) ||
        ck_h2d(c->d_grid_mem + CH_NTAB * (size_t)n, tab + 21 * (size_t)n, CH_NPOW10 * sizeof(double), c->stream) ||
        ck_stream_sync(c->stream)) {
        free(tab);
        return dev_fail("grid upload");
    }
    free(tab);
    d = c->d_grid_mem;
    c->dgrid.n = n;
    c->dgrid.j_pp = c->g.j_pp;
    c->dgrid.pp = d + 0 * (size_t)n;
    c->dgrid.pdp = d + 1 * (size_t)n;
    c->dgrid.pE = d + 2 * (size_t)n;
    c->dgrid.p_p2 = d + 3 * (size_t)n;
    c->dgrid.p_expcut = d + 4 * (size_t)n;
    c->dgrid.p_sig = d + 5 * (size_t)n;
    c->dgrid.ep = d + 6 * (size_t)n;
    c->dgrid.edp = d + 7 * (size_t)n;
    c->dgrid.eE = d + 8 * (size_t)n;
    c->dgrid.e_p2 = d + 9 * (size_t)n;
    c->dgrid.e_expcut = d + 10 * (size_t)n;
    c->dgrid.p_ecr_w = d + 11 * (size_t)n;
    c->dgrid.e_ecr_w = d + 12 * (size_t)n;
    c->dgrid.p_h = d + 13 * (size_t)n;
    c->dgrid.p_inv_h = d + 14 * (size_t)n;
    c->dgrid.p_inv_dx = d + 15 * (size_t)n;
    c->dgrid.p_2op = d + 16 * (size_t)n;
    c->dgrid.e_h = d + 17 * (size_t)n;
    c->dgrid.e_inv_h = d + 18 * (size_t)n;
    c->dgrid.e_inv_dx = d + 19 * (size_t)n;
    c->dgrid.e_2op = d + 20 * (size_t)n;
    c->dgrid.p_Ep = d + 21 * (size_t)n;
    c->dgrid.p_i2mk = d + 22 * (size_t)n;
    c->dgrid.p_ib2 = d + 23 * (size_t)n;
    c->dgrid.p_expE = d + 24 * (size_t)n;
    c->dgrid.e_ib2 = d + 25 * (size_t)n;
    c->dgrid.e_lnp = d + 26 * (size_t)n;
    c->dgrid.zeros = d + 27 * (size_t)n;
    c->dgrid.pow10 = d + CH_NTAB * (size_t)n;
    c->dgrid.ppm1 = c->g.ppm1;
    c->dgrid.ppp1 = c->g.ppp1;
    c->dgrid.epm1 = c->g.epm1;
    c->dgrid.epp1 = c->g.epp1;
    if (ck_grid_loss_tables(c->dgrid, c->prm.epmax, c->stream) || ck_stream_sync(c->stream))
        return dev_fail("grid loss tables");
    return 0;
}

static void default_tiles(CromaHipContext *c)
{
    const int n = c->g.n;
    /* rows per lane: the spectrum and one factor set stay in registers;
     * two waves per SIMD hide the cross-lane latency of the substep scans */
    /* measured defaults (the autotuner refines them per workload) */
    c->tile[0].rows_per_lane = (n == 64) ? 4 : (n == 128) ? 8 : (n == 256) ? 8 : 16;
    c->tile[0].waves_per_simd = (n <= 128) ? 1 : 2;
    c->tile[0].chain = 1;
    c->tile[0].secondary_gemm = 0;
    c->tile[1].rows_per_lane = (n <= 256) ? 4 : 8;
    c->tile[1].waves_per_simd = (n == 64) ? 1 : 2;
    c->tile[1].chain = 1;
    c->tile[1].secondary_gemm = 1;
}

static int tile_ok(int n, const CromaHipTile *t)
{
    return ck_fused_supported(n, t->rows_per_lane) && t->waves_per_simd >= 1 && t->waves_per_simd <= 2 &&
           t->chain >= 1 && t->chain <= 4096 && (t->secondary_gemm == 0 || t->secondary_gemm == 1);
}

CromaHipContext *croma_hip_context_create(const CromaHipGridSpec *gs, const CromaHipParams *prm)
{
    CromaHipContext *c;
    CromaHipParams def;
    if (!gs || !(gs->n == 64 || gs->n == 128 || gs->n == 256 || gs->n == 512)) {
        ch_set_error("grid size must be 64, 128, 256 or 512");
        return 0;
    }
    if (!prm) {
        croma_hip_default_params(&def);
        prm = &def;
    }
    if (!valid_params(prm)) {
        ch_set_error("invalid parameters");
        return 0;
    }
    if (ck_device_count() <= 0) {
        ch_set_error("no HIP device visible");
        return 0;
    }
    c = (CromaHipContext *)calloc(1, sizeof(*c));
    if (!c) return 0;
    pthread_mutex_init(&c->ic_lock, 0);
    c->device = ck_get_device();
    if (c->device < 0) c->device = 0;
    c->cu_count = ck_device_cu_count(c->device);
    if (c->cu_count <= 0) c->cu_count = 256;
    if (ch_grid_build(&c->g, gs->n, gs->pmin, gs->pmax, gs->pemin, gs->pemax) != 0) {
        ch_set_error("grid build failed");
        free(c);
        return 0;
    }
    c->uprm = *prm;
    ch_params_to_ck(prm, &c->prm);
    c->stream = ck_stream_create();
    c->d_counter = (int *)ck_malloc(sizeof(int) * 64);
    c->d_err = (int *)ck_malloc(sizeof(int) * 64);
    if (!c->stream || !c->d_counter || !c->d_err || ctx_upload_grid(c) != 0) {
        croma_hip_context_destroy(c);
        return 0;
    }
    if (ck_fill_i32(c->d_err, 0, 64, c->stream) || ck_stream_sync(c->stream)) {
        dev_fail("context init");
        croma_hip_context_destroy(c);
        return 0;
    }
    default_tiles(c);
    c->use_graphs = 1;
    c->pipeline = 0;
    c->nsub_device = 1;
    return c;
}

/* Graph executables are destroyed only after the stream has drained. */
static void retire_graph(CromaHipContext *c, void *ex)
{
    if (!ex) return;
    if (c->nretired == c->retired_cap) {
        const int cap = c->retired_cap ? 2 * c->retired_cap : 8;
        void **nr = (void **)realloc(c->retired, (size_t)cap * sizeof(void *));
        if (!nr) {
            ck_stream_sync(c->stream);
            ck_graph_destroy(ex);
            return;
        }
        c->retired = nr;
        c->retired_cap = cap;
    }
    c->retired[c->nretired++] = ex;
}

/* Call with the stream drained. */
static void flush_retired(CromaHipContext *c)
{
    int i;
    for (i = 0; i < c->nretired; i++) ck_graph_destroy(c->retired[i]);
    c->nretired = 0;
}

void croma_hip_context_destroy(CromaHipContext *c)
{
    if (!c) return;
    if (c->stream) ck_stream_sync(c->stream);
    croma_hip_batch_destroy(c->xb);
    ck_free(c->x_ws);
    ck_graph_destroy(c->x_graph.exec);
    flush_retired(c);
    free(c->retired);
    ck_free(c->d_grid_mem);
    ck_free(c->d_counter);
    ck_free(c->d_err);
    ck_free(c->d_fqe);
    ck_free(c->d_ic_last);
    ck_free(c->pws_mem);
    free(c->h_npmin);
    if (c->stream_e) ck_stream_destroy(c->stream_e);
    if (c->stream) ck_stream_destroy(c->stream);
    pthread_mutex_destroy(&c->ic_lock);
    free(c);
}

int croma_hip_context_set_params(CromaHipContext *c, const CromaHipParams *prm)
{
    const double old_epmax = c ? c->prm.epmax : 0.0;
    if (!c || !valid_params(prm)) return CROMA_HIP_EINVAL;
    c->uprm = *prm;
    ch_params_to_ck(prm, &c->prm);
    c->ic_last_valid = 0;
    /* the D_pp cutoff table depends on E_pmax */
    if (c->prm.epmax != old_epmax &&
        (ck_stream_sync(c->stream) || ck_grid_loss_tables(c->dgrid, c->prm.epmax, c->stream) ||
         ck_stream_sync(c->stream)))
        return dev_fail("grid loss tables");
    return 0;
}

int croma_hip_context_set_tile(CromaHipContext *c, int protons, const CromaHipTile *t)
{
    if (!c || !t || protons < -1 || protons > 1) return CROMA_HIP_EINVAL;
    if (!tile_ok(c->g.n, t)) return CROMA_HIP_EUNSUPPORTED;
    if (protons != 1) c->tile[0] = *t;
    if (protons != 0) c->tile[1] = *t;
    return 0;
}

int croma_hip_context_get_tile(const CromaHipContext *c, int protons, CromaHipTile *t)
{
    if (!c || !t) return CROMA_HIP_EINVAL;
    *t = c->tile[protons ? 1 : 0];
    return 0;
}

int croma_hip_context_set_coeff_precision(CromaHipContext *c, int precision)
{
    if (!c || (precision != CROMA_HIP_COEFF_FP64 && precision != CROMA_HIP_COEFF_FP32)) return CROMA_HIP_EINVAL;
    ck_stream_sync(c->stream);
    if (c->stream_e) ck_stream_sync(c->stream_e);
    c->coeff_fp32 = precision == CROMA_HIP_COEFF_FP32;
    return 0;
}

int ch_ctx_coeff_fp32(const CromaHipContext *c) { return c->coeff_fp32; }

int croma_hip_context_set_nsub_device(CromaHipContext *c, int enable)
{
    if (!c || (enable != 0 && enable != 1)) return CROMA_HIP_EINVAL;
    ck_stream_sync(c->stream);
    c->nsub_device = enable;
    return 0;
}

int croma_hip_context_set_pipeline(CromaHipContext *c, int enable)
{
    if (!c || (enable != 0 && enable != 1)) return CROMA_HIP_EINVAL;
    ck_stream_sync(c->stream);
    if (c->stream_e) ck_stream_sync(c->stream_e);
    c->pipeline = enable;
    return 0;
}

int croma_hip_context_set_graphs(CromaHipContext *c, int enable)
{
    if (!c || (enable != 0 && enable != 1)) return CROMA_HIP_EINVAL;
    ck_stream_sync(c->stream);
    c->use_graphs = enable;
    return 0;
}

int croma_hip_context_grid(const CromaHipContext *c, int species, double *p, double *dp, double *E, double *pm1,
                           double *pp1)
{
    const size_t nb = c ? (size_t)c->g.n * sizeof(double) : 0;
    if (!c) return CROMA_HIP_EINVAL;
    if (species == 0) {
        if (p) memcpy(p, c->g.pp, nb);
        if (dp) memcpy(dp, c->g.pdp, nb);
        if (E) memcpy(E, c->g.pE, nb);
        if (pm1) *pm1 = c->g.ppm1;
        if (pp1) *pp1 = c->g.ppp1;
    } else {
        if (p) memcpy(p, c->g.ep, nb);
        if (dp) memcpy(dp, c->g.edp, nb);
        if (E) memcpy(E, c->g.eE, nb);
        if (pm1) *pm1 = c->g.epm1;
        if (pp1) *pp1 = c->g.epp1;
    }
    return 0;
}

int croma_hip_ic_cooling_rows_device(CromaHipContext *c, int nz, const double *z, double *rows, double *ghosts)
{
    const int n = c ? c->g.n : 0;
    double *h = 0, *d = 0;
    int q, rc = 0, err = 0;
    if (!c || nz <= 0 || !z || !rows) return CROMA_HIP_EINVAL;
    h = (double *)malloc((size_t)nz * 3 * sizeof(double));
    d = (double *)ck_malloc((size_t)nz * (3 + n + 2) * sizeof(double) + sizeof(int));
    if (!h || !d) {
        free(h);
        ck_free(d);
        return dev_fail("IC rows alloc");
    }
    for (q = 0; q < nz; q++) {
        h[3 * q] = z[q];
        ch_ic_bounds(z[q], &h[3 * q + 1], &h[3 * q + 2]);
    }
    {
        double *d_zb = d, *d_rows = d + (size_t)nz * 3, *d_gh = d_rows + (size_t)nz * n;
        int *d_err = (int *)(d_gh + (size_t)nz * 2);
        if (ck_fill_i32(d_err, 0, 1, c->stream) || ck_h2d(d_zb, h, (size_t)nz * 3 * sizeof(double), c->stream) ||
            ck_ic_rows(n, c->dgrid.ep, c->g.epm1, c->g.epp1, nz, d_zb, d_rows, d_gh, d_err, c->stream) ||
            ck_d2h(rows, d_rows, (size_t)nz * n * sizeof(double), c->stream) ||
            (ghosts && ck_d2h(ghosts, d_gh, (size_t)nz * 2 * sizeof(double), c->stream)) ||
            ck_d2h(&err, d_err, sizeof(int), c->stream) || ck_stream_sync(c->stream))
            rc = dev_fail("IC rows");
    }
    if (!rc && err) {
        ch_set_error("IC cooling row: CMB bounds exceed the device quadrature size");
        rc = CROMA_HIP_EUNSUPPORTED;
    }
    free(h);
    ck_free(d);
    return rc;
}

/* IC cooling rate at the first and last electron bins at z (cached). */
void ch_ctx_ic_ends(CromaHipContext *c, double z, double ends[2])
{
    int k, bins[2];
    pthread_mutex_lock(&c->ic_lock);
    for (k = 0; k < CH_IC_ENDS; k++)
        if (c->ic_end[k].valid && c->ic_end[k].z == z) {
            ends[0] = c->ic_end[k].v[0];
            ends[1] = c->ic_end[k].v[1];
            pthread_mutex_unlock(&c->ic_lock);
            return;
        }
    pthread_mutex_unlock(&c->ic_lock);
    bins[0] = 0;
    bins[1] = c->g.n - 1;
    ch_ic_cooling_bins(&c->g, z, 2, bins, ends);
    pthread_mutex_lock(&c->ic_lock);
    k = c->ic_end_next;
    c->ic_end_next = (k + 1) % CH_IC_ENDS;
    c->ic_end[k].z = z;
    c->ic_end[k].v[0] = ends[0];
    c->ic_end[k].v[1] = ends[1];
    c->ic_end[k].valid = 1;
    pthread_mutex_unlock(&c->ic_lock);
}

int croma_hip_ic_cooling_row(const CromaHipContext *cc, double z, double *row, double *m1, double *p1)
{
    CromaHipContext *c = (CromaHipContext *)cc;
    int k, slot;
    if (!c || !row) return CROMA_HIP_EINVAL;
    pthread_mutex_lock(&c->ic_lock);
    for (k = 0; k < CH_IC_CACHE; k++)
        if (c->ic[k].valid && c->ic[k].z == z) {
            memcpy(row, c->ic[k].row, (size_t)c->g.n * sizeof(double));
            if (m1) *m1 = c->ic[k].m1;
            if (p1) *p1 = c->ic[k].p1;
            pthread_mutex_unlock(&c->ic_lock);
            return 0;
        }
    pthread_mutex_unlock(&c->ic_lock);
    {
        double r[CH_NMAX], a, b;
        ch_ic_cooling_row(&c->g, z, r, &a, &b);
        pthread_mutex_lock(&c->ic_lock);
        slot = c->ic_next;
        c->ic_next = (c->ic_next + 1) % CH_IC_CACHE;
        memcpy(c->ic[slot].row, r, (size_t)c->g.n * sizeof(double));
        c->ic[slot].z = z;
        c->ic[slot].m1 = a;
        c->ic[slot].p1 = b;
        c->ic[slot].valid = 1;
        pthread_mutex_unlock(&c->ic_lock);
        memcpy(row, r, (size_t)c->g.n * sizeof(double));
        if (m1) *m1 = a;
        if (p1) *p1 = b;
    }
    return 0;
}

int croma_hip_nsub_estimate(const CromaHipContext *c, int ntr, double dt_snap, const double *n_gas,
                            const double *kbt, const double *b_field, const double *divv, double ric_lo,
                            double ric_hi, const double *l_turb, const double *dv, int *nsub, int *n_on)
{
    if (!c || ntr < 0 || !n_gas || !kbt || !b_field || !divv || !l_turb || !dv || !nsub || !n_on)
        return CROMA_HIP_EINVAL;
    return ch_nsub_estimate(&c->g, &c->prm, ntr, dt_snap, ric_lo, ric_hi, n_gas, kbt, b_field, divv, l_turb, dv,
                            nsub, n_on) == 0
               ? 0
               : CROMA_HIP_EINVAL;
}

int croma_hip_nsub_estimate_device(const CromaHipContext *c, int ntr, double dt_snap, const double *n_gas,
                                   const double *kbt, const double *b_field, const double *divv, double ric_lo,
                                   double ric_hi, const double *l_turb, const double *dv, int *nsub, int *n_on)
{
    const double *src[6] = {n_gas, kbt, b_field, divv, l_turb, dv};
    double *d;
    int *di, k, rc = 0;
    if (!c || ntr < 0 || !n_gas || !kbt || !b_field || !divv || !l_turb || !dv || !nsub || !n_on || !(dt_snap > 0.0))
        return CROMA_HIP_EINVAL;
    if (ntr == 0) return 0;
    d = (double *)ck_malloc((size_t)ntr * (6 * sizeof(double) + 2 * sizeof(int)));
    if (!d) return dev_fail("nsub buffers");
    di = (int *)(d + (size_t)6 * ntr);
    for (k = 0; k < 6 && !rc; k++)
        if (ck_h2d(d + (size_t)k * ntr, src[k], (size_t)ntr * sizeof(double), c->stream)) rc = 1;
    if (rc || ck_nsub_arrays(c->dgrid, c->prm, ntr, dt_snap, ric_lo, ric_hi, d, d + ntr, d + 2 * (size_t)ntr,
                             d + 3 * (size_t)ntr, d + 4 * (size_t)ntr, d + 5 * (size_t)ntr, di, di + ntr, c->stream) ||
        ck_d2h(nsub, di, (size_t)ntr * sizeof(int), c->stream) ||
        ck_d2h(n_on, di + ntr, (size_t)ntr * sizeof(int), c->stream) || ck_stream_sync(c->stream))
        rc = dev_fail("device nsub");
    ck_free(d);
    return rc;
}

int ch_ctx_fqe(CromaHipContext *c)
{
    const int n = c->g.n;
    const size_t nn = (size_t)n * n;
    double *d_beta = 0;
    int *d_npmin = 0, jp0, je, rc = 0;
    if (c->fqe_ready) return 0;
    if (!c->h_npmin) c->h_npmin = (int *)malloc((size_t)n * sizeof(int));
    if (!c->h_npmin) return CROMA_HIP_ENOMEM;
    if (!c->d_fqe) c->d_fqe = (double *)ck_malloc((nn + (size_t)2 * n) * sizeof(double));
    if (!c->d_fqe) return dev_fail("fqe alloc");
    d_beta = c->d_fqe + nn;
    d_npmin = (int *)(d_beta + n);
    if (ck_h2d(d_beta, c->g.beta_p, (size_t)n * sizeof(double), c->stream) ||
        ck_secondary_weights(n, c->dgrid.eE, c->dgrid.pp, d_beta, c->dgrid.pdp, c->d_fqe, d_npmin, c->stream) ||
        ck_d2h(c->h_npmin, d_npmin, (size_t)n * sizeof(int), c->stream) || ck_stream_sync(c->stream))
        rc = dev_fail("fqe build");
    if (rc) return rc;
    jp0 = n;
    for (je = 0; je < n; je++) {
        const int v = c->h_npmin[je] < 0 ? 0 : c->h_npmin[je];
        if (v < jp0) jp0 = v;
    }
    c->jp0 = jp0;
    c->fqe_ready = 1;
    return 0;
}

int croma_hip_secondary_weights(const CromaHipContext *cc, double *fqe_t, int *np_min_qe)
{
    CromaHipContext *c = (CromaHipContext *)cc;
    int rc;
    if (!c || !fqe_t || !np_min_qe) return CROMA_HIP_EINVAL;
    rc = ch_ctx_fqe(c);
    if (rc) return rc;
    if (ck_d2h(fqe_t, c->d_fqe, (size_t)c->g.n * c->g.n * sizeof(double), c->stream) || ck_stream_sync(c->stream))
        return dev_fail("fqe download");
    memcpy(np_min_qe, c->h_npmin, (size_t)c->g.n * sizeof(int));
    return 0;
}

int croma_hip_emission_kernel_table(const CromaHipContext *cc, int kind, int nbins, double lo, double hi, double z,
                                    double *table, double *elapsed_ms)
{
    CromaHipContext *c = (CromaHipContext *)cc;
    const int n = c ? c->g.n : 0;
    const double *mom;
    double *d_out;
    void *e0 = 0, *e1 = 0;
    int rc = 0;
    if (!c || !table || nbins <= 0 || nbins > 65535 || kind < CROMA_HIP_TABLE_GAMMA || kind > CROMA_HIP_TABLE_IC ||
        !(hi > lo) || !isfinite(lo) || !isfinite(hi) || !isfinite(z))
        return CROMA_HIP_EINVAL;
    mom = kind == CROMA_HIP_TABLE_IC ? c->dgrid.ep : c->dgrid.pp;
    d_out = (double *)ck_malloc((size_t)nbins * n * sizeof(double));
    if (!d_out) return dev_fail("kernel table alloc");
    if (elapsed_ms) {
        e0 = ck_event_create();
        e1 = ck_event_create();
    }
    if ((e0 && ck_event_record(e0, c->stream)) ||
        ck_emission_kernel_table(kind, n, nbins, lo, hi, z, mom, d_out, c->stream) ||
        (e1 && ck_event_record(e1, c->stream)) ||
        ck_d2h(table, d_out, (size_t)nbins * n * sizeof(double), c->stream) || ck_stream_sync(c->stream))
        rc = dev_fail("kernel table");
    if (!rc && elapsed_ms) *elapsed_ms = (e0 && e1) ? (double)ck_event_elapsed_ms(e0, e1) : -1.0;
    if (e0) ck_event_destroy(e0);
    if (e1) ck_event_destroy(e1);
    ck_free(d_out);
    return rc;
}

/* Accessors for the serving loop (croma_serve.c). */
const ChGrid *ch_ctx_grid(const CromaHipContext *c) { return &c->g; }
const CkGrid *ch_ctx_dgrid(const CromaHipContext *c) { return &c->dgrid; }
const CkParams *ch_ctx_prm(const CromaHipContext *c) { return &c->prm; }
int ch_ctx_cu_count(const CromaHipContext *c) { return c->cu_count; }
CromaHipTile ch_ctx_tile(const CromaHipContext *c, int protons) { return c->tile[protons ? 1 : 0]; }
const double *ch_ctx_d_fqe(const CromaHipContext *c) { return c->d_fqe; }
/* np_min of the secondary weights on the device (after the weights and beta_p) */
const int *ch_ctx_d_npmin(const CromaHipContext *c) { return c->d_fqe ? (const int *)(c->d_fqe + (size_t)c->g.n * c->g.n + c->g.n) : 0; }
int ch_ctx_jp0(const CromaHipContext *c) { return c->jp0; }

/* ------------------------------------------------------------------ */
/* task ordering: counting sort of tracers by descending cost, cost =
 * sum over the task's snapshots of nsub x species + 3 x segments (the
 * coefficient builds), bucketed on a log scale (1/256 octave). */

int ch_cost_bucket(double cost)
{
    int k;
    if (cost < 1.0) return 0;
    k = (int)(log2(cost) * 256.0) + 1;
    return k >= CH_HIST ? CH_HIST - 1 : k;
}

static void order_tasks(const CkParams *prm, const int *allow, int nslot, int ntr, const int *nsub, int protons,
                        CkTask *out, int *hist_buf, int *key_buf)
{
    int i, s;
    long pos = 0;
    memset(hist_buf, 0, (size_t)CH_HIST * sizeof(int));
    for (i = 0; i < ntr; i++) {
        double cost = 0.0;
        for (s = 0; s < nslot; s++) {
            const int ns = nsub[(size_t)s * ntr + i];
            cost += (double)ns * (protons ? 2 : 1) + 3.0 * ch_choose_segments(prm, ns, allow[s]);
        }
        key_buf[i] = ch_cost_bucket(cost);
        hist_buf[key_buf[i]]++;
    }
    for (i = CH_HIST - 1; i >= 0; i--) {
        const int cnt = hist_buf[i];
        hist_buf[i] = (int)pos;
        pos += cnt;
    }
    for (i = 0; i < ntr; i++) {
        CkTask *t = &out[hist_buf[key_buf[i]]++];
        t->slot = 0;
        t->tracer = i;
        t->nslot = nslot;
    }
}

/* ------------------------------------------------------------------ */
/* Launch graphs.  The launches of a window (counter fill, fused kernel,
 * GEMM, fused kernel) depend only on the key: kernel shape, buffer
 * addresses and GEMM rows.  The per-window data travel in device buffers
 * uploaded before the launch, so a window whose key equals the previous
 * window's of the same buffer set replays the captured graph.  A new key
 * runs as plain launches once; its first repeat is captured. */
typedef struct {
    long long launches, captures, failures;
} GraphStats;

typedef int (*ch_issue_fn)(void *arg);

static int graph_issue(CromaHipContext *c, void *stream, ChGraph *g, const long key[CH_GKEY], ch_issue_fn fn,
                       void *arg, GraphStats *gs)
{
    void *ex;
    int rc;
    const int same = g->state != 0 && !memcmp(g->key, key, sizeof(g->key));
    if (!c->use_graphs) return fn(arg);
    if (!same) {
        retire_graph(c, g->exec);
        g->exec = 0;
        memcpy(g->key, key, sizeof(g->key));
        g->state = 1;
        return fn(arg);
    }
    if (g->state == -1) return fn(arg);
    if (g->state == 1) {
        if (ck_graph_begin(stream)) {
            g->state = -1;
            gs->failures++;
            return fn(arg);
        }
        rc = fn(arg);
        ex = ck_graph_end(stream);
        if (rc || !ex) {
            /* nothing captured ran: issue the window as plain launches */
            ck_graph_destroy(ex);
            g->state = -1;
            gs->failures++;
            return fn(arg);
        }
        g->exec = ex;
        g->state = 2;
        gs->captures++;
    }
    if (ck_graph_launch(g->exec, stream)) return dev_fail("graph launch");
    gs->launches++;
    return 0;
}

static void graph_reset(CromaHipContext *c, ChGraph *g)
{
    retire_graph(c, g->exec);
    memset(g, 0, sizeof(*g));
}

/* ------------------------------------------------------------------ */
/* A window is the unit of one fused launch: W consecutive snapshots, one
 * queue task per tracer advancing it over all of them.  Two window buffer
 * sets alternate so the host prepares window w+1 while the GPU runs w. */
typedef struct {
    int cap;          /* snapshots the buffers hold */
    int *d_nsub, *d_non;
    CkTask *d_tasks;
    CkSlot *d_slot;
    int *h_nsub, *h_non;     /* pinned [cap][ntr] */
    CkTask *h_tasks;         /* pinned [ntr] */
    CkSlot *h_slot;          /* pinned [cap] */
    int count;               /* snapshots prepared */
    int *d_segoff, *h_segoff; /* split launches: [cap][ntr] first segment row per tracer */
    long segrows;            /* segment rows of the window */
    long segq;               /* segrows rounded up to the batch's GEMM row quantum */
    int split;               /* the window runs as proton pass + GEMM + electron pass */
    void *ev_up;
    ChGraph graph;           /* launches of the last window shape run from this buffer set */
    /* pipelined split windows: the GEMM and the electron pass have graphs
     * of their own (they are issued after waits on the other stream) */
    ChGraph graph_g, graph_e;
    void *ev_g, *ev_e;       /* contraction done, electron pass done (electron stream) */
    int dev_nsub;            /* nsub, n_on, order and segment rows are built on the device */
    int *d_key, *d_segloc, *d_blk, *d_hist; /* their work buffers */
    int nosrc;               /* electron_only window whose snapshots have no source rows */
} WinBuf;

struct CromaHipBatch {
    CromaHipContext *ctx;
    int ntr, species, nsnap_res, protons;
    double *d_crp, *d_cre, *d_qpi, *d_qepri, *d_mass;
    unsigned char *d_dis;
    int has_qpi, has_qepri, has_mass, has_dis;
    double *d_bg;  /* [nsnap_res][8][ntr] */
    double *h_bg;  /* host copy for nsub */
    double *d_src; /* optional [nsnap_res][2][ntr*n] per-snapshot sources */
    unsigned char *src_set;
    WinBuf win[2];
    double *d_segrec, *d_seginje; /* split launches: [segrows_cap][n] */
    long segrows_cap;
    long seg_quantum;             /* GEMM rows are rounded up to a multiple of this */
    int *hist, *keys;
    unsigned long long *d_cells; /* sum nsub of the device-estimated windows of a run */
    void *ev_a, *ev_b;
    void **ev_k; /* per-launch timing events, 2 per window */
    int ev_k_cap;
    /* IC cooling rows of every redshift a run visits, row q at z[icr_s0 + q],
     * built on the device in one launch at the start of the run */
    double *d_icr, *d_icz, *h_icz; /* [icr_cap][n], [icr_cap][3] (z, CMB e0 bounds), pinned copy */
    int icr_cap, icr_s0;
    /* plain path */
    int *d_idx;
    int *h_idx;
};

static void win_free(CromaHipContext *c, WinBuf *w)
{
    graph_reset(c, &w->graph);
    graph_reset(c, &w->graph_g);
    graph_reset(c, &w->graph_e);
    ck_free(w->d_nsub);
    ck_free(w->d_non);
    ck_free(w->d_tasks);
    ck_free(w->d_slot);
    ck_free(w->d_segoff);
    ck_free(w->d_key);
    ck_free(w->d_segloc);
    ck_free(w->d_blk);
    ck_free(w->d_hist);
    ck_host_free(w->h_segoff);
    ck_host_free(w->h_nsub);
    ck_host_free(w->h_non);
    ck_host_free(w->h_tasks);
    ck_host_free(w->h_slot);
    ck_event_destroy(w->ev_up);
    ck_event_destroy(w->ev_g);
    ck_event_destroy(w->ev_e);
    memset(w, 0, sizeof(*w));
}

static int win_reserve(CromaHipContext *c, WinBuf *w, int cap, int ntr)
{
    if (w->cap >= cap) return 0;
    if (w->ev_up) ck_event_sync(w->ev_up);
    ck_stream_sync(c->stream);
    if (c->stream_e) ck_stream_sync(c->stream_e);
    win_free(c, w);
    w->d_nsub = (int *)ck_malloc((size_t)cap * ntr * sizeof(int));
    w->d_non = (int *)ck_malloc((size_t)cap * ntr * sizeof(int));
    w->d_tasks = (CkTask *)ck_malloc((size_t)ntr * sizeof(CkTask));
    w->d_slot = (CkSlot *)ck_malloc((size_t)cap * sizeof(CkSlot));
    w->h_nsub = (int *)ck_host_alloc((size_t)cap * ntr * sizeof(int));
    w->h_non = (int *)ck_host_alloc((size_t)cap * ntr * sizeof(int));
    w->h_tasks = (CkTask *)ck_host_alloc((size_t)ntr * sizeof(CkTask));
    w->h_slot = (CkSlot *)ck_host_alloc((size_t)cap * sizeof(CkSlot));
    w->d_segoff = (int *)ck_malloc((size_t)cap * ntr * sizeof(int));
    w->h_segoff = (int *)ck_host_alloc((size_t)cap * ntr * sizeof(int));
    w->d_key = (int *)ck_malloc((size_t)ntr * sizeof(int));
    w->d_segloc = (int *)ck_malloc((size_t)ntr * sizeof(int));
    w->d_blk = (int *)ck_malloc(((size_t)ntr / 256 + 2) * sizeof(int));
    w->d_hist = (int *)ck_malloc((size_t)CK_NSUB_HIST * sizeof(int));
    w->ev_up = ck_event_create();
    w->ev_g = ck_event_create();
    w->ev_e = ck_event_create();
    if (!w->d_nsub || !w->d_non || !w->d_tasks || !w->d_slot || !w->h_nsub || !w->h_non ||
        !w->h_tasks || !w->h_slot || !w->ev_up || !w->ev_g || !w->ev_e || !w->d_segoff || !w->h_segoff ||
        !w->d_key || !w->d_segloc || !w->d_blk || !w->d_hist) {
        win_free(c, w);
        return dev_fail("window buffers");
    }
    w->cap = cap;
    return 0;
}

void croma_hip_batch_destroy(CromaHipBatch *b)
{
    int k;
    if (!b) return;
    if (b->ctx && b->ctx->stream) ck_stream_sync(b->ctx->stream);
    if (b->ctx && b->ctx->stream_e) ck_stream_sync(b->ctx->stream_e);
    for (k = 0; k < 2; k++) win_free(b->ctx, &b->win[k]);
    if (b->ctx) flush_retired(b->ctx);
    ck_free(b->d_crp);
    ck_free(b->d_cre);
    ck_free(b->d_qpi);
    ck_free(b->d_qepri);
    ck_free(b->d_mass);
    ck_free(b->d_dis);
    ck_free(b->d_bg);
    ck_free(b->d_src);
    ck_free(b->d_idx);
    ck_free(b->d_segrec);
    ck_free(b->d_seginje);
    ck_free(b->d_cells);
    ck_free(b->d_icr);
    ck_free(b->d_icz);
    ck_host_free(b->h_icz);
    free(b->h_bg);
    free(b->src_set);
    free(b->hist);
    free(b->keys);
    ck_host_free(b->h_idx);
    ck_event_destroy(b->ev_a);
    ck_event_destroy(b->ev_b);
    for (k = 0; k < b->ev_k_cap; k++) ck_event_destroy(b->ev_k[k]);
    free(b->ev_k);
    free(b);
}

CromaHipBatch *croma_hip_batch_create(CromaHipContext *c, int ntr, int species, int nsnap_res)
{
    CromaHipBatch *b;
    const int n = c ? c->g.n : 0;
    const size_t sn = (size_t)ntr * n * sizeof(double);
    int k;
    if (!c || ntr <= 0 || nsnap_res <= 0 ||
        !(species == CROMA_HIP_SPECIES_ELECTRON || species == CROMA_HIP_SPECIES_PROTON ||
          species == CROMA_HIP_SPECIES_ELECTRON_PROTON)) {
        ch_set_error("invalid batch arguments");
        return 0;
    }
    b = (CromaHipBatch *)calloc(1, sizeof(*b));
    if (!b) return 0;
    b->ctx = c;
    b->ntr = ntr;
    b->species = species;
    b->protons = (species != CROMA_HIP_SPECIES_ELECTRON);
    b->nsnap_res = nsnap_res;
    b->d_cre = (double *)ck_malloc(sn);
    b->d_crp = b->protons ? (double *)ck_malloc(sn) : 0;
    b->d_bg = (double *)ck_malloc((size_t)nsnap_res * CK_NBG * ntr * sizeof(double));
    b->h_bg = (double *)malloc((size_t)nsnap_res * CK_NBG * ntr * sizeof(double));
    b->hist = (int *)malloc(CH_HIST * sizeof(int));
    b->keys = (int *)malloc((size_t)ntr * sizeof(int));
    b->src_set = (unsigned char *)calloc((size_t)nsnap_res, 1);
    b->ev_a = ck_event_create();
    b->ev_b = ck_event_create();
    b->d_cells = (unsigned long long *)ck_malloc(256);
    if (!b->d_cre || (b->protons && !b->d_crp) || !b->d_bg || !b->h_bg || !b->hist || !b->keys || !b->src_set ||
        !b->ev_a || !b->ev_b || !b->d_cells)
        goto fail;
    for (k = 0; k < 2; k++)
        if (win_reserve(c, &b->win[k], 1, ntr)) goto fail;
    if (b->protons && ch_ctx_fqe(c) != 0) goto fail;
    return b;
fail:
    if (!g_err[0]) ch_set_error("batch allocation failed: %s", ck_last_error());
    croma_hip_batch_destroy(b);
    return 0;
}

static int upload_opt(double **dst, const double *src, size_t bytes, void *stream)
{
    if (!src) return 0;
    if (!*dst) *dst = (double *)ck_malloc(bytes);
    if (!*dst) return -1;
    return ck_h2d(*dst, src, bytes, stream);
}

int croma_hip_batch_set_state(CromaHipBatch *b, const double *crp, const double *cre, const double *qpi,
                              const double *qepri, const double *mass, const unsigned char *dis)
{
    CromaHipContext *c;
    size_t sn;
    if (!b || !cre) return CROMA_HIP_EINVAL;
    c = b->ctx;
    sn = (size_t)b->ntr * c->g.n * sizeof(double);
    if (!b->protons && (crp || qpi)) {
        /* electron_only requires crp == 0 and qpi == 0 exactly */
        size_t i, m = (size_t)b->ntr * c->g.n;
        for (i = 0; i < m; i++)
            if ((crp && crp[i] != 0.0) || (qpi && qpi[i] != 0.0)) {
                ch_set_error("electron_only batch given non-zero proton state or source");
                return CROMA_HIP_EINVAL;
            }
    }
    if (ck_h2d(b->d_cre, cre, sn, c->stream)) return dev_fail("state upload");
    if (b->protons) {
        if (crp) {
            if (ck_h2d(b->d_crp, crp, sn, c->stream)) return dev_fail("state upload");
        } else if (ck_fill_f64(b->d_crp, 0.0, (size_t)b->ntr * c->g.n, c->stream)) {
            return dev_fail("state clear");
        }
        if (upload_opt(&b->d_qpi, qpi, sn, c->stream)) return dev_fail("qpi upload");
        b->has_qpi = qpi != 0;
    }
    if (upload_opt(&b->d_qepri, qepri, sn, c->stream)) return dev_fail("qepri upload");
    b->has_qepri = qepri != 0;
    if (upload_opt(&b->d_mass, mass, (size_t)b->ntr * sizeof(double), c->stream)) return dev_fail("mass upload");
    b->has_mass = mass != 0;
    if (dis) {
        if (!b->d_dis) b->d_dis = (unsigned char *)ck_malloc((size_t)b->ntr);
        if (!b->d_dis || ck_h2d(b->d_dis, dis, (size_t)b->ntr, c->stream)) return dev_fail("dis upload");
    }
    b->has_dis = dis != 0;
    return ck_stream_sync(c->stream) ? dev_fail("state sync") : 0;
}

int ch_batch_state(CromaHipBatch *b, ChBatchState *out)
{
    if (!b || !out) return CROMA_HIP_EINVAL;
    out->d_cre = b->d_cre;
    out->d_crp = b->protons ? b->d_crp : 0;
    out->ntr = b->ntr;
    out->n = b->ctx->g.n;
    out->protons = b->protons;
    out->stream = b->ctx->stream;
    return 0;
}

int croma_hip_batch_get_state(CromaHipBatch *b, double *crp, double *cre)
{
    CromaHipContext *c;
    size_t sn;
    if (!b) return CROMA_HIP_EINVAL;
    c = b->ctx;
    sn = (size_t)b->ntr * c->g.n * sizeof(double);
    if (cre && ck_d2h(cre, b->d_cre, sn, c->stream)) return dev_fail("state download");
    if (crp) {
        if (b->protons) {
            if (ck_d2h(crp, b->d_crp, sn, c->stream)) return dev_fail("state download");
        } else {
            memset(crp, 0, sn);
        }
    }
    return ck_stream_sync(c->stream) ? dev_fail("state sync") : 0;
}

int croma_hip_batch_set_background(CromaHipBatch *b, int isnap, const double *const fields[8])
{
    int f;
    if (!b || !fields || isnap < 0 || isnap >= b->nsnap_res) return CROMA_HIP_EINVAL;
    for (f = 0; f < CK_NBG; f++) {
        const size_t off = ((size_t)isnap * CK_NBG + f) * b->ntr;
        if (!fields[f]) return CROMA_HIP_EINVAL;
        memcpy(b->h_bg + off, fields[f], (size_t)b->ntr * sizeof(double));
        if (ck_h2d(b->d_bg + off, b->h_bg + off, (size_t)b->ntr * sizeof(double), b->ctx->stream))
            return dev_fail("background upload");
    }
    return ck_stream_sync(b->ctx->stream) ? dev_fail("background sync") : 0;
}

int croma_hip_batch_set_sources(CromaHipBatch *b, int isnap, const double *qpi, const double *qepri)
{
    const size_t m = b ? (size_t)b->ntr * b->ctx->g.n : 0;
    if (!b || isnap < 0 || isnap >= b->nsnap_res) return CROMA_HIP_EINVAL;
    if (!qpi && !qepri) {
        b->src_set[isnap] = 0;
        return 0;
    }
    if (!b->d_src) {
        b->d_src = (double *)ck_malloc((size_t)b->nsnap_res * 2 * m * sizeof(double));
        if (!b->d_src) return dev_fail("source alloc");
    }
    {
        double *dp = b->d_src + (size_t)isnap * 2 * m;
        if (qpi) {
            if (!b->protons) {
                size_t i;
                for (i = 0; i < m; i++)
                    if (qpi[i] != 0.0) {
                        ch_set_error("electron_only batch given a proton source");
                        return CROMA_HIP_EINVAL;
                    }
            }
            if (ck_h2d(dp, qpi, m * sizeof(double), b->ctx->stream)) return dev_fail("source upload");
        } else if (ck_fill_f64(dp, 0.0, m, b->ctx->stream)) {
            return dev_fail("source clear");
        }
        if (qepri) {
            if (ck_h2d(dp + m, qepri, m * sizeof(double), b->ctx->stream)) return dev_fail("source upload");
        } else if (ck_fill_f64(dp + m, 0.0, m, b->ctx->stream)) {
            return dev_fail("source clear");
        }
    }
    b->src_set[isnap] = 1;
    return ck_stream_sync(b->ctx->stream) ? dev_fail("source sync") : 0;
}

/* Host-side preparation of snapshots [s, s + count) into window w. */
typedef struct {
    double ic_ms, nsub_ms, order_ms;
    long long cell_steps, species_steps;
} PrepStats;

static int prepare_window(CromaHipBatch *b, WinBuf *w, int s_first, int count, int nsnap_total, const double *z,
                          const double *dt_snap, double elapsed, int split, int dev_nsub, PrepStats *ps)
{
    CromaHipContext *c = b->ctx;
    const int n = c->g.n, ntr = b->ntr;
    int allow[64], *allow_v = allow;
    double t0;
    int q, i, f;
    long long cs = 0;

    if (count > 64) {
        allow_v = (int *)malloc((size_t)count * sizeof(int));
        if (!allow_v) return CROMA_HIP_ENOMEM;
    }
    w->split = split;
    w->dev_nsub = dev_nsub;
    w->segrows = 0;
    w->nosrc = !b->protons;
    /* the IC rows come from the run's table (run_ic_rows); the host only
     * evaluates the two end bins nsub needs (same arithmetic as the
     * reference row) */
    for (q = 0; q < count; q++) {
        const int s = s_first + q;
        const int al = (s + 1 < nsnap_total) ? 1 : 0;
        const int sn = (s + 1 < b->nsnap_res && al) ? s + 1 : s;
        const double *hb = b->h_bg + (size_t)s * CK_NBG * ntr;
        int *nsub = w->h_nsub + (size_t)q * ntr, *non = w->h_non + (size_t)q * ntr;
        CkSlot *S = &w->h_slot[q];
        double ends[2];
        if (al && s + 1 >= b->nsnap_res) {
            ch_set_error("snapshot %d needs background slot %d (resident %d)", s, s + 1, b->nsnap_res);
            if (allow_v != allow) free(allow_v);
            return CROMA_HIP_EINVAL;
        }
        allow_v[q] = al;
        t0 = ch_now_s();
        ch_ctx_ic_ends(c, z[s], ends);
        ps->ic_ms += 1e3 * (ch_now_s() - t0);
        if (!dev_nsub) {
            t0 = ch_now_s();
            ch_nsub_estimate(&c->g, &c->prm, ntr, dt_snap[s], ends[0], ends[1], hb + 0 * (size_t)ntr,
                             hb + 1 * (size_t)ntr, hb + 2 * (size_t)ntr, hb + 3 * (size_t)ntr, hb + 4 * (size_t)ntr,
                             hb + 5 * (size_t)ntr, nsub, non);
            ps->nsub_ms += 1e3 * (ch_now_s() - t0);
            for (i = 0; i < ntr; i++) cs += nsub[i];
        }

        memset(S, 0, sizeof(*S));
        S->crp = b->d_crp;
        S->cre = b->d_cre;
        if (b->d_src && b->src_set[s]) {
            const size_t m = (size_t)ntr * n;
            S->qpi = b->protons ? b->d_src + (size_t)s * 2 * m : 0;
            S->qepri = b->d_src + (size_t)s * 2 * m + m;
        } else {
            S->qpi = (b->protons && b->has_qpi) ? b->d_qpi : 0;
            S->qepri = b->has_qepri ? b->d_qepri : 0;
        }
        if (S->qepri) w->nosrc = 0;
        S->mass = b->has_mass ? b->d_mass : 0;
        S->dis = b->has_dis ? b->d_dis : 0;
        for (f = 0; f < CK_NBG; f++) {
            S->bgc[f] = b->d_bg + ((size_t)s * CK_NBG + f) * ntr;
            S->bgn[f] = b->d_bg + ((size_t)sn * CK_NBG + f) * ntr;
        }
        S->ric_c = b->d_icr + (size_t)(s - b->icr_s0) * n;
        S->ric_n = al ? S->ric_c + n : S->ric_c;
        S->nsub = w->d_nsub + (size_t)q * ntr;
        S->n_on = w->d_non + (size_t)q * ntr;
        S->fqe_t = b->protons ? c->d_fqe : 0;
        S->jp0 = c->jp0;
        S->ntr = ntr;
        S->allow_interp = al;
        S->protons = b->protons;
        S->dt_snap = dt_snap[s];
        S->elapsed = elapsed;
        S->prm = c->prm;
        S->ric_lo = ends[0];
        S->ric_hi = ends[1];
        if (split) {
            int *off = w->h_segoff + (size_t)q * ntr;
            if (!dev_nsub)
                for (i = 0; i < ntr; i++) {
                    off[i] = (int)w->segrows;
                    w->segrows += ch_choose_segments(&c->prm, nsub[i], al);
                }
            S->seg_off = w->d_segoff + (size_t)q * ntr;
            S->seg_rec = b->d_segrec;
            S->seg_inje = b->d_seginje;
        }
        elapsed += dt_snap[s];
    }
    if (!dev_nsub) {
        w->segq = split ? (w->segrows + b->seg_quantum - 1) / b->seg_quantum * b->seg_quantum : 0;
        t0 = ch_now_s();
        order_tasks(&c->prm, allow_v, count, ntr, w->h_nsub, b->protons, w->h_tasks, b->hist, b->keys);
        ps->order_ms += 1e3 * (ch_now_s() - t0);
        ps->cell_steps += cs;
        ps->species_steps += cs * (b->protons ? 2 : 1);
    } else {
        /* the device assigns the segment rows: the contraction's launch
         * covers the batch's capacity and stops at the device-side count */
        w->segq = split ? b->segrows_cap : 0;
    }
    w->count = count;
    if (allow_v != allow) free(allow_v);
    return 0;
}

/* IC cooling rows at z[s0] .. z[s1 - 1] (+ z[s1] when snapshot s1 - 1 has a
 * successor) for the whole run, in one launch: one row alone is a few hundred
 * blocks that cannot fill the device, so the rows of a run are built
 * together.  A first row at the (z, lo, hi) of the previous run's last row is
 * copied rather than rebuilt (the rows depend on z and the grid only). */
static int run_ic_rows(CromaHipBatch *b, int s0, int s1, int nsnap_total, const double *z)
{
    CromaHipContext *c = b->ctx;
    const int n = c->g.n, nz = s1 - s0 + ((s1 < nsnap_total) ? 1 : 0);
    int q, reuse;
    if (nz > b->icr_cap) {
        ck_stream_sync(c->stream);
        ck_free(b->d_icr);
        ck_free(b->d_icz);
        ck_host_free(b->h_icz);
        b->d_icr = (double *)ck_malloc((size_t)nz * n * sizeof(double));
        b->d_icz = (double *)ck_malloc((size_t)nz * 3 * sizeof(double));
        b->h_icz = (double *)ck_host_alloc((size_t)nz * 3 * sizeof(double));
        b->icr_cap = (b->d_icr && b->d_icz && b->h_icz) ? nz : 0;
        if (!b->icr_cap) return dev_fail("IC row table");
    }
    if (!c->d_ic_last && !(c->d_ic_last = (double *)ck_malloc((size_t)n * sizeof(double))))
        return dev_fail("IC row cache");
    for (q = 0; q < nz; q++) {
        b->h_icz[3 * q] = z[s0 + q];
        ch_ic_bounds(z[s0 + q], &b->h_icz[3 * q + 1], &b->h_icz[3 * q + 2]);
    }
    reuse = c->ic_last_valid && c->ic_last_zb[0] == b->h_icz[0] && c->ic_last_zb[1] == b->h_icz[1] &&
            c->ic_last_zb[2] == b->h_icz[2];
    if (ck_h2d(b->d_icz, b->h_icz, (size_t)nz * 3 * sizeof(double), c->stream) ||
        (reuse && ck_d2d(b->d_icr, c->d_ic_last, (size_t)n * sizeof(double), c->stream)) ||
        (nz > reuse && ck_ic_rows(n, c->dgrid.ep, c->g.epm1, c->g.epp1, nz - reuse, b->d_icz + 3 * reuse,
                                  b->d_icr + (size_t)reuse * n, 0, c->d_err, c->stream)) ||
        ck_d2d(c->d_ic_last, b->d_icr + (size_t)(nz - 1) * n, (size_t)n * sizeof(double), c->stream))
        return dev_fail("IC rows");
    memcpy(c->ic_last_zb, b->h_icz + 3 * (nz - 1), sizeof(c->ic_last_zb));
    c->ic_last_valid = 1;
    b->icr_s0 = s0;
    return 0;
}

static int upload_window(CromaHipBatch *b, WinBuf *w)
{
    CromaHipContext *c = b->ctx;
    const int ntr = b->ntr, q = w->count;
    const int host = !w->dev_nsub;
    if ((host && ck_h2d(w->d_nsub, w->h_nsub, (size_t)q * ntr * sizeof(int), c->stream)) ||
        (host && ck_h2d(w->d_non, w->h_non, (size_t)q * ntr * sizeof(int), c->stream)) ||
        (host && ck_h2d(w->d_tasks, w->h_tasks, (size_t)ntr * sizeof(CkTask), c->stream)) ||
        ck_h2d(w->d_slot, w->h_slot, (size_t)q * sizeof(CkSlot), c->stream) ||
        (host && w->split && ck_h2d(w->d_segoff, w->h_segoff, (size_t)q * ntr * sizeof(int), c->stream)) ||
        ck_event_record(w->ev_up, c->stream))
        return dev_fail("window upload");
    return 0;
}

/* A fused pass fills the device (cu_count * waves_per_simd blocks of four
 * waves), or half of it when it runs beside the other species' pass of the
 * neighbouring window; the two passes then take their tasks from separate
 * queue counters. */
static int launch_mode(CromaHipContext *c, int protons, int mode, const CkSlot *d_slot, const CkTask *d_tasks,
                       int ntasks, int shared, int *counter, void *stream)
{
    const CromaHipTile *t = &c->tile[protons ? 1 : 0];
    const int nfull = c->cu_count * t->waves_per_simd;
    const int nblocks = shared ? (nfull + 1) / 2 : nfull;
    if (c->coeff_fp32 && !ck_fused_f32_supported(c->g.n, t->rows_per_lane, t->waves_per_simd, mode)) {
        ch_set_error("FP32 coefficients: no instance for n=%d, %d rows per lane, %d waves per SIMD, mode %d",
                     c->g.n, t->rows_per_lane, t->waves_per_simd, mode);
        return CROMA_HIP_EUNSUPPORTED;
    }
    if (ck_fill_i32(counter, 0, 1, stream)) return dev_fail("counter reset");
    if (ck_launch_fused(c->g.n, t->rows_per_lane, t->waves_per_simd, nblocks,
                        mode + (c->coeff_fp32 ? CK_LAUNCH_F32 : 0), c->dgrid, d_slot, d_tasks, ntasks, counter,
                        c->d_err, stream))
        return dev_fail("fused launch");
    return 0;
}

/* nosrc: an electron_only launch whose slots have no source rows (the
 * instance without source terms; the same results) */
static int launch_tasks(CromaHipContext *c, int protons, int nosrc, const CkSlot *d_slot, const CkTask *d_tasks,
                        int ntasks)
{
    return launch_mode(c, protons, protons ? 1 : (nosrc ? CK_LAUNCH_NOSRC : 0), d_slot, d_tasks, ntasks, 0,
                       c->d_counter, c->stream);
}

/* queue counter of the electron pass of a pipelined split window (its own
 * 128-byte line of the counter block) */
#define CH_CTR_E 32

/* Split two-species window: proton pass recording the segment-start rows,
 * the contraction of every row against the secondary weights
 * (croma_secondary.hip), then the electron pass.  With the schedule built on
 * the device the contraction reads the row count from the window's segment
 * total (d_blk[nblk], written by the ordering kernels) and w->segq is the
 * batch's capacity; with the host schedule w->segq is the count rounded up
 * to the row quantum (rows past the count hold stale finite rows whose
 * products are never read).  Either way the launch depends only on w->segq,
 * which keys the window's graph. */
static int launch_gemm(CromaHipBatch *b, WinBuf *w)
{
    CromaHipContext *c = b->ctx;
    const int nblk = (b->ntr + 255) / 256;
    if (w->segq > 0 && ck_sec_contract(c->g.n, c->d_fqe, ch_ctx_d_npmin(c), b->d_segrec, b->d_seginje, w->segq,
                                       w->dev_nsub ? w->d_blk + nblk : 0, 8 * c->cu_count, c->stream))
        return dev_fail("secondary contraction");
    return 0;
}

static int launch_split(CromaHipBatch *b, WinBuf *w)
{
    CromaHipContext *c = b->ctx;
    int rc = launch_mode(c, 1, 2, w->d_slot, w->d_tasks, b->ntr, 0, c->d_counter, c->stream);
    if (!rc) rc = launch_gemm(b, w);
    if (!rc) rc = launch_mode(c, 1, 3, w->d_slot, w->d_tasks, b->ntr, 0, c->d_counter, c->stream);
    return rc;
}

typedef struct {
    CromaHipBatch *b;
    WinBuf *w;
    int pipe;   /* pipelined split window: this issue is the proton pass only */
    int shared; /* the pass runs beside the other species' pass */
} WinIssue;

static int issue_window(void *arg)
{
    WinIssue *wi = (WinIssue *)arg;
    CromaHipBatch *b = wi->b;
    WinBuf *w = wi->w;
    CromaHipContext *c = b->ctx;
    if (w->dev_nsub && ck_nsub_order(c->dgrid, w->d_slot, w->count, b->ntr, b->protons, w->split, w->d_tasks,
                                     w->d_key, w->d_segloc, w->d_blk, w->d_hist, b->d_cells, c->stream))
        return dev_fail("device nsub");
    if (wi->pipe) return launch_mode(c, 1, 2, w->d_slot, w->d_tasks, b->ntr, wi->shared, c->d_counter, c->stream);
    if (w->split) return launch_split(b, w);
    return launch_tasks(c, b->protons, w->nosrc, w->d_slot, w->d_tasks, b->ntr);
}

static int issue_gemm(void *arg)
{
    WinIssue *wi = (WinIssue *)arg;
    return launch_gemm(wi->b, wi->w);
}

static int issue_epass(void *arg)
{
    WinIssue *wi = (WinIssue *)arg;
    CromaHipContext *c = wi->b->ctx;
    return launch_mode(c, 1, 3, wi->w->d_slot, wi->w->d_tasks, wi->b->ntr, wi->shared, c->d_counter + CH_CTR_E,
                       c->stream_e);
}

static int launch_window(CromaHipBatch *b, WinBuf *w, WinBuf *prev, int last, GraphStats *gs)
{
    CromaHipContext *c = b->ctx;
    const CromaHipTile *t = &c->tile[b->protons ? 1 : 0];
    const int pipe = w->split && c->pipeline;
    WinIssue wi;
    long key[CH_GKEY];
    int rc;
    memset(key, 0, sizeof(key));
    key[0] = w->split | (pipe << 1) | ((pipe && prev) << 2);
    key[1] = w->segq;
    key[2] = t->rows_per_lane;
    key[3] = t->waves_per_simd;
    key[4] = (long)(size_t)b->d_segrec;
    key[5] = b->protons | ((long)c->coeff_fp32 << 1) | ((long)w->nosrc << 2);
    key[6] = c->jp0 | ((long)w->dev_nsub << 20) | ((long)w->count << 21);
    key[7] = b->ntr;
    wi.b = b;
    wi.w = w;
    wi.pipe = pipe;
    wi.shared = prev != 0;
    if (!pipe) return graph_issue(c, c->stream, &w->graph, key, issue_window, &wi, gs);
    /* Pipelined split window w on the library stream L and the electron
     * stream E:
     *   L: nsub(w), proton pass(w)    beside E: electron pass(w-1)
     *   L: wait E done(w-1), GEMM(w), record G(w)
     *   E: wait G(w), electron pass(w), record E done(w)
     * The proton pass of w+1 writes the segment rows only after GEMM(w) read
     * them (both on L); GEMM(w+1) overwrites the GEMM rows only after the
     * electron pass of w read them (the wait on E); the upload of w+1 into
     * the buffer set of w-1 follows the wait on its electron pass.  The last
     * window's electron pass runs alone and L waits for it. */
    if (!c->stream_e && !(c->stream_e = ck_stream_create())) return dev_fail("electron stream");
    if ((rc = graph_issue(c, c->stream, &w->graph, key, issue_window, &wi, gs)) != 0) return rc;
    if (prev && ck_stream_wait_event(c->stream, prev->ev_e)) return dev_fail("stream wait");
    if ((rc = graph_issue(c, c->stream, &w->graph_g, key, issue_gemm, &wi, gs)) != 0) return rc;
    if (ck_event_record(w->ev_g, c->stream) || ck_stream_wait_event(c->stream_e, w->ev_g))
        return dev_fail("stream wait");
    wi.shared = !last;
    key[0] |= (long)wi.shared << 3;
    if ((rc = graph_issue(c, c->stream_e, &w->graph_e, key, issue_epass, &wi, gs)) != 0) return rc;
    if (ck_event_record(w->ev_e, c->stream_e) || (last && ck_stream_wait_event(c->stream, w->ev_e)))
        return dev_fail("stream wait");
    return 0;
}

/* Plain path: reference GPU semantics (per bucket of equal (nsub, n_on),
 * per coefficient segment: lerp, coefficient kernels, secondary, Thomas). */
static int ensure_plain_ws(CromaHipContext *c, int count)
{
    const int n = c->g.n;
    size_t per, tot;
    char *m;
    if (c->pws_cap >= count) return 0;
    ck_stream_sync(c->stream);
    ck_free(c->pws_mem);
    c->pws_mem = 0;
    c->pws_cap = 0;
    per = (size_t)count * n * sizeof(double);
    tot = per * (7 + 12 + 1 + 1 + 4) + (size_t)CK_NBG * count * sizeof(double) + (size_t)n * sizeof(double) +
          (size_t)count * sizeof(double) + (size_t)count + 8192;
    c->pws_mem = ck_malloc(tot);
    if (!c->pws_mem) return dev_fail("plain workspace");
    m = (char *)c->pws_mem;
#define TAKE(field, bytes)                                                                                             \
    do {                                                                                                               \
        c->pws.field = (void *)m;                                                                                      \
        m += ((bytes) + 255) & ~(size_t)255;                                                                           \
    } while (0)
    TAKE(radp_p, per);
    TAKE(tl_p, per);
    TAKE(radp_e, per);
    TAKE(dpp_on_p, per);
    TAKE(dpp_on_e, per);
    TAKE(dpp_off_p, per);
    TAKE(dpp_off_e, per);
    TAKE(a, 4 * per);
    TAKE(b, 4 * per);
    TAKE(c, 4 * per);
    TAKE(inje, per);
    TAKE(cprime, per);
    TAKE(xp, per);
    TAKE(xe, per);
    TAKE(qp, per);
    TAKE(qe, per);
    TAKE(bg, (size_t)CK_NBG * count * sizeof(double));
    TAKE(ric, (size_t)n * sizeof(double));
    TAKE(mass, (size_t)count * sizeof(double));
    TAKE(dis, (size_t)count);
#undef TAKE
    c->pws.ric_js = 1;
    c->pws.ric_is = 0;
    c->pws_cap = count;
    return 0;
}

int ch_ctx_plain_ws(CromaHipContext *c, int count, CkPlainWs *out)
{
    const int rc = ensure_plain_ws(c, count);
    if (!rc) *out = c->pws;
    return rc;
}

void *ch_ctx_stream(const CromaHipContext *c) { return c->stream; }
int *ch_ctx_d_err(const CromaHipContext *c) { return c->d_err; }

/* Reads and clears the context's error word (after the stream drains):
 * CROMA_HIP_EPIVOT / CROMA_HIP_ENONFINITE or 0. */
int ch_ctx_take_err(CromaHipContext *c)
{
    int err = 0;
    if (ck_d2h(&err, c->d_err, sizeof(int), c->stream) || ck_stream_sync(c->stream)) return dev_fail("error word");
    if (!err) return 0;
    if (ck_fill_i32(c->d_err, 0, 1, c->stream) || ck_stream_sync(c->stream)) return dev_fail("error word");
    ch_set_error((err & 1) ? "Thomas pivot below 1e-300" : "non-finite spectrum");
    return (err & 1) ? CROMA_HIP_EPIVOT : CROMA_HIP_ENONFINITE;
}

/* LSD radix sort of (key, index) pairs on 64-bit keys, 16 bits per pass. */
static int radix_sort_pairs(unsigned long long *key, int *idx, long n)
{
    unsigned long long *k2 = (unsigned long long *)malloc((size_t)n * sizeof(*k2));
    int *i2 = (int *)malloc((size_t)n * sizeof(int));
    long *cnt = (long *)malloc(65536 * sizeof(long));
    int pass;
    if (!k2 || !i2 || !cnt) {
        free(k2);
        free(i2);
        free(cnt);
        return -1;
    }
    for (pass = 0; pass < 4; pass++) {
        const int sh = 16 * pass;
        long i, pos = 0;
        memset(cnt, 0, 65536 * sizeof(long));
        for (i = 0; i < n; i++) cnt[(key[i] >> sh) & 0xffff]++;
        for (i = 0; i < 65536; i++) {
            const long c = cnt[i];
            cnt[i] = pos;
            pos += c;
        }
        for (i = 0; i < n; i++) {
            const long d = cnt[(key[i] >> sh) & 0xffff]++;
            k2[d] = key[i];
            i2[d] = idx[i];
        }
        memcpy(key, k2, (size_t)n * sizeof(*key));
        memcpy(idx, i2, (size_t)n * sizeof(int));
    }
    free(k2);
    free(i2);
    free(cnt);
    return 0;
}

/* Plain path: the reference GPU semantics.  Tracers are bucketed by
 * (nsub, n_on) (tracer_fp_bucket.c:208-262); every bucket runs its
 * coefficient segments with separate launches: background lerp, the
 * coefficient kernel, the secondary source and the thread-per-tracer Thomas
 * substeps (fp_cuda_tracer_backend.cu:335-830). */
static int run_plain_snapshot(CromaHipBatch *b, WinBuf *w)
{
    CromaHipContext *c = b->ctx;
    const CkSlot *S = &w->h_slot[0];
    const int n = c->g.n, ntr = b->ntr;
    const int *nsub = w->h_nsub, *non = w->h_non;
    unsigned long long *key;
    long i0;
    int maxb = 0;
    if (!b->h_idx) {
        b->h_idx = (int *)ck_host_alloc((size_t)ntr * sizeof(int));
        b->d_idx = (int *)ck_malloc((size_t)ntr * sizeof(int));
        if (!b->h_idx || !b->d_idx) return dev_fail("plain index");
    }
    /* the previous snapshot's kernels may still read d_idx / h_idx */
    if (ck_stream_sync(c->stream)) return dev_fail("plain sync");
    key = (unsigned long long *)malloc((size_t)ntr * sizeof(*key));
    if (!key) return CROMA_HIP_ENOMEM;
    for (i0 = 0; i0 < ntr; i0++) {
        key[i0] = ((unsigned long long)(unsigned)nsub[i0] << 32) | (unsigned)non[i0];
        b->h_idx[i0] = (int)i0;
    }
    if (radix_sort_pairs(key, b->h_idx, ntr)) {
        free(key);
        return CROMA_HIP_ENOMEM;
    }
    for (i0 = 0; i0 < ntr;) {
        long i1 = i0 + 1;
        while (i1 < ntr && key[i1] == key[i0]) i1++;
        if (i1 - i0 > maxb) maxb = (int)(i1 - i0);
        i0 = i1;
    }
    if (ensure_plain_ws(c, maxb) || ck_h2d(b->d_idx, b->h_idx, (size_t)ntr * sizeof(int), c->stream)) {
        free(key);
        return dev_fail("plain setup");
    }
    for (i0 = 0; i0 < ntr;) {
        long i1 = i0 + 1;
        int count, ns, on, nseg, iseg, on_start, on_end;
        const int *didx;
        while (i1 < ntr && key[i1] == key[i0]) i1++;
        count = (int)(i1 - i0);
        ns = (int)(key[i0] >> 32);
        on = (int)(key[i0] & 0xffffffffu);
        didx = b->d_idx + i0;
        nseg = ch_choose_segments(&c->prm, ns, S->allow_interp);
        on_start = (ns - on) / 2;
        on_end = on_start + on;
        if (ck_plain_gather(n, count, didx, S, c->pws, c->stream)) {
            free(key);
            return dev_fail("plain gather");
        }
        for (iseg = 0; iseg < nseg; iseg++) {
            const int fb = (int)(((long long)iseg * ns) / nseg);
            const int fe = (int)(((long long)(iseg + 1) * ns) / nseg);
            const int steps = fe - fb;
            const double alpha = (nseg > 1) ? ((double)iseg + 0.5) / (double)nseg : 0.0;
            const double dt = S->dt_snap / (double)ns;
            int sob = on_start > fb ? on_start : fb, soe = on_end < fe ? on_end : fe, windowed;
            double tacc_d = c->prm.t_acc_direct_gyr;
            if (steps <= 0) continue;
            if (soe < sob) soe = sob;
            windowed = (soe - sob) < steps;
            {
                const double tmid = S->elapsed + ((double)fb + 0.5 * (double)steps) * dt;
                const CkParams *p = &c->prm;
                const int direct_on = p->dpp_model == 4 && isfinite(p->t_acc_direct_gyr) &&
                                      p->t_acc_direct_gyr > 0.0 && isfinite(p->t_off_gyr) && p->t_off_gyr > 0.0 &&
                                      isfinite(p->t_acc_off_gyr) && p->t_acc_off_gyr > 0.0;
                if (direct_on && tmid < p->t_off_gyr) tacc_d = p->t_acc_off_gyr;
            }
            if (ck_plain_lerp(n, count, didx, S, alpha, c->pws, c->stream) ||
                ck_plain_coeff(n, count, c->dgrid, c->prm, dt, tacc_d, b->protons, c->pws, c->d_err, c->stream) ||
                ck_plain_secondary(n, count, b->protons ? c->d_fqe : 0, c->jp0, c->pws, c->stream) ||
                ck_plain_thomas(n, count, steps, dt, windowed, sob - fb, sob - fb + (soe - sob), b->protons, c->pws,
                                c->d_err, c->stream)) {
                free(key);
                return dev_fail("plain segment");
            }
        }
        if (ck_plain_scatter(n, count, didx, S, c->pws, c->stream)) {
            free(key);
            return dev_fail("plain scatter");
        }
        i0 = i1;
    }
    free(key);
    return 0;
}

int croma_hip_batch_sync(CromaHipBatch *b)
{
    int err = 0;
    if (!b) return CROMA_HIP_EINVAL;
    if (ck_d2h(&err, b->ctx->d_err, sizeof(int), b->ctx->stream) || ck_stream_sync(b->ctx->stream))
        return dev_fail("batch sync");
    flush_retired(b->ctx);
    if (err) {
        ck_fill_i32(b->ctx->d_err, 0, 1, b->ctx->stream);
        ck_stream_sync(b->ctx->stream);
        if (err & 1) {
            ch_set_error("Thomas pivot below 1e-300");
            return CROMA_HIP_EPIVOT;
        }
        if (err & 4) {
            ch_set_error("IC cooling row: CMB bounds exceed the device quadrature size");
            return CROMA_HIP_EUNSUPPORTED;
        }
        ch_set_error("non-finite spectrum");
        return CROMA_HIP_ENONFINITE;
    }
    return 0;
}

int croma_hip_batch_run(CromaHipBatch *b, int s0, int s1, int nsnap_total, const double *z, const double *dt_snap,
                        double elapsed, int path, CromaHipTimes *times)
{
    CromaHipContext *c;
    PrepStats ps;
    GraphStats gst;
    double t_start, el = elapsed;
    int s, rc = 0, k = 0, wsz, nwin, iw, split, dev_nsub;
    float dev_ms = 0.0f;
    double kern_ms = 0.0;
    if (!b || !z || !dt_snap || s0 < 0 || s1 < s0 || s1 > b->nsnap_res || nsnap_total < s1 ||
        !(path == CROMA_HIP_PATH_FUSED || path == CROMA_HIP_PATH_PLAIN))
        return CROMA_HIP_EINVAL;
    if (s1 == s0) return 0;
    c = b->ctx;
    wsz = (path == CROMA_HIP_PATH_FUSED) ? c->tile[b->protons ? 1 : 0].chain : 1;
    if (wsz > s1 - s0) wsz = s1 - s0;
    /* the split launches need the species decoupled within a segment: no eta cap */
    split = path == CROMA_HIP_PATH_FUSED && b->protons && c->tile[1].secondary_gemm &&
            !(c->prm.dpp_model != 3 && c->prm.eta_dpp_cap > 0.0);
    /* the plain path keeps the reference's host-side buckets */
    dev_nsub = path == CROMA_HIP_PATH_FUSED && c->nsub_device;
    if (split) {
        const long segmax = c->prm.coeff_interp_max_segments > 1 ? c->prm.coeff_interp_max_segments : 1;
        long need = (long)wsz * b->ntr * segmax, q = 256;
        while (q < need / 64) q <<= 1;
        need = (need + q - 1) / q * q;
        if (need > b->segrows_cap) {
            const size_t cnt = (size_t)need * c->g.n;
            ck_stream_sync(c->stream);
            ck_free(b->d_segrec);
            ck_free(b->d_seginje);
            b->d_segrec = (double *)ck_malloc(cnt * sizeof(double));
            b->d_seginje = (double *)ck_malloc(cnt * sizeof(double));
            b->segrows_cap = (b->d_segrec && b->d_seginje) ? need : 0;
            if (!b->segrows_cap) return dev_fail("segment rows");
            /* rows past a window's segment count enter the rounded GEMM: keep them finite */
            if (ck_fill_f64(b->d_segrec, 0.0, cnt, c->stream) || ck_fill_f64(b->d_seginje, 0.0, cnt, c->stream))
                return dev_fail("segment rows");
        }
        b->seg_quantum = q;
    }
    nwin = (s1 - s0 + wsz - 1) / wsz;
    for (k = 0; k < 2; k++)
        if ((rc = win_reserve(c, &b->win[k], wsz, b->ntr)) != 0) return rc;
    if (path == CROMA_HIP_PATH_FUSED && b->ev_k_cap < 2 * nwin) {
        int need = 2 * nwin, j;
        void **ne = (void **)realloc(b->ev_k, (size_t)need * sizeof(void *));
        if (!ne) return CROMA_HIP_ENOMEM;
        b->ev_k = ne;
        for (j = b->ev_k_cap; j < need; j++) {
            b->ev_k[j] = ck_event_create();
            if (!b->ev_k[j]) {
                b->ev_k_cap = j;
                return dev_fail("event");
            }
        }
        b->ev_k_cap = need;
    }
    memset(&ps, 0, sizeof(ps));
    memset(&gst, 0, sizeof(gst));
    t_start = ch_now_s();
    if (ck_event_record(b->ev_a, c->stream)) return dev_fail("event");
    if (dev_nsub && ck_fill_i32((int *)b->d_cells, 0, 2, c->stream)) return dev_fail("cell counter");
    if ((rc = run_ic_rows(b, s0, s1, nsnap_total, z)) != 0) return rc;
    rc = prepare_window(b, &b->win[0], s0, (s1 - s0 < wsz) ? s1 - s0 : wsz, nsnap_total, z, dt_snap, el, split,
                        dev_nsub, &ps);
    if (rc) return rc;
    for (iw = 0, s = s0; s < s1; iw++) {
        WinBuf *w = &b->win[iw & 1];
        const int cnt = w->count;
        if ((rc = upload_window(b, w)) != 0) return rc;
        if (path == CROMA_HIP_PATH_FUSED) {
            if (ck_event_record(b->ev_k[2 * iw], c->stream)) return dev_fail("event");
            rc = launch_window(b, w, iw ? &b->win[(iw + 1) & 1] : 0, s + cnt >= s1, &gst);
            if (!rc && ck_event_record(b->ev_k[2 * iw + 1], c->stream)) return dev_fail("event");
        } else {
            rc = run_plain_snapshot(b, w);
        }
        if (rc) return rc;
        for (k = 0; k < cnt; k++) el += dt_snap[s + k];
        s += cnt;
        if (s < s1) {
            /* the other window's host buffers were consumed by its upload */
            WinBuf *o = &b->win[(iw + 1) & 1];
            if (ck_event_sync(o->ev_up)) return dev_fail("upload event");
            rc = prepare_window(b, o, s, (s1 - s < wsz) ? s1 - s : wsz, nsnap_total, z, dt_snap, el, split, dev_nsub,
                                &ps);
            if (rc) return rc;
        }
    }
    if (ck_event_record(b->ev_b, c->stream)) return dev_fail("event");
    rc = croma_hip_batch_sync(b);
    if (!rc && dev_nsub) {
        unsigned long long cells = 0;
        if (ck_d2h(&cells, b->d_cells, sizeof(cells), c->stream) || ck_stream_sync(c->stream))
            return dev_fail("cell counter");
        ps.cell_steps = (long long)cells;
        ps.species_steps = (long long)cells * (b->protons ? 2 : 1);
    }
    dev_ms = ck_event_elapsed_ms(b->ev_a, b->ev_b);
    if (path == CROMA_HIP_PATH_FUSED)
        for (k = 0; k < nwin; k++) kern_ms += ck_event_elapsed_ms(b->ev_k[2 * k], b->ev_k[2 * k + 1]);
    if (times) {
        times->kernel_ms += kern_ms;
        if (path == CROMA_HIP_PATH_FUSED) times->kernel_launches += nwin;
        times->ic_rows_ms += ps.ic_ms;
        times->nsub_estimate_ms += ps.nsub_ms;
        times->bucket_build_ms += ps.order_ms;
        times->solve_ms += dev_ms;
        times->total_ms += 1e3 * (ch_now_s() - t_start);
        times->cell_steps += ps.cell_steps;
        times->species_steps += ps.species_steps;
        times->graph_launches += gst.launches;
        times->graph_captures += gst.captures;
        times->graph_failures += gst.failures;
    }
    return rc;
}

typedef struct {
    CromaHipContext *c;
    int protons, ntr;
    const CkSlot *d_slot;
    const CkTask *d_tasks;
} XIssue;

static int issue_tasks(void *arg)
{
    XIssue *x = (XIssue *)arg;
    return launch_tasks(x->c, x->protons, 0, x->d_slot, x->d_tasks, x->ntr);
}

/* ------------------------------------------------------------------ */
/* Snapshot-level drop-in for tracer_fp_batch_execute_parallel          */
int croma_hip_execute_snapshot(
    CromaHipContext *c, int ntr, const int *nsubsteps, const int *n_onsteps, const int *base_nsteps,
    const int *fp_cadence_steps, int integration_mode, int allow, double elapsed, double dt_snap, double z_curr,
    double z_next, double t_acc_direct, const double *n_gas, const double *kbt, const double *bf,
    const double *divv, const double *lturb, const double *dv, const double *cs, const double *beta,
    const double *rad_ic_zero, const double *rad_ic_m1_zero, const double *rad_ic_p1_zero, const double *n_gas_n,
    const double *kbt_n, const double *bf_n, const double *divv_n, const double *lturb_n, const double *dv_n,
    const double *cs_n, const double *beta_n, const double *row_next, double m1n, double p1n, const double *qpi,
    const double *qepri, const double *fqe_flat, const int *np_min_qe, const double *mass,
    const unsigned char *dis, double *crp, double *cre, CromaHipTimes *times, int *bucket_calls_out,
    int *min_bucket_out, int *max_bucket_out)
{
    CromaHipBatch *b = 0;
    const double *cur[8] = {n_gas, kbt, bf, divv, lturb, dv, cs, beta};
    const double *nxt[8] = {n_gas_n, kbt_n, bf_n, divv_n, lturb_n, dv_n, cs_n, beta_n};
    int n, i, j, rc = 0, protons = 0, minb = 0, maxb = 0;
    double t0 = ch_now_s();
    size_t m;
    CkSlot *S;
    void *d_ws = 0;
    int *d_nsub = 0, *d_non = 0;
    CkTask *d_tasks = 0;
    CkSlot *d_slot = 0;
    double *d_ric = 0, *d_fqe = 0;
    CkParams prm;
    double ric2[2 * CH_NMAX]; /* staging of the caller's rows; read by the async upload before the final sync */
    (void)z_curr;
    (void)z_next;
    if (!c || ntr <= 0 || !nsubsteps || !n_onsteps || !base_nsteps || !fp_cadence_steps || !rad_ic_zero ||
        !rad_ic_m1_zero || !rad_ic_p1_zero || !row_next || !qpi || !qepri || !fqe_flat || !np_min_qe || !mass ||
        !crp || !cre)
        return CROMA_HIP_EINVAL;
    for (i = 0; i < 8; i++)
        if (!cur[i] || !nxt[i]) return CROMA_HIP_EINVAL;
    if (integration_mode != 0) {
        ch_set_error("only the quantized integration mode exists");
        return CROMA_HIP_EUNSUPPORTED;
    }
    n = c->g.n;
    m = (size_t)ntr * n;
    for (i = 0; i < ntr; i++) {
        if (base_nsteps[i] != nsubsteps[i] || fp_cadence_steps[i] != 1) {
            ch_set_error("base_nsteps must equal nsubsteps and cadence must be 1");
            return CROMA_HIP_EUNSUPPORTED;
        }
        if (nsubsteps[i] <= 0) return CROMA_HIP_EINVAL;
        if (rad_ic_m1_zero[i] != rad_ic_m1_zero[0] || rad_ic_p1_zero[i] != rad_ic_p1_zero[0]) {
            ch_set_error("rad_ic edges must be the broadcast snapshot row");
            return CROMA_HIP_EUNSUPPORTED;
        }
    }
    for (j = 0; j < n; j++)
        for (i = 1; i < ntr; i++)
            if (rad_ic_zero[(size_t)j * ntr + i] != rad_ic_zero[(size_t)j * ntr]) {
                ch_set_error("rad_ic_zero must be the broadcast snapshot row");
                return CROMA_HIP_EUNSUPPORTED;
            }
    for (i = 0; i < (int)m && !protons; i++)
        if (crp[i] != 0.0 || qpi[i] != 0.0) protons = 1;

    /* the batch and the workspace persist between calls of equal shape, as
     * the reference's grow-only pipeline workspace does */
    if (c->xb && (c->xb->ntr != ntr || c->xb->protons != protons)) {
        croma_hip_batch_destroy(c->xb);
        c->xb = 0;
    }
    if (!c->xb &&
        !(c->xb = croma_hip_batch_create(c, ntr, protons ? CROMA_HIP_SPECIES_ELECTRON_PROTON : CROMA_HIP_SPECIES_ELECTRON,
                                         2)))
        return CROMA_HIP_ENOMEM;
    b = c->xb;
    rc = croma_hip_batch_set_state(b, protons ? crp : 0, cre, protons ? qpi : 0, qepri, mass, dis);
    if (!rc) rc = croma_hip_batch_set_background(b, 0, cur);
    if (!rc) rc = croma_hip_batch_set_background(b, 1, nxt);
    if (rc) goto done;

    /* caller-provided schedule, IC rows and secondary weights */
    {
        const size_t need = (size_t)ntr * (2 * sizeof(int) + sizeof(CkTask)) + sizeof(CkSlot) +
                            2 * (size_t)n * sizeof(double) + (size_t)n * n * sizeof(double) + 8192;
        if (need > c->x_ws_bytes) {
            ck_stream_sync(c->stream);
            ck_graph_destroy(c->x_graph.exec);
            memset(&c->x_graph, 0, sizeof(c->x_graph));
            ck_free(c->x_ws);
            c->x_ws = ck_malloc(need);
            c->x_ws_bytes = c->x_ws ? need : 0;
        }
        d_ws = c->x_ws;
    }
    if (!d_ws) {
        rc = dev_fail("execute workspace");
        goto done;
    }
    {
        char *p = (char *)d_ws;
        d_nsub = (int *)p;
        p += (((size_t)ntr * sizeof(int)) + 255) & ~(size_t)255;
        d_non = (int *)p;
        p += (((size_t)ntr * sizeof(int)) + 255) & ~(size_t)255;
        d_tasks = (CkTask *)p;
        p += (((size_t)ntr * sizeof(CkTask)) + 255) & ~(size_t)255;
        d_slot = (CkSlot *)p;
        p += (sizeof(CkSlot) + 255) & ~(size_t)255;
        d_ric = (double *)p;
        p += ((2 * (size_t)n * sizeof(double)) + 255) & ~(size_t)255;
        d_fqe = (double *)p;
    }
    {
        WinBuf *w = &b->win[0];
        int allow1 = allow;
        int jp0 = n;
        for (j = 0; j < n; j++) ric2[j] = rad_ic_zero[(size_t)j * ntr];
        memcpy(ric2 + n, row_next, (size_t)n * sizeof(double));
        for (j = 0; j < n; j++) {
            int v = np_min_qe[j] < 0 ? 0 : np_min_qe[j];
            if (v < jp0) jp0 = v;
        }
        memcpy(w->h_nsub, nsubsteps, (size_t)ntr * sizeof(int));
        memcpy(w->h_non, n_onsteps, (size_t)ntr * sizeof(int));
        ch_params_to_ck(&c->uprm, &prm);
        prm.t_acc_direct_gyr = t_acc_direct;
        order_tasks(&prm, &allow1, 1, ntr, nsubsteps, protons, w->h_tasks, b->hist, b->keys);
        S = w->h_slot;
        memset(S, 0, sizeof(*S));
        S->crp = b->d_crp;
        S->cre = b->d_cre;
        S->qpi = protons ? b->d_qpi : 0;
        S->qepri = b->d_qepri;
        S->mass = b->d_mass;
        S->dis = b->has_dis ? b->d_dis : 0;
        for (i = 0; i < CK_NBG; i++) {
            S->bgc[i] = b->d_bg + (size_t)i * ntr;
            S->bgn[i] = b->d_bg + ((size_t)CK_NBG + i) * ntr;
        }
        S->ric_c = d_ric;
        S->ric_n = d_ric + n;
        S->nsub = d_nsub;
        S->n_on = d_non;
        S->fqe_t = protons ? d_fqe : 0;
        S->jp0 = jp0;
        S->ntr = ntr;
        S->allow_interp = allow;
        S->protons = protons;
        S->dt_snap = dt_snap;
        S->elapsed = elapsed;
        S->prm = prm;
        if (ck_h2d(d_nsub, w->h_nsub, (size_t)ntr * sizeof(int), c->stream) ||
            ck_h2d(d_non, w->h_non, (size_t)ntr * sizeof(int), c->stream) ||
            ck_h2d(d_tasks, w->h_tasks, (size_t)ntr * sizeof(CkTask), c->stream) ||
            ck_h2d(d_slot, S, sizeof(CkSlot), c->stream) ||
            ck_h2d(d_ric, ric2, 2 * (size_t)n * sizeof(double), c->stream) ||
            (protons && ck_h2d(d_fqe, fqe_flat, (size_t)n * n * sizeof(double), c->stream))) {
            rc = dev_fail("execute upload");
            goto done;
        }
        {
            const CromaHipTile *t = &c->tile[protons ? 1 : 0];
            XIssue xi;
            GraphStats gst;
            long key[CH_GKEY];
            memset(key, 0, sizeof(key));
            memset(&gst, 0, sizeof(gst));
            key[0] = protons;
            key[1] = ntr;
            key[2] = t->rows_per_lane;
            key[3] = t->waves_per_simd;
            key[4] = (long)(size_t)d_ws;
            xi.c = c;
            xi.protons = protons;
            xi.d_slot = d_slot;
            xi.d_tasks = d_tasks;
            xi.ntr = ntr;
            if ((rc = graph_issue(c, c->stream, &c->x_graph, key, issue_tasks, &xi, &gst)) != 0) goto done;
            if (times) {
                times->graph_launches += gst.launches;
                times->graph_captures += gst.captures;
                times->graph_failures += gst.failures;
            }
        }
    }
    rc = croma_hip_batch_sync(b);
    if (!rc) rc = croma_hip_batch_get_state(b, protons ? crp : 0, cre);
    if (!rc && !protons) memset(crp, 0, m * sizeof(double));
    for (i = 0; i < ntr; i++) {
        if (minb == 0 || nsubsteps[i] < minb) minb = nsubsteps[i];
        if (nsubsteps[i] > maxb) maxb = nsubsteps[i];
        if (times) times->cell_steps += nsubsteps[i];
    }
    if (bucket_calls_out) *bucket_calls_out = 1;
    if (min_bucket_out) *min_bucket_out = ntr;
    if (max_bucket_out) *max_bucket_out = ntr;
    if (times) times->total_ms += 1e3 * (ch_now_s() - t0);
done:
    ck_stream_sync(c->stream);
    flush_retired(c);
    return rc;
}

/* ------------------------------------------------------------------ */
int croma_hip_solve_cc_batch(int nsys, int nrow, double dt, const double *a, const double *bb, const double *cc,
                             const double *src, const double *st_in, double *st_out, double *elapsed_ms)
{
    const size_t m = (size_t)nsys * nrow;
    void *s = 0, *e0 = 0, *e1 = 0;
    double *d = 0;
    int *derr = 0, err = 0, rc = 0;
    if (nsys <= 0 || nrow <= 0 || !a || !bb || !cc || !src || !st_in || !st_out) return CROMA_HIP_EINVAL;
    s = ck_stream_create();
    e0 = ck_event_create();
    e1 = ck_event_create();
    d = (double *)ck_malloc(m * 7 * sizeof(double));
    derr = (int *)ck_malloc(sizeof(int));
    if (!s || !e0 || !e1 || !d || !derr) {
        rc = dev_fail("solve alloc");
        goto out;
    }
    if (ck_fill_i32(derr, 0, 1, s) || ck_h2d(d, a, m * 8, s) || ck_h2d(d + m, bb, m * 8, s) ||
        ck_h2d(d + 2 * m, cc, m * 8, s) || ck_h2d(d + 3 * m, src, m * 8, s) || ck_h2d(d + 4 * m, st_in, m * 8, s) ||
        ck_event_record(e0, s) ||
        ck_solve_cc_batch(nsys, nrow, dt, d, d + m, d + 2 * m, d + 3 * m, d + 4 * m, d + 5 * m, d + 6 * m, derr, s) ||
        ck_event_record(e1, s) || ck_d2h(st_out, d + 5 * m, m * 8, s) || ck_d2h(&err, derr, sizeof(int), s) ||
        ck_stream_sync(s)) {
        rc = dev_fail("solve");
        goto out;
    }
    if (elapsed_ms) *elapsed_ms = ck_event_elapsed_ms(e0, e1);
    if (err) rc = CROMA_HIP_EPIVOT;
out:
    ck_free(d);
    ck_free(derr);
    ck_event_destroy(e0);
    ck_event_destroy(e1);
    ck_stream_destroy(s);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Autotuner: times every launch shape of the fused kernel on a seeded
 * synthetic probe (backgrounds drawn like the benchmark workloads) and keeps
 * the fastest.  Ties within 2% go to the larger rows_per_lane (fewer lanes
 * per tracer leave more tracers per wave for the tail). */
static double at_uniform(unsigned long long *s)
{
    /* splitmix64 */
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static double at_normal(unsigned long long *s)
{
    const double u1 = at_uniform(s), u2 = at_uniform(s);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

/* Whether every fused launch mode a tile runs has an FP32-coefficient
 * instance (electrons: mode 0; two species: the split passes 2 and 3, or the
 * in-kernel contraction 1, which the eta cap also selects). */
static int tile_f32_ok(const CromaHipContext *c, int protons, int r, int w, int g)
{
    const int n = c->g.n;
    if (!protons) return ck_fused_f32_supported(n, r, w, 0);
    if (!g || c->prm.eta_dpp_cap > 0.0) return ck_fused_f32_supported(n, r, w, 1);
    return ck_fused_f32_supported(n, r, w, 2) && ck_fused_f32_supported(n, r, w, 3);
}

/* Candidate tiles of the autotuners: rows per lane x waves per SIMD (x the
 * split secondary contraction for two species), one snapshot per task. */
static int tile_candidates(const CromaHipContext *c, int protons, CromaHipTile *out)
{
    static const int rows[4] = {2, 4, 8, 16};
    const int n = c->g.n;
    int r, w, g, k = 0;
    for (g = protons ? 1 : 0; g >= 0; g--)
        for (r = 0; r < 4; r++) {
            if (!ck_fused_supported(n, rows[r])) continue;
            for (w = 1; w <= 2; w++) {
                if (c->coeff_fp32 && !tile_f32_ok(c, protons, rows[r], w, g)) continue;
                out[k].rows_per_lane = rows[r];
                out[k].waves_per_simd = w;
                out[k].chain = 1;
                out[k].secondary_gemm = g;
                k++;
            }
        }
    return k;
}

/* Ties within 2% go to more rows per lane (fewer cross-lane levels). */
static int tile_better(double ms, const CromaHipTile *t, double best_ms, const CromaHipTile *bt)
{
    return ms < best_ms * 0.98 || (ms < best_ms * 1.02 && t->rows_per_lane > bt->rows_per_lane);
}

int croma_hip_batch_autotune(CromaHipBatch *b, int s, int nsnap_total, const double *z, const double *dt_snap,
                             double elapsed, CromaHipTile *best)
{
    CromaHipContext *c;
    CromaHipTile cand[16], keep, bt;
    double best_ms = 1e300, *d_save = 0;
    size_t m;
    int k, nc, rep, rc = 0;
    if (!b || !z || !dt_snap || !best || s < 0 || s + 1 > b->nsnap_res || nsnap_total < s + 1)
        return CROMA_HIP_EINVAL;
    c = b->ctx;
    m = (size_t)b->ntr * c->g.n;
    keep = c->tile[b->protons ? 1 : 0];
    bt = keep;
    /* every trial starts from the batch's current state */
    d_save = (double *)ck_malloc(m * sizeof(double) * (b->protons ? 2 : 1));
    if (!d_save) return dev_fail("autotune state copy");
    if (ck_d2d(d_save, b->d_cre, m * sizeof(double), c->stream) ||
        (b->protons && ck_d2d(d_save + m, b->d_crp, m * sizeof(double), c->stream))) {
        rc = dev_fail("autotune state copy");
        goto out;
    }
    nc = tile_candidates(c, b->protons, cand);
    for (k = 0; k < nc; k++) {
        double ms = 0.0;
        c->tile[b->protons ? 1 : 0] = cand[k];
        /* the first run of a shape also loads its code and captures nothing
         * yet: a short snapshot is timed on its second run */
        for (rep = 0; rep < 2; rep++) {
            CromaHipTimes tm;
            memset(&tm, 0, sizeof(tm));
            if (ck_d2d(b->d_cre, d_save, m * sizeof(double), c->stream) ||
                (b->protons && ck_d2d(b->d_crp, d_save + m, m * sizeof(double), c->stream))) {
                rc = dev_fail("autotune state restore");
                goto out;
            }
            if ((rc = croma_hip_batch_run(b, s, s + 1, nsnap_total, z, dt_snap, elapsed, CROMA_HIP_PATH_FUSED, &tm)) !=
                0)
                goto out;
            ms = tm.kernel_ms;
            if (ms > 50.0) break;
        }
        if (tile_better(ms, &cand[k], best_ms, &bt)) {
            if (ms < best_ms) best_ms = ms;
            bt = cand[k];
        }
    }
    /* snapshots per task: one launch per snapshot against one task over two
     * snapshots, timed on snapshots s and s + 1 (the second needs its
     * successor resident unless it is the last) */
    if (s + 2 <= nsnap_total && s + 2 <= b->nsnap_res && (s + 2 == nsnap_total || s + 2 < b->nsnap_res)) {
        double cms[2] = {0.0, 0.0};
        int ch;
        for (ch = 0; ch < 2; ch++) {
            c->tile[b->protons ? 1 : 0] = bt;
            c->tile[b->protons ? 1 : 0].chain = ch + 1;
            for (rep = 0; rep < 2; rep++) {
                CromaHipTimes tm;
                memset(&tm, 0, sizeof(tm));
                if (ck_d2d(b->d_cre, d_save, m * sizeof(double), c->stream) ||
                    (b->protons && ck_d2d(b->d_crp, d_save + m, m * sizeof(double), c->stream))) {
                    rc = dev_fail("autotune state restore");
                    goto out;
                }
                if ((rc = croma_hip_batch_run(b, s, s + 2, nsnap_total, z, dt_snap, elapsed, CROMA_HIP_PATH_FUSED,
                                              &tm)) != 0)
                    goto out;
                cms[ch] = tm.kernel_ms;
                if (cms[ch] > 100.0) break;
            }
        }
        bt.chain = (cms[1] < 0.99 * cms[0]) ? 2 : 1;
    }
out:
    if (d_save) {
        if (ck_d2d(b->d_cre, d_save, m * sizeof(double), c->stream) ||
            (b->protons && ck_d2d(b->d_crp, d_save + m, m * sizeof(double), c->stream)) || ck_stream_sync(c->stream))
            if (!rc) rc = dev_fail("autotune state restore");
        ck_free(d_save);
    }
    c->tile[b->protons ? 1 : 0] = rc ? keep : bt;
    if (!rc) *best = bt;
    return rc;
}

int croma_hip_autotune(CromaHipContext *c, int protons, int ntracer, CromaHipTile *best)
{
    const int n = c ? c->g.n : 0;
    const int nsnap = 3;
    double *bg = 0, *cre = 0, *crp = 0, z[4], dt[3];
    double best_ms = 1e300;
    CromaHipTile keep, bt, cand[16];
    CromaHipBatch *b = 0;
    unsigned long long seed = 0x5EEDULL + (unsigned long long)n;
    int i, s, f, k, nc, rc = 0;
    if (!c || ntracer <= 0 || !best) return CROMA_HIP_EINVAL;
    keep = c->tile[protons ? 1 : 0];
    bt = keep;
    bg = (double *)malloc((size_t)(nsnap + 1) * CK_NBG * ntracer * sizeof(double));
    cre = (double *)malloc((size_t)ntracer * n * sizeof(double));
    crp = (double *)malloc((size_t)ntracer * n * sizeof(double));
    if (!bg || !cre || !crp) {
        rc = CROMA_HIP_ENOMEM;
        goto out;
    }
    for (i = 0; i < ntracer; i++) {
        const double ln = -5.0 + 3.0 * at_uniform(&seed), lt = 6.0 + 2.0 * at_uniform(&seed);
        const double lb = -2.0 + 3.0 * at_uniform(&seed), curl = 3.0e-17 * exp(at_normal(&seed));
        const double amp = pow(10.0, at_uniform(&seed) - 0.5);
        int j;
        for (s = 0; s <= nsnap; s++) {
            const double drift = 0.02 * s;
            const double ngas = pow(10.0, ln + drift), temp = pow(10.0, lt), bf = pow(10.0, lb - drift) * 1.0e-6;
            const double cs = sqrt(5.0 / 3.0 * 8.62e-5 * temp / (0.59 * 0.938 * 1.0e9)) * 2.999e10;
            double dv = curl * 0.15 * 3.0857e24;
            double *o = bg + (size_t)s * CK_NBG * ntracer;
            if (dv / cs > 0.5) dv = 0.5 * cs;
            o[CK_BG_NGAS * (size_t)ntracer + i] = ngas;
            o[CK_BG_KBT * (size_t)ntracer + i] = 8.62e-5 * temp * 1.0e-9;
            o[CK_BG_B * (size_t)ntracer + i] = bf;
            o[CK_BG_DIVV * (size_t)ntracer + i] = 1.0e-17 * at_normal(&seed) * 3.1536e16;
            o[CK_BG_LTURB * (size_t)ntracer + i] = 0.15;
            o[CK_BG_DV * (size_t)ntracer + i] = dv;
            o[CK_BG_CS * (size_t)ntracer + i] = cs;
            o[CK_BG_BETA * (size_t)ntracer + i] = ngas * 8.62e-5 * temp * 1.6e-12 / (bf * bf / (8.0 * 3.14159265358979323846));
        }
        for (j = 0; j < n; j++) {
            cre[(size_t)i * n + j] = amp * pow(c->g.ep[j], -2.2) * exp(-c->g.ep[j] / 1.0e4);
            crp[(size_t)i * n + j] = 100.0 * amp * pow(c->g.pp[j], -2.2) * exp(-c->g.pp[j] / 1.0e6);
        }
    }
    for (s = 0; s <= nsnap; s++) z[s] = 0.5 - 0.01 * s;
    for (s = 0; s < nsnap; s++) dt[s] = 0.077;
    b = croma_hip_batch_create(c, ntracer, protons ? CROMA_HIP_SPECIES_ELECTRON_PROTON : CROMA_HIP_SPECIES_ELECTRON,
                               nsnap + 1);
    if (!b) {
        rc = CROMA_HIP_ENOMEM;
        goto out;
    }
    for (s = 0; s <= nsnap; s++) {
        const double *fl[CK_NBG];
        for (f = 0; f < CK_NBG; f++) fl[f] = bg + ((size_t)s * CK_NBG + f) * ntracer;
        if ((rc = croma_hip_batch_set_background(b, s, fl)) != 0) goto out;
    }
    nc = tile_candidates(c, protons, cand);
    for (k = 0; k < nc; k++) {
        CromaHipTimes tm;
        c->tile[protons ? 1 : 0] = cand[k];
        memset(&tm, 0, sizeof(tm));
        if ((rc = croma_hip_batch_set_state(b, protons ? crp : 0, cre, 0, 0, 0, 0)) != 0) goto out;
        /* first snapshot warms the instance, the next two are timed */
        if ((rc = croma_hip_batch_run(b, 0, 1, nsnap + 1, z, dt, 0.0, CROMA_HIP_PATH_FUSED, 0)) != 0) goto out;
        if ((rc = croma_hip_batch_run(b, 1, nsnap, nsnap + 1, z, dt, dt[0], CROMA_HIP_PATH_FUSED, &tm)) != 0) goto out;
        if (tile_better(tm.kernel_ms, &cand[k], best_ms, &bt)) {
            if (tm.kernel_ms < best_ms) best_ms = tm.kernel_ms;
            bt = cand[k];
        }
    }
out:
    croma_hip_batch_destroy(b);
    free(bg);
    free(cre);
    free(crp);
    c->tile[protons ? 1 : 0] = rc ? keep : bt;
    if (!rc) *best = bt;
    return rc;
}
