/*
 * croma_physics.c -- host-side physics tables for the HIP backend:
 * momentum grids and per-grid caches, the IC cooling row, the nsub
 * estimator, coefficient-segment choice, and the hadronic / IC / synchrotron
 * emission kernel tables.  These run on the host once per grid or once per
 * snapshot (they are identical for every tracer); the per-tracer work runs
 * on the GPU.
 *
 * The arithmetic follows the reference CPU path in the same evaluation order
 * (FP64, no FMA contraction; build with -ffp-contract=off) so that integer
 * outputs such as nsub agree exactly with the reference:
 *   momentum_bin            FP_Coef.c:291-330
 *   grid caches             FP_Coef.c:77-154
 *   pp cross section        HADRONIC.c:342-431
 *   IC cooling              FP_Coef.c:361-539
 *   nsub                    tracer_fp_nsub.c:37-282
 *   segments                tracer_fp_setup.c:73-95
 *   secondary weights       fp_shared_core.c:1453-1506, HADRONIC.c:433-533
 *   gamma / nu / IC kernels fp_shared_core.c:1724-1807, HADRONIC.c:139-289,
 *                           FP_Coef.c:553-582
 */
#include "croma_host.h"
#include "croma_hadronic.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define C_LIGHT 2.999e+10
#define GEV_ERG 1.6e-3
#define GYR_S 3.1536e+16
#define H_PLANCK 6.63e-27
#define CH_MP 0.938
#define CH_ME 0.511e-3
#define MB_CM2 1.0e-27
#define CH_MPI0 1.35e-1
#define CH_MPIC 1.40e-1
#define CH_MMU 1.05e-1

/* ------------------------------------------------------------------ */
/* tiny pthread parallel-for (keeps OpenMP out of the shared library)  */
typedef struct {
    void (*fn)(void *, long, long);
    void *arg;
    long begin, end;
} ChPfJob;

static void *ch_pf_thread(void *p)
{
    ChPfJob *j = (ChPfJob *)p;
    j->fn(j->arg, j->begin, j->end);
    return 0;
}

int ch_host_threads(void)
{
    static int cached = 0;
    if (cached == 0) {
        const char *env = getenv("CROMA_HIP_HOST_THREADS");
        long v = env ? strtol(env, 0, 10) : 0;
        if (v <= 0) {
            v = sysconf(_SC_NPROCESSORS_ONLN);
            if (v <= 0) v = 8;
            if (v > 16) v = 16;
        }
        if (v > 64) v = 64;
        cached = (int)v;
    }
    return cached;
}

void ch_parallel_for(long n, long min_chunk, void (*fn)(void *, long, long), void *arg)
{
    int nthr = ch_host_threads(), t;
    pthread_t th[64];
    int started[64];
    ChPfJob jobs[64];
    long chunk;
    if (n <= 0) return;
    if (min_chunk < 1) min_chunk = 1;
    if ((long)nthr * min_chunk > n) nthr = (int)((n + min_chunk - 1) / min_chunk);
    if (nthr <= 1) {
        fn(arg, 0, n);
        return;
    }
    chunk = (n + nthr - 1) / nthr;
    for (t = 0; t < nthr; t++) {
        jobs[t].fn = fn;
        jobs[t].arg = arg;
        jobs[t].begin = t * chunk < n ? t * chunk : n;
        jobs[t].end = (t + 1) * chunk < n ? (t + 1) * chunk : n;
        started[t] = 0;
    }
    for (t = 1; t < nthr; t++) {
        if (jobs[t].begin >= jobs[t].end) continue;
        if (pthread_create(&th[t], 0, ch_pf_thread, &jobs[t]) == 0) started[t] = 1;
        else fn(arg, jobs[t].begin, jobs[t].end);
    }
    fn(arg, jobs[0].begin, jobs[0].end);
    for (t = 1; t < nthr; t++)
        if (started[t]) pthread_join(th[t], 0);
}

/* ------------------------------------------------------------------ */
/* pp cross section (Kamae 2006 fit, Blattnig 2000 charged ratios)     */
double ch_pp_pion_cross_section(int kind, double p) { return chk_pp_pion_cross_section(kind, p); }
/* ------------------------------------------------------------------ */
static void ch_grid_axis(int n, double lo, double hi, double mass, double *p, double *dp, double *E,
                         double *pm1, double *pp1)
{
    const double step = (hi - lo) / (double)n;
    double a = lo - step;
    int j;
    *pm1 = pow(10.0, a);
    for (j = 0; j < n; j++) {
        a = a + step;
        p[j] = pow(10.0, a);
        E[j] = mass * sqrt(1.0 + p[j] * p[j]);
    }
    a = a + step;
    *pp1 = pow(10.0, a);
    for (j = 1; j < n - 1; j++) dp[j] = 0.5 * (p[j + 1] - p[j - 1]);
    dp[0] = 0.5 * (p[1] - *pm1);
    dp[n - 1] = 0.5 * (pow(10.0, a + step) - p[n - 2]);
}

int ch_grid_build(ChGrid *g, int n, double pmin, double pmax, double pemin, double pemax)
{
    int j, j_pp = 0;
    double p_pp;
    if (!g || n < 8 || n > CH_NMAX) return -1;
    memset(g, 0, sizeof(*g));
    g->n = n;
    g->pmin = pmin;
    g->pmax = pmax;
    g->pemin = pemin;
    g->pemax = pemax;
    ch_grid_axis(n, pmin, pmax, CH_MP, g->pp, g->pdp, g->pE, &g->ppm1, &g->ppp1);
    ch_grid_axis(n, pemin, pemax, CH_ME, g->ep, g->edp, g->eE, &g->epm1, &g->epp1);
    for (j = 0; j < n; j++) {
        g->beta_p[j] = g->pp[j] / sqrt(1.0 + g->pp[j] * g->pp[j]);
        g->p_p2[j] = g->pp[j] * g->pp[j];
        g->p_expcut[j] = exp(-1.0 / g->pp[j]);
        g->sigma_pp[j] = ch_pp_pion_cross_section(2, CH_MP * g->pp[j]);
        g->e_p2[j] = g->ep[j] * g->ep[j];
        g->e_expcut[j] = exp(-1.0 / g->ep[j]);
    }
    /* the threshold scan stops at n-100 and re-arms while j_pp == 0 */
    for (j = 0; j < n - 100; j++)
        if (j_pp == 0 && g->sigma_pp[j] > 0.0) j_pp = j;
    j_pp += 2;
    if (j_pp < 0) j_pp = 0;
    if (j_pp >= n) j_pp = n - 1;
    g->j_pp = j_pp;
    p_pp = g->pp[j_pp];
    for (j = 0; j < n; j++) {
        double sig = (j > j_pp) ? g->sigma_pp[j] : g->sigma_pp[j_pp];
        g->sigmoid[j] = 1.0 / (1.0 + exp(-5.0 * (g->pp[j] - p_pp)));
        sig *= g->sigmoid[j];
        g->p_sig[j] = sig;
        g->p_ecr_w[j] = 0.0;
        g->e_ecr_w[j] = 0.0;
        if (n > 4 && j >= 2 && j < n - 2) {
            const double kp = g->pE[j] - CH_MP, ke = g->eE[j] - CH_ME;
            if (kp > 0.0) g->p_ecr_w[j] = kp * GEV_ERG;
            if (ke > 0.0) g->e_ecr_w[j] = ke * GEV_ERG;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* loss-rate endpoints for nsub                                        */
static double ch_coul_kernel(double x) { return 0.5 * sqrt(M_PI) * erf(sqrt(x)) - sqrt(x) * exp(-x); }

double ch_b_coulomb_p(double nicm, double p, double kBT)
{
    const double p2 = p * p, gamma = sqrt(1.0 + p2), betap2 = p2 / (1.0 + p2);
    const double kinetic = gamma - 1.0, Ep = gamma * CH_MP - CH_MP;
    const double xp = Ep / (kBT), xe = xp * (CH_ME / CH_MP);
    const double sum = ch_coul_kernel(xe) + ch_coul_kernel(xp) * (CH_ME / CH_MP);
    const double thermal = 1.0 - kBT / (2.0 * CH_MP * kinetic);
    if (thermal < 0.0) return 0.0;
    return GYR_S * C_LIGHT / (CH_MP * GEV_ERG) * sum * nicm * (1.0 / betap2) * thermal * 3.5e-29;
}

static double ch_b_synch(double p, double B)
{
    const double b_micro = B * 1.0e+6, b_ratio = b_micro / 3.2;
    return 4.8e-4 * p * p * b_ratio * b_ratio * 9.11e-28 * 3.0e+10 * 3.1536e+16;
}

static double ch_b_coulomb_e(double nicm, double pe)
{
    const double pe2 = pe * pe, beta2 = pe2 / (1.0 + pe2), A = 3.05e-29;
    return A * nicm * (1.0 + log(pe / nicm) / 74.8) / (CH_ME * GEV_ERG) * C_LIGHT * GYR_S / beta2;
}

static double ch_quantize_dt(double t)
{
    static const double ladder[10] = {1.0, 1.25, 1.6, 2.0, 2.5, 3.2, 4.0, 5.0, 6.4, 8.0};
    double e10, scale, norm, chosen = ladder[0];
    int i;
    if (!(t > 0.0) || !isfinite(t)) return t;
    e10 = floor(log10(t));
    scale = pow(10.0, e10);
    norm = t / scale;
    for (i = 0; i < 10; i++) {
        if (ladder[i] <= norm) chosen = ladder[i];
        else break;
    }
    return chosen * scale;
}

static double ch_quantize_on(double f)
{
    static const double bins[5] = {1.0 / 8.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 2.0, 1.0};
    double chosen = bins[0], best;
    int i;
    if (!(f > 0.0) || !isfinite(f)) return 1.0;
    if (f >= 1.0) return 1.0;
    best = fabs(f - chosen);
    for (i = 1; i < 5; i++) {
        const double d = fabs(f - bins[i]);
        if (d < best) {
            chosen = bins[i];
            best = d;
        }
    }
    return chosen;
}

typedef struct {
    const ChGrid *g;
    const CkParams *prm;
    double dt_snap, ric_lo, ric_hi;
    const double *n_gas, *kbt, *b_field, *divv, *l_turb, *dv;
    int *nsub, *n_on;
} ChNsubJob;

static void ch_nsub_range(void *arg, long i0, long i1)
{
    const ChNsubJob *J = (const ChNsubJob *)arg;
    const ChGrid *g = J->g;
    const int n = g->n;
    long i;
    for (i = i0; i < i1; i++) {
        const double lo_p = ch_b_coulomb_p(J->n_gas[i], g->pp[0], J->kbt[i]) + 1.0 / 3.0 * J->divv[i] * g->pp[0];
        const double hi_p = ch_b_coulomb_p(J->n_gas[i], g->pp[n - 1], J->kbt[i]) + 1.0 / 3.0 * J->divv[i] * g->pp[n - 1];
        const double lo_e = ch_b_synch(g->ep[0], J->b_field[i]) + J->ric_lo + ch_b_coulomb_e(J->n_gas[i], g->ep[0]) +
                            1.0 / 3.0 * J->divv[i] * g->ep[0];
        const double hi_e = ch_b_synch(g->ep[n - 1], J->b_field[i]) + J->ric_hi +
                            ch_b_coulomb_e(J->n_gas[i], g->ep[n - 1]) + 1.0 / 3.0 * J->divv[i] * g->ep[n - 1];
        const double te0 = g->ep[0] / lo_e, tp0 = g->pp[0] / lo_p;
        const double teN = g->ep[n - 1] / hi_e, tpN = g->pp[n - 1] / hi_p;
        const double tmin = fmin(fabs(tp0), fmin(fabs(teN), fmin(fabs(te0), fabs(tpN))));
        double target = J->dt_snap / 10.0, used, raw, on_fraction = 1.0;
        int ns, non;
        if (isfinite(tmin) && tmin > 0.0) {
            const double st = J->prm->tracer_nsub_safety * tmin;
            if (st < target) target = st;
        }
        if (target <= 0.0 || !isfinite(target)) target = J->dt_snap / 10.0;
        used = ch_quantize_dt(target);
        if (!(used > 0.0) || !isfinite(used) || used > target) used = target;
        raw = ceil(J->dt_snap / used);
        if (!isfinite(raw) || raw > 1000000.0) raw = 1000000.0;
        if (raw < 10.0) raw = 10.0;
        ns = (int)raw;
        if (J->prm->tracer_nsub_max > 0 && ns > J->prm->tracer_nsub_max) ns = J->prm->tracer_nsub_max;
        if (J->prm->reacc_window_mode == 1 && J->prm->f_eddy > 0.0 && J->dv[i] > 0.0) {
            const double t_eddy = J->l_turb[i] * 3.0857e24 / J->dv[i] / 3.1536e16;
            if (J->prm->f_eddy * t_eddy < J->dt_snap)
                on_fraction = ch_quantize_on(J->prm->f_eddy * t_eddy / J->dt_snap);
        }
        non = (int)floor((double)ns * on_fraction + 0.5);
        if (non < 1) non = 1;
        if (non > ns) non = ns;
        J->nsub[i] = ns;
        J->n_on[i] = non;
    }
}

int ch_nsub_estimate(const ChGrid *g, const CkParams *prm, long ntr, double dt_snap, double ric_lo,
                     double ric_hi, const double *n_gas, const double *kbt, const double *b_field,
                     const double *divv, const double *l_turb, const double *dv, int *nsub, int *n_on)
{
    ChNsubJob J;
    if (!g || !prm || ntr < 0 || !(dt_snap > 0.0)) return -1;
    J.g = g;
    J.prm = prm;
    J.dt_snap = dt_snap;
    J.ric_lo = ric_lo;
    J.ric_hi = ric_hi;
    J.n_gas = n_gas;
    J.kbt = kbt;
    J.b_field = b_field;
    J.divv = divv;
    J.l_turb = l_turb;
    J.dv = dv;
    J.nsub = nsub;
    J.n_on = n_on;
    ch_parallel_for(ntr, 4096, ch_nsub_range, &J);
    return 0;
}

int ch_choose_segments(const CkParams *prm, int nsub, int allow)
{
    int segments = 1, max_allowed, maxseg = prm->coeff_interp_max_segments;
    if (!allow || nsub <= 1) return 1;
    if (prm->coeff_interp_min_steps <= 0) return 1;
    if (maxseg < 1) maxseg = 1;
    max_allowed = nsub / prm->coeff_interp_min_steps;
    if (max_allowed < 1) return 1;
    while (segments <= 1073741823 && segments * 2 <= maxseg && segments * 2 <= max_allowed) segments *= 2;
    if (segments > nsub) segments = nsub;
    if (segments < 1) segments = 1;
    return segments;
}

/* ------------------------------------------------------------------ */
/* IC cooling row                                                      */
static double ch_ic_kernel(double eps0, double gamma, double eps) { return chk_ic_kernel(eps0, gamma, eps); }
static double ch_cmb_n(double eps, double z) { return chk_cmb_n(eps, z); }

static double ch_cmb_dn(double eps, double z)
{
    const double h_PL = 6.63e-27, T0 = 2.757, meg = 9.11e-28, k = 1.38 * pow(10.0, -16.0);
    const double Tcmb = T0 * (1.0 + z);
    return 8.0 * M_PI * C_LIGHT * C_LIGHT * C_LIGHT * pow(meg / h_PL, 3.0) * eps /
           (exp(eps * meg * C_LIGHT * C_LIGHT / (k * Tcmb)) - 1.0) *
           (2.0 - eps * meg * C_LIGHT * C_LIGHT * exp(eps * meg * C_LIGHT * C_LIGHT / (k * Tcmb)) /
                      (k * Tcmb * (exp(eps * meg * C_LIGHT * C_LIGHT / (k * Tcmb)) - 1.0)));
}

static void ch_cmb_bounds(double z, double *lo, double *hi)
{
    const double delta = pow(10.0, -5.5), meg = 9.11 * pow(10.0, -28.0), c = 3.0e+10, T0 = 2.725;
    const double k = 1.38 * pow(10.0, -16.0);
    const double epsp = k / (meg * c * c) * T0 * (1.0 + z);
    const double peak = ch_cmb_n(epsp, z);
    double eps = 1.0, epsnew = 0.01 * epsp;
    while (fabs(epsnew - eps) > delta * eps) {
        eps = epsnew;
        epsnew = eps - (ch_cmb_n(eps, z) - peak * delta) / ch_cmb_dn(eps, z);
    }
    *lo = epsnew;
    eps = 1.0;
    epsnew = 10.0 * epsp;
    while (fabs(epsnew - eps) > delta * eps) {
        eps = epsnew;
        epsnew = eps - (ch_cmb_n(eps, z) - peak * delta) / ch_cmb_dn(eps, z);
    }
    *hi = epsnew;
}

static double ch_b_ic_inner(double eps0, double gamma, double emin, double emax)
{
    const double da = 0.50e-1;
    const int N = (int)round((log10(emax) - log10(emin)) / da);
    double epsnew = emin, a = log10(emin), newC = ch_ic_kernel(eps0, gamma, epsnew), sum = 0.0;
    int i;
    for (i = 0; i < N; i++) {
        double eps, C, deps;
        a = a + da;
        eps = epsnew;
        epsnew = pow(10.0, a);
        deps = epsnew - eps;
        C = newC;
        newC = ch_ic_kernel(eps0, gamma, epsnew);
        sum = sum + 0.5 * (eps * C + epsnew * newC) * deps;
    }
    return sum;
}

static double ch_b_ic(double gamma, double e0min, double e0max, double z)
{
    const double da = 0.50e-1;
    const int N = (int)round((log10(e0max) - log10(e0min)) / da);
    double a = log10(e0min), e0new = e0min, sum = 0.0;
    double emax = gamma * 4.0 * e0new * gamma / (1.0 + 4.0 * e0new * gamma);
    double IC2new = ch_b_ic_inner(e0new, gamma, e0new, emax);
    double n0new = ch_cmb_n(e0new, z);
    int i;
    for (i = 0; i < N; i++) {
        double e0, nn, IC2, deps0;
        a = a + da;
        e0 = e0new;
        nn = n0new;
        IC2 = IC2new;
        e0new = pow(10.0, a);
        n0new = ch_cmb_n(e0new, z);
        deps0 = e0new - e0;
        emax = gamma * 4.0 * e0new * gamma / (1.0 + 4.0 * e0new * gamma);
        IC2new = ch_b_ic_inner(e0new, gamma, e0new, emax);
        sum = sum + 0.5 * (nn * IC2 + n0new * IC2new) * deps0;
    }
    return sum * GYR_S;
}

typedef struct {
    const ChGrid *g;
    double z, lo, hi;
    double *row; /* n + 2: bins, then m1, p1 */
} ChIcJob;

static void ch_ic_range(void *arg, long i0, long i1)
{
    ChIcJob *J = (ChIcJob *)arg;
    const int n = J->g->n;
    long i;
    for (i = i0; i < i1; i++) {
        const double p = (i < n) ? J->g->ep[i] : (i == n ? J->g->epm1 : J->g->epp1);
        J->row[i] = ch_b_ic(sqrt(1.0 + p * p), J->lo, J->hi, J->z);
    }
}

void ch_ic_bounds(double z, double *e0min, double *e0max) { ch_cmb_bounds(z, e0min, e0max); }

typedef struct {
    const ChGrid *g;
    double z, lo, hi;
    const int *bins;
    double *out;
} ChIcBinsJob;

static void ch_ic_bins_range(void *arg, long i0, long i1)
{
    ChIcBinsJob *J = (ChIcBinsJob *)arg;
    long i;
    for (i = i0; i < i1; i++) {
        const double p = J->g->ep[J->bins[i]];
        J->out[i] = ch_b_ic(sqrt(1.0 + p * p), J->lo, J->hi, J->z);
    }
}

int ch_ic_cooling_bins(const ChGrid *g, double z, int count, const int *bins, double *out)
{
    ChIcBinsJob J;
    int i;
    if (!g || count < 0 || !bins || !out) return -1;
    for (i = 0; i < count; i++)
        if (bins[i] < 0 || bins[i] >= g->n) return -1;
    J.g = g;
    J.z = z;
    J.bins = bins;
    J.out = out;
    ch_cmb_bounds(z, &J.lo, &J.hi);
    ch_parallel_for(count, 1, ch_ic_bins_range, &J);
    return 0;
}

int ch_ic_cooling_row(const ChGrid *g, double z, double *row, double *m1, double *p1)
{
    ChIcJob J;
    double buf[CH_NMAX + 2];
    if (!g || !row) return -1;
    J.g = g;
    J.z = z;
    J.row = buf;
    ch_cmb_bounds(z, &J.lo, &J.hi);
    ch_parallel_for(g->n + 2, 4, ch_ic_range, &J);
    memcpy(row, buf, (size_t)g->n * sizeof(double));
    if (m1) *m1 = buf[g->n];
    if (p1) *p1 = buf[g->n + 1];
    return 0;
}
