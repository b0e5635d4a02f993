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
double ch_pp_pion_cross_section(int kind, double p)
{
    const double a0 = 0.1176, a1 = 0.3829, a2 = 23.10, a3 = 6.454, a4 = -5.764, a5 = -23.63,
                 a6 = 94.75, a7 = 0.02667, b0 = 11.34, b1 = 23.72, c0 = 28.5, c1 = -6.133,
                 c2 = 1.464, d0 = 0.3522, d1 = 0.1530, d2 = 1.498, d3 = 2.0, d4 = 30.0,
                 d5 = 3.155, d6 = 1.042, e0 = 5.922, e1 = 1.632, f0 = 0.0834, f1 = 9.5, f2 = 5.5,
                 f3 = 1.68, f4 = 3134.0, g0 = 0.0004257, g1 = 4.5, g2 = 7.0, g3 = 2.1,
                 g4 = 503.5;
    const double Ep = sqrt(CH_MP * CH_MP + p * p), Tp = Ep - CH_MP;
    const double x = log10(p), x2 = x * x, x3 = x2 * x;
    double nd, df, dl, rs, tot, pi0, pip, pim;

    if (abs(kind) > 2) return 0.0;
    if (p < 1.0) nd = 0.0;
    else if (p < 1.3) nd = 0.57 * pow(x / a0, 1.2) * (a2 + a3 * x2 + a4 * x3 + a5 * exp(-a6 * pow(x + a7, 2.0)));
    else if (p < 2.4) nd = (b0 * fabs(a1 - x) + b1 * fabs(a0 - x)) / (a1 - a0);
    else if (p < 10.0) nd = a2 + a3 * x2 + a4 * x3 + a5 * exp(-a6 * pow(x + a7, 2.0));
    else nd = c0 + c1 * x + c2 * x2;

    if (p < 2.25) df = 0.0;
    else if (p < 3.2) df = sqrt(fabs(x - d0) / d1) * (d2 + d3 * log10(d4 * (x - 0.25)) + d5 * x2 - d6 * x3);
    else if (p < 100.0) df = d2 + d3 * log10(d4 * (x - 0.25)) + d5 * x2 - d6 * x3;
    else df = e0 + e1 * x;

    if (Ep < 1.4) dl = 0.0;
    else if (Ep < 1.6) dl = f0 * pow(Ep, 10.0);
    else if (Ep < 1.8) dl = f1 * exp(-f2 * pow(Ep - f3, 2.0));
    else if (Ep < 10.0) dl = f4 * pow(Ep, -10.0);
    else dl = 0.0;

    if (Ep < 1.6) rs = 0.0;
    else if (Ep < 1.9) rs = g0 * pow(Ep, 14.0);
    else if (Ep < 2.3) rs = g1 * exp(-g2 * pow(Ep - g3, 2.0));
    else if (Ep < 20.0) rs = g4 * pow(Ep, -6.0);
    else rs = 0.0;

    tot = (nd + df + dl + rs) * MB_CM2;
    pi0 = 1.0 / (0.007 + 0.1 * log(Tp) / Tp + 0.3 / Tp / Tp);
    pip = 1.0 / (0.00717 + 0.0652 * log(Tp) / Tp + 0.162 / Tp / Tp);
    pim = 1.0 / (0.00456 + 0.0846 / pow(Tp, 0.5) + 0.577 / pow(Tp, 1.5));
    if (kind == 0) return tot;
    if (kind == 1) return tot * pip / pi0;
    if (kind == -1) return tot * pim / pi0;
    return tot * (1.0 + pip / pi0 + pim / pi0);
}

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
static double ch_ic_kernel(double eps0, double gamma, double eps)
{
    const double re = 2.82e-13;
    double kappa = eps / (4.0 * eps0 * gamma * (gamma - eps)), C;
    if (kappa <= 0.0) return 0.0;
    C = 2.0 * kappa * log(kappa) + (1.0 + 2.0 * kappa) * (1.0 - kappa) +
        pow(4.0 * eps0 * kappa * gamma, 2.0) / (2.0 * (1.0 + 4.0 * eps0 * gamma * kappa)) * (1.0 - kappa);
    return 2.0 * M_PI * re * re * C_LIGHT / (gamma * gamma * eps0) * C;
}

static double ch_cmb_n(double eps, double z)
{
    const double T0 = 2.757, meg = 9.11e-28, k = 1.38e-16, tiny = 1.0e-40;
    const double Tcmb = T0 * (1.0 + z);
    if (eps * meg * C_LIGHT * C_LIGHT / (k * Tcmb) > 90.0 * log(10.0)) return 0;
    if (exp(eps * meg * C_LIGHT * C_LIGHT / (k * Tcmb)) - 1.0 < tiny) {
        const double x = eps * meg * C_LIGHT * C_LIGHT / (k * Tcmb);
        return 8.0 * M_PI * C_LIGHT * C_LIGHT * C_LIGHT * pow(meg / H_PLANCK, 3.0) * eps /
               (meg * C_LIGHT * C_LIGHT * (1.0 + 0.5 * x + x / 6.0) / (k * Tcmb));
    }
    return 8.0 * M_PI * C_LIGHT * C_LIGHT * C_LIGHT * pow(meg / H_PLANCK, 3.0) * pow(eps, 2.0) /
           (exp(eps * meg * C_LIGHT * C_LIGHT / (k * Tcmb)) - 1.0);
}

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

/* ------------------------------------------------------------------ */
/* Hadronic kernels (secondary electrons, gamma, neutrino)            */
static double ch_FpiQ(int m, double Epi, double Ep)
{
    const double L = log(0.001 * Ep), x = Epi / Ep;
    double mpi = (m == 1) ? 1.40e-1 : 1.35e-1;
    double Bpi, alpha, r, xa, A;
    if (x <= mpi / Ep) return 0.0;
    if (x > 1.0) return 0.0;
    Bpi = 5.58 + 0.78 * L + 0.10 * L * L;
    alpha = 0.89 / (sqrt(Bpi) * (1.0 - exp(-0.33 * Bpi)));
    r = 3.1 / (pow(Bpi, 1.5));
    xa = pow(x, alpha);
    A = pow((1.0 - xa) / pow(1.0 + r * xa, 3.0), 4.0);
    A = A * (1.0 / (1.0 - xa) + 3.0 * r / (1.0 + r * xa)) * sqrt(1.0 - mpi / (x * Ep));
    return 4.0 * alpha * Bpi * xa * A / x / Ep;
}

static double ch_dsigma(int i, double Epi, double Ep)
{
    const double gp = Ep / CH_MP, gbp = sqrt(gp * gp - 1.0), pp = gbp * CH_MP;
    return ch_pp_pion_cross_section(i, pp) * ch_FpiQ(0, Epi, Ep) / sqrt(1.0 + 1.0 / (gbp * gbp));
}

static double ch_Fe(double Ee, double Epi)
{
    const double mpi = 1.3957e-1, mmu = 1.05658e-1, bmu = 0.2714;
    const double gpi = Epi / mpi, xi = mmu * mmu / (mpi * mpi);
    const double Emu = (mpi * mpi - mmu * mmu) / (mpi * mpi) * Epi / (2.0 * bmu);
    const double bm = sqrt(1.0 - mmu * mmu / (Emu * Emu));
    const double lam = 2.0 * mpi * mpi * Ee * bmu / ((mpi * mpi - mmu * mmu) * Epi);
    const double l2 = lam * lam, l3 = l2 * lam;
    const double P = 1.0 / bm * (2.0 * Epi * xi / (Emu * (1.0 - xi)) - (1.0 + xi) / (1.0 - xi));
    if (lam > 1.0 || gpi == 1.0 || lam < 0.0) return 0.0;
    if ((1.0 - bm) / (1.0 + bm) < lam)
        return (5.0 - 9.0 * l2 + 4.0 * l3 - P / bm * (1.0 - (6.0 * bm + 3.0) * l2 + (6.0 * bm + 2.0) * l3)) / 12.0;
    if (lam < (1.0 - bm) / (1.0 + bm))
        return l2 * bm / pow(1.0 - bm, 2.0) * (3.0 - 2.0 / 3.0 * lam * (3.0 + bm * bm) / (1.0 - bm)) -
               P / (1.0 - bm) *
                   (l2 * (1.0 + bm) - 2.0 * l2 / (1.0 - bm) * (0.5 + lam * (1.0 + bm)) +
                    2.0 * l3 * (3.0 + bm * bm) / (3.0 * pow(1.0 - bm, 2.0)));
    return 0.0;
}

static double ch_lorentz_beta(double gamma)
{
    const double b2 = gamma * gamma - 1.0;
    if (b2 < 0.0) return 0.0;
    return sqrt(b2) / gamma;
}

static double ch_qe_kernel(double Ee, double p)
{
    const double bmu = 0.2714, gp = sqrt(p * p + 1.0), Ep = gp * CH_MP, Emax = gp * CH_MP;
    const double r = 0.5 * (CH_MPIC * CH_MPIC - CH_MMU * CH_MMU) / (bmu * CH_MPIC * CH_MPIC);
    const double da = 0.05;
    double Emin = Ee / r, a, f, sum = 0.0, Epi, bmu_v;
    int i, Na;
    if (Emax < Emin || (Ep - CH_MP) < 0.289) return 0;
    Emin = fmax(Emin, CH_MPIC);
    Na = (int)fmax(50, round((log10(Emax) - log10(Emin)) / da));
    a = log10(Emin);
    Epi = pow(10.0, a);
    bmu_v = ch_lorentz_beta(Epi * r / CH_MMU);
    f = (bmu_v == 0.0) ? 0.0 : (ch_dsigma(1, Epi, Ep) + ch_dsigma(-1, Epi, Ep)) * ch_Fe(Ee, Epi) / (bmu_v * Epi);
    for (i = 1; i < Na; i++) {
        const double Epre = Epi, fpre = f;
        a = a + da;
        Epi = pow(10.0, a);
        bmu_v = ch_lorentz_beta(Epi * r / CH_MMU);
        f = (ch_dsigma(1, Epi, Ep) + ch_dsigma(-1, Epi, Ep)) * ch_Fe(Ee, Epi) / (bmu_v * Epi);
        sum = sum + 0.5 * (f + fpre) * (Epi - Epre);
        if (sum < 0.0 || isnan(sum)) return 0.0;
    }
    return 8.0 * bmu * CH_MPIC * CH_MPIC * (C_LIGHT * CH_ME) * sum / (CH_MPIC * CH_MPIC - CH_MMU * CH_MMU);
}

typedef struct {
    const ChGrid *g;
    double *fqe_t;
    int *np_min;
} ChSecJob;

static void ch_sec_range(void *arg, long i0, long i1)
{
    ChSecJob *J = (ChSecJob *)arg;
    const ChGrid *g = J->g;
    const int n = g->n;
    double vals[CH_NMAX];
    long je;
    int jp;
    for (je = i0; je < i1; je++) {
        int jp0, npmin = 1;
        for (jp = 0; jp < n; jp++) {
            double v = ch_qe_kernel(g->eE[je], g->pp[jp]) * g->beta_p[jp];
            if (isnan(v)) v = 0.0;
            vals[jp] = v;
        }
        for (jp = 1; jp < n; jp++)
            if (vals[jp] < 1.0e-80) npmin = jp;
        jp0 = npmin < 1 ? 1 : (npmin > n ? n : npmin);
        for (jp = 0; jp < n; jp++) {
            double w = 0.0;
            if (jp0 < n) {
                if (jp == jp0 - 1) w = 0.5 * vals[jp] * g->pdp[jp0];
                else if (jp >= jp0 && jp < n - 1) w = 0.5 * vals[jp] * (g->pdp[jp] + g->pdp[jp + 1]);
                else if (jp == n - 1 && jp >= jp0) w = 0.5 * vals[jp] * g->pdp[jp];
            }
            J->fqe_t[(size_t)jp * n + je] = w;
        }
        J->np_min[je] = (jp0 < n) ? (jp0 - 1) : n;
    }
}

int ch_secondary_weights(const ChGrid *g, double *fqe_t, int *np_min_qe, int *jp0_out)
{
    ChSecJob J;
    int je, jp0;
    J.g = g;
    J.fqe_t = fqe_t;
    J.np_min = np_min_qe;
    ch_parallel_for(g->n, 1, ch_sec_range, &J);
    jp0 = g->n;
    for (je = 0; je < g->n; je++) {
        int v = np_min_qe[je] < 0 ? 0 : np_min_qe[je];
        if (v < jp0) jp0 = v;
    }
    if (jp0_out) *jp0_out = jp0;
    return 0;
}

double ch_gamma_kernel(double Eg, double p)
{
    const int Na = 200;
    const double gp = sqrt(p * p + 1.0), Ep = gp * CH_MP, Emax = Ep;
    double Emin = Eg + CH_MPI0 * CH_MPI0 / (4.0 * Eg), a, da, Epi, f, sum = 0.0;
    int i;
    if (Emax < Emin) return 0;
    if (Emin < CH_MPI0) Emin = CH_MPI0;
    if ((Ep - CH_MP) < 0.285) return 0;
    da = (log10(Emax) - log10(Emin)) / (double)Na;
    a = log10(Emin);
    Epi = pow(10.0, a);
    f = ch_dsigma(0, Epi, Ep) / sqrt(Epi * Epi - CH_MPI0 * CH_MPI0);
    for (i = 1; i < Na; i++) {
        const double Epre = Epi, fpre = f;
        a += da;
        Epi = pow(10.0, a);
        f = ch_dsigma(0, Epi, Ep) / sqrt(Epi * Epi - CH_MPI0 * CH_MPI0);
        sum += 0.5 * (f + fpre) * (Epi - Epre);
    }
    return 2.0 * C_LIGHT * sum / (4.0 * M_PI);
}

static double ch_fe(double x)
{
    const double r = CH_MMU * CH_MMU / (CH_MPIC * CH_MPIC), x2 = x * x, r2 = r * r;
    if (x > r) return (3.0 - 2.0 * r) * (9.0 * x2 - 6.0 * log(x) - 4.0 * x2 * x - 5.0) / (9.0 * pow(1.0 - r, 2.0));
    return (3.0 - 2.0 * r) * (9.0 * r2 - 6.0 * log(r) - 4.0 * r2 * r - 5.0) / (9.0 * pow(1.0 - r, 2.0)) +
           (1.0 + 2.0 * r) * (r - x) * (9.0 * (r + x) - 4.0 * (r2 + r * x + x2)) / (9.0 * r2);
}

static double ch_fnue(double x)
{
    const double r = 0.573, x2 = x * x, r2 = r * r;
    if (x > r)
        return 2.0 * (1.0 - x) * (6.0 * pow(1.0 - x, 2.0) + r * (5.0 + 5.0 * x - 4.0 * x2) + 6.0 * r * log(x) / (1.0 - x)) /
               (3.0 * pow(1.0 - r, 2.0));
    return 2.0 * ((1.0 - r) * (6.0 - 7.0 * r + 11.0 * r2 - 4.0 * r2 * r) + 6.0 * r * log(r)) / (3.0 * pow(1.0 - r, 2.0)) +
           2.0 * (r - x) * (7.0 * r2 - 4.0 * r2 * r + 7.0 * x * r - 4.0 * x * r2 - 2.0 * x2 - 4.0 * x2 * r) / (3.0 * r2);
}

static double ch_lepton_integral(int kind, int sgn, double E, double Ep, double dx, double scale_x)
{
    double x = 0.0, f = 0.0, sum = 0.0;
    int i;
    for (i = 1; i < 200; i++) {
        const double fpre = f;
        double Epi;
        x += dx;
        Epi = E / x;
        if (kind == 1) f = ch_dsigma(sgn, Epi, Ep) * ch_fnue(x) / x;
        else if (kind == 2) f = ch_dsigma(sgn, Epi, Ep) * ch_fe(x) / x;
        else f = ch_dsigma(sgn, Epi, Ep) / x;
        if (i == 1) sum += f * dx;
        else sum += 0.5 * (f + fpre) * dx;
    }
    (void)scale_x;
    return sum;
}

static double ch_ekernel(int kind, double E, double p)
{
    const double gp = sqrt(p * p + 1.0), Ep = gp * CH_MP, Emax = gp * CH_MP;
    double Emin = 2.0 * E * CH_MPIC * CH_MPIC / (CH_MPIC * CH_MPIC + CH_MMU * CH_MMU);
    if (Emax < Emin) return 0;
    if (Emin < CH_MPIC) Emin = CH_MPIC;
    if ((Ep - CH_MP) < 0.289) return 0;
    return C_LIGHT * (ch_lepton_integral(kind, 1, E, Ep, 1.0 / 200.0, 1.0) +
                      ch_lepton_integral(kind, -1, E, Ep, 1.0 / 200.0, 1.0)) / (4.0 * M_PI);
}

static double ch_numu1(double Enu, double p)
{
    const double lam = 1.0 - CH_MMU * CH_MMU / (CH_MPIC * CH_MPIC);
    const double gp = sqrt(p * p + 1.0), Ep = gp * CH_MP, Emax = gp * CH_MP;
    double Emin = Enu / lam;
    if (Emax < Emin) return 0;
    if (Emin < CH_MPIC) Emin = CH_MPIC;
    if ((Ep - CH_MP) < 0.29) return 0;
    return C_LIGHT * (ch_lepton_integral(0, 1, Enu, Ep, lam / 200.0, 1.0) +
                      ch_lepton_integral(0, -1, Enu, Ep, lam / 200.0, 1.0)) / lam / (4.0 * M_PI);
}

double ch_nu_kernel(double E, double p) { return ch_ekernel(2, E, p) + ch_ekernel(1, E, p) + ch_numu1(E, p); }

double ch_ic_emissivity_kernel(double nu, double gamma, double z)
{
    const int N = 400;
    const double meg = 9.11e-28, c = 3.0e+10;
    const double eps = H_PLANCK * nu / (meg * c * c);
    double e0max = eps, e0min = eps / (4.0 * gamma * (gamma - eps)), a, da, Cn, nn, e0n, sum = 0.0;
    int i;
    if (e0min < 0) e0min = eps / (gamma * gamma);
    da = (log10(e0max) - log10(e0min)) / (double)N;
    a = log10(e0min);
    e0n = pow(10.0, a);
    Cn = ch_ic_kernel(e0n, gamma, eps);
    nn = ch_cmb_n(e0n, z);
    for (i = 0; i < N; i++) {
        const double e0 = e0n, C = Cn, n0 = nn;
        a = a + da;
        e0n = pow(10.0, a);
        nn = ch_cmb_n(e0n, z);
        Cn = ch_ic_kernel(e0n, gamma, eps);
        sum = sum + 0.5 * (C * n0 + Cn * nn) * (e0n - e0);
    }
    return H_PLANCK * eps * sum / (4.0 * M_PI);
}

typedef struct {
    const ChGrid *g;
    int kind; /* 0 gamma, 1 nu, 2 IC */
    int nbins;
    double lo, hi, z;
    double *out; /* [nbins * n] */
} ChTabJob;

static void ch_tab_range(void *arg, long i0, long i1)
{
    ChTabJob *J = (ChTabJob *)arg;
    const int n = J->g->n;
    const double d = (J->hi - J->lo) / (double)J->nbins;
    long k;
    for (k = i0; k < i1; k++) {
        /* same accumulated log-grid as the reference loops */
        double a = J->lo - d, e;
        long kk;
        int j;
        for (kk = 0; kk <= k; kk++) a += d;
        e = pow(10.0, a);
        for (j = 0; j < n; j++) {
            double v;
            if (J->kind == 0) v = ch_gamma_kernel(e, J->g->pp[j]);
            else if (J->kind == 1) v = ch_nu_kernel(e, J->g->pp[j]);
            else v = ch_ic_emissivity_kernel(e, sqrt(1.0 + J->g->ep[j] * J->g->ep[j]), J->z);
            J->out[(size_t)k * n + j] = v;
        }
    }
}

int ch_emission_table(const ChGrid *g, int kind, int nbins, double lo, double hi, double z, double *out)
{
    ChTabJob J;
    J.g = g;
    J.kind = kind;
    J.nbins = nbins;
    J.lo = lo;
    J.hi = hi;
    J.z = z;
    J.out = out;
    ch_parallel_for(nbins, 1, ch_tab_range, &J);
    return 0;
}
