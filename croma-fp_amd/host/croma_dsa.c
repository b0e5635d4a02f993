/*
 * croma_dsa.c -- DSA source builder: shock-accelerated CR injection per
 * tracer and snapshot (tracer_fp_apply_dsa_tracer_injection_snapshot,
 * tracer_fp_dsa.c:184-505), with the injection models of DSA_MODELS.c and
 * the reacceleration of the pre-shock spectrum (positive delta, Drury 1983
 * convolution; tracer_fp_dsa_reacc.c).  Host C; the rows it produces are the
 * qpi / qepri (source mode) or state increments (state mode) the solve
 * consumes.
 *
 * The arithmetic follows the reference expression by expression (FP64, no
 * contraction), so the rows equal the reference's bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "croma_host.h"

/* CGS constants of the injection models (DSA_MODELS.c:9-19) and M_sun
 * (CONSTANTS.c:6) */
#define DS_ME 9.1093837015e-28
#define DS_MP 1.67262192369e-24
#define DS_KB 1.380649e-16
#define DS_C 2.99792458e10
#define DS_4PI 12.566370614359172
#define DS_MSUN 1.99e+33
#define DS_NSIMPSON 4000
/* the reacceleration's kinetic energies use the solver's c (CONSTANTS.c:3)
 * and the driver's own masses and k_B (tracer_fp_dsa.c:22-24) */
#define RA_C 2.999e+10
#define RA_MP 1.67262192369e-24
#define RA_ME 9.1093837015e-28
#define RA_KB 1.380649e-16

enum { FAM_KR07, FAM_KR13, FAM_RYU19 };

/* efficiency fits eta(M) = sum_k b_k (M - 1)^k / M^4 with their branches;
 * x_cr0: the CR pressure fraction at which a reacceleration model takes
 * over from the initial one */
typedef struct {
    int family;
    int reacc;
    double x_cr0;
    double b[5];
    double mach_crit;
    double scale;
} DsModel;

/* the reference's model table (DSA_MODELS.c:287-354), DSAModelID order */
static const DsModel ds_models[8] = {
    {FAM_KR07, 0, 0.0, {5.46, -9.78, 4.17, -0.33, 0.57}, 1.0, 1.0},
    {FAM_KR07, 1, 0.3, {0.24, -1.56, 2.8, 0.512, 0.557}, 1.0, 1.0},
    {FAM_KR13, 0, 0.0, {-2.87, 9.67, -8.88, 1.94, 0.18}, 2.0, 1.0},
    {FAM_KR13, 1, 0.05, {-0.72, 2.73, -3.29, 1.34, 0.19}, 2.0, 1.0},
    {FAM_KR13, 0, 0.0, {-2.87, 9.67, -8.88, 1.94, 0.18}, 2.0, 0.5},
    {FAM_KR13, 1, 0.05, {-0.72, 2.73, -3.29, 1.34, 0.19}, 2.0, 0.5},
    {FAM_RYU19, 0, 0.0, {-1.5255, 2.4026, -1.2534, 0.2215, 0.0336}, 2.25, 1.0},
    {FAM_RYU19, 1, 0.05, {0.3965, -0.21898, -0.2074, 0.1319, 0.0351}, 2.25, 1.0},
};

void croma_hip_dsa_default_params(CromaHipDsaParams *p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->eta_model = CROMA_HIP_DSA_RYU19;
    p->inject_species = CROMA_HIP_DSA_SPECIES_BOTH;
    p->chi_p = 3.5;
    p->chi_e = 3.5;
    p->k_ep = 1.0e-2;
    p->pmax_p = 1.0e6;
    p->pmax_e = 1.0e5;
    p->gamma_gas = 5.0 / 3.0;
    p->min_mach = 1.5;
    p->mu_mol = 0.59;
    p->pmin_dsa_p_log10 = -3.0;
    p->pmin_dsa_e_log10 = -3.0;
    p->reacc_mode = CROMA_HIP_DSA_REACC_OFF;
    p->eta_model_reacc = CROMA_HIP_DSA_RYU19R;
    p->reacc_eta_cap = 0.0;
    p->xcr_pmin_pmc = 1.0;
}

static double ds_poly(double mach, const double b[5])
{
    double x;
    if (mach <= 0.0) return 0.0;
    x = mach - 1.0;
    return (b[0] + b[1] * x + b[2] * x * x + b[3] * x * x * x + b[4] * x * x * x * x) /
           (mach * mach * mach * mach);
}

double croma_hip_dsa_eta(double mach, int id)
{
    const DsModel *m;
    double eta;
    if (id < 0 || id >= 8) return 0.0;
    m = &ds_models[id];
    if (mach < m->mach_crit) return 0.0;
    if (m->family == FAM_KR07 && !m->reacc) {
        eta = (mach <= 2.0) ? 1.96e-3 * (mach * mach - 1.0) : ds_poly(mach, m->b);
    } else if (m->family == FAM_KR07) {
        if (mach <= 1.5) {
            /* low-Mach reacceleration branch (KR07 eq. for delta_0), gamma = 5/3 */
            const double g = 5.0 / 3.0;
            const double m2 = mach * mach;
            const double xs = (g + 1.0) / (g - 1.0 + 2.0 / m2);
            const double d0 =
                2.0 * (((2.0 * g * m2 - g + 1.0) / (g + 1.0)) - pow(xs, g)) / (g * (g - 1.0) * m2 * xs);
            eta = 1.025 * d0;
        } else {
            eta = ds_poly(mach, m->b);
        }
    } else if (m->family == FAM_KR13 && !m->reacc) {
        if (mach <= 5.0) eta = -0.0005950569221922047 + 1.880258286365841e-5 * pow(mach, 5.334076006529829);
        else if (mach <= 15.0) eta = ds_poly(mach, m->b);
        else eta = 0.21152;
    } else if (m->family == FAM_KR13) {
        eta = (mach <= 17.7) ? ds_poly(mach, m->b) : 0.2055;
    } else {
        eta = (mach <= 34.0) ? ds_poly(mach, m->b) : 0.0348;
    }
    eta *= m->scale;
    return eta < 0.0 ? 0.0 : eta;
}

/* ---- reacceleration of the pre-shock spectrum (tracer_fp_dsa_reacc.c,
 * driven as tracer_fp_dsa.c:322-418 does) ---------------------------- */

/* kinetic energy [erg] of momentum p (m c) */
static double ra_ekin(double p, int electron)
{
    const double m = electron ? RA_ME : RA_MP;
    return m * RA_C * RA_C * (sqrt(1.0 + p * p) - 1.0);
}

/* eta weighted between the initial model and the reacceleration model by
 * x_cr / x_cr0 of the latter, clipped to [0, 1] (DSA_MODELS.c:474-522) */
double croma_hip_dsa_eta_interpolated(double mach, int initial_id, int reacc_id, double x_cr)
{
    double e0, e1, x0, w;
    if (initial_id < 0 || initial_id >= 8 || reacc_id < 0 || reacc_id >= 8) return 0.0;
    e0 = croma_hip_dsa_eta(mach, initial_id);
    e1 = croma_hip_dsa_eta(mach, reacc_id);
    x0 = ds_models[reacc_id].x_cr0;
    if (x0 <= 0.0) return e0;
    w = x_cr / x0;
    if (w < 0.0) w = 0.0;
    if (w > 1.0) w = 1.0;
    return (1.0 - w) * e0 + w * e1;
}

/* kinetic energy density of a state row above pmin in a volume
 * (tracer_fp_dsa.c:120-139) */
static double ra_state_edens(const double *p, const double *dp, const double *state, int n, double vol, double pmin,
                             int electron)
{
    double e = 0.0;
    int j;
    if (!(vol > 0.0)) return 0.0;
    for (j = 0; j < n; j++) {
        const double ek = ra_ekin(p[j], electron);
        if (p[j] < pmin) continue;
        e += ek * state[j] * dp[j];
    }
    return e / vol;
}

/* the energy density of a delta row over bins 2..n-3 above pmin
 * (tracer_dsa_reacc_delta_energy_density_1d) */
static double ra_delta_edens(const double *p, const double *dp, const double *delta, int n, double pmin, int electron)
{
    double e = 0.0;
    int j;
    if (n <= 4) return 0.0;
    for (j = 2; j <= n - 3; j++) {
        if (!(p[j] > 0.0) || p[j] < pmin || !isfinite(delta[j])) continue;
        e += ra_ekin(p[j], electron) * delta[j] * dp[j];
    }
    return e;
}

/* positive delta: the pre-shock spectrum f seeds the test-particle solution
 * 4 pi p^2 q p^-q int p'^(q-1) f dp' from the second bin on; where that
 * exceeds the current density the excess is the delta (tracer_dsa_reacc_
 * positive_delta_1d).  Returns the delta's energy density. */
static double ra_positive_delta(const double *p, const double *dp, const double *state, int n, double vd, double q,
                                double pmin, double *delta, int electron)
{
    double acc = 0.0, e = 0.0;
    int j;
    for (j = 0; j < n; j++) delta[j] = 0.0;
    if (n <= 0 || !(vd > 0.0) || !(q > 0.0)) return 0.0;
    for (j = (n > 1); j < n; j++) {
        const double pj = p[j];
        double cur, f, re;
        if (!(pj > 0.0) || pj < pmin) continue;
        cur = state[j] / vd;
        f = cur / (4.0 * M_PI * pj * pj);
        acc += pow(pj, q - 1.0) * f * dp[j];
        re = 4.0 * M_PI * pj * pj * q * pow(pj, -q) * acc;
        if (re > cur) {
            delta[j] = re - cur;
            e += ra_ekin(pj, electron) * delta[j] * dp[j];
        }
    }
    return e;
}

/* Drury (1983): the upstream spectrum convolved to (a + 2) p^-a int p'^(a-1)
 * N_up dp', a = q - 2, minus the current downstream spectrum, the delta not
 * below minus the current value (tracer_dsa_reacc_convolution_delta_1d).
 * Returns the delta's energy density over every bin above pmin. */
static double ra_drury(const double *p, const double *dp, const double *state, int n, double vu, double vd, double q,
                       double pmin, double *delta, int electron)
{
    const double a = q - 2.0;
    double acc = 0.0, e = 0.0;
    int j;
    for (j = 0; j < n; j++) delta[j] = 0.0;
    if (n <= 0 || !(vu > 0.0) || !(vd > 0.0) || !(q > 0.0)) return 0.0;
    for (j = (n > 1); j < n; j++) {
        const double pj = p[j];
        double seed, cur, d;
        if (!(pj > 0.0)) continue;
        seed = state[j] / vu;
        cur = state[j] / vd;
        if (pj < pmin) continue;
        acc += pow(pj, a - 1.0) * seed * dp[j];
        d = (a + 2.0) * pow(pj, -a) * acc - cur;
        if (!isfinite(d)) continue;
        if (d < -cur) d = -cur;
        delta[j] = d;
        e += ra_ekin(pj, electron) * d * dp[j];
    }
    return e;
}

/* value at p_mc of a row, linear between the bins that bracket it inside
 * bins 1..n-2, 0 outside (tracer_fp_dsa_reacc.c:24-50) */
static double ra_interp(const double *p, const double *row, int n, double x)
{
    int lo = 1, hi = n - 2;
    if (n <= 2 || !(x > 0.0) || x < p[lo] || x > p[hi]) return 0.0;
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (p[mid] <= x) lo = mid;
        else hi = mid;
    }
    if (p[hi] <= p[lo]) return row[lo];
    return row[lo] + (row[hi] - row[lo]) * (x - p[lo]) / (p[hi] - p[lo]);
}

/* pure adiabatic compression by r: p -> r^(1/3) p, density x r^(2/3), minus
 * the current downstream spectrum (tracer_dsa_reacc_adiabatic_delta_1d);
 * sd: n doubles of work.  Returns the delta's energy density. */
static double ra_adiabatic(const double *p, const double *dp, const double *state, int n, double vu, double vd,
                           double r, double pmin, double *delta, double *sd, int electron)
{
    double ps, norm;
    int j;
    for (j = 0; j < n; j++) delta[j] = 0.0;
    if (n <= 0 || !(vu > 0.0) || !(vd > 0.0) || !(r > 0.0) || !isfinite(r)) return 0.0;
    ps = cbrt(r);
    norm = ps * ps;
    for (j = 0; j < n; j++) sd[j] = state[j] / vu;
    for (j = 0; j < n; j++) {
        double cur, ad;
        if (!(p[j] > 0.0) || p[j] < pmin) continue;
        cur = state[j] / vd;
        ad = norm * ra_interp(p, sd, n, p[j] / ps);
        if (!isfinite(cur) || !isfinite(ad)) continue;
        delta[j] = ad - cur;
    }
    return ra_delta_edens(p, dp, delta, n, pmin, electron);
}

static double ds_ratio(double mach, double g)
{
    double m2, num, den;
    if (mach <= 1.0 || g <= 1.0) return 1.0;
    m2 = mach * mach;
    num = (g + 1.0) * m2;
    den = (g - 1.0) * m2 + 2.0;
    if (den <= 0.0) return 1.0;
    return num / den;
}

/* q of f(p) ~ p^-q per d^3p */
static double ds_slope(double mach, double g)
{
    double r;
    if (mach <= 1.0) return 0.0;
    r = ds_ratio(mach, g);
    if (r <= 1.0) return 0.0;
    return 3.0 * r / (r - 1.0);
}

/* chi * sqrt(2 k T / (m c^2)) */
static double ds_pinj(double temp, double chi, double mass)
{
    double theta;
    if (chi <= 0.0 || temp <= 0.0) return 0.0;
    theta = 2.0 * DS_KB * temp / (mass * DS_C * DS_C);
    if (theta <= 0.0) return 0.0;
    return chi * sqrt(theta);
}

static double ds_shape(double p, double pinj, double q, double pmax)
{
    double ri, rm, ex;
    if (p <= 0.0 || pinj <= 0.0 || pmax <= 0.0 || p < pinj) return 0.0;
    ri = p / pinj;
    rm = p / pmax;
    ex = -rm * rm;
    if (ex < -700.0) ex = -700.0;
    if (ex > 700.0) ex = 700.0;
    return pow(ri, -q) * exp(ex);
}

/* (gamma - 1) p^2 shape(p): kinetic energy per particle in m_p c^2 */
static double ds_energy_integrand(double p, double pinj, double q, double pmax)
{
    double ke;
    if (p <= 0.0 || pinj <= 0.0 || pmax <= 0.0 || p < pinj) return 0.0;
    ke = sqrt(1.0 + p * p) - 1.0;
    return ke * p * p * ds_shape(p, pinj, q, pmax);
}

/* kinetic energy density per unit A_p: 4 pi m_p c^2 int (gamma - 1) p^2 shape dp, Simpson */
static double ds_edens_unit(double pinj, double q, double pmax)
{
    const int ns = DS_NSIMPSON;
    double h, sum;
    int i;
    if (pinj <= 0.0 || pmax <= 0.0 || pmax <= pinj) return 0.0;
    h = (pmax - pinj) / (double)ns;
    sum = ds_energy_integrand(pinj, pinj, q, pmax) + ds_energy_integrand(pmax, pinj, q, pmax);
    for (i = 1; i < ns; i++)
        sum += ((i % 2 == 0) ? 2.0 : 4.0) * ds_energy_integrand(pinj + h * (double)i, pinj, q, pmax);
    return DS_4PI * DS_MP * DS_C * DS_C * (sum * h / 3.0);
}

typedef struct {
    double q, a, pinj, pmax;
} DsSpec;

static int ds_proton_spec(const CromaHipDsaParams *P, double mach, double temp, double rho1, double u1, double u2,
                          DsSpec *s)
{
    double eta, q, pinj, unit, target;
    if (mach <= 1.0 || P->gamma_gas <= 1.0) return 0;
    if (rho1 <= 0.0 || u1 <= 0.0 || u2 <= 0.0 || temp <= 0.0) return 0;
    if (P->chi_p <= 0.0 || P->pmax_p <= 0.0) return 0;
    eta = croma_hip_dsa_eta(mach, P->eta_model);
    q = ds_slope(mach, P->gamma_gas);
    pinj = ds_pinj(temp, P->chi_p, DS_MP);
    if (eta <= 0.0 || q <= 0.0 || pinj <= 0.0 || P->pmax_p <= pinj) return 0;
    /* normalisation: eta of the upstream kinetic energy flux, per u2 */
    unit = ds_edens_unit(pinj, q, P->pmax_p);
    if (unit <= 0.0) return 0;
    target = eta * (0.5 * rho1 * u1 * u1 * u1) / u2;
    s->a = target / unit;
    if (s->a <= 0.0) return 0;
    s->q = q;
    s->pinj = pinj;
    s->pmax = P->pmax_p;
    return 1;
}

/* electrons at K_ep of the protons at the proton injection momentum
 * (suprathermal electron injection, DSA_MODELS.c:1022-1070) */
static int ds_electron_spec(const CromaHipDsaParams *P, double temp, const DsSpec *sp, DsSpec *se)
{
    const double chi_e = P->chi_e > 0.0 ? P->chi_e : P->chi_p;
    double pinj, ratio, sh_p, sh_e, mf;
    if (temp <= 0.0 || P->pmax_e <= 0.0 || chi_e <= 0.0) return 0;
    pinj = ds_pinj(temp, chi_e, DS_ME);
    if (pinj <= 0.0 || P->pmax_e <= pinj) return 0;
    if (P->k_ep <= 0.0 || sp->a <= 0.0) return 0;
    ratio = DS_MP / DS_ME;
    sh_p = ds_shape(sp->pinj, sp->pinj, sp->q, sp->pmax);
    sh_e = ds_shape(ratio * sp->pinj, pinj, sp->q, P->pmax_e);
    if (sh_p <= 0.0 || sh_e <= 0.0) return 0;
    mf = 1.0 / ((DS_MP / DS_ME) * (DS_MP / DS_ME) * (DS_MP / DS_ME));
    se->a = P->k_ep * mf * sp->a * sh_p / sh_e;
    if (se->a <= 0.0) return 0;
    se->q = sp->q;
    se->pinj = pinj;
    se->pmax = P->pmax_e;
    return 1;
}

/* 4 pi p^2 f(p) on the FP bins: the DSA grid extends the FP grid's log
 * spacing down to pmin_dsa; FP bin j is DSA bin j + n_extra, whose momentum
 * is recomputed from the extended grid's start (DSA_MODELS.c:208-285,
 * 1396-1488) */
static int ds_rows(const double *p, int n, double pmin_dsa, const DsSpec *s, double *row)
{
    double dlogp, pmin_fp, start;
    int nx = 0, j;
    if (p[0] <= 0.0 || p[1] <= 0.0) return 0;
    dlogp = log10(p[1]) - log10(p[0]);
    if (dlogp <= 0.0) return 0;
    pmin_fp = log10(p[0]);
    if (pmin_dsa < pmin_fp) {
        nx = (int)ceil((pmin_fp - pmin_dsa) / dlogp - 1.0e-12);
        if (nx < 0) nx = 0;
    }
    start = pmin_fp - dlogp * (double)nx;
    for (j = 0; j < n; j++) {
        const double pj = pow(10.0, start + dlogp * (double)(j + nx));
        double f = 0.0;
        if (pj > 0.0 && pj >= s->pinj && pj <= s->pmax)
            f = s->a * pow(pj / s->pinj, -s->q) * exp(-(pj / s->pmax) * (pj / s->pmax));
        row[j] = (f <= 0.0) ? 0.0 : DS_4PI * pj * pj * f;
    }
    return 1;
}

static int ds_grid(const CromaHipGridSpec *gs, ChGrid **out)
{
    ChGrid *g;
    if (!gs || gs->n < 2 || gs->n > CH_NMAX) return CROMA_HIP_EINVAL;
    g = (ChGrid *)malloc(sizeof(*g));
    if (!g) return CROMA_HIP_ENOMEM;
    if (ch_grid_build(g, gs->n, gs->pmin, gs->pmax, gs->pemin, gs->pemax) != 0) {
        free(g);
        return CROMA_HIP_EINVAL;
    }
    *out = g;
    return 0;
}

/* both rows of one shock; returns flags bit 0 proton, bit 1 electron */
static int ds_shock(const ChGrid *g, const CromaHipDsaParams *P, double mach, double temp, double rho1, double u1,
                    double *sp_row, double *se_row)
{
    DsSpec sp, se;
    const double r = ds_ratio(mach, P->gamma_gas);
    int fl = 0;
    memset(sp_row, 0, (size_t)g->n * sizeof(double));
    memset(se_row, 0, (size_t)g->n * sizeof(double));
    if (ds_proton_spec(P, mach, temp, rho1, u1, u1 / r, &sp) && ds_rows(g->pp, g->n, P->pmin_dsa_p_log10, &sp, sp_row)) {
        fl |= 1;
        if (ds_electron_spec(P, temp, &sp, &se) && ds_rows(g->ep, g->n, P->pmin_dsa_e_log10, &se, se_row)) fl |= 2;
    }
    return fl;
}

int croma_hip_dsa_shock_sources(const CromaHipGridSpec *grid, const CromaHipDsaParams *p, double mach,
                                double temp1_K, double rho1_cgs, double u1_cgs, double *source_p,
                                double *source_e, int *flags)
{
    ChGrid *g = 0;
    double u1 = u1_cgs;
    int rc, fl;
    if (!p || !source_p || !source_e || !flags) return CROMA_HIP_EINVAL;
    if ((rc = ds_grid(grid, &g)) != 0) return rc;
    if (!(u1 > 0.0)) u1 = mach * sqrt(p->gamma_gas * DS_KB * temp1_K / (p->mu_mol * DS_MP));
    fl = ds_shock(g, p, mach, temp1_K, rho1_cgs, u1, source_p, source_e);
    flags[0] = fl & 1;
    flags[1] = (fl >> 1) & 1;
    free(g);
    return 0;
}

typedef struct {
    const ChGrid *g;
    const CromaHipDsaParams *P;
    int mode;
    double inv_dt;
    const double *mach, *temp1, *rho1, *u1, *flux, *side, *mass;
    double *crp, *cre, *qpi, *qepri;
    unsigned char *dis;
    int ninj[64];
} DsJob;

/* resolve one tracer's shock (tracer_fp_dsa.c:26-111); 0 = no injection */
static int ds_resolve(const DsJob *J, long i, double *mach, double *temp, double *rho1, double *u1)
{
    const CromaHipDsaParams *P = J->P;
    double m = J->mach[i], t, r, u;
    if (J->flux) {
        const double f = J->flux[i];
        r = J->rho1[i];
        if (!(m >= P->min_mach) || !isfinite(m) || !(r > 0.0) || !isfinite(r)) return 0;
        if (!(f > 0.0) || !isfinite(f)) return 0;
        u = cbrt(2.0 * f / r);
        if (!(u > 0.0) || !isfinite(u)) return 0;
        t = (P->mu_mol * DS_MP / (P->gamma_gas * DS_KB)) * (u * u) / (m * m);
        if (!(t > 0.0) || !isfinite(t)) return 0;
    } else {
        t = J->temp1[i];
        if (!(m >= P->min_mach) || !(t > 0.0) || !isfinite(m) || !isfinite(t)) return 0;
        if (J->side) {
            const int sc = (int)floor(J->side[i] + 0.5);
            if (sc == -1 || sc == 99) return 0;
        }
        r = J->rho1[i];
        if (!(r > 0.0) || !isfinite(r)) return 0;
        u = (J->u1 && J->u1[i] > 0.0) ? J->u1[i] : m * sqrt(P->gamma_gas * DS_KB * t / (P->mu_mol * DS_MP));
        if (!(u > 0.0) || !isfinite(u)) return 0;
    }
    *mach = m;
    *temp = t;
    *rho1 = r;
    *u1 = u;
    return 1;
}

/* The reacceleration rows of one spectrum on a grid (the functions of
 * include/tracer_fp_dsa_reacc.h): kind CROMA_HIP_DSA_REACC_KIND_*;
 * species 0 protons, 1 electrons; returns the delta's energy density. */
int croma_hip_dsa_reacc_delta(const CromaHipGridSpec *grid, int kind, int species, const double *state,
                              double volume_upstream, double volume_downstream, double q, double compression_ratio,
                              double pmin_mc, double *delta, double *edens)
{
    ChGrid *g = 0;
    const double *p, *dp;
    double *sd = 0;
    int rc;
    if (!state || !delta || !edens || species < 0 || species > 1 || kind < 0 || kind > 3) return CROMA_HIP_EINVAL;
    if ((rc = ds_grid(grid, &g)) != 0) return rc;
    p = species ? g->ep : g->pp;
    dp = species ? g->edp : g->pdp;
    if (kind == CROMA_HIP_DSA_REACC_KIND_POSITIVE) {
        *edens = ra_positive_delta(p, dp, state, g->n, volume_downstream, q, pmin_mc, delta, species);
    } else if (kind == CROMA_HIP_DSA_REACC_KIND_DRURY) {
        *edens = ra_drury(p, dp, state, g->n, volume_upstream, volume_downstream, q, pmin_mc, delta, species);
    } else if (kind == CROMA_HIP_DSA_REACC_KIND_ADIABATIC) {
        sd = (double *)malloc((size_t)g->n * sizeof(double));
        if (!sd) {
            free(g);
            return CROMA_HIP_ENOMEM;
        }
        *edens = ra_adiabatic(p, dp, state, g->n, volume_upstream, volume_downstream, compression_ratio, pmin_mc,
                              delta, sd, species);
    } else {
        /* energy density of a given delta row (delta is the input here) */
        *edens = ra_delta_edens(p, dp, state, g->n, pmin_mc, species);
    }
    free(sd);
    free(g);
    return 0;
}

/* Reacceleration delta rows of one tracer from its pre-shock state
 * (tracer_fp_dsa.c:322-418): dpr / der [n] and the scale the driver applies
 * to them; *energy is the reacceleration energy [erg]. */
static double ds_reacc(const DsJob *J, long i, double mach, double temp, double rho1, double u1, double ratio,
                       double q, double ppm, double pem, double mass_cgs, double vu, double vd, double *dpr,
                       double *der, double *adp, double *ade, double *sd, double *energy)
{
    const ChGrid *g = J->g;
    const CromaHipDsaParams *P = J->P;
    const int n = g->n;
    const double *xp = J->crp + (size_t)i * n, *xe = J->cre + (size_t)i * n;
    double scale = 0.0;
    int j;
    *energy = 0.0;
    for (j = 0; j < n; j++) dpr[j] = der[j] = 0.0;
    if (P->reacc_mode == CROMA_HIP_DSA_REACC_POSITIVE_DELTA) {
        /* extra efficiency from the CR pressure already present */
        double x_cr = 0.0, extra, budget;
        {
            const double pcr = ra_state_edens(g->pp, g->pdp, xp, n, vd, P->xcr_pmin_pmc, 0) / 3.0;
            const double pth = rho1 * RA_KB * temp / (P->mu_mol * RA_MP);
            if (pth > 0.0 && isfinite(pcr) && isfinite(pth) && pcr > 0.0) x_cr = pcr / pth;
        }
        extra = croma_hip_dsa_eta_interpolated(mach, P->eta_model, P->eta_model_reacc, x_cr) -
                croma_hip_dsa_eta(mach, P->eta_model);
        if (extra < 0.0 || !isfinite(extra)) extra = 0.0;
        budget = extra * 0.5 * mass_cgs * u1 * u1;
        if (budget > 0.0) {
            const double ep = ra_positive_delta(g->pp, g->pdp, xp, n, vd, q, ppm, dpr, 0);
            const double ee = ra_positive_delta(g->ep, g->edp, xe, n, vd, q, pem, der, 1);
            *energy = (ep + ee) * vd;
            if (*energy > 0.0 && isfinite(*energy)) {
                scale = (*energy > budget) ? budget / *energy : 1.0;
            } else {
                for (j = 0; j < n; j++) dpr[j] = der[j] = 0.0;
            }
        }
    } else if (P->reacc_mode == CROMA_HIP_DSA_REACC_DRURY83) {
        double cp, ce;
        ra_drury(g->pp, g->pdp, xp, n, vu, vd, q, ppm, dpr, 0);
        ra_drury(g->ep, g->edp, xe, n, vu, vd, q, pem, der, 1);
        cp = ra_delta_edens(g->pp, g->pdp, dpr, n, ppm, 0);
        ce = ra_delta_edens(g->ep, g->edp, der, n, pem, 1);
        *energy = (cp + ce) * vd;
        scale = 1.0;
        if (P->reacc_eta_cap > 0.0 && isfinite(P->reacc_eta_cap)) {
            /* the part beyond adiabatic compression is capped at
             * reacc_eta_cap of the upstream kinetic energy */
            const double budget = P->reacc_eta_cap * 0.5 * mass_cgs * u1 * u1;
            const double ap = ra_adiabatic(g->pp, g->pdp, xp, n, vu, vd, ratio, ppm, adp, sd, 0);
            const double ae = ra_adiabatic(g->ep, g->edp, xe, n, vu, vd, ratio, pem, ade, sd, 1);
            const double nonad = (cp + ce - ap - ae) * vd;
            if (nonad > 0.0 && isfinite(nonad) && budget > 0.0 && isfinite(budget)) {
                if (nonad > budget) scale = budget / nonad;
                for (j = 0; j < n; j++) dpr[j] = adp[j] + scale * (dpr[j] - adp[j]);
                for (j = 0; j < n; j++) der[j] = ade[j] + scale * (der[j] - ade[j]);
                *energy = (ap + ae + scale * (cp + ce - ap - ae)) * vd;
            }
        }
        if (J->dis) J->dis[i] = 1;
    }
    return scale;
}

static void ds_range(void *arg, long i0, long i1)
{
    DsJob *J = (DsJob *)arg;
    const ChGrid *g = J->g;
    const CromaHipDsaParams *P = J->P;
    const int n = g->n;
    const int inj_p = P->inject_species != CROMA_HIP_DSA_SPECIES_ELECTRON;
    const int inj_e = P->inject_species != CROMA_HIP_DSA_SPECIES_PROTON;
    const int reacc = P->reacc_mode == CROMA_HIP_DSA_REACC_POSITIVE_DELTA || P->reacc_mode == CROMA_HIP_DSA_REACC_DRURY83;
    double *sp = (double *)malloc(7 * (size_t)n * sizeof(double)), *se, *dpr, *der, *adp, *ade, *sd;
    long i, cnt = 0;
    if (!sp) return;
    se = sp + n;
    dpr = se + n;
    der = dpr + n;
    adp = der + n;
    ade = adp + n;
    sd = ade + n;
    for (i = i0; i < i1; i++) {
        double mach, temp, rho1, u1, ratio, mass_cgs, vd, vu, q, ppm, pem, rscale = 0.0, renergy = 0.0;
        int fl, j, applied = 0;
        if (!ds_resolve(J, i, &mach, &temp, &rho1, &u1)) continue;
        ratio = ds_ratio(mach, P->gamma_gas);
        if (!(ratio > 0.0) || !isfinite(ratio)) continue;
        q = ds_slope(mach, P->gamma_gas);
        ppm = ds_pinj(temp, P->chi_p, DS_MP);
        pem = ds_pinj(temp, P->chi_e > 0.0 ? P->chi_e : P->chi_p, DS_ME);
        if (!(q > 0.0) || !isfinite(q) || !(ppm > 0.0) || !isfinite(ppm) || !(pem > 0.0) || !isfinite(pem))
            continue;
        fl = ds_shock(g, P, mach, temp, rho1, u1, sp, se);
        mass_cgs = J->mass[i] * DS_MSUN;
        vu = mass_cgs / rho1;
        vd = mass_cgs / (ratio * rho1);
        if (!(vu > 0.0) || !isfinite(vu) || !(vd > 0.0) || !isfinite(vd)) continue;
        /* the reacceleration reads the state before this snapshot's
         * injection is added */
        if (reacc) rscale = ds_reacc(J, i, mach, temp, rho1, u1, ratio, q, ppm, pem, mass_cgs, vu, vd, dpr, der, adp,
                                     ade, sd, &renergy);
        if (J->mode == CROMA_HIP_DSA_STATE) {
            if (inj_p && (fl & 1)) {
                for (j = 0; j < n; j++) J->crp[i * n + j] += sp[j] * vd;
                applied = 1;
            }
            if (inj_e && (fl & 3) == 3) {
                for (j = 0; j < n; j++) J->cre[i * n + j] += se[j] * vd;
                applied = 1;
            }
        } else {
            if (inj_p && (fl & 1)) {
                for (j = 0; j < n; j++) J->qpi[i * n + j] += sp[j] * vd * J->inv_dt;
                applied = 1;
            }
            if (inj_e && (fl & 3) == 3) {
                for (j = 0; j < n; j++) J->qepri[i * n + j] += se[j] * vd * J->inv_dt;
                applied = 1;
            }
        }
        if (reacc) {
            /* the Drury83 deltas carry their cap already */
            const double as = (P->reacc_mode == CROMA_HIP_DSA_REACC_DRURY83) ? 1.0 : rscale;
            if (J->mode == CROMA_HIP_DSA_STATE) {
                for (j = 0; j < n; j++) J->crp[i * n + j] += dpr[j] * vd * as;
                for (j = 0; j < n; j++) J->cre[i * n + j] += der[j] * vd * as;
            } else {
                for (j = 0; j < n; j++) J->qpi[i * n + j] += dpr[j] * vd * as * J->inv_dt;
                for (j = 0; j < n; j++) J->qepri[i * n + j] += der[j] * vd * as * J->inv_dt;
            }
            if (rscale > 0.0 && isfinite(renergy) && fabs(renergy) > 0.0) applied = 1;
        }
        cnt += applied;
    }
    free(sp);
    __atomic_fetch_add(&J->ninj[0], (int)cnt, __ATOMIC_RELAXED);
}

int croma_hip_dsa_inject_reacc(const CromaHipGridSpec *grid, const CromaHipDsaParams *p, int mode, int ntracer,
                               double dt_gyr, const double *mach, const double *temp1_K, const double *rho1_cgs,
                               const double *u1_cgs, const double *kinetic_flux, const double *side_code,
                               const double *tracer_mass, double *crp, double *cre, double *qpi, double *qepri,
                               unsigned char *disable_adiabatic, int *ninjected)
{
    ChGrid *g = 0;
    DsJob J;
    int rc;
    if (ninjected) *ninjected = 0;
    if (!p || ntracer < 0 || !mach || !rho1_cgs || !tracer_mass || (!kinetic_flux && !temp1_K)) return CROMA_HIP_EINVAL;
    if (p->reacc_mode < CROMA_HIP_DSA_REACC_OFF || p->reacc_mode > CROMA_HIP_DSA_REACC_DRURY83 ||
        p->eta_model < 0 || p->eta_model >= 8 || p->eta_model_reacc < 0 || p->eta_model_reacc >= 8)
        return CROMA_HIP_EINVAL;
    /* state mode and any reacceleration read or write the state rows
     * (tracer_fp_dsa.c:214-221) */
    if ((mode == CROMA_HIP_DSA_STATE || p->reacc_mode != CROMA_HIP_DSA_REACC_OFF) && (!crp || !cre))
        return CROMA_HIP_EINVAL;
    if (mode == CROMA_HIP_DSA_SOURCE && (!qpi || !qepri || !(dt_gyr > 0.0) || !isfinite(dt_gyr)))
        return CROMA_HIP_EINVAL;
    if (mode != CROMA_HIP_DSA_STATE && mode != CROMA_HIP_DSA_SOURCE) return CROMA_HIP_EINVAL;
    if (ntracer == 0) return 0;
    if ((rc = ds_grid(grid, &g)) != 0) return rc;
    memset(&J, 0, sizeof(J));
    J.g = g;
    J.P = p;
    J.mode = mode;
    J.inv_dt = (mode == CROMA_HIP_DSA_SOURCE) ? 1.0 / dt_gyr : 0.0;
    J.mach = mach;
    J.temp1 = temp1_K;
    J.rho1 = rho1_cgs;
    J.u1 = u1_cgs;
    J.flux = kinetic_flux;
    J.side = side_code;
    J.mass = tracer_mass;
    J.crp = crp;
    J.cre = cre;
    J.qpi = qpi;
    J.qepri = qepri;
    J.dis = disable_adiabatic;
    ch_parallel_for(ntracer, 16, ds_range, &J);
    if (ninjected) *ninjected = J.ninj[0];
    free(g);
    return 0;
}

int croma_hip_dsa_inject(const CromaHipGridSpec *grid, const CromaHipDsaParams *p, int mode, int ntracer,
                         double dt_gyr, const double *mach, const double *temp1_K, const double *rho1_cgs,
                         const double *u1_cgs, const double *kinetic_flux, const double *side_code,
                         const double *tracer_mass, double *crp, double *cre, double *qpi, double *qepri,
                         int *ninjected)
{
    return croma_hip_dsa_inject_reacc(grid, p, mode, ntracer, dt_gyr, mach, temp1_K, rho1_cgs, u1_cgs, kinetic_flux,
                                      side_code, tracer_mass, crp, cre, qpi, qepri, 0, ninjected);
}
