/*
 * croma_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for CROMA's
 * per-tracer Fokker-Planck update.  Every function restates the reference
 * CPU path operation for operation (same evaluation order, FP64, no FMA
 * contraction -- build with -ffp-contract=off) so it can be compared
 * bit-for-bit with the partially compiled reference in oracle/_ref.
 *
 * Pinning (how each part is checked; see tests/test_oracle_*.py):
 *   grid (momentum_bin)            compiled reference, bit-exact; closed form
 *   electron losses                compiled reference (prepare_cre_losses_1d)
 *   b_Coulomb_p/e, b_synch         compiled reference (FP_Coef.c)
 *   t_acc ASA / TTD                compiled reference (accelerationtime_*)
 *   IC cooling row                 compiled reference (prepare_ic_cooling_row)
 *   nsub endpoints / estimate      compiled reference (tracer_fp_nsub.c)
 *   Thomas / PCR                   compiled reference + dense LU brute force
 *   secondary source contraction   compiled reference (given weights)
 *   IC kernel + IC emission        compiled reference
 *   synch table + emission         compiled reference (given the F(x) table)
 *   face weights                   Bernoulli identities W+ - W- = w, W(0)=1,
 *                                  series/exact continuity, 1/(e^w-1) limits
 *   CC assembly                    exact discrete Boltzmann equilibrium of the
 *                                  Chang-Cooper scheme, column sums equal
 *                                  dx_j(1+dt/T_j), M-matrix signs
 *   D_pp fill                      off floor p^2/4e5 closed form, p^2 scaling,
 *                                  eta-cap energy ratio
 *   evolve / snapshot              steady state L x* = Q, conservation under
 *                                  pure drift, zero stays zero, batch/thread
 *                                  invariance
 *   pp cross section, sigma_pp     parity unpinned (HADRONIC.c needs GSL
 *   tables, QeKernel_BB05,         headers absent from this image): checked
 *   gamma / neutrino kernels       only for continuity, sign and thresholds
 *   F(x) table                     parity unpinned vs GSL; checked against the
 *                                  x->0 and x->inf asymptotes
 */
#include "croma_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Physical constants: CONSTANTS.c:3-27. */
static const double K_c = 2.999e+10;
static const double K_GeV = 1.6e-3;
static const double K_eV = 1.6e-12;
static const double K_Msun = 1.99e+33;
static const double K_Mpc = 3.0857e+24;
static const double K_kpc = 3.0857e+21;
static const double K_Gyr = 3.1536e+16;
static const double K_planck = 6.63e-27;
static const double K_mp = 0.938;
static const double K_me = 0.511e-3;
static const double K_mb = 1.0e-27;
static const double K_mpi0 = 1.35e-1;
static const double K_mpic = 1.40e-1;
static const double K_mmu = 1.05e-1;
static const double K_mu_mol = 0.59;
static const double K_kB = 8.62e-5;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct OrcCtx {
    int n;
    /* proton grid */
    double *pp, *pdp, *pE, *beta_p;
    double ppm1, ppp1;
    /* electron grid */
    double *ep, *edp, *eE;
    double epm1, epp1;
    /* caches (FP_Coef.c:29-58) */
    double *p_p2, *p_expcut, *p_sigma, *p_sigmoid;
    double *e_p2, *e_expcut;
    double p_pm1_p2, p_pp1_p2, p_pm1_expcut, p_pp1_expcut;
    double e_pm1_p2, e_pp1_p2, e_pm1_expcut, e_pp1_expcut;
    int j_pp;
};

void orc_default_params(OrcParams *prm)
{
    if (prm == 0) return;
    memset(prm, 0, sizeof(*prm));
    prm->dpp_model = ORC_DPP_OFF;
    prm->ttd_tacc_model = 1;
    prm->reacc_window_mode = 0;
    prm->tracer_nsub_max = 0;
    prm->coeff_interp_min_steps = 4;
    prm->coeff_interp_max_segments = 16;
    prm->psi = 0.5;
    prm->mach_limit = 0.5;
    prm->eta_dpp_cap = 0.0;
    prm->t_acc_direct_gyr = 0.0;
    prm->t_off_gyr = 0.0;
    prm->t_acc_off_gyr = 3.0;
    prm->f_eddy = 2.0;
    prm->tracer_nsub_safety = 0.25;
    prm->epmax = 1.0e7;
}

/* ---------------------------------------------------------------------- */
/* pp cross section, Kamae et al. 2006 + Blattnig 2000 (HADRONIC.c:342-431) */
double orc_pp_pion_cross_section(int kind, double p)
{
    double Ep, Tp, x, x2, x3, nondiff = 0.0, diff = 0.0, Delta = 0.0, res = 0.0;
    double tot, pi0, piplus, piminus, sig = 0.0;
    const double a0 = 0.1176, a1 = 0.3829, a2 = 23.10, a3 = 6.454, a4 = -5.764,
                 a5 = -23.63, a6 = 94.75, a7 = 0.02667;
    const double b0 = 11.34, b1 = 23.72;
    const double c0 = 28.5, c1 = -6.133, c2 = 1.464;
    const double d0 = 0.3522, d1 = 0.1530, d2 = 1.498, d3 = 2.0, d4 = 30.0, d5 = 3.155,
                 d6 = 1.042;
    const double e0 = 5.922, e1 = 1.632;
    const double f0 = 0.0834, f1 = 9.5, f2 = 5.5, f3 = 1.68, f4 = 3134.0;
    const double g0 = 0.0004257, g1 = 4.5, g2 = 7.0, g3 = 2.1, g4 = 503.5;

    if (abs(kind) > 2) return 0.0;
    Ep = sqrt(K_mp * K_mp + p * p);
    Tp = Ep - K_mp;
    x = log10(p);
    x2 = x * x;
    x3 = x2 * x;

    if (p < 1.0) {
        nondiff = 0.0;
    } else if (p >= 1.0 && p < 1.3) {
        nondiff = 0.57 * pow(x / a0, 1.2) *
                  (a2 + a3 * x2 + a4 * x3 + a5 * exp(-a6 * pow(x + a7, 2.0)));
    } else if (p >= 1.3 && p < 2.4) {
        nondiff = (b0 * fabs(a1 - x) + b1 * fabs(a0 - x)) / (a1 - a0);
    } else if (p >= 2.4 && p < 10.0) {
        nondiff = a2 + a3 * x2 + a4 * x3 + a5 * exp(-a6 * pow(x + a7, 2.0));
    } else {
        nondiff = c0 + c1 * x + c2 * x2;
    }

    if (p < 2.25) {
        diff = 0.0;
    } else if (p >= 2.25 && p < 3.2) {
        diff = sqrt(fabs(x - d0) / d1) * (d2 + d3 * log10(d4 * (x - 0.25)) + d5 * x2 - d6 * x3);
    } else if (p >= 3.2 && p < 100.0) {
        diff = d2 + d3 * log10(d4 * (x - 0.25)) + d5 * x2 - d6 * x3;
    } else {
        diff = e0 + e1 * x;
    }

    if (Ep < 1.4) {
        Delta = 0.0;
    } else if (Ep >= 1.4 && Ep < 1.6) {
        Delta = f0 * pow(Ep, 10.0);
    } else if (Ep >= 1.6 && Ep < 1.8) {
        Delta = f1 * exp(-f2 * pow(Ep - f3, 2.0));
    } else if (Ep >= 1.8 && Ep < 10.0) {
        Delta = f4 * pow(Ep, -10.0);
    } else {
        Delta = 0.0;
    }

    if (Ep < 1.6) {
        res = 0.0;
    } else if (Ep >= 1.6 && Ep < 1.9) {
        res = g0 * pow(Ep, 14.0);
    } else if (Ep >= 1.9 && Ep < 2.3) {
        res = g1 * exp(-g2 * pow(Ep - g3, 2.0));
    } else if (Ep >= 2.3 && Ep < 20.0) {
        res = g4 * pow(Ep, -6.0);
    } else {
        res = 0.0;
    }

    tot = (nondiff + diff + Delta + res) * K_mb;
    pi0 = 1.0 / (0.007 + 0.1 * log(Tp) / Tp + 0.3 / Tp / Tp);
    piplus = 1.0 / (0.00717 + 0.0652 * log(Tp) / Tp + 0.162 / Tp / Tp);
    piminus = 1.0 / (0.00456 + 0.0846 / pow(Tp, 0.5) + 0.577 / pow(Tp, 1.5));

    if (kind == 0) sig = tot;
    else if (kind == 1) sig = tot * piplus / pi0;
    else if (kind == -1) sig = tot * piminus / pi0;
    else if (kind == 2) sig = tot * (1.0 + piplus / pi0 + piminus / pi0);
    return sig;
}

/* ---------------------------------------------------------------------- */
/* Grids: momentum_bin (FP_Coef.c:291-330); caches FP_Coef.c:77-154.       */
static void orc_build_grid(int n, double lo, double hi, double mass, double *p,
                           double *dp, double *E, double *pm1, double *pp1)
{
    double a, da;
    int j;

    da = (hi - lo) / (double)n;
    a = lo - da;
    *pm1 = pow(10.0, a);
    for (j = 0; j < n; j++) {
        a = a + da;
        p[j] = pow(10.0, a);
        E[j] = mass * sqrt(1.0 + p[j] * p[j]);
    }
    a = a + da;
    *pp1 = pow(10.0, a);
    for (j = 1; j < n - 1; j++) dp[j] = 0.5 * (p[j + 1] - p[j - 1]);
    dp[0] = 0.5 * (p[1] - *pm1);
    /* the last width reaches two bins past the grid (FP_Coef.c:306, 327) */
    dp[n - 1] = 0.5 * (pow(10.0, a + da) - p[n - 2]);
}

OrcCtx *orc_ctx_create(int n, double pmin, double pmax, double pemin, double pemax)
{
    OrcCtx *ctx;
    int j, j_pp = 0;
    const double k_sigmoid = 5.0;
    double p_pp;

    if (n < 4) return 0;
    ctx = (OrcCtx *)calloc(1, sizeof(*ctx));
    if (!ctx) return 0;
    ctx->n = n;
#define ORC_ALLOC(f) ctx->f = (double *)calloc((size_t)n, sizeof(double))
    ORC_ALLOC(pp); ORC_ALLOC(pdp); ORC_ALLOC(pE); ORC_ALLOC(beta_p);
    ORC_ALLOC(ep); ORC_ALLOC(edp); ORC_ALLOC(eE);
    ORC_ALLOC(p_p2); ORC_ALLOC(p_expcut); ORC_ALLOC(p_sigma); ORC_ALLOC(p_sigmoid);
    ORC_ALLOC(e_p2); ORC_ALLOC(e_expcut);
#undef ORC_ALLOC

    orc_build_grid(n, pmin, pmax, K_mp, ctx->pp, ctx->pdp, ctx->pE, &ctx->ppm1, &ctx->ppp1);
    for (j = 0; j < n; j++) ctx->beta_p[j] = ctx->pp[j] / sqrt(1.0 + ctx->pp[j] * ctx->pp[j]);
    orc_build_grid(n, pemin, pemax, K_me, ctx->ep, ctx->edp, ctx->eE, &ctx->epm1, &ctx->epp1);

    for (j = 0; j < n; j++) {
        ctx->p_p2[j] = ctx->pp[j] * ctx->pp[j];
        ctx->p_expcut[j] = exp(-1.0 / ctx->pp[j]);
        ctx->p_sigma[j] = orc_pp_pion_cross_section(2, K_mp * ctx->pp[j]);
    }
    /* j_pp scan stops at n-100 and re-arms while j_pp == 0 (FP_Coef.c:99-102) */
    for (j = 0; j < n - 100; j++) {
        if (j_pp == 0 && ctx->p_sigma[j] > 0.0) j_pp = j;
    }
    j_pp += 2;
    if (j_pp < 0) j_pp = 0;
    if (j_pp >= n) j_pp = n - 1;
    ctx->j_pp = j_pp;
    p_pp = ctx->pp[j_pp];
    for (j = 0; j < n; j++)
        ctx->p_sigmoid[j] = 1.0 / (1.0 + exp(-k_sigmoid * (ctx->pp[j] - p_pp)));
    ctx->p_pm1_p2 = ctx->ppm1 * ctx->ppm1;
    ctx->p_pp1_p2 = ctx->ppp1 * ctx->ppp1;
    ctx->p_pm1_expcut = exp(-1.0 / ctx->ppm1);
    ctx->p_pp1_expcut = exp(-1.0 / ctx->ppp1);

    for (j = 0; j < n; j++) {
        ctx->e_p2[j] = ctx->ep[j] * ctx->ep[j];
        ctx->e_expcut[j] = exp(-1.0 / ctx->ep[j]);
    }
    ctx->e_pm1_p2 = ctx->epm1 * ctx->epm1;
    ctx->e_pp1_p2 = ctx->epp1 * ctx->epp1;
    ctx->e_pm1_expcut = exp(-1.0 / ctx->epm1);
    ctx->e_pp1_expcut = exp(-1.0 / ctx->epp1);
    return ctx;
}

void orc_ctx_free(OrcCtx *ctx)
{
    if (!ctx) return;
    free(ctx->pp); free(ctx->pdp); free(ctx->pE); free(ctx->beta_p);
    free(ctx->ep); free(ctx->edp); free(ctx->eE);
    free(ctx->p_p2); free(ctx->p_expcut); free(ctx->p_sigma); free(ctx->p_sigmoid);
    free(ctx->e_p2); free(ctx->e_expcut);
    free(ctx);
}

int orc_ctx_n(const OrcCtx *ctx) { return ctx ? ctx->n : 0; }

void orc_ctx_grid(const OrcCtx *ctx, int species, double *p, double *dp, double *E,
                  double *pm1, double *pp1)
{
    const size_t bytes = (size_t)ctx->n * sizeof(double);
    if (species == 0) {
        if (p) memcpy(p, ctx->pp, bytes);
        if (dp) memcpy(dp, ctx->pdp, bytes);
        if (E) memcpy(E, ctx->pE, bytes);
        if (pm1) *pm1 = ctx->ppm1;
        if (pp1) *pp1 = ctx->ppp1;
    } else {
        if (p) memcpy(p, ctx->ep, bytes);
        if (dp) memcpy(dp, ctx->edp, bytes);
        if (E) memcpy(E, ctx->eE, bytes);
        if (pm1) *pm1 = ctx->epm1;
        if (pp1) *pp1 = ctx->epp1;
    }
}

void orc_ctx_beta_p(const OrcCtx *ctx, double *beta_p)
{
    memcpy(beta_p, ctx->beta_p, (size_t)ctx->n * sizeof(double));
}

int orc_ctx_j_pp(const OrcCtx *ctx) { return ctx->j_pp; }

void orc_ctx_sigma_pp(const OrcCtx *ctx, double *sigma, double *sigmoid)
{
    if (sigma) memcpy(sigma, ctx->p_sigma, (size_t)ctx->n * sizeof(double));
    if (sigmoid) memcpy(sigmoid, ctx->p_sigmoid, (size_t)ctx->n * sizeof(double));
}

/* ---------------------------------------------------------------------- */
/* Loss rates (FP_Coef.c:438-469, fp_shared_core.c:729-842)                */
static double orc_coulomb_kernel(double x)
{
    return 0.5 * sqrt(M_PI) * erf(sqrt(x)) - sqrt(x) * exp(-x);
}

double orc_b_coulomb_p(double nicm, double p, double kBT)
{
    const double p2 = p * p;
    const double gamma = sqrt(1.0 + p2);
    const double betap2 = p2 / (1.0 + p2);
    const double kinetic = gamma - 1.0;
    const double Ep = gamma * K_mp - K_mp;
    double xp = Ep / (kBT), xe = xp * (K_me / K_mp);
    double Sum = orc_coulomb_kernel(xe) + orc_coulomb_kernel(xp) * (K_me / K_mp);
    double thermal = 1.0 - kBT / (2.0 * K_mp * kinetic);
    if (thermal < 0.0) return 0.0;
    return K_Gyr * K_c / (K_mp * K_GeV) * Sum * nicm * (1.0 / betap2) * thermal * 3.5e-29;
}

double orc_b_synch(double p, double B)
{
    const double b_micro = B * 1.0e+6;
    const double b_ratio = b_micro / 3.2;
    return 4.8e-4 * p * p * b_ratio * b_ratio * 9.11e-28 * 3.0e+10 * 3.1536e+16;
}

double orc_b_coulomb_e(double nicm, double pe)
{
    const double pe2 = pe * pe;
    double beta2 = pe2 / (1.0 + pe2);
    double A = 3.05e-29;
    return A * nicm * (1.0 + log(pe / nicm) / 74.8) / (K_me * K_GeV) * K_c * K_Gyr / beta2;
}

void orc_losses_e(const OrcCtx *ctx, double n_gas, double b_field, double divv,
                  const double *rad_ic, double rad_ic_m1, double rad_ic_p1,
                  double *radp, double *tloss, double *radpm1, double *radpp1)
{
    const int n = ctx->n;
    const double b_micro = b_field * 1.0e+6;
    const double b_ratio = b_micro / 3.2;
    const double synch_coeff = 4.8e-4 * b_ratio * b_ratio * 9.11e-28 * 3.0e+10 * 3.1536e+16;
    const double coulomb_coeff = 3.05e-29 * n_gas * K_c * K_Gyr / (K_me * K_GeV);
    const double adiabatic_prefactor = 1.0 / 3.0 * divv;
    int j;

    for (j = 0; j < n; j++) {
        const double pe = ctx->ep[j];
        const double pe2 = ctx->e_p2[j];
        const double beta2 = pe2 / (1.0 + pe2);
        const double synch = synch_coeff * pe2;
        const double coulomb = coulomb_coeff * (1.0 + log(pe / n_gas) / 74.8) / beta2;
        radp[j] = synch + rad_ic[j] + coulomb + adiabatic_prefactor * pe;
        tloss[j] = 1.0e+100;
    }
    {
        const double beta2_m1 = ctx->e_pm1_p2 / (1.0 + ctx->e_pm1_p2);
        const double beta2_p1 = ctx->e_pp1_p2 / (1.0 + ctx->e_pp1_p2);
        const double coulomb_m1 = coulomb_coeff * (1.0 + log(ctx->epm1 / n_gas) / 74.8) / beta2_m1;
        const double coulomb_p1 = coulomb_coeff * (1.0 + log(ctx->epp1 / n_gas) / 74.8) / beta2_p1;
        *radpm1 = synch_coeff * ctx->e_pm1_p2 + rad_ic_m1 + coulomb_m1;
        *radpp1 = synch_coeff * ctx->e_pp1_p2 + rad_ic_p1 + coulomb_p1;
    }
}

void orc_losses_p(const OrcCtx *ctx, double n_gas, double kbt, double divv,
                  double *radp, double *tloss, double *radpm1, double *radpp1)
{
    const int n = ctx->n;
    const double kappa_pp = 0.5;
    const double tloss_prefactor = 1.0 / (kappa_pp * K_c * n_gas) / K_Gyr;
    const double adiabatic_prefactor = 1.0 / 3.0 * divv;
    const int j_pp = ctx->j_pp;
    int j;

    for (j = 0; j < n; j++) {
        double sig_pp;
        const double p = ctx->pp[j];
        radp[j] = orc_b_coulomb_p(n_gas, p, kbt) + adiabatic_prefactor * p;
        sig_pp = (j > j_pp) ? ctx->p_sigma[j] : ctx->p_sigma[j_pp];
        sig_pp *= ctx->p_sigmoid[j];
        tloss[j] = tloss_prefactor / sig_pp;
    }
    *radpm1 = orc_b_coulomb_p(n_gas, ctx->ppm1, kbt);
    *radpp1 = orc_b_coulomb_p(n_gas, ctx->ppp1, kbt);
}

/* ---------------------------------------------------------------------- */
/* Momentum diffusion (FP_Coef.c:636-849; fp_shared_core.c:974-1081, 1241-1317) */
static double orc_I_theta(double x)
{
    return pow(x, 4.0) / 4.0 + pow(x, 2.0) - (1.0 + 2.0 * x * x) * log(x) - 5.0 / 4.0;
}

double orc_tacc_ttd(double L, double cs, double Ms, double v_turb_com, int ttd_model)
{
    double x = cs / K_c;
    if (ttd_model == 1) {
        const double t_norm_gyr = 2.5e-3;
        const double l_norm_cm = 300.0 * K_kpc;
        const double cs_norm = 1500.0e5;
        double theta = orc_I_theta(x);
        if (!(x > 0.0) || !(theta > 0.0) || !(Ms > 0.0) || !(L > 0.0) || !(cs > 0.0))
            return 0.0;
        return t_norm_gyr / (x * theta) * pow(Ms / 0.5, -4.0) * (L / l_norm_cm) *
               pow(cs / cs_norm, -1.0);
    }
    {
        double kL = 2.0 * M_PI / L;
        double k_cut_th = 1.04e+4 * pow(Ms, 4.0) * kL;
        double k_cut = k_cut_th;
        double kWk_L = 0.5 * pow(v_turb_com, 2.0);
        double kWk_cut = kWk_L * pow(k_cut / kL, -0.5);
        return K_c / M_PI / orc_I_theta(x) / k_cut / kWk_cut / K_Gyr;
    }
}

double orc_tacc_asa(double L, double cs, double Ms, double beta_pl, double psi)
{
    return sqrt(6.0 / 5.0) / 12.0 * (K_c / cs / cs) * L / sqrt(beta_pl) / pow(Ms, 3.0) *
           pow(psi, 3.0) / K_Gyr;
}

static double orc_effective_mach(double dv, double cs, double mach_limit)
{
    double mach;
    if (cs <= 0.0) return 0.0;
    mach = dv / cs;
    if (mach < 0.0) mach = 0.0;
    if (mach_limit > 0.0 && mach > mach_limit) mach = mach_limit;
    return mach;
}

static void orc_dpp_off(const OrcCtx *ctx, double *dpp, double *dppe, double *edges)
{
    const int n = ctx->n;
    const double t_acc = 1.0e+5;
    const double inv_4_tacc = 1.0 / (4.0 * t_acc);
    int j;
    for (j = 0; j < n; j++) dpp[j] = ctx->p_p2[j] * inv_4_tacc;
    edges[1] = ctx->p_pp1_p2 * inv_4_tacc;
    edges[0] = ctx->p_pm1_p2 * inv_4_tacc;
    for (j = 0; j < n; j++) dppe[j] = ctx->e_p2[j] * inv_4_tacc;
    edges[3] = ctx->e_pp1_p2 * inv_4_tacc;
    edges[2] = ctx->e_pm1_p2 * inv_4_tacc;
}

static void orc_dpp_fill(const OrcCtx *ctx, double inv, double epmax, int with_cut,
                         double *dpp, double *dppe, double *edges)
{
    /* with_cut: the q=2 fill (FP_Coef.c:680-713) and the direct fill (715-748)
     * share the same form; only the prefactor inv differs. */
    const int n = ctx->n;
    int j;
    (void)with_cut;
    for (j = 0; j < n; j++)
        dpp[j] = ctx->p_p2[j] * inv * exp(-ctx->pE[j] / epmax) * ctx->p_expcut[j];
    edges[1] = ctx->p_pp1_p2 * inv * exp(-K_mp * sqrt(1.0 + ctx->ppp1 * ctx->ppp1) / epmax) *
               ctx->p_pp1_expcut;
    edges[0] = ctx->p_pm1_p2 * inv * exp(-K_mp * sqrt(1.0 + ctx->ppm1 * ctx->ppm1) / epmax) *
               ctx->p_pm1_expcut;
    for (j = 0; j < n; j++) dppe[j] = ctx->e_p2[j] * inv * ctx->e_expcut[j];
    edges[3] = ctx->e_pp1_p2 * inv * ctx->e_pp1_expcut;
    edges[2] = ctx->e_pm1_p2 * inv * ctx->e_pm1_expcut;
}

static double orc_tacc_for_cap(const OrcParams *prm, int model, double l_turb_mpc,
                               double dv, double cs, double beta_pl, double t_acc_direct)
{
    double mach;
    if (!(dv > 0.0) || !(l_turb_mpc > 0.0) || !(cs > 0.0)) return 0.0;
    mach = dv / cs;
    if (mach < 0.0) mach = 0.0;
    if (prm->mach_limit > 0.0 && mach > prm->mach_limit) mach = prm->mach_limit;
    if (mach <= 0.0) return 0.0;
    if (model == ORC_DPP_ASA) {
        if (!(beta_pl > 0.0)) return 0.0;
        return orc_tacc_asa(l_turb_mpc * K_Mpc, cs, mach, beta_pl, prm->psi);
    }
    if (model == ORC_DPP_TTD) return orc_tacc_ttd(l_turb_mpc * K_Mpc, cs, mach, dv, prm->ttd_tacc_model);
    if (model == ORC_DPP_DIRECT)
        return (t_acc_direct > 0.0 && isfinite(t_acc_direct)) ? t_acc_direct : 0.0;
    return 0.0;
}

static void orc_dpp_impl(const OrcCtx *ctx, const OrcParams *prm, int model,
                         double l_turb_mpc, double dv, double cs, double beta_pl,
                         double mass, const double *crp_state, const double *cre_state,
                         double t_acc_direct, double *dpp, double *dppe, double *edges)
{
    const int n = ctx->n;
    int j;

    if (!(dv > 0.0) || !isfinite(dv) || model == ORC_DPP_OFF) {
        orc_dpp_off(ctx, dpp, dppe, edges);
    } else if (model == ORC_DPP_ASA) {
        const double mach = orc_effective_mach(dv, cs, prm->mach_limit);
        if (mach <= 0.0 || beta_pl <= 0.0) {
            orc_dpp_off(ctx, dpp, dppe, edges);
        } else {
            const double t_acc = orc_tacc_asa(l_turb_mpc * K_Mpc, cs, mach, beta_pl, prm->psi);
            const double inv_q_tacc = 1.0 / ((2.0 + 2.0) * t_acc);
            orc_dpp_fill(ctx, inv_q_tacc, prm->epmax, 1, dpp, dppe, edges);
        }
    } else if (model == ORC_DPP_TTD) {
        const double mach = orc_effective_mach(dv, cs, prm->mach_limit);
        if (mach <= 0.0 || dv <= 0.0) {
            orc_dpp_off(ctx, dpp, dppe, edges);
        } else {
            const double t_acc = orc_tacc_ttd(l_turb_mpc * K_Mpc, cs, mach, dv, prm->ttd_tacc_model);
            const double inv_q_tacc = 1.0 / ((2.0 + 2.0) * t_acc);
            orc_dpp_fill(ctx, inv_q_tacc, prm->epmax, 1, dpp, dppe, edges);
        }
    } else if (model == ORC_DPP_DIRECT) {
        if (!(t_acc_direct > 0.0) || !isfinite(t_acc_direct)) {
            orc_dpp_off(ctx, dpp, dppe, edges);
        } else {
            const double inv_4_tacc = 1.0 / (4.0 * t_acc_direct);
            orc_dpp_fill(ctx, inv_4_tacc, prm->epmax, 0, dpp, dppe, edges);
        }
    } else {
        orc_dpp_off(ctx, dpp, dppe, edges);
    }

    /* eta cap (fp_shared_core.c:974-1081), only for the on-set model */
    if (model != ORC_DPP_OFF) {
        double t_acc_gyr, ecr = 0.0, q_turb, eta, scale;
        if (!(prm->eta_dpp_cap > 0.0) || !isfinite(prm->eta_dpp_cap)) return;
        if (!(mass > 0.0) || !(l_turb_mpc > 0.0) || !(dv > 0.0)) return;
        if (crp_state == 0 && cre_state == 0) return;
        t_acc_gyr = orc_tacc_for_cap(prm, model, l_turb_mpc, dv, cs, beta_pl, t_acc_direct);
        if (!(t_acc_gyr > 0.0) || !isfinite(t_acc_gyr)) return;
        if (crp_state != 0 && n > 4) {
            for (j = 2; j < n - 2; j++) {
                const double ekin = ctx->pE[j] - K_mp;
                if (ekin > 0.0 && isfinite(crp_state[j]))
                    ecr += ekin * K_GeV * crp_state[j] * ctx->pdp[j];
            }
        }
        if (cre_state != 0 && n > 4) {
            for (j = 2; j < n - 2; j++) {
                const double ekin = ctx->eE[j] - K_me;
                if (ekin > 0.0 && isfinite(cre_state[j]))
                    ecr += ekin * K_GeV * cre_state[j] * ctx->edp[j];
            }
        }
        q_turb = 0.5 * mass * K_Msun * dv * dv * dv / (l_turb_mpc * K_Mpc);
        if (!(ecr > 0.0) || !(q_turb > 0.0) || !isfinite(ecr) || !isfinite(q_turb)) return;
        eta = ecr / (q_turb * t_acc_gyr * K_Gyr);
        if (!(eta > prm->eta_dpp_cap) || !isfinite(eta)) return;
        scale = prm->eta_dpp_cap / eta;
        for (j = 0; j < n; j++) dpp[j] *= scale;
        for (j = 0; j < n; j++) dppe[j] *= scale;
        edges[0] *= scale;
        edges[1] *= scale;
        edges[2] *= scale;
        edges[3] *= scale;
    }
}

void orc_dpp(const OrcCtx *ctx, const OrcParams *prm, int model,
             double l_turb_mpc, double dv, double cs, double beta_pl,
             double mass_msun, const double *crp_state, const double *cre_state,
             double *dpp, double *dppe, double *edges)
{
    orc_dpp_impl(ctx, prm, model, l_turb_mpc, dv, cs, beta_pl, mass_msun, crp_state,
                 cre_state, prm->t_acc_direct_gyr, dpp, dppe, edges);
}

/* ---------------------------------------------------------------------- */
/* Chang-Cooper (Chang_Cooper.c:22-184; fp_shared_core.c:660-704)          */
void orc_face_weights(double w, double *delta, double *wp, double *wm)
{
    *delta = 1.0 / w - 1.0 / (exp(w) - 1.0);
    if (fabs(w) < 0.1) {
        double denom = 1.0 + pow(w, 2.0) / 24.0 + pow(w, 4.0) / 1920.0;
        *wp = exp(0.5 * w) / denom;
        *wm = exp(-0.5 * w) / denom;
    } else {
        double aw = fabs(w);
        double denom = 1.0 - exp(-aw);
        *wp = aw * exp(-0.5 * aw + 0.5 * w) / denom;
        *wm = aw * exp(-0.5 * aw - 0.5 * w) / denom;
    }
}

static void orc_face(double bl, double br, double cl, double cr, double xl, double xr,
                     double *delta, double *wp, double *wm, double *w_out)
{
    double w;
    *delta = 0.0;
    *wp = 0.0;
    *wm = 0.0;
    if (w_out) *w_out = 0.0;
    if (fabs(cl) <= 1.0e-200) return;
    w = 0.5 * (bl + br) / (0.5 * (cl + cr)) * (xr - xl);
    orc_face_weights(w, delta, wp, wm);
    if (w_out) *w_out = w;
}

void orc_cc_build(int n, double dt, const double *x, const double *dx,
                  double xm1, double xp1, const double *radp, const double *dpp,
                  const double *tloss, double radpm1, double radpp1,
                  double dppm1, double dppp1, double *a, double *b, double *c)
{
    const int nm = n - 1;
    double *B = (double *)malloc((size_t)n * sizeof(double));
    double *delta = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *Wp = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *Wm = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *w = (double *)malloc((size_t)(n + 1) * sizeof(double));
    const double *C = dpp;
    const double *T = tloss;
    double bm1, bp1, Wpm1 = 0.0, Wmm1 = 0.0;
    int m;

    for (m = 0; m < n; m++) B[m] = radp[m] - 2.0 / x[m] * dpp[m];
    bm1 = radpm1 - 2.0 / xm1 * dppm1;
    bp1 = radpp1 - 2.0 / xp1 * dppp1;

    for (m = 0; m < nm; m++)
        orc_face(B[m], B[m + 1], C[m], C[m + 1], x[m], x[m + 1], &delta[m], &Wp[m], &Wm[m], &w[m]);
    /* the left ghost face overwrites delta[0] (Chang_Cooper.c:205-208) */
    if (fabs(C[0]) > 1.0e-200) orc_face(bm1, B[0], dppm1, C[0], xm1, x[0], &delta[0], &Wpm1, &Wmm1, 0);
    orc_face(B[nm], bp1, C[nm], dppp1, x[nm], xp1, &delta[nm], &Wp[nm], &Wm[nm], &w[nm]);

    for (m = 1; m < nm; m++) {
        if (fabs(C[m]) <= 1.0e-200) delta[m - 1] = 0.0;
        if (fabs(C[m]) > 1.0e-200) {
            double dxm_12 = x[m] - x[m - 1];
            double dxm12 = x[m + 1] - x[m];
            double cm_12 = 0.5 * (C[m] + C[m - 1]);
            double cm12 = 0.5 * (C[m] + C[m + 1]);
            a[m] = dt / (1.0 * dx[m]) * cm_12 / dxm_12 * Wm[m - 1];
            b[m] = 1.0 + dt / (1.0 * dx[m]) * (cm_12 / dxm_12 * Wp[m - 1] + cm12 / dxm12 * Wm[m]) +
                   dt / T[m];
            c[m] = dt / (1.0 * dx[m]) * cm12 / dxm12 * Wp[m];
        } else {
            double deltam_12 = delta[m - 1];
            double deltam12 = delta[m];
            double bm_12 = 0.5 * (B[m] + B[m - 1]);
            double bm12 = 0.5 * (B[m] + B[m + 1]);
            a[m] = -deltam_12 * bm_12 * dt / (1.0 * dx[m]);
            b[m] = 1.0 + dt / (1.0 * dx[m]) * ((1.0 - deltam_12) * bm_12 - deltam12 * bm12) + dt / T[m];
            c[m] = dt / (1.0 * dx[m]) * (1.0 - delta[m]) * bm12;
        }
    }

    /* left boundary: zero-flux ghost face (Chang_Cooper.c:106-141) */
    if (fabs(C[0]) > 1.0e-200) {
        double c_12 = 0.0;
        a[0] = dt / (1.0 * dx[0]) * c_12 / (x[0] - xm1) * Wmm1;
        b[0] = 1.0 + dt / (1.0 * dx[0]) *
                         (c_12 / (x[0] - xm1) * Wpm1 + 0.5 * (C[0] + C[1]) / (x[1] - x[0]) * Wm[0]) +
               dt / T[0];
        c[0] = dt / (1.0 * dx[0]) * 0.5 * (C[0] + C[1]) / (x[1] - x[0]) * Wp[0];
    } else {
        double b_12 = 0.0;
        a[0] = -delta[0] * b_12 * dt / (1.0 * dx[0]);
        b[0] = 1.0 + dt / (1.0 * dx[0]) * b_12 + dt / T[0];
        c[0] = dt / (1.0 * dx[0]) * (1.0 - delta[0]) * 0.5 * (B[0] + B[1]);
    }
    /* right boundary (Chang_Cooper.c:143-184) */
    if (fabs(C[nm]) > 1.0e-200) {
        a[nm] = dt / (1.0 * dx[nm]) * (0.5 * (C[nm] + C[nm - 1])) / (x[nm] - x[nm - 1]) * Wm[nm - 1];
        b[nm] = 1.0 + dt / (1.0 * dx[nm]) * ((0.5 * (C[nm] + C[nm - 1])) / (x[nm] - x[nm - 1]) * Wp[nm - 1]) +
                dt / T[nm];
        c[nm] = 0.0;
    } else {
        a[nm] = -delta[nm - 1] * 0.5 * (B[nm] + B[nm - 1]) * dt / (1.0 * dx[nm]);
        b[nm] = 1.0 + dt / (1.0 * dx[nm]) * (0.5 * (B[nm - 1] + B[nm])) + dt / T[nm];
        c[nm] = 0.0;
    }
    free(B); free(delta); free(Wp); free(Wm); free(w);
}

/* ---------------------------------------------------------------------- */
/* Tridiagonal solvers (fp_shared_core.c:226-257, 430-547)                 */
static int orc_thomas_ws(int nrow, const double *a, const double *b, const double *c,
                         double *rhs, double *cprime)
{
    int i;
    if (nrow <= 0 || !a || !b || !c || !rhs || !cprime) return -1;
    if (fabs(b[0]) <= 1.0e-300) return -2;
    cprime[0] = -c[0] / b[0];
    rhs[0] /= b[0];
    for (i = 1; i < nrow; i++) {
        double denom = b[i] + a[i] * cprime[i - 1];
        if (fabs(denom) <= 1.0e-300) return -2;
        cprime[i] = (i == nrow - 1) ? 0.0 : -c[i] / denom;
        rhs[i] = (rhs[i] + a[i] * rhs[i - 1]) / denom;
    }
    for (i = nrow - 2; i >= 0; i--) rhs[i] -= cprime[i] * rhs[i + 1];
    return 0;
}

int orc_thomas(int n, const double *a, const double *b, const double *c, double *rhs)
{
    double *cp;
    int ierr;
    if (n <= 0) return -1;
    cp = (double *)malloc((size_t)n * sizeof(double));
    if (!cp) return -3;
    ierr = orc_thomas_ws(n, a, b, c, rhs, cp);
    free(cp);
    return ierr;
}

int orc_pcr_batch(int nsys, int n, const double *ain, const double *bin, const double *cin,
                  const double *rhsin, double *x)
{
    const size_t nall = (size_t)nsys * (size_t)n;
    double *a = (double *)malloc(nall * sizeof(double)), *b = (double *)malloc(nall * sizeof(double));
    double *c = (double *)malloc(nall * sizeof(double)), *r = (double *)malloc(nall * sizeof(double));
    double *a2 = (double *)malloc(nall * sizeof(double)), *b2 = (double *)malloc(nall * sizeof(double));
    double *c2 = (double *)malloc(nall * sizeof(double)), *r2 = (double *)malloc(nall * sizeof(double));
    int stride, s, i, ret = 0;
    size_t k;

    if (n <= 0 || (n & (n - 1)) != 0) { ret = -4; goto done; }
    for (k = 0; k < nall; k++) {
        a[k] = -ain[k];
        b[k] = bin[k];
        c[k] = -cin[k];
        r[k] = rhsin[k];
    }
    for (stride = 1; stride < n; stride <<= 1) {
        for (s = 0; s < nsys; s++) {
            const size_t off = (size_t)s * (size_t)n;
            for (i = 0; i < n; i++) {
                const int left = i - stride, right = i + stride;
                double next_a = 0.0, next_b = b[off + i], next_c = 0.0, next_r = r[off + i];
                if (left >= 0) {
                    const double denom = b[off + left];
                    double ls;
                    if (fabs(denom) <= 1.0e-300) { ret = -2; goto done; }
                    ls = a[off + i] / denom;
                    next_a = -a[off + left] * ls;
                    next_b -= c[off + left] * ls;
                    next_r -= r[off + left] * ls;
                }
                if (right < n) {
                    const double denom = b[off + right];
                    double rs;
                    if (fabs(denom) <= 1.0e-300) { ret = -2; goto done; }
                    rs = c[off + i] / denom;
                    next_c = -c[off + right] * rs;
                    next_b -= a[off + right] * rs;
                    next_r -= r[off + right] * rs;
                }
                a2[off + i] = next_a;
                b2[off + i] = next_b;
                c2[off + i] = next_c;
                r2[off + i] = next_r;
            }
        }
        { double *t; t = a; a = a2; a2 = t; t = b; b = b2; b2 = t; t = c; c = c2; c2 = t; t = r; r = r2; r2 = t; }
    }
    for (k = 0; k < nall; k++) {
        if (fabs(b[k]) <= 1.0e-300) { ret = -2; goto done; }
        x[k] = r[k] / b[k];
    }
done:
    free(a); free(b); free(c); free(r); free(a2); free(b2); free(c2); free(r2);
    return ret;
}

/* ---------------------------------------------------------------------- */
/* Secondary electrons (fp_shared_core.c:1453-1672; HADRONIC.c:433-533)    */
static double orc_lorentz_beta(double gamma)
{
    double beta_squared = gamma * gamma - 1.0;
    if (beta_squared < 0.0) return 0.0;
    return sqrt(beta_squared) / gamma;
}

static double orc_FpiQ(int m, double Epi, double Ep)
{
    double L = log(0.001 * Ep), x = Epi / Ep, mpi = 1.35e-1;
    double xmin, Bpi, alpha, r, xa, A;
    if (m == 1) mpi = 1.40e-1;
    if (m == 0) mpi = 1.35e-1;
    xmin = mpi / Ep;
    if (x <= xmin) return 0.0;
    if (x > 1.0) return 0.0;
    Bpi = 5.58 + 0.78 * L + 0.10 * L * L;
    alpha = 0.89 / (sqrt(Bpi) * (1.0 - exp(-0.33 * Bpi)));
    r = 3.1 / (pow(Bpi, 1.5));
    xa = pow(x, alpha);
    A = pow((1.0 - xa) / pow(1.0 + r * xa, 3.0), 4.0);
    A = A * (1.0 / (1.0 - xa) + 3.0 * r / (1.0 + r * xa)) * sqrt(1.0 - mpi / (x * Ep));
    return 4.0 * alpha * Bpi * xa * A / x / Ep;
}

static double orc_dsigma(int i, double Epi, double Ep)
{
    double gp = Ep / K_mp, gbp = sqrt(gp * gp - 1.0), pp = gbp * K_mp;
    double Fpi = orc_FpiQ(0, Epi, Ep);
    double s = orc_pp_pion_cross_section(i, pp);
    return s * Fpi / sqrt(1.0 + 1.0 / (gbp * gbp));
}

static double orc_Fe(double Ee, double Epi)
{
    double mpi = 1.3957e-1, mmu = 1.05658e-1, bmu = 0.2714;
    double gpi = Epi / mpi;
    double xi = mmu * mmu / (mpi * mpi);
    double Emu = (mpi * mpi - mmu * mmu) / (mpi * mpi) * Epi / (2.0 * bmu);
    double betamu = sqrt(1.0 - mmu * mmu / (Emu * Emu));
    double lambda = 2.0 * mpi * mpi * Ee * bmu / ((mpi * mpi - mmu * mmu) * Epi);
    double P, X, lambda2 = lambda * lambda, lambda3 = lambda2 * lambda;
    P = 1.0 / betamu * (2.0 * Epi * xi / (Emu * (1.0 - xi)) - (1.0 + xi) / (1.0 - xi));
    if (lambda > 1.0) return 0.0;
    if (gpi == 1.0) return 0.0;
    if (lambda < 0.0) return 0.0;
    if ((1.0 - betamu) / (1.0 + betamu) < lambda) {
        X = 5.0 - 9.0 * lambda2 + 4.0 * lambda3 -
            P / betamu * (1.0 - (6.0 * betamu + 3.0) * lambda2 + (6.0 * betamu + 2.0) * lambda3);
        return X / 12.0;
    } else if (lambda < (1.0 - betamu) / (1.0 + betamu)) {
        X = lambda2 * betamu / pow(1.0 - betamu, 2.0) *
                (3.0 - 2.0 / 3.0 * lambda * (3.0 + betamu * betamu) / (1.0 - betamu)) -
            P / (1.0 - betamu) *
                (lambda2 * (1.0 + betamu) - 2.0 * lambda2 / (1.0 - betamu) * (0.5 + lambda * (1.0 + betamu)) +
                 2.0 * lambda3 * (3.0 + betamu * betamu) / (3.0 * pow(1.0 - betamu, 2.0)));
        return X;
    }
    return 0.0;
}

double orc_qe_kernel_bb05(double Ee, double p)
{
    const double bmu = 0.2714;
    double gp = sqrt(p * p + 1.0), Ep = gp * K_mp, Emax = gp * K_mp, Emu, betamu;
    double Integral = 0.0;
    double r = 0.5 * (K_mpic * K_mpic - K_mmu * K_mmu) / (bmu * K_mpic * K_mpic);
    double Emin = Ee / r;
    double a, amin, amax, da = 0.05, f, fpre, Epi, Epipre, dEpi;
    int i, Na;

    if (Emax < Emin || (Ep - K_mp) < 0.289) return 0;
    Emin = fmax(Emin, K_mpic);
    amin = log10(Emin);
    amax = log10(Emax);
    Na = (int)fmax(50, round((amax - amin) / da));
    a = amin;
    Epi = pow(10.0, a);
    Emu = Epi * r;
    betamu = orc_lorentz_beta(Emu / K_mmu);
    if (betamu == 0.0) f = 0.0;
    else f = (orc_dsigma(1, Epi, Ep) + orc_dsigma(-1, Epi, Ep)) * orc_Fe(Ee, Epi) / (betamu * Epi);
    for (i = 1; i < Na; i++) {
        Epipre = Epi;
        a = a + da;
        Epi = pow(10.0, a);
        dEpi = Epi - Epipre;
        fpre = f;
        Emu = Epi * r;
        betamu = orc_lorentz_beta(Emu / K_mmu);
        f = (orc_dsigma(1, Epi, Ep) + orc_dsigma(-1, Epi, Ep)) * orc_Fe(Ee, Epi) / (betamu * Epi);
        Integral = Integral + 0.5 * (f + fpre) * dEpi;
        if (Integral < 0.0 || isnan(Integral)) return 0.0;
    }
    return 8.0 * bmu * K_mpic * K_mpic * (K_c * K_me) * Integral / (K_mpic * K_mpic - K_mmu * K_mmu);
}

void orc_secondary_kernel(const OrcCtx *ctx, double *fqe_t, int *np_min_qe)
{
    const int n = ctx->n;
    double *values = (double *)malloc((size_t)n * sizeof(double));
    int je, jp;
    for (je = 0; je < n; je++) {
        int jp0;
        for (jp = 0; jp < n; jp++) {
            double val = orc_qe_kernel_bb05(ctx->eE[je], ctx->pp[jp]) * ctx->beta_p[jp];
            if (isnan(val)) val = 0.0;
            values[jp] = val;
        }
        np_min_qe[je] = 1;
        for (jp = 1; jp < n; jp++)
            if (values[jp] < 1.0e-80) np_min_qe[je] = jp;
        jp0 = np_min_qe[je];
        if (jp0 < 1) jp0 = 1;
        if (jp0 > n) jp0 = n;
        for (jp = 0; jp < n; jp++) {
            double weight = 0.0;
            const double val = values[jp];
            if (jp0 < n) {
                if (jp == jp0 - 1) weight = 0.5 * val * ctx->pdp[jp0];
                else if (jp >= jp0 && jp < n - 1) weight = 0.5 * val * (ctx->pdp[jp] + ctx->pdp[jp + 1]);
                else if (jp == n - 1 && jp >= jp0) weight = 0.5 * val * ctx->pdp[jp];
            }
            fqe_t[(size_t)jp * (size_t)n + (size_t)je] = weight;
        }
        np_min_qe[je] = (jp0 < n) ? (jp0 - 1) : n;
    }
    free(values);
}

void orc_secondary_source(int n, int ntracer, const double *n_gas, const double *crp,
                          const int *np_min_qe, const double *fqe_t,
                          const double *qepri, double *inje)
{
    int it, je, jp, jp0 = n;
    double *q = (double *)malloc((size_t)n * sizeof(double));
    for (je = 0; je < n; je++) {
        int v = np_min_qe[je];
        if (v < 0) v = 0;
        if (v < jp0) jp0 = v;
    }
    if (jp0 > n) jp0 = n;
    for (it = 0; it < ntracer; it++) {
        const double *cr = crp + (size_t)it * (size_t)n;
        for (je = 0; je < n; je++) q[je] = 0.0;
        for (jp = jp0; jp < n; jp++) {
            const double nn = cr[jp];
            const double *wrow = fqe_t + (size_t)jp * (size_t)n;
            for (je = 0; je < n; je++) q[je] += nn * wrow[je];
        }
        for (je = 0; je < n; je++) {
            const double qpri = qepri ? qepri[(size_t)it * (size_t)n + je] : 0.0;
            inje[(size_t)it * (size_t)n + je] = q[je] * K_Gyr * n_gas[it] + qpri;
        }
    }
    free(q);
}

/* ---------------------------------------------------------------------- */
/* One coefficient segment (tracer_fp_cpu.c:88-196, tracer_fp_coef.c:218-359) */
int orc_evolve(const OrcCtx *ctx, const OrcParams *prm, int ntracer, int nstep,
               double dt, int windowed, int on_start, int on_end, double t_acc_direct,
               const double *n_gas, const double *kbt, const double *b_field,
               const double *divv, const double *l_turb, const double *dv,
               const double *cs, const double *beta_pl,
               const double *rad_ic, const double *rad_ic_m1, const double *rad_ic_p1,
               const double *qpi, const double *qepri,
               const double *fqe_t, const int *np_min_qe,
               const double *mass, const unsigned char *disable_adiabatic,
               double *crp, double *cre)
{
    const int n = ctx->n;
    const size_t nn = (size_t)n;
    double *buf = (double *)calloc(nn * 24, sizeof(double));
    double *prad = buf, *ptl = buf + nn, *erad = buf + 2 * nn, *etl = buf + 3 * nn;
    double *dpp = buf + 4 * nn, *dppe = buf + 5 * nn, *dppo = buf + 6 * nn, *dppeo = buf + 7 * nn;
    double *pa = buf + 8 * nn, *pb = buf + 9 * nn, *pc = buf + 10 * nn;
    double *pao = buf + 11 * nn, *pbo = buf + 12 * nn, *pco = buf + 13 * nn;
    double *ea = buf + 14 * nn, *eb = buf + 15 * nn, *ec = buf + 16 * nn;
    double *eao = buf + 17 * nn, *ebo = buf + 18 * nn, *eco = buf + 19 * nn;
    double *ric = buf + 20 * nn, *inje = buf + 21 * nn, *cp = buf + 22 * nn;
    int it, j, istep, ierr = 0;

    if (!buf) return -1;
    for (it = 0; it < ntracer && ierr == 0; it++) {
        double *xp = crp + (size_t)it * nn;
        double *xe = cre + (size_t)it * nn;
        const double dvl = (disable_adiabatic != 0 && disable_adiabatic[it] != 0) ? 0.0 : divv[it];
        double prm1, prp1, erm1, erp1, ed[4], edo[4];

        for (j = 0; j < n; j++) ric[j] = rad_ic[(size_t)j * (size_t)ntracer + (size_t)it];
        orc_losses_p(ctx, n_gas[it], kbt[it], dvl, prad, ptl, &prm1, &prp1);
        orc_losses_e(ctx, n_gas[it], b_field[it], dvl, ric, rad_ic_m1[it], rad_ic_p1[it],
                     erad, etl, &erm1, &erp1);
        orc_dpp_impl(ctx, prm, prm->dpp_model, l_turb[it], dv[it], cs[it], beta_pl[it],
                     mass[it], xp, xe, t_acc_direct, dpp, dppe, ed);
        orc_dpp_impl(ctx, prm, ORC_DPP_OFF, l_turb[it], dv[it], cs[it], beta_pl[it], 0.0, 0, 0,
                     t_acc_direct, dppo, dppeo, edo);
        orc_cc_build(n, dt, ctx->pp, ctx->pdp, ctx->ppm1, ctx->ppp1, prad, dpp, ptl, prm1, prp1,
                     ed[0], ed[1], pa, pb, pc);
        orc_cc_build(n, dt, ctx->ep, ctx->edp, ctx->epm1, ctx->epp1, erad, dppe, etl, erm1, erp1,
                     ed[2], ed[3], ea, eb, ec);
        orc_cc_build(n, dt, ctx->pp, ctx->pdp, ctx->ppm1, ctx->ppp1, prad, dppo, ptl, prm1, prp1,
                     edo[0], edo[1], pao, pbo, pco);
        orc_cc_build(n, dt, ctx->ep, ctx->edp, ctx->epm1, ctx->epp1, erad, dppeo, etl, erm1, erp1,
                     edo[2], edo[3], eao, ebo, eco);
        /* secondary source refreshed once per segment, from the segment-start
         * proton state (tracer_fp_cpu.c:142-147) */
        orc_secondary_source(n, 1, n_gas + it, xp, np_min_qe, fqe_t,
                             qepri ? qepri + (size_t)it * nn : 0, inje);

        for (istep = 0; istep < nstep; istep++) {
            const int on = !windowed || (istep >= on_start && istep < on_end);
            const double *qp = qpi + (size_t)it * nn;
            for (j = 0; j < n; j++) xp[j] = dt * qp[j] + xp[j];
            ierr = orc_thomas_ws(n, on ? pa : pao, on ? pb : pbo, on ? pc : pco, xp, cp);
            if (ierr) break;
            for (j = 0; j < n; j++) xe[j] = dt * inje[j] + xe[j];
            ierr = orc_thomas_ws(n, on ? ea : eao, on ? eb : ebo, on ? ec : eco, xe, cp);
            if (ierr) break;
        }
    }
    free(buf);
    return ierr;
}

/* ---------------------------------------------------------------------- */
/* nsub (tracer_fp_nsub.c:37-282) and segments (tracer_fp_setup.c:73-95)   */
static double orc_quantize_dt_sub(double t)
{
    static const double ladder[] = {1.0, 1.25, 1.6, 2.0, 2.5, 3.2, 4.0, 5.0, 6.4, 8.0};
    double e10, scale, norm, chosen;
    int i;
    if (!(t > 0.0) || !isfinite(t)) return t;
    e10 = floor(log10(t));
    scale = pow(10.0, e10);
    norm = t / scale;
    chosen = ladder[0];
    for (i = 0; i < 10; i++) {
        if (ladder[i] <= norm) chosen = ladder[i];
        else break;
    }
    return chosen * scale;
}

static double orc_quantize_on_fraction(double f)
{
    static const double bins[] = {1.0 / 8.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 2.0, 1.0};
    double chosen, best;
    int i;
    if (!(f > 0.0) || !isfinite(f)) return 1.0;
    if (f >= 1.0) return 1.0;
    chosen = bins[0];
    best = fabs(f - chosen);
    for (i = 1; i < 5; i++) {
        const double d = fabs(f - bins[i]);
        if (d < best) { chosen = bins[i]; best = d; }
    }
    return chosen;
}

void orc_nsub_endpoints(const OrcCtx *ctx, int ntracer, const double *n_gas,
                        const double *kbt, const double *b_field, const double *divv,
                        double rad_ic_lo, double rad_ic_hi,
                        double *crp_lo, double *crp_hi, double *cre_lo, double *cre_hi)
{
    const int n = ctx->n;
    int it;
    for (it = 0; it < ntracer; it++) {
        crp_lo[it] = orc_b_coulomb_p(n_gas[it], ctx->pp[0], kbt[it]) + 1.0 / 3.0 * divv[it] * ctx->pp[0];
        crp_hi[it] = orc_b_coulomb_p(n_gas[it], ctx->pp[n - 1], kbt[it]) + 1.0 / 3.0 * divv[it] * ctx->pp[n - 1];
        cre_lo[it] = orc_b_synch(ctx->ep[0], b_field[it]) + rad_ic_lo +
                     orc_b_coulomb_e(n_gas[it], ctx->ep[0]) + 1.0 / 3.0 * divv[it] * ctx->ep[0];
        cre_hi[it] = orc_b_synch(ctx->ep[n - 1], b_field[it]) + rad_ic_hi +
                     orc_b_coulomb_e(n_gas[it], ctx->ep[n - 1]) + 1.0 / 3.0 * divv[it] * ctx->ep[n - 1];
    }
}

int orc_nsub_compute(const OrcCtx *ctx, const OrcParams *prm, int ntracer, double dt_snap,
                     const double *crp_lo, const double *crp_hi,
                     const double *cre_lo, const double *cre_hi,
                     const double *l_turb, const double *dv,
                     int *nsub_out, int *target_out, int *n_on_out)
{
    const int n = ctx->n;
    int it;
    if (ntracer <= 0 || dt_snap <= 0.0) return -1;
    for (it = 0; it < ntracer; it++) {
        const double te0 = ctx->ep[0] / cre_lo[it];
        const double tp0 = ctx->pp[0] / crp_lo[it];
        const double teN = ctx->ep[n - 1] / cre_hi[it];
        const double tpN = ctx->pp[n - 1] / crp_hi[it];
        const double tmin = fmin(fabs(tp0), fmin(fabs(teN), fmin(fabs(te0), fabs(tpN))));
        int nsub, n_on;
        double target = dt_snap / 10.0, used, tnsub, raw, on_fraction = 1.0;
        if (isfinite(tmin) && tmin > 0.0) {
            const double st = prm->tracer_nsub_safety * tmin;
            if (st < target) target = st;
        }
        if (target <= 0.0 || !isfinite(target)) target = dt_snap / 10.0;
        tnsub = ceil(dt_snap / target);
        if (!isfinite(tnsub) || tnsub > 1000000.0) tnsub = 1000000.0;
        if (tnsub < 10.0) tnsub = 10.0;
        used = orc_quantize_dt_sub(target);
        if (!(used > 0.0) || !isfinite(used) || used > target) used = target;
        raw = ceil(dt_snap / used);
        if (!isfinite(raw) || raw > 1000000.0) raw = 1000000.0;
        if (raw < 10.0) raw = 10.0;
        nsub = (int)raw;
        if (prm->tracer_nsub_max > 0 && nsub > prm->tracer_nsub_max) nsub = prm->tracer_nsub_max;
        if (prm->reacc_window_mode == 1 && prm->f_eddy > 0.0 && dv[it] > 0.0) {
            const double t_eddy = l_turb[it] * 3.0857e24 / dv[it] / 3.1536e16;
            if (prm->f_eddy * t_eddy < dt_snap)
                on_fraction = orc_quantize_on_fraction(prm->f_eddy * t_eddy / dt_snap);
        }
        n_on = (int)floor((double)nsub * on_fraction + 0.5);
        if (n_on < 1) n_on = 1;
        if (n_on > nsub) n_on = nsub;
        nsub_out[it] = nsub;
        if (target_out) target_out[it] = (int)tnsub;
        n_on_out[it] = n_on;
    }
    return 0;
}

int orc_choose_segments(const OrcParams *prm, int nsub, int allow)
{
    int segments = 1, max_allowed, maxseg = prm->coeff_interp_max_segments;
    if (!allow || nsub <= 1) return 1;
    if (prm->coeff_interp_min_steps <= 0) return 1;
    if (maxseg < 1) maxseg = 1;
    max_allowed = nsub / prm->coeff_interp_min_steps;
    if (max_allowed < 1) return 1;
    while (segments <= 1073741823 && segments * 2 <= maxseg && segments * 2 <= max_allowed)
        segments *= 2;
    if (segments > nsub) segments = nsub;
    if (segments < 1) segments = 1;
    return segments;
}

/* ---------------------------------------------------------------------- */
/* Snapshot execution: per-tracer semantics of tracer_fp_batch_execute_bucket
 * (tracer_fp_batch.c:651-827) with LPT thread assignment (:829-1126).     */
typedef struct { int tracer; long long weight; } OrcTask;

static int orc_task_cmp(const void *l, const void *r)
{
    const OrcTask *a = (const OrcTask *)l, *b = (const OrcTask *)r;
    if (a->weight < b->weight) return 1;
    if (a->weight > b->weight) return -1;
    return (a->tracer > b->tracer) - (a->tracer < b->tracer);
}

static double orc_lerp(double cur, double nxt, double alpha)
{
    if (alpha <= 0.0) return cur;
    if (alpha >= 1.0) return nxt;
    return cur + alpha * (nxt - cur);
}

static int orc_run_tracer(const OrcCtx *ctx, const OrcParams *prm, int it, int ntracer,
                          int nsub, int n_on, int allow_interp, double elapsed_gyr,
                          double dt_snap, const double *const *bgc, const double *rad_ic_zero,
                          const double *rm1z, const double *rp1z, const double *const *bgn,
                          const double *row_next, double m1n, double p1n,
                          const double *qpi, const double *qepri, const double *fqe_t,
                          const int *np_min_qe, const double *mass,
                          const unsigned char *dis, double *crp, double *cre, double *ric)
{
    const int n = ctx->n;
    const int nseg = orc_choose_segments(prm, nsub, allow_interp);
    const int on_start = (nsub - n_on) / 2;
    const int on_end = on_start + n_on;
    const double base_dt = dt_snap / (double)nsub;
    const size_t nn = (size_t)n;
    int iseg, j, k;

    for (iseg = 0; iseg < nseg; iseg++) {
        const double alpha = (nseg > 1) ? ((double)iseg + 0.5) / (double)nseg : 0.0;
        const int fb = (int)(((long long)iseg * (long long)nsub) / (long long)nseg);
        const int fe = (int)(((long long)(iseg + 1) * (long long)nsub) / (long long)nseg);
        const int steps = fe - fb;
        double bg[8], m1, p1, tacc_d;
        int sob, soe, windowed, ls, le;
        if (steps <= 0) continue;
        for (k = 0; k < 8; k++) bg[k] = (nseg > 1) ? orc_lerp(bgc[k][it], bgn[k][it], alpha) : bgc[k][it];
        for (j = 0; j < n; j++) {
            const double cur = rad_ic_zero[(size_t)j * (size_t)ntracer + (size_t)it];
            ric[j] = (nseg > 1) ? orc_lerp(cur, row_next[j], alpha) : cur;
        }
        m1 = (nseg > 1) ? orc_lerp(rm1z[it], m1n, alpha) : rm1z[it];
        p1 = (nseg > 1) ? orc_lerp(rp1z[it], p1n, alpha) : rp1z[it];
        sob = (on_start > fb) ? on_start : fb;
        soe = (on_end < fe) ? on_end : fe;
        if (soe < sob) soe = sob;
        windowed = ((soe - sob) < steps) ? 1 : 0;
        ls = sob - fb;
        le = ls + (soe - sob);
        {
            const double tmid = elapsed_gyr + ((double)fb + 0.5 * (double)steps) * base_dt;
            const int direct_on = prm->dpp_model == ORC_DPP_DIRECT && isfinite(prm->t_acc_direct_gyr) &&
                                  prm->t_acc_direct_gyr > 0.0 && isfinite(prm->t_off_gyr) &&
                                  prm->t_off_gyr > 0.0 && isfinite(prm->t_acc_off_gyr) &&
                                  prm->t_acc_off_gyr > 0.0;
            tacc_d = (direct_on && tmid < prm->t_off_gyr) ? prm->t_acc_off_gyr : prm->t_acc_direct_gyr;
        }
        if (orc_evolve(ctx, prm, 1, steps, dt_snap / (double)nsub, windowed, ls, le, tacc_d,
                       &bg[0], &bg[1], &bg[2], &bg[3], &bg[4], &bg[5], &bg[6], &bg[7],
                       ric, &m1, &p1, qpi + (size_t)it * nn, qepri + (size_t)it * nn,
                       fqe_t, np_min_qe, mass + it, dis ? dis + it : 0,
                       crp + (size_t)it * nn, cre + (size_t)it * nn) != 0)
            return -1;
    }
    return 0;
}

int orc_execute_snapshot(const OrcCtx *ctx, const OrcParams *prm, int ntracer,
                         const int *nsub, const int *n_on, int allow_interp,
                         double elapsed_gyr, double dt_snap,
                         const double *const *bg_curr, const double *rad_ic_zero,
                         const double *rad_ic_m1_zero, const double *rad_ic_p1_zero,
                         const double *const *bg_next, const double *rad_ic_row_next,
                         double rad_ic_m1_next, double rad_ic_p1_next,
                         const double *qpi, const double *qepri,
                         const double *fqe_t, const int *np_min_qe,
                         const double *mass, const unsigned char *disable_adiabatic,
                         double *crp, double *cre, int nthreads)
{
    OrcTask *tasks;
    long long *loads;
    int *owner;
    int nthr = 1, it, err = 0;

#ifdef _OPENMP
    nthr = (nthreads > 0) ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
#endif
    if (nthr < 1) nthr = 1;
    tasks = (OrcTask *)calloc((size_t)ntracer, sizeof(OrcTask));
    loads = (long long *)calloc((size_t)nthr, sizeof(long long));
    owner = (int *)calloc((size_t)ntracer, sizeof(int));
    if (!tasks || !loads || !owner) { free(tasks); free(loads); free(owner); return -1; }
    for (it = 0; it < ntracer; it++) {
        tasks[it].tracer = it;
        tasks[it].weight = (long long)nsub[it] * (long long)orc_choose_segments(prm, nsub[it], allow_interp);
    }
    qsort(tasks, (size_t)ntracer, sizeof(OrcTask), orc_task_cmp);
    for (it = 0; it < ntracer; it++) {
        int best = 0, t;
        for (t = 1; t < nthr; t++)
            if (loads[t] < loads[best]) best = t;
        owner[it] = best;
        loads[best] += tasks[it].weight;
    }
#pragma omp parallel num_threads(nthr) reduction(| : err)
    {
        int tid = 0, k;
        double *ric = (double *)malloc((size_t)ctx->n * sizeof(double));
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        for (k = 0; k < ntracer && ric; k++) {
            const int tr = tasks[k].tracer;
            if (owner[k] != tid) continue;
            if (orc_run_tracer(ctx, prm, tr, ntracer, nsub[tr], n_on[tr], allow_interp,
                               elapsed_gyr, dt_snap, bg_curr, rad_ic_zero, rad_ic_m1_zero,
                               rad_ic_p1_zero, bg_next, rad_ic_row_next, rad_ic_m1_next,
                               rad_ic_p1_next, qpi, qepri, fqe_t, np_min_qe, mass,
                               disable_adiabatic, crp, cre, ric) != 0)
                err |= 1;
        }
        if (!ric) err |= 1;
        free(ric);
    }
    free(tasks); free(loads); free(owner);
    return err ? -1 : 0;
}

/* ---------------------------------------------------------------------- */
/* IC cooling (FP_Coef.c:361-539)                                          */
static double orc_ic_kernel(double eps0, double gamma, double eps)
{
    const double re = 2.82e-13;
    double kappa, C = 0.0;
    kappa = eps / (4.0 * eps0 * gamma * (gamma - eps));
    if (kappa <= 0.0) {
        C = 0.0;
    } else {
        C = 2.0 * kappa * log(kappa) + (1.0 + 2.0 * kappa) * (1.0 - kappa) +
            pow(4.0 * eps0 * kappa * gamma, 2.0) / (2.0 * (1.0 + 4.0 * eps0 * gamma * kappa)) * (1.0 - kappa);
        C = 2.0 * M_PI * re * re * K_c / (gamma * gamma * eps0) * C;
    }
    return C;
}

static double orc_cmb_density(double eps, double z)
{
    const double T0 = 2.757, meg = 9.11e-28, k = 1.38e-16, epsilon = 1.0e-40;
    double Tcmb = T0 * (1.0 + z), x;
    if (eps * meg * K_c * K_c / (k * Tcmb) > 90.0 * log(10.0)) return 0;
    if (exp(eps * meg * K_c * K_c / (k * Tcmb)) - 1.0 < epsilon) {
        x = eps * meg * K_c * K_c / (k * Tcmb);
        return 8.0 * M_PI * K_c * K_c * K_c * pow(meg / K_planck, 3.0) * eps /
               (meg * K_c * K_c * (1.0 + 0.5 * x + x / 6.0) / (k * Tcmb));
    }
    return 8.0 * M_PI * K_c * K_c * K_c * pow(meg / K_planck, 3.0) * pow(eps, 2.0) /
           (exp(eps * meg * K_c * K_c / (k * Tcmb)) - 1.0);
}

static double orc_cmb_density_deriv(double eps, double z)
{
    const double h_PL = 6.63e-27, T0 = 2.757, meg = 9.11e-28;
    const double k = 1.38 * pow(10.0, -16.0);
    const double Tcmb = T0 * (1.0 + z);
    return 8.0 * M_PI * K_c * K_c * K_c * pow(meg / h_PL, 3.0) * eps /
           (exp(eps * meg * K_c * K_c / (k * Tcmb)) - 1.0) *
           (2.0 - eps * meg * K_c * K_c * exp(eps * meg * K_c * K_c / (k * Tcmb)) /
                      (k * Tcmb * (exp(eps * meg * K_c * K_c / (k * Tcmb)) - 1.0)));
}

static void orc_cmb_bounds(double z, double *emin, double *emax)
{
    double eps, epsnew, delta = pow(10.0, -5.5), peak, epsp;
    const double meg = 9.11 * pow(10.0, -28.0), c = 3.0e+10, T0 = 2.725;
    const double k = 1.38 * pow(10.0, -16.0);
    epsp = k / (meg * c * c) * T0 * (1.0 + z);
    peak = orc_cmb_density(epsp, z);
    epsnew = 0.01 * epsp;
    eps = 1.0;
    while (fabs(epsnew - eps) > delta * eps) {
        eps = epsnew;
        epsnew = eps - (orc_cmb_density(eps, z) - peak * delta) / orc_cmb_density_deriv(eps, z);
    }
    *emin = epsnew;
    epsnew = 10.0 * epsp;
    eps = 1.0;
    while (fabs(epsnew - eps) > delta * eps) {
        eps = epsnew;
        epsnew = eps - (orc_cmb_density(eps, z) - peak * delta) / orc_cmb_density_deriv(eps, z);
    }
    *emax = epsnew;
}

static double orc_b_ic_integrand(double eps0, double gamma, double emin, double emax, double z)
{
    int i, N;
    double eps, epsnew, deps, Integral = 0.0, a, da = 0.50e-1, C, newC;
    (void)z;
    N = (int)round((log10(emax) - log10(emin)) / da);
    epsnew = emin;
    a = log10(emin);
    newC = orc_ic_kernel(eps0, gamma, epsnew);
    for (i = 0; i < N; i++) {
        a = a + da;
        eps = epsnew;
        epsnew = pow(10.0, a);
        deps = epsnew - eps;
        C = newC;
        newC = orc_ic_kernel(eps0, gamma, epsnew);
        Integral = Integral + 0.5 * (eps * C + epsnew * newC) * deps;
    }
    return Integral;
}

static double orc_b_ic(double gamma, double e0min, double e0max, double z)
{
    int i, N;
    double emin, emax, Integral = 0.0, deps0, e0, e0new, a, da = 0.50e-1, n0new, IC2, IC2new, nn;
    N = (int)round((log10(e0max) - log10(e0min)) / da);
    a = log10(e0min);
    e0new = e0min;
    emin = e0new;
    emax = gamma * 4.0 * e0new * gamma / (1.0 + 4.0 * e0new * gamma);
    IC2new = orc_b_ic_integrand(e0new, gamma, emin, emax, z);
    n0new = orc_cmb_density(e0new, z);
    for (i = 0; i < N; i++) {
        a = a + da;
        e0 = e0new;
        nn = n0new;
        IC2 = IC2new;
        e0new = pow(10.0, a);
        n0new = orc_cmb_density(e0new, z);
        deps0 = e0new - e0;
        emin = e0new;
        emax = gamma * 4.0 * e0new * gamma / (1.0 + 4.0 * e0new * gamma);
        IC2new = orc_b_ic_integrand(e0new, gamma, emin, emax, z);
        Integral = Integral + 0.5 * (nn * IC2 + n0new * IC2new) * deps0;
    }
    return Integral * K_Gyr;
}

int orc_ic_cooling_row(const OrcCtx *ctx, double z, double *row, double *m1, double *p1)
{
    double e0min, e0max;
    int j;
    orc_cmb_bounds(z, &e0min, &e0max);
    for (j = 0; j < ctx->n; j++) row[j] = orc_b_ic(sqrt(1.0 + ctx->ep[j] * ctx->ep[j]), e0min, e0max, z);
    *m1 = orc_b_ic(sqrt(1.0 + ctx->epm1 * ctx->epm1), e0min, e0max, z);
    *p1 = orc_b_ic(sqrt(1.0 + ctx->epp1 * ctx->epp1), e0min, e0max, z);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Emission kernels (FP_Coef.c:540-582; HADRONIC.c:139-289)               */
double orc_ic_emissivity_kernel(double nu, double gamma, double z)
{
    int i, N = 400;
    double Integral = 0.0, eps;
    const double meg = 9.11e-28, c = 3.0e+10;
    double a, da, C, Cnew, eps0, eps0new, deps0, nn, n0new, e0min, e0max;
    eps = K_planck * nu / (meg * c * c);
    e0max = eps;
    e0min = eps / (4.0 * gamma * (gamma - eps));
    if (e0min < 0) e0min = eps / (gamma * gamma);
    da = (log10(e0max) - log10(e0min)) / (double)N;
    a = log10(e0min);
    eps0new = pow(10.0, a);
    Cnew = orc_ic_kernel(eps0new, gamma, eps);
    n0new = orc_cmb_density(eps0new, z);
    for (i = 0; i < N; i++) {
        a = a + da;
        eps0 = eps0new;
        eps0new = pow(10.0, a);
        deps0 = eps0new - eps0;
        C = Cnew;
        nn = n0new;
        n0new = orc_cmb_density(eps0new, z);
        Cnew = orc_ic_kernel(eps0new, gamma, eps);
        Integral = Integral + 0.5 * (C * nn + Cnew * n0new) * deps0;
    }
    return K_planck * eps * Integral / (4.0 * M_PI);
}

double orc_gamma_kernel(double Egamma, double p)
{
    int i, Na = 200;
    double gp = sqrt(p * p + 1.0), Ep = gp * K_mp;
    double Emax = Ep, Emin = Egamma + K_mpi0 * K_mpi0 / (4.0 * Egamma);
    double dEpi, Epi, Integral = 0.0, a, amin, amax, da, f, fpre, Epipre;
    if (Emax < Emin) return 0;
    if (Emin < K_mpi0) Emin = K_mpi0;
    if ((Ep - K_mp) < 0.285) return 0;
    amin = log10(Emin);
    amax = log10(Emax);
    da = (amax - amin) / (double)Na;
    a = amin;
    Epi = pow(10.0, a);
    f = orc_dsigma(0, Epi, Ep) / sqrt(Epi * Epi - K_mpi0 * K_mpi0);
    for (i = 1; i < Na; i++) {
        Epipre = Epi;
        fpre = f;
        a += da;
        Epi = pow(10.0, a);
        dEpi = Epi - Epipre;
        f = orc_dsigma(0, Epi, Ep) / sqrt(Epi * Epi - K_mpi0 * K_mpi0);
        Integral += 0.5 * (f + fpre) * dEpi;
    }
    return 2.0 * K_c * Integral / (4.0 * M_PI);
}

static double orc_fe(double x)
{
    double r = K_mmu * K_mmu / (K_mpic * K_mpic), x2 = x * x, r2 = r * r;
    double g = (3.0 - 2.0 * r) * (9.0 * x2 - 6.0 * log(x) - 4.0 * x2 * x - 5.0) / (9.0 * pow(1.0 - r, 2.0));
    double h1 = (3.0 - 2.0 * r) * (9.0 * r2 - 6.0 * log(r) - 4.0 * r2 * r - 5.0) / (9.0 * pow(1.0 - r, 2.0));
    double h2 = (1.0 + 2.0 * r) * (r - x) * (9.0 * (r + x) - 4.0 * (r2 + r * x + x2)) / (9.0 * r2);
    if (x > r) return g;
    return h1 + h2;
}

static double orc_fnue(double x)
{
    double r = 0.573, x2 = x * x, r2 = r * r;
    double g = 2.0 * (1.0 - x) *
               (6.0 * pow(1.0 - x, 2.0) + r * (5.0 + 5.0 * x - 4.0 * x2) + 6.0 * r * log(x) / (1.0 - x)) /
               (3.0 * pow(1.0 - r, 2.0));
    double h1 = 2.0 * ((1.0 - r) * (6.0 - 7.0 * r + 11.0 * r2 - 4.0 * r2 * r) + 6.0 * r * log(r)) /
                (3.0 * pow(1.0 - r, 2.0));
    double h2 = 2.0 * (r - x) * (7.0 * r2 - 4.0 * r2 * r + 7.0 * x * r - 4.0 * x * r2 - 2.0 * x2 - 4.0 * x2 * r) /
                (3.0 * r2);
    if (x > r) return g;
    return h1 + h2;
}

/* kind: 1 = nu_e, 2 = nu_mu (HADRONIC.c:179-235) */
static double orc_ekernel(int kind, double E, double p)
{
    double gp = sqrt(p * p + 1.0), Ep = gp * K_mp;
    double Emax = gp * K_mp;
    double Emin = 2.0 * E * K_mpic * K_mpic / (K_mpic * K_mpic + K_mmu * K_mmu);
    double Epi, Integral = 0.0, Integral_2 = 0.0, x, dx = 1.0 / 200.0, f, fpre;
    int i;
    if (Emax < Emin) return 0;
    if (Emin < K_mpic) Emin = K_mpic;
    if ((Ep - K_mp) < 0.289) return 0;
    x = 0.0;
    f = 0.0;
    for (i = 1; i < 200; i++) {
        x += dx;
        Epi = E / x;
        fpre = f;
        if (kind == 1) f = orc_dsigma(1, Epi, Ep) * orc_fnue(x) / x;
        else f = orc_dsigma(1, Epi, Ep) * orc_fe(x) / x;
        if (i == 1) Integral += f * dx;
        else Integral += 0.5 * (f + fpre) * dx;
    }
    x = 0.0;
    f = 0.0;
    for (i = 1; i < 200; i++) {
        x += dx;
        Epi = E / x;
        fpre = f;
        if (kind == 1) f = orc_dsigma(-1, Epi, Ep) * orc_fnue(x) / x;
        else f = orc_dsigma(-1, Epi, Ep) * orc_fe(x) / x;
        if (i == 1) Integral_2 += f * dx;
        else Integral_2 += 0.5 * (f + fpre) * dx;
    }
    return K_c * (Integral + Integral_2) / (4.0 * M_PI);
}

static double orc_numu1_kernel(double Enu, double p)
{
    double lambda = 1.0 - K_mmu * K_mmu / (K_mpic * K_mpic);
    double gp = sqrt(p * p + 1.0), Ep = gp * K_mp, Emax = gp * K_mp, Emin = Enu / lambda;
    double Epi, Integral = 0.0, Integral_2 = 0.0, f, fpre, x, dx = lambda / 200.0;
    int i;
    if (Emax < Emin) return 0;
    if (Emin < K_mpic) Emin = K_mpic;
    if ((Ep - K_mp) < 0.29) return 0;
    x = 0.0;
    f = 0.0;
    for (i = 1; i < 200; i++) {
        x += dx;
        Epi = Enu / x;
        fpre = f;
        f = orc_dsigma(1, Epi, Ep) / x;
        if (i == 1) Integral += f * dx;
        else Integral += 0.5 * (f + fpre) * dx;
    }
    x = 0.0;
    f = 0.0;
    for (i = 1; i < 200; i++) {
        x += dx;
        Epi = Enu / x;
        fpre = f;
        f = orc_dsigma(-1, Epi, Ep) / x;
        if (i == 1) Integral_2 += f * dx;
        else Integral_2 += 0.5 * (f + fpre) * dx;
    }
    return K_c * (Integral + Integral_2) / lambda / (4.0 * M_PI);
}

double orc_nu_kernel(double enu, double p)
{
    return orc_ekernel(2, enu, p) + orc_ekernel(1, enu, p) + orc_numu1_kernel(enu, p);
}

void orc_gamma_kernel_flat(const OrcCtx *ctx, int nbins, double emin, double emax, double *fga)
{
    double d = (emax - emin) / (double)nbins, a = emin - d;
    int k, j;
    for (k = 0; k < nbins; k++) {
        double e;
        a += d;
        e = pow(10.0, a);
        for (j = 0; j < ctx->n; j++) fga[(size_t)k * ctx->n + j] = orc_gamma_kernel(e, ctx->pp[j]);
    }
}

void orc_nu_kernel_flat(const OrcCtx *ctx, int nbins, double emin, double emax, double *fnu)
{
    double d = (emax - emin) / (double)nbins, a = emin - d;
    int k, j;
    for (k = 0; k < nbins; k++) {
        double e;
        a += d;
        e = pow(10.0, a);
        for (j = 0; j < ctx->n; j++) fnu[(size_t)k * ctx->n + j] = orc_nu_kernel(e, ctx->pp[j]);
    }
}

void orc_ic_kernel_flat(const OrcCtx *ctx, int nbins, double numin, double numax, double z,
                        double *fic)
{
    double d = (numax - numin) / (double)nbins, a = numin - d;
    int k, j;
    for (k = 0; k < nbins; k++) {
        double nu;
        a += d;
        nu = pow(10.0, a);
        for (j = 0; j < ctx->n; j++) {
            const double gamma = sqrt(1.0 + ctx->ep[j] * ctx->ep[j]);
            fic[(size_t)k * ctx->n + j] = orc_ic_emissivity_kernel(nu, gamma, z);
        }
    }
}

void orc_emission_ic(int n, int ntracer, int nbins, const double *cre, const double *dp,
                     const double *fic, double *eps)
{
    int it, k, j;
    for (it = 0; it < ntracer; it++) {
        const double *cl = cre + (size_t)it * n;
        for (k = 0; k < nbins; k++) {
            const double *row = fic + (size_t)k * n;
            double s = 0.0;
            for (j = 0; j < n; j++) s += row[j] * cl[j] * dp[j];
            eps[(size_t)it * nbins + k] = s;
        }
    }
}

/* gamma / neutrino: eps = E * Q_one(fk_row, N, beta, dp) * n_gas
 * (fp_shared_core.c:2166-2243, EMISSION.c:297-308) */
void orc_emission_hadronic(int n, int ntracer, int nbins, double emin, double emax,
                           const double *n_gas, const double *crp, const double *beta_p,
                           const double *dp, const double *fk, double *eps)
{
    double d = (emax - emin) / (double)nbins, a = emin - d;
    int k, it, i;
    for (k = 0; k < nbins; k++) {
        const double *row = fk + (size_t)k * n;
        double e;
        a += d;
        e = pow(10.0, a);
        for (it = 0; it < ntracer; it++) {
            const double *Np = crp + (size_t)it * n;
            double Integral = 0.0, fn, fpre = beta_p[0] * Np[0] * row[0];
            for (i = 1; i < n; i++) {
                fn = beta_p[i] * Np[i] * row[i];
                Integral += (fpre + fn) * dp[i];
                fpre = fn;
            }
            eps[(size_t)it * nbins + k] = e * (Integral * 0.5) * n_gas[it];
        }
    }
}

/* Synchrotron tables (tracer_fp_synch.c:87-121, Synchrotron.c:9-63) */
void orc_synch_logy(const OrcCtx *ctx, int nfreq, double nu_min, double nu_max, int ntheta,
                    double *logy, double *pitch_weight)
{
    double *theta = (double *)malloc((size_t)ntheta * sizeof(double));
    double *dtheta = (double *)malloc((size_t)ntheta * sizeof(double));
    double *nus = (double *)malloc((size_t)nfreq * sizeof(double));
    const double dnu = (nu_max - nu_min) / (double)nfreq;
    double a = nu_min - dnu;
    const double c0 = 4.194284e+06;
    int i, j, k;
    for (k = 0; k < ntheta; k++) {
        theta[k] = (k + 0.5) * M_PI / (double)ntheta;
        dtheta[k] = M_PI / (double)ntheta;
        pitch_weight[k] = sin(theta[k]) * sin(theta[k]) * dtheta[k];
    }
    for (i = 0; i < nfreq; i++) {
        a += dnu;
        nus[i] = pow(10.0, a);
    }
    for (i = 0; i < nfreq; i++)
        for (j = 0; j < ctx->n; j++)
            for (k = 0; k < ntheta; k++) {
                const double nu_crit = 1.0 * (1.0 + pow(ctx->ep[j], 2.0)) * c0 * sin(theta[k]);
                logy[((size_t)i * ctx->n + j) * ntheta + k] = log10(nus[i] / (1.0e-6 * nu_crit));
            }
    free(theta); free(dtheta); free(nus);
}

void orc_synch_kernel_table(int n, int nfreq, int ntheta, int nlogb, double logb_min,
                            double inv_dlogb, int nx_tab, double xmin,
                            const double *logx_tab, const double *logfx_tab,
                            const double *fx_tab, const double *logy,
                            const double *pitch_weight, double *table)
{
    int ib, nf, je, k;
    const double inv_dlogx = (nx_tab > 1) ? 1.0 / (logx_tab[1] - logx_tab[0]) : 0.0;
    for (ib = 0; ib < nlogb; ib++) {
        const double logb = logb_min + (double)ib / inv_dlogb;
        for (nf = 0; nf < nfreq; nf++)
            for (je = 0; je < n; je++) {
                double int_th = 0.0;
                for (k = 0; k < ntheta; k++) {
                    double Fx = 0.0;
                    const double logx = logy[((size_t)nf * n + je) * ntheta + k] - logb;
                    if (logx <= logx_tab[nx_tab - 1]) {
                        const int ix = (int)floor((logx - xmin) * inv_dlogx);
                        if (ix >= 0) {
                            if (ix >= nx_tab - 1) {
                                Fx = fx_tab[nx_tab - 1];
                            } else {
                                const double frac = (logx - logx_tab[ix]) * inv_dlogx;
                                Fx = pow(10.0, logfx_tab[ix] + frac * (logfx_tab[ix + 1] - logfx_tab[ix]));
                            }
                        }
                    }
                    int_th += Fx * pitch_weight[k];
                }
                table[((size_t)ib * nfreq + nf) * n + je] = int_th;
            }
    }
}

void orc_emission_synch(int n, int ntracer, int nfreq, int nlogb, double logb_min,
                        double inv_dlogb, const double *table, const double *cre,
                        const double *dp, const double *b_field, double *eps)
{
    int it, nf, je;
    for (it = 0; it < ntracer; it++) {
        const double *cl = cre + (size_t)it * n;
        double bval = b_field[it], logb, u, t;
        int ib;
        if (!(bval > 0.0) || !isfinite(bval)) bval = 1.0e-30;
        logb = log10(bval / 1.0e-6);
        u = (logb - logb_min) * inv_dlogb;
        ib = (int)floor(u);
        if (ib < 0) ib = 0;
        if (ib > nlogb - 2) ib = nlogb - 2;
        t = u - (double)ib;
        if (t < 0.0) t = 0.0;
        if (t > 1.0) t = 1.0;
        for (nf = 0; nf < nfreq; nf++) {
            const double A = sqrt(3.0) * pow(4.8e-10, 3.0) * b_field[it] / (2.0 * 9.11e-28 * K_c * K_c) / (4.0 * M_PI);
            double integral = 0.0;
            for (je = 2; je < n - 1 && cl[je + 1] > 0.0; je++) {
                const double k0 = table[((size_t)ib * nfreq + nf) * n + je];
                const double k1 = table[((size_t)(ib + 1) * nfreq + nf) * n + je];
                double K;
                if (t <= 0.0) K = k0;
                else if (t >= 1.0) K = k1;
                else if (k0 > 0.0 && k1 > 0.0) K = exp((1.0 - t) * log(k0) + t * log(k1));
                else K = (1.0 - t) * k0 + t * k1;
                integral += cl[je] * K * dp[je];
            }
            eps[(size_t)it * nfreq + nf] = A * integral;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Pieces of the hadronic kernels, exported so the tests can check them    */
/* against sum rules (pion / lepton counting) rather than against a copy   */
/* of their own formulas.                                                  */
double orc_hadron_dsigma(int kind, double epi, double ep) { return orc_dsigma(kind, epi, ep); }
double orc_hadron_fpi(double epi, double ep) { return orc_FpiQ(0, epi, ep); }
double orc_hadron_fe(double x) { return orc_fe(x); }
double orc_hadron_fnue(double x) { return orc_fnue(x); }
double orc_hadron_Fe(double ee, double epi) { return orc_Fe(ee, epi); }
double orc_hadron_ekernel(int kind, double e, double p) { return orc_ekernel(kind, e, p); }
double orc_hadron_numu1(double e, double p) { return orc_numu1_kernel(e, p); }
