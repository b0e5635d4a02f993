/*
 * ref_harness.c -- thin wrappers that call the reference's own C functions
 * (compiled from /root/reference/src by oracle/ref.mk) so the oracle
 * restatement can be checked against them.  TEST INFRASTRUCTURE ONLY.
 *
 * Only reference files that compile with this image's toolchain are used
 * (fp_shared_core.c, FP_Coef.c, CONSTANTS.c, params.c, tracer_fp_nsub.c,
 * DSA_MODELS.c, tracer_fp_dsa_reacc.c).
 * Chang_Cooper.c, HADRONIC.c, Synchrotron.c, EMISSION.c and the batch driver
 * need GSL / MPI headers that are absent, so nothing here reaches them (the
 * synchrotron emission driver is left out: its fallback branch calls
 * Synchrotron.c); the
 * link uses --gc-sections and -z defs, which fails if any wrapper does.
 *
 * Grid size comes from -Dnp=N -Dnpe=N (the reference's compile-time sizes).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "params.h"
#include "CONSTANTS.h"
#include "FP_Coef.h"
#include "fp_shared_core.h"
#include "tracer_fp_nsub.h"
#include "tracer_fp_types.h"
#include "DSA_MODELS.h"

#include <stddef.h>

#define REF_API __attribute__((visibility("default")))

static CRspectrum g_crp, g_cre;
static double g_bp[np], g_gamma2e[npe];

REF_API int ref_n(void) { return np; }

REF_API void ref_grid(double pmin_, double pmax_, double pemin_, double pemax_,
                      double *pp, double *pdp, double *pE, double *pghost,
                      double *ep, double *edp, double *eE, double *eghost, double *bp)
{
    pmin = pmin_;
    pmax = pmax_;
    pemin = pemin_;
    pemax = pemax_;
    memset(&g_crp, 0, sizeof(g_crp));
    memset(&g_cre, 0, sizeof(g_cre));
    momentum_bin(&g_crp, &g_cre, g_bp, g_gamma2e);
    memcpy(pp, g_crp.p, sizeof(g_crp.p));
    memcpy(pdp, g_crp.dp, sizeof(g_crp.dp));
    memcpy(pE, g_crp.E, sizeof(g_crp.E));
    pghost[0] = g_crp.pm1;
    pghost[1] = g_crp.pp1;
    memcpy(ep, g_cre.p, sizeof(g_cre.p));
    memcpy(edp, g_cre.dp, sizeof(g_cre.dp));
    memcpy(eE, g_cre.E, sizeof(g_cre.E));
    eghost[0] = g_cre.pm1;
    eghost[1] = g_cre.pp1;
    memcpy(bp, g_bp, sizeof(g_bp));
}

REF_API void ref_losses_e(double n_gas, double b_field, double divv, const double *rad_ic,
                          double rad_ic_m1, double rad_ic_p1, double *radp, double *tloss,
                          double *edges)
{
    double invtloss[npe];
    FPloss loss;
    loss.radp = radp;
    loss.tloss = tloss;
    loss.invtloss = invtloss;
    prepare_cre_losses_1d(n_gas, b_field, divv, rad_ic, rad_ic_m1, rad_ic_p1, &g_cre, &loss);
    edges[0] = loss.radpm1;
    edges[1] = loss.radpp1;
}

REF_API int ref_ic_row(double z, double *row, double *edges)
{
    return prepare_ic_cooling_row(z, &g_cre, row, &edges[0], &edges[1]);
}

REF_API double ref_b_coulomb_p(double n, double p, double kbt) { return b_Coulomb_p(n, p, kbt); }
REF_API double ref_b_coulomb_e(double n, double p) { return b_Coulomb_e(n, p); }
REF_API double ref_b_synch(double p, double b) { return b_synch(p, b); }

REF_API double ref_tacc_asa(double l_cm, double cs, double mach, double beta, double psi_)
{
    psi = psi_;
    return accelerationtime_ASA(l_cm, cs, mach, beta);
}

REF_API double ref_tacc_ttd(double l_cm, double cs, double mach, double dv, int model)
{
    ttd_tacc_model = model;
    return accelerationtime_TTD(l_cm, cs, mach, dv);
}

REF_API double ref_c_sound(double t) { return c_sound(t); }

REF_API int ref_nsub(int ntr, double dt_snap, const double *n_gas, const double *kbt,
                     const double *b_field, const double *divv, const double *rad_ic_mm,
                     const double *l_turb, const double *dv, double safety, int nsub_max,
                     int window_mode, double f_eddy_, int *nsub, int *target, int *n_on,
                     double *endpoints)
{
    tracer_nsub_safety = safety;
    tracer_nsub_max = nsub_max;
    reacc_window_mode = window_mode;
    f_eddy = f_eddy_;
    if (tracer_fp_prepare_nsub_loss_endpoints(ntr, &g_crp, &g_cre, n_gas, kbt, b_field, divv,
                                              rad_ic_mm, endpoints, endpoints + ntr,
                                              endpoints + 2 * ntr, endpoints + 3 * ntr) != 0)
        return -1;
    return tracer_fp_compute_nsubsteps_from_endpoints(
        ntr, dt_snap, &g_crp, &g_cre, endpoints, endpoints + ntr, endpoints + 2 * ntr,
        endpoints + 3 * ntr, l_turb, dv, nsub, target, n_on, 0, 0, 0, 0);
}

REF_API int ref_thomas(int n, const double *a, const double *b, const double *c, double *rhs)
{
    return fp_thomas_solve_cpu(n, a, b, c, rhs);
}

REF_API int ref_solve_cc_batch(int nsys, int nrow, double dt, const double *a, const double *b,
                               const double *c, const double *src, double *state)
{
    FpChangCooperSolveBatchInput in;
    memset(&in, 0, sizeof(in));
    in.nsys = nsys;
    in.nrow = nrow;
    in.dt = dt;
    in.a_batch = a;
    in.b_batch = b;
    in.c_batch = c;
    in.source_batch = src;
    return solve_cc_batch_cpu_timed(&in, state, 0);
}

REF_API int ref_pcr(int nsys, int nrow, const double *a, const double *b, const double *c,
                    const double *rhs, double *x)
{
    FpTridiagBatchInput in;
    in.nsys = nsys;
    in.nrow = nrow;
    in.a_batch = a;
    in.b_batch = b;
    in.c_batch = c;
    in.rhs_batch = rhs;
    return solve_tridiagonal_batch_pcr_cpu(&in, x);
}

REF_API void ref_secondary_source(int ncell, const double *n_gas, const double *crp,
                                  const int *np_min_qe, const double *fqe_t,
                                  const double *qepri, double *inje)
{
    double *qint = (double *)malloc((size_t)ncell * npe * sizeof(double));
    prepare_secondary_sources_cell_major(ncell, n_gas, crp, g_crp.dp, np_min_qe, fqe_t, qepri,
                                         qint, inje);
    free(qint);
}

REF_API void ref_ic_kernel_flat(int nbins, double numin, double numax, double z, double *fic)
{
    prepare_ic_kernel_flat(&g_cre, nbins, numin, numax, z, fic);
}

REF_API int ref_ic_emission(int ncell, int nbins, const double *cre, const double *fic, double *eps)
{
    FpIcEmissionBatchInput in;
    memset(&in, 0, sizeof(in));
    in.ncell = ncell;
    in.nbins = nbins;
    in.cre_batch = cre;
    in.fic_flat = fic;
    return prepare_ic_emission_cell_major(&in, &g_cre, eps);
}

/* logy is [nfreq][npe][ntheta] flat; the reference wants double*** planes. */
static double ***ref_logy_planes(int nfreq, int ntheta, const double *logy)
{
    double ***planes = (double ***)malloc((size_t)nfreq * sizeof(double **));
    double **rows = (double **)malloc((size_t)nfreq * npe * sizeof(double *));
    int i, j;
    for (i = 0; i < nfreq; i++) {
        planes[i] = rows + (size_t)i * npe;
        for (j = 0; j < npe; j++) planes[i][j] = (double *)logy + ((size_t)i * npe + j) * ntheta;
    }
    return planes;
}

static void ref_free_planes(double ***planes)
{
    free(planes[0]);
    free(planes);
}

REF_API int ref_synch_table(int nfreq, int ntheta, int nlogb, double logb_min, double inv_dlogb,
                            int nx_tab, double xmin, const double *logx_tab,
                            const double *logfx_tab, const double *fx_tab, const double *logy,
                            const double *pitch_weight, double *table)
{
    FpSynchEmissionBatchInput in;
    int rc;
    memset(&in, 0, sizeof(in));
    in.nfreq = nfreq;
    in.nx_tab = nx_tab;
    in.ntheta_pitch = ntheta;
    in.nlogb = nlogb;
    in.xmin = xmin;
    in.logb_min = logb_min;
    in.inv_dlogb = inv_dlogb;
    in.fx_tab = fx_tab;
    in.logfx_tab = logfx_tab;
    in.logx_tab = logx_tab;
    in.logy = ref_logy_planes(nfreq, ntheta, logy);
    in.pitch_weight = pitch_weight;
    rc = prepare_synch_pitch_kernel_table(&in, table);
    ref_free_planes(in.logy);
    return rc;
}

/* Byte layout of the reference's checkpoint header TracerFpCheckpointHeader
 * (include/tracer_fp_types.h:356-371): out[0] = sizeof, out[1..14] = offsetof
 * of magic, version, mpi_rank, world_size, ntracer, ntracer_global, nsnap,
 * nfreq, np_bins, npe_bins, nsnp_i_value, nsnp_f_value, completed_snapshot,
 * tracer_start; out[15] = sizeof(magic).  Returns the number written. */
REF_API int ref_ckpt_header_layout(long *out)
{
    out[0] = (long)sizeof(TracerFpCheckpointHeader);
    out[1] = (long)offsetof(TracerFpCheckpointHeader, magic);
    out[2] = (long)offsetof(TracerFpCheckpointHeader, version);
    out[3] = (long)offsetof(TracerFpCheckpointHeader, mpi_rank);
    out[4] = (long)offsetof(TracerFpCheckpointHeader, world_size);
    out[5] = (long)offsetof(TracerFpCheckpointHeader, ntracer);
    out[6] = (long)offsetof(TracerFpCheckpointHeader, ntracer_global);
    out[7] = (long)offsetof(TracerFpCheckpointHeader, nsnap);
    out[8] = (long)offsetof(TracerFpCheckpointHeader, nfreq);
    out[9] = (long)offsetof(TracerFpCheckpointHeader, np_bins);
    out[10] = (long)offsetof(TracerFpCheckpointHeader, npe_bins);
    out[11] = (long)offsetof(TracerFpCheckpointHeader, nsnp_i_value);
    out[12] = (long)offsetof(TracerFpCheckpointHeader, nsnp_f_value);
    out[13] = (long)offsetof(TracerFpCheckpointHeader, completed_snapshot);
    out[14] = (long)offsetof(TracerFpCheckpointHeader, tracer_start);
    out[15] = (long)sizeof(((TracerFpCheckpointHeader *)0)->magic);
    return 16;
}

/* DSA source rows of one shock from the reference's own builders
 * (dsa_init_grid with the pmin_dsa of tracer_fp.c:841, dsa_build_proton /
 * electron_source_fp), with the shock state and injection model filled as
 * tracer_fp_dsa.c:240-312 does (non-merged input branch, reacceleration off,
 * suprathermal electron injection).  eta_model: DSAModelID.  flags[0] /
 * flags[1] = fresh proton / electron source built; spec[0..4] = proton q, A_f,
 * p_inj, p_max, electron A_f. */
REF_API int ref_dsa_sources(int eta_model, double chi_p, double chi_e, double k_ep, double pmax_p, double pmax_e,
                            double gamma_gas, double mach, double temp1, double rho1, double u1,
                            double *source_p, double *source_e, int *flags, double *spec)
{
    static DSAGrid grid;
    static int grid_ready = 0;
    DSAInjectionModel model;
    DSAShockState shock;
    DSASpectrum sp, se;
    double ratio;
    if (!grid_ready) {
        memset(&grid, 0, sizeof(grid));
        if (!dsa_init_grid(&grid, &g_crp, &g_cre, -3.0, -3.0)) return -1;
        grid_ready = 1;
    }
    memset(&model, 0, sizeof(model));
    model.eta_model_initial = (DSAModelID)eta_model;
    model.chi_p = chi_p;
    model.chi_e = (chi_e > 0.0) ? chi_e : chi_p;
    model.K_ep = k_ep;
    model.p_dsamax_p_mc = pmax_p;
    model.p_dsamax_e_mc = pmax_e;
    model.electron_mode = DSA_ELECTRON_INJ_SUPRATHERMAL;
    memset(source_p, 0, sizeof(double) * np);
    memset(source_e, 0, sizeof(double) * npe);
    memset(&sp, 0, sizeof(sp));
    memset(&se, 0, sizeof(se));
    flags[0] = flags[1] = 0;
    ratio = dsa_compression_ratio(mach, gamma_gas);
    memset(&shock, 0, sizeof(shock));
    shock.mach = mach;
    shock.gamma_gas = gamma_gas;
    shock.rho1 = rho1;
    shock.u1 = u1;
    shock.u2 = u1 / ratio;
    shock.temperature = temp1;
    flags[0] = dsa_build_proton_source_fp(&shock, &model, &grid, &g_crp, source_p, &sp);
    if (flags[0]) flags[1] = dsa_build_electron_source_fp(&shock, &model, &grid, &g_cre, &sp, source_e, &se);
    spec[0] = sp.q;
    spec[1] = sp.A_f;
    spec[2] = sp.p_inj_mc;
    spec[3] = sp.p_max_mc;
    spec[4] = se.A_f;
    return 0;
}

REF_API double ref_dsa_eta(double mach, int eta_model) { return dsa_eta_from_id(mach, (DSAModelID)eta_model); }

/* Reacceleration pieces (tracer_fp_dsa_reacc.c, DSA_MODELS.c:474-522) on the
 * harness grid: kind 0 positive delta (volume_downstream = v2), 1 Drury83
 * convolution (v1 upstream, v2 downstream), 2 adiabatic compression delta
 * (compression ratio r); species 0 protons (proton grid), 1 electrons
 * (electron grid).  Returns the function's energy density; delta [n]. */
#include "tracer_fp_dsa_reacc.h"
REF_API double ref_dsa_reacc(int kind, int species, const double *state, double v1, double v2, double q, double r,
                             double pmin_mc, double *delta)
{
    const CRspectrum *g = species ? &g_cre : &g_crp;
    const int n = species ? npe : np;
    const DSASpecies sp = species ? DSA_SPECIES_ELECTRON : DSA_SPECIES_PROTON;
    if (kind == 0) return tracer_dsa_reacc_positive_delta_1d(g, state, n, v2, q, pmin_mc, delta, sp);
    if (kind == 1) return tracer_dsa_reacc_convolution_delta_1d(g, state, n, v1, v2, q, pmin_mc, delta, sp);
    return tracer_dsa_reacc_adiabatic_delta_1d(g, state, n, v1, v2, r, pmin_mc, delta, sp);
}

/* the energy density of a delta row (tracer_dsa_reacc_delta_energy_density_1d) */
REF_API double ref_dsa_reacc_edens(int species, const double *delta, double pmin_mc)
{
    const CRspectrum *g = species ? &g_cre : &g_crp;
    return tracer_dsa_reacc_delta_energy_density_1d(g, delta, species ? npe : np, pmin_mc,
                                                    species ? DSA_SPECIES_ELECTRON : DSA_SPECIES_PROTON);
}

REF_API double ref_dsa_eta_interpolated(double mach, int initial_id, int reacc_id, double x_cr)
{
    return dsa_eta_interpolated(mach, (DSAModelID)initial_id, (DSAModelID)reacc_id, x_cr);
}
