/*
 * croma_oracle.h -- CPU oracle for CROMA's per-tracer Fokker-Planck update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C (C11 + OpenMP, FP64) restatement
 * of the reference CPU path (tracer_fp_cpu.c -> Chang_Cooper.c) used as the
 * checker for the HIP library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * or calls it.
 *
 * Unlike the reference, the momentum-grid size n is a runtime value (the
 * reference fixes np == npe at compile time, params.h:21-27); every array
 * below with "n" entries is one species' grid.  Layouts follow the reference
 * names: "cell-major" = [tracer][bin], "momentum-major" = [bin][tracer].
 *
 * Pinning status of each part (how a plausible mistake is caught) is listed
 * at the top of croma_oracle.c.
 */
#ifndef CROMA_ORACLE_H
#define CROMA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Momentum-diffusion models, same numbering as fp_shared_core.h:245-251. */
enum {
    ORC_DPP_AUTO = 0,
    ORC_DPP_ASA = 1,
    ORC_DPP_TTD = 2,
    ORC_DPP_OFF = 3,
    ORC_DPP_DIRECT = 4
};

/* Physics / scheduling parameters the reference reads from params.c globals.
 * Field names follow the reference key names. */
typedef struct {
    int dpp_model;              /* ORC_DPP_* (dpp_mode_spec)               */
    int ttd_tacc_model;         /* 0: spectrum, 1: brunetti16 (default)     */
    int reacc_window_mode;      /* 0: always on, 1: centred window          */
    int tracer_nsub_max;        /* 0 = no cap                                */
    int coeff_interp_min_steps; /* default 4                                 */
    int coeff_interp_max_segments; /* default 16                             */
    double psi;                 /* default 0.5                               */
    double mach_limit;          /* default 0.5                               */
    double eta_dpp_cap;         /* <= 0 disables                             */
    double t_acc_direct_gyr;    /* direct_tacc model                         */
    double t_off_gyr;           /* direct_tacc switch-off time               */
    double t_acc_off_gyr;       /* t_acc before t_off                        */
    double f_eddy;              /* window length factor                      */
    double tracer_nsub_safety;  /* CFL safety                                */
    double epmax;               /* proton D_pp energy cutoff [GeV], 1e7      */
} OrcParams;

typedef struct OrcCtx OrcCtx;

/* Grids + per-grid caches (momentum_bin FP_Coef.c:291-330 and the
 * momentumdiff grid cache FP_Coef.c:77-154). */
OrcCtx *orc_ctx_create(int n, double pmin, double pmax, double pemin, double pemax);
void orc_ctx_free(OrcCtx *ctx);
int orc_ctx_n(const OrcCtx *ctx);
/* species 0 = proton, 1 = electron.  Copies p, dp, E and ghosts out. */
void orc_ctx_grid(const OrcCtx *ctx, int species, double *p, double *dp, double *E,
                  double *pm1, double *pp1);
void orc_ctx_beta_p(const OrcCtx *ctx, double *beta_p);
int orc_ctx_j_pp(const OrcCtx *ctx);
void orc_ctx_sigma_pp(const OrcCtx *ctx, double *sigma, double *sigmoid);

void orc_default_params(OrcParams *prm);

/* ---------------- physics pieces ---------------- */
double orc_pp_pion_cross_section(int kind, double p);
double orc_b_coulomb_p(double n_gas, double p, double kbt);
double orc_b_coulomb_e(double n_gas, double pe);
double orc_b_synch(double p, double b_gauss);
double orc_tacc_asa(double l_cm, double cs, double mach, double beta_pl, double psi);
double orc_tacc_ttd(double l_cm, double cs, double mach, double dv, int ttd_model);

void orc_losses_e(const OrcCtx *ctx, double n_gas, double b_field, double divv,
                  const double *rad_ic, double rad_ic_m1, double rad_ic_p1,
                  double *radp, double *tloss, double *radpm1, double *radpp1);
void orc_losses_p(const OrcCtx *ctx, double n_gas, double kbt, double divv,
                  double *radp, double *tloss, double *radpm1, double *radpp1);

/* On-set D_pp of `model` (eta cap applied when model != OFF) or the off floor.
 * edges = {dppm1, dppp1, dppem1, dppep1}. */
void orc_dpp(const OrcCtx *ctx, const OrcParams *prm, int model,
             double l_turb_mpc, double dv, double cs, double beta_pl,
             double mass_msun, const double *crp_state, const double *cre_state,
             double *dpp, double *dppe, double *edges);

/* Chang-Cooper face weights (Chang_Cooper.c:22-40). */
void orc_face_weights(double w, double *delta, double *wp, double *wm);
/* Chang-Cooper assembly from loss / diffusion terms with A = 1
 * (fp_shared_core.c:660-704, 2358-2431 -> Chang_Cooper.c:42-184). */
void orc_cc_build(int n, double dt, const double *x, const double *dx,
                  double xm1, double xp1, const double *radp, const double *dpp,
                  const double *tloss, double radpm1, double radpp1,
                  double dppm1, double dppp1,
                  double *a, double *b, double *c);

/* Thomas on b_i x_i - a_i x_{i-1} - c_i x_{i+1} = r_i, in place
 * (fp_shared_core.c:226-257).  Returns 0, -1 or -2 (pivot). */
int orc_thomas(int n, const double *a, const double *b, const double *c, double *rhs);
/* Batched PCR (fp_shared_core.c:430-547), cell-major, n power of two. */
int orc_pcr_batch(int nsys, int n, const double *a, const double *b, const double *c,
                  const double *rhs, double *x);

/* One coefficient segment of tracer_fp_cpu_evolve (tracer_fp_cpu.c:88-196)
 * for ntracer tracers; rad_ic is momentum-major [n * ntracer]; states,
 * qpi, qepri cell-major.  fqe_t is the transposed weight [n_p * n_e]. */
int orc_evolve(const OrcCtx *ctx, const OrcParams *prm, int ntracer, int nstep,
               double dt, int windowed, int on_start, int on_end,
               double t_acc_direct,
               const double *n_gas, const double *kbt, const double *b_field,
               const double *divv, const double *l_turb, const double *dv,
               const double *cs, const double *beta_pl,
               const double *rad_ic, const double *rad_ic_m1, const double *rad_ic_p1,
               const double *qpi, const double *qepri,
               const double *fqe_t, const int *np_min_qe,
               const double *mass, const unsigned char *disable_adiabatic,
               double *crp, double *cre);

/* nsub estimator (tracer_fp_nsub.c:37-282). */
void orc_nsub_endpoints(const OrcCtx *ctx, int ntracer, const double *n_gas,
                        const double *kbt, const double *b_field, const double *divv,
                        double rad_ic_lo, double rad_ic_hi,
                        double *crp_lo, double *crp_hi, double *cre_lo, double *cre_hi);
int orc_nsub_compute(const OrcCtx *ctx, const OrcParams *prm, int ntracer, double dt_snap,
                     const double *crp_lo, const double *crp_hi,
                     const double *cre_lo, const double *cre_hi,
                     const double *l_turb, const double *dv,
                     int *nsub, int *target_nsub, int *n_on);
int orc_choose_segments(const OrcParams *prm, int nsub, int allow_interp);

/* Snapshot-level update of every tracer (semantics of
 * tracer_fp_batch_execute_parallel, tracer_fp_batch.c:829-1126, with the
 * same longest-processing-time-first thread assignment).  rad_ic_zero is
 * momentum-major [n * ntracer]; rad_ic_row_next a broadcast row [n].
 * nthreads <= 0 uses the OpenMP default. */
int orc_execute_snapshot(const OrcCtx *ctx, const OrcParams *prm, int ntracer,
                         const int *nsub, const int *n_on, int allow_interp,
                         double elapsed_gyr, double dt_snap,
                         const double *const *bg_curr, /* 8 arrays [ntracer] */
                         const double *rad_ic_zero, const double *rad_ic_m1_zero,
                         const double *rad_ic_p1_zero,
                         const double *const *bg_next, /* 8 arrays [ntracer] */
                         const double *rad_ic_row_next, double rad_ic_m1_next,
                         double rad_ic_p1_next,
                         const double *qpi, const double *qepri,
                         const double *fqe_t, const int *np_min_qe,
                         const double *mass, const unsigned char *disable_adiabatic,
                         double *crp, double *cre, int nthreads);

/* IC cooling row at redshift z (fp_shared_core.c:844-864 -> FP_Coef.c:361-539). */
int orc_ic_cooling_row(const OrcCtx *ctx, double z, double *row, double *m1, double *p1);

/* Secondary-electron weights (fp_shared_core.c:1453-1506, transposed layout)
 * and source (fp_shared_core.c:1591-1672). */
void orc_secondary_kernel(const OrcCtx *ctx, double *fqe_t, int *np_min_qe);
double orc_qe_kernel_bb05(double ee, double p);
void orc_secondary_source(int n, int ntracer, const double *n_gas, const double *crp,
                          const int *np_min_qe, const double *fqe_t,
                          const double *qepri, double *inje);

/* Emission kernels and contractions (fp_shared_core.c:1724-1807, 1966-2243). */
double orc_gamma_kernel(double egamma, double p);
double orc_nu_kernel(double enu, double p);
double orc_ic_emissivity_kernel(double nu, double gamma, double z);
void orc_gamma_kernel_flat(const OrcCtx *ctx, int nbins, double emin, double emax, double *fga);
void orc_nu_kernel_flat(const OrcCtx *ctx, int nbins, double emin, double emax, double *fnu);
void orc_ic_kernel_flat(const OrcCtx *ctx, int nbins, double numin, double numax, double z,
                        double *fic);
void orc_emission_ic(int n, int ntracer, int nbins, const double *cre, const double *dp,
                     const double *fic, double *eps);
void orc_emission_hadronic(int n, int ntracer, int nbins, double emin, double emax,
                           const double *n_gas, const double *crp, const double *beta_p,
                           const double *dp, const double *fk, double *eps);
/* Synchrotron: pitch-kernel table (fp_shared_core.c:2008-2052) built from
 * the F(x) table, then the per-tracer emission (fp_shared_core.c:2070-2129). */
void orc_synch_logy(const OrcCtx *ctx, int nfreq, double nu_min, double nu_max,
                    int ntheta, double *logy, double *pitch_weight);
void orc_synch_kernel_table(int n, int nfreq, int ntheta, int nlogb, double logb_min,
                            double inv_dlogb, int nx_tab, double xmin,
                            const double *logx_tab, const double *logfx_tab,
                            const double *fx_tab, const double *logy,
                            const double *pitch_weight, double *table);
void orc_emission_synch(int n, int ntracer, int nfreq, int nlogb, double logb_min,
                        double inv_dlogb, const double *table, const double *cre,
                        const double *dp, const double *b_field, double *eps);


/* Pieces of the hadronic kernels (HADRONIC.c:139-498) for the sum-rule
 * checks: dsigma (pion production spectrum x cross section, :291-300), the
 * QGSJET pion spectrum FpiQ (:326-336), the Kelner decay spectra fe / fnue
 * (:302-322), the Brunetti-Blasi muon-decay electron spectrum Fe (:500-533),
 * eKernel for nu_e (kind 1) / nu_mu from muon decay (kind 2) (:179-235) and
 * numu1Kernel (:239-289). */
double orc_hadron_dsigma(int kind, double epi, double ep);
double orc_hadron_fpi(double epi, double ep);
double orc_hadron_fe(double x);
double orc_hadron_fnue(double x);
double orc_hadron_Fe(double ee, double epi);
double orc_hadron_ekernel(int kind, double e, double p);
double orc_hadron_numu1(double e, double p);

#ifdef __cplusplus
}
#endif

#endif
