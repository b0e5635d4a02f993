/*
 * croma_hip_b1.h -- the reference's GPU backend surface (include/fp_cuda_backend.h:10-159)
 * served by libcroma_hip.so, with the reference's own structs.
 *
 * The reference passes its structs by pointer; the typedefs below restate
 * their layouts field for field (same order, types and padding on LP64), so
 * a reference translation unit can hand its FpCudaPipelineInput* etc. to
 * these functions through a cast.  Their byte layouts are pinned against the
 * reference headers compiled at every grid size (sizeof and each offsetof;
 * croma_hip_b1_layout below exports ours).  CRspectrum (params.h:52-61) has
 * arrays of the compile-time bin count np, so every function that reads one
 * takes np (and npe) as its first arguments: the reference's compile-time
 * values, which the integration shim passes from its own macros.
 *
 * Conventions kept from the reference: 0 on success, -1 on invalid input,
 * host arrays caller-owned and read during the call, calls synchronous.
 * Changed: device errors return -1 (with croma_hip_last_error) instead of
 * exit(2); the physics parameters the reference reads from globals
 * (dpp_mode_spec, eta_dpp_cap, ttd_tacc_model, t_acc_direct_gyr, ...) are set
 * once with croma_hip_b1_set_params.
 */
#ifndef CROMA_HIP_B1_H
#define CROMA_HIP_B1_H

#include <stddef.h>

#include "croma_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FpCudaOutputSchedule (fp_cuda_backend.h:12-20) */
typedef struct {
    int emit_final_only;
    int emit_all_steps;
    int emit_stride;
    int physical_min;
    int physical_max;
    const int *selected_steps;
    int nselected_steps;
} CromaB1OutputSchedule;

/* FpCoeffBatchInput (fp_shared_core.h:366-407) */
typedef struct {
    int ncell;
    double z;
    double dt_gyr;
    double epmax;
    const double *n_gas;
    const double *kbt;
    const double *b_field;
    const double *divv_gyr;
    const unsigned char *disable_adiabatic;
    const double *l_turb_mpc;
    const double *dv_imc;
    const double *cs;
    const double *beta_pl;
    const double *tracer_mass_msun;
    const double *crp_state_cell_major;
    const double *cre_state_cell_major;
    const double *rad_ic_batch; /* [npe * ncell] momentum-major */
    const double *rad_ic_m1;
    const double *rad_ic_p1;
    int device_interp_background;
    double interp_alpha;
    const double *n_gas_next;
    const double *kbt_next;
    const double *b_field_next;
    const double *divv_gyr_next;
    const double *l_turb_mpc_next;
    const double *dv_imc_next;
    const double *cs_next;
    const double *beta_pl_next;
    const double *rad_ic_batch_next;
    const double *rad_ic_m1_next;
    const double *rad_ic_p1_next;
} CromaB1CoeffBatchInput;

/* FpSynchEmissionBatchInput (fp_shared_core.h:466-490) */
typedef struct {
    int ncell;
    int nfreq;
    int nx_tab;
    int ntheta_pitch;
    int nlogb;
    double z;
    double xmin;
    double logb_min;
    double inv_dlogb;
    const double *fx_tab;
    const double *logfx_tab;
    const double *logx_tab;
    double ***logy;                   /* [nfreq][npe][ntheta_pitch] */
    const double *pitch_kernel_table; /* [nlogb * nfreq * npe] or NULL */
    const double *b_dyn;
    const double *logb;
    const double *cre_batch;
    const double *gamma2e;
    const double *theta;
    const double *dtheta;
    const double *pitch_weight;
    const double *nus;
} CromaB1SynchEmissionBatchInput;

/* FpGammaEmissionBatchInput / FpNeutrinoEmissionBatchInput (fp_shared_core.h:505-530) */
typedef struct {
    int ncell;
    int nbins;
    double z;
    double e_min; /* log10 GeV */
    double e_max;
    const double *n_gas;
    const double *crp_batch;
    const double *beta_p;
    const double *f_flat; /* [nbins * np] */
} CromaB1HadronicEmissionBatchInput;

/* FpChangCooperSolveBatchInput (fp_shared_core.h:680-690) */
typedef struct {
    int nsys;
    int nrow;
    double dt;
    const double *a_batch;
    const double *b_batch;
    const double *c_batch;
    const double *source_batch;
} CromaB1SolveBatchInput;

typedef struct CromaB1Workspace CromaB1Workspace; /* FpCudaPipelineWorkspace */

/* FpCudaPipelineInput (fp_cuda_backend.h:27-67) */
typedef struct {
    int ncell;
    int nstep;
    int fp_cadence_steps;
    int diff_cadence_steps;
    int adv_cadence_steps;
    int use_windowed_reacc;
    int on_start_step;
    int on_end_step;
    int nfreq;
    int nx_tab;
    double dt;
    double epmax;
    double xmin;
    double psi_value;
    double mach_limit_value;
    int transport_mode; /* FpCudaTransportMode */
    const void *crp_grid; /* const CRspectrum * */
    const void *cre_grid;
    const CromaB1CoeffBatchInput *coeff_in;
    const CromaB1SynchEmissionBatchInput *synch_in;
    const CromaB1HadronicEmissionBatchInput *gamma_in;
    const CromaB1HadronicEmissionBatchInput *neutrino_in;
    const double *qpi_batch;
    const double *qepri_batch;
    const double *beta_p;
    const double *crp_init;
    const double *cre_init;
    double *crp_state_out;
    double *cre_state_out;
    double *eps_syn_out;
    double *eps_gamma_out;
    double *eps_nu_out;
    const int *rank_state_indices;
    int gather_rank_state;
    int scatter_rank_state;
    CromaB1Workspace *workspace;
} CromaB1PipelineInput;

/* FpCudaPipelineTimes (fp_cuda_backend.h:69-90) */
typedef struct {
    double setup_ms;
    double h2d_ms;
    double device_interp_ms;
    double coeff_ms;
    double secondary_ms;
    double solve_ms;
    double diff_ms;
    double adv_ms;
    double synch_ms;
    double gamma_ms;
    double neutrino_ms;
    double d2h_ms;
    double other_ms;
    double total_ms;
    int emitted_synch;
    int emitted_gamma;
    int emitted_neutrino;
    int last_synch_step;
    int used_pcr_proton;
    int used_pcr_electron;
} CromaB1PipelineTimes;

/* Physics parameters the reference reads from globals (params.c:67-88). */
int croma_hip_b1_set_params(const CromaHipParams *prm);

/* cuda_pipeline_workspace_create / _destroy (fp_cuda_backend.h:121-122) */
CromaB1Workspace *croma_hip_b1_workspace_create(void);
void croma_hip_b1_workspace_destroy(CromaB1Workspace *ws);

/* cuda_rank_state_upload / _download (fp_cuda_backend.h:124-136): the
 * rank's whole state [ncell * np], [ncell * npe] cell-major. */
int croma_hip_b1_rank_state_upload(int np_, int npe_, CromaB1Workspace *ws, int ncell, const double *qpi,
                                   const double *qepri, const double *crp, const double *cre, double *elapsed_ms);
int croma_hip_b1_rank_state_download(int np_, int npe_, CromaB1Workspace *ws, int ncell, double *crp, double *cre,
                                     double *elapsed_ms);

/* run_tracer_pipeline (fp_cuda_backend.h:138-145): one bucket segment --
 * coefficients of the bucket's backgrounds (optionally interpolated with
 * coeff_in->interp_alpha), the secondary source of its protons, then nstep
 * implicit substeps of both species with the on / off D_pp window
 * [on_start_step, on_end_step).  State comes from the inputs
 * (qpi_batch, qepri_batch, crp_init, cre_init: all or none), from the rank
 * state (gather_rank_state with rank_state_indices) or stays resident in the
 * workspace from the previous call of the same ncell; it goes to
 * crp_state_out / cre_state_out when given and to the rank state with
 * scatter_rank_state.  As in the reference, gamma / neutrino outputs are
 * rejected; emission inside the pipeline (a schedule that emits) and
 * cadences other than 1 are not served (-1): the call site passes an empty
 * schedule and unit cadences (include/tracer_fp_solve.inc:1418-1471). */
int croma_hip_b1_run_tracer_pipeline(int np_, int npe_, const CromaB1PipelineInput *in,
                                     const CromaB1OutputSchedule *schedule, double *crp_state_out,
                                     double *cre_state_out, double *eps_syn_out, double *eps_gamma_out,
                                     double *eps_nu_out, CromaB1PipelineTimes *times);
/* run_tracer_pipeline_group (fp_cuda_backend.h:147-150): the inputs one after
 * the other, times summed. */
int croma_hip_b1_run_tracer_pipeline_group(int np_, int npe_, const CromaB1PipelineInput *inputs, int ninput,
                                           const CromaB1OutputSchedule *schedule, CromaB1PipelineTimes *times);

/* solve_cc_cuda_batch (fp_cuda_backend.h:112-116): Thomas on the device;
 * *used_pcr_out = 0. */
int croma_hip_b1_solve_cc_batch(const CromaB1SolveBatchInput *in, const double *state_init, double *state_out,
                                double *elapsed_ms, int *used_pcr_out);
/* fp_cuda_tridiag_uses_pcr / fp_cuda_tridiag_solver_name (fp_cuda_backend.h:118-119):
 * this backend's batched solve is Thomas at every size (0, "Thomas"). */
int croma_hip_b1_tridiag_uses_pcr(int nrow);
const char *croma_hip_b1_tridiag_solver_name(int nrow);

/* prepare_{synch,gamma,neutrino}_emission_cuda_batch (fp_cuda_backend.h:97-110),
 * FP64.  Synchrotron needs the pitch-kernel table: in->pitch_kernel_table
 * when given, else built on the device from logy / fx_tab when nlogb > 1 and
 * inv_dlogb > 0; the per-cell direct form (no table) returns -1. */
int croma_hip_b1_prepare_synch_emission(int np_, int npe_, const CromaB1SynchEmissionBatchInput *in,
                                        const void *cre_grid, double *eps_syn_out, double *elapsed_ms);
int croma_hip_b1_prepare_gamma_emission(int np_, int npe_, const CromaB1HadronicEmissionBatchInput *in,
                                        const void *crp_grid, double *eps_gamma_out, double *elapsed_ms);
int croma_hip_b1_prepare_neutrino_emission(int np_, int npe_, const CromaB1HadronicEmissionBatchInput *in,
                                           const void *crp_grid, double *eps_nu_out, double *elapsed_ms);

/* Byte layout of the structs above: sizeof (field "") or offsetof(field) of
 * type name (the reference's names: "FpCudaPipelineInput", ...); (size_t)-1
 * for an unknown name. */
size_t croma_hip_b1_layout(const char *type_name, const char *field);

#ifdef __cplusplus
}
#endif
#endif
