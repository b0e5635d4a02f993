/*
 * ref_layout.c -- TEST INFRASTRUCTURE ONLY.  Prints the byte layout (sizeof
 * and offsetof) of the reference's GPU-backend structs as compiled from the
 * reference's own headers (include/fp_cuda_backend.h:12-90,
 * include/fp_shared_core.h:366-407, 466-530, 680-690, include/params.h:52-61)
 * at the grid size given by -Dnp / -Dnpe.  Built by oracle/ref.mk into
 * oracle/_ref/; tests/golden/make_b1_layout.py turns its output into the
 * fixture the layout test reads.  One line per entry: type field offset
 * (field "-" is sizeof).
 */
#include <stddef.h>
#include <stdio.h>

#include "fp_cuda_backend.h"

#define S(T) printf("%s - %zu\n", #T, sizeof(T))
#define F(T, f) printf("%s %s %zu\n", #T, #f, offsetof(T, f))

int main(void)
{
    S(FpCudaOutputSchedule);
    F(FpCudaOutputSchedule, emit_final_only); F(FpCudaOutputSchedule, emit_all_steps);
    F(FpCudaOutputSchedule, emit_stride); F(FpCudaOutputSchedule, physical_min);
    F(FpCudaOutputSchedule, physical_max); F(FpCudaOutputSchedule, selected_steps);
    F(FpCudaOutputSchedule, nselected_steps);
    S(FpCudaPipelineInput);
    F(FpCudaPipelineInput, ncell); F(FpCudaPipelineInput, nstep); F(FpCudaPipelineInput, fp_cadence_steps);
    F(FpCudaPipelineInput, diff_cadence_steps); F(FpCudaPipelineInput, adv_cadence_steps);
    F(FpCudaPipelineInput, use_windowed_reacc); F(FpCudaPipelineInput, on_start_step);
    F(FpCudaPipelineInput, on_end_step); F(FpCudaPipelineInput, nfreq); F(FpCudaPipelineInput, nx_tab);
    F(FpCudaPipelineInput, dt); F(FpCudaPipelineInput, epmax); F(FpCudaPipelineInput, xmin);
    F(FpCudaPipelineInput, psi_value); F(FpCudaPipelineInput, mach_limit_value);
    F(FpCudaPipelineInput, transport_mode); F(FpCudaPipelineInput, crp_grid); F(FpCudaPipelineInput, cre_grid);
    F(FpCudaPipelineInput, coeff_in); F(FpCudaPipelineInput, synch_in); F(FpCudaPipelineInput, gamma_in);
    F(FpCudaPipelineInput, neutrino_in); F(FpCudaPipelineInput, qpi_batch); F(FpCudaPipelineInput, qepri_batch);
    F(FpCudaPipelineInput, beta_p); F(FpCudaPipelineInput, crp_init); F(FpCudaPipelineInput, cre_init);
    F(FpCudaPipelineInput, crp_state_out); F(FpCudaPipelineInput, cre_state_out);
    F(FpCudaPipelineInput, eps_syn_out); F(FpCudaPipelineInput, eps_gamma_out); F(FpCudaPipelineInput, eps_nu_out);
    F(FpCudaPipelineInput, rank_state_indices); F(FpCudaPipelineInput, gather_rank_state);
    F(FpCudaPipelineInput, scatter_rank_state); F(FpCudaPipelineInput, workspace);
    S(FpCudaPipelineTimes);
    F(FpCudaPipelineTimes, setup_ms); F(FpCudaPipelineTimes, total_ms); F(FpCudaPipelineTimes, emitted_synch);
    F(FpCudaPipelineTimes, last_synch_step); F(FpCudaPipelineTimes, used_pcr_proton);
    F(FpCudaPipelineTimes, used_pcr_electron);
    S(FpCoeffBatchInput);
    F(FpCoeffBatchInput, ncell); F(FpCoeffBatchInput, z); F(FpCoeffBatchInput, dt_gyr); F(FpCoeffBatchInput, epmax);
    F(FpCoeffBatchInput, n_gas); F(FpCoeffBatchInput, kbt); F(FpCoeffBatchInput, b_field);
    F(FpCoeffBatchInput, divv_gyr); F(FpCoeffBatchInput, disable_adiabatic); F(FpCoeffBatchInput, l_turb_mpc);
    F(FpCoeffBatchInput, dv_imc); F(FpCoeffBatchInput, cs); F(FpCoeffBatchInput, beta_pl);
    F(FpCoeffBatchInput, tracer_mass_msun); F(FpCoeffBatchInput, crp_state_cell_major);
    F(FpCoeffBatchInput, cre_state_cell_major); F(FpCoeffBatchInput, rad_ic_batch); F(FpCoeffBatchInput, rad_ic_m1);
    F(FpCoeffBatchInput, rad_ic_p1); F(FpCoeffBatchInput, device_interp_background);
    F(FpCoeffBatchInput, interp_alpha); F(FpCoeffBatchInput, n_gas_next); F(FpCoeffBatchInput, kbt_next);
    F(FpCoeffBatchInput, b_field_next); F(FpCoeffBatchInput, divv_gyr_next); F(FpCoeffBatchInput, l_turb_mpc_next);
    F(FpCoeffBatchInput, dv_imc_next); F(FpCoeffBatchInput, cs_next); F(FpCoeffBatchInput, beta_pl_next);
    F(FpCoeffBatchInput, rad_ic_batch_next); F(FpCoeffBatchInput, rad_ic_m1_next);
    F(FpCoeffBatchInput, rad_ic_p1_next);
    S(FpSynchEmissionBatchInput);
    F(FpSynchEmissionBatchInput, ncell); F(FpSynchEmissionBatchInput, nlogb); F(FpSynchEmissionBatchInput, z);
    F(FpSynchEmissionBatchInput, inv_dlogb); F(FpSynchEmissionBatchInput, fx_tab); F(FpSynchEmissionBatchInput, logy);
    F(FpSynchEmissionBatchInput, pitch_kernel_table); F(FpSynchEmissionBatchInput, b_dyn);
    F(FpSynchEmissionBatchInput, cre_batch); F(FpSynchEmissionBatchInput, nus);
    S(FpGammaEmissionBatchInput);
    F(FpGammaEmissionBatchInput, ncell); F(FpGammaEmissionBatchInput, nbins); F(FpGammaEmissionBatchInput, z);
    F(FpGammaEmissionBatchInput, egamma_min); F(FpGammaEmissionBatchInput, egamma_max);
    F(FpGammaEmissionBatchInput, n_gas); F(FpGammaEmissionBatchInput, crp_batch); F(FpGammaEmissionBatchInput, beta_p);
    F(FpGammaEmissionBatchInput, fga_flat);
    S(FpNeutrinoEmissionBatchInput);
    F(FpNeutrinoEmissionBatchInput, enu_min); F(FpNeutrinoEmissionBatchInput, enu_max);
    F(FpNeutrinoEmissionBatchInput, fnu_flat);
    S(FpChangCooperSolveBatchInput);
    F(FpChangCooperSolveBatchInput, nsys); F(FpChangCooperSolveBatchInput, nrow); F(FpChangCooperSolveBatchInput, dt);
    F(FpChangCooperSolveBatchInput, a_batch); F(FpChangCooperSolveBatchInput, source_batch);
    S(CRspectrum);
    F(CRspectrum, p); F(CRspectrum, dp); F(CRspectrum, pp1); F(CRspectrum, pm1); F(CRspectrum, E);
    return 0;
}
