"""ctypes mirror of include/croma_hip_b1.h: the reference's GPU-backend structs
(fp_cuda_backend.h:12-90, fp_shared_core.h:366-407, 466-530, 680-690) and the
entry points that take them.  Field order and types follow the header, so
these Structures have the reference's byte layouts."""
from __future__ import annotations

import ctypes as C

import numpy as np

import croma_hip as ch

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)


class OutputSchedule(C.Structure):
    _fields_ = [("emit_final_only", C.c_int), ("emit_all_steps", C.c_int), ("emit_stride", C.c_int),
                ("physical_min", C.c_int), ("physical_max", C.c_int), ("selected_steps", _ip),
                ("nselected_steps", C.c_int)]


_COEFF_PTRS = ["n_gas", "kbt", "b_field", "divv_gyr"]
_COEFF_PTRS2 = ["l_turb_mpc", "dv_imc", "cs", "beta_pl", "tracer_mass_msun", "crp_state_cell_major",
                "cre_state_cell_major", "rad_ic_batch", "rad_ic_m1", "rad_ic_p1"]
_NEXT = ["n_gas_next", "kbt_next", "b_field_next", "divv_gyr_next", "l_turb_mpc_next", "dv_imc_next", "cs_next",
         "beta_pl_next", "rad_ic_batch_next", "rad_ic_m1_next", "rad_ic_p1_next"]


class CoeffBatchInput(C.Structure):
    _fields_ = ([("ncell", C.c_int), ("z", C.c_double), ("dt_gyr", C.c_double), ("epmax", C.c_double)] +
                [(f, _dp) for f in _COEFF_PTRS] + [("disable_adiabatic", _up)] +
                [(f, _dp) for f in _COEFF_PTRS2] +
                [("device_interp_background", C.c_int), ("interp_alpha", C.c_double)] + [(f, _dp) for f in _NEXT])


class SynchEmissionBatchInput(C.Structure):
    _fields_ = [("ncell", C.c_int), ("nfreq", C.c_int), ("nx_tab", C.c_int), ("ntheta_pitch", C.c_int),
                ("nlogb", C.c_int), ("z", C.c_double), ("xmin", C.c_double), ("logb_min", C.c_double),
                ("inv_dlogb", C.c_double), ("fx_tab", _dp), ("logfx_tab", _dp), ("logx_tab", _dp),
                ("logy", C.c_void_p), ("pitch_kernel_table", _dp), ("b_dyn", _dp), ("logb", _dp),
                ("cre_batch", _dp), ("gamma2e", _dp), ("theta", _dp), ("dtheta", _dp), ("pitch_weight", _dp),
                ("nus", _dp)]


class HadronicEmissionBatchInput(C.Structure):
    _fields_ = [("ncell", C.c_int), ("nbins", C.c_int), ("z", C.c_double), ("e_min", C.c_double),
                ("e_max", C.c_double), ("n_gas", _dp), ("crp_batch", _dp), ("beta_p", _dp), ("f_flat", _dp)]


class SolveBatchInput(C.Structure):
    _fields_ = [("nsys", C.c_int), ("nrow", C.c_int), ("dt", C.c_double), ("a_batch", _dp), ("b_batch", _dp),
                ("c_batch", _dp), ("source_batch", _dp)]


class PipelineInput(C.Structure):
    _fields_ = ([(f, C.c_int) for f in ("ncell", "nstep", "fp_cadence_steps", "diff_cadence_steps",
                                        "adv_cadence_steps", "use_windowed_reacc", "on_start_step", "on_end_step",
                                        "nfreq", "nx_tab")] +
                [(f, C.c_double) for f in ("dt", "epmax", "xmin", "psi_value", "mach_limit_value")] +
                [("transport_mode", C.c_int), ("crp_grid", C.c_void_p), ("cre_grid", C.c_void_p),
                 ("coeff_in", C.POINTER(CoeffBatchInput)), ("synch_in", C.POINTER(SynchEmissionBatchInput)),
                 ("gamma_in", C.POINTER(HadronicEmissionBatchInput)),
                 ("neutrino_in", C.POINTER(HadronicEmissionBatchInput))] +
                [(f, _dp) for f in ("qpi_batch", "qepri_batch", "beta_p", "crp_init", "cre_init", "crp_state_out",
                                    "cre_state_out", "eps_syn_out", "eps_gamma_out", "eps_nu_out")] +
                [("rank_state_indices", _ip), ("gather_rank_state", C.c_int), ("scatter_rank_state", C.c_int),
                 ("workspace", C.c_void_p)])


class PipelineTimes(C.Structure):
    _fields_ = ([(f, C.c_double) for f in ("setup_ms", "h2d_ms", "device_interp_ms", "coeff_ms", "secondary_ms",
                                           "solve_ms", "diff_ms", "adv_ms", "synch_ms", "gamma_ms", "neutrino_ms",
                                           "d2h_ms", "other_ms", "total_ms")] +
                [(f, C.c_int) for f in ("emitted_synch", "emitted_gamma", "emitted_neutrino", "last_synch_step",
                                        "used_pcr_proton", "used_pcr_electron")])


def crspectrum(p, dp, pm1, pp1, E):
    """A CRspectrum (params.h:52-61) of n bins as a float64 buffer:
    p[n], dp[n], pp1, pm1, E[n], N[n], N_pre[n], N_ave[n]."""
    n = len(p)
    buf = np.zeros(6 * n + 2)
    buf[:n] = p
    buf[n:2 * n] = dp
    buf[2 * n] = pp1
    buf[2 * n + 1] = pm1
    buf[2 * n + 2:3 * n + 2] = E
    return buf


def lib():
    L = ch.lib()
    L.croma_hip_b1_workspace_create.restype = C.c_void_p
    L.croma_hip_b1_workspace_destroy.argtypes = [C.c_void_p]
    L.croma_hip_b1_workspace_destroy.restype = None
    L.croma_hip_b1_run_tracer_pipeline.argtypes = [C.c_int, C.c_int, C.POINTER(PipelineInput),
                                                   C.POINTER(OutputSchedule), _dp, _dp, _dp, _dp, _dp,
                                                   C.POINTER(PipelineTimes)]
    L.croma_hip_b1_rank_state_upload.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.croma_hip_b1_rank_state_download.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, _dp, _dp, _dp]
    L.croma_hip_b1_set_params.argtypes = [C.POINTER(ch.Params)]
    L.croma_hip_b1_solve_cc_batch.argtypes = [C.POINTER(SolveBatchInput), _dp, _dp, _dp, _ip]
    L.croma_hip_b1_tridiag_solver_name.restype = C.c_char_p
    L.croma_hip_b1_prepare_gamma_emission.argtypes = [C.c_int, C.c_int, C.POINTER(HadronicEmissionBatchInput),
                                                      C.c_void_p, _dp, _dp]
    L.croma_hip_b1_prepare_neutrino_emission.argtypes = L.croma_hip_b1_prepare_gamma_emission.argtypes
    L.croma_hip_b1_prepare_synch_emission.argtypes = [C.c_int, C.c_int, C.POINTER(SynchEmissionBatchInput),
                                                      C.c_void_p, _dp, _dp]
    return L


def ptr(a):
    """float64 / int32 / uint8 pointer of a C-contiguous array (None -> NULL)."""
    return ch._p(a)
