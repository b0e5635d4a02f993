/*
 * croma_hip.h -- C ABI of libcroma_hip.so, the MI355X (gfx950) backend for
 * CROMA's per-tracer Fokker-Planck update.
 *
 * Conventions (kept from the reference backend): every function returns 0 on
 * success; caller-owned host arrays are read (or written) only during the
 * call; handles are opaque.  Changed from the reference: errors are negative
 * return codes (the library never calls exit()), physics parameters live in
 * a per-handle struct instead of process globals, and the grid size n is a
 * runtime value (64, 128, 256 or 512; both species use the same n as the
 * reference's CRspectrum layout requires, params.h:52-61).
 *
 * Array layouts use the reference names: "cell-major" = [tracer][bin];
 * "momentum-major" = [bin][tracer].
 */
#ifndef CROMA_HIP_H
#define CROMA_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------- */
#define CROMA_HIP_OK 0
#define CROMA_HIP_EINVAL (-1)      /* invalid argument (reference: -1)      */
#define CROMA_HIP_EPIVOT (-2)      /* Thomas pivot |denom| <= 1e-300 (-2)   */
#define CROMA_HIP_ENOMEM (-3)      /* host or device allocation failed      */
#define CROMA_HIP_EDEVICE (-4)     /* HIP runtime / kernel launch failure   */
#define CROMA_HIP_ENONFINITE (-5)  /* non-finite spectrum after a snapshot  */
#define CROMA_HIP_EBUDGET (-6)     /* request exceeds the server HBM budget */
#define CROMA_HIP_ECANCELED (-7)   /* request was cancelled                 */
#define CROMA_HIP_EUNSUPPORTED (-8)/* configuration this build does not run */
#define CROMA_HIP_EBUSY (-9)       /* result not ready yet                   */
#define CROMA_HIP_EIO (-10)        /* file error or file that does not match */

/* Backend selector value for the reference's `backend = auto|cpu|cuda`
 * switch (tracer_fp_setup.c:401-416, include/tracer_fp_types.h:173-177). */
#define CROMA_HIP_BACKEND_NAME "hip"
int croma_hip_backend_parse(const char *spec); /* 1 if spec names this backend */

/* ---- device binding (fp_cuda_backend.h:92-95) ------------------------ */
/* Replaces cuda_pipeline_is_available. */
int croma_hip_is_available(void);
/* Replaces cuda_bind_local_rank: binds device local_rank % device_count. */
int croma_hip_bind_local_rank(int local_rank, int *device_out);
/* Replaces cuda_get_binding_state. */
int croma_hip_get_binding_state(int *device_count_out, int *current_device_out);
/* Replaces cuda_get_current_device_id. */
int croma_hip_get_current_device_id(char *device_id_out, int device_id_out_len);
const char *croma_hip_last_error(void);

/* ---- configuration -------------------------------------------------- */
typedef struct {
    int n;          /* bins per species: 64, 128, 256 or 512 */
    double pmin;    /* log10 proton p/(m_p c) bounds (params.c:51) */
    double pmax;
    double pemin;   /* log10 electron p/(m_e c) bounds (params.c:52) */
    double pemax;
} CromaHipGridSpec;

/* Physics and scheduling parameters; names follow the reference keys
 * (params.c:67-88).  dpp_model: 1 asa, 2 ttd, 3 off, 4 direct_tacc. */
typedef struct {
    int dpp_model;
    int ttd_tacc_model;          /* 0 spectrum, 1 brunetti16 */
    int reacc_window_mode;       /* 0 always on, 1 centred window */
    int tracer_nsub_max;         /* 0 = no cap */
    int coeff_interp_min_steps;  /* 4 */
    int coeff_interp_max_segments; /* 16 */
    double psi;                  /* 0.5 */
    double mach_limit;           /* 0.5 */
    double eta_dpp_cap;          /* <= 0 disables */
    double t_acc_direct_gyr;
    double t_off_gyr;
    double t_acc_off_gyr;        /* 3.0 */
    double f_eddy;               /* 2.0 */
    double tracer_nsub_safety;   /* 0.25 */
    double epmax;                /* 1e7 GeV */
} CromaHipParams;

void croma_hip_default_params(CromaHipParams *prm);

/* Execution path. */
#define CROMA_HIP_PATH_FUSED 0   /* persistent lane-group kernel (product) */
#define CROMA_HIP_PATH_PLAIN 1   /* one tracer per thread, per-bucket launches */

/* Launch shape of the fused kernel. */
typedef struct {
    int rows_per_lane;    /* 2, 4, 8 or 16 spectrum rows per lane: n / rows_per_lane lanes (4..64) per tracer */
    int waves_per_simd;   /* 1 or 2: register budget of the kernel and resident persistent waves per SIMD */
    int chain;            /* batch runs: snapshots advanced per queue task (1 = one launch per snapshot) */
    int secondary_gemm;   /* two species: 1 = proton pass, one FP64 MFMA contraction of all
                           * segment-start proton rows against the secondary weights, electron
                           * pass; 0 = contraction inside the fused kernel (always used when the
                           * eta cap couples the species) */
} CromaHipTile;

/* Stage timers, the subset of TracerFpGpuTimes (include/tracer_fp_types.h:10-57)
 * this backend fills. */
typedef struct {
    double nsub_estimate_ms;
    double bucket_build_ms;   /* task ordering / bucketing */
    double ic_rows_ms;        /* IC cooling rows (host) */
    double h2d_ms;
    double solve_ms;          /* device time of the snapshot kernels */
    double d2h_ms;
    double total_ms;
    long long cell_steps;     /* tracer-substeps: sum of nsub over tracers */
    long long species_steps;  /* sum of nsub x active species */
    double kernel_ms;         /* device time inside the fused snapshot kernels */
    long long kernel_launches;/* fused snapshot kernel launches timed in kernel_ms */
    long long graph_launches; /* windows replayed from a captured hipGraph */
    long long graph_captures; /* hipGraphs captured */
    long long graph_failures; /* captures the runtime refused (those windows ran as plain launches) */
} CromaHipTimes;

/* ---- context: device, grid tables, streams (cuda_pipeline_workspace_*,
 * fp_cuda_backend.h:121-122) ----------------------------------------- */
typedef struct CromaHipContext CromaHipContext;

CromaHipContext *croma_hip_context_create(const CromaHipGridSpec *grid, const CromaHipParams *prm);
void croma_hip_context_destroy(CromaHipContext *ctx);
int croma_hip_context_set_params(CromaHipContext *ctx, const CromaHipParams *prm);
/* protons: 0 electron_only tile, 1 electron_proton tile, -1 both. */
int croma_hip_context_set_tile(CromaHipContext *ctx, int protons, const CromaHipTile *tile);
int croma_hip_context_get_tile(const CromaHipContext *ctx, int protons, CromaHipTile *tile);
/* 1 (default): a snapshot window whose launch shape repeats the previous
 * window of its buffer set replays a captured hipGraph of its launches
 * (fill, fused kernel, contraction, fused kernel); 0: plain stream launches. No
 * reference counterpart (the reference launches per bucket and segment). */
int croma_hip_context_set_graphs(CromaHipContext *ctx, int enable);
/* 1: in batch runs with the split secondary contraction, the electron pass of each
 * snapshot window runs on a second stream beside the proton pass of the next
 * window, each on half the device; 0 (default): the passes of a window run
 * in turn on the full device.  The spectra are the same either way; which is
 * faster depends on the tile.  No reference counterpart (the reference runs
 * one species pass at a time). */
int croma_hip_context_set_pipeline(CromaHipContext *ctx, int enable);
/* 1 (default): batch runs estimate nsub (tracer_fp_nsub.c:94-282), order the
 * queue by cost and assign the split segment rows on the device, without a
 * host pass over the tracers; 0: on the host (the same integers). */
int croma_hip_context_set_nsub_device(CromaHipContext *ctx, int enable);
/* Precision of the coefficient evaluation of the fused path (the FP32 mode of
 * north-star subsystem (a); the reference evaluates FP_Coef.c:438-849 and
 * fp_shared_core.c:729-1081 in FP64).  CROMA_HIP_COEFF_FP64 (default):
 * spectra within 1e-10 of the reference CPU path.  CROMA_HIP_COEFF_FP32: loss
 * rates, D_pp rows, the drift B and diffusion C of every row and the face
 * averages and Peclet numbers in FP32, with the Bernoulli face weights, the
 * rows a, b, c, the Thomas factorisation and every substep in FP64; spectra
 * then carry the FP32 coefficients' relative error (measured against the
 * FP64 oracle in the tests).  Applies to batch runs and the drop-in; the plain
 * path stays FP64. */
#define CROMA_HIP_COEFF_FP64 0
#define CROMA_HIP_COEFF_FP32 1
int croma_hip_context_set_coeff_precision(CromaHipContext *ctx, int precision);
/* Grid arrays (momentum_bin, FP_Coef.c:291-330); species 0 proton, 1 electron. */
int croma_hip_context_grid(const CromaHipContext *ctx, int species, double *p, double *dp, double *E,
                           double *pm1, double *pp1);
/* Autotune the fused kernel's tile for (protons) on a synthetic probe of
 * ntracer tracers; stores the best tile in the context and returns it. */
int croma_hip_autotune(CromaHipContext *ctx, int protons, int ntracer, CromaHipTile *best);

/* ---- snapshot-level drop-in ----------------------------------------- */
/* Same parameters and meaning as tracer_fp_batch_execute_parallel
 * (include/tracer_fp_batch.h:152-197, tracer_fp_batch.c:829-876), with the
 * grids held by ctx.  rad_ic_zero is momentum-major [n * ntracer] and must be
 * the broadcast row the reference builds (tracer_fp_synch.c:284-291).
 * fqe_flat is the transposed secondary weight [n_p * n_e] with np_min_qe.
 * base_nsteps / fp_cadence_steps / integration_mode are accepted for ABI
 * compatibility; only the quantized mode (base == nsub, cadence 1) exists.
 * Host arrays in, host arrays out; runs the fused path. */
int croma_hip_execute_snapshot(
    CromaHipContext *ctx, int ntracer, const int *nsubsteps, const int *n_onsteps,
    const int *base_nsteps, const int *fp_cadence_steps, int integration_mode,
    int allow_snapshot_interp, double elapsed_gyr, double dt_snap, double z_curr, double z_next,
    double requested_t_acc_direct_gyr,
    const double *n_gas_snap, const double *kbt_snap, const double *b_field_snap,
    const double *divv_snap, const double *lturb_snap, const double *dv_snap,
    const double *cs_snap, const double *beta_snap,
    const double *rad_ic_zero, const double *rad_ic_m1_zero, const double *rad_ic_p1_zero,
    const double *n_gas_next_snap, const double *kbt_next_snap, const double *b_field_next_snap,
    const double *divv_next_snap, const double *lturb_next_snap, const double *dv_next_snap,
    const double *cs_next_snap, const double *beta_next_snap,
    const double *rad_ic_row_next, double rad_ic_m1_next_val, double rad_ic_p1_next_val,
    const double *qpi_batch, const double *qepri_batch, const double *fqe_flat,
    const int *np_min_qe, const double *tracer_mass, const unsigned char *disable_adiabatic,
    double *crp_state, double *cre_state, CromaHipTimes *times,
    int *bucket_calls_out, int *min_bucket_size_out, int *max_bucket_size_out);

/* nsub estimate from the current snapshot (tracer_fp_prepare_nsub_loss_endpoints
 * + tracer_fp_compute_nsubsteps_from_endpoints, tracer_fp_nsub.c:94-282);
 * rad_ic_lo / rad_ic_hi are the IC cooling row at bins 0 and n-1. */
int croma_hip_nsub_estimate(const CromaHipContext *ctx, int ntracer, double dt_snap,
                            const double *n_gas, const double *kbt, const double *b_field,
                            const double *divv, double rad_ic_lo, double rad_ic_hi,
                            const double *l_turb, const double *dv, int *nsubsteps, int *n_onsteps);
/* The same estimate computed on the device (the batch runs' device step
 * before the solve; host arrays in and out). */
int croma_hip_nsub_estimate_device(const CromaHipContext *ctx, int ntracer, double dt_snap,
                                   const double *n_gas, const double *kbt, const double *b_field,
                                   const double *divv, double rad_ic_lo, double rad_ic_hi,
                                   const double *l_turb, const double *dv, int *nsubsteps, int *n_onsteps);
/* IC cooling row at redshift z (prepare_ic_cooling_row, fp_shared_core.c:844-864). */
int croma_hip_ic_cooling_row(const CromaHipContext *ctx, double z, double *row, double *m1, double *p1);
/* The same rows built on the device (the batch path's rows): rows [nz][n],
 * ghosts [nz][2] (bins at epm1, epp1; may be NULL). */
int croma_hip_ic_cooling_rows_device(CromaHipContext *ctx, int nz, const double *z, double *rows, double *ghosts);
/* Secondary-electron weights, transposed layout (prepare_secondary_kernel_flat,
 * fp_shared_core.c:1453-1506), built on the device once per context
 * (QeKernel_BB05 quadrature per entry, HADRONIC.c:436-498) and kept there for
 * the secondary-source contraction; this call copies them out. */
int croma_hip_secondary_weights(const CromaHipContext *ctx, double *fqe_t, int *np_min_qe);
/* Emission kernel tables [nbins][n] built on the device over the log10 range
 * lo..hi (the reference builds them on the CPU: prepare_gamma_kernel_flat,
 * prepare_neutrino_kernel_flat, prepare_ic_kernel_flat, fp_shared_core.c:
 * 1724-1804): kind CROMA_HIP_TABLE_GAMMA (E_gamma in GeV, proton grid),
 * CROMA_HIP_TABLE_NEUTRINO (E_nu in GeV, proton grid) or CROMA_HIP_TABLE_IC
 * (nu in Hz at redshift z, electron grid).  table: host [nbins][n]. */
#define CROMA_HIP_TABLE_GAMMA 0
#define CROMA_HIP_TABLE_NEUTRINO 1
#define CROMA_HIP_TABLE_IC 2
int croma_hip_emission_kernel_table(const CromaHipContext *ctx, int kind, int nbins, double lo, double hi,
                                    double z, double *table, double *elapsed_ms);

/* ---- device-resident batch (cuda_rank_state_upload/download +
 * run_tracer_pipeline, fp_cuda_backend.h:124-150) --------------------- */
typedef struct CromaHipBatch CromaHipBatch;

#define CROMA_HIP_SPECIES_ELECTRON 1          /* electron_only seed: crp == 0, only electrons evolve */
#define CROMA_HIP_SPECIES_PROTON 2            /* proton_only seed: both species evolve (secondaries) */
#define CROMA_HIP_SPECIES_ELECTRON_PROTON 3   /* electron_proton seed: both species evolve */

/* nsnap_resident: number of snapshot background sets kept in HBM. */
CromaHipBatch *croma_hip_batch_create(CromaHipContext *ctx, int ntracer, int species, int nsnap_resident);
void croma_hip_batch_destroy(CromaHipBatch *b);
/* Upload state and per-tracer constants (cuda_rank_state_upload). crp may be
 * NULL for electron_only; qpi/qepri may be NULL (zero). */
int croma_hip_batch_set_state(CromaHipBatch *b, const double *crp, const double *cre, const double *qpi,
                              const double *qepri, const double *tracer_mass,
                              const unsigned char *disable_adiabatic);
/* Download state (cuda_rank_state_download). */
int croma_hip_batch_get_state(CromaHipBatch *b, double *crp, double *cre);
/* Backgrounds of snapshot slot isnap (8 arrays [ntracer]: n_gas [cm^-3],
 * kbt [GeV], b_field [G], divv [1/Gyr], l_turb [Mpc], dv [cm/s], cs [cm/s], beta_pl). */
int croma_hip_batch_set_background(CromaHipBatch *b, int isnap, const double *const fields[8]);
/* Per-snapshot sources (DSA source mode adds S/dt into qpi/qepri,
 * tracer_fp_dsa.c:470-505); NULL clears the snapshot's sources. */
int croma_hip_batch_set_sources(CromaHipBatch *b, int isnap, const double *qpi, const double *qepri);
/* Run snapshots [isnap_begin, isnap_end): z[] and dt_snap[] indexed by
 * snapshot (z needs isnap_end + 1 entries).  Snapshot s interpolates towards
 * s+1 when s + 1 < nsnap_total (allow_snapshot_interp, solve.inc:10-12). */
int croma_hip_batch_run(CromaHipBatch *b, int isnap_begin, int isnap_end, int nsnap_total, const double *z,
                        const double *dt_snap, double elapsed_gyr, int path, CromaHipTimes *times);
/* Wait for the batch's device work. */
int croma_hip_batch_sync(CromaHipBatch *b);
/* Autotune the fused kernel's tile on the batch's own snapshot s: every
 * candidate (rows per lane x waves per SIMD, x split secondary for two
 * species) advances snapshot s from the current state, which is restored
 * after each trial and at the end.  Then, where snapshot s + 1 is resident
 * with its successor, the winner advances snapshots s and s + 1 with one
 * and with two snapshots per queue task, and keeps two if that is more than
 * 1% faster.  Stores the fastest tile in the context and returns it.  No
 * reference counterpart (the reference has no tiles). */
int croma_hip_batch_autotune(CromaHipBatch *b, int isnap, int nsnap_total, const double *z,
                             const double *dt_snap, double elapsed_gyr, CromaHipTile *best);

/* Batched Chang-Cooper solve (solve_cc_cuda_batch, fp_cuda_backend.h:112-116):
 * state_out = Thomas(a, b, c; dt*source + state_init) per system, cell-major
 * [nsys * nrow] host arrays. */
int croma_hip_solve_cc_batch(int nsys, int nrow, double dt, const double *a, const double *b,
                             const double *c, const double *source, const double *state_init,
                             double *state_out, double *elapsed_ms);

/* Allocation-free variant of croma_hip_solve_cc_batch.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernel and arithmetic, but a, b, c, source, state_init
 * and state_out are DEVICE pointers ([nsys * nrow] cell-major FP64, caller
 * owned; state_out may equal state_init), the work runs on `stream` (a
 * hipStream_t; NULL is the null stream) and the call returns once that
 * stream has drained.  The solver's scratch (the Thomas c' workspace and the
 * pivot-failure word) is a static device array of the library, one per GPU,
 * held until the process ends: nsys * nrow may not exceed
 * CROMA_HIP_NOALLOC_MAX_CELLS (32 MiB of c'), else CROMA_HIP_EINVAL.  Calls
 * that overlap on one GPU share that array and race.  Returns 0,
 * CROMA_HIP_EPIVOT, CROMA_HIP_EINVAL or CROMA_HIP_EDEVICE. */
#define CROMA_HIP_NOALLOC_MAX_CELLS (4 * 1024 * 1024)
int croma_hip_solve_cc_batch_noalloc(int nsys, int nrow, double dt, const double *a, const double *b,
                                     const double *c, const double *source, const double *state_init,
                                     double *state_out, void *stream);

/* ---- emission (prepare_{synch,gamma,neutrino}_emission_cuda_batch,
 * fp_cuda_backend.h:97-110, plus IC which the reference runs on the CPU) */
#define CROMA_HIP_EMIS_FP32 0        /* FP32 VALU contraction */
#define CROMA_HIP_EMIS_MFMA_XF32 1   /* split-BF16 x3 on MFMA, FP32 accumulate */
#define CROMA_HIP_EMIS_MFMA_BF16 2   /* single BF16 pass on MFMA (fastest, ~1e-3) */
#define CROMA_HIP_EMIS_FP64 3        /* FP64 reference on the GPU */

/* IC: eps[i][k] = sum_j fic[k][j] cre[i][j] dp[j]  (fp_shared_core.c:2131-2157). */
int croma_hip_emission_ic(const CromaHipContext *ctx, int ntracer, int nbins, const double *cre,
                          const double *fic_flat, double *eps, int mode, double *elapsed_ms);
/* gamma / neutrino: eps[i][k] = E_k * n_gas[i] * 0.5 * sum trapezoid(beta N f) dp
 * (fp_shared_core.c:2166-2243, EMISSION.c:297-308).  emin/emax log10 GeV. */
int croma_hip_emission_hadronic(const CromaHipContext *ctx, int ntracer, int nbins, double emin,
                                double emax, const double *n_gas, const double *crp,
                                const double *f_flat, double *eps, int mode, double *elapsed_ms);
/* Synchrotron pitch-kernel table [nlogb][nfreq][n], built on the device
 * (prepare_synch_pitch_kernel_table, fp_shared_core.c:2008-2052): the inputs
 * are the reference's SynchData tables (tracer_fp_synch.c:87-121): the log
 * F(x) table logx_tab / logfx_tab / fx_tab of nx_tab nodes from xmin, logy
 * [nfreq][n][ntheta] and the pitch weights [ntheta]. */
int croma_hip_synch_kernel_table(const CromaHipContext *ctx, int nfreq, int ntheta, int nlogb, double logb_min,
                                 double inv_dlogb, int nx_tab, double xmin, const double *logx_tab,
                                 const double *logfx_tab, const double *fx_tab, const double *logy,
                                 const double *pitch_weight, double *table, double *elapsed_ms);
/* Synchrotron from a pitch-kernel table [nlogb][nfreq][n] with log
 * interpolation in log B (fp_shared_core.c:2070-2129). */
int croma_hip_emission_synch(const CromaHipContext *ctx, int ntracer, int nfreq, int nlogb,
                             double logb_min, double inv_dlogb, const double *table, const double *cre,
                             const double *b_field, double *eps, double *elapsed_ms);
/* The same with a precision mode: CROMA_HIP_EMIS_FP64 (as above) or
 * CROMA_HIP_EMIS_FP32 (terms as 2^(log2|N dp| + log2 K) summed in FP32
 * relative to a running largest exponent, float log2 table, hardware exp2;
 * emissivities within 1e-4).  Replaces
 * prepare_synch_emission_cuda_batch (fp_cuda_backend.h:97-100) in its FP32
 * form. */
int croma_hip_emission_synch_mode(const CromaHipContext *ctx, int ntracer, int nfreq, int nlogb,
                                  double logb_min, double inv_dlogb, const double *table, const double *cre,
                                  const double *b_field, double *eps, int mode, double *elapsed_ms);

/* Synchrotron on MFMA with the table's device form kept across calls: a plan
 * builds the Chebyshev coefficients of the log-B interpolation (8 terms per
 * plane interval, bin and frequency, split into two BF16 parts) and the float
 * log2 table of the range fix-up once; each run then sorts its tracers by
 * plane, contracts them on MFMA (split-BF16 x3, FP32 accumulate, 2^59 operand
 * scales) and recomputes in the FP32 log-domain form every tracer whose
 * emissivities within 25 decades of its peak fall below the contraction's
 * range (emissivities within 1e-4 of the FP64 sum).  Needs n <= 256 and
 * planes no coarser than 0.02 in log10 B.  The same computation as
 * croma_hip_emission_synch_mode(CROMA_HIP_EMIS_MFMA_XF32), which builds and
 * drops a plan per call.  A plan belongs to one thread at a time. */
typedef struct CromaHipSynchPlan CromaHipSynchPlan;
CromaHipSynchPlan *croma_hip_synch_plan_create(const CromaHipContext *ctx, int nfreq, int nlogb, double logb_min,
                                               double inv_dlogb, const double *table, double *elapsed_ms);
int croma_hip_synch_plan_run(CromaHipSynchPlan *plan, int ntracer, const double *cre, const double *b_field,
                             double *eps, double *elapsed_ms);
void croma_hip_synch_plan_destroy(CromaHipSynchPlan *plan);

/* ---- output and checkpoint files (tracer_fp_output.c, tracer_fp_restart.c)
 * The reference's byte layouts, float64 native-endian:
 *   CROMA_HIP_OUT_MAP   <dir>/<stem>_core%02d.bin       [tracer][slot][nbin]
 *                       (tracer_fp_map_open :613-664; the same file and
 *                       layout as the buffered CRE/CRP_core%02d.bin writer
 *                       output_CRspectrum_buffer_node :38-72)
 *   CROMA_HIP_OUT_TILE  <dir>/<stem>_tile_core%02d.bin  [slot][tracer][nbin]
 *                       (tracer_fp_tile_open :353-416)
 * Spectra use nslot = nsnap + 1 (slot 0 = initial state), emissivities
 * nslot = nsnap; stems CRE, CRP, eSyn, eIC, eGamma, eNu.  Files are sized
 * at open (resume_existing: the file must exist with that size). -------- */
#define CROMA_HIP_OUT_MAP 0
#define CROMA_HIP_OUT_TILE 1
typedef struct CromaHipOutput CromaHipOutput;
CromaHipOutput *croma_hip_output_open(const char *dir, const char *stem, int rank, int mode, int ntracer,
                                      int nbin, int nslot, int resume_existing);
/* Rows [ntracer][nbin] (host) into slot.  rho_gcc and tracer_mass both
 * non-NULL select per-cm^3 output: rows scaled by rho / (mass M_sun), 0 for a
 * tracer without positive rho and mass (output_volume_factor :175-188). */
int croma_hip_output_write(CromaHipOutput *out, int slot, const double *rows, const double *rho_gcc,
                           const double *tracer_mass);
/* The batch's device-resident spectra into slot of cre_out (and crp_out,
 * may be NULL): staged through pinned host memory in chunks, the copy of
 * one chunk overlapping the file write of the previous one
 * (tracer_fp_write_cr_map :249-296, tracer_fp_tile_write_cr :439-477). */
int croma_hip_batch_write_spectra(CromaHipBatch *b, CromaHipOutput *cre_out, CromaHipOutput *crp_out, int slot,
                                  const double *rho_gcc, const double *tracer_mass);
int croma_hip_output_close(CromaHipOutput *out);
/* <dir>/tracerid_core%02d.txt, one "%09d" id per line (output_tracerid_node :74-96). */
int croma_hip_write_tracer_ids(const char *dir, int rank, int ntracer, const int *ids);

/* Checkpoint TFPCHK1 (tracer_fp_ckpt_write / _read, tracer_fp_restart.c:136-265):
 * <dir>/checkpoint_rank%02d.bin = the 72-byte header (TracerFpCheckpointHeader,
 * include/tracer_fp_types.h:356-371), int ids[ntracer], crp[ntracer * np_bins],
 * cre[ntracer * npe_bins]; written to a .tmp file, synced and renamed.  Reads
 * check every header field and the tracer order against meta / ids. */
typedef struct {
    int mpi_rank, world_size;
    int ntracer, ntracer_global;
    int nsnap, nfreq;
    int np_bins, npe_bins;
    int nsnp_i, nsnp_f;
    int completed_snapshot; /* written; ignored by reads */
    long tracer_start;
} CromaHipCheckpointMeta;
int croma_hip_checkpoint_write(const char *dir, const CromaHipCheckpointMeta *meta, const int *ids,
                               const double *crp, const double *cre);
int croma_hip_checkpoint_read(const char *dir, const CromaHipCheckpointMeta *meta, const int *ids, double *crp,
                              double *cre, int *completed_snapshot);
/* The same from / into a batch's device state (np_bins = npe_bins = the grid). */
int croma_hip_batch_checkpoint_write(CromaHipBatch *b, const char *dir, const CromaHipCheckpointMeta *meta,
                                     const int *ids);
int croma_hip_batch_checkpoint_read(CromaHipBatch *b, const char *dir, const CromaHipCheckpointMeta *meta,
                                    const int *ids, int *completed_snapshot);

/* ---- DSA source builder (tracer_fp_apply_dsa_tracer_injection_snapshot,
 * tracer_fp_dsa.c:184-505; DSA_MODELS.c:394-1558) -----------------------
 * Shock-accelerated CR injection per tracer, with or without
 * reacceleration of the pre-shock spectrum: the proton spectrum A_p (p/p_inj)^-q exp(-(p/p_max)^2) normalised to
 * eta(M) of the upstream kinetic energy flux, the electrons at K_ep of the
 * protons at p_inj,p (suprathermal injection), on the FP grid as 4 pi p^2 f. */
#define CROMA_HIP_DSA_KR07 0   /* eta models, DSAModelID order (DSA_MODELS.h:37-47) */
#define CROMA_HIP_DSA_KR07R 1
#define CROMA_HIP_DSA_KR13 2
#define CROMA_HIP_DSA_KR13R 3
#define CROMA_HIP_DSA_CS14 4
#define CROMA_HIP_DSA_CS14R 5
#define CROMA_HIP_DSA_RYU19 6
#define CROMA_HIP_DSA_RYU19R 7
#define CROMA_HIP_DSA_SPECIES_BOTH 0
#define CROMA_HIP_DSA_SPECIES_ELECTRON 1
#define CROMA_HIP_DSA_SPECIES_PROTON 2
#define CROMA_HIP_DSA_STATE 1   /* add the fresh spectrum to the state (TRACER_STATE) */
#define CROMA_HIP_DSA_SOURCE 2  /* add spectrum / dt to qpi, qepri (TRACER_SOURCE) */
typedef struct {
    int eta_model;          /* DSAEtaModelInitial (params.c:97): RYU19 */
    int inject_species;     /* DSAInjectSpecies: both */
    double chi_p, chi_e;    /* DSAChiP, DSAChiE: 3.5, 3.5 (chi_e <= 0 uses chi_p) */
    double k_ep;            /* DSAKep: 1e-2 */
    double pmax_p, pmax_e;  /* DSAPmaxPmc, DSAPmaxEmc: 1e6, 1e5 (p / m c) */
    double gamma_gas;       /* DSAGammaGas: 5/3 */
    double min_mach;        /* DSAMinMach: 1.5 */
    double mu_mol;          /* mu_mol: 0.59 */
    double pmin_dsa_p_log10, pmin_dsa_e_log10; /* DSA grid start (tracer_fp.c:841): -3, -3 */
    /* reacceleration of the tracer's pre-shock spectrum (TracerDsaReaccMode,
     * include/tracer_fp_types.h:321-325; tracer_fp_dsa.c:322-418) */
    int reacc_mode;         /* CROMA_HIP_DSA_REACC_*: off */
    int eta_model_reacc;    /* DSAEtaModelReacc (params.c:98): RYU19R */
    double reacc_eta_cap;   /* DSAReaccEtaCap (params.c:110): 0 (<= 0: no cap) */
    double xcr_pmin_pmc;    /* DSAXcrPminPmc (params.c:109): 1 */
} CromaHipDsaParams;
#define CROMA_HIP_DSA_REACC_OFF 0
#define CROMA_HIP_DSA_REACC_POSITIVE_DELTA 1 /* eta interpolated in X_cr, extra energy as a positive delta */
#define CROMA_HIP_DSA_REACC_DRURY83 2        /* Drury 1983 convolution of the pre-shock spectrum */
void croma_hip_dsa_default_params(CromaHipDsaParams *p);
/* eta(M) of a model (dsa_eta_from_id). */
double croma_hip_dsa_eta(double mach, int eta_model);
/* eta between an initial and a reacceleration model, weighted by
 * x_cr / x_cr0 of the latter clipped to [0, 1] (dsa_eta_interpolated,
 * DSA_MODELS.c:502-522). */
double croma_hip_dsa_eta_interpolated(double mach, int initial_model, int reacc_model, double x_cr);
/* Source rows of one shock on a grid: source_p [n], source_e [n] (per unit
 * volume, 4 pi p^2 f), flags[0] / [1] = proton / electron source built.
 * u1 <= 0 derives it from mach and temp1. */
int croma_hip_dsa_shock_sources(const CromaHipGridSpec *grid, const CromaHipDsaParams *p, double mach,
                                double temp1_K, double rho1_cgs, double u1_cgs, double *source_p,
                                double *source_e, int *flags);
/* Per-tracer injection of one snapshot with shock data (the non-merged
 * input branch: mach, pre-shock temperature [K], density [g cm^-3], upstream
 * speed [cm/s] or NULL, side code or NULL (-1 and 99 skip)); or, with
 * kinetic_flux non-NULL, the merged branch (u1 = (2 F / rho1)^(1/3), temp1
 * from u1 and M).  tracer_mass [M_sun].  mode STATE adds V_down S to
 * crp / cre [ntracer * n]; SOURCE adds V_down S / dt_gyr to qpi / qepri.
 * Tracers with mach NaN or below min_mach are untouched.  *ninjected counts
 * the tracers that received a source. */
int croma_hip_dsa_inject(const CromaHipGridSpec *grid, const CromaHipDsaParams *p, int mode, int ntracer,
                         double dt_gyr, const double *mach, const double *temp1_K, const double *rho1_cgs,
                         const double *u1_cgs, const double *kinetic_flux, const double *side_code,
                         const double *tracer_mass, double *crp, double *cre, double *qpi, double *qepri,
                         int *ninjected);
/* One reacceleration row of a spectrum on a grid (tracer_fp_dsa_reacc.h:
 * tracer_dsa_reacc_positive_delta_1d, _convolution_delta_1d,
 * _adiabatic_delta_1d, _delta_energy_density_1d).  species 0 protons (the
 * proton grid), 1 electrons.  kind ENERGY takes the delta row as state and
 * only returns its energy density.  *edens: the function's return value. */
#define CROMA_HIP_DSA_REACC_KIND_POSITIVE 0
#define CROMA_HIP_DSA_REACC_KIND_DRURY 1
#define CROMA_HIP_DSA_REACC_KIND_ADIABATIC 2
#define CROMA_HIP_DSA_REACC_KIND_ENERGY 3
int croma_hip_dsa_reacc_delta(const CromaHipGridSpec *grid, int kind, int species, const double *state,
                              double volume_upstream, double volume_downstream, double q, double compression_ratio,
                              double pmin_mc, double *delta, double *edens);
/* croma_hip_dsa_inject with reacceleration (p->reacc_mode, tracer_fp_dsa.c:312-505 with
 * tracer_fp_dsa_reacc.c): the tracer's current spectra crp / cre (required
 * for any reacceleration mode, also in source mode; read before this
 * snapshot's injection is added) give the reacceleration delta, added like
 * the fresh source (state: V_down delta x scale; source: / dt).  Positive
 * delta: eta interpolated between the initial and the reacceleration model
 * in X_cr = P_cr / P_th (protons above xcr_pmin_pmc), the extra energy
 * (eta_interp - eta) M u1^2 / 2 caps the delta's energy.  Drury83: the
 * convolution delta, its non-adiabatic energy capped at
 * reacc_eta_cap M u1^2 / 2 when reacc_eta_cap > 0; each such tracer gets
 * disable_adiabatic[i] = 1 (NULL: not written), the flag the solve's
 * per-tracer dis input takes (include/tracer_fp_solve.inc:80-94). */
int croma_hip_dsa_inject_reacc(const CromaHipGridSpec *grid, const CromaHipDsaParams *p, int mode, int ntracer,
                               double dt_gyr, const double *mach, const double *temp1_K, const double *rho1_cgs,
                               const double *u1_cgs, const double *kinetic_flux, const double *side_code,
                               const double *tracer_mass, double *crp, double *cre, double *qpi, double *qepri,
                               unsigned char *disable_adiabatic, int *ninjected);

/* ---- serving: continuous batching of tracer chunks ------------------ */
typedef struct CromaHipServer CromaHipServer;

typedef struct {
    int ntracer;
    int n;                 /* grid size of this chunk */
    int species;           /* CROMA_HIP_SPECIES_* */
    int dpp_model;         /* overrides params.dpp_model for this chunk */
    int nsnap;             /* snapshots to advance */
    const double *z;       /* [nsnap + 1] */
    const double *dt_snap; /* [nsnap] */
    const double *bg;      /* [(nsnap + 1)][8][ntracer] backgrounds */
    const double *crp;     /* [ntracer * n] or NULL */
    const double *cre;     /* [ntracer * n] */
    const double *qpi;     /* [ntracer * n] or NULL */
    const double *qepri;   /* [ntracer * n] or NULL */
    const double *mass;    /* [ntracer] or NULL */
} CromaHipChunk;

typedef struct {
    double submit_s;       /* host submit time (monotonic clock) */
    double admit_s;        /* admitted into HBM */
    double done_s;         /* result visible on the host */
    long long cell_steps;  /* tracer-substeps of the request */
    int status;            /* 0 ok, else CROMA_HIP_E* */
} CromaHipChunkStats;

typedef struct {
    long long submitted, completed, cancelled, failed;
    long long cell_steps;
    size_t budget_bytes, in_use_bytes, peak_bytes;
    int active, queued;
    long long iterations;  /* scheduler iterations (batched launches) */
} CromaHipServeStats;

CromaHipServer *croma_hip_serve_create(const CromaHipParams *prm, size_t hbm_budget_bytes,
                                       double pmin, double pmax, double pemin, double pemax);
void croma_hip_serve_destroy(CromaHipServer *s);
/* Copies the chunk's inputs; returns the request id in *id_out. */
int croma_hip_serve_submit(CromaHipServer *s, const CromaHipChunk *chunk, long long *id_out);
/* 1 done, 0 pending, < 0 error code of the request. */
int croma_hip_serve_poll(CromaHipServer *s, long long id);
/* Blocks until done; copies the final spectra (crp may be NULL). */
int croma_hip_serve_wait(CromaHipServer *s, long long id, double *crp, double *cre, CromaHipChunkStats *st);
int croma_hip_serve_cancel(CromaHipServer *s, long long id);
int croma_hip_serve_stats(CromaHipServer *s, CromaHipServeStats *out);
/* Memory model used for admission (bytes of HBM a chunk holds). */
size_t croma_hip_chunk_bytes(int ntracer, int n, int species, int nsnap);

#ifdef __cplusplus
}
#endif

#endif
