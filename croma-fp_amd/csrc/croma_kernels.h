/*
 * croma_kernels.h -- thin C ABI between the C host runtime (the host directory) and the
 * HIP kernels (csrc/croma_kernels.hip).  Plain structs of device pointers and
 * scalars; no HIP or torch types.  Internal to libcroma_hip.so.
 */
#ifndef CROMA_KERNELS_H
#define CROMA_KERNELS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CK_MAX_N 512
/* added to a fused launch mode: the FP32-coefficient instance */
#define CK_LAUNCH_F32 4
/* added to launch mode 0 (electron_only): no slot of the launch has source
 * rows (qepri absent, no secondaries), so the instance forms no source terms */
#define CK_LAUNCH_NOSRC 8
#define CK_NBG 8 /* n_gas, kbt, b_field, divv, l_turb, dv, cs, beta_pl */

enum { CK_BG_NGAS = 0, CK_BG_KBT, CK_BG_B, CK_BG_DIVV, CK_BG_LTURB, CK_BG_DV, CK_BG_CS, CK_BG_BETA };

/* Physics parameters (reference key names, params.c:67-88). */
typedef struct {
    int dpp_model;       /* 1 asa, 2 ttd, 3 off, 4 direct */
    int ttd_tacc_model;  /* 0 spectrum, 1 brunetti16 */
    int reacc_window_mode;
    int tracer_nsub_max;
    int coeff_interp_min_steps;
    int coeff_interp_max_segments;
    double psi;
    double mach_limit;
    double eta_dpp_cap;
    double t_acc_direct_gyr;
    double t_off_gyr;
    double t_acc_off_gyr;
    double f_eddy;
    double tracer_nsub_safety;
    double epmax;
} CkParams;

/* Device-resident grid tables for one grid size n (both species). */
typedef struct {
    int n;
    int j_pp;
    const double *pp, *pdp, *pE, *p_p2, *p_expcut, *p_sig; /* p_sig: sigma_pp(j or j_pp) * sigmoid */
    const double *ep, *edp, *eE, *e_p2, *e_expcut;
    const double *p_ecr_w, *e_ecr_w; /* eta-cap energy weights (E-m) GeV dp on bins 2..n-3, else 0 */
    /* derived tables of the Chang-Cooper build: face width h_j = x_{j+1} - x_j
     * (0 for j = n-1), 1/h_j, 1/dx_j and 2/p_j */
    const double *p_h, *p_inv_h, *p_inv_dx, *p_2op;
    const double *e_h, *e_inv_h, *e_inv_dx, *e_2op;
    /* per-bin factors of the loss rates and D_pp, built on the device:
     * proton E_kin = gamma m_p - m_p, 1 / (2 m_p (gamma - 1)), 1 / beta^2,
     * exp(-E / E_pmax) (rebuilt when the parameters change); electron 1 / beta^2
     * and ln p */
    const double *p_Ep, *p_i2mk, *p_ib2, *p_expE, *e_ib2, *e_lnp;
    const double *zeros; /* n zeros: the rows of an absent source */
    const double *pow10; /* 10^k for k = -40..40 (index k + 40), as the host's pow gives them */
    double ppm1, ppp1, epm1, epp1;
} CkGrid;

/* One request (or the whole batch) in a launch.  Arrays are device pointers. */
typedef struct {
    double *crp;                  /* [ntr * n] cell-major, updated in place */
    double *cre;                  /* [ntr * n] */
    const double *qpi;            /* [ntr * n] or NULL (all zero) */
    const double *qepri;          /* [ntr * n] or NULL */
    const double *mass;           /* [ntr] or NULL */
    const unsigned char *dis;     /* [ntr] disable_adiabatic or NULL */
    const double *bgc[CK_NBG];    /* current snapshot backgrounds [ntr] */
    const double *bgn[CK_NBG];    /* next snapshot backgrounds [ntr] */
    const double *ric_c;          /* IC cooling row at z_curr [n] */
    const double *ric_n;          /* IC cooling row at z_next [n] */
    const int *nsub;              /* [ntr] */
    const int *n_on;              /* [ntr] */
    const double *fqe_t;          /* [n * n] secondary weights (transposed) or NULL */
    int jp0;                      /* first contributing proton bin */
    int ntr;
    int allow_interp;
    int protons;                  /* proton species active (else crp == 0 exactly) */
    double dt_snap;
    double elapsed;
    double *out_e;                /* optional post-snapshot electron state [ntr * n] */
    double *out_p;                /* optional post-snapshot proton state [ntr * n] */
    /* split launches (modes 2 and 3): segment rows, tracer t's segment s at
     * row seg_off[t] + s; seg_rec [rows][n] segment-start protons (written
     * by mode 2), seg_inje [rows][n] = sum_p seg_rec[row][p] fqe_t[p][e] */
    double *seg_rec;
    const double *seg_inje;
    /* IC cooling rate at bins 0 and n-1 of z_curr (the nsub estimate's endpoints) */
    double ric_lo, ric_hi;
    const int *seg_off;
    CkParams prm;
} CkSlot;

/* One unit of the persistent queue: tracer `tracer` of the request whose
 * first snapshot is slots[slot], advanced over nslot consecutive slots
 * (snapshots) in one go.  State lives in slots[slot].crp / .cre. */
typedef struct {
    int slot;
    int tracer;
    int nslot;
} CkTask;

/* Runtime helpers (thin wrappers over the HIP runtime). */
int ck_device_count(void);
int ck_set_device(int dev);
int ck_get_device(void);
int ck_device_name(int dev, char *buf, int len);
int ck_device_cu_count(int dev);
int ck_mem_info(size_t *free_b, size_t *total_b);
void *ck_malloc(size_t bytes);
void *ck_host_alloc(size_t bytes);
int ck_free(void *p);
int ck_host_free(void *p);
void *ck_stream_create(void);
int ck_stream_destroy(void *s);
int ck_stream_sync(void *s);
void *ck_event_create(void);
int ck_event_destroy(void *e);
int ck_event_record(void *e, void *s);
int ck_event_query(void *e);          /* 1 done, 0 pending, <0 error */
int ck_event_sync(void *e);
float ck_event_elapsed_ms(void *a, void *b);
int ck_stream_wait_event(void *s, void *e);
int ck_h2d(void *dst, const void *src, size_t bytes, void *stream);
int ck_d2h(void *dst, const void *src, size_t bytes, void *stream);
int ck_d2d(void *dst, const void *src, size_t bytes, void *stream);
int ck_device_sync(void);
const char *ck_last_error(void);

/* Graph capture of a stream's work (one snapshot's launch sequence). */
int ck_graph_begin(void *stream);
void *ck_graph_end(void *stream);          /* returns executable graph or NULL */
int ck_graph_launch(void *graph_exec, void *stream);
int ck_graph_destroy(void *graph_exec);

/* Secondary-electron contraction of the split two-species launches
 * (croma_secondary.hip): inje[r][je] = sum_{jp >= np_min[je]} fqe_t[jp][je]
 * rec[r][jp] for r < rows, rows = *d_rows (device count, capped at rows_cap)
 * or rows_cap when d_rows is NULL; FP64 MFMA, at most nblocks blocks looping
 * over row tiles.  n in {64, 128, 256, 512}. */
int ck_sec_contract(int n, const double *fqe_t, const int *np_min, const double *rec, double *inje, long rows_cap,
                    const int *d_rows, int nblocks, void *stream);

/* Inverse-Compton cooling rows on the device: for each of nz redshifts,
 * zb[3 iz .. 3 iz + 2] = {z, e0min, e0max} (CMB bounds from the host),
 * rows[iz * n + j] for the n electron bins and ghosts[2 iz], ghosts[2 iz + 1]
 * for the ghost momenta epm1, epp1 (ghosts may be NULL).  err |= 4 when the
 * CMB bounds span more nodes than the kernel holds. */
int ck_ic_rows(int n, const double *ep_dev, double epm1, double epp1, int nz, const double *zb_dev,
               double *rows_dev, double *ghosts_dev, int *err_dev, void *stream);

/* Substep counts, cost order and split segment rows of one window on the
 * device: for every tracer and slot q < count, slots[q].nsub / n_on (and
 * seg_off when split) are written, tasks[] receives the tracers in
 * descending cost buckets, *cells accumulates sum nsub.  Work buffers:
 * key [ntr], segloc [ntr], blk [(ntr + 255) / 256 + 1], hist [CK_NSUB_HIST].
 * The slot array is read on the device (d_slot) and on the host (h_slot,
 * for the scalars of the launch). */
#define CK_NSUB_HIST 8192
int ck_nsub_order(CkGrid g, const CkSlot *d_slot, int count, int ntr, int protons, int split, CkTask *tasks,
                  int *key, int *segloc, int *blk, int *hist, unsigned long long *cells, void *stream);
/* The nsub estimate alone for device arrays of ntr tracers. */
int ck_nsub_arrays(CkGrid g, CkParams prm, int ntr, double dt_snap, double ric_lo, double ric_hi,
                   const double *n_gas, const double *kbt, const double *b_field, const double *divv,
                   const double *l_turb, const double *dv, int *nsub, int *n_on, void *stream);

/* Fills the per-bin loss factors of g (p_Ep, p_i2mk, p_ib2, p_expE, e_ib2;
 * writable device arrays of g.n) for the given E_pmax [GeV]. */
int ck_grid_loss_tables(CkGrid g, double epmax, void *stream);

/* Fill kernel (used instead of memset so it is safe under graph capture). */
int ck_fill_i32(int *dst, int value, size_t count, void *stream);
int ck_fill_f64(double *dst, double value, size_t count, void *stream);
int ck_fill_u8(unsigned char *dst, unsigned char value, size_t count, void *stream);

/* Fused path: persistent cost-ordered lane-group kernel.  rows_per_lane is
 * 2, 4, 8 or 16 (n / rows_per_lane lanes per tracer, 4..64); occupancy 1 or 2
 * waves per SIMD selects the register budget (launch cu_count * occupancy
 * blocks of 4 waves for one resident wave per SIMD and occupancy step).
 * mode: 0 electron_only, 1 both species (secondaries contracted in the
 * kernel), 2 protons only recording segment-start rows, 3 electrons only
 * with secondaries from seg_inje; every slot of one launch must agree.  The
 * counter must be zero on entry. */
/* 1 if the FP32-coefficient instance of this launch shape and mode (0..3) exists */
int ck_fused_f32_supported(int n, int r, int occ, int mode);
/* LDS layout check of a fused instance (tb: 8 FP64 / 4 FP32 tables); 0 if in bounds */
int ck_fused_lds_check(int n, int r, int occ, int mode, int tb, long *info);
int ck_launch_fused(int n, int rows_per_lane, int occupancy, int nblocks, int mode,
                    CkGrid grid, const CkSlot *slots_dev, const CkTask *tasks_dev, int ntasks,
                    int *counter_dev, int *err_dev, void *stream);
int ck_fused_supported(int n, int rows_per_lane);

/* Plain path (one tracer per thread, unfused coefficient kernels, every
 * per-bin array in HBM, momentum-major: bin j of the bucket's tracer i at
 * j * count + i).  idx[count] selects tracers of the slot (a bucket). */
typedef struct {
    double *radp_p, *tl_p, *radp_e, *dpp_on_p, *dpp_on_e, *dpp_off_p, *dpp_off_e; /* [n][count] */
    double *a, *b, *c;         /* [4][n][count]: p-on, e-on, p-off, e-off */
    double *inje;              /* [n][count] */
    double *cprime;            /* [n][count] */
    double *bg;                /* [CK_NBG][count] interpolated backgrounds */
    double *ric;               /* IC cooling rate of bin j, tracer i at j * ric_js + i * ric_is:
                                * one row [n] (ric_js 1, ric_is 0) or per tracer [n][count] */
    long ric_js, ric_is;
    double *xp, *xe;           /* [count*n] gathered state */
    double *qp, *qe;           /* [count*n] gathered sources */
    double *mass;              /* [count] */
    unsigned char *dis;        /* [count] */
} CkPlainWs;

/* out[i] = cur[i] + alpha (nxt[i] - cur[i]) (the reference's lerp), i < count */
int ck_lerp_f64(size_t count, const double *cur, const double *nxt, double alpha, double *out, void *stream);
int ck_plain_gather(int n, int count, const int *idx_dev, const CkSlot *slot_host, CkPlainWs ws,
                    void *stream);
int ck_plain_lerp(int n, int count, const int *idx_dev, const CkSlot *slot_host, double alpha,
                  CkPlainWs ws, void *stream);
int ck_plain_coeff(int n, int count, CkGrid grid, CkParams prm, double dt, double t_acc_direct,
                   int protons, CkPlainWs ws, int *err_dev, void *stream);
int ck_plain_secondary(int n, int count, const double *fqe_t, int jp0, CkPlainWs ws, void *stream);
int ck_plain_thomas(int n, int count, int nstep, double dt, int windowed, int on_start, int on_end,
                    int protons, CkPlainWs ws, int *err_dev, void *stream);
int ck_plain_scatter(int n, int count, const int *idx_dev, const CkSlot *slot_host, CkPlainWs ws,
                     void *stream);

/* Batched tridiagonal Chang-Cooper solve (solve_cc_cuda_batch semantics):
 * state_out = Thomas(a,b,c; dt*Q + state_in), cell-major, one system per thread. */
int ck_solve_cc_batch(int nsys, int nrow, double dt, const double *a, const double *b,
                      const double *c, const double *q, const double *state_in, double *state_out,
                      double *cprime_ws, int *err_dev, void *stream);

/* Emission contractions: out[i][k] = row_scale[i] * sum_j spectra[i][j] kmat[j][k].
 * spectra [ntr x K] FP64; the kernel matrix in bin-major layout [nbins_pad][K]
 * (nbins_pad a multiple of 64, zero padded) as FP64 (kt64), FP32 scaled by
 * 2^-col_shift[k] (kt32) and that FP32 value split into BF16 halves (kt_hi,
 * kt_lo); row_scale [ntr] or NULL; out [ntr x nbins] FP64.  mode 0 FP32
 * VALU, 1 split-BF16 x3 on MFMA, 2 one BF16 pass on MFMA, 3 FP64 VALU. */
int ck_emission_contract(int ntr, int K, int nbins, int nbins_pad, const double *spectra, const uint16_t *kt_hi,
                         const uint16_t *kt_lo, const float *kt32, const double *kt64, const int *col_shift,
                         const double *row_scale, double *out, int mode, void *stream);
/* Synchrotron pitch-kernel table [nlogb][nfreq][n] from logy [nfreq][n][ntheta]. */
int ck_synch_table(int n, int nfreq, int ntheta, int nlogb, double logb_min, double inv_dlogb, int nx_tab,
                   double xmin, const double *logx_tab, const double *logfx_tab, const double *fx_tab,
                   const double *logy, const double *pitch_weight, double *table, void *stream);
/* Synchrotron on MFMA.  Prepare (once per table): Chebyshev coefficients of
 * the log-B interpolation per plane interval (kt_hi / kt_lo: (nlogb - 1) x
 * 2 ceil(nfreq / 16) column blocks of 64 columns x n bins, bf16, in the
 * contraction's LDS order; col_shift [(nlogb - 1) * nfreq]) and the float
 * log2 table lt [nlogb * nfreq * n] of the FP32 fix-up.  Run (per call):
 * tracers grouped by plane (work: 2 ntr + 2 nlogb + 8 ints), the fused
 * split-BF16 x3 contraction with its range check, then the tracers whose
 * emissivities fall below its range (fix: count + list, ntr + 1 ints) again
 * by the FP32 log-domain kernel.  n in {64, 128, 256}. */
int ck_synch_mfma_prepare(int n, int nfreq, int nlogb, const double *table, const double *dp, uint16_t *kt_hi,
                          uint16_t *kt_lo, int *col_shift, float *lt, void *stream);
int ck_synch_mfma_run(int ntr, int n, int nfreq, int nlogb, double logb_min, double inv_dlogb, const double *table,
                      const uint16_t *kt_hi, const uint16_t *kt_lo, const int *col_shift, const float *lt, int *work,
                      int *fix, const double *cre, const double *dp, const double *b_field, double *eps, void *stream);
/* Grid-constant tables from the shared physics header (croma_tables.hip):
 * secondary-electron weights fqe_t[jp][je] and np_min[je] from the electron
 * energies, proton momenta, beta_p and proton bin widths; gamma (kind 0),
 * neutrino (1) or IC emissivity (2) kernel rows out[nbins][n] over the log10
 * range lo..hi (mom: proton momenta for kinds 0/1, electron momenta for 2). */
int ck_secondary_weights(int n, const double *e_energy, const double *pp, const double *beta_p,
                         const double *pdp, double *fqe_t, int *np_min, void *stream);
int ck_emission_kernel_table(int kind, int n, int nbins, double lo, double hi, double z,
                             const double *mom, double *out, void *stream);
/* Synchrotron: per-tracer log-B interpolated table contraction (VALU, FP64 or FP32).
 * logt: workspace of nlogb * nfreq * n doubles, filled here with the table's
 * logarithms (FP64: natural log, NULL computes them per element; FP32: float
 * log2 in its first half, required).
 * work (NULL: tracers in input order): 2 * ntr + nlogb ints, used to run the
 * tracers grouped by table plane. */
int ck_emission_synch(int ntr, int n, int nfreq, int nlogb, double logb_min, double inv_dlogb,
                      const double *table, double *logt, int *work, const double *cre, const double *dp,
                      const double *b_field, double *eps, int fp64, void *stream);

#ifdef __cplusplus
}
#endif

#endif
