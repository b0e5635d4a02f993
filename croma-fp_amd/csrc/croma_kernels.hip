// croma_kernels.hip -- CDNA4 (gfx950) kernels for CROMA's per-tracer
// Fokker-Planck update, plus the thin runtime wrappers the C host code uses.
//
// Fused path (k_fused): one lane group of G = n/R lanes owns one tracer for a
// whole snapshot.  Each lane keeps R consecutive momentum rows of the
// spectrum in registers.  Per coefficient segment the group rebuilds the
// Chang-Cooper matrix in FP64 (losses, D_pp, face weights fused into the
// assembly), factors it once (Thomas LU; the pivots come from a 2x2
// projective scan across lanes followed by the exact in-lane recurrence),
// and then runs every substep as two affine scans (forward elimination and
// back substitution) whose coefficients are all non-negative.  Tracers are
// pulled from a cost-ordered task list through one atomic counter.
//
// Plain path (k_plain_*): the reference GPU semantics -- one tracer per
// thread, cell-major a/b/c arrays in HBM, unfused coefficient kernels,
// per-bucket and per-segment launches (fp_cuda_solver.cu:104-141,
// fp_cuda_coeff.cu:379-803) -- written fresh as the baseline.
//
// Physics follows the reference CPU path operation for operation:
// FP_Coef.c:438-849, fp_shared_core.c:729-1081, Chang_Cooper.c:22-184,
// fp_shared_core.c:226-257, 564-615, 1591-1672, tracer_fp_cpu.c:88-196,
// tracer_fp_batch.c:651-827.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "croma_device.h"


static thread_local char g_errbuf[256];

static int hip_ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    snprintf(g_errbuf, sizeof(g_errbuf), "%s: %s", what, hipGetErrorString(e));
    return -1;
}

extern "C" int ck_launch_fused_n64(int, int, int, int, CkGrid, const CkSlot *, const CkTask *, int, int *, int *, void *);
extern "C" int ck_launch_fused_n128(int, int, int, int, CkGrid, const CkSlot *, const CkTask *, int, int *, int *, void *);
extern "C" int ck_launch_fused_n256(int, int, int, int, CkGrid, const CkSlot *, const CkTask *, int, int *, int *, void *);
extern "C" int ck_launch_fused_n512(int, int, int, int, CkGrid, const CkSlot *, const CkTask *, int, int *, int *, void *);

extern "C" int ck_fused_supported(int n, int r)
{
    if (!(n == 64 || n == 128 || n == 256 || n == 512)) return 0;
    if (!(r == 2 || r == 4 || r == 8 || r == 16)) return 0;
    if (n / r > 64 || n / r < 4) return 0;
    return 1;
}

// Host-side check of one fused instance's LDS use.  Per lane group the
// slice of fused_lds_doubles(N, R, MODE, PK) doubles holds, in this order,
// the segment's loss rates (row j at j + j / R; a lane reads its own rows
// at gl (R + 1) + k), the parked spectrum (PK; rows at gl (R + 1) + k), the
// EP rows (MODE 1; row j at j) and the 2 CK_NBG backgrounds.  info[0..5]:
// largest index of each region's accesses (rates, own rates, parked, EP,
// backgrounds) and the slice size; info[6] the bytes a CU holds for the
// blocks its register budget admits.  Returns 0 when every index lies in the
// slice and the blocks fit the 160 KiB of LDS.
extern "C" int ck_fused_lds_check(int n, int r, int occ, int mode, int tb, long *info)
{
    if (!ck_fused_supported(n, r) || occ < 1 || occ > 2 || mode < 0 || mode > 3 || !(tb == 4 || tb == 8)) return -1;
    const int WPB = 4, G = n / r;
    const bool LG = fused_use_lg(n, r, WPB, mode, occ, tb);
    const bool PK = fused_use_pk(n, r, WPB, mode, occ, tb);
    const long D = fused_lds_doubles(n, r, mode, PK);
    const long nrates = n + n / r + 1;
    const long rates_max = (long)(n - 1) + (n - 1) / r;
    const long own_rates_max = (long)(G - 1) * (r + 1) + r - 1;
    const long pk_base = nrates;
    const long parked_max = PK ? pk_base + (long)(G - 1) * (r + 1) + r - 1 : -1;
    const long row_base = pk_base + (PK ? (long)G * (r + 1) : 0);
    const long ep_max = (mode == CK_MODE_EP) ? row_base + n - 1 : -1;
    const long bg_max = row_base + (mode == CK_MODE_EP ? n : 0) + 2 * CK_NBG - 1;
    const long blocks = (occ * 4 / WPB > 0) ? occ * 4 / WPB : 1;
    const long bytes = (8L * WPB * (64 / G) * D + (LG ? (long)tb * fused_grid_tables(mode) * nrates : 0)) * blocks;
    if (info) {
        info[0] = rates_max;
        info[1] = own_rates_max;
        info[2] = parked_max;
        info[3] = ep_max;
        info[4] = bg_max;
        info[5] = D;
        info[6] = bytes;
    }
    const bool ok = rates_max < nrates && own_rates_max < nrates && parked_max < row_base && ep_max < D &&
                    bg_max < D && bytes <= 160L * 1024 && (LG || tb == 8);
    return ok ? 0 : 1;
}

// Whether the FP32-coefficient instance of (n, r, occ, mode) exists: its
// float grid tables must fit the LDS beside the lane groups' regions.
extern "C" int ck_fused_f32_supported(int n, int r, int occ, int mode)
{
    if (mode == CK_LAUNCH_NOSRC) mode = 0;
    if (!ck_fused_supported(n, r) || occ < 1 || occ > 2 || mode < 0 || mode > 3) return 0;
    return fused_use_lg(n, r, 4, mode, occ, 4) ? 1 : 0;
}

extern "C" int ck_launch_fused(int n, int r, int occ, int nblocks, int mode, CkGrid grid, const CkSlot *slots,
                               const CkTask *tasks, int ntasks, int *counter, int *err, void *stream)
{
    int rc = -1;
    if (!ck_fused_supported(n, r) || ntasks < 0 || nblocks <= 0) return -1;
    if (occ < 1 || occ > 2 || mode < 0 || (mode > 7 && (mode & ~CK_LAUNCH_F32) != CK_LAUNCH_NOSRC)) return -1;
    if (ntasks == 0) return 0;
    switch (n) {
    case 64: rc = ck_launch_fused_n64(r, occ, nblocks, mode, grid, slots, tasks, ntasks, counter, err, stream); break;
    case 128: rc = ck_launch_fused_n128(r, occ, nblocks, mode, grid, slots, tasks, ntasks, counter, err, stream); break;
    case 256: rc = ck_launch_fused_n256(r, occ, nblocks, mode, grid, slots, tasks, ntasks, counter, err, stream); break;
    case 512: rc = ck_launch_fused_n512(r, occ, nblocks, mode, grid, slots, tasks, ntasks, counter, err, stream); break;
    }
    if (rc) hip_ok(hipGetLastError(), "k_fused launch");
    return rc;
}

// ---------------------------------------------------------------------------
// Plain path kernels: one tracer per thread, unfused coefficient kernels,
// every per-bin array materialised in HBM.  The workspace is momentum-major
// ([bin][tracer of the bucket], stride `count`), so the threads of a wave
// (consecutive tracers) touch consecutive addresses at every bin; gather and
// scatter transpose from and to the batch's cell-major spectra.
// ---------------------------------------------------------------------------
__global__ void k_plain_gather(int n, int count, const int *__restrict__ idx, const double *__restrict__ crp,
                               const double *__restrict__ cre, const double *__restrict__ qpi,
                               const double *__restrict__ qepri, const double *__restrict__ mass,
                               const unsigned char *__restrict__ dis, CkPlainWs ws)
{
    const size_t tot = (size_t)count * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(i / count), c = (int)(i % count);
        const size_t src = (size_t)idx[c] * n + j;
        ws.xp[i] = crp ? crp[src] : 0.0;
        ws.xe[i] = cre[src];
        ws.qp[i] = qpi ? qpi[src] : 0.0;
        ws.qe[i] = qepri ? qepri[src] : 0.0;
        if (j == 0) {
            ws.mass[c] = mass ? mass[idx[c]] : 0.0;
            ws.dis[c] = dis ? dis[idx[c]] : 0;
        }
    }
}

__global__ void k_plain_scatter(int n, int count, const int *__restrict__ idx, double *__restrict__ crp,
                                double *__restrict__ cre, CkPlainWs ws)
{
    const size_t tot = (size_t)count * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(i / count), c = (int)(i % count);
        const size_t dst = (size_t)idx[c] * n + j;
        if (crp) crp[dst] = ws.xp[i];
        cre[dst] = ws.xe[i];
    }
}

struct BgPtrs {
    const double *c[CK_NBG];
    const double *x[CK_NBG];
};

__global__ void k_plain_lerp(int n, int count, const int *__restrict__ idx, BgPtrs bp, const double *ric_c,
                             const double *ric_n, double alpha, int interp, CkPlainWs ws)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const int t = idx[i];
#pragma unroll
        for (int f = 0; f < CK_NBG; ++f)
            ws.bg[(size_t)f * count + i] = interp ? d_lerp(bp.c[f][t], bp.x[f][t], alpha) : bp.c[f][t];
    }
    if (i < n && ric_c) ws.ric[i] = interp ? d_lerp(ric_c[i], ric_n[i], alpha) : ric_c[i];
}

__global__ void k_lerp_f64(size_t count, const double *__restrict__ cur, const double *__restrict__ nxt, double alpha,
                           double *__restrict__ out)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        out[i] = d_lerp(cur[i], nxt[i], alpha);
}

// losses + D_pp (on / off) + Chang-Cooper assembly, one tracer per thread;
// bin j of tracer i at j * count + i
__global__ void k_plain_coeff(int n, int count, CkGrid g, CkParams prm, double dt, double t_acc_direct,
                              int protons, CkPlainWs ws, int *err)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const size_t st = (size_t)count;
    const double n_gas = ws.bg[0 * st + i], kbt = ws.bg[1 * st + i], bfield = ws.bg[2 * st + i];
    const double divv = ws.dis[i] ? 0.0 : ws.bg[3 * st + i];
    const double lturb = ws.bg[4 * st + i], dv = ws.bg[5 * st + i], cs = ws.bg[6 * st + i];
    const double beta = ws.bg[7 * st + i];
    const double b_micro = bfield * 1.0e+6;
    const double b_ratio = b_micro / 3.2;
    const double synch_coeff = 4.8e-4 * b_ratio * b_ratio * 9.11e-28 * 3.0e+10 * 3.1536e+16;
    const double coul_coeff = 3.05e-29 * n_gas * K_C * K_GYR / (K_ME * K_GEV);
    const double adia = 1.0 / 3.0 * divv;
    const double tl_pref = 1.0 / (0.5 * K_C * n_gas) / K_GYR;
    for (int j = 0; j < n; ++j) {
        const size_t o = (size_t)j * st + i;
        const double pe = g.ep[j], pe2 = g.e_p2[j];
        const double beta2 = pe2 / (1.0 + pe2);
        ws.radp_e[o] = synch_coeff * pe2 + ws.ric[(size_t)j * ws.ric_js + (size_t)i * ws.ric_is] + coul_coeff * (1.0 + log(pe / n_gas) / 74.8) / beta2 +
                       adia * pe;
        if (protons) {
            ws.radp_p[o] = d_b_coulomb_p(n_gas, g.pp[j], kbt) + adia * g.pp[j];
            ws.tl_p[o] = tl_pref / g.p_sig[j];
        }
    }
    double inv_on = 0.0;
    const int mode_on = d_dpp_model(prm, prm.dpp_model, lturb, dv, cs, beta, t_acc_direct, &inv_on);
    double scale = 1.0;
    if (prm.dpp_model != 3 && prm.eta_dpp_cap > 0.0) {
        double ecr = 0.0;
        if (protons)
            for (int j = 2; j < n - 2; ++j) {
                const double w = g.p_ecr_w[j], v = ws.xp[(size_t)j * st + i];
                if (w > 0.0 && isfinite(v)) ecr += w * v * g.pdp[j];
            }
        for (int j = 2; j < n - 2; ++j) {
            const double w = g.e_ecr_w[j], v = ws.xe[(size_t)j * st + i];
            if (w > 0.0 && isfinite(v)) ecr += w * v * g.edp[j];
        }
        scale = d_eta_scale(prm, prm.dpp_model, lturb, dv, cs, beta, ws.mass[i], t_acc_direct, ecr);
    }
    for (int j = 0; j < n; ++j) {
        const size_t o = (size_t)j * st + i;
        const double off_e = g.e_p2[j] * (1.0 / (4.0 * 1.0e+5));
        const double on_e = mode_on ? g.e_p2[j] * inv_on * g.e_expcut[j] : off_e;
        ws.dpp_off_e[o] = off_e;
        ws.dpp_on_e[o] = on_e * scale;
        if (protons) {
            const double off_p = g.p_p2[j] * (1.0 / (4.0 * 1.0e+5));
            const double on_p = mode_on ? g.p_p2[j] * inv_on * exp(-g.pE[j] / prm.epmax) * g.p_expcut[j] : off_p;
            ws.dpp_off_p[o] = off_p;
            ws.dpp_on_p[o] = on_p * scale;
        }
    }
    // assembly: sets 0 p-on, 1 e-on, 2 p-off, 3 e-off
    const size_t plane = (size_t)count * n;
    for (int set = 0; set < 4; ++set) {
        const int sp = set & 1; // 0 proton, 1 electron
        if (sp == 0 && !protons) continue;
        const double *xg = sp ? g.ep : g.pp;
        const double *dxg = sp ? g.edp : g.pdp;
        const double *radp = sp ? ws.radp_e + i : ws.radp_p + i;
        const double *D = (set < 2) ? (sp ? ws.dpp_on_e + i : ws.dpp_on_p + i)
                                    : (sp ? ws.dpp_off_e + i : ws.dpp_off_p + i);
        double *a = ws.a + set * plane + i, *b = ws.b + set * plane + i, *c = ws.c + set * plane + i;
        const double *tl = ws.tl_p + i;
        double Wp_prev = 0.0, Wm_prev = 0.0, w_prev = 0.0;
        for (int m = 0; m < n; ++m) {
            const size_t om = (size_t)m * st;
            const double Dm = D[om];
            const double Bm = radp[om] - 2.0 / xg[m] * Dm;
            double Bl = 0.0, Br = 0.0, xl = 0.0, xr = 0.0, Cl = 0.0, Cr = 0.0;
            if (m > 0) { Bl = radp[om - st] - 2.0 / xg[m - 1] * D[om - st]; Cl = D[om - st]; xl = xg[m - 1]; }
            if (m < n - 1) { Br = radp[om + st] - 2.0 / xg[m + 1] * D[om + st]; Cr = D[om + st]; xr = xg[m + 1]; }
            double Wp = 0.0, Wm = 0.0, w = 0.0;
            if (m < n - 1) d_face(Bm, Br, Dm, Cr, xg[m], xr, Wp, Wm, w);
            const int kind = (m == 0) ? 1 : ((m == n - 1) ? 2 : 0);
            const double dtT = sp ? dt / 1.0e+100 : dt / tl[om];
            double delta_l = 0.0;
            if (kind == 2 && !(fabs(Dm) > 1.0e-200) && fabs(Cl) > 1.0e-200) delta_l = d_face_delta(w_prev);
            d_cc_row(kind, dt, xg[m], dxg[m], Bm, Dm, dtT, xl, Bl, Cl, Wp_prev, Wm_prev, delta_l, xr, Br, Cr, Wp,
                     Wm, a[om], b[om], c[om]);
            Wp_prev = Wp;
            Wm_prev = Wm;
            w_prev = w;
        }
    }
    (void)err;
}

__global__ void k_plain_secondary(int n, int count, const double *__restrict__ fqe_t, int jp0, CkPlainWs ws)
{
    const size_t tot = (size_t)count * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const int je = (int)(i / count), c = (int)(i % count);
        double q = 0.0;
        if (fqe_t)
            for (int jp = jp0; jp < n; ++jp) q += ws.xp[(size_t)jp * count + c] * fqe_t[(size_t)jp * n + je];
        ws.inje[i] = q * K_GYR * ws.bg[(size_t)c] + ws.qe[i];
    }
}

// Thomas with the rhs update (solve_cc_batch_cpu_timed semantics) for one
// system of n rows at stride st, state in place; cprime scratch in HBM.
__device__ __forceinline__ int d_thomas_cc(int n, double dt, const double *a, const double *b, const double *c,
                                           const double *q, double *x, double *cp, size_t st = 1)
{
    for (int j = 0; j < n; ++j) x[j * st] = dt * q[j * st] + x[j * st];
    if (fabs(b[0]) <= 1.0e-300) return ERR_PIVOT;
    cp[0] = -c[0] / b[0];
    x[0] /= b[0];
    for (int j = 1; j < n; ++j) {
        const size_t o = j * st;
        const double den = b[o] + a[o] * cp[o - st];
        if (fabs(den) <= 1.0e-300) return ERR_PIVOT;
        cp[o] = (j == n - 1) ? 0.0 : -c[o] / den;
        x[o] = (x[o] + a[o] * x[o - st]) / den;
    }
    for (int j = n - 2; j >= 0; --j) x[j * st] -= cp[j * st] * x[(j + 1) * st];
    return 0;
}

__global__ void k_plain_thomas(int n, int count, int nstep, double dt, int windowed, int on_start, int on_end,
                               int protons, CkPlainWs ws, int *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const size_t plane = (size_t)count * n, st = (size_t)count;
    int bad = 0;
    for (int s = 0; s < nstep && !bad; ++s) {
        const int on = !windowed || (s >= on_start && s < on_end);
        const size_t ps = (on ? 0 : 2) * plane + i, es = (on ? 1 : 3) * plane + i;
        if (protons)
            bad |= d_thomas_cc(n, dt, ws.a + ps, ws.b + ps, ws.c + ps, ws.qp + i, ws.xp + i, ws.cprime + i, st);
        bad |= d_thomas_cc(n, dt, ws.a + es, ws.b + es, ws.c + es, ws.inje + i, ws.xe + i, ws.cprime + i, st);
    }
    if (bad) atomicOr(err, bad);
}

__global__ void k_solve_cc_batch(int nsys, int nrow, double dt, const double *a, const double *b, const double *c,
                                 const double *q, const double *xin, double *xout, double *cp, int *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsys) return;
    const size_t off = (size_t)i * nrow;
    for (int j = 0; j < nrow; ++j) xout[off + j] = xin[off + j];
    const int bad = d_thomas_cc(nrow, dt, a + off, b + off, c + off, q + off, xout + off, cp + off);
    if (bad) atomicOr(err, bad);
}

static int grid1d(size_t work, int block)
{
    size_t nb = (work + block - 1) / block;
    if (nb > 65535u * 32u) nb = 65535u * 32u;
    if (nb < 1) nb = 1;
    return (int)nb;
}

extern "C" int ck_plain_gather(int n, int count, const int *idx, const CkSlot *S, CkPlainWs ws, void *stream)
{
    hipLaunchKernelGGL(k_plain_gather, dim3(grid1d((size_t)count * n, 256)), dim3(256), 0, (hipStream_t)stream, n,
                       count, idx, S->crp, S->cre, S->qpi, S->qepri, S->mass, S->dis, ws);
    return hip_ok(hipGetLastError(), "k_plain_gather");
}

extern "C" int ck_plain_scatter(int n, int count, const int *idx, const CkSlot *S, CkPlainWs ws, void *stream)
{
    hipLaunchKernelGGL(k_plain_scatter, dim3(grid1d((size_t)count * n, 256)), dim3(256), 0, (hipStream_t)stream,
                       n, count, idx, S->crp, S->cre, ws);
    return hip_ok(hipGetLastError(), "k_plain_scatter");
}

extern "C" int ck_lerp_f64(size_t count, const double *cur, const double *nxt, double alpha, double *out,
                           void *stream)
{
    if (count == 0) return 0;
    hipLaunchKernelGGL(k_lerp_f64, dim3(grid1d(count, 256)), dim3(256), 0, (hipStream_t)stream, count, cur, nxt,
                       alpha, out);
    return hip_ok(hipGetLastError(), "k_lerp_f64");
}

extern "C" int ck_plain_lerp(int n, int count, const int *idx, const CkSlot *S, double alpha, CkPlainWs ws,
                             void *stream)
{
    BgPtrs bp;
    for (int f = 0; f < CK_NBG; ++f) {
        bp.c[f] = S->bgc[f];
        bp.x[f] = S->bgn[f];
    }
    const int work = count > n ? count : n;
    hipLaunchKernelGGL(k_plain_lerp, dim3(grid1d((size_t)work, 128)), dim3(128), 0, (hipStream_t)stream, n, count,
                       idx, bp, S->ric_c, S->ric_n, alpha, alpha > 0.0 ? 1 : 0, ws);
    return hip_ok(hipGetLastError(), "k_plain_lerp");
}

extern "C" int ck_plain_coeff(int n, int count, CkGrid grid, CkParams prm, double dt, double t_acc_direct,
                              int protons, CkPlainWs ws, int *err, void *stream)
{
    hipLaunchKernelGGL(k_plain_coeff, dim3(grid1d((size_t)count, 128)), dim3(128), 0, (hipStream_t)stream, n,
                       count, grid, prm, dt, t_acc_direct, protons, ws, err);
    return hip_ok(hipGetLastError(), "k_plain_coeff");
}

extern "C" int ck_plain_secondary(int n, int count, const double *fqe_t, int jp0, CkPlainWs ws, void *stream)
{
    hipLaunchKernelGGL(k_plain_secondary, dim3(grid1d((size_t)count * n, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, count, fqe_t, jp0, ws);
    return hip_ok(hipGetLastError(), "k_plain_secondary");
}

extern "C" int ck_plain_thomas(int n, int count, int nstep, double dt, int windowed, int on_start, int on_end,
                               int protons, CkPlainWs ws, int *err, void *stream)
{
    hipLaunchKernelGGL(k_plain_thomas, dim3(grid1d((size_t)count, 64)), dim3(64), 0, (hipStream_t)stream, n, count,
                       nstep, dt, windowed, on_start, on_end, protons, ws, err);
    return hip_ok(hipGetLastError(), "k_plain_thomas");
}

extern "C" int ck_solve_cc_batch(int nsys, int nrow, double dt, const double *a, const double *b, const double *c,
                                 const double *q, const double *xin, double *xout, double *cp, int *err,
                                 void *stream)
{
    hipLaunchKernelGGL(k_solve_cc_batch, dim3(grid1d((size_t)nsys, 64)), dim3(64), 0, (hipStream_t)stream, nsys,
                       nrow, dt, a, b, c, q, xin, xout, cp, err);
    return hip_ok(hipGetLastError(), "k_solve_cc_batch");
}

// ---------------------------------------------------------------------------
// Fill kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T *dst, T v, size_t count)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

// Per-bin factors of the segment build (loss rates and the D_pp cutoff).
__global__ void k_grid_loss_tables(CkGrid g, double inv_epmax)
{
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.n) return;
    const double p = g.pp[j];
    const double p2 = p * p;
    const double gamma = sqrt(1.0 + p2);
    const double betap2 = p2 / (1.0 + p2);
    const double kinetic = gamma - 1.0;
    const_cast<double *>(g.p_Ep)[j] = gamma * K_MP - K_MP;
    const_cast<double *>(g.p_i2mk)[j] = 1.0 / (2.0 * K_MP * kinetic);
    const_cast<double *>(g.p_ib2)[j] = 1.0 / betap2;
    const_cast<double *>(g.p_expE)[j] = exp(-g.pE[j] * inv_epmax);
    const double pe2 = g.e_p2[j];
    const_cast<double *>(g.e_ib2)[j] = 1.0 / (pe2 / (1.0 + pe2));
    const_cast<double *>(g.e_lnp)[j] = log(g.ep[j]);
    const_cast<double *>(g.zeros)[j] = 0.0;
}

extern "C" int ck_grid_loss_tables(CkGrid g, double epmax, void *stream)
{
    if (g.n <= 0 || !(epmax > 0.0)) return -1;
    hipLaunchKernelGGL(k_grid_loss_tables, dim3((g.n + 255) / 256), dim3(256), 0, (hipStream_t)stream, g,
                       1.0 / epmax);
    return hip_ok(hipGetLastError(), "k_grid_loss_tables");
}

extern "C" int ck_fill_i32(int *dst, int v, size_t count, void *stream)
{
    if (count == 0) return 0;
    hipLaunchKernelGGL(k_fill<int>, dim3(grid1d(count, 256)), dim3(256), 0, (hipStream_t)stream, dst, v, count);
    return hip_ok(hipGetLastError(), "k_fill_i32");
}

extern "C" int ck_fill_u8(unsigned char *dst, unsigned char v, size_t count, void *stream)
{
    if (count == 0) return 0;
    hipLaunchKernelGGL(k_fill<unsigned char>, dim3(grid1d(count, 256)), dim3(256), 0, (hipStream_t)stream, dst, v,
                       count);
    return hip_ok(hipGetLastError(), "k_fill_u8");
}

extern "C" int ck_fill_f64(double *dst, double v, size_t count, void *stream)
{
    if (count == 0) return 0;
    hipLaunchKernelGGL(k_fill<double>, dim3(grid1d(count, 256)), dim3(256), 0, (hipStream_t)stream, dst, v, count);
    return hip_ok(hipGetLastError(), "k_fill_f64");
}

// ---------------------------------------------------------------------------
// Runtime wrappers
// ---------------------------------------------------------------------------
extern "C" const char *ck_last_error(void) { return g_errbuf; }

extern "C" int ck_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int ck_set_device(int d) { return hip_ok(hipSetDevice(d), "hipSetDevice"); }

extern "C" int ck_get_device(void)
{
    int d = -1;
    if (hipGetDevice(&d) != hipSuccess) return -1;
    return d;
}

extern "C" int ck_device_name(int dev, char *buf, int len)
{
    hipDeviceProp_t p;
    if (hip_ok(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties")) return -1;
    snprintf(buf, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
    return 0;
}

extern "C" int ck_device_cu_count(int dev)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return 0;
    return p.multiProcessorCount;
}

extern "C" int ck_mem_info(size_t *f, size_t *t) { return hip_ok(hipMemGetInfo(f, t), "hipMemGetInfo"); }

extern "C" void *ck_malloc(size_t bytes)
{
    void *p = nullptr;
    if (bytes == 0) bytes = 8;
    if (hip_ok(hipMalloc(&p, bytes), "hipMalloc")) return nullptr;
    return p;
}

extern "C" void *ck_host_alloc(size_t bytes)
{
    void *p = nullptr;
    if (bytes == 0) bytes = 8;
    if (hip_ok(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc")) return nullptr;
    return p;
}

extern "C" int ck_free(void *p) { return p ? hip_ok(hipFree(p), "hipFree") : 0; }
extern "C" int ck_host_free(void *p) { return p ? hip_ok(hipHostFree(p), "hipHostFree") : 0; }

extern "C" void *ck_stream_create(void)
{
    hipStream_t s = nullptr;
    if (hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate")) return nullptr;
    return (void *)s;
}

extern "C" int ck_stream_destroy(void *s) { return s ? hip_ok(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy") : 0; }
extern "C" int ck_stream_sync(void *s) { return hip_ok(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize"); }

extern "C" void *ck_event_create(void)
{
    hipEvent_t e = nullptr;
    if (hip_ok(hipEventCreate(&e), "hipEventCreate")) return nullptr;
    return (void *)e;
}

extern "C" int ck_event_destroy(void *e) { return e ? hip_ok(hipEventDestroy((hipEvent_t)e), "hipEventDestroy") : 0; }
extern "C" int ck_event_record(void *e, void *s)
{
    return hip_ok(hipEventRecord((hipEvent_t)e, (hipStream_t)s), "hipEventRecord");
}

extern "C" int ck_event_query(void *e)
{
    const hipError_t r = hipEventQuery((hipEvent_t)e);
    if (r == hipSuccess) return 1;
    if (r == hipErrorNotReady) return 0;
    hip_ok(r, "hipEventQuery");
    return -1;
}

extern "C" int ck_event_sync(void *e) { return hip_ok(hipEventSynchronize((hipEvent_t)e), "hipEventSynchronize"); }

extern "C" float ck_event_elapsed_ms(void *a, void *b)
{
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, (hipEvent_t)a, (hipEvent_t)b) != hipSuccess) return -1.0f;
    return ms;
}

extern "C" int ck_stream_wait_event(void *s, void *e)
{
    return hip_ok(hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)e, 0), "hipStreamWaitEvent");
}

extern "C" int ck_h2d(void *dst, const void *src, size_t bytes, void *s)
{
    if (bytes == 0) return 0;
    return hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s), "hipMemcpyAsync H2D");
}

extern "C" int ck_d2h(void *dst, const void *src, size_t bytes, void *s)
{
    if (bytes == 0) return 0;
    return hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s), "hipMemcpyAsync D2H");
}

extern "C" int ck_d2d(void *dst, const void *src, size_t bytes, void *s)
{
    if (bytes == 0) return 0;
    return hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s), "hipMemcpyAsync D2D");
}

extern "C" int ck_device_sync(void) { return hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

extern "C" int ck_graph_begin(void *s)
{
    return hip_ok(hipStreamBeginCapture((hipStream_t)s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
}

extern "C" void *ck_graph_end(void *s)
{
    hipGraph_t gph = nullptr;
    hipGraphExec_t ex = nullptr;
    if (hip_ok(hipStreamEndCapture((hipStream_t)s, &gph), "hipStreamEndCapture")) return nullptr;
    if (hip_ok(hipGraphInstantiate(&ex, gph, nullptr, nullptr, 0), "hipGraphInstantiate")) {
        (void)hipGraphDestroy(gph);
        return nullptr;
    }
    (void)hipGraphDestroy(gph);
    return (void *)ex;
}

extern "C" int ck_graph_launch(void *ex, void *s)
{
    return hip_ok(hipGraphLaunch((hipGraphExec_t)ex, (hipStream_t)s), "hipGraphLaunch");
}

extern "C" int ck_graph_destroy(void *ex)
{
    return ex ? hip_ok(hipGraphExecDestroy((hipGraphExec_t)ex), "hipGraphExecDestroy") : 0;
}

