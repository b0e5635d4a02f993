// Grid-constant kernel tables built on the device from the shared physics
// header (croma_hadronic.h), one table entry per thread:
//   secondary-electron weights fqe_t[jp][je]   fp_shared_core.c:1453-1506 (QeKernel_BB05, HADRONIC.c:436-498)
//   gamma / neutrino / IC kernel tables [k][j]  fp_shared_core.c:1724-1804 (prepare_*_kernel_flat),
//                                              HADRONIC.c:139-289, FP_Coef.c:553-582
// Every entry is an independent FP64 quadrature of 50..400 nodes (pow /
// log10 / exp per node), so the tables are FP64-VALU bound and need no LDS;
// threads of a wave walk consecutive columns so each store is coalesced.
#include <hip/hip_runtime.h>

#include "croma_hadronic.h"
#include "croma_kernels.h"

namespace {

constexpr int kTabBlock = 128;

// vals[jp][je] = QeKernel(E_e[je], p[jp]) * beta_p[jp], NaN -> 0
// (fp_shared_core.c:1464-1471).
__global__ void __launch_bounds__(kTabBlock) k_fqe_vals(int n, const double *__restrict__ e_energy,
                                                         const double *__restrict__ pp,
                                                         const double *__restrict__ beta_p, double *vals)
{
    const int je = blockIdx.x * kTabBlock + threadIdx.x;
    const int jp = blockIdx.y;
    if (je >= n) return;
    double v = chk_qe_kernel(e_energy[je], pp[jp]) * beta_p[jp];
    if (isnan(v)) v = 0.0;
    vals[(size_t)jp * n + je] = v;
}

// Per electron bin je: jp0 = 1 + the last proton bin whose value is below
// 1e-80 (at least 1), then the trapezoid weights over proton bins jp0-1..n-1
// replace the values in place (fp_shared_core.c:1473-1503).  np_min[je] is
// jp0 - 1, or n when no proton bin contributes.
__global__ void __launch_bounds__(kTabBlock) k_fqe_weights(int n, const double *__restrict__ pdp, double *fqe,
                                                            int *np_min)
{
    const int je = blockIdx.x * kTabBlock + threadIdx.x;
    if (je >= n) return;
    int npmin = 1;
    for (int jp = 1; jp < n; jp++)
        if (fqe[(size_t)jp * n + je] < 1.0e-80) npmin = jp;
    const int jp0 = npmin < 1 ? 1 : (npmin > n ? n : npmin);
    for (int jp = 0; jp < n; jp++) {
        const size_t at = (size_t)jp * n + je;
        const double v = fqe[at];
        double w = 0.0;
        if (jp0 < n) {
            if (jp == jp0 - 1) w = 0.5 * v * pdp[jp0];
            else if (jp >= jp0 && jp < n - 1) w = 0.5 * v * (pdp[jp] + pdp[jp + 1]);
            else if (jp == n - 1 && jp >= jp0) w = 0.5 * v * pdp[jp];
        }
        fqe[at] = w;
    }
    np_min[je] = (jp0 < n) ? (jp0 - 1) : n;
}

// out[k][j] for k < nbins on the accumulated log grid a_k = lo + (k+1) d - d
// summed term by term as the reference's loops do (so a_k is bit-identical
// to theirs, fp_shared_core.c:1735-1741); kind 0 gamma(E_k, p_j), 1 neutrino(E_k, p_j), 2 IC
// emissivity(nu_k, gamma_j = sqrt(1 + p_j^2), z).
__global__ void __launch_bounds__(kTabBlock) k_emis_table(int kind, int n, int nbins, double lo, double d, double z,
                                                           const double *__restrict__ mom, double *out)
{
    const int j = blockIdx.x * kTabBlock + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= n || k >= nbins) return;
    double a = lo - d;
    for (int kk = 0; kk <= k; kk++) a += d;
    const double e = pow(10.0, a);
    const double p = mom[j];
    double v;
    if (kind == 0) v = chk_gamma_kernel(e, p);
    else if (kind == 1) v = chk_nu_kernel(e, p);
    else v = chk_ic_emissivity_kernel(e, sqrt(1.0 + p * p), z);
    out[(size_t)k * n + j] = v;
}

}  // namespace

extern "C" int ck_secondary_weights(int n, const double *e_energy, const double *pp, const double *beta_p,
                                    const double *pdp, double *fqe_t, int *np_min, void *stream)
{
    if (n <= 0 || n > 65535) return 1;
    const hipStream_t st = (hipStream_t)stream;
    const unsigned gx = (unsigned)((n + kTabBlock - 1) / kTabBlock);
    hipLaunchKernelGGL(k_fqe_vals, dim3(gx, (unsigned)n), dim3(kTabBlock), 0, st, n, e_energy, pp, beta_p, fqe_t);
    hipLaunchKernelGGL(k_fqe_weights, dim3(gx), dim3(kTabBlock), 0, st, n, pdp, fqe_t, np_min);
    return hipGetLastError() != hipSuccess;
}

extern "C" int ck_emission_kernel_table(int kind, int n, int nbins, double lo, double hi, double z,
                                        const double *mom, double *out, void *stream)
{
    if (n <= 0 || nbins <= 0 || nbins > 65535 || kind < 0 || kind > 2) return 1;
    const unsigned gx = (unsigned)((n + kTabBlock - 1) / kTabBlock);
    hipLaunchKernelGGL(k_emis_table, dim3(gx, (unsigned)nbins), dim3(kTabBlock), 0, (hipStream_t)stream, kind, n,
                       nbins, lo, (hi - lo) / (double)nbins, z, mom, out);
    return hipGetLastError() != hipSuccess;
}
