// croma_ic.hip -- inverse-Compton cooling rows on the device.
//
// rad_IC_cool (FP_Coef.c:361-375) -> b_IC (:485-516) -> b_IC_integrand
// (:517-539) with IC_kernel (:378-388) and the CMB photon density (:390-405):
// for every electron bin (Lorentz factor gamma) a trapezoid rule over the CMB
// photon energy e0 (log step 0.05 between the bounds of :413-436, computed on
// the host) of n_cmb(e0) * I(e0), where I(e0) is itself a trapezoid rule over
// the scattered energy eps in [e0, 4 gamma^2 e0 / (1 + 4 e0 gamma)].
//
// One block per (bin, redshift): the threads evaluate the outer nodes in
// parallel (each runs its inner integral), then one thread sums the outer
// trapezoid in the reference's order.  Log abscissae are accumulated by
// repeated addition exactly as the reference does.
#include <hip/hip_runtime.h>
#include <math.h>

#include "croma_kernels.h"

#define IC_C_LIGHT 2.999e+10
#define IC_H_PLANCK 6.63e-27
#define IC_GYR 3.1536e+16
#define IC_PI 3.14159265358979323846
#define IC_MAX_NODES 2048

__device__ static double ic_kernel(double eps0, double gamma, double eps)
{
#pragma clang fp contract(off)
    const double re = 2.82e-13;
    const double kappa = eps / (4.0 * eps0 * gamma * (gamma - eps));
    if (kappa <= 0.0) return 0.0;
    const double q = 4.0 * eps0 * kappa * gamma;
    const double C = 2.0 * kappa * log(kappa) + (1.0 + 2.0 * kappa) * (1.0 - kappa) +
                     (q * q) / (2.0 * (1.0 + 4.0 * eps0 * gamma * kappa)) * (1.0 - kappa);
    return 2.0 * IC_PI * re * re * IC_C_LIGHT / (gamma * gamma * eps0) * C;
}

__device__ static double ic_cmb_n(double eps, double z)
{
#pragma clang fp contract(off)
    const double T0 = 2.757, meg = 9.11e-28, k = 1.38e-16, tiny = 1.0e-40;
    const double Tcmb = T0 * (1.0 + z);
    const double x = eps * meg * IC_C_LIGHT * IC_C_LIGHT / (k * Tcmb);
    const double r = meg / IC_H_PLANCK;
    const double pre = 8.0 * IC_PI * IC_C_LIGHT * IC_C_LIGHT * IC_C_LIGHT * pow(r, 3.0);
    if (x > 90.0 * log(10.0)) return 0.0;
    const double em1 = exp(x) - 1.0;
    if (em1 < tiny)
        return pre * eps / (meg * IC_C_LIGHT * IC_C_LIGHT * (1.0 + 0.5 * x + x / 6.0) / (k * Tcmb));
    return pre * (eps * eps) / em1;
}

// 10^a at the inner nodes as exp(a ln 10): the node follows the same rounded
// abscissa a; the rounding of a ln 10 moves the result by at most
// |a ln 10| 2^-53 relative (|a| < 12 here, below 3e-15), and exp costs a
// fraction of pow in the quadrature's inner loop.
__device__ static double ic_exp10(double a) { return exp(a * 2.302585092994045684); }

__device__ static double ic_inner(double eps0, double gamma, double emin, double emax)
{
#pragma clang fp contract(off)
    const double da = 0.50e-1;
    const int N = (int)round((log10(emax) - log10(emin)) / da);
    double epsnew = emin, a = log10(emin), newC = ic_kernel(eps0, gamma, epsnew), sum = 0.0;
    for (int i = 0; i < N; ++i) {
        a = a + da;
        const double eps = epsnew;
        epsnew = ic_exp10(a);
        const double deps = epsnew - eps;
        const double C = newC;
        newC = ic_kernel(eps0, gamma, epsnew);
        sum = sum + 0.5 * (eps * C + epsnew * newC) * deps;
    }
    return sum;
}

// zb[iz] = {z, e0min, e0max}; rows[iz * n + j] for j < n, ghosts[iz * 2 + {0, 1}]
__global__ __launch_bounds__(256) void k_ic_rows(int n, const double *__restrict__ ep, double epm1, double epp1,
                                                 const double *__restrict__ zb, double *__restrict__ rows,
                                                 double *__restrict__ ghosts, int *__restrict__ err)
{
#pragma clang fp contract(off)
    __shared__ double s_f[IC_MAX_NODES + 1];
    __shared__ double s_e[IC_MAX_NODES + 1];
    const int j = blockIdx.x;  // 0..n+1
    const int iz = blockIdx.y;
    const double z = zb[3 * iz], lo = zb[3 * iz + 1], hi = zb[3 * iz + 2];
    const double p = (j < n) ? ep[j] : (j == n ? epm1 : epp1);
    const double gamma = sqrt(1.0 + p * p);
    const double da = 0.50e-1;
    const int N = (int)round((log10(hi) - log10(lo)) / da);
    if (N > IC_MAX_NODES || N < 0) {
        if (threadIdx.x == 0) atomicOr(err, 4);
        return;
    }
    for (int i = threadIdx.x; i <= N; i += blockDim.x) {
        double a = log10(lo);
        for (int k = 0; k < i; ++k) a = a + da;
        const double e0 = (i == 0) ? lo : pow(10.0, a);
        const double emax = gamma * 4.0 * e0 * gamma / (1.0 + 4.0 * e0 * gamma);
        s_f[i] = ic_cmb_n(e0, z) * ic_inner(e0, gamma, e0, emax);
        s_e[i] = e0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int i = 0; i < N; ++i) sum = sum + 0.5 * (s_f[i] + s_f[i + 1]) * (s_e[i + 1] - s_e[i]);
        const double v = sum * IC_GYR;
        if (j < n) rows[(size_t)iz * n + j] = v;
        else if (ghosts) ghosts[2 * iz + (j - n)] = v;
    }
}

extern "C" int ck_ic_rows(int n, const double *ep_dev, double epm1, double epp1, int nz, const double *zb_dev,
                          double *rows_dev, double *ghosts_dev, int *err_dev, void *stream)
{
    if (n <= 0 || nz <= 0) return nz == 0 ? 0 : -1;
    hipLaunchKernelGGL(k_ic_rows, dim3(n + 2, nz), dim3(256), 0, (hipStream_t)stream, n, ep_dev, epm1, epp1, zb_dev,
                       rows_dev, ghosts_dev, err_dev);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
