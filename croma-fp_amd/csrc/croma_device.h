// croma_device.h -- device-side physics and the lane-group solver shared by
// the fused kernel instantiations (croma_fused_n*.hip) and the plain-path
// kernels (croma_kernels.hip).  Restates the reference CPU path:
// FP_Coef.c:438-849, fp_shared_core.c:729-1081, Chang_Cooper.c:22-184,
// fp_shared_core.c:226-257, tracer_fp_cpu.c:88-196, tracer_fp_batch.c:651-827.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "croma_kernels.h"

// Physical constants (CONSTANTS.c:3-27).
#define K_C 2.999e+10
#define K_GEV 1.6e-3
#define K_MSUN 1.99e+33
#define K_MPC 3.0857e+24
#define K_KPC 3.0857e+21
#define K_GYR 3.1536e+16
#define K_MP 0.938
#define K_ME 0.511e-3
#define K_PI 3.14159265358979323846

#define ERR_PIVOT 1
#define ERR_NONFINITE 2

// ---------------------------------------------------------------------------
// Device physics (restated from the reference, contraction off so that the
// coefficient arithmetic follows the same rounding as the CPU path).
// ---------------------------------------------------------------------------

__device__ __forceinline__ double d_coulomb_kernel(double x)
{
#pragma clang fp contract(off)
    return 0.5 * sqrt(K_PI) * erf(sqrt(x)) - sqrt(x) * exp(-x);
}

// b_Coulomb_p (FP_Coef.c:438-453)
__device__ double d_b_coulomb_p(double nicm, double p, double kBT)
{
#pragma clang fp contract(off)
    const double p2 = p * p;
    const double gamma = sqrt(1.0 + p2);
    const double betap2 = p2 / (1.0 + p2);
    const double kinetic = gamma - 1.0;
    const double Ep = gamma * K_MP - K_MP;
    const double xp = Ep / kBT;
    const double xe = xp * (K_ME / K_MP);
    const double sum = d_coulomb_kernel(xe) + d_coulomb_kernel(xp) * (K_ME / K_MP);
    const double thermal = 1.0 - kBT / (2.0 * K_MP * kinetic);
    if (thermal < 0.0) return 0.0;
    return K_GYR * K_C / (K_MP * K_GEV) * sum * nicm * (1.0 / betap2) * thermal * 3.5e-29;
}

__device__ double d_I_theta(double x)
{
#pragma clang fp contract(off)
    return pow(x, 4.0) / 4.0 + pow(x, 2.0) - (1.0 + 2.0 * x * x) * log(x) - 5.0 / 4.0;
}

// accelerationtime_TTD / _ASA (FP_Coef.c:634-666)
__device__ double d_tacc_ttd(double L, double cs, double Ms, double v, int model)
{
#pragma clang fp contract(off)
    const double x = cs / K_C;
    if (model == 1) {
        const double theta = d_I_theta(x);
        if (!(x > 0.0) || !(theta > 0.0) || !(Ms > 0.0) || !(L > 0.0) || !(cs > 0.0)) return 0.0;
        return 2.5e-3 / (x * theta) * pow(Ms / 0.5, -4.0) * (L / (300.0 * K_KPC)) *
               pow(cs / 1500.0e5, -1.0);
    }
    const double kL = 2.0 * K_PI / L;
    const double k_cut = 1.04e+4 * pow(Ms, 4.0) * kL;
    const double kWk_L = 0.5 * pow(v, 2.0);
    const double kWk_cut = kWk_L * pow(k_cut / kL, -0.5);
    return K_C / K_PI / d_I_theta(x) / k_cut / kWk_cut / K_GYR;
}

__device__ double d_tacc_asa(double L, double cs, double Ms, double beta, double psi)
{
#pragma clang fp contract(off)
    return sqrt(6.0 / 5.0) / 12.0 * (K_C / cs / cs) * L / sqrt(beta) / pow(Ms, 3.0) * pow(psi, 3.0) /
           K_GYR;
}

// Momentum-diffusion prefactor for one tracer-segment (prepare_momentumdiff_cell_major,
// fp_shared_core.c:1241-1317 and FP_Coef.c:680-849).  Returns 0 for the off floor
// (D = p^2/(4e5)) and 1 for the filled form D = p^2 * inv * [exp(-E/Epmax)] * exp(-1/p).
__device__ int d_dpp_model(const CkParams &prm, int model, double l_turb, double dv, double cs,
                           double beta, double t_acc_direct, double *inv)
{
#pragma clang fp contract(off)
    if (!(dv > 0.0) || !isfinite(dv) || model == 3) return 0;
    double mach = 0.0;
    if (cs > 0.0) {
        mach = dv / cs;
        if (mach < 0.0) mach = 0.0;
        if (prm.mach_limit > 0.0 && mach > prm.mach_limit) mach = prm.mach_limit;
    }
    if (model == 1) {
        if (mach <= 0.0 || beta <= 0.0) return 0;
        const double t = d_tacc_asa(l_turb * K_MPC, cs, mach, beta, prm.psi);
        *inv = 1.0 / ((2.0 + 2.0) * t);
        return 1;
    }
    if (model == 2) {
        if (mach <= 0.0 || dv <= 0.0) return 0;
        const double t = d_tacc_ttd(l_turb * K_MPC, cs, mach, dv, prm.ttd_tacc_model);
        *inv = 1.0 / ((2.0 + 2.0) * t);
        return 1;
    }
    if (model == 4) {
        if (!(t_acc_direct > 0.0) || !isfinite(t_acc_direct)) return 0;
        *inv = 1.0 / (4.0 * t_acc_direct);
        return 1;
    }
    return 0;
}

// t_acc used by the eta cap (fp_shared_core.c:940-972).
__device__ double d_tacc_for_cap(const CkParams &prm, int model, double l_turb, double dv, double cs,
                                 double beta, double t_acc_direct)
{
#pragma clang fp contract(off)
    if (!(dv > 0.0) || !(l_turb > 0.0) || !(cs > 0.0)) return 0.0;
    double mach = dv / cs;
    if (mach < 0.0) mach = 0.0;
    if (prm.mach_limit > 0.0 && mach > prm.mach_limit) mach = prm.mach_limit;
    if (mach <= 0.0) return 0.0;
    if (model == 1) {
        if (!(beta > 0.0)) return 0.0;
        return d_tacc_asa(l_turb * K_MPC, cs, mach, beta, prm.psi);
    }
    if (model == 2) return d_tacc_ttd(l_turb * K_MPC, cs, mach, dv, prm.ttd_tacc_model);
    if (model == 4) return (t_acc_direct > 0.0 && isfinite(t_acc_direct)) ? t_acc_direct : 0.0;
    return 0.0;
}

// Eta-cap scale given the CR energy of the segment-start state (0 disables).
__device__ double d_eta_scale(const CkParams &prm, int model, double l_turb, double dv, double cs,
                              double beta, double mass, double t_acc_direct, double ecr)
{
#pragma clang fp contract(off)
    if (!(prm.eta_dpp_cap > 0.0) || !isfinite(prm.eta_dpp_cap)) return 1.0;
    if (!(mass > 0.0) || !(l_turb > 0.0) || !(dv > 0.0)) return 1.0;
    const double t = d_tacc_for_cap(prm, model, l_turb, dv, cs, beta, t_acc_direct);
    if (!(t > 0.0) || !isfinite(t)) return 1.0;
    const double q = 0.5 * mass * K_MSUN * dv * dv * dv / (l_turb * K_MPC);
    if (!(ecr > 0.0) || !(q > 0.0) || !isfinite(ecr) || !isfinite(q)) return 1.0;
    const double eta = ecr / (q * t * K_GYR);
    if (!(eta > prm.eta_dpp_cap) || !isfinite(eta)) return 1.0;
    return prm.eta_dpp_cap / eta;
}

// Chang-Cooper face weights W+ = B(-w), W- = B(w) (Chang_Cooper.c:22-40).
// For |w| >= 0.1 the reference's two exponentials reduce exactly to one:
// one argument is exactly 0 and the other exactly -|w|.
__device__ __forceinline__ void d_face_w(double w, double &wp, double &wm)
{
#pragma clang fp contract(off)
    if (fabs(w) < 0.1) {
        const double w2 = w * w;
        const double denom = 1.0 + w2 / 24.0 + (w2 * w2) / 1920.0;
        wp = exp(0.5 * w) / denom;
        wm = exp(-0.5 * w) / denom;
    } else {
        const double aw = fabs(w);
        const double E = exp(-aw);
        const double denom = 1.0 - E;
        if (w > 0.0) {
            wp = aw * 1.0 / denom;
            wm = aw * E / denom;
        } else {
            wp = aw * E / denom;
            wm = aw * 1.0 / denom;
        }
    }
}

__device__ __forceinline__ double d_face_delta(double w)
{
#pragma clang fp contract(off)
    return 1.0 / w - 1.0 / (exp(w) - 1.0);
}

// Face weight w between two rows (cc_prepare_face_weight, Chang_Cooper.c:42-66).
__device__ __forceinline__ bool d_face(double bl, double br, double cl, double cr, double xl, double xr,
                                       double &wp, double &wm, double &w)
{
#pragma clang fp contract(off)
    wp = 0.0;
    wm = 0.0;
    w = 0.0;
    if (fabs(cl) <= 1.0e-200) return false;
    w = 0.5 * (bl + br) / (0.5 * (cl + cr)) * (xr - xl);
    d_face_w(w, wp, wm);
    return true;
}

// One Chang-Cooper row (fp_shared_core.c:660-704 -> Chang_Cooper.c:68-184),
// A = 1.  Inputs: this row (x, dx, B, C, dtT = dt/T), left neighbour
// (xl, Bl, Cl, face weights Wpl/Wml, face valid) and right neighbour
// (xr, Br, Cr, Wpr/Wmr).  kind: 0 interior, 1 left boundary, 2 right boundary.
// delta_l: delta of the left face (used only by the right-boundary C ~ 0 branch).
__device__ __forceinline__ void d_cc_row(int kind, double dt, double x, double dx, double B, double C,
                                         double dtT, double xl, double Bl, double Cl, double Wpl,
                                         double Wml, double delta_l, double xr, double Br, double Cr,
                                         double Wpr, double Wmr, double &a, double &b, double &c)
{
#pragma clang fp contract(off)
    const double t1 = dt / (1.0 * dx);
    if (kind == 0) {
        if (fabs(C) > 1.0e-200) {
            const double dxm_12 = x - xl;
            const double dxm12 = xr - x;
            const double cm_12 = 0.5 * (C + Cl);
            const double cm12 = 0.5 * (C + Cr);
            a = t1 * cm_12 / dxm_12 * Wml;
            b = 1.0 + t1 * (cm_12 / dxm_12 * Wpl + cm12 / dxm12 * Wmr) + dtT;
            c = t1 * cm12 / dxm12 * Wpr;
        } else {
            // delta of both adjacent faces is zero here (Chang_Cooper.c:216-219)
            const double bm_12 = 0.5 * (B + Bl);
            const double bm12 = 0.5 * (B + Br);
            a = -0.0 * bm_12 * dt / (1.0 * dx);
            b = 1.0 + t1 * ((1.0 - 0.0) * bm_12 - 0.0 * bm12) + dtT;
            c = t1 * (1.0 - 0.0) * bm12;
        }
    } else if (kind == 1) {
        if (fabs(C) > 1.0e-200) {
            a = 0.0;
            b = 1.0 + t1 * (0.0 + 0.5 * (C + Cr) / (xr - x) * Wmr) + dtT;
            c = t1 * 0.5 * (C + Cr) / (xr - x) * Wpr;
        } else {
            a = -0.0 * 0.0 * dt / (1.0 * dx);
            b = 1.0 + t1 * 0.0 + dtT;
            c = t1 * (1.0 - 0.0) * 0.5 * (B + Br);
        }
    } else {
        if (fabs(C) > 1.0e-200) {
            a = t1 * (0.5 * (C + Cl)) / (x - xl) * Wml;
            b = 1.0 + t1 * ((0.5 * (C + Cl)) / (x - xl) * Wpl) + dtT;
        } else {
            a = -delta_l * 0.5 * (B + Bl) * dt / (1.0 * dx);
            b = 1.0 + t1 * (0.5 * (Bl + B)) + dtT;
        }
        c = 0.0;
    }
}

// tracer_fp_choose_coeff_interp_segments (tracer_fp_setup.c:73-95)
__device__ __host__ __forceinline__ int d_choose_segments(int nsub, int allow, int min_steps, int max_seg)
{
    int segments = 1;
    if (!allow || nsub <= 1) return 1;
    if (min_steps <= 0) return 1;
    if (max_seg < 1) max_seg = 1;
    const int max_allowed = nsub / min_steps;
    if (max_allowed < 1) return 1;
    while (segments <= 1073741823 && segments * 2 <= max_seg && segments * 2 <= max_allowed) segments *= 2;
    if (segments > nsub) segments = nsub;
    if (segments < 1) segments = 1;
    return segments;
}

__device__ __forceinline__ double d_lerp(double cur, double nxt, double alpha)
{
#pragma clang fp contract(off)
    if (alpha <= 0.0) return cur;
    if (alpha >= 1.0) return nxt;
    return cur + alpha * (nxt - cur);
}

// ---------------------------------------------------------------------------
// Lane-group cross-lane helpers.  G lanes per group, groups aligned in the wave.
// Out-of-group sources are masked by the callers (zero coefficients), so the
// raw shuffles below may read a neighbouring group's lane.
// ---------------------------------------------------------------------------
template <int G>
__device__ __forceinline__ double g_up(double v, int s)
{
    return __shfl_up(v, (unsigned)s, G);
}
template <int G>
__device__ __forceinline__ double g_down(double v, int s)
{
    return __shfl_down(v, (unsigned)s, G);
}
template <int G>
__device__ __forceinline__ double g_sum(double v)
{
    for (int s = G / 2; s > 0; s >>= 1) v += __shfl_xor(v, s, G);
    return v;
}

template <int N>
struct LogG {
    static constexpr int v = (N <= 1) ? 0 : 1 + LogG<N / 2>::v;
};

// Per-species factorised operator held in registers for one segment.
template <int R, int G>
struct Factor {
    double al[R];   // a_i / denom_i            (forward multiplier)
    double iv[R];   // 1 / denom_i
    double P[R];    // prefix product of al inside the lane
    double g[R];    // -cprime_i = c_i / denom_i (backward multiplier)
    double Q[R];    // suffix product of g inside the lane
    double Af[LogG<G>::v + 1];  // forward scan coefficients per step (0 = masked)
    double Qb[LogG<G>::v + 1];  // backward scan coefficients per step
};

// Thomas LU factorisation of the group's system (fp_shared_core.c:226-257):
// denom_i = b_i + a_i cprime_{i-1}, cprime_i = -c_i / denom_i.
template <int R, int G>
__device__ void factorize(int gl, int n, const double (&a)[R], const double (&b)[R], const double (&c)[R],
                          Factor<R, G> &F, int &bad)
{
    // lane product of M_i = [[0, -c_i], [a_i, b_i]] acting on (u, v), cprime = u / v
    double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double n00 = -c[k] * m10, n01 = -c[k] * m11;
        const double n10 = a[k] * m00 + b[k] * m10, n11 = a[k] * m01 + b[k] * m11;
        m00 = n00; m01 = n01; m10 = n10; m11 = n11;
        const double mx = fmax(fmax(fabs(m00), fabs(m01)), fmax(fabs(m10), fabs(m11)));
        if (mx > 0.0 && isfinite(mx)) {
            int e;
            frexp(mx, &e);
            m00 = ldexp(m00, -e); m01 = ldexp(m01, -e); m10 = ldexp(m10, -e); m11 = ldexp(m11, -e);
        }
    }
    // inclusive scan S_l = T_l S_{l-1} (exact power-of-two renormalisation)
#pragma unroll
    for (int s = 1; s < G; s <<= 1) {
        const double o00 = g_up<G>(m00, s), o01 = g_up<G>(m01, s);
        const double o10 = g_up<G>(m10, s), o11 = g_up<G>(m11, s);
        if (gl >= s) {
            const double n00 = m00 * o00 + m01 * o10, n01 = m00 * o01 + m01 * o11;
            const double n10 = m10 * o00 + m11 * o10, n11 = m10 * o01 + m11 * o11;
            m00 = n00; m01 = n01; m10 = n10; m11 = n11;
            const double mx = fmax(fmax(fabs(m00), fabs(m01)), fmax(fabs(m10), fabs(m11)));
            if (mx > 0.0 && isfinite(mx)) {
                int e;
                frexp(mx, &e);
                m00 = ldexp(m00, -e); m01 = ldexp(m01, -e); m10 = ldexp(m10, -e); m11 = ldexp(m11, -e);
            }
        }
    }
    const double uin = g_up<G>(m01, 1), vin = g_up<G>(m11, 1);
    double cprev = (gl == 0) ? 0.0 : uin / vin;
    const int base = gl * R;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double den = b[k] + a[k] * cprev;
        if (!(fabs(den) > 1.0e-300)) bad |= ERR_PIVOT;
        const double inv = 1.0 / den;
        const double cp = (base + k == n - 1) ? 0.0 : -c[k] * inv;
        F.iv[k] = inv;
        F.al[k] = a[k] * inv;
        F.g[k] = -cp;
        cprev = cp;
    }
    double pr = 1.0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        pr *= F.al[k];
        F.P[k] = pr;
    }
    double sf = 1.0;
#pragma unroll
    for (int k = R - 1; k >= 0; --k) {
        sf *= F.g[k];
        F.Q[k] = sf;
    }
    // scan coefficient schedules (masked outside the group)
    double A = F.P[R - 1];
    double Bq = F.Q[0];
    int st = 0;
#pragma unroll
    for (int s = 1; s < G; s <<= 1, ++st) {
        const double Ao = g_up<G>(A, s);
        const double Bo = g_down<G>(Bq, s);
        F.Af[st] = (gl >= s) ? A : 0.0;
        F.Qb[st] = (gl + s < G) ? Bq : 0.0;
        if (gl >= s) A = A * Ao;
        if (gl + s < G) Bq = Bq * Bo;
    }
}

// Solve M x = r for the group, in place; rhs = x + src (src = dt*Q).
template <int R, int G>
__device__ __forceinline__ void apply_solve(int gl, const Factor<R, G> &F, double (&x)[R], const double (&src)[R])
{
    double L[R];
    double prev = 0.0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double r = src[k] + x[k];
        prev = fma(F.al[k], prev, F.iv[k] * r);
        L[k] = prev;
    }
    double B = prev;
    int st = 0;
#pragma unroll
    for (int s = 1; s < G; s <<= 1, ++st) B = fma(F.Af[st], g_up<G>(B, s), B);
    double Din = g_up<G>(B, 1);
    if (gl == 0) Din = 0.0;
#pragma unroll
    for (int k = 0; k < R; ++k) L[k] = fma(F.P[k], Din, L[k]);
    // back substitution x_k = d_k + g_k x_{k+1}
    double M = 0.0;
#pragma unroll
    for (int k = R - 1; k >= 0; --k) {
        M = fma(F.g[k], M, L[k]);
        x[k] = M;
    }
    double Xb = M;
    st = 0;
#pragma unroll
    for (int s = 1; s < G; s <<= 1, ++st) Xb = fma(F.Qb[st], g_down<G>(Xb, s), Xb);
    double Xin = g_down<G>(Xb, 1);
    if (gl == G - 1) Xin = 0.0;
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = fma(F.Q[k], Xin, x[k]);
}

// Build one species' Chang-Cooper rows for the lane (needs neighbour lanes).
template <int R, int G>
__device__ void build_cc(int gl, int n, double dt, const double *__restrict__ xg,
                         const double *__restrict__ dxg, const double (&B)[R], const double (&C)[R],
                         const double (&dtT)[R], double (&a)[R], double (&b)[R], double (&c)[R])
{
    const int base = gl * R;
    double x[R], dx[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        x[k] = xg[base + k];
        dx[k] = dxg[base + k];
    }
    const double Bn = g_down<G>(B[0], 1), Cn = g_down<G>(C[0], 1);
    const double xn = (gl < G - 1) ? xg[base + R] : 0.0;
    // faces base+k (rows base+k, base+k+1); the group's last face is the ghost face (unused)
    double Wp[R], Wm[R], Wd[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double br = (k < R - 1) ? B[k + 1] : Bn;
        const double cr = (k < R - 1) ? C[k + 1] : Cn;
        const double xr = (k < R - 1) ? x[k + 1] : xn;
        double w;
        if (base + k < n - 1) {
            const bool ok = d_face(B[k], br, C[k], cr, x[k], xr, Wp[k], Wm[k], w);
            Wd[k] = w;
            (void)ok;
        } else {
            Wp[k] = 0.0;
            Wm[k] = 0.0;
            Wd[k] = 0.0;
        }
    }
    const double Wpl0 = g_up<G>(Wp[R - 1], 1), Wml0 = g_up<G>(Wm[R - 1], 1);
    const double Wdl0 = g_up<G>(Wd[R - 1], 1);
    const double Bl0 = g_up<G>(B[R - 1], 1), Cl0 = g_up<G>(C[R - 1], 1);
    const double xl0 = (gl > 0) ? xg[base - 1] : 0.0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int m = base + k;
        const int kind = (m == 0) ? 1 : ((m == n - 1) ? 2 : 0);
        const double xl = (k > 0) ? x[k - 1] : xl0;
        const double Bl = (k > 0) ? B[k - 1] : Bl0;
        const double Cl = (k > 0) ? C[k - 1] : Cl0;
        const double Wpl = (k > 0) ? Wp[k - 1] : Wpl0;
        const double Wml = (k > 0) ? Wm[k - 1] : Wml0;
        const double wl = (k > 0) ? Wd[k - 1] : Wdl0;
        const double xr = (k < R - 1) ? x[k + 1] : xn;
        const double Br = (k < R - 1) ? B[k + 1] : Bn;
        const double Cr = (k < R - 1) ? C[k + 1] : Cn;
        double delta_l = 0.0;
        if (kind == 2 && !(fabs(C[k]) > 1.0e-200) && fabs(Cl) > 1.0e-200) delta_l = d_face_delta(wl);
        d_cc_row(kind, dt, x[k], dx[k], B[k], C[k], dtT[k], xl, Bl, Cl, Wpl, Wml, delta_l, xr, Br, Cr, Wp[k],
                 Wm[k], a[k], b[k], c[k]);
    }
}

// ---------------------------------------------------------------------------
// Fused persistent kernel
// ---------------------------------------------------------------------------
template <int N, int R, int WPB, bool PROT>
__global__ __launch_bounds__(WPB * 64) void k_fused(CkGrid g, const CkSlot *__restrict__ slots,
                                                    const CkTask *__restrict__ tasks, int ntasks,
                                                    int *__restrict__ counter, int *__restrict__ err)
{
    constexpr int G = N / R;
    constexpr int GPW = 64 / G;
    __shared__ double s_crp[PROT ? WPB : 1][PROT ? GPW : 1][PROT ? N : 1];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int grp = lane / G;
    const int gl = lane % G;
    const int base = gl * R;
    int bad = 0;

    for (;;) {
        // wave-uniform grab of GPW consecutive tasks (cost-ordered list)
        int tbase = 0;
        if (lane == 0) tbase = atomicAdd(counter, GPW);
        tbase = __shfl(tbase, 0, 64);
        if (tbase >= ntasks) break;
        const int tid = tbase + grp;
        if (tid < ntasks) {
        const CkTask task = tasks[tid];
        const CkSlot &S = slots[task.slot];
        const int t = task.tracer;
        const CkParams &prm = S.prm;
        const int nsub = S.nsub[t];
        const int non = S.n_on[t];
        const int nseg = d_choose_segments(nsub, S.allow_interp, prm.coeff_interp_min_steps,
                                           prm.coeff_interp_max_segments);
        const int on_start = (nsub - non) / 2;
        const int on_end = on_start + non;
        const double dt = S.dt_snap / (double)nsub;
        const double base_dt = dt;
        constexpr bool prot = PROT;
        const size_t roff = (size_t)t * N + base;
        const double mass = S.mass ? S.mass[t] : 0.0;
        const bool dis = S.dis ? (S.dis[t] != 0) : false;

        double xe[R], xp[R], qp[R], qe[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            xe[k] = S.cre[roff + k];
            xp[k] = prot ? S.crp[roff + k] : 0.0;
            qp[k] = (prot && S.qpi) ? S.qpi[roff + k] : 0.0;
            qe[k] = S.qepri ? S.qepri[roff + k] : 0.0;
        }
        double bgc[CK_NBG], bgn[CK_NBG];
#pragma unroll
        for (int f = 0; f < CK_NBG; ++f) {
            bgc[f] = S.bgc[f][t];
            bgn[f] = S.bgn[f][t];
        }

        for (int iseg = 0; iseg < nseg; ++iseg) {
            const int fb = (int)(((long long)iseg * nsub) / nseg);
            const int fe = (int)(((long long)(iseg + 1) * nsub) / nseg);
            const int steps = fe - fb;
            if (steps <= 0) continue;
            const double alpha = (nseg > 1) ? ((double)iseg + 0.5) / (double)nseg : 0.0;
            double bg[CK_NBG];
#pragma unroll
            for (int f = 0; f < CK_NBG; ++f) bg[f] = (nseg > 1) ? d_lerp(bgc[f], bgn[f], alpha) : bgc[f];
            const double n_gas = bg[CK_BG_NGAS], kbt = bg[CK_BG_KBT], bfield = bg[CK_BG_B];
            const double divv = dis ? 0.0 : bg[CK_BG_DIVV];
            const double lturb = bg[CK_BG_LTURB], dv = bg[CK_BG_DV], cs = bg[CK_BG_CS], beta = bg[CK_BG_BETA];
            int sob = (on_start > fb) ? on_start : fb;
            int soe = (on_end < fe) ? on_end : fe;
            if (soe < sob) soe = sob;
            const bool windowed = (soe - sob) < steps;
            const int ls = sob - fb, le = ls + (soe - sob);
            double tacc_d = prm.t_acc_direct_gyr;
            {
                const double tmid = S.elapsed + ((double)fb + 0.5 * (double)steps) * base_dt;
                const bool direct_on = prm.dpp_model == 4 && isfinite(prm.t_acc_direct_gyr) &&
                                       prm.t_acc_direct_gyr > 0.0 && isfinite(prm.t_off_gyr) &&
                                       prm.t_off_gyr > 0.0 && isfinite(prm.t_acc_off_gyr) &&
                                       prm.t_acc_off_gyr > 0.0;
                if (direct_on && tmid < prm.t_off_gyr) tacc_d = prm.t_acc_off_gyr;
            }

            // --- losses (prepare_cre/crp_losses_1d) ---
            double radpe[R], radpp[R], dtTp[R], dtTe[R];
            {
#pragma clang fp contract(off)
                const double b_micro = bfield * 1.0e+6;
                const double b_ratio = b_micro / 3.2;
                const double synch_coeff = 4.8e-4 * b_ratio * b_ratio * 9.11e-28 * 3.0e+10 * 3.1536e+16;
                const double coul_coeff = 3.05e-29 * n_gas * K_C * K_GYR / (K_ME * K_GEV);
                const double adia = 1.0 / 3.0 * divv;
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int j = base + k;
                    const double pe = g.ep[j];
                    const double pe2 = g.e_p2[j];
                    const double beta2 = pe2 / (1.0 + pe2);
                    const double ric = (nseg > 1) ? d_lerp(S.ric_c[j], S.ric_n[j], alpha) : S.ric_c[j];
                    const double coul = coul_coeff * (1.0 + log(pe / n_gas) / 74.8) / beta2;
                    radpe[k] = synch_coeff * pe2 + ric + coul + adia * pe;
                    dtTe[k] = dt / 1.0e+100;
                }
                if (prot) {
                    const double tl_pref = 1.0 / (0.5 * K_C * n_gas) / K_GYR;
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        const int j = base + k;
                        const double p = g.pp[j];
                        radpp[k] = d_b_coulomb_p(n_gas, p, kbt) + adia * p;
                        dtTp[k] = dt / (tl_pref / g.p_sig[j]);
                    }
                }
            }
            // --- momentum diffusion: on-set prefactor and eta cap ---
            double inv_on = 0.0;
            const int mode_on = d_dpp_model(prm, prm.dpp_model, lturb, dv, cs, beta, tacc_d, &inv_on);
            double scale = 1.0;
            if (prm.dpp_model != 3 && prm.eta_dpp_cap > 0.0) {
                double part = 0.0;
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int j = base + k;
                    const double wp_ = g.p_ecr_w[j];
                    const double we_ = g.e_ecr_w[j];
                    if (prot && wp_ > 0.0 && isfinite(xp[k])) part += wp_ * xp[k] * g.pdp[j];
                    if (we_ > 0.0 && isfinite(xe[k])) part += we_ * xe[k] * g.edp[j];
                }
                const double ecr = g_sum<G>(part);
                scale = d_eta_scale(prm, prm.dpp_model, lturb, dv, cs, beta, mass, tacc_d, ecr);
            }
            // --- secondary electrons from the segment-start protons ---
            double inje[R];
            if (prot && S.fqe_t) {
#pragma unroll
                for (int k = 0; k < R; ++k) s_crp[wave][grp][base + k] = xp[k];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                double acc[R];
#pragma unroll
                for (int k = 0; k < R; ++k) acc[k] = 0.0;
                for (int jp = S.jp0; jp < N; ++jp) {
                    const double nn = s_crp[wave][grp][jp];
                    const double *wrow = S.fqe_t + (size_t)jp * N + base;
#pragma unroll
                    for (int k = 0; k < R; ++k) acc[k] += nn * wrow[k];
                }
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int k = 0; k < R; ++k) inje[k] = acc[k] * K_GYR * n_gas + qe[k];
            } else {
#pragma unroll
                for (int k = 0; k < R; ++k) inje[k] = qe[k];
            }
            double srcp[R], srce[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                srcp[k] = dt * qp[k];
                srce[k] = dt * inje[k];
            }

            // --- steps, in up to three runs of constant coefficient set ---
            int s0 = 0;
            while (s0 < steps) {
                const bool on = !windowed || (s0 >= ls && s0 < le);
                int s1;
                if (!windowed) s1 = steps;
                else if (s0 < ls) s1 = ls;
                else if (s0 < le) s1 = le;
                else s1 = steps;
                // a segment wholly before the on-window has ls > steps
                if (s1 > steps) s1 = steps;
                const int use_model = on ? mode_on : 0;
                // assemble + factorise electrons (and protons)
                Factor<R, G> Fe, Fp;
                {
                    double Bv[R], Cv[R], a[R], b[R], c[R];
#pragma unroll
                    for (int k = 0; k < R; ++k) {
#pragma clang fp contract(off)
                        const int j = base + k;
                        double D;
                        if (use_model) D = g.e_p2[j] * inv_on * g.e_expcut[j];
                        else D = g.e_p2[j] * (1.0 / (4.0 * 1.0e+5));
                        if (on) D *= scale;
                        Cv[k] = D;
                        Bv[k] = radpe[k] - 2.0 / g.ep[j] * D;
                    }
                    build_cc<R, G>(gl, N, dt, g.ep, g.edp, Bv, Cv, dtTe, a, b, c);
                    factorize<R, G>(gl, N, a, b, c, Fe, bad);
                    if (prot) {
#pragma unroll
                        for (int k = 0; k < R; ++k) {
#pragma clang fp contract(off)
                            const int j = base + k;
                            double D;
                            if (use_model) D = g.p_p2[j] * inv_on * exp(-g.pE[j] / prm.epmax) * g.p_expcut[j];
                            else D = g.p_p2[j] * (1.0 / (4.0 * 1.0e+5));
                            if (on) D *= scale;
                            Cv[k] = D;
                            Bv[k] = radpp[k] - 2.0 / g.pp[j] * D;
                        }
                        build_cc<R, G>(gl, N, dt, g.pp, g.pdp, Bv, Cv, dtTp, a, b, c);
                        factorize<R, G>(gl, N, a, b, c, Fp, bad);
                    }
                }
                for (int s = s0; s < s1; ++s) {
                    if (prot) apply_solve<R, G>(gl, Fp, xp, srcp);
                    apply_solve<R, G>(gl, Fe, xe, srce);
                }
                s0 = s1;
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (!isfinite(xe[k]) || !isfinite(xp[k])) bad |= ERR_NONFINITE;
            S.cre[roff + k] = xe[k];
            if (prot) S.crp[roff + k] = xp[k];
        }
        }
    }
    if (bad) atomicOr(err, bad);
}

