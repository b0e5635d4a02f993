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
    // For x >= 40, erf(sqrt x) rounds to 1 and sqrt(x) exp(-x) < 3e-17 is
    // below half an ulp of 0.5 sqrt(pi): the expression equals that constant
    // exactly (the thermal proton term of every bin well above kT).
    if (x >= 40.0) return 0.5 * sqrt(K_PI);
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

// The same rate from per-bin factors of the segment build (Ep = gamma m_p -
// m_p, i2mk = 1 / (2 m_p (gamma - 1)), ib2 = 1 / beta^2) with the divisions
// by kBT done as one reciprocal: within a few ulp of d_b_coulomb_p.
__device__ __forceinline__ double d_b_coulomb_p_tab(double nicm, double kBT, double ikbt, double Ep, double i2mk,
                                                    double ib2)
{
#pragma clang fp contract(off)
    const double xp = Ep * ikbt;
    const double xe = xp * (K_ME / K_MP);
    const double sum = d_coulomb_kernel(xe) + d_coulomb_kernel(xp) * (K_ME / K_MP);
    const double thermal = 1.0 - kBT * i2mk;
    if (thermal < 0.0) return 0.0;
    return K_GYR * K_C / (K_MP * K_GEV) * sum * nicm * ib2 * thermal * 3.5e-29;
}

// FP32 form of d_b_coulomb_p_tab (the FP32-coefficient instances).
__device__ __forceinline__ float d_coulomb_kernelf(float x)
{
    if (x >= 40.0f) return (float)(0.5 * 1.7724538509055160273);
    const float sx = sqrtf(x);
    return (float)(0.5 * 1.7724538509055160273) * erff(sx) - sx * __expf(-x);
}

__device__ __forceinline__ float d_b_coulomb_p_tabf(float nicm, float kBT, float ikbt, float Ep, float i2mk, float ib2)
{
#pragma clang fp contract(off)
    const float xp = Ep * ikbt;
    const float xe = xp * (float)(K_ME / K_MP);
    const float sum = d_coulomb_kernelf(xe) + d_coulomb_kernelf(xp) * (float)(K_ME / K_MP);
    const float thermal = 1.0f - kBT * i2mk;
    if (thermal < 0.0f) return 0.0f;
    return (float)(K_GYR * K_C / (K_MP * K_GEV) * 3.5e-29) * sum * nicm * ib2 * thermal;
}

// Integer powers below are products (pow with an integer exponent differs
// from them by an ulp or two; these enter t_acc, and through it D_pp, only
// multiplicatively).
__device__ double d_I_theta(double x)
{
#pragma clang fp contract(off)
    const double x2 = x * x;
    return x2 * x2 / 4.0 + x2 - (1.0 + 2.0 * x * x) * log(x) - 5.0 / 4.0;
}

// accelerationtime_TTD / _ASA (FP_Coef.c:634-666)
__device__ double d_tacc_ttd(double L, double cs, double Ms, double v, int model)
{
#pragma clang fp contract(off)
    const double x = cs / K_C;
    if (model == 1) {
        const double theta = d_I_theta(x);
        if (!(x > 0.0) || !(theta > 0.0) || !(Ms > 0.0) || !(L > 0.0) || !(cs > 0.0)) return 0.0;
        const double m2 = (Ms / 0.5) * (Ms / 0.5);
        return 2.5e-3 / (x * theta) * (1.0 / (m2 * m2)) * (L / (300.0 * K_KPC)) * (1.0 / (cs / 1500.0e5));
    }
    const double kL = 2.0 * K_PI / L;
    const double k_cut = 1.04e+4 * ((Ms * Ms) * (Ms * Ms)) * kL;
    const double kWk_L = 0.5 * (v * v);
    const double kWk_cut = kWk_L * (1.0 / sqrt(k_cut / kL));
    return K_C / K_PI / d_I_theta(x) / k_cut / kWk_cut / K_GYR;
}

__device__ double d_tacc_asa(double L, double cs, double Ms, double beta, double psi)
{
#pragma clang fp contract(off)
    return sqrt(6.0 / 5.0) / 12.0 * (K_C / cs / cs) * L / sqrt(beta) / (Ms * Ms * Ms) * (psi * psi * psi) /
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

// ---------------------------------------------------------------------------
// Substep count of one tracer from its current backgrounds
// (tracer_fp_prepare_nsub_loss_endpoints, tracer_fp_nsub.c:242-282;
// compute_nsubsteps_core :94-187; ladders :37-92): loss times p / b(p) at
// bins 0 and n-1 of both species, dt_sub = safety * min, quantized down to
// the mantissa ladder, nsub = max(10, ceil(dt_snap / dt_sub)), optional cap;
// the on-window from f_eddy t_eddy quantized to {1/8, 1/4, 1/3, 1/2, 1}.
// The loss terms are the endpoint forms of the reference (b_synch,
// b_Coulomb_e, b_Coulomb_p; FP_Coef.c:438-469), association as written.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double d_nsub_b_synch(double p, double B)
{
#pragma clang fp contract(off)
    const double b_micro = B * 1.0e+6, b_ratio = b_micro / 3.2;
    return 4.8e-4 * p * p * b_ratio * b_ratio * 9.11e-28 * 3.0e+10 * 3.1536e+16;
}

__device__ __forceinline__ double d_nsub_b_coulomb_e(double nicm, double pe)
{
#pragma clang fp contract(off)
    const double pe2 = pe * pe, beta2 = pe2 / (1.0 + pe2), A = 3.05e-29;
    return A * nicm * (1.0 + log(pe / nicm) / 74.8) / (K_ME * K_GEV) * K_C * K_GYR / beta2;
}

__device__ __forceinline__ double d_quantize_dt(double t, const double *pow10)
{
#pragma clang fp contract(off)
    const double ladder[10] = {1.0, 1.25, 1.6, 2.0, 2.5, 3.2, 4.0, 5.0, 6.4, 8.0};
    if (!(t > 0.0) || !isfinite(t)) return t;
    const double e10 = floor(log10(t));
    const double scale = (e10 >= -40.0 && e10 <= 40.0) ? pow10[(int)e10 + 40] : pow(10.0, e10);
    const double norm = t / scale;
    double chosen = ladder[0];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        if (ladder[i] <= norm) chosen = ladder[i];
        else break;
    }
    return chosen * scale;
}

__device__ __forceinline__ double d_quantize_on(double f)
{
#pragma clang fp contract(off)
    const double bins[5] = {1.0 / 8.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 2.0, 1.0};
    if (!(f > 0.0) || !isfinite(f)) return 1.0;
    if (f >= 1.0) return 1.0;
    double chosen = bins[0], best = fabs(f - chosen);
#pragma unroll
    for (int i = 1; i < 5; ++i) {
        const double d = fabs(f - bins[i]);
        if (d < best) {
            chosen = bins[i];
            best = d;
        }
    }
    return chosen;
}

__device__ __forceinline__ void d_nsub_one(const CkGrid &g, const CkParams &prm, double dt_snap, double ric_lo,
                                           double ric_hi, double n_gas, double kbt, double bf, double divv,
                                           double lturb, double dv, int &ns_out, int &non_out)
{
#pragma clang fp contract(off)
    const int n = g.n;
    const double pp0 = g.pp[0], ppN = g.pp[n - 1], ep0 = g.ep[0], epN = g.ep[n - 1];
    const double lo_p = d_b_coulomb_p(n_gas, pp0, kbt) + 1.0 / 3.0 * divv * pp0;
    const double hi_p = d_b_coulomb_p(n_gas, ppN, kbt) + 1.0 / 3.0 * divv * ppN;
    const double lo_e = d_nsub_b_synch(ep0, bf) + ric_lo + d_nsub_b_coulomb_e(n_gas, ep0) + 1.0 / 3.0 * divv * ep0;
    const double hi_e = d_nsub_b_synch(epN, bf) + ric_hi + d_nsub_b_coulomb_e(n_gas, epN) + 1.0 / 3.0 * divv * epN;
    const double te0 = ep0 / lo_e, tp0 = pp0 / lo_p;
    const double teN = epN / hi_e, tpN = ppN / hi_p;
    const double tmin = fmin(fabs(tp0), fmin(fabs(teN), fmin(fabs(te0), fabs(tpN))));
    double target = dt_snap / 10.0, on_fraction = 1.0;
    if (isfinite(tmin) && tmin > 0.0) {
        const double st = prm.tracer_nsub_safety * tmin;
        if (st < target) target = st;
    }
    if (target <= 0.0 || !isfinite(target)) target = dt_snap / 10.0;
    double used = d_quantize_dt(target, g.pow10);
    if (!(used > 0.0) || !isfinite(used) || used > target) used = target;
    double raw = ceil(dt_snap / used);
    if (!isfinite(raw) || raw > 1000000.0) raw = 1000000.0;
    if (raw < 10.0) raw = 10.0;
    int ns = (int)raw;
    if (prm.tracer_nsub_max > 0 && ns > prm.tracer_nsub_max) ns = prm.tracer_nsub_max;
    if (prm.reacc_window_mode == 1 && prm.f_eddy > 0.0 && dv > 0.0) {
        const double t_eddy = lturb * 3.0857e24 / dv / 3.1536e16;
        if (prm.f_eddy * t_eddy < dt_snap) on_fraction = d_quantize_on(prm.f_eddy * t_eddy / dt_snap);
    }
    int non = (int)floor((double)ns * on_fraction + 0.5);
    if (non < 1) non = 1;
    if (non > ns) non = ns;
    ns_out = ns;
    non_out = non;
}

// Cost bucket of the task order: 256 buckets per octave of the cost.
__device__ __forceinline__ int d_cost_bucket(double cost)
{
    if (cost < 1.0) return 0;
    const int k = (int)(log2(cost) * 256.0) + 1;
    return k >= CK_NSUB_HIST ? CK_NSUB_HIST - 1 : k;
}

// Branch-free (selects), so loads of both ends can issue together.
__device__ __forceinline__ double d_lerp(double cur, double nxt, double alpha)
{
#pragma clang fp contract(off)
    const double v = cur + alpha * (nxt - cur);
    return (alpha <= 0.0) ? cur : ((alpha >= 1.0) ? nxt : v);
}

// ---------------------------------------------------------------------------
// Lane-group cross-lane helpers.  G lanes per group, groups aligned in the
// wave.  Shifts inside 16-lane rows use DPP (row_shr / row_shl), across rows
// DPP row_bcast, wave shifts or readlane (the per-segment group sum of the
// eta cap uses __shfl_xor).  A lane whose
// source lies outside its group reads a neighbouring group's (finite) value,
// or 0, and every caller discards it or multiplies it by a zero coefficient.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ double dpp64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
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

// value of lane L (a wave-uniform lane index) as a scalar
__device__ __forceinline__ double rdlane64(double v, int L)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), L);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), L);
    return __hiloint2double(hi, lo);
}

// lane l - S / l + S inside the 16-lane row (0 where that lane is outside the row)
template <int S>
__device__ __forceinline__ double row_up(double v) { return dpp64<0x110 + S>(v); }
template <int S>
__device__ __forceinline__ double row_dn(double v) { return dpp64<0x100 + S>(v); }
// the value at lane 15 of the row before (row_bcast:15; row 0: 0).  Rows 1
// and 3 use it; rows 0 and 2 multiply it by a zero multiplier (it is finite).
__device__ __forceinline__ double bcast15(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x142, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x142, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
// lane l - 1 / l + 1 across the whole wave (wave_shr:1 / wave_shl:1)
__device__ __forceinline__ double wave_up1(double v) { return dpp64<0x138>(v); }
__device__ __forceinline__ double wave_dn1(double v) { return dpp64<0x130>(v); }

// Factorised Chang-Cooper operator of one species for one coefficient run,
// held in registers.  One implicit substep is x <- M^{-1} (x + src):
//   forward   d_i = al_i d_{i-1} + iv_i x_i + vs_i
//   backward  x_i = d_i + g_i x_{i+1}
// with the lane-level affine scans over the group (Af, Qb: products of al
// resp. g over lane ranges, zero where the source lane is outside the group).
template <int R, int G>
struct Factor {
    static constexpr int LV = LogG<G>::v < 4 ? LogG<G>::v : 4;  // scan levels inside a 16-lane row
    double al[R];   // a_i / denom_i
    double iv[R];   // 1 / denom_i
    double vs[R];   // src_i / denom_i
    double g[R];    // c_i / denom_i (= -cprime_i of the reference)
    double Af[LV + 1];  // forward scan multipliers, zero where the source lane is outside the row / group
    double Qb[LV + 1];  // backward scan multipliers
    double PA, PQ;      // groups wider than a row: in-row inclusive products (prefix of A, suffix of Q)
    double PQ0, PQ2;    // two-row groups: PQ in rows 0 / 2 only (0 elsewhere)
    double Tf[4], Tb[4];  // and (one group per wave) the row totals of those products
};

// Multipliers of the in-row scan levels (and the running in-row products).
template <int G, int ST, int R>
__device__ __forceinline__ void scan_coeffs(int gl, int rl, double &A, double &Bq, Factor<R, G> &F)
{
    if constexpr (ST < Factor<R, G>::LV) {
        constexpr int s = 1 << ST;
        const double Ao = row_up<s>(A);
        const double Bo = row_dn<s>(Bq);
        // a source lane outside the row (or, for groups narrower than a row,
        // outside the group) contributes nothing
        const bool okf = (G >= 16) ? (rl >= s) : (gl >= s);
        const bool okb = (G >= 16) ? (rl + s <= 15) : (gl + s < G);
        F.Af[ST] = okf ? A : 0.0;
        F.Qb[ST] = okb ? Bq : 0.0;
        if (okf) A = A * Ao;
        if (okb) Bq = Bq * Bo;
        scan_coeffs<G, ST + 1, R>(gl, rl, A, Bq, F);
    }
}

// Per-row inputs of a build (loss rates, sources) and the parked spectrum are
// read from memory once per run at an offset that passes through an empty asm:
// the compiler can neither hoist those loads out of the run loop nor forward
// the parked values, so none of them holds registers across the substeps or
// the build.
__device__ __forceinline__ int run_opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Source rows of one run: src_k = scale * (inj_k * ka * kb + q_k) (q: the
// tracer's direct source rows, or rows already scaled in LDS; inj: the
// contracted secondary rows of the segment); an absent one points at the
// lane's rows of the grid's zero table.
struct SrcRows {
    const double *q;
    const double *inj;
    double ka, kb, scale;
};

// exp(x) for x <= 0 (the face factor exp(-|w|/2)): Cody-Waite reduction
// x = n ln2 + r (ln2 split in fdlibm's hi / lo parts), |r| <= ln2 / 2, the
// degree-12 Taylor polynomial (truncation below 2e-16 relative) and a scale by
// 2^n; 0 below -1100.  The same steps in C, run against the C library's exp
// on 2e7 points of [-700, 0], stay within 2 ulp over the normal range.
__device__ __forceinline__ double d_exp_nonpos(double x)
{
    x = fmax(x, -1100.0);
    const double n = rint(x * 1.4426950408889634);
    double r = fma(-n, 6.93147180369123816490e-01, x);
    r = fma(-n, 1.90821492927058770002e-10, r);
    double p = 1.0 / 479001600.0;
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}

// a / b by the hardware reciprocal, one Newton step and one residual
// correction (the step squares the reciprocal's relative error, the
// correction squares the quotient's): the division sequence without the
// operand scaling and fix-up for denormal, infinite or zero divisors, which
// the build's divisors (face diffusion averages above 1e-200, Bernoulli
// denominators in (0, 1], pivots above 1e-300 in magnitude) never are.
// Within an ulp of the IEEE quotient.
__device__ __forceinline__ double d_div_normal(double a, double b)
{
    double r = __builtin_amdgcn_rcp(b);
    const double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    const double q = a * r;
    return fma(r, fma(-b, q, a), q);
}

// 2x2 projective maps of the pivot recurrence: m <- m o (matrix product),
// and renormalisation by a power of two (exact; the map is unchanged).
__device__ __forceinline__ void pm_mul(double &m00, double &m01, double &m10, double &m11, double o00, double o01,
                                       double o10, double o11)
{
    const double n00 = m00 * o00 + m01 * o10, n01 = m00 * o01 + m01 * o11;
    const double n10 = m10 * o00 + m11 * o10, n11 = m10 * o01 + m11 * o11;
    m00 = n00; m01 = n01; m10 = n10; m11 = n11;
}
__device__ __forceinline__ void pm_renorm(double &m00, double &m01, double &m10, double &m11)
{
    const double mx = fmax(fmax(fabs(m00), fabs(m01)), fmax(fabs(m10), fabs(m11)));
    if (mx > 0.0 && isfinite(mx)) {
        const int e = __builtin_amdgcn_frexp_exp(mx);
        m00 = ldexp(m00, -e); m01 = ldexp(m01, -e); m10 = ldexp(m10, -e); m11 = ldexp(m11, -e);
    }
}
// In-row Hillis-Steele levels of the pivot-map scan (shift 1, 2, 4, 8 by DPP;
// a source lane outside the row, or outside a group narrower than a row,
// contributes nothing).
template <int G, int ST>
__device__ __forceinline__ void pm_row_scan(int gl, int rl, double &m00, double &m01, double &m10, double &m11)
{
    constexpr int LV = LogG<G>::v < 4 ? LogG<G>::v : 4;
    if constexpr (ST < LV) {
        constexpr int s = 1 << ST;
        const double o00 = row_up<s>(m00), o01 = row_up<s>(m01);
        const double o10 = row_up<s>(m10), o11 = row_up<s>(m11);
        if ((G >= 16) ? (rl >= s) : (gl >= s)) {
            pm_mul(m00, m01, m10, m11, o00, o01, o10, o11);
            pm_renorm(m00, m01, m10, m11);
        }
        pm_row_scan<G, ST + 1>(gl, rl, m00, m01, m10, m11);
    }
}

// Thomas LU of the rows held in F (on entry al = a, iv = b, g = c), as in
// fp_thomas_solve_ws (fp_shared_core.c:226-257): denom_i = b_i + a_i cprime_{i-1},
// cprime_i = -c_i / denom_i, pivot guard |denom| <= 1e-300.  The pivot
// entering each lane comes from a scan of 2x2 projective maps across the
// group's lanes (renormalised by powers of two, so exact up to rounding of
// the products); inside the lane the reference recurrence is applied.
template <int R, int G, bool NOSRC = false>
__device__ __forceinline__ void factorize(int gl, int n, const double (&sv)[R], Factor<R, G> &F, int &bad)
{
    double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double a = F.al[k], b = F.iv[k], c = F.g[k];
        const double n00 = -c * m10, n01 = -c * m11;
        const double n10 = a * m00 + b * m10, n11 = a * m01 + b * m11;
        m00 = n00; m01 = n01; m10 = n10; m11 = n11;
        // renormalised every second row: two rows grow the entries by at most
        // the product of two rows' coefficient sums, far inside the range
        if ((k & 1) || k == R - 1) pm_renorm(m00, m01, m10, m11);
    }
    // inclusive prefix over the group's lanes, P_l = M_l ... M_0: Hillis-Steele
    // levels inside 16-lane rows (DPP), then the rows of a wider group take the
    // full prefix at the last lane of the row before (readlane)
    const int rl = (int)(threadIdx.x & 15);
    pm_row_scan<G, 0>(gl, rl, m00, m01, m10, m11);
    if constexpr (G > 16) {
        const int row = (int)((threadIdx.x & 63) >> 4);
        double c00, c01, c10, c11;  // full prefix entering this lane's row
        if constexpr (G == 32) {
            // rows (0, 1) and (2, 3) form groups: odd rows take lane 15 / 47
            const double a00 = rdlane64(m00, 15), a01 = rdlane64(m01, 15);
            const double a10 = rdlane64(m10, 15), a11 = rdlane64(m11, 15);
            const double b00 = rdlane64(m00, 47), b01 = rdlane64(m01, 47);
            const double b10 = rdlane64(m10, 47), b11 = rdlane64(m11, 47);
            const bool hi = row >= 2;
            c00 = hi ? b00 : a00; c01 = hi ? b01 : a01; c10 = hi ? b10 : a10; c11 = hi ? b11 : a11;
        } else {
            // one group per wave: carries through rows 0, 1, 2
            double f00 = rdlane64(m00, 15), f01 = rdlane64(m01, 15);
            double f10 = rdlane64(m10, 15), f11 = rdlane64(m11, 15);  // full prefix at lane 15
            c00 = f00; c01 = f01; c10 = f10; c11 = f11;
#pragma unroll
            for (int r = 1; r < 3; ++r) {
                double u00 = rdlane64(m00, 16 * r + 15), u01 = rdlane64(m01, 16 * r + 15);
                double u10 = rdlane64(m10, 16 * r + 15), u11 = rdlane64(m11, 16 * r + 15);
                pm_mul(u00, u01, u10, u11, f00, f01, f10, f11);
                pm_renorm(u00, u01, u10, u11);
                f00 = u00; f01 = u01; f10 = u10; f11 = u11;
                if (row == r + 1) { c00 = f00; c01 = f01; c10 = f10; c11 = f11; }
            }
        }
        if (row & (G / 16 - 1)) {
            pm_mul(m00, m01, m10, m11, c00, c01, c10, c11);
            pm_renorm(m00, m01, m10, m11);
        }
    }
    const double uin = wave_up1(m01), vin = wave_up1(m11);
    double cprev = (gl == 0) ? 0.0 : uin / vin;
    const int base = gl * R;
    double A = 1.0, Bq = 1.0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double a = F.al[k], b = F.iv[k], c = F.g[k];
        const double den = b + a * cprev;
        if (!(fabs(den) > 1.0e-300)) bad |= ERR_PIVOT;
        const double inv = d_div_normal(1.0, den);
        const double cp = (base + k == n - 1) ? 0.0 : -c * inv;
        F.iv[k] = inv;
        // the group's first row has no predecessor: its multiplier is 0, so
        // the value a substep shifts in from the lane before (another
        // tracer's lanes, or 0) drops out without a select
        F.al[k] = (k == 0 && gl == 0) ? 0.0 : a * inv;
        if constexpr (!NOSRC) F.vs[k] = inv * sv[k];
        F.g[k] = -cp;
        A *= F.al[k];
        Bq *= F.g[k];
        cprev = cp;
    }
    // Scan multipliers of the substep solve.  Hillis-Steele levels run
    // inside 16-lane rows (DPP); a group spanning several rows completes the
    // scan with per-row carries (see apply_solve).
    scan_coeffs<G, 0>(gl, rl, A, Bq, F);
    // two-row groups: the carry multiplier is used in rows 1 and 3 only
    F.PA = (G == 32 && !((threadIdx.x >> 4) & 1)) ? 0.0 : A;
    F.PQ = Bq;
    if constexpr (G == 32) {
        const int row = (int)((threadIdx.x & 63) >> 4);
        F.PQ0 = (row == 0) ? Bq : 0.0;
        F.PQ2 = (row == 2) ? Bq : 0.0;
    }
    if constexpr (G == 64) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            F.Tf[r] = rdlane64(A, 16 * r + 15);
            F.Tb[r] = rdlane64(Bq, 16 * r);
        }
    }
}

// In-row Hillis-Steele levels of the forward (prefix) and backward (suffix)
// affine scans: B_l += mult_l * B_{l -/+ s}.
template <int G, int ST, int LV>
__device__ __forceinline__ void scan_fwd(const double *Af, double &B)
{
    if constexpr (ST < LV) {
        B = fma(Af[ST], row_up<(1 << ST)>(B), B);
        scan_fwd<G, ST + 1, LV>(Af, B);
    }
}

template <int G, int ST, int LV>
__device__ __forceinline__ void scan_bwd(const double *Qb, double &B)
{
    if constexpr (ST < LV) {
        B = fma(Qb[ST], row_dn<(1 << ST)>(B), B);
        scan_bwd<G, ST + 1, LV>(Qb, B);
    }
}

// One implicit substep: x <- M^{-1} (x + src) for the group, in place.
// NOSRC: a run without sources (src = 0; vs is not formed).
template <int R, int G, bool NOSRC = false>
__device__ __forceinline__ void apply_solve(int gl, const Factor<R, G> &F, double (&x)[R])
{
    double t[R];
    // the lane-local prefixes start from the first row itself (al * 0 + t is
    // t exactly for the finite multipliers of a factorised run)
    t[0] = NOSRC ? F.iv[0] * x[0] : fma(F.iv[0], x[0], F.vs[0]);
    double L = t[0];
#pragma unroll
    for (int k = 1; k < R; ++k) {
        t[k] = NOSRC ? F.iv[k] * x[k] : fma(F.iv[k], x[k], F.vs[k]);
        L = fma(F.al[k], L, t[k]);
    }
    scan_fwd<G, 0, Factor<R, G>::LV>(F.Af, L);
    if constexpr (G == 32) {
        // rows 1 and 3 continue the groups begun in rows 0 and 2: they take
        // the inclusive value at the last lane of the row before (PA is 0 in
        // rows 0 and 2)
        L = fma(F.PA, bcast15(L), L);
    } else if constexpr (G == 64) {
        // carries through the four rows of the group: the running value
        // entering row r is T_{r-1} + Tf_{r-1} C_{r-1} (T: inclusive in-row
        // value at the row's last lane)
        const double T0 = rdlane64(L, 15), T1 = rdlane64(L, 31), T2 = rdlane64(L, 47);
        const double C1 = T0;
        const double C2 = fma(F.Tf[1], C1, T1);
        const double C3 = fma(F.Tf[2], C2, T2);
        const int row = (int)((threadIdx.x & 63) >> 4);
        const double C = (row == 1) ? C1 : (row == 2) ? C2 : (row == 3) ? C3 : 0.0;
        L = fma(F.PA, C, L);
    }
    // the first lane of the group multiplies the shifted-in value by al[0] = 0
    double d = wave_up1(L);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        d = fma(F.al[k], d, t[k]);
        t[k] = d;
    }
    double M = t[R - 1];
#pragma unroll
    for (int k = R - 2; k >= 0; --k) M = fma(F.g[k], M, t[k]);
    scan_bwd<G, 0, Factor<R, G>::LV>(F.Qb, M);
    if constexpr (G == 32) {
        // rows 0 and 2 take the value at the first lane of the row after:
        // the wave-uniform values at lanes 16 and 48 times multipliers that
        // are zero outside rows 0 resp. 2
        const double U1 = rdlane64(M, 16), U3 = rdlane64(M, 48);
        M = fma(F.PQ0, U1, fma(F.PQ2, U3, M));
    } else if constexpr (G == 64) {
        const double U1 = rdlane64(M, 16), U2 = rdlane64(M, 32), U3 = rdlane64(M, 48);
        const double D2 = U3;
        const double D1 = fma(F.Tb[2], D2, U2);
        const double D0 = fma(F.Tb[1], D1, U1);
        const int row = (int)((threadIdx.x & 63) >> 4);
        const double D = (row == 0) ? D0 : (row == 1) ? D1 : (row == 2) ? D2 : 0.0;
        M = fma(F.PQ, D, M);
    }
    // the group's last row has g = 0 (the system's last row), so the value
    // shifted in from the lane after drops out without a select
    double xn = wave_dn1(M);
#pragma unroll
    for (int k = R - 1; k >= 0; --k) {
        xn = fma(F.g[k], xn, t[k]);
        x[k] = xn;
    }
}

// ---------------------------------------------------------------------------
// Coefficient build of one species for one run
// ---------------------------------------------------------------------------

// Per-species view of the per-row grid tables of the build and the loss
// rates: the device tables (row j at j), or their copies in LDS (row j at
// j + (j >> ps), the padding of the rate rows).  ps is a literal at every
// construction, so the index arithmetic folds away.  T is the coefficient
// type: double, or float in the FP32-coefficient instances (whose tables are
// float copies in LDS).
template <typename T>
struct SpGridT {
    const T *x, *h, *inv_h, *inv_dx, *two_over_p, *p2, *expcut, *ib2;
    const T *lnp;                     // electrons only
    const T *expE, *sig, *Ep, *i2mk;  // protons only
    int ps;                           // < 0: device tables
    __device__ __forceinline__ int ix(int j) const { return ps < 0 ? j : j + (j >> ps); }
};
using SpGrid = SpGridT<double>;

// Tables of one species, in the order of their LDS copies: the common set,
// then ln p (electrons, GT_NE tables) or the proton factors (GT_NP tables).
enum { GT_X, GT_H, GT_INVH, GT_INVDX, GT_2OP, GT_P2, GT_EXPCUT, GT_IB2, GT_NC,
       GT_LNP = GT_NC, GT_NE,
       GT_EXPE = GT_NC, GT_SIG, GT_EP, GT_I2MK, GT_NP };

__device__ __forceinline__ const double *sp_table(const CkGrid &g, bool elec, int t)
{
    switch (t) {
    case GT_X: return elec ? g.ep : g.pp;
    case GT_H: return elec ? g.e_h : g.p_h;
    case GT_INVH: return elec ? g.e_inv_h : g.p_inv_h;
    case GT_INVDX: return elec ? g.e_inv_dx : g.p_inv_dx;
    case GT_2OP: return elec ? g.e_2op : g.p_2op;
    case GT_P2: return elec ? g.e_p2 : g.p_p2;
    case GT_EXPCUT: return elec ? g.e_expcut : g.p_expcut;
    case GT_IB2: return elec ? g.e_ib2 : g.p_ib2;
    case GT_EXPE: return elec ? g.e_lnp : g.p_expE;
    case GT_SIG: return g.p_sig;
    case GT_EP: return g.p_Ep;
    default: return g.p_i2mk;
    }
}

// tab: the device tables (ps < 0, T = double only) or the species' LDS
// copies, np elements apart
template <typename T>
__device__ __forceinline__ SpGridT<T> sp_grid(const CkGrid &g, bool elec, const T *tab, int np, int ps)
{
    SpGridT<T> s;
    const T *t[GT_NP];
#pragma unroll
    for (int i = 0; i < GT_NP; ++i) {
        if (i >= (elec ? GT_NE : GT_NP)) t[i] = nullptr;
        else if constexpr (sizeof(T) == sizeof(double)) t[i] = (ps < 0) ? sp_table(g, elec, i) : tab + (size_t)i * np;
        else t[i] = tab + (size_t)i * np;
    }
    s.x = t[GT_X]; s.h = t[GT_H]; s.inv_h = t[GT_INVH]; s.inv_dx = t[GT_INVDX]; s.two_over_p = t[GT_2OP];
    s.p2 = t[GT_P2]; s.expcut = t[GT_EXPCUT]; s.ib2 = t[GT_IB2];
    s.lnp = elec ? t[GT_LNP] : nullptr;
    s.expE = elec ? nullptr : t[GT_EXPE];
    s.sig = elec ? nullptr : t[GT_SIG];
    s.Ep = elec ? nullptr : t[GT_EP];
    s.i2mk = elec ? nullptr : t[GT_I2MK];
    s.ps = ps;
    return s;
}

// Face between rows j and j+1 (cc_prepare_face_weight + cc_eval_face_weights,
// Chang_Cooper.c:22-66): w = (B/C)_face * h, W+ = B(-w), W- = B(w); both zero
// when |C_j| <= 1e-200.  Fh = C_face / h.  The Bernoulli pair is formed from
// one exponential h2 = exp(-|w|/2):
//   |w| <  0.1: W_big = 1 / (h2 * den), den = 1 + w^2/24 + w^4/1920
//   |w| >= 0.1: W_big = |w| / (1 - h2^2)
//   W_small = W_big * h2^2; W+ = W_big for w > 0 (else W_small).
struct Face {
    double bm, Fh, wp, wm, w;
};

// FP32 coefficients (T = float): the face averages, diffusion flux factor
// and Peclet number w come from float rows; the Bernoulli weights of w are
// still evaluated in FP64 (the stability-critical part of the weights).
template <typename T>
__device__ __forceinline__ Face d_face2(T Bl, T Br, T Cl, T Cr, T h, T inv_h)
{
#pragma clang fp contract(off)
    Face f;
    double wv;
    bool live;
    if constexpr (sizeof(T) == sizeof(double)) {
        f.bm = 0.5 * (Bl + Br);
        const double cm = 0.5 * (Cl + Cr);
        f.Fh = cm * inv_h;
        live = fabs(Cl) > 1.0e-200;
        wv = live ? d_div_normal(f.bm, cm) * h : 0.0;
    } else {
        const float bm = 0.5f * (Bl + Br);
        const float cm = 0.5f * (Cl + Cr);
        f.bm = (double)bm;
        f.Fh = (double)(cm * inv_h);
        live = fabs((double)Cl) > 1.0e-200;
        wv = live ? (double)(bm * __builtin_amdgcn_rcpf(cm) * h) : 0.0;
    }
    f.wp = 0.0;
    f.wm = 0.0;
    f.w = 0.0;
    if (live) {
        const double w = wv;
        const double aw = fabs(w);
        const double h2 = d_exp_nonpos(-0.5 * aw);
        const double E = h2 * h2;
        const bool small = aw < 0.1;
        const double w2 = w * w;
        const double den = small ? h2 * (1.0 + w2 * (1.0 / 24.0) + (w2 * w2) * (1.0 / 1920.0)) : 1.0 - E;
        const double big = d_div_normal(small ? 1.0 : aw, den);
        const double sm = big * E;
        f.wp = (w > 0.0) ? big : sm;
        f.wm = (w > 0.0) ? sm : big;
        f.w = w;
    }
    return f;
}

// Row j of the Chang-Cooper system (fp_shared_core.c:660-704 ->
// Chang_Cooper.c:68-184), A = 1, from its two faces.  kind: 0 interior, 1 left
// boundary, 2 right boundary.  Output as positive coupling magnitudes
// b x_j - a x_{j-1} - c x_{j+1}.
__device__ __forceinline__ void d_cc_row2(int kind, double t1, double C, double dtT, const Face &fl,
                                          const Face &fr, double Cl, double &a, double &b, double &c)
{
#pragma clang fp contract(off)
    const bool cz = !(fabs(C) > 1.0e-200);
    if (kind == 0) {
        if (!cz) {
            a = t1 * fl.Fh * fl.wm;
            b = 1.0 + t1 * (fl.Fh * fl.wp + fr.Fh * fr.wm) + dtT;
            c = t1 * fr.Fh * fr.wp;
        } else {
            a = 0.0;
            b = 1.0 + t1 * fl.bm + dtT;
            c = t1 * fr.bm;
        }
    } else if (kind == 1) {
        a = 0.0;
        if (!cz) {
            b = 1.0 + t1 * (fr.Fh * fr.wm) + dtT;
            c = t1 * fr.Fh * fr.wp;
        } else {
            b = 1.0 + dtT;
            c = t1 * fr.bm;
        }
    } else {
        c = 0.0;
        if (!cz) {
            a = t1 * fl.Fh * fl.wm;
            b = 1.0 + t1 * (fl.Fh * fl.wp) + dtT;
        } else {
            // delta of the left face (Chang_Cooper.c:143-184): nonzero only
            // when that face carries weights
            const double delta = (fabs(Cl) > 1.0e-200) ? d_face_delta(fl.w) : 0.0;
            a = -delta * fl.bm * t1;
            b = 1.0 + t1 * fl.bm + dtT;
        }
    }
}

// Run-level coefficient inputs: D(p) = p^2 * Dshape(p) * dfac (+ eta scale),
// B = radp - (2/p) D, C = D; dt/T from dtT_fac * sig (protons).
struct RunCoef {
    double dt;
    double dfac;     // inv_on * scale (model) or scale / 4e5 (floor)
    bool model;      // filled form (p^2 inv exp(-E/Epmax) exp(-1/p)) vs floor p^2/4e5
    double inv_epmax;
    double dtT_fac;  // dt * 0.5 c n_gas Gyr (protons), 0 for electrons
};

template <bool ELEC, typename T>
__device__ __forceinline__ void row_BC(const SpGridT<T> &sg, const RunCoef &rc, int j, T radp, T &B, T &C)
{
#pragma clang fp contract(off)
    const int i = sg.ix(j);
    const T dfac = (T)rc.dfac;
    T D;
    if (!rc.model) D = sg.p2[i] * dfac;
    else if (ELEC) D = sg.p2[i] * dfac * sg.expcut[i];
    else D = sg.p2[i] * dfac * sg.expE[i] * sg.expcut[i];
    C = D;
    B = radp - sg.two_over_p[i] * D;
}

// lane l + 1 / l - 1 across the wave for the row inputs of the build (one
// 32-bit DPP move for float rows, two for double)
template <typename T>
__device__ __forceinline__ T wave_dn1_t(T v)
{
    if constexpr (sizeof(T) == sizeof(double)) return wave_dn1(v);
    else return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
template <typename T>
__device__ __forceinline__ T wave_up1_t(T v)
{
    if constexpr (sizeof(T) == sizeof(double)) return wave_up1(v);
    else return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}

// Builds the rows of the lane into F (al = a, iv = b, g = c) and factorises.
// The rows' coefficients (B, C and the face inputs) are of type T; the
// face weights, the rows a, b, c and the factorisation are FP64.
template <bool ELEC, int R, int G, typename T, bool NOSRC = false>
__device__ __forceinline__ void build_factor(int gl, int n, const SpGridT<T> &sg, const RunCoef &rc,
                                             const T *radp, const SrcRows &src, Factor<R, G> &F, int &bad)
{
#pragma clang fp contract(off)
    const int base = gl * R;
    const int o = run_opaque(0);
    T B0, C0, BL, CL;
    row_BC<ELEC, T>(sg, rc, base, radp[o], B0, C0);
    row_BC<ELEC, T>(sg, rc, base + R - 1, radp[o + R - 1], BL, CL);
    // neighbour lanes' edge rows (DPP across the wave; a group's edge lanes
    // read the next group's values and do not use them)
    const T Bn = wave_dn1_t(B0), Cn = wave_dn1_t(C0);
    const T Bp = wave_up1_t(BL), Cp = wave_up1_t(CL);
    Face fl = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (gl > 0) fl = d_face2<T>(Bp, B0, Cp, C0, sg.h[sg.ix(base - 1)], sg.inv_h[sg.ix(base - 1)]);
    double Cl = (double)Cp;
    T Bk = B0, Ck = C0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int j = base + k;
        const bool last = (j == n - 1);
        T Br, Cr;
        if (k == R - 1) { Br = Bn; Cr = Cn; }
        else if (k == R - 2) { Br = BL; Cr = CL; }
        else row_BC<ELEC, T>(sg, rc, j + 1, radp[o + k + 1], Br, Cr);
        Face fr = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (!last) fr = d_face2<T>(Bk, Br, Ck, Cr, sg.h[sg.ix(j)], sg.inv_h[sg.ix(j)]);
        const int kind = (j == 0) ? 1 : (last ? 2 : 0);
        const double t1 = rc.dt * (double)sg.inv_dx[sg.ix(j)];
        double dtT = 0.0;
        if constexpr (!ELEC) dtT = (double)((T)rc.dtT_fac * sg.sig[sg.ix(j)]);
        double a, b, c;
        d_cc_row2(kind, t1, (double)Ck, dtT, fl, fr, Cl, a, b, c);
        F.al[k] = a;
        F.iv[k] = b;
        F.g[k] = c;
        fl = fr;
        Cl = (double)Ck;
        Bk = Br;
        Ck = Cr;
    }
    // the run's sources after the rows (absent rows point at the grid's zero
    // row: no branches), so they hold no registers during the row build;
    // their loads overlap the factorisation's pivot scan
    double sv[R];
#pragma unroll
    for (int k = 0; k < R; ++k) sv[k] = NOSRC ? 0.0 : src.scale * (src.inj[o + k] * src.ka * src.kb + src.q[o + k]);
    factorize<R, G, NOSRC>(gl, n, sv, F, bad);
}

// ---------------------------------------------------------------------------
// Fused persistent kernel
// ---------------------------------------------------------------------------
struct SegScalars {
    double n_gas, kbt, bfield, divv, lturb, dv, cs, beta;
};

// Electron loss rates of the lane's rows for one segment
// (prepare_cre_losses_1d, fp_shared_core.c:780-842).
template <int R, typename T>
__device__ __forceinline__ void seg_radp_e(const SpGridT<T> &sg, const CkSlot &S, int base, int nseg, double alpha,
                                           const SegScalars &q, T *radp)
{
#pragma clang fp contract(off)
    {
        const double b_micro = q.bfield * 1.0e+6;
        const double b_ratio = b_micro / 3.2;
        const T synch_coeff = (T)(4.8e-4 * b_ratio * b_ratio * 9.11e-28 * 3.0e+10 * 3.1536e+16);
        const T coul_coeff = (T)(3.05e-29 * q.n_gas * K_C * K_GYR / (K_ME * K_GEV));
        const T adia = (T)(1.0 / 3.0 * q.divv);
        // ln(p / n_gas) as ln p (table) + ln(1 / n_gas), one log per segment
        const T lning = (T)log(1.0 / q.n_gas);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int j = base + k;
            const T pe = sg.x[sg.ix(j)];
            const T pe2 = sg.p2[sg.ix(j)];
            // both rows loaded unconditionally (ric_n is always a valid row)
            // alpha is 0 without interpolation; ric_n is always a valid row
            const T ric = (T)d_lerp(S.ric_c[j], S.ric_n[j], alpha);
            const T coul = coul_coeff * ((T)1.0 + (sg.lnp[sg.ix(j)] + lning) * (T)(1.0 / 74.8)) * sg.ib2[sg.ix(j)];
            radp[k] = synch_coeff * pe2 + ric + coul + adia * pe;
        }
    }
}

// Proton loss rates (prepare_crp_losses_1d_with_threshold,
// fp_shared_core.c:729-778) evaluated in the strided row order m G + gl and
// left in LDS, where the lane owning each row reads it during the builds.  The thermal (erf)
// branch of the Coulomb kernel is taken only by the lowest bins; in the
// block order the lane holding them would make the whole wave take the
// branch for each of its R rows, in the strided order the wave takes it
// ceil(rows / G) times.  s_tr holds N + N / R doubles (row j at j + j / R:
// one pad slot per block of R rows against LDS bank conflicts on the reads).
template <int N, int R, int G, typename T>
__device__ __forceinline__ void seg_radp_p(const SpGridT<T> &sg, int gl, const SegScalars &q, T *s_tr)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int m = 0; m < R; ++m) {
        const int j = m * G + gl;
        const int i = sg.ix(j);
        if constexpr (sizeof(T) == sizeof(double)) {
            const double adia = 1.0 / 3.0 * q.divv;
            const double ikbt = 1.0 / q.kbt;
            s_tr[j + j / R] =
                d_b_coulomb_p_tab(q.n_gas, q.kbt, ikbt, sg.Ep[i], sg.i2mk[i], sg.ib2[i]) + adia * sg.x[i];
        } else {
            const float adia = (float)(1.0 / 3.0 * q.divv);
            s_tr[j + j / R] = d_b_coulomb_p_tabf((float)q.n_gas, (float)q.kbt, (float)(1.0 / q.kbt), sg.Ep[i],
                                                 sg.i2mk[i], sg.ib2[i]) +
                              adia * sg.x[i];
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The lane's R spectrum values move to LDS (own slots, stride R + 1) while a
// build runs and come back before its substeps.
// own null: the kernel instance keeps the spectrum in registers throughout.
template <int R>
__device__ __forceinline__ void park_rows(double *own, const double (&x)[R])
{
    if (!own) return;
#pragma unroll
    for (int k = 0; k < R; ++k) own[k] = x[k];
}
template <int R>
__device__ __forceinline__ void unpark_rows(const double *own, double (&x)[R])
{
    if (!own) return;
    const int o = run_opaque(0);
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = own[o + k];
}

// All runs of one species over one segment: up to three runs of constant
// coefficients (before / inside / after the on-window, tracer_fp_cpu.c:150-157).
// On entry the spectrum is parked (park_rows); on exit it is in x.
template <bool ELEC, int N, int R, int G, typename T, bool NOSRC = false>
__device__ __forceinline__ void species_segment(int gl, const SpGridT<T> &sg, int steps, bool windowed, int ls, int le,
                                                RunCoef rc, double dfac_on, double dfac_off, bool mode_on,
                                                const T *radp, const SrcRows &src, double *xpark,
                                                double (&x)[R], int &bad)
{
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
        rc.model = on && mode_on;
        rc.dfac = on ? dfac_on : dfac_off;
        if (s0 > 0) park_rows<R>(xpark, x);
        Factor<R, G> F;
        build_factor<ELEC, R, G, T, NOSRC>(gl, N, sg, rc, radp, src, F, bad);
        unpark_rows<R>(xpark, x);
        // substeps two per loop trip (one loop branch per pair)
        int s = s0;
        for (; s + 1 < s1; s += 2) {
            apply_solve<R, G, NOSRC>(gl, F, x);
            apply_solve<R, G, NOSRC>(gl, F, x);
        }
        if (s < s1) apply_solve<R, G, NOSRC>(gl, F, x);
        s0 = s1;
    }
}

// One snapshot of one tracer (tracer_fp_batch_execute_bucket,
// tracer_fp_batch.c:651-827 -> tracer_fp_cpu_evolve, tracer_fp_cpu.c:88-196).
// Within a segment the proton and electron updates are independent (the
// electron source uses the segment-start protons), so the species run one
// after the other and only one factor set is live.
// Launch modes of the fused kernel.
enum { CK_MODE_E = 0,      // electron_only: protons are exactly zero
       CK_MODE_EP = 1,     // both species, secondaries contracted in the kernel
       CK_MODE_P_REC = 2,  // protons only; records the segment-start proton rows
       CK_MODE_E_SEC = 3,  // electrons only; secondaries from the contracted rows
       CK_MODE_F32 = CK_LAUNCH_F32,  // added to a mode: FP32 coefficient evaluation (losses, D_pp, B, C, face inputs)
       CK_MODE_E_NOSRC = CK_LAUNCH_NOSRC  // electron_only with no source rows in the launch (src = 0)
};

// LDS doubles per lane group of the fused kernel: loss rates (padded), the
// parked spectrum (PK), in mode EP the contraction's protons / electron
// sources, and the tracer's backgrounds at two snapshots.
__host__ __device__ constexpr int fused_lds_doubles(int N, int R, int MODE, bool PK)
{
    return N + N / R + 1 + (PK ? (N / R) * (R + 1) : 0) + (MODE == CK_MODE_EP ? N : 0) + 2 * CK_NBG;
}

// Grid tables copied to LDS per block (electron set, proton set or both).
__host__ __device__ constexpr int fused_grid_tables(int MODE)
{
    return MODE == CK_MODE_EP ? GT_NE + GT_NP : MODE == CK_MODE_P_REC ? GT_NP : GT_NE;
}

// Whether the blocks that the register budget lets share a CU (OCC waves per
// SIMD, WPB waves per block) fit its 160 KiB of LDS with the grid tables (LG,
// TB bytes per entry: 8 for FP64 coefficients, 4 for FP32) and the parked
// spectra (PK).
__host__ __device__ constexpr bool fused_lds_fits(int N, int R, int WPB, int MODE, int OCC, bool LG, bool PK,
                                                  int TB = 8)
{
    return ((size_t)8 * WPB * (64 / (N / R)) * fused_lds_doubles(N, R, MODE, PK) +
            (LG ? (size_t)TB * fused_grid_tables(MODE) * (N + N / R + 1) : 0)) *
               (OCC * 4 / WPB > 0 ? OCC * 4 / WPB : 1) <=
           (size_t)160 * 1024;
}
// LDS use of an instance: the grid tables take precedence over parking (the
// tables cut the build's load latency, parking its register pressure).
// FP32-coefficient instances exist only where their float tables fit.
__host__ __device__ constexpr bool fused_use_lg(int N, int R, int WPB, int MODE, int OCC, int TB = 8)
{
    return fused_lds_fits(N, R, WPB, MODE, OCC, true, false, TB);
}
__host__ __device__ constexpr bool fused_use_pk(int N, int R, int WPB, int MODE, int OCC, int TB = 8)
{
    return fused_lds_fits(N, R, WPB, MODE, OCC, fused_use_lg(N, R, WPB, MODE, OCC, TB), true, TB);
}

template <int N, int R, int MODE, bool PK, typename T>
__device__ __forceinline__ void snapshot_update(const CkGrid &g, const SpGridT<T> &sge, const SpGridT<T> &sgp,
                                                const CkSlot &S, int t, int gl, double *s_grp, double (&xe)[R],
                                                double (&xp)[R], int &bad)
{
    // The group's LDS region (fused_lds_doubles): the loss rates of the
    // segment (N + N / R + 1 doubles, row j at j + j / R), the parked spectrum
    // (G (R + 1) doubles, the lane's rows at gl (R + 1)), and in mode EP N
    // doubles for the segment-start protons of the contraction, then the
    // electron sources.  Rates and sources are read once per run (sources
    // without additions straight from the tracer's rows), so they take no
    // registers during the substeps, and the spectrum takes none during builds.
    constexpr int G = N / R;
    constexpr bool PROT = (MODE == CK_MODE_EP || MODE == CK_MODE_P_REC);
    constexpr bool ELEC = (MODE != CK_MODE_P_REC);
    const int base = gl * R;
    const CkParams &prm = S.prm;
    const int nsub = S.nsub[t];
    const int non = S.n_on[t];
    const int nseg =
        d_choose_segments(nsub, S.allow_interp, prm.coeff_interp_min_steps, prm.coeff_interp_max_segments);
    const int on_start = (nsub - non) / 2;
    const int on_end = on_start + non;
    const double dt = S.dt_snap / (double)nsub;
    const size_t roff = (size_t)t * N + base;
    const double mass = S.mass ? S.mass[t] : 0.0;
    const bool dis = S.dis ? (S.dis[t] != 0) : false;
    const bool sec = (MODE == CK_MODE_EP) && S.fqe_t;
    const int row0 = (MODE == CK_MODE_P_REC || MODE == CK_MODE_E_SEC) ? S.seg_off[t] : 0;
    // loss rates in the coefficient type (float rates use the region's first half)
    T *const s_rad = (T *)s_grp;
    const T *const own_rad = (T *)s_grp + gl * (R + 1);
    double *const own_x = PK ? s_grp + (N + N / R + 1) + gl * (R + 1) : nullptr;
    double *const s_row = s_grp + (N + N / R + 1) + (PK ? G * (R + 1) : 0);
    double *const own_src = s_row + base;
    double *const s_bg = s_row + (MODE == CK_MODE_EP ? N : 0);
    // the tracer's backgrounds at this snapshot and the next, one round trip
    // for the group, then read from LDS by every segment
    for (int i = gl; i < 2 * CK_NBG; i += G) s_bg[i] = (i < CK_NBG ? S.bgc[i] : S.bgn[i - CK_NBG])[t];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    for (int iseg = 0; iseg < nseg; ++iseg) {
        const int fb = (int)(((long long)iseg * nsub) / nseg);
        const int fe = (int)(((long long)(iseg + 1) * nsub) / nseg);
        const int steps = fe - fb;
        if (steps <= 0) continue;
        const double alpha = (nseg > 1) ? ((double)iseg + 0.5) / (double)nseg : 0.0;
        SegScalars q;
        {
            // alpha is 0 (the current snapshot's value) without interpolation
            double v[CK_NBG];
#pragma unroll
            for (int f = 0; f < CK_NBG; ++f) v[f] = d_lerp(s_bg[f], s_bg[CK_NBG + f], alpha);
            q.n_gas = v[CK_BG_NGAS];
            q.kbt = v[CK_BG_KBT];
            q.bfield = v[CK_BG_B];
            q.divv = dis ? 0.0 : v[CK_BG_DIVV];
            q.lturb = v[CK_BG_LTURB];
            q.dv = v[CK_BG_DV];
            q.cs = v[CK_BG_CS];
            q.beta = v[CK_BG_BETA];
        }
        int sob = (on_start > fb) ? on_start : fb;
        int soe = (on_end < fe) ? on_end : fe;
        if (soe < sob) soe = sob;
        const bool windowed = (soe - sob) < steps;
        const int ls = sob - fb, le = ls + (soe - sob);
        double tacc_d = prm.t_acc_direct_gyr;
        {
            const double tmid = S.elapsed + ((double)fb + 0.5 * (double)steps) * dt;
            const bool direct_on = prm.dpp_model == 4 && isfinite(prm.t_acc_direct_gyr) &&
                                   prm.t_acc_direct_gyr > 0.0 && isfinite(prm.t_off_gyr) && prm.t_off_gyr > 0.0 &&
                                   isfinite(prm.t_acc_off_gyr) && prm.t_acc_off_gyr > 0.0;
            if (direct_on && tmid < prm.t_off_gyr) tacc_d = prm.t_acc_off_gyr;
        }
        double inv_on = 0.0;
        const bool mode_on = d_dpp_model(prm, prm.dpp_model, q.lturb, q.dv, q.cs, q.beta, tacc_d, &inv_on) != 0;
        double scale = 1.0;
        if (prm.dpp_model != 3 && prm.eta_dpp_cap > 0.0) {
            double part = 0.0;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int j = base + k;
                const double wp_ = g.p_ecr_w[j];
                const double we_ = g.e_ecr_w[j];
                if (PROT && wp_ > 0.0 && isfinite(xp[k])) part += wp_ * xp[k] * g.pdp[j];
                if (we_ > 0.0 && isfinite(xe[k])) part += we_ * xe[k] * g.edp[j];
            }
            const double ecr = g_sum<G>(part);
            scale = d_eta_scale(prm, prm.dpp_model, q.lturb, q.dv, q.cs, q.beta, mass, tacc_d, ecr);
        }
        RunCoef rc;
        rc.dt = dt;
        rc.model = false;
        rc.dfac = 0.0;
        rc.inv_epmax = 1.0 / prm.epmax;
        rc.dtT_fac = 0.0;
        // D prefactors: on-window uses the model form times the eta scale (or
        // the floor times the scale), outside the floor p^2/4e5
        const double dfac_on = mode_on ? inv_on * scale : (1.0 / (4.0 * 1.0e+5)) * scale;
        const double dfac_off = 1.0 / (4.0 * 1.0e+5);

        if (PROT) {
            // segment-start protons feed the electron secondaries
            if (sec) {
#pragma unroll
                for (int k = 0; k < R; ++k) s_row[base + k] = xp[k];
            }
            if (MODE == CK_MODE_P_REC) {
                double *rec = S.seg_rec + (size_t)(row0 + iseg) * N + base;
#pragma unroll
                for (int k = 0; k < R; ++k) rec[k] = xp[k];
            }
            park_rows<R>(own_x, xp);
            seg_radp_p<N, R, G, T>(sgp, gl, q, s_rad);
            RunCoef rp = rc;
            rp.dtT_fac = dt * (0.5 * K_C * q.n_gas) * K_GYR;
            const SrcRows sp = {S.qpi ? S.qpi + roff : g.zeros + base, g.zeros + base, 0.0, 0.0, dt};
            species_segment<false, N, R, G, T>(gl, sgp, steps, windowed, ls, le, rp, dfac_on, dfac_off, mode_on, own_rad,
                                            sp, own_x, xp, bad);
        }
        if (ELEC && MODE == CK_MODE_EP) {
            double src[R];
#pragma unroll
            for (int k = 0; k < R; ++k) src[k] = S.qepri ? S.qepri[roff + k] : 0.0;
            if (sec) {
                // inje = Gyr n_gas sum_p N_p fqe[p, e] + qepri (tracer_fp_cpu.c:142-147)
                double acc[R];
#pragma unroll
                for (int k = 0; k < R; ++k) acc[k] = 0.0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (int jp = S.jp0; jp < N; ++jp) {
                    const double nn = s_row[jp];
                    const double *wrow = S.fqe_t + (size_t)jp * N + base;
#pragma unroll
                    for (int k = 0; k < R; ++k) acc[k] += nn * wrow[k];
                }
                // every lane of the group has read the protons before the
                // sources overwrite them
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int k = 0; k < R; ++k) src[k] = acc[k] * K_GYR * q.n_gas + src[k];
            }
#pragma unroll
            for (int k = 0; k < R; ++k) own_src[k] = dt * src[k];
            park_rows<R>(own_x, xe);
            seg_radp_e<R, T>(sge, S, base, nseg, alpha, q, (T *)s_grp + gl * (R + 1));
            const SrcRows se = {own_src, g.zeros + base, 0.0, 0.0, 1.0};
            species_segment<true, N, R, G, T>(gl, sge, steps, windowed, ls, le, rc, dfac_on, dfac_off, mode_on, own_rad,
                                           se, own_x, xe, bad);
        } else if (ELEC) {
            // mode E_SEC: the same source as the in-kernel contraction, with
            // the proton contraction done for all tracers at once outside the
            // kernel (seg_inje rows)
            const double *inj =
                (MODE == CK_MODE_E_SEC) ? S.seg_inje + (size_t)(row0 + iseg) * N + base : g.zeros + base;
            park_rows<R>(own_x, xe);
            seg_radp_e<R, T>(sge, S, base, nseg, alpha, q, (T *)s_grp + gl * (R + 1));
            const SrcRows se = {S.qepri ? S.qepri + roff : g.zeros + base, inj, K_GYR, q.n_gas, dt};
            species_segment<true, N, R, G, T, MODE == CK_MODE_E_NOSRC>(gl, sge, steps, windowed, ls, le, rc, dfac_on,
                                                                      dfac_off, mode_on, own_rad, se, own_x, xe, bad);
        }
    }
}

template <int N, int R, int WPB, int MODE, bool LG, bool PK, typename T>
__device__ __forceinline__ void k_fused_body(const CkGrid &g, const CkSlot *__restrict__ slots,
                                             const CkTask *__restrict__ tasks, int ntasks,
                                             int *__restrict__ counter, int *__restrict__ err)
{
    constexpr bool PROT = (MODE == CK_MODE_EP || MODE == CK_MODE_P_REC);
    constexpr bool ELEC = (MODE != CK_MODE_P_REC);
    constexpr int G = N / R;
    constexpr int GPW = 64 / G;
    __shared__ double s_lds[WPB][GPW][fused_lds_doubles(N, R, MODE, PK)];
    constexpr int NP = N + N / R + 1;
    constexpr int PS = LogG<R>::v;
    static_assert(LG || sizeof(T) == sizeof(double), "FP32-coefficient instances need their tables in LDS");
    __shared__ T s_tab[LG ? fused_grid_tables(MODE) : 1][LG ? NP : 1];
    if constexpr (LG) {
        // the launch's grid tables, row j at j + j / R: electron set first
        constexpr int NT = fused_grid_tables(MODE);
        for (int i = threadIdx.x; i < NT * N; i += WPB * 64) {
            const int tb = i / N, j = i - tb * N;
            const bool el = ELEC && tb < GT_NE;
            const int ti = el ? tb : (ELEC ? tb - GT_NE : tb);
            s_tab[tb][j + (j >> PS)] = (T)sp_table(g, el, ti)[j];
        }
        __syncthreads();
    }
    const SpGridT<T> sge =
        LG ? sp_grid<T>(g, true, &s_tab[0][0], NP, PS) : sp_grid<T>(g, true, nullptr, 0, -1);
    const SpGridT<T> sgp = LG ? sp_grid<T>(g, false, &s_tab[0][0] + ((LG && ELEC) ? GT_NE * NP : 0), NP, PS)
                              : sp_grid<T>(g, false, nullptr, 0, -1);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int grp = lane / G;
    const int gl = lane % G;
    const int base = gl * R;
    double *s_grp = &s_lds[wave][grp][0];
    int bad = 0;

    for (;;) {
        // wave-uniform grab of GPW consecutive tasks (cost-ordered list)
        int tbase = 0;
        if (lane == 0) tbase = atomicAdd(counter, GPW);
        tbase = __shfl(tbase, 0, 64);
        if (tbase >= ntasks) break;
        const int tid = tbase + grp;
        // groups past the end of the list repeat the wave's first task (the
        // same data through the same instructions) and store nothing: the
        // solve's cross-lane reads multiply a neighbouring group's values by
        // zero, so every lane must hold finite values of a real tracer, not
        // registers that were never written in this launch
        const bool idle = tid >= ntasks;
        {
            const CkTask task = tasks[idle ? tbase : tid];
            const CkSlot &S0 = slots[task.slot];
            const int t = task.tracer;
            const size_t roff = (size_t)t * N + base;
            double xe[R], xp[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                xe[k] = ELEC ? S0.cre[roff + k] : 0.0;
                xp[k] = PROT ? S0.crp[roff + k] : 0.0;
            }
            // a task advances one tracer over nslot consecutive snapshots
            for (int sl = task.slot; sl < task.slot + task.nslot; ++sl) {
                const CkSlot &S = slots[sl];
                snapshot_update<N, R, MODE, PK, T>(g, sge, sgp, S, t, gl, s_grp, xe, xp, bad);
                if (ELEC && S.out_e && !idle) {
#pragma unroll
                    for (int k = 0; k < R; ++k) S.out_e[roff + k] = xe[k];
                }
                if (PROT && S.out_p && !idle) {
#pragma unroll
                    for (int k = 0; k < R; ++k) S.out_p[roff + k] = xp[k];
                }
            }
            if (!idle) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    if (!isfinite(xe[k]) || !isfinite(xp[k])) bad |= ERR_NONFINITE;
                    if (ELEC) S0.cre[roff + k] = xe[k];
                    if (PROT) S0.crp[roff + k] = xp[k];
                }
            }
        }
    }
    if (bad) atomicOr(err, bad);
}

// The kernel: WPB waves per block, register budget set for OCC waves per SIMD
// (the solve is latency-bound, so the budget trades unrolled state against
// waves available to hide cross-lane and memory latency).  T: the
// coefficient type (double; float for the FP32-coefficient instances).
template <int N, int R, int WPB, int MODE, int OCC, typename T = double>
__global__ __launch_bounds__(WPB * 64) __attribute__((amdgpu_waves_per_eu(OCC, 8))) void k_fused(
    CkGrid g, const CkSlot *__restrict__ slots, const CkTask *__restrict__ tasks, int ntasks,
    int *__restrict__ counter, int *__restrict__ err)
{
    constexpr int TB = (int)sizeof(T);
    k_fused_body<N, R, WPB, MODE, fused_use_lg(N, R, WPB, MODE, OCC, TB), fused_use_pk(N, R, WPB, MODE, OCC, TB),
                 T>(g, slots, tasks, ntasks, counter, err);
}
