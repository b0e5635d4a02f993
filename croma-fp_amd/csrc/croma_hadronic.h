/*
 * croma_hadronic.h -- the emission and secondary-source kernels' physics,
 * compiled both by the host C runtime (gcc, -ffp-contract=off) and by the
 * device tables (hipcc, contraction off inside every function), so the host
 * and device tables follow one arithmetic:
 *   pp cross section (Kamae 2006 fit, Blattnig 2000 charged ratios)  HADRONIC.c:342-431
 *   pion spectra, muon decay, Qe kernel                             HADRONIC.c:433-533
 *   gamma kernel                                                    HADRONIC.c:139-174
 *   electron / muon neutrino kernels                                HADRONIC.c:179-322
 *   IC scattering kernel, CMB photon density                        FP_Coef.c:378-405
 *   IC emissivity kernel (N = 400 trapezoid)                        FP_Coef.c:553-582
 */
#ifndef CROMA_HADRONIC_H
#define CROMA_HADRONIC_H

#include <math.h>

#if defined(__HIPCC__)
#define CHK_FN __host__ __device__ static inline
#else
#define CHK_FN static inline
#endif
#if defined(__clang__)
#define CHK_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define CHK_NOCONTRACT
#endif

/* constants (CONSTANTS.c, HADRONIC.c) */
#define CHK_PI 3.14159265358979323846
#define CHK_C 2.999e+10
#define CHK_H 6.63e-27
#define CHK_MP 0.938
#define CHK_ME 0.511e-3
#define CHK_MB 1.0e-27
#define CHK_MPI0 1.35e-1
#define CHK_MPIC 1.40e-1
#define CHK_MMU 1.05e-1

CHK_FN double chk_pp_pion_cross_section(int kind, double p)
{
    CHK_NOCONTRACT
    const double a0 = 0.1176, a1 = 0.3829, a2 = 23.10, a3 = 6.454, a4 = -5.764, a5 = -23.63,
                 a6 = 94.75, a7 = 0.02667, b0 = 11.34, b1 = 23.72, c0 = 28.5, c1 = -6.133,
                 c2 = 1.464, d0 = 0.3522, d1 = 0.1530, d2 = 1.498, d3 = 2.0, d4 = 30.0,
                 d5 = 3.155, d6 = 1.042, e0 = 5.922, e1 = 1.632, f0 = 0.0834, f1 = 9.5, f2 = 5.5,
                 f3 = 1.68, f4 = 3134.0, g0 = 0.0004257, g1 = 4.5, g2 = 7.0, g3 = 2.1,
                 g4 = 503.5;
    const double Ep = sqrt(CHK_MP * CHK_MP + p * p), Tp = Ep - CHK_MP;
    const double x = log10(p), x2 = x * x, x3 = x2 * x;
    double nd, df, dl, rs, tot, pi0, pip, pim;

    if (kind > 2 || kind < -2) return 0.0;
    if (p < 1.0) nd = 0.0;
    else if (p < 1.3) nd = 0.57 * pow(x / a0, 1.2) * (a2 + a3 * x2 + a4 * x3 + a5 * exp(-a6 * pow(x + a7, 2.0)));
    else if (p < 2.4) nd = (b0 * fabs(a1 - x) + b1 * fabs(a0 - x)) / (a1 - a0);
    else if (p < 10.0) nd = a2 + a3 * x2 + a4 * x3 + a5 * exp(-a6 * pow(x + a7, 2.0));
    else nd = c0 + c1 * x + c2 * x2;

    if (p < 2.25) df = 0.0;
    else if (p < 3.2) df = sqrt(fabs(x - d0) / d1) * (d2 + d3 * log10(d4 * (x - 0.25)) + d5 * x2 - d6 * x3);
    else if (p < 100.0) df = d2 + d3 * log10(d4 * (x - 0.25)) + d5 * x2 - d6 * x3;
    else df = e0 + e1 * x;

    if (Ep < 1.4) dl = 0.0;
    else if (Ep < 1.6) dl = f0 * pow(Ep, 10.0);
    else if (Ep < 1.8) dl = f1 * exp(-f2 * pow(Ep - f3, 2.0));
    else if (Ep < 10.0) dl = f4 * pow(Ep, -10.0);
    else dl = 0.0;

    if (Ep < 1.6) rs = 0.0;
    else if (Ep < 1.9) rs = g0 * pow(Ep, 14.0);
    else if (Ep < 2.3) rs = g1 * exp(-g2 * pow(Ep - g3, 2.0));
    else if (Ep < 20.0) rs = g4 * pow(Ep, -6.0);
    else rs = 0.0;

    tot = (nd + df + dl + rs) * CHK_MB;
    pi0 = 1.0 / (0.007 + 0.1 * log(Tp) / Tp + 0.3 / Tp / Tp);
    pip = 1.0 / (0.00717 + 0.0652 * log(Tp) / Tp + 0.162 / Tp / Tp);
    pim = 1.0 / (0.00456 + 0.0846 / pow(Tp, 0.5) + 0.577 / pow(Tp, 1.5));
    if (kind == 0) return tot;
    if (kind == 1) return tot * pip / pi0;
    if (kind == -1) return tot * pim / pi0;
    return tot * (1.0 + pip / pi0 + pim / pi0);
}


CHK_FN double chk_ic_kernel(double eps0, double gamma, double eps)
{
    CHK_NOCONTRACT
    const double re = 2.82e-13;
    double kappa = eps / (4.0 * eps0 * gamma * (gamma - eps)), C;
    if (kappa <= 0.0) return 0.0;
    C = 2.0 * kappa * log(kappa) + (1.0 + 2.0 * kappa) * (1.0 - kappa) +
        pow(4.0 * eps0 * kappa * gamma, 2.0) / (2.0 * (1.0 + 4.0 * eps0 * gamma * kappa)) * (1.0 - kappa);
    return 2.0 * CHK_PI * re * re * CHK_C / (gamma * gamma * eps0) * C;
}

CHK_FN double chk_cmb_n(double eps, double z)
{
    CHK_NOCONTRACT
    const double T0 = 2.757, meg = 9.11e-28, k = 1.38e-16, tiny = 1.0e-40;
    const double Tcmb = T0 * (1.0 + z);
    if (eps * meg * CHK_C * CHK_C / (k * Tcmb) > 90.0 * log(10.0)) return 0;
    if (exp(eps * meg * CHK_C * CHK_C / (k * Tcmb)) - 1.0 < tiny) {
        const double x = eps * meg * CHK_C * CHK_C / (k * Tcmb);
        return 8.0 * CHK_PI * CHK_C * CHK_C * CHK_C * pow(meg / CHK_H, 3.0) * eps /
               (meg * CHK_C * CHK_C * (1.0 + 0.5 * x + x / 6.0) / (k * Tcmb));
    }
    return 8.0 * CHK_PI * CHK_C * CHK_C * CHK_C * pow(meg / CHK_H, 3.0) * pow(eps, 2.0) /
           (exp(eps * meg * CHK_C * CHK_C / (k * Tcmb)) - 1.0);
}


CHK_FN double chk_FpiQ(int m, double Epi, double Ep)
{
    CHK_NOCONTRACT
    const double L = log(0.001 * Ep), x = Epi / Ep;
    double mpi = (m == 1) ? 1.40e-1 : 1.35e-1;
    double Bpi, alpha, r, xa, A;
    if (x <= mpi / Ep) return 0.0;
    if (x > 1.0) return 0.0;
    Bpi = 5.58 + 0.78 * L + 0.10 * L * L;
    alpha = 0.89 / (sqrt(Bpi) * (1.0 - exp(-0.33 * Bpi)));
    r = 3.1 / (pow(Bpi, 1.5));
    xa = pow(x, alpha);
    A = pow((1.0 - xa) / pow(1.0 + r * xa, 3.0), 4.0);
    A = A * (1.0 / (1.0 - xa) + 3.0 * r / (1.0 + r * xa)) * sqrt(1.0 - mpi / (x * Ep));
    return 4.0 * alpha * Bpi * xa * A / x / Ep;
}

CHK_FN double chk_dsigma(int i, double Epi, double Ep)
{
    CHK_NOCONTRACT
    const double gp = Ep / CHK_MP, gbp = sqrt(gp * gp - 1.0), pp = gbp * CHK_MP;
    return chk_pp_pion_cross_section(i, pp) * chk_FpiQ(0, Epi, Ep) / sqrt(1.0 + 1.0 / (gbp * gbp));
}

CHK_FN double chk_Fe(double Ee, double Epi)
{
    CHK_NOCONTRACT
    const double mpi = 1.3957e-1, mmu = 1.05658e-1, bmu = 0.2714;
    const double gpi = Epi / mpi, xi = mmu * mmu / (mpi * mpi);
    const double Emu = (mpi * mpi - mmu * mmu) / (mpi * mpi) * Epi / (2.0 * bmu);
    const double bm = sqrt(1.0 - mmu * mmu / (Emu * Emu));
    const double lam = 2.0 * mpi * mpi * Ee * bmu / ((mpi * mpi - mmu * mmu) * Epi);
    const double l2 = lam * lam, l3 = l2 * lam;
    const double P = 1.0 / bm * (2.0 * Epi * xi / (Emu * (1.0 - xi)) - (1.0 + xi) / (1.0 - xi));
    if (lam > 1.0 || gpi == 1.0 || lam < 0.0) return 0.0;
    if ((1.0 - bm) / (1.0 + bm) < lam)
        return (5.0 - 9.0 * l2 + 4.0 * l3 - P / bm * (1.0 - (6.0 * bm + 3.0) * l2 + (6.0 * bm + 2.0) * l3)) / 12.0;
    if (lam < (1.0 - bm) / (1.0 + bm))
        return l2 * bm / pow(1.0 - bm, 2.0) * (3.0 - 2.0 / 3.0 * lam * (3.0 + bm * bm) / (1.0 - bm)) -
               P / (1.0 - bm) *
                   (l2 * (1.0 + bm) - 2.0 * l2 / (1.0 - bm) * (0.5 + lam * (1.0 + bm)) +
                    2.0 * l3 * (3.0 + bm * bm) / (3.0 * pow(1.0 - bm, 2.0)));
    return 0.0;
}

CHK_FN double chk_lorentz_beta(double gamma)
{
    CHK_NOCONTRACT
    const double b2 = gamma * gamma - 1.0;
    if (b2 < 0.0) return 0.0;
    return sqrt(b2) / gamma;
}

CHK_FN double chk_qe_kernel(double Ee, double p)
{
    CHK_NOCONTRACT
    const double bmu = 0.2714, gp = sqrt(p * p + 1.0), Ep = gp * CHK_MP, Emax = gp * CHK_MP;
    const double r = 0.5 * (CHK_MPIC * CHK_MPIC - CHK_MMU * CHK_MMU) / (bmu * CHK_MPIC * CHK_MPIC);
    const double da = 0.05;
    double Emin = Ee / r, a, f, sum = 0.0, Epi, bmu_v;
    int i, Na;
    if (Emax < Emin || (Ep - CHK_MP) < 0.289) return 0;
    Emin = fmax(Emin, CHK_MPIC);
    Na = (int)fmax(50, round((log10(Emax) - log10(Emin)) / da));
    a = log10(Emin);
    Epi = pow(10.0, a);
    bmu_v = chk_lorentz_beta(Epi * r / CHK_MMU);
    f = (bmu_v == 0.0) ? 0.0 : (chk_dsigma(1, Epi, Ep) + chk_dsigma(-1, Epi, Ep)) * chk_Fe(Ee, Epi) / (bmu_v * Epi);
    for (i = 1; i < Na; i++) {
        const double Epre = Epi, fpre = f;
        a = a + da;
        Epi = pow(10.0, a);
        bmu_v = chk_lorentz_beta(Epi * r / CHK_MMU);
        f = (chk_dsigma(1, Epi, Ep) + chk_dsigma(-1, Epi, Ep)) * chk_Fe(Ee, Epi) / (bmu_v * Epi);
        sum = sum + 0.5 * (f + fpre) * (Epi - Epre);
        if (sum < 0.0 || isnan(sum)) return 0.0;
    }
    return 8.0 * bmu * CHK_MPIC * CHK_MPIC * (CHK_C * CHK_ME) * sum / (CHK_MPIC * CHK_MPIC - CHK_MMU * CHK_MMU);
}


CHK_FN double chk_gamma_kernel(double Eg, double p)
{
    CHK_NOCONTRACT
    const int Na = 200;
    const double gp = sqrt(p * p + 1.0), Ep = gp * CHK_MP, Emax = Ep;
    double Emin = Eg + CHK_MPI0 * CHK_MPI0 / (4.0 * Eg), a, da, Epi, f, sum = 0.0;
    int i;
    if (Emax < Emin) return 0;
    if (Emin < CHK_MPI0) Emin = CHK_MPI0;
    if ((Ep - CHK_MP) < 0.285) return 0;
    da = (log10(Emax) - log10(Emin)) / (double)Na;
    a = log10(Emin);
    Epi = pow(10.0, a);
    f = chk_dsigma(0, Epi, Ep) / sqrt(Epi * Epi - CHK_MPI0 * CHK_MPI0);
    for (i = 1; i < Na; i++) {
        const double Epre = Epi, fpre = f;
        a += da;
        Epi = pow(10.0, a);
        f = chk_dsigma(0, Epi, Ep) / sqrt(Epi * Epi - CHK_MPI0 * CHK_MPI0);
        sum += 0.5 * (f + fpre) * (Epi - Epre);
    }
    return 2.0 * CHK_C * sum / (4.0 * CHK_PI);
}

CHK_FN double chk_fe(double x)
{
    CHK_NOCONTRACT
    const double r = CHK_MMU * CHK_MMU / (CHK_MPIC * CHK_MPIC), x2 = x * x, r2 = r * r;
    if (x > r) return (3.0 - 2.0 * r) * (9.0 * x2 - 6.0 * log(x) - 4.0 * x2 * x - 5.0) / (9.0 * pow(1.0 - r, 2.0));
    return (3.0 - 2.0 * r) * (9.0 * r2 - 6.0 * log(r) - 4.0 * r2 * r - 5.0) / (9.0 * pow(1.0 - r, 2.0)) +
           (1.0 + 2.0 * r) * (r - x) * (9.0 * (r + x) - 4.0 * (r2 + r * x + x2)) / (9.0 * r2);
}

CHK_FN double chk_fnue(double x)
{
    CHK_NOCONTRACT
    const double r = 0.573, x2 = x * x, r2 = r * r;
    if (x > r)
        return 2.0 * (1.0 - x) * (6.0 * pow(1.0 - x, 2.0) + r * (5.0 + 5.0 * x - 4.0 * x2) + 6.0 * r * log(x) / (1.0 - x)) /
               (3.0 * pow(1.0 - r, 2.0));
    return 2.0 * ((1.0 - r) * (6.0 - 7.0 * r + 11.0 * r2 - 4.0 * r2 * r) + 6.0 * r * log(r)) / (3.0 * pow(1.0 - r, 2.0)) +
           2.0 * (r - x) * (7.0 * r2 - 4.0 * r2 * r + 7.0 * x * r - 4.0 * x * r2 - 2.0 * x2 - 4.0 * x2 * r) / (3.0 * r2);
}

CHK_FN double chk_lepton_integral(int kind, int sgn, double E, double Ep, double dx, double scale_x)
{
    CHK_NOCONTRACT
    double x = 0.0, f = 0.0, sum = 0.0;
    int i;
    for (i = 1; i < 200; i++) {
        const double fpre = f;
        double Epi;
        x += dx;
        Epi = E / x;
        if (kind == 1) f = chk_dsigma(sgn, Epi, Ep) * chk_fnue(x) / x;
        else if (kind == 2) f = chk_dsigma(sgn, Epi, Ep) * chk_fe(x) / x;
        else f = chk_dsigma(sgn, Epi, Ep) / x;
        if (i == 1) sum += f * dx;
        else sum += 0.5 * (f + fpre) * dx;
    }
    (void)scale_x;
    return sum;
}

CHK_FN double chk_ekernel(int kind, double E, double p)
{
    CHK_NOCONTRACT
    const double gp = sqrt(p * p + 1.0), Ep = gp * CHK_MP, Emax = gp * CHK_MP;
    double Emin = 2.0 * E * CHK_MPIC * CHK_MPIC / (CHK_MPIC * CHK_MPIC + CHK_MMU * CHK_MMU);
    if (Emax < Emin) return 0;
    if (Emin < CHK_MPIC) Emin = CHK_MPIC;
    if ((Ep - CHK_MP) < 0.289) return 0;
    return CHK_C * (chk_lepton_integral(kind, 1, E, Ep, 1.0 / 200.0, 1.0) +
                      chk_lepton_integral(kind, -1, E, Ep, 1.0 / 200.0, 1.0)) / (4.0 * CHK_PI);
}

CHK_FN double chk_numu1(double Enu, double p)
{
    CHK_NOCONTRACT
    const double lam = 1.0 - CHK_MMU * CHK_MMU / (CHK_MPIC * CHK_MPIC);
    const double gp = sqrt(p * p + 1.0), Ep = gp * CHK_MP, Emax = gp * CHK_MP;
    double Emin = Enu / lam;
    if (Emax < Emin) return 0;
    if (Emin < CHK_MPIC) Emin = CHK_MPIC;
    if ((Ep - CHK_MP) < 0.29) return 0;
    return CHK_C * (chk_lepton_integral(0, 1, Enu, Ep, lam / 200.0, 1.0) +
                      chk_lepton_integral(0, -1, Enu, Ep, lam / 200.0, 1.0)) / lam / (4.0 * CHK_PI);
}

CHK_FN double chk_nu_kernel(double E, double p)
{
    CHK_NOCONTRACT
    return chk_ekernel(2, E, p) + chk_ekernel(1, E, p) + chk_numu1(E, p);
}

CHK_FN double chk_ic_emissivity_kernel(double nu, double gamma, double z)
{
    CHK_NOCONTRACT
    const int N = 400;
    const double meg = 9.11e-28, c = 3.0e+10;
    const double eps = CHK_H * nu / (meg * c * c);
    double e0max = eps, e0min = eps / (4.0 * gamma * (gamma - eps)), a, da, Cn, nn, e0n, sum = 0.0;
    int i;
    if (e0min < 0) e0min = eps / (gamma * gamma);
    da = (log10(e0max) - log10(e0min)) / (double)N;
    a = log10(e0min);
    e0n = pow(10.0, a);
    Cn = chk_ic_kernel(e0n, gamma, eps);
    nn = chk_cmb_n(e0n, z);
    for (i = 0; i < N; i++) {
        const double e0 = e0n, C = Cn, n0 = nn;
        a = a + da;
        e0n = pow(10.0, a);
        nn = chk_cmb_n(e0n, z);
        Cn = chk_ic_kernel(e0n, gamma, eps);
        sum = sum + 0.5 * (C * n0 + Cn * nn) * (e0n - e0);
    }
    return CHK_H * eps * sum / (4.0 * CHK_PI);
}

#endif
