"""Pins of the oracle functions the compiled reference cannot reach.

HADRONIC.c, Chang_Cooper.c, Synchrotron.c, EMISSION.c and the batch driver
(tracer_fp_batch.c, tracer_fp_setup.c) need GSL or MPI headers that this image
lacks, so the oracle's restatements of them are checked here against what the
mathematics fixes instead:

  hadronic kernels   particle-counting sum rules: two photons per neutral
                     pion, one muon neutrino per charged pion decay, one
                     electron / electron neutrino / muon neutrino per muon
                     decay (the Kelner decay spectra and the Brunetti-Blasi
                     muon-decay spectrum integrate to 1 and 1/4)
  pp cross section   an empirical fit (Kamae et al. 2006, Blattnig 2000): its
                     pieces join at their break points, so a mistyped or
                     transposed coefficient opens a jump; the production
                     threshold p = 1 GeV/c and the j_pp / sigmoid cache rule
  D_pp fill          the magnitude p^2 exp(-E/E_max) exp(-1/p) / (4 t_acc)
                     with t_acc from the accelerationtime_* functions that are
                     pinned bit-exactly against the compiled reference; the
                     eta cap from an energy integral of a one-bin state
  secondary weights  the weight rule integrates Qe beta over the proton grid
  gamma / nu and     the contractions reduce to quadratures of smooth
  synchrotron        integrands; the synchrotron sum's truncation, bin range,
  contractions       log-B interpolation and A(B) prefactor
  segment / window   decomposition identities of the snapshot update: a
  composite          segmented snapshot equals its segments run as separate
                     snapshots at the midpoint backgrounds, the on-window is
                     centred, one segment ignores the next snapshot, equal
                     backgrounds make the segment count irrelevant

Each check is chosen so that a plausible mistake -- a dropped factor, a wrong
index or branch, an off-by-one range -- fails it.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import pyoracle as po
import synth

MP = 0.938          # CONSTANTS.c
ME = 0.511e-3
C_LIGHT = 2.999e10
MPI0 = 0.135
GEV_ERG = 1.6e-3
MSUN_G = 1.99e33
MPC_CM = 3.0857e24
GYR_S = 3.1536e16


def _log_trapz(fn, a, b, n):
    x = np.logspace(math.log10(a), math.log10(b), n)
    y = np.array([fn(v) for v in x])
    return float(np.trapezoid(y, x))


def _e_proton(p):
    return math.sqrt(p * p + 1.0) * MP


def _pion_yield(kinds, ep, n=4000):
    """int dsigma dE_pi over the pion energies a proton of energy ep can make
    (FpiQ vanishes outside (m_pi0, E_p))."""
    L = po.lib()
    return _log_trapz(lambda e: sum(L.orc_hadron_dsigma(k, e, ep) for k in kinds), MPI0, ep * (1.0 - 1e-9), n)


# ------------------------------------------------------------- decay spectra
def test_kelner_decay_spectra_normalised():
    """Kelner et al. 2006 eqs. 36 and 40 (HADRONIC.c:302-322): one lepton per
    decay, continuous at x = r, vanishing at x = 1."""
    L = po.lib()
    r_mu = 1.05e-1 ** 2 / 1.40e-1 ** 2
    for f, r in ((L.orc_hadron_fe, r_mu), (L.orc_hadron_fnue, 0.573)):
        total = quad(f, 0.0, r, limit=200)[0] + quad(f, r, 1.0, limit=200)[0]
        assert abs(total - 1.0) < 1e-10
        assert abs(f(r * (1 - 1e-12)) - f(r * (1 + 1e-12))) < 1e-9
        assert abs(f(1.0 - 1e-12)) < 1e-9


def test_muon_decay_electron_spectrum_normalised():
    """Brunetti & Blasi 2005 / Blasi 1999 (HADRONIC.c:500-533): over
    lambda = E_e / E_e,max the spectrum integrates to 1/4 for a relativistic
    muon (unpolarised part (5 - 9 l^2 + 4 l^3) / 12 -> 1/4, the polarisation
    term -> 0), and to less for a slow one."""
    L = po.lib()
    mpi, mmu, bmu = 1.3957e-1, 1.05658e-1, 0.2714
    lam = np.linspace(1e-9, 1.0, 100001)
    vals = {}
    for epi in (0.3, 10.0, 1.0e3):
        k = 2.0 * mpi * mpi * bmu / ((mpi * mpi - mmu * mmu) * epi)
        fe = np.array([L.orc_hadron_Fe(v / k, epi) for v in lam])
        assert np.all(fe >= -1e-12)
        vals[epi] = float(np.trapezoid(fe, lam))
    assert abs(vals[1.0e3] - 0.25) < 2e-5
    assert abs(vals[10.0] - 0.25) < 1e-4
    assert vals[0.3] < vals[10.0]


# --------------------------------------------------------------- sum rules
P_SUM = (30.0, 1.0e3, 1.0e5)


@pytest.mark.parametrize("p", P_SUM)
def test_gamma_kernel_two_photons_per_neutral_pion(p):
    """gammaKernel (HADRONIC.c:139-174): a pi0 of energy E decays into two
    photons spread flat over a box of width p_pi, so
    int K_gamma dE_gamma = (2 c / 4 pi) int dsigma_0 dE_pi."""
    ep = _e_proton(p)
    lhs = _log_trapz(lambda e: po.lib().orc_gamma_kernel(e, p), 1e-6, ep * (1 - 1e-9), 3000)
    rhs = 2.0 * C_LIGHT / (4.0 * math.pi) * _pion_yield((0,), ep)
    assert abs(lhs / rhs - 1.0) < 3e-3


@pytest.mark.parametrize("p", P_SUM)
def test_numu_kernel_one_neutrino_per_charged_pion(p):
    """numu1Kernel (HADRONIC.c:239-289): pi -> mu nu_mu puts the neutrino flat
    in x = E_nu / E_pi over (0, lambda); the reference's x-quadrature covers
    199 of its 200 steps, so int K dE_nu = 0.995 (c / 4 pi) int dsigma_+- dE_pi."""
    ep = _e_proton(p)
    lhs = _log_trapz(lambda e: po.lib().orc_hadron_numu1(e, p), 1e-6, ep * (1 - 1e-9), 2000)
    rhs = C_LIGHT / (4.0 * math.pi) * _pion_yield((1, -1), ep)
    assert abs(lhs / rhs - 0.995) < 1e-3


@pytest.mark.parametrize("kind", (1, 2))
@pytest.mark.parametrize("p", (1.0e3,))
def test_muon_decay_neutrino_kernels_one_lepton_per_pion(kind, p):
    """eKernel for nu_e (fnue) and for the muon-decay nu_mu (fe)
    (HADRONIC.c:179-235): with spectra normalised to one lepton,
    int K dE = (c / 4 pi) int dsigma_+- dE_pi."""
    ep = _e_proton(p)
    lhs = _log_trapz(lambda e: po.lib().orc_hadron_ekernel(kind, e, p), 1e-6, ep * (1 - 1e-9), 2000)
    rhs = C_LIGHT / (4.0 * math.pi) * _pion_yield((1, -1), ep)
    assert abs(lhs / rhs - 1.0) < 2e-3


@pytest.mark.parametrize("p", P_SUM)
def test_secondary_electron_kernel_one_electron_per_charged_pion(p):
    """QeKernel_BB05 (HADRONIC.c:436-498), per unit p_e = p / (m_e c): one
    electron per charged pion, int Qe dp_e = c int dsigma_+- dE_pi, up to the
    reference's mix of pion / muon masses (0.140 / 0.105 GeV in the prefactor,
    0.13957 / 0.105658 in Fe: up to 2.4% low) and its 0.05-dex pion-energy grid."""
    ep = _e_proton(p)
    lhs = _log_trapz(lambda pe: po.lib().orc_qe_kernel_bb05(ME * math.sqrt(1.0 + pe * pe), p), 1e-2,
                     ep / ME, 3000)
    rhs = C_LIGHT * _pion_yield((1, -1), ep)
    assert 0.97 < lhs / rhs < 1.0


def test_hadronic_kernels_vanish_below_threshold():
    L = po.lib()
    # kinetic energies below 0.285 / 0.289 / 0.29 GeV (HADRONIC.h:10-14)
    for tk, fn in ((0.28, lambda e, p: L.orc_gamma_kernel(e, p)),
                   (0.285, lambda e, p: L.orc_hadron_ekernel(1, e, p)),
                   (0.285, lambda e, p: L.orc_qe_kernel_bb05(e, p)),
                   (0.285, lambda e, p: L.orc_hadron_numu1(e, p))):
        p = math.sqrt(((MP + tk) / MP) ** 2 - 1.0)
        assert all(fn(e, p) == 0.0 for e in np.logspace(-3, 0, 10))


# ---------------------------------------------------------- pp cross section
# (momentum break points of the nondiffractive / diffractive pieces, total
# energy break points of the Delta and res pieces; tolerated relative jump)
SIGMA_BREAKS_P = ((2.4, 1e-2), (10.0, 4e-3), (3.2, 1e-4), (100.0, 3e-3))
SIGMA_BREAKS_E = ((1.6, 3e-2), (1.8, 1e-4), (1.9, 1e-4), (2.3, 1e-4), (10.0, 1e-6), (20.0, 1e-5))


def test_pp_cross_section_pieces_join():
    """Kamae et al. 2006 (errata 2007) fit pieces join at their break points:
    a mistyped or transposed coefficient of any piece opens a jump there.
    (The fit has two jumps of its own: the low-momentum piece is scaled to
    join the next one at log10 p = a0 = 0.1176 while the switch is at
    p = 1.3, 5% apart; the Delta resonance turns on at E_p = 1.4 GeV.)"""
    f = po.lib().orc_pp_pion_cross_section
    eps = 1e-9
    for kind in (0, 2):
        for p, tol in SIGMA_BREAKS_P:
            lo, hi = f(kind, p * (1 - eps)), f(kind, p * (1 + eps))
            assert abs(hi - lo) <= tol * hi, (kind, p, lo, hi)
        for e, tol in SIGMA_BREAKS_E:
            p = math.sqrt(e * e - MP * MP)
            lo, hi = f(kind, p * (1 - eps)), f(kind, p * (1 + eps))
            assert abs(hi - lo) <= tol * hi, (kind, e, lo, hi)
        lo, hi = f(kind, 1.3 * (1 - eps)), f(kind, 1.3 * (1 + eps))
        assert 0.0 < (hi - lo) / hi < 0.06


def test_pp_cross_section_threshold_and_grid_cache():
    """sigma vanishes up to p = 1 GeV/c (log10 p = 0 kills the nondiffractive
    piece; the others start higher) and is positive above.  The grid cache
    (FP_Coef.c:77-116): sigma_j = sigma_2(m_p p_j), j_pp = 2 + the first bin
    below n - 100 with sigma > 0 (2 when n <= 100), sigmoid 1/(1+e^{-5(p-p_jpp)})."""
    f = po.lib().orc_pp_pion_cross_section
    assert f(2, 1.0) == 0.0 and f(2, 0.999) == 0.0 and f(2, 1.0 + 1e-9) > 0.0
    for n in (64, 128, 256, 512):
        o = po.Oracle(n)
        sig, smd = o.sigma_pp()
        assert np.array_equal(sig, np.array([f(2, MP * p) for p in o.pp]))
        first = next((j for j in range(max(n - 100, 0)) if MP * o.pp[j] > 1.0), None)
        want = 2 if (first is None or n <= 100) else first + 2
        assert o.j_pp == want
        assert smd[o.j_pp] == 0.5
        assert np.all(np.diff(smd) >= 0.0)
        assert np.allclose(smd, 1.0 / (1.0 + np.exp(-5.0 * (o.pp - o.pp[o.j_pp]))), rtol=1e-14)


# ------------------------------------------------------------------- D_pp
@pytest.mark.parametrize("model", (po.DPP_ASA, po.DPP_TTD, po.DPP_DIRECT))
def test_dpp_magnitude_from_acceleration_time(model):
    """D_pp = p^2 exp(-E/E_max) exp(-1/p) / ((q + 2) t_acc), q = 2
    (FP_Coef.c:680-849); electrons without the E_max cutoff.  t_acc comes from
    accelerationtime_ASA / _TTD (bit-exact against the compiled reference) or
    the direct value; a wrong (q + 2) or a dropped cutoff fails."""
    o = po.Oracle(128)
    l_turb, dv, cs, beta = 0.15, 3.0e6, 1.0e7, 50.0
    prm = po.default_params(dpp_model=model, t_acc_direct_gyr=0.37)
    d, de, edges = o.dpp(prm, model, l_turb, dv, cs, beta)
    mach = min(dv / cs, prm.mach_limit)
    L = po.lib()
    if model == po.DPP_ASA:
        t_acc = L.orc_tacc_asa(l_turb * MPC_CM, cs, mach, beta, prm.psi)
    elif model == po.DPP_TTD:
        t_acc = L.orc_tacc_ttd(l_turb * MPC_CM, cs, mach, dv, prm.ttd_tacc_model)
    else:
        t_acc = 0.37
    want_p = o.pp ** 2 * np.exp(-o.pE / prm.epmax) * np.exp(-1.0 / o.pp) / (4.0 * t_acc)
    want_e = o.ep ** 2 * np.exp(-1.0 / o.ep) / (4.0 * t_acc)
    assert np.max(np.abs(d / want_p - 1.0)) < 1e-13
    assert np.max(np.abs(de / want_e - 1.0)) < 1e-13
    # ghost cells: {dppm1, dppp1, dppem1, dppep1}
    pm1, pp1 = o.pghost
    em1, ep1 = o.eghost
    gp = [p ** 2 * math.exp(-MP * math.sqrt(1 + p * p) / prm.epmax) * math.exp(-1.0 / p) / (4.0 * t_acc)
          for p in (pm1, pp1)]
    ge = [p ** 2 * math.exp(-1.0 / p) / (4.0 * t_acc) for p in (em1, ep1)]
    assert np.allclose(edges, gp + ge, rtol=1e-13, atol=0.0)


def test_dpp_eta_cap_from_one_bin_energy():
    """eta cap (fp_shared_core.c:974-1081): eta = E_CR / (Q_turb t_acc) with
    E_CR = sum over bins 2 .. n-3 of (E - m) N dp [erg], Q_turb =
    M dv^3 / (2 L); D is scaled by cap / eta when eta exceeds the cap.  A
    state with one occupied bin has a known E_CR; bins 1 and n-2 are outside
    the sum."""
    n = 64
    o = po.Oracle(n)
    l_turb, dv, cs, beta, mass = 0.15, 3.0e6, 1.0e7, 50.0, 1.0e7
    base = po.default_params(dpp_model=po.DPP_TTD)
    d0, de0, _ = o.dpp(base, po.DPP_TTD, l_turb, dv, cs, beta)
    mach = min(dv / cs, base.mach_limit)
    t_acc = po.lib().orc_tacc_ttd(l_turb * MPC_CM, cs, mach, dv, base.ttd_tacc_model)
    q_turb = 0.5 * mass * MSUN_G * dv ** 3 / (l_turb * MPC_CM)
    jp, je = 17, 40
    crp = np.zeros(n)
    cre = np.zeros(n)
    crp[jp] = 1.0e-3
    cre[je] = 2.0e-2
    ecr = (o.pE[jp] - MP) * GEV_ERG * crp[jp] * o.pdp[jp] + (o.eE[je] - ME) * GEV_ERG * cre[je] * o.edp[je]
    eta = ecr / (q_turb * t_acc * GYR_S)
    cap = eta / 3.0
    prm = po.default_params(dpp_model=po.DPP_TTD, eta_dpp_cap=cap)
    d, de, _ = o.dpp(prm, po.DPP_TTD, l_turb, dv, cs, beta, mass=mass, crp=crp, cre=cre)
    assert np.allclose(d / d0, cap / eta, rtol=1e-12) and np.allclose(de / de0, cap / eta, rtol=1e-12)
    # below the cap: untouched
    prm_hi = po.default_params(dpp_model=po.DPP_TTD, eta_dpp_cap=eta * 1.01)
    d2, _, _ = o.dpp(prm_hi, po.DPP_TTD, l_turb, dv, cs, beta, mass=mass, crp=crp, cre=cre)
    assert np.array_equal(d2, d0)
    # energy in bins 1 and n-2 does not count
    for j in (1, n - 2):
        edge = np.zeros(n)
        edge[j] = 1.0e6
        d3, _, _ = o.dpp(prm, po.DPP_TTD, l_turb, dv, cs, beta, mass=mass, crp=edge, cre=edge)
        assert np.array_equal(d3, d0)


# ------------------------------------------------------- quadrature weights
@pytest.mark.parametrize("n", (64, 128, 256))
def test_secondary_weights_rule(n):
    """prepare_secondary_kernel_flat (fp_shared_core.c:1453-1506): the weight
    of proton bin jp for electron bin je is Qe(E_e, p_jp) beta_jp times a
    quadrature weight.  Divided by Qe beta, the weights of one electron bin
    above jp0 telescope to sum(dp[jp0 + 1 ..]) - dp[jp0 + 1] / 2 (a dropped half
    weight at the top, or a range off by one, changes that), and nothing below
    bin jp0 - 1 carries weight (np_min_qe = jp0 - 1 is the first contributing
    bin, jp0 the last with Qe beta below 1e-80)."""
    o = po.Oracle(n)
    fqe, npmin = o.secondary_kernel()
    W = fqe.reshape(n, n)
    beta = o.pp / np.sqrt(1.0 + o.pp ** 2)
    for je in range(0, n, max(1, n // 8)):
        lo = int(npmin[je])
        if lo >= n:
            assert np.all(W[:, je] == 0.0)
            continue
        val = np.array([po.lib().orc_qe_kernel_bb05(o.eE[je], p) for p in o.pp]) * beta
        assert np.all(W[:lo, je] == 0.0)
        # bins lo and jp0 = lo + 1 hold Qe beta < 1e-80 by construction; above
        # them every value is positive and the weights telescope
        jp0 = lo + 1
        assert np.all(val[jp0 + 1:] > 0.0)
        total = float(np.sum(W[jp0 + 1:, je] / val[jp0 + 1:]))
        want = float(np.sum(o.pdp[jp0 + 1:])) - 0.5 * o.pdp[jp0 + 1]
        assert abs(total - want) <= 1e-12 * want


def test_secondary_weights_converge_at_first_order():
    """The weighted sum approaches int Qe beta dp over the kernel's support as
    the grid refines, with error halving per doubling of n: first order, as
    the half-bin shift of the rule makes it (a centred rule would converge at
    second order)."""
    ratios = []
    for n in (64, 128, 256, 512):
        o = po.Oracle(n)
        fqe, npmin = o.secondary_kernel()
        W = fqe.reshape(n, n)
        je = n // 2
        lo = int(npmin[je])
        x = np.logspace(math.log10(o.pp[lo + 1]), math.log10(o.pp[-1]), 4000)
        y = np.array([po.lib().orc_qe_kernel_bb05(o.eE[je], v) for v in x]) * x / np.sqrt(1.0 + x * x)
        ratios.append(W[:, je].sum() / np.trapezoid(y, x))
    err = np.array(ratios) - 1.0
    assert np.all(err > 0.0)
    assert np.all((err[:-1] / err[1:] > 1.7) & (err[:-1] / err[1:] < 2.7)), ratios


def test_hadronic_contraction_is_the_shifted_quadrature():
    """Q_one (EMISSION.c:297-308) through prepare_gamma/neutrino_emission
    (fp_shared_core.c:2166-2243): eps = E_k n_gas * 0.5 sum_{i>=1} (f_{i-1} +
    f_i) dp_i, f = beta N K.  Linear in the spectrum and in n_gas; a spectrum
    with beta N K = 1 gives the telescoped weight sum(dp[1 .. n-1]); bin 0
    enters with dp_1 / 2 only."""
    n, nbins = 64, 4
    o = po.Oracle(n)
    emin, emax = -1.0, 3.0
    fk = np.abs(np.random.default_rng(2).standard_normal(nbins * n)) + 0.1
    e = 10.0 ** (emin + (np.arange(nbins) + 1) * (emax - emin) / nbins - (emax - emin) / nbins)
    spec = np.zeros((nbins, n))
    for k in range(nbins):
        spec[k] = 1.0 / (o.beta_p * fk[k * n:(k + 1) * n])
    n_gas = np.array([2.0])
    for k in range(nbins):
        eps = po.emission_hadronic(n, nbins, emin, emax, n_gas, spec[k], o.beta_p, o.pdp, fk).reshape(1, nbins)
        want = e[k] * 2.0 * float(np.sum(o.pdp[1:]))
        assert abs(eps[0, k] - want) <= 1e-13 * want
    # bin 0 alone: half of dp_1
    one = np.zeros(n)
    one[0] = 1.0
    eps = po.emission_hadronic(n, nbins, emin, emax, np.array([1.0]), one, o.beta_p, o.pdp, fk).reshape(1, nbins)
    assert np.allclose(eps[0], e * 0.5 * o.beta_p[0] * fk.reshape(nbins, n)[:, 0] * o.pdp[1], rtol=1e-14)


# ------------------------------------------------------------ synchrotron
def _synch_setup(n=64, nfreq=3, nlogb=8, seed=4):
    rng = np.random.default_rng(seed)
    o = po.Oracle(n)
    table = np.exp(rng.uniform(-3.0, 3.0, nlogb * nfreq * n))
    return o, table


def test_synch_contraction_on_a_plane():
    """prepare_synch_emission_cell_major (fp_shared_core.c:2070-2129): with B
    on a table plane (t = 0) the emissivity is A(B) sum_{je=2}^{n-2} N dp K,
    A(B) = sqrt(3) e^3 B / (2 m_e c^2) / (4 pi), e = 4.8e-10, m_e = 9.11e-28."""
    n, nfreq, nlogb = 64, 3, 8
    o, table = _synch_setup(n, nfreq, nlogb)
    logb_min, inv_dlogb = -3.0, 1.0
    cre = o.ep ** -2.2
    b = np.array([1.0e-6, 1.0e-7])        # log10(B / 1 uG) = 0, -1: planes 3 and 2
    eps = po.emission_synch(n, nfreq, nlogb, logb_min, inv_dlogb, table, np.tile(cre, len(b)), o.edp,
                            b).reshape(len(b), nfreq)
    T = table.reshape(nlogb, nfreq, n)
    for i, ib in enumerate((3, 2)):
        A = math.sqrt(3.0) * 4.8e-10 ** 3 * b[i] / (2.0 * 9.11e-28 * C_LIGHT ** 2) / (4.0 * math.pi)
        want = A * np.array([np.sum(cre[2:n - 1] * T[ib, k, 2:n - 1] * o.edp[2:n - 1]) for k in range(nfreq)])
        assert np.allclose(eps[i], want, rtol=1e-13, atol=0.0)


def test_synch_contraction_truncation_and_bin_range():
    """The sum runs from bin 2 and stops before the first bin je whose
    successor N[je + 1] is not positive; bins 0 and 1 never enter and bin
    n - 1 only as the successor test."""
    n, nfreq, nlogb = 64, 3, 8
    o, table = _synch_setup(n, nfreq, nlogb)
    args = (n, nfreq, nlogb, -3.0, 1.0, table)
    b = np.array([3.3e-6])
    cre = o.ep ** -2.0
    base = po.emission_synch(*args, cre, o.edp, b)
    alt = cre.copy()
    alt[0] = 5.0
    alt[1] = 0.0
    alt[n - 1] = 123.0
    assert np.array_equal(po.emission_synch(*args, alt, o.edp, b), base)
    cut = cre.copy()
    cut[30] = 0.0                      # bin 29 (empty successor) and all above drop out
    trunc = cre.copy()
    trunc[30:] = 0.0
    got = po.emission_synch(*args, cut, o.edp, b)
    want = po.emission_synch(*args, trunc, o.edp, b)
    assert np.array_equal(got, want)
    assert np.all(got < base)


def test_synch_contraction_log_interpolation_in_b():
    """Between planes the kernel is interpolated in log K (geometric mean at
    t = 1/2); where either plane is zero, linearly."""
    n, nfreq, nlogb = 64, 1, 4
    o = po.Oracle(n)
    table = np.zeros((nlogb, nfreq, n))
    table[1, 0, :] = 1.0
    table[2, 0, :] = 4.0
    table[3, 0, :] = 0.0
    cre = np.ones(n)
    logb_min, inv_dlogb = -3.0, 1.0
    # log10(B / uG) = -1.5: planes 1, 2 at t = 1/2 -> K = 2
    b = np.array([1.0e-6 * 10 ** -1.5])
    eps = po.emission_synch(n, nfreq, nlogb, logb_min, inv_dlogb, table.ravel(), cre, o.edp, b).reshape(1, 1)
    A = math.sqrt(3.0) * 4.8e-10 ** 3 * b[0] / (2.0 * 9.11e-28 * C_LIGHT ** 2) / (4.0 * math.pi)
    s = float(np.sum(o.edp[2:n - 1]))
    assert abs(eps[0, 0] - A * 2.0 * s) <= 1e-12 * A * 2.0 * s
    # planes 2 (K = 4) and 3 (K = 0) at t = 1/2: linear, K = 2
    b = np.array([1.0e-6 * 10 ** -0.5])
    eps = po.emission_synch(n, nfreq, nlogb, logb_min, inv_dlogb, table.ravel(), cre, o.edp, b).reshape(1, 1)
    A = math.sqrt(3.0) * 4.8e-10 ** 3 * b[0] / (2.0 * 9.11e-28 * C_LIGHT ** 2) / (4.0 * math.pi)
    assert abs(eps[0, 0] - A * 2.0 * s) <= 1e-12 * A * 2.0 * s


# ------------------------------------------------ segment / window composite
def _snap_inputs(w, s, o):
    row, m1, p1 = o.ic_cooling_row(w.z[s])
    rown, m1n, p1n = o.ic_cooling_row(w.z[s + 1])
    return row, m1, p1, rown, m1n, p1n


def _run_snap(o, prm, w, nsub, non, allow, bgc, bgn, row, m1, p1, rown, m1n, p1n, dt_snap, crp, cre,
              elapsed=0.0, secondary=None):
    zeros = np.zeros(w.ntracer * w.n)
    fqe, npmin = secondary if secondary is not None else (np.zeros(w.n * w.n), np.full(w.n, w.n, np.int32))
    o.execute_snapshot(prm, nsub, non, allow, elapsed, dt_snap, bgc, row, m1, p1, bgn, rown, m1n, p1n, zeros,
                       zeros, fqe, npmin, w.mass, None, crp, cre)


def _lerp(a, b, t):
    return a + t * (b - a)


def test_segment_count_irrelevant_for_equal_backgrounds():
    """split_range / tracer_fp_batch_execute_bucket (tracer_fp_batch.c:63-79,
    651-827): with the next snapshot's backgrounds and IC row equal to the
    current ones, every segment rebuilds the same operator, so the result
    cannot depend on the segment count -- unless the steps do not add up to
    nsub, dt depends on the split, or the on-window moves with it."""
    w = synth.make_workload(n=64, ntracer=6, nsnap=2, species=1, dpp_model=2, seed=12, z0=0.02)
    o = po.Oracle(64)
    row, m1, p1, _, _, _ = _snap_inputs(w, 0, o)
    nsub = np.array([37, 64, 100, 129, 250, 16], np.int32)
    non = np.array([9, 64, 33, 40, 125, 4], np.int32)
    outs = []
    for maxseg in (1, 2, 4, 16):
        prm = po.default_params(dpp_model=2, reacc_window_mode=1, coeff_interp_max_segments=maxseg)
        crp, cre = w.crp0.copy(), w.cre0.copy()
        _run_snap(o, prm, w, nsub, non, 1, w.bg[0], w.bg[0], row, m1, p1, row, m1, p1, w.dt_snap[0], crp, cre)
        outs.append(cre)
    for c in outs[1:]:
        assert np.array_equal(c, outs[0])
    assert not np.array_equal(outs[0], w.cre0)


def test_one_segment_uses_the_current_snapshot():
    """With one segment (alpha = 0) the next snapshot's values never enter."""
    w = synth.make_workload(n=64, ntracer=4, nsnap=2, species=3, dpp_model=1, seed=13, z0=0.02)
    o = po.Oracle(64)
    row, m1, p1, rown, m1n, p1n = _snap_inputs(w, 0, o)
    nsub = np.array([20, 31, 7, 50], np.int32)
    prm = po.default_params(dpp_model=1, coeff_interp_max_segments=1)
    sec = o.secondary_kernel()
    a = (w.crp0.copy(), w.cre0.copy())
    b = (w.crp0.copy(), w.cre0.copy())
    _run_snap(o, prm, w, nsub, nsub, 1, w.bg[0], w.bg[1], row, m1, p1, rown, m1n, p1n, w.dt_snap[0], *a,
              secondary=sec)
    _run_snap(o, prm, w, nsub, nsub, 1, w.bg[0], w.bg[0] * 3.0, row, m1, p1, 2.0 * rown, m1n, p1n,
              w.dt_snap[0], *b, secondary=sec)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("species", (1, 3))
def test_two_segments_are_two_midpoint_snapshots(species):
    """Two coefficient segments (tracer_fp_batch.c:718-773): the first nsub/2
    steps use the backgrounds and IC row interpolated at alpha = 1/4, the rest
    at 3/4, and (two species) the secondary source is refreshed from the
    protons at each segment start (tracer_fp_cpu.c:140-147).  The same
    result comes from two unsegmented snapshots of nsub/2 steps each at those
    backgrounds, with dt unchanged (dt_snap / 2 over nsub / 2)."""
    w = synth.make_workload(n=64, ntracer=5, nsnap=2, species=species, dpp_model=1, seed=14, z0=0.02)
    o = po.Oracle(64)
    row, m1, p1, rown, m1n, p1n = _snap_inputs(w, 0, o)
    nsub = np.array([8, 12, 20, 40, 64], np.int32)   # even, >= 8: two segments at max 2
    prm = po.default_params(dpp_model=1, coeff_interp_max_segments=2)
    sec = o.secondary_kernel() if species == 3 else None
    assert all(o.choose_segments(prm, int(v), 1) == 2 for v in nsub)
    seg = (w.crp0.copy(), w.cre0.copy())
    _run_snap(o, prm, w, nsub, nsub, 1, w.bg[0], w.bg[1], row, m1, p1, rown, m1n, p1n, w.dt_snap[0], *seg,
              secondary=sec)
    two = (w.crp0.copy(), w.cre0.copy())
    half = nsub // 2
    for t in (0.25, 0.75):
        bg = _lerp(w.bg[0], w.bg[1], t)
        r = _lerp(row, rown, t)
        mm, pp = _lerp(m1, m1n, t), _lerp(p1, p1n, t)
        _run_snap(o, prm, w, half, half, 0, bg, bg, r, mm, pp, r, mm, pp, w.dt_snap[0] / 2.0, *two, secondary=sec)
    assert np.array_equal(seg[1], two[1])
    assert np.array_equal(seg[0], two[0])


def test_on_window_is_centred():
    """The re-acceleration window (tracer_fp_batch.c:701-706, 784-797): n_on
    steps with the on-set D_pp starting at step (nsub - n_on) / 2, off-floor
    steps before and after.  The same result comes from three unwindowed
    snapshots: floor (D_pp model off), model, floor, each at dt = dt_snap /
    nsub (dyadic dt_snap, so every quotient is exact)."""
    w = synth.make_workload(n=64, ntracer=3, nsnap=2, species=1, dpp_model=2, seed=15, z0=0.02)
    o = po.Oracle(64)
    row, m1, p1, _, _, _ = _snap_inputs(w, 0, o)
    nsub = np.array([41, 100, 17], np.int32)
    non = np.array([10, 33, 17], np.int32)
    d = 2.0 ** -10
    on = po.default_params(dpp_model=2, reacc_window_mode=1)
    win = (w.crp0.copy(), w.cre0.copy())
    for i in range(3):
        sl = slice(i * 64, (i + 1) * 64)
        wi = w.subset([i])
        part = (win[0][sl].copy(), win[1][sl].copy())
        _run_snap(o, on, wi, nsub[i:i + 1], non[i:i + 1], 0, wi.bg[0], wi.bg[0], row, m1, p1, row, m1, p1,
                  float(nsub[i]) * d, *part)
        win[1][sl] = part[1]
    ref = w.cre0.copy()
    for i in range(3):
        wi = w.subset([i])
        a = (int(nsub[i]) - int(non[i])) // 2
        parts = ((a, 3), (int(non[i]), 2), (int(nsub[i]) - a - int(non[i]), 3))
        x = (np.zeros(64), ref[i * 64:(i + 1) * 64].copy())
        for steps, model in parts:
            if steps == 0:
                continue
            prm = po.default_params(dpp_model=model)
            k = np.array([steps], np.int32)
            _run_snap(o, prm, wi, k, k, 0, wi.bg[0], wi.bg[0], row, m1, p1, row, m1, p1, steps * d, *x)
        ref[i * 64:(i + 1) * 64] = x[1]
    assert np.array_equal(win[1], ref)


def test_direct_tacc_switch_time():
    """direct_tacc (tracer_fp_batch.c:47-60, cpu_direct_tacc_value): a segment
    whose midpoint lies before t_off uses t_acc_off, after it the direct value."""
    w = synth.make_workload(n=64, ntracer=3, nsnap=2, species=1, dpp_model=4, seed=16, z0=0.02)
    o = po.Oracle(64)
    row, m1, p1, _, _, _ = _snap_inputs(w, 0, o)
    nsub = np.array([20, 40, 33], np.int32)
    base = dict(dpp_model=4, t_acc_direct_gyr=0.2, t_acc_off_gyr=1.5, coeff_interp_max_segments=1)

    def run(prm, elapsed):
        x = (w.crp0.copy(), w.cre0.copy())
        _run_snap(o, prm, w, nsub, nsub, 0, w.bg[0], w.bg[0], row, m1, p1, row, m1, p1, w.dt_snap[0], *x,
                  elapsed=elapsed)
        return x[1]
    before = run(po.default_params(t_off_gyr=10.0, **base), 1.0)
    after = run(po.default_params(t_off_gyr=0.5, **base), 1.0)
    plain_off = run(po.default_params(**dict(base, t_acc_direct_gyr=1.5)), 1.0)
    plain_on = run(po.default_params(**base), 1.0)
    assert np.array_equal(before, plain_off) and np.array_equal(after, plain_on)
    assert not np.array_equal(before, after)
