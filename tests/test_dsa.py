"""DSA source builder (host C, no GPU needed).

Pinned against the compiled reference: DSA_MODELS.c builds in this container
(oracle/ref.mk) and tests/golden/make_golden.py records, for seeded shocks
over all eta models, the source rows of dsa_build_proton_source_fp /
dsa_build_electron_source_fp with the shock state and model the reference
driver fills (tracer_fp_dsa.c:240-312), and eta(M) of every model.  The
library must reproduce them bit for bit.

The per-tracer driver (tracer_fp_dsa.c:184-505) needs MPI headers and is not
compiled here; its arithmetic (input resolution, downstream volume, S V / dt,
species selection, skips) is checked through the relations it must obey."""
import os

import numpy as np
import pytest

import croma_hip as ch

GOLD = os.path.join(os.path.dirname(__file__), "golden")
MSUN = 1.99e33


@pytest.fixture(scope="module", params=[64, 256])
def gold(request):
    return request.param, np.load(os.path.join(GOLD, f"dsa_n{request.param}.npz"))


def test_eta_models_bit_exact(gold):
    _, d = gold
    for k in range(8):
        got = np.array([ch.dsa_eta(m, k) for m in d["eta_mach"]])
        assert np.array_equal(got, d["eta"][k]), f"model {k}"


def test_shock_sources_bit_exact(gold):
    n, d = gold
    for i in range(len(d["mach"])):
        prm = ch.dsa_default_params(eta_model=int(d["model"][i]), chi_p=float(d["chi"][i]), chi_e=float(d["chie"][i]),
                                    k_ep=float(d["kep"][i]), pmax_e=float(d["pmax_e"][i]))
        sp, se, fl = ch.dsa_shock_sources(n, prm, d["mach"][i], d["temp"][i], d["rho"][i], d["u1"][i])
        assert fl == tuple(d["flags"][i]), f"shock {i}"
        assert np.array_equal(sp, d["source_p"][i]), f"shock {i} protons"
        assert np.array_equal(se, d["source_e"][i]), f"shock {i} electrons"


def _shocks(rng, nt):
    mach = 10 ** rng.uniform(np.log10(1.2), 1.0, nt)
    temp = 10 ** rng.uniform(6.0, 8.0, nt)
    rho = 10 ** rng.uniform(-29.0, -26.0, nt)
    mass = 10 ** rng.uniform(6.0, 8.0, nt)
    return mach, temp, rho, mass


def _u1(mach, temp, prm):
    return mach * np.sqrt(prm.gamma_gas * 1.380649e-16 * temp / (prm.mu_mol * 1.67262192369e-24))


def test_source_mode_rows(gold):
    """qpi / qepri += S V_down / dt per tracer, V_down = M M_sun / (r rho1);
    electrons only where the proton source exists."""
    n = gold[0]
    rng = np.random.default_rng(11)
    nt, dt = 40, 0.077
    prm = ch.dsa_default_params()
    mach, temp, rho, mass = _shocks(rng, nt)
    mach[3] = np.nan  # no shock
    mach[4] = 1.4     # below DSAMinMach
    qpi, qepri = np.full(nt * n, 0.5), np.full(nt * n, 0.25)
    ninj = ch.dsa_inject(n, prm, ch.DSA_SOURCE, mach, rho, mass, temp1=temp, dt_gyr=dt, qpi=qpi, qepri=qepri)
    g = 5.0 / 3.0
    count = 0
    for i in range(nt):
        row_p, row_e = qpi[i * n:(i + 1) * n] - 0.5, qepri[i * n:(i + 1) * n] - 0.25
        if i in (3, 4):
            assert np.all(row_p == 0) and np.all(row_e == 0)
            continue
        sp, se, fl = ch.dsa_shock_sources(n, prm, mach[i], temp[i], rho[i], _u1(mach[i], temp[i], prm))
        m2 = mach[i] ** 2
        vd = mass[i] * MSUN / ((g + 1) * m2 / ((g - 1) * m2 + 2) * rho[i])
        np.testing.assert_allclose(row_p, sp * vd / dt, rtol=1e-12, atol=0)
        np.testing.assert_allclose(row_e, se * vd / dt if fl[1] else 0.0, rtol=1e-12, atol=0)
        count += int(fl[0])
    assert ninj == count


def test_state_mode_species_and_skips(gold):
    n = gold[0]
    rng = np.random.default_rng(12)
    nt = 24
    mach, temp, rho, mass = _shocks(rng, nt)
    side = np.zeros(nt)
    side[0], side[1] = -1.0, 99.0  # skipped sides
    base = dict(temp1=temp, side=side)
    full = ch.dsa_default_params()
    crp, cre = np.zeros(nt * n), np.zeros(nt * n)
    ch.dsa_inject(n, full, ch.DSA_STATE, mach, rho, mass, crp=crp, cre=cre, **base)
    assert np.all(crp[:2 * n] == 0) and np.all(cre[:2 * n] == 0)
    assert np.any(crp > 0) and np.any(cre > 0)
    # source mode = state mode / dt, exactly linear in the time step
    qpi, qepri = np.zeros(nt * n), np.zeros(nt * n)
    ch.dsa_inject(n, full, ch.DSA_SOURCE, mach, rho, mass, dt_gyr=0.5, qpi=qpi, qepri=qepri, **base)
    np.testing.assert_allclose(qpi, crp * 2.0, rtol=1e-15, atol=0)
    # species selection
    for species, proton, electron in ((1, 0, 1), (2, 1, 0)):  # electron only, proton only
        prm = ch.dsa_default_params(inject_species=species)
        a, b = np.zeros(nt * n), np.zeros(nt * n)
        ch.dsa_inject(n, prm, ch.DSA_STATE, mach, rho, mass, crp=a, cre=b, **base)
        assert np.array_equal(a, crp if proton else np.zeros_like(crp))
        assert np.array_equal(b, cre if electron else np.zeros_like(cre))


def test_merged_input_branch(gold):
    """Kinetic-flux input: u1 = (2 F / rho1)^(1/3), T from u1 and M, then the
    same builder (tracer_fp_dsa.c:49-71)."""
    n = gold[0]
    rng = np.random.default_rng(13)
    nt = 16
    mach, temp, rho, mass = _shocks(rng, nt)
    prm = ch.dsa_default_params()
    flux = 10 ** rng.uniform(-6.0, -2.0, nt)
    crp, cre = np.zeros(nt * n), np.zeros(nt * n)
    ch.dsa_inject(n, prm, ch.DSA_STATE, mach, rho, mass, flux=flux, crp=crp, cre=cre)
    for i in range(nt):
        u1 = np.cbrt(2.0 * flux[i] / rho[i])
        t = (prm.mu_mol * 1.67262192369e-24 / (prm.gamma_gas * 1.380649e-16)) * u1 ** 2 / mach[i] ** 2
        sp, se, fl = ch.dsa_shock_sources(n, prm, mach[i], t, rho[i], u1)
        g = prm.gamma_gas
        m2 = mach[i] ** 2
        vd = mass[i] * MSUN / ((g + 1) * m2 / ((g - 1) * m2 + 2) * rho[i])
        np.testing.assert_allclose(crp[i * n:(i + 1) * n], sp * vd, rtol=1e-9, atol=0)


# ---- reacceleration (tracer_fp_dsa_reacc.c, tracer_fp_dsa.c:322-505) ----
# The row functions compile in this container (oracle/ref.mk) and
# tests/golden/make_golden.py records their rows for seeded pre-shock
# spectra: the library must reproduce them bit for bit.  The driver that
# composes them needs MPI headers; its composition is checked by relations.

KB, MP, ME, CD = 1.380649e-16, 1.67262192369e-24, 9.1093837015e-28, 2.99792458e10


@pytest.fixture(scope="module", params=[64, 256])
def reacc_gold(request):
    return request.param, np.load(os.path.join(GOLD, f"dsa_reacc_n{request.param}.npz"))


def test_reacc_rows_bit_exact(reacc_gold):
    n, d = reacc_gold
    for i in range(len(d["mach"])):
        for kind in (ch.DSA_REACC_KIND_POSITIVE, ch.DSA_REACC_KIND_DRURY, ch.DSA_REACC_KIND_ADIABATIC):
            for sp in (0, 1):
                dl, ed = ch.dsa_reacc_delta(n, kind, sp, d["state"][i, sp], d["v1"][i], d["v2"][i], d["q"][i],
                                            d["r"][i], d["pmin"][i])
                assert np.array_equal(dl, d["delta"][i, kind, sp]), (i, kind, sp)
                assert ed == d["edens"][i, kind, sp], (i, kind, sp)
                _, e2 = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_ENERGY, sp, d["delta"][i, kind, sp], 0.0, 0.0, 0.0,
                                           0.0, d["pmin"][i])
                assert e2 == d["edens_of"][i, kind, sp], (i, kind, sp)
    # the cases exercise gaps, empty spectra and the Drury floor -current
    assert np.any(d["delta"][:, 1] < 0) and np.any(d["delta"][:, 0] > 0)


def test_eta_interpolated_bit_exact(reacc_gold):
    _, d = reacc_gold
    for k, (a, b) in enumerate(d["eta_pairs"]):
        got = np.array([[ch.dsa_eta_interpolated(m, a, b, x) for x in d["eta_x"]] for m in d["eta_mach"]])
        assert np.array_equal(got, d["eta_interp"][k]), (a, b)


def _pre_state(rng, nt, n):
    ax_p = 10 ** np.linspace(-1.0, 8.0, n)
    ax_e = 10 ** np.linspace(0.0, 6.0, n)
    s = rng.uniform(2.2, 3.2, nt)[:, None]
    crp = (10 ** rng.uniform(55, 62, nt)[:, None] * ax_p[None, :] ** -s).ravel()
    cre = (10 ** rng.uniform(50, 57, nt)[:, None] * ax_e[None, :] ** -s).ravel()
    return crp, cre


def _derived(prm, mach, temp, rho, mass):
    g = prm.gamma_gas
    r = (g + 1) * mach ** 2 / ((g - 1) * mach ** 2 + 2)
    vu = mass * MSUN / rho
    vd = mass * MSUN / (r * rho)
    q = 3.0 * r / (r - 1.0)
    ppm = prm.chi_p * np.sqrt(2.0 * KB * temp / (MP * CD * CD))
    pem = prm.chi_e * np.sqrt(2.0 * KB * temp / (ME * CD * CD))
    return r, vu, vd, q, ppm, pem


@pytest.mark.parametrize("n", [64, 256])
def test_drury83_composition_and_flags(n):
    """Drury83 without a cap adds V_down times the convolution delta of the
    pre-shock state to the fresh injection, in state and source mode, and
    flags exactly the reaccelerated tracers for the solve's dis input."""
    rng = np.random.default_rng(21)
    nt = 12
    mach, temp, rho, mass = _shocks(rng, nt)
    mach = np.maximum(mach, 2.0)
    mach[2] = np.nan  # no shock: untouched, not flagged
    crp0, cre0 = _pre_state(rng, nt, n)
    off = ch.dsa_default_params()
    dru = ch.dsa_default_params(reacc_mode=ch.DSA_REACC_DRURY83)
    a_p, a_e = crp0.copy(), cre0.copy()
    ch.dsa_inject(n, off, ch.DSA_STATE, mach, rho, mass, temp1=temp, crp=a_p, cre=a_e)
    b_p, b_e = crp0.copy(), cre0.copy()
    dis = np.zeros(nt, np.uint8)
    ch.dsa_inject(n, dru, ch.DSA_STATE, mach, rho, mass, temp1=temp, crp=b_p, cre=b_e, dis=dis)
    assert list(dis) == [0 if i == 2 else 1 for i in range(nt)]
    for i in range(nt):
        sl = slice(i * n, (i + 1) * n)
        if i == 2:
            assert np.array_equal(b_p[sl], crp0[sl]) and np.array_equal(b_e[sl], cre0[sl])
            continue
        r, vu, vd, q, ppm, pem = _derived(dru, mach[i], temp[i], rho[i], mass[i])
        dp_, _ = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_DRURY, 0, crp0[sl], vu, vd, q, r, ppm)
        de_, _ = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_DRURY, 1, cre0[sl], vu, vd, q, r, pem)
        # (b - a) cancels against the state it was added to: absolute
        # tolerance a few ulps of the state
        np.testing.assert_allclose(b_p[sl] - a_p[sl], dp_ * vd, rtol=1e-9, atol=1e-14 * np.max(np.abs(a_p[sl])))
        np.testing.assert_allclose(b_e[sl] - a_e[sl], de_ * vd, rtol=1e-9, atol=1e-14 * np.max(np.abs(a_e[sl])))
    # source mode: the same increments per unit time, read from the state
    qpi, qe = np.zeros(nt * n), np.zeros(nt * n)
    s_p, s_e = crp0.copy(), cre0.copy()
    ch.dsa_inject(n, dru, ch.DSA_SOURCE, mach, rho, mass, temp1=temp, dt_gyr=0.25, crp=s_p, cre=s_e, qpi=qpi,
                  qepri=qe)
    assert np.array_equal(s_p, crp0) and np.array_equal(s_e, cre0)
    np.testing.assert_allclose(qpi * 0.25, b_p - crp0, rtol=1e-12, atol=1e-12 * np.max(np.abs(b_p - crp0)))


def test_drury83_cap_holds_the_budget():
    """With reacc_eta_cap the applied Drury83 delta is the adiabatic rows
    plus one common scale s of the convolution's excess over them, in both
    species, and s times the excess energy (convolution rows over bins
    2..n-3 minus the adiabatic energy, times the downstream volume) is
    reacc_eta_cap M u1^2 / 2 where the excess is larger, s = 1 otherwise
    (tracer_fp_dsa.c:361-418).  Each tracer gets a cap at 0.3 or 3 times
    its own excess, so s is recovered from the rows without cancellation."""
    n = 128
    rng = np.random.default_rng(22)
    nt = 10
    mach, temp, rho, mass = _shocks(rng, nt)
    mach = np.maximum(mach, 2.5)
    crp0, cre0 = _pre_state(rng, nt, n)
    crp0 *= 1e3
    none = ch.dsa_default_params(inject_species=1)
    hit = 0
    for i in range(nt):
        sl = slice(i * n, (i + 1) * n)
        one = slice(i, i + 1)
        r, vu, vd, q, ppm, pem = _derived(none, mach[i], temp[i], rho[i], mass[i])
        u1 = _u1(mach[i], temp[i], none)
        kinetic = 0.5 * mass[i] * MSUN * u1 * u1
        ad_p_rows, ad_p = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_ADIABATIC, 0, crp0[sl], vu, vd, q, r, ppm)
        ad_e_rows, ad_e = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_ADIABATIC, 1, cre0[sl], vu, vd, q, r, pem)
        dv_p, _ = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_DRURY, 0, crp0[sl], vu, vd, q, r, ppm)
        dv_e, _ = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_DRURY, 1, cre0[sl], vu, vd, q, r, pem)
        _, cv_p = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_ENERGY, 0, dv_p, 0, 0, 0, 0, ppm)
        _, cv_e = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_ENERGY, 1, dv_e, 0, 0, 0, 0, pem)
        nonad = (cv_p + cv_e - ad_p - ad_e) * vd
        assert nonad > 0.0
        cap = (0.3 if i % 2 == 0 else 3.0) * nonad / kinetic
        budget = cap * kinetic
        prm = ch.dsa_default_params(reacc_mode=ch.DSA_REACC_DRURY83, reacc_eta_cap=cap, inject_species=1)
        a_p, a_e = crp0[sl].copy(), cre0[sl].copy()
        ch.dsa_inject(n, none, ch.DSA_STATE, mach[one], rho[one], mass[one], temp1=temp[one], crp=a_p, cre=a_e)
        b_p, b_e = crp0[sl].copy(), cre0[sl].copy()
        ch.dsa_inject(n, prm, ch.DSA_STATE, mach[one], rho[one], mass[one], temp1=temp[one], crp=b_p, cre=b_e)
        app = np.concatenate([(b_p - a_p) / vd, (b_e - a_e) / vd])
        excess = np.concatenate([dv_p - ad_p_rows, dv_e - ad_e_rows])
        base = np.concatenate([ad_p_rows, ad_e_rows])
        # one scale for every bin of both species (least squares over the rows)
        s = float(np.dot(app - base, excess) / np.dot(excess, excess))
        scale_rows = max(np.abs(app).max(), np.abs(base).max())
        np.testing.assert_allclose(app, base + s * excess, rtol=0, atol=1e-9 * scale_rows)
        if nonad > budget:
            hit += 1
            assert s * nonad == pytest.approx(budget, rel=1e-6)
        else:
            assert s == pytest.approx(1.0, rel=1e-9)
    assert hit == nt // 2


def test_positive_delta_budget():
    """Positive delta: nothing without a CR pressure (X_cr = 0 leaves eta at
    the initial model); otherwise the applied delta's energy is
    min(E_delta, (eta_interp - eta) M u1^2 / 2)."""
    n = 128
    rng = np.random.default_rng(23)
    nt = 10
    mach, temp, rho, mass = _shocks(rng, nt)
    mach = np.maximum(mach, 3.0)
    prm = ch.dsa_default_params(reacc_mode=ch.DSA_REACC_POSITIVE_DELTA, inject_species=1)
    none = ch.dsa_default_params(inject_species=1)
    zero_p, zero_e = np.zeros(nt * n), np.zeros(nt * n)
    ch.dsa_inject(n, prm, ch.DSA_STATE, mach, rho, mass, temp1=temp, crp=zero_p, cre=zero_e)
    ref_p, ref_e = np.zeros(nt * n), np.zeros(nt * n)
    ch.dsa_inject(n, none, ch.DSA_STATE, mach, rho, mass, temp1=temp, crp=ref_p, cre=ref_e)
    assert np.array_equal(zero_p, ref_p) and np.array_equal(zero_e, ref_e)
    crp0, cre0 = _pre_state(rng, nt, n)
    crp0 *= 1e4
    a_p, a_e = crp0.copy(), cre0.copy()
    ch.dsa_inject(n, none, ch.DSA_STATE, mach, rho, mass, temp1=temp, crp=a_p, cre=a_e)
    b_p, b_e = crp0.copy(), cre0.copy()
    ch.dsa_inject(n, prm, ch.DSA_STATE, mach, rho, mass, temp1=temp, crp=b_p, cre=b_e)
    moved = 0
    for i in range(nt):
        sl = slice(i * n, (i + 1) * n)
        r, vu, vd, q, ppm, pem = _derived(prm, mach[i], temp[i], rho[i], mass[i])
        u1 = _u1(mach[i], temp[i], prm)
        _, e_p = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_POSITIVE, 0, crp0[sl], vu, vd, q, r, ppm)
        _, e_e = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_POSITIVE, 1, cre0[sl], vu, vd, q, r, pem)
        energy = (e_p + e_e) * vd
        # X_cr = (E_cr / 3) / P_th, protons above xcr_pmin_pmc, E from the
        # state's kinetic energies (the solver's c) in the downstream volume
        pcr = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_ENERGY, 0, crp0[sl] / vd, 0, 0, 0, 0,
                                 prm.xcr_pmin_pmc)[1] / 3.0
        pth = rho[i] * KB * temp[i] / (prm.mu_mol * MP)
        x_cr = pcr / pth
        extra = max(ch.dsa_eta_interpolated(mach[i], prm.eta_model, prm.eta_model_reacc, x_cr)
                    - ch.dsa_eta(mach[i], prm.eta_model), 0.0)
        budget = extra * 0.5 * mass[i] * MSUN * u1 * u1
        dp_app = (b_p[sl] - a_p[sl]) / vd
        de_app = (b_e[sl] - a_e[sl]) / vd
        assert np.all(dp_app >= -1e-14 * np.max(np.abs(a_p[sl] / vd)))
        if budget > 0.0 and energy > 0.0:
            moved += 1
        # the applied rows are the positive delta rows scaled by
        # min(1, budget / E_delta)
        scale = min(1.0, budget / energy) if energy > 0 and budget > 0 else 0.0
        dp_full, _ = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_POSITIVE, 0, crp0[sl], vu, vd, q, r, ppm)
        de_full, _ = ch.dsa_reacc_delta(n, ch.DSA_REACC_KIND_POSITIVE, 1, cre0[sl], vu, vd, q, r, pem)
        np.testing.assert_allclose(dp_app, scale * dp_full, rtol=1e-9, atol=1e-14 * np.max(np.abs(a_p[sl] / vd)))
        np.testing.assert_allclose(de_app, scale * de_full, rtol=1e-9, atol=1e-14 * np.max(np.abs(a_e[sl] / vd)))
        if 0.0 < budget < energy:
            # the applied energy (the driver's measure over all bins) is the budget
            assert scale * energy == pytest.approx(budget, rel=1e-12)
    assert moved > 0


def test_reacc_needs_state():
    n = 64
    rng = np.random.default_rng(24)
    mach, temp, rho, mass = _shocks(rng, 4)
    prm = ch.dsa_default_params(reacc_mode=ch.DSA_REACC_DRURY83)
    with pytest.raises(ch.CromaHipError):
        ch.dsa_inject(n, prm, ch.DSA_SOURCE, mach, rho, mass, temp1=temp, dt_gyr=1.0, qpi=np.zeros(4 * n),
                      qepri=np.zeros(4 * n))
