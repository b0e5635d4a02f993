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
