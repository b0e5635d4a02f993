"""Oracle restatement vs the reference's own functions (golden fixtures made by
tests/golden/make_golden.py from the partially compiled reference)."""
import os

import numpy as np
import pytest

import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden")
SIZES = (64, 128, 256, 512)


def load(n):
    return np.load(os.path.join(GOLD, f"ref_n{n}.npz"))


@pytest.fixture(scope="module", params=SIZES)
def case(request):
    n = request.param
    g = load(n)
    b = g["grid_bounds"]
    return n, g, po.Oracle(n, *b)


def test_grid_bit_exact(case):
    n, g, o = case
    for k in ("pp", "pdp", "pE"):
        assert np.array_equal(getattr(o, k), g[k]), k
    for k in ("ep", "edp", "eE"):
        assert np.array_equal(getattr(o, k), g[k]), k
    assert np.array_equal(np.array(o.pghost), g["pghost"])
    assert np.array_equal(np.array(o.eghost), g["eghost"])
    assert np.array_equal(o.beta_p, g["beta_p"])


def test_ic_cooling_rows_bit_exact(case):
    n, g, o = case
    for i, z in enumerate(g["ic_z"]):
        row, m1, p1 = o.ic_cooling_row(z)
        assert np.array_equal(row, g["ic_rows"][i])
        assert (m1, p1) == tuple(g["ic_edges"][i])


def test_electron_losses_bit_exact(case):
    n, g, o = case
    row, e = g["ic_rows"][1], g["ic_edges"][1]
    for i in range(len(g["le_n"])):
        radp, tl, m1, p1 = o.losses_e(g["le_n"][i], g["le_b"][i], g["le_divv"][i], row, e[0], e[1])
        assert np.array_equal(radp, g["le_radp"][i])
        assert np.array_equal(tl, g["le_tloss"][i])
        assert (m1, p1) == tuple(g["le_edges"][i])


def test_scalar_rates_bit_exact(case):
    n, g, o = case
    for a, p, k, want in zip(g["s_n"], g["s_p"], g["s_kbt"], g["s_coul_p"]):
        assert po.scalar("orc_b_coulomb_p", a, p, k) == want
    for a, p, want in zip(g["s_n"], g["s_p"], g["s_coul_e"]):
        assert po.scalar("orc_b_coulomb_e", a, p) == want
    for p, b, want in zip(g["s_p"], g["s_b"], g["s_synch"]):
        assert po.scalar("orc_b_synch", p, b) == want


def test_acceleration_times_bit_exact(case):
    n, g, o = case
    for l, cs, m, beta, dv, asa, t1, t0 in zip(g["t_l"], g["t_cs"], g["t_mach"], g["t_beta"], g["t_dv"],
                                              g["t_asa"], g["t_ttd1"], g["t_ttd0"]):
        assert po.scalar("orc_tacc_asa", l, cs, m, beta, 0.5) == asa
        assert po.scalar("orc_tacc_ttd", l, cs, m, dv, 1) == t1
        assert po.scalar("orc_tacc_ttd", l, cs, m, dv, 0) == t0


@pytest.mark.parametrize("mode", [0, 1])
def test_nsub_bit_exact(case, mode):
    n, g, o = case
    prm = po.default_params(tracer_nsub_safety=0.5, reacc_window_mode=mode, f_eddy=2.0)
    row = g["ic_rows"][1]
    ns, tg, non, endp = o.nsub(prm, 0.077, g["ns_n"], g["ns_kbt"], g["ns_b"], g["ns_divv"], row[0], row[-1],
                               g["ns_lturb"], g["ns_dv"])
    tag = f"w{mode}"
    assert np.array_equal(endp, g[f"nsub_endp_{tag}"])
    assert np.array_equal(ns, g[f"nsub_{tag}"])
    assert np.array_equal(tg, g[f"nsub_target_{tag}"])
    assert np.array_equal(non, g[f"nsub_non_{tag}"])


def test_thomas_pcr_bit_exact(case):
    n, g, o = case
    for s in range(g["tri_a"].shape[0]):
        rc, x = po.thomas(g["tri_a"][s], g["tri_b"][s], g["tri_c"][s], g["tri_rhs"][s])
        assert rc == 0
        assert np.array_equal(x, g["tri_thomas"][s])
    rc, x = po.pcr_batch(g["tri_a"], g["tri_b"], g["tri_c"], g["tri_rhs"])
    assert rc == 0
    assert np.array_equal(x, g["tri_pcr"])


def test_secondary_source_bit_exact(case):
    n, g, o = case
    inje = po.secondary_source(n, g["sec_n"], g["sec_crp"], g["sec_npmin"], g["sec_fqe"], g["sec_qepri"])
    assert np.array_equal(inje.reshape(g["sec_inje"].shape), g["sec_inje"])


def test_ic_kernel_and_emission_bit_exact(case):
    n, g, o = case
    fic = po.ic_kernel_flat(o, 8, 13.0, 28.0, 0.3)
    assert np.array_equal(fic, g["ic_fic"])
    eps = po.emission_ic(n, g["ic_cre"], o.edp, fic, 8)
    assert np.array_equal(eps, g["ic_eps"])


def test_synch_kernel_table_bit_exact(case):
    n, g, o = case
    fx = np.load(os.path.join(GOLD, "fx_table.npz"))
    nfreq, nth, nlogb = (int(v) for v in g["sy_dims"])
    table = po.synch_kernel_table(n, nfreq, nth, nlogb, -6.0, float(g["sy_inv_dlogb"][0]), fx["logx"],
                                  fx["logfx"], fx["fx"], float(fx["xmin"][0]), g["sy_logy"], g["sy_pw"])
    assert np.array_equal(table, g["sy_table"])
