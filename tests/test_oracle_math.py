"""Mathematical pins of the CPU oracle, for the parts the compiled reference
fixtures leave open (Chang-Cooper face weights and assembly, D_pp closed
forms, solver brute force, evolution invariants) and plausibility checks of
the parts that stay parity-unpinned (pp cross section, QeKernel_BB05,
gamma / neutrino kernels, the F(x) table).  Each check is chosen so that a
plausible mistake -- a dropped term, a wrong sign or index, swapped
operands -- fails it."""
import math
import os

import numpy as np
import pytest

import pyoracle as po
import run_oracle
import synth

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- face weights
W_GRID = np.concatenate([-np.logspace(-8, 2.6, 200), [0.0], np.logspace(-8, 2.6, 200), [0.0999999, 0.1000001]])


def test_face_weights_bernoulli_identities():
    for w in W_GRID:
        d, wp, wm = po.face_weights(w) if w != 0.0 else (0.5, *po.face_weights(w)[1:])
        assert wp > 0.0 and wm > 0.0
        # W+ - W- = w (B(-w) - B(w) = w); the |w| < 0.1 branch is a series
        # (W+ and W- are O(1): their difference carries an absolute rounding of a few ulp of 1)
        tol = 1e-12 if abs(w) >= 0.1 else 1e-9
        assert abs((wp - wm) - w) <= tol * abs(w) + 4.0 * np.finfo(float).eps * max(wp, 1.0)
        # mirror symmetry W+(w) = W-(-w), bit for bit
        _, wp2, wm2 = po.face_weights(-w)
        assert wp == wm2 and wm == wp2
        # exact Bernoulli form B(w) = w / (e^w - 1)
        if abs(w) >= 1e-3 and abs(w) < 600:
            assert abs(wm - w / math.expm1(w)) <= (5e-12 if abs(w) < 0.1 else 1e-13) * wm


def test_face_weights_zero_and_branch_continuity():
    _, wp, wm = po.face_weights(0.0)
    assert wp == 1.0 and wm == 1.0
    lo = po.face_weights(0.1 - 1e-12)
    hi = po.face_weights(0.1 + 1e-12)
    # the series (1 + w^2/24 + w^4/1920) truncates at O(w^6): ~4e-12 relative at |w| = 0.1
    assert abs(lo[1] - hi[1]) <= 5e-12 * hi[1] and abs(lo[2] - hi[2]) <= 5e-12 * hi[2]


def test_delta_small_w_limit():
    for w in (1e-3, -1e-3, 1e-2):
        d, _, _ = po.face_weights(w)
        assert abs(d - (0.5 - w / 12.0)) <= 1e-8


# ------------------------------------------------------------- CC assembly
def _assembly(n=64, seed=0, finite_loss=False):
    rng = np.random.default_rng(seed)
    o = po.Oracle(n)
    x, dx, (xm1, xp1) = o.ep, o.edp, o.eghost
    dpp = x ** 2 * 10 ** rng.uniform(-6, -3, n)            # C > 0
    radp = (1e-3 * x ** 2 + 1e-2) * rng.uniform(0.2, 2.0, n)  # drift of either sign after -2C/p
    tloss = 10 ** rng.uniform(-1, 1, n) if finite_loss else np.full(n, 1e100)
    dt = 0.05
    a, b, c = po.cc_build(n, dt, x, dx, xm1, xp1, radp, dpp, tloss, radp[0], radp[-1], dpp[0], dpp[-1])
    B = radp - 2.0 / x * dpp
    return dict(n=n, x=x, dx=dx, a=a, b=b, c=c, B=B, C=dpp, dt=dt, tloss=tloss)


def _apply(A, N):
    """(M N)_i = b_i N_i - a_i N_{i-1} - c_i N_{i+1}."""
    out = A["b"] * N
    out[1:] -= A["a"][1:] * N[:-1]
    out[:-1] -= A["c"][:-1] * N[1:]
    return out


def test_cc_signs_and_boundaries():
    A = _assembly()
    assert np.all(A["a"] >= 0.0) and np.all(A["c"] >= 0.0) and np.all(A["b"] >= 1.0)
    assert A["a"][0] == 0.0 and A["c"][-1] == 0.0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_cc_conserves_number_with_zero_flux_edges(seed):
    # telescoping fluxes: sum_i dx_i (M N)_i = sum_i dx_i N_i (1 + dt / T_i)
    A = _assembly(seed=seed, finite_loss=bool(seed % 2))
    rng = np.random.default_rng(seed + 10)
    N = 10 ** rng.uniform(-3, 3, A["n"])
    lhs = np.sum(A["dx"] * _apply(A, N))
    rhs = np.sum(A["dx"] * N * (1.0 + A["dt"] / A["tloss"]))
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs)


def test_cc_discrete_equilibrium_is_stationary():
    # Chang-Cooper's defining property: the spectrum with N_{j+1}/N_j =
    # W-_j / W+_j (= e^{-w_j}) carries zero flux through every face, so one
    # implicit step leaves it unchanged (no losses).  Drift and diffusion are
    # balanced so that |w| spans both weight branches without underflow.
    n = 64
    rng = np.random.default_rng(4)
    o = po.Oracle(n)
    x, dx, (xm1, xp1) = o.ep, o.edp, o.eghost
    C = x ** 2 * 10 ** rng.uniform(-6, -3, n)
    radp = 2.0 / x * C + 4.0 * C / x * rng.uniform(-1.0, 1.0, n)
    a, b, c = po.cc_build(n, 0.05, x, dx, xm1, xp1, radp, C, np.full(n, 1e100), radp[0], radp[-1], C[0], C[-1])
    A = dict(a=a, b=b, c=c)
    B = radp - 2.0 / x * C
    N = np.empty(n)
    N[0] = 1.0
    for j in range(n - 1):
        w = 0.5 * (B[j] + B[j + 1]) / (0.5 * (C[j] + C[j + 1])) * (x[j + 1] - x[j])
        _, wp, wm = po.face_weights(w)
        N[j + 1] = N[j] * wm / wp
    assert N.min() > 1e-250
    ws = [abs(0.5 * (B[j] + B[j + 1]) / (0.5 * (C[j] + C[j + 1])) * (x[j + 1] - x[j])) for j in range(n - 1)]
    assert min(ws) < 0.1 < max(ws)
    got = _apply(A, N)
    assert np.max(np.abs(got - N) / N) <= 1e-10


# ------------------------------------------------------------------ solver
@pytest.mark.parametrize("n", [1, 2, 5, 16])
def test_thomas_matches_dense_solve(n):
    rng = np.random.default_rng(n)
    a, c = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    a[0] = 0.0
    c[-1] = 0.0
    b = 1.0 + a + c + rng.uniform(0, 1, n)
    r = rng.uniform(-1, 1, n)
    M = np.diag(b) - np.diag(a[1:], -1) - np.diag(c[:-1], 1)
    rc, x = po.thomas(a, b, c, r)
    assert rc == 0
    assert np.allclose(x, np.linalg.solve(M, r), rtol=1e-13, atol=1e-15)


def test_pcr_matches_thomas():
    rng = np.random.default_rng(9)
    nsys, n = 6, 64
    a, c = rng.uniform(0, 1, (nsys, n)), rng.uniform(0, 1, (nsys, n))
    a[:, 0] = 0.0
    c[:, -1] = 0.0
    b = 1.0 + a + c + rng.uniform(0, 1, (nsys, n))
    r = rng.uniform(0, 1, (nsys, n))
    rc, xp = po.pcr_batch(a, b, c, r)
    assert rc == 0
    for s in range(nsys):
        _, xt = po.thomas(a[s], b[s], c[s], r[s])
        assert np.allclose(xp[s], xt, rtol=1e-12)


def test_thomas_pivot_guard():
    rc, _ = po.thomas(np.zeros(3), np.array([1.0, 0.0, 1.0]), np.zeros(3), np.ones(3))
    assert rc == -2


# ------------------------------------------------------------------- D_pp
def test_dpp_off_floor_and_model_shape():
    o = po.Oracle(128)
    prm = po.default_params(dpp_model=3)
    d, de, _ = o.dpp(prm, 3, 0.15, 1e6, 1e7, 50.0)
    assert np.allclose(d, o.pp ** 2 / 4.0e5, rtol=1e-15) and np.allclose(de, o.ep ** 2 / 4.0e5, rtol=1e-15)
    prm = po.default_params(dpp_model=1)
    d, de, _ = o.dpp(prm, 1, 0.15, 3e6, 1e7, 50.0)
    # D = p^2 exp(-E/E_max) exp(-1/p) / (4 t_acc): the ratio to the shape is one constant
    shape_p = o.pp ** 2 * np.exp(-o.pE / prm.epmax) * np.exp(-1.0 / o.pp)
    shape_e = o.ep ** 2 * np.exp(-1.0 / o.ep)
    for dd, sh in ((d, shape_p), (de, shape_e)):
        m = sh > 0
        r = dd[m] / sh[m]
        assert np.max(np.abs(r / r[0] - 1.0)) <= 1e-12


def test_tacc_asa_scalings():
    # t_acc ~ L / Ms^3 at fixed c_s, beta (ASA): doubling L doubles it, doubling Ms divides by 8
    f = po.lib().orc_tacc_asa
    t0 = f(3e23, 1e7, 0.3, 50.0, 0.5)
    assert abs(f(6e23, 1e7, 0.3, 50.0, 0.5) / t0 - 2.0) <= 1e-14
    assert abs(f(3e23, 1e7, 0.6, 50.0, 0.5) / t0 - 0.125) <= 1e-14


# ---------------------------------------------------------- evolution
def _small(species=3, nt=8):
    return synth.make_workload(n=64, ntracer=nt, nsnap=2, species=species, dpp_model=1, seed=3, z0=0.02)


def test_zero_state_without_sources_stays_zero():
    w = _small(species=1)
    w.cre0[:] = 0.0
    crp, cre = run_oracle.run_workload(w)
    assert np.all(cre == 0.0) and np.all(crp == 0.0)


def test_thread_count_invariance():
    w = _small()
    a = run_oracle.run_workload(w, nthreads=1)
    b = run_oracle.run_workload(w, nthreads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_tracer_order_invariance():
    # tracers never interact: a permuted batch gives the permuted result
    w = _small(nt=10)
    perm = np.random.default_rng(1).permutation(w.ntracer)
    a = run_oracle.run_workload(w)
    b = run_oracle.run_workload(w.subset(perm))
    n = w.n
    rows = (perm[:, None] * n + np.arange(n)[None, :]).ravel()
    assert np.array_equal(a[1][rows], b[1]) and np.array_equal(a[0][rows], b[0])


# ------------------------------------------- parity-unpinned plausibility
def test_pp_cross_section_threshold_and_positivity():
    f = po.lib().orc_pp_pion_cross_section
    p = np.logspace(-1, 7, 400)
    s = np.array([f(0, v) for v in p])
    assert np.all(np.isfinite(s)) and np.all(s >= 0.0)
    assert s[-1] > 0.0 and s[0] == 0.0  # below the pion threshold nothing


def test_hadronic_kernels_thresholds():
    g = po.lib().orc_gamma_kernel
    nu = po.lib().orc_nu_kernel
    # below the kinetic threshold E_p - m_p < 0.285 GeV the gamma kernel vanishes
    p_lo = math.sqrt((0.938 + 0.2) ** 2 / 0.938 ** 2 - 1.0)
    assert g(0.1, p_lo) == 0.0
    vals = [g(e, 100.0) for e in np.logspace(-1, 1.5, 20)]
    assert all(v >= 0.0 and math.isfinite(v) for v in vals) and max(vals) > 0.0
    assert all(math.isfinite(nu(e, 100.0)) and nu(e, 100.0) >= 0.0 for e in np.logspace(-1, 1.5, 20))


def test_fx_table_asymptotes():
    d = np.load(os.path.join(GOLD, "fx_table.npz"))
    x = 10 ** d["logx"]
    fx = d["fx"]
    small = x < 1e-6
    # F(x) -> 4 pi / (sqrt(3) Gamma(1/3)) (x/2)^{1/3} = 2.1495 x^{1/3} for x -> 0
    c0 = 4.0 * math.pi / (math.sqrt(3.0) * math.gamma(1.0 / 3.0)) * 2.0 ** (-1.0 / 3.0)
    assert np.max(np.abs(fx[small] / (c0 * x[small] ** (1.0 / 3.0)) - 1.0)) < 2e-3
    big = (x > 30) & (x < 300)
    asym = math.sqrt(math.pi / 2.0) * np.sqrt(x[big]) * np.exp(-x[big]) * (1.0 + 55.0 / (72.0 * x[big]))
    assert np.max(np.abs(fx[big] / asym - 1.0)) < 2e-3
