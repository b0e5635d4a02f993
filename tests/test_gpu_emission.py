"""Emission contractions on the device vs the reference / oracle.

Tolerances (stated per mode): FP64 1e-12 relative; FP32 VALU and split-BF16
x3 on MFMA 1e-4 relative (the emissivity bar); one BF16 pass 2e-2 relative
(its 8-bit mantissa, documented as the fast lossy mode).  IC is checked
against the compiled reference's own kernel table and contraction (golden
fixtures); hadronic and synchrotron against the oracle restatement."""
import os

import numpy as np
import pytest

import croma_hip as ch
import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = {ch.EMIS_FP64: 1e-12, ch.EMIS_FP32: 1e-4, ch.EMIS_MFMA_XF32: 1e-4, ch.EMIS_MFMA_BF16: 2e-2}
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(got, want):
    """Max relative error over the entries within 25 decades of their
    tracer's largest emissivity (and above the FP64 denormal range): deeper
    entries are sums over the exponential tails of the spectrum, below what
    a reduced-precision contraction resolves and irrelevant to the emissivity."""
    want = np.atleast_2d(np.asarray(want))
    got = np.asarray(got).reshape(want.shape)
    peak = np.max(np.abs(want), axis=1, keepdims=True)
    m = np.abs(want) > np.maximum(1e-280, 1e-25 * peak)
    return float(np.max(np.abs(got[m] - want[m]) / np.abs(want[m])))


@pytest.fixture(scope="module")
def ctx128():
    c = ch.Context(128, ch.default_params())
    yield c
    c.close()


@pytest.mark.parametrize("mode", list(TOL))
def test_ic_matches_reference_fixture(ctx128, mode):
    d = np.load(os.path.join(GOLD, "ref_n128.npz"))
    nb = d["ic_fic"].size // 128
    eps, _ = ctx128.emission_ic(d["ic_cre"], d["ic_fic"], nb, mode)
    err = rel(eps, d["ic_eps"].reshape(eps.shape))
    print(f"[emission] ic mode={mode} max rel err {err:.3e}", flush=True)
    assert err < TOL[mode]


@pytest.mark.parametrize("mode", list(TOL))
def test_ic_many_tracers(ctx128, mode):
    o = po.Oracle(128)
    rng = np.random.default_rng(3)
    nb = 64
    fic = po.ic_kernel_flat(o, nb, 8.0, 28.0, 0.4)
    nt = 777  # not a multiple of the 64-tracer tile
    q = rng.uniform(2.0, 3.5, nt)
    cut = 10 ** rng.uniform(2, 6, nt)
    cre = (o.ep[None, :] ** -q[:, None]) * np.exp(-o.ep[None, :] / cut[:, None]) * 10 ** rng.uniform(-3, 3, nt)[:, None]
    ref = po.emission_ic(128, cre.ravel(), o.edp, fic, nb).reshape(nt, nb)
    eps, ms = ctx128.emission_ic(cre, fic, nb, mode)
    err = rel(eps, ref)
    print(f"[emission] ic 777x64 mode={mode} max rel err {err:.3e} ({ms:.3f} ms)", flush=True)
    assert err < TOL[mode]


@pytest.mark.parametrize("mode", list(TOL))
def test_hadronic_matches_oracle(ctx128, mode):
    o = po.Oracle(128)
    rng = np.random.default_rng(4)
    nb, emin, emax = 40, -1.0, 5.0
    fk = po.gamma_kernel_flat(o, nb, emin, emax)
    nt = 300
    q = rng.uniform(2.0, 3.0, nt)
    crp = (o.pp[None, :] ** -q[:, None]) * np.exp(-o.pp[None, :] / 1e6)
    ng = 10 ** rng.uniform(-5, -2, nt)
    ref = po.emission_hadronic(128, nb, emin, emax, ng, crp.ravel(), o.beta_p, o.pdp, fk).reshape(nt, nb)
    eps, _ = ctx128.emission_hadronic(crp, ng, fk, nb, emin, emax, mode)
    err = rel(eps, ref)
    print(f"[emission] gamma mode={mode} max rel err {err:.3e}", flush=True)
    assert err < TOL[mode]


@pytest.mark.parametrize("mode", [ch.EMIS_FP64, ch.EMIS_FP32])
def test_synch_matches_oracle(ctx128, mode):
    """FP64 and FP32 modes against the oracle on the compiled reference's
    table; the FP32 mode also on spectra spanning 1e-300..1e+280 in scale
    (per-tracer power-of-two scaling keeps them inside float's range)."""
    d = np.load(os.path.join(GOLD, "ref_n128.npz"))
    nfreq, _, nlogb = (int(v) for v in d["sy_dims"])
    inv_dlogb = float(d["sy_inv_dlogb"][0])
    o = po.Oracle(128)
    rng = np.random.default_rng(5)
    nt = 130
    cre = (o.ep[None, :] ** -2.4) * np.exp(-o.ep[None, :] / 10 ** rng.uniform(3, 5, nt)[:, None])
    cre[5, 70:] = 0.0  # truncated spectrum: the sum stops at the first empty successor
    cre[7] = 0.0  # an empty spectrum
    cre[9] *= 1e-300
    cre[11] *= 1e280
    b = 10 ** rng.uniform(-8.5, -4.5, nt)
    ref = po.emission_synch(128, nfreq, nlogb, -6.0, inv_dlogb, d["sy_table"], cre.ravel(), o.edp, b)
    eps, ms = ctx128.emission_synch(cre, b, d["sy_table"], nfreq, nlogb, -6.0, inv_dlogb, mode=mode)
    assert np.all(eps[7] == 0.0)
    err = rel(eps, ref.reshape(eps.shape))
    print(f"[emission] synch mode={mode} max rel err {err:.3e} ({ms:.3f} ms)", flush=True)
    assert err < TOL[mode]


@pytest.mark.parametrize("n", [64, 256, 512])
def test_synch_fp32_matches_fp64_on_device_tables(n):
    """FP32 against FP64 at every spectrum length, on a device-built table
    (32 log-B planes x 64 frequencies) and 2000 tracers: 1e-4."""
    fx = np.load(os.path.join(GOLD, "fx_table.npz"))
    o = po.Oracle(n)
    logy, pw = po.synch_logy(o, 64, 7.0, 11.0, 16)
    rng = np.random.default_rng(17)
    nt = 2000
    q = rng.uniform(2.0, 3.5, nt)
    cre = (o.ep[None, :] ** -q[:, None]) * np.exp(-o.ep[None, :] / 10 ** rng.uniform(2, 6, nt)[:, None])
    cre *= 10 ** rng.uniform(-5, 5, nt)[:, None]
    b = 10 ** rng.uniform(-8.0, -5.0, nt)
    ctx = ch.Context(n, ch.default_params())
    try:
        tab, _ = ctx.synch_kernel_table(64, 16, 32, -2.0, 8.0, fx["logx"], fx["logfx"], fx["fx"],
                                        float(fx["xmin"][0]), logy, pw)
        e64, _ = ctx.emission_synch(cre, b, tab, 64, 32, -2.0, 8.0, mode=ch.EMIS_FP64)
        e32, ms = ctx.emission_synch(cre, b, tab, 64, 32, -2.0, 8.0, mode=ch.EMIS_FP32)
    finally:
        ctx.close()
    err = rel(e32, e64)
    print(f"[emission] synch n={n} FP32 vs FP64 max rel err {err:.3e} ({ms:.3f} ms)", flush=True)
    assert err < 1e-4


@pytest.mark.parametrize("n", [64, 256, 512])
@pytest.mark.parametrize("mode", [ch.EMIS_MFMA_XF32, ch.EMIS_MFMA_BF16, ch.EMIS_FP64])
def test_ic_grid_sizes(n, mode):
    """Every spectrum length: K = 64..256 run the streaming MFMA kernel (the
    kernel table resident in LDS), K = 512 the LDS-tiled one."""
    o = po.Oracle(n)
    rng = np.random.default_rng(6)
    nb = 64
    fic = po.ic_kernel_flat(o, nb, 8.0, 28.0, 0.4)
    nt = 1000 + 13
    q = rng.uniform(2.0, 3.5, nt)
    cut = 10 ** rng.uniform(2, 6, nt)
    cre = (o.ep[None, :] ** -q[:, None]) * np.exp(-o.ep[None, :] / cut[:, None]) * 10 ** rng.uniform(-3, 3, nt)[:, None]
    cre[7] = 0.0  # an empty spectrum
    ref = po.emission_ic(n, cre.ravel(), o.edp, fic, nb).reshape(nt, nb)
    ctx = ch.Context(n, ch.default_params())
    try:
        eps, ms = ctx.emission_ic(cre, fic, nb, mode)
    finally:
        ctx.close()
    assert np.all(eps[7] == 0.0)
    err = rel(eps, ref)
    print(f"[emission] ic n={n} mode={mode} max rel err {err:.3e} ({ms:.3f} ms)", flush=True)
    assert err < TOL[mode]


@pytest.mark.parametrize("n", [64, 128, 256, 512])
def test_synch_table_on_device_matches_reference(n):
    """Device pitch-kernel table vs the compiled reference's table (fixture)
    and, at a larger shape, vs the oracle: 1e-13 relative (pow of the device
    libm against the host's, summed over the pitch angles)."""
    d = np.load(os.path.join(GOLD, f"ref_n{n}.npz"))
    fx = np.load(os.path.join(GOLD, "fx_table.npz"))
    nfreq, nth, nlogb = (int(v) for v in d["sy_dims"])
    inv = float(d["sy_inv_dlogb"][0])
    pmin, pmax, pemin, pemax = (float(v) for v in d["grid_bounds"])
    ctx = ch.Context(n, ch.default_params(), pmin=pmin, pmax=pmax, pemin=pemin, pemax=pemax)
    try:
        tab, ms = ctx.synch_kernel_table(nfreq, nth, nlogb, -6.0, inv, fx["logx"], fx["logfx"], fx["fx"],
                                         float(fx["xmin"][0]), d["sy_logy"], d["sy_pw"])
        ref = d["sy_table"]
        m = ref != 0
        assert np.array_equal(tab[~m], ref[~m])
        err = float(np.max(np.abs(tab[m] - ref[m]) / np.abs(ref[m]))) if m.any() else 0.0
        print(f"[emission] synch table n={n} vs reference: max rel err {err:.3e}", flush=True)
        assert err < 1e-13
        # a larger table: 32 log-B planes x 64 frequencies x 24 pitch angles
        o = po.Oracle(n, pmin, pmax, pemin, pemax)
        logy, pw = po.synch_logy(o, 64, 7.0, 11.0, 24)
        big_ref = po.synch_kernel_table(n, 64, 24, 32, -2.0, 8.0, fx["logx"], fx["logfx"], fx["fx"],
                                        float(fx["xmin"][0]), logy, pw)
        big, ms = ctx.synch_kernel_table(64, 24, 32, -2.0, 8.0, fx["logx"], fx["logfx"], fx["fx"],
                                         float(fx["xmin"][0]), logy, pw)
    finally:
        ctx.close()
    m = big_ref != 0
    assert np.array_equal(big[~m], big_ref[~m])
    err = float(np.max(np.abs(big[m] - big_ref[m]) / np.abs(big_ref[m])))
    print(f"  32x64x{n} table: max rel err {err:.3e} ({ms:.3f} ms)", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("n", [64, 128, 256])
def test_synch_mfma_matches_oracle(n):
    """Synchrotron on MFMA (Chebyshev expansion of the log-B interpolation,
    plane-grouped split-BF16 x3 contraction) against the oracle's FP64 sum on
    a device-built table of 512 planes over 6 decades in B: 1e-4 over the
    entries within 25 decades of each tracer's peak; truncated, empty and
    extreme-scale spectra included."""
    fx = np.load(os.path.join(GOLD, "fx_table.npz"))
    o = po.Oracle(n)
    nfreq, ntheta, nlogb, logb_min, inv = 64, 16, 512, -3.0, 512 / 6.0
    logy, pw = po.synch_logy(o, nfreq, 7.0, 11.0, ntheta)
    rng = np.random.default_rng(29)
    nt = 3000
    q = rng.uniform(2.0, 3.5, nt)
    cre = (o.ep[None, :] ** -q[:, None]) * np.exp(-o.ep[None, :] / 10 ** rng.uniform(2, 6, nt)[:, None])
    cre *= 10 ** rng.uniform(-5, 5, nt)[:, None]
    cre[5, n // 2:] = 0.0   # truncated: the sum stops at the first empty successor
    cre[7] = 0.0            # empty
    cre[9] *= 1e-280
    cre[11] *= 1e250
    b = 10 ** rng.uniform(-8.0, -5.0, nt)
    ctx = ch.Context(n, ch.default_params())
    try:
        tab, _ = ctx.synch_kernel_table(nfreq, ntheta, nlogb, logb_min, inv, fx["logx"], fx["logfx"], fx["fx"],
                                        float(fx["xmin"][0]), logy, pw)
        eps, ms = ctx.emission_synch(cre, b, tab, nfreq, nlogb, logb_min, inv, mode=ch.EMIS_MFMA_XF32)
    finally:
        ctx.close()
    sample = np.r_[np.arange(16), np.sort(np.random.default_rng(30).choice(np.arange(16, nt), 400, replace=False))]
    ref = po.emission_synch(n, nfreq, nlogb, logb_min, inv, tab, cre[sample].ravel(), o.edp,
                            b[sample]).reshape(len(sample), nfreq)
    assert np.all(eps[7] == 0.0)
    err = rel(eps[sample], ref)
    print(f"[emission] synch MFMA n={n} max rel err {err:.3e} ({ms:.3f} ms)", flush=True)
    assert err < 1e-4


def test_synch_plan_reuse_matches_one_shot():
    """A plan built once serves calls of different sizes (its buffers grow,
    then shrink to a prefix) with results identical to the one-shot mode."""
    fx = np.load(os.path.join(GOLD, "fx_table.npz"))
    n = 128
    o = po.Oracle(n)
    nfreq, ntheta, nlogb, logb_min, inv = 32, 12, 256, -3.0, 256 / 4.0
    logy, pw = po.synch_logy(o, nfreq, 7.0, 11.0, ntheta)
    rng = np.random.default_rng(41)
    nt = 2000
    q = rng.uniform(2.0, 3.5, nt)
    cre = (o.ep[None, :] ** -q[:, None]) * np.exp(-o.ep[None, :] / 10 ** rng.uniform(2, 6, nt)[:, None])
    b = 10 ** rng.uniform(-7.5, -5.5, nt)
    ctx = ch.Context(n, ch.default_params())
    try:
        tab, _ = ctx.synch_kernel_table(nfreq, ntheta, nlogb, logb_min, inv, fx["logx"], fx["logfx"], fx["fx"],
                                        float(fx["xmin"][0]), logy, pw)
        one, _ = ctx.emission_synch(cre, b, tab, nfreq, nlogb, logb_min, inv, mode=ch.EMIS_MFMA_XF32)
        plan = ch.SynchPlan(ctx, tab, nfreq, nlogb, logb_min, inv)
        try:
            small, _ = plan.run(cre[:300], b[:300])
            big, ms = plan.run(cre, b)
            again, _ = plan.run(cre[:300], b[:300])
        finally:
            plan.close()
    finally:
        ctx.close()
    print(f"[emission] synch plan: prepare {plan.prepare_ms:.3f} ms, run {ms:.3f} ms", flush=True)
    assert np.array_equal(big, one)
    assert np.array_equal(small, one[:300]) and np.array_equal(again, one[:300])


def test_synch_large_batch_equals_its_subsets():
    """300,000 tracers: the plane sort then gives each block of its count and
    place kernels a chunk of several hundred tracers.  Every tracer's
    emissivities depend on its own spectrum and field only, so the large
    batch equals a 1,500-tracer subset run alone bit for bit, in the FP64
    and the MFMA modes (each tracer listed once, none lost), and the FP64
    subset equals the oracle."""
    fx = np.load(os.path.join(GOLD, "fx_table.npz"))
    n = 64
    o = po.Oracle(n)
    nfreq, ntheta, nlogb, logb_min, inv = 32, 8, 512, -3.0, 512 / 6.0
    logy, pw = po.synch_logy(o, nfreq, 7.0, 11.0, ntheta)
    rng = np.random.default_rng(53)
    nt = 300_000
    q = rng.uniform(2.0, 3.5, nt)
    cre = (o.ep[None, :] ** -q[:, None]) * np.exp(-o.ep[None, :] / 10 ** rng.uniform(2, 6, nt)[:, None])
    b = 10 ** rng.uniform(-8.0, -5.0, nt)
    idx = np.sort(rng.choice(nt, 1500, replace=False))
    ctx = ch.Context(n, ch.default_params())
    try:
        tab, _ = ctx.synch_kernel_table(nfreq, ntheta, nlogb, logb_min, inv, fx["logx"], fx["logfx"], fx["fx"],
                                        float(fx["xmin"][0]), logy, pw)
        big64, _ = ctx.emission_synch(cre, b, tab, nfreq, nlogb, logb_min, inv, mode=ch.EMIS_FP64)
        sub64, _ = ctx.emission_synch(cre[idx], b[idx], tab, nfreq, nlogb, logb_min, inv, mode=ch.EMIS_FP64)
        plan = ch.SynchPlan(ctx, tab, nfreq, nlogb, logb_min, inv)
        try:
            bigm, ms = plan.run(cre, b)
            subm, _ = plan.run(cre[idx], b[idx])
        finally:
            plan.close()
    finally:
        ctx.close()
    print(f"[emission] synch MFMA run on {nt} tracers x {n} bins: {ms:.3f} ms", flush=True)
    assert np.array_equal(big64[idx], sub64)
    assert np.array_equal(bigm[idx], subm)
    assert np.all(np.isfinite(bigm))
    ref = po.emission_synch(n, nfreq, nlogb, logb_min, inv, tab, cre[idx[:40]].ravel(), o.edp,
                            b[idx[:40]]).reshape(40, nfreq)
    assert rel(sub64[:40], ref) < 1e-12


def test_synch_mfma_refuses_coarse_planes():
    """Planes coarser than 0.02 in log10 B leave the Chebyshev truncation
    unbounded: the MFMA mode refuses them (the FP32 VALU mode serves them)."""
    fx = np.load(os.path.join(GOLD, "fx_table.npz"))
    o = po.Oracle(128)
    logy, pw = po.synch_logy(o, 16, 7.0, 11.0, 8)
    ctx = ch.Context(128, ch.default_params())
    try:
        tab, _ = ctx.synch_kernel_table(16, 8, 32, -3.0, 32 / 6.0, fx["logx"], fx["logfx"], fx["fx"],
                                        float(fx["xmin"][0]), logy, pw)
        with pytest.raises(ch.CromaHipError) as ei:
            ctx.emission_synch(np.ones((4, 128)), np.full(4, 1e-6), tab, 16, 32, -3.0, 32 / 6.0,
                               mode=ch.EMIS_MFMA_XF32)
        assert ei.value.rc == ch.EUNSUPPORTED
    finally:
        ctx.close()
