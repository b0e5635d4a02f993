"""The benchmark configurations on the device against the CPU oracle.

- c0 at full size: 1,024 electron tracers x 128 bins over all 100
  snapshots z = 1 -> 0 (per-step rounding accumulates over the horizon),
  fused and plain paths: |hip - cpu| <= 1e-30 + 1e-10 |cpu| (the reference
  comparator, analysis/compare_tracer_outputs.py:126-127).
- The direct acceleration-time D_pp model (dpp_model 4, t_acc switching at
  t_off: FP_Coef.c:759-849, fp_cuda_tracer_backend.cu:549-551) and per-tracer
  disable_adiabatic flags (fp_shared_core.c:1360, 1427), both species.
- c3 emission: 1,000,000 tracers x 256 bins -> 64 bins with the real
  kernels (IC at a redshift drawn from U(0, 1), pp gamma, pp neutrino,
  synchrotron on a device-built 512-plane table) in every precision mode;
  2,000 sampled tracers against the oracle's contractions.
- c4 serving: a Poisson stream of chunks whose sizes span 256..65,536
  tracers, mixed grids, species and D_pp, 10 snapshots each; every chunk
  checked on a tracer sample against the oracle, and the largest chunk
  bit-identical alone and in the mix.
- DSA injection with reacceleration (Drury83 with its disable-adiabatic
  flags, positive delta) feeding a two-species batch.
"""
import os
import time

import numpy as np
import pytest

import croma_hip as ch
import pyoracle as po
import run_oracle
import serve_stream
import synth
from workloads import hip_params, rel_err, run_hip, small_workload

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EMIS_TOL = {ch.EMIS_FP64: 1e-12, ch.EMIS_FP32: 1e-4, ch.EMIS_MFMA_XF32: 1e-4, ch.EMIS_MFMA_BF16: 2e-2}


@pytest.fixture(scope="module")
def c0():
    w = synth.make_workload(seed=2026, **synth.CONFIGS["c0"])
    print(f"[c0] oracle: {w.ntracer} tracers x {w.n} bins x {w.nsnap} snapshots", flush=True)
    t0 = time.perf_counter()
    crp_ref, cre_ref = run_oracle.run_workload(w, run_oracle.oracle_params(w), nthreads=16)
    print(f"[c0] oracle {time.perf_counter() - t0:.1f} s", flush=True)
    return w, cre_ref, crp_ref


@pytest.mark.parametrize("path", [ch.PATH_FUSED, ch.PATH_PLAIN], ids=["fused", "plain"])
def test_c0_full_horizon(c0, path):
    w, cre_ref, crp_ref = c0
    print(f"[c0] device path={path}", flush=True)
    ctx = ch.Context(w.n, hip_params(w))
    try:
        crp, cre, t = run_hip(w, ctx, path=path)
    finally:
        ctx.close()
    err, ok = rel_err(cre, cre_ref)
    print(f"  device substeps {t.cell_steps}; max rel err {err:.3e}", flush=True)
    assert ok, f"max rel err {err:.3e}"
    assert np.all(crp == 0.0) and np.all(crp_ref == 0.0)


DIRECT = dict(dpp_model=4, t_acc_direct_gyr=0.35, t_off_gyr=0.2, t_acc_off_gyr=1.5)


@pytest.mark.parametrize("species", [1, 3], ids=["e", "ep"])
@pytest.mark.parametrize("case", ["direct", "dis", "direct+dis"])
@pytest.mark.parametrize("path", [ch.PATH_FUSED, ch.PATH_PLAIN], ids=["fused", "plain"])
def test_dpp_direct_and_disable_adiabatic(species, case, path):
    n = 128
    w = small_workload(n, 96, 3, species, 1, seed=41, dsa_frac=0.2 if species == 3 else 0.0)
    extra = dict(DIRECT) if "direct" in case else {}
    dis = None
    if "dis" in case:
        dis = (np.random.default_rng(42).random(w.ntracer) < 0.4).astype(np.uint8)
    print(f"[dpp/dis] species={species} case={case} path={path}", flush=True)
    crp_ref, cre_ref = run_oracle.run_workload(w, run_oracle.oracle_params(w, **extra), dis=dis)
    ctx = ch.Context(n, hip_params(w, **extra))
    try:
        crp, cre, _ = run_hip(w, ctx, path=path, dis=dis)
    finally:
        ctx.close()
    err, ok = rel_err(cre, cre_ref)
    print(f"  electron max rel err {err:.3e}", flush=True)
    assert ok
    if species == 3:
        perr, pok = rel_err(crp, crp_ref)
        print(f"  proton max rel err {perr:.3e}", flush=True)
        assert pok
    # the flags and the direct model change the answer (the case is not vacuous)
    crp0, cre0 = run_oracle.run_workload(w, run_oracle.oracle_params(w))
    assert not np.array_equal(cre0, cre_ref)


def _emis_rel(got, want):
    """Max relative error over entries within 25 decades of their tracer's
    largest emissivity (deeper entries are sums over exponential tails)."""
    peak = np.max(np.abs(want), axis=1, keepdims=True)
    m = np.abs(want) > np.maximum(1e-280, 1e-25 * peak)
    return float(np.max(np.abs(got[m] - want[m]) / np.abs(want[m])))


@pytest.fixture(scope="module")
def c3():
    """The c3 spectra: 1e6 tracers x 256 bins, power laws q ~ U(2, 3.5) with
    log-uniform cutoffs on the electron and proton grids."""
    n, nt = 256, 1_000_000
    rng = np.random.default_rng(2026)
    o = po.Oracle(n)
    q = rng.uniform(2.0, 3.5, nt)
    cut_e = 10.0 ** rng.uniform(2.0, 6.0, nt)
    cut_p = 10.0 ** rng.uniform(3.0, 8.0, nt)
    cre = np.exp(-q[:, None] * np.log(o.ep)[None, :] - o.ep[None, :] / cut_e[:, None])
    crp = np.exp(-q[:, None] * np.log(o.pp)[None, :] - o.pp[None, :] / cut_p[:, None])
    ng = 10.0 ** rng.uniform(-5.0, -2.0, nt)
    b = 10.0 ** rng.uniform(-8.0, -5.0, nt)
    sample = np.sort(rng.choice(nt, 2000, replace=False))
    ctx = ch.Context(n, ch.default_params())
    yield dict(n=n, o=o, cre=cre, crp=crp, ng=ng, b=b, sample=sample, ctx=ctx, z=float(rng.uniform(0.0, 1.0)))
    ctx.close()


@pytest.mark.parametrize("mode", list(EMIS_TOL))
@pytest.mark.parametrize("kind", ["ic", "gamma", "nu"])
def test_c3_emission_sampled(c3, kind, mode):
    n, o, s, nb = c3["n"], c3["o"], c3["sample"], 64
    if kind == "ic":
        fk = po.ic_kernel_flat(o, nb, 8.0, 28.0, c3["z"])
        eps, ms = c3["ctx"].emission_ic(c3["cre"], fk, nb, mode)
        ref = po.emission_ic(n, c3["cre"][s].ravel(), o.edp, fk, nb).reshape(len(s), nb)
    else:
        emin, emax = -1.0, 5.0
        fk = (po.gamma_kernel_flat if kind == "gamma" else po.nu_kernel_flat)(o, nb, emin, emax)
        eps, ms = c3["ctx"].emission_hadronic(c3["crp"], c3["ng"], fk, nb, emin, emax, mode)
        ref = po.emission_hadronic(n, nb, emin, emax, c3["ng"][s], c3["crp"][s].ravel(), o.beta_p, o.pdp,
                                   fk).reshape(len(s), nb)
    err = _emis_rel(eps[s], ref)
    print(f"[c3] {kind} mode={mode} 1e6 x 256 -> 64: {ms:.3f} ms, sampled max rel err {err:.3e}", flush=True)
    assert np.isfinite(eps).all()
    assert err < EMIS_TOL[mode]


@pytest.mark.parametrize("mode", [ch.EMIS_FP64, ch.EMIS_FP32, ch.EMIS_MFMA_XF32])
def test_c3_synch_sampled(c3, mode):
    n, o, s, nfreq = c3["n"], c3["o"], c3["sample"], 64
    nlogb, logb_min, inv_dlogb, ntheta = 512, -3.0, 512 / 6.0, 32
    fx = np.load(os.path.join(GOLD, "fx_table.npz"))
    logy, pw = po.synch_logy(o, nfreq, 7.0, 11.0, ntheta)
    tab, _ = c3["ctx"].synch_kernel_table(nfreq, ntheta, nlogb, logb_min, inv_dlogb, fx["logx"], fx["logfx"],
                                          fx["fx"], float(fx["xmin"][0]), logy, pw)
    eps, ms = c3["ctx"].emission_synch(c3["cre"], c3["b"], tab, nfreq, nlogb, logb_min, inv_dlogb, mode=mode)
    ref = po.emission_synch(n, nfreq, nlogb, logb_min, inv_dlogb, tab, c3["cre"][s].ravel(), o.edp,
                            c3["b"][s]).reshape(len(s), nfreq)
    err = _emis_rel(eps[s], ref)
    print(f"[c3] synch mode={mode} 1e6 x 256 -> 64: {ms:.3f} ms, sampled max rel err {err:.3e}", flush=True)
    assert np.isfinite(eps).all()
    assert err < EMIS_TOL[mode]


def test_c4_stream_chunks_match_oracle():
    specs = serve_stream.chunk_specs(10, seed=77)
    specs[0]["ntracer"] = 65536
    specs[1]["ntracer"] = 256
    rng = np.random.default_rng(78)
    srv = ch.Server(48 * 2 ** 30, serve_stream.server_params())
    try:
        ids = []
        for sp in specs:
            print(f"[c4] submit {sp}", flush=True)
            ids.append(srv.submit(**serve_stream.make_chunk(sp)[1]))
            time.sleep(float(rng.exponential(0.02)))
        results = [srv.wait(i) for i in ids]
        print("[c4] largest chunk alone", flush=True)
        alone = srv.wait(srv.submit(**serve_stream.make_chunk(specs[0])[1]))
    finally:
        srv.close()
    assert np.array_equal(alone[0], results[0][0]) and np.array_equal(alone[1], results[0][1])
    for sp, (crp, cre, st) in zip(specs, results):
        w, _ = serve_stream.make_chunk(sp)
        idx = np.sort(np.random.default_rng(sp["seed"]).choice(w.ntracer, min(16, w.ntracer), replace=False))
        sub = w.subset(idx)
        crp_ref, cre_ref = run_oracle.run_workload(sub, run_oracle.oracle_params(sub), s0=0, s1=sp["nsnap"])
        rows = (idx[:, None] * w.n + np.arange(w.n)[None, :]).ravel()
        err, ok = rel_err(cre[rows], cre_ref)
        msg = f"  chunk n={w.n} ntr={w.ntracer} species={sp['species']} dpp={sp['dpp']}: electron {err:.3e}"
        if sp["species"] != ch.SPECIES_ELECTRON:
            perr, pok = rel_err(crp[rows], crp_ref)
            msg += f", proton {perr:.3e}"
            ok = ok and pok
        print(msg + f", latency {1e3 * st['latency_s']:.1f} ms", flush=True)
        assert ok


@pytest.mark.parametrize("mode", ["drury83", "positive_delta"])
def test_dsa_reacceleration_feeds_the_solve(mode):
    """DSA injection with reacceleration of the pre-shock spectra (state
    mode: the fresh spectra and the reacceleration deltas added to crp / cre;
    Drury83 also sets each shocked tracer's disable-adiabatic flag, the
    reference driver's dis input, include/tracer_fp_solve.inc:80-94) feeding
    a two-species batch: HIP against the oracle on the same injected state
    and flags, at the reference comparator's bar."""
    n = 128
    w = small_workload(n, 96, 2, 3, 1, seed=47)
    rng = np.random.default_rng(48)
    nt = w.ntracer
    mach = np.where(rng.random(nt) < 0.5, 10 ** rng.uniform(np.log10(2.0), 1.0, nt), np.nan)
    temp = 10 ** rng.uniform(6.0, 8.0, nt)
    rho = 10 ** rng.uniform(-29.0, -26.0, nt)
    if mode == "drury83":
        prm = ch.dsa_default_params(reacc_mode=ch.DSA_REACC_DRURY83, reacc_eta_cap=1e-3)
    else:
        prm = ch.dsa_default_params(reacc_mode=ch.DSA_REACC_POSITIVE_DELTA)
    crp, cre = w.crp0.copy(), w.cre0.copy()
    # the state carries a CR population before the shock (the positive
    # delta's budget grows with the proton pressure)
    crp *= 1e3
    dis = np.zeros(nt, np.uint8)
    ninj = ch.dsa_inject(n, prm, ch.DSA_STATE, mach, rho, w.mass, temp1=temp, crp=crp, cre=cre, dis=dis)
    print(f"[dsa reacc] mode={mode} injected {ninj} of {nt}, flags {int(dis.sum())}", flush=True)
    assert ninj > 0
    assert (int(dis.sum()) > 0) == (mode == "drury83")
    w.crp0, w.cre0 = crp, cre
    crp_ref, cre_ref = run_oracle.run_workload(w, run_oracle.oracle_params(w), dis=dis)
    ctx = ch.Context(n, hip_params(w))
    try:
        crp_h, cre_h, _ = run_hip(w, ctx, dis=dis)
    finally:
        ctx.close()
    err, ok = rel_err(cre_h, cre_ref)
    perr, pok = rel_err(crp_h, crp_ref)
    print(f"  electron max rel err {err:.3e}, proton {perr:.3e}", flush=True)
    assert ok and pok
