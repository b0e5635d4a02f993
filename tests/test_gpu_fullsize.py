"""Parity at the full benchmark sizes.

The whole workload runs on the device; the oracle then checks what it can
afford at that size:
- a seeded sample of tracers, rerun by the oracle on the same inputs
  (|hip - cpu| <= 1e-30 + 1e-10 |cpu|, analysis/compare_tracer_outputs.py:126-127);
- size-independent properties of every tracer: finite, non-negative spectra
  (the Chang-Cooper update preserves positivity), and the tracer-substep
  count equal to the sum of the oracle's integer nsub over all tracers."""
import numpy as np
import pytest

import croma_hip as ch
import pyoracle as po
import run_oracle
import synth
from workloads import hip_params, rel_err

pytestmark = pytest.mark.gpu


def _run_full(w, s1, tile):
    ctx = ch.Context(w.n, hip_params(w))
    ctx.set_tile(tile[0], tile[1], tile[2], secondary_gemm=tile[3])
    b = ch.Batch(ctx, w.ntracer, w.species, w.nsnap)
    try:
        b.set_state(w.crp0 if w.protons else None, w.cre0, mass=w.mass)
        for s in range(w.nsnap):
            b.set_background(s, w.bg[s])
            if s in w.qpi or s in w.qepri:
                b.set_sources(s, w.qpi.get(s), w.qepri.get(s))
        t = b.run(0, s1, w.nsnap, w.z, w.dt_snap)
        crp, cre = b.get_state()
    finally:
        b.close()
        ctx.close()
    return crp, cre, t


def _oracle_nsub_total(w, s1):
    o = po.Oracle(w.n, w.pmin, w.pmax, w.pemin, w.pemax)
    prm = run_oracle.oracle_params(w)
    total = 0
    for s in range(s1):
        row = o.ic_cooling_row(w.z[s])[0]
        bg = w.bg[s]
        ns, _, _, _ = o.nsub(prm, w.dt_snap[s], bg[0], bg[1], bg[2], bg[3], row[0], row[-1], bg[4], bg[5])
        total += int(ns.sum())
    return total


# the bench's tiles: rows per lane, waves per SIMD, snapshots per task, split GEMM
@pytest.mark.parametrize("name,s1,nsample,tile", [("c2", 2, 48, (16, 2, 2, 0)), ("c1", 3, 48, (8, 2, 2, 1))])
def test_full_size_sample_and_invariants(name, s1, nsample, tile):
    w = synth.make_workload(seed=2026, **synth.CONFIGS[name])
    print(f"[fullsize] {name}: {w.ntracer} tracers x {w.n} bins, snapshots 0..{s1 - 1}, tile {tile}", flush=True)
    crp, cre, t = _run_full(w, s1, tile)
    print(f"  device done: {t.cell_steps} tracer-substeps", flush=True)
    assert t.cell_steps == _oracle_nsub_total(w, s1)
    for a in (cre, crp):
        assert np.isfinite(a).all()
        assert (a >= 0.0).all()
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(w.ntracer, nsample, replace=False))
    sub = w.subset(idx)
    crp_ref, cre_ref = run_oracle.run_workload(sub, run_oracle.oracle_params(sub), s0=0, s1=s1)
    n = w.n
    rows = (idx[:, None] * n + np.arange(n)[None, :]).ravel()
    e_err, e_ok = rel_err(cre[rows], cre_ref)
    print(f"  sample of {nsample}: electron max rel err {e_err:.3e}", flush=True)
    assert e_ok
    if w.protons:
        p_err, p_ok = rel_err(crp[rows], crp_ref)
        print(f"  proton max rel err {p_err:.3e}", flush=True)
        assert p_ok
