"""Launch graphs and the snapshot-level drop-in.

1. A snapshot window whose launch shape repeats is replayed from a captured
   hipGraph.  The same windows run with graphs on and off must give
   bit-identical spectra, on every replay, with host copies and
   synchronisations of the program's own between the replays.
2. Pipelined split windows (the electron pass of a window beside the next
   window's proton pass, on a second stream) give the same spectra as the
   passes run in turn, graphs on and off, over runs of several windows.
3. croma_hip_execute_snapshot (the parameter list of
   tracer_fp_batch_execute_parallel, tracer_fp_batch.c:829-876) called per
   snapshot through the C ABI, as the reference driver calls it
   (include/tracer_fp_solve.inc:631-673), against the oracle's same entry
   point on the same inputs: |hip - cpu| <= 1e-30 + 1e-10 |cpu|
   (analysis/compare_tracer_outputs.py:126-127)."""
import numpy as np
import pytest

import croma_hip as ch
import pyoracle as po
import run_oracle
from workloads import hip_params, rel_err, small_workload

pytestmark = pytest.mark.gpu

GRAPH_CASES = [
    # n, ntracer, nsnap, species, dpp, dsa, rows, waves, secondary_gemm
    (128, 96, 4, 3, 1, 0.3, 4, 2, 1),
    (128, 96, 4, 3, 3, 0.0, 8, 1, 0),
    (256, 64, 3, 1, 2, 0.0, 8, 2, 0),
]


def _batch(ctx, w):
    b = ch.Batch(ctx, w.ntracer, w.species, w.nsnap)
    b.set_state(w.crp0 if w.protons else None, w.cre0, mass=w.mass)
    for s in range(w.nsnap):
        b.set_background(s, w.bg[s])
        if s in w.qpi or s in w.qepri:
            b.set_sources(s, w.qpi.get(s), w.qepri.get(s))
    return b


@pytest.mark.parametrize("graphs", [True, False])
def test_pipelined_split_windows_equal_serial(graphs):
    n, nt, ns = 128, 96, 8
    w = small_workload(n, nt, ns, 3, 1, dsa_frac=0.3, seed=29)
    prm = hip_params(w)
    cp, cs = ch.Context(n, prm), ch.Context(n, prm)
    cp.set_pipeline(True)
    for c in (cp, cs):
        c.set_graphs(graphs)
        c.set_tile(8, 2, 1, secondary_gemm=1)
    bp, bs = _batch(cp, w), _batch(cs, w)
    try:
        el = 0.0
        # windows 0..3, then 3 single-window runs, then 4..7 in one run
        for s0, s1 in ((0, 3), (3, 4), (4, 8)):
            print(f"[pipeline] graphs={graphs} snapshots {s0}..{s1 - 1}", flush=True)
            tp = bp.run(s0, s1, ns, w.z, w.dt_snap, elapsed=el)
            ts = bs.run(s0, s1, ns, w.z, w.dt_snap, elapsed=el)
            assert tp.cell_steps == ts.cell_steps
            for a, b in zip(bp.get_state(), bs.get_state()):
                assert np.array_equal(a, b), f"snapshots {s0}..{s1 - 1}"
            el += float(np.sum(w.dt_snap[s0:s1]))
    finally:
        for o in (bp, bs, cp, cs):
            o.close()


@pytest.mark.parametrize("case", GRAPH_CASES, ids=[f"n{c[0]}-sp{c[3]}-gemm{c[8]}" for c in GRAPH_CASES])
def test_graph_replays_equal_plain_launches(case):
    n, nt, ns, sp, dpp, dsa, r, occ, gemm = case
    w = small_workload(n, nt, ns, sp, dpp, dsa_frac=dsa, seed=23)
    prm = hip_params(w)
    cg, ce = ch.Context(n, prm), ch.Context(n, prm)
    ce.set_graphs(False)
    for c in (cg, ce):
        c.set_tile(r, occ, 1, secondary_gemm=gemm)
    bg, be = _batch(cg, w), _batch(ce, w)
    launches = captures = failures = 0
    try:
        el = 0.0
        for s in range(ns):
            # the same snapshot three times: plain launches, capture + launch, replay
            for rep in range(3):
                print(f"[graph] n={n} species={sp} snapshot={s} rep={rep}", flush=True)
                tg = bg.run(s, s + 1, ns, w.z, w.dt_snap, elapsed=el)
                te = be.run(s, s + 1, ns, w.z, w.dt_snap, elapsed=el)
                assert te.graph_launches == 0 and te.graph_captures == 0
                launches += tg.graph_launches
                captures += tg.graph_captures
                failures += tg.graph_failures
                crp_g, cre_g = bg.get_state()
                crp_e, cre_e = be.get_state()
                assert np.isfinite(cre_g).all() and np.isfinite(crp_g).all()
                assert np.array_equal(cre_g, cre_e), f"electron spectra differ at snapshot {s} rep {rep}"
                assert np.array_equal(crp_g, crp_e), f"proton spectra differ at snapshot {s} rep {rep}"
            el += float(w.dt_snap[s])
    finally:
        bg.close()
        be.close()
        cg.close()
        ce.close()
    print(f"  graph launches {launches}, captures {captures}, refused {failures}", flush=True)
    assert failures == 0
    assert captures >= 1 and launches >= ns


DROPIN_CASES = [
    # n, ntracer, nsnap, species, dpp, dsa
    (128, 64, 3, 1, 1, 0.0),
    (256, 48, 3, 3, 1, 0.3),
]


@pytest.mark.parametrize("case", DROPIN_CASES, ids=[f"n{c[0]}-sp{c[3]}" for c in DROPIN_CASES])
def test_execute_snapshot_matches_oracle(case):
    n, nt, ns, sp, dpp, dsa = case
    w = small_workload(n, nt, ns, sp, dpp, dsa_frac=dsa, seed=29)
    o = po.Oracle(w.n, w.pmin, w.pmax, w.pemin, w.pemax)
    prm_o = run_oracle.oracle_params(w)
    ctx = ch.Context(n, hip_params(w))
    fqe, npmin = (o.secondary_kernel() if w.protons else (np.zeros(n * n), np.full(n, n, np.int32)))
    zeros = np.zeros(nt * n)
    crp_h, cre_h = w.crp0.copy(), w.cre0.copy()
    crp_o, cre_o = w.crp0.copy(), w.cre0.copy()
    graph_launches = 0
    try:
        el = 0.0
        for s in range(ns):
            allow = s + 1 < ns
            row, m1, p1 = o.ic_cooling_row(w.z[s])
            rown, m1n, p1n = o.ic_cooling_row(w.z[s + 1]) if allow else (row, m1, p1)
            bgc = w.bg[s]
            bgn = w.bg[s + 1] if allow else w.bg[s]
            nsub, _, non, _ = o.nsub(prm_o, w.dt_snap[s], bgc[0], bgc[1], bgc[2], bgc[3], row[0], row[-1], bgc[4],
                                     bgc[5])
            qpi, qepri = w.qpi.get(s, zeros), w.qepri.get(s, zeros)
            print(f"[dropin] n={n} species={sp} snapshot={s}", flush=True)
            t = ctx.execute_snapshot(nsub, non, allow, el, w.dt_snap[s], bgc, row, m1, p1, bgn, rown, m1n, p1n, qpi,
                                     qepri, fqe, npmin, w.mass, None, crp_h, cre_h)
            graph_launches += t.graph_launches
            assert t.graph_failures == 0
            o.execute_snapshot(prm_o, nsub, non, allow, el, w.dt_snap[s], bgc, row, m1, p1, bgn, rown, m1n, p1n, qpi,
                               qepri, fqe, npmin, w.mass, None, crp_o, cre_o)
            e_err, e_ok = rel_err(cre_h, cre_o)
            assert e_ok, f"snapshot {s}: electron max rel err {e_err:.3e}"
            if w.protons:
                p_err, p_ok = rel_err(crp_h, crp_o)
                assert p_ok, f"snapshot {s}: proton max rel err {p_err:.3e}"
            else:
                assert np.all(crp_h == 0.0)
            el += float(w.dt_snap[s])
    finally:
        ctx.close()
    # the third call of equal shape replays the graph captured by the second
    assert graph_launches >= ns - 1
