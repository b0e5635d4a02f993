"""The step before the solve on the device: nsub / n_on estimates, the cost
order of the queue and the split segment rows.

1. nsub and n_on from the device estimate equal the oracle's
   (tracer_fp_nsub.c:94-282 restated) integer for integer, on the
   backgrounds of the bulk-throughput and batched workloads at 10^5 tracers,
   with and without the eddy-time on-window.
2. Batch runs with the estimate and the ordering on the device give the same
   spectra and the same tracer-substep count as with them on the host:
   bit-identical where the launches are the same; within 1e-13 relative for
   the split launches, whose GEMM then covers every row the batch can hold."""
import numpy as np
import pytest

import croma_hip as ch
import pyoracle as po
import run_oracle
import synth
from workloads import hip_params, small_workload

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nt,nsnap,dpp,extra", [
    (512, 100_000, 20, 2, {}),
    (256, 100_000, 100, 1, {}),
    (128, 50_000, 100, 1, {"reacc_window_mode": 1, "f_eddy": 0.02}),
])
def test_device_nsub_equals_oracle(n, nt, nsnap, dpp, extra):
    w = synth.make_workload(n=64, ntracer=nt, nsnap=nsnap, species=1, dpp_model=dpp, seed=41)
    o = po.Oracle(n, w.pmin, w.pmax, w.pemin, w.pemax)
    prm_o = run_oracle.oracle_params(w, **extra)
    ctx = ch.Context(n, hip_params(w, **extra))
    try:
        for s in (0, nsnap // 2, nsnap - 1):
            bg = w.bg[s]
            row = o.ic_cooling_row(w.z[s])[0]
            print(f"[nsub] n={n} tracers={nt} snapshot={s} {extra}", flush=True)
            ns_o, _, non_o, _ = o.nsub(prm_o, w.dt_snap[s], bg[0], bg[1], bg[2], bg[3], row[0], row[-1], bg[4], bg[5])
            ns_d, non_d = ctx.nsub_device(w.dt_snap[s], bg[0], bg[1], bg[2], bg[3], row[0], row[-1], bg[4], bg[5])
            bad = np.flatnonzero((ns_o != ns_d) | (non_o != non_d))
            assert bad.size == 0, f"snapshot {s}: {bad.size} tracers differ, first {bad[:5]}: " \
                                  f"{ns_o[bad[:5]]} vs {ns_d[bad[:5]]}"
            print(f"  nsub range {ns_o.min()}..{ns_o.max()}, n_on range {non_o.min()}..{non_o.max()}", flush=True)
    finally:
        ctx.close()


CASES = [
    # n, ntracer, nsnap, species, dpp, dsa, rows, waves, chain, gemm
    (128, 200, 3, 1, 1, 0.0, 8, 1, 1, 0),
    (256, 160, 4, 3, 1, 0.3, 8, 2, 2, 0),
    (256, 160, 4, 3, 1, 0.3, 4, 2, 1, 1),
    (64, 300, 3, 3, 3, 0.2, 4, 1, 3, 1),
]


def _run(ctx, w):
    b = ch.Batch(ctx, w.ntracer, w.species, w.nsnap)
    try:
        b.set_state(w.crp0 if w.protons else None, w.cre0, mass=w.mass)
        for s in range(w.nsnap):
            b.set_background(s, w.bg[s])
            if s in w.qpi or s in w.qepri:
                b.set_sources(s, w.qpi.get(s), w.qepri.get(s))
        t = b.run(0, w.nsnap, w.nsnap, w.z, w.dt_snap)
        crp, cre = b.get_state()
    finally:
        b.close()
    return crp, cre, t


@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}-sp{c[3]}-chain{c[8]}-gemm{c[9]}" for c in CASES])
def test_device_order_equals_host(case):
    n, nt, ns, sp, dpp, dsa, r, occ, chain, gemm = case
    w = small_workload(n, nt, ns, sp, dpp, dsa_frac=dsa, seed=43)
    ctx = ch.Context(n, hip_params(w))
    try:
        ctx.set_tile(r, occ, chain, secondary_gemm=gemm)
        print(f"[order] n={n} species={sp} chain={chain} gemm={gemm} host", flush=True)
        ctx.set_nsub_device(False)
        crp_h, cre_h, th = _run(ctx, w)
        print(f"[order] n={n} species={sp} chain={chain} gemm={gemm} device", flush=True)
        ctx.set_nsub_device(True)
        crp_d, cre_d, td = _run(ctx, w)
    finally:
        ctx.close()
    assert th.cell_steps == td.cell_steps and th.species_steps == td.species_steps
    assert td.nsub_estimate_ms == 0.0 and td.bucket_build_ms == 0.0
    if gemm and sp == 3:
        for got, want in ((cre_d, cre_h), (crp_d, crp_h)):
            assert np.all(np.abs(got - want) <= 1e-300 + 1e-13 * np.abs(want))
    else:
        assert np.array_equal(cre_d, cre_h) and np.array_equal(crp_d, crp_h)
