"""Every launch shape of the fused kernel (rows per lane x waves per SIMD x
snapshots per task x secondary contraction in the kernel or as one GEMM
between a proton and an electron pass) against the CPU oracle, on small seeded workloads of both
species sets.  Same tolerance as the main parity test: |hip - cpu| <= 1e-30 +
1e-10 |cpu| (analysis/compare_tracer_outputs.py:126-127)."""
import numpy as np
import pytest

import croma_hip as ch
import run_oracle
from workloads import hip_params, rel_err, run_hip, small_workload

pytestmark = pytest.mark.gpu

CASES = [
    # n, ntracer, nsnap, species, dpp, extra params, dsa
    (256, 40, 3, 3, 1, {}, 0.3),
    (128, 48, 3, 1, 1, {"reacc_window_mode": 1, "f_eddy": 0.02}, 0.0),
    (512, 24, 2, 1, 2, {}, 0.0),
    (64, 48, 3, 3, 3, {}, 0.2),
]
ROWS = {64: (2, 4, 8, 16), 128: (2, 4, 8, 16), 256: (4, 8, 16), 512: (8, 16)}


@pytest.fixture(scope="module")
def refs():
    return {}


@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}-sp{c[3]}-dpp{c[4]}" for c in CASES])
def test_all_tiles_match_oracle(refs, case):
    n, nt, ns, sp, dpp, extra, dsa = case
    w = small_workload(n, nt, ns, sp, dpp, dsa_frac=dsa, seed=19)
    crp_ref, cre_ref = run_oracle.run_workload(w, run_oracle.oracle_params(w, **extra))
    ctx = ch.Context(n, hip_params(w, **extra))
    worst = 0.0
    try:
        for r in ROWS[n]:
            for occ in (1, 2):
                for chain, gemm in ((1, 0), (ns, 0), (1, 1), (ns, 1)):
                    if gemm and sp == 1:
                        continue  # the split launch exists for two species only
                    print(f"[tiles] n={n} species={sp} rows={r} waves={occ} chain={chain} gemm={gemm}", flush=True)
                    ctx.set_tile(r, occ, chain, secondary_gemm=gemm)
                    crp, cre, _ = run_hip(w, ctx, path=ch.PATH_FUSED)
                    e_err, e_ok = rel_err(cre, cre_ref)
                    assert e_ok, f"rows={r} waves={occ} chain={chain}: electron max rel err {e_err:.3e}"
                    worst = max(worst, e_err)
                    if sp == 3:
                        p_err, p_ok = rel_err(crp, crp_ref)
                        assert p_ok, f"rows={r} waves={occ} chain={chain}: proton max rel err {p_err:.3e}"
                        worst = max(worst, p_err)
                    else:
                        assert np.all(crp == 0.0)
    finally:
        ctx.close()
    print(f"  worst rel err {worst:.3e}", flush=True)


@pytest.mark.parametrize("sp,gemm", [(3, 1), (3, 0), (1, 0)])
def test_two_snapshot_tasks_match_oracle(sp, gemm):
    """Two snapshots per queue task over five snapshots (windows of 2, 2, 1),
    as the workload autotuner picks for the bench workloads."""
    n, nt, ns = 128, 96, 5
    w = small_workload(n, nt, ns, sp, 1, dsa_frac=0.2, seed=37)
    crp_ref, cre_ref = run_oracle.run_workload(w, run_oracle.oracle_params(w))
    ctx = ch.Context(n, hip_params(w))
    try:
        ctx.set_tile(8, 2, 2, secondary_gemm=gemm)
        print(f"[tiles] two snapshots per task, species={sp} gemm={gemm}", flush=True)
        crp, cre, _ = run_hip(w, ctx, path=ch.PATH_FUSED)
    finally:
        ctx.close()
    e_err, e_ok = rel_err(cre, cre_ref)
    assert e_ok, f"electron max rel err {e_err:.3e}"
    if sp == 3:
        p_err, p_ok = rel_err(crp, crp_ref)
        assert p_ok, f"proton max rel err {p_err:.3e}"


def test_batch_autotune_restores_state():
    """The workload autotuner times every tile on the batch's own snapshot and
    leaves the state as it found it; the chosen tile then meets the parity bar."""
    n, nt, ns = 128, 64, 3
    w = small_workload(n, nt, ns, 3, 1, dsa_frac=0.2, seed=31)
    crp_ref, cre_ref = run_oracle.run_workload(w, run_oracle.oracle_params(w))
    ctx = ch.Context(n, hip_params(w))
    b = ch.Batch(ctx, w.ntracer, w.species, w.nsnap)
    try:
        b.set_state(w.crp0, w.cre0, mass=w.mass)
        for s in range(w.nsnap):
            b.set_background(s, w.bg[s])
            if s in w.qpi or s in w.qepri:
                b.set_sources(s, w.qpi.get(s), w.qepri.get(s))
        print("[tiles] batch autotune", flush=True)
        tile = b.autotune(0, w.nsnap, w.z, w.dt_snap)
        assert tile == ctx.get_tile(True)
        assert tile[2] in (1, 2)
        crp, cre = b.get_state()
        assert np.array_equal(crp, w.crp0) and np.array_equal(cre, w.cre0)
        b.run(0, ns, ns, w.z, w.dt_snap)
        crp, cre = b.get_state()
    finally:
        b.close()
        ctx.close()
    e_err, e_ok = rel_err(cre, cre_ref)
    p_err, p_ok = rel_err(crp, crp_ref)
    assert e_ok and p_ok, f"tile {tile}: rel err e {e_err:.3e} p {p_err:.3e}"
