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
