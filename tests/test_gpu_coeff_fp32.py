"""FP32 coefficient evaluation (croma_hip_context_set_coeff_precision) against
the FP64 oracle.

In the FP32 mode the loss rates, D_pp rows, the drift B and diffusion C of
every row and the face averages / Peclet numbers are float; the Bernoulli
face weights, the rows a, b, c, the factorisation and every substep stay
FP64.  The spectra then carry the float coefficients' error, which the
reference comparator's 1e-10 bar does not admit; each case states its
measured maximum relative error over the entries above the 1e-30 floor and
within 30 decades of the tracer's peak (deeper entries are exponential tails
whose relative error grows with the depth of the cutoff) and asserts it
stays below 1e-4.  Switching back to FP64 gives the FP64 path's spectra bit
for bit."""
import numpy as np
import pytest

import croma_hip as ch
import run_oracle
from workloads import hip_params, rel_err, run_hip, small_workload

pytestmark = pytest.mark.gpu

CASES = [
    # n, ntracer, nsnap, species, dpp, dsa   (c0-, c1- and c2-shaped workloads)
    (128, 256, 4, 1, 3, 0.0),
    (256, 192, 3, 3, 1, 0.05),
    (512, 128, 2, 1, 2, 0.0),
]


def fp32_err(got, want, floor=1e-30, decades=30.0):
    got = np.asarray(got).reshape(-1)
    want = np.asarray(want).reshape(-1)
    peak = np.max(np.abs(want))
    m = np.abs(want) > max(floor, peak * 10.0 ** -decades)
    return float(np.max(np.abs(got[m] - want[m]) / np.abs(want[m]))) if m.any() else 0.0


@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}-sp{c[3]}-dpp{c[4]}" for c in CASES])
def test_fp32_coefficients_error_vs_fp64_oracle(case):
    n, nt, ns, sp, dpp, dsa = case
    w = small_workload(n, nt, ns, sp, dpp, seed=23, dsa_frac=dsa)
    print(f"[fp32] n={n} ntracer={nt} nsnap={ns} species={sp} dpp={dpp}", flush=True)
    crp_ref, cre_ref = run_oracle.run_workload(w, run_oracle.oracle_params(w))
    ctx = ch.Context(n, hip_params(w))
    try:
        crp64, cre64, _ = run_hip(w, ctx)
        ctx.set_coeff_precision(True)
        crp32, cre32, t32 = run_hip(w, ctx)
        ctx.set_coeff_precision(False)
        crp64b, cre64b, _ = run_hip(w, ctx)
    finally:
        ctx.close()
    assert np.array_equal(cre64, cre64b) and np.array_equal(crp64, crp64b)
    _, ok64 = rel_err(cre64, cre_ref)
    assert ok64
    e_err = fp32_err(cre32, cre_ref)
    msg = f"  FP32 coefficients: electron max rel err {e_err:.3e}"
    errs = [e_err]
    if sp == 3:
        p_err = fp32_err(crp32, crp_ref)
        msg += f", proton {p_err:.3e}"
        errs.append(p_err)
    print(msg + f" (substeps {t32.cell_steps})", flush=True)
    assert np.isfinite(cre32).all()
    assert max(errs) < 1e-4
    assert max(errs) > 0.0  # the FP32 instance ran (its rounding differs from FP64)
