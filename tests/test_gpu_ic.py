"""Device-built inverse-Compton cooling rows vs the compiled reference's rows
(golden fixtures written by the reference's prepare_ic_cooling_row,
fp_shared_core.c:844-864) and vs the host restatement.  The device evaluates
the same quadrature with its own libm, so agreement is to rounding (the
fixtures' rows are smooth positive rates; 1e-12 relative is ~1e4 ulp of
headroom for the ~1e4-term sums)."""
import numpy as np
import pytest

import croma_hip as ch

pytestmark = pytest.mark.gpu

GOLD = "tests/golden/ref_n{}.npz"


@pytest.mark.parametrize("n", [64, 128, 256, 512])
def test_device_ic_rows_match_reference(n):
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", f"ref_n{n}.npz"))
    pmin, pmax, pemin, pemax = (float(v) for v in d["grid_bounds"])
    print(f"[ic] n={n} z={d['ic_z'].tolist()}", flush=True)
    ctx = ch.Context(n, ch.default_params(), pmin=pmin, pmax=pmax, pemin=pemin, pemax=pemax)
    try:
        rows, gh = ctx.ic_cooling_rows_device(d["ic_z"])
        host = np.array([ctx.ic_cooling_row(z)[0] for z in d["ic_z"]])
    finally:
        ctx.close()
    ref = d["ic_rows"]
    err = np.max(np.abs(rows - ref) / np.abs(ref))
    print(f"  max rel err vs reference {err:.3e}", flush=True)
    assert err < 1e-12
    assert np.max(np.abs(gh - d["ic_edges"]) / np.abs(d["ic_edges"])) < 1e-12
    assert np.array_equal(host, ref)
