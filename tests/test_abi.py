"""The C ABI library: it loads without a GPU, exports every function that
the headers under include/ declare, and its host-side logic (grid tables, nsub
schedule, IC end bins, admission memory model, backend parsing) matches the
oracle / the compiled reference fixtures.  No device calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "croma-fp_amd", "libcroma_hip.so")
HDRS = [os.path.join(REPO, "include", f) for f in sorted(os.listdir(os.path.join(REPO, "include")))
        if f.endswith(".h")]


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    return C.CDLL(LIB)


def declared():
    names = set()
    for h in HDRS:
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(croma_hip_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_exports_every_declared_function(lib):
    names = declared()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared under include/ but not exported: {missing}"


def test_backend_parse(lib):
    lib.croma_hip_backend_parse.argtypes = [C.c_char_p]
    assert lib.croma_hip_backend_parse(b"hip") == 1
    assert lib.croma_hip_backend_parse(b"cpu") == 0
    assert lib.croma_hip_backend_parse(None) == 0


def test_chunk_bytes_model(lib):
    f = lib.croma_hip_chunk_bytes
    f.restype = C.c_size_t
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    e = f(1000, 128, 1, 10)
    ep = f(1000, 128, 3, 10)
    # spectra + sources: 8 B per bin per species each, backgrounds 8 fields per snapshot
    assert e >= 2 * 1000 * 128 * 8 + 11 * 8 * 1000 * 8
    assert ep - e >= 2 * 1000 * 128 * 8
    assert f(2000, 128, 1, 10) > e and f(1000, 128, 1, 20) > e
    assert f(0, 128, 1, 10) == 0


class _Grid(C.Structure):
    # ChGrid (croma-fp_amd/host/croma_host.h), CH_NMAX = 512
    _fields_ = [("n", C.c_int), ("j_pp", C.c_int), ("bounds", C.c_double * 4)] + [
        (name, C.c_double * 512) for name in
        ("pp", "pdp", "pE", "beta_p", "p_p2", "p_expcut", "sigma_pp", "sigmoid", "p_sig", "p_ecr_w", "ep", "edp",
         "eE", "e_p2", "e_expcut", "e_ecr_w")] + [("ppm1", C.c_double), ("ppp1", C.c_double), ("epm1", C.c_double),
                                                  ("epp1", C.c_double)]


@pytest.mark.parametrize("n", [64, 128, 256, 512])
def test_host_grid_and_ic_end_bins_match_reference(lib, n):
    d = np.load(os.path.join(REPO, "tests", "golden", f"ref_n{n}.npz"))
    pmin, pmax, pemin, pemax = (float(v) for v in d["grid_bounds"])
    g = _Grid()
    lib.ch_grid_build.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    assert lib.ch_grid_build(C.byref(g), n, pmin, pmax, pemin, pemax) == 0
    for key in ("pp", "pdp", "pE", "ep", "edp", "eE"):
        assert np.array_equal(np.ctypeslib.as_array(getattr(g, key))[:n], d[key]), key
    bins = (C.c_int * 2)(0, n - 1)
    out = (C.c_double * 2)()
    lib.ch_ic_cooling_bins.argtypes = [C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(C.c_double)]
    for iz, z in enumerate(d["ic_z"]):
        assert lib.ch_ic_cooling_bins(C.byref(g), C.c_double(z), 2, bins, out) == 0
        assert out[0] == d["ic_rows"][iz][0] and out[1] == d["ic_rows"][iz][n - 1]


def test_host_nsub_matches_oracle(lib):
    import sys
    sys.path[:0] = [os.path.join(REPO, "croma-fp_amd"), os.path.join(REPO, "oracle")]
    import pyoracle as po
    import run_oracle
    import synth

    w = synth.make_workload(n=128, ntracer=300, nsnap=2, species=3, dpp_model=1, seed=5)
    o = po.Oracle(w.n, w.pmin, w.pmax, w.pemin, w.pemax)
    prm = run_oracle.oracle_params(w)
    row = o.ic_cooling_row(w.z[0])[0]
    bgc = w.bg[0]
    ns_ref, _, non_ref, _ = o.nsub(prm, w.dt_snap[0], bgc[0], bgc[1], bgc[2], bgc[3], row[0], row[-1], bgc[4],
                                   bgc[5])
    g = _Grid()
    lib.ch_grid_build.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    lib.ch_grid_build(C.byref(g), w.n, w.pmin, w.pmax, w.pemin, w.pemax)

    class _Prm(C.Structure):  # CkParams (croma-fp_amd/csrc/croma_kernels.h)
        _fields_ = [(k, C.c_int) for k in ("dpp_model", "ttd_tacc_model", "reacc_window_mode", "tracer_nsub_max",
                                           "coeff_interp_min_steps", "coeff_interp_max_segments")] + [
            (k, C.c_double) for k in ("psi", "mach_limit", "eta_dpp_cap", "t_acc_direct_gyr", "t_off_gyr",
                                      "t_acc_off_gyr", "f_eddy", "tracer_nsub_safety", "epmax")]

    p = _Prm(1, 1, 0, 0, 4, 16, 0.5, 0.5, 0.0, 0.0, 0.0, 3.0, 2.0, 0.5, 1.0e7)
    ns = np.zeros(w.ntracer, np.int32)
    non = np.zeros(w.ntracer, np.int32)
    dp = C.POINTER(C.c_double)
    f = [np.ascontiguousarray(bgc[k]) for k in range(6)]
    lib.ch_nsub_estimate.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_double, C.c_double, C.c_double] + \
        [dp] * 6 + [C.POINTER(C.c_int)] * 2
    rc = lib.ch_nsub_estimate(C.byref(g), C.byref(p), w.ntracer, w.dt_snap[0], row[0], row[-1],
                              *[a.ctypes.data_as(dp) for a in f], ns.ctypes.data_as(C.POINTER(C.c_int)),
                              non.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == 0
    assert np.array_equal(ns, np.asarray(ns_ref)) and np.array_equal(non, np.asarray(non_ref))
