"""The reference-struct surface (include/croma_hip_b1.h) on the device against
the CPU oracle.

run_tracer_pipeline is the reference's bucket-segment unit
(include/tracer_fp_solve.inc:1418-1471): one call advances a bucket of
tracers with equal nsub over one coefficient segment.  The oracle's snapshot
update of the same tracers with that nsub is the same arithmetic when the
snapshot is one segment (allow_interp 0) or, with two segments, when the two
pipeline calls carry the segment midpoints alpha = 1/4, 3/4 through
device_interp_background and the second continues the first's resident
state.  Spectra within |hip - cpu| <= 1e-30 + 1e-10 |cpu| (the reference
comparator).  Also: the rank-state gather / scatter path, the batched solve
(solve_cc_cuda_batch) and the gamma / neutrino emission entry points."""
import ctypes as C

import numpy as np
import pytest

import croma_b1 as b1
import croma_hip as ch
import pyoracle as po
from workloads import hip_params, rel_err, small_workload

pytestmark = pytest.mark.gpu

N = 128


@pytest.fixture(scope="module")
def env():
    L = b1.lib()
    w = small_workload(N, 96, 1, 3, 1, seed=61, dsa_frac=0.3)
    prm = hip_params(w)
    assert L.croma_hip_b1_set_params(C.byref(prm)) == 0
    o = po.Oracle(N, w.pmin, w.pmax, w.pemin, w.pemax)
    ctx = ch.Context(N, prm, w.pmin, w.pmax, w.pemin, w.pemax)
    gp = ctx.grid(0)
    ge = ctx.grid(1)
    ctx.close()
    crp_grid = b1.crspectrum(gp[0], gp[1], gp[3][0], gp[3][1], gp[2])
    cre_grid = b1.crspectrum(ge[0], ge[1], ge[3][0], ge[3][1], ge[2])
    ws = L.croma_hip_b1_workspace_create()
    yield dict(L=L, w=w, prm=prm, o=o, crp_grid=crp_grid, cre_grid=cre_grid, ws=ws,
               fqe=o.secondary_kernel())
    L.croma_hip_b1_workspace_destroy(ws)


def _oracle(e, nsub, n_on, allow, idx=None, dis=None, mseg=None):
    w, o = e["w"], e["o"]
    sub = w.subset(idx) if idx is not None else w
    kw = dict(dpp_model=w.dpp_model, tracer_nsub_safety=0.5)
    if mseg:
        kw["coeff_interp_max_segments"] = mseg
    prm = po.default_params(**kw)
    row, m1, p1 = o.ic_cooling_row(w.z[0])
    rown, m1n, p1n = o.ic_cooling_row(w.z[1]) if allow else (row, m1, p1)
    crp, cre = sub.crp0.copy(), sub.cre0.copy()
    nt = sub.ntracer
    zeros = np.zeros(nt * N)
    bgn = sub.bg[1] if allow else sub.bg[0]
    o.execute_snapshot(prm, np.full(nt, nsub, np.int32), np.full(nt, n_on, np.int32), allow, 0.0, w.dt_snap[0],
                       sub.bg[0], row, m1, p1, bgn, rown, m1n, p1n, sub.qpi.get(0, zeros), sub.qepri.get(0, zeros),
                       e["fqe"][0], e["fqe"][1], sub.mass, dis, crp, cre)
    return crp, cre


class Bucket:
    """Host arrays of one pipeline call (kept alive while the call runs)."""

    def __init__(self, e, sub, alpha=None, dis=None):
        w, o = e["w"], e["o"]
        nt = sub.ntracer
        self.keep = []
        row = o.ic_cooling_row(w.z[0])[0]
        ric = np.ascontiguousarray(np.repeat(row[:, None], nt, axis=1))  # [npe][ncell]
        ci = b1.CoeffBatchInput()
        ci.ncell = nt
        ci.z = w.z[0]
        ci.epmax = e["prm"].epmax
        names = ["n_gas", "kbt", "b_field", "divv_gyr", "l_turb_mpc", "dv_imc", "cs", "beta_pl"]
        for k, f in enumerate(names):
            a = np.ascontiguousarray(sub.bg[0][k])
            self.keep.append(a)
            setattr(ci, f, b1.ptr(a))
        self.keep += [ric, sub.mass]
        ci.rad_ic_batch = b1.ptr(ric)
        ci.tracer_mass_msun = b1.ptr(sub.mass)
        if dis is not None:
            self.keep.append(dis)
            ci.disable_adiabatic = b1.ptr(dis)
        if alpha is not None:
            rown = o.ic_cooling_row(w.z[1])[0]
            ricn = np.ascontiguousarray(np.repeat(rown[:, None], nt, axis=1))
            self.keep.append(ricn)
            ci.device_interp_background = 1
            ci.interp_alpha = alpha
            for k, f in enumerate(names):
                a = np.ascontiguousarray(sub.bg[1][k])
                self.keep.append(a)
                setattr(ci, f + "_next", b1.ptr(a))
            ci.rad_ic_batch_next = b1.ptr(ricn)
        self.ci = ci


def _pipeline(e, bucket, nt, nstep, dt, window=None, state=None, out=True, rank=None):
    p = b1.PipelineInput()
    p.ncell = nt
    p.nstep = nstep
    p.fp_cadence_steps = p.diff_cadence_steps = p.adv_cadence_steps = 1
    if window:
        p.use_windowed_reacc = 1
        p.on_start_step, p.on_end_step = window
    p.dt = dt
    p.epmax = e["prm"].epmax
    p.psi_value = e["prm"].psi
    p.mach_limit_value = e["prm"].mach_limit
    p.crp_grid = e["crp_grid"].ctypes.data
    p.cre_grid = e["cre_grid"].ctypes.data
    p.coeff_in = C.pointer(bucket.ci)
    keep = []
    if state is not None:
        for f, a in zip(("crp_init", "cre_init", "qpi_batch", "qepri_batch"), state):
            a = np.ascontiguousarray(a, np.float64)
            keep.append(a)
            setattr(p, f, b1.ptr(a))
    if rank is not None:
        idx = np.ascontiguousarray(rank, np.int32)
        keep.append(idx)
        p.rank_state_indices = b1.ptr(idx)
        p.gather_rank_state = 1
        p.scatter_rank_state = 1
    p.workspace = e["ws"]
    crp = np.zeros(nt * N) if out else None
    cre = np.zeros(nt * N) if out else None
    t = b1.PipelineTimes()
    rc = e["L"].croma_hip_b1_run_tracer_pipeline(N, N, C.byref(p), None, b1.ptr(crp), b1.ptr(cre), None, None, None,
                                                  C.byref(t))
    assert rc == 0, ch.last_error()
    return crp, cre, t


def _check(crp, cre, crp_ref, cre_ref, what):
    e_err, e_ok = rel_err(cre, cre_ref)
    p_err, p_ok = rel_err(crp, crp_ref)
    print(f"[b1] {what}: electron {e_err:.3e}, proton {p_err:.3e}", flush=True)
    assert e_ok and p_ok


@pytest.mark.parametrize("dis_frac", [0.0, 0.4])
def test_pipeline_one_segment_windowed_matches_oracle(env, dis_frac):
    w = env["w"]
    nsub, n_on = 40, 16
    dis = (np.random.default_rng(7).random(w.ntracer) < dis_frac).astype(np.uint8) if dis_frac else None
    crp_ref, cre_ref = _oracle(env, nsub, n_on, 0, dis=dis)
    bucket = Bucket(env, w, dis=dis)
    ls = (nsub - n_on) // 2
    zeros = np.zeros(w.ntracer * N)
    crp, cre, t = _pipeline(env, bucket, w.ntracer, nsub, w.dt_snap[0] / nsub, window=(ls, ls + n_on),
                            state=(w.crp0, w.cre0, w.qpi.get(0, zeros), w.qepri.get(0, zeros)))
    _check(crp, cre, crp_ref, cre_ref, f"one segment, window [{ls}, {ls + n_on}), dis {dis_frac}")
    assert t.total_ms > 0.0


def test_pipeline_two_segments_interp_and_resident_state(env):
    w = env["w"]
    nsub = 40
    crp_ref, cre_ref = _oracle(env, nsub, nsub, 1, mseg=2)
    zeros = np.zeros(w.ntracer * N)
    dt = w.dt_snap[0] / nsub
    _pipeline(env, Bucket(env, w, alpha=0.25), w.ntracer, nsub // 2, dt,
              state=(w.crp0, w.cre0, w.qpi.get(0, zeros), w.qepri.get(0, zeros)), out=False)
    crp, cre, _ = _pipeline(env, Bucket(env, w, alpha=0.75), w.ntracer, nsub // 2, dt)
    _check(crp, cre, crp_ref, cre_ref, "two segments, device interpolation, resident state")


def test_pipeline_rank_state_gather_scatter(env):
    w = env["w"]
    L = env["L"]
    nt = w.ntracer
    zeros = np.zeros(nt * N)
    qpi = w.qpi.get(0, zeros).copy()
    qepri = w.qepri.get(0, zeros).copy()
    crp0, cre0 = w.crp0.copy(), w.cre0.copy()
    assert L.croma_hip_b1_rank_state_upload(N, N, env["ws"], nt, b1.ptr(qpi), b1.ptr(qepri), b1.ptr(crp0),
                                            b1.ptr(cre0), None) == 0
    idx = np.arange(1, nt, 2)  # the bucket: odd tracers of the rank
    sub = w.subset(idx)
    nsub, n_on = 24, 24
    crp_ref, cre_ref = _oracle(env, nsub, n_on, 0, idx=idx)
    _pipeline(env, Bucket(env, sub), len(idx), nsub, w.dt_snap[0] / nsub, out=False, rank=idx)
    crp_r, cre_r = np.zeros(nt * N), np.zeros(nt * N)
    assert L.croma_hip_b1_rank_state_download(N, N, env["ws"], nt, b1.ptr(crp_r), b1.ptr(cre_r), None) == 0
    rows = (idx[:, None] * N + np.arange(N)[None, :]).ravel()
    _check(crp_r[rows], cre_r[rows], crp_ref, cre_ref, "rank gather / scatter")
    other = np.setdiff1d(np.arange(nt * N), rows)
    assert np.array_equal(cre_r[other], cre0[other]) and np.array_equal(crp_r[other], crp0[other])
    # an index outside the rank state is refused on the host, before any launch
    with pytest.raises(AssertionError):
        _pipeline(env, Bucket(env, sub), len(idx), nsub, w.dt_snap[0] / nsub, out=False,
                  rank=np.where(np.arange(len(idx)) == 3, nt, idx))
    assert "outside the rank state" in ch.last_error()


def test_solve_and_tridiag_selector(env):
    L = env["L"]
    rng = np.random.default_rng(5)
    nsys, nrow, dt = 300, N, 0.02
    a = rng.uniform(0, 1, (nsys, nrow))
    c = rng.uniform(0, 1, (nsys, nrow))
    b = 1.0 + a + c
    q = rng.uniform(0, 1, (nsys, nrow))
    x = rng.uniform(0, 1, (nsys, nrow))
    out = np.zeros_like(x)
    s = b1.SolveBatchInput(nsys, nrow, dt, b1.ptr(a), b1.ptr(b), b1.ptr(c), b1.ptr(q))
    used = C.c_int(7)
    assert L.croma_hip_b1_solve_cc_batch(C.byref(s), b1.ptr(x), b1.ptr(out), None, C.byref(used)) == 0
    assert used.value == 0
    assert L.croma_hip_b1_tridiag_uses_pcr(nrow) == 0
    assert L.croma_hip_b1_tridiag_solver_name(nrow) == b"Thomas"
    for i in range(0, nsys, 37):
        _, xo = po.thomas(a[i], b[i], c[i], dt * q[i] + x[i])
        assert np.all(np.abs(out[i] - xo) <= 1e-12 * np.abs(xo))


@pytest.mark.parametrize("kind", ["gamma", "nu"])
def test_hadronic_emission_entry_points(env, kind):
    L, o = env["L"], env["o"]
    nb, emin, emax = 32, -1.0, 5.0
    fk = (po.gamma_kernel_flat if kind == "gamma" else po.nu_kernel_flat)(o, nb, emin, emax)
    rng = np.random.default_rng(9)
    nt = 200
    crp = np.ascontiguousarray((o.pp[None, :] ** -rng.uniform(2, 3, nt)[:, None]) * np.exp(-o.pp[None, :] / 1e6))
    ng = 10 ** rng.uniform(-5, -2, nt)
    ref = po.emission_hadronic(N, nb, emin, emax, ng, crp.ravel(), o.beta_p, o.pdp, fk).reshape(nt, nb)
    inp = b1.HadronicEmissionBatchInput(nt, nb, 0.0, emin, emax, b1.ptr(ng), b1.ptr(crp.ravel()), b1.ptr(o.beta_p),
                                        b1.ptr(fk))
    eps = np.zeros(nt * nb)
    fn = L.croma_hip_b1_prepare_gamma_emission if kind == "gamma" else L.croma_hip_b1_prepare_neutrino_emission
    assert fn(N, N, C.byref(inp), env["crp_grid"].ctypes.data, b1.ptr(eps), None) == 0
    eps = eps.reshape(nt, nb)
    m = np.abs(ref) > 1e-280
    err = float(np.max(np.abs(eps[m] - ref[m]) / np.abs(ref[m])))
    print(f"[b1] {kind} emission max rel err {err:.3e}", flush=True)
    assert err < 1e-12
