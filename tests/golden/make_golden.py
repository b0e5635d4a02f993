"""Write the golden fixtures under tests/golden/.

Runs only where the reference sources exist (they do not travel to the GPU
box).  It

  1. builds the partial reference libraries (oracle/ref.mk -> oracle/_ref/),
  2. calls the reference's own functions through oracle/ref_harness.c on
     seeded inputs, and
  3. stores inputs + outputs as small .npz fixtures (data only).

It also tabulates the synchrotron function F(x) = x * int_x^inf K_{5/3}(t) dt
on the reference's grid (logx in [-25, 2.5), 2750 points,
tracer_fp_synch.c:53-56, 95-97) with scipy, using the reference's own
asymptotic branch for x > 300 (Synchrotron.c:36).  GSL is absent here, so
F(x) itself is checked only against its asymptotes (parity unpinned vs GSL).

It also records the byte layout of the reference's checkpoint header
(ckpt_header_layout.json) from the compiled struct.

and the DSA source rows of seeded shocks (dsa_n{64,256}.npz) and the
reacceleration rows of seeded pre-shock spectra (dsa_reacc_n{64,256}.npz).

Usage:  python tests/golden/make_golden.py [--layout-only | --dsa-only]
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_DIR = os.path.join(REPO, "oracle", "_ref")
SIZES = (64, 128, 256, 512)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def P(a):
    return a.ctypes.data_as(_ip if a.dtype == np.int32 else _dp)


def fx_table():
    from scipy import integrate, special

    nx, xmin, xmax = 2750, -25.0, 2.5
    dx = (xmax - xmin) / nx
    logx = np.array([xmin + dx * i for i in range(nx)])
    logfx = np.zeros(nx)
    for i, lx in enumerate(logx):
        x = 10.0 ** lx
        if x > 300.0:
            F = np.sqrt(0.5 * np.pi) * np.exp(-x) * x ** 0.5
        else:
            upper = np.log(max(60.0, 2 * x) / x)
            val, _ = integrate.quad(lambda u: special.kv(5.0 / 3.0, x * np.exp(u)) * x * np.exp(u),
                                    0.0, upper, epsabs=0.0, epsrel=1e-13, limit=400)
            F = x * val
        logfx[i] = np.log10(F)
    fx = 10.0 ** logfx
    return logx, logfx, fx, xmin


def build_ref():
    subprocess.check_call(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"])


def ref_lib(n):
    return C.CDLL(os.path.join(REF_DIR, f"libcroma_ref_n{n}.so"))


def random_backgrounds(rng, nt):
    n_gas = 10 ** rng.uniform(-5, -2, nt)
    temp = 10 ** rng.uniform(6, 8, nt)
    kbt = 8.62e-5 * temp * 1.0e-9
    b = 10 ** rng.uniform(-2, 1, nt) * 1e-6
    divv = rng.normal(0, 1e-17, nt) * 3.1536e16
    l_turb = np.full(nt, 0.15)
    cs = np.sqrt(5.0 / 3.0 * 8.62e-5 * temp / (0.59 * 0.938 * 1.0e9)) * 2.999e10
    dv = np.minimum(3e-17 * np.exp(rng.normal(0, 1, nt)) * 0.15 * 3.0857e24, 0.5 * cs)
    return n_gas, kbt, b, divv, l_turb, dv, cs


def make_size(n, logx, logfx, fx, xmin, rng):
    ref = ref_lib(n)
    out = {}
    pp, pdp, pE, ep, edp, eE, bp = (np.zeros(n) for _ in range(7))
    pg, eg = np.zeros(2), np.zeros(2)
    grids = (-1.0, 8.0, 0.0, 6.0)
    ref.ref_grid(*(C.c_double(v) for v in grids), P(pp), P(pdp), P(pE), P(pg), P(ep), P(edp), P(eE),
                 P(eg), P(bp))
    out.update(grid_bounds=np.array(grids), pp=pp, pdp=pdp, pE=pE, pghost=pg, ep=ep, edp=edp, eE=eE,
               eghost=eg, beta_p=bp)

    zs = np.array([0.0, 0.5, 1.0])
    rows = np.zeros((3, n))
    redges = np.zeros((3, 2))
    for i, z in enumerate(zs):
        ref.ref_ic_row(C.c_double(z), P(rows[i]), P(redges[i]))
    out.update(ic_z=zs, ic_rows=rows, ic_edges=redges)

    nt = 16
    n_gas, kbt, b, divv, l_turb, dv, cs = random_backgrounds(rng, nt)
    radp = np.zeros((nt, n))
    tl = np.zeros((nt, n))
    ed = np.zeros((nt, 2))
    for i in range(nt):
        ref.ref_losses_e(C.c_double(n_gas[i]), C.c_double(b[i]), C.c_double(divv[i]), P(rows[1]),
                         C.c_double(redges[1, 0]), C.c_double(redges[1, 1]), P(radp[i]), P(tl[i]), P(ed[i]))
    out.update(le_n=n_gas, le_b=b, le_divv=divv, le_radp=radp, le_tloss=tl, le_edges=ed)

    # scalar loss / acceleration formulas over a spread of arguments
    sp = 10 ** rng.uniform(-1, 8, 64)
    sn = 10 ** rng.uniform(-5, -2, 64)
    sk = 10 ** rng.uniform(6, 8, 64) * 8.62e-14
    sb = 10 ** rng.uniform(-2, 1, 64) * 1e-6
    out.update(s_p=sp, s_n=sn, s_kbt=sk, s_b=sb)
    ref.ref_b_coulomb_p.restype = C.c_double
    ref.ref_b_coulomb_e.restype = C.c_double
    ref.ref_b_synch.restype = C.c_double
    ref.ref_tacc_asa.restype = C.c_double
    ref.ref_tacc_ttd.restype = C.c_double
    out["s_coul_p"] = np.array([ref.ref_b_coulomb_p(C.c_double(a), C.c_double(p), C.c_double(k))
                                for a, p, k in zip(sn, sp, sk)])
    out["s_coul_e"] = np.array([ref.ref_b_coulomb_e(C.c_double(a), C.c_double(p)) for a, p in zip(sn, sp)])
    out["s_synch"] = np.array([ref.ref_b_synch(C.c_double(p), C.c_double(bb)) for p, bb in zip(sp, sb)])
    tl_cm = 10 ** rng.uniform(-2, 0, 32) * 3.0857e24
    tcs = 10 ** rng.uniform(6.5, 8, 32)
    tm = rng.uniform(0.01, 0.5, 32)
    tbeta = 10 ** rng.uniform(1, 3, 32)
    tdv = tm * tcs
    out.update(t_l=tl_cm, t_cs=tcs, t_mach=tm, t_beta=tbeta, t_dv=tdv)
    out["t_asa"] = np.array([ref.ref_tacc_asa(C.c_double(a), C.c_double(b_), C.c_double(c_), C.c_double(d),
                                              C.c_double(0.5)) for a, b_, c_, d in zip(tl_cm, tcs, tm, tbeta)])
    out["t_ttd1"] = np.array([ref.ref_tacc_ttd(C.c_double(a), C.c_double(b_), C.c_double(c_), C.c_double(d), 1)
                              for a, b_, c_, d in zip(tl_cm, tcs, tm, tdv)])
    out["t_ttd0"] = np.array([ref.ref_tacc_ttd(C.c_double(a), C.c_double(b_), C.c_double(c_), C.c_double(d), 0)
                              for a, b_, c_, d in zip(tl_cm, tcs, tm, tdv)])

    # nsub estimator: 256 tracers, both window modes
    nt = 256
    n_gas, kbt, b, divv, l_turb, dv, cs = random_backgrounds(rng, nt)
    ric_mm = np.ascontiguousarray(np.repeat(rows[1][:, None], nt, axis=1))
    for mode, tag in ((0, "w0"), (1, "w1")):
        ns, tg, non = (np.zeros(nt, np.int32) for _ in range(3))
        endp = np.zeros(4 * nt)
        rc = ref.ref_nsub(nt, C.c_double(0.077), P(n_gas), P(kbt), P(b), P(divv), P(ric_mm), P(l_turb), P(dv),
                          C.c_double(0.5), 0, mode, C.c_double(2.0), P(ns), P(tg), P(non), P(endp))
        assert rc == 0
        out[f"nsub_{tag}"] = ns
        out[f"nsub_target_{tag}"] = tg
        out[f"nsub_non_{tag}"] = non
        out[f"nsub_endp_{tag}"] = endp.reshape(4, nt)
    out.update(ns_n=n_gas, ns_kbt=kbt, ns_b=b, ns_divv=divv, ns_lturb=l_turb, ns_dv=dv)

    # tridiagonal solvers on random column-dominant M-matrices
    nsys = 8
    a = rng.uniform(0, 1, (nsys, n)) * 10 ** rng.uniform(-3, 2, (nsys, 1))
    c = rng.uniform(0, 1, (nsys, n)) * 10 ** rng.uniform(-3, 2, (nsys, 1))
    a[:, 0] = 0.0
    c[:, -1] = 0.0
    bdiag = 1.0 + np.roll(a, -1, axis=1) + np.roll(c, 1, axis=1) + rng.uniform(0, 1, (nsys, n))
    rhs = 10 ** rng.uniform(-30, 0, (nsys, n))
    th = rhs.copy()
    for s in range(nsys):
        assert ref.ref_thomas(n, P(a[s].copy()), P(bdiag[s].copy()), P(c[s].copy()), P(th[s])) == 0
    pcr = np.zeros_like(rhs)
    assert ref.ref_pcr(nsys, n, P(a), P(bdiag), P(c), P(rhs), P(pcr)) == 0
    q = 10 ** rng.uniform(-5, 0, (nsys, n))
    st = rhs.copy()
    assert ref.ref_solve_cc_batch(nsys, n, C.c_double(0.003), P(a), P(bdiag), P(c), P(q), P(st)) == 0
    out.update(tri_a=a, tri_b=bdiag, tri_c=c, tri_rhs=rhs, tri_thomas=th, tri_pcr=pcr, tri_q=q, tri_cc=st)

    # secondary source contraction with synthetic weights
    nt = 8
    fqe = rng.uniform(0, 1e-20, (n, n))
    npmin = rng.integers(0, n // 2, n).astype(np.int32)
    for je in range(n):
        fqe[: max(npmin[je], 0), je] = 0.0
    crp = 10 ** rng.uniform(-10, 0, (nt, n))
    qepri = 10 ** rng.uniform(-12, -6, (nt, n))
    ng = 10 ** rng.uniform(-5, -2, nt)
    inje = np.zeros((nt, n))
    ref.ref_secondary_source(nt, P(ng), P(crp), P(npmin), P(np.ascontiguousarray(fqe.ravel())), P(qepri),
                             P(inje))
    out.update(sec_fqe=fqe.ravel(), sec_npmin=npmin, sec_crp=crp, sec_qepri=qepri, sec_n=ng, sec_inje=inje)

    # IC emission kernel + contraction (8 frequency bins, z = 0.3)
    nb = 8
    fic = np.zeros(nb * n)
    ref.ref_ic_kernel_flat(nb, C.c_double(13.0), C.c_double(28.0), C.c_double(0.3), P(fic))
    cre = 10 ** rng.uniform(-20, 0, (4, n))
    eps = np.zeros(4 * nb)
    assert ref.ref_ic_emission(4, nb, P(cre), P(fic), P(eps)) == 0
    out.update(ic_fic=fic, ic_cre=cre, ic_eps=eps)

    # synchrotron pitch-kernel table on a small (nfreq, ntheta, nlogb) grid
    nfreq, nth, nlogb = 4, 16, 8
    logy = np.zeros(nfreq * n * nth)
    pw = np.zeros(nth)
    theta = (np.arange(nth) + 0.5) * np.pi / nth
    dth = np.pi / nth
    pw[:] = np.sin(theta) * np.sin(theta) * dth
    nus = np.array([10 ** (5.5 + (10.5 - 5.5) / nfreq * (i)) for i in range(nfreq)])
    for i in range(nfreq):
        for j in range(n):
            for k in range(nth):
                nu_c = 1.0 * (1.0 + ep[j] ** 2) * 4.194284e06 * np.sin(theta[k])
                logy[(i * n + j) * nth + k] = np.log10(nus[i] / (1.0e-6 * nu_c))
    table = np.zeros(nlogb * nfreq * n)
    inv_dlogb = (nlogb - 1) / (2.0 - (-6.0))
    assert ref.ref_synch_table(nfreq, nth, nlogb, C.c_double(-6.0), C.c_double(inv_dlogb), len(logx),
                               C.c_double(xmin), P(logx), P(logfx), P(fx), P(logy), P(pw), P(table)) == 0
    out.update(sy_logy=logy, sy_pw=pw, sy_table=table, sy_dims=np.array([nfreq, nth, nlogb]),
               sy_inv_dlogb=np.array([inv_dlogb]))
    return out


def ckpt_layout():
    """sizeof / offsetof of the reference's TracerFpCheckpointHeader
    (include/tracer_fp_types.h:356-371), compiled into the harness."""
    import json
    lib = ref_lib(128)
    out = (C.c_long * 16)()
    cnt = lib.ref_ckpt_header_layout(out)
    names = ["sizeof", "magic", "version", "mpi_rank", "world_size", "ntracer", "ntracer_global", "nsnap", "nfreq",
             "np_bins", "npe_bins", "nsnp_i_value", "nsnp_f_value", "completed_snapshot", "tracer_start",
             "sizeof_magic"]
    with open(os.path.join(HERE, "ckpt_header_layout.json"), "w") as f:
        json.dump({k: int(out[i]) for i, k in enumerate(names[:cnt])}, f, indent=1)
        f.write("\n")


def dsa_fixture(n, rng):
    """DSA source rows of seeded shocks from the reference's own builders
    (DSA_MODELS.c through ref_dsa_sources) and eta(M) of every model."""
    lib = ref_lib(n)
    pp, pdp, pE = np.zeros(n), np.zeros(n), np.zeros(n)
    ep, edp, eE = np.zeros(n), np.zeros(n), np.zeros(n)
    pg, eg, bp = np.zeros(2), np.zeros(2), np.zeros(n)
    d = C.c_double
    lib.ref_grid(d(-1.0), d(8.0), d(0.0), d(6.0), P(pp), P(pdp), P(pE), P(pg), P(ep), P(edp), P(eE), P(eg), P(bp))
    nsh = 48
    mach = 10 ** rng.uniform(np.log10(1.2), np.log10(40.0), nsh)
    temp = 10 ** rng.uniform(5.0, 8.5, nsh)
    rho = 10 ** rng.uniform(-30.0, -25.0, nsh)
    u1 = mach * np.sqrt(5.0 / 3.0 * 1.380649e-16 * temp / (0.59 * 1.67262192369e-24)) * rng.uniform(0.8, 1.2, nsh)
    # (eta model, chi_p, chi_e, K_ep, pmax_p, pmax_e) per shock
    models = rng.integers(0, 8, nsh)
    chi = rng.choice([3.0, 3.5, 4.0], nsh)
    chie = rng.choice([0.0, 3.5], nsh)
    kep = rng.choice([1e-2, 1e-3], nsh)
    pmax_e = rng.choice([1e5, 1e4], nsh)
    sp, se = np.zeros((nsh, n)), np.zeros((nsh, n))
    flags = np.zeros((nsh, 2), np.int32)
    spec = np.zeros((nsh, 5))
    for i in range(nsh):
        fl = (C.c_int * 2)()
        row_p, row_e, sv = np.zeros(n), np.zeros(n), np.zeros(5)
        lib.ref_dsa_sources(int(models[i]), d(chi[i]), d(chie[i]), d(kep[i]), d(1e6), d(pmax_e[i]), d(5.0 / 3.0),
                            d(mach[i]), d(temp[i]), d(rho[i]), d(u1[i]), P(row_p), P(row_e), fl, P(sv))
        sp[i], se[i], flags[i], spec[i] = row_p, row_e, (fl[0], fl[1]), sv
    lib.ref_dsa_eta.restype = C.c_double
    mgrid = np.concatenate([np.linspace(1.0, 3.0, 81), np.linspace(3.0, 40.0, 149)])
    eta = np.array([[lib.ref_dsa_eta(d(m), k) for m in mgrid] for k in range(8)])
    return dict(mach=mach, temp=temp, rho=rho, u1=u1, model=models.astype(np.int32), chi=chi, chie=chie, kep=kep,
                pmax_e=pmax_e, source_p=sp, source_e=se, flags=flags, spec=spec, eta_mach=mgrid, eta=eta)


def reacc_fixture(n, rng):
    """Reacceleration rows (tracer_fp_dsa_reacc.c through ref_dsa_reacc /
    ref_dsa_reacc_edens) of seeded pre-shock spectra on the harness grid, and
    eta interpolated between initial and reacceleration models
    (DSA_MODELS.c:474-522 through ref_dsa_eta_interpolated)."""
    lib = ref_lib(n)
    pp, pdp, pE = np.zeros(n), np.zeros(n), np.zeros(n)
    ep, edp, eE = np.zeros(n), np.zeros(n), np.zeros(n)
    pg, eg, bp = np.zeros(2), np.zeros(2), np.zeros(n)
    d = C.c_double
    lib.ref_grid(d(-1.0), d(8.0), d(0.0), d(6.0), P(pp), P(pdp), P(pE), P(pg), P(ep), P(edp), P(eE), P(eg), P(bp))
    lib.ref_dsa_reacc.restype = C.c_double
    lib.ref_dsa_reacc_edens.restype = C.c_double
    lib.ref_dsa_eta_interpolated.restype = C.c_double
    ncase = 24
    mach = 10 ** rng.uniform(np.log10(1.5), np.log10(30.0), ncase)
    g = 5.0 / 3.0
    r = (g + 1.0) * mach ** 2 / ((g - 1.0) * mach ** 2 + 2.0)
    q = 3.0 * r / (r - 1.0)
    v1 = 10 ** rng.uniform(60.0, 70.0, ncase)
    v2 = v1 / r
    pmin = 10 ** rng.uniform(-3.0, 0.5, ncase)
    state = np.zeros((ncase, 2, n))
    for i in range(ncase):
        for sp, axis in ((0, pp), (1, ep)):
            s_ = rng.uniform(2.0, 3.5)
            cut = 10 ** rng.uniform(1.0, 6.0)
            state[i, sp] = 10 ** rng.uniform(50, 60) * axis ** -s_ * np.exp(-axis / cut)
        if i % 6 == 5:
            state[i, :, n // 3: n // 2] = 0.0  # a gap
        if i % 8 == 7:
            state[i] = 0.0  # empty
    delta = np.zeros((ncase, 3, 2, n))
    edens = np.zeros((ncase, 3, 2))
    edens_of = np.zeros((ncase, 3, 2))
    for i in range(ncase):
        for kind in range(3):
            for sp in range(2):
                row = np.ascontiguousarray(state[i, sp])
                out = np.zeros(n)
                edens[i, kind, sp] = lib.ref_dsa_reacc(kind, sp, P(row), d(v1[i]), d(v2[i]), d(q[i]), d(r[i]),
                                                       d(pmin[i]), P(out))
                delta[i, kind, sp] = out
                edens_of[i, kind, sp] = lib.ref_dsa_reacc_edens(sp, P(out), d(pmin[i]))
    ms = np.concatenate([np.linspace(1.0, 3.0, 41), np.linspace(3.0, 40.0, 75)])
    xs = np.array([-0.1, 0.0, 0.01, 0.03, 0.05, 0.1, 0.3, 1.0])
    pairs = np.array([(0, 1), (2, 3), (4, 5), (6, 7), (6, 1), (0, 0)], np.int32)
    eta_i = np.array([[[lib.ref_dsa_eta_interpolated(d(m), int(a), int(b), d(x)) for x in xs] for m in ms]
                      for a, b in pairs])
    return dict(mach=mach, r=r, q=q, v1=v1, v2=v2, pmin=pmin, state=state, delta=delta, edens=edens,
                edens_of=edens_of, eta_mach=ms, eta_x=xs, eta_pairs=pairs, eta_interp=eta_i)


def main():
    if not os.path.isdir("/root/reference"):
        print("reference sources absent; nothing to do")
        return 0
    build_ref()
    ckpt_layout()
    if "--layout-only" in sys.argv:
        return 0
    drng = np.random.default_rng(8080)
    for n in (64, 256):
        np.savez_compressed(os.path.join(HERE, f"dsa_n{n}.npz"), **dsa_fixture(n, drng))
    rrng = np.random.default_rng(8181)
    for n in (64, 256):
        np.savez_compressed(os.path.join(HERE, f"dsa_reacc_n{n}.npz"), **reacc_fixture(n, rrng))
    if "--dsa-only" in sys.argv:
        return 0
    logx, logfx, fx, xmin = fx_table()
    np.savez_compressed(os.path.join(HERE, "fx_table.npz"), logx=logx, logfx=logfx, fx=fx,
                        xmin=np.array([xmin]))
    rng = np.random.default_rng(20261202)
    for n in SIZES:
        out = make_size(n, logx, logfx, fx, xmin, rng)
        np.savez_compressed(os.path.join(HERE, f"ref_n{n}.npz"), **out)
        print("wrote", n)
    return 0


if __name__ == "__main__":
    sys.exit(main())
