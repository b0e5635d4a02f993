"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

DPP_AUTO, DPP_ASA, DPP_TTD, DPP_OFF, DPP_DIRECT = 0, 1, 2, 3, 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)


class OrcParams(C.Structure):
    _fields_ = [
        ("dpp_model", C.c_int),
        ("ttd_tacc_model", C.c_int),
        ("reacc_window_mode", C.c_int),
        ("tracer_nsub_max", C.c_int),
        ("coeff_interp_min_steps", C.c_int),
        ("coeff_interp_max_segments", C.c_int),
        ("psi", C.c_double),
        ("mach_limit", C.c_double),
        ("eta_dpp_cap", C.c_double),
        ("t_acc_direct_gyr", C.c_double),
        ("t_off_gyr", C.c_double),
        ("t_acc_off_gyr", C.c_double),
        ("f_eddy", C.c_double),
        ("tracer_nsub_safety", C.c_double),
        ("epmax", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_ctx_create.restype = C.c_void_p
        _lib.orc_ctx_create.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        _lib.orc_ctx_free.argtypes = [C.c_void_p]
        for name in ("orc_b_coulomb_p", "orc_b_coulomb_e", "orc_b_synch", "orc_tacc_asa",
                     "orc_tacc_ttd", "orc_pp_pion_cross_section", "orc_qe_kernel_bb05",
                     "orc_gamma_kernel", "orc_nu_kernel", "orc_ic_emissivity_kernel"):
            getattr(_lib, name).restype = C.c_double
        _lib.orc_b_coulomb_p.argtypes = [C.c_double] * 3
        _lib.orc_b_coulomb_e.argtypes = [C.c_double] * 2
        _lib.orc_b_synch.argtypes = [C.c_double] * 2
        _lib.orc_tacc_asa.argtypes = [C.c_double] * 5
        _lib.orc_tacc_ttd.argtypes = [C.c_double] * 4 + [C.c_int]
        _lib.orc_pp_pion_cross_section.argtypes = [C.c_int, C.c_double]
        _lib.orc_qe_kernel_bb05.argtypes = [C.c_double] * 2
        _lib.orc_gamma_kernel.argtypes = [C.c_double] * 2
        _lib.orc_nu_kernel.argtypes = [C.c_double] * 2
        _lib.orc_ic_emissivity_kernel.argtypes = [C.c_double] * 3
        _lib.orc_ctx_j_pp.argtypes = [C.c_void_p]
        for name, args in (("orc_hadron_dsigma", [C.c_int, C.c_double, C.c_double]),
                           ("orc_hadron_fpi", [C.c_double] * 2), ("orc_hadron_fe", [C.c_double]),
                           ("orc_hadron_fnue", [C.c_double]), ("orc_hadron_Fe", [C.c_double] * 2),
                           ("orc_hadron_ekernel", [C.c_int, C.c_double, C.c_double]),
                           ("orc_hadron_numu1", [C.c_double] * 2)):
            getattr(_lib, name).restype = C.c_double
            getattr(_lib, name).argtypes = args
    return _lib


def _p(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    if a.dtype == np.uint8:
        return a.ctypes.data_as(_up)
    raise TypeError(a.dtype)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def default_params(**kw) -> OrcParams:
    prm = OrcParams()
    lib().orc_default_params(C.byref(prm))
    for k, v in kw.items():
        setattr(prm, k, v)
    return prm


class Oracle:
    """Grid context for one grid size n (both species share n)."""

    def __init__(self, n, pmin=-1.0, pmax=8.0, pemin=0.0, pemax=6.0):
        self.n = n
        self.h = lib().orc_ctx_create(n, pmin, pmax, pemin, pemax)
        if not self.h:
            raise RuntimeError("orc_ctx_create failed")
        self.pp, self.pdp, self.pE, self.pghost = self.grid(0)
        self.ep, self.edp, self.eE, self.eghost = self.grid(1)
        self.beta_p = np.zeros(n)
        lib().orc_ctx_beta_p(C.c_void_p(self.h), _p(self.beta_p))

    def __del__(self):
        try:
            if self.h:
                lib().orc_ctx_free(C.c_void_p(self.h))
        except Exception:
            pass

    def grid(self, species):
        n = self.n
        p, dp, E = np.zeros(n), np.zeros(n), np.zeros(n)
        m1, p1 = C.c_double(), C.c_double()
        lib().orc_ctx_grid(C.c_void_p(self.h), species, _p(p), _p(dp), _p(E), C.byref(m1), C.byref(p1))
        return p, dp, E, (m1.value, p1.value)

    @property
    def j_pp(self):
        return lib().orc_ctx_j_pp(C.c_void_p(self.h))

    def sigma_pp(self):
        s, g = np.zeros(self.n), np.zeros(self.n)
        lib().orc_ctx_sigma_pp(C.c_void_p(self.h), _p(s), _p(g))
        return s, g

    def losses_e(self, n_gas, b_field, divv, rad_ic, m1, p1):
        radp, tl = np.zeros(self.n), np.zeros(self.n)
        e0, e1 = C.c_double(), C.c_double()
        lib().orc_losses_e(C.c_void_p(self.h), C.c_double(n_gas), C.c_double(b_field),
                           C.c_double(divv), _p(f64(rad_ic)), C.c_double(m1), C.c_double(p1),
                           _p(radp), _p(tl), C.byref(e0), C.byref(e1))
        return radp, tl, e0.value, e1.value

    def losses_p(self, n_gas, kbt, divv):
        radp, tl = np.zeros(self.n), np.zeros(self.n)
        e0, e1 = C.c_double(), C.c_double()
        lib().orc_losses_p(C.c_void_p(self.h), C.c_double(n_gas), C.c_double(kbt), C.c_double(divv),
                           _p(radp), _p(tl), C.byref(e0), C.byref(e1))
        return radp, tl, e0.value, e1.value

    def dpp(self, prm, model, l_turb, dv, cs, beta, mass=0.0, crp=None, cre=None):
        d, de, ed = np.zeros(self.n), np.zeros(self.n), np.zeros(4)
        lib().orc_dpp(C.c_void_p(self.h), C.byref(prm), model, C.c_double(l_turb), C.c_double(dv),
                      C.c_double(cs), C.c_double(beta), C.c_double(mass),
                      _p(f64(crp)) if crp is not None else None,
                      _p(f64(cre)) if cre is not None else None, _p(d), _p(de), _p(ed))
        return d, de, ed

    def ic_cooling_row(self, z):
        row = np.zeros(self.n)
        m1, p1 = C.c_double(), C.c_double()
        lib().orc_ic_cooling_row(C.c_void_p(self.h), C.c_double(z), _p(row), C.byref(m1), C.byref(p1))
        return row, m1.value, p1.value

    def nsub(self, prm, dt_snap, n_gas, kbt, b_field, divv, ric_lo, ric_hi, l_turb, dv):
        nt = len(n_gas)
        lo_p, hi_p, lo_e, hi_e = (np.zeros(nt) for _ in range(4))
        lib().orc_nsub_endpoints(C.c_void_p(self.h), nt, _p(f64(n_gas)), _p(f64(kbt)), _p(f64(b_field)),
                                 _p(f64(divv)), C.c_double(ric_lo), C.c_double(ric_hi),
                                 _p(lo_p), _p(hi_p), _p(lo_e), _p(hi_e))
        ns, tg, non = (np.zeros(nt, np.int32) for _ in range(3))
        rc = lib().orc_nsub_compute(C.c_void_p(self.h), C.byref(prm), nt, C.c_double(dt_snap),
                                    _p(lo_p), _p(hi_p), _p(lo_e), _p(hi_e), _p(f64(l_turb)), _p(f64(dv)),
                                    _p(ns), _p(tg), _p(non))
        if rc != 0:
            raise RuntimeError("orc_nsub_compute failed")
        return ns, tg, non, np.stack([lo_p, hi_p, lo_e, hi_e])

    def choose_segments(self, prm, nsub, allow):
        return lib().orc_choose_segments(C.byref(prm), int(nsub), int(allow))

    def secondary_kernel(self):
        fqe = np.zeros(self.n * self.n)
        npmin = np.zeros(self.n, np.int32)
        lib().orc_secondary_kernel(C.c_void_p(self.h), _p(fqe), _p(npmin))
        return fqe, npmin

    def execute_snapshot(self, prm, nsub, n_on, allow_interp, elapsed, dt_snap, bg_curr, rad_ic_row,
                         ric_m1, ric_p1, bg_next, rad_ic_row_next, ric_m1n, ric_p1n, qpi, qepri,
                         fqe, npmin, mass, dis, crp, cre, nthreads=0):
        """bg_curr / bg_next: (8, ntracer) arrays in the order
        n_gas, kbt, b_field, divv, l_turb, dv, cs, beta_pl.  crp/cre updated in place."""
        nt = len(nsub)
        n = self.n
        ric_zero = np.ascontiguousarray(np.repeat(f64(rad_ic_row)[:, None], nt, axis=1))
        m1z = np.full(nt, ric_m1)
        p1z = np.full(nt, ric_p1)
        bgc = [f64(bg_curr[k]) for k in range(8)]
        bgn = [f64(bg_next[k]) for k in range(8)]
        arr_c = (_dp * 8)(*[_p(a) for a in bgc])
        arr_n = (_dp * 8)(*[_p(a) for a in bgn])
        assert crp.dtype == np.float64 and crp.flags.c_contiguous and crp.size == nt * n
        assert cre.dtype == np.float64 and cre.flags.c_contiguous and cre.size == nt * n
        rc = lib().orc_execute_snapshot(
            C.c_void_p(self.h), C.byref(prm), nt, _p(np.ascontiguousarray(nsub, np.int32)),
            _p(np.ascontiguousarray(n_on, np.int32)), int(allow_interp), C.c_double(elapsed),
            C.c_double(dt_snap), arr_c, _p(ric_zero), _p(m1z), _p(p1z), arr_n,
            _p(f64(rad_ic_row_next)), C.c_double(ric_m1n), C.c_double(ric_p1n), _p(f64(qpi)),
            _p(f64(qepri)), _p(f64(fqe)), _p(np.ascontiguousarray(npmin, np.int32)), _p(f64(mass)),
            _p(np.ascontiguousarray(dis, np.uint8)) if dis is not None else None,
            _p(crp), _p(cre), int(nthreads))
        if rc != 0:
            raise RuntimeError("orc_execute_snapshot failed")


def face_weights(w):
    d, wp, wm = C.c_double(), C.c_double(), C.c_double()
    lib().orc_face_weights(C.c_double(w), C.byref(d), C.byref(wp), C.byref(wm))
    return d.value, wp.value, wm.value


def cc_build(n, dt, x, dx, xm1, xp1, radp, dpp, tloss, radpm1, radpp1, dppm1, dppp1):
    a, b, c = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_cc_build(n, C.c_double(dt), _p(f64(x)), _p(f64(dx)), C.c_double(xm1), C.c_double(xp1),
                       _p(f64(radp)), _p(f64(dpp)), _p(f64(tloss)), C.c_double(radpm1),
                       C.c_double(radpp1), C.c_double(dppm1), C.c_double(dppp1), _p(a), _p(b), _p(c))
    return a, b, c


def thomas(a, b, c, rhs):
    r = f64(rhs).copy()
    rc = lib().orc_thomas(len(r), _p(f64(a)), _p(f64(b)), _p(f64(c)), _p(r))
    return rc, r


def pcr_batch(a, b, c, rhs):
    nsys, n = a.shape
    x = np.zeros_like(f64(rhs))
    rc = lib().orc_pcr_batch(nsys, n, _p(f64(a)), _p(f64(b)), _p(f64(c)), _p(f64(rhs)), _p(x))
    return rc, x


def secondary_source(n, n_gas, crp, npmin, fqe, qepri):
    nt = len(n_gas)
    inje = np.zeros(nt * n)
    lib().orc_secondary_source(n, nt, _p(f64(n_gas)), _p(f64(crp)), _p(np.ascontiguousarray(npmin, np.int32)),
                               _p(f64(fqe)), _p(f64(qepri)) if qepri is not None else None, _p(inje))
    return inje


def scalar(name, *args):
    return getattr(lib(), name)(*args)


def ic_kernel_flat(o: Oracle, nbins, numin, numax, z):
    fic = np.zeros(nbins * o.n)
    lib().orc_ic_kernel_flat(C.c_void_p(o.h), nbins, C.c_double(numin), C.c_double(numax), C.c_double(z), _p(fic))
    return fic


def emission_ic(n, cre, dp, fic, nbins):
    nt = cre.size // n
    eps = np.zeros(nt * nbins)
    lib().orc_emission_ic(n, nt, nbins, _p(f64(cre)), _p(f64(dp)), _p(f64(fic)), _p(eps))
    return eps


def gamma_kernel_flat(o: Oracle, nbins, emin, emax):
    f = np.zeros(nbins * o.n)
    lib().orc_gamma_kernel_flat(C.c_void_p(o.h), nbins, C.c_double(emin), C.c_double(emax), _p(f))
    return f


def nu_kernel_flat(o: Oracle, nbins, emin, emax):
    f = np.zeros(nbins * o.n)
    lib().orc_nu_kernel_flat(C.c_void_p(o.h), nbins, C.c_double(emin), C.c_double(emax), _p(f))
    return f


def emission_hadronic(n, nbins, emin, emax, n_gas, crp, beta_p, dp, fk):
    nt = len(n_gas)
    eps = np.zeros(nt * nbins)
    lib().orc_emission_hadronic(n, nt, nbins, C.c_double(emin), C.c_double(emax), _p(f64(n_gas)), _p(f64(crp)),
                                _p(f64(beta_p)), _p(f64(dp)), _p(f64(fk)), _p(eps))
    return eps


def synch_logy(o: Oracle, nfreq, nu_min, nu_max, ntheta):
    logy = np.zeros(nfreq * o.n * ntheta)
    pw = np.zeros(ntheta)
    lib().orc_synch_logy(C.c_void_p(o.h), nfreq, C.c_double(nu_min), C.c_double(nu_max), ntheta, _p(logy), _p(pw))
    return logy, pw


def synch_kernel_table(n, nfreq, ntheta, nlogb, logb_min, inv_dlogb, logx, logfx, fx, xmin, logy, pw):
    table = np.zeros(nlogb * nfreq * n)
    lib().orc_synch_kernel_table(n, nfreq, ntheta, nlogb, C.c_double(logb_min), C.c_double(inv_dlogb), len(logx),
                                 C.c_double(xmin), _p(f64(logx)), _p(f64(logfx)), _p(f64(fx)), _p(f64(logy)),
                                 _p(f64(pw)), _p(table))
    return table


def emission_synch(n, nfreq, nlogb, logb_min, inv_dlogb, table, cre, dp, b_field):
    nt = len(b_field)
    eps = np.zeros(nt * nfreq)
    lib().orc_emission_synch(n, nt, nfreq, nlogb, C.c_double(logb_min), C.c_double(inv_dlogb), _p(f64(table)),
                             _p(f64(cre)), _p(f64(dp)), _p(f64(b_field)), _p(eps))
    return eps
