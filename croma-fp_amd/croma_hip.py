"""ctypes binding of libcroma_hip.so (the C ABI declared in include/croma_hip.h).

This is the Python-side mirror of the reference's C entry points for the
per-tracer Fokker-Planck update; the arithmetic runs in the HIP library.
Importing this module loads the in-tree library and raises if it is missing:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcroma_hip.so")

OK, EINVAL, EPIVOT, ENOMEM, EDEVICE, ENONFINITE, EBUDGET, ECANCELED, EUNSUPPORTED, EBUSY = (
    0, -1, -2, -3, -4, -5, -6, -7, -8, -9)
PATH_FUSED, PATH_PLAIN = 0, 1
SPECIES_ELECTRON, SPECIES_PROTON, SPECIES_ELECTRON_PROTON = 1, 2, 3
DPP_ASA, DPP_TTD, DPP_OFF, DPP_DIRECT = 1, 2, 3, 4
EMIS_FP32, EMIS_MFMA_XF32, EMIS_MFMA_BF16, EMIS_FP64 = 0, 1, 2, 3
TABLE_GAMMA, TABLE_NEUTRINO, TABLE_IC = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)


class CromaHipError(RuntimeError):
    def __init__(self, rc, what):
        super().__init__(f"{what} failed (rc={rc}): {last_error()}")
        self.rc = rc


class GridSpec(C.Structure):
    _fields_ = [("n", C.c_int), ("pmin", C.c_double), ("pmax", C.c_double),
                ("pemin", C.c_double), ("pemax", C.c_double)]


class Params(C.Structure):
    _fields_ = [
        ("dpp_model", C.c_int), ("ttd_tacc_model", C.c_int), ("reacc_window_mode", C.c_int),
        ("tracer_nsub_max", C.c_int), ("coeff_interp_min_steps", C.c_int),
        ("coeff_interp_max_segments", C.c_int), ("psi", C.c_double), ("mach_limit", C.c_double),
        ("eta_dpp_cap", C.c_double), ("t_acc_direct_gyr", C.c_double), ("t_off_gyr", C.c_double),
        ("t_acc_off_gyr", C.c_double), ("f_eddy", C.c_double), ("tracer_nsub_safety", C.c_double),
        ("epmax", C.c_double),
    ]


class Tile(C.Structure):
    _fields_ = [("rows_per_lane", C.c_int), ("waves_per_simd", C.c_int), ("chain", C.c_int),
                ("secondary_gemm", C.c_int)]


class Times(C.Structure):
    _fields_ = [
        ("nsub_estimate_ms", C.c_double), ("bucket_build_ms", C.c_double), ("ic_rows_ms", C.c_double),
        ("h2d_ms", C.c_double), ("solve_ms", C.c_double), ("d2h_ms", C.c_double), ("total_ms", C.c_double),
        ("cell_steps", C.c_longlong), ("species_steps", C.c_longlong), ("kernel_ms", C.c_double),
        ("kernel_launches", C.c_longlong), ("graph_launches", C.c_longlong),
        ("graph_captures", C.c_longlong), ("graph_failures", C.c_longlong),
    ]


class Chunk(C.Structure):
    _fields_ = [
        ("ntracer", C.c_int), ("n", C.c_int), ("species", C.c_int), ("dpp_model", C.c_int), ("nsnap", C.c_int),
        ("z", _dp), ("dt_snap", _dp), ("bg", _dp), ("crp", _dp), ("cre", _dp), ("qpi", _dp), ("qepri", _dp),
        ("mass", _dp),
    ]


class ChunkStats(C.Structure):
    _fields_ = [("submit_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
                ("cell_steps", C.c_longlong), ("status", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("cancelled", C.c_longlong),
                ("failed", C.c_longlong), ("cell_steps", C.c_longlong), ("budget_bytes", C.c_size_t),
                ("in_use_bytes", C.c_size_t), ("peak_bytes", C.c_size_t), ("active", C.c_int),
                ("queued", C.c_int), ("iterations", C.c_longlong)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.croma_hip_last_error.restype = C.c_char_p
        L.croma_hip_context_create.restype = C.c_void_p
        L.croma_hip_context_create.argtypes = [C.POINTER(GridSpec), C.POINTER(Params)]
        for name in ("croma_hip_context_destroy", "croma_hip_batch_destroy"):
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = None
        L.croma_hip_batch_create.restype = C.c_void_p
        L.croma_hip_batch_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.croma_hip_synch_plan_create.restype = C.c_void_p
        L.croma_hip_synch_plan_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.croma_hip_synch_plan_run.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.croma_hip_synch_plan_destroy.argtypes = [C.c_void_p]
        L.croma_hip_synch_plan_destroy.restype = None
        _lib = L
    return _lib


def last_error() -> str:
    try:
        v = lib().croma_hip_last_error()
        return v.decode() if v else ""
    except Exception:  # pragma: no cover
        return ""


def _chk(rc, what):
    if rc != 0:
        raise CromaHipError(rc, what)


def _p(a):
    if a is None:
        return None
    assert a.flags.c_contiguous, "arrays must be C-contiguous"
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    if a.dtype == np.uint8:
        return a.ctypes.data_as(_up)
    raise TypeError(a.dtype)


def f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def default_params(**kw) -> Params:
    p = Params()
    lib().croma_hip_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def is_available() -> bool:
    return bool(lib().croma_hip_is_available())


def device_name() -> str:
    buf = C.create_string_buffer(256)
    _chk(lib().croma_hip_get_current_device_id(buf, 256), "device id")
    return buf.value.decode()


class SynchPlan:
    """Synchrotron on MFMA with the table's device form built once
    (croma_hip_synch_plan_*)."""

    def __init__(self, ctx, table, nfreq, nlogb, logb_min, inv_dlogb):
        ms = C.c_double()
        self.n, self.nfreq = ctx.n, int(nfreq)
        self._table = f64(table)
        self.h = lib().croma_hip_synch_plan_create(C.c_void_p(ctx.h), self.nfreq, int(nlogb), C.c_double(logb_min),
                                                   C.c_double(inv_dlogb), _p(self._table), C.byref(ms))
        if not self.h:
            raise CromaHipError(-1, "synch_plan_create")
        self.prepare_ms = ms.value

    def run(self, cre, b_field):
        """Returns (eps [ntr, nfreq], device ms of the run)."""
        cre = f64(cre)
        nt = len(b_field)
        eps = np.zeros(nt * self.nfreq)
        ms = C.c_double()
        _chk(lib().croma_hip_synch_plan_run(C.c_void_p(self.h), nt, _p(cre), _p(f64(b_field)), _p(eps), C.byref(ms)),
             "synch_plan_run")
        return eps.reshape(nt, self.nfreq), ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().croma_hip_synch_plan_destroy(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        self.close()


class Context:
    """Device, grid tables and streams for one grid size n."""

    def __init__(self, n, params=None, pmin=-1.0, pmax=8.0, pemin=0.0, pemax=6.0):
        gs = GridSpec(n, pmin, pmax, pemin, pemax)
        self.n = n
        self.params = params if params is not None else default_params()
        self.h = lib().croma_hip_context_create(C.byref(gs), C.byref(self.params))
        if not self.h:
            raise CromaHipError(-1, "context create")

    def close(self):
        if getattr(self, "h", None):
            lib().croma_hip_context_destroy(C.c_void_p(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        self.params = params
        _chk(lib().croma_hip_context_set_params(C.c_void_p(self.h), C.byref(params)), "set_params")

    def set_tile(self, rows_per_lane, waves_per_simd, chain=1, protons=-1, secondary_gemm=1):
        """protons: 0 electron_only tile, 1 electron_proton tile, -1 both."""
        t = Tile(rows_per_lane, waves_per_simd, chain, secondary_gemm)
        _chk(lib().croma_hip_context_set_tile(C.c_void_p(self.h), int(protons), C.byref(t)), "set_tile")

    def set_coeff_precision(self, fp32):
        """FP32 (True) or FP64 (False, default) coefficient evaluation of the fused path."""
        _chk(lib().croma_hip_context_set_coeff_precision(C.c_void_p(self.h), 1 if fp32 else 0),
             "set_coeff_precision")

    def set_graphs(self, enable):
        """True: repeated snapshot windows replay a captured hipGraph of their launches."""
        _chk(lib().croma_hip_context_set_graphs(C.c_void_p(self.h), int(bool(enable))), "set_graphs")

    def set_pipeline(self, enable):
        """True: split windows run the electron pass beside the next window's proton pass."""
        _chk(lib().croma_hip_context_set_pipeline(C.c_void_p(self.h), int(bool(enable))), "set_pipeline")

    def get_tile(self, protons=False):
        t = Tile()
        _chk(lib().croma_hip_context_get_tile(C.c_void_p(self.h), int(protons), C.byref(t)), "get_tile")
        return (t.rows_per_lane, t.waves_per_simd, t.chain, t.secondary_gemm)

    def autotune(self, protons=False, ntracer=16384):
        t = Tile()
        _chk(lib().croma_hip_autotune(C.c_void_p(self.h), int(protons), int(ntracer), C.byref(t)), "autotune")
        return (t.rows_per_lane, t.waves_per_simd, t.chain, t.secondary_gemm)

    def grid(self, species):
        n = self.n
        p, dp, E = np.zeros(n), np.zeros(n), np.zeros(n)
        m1, p1 = C.c_double(), C.c_double()
        _chk(lib().croma_hip_context_grid(C.c_void_p(self.h), species, _p(p), _p(dp), _p(E), C.byref(m1),
                                          C.byref(p1)), "grid")
        return p, dp, E, (m1.value, p1.value)

    def ic_cooling_row(self, z):
        row = np.zeros(self.n)
        m1, p1 = C.c_double(), C.c_double()
        _chk(lib().croma_hip_ic_cooling_row(C.c_void_p(self.h), C.c_double(z), _p(row), C.byref(m1), C.byref(p1)),
             "ic_cooling_row")
        return row, m1.value, p1.value

    def ic_cooling_rows_device(self, zs):
        zs = f64(zs)
        rows = np.zeros((len(zs), self.n))
        gh = np.zeros((len(zs), 2))
        _chk(lib().croma_hip_ic_cooling_rows_device(C.c_void_p(self.h), len(zs), _p(zs), _p(rows), _p(gh)),
             "ic_cooling_rows_device")
        return rows, gh

    def emission_ic(self, cre, fic, nbins, mode=EMIS_MFMA_XF32):
        """eps[i][k] = sum_j fic[k][j] cre[i][j] dp[j]; returns (eps [ntr, nbins], device ms)."""
        cre = f64(cre)
        nt = cre.size // self.n
        eps = np.zeros(nt * nbins)
        ms = C.c_double()
        _chk(lib().croma_hip_emission_ic(C.c_void_p(self.h), nt, int(nbins), _p(cre), _p(f64(fic)), _p(eps),
                                         int(mode), C.byref(ms)), "emission_ic")
        return eps.reshape(nt, nbins), ms.value

    def emission_hadronic(self, crp, n_gas, fk, nbins, emin, emax, mode=EMIS_MFMA_XF32):
        crp = f64(crp)
        nt = len(n_gas)
        eps = np.zeros(nt * nbins)
        ms = C.c_double()
        _chk(lib().croma_hip_emission_hadronic(C.c_void_p(self.h), nt, int(nbins), C.c_double(emin),
                                               C.c_double(emax), _p(f64(n_gas)), _p(crp), _p(f64(fk)), _p(eps),
                                               int(mode), C.byref(ms)), "emission_hadronic")
        return eps.reshape(nt, nbins), ms.value

    def synch_kernel_table(self, nfreq, ntheta, nlogb, logb_min, inv_dlogb, logx, logfx, fx, xmin, logy, pw):
        """Pitch-kernel table [nlogb * nfreq * n] built on the device; returns (table, device ms)."""
        table = np.zeros(nlogb * nfreq * self.n)
        ms = C.c_double()
        _chk(lib().croma_hip_synch_kernel_table(C.c_void_p(self.h), int(nfreq), int(ntheta), int(nlogb),
                                                C.c_double(logb_min), C.c_double(inv_dlogb), len(logx),
                                                C.c_double(xmin), _p(f64(logx)), _p(f64(logfx)), _p(f64(fx)),
                                                _p(f64(logy)), _p(f64(pw)), _p(table), C.byref(ms)),
             "synch_kernel_table")
        return table, ms.value

    def emission_synch(self, cre, b_field, table, nfreq, nlogb, logb_min, inv_dlogb, mode=EMIS_FP64):
        nt = len(b_field)
        eps = np.zeros(nt * nfreq)
        ms = C.c_double()
        _chk(lib().croma_hip_emission_synch_mode(C.c_void_p(self.h), nt, int(nfreq), int(nlogb),
                                                 C.c_double(logb_min), C.c_double(inv_dlogb), _p(f64(table)),
                                                 _p(f64(cre)), _p(f64(b_field)), _p(eps), int(mode), C.byref(ms)),
             "emission_synch")
        return eps.reshape(nt, nfreq), ms.value

    def nsub(self, dt_snap, n_gas, kbt, b_field, divv, ric_lo, ric_hi, l_turb, dv):
        nt = len(n_gas)
        ns = np.zeros(nt, np.int32)
        non = np.zeros(nt, np.int32)
        _chk(lib().croma_hip_nsub_estimate(C.c_void_p(self.h), nt, C.c_double(dt_snap), _p(f64(n_gas)),
                                           _p(f64(kbt)), _p(f64(b_field)), _p(f64(divv)), C.c_double(ric_lo),
                                           C.c_double(ric_hi), _p(f64(l_turb)), _p(f64(dv)), _p(ns), _p(non)),
             "nsub")
        return ns, non

    def nsub_device(self, dt_snap, n_gas, kbt, b_field, divv, ric_lo, ric_hi, l_turb, dv):
        """The nsub estimate computed on the device."""
        nt = len(n_gas)
        ns = np.zeros(nt, np.int32)
        non = np.zeros(nt, np.int32)
        _chk(lib().croma_hip_nsub_estimate_device(C.c_void_p(self.h), nt, C.c_double(dt_snap), _p(f64(n_gas)),
                                                  _p(f64(kbt)), _p(f64(b_field)), _p(f64(divv)), C.c_double(ric_lo),
                                                  C.c_double(ric_hi), _p(f64(l_turb)), _p(f64(dv)), _p(ns), _p(non)),
             "nsub_device")
        return ns, non

    def set_nsub_device(self, enable):
        """True: batch runs estimate nsub and order the queue on the device."""
        _chk(lib().croma_hip_context_set_nsub_device(C.c_void_p(self.h), int(bool(enable))), "set_nsub_device")

    def emission_kernel_table(self, kind, nbins, lo, hi, z=0.0):
        """Gamma / neutrino / IC kernel rows [nbins, n] built on the device
        (kind TABLE_GAMMA, TABLE_NEUTRINO or TABLE_IC); returns (table, device ms)."""
        table = np.zeros(int(nbins) * self.n)
        ms = C.c_double()
        _chk(lib().croma_hip_emission_kernel_table(C.c_void_p(self.h), int(kind), int(nbins), C.c_double(lo),
                                                   C.c_double(hi), C.c_double(z), _p(table), C.byref(ms)),
             "emission_kernel_table")
        return table.reshape(int(nbins), self.n), ms.value

    def secondary_weights(self):
        fqe = np.zeros(self.n * self.n)
        npmin = np.zeros(self.n, np.int32)
        _chk(lib().croma_hip_secondary_weights(C.c_void_p(self.h), _p(fqe), _p(npmin)), "secondary_weights")
        return fqe, npmin

    def execute_snapshot(self, nsub, n_on, allow_interp, elapsed, dt_snap, bg_curr, rad_ic_row, ric_m1, ric_p1,
                         bg_next, rad_ic_row_next, ric_m1n, ric_p1n, qpi, qepri, fqe, npmin, mass, dis, crp, cre,
                         t_acc_direct=0.0):
        """Drop-in for tracer_fp_batch_execute_parallel (host arrays; crp/cre updated in place)."""
        nt = len(nsub)
        n = self.n
        ns = np.ascontiguousarray(nsub, np.int32)
        no = np.ascontiguousarray(n_on, np.int32)
        cad = np.ones(nt, np.int32)
        ric_zero = np.ascontiguousarray(np.repeat(f64(rad_ic_row)[:, None], nt, axis=1))
        m1z, p1z = np.full(nt, ric_m1), np.full(nt, ric_p1)
        bgc = [f64(bg_curr[k]) for k in range(8)]
        bgn = [f64(bg_next[k]) for k in range(8)]
        times = Times()
        bc, mn, mx = C.c_int(), C.c_int(), C.c_int()
        assert crp.size == nt * n and cre.size == nt * n
        rc = lib().croma_hip_execute_snapshot(
            C.c_void_p(self.h), nt, _p(ns), _p(no), _p(ns), _p(cad), 0, int(allow_interp), C.c_double(elapsed),
            C.c_double(dt_snap), C.c_double(0.0), C.c_double(0.0), C.c_double(t_acc_direct),
            *[_p(a) for a in bgc], _p(ric_zero), _p(m1z), _p(p1z), *[_p(a) for a in bgn],
            _p(f64(rad_ic_row_next)), C.c_double(ric_m1n), C.c_double(ric_p1n), _p(f64(qpi)), _p(f64(qepri)),
            _p(f64(fqe)), _p(np.ascontiguousarray(npmin, np.int32)), _p(f64(mass)),
            _p(np.ascontiguousarray(dis, np.uint8)) if dis is not None else None, _p(crp), _p(cre),
            C.byref(times), C.byref(bc), C.byref(mn), C.byref(mx))
        _chk(rc, "execute_snapshot")
        return times


class Batch:
    """Device-resident tracers (spectra, sources, backgrounds of nsnap snapshots)."""

    def __init__(self, ctx: Context, ntracer, species, nsnap):
        self.ctx, self.ntracer, self.species, self.nsnap = ctx, ntracer, species, nsnap
        self.h = lib().croma_hip_batch_create(C.c_void_p(ctx.h), ntracer, species, nsnap)
        if not self.h:
            raise CromaHipError(-1, "batch create")

    def close(self):
        if getattr(self, "h", None):
            lib().croma_hip_batch_destroy(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, crp, cre, qpi=None, qepri=None, mass=None, dis=None):
        _chk(lib().croma_hip_batch_set_state(C.c_void_p(self.h), _p(f64(crp)), _p(f64(cre)), _p(f64(qpi)),
                                             _p(f64(qepri)), _p(f64(mass)),
                                             _p(np.ascontiguousarray(dis, np.uint8)) if dis is not None else None),
             "set_state")

    def get_state(self):
        n = self.ctx.n
        crp = np.zeros(self.ntracer * n)
        cre = np.zeros(self.ntracer * n)
        _chk(lib().croma_hip_batch_get_state(C.c_void_p(self.h), _p(crp), _p(cre)), "get_state")
        return crp, cre

    def set_background(self, isnap, fields):
        arrs = [f64(fields[k]) for k in range(8)]
        arr = (_dp * 8)(*[_p(a) for a in arrs])
        _chk(lib().croma_hip_batch_set_background(C.c_void_p(self.h), int(isnap), arr), "set_background")

    def set_sources(self, isnap, qpi, qepri):
        _chk(lib().croma_hip_batch_set_sources(C.c_void_p(self.h), int(isnap), _p(f64(qpi)), _p(f64(qepri))),
             "set_sources")

    def run(self, s0, s1, nsnap_total, z, dt_snap, elapsed=0.0, path=PATH_FUSED):
        t = Times()
        _chk(lib().croma_hip_batch_run(C.c_void_p(self.h), int(s0), int(s1), int(nsnap_total), _p(f64(z)),
                                       _p(f64(dt_snap)), C.c_double(elapsed), int(path), C.byref(t)), "batch_run")
        return t

    def sync(self):
        _chk(lib().croma_hip_batch_sync(C.c_void_p(self.h)), "batch_sync")

    def checkpoint_write(self, ckpt_dir, meta, ids):
        ids = np.ascontiguousarray(ids, np.int32)
        _chk(_out_lib().croma_hip_batch_checkpoint_write(C.c_void_p(self.h), str(ckpt_dir).encode(), C.byref(meta),
                                                         _p(ids)), "batch_checkpoint_write")

    def checkpoint_read(self, ckpt_dir, meta, ids):
        """Loads the checkpoint's spectra into the batch; returns the completed snapshot."""
        ids = np.ascontiguousarray(ids, np.int32)
        done = C.c_int()
        _chk(_out_lib().croma_hip_batch_checkpoint_read(C.c_void_p(self.h), str(ckpt_dir).encode(), C.byref(meta),
                                                        _p(ids), C.byref(done)), "batch_checkpoint_read")
        return done.value

    def autotune(self, isnap, nsnap_total, z, dt_snap, elapsed=0.0):
        """Tile autotuned on snapshot isnap of this batch (state restored); returns the tile tuple."""
        t = Tile()
        _chk(lib().croma_hip_batch_autotune(C.c_void_p(self.h), int(isnap), int(nsnap_total), _p(f64(z)),
                                            _p(f64(dt_snap)), C.c_double(elapsed), C.byref(t)), "batch_autotune")
        return (t.rows_per_lane, t.waves_per_simd, t.chain, t.secondary_gemm)


def solve_cc_batch(a, b, c, source, state, dt):
    nsys, nrow = a.shape
    out = np.zeros_like(f64(state))
    ms = C.c_double()
    _chk(lib().croma_hip_solve_cc_batch(nsys, nrow, C.c_double(dt), _p(f64(a)), _p(f64(b)), _p(f64(c)),
                                        _p(f64(source)), _p(f64(state)), _p(out), C.byref(ms)), "solve_cc_batch")
    return out, ms.value


def solve_cc_batch_noalloc(nsys, nrow, dt, a_ptr, b_ptr, c_ptr, source_ptr, state_in_ptr, state_out_ptr, stream=0):
    """Allocation-free batched solve on caller-owned DEVICE arrays (addresses
    as integers, e.g. torch tensors' data_ptr()); returns the library's code
    (0, or raises on error)."""
    L = lib()
    fn = L.croma_hip_solve_cc_batch_noalloc
    fn.argtypes = [C.c_int, C.c_int, C.c_double] + [C.c_void_p] * 6 + [C.c_void_p]
    fn.restype = C.c_int
    _chk(fn(nsys, nrow, dt, a_ptr, b_ptr, c_ptr, source_ptr, state_in_ptr, state_out_ptr, stream or None),
         "solve_cc_batch_noalloc")


NOALLOC_MAX_CELLS = 4 * 1024 * 1024


class Server:
    """Request loop (croma_hip_serve_*): chunks of tracers are submitted at any
    time, admitted against an HBM budget and advanced together on the device."""

    def __init__(self, budget_bytes, params=None, pmin=-1.0, pmax=8.0, pemin=0.0, pemax=6.0):
        L = lib()
        L.croma_hip_serve_create.restype = C.c_void_p
        L.croma_hip_serve_create.argtypes = [C.POINTER(Params), C.c_size_t, C.c_double, C.c_double, C.c_double,
                                             C.c_double]
        L.croma_hip_serve_destroy.argtypes = [C.c_void_p]
        L.croma_hip_serve_destroy.restype = None
        L.croma_hip_serve_submit.argtypes = [C.c_void_p, C.POINTER(Chunk), C.POINTER(C.c_longlong)]
        L.croma_hip_serve_poll.argtypes = [C.c_void_p, C.c_longlong]
        L.croma_hip_serve_wait.argtypes = [C.c_void_p, C.c_longlong, _dp, _dp, C.POINTER(ChunkStats)]
        L.croma_hip_serve_cancel.argtypes = [C.c_void_p, C.c_longlong]
        L.croma_hip_serve_stats.argtypes = [C.c_void_p, C.POINTER(ServeStats)]
        L.croma_hip_chunk_bytes.restype = C.c_size_t
        L.croma_hip_chunk_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        self.params = params if params is not None else default_params()
        self.h = L.croma_hip_serve_create(C.byref(self.params), int(budget_bytes), pmin, pmax, pemin, pemax)
        if not self.h:
            raise CromaHipError(-1, "serve_create")
        self.shapes = {}

    def close(self):
        if getattr(self, "h", None):
            lib().croma_hip_serve_destroy(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, n, species, dpp_model, z, dt_snap, bg, cre=None, crp=None, qpi=None, qepri=None, mass=None):
        """bg: [(nsnap + 1), 8, ntracer]; spectra [ntracer * n]. Returns the request id."""
        z, dt_snap, bg = f64(z), f64(dt_snap), f64(bg)
        nsnap = len(dt_snap)
        ntr = bg.shape[2]
        assert bg.shape == (nsnap + 1, 8, ntr) and len(z) == nsnap + 1
        arrs = [None if a is None else f64(a) for a in (crp, cre, qpi, qepri, mass)]
        ch = Chunk(ntr, n, species, dpp_model, nsnap, _p(z), _p(dt_snap), _p(bg), *[_p(a) for a in arrs])
        rid = C.c_longlong()
        _chk(lib().croma_hip_serve_submit(C.c_void_p(self.h), C.byref(ch), C.byref(rid)), "serve_submit")
        self.shapes[rid.value] = (ntr, n)
        return rid.value

    def poll(self, rid):
        return lib().croma_hip_serve_poll(C.c_void_p(self.h), C.c_longlong(rid))

    def wait(self, rid):
        """Blocks until the request is done; returns (crp, cre, stats dict)."""
        ntr, n = self.shapes[rid]
        crp = np.zeros(ntr * n)
        cre = np.zeros(ntr * n)
        st = ChunkStats()
        rc = lib().croma_hip_serve_wait(C.c_void_p(self.h), C.c_longlong(rid), _p(crp), _p(cre), C.byref(st))
        stats = {"submit_s": st.submit_s, "admit_s": st.admit_s, "done_s": st.done_s, "cell_steps": st.cell_steps,
                 "status": st.status, "latency_s": st.done_s - st.submit_s}
        _chk(rc, "serve_wait")
        return crp, cre, stats

    def cancel(self, rid):
        _chk(lib().croma_hip_serve_cancel(C.c_void_p(self.h), C.c_longlong(rid)), "serve_cancel")

    def stats(self):
        st = ServeStats()
        _chk(lib().croma_hip_serve_stats(C.c_void_p(self.h), C.byref(st)), "serve_stats")
        return {k: getattr(st, k) for k, _ in ServeStats._fields_}


OUT_MAP, OUT_TILE = 0, 1

DSA_STATE, DSA_SOURCE = 1, 2
DSA_REACC_OFF, DSA_REACC_POSITIVE_DELTA, DSA_REACC_DRURY83 = 0, 1, 2
DSA_REACC_KIND_POSITIVE, DSA_REACC_KIND_DRURY, DSA_REACC_KIND_ADIABATIC, DSA_REACC_KIND_ENERGY = 0, 1, 2, 3
DSA_KR07, DSA_KR07R, DSA_KR13, DSA_KR13R, DSA_CS14, DSA_CS14R, DSA_RYU19, DSA_RYU19R = range(8)


class DsaParams(C.Structure):
    _fields_ = [("eta_model", C.c_int), ("inject_species", C.c_int), ("chi_p", C.c_double), ("chi_e", C.c_double),
                ("k_ep", C.c_double), ("pmax_p", C.c_double), ("pmax_e", C.c_double), ("gamma_gas", C.c_double),
                ("min_mach", C.c_double), ("mu_mol", C.c_double), ("pmin_dsa_p_log10", C.c_double),
                ("pmin_dsa_e_log10", C.c_double), ("reacc_mode", C.c_int), ("eta_model_reacc", C.c_int),
                ("reacc_eta_cap", C.c_double), ("xcr_pmin_pmc", C.c_double)]


def dsa_default_params(**kw):
    p = DsaParams()
    lib().croma_hip_dsa_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def dsa_eta(mach, model):
    L = lib()
    L.croma_hip_dsa_eta.restype = C.c_double
    L.croma_hip_dsa_eta.argtypes = [C.c_double, C.c_int]
    return L.croma_hip_dsa_eta(float(mach), int(model))


def dsa_shock_sources(n, params, mach, temp1, rho1, u1=0.0, pmin=-1.0, pmax=8.0, pemin=0.0, pemax=6.0):
    """Source rows (4 pi p^2 f per unit volume) of one shock: (proton, electron, flags)."""
    gs = GridSpec(n, pmin, pmax, pemin, pemax)
    sp, se = np.zeros(n), np.zeros(n)
    fl = (C.c_int * 2)()
    _chk(lib().croma_hip_dsa_shock_sources(C.byref(gs), C.byref(params), C.c_double(mach), C.c_double(temp1),
                                           C.c_double(rho1), C.c_double(u1), _p(sp), _p(se), fl),
         "dsa_shock_sources")
    return sp, se, (fl[0], fl[1])


def dsa_inject(n, params, mode, mach, rho1, mass, temp1=None, u1=None, flux=None, side=None, dt_gyr=0.0,
               crp=None, cre=None, qpi=None, qepri=None, pmin=-1.0, pmax=8.0, pemin=0.0, pemax=6.0, dis=None):
    """Adds one snapshot's DSA injection (and, with params.reacc_mode, the
    reacceleration of crp / cre) in place; returns the injected count.  dis:
    uint8 [ntracer] receiving the Drury83 tracers' disable_adiabatic flags."""
    gs = GridSpec(n, pmin, pmax, pemin, pemax)
    nt = len(mach)
    ninj = C.c_int()
    opt = lambda a: _p(f64(a)) if a is not None else None
    for a in (crp, cre, qpi, qepri):
        assert a is None or (a.dtype == np.float64 and a.flags.c_contiguous and a.size == nt * n)
    assert dis is None or (dis.dtype == np.uint8 and dis.size == nt)
    _chk(lib().croma_hip_dsa_inject_reacc(C.byref(gs), C.byref(params), int(mode), nt, C.c_double(dt_gyr),
                                          _p(f64(mach)), opt(temp1), _p(f64(rho1)), opt(u1), opt(flux), opt(side),
                                          _p(f64(mass)), _p(crp) if crp is not None else None,
                                          _p(cre) if cre is not None else None, _p(qpi) if qpi is not None else None,
                                          _p(qepri) if qepri is not None else None, _p(dis) if dis is not None else None,
                                          C.byref(ninj)), "dsa_inject")
    return ninj.value


def dsa_reacc_delta(n, kind, species, state, v1, v2, q, r, pmin_mc, pmin=-1.0, pmax=8.0, pemin=0.0, pemax=6.0):
    """One reacceleration row of a spectrum (croma_hip_dsa_reacc_delta): (delta, energy density)."""
    gs = GridSpec(n, pmin, pmax, pemin, pemax)
    delta = np.zeros(n)
    ed = C.c_double()
    _chk(lib().croma_hip_dsa_reacc_delta(C.byref(gs), int(kind), int(species), _p(f64(state)), C.c_double(v1),
                                         C.c_double(v2), C.c_double(q), C.c_double(r), C.c_double(pmin_mc),
                                         _p(delta), C.byref(ed)), "dsa_reacc_delta")
    return delta, ed.value


def dsa_eta_interpolated(mach, initial_model, reacc_model, x_cr):
    L = lib()
    L.croma_hip_dsa_eta_interpolated.restype = C.c_double
    L.croma_hip_dsa_eta_interpolated.argtypes = [C.c_double, C.c_int, C.c_int, C.c_double]
    return L.croma_hip_dsa_eta_interpolated(float(mach), int(initial_model), int(reacc_model), float(x_cr))


class CheckpointMeta(C.Structure):
    _fields_ = [("mpi_rank", C.c_int), ("world_size", C.c_int), ("ntracer", C.c_int), ("ntracer_global", C.c_int),
                ("nsnap", C.c_int), ("nfreq", C.c_int), ("np_bins", C.c_int), ("npe_bins", C.c_int),
                ("nsnp_i", C.c_int), ("nsnp_f", C.c_int), ("completed_snapshot", C.c_int),
                ("tracer_start", C.c_long)]


def _out_lib():
    L = lib()
    if not getattr(L, "_out_ready", False):
        L.croma_hip_output_open.restype = C.c_void_p
        L.croma_hip_output_open.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int]
        L.croma_hip_output_write.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.croma_hip_output_close.argtypes = [C.c_void_p]
        L.croma_hip_batch_write_spectra.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _dp, _dp]
        L.croma_hip_write_tracer_ids.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int)]
        ip = C.POINTER(C.c_int)
        L.croma_hip_checkpoint_write.argtypes = [C.c_char_p, C.POINTER(CheckpointMeta), ip, _dp, _dp]
        L.croma_hip_checkpoint_read.argtypes = [C.c_char_p, C.POINTER(CheckpointMeta), ip, _dp, _dp, ip]
        L.croma_hip_batch_checkpoint_write.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(CheckpointMeta), ip]
        L.croma_hip_batch_checkpoint_read.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(CheckpointMeta), ip, ip]
        L._out_ready = True
    return L


class Output:
    """A spectra / emissivity file in the reference layout (OUT_MAP: [tracer][slot][nbin]
    as <stem>_coreNN.bin; OUT_TILE: [slot][tracer][nbin] as <stem>_tile_coreNN.bin)."""

    def __init__(self, out_dir, stem, rank, mode, ntracer, nbin, nslot, resume=False):
        self.ntracer, self.nbin, self.nslot = ntracer, nbin, nslot
        self.h = _out_lib().croma_hip_output_open(str(out_dir).encode(), stem.encode(), int(rank), int(mode),
                                                  int(ntracer), int(nbin), int(nslot), int(bool(resume)))
        if not self.h:
            raise CromaHipError(-10, f"output_open: {last_error()}")

    def write(self, slot, rows, rho=None, mass=None):
        rows = f64(rows)
        assert rows.size == self.ntracer * self.nbin
        _chk(_out_lib().croma_hip_output_write(C.c_void_p(self.h), int(slot), _p(rows),
                                               _p(f64(rho)) if rho is not None else None,
                                               _p(f64(mass)) if mass is not None else None), "output_write")

    def write_batch(self, batch, slot, crp_out=None, rho=None, mass=None):
        """The batch's device spectra (electrons here, protons into crp_out)."""
        _chk(_out_lib().croma_hip_batch_write_spectra(C.c_void_p(batch.h), C.c_void_p(self.h),
                                                      C.c_void_p(crp_out.h) if crp_out is not None else None,
                                                      int(slot), _p(f64(rho)) if rho is not None else None,
                                                      _p(f64(mass)) if mass is not None else None),
             "batch_write_spectra")

    def close(self):
        if self.h:
            h, self.h = self.h, None
            _chk(_out_lib().croma_hip_output_close(C.c_void_p(h)), "output_close")


def write_tracer_ids(out_dir, rank, ids):
    ids = np.ascontiguousarray(ids, np.int32)
    _chk(_out_lib().croma_hip_write_tracer_ids(str(out_dir).encode(), int(rank), ids.size, _p(ids)),
         "write_tracer_ids")


def checkpoint_write(ckpt_dir, meta, ids, crp, cre):
    ids = np.ascontiguousarray(ids, np.int32)
    _chk(_out_lib().croma_hip_checkpoint_write(str(ckpt_dir).encode(), C.byref(meta), _p(ids), _p(f64(crp)),
                                               _p(f64(cre))), "checkpoint_write")


def checkpoint_read(ckpt_dir, meta, ids):
    ids = np.ascontiguousarray(ids, np.int32)
    crp = np.zeros(meta.ntracer * meta.np_bins)
    cre = np.zeros(meta.ntracer * meta.npe_bins)
    done = C.c_int()
    _chk(_out_lib().croma_hip_checkpoint_read(str(ckpt_dir).encode(), C.byref(meta), _p(ids), _p(crp), _p(cre),
                                              C.byref(done)), "checkpoint_read")
    return crp, cre, done.value


def chunk_bytes(ntracer, n, species, nsnap):
    L = lib()
    L.croma_hip_chunk_bytes.restype = C.c_size_t
    L.croma_hip_chunk_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    return int(L.croma_hip_chunk_bytes(ntracer, n, species, nsnap))
