"""Allocation-free batched Chang-Cooper solve (croma_hip_solve_cc_batch_noalloc)
against the allocating entry point (croma_hip_solve_cc_batch) on the same
systems: bit-identical states on two calls in a row, the second smaller than
the first, plus the oracle's Thomas sweep (fp_shared_core.c:226-257) on a
sample of systems within 1e-12, the size limit and the pivot-failure word
being cleared between calls.  The caller's device arrays come from the HIP
runtime the library itself links (hipMalloc / hipMemcpy through ctypes)."""
import ctypes as C

import numpy as np
import pytest

import croma_hip as ch
import pyoracle as po

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceArrays:
    """Caller-owned device copies of host float64 arrays."""

    def __init__(self, *arrays):
        ch.lib()  # loads the HIP runtime the library links
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []
        for a in arrays:
            a = np.ascontiguousarray(a, dtype=np.float64)
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), a.nbytes) == 0
            self.ptrs.append(p.value)
            assert self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, H2D) == 0

    def alloc_like(self, a):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), a.nbytes) == 0
        self.ptrs.append(p.value)
        return p.value

    def read(self, ptr, like):
        out = np.empty_like(like, dtype=np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _systems(nsys, nrow, seed):
    """Chang-Cooper-shaped systems: non-negative couplings a, c and a
    dominant diagonal b = 1 + dt/T + a + c (rows b x_j - a x_{j-1} - c x_{j+1})."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, 2.0, (nsys, nrow)) * 10.0 ** rng.uniform(-3, 2, (nsys, 1))
    c = rng.uniform(0.0, 2.0, (nsys, nrow)) * 10.0 ** rng.uniform(-3, 2, (nsys, 1))
    a[:, 0] = 0.0
    c[:, -1] = 0.0
    b = 1.0 + rng.uniform(0.0, 0.1, (nsys, nrow)) + a + c
    q = 10.0 ** rng.uniform(-6, 0, (nsys, nrow))
    x = 10.0 ** rng.uniform(-20, 0, (nsys, nrow))
    return a, b, c, q, x


def _noalloc(a, b, c, q, x, dt, in_place=False):
    dev = DeviceArrays(a, b, c, q, x)
    try:
        out = dev.ptrs[4] if in_place else dev.alloc_like(x)
        ch.solve_cc_batch_noalloc(a.shape[0], a.shape[1], dt, *dev.ptrs[:5], out, 0)
        return dev.read(out, x)
    finally:
        dev.close()


def test_noalloc_matches_allocating_entry_point_twice():
    dt = 0.0137
    for call, (nsys, nrow, seed) in enumerate(((4096, 512, 1), (777, 128, 2))):
        print(f"noalloc call {call}: nsys={nsys} nrow={nrow}", flush=True)
        a, b, c, q, x = _systems(nsys, nrow, seed)
        ref, _ = ch.solve_cc_batch(a, b, c, q, x, dt)
        got = _noalloc(a, b, c, q, x, dt, in_place=(call == 1))
        assert np.array_equal(got, ref), f"call {call}: max diff {np.max(np.abs(got - ref))}"
        for s in range(0, nsys, max(1, nsys // 16)):
            rc, xo = po.thomas(a[s], b[s], c[s], dt * q[s] + x[s])
            assert rc == 0
            assert np.all(np.abs(got[s] - xo) <= 1e-30 + 1e-12 * np.abs(xo)), f"system {s}"


def test_noalloc_limits_and_pivot_word():
    a, b, c, q, x = _systems(64, 64, 3)
    dev = DeviceArrays(a, b, c, q, x)
    try:
        # more cells than the static scratch holds: rejected before any launch
        with pytest.raises(ch.CromaHipError):
            ch.solve_cc_batch_noalloc(ch.NOALLOC_MAX_CELLS // 64 + 1, 64, 0.1, *dev.ptrs[:5], dev.ptrs[4])
    finally:
        dev.close()
    # a zero first pivot is reported (the kernel's own guard) ...
    b_bad = b.copy()
    b_bad[5, 0] = 0.0
    with pytest.raises(ch.CromaHipError):
        _noalloc(a, b_bad, c, q, x, 0.1)
    # ... and the failure word starts at zero again on the next call
    ref, _ = ch.solve_cc_batch(a, b, c, q, x, 0.1)
    assert np.array_equal(_noalloc(a, b, c, q, x, 0.1), ref)
