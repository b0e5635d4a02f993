"""The reference-struct surface (include/croma_hip_b1.h) has the byte layout of
the reference's own structs: every sizeof and offsetof that
croma_hip_b1_layout reports equals the value compiled from the reference
headers (tests/golden/b1_layout.json, oracle/ref_layout.c) at every grid
size, and the CRspectrum view the library reads (p at 0, dp at 8 np, pp1 at
16 np, pm1 at 16 np + 8, E at 16 np + 16) is the reference's."""
import ctypes as C
import json
import os

import pytest

import croma_hip as ch

GOLD = os.path.join(os.path.dirname(__file__), "golden", "b1_layout.json")


def _lib():
    L = ch.lib()
    L.croma_hip_b1_layout.restype = C.c_size_t
    L.croma_hip_b1_layout.argtypes = [C.c_char_p, C.c_char_p]
    return L


@pytest.mark.parametrize("n", ["64", "128", "256", "512"])
def test_b1_struct_layouts_match_reference(n):
    ref = json.load(open(GOLD))[n]
    L = _lib()
    checked = 0
    for typ, fields in ref.items():
        if typ == "CRspectrum":
            continue
        for field, off in fields.items():
            got = L.croma_hip_b1_layout(typ.encode(), field.encode())
            assert got == off, f"{typ}.{field or 'sizeof'}: library {got}, reference {off}"
            checked += 1
    assert checked > 100


@pytest.mark.parametrize("n", ["64", "128", "256", "512"])
def test_crspectrum_view(n):
    ref = json.load(open(GOLD))[n]["CRspectrum"]
    k = int(n)
    assert ref == {"": 8 * (6 * k + 2), "p": 0, "dp": 8 * k, "pp1": 16 * k, "pm1": 16 * k + 8, "E": 16 * k + 16}


def test_unknown_names_rejected():
    L = _lib()
    assert L.croma_hip_b1_layout(b"FpCudaPipelineInput", b"no_such_field") == C.c_size_t(-1).value
    assert L.croma_hip_b1_layout(b"NoSuchType", b"") == C.c_size_t(-1).value


@pytest.mark.timeout(60, method="thread")
def test_emission_entry_points_return_without_a_device():
    """The emission entry points take their context lock and then the
    parameter lock; on a host without a device they must return -1 (no
    context), not block on their own lock."""
    import numpy as np

    import croma_b1 as b1

    if ch.is_available():
        pytest.skip("a device is visible: the device tests cover these entry points")
    L = b1.lib()
    n = 64
    spec = b1.crspectrum(np.logspace(-1, 8, n), np.ones(n), 0.05, 1.1e8, np.ones(n))
    ng = np.ones(4)
    crp = np.ones(4 * n)
    f = np.ones(8 * n)
    eps = np.zeros(4 * 8)
    inp = b1.HadronicEmissionBatchInput(4, 8, 0.0, -1.0, 5.0, b1.ptr(ng), b1.ptr(crp), b1.ptr(spec[:n]), b1.ptr(f))
    assert L.croma_hip_b1_prepare_gamma_emission(n, n, C.byref(inp), spec.ctypes.data, b1.ptr(eps), None) == -1
    assert L.croma_hip_b1_prepare_neutrino_emission(n, n, C.byref(inp), spec.ctypes.data, b1.ptr(eps), None) == -1
