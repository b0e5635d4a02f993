"""Output and checkpoint files in the reference's byte layouts (host-array
paths of the library; no GPU needed).

The layouts are fixed by the reference's writers and readers:
- mapped / buffered spectra: <stem>_coreNN.bin, float64 [tracer][slot][nbin]
  (tracer_fp_output.c:38-72, 613-664; analysis/compare_tracer_outputs.py:300-330
  reads ndouble / ntracer columns per row);
- tile: <stem>_tile_coreNN.bin, float64 [slot][tracer][nbin] (:353-416);
- tracer ids: tracerid_coreNN.txt, "%09d\\n" per tracer (:74-96);
- checkpoint TFPCHK1: 72-byte header (include/tracer_fp_types.h:356-371:
  char[16] magic, 12 int, long tracer_start), int ids, crp, cre
  (tracer_fp_restart.c:136-265)."""
import os
import struct

import numpy as np
import pytest

import croma_hip as ch

HDR = struct.Struct("<16s12iq")


def _rows(rng, nt, nbin):
    return rng.lognormal(0.0, 3.0, size=(nt, nbin))


@pytest.mark.parametrize("mode", [ch.OUT_MAP, ch.OUT_TILE])
def test_spectra_file_layout(tmp_path, mode):
    rng = np.random.default_rng(1)
    nt, nbin, nslot = 37, 16, 4
    slots = [_rows(rng, nt, nbin) for _ in range(nslot)]
    o = ch.Output(tmp_path, "CRE", 3, mode, nt, nbin, nslot)
    for s in (2, 0, 3, 1):  # any order
        o.write(s, slots[s])
    o.close()
    name = "CRE_tile_core03.bin" if mode == ch.OUT_TILE else "CRE_core03.bin"
    raw = np.fromfile(tmp_path / name, dtype="<f8")
    assert raw.size == nt * nbin * nslot
    if mode == ch.OUT_TILE:
        got = raw.reshape(nslot, nt, nbin)
        for s in range(nslot):
            assert np.array_equal(got[s], slots[s])
    else:
        # the comparator's view: one row of ncol = ndouble / ntracer per tracer
        got = raw.reshape(nt, nslot * nbin)
        for s in range(nslot):
            assert np.array_equal(got[:, s * nbin:(s + 1) * nbin], slots[s])


def test_per_cc_factor(tmp_path):
    """rho / (mass M_sun) per tracer; 0 where rho or mass is not positive."""
    rng = np.random.default_rng(2)
    nt, nbin = 6, 5
    rows = _rows(rng, nt, nbin)
    rho = np.array([1e-27, 0.0, 3e-26, -1.0, 2e-28, 5e-27])
    mass = np.array([1e7, 1e7, 0.0, 1e7, 2e6, 4e7])
    o = ch.Output(tmp_path, "eSyn", 0, ch.OUT_MAP, nt, nbin, 1)
    o.write(0, rows, rho=rho, mass=mass)
    o.close()
    got = np.fromfile(tmp_path / "eSyn_core00.bin").reshape(nt, nbin)
    ok = (rho > 0) & (mass > 0)
    fac = np.zeros(nt)
    fac[ok] = rho[ok] / (mass[ok] * 1.99e33)
    assert np.array_equal(got, rows * fac[:, None])


def test_resume_checks_size_and_slot_range(tmp_path):
    nt, nbin, nslot = 4, 8, 3
    o = ch.Output(tmp_path, "CRP", 1, ch.OUT_MAP, nt, nbin, nslot)
    with pytest.raises(ch.CromaHipError):
        o.write(nslot, np.zeros((nt, nbin)))
    o.close()
    with pytest.raises(ch.CromaHipError):
        ch.Output(tmp_path, "CRP", 1, ch.OUT_MAP, nt, nbin, nslot + 1, resume=True)
    o = ch.Output(tmp_path, "CRP", 1, ch.OUT_MAP, nt, nbin, nslot, resume=True)
    o.write(1, np.ones((nt, nbin)))
    o.close()
    got = np.fromfile(tmp_path / "CRP_core01.bin").reshape(nt, nslot, nbin)
    assert np.all(got[:, 1] == 1.0) and np.all(got[:, 0] == 0.0) and np.all(got[:, 2] == 0.0)


def test_tracer_id_file(tmp_path):
    ids = np.array([5, 123456789, 0, 42])
    ch.write_tracer_ids(tmp_path, 7, ids)
    text = (tmp_path / "tracerid_core07.txt").read_text()
    assert text == "000000005\n123456789\n000000000\n000000042\n"


def _meta(nt, n, **kw):
    m = ch.CheckpointMeta(mpi_rank=2, world_size=4, ntracer=nt, ntracer_global=4 * nt, nsnap=100, nfreq=64,
                          np_bins=n, npe_bins=n, nsnp_i=10, nsnp_f=110, completed_snapshot=17,
                          tracer_start=123456789012)
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def test_checkpoint_bytes_and_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    nt, n = 9, 12
    ids = np.arange(100, 100 + nt)
    crp, cre = rng.normal(size=nt * n), rng.normal(size=nt * n)
    ch.checkpoint_write(tmp_path, _meta(nt, n), ids, crp, cre)
    path = tmp_path / "checkpoint_rank02.bin"
    assert sorted(os.listdir(tmp_path)) == ["checkpoint_rank02.bin"]  # the .tmp was renamed
    data = path.read_bytes()
    assert len(data) == 72 + 4 * nt + 8 * nt * 2 * n
    h = HDR.unpack(data[:72])
    assert h[0] == b"TFPCHK1" + b"\0" * 9
    assert h[1:] == (1, 2, 4, nt, 4 * nt, 100, 64, n, n, 10, 110, 17, 123456789012)
    assert np.array_equal(np.frombuffer(data, "<i4", nt, 72), ids)
    assert np.array_equal(np.frombuffer(data, "<f8", nt * n, 72 + 4 * nt), crp)
    assert np.array_equal(np.frombuffer(data, "<f8", nt * n, 72 + 4 * nt + 8 * nt * n), cre)
    crp2, cre2, done = ch.checkpoint_read(tmp_path, _meta(nt, n, completed_snapshot=0), ids)
    assert done == 17 and np.array_equal(crp2, crp) and np.array_equal(cre2, cre)


@pytest.mark.parametrize("field,value", [("world_size", 8), ("nsnap", 99), ("np_bins", 13), ("nsnp_i", 11),
                                         ("tracer_start", 5)])
def test_checkpoint_rejects_other_run(tmp_path, field, value):
    nt, n = 3, 4
    ids = np.arange(nt)
    ch.checkpoint_write(tmp_path, _meta(nt, n), ids, np.zeros(nt * n), np.zeros(nt * n))
    bad = _meta(nt, n, **{field: value})
    if field == "np_bins":
        bad.npe_bins = value
    with pytest.raises(ch.CromaHipError):
        ch.checkpoint_read(tmp_path, bad, ids)


def test_checkpoint_rejects_tracer_order_and_bad_magic(tmp_path):
    nt, n = 3, 4
    ids = np.arange(nt)
    ch.checkpoint_write(tmp_path, _meta(nt, n), ids, np.zeros(nt * n), np.zeros(nt * n))
    with pytest.raises(ch.CromaHipError):
        ch.checkpoint_read(tmp_path, _meta(nt, n), ids[::-1].copy())
    path = tmp_path / "checkpoint_rank02.bin"
    data = bytearray(path.read_bytes())
    data[6] = ord("2")  # TFPCHK2
    path.write_bytes(bytes(data))
    with pytest.raises(ch.CromaHipError):
        ch.checkpoint_read(tmp_path, _meta(nt, n), ids)


def test_checkpoint_header_matches_reference_struct():
    """The header bytes follow the compiled reference struct (sizeof and every
    offsetof, recorded by tests/golden/make_golden.py from the reference's
    include/tracer_fp_types.h)."""
    import json
    lay = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ckpt_header_layout.json")))
    assert lay["sizeof"] == HDR.size == 72
    assert lay["sizeof_magic"] == 16 and lay["magic"] == 0
    ints = ["version", "mpi_rank", "world_size", "ntracer", "ntracer_global", "nsnap", "nfreq", "np_bins",
            "npe_bins", "nsnp_i_value", "nsnp_f_value", "completed_snapshot"]
    assert [lay[k] for k in ints] == [16 + 4 * i for i in range(12)]
    assert lay["tracer_start"] == 64
