"""Files written from the batch's device-resident state, and resume.

1. Spectra written per snapshot straight from the device (both layouts, with
   and without the per-cm^3 factor) equal the state the batch reports, slot
   for slot (slot 0 = initial state, as the reference writes it).
2. A checkpoint written from the device after k snapshots and read into a
   fresh batch resumes to the same spectra as the uninterrupted run, bit for
   bit (tracer_fp_ckpt_write / _read, tracer_fp_restart.c:136-265)."""
import numpy as np
import pytest

import croma_hip as ch
from workloads import hip_params, small_workload

pytestmark = pytest.mark.gpu


def _batch(ctx, w, state=True):
    b = ch.Batch(ctx, w.ntracer, w.species, w.nsnap)
    b.set_state(w.crp0 if w.protons else None, w.cre0, mass=w.mass)
    for s in range(w.nsnap):
        b.set_background(s, w.bg[s])
        if s in w.qpi or s in w.qepri:
            b.set_sources(s, w.qpi.get(s), w.qepri.get(s))
    return b


@pytest.mark.parametrize("mode", [ch.OUT_MAP, ch.OUT_TILE])
@pytest.mark.parametrize("per_cc", [False, True])
def test_spectra_written_from_device(tmp_path, mode, per_cc):
    n, nt, ns = 128, 3000, 3
    w = small_workload(n, nt, ns, 3, 1, dsa_frac=0.2, seed=51)
    rng = np.random.default_rng(9)
    rho = rng.uniform(1e-28, 1e-25, nt) if per_cc else None
    rho_kw = dict(rho=rho, mass=w.mass) if per_cc else {}
    ctx = ch.Context(n, hip_params(w))
    b = _batch(ctx, w)
    oe = ch.Output(tmp_path, "CRE", 0, mode, nt, n, ns + 1)
    op = ch.Output(tmp_path, "CRP", 0, mode, nt, n, ns + 1)
    states = []
    try:
        for s in range(ns + 1):
            if s > 0:
                print(f"[output] mode={mode} per_cc={per_cc} snapshot {s}", flush=True)
                b.run(s - 1, s, ns, w.z, w.dt_snap, elapsed=float(np.sum(w.dt_snap[:s - 1])))
            oe.write_batch(b, s, crp_out=op, **rho_kw)
            states.append(b.get_state())
    finally:
        oe.close()
        op.close()
        b.close()
        ctx.close()
    fac = (rho / (w.mass * 1.99e33))[:, None] if per_cc else 1.0
    for stem, k in (("CRE", 1), ("CRP", 0)):
        name = f"{stem}_tile_core00.bin" if mode == ch.OUT_TILE else f"{stem}_core00.bin"
        raw = np.fromfile(tmp_path / name)
        got = raw.reshape(ns + 1, nt, n) if mode == ch.OUT_TILE else raw.reshape(nt, ns + 1, n).transpose(1, 0, 2)
        for s in range(ns + 1):
            assert np.array_equal(got[s], states[s][k].reshape(nt, n) * fac), f"{stem} slot {s}"


@pytest.mark.parametrize("species,gemm", [(1, 0), (3, 1)])
def test_checkpoint_resume_is_bit_identical(tmp_path, species, gemm):
    n, nt, ns, k = 128, 200, 5, 2
    w = small_workload(n, nt, ns, species, 1, dsa_frac=0.2, seed=53)
    ctx = ch.Context(n, hip_params(w))
    ctx.set_tile(8, 1, 1, secondary_gemm=gemm)
    ids = np.arange(1000, 1000 + nt)
    meta = ch.CheckpointMeta(mpi_rank=0, world_size=1, ntracer=nt, ntracer_global=nt, nsnap=ns, nfreq=64,
                             np_bins=n, npe_bins=n, nsnp_i=0, nsnp_f=ns, completed_snapshot=k - 1, tracer_start=0)
    try:
        print(f"[ckpt] species={species} uninterrupted", flush=True)
        b = _batch(ctx, w)
        b.run(0, ns, ns, w.z, w.dt_snap)
        want = b.get_state()
        b.close()
        print(f"[ckpt] species={species} first {k} snapshots + checkpoint", flush=True)
        b = _batch(ctx, w)
        b.run(0, k, ns, w.z, w.dt_snap)
        b.checkpoint_write(tmp_path, meta, ids)
        b.close()
        print(f"[ckpt] species={species} resume", flush=True)
        b = _batch(ctx, w)
        b.set_state(np.zeros(nt * n) if w.protons else None, np.zeros(nt * n), mass=w.mass)
        done = b.checkpoint_read(tmp_path, meta, ids)
        assert done == k - 1
        b.run(done + 1, ns, ns, w.z, w.dt_snap, elapsed=float(np.sum(w.dt_snap[:done + 1])))
        got = b.get_state()
        b.close()
    finally:
        ctx.close()
    assert np.array_equal(got[1], want[1]) and np.array_equal(got[0], want[0])
