"""Request loop (croma_hip_serve_*): results are the same whether a chunk runs
alone or in a mix of chunks of other sizes, grids, species and D_pp models
served concurrently; every result matches the CPU oracle (|hip - cpu| <=
1e-30 + 1e-10 |cpu|); admission respects the HBM budget; cancellation."""
import numpy as np
import pytest

import croma_hip as ch
import run_oracle
import serve_stream
from workloads import rel_err

pytestmark = pytest.mark.gpu

MIX = [
    dict(ntracer=96, n=128, species=ch.SPECIES_ELECTRON, dpp=ch.DPP_ASA, nsnap=3, seed=11, z0=0.4),
    dict(ntracer=40, n=256, species=ch.SPECIES_ELECTRON_PROTON, dpp=ch.DPP_TTD, nsnap=2, seed=12, z0=0.8),
    dict(ntracer=64, n=64, species=ch.SPECIES_PROTON, dpp=ch.DPP_OFF, nsnap=3, seed=13, z0=0.3),
    dict(ntracer=24, n=512, species=ch.SPECIES_ELECTRON, dpp=ch.DPP_TTD, nsnap=2, seed=14, z0=0.6),
    dict(ntracer=200, n=128, species=ch.SPECIES_ELECTRON, dpp=ch.DPP_OFF, nsnap=2, seed=15, z0=0.2),
]


def oracle_result(spec):
    w, _ = serve_stream.make_chunk(spec)
    return run_oracle.run_workload(w, run_oracle.oracle_params(w), s0=0, s1=spec["nsnap"])


@pytest.fixture(scope="module")
def server():
    s = ch.Server(2 * 2 ** 30, serve_stream.server_params())
    yield s
    s.close()


def test_alone_equals_mix_and_oracle(server):
    alone = []
    for spec in MIX:
        print(f"[serve] alone {spec}", flush=True)
        _, kw = serve_stream.make_chunk(spec)
        rid = server.submit(**kw)
        crp, cre, st = server.wait(rid)
        alone.append((crp, cre))
        # the served chunk's nsub schedule is built on the device: its
        # tracer-substeps equal the batch path's for the same chunk
        seq = serve_stream.run_sequential([spec])
        assert st["cell_steps"] == seq["tracer_substeps"] > 0, (st["cell_steps"], seq["tracer_substeps"])
    print("[serve] mix", flush=True)
    ids = [server.submit(**serve_stream.make_chunk(spec)[1]) for spec in MIX]
    mixed = [server.wait(i)[:2] for i in ids]
    for spec, (pa, ea), (pm, em) in zip(MIX, alone, mixed):
        assert np.array_equal(ea, em) and np.array_equal(pa, pm), f"alone != mix for {spec}"
        crp_ref, cre_ref = oracle_result(spec)
        e_err, e_ok = rel_err(em, cre_ref)
        print(f"  n={spec['n']} species={spec['species']} electron max rel err {e_err:.3e}", flush=True)
        assert e_ok
        if spec["species"] != ch.SPECIES_ELECTRON:
            p_err, p_ok = rel_err(pm, crp_ref)
            print(f"    proton max rel err {p_err:.3e}", flush=True)
            assert p_ok
        else:
            assert np.all(pm == 0.0)
    st = server.stats()
    assert st["completed"] == 2 * len(MIX) and st["failed"] == 0


def test_budget_admission_and_rejection():
    spec = dict(ntracer=512, n=128, species=ch.SPECIES_ELECTRON, dpp=ch.DPP_ASA, nsnap=2, seed=21, z0=0.5)
    need = ch.chunk_bytes(512, 128, ch.SPECIES_ELECTRON, 2)
    # room for one such chunk at a time: later chunks wait for memory
    s = ch.Server(int(need * 1.5), serve_stream.server_params())
    try:
        _, kw = serve_stream.make_chunk(spec)
        ids = [s.submit(**kw) for _ in range(3)]
        outs = [s.wait(i) for i in ids]
        for o in outs[1:]:
            assert np.array_equal(o[1], outs[0][1])
        st = s.stats()
        assert st["peak_bytes"] <= st["budget_bytes"]
        assert st["peak_bytes"] < 2 * need
        big = dict(spec, ntracer=4096)
        with pytest.raises(ch.CromaHipError) as ei:
            s.submit(**serve_stream.make_chunk(big)[1])
        assert ei.value.rc == ch.EBUDGET
    finally:
        s.close()


def test_cancel_queued():
    spec = dict(ntracer=2048, n=64, species=ch.SPECIES_ELECTRON, dpp=ch.DPP_OFF, nsnap=2, seed=31, z0=0.5)
    need = ch.chunk_bytes(2048, 64, ch.SPECIES_ELECTRON, 2)
    # one chunk at a time: the last submission is still queued when cancelled
    s = ch.Server(int(need * 1.5), serve_stream.server_params())
    try:
        _, kw = serve_stream.make_chunk(spec)
        ids = [s.submit(**kw) for _ in range(4)]
        s.cancel(ids[-1])
        with pytest.raises(ch.CromaHipError) as ei:
            s.wait(ids[-1])
        assert ei.value.rc == ch.ECANCELED
        for i in ids[:-1]:
            s.wait(i)
        assert s.stats()["cancelled"] == 1
    finally:
        s.close()


def test_many_small_odd_chunks():
    """120 concurrent chunks with odd tracer counts (33..301) over all grid
    sizes: the last wave of most launches holds lane groups past the end of
    its task list, which must not disturb the live groups beside them (the
    solve reads across groups and multiplies by zero).  Every chunk
    finishes with finite spectra; chunks run alone equal their served
    results, and two of them equal the oracle."""
    rng = np.random.default_rng(77)
    specs = [
        dict(ntracer=33, n=64, species=ch.SPECIES_ELECTRON, dpp=ch.DPP_TTD, nsnap=2, seed=901, z0=0.5),
        dict(ntracer=65, n=128, species=ch.SPECIES_ELECTRON_PROTON, dpp=ch.DPP_ASA, nsnap=2, seed=902, z0=0.7),
    ]
    for i in range(118):
        specs.append(dict(ntracer=int(2 * rng.integers(16, 151) + 1), n=int(rng.choice([64, 128, 256, 512])),
                          species=int(rng.choice([ch.SPECIES_ELECTRON, ch.SPECIES_ELECTRON_PROTON])),
                          dpp=int(rng.choice([ch.DPP_TTD, ch.DPP_ASA, ch.DPP_OFF])), nsnap=2, seed=1000 + i,
                          z0=float(rng.uniform(0.2, 0.8))))
    s = ch.Server(2 * 2 ** 30, serve_stream.server_params())
    try:
        print(f"[serve] {len(specs)} odd-sized chunks", flush=True)
        ids = [s.submit(**serve_stream.make_chunk(sp)[1]) for sp in specs]
        served = [s.wait(i)[:2] for i in ids]
        assert s.stats()["failed"] == 0
        for crp, cre in served:
            assert np.all(np.isfinite(cre)) and np.all(np.isfinite(crp))
        for k in (0, 1, 2, 3):
            crp, cre, _ = s.wait(s.submit(**serve_stream.make_chunk(specs[k])[1]))
            assert np.array_equal(cre, served[k][1]) and np.array_equal(crp, served[k][0]), specs[k]
    finally:
        s.close()
    for k in (0, 1):
        crp_ref, cre_ref = oracle_result(specs[k])
        e_err, e_ok = rel_err(served[k][1], cre_ref)
        print(f"  chunk {k}: electron max rel err {e_err:.3e}", flush=True)
        assert e_ok
        if specs[k]["species"] != ch.SPECIES_ELECTRON:
            p_err, p_ok = rel_err(served[k][0], crp_ref)
            assert p_ok
