"""Shared helpers for the parity tests: small seeded workloads and the two
drivers (CPU oracle, HIP library) over the same inputs."""
from __future__ import annotations

import numpy as np

import synth


def small_workload(n, ntracer, nsnap, species, dpp_model, seed=7, dsa_frac=0.0):
    # z0 keeps snapshot durations close to the 100-snapshot configurations
    return synth.make_workload(n=n, ntracer=ntracer, nsnap=nsnap, species=species, dpp_model=dpp_model,
                               seed=seed, z0=0.01 * nsnap, dsa_frac=dsa_frac)


def hip_params(w, **kw):
    import croma_hip as ch
    base = dict(dpp_model=w.dpp_model, tracer_nsub_safety=0.5)
    base.update(kw)
    return ch.default_params(**base)


def run_hip(w, ctx, path=0, s0=0, s1=None, crp=None, cre=None, dis=None):
    import croma_hip as ch
    s1 = w.nsnap if s1 is None else s1
    b = ch.Batch(ctx, w.ntracer, w.species, w.nsnap)
    try:
        b.set_state(crp if crp is not None else (w.crp0 if w.protons else None),
                    cre if cre is not None else w.cre0, mass=w.mass, dis=dis)
        for s in range(w.nsnap):
            b.set_background(s, w.bg[s])
            if s in w.qpi or s in w.qepri:
                b.set_sources(s, w.qpi.get(s), w.qepri.get(s))
        t = b.run(s0, s1, w.nsnap, w.z, w.dt_snap, elapsed=float(np.sum(w.dt_snap[:s0])), path=path)
        crp_o, cre_o = b.get_state()
    finally:
        b.close()
    return crp_o, cre_o, t


def rel_err(got, want, floor=1e-30):
    """max |got - want| / |want| over entries with |want| > floor; also the
    reference comparator's pass flag |got - want| <= 1e-30 + 1e-10 |want|."""
    got = np.asarray(got)
    want = np.asarray(want)
    mask = np.abs(want) > floor
    err = np.max(np.abs(got[mask] - want[mask]) / np.abs(want[mask])) if mask.any() else 0.0
    ok = bool(np.all(np.abs(got - want) <= 1e-30 + 1e-10 * np.abs(want))) and bool(np.isfinite(got).all())
    return float(err), ok
