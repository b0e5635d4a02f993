"""Multi-snapshot driver of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Follows the reference snapshot loop (include/tracer_fp_solve.inc:1-1780):
per snapshot s the IC cooling rows at z_s (and z_{s+1} when a next snapshot
exists), the nsub estimate from the current backgrounds, then the
snapshot-level update of every tracer.
"""
from __future__ import annotations

import numpy as np

import pyoracle as po


def oracle_params(w, **kw):
    base = dict(dpp_model=w.dpp_model, tracer_nsub_safety=0.5)
    base.update(kw)
    return po.default_params(**base)


def run_workload(w, prm=None, s0=0, s1=None, nthreads=0, oracle=None, crp=None, cre=None, return_nsub=False,
                 secondary=None, ic_rows=None, dis=None):
    """Advance the workload's spectra over snapshots [s0, s1); returns (crp, cre[, nsub list]).

    secondary: precomputed (fqe_t, np_min_qe) of the grid; ic_rows: dict
    z -> (row, m1, p1) of precomputed IC cooling rows (both are per-grid or
    per-snapshot constants shared by every tracer); dis: per-tracer
    disable_adiabatic flags [ntracer] (uint8) or None."""
    o = oracle or po.Oracle(w.n, w.pmin, w.pmax, w.pemin, w.pemax)
    prm = prm or oracle_params(w)
    s1 = w.nsnap if s1 is None else s1
    crp = (w.crp0 if crp is None else crp).copy()
    cre = (w.cre0 if cre is None else cre).copy()
    if secondary is not None:
        fqe, npmin = secondary
    else:
        fqe, npmin = (o.secondary_kernel() if w.protons else (np.zeros(w.n * w.n), np.full(w.n, w.n, np.int32)))
    rows = ic_rows if ic_rows is not None else {}

    def ic(z):
        if z not in rows:
            rows[z] = o.ic_cooling_row(z)
        return rows[z]
    zeros = np.zeros(w.ntracer * w.n)
    elapsed = float(np.sum(w.dt_snap[:s0]))
    nsubs = []
    for s in range(s0, s1):
        allow = s + 1 < w.nsnap
        row, m1, p1 = ic(w.z[s])
        if allow:
            rown, m1n, p1n = ic(w.z[s + 1])
        else:
            rown, m1n, p1n = row, m1, p1
        bgc = w.bg[s]
        bgn = w.bg[s + 1] if allow else w.bg[s]
        ns, _, non, _ = o.nsub(prm, w.dt_snap[s], bgc[0], bgc[1], bgc[2], bgc[3], row[0], row[-1], bgc[4], bgc[5])
        nsubs.append(ns)
        qpi = w.qpi.get(s, zeros)
        qepri = w.qepri.get(s, zeros)
        o.execute_snapshot(prm, ns, non, allow, elapsed, w.dt_snap[s], bgc, row, m1, p1, bgn, rown, m1n, p1n, qpi,
                           qepri, fqe, npmin, w.mass, dis, crp, cre, nthreads=nthreads)
        elapsed += w.dt_snap[s]
    if return_nsub:
        return crp, cre, nsubs
    return crp, cre
