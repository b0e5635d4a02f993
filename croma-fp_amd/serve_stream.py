"""Served-stream measurement: a Poisson arrival stream of tracer chunks
through the request loop (croma_hip_serve_*), reporting chunk latency
percentiles and throughput.

Chunk mix: sizes log-uniform in [256, 65536] tracers, grids 64/128/256/512
bins, seed species electron_only / proton_only / electron_proton, D_pp
asa / ttd / off, each chunk advancing 10 snapshots; backgrounds drawn as in
the synthetic workloads (synth.make_workload).  Latency = submit -> final
spectra on the host.  Throughput = sum of tracer-substeps / makespan.

The same chunks, run one at a time through the batch API (no overlap
between chunks), give the "sequential" comparison; a sample run through the
plain one-tracer-per-thread path gives the plain rate.
"""
from __future__ import annotations

import math
import sys
import threading
import time

import numpy as np

import croma_hip as ch
import synth

NS = (64, 128, 256, 512)
SPECIES = (ch.SPECIES_ELECTRON, ch.SPECIES_PROTON, ch.SPECIES_ELECTRON_PROTON)
DPP = (ch.DPP_ASA, ch.DPP_TTD, ch.DPP_OFF)


def chunk_specs(count, seed, min_tr=256, max_tr=65536, nsnap=10):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        out.append(dict(ntracer=int(round(10.0 ** rng.uniform(math.log10(min_tr), math.log10(max_tr)))),
                        n=int(rng.choice(NS)), species=int(rng.choice(SPECIES)), dpp=int(rng.choice(DPP)),
                        nsnap=nsnap, seed=int(seed * 100003 + i), z0=float(rng.uniform(0.1, 1.0))))
    return out


def make_chunk(spec):
    """Synthetic chunk: returns (workload with nsnap + 1 background sets, submit kwargs)."""
    k = spec["nsnap"]
    w = synth.make_workload(n=spec["n"], ntracer=spec["ntracer"], nsnap=k + 1, species=spec["species"],
                            dpp_model=spec["dpp"], seed=spec["seed"], z0=spec["z0"])
    kw = dict(n=w.n, species=spec["species"], dpp_model=spec["dpp"], z=w.z[:k + 1], dt_snap=w.dt_snap[:k],
              bg=np.ascontiguousarray(w.bg[:k + 1]), cre=w.cre0, crp=w.crp0 if w.protons else None,
              mass=w.mass)
    return w, kw


def server_params():
    return ch.default_params(tracer_nsub_safety=0.5)


def warm(server, seed=99):
    """One small chunk per launch class, so grid tables and secondary
    weights exist before timing."""
    ids = []
    for n in NS:
        for sp in (ch.SPECIES_ELECTRON, ch.SPECIES_ELECTRON_PROTON):
            _, kw = make_chunk(dict(ntracer=64, n=n, species=sp, dpp=ch.DPP_ASA, nsnap=2, seed=seed, z0=0.5))
            ids.append(server.submit(**kw))
    for i in ids:
        server.wait(i)


def run_stream(chm=ch, count=1000, seconds=240.0, seed=3, rate=None, budget_gb=64.0, max_tracers=65536,
               load=0.7):
    """Poisson stream of `count` chunks through the server (stopping early
    once `seconds` have passed); returns a summary dict.  Chunks are
    generated ahead of their arrival by a producer thread (at most 32
    pending), so generation cost does not delay arrivals; the arrival rate
    is set so the offered load is `load` times what the server sustains on
    a closed-loop probe of the stream's first 16 chunks."""
    import queue

    specs = chunk_specs(count, seed, max_tr=max_tracers)
    srv = ch.Server(int(budget_gb * 2 ** 30), server_params())
    try:
        warm(srv)
        if rate is None:
            probe = [make_chunk(s)[1] for s in specs[:16]]
            t0 = time.perf_counter()
            ids = [srv.submit(**kw) for kw in probe]
            for i in ids:
                srv.wait(i)
            per_chunk_s = (time.perf_counter() - t0) / len(ids)
            rate = load / per_chunk_s
            del probe
        rng = np.random.default_rng(seed + 1)
        ready = queue.Queue(maxsize=32)
        stop = threading.Event()
        gen_times = []

        def produce():
            for sp in specs:
                if stop.is_set():
                    break
                g0 = time.perf_counter()
                kw = make_chunk(sp)[1]
                gen_times.append(time.perf_counter() - g0)
                ready.put(kw)
            ready.put(None)

        producer = threading.Thread(target=produce, daemon=True)
        producer.start()
        done, waiters = [], []

        def collect(rid):
            _, _, st = srv.wait(rid)
            done.append(st)

        t_start = time.perf_counter()
        t_next = t_start
        late = 0
        while time.perf_counter() - t_start < seconds:
            kw = ready.get()
            if kw is None:
                break
            now = time.perf_counter()
            if now < t_next:
                time.sleep(t_next - now)
            elif now - t_next > 0.05:
                late += 1
            rid = srv.submit(**kw)
            del kw
            th = threading.Thread(target=collect, args=(rid,), daemon=True)
            th.start()
            waiters.append(th)
            t_next += rng.exponential(1.0 / rate)
            if len(waiters) % 100 == 0:
                print(f"[serve] {len(waiters)} chunks submitted, {len(done)} done, "
                      f"{time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)
        stop.set()
        while producer.is_alive():
            try:
                ready.get(timeout=0.1)
            except queue.Empty:
                pass
        for th in waiters:
            th.join()
        makespan = max(d["done_s"] for d in done) - min(d["submit_s"] for d in done)
        lat = np.array([d["latency_s"] for d in done]) * 1e3
        steps_total = sum(d["cell_steps"] for d in done)
        st = srv.stats()
    finally:
        srv.close()
    return {"chunks": len(done), "arrival_rate_per_s": rate, "offered_load": load,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "throughput_substeps_per_s": steps_total / makespan, "makespan_s": makespan,
            "tracer_substeps": int(steps_total), "launches": int(st["iterations"]),
            "peak_hbm_bytes": int(st["peak_bytes"]), "budget_bytes": int(st["budget_bytes"]),
            "arrivals_late_by_generation": late,
            "chunk_gen_ms_mean": 1e3 * float(np.mean(gen_times)) if gen_times else 0.0}


def run_closed_loop(specs, budget_gb=64.0):
    """Server capacity: every chunk submitted at once (admitted as the HBM
    budget allows); throughput = sum of tracer-substeps / makespan."""
    payloads = [make_chunk(s)[1] for s in specs]
    srv = ch.Server(int(budget_gb * 2 ** 30), server_params())
    try:
        warm(srv)
        t0 = time.perf_counter()
        ids = [srv.submit(**kw) for kw in payloads]
        steps = sum(srv.wait(i)[2]["cell_steps"] for i in ids)
        wall = time.perf_counter() - t0
    finally:
        srv.close()
    return {"chunks": len(specs), "throughput_substeps_per_s": steps / wall, "tracer_substeps": int(steps),
            "wall_s": wall}


def run_sequential(specs, path=ch.PATH_FUSED, params=None, verbose=False, budget_s=None):
    """The same chunks one at a time through the batch API (no overlap).
    budget_s: stop after the chunk that takes the total past it (the result
    counts the chunks run, a prefix of specs)."""
    params = params or server_params()
    ctxs = {}
    steps, wall, done = 0, 0.0, 0
    try:
        for i, spec in enumerate(specs):
            if budget_s is not None and done and wall >= budget_s:
                break
            w, kw = make_chunk(spec)
            if w.n not in ctxs:
                ctxs[w.n] = ch.Context(w.n, params)
            ctx = ctxs[w.n]
            prm = ch.default_params(tracer_nsub_safety=0.5, dpp_model=spec["dpp"])
            ctx.set_params(prm)
            k = spec["nsnap"]
            t0 = time.perf_counter()
            b = ch.Batch(ctx, w.ntracer, spec["species"], k + 1)
            b.set_state(kw["crp"], kw["cre"], mass=kw["mass"])
            for s in range(k + 1):
                b.set_background(s, w.bg[s])
            t = b.run(0, k, k + 1, w.z, w.dt_snap, path=path)
            b.get_state()
            b.close()
            dt = time.perf_counter() - t0
            wall += dt
            steps += t.cell_steps
            done += 1
            if verbose:
                print(f"[serve] sequential {'plain' if path == ch.PATH_PLAIN else 'fused'} chunk {i + 1}/{len(specs)}: "
                      f"{w.ntracer} tracers x {w.n} bins, {dt:.2f} s", file=sys.stderr, flush=True)
    finally:
        for c in ctxs.values():
            c.close()
    return {"chunks": done, "throughput_substeps_per_s": steps / wall, "tracer_substeps": int(steps),
            "wall_s": wall}
