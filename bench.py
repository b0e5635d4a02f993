"""Benchmark of the per-tracer Fokker-Planck update on one MI355X.

A step is one snapshot update of every tracer of the workload (coefficient
segments, Chang-Cooper assembly and all nsub implicit substeps of every
active species).  The default workload is the bulk-throughput
configuration: 1,000,000 electron tracers, 512 momentum bins, TTD D_pp,
synchrotron + IC + Coulomb + adiabatic losses, 20 snapshots (z = 1 -> 0).
W warmup snapshots (0 .. W-1) run untimed, the spectra are restored to
their initial state, then K snapshots are timed from snapshot 0 with the
device synchronized on both sides (timed snapshot k is the workload's
snapshot k mod 20, the state carrying on).  Inputs (spectra, all snapshot
backgrounds, sources) are resident in HBM before the timed region starts.

The JSON line reports the fused (persistent, cost-ordered) path's
tracer-substeps/s; the plain one-tracer-per-thread path on the same
tracers; the served stream (p50/p99 chunk latency over >= 1000 chunks, and
the same chunks run one at a time through the batch API); the roofline of
the fused kernel; the emission contractions; and the CPU oracle timed on a
bounded sample of the same workload (in a child process with OpenMP bound
to physical cores).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(REPO, "croma-fp_amd"), os.path.join(REPO, "oracle"), os.path.join(REPO, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# FP64 flops per row and species-substep of the solve (reference Thomas: rhs
# add 1, forward elimination 3 + 2, back substitution 2 + 2 in the scan form).
FP64_FLOP_PER_ROW = 10


def progress(msg):
    """Stage line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def load_traffic(name):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE, from
    the committed summary profiles/<name>.txt ("bytes_per_launch = ...")."""
    path = os.path.join(REPO, "profiles", name + ".txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("bytes_per_launch"):
            try:
                return float(line.split("=")[1].split()[0])
            except (IndexError, ValueError):
                return None
    return None


def make_batch(ch, ctx, w):
    b = ch.Batch(ctx, w.ntracer, w.species, w.nsnap)
    b.set_state(w.crp0 if w.protons else None, w.cre0, mass=w.mass)
    for s in range(w.nsnap):
        b.set_background(s, w.bg[s])
        if s in w.qpi or s in w.qepri:
            b.set_sources(s, w.qpi.get(s), w.qepri.get(s))
    return b


def run_snapshots(b, w, k0, count, path, elapsed):
    """Runs `count` snapshots from timed index k0; timed snapshot k is the
    workload's snapshot k mod nsnap (the state carries on).  Returns the
    summed counters and the elapsed time after the run."""
    import croma_hip as ch

    tot = ch.Times()
    k, k1 = k0, k0 + count
    while k < k1:
        s = k % w.nsnap
        s_end = min(w.nsnap, s + (k1 - k))
        t = b.run(s, s_end, w.nsnap, w.z, w.dt_snap, elapsed=elapsed, path=path)
        for f, _ in ch.Times._fields_:
            setattr(tot, f, getattr(tot, f) + getattr(t, f))
        elapsed += float(np.sum(w.dt_snap[s:s_end]))
        k += s_end - s
    return tot, elapsed


def run_path(ch, ctx, w, path, warmup, steps, batch=None, final_state=None):
    """W untimed warmup snapshots from snapshot 0, the initial state restored,
    then `steps` snapshots timed from snapshot 0 with the library's stream
    synchronized before and after.  final_state (a list) receives the
    electron spectra after the timed run."""
    b = batch if batch is not None else make_batch(ch, ctx, w)
    try:
        if warmup:
            run_snapshots(b, w, 0, warmup, path, 0.0)
            b.sync()
            b.set_state(w.crp0 if w.protons else None, w.cre0, mass=w.mass)
        b.sync()
        t0 = time.perf_counter()
        t, _ = run_snapshots(b, w, 0, steps, path, 0.0)
        b.sync()
        wall = time.perf_counter() - t0
        if final_state is not None:
            final_state.append(b.get_state())
    finally:
        b.close()
    return wall, t


def fp32_error(got, want, ntr, n, floor=1e-30, decades=30.0):
    """Max relative error of the FP32-coefficient spectra against the FP64
    ones over entries above the 1e-30 floor and within 30 decades of their
    tracer's peak."""
    g = np.asarray(got).reshape(ntr, n)
    r = np.asarray(want).reshape(ntr, n)
    peak = np.max(np.abs(r), axis=1, keepdims=True)
    m = np.abs(r) > np.maximum(floor, peak * 10.0 ** -decades)
    return float(np.max(np.abs(g[m] - r[m]) / np.abs(r[m]))) if m.any() else 0.0


def coeff_fp32_run(ch, ctx, w, steps, fp64_state=None, fp64_rate=None, sample=20000):
    """The same timed snapshots with FP32 coefficient evaluation (FP64
    weights, factorisation and substeps); its rate and its spectra's error
    against the FP64 path's on a tracer sample."""
    ctx.set_coeff_precision(True)
    try:
        st = []
        wall, t = run_path(ch, ctx, w, ch.PATH_FUSED, 0, steps, final_state=st)
    except ch.CromaHipError as exc:
        return {"error": str(exc)}
    finally:
        ctx.set_coeff_precision(False)
    out = {"value": t.cell_steps / wall, "unit": "tracer-substeps/s", "snapshots": steps,
           "ms_per_step": 1e3 * wall / steps}
    if fp64_rate:
        out["vs_fp64"] = out["value"] / fp64_rate
    if fp64_state is not None:
        idx = np.sort(np.random.default_rng(1).choice(w.ntracer, min(sample, w.ntracer), replace=False))
        rows = (idx[:, None] * w.n + np.arange(w.n)[None, :]).ravel()
        out["max_rel_err_vs_fp64"] = fp32_error(st[0][1][rows], fp64_state[1][rows], len(idx), w.n)
        if w.protons:
            out["max_rel_err_vs_fp64_protons"] = fp32_error(st[0][0][rows], fp64_state[0][rows], len(idx), w.n)
        out["error_sample_tracers"] = int(len(idx))
    return out


def plain_comparison(ch, ctx, w, snapshots, ntracer=0, slice_target_s=60.0):
    """The plain one-tracer-per-thread path against the fused path on the
    same tracers and snapshots (from snapshot 0).  A 65,536-tracer probe
    sizes the run: the tracers are then processed in consecutive slices of
    about slice_target_s seconds each (tracers never interact, so slices do
    the same work as one launch set; each slice prints a progress line)."""
    nt = min(ntracer, w.ntracer) if ntracer else w.ntracer
    probe_n = min(65536, nt)
    pw, pt = run_path(ch, ctx, w.subset(np.arange(probe_n)), ch.PATH_PLAIN, 0, snapshots)
    rate_probe = pt.cell_steps / pw
    per_tracer_s = pw / probe_n
    slice_n = max(probe_n, int(slice_target_s / max(per_tracer_s, 1e-12)))
    progress(f"plain probe: {probe_n} tracers {pw:.1f} s ({rate_probe:.4g} tracer-substeps/s); "
             f"slices of {min(slice_n, nt)} tracers")
    p_wall, p_steps, f_wall, f_steps = 0.0, 0, 0.0, 0
    nslice = 0
    for lo in range(0, nt, slice_n):
        sub = w.subset(np.arange(lo, min(nt, lo + slice_n)))
        sw, st = run_path(ch, ctx, sub, ch.PATH_PLAIN, 0, snapshots)
        fw, ft = run_path(ch, ctx, sub, ch.PATH_FUSED, 0, snapshots)
        p_wall += sw
        p_steps += st.cell_steps
        f_wall += fw
        f_steps += ft.cell_steps
        nslice += 1
        progress(f"plain slice {nslice}: tracers {lo}..{min(nt, lo + slice_n) - 1}: plain {sw:.1f} s, "
                 f"fused {fw:.2f} s")
        del sub
    species = 2 if w.protons else 1
    return {"value": p_steps / p_wall, "snapshots": snapshots, "ntracer": nt, "slices": nslice,
            "wall_s": p_wall, "tracer_substeps": int(p_steps), "fused_same_snapshots": f_steps / f_wall,
            "speedup_fused_vs_plain": (f_steps / f_wall) / (p_steps / p_wall),
            "plain_alg_hbm_gbs": 48.0 * w.n * species * p_steps / p_wall / 1e9,
            "probe": {"ntracer": probe_n, "value": rate_probe}}


def secondary_workload(ch, name, steps=40, warmup=5):
    """Another configuration's fused rate (FP64 and FP32 coefficients) on the
    same device, autotuned like the headline: a line of the result, not the
    headline."""
    import synth
    from workloads import hip_params

    w = synth.make_workload(seed=2026, **synth.CONFIGS[name])
    steps = min(steps, w.nsnap)
    ctx = ch.Context(w.n, hip_params(w))
    try:
        b = make_batch(ch, ctx, w)
        tile = b.autotune(0, w.nsnap, w.z, w.dt_snap)
        fin = []
        wall, t = run_path(ch, ctx, w, ch.PATH_FUSED, warmup, steps, batch=b, final_state=fin)
        rate = t.cell_steps / wall
        out = {"workload": f"{name}: {w.ntracer} tracers x {w.n} bins, "
                           f"{'electron_proton' if w.protons else 'electron_only'}, {w.nsnap} snapshots",
               "value": rate, "unit": "tracer-substeps/s", "snapshots_timed": steps, "ms_per_step": 1e3 * wall / steps,
               "tile": {"rows_per_lane": tile[0], "waves_per_simd": tile[1], "snapshots_per_task": tile[2],
                        "secondary_gemm": tile[3]}}
        out["coeff_fp32"] = coeff_fp32_run(ch, ctx, w, steps, fin[0], rate)
    finally:
        ctx.close()
    return out


def physical_cores_in_affinity():
    """Distinct (package, core) pairs among the CPUs this process may run on."""
    cpus = sorted(os.sched_getaffinity(0))
    seen = set()
    for cpu in cpus:
        base = f"/sys/devices/system/cpu/cpu{cpu}/topology/"
        try:
            pkg = open(base + "physical_package_id").read().strip()
            core = open(base + "core_id").read().strip()
        except OSError:
            return len(cpus), len(cpus)
        seen.add((pkg, core))
    return len(cpus), len(seen)


def lscpu_summary():
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        out = ""
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            info[k.strip()] = v.strip()
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU max MHz": "max_mhz", "CPU MHz": "mhz",
            "CPU(s)": "logical_cpus"}
    return {v: info[k] for k, v in keys.items() if k in info}


def cpu_baseline(w, warmup, seconds_target=15.0, max_threads=16):
    """Oracle (C + OpenMP restatement of the reference CPU path) on a bounded
    tracer sample of the same workload and snapshots, timed in a child
    process with OMP_NUM_THREADS = the physical cores this process may use
    (at most max_threads), OMP_PROC_BIND=close and OMP_PLACES=cores.  The
    child never touches the GPU."""
    logical, physical = physical_cores_in_affinity()
    nthreads = max(1, min(max_threads, physical))
    rng = np.random.default_rng(5)
    nsample = max(nthreads * 48, 64)
    idx = np.sort(rng.choice(w.ntracer, size=min(nsample, w.ntracer), replace=False))
    sub = w.subset(idx)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "sample.npz")
        arrs = dict(n=sub.n, ntracer=sub.ntracer, nsnap=sub.nsnap, species=sub.species, dpp_model=sub.dpp_model,
                    z=sub.z, dt_snap=sub.dt_snap, bg=sub.bg, crp0=sub.crp0, cre0=sub.cre0, mass=sub.mass,
                    grid=np.array([sub.pmin, sub.pmax, sub.pemin, sub.pemax]))
        for s, v in sub.qpi.items():
            arrs[f"qpi_{s}"] = v
        for s, v in sub.qepri.items():
            arrs[f"qepri_{s}"] = v
        np.savez(path, **arrs)
        env = dict(os.environ, OMP_NUM_THREADS=str(nthreads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        res = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", path,
                              str(warmup), str(seconds_target), str(nthreads)],
                             capture_output=True, text=True, env=env, timeout=600)
        if res.returncode != 0:
            return {"error": res.stderr[-2000:]}
        child = json.loads(res.stdout.strip().splitlines()[-1])
    cpu = lscpu_summary()
    cpu.update({"affinity_logical_cpus": logical, "affinity_physical_cores": physical,
                "omp": {"OMP_NUM_THREADS": nthreads, "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}})
    return {"value": child["steps"] / child["seconds"], "unit": "tracer-substeps/s", "cores": nthreads,
            "kind": "port", "cpu": cpu,
            "sample": f"{len(idx)} of {w.ntracer} tracers, snapshots {warmup}..{child['last_snap']} "
                      f"({child['steps']} tracer-substeps in {child['seconds']:.1f} s, OpenMP {nthreads} threads "
                      f"on {nthreads} physical cores)"}


def cpu_baseline_child(path, warmup, seconds_target, nthreads):
    """Child side of cpu_baseline: loads the sample and runs the oracle
    snapshot by snapshot until about seconds_target of CPU work is timed."""
    import pyoracle as po
    import run_oracle
    import synth

    d = np.load(path)
    grid = d["grid"]
    sub = synth.Workload(int(d["n"]), int(d["ntracer"]), int(d["nsnap"]), int(d["species"]), int(d["dpp_model"]),
                         d["z"], d["dt_snap"], d["bg"], d["crp0"], d["cre0"], d["mass"], pmin=float(grid[0]),
                         pmax=float(grid[1]), pemin=float(grid[2]), pemax=float(grid[3]))
    sub.qpi = {int(k[4:]): d[k] for k in d.files if k.startswith("qpi_")}
    sub.qepri = {int(k[6:]): d[k] for k in d.files if k.startswith("qepri_")}
    prm = run_oracle.oracle_params(sub)
    o = po.Oracle(sub.n, sub.pmin, sub.pmax, sub.pemin, sub.pemax)
    secondary = o.secondary_kernel() if sub.protons else (np.zeros(sub.n * sub.n), np.full(sub.n, sub.n, np.int32))
    rows = {}
    crp, cre = sub.crp0.copy(), sub.cre0.copy()
    steps_done, t_total, s = 0, 0.0, warmup % sub.nsnap
    while s < sub.nsnap and (t_total < seconds_target or steps_done == 0):
        for zz in (sub.z[s], sub.z[min(s + 1, sub.nsnap)]):
            if zz not in rows:
                rows[zz] = o.ic_cooling_row(zz)
        t0 = time.perf_counter()
        crp, cre, ns = run_oracle.run_workload(sub, prm, s0=s, s1=s + 1, nthreads=nthreads, oracle=o, crp=crp,
                                               cre=cre, return_nsub=True, secondary=secondary, ic_rows=rows)
        t_total += time.perf_counter() - t0
        steps_done += int(ns[0].sum())
        s += 1
    print(json.dumps({"steps": steps_done, "seconds": t_total, "last_snap": s - 1}))


def emission_bench(ch, ntracer=1_000_000, n=256, nbins=64, seed=11):
    """Emission contractions on the spectra workload of 1e6 tracers x 256
    bins (power laws q ~ U(2, 3.5), log-uniform cutoffs) onto 64 bins: IC
    and pp gamma/neutrino (one GEMM-shaped kernel each, per precision mode)
    and synchrotron (log-B table interpolation).  Kernel tables are
    synthetic positive tables of the reference's shapes (the contraction's
    cost does not depend on their values).  Device time per kernel from HIP
    events on the library's stream; bytes = FP64 spectrum row + FP64 output
    row per tracer; flops = 2 K nbins per tracer (the split-BF16 mode issues
    three MFMA passes of them)."""
    rng = np.random.default_rng(seed)
    ctx = ch.Context(n, ch.default_params())
    out = {"workload": f"{ntracer} tracers x {n} bins -> {nbins} bins", "data": "synthetic"}
    try:
        pe = ctx.grid(1)[0]
        pp = ctx.grid(0)[0]
        q = rng.uniform(2.0, 3.5, ntracer)
        cut = 10.0 ** rng.uniform(2.0, 6.0, ntracer)
        spec = np.exp(-q[:, None] * np.log(pe)[None, :] - pe[None, :] / cut[:, None])
        centres = np.linspace(np.log10(pe[0]), np.log10(pe[-1]), nbins)
        lp = np.log10(pe)
        fic = np.exp(-0.5 * ((lp[None, :] - centres[:, None]) / 0.7) ** 2).ravel()
        lpp = np.log10(pp)
        fk = np.exp(-0.5 * ((lpp[None, :] - centres[:, None] - 1.0) / 0.9) ** 2).ravel()
        n_gas = 10.0 ** rng.uniform(-5.0, -2.0, ntracer)
        modes = {"fp64": ch.EMIS_FP64, "fp32": ch.EMIS_FP32, "mfma_split_bf16x3": ch.EMIS_MFMA_XF32,
                 "mfma_bf16": ch.EMIS_MFMA_BF16}
        byts = ntracer * (8.0 * n + 8.0 * nbins)
        flops = 2.0 * n * nbins * ntracer
        peaks = {"fp64": ("fp64 valu", 78.6), "fp32": ("fp32 valu", 157.3),
                 "mfma_split_bf16x3": ("bf16 mfma", 2500.0), "mfma_bf16": ("bf16 mfma", 2500.0)}
        res = {}
        for kname in ("ic", "hadronic"):
            res[kname] = {}
            for mname, mode in modes.items():
                ms = []
                for _ in range(2):
                    if kname == "ic":
                        _, t = ctx.emission_ic(spec, fic, nbins, mode)
                    else:
                        _, t = ctx.emission_hadronic(spec, n_gas, fk, nbins, -1.0, 5.0, mode)
                    ms.append(t)
                t = ms[-1]
                passes = 3.0 if mode == ch.EMIS_MFMA_XF32 else 1.0
                unit, peak = peaks[mname]
                res[kname][mname] = {"ms": t, "hbm_gbs": byts / (t * 1e-3) / 1e9,
                                     "hbm_frac": byts / (t * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                     "tflops_issued": passes * flops / (t * 1e-3) / 1e12,
                                     "peak_tflops": peak, "peak_unit": unit,
                                     "compute_frac": passes * flops / (t * 1e-3) / 1e12 / peak}
        # synchrotron: the pitch-kernel table (512 log-B planes x 64 frequencies
        # x n bins, 32 pitch angles) is built on the device from the F(x)
        # table, then each tracer interpolates it at its own B
        nlogb, logb_min, inv_dlogb, ntheta = 512, -3.0, 512 / 6.0, 32
        fx = np.load(os.path.join(REPO, "tests", "golden", "fx_table.npz"))
        theta = (np.arange(ntheta) + 0.5) * np.pi / ntheta
        pw = np.sin(theta) ** 2 * (np.pi / ntheta)
        nus = 10.0 ** np.linspace(7.0, 11.0, nbins)
        logy = np.log10(nus[:, None, None] / (1.0e-6 * (1.0 + pe ** 2)[None, :, None] * 4.194284e6 *
                                              np.sin(theta)[None, None, :]))
        table, ms_tab = ctx.synch_kernel_table(nbins, ntheta, nlogb, logb_min, inv_dlogb, fx["logx"], fx["logfx"],
                                               fx["fx"], float(fx["xmin"][0]), logy.ravel(), pw)
        bfield = 10.0 ** rng.uniform(-8.0, -5.0, ntracer)
        res["synch"] = {}
        for mname, mode in (("fp64", ch.EMIS_FP64), ("fp32", ch.EMIS_FP32),
                            ("mfma_split_bf16x3", ch.EMIS_MFMA_XF32)):
            plan = ch.SynchPlan(ctx, table, nbins, nlogb, logb_min, inv_dlogb) if mode == ch.EMIS_MFMA_XF32 else None
            try:
                ms = []
                for _ in range(2):
                    if plan is not None:
                        _, t = plan.run(spec, bfield)
                    else:
                        _, t = ctx.emission_synch(spec, bfield, table, nbins, nlogb, logb_min, inv_dlogb, mode=mode)
                    ms.append(t)
            finally:
                if plan is not None:
                    plan.close()
            t = ms[-1]
            res["synch"][mname] = {"ms": t, "hbm_gbs": (byts + 8.0 * ntracer) / (t * 1e-3) / 1e9,
                                   "hbm_frac": (byts + 8.0 * ntracer) / (t * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                   "table_build_ms": ms_tab}
            if plan is not None:
                # the Chebyshev coefficient tables of the 511 plane intervals
                # are built once per table (the plan); each run sorts the
                # tracers by plane, contracts them and fixes up the tracers
                # beyond the contraction's range
                res["synch"][mname]["plan_prepare_ms"] = plan.prepare_ms
                res["synch"][mname]["includes"] = "plane sort, contraction, range check, FP32 fix-up"
                res["synch"][mname]["mfma_tflops_issued"] = 3.0 * 2.0 * n * nbins * 8 * ntracer / (t * 1e-3) / 1e12
        out["kernels"] = res
        out["bytes_per_call"] = byts
        out["flops_per_call"] = flops
    finally:
        ctx.close()
    return out


def served_latency(ch, chunks=1000, seconds=240.0, seed=3, compare_chunks=48, plain_chunks=6, plain_max_tracers=2048,
                   plain_budget_s=20.0):
    """Poisson stream of mixed chunks through the serving loop, then the
    server's closed-loop capacity on the stream's first chunks against the
    same chunks run one at a time through the batch API; and, on the first
    chunks of at most plain_max_tracers tracers, the same against the plain
    one-tracer-per-thread path run one chunk at a time."""
    import serve_stream

    out = serve_stream.run_stream(ch, count=chunks, seconds=seconds, seed=seed)
    progress(f"served: {out['chunks']} chunks, p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms")
    specs = serve_stream.chunk_specs(compare_chunks, seed)
    out["closed_loop"] = serve_stream.run_closed_loop(specs)
    out["sequential_fused"] = serve_stream.run_sequential(specs, path=ch.PATH_FUSED, verbose=True)
    out["served_vs_sequential_fused"] = (out["closed_loop"]["throughput_substeps_per_s"] /
                                         out["sequential_fused"]["throughput_substeps_per_s"])
    progress(f"served closed loop vs sequential fused: {out['served_vs_sequential_fused']:.3f}")
    small = [s for s in serve_stream.chunk_specs(20 * plain_chunks, seed)
             if s["ntracer"] <= plain_max_tracers][:plain_chunks]
    # the plain path launches per bucket and substep (the reference's
    # structure): its chunks run until plain_budget_s is spent, and the
    # server's closed loop then runs the same chunks
    seq = serve_stream.run_sequential(small, path=ch.PATH_PLAIN, verbose=True, budget_s=plain_budget_s)
    small = small[:seq["chunks"]]
    ps = {"chunks": len(small), "max_tracers": plain_max_tracers, "budget_s": plain_budget_s,
          "closed_loop": serve_stream.run_closed_loop(small), "sequential_plain": seq}
    ps["served_vs_plain"] = (ps["closed_loop"]["throughput_substeps_per_s"] /
                             ps["sequential_plain"]["throughput_substeps_per_s"])
    out["plain_set"] = ps
    progress(f"served closed loop vs sequential plain: {ps['served_vs_plain']:.3f}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=["c0", "c1", "c2"])
    ap.add_argument("--ntracer", type=int, default=0, help="override the configuration's tracer count")
    ap.add_argument("--plain-steps", type=int, default=1, help="snapshots timed on the plain path (0 skips)")
    ap.add_argument("--plain-ntracer", type=int, default=0,
                    help="run the plain-vs-fused comparison on the first N tracers (0: all of them)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-coefficient run")
    ap.add_argument("--no-secondary", action="store_true", help="skip the batched uneven-work (c1) lines")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-chunks", type=int, default=1000)
    ap.add_argument("--serve-seconds", type=float, default=240.0, help="upper bound on the served stream's length")
    ap.add_argument("--emission-ntracer", type=int, default=1_000_000,
                    help="tracers of the emission contraction workload (0 skips it)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile instead of autotuning it")
    ap.add_argument("--no-graphs", action="store_true", help="plain stream launches instead of replayed hipGraphs")
    ap.add_argument("--tile", default="", help="rows_per_lane,waves_per_simd,chain,secondary_gemm (skips the autotuner)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (scale of the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--cpu-baseline-child", nargs=4, metavar=("SAMPLE", "WARMUP", "SECONDS", "THREADS"),
                    help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_baseline_child:
        path, warm, secs, thr = args.cpu_baseline_child
        cpu_baseline_child(path, int(warm), float(secs), int(thr))
        return

    if args.copy_calibration:
        import torch

        n = 256 * 1024 * 1024 // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_dispatch": n * 8, "dispatches": 3}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; the tracer update shards across GPUs "
                                   "without communication, but --gpus > 1 is not run here"}))
        sys.exit(2)

    import croma_hip as ch
    import synth
    from workloads import hip_params

    if not ch.is_available():
        raise SystemExit("no HIP device visible")
    cfg = dict(synth.CONFIGS[args.config])
    if args.ntracer:
        cfg["ntracer"] = args.ntracer
    w = synth.make_workload(seed=2026, **cfg)
    progress(f"workload {args.config}: {w.ntracer} tracers x {w.n} bins, {w.nsnap} snapshots")
    steps = args.steps
    ctx = ch.Context(w.n, hip_params(w))
    if args.no_graphs:
        ctx.set_graphs(False)
    t_tune = 0.0
    if args.tile:
        r_, w_, c_, g_ = (int(v) for v in args.tile.split(","))
        ctx.set_tile(r_, w_, c_, protons=int(w.protons), secondary_gemm=g_)
        tile = ctx.get_tile(w.protons)
    elif args.no_tune:
        tile = ctx.get_tile(w.protons)
    batch = make_batch(ch, ctx, w)
    if not args.tile and not args.no_tune:
        # trials of every tile on the workload's first snapshot (state restored)
        t_tune = time.perf_counter()
        tile = batch.autotune(0, w.nsnap, w.z, w.dt_snap)
        t_tune = time.perf_counter() - t_tune

    progress(f"tile {tile}")
    fp64_final = []
    wall, t = run_path(ch, ctx, w, ch.PATH_FUSED, args.warmup, steps, batch=batch, final_state=fp64_final)
    value = t.cell_steps / wall
    progress(f"fused: {steps} snapshots in {wall:.2f} s ({value:.4g} tracer-substeps/s)")
    species = 2 if w.protons else 1
    dpp_name = ["", "asa", "ttd", "off", "direct"][w.dpp_model]

    out = {
        "metric": "tracer-substeps/s (fused batched path) + p50/p99 chunk latency, served vs plain, 1 MI355X",
        "value": value,
        "unit": "tracer-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded log-uniform backgrounds, power-law spectra"
                + (", DSA-like sources)" if w.qepri else ")"),
        "config": {"workload": f"{args.config}: {w.ntracer} tracers x {w.n} bins, "
                               f"{'electron_proton' if w.protons else 'electron_only'}, "
                               f"dpp={dpp_name}, {w.nsnap} snapshots z=1->0",
                   "ntracer": w.ntracer, "n": w.n, "species": species, "snapshots_timed": steps,
                   "timed_from_snapshot": 0, "warmup_then_state_restored": bool(args.warmup),
                   "tile": {"rows_per_lane": tile[0], "waves_per_simd": tile[1], "snapshots_per_task": tile[2],
                            "secondary_gemm": tile[3], "autotune_s": round(t_tune, 2)}},
        "tracer_substeps": int(t.cell_steps),
        "species_substeps": int(t.species_steps),
        "host_ms": {"ic_rows": t.ic_rows_ms, "nsub": t.nsub_estimate_ms, "order": t.bucket_build_ms},
        "graphs": {"windows_replayed": int(t.graph_launches), "captured": int(t.graph_captures),
                   "refused": int(t.graph_failures)},
    }

    # roofline of the dominant kernel (the fused launch window), from HIP
    # events on the library's stream.  The binding roof is FP64 VALU issue:
    # algorithmic flops 10 n per species-substep.  HBM: algorithmic bytes
    # 48 n per species-substep (the reference's materialised solve) and the
    # measured FETCH_SIZE + WRITE_SIZE bytes of the same window.
    launches = max(int(t.kernel_launches), 1)
    kern_ms = t.kernel_ms / launches
    flops = FP64_FLOP_PER_ROW * w.n * t.species_steps / launches
    alg_bytes = 48.0 * w.n * t.species_steps / launches
    tflops = flops / (kern_ms * 1e-3) / 1e12
    traffic = load_traffic(f"k_fused_{args.config}_traffic")
    split = bool(w.protons and tile[3])
    roof = {"bound": "fp64_valu", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tflops / FP64_PEAK_TFLOPS, "traffic": traffic,
            "kernel": ("k_fused proton pass + k_sec_contract + k_fused electron pass" if split else "k_fused")
                      + f" (one launch window of {tile[2]} snapshot(s))",
            "avg_launch_ms": kern_ms, "alg_flops_per_launch": flops,
            "hbm": {"alg_bytes_per_launch": alg_bytes, "measured_bytes_per_launch": traffic,
                    "measured_gbs": traffic / (kern_ms * 1e-3) / 1e9 if traffic else None,
                    "measured_frac": traffic / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                    "peak_gbs": HBM_PEAK_GBS}}
    assert roof["frac"] <= 1.0, roof
    if roof["hbm"]["measured_frac"] is not None:
        assert roof["hbm"]["measured_frac"] <= 1.0, roof
    out["roofline"] = roof

    if not args.no_fp32:
        progress("FP32 coefficients")
        out["coeff_fp32"] = coeff_fp32_run(ch, ctx, w, steps, fp64_final[0], value)
        progress(f"FP32 coefficients: {out['coeff_fp32']}")
    del fp64_final

    if args.plain_steps > 0:
        out["plain"] = plain_comparison(ch, ctx, w, args.plain_steps, args.plain_ntracer)
    ctx.close()
    del batch

    if not args.no_secondary and args.config != "c1":
        progress("batched uneven work (c1)")
        out["c1"] = secondary_workload(ch, "c1")
        progress(f"c1: {out['c1']['value']:.4g} tracer-substeps/s")

    if args.emission_ntracer > 0:
        try:
            progress("emission")
            out["emission"] = emission_bench(ch, ntracer=args.emission_ntracer)
        except (ch.CromaHipError, MemoryError) as exc:
            out["emission"] = {"error": str(exc)}

    if not args.no_serve:
        progress("served stream")
        try:
            out["served"] = served_latency(ch, chunks=args.serve_chunks, seconds=args.serve_seconds)
        except ch.CromaHipError as exc:
            # a failed request is reported in the line, not in place of it
            out["served"] = {"error": str(exc)}

    if not args.no_cpu:
        progress("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(w, args.warmup)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
