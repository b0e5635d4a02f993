"""Benchmark of the per-tracer Fokker-Planck update on one MI355X.

A step is one snapshot update of every tracer of the workload (coefficient
segments, Chang-Cooper assembly and all nsub implicit substeps of both active
species).  The default workload is the batched uneven-work configuration:
10,000 electron_proton tracers, 256 momentum bins, ASA D_pp, DSA-like sources
on 5% of (tracer, snapshot) pairs, 100 snapshots (z = 1 -> 0).  W warmup
snapshots run untimed, then K snapshots are timed with the device
synchronized on both sides.  Inputs (spectra, all snapshot backgrounds,
sources) are resident in HBM before the timed region starts.

The JSON line reports the fused (persistent, cost-ordered) path's
tracer-substeps/s; the plain one-tracer-per-thread path on the same
snapshots; a served-stream latency summary; the roofline of the fused
kernel; and the CPU oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
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


def run_path(ch, ctx, w, path, warmup, steps, batch=None):
    b = batch if batch is not None else make_batch(ch, ctx, w)
    try:
        if warmup:
            b.run(0, warmup, w.nsnap, w.z, w.dt_snap, path=path)
        b.sync()
        t0 = time.perf_counter()
        t = b.run(warmup, warmup + steps, w.nsnap, w.z, w.dt_snap, elapsed=float(np.sum(w.dt_snap[:warmup])),
                  path=path)
        b.sync()
        wall = time.perf_counter() - t0
    finally:
        b.close()
    return wall, t


def cpu_baseline(w, warmup, seconds_target=15.0):
    """Oracle (C + OpenMP restatement of the reference CPU path) on a bounded
    tracer sample of the same workload and snapshots.  The per-grid secondary
    weights and the per-snapshot IC cooling rows (identical for every tracer)
    are computed before timing, so the sample measures the per-tracer rate."""
    import pyoracle as po
    import run_oracle

    nthreads = os.cpu_count() or 1
    rng = np.random.default_rng(5)
    nsample = max(nthreads * 48, 64)
    idx = np.sort(rng.choice(w.ntracer, size=min(nsample, w.ntracer), replace=False))
    sub = w.subset(idx)
    prm = run_oracle.oracle_params(sub)
    o = po.Oracle(sub.n, sub.pmin, sub.pmax, sub.pemin, sub.pemax)
    secondary = o.secondary_kernel() if sub.protons else (np.zeros(sub.n * sub.n), np.full(sub.n, sub.n, np.int32))
    rows = {}
    crp, cre = sub.crp0.copy(), sub.cre0.copy()
    steps_done, t_total, s = 0, 0.0, warmup
    while s < w.nsnap and (t_total < seconds_target or s == warmup):
        for zz in (sub.z[s], sub.z[min(s + 1, sub.nsnap)]):
            if zz not in rows:
                rows[zz] = o.ic_cooling_row(zz)
        t0 = time.perf_counter()
        crp, cre, ns = run_oracle.run_workload(sub, prm, s0=s, s1=s + 1, nthreads=nthreads, oracle=o, crp=crp,
                                               cre=cre, return_nsub=True, secondary=secondary, ic_rows=rows)
        t_total += time.perf_counter() - t0
        steps_done += int(ns[0].sum())
        s += 1
    return {"value": steps_done / t_total, "unit": "tracer-substeps/s", "cores": nthreads, "kind": "port",
            "sample": f"{len(idx)} of {w.ntracer} tracers, snapshots {warmup}..{s - 1} ({steps_done} "
                      f"tracer-substeps in {t_total:.1f} s, OpenMP {nthreads} threads)"}


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
        for mname, mode in (("fp64", ch.EMIS_FP64), ("fp32", ch.EMIS_FP32)):
            ms = []
            for _ in range(2):
                _, t = ctx.emission_synch(spec, bfield, table, nbins, nlogb, logb_min, inv_dlogb, mode=mode)
                ms.append(t)
            t = ms[-1]
            res["synch"][mname] = {"ms": t, "hbm_gbs": (byts + 8.0 * ntracer) / (t * 1e-3) / 1e9,
                                   "hbm_frac": (byts + 8.0 * ntracer) / (t * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                   "table_build_ms": ms_tab}
        out["kernels"] = res
        out["bytes_per_call"] = byts
        out["flops_per_call"] = flops
    finally:
        ctx.close()
    return out


def served_latency(ch, seconds=20.0, seed=3):
    """Short Poisson stream of mixed chunks through the serving loop."""
    import serve_stream

    return serve_stream.run_stream(ch, seconds=seconds, seed=seed)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c1", choices=["c0", "c1", "c2"])
    ap.add_argument("--ntracer", type=int, default=0, help="override the configuration's tracer count")
    ap.add_argument("--plain-steps", type=int, default=3, help="snapshots timed on the plain path (0 skips)")
    ap.add_argument("--plain-ntracer", type=int, default=0,
                    help="run the plain-vs-fused comparison on the first N tracers (rates are per tracer-substep)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--emission-ntracer", type=int, default=1_000_000,
                    help="tracers of the emission contraction workload (0 skips it)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile instead of autotuning it")
    ap.add_argument("--no-graphs", action="store_true", help="plain stream launches instead of replayed hipGraphs")
    ap.add_argument("--tile", default="", help="rows_per_lane,waves_per_simd,chain,secondary_gemm (skips the autotuner)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (scale of the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

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
    steps = min(args.steps, w.nsnap - args.warmup)
    ctx = ch.Context(w.n, hip_params(w))
    if args.no_graphs:
        ctx.set_graphs(False)
    t_tune = time.perf_counter()
    if args.tile:
        r_, w_, c_, g_ = (int(v) for v in args.tile.split(","))
        ctx.set_tile(r_, w_, c_, protons=int(w.protons), secondary_gemm=g_)
        tile = ctx.get_tile(w.protons)
    elif args.no_tune:
        tile = ctx.get_tile(w.protons)
    t_tune = time.perf_counter() - t_tune
    batch = make_batch(ch, ctx, w)
    if not args.tile and not args.no_tune:
        # trials of every tile on the workload's first snapshot (state restored)
        t_tune = time.perf_counter()
        tile = batch.autotune(0, w.nsnap, w.z, w.dt_snap)
        t_tune = time.perf_counter() - t_tune

    progress(f"tile {tile}")
    wall, t = run_path(ch, ctx, w, ch.PATH_FUSED, args.warmup, steps, batch=batch)
    value = t.cell_steps / wall
    progress(f"fused: {steps} snapshots in {wall:.2f} s")
    species = 2 if w.protons else 1

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
        "data": "synthetic (seeded log-uniform backgrounds, power-law spectra, DSA-like sources)",
        "config": {"workload": f"{args.config}: {w.ntracer} tracers x {w.n} bins, "
                               f"{'electron_proton' if w.protons else 'electron_only'}, "
                               f"dpp={['', 'asa', 'ttd', 'off'][w.dpp_model]}, {w.nsnap} snapshots z=1->0",
                   "ntracer": w.ntracer, "n": w.n, "species": species, "snapshots_timed": steps,
                   "tile": {"rows_per_lane": tile[0], "waves_per_simd": tile[1], "snapshots_per_task": tile[2],
                            "secondary_gemm": tile[3], "autotune_s": round(t_tune, 2)}},
        "tracer_substeps": int(t.cell_steps),
        "species_substeps": int(t.species_steps),
        "host_ms": {"ic_rows": t.ic_rows_ms, "nsub": t.nsub_estimate_ms, "order": t.bucket_build_ms},
        "graphs": {"windows_replayed": int(t.graph_launches), "captured": int(t.graph_captures),
                   "refused": int(t.graph_failures)},
    }

    # roofline of the dominant kernel (k_fused), from per-launch HIP events on
    # the library's stream: algorithmic bytes 48 n per species-substep
    kern_ms = t.kernel_ms / max(t.kernel_launches, 1)
    alg_bytes = 48.0 * w.n * t.species_steps / max(t.kernel_launches, 1)
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    flops = FP64_FLOP_PER_ROW * w.n * t.species_steps / max(t.kernel_launches, 1)
    split = bool(w.protons and tile[3])
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(f"k_fused_{args.config}_traffic"),
                       "kernel": ("k_fused proton pass + secondary GEMM + k_fused electron pass"
                                  if split else "k_fused") + f" (one launch window of {tile[2]} snapshot(s))",
                       "avg_launch_ms": kern_ms, "alg_bytes_per_launch": alg_bytes,
                       "fp64_valu": {"achieved_tflops": flops / (kern_ms * 1e-3) / 1e12, "peak_tflops": FP64_PEAK_TFLOPS,
                                     "frac": flops / (kern_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}}

    if args.plain_steps > 0:
        wp = w.subset(np.arange(min(args.plain_ntracer, w.ntracer))) if args.plain_ntracer else w
        pw, pt = run_path(ch, ctx, wp, ch.PATH_PLAIN, args.warmup, min(args.plain_steps, steps))
        progress(f"plain: {pt.cell_steps / pw:.4g} tracer-substeps/s")
        fw, ft = run_path(ch, ctx, wp, ch.PATH_FUSED, args.warmup, min(args.plain_steps, steps))
        out["plain"] = {"value": pt.cell_steps / pw, "snapshots": min(args.plain_steps, steps), "ntracer": wp.ntracer,
                        "fused_same_snapshots": ft.cell_steps / fw,
                        "speedup_fused_vs_plain": (ft.cell_steps / fw) / (pt.cell_steps / pw)}
    ctx.close()

    if args.emission_ntracer > 0:
        try:
            progress("emission")
            out["emission"] = emission_bench(ch, ntracer=args.emission_ntracer)
        except (ch.CromaHipError, MemoryError) as exc:
            out["emission"] = {"error": str(exc)}

    if not args.no_serve:
        progress("served stream")
        try:
            out["served"] = served_latency(ch, seconds=args.serve_seconds)
        except (ImportError, AttributeError, OSError) as exc:  # serving loop not built
            out["served"] = {"error": str(exc)}

    if not args.no_cpu:
        progress("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(w, args.warmup)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
