#!/usr/bin/env python3
"""Benchmark of the 3D Yee FDTD time-stepping path on one MI355X.

Workload (default): 512^3 FP32 sphere medium (seed 42: 200 spheres, radius
U[4,20] cells, eps_r U[1,12], vacuum background), dx = 1, dt = 0.99/sqrt(3),
PEC walls, Gaussian-modulated Jz point source at the centre
J(n) = exp(-((n-30)/10)^2) sin(2 pi 0.05 n), FP64 DFT monitor of Ex and Ez on
the z = 256 plane at 0.03/0.05/0.07 cycles/step, 1000 time steps.

One benchmark "step" = one complete 1000-time-step run of that workload on
device-resident inputs (fields reset to zero before each run, outside the
timed region). value = 512^3 * 1000 * K / (wall time of the K timed runs).

Prints one JSON line (see the keys below). ``--gpus`` above 1 is refused: this
build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic bytes per cell-update, one fixed basis per field format: 6 fields
# x (read + write) + one per-cell material coefficient of the field's width
# (FP32: 48 + 4 = 52 B; binary16: 24 + 2 = 26 B). The 1-byte material id form
# moves 3 (FP32) or 1 (binary16) bytes less; the basis does not follow it, so
# dense and indexed runs stay comparable.
FIELD_BYTES = {"f32": 4, "f16": 2}


def bytes_alg(storage):
    return 6 * FIELD_BYTES[storage] * 2 + FIELD_BYTES[storage]


F16_SCALE_EXP = -8  # binary16 range of the stored fields: ~2.4e-7 .. 256
HBM_PEAK_GBS = 8000.0
# one TILED launch of B steps: the lean kernel over most tiles and, on a second
# stream at the same time, the full kernel over the tiles with the source or
# the monitored plane; launch times span both
TILED_KERNELS = "zstream_lean_kernel + zstream_kernel (concurrent pair per launch)"
PLAIN_KERNELS = "plain_e_kernel + plain_h_kernel (one pair per step)"
FUSED_KERNELS = "fused_step_kernel (one launch per step)"
# the headline metric of the 512^3 workload
METRIC = "Gcell-updates/s at 512\u00b3 + % HBM roofline; chosen vs plain speedup at fixed error"


def parse_traffic(path, want, key="bytes_per_launch"):
    """HBM bytes per launch of a kernel from a committed rocprofv3 PMC summary
    (FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel calibration). Used only
    when the summary holds every line of `want` (the tiling, size, storage and
    material it was measured for)."""
    if not os.path.exists(path):
        return None
    seen, val = set(), None
    with open(path) as f:
        for ln in f:
            if ln.strip() in want:
                seen.add(ln.strip())
            if ln.startswith(key + ":"):
                val = float(ln.split(":", 1)[1])
    return val if seen == set(want) else None


def tiling_line(tiling):
    return "tiling: " + " ".join(f"{k}={tiling[k]}" for k in ("blt", "rows", "waves", "zchunk"))


def measured_hbm(traffic, launches, kernel_ms, cell_updates):
    """Real HBM use from the measured bytes per launch: bytes of all launches
    of a run over their summed duration, and bytes per cell-update."""
    if traffic is None or not kernel_ms:
        return None
    gbs = traffic * launches / (kernel_ms * 1e-3) / 1e9
    return {"GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
            "bytes_per_cell_update": round(traffic * launches / cell_updates, 2)}


def build_problem(gd, n, timesteps, seed=42):
    dt = np.float32(0.99 / math.sqrt(3.0))
    eps = gd.sphere_medium(n, n, n, seed, 200, 4.0, 20.0, 1.0, 12.0)
    src = gd.gaussian_pulse_table(timesteps, 30.0, 10.0, 0.05)
    return dt, eps, src


def cpu_baseline(gd, n, eps, dt, src, steps, threads):
    """Oracle (C port of update_FDTD_omp, gdiamond_omp.hpp:110-172) on the host
    cores for a bounded number of steps of the same workload."""
    from oracle import oracle as O
    cb, db = gd.iso_cb_from_eps(eps, dt, 1.0)
    shape = (n, n, n)
    f = O.zero_fields(shape)
    c = n // 2
    # warm the pages once (first touch) outside the timing
    O.run(shape, f, cb=cb, db=db, dx=1.0, sources=[(2, c, c, c, src)], nsteps=1,
          threads=threads)
    t0 = time.perf_counter()
    O.run(shape, f, cb=cb, db=db, dx=1.0, sources=[(2, c, c, c, src)], n0=1, nsteps=steps,
          threads=threads)
    dt_s = time.perf_counter() - t0
    return n ** 3 * steps / dt_s / 1e9


def host_cpu():
    """CPU model name of the host, from /proc/cpuinfo (None where absent)."""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


ERROR_BOUND = 1e-3  # relative DFT-spectrum error of a form counted as "at fixed error"


def spectral_error(re, im, ref_re, ref_im):
    """max over (component, frequency) of ||X - X_ref||_2 / ||X_ref||_2 on the
    monitor plane."""
    X = np.asarray(re, np.float64) + 1j * np.asarray(im, np.float64)
    R = np.asarray(ref_re, np.float64) + 1j * np.asarray(ref_im, np.float64)
    return float(np.max(np.linalg.norm(X - R, axis=-1) / np.linalg.norm(R, axis=-1)))


def spectra(ctx, gd):
    out = [ctx.dft_get(c) for c in (gd.EX, gd.EZ)]
    return np.stack([o[0] for o in out]), np.stack([o[1] for o in out])


def error_leg(gd, n, T, eps, dt, src, f_cps, tiled_spec, tiled_ms):
    """The same workload in other forms against an FP64 PLAIN reference run:
    DFT-spectrum error and run time of each; the chosen form is the fastest
    within ERROR_BOUND, compared with the FP32 PLAIN and the FP32 TILED forms."""
    c = n // 2

    def one(storage, variant, scale=None):
        ctx = gd.Context(n, n, n, dx=1.0)
        try:
            if storage != gd.F32:
                ctx.set_storage(storage, scale)
            ctx.set_isotropic_eps(eps, dt)
            ctx.add_source(gd.EZ, c, c, c, src)
            ctx.dft_config([gd.EX, gd.EZ], axis=2, index=c, f_cps=f_cps)
            ctx.run(T, variant)
            ms = ctx.last_run_stats()["total_ms"]
            return spectra(ctx, gd), ms
        finally:
            ctx.close()

    ref, ref_ms = one(gd.F64, gd.PLAIN)
    print(f"# fp64 plain reference {ref_ms:.0f} ms", file=sys.stderr, flush=True)
    forms = {"f32-tiled": (tiled_spec, tiled_ms)}
    forms["f32-plain"] = one(gd.F32, gd.PLAIN)
    forms["f32-fused"] = one(gd.F32, gd.FUSED)
    forms["f16-tiled"] = one(gd.F16, gd.TILED, [F16_SCALE_EXP] * 6)
    rep = {}
    for name, (spec, ms) in forms.items():
        rep[name] = {"ms_per_run": round(ms, 2), "GCUPS": round(n ** 3 * T / (ms * 1e-3) / 1e9, 2),
                     "spectrum_err": float(f"{spectral_error(spec[0], spec[1], *ref):.3e}")}
        print(f"# {name}: {rep[name]}", file=sys.stderr, flush=True)
    ok = [k for k in rep if rep[k]["spectrum_err"] <= ERROR_BOUND]
    chosen = min(ok, key=lambda k: rep[k]["ms_per_run"]) if ok else None

    def speedup(base):
        return round(rep[base]["ms_per_run"] / rep[chosen]["ms_per_run"], 3) if chosen else None

    # two baselines: the untiled FP32 path (the reference's plain GPU engine)
    # and the FP32 tiled path (the plain-precision form of the tiled engine)
    return {"bound": ERROR_BOUND, "reference": "fp64 plain, same workload",
            "error": "max over Ex,Ez and 3 frequencies of ||X-X64||_2/||X64||_2 on z=n/2",
            "forms": rep, "plain": "f32-plain", "chosen": chosen,
            "speedup_chosen_vs_plain": speedup("f32-plain"),
            "speedup_chosen_vs_f32_tiled": speedup("f32-tiled")}


def run_precision(gd, n=1024, steps=2, warmup=1, T=2000, error_leg=True):
    """Precision form: FP16 field storage (scale 2^F16_SCALE_EXP, exact FP32 Cb
    per cell), FP32 arithmetic, FP64 DFT, TILED kernel.
    Throughput on the n^3 sphere medium (seed 43) for 2000 steps; error of
    the FP16, BF16 and FP32 forms against an FP64 run of the identical problem
    at 256^3 (same generator and seed, 2000 steps). Returns the record."""
    dt = np.float32(0.99 / math.sqrt(3.0))
    src = gd.gaussian_pulse_table(T, 30.0, 10.0, 0.05)
    f_cps = np.array([0.03, 0.05, 0.07])

    def make(n, storage, eps):
        ctx = gd.Context(n, n, n, dx=1.0)
        if storage == gd.F16:
            ctx.set_storage(gd.F16, [F16_SCALE_EXP] * 6)
        elif storage == gd.BF16:
            ctx.set_storage(gd.BF16)  # float's exponent range: no scale needed
        elif storage == gd.F64:
            ctx.set_storage(gd.F64)
        ctx.set_isotropic_eps(eps, dt)
        c = n // 2
        ctx.add_source(gd.EZ, c, c, c, src)
        ctx.dft_config([gd.EX, gd.EZ], axis=2, index=c, f_cps=f_cps)
        return ctx

    # error at 256^3
    ne = 256
    eps_s = gd.sphere_medium(ne, ne, ne, 43, 200, 4.0, 20.0, 1.0, 12.0)
    res = {}
    ref = None
    forms = (("f64-plain", gd.F64, gd.PLAIN), ("f32-tiled", gd.F32, gd.TILED),
             ("f16-tiled", gd.F16, gd.TILED), ("f16-plain", gd.F16, gd.PLAIN),
             ("bf16-tiled", gd.BF16, gd.TILED))
    for name, storage, variant in (forms if error_leg else ()):
        ctx = make(ne, storage, eps_s)
        ctx.run(T, variant)
        spec = spectra(ctx, gd)
        ms = ctx.last_run_stats()["total_ms"]
        ctx.close()
        if ref is None:
            ref = spec
        res[name] = {"ms_per_run": round(ms, 2),
                     "spectrum_err": float(f"{spectral_error(spec[0], spec[1], *ref):.3e}")}
        print(f"# 256^3 {name}: {res[name]}", file=sys.stderr, flush=True)
    del eps_s

    # throughput at n^3
    t0 = time.perf_counter()
    eps = gd.sphere_medium(n, n, n, 43, 200, 4.0, 20.0, 1.0, 12.0)
    ctx = make(n, gd.F16, eps)
    del eps
    fmt, ntable = ctx.get_material_format()
    material = "indexed" if fmt == gd.MAT_INDEXED else "dense"
    ctx.set_launch_timing(True)
    tiling = ctx.plan_tiling()
    print(f"# {n}^3 setup {time.perf_counter() - t0:.1f}s tiling={tiling} material={fmt},{ntable}",
          file=sys.stderr, flush=True)
    for _ in range(warmup):
        ctx.reset()
        ctx.run(T, gd.TILED)
    ms, kms = [], []
    nlaunch = 0
    for _ in range(steps):
        ctx.reset()
        ctx.run(T, gd.TILED)
        st = ctx.last_run_stats()
        ms.append(st["total_ms"])
        kms.append(st["kernel_ms"])
        nlaunch = st["nlaunch"]
        print(f"# run {st}", file=sys.stderr, flush=True)
    re, im = ctx.dft_get(gd.EZ)
    finite = bool(np.isfinite(re).all() and np.isfinite(im).all())
    ctx.close()
    cells = float(n) ** 3
    value = cells * T / (np.mean(ms) * 1e-3) / 1e9
    b_alg = bytes_alg("f16")
    achieved = b_alg * cells * T / (np.mean(kms) * 1e-3) / 1e9
    traffic_name = f"zstream_{n}cube_f16_{material}_traffic.txt"
    traffic = parse_traffic(os.path.join(ROOT, "profiles", traffic_name),
                            {tiling_line(tiling), "storage: f16", f"material: {material}",
                             f"grid: {n}^3"})
    return {
        "metric": f"Gcell-updates/s at {n}^3, FP16 field storage",
        "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(float(np.mean(ms)), 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f16",
        "data": "synthetic",
        "config": {"workload": f"fdtd3d-yee-spheres-{n}^3-{T}steps-f16", "grid": [n, n, n],
                   "timesteps": T, "medium": "200 spheres seed 43, eps 1-12",
                   "storage": f"binary16 x 2^{F16_SCALE_EXP}", "tiling": tiling,
                   "material": material, "dft_finite": finite,
                   "error_256cube": {"reference": "f64-plain", "forms": res,
                                     "bound": ERROR_BOUND}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": TILED_KERNELS, "bytes_alg_per_cell_update": b_alg,
                     "mean_launch_ms": round(float(np.mean(kms)) / max(1, nlaunch), 4),
                     "launches_per_run": nlaunch,
                     "measured": measured_hbm(traffic, nlaunch, float(np.mean(kms)), cells * T),
                     "traffic_source": "profiles/" + traffic_name if traffic else None},
    }


def run_learned(args, gd):
    """Learned form: 64 seeded fine/coarse solver pairs (fine 96^3, coarse
    48^3, 8 snapshots each) train the corrector; a second pass fine-tunes it on
    rollout samples (the corrected coarse solver's own trajectories, 8
    intervals from a restricted fine state). Then the fine 512^3 x 2000-step
    run (eps 1-6, seed 6000) is the reference for the coarse 256^3 x
    1000-step run without the corrector, with either corrector (applied in
    full every 4 coarse steps), and with the exact correction (the restricted
    fine state itself every 4 coarse steps: the bound of any corrector
    applied that often). Relative DFT-spectrum errors of Ex, Ey on the mid
    plane per frequency and per eps bin of the coarse plane, and run times."""
    from gdiamond_amd import learned
    say = lambda s: print("# " + s, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    raw = learned.make_samples(npairs=args.learned_pairs, n=96, nsnap=8, log=say)
    gen_s = time.perf_counter() - t0
    dx_over_dt = float(np.float32(2.0) / (np.float32(2.0) * learned.DT_FINE))
    cfg = learned.fit_scales(raw, dx_over_dt)
    base = learned.baseline_loss(raw, cfg)
    t0 = time.perf_counter()
    params1, hist1 = learned.train(raw, cfg, epochs=args.learned_epochs, lr=1e-3, lr_min=1e-4,
                                   log=say)
    train_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    roll = learned.make_rollout_samples(params1, cfg, npairs=args.learned_pairs, n=96, log=say)
    roll_base = learned.baseline_loss(roll, cfg)
    params2, hist2 = learned.train(raw + roll, cfg, epochs=args.learned_epochs, lr=3e-4,
                                   lr_min=3e-5, params0=params1, log=say)
    rollout_s = time.perf_counter() - t0
    del raw, roll
    if args.learned_params_out:
        d = os.path.dirname(os.path.abspath(args.learned_params_out))
        os.makedirs(d, exist_ok=True)
        np.savez(args.learned_params_out, params=params2, cfg=np.asarray(cfg))

    n = args.size
    T = args.timesteps * 2  # fine steps (2000 at the default 1000)
    r1 = learned.compare(n, T, params1, cfg, seed=6000, nspheres=200, log=say)
    r2 = learned.compare(n, T, params2, cfg, seed=6000, nspheres=200, log=say)
    ideal = learned.ideal_correction(n, T, seed=6000, nspheres=200, log=say)

    def worst(e):
        v = np.asarray(e, np.float64)
        return float(np.max(v)) if np.all(np.isfinite(v)) else float("inf")

    forms = {"coarse": worst(r1["coarse_err"]), "learned": worst(r1["learned_err"]),
             "learned_rollout": worst(r2["learned_err"]), "exact_correction": worst(ideal["ideal_err"])}
    best = min(("learned", "learned_rollout"), key=lambda k: forms[k])
    best_ms = r1["learned_ms"] if best == "learned" else r2["learned_ms"]
    chosen = best if forms[best] <= ERROR_BOUND else "fine (fallback)"
    out = {
        "metric": "time-to-solution and DFT-spectrum error, learned coarse grid vs fine grid",
        "value": round(r1["fine_ms"] / best_ms, 3),
        "unit": "x faster than the fine run",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(best_ms, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16",
        "data": "synthetic",
        "config": {"workload": f"fdtd3d-yee-learned-{n}^3-{T}steps",
                   "fine": f"{n}^3 x {T} steps, dx 1", "coarse": f"{n // 2}^3 x {T // 2} steps, dx 2",
                   "medium": "200 spheres seed 6000, eps 1-6, radius 4-20 fine cells",
                   "network": "3x3x3 conv 12-16-16-12, BF16 MFMA, FP32 accumulate",
                   "training": {"pairs": args.learned_pairs, "fine": "96^3", "snapshots": 8,
                                "epochs": args.learned_epochs, "gen_s": round(gen_s, 1),
                                "train_s": round(train_s, 1), "rollout_s": round(rollout_s, 1),
                                "loss_first": hist1[0], "loss_last": hist1[-1],
                                "zero_correction_loss": base,
                                "rollout_zero_correction_loss": roll_base,
                                "rollout_loss_first": hist2[0], "rollout_loss_last": hist2[-1]},
                   "fine_ms": round(r1["fine_ms"], 1), "coarse_ms": round(r1["coarse_ms"], 1),
                   "learned_ms": round(r1["learned_ms"], 1),
                   "learned_rollout_ms": round(r2["learned_ms"], 1),
                   "spectrum_err_max": forms, "error_bound": ERROR_BOUND, "chosen_at_bound": chosen,
                   "errors": {"frequencies_cycles_per_fine_step": [0.03, 0.05, 0.07],
                              "components": ["Ex", "Ey"],
                              "eps_bins": [list(b) for b in learned.EPS_BINS],
                              "coarse": r1["coarse_err"], "coarse_eps_bins": r1["coarse_err_eps_bins"],
                              "learned": r1["learned_err"],
                              "learned_eps_bins": r1["learned_err_eps_bins"],
                              "learned_rollout": r2["learned_err"],
                              "learned_rollout_eps_bins": r2["learned_err_eps_bins"],
                              "exact_correction": ideal["ideal_err"],
                              "exact_correction_eps_bins": ideal["ideal_err_eps_bins"]}},
    }
    print(json.dumps(out), flush=True)


def run_batch(gd, batch_size=256, T=500, steps=2, warmup=1,
              modes=("graph", "persistent", "per_instance")):
    """Batched workload: `batch_size` independent FP32 problems, cube edge drawn
    uniformly from {32,48,64,96,128,160} (numpy default_rng(2024)), each with
    its own sphere medium (gd_sphere_medium seed 1000+i, 200 spheres, radius
    U[4,20], eps U[1,12]), a Gaussian Jz point source at its centre with its
    own frequency U[0.02,0.08] cycles/step, T steps. Tilings per size class
    from the autotuner. One step of the benchmark = one full run of every
    instance; value = total cell-updates / mean run time (device events around
    the run, fields reset outside). Returns the record."""
    dt = np.float32(0.99 / math.sqrt(3.0))
    rng = np.random.default_rng(2024)
    sizes = rng.choice([32, 48, 64, 96, 128, 160], size=batch_size)
    freqs = rng.uniform(0.02, 0.08, size=batch_size)
    dims = [(int(s), int(s), int(s)) for s in sizes]
    t0 = time.perf_counter()
    b = gd.Batch(dims, dx=1.0)
    indexed = True  # launches run the 1-byte material ids when every instance has them
    for i, (s, f) in enumerate(zip(sizes, freqs)):
        s = int(s)
        eps = gd.sphere_medium(s, s, s, 1000 + i, 200, 4.0, 20.0, 1.0, 12.0)
        b.set_isotropic_eps(i, eps, dt)
        indexed = indexed and np.unique(eps).size <= 256
        b.add_source(i, gd.EZ, s // 2, s // 2, s // 2, gd.gaussian_pulse_table(T, 30.0, 10.0, f))
    cells = float(sum(int(s) ** 3 for s in sizes))
    print(f"# batch setup {time.perf_counter() - t0:.1f}s, {cells:.3e} cells, "
          f"{b.device_bytes() / 2**30:.1f} GiB", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    classes = b.autotune(reps=2)
    tune_s = time.perf_counter() - t0
    print(f"# autotune {tune_s:.1f}s: {classes}", file=sys.stderr, flush=True)

    def timed(mode, nrun, nwarm):
        for _ in range(nwarm):
            b.reset()
            b.run(T, mode)
        ms = []
        for _ in range(nrun):
            b.reset()
            b.run(T, mode)
            ms.append(b.last_run_stats()["total_ms"])
        return float(np.mean(ms)), b.last_run_stats()["nlaunch"]

    res = {}
    for name, mode, nrun in (("graph", gd.BATCH_GRAPH, steps),
                             ("persistent", gd.BATCH_PERSISTENT, 1),
                             ("per_instance", gd.BATCH_PER_INSTANCE, 1)):
        if name not in modes and name != "graph":
            continue
        ms, nl = timed(mode, nrun, warmup)
        res[name] = {"ms_per_run": round(ms, 2), "GCUPS": round(cells * T / (ms * 1e-3) / 1e9, 2),
                     "launches_per_run": nl}
        print(f"# {name}: {res[name]}", file=sys.stderr, flush=True)
    b.reset()
    b.run(T, gd.BATCH_GRAPH)
    fin = b.get_fields(0)
    finite = bool(all(np.isfinite(x).all() for x in fin))
    b.close()
    g = res["graph"]
    b_alg = bytes_alg("f32")
    achieved = b_alg * cells * T / (g["ms_per_run"] * 1e-3) / 1e9
    material = "indexed" if indexed else "dense"
    # the persistent launches of one run: per launch group, one launch per
    # blt steps (one dispatch holding both tile builds), plus the advance
    nlaunch = g["launches_per_run"] - 1
    # measured bytes of a whole run (the autotuner's picks, and with them the
    # number of launches, can differ between runs); per launch for this run
    traffic_name = f"zstream_batch_{batch_size}inst_traffic.txt"
    run_bytes = parse_traffic(os.path.join(ROOT, "profiles", traffic_name),
                              {f"instances: {batch_size}", f"timesteps: {T}",
                               f"material: {material}"}, key="bytes_per_run")
    traffic = run_bytes / nlaunch if run_bytes else None
    return {
        "metric": "Gcell-updates/s, batched independent problems",
        "value": g["GCUPS"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": g["ms_per_run"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"fdtd3d-yee-batch-{batch_size}x-{T}steps",
                   "sizes": "cube edge from {32,48,64,96,128,160}, default_rng(2024)",
                   "cells": cells, "timesteps": T, "classes": classes, "material": material,
                   "autotune_s": round(tune_s, 1), "modes": res, "finite": finite,
                   "speedup_graph_vs_per_instance":
                       round(res["per_instance"]["ms_per_run"] / g["ms_per_run"], 3)
                       if "per_instance" in res else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "zstream_batch_kernel (one dispatch per launch; launch groups "
                               "run on concurrent streams, so run time / launches is below the "
                               "mean dispatch duration)",
                     "bytes_alg_per_cell_update": b_alg,
                     "run_ms_per_launch": round(g["ms_per_run"] / max(1, nlaunch), 4),
                     "launches_per_run": nlaunch,
                     "measured": measured_hbm(traffic, nlaunch, g["ms_per_run"], cells * T),
                     "traffic_source": "profiles/" + traffic_name if traffic else None},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed runs of the whole workload")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--timesteps", type=int, default=1000)
    ap.add_argument("--variant", choices=["tiled", "plain", "fused"], default="tiled")
    ap.add_argument("--storage", choices=["f32", "f16"], default="f32",
                    help="field storage: FP32, or binary16 with scale 2^%d" % F16_SCALE_EXP)
    ap.add_argument("--material", choices=["auto", "dense", "indexed"], default="auto",
                    help="device form of the isotropic Cb (auto: indexed for <= 256 values)")
    ap.add_argument("--blt", type=int, default=0)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=120,
                    help="time steps of the CPU baseline sample (about 6-10 s on 16 cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error-leg", action="store_true",
                    help="skip the fixed-error comparison (FP64 reference, FP32 PLAIN, FP16)")
    ap.add_argument("--workload", choices=["single", "batch", "precision", "learned"],
                    default="single",
                    help="single: the 512^3 problem; batch: 256 independent problems; "
                         "precision: FP16 storage at 1024^3 with its error at 256^3")
    ap.add_argument("--no-legs", dest="legs", action="store_false",
                    help="skip the precision (1024^3 FP16) and batch legs of the default run")
    ap.add_argument("--batch-size", type=int, default=256)
    ap.add_argument("--learned-pairs", type=int, default=64)
    ap.add_argument("--learned-epochs", type=int, default=6)
    ap.add_argument("--learned-params-out", default="",
                    help="write the trained corrector parameters to this .npz file")
    ap.add_argument("--batch-timesteps", type=int, default=500)
    ap.add_argument("--batch-modes", default="graph,persistent,per_instance",
                    help="batch run modes to time (comma-separated; graph always runs)")
    ap.add_argument("--precision-timesteps", type=int, default=2000)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only the copy kernel used to calibrate HBM counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not a measured path)"}))
        sys.exit(2)

    import gdiamond_amd as gd

    if args.workload == "batch":
        print(json.dumps(run_batch(gd, args.batch_size, args.batch_timesteps, args.steps,
                                   args.warmup, tuple(args.batch_modes.split(",")))), flush=True)
        return
    if args.workload == "precision":
        n = args.size if args.size != 512 else 1024
        print(json.dumps(run_precision(gd, n, args.steps, args.warmup, args.precision_timesteps,
                                       not args.no_error_leg)), flush=True)
        return
    if args.workload == "learned":
        run_learned(args, gd)
        return

    n, T = args.size, args.timesteps
    if args.calibrate_copy:
        ctx = gd.Context(3, 3, 3)
        nfl = 256 * 1024 * 1024
        ms = ctx.calibration_copy(nfl, reps=5)
        print(json.dumps({"copy_floats": nfl, "bytes_read": 4 * nfl, "bytes_written": 4 * nfl,
                          "ms": ms, "GBps": 8 * nfl / ms / 1e6}))
        return

    t_setup = time.perf_counter()
    dt, eps, src = build_problem(gd, n, T)
    ctx = gd.Context(n, n, n, dx=1.0)
    if args.storage == "f16":
        ctx.set_storage(gd.F16, [F16_SCALE_EXP] * 6)
    ctx.set_material_format({"auto": gd.MAT_AUTO, "dense": gd.MAT_DENSE,
                             "indexed": gd.MAT_INDEXED}[args.material])
    ctx.set_isotropic_eps(eps, dt)
    mat_fmt, ntable = ctx.get_material_format()
    material = "indexed" if mat_fmt == gd.MAT_INDEXED else "dense"
    c = n // 2
    ctx.add_source(gd.EZ, c, c, c, src)
    f_cps = np.array([0.03, 0.05, 0.07])
    ctx.dft_config([gd.EX, gd.EZ], axis=2, index=c, f_cps=f_cps)
    ctx.set_launch_timing(True)
    variant = {"tiled": gd.TILED, "plain": gd.PLAIN, "fused": gd.FUSED}[args.variant]
    tiling = None
    if variant == gd.TILED:
        tiling = ctx.plan_tiling()
        for k in ("blt", "rows", "waves", "zchunk"):
            v = getattr(args, k)
            if v:
                tiling[k] = v
    setup_s = time.perf_counter() - t_setup
    print(f"# setup {setup_s:.1f}s tiling={tiling}", file=sys.stderr, flush=True)

    for w in range(args.warmup):
        ctx.reset()
        ctx.run(T, variant, tiling)
        print(f"# warmup {w} {ctx.last_run_stats()}", file=sys.stderr, flush=True)

    wall = []
    kernel_ms = []
    nlaunch = 0
    for s in range(args.steps):
        ctx.reset()
        t0 = time.perf_counter()
        ctx.run(T, variant, tiling)
        wall.append(time.perf_counter() - t0)
        st = ctx.last_run_stats()
        kernel_ms.append(st["kernel_ms"])
        nlaunch = st["nlaunch"]
        print(f"# step {s} wall {wall[-1]*1e3:.1f} ms {st}", file=sys.stderr, flush=True)

    # sanity: the monitor saw the pulse, fields are finite
    re, im = ctx.dft_get(gd.EZ)
    finite = bool(np.isfinite(re).all() and np.isfinite(im).all())
    run_spec = spectra(ctx, gd)
    run_ms = ctx.last_run_stats()["total_ms"]

    # The reference's GPU drivers take host arrays: each call copies the
    # coefficient arrays to the device (gdiamond_gpu.cuh:878-889), steps, and
    # copies the six fields back (gdiamond_gpu.cuh:921-926). The same round
    # trip through the library (eps -> Cb and the material index from host
    # memory, one run, six fields out to pageable host arrays) is timed once
    # here and reported beside `value`, never as it.
    host_io = None
    if variant == gd.TILED:
        ctx.reset()  # fields zeroed on the device, step 0, monitor sums cleared
        t0 = time.perf_counter()
        ctx.set_isotropic_eps(eps, dt)
        t1 = time.perf_counter()
        ctx.run(T, variant, tiling)
        t2 = time.perf_counter()
        back = ctx.get_fields()
        t3 = time.perf_counter()
        del back
        host_io = {"value": round(n ** 3 * T / (t3 - t0) / 1e9, 3), "unit": "Gcell-updates/s",
                   "in_s": round(t1 - t0, 4), "run_s": round(t2 - t1, 4),
                   "out_s": round(t3 - t2, 4),
                   "sample": "one run with the material taken from host eps and the six "
                             "fields copied back to pageable host arrays (PCIe-inclusive)"}
        print(f"# host round trip {host_io}", file=sys.stderr, flush=True)

    cells = n ** 3
    total = sum(wall)
    value = cells * T * args.steps / total / 1e9
    ms_per_step = total / args.steps * 1e3
    mean_launch_ms = sum(kernel_ms) / max(1, nlaunch * args.steps)
    # algorithmic bytes of all launches of one run / their summed duration;
    # with B steps per TILED launch this is an equivalent bandwidth and can
    # exceed the peak. `measured` is the real HBM use from the committed
    # FETCH_SIZE / WRITE_SIZE summary of the same launch.
    b_alg = bytes_alg(args.storage)
    achieved = b_alg * cells * T / (np.mean(kernel_ms) * 1e-3) / 1e9
    if variant == gd.TILED:
        traffic_name = f"zstream_{n}cube_{args.storage}_{material}_traffic.txt"
        want = {tiling_line(tiling), f"storage: {args.storage}", f"material: {material}",
                f"grid: {n}^3"}
    else:
        traffic_name = f"{args.variant}_{n}cube_{args.storage}_{material}_traffic.txt"
        want = {f"storage: {args.storage}", f"material: {material}", f"grid: {n}^3"}
    traffic = parse_traffic(os.path.join(ROOT, "profiles", traffic_name), want)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": {gd.TILED: TILED_KERNELS, gd.PLAIN: PLAIN_KERNELS,
                           gd.FUSED: FUSED_KERNELS}[variant],
                "bytes_alg_per_cell_update": b_alg,
                "mean_launch_ms": round(mean_launch_ms, 4),
                "launches_per_run": nlaunch,
                "measured": measured_hbm(traffic, nlaunch, float(np.mean(kernel_ms)), cells * T),
                "traffic_source": "profiles/" + traffic_name if traffic else None}

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.storage,
        "data": "synthetic",
        "config": {"workload": f"fdtd3d-yee-spheres-{n}^3-{T}steps-dft3",
                   "grid": [n, n, n], "timesteps": T, "medium": "200 spheres seed 42, eps 1-12",
                   "source": "Gaussian Jz at centre", "dft": "Ex,Ez z=256 f=0.03/0.05/0.07 fp64",
                   "variant": args.variant, "storage": args.storage,
                   "material": f"{material} ({ntable} values)" if ntable else material,
                   "tiling": tiling, "parallelism": "single-gpu",
                   "dft_finite": finite},
        "roofline": roofline,
    }
    if host_io:
        out["host_round_trip"] = host_io
    if (not args.no_error_leg and args.storage == "f32" and variant == gd.TILED and
            args.material != "indexed"):
        ctx.close()
        out["fixed_error"] = error_leg(gd, n, T, eps, dt, src, f_cps, run_spec, run_ms)
    if not args.no_cpu_baseline and args.storage == "f32":
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(gd, n, eps, dt, src, args.cpu_steps, threads)
        out["cpu_baseline"] = {"value": round(cpu, 4), "unit": "Gcell-updates/s",
                               "cores": threads, "kind": "port",
                               "sample": f"same {n}^3 workload, steps 1..{args.cpu_steps} of {T} "
                                         f"(after one untimed first-touch step), OpenMP "
                                         f"collapse(2) port of update_FDTD_omp",
                               "host": host_cpu(), "nproc": os.cpu_count()}
        seq_steps = 4
        seq = cpu_baseline(gd, n, eps, dt, src, seq_steps, 1)
        out["cpu_baseline"]["single_thread"] = {
            "value": round(seq, 4), "unit": "Gcell-updates/s",
            "sample": f"steps 1..{seq_steps}, one thread (the update_FDTD_seq loop)"}
    # the precision (1024^3 FP16 storage) and batch (256 instances) workloads
    # measured in the same default run, each with its own roofline
    if (args.legs and args.workload == "single" and n == 512 and args.storage == "f32" and
            variant == gd.TILED):
        ctx.close()
        out["precision"] = run_precision(gd, 1024, steps=2, warmup=1)
        out["batch"] = run_batch(gd, 256, 500, steps=2, warmup=1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
