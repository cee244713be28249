#!/usr/bin/env python3
"""Benchmark of the k-space first-order time loop on one MI355X.

Default workload: 3D 256x256x256 FP32, dx = 1e-4 m, PML 10, heterogeneous seeded Gaussian random
fields c in [1400, 1600] m/s (seed 3) and rho in [900, 1100] kg/m^3 (seed 4), correlation 20
cells; power-law absorption alpha = 0.5 dB/(MHz^y cm), y = 1.1 with dispersion; linear; p0 = ball
of radius 8 cells, 1 Pa, smoothed; dt from CFL 0.3 on max(c); nt = 1000; sensor = plane x = 168
recording p and uz.

A "step" is one time step of the loop. `value` is Mcell-steps/s = N * K / elapsed / 1e6 over K
timed steps with every input resident in HBM. One JSON line is printed.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "k-wave-swift_amd"))

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12


def workload(name, dtype):
    from kwavehip import (AbsorptionMode, DTypePrecision, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource,
                          RecordField, SimulationOptions, makeBall, makeCircle, makeDisc, random_field)
    dt_ = DTypePrecision.float64 if dtype == "f64" else DTypePrecision.float32
    if name == "c2":
        from kwavehip import workloads
        grid, med, src, sens, opts = workloads.throughput_3d(n=256, steps=1000, dtype=dt_)
        info = {"workload": "3D 256^3 heterogeneous power-law (y=1.1) p0-ball, plane sensor p+uz",
                "grid": [256, 256, 256], "pml": 10, "nt": 1000}
        # byte model per step: s*N*(2*d*T + S), T = 15 transforms, S = 27 other full-grid arrays
        model = {"T": 15, "S": 27, "d": 3}
    elif name == "slab":
        # one x slab of a 512^3 grid split over 8 GPUs: the per-GPU shape of the local z / y passes
        from kwavehip import workloads
        grid, med, src, sens, opts = workloads.throughput_3d(steps=1000, dtype=dt_, plane=40, shape=(64, 512, 512))
        info = {"workload": "3D 64x512x512 (one x slab of 512^3 over 8 GPUs) heterogeneous power-law (y=1.1) "
                            "p0-ball, plane sensor p+uz", "grid": [64, 512, 512], "pml": 10, "nt": 1000}
        model = {"T": 15, "S": 27, "d": 3}
    elif name == "c0":
        n, dx = 128, 1e-4
        grid = KWaveGrid(n, dx, n, dx)
        grid.makeTime(1500, cfl=0.3, tEnd=4e-6)
        med = KWaveMedium(1500.0, 1000.0)
        src = KWaveSource(p0=makeDisc(n, n, radius=5))
        sens = KWaveSensor(mask=makeCircle(n, n, radius=40), record=RecordField.p | RecordField.pFinal | RecordField.pMax)
        opts = SimulationOptions(pmlSize=20, smoothP0=True, dtype=dt_)
        info = {"workload": "2D 128^2 homogeneous p0-disc IVP, circle sensor", "grid": [n, n], "pml": 20, "nt": grid.nt}
        model = {"T": 7, "S": 5, "d": 2}
    else:
        raise SystemExit(f"unknown workload {name}")
    return grid, med, src, sens, opts, info, model


def measured_traffic(workload, dtype, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), or None when that kernel was not counted."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    rec = table.get(f"{workload}/{dtype}/{kernel}")
    if not rec:
        return None
    return {"bytes": rec["fetch_bytes"] + rec["write_bytes"], "source": rec["summary"]}


def precision_check(grid, med, src, sens, opts):
    """The whole workload (all nt steps) in FP32 and in FP64 on the GPU: the FP64 throughput and
    the FP32 error against FP64 (relL2 of the sensor series; pFinal residual normalised by
    max_t |p| at the sensor, as the FP32 reference tests do)."""
    import dataclasses
    from kwavehip import DTypePrecision, Simulation
    outs, rate = {}, {}
    for name, dt in (("f32", DTypePrecision.float32), ("f64", DTypePrecision.float64)):
        sim = Simulation(grid, med, src, sens, dataclasses.replace(opts, dtype=dt))
        sim.launch(0, 1)
        sim.sync()
        t0 = time.perf_counter()
        sim.launch(1, grid.nt)
        sim.sync()
        rate[name] = grid.totalGridPoints * (grid.nt - 1) / (time.perf_counter() - t0) / 1e6
        outs[name] = sim.outputs()
        sim.close()
    o32, o64 = outs["f32"], outs["f64"]
    def rel_l2(a, b):
        a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
        return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b * b)), 1e-300))
    err = {}
    for f in ("p", "ux", "uy", "uz"):
        a, b = getattr(o32, f), getattr(o64, f)
        if a is not None and b is not None:
            err[f"{f}_relL2"] = rel_l2(a, b)
    if o32.pFinal is not None and o64.p is not None:
        d = np.asarray(o32.pFinal, np.float64) - o64.pFinal
        err["pFinal_rms_over_max_p"] = float(np.sqrt(np.mean(d * d)) / max(np.abs(o64.p).max(), 1e-300))
    return {"steps": grid.nt, "f64_Mcell_steps_per_s": round(rate["f64"], 2),
            "f32_Mcell_steps_per_s_full_run": round(rate["f32"], 2),
            "error_fp32_vs_fp64": {k: float(f"{v:.3e}") for k, v in err.items()}}


def scan_line_cases(n_runs=48, nt=800, seed=7):
    """Per-scan-line transmits (kwavehip.workloads.scan_line_batch): 48 independent 3D FP32 runs,
    sizes drawn from {96^3, 128x96x96, 160x128x128}, speckle media, a 32-element linear array at
    x = 10 with per-run steering and focus, nt = 800, receive on the active voxels (p), combined
    per element."""
    from kwavehip import workloads
    return workloads.scan_line_batch(n_runs, nt, seed)


def batch_comparison(n_runs=48, nt=800):
    """Batched runs against the same runs launched one after another (kwh_sim_run each), both
    timed from device-resident inputs to combined traces, every run in its default mode (the y-z
    plane kernels, kwh_sim_set_plane, which cover every run of this workload). Batched modes:
    fused (one launch per pass of the step over all runs of a grid shape, each shape on its own
    stream, the batched steps replayed from hipGraphs), graphs (each run replays its own step
    graph on its own stream), streams (each run enqueued on its own stream). Sequential
    baselines: plain (default mode), plain_autotuned (autotuned launch schedule). The split
    passes (kwh_sim_set_plane(0)) are reported as plain_split / fused_split."""
    from kwavehip import Simulation, SimulationBatch
    cases = scan_line_cases(n_runs, nt)
    cells = sum(c[0][0].totalGridPoints for c in cases) * nt

    def sequential(split=False, tune=False):
        total, traces = 0.0, []
        for args, tr in cases:
            sim = Simulation(*args)
            if tune:
                sim.autotune()   # tuned once per grid shape, outside the timing
            if split:
                sim.set_plane(False)
            t0 = time.perf_counter()
            sim.run()
            traces.append(sim.transducer_traces(tr))
            total += time.perf_counter() - t0
            sim.close()
        return total, traces

    def rel_l2(a, b):
        d = np.asarray(a, np.float64) - b
        return float(np.sqrt(np.sum(d * d) / max(np.sum(np.asarray(b, np.float64) ** 2), 1e-300)))

    plain_s, ref = sequential()
    res = {"runs": n_runs, "steps": nt, "cells_total": int(cells // nt),
           "plain_Mcell_steps_per_s": round(cells / plain_s / 1e6, 1)}
    tuned_s, _ = sequential(tune=True)
    res["plain_autotuned_Mcell_steps_per_s"] = round(cells / tuned_s / 1e6, 1)
    split_s, ref_split = sequential(split=True)
    res["plain_split_Mcell_steps_per_s"] = round(cells / split_s / 1e6, 1)
    for mode, use_graph, fused, split in (("fused", True, True, False), ("fused_split", True, True, True),
                                          ("graphs", True, False, False), ("streams", False, False, False)):
        batch = SimulationBatch([c[0] for c in cases])
        if split:
            for sim in batch.sims:
                sim.set_plane(False)
        t0 = time.perf_counter()
        batch.run(use_graph=use_graph, fused=fused)
        traces = [sim.transducer_traces(tr) for sim, (_, tr) in zip(batch.sims, cases)]
        batch_s = time.perf_counter() - t0
        # bitwise against the sequential runs in the same mode; relL2 against the plain runs
        same = all(np.array_equal(a, b) for a, b in zip(traces, ref_split if split else ref))
        batch.close()
        res[f"batched_{mode}_Mcell_steps_per_s"] = round(cells / batch_s / 1e6, 1)
        res[f"speedup_{mode}"] = round((split_s if split else plain_s) / batch_s, 3)
        res[f"traces_identical_{mode}"] = bool(same)
        res[f"traces_relL2_vs_plain_{mode}"] = float(f"{max(rel_l2(a, b) for a, b in zip(traces, ref)):.3e}")
        if mode == "fused":
            res["speedup_fused_vs_autotuned_plain"] = round(tuned_s / batch_s, 3)
            res["speedup_fused_vs_plain_split"] = round(split_s / batch_s, 3)
    res["speedup"] = res["speedup_fused"]
    return res


def adjoint_case():
    """2D sound-speed / photoacoustic inversion (kwavehip.workloads.inversion_2d): 512^2, 1500
    steps, 256-point Cartesian circle, L = 1/2 ||p - p_obs||^2 with p_obs from a 1% perturbation."""
    from kwavehip import workloads
    return workloads.inversion_2d(n=512, steps=1500)


def adjoint_comparison(n_dirs=16):
    """Adjoint time (reverse sweep incl. recompute) over forward time, and the c-gradient against
    central differences of the FP64 forward along random directions that leave max(c) alone."""
    from kwavehip import DTypePrecision, KWaveMedium, KWaveSource, Simulation
    import dataclasses
    grid, c, c_obs, p0, sens, opts = adjoint_case()
    def forward(cc, dtype=DTypePrecision.float32, ckpt=0):
        sim = Simulation(grid, KWaveMedium(cc, np.float32(1000.0)), KWaveSource(p0=p0), sens,
                         dataclasses.replace(opts, dtype=dtype))
        if ckpt:
            sim.set_checkpointing(ckpt)
        t0 = time.perf_counter()
        sim.run()
        el = time.perf_counter() - t0
        return sim, sim.outputs().p.astype(np.float64), el
    obs, y_obs, _ = forward(c_obs)
    obs.close()
    plain, _, _ = forward(c)
    plain.close()
    _, _, t_fwd = forward(c)
    sim, y, t_fwd_ckpt = forward(c, ckpt=32)
    t0 = time.perf_counter()
    gp0, gc = sim.vjp((y - y_obs).astype(np.float32), checkpoint_every=32)
    t_adj = time.perf_counter() - t0
    sim.close()
    # central differences of the FP64 forward (c enters as float32: steps of 0.25 m/s are exact)
    sim64, y64, _ = forward(c, DTypePrecision.float64)
    sim64.close()
    c64 = c.astype(np.float64)
    keep = c64 < c64.max() - 20.0
    g64 = gc.astype(np.float64)
    scale = float(np.sqrt(np.sum(np.where(keep, g64, 0.0) ** 2)))   # size of <g, v> for random signs v
    rng = np.random.default_rng(14)
    errs, rel = [], []
    for _ in range(n_dirs):
        v = np.where(keep, np.sign(rng.standard_normal(c.shape)), 0.0)
        h = 0.125
        lp = 0.5 * np.sum((forward((c64 + h * v).astype(np.float32), DTypePrecision.float64)[1] - y_obs) ** 2)
        lm = 0.5 * np.sum((forward((c64 - h * v).astype(np.float32), DTypePrecision.float64)[1] - y_obs) ** 2)
        fd = (lp - lm) / (2 * h)
        an = float(np.sum(g64 * v))
        errs.append(abs(fd - an) / scale)
        rel.append(abs(fd - an) / max(abs(fd), 1e-300))
    return {"grid": [512, 512], "steps": grid.nt, "checkpoint_every": 32,
            "t_forward_s": round(t_fwd, 4), "t_forward_checkpointed_s": round(t_fwd_ckpt, 4),
            "t_adjoint_s": round(t_adj, 4), "adjoint_over_forward": round(t_adj / t_fwd, 3),
            "fd_directions": n_dirs, "grad_c_fd_max_err_over_grad_norm": float(f"{max(errs):.3e}"),
            "grad_c_fd_median_rel_err": float(f"{float(np.median(rel)):.3e}"),
            "grad_dtype": "f32", "fd_dtype": "f64"}


# The reference's own k-space benchmark scenarios (Sources/KWaveBench/main.swift:84-222) with the
# wall times read off its published chart (Sources/KWaveBench/k-wave-python vs swift.png, MLX on an
# unspecified Apple-silicon GPU, +-15%), as Mcell-steps/s; None where only a speedup is legible.
REF_PUBLISHED = {"kspace2d_64": 18.0, "kspace2d_128": 70.0, "kspace2d_256": 160.0, "kspace2d_512": 224.0,
                 "kspace2d_1024": 283.0, "kspace2d_256_hetero": None, "kspace2d_256_absorbing": None,
                 "kspace2d_256_source": None, "kspace3d_32": 65.0, "kspace3d_64": 125.0, "kspace3d_128": 137.0}


def reference_scenarios(repeats=5):
    """Each scenario as KWaveBench times it: the whole solver call (setup, the loop, read-back of
    the output), one warmup then the median of `repeats`; dx = 1e-4 m, c = 1500 m/s,
    rho = 1000 kg/m^3, dt from CFL 0.3, PML inside (20 in 2D, 10 in 3D), smoothed p0."""
    from kwavehip import (AbsorptionMode, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField,
                          SimulationOptions, SourceMode, kspaceFirstOrder, makeBall, makeDisc, toneBurst)

    def step_map(shape, base):                     # main.swift:77-83
        m = np.full(shape, base, np.float32)
        m[shape[0] // 2:] = np.float32(base * 1.2)
        return m

    def case2d(n, nt, kind=""):
        grid = KWaveGrid(n, 1e-4, n, 1e-4)
        grid.makeTime(1500.0, cfl=0.3)
        grid.setTime(nt, grid.dt)
        med = KWaveMedium(1500.0, 1000.0)
        if kind == "hetero":
            med = KWaveMedium(step_map((n, n), 1500.0), step_map((n, n), 1000.0))
        elif kind == "absorbing":
            med = KWaveMedium(1500.0, 1000.0, alphaCoeff=0.75, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw)
        if kind == "source":                       # main.swift:127-163
            pm = np.zeros((n, n), np.float32)
            pm[20, :] = 1
            sm = np.zeros((n, n), np.float32)
            sm[n - 20, :] = 1
            src = KWaveSource(pMask=pm, p=toneBurst(1 / grid.dt, 1e6, 5).astype(np.float32),
                              pMode=SourceMode.additive)
            return grid, med, src, KWaveSensor(mask=sm, record=RecordField.p), SimulationOptions(pmlSize=20,
                                                                                                 smoothP0=False)
        src = KWaveSource(p0=makeDisc(n, n, radius=n / 16))
        return grid, med, src, KWaveSensor(record=RecordField.pFinal), SimulationOptions(pmlSize=20, smoothP0=True)

    def case3d(n, nt):
        grid = KWaveGrid(n, 1e-4, n, 1e-4, n, 1e-4)
        grid.makeTime(1500.0, cfl=0.3)
        grid.setTime(nt, grid.dt)
        src = KWaveSource(p0=makeBall(n, n, n, radius=n / 8))
        return (grid, KWaveMedium(1500.0, 1000.0), src, KWaveSensor(record=RecordField.pFinal),
                SimulationOptions(pmlSize=10, smoothP0=True))

    cases = [("kspace2d_64", case2d(64, 256)), ("kspace2d_128", case2d(128, 256)), ("kspace2d_256", case2d(256, 256)),
             ("kspace2d_512", case2d(512, 256)), ("kspace2d_1024", case2d(1024, 256)),
             ("kspace2d_256_hetero", case2d(256, 256, "hetero")), ("kspace2d_256_absorbing", case2d(256, 256, "absorbing")),
             ("kspace2d_256_source", case2d(256, 256, "source")), ("kspace3d_32", case3d(32, 128)),
             ("kspace3d_64", case3d(64, 128)), ("kspace3d_128", case3d(128, 128))]
    out = {}
    for name, (grid, med, src, sens, opts) in cases:
        kspaceFirstOrder(grid, med, src, sens, opts)
        secs = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            kspaceFirstOrder(grid, med, src, sens, opts)
            secs.append(time.perf_counter() - t0)
        med_s = sorted(secs)[len(secs) // 2]
        v = grid.totalGridPoints * grid.nt / med_s / 1e6
        ref = REF_PUBLISHED[name]
        out[name] = {"median_s": round(med_s, 5), "Mcell_steps_per_s": round(v, 1), "reference_mlx_gpu": ref,
                     "ratio": round(v / ref, 1) if ref else None}
        print(f"[bench] {name}: {med_s * 1e3:.2f} ms, {v:.0f} Mcell-steps/s", file=sys.stderr, flush=True)
    return out


def device_copy_bandwidth():
    """Event-timed bandwidth of 1 GiB -> 1 GiB device copies (kwavehip/copy_probe, a child
    process): the plain grid-stride copy (its PMC counts calibrate FETCH_SIZE / WRITE_SIZE to
    bytes) and the nontemporal one-pass copy (the fastest copy found). Reference points beside
    the solver's kernels, not ceilings."""
    import subprocess
    exe = os.path.join(ROOT, "k-wave-swift_amd", "kwavehip", "copy_probe")
    if not os.path.exists(exe):
        print("[bench] copy_probe not built: device copy bandwidth not measured", file=sys.stderr, flush=True)
        return None
    res = subprocess.run([exe, "20"], capture_output=True, text=True, timeout=120)
    if res.returncode != 0:
        print(f"[bench] copy_probe failed: {res.stderr.strip()}", file=sys.stderr, flush=True)
        return None
    recs = {r["kernel"]: r for r in (json.loads(l) for l in res.stdout.strip().splitlines() if l.startswith("{"))}
    rec, nt = recs["copy_kernel"], recs.get("copy_nt_kernel")
    out = {"GBps": rec["GBps"], "frac_of_8TBps": round(rec["GBps"] / 8000.0, 4), "avg_us": rec["avg_us"],
           "bytes_per_launch": rec["bytes_per_launch"], "kernel": "copy_kernel (k-wave-swift_amd/csrc/copy_probe.hip)",
           "note": "plain grid-stride copy (the PMC calibration kernel)"}
    if nt:
        out["nontemporal_copy"] = {"GBps": nt["GBps"], "frac_of_8TBps": round(nt["GBps"] / 8000.0, 4),
                                   "avg_us": nt["avg_us"], "kernel": "copy_nt_kernel",
                                   "note": "one pass, 16 B per lane, nontemporal loads and stores: the fastest "
                                           "copy found; not a ceiling (strided_inv sustains more by PMC)"}
    return out


def cpu_baseline(grid, med, src, sens, opts, steps, workers, dtype=np.float64):
    """Time the oracle (numpy/scipy pocketfft, `workers` threads) computing in `dtype` (float32:
    state, operators and FFTs all single precision): 1 warmup step (the t = 0 step, which also
    applies p0), then `steps` timed steps; the rate is taken from the median step time."""
    from oracle import kwave_oracle as O
    out = O.kspace_first_order(grid, med, src, sens, opts, dtype=dtype, workers=workers, t_end=steps + 1)
    sec = float(np.median(out["_step_seconds"][1:]))
    return grid.totalGridPoints / sec / 1e6, sec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)   # at least 2: step 0 is never timed
    ap.add_argument("--workload", default="c2")   # c2 (default), c0, slab, or refbench (the reference's scenarios)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    args = ap.parse_args()
    args.warmup = max(args.warmup, 2)   # step 0 (code-object loading, t = 0 kernels) plus one steady step
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "refbench":
        res = reference_scenarios()
        print(json.dumps({"metric": "Mcell-steps/s per KWaveBench scenario (whole solver call, median of 5)",
                          "unit": "Mcell-steps/s", "n_gpus": 1, "dtype": "f32", "data": "synthetic",
                          "scenarios": res}), flush=True)
        return

    from kwavehip import Simulation
    t_setup = time.perf_counter()
    grid, med, src, sens, opts, info, model = workload(args.workload, args.dtype)
    total = args.warmup + args.steps
    if total > grid.nt:
        grid.setTime(total, grid.dt)
    media_s = time.perf_counter() - t_setup
    t_create = time.perf_counter()
    sim = Simulation(grid, med, src, sens, opts)   # host -> device upload + operator setup
    create_s = time.perf_counter() - t_create
    setup_s = media_s + create_s
    print(f"[bench] setup done in {setup_s:.1f}s ({info['workload']})", file=sys.stderr, flush=True)

    # Warmup: step 0 untimed (it loads the code objects and runs the one-time t = 0 kernels), then
    # every kernel launch of the remaining warmup steps timed with HIP events, to name the
    # dominant kernel of the steady-state step.
    timing = not args.no_kernel_timing
    sim.launch(0, 1)
    sim.sync()
    if timing:
        sim.set_kernel_timing(True)
    sim.launch(1, args.warmup)
    sim.sync()
    warm_stats = sim.kernel_stats() if timing else []
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    # the dominant kernel: the largest share of the step; a near-tie (within 10%) goes to the kernel
    # with fewer launches per step, whose events perturb the timed region least
    dom_name = None
    if warm_stats:
        top = max(x["total_ms"] for x in warm_stats)
        near = [x for x in warm_stats if x["total_ms"] >= 0.9 * top]
        dom_name = min(near, key=lambda x: (x["launches"], -x["total_ms"]))["name"]
    # Timed region: events only around the dominant kernel's launches (on the simulation stream).
    if dom_name:
        sim.set_kernel_timing(True, only=dom_name)
    else:
        sim.set_kernel_timing(False)
    sim.sync()
    t0 = time.perf_counter()
    sim.launch(args.warmup, total)
    sim.sync()
    elapsed = time.perf_counter() - t0
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f}s", file=sys.stderr, flush=True)
    stats = sim.kernel_stats() if dom_name else []
    t_read = time.perf_counter()
    outputs = sim.outputs()                         # device -> host copy of the recorded series
    read_s = time.perf_counter() - t_read
    N = grid.totalGridPoints
    value = N * args.steps / elapsed / 1e6
    ms_step = elapsed / args.steps * 1e3
    s = 8 if args.dtype == "f64" else 4
    model_bytes = s * N * (2 * model["d"] * model["T"] + model["S"])

    roof = None
    kernels = []
    for st in warm_stats:
        avg_ms = st["total_ms"] / max(st["launches"], 1)
        kernels.append({"name": st["name"], "launches": st["launches"], "avg_us": round(avg_ms * 1e3, 2),
                        "ms_per_step": round(st["total_ms"] / max(args.warmup - 1, 1), 4),
                        "GBps": round(st["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9, 1) if avg_ms > 0 else None})
    dom = next((st for st in stats if st["name"] == dom_name and st["launches"]), None)
    if dom:
        avg_s = dom["total_ms"] / dom["launches"] * 1e-3
        ach = dom["bytes_per_launch"] / avg_s / 1e9
        traffic = measured_traffic(args.workload, args.dtype, dom["name"])
        roof = {"bound": "hbm", "kernel": dom["name"], "achieved": round(ach, 1), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": round(ach * 1e9 / HBM_PEAK, 4),
                "traffic": traffic["bytes"] if traffic else None,
                "traffic_source": traffic["source"] if traffic else None,
                "bytes_per_launch": dom["bytes_per_launch"], "avg_launch_us": round(avg_s * 1e6, 2),
                "launches_timed": dom["launches"]}
    step_model = {"bytes_per_step": model_bytes, "GBps": round(model_bytes / (ms_step * 1e-3) / 1e9, 1),
                  "frac_of_8TBps": round(model_bytes / (ms_step * 1e-3) / HBM_PEAK, 4),
                  "note": "algorithmic byte model of the step (transform passes + full-grid arrays), not measured "
                          "traffic; see step_measured for the DRAM bytes the counters saw"}
    # DRAM bytes per step from the committed PMC summary (every counted kernel x its launches per
    # step in the warmup steps) at the measured step time
    step_measured = None
    if warm_stats:
        per_step, missing = 0, []
        for st in warm_stats:
            tr = measured_traffic(args.workload, args.dtype, st["name"])
            if tr is None:
                missing.append(st["name"])
                continue
            per_step += tr["bytes"] * st["launches"] / max(args.warmup - 1, 1)
        if per_step:
            step_measured = {"bytes_per_step": int(per_step),
                             "GBps": round(per_step / (ms_step * 1e-3) / 1e9, 1),
                             "frac_of_8TBps": round(per_step / (ms_step * 1e-3) / HBM_PEAK, 4),
                             "source": "profiles/pmc_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE per launch)",
                             "kernels_without_counts": missing}
    sim.close()

    del outputs
    fp64 = None
    if args.dtype == "f32" and not args.no_fp64_check:
        print(f"[bench] full {grid.nt}-step runs in FP32 and FP64 for the precision comparison", file=sys.stderr,
              flush=True)
        fp64 = precision_check(grid, med, src, sens, opts)

    batch = None
    if not args.no_batch and args.workload == "c2":
        print("[bench] batched scan lines: 48 runs, batched vs sequential", file=sys.stderr, flush=True)
        batch = batch_comparison()
        print(f"[bench] batch speedup {batch['speedup']}", file=sys.stderr, flush=True)

    adjoint = None
    if not args.no_adjoint and args.workload == "c2":
        print("[bench] adjoint: 2D 512^2, 1500 steps, checkpoints every 32", file=sys.stderr, flush=True)
        adjoint = adjoint_comparison()
        print(f"[bench] adjoint/forward {adjoint['adjoint_over_forward']}", file=sys.stderr, flush=True)

    copy_bw = device_copy_bandwidth()
    if copy_bw:
        print(f"[bench] device copy {copy_bw['GBps']:.0f} GB/s", file=sys.stderr, flush=True)

    cpu = cpu_other = None
    if not args.no_cpu_baseline:
        # every core of the process's share: os.cpu_count(), capped at 16, the CPU share of one GPU on
        # the GPU boxes (their os.cpu_count() reports the whole node)
        ncpu = os.cpu_count() or 1
        workers = min(ncpu, 16)
        legs = {}
        for tag, npdt in (("f32", np.float32), ("f64", np.float64)):
            print(f"[bench] cpu baseline: {tag} oracle, {args.cpu_steps} steps, {workers} threads", file=sys.stderr,
                  flush=True)
            v, sec = cpu_baseline(grid, med, src, sens, opts, args.cpu_steps, workers, npdt)
            legs[tag] = {"value": round(v, 3), "unit": "Mcell-steps/s", "cores": workers, "kind": "port",
                         "dtype": tag,
                         "cores_note": f"{workers} threads = min(os.cpu_count() = {ncpu}, 16: the CPU share of "
                                       f"one GPU on the GPU box)",
                         "sample": f"steps 1..{args.cpu_steps} of the same workload after the t = 0 step as "
                                   f"warmup, median step {sec:.2f} s, numpy/scipy oracle in "
                                   f"{'float32' if tag == 'f32' else 'float64'}"}
            print(f"[bench] cpu {tag}: {v:.2f} Mcell-steps/s", file=sys.stderr, flush=True)
        # the headline baseline computes in the run's own precision; the other one is reported beside it
        cpu = legs[args.dtype]
        cpu_other = legs["f64" if args.dtype == "f32" else "f32"]

    line = {
        "metric": "Mcell-steps/s 3D 256^3 FP32 (err vs FP64); adjoint/forward time; batched vs plain",
        "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.dtype, "data": "synthetic (seeded random media)",
        "config": dict(info, dtype=args.dtype),
        "roofline": roof, "step_model": step_model, "step_measured": step_measured, "device_copy": copy_bw, "kernels": kernels, "cpu_baseline": cpu,
        "cpu_baseline_other_dtype": cpu_other, "fp64_gpu": fp64,
        "batched_scan_lines": batch, "adjoint": adjoint,
        "setup_s": round(setup_s, 2),
        # whole job seen from the host boundary: upload + setup, the timed steps, read-back
        "host_boundary": {"create_s": round(create_s, 3), "read_s": round(read_s, 3),
                          "Mcell_steps_per_s": round(N * args.steps / (create_s + elapsed + read_s) / 1e6, 2)},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
