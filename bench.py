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
        n, dx = 256, 1e-4
        c = random_field((n, n, n), 3, 20.0, 1400.0, 1600.0)
        rho = random_field((n, n, n), 4, 20.0, 900.0, 1100.0)
        med = KWaveMedium(c, rho, alphaCoeff=0.5, alphaPower=1.1, alphaMode=AbsorptionMode.powerLaw)
        grid = KWaveGrid(n, dx, n, dx, n, dx)
        grid.setTime(1000, 0.3 * dx / float(c.max()))
        src = KWaveSource(p0=makeBall(n, n, n, radius=8))
        mask = np.zeros((n, n, n), np.float32)
        mask[168, :, :] = 1
        sens = KWaveSensor(mask=mask, record=RecordField.p | RecordField.uz)
        opts = SimulationOptions(pmlSize=10, smoothP0=True, dtype=dt_)
        info = {"workload": "3D 256^3 heterogeneous power-law (y=1.1) p0-ball, plane sensor p+uz",
                "grid": [n, n, n], "pml": 10, "nt": 1000}
        # byte model per step: s*N*(2*d*T + S), T = 15 transforms, S = 27 other full-grid arrays
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


def cpu_baseline(grid, med, src, sens, opts, steps, workers):
    """Time the float64 oracle (numpy/scipy pocketfft, `workers` threads) on `steps` steps."""
    from oracle import kwave_oracle as O
    out = O.kspace_first_order(grid, med, src, sens, opts, workers=workers, t_end=steps)
    sec = out["_loop_seconds"]
    return grid.totalGridPoints * steps / sec / 1e6, sec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

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

    # Warmup: every kernel launch timed with HIP events, to name the dominant kernel.
    timing = not args.no_kernel_timing
    if timing:
        sim.set_kernel_timing(True)
    sim.launch(0, args.warmup)
    sim.sync()
    warm_stats = sim.kernel_stats() if timing else []
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    dom_name = max(warm_stats, key=lambda x: x["total_ms"])["name"] if warm_stats else None
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
                        "ms_per_step": round(st["total_ms"] / max(args.warmup, 1), 4),
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
                  "frac_of_8TBps": round(model_bytes / (ms_step * 1e-3) / HBM_PEAK, 4)}
    sim.close()

    del outputs
    fp64 = None
    if args.dtype == "f32" and not args.no_fp64_check:
        print(f"[bench] full {grid.nt}-step runs in FP32 and FP64 for the precision comparison", file=sys.stderr,
              flush=True)
        fp64 = precision_check(grid, med, src, sens, opts)

    cpu = None
    if not args.no_cpu_baseline:
        workers = min(os.cpu_count() or 1, 16)
        print(f"[bench] cpu baseline: float64 oracle, {args.cpu_steps} steps, {workers} threads", file=sys.stderr,
              flush=True)
        v, sec = cpu_baseline(grid, med, src, sens, opts, args.cpu_steps, workers)
        cpu = {"value": round(v, 3), "unit": "Mcell-steps/s", "cores": workers, "kind": "port",
               "sample": f"first {args.cpu_steps} steps of the same workload, float64 numpy/scipy oracle, "
                         f"{sec:.1f}s of loop time"}

    line = {
        "metric": "Mcell-steps/s 3D 256^3 FP32 (err vs FP64); adjoint/forward time; batched vs plain",
        "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.dtype, "data": "synthetic (seeded random media)",
        "config": dict(info, dtype=args.dtype),
        "roofline": roof, "step_model": step_model, "kernels": kernels, "cpu_baseline": cpu, "fp64_gpu": fp64,
        "setup_s": round(setup_s, 2),
        # whole job seen from the host boundary: upload + setup, the timed steps, read-back
        "host_boundary": {"create_s": round(create_s, 3), "read_s": round(read_s, 3),
                          "Mcell_steps_per_s": round(N * args.steps / (create_s + elapsed + read_s) / 1e6, 2)},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
