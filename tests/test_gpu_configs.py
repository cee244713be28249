"""GPU parity at the workloads' own sizes (kwavehip.workloads), through the C ABI, against the
float64 oracle on identical float32 inputs:

- 512^2 heterogeneous absorbing nonlinear run (power law y = 1.5, B/A = 6, 64-point Hann line
  source of 1 MPa), the full 2000 steps: full-grid pMax / pRms and the 128-point line series, in
  float32 and float64;
- per-scan-line transmits at 96^3, 128x96x96 and 160x128x128 (32-element array, delayed velocity
  drive), 120 steps each against the oracle, receive traces combined on the device, and the three
  shapes as one fused batch bitwise equal to the single runs;
- 512^2 sound-speed inversion: the adjoint gradient against the float64 adjoint oracle (600 steps),
  and over the full 1500 steps against central differences of the float64 forward on 16 random
  directions;
- the float64 length-256 register kernels: 256^2 and the 256^3 power-law workload against the
  oracle at 1e-9.
Tolerances are the reference's parity tolerances (ParityTests.swift:54-58): float32 relL2 <= 1e-4
and relMax <= 1e-3, float64 relL2 <= 1e-9.
"""
import ctypes as C
import dataclasses
import math
import os

import numpy as np
import pytest

from oracle import kwave_adjoint as A
from oracle import kwave_oracle as O
from kwavehip import (DTypePrecision, KWaveMedium, KWaveSource, Simulation, SimulationBatch, kspaceFirstOrder,
                      workloads as W)
from kwavehip import _lib as L

pytestmark = pytest.mark.gpu
F32, F64 = DTypePrecision.float32, DTypePrecision.float64
WORKERS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = L.load()
    h = C.c_void_p()
    if lib.kwh_ctx_create(0, C.byref(h)) != 0:
        pytest.skip("no HIP device")
    lib.kwh_ctx_destroy(h)


def metrics(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = a - b
    return (math.sqrt(np.mean(d * d)) / max(math.sqrt(np.mean(b * b)), 1e-300),
            np.abs(d).max() / max(np.abs(b).max(), 1e-300))


def assert_close(name, got, ref, dtype, scale=None):
    assert got is not None and got.shape == ref.shape, (name, None if got is None else got.shape, ref.shape)
    assert np.all(np.isfinite(got)), name
    rl2, rmax = metrics(got, ref)
    if scale is not None:
        d = np.asarray(got, np.float64) - ref
        rl2, rmax = math.sqrt(np.mean(d * d)) / scale, np.abs(d).max() / scale
    l2, mx = (1e-4, 1e-3) if dtype == F32 else (1e-9, 1e-8)
    print(f"  {name:10s} relL2={rl2:.3e} relMax={rmax:.3e}", flush=True)
    assert rl2 <= l2 and rmax <= mx, (name, rl2, rmax)


def _progress(label, every):
    def cb(t):
        if (t + 1) % every == 0:
            print(f"  oracle {label}: step {t + 1}", flush=True)
    return cb


# ----------------------------------------------------------- 512^2 heterogeneous nonlinear run
@pytest.fixture(scope="module")
def nonlinear_2d():
    g, med, src, full, line, opts = W.heterogeneous_nonlinear_2d()
    print(f"\n[case] oracle 512^2 nonlinear, {g.nt} steps, {WORKERS} threads", flush=True)
    ref = O.kspace_first_order(g, med, src, [full, line], opts, workers=WORKERS,
                               progress=_progress("512^2", 250))
    return (g, med, src, full, line, opts), ref


@pytest.mark.parametrize("dtype", [F32, F64])
def test_nonlinear_2d_512_full_run(nonlinear_2d, dtype):
    (g, med, src, full, line, opts), (rfull, rline) = nonlinear_2d
    print(f"\n[case] 512^2 nonlinear {dtype.value}, {g.nt} steps", flush=True)
    o = dataclasses.replace(opts, dtype=dtype)
    a = kspaceFirstOrder(g, med, src, full, o)
    b = kspaceFirstOrder(g, med, src, line, o)
    assert float(np.abs(rline["p"]).max()) > 1e4            # the burst reached the line
    for f in ("pMax", "pRms"):
        assert_close(f, getattr(a, f), rfull[f], dtype)
    assert_close("line p", b.p, rline["p"], dtype)
    assert_close("pFinal", a.pFinal, rfull["pFinal"], dtype, scale=float(np.abs(rfull["pMax"]).max()))


# ------------------------------------------------------------------------ scan-line transmits
SCAN_STEPS = 120


def host_rows(src):
    """The transducer drive as the reference builds it: one row per active voxel, the burst
    delayed and apodised per voxel (KWaveTransducer.swift:185-258); the library synthesises the
    same rows on the device from the single burst, the delays and the scales."""
    row = np.asarray(src.ux, np.float32)
    d = np.asarray(src.uDelay)
    s = np.asarray(src.uScale, np.float32)
    rows = np.zeros((d.size, row.size + int(d.max())), np.float32)
    for k in range(d.size):
        rows[k, d[k]:d[k] + row.size] = s[k] * row
    return KWaveSource(uMask=src.uMask, ux=rows, uMode=src.uMode)


@pytest.mark.parametrize("k", range(3))
def test_scan_line_transmit_full_size(k):
    shape = W.SCAN_LINE_SHAPES[k]
    args, tr = W.scan_line_transmit(k, shape, steps=SCAN_STEPS)
    print(f"\n[case] scan line {shape}, {SCAN_STEPS} steps", flush=True)
    sim = Simulation(*args)
    sim.run()
    got = sim.output(L.OUT_P)
    traces = sim.transducer_traces(tr)
    sim.close()
    grid, med, src, sens, opts = args
    ref = O.kspace_first_order(grid, med, host_rows(src), sens, opts, workers=WORKERS)
    assert float(np.abs(ref["p"]).max()) > 0
    assert_close("p", got, ref["p"], F32)
    assert_close("traces", traces, tr.combineSensorData(ref["p"]), F32)


def test_scan_line_shapes_fused_batch_equals_single_runs():
    cases = [W.scan_line_transmit(10 + k, s, steps=SCAN_STEPS) for k, s in enumerate(W.SCAN_LINE_SHAPES * 2)]
    print(f"\n[case] fused batch of {len(cases)} full-size transmits, {SCAN_STEPS} steps", flush=True)
    single = []
    for args, tr in cases:
        sim = Simulation(*args)
        sim.run()
        single.append(sim.transducer_traces(tr))
        sim.close()
    batch = SimulationBatch([c[0] for c in cases])
    batch.run(use_graph=True, fused=True)
    for i, (sim, (_, tr)) in enumerate(zip(batch.sims, cases)):
        assert np.array_equal(sim.transducer_traces(tr), single[i]), i
    batch.close()


# ------------------------------------------------------------------- 512^2 adjoint gradient
def _forward(grid, c, p0, sens, opts, dtype, ckpt=0):
    sim = Simulation(grid, KWaveMedium(c, np.float32(1000.0)), KWaveSource(p0=p0), sens,
                     dataclasses.replace(opts, dtype=dtype))
    if ckpt:
        sim.set_checkpointing(ckpt)
    sim.run()
    return sim, sim.outputs().p.astype(np.float64)


def test_inversion_gradient_512_against_adjoint_oracle():
    grid, c, c_obs, p0, sens, opts = W.inversion_2d(steps=600)
    print(f"\n[case] 512^2 adjoint vs oracle, {grid.nt} steps", flush=True)
    obs, y_obs = _forward(grid, c_obs, p0, sens, opts, F64)
    obs.close()
    s64, y64 = _forward(grid, c, p0, sens, opts, F64, ckpt=32)
    dldy = y64 - y_obs
    assert np.abs(dldy).max() > 1e-3 * np.abs(y64).max()     # the sensors see the perturbation
    gp0, gc = s64.vjp(dldy, checkpoint_every=32)
    s64.close()
    s32, _ = _forward(grid, c, p0, sens, opts, F32, ckpt=32)
    fp0, fc = s32.vjp(dldy.astype(np.float32), checkpoint_every=32)
    s32.close()
    ops = A.LosslessOperators(grid, np.float32(1000.0), sens.mask, opts, float(np.max(c)), workers=WORKERS)
    rp0, rc = A.vjp(ops, np.asarray(c, np.float64), np.asarray(p0, np.float64), dldy)
    for name, got, ref, dt in (("grad_p0 64", gp0, rp0, F64), ("grad_c 64", gc, rc, F64),
                               ("grad_p0 32", fp0, rp0, F32), ("grad_c 32", fc, rc, F32)):
        rl2, _ = metrics(got, ref)
        print(f"  {name:10s} relL2={rl2:.3e}", flush=True)
        assert rl2 <= (1e-9 if dt == F64 else 1e-4), (name, rl2)


def test_inversion_gradient_512_against_central_differences():
    """The float32 c-gradient over the full 1500 steps against central differences of the float64
    forward along 16 random sign directions that leave max(c) (hence c_ref) alone; c enters as
    float32, so steps of 0.125 m/s are exact."""
    grid, c, c_obs, p0, sens, opts = W.inversion_2d()
    print(f"\n[case] 512^2 adjoint vs central differences, {grid.nt} steps, 16 directions", flush=True)
    obs, y_obs = _forward(grid, c_obs, p0, sens, opts, F64)
    obs.close()
    s32, y = _forward(grid, c, p0, sens, opts, F32, ckpt=32)
    _, gc = s32.vjp((y - y_obs).astype(np.float32), checkpoint_every=32)
    s32.close()
    c64 = c.astype(np.float64)
    keep = c64 < c64.max() - 20.0
    g64 = gc.astype(np.float64)
    scale = float(np.sqrt(np.sum(np.where(keep, g64, 0.0) ** 2)))
    rng = np.random.default_rng(14)
    errs, rels = [], []
    for i in range(16):
        v = np.where(keep, np.sign(rng.standard_normal(c.shape)), 0.0)
        h = 0.125
        loss = []
        for sgn in (1, -1):
            sim, yy = _forward(grid, (c64 + sgn * h * v).astype(np.float32), p0, sens, opts, F64)
            sim.close()
            loss.append(0.5 * np.sum((yy - y_obs) ** 2))
        fd = (loss[0] - loss[1]) / (2 * h)
        an = float(np.sum(g64 * v))
        errs.append(abs(fd - an) / scale)
        rels.append(abs(fd - an) / max(abs(fd), 1e-300))
        print(f"  direction {i:2d}: fd={fd:+.6e} adjoint={an:+.6e} rel={rels[-1]:.2e}", flush=True)
    print(f"  max |fd - adj| / ||g|| = {max(errs):.3e}; per-direction relative error max {max(rels):.3e}, "
          f"median {float(np.median(rels)):.3e}", flush=True)
    assert max(errs) <= 1e-3


# ---------------------------------------------------------------- float64 length-256 kernels
def test_fp64_256_2d_heterogeneous_absorbing():
    """256^2: the float64 length-256 register kernels as contiguous and middle axis."""
    g, med, src, full, line, opts = W.heterogeneous_nonlinear_2d(n=256, steps=500)
    print(f"\n[case] 256^2 nonlinear float64, {g.nt} steps", flush=True)
    rfull, rline = O.kspace_first_order(g, med, src, [full, line], opts, workers=WORKERS)
    o = dataclasses.replace(opts, dtype=F64)
    a = kspaceFirstOrder(g, med, src, full, o)
    b = kspaceFirstOrder(g, med, src, line, o)
    for f in ("pMax", "pRms"):
        assert_close(f, getattr(a, f), rfull[f], F64)
    assert_close("line p", b.p, rline["p"], F64)


@pytest.fixture(scope="module")
def throughput_3d_short():
    args = W.throughput_3d(steps=6, plane=136)     # a plane close enough to the ball for 6 steps
    print(f"\n[case] oracle 256^3 power law, 6 steps, {WORKERS} threads", flush=True)
    return args, O.kspace_first_order(*args[:4], args[4], workers=WORKERS)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_3d_256_power_law_short(throughput_3d_short, dtype):
    """The 256^3 workload (random-field media, power law y = 1.1, p0 ball, plane p + uz): the
    float32 and float64 length-256 register kernels on every axis."""
    (g, med, src, sens, opts), ref = throughput_3d_short
    print(f"\n[case] 256^3 {dtype.value}, 6 steps", flush=True)
    out = kspaceFirstOrder(g, med, src, sens, dataclasses.replace(opts, dtype=dtype))
    for f in ("p", "uz", "pFinal"):
        assert_close(f, getattr(out, f), ref[f], dtype)
