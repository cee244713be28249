"""GPU: transducer drive synthesised on the device, receive combination on the device, and the
batched runner (kwh_batch_*) against single runs, through the C ABI."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (AbsorptionMode, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, KWaveTransducer, KWaveTransducerGeometry,
                      RecordField, Simulation, SimulationBatch, SimulationOptions, SourceMode, toneBurst)
from kwavehip import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = L.load()
    h = C.c_void_p()
    if lib.kwh_ctx_create(0, C.byref(h)) != 0:
        pytest.skip("no HIP device")
    lib.kwh_ctx_destroy(h)


def scan_line_case(shape, seed, steer, focus, nt=140, dx=1e-4):
    """One transmit of a 16-element array (width 2, length 8) at x = 4: speckle medium, delayed
    velocity drive, sensor on the active voxels recording p."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    c = (1540.0 * (1 + rng.uniform(-0.02, 0.02, shape))).astype(np.float32)
    rho = (1000.0 * (1 + rng.uniform(-0.02, 0.02, shape))).astype(np.float32)
    grid = KWaveGrid(nx, dx, ny, dx, nz, dx)
    grid.setTime(nt, 0.3 * dx / float(c.max()))
    geom = KWaveTransducerGeometry(numberElements=16, elementWidth=2, elementLength=8, elementSpacing=0,
                                   position=(4, (ny - 32) // 2, (nz - 8) // 2), dy=dx, dz=dx)
    burst = (0.5 * toneBurst(1 / grid.dt, 2e6, 2)).astype(np.float32)
    tr = KWaveTransducer(geom, shape, grid.dt, 1540.0, focusDistance=focus, elevationFocusDistance=6e-3,
                         steeringAngle=steer, transmitApodization="hanning", inputSignal=burst)
    mask, row, delays, scales = tr.velocity_source()
    src = KWaveSource(uMask=mask, ux=row, uDelay=delays, uScale=scales, uMode=SourceMode.additive)
    sens = KWaveSensor(mask=mask, record=RecordField.p)
    return grid, KWaveMedium(c, rho), src, sens, SimulationOptions(pmlSize=6, smoothP0=False), tr


def expanded(src, grid):
    """The same drive with the per-point rows built on the host."""
    row = np.asarray(src.ux, np.float32)
    d = np.asarray(src.uDelay)
    s = np.asarray(src.uScale, np.float32)
    L_ = row.size + int(d.max())
    rows = np.zeros((d.size, L_), np.float32)
    for k in range(d.size):
        rows[k, d[k]:d[k] + row.size] = s[k] * row
    return KWaveSource(uMask=src.uMask, ux=rows, uMode=src.uMode)


def test_device_drive_synthesis_matches_host_rows_and_oracle():
    grid, med, src, sens, opts, tr = scan_line_case((32, 48, 24), 1, 12.0, 8e-3, nt=90)
    print("\n[case] delayed-copy drive, 32x48x24", flush=True)
    a = Simulation(grid, med, src, sens, opts)
    a.run()
    pa = a.outputs().p
    a.close()
    host = expanded(src, grid)
    b = Simulation(grid, med, host, sens, opts)
    b.run()
    pb = b.outputs().p
    b.close()
    assert np.array_equal(pa, pb)
    ref = O.kspace_first_order(grid, med, host, sens, opts, workers=8)["p"]
    d = np.asarray(pa, np.float64) - ref
    rl2 = math.sqrt(np.mean(d * d)) / math.sqrt(np.mean(ref * ref))
    print(f"  p relL2={rl2:.3e}", flush=True)
    assert rl2 < 1e-4


def test_element_traces_match_host_combination():
    grid, med, src, sens, opts, tr = scan_line_case((32, 48, 24), 2, -8.0, np.inf, nt=80)
    print("\n[case] device receive combination", flush=True)
    sim = Simulation(grid, med, src, sens, opts)
    sim.run()
    p = sim.outputs().p
    dev = sim.transducer_traces(tr)
    sim.close()
    host = tr.combineSensorData(p)
    scale = float(np.abs(host).max())
    assert dev.shape == host.shape
    assert np.abs(dev - host).max() <= 1e-6 * scale


@pytest.mark.parametrize("use_graph,fused", [(True, True), (False, True), (True, False), (False, False)])
def test_batch_of_uneven_sizes_equals_single_runs(use_graph, fused):
    """Three sims of different sizes and drives (their sources switch off at different steps and
    their runs end at different steps, so the fused launches regroup mid-run): bitwise equal to
    separate runs."""
    cases = [scan_line_case((32, 48, 24), 3, 0.0, 8e-3, nt=70), scan_line_case((40, 40, 24), 4, 15.0, np.inf, nt=90),
             scan_line_case((36, 48, 32), 5, -10.0, 6e-3, nt=60)]
    print(f"\n[case] batch of 3 uneven sims, graph={use_graph} fused={fused}", flush=True)
    singles = []
    for grid, med, src, sens, opts, _ in cases:
        s = Simulation(grid, med, src, sens, opts)
        s.run()
        singles.append(s.outputs())
        s.close()
    batch = SimulationBatch([c[:5] for c in cases])
    batch.run(use_graph=use_graph, fused=fused)
    for i, (sim, ref) in enumerate(zip(batch.sims, singles)):
        out = sim.outputs()
        assert out.p.shape == ref.p.shape, i
        assert np.array_equal(out.p, ref.p), (i, float(np.abs(out.p - ref.p).max()))
        assert np.array_equal(out.pFinal, ref.pFinal), i
        tr_out = sim.transducer_traces(cases[i][5])
        assert tr_out.shape == (16, cases[i][0].nt)
    batch.close()


def ivp_case(shape, seed, nt, absorbing=False, nonlinear=False):
    """2D initial-value problem on a random medium, binary line sensor recording p / pMax / uy."""
    rng = np.random.default_rng(seed)
    nx, ny = shape
    c = (1500.0 * (1 + rng.uniform(-0.05, 0.05, shape))).astype(np.float32)
    rho = (1000.0 * (1 + rng.uniform(-0.05, 0.05, shape))).astype(np.float32)
    grid = KWaveGrid(nx, 1e-4, ny, 1e-4)
    grid.setTime(nt, 0.3 * 1e-4 / float(c.max()))
    p0 = np.zeros(shape, np.float32)
    p0[nx // 2 - 3:nx // 2 + 3, ny // 3:ny // 3 + 5] = 1.0
    mask = np.zeros(shape, np.float32)
    mask[nx // 4, 2:ny - 2] = 1
    kw = {}
    if absorbing:
        kw.update(alphaCoeff=0.75, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw)
    if nonlinear:
        kw.update(bOnA=6.0)
    med = KWaveMedium(c, rho, **kw)
    sens = KWaveSensor(mask=mask, record=RecordField.p | RecordField.pMax | RecordField.uy)
    return grid, med, KWaveSource(p0=p0), sens, SimulationOptions(pmlSize=8, smoothP0=False)


def test_fused_batch_register_resident_mix():
    """Fused launches over sims of the register-resident path (64^2, 256^2), a generic size,
    absorbing and nonlinear media: sims with equal launch sequences share launches, the rest run
    in their own groups; every output bitwise equal to separate runs."""
    cases = [ivp_case((64, 64), 1, 60), ivp_case((64, 64), 2, 70), ivp_case((256, 256), 3, 40),
             ivp_case((48, 64), 4, 50), ivp_case((64, 64), 5, 60, absorbing=True),
             ivp_case((64, 64), 6, 55, absorbing=True, nonlinear=True), ivp_case((256, 256), 7, 30, absorbing=True)]
    print("\n[case] fused batch, register-resident / generic / absorbing / nonlinear mix", flush=True)
    singles = []
    for case in cases:
        s = Simulation(*case)
        s.run()
        singles.append(s.outputs())
        s.close()
    batch = SimulationBatch(cases)
    batch.run(use_graph=True, fused=True)
    for i, (sim, ref) in enumerate(zip(batch.sims, singles)):
        out = sim.outputs()
        for f in ("p", "pMax", "uy", "pFinal"):
            a, b = getattr(out, f), getattr(ref, f)
            assert np.array_equal(a, b), (i, f, float(np.abs(a - b).max()))
    batch.close()
